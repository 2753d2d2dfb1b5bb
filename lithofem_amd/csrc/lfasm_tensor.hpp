// Precontracted element matrices for constant-coefficient elements.
//
// Outside the PML the stretch is s = 1 at every quadrature point
// (solver/problem_common.inc:69-98), so Lambda = I and the five integrators
// of the reference (solver/asm_gpu.inc:70-214) have element-constant
// coefficients.  Their quadrature sums then factor into element
// coefficients times reference-element tensors that depend only on the
// basis and the rule:
//
//   curl-curl   sum_ab (G_ab / detJ) Rcc_ab[k,l],   G = J^T J
//   mass + KLK  sum_ab Mc_ab Rm_ab[k,l],
//               Mc = detJ Ji (-k0^2 eps I - K^2) Ji^T   (Ji = J^{-1})
//   CROSS0+1    i sum_ab X_ab Q_ab[k,l],           X = Ji K J
//
// with Rcc_ab = sum_q w_q (c_ka c_lb + c_kb c_la) over the curl-curl rule
// (one term when a = b), Rm_ab the same with the value shapes on the mass
// rule, Q_ab = sum_q w_q (v_ka c_lb - v_la c_kb) on the mass rule (CROSS1 =
// -CROSS0^T), K the cross-product matrix of k_par.  21 real tensors of
// nd x nd; an element matrix is 27 multiply-adds per entry instead of the
// O(nq nd^2) quadrature contraction, so the kernel is bound by the HBM
// write of Ae.  This is the reference-tensor form of the element integral
// (Kirby and Logg, "A compiler for variational forms", ACM TOMS 32, 2006).
//
// Layout: thread <-> 2x2 block (rows 2rp, 2rp+1; columns 2cp, 2cp+1), so
// each face-DOF pair sits in one thread and the dual face transform is
// applied in registers before the single write of the block.
#pragma once

#include "lfasm_device.hpp"

namespace lfasm {

constexpr int NTENS = 21;     // 6 curl-curl, 6 mass, 9 cross reference tensors
constexpr int TEN_EPC = 64;   // elements per workgroup (coefficients staged in LDS)
constexpr int TEN_NC = 28;    // doubles per element: 6 cc, 6 complex mass, 9 cross, pad

// symmetric index pairs (a, b) of the six-term tensors:
// t = 0..5 -> (0,0) (1,1) (2,2) (0,1) (0,2) (1,2)
__host__ __device__ constexpr int sym_a(int t) { return t < 3 ? t : (t == 5 ? 1 : 0); }
__host__ __device__ constexpr int sym_b(int t) { return t < 3 ? t : (t == 3 ? 1 : 2); }

// R[((blk * NTENS) + t) * 4 + s]: tensor t at entry (2rp + s/2, 2cp + s%2)
// of block blk = rp * npair + cp; zero outside nd x nd.
__global__ void k_ref_tensor(int nd, int npair, RuleDev rc, RuleDev rm, double *__restrict__ R)
{
   const int total = npair * npair * NTENS * 4;
   for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total;
        idx += gridDim.x * blockDim.x)
   {
      const int s = idx & 3;
      const int t = (idx >> 2) % NTENS;
      const int blk = (idx >> 2) / NTENS;
      const int rp = blk / npair, cp = blk - rp * npair;
      const int k = 2 * rp + (s >> 1), l = 2 * cp + (s & 1);
      double acc = 0.0;
      if (k < nd && l < nd)
      {
         if (t < 12)
         {
            const RuleDev &Q = t < 6 ? rc : rm;
            const double *tab = t < 6 ? rc.cs : rm.vs;
            const int a = sym_a(t % 6), b = sym_b(t % 6);
            for (int q = 0; q < Q.nq; ++q)
            {
               const double *T = tab + (size_t)q * nd * 3;
               double v = T[k * 3 + a] * T[l * 3 + b];
               if (a != b) { v += T[k * 3 + b] * T[l * 3 + a]; }
               acc += Q.w[q] * v;
            }
         }
         else
         {
            const int a = (t - 12) / 3, b = (t - 12) % 3;
            for (int q = 0; q < rm.nq; ++q)
            {
               const double *V = rm.vs + (size_t)q * nd * 3;
               const double *C = rm.cs + (size_t)q * nd * 3;
               acc += rm.w[q] * (V[k * 3 + a] * C[l * 3 + b] - V[l * 3 + a] * C[k * 3 + b]);
            }
         }
      }
      R[idx] = acc;
   }
}

struct TensorParams {
   int nd, p, npair, nb;   // nb = npair^2 blocks of 2x2
   int nlist;              // number of elements in elist
   const int *elist;       // constant-coefficient elements (output Ae[elist[i]])
   const double *R;        // packed reference tensors (k_ref_tensor)
   const double *J, *Ji, *detJ;
   const int *attr, *fo, *pid, *eps_off;
   const double2 *eps_tab, *eps_elem;
   const PmlDev *pml_tab;
   double TInv[24];
   int use_cc, use_mass, use_klk, use_cross;
   int apply_te;
   double2 *Ae;
};

// Element coefficients (27 doubles) of element e into c[0..27].
__device__ inline void tensor_coef(const TensorParams &P, int e, double *c)
{
   double J[9], Ji[9];
#pragma unroll
   for (int i = 0; i < 9; ++i)
   {
      J[i] = P.J[(size_t)e * 9 + i];
      Ji[i] = P.Ji[(size_t)e * 9 + i];
   }
   const double det = P.detJ[e];
   const int pid = P.pid ? P.pid[e] : 0;
   const PmlDev pm = P.pml_tab[pid];
   const double2 eps = P.eps_elem ? P.eps_elem[e] : P.eps_tab[P.eps_off[pid] + P.attr[e] - 1];
   const double k2 = pm.k0 * pm.k0;
   const double kx = pm.kx, ky = pm.ky;
   // -K^2 = |k|^2 I - k k^T for k = (kx, ky, 0)
   const double nk2[9] = {ky * ky, -kx * ky, 0.0, -kx * ky, kx * kx, 0.0, 0.0, 0.0,
                          kx * kx + ky * ky};
   const double K[9] = {0.0, 0.0, ky, 0.0, 0.0, -kx, -ky, kx, 0.0};
#pragma unroll
   for (int t = 0; t < 6; ++t)
   {
      const int a = sym_a(t), b = sym_b(t);
      double G = 0.0, H = 0.0, L = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
      {
         G += J[i * 3 + a] * J[i * 3 + b];
         H += Ji[a * 3 + i] * Ji[b * 3 + i];
#pragma unroll
         for (int j = 0; j < 3; ++j) { L += Ji[a * 3 + i] * nk2[i * 3 + j] * Ji[b * 3 + j]; }
      }
      c[t] = P.use_cc ? G / det : 0.0;
      const double mr = (P.use_mass ? -k2 * eps.x * H : 0.0) + (P.use_klk ? L : 0.0);
      const double mi = P.use_mass ? -k2 * eps.y * H : 0.0;
      c[6 + 2 * t] = det * mr;
      c[7 + 2 * t] = det * mi;
   }
#pragma unroll
   for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
      {
         double X = 0.0;
#pragma unroll
         for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) { X += Ji[a * 3 + r] * K[r * 3 + q] * J[q * 3 + b]; }
         c[18 + a * 3 + b] = P.use_cross ? X : 0.0;
      }
   c[27] = 0.0;
}

// Face of a DOF pair starting at even index d0 (-1 if not a face pair).
__device__ __forceinline__ int pair_face(int d0, int p)
{
   const int foff = 6 * p, ntd = p * (p - 1);
   return (p >= 2 && d0 >= foff && d0 < foff + 4 * ntd) ? (d0 - foff) / ntd : -1;
}

// grid.x: chunks of TEN_EPC list elements; grid.y: slices of 256 blocks
// (nb > 128), else `epar` = 256 / nb elements run side by side in one
// workgroup.
__global__ void __launch_bounds__(256) k_local_tensor(TensorParams P, int epar)
{
   __shared__ double coef[TEN_EPC * TEN_NC];
   __shared__ int sfo[TEN_EPC * 4];
   __shared__ int sel[TEN_EPC];
   const int tid = threadIdx.x;
   const int c0 = blockIdx.x * TEN_EPC;
   const int cnt = min(TEN_EPC, P.nlist - c0);
   if (tid < cnt)
   {
      const int e = P.elist[c0 + tid];
      sel[tid] = e;
      tensor_coef(P, e, coef + tid * TEN_NC);
      if (P.p >= 2)
      {
#pragma unroll
         for (int f = 0; f < 4; ++f) { sfo[tid * 4 + f] = P.fo[(size_t)e * 4 + f]; }
      }
   }
   __syncthreads();
   const int grp = epar > 1 ? tid / P.nb : 0;
   const int blk = epar > 1 ? tid - grp * P.nb : blockIdx.y * 256 + tid;
   if (grp >= epar || blk >= P.nb) { return; }

   double r[NTENS][4];
   const double *Rb = P.R + (size_t)blk * NTENS * 4;
#pragma unroll
   for (int t = 0; t < NTENS; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) { r[t][s] = Rb[t * 4 + s]; }

   const int nd = P.nd;
   const int rp = blk / P.npair, cp = blk - rp * P.npair;
   const int k0 = 2 * rp, l0 = 2 * cp;
   const int fr = P.apply_te ? pair_face(k0, P.p) : -1;
   const int fc = P.apply_te ? pair_face(l0, P.p) : -1;

   for (int i = grp; i < cnt; i += epar)
   {
      const double *c = coef + i * TEN_NC;
      double re[4], im[4];
#pragma unroll
      for (int s = 0; s < 4; ++s)
      {
         double x = 0.0, y = 0.0;
#pragma unroll
         for (int t = 0; t < 6; ++t)
         {
            x += c[t] * r[t][s];
            x += c[6 + 2 * t] * r[6 + t][s];
            y += c[7 + 2 * t] * r[6 + t][s];
         }
#pragma unroll
         for (int t = 0; t < 9; ++t) { y += c[18 + t] * r[12 + t][s]; }
         re[s] = x;
         im[s] = y;
      }
      // dual face transform: columns, then rows (solver/asm_gpu.inc:221-251)
      if (fc >= 0)
      {
         const double *Ti = P.TInv + 4 * sfo[i * 4 + fc];
#pragma unroll
         for (int row = 0; row < 2; ++row)
         {
            const double ar = re[2 * row], ai = im[2 * row];
            const double br = re[2 * row + 1], bi = im[2 * row + 1];
            re[2 * row] = Ti[0] * ar + Ti[1] * br;
            im[2 * row] = Ti[0] * ai + Ti[1] * bi;
            re[2 * row + 1] = Ti[2] * ar + Ti[3] * br;
            im[2 * row + 1] = Ti[2] * ai + Ti[3] * bi;
         }
      }
      if (fr >= 0)
      {
         const double *Ti = P.TInv + 4 * sfo[i * 4 + fr];
#pragma unroll
         for (int col = 0; col < 2; ++col)
         {
            const double ar = re[col], ai = im[col];
            const double br = re[2 + col], bi = im[2 + col];
            re[col] = Ti[0] * ar + Ti[1] * br;
            im[col] = Ti[0] * ai + Ti[1] * bi;
            re[2 + col] = Ti[2] * ar + Ti[3] * br;
            im[2 + col] = Ti[2] * ai + Ti[3] * bi;
         }
      }
      double2 *A = P.Ae + (size_t)sel[i] * nd * nd;
#pragma unroll
      for (int s = 0; s < 4; ++s)
      {
         const int k = k0 + (s >> 1), l = l0 + (s & 1);
         if (k < nd && l < nd) { A[(size_t)k * nd + l] = make_double2(re[s], im[s]); }
      }
   }
}

}  // namespace lfasm
