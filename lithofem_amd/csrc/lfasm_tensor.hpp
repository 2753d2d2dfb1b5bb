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
      const int k0 = 2 * rp + (s >> 1), l0 = 2 * cp + (s & 1);
      // computed for k <= l and mirrored, so R_cc, R_m are exactly
      // symmetric and Q exactly antisymmetric (the fused row kernels store
      // one triangle)
      const int k = min(k0, l0), l = max(k0, l0);
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
            if (k0 > l0) { acc = -acc; }
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
   int eps_unit;            // 1: eps = 1 in the mass term (unit-eps mass matrices)
   int out_compact;         // 1: list entry i writes Ae[i] (else Ae[elist[i]])
   double2 *Ae;
   int lda;                 // row stride of the element matrices in Ae (>= nd)
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
   const double2 eps = P.eps_unit ? make_double2(1.0, 0.0)
                       : P.eps_elem ? P.eps_elem[e] : P.eps_tab[P.eps_off[pid] + P.attr[e] - 1];
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

// Dual face transform of one DOF pair (lo, hi) with the 2x2 block Ti, with
// explicit FMAs so every kernel that applies it rounds the same way.
__device__ __forceinline__ double pair_lo(const double *Ti, double lo, double hi)
{
   return __builtin_fma(Ti[1], hi, Ti[0] * lo);
}
__device__ __forceinline__ double pair_hi(const double *Ti, double lo, double hi)
{
   return __builtin_fma(Ti[3], hi, Ti[2] * lo);
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
      sel[tid] = P.out_compact ? c0 + tid : e;
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
            x = __builtin_fma(c[t], r[t][s], x);
            x = __builtin_fma(c[6 + 2 * t], r[6 + t][s], x);
            y = __builtin_fma(c[7 + 2 * t], r[6 + t][s], y);
         }
#pragma unroll
         for (int t = 0; t < 9; ++t) { y = __builtin_fma(c[18 + t], r[12 + t][s], y); }
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
            re[2 * row] = pair_lo(Ti, ar, br);
            im[2 * row] = pair_lo(Ti, ai, bi);
            re[2 * row + 1] = pair_hi(Ti, ar, br);
            im[2 * row + 1] = pair_hi(Ti, ai, bi);
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
            re[col] = pair_lo(Ti, ar, br);
            im[col] = pair_lo(Ti, ai, bi);
            re[2 + col] = pair_hi(Ti, ar, br);
            im[2 + col] = pair_hi(Ti, ai, bi);
         }
      }
      double2 *A = P.Ae + (size_t)sel[i] * nd * P.lda;
#pragma unroll
      for (int s = 0; s < 4; ++s)
      {
         const int k = k0 + (s >> 1), l = l0 + (s & 1);
         if (k < nd && l < nd) { A[(size_t)k * P.lda + l] = make_double2(re[s], im[s]); }
      }
   }
}

// Column-per-lane form for padded element matrices (row stride lda a
// multiple of 8 entries, 128-byte aligned rows): a thread owns rows 2rp and
// 2rp+1 at one column l < lda (pad columns are written as zeros), so the
// lanes of a wave cover whole row segments and every store instruction
// writes whole 128-byte lines -- which lets the stores be non-temporal.
// The column face pair (l, l^1) sits in lanes (t, t^1).  Every entry takes
// exactly the FMAs of k_local_tensor, so both give bit-identical values.
typedef double ten_dv2 __attribute__((ext_vector_type(2)));
constexpr int TEN21_EPC = 16;   // elements per workgroup

// The element coefficients come precomputed from k_tensor_coef (28 doubles
// per element, the face orientations packed into the last) and are read
// with wave-uniform loads: they sit in scalar registers and feed the FMAs
// directly, so the vector registers hold only the 42 tensor values and the
// entries (4 waves per SIMD) and no LDS staging or barrier sits between
// element batches.  Measured at p = 3 on 58,752 elements: 0.43 ms against
// 0.48 ms for coefficients staged in LDS by each workgroup (64 elements per
// workgroup; fewer made the staging version slower).
__global__ void __launch_bounds__(256) k_local_tensor21(TensorParams P,
                                                        const int *__restrict__ elist,
                                                        const double *__restrict__ coef,
                                                        double2 *__restrict__ Ae)
{
   const int tid = threadIdx.x;
   const int nd = P.nd, lda = P.lda;
   // face-transform table in LDS (its entries are picked per lane)
   __shared__ double TT[24];
   if (tid < 24) { TT[tid] = P.TInv[tid]; }
   __syncthreads();
   // waves with no entry of the element matrices (past the last row pair)
   if (__builtin_amdgcn_readfirstlane((int)(blockIdx.y * 256 + (tid & ~63))) >= P.npair * lda)
   {
      return;
   }
   const int blk = blockIdx.y * 256 + tid;
   const bool live = blk < P.npair * lda;
   const int bb = live ? blk : 0;
   const int rp = bb / lda, l = bb - rp * lda;
   const int k0 = 2 * rp, odd = l & 1;
   const bool col_ok = l < nd;
   const double *Rb = P.R + ((size_t)rp * P.npair + (col_ok ? l >> 1 : 0)) * NTENS * 4;
   double r[NTENS][2];
#pragma unroll
   for (int t = 0; t < NTENS; ++t)
   {
      r[t][0] = col_ok ? Rb[t * 4 + odd] : 0.0;
      r[t][1] = col_ok ? Rb[t * 4 + 2 + odd] : 0.0;
   }
   const int fr = P.apply_te ? pair_face(k0, P.p) : -1;
   const int fc = (P.apply_te && col_ok) ? pair_face(l - odd, P.p) : -1;
   const bool st0 = live && k0 < nd, st1 = live && k0 + 1 < nd;

   // chunks of TEN21_EPC list elements, grid-stride over blockIdx.x (a grid
   // narrower than the chunk count keeps the kernel to a few workgroups per
   // CU, beside the quadrature kernel)
   const int nchunk = (P.nlist + TEN21_EPC - 1) / TEN21_EPC;
   for (int cx = blockIdx.x; cx < nchunk; cx += gridDim.x)
   {
   const int c0 = cx * TEN21_EPC;
   const int cnt = min(TEN21_EPC, P.nlist - c0);
   for (int i = 0; i < cnt; ++i)
   {
      const int e = elist[c0 + i];
      const double *c = coef + (size_t)e * TEN_NC;
      double re[2], im[2];
#pragma unroll
      for (int s = 0; s < 2; ++s)
      {
         double x = 0.0, y = 0.0;
#pragma unroll
         for (int t = 0; t < 6; ++t)
         {
            x = __builtin_fma(c[t], r[t][s], x);
            x = __builtin_fma(c[6 + 2 * t], r[6 + t][s], x);
            y = __builtin_fma(c[7 + 2 * t], r[6 + t][s], y);
         }
#pragma unroll
         for (int t = 0; t < 9; ++t) { y = __builtin_fma(c[18 + t], r[12 + t][s], y); }
         re[s] = x;
         im[s] = y;
      }
      const int packed = (int)c[27];
      const double *Tc = TT + 4 * (fc >= 0 ? (packed >> (3 * fc)) & 7 : 0);
#pragma unroll
      for (int s = 0; s < 2; ++s)
      {
         const double ur = lane_xor1(re[s]), ui = lane_xor1(im[s]);
         if (fc >= 0)
         {
            if (!odd)
            {
               re[s] = pair_lo(Tc, re[s], ur);
               im[s] = pair_lo(Tc, im[s], ui);
            }
            else
            {
               re[s] = pair_hi(Tc, ur, re[s]);
               im[s] = pair_hi(Tc, ui, im[s]);
            }
         }
      }
      if (fr >= 0)
      {
         const double *Ti = TT + 4 * ((packed >> (3 * fr)) & 7);
         const double ar = re[0], ai = im[0], br = re[1], bi = im[1];
         re[0] = pair_lo(Ti, ar, br);
         im[0] = pair_lo(Ti, ai, bi);
         re[1] = pair_hi(Ti, ar, br);
         im[1] = pair_hi(Ti, ai, bi);
      }
      ten_dv2 *A = reinterpret_cast<ten_dv2 *>(Ae) + (size_t)(P.out_compact ? c0 + i : e) * nd * lda;
      const ten_dv2 v0 = col_ok ? ten_dv2{re[0], im[0]} : ten_dv2{0.0, 0.0};
      const ten_dv2 v1 = col_ok ? ten_dv2{re[1], im[1]} : ten_dv2{0.0, 0.0};
      if (st0) { __builtin_nontemporal_store(v0, A + (size_t)k0 * lda + l); }
      if (st1) { __builtin_nontemporal_store(v1, A + (size_t)(k0 + 1) * lda + l); }
   }
   }
}

}  // namespace lfasm

namespace lfasm {

// ---- reference-tensor rows fused into the row scatter ----------------------
//
// For constant-coefficient elements the element matrix is never written to
// HBM: the CSR row kernel evaluates each contributor row (e, a) from the 27
// element coefficients and the reference tensors held in LDS.  Rows are
// grouped by the type of their local index -- edge [0, 6p), face
// [6p, 6p + 4p(p-1)), interior -- because every element sees a global DOF
// with a local index of the same type; one launch per type keeps only the
// tensor rows of that type in LDS: R_t[a][b] for a in the type's range, all
// b, with the in-range diagonal block stored as a triangle (R_cc, R_m
// symmetric, Q antisymmetric).  Face rows come in pairs (the two DOFs of a
// face point, local a and a+1 in every element) that the dual face
// transform mixes: a unit of work is such a pair of CSR rows, so each
// tensor row is read once.  PML elements' matrices come from Ae as in
// k_scatter_rows.  Contributors are summed in ascending element order, as
// in k_scatter_rows, and each entry is formed with the same FMAs as in
// k_local_tensor, so both paths give bit-identical values.
//
// One workgroup per CU holds the slice, so each wave keeps many loads in
// flight itself: a unit's header is one 32-byte load, its contributor list
// one vector load (class and row signs packed into each entry), the next
// unit's header and list are fetched while the current one is processed,
// and the loads of FG contributors (coefficients as one vector load with a
// lane per coefficient, broadcast by readlane) are issued before their
// in-order accumulation.

// Element coefficients of the tensor path, plus the face orientations
// packed into the spare slot 27 (fo_f in bits 3f..3f+2).
__global__ void __launch_bounds__(256) k_tensor_coef(TensorParams P, double *__restrict__ coef)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= P.nlist) { return; }
   const int e = P.elist[i];
   double c[TEN_NC];
   tensor_coef(P, e, c);
   int packed = 0;
   if (P.p >= 2)
   {
#pragma unroll
      for (int f = 0; f < 4; ++f) { packed |= P.fo[(size_t)e * 4 + f] << (3 * f); }
   }
   c[27] = (double)packed;
#pragma unroll
   for (int t = 0; t < TEN_NC; ++t) { coef[(size_t)e * TEN_NC + t] = c[t]; }
}

// One unit of the fused row kernels: CSR row i0 and its face partner i1
// (-1 if none), their value ranges, and the contributor list of i0
// (i1's list is i0's with a -> a + 1).  flags: bit0 row i0 essential,
// bit1 row i1 essential.
struct alignas(16) UnitHdr {
   int i0, i1, base0, len0, base1, len1, c0, nc_flags;   // nc in bits 8.., flags in 0..7
};

// Contributor entry of the fused kernels: e * nd + a in the low 29 bits,
// bit 31 = PML element (Ae from memory), bit 30 = sign(e, a) < 0,
// bit 29 = sign(e, a + 1) < 0.
constexpr int CE_MASK = (1 << 29) - 1;

struct FusedParams {
   int nd, p, npair;
   int r0, nR, no, S;          // local rows [r0, r0 + nR); no = nd - nR; S = per-tensor stride
   int face;                   // 1: face rows (row transform, pair units)
   int nunit;
   const UnitHdr *hdr;
   int rowcap;                 // longest CSR row of this type
   const int *Jc;
   int eliminate;
   const uint16_t *slot;
   int sld;                    // slot row stride (entries)
   const unsigned char *eflag;
   double TInv[24];
   const double2 *Ae;
   int lda;                    // row stride of the element matrices in Ae
   const double *R;            // blocked reference tensors (k_ref_tensor)
   double2 *vals;
};

constexpr int FG = 4;          // contributors whose loads are issued together
constexpr int FUSED_MAX_WAVES = 12;   // launch bound of k_fused_rows (768 threads)

__device__ __forceinline__ int tri_off(int i, int j, int n) { return i * n - (i * (i - 1)) / 2 + (j - i); }

// LDS offset of R_t[r0 + ar][b] within one tensor's slice; `swapped` marks
// a lower-triangle entry read from its mirror (negated for Q).
__device__ __forceinline__ int slice_off(const FusedParams &F, int ar, int b, bool &swapped)
{
   const int br = b - F.r0;
   swapped = false;
   if (br < 0 || br >= F.nR) { return ar * F.no + (br < 0 ? b : b - F.nR); }
   if (br >= ar) { return F.nR * F.no + tri_off(ar, br, F.nR); }
   swapped = true;
   return F.nR * F.no + tri_off(br, ar, F.nR);
}

__device__ __forceinline__ double ref_entry(const double *R, int npair, int t, int k, int l)
{
   const int blk = (k >> 1) * npair + (l >> 1);
   return R[((size_t)blk * NTENS + t) * 4 + (k & 1) * 2 + (l & 1)];
}

// value of lane `src` (wave-uniform) of a double held one per lane
__device__ __forceinline__ double bcast(double v, int src)
{
   const long long x = __double_as_longlong(v);
   const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)x, src);
   const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), src);
   return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Row ar of the element matrix at this lane's column (before any face
// transform) from coefficients held one per lane in cl; same FMA order as
// k_local_tensor.
__device__ __forceinline__ void tensor_row(const double *Rs, int S, int off, double qs,
                                           double cl, double &x, double &y)
{
   x = 0.0;
   y = 0.0;
#pragma unroll
   for (int t = 0; t < 6; ++t)
   {
      const double rc = Rs[t * S + off];
      const double rm = Rs[(6 + t) * S + off];
      x = __builtin_fma(bcast(cl, t), rc, x);
      x = __builtin_fma(bcast(cl, 6 + 2 * t), rm, x);
      y = __builtin_fma(bcast(cl, 7 + 2 * t), rm, y);
   }
#pragma unroll
   for (int t = 0; t < 9; ++t)
   {
      y = __builtin_fma(bcast(cl, 18 + t), qs * Rs[(12 + t) * S + off], y);
   }
}

__device__ __forceinline__ UnitHdr load_hdr(const UnitHdr *h, int u)
{
   const int4 *p = reinterpret_cast<const int4 *>(h + u);
   const int4 a = p[0], b = p[1];
   return UnitHdr{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// coef and rc_enc are separate restrict arguments so that their
// wave-uniform loads may go through the scalar cache.
__global__ void __launch_bounds__(768) k_fused_rows(FusedParams F,
                                                     const double *__restrict__ coef,
                                                     const int *__restrict__ rc_enc)
{
   extern __shared__ double smem[];
   double *Rs = smem;
   const int nthr = blockDim.x;
   {
      const int per_t = F.nR * F.nd;
      for (int idx = threadIdx.x; idx < NTENS * per_t; idx += nthr)
      {
         const int t = idx / per_t;
         const int rem = idx - t * per_t;
         const int ar = rem / F.nd, b = rem - ar * F.nd;
         bool sw;
         const int off = slice_off(F, ar, b, sw);
         if (!sw) { Rs[(size_t)t * F.S + off] = ref_entry(F.R, F.npair, t, F.r0 + ar, b); }
      }
   }
   __syncthreads();
   const int nw = nthr >> 6;
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int nbuf = F.face ? 2 : 1;
   double2 *buf0 = reinterpret_cast<double2 *>(Rs + (size_t)NTENS * F.S) +
                   (size_t)wave * nbuf * F.rowcap;
   double2 *buf1 = buf0 + F.rowcap;
   const int nd = F.nd, p = F.p;
   const int foff = 6 * p, ntd = p * (p - 1);
   const unsigned char ess_bit = F.eliminate ? 2 : 0;
   const int b = lane;                         // nd <= 64 on this path
   const bool lane_live = b < nd;
   const int bb = lane_live ? b : 0;
   const int bface = (b >= foff && b < foff + 4 * ntd) ? (b - foff) / ntd : -1;
   const bool bodd = ((b - foff) & 1) != 0;
   const int stride = gridDim.x * nw;

   int u = blockIdx.x * nw + wave;
   if (u >= F.nunit) { return; }
   UnitHdr H = load_hdr(F.hdr, u);
   // rc_enc carries 64 trailing pad entries: the list load needs no guard
   int encL = rc_enc[H.c0 + lane];
   for (;;)
   {
      const int un = u + stride;
      const bool more = un < F.nunit;
      const UnitHdr Hn = load_hdr(F.hdr, more ? un : u);
      const int nc = H.nc_flags >> 8;
      const bool pair = H.i1 >= 0;
      const bool ess0 = F.eliminate && (H.nc_flags & 1);
      const bool ess1 = F.eliminate && (H.nc_flags & 2);
      for (int s = lane; s < H.len0; s += 64) { buf0[s] = make_double2(0.0, 0.0); }
      for (int s = lane; s < H.len1; s += 64) { buf1[s] = make_double2(0.0, 0.0); }
      wave_sync_lds();
      const bool work = !ess0 || (pair && !ess1);
      int encN = 0;
      bool have_next = false;
      for (int g0 = 0; work && g0 < nc; g0 += FG)
      {
         int enc[FG], sl0[FG], sl1[FG];
         unsigned char fb[FG];
         double cl[FG];
         double2 ap0[FG], ap1[FG];
#pragma unroll
         for (int k = 0; k < FG; ++k)
         {
            // a past-the-end slot of the group repeats the last contributor
            const int kk = min(g0 + k, nc - 1);
            enc[k] = __builtin_amdgcn_readlane(encL, kk);
            const int ea = enc[k] & CE_MASK;
            const int e = ea / nd;
            fb[k] = F.eflag[(size_t)e * nd + bb];
            sl0[k] = F.slot[(size_t)ea * F.sld + bb];
            sl1[k] = pair ? F.slot[(size_t)(ea + 1) * F.sld + bb] : 0;
            cl[k] = coef[(size_t)e * TEN_NC + min(lane, TEN_NC - 1)];
            ap0[k] = make_double2(0.0, 0.0);
            ap1[k] = make_double2(0.0, 0.0);
            if (enc[k] < 0)
            {
               const dbl2v x = __builtin_nontemporal_load(
                  reinterpret_cast<const dbl2v *>(F.Ae) + (size_t)ea * F.lda + bb);
               ap0[k] = make_double2(x.x, x.y);
               if (pair)
               {
                  const dbl2v y = __builtin_nontemporal_load(
                     reinterpret_cast<const dbl2v *>(F.Ae) + (size_t)(ea + 1) * F.lda + bb);
                  ap1[k] = make_double2(y.x, y.y);
               }
            }
         }
         if (!have_next)
         {
            encN = rc_enc[Hn.c0 + lane];
            have_next = true;
         }
#pragma unroll
         for (int k = 0; k < FG; ++k)
         {
            if (g0 + k >= nc) { break; }
            const int ea = enc[k] & CE_MASK;
            const int e = ea / nd, a = ea - e * nd;
            double2 v0 = ap0[k], v1 = ap1[k];
            if (enc[k] >= 0)
            {
               const int ar = a - F.r0;
               const int a_lo = F.face ? F.r0 + (ar & ~1) : a;
               double xr, xi, yr = 0.0, yi = 0.0;
               bool sw;
               int off = slice_off(F, a_lo - F.r0, bb, sw);
               tensor_row(Rs, F.S, off, sw ? -1.0 : 1.0, cl[k], xr, xi);
               if (F.face)
               {
                  off = slice_off(F, a_lo + 1 - F.r0, bb, sw);
                  tensor_row(Rs, F.S, off, sw ? -1.0 : 1.0, cl[k], yr, yi);
               }
               if (!lane_live) { xr = xi = yr = yi = 0.0; }
               const int fop = (int)bcast(cl[k], 27);
               // dual face transform: columns (lane pairs), then rows
               {
                  const double pxr = __shfl_xor(xr, 1), pxi = __shfl_xor(xi, 1);
                  const double pyr = __shfl_xor(yr, 1), pyi = __shfl_xor(yi, 1);
                  if (bface >= 0)
                  {
                     const double *Ti = F.TInv + 4 * ((fop >> (3 * bface)) & 7);
                     if (!bodd)
                     {
                        xr = pair_lo(Ti, xr, pxr);
                        xi = pair_lo(Ti, xi, pxi);
                        yr = pair_lo(Ti, yr, pyr);
                        yi = pair_lo(Ti, yi, pyi);
                     }
                     else
                     {
                        xr = pair_hi(Ti, pxr, xr);
                        xi = pair_hi(Ti, pxi, xi);
                        yr = pair_hi(Ti, pyr, yr);
                        yi = pair_hi(Ti, pyi, yi);
                     }
                  }
               }
               if (F.face)
               {
                  const int f = (a_lo - foff) / ntd;
                  const double *Ti = F.TInv + 4 * ((fop >> (3 * f)) & 7);
                  const double lr = pair_lo(Ti, xr, yr), li = pair_lo(Ti, xi, yi);
                  const double hr = pair_hi(Ti, xr, yr), hi = pair_hi(Ti, xi, yi);
                  if (pair || a == a_lo)
                  {
                     v0 = make_double2(lr, li);
                     v1 = make_double2(hr, hi);
                  }
                  else
                  {
                     v0 = make_double2(hr, hi);
                  }
               }
               else
               {
                  v0 = make_double2(xr, xi);
               }
            }
            const bool use = lane_live && (fb[k] & ess_bit) == 0;
            if (use && !ess0)
            {
               double2 t = buf0[sl0[k]];
               if (((enc[k] >> 30) ^ fb[k]) & 1) { t.x -= v0.x; t.y -= v0.y; }
               else { t.x += v0.x; t.y += v0.y; }
               buf0[sl0[k]] = t;
            }
            if (pair && use && !ess1)
            {
               double2 t = buf1[sl1[k]];
               if (((enc[k] >> 29) ^ fb[k]) & 1) { t.x -= v1.x; t.y -= v1.y; }
               else { t.x += v1.x; t.y += v1.y; }
               buf1[sl1[k]] = t;
            }
            wave_sync_lds();
         }
      }
      if (!have_next) { encN = rc_enc[Hn.c0 + lane]; }
      if (ess0)
      {
         for (int s = lane; s < H.len0; s += 64)
         {
            if (F.Jc[H.base0 + s] == H.i0) { buf0[s] = make_double2(1.0, 0.0); }
         }
      }
      if (ess1)
      {
         for (int s = lane; s < H.len1; s += 64)
         {
            if (F.Jc[H.base1 + s] == H.i1) { buf1[s] = make_double2(1.0, 0.0); }
         }
      }
      wave_sync_lds();
      for (int s = lane; s < H.len0; s += 64) { F.vals[H.base0 + s] = buf0[s]; }
      for (int s = lane; s < H.len1; s += 64) { F.vals[H.base1 + s] = buf1[s]; }
      wave_sync_lds();
      if (!more) { break; }
      u = un;
      H = Hn;
      encL = encN;
   }
}

}  // namespace lfasm
