// Quadrature kernel for nd <= 48 that uses the symmetry of the element
// matrix (solver/asm_gpu.inc:70-214 computes every entry).
//
// Ae = sum_q [v c]^T D_q [v c] over the mass rule plus the curl-curl sum.
// The curl-curl, mass and -K Lam^-1 K coefficients are symmetric; the
// MixedVectorCurl / MixedVectorWeakCurl pair couples v and c with
// antisymmetric coefficients (CROSS1 = -CROSS0^T, the C3..C5 rows of
// mass_coef are minus the transpose of the c-terms of C0..C2).  So
//     Ae = S + P - P^T,
// S = sum_cc c^T (w Lam^-1) c + sum_m v^T Dvv v   (symmetric),
// P = sum_m v^T X c                                (the CROSS0 coupling),
// and only the tiles of S on and above the diagonal are formed.  With three
// 16-wide tiles (nd <= 48) a wavefront owns the tile pair {(ti, tj), (tj, ti)}
// and one diagonal tile td:
//     Ae(ti, tj) = S(ti, tj) + P(ti, tj) - P(tj, ti)^T
//     Ae(tj, ti) = (S(ti, tj) + P(tj, ti)^T - P(ti, tj))^T
//     Ae(td, td) = S(td, td) + P(td, td) - P(td, td)^T
// P(tj, ti)^T is accumulated directly in the (ti, tj) orientation (the MFMA
// operands swapped), so the pair needs one transpose.
// The K stream stages per mass point only the value rows v (the curl rows
// enter through X c), and per row the symmetric coefficient row CS.  X has
// rank 2: its first two rows are i wb (c0y, -c0x) times e_2^T (k_par x
// couples only the z component of the curl into the transverse values), so
//     v^T X c = (c0y v0 - c0x v1) (i wb c2) + v2 (X c)_2,
// and P takes two staged rows per mass point (AP = (u, v2), CP = the two
// coefficient rows) instead of three.  180 MFMAs per wavefront at p = 3
// with k_par (cc rule 11 points, mass rule 24) instead of 270 for the full
// matrix.  Masks holding only one of CROSS0 / CROSS1 are not antisymmetric
// and go to the general kernels.
#pragma once

#include "lfasm_device.hpp"

namespace lfasm {

constexpr int SYM_NP = 48;                        // 3 tiles of 16 columns
constexpr int SYM_ROWS = 12;                      // 4 points x 3 rows per chunk
constexpr int SYM_PROWS = 8;                      // 4 points x 2 rows of the P operands
// A, CSr, CSi (SYM_ROWS rows each), AP, CPr, CPi (SYM_PROWS rows each)
constexpr int SYM_STG = (3 * SYM_ROWS + 3 * SYM_PROWS) * SYM_NP;
constexpr int SYM_FCS = 2 * SYM_NP;               // face-transform pair per local index
constexpr int SYM_THREADS = 192;                  // 3 wavefronts per element

struct SymShape {
   int nq_cc, nq_m, nch_cc, nch_m;
   int cross;   // CROSS0 and CROSS1 both on
};

__device__ __forceinline__ SymShape sym_shape(const LocalParams &P)
{
   SymShape s;
   s.nq_cc = P.use_cc ? P.rc.nq : 0;
   s.nq_m = (P.mflags & 15) ? P.rm.nq : 0;
   s.nch_cc = (s.nq_cc + 3) / 4;
   s.nch_m = (s.nq_m + 3) / 4;
   s.cross = (P.mflags & 6) == 6;
   return s;
}

// One (point, basis k) item of a chunk: thread t forms point t / 48, basis
// t % 48 (192 threads, 4 points x 48 columns).
struct SymItem {
   double h[6];
   int q, k, row0, prow0;
   bool valid, live, mass;
};

__device__ __forceinline__ SymItem sym_load_item(const LocalParams &P, const SymShape &ss, int ch,
                                                 int tid)
{
   SymItem it;
#pragma unroll
   for (int i = 0; i < 6; ++i) { it.h[i] = 0.0; }
   it.mass = ch >= ss.nch_cc;
   const int c = it.mass ? ch - ss.nch_cc : ch;
   const int nq = it.mass ? ss.nq_m : ss.nq_cc;
   const int qi = tid / SYM_NP;
   it.k = tid - qi * SYM_NP;
   const int q = c * 4 + qi;
   it.valid = q < nq;
   it.q = it.valid ? q : 0;
   it.row0 = qi * 3;
   it.prow0 = qi * 2;
   it.live = it.valid && it.k < P.nd;
   if (it.live)
   {
      const RuleDev &r = it.mass ? P.rm : P.rc;
      const double *t0 = (it.mass ? r.vs : r.cs) + ((size_t)it.q * P.nd + it.k) * 3;
      it.h[0] = t0[0];
      it.h[1] = t0[1];
      it.h[2] = t0[2];
      if (it.mass && ss.cross)
      {
         const double *t1 = r.cs + ((size_t)it.q * P.nd + it.k) * 3;
         it.h[3] = t1[0];
         it.h[4] = t1[1];
         it.h[5] = t1[2];
      }
   }
   return it;
}

// Staging rows of one item: A (shape rows), CS (symmetric coefficients),
// AP / CP (the two rank-one pairs of the cross coupling, mass rows with the
// cross terms only).
__device__ __forceinline__ void sym_form_item(const LocalParams &P, const double *ec,
                                              const double *pj, const SymShape &ss,
                                              const double *qcoef, const SymItem &it, double *stg)
{
   constexpr int RN = SYM_ROWS * SYM_NP, RP = SYM_PROWS * SYM_NP;
   double *A = stg, *CSr = stg + RN, *CSi = stg + 2 * RN;
   double *AP = stg + 3 * RN, *CPr = AP + RP, *CPi = AP + 2 * RP;
   const int k = it.k, r0 = it.row0, p0 = it.prow0;
   if (!it.mass)
   {
      const double *qc = qcoef + (size_t)it.q * 4;
      double c0 = 0.0, c1 = 0.0, c2 = 0.0;
      if (it.live) { piola_c(pj, it.h[0], it.h[1], it.h[2], c0, c1, c2); }
      const double war = it.valid ? qc[0] : 0.0, wai = it.valid ? qc[1] : 0.0;
      const double wbr = it.valid ? qc[2] : 0.0, wbi = it.valid ? qc[3] : 0.0;
      A[(r0 + 0) * SYM_NP + k] = c0;
      A[(r0 + 1) * SYM_NP + k] = c1;
      A[(r0 + 2) * SYM_NP + k] = c2;
      CSr[(r0 + 0) * SYM_NP + k] = war * c0;
      CSi[(r0 + 0) * SYM_NP + k] = wai * c0;
      CSr[(r0 + 1) * SYM_NP + k] = war * c1;
      CSi[(r0 + 1) * SYM_NP + k] = wai * c1;
      CSr[(r0 + 2) * SYM_NP + k] = wbr * c2;
      CSi[(r0 + 2) * SYM_NP + k] = wbi * c2;
      return;
   }
   const double *qc = qcoef + (size_t)(P.rc.nq + it.q) * 4;
   double v0 = 0.0, v1 = 0.0, v2 = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
   if (it.live)
   {
      piola_v(pj, it.h[0], it.h[1], it.h[2], v0, v1, v2);
      if (ss.cross) { piola_c(pj, it.h[3], it.h[4], it.h[5], c0, c1, c2); }
   }
   const double2 wa = it.valid ? make_double2(qc[0], qc[1]) : make_double2(0.0, 0.0);
   const double2 wb = it.valid ? make_double2(qc[2], qc[3]) : make_double2(0.0, 0.0);
   const double *mc = ec + 18;   // m0r m0i m1r m1i m2r m2i kxy c0x c0y c1x c1y
   const double2 S0 = cmul(wb, make_double2(mc[0] * v0 - mc[6] * v1, mc[1] * v0));
   const double2 S1 = cmul(wb, make_double2(mc[2] * v1 - mc[6] * v0, mc[3] * v1));
   const double2 S2 = cmul(wa, make_double2(mc[4] * v2, mc[5] * v2));
   A[(r0 + 0) * SYM_NP + k] = v0;
   A[(r0 + 1) * SYM_NP + k] = v1;
   A[(r0 + 2) * SYM_NP + k] = v2;
   CSr[(r0 + 0) * SYM_NP + k] = S0.x;
   CSi[(r0 + 0) * SYM_NP + k] = S0.y;
   CSr[(r0 + 1) * SYM_NP + k] = S1.x;
   CSi[(r0 + 1) * SYM_NP + k] = S1.y;
   CSr[(r0 + 2) * SYM_NP + k] = S2.x;
   CSi[(r0 + 2) * SYM_NP + k] = S2.y;
   if (ss.cross)
   {
      // the c-terms of C0..C2 of mass_coef, X c = (i wb c0y c2, -i wb c0x c2,
      // i wa (c0x c1 - c0y c0)), as two rank-one pairs:
      // v0 (X c)_0 + v1 (X c)_1 = u (i wb c2) with u = c0y v0 - c0x v1
      const double u = mc[8] * v0 - mc[7] * v1;
      const double2 Y = make_double2(-wb.y * c2, wb.x * c2);
      const double2 X2 = cmul(wa, make_double2(0.0, -mc[8] * c0 + mc[7] * c1));
      AP[(p0 + 0) * SYM_NP + k] = u;
      AP[(p0 + 1) * SYM_NP + k] = v2;
      CPr[(p0 + 0) * SYM_NP + k] = Y.x;
      CPi[(p0 + 0) * SYM_NP + k] = Y.y;
      CPr[(p0 + 1) * SYM_NP + k] = X2.x;
      CPi[(p0 + 1) * SYM_NP + k] = X2.y;
   }
}

// Transpose of a 16 x 16 MFMA output tile (lane l = m + 16 g holds rows
// 4v + g of column m in register v) through a per-wavefront LDS scratch of
// 16 x 17 doubles.
__device__ __forceinline__ dbl4 sym_transpose(const dbl4 x, double *tr, int g4, int m16)
{
#pragma unroll
   for (int v = 0; v < 4; ++v) { tr[(4 * v + g4) * 17 + m16] = x[v]; }
   wave_sync_lds();
   dbl4 y;
#pragma unroll
   for (int v = 0; v < 4; ++v) { y[v] = tr[m16 * 17 + 4 * v + g4]; }
   wave_sync_lds();
   return y;
}

// Face-transform coefficients of element e for every local index d < 48:
// fcs[2d] (own weight), fcs[2d + 1] (partner weight); (1, 0) off the faces.
// Threads tid < SYM_NP of the workgroup.
__device__ __forceinline__ void sym_face_coef(const LocalParams &P, int e, int tid, double *fcs)
{
   if (tid >= SYM_NP) { return; }
   const int p = P.p, foff = 6 * p, ntd = p * (p - 1);
   double cs = 1.0, cp = 0.0;
   if (P.apply_te && p >= 2 && tid >= foff && tid < foff + 4 * ntd && tid < P.nd)
   {
      const double *Ti = P.TInv + 4 * P.fo[(size_t)e * 4 + (tid - foff) / ntd];
      const bool even = ((tid - foff) & 1) == 0;
      cs = even ? Ti[0] : Ti[3];
      cp = even ? Ti[1] : Ti[2];
   }
   fcs[2 * tid] = cs;
   fcs[2 * tid + 1] = cp;
}

// Per-element setup of the symmetric kernel, formed ahead of it by
// k_sym_prep: [element constants (EC_SIZE) | per-point factors (4 nqt) |
// face-transform coefficients (SYM_FCS)] doubles, the kernel's LDS layout.
// The record of list entry io is written into the first doubles of that
// entry's own element-matrix slot in Ae, which the quadrature kernel reads
// (and stages in LDS) before it writes the entry's matrix; the slot holds
// nd * lda complex entries, far more than the record.  This takes the
// dependent geometry/material loads, the PML profile (pow) and the
// divisions off the MFMA kernel's critical path.
__host__ __device__ constexpr int sym_rec_size(int nqt) { return EC_SIZE + 4 * nqt + SYM_FCS; }

__device__ __forceinline__ double *sym_rec(const LocalParams &P, int io, int e)
{
   return elem_rec(P, io, e);
}

// One wavefront per list entry (nqt points over the 64 lanes); also resets
// the work counter of the persistent quadrature kernel (when given).
__global__ void __launch_bounds__(64) k_sym_prep(LocalParams P, int nqt, int *queue)
{
   const int io = blockIdx.x;
   if (queue && io == 0 && threadIdx.x == 0) { *queue = 0; }
   const int e = P.elist ? P.elist[io] : io;
   const int tid = threadIdx.x;
   double *rec = sym_rec(P, io, e);
   ElemGeo g;
   load_geo(P, e, g);
   for (int q = tid; q < nqt; q += 64) { q_coef(P, g, q, rec + EC_SIZE + 4 * q); }
   if (tid == 0) { store_elem_const(g, mass_coef(P, g.pm, g.eps), rec); }
   sym_face_coef(P, e, tid, rec + EC_SIZE + 4 * nqt);
}

// Dual face transform (columns, then rows; solver/asm_gpu.inc:221-251) with
// the element's coefficients fcs, and store of output tile (I, J) of
// element e.
__device__ __forceinline__ void sym_tile_out(const LocalParams &P, const double *fcs, dbl4 (&t)[2],
                                             int e, int io, int I, int J, int g4, int m16)
{
   const int nd = P.nd, p = P.p;
   const int foff = 6 * p, fend = 6 * p + 4 * p * (p - 1);
   const bool te = P.apply_te && p >= 2;
   const bool face_cols = te && 16 * J + 16 > foff && 16 * J < fend;
   const bool face_rows = te && 16 * I + 16 > foff && 16 * I < fend;
   {
      if (face_cols)
      {
         const int col = 16 * J + m16;
         const double cs = fcs[2 * col], cp = fcs[2 * col + 1];
#pragma unroll
         for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
               const double x = t[c][v];
               const double y = lane_xor1(x);
               t[c][v] = cs * x + cp * y;
            }
      }
      if (face_rows)
      {
#pragma unroll
         for (int v = 0; v < 4; ++v)
         {
            const int row = 16 * I + g4 + 4 * v;
            const double cs = fcs[2 * row], cp = fcs[2 * row + 1];
#pragma unroll
            for (int c = 0; c < 2; ++c)
            {
               const double x = t[c][v];
               const double y = lane_xor16(x);
               t[c][v] = cs * x + cp * y;
            }
         }
      }
   }
   double2 *Ae = P.Ae + (size_t)(P.out_by_elem ? e : io) * nd * P.lda;
   const bool lines = (P.lda & 7) == 0;
   const int col = 16 * J + m16;
#pragma unroll
   for (int v = 0; v < 4; ++v)
   {
      const int row = 16 * I + g4 + 4 * v;
      if (row < nd && col < P.lda)
      {
         if (lines)
         {
            __builtin_nontemporal_store(dbl2v{t[0][v], t[1][v]},
                                        reinterpret_cast<dbl2v *>(Ae) + (size_t)row * P.lda + col);
         }
         else if (col < nd)
         {
            Ae[(size_t)row * P.lda + col] = make_double2(t[0][v], t[1][v]);
         }
      }
   }
}

// One element per workgroup of 3 wavefronts at a time.  NBUF = 2:
// double-buffered staging as in k_local_mfma_pipe (the table loads of chunk
// c+1 are issued before the MFMAs of chunk c, its rows formed after them);
// NBUF = 1: one staging buffer (half the LDS, more workgroups per CU), the
// next chunk formed between two barriers.
// PRIO = 1: the wavefronts raise their issue priority while they run the
// MFMAs of a chunk and drop it for the staging, so that on a SIMD shared by
// several elements the matrix pipe is fed first; PRIO = 2: also raised for
// the epilogue (its FP64 face-transform FMAs share the SIMD's FP64 issue
// with the other wavefronts' MFMAs).
// PERSIST: a grid of a few workgroups per CU; each takes list entries from
// `queue` (reset to 0 by k_sym_prep) and loads the next entry's setup record
// into registers while it computes the current one, so neither the record
// load nor a workgroup launch sits between two elements.
constexpr int SYM_PF = 2;   // prefetch registers per thread (record <= SYM_PF * SYM_THREADS)

template <int NBUF, int WAVES_PER_EU, int PRIO = 0, bool PERSIST = false>
__global__ void __launch_bounds__(SYM_THREADS)
__attribute__((amdgpu_waves_per_eu(WAVES_PER_EU)))
k_local_mfma_sym(LocalParams P, int nqt, int *queue)
{
   constexpr int RN = SYM_ROWS * SYM_NP, RP = SYM_PROWS * SYM_NP;
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   __shared__ int s_next;

   extern __shared__ double smem[];
   double *ec = smem;
   double *qcoef = ec + EC_SIZE;
   double *fcs = qcoef + nqt * 4;
   double *buf = fcs + SYM_FCS;
   const int rs = sym_rec_size(nqt);
   int io = blockIdx.x;   // grid <= nout
   int e = P.elist ? P.elist[io] : io;
   // the element's setup record (k_sym_prep), read before any output of
   // this element is written over it
   {
      const double *rec = sym_rec(P, io, e);
      for (int i = threadIdx.x; i < rs; i += SYM_THREADS) { ec[i] = rec[i]; }
   }
   const SymShape ss = sym_shape(P);
   // the element's Piola constants (J / detJ, J^-1) in scalar registers
   double pj[18];
   {
      const double *rec = sym_rec(P, io, e);
#pragma unroll
      for (int i = 0; i < 18; ++i) { pj[i] = rec[i]; }
   }
   __syncthreads();

   // tiles of this wavefront: the pair (ti, tj) and the diagonal td
   const int ti = wave == 2 ? 1 : 0;
   const int tj = wave == 0 ? 1 : 2;
   const int td = 2 - wave;
   const int nch = ss.nch_cc + ss.nch_m;
   for (;;)
   {
      // the thread index made opaque to the loop over elements: without it
      // the compiler keeps every per-thread address of the element's work
      // live across the loop (spills)
      int tid_v = threadIdx.x;
      if (PERSIST) { asm volatile("" : "+v"(tid_v)); }
      const int tid = tid_v;
      const int lane = tid & 63;
      const int g4 = lane >> 4, m16 = lane & 15;
      if (PERSIST && tid == 0) { s_next = (int)gridDim.x + atomicAdd(queue, 1); }
      dbl4 sij[2], pij[2], pjiT[2], sdd[2], pdd[2];
#pragma unroll
      for (int c = 0; c < 2; ++c)
      {
         sij[c] = dbl4{0.0, 0.0, 0.0, 0.0};
         pij[c] = dbl4{0.0, 0.0, 0.0, 0.0};
         pjiT[c] = dbl4{0.0, 0.0, 0.0, 0.0};
         sdd[c] = dbl4{0.0, 0.0, 0.0, 0.0};
         pdd[c] = dbl4{0.0, 0.0, 0.0, 0.0};
      }
      if (nch > 0) { sym_form_item(P, ec, pj, ss, qcoef, sym_load_item(P, ss, 0, tid), buf); }
      __syncthreads();
      // next list entry (read after the barrier that follows its fetch)
      const int io_n = PERSIST ? __builtin_amdgcn_readfirstlane(s_next) : P.nout;
      const bool has_next = io_n < P.nout;
      double pf[SYM_PF];
#pragma unroll
      for (int k = 0; k < SYM_PF; ++k) { pf[k] = 0.0; }
      if (PERSIST && has_next)
      {
         const double *rn = sym_rec(P, io_n, P.elist ? P.elist[io_n] : io_n);
#pragma unroll
         for (int k = 0; k < SYM_PF; ++k)
         {
            if (tid + k * SYM_THREADS < rs) { pf[k] = rn[tid + k * SYM_THREADS]; }
         }
      }
      for (int ch = 0; ch < nch; ++ch)
      {
         const double *A = buf + (NBUF == 2 ? (ch & 1) * SYM_STG : 0);
         const double *CSr = A + RN, *CSi = A + 2 * RN;
         const double *AP = A + 3 * RN, *CPr = AP + RP, *CPi = AP + 2 * RP;
         const bool more = ch + 1 < nch;
         const SymItem nxt = sym_load_item(P, ss, more ? ch + 1 : ch, tid);
         const bool cross = ss.cross && ch >= ss.nch_cc;   // wave-uniform
         if (PRIO) { __builtin_amdgcn_s_setprio(2); }
#pragma unroll
         for (int ks = 0; ks < SYM_ROWS / 4; ++ks)
         {
            const int o = (ks * 4 + g4) * SYM_NP + m16;
            const double ai = A[o + 16 * ti], ad = A[o + 16 * td];
            const double sjr = CSr[o + 16 * tj], sji = CSi[o + 16 * tj];
            const double sdr = CSr[o + 16 * td], sdi = CSi[o + 16 * td];
            sij[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, sjr, sij[0], 0, 0, 0);
            sij[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, sji, sij[1], 0, 0, 0);
            sdd[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, sdr, sdd[0], 0, 0, 0);
            sdd[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, sdi, sdd[1], 0, 0, 0);
         }
         if (cross)
         {
#pragma unroll
            for (int ks = 0; ks < SYM_PROWS / 4; ++ks)
            {
               const int o = (ks * 4 + g4) * SYM_NP + m16;
               const double ai = AP[o + 16 * ti], aj = AP[o + 16 * tj], ad = AP[o + 16 * td];
               const double pjr = CPr[o + 16 * tj], pjq = CPi[o + 16 * tj];
               const double pir = CPr[o + 16 * ti], piq = CPi[o + 16 * ti];
               const double pdr = CPr[o + 16 * td], pdq = CPi[o + 16 * td];
               pij[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, pjr, pij[0], 0, 0, 0);
               pij[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, pjq, pij[1], 0, 0, 0);
               // P(tj, ti)^T in the (ti, tj) orientation: operands swapped
               pjiT[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(pir, aj, pjiT[0], 0, 0, 0);
               pjiT[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(piq, aj, pjiT[1], 0, 0, 0);
               pdd[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, pdr, pdd[0], 0, 0, 0);
               pdd[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, pdq, pdd[1], 0, 0, 0);
            }
         }
         if (PRIO) { __builtin_amdgcn_s_setprio(0); }
         if (NBUF == 1) { __syncthreads(); }
         if (more)
         {
            sym_form_item(P, ec, pj, ss, qcoef, nxt, buf + (NBUF == 2 ? ((ch + 1) & 1) * SYM_STG : 0));
         }
         __syncthreads();
      }

      // the staging is free now: per-wavefront transpose scratch
      if (PRIO == 2) { __builtin_amdgcn_s_setprio(2); }
      double *tr = buf + wave * 16 * 17;
      // Ae(ti, tj) = S + P(ti, tj) - P(tj, ti)^T, and
      // Ae(tj, ti) = (S + P(tj, ti)^T - P(ti, tj))^T: one transpose per pair
      dbl4 t1[2], t2[2], t3[2];
#pragma unroll
      for (int c = 0; c < 2; ++c)
      {
         t1[c] = sij[c] + pij[c] - pjiT[c];
         t2[c] = sym_transpose(sij[c] + pjiT[c] - pij[c], tr, g4, m16);
         const dbl4 pddT = sym_transpose(pdd[c], tr, g4, m16);
         t3[c] = sdd[c] + pdd[c] - pddT;
      }
      sym_tile_out(P, fcs, t1, e, io, ti, tj, g4, m16);
      sym_tile_out(P, fcs, t2, e, io, tj, ti, g4, m16);
      sym_tile_out(P, fcs, t3, e, io, td, td, g4, m16);
      if (PRIO == 2) { __builtin_amdgcn_s_setprio(0); }
      if (!PERSIST || !has_next) { break; }
      // every wavefront is done with this element's LDS (face coefficients,
      // transpose scratch): the prefetched record replaces it
      __syncthreads();
#pragma unroll
      for (int k = 0; k < SYM_PF; ++k)
      {
         if (tid + k * SYM_THREADS < rs) { ec[tid + k * SYM_THREADS] = pf[k]; }
      }
      __syncthreads();
      io = io_n;
      e = P.elist ? P.elist[io] : io;
   }
}

}  // namespace lfasm
