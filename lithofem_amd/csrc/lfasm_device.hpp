// Device code of the gfx950 Nedelec assembly: element matrices (MFMA FP64
// for 16-wide tiles, VALU for the generic path), dual face-DOF transform,
// deterministic row-segmented scatter with fused essential elimination,
// CSR slot map, SpMV.
//
// Mathematical content follows the reference assembly
// (solver/asm_gpu.inc:70-214 local matrices, :221-251 face transform,
// :434-478 scatter/elimination, :481-493 SpMV; coefficient definitions in
// solver/problem_common.inc:69-98,150-216).  The element matrix is computed
// as the quadrature contraction Ae = B^T C, where B stacks the physical
// shape components of every basis function at every quadrature point
// (K rows) and C = (w D) B applies the per-point complex coefficient
// matrix D of the five integrators; the K stream is processed in chunks of
// 12 rows staged in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lfasm {

constexpr int BLOCK = 256;        // 4 wavefronts of 64 lanes
constexpr int CHUNK_ROWS = 12;    // K rows staged per chunk (3 MFMA k-steps)

struct RuleDev {
   int nq;
   const double *ip, *w, *vs, *cs;
};

struct PmlDev {
   double z_min, z_max, thick, order, smax_bot, smax_top, k0, kx, ky;
};

struct LocalParams {
   int ne, nd, p;
   const double *J, *Ji, *detJ, *v0;
   const int *attr;
   const double2 *eps_tab;
   const double2 *eps_elem;
   const int *fo;
   RuleDev rc, rm;
   const PmlDev *pml_tab;   // per problem (one entry for a single mesh)
   const int *pid;          // element -> problem id (NULL: problem 0)
   const int *eps_off;      // problem -> offset of its region table in eps_tab
   double TInv[24];
   int use_cc;       // curl-curl stream on the cc rule
   int mflags;       // bit0 mass, bit1 cross0, bit2 cross1, bit3 klk
   int apply_te;
   double2 *Ae;
   int nout;             // number of element matrices written (launch size)
   const int *elist;     // output i -> element elist[i] (NULL: element i)
   int eps_unit;         // 1: eps = 1 in the mass term (unit-eps mass matrices)
   int out_by_elem;      // 1: write element elist[i]'s matrix to Ae[elist[i]], not Ae[i]
   int lda;              // row stride of the element matrices in Ae (>= nd)
};

// ---- complex helpers ------------------------------------------------------

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
   return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// 1/s as conj(s)/|s|^2 (the reference device path, solver/asm_gpu.inc:23-27)
__device__ __forceinline__ double2 cinv(double2 a)
{
   const double d = a.x * a.x + a.y * a.y;
   return make_double2(a.x / d, -a.y / d);
}

// PML stretch s(z) (solver/asm_gpu.inc:40-48)
__device__ __forceinline__ double2 pml_s(const PmlDev &p, double z)
{
   double d = -1.0, smax = 0.0;
   if (z < p.z_min) { d = (p.z_min - z) / p.thick; smax = p.smax_bot; }
   else if (z > p.z_max) { d = (z - p.z_max) / p.thick; smax = p.smax_top; }
   if (d < 0.0) { return make_double2(1.0, 0.0); }
   if (d > 1.0) { d = 1.0; }
   return make_double2(1.0, smax * pow(d, p.order));
}

// ---- per-element coefficient setup ----------------------------------------
// Element-uniform data of one element, loaded by every lane of the group.
struct ElemGeo {
   double Jd[9];    // J / detJ (contravariant Piola)
   double Ji[9];    // row-major J^{-1}
   double J6, J7, J8, v0z, detJ;
   double2 eps;
   PmlDev pm;       // PML / Bloch parameters of the element's problem
};

__device__ __forceinline__ void load_geo(const LocalParams &P, int e, ElemGeo &g)
{
   const double det = P.detJ[e];
   const double inv = 1.0 / det;
#pragma unroll
   for (int i = 0; i < 9; ++i)
   {
      g.Jd[i] = P.J[(size_t)e * 9 + i] * inv;
      g.Ji[i] = P.Ji[(size_t)e * 9 + i];
   }
   g.J6 = P.J[(size_t)e * 9 + 6];
   g.J7 = P.J[(size_t)e * 9 + 7];
   g.J8 = P.J[(size_t)e * 9 + 8];
   g.v0z = P.v0[(size_t)e * 3 + 2];
   g.detJ = det;
   const int pid = P.pid ? P.pid[e] : 0;
   g.pm = P.pml_tab[pid];
   g.eps = P.eps_unit ? make_double2(1.0, 0.0)
           : P.eps_elem ? P.eps_elem[e] : P.eps_tab[P.eps_off[pid] + P.attr[e] - 1];
}

// Per quadrature point q (cc rule first, then mass rule): w*Lam^-1 diagonal
// factors wa = w/s and wb = w*s, with w = weight*detJ.
__device__ __forceinline__ void q_coef(const LocalParams &P, const ElemGeo &g, int q,
                                       double *qc)
{
   const RuleDev &R = (q < P.rc.nq) ? P.rc : P.rm;
   const int qq = (q < P.rc.nq) ? q : q - P.rc.nq;
   const double *xi = R.ip + (size_t)qq * 3;
   const double z = g.v0z + g.J6 * xi[0] + g.J7 * xi[1] + g.J8 * xi[2];
   const double2 s = pml_s(g.pm, z);
   const double2 si = cinv(s);
   const double w = R.w[qq] * g.detJ;
   qc[0] = w * si.x;
   qc[1] = w * si.y;
   qc[2] = w * s.x;
   qc[3] = w * s.y;
}

// Row counts of the two K streams.
struct StreamShape {
   int nq_cc, nq_m;      // quadrature points used per stream (0 if off)
   int rows_m;           // rows per mass-rule point: 3 (v only) or 6 (v and curl)
   int qpc_cc, qpc_m;    // points per 12-row chunk
   int nch_cc, nch_m;    // chunks per stream
};

__host__ __device__ inline StreamShape stream_shape(int use_cc, int mflags, int nq_cc,
                                                    int nq_m)
{
   StreamShape s;
   s.nq_cc = use_cc ? nq_cc : 0;
   s.nq_m = (mflags & 15) ? nq_m : 0;
   s.rows_m = (mflags & 6) ? 6 : 3;
   s.qpc_cc = 4;
   s.qpc_m = CHUNK_ROWS / s.rows_m;
   s.nch_cc = (s.nq_cc + s.qpc_cc - 1) / s.qpc_cc;
   s.nch_m = (s.nq_m + s.qpc_m - 1) / s.qpc_m;
   return s;
}

// Mass-rule row coefficients (element constants).  With flags fm (mass),
// f0 (cross0), f1 (cross1), fk (klk) the C rows of one point are
//   C0 = wb ((fm me + fk ky^2) v0 - fk kx ky v1 + i f0 ky c2)
//   C1 = wb ((fm me + fk kx^2) v1 - fk kx ky v0 - i f0 kx c2)
//   C2 = wa ((fm me + fk (kx^2+ky^2)) v2 - i f0 ky c0 + i f0 kx c1)
//   C3 = wa (i f1 ky v2)
//   C4 = wa (-i f1 kx v2)
//   C5 = wb (-i f1 ky v0 + i f1 kx v1)
// with me = -k0^2 eps, wa = w/s, wb = w s: the mass, MixedVectorCurl,
// MixedVectorWeakCurl and -K Lam^-1 K terms of solver/asm_gpu.inc:140-203
// (K = [[0,0,ky],[0,0,-kx],[-ky,kx,0]], Lam^-1 = diag(1/s, 1/s, s)).
struct MassCoef {
   double m0r, m0i, m1r, m1i, m2r, m2i, kxy, c0x, c0y, c1x, c1y;
};

__device__ __forceinline__ MassCoef mass_coef(const LocalParams &P, const PmlDev &pm,
                                              double2 eps)
{
   const double fm = (P.mflags & 1) ? 1.0 : 0.0;
   const double f0 = (P.mflags & 2) ? 1.0 : 0.0;
   const double f1 = (P.mflags & 4) ? 1.0 : 0.0;
   const double fk = (P.mflags & 8) ? 1.0 : 0.0;
   const double kx = pm.kx, ky = pm.ky;
   const double k2 = pm.k0 * pm.k0;
   const double mer = -k2 * eps.x * fm, mei = -k2 * eps.y * fm;
   MassCoef c;
   c.m0r = mer + fk * ky * ky;
   c.m0i = mei;
   c.m1r = mer + fk * kx * kx;
   c.m1i = mei;
   c.m2r = mer + fk * (kx * kx + ky * ky);
   c.m2i = mei;
   c.kxy = fk * kx * ky;
   c.c0x = f0 * kx;
   c.c0y = f0 * ky;
   c.c1x = f1 * kx;
   c.c1y = f1 * ky;
   return c;
}

// Element constants kept in LDS during the K stream (EC_SIZE doubles):
// [0..8] J/detJ, [9..17] J^{-1}, [18..28] MassCoef.
constexpr int EC_SIZE = 32;

__device__ __forceinline__ void store_elem_const(const ElemGeo &g, const MassCoef &mc,
                                                 double *ec)
{
#pragma unroll
   for (int i = 0; i < 9; ++i)
   {
      ec[i] = g.Jd[i];
      ec[9 + i] = g.Ji[i];
   }
   ec[18] = mc.m0r; ec[19] = mc.m0i; ec[20] = mc.m1r; ec[21] = mc.m1i;
   ec[22] = mc.m2r; ec[23] = mc.m2i; ec[24] = mc.kxy; ec[25] = mc.c0x;
   ec[26] = mc.c0y; ec[27] = mc.c1x; ec[28] = mc.c1y;
}

// Element setup for one element group: per-point factors into qcoef (nqt x 4)
// and the element constants into ec.  Threads tid < nthr of the group.
__device__ __forceinline__ void elem_setup(const LocalParams &P, int e, int nqt, int tid,
                                           int nthr, double *qcoef, double *ec)
{
   ElemGeo g;
   load_geo(P, e, g);
   for (int q = tid; q < nqt; q += nthr) { q_coef(P, g, q, qcoef + q * 4); }
   if (tid == 0) { store_elem_const(g, mass_coef(P, g.pm, g.eps), ec); }
}

// Contravariant Piola c = (J/detJ) chat, covariant v = J^{-T} vhat.
__device__ __forceinline__ void piola_c(const double *ec, double h0, double h1, double h2,
                                        double &c0, double &c1, double &c2)
{
   c0 = ec[0] * h0 + ec[1] * h1 + ec[2] * h2;
   c1 = ec[3] * h0 + ec[4] * h1 + ec[5] * h2;
   c2 = ec[6] * h0 + ec[7] * h1 + ec[8] * h2;
}

__device__ __forceinline__ void piola_v(const double *ec, double a0, double a1, double a2,
                                        double &v0, double &v1, double &v2)
{
   v0 = a0 * ec[9] + a1 * ec[12] + a2 * ec[15];
   v1 = a0 * ec[10] + a1 * ec[13] + a2 * ec[16];
   v2 = a0 * ec[11] + a1 * ec[14] + a2 * ec[17];
}

// Form one (point, basis k) item of a chunk into the LDS staging rows
// A[row][k], Cr[row][k], Ci[row][k] (row stride NP).  qc = per-point
// factors (wa.re, wa.im, wb.re, wb.im).  valid == false writes zero rows.
__device__ __forceinline__ void form_cc_raw(const double *pj, const double *qc, bool valid,
                                            bool live, double h0, double h1, double h2,
                                            double *A, double *Cr, double *Ci, int row0, int k,
                                            int NP)
{
   double c0 = 0.0, c1 = 0.0, c2 = 0.0;
   if (live) { piola_c(pj, h0, h1, h2, c0, c1, c2); }
   const double war = valid ? qc[0] : 0.0, wai = valid ? qc[1] : 0.0;
   const double wbr = valid ? qc[2] : 0.0, wbi = valid ? qc[3] : 0.0;
   A[(row0 + 0) * NP + k] = c0;
   A[(row0 + 1) * NP + k] = c1;
   A[(row0 + 2) * NP + k] = c2;
   Cr[(row0 + 0) * NP + k] = war * c0;
   Ci[(row0 + 0) * NP + k] = wai * c0;
   Cr[(row0 + 1) * NP + k] = war * c1;
   Ci[(row0 + 1) * NP + k] = wai * c1;
   Cr[(row0 + 2) * NP + k] = wbr * c2;
   Ci[(row0 + 2) * NP + k] = wbi * c2;
}

__device__ __forceinline__ void form_cc_item(const LocalParams &P, const double *pj,
                                             const double *qc, int q, int k, bool valid,
                                             double *A, double *Cr, double *Ci, int row0,
                                             int NP)
{
   const bool live = valid && k < P.nd;
   double h0 = 0.0, h1 = 0.0, h2 = 0.0;
   if (live)
   {
      const double *cs = P.rc.cs + ((size_t)q * P.nd + k) * 3;
      h0 = cs[0];
      h1 = cs[1];
      h2 = cs[2];
   }
   form_cc_raw(pj, qc, valid, live, h0, h1, h2, A, Cr, Ci, row0, k, NP);
}

__device__ __forceinline__ void form_mass_raw(const double *pj, const double *ec,
                                              const double *qc, bool valid, bool live, int rows,
                                              const double *hv, const double *hc, double *A,
                                              double *Cr, double *Ci, int row0, int k, int NP)
{
   double v0 = 0.0, v1 = 0.0, v2 = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
   if (live)
   {
      piola_v(pj, hv[0], hv[1], hv[2], v0, v1, v2);
      if (rows == 6) { piola_c(pj, hc[0], hc[1], hc[2], c0, c1, c2); }
   }
   const double2 wa = valid ? make_double2(qc[0], qc[1]) : make_double2(0.0, 0.0);
   const double2 wb = valid ? make_double2(qc[2], qc[3]) : make_double2(0.0, 0.0);
   const double *mc = ec + 18;   // m0r m0i m1r m1i m2r m2i kxy c0x c0y c1x c1y
   const double2 al0 = make_double2(mc[0] * v0 - mc[6] * v1, mc[1] * v0 + mc[8] * c2);
   const double2 al1 = make_double2(mc[2] * v1 - mc[6] * v0, mc[3] * v1 - mc[7] * c2);
   const double2 al2 = make_double2(mc[4] * v2, mc[5] * v2 - mc[8] * c0 + mc[7] * c1);
   const double2 C0 = cmul(wb, al0), C1 = cmul(wb, al1), C2 = cmul(wa, al2);
   A[(row0 + 0) * NP + k] = v0;
   A[(row0 + 1) * NP + k] = v1;
   A[(row0 + 2) * NP + k] = v2;
   Cr[(row0 + 0) * NP + k] = C0.x;
   Ci[(row0 + 0) * NP + k] = C0.y;
   Cr[(row0 + 1) * NP + k] = C1.x;
   Ci[(row0 + 1) * NP + k] = C1.y;
   Cr[(row0 + 2) * NP + k] = C2.x;
   Ci[(row0 + 2) * NP + k] = C2.y;
   if (rows == 6)
   {
      const double2 al3 = make_double2(0.0, mc[10] * v2);
      const double2 al4 = make_double2(0.0, -mc[9] * v2);
      const double2 al5 = make_double2(0.0, -mc[10] * v0 + mc[9] * v1);
      const double2 C3 = cmul(wa, al3), C4 = cmul(wa, al4), C5 = cmul(wb, al5);
      A[(row0 + 3) * NP + k] = c0;
      A[(row0 + 4) * NP + k] = c1;
      A[(row0 + 5) * NP + k] = c2;
      Cr[(row0 + 3) * NP + k] = C3.x;
      Ci[(row0 + 3) * NP + k] = C3.y;
      Cr[(row0 + 4) * NP + k] = C4.x;
      Ci[(row0 + 4) * NP + k] = C4.y;
      Cr[(row0 + 5) * NP + k] = C5.x;
      Ci[(row0 + 5) * NP + k] = C5.y;
   }
}

__device__ __forceinline__ void form_mass_item(const LocalParams &P, const double *pj,
                                               const double *ec, const double *qc, int q, int k,
                                               bool valid,
                                               int rows, double *A, double *Cr, double *Ci,
                                               int row0, int NP)
{
   const bool live = valid && k < P.nd;
   double hv[3] = {0.0, 0.0, 0.0}, hc[3] = {0.0, 0.0, 0.0};
   if (live)
   {
      const double *vs = P.rm.vs + ((size_t)q * P.nd + k) * 3;
      hv[0] = vs[0];
      hv[1] = vs[1];
      hv[2] = vs[2];
      if (rows == 6)
      {
         const double *cs = P.rm.cs + ((size_t)q * P.nd + k) * 3;
         hc[0] = cs[0];
         hc[1] = cs[1];
         hc[2] = cs[2];
      }
   }
   form_mass_raw(pj, ec, qc, valid, live, rows, hv, hc, A, Cr, Ci, row0, k, NP);
}

// One item (point, basis k) of chunk `ch` per thread, split into the
// reference-table loads and the forming, so a kernel can issue the loads
// of chunk c+1 before the MFMAs of chunk c.  Valid when the chunk has at
// most one item per thread (NP = 16 NT columns, 64 NT threads).
struct ChunkItem {
   double h[6];
   int q, k, row0;
   bool has, valid, live, mass;
};

__device__ __forceinline__ ChunkItem load_chunk_item(const LocalParams &P, const StreamShape &ss,
                                                     int ch, int tid, int NP)
{
   ChunkItem it;
#pragma unroll
   for (int i = 0; i < 6; ++i) { it.h[i] = 0.0; }
   it.mass = ch >= ss.nch_cc;
   const int qpc = it.mass ? ss.qpc_m : ss.qpc_cc;
   const int nq = it.mass ? ss.nq_m : ss.nq_cc;
   const int c = it.mass ? ch - ss.nch_cc : ch;
   const int rows = it.mass ? ss.rows_m : 3;
   it.has = tid < qpc * NP;
   const int qi = tid / NP;
   it.k = tid - qi * NP;
   const int q = c * qpc + qi;
   it.valid = it.has && q < nq;
   it.q = it.valid ? q : 0;
   it.row0 = qi * rows;
   it.live = it.valid && it.k < P.nd;
   if (it.live)
   {
      const RuleDev &r = it.mass ? P.rm : P.rc;
      const double *t0 = (it.mass ? r.vs : r.cs) + ((size_t)it.q * P.nd + it.k) * 3;
      it.h[0] = t0[0];
      it.h[1] = t0[1];
      it.h[2] = t0[2];
      if (it.mass && rows == 6)
      {
         const double *t1 = r.cs + ((size_t)it.q * P.nd + it.k) * 3;
         it.h[3] = t1[0];
         it.h[4] = t1[1];
         it.h[5] = t1[2];
      }
   }
   return it;
}

// pj: the element's Piola constants (J / detJ, J^-1; registers in the MFMA
// kernels), ec: the element constants in LDS (mass coefficients).
__device__ __forceinline__ void form_chunk_item(const LocalParams &P, const double *pj,
                                                const double *ec, const StreamShape &ss,
                                                const double *qcoef, const ChunkItem &it,
                                                double *A, double *Cr, double *Ci, int NP)
{
   if (!it.has) { return; }
   if (!it.mass)
   {
      form_cc_raw(pj, qcoef + (size_t)it.q * 4, it.valid, it.live, it.h[0], it.h[1], it.h[2],
                  A, Cr, Ci, it.row0, it.k, NP);
   }
   else
   {
      form_mass_raw(pj, ec, qcoef + (size_t)(P.rc.nq + it.q) * 4, it.valid, it.live, ss.rows_m,
                    it.h, it.h + 3, A, Cr, Ci, it.row0, it.k, NP);
   }
}

// Form chunk `ch` of the combined K stream (cc chunks first, then mass
// chunks) for the element group: items (point, basis) strided over nthr
// threads starting at tid.  Staging: A, Cr, Ci of CHUNK_ROWS x NP each.
__device__ __forceinline__ void form_chunk(const LocalParams &P, const double *pj,
                                           const double *ec, const StreamShape &ss,
                                           const double *qcoef,
                                           int ch, int tid, int nthr, double *A, double *Cr,
                                           double *Ci, int NP)
{
   if (ch < ss.nch_cc)
   {
      const int nitem = ss.qpc_cc * NP;
      for (int it = tid; it < nitem; it += nthr)
      {
         const int qi = it / NP, k = it - qi * NP;
         const int q = ch * ss.qpc_cc + qi;
         const bool valid = q < ss.nq_cc;
         const int qs = valid ? q : 0;
         form_cc_item(P, pj, qcoef + (size_t)qs * 4, qs, k, valid, A, Cr, Ci, qi * 3, NP);
      }
   }
   else
   {
      const int c = ch - ss.nch_cc;
      const int nitem = ss.qpc_m * NP;
      for (int it = tid; it < nitem; it += nthr)
      {
         const int qi = it / NP, k = it - qi * NP;
         const int q = c * ss.qpc_m + qi;
         const bool valid = q < ss.nq_m;
         const int qs = valid ? q : 0;
         form_mass_item(P, pj, ec, qcoef + (size_t)(P.rc.nq + qs) * 4, qs, k, valid, ss.rows_m,
                        A, Cr, Ci, qi * ss.rows_m, NP);
      }
   }
}

// ---- MFMA local-matrix kernel ---------------------------------------------
// v_mfma_f64_16x16x4_f64: A 16x4, B 4x16, D 16x16.  Lane l holds
// A[l%16][l/16], B[l/16][l%16] and D[l/16 + 4v][l%16], v = 0..3 (the FP64
// output interleaves rows across the four 16-lane groups).
// Ae[k][l] = sum_K Bst[K][k] C[K][l]: A operand = staged shape rows,
// B operand = staged coefficient rows (re and im separately).
//
// NT = column tiles (nd padded to 16 NT), RT = row tiles per wavefront,
// WPE = NT / RT wavefronts per element, EPB elements per block; the block
// has 64 WPE EPB threads.  WAVES_PER_EU bounds the register budget.

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2v __attribute__((ext_vector_type(2)));

// Per-element setup record of the MFMA quadrature kernels, formed ahead of
// them by a one-wavefront-per-element kernel: [element constants (EC_SIZE) |
// per-point factors (4 per rule point)] doubles, the kernels' LDS layout.
// The record of list entry io sits in the first doubles of that entry's
// own element-matrix slot in Ae (nd * lda complex entries), which the
// quadrature kernel reads before it writes the entry's matrix: the
// dependent geometry/material loads, the PML profile (pow) and the
// divisions leave the MFMA kernels.
__device__ __forceinline__ double *elem_rec(const LocalParams &P, int io, int e)
{
   return reinterpret_cast<double *>(P.Ae + (size_t)(P.out_by_elem ? e : io) * P.nd * P.lda);
}

__global__ void __launch_bounds__(64) k_elem_prep(LocalParams P, int nqt)
{
   const int io = blockIdx.x;
   const int e = P.elist ? P.elist[io] : io;
   elem_setup(P, e, nqt, threadIdx.x, 64, elem_rec(P, io, e) + EC_SIZE, elem_rec(P, io, e));
}

// Partner values of the face-pair transforms, moved by VALU lane
// operations instead of the LDS crossbar (ds_bpermute): lane ^ 1 by a DPP
// quad permutation [1, 0, 3, 2]; lane ^ 16 by v_permlane16_swap, which
// swaps the odd 16-lane rows of its first operand with the even rows of its
// second -- with the same register as both, the first result holds rows
// [x0 x0 x2 x2] and the second [x1 x1 x3 x3], so a lane of an even row takes
// the second and a lane of an odd row the first (lane layout checked on the
// device).  Every lane of the wavefront must be active.
__device__ __forceinline__ double lane_xor1(double x)
{
   const long long v = __double_as_longlong(x);
   const int lo = __builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
   const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), 0xB1, 0xF, 0xF, false);
   return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double lane_xor16(double x)
{
   const long long v = __double_as_longlong(x);
   const unsigned lo = (unsigned)v, hi = (unsigned)((unsigned long long)v >> 32);
   const auto pl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
   const auto ph = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
   const bool odd = ((__lane_id() >> 4) & 1) != 0;
   const unsigned rl = odd ? pl[0] : pl[1], rh = odd ? ph[0] : ph[1];
   return __longlong_as_double((long long)(((unsigned long long)rh << 32) | rl));
}

// Face transform and store of the MFMA accumulators of one wavefront
// (row tiles wsub*RT .. wsub*RT+RT-1, all NT column tiles), one 16 x 16
// tile at a time: its columns, then its rows, then its store, so that only
// one tile's partner values are live at once (a face DOF pair (d0, d0+1)
// starts at an even index and never straddles two tiles).
template <int NT, int RT>
__device__ __forceinline__ void mfma_epilogue(const LocalParams &P, dbl4 (&acc)[RT][NT][2],
                                              int e, int io, bool live, int wsub, int g4,
                                              int m16)
{
   // dual face-DOF transform TInv(fo)^T A TInv: columns, then rows
   // (solver/asm_gpu.inc:221-251)
   const int nd = P.nd, p = P.p;
   const bool te = P.apply_te && p >= 2;
   const int foff = 6 * p, ntd = p * (p - 1);
   const int *fo = P.fo + (size_t)e * 4;
   double2 *Ae = P.Ae + (size_t)(P.out_by_elem ? e : io) * nd * P.lda;
   // rows padded to whole 128-byte lines: the 16 columns of a tile row are
   // two whole lines (pad columns carry the zero accumulators), so the
   // stores go out non-temporal; unpadded rows keep ordinary stores
   const bool lines = (P.lda & 7) == 0;
#pragma unroll
   for (int t = 0; t < RT; ++t)
   {
      // a tile with no face DOF among its 16 indices has weights (1, 0)
      // throughout: its pass is the identity and is skipped (wave-uniform)
      const int I = wsub * RT + t;
      const bool face_rows = te && 16 * I + 16 > foff && 16 * I < foff + 4 * ntd;
      // row weights of this row tile: row = 16 I + g4 + 4 v, so the pair
      // (d0, d0+1) of a face sits in lanes l (g4 even) and l ^ 16 (g4
      // odd), same register
      double rcs[4], rcp[4];
#pragma unroll
      for (int v = 0; v < 4; ++v)
      {
         const int row = 16 * (wsub * RT + t) + g4 + 4 * v;
         rcs[v] = 1.0;
         rcp[v] = 0.0;
         if (te && row >= foff && row < foff + 4 * ntd)
         {
            const double *Ti = P.TInv + 4 * fo[(row - foff) / ntd];
            const bool even = ((row - foff) & 1) == 0;
            rcs[v] = even ? Ti[0] : Ti[3];
            rcp[v] = even ? Ti[1] : Ti[2];
         }
      }
#pragma unroll
      for (int j = 0; j < NT; ++j)
      {
         const int col = 16 * j + m16;
         const bool face_cols = te && 16 * j + 16 > foff && 16 * j < foff + 4 * ntd;
         if (face_cols)
         {
            double cs = 1.0, cp = 0.0;
            if (col >= foff && col < foff + 4 * ntd)
            {
               const double *Ti = P.TInv + 4 * fo[(col - foff) / ntd];
               const bool even = ((col - foff) & 1) == 0;
               cs = even ? Ti[0] : Ti[3];
               cp = even ? Ti[1] : Ti[2];
            }
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
               for (int v = 0; v < 4; ++v)
               {
                  const double x = acc[t][j][c][v];
                  const double y = lane_xor1(x);
                  acc[t][j][c][v] = cs * x + cp * y;
               }
         }
         if (face_rows)
         {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
               for (int c = 0; c < 2; ++c)
               {
                  const double x = acc[t][j][c][v];
                  const double y = lane_xor16(x);
                  acc[t][j][c][v] = rcs[v] * x + rcp[v] * y;
               }
         }
         if (!live) { continue; }
#pragma unroll
         for (int v = 0; v < 4; ++v)
         {
            const int row = 16 * (wsub * RT + t) + g4 + 4 * v;
            if (row < nd && col < P.lda)
            {
               if (lines)
               {
                  __builtin_nontemporal_store(dbl2v{acc[t][j][0][v], acc[t][j][1][v]},
                                              reinterpret_cast<dbl2v *>(Ae) + (size_t)row * P.lda + col);
               }
               else if (col < nd)
               {
                  Ae[(size_t)row * P.lda + col] = make_double2(acc[t][j][0][v], acc[t][j][1][v]);
               }
            }
         }
      }
   }
}

template <int NT, int RT, int EPB, int WAVES_PER_EU>
__global__ void __launch_bounds__(64 * (NT / RT) * EPB)
__attribute__((amdgpu_waves_per_eu(WAVES_PER_EU)))
k_local_mfma(LocalParams P, int nqt)
{
   constexpr int NP = 16 * NT;
   constexpr int WPE = NT / RT;
   static_assert(NT % RT == 0, "row tiles must divide the tile count");
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int slot = wave / WPE, wsub = wave - slot * WPE;
   const int i_raw = blockIdx.x * EPB + slot;
   const bool live = i_raw < P.nout;
   const int io = live ? i_raw : P.nout - 1;
   const int e = P.elist ? P.elist[io] : io;

   extern __shared__ double smem[];
   const int slot_doubles = EC_SIZE + nqt * 4 + 3 * CHUNK_ROWS * NP;
   double *ec = smem + (size_t)slot * slot_doubles;
   double *qcoef = ec + EC_SIZE;
   double *A = qcoef + nqt * 4;
   double *Cr = A + CHUNK_ROWS * NP;
   double *Ci = Cr + CHUNK_ROWS * NP;
   const int gtid = wsub * 64 + lane;   // thread index within the element group

   // the element's setup record (k_elem_prep) into LDS, its Piola constants
   // into scalar registers
   double pj[18];
   {
      const double *rec = elem_rec(P, io, e);
      for (int i = gtid; i < EC_SIZE + 4 * nqt; i += WPE * 64) { ec[i] = rec[i]; }
#pragma unroll
      for (int i = 0; i < 18; ++i) { pj[i] = rec[i]; }
   }
   const StreamShape ss = stream_shape(P.use_cc, P.mflags, P.rc.nq, P.rm.nq);
   __syncthreads();

   dbl4 acc[RT][NT][2];
#pragma unroll
   for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int j = 0; j < NT; ++j)
      {
         acc[t][j][0] = dbl4{0.0, 0.0, 0.0, 0.0};
         acc[t][j][1] = dbl4{0.0, 0.0, 0.0, 0.0};
      }

   const int g4 = lane >> 4, m16 = lane & 15;
   const int nch = ss.nch_cc + ss.nch_m;
   for (int ch = 0; ch < nch; ++ch)
   {
      form_chunk(P, pj, ec, ss, qcoef, ch, gtid, WPE * 64, A, Cr, Ci, NP);
      __syncthreads();
#pragma unroll 1
      for (int ks = 0; ks < CHUNK_ROWS / 4; ++ks)
      {
         const int row = ks * 4 + g4;
         double a[RT], br[NT], bi[NT];
#pragma unroll
         for (int t = 0; t < RT; ++t) { a[t] = A[row * NP + 16 * (wsub * RT + t) + m16]; }
#pragma unroll
         for (int j = 0; j < NT; ++j)
         {
            br[j] = Cr[row * NP + 16 * j + m16];
            bi[j] = Ci[row * NP + 16 * j + m16];
         }
#pragma unroll
         for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int j = 0; j < NT; ++j)
            {
               acc[t][j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], br[j], acc[t][j][0], 0, 0, 0);
               acc[t][j][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], bi[j], acc[t][j][1], 0, 0, 0);
            }
      }
      __syncthreads();
   }

   mfma_epilogue<NT, RT>(P, acc, e, io, live, wsub, g4, m16);
}

// Pipelined variant: one row tile per wavefront (NT wavefronts per
// element), one (point, basis) item per thread and chunk, double-buffered
// staging: the MFMAs of chunk c and the forming of chunk c+1 share one
// barrier interval, so table loads and coefficient arithmetic overlap the
// matrix-core work.
template <int NT, int EPB, int WAVES_PER_EU>
__global__ void __launch_bounds__(64 * NT * EPB)
__attribute__((amdgpu_waves_per_eu(WAVES_PER_EU)))
k_local_mfma_pipe(LocalParams P, int nqt)
{
   constexpr int NP = 16 * NT;
   constexpr int STG = 3 * CHUNK_ROWS * NP;     // one staging buffer (A, Cr, Ci)
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int slot = wave / NT, wsub = wave - slot * NT;
   const int i_raw = blockIdx.x * EPB + slot;
   const bool live = i_raw < P.nout;
   const int io = live ? i_raw : P.nout - 1;
   const int e = P.elist ? P.elist[io] : io;

   extern __shared__ double smem[];
   const int slot_doubles = EC_SIZE + nqt * 4 + 2 * STG;
   double *ec = smem + (size_t)slot * slot_doubles;
   double *qcoef = ec + EC_SIZE;
   double *buf = qcoef + nqt * 4;
   const int gtid = wsub * 64 + lane;

   // the element's setup record (k_elem_prep) into LDS, its Piola constants
   // into scalar registers
   double pj[18];
   {
      const double *rec = elem_rec(P, io, e);
      for (int i = gtid; i < EC_SIZE + 4 * nqt; i += NT * 64) { ec[i] = rec[i]; }
#pragma unroll
      for (int i = 0; i < 18; ++i) { pj[i] = rec[i]; }
   }
   const StreamShape ss = stream_shape(P.use_cc, P.mflags, P.rc.nq, P.rm.nq);
   __syncthreads();

   dbl4 acc[1][NT][2];
#pragma unroll
   for (int j = 0; j < NT; ++j)
   {
      acc[0][j][0] = dbl4{0.0, 0.0, 0.0, 0.0};
      acc[0][j][1] = dbl4{0.0, 0.0, 0.0, 0.0};
   }
   const int g4 = lane >> 4, m16 = lane & 15;
   const int nch = ss.nch_cc + ss.nch_m;
   if (nch > 0)
   {
      form_chunk(P, pj, ec, ss, qcoef, 0, gtid, NT * 64, buf, buf + CHUNK_ROWS * NP,
                 buf + 2 * CHUNK_ROWS * NP, NP);
   }
   __syncthreads();
   for (int ch = 0; ch < nch; ++ch)
   {
      const double *A = buf + (ch & 1) * STG;
      const double *Cr = A + CHUNK_ROWS * NP;
      const double *Ci = Cr + CHUNK_ROWS * NP;
      // the reference-table loads of the next chunk go out before this
      // chunk's MFMAs (the last chunk reloads its own item, unused)
      const bool more = ch + 1 < nch;
      const ChunkItem nxt = load_chunk_item(P, ss, more ? ch + 1 : ch, gtid, NP);
#pragma unroll
      for (int ks = 0; ks < CHUNK_ROWS / 4; ++ks)
      {
         const int row = ks * 4 + g4;
         const double a = A[row * NP + 16 * wsub + m16];
         double br[NT], bi[NT];
#pragma unroll
         for (int j = 0; j < NT; ++j)
         {
            br[j] = Cr[row * NP + 16 * j + m16];
            bi[j] = Ci[row * NP + 16 * j + m16];
         }
#pragma unroll
         for (int j = 0; j < NT; ++j)
         {
            acc[0][j][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, br[j], acc[0][j][0], 0, 0, 0);
            acc[0][j][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bi[j], acc[0][j][1], 0, 0, 0);
         }
      }
      if (more)
      {
         double *An = buf + ((ch + 1) & 1) * STG;
         form_chunk_item(P, pj, ec, ss, qcoef, nxt, An, An + CHUNK_ROWS * NP,
                         An + 2 * CHUNK_ROWS * NP, NP);
      }
      __syncthreads();
   }
   mfma_epilogue<NT, 1>(P, acc, e, io, live, wsub, g4, m16);
}

// ---- generic VALU local-matrix kernel (any nd) -----------------------------
// One block per element; the same K-stream staging; each thread accumulates
// the entries idx = tid + BLOCK*i of the nd x nd matrix.  The face-DOF
// transform runs afterwards in k_te.

template <int NACC, int NTHR>
__global__ void __launch_bounds__(NTHR)
k_local_generic(LocalParams P, int nqt, int NP)
{
   const int io = blockIdx.x;
   const int e = P.elist ? P.elist[io] : io;
   const int tid = threadIdx.x;
   extern __shared__ double smem[];
   double *ec = smem;
   double *qcoef = ec + EC_SIZE;
   double *A = qcoef + nqt * 4;
   double *Cr = A + CHUNK_ROWS * NP;
   double *Ci = Cr + CHUNK_ROWS * NP;

   elem_setup(P, e, nqt, tid, NTHR, qcoef, ec);
   const StreamShape ss = stream_shape(P.use_cc, P.mflags, P.rc.nq, P.rm.nq);
   __syncthreads();

   const int nd = P.nd, n2 = nd * nd;
   double ar[NACC], ai[NACC];
#pragma unroll
   for (int i = 0; i < NACC; ++i) { ar[i] = 0.0; ai[i] = 0.0; }

   const int nch = ss.nch_cc + ss.nch_m;
   for (int ch = 0; ch < nch; ++ch)
   {
      form_chunk(P, ec, ec, ss, qcoef, ch, tid, NTHR, A, Cr, Ci, NP);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NACC; ++i)
      {
         const int idx = tid + NTHR * i;
         if (idx < n2)
         {
            const int k = idx / nd, l = idx - k * nd;
            double sr = ar[i], si = ai[i];
            for (int r = 0; r < CHUNK_ROWS; ++r)
            {
               const double a = A[r * NP + k];
               sr += a * Cr[r * NP + l];
               si += a * Ci[r * NP + l];
            }
            ar[i] = sr;
            ai[i] = si;
         }
      }
      __syncthreads();
   }
   double2 *Ae = P.Ae + (size_t)(P.out_by_elem ? e : io) * nd * P.lda;
#pragma unroll
   for (int i = 0; i < NACC; ++i)
   {
      const int idx = tid + NTHR * i;
      if (idx < n2)
      {
         const int k = idx / nd, l = idx - k * nd;
         Ae[(size_t)k * P.lda + l] = make_double2(ar[i], ai[i]);
      }
   }
}

// Dual face-DOF transform in place on Ae (solver/asm_gpu.inc:221-251):
// one block per element, columns then rows.
__global__ void k_te(int nd, int lda, int p, const int *__restrict__ fo,
                     const int *__restrict__ elist, int out_by_elem, const double *TInvG,
                     double2 *__restrict__ Ae)
{
   const int io = blockIdx.x;
   const int e = elist ? elist[io] : io;
   double2 *A = Ae + (size_t)(out_by_elem ? e : io) * nd * lda;
   const int foff = 6 * p;
   const int ntd = p * (p - 1);
   for (int axis = 0; axis < 2; ++axis)
   {
      for (int f = 0; f < 4; ++f)
      {
         const double *Ti = TInvG + 4 * fo[(size_t)e * 4 + f];
         for (int i = 0; i < ntd / 2; ++i)
         {
            const int d0 = foff + f * ntd + 2 * i;
            for (int r = threadIdx.x; r < nd; r += blockDim.x)
            {
               double2 &a = axis == 0 ? A[(size_t)r * lda + d0] : A[(size_t)d0 * lda + r];
               double2 &b = axis == 0 ? A[(size_t)r * lda + d0 + 1] : A[(size_t)(d0 + 1) * lda + r];
               const double2 na = make_double2(Ti[0] * a.x + Ti[1] * b.x, Ti[0] * a.y + Ti[1] * b.y);
               const double2 nb = make_double2(Ti[2] * a.x + Ti[3] * b.x, Ti[2] * a.y + Ti[3] * b.y);
               a = na;
               b = nb;
            }
         }
      }
      __syncthreads();
   }
}

// ---- CSR slot map ----------------------------------------------------------
// slot[(e*nd + a)*sld + b] = position of column dof[e][b] inside row
// dof[e][a] (lower bound in the sorted row, as solver/asm_gpu.inc:447-453
// and solver/asm_data.inc:393).  A missing column sets *bad.
__global__ void k_slot_map(int ne, int nd, int sld, const int *__restrict__ dof,
                           const int *__restrict__ Ip, const int *__restrict__ Jc,
                           uint16_t *__restrict__ slot, int *__restrict__ bad)
{
   const size_t n2 = (size_t)nd * nd;
   const size_t total = (size_t)ne * n2;
   for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
        idx += (size_t)gridDim.x * blockDim.x)
   {
      const size_t e = idx / n2;
      const int ab = (int)(idx - e * n2);
      const int a = ab / nd, b = ab - a * nd;
      const int r = dof[e * nd + a], c = dof[e * nd + b];
      int lo = Ip[r], hi = Ip[r + 1];
      const int base = lo;
      while (lo < hi)
      {
         const int mid = (lo + hi) >> 1;
         if (Jc[mid] < c) { lo = mid + 1; }
         else { hi = mid; }
      }
      const size_t o = (e * nd + a) * (size_t)sld + b;   // rows of sld entries
      if (lo >= Ip[r + 1] || Jc[lo] != c) { *bad = 1; slot[o] = 0; }
      else { slot[o] = (uint16_t)(lo - base); }
   }
}

// Slot rows regrouped per CSR row as bytes (every row has at most 256
// entries), in the row's contributor order (rc_list): block of row r =
// rslot[rsoff[r] ..], contributor c - c0 at (c - c0) * nd.  One wavefront
// per row.
__global__ void __launch_bounds__(BLOCK) k_row_slots(int n, int nd, int sld,
                                                     const int *__restrict__ rc_ptr,
                                                     const int *__restrict__ rc_list,
                                                     const uint16_t *__restrict__ slot,
                                                     const int *__restrict__ rsoff,
                                                     uint8_t *__restrict__ rslot)
{
   const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   if (r >= n) { return; }
   const int c0 = rc_ptr[r], c1 = rc_ptr[r + 1];
   uint8_t *dst = rslot + rsoff[r];
   for (int c = c0; c < c1; ++c)
   {
      const int ea = rc_list[c];
      for (int b = lane; b < nd; b += 64) { dst[(size_t)(c - c0) * nd + b] = (uint8_t)slot[(size_t)ea * sld + b]; }
   }
}

// ---- deterministic row-segmented scatter + elimination ----------------------
// One wavefront per CSR row r.  The row is accumulated in an LDS buffer from
// its contributors (element, local index) in ascending element order --
// the summation order of the reference's CPU reassembler
// (solver/asm_data.inc:381-400) -- then written once.  With `eliminate`,
// essential rows become (0 ... 1 ... 0) and essential columns 0
// (solver/asm_gpu.inc:463-478).
struct ScatterParams {
   int n, nd;
   const int *Ip, *Jc;
   const unsigned char *eflag;  // per (e, b): bit0 sign < 0, bit1 dof essential
   const uint64_t *emask;       // nd <= 64: per e, {sign < 0 bits, essential bits} (else NULL)
   const int *rc_ptr, *rc_list;
   const unsigned char *is_ess;
   const double2 *Ae;
   int lda;                     // row stride of the element matrices in Ae
   double2 *vals;
   int eliminate;
   int maxrow;
   const uint16_t *slot;        // element-major slot map (rows of sld entries)
   int sld;
   const int *rsoff;            // per row: start of its slot block in rslot
   const uint8_t *rslot;        // the row's contributors' slot rows as bytes, in
                                // contributor order (rows of at most 256 entries)
   const int *rows;             // rows to form (NULL: rows 0 .. nrows-1)
   int nrows;
};

__device__ __forceinline__ void wave_sync_lds()
{
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Contributors are processed in groups of CG: the loads of a group are
// issued together, the LDS accumulation then runs contributor by
// contributor (ascending element order) so the sums are deterministic.
constexpr int SCATTER_CG = 4;   // 6 and 8 measured slower at p = 3

// M contributors (list entries cg .. cg+M-1) of one CSR row into the row
// buffer: every load of the group is issued before the first use, with no
// branch between them, so they are in flight together; the accumulation
// then runs in ascending element order (deterministic sums).
template <int M, bool BYTE_SLOTS>
__device__ __forceinline__ void scatter_group(const ScatterParams &S, double2 *buf, int cg, int lane,
                                              unsigned ess_bit, const uint8_t *sblk)
{
   const int nd = S.nd;
   int ea[M];   // e*nd + a
#pragma unroll
   for (int u = 0; u < M; ++u) { ea[u] = S.rc_list[cg + u]; }
   for (int b0 = 0; b0 < nd; b0 += 64)
   {
      const int b = b0 + lane;
      const int bc = min(b, nd - 1);   // lanes past the row load a valid entry, unused
      dbl2v x[M];
      int sl[M];
      unsigned fa[M], fb[M];
#pragma unroll
      for (int u = 0; u < M; ++u)
      {
         x[u] = __builtin_nontemporal_load(reinterpret_cast<const dbl2v *>(S.Ae) +
                                           (size_t)ea[u] * S.lda + bc);
      }
      // byte slots: the group's slot rows, contiguous in the row's block
      // (the element-major 16-bit map, where a row's slot rows lie apart,
      // ran 0.97 against 0.874 ms at 1.35M DOF, p = 3; at p = 4, rows above
      // 256 entries, a per-row 16-bit copy was slower than the map)
#pragma unroll
      for (int u = 0; u < M; ++u)
      {
         if constexpr (BYTE_SLOTS) { sl[u] = sblk[u * nd + bc]; }
         else { sl[u] = S.slot[(size_t)ea[u] * S.sld + bc]; }
      }
#pragma unroll
      for (int u = 0; u < M; ++u)
      {
         const int e = ea[u] / nd;
         if (S.emask)
         {
            // two uniform 8-byte loads instead of a byte per lane
            const uint64_t sm = S.emask[2 * (size_t)e], em = S.emask[2 * (size_t)e + 1];
            fa[u] = (unsigned)((sm >> (ea[u] - e * nd)) & 1);
            fb[u] = (unsigned)(((sm >> bc) & 1) | (((em >> bc) & 1) << 1));
         }
         else
         {
            fa[u] = S.eflag[ea[u]];
            fb[u] = S.eflag[(size_t)e * nd + bc];
         }
      }
#pragma unroll
      for (int u = 0; u < M; ++u)
      {
         if (b < nd && (fb[u] & ess_bit) == 0)
         {
            const bool neg = ((fa[u] ^ fb[u]) & 1) != 0;
            double2 t = buf[sl[u]];
            t.x += neg ? -x[u].x : x[u].x;
            t.y += neg ? -x[u].y : x[u].y;
            buf[sl[u]] = t;
         }
         wave_sync_lds();
      }
   }
}

// nd <= 32: two contributors per load instruction -- lanes 0-31 carry
// contributor cg + 2u, lanes 32-63 contributor cg + 2u + 1 -- for the
// first m (<= 2 NP) contributors of the group starting at list entry cg;
// lanes of a half without a contributor issue no loads.  The accumulation
// still runs contributor by contributor in list order (deterministic sums,
// the same as scatter_group's).  At p = 2 a row's 20 entries filled a third
// of a load instruction's lanes.
template <int NP, bool BYTE_SLOTS>
__device__ __forceinline__ void scatter_group2(const ScatterParams &S, double2 *buf, int cg, int m,
                                               int lane, unsigned ess_bit, const uint8_t *sblk)
{
   const int nd = S.nd;
   const int half = lane >> 5;
   const int b = lane & 31;
   const int bc = min(b, nd - 1);      // lanes past the row load a valid entry, unused
   int ea[NP];
   bool valid[NP];
   dbl2v x[NP];
   int sl[NP];
   unsigned fa[NP], fb[NP];
#pragma unroll
   for (int u = 0; u < NP; ++u)
   {
      valid[u] = 2 * u + half < m;
      ea[u] = valid[u] ? S.rc_list[cg + 2 * u + half] : 0;
   }
#pragma unroll
   for (int u = 0; u < NP; ++u)
   {
      x[u] = dbl2v{0.0, 0.0};
      if (valid[u])
      {
         x[u] = __builtin_nontemporal_load(reinterpret_cast<const dbl2v *>(S.Ae) +
                                           (size_t)ea[u] * S.lda + bc);
      }
   }
#pragma unroll
   for (int u = 0; u < NP; ++u)
   {
      sl[u] = 0;
      if (valid[u])
      {
         if constexpr (BYTE_SLOTS) { sl[u] = sblk[(2 * u + half) * nd + bc]; }
         else { sl[u] = S.slot[(size_t)ea[u] * S.sld + bc]; }
      }
   }
#pragma unroll
   for (int u = 0; u < NP; ++u)
   {
      fa[u] = 0;
      fb[u] = 2;   // treated as essential: never added
      if (valid[u])
      {
         const int e = ea[u] / nd;
         if (S.emask)
         {
            const uint64_t sm = S.emask[2 * (size_t)e], em = S.emask[2 * (size_t)e + 1];
            fa[u] = (unsigned)((sm >> (ea[u] - e * nd)) & 1);
            fb[u] = (unsigned)(((sm >> bc) & 1) | (((em >> bc) & 1) << 1));
         }
         else
         {
            fa[u] = S.eflag[ea[u]];
            fb[u] = S.eflag[(size_t)e * nd + bc];
         }
      }
   }
#pragma unroll
   for (int u = 0; u < NP; ++u)
   {
#pragma unroll
      for (int h = 0; h < 2; ++h)
      {
         if (half == h && valid[u] && b < nd && (fb[u] & ess_bit) == 0)
         {
            const bool neg = ((fa[u] ^ fb[u]) & 1) != 0;
            double2 t = buf[sl[u]];
            t.x += neg ? -x[u].x : x[u].x;
            t.y += neg ? -x[u].y : x[u].y;
            buf[sl[u]] = t;
         }
         wave_sync_lds();
      }
   }
}

// scatter_group2<np> for the pairs of a group of m <= 2 NP contributors
// (np = ceil(m / 2), wave-uniform)
template <int NP, bool BYTE_SLOTS>
__device__ __forceinline__ void scatter_dispatch2(const ScatterParams &S, double2 *buf, int cg, int m,
                                                  int lane, unsigned ess_bit, const uint8_t *sblk)
{
   if constexpr (NP == 1) { scatter_group2<1, BYTE_SLOTS>(S, buf, cg, m, lane, ess_bit, sblk); }
   else
   {
      if ((m + 1) / 2 == NP) { scatter_group2<NP, BYTE_SLOTS>(S, buf, cg, m, lane, ess_bit, sblk); }
      else { scatter_dispatch2<NP - 1, BYTE_SLOTS>(S, buf, cg, m, lane, ess_bit, sblk); }
   }
}

// scatter_group<m> for a group of m <= M contributors (m is wave-uniform)
template <int M, bool BYTE_SLOTS>
__device__ __forceinline__ void scatter_dispatch(const ScatterParams &S, double2 *buf, int cg, int m,
                                                 int lane, unsigned ess_bit, const uint8_t *sblk)
{
   if constexpr (M == 1) { scatter_group<1, BYTE_SLOTS>(S, buf, cg, lane, ess_bit, sblk); }
   else
   {
      if (m == M) { scatter_group<M, BYTE_SLOTS>(S, buf, cg, lane, ess_bit, sblk); }
      else { scatter_dispatch<M - 1, BYTE_SLOTS>(S, buf, cg, m, lane, ess_bit, sblk); }
   }
}

constexpr int SCATTER_NP = 4;   // contributor pairs per group of the packed form (nd <= 32)

// PACK2 (nd <= 32): two contributors per load instruction (scatter_group2)
// Row r of the CSR values by one wavefront (buf: its LDS row buffer).
template <bool BYTE_SLOTS, bool PACK2>
__device__ __forceinline__ void scatter_row(const ScatterParams &S, double2 *buf, int r, int lane)
{
   const int base = S.Ip[r];
   const int len = S.Ip[r + 1] - base;
   for (int s = lane; s < len; s += 64) { buf[s] = make_double2(0.0, 0.0); }
   wave_sync_lds();
   const bool row_ess = S.eliminate && S.is_ess[r];
   const unsigned ess_bit = S.eliminate ? 2u : 0u;
   if (!row_ess)
   {
      const int c0 = S.rc_ptr[r], c1 = S.rc_ptr[r + 1];
      const uint8_t *sblk = BYTE_SLOTS ? S.rslot + S.rsoff[r] : nullptr;
      if constexpr (PACK2)
      {
         for (int cg = c0; cg < c1; cg += 2 * SCATTER_NP)
         {
            scatter_dispatch2<SCATTER_NP, BYTE_SLOTS>(
               S, buf, cg, min(2 * SCATTER_NP, c1 - cg), lane, ess_bit,
               BYTE_SLOTS ? sblk + (size_t)(cg - c0) * S.nd : nullptr);
         }
      }
      else
      {
         for (int cg = c0; cg < c1; cg += SCATTER_CG)
         {
            scatter_dispatch<SCATTER_CG, BYTE_SLOTS>(S, buf, cg, min(SCATTER_CG, c1 - cg), lane,
                                                     ess_bit,
                                                     BYTE_SLOTS ? sblk + (size_t)(cg - c0) * S.nd
                                                                : nullptr);
         }
      }
   }
   else
   {
      for (int s = lane; s < len; s += 64)
      {
         if (S.Jc[base + s] == r) { buf[s] = make_double2(1.0, 0.0); }
      }
      wave_sync_lds();
   }
   // each store instruction writes 64 consecutive entries (whole lines):
   // non-temporal stores measured faster here (1.057 vs 1.138 ms at 1.35M
   // DOF)
   for (int s = lane; s < len; s += 64)
   {
      const double2 t = buf[s];
      __builtin_nontemporal_store(dbl2v{t.x, t.y}, reinterpret_cast<dbl2v *>(S.vals) + base + s);
   }
}

// Every row, one per wavefront, 4 rows per workgroup.
template <bool BYTE_SLOTS, bool PACK2 = false>
__global__ void __launch_bounds__(BLOCK) k_scatter_rows(ScatterParams S)
{
   extern __shared__ double2 rowbuf[];
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int r = blockIdx.x * 4 + wave;
   if (r >= S.n) { return; }
   scatter_row<BYTE_SLOTS, PACK2>(S, rowbuf + (size_t)wave * S.maxrow, r, lane);
}

// The rows of the list S.rows (S.nrows of them), one per wavefront.
template <bool BYTE_SLOTS, bool PACK2 = false>
__global__ void __launch_bounds__(BLOCK) k_scatter_row_list(ScatterParams S)
{
   extern __shared__ double2 rowbuf[];
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int ri = blockIdx.x * 4 + wave;
   if (ri >= S.nrows) { return; }
   scatter_row<BYTE_SLOTS, PACK2>(S, rowbuf + (size_t)wave * S.maxrow, S.rows[ri], lane);
}

// Rows with at least one contributor outside the constant-coefficient
// (tensor-path) element class: flag[r] = 1.
__global__ void k_row_quad_flag(int n, int nd, const int *__restrict__ rc_ptr,
                                const int *__restrict__ rc_list,
                                const unsigned char *__restrict__ eclass,
                                unsigned char *__restrict__ flag)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= n) { return; }
   unsigned char f = 0;
   for (int k = rc_ptr[r]; k < rc_ptr[r + 1]; ++k)
   {
      if (!eclass[rc_list[k] / nd]) { f = 1; }
   }
   flag[r] = f;
}

// ---- complex CSR SpMV (solver/asm_gpu.inc:481-493) ------------------------
__global__ void k_spmv(int n, const int *__restrict__ Ip, const int *__restrict__ Jc,
                       const double2 *__restrict__ vals, const double2 *__restrict__ x,
                       double2 *__restrict__ y)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= n) { return; }
   double yr = 0.0, yi = 0.0;
   for (int k = Ip[r]; k < Ip[r + 1]; ++k)
   {
      const double2 a = vals[k], b = x[Jc[k]];
      yr += a.x * b.x - a.y * b.y;
      yi += a.x * b.y + a.y * b.x;
   }
   y[r] = make_double2(yr, yi);
}

__global__ void k_fill_zero(double2 *p, size_t n)
{
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (size_t)gridDim.x * blockDim.x)
   {
      p[i] = make_double2(0.0, 0.0);
   }
}

} // namespace lfasm
