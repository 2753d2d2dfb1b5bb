// Batched and adjoint device code of the gfx950 Nedelec assembly.  None of
// these have a reference counterpart; they build on the reference's
// assembly (solver/asm_gpu.inc) and RHS definition
// (solver/problem_common.inc:218-266):
//
// * eps variants: A(eps_v) = A(eps_base) + sum_{e in params} deps_{v,e} Mhat_e,
//   where Mhat_e is the face-transformed mass element matrix at unit eps
//   (the mass term is linear in eps); one wavefront per CSR row writes the
//   row of every variant, contributions in ascending element order;
// * batched RHS: b_k = int k0^2 deps_region E_inc(z) . v_k over scattered-
//   field source (frustum) regions, for many incident fields at once;
// * eps sensitivity: s_e = lam_e^T (dA_e/deps_e u_e - db_e/deps_e)
//   = -k0^2 int L(x) . (Lam U(x) + f_e E_inc(z)) with L, U the adjoint and
//   forward fields of the element.
#pragma once

#include "lfasm_device.hpp"

namespace lfasm {

// ---- eps variants -------------------------------------------------------------

struct VariantParams {
   int n, nd, nvar, nparam;
   int64_t nnz;
   const int *Ip;
   const unsigned char *eflag;   // per (e, b): bit0 sign < 0, bit1 essential
   const int *prc_ptr, *prc_list;  // per row: param contributors pi*nd + a, ascending element
   const int *pelem;             // param index -> element
   const uint16_t *slot;         // element slot map (rows of sld entries)
   int sld;
   const unsigned char *is_ess;
   const double2 *Mhat;          // nparam x nd x nd
   const double2 *deps;          // nvar x nparam
   const double2 *depsT;         // nparam x nvar (the same values, parameter-major)
   const double2 *base;          // nnz
   double2 *out;                 // nvar x nnz
   int eliminate, maxrow;
   const unsigned char *vcopied; // per row: 1 = written by k_variant_copy (or NULL)
};

// Register-resident variant of k_variant_rows for nd <= 64 and at most
// VMAXC parameter contributors per row: each lane keeps its column's
// (signed) unit-eps entry, slot and use flag of every contributor, so the
// Mhat rows are read once per row instead of once per variant; the
// variants then only rescale and accumulate.  Same sums, same order.
constexpr int VMAXC = 16;

// value of lane `src` (wave-uniform) of a double held one per lane
__device__ __forceinline__ double lane_bcast(double v, int src)
{
   const long long x = __double_as_longlong(v);
   const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)x, src);
   const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), src);
   return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// nvar x nparam -> nparam x nvar (32 x 32 tiles through LDS)
__global__ void __launch_bounds__(256) k_transpose_deps(const double2 *__restrict__ in,
                                                        double2 *__restrict__ out, int nvar,
                                                        int nparam)
{
   __shared__ double2 tile[32][33];
   const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
   const int p0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
   for (int j = ty; j < 32; j += 8)
   {
      const int v = v0 + j, pi = p0 + tx;
      if (v < nvar && pi < nparam) { tile[j][tx] = in[(size_t)v * nparam + pi]; }
   }
   __syncthreads();
   for (int j = ty; j < 32; j += 8)
   {
      const int pi = p0 + j, v = v0 + tx;
      if (v < nvar && pi < nparam) { out[(size_t)pi * nvar + v] = tile[tx][j]; }
   }
}

// Long runs of consecutive rows without a parameter contributor are the
// base rows in every variant and contiguous ranges of the CSR values:
// copied here in chunks (one workgroup per chunk, whole-line non-temporal
// stores except at the chunk ends).  The row kernels skip those rows
// (vcopied[r] = 1) and keep the short runs, whose stores share their first
// and last lines with neighbouring rows written at about the same time.
__global__ void __launch_bounds__(256) k_variant_copy(const int2 *__restrict__ chunks,
                                                      const double2 *__restrict__ base,
                                                      double2 *__restrict__ out, int64_t nnz,
                                                      int nvar)
{
   const int2 c = chunks[blockIdx.x];
   for (int i = c.x + (int)threadIdx.x; i < c.y; i += 256)
   {
      const double2 x = base[i];
      for (int v = 0; v < nvar; ++v)
      {
         __builtin_nontemporal_store(dbl2v{x.x, x.y},
                                     reinterpret_cast<dbl2v *>(out + (size_t)v * nnz + i));
      }
   }
}

__global__ void __launch_bounds__(BLOCK) k_variant_rows_reg(VariantParams V)
{
   extern __shared__ double2 vbuf[];
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int r = blockIdx.x * 4 + wave;
   if (r >= V.n) { return; }
   double2 *row0 = vbuf + (size_t)wave * 2 * V.maxrow;
   double2 *buf = row0 + V.maxrow;
   if (V.vcopied && V.vcopied[r]) { return; }   // k_variant_copy
   const int base = V.Ip[r];
   const int len = V.Ip[r + 1] - base;
   const int c0 = V.prc_ptr[r], c1 = V.prc_ptr[r + 1];
   const bool row_fixed = c0 == c1 || (V.eliminate && V.is_ess[r]);
   if (row_fixed)
   {
      // the base row, 64 entries at a time, held in registers across the
      // variants (the stores may alias the loads for the compiler)
      for (int s0 = 0; s0 < len; s0 += 64)
      {
         const int s = s0 + lane;
         const double2 x = s < len ? V.base[base + s] : make_double2(0.0, 0.0);
         if (s < len)
         {
            for (int v = 0; v < V.nvar; ++v) { V.out[(size_t)v * V.nnz + base + s] = x; }
         }
      }
      return;
   }
   for (int s = lane; s < len; s += 64) { row0[s] = V.base[base + s]; }
   const int nd = V.nd;
   const int nc = c1 - c0;                    // <= VMAXC (host-checked)
   const unsigned char ess_bit = V.eliminate ? 2 : 0;
   const int b = lane;
   const bool live = b < nd;
   double2 m[VMAXC];
   int sl[VMAXC], pidx[VMAXC];
   bool use[VMAXC];
#pragma unroll
   for (int k = 0; k < VMAXC; ++k)
   {
      m[k] = make_double2(0.0, 0.0);
      sl[k] = 0;
      pidx[k] = 0;
      use[k] = false;
      if (k < nc)
      {
         const int pa = V.prc_list[c0 + k];
         const int pi = pa / nd, a = pa - pi * nd;
         const int e = V.pelem[pi];
         const int ea = e * nd + a;
         pidx[k] = pi;
         if (live)
         {
            const unsigned char fa = V.eflag[ea];
            const unsigned char fb = V.eflag[(size_t)e * nd + b];
            use[k] = (fb & ess_bit) == 0;
            if (use[k])
            {
               sl[k] = V.slot[(size_t)ea * V.sld + b];
               const double2 x = V.Mhat[((size_t)pi * nd + a) * nd + b];
               m[k] = ((fa ^ fb) & 1) ? make_double2(-x.x, -x.y) : x;
            }
         }
      }
   }
   // variants in blocks of 64: lane j holds the eps offsets of variant
   // v0 + j for every contributor (one coalesced load each from the
   // parameter-major copy), broadcast by readlane inside the variant loop
   for (int v0 = 0; v0 < V.nvar; v0 += 64)
   {
      double2 dk[VMAXC];
#pragma unroll
      for (int k = 0; k < VMAXC; ++k)
      {
         dk[k] = (k < nc && v0 + lane < V.nvar)
                    ? V.depsT[(size_t)pidx[k] * V.nvar + v0 + lane] : make_double2(0.0, 0.0);
      }
      const int v1 = min(v0 + 64, V.nvar);
      for (int v = v0; v < v1; ++v)
      {
         wave_sync_lds();
         for (int s = lane; s < len; s += 64) { buf[s] = row0[s]; }
         wave_sync_lds();
#pragma unroll
         for (int k = 0; k < VMAXC; ++k)
         {
            if (k < nc)
            {
               const double2 d = make_double2(lane_bcast(dk[k].x, v - v0), lane_bcast(dk[k].y, v - v0));
               const double2 t = cmul(d, m[k]);
               if (use[k])
               {
                  double2 x = buf[sl[k]];
                  x.x += t.x;
                  x.y += t.y;
                  buf[sl[k]] = x;
               }
               wave_sync_lds();
            }
         }
         for (int s = lane; s < len; s += 64) { V.out[(size_t)v * V.nnz + base + s] = buf[s]; }
      }
   }
}

// Accumulating form of k_variant_rows_reg for rows of at most 64 NS entries
// and nd <= 64.  The base row stays in registers (lane l holds entries l,
// l + 64, ...); per variant the wave writes it into its row buffer, the
// lanes of each contributor add their scaled unit-eps entries with LDS
// floating-point atomics, and the row is stored.  One contributor's columns
// are distinct slots and the contributors are issued in ascending element
// order, so every slot sums base + its contributions in the same order as
// above, without a wait between contributors.  The eps offsets are
// wave-uniform loads through the scalar cache (parameter-major copy).  Lanes
// without an entry (b >= nd, eliminated columns) add zero to a slot of their
// own past the row (64 sink entries per wave).
typedef const __attribute__((address_space(4))) double *const_dbl_ptr;

template <int NS, int NC = VMAXC>
__global__ void __launch_bounds__(BLOCK) k_variant_rows_acc(VariantParams V)
{
   extern __shared__ double2 vbuf[];
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int r = blockIdx.x * 4 + wave;
   if (r >= V.n) { return; }
   double2 *buf = vbuf + (size_t)wave * (V.maxrow + 64);
   if (V.vcopied && V.vcopied[r]) { return; }   // k_variant_copy
   const int base = V.Ip[r];
   const int len = V.Ip[r + 1] - base;
   const int c0 = V.prc_ptr[r], c1 = V.prc_ptr[r + 1];
   double2 x[NS];
#pragma unroll
   for (int c = 0; c < NS; ++c)
   {
      const int s = lane + 64 * c;
      x[c] = s < len ? V.base[base + s] : make_double2(0.0, 0.0);
   }
   if (c0 == c1 || (V.eliminate && V.is_ess[r]))
   {
      for (int v = 0; v < V.nvar; ++v)
      {
         double2 *o = V.out + (size_t)v * V.nnz + base;
#pragma unroll
         for (int c = 0; c < NS; ++c)
         {
            if (lane + 64 * c < len) { o[lane + 64 * c] = x[c]; }
         }
      }
      return;
   }
   const int nd = V.nd;
   const int nc = c1 - c0;                    // <= NC (host-checked)
   const unsigned char ess_bit = V.eliminate ? 2 : 0;
   const int b = lane;
   double2 m[NC];
   int sl[NC], pidx[NC];
#pragma unroll
   for (int k = 0; k < NC; ++k)
   {
      m[k] = make_double2(0.0, 0.0);
      sl[k] = V.maxrow + lane;
      pidx[k] = 0;
      if (k < nc)
      {
         const int pa = __builtin_amdgcn_readfirstlane(V.prc_list[c0 + k]);
         const int pi = pa / nd, a = pa - pi * nd;
         const int e = V.pelem[pi];
         const int ea = e * nd + a;
         pidx[k] = pi;
         if (b < nd)
         {
            const unsigned char fa = V.eflag[ea];
            const unsigned char fb = V.eflag[(size_t)e * nd + b];
            if ((fb & ess_bit) == 0)
            {
               sl[k] = V.slot[(size_t)ea * V.sld + b];
               const double2 y = V.Mhat[((size_t)pi * nd + a) * nd + b];
               m[k] = ((fa ^ fb) & 1) ? make_double2(-y.x, -y.y) : y;
            }
         }
      }
   }
   const const_dbl_ptr dT = (const_dbl_ptr)V.depsT;
   for (int v = 0; v < V.nvar; ++v)
   {
      wave_sync_lds();
#pragma unroll
      for (int c = 0; c < NS; ++c)
      {
         if (lane + 64 * c < len) { buf[lane + 64 * c] = x[c]; }
      }
      // contributors in groups of 4: the group's offsets are loaded together
      // (pidx = 0 past the row's contributors keeps the addresses valid)
#pragma unroll
      for (int k4 = 0; k4 < NC; k4 += 4)
      {
         if (k4 < nc)
         {
            double2 d[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
               const size_t i = 2 * ((size_t)pidx[k4 + j] * V.nvar + v);
               d[j] = make_double2(dT[i], dT[i + 1]);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
               if (k4 + j < nc)
               {
                  const double2 t = cmul(d[j], m[k4 + j]);
                  double *q = reinterpret_cast<double *>(buf + sl[k4 + j]);
                  __hip_atomic_fetch_add(q, t.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                  __hip_atomic_fetch_add(q + 1, t.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
               }
            }
         }
      }
      wave_sync_lds();
      double2 *o = V.out + (size_t)v * V.nnz + base;
#pragma unroll
      for (int c = 0; c < NS; ++c)
      {
         if (lane + 64 * c < len) { o[lane + 64 * c] = buf[lane + 64 * c]; }
      }
   }
}

__global__ void __launch_bounds__(BLOCK) k_variant_rows(VariantParams V)
{
   extern __shared__ double2 vbuf[];
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int r = blockIdx.x * 4 + wave;
   if (r >= V.n) { return; }
   double2 *row0 = vbuf + (size_t)wave * 2 * V.maxrow;
   double2 *buf = row0 + V.maxrow;
   if (V.vcopied && V.vcopied[r]) { return; }   // k_variant_copy
   const int base = V.Ip[r];
   const int len = V.Ip[r + 1] - base;
   const int c0 = V.prc_ptr[r], c1 = V.prc_ptr[r + 1];
   const bool row_fixed = c0 == c1 || (V.eliminate && V.is_ess[r]);
   if (row_fixed)
   {
      // the base row, 64 entries at a time, held in registers across the
      // variants (the stores may alias the loads for the compiler)
      for (int s0 = 0; s0 < len; s0 += 64)
      {
         const int s = s0 + lane;
         const double2 x = s < len ? V.base[base + s] : make_double2(0.0, 0.0);
         if (s < len)
         {
            for (int v = 0; v < V.nvar; ++v) { V.out[(size_t)v * V.nnz + base + s] = x; }
         }
      }
      return;
   }
   for (int s = lane; s < len; s += 64) { row0[s] = V.base[base + s]; }
   const int nd = V.nd;
   const unsigned char ess_bit = V.eliminate ? 2 : 0;
   for (int v = 0; v < V.nvar; ++v)
   {
      wave_sync_lds();
      for (int s = lane; s < len; s += 64) { buf[s] = row0[s]; }
      wave_sync_lds();
      for (int c = c0; c < c1; ++c)
      {
         const int pa = V.prc_list[c];
         const int pi = pa / nd, a = pa - pi * nd;
         const int e = V.pelem[pi];
         const int ea = e * nd + a;
         const double2 d = V.deps[(size_t)v * V.nparam + pi];
         const unsigned char fa = V.eflag[ea];
         for (int b = lane; b < nd; b += 64)
         {
            const unsigned char fb = V.eflag[(size_t)e * nd + b];
            if (fb & ess_bit) { continue; }
            const int s = V.slot[(size_t)ea * V.sld + b];
            const double2 m = V.Mhat[((size_t)pi * nd + a) * nd + b];
            double2 t = cmul(d, m);
            if ((fa ^ fb) & 1) { t.x = -t.x; t.y = -t.y; }
            double2 x = buf[s];
            x.x += t.x;
            x.y += t.y;
            buf[s] = x;
         }
         wave_sync_lds();
      }
      for (int s = lane; s < len; s += 64) { V.out[(size_t)v * V.nnz + base + s] = buf[s]; }
   }
}

// ---- angle sweeps through the k_par structure --------------------------------
//
// K(kx, ky) is linear in (kx, ky), CROSS0 + CROSS1 linear in K and KLK
// quadratic, so for one mesh and wavelength the system matrix of every
// incidence angle is
//   A(kx, ky) = B0 + kx Bx + ky By + kx^2 Bxx + ky^2 Byy + kx ky (B11 - Bxx - Byy)
// with B0 the curl-curl + eps-mass matrix, Bx / By the cross terms at
// k_par = (1, 0) / (0, 1), Bxx / Byy / B11 the KLK term at (1, 0) / (0, 1) /
// (1, 1); the host passes per angle one coefficient per basis matrix.  One
// pass over the nonzeros reads the nb basis values once and writes every
// angle's value (sums in basis order).

struct KparCombine {
   int64_t nnz;
   int nb, nang;
   const double2 *B[6];
   const double *coef;   // nang x nb
   double2 *out;         // nang x nnz
};

__global__ void __launch_bounds__(256) k_kpar_combine(KparCombine K)
{
   for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < K.nnz;
        i += (int64_t)gridDim.x * 256)
   {
      double2 b[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) { b[q] = q < K.nb ? K.B[q][i] : make_double2(0.0, 0.0); }
      for (int j = 0; j < K.nang; ++j)
      {
         const double *c = K.coef + (size_t)j * K.nb;
         double re = b[0].x * c[0], im = b[0].y * c[0];
#pragma unroll
         for (int q = 1; q < 6; ++q)
         {
            if (q < K.nb)
            {
               re = __builtin_fma(c[q], b[q].x, re);
               im = __builtin_fma(c[q], b[q].y, im);
            }
         }
         __builtin_nontemporal_store(dbl2v{re, im},
                                     reinterpret_cast<dbl2v *>(K.out + (size_t)j * K.nnz + i));
      }
   }
}

// The basis matrices' scatter and the combination in one pass: each
// wavefront forms the row of all NB element-matrix sets at once (one LDS row
// buffer per set, contributions in contributor order, as scatter_row), then
// writes every angle's row as the combination of the NB row sums (the
// arithmetic of k_kpar_combine), so neither the basis CSR values nor their
// second read reach HBM.  Same bits as scatter + k_kpar_combine +
// k_kpar_ess.  nd <= 64 (per-element bit masks), byte slot blocks.
struct MultiScatter {
   ScatterParams S;
   int nang;
   const double2 *Ae[6];
   const double *coef;   // nang x NB
   double2 *out;         // nang x nnz
   int64_t nnz;
};

template <int NB, int CG>
__global__ void __launch_bounds__(BLOCK) k_scatter_multi(MultiScatter M)
{
   extern __shared__ double2 rowbuf[];
   const ScatterParams &S = M.S;
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int r = blockIdx.x * 4 + wave;
   if (r >= S.n) { return; }
   const int ld = S.maxrow;
   double2 *buf = rowbuf + (size_t)wave * NB * ld;   // set q at buf + q * ld
   const int base = S.Ip[r];
   const int len = S.Ip[r + 1] - base;
   const int nd = S.nd;
#pragma unroll
   for (int q = 0; q < NB; ++q)
   {
      for (int s = lane; s < len; s += 64) { buf[q * ld + s] = make_double2(0.0, 0.0); }
   }
   wave_sync_lds();
   const bool row_ess = S.eliminate && S.is_ess[r];
   const unsigned ess_bit = S.eliminate ? 2u : 0u;
   if (!row_ess)
   {
      const int c0 = S.rc_ptr[r], c1 = S.rc_ptr[r + 1];
      const uint8_t *sblk = S.rslot + S.rsoff[r];
      const int b = lane;
      const int bc = min(b, nd - 1);   // lanes past the row load a valid entry, unused
      for (int cg = c0; cg < c1; cg += CG)
      {
         const int m = min(CG, c1 - cg);
         int ea[CG];
#pragma unroll
         for (int u = 0; u < CG; ++u) { ea[u] = S.rc_list[cg + min(u, m - 1)]; }
         dbl2v x[NB][CG];
#pragma unroll
         for (int q = 0; q < NB; ++q)
         {
#pragma unroll
            for (int u = 0; u < CG; ++u)
            {
               x[q][u] = __builtin_nontemporal_load(reinterpret_cast<const dbl2v *>(M.Ae[q]) +
                                                    (size_t)ea[u] * S.lda + bc);
            }
         }
         int sl[CG];
         unsigned fa[CG], fb[CG];
#pragma unroll
         for (int u = 0; u < CG; ++u)
         {
            sl[u] = sblk[(size_t)(cg - c0 + min(u, m - 1)) * nd + bc];
            const int e = ea[u] / nd;
            const uint64_t sm = S.emask[2 * (size_t)e], em = S.emask[2 * (size_t)e + 1];
            fa[u] = (unsigned)((sm >> (ea[u] - e * nd)) & 1);
            fb[u] = (unsigned)(((sm >> bc) & 1) | (((em >> bc) & 1) << 1));
         }
#pragma unroll
         for (int u = 0; u < CG; ++u)
         {
            if (u < m)
            {
               if (b < nd && (fb[u] & ess_bit) == 0)
               {
                  const bool neg = ((fa[u] ^ fb[u]) & 1) != 0;
#pragma unroll
                  for (int q = 0; q < NB; ++q)
                  {
                     double2 t = buf[q * ld + sl[u]];
                     t.x += neg ? -x[q][u].x : x[q][u].x;
                     t.y += neg ? -x[q][u].y : x[q][u].y;
                     buf[q * ld + sl[u]] = t;
                  }
               }
               wave_sync_lds();
            }
         }
      }
   }
   for (int j = 0; j < M.nang; ++j)
   {
      const double *c = M.coef + (size_t)j * NB;
      for (int s = lane; s < len; s += 64)
      {
         double re, im;
         if (row_ess)
         {
            re = S.Jc[base + s] == r ? 1.0 : 0.0;
            im = 0.0;
         }
         else
         {
            re = buf[s].x * c[0];
            im = buf[s].y * c[0];
#pragma unroll
            for (int q = 1; q < NB; ++q)
            {
               re = __builtin_fma(c[q], buf[q * ld + s].x, re);
               im = __builtin_fma(c[q], buf[q * ld + s].y, im);
            }
         }
         __builtin_nontemporal_store(dbl2v{re, im},
                                     reinterpret_cast<dbl2v *>(M.out + (size_t)j * M.nnz + base + s));
      }
   }
}

// Eliminated rows: every basis matrix holds the identity row there, so the
// combination is replaced by B0's row (0 ... 1 ... 0) in every angle.
__global__ void k_kpar_ess(int n, const int *__restrict__ Ip, const unsigned char *__restrict__ ess,
                           const double2 *__restrict__ B0, double2 *__restrict__ out, int64_t nnz,
                           int nang)
{
   for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
   {
      if (!ess[r]) { continue; }
      for (int s = Ip[r]; s < Ip[r + 1]; ++s)
      {
         for (int j = 0; j < nang; ++j) { out[(size_t)j * nnz + s] = B0[s]; }
      }
   }
}

// ---- incident field ---------------------------------------------------------------

// One slab wave (the reference's SlabWave, solver/problem_common.inc:11-16):
// E(z) = A exp(i qA (z - zA)) + B exp(i qB (z - zB)).
struct SlabWaveDev {
   double zA, zB;
   double2 qA, qB;
   double2 A[3], B[3];
};

__device__ __forceinline__ double2 cexpi(double2 q, double dz)
{
   // exp(i q dz) = exp(-Im(q) dz) (cos(Re(q) dz) + i sin(Re(q) dz))
   const double mag = exp(-q.y * dz);
   double sn, cs;
   sincos(q.x * dz, &sn, &cs);
   return make_double2(mag * cs, mag * sn);
}

__device__ __forceinline__ void incident_at(const SlabWaveDev &w, double z, double2 E[3])
{
   const double2 up = cexpi(w.qA, z - w.zA);
   const double2 dn = cexpi(w.qB, z - w.zB);
#pragma unroll
   for (int i = 0; i < 3; ++i)
   {
      const double2 a = cmul(w.A[i], up), b = cmul(w.B[i], dn);
      E[i] = make_double2(a.x + b.x, a.y + b.y);
   }
}

struct RhsParams {
   int nd, p, nrhs, nslab;
   int nsrc;                      // source elements (list below)
   const int *src_elem;           // source element ids
   const int *attr, *pid, *eps_off;
   const double *J, *Ji, *detJ, *v0;
   const int *fo;
   RuleDev rule;                  // quadrature rule of the linear form
   const PmlDev *pml_tab;         // k0 per problem
   const int *region_slab;        // per region (concatenated per problem)
   const double2 *region_deps;    // per region: eps - eps_bg (source regions)
   const int *region_src;         // per region: 1 for scattered-field source regions
   const double2 *eps_tab;        // region eps (concatenated per problem)
   const double2 *eps_elem;       // optional per-element eps
   const SlabWaveDev *waves;      // [pid][rhs][slab]
   double TInv[24];
   double2 *bloc;                 // nsrc x nrhs x nd
};

// One wavefront per source element: the fields at the rule points are staged
// in LDS, each lane then integrates one basis function against every
// incident field, and the dual face transform is applied to the vector.
__global__ void __launch_bounds__(64) k_rhs_local(RhsParams R)
{
   extern __shared__ double2 efield[];        // nq x nrhs x 3
   const int fi = blockIdx.x;
   const int lane = threadIdx.x;
   const int e = R.src_elem[fi];
   const int pid = R.pid ? R.pid[e] : 0;
   const int reg = R.eps_off[pid] + R.attr[e] - 1;
   const int slab = R.region_slab[reg];
   // source strength eps_e - eps_bg (eps_bg = region eps - region deps,
   // region deps = 0 outside source regions); without per-element eps this
   // is the region's eps - eps_bg itself
   double2 de = R.region_src[reg] ? R.region_deps[reg] : make_double2(0.0, 0.0);
   if (R.eps_elem)
   {
      const double2 ee = R.eps_elem[e], er = R.eps_tab[reg];
      de = make_double2(ee.x - (er.x - de.x), ee.y - (er.y - de.y));
   }
   const double k2 = R.pml_tab[pid].k0 * R.pml_tab[pid].k0;
   const double det = R.detJ[e];
   const double *J = R.J + (size_t)e * 9;
   const double *Ji = R.Ji + (size_t)e * 9;
   const double v0z = R.v0[(size_t)e * 3 + 2];
   const int nq = R.rule.nq, nrhs = R.nrhs;
   // f(z_q) = w det k0^2 deps E_inc(z_q) for every (q, rhs)
   for (int it = lane; it < nq * nrhs; it += 64)
   {
      const int q = it / nrhs, r = it - q * nrhs;
      const double *xi = R.rule.ip + (size_t)q * 3;
      const double z = v0z + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
      double2 E[3];
      incident_at(R.waves[((size_t)pid * nrhs + r) * R.nslab + slab], z, E);
      const double2 f = make_double2(R.rule.w[q] * det * k2 * de.x, R.rule.w[q] * det * k2 * de.y);
#pragma unroll
      for (int i = 0; i < 3; ++i) { efield[(size_t)it * 3 + i] = cmul(f, E[i]); }
   }
   __syncthreads();
   const int nd = R.nd, p = R.p;
   for (int k0 = 0; k0 < nd; k0 += 64)
   {
      const int k = k0 + lane;
      const bool valid = k < nd;
      for (int r = 0; r < nrhs; ++r)
      {
         double2 acc = make_double2(0.0, 0.0);
         if (valid)
         {
            for (int q = 0; q < nq; ++q)
            {
               const double *vs = R.rule.vs + ((size_t)q * nd + k) * 3;
               const double a0 = vs[0], a1 = vs[1], a2 = vs[2];
               const double v[3] = {a0 * Ji[0] + a1 * Ji[3] + a2 * Ji[6],
                                    a0 * Ji[1] + a1 * Ji[4] + a2 * Ji[7],
                                    a0 * Ji[2] + a1 * Ji[5] + a2 * Ji[8]};
               const double2 *f = efield + ((size_t)q * nrhs + r) * 3;
#pragma unroll
               for (int i = 0; i < 3; ++i)
               {
                  acc.x += v[i] * f[i].x;
                  acc.y += v[i] * f[i].y;
               }
            }
         }
         // dual face transform of the element vector (pairs in lanes k, k^1)
         if (p >= 2)
         {
            const int foff = 6 * p, ntd = p * (p - 1);
            double cs = 1.0, cp = 0.0;
            if (valid && k >= foff && k < foff + 4 * ntd)
            {
               const double *Ti = R.TInv + 4 * R.fo[(size_t)e * 4 + (k - foff) / ntd];
               const bool even = ((k - foff) & 1) == 0;
               cs = even ? Ti[0] : Ti[3];
               cp = even ? Ti[1] : Ti[2];
            }
            const double yx = __shfl_xor(acc.x, 1), yy = __shfl_xor(acc.y, 1);
            acc = make_double2(cs * acc.x + cp * yx, cs * acc.y + cp * yy);
         }
         if (valid) { R.bloc[((size_t)fi * nrhs + r) * nd + k] = acc; }
      }
   }
}

// b[rhs][row] = sum over the row's contributors (ascending element) of
// sign * bloc; essential rows are zero with `eliminate`.
__global__ void k_rhs_rows(int n, int nd, int nrhs, const int *__restrict__ rc_ptr,
                           const int *__restrict__ rc_list, const int *__restrict__ src_index,
                           const unsigned char *__restrict__ eflag,
                           const unsigned char *__restrict__ is_ess, int eliminate,
                           const double2 *__restrict__ bloc, double2 *__restrict__ b)
{
   const int r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= n) { return; }
   const bool zero = eliminate && is_ess[r];
   for (int q = 0; q < nrhs; ++q)
   {
      double2 acc = make_double2(0.0, 0.0);
      if (!zero)
      {
         for (int c = rc_ptr[r]; c < rc_ptr[r + 1]; ++c)
         {
            const int ea = rc_list[c];
            const int e = ea / nd, a = ea - e * nd;
            const int fi = src_index[e];
            if (fi < 0) { continue; }
            const double2 v = bloc[((size_t)fi * nrhs + q) * nd + a];
            if (eflag[ea] & 1) { acc.x -= v.x; acc.y -= v.y; }
            else { acc.x += v.x; acc.y += v.y; }
         }
      }
      b[(size_t)q * n + r] = acc;
   }
}

// ---- eps sensitivity (adjoint) --------------------------------------------------

struct SensParams {
   int nd, p, nparam, nrhs, nslab, rhs_col;
   const int *pelem;
   const int *dof, *attr, *pid, *eps_off, *fo;
   const unsigned char *eflag;
   const double *J, *Ji, *detJ, *v0;
   RuleDev rm;                    // mass rule
   const PmlDev *pml_tab;
   const int *region_slab;        // may be NULL when rhs_col < 0
   const int *region_src;         // 1 for scattered-field source regions
   const SlabWaveDev *waves;
   double TInv[24];
   const double2 *lam, *u;        // global adjoint / forward vectors
   double2 *s;                    // nparam
};

// One wavefront per parameter element: gather the local vectors (sign and
// face-pair transform), then lane q evaluates L(x_q), U(x_q) and the
// integrand at point q; a wave reduction sums the points.
__global__ void __launch_bounds__(64) k_eps_sens(SensParams S)
{
   extern __shared__ double2 loc[];            // 2 x nd
   const int pi = blockIdx.x;
   const int lane = threadIdx.x;
   const int e = S.pelem[pi];
   const int nd = S.nd, p = S.p;
   double2 *lam = loc, *uu = loc + nd;
   for (int k0 = 0; k0 < nd; k0 += 64)
   {
      const int k = k0 + lane;
      const bool valid = k < nd;
      double2 l = make_double2(0.0, 0.0), u = make_double2(0.0, 0.0);
      if (valid)
      {
         const int d = S.dof[(size_t)e * nd + k];
         const double sg = (S.eflag[(size_t)e * nd + k] & 1) ? -1.0 : 1.0;
         l = S.lam[d];
         u = S.u[d];
         l.x *= sg; l.y *= sg;
         u.x *= sg; u.y *= sg;
      }
      if (p >= 2)
      {
         // local pair = M g, M = [[t0, t2], [t1, t3]] (TInv column-major)
         const int foff = 6 * p, ntd = p * (p - 1);
         double cs = 1.0, cp = 0.0;
         if (valid && k >= foff && k < foff + 4 * ntd)
         {
            const double *Ti = S.TInv + 4 * S.fo[(size_t)e * 4 + (k - foff) / ntd];
            const bool even = ((k - foff) & 1) == 0;
            cs = even ? Ti[0] : Ti[3];
            cp = even ? Ti[2] : Ti[1];
         }
         const double lx = __shfl_xor(l.x, 1), ly = __shfl_xor(l.y, 1);
         const double ux = __shfl_xor(u.x, 1), uy = __shfl_xor(u.y, 1);
         l = make_double2(cs * l.x + cp * lx, cs * l.y + cp * ly);
         u = make_double2(cs * u.x + cp * ux, cs * u.y + cp * uy);
      }
      if (valid)
      {
         lam[k] = l;
         uu[k] = u;
      }
   }
   __syncthreads();
   const int pid = S.pid ? S.pid[e] : 0;
   const PmlDev pm = S.pml_tab[pid];
   const int reg = S.eps_off[pid] + S.attr[e] - 1;
   // eps_e enters the element's source strength eps_e - eps_bg linearly in
   // any region, so db/d eps_e carries the incident field for every
   // parameter element
   const bool src = S.rhs_col >= 0;
   const double det = S.detJ[e];
   const double *J = S.J + (size_t)e * 9;
   const double *Ji = S.Ji + (size_t)e * 9;
   const double v0z = S.v0[(size_t)e * 3 + 2];
   double2 acc = make_double2(0.0, 0.0);
   for (int q = lane; q < S.rm.nq; q += 64)
   {
      double2 L[3] = {}, U[3] = {};
      for (int k = 0; k < nd; ++k)
      {
         const double *vs = S.rm.vs + ((size_t)q * nd + k) * 3;
         const double a0 = vs[0], a1 = vs[1], a2 = vs[2];
         const double v[3] = {a0 * Ji[0] + a1 * Ji[3] + a2 * Ji[6],
                              a0 * Ji[1] + a1 * Ji[4] + a2 * Ji[7],
                              a0 * Ji[2] + a1 * Ji[5] + a2 * Ji[8]};
         const double2 lk = lam[k], uk = uu[k];
#pragma unroll
         for (int i = 0; i < 3; ++i)
         {
            L[i].x += v[i] * lk.x; L[i].y += v[i] * lk.y;
            U[i].x += v[i] * uk.x; U[i].y += v[i] * uk.y;
         }
      }
      const double *xi = S.rm.ip + (size_t)q * 3;
      const double z = v0z + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
      const double2 s = pml_s(pm, z);
      const double2 lamd[3] = {s, s, cinv(s)};   // Lam = diag(s, s, 1/s)
      double2 E[3] = {};
      if (src)
      {
         incident_at(S.waves[((size_t)pid * S.nrhs + S.rhs_col) * S.nslab + S.region_slab[reg]], z, E);
      }
      const double w = S.rm.w[q] * det;
      double2 t = make_double2(0.0, 0.0);
#pragma unroll
      for (int i = 0; i < 3; ++i)
      {
         const double2 lu = cmul(lamd[i], U[i]);
         const double2 f = make_double2(lu.x + E[i].x, lu.y + E[i].y);
         const double2 g = cmul(L[i], f);
         t.x += g.x;
         t.y += g.y;
      }
      acc.x += w * t.x;
      acc.y += w * t.y;
   }
   // wave reduction over the points
   for (int off = 32; off > 0; off >>= 1)
   {
      acc.x += __shfl_xor(acc.x, off);
      acc.y += __shfl_xor(acc.y, off);
   }
   if (lane == 0)
   {
      const double k2 = pm.k0 * pm.k0;
      S.s[pi] = make_double2(-k2 * acc.x, -k2 * acc.y);
   }
}

}  // namespace lfasm
