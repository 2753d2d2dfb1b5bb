// Device-side extraction (the host loop of solver/asm_data.inc:99-173 and the
// per-row contributor lists of the deterministic scatter):
//
// * affine geometry of every element from its four vertex coordinates (the
//   element's own vertex order, unwrapped across periodic boundaries -- the
//   points the reference's element transformation maps the reference
//   vertices to): J = [x1 - x0, x2 - x0, x3 - x0] (row-major), detJ by
//   cofactors of the first row, J^-1 = adj(J) / detJ, v0 = x0
//   (asm_data.inc:152-171 reads the same from MFEM's ElementTransformation);
// * per-row contributor lists (element e, local index a) in ascending
//   element order, by a device counting sort of the element dof table;
// * per-(element, local dof) sign / essential flags and their bit masks.
//
// The geometry is evaluated without floating-point contraction so that the
// oracle's plain-C twin (orc_geometry) reproduces it bit for bit.
#pragma once

#include "lfasm_device.hpp"

namespace lfasm {

__global__ void k_geometry(int ne, const double *__restrict__ xyz, double *__restrict__ J,
                           double *__restrict__ Ji, double *__restrict__ detJ,
                           double *__restrict__ v0, int *__restrict__ bad_first)
{
#pragma clang fp contract(off)
   for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x)
   {
      const double *x = xyz + (size_t)e * 12;
      double a[9];
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) { a[i * 3 + j] = x[(j + 1) * 3 + i] - x[i]; }
      // cofactors of the first row, then the adjugate
      const double c00 = a[4] * a[8] - a[5] * a[7];
      const double c01 = a[5] * a[6] - a[3] * a[8];
      const double c02 = a[3] * a[7] - a[4] * a[6];
      const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
      double inv[9];
      inv[0] = c00 / det;
      inv[3] = c01 / det;
      inv[6] = c02 / det;
      inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
      inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
      inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
      inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
      inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
      inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
      for (int i = 0; i < 9; ++i)
      {
         J[(size_t)e * 9 + i] = a[i];
         Ji[(size_t)e * 9 + i] = inv[i];
      }
      detJ[e] = det;
      for (int d = 0; d < 3; ++d) { v0[(size_t)e * 3 + d] = x[d]; }
      if (!(det > 0.0)) { atomicMin(bad_first, e); }
   }
}

// number of contributors of every row (atomic histogram of the dof table)
__global__ void k_contrib_count(int64_t nk, const int *__restrict__ dof, int *__restrict__ cnt)
{
   for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk;
        i += (int64_t)gridDim.x * blockDim.x)
   {
      atomicAdd(cnt + dof[i], 1);
   }
}

// place every (element, local index) = i into its row's segment (arbitrary
// order inside the segment; k_contrib_sort restores ascending order)
__global__ void k_contrib_fill(int64_t nk, const int *__restrict__ dof,
                               const int *__restrict__ rc_ptr, int *__restrict__ fill,
                               int *__restrict__ rc_list)
{
   for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk;
        i += (int64_t)gridDim.x * blockDim.x)
   {
      const int r = dof[i];
      const int pos = atomicAdd(fill + r, 1);
      rc_list[rc_ptr[r] + pos] = (int)i;
   }
}

// ascending order inside each row's segment (insertion sort; a row holds a
// few tens of contributors), and the largest segment
__global__ void k_contrib_sort(int n, const int *__restrict__ rc_ptr, int *__restrict__ rc_list,
                               int *__restrict__ maxc)
{
   for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
   {
      const int c0 = rc_ptr[r], c1 = rc_ptr[r + 1];
      for (int c = c0 + 1; c < c1; ++c)
      {
         const int v = rc_list[c];
         int j = c - 1;
         while (j >= c0 && rc_list[j] > v)
         {
            rc_list[j + 1] = rc_list[j];
            --j;
         }
         rc_list[j + 1] = v;
      }
      atomicMax(maxc, c1 - c0);
   }
}

// bit0: sign < 0, bit1: the dof is essential
__global__ void k_eflag(int64_t nk, const double *__restrict__ sign, const int *__restrict__ dof,
                        const unsigned char *__restrict__ ess, unsigned char *__restrict__ eflag)
{
   for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk;
        i += (int64_t)gridDim.x * blockDim.x)
   {
      eflag[i] = (unsigned char)((sign[i] < 0.0 ? 1 : 0) | (ess[dof[i]] ? 2 : 0));
   }
}

// nd <= 64: per element the sign and essential bits as two 64-bit masks
// (one wavefront per element, lane b holds local dof b)
__global__ void k_emask(int ne, int nd, const unsigned char *__restrict__ eflag,
                        uint64_t *__restrict__ emask)
{
   const int lane = threadIdx.x & 63;
   const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
   const int nw = (gridDim.x * blockDim.x) >> 6;
   for (int e = w; e < ne; e += nw)
   {
      const unsigned char f = lane < nd ? eflag[(size_t)e * nd + lane] : 0;
      const uint64_t s = __ballot(f & 1), es = __ballot(f & 2);
      if (lane == 0)
      {
         emask[2 * (size_t)e] = s;
         emask[2 * (size_t)e + 1] = es;
      }
   }
}

// row r's slot-block length: its contributor count x nd
__global__ void k_slot_block_len(int n, int nd, const int *__restrict__ rc_ptr,
                                 int64_t *__restrict__ len)
{
   for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
   {
      len[r] = (int64_t)(rc_ptr[r + 1] - rc_ptr[r]) * nd;
   }
}

__global__ void k_widen(int n, const int *__restrict__ src, int64_t *__restrict__ dst)
{
   for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
   {
      dst[i] = src[i];
   }
}

__global__ void k_narrow(int64_t n, const int64_t *__restrict__ src, int *__restrict__ dst)
{
   for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (int64_t)gridDim.x * blockDim.x)
   {
      dst[i] = (int)src[i];
   }
}

}  // namespace lfasm
