// Device CSR symbolic phase: the sorted-column pattern of the union of the
// element dof x dof graphs (what solver/asm_data.inc:178-210 builds on the
// host, serially), from the per-row contributor lists.
//
// One wavefront per row: the columns dof[e][b] of every contributor (e, a)
// of the row are gathered into LDS (padded to a power of two with INT_MAX),
// sorted by a bitonic network, and the distinct values are counted (pass 0)
// or written in ascending order (pass 1, after an exclusive scan of the
// counts gave the row pointers).
#pragma once

#include "lfasm_device.hpp"

#include <climits>

namespace lfasm {

constexpr int PAT_WAVES = 4;   // rows per workgroup

__global__ void __launch_bounds__(64 * PAT_WAVES)
k_csr_rows(int n, int nd, const int *__restrict__ dof, const int *__restrict__ rc_ptr,
           const int *__restrict__ rc_list, int cap, int pass, int *__restrict__ cnt_or_ptr,
           int *__restrict__ Jcol)
{
   extern __shared__ int keys_all[];
   const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
   const int lane = threadIdx.x & 63;
   const int r = blockIdx.x * PAT_WAVES + wave;
   if (r >= n) { return; }
   int *keys = keys_all + (size_t)wave * cap;
   const int c0 = rc_ptr[r], c1 = rc_ptr[r + 1];
   const int m = (c1 - c0) * nd;               // candidates
   int N = 64;
   while (N < m) { N <<= 1; }                  // host guarantees N <= cap
   for (int i = lane; i < N; i += 64)
   {
      int v = INT_MAX;
      if (i < m)
      {
         const int c = c0 + i / nd, b = i - (i / nd) * nd;
         const int e = rc_list[c] / nd;
         v = dof[(size_t)e * nd + b];
      }
      keys[i] = v;
   }
   wave_sync_lds();
   for (int k = 2; k <= N; k <<= 1)
   {
      for (int j = k >> 1; j > 0; j >>= 1)
      {
         for (int i = lane; i < N; i += 64)
         {
            const int l = i ^ j;
            if (l > i)
            {
               const int a = keys[i], b = keys[l];
               const bool up = (i & k) == 0;
               if ((a > b) == up)
               {
                  keys[i] = b;
                  keys[l] = a;
               }
            }
         }
         wave_sync_lds();
      }
   }
   // distinct values in ascending order
   int base = pass ? cnt_or_ptr[r] : 0;
   int count = 0;
   for (int i0 = 0; i0 < N; i0 += 64)
   {
      const int i = i0 + lane;
      const int v = keys[i];
      const bool head = v != INT_MAX && (i == 0 || keys[i - 1] != v);
      const unsigned long long mask = __ballot(head);
      if (pass && head)
      {
         const unsigned long long below = mask & ((1ull << lane) - 1ull);
         Jcol[base + count + __popcll(below)] = v;
      }
      count += __popcll(mask);
   }
   if (!pass && lane == 0) { cnt_or_ptr[r] = count; }
}

}  // namespace lfasm
