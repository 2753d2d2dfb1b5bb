// On-device peak probes: FP64 MFMA throughput (v_mfma_f64_16x16x4f64),
// FP64 VALU FMA throughput, and HBM copy bandwidth.  Used to report
// measured ceilings next to the published ones.
#pragma once

#include <hip/hip_runtime.h>

namespace lfasm {

typedef double pdbl4 __attribute__((ext_vector_type(4)));

// Each wavefront runs `iters` rounds of 8 independent MFMA chains.
__global__ void __launch_bounds__(256) k_peak_mfma(int iters, double *out)
{
   const int lane = threadIdx.x & 63;
   const double a = 1.0 + 1e-9 * lane, b = 1.0 - 1e-9 * lane;
   pdbl4 acc[8];
#pragma unroll
   for (int i = 0; i < 8; ++i) { acc[i] = pdbl4{0.0, 0.0, 0.0, 0.0}; }
   for (int it = 0; it < iters; ++it)
   {
#pragma unroll
      for (int i = 0; i < 8; ++i)
      {
         acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
      }
   }
   double s = 0.0;
#pragma unroll
   for (int i = 0; i < 8; ++i) { s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3]; }
   if (s == -1.0) { out[blockIdx.x] = s; }   // never true; keeps the work live
}

// Each thread runs `iters` rounds of 8 independent FMA chains.
__global__ void __launch_bounds__(256) k_peak_valu(int iters, double y, double *out)
{
   double x[8];
#pragma unroll
   for (int i = 0; i < 8; ++i) { x[i] = 1.0 + i * 1e-3 + threadIdx.x * 1e-9; }
   for (int it = 0; it < iters; ++it)
   {
#pragma unroll
      for (int i = 0; i < 8; ++i) { x[i] = __fma_rn(x[i], y, 1e-12); }
   }
   double s = 0.0;
#pragma unroll
   for (int i = 0; i < 8; ++i) { s += x[i]; }
   if (s == -1.0) { out[blockIdx.x] = s; }
}

__global__ void __launch_bounds__(256) k_peak_copy(const double2 *__restrict__ src,
                                                   double2 *__restrict__ dst, size_t n)
{
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (size_t)gridDim.x * blockDim.x)
   {
      dst[i] = src[i];
   }
}

// Write-only and read-only streams of double2, four per thread per
// iteration (the write pattern of the element-matrix kernels, the read
// side of the scatter).
__global__ void __launch_bounds__(256) k_peak_write(double2 *__restrict__ dst, size_t n)
{
   const size_t stride = (size_t)gridDim.x * blockDim.x;
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 4 * stride)
   {
#pragma unroll
      for (int u = 0; u < 4; ++u)
      {
         const size_t j = i + u * stride;
         if (j < n) { dst[j] = make_double2((double)j, 0.0); }
      }
   }
}

__global__ void __launch_bounds__(256) k_peak_read(const double2 *__restrict__ src, size_t n,
                                                   double *__restrict__ out)
{
   const size_t stride = (size_t)gridDim.x * blockDim.x;
   double acc = 0.0;
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 4 * stride)
   {
#pragma unroll
      for (int u = 0; u < 4; ++u)
      {
         const size_t j = i + u * stride;
         if (j < n) { const double2 v = src[j]; acc += v.x + v.y; }
      }
   }
   if (acc == 12345.678) { out[0] = acc; }   // keeps the loads; never true for zeros
}

// Launch-shape and store-mode variants of the streams above (U accesses
// per thread per iteration, ordinary or non-temporal): the ceiling taken is
// the best of them, since which pattern reaches the most depends on the
// access (measured on MI355X: copy best with one access per thread and a
// 1,024-block grid, writes and reads best non-temporal).
typedef double pk_dv2 __attribute__((ext_vector_type(2)));

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_peak_copy_v(const pk_dv2 *__restrict__ s,
                                                     pk_dv2 *__restrict__ d, size_t n)
{
   const size_t st = (size_t)gridDim.x * blockDim.x;
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += U * st)
   {
      pk_dv2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
         const size_t j = i + u * st;
         v[u] = j < n ? (NT ? __builtin_nontemporal_load(s + j) : s[j]) : pk_dv2{0.0, 0.0};
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
         const size_t j = i + u * st;
         if (j < n)
         {
            if (NT) { __builtin_nontemporal_store(v[u], d + j); }
            else { d[j] = v[u]; }
         }
      }
   }
}

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_peak_write_v(pk_dv2 *__restrict__ d, size_t n)
{
   const size_t st = (size_t)gridDim.x * blockDim.x;
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += U * st)
   {
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
         const size_t j = i + u * st;
         if (j < n)
         {
            const pk_dv2 v = {(double)j, 0.0};
            if (NT) { __builtin_nontemporal_store(v, d + j); }
            else { d[j] = v; }
         }
      }
   }
}

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_peak_read_v(const pk_dv2 *__restrict__ s, size_t n,
                                                     double *__restrict__ out)
{
   const size_t st = (size_t)gridDim.x * blockDim.x;
   double acc = 0.0;
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += U * st)
   {
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
         const size_t j = i + u * st;
         if (j < n)
         {
            const pk_dv2 v = NT ? __builtin_nontemporal_load(s + j) : s[j];
            acc += v.x + v.y;
         }
      }
   }
   if (acc == 12345.678) { out[0] = acc; }
}

}  // namespace lfasm
