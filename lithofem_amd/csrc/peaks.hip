// Device ceilings for the benchmark's roofline context (bench.py --mode
// peaks): FP64 MFMA (v_mfma_f64_16x16x4f64) and FP64 VALU FMA throughput over
// launch shapes, the shader clock under each load (clock64 against the
// constant-rate wall clock), and HBM copy / write / read streams.  Built as
// its own library (lithofem_amd/libpeaks.so); the assembly library does not
// contain it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace lfpeaks {

typedef double pdbl4 __attribute__((ext_vector_type(4)));

// CHAINS independent MFMA accumulators per wave, `iters` rounds; lane 0 of
// each block's first wave records clock64 / wall_clock64 at start and end.
template <int CHAINS>
__global__ void k_peak_mfma(int iters, double *__restrict__ sink,
                            long long *__restrict__ clk)
{
   const int lane = threadIdx.x & 63;
   const double a = 1.0 + 1e-9 * lane, b = 1.0 - 1e-9 * lane;
   pdbl4 acc[CHAINS];
#pragma unroll
   for (int i = 0; i < CHAINS; ++i) { acc[i] = pdbl4{0.0, 0.0, 0.0, 0.0}; }
   const long long c0 = clock64(), w0 = wall_clock64();
   for (int it = 0; it < iters; ++it)
   {
#pragma unroll
      for (int i = 0; i < CHAINS; ++i)
      {
         acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
      }
   }
   double s = 0.0;
#pragma unroll
   for (int i = 0; i < CHAINS; ++i) { s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3]; }
   const long long c1 = clock64(), w1 = wall_clock64();
   if (threadIdx.x == 0)
   {
      clk[4 * blockIdx.x + 0] = c0;
      clk[4 * blockIdx.x + 1] = c1;
      clk[4 * blockIdx.x + 2] = w0;
      clk[4 * blockIdx.x + 3] = w1;
   }
   if (s == -1.0) { sink[blockIdx.x] = s; }   // never true; keeps the work live
}

// 8 independent FMA chains per thread, `iters` rounds.
__global__ void k_peak_valu(int iters, double y, double *__restrict__ sink,
                            long long *__restrict__ clk)
{
   double x[8];
#pragma unroll
   for (int i = 0; i < 8; ++i) { x[i] = 1.0 + i * 1e-3 + threadIdx.x * 1e-9; }
   const long long c0 = clock64(), w0 = wall_clock64();
   for (int it = 0; it < iters; ++it)
   {
#pragma unroll
      for (int i = 0; i < 8; ++i) { x[i] = __fma_rn(x[i], y, 1e-12); }
   }
   double s = 0.0;
#pragma unroll
   for (int i = 0; i < 8; ++i) { s += x[i]; }
   const long long c1 = clock64(), w1 = wall_clock64();
   if (threadIdx.x == 0)
   {
      clk[4 * blockIdx.x + 0] = c0;
      clk[4 * blockIdx.x + 1] = c1;
      clk[4 * blockIdx.x + 2] = w0;
      clk[4 * blockIdx.x + 3] = w1;
   }
   if (s == -1.0) { sink[blockIdx.x] = s; }
}

// Does FP64 VALU work overlap FP64 MFMA work?  Each wave runs `iters`
// rounds of NM MFMAs on 4 accumulator chains followed by NV independent FMAs
// on 8 chains; with the two kinds overlapping, the time of the mix is the
// larger of the two parts, with a shared pipe it is their sum.
template <int NM, int NV>
__global__ void __launch_bounds__(256) k_peak_mix(int iters, double y, double *__restrict__ sink)
{
   const int lane = threadIdx.x & 63;
   const double a = 1.0 + 1e-9 * lane, b = 1.0 - 1e-9 * lane;
   pdbl4 acc[4];
#pragma unroll
   for (int i = 0; i < 4; ++i) { acc[i] = pdbl4{0.0, 0.0, 0.0, 0.0}; }
   double x[8];
#pragma unroll
   for (int i = 0; i < 8; ++i) { x[i] = 1.0 + i * 1e-3 + threadIdx.x * 1e-9; }
   for (int it = 0; it < iters; ++it)
   {
#pragma unroll
      for (int m = 0; m < NM; ++m) { acc[m & 3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[m & 3], 0, 0, 0); }
#pragma unroll
      for (int v = 0; v < NV; ++v) { x[v & 7] = __fma_rn(x[v & 7], y, 1e-12); }
   }
   double s = 0.0;
#pragma unroll
   for (int i = 0; i < 4; ++i) { s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3]; }
#pragma unroll
   for (int i = 0; i < 8; ++i) { s += x[i]; }
   if (s == -1.0) { sink[blockIdx.x] = s; }
}

__global__ void k_fill_zero(double2 *p, size_t n)
{
   for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (size_t)gridDim.x * blockDim.x)
   {
      p[i] = make_double2(0.0, 0.0);
   }
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

}  // namespace lfpeaks

namespace {

struct Ctx
{
   hipStream_t s = nullptr;
   hipEvent_t e0 = nullptr, e1 = nullptr;
   double wall_hz = 0.0;
   bool ok = true;

   bool check(hipError_t e)
   {
      if (e != hipSuccess) { ok = false; }
      return ok;
   }

   // elapsed ms of the second of two launches
   template <class F>
   float time2(F launch)
   {
      float ms = 0.0f;
      launch();
      check(hipEventRecord(e0, s));
      launch();
      check(hipEventRecord(e1, s));
      if (check(hipEventSynchronize(e1))) { check(hipEventElapsedTime(&ms, e0, e1)); }
      return ms;
   }

   // mean shader clock (GHz) of the blocks of the last launch
   double clock_ghz(const long long *clk_dev, int nblk)
   {
      std::vector<long long> c((size_t)4 * nblk);
      if (!check(hipMemcpy(c.data(), clk_dev, c.size() * sizeof(long long), hipMemcpyDeviceToHost)))
      {
         return 0.0;
      }
      double sum = 0.0;
      int cnt = 0;
      for (int b = 0; b < nblk; ++b)
      {
         const double dc = (double)(c[4 * b + 1] - c[4 * b]), dw = (double)(c[4 * b + 3] - c[4 * b + 2]);
         if (dw > 0) { sum += dc / dw * wall_hz / 1e9; ++cnt; }
      }
      return cnt ? sum / cnt : 0.0;
   }
};

template <int CHAINS>
void mfma_case(Ctx &x, int ncu, int bpc, double *sink, long long *clk, double *row)
{
   const int grid = ncu * bpc;
   const double waves = (double)grid * 4;
   const int iters = (int)std::max(100.0, 3e11 / (waves * CHAINS * 2048.0));
   const float ms = x.time2([&] {
      hipLaunchKernelGGL(lfpeaks::k_peak_mfma<CHAINS>, dim3(grid), dim3(256), 0, x.s, iters, sink, clk);
   });
   row[0] = CHAINS;
   row[1] = bpc;
   row[2] = ms > 0 ? waves * iters * CHAINS * 2048.0 / (ms * 1e-3) / 1e12 : 0.0;
   row[3] = x.clock_ghz(clk, grid);
}

}  // namespace

extern "C" {

/* out (nout >= 11 + 4 * 16 + 3 doubles):
 *  [0] best FP64 MFMA TFLOP/s, [1] shader clock during it (GHz),
 *  [2] its chains per wave, [3] its waves per SIMD,
 *  [4] FP64 VALU FMA TFLOP/s, [5] shader clock during it (GHz),
 *  [6] HBM copy GB/s (read + write bytes), [7] write-only GB/s, [8] read-only GB/s,
 *  [9] compute units, [10] reported maximum clock (GHz),
 *  [11 + 4 i ...] the MFMA sweep: chains, waves per SIMD, TFLOP/s, clock GHz
 *  (i < 16), then [11 + 64 ..] the overlap probe: ms of MFMAs alone, FMAs
 *  alone, both interleaved.
 * Each stream figure is the best over launch shapes and ordinary /
 * non-temporal accesses of 2 GiB buffers.  Returns the number of sweep rows
 * or a negative HIP error. */
int lfpeaks_measure(int device, double *out, int nout)
{
   if (!out || nout < 11 + 4 * 16 + 3) { return -1; }
   for (int i = 0; i < nout; ++i) { out[i] = 0.0; }
   if (hipSetDevice(device) != hipSuccess) { return -2; }
   int ncu = 0, wall_khz = 0, clk_khz = 0;
   if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
       hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
       hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, device) != hipSuccess)
   {
      return -2;
   }
   Ctx x;
   x.wall_hz = wall_khz * 1e3;
   double *sink = nullptr;
   long long *clk = nullptr;
   double2 *a = nullptr, *b = nullptr;
   const size_t nbytes = (size_t)1 << 31;
   const size_t n2 = nbytes / sizeof(double2);
   int rows = 0;
   if (x.check(hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking)) && x.check(hipEventCreate(&x.e0)) &&
       x.check(hipEventCreate(&x.e1)) && x.check(hipMalloc((void **)&sink, 65536 * sizeof(double))) &&
       x.check(hipMalloc((void **)&clk, 4 * 65536 * sizeof(long long))) &&
       x.check(hipMalloc((void **)&a, nbytes)) && x.check(hipMalloc((void **)&b, nbytes)))
   {
      out[9] = ncu;
      out[10] = clk_khz / 1e6;
      // FP64 MFMA: chains per wave x waves per SIMD (blocks of 4 waves per CU)
      for (int bpc : {1, 2, 4, 8})
      {
         double *r = out + 11 + 4 * rows;
         mfma_case<4>(x, ncu, bpc, sink, clk, r);
         ++rows;
         r = out + 11 + 4 * rows;
         mfma_case<8>(x, ncu, bpc, sink, clk, r);
         ++rows;
         r = out + 11 + 4 * rows;
         mfma_case<16>(x, ncu, bpc, sink, clk, r);
         ++rows;
      }
      for (int i = 0; i < rows; ++i)
      {
         const double *r = out + 11 + 4 * i;
         if (r[2] > out[0]) { out[0] = r[2]; out[1] = r[3]; out[2] = r[0]; out[3] = r[1]; }
      }
      {
         const int grid = ncu * 8, iters = 20000;
         const float ms = x.time2([&] {
            hipLaunchKernelGGL(lfpeaks::k_peak_valu, dim3(grid), dim3(256), 0, x.s, iters, 0.999999, sink, clk);
         });
         out[4] = ms > 0 ? (double)grid * 256 * iters * 8 * 2.0 / (ms * 1e-3) / 1e12 : 0.0;
         out[5] = x.clock_ghz(clk, grid);
      }
      {
         // MFMA / VALU overlap: 8 MFMAs (512 pipe cycles) and 128 FMAs
         // (512 cycles) per round, alone and mixed, 2 waves per SIMD
         const int grid = ncu * 2, iters = 4000;
         out[11 + 4 * 16] = x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_mix<8, 0>), dim3(grid), dim3(256), 0, x.s, iters, 0.999999, sink); });
         out[12 + 4 * 16] = x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_mix<0, 128>), dim3(grid), dim3(256), 0, x.s, iters, 0.999999, sink); });
         out[13 + 4 * 16] = x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_mix<8, 128>), dim3(grid), dim3(256), 0, x.s, iters, 0.999999, sink); });
      }
      hipLaunchKernelGGL(lfpeaks::k_fill_zero, dim3(4096), dim3(256), 0, x.s, a, n2);
      const lfpeaks::pk_dv2 *va = reinterpret_cast<const lfpeaks::pk_dv2 *>(a);
      lfpeaks::pk_dv2 *vb = reinterpret_cast<lfpeaks::pk_dv2 *>(b);
      const double nb = (double)nbytes;
      auto best = [&](double &slot, double bytes, float ms) {
         if (ms > 0) { slot = std::max(slot, bytes / (ms * 1e-3) / 1e9); }
      };
      for (int g : {4, 16, 64})
      {
         const dim3 grid((unsigned)(ncu * g));
         best(out[6], 2 * nb, x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_copy_v<1, false>), grid, dim3(256), 0, x.s, va, vb, n2); }));
         best(out[6], 2 * nb, x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_copy_v<4, true>), grid, dim3(256), 0, x.s, va, vb, n2); }));
         best(out[7], nb, x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_write_v<4, false>), grid, dim3(256), 0, x.s, vb, n2); }));
         best(out[7], nb, x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_write_v<4, true>), grid, dim3(256), 0, x.s, vb, n2); }));
         best(out[8], nb, x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_read_v<4, false>), grid, dim3(256), 0, x.s, va, n2, sink); }));
         best(out[8], nb, x.time2([&] { hipLaunchKernelGGL((lfpeaks::k_peak_read_v<4, true>), grid, dim3(256), 0, x.s, va, n2, sink); }));
      }
   }
   if (a) { (void)hipFree(a); }
   if (b) { (void)hipFree(b); }
   if (sink) { (void)hipFree(sink); }
   if (clk) { (void)hipFree(clk); }
   if (x.e0) { (void)hipEventDestroy(x.e0); }
   if (x.e1) { (void)hipEventDestroy(x.e1); }
   if (x.s) { (void)hipStreamDestroy(x.s); }
   return x.ok ? rows : -2;
}

}  // extern "C"
