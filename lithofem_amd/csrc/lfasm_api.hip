// Host side of the gfx950 Nedelec assembly: the C-ABI of include/lithofem_asm.h.
//
// One lfasm_handle replaces one AsmGpuGlobal object of the reference
// (solver/asm_gpu.inc:271-599): it owns every device buffer of one mesh and
// one HIP stream on its device.  Errors are returned as status codes with a
// message string instead of printf + bool (solver/asm_gpu.inc:255-263).
#include "lfasm_device.hpp"
#include "lfasm_batch.hpp"
#include "lfasm_tensor.hpp"
#include "lfasm_pattern.hpp"
#include "lfasm_sym.hpp"
#include "lfasm_extract.hpp"
#include "../../include/lithofem_asm.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace lfasm;

namespace {

struct DevRule {
   int nq = 0;
   double *ip = nullptr, *w = nullptr, *vs = nullptr, *cs = nullptr;
};

enum {
   T_LOCAL = 0, T_SCATTER = 1, T_BUILD = 2, T_BATCH = 3, T_ADJ = 4,
   T_LOCAL_TENSOR = 5, T_LOCAL_QUAD = 6, T_STEP = 7, T_PATTERN = 8, T_GEOM = 9, T_N = 10
};

}  // namespace

// ---- static workspace of lfasm_residual_static --------------------------------
// The entry point that allocates nothing keeps its device memory in this
// file-scope array: x and b staged split re/im (4 n doubles for n <= RS_MAX_N
// dofs), then RS_NPART per-block partial sums (2 doubles each) and the final
// pair.  Every GPU that loads the library holds its own copy for the life of
// the process; its address on each device comes from hipGetSymbolAddress at
// lfasm_create (host code never passes the array's name to a kernel).
namespace {
constexpr int RS_MAX_N = 1572864;            // 1.5 * 2^20 dofs
constexpr int RS_NPART = 2048;
constexpr size_t RS_PART_OFF = 4 * (size_t)RS_MAX_N;
constexpr size_t RS_OUT_OFF = RS_PART_OFF + 2 * (size_t)RS_NPART;
constexpr size_t RS_WORDS = RS_OUT_OFF + 2;
constexpr int RS_MAX_DEVICES = 64;
static_assert(RS_MAX_N == LFASM_RESIDUAL_STATIC_MAX_N, "workspace size and header disagree");

std::mutex g_scratch_mu;
double *g_scratch_ptr[RS_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[RS_WORDS];

namespace {

// Per block: |A x - b|^2 and |b|^2 over a grid-stride set of rows (one row
// per thread, ascending column order as k_spmv), tree-reduced in LDS; the
// block's pair goes to part[2 b], part[2 b + 1].
__global__ void __launch_bounds__(256) k_resid_partial(int n, const int *__restrict__ Ip,
                                                       const int *__restrict__ Jc,
                                                       const double2 *__restrict__ vals,
                                                       const double *__restrict__ ws,
                                                       double *__restrict__ part)
{
   const double *xr = ws, *xi = ws + n, *br = ws + 2 * (size_t)n, *bi = ws + 3 * (size_t)n;
   double rn = 0.0, bn = 0.0;
   for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
   {
      double yr = 0.0, yi = 0.0;
      for (int k = Ip[r]; k < Ip[r + 1]; ++k)
      {
         const double2 a = vals[k];
         const int c = Jc[k];
         yr += a.x * xr[c] - a.y * xi[c];
         yi += a.x * xi[c] + a.y * xr[c];
      }
      const double dr = yr - br[r], di = yi - bi[r];
      rn += dr * dr + di * di;
      bn += br[r] * br[r] + bi[r] * bi[r];
   }
   __shared__ double s_rn[256], s_bn[256];
   s_rn[threadIdx.x] = rn;
   s_bn[threadIdx.x] = bn;
   __syncthreads();
   for (int h = 128; h > 0; h >>= 1)
   {
      if ((int)threadIdx.x < h)
      {
         s_rn[threadIdx.x] += s_rn[threadIdx.x + h];
         s_bn[threadIdx.x] += s_bn[threadIdx.x + h];
      }
      __syncthreads();
   }
   if (threadIdx.x == 0)
   {
      part[2 * blockIdx.x] = s_rn[0];
      part[2 * blockIdx.x + 1] = s_bn[0];
   }
}

// One block: the nb partial pairs summed in a fixed order into out[0..1].
__global__ void __launch_bounds__(256) k_resid_final(int nb, const double *__restrict__ part,
                                                     double *__restrict__ out)
{
   double rn = 0.0, bn = 0.0;
   for (int b = threadIdx.x; b < nb; b += blockDim.x)
   {
      rn += part[2 * b];
      bn += part[2 * b + 1];
   }
   __shared__ double s_rn[256], s_bn[256];
   s_rn[threadIdx.x] = rn;
   s_bn[threadIdx.x] = bn;
   __syncthreads();
   for (int h = 128; h > 0; h >>= 1)
   {
      if ((int)threadIdx.x < h)
      {
         s_rn[threadIdx.x] += s_rn[threadIdx.x + h];
         s_bn[threadIdx.x] += s_bn[threadIdx.x + h];
      }
      __syncthreads();
   }
   if (threadIdx.x == 0)
   {
      out[0] = s_rn[0];
      out[1] = s_bn[0];
   }
}

// Device address of `scratch` on `device` (the current device), resolved
// once per device and kept.
hipError_t scratch_address(int device, double **out)
{
   *out = nullptr;
   if (device < 0 || device >= RS_MAX_DEVICES) { return hipErrorInvalidDevice; }
   std::lock_guard<std::mutex> lock(g_scratch_mu);
   if (!g_scratch_ptr[device])
   {
      double *ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess) { return e; }
      g_scratch_ptr[device] = ptr;
   }
   *out = g_scratch_ptr[device];
   return hipSuccess;
}

}  // namespace

struct lfasm_handle {
   int device = 0;
   double *scratch = nullptr;         // this device's address of the static workspace
   hipStream_t stream = nullptr;
   // captured assembly step (element matrices + scatter)
   hipGraph_t graph = nullptr;
   hipGraphExec_t graph_exec = nullptr;
   bool graph_dirty = true;
   bool g_have = false;                       // capture parameters set
   // pinned staging of the host handoff (two chunks, allocated on first use)
   double2 *pin[2] = {nullptr, nullptr};
   hipEvent_t pin_ev[2] = {nullptr, nullptr};
   int g_mask_cc = 0, g_mask_mass = 0, g_apply_te = 0, g_elim = 0;
   std::string err;
   // mesh
   int ne = 0, nd = 0, p = 0, height = 0, nreg = 0;
   bool inited = false, global_built = false, local_valid = false;
   double *J = nullptr, *Ji = nullptr, *detJ = nullptr, *v0 = nullptr;
   int *attr = nullptr, *fo = nullptr, *dof = nullptr;
   double *sign = nullptr;
   double2 *eps_tab = nullptr, *eps_elem = nullptr;
   // host copies: region eps table (concatenated per problem) and the
   // per-element eps (empty when the region table is in use)
   std::vector<double2> eps_tab_host, eps_elem_host;
   DevRule rc, rm;
   PmlDev *pml_tab = nullptr;
   int *pid = nullptr, *eps_off = nullptr;
   bool any_kpar = false;
   int nmesh = 0;
   std::vector<int64_t> elem_off, row_off;
   double TInv[24];
   double *TInv_dev = nullptr;
   double2 *Ae = nullptr;
   int lda = 0;                       // row stride of the element matrices in Ae
   // constant-coefficient elements (every vertex between z_min and z_max of
   // the element's problem) take the reference-tensor kernel; the others the
   // quadrature kernels
   int nconst = 0, nquad = 0;
   int *clist = nullptr, *qlist = nullptr;
   double *Rten = nullptr;
   int tensor_mode = 1;       // 0 off, 1 materialised Ae (default), 2 fused into the scatter
   unsigned char *eclass = nullptr;   // 1: constant-coefficient element
   double *coef = nullptr;            // ne x TEN_NC element coefficients (fused mode)
   // fused row units per local-index type (edge, face, interior)
   bool fused_ok = false;
   bool fused_tried = false;          // the fused units were attempted for this pattern
   UnitHdr *units[3] = {nullptr, nullptr, nullptr};
   int *rc_enc = nullptr;             // contributor entries of the fused kernels
   std::vector<unsigned char> eclass_host;
   int nunit[3] = {0, 0, 0}, rowcap[3] = {0, 0, 0}, fwaves[3] = {0, 0, 0};
   int local_mode = 0;        // last compute_local: 0 quadrature, 1 tensor + Ae, 2 fused
   int cus = 0;               // compute units of the device
   size_t lds_max = 0;        // LDS bytes one workgroup may hold
   LocalParams last_P{};
   // global
   int64_t nnz = 0;
   int maxrow = 0;
   int *Ip = nullptr, *Jc = nullptr, *rc_ptr = nullptr, *rc_list = nullptr, *bad = nullptr;
   uint16_t *slot = nullptr;
   int sld = 0;                       // slot row stride (entries)
   int *rsoff = nullptr;              // per row: start of its slot block in rslot
   uint8_t *rslot = nullptr;          // byte slot rows in each row's contributor order
                                      // (built when every row has at most 256 entries)
   unsigned char *ess = nullptr, *eflag = nullptr;
   uint64_t *emask = nullptr;         // nd <= 64: per element sign and essential bit masks
   double2 *vals = nullptr;
   // eps variants
   int nparam = 0, nvar = 0;
   int *pelem = nullptr, *prc_ptr = nullptr, *prc_list = nullptr;
   double2 *Mhat = nullptr, *var_vals = nullptr;
   // angle sweeps (lfasm_angles_assemble)
   int nang = 0, nkb = 0;
   double2 *ang_vals = nullptr;       // nang x nnz
   double2 *kbasis = nullptr;         // nkb x nnz basis matrices
   double2 *kAe = nullptr;            // nkae element-matrix sets of the basis terms 1..
   int nkae = 0;
   double *d_kcoef = nullptr;         // nang x 6 combination coefficients
   int kcoef_cap = 0;
   std::vector<PmlDev> pml_host;      // host copy of pml_tab (k_par of each mesh)
   std::vector<int> pelem_host;
   std::vector<int> dof_host;         // element dof table (host copy)
   int pmaxc = 0;                     // most parameter contributors of one row
   int variant_kernel = 0;            // 0 automatic, 1 accumulating, 2 register, 3 general
   int2 *vfix = nullptr;              // chunks of long runs of rows without parameter contributors
   int nvfix = 0;
   unsigned char *vcopied = nullptr;  // per row: inside such a run
   std::vector<int> I_host;           // CSR row pointer (host copy)
   double2 *d_deps = nullptr;         // eps offsets of the variants (device)
   double2 *d_depsT = nullptr;        // the same, parameter-major
   size_t d_deps_cap = 0;
   // incident fields / RHS
   int inc_nrhs = 0, inc_nslab = 0;
   SlabWaveDev *waves = nullptr;
   int *region_slab = nullptr, *region_src = nullptr;
   std::vector<int> region_src_host;
   std::vector<double2> region_deps_host;
   double2 *region_deps = nullptr, *rhs = nullptr;
   // source elements of the incident set (built at the first RHS assembly
   // after lfasm_set_incident) and the per-element RHS staging
   int rhs_nsrc = -1;
   int *rhs_src = nullptr, *rhs_index = nullptr;
   double2 *rhs_bloc = nullptr;
   size_t rhs_cap = 0;
   // timing
   hipEvent_t ev[2 * T_N] = {};
   bool ev_used[T_N] = {};
   int local_kernel = 0;   // tuned choice used by kernel selector 0
   int *sym_queue = nullptr;  // work counter of the persistent quadrature kernel
   // concurrent element kernels: the tensor kernel on `stream` with at most
   // conc_grid workgroups along the element chunks, the quadrature kernel on
   // stream2 (both default priority), forked and joined by events
   int conc_grid = 0;
   int quad_res = 0;       // quadrature workgroups per CU beside the other kernels (0 = no cap)
   // captured step only (scat_overlap = 1): rows whose contributors are all
   // tensor-path elements (rows_t) are scattered on the tensor branch while
   // the quadrature kernel runs; the other rows (rows_q) after the join
   int scat_overlap = 0;
   int *rows_t = nullptr, *rows_q = nullptr;
   int nrows_t = 0, nrows_q = 0;
   bool rows_split = false;
   hipStream_t stream2 = nullptr;
   hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_tensor = nullptr;

   int fail(int code, const char *fmt, ...)
   {
      char buf[512];
      va_list ap;
      va_start(ap, fmt);
      vsnprintf(buf, sizeof(buf), fmt, ap);
      va_end(ap);
      err = buf;
      return code;
   }
};

#define LF_HIP(call)                                                             \
   do {                                                                          \
      hipError_t e_ = (call);                                                    \
      if (e_ != hipSuccess) {                                                    \
         return h->fail(LFASM_E_HIP, "HIP error '%s' at %s:%d (%s)",             \
                        hipGetErrorString(e_), __FILE__, __LINE__, #call);       \
      }                                                                          \
   } while (0)

#define LF_TRY(call)                                                             \
   do {                                                                          \
      hipError_t e_ = (call);                                                    \
      if (e_ != hipSuccess) { return e_; }                                       \
   } while (0)

namespace {

template <typename T>
hipError_t upload(lfasm_handle *h, T *&dst, const T *src, size_t n)
{
   if (dst) { (void)hipFree(dst); dst = nullptr; }
   hipError_t e = hipMalloc((void **)&dst, std::max<size_t>(n, 1) * sizeof(T));
   if (e != hipSuccess) { return e; }
   if (n == 0) { return hipSuccess; }
   e = hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, h->stream);
   if (e != hipSuccess) { return e; }
   return hipStreamSynchronize(h->stream);
}

template <typename T>
void dfree(T *&p)
{
   if (p) { (void)hipFree((void *)p); p = nullptr; }
}

// The captured step holds buffer addresses and launch choices: every call
// that changes either drops it (the next lfasm_step_launch captures anew).
void drop_graph(lfasm_handle *h)
{
   // a launched step may still read the buffers the caller is about to change
   if (h->graph_exec && h->stream) { (void)hipStreamSynchronize(h->stream); }
   if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
   if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
   h->graph_dirty = true;
}

void release_mesh(lfasm_handle *h)
{
   drop_graph(h);
   dfree(h->J); dfree(h->Ji); dfree(h->detJ); dfree(h->v0);
   dfree(h->attr); dfree(h->fo); dfree(h->dof); dfree(h->sign);
   dfree(h->eps_tab); dfree(h->eps_elem); dfree(h->TInv_dev); dfree(h->Ae);
   dfree(h->clist); dfree(h->qlist); dfree(h->Rten); dfree(h->eclass); dfree(h->coef);
   h->nconst = h->nquad = 0;
   dfree(h->pml_tab); dfree(h->pid); dfree(h->eps_off);
   dfree(h->waves); dfree(h->region_slab); dfree(h->region_src); dfree(h->region_deps);
   dfree(h->rhs_src); dfree(h->rhs_index); dfree(h->rhs_bloc);
   h->rhs_nsrc = -1;
   h->inc_nrhs = 0;
   h->eps_tab_host.clear();
   h->eps_elem_host.clear();
   h->region_src_host.clear();
   h->region_deps_host.clear();
   for (DevRule *r : {&h->rc, &h->rm})
   {
      dfree(r->ip); dfree(r->w); dfree(r->vs); dfree(r->cs);
      r->nq = 0;
   }
   h->inited = false;
   h->local_valid = false;
}

void release_global(lfasm_handle *h)
{
   drop_graph(h);
   dfree(h->Ip); dfree(h->Jc); dfree(h->rc_ptr); dfree(h->rc_list);
   for (int t = 0; t < 3; ++t)
   {
      dfree(h->units[t]);
      h->nunit[t] = h->rowcap[t] = h->fwaves[t] = 0;
   }
   dfree(h->rc_enc);
   h->fused_ok = false;
   h->fused_tried = false;
   dfree(h->bad); dfree(h->slot); dfree(h->ess); dfree(h->eflag); dfree(h->vals);
   dfree(h->rsoff); dfree(h->rslot);
   dfree(h->emask);
   dfree(h->rows_t); dfree(h->rows_q);
   h->nrows_t = h->nrows_q = 0;
   h->rows_split = false;
   dfree(h->pelem); dfree(h->prc_ptr); dfree(h->prc_list); dfree(h->Mhat); dfree(h->var_vals);
   dfree(h->ang_vals); dfree(h->kbasis); dfree(h->d_kcoef); dfree(h->kAe);
   h->nang = h->nkb = h->kcoef_cap = h->nkae = 0;
   dfree(h->rhs); dfree(h->d_deps); dfree(h->d_depsT); dfree(h->vfix); dfree(h->vcopied);
   h->nvfix = 0;
   h->rhs_cap = 0;
   h->d_deps_cap = 0;
   h->nparam = h->nvar = 0;
   h->global_built = false;
}

int check_rule(lfasm_handle *h, const lfasm_rule &r, const char *name)
{
   if (r.nq <= 0 || !r.ip || !r.w || !r.vshape || !r.curl)
   {
      return h->fail(LFASM_E_ARG, "rule %s: nq must be > 0 and all tables set", name);
   }
   return LFASM_OK;
}

hipError_t upload_rule(lfasm_handle *h, DevRule &d, const lfasm_rule &r, int nd)
{
   d.nq = r.nq;
   hipError_t e;
   if ((e = upload(h, d.ip, r.ip, (size_t)r.nq * 3)) != hipSuccess) { return e; }
   if ((e = upload(h, d.w, r.w, (size_t)r.nq)) != hipSuccess) { return e; }
   if ((e = upload(h, d.vs, r.vshape, (size_t)r.nq * nd * 3)) != hipSuccess) { return e; }
   return upload(h, d.cs, r.curl, (size_t)r.nq * nd * 3);
}

LocalParams local_params(lfasm_handle *h, int mask_cc, int mask_mass, int apply_te,
                         double2 *Ae)
{
   LocalParams P;
   P.ne = h->ne;
   P.nd = h->nd;
   P.p = h->p;
   P.J = h->J;
   P.Ji = h->Ji;
   P.detJ = h->detJ;
   P.v0 = h->v0;
   P.attr = h->attr;
   P.eps_tab = h->eps_tab;
   P.eps_elem = h->eps_elem;
   P.fo = h->fo;
   P.rc = RuleDev{h->rc.nq, h->rc.ip, h->rc.w, h->rc.vs, h->rc.cs};
   P.rm = RuleDev{h->rm.nq, h->rm.ip, h->rm.w, h->rm.vs, h->rm.cs};
   P.pml_tab = h->pml_tab;
   P.pid = h->pid;
   P.eps_off = h->eps_off;
   std::memcpy(P.TInv, h->TInv, sizeof(P.TInv));
   P.use_cc = (mask_cc & LFASM_CURLCURL) ? 1 : 0;
   int mf = 0;
   if (mask_mass & LFASM_MASS) { mf |= 1; }
   if (h->any_kpar)
   {
      // the k-parallel integrators exist only when k_par != 0
      // (solver/lithofem_solve.cpp:695-704)
      if (mask_mass & LFASM_CROSS0) { mf |= 2; }
      if (mask_mass & LFASM_CROSS1) { mf |= 4; }
      if (mask_mass & LFASM_KLK) { mf |= 8; }
   }
   P.mflags = mf;
   P.apply_te = apply_te && h->p >= 2;
   P.Ae = Ae;
   P.lda = h->lda;
   P.nout = h->ne;
   P.elist = nullptr;
   P.eps_unit = 0;
   P.out_by_elem = 0;
   return P;
}

// Setup records of the general MFMA kernels (k_elem_prep), one wavefront
// per list entry; each record sits in its entry's own element-matrix slot.
hipError_t launch_elem_prep(lfasm_handle *h, const LocalParams &P, hipStream_t st)
{
   const int nqt = P.rc.nq + P.rm.nq;
   if ((size_t)(EC_SIZE + 4 * nqt) > 2 * (size_t)P.nd * P.lda) { return hipErrorInvalidValue; }
   hipLaunchKernelGGL(k_elem_prep, dim3((unsigned)P.nout), dim3(64), 0, st, P, nqt);
   return hipGetLastError();
}

constexpr size_t LDS_MAX = 163840;   // LDS bytes per CU on gfx950

// LDS request of an MFMA quadrature kernel.  Launched on the second stream
// (beside the tensor kernel and the row scatter) with a residency cap
// quad_res > 0, the request is raised so that at most quad_res workgroups fit
// on one CU: the kernel then holds fewer of the CU's wave slots and vector
// registers, which the HBM-bound kernels beside it need to keep enough
// loads and stores in flight.  The kernel uses only its own part.
size_t resident_shmem(const lfasm_handle *h, size_t shmem, hipStream_t st)
{
   if (st != h->stream2 || h->quad_res <= 0) { return shmem; }
   const size_t need = LDS_MAX / (size_t)(h->quad_res + 1) + 1024;
   return std::max(shmem, std::min(need, h->lds_max));
}

template <int NT, int RT, int EPB, int WPEU>
hipError_t launch_mfma(lfasm_handle *h, const LocalParams &P, hipStream_t st)
{
   constexpr int NP = 16 * NT;
   constexpr int WPE = NT / RT;
   constexpr int THREADS = 64 * WPE * EPB;
   const int nqt = P.rc.nq + P.rm.nq;
   const size_t shmem = resident_shmem(
      h, (size_t)EPB * (EC_SIZE + nqt * 4 + 3 * CHUNK_ROWS * NP) * sizeof(double), st);
   hipError_t e = hipFuncSetAttribute((const void *)k_local_mfma<NT, RT, EPB, WPEU>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
   if (e != hipSuccess) { return e; }
   LF_TRY(launch_elem_prep(h, P, st));
   const int grid = (P.nout + EPB - 1) / EPB;
   hipLaunchKernelGGL((k_local_mfma<NT, RT, EPB, WPEU>), dim3(grid), dim3(THREADS), shmem,
                      st, P, nqt);
   return hipGetLastError();
}

template <int NACC, int NTHR>
hipError_t launch_generic(lfasm_handle *h, const LocalParams &P, hipStream_t st)
{
   const int nqt = P.rc.nq + P.rm.nq;
   const int NP = P.nd;
   const size_t shmem = (size_t)(EC_SIZE + nqt * 4 + 3 * CHUNK_ROWS * NP) * sizeof(double);
   hipError_t e = hipFuncSetAttribute((const void *)k_local_generic<NACC, NTHR>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
   if (e != hipSuccess) { return e; }
   hipLaunchKernelGGL((k_local_generic<NACC, NTHR>), dim3(P.nout), dim3(NTHR), shmem, st,
                      P, nqt, NP);
   e = hipGetLastError();
   if (e != hipSuccess || !P.apply_te) { return e; }
   hipLaunchKernelGGL(k_te, dim3(P.nout), dim3(64), 0, st, P.nd, P.lda, P.p, P.fo,
                      P.elist, P.out_by_elem, h->TInv_dev, P.Ae);
   return hipGetLastError();
}

template <int NT, int EPB, int WPEU>
hipError_t launch_mfma_pipe(lfasm_handle *h, const LocalParams &P, hipStream_t st)
{
   constexpr int NP = 16 * NT;
   constexpr int THREADS = 64 * NT * EPB;
   const int nqt = P.rc.nq + P.rm.nq;
   const size_t shmem = resident_shmem(
      h, (size_t)EPB * (EC_SIZE + nqt * 4 + 2 * 3 * CHUNK_ROWS * NP) * sizeof(double), st);
   hipError_t e = hipFuncSetAttribute((const void *)k_local_mfma_pipe<NT, EPB, WPEU>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
   if (e != hipSuccess) { return e; }
   LF_TRY(launch_elem_prep(h, P, st));
   const int grid = (P.nout + EPB - 1) / EPB;
   hipLaunchKernelGGL((k_local_mfma_pipe<NT, EPB, WPEU>), dim3(grid), dim3(THREADS), shmem,
                      st, P, nqt);
   return hipGetLastError();
}

// Symmetric-split kernel (nd <= 48 with 3 tiles; CROSS0 and CROSS1 both on
// or both off), after its setup-record kernel.  PERSIST: as many workgroups
// as fit on the device at once, fed from the handle's work counter.
template <int NBUF, int WPEU, int PRIO = 0, bool PERSIST = false>
hipError_t launch_mfma_sym(lfasm_handle *h, const LocalParams &P, hipStream_t st)
{
   const int nqt = P.rc.nq + P.rm.nq;
   // each entry's setup record sits in its own element-matrix slot
   if ((size_t)sym_rec_size(nqt) > 2 * (size_t)P.nd * P.lda) { return hipErrorInvalidValue; }
   // the persistent form prefetches a record in SYM_PF registers per thread
   if (PERSIST && sym_rec_size(nqt) > SYM_PF * SYM_THREADS)
   {
      return launch_mfma_sym<NBUF, WPEU, PRIO, false>(h, P, st);
   }
   const size_t shmem = resident_shmem(
      h, (size_t)(EC_SIZE + nqt * 4 + SYM_FCS + NBUF * SYM_STG) * sizeof(double), st);
   auto kern = k_local_mfma_sym<NBUF, WPEU, PRIO, PERSIST>;
   hipError_t e = hipFuncSetAttribute((const void *)kern,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
   if (e != hipSuccess) { return e; }
   int grid = P.nout;
   if (PERSIST)
   {
      int per_cu = 0;
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, SYM_THREADS,
                                                       shmem);
      if (e != hipSuccess) { return e; }
      grid = std::max(1, std::min(P.nout, std::max(1, per_cu) * h->cus));
   }
   hipLaunchKernelGGL(k_sym_prep, dim3((unsigned)P.nout), dim3(64), 0, st, P, nqt,
                      PERSIST ? h->sym_queue : nullptr);
   e = hipGetLastError();
   if (e != hipSuccess) { return e; }
   hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(SYM_THREADS), shmem, st, P, nqt,
                      h->sym_queue);
   return hipGetLastError();
}

bool sym_applies(const LocalParams &P)
{
   return P.nd > 32 && P.nd <= 48 && ((P.mflags >> 1) & 1) == ((P.mflags >> 2) & 1);
}

// MFMA launch variants per tile count (element-per-block count, waves per
// EU, pipelined or not, symmetric split).  Variant 0 of each list is the
// default; the launch autotuner (lithofem_amd.tune) times all of them.
constexpr int N_MFMA_VARIANTS = LFASM_N_MFMA_VARIANTS;

hipError_t launch_mfma_variant(lfasm_handle *h, const LocalParams &P, int nt, int v,
                               hipStream_t st)
{
   switch (nt * 16 + v)
   {
   case 1 * 16 + 0: return launch_mfma_pipe<1, 4, 2>(h, P, st);
   case 1 * 16 + 1: return launch_mfma_pipe<1, 2, 2>(h, P, st);
   case 1 * 16 + 2: return launch_mfma<1, 1, 4, 2>(h, P, st);
   case 1 * 16 + 3: return launch_mfma_pipe<1, 4, 4>(h, P, st);
   case 1 * 16 + 4: return launch_mfma_pipe<1, 4, 2>(h, P, st);
   case 1 * 16 + 5: return launch_mfma_pipe<1, 2, 2>(h, P, st);
   case 1 * 16 + 6: return launch_mfma_pipe<1, 4, 2>(h, P, st);
   case 1 * 16 + 7: return launch_mfma_pipe<1, 4, 2>(h, P, st);
   case 2 * 16 + 0: return launch_mfma<2, 2, 4, 2>(h, P, st);
   case 2 * 16 + 1: return launch_mfma_pipe<2, 1, 2>(h, P, st);
   case 2 * 16 + 2: return launch_mfma<2, 1, 2, 2>(h, P, st);
   case 2 * 16 + 3: return launch_mfma_pipe<2, 2, 2>(h, P, st);
   case 2 * 16 + 4: return launch_mfma_pipe<2, 2, 2>(h, P, st);
   case 2 * 16 + 5: return launch_mfma_pipe<2, 1, 2>(h, P, st);
   case 2 * 16 + 6: return launch_mfma<2, 2, 4, 2>(h, P, st);
   case 2 * 16 + 7: return launch_mfma<2, 2, 4, 2>(h, P, st);
   case 3 * 16 + 0: return sym_applies(P) ? launch_mfma_sym<1, 3>(h, P, st) : launch_mfma_pipe<3, 1, 4>(h, P, st);
   case 3 * 16 + 1: return sym_applies(P) ? launch_mfma_sym<2, 3, 1>(h, P, st) : launch_mfma_pipe<3, 2, 2>(h, P, st);
   case 3 * 16 + 2: return launch_mfma_pipe<3, 1, 4>(h, P, st);
   case 3 * 16 + 3: return launch_mfma<3, 1, 1, 2>(h, P, st);
   case 3 * 16 + 4: return sym_applies(P) ? launch_mfma_sym<1, 3, 2>(h, P, st) : launch_mfma_pipe<3, 1, 4>(h, P, st);
   case 3 * 16 + 5: return sym_applies(P) ? launch_mfma_sym<1, 3, 1>(h, P, st) : launch_mfma_pipe<3, 1, 2>(h, P, st);
   case 3 * 16 + 6: return sym_applies(P) ? launch_mfma_sym<1, 3, 1, true>(h, P, st) : launch_mfma_pipe<3, 1, 4>(h, P, st);
   case 3 * 16 + 7: return sym_applies(P) ? launch_mfma_sym<1, 3, 2, true>(h, P, st) : launch_mfma_pipe<3, 1, 4>(h, P, st);
   case 6 * 16 + 0: return launch_mfma_pipe<6, 1, 2>(h, P, st);
   case 6 * 16 + 1: return launch_mfma_pipe<6, 1, 1>(h, P, st);
   case 6 * 16 + 2: return launch_mfma<6, 1, 1, 2>(h, P, st);
   case 6 * 16 + 3: return launch_mfma_pipe<6, 1, 2>(h, P, st);
   case 6 * 16 + 4: return launch_mfma_pipe<6, 1, 2>(h, P, st);
   case 6 * 16 + 5: return launch_mfma_pipe<6, 1, 1>(h, P, st);
   case 6 * 16 + 6: return launch_mfma_pipe<6, 1, 2>(h, P, st);
   case 6 * 16 + 7: return launch_mfma_pipe<6, 1, 2>(h, P, st);
   default: return hipErrorInvalidValue;
   }
}

// Kernel choice: `kernel` = 0 automatic (the handle's tuned variant, else
// MFMA variant 0 for nd > 16 -- at p = 2 on 57,660 PML elements 0.61 ms
// against 1.05 ms for the VALU kernel and its separate face transform --
// and the VALU kernel for nd <= 16), 1 = VALU kernel, 2 + v = MFMA variant v.
hipError_t launch_local(lfasm_handle *h, const LocalParams &P, int kernel, hipStream_t st)
{
   const int nd = P.nd;
   // 16-wide column tiles: p = 1..4 (nd 6, 20, 45, 84) take 1, 2, 3, 6
   const int nt = nd <= 16 ? 1 : nd <= 32 ? 2 : nd <= 48 ? 3 : 6;
   if (kernel >= 2 && kernel < 2 + N_MFMA_VARIANTS) { return launch_mfma_variant(h, P, nt, kernel - 2, st); }
   if (kernel == 0 && h->local_kernel != 0) { return launch_local(h, P, h->local_kernel, st); }
   if (kernel == 0 && nd > 16) { return launch_mfma_variant(h, P, nt, 0, st); }
   const int n2 = nd * nd;
   if (n2 <= 256) { return launch_generic<1, 256>(h, P, st); }
   if (n2 <= 512) { return launch_generic<2, 256>(h, P, st); }
   if (n2 <= 2048) { return launch_generic<4, 512>(h, P, st); }
   if (n2 <= 8192) { return launch_generic<8, 1024>(h, P, st); }
   return hipErrorInvalidValue;
}

// Geometry of the LDS slice of local-index type `t` (0 edge, 1 face, 2 interior).
void slice_geometry(int nd, int p, int t, int &r0, int &nR)
{
   const int ne_ = 6 * p, nf = 4 * p * (p - 1);
   r0 = t == 0 ? 0 : t == 1 ? ne_ : ne_ + nf;
   nR = t == 0 ? ne_ : t == 1 ? nf : nd - ne_ - nf;
}

size_t slice_bytes(int nd, int nR)
{
   const size_t S = (size_t)nR * (nd - nR) + (size_t)nR * (nR + 1) / 2;
   return (size_t)NTENS * S * sizeof(double);
}

// Row units of the fused scatter: every CSR row is typed by the local index
// its contributors hold (the same type in every element); face rows whose
// partner row (local a+1 in every contributing element, same elements in
// the same order) exists form one unit.  A pattern that does not fit (mixed
// types, more than 64 contributors in a row, rows too long for LDS) keeps
// the fused path off; the materialised path still works.
int build_fused_units(lfasm_handle *h, const int *I, const std::vector<int> &dof,
                      const std::vector<int> &rc_ptr, const std::vector<int> &rc_list,
                      const std::vector<double> &sign, const std::vector<unsigned char> &ess)
{
   const int n = h->height, nd = h->nd, p = h->p;
   if (nd > 64 || (int64_t)h->ne * nd > CE_MASK) { return LFASM_OK; }
   std::vector<int> type((size_t)n, 0);
   int r0s[3], nRs[3];
   for (int t = 0; t < 3; ++t) { slice_geometry(nd, p, t, r0s[t], nRs[t]); }
   auto type_of = [&](int a) { return a < r0s[1] ? 0 : a < r0s[2] ? 1 : 2; };
   for (int r = 0; r < n; ++r)
   {
      const int c0 = rc_ptr[(size_t)r], c1 = rc_ptr[(size_t)r + 1];
      if (c1 - c0 > 64) { return LFASM_OK; }
      const int t = c0 < c1 ? type_of(rc_list[(size_t)c0] % nd) : 0;
      for (int c = c0; c < c1; ++c)
      {
         if (type_of(rc_list[(size_t)c] % nd) != t) { return LFASM_OK; }
      }
      type[(size_t)r] = t;
   }
   std::vector<UnitHdr> units[3];
   std::vector<char> taken((size_t)n, 0);
   int rowcap[3] = {0, 0, 0};
   for (int r = 0; r < n; ++r)
   {
      const int t = type[(size_t)r];
      rowcap[t] = std::max(rowcap[t], I[r + 1] - I[r]);
      if (taken[(size_t)r]) { continue; }
      taken[(size_t)r] = 1;
      int partner = -1;
      const int c0 = rc_ptr[(size_t)r], c1 = rc_ptr[(size_t)r + 1];
      if (t == 1 && c0 < c1)
      {
         const int ea0 = rc_list[(size_t)c0];
         const int a0 = ea0 % nd;
         if (((a0 - r0s[1]) & 1) == 0)
         {
            const int j = dof[(size_t)ea0 + 1];
            bool ok = j != r && !taken[(size_t)j] &&
                      rc_ptr[(size_t)j + 1] - rc_ptr[(size_t)j] == c1 - c0;
            for (int c = c0; ok && c < c1; ++c)
            {
               const int ea = rc_list[(size_t)c];
               ok = ((ea % nd - r0s[1]) & 1) == 0 &&
                    rc_list[(size_t)(rc_ptr[(size_t)j] + (c - c0))] == ea + 1;
            }
            if (ok) { partner = j; }
         }
      }
      if (partner >= 0) { taken[(size_t)partner] = 1; }
      UnitHdr u;
      u.i0 = r;
      u.i1 = partner;
      u.base0 = I[r];
      u.len0 = I[r + 1] - I[r];
      u.base1 = partner >= 0 ? I[partner] : 0;
      u.len1 = partner >= 0 ? I[partner + 1] - I[partner] : 0;
      u.c0 = c0;
      u.nc_flags = ((c1 - c0) << 8) | (ess[(size_t)r] ? 1 : 0) |
                   (partner >= 0 && ess[(size_t)partner] ? 2 : 0);
      units[t].push_back(u);
   }
   // contributor entries: e*nd + a, PML bit, signs of rows a and a+1
   std::vector<int> enc(rc_list.size() + 64, 0);   // 64 pad entries: unguarded list loads
   for (size_t c = 0; c < rc_list.size(); ++c)
   {
      const int ea = rc_list[c];
      const int a = ea % nd;
      unsigned v = (unsigned)ea;
      if (!h->eclass_host[(size_t)(ea / nd)]) { v |= 1u << 31; }
      if (sign[(size_t)ea] < 0.0) { v |= 1u << 30; }
      if (a + 1 < nd && sign[(size_t)ea + 1] < 0.0) { v |= 1u << 29; }
      enc[c] = (int)v;
   }
   // launch shape: as many waves as the LDS left by the tensor slice holds
   int waves[3];
   for (int t = 0; t < 3; ++t)
   {
      waves[t] = 0;
      if (units[t].empty()) { continue; }
      const size_t rb = slice_bytes(nd, nRs[t]);
      const size_t per_wave = (size_t)(t == 1 ? 2 : 1) * std::max(rowcap[t], 1) * sizeof(double2);
      if (rb + 4 * per_wave > h->lds_max) { return LFASM_OK; }
      waves[t] = (int)std::min<size_t>(FUSED_MAX_WAVES, (h->lds_max - rb) / per_wave);
   }
   for (int t = 0; t < 3; ++t)
   {
      LF_HIP(upload(h, h->units[t], units[t].data(), units[t].size()));
      h->nunit[t] = (int)units[t].size();
      h->rowcap[t] = std::max(rowcap[t], 1);
      h->fwaves[t] = waves[t];
   }
   LF_HIP(upload(h, h->rc_enc, enc.data(), enc.size()));
   h->fused_ok = true;
   return LFASM_OK;
}

// Element coefficients of the constant-coefficient elements (fused mode).
hipError_t launch_tensor_coef(lfasm_handle *h, const TensorParams &T)
{
   hipLaunchKernelGGL(k_tensor_coef, dim3((unsigned)((T.nlist + 255) / 256)), dim3(256), 0,
                      h->stream, T, h->coef);
   return hipGetLastError();
}

// The three fused row launches (edge, face, interior rows).
hipError_t launch_fused_rows(lfasm_handle *h, int eliminate)
{
   for (int t = 0; t < 3; ++t)
   {
      if (h->nunit[t] == 0) { continue; }
      FusedParams F;
      F.nd = h->nd;
      F.p = h->p;
      F.npair = (h->nd + 1) / 2;
      slice_geometry(h->nd, h->p, t, F.r0, F.nR);
      F.no = h->nd - F.nR;
      F.S = F.nR * F.no + F.nR * (F.nR + 1) / 2;
      F.face = t == 1 ? 1 : 0;
      F.nunit = h->nunit[t];
      F.hdr = h->units[t];
      F.rowcap = h->rowcap[t];
      F.Jc = h->Jc;
      F.eliminate = eliminate ? 1 : 0;
      F.slot = h->slot;
      F.sld = h->sld;
      F.eflag = h->eflag;
      std::memcpy(F.TInv, h->TInv, sizeof(F.TInv));
      F.Ae = h->Ae;
      F.lda = h->lda;
      F.R = h->Rten;
      F.vals = h->vals;
      const int nw = h->fwaves[t];
      const size_t shmem = slice_bytes(h->nd, F.nR) +
                           (size_t)nw * (F.face ? 2 : 1) * F.rowcap * sizeof(double2);
      LF_TRY(hipFuncSetAttribute((const void *)k_fused_rows,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
      const int grid = std::max(1, std::min(h->cus, (F.nunit + nw - 1) / nw));
      hipLaunchKernelGGL(k_fused_rows, dim3(grid), dim3(64 * nw), shmem, h->stream, F,
                         (const double *)h->coef, (const int *)h->rc_enc);
      LF_TRY(hipGetLastError());
   }
   return hipSuccess;
}

// Reference-tensor kernel over the handle's constant-coefficient elements.
TensorParams tensor_params(lfasm_handle *h, const LocalParams &P)
{
   TensorParams T;
   T.nd = h->nd;
   T.p = h->p;
   T.npair = (h->nd + 1) / 2;
   T.nb = T.npair * T.npair;
   T.nlist = h->nconst;
   T.elist = h->clist;
   T.R = h->Rten;
   T.J = h->J;
   T.Ji = h->Ji;
   T.detJ = h->detJ;
   T.attr = h->attr;
   T.fo = h->fo;
   T.pid = h->pid;
   T.eps_off = h->eps_off;
   T.eps_tab = h->eps_tab;
   T.eps_elem = h->eps_elem;
   T.pml_tab = h->pml_tab;
   std::memcpy(T.TInv, h->TInv, sizeof(T.TInv));
   T.use_cc = P.use_cc;
   T.use_mass = P.mflags & 1;
   T.use_klk = (P.mflags >> 3) & 1;
   T.use_cross = (P.mflags >> 1) & 1;
   T.apply_te = P.apply_te;
   T.eps_unit = P.eps_unit;
   T.out_compact = 0;
   T.Ae = P.Ae;
   T.lda = P.lda;
   return T;
}

hipError_t launch_tensor(lfasm_handle *h, const LocalParams &P)
{
   const TensorParams T = tensor_params(h, P);
   if (T.lda % 8 == 0 && T.lda > T.nd && !T.out_compact)
   {
      // padded rows: the column-per-lane kernel with whole-line stores
      // element coefficients first, then the entries
      LF_TRY(launch_tensor_coef(h, T));
      const int nchunk = (h->nconst + TEN21_EPC - 1) / TEN21_EPC;
      const dim3 grid((unsigned)(h->conc_grid > 0 ? std::min(nchunk, h->conc_grid) : nchunk),
                      (unsigned)((T.npair * T.lda + 255) / 256));
      hipLaunchKernelGGL(k_local_tensor21, grid, dim3(256), 0, h->stream, T, T.elist,
                         (const double *)h->coef, T.Ae);
      return hipGetLastError();
   }
   const int epar = T.nb > 128 ? 1 : 256 / T.nb;
   const dim3 grid((unsigned)((h->nconst + TEN_EPC - 1) / TEN_EPC),
                   epar > 1 ? 1u : (unsigned)((T.nb + 255) / 256));
   hipLaunchKernelGGL(k_local_tensor, grid, dim3(256), 0, h->stream, T, epar);
   return hipGetLastError();
}

// Device complex array -> host re / im arrays, in chunks of 64 MB through
// the handle's two pinned buffers: the copy of chunk k overlaps the
// host-side split of chunk k-1 (threads).  Either output may be null.
int download_split(lfasm_handle *h, const double2 *dev, size_t n, double *re, double *im)
{
   constexpr size_t CH = (size_t)1 << 22;
   for (int k = 0; k < 2; ++k)
   {
      if (!h->pin[k]) { LF_HIP(hipHostMalloc((void **)&h->pin[k], CH * sizeof(double2))); }
      if (!h->pin_ev[k]) { LF_HIP(hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming)); }
   }
   const size_t nch = (n + CH - 1) / CH;
   const unsigned nthr = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
   auto split = [&](size_t c) {
      const double2 *src = h->pin[c & 1];
      const size_t off = c * CH, len = std::min(CH, n - off);
      std::vector<std::thread> pool;
      for (unsigned t = 0; t < nthr; ++t)
      {
         pool.emplace_back([=]() {
            const size_t a = len * t / nthr, b = len * (t + 1) / nthr;
            for (size_t k = a; k < b; ++k)
            {
               if (re) { re[off + k] = src[k].x; }
               if (im) { im[off + k] = src[k].y; }
            }
         });
      }
      for (auto &th : pool) { th.join(); }
   };
   for (size_t c = 0; c <= nch; ++c)
   {
      if (c < nch)
      {
         const size_t off = c * CH, len = std::min(CH, n - off);
         LF_HIP(hipMemcpyAsync(h->pin[c & 1], dev + off, len * sizeof(double2),
                               hipMemcpyDeviceToHost, h->stream));
         LF_HIP(hipEventRecord(h->pin_ev[c & 1], h->stream));
      }
      if (c > 0)
      {
         LF_HIP(hipEventSynchronize(h->pin_ev[(c - 1) & 1]));
         split(c - 1);
      }
   }
   return LFASM_OK;
}

// Device rows of `lda` complex entries -> host rows of `nd` (the padding
// dropped), in chunks through the handle's two pinned 64 MB buffers: the
// copy of chunk k overlaps the host-side repacking of chunk k-1 (threads).
int download_unpad(lfasm_handle *h, const double2 *dev, size_t rows, int nd, int lda, double *out)
{
   constexpr size_t CH = (size_t)1 << 22;            // double2 entries per buffer
   for (int k = 0; k < 2; ++k)
   {
      if (!h->pin[k]) { LF_HIP(hipHostMalloc((void **)&h->pin[k], CH * sizeof(double2))); }
      if (!h->pin_ev[k]) { LF_HIP(hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming)); }
   }
   const size_t rch = std::max<size_t>(1, CH / (size_t)lda);   // rows per chunk
   const size_t nch = (rows + rch - 1) / rch;
   const unsigned nthr = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
   auto repack = [&](size_t c) {
      const double2 *src = h->pin[c & 1];
      const size_t r0 = c * rch, nr = std::min(rch, rows - r0);
      std::vector<std::thread> pool;
      for (unsigned t = 0; t < nthr; ++t)
      {
         pool.emplace_back([=]() {
            for (size_t r = nr * t / nthr; r < nr * (t + 1) / nthr; ++r)
            {
               std::memcpy(out + 2 * (r0 + r) * nd, src + r * lda, (size_t)nd * sizeof(double2));
            }
         });
      }
      for (auto &th : pool) { th.join(); }
   };
   for (size_t c = 0; c <= nch; ++c)
   {
      if (c < nch)
      {
         const size_t r0 = c * rch, nr = std::min(rch, rows - r0);
         LF_HIP(hipMemcpyAsync(h->pin[c & 1], dev + r0 * lda, nr * lda * sizeof(double2),
                               hipMemcpyDeviceToHost, h->stream));
         LF_HIP(hipEventRecord(h->pin_ev[c & 1], h->stream));
      }
      if (c > 0)
      {
         LF_HIP(hipEventSynchronize(h->pin_ev[(c - 1) & 1]));
         repack(c - 1);
      }
   }
   return LFASM_OK;
}

hipError_t ev_begin(lfasm_handle *h, int t, hipStream_t st = nullptr)
{
   h->ev_used[t] = true;
   return hipEventRecord(h->ev[2 * t], st ? st : h->stream);
}

hipError_t ev_end(lfasm_handle *h, int t, hipStream_t st = nullptr)
{
   return hipEventRecord(h->ev[2 * t + 1], st ? st : h->stream);
}


// Element matrices of the whole mesh, enqueued on h->stream without a
// host wait: with the tensor path, the reference-tensor kernel over the
// constant-coefficient elements, then the quadrature kernel over the PML
// elements (both fill the device, so running them on two streams gains
// nothing).  `timed` records the timing events (never inside a graph
// capture).
hipError_t enqueue_local(lfasm_handle *h, const LocalParams &P, int kernel, bool timed)
{
   // the tensor kernel sums CROSS0 and CROSS1 as one antisymmetric term, so
   // a mask holding only one of them goes through the quadrature kernels
   const bool one_cross = ((P.mflags >> 1) & 1) != ((P.mflags >> 2) & 1);
   const bool split = h->tensor_mode > 0 && !one_cross && h->nconst > 0;
   // fused: element coefficients only; the row kernels evaluate the tensor
   // rows (needs the row units of build_global and the face transform on)
   const bool fused = split && h->tensor_mode == 2 && h->fused_ok && h->global_built &&
                      (P.apply_te || h->p < 2);
   h->last_P = P;
   h->local_mode = fused ? 2 : split ? 1 : 0;
   if (timed)
   {
      h->ev_used[T_LOCAL_TENSOR] = h->ev_used[T_LOCAL_QUAD] = false;
      LF_TRY(ev_begin(h, T_LOCAL));
   }
   if (!P.use_cc && !P.mflags)
   {
      const size_t n = (size_t)h->ne * h->nd * h->lda;
      hipLaunchKernelGGL(k_fill_zero, dim3(4096), dim3(BLOCK), 0, h->stream, P.Ae, n);
      LF_TRY(hipGetLastError());
   }
   else if (split && !fused && h->conc_grid > 0 && h->nquad > 0)
   {
      // tensor kernel and quadrature kernel side by side
      LocalParams Pq = P;
      Pq.elist = h->qlist;
      Pq.nout = h->nquad;
      Pq.out_by_elem = 1;
      LF_TRY(hipEventRecord(h->ev_fork, h->stream));
      LF_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
      if (timed) { LF_TRY(ev_begin(h, T_LOCAL_TENSOR)); }
      LF_TRY(launch_tensor(h, P));
      if (timed) { LF_TRY(ev_end(h, T_LOCAL_TENSOR)); }
      if (timed) { LF_TRY(ev_begin(h, T_LOCAL_QUAD, h->stream2)); }
      LF_TRY(launch_local(h, Pq, kernel, h->stream2));
      if (timed) { LF_TRY(ev_end(h, T_LOCAL_QUAD, h->stream2)); }
      LF_TRY(hipEventRecord(h->ev_join, h->stream2));
      LF_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
   }
   else if (split)
   {
      if (timed) { LF_TRY(ev_begin(h, T_LOCAL_TENSOR)); }
      if (fused) { LF_TRY(launch_tensor_coef(h, tensor_params(h, P))); }
      else { LF_TRY(launch_tensor(h, P)); }
      if (timed) { LF_TRY(ev_end(h, T_LOCAL_TENSOR)); }
      if (h->nquad > 0)
      {
         LocalParams Pq = P;
         Pq.elist = h->qlist;
         Pq.nout = h->nquad;
         Pq.out_by_elem = 1;
         if (timed) { LF_TRY(ev_begin(h, T_LOCAL_QUAD)); }
         LF_TRY(launch_local(h, Pq, kernel, h->stream));
         if (timed) { LF_TRY(ev_end(h, T_LOCAL_QUAD)); }
      }
   }
   else
   {
      LF_TRY(launch_local(h, P, kernel, h->stream));
   }
   if (timed) { LF_TRY(ev_end(h, T_LOCAL)); }
   return hipSuccess;
}

// Scatter + elimination of the element matrices into the CSR values.
hipError_t enqueue_scatter_rows(lfasm_handle *h, int eliminate, bool timed, const int *rows,
                                int nrows, hipStream_t st);

hipError_t enqueue_scatter(lfasm_handle *h, int eliminate, bool timed)
{
   return enqueue_scatter_rows(h, eliminate, timed, nullptr, h->height, h->stream);
}

// Scatter + elimination of the rows `rows` (NULL: all rows) on stream st.
hipError_t enqueue_scatter_rows(lfasm_handle *h, int eliminate, bool timed, const int *rows,
                                int nrows, hipStream_t st)
{
   if (h->local_mode == 2)
   {
      if (timed) { LF_TRY(ev_begin(h, T_SCATTER)); }
      LF_TRY(launch_fused_rows(h, eliminate));
      if (timed) { LF_TRY(ev_end(h, T_SCATTER)); }
      return hipSuccess;
   }
   ScatterParams S;
   S.n = h->height;
   S.nd = h->nd;
   S.Ip = h->Ip;
   S.Jc = h->Jc;
   S.eflag = h->eflag;
   S.emask = h->emask;
   S.rc_ptr = h->rc_ptr;
   S.rc_list = h->rc_list;
   S.is_ess = h->ess;
   S.Ae = h->Ae;
   S.lda = h->lda;
   S.vals = h->vals;
   S.eliminate = eliminate ? 1 : 0;
   S.maxrow = h->maxrow;
   S.slot = h->slot;
   S.sld = h->sld;
   S.rsoff = h->rsoff;
   S.rslot = h->rslot;
   S.rows = rows;
   S.nrows = nrows;
   if (nrows <= 0) { return hipSuccess; }
   const size_t shmem = (size_t)4 * h->maxrow * sizeof(double2);
   const bool bytes = h->rslot != nullptr;
   // nd <= 32: two contributors per load instruction
   const bool pack2 = h->nd <= 32;
   const bool list = rows != nullptr;
   const void *kern =
      list ? (bytes ? (pack2 ? (const void *)k_scatter_row_list<true, true>
                             : (const void *)k_scatter_row_list<true, false>)
                    : (pack2 ? (const void *)k_scatter_row_list<false, true>
                             : (const void *)k_scatter_row_list<false, false>))
           : (bytes ? (pack2 ? (const void *)k_scatter_rows<true, true>
                             : (const void *)k_scatter_rows<true, false>)
                    : (pack2 ? (const void *)k_scatter_rows<false, true>
                             : (const void *)k_scatter_rows<false, false>));
   LF_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
   if (timed) { LF_TRY(ev_begin(h, T_SCATTER, st)); }
   const dim3 grid((unsigned)((nrows + 3) / 4));
   void *args[] = {&S};
   LF_TRY(hipLaunchKernel(kern, grid, dim3(BLOCK), args, shmem, st));
   LF_TRY(hipGetLastError());
   if (timed) { LF_TRY(ev_end(h, T_SCATTER, st)); }
   return hipSuccess;
}

// Row lists of the overlapped captured step: rows formed only by
// tensor-path elements and the others, in ascending row order (built once
// per pattern and element classification, outside any capture).
hipError_t prepare_row_split(lfasm_handle *h)
{
   if (h->rows_split) { return hipSuccess; }
   std::vector<unsigned char> flag((size_t)h->height);
   unsigned char *d_flag = nullptr;
   LF_TRY(hipMalloc((void **)&d_flag, std::max<size_t>(flag.size(), 1)));
   hipLaunchKernelGGL(k_row_quad_flag, dim3((unsigned)((h->height + 255) / 256)), dim3(256), 0,
                      h->stream, h->height, h->nd, h->rc_ptr, h->rc_list, h->eclass, d_flag);
   hipError_t e = hipGetLastError();
   if (e == hipSuccess)
   {
      e = hipMemcpyAsync(flag.data(), d_flag, flag.size(), hipMemcpyDeviceToHost, h->stream);
   }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   (void)hipFree(d_flag);
   LF_TRY(e);
   std::vector<int> rt, rq;
   for (int r = 0; r < h->height; ++r) { (flag[(size_t)r] ? rq : rt).push_back(r); }
   h->nrows_t = (int)rt.size();
   h->nrows_q = (int)rq.size();
   LF_TRY(upload(h, h->rows_t, rt.data(), rt.size()));
   LF_TRY(upload(h, h->rows_q, rq.data(), rq.size()));
   h->rows_split = true;
   return hipSuccess;
}

// The overlapped step applies: materialised tensor path with PML elements,
// both element kernels side by side, a scatter cap, and a mask that the
// tensor kernel takes (not exactly one of CROSS0 / CROSS1).
bool step_overlaps(const lfasm_handle *h, const LocalParams &P)
{
   const bool one_cross = ((P.mflags >> 1) & 1) != ((P.mflags >> 2) & 1);
   return h->tensor_mode == 1 && h->nconst > 0 && h->nquad > 0 && h->conc_grid > 0 &&
          h->scat_overlap && !one_cross && (P.use_cc || P.mflags);
}

// Captured step with the scatters beside the element kernels: the tensor
// kernel, then the scatter of rows_t (rows formed only by tensor-path
// elements) on the handle's stream; the quadrature kernel on the second
// stream, then -- once the tensor kernel is done too -- the scatter of
// rows_q there, beside the rows_t scatter; joined at the end.  Each CSR row
// is written by exactly one of the two scatters.
hipError_t enqueue_step_overlap(lfasm_handle *h, const LocalParams &P, int eliminate)
{
   h->last_P = P;
   h->local_mode = 1;
   LocalParams Pq = P;
   Pq.elist = h->qlist;
   Pq.nout = h->nquad;
   Pq.out_by_elem = 1;
   LF_TRY(hipEventRecord(h->ev_fork, h->stream));
   LF_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
   LF_TRY(launch_tensor(h, P));
   LF_TRY(hipEventRecord(h->ev_tensor, h->stream));
   LF_TRY(enqueue_scatter_rows(h, eliminate, false, h->rows_t, h->nrows_t, h->stream));
   LF_TRY(launch_local(h, Pq, 0, h->stream2));
   LF_TRY(hipStreamWaitEvent(h->stream2, h->ev_tensor, 0));
   LF_TRY(enqueue_scatter_rows(h, eliminate, false, h->rows_q, h->nrows_q, h->stream2));
   LF_TRY(hipEventRecord(h->ev_join, h->stream2));
   return hipStreamWaitEvent(h->stream, h->ev_join, 0);
}

}  // namespace

// ---- C-ABI -------------------------------------------------------------------

extern "C" {

int lfasm_create(int device, lfasm_handle **out)
{
   if (!out) { return LFASM_E_ARG; }
   *out = nullptr;
   lfasm_handle *h = new lfasm_handle();
   h->device = device;
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
   {
      *out = h;
      return h->fail(LFASM_E_HIP, "no HIP device");
   }
   if (device < 0 || device >= ndev)
   {
      *out = h;
      return h->fail(LFASM_E_ARG, "device %d out of range (%d devices)", device, ndev);
   }
   *out = h;
   LF_HIP(hipSetDevice(device));
   // the static workspace of lfasm_residual_static: its address on this
   // device, with the device current (once per device)
   LF_HIP(scratch_address(device, &h->scratch));
   LF_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
   LF_HIP(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, device));
   {
      int lds = 0;
      LF_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
      h->lds_max = std::min<size_t>((size_t)lds, LDS_MAX);
   }

   for (int i = 0; i < 2 * T_N; ++i) { LF_HIP(hipEventCreate(&h->ev[i])); }
   LF_HIP(hipMalloc((void **)&h->sym_queue, 256));
   LF_HIP(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
   LF_HIP(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
   LF_HIP(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
   LF_HIP(hipEventCreateWithFlags(&h->ev_tensor, hipEventDisableTiming));
   return LFASM_OK;
}

void lfasm_destroy(lfasm_handle *h)
{
   if (!h) { return; }
   (void)hipSetDevice(h->device);
   if (h->stream) { (void)hipStreamSynchronize(h->stream); }
   release_global(h);
   release_mesh(h);
   for (int i = 0; i < 2 * T_N; ++i)
   {
      if (h->ev[i]) { (void)hipEventDestroy(h->ev[i]); }
   }
   drop_graph(h);
   for (int k = 0; k < 2; ++k)
   {
      if (h->pin[k]) { (void)hipHostFree(h->pin[k]); }
      if (h->pin_ev[k]) { (void)hipEventDestroy(h->pin_ev[k]); }
   }
   dfree(h->sym_queue);
   if (h->ev_fork) { (void)hipEventDestroy(h->ev_fork); }
   if (h->ev_join) { (void)hipEventDestroy(h->ev_join); }
   if (h->ev_tensor) { (void)hipEventDestroy(h->ev_tensor); }
   if (h->stream2) { (void)hipStreamDestroy(h->stream2); }
   if (h->stream) { (void)hipStreamDestroy(h->stream); }
   delete h;
}

const char *lfasm_last_error(const lfasm_handle *h) { return h ? h->err.c_str() : "null handle"; }

double lfasm_pml_smax(double target, double order, double k0, double n_med, double thick)
{
   return -std::log(target) * (order + 1) / (2.0 * k0 * n_med * thick);
}

}  // extern "C"

namespace {

// geometry: the mesh carries J, Jinv, detJ, v0 (else they come from vertex
// coordinates on the device and are checked there)
int validate_mesh(lfasm_handle *h, const lfasm_mesh *m, int idx, bool geometry = true)
{
   if (m->ne <= 0 || m->nd <= 0 || m->p <= 0 || m->height <= 0 || m->nreg <= 0)
   {
      return h->fail(LFASM_E_ARG, "mesh %d: ne, nd, p, height, nreg must be positive", idx);
   }
   if (m->p > 4 || m->nd != m->p * (m->p + 2) * (m->p + 3) / 2)
   {
      return h->fail(LFASM_E_UNSUPPORTED, "mesh %d: order %d / nd %d outside the kernels (p <= 4)",
                     idx, m->p, m->nd);
   }
   if (!m->dof || !m->sign || !m->attr || !m->eps_tab || !m->TInv || (m->p >= 2 && !m->fo) ||
       (geometry && (!m->J || !m->Jinv || !m->detJ || !m->v0)))
   {
      return h->fail(LFASM_E_ARG, "mesh %d: missing mesh array", idx);
   }
   int rc;
   if ((rc = check_rule(h, m->rule_cc, "cc")) != LFASM_OK) { return rc; }
   if ((rc = check_rule(h, m->rule_mass, "mass")) != LFASM_OK) { return rc; }
   // host-side validation of everything the kernels index with
   for (size_t i = 0; i < (size_t)m->ne * m->nd; ++i)
   {
      if (m->dof[i] < 0 || m->dof[i] >= m->height)
      {
         return h->fail(LFASM_E_ARG, "mesh %d: dof[%zu] = %d outside [0, %d)", idx, i, m->dof[i],
                        m->height);
      }
   }
   // the local dofs of an element are distinct global dofs: the row scatter
   // accumulates one contributor row's entries in one LDS instruction, where
   // two equal columns would lose a contribution
   {
      std::vector<int> stamp((size_t)m->height, -1);
      for (int e = 0; e < m->ne; ++e)
      {
         const int *ed = m->dof + (size_t)e * m->nd;
         for (int a = 0; a < m->nd; ++a)
         {
            if (stamp[(size_t)ed[a]] == e)
            {
               return h->fail(LFASM_E_ARG, "mesh %d: element %d holds global dof %d twice", idx,
                              e, ed[a]);
            }
            stamp[(size_t)ed[a]] = e;
         }
      }
   }
   for (int e = 0; e < m->ne; ++e)
   {
      if (m->attr[e] < 1 || m->attr[e] > m->nreg)
      {
         return h->fail(LFASM_E_ARG, "mesh %d: attr[%d] = %d outside [1, %d]", idx, e, m->attr[e],
                        m->nreg);
      }
      if (geometry && !(m->detJ[e] > 0.0))
      {
         return h->fail(LFASM_E_ARG, "mesh %d: detJ[%d] <= 0 (expected positively oriented tet)",
                        idx, e);
      }
   }
   if (m->p >= 2)
   {
      for (size_t i = 0; i < (size_t)m->ne * 4; ++i)
      {
         if (m->fo[i] < 0 || m->fo[i] > 5)
         {
            return h->fail(LFASM_E_ARG, "mesh %d: fo[%zu] = %d outside 0..5", idx, i, m->fo[i]);
         }
      }
   }
   return LFASM_OK;
}

bool same_rule(const lfasm_rule &a, const lfasm_rule &b, int nd)
{
   if (a.nq != b.nq) { return false; }
   const size_t n3 = (size_t)a.nq * 3, nt = (size_t)a.nq * nd * 3;
   return std::memcmp(a.ip, b.ip, n3 * sizeof(double)) == 0 &&
          std::memcmp(a.w, b.w, (size_t)a.nq * sizeof(double)) == 0 &&
          std::memcmp(a.vshape, b.vshape, nt * sizeof(double)) == 0 &&
          std::memcmp(a.curl, b.curl, nt * sizeof(double)) == 0;
}

template <typename T>
void append(std::vector<T> &dst, const T *src, size_t n)
{
   dst.insert(dst.end(), src, src + n);
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// lfasm_init_batch / lfasm_init_vertices: xyz (single mesh only) = element
// vertex coordinates, ne x 4 x 3, from which the geometry is computed on the
// device instead of taken from the mesh's J, Jinv, detJ, v0.
int init_impl(lfasm_handle *h, const lfasm_mesh *ms, int nmesh, const double *xyz)
{
   if (!h) { return LFASM_E_STATE; }
   if (std::getenv("LITHOFEM_ASM_FORCE_FAIL"))
   {
      return h->fail(LFASM_E_FORCED, "forced failure injected (LITHOFEM_ASM_FORCE_FAIL)");
   }
   // the input checks run on the host before any device state is needed
   if (!ms || nmesh <= 0) { return h->fail(LFASM_E_ARG, "no mesh"); }
   int rc;
   int64_t ne_tot = 0, n_tot = 0;
   for (int i = 0; i < nmesh; ++i)
   {
      if ((rc = validate_mesh(h, &ms[i], i, xyz == nullptr)) != LFASM_OK) { return rc; }
      if (ms[i].p != ms[0].p || !same_rule(ms[i].rule_cc, ms[0].rule_cc, ms[0].nd) ||
          !same_rule(ms[i].rule_mass, ms[0].rule_mass, ms[0].nd) ||
          std::memcmp(ms[i].TInv, ms[0].TInv, 24 * sizeof(double)) != 0)
      {
         return h->fail(LFASM_E_ARG, "mesh %d: batched meshes need the same order, rules and TInv", i);
      }
      ne_tot += ms[i].ne;
      n_tot += ms[i].height;
   }
   if (ne_tot > INT32_MAX || n_tot > INT32_MAX || ne_tot * ms[0].nd > INT32_MAX)
   {
      return h->fail(LFASM_E_UNSUPPORTED, "batch exceeds int32 element/dof indexing");
   }
   if (!h->stream) { return h->fail(LFASM_E_STATE, "no device stream (lfasm_create failed)"); }
   LF_HIP(hipSetDevice(h->device));
   release_global(h);
   release_mesh(h);
   const lfasm_mesh &m0 = ms[0];
   h->ne = (int)ne_tot;
   h->nd = m0.nd;
   h->p = m0.p;
   h->height = (int)n_tot;
   h->nmesh = nmesh;
   h->elem_off.assign(1, 0);
   h->row_off.assign(1, 0);
   const size_t nd = (size_t)m0.nd;
   std::vector<PmlDev> pml;
   std::vector<int> eps_off;
   std::vector<double> eps_tab;
   bool kpar = false;
   for (int i = 0; i < nmesh; ++i)
   {
      const lfasm_mesh &m = ms[i];
      pml.push_back(PmlDev{m.pml.z_min, m.pml.z_max, m.pml.thick, m.pml.order, m.pml.smax_bot,
                           m.pml.smax_top, m.pml.k0, m.pml.kx, m.pml.ky});
      kpar = kpar || m.pml.kx != 0.0 || m.pml.ky != 0.0;
      eps_off.push_back((int)(eps_tab.size() / 2));
      append(eps_tab, m.eps_tab, (size_t)m.nreg * 2);
      h->elem_off.push_back(h->elem_off.back() + m.ne);
      h->row_off.push_back(h->row_off.back() + m.height);
   }
   h->any_kpar = kpar;
   h->nreg = (int)(eps_tab.size() / 2);
   const size_t ne = (size_t)ne_tot;
   if (nmesh == 1 && xyz)
   {
      // affine geometry from the vertex coordinates, on the device
      double *dx = nullptr;
      int *bad = nullptr;
      int bad_first = INT32_MAX;
      hipError_t e = upload(h, dx, xyz, ne * 12);
      if (e == hipSuccess) { e = hipMalloc((void **)&h->J, ne * 9 * sizeof(double)); }
      if (e == hipSuccess) { e = hipMalloc((void **)&h->Ji, ne * 9 * sizeof(double)); }
      if (e == hipSuccess) { e = hipMalloc((void **)&h->detJ, ne * sizeof(double)); }
      if (e == hipSuccess) { e = hipMalloc((void **)&h->v0, ne * 3 * sizeof(double)); }
      if (e == hipSuccess) { e = upload(h, bad, &bad_first, 1); }
      if (e == hipSuccess) { e = ev_begin(h, T_GEOM); }
      if (e == hipSuccess)
      {
         hipLaunchKernelGGL(k_geometry, dim3((unsigned)std::min<size_t>((ne + 255) / 256, 8192)),
                            dim3(256), 0, h->stream, (int)ne, (const double *)dx, h->J, h->Ji,
                            h->detJ, h->v0, bad);
         e = hipGetLastError();
      }
      if (e == hipSuccess) { e = ev_end(h, T_GEOM); }
      if (e == hipSuccess) { e = hipMemcpyAsync(&bad_first, bad, sizeof(int), hipMemcpyDeviceToHost, h->stream); }
      if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
      if (dx) { (void)hipFree(dx); }
      if (bad) { (void)hipFree(bad); }
      LF_HIP(e);
      if (bad_first != INT32_MAX)
      {
         release_mesh(h);
         return h->fail(LFASM_E_ARG, "element %d: detJ <= 0 (expected positively oriented tet)",
                        bad_first);
      }
   }
   if (nmesh == 1)
   {
      if (!xyz)
      {
         LF_HIP(upload(h, h->J, m0.J, ne * 9));
         LF_HIP(upload(h, h->Ji, m0.Jinv, ne * 9));
         LF_HIP(upload(h, h->detJ, m0.detJ, ne));
         LF_HIP(upload(h, h->v0, m0.v0, ne * 3));
      }
      LF_HIP(upload(h, h->attr, m0.attr, ne));
      LF_HIP(upload(h, h->dof, m0.dof, ne * nd));
      h->dof_host.assign(m0.dof, m0.dof + ne * nd);
      LF_HIP(upload(h, h->sign, m0.sign, ne * nd));
      if (m0.p >= 2) { LF_HIP(upload(h, h->fo, m0.fo, ne * 4)); }
      if (m0.eps_elem)
      {
         LF_HIP(upload(h, h->eps_elem, (const double2 *)m0.eps_elem, ne));
         h->eps_elem_host.assign((const double2 *)m0.eps_elem, (const double2 *)m0.eps_elem + ne);
      }
   }
   else
   {
      // block-diagonal union: element arrays concatenated, dofs offset by
      // the preceding meshes' dof counts, one problem id per element
      std::vector<double> J, Ji, det, v0, sign, eel;
      std::vector<int> attr, dof, fo, pid;
      bool any_eel = false;
      for (int i = 0; i < nmesh; ++i) { any_eel = any_eel || ms[i].eps_elem; }
      for (int i = 0; i < nmesh; ++i)
      {
         const lfasm_mesh &m = ms[i];
         const size_t mne = (size_t)m.ne;
         append(J, m.J, mne * 9);
         append(Ji, m.Jinv, mne * 9);
         append(det, m.detJ, mne);
         append(v0, m.v0, mne * 3);
         append(attr, m.attr, mne);
         append(sign, m.sign, mne * nd);
         const int off = (int)h->row_off[(size_t)i];
         for (size_t k = 0; k < mne * nd; ++k) { dof.push_back(m.dof[k] + off); }
         if (m.p >= 2) { append(fo, m.fo, mne * 4); }
         pid.insert(pid.end(), mne, i);
         if (any_eel)
         {
            for (size_t e = 0; e < mne; ++e)
            {
               if (m.eps_elem) { append(eel, m.eps_elem + 2 * e, 2); }
               else { append(eel, m.eps_tab + 2 * (m.attr[e] - 1), 2); }
            }
         }
      }
      LF_HIP(upload(h, h->J, J.data(), J.size()));
      LF_HIP(upload(h, h->Ji, Ji.data(), Ji.size()));
      LF_HIP(upload(h, h->detJ, det.data(), det.size()));
      LF_HIP(upload(h, h->v0, v0.data(), v0.size()));
      LF_HIP(upload(h, h->attr, attr.data(), attr.size()));
      LF_HIP(upload(h, h->dof, dof.data(), dof.size()));
      h->dof_host = dof;
      LF_HIP(upload(h, h->sign, sign.data(), sign.size()));
      if (m0.p >= 2) { LF_HIP(upload(h, h->fo, fo.data(), fo.size())); }
      LF_HIP(upload(h, h->pid, pid.data(), pid.size()));
      if (any_eel)
      {
         LF_HIP(upload(h, h->eps_elem, (const double2 *)eel.data(), eel.size() / 2));
         h->eps_elem_host.assign((const double2 *)eel.data(),
                                 (const double2 *)eel.data() + eel.size() / 2);
      }
   }
   LF_HIP(upload_rule(h, h->rc, m0.rule_cc, m0.nd));
   LF_HIP(upload_rule(h, h->rm, m0.rule_mass, m0.nd));
   LF_HIP(upload(h, h->eps_tab, (const double2 *)eps_tab.data(), eps_tab.size() / 2));
   h->eps_tab_host.assign((const double2 *)eps_tab.data(),
                          (const double2 *)eps_tab.data() + eps_tab.size() / 2);
   LF_HIP(upload(h, h->eps_off, eps_off.data(), eps_off.size()));
   LF_HIP(upload(h, h->pml_tab, pml.data(), pml.size()));
   h->pml_host = pml;
   std::memcpy(h->TInv, m0.TInv, sizeof(h->TInv));
   LF_HIP(upload(h, h->TInv_dev, h->TInv, 24));
   // element-matrix rows padded to whole 128-byte lines (8 entries)
   h->lda = (int)((nd + 7) / 8 * 8);
   LF_HIP(hipMalloc((void **)&h->Ae, ne * nd * (size_t)h->lda * sizeof(double2)));
   // element classes: s(z) = 1 on the closed slab [z_min, z_max], so an
   // element whose four vertices lie in it has Lambda = I at every
   // quadrature point (the points are convex combinations of the vertices)
   {
      std::vector<int> cl, ql;
      int64_t base = 0;
      for (int i = 0; i < nmesh; ++i)
      {
         const lfasm_mesh &m = ms[i];
         for (int e = 0; e < m.ne; ++e)
         {
            bool inside = true;
            for (int v = 0; v < 4; ++v)
            {
               const double z = xyz ? xyz[(size_t)e * 12 + v * 3 + 2]
                                    : (v == 0 ? m.v0[(size_t)e * 3 + 2]
                                              : m.v0[(size_t)e * 3 + 2] + m.J[(size_t)e * 9 + 6 + v - 1]);
               inside = inside && z >= m.pml.z_min && z <= m.pml.z_max;
            }
            (inside ? cl : ql).push_back((int)(base + e));
         }
         base += m.ne;
      }
      h->nconst = (int)cl.size();
      h->nquad = (int)ql.size();
      LF_HIP(upload(h, h->clist, cl.data(), cl.size()));
      LF_HIP(upload(h, h->qlist, ql.data(), ql.size()));
      std::vector<unsigned char> ec(ne, 0);
      for (int e : cl) { ec[(size_t)e] = 1; }
      LF_HIP(upload(h, h->eclass, ec.data(), ec.size()));
      h->eclass_host = ec;
      LF_HIP(hipMalloc((void **)&h->coef, ne * TEN_NC * sizeof(double)));
      const int npair = (m0.nd + 1) / 2;
      const size_t nR = (size_t)npair * npair * NTENS * 4;
      LF_HIP(hipMalloc((void **)&h->Rten, nR * sizeof(double)));
      const RuleDev rc{h->rc.nq, h->rc.ip, h->rc.w, h->rc.vs, h->rc.cs};
      const RuleDev rm{h->rm.nq, h->rm.ip, h->rm.w, h->rm.vs, h->rm.cs};
      hipLaunchKernelGGL(k_ref_tensor, dim3((unsigned)std::min<size_t>((nR + 255) / 256, 4096)),
                         dim3(256), 0, h->stream, m0.nd, npair, rc, rm, h->Rten);
      LF_HIP(hipGetLastError());
      LF_HIP(hipStreamSynchronize(h->stream));
   }
   h->inited = true;
   return LFASM_OK;
}

}  // namespace

extern "C" {

int lfasm_init_batch(lfasm_handle *h, const lfasm_mesh *ms, int nmesh)
{
   return init_impl(h, ms, nmesh, nullptr);
}

int lfasm_init(lfasm_handle *h, const lfasm_mesh *m) { return init_impl(h, m, 1, nullptr); }

int lfasm_init_vertices(lfasm_handle *h, const lfasm_mesh *m, const double *elem_xyz)
{
   if (!h) { return LFASM_E_STATE; }
   if (!elem_xyz) { return h->fail(LFASM_E_ARG, "vertex coordinates missing"); }
   return init_impl(h, m, 1, elem_xyz);
}

int lfasm_download_geometry(lfasm_handle *h, double *J, double *Jinv, double *detJ, double *v0)
{
   if (!h || !h->inited) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   const size_t ne = (size_t)h->ne;
   if (J) { LF_HIP(hipMemcpyAsync(J, h->J, ne * 9 * sizeof(double), hipMemcpyDeviceToHost, h->stream)); }
   if (Jinv) { LF_HIP(hipMemcpyAsync(Jinv, h->Ji, ne * 9 * sizeof(double), hipMemcpyDeviceToHost, h->stream)); }
   if (detJ) { LF_HIP(hipMemcpyAsync(detJ, h->detJ, ne * sizeof(double), hipMemcpyDeviceToHost, h->stream)); }
   if (v0) { LF_HIP(hipMemcpyAsync(v0, h->v0, ne * 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream)); }
   LF_HIP(hipStreamSynchronize(h->stream));
   return LFASM_OK;
}

int lfasm_download_contributors(lfasm_handle *h, int *rc_ptr, int *rc_list)
{
   if (!h || !h->global_built) { return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   if (rc_ptr)
   {
      LF_HIP(hipMemcpyAsync(rc_ptr, h->rc_ptr, ((size_t)h->height + 1) * sizeof(int),
                            hipMemcpyDeviceToHost, h->stream));
   }
   if (rc_list)
   {
      LF_HIP(hipMemcpyAsync(rc_list, h->rc_list, (size_t)h->ne * h->nd * sizeof(int),
                            hipMemcpyDeviceToHost, h->stream));
   }
   LF_HIP(hipStreamSynchronize(h->stream));
   return LFASM_OK;
}

int lfasm_batch_offsets(const lfasm_handle *h, int64_t *elem_off, int64_t *row_off)
{
   if (!h || !h->inited) { return LFASM_E_STATE; }
   for (int i = 0; i <= h->nmesh; ++i)
   {
      if (elem_off) { elem_off[i] = h->elem_off[(size_t)i]; }
      if (row_off) { row_off[i] = h->row_off[(size_t)i]; }
   }
   return h->nmesh;
}

int lfasm_set_eps_elem(lfasm_handle *h, const double *eps)
{
   if (!h || !h->inited) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   drop_graph(h);
   // the RHS source elements depend on the per-element eps: rebuilt at the
   // next lfasm_assemble_rhs
   h->rhs_nsrc = -1;
   if (!eps)
   {
      dfree(h->eps_elem);
      h->eps_elem_host.clear();
      return LFASM_OK;
   }
   LF_HIP(upload(h, h->eps_elem, (const double2 *)eps, (size_t)h->ne));
   h->eps_elem_host.assign((const double2 *)eps, (const double2 *)eps + h->ne);
   return LFASM_OK;
}

}  // extern "C"

namespace {

// Exclusive prefix sum of n int64 values into n + 1 (device, hipCUB);
// returns the total through *total (host).
hipError_t scan64(lfasm_handle *h, const int64_t *in, int64_t *out, int n, int64_t *total)
{
   size_t tmp_bytes = 0;
   hipError_t e = hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, in, out + 1, n, h->stream);
   if (e != hipSuccess) { return e; }
   void *tmp = nullptr;
   if ((e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1))) != hipSuccess) { return e; }
   e = hipMemsetAsync(out, 0, sizeof(int64_t), h->stream);
   if (e == hipSuccess) { e = hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, in, out + 1, n, h->stream); }
   if (e == hipSuccess)
   {
      e = hipMemcpyAsync(total, out + n, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream);
   }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   (void)hipFree(tmp);
   return e;
}

// The fused row units (tensor path mode 2) from host copies of the pattern
// pieces they need; built only when that mode is selected.
int ensure_fused_units(lfasm_handle *h)
{
   if (h->fused_tried || !h->global_built || !h->coef || h->p > 3) { return LFASM_OK; }
   h->fused_tried = true;
   const int n = h->height;
   const size_t nk = (size_t)h->ne * h->nd;
   std::vector<int> rc_ptr((size_t)n + 1), rc_list(nk);
   std::vector<double> sign(nk);
   std::vector<unsigned char> ess((size_t)n);
   LF_HIP(hipMemcpy(rc_ptr.data(), h->rc_ptr, rc_ptr.size() * sizeof(int), hipMemcpyDeviceToHost));
   LF_HIP(hipMemcpy(rc_list.data(), h->rc_list, nk * sizeof(int), hipMemcpyDeviceToHost));
   LF_HIP(hipMemcpy(sign.data(), h->sign, nk * sizeof(double), hipMemcpyDeviceToHost));
   LF_HIP(hipMemcpy(ess.data(), h->ess, ess.size(), hipMemcpyDeviceToHost));
   return build_fused_units(h, h->I_host.data(), h->dof_host, rc_ptr, rc_list, sign, ess);
}

// Everything of build_global after the pattern (h->Ip, h->Jc, h->rc_ptr,
// h->rc_list on the device, h->I_host on the host): row-length limits,
// essential mask, sign / essential flags and masks, slot map, the per-row
// byte slot blocks -- all on the device.
int finish_global(lfasm_handle *h, const unsigned char *is_ess)
{
   const int n = h->height;
   const size_t ne = (size_t)h->ne, nd = (size_t)h->nd;
   const int64_t nk = (int64_t)(ne * nd);
   const std::vector<int> &I = h->I_host;
   int maxrow = 0;
   for (int r = 0; r < n; ++r) { maxrow = std::max(maxrow, I[(size_t)r + 1] - I[(size_t)r]); }
   if (maxrow > 65535)
   {
      release_global(h);
      return h->fail(LFASM_E_UNSUPPORTED, "row length %d > 65535", maxrow);
   }
   // the row scatter keeps 4 rows per workgroup in LDS
   int lds_max = 0;
   LF_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
   if ((size_t)4 * maxrow * sizeof(double2) > (size_t)lds_max)
   {
      release_global(h);
      return h->fail(LFASM_E_UNSUPPORTED, "row length %d > %d (LDS row buffers of the scatter)",
                     maxrow, lds_max / (int)(4 * sizeof(double2)));
   }
   h->maxrow = maxrow;
   if (is_ess) { LF_HIP(upload(h, h->ess, is_ess, (size_t)n)); }
   else
   {
      LF_HIP(hipMalloc((void **)&h->ess, (size_t)std::max(n, 1)));
      LF_HIP(hipMemsetAsync(h->ess, 0, (size_t)n, h->stream));
   }
   const unsigned grid = (unsigned)std::min<int64_t>((nk + 255) / 256, 8192);
   // per (element, local dof) flags: bit0 sign < 0, bit1 essential dof
   LF_HIP(hipMalloc((void **)&h->eflag, (size_t)std::max<int64_t>(nk, 1)));
   hipLaunchKernelGGL(k_eflag, dim3(grid), dim3(256), 0, h->stream, nk, h->sign, h->dof, h->ess,
                      h->eflag);
   LF_HIP(hipGetLastError());
   if (nd <= 64)
   {
      LF_HIP(hipMalloc((void **)&h->emask, 2 * ne * sizeof(uint64_t)));
      hipLaunchKernelGGL(k_emask, dim3((unsigned)std::min<size_t>((ne + 3) / 4, 8192)), dim3(256),
                         0, h->stream, h->ne, h->nd, h->eflag, h->emask);
      LF_HIP(hipGetLastError());
   }
   LF_HIP(hipMalloc((void **)&h->vals, (size_t)h->nnz * sizeof(double2)));
   // rows unpadded: padding them to aligned 128-byte lines measured no
   // faster (1.132 vs 1.135 ms at 1.35M DOF) for 42% more memory
   h->sld = (int)nd;
   LF_HIP(hipMalloc((void **)&h->slot, ne * nd * (size_t)h->sld * sizeof(uint16_t)));
   LF_HIP(hipMalloc((void **)&h->bad, sizeof(int)));
   LF_HIP(hipMemsetAsync(h->bad, 0, sizeof(int), h->stream));
   LF_HIP(ev_begin(h, T_BUILD));
   hipLaunchKernelGGL(k_slot_map, dim3(2048), dim3(BLOCK), 0, h->stream, h->ne, h->nd, h->sld,
                      h->dof, h->Ip, h->Jc, h->slot, h->bad);
   LF_HIP(hipGetLastError());
   LF_HIP(ev_end(h, T_BUILD));
   int bad = 0;
   LF_HIP(hipMemcpyAsync(&bad, h->bad, sizeof(int), hipMemcpyDeviceToHost, h->stream));
   LF_HIP(hipStreamSynchronize(h->stream));
   if (bad)
   {
      release_global(h);
      return h->fail(LFASM_E_PATTERN, "an element (row, col) pair is missing from the CSR pattern");
   }
   if (maxrow <= 256)
   {
      // byte slot rows regrouped per CSR row, in the row's contributor order
      // (every row has at most 256 entries): the scatter reads a row's slot
      // rows from one contiguous block
      int64_t *len = nullptr, *off = nullptr;
      int64_t total = 0;
      hipError_t e = hipMalloc((void **)&len, (size_t)n * sizeof(int64_t));
      if (e == hipSuccess) { e = hipMalloc((void **)&off, ((size_t)n + 1) * sizeof(int64_t)); }
      if (e == hipSuccess)
      {
         hipLaunchKernelGGL(k_slot_block_len, dim3((unsigned)std::min(8192, (n + 255) / 256)),
                            dim3(256), 0, h->stream, n, h->nd, h->rc_ptr, len);
         e = hipGetLastError();
      }
      if (e == hipSuccess) { e = scan64(h, len, off, n, &total); }
      if (e == hipSuccess && total <= INT32_MAX)
      {
         e = hipMalloc((void **)&h->rsoff, ((size_t)n + 1) * sizeof(int));
         if (e == hipSuccess)
         {
            hipLaunchKernelGGL(k_narrow, dim3((unsigned)std::min(8192, (n + 256) / 256)), dim3(256),
                               0, h->stream, (int64_t)n + 1, (const int64_t *)off, h->rsoff);
            e = hipGetLastError();
         }
         if (e == hipSuccess) { e = hipMalloc((void **)&h->rslot, std::max<size_t>((size_t)total, 1)); }
      }
      if (len) { (void)hipFree(len); }
      if (off) { (void)hipFree(off); }
      LF_HIP(e);
      if (total > INT32_MAX)
      {
         release_global(h);
         return h->fail(LFASM_E_UNSUPPORTED, "slot blocks exceed 2^31 entries");
      }
      hipLaunchKernelGGL(k_row_slots, dim3((unsigned)((n + 3) / 4)), dim3(BLOCK), 0, h->stream, n,
                         (int)nd, h->sld, (const int *)h->rc_ptr, (const int *)h->rc_list,
                         (const uint16_t *)h->slot, (const int *)h->rsoff, h->rslot);
      LF_HIP(hipGetLastError());
   }
   h->global_built = true;
   if (h->tensor_mode == 2)
   {
      const int rc = ensure_fused_units(h);
      if (rc != LFASM_OK) { release_global(h); return rc; }
   }
   return LFASM_OK;
}

// Per-row contributor lists (ascending element, then local index) of the
// element dof table on the device: histogram, scan, scattered fill, and a
// sort of each row's segment.  Returns the largest row segment.
int device_contributors(lfasm_handle *h, int *maxc_out)
{
   const int n = h->height;
   const int64_t nk = (int64_t)h->ne * h->nd;
   const unsigned grid = (unsigned)std::min<int64_t>((nk + 255) / 256, 8192);
   const unsigned rgrid = (unsigned)std::min(8192, (n + 255) / 256);
   int *cnt = nullptr, *maxc = nullptr;
   int64_t *cnt64 = nullptr, *ptr64 = nullptr;
   int64_t total = 0;
   hipError_t e = hipMalloc((void **)&cnt, ((size_t)n + 1) * sizeof(int));
   if (e == hipSuccess) { e = hipMalloc((void **)&maxc, sizeof(int)); }
   if (e == hipSuccess) { e = hipMalloc((void **)&cnt64, (size_t)n * sizeof(int64_t)); }
   if (e == hipSuccess) { e = hipMalloc((void **)&ptr64, ((size_t)n + 1) * sizeof(int64_t)); }
   if (e == hipSuccess) { e = hipMalloc((void **)&h->rc_ptr, ((size_t)n + 1) * sizeof(int)); }
   if (e == hipSuccess) { e = hipMalloc((void **)&h->rc_list, (size_t)std::max<int64_t>(nk, 1) * sizeof(int)); }
   if (e == hipSuccess) { e = hipMemsetAsync(cnt, 0, ((size_t)n + 1) * sizeof(int), h->stream); }
   if (e == hipSuccess) { e = hipMemsetAsync(maxc, 0, sizeof(int), h->stream); }
   if (e == hipSuccess)
   {
      hipLaunchKernelGGL(k_contrib_count, dim3(grid), dim3(256), 0, h->stream, nk, h->dof, cnt);
      e = hipGetLastError();
   }
   if (e == hipSuccess)
   {
      // the counts widened to 64 bits for the scan (nk fits in int32, checked at init)
      hipLaunchKernelGGL(k_widen, dim3(rgrid), dim3(256), 0, h->stream, n, (const int *)cnt, cnt64);
      e = hipGetLastError();
   }
   if (e == hipSuccess) { e = scan64(h, cnt64, ptr64, n, &total); }
   if (e == hipSuccess)
   {
      hipLaunchKernelGGL(k_narrow, dim3((unsigned)std::min(8192, (n + 256) / 256)), dim3(256), 0,
                         h->stream, (int64_t)n + 1, (const int64_t *)ptr64, h->rc_ptr);
      e = hipGetLastError();
   }
   if (e == hipSuccess) { e = hipMemsetAsync(cnt, 0, ((size_t)n + 1) * sizeof(int), h->stream); }
   if (e == hipSuccess)
   {
      hipLaunchKernelGGL(k_contrib_fill, dim3(grid), dim3(256), 0, h->stream, nk, h->dof,
                         (const int *)h->rc_ptr, cnt, h->rc_list);
      e = hipGetLastError();
   }
   if (e == hipSuccess)
   {
      hipLaunchKernelGGL(k_contrib_sort, dim3(rgrid), dim3(256), 0, h->stream, n,
                         (const int *)h->rc_ptr, h->rc_list, maxc);
      e = hipGetLastError();
   }
   if (e == hipSuccess) { e = hipMemcpyAsync(maxc_out, maxc, sizeof(int), hipMemcpyDeviceToHost, h->stream); }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   for (void *q : {(void *)cnt, (void *)maxc, (void *)cnt64, (void *)ptr64})
   {
      if (q) { (void)hipFree(q); }
   }
   LF_HIP(e);
   if (total != nk) { return h->fail(LFASM_E_STATE, "contributor count %lld != %lld", (long long)total, (long long)nk); }
   return LFASM_OK;
}

}  // namespace

extern "C" {

int lfasm_build_global(lfasm_handle *h, const int *I, const int *Jcol, int64_t nnz,
                       const unsigned char *is_ess)
{
   if (!h || !h->inited) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   if (!I || !Jcol || nnz <= 0 || nnz > INT32_MAX)
   {
      return h->fail(LFASM_E_ARG, "CSR pattern missing or nnz out of int32 range");
   }
   const int n = h->height;
   if (I[0] != 0 || I[n] != nnz) { return h->fail(LFASM_E_ARG, "I[0] != 0 or I[n] != nnz"); }
   for (int r = 0; r < n; ++r)
   {
      if (I[r + 1] < I[r]) { return h->fail(LFASM_E_ARG, "I not monotone at row %d", r); }
   }
   for (int64_t k = 0; k < nnz; ++k)
   {
      if (Jcol[k] < 0 || Jcol[k] >= n) { return h->fail(LFASM_E_ARG, "Jcol[%lld] out of range", (long long)k); }
   }
   LF_HIP(hipSetDevice(h->device));
   release_global(h);
   h->nnz = nnz;
   h->I_host.assign(I, I + (size_t)n + 1);
   LF_HIP(upload(h, h->Ip, I, (size_t)n + 1));
   LF_HIP(upload(h, h->Jc, Jcol, (size_t)nnz));
   int maxc = 0;
   const int rc = device_contributors(h, &maxc);
   if (rc != LFASM_OK) { release_global(h); return rc; }
   return finish_global(h, is_ess);
}

int lfasm_build_global_device(lfasm_handle *h, const unsigned char *is_ess)
{
   if (!h || !h->inited) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   release_global(h);
   const int n = h->height, nd = h->nd;
   int maxc = 0;
   int rc = device_contributors(h, &maxc);
   if (rc != LFASM_OK) { release_global(h); return rc; }
   int cap = 64;
   while (cap < maxc * nd) { cap <<= 1; }
   const size_t shmem = (size_t)PAT_WAVES * cap * sizeof(int);
   if (shmem > h->lds_max)
   {
      release_global(h);
      return h->fail(LFASM_E_UNSUPPORTED, "%d contributors x %d dofs per row exceed the LDS sort",
                     maxc, nd);
   }
   LF_HIP(hipMalloc((void **)&h->Ip, ((size_t)n + 1) * sizeof(int)));
   LF_HIP(hipFuncSetAttribute((const void *)k_csr_rows,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
   const dim3 grid((unsigned)((n + PAT_WAVES - 1) / PAT_WAVES));
   int *cnt = nullptr;
   int64_t *cnt64 = nullptr, *ptr64 = nullptr;
   int64_t nnz = 0;
   hipError_t e = hipMalloc((void **)&cnt, (size_t)n * sizeof(int));
   if (e == hipSuccess) { e = hipMalloc((void **)&cnt64, (size_t)n * sizeof(int64_t)); }
   if (e == hipSuccess) { e = hipMalloc((void **)&ptr64, ((size_t)n + 1) * sizeof(int64_t)); }
   if (e == hipSuccess) { e = ev_begin(h, T_PATTERN); }
   if (e == hipSuccess)
   {
      hipLaunchKernelGGL(k_csr_rows, grid, dim3(64 * PAT_WAVES), shmem, h->stream, n, nd, h->dof,
                         h->rc_ptr, h->rc_list, cap, 0, cnt, (int *)nullptr);
      e = hipGetLastError();
   }
   if (e == hipSuccess)
   {
      hipLaunchKernelGGL(k_widen, dim3((unsigned)std::min(8192, (n + 255) / 256)), dim3(256), 0,
                         h->stream, n, (const int *)cnt, cnt64);
      e = hipGetLastError();
   }
   // row pointers: a device scan of the counts (64-bit, checked against int32)
   if (e == hipSuccess) { e = scan64(h, cnt64, ptr64, n, &nnz); }
   if (e == hipSuccess && nnz <= INT32_MAX)
   {
      hipLaunchKernelGGL(k_narrow, dim3((unsigned)std::min(8192, (n + 256) / 256)), dim3(256), 0,
                         h->stream, (int64_t)n + 1, (const int64_t *)ptr64, h->Ip);
      e = hipGetLastError();
      if (e == hipSuccess) { e = hipMalloc((void **)&h->Jc, (size_t)std::max<int64_t>(nnz, 1) * sizeof(int)); }
      if (e == hipSuccess)
      {
         hipLaunchKernelGGL(k_csr_rows, grid, dim3(64 * PAT_WAVES), shmem, h->stream, n, nd, h->dof,
                            h->rc_ptr, h->rc_list, cap, 1, h->Ip, h->Jc);
         e = hipGetLastError();
      }
      if (e == hipSuccess) { e = ev_end(h, T_PATTERN); }
   }
   for (void *q : {(void *)cnt, (void *)cnt64, (void *)ptr64})
   {
      if (q) { (void)hipFree(q); }
   }
   if (e != hipSuccess) { release_global(h); }
   LF_HIP(e);
   if (nnz > INT32_MAX) { release_global(h); return h->fail(LFASM_E_UNSUPPORTED, "nnz > int32"); }
   h->nnz = nnz;
   // the host copy of the row pointers (n + 1 ints) kept for the variants
   h->I_host.resize((size_t)n + 1);
   LF_HIP(hipMemcpyAsync(h->I_host.data(), h->Ip, ((size_t)n + 1) * sizeof(int),
                         hipMemcpyDeviceToHost, h->stream));
   LF_HIP(hipStreamSynchronize(h->stream));
   return finish_global(h, is_ess);
}

int lfasm_download_pattern(lfasm_handle *h, int *I, int *Jcol)
{
   if (!h || !h->global_built) { return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   if (I)
   {
      LF_HIP(hipMemcpyAsync(I, h->Ip, ((size_t)h->height + 1) * sizeof(int), hipMemcpyDeviceToHost,
                            h->stream));
   }
   if (Jcol)
   {
      LF_HIP(hipMemcpyAsync(Jcol, h->Jc, (size_t)h->nnz * sizeof(int), hipMemcpyDeviceToHost,
                            h->stream));
   }
   LF_HIP(hipStreamSynchronize(h->stream));
   return LFASM_OK;
}

int lfasm_compute_local_kernel(lfasm_handle *h, int mask_cc, int mask_mass, int apply_te,
                               int kernel)
{
   if (!h || !h->inited) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   if (kernel < 0 || kernel >= 2 + N_MFMA_VARIANTS)
   {
      return h->fail(LFASM_E_ARG, "kernel selector %d outside [0, %d)", kernel, 2 + N_MFMA_VARIANTS);
   }
   LF_HIP(hipSetDevice(h->device));
   const LocalParams P = local_params(h, mask_cc, mask_mass, apply_te, h->Ae);
   LF_HIP(enqueue_local(h, P, kernel, true));
   LF_HIP(hipStreamSynchronize(h->stream));
   h->local_valid = true;
   return LFASM_OK;
}

int lfasm_set_local_kernel(lfasm_handle *h, int kernel)
{
   if (!h) { return LFASM_E_STATE; }
   if (kernel < 0 || kernel >= 2 + N_MFMA_VARIANTS)
   {
      return h->fail(LFASM_E_ARG, "kernel selector %d outside [0, %d)", kernel, 2 + N_MFMA_VARIANTS);
   }
   h->local_kernel = kernel;
   drop_graph(h);
   return LFASM_OK;
}

int lfasm_set_element_concurrency(lfasm_handle *h, int grid, int scatter_overlap)
{
   if (!h) { return LFASM_E_STATE; }
   if (grid < 0 || scatter_overlap < 0 || scatter_overlap > 1)
   {
      return h->fail(LFASM_E_ARG, "concurrency grid %d / scatter overlap %d out of range", grid,
                     scatter_overlap);
   }
   h->conc_grid = grid;
   h->scat_overlap = scatter_overlap;
   drop_graph(h);
   return LFASM_OK;
}

int lfasm_set_quad_residency(lfasm_handle *h, int wg_per_cu)
{
   if (!h) { return LFASM_E_STATE; }
   if (wg_per_cu < 0 || wg_per_cu > 8)
   {
      return h->fail(LFASM_E_ARG, "quadrature residency %d out of range", wg_per_cu);
   }
   h->quad_res = wg_per_cu;
   drop_graph(h);
   return LFASM_OK;
}

int lfasm_set_variant_kernel(lfasm_handle *h, int kernel)
{
   if (!h) { return LFASM_E_STATE; }
   if (kernel < 0 || kernel > 3) { return h->fail(LFASM_E_ARG, "variant kernel %d out of range", kernel); }
   h->variant_kernel = kernel;
   return LFASM_OK;
}

int lfasm_set_tensor_path(lfasm_handle *h, int mode)
{
   if (!h || !h->inited) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   if (mode < 0 || mode > 2) { return h->fail(LFASM_E_ARG, "tensor path mode %d outside 0..2", mode); }
   h->tensor_mode = mode;
   drop_graph(h);
   if (mode == 2)
   {
      const int rc = ensure_fused_units(h);
      if (rc != LFASM_OK) { return rc; }
   }
   return mode ? h->nconst : 0;
}

int lfasm_step_capture(lfasm_handle *h, int mask_cc, int mask_mass, int apply_te, int eliminate)
{
   if (!h || !h->global_built)
   {
      return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE;
   }
   LF_HIP(hipSetDevice(h->device));
   drop_graph(h);
   h->g_have = true;
   h->g_mask_cc = mask_cc;
   h->g_mask_mass = mask_mass;
   h->g_apply_te = apply_te;
   h->g_elim = eliminate;
   const LocalParams P = local_params(h, mask_cc, mask_mass, apply_te, h->Ae);
   const bool overlap = step_overlaps(h, P);
   if (overlap) { LF_HIP(prepare_row_split(h)); }
   LF_HIP(hipStreamSynchronize(h->stream));
   LF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
   hipError_t e = hipSuccess;
   if (overlap) { e = enqueue_step_overlap(h, P, eliminate); }
   else
   {
      e = enqueue_local(h, P, 0, false);
      if (e == hipSuccess) { e = enqueue_scatter(h, eliminate, false); }
   }
   hipGraph_t g = nullptr;
   const hipError_t e2 = hipStreamEndCapture(h->stream, &g);
   if (e == hipSuccess) { e = e2; }
   if (e == hipSuccess) { h->graph = g; }
   else if (g) { (void)hipGraphDestroy(g); }
   LF_HIP(e);
   LF_HIP(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
   h->graph_dirty = false;
   return LFASM_OK;
}

int lfasm_step_launch(lfasm_handle *h, int sync)
{
   if (!h || !h->global_built)
   {
      return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE;
   }
   if (!h->g_have) { return h->fail(LFASM_E_STATE, "lfasm_step_capture first"); }
   LF_HIP(hipSetDevice(h->device));
   if (h->graph_dirty)
   {
      const int rc = lfasm_step_capture(h, h->g_mask_cc, h->g_mask_mass, h->g_apply_te, h->g_elim);
      if (rc != LFASM_OK) { return rc; }
   }
   LF_HIP(ev_begin(h, T_STEP));
   LF_HIP(hipGraphLaunch(h->graph_exec, h->stream));
   LF_HIP(ev_end(h, T_STEP));
   h->local_valid = true;
   if (sync) { LF_HIP(hipStreamSynchronize(h->stream)); }
   return LFASM_OK;
}

int lfasm_compute_local(lfasm_handle *h, int mask_cc, int mask_mass, int apply_te)
{
   return lfasm_compute_local_kernel(h, mask_cc, mask_mass, apply_te, 0);
}

int lfasm_scatter_eliminate(lfasm_handle *h, int eliminate)
{
   if (!h || !h->global_built || !h->local_valid)
   {
      return h ? h->fail(LFASM_E_STATE, "build_global and compute_local first") : LFASM_E_STATE;
   }
   LF_HIP(hipSetDevice(h->device));
   LF_HIP(enqueue_scatter(h, eliminate, true));
   LF_HIP(hipStreamSynchronize(h->stream));
   return LFASM_OK;
}

int lfasm_device_csr(lfasm_handle *h, int **I, int **Jcol, void **vals)
{
   if (!h || !h->global_built) { return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE; }
   if (I) { *I = h->Ip; }
   if (Jcol) { *Jcol = h->Jc; }
   if (vals) { *vals = (void *)h->vals; }
   return LFASM_OK;
}

int lfasm_download_values(lfasm_handle *h, double *re, double *im)
{
   if (!h || !h->global_built) { return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   return download_split(h, h->vals, (size_t)h->nnz, re, im);
}

int lfasm_download_local(lfasm_handle *h, double *Ae)
{
   if (!h || !h->inited || !Ae) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   std::printf("asm_gpu_test_download (test-only D2H)\n");
   std::fflush(stdout);
   LF_HIP(hipSetDevice(h->device));
   if (h->local_valid && h->local_mode == 2)
   {
      // the fused path never wrote the constant-coefficient elements' Ae
      LF_HIP(launch_tensor(h, h->last_P));
   }
   return download_unpad(h, h->Ae, (size_t)h->ne * h->nd, h->nd, h->lda, Ae);
}

int lfasm_residual(lfasm_handle *h, const double *X2n, const double *B2n, double *rel)
{
   if (!h || !h->global_built) { return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE; }
   if (!X2n || !B2n || !rel) { return h->fail(LFASM_E_ARG, "null vector"); }
   LF_HIP(hipSetDevice(h->device));
   const int n = h->height;
   std::vector<double2> hx((size_t)n), hy((size_t)n);
   for (int i = 0; i < n; ++i) { hx[(size_t)i] = make_double2(X2n[i], X2n[n + i]); }
   double2 *dx = nullptr, *dy = nullptr;
   LF_HIP(hipMalloc((void **)&dx, (size_t)n * sizeof(double2)));
   hipError_t e = hipMalloc((void **)&dy, (size_t)n * sizeof(double2));
   if (e == hipSuccess)
   {
      e = hipMemcpyAsync(dx, hx.data(), (size_t)n * sizeof(double2), hipMemcpyHostToDevice, h->stream);
   }
   if (e == hipSuccess)
   {
      hipLaunchKernelGGL(k_spmv, dim3((n + 255) / 256), dim3(256), 0, h->stream, n, h->Ip, h->Jc,
                         h->vals, dx, dy);
      e = hipGetLastError();
   }
   if (e == hipSuccess)
   {
      e = hipMemcpyAsync(hy.data(), dy, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, h->stream);
   }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   (void)hipFree(dx);
   if (dy) { (void)hipFree(dy); }
   LF_HIP(e);
   double rn = 0.0, bn = 0.0;
   for (int i = 0; i < n; ++i)
   {
      const double dr = hy[(size_t)i].x - B2n[i], di = hy[(size_t)i].y - B2n[n + i];
      rn += dr * dr + di * di;
      bn += B2n[i] * B2n[i] + B2n[n + i] * B2n[n + i];
   }
   *rel = bn > 0.0 ? std::sqrt(rn / bn) : 0.0;
   return LFASM_OK;
}

int lfasm_residual_static(lfasm_handle *h, int64_t len, const double *X2n, const double *B2n,
                          double *rel)
{
   if (!h) { return LFASM_E_STATE; }
   if (!X2n || !B2n || !rel) { return h->fail(LFASM_E_ARG, "null vector"); }
   if (len <= 0 || (len & 1)) { return h->fail(LFASM_E_ARG, "vector length %lld is not 2n", (long long)len); }
   const int64_t n = len / 2;
   if (n > RS_MAX_N)
   {
      return h->fail(LFASM_E_ARG,
                     "n = %lld dofs exceeds the static workspace of lfasm_residual_static "
                     "(at most %d dofs; lfasm_residual has no limit)", (long long)n, RS_MAX_N);
   }
   if (!h->global_built) { return h->fail(LFASM_E_STATE, "build_global first"); }
   if (n != h->height)
   {
      return h->fail(LFASM_E_ARG, "vector length 2 x %lld does not match the system (n = %d)",
                     (long long)n, h->height);
   }
   if (!h->scratch) { return h->fail(LFASM_E_STATE, "static workspace not resolved on this device"); }
   LF_HIP(hipSetDevice(h->device));
   double *ws = h->scratch;
   // [x_re; x_im] and [b_re; b_im] as the caller holds them
   LF_HIP(hipMemcpyAsync(ws, X2n, (size_t)len * sizeof(double), hipMemcpyHostToDevice, h->stream));
   LF_HIP(hipMemcpyAsync(ws + len, B2n, (size_t)len * sizeof(double), hipMemcpyHostToDevice,
                         h->stream));
   // every partial the final kernel reads is written by this call: nothing
   // in the workspace needs clearing
   const int nb = (int)std::min<int64_t>((n + 255) / 256, RS_NPART);
   hipLaunchKernelGGL(k_resid_partial, dim3((unsigned)nb), dim3(256), 0, h->stream, (int)n, h->Ip,
                      h->Jc, h->vals, (const double *)ws, ws + RS_PART_OFF);
   LF_HIP(hipGetLastError());
   hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(256), 0, h->stream, nb,
                      (const double *)(ws + RS_PART_OFF), ws + RS_OUT_OFF);
   LF_HIP(hipGetLastError());
   double out[2] = {0.0, 0.0};
   LF_HIP(hipMemcpyAsync(out, ws + RS_OUT_OFF, sizeof(out), hipMemcpyDeviceToHost, h->stream));
   LF_HIP(hipStreamSynchronize(h->stream));
   *rel = out[1] > 0.0 ? std::sqrt(out[0] / out[1]) : 0.0;
   return LFASM_OK;
}

int lfasm_sizes(const lfasm_handle *h, int *ne, int *nd, int *n, int64_t *nnz)
{
   if (!h) { return LFASM_E_STATE; }
   if (ne) { *ne = h->ne; }
   if (nd) { *nd = h->nd; }
   if (n) { *n = h->height; }
   if (nnz) { *nnz = h->nnz; }
   return LFASM_OK;
}

int lfasm_timings(const lfasm_handle *hc, double *ms, int n)
{
   lfasm_handle *h = const_cast<lfasm_handle *>(hc);
   if (!h || !ms) { return LFASM_E_ARG; }
   for (int t = 0; t < n && t < T_N; ++t)
   {
      ms[t] = 0.0;
      if (!h->ev_used[t]) { continue; }
      LF_HIP(hipEventSynchronize(h->ev[2 * t + 1]));
      float f = 0.0f;
      LF_HIP(hipEventElapsedTime(&f, h->ev[2 * t], h->ev[2 * t + 1]));
      ms[t] = f;
   }
   return LFASM_OK;
}


// ---- eps variants ---------------------------------------------------------------

}  // extern "C"

namespace {

// Host-to-device copy of pageable memory through the handle's two pinned
// 64 MB buffers: the host copy of chunk k+1 (threads) overlaps the DMA of
// chunk k.  Ordered on the handle's stream.
int upload_pinned(lfasm_handle *h, void *dst, const void *src, size_t bytes, hipStream_t st = nullptr)
{
   if (!st) { st = h->stream; }
   constexpr size_t CH = ((size_t)1 << 22) * sizeof(double2);
   for (int k = 0; k < 2; ++k)
   {
      if (!h->pin[k]) { LF_HIP(hipHostMalloc((void **)&h->pin[k], CH)); }
      if (!h->pin_ev[k]) { LF_HIP(hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming)); }
   }
   const unsigned nthr = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
   const size_t nch = (bytes + CH - 1) / CH;
   for (size_t c = 0; c < nch; ++c)
   {
      const size_t off = c * CH, len = std::min(CH, bytes - off);
      char *pin = reinterpret_cast<char *>(h->pin[c & 1]);
      LF_HIP(hipEventSynchronize(h->pin_ev[c & 1]));      // its previous DMA is done
      const char *s = reinterpret_cast<const char *>(src) + off;
      std::vector<std::thread> pool;
      for (unsigned t = 0; t < nthr; ++t)
      {
         pool.emplace_back([=]() {
            const size_t a = len * t / nthr, b = len * (t + 1) / nthr;
            std::memcpy(pin + a, s + a, b - a);
         });
      }
      for (auto &th : pool) { th.join(); }
      LF_HIP(hipMemcpyAsync(reinterpret_cast<char *>(dst) + off, pin, len, hipMemcpyHostToDevice, st));
      LF_HIP(hipEventRecord(h->pin_ev[c & 1], st));
   }
   return LFASM_OK;
}

// Unit-eps mass matrices (face-transformed) of the parameter elements:
// the reference-tensor kernel when every one of them has constant
// coefficients, else the quadrature kernels.
int variants_mhat(lfasm_handle *h)
{
   const int nparam = h->nparam;
   LocalParams P = local_params(h, 0, LFASM_MASS, 1, h->Mhat);
   P.lda = h->nd;                     // Mhat rows are unpadded
   P.nout = nparam;
   P.elist = h->pelem;
   P.eps_unit = 1;
   bool all_const = h->Rten != nullptr;
   for (int i = 0; all_const && i < nparam; ++i)
   {
      all_const = h->eclass_host[(size_t)h->pelem_host[(size_t)i]] != 0;
   }
   if (all_const)
   {
      TensorParams T = tensor_params(h, P);
      T.nlist = nparam;
      T.elist = h->pelem;
      T.out_compact = 1;
      const int epar = T.nb > 128 ? 1 : 256 / T.nb;
      const dim3 grid((unsigned)((nparam + TEN_EPC - 1) / TEN_EPC),
                      epar > 1 ? 1u : (unsigned)((T.nb + 255) / 256));
      hipLaunchKernelGGL(k_local_tensor, grid, dim3(256), 0, h->stream, T, epar);
      LF_HIP(hipGetLastError());
   }
   else
   {
      LF_HIP(launch_local(h, P, h->nd >= 45 ? 2 : 1, h->stream));
   }
   LF_HIP(hipStreamSynchronize(h->stream));
   return LFASM_OK;
}

}  // namespace

extern "C" {

int lfasm_variants_prepare(lfasm_handle *h, int nparam, const int *elems)
{
   if (!h || !h->global_built) { return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE; }
   if (nparam <= 0 || !elems) { return h->fail(LFASM_E_ARG, "no parameter elements"); }
   for (int i = 0; i < nparam; ++i)
   {
      if (elems[i] < 0 || elems[i] >= h->ne || (i > 0 && elems[i] <= elems[i - 1]))
      {
         return h->fail(LFASM_E_ARG, "parameter elements must be sorted, unique, in [0, ne)");
      }
   }
   LF_HIP(hipSetDevice(h->device));
   const size_t nd = (size_t)h->nd;
   // the same parameter set again: its contributor lists and buffers stand,
   // only the unit-eps matrices are formed anew
   const bool same = h->nparam == nparam && h->Mhat &&
                     std::equal(elems, elems + nparam, h->pelem_host.begin());
   if (same) { return variants_mhat(h); }
   dfree(h->pelem); dfree(h->prc_ptr); dfree(h->prc_list); dfree(h->Mhat); dfree(h->vfix);
   dfree(h->vcopied);
   h->nparam = 0;
   h->nvfix = 0;
   h->pelem_host.assign(elems, elems + nparam);
   LF_HIP(upload(h, h->pelem, elems, (size_t)nparam));
   // per-row parameter contributors, ascending element order
   const std::vector<int> &dof = h->dof_host;
   std::vector<int> ptr((size_t)h->height + 1, 0), lst((size_t)nparam * nd);
   for (int i = 0; i < nparam; ++i)
      for (size_t a = 0; a < nd; ++a) { ptr[(size_t)dof[(size_t)elems[i] * nd + a] + 1]++; }
   for (int r = 0; r < h->height; ++r) { ptr[(size_t)r + 1] += ptr[(size_t)r]; }
   {
      std::vector<int> pos(ptr.begin(), ptr.end() - 1);
      for (int i = 0; i < nparam; ++i)
         for (size_t a = 0; a < nd; ++a)
         {
            lst[(size_t)pos[(size_t)dof[(size_t)elems[i] * nd + a]]++] = (int)(i * nd + a);
         }
   }
   h->pmaxc = 0;
   for (int r = 0; r < h->height; ++r) { h->pmaxc = std::max(h->pmaxc, ptr[(size_t)r + 1] - ptr[(size_t)r]); }
   LF_HIP(upload(h, h->prc_ptr, ptr.data(), ptr.size()));
   LF_HIP(upload(h, h->prc_list, lst.data(), lst.size()));
   {
      // runs of at least VCOPY_MIN values in consecutive rows without a
      // parameter contributor go to k_variant_copy, in chunks (64 variants on
      // 271k tets, p = 2: 20.05 ms against 20.8 ms with every row in the row
      // kernel; 15.5 against 21.3 ms with 16 parameter elements; thresholds
      // of 2k and 8k values were slower on the full parameter set)
      constexpr int VCOPY_MIN = 32768, CH = 4096;
      std::vector<int2> fx;
      std::vector<unsigned char> cp((size_t)h->height, 0);
      const int n = h->height;
      for (int r = 0; r < n;)
      {
         if (ptr[(size_t)r + 1] != ptr[(size_t)r]) { ++r; continue; }
         int r1 = r;
         while (r1 < n && ptr[(size_t)r1 + 1] == ptr[(size_t)r1]) { ++r1; }
         const int i0 = h->I_host[(size_t)r], i1 = h->I_host[(size_t)r1];
         if (i1 - i0 >= VCOPY_MIN)
         {
            for (int a = i0; a < i1; a += CH) { fx.push_back(make_int2(a, std::min(a + CH, i1))); }
            std::fill(cp.begin() + r, cp.begin() + r1, (unsigned char)1);
         }
         r = r1;
      }
      h->nvfix = (int)fx.size();
      if (!fx.empty()) { LF_HIP(upload(h, h->vfix, fx.data(), fx.size())); }
      LF_HIP(upload(h, h->vcopied, cp.data(), cp.size()));
   }
   LF_HIP(hipMalloc((void **)&h->Mhat, (size_t)nparam * nd * nd * sizeof(double2)));
   h->nparam = nparam;
   return variants_mhat(h);
}

int lfasm_variants_assemble(lfasm_handle *h, int nvar, const double *deps, int eliminate)
{
   if (!h || !h->nparam) { return h ? h->fail(LFASM_E_STATE, "variants_prepare first") : LFASM_E_STATE; }
   if (nvar <= 0 || !deps) { return h->fail(LFASM_E_ARG, "no variants"); }
   LF_HIP(hipSetDevice(h->device));
   {
      // the variant row kernels keep 2 rows per wavefront, 4 wavefronts, in LDS
      int lds_max = 0;
      LF_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
      if ((size_t)8 * h->maxrow * sizeof(double2) > (size_t)lds_max)
      {
         return h->fail(LFASM_E_UNSUPPORTED, "row length %d > %d (LDS row buffers of the variant kernel)",
                        h->maxrow, lds_max / (int)(8 * sizeof(double2)));
      }
   }
   if (nvar != h->nvar)
   {
      dfree(h->var_vals);
      LF_HIP(hipMalloc((void **)&h->var_vals, (size_t)nvar * h->nnz * sizeof(double2)));
      h->nvar = nvar;
   }
   const size_t ndeps = (size_t)nvar * h->nparam;
   if (ndeps > h->d_deps_cap)
   {
      dfree(h->d_deps);
      dfree(h->d_depsT);
      h->d_deps_cap = 0;
      LF_HIP(hipMalloc((void **)&h->d_deps, ndeps * sizeof(double2)));
      LF_HIP(hipMalloc((void **)&h->d_depsT, ndeps * sizeof(double2)));
      h->d_deps_cap = ndeps;
   }
   const double2 *d_deps = h->d_deps;
   VariantParams V;
   V.n = h->height;
   V.nd = h->nd;
   V.nvar = nvar;
   V.nparam = h->nparam;
   V.nnz = h->nnz;
   V.Ip = h->Ip;
   V.eflag = h->eflag;
   V.prc_ptr = h->prc_ptr;
   V.prc_list = h->prc_list;
   V.pelem = h->pelem;
   V.slot = h->slot;
   V.sld = h->sld;
   V.is_ess = h->ess;
   V.Mhat = h->Mhat;
   V.deps = d_deps;
   V.depsT = h->d_depsT;
   V.base = h->vals;
   V.out = h->var_vals;
   V.eliminate = eliminate ? 1 : 0;
   V.maxrow = h->maxrow;
   V.vcopied = h->vcopied;
   // kernel form: the accumulating one for rows of at most 256 entries, the
   // register one for longer rows, the general one for nd > 64 or more than
   // VMAXC contributors in a row
   const bool reg_ok = h->nd <= 64 && h->pmaxc <= VMAXC;
   const bool acc_ok = reg_ok && h->maxrow <= 256;
   int vk = h->variant_kernel;
   if (vk == 0) { vk = acc_ok ? 1 : reg_ok ? 2 : 3; }
   if ((vk == 1 && !acc_ok) || (vk == 2 && !reg_ok))
   {
      return h->fail(LFASM_E_UNSUPPORTED, "variant kernel %d: nd %d, %d contributors, row length %d", vk,
                     h->nd, h->pmaxc, h->maxrow);
   }
   const size_t shmem = vk == 1 ? (size_t)4 * (h->maxrow + 64) * sizeof(double2)
                                : (size_t)4 * 2 * h->maxrow * sizeof(double2);
   const int ns = (h->maxrow + 63) / 64;
   // the accumulating form holds at most 8 contributors per row in fewer
   // registers when the mesh allows (more wavefronts per SIMD)
   const bool nc8 = h->pmaxc <= 8;
   const void *kfun = vk == 3 ? (const void *)k_variant_rows
                      : vk == 2 ? (const void *)k_variant_rows_reg
                      : nc8 ? (ns <= 1 ? (const void *)k_variant_rows_acc<1, 8>
                               : ns == 2 ? (const void *)k_variant_rows_acc<2, 8>
                               : ns == 3 ? (const void *)k_variant_rows_acc<3, 8>
                                         : (const void *)k_variant_rows_acc<4, 8>)
                      : ns <= 1 ? (const void *)k_variant_rows_acc<1>
                      : ns == 2 ? (const void *)k_variant_rows_acc<2>
                      : ns == 3 ? (const void *)k_variant_rows_acc<3>
                                : (const void *)k_variant_rows_acc<4>;
   hipError_t e = hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
   if (e == hipSuccess) { e = ev_begin(h, T_BATCH); }
   // the copy of parameter-free runs needs no eps offsets: it runs while
   // the offsets go up on the second stream
   if (e == hipSuccess && h->nvfix > 0)
   {
      hipLaunchKernelGGL(k_variant_copy, dim3((unsigned)h->nvfix), dim3(256), 0, h->stream,
                         (const int2 *)h->vfix, (const double2 *)h->vals, h->var_vals, h->nnz, nvar);
      e = hipGetLastError();
   }
   LF_HIP(e);
   {
      const int rc = upload_pinned(h, h->d_deps, deps, ndeps * sizeof(double2), h->stream2);
      if (rc != LFASM_OK) { return rc; }
   }
   LF_HIP(hipEventRecord(h->ev_join, h->stream2));
   LF_HIP(hipStreamWaitEvent(h->stream, h->ev_join, 0));
   if (e == hipSuccess && vk != 3)
   {
      const dim3 tg((unsigned)((h->nparam + 31) / 32), (unsigned)((nvar + 31) / 32));
      hipLaunchKernelGGL(k_transpose_deps, tg, dim3(256), 0, h->stream, d_deps, h->d_depsT, nvar,
                         h->nparam);
      e = hipGetLastError();
   }
   if (e == hipSuccess)
   {
      const dim3 grid((unsigned)((h->height + 3) / 4));
      if (vk == 3) { hipLaunchKernelGGL(k_variant_rows, grid, dim3(BLOCK), shmem, h->stream, V); }
      else if (vk == 2) { hipLaunchKernelGGL(k_variant_rows_reg, grid, dim3(BLOCK), shmem, h->stream, V); }
      else
      {
         void *vargs[] = {&V};
         e = hipLaunchKernel(kfun, grid, dim3(BLOCK), vargs, shmem, h->stream);
      }
      if (e == hipSuccess) { e = hipGetLastError(); }
   }
   if (e == hipSuccess) { e = ev_end(h, T_BATCH); }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   LF_HIP(e);
   return LFASM_OK;
}

// k_par of the handle's (single) mesh, on the device table and the host copy
hipError_t set_kpar(lfasm_handle *h, double kx, double ky)
{
   h->pml_host[0].kx = kx;
   h->pml_host[0].ky = ky;
   h->any_kpar = kx != 0.0 || ky != 0.0;
   LF_TRY(hipMemcpyAsync(h->pml_tab, h->pml_host.data(), sizeof(PmlDev), hipMemcpyHostToDevice,
                         h->stream));
   return hipStreamSynchronize(h->stream);
}

int lfasm_angles_assemble(lfasm_handle *h, int nang, const double *kxy, int mask_cc, int mask_mass,
                          int apply_te, int eliminate)
{
   if (!h || !h->inited || !h->global_built)
   {
      return h ? h->fail(LFASM_E_STATE, "init and build_global first") : LFASM_E_STATE;
   }
   if (h->nmesh != 1) { return h->fail(LFASM_E_UNSUPPORTED, "angle sweeps take a single-mesh handle"); }
   if (nang <= 0 || !kxy) { return h->fail(LFASM_E_ARG, "no angles"); }
   LF_HIP(hipSetDevice(h->device));
   drop_graph(h);
   bool any_ky = false;
   for (int j = 0; j < nang; ++j) { any_ky = any_ky || kxy[2 * j + 1] != 0.0; }
   // basis matrices: (mask_cc, mask_mass, kx, ky) of each
   struct Basis { int mcc, mm; double kx, ky; };
   const int cross = mask_mass & (LFASM_CROSS0 | LFASM_CROSS1), klk = mask_mass & LFASM_KLK;
   std::vector<Basis> bs;
   bs.push_back({mask_cc, mask_mass & LFASM_MASS, 0.0, 0.0});
   const int ix = cross ? (int)bs.size() : -1;
   if (cross) { bs.push_back({0, cross, 1.0, 0.0}); }
   const int ixx = klk ? (int)bs.size() : -1;
   if (klk) { bs.push_back({0, klk, 1.0, 0.0}); }
   const int iy = (cross && any_ky) ? (int)bs.size() : -1;
   if (iy >= 0) { bs.push_back({0, cross, 0.0, 1.0}); }
   const int iyy = (klk && any_ky) ? (int)bs.size() : -1;
   if (iyy >= 0) { bs.push_back({0, klk, 0.0, 1.0}); }
   const int i11 = (klk && any_ky) ? (int)bs.size() : -1;
   if (i11 >= 0) { bs.push_back({0, klk, 1.0, 1.0}); }
   const int nb = (int)bs.size();
   // combination coefficients per angle (basis order)
   std::vector<double> c((size_t)nang * nb, 0.0);
   for (int j = 0; j < nang; ++j)
   {
      const double kx = kxy[2 * j], ky = kxy[2 * j + 1];
      double *cj = c.data() + (size_t)j * nb;
      cj[0] = 1.0;
      if (ix >= 0) { cj[ix] = kx; }
      if (iy >= 0) { cj[iy] = ky; }
      if (ixx >= 0) { cj[ixx] = i11 >= 0 ? kx * kx - kx * ky : kx * kx; }
      if (iyy >= 0) { cj[iyy] = ky * ky - kx * ky; }
      if (i11 >= 0) { cj[i11] = kx * ky; }
   }
   const size_t nnz = (size_t)h->nnz;
   if (nang != h->nang)
   {
      dfree(h->ang_vals);
      h->nang = 0;
      LF_HIP(hipMalloc((void **)&h->ang_vals, (size_t)nang * nnz * sizeof(double2)));
      h->nang = nang;
   }
   // one pass for all basis terms (their element matrices in separate
   // sets, scatter and combination fused) when the row buffers of all terms
   // fit in LDS; else each term's CSR values, then the combination
   const size_t ae_elems = (size_t)h->ne * h->nd * h->lda;
   const size_t shmem_multi = (size_t)4 * nb * h->maxrow * sizeof(double2);
   const bool multi = h->rslot && h->emask && h->nd <= 64 && h->tensor_mode != 2 &&
                      shmem_multi <= h->lds_max;
   if (multi && nb - 1 > h->nkae)
   {
      dfree(h->kAe);
      h->nkae = 0;
      LF_HIP(hipMalloc((void **)&h->kAe, (size_t)(nb - 1) * ae_elems * sizeof(double2)));
      h->nkae = nb - 1;
   }
   if (!multi && nb > h->nkb)
   {
      dfree(h->kbasis);
      h->nkb = 0;
      LF_HIP(hipMalloc((void **)&h->kbasis, (size_t)nb * nnz * sizeof(double2)));
      h->nkb = nb;
   }
   if ((int)c.size() > h->kcoef_cap)
   {
      dfree(h->d_kcoef);
      h->kcoef_cap = 0;
      LF_HIP(hipMalloc((void **)&h->d_kcoef, c.size() * sizeof(double)));
      h->kcoef_cap = (int)c.size();
   }
   LF_HIP(hipMemcpyAsync(h->d_kcoef, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice,
                         h->stream));
   const double kx0 = h->pml_host[0].kx, ky0 = h->pml_host[0].ky;
   hipError_t e = ev_begin(h, T_BATCH);
   double2 *vals = h->vals;
   for (int b = 0; b < nb && e == hipSuccess; ++b)
   {
      e = set_kpar(h, bs[b].kx, bs[b].ky);
      if (e != hipSuccess) { break; }
      double2 *ae = (multi && b > 0) ? h->kAe + (size_t)(b - 1) * ae_elems : h->Ae;
      const LocalParams P = local_params(h, bs[b].mcc, bs[b].mm, apply_te, ae);
      e = enqueue_local(h, P, 0, false);
      if (e != hipSuccess || multi) { continue; }
      // this term's CSR values into its own buffer (the handle's value
      // array swapped for it)
      h->vals = h->kbasis + (size_t)b * nnz;
      e = enqueue_scatter(h, eliminate, false);
      h->vals = vals;
   }
   if (e == hipSuccess && multi)
   {
      MultiScatter M;
      ScatterParams &S = M.S;
      S.n = h->height;
      S.nd = h->nd;
      S.Ip = h->Ip;
      S.Jc = h->Jc;
      S.eflag = h->eflag;
      S.emask = h->emask;
      S.rc_ptr = h->rc_ptr;
      S.rc_list = h->rc_list;
      S.is_ess = h->ess;
      S.Ae = h->Ae;
      S.lda = h->lda;
      S.vals = nullptr;
      S.eliminate = eliminate ? 1 : 0;
      S.maxrow = h->maxrow;
      S.slot = h->slot;
      S.sld = h->sld;
      S.rsoff = h->rsoff;
      S.rslot = h->rslot;
      S.rows = nullptr;
      S.nrows = h->height;
      M.nang = nang;
      for (int q = 0; q < 6; ++q)
      {
         const int qq = std::min(q, nb - 1);
         M.Ae[q] = qq == 0 ? h->Ae : h->kAe + (size_t)(qq - 1) * ae_elems;
      }
      M.coef = h->d_kcoef;
      M.out = h->ang_vals;
      M.nnz = (int64_t)nnz;
      const void *kf = nb == 1 ? (const void *)k_scatter_multi<1, 4>
                       : nb == 2 ? (const void *)k_scatter_multi<2, 4>
                       : nb == 3 ? (const void *)k_scatter_multi<3, 4>
                       : nb == 4 ? (const void *)k_scatter_multi<4, 2>
                       : nb == 5 ? (const void *)k_scatter_multi<5, 2>
                                 : (const void *)k_scatter_multi<6, 2>;
      e = hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem_multi);
      if (e == hipSuccess)
      {
         void *margs[] = {&M};
         e = hipLaunchKernel(kf, dim3((unsigned)((h->height + 3) / 4)), dim3(BLOCK), margs, shmem_multi,
                             h->stream);
      }
      if (e == hipSuccess) { e = hipGetLastError(); }
   }
   if (e == hipSuccess && !multi)
   {
      KparCombine K;
      K.nnz = (int64_t)nnz;
      K.nb = nb;
      K.nang = nang;
      for (int q = 0; q < 6; ++q) { K.B[q] = h->kbasis + (size_t)std::min(q, nb - 1) * nnz; }
      K.coef = h->d_kcoef;
      K.out = h->ang_vals;
      const unsigned grid = (unsigned)std::min<size_t>((nnz + 255) / 256, (size_t)h->cus * 16);
      hipLaunchKernelGGL(k_kpar_combine, dim3(std::max(1u, grid)), dim3(256), 0, h->stream, K);
      e = hipGetLastError();
   }
   if (e == hipSuccess && eliminate && !multi)
   {
      hipLaunchKernelGGL(k_kpar_ess, dim3((unsigned)std::min(8192, (h->height + 255) / 256)), dim3(256),
                         0, h->stream, h->height, (const int *)h->Ip, (const unsigned char *)h->ess,
                         (const double2 *)h->kbasis, h->ang_vals, (int64_t)nnz, nang);
      e = hipGetLastError();
   }
   if (e == hipSuccess) { e = ev_end(h, T_BATCH); }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   // the handle's own k_par back; its element matrices now hold the last
   // basis term, so the next scatter needs a compute_local first
   const hipError_t e2 = set_kpar(h, kx0, ky0);
   h->local_valid = false;
   LF_HIP(e);
   LF_HIP(e2);
   return LFASM_OK;
}

int lfasm_angle_values(lfasm_handle *h, int j, void **dev_vals)
{
   if (!h || j < 0 || j >= h->nang || !dev_vals) { return h ? h->fail(LFASM_E_ARG, "angle index") : LFASM_E_STATE; }
   *dev_vals = (void *)(h->ang_vals + (size_t)j * h->nnz);
   return LFASM_OK;
}

int lfasm_download_angle(lfasm_handle *h, int j, double *re, double *im)
{
   if (!h || j < 0 || j >= h->nang) { return h ? h->fail(LFASM_E_ARG, "angle index") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   return download_split(h, h->ang_vals + (size_t)j * h->nnz, (size_t)h->nnz, re, im);
}

int lfasm_variant_values(lfasm_handle *h, int v, void **dev_vals)
{
   if (!h || v < 0 || v >= h->nvar || !dev_vals) { return h ? h->fail(LFASM_E_ARG, "variant index") : LFASM_E_STATE; }
   *dev_vals = (void *)(h->var_vals + (size_t)v * h->nnz);
   return LFASM_OK;
}

int lfasm_download_variant(lfasm_handle *h, int v, double *re, double *im)
{
   if (!h || v < 0 || v >= h->nvar) { return h ? h->fail(LFASM_E_ARG, "variant index") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   std::vector<double2> t((size_t)h->nnz);
   LF_HIP(hipMemcpyAsync(t.data(), h->var_vals + (size_t)v * h->nnz, t.size() * sizeof(double2),
                         hipMemcpyDeviceToHost, h->stream));
   LF_HIP(hipStreamSynchronize(h->stream));
   for (size_t k = 0; k < t.size(); ++k)
   {
      if (re) { re[k] = t[k].x; }
      if (im) { im[k] = t[k].y; }
   }
   return LFASM_OK;
}

// ---- incident fields, RHS, sensitivity ---------------------------------------

int lfasm_set_incident(lfasm_handle *h, const lfasm_incident *inc)
{
   if (!h || !h->inited) { return h ? h->fail(LFASM_E_STATE, "init first") : LFASM_E_STATE; }
   if (!inc || inc->nrhs <= 0 || inc->nslab <= 0 || !inc->waves || !inc->region_slab ||
       !inc->region_deps || !inc->region_src)
   {
      return h->fail(LFASM_E_ARG, "incomplete incident description");
   }
   for (int r = 0; r < h->nreg; ++r)
   {
      if (inc->region_slab[r] < 0 || inc->region_slab[r] >= inc->nslab)
      {
         return h->fail(LFASM_E_ARG, "region %d slab %d outside [0, %d)", r, inc->region_slab[r],
                        inc->nslab);
      }
   }
   LF_HIP(hipSetDevice(h->device));
   const size_t nw = (size_t)h->nmesh * inc->nrhs * inc->nslab;
   std::vector<SlabWaveDev> w(nw);
   for (size_t i = 0; i < nw; ++i)
   {
      const double *x = inc->waves + 18 * i;
      w[i].zA = x[0];
      w[i].zB = x[1];
      w[i].qA = make_double2(x[2], x[3]);
      w[i].qB = make_double2(x[4], x[5]);
      for (int c = 0; c < 3; ++c)
      {
         w[i].A[c] = make_double2(x[6 + 2 * c], x[7 + 2 * c]);
         w[i].B[c] = make_double2(x[12 + 2 * c], x[13 + 2 * c]);
      }
   }
   LF_HIP(upload(h, h->waves, w.data(), w.size()));
   LF_HIP(upload(h, h->region_slab, inc->region_slab, (size_t)h->nreg));
   LF_HIP(upload(h, h->region_src, inc->region_src, (size_t)h->nreg));
   LF_HIP(upload(h, h->region_deps, (const double2 *)inc->region_deps, (size_t)h->nreg));
   h->inc_nrhs = inc->nrhs;
   h->inc_nslab = inc->nslab;
   h->region_src_host.assign(inc->region_src, inc->region_src + h->nreg);
   h->region_deps_host.assign((const double2 *)inc->region_deps,
                              (const double2 *)inc->region_deps + h->nreg);
   dfree(h->rhs_src); dfree(h->rhs_index); dfree(h->rhs_bloc);
   h->rhs_nsrc = -1;
   return LFASM_OK;
}

int lfasm_assemble_rhs(lfasm_handle *h, int eliminate)
{
   if (!h || !h->global_built || !h->inc_nrhs)
   {
      return h ? h->fail(LFASM_E_STATE, "build_global and set_incident first") : LFASM_E_STATE;
   }
   LF_HIP(hipSetDevice(h->device));
   const int nrhs = h->inc_nrhs, nd = h->nd;
   if (h->rhs_nsrc < 0)
   {
      // source elements: those with a nonzero source strength eps - eps_bg
      // (k_rhs_local): elements of source regions, and with per-element
      // eps every element whose eps differs from its slab background
      std::vector<int> attr((size_t)h->ne), pid((size_t)h->ne, 0), eps_off((size_t)h->nmesh);
      LF_HIP(hipMemcpy(attr.data(), h->attr, attr.size() * sizeof(int), hipMemcpyDeviceToHost));
      if (h->pid) { LF_HIP(hipMemcpy(pid.data(), h->pid, pid.size() * sizeof(int), hipMemcpyDeviceToHost)); }
      LF_HIP(hipMemcpy(eps_off.data(), h->eps_off, eps_off.size() * sizeof(int), hipMemcpyDeviceToHost));
      std::vector<int> src, index((size_t)h->ne, -1);
      for (int e = 0; e < h->ne; ++e)
      {
         const int reg = eps_off[(size_t)pid[(size_t)e]] + attr[(size_t)e] - 1;
         double2 de = h->region_src_host[(size_t)reg] ? h->region_deps_host[(size_t)reg]
                                                      : make_double2(0.0, 0.0);
         if (!h->eps_elem_host.empty())
         {
            const double2 ee = h->eps_elem_host[(size_t)e], er = h->eps_tab_host[(size_t)reg];
            de = make_double2(ee.x - (er.x - de.x), ee.y - (er.y - de.y));
         }
         if (de.x != 0.0 || de.y != 0.0) { index[(size_t)e] = (int)src.size(); src.push_back(e); }
      }
      LF_HIP(upload(h, h->rhs_index, index.data(), index.size()));
      LF_HIP(upload(h, h->rhs_src, src.data(), src.size()));
      dfree(h->rhs_bloc);
      if (!src.empty())
      {
         LF_HIP(hipMalloc((void **)&h->rhs_bloc, src.size() * nrhs * nd * sizeof(double2)));
      }
      h->rhs_nsrc = (int)src.size();
   }
   const size_t nrhs_entries = (size_t)nrhs * h->height;
   if (h->rhs_cap != nrhs_entries)
   {
      dfree(h->rhs);
      h->rhs_cap = 0;
      LF_HIP(hipMalloc((void **)&h->rhs, nrhs_entries * sizeof(double2)));
      h->rhs_cap = nrhs_entries;
   }
   const int nsrc = h->rhs_nsrc;
   const int *d_src = h->rhs_src, *d_index = h->rhs_index;
   double2 *bloc = h->rhs_bloc;
   hipError_t e = hipSuccess;
   if (e == hipSuccess) { e = ev_begin(h, T_BATCH); }
   if (e == hipSuccess && nsrc > 0)
   {
      RhsParams R;
      R.nd = nd;
      R.p = h->p;
      R.nrhs = nrhs;
      R.nslab = h->inc_nslab;
      R.nsrc = nsrc;
      R.src_elem = d_src;
      R.attr = h->attr;
      R.pid = h->pid;
      R.eps_off = h->eps_off;
      R.J = h->J;
      R.Ji = h->Ji;
      R.detJ = h->detJ;
      R.v0 = h->v0;
      R.fo = h->fo;
      R.rule = RuleDev{h->rm.nq, h->rm.ip, h->rm.w, h->rm.vs, h->rm.cs};
      R.pml_tab = h->pml_tab;
      R.region_slab = h->region_slab;
      R.region_deps = h->region_deps;
      R.region_src = h->region_src;
      R.eps_tab = h->eps_tab;
      R.eps_elem = h->eps_elem;
      R.waves = h->waves;
      std::memcpy(R.TInv, h->TInv, sizeof(R.TInv));
      R.bloc = bloc;
      const size_t shmem = (size_t)h->rm.nq * nrhs * 3 * sizeof(double2);
      e = hipFuncSetAttribute((const void *)k_rhs_local, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)shmem);
      if (e == hipSuccess)
      {
         hipLaunchKernelGGL(k_rhs_local, dim3((unsigned)nsrc), dim3(64), shmem, h->stream, R);
         e = hipGetLastError();
      }
   }
   if (e == hipSuccess)
   {
      if (nsrc == 0)
      {
         hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, h->stream, h->rhs,
                            (size_t)nrhs * h->height);
      }
      else
      {
         hipLaunchKernelGGL(k_rhs_rows, dim3((h->height + 255) / 256), dim3(256), 0, h->stream,
                            h->height, nd, nrhs, h->rc_ptr, h->rc_list, d_index, h->eflag, h->ess,
                            eliminate ? 1 : 0, bloc, h->rhs);
      }
      e = hipGetLastError();
   }
   if (e == hipSuccess) { e = ev_end(h, T_BATCH); }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   LF_HIP(e);
   return LFASM_OK;
}

int lfasm_download_rhs(lfasm_handle *h, double *re, double *im)
{
   if (!h || !h->rhs) { return h ? h->fail(LFASM_E_STATE, "assemble_rhs first") : LFASM_E_STATE; }
   LF_HIP(hipSetDevice(h->device));
   return download_split(h, h->rhs, (size_t)h->inc_nrhs * h->height, re, im);
}

int lfasm_eps_sensitivity(lfasm_handle *h, int nparam, const int *elems, const double *lam,
                          const double *u, int rhs_col, double *s_out)
{
   if (!h || !h->global_built) { return h ? h->fail(LFASM_E_STATE, "build_global first") : LFASM_E_STATE; }
   if (nparam <= 0 || !elems || !lam || !u || !s_out) { return h->fail(LFASM_E_ARG, "null argument"); }
   if (rhs_col >= 0 && rhs_col >= h->inc_nrhs) { return h->fail(LFASM_E_ARG, "rhs_col without incident field"); }
   for (int i = 0; i < nparam; ++i)
   {
      if (elems[i] < 0 || elems[i] >= h->ne) { return h->fail(LFASM_E_ARG, "element %d out of range", elems[i]); }
   }
   LF_HIP(hipSetDevice(h->device));
   int *d_el = nullptr;
   double2 *d_lam = nullptr, *d_u = nullptr, *d_s = nullptr;
   hipError_t e = upload(h, d_el, elems, (size_t)nparam);
   if (e == hipSuccess) { e = upload(h, d_lam, (const double2 *)lam, (size_t)h->height); }
   if (e == hipSuccess) { e = upload(h, d_u, (const double2 *)u, (size_t)h->height); }
   if (e == hipSuccess) { e = hipMalloc((void **)&d_s, (size_t)nparam * sizeof(double2)); }
   if (e == hipSuccess)
   {
      SensParams S;
      S.nd = h->nd;
      S.p = h->p;
      S.nparam = nparam;
      S.nrhs = h->inc_nrhs;
      S.nslab = h->inc_nslab;
      S.rhs_col = rhs_col;
      S.pelem = d_el;
      S.dof = h->dof;
      S.attr = h->attr;
      S.pid = h->pid;
      S.eps_off = h->eps_off;
      S.fo = h->fo;
      S.eflag = h->eflag;
      S.J = h->J;
      S.Ji = h->Ji;
      S.detJ = h->detJ;
      S.v0 = h->v0;
      S.rm = RuleDev{h->rm.nq, h->rm.ip, h->rm.w, h->rm.vs, h->rm.cs};
      S.pml_tab = h->pml_tab;
      S.region_slab = h->region_slab;
      S.region_src = h->region_src;
      S.waves = h->waves;
      std::memcpy(S.TInv, h->TInv, sizeof(S.TInv));
      S.lam = d_lam;
      S.u = d_u;
      S.s = d_s;
      e = ev_begin(h, T_ADJ);
      if (e == hipSuccess)
      {
         hipLaunchKernelGGL(k_eps_sens, dim3(nparam), dim3(64), 2 * h->nd * sizeof(double2),
                            h->stream, S);
         e = hipGetLastError();
      }
      if (e == hipSuccess) { e = ev_end(h, T_ADJ); }
   }
   if (e == hipSuccess)
   {
      e = hipMemcpyAsync(s_out, d_s, (size_t)nparam * sizeof(double2), hipMemcpyDeviceToHost, h->stream);
   }
   if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
   if (d_el) { (void)hipFree(d_el); }
   if (d_lam) { (void)hipFree(d_lam); }
   if (d_u) { (void)hipFree(d_u); }
   if (d_s) { (void)hipFree(d_s); }
   LF_HIP(e);
   return LFASM_OK;
}

void *lfasm_stream(lfasm_handle *h) { return h ? (void *)h->stream : nullptr; }

// build_csr_symbolic (solver/asm_data.inc:178-210) on host threads: rows are
// split into contiguous ranges; each row's column list is the sorted union
// of the dof lists of the elements that contain the row.
int64_t lfasm_csr_symbolic(int ne, int nd, int height, const int *dof, int *I, int **Jcol)
{
   if (ne <= 0 || nd <= 0 || height <= 0 || !dof || !I || !Jcol) { return LFASM_E_ARG; }
   const size_t nn = (size_t)ne * nd;
   for (size_t i = 0; i < nn; ++i)
   {
      if (dof[i] < 0 || dof[i] >= height) { return LFASM_E_ARG; }
   }
   std::vector<int> cptr((size_t)height + 1, 0), elist(nn);
   for (size_t i = 0; i < nn; ++i) { cptr[(size_t)dof[i] + 1]++; }
   for (int r = 0; r < height; ++r) { cptr[(size_t)r + 1] += cptr[r]; }
   {
      std::vector<int> pos(cptr.begin(), cptr.end() - 1);
      for (size_t i = 0; i < nn; ++i) { elist[(size_t)pos[dof[i]]++] = (int)(i / nd); }
   }
   unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
   std::vector<int> len((size_t)height, 0);
   auto row_cols = [&](int r, std::vector<int> &tmp) {
      tmp.clear();
      for (int c = cptr[r]; c < cptr[(size_t)r + 1]; ++c)
      {
         const int *ed = dof + (size_t)elist[c] * nd;
         tmp.insert(tmp.end(), ed, ed + nd);
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
   };
   auto run = [&](auto body) {
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nth; ++t)
      {
         th.emplace_back([&, t]() {
            std::vector<int> tmp;
            const int r0 = (int)((int64_t)height * t / nth), r1 = (int)((int64_t)height * (t + 1) / nth);
            for (int r = r0; r < r1; ++r) { body(r, tmp); }
         });
      }
      for (auto &x : th) { x.join(); }
   };
   run([&](int r, std::vector<int> &tmp) { row_cols(r, tmp); len[(size_t)r] = (int)tmp.size(); });
   int64_t nnz = 0;
   I[0] = 0;
   for (int r = 0; r < height; ++r)
   {
      nnz += len[(size_t)r];
      if (nnz > INT32_MAX) { return LFASM_E_UNSUPPORTED; }
      I[r + 1] = (int)nnz;
   }
   int *J = (int *)std::malloc(sizeof(int) * (size_t)std::max<int64_t>(nnz, 1));
   if (!J) { return LFASM_E_ARG; }
   run([&](int r, std::vector<int> &tmp) {
      row_cols(r, tmp);
      std::copy(tmp.begin(), tmp.end(), J + I[r]);
   });
   *Jcol = J;
   return nnz;
}

void lfasm_free(void *p) { std::free(p); }

}  // extern "C"
