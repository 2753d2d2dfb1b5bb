/*
 * lithofem_asm.h -- C-ABI of the MI355X (gfx950) Nedelec system-matrix
 * assembly.  Plain C: opaque handle, plain pointers and sizes, int status.
 *
 * Every entry point below replaces a member of the reference's CUDA
 * assembly classes AsmGpu / AsmGpuGlobal (solver/asm_gpu.inc:271-599), which
 * the reference's solve driver calls at solver/lithofem_solve.cpp:783-792
 * (Init, BuildGlobal, ComputeLocal, ScatterEliminate), :873-875 (device CSR
 * handoff) and :910 (Residual).  The driver-side shim fills lfasm_mesh from
 * the reference's AsmData (solver/asm_data.inc:77-97) and Problem
 * (solver/problem_common.inc:40-65) exactly as AsmGpu::Init reads them.
 *
 * Status codes: 0 = OK; negative = error class (LFASM_E_*).  After an error
 * the handle's lfasm_last_error() string says what failed; the caller is
 * expected to fall back to its CPU assembly, as the reference driver does
 * (solver/lithofem_solve.cpp:796-799).
 *
 * Threading: one handle per host thread; each handle owns one HIP stream on
 * its device and every device call of the handle is ordered on it.  Calls
 * that return host data synchronise that stream.
 */
#ifndef LITHOFEM_ASM_H
#define LITHOFEM_ASM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
   LFASM_OK = 0,
   LFASM_E_ARG = -1,      /* invalid argument or size */
   LFASM_E_HIP = -2,      /* HIP runtime error (message in last_error) */
   LFASM_E_FORCED = -3,   /* LITHOFEM_ASM_FORCE_FAIL set in the environment */
   LFASM_E_UNSUPPORTED = -4, /* element order or size outside the kernels */
   LFASM_E_STATE = -5,    /* call order (e.g. scatter before build_global) */
   LFASM_E_PATTERN = -6   /* an element (row, col) pair is missing from the CSR */
};

/* integrator selection bits, as the reference's mask enum
 * (solver/asm_gpu.inc:51-58) */
enum {
   LFASM_CURLCURL = 1,
   LFASM_MASS = 2,
   LFASM_CROSS0 = 4,
   LFASM_CROSS1 = 8,
   LFASM_KLK = 16
};

typedef struct lfasm_handle lfasm_handle;

/* One reference integration rule (solver/asm_data.inc:38-45). */
typedef struct {
   int nq;
   const double *ip;       /* nq*3 reference points */
   const double *w;        /* nq weights */
   const double *vshape;   /* nq*nd*3, [(q*nd + k)*3 + d] */
   const double *curl;     /* nq*nd*3 */
} lfasm_rule;

/* PML and Bloch parameters: the fields of the reference's AsmPmlParams
 * (solver/asm_gpu.inc:31-36), smax already per side. */
typedef struct {
   double z_min, z_max, thick, order;
   double smax_bot, smax_top;
   double k0, kx, ky;
} lfasm_pml;

/* Mesh + discretisation data (solver/asm_data.inc:77-97 plus the region eps
 * table and PML parameters AsmGpu::Init takes from prob,
 * solver/asm_gpu.inc:296-319).  All arrays are host pointers, copied during
 * lfasm_init. */
typedef struct {
   int ne, nd, p, height;
   const int *dof;          /* ne*nd decoded global dof (d<0 -> -1-d) */
   const double *sign;      /* ne*nd  +/-1 */
   const int *fo;           /* ne*4 face orientations 0..5 (p >= 2; else NULL) */
   const int *attr;         /* ne 1-based region attribute */
   const double *J;         /* ne*9 row-major Jacobian */
   const double *Jinv;      /* ne*9 row-major inverse Jacobian */
   const double *detJ;      /* ne */
   const double *v0;        /* ne*3 element origin */
   lfasm_rule rule_cc;      /* CurlCurl rule (order 2p-2) */
   lfasm_rule rule_mass;    /* mass / mixed rule (order 2p) */
   int nreg;
   const double *eps_tab;   /* nreg interleaved complex, by attribute-1 */
   const double *eps_elem;  /* optional ne interleaved complex (overrides eps_tab) */
   lfasm_pml pml;
   const double *TInv;      /* 24 doubles: solver/asm_data.inc:26-34 layout */
} lfasm_mesh;

/* ---- lifetime (AsmGpu constructor / Release, solver/asm_gpu.inc:368-376) */
int lfasm_create(int device, lfasm_handle **out);
void lfasm_destroy(lfasm_handle *h);
const char *lfasm_last_error(const lfasm_handle *h);

/* smax = -ln(target) (order+1) / (2 k0 n_med thick)
 * (solver/asm_gpu.inc:316-319); helper for the driver-side shim. */
double lfasm_pml_smax(double target, double order, double k0, double n_med, double thick);

/* AsmGpu::Init (solver/asm_gpu.inc:276-322): upload geometry, tables,
 * coefficients; allocate the local-matrix buffer.  Honours
 * LITHOFEM_ASM_FORCE_FAIL (returns LFASM_E_FORCED before any device call). */
int lfasm_init(lfasm_handle *h, const lfasm_mesh *m);

/* AsmGpu::Init with the element geometry extracted on the device
 * (solver/asm_data.inc:152-171 reads J, detJ, J^-1 and x(0) from MFEM's
 * element transformation on the host): elem_xyz holds every element's four
 * vertex coordinates (ne x 4 x 3, the element's local vertex order,
 * unwrapped across periodic boundaries), from which the device forms
 * J = [x1 - x0, x2 - x0, x3 - x0] (row-major), detJ (cofactor expansion of
 * the first row), J^-1 = adj(J) / detJ and v0 = x0.  m's J, Jinv, detJ, v0 are
 * not read (may be NULL); the other fields as lfasm_init.  An element with
 * detJ <= 0 is refused (LFASM_E_ARG). */
int lfasm_init_vertices(lfasm_handle *h, const lfasm_mesh *m, const double *elem_xyz);

/* Host copies of the element geometry the handle holds (ne x 9, ne x 9, ne,
 * ne x 3); any pointer may be NULL. */
int lfasm_download_geometry(lfasm_handle *h, double *J, double *Jinv, double *detJ, double *v0);

/* Batched multi-mesh init (no reference counterpart): the meshes (same
 * order, same rule tables and TInv; each with its own geometry, region
 * table, PML and k_par) are assembled as one block-diagonal system whose
 * elements and dofs are the meshes' in order.  build_global then takes the
 * union pattern (each mesh's CSR with rows/columns offset by the preceding
 * meshes' dof counts); every kernel launch covers all meshes at once. */
int lfasm_init_batch(lfasm_handle *h, const lfasm_mesh *meshes, int nmesh);

/* Element and dof offsets of the batched meshes (nmesh+1 entries each);
 * returns nmesh. */
int lfasm_batch_offsets(const lfasm_handle *h, int64_t *elem_off, int64_t *row_off);

/* Per-element complex eps (no reference counterpart: the reference reads
 * region eps only, solver/asm_gpu.inc:96).  eps = ne interleaved complex, or
 * NULL to return to the region table. */
int lfasm_set_eps_elem(lfasm_handle *h, const double *eps);

/* AsmGpuGlobal::BuildGlobal (solver/asm_gpu.inc:501-514): CSR pattern
 * (I: height+1, Jcol: nnz, sorted columns) and essential mask (height
 * bytes, 1 = essential).  Also builds the element-to-slot map and the
 * per-row contributor lists of the deterministic scatter. */
int lfasm_build_global(lfasm_handle *h, const int *I, const int *Jcol, int64_t nnz,
                       const unsigned char *is_ess);

/* build_csr_symbolic (solver/asm_data.inc:178-210) + AsmGpuGlobal::BuildGlobal
 * in one call, with the CSR pattern built on the device from the element dof
 * table: sorted-column rows of the union of the element dof x dof graphs,
 * the same pattern as the host routine (and as MFEM's after
 * SortColumnIndices).  Then as lfasm_build_global.  lfasm_sizes reports nnz;
 * lfasm_device_csr / lfasm_download_pattern return the pattern. */
int lfasm_build_global_device(lfasm_handle *h, const unsigned char *is_ess);

/* Host copy of the CSR pattern (I: height+1, Jcol: nnz); either may be NULL. */
int lfasm_download_pattern(lfasm_handle *h, int *I, int *Jcol);

/* Host copy of the per-row contributor lists of the scatter (built on the
 * device by both build_global calls): rc_ptr height+1, rc_list ne*nd entries
 * e*nd + a of the elements holding each row's dof, ascending; either may be
 * NULL. */
int lfasm_download_contributors(lfasm_handle *h, int *rc_ptr, int *rc_list);

/* AsmGpu::ComputeLocal (solver/asm_gpu.inc:326-351): element matrices of
 * the CurlCurlIntegrator (mask_cc & LFASM_CURLCURL) on the 2p-2 rule plus
 * the masked mass / mixed / KLK integrators (mask_mass) on the 2p rule,
 * then (apply_te, p >= 2) the dual face-DOF transform.  Ae is overwritten. */
int lfasm_compute_local(lfasm_handle *h, int mask_cc, int mask_mass, int apply_te);

/* lfasm_compute_local with an explicit kernel choice (0 = automatic:
 * FP64 MFMA tiles for nd > 16, VALU otherwise; 1 = VALU kernel;
 * 2 + v = MFMA launch variant v, v < LFASM_N_MFMA_VARIANTS; for
 * 32 < nd <= 48 variant 0 is the symmetric split Ae = S + P - P^T when
 * CROSS0 and CROSS1 are both on or both off).  No reference counterpart:
 * used by the parity tests to check every kernel and by the launch
 * autotuner. */
int lfasm_compute_local_kernel(lfasm_handle *h, int mask_cc, int mask_mass, int apply_te,
                               int kernel);

/* Number of MFMA launch variants (kernel selectors 2 .. 2+N-1). */
#define LFASM_N_MFMA_VARIANTS 8

/* Make `kernel` (a selector of lfasm_compute_local_kernel) the one that
 * selector 0 uses on this handle -- the launch autotuner's result.  0
 * restores the built-in default.  No reference counterpart. */
int lfasm_set_local_kernel(lfasm_handle *h, int kernel);

/* Element-matrix path selection (no reference counterpart).  Elements
 * whose four vertices lie in [z_min, z_max] of their problem -- Lambda = I
 * at every quadrature point, so every coefficient is constant over the
 * element -- can take the reference-tensor path: Ae = sum of 21
 * precomputed reference-element tensors (built from the same rule tables)
 * times element coefficients from J, eps and k_par.  The other elements
 * take the quadrature kernels that lfasm_compute_local_kernel's selector
 * names.  mode 0: quadrature kernels for every element; 1 (default):
 * reference-tensor kernel writing Ae; 2: the CSR row kernels of
 * lfasm_scatter_eliminate evaluate those elements' rows directly from the
 * tensors (their Ae is never written -- it saves that HBM, but the row
 * kernels then run one LDS-heavy workgroup per CU and are slower than mode
 * 1 at 1.3M DOF; lfasm_download_local computes Ae on demand), for p <= 3
 * with the face transform on, else as mode 1.  Modes 1 and 2 give
 * bit-identical values.  Masks
 * holding exactly one of LFASM_CROSS0 / LFASM_CROSS1 always use the
 * quadrature kernels.  Returns the number of elements on the tensor path
 * (0 in mode 0). */
int lfasm_set_tensor_path(lfasm_handle *h, int mode);

/* Run the two element-matrix kernels of the tensor path side by side (no
 * reference counterpart).  grid = 0 (default): the reference-tensor kernel,
 * then the quadrature kernel, on the handle's stream.  grid > 0: the
 * tensor kernel with at most `grid` workgroups along its element chunks
 * (each workgroup walks the chunks with that stride) on the handle's
 * stream, and the quadrature kernel on a second stream of the handle
 * (default priority), forked and joined by events, so that the HBM-bound
 * tensor kernel holds only part of each CU and the MFMA-bound quadrature
 * kernel the rest.  Values are bit-identical to grid = 0 (the same
 * kernels, disjoint outputs).  scatter_overlap = 1 (with grid > 0) changes
 * the captured step only (lfasm_step_capture): the CSR rows whose
 * contributors are all tensor-path elements are scattered on the tensor
 * branch right after the tensor kernel, while the quadrature kernel still
 * runs; the other rows are scattered after the join.  Every row is formed
 * by exactly one scatter, in the same contributor order, so the values are
 * again bit-identical.  The launch autotuner chooses both on the handle's
 * mesh.  Drops the captured step. */
int lfasm_set_element_concurrency(lfasm_handle *h, int grid, int scatter_overlap);

/* Residency cap of the MFMA quadrature kernels when they run beside the
 * tensor kernel and the row scatter (no reference counterpart): with
 * wg_per_cu > 0 (at most 8) and element kernels side by side
 * (lfasm_set_element_concurrency grid > 0), a quadrature launch asks for
 * enough LDS that at most wg_per_cu of its workgroups fit on one CU, so the
 * HBM-bound kernels beside it keep more wave slots; 0 = no cap.  Launches in
 * sequence are not affected, nor are values.  The launch autotuner chooses
 * it with the concurrency.  Drops the captured step. */
int lfasm_set_quad_residency(lfasm_handle *h, int wg_per_cu);

/* Captured assembly step (no reference counterpart; the reference launches
 * ComputeLocal and ScatterEliminate one kernel at a time,
 * solver/lithofem_solve.cpp:789-790).  capture records the whole step --
 * element matrices (both kernel paths) and scatter + elimination -- into a
 * HIP graph on the handle's stream; launch replays it (sync = 1 waits for
 * it).  Calls that change
 * buffers or launch choices (init, build_global, set_eps_elem,
 * set_local_kernel, set_tensor_path) drop the graph, and the next launch
 * captures it again with the same masks.  The step holds no memset nodes. */
int lfasm_step_capture(lfasm_handle *h, int mask_cc, int mask_mass, int apply_te, int eliminate);
int lfasm_step_launch(lfasm_handle *h, int sync);

/* AsmGpuGlobal::ScatterEliminate (solver/asm_gpu.inc:517-530): CSR values
 * from the local matrices by a deterministic per-row segmented reduction
 * (contributions summed in ascending element order), then (eliminate)
 * essential rows/columns -> 0 with unit diagonal. */
int lfasm_scatter_eliminate(lfasm_handle *h, int eliminate);

/* AsmGpuGlobal::DeviceI / DeviceJ / DeviceValues (solver/asm_gpu.inc:
 * 579-583): borrowed device pointers (values interleaved complex double). */
int lfasm_device_csr(lfasm_handle *h, int **I, int **Jcol, void **vals);

/* Production download of the CSR values for the host direct solver, split
 * re/im in (I, Jcol) order (the layout of MFEM's ComplexSparseMatrix
 * real()/imag() after SortColumnIndices). */
int lfasm_download_values(lfasm_handle *h, double *re, double *im);

/* AsmGpu::DownloadLocal (solver/asm_gpu.inc:355-362): test-only D2H of the
 * ne*nd*nd interleaved local matrices; prints the marker line
 * "asm_gpu_test_download (test-only D2H)". */
int lfasm_download_local(lfasm_handle *h, double *Ae);

/* AsmGpuGlobal::Residual (solver/asm_gpu.inc:533-565): ||A X - B|| / ||B||
 * with X, B as 2n [re; im] host vectors. */
int lfasm_residual(lfasm_handle *h, const double *X2n, const double *B2n, double *rel);

/* lfasm_residual without allocations (no reference counterpart; for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * lfasm_residual (||A X - B|| / ||B||, X and B 2n [re; im] host vectors of
 * len = 2n doubles; the norms are reduced on the device, so the last bits
 * may differ), with no hipMalloc / hipFree: X, B and the partial sums live
 * in the library's static device workspace (one per GPU, about 48 MB, held
 * until the process ends).  n is limited to 1,572,864 dofs (LFASM_E_ARG
 * beyond).  Every call on a device shares that device's workspace: calls
 * that overlap on one GPU (several host threads or handles at once) race. */
#define LFASM_RESIDUAL_STATIC_MAX_N 1572864
int lfasm_residual_static(lfasm_handle *h, int64_t len, const double *X2n, const double *B2n,
                          double *rel);

/* n(), nnz(), ne(), nd() of the reference classes. */
int lfasm_sizes(const lfasm_handle *h, int *ne, int *nd, int *n, int64_t *nnz);

/* Device-event timings (milliseconds) of the last calls, in this order:
 * [0] compute_local kernels, [1] scatter_eliminate kernels,
 * [2] build_global device part, [3] last batched call (variants_assemble or
 * assemble_rhs kernels), [4] last adjoint call,
 * [5] reference-tensor part and [6] quadrature part of the last
 * compute_local (0 when that part did not run), [7] last lfasm_step_launch,
 * [8] lfasm_build_global_device's pattern phase (both kernel passes and the
 * row-pointer scan between them), [9] lfasm_init_vertices' geometry kernel. */
int lfasm_timings(const lfasm_handle *h, double *ms, int n);

/* ---- eps variants (no reference counterpart) ------------------------------
 * Many per-element eps sets on one mesh by delta assembly: the mass term is
 * linear in eps, so A(eps_v) = A_base + sum_e deps_{v,e} Mhat_e with Mhat_e
 * the face-transformed unit-eps mass matrix of parameter element e.
 * prepare: parameter elements (sorted, unique element ids); computes Mhat
 * (called again with the same elements it keeps the per-row parameter
 * contributor lists and forms only Mhat anew).
 * assemble: A_base = the handle's current CSR values (compute_local +
 * scatter_eliminate with the same `eliminate`); deps = nvar x nparam
 * interleaved complex; writes nvar value arrays on the device. */
int lfasm_variants_prepare(lfasm_handle *h, int nparam, const int *elems);
int lfasm_variants_assemble(lfasm_handle *h, int nvar, const double *deps, int eliminate);
/* Variant kernel form (no reference counterpart): 0 = automatic (the
 * accumulating form when nd <= 64, at most 16 parameter contributors per
 * row and rows of at most 256 entries; else the register form when nd <= 64
 * and at most 16 contributors; else the general form), 1 = accumulating
 * (base row in registers, contributions added with LDS atomics),
 * 2 = register (contributor entries in registers, read-modify-write of the
 * row buffer), 3 = general.  Every form sums each entry as base + the
 * contributions in ascending element order.  assemble fails with
 * LFASM_E_UNSUPPORTED when a forced form does not fit the mesh. */
int lfasm_set_variant_kernel(lfasm_handle *h, int kernel);
/* ---- angle sweeps (no reference counterpart; the reference assembles one
 * matrix per incidence angle, solver/lithofem_solve.cpp:744-808) ----------
 * For one mesh and wavelength the k_par terms are polynomial in
 * k_par = (kx, ky): CROSS0 + CROSS1 (i K Lambda^-1 + i Lambda^-1 K) are linear
 * in K and KLK (-K Lambda^-1 K) quadratic, with K linear in (kx, ky).  So
 *   A(kx, ky) = B0 + kx Bx + ky By + kx^2 Bxx + ky^2 Byy + kx ky (B11 - Bxx - Byy)
 * where B0 holds the curl-curl and eps-mass terms of the masks, Bx / By the
 * cross terms at k_par = (1, 0) / (0, 1) and Bxx / Byy / B11 the KLK term at
 * (1, 0) / (0, 1) / (1, 1) (By, Byy, B11 only when some ky != 0).
 * angles_assemble forms those basis matrices (element matrices + scatter +
 * `eliminate`, on a single-mesh handle after build_global), then every
 * angle's CSR values in one pass over the nonzeros; kxy = nang x (kx, ky).
 * Eliminated rows are (0 ... 1 ... 0) in every angle.  The values equal a
 * plain assembly at that k_par up to rounding (the same terms, summed in a
 * different order).  The handle's own k_par is restored; its element
 * matrices are left holding a basis term, so call compute_local before the
 * next scatter_eliminate. */
int lfasm_angles_assemble(lfasm_handle *h, int nang, const double *kxy, int mask_cc, int mask_mass,
                          int apply_te, int eliminate);
int lfasm_angle_values(lfasm_handle *h, int j, void **dev_vals);
int lfasm_download_angle(lfasm_handle *h, int j, double *re, double *im);

int lfasm_variant_values(lfasm_handle *h, int v, void **dev_vals);
int lfasm_download_variant(lfasm_handle *h, int v, double *re, double *im);

/* ---- incident fields, batched RHS and eps sensitivity ----------------------
 * Incident envelope per slab (the reference's SlabWave,
 * solver/problem_common.inc:11-16,218-236): E(z) = A exp(i qA (z-zA)) +
 * B exp(i qB (z-zB)).  waves: [problem][rhs][slab] x 18 doubles
 * (zA, zB, qA re/im, qB re/im, A[3] re/im, B[3] re/im).  Region arrays are
 * concatenated per problem like the eps tables. */
typedef struct {
   int nrhs;                  /* incident fields per problem */
   int nslab;
   const double *waves;
   const int *region_slab;    /* slab index of each region */
   const double *region_deps; /* interleaved complex eps - eps_bg (source strength) */
   const int *region_src;     /* 1: scattered-field source region (frustum) */
} lfasm_incident;

int lfasm_set_incident(lfasm_handle *h, const lfasm_incident *inc);

/* RHS of every incident field (ScatterSource, solver/problem_common.inc:
 * 238-266): b_k = int k0^2 deps E(z) . v_k on the mass rule, dual face
 * transform, sign scatter (ascending element order); essential rows zeroed
 * with `eliminate` (solver/lithofem_solve.cpp:732-742).  deps is the
 * region's eps - eps_bg in source regions (0 elsewhere), as the reference;
 * with per-element eps (lfasm_set_eps_elem) it is eps_e - eps_bg in every
 * region, eps_bg = region eps - region deps, so an element whose eps
 * differs from its slab background is a source wherever it lies.  Device
 * result nrhs x n complex. */
int lfasm_assemble_rhs(lfasm_handle *h, int eliminate);
int lfasm_download_rhs(lfasm_handle *h, double *re, double *im);

/* Adjoint sensitivity per parameter element (sorted unique element ids):
 * s_e = lam^T (dA/deps_e u - db/deps_e) with lam, u global complex vectors
 * (n interleaved, host); db/deps_e is the element RHS of incident field
 * rhs_col of the element's problem at unit source strength, in every
 * region (eps_e enters eps_e - eps_bg linearly; rhs_col < 0: matrix term
 * only).  s_out: nparam interleaved complex. */
int lfasm_eps_sensitivity(lfasm_handle *h, int nparam, const int *elems, const double *lam,
                          const double *u, int rhs_col, double *s_out);

/* The handle's HIP stream (hipStream_t) for callers that record their own
 * events or order their own work after the assembly. */
void *lfasm_stream(lfasm_handle *h);

/* build_csr_symbolic (solver/asm_data.inc:178-210), host C++ (multi-thread):
 * I (height+1) is filled; *Jcol is allocated with malloc and must be freed
 * with lfasm_free.  Returns nnz or a negative status. */
int64_t lfasm_csr_symbolic(int ne, int nd, int height, const int *dof, int *I, int **Jcol);
void lfasm_free(void *p);


#ifdef __cplusplus
}
#endif
#endif
