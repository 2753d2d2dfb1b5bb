"""Python mirror of the reference's AsmGpu / AsmGpuGlobal interface over the
C-ABI of include/lithofem_asm.h (liblfasm.so, built by lithofem_amd.build).

Method names follow solver/asm_gpu.inc:271-599 (Init, ComputeLocal,
BuildGlobal, ScatterEliminate, Residual, DeviceI/J/Values, DownloadLocal,
DownloadValues) in snake_case.  Errors raise :class:`AsmGpuError` carrying
the library's status code and message -- the caller falls back to CPU
assembly exactly where the reference driver does
(solver/lithofem_solve.cpp:796-799).  There is no CPU fallback inside this
module: without the built library every call raises.
"""

from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

from .nedelec import ND_TINV

LIB_PATH = Path(__file__).resolve().parent / "liblfasm.so"

CURLCURL, MASS, CROSS0, CROSS1, KLK = 1, 2, 4, 8, 16
MASS_ALL = MASS | CROSS0 | CROSS1 | KLK

E_ARG, E_HIP, E_FORCED, E_UNSUPPORTED, E_STATE, E_PATTERN = -1, -2, -3, -4, -5, -6

# most dofs lfasm_residual_static's static workspace holds
RESIDUAL_STATIC_MAX_N = 1572864

# kernel selector of lfasm_compute_local_kernel
KERNEL_AUTO, KERNEL_VALU = 0, 1
N_MFMA_VARIANTS = 8


def kernel_mfma(variant: int = 0) -> int:
    return 2 + variant


class AsmGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lfasm error {code}: {msg}")
        self.code = code


class _Rule(ct.Structure):
    _fields_ = [("nq", ct.c_int), ("ip", ct.c_void_p), ("w", ct.c_void_p),
                ("vshape", ct.c_void_p), ("curl", ct.c_void_p)]


class _Pml(ct.Structure):
    _fields_ = [(n, ct.c_double) for n in
                ("z_min", "z_max", "thick", "order", "smax_bot", "smax_top", "k0", "kx", "ky")]


class _Incident(ct.Structure):
    _fields_ = [("nrhs", ct.c_int), ("nslab", ct.c_int), ("waves", ct.c_void_p),
                ("region_slab", ct.c_void_p), ("region_deps", ct.c_void_p),
                ("region_src", ct.c_void_p)]


class _Mesh(ct.Structure):
    _fields_ = [("ne", ct.c_int), ("nd", ct.c_int), ("p", ct.c_int), ("height", ct.c_int),
                ("dof", ct.c_void_p), ("sign", ct.c_void_p), ("fo", ct.c_void_p),
                ("attr", ct.c_void_p), ("J", ct.c_void_p), ("Jinv", ct.c_void_p),
                ("detJ", ct.c_void_p), ("v0", ct.c_void_p), ("rule_cc", _Rule),
                ("rule_mass", _Rule), ("nreg", ct.c_int), ("eps_tab", ct.c_void_p),
                ("eps_elem", ct.c_void_p), ("pml", _Pml), ("TInv", ct.c_void_p)]


_lib = None


def lib():
    """Load liblfasm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise AsmGpuError(E_STATE, f"{LIB_PATH} not built (python -m lithofem_amd.build)")
    # One HIP runtime per process: when PyTorch-ROCm is installed its
    # bundled runtime is loaded first, and liblfasm.so binds to it (same
    # soname) instead of bringing a second copy next to it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ct.CDLL(str(LIB_PATH))
    vp, i, i64, d = ct.c_void_p, ct.c_int, ct.c_int64, ct.c_double
    sig = {
        "lfasm_create": ([i, ct.POINTER(vp)], i),
        "lfasm_destroy": ([vp], None),
        "lfasm_last_error": ([vp], ct.c_char_p),
        "lfasm_pml_smax": ([d, d, d, d, d], d),
        "lfasm_init": ([vp, vp], i),
        "lfasm_set_eps_elem": ([vp, vp], i),
        "lfasm_build_global": ([vp, vp, vp, i64, vp], i),
        "lfasm_compute_local": ([vp, i, i, i], i),
        "lfasm_compute_local_kernel": ([vp, i, i, i, i], i),
        "lfasm_set_local_kernel": ([vp, i], i),
        "lfasm_scatter_eliminate": ([vp, i], i),
        "lfasm_device_csr": ([vp, vp, vp, vp], i),
        "lfasm_download_values": ([vp, vp, vp], i),
        "lfasm_download_local": ([vp, vp], i),
        "lfasm_residual": ([vp, vp, vp, vp], i),
        "lfasm_residual_static": ([vp, i64, vp, vp, vp], i),
        "lfasm_sizes": ([vp, vp, vp, vp, vp], i),
        "lfasm_timings": ([vp, vp, i], i),
        "lfasm_stream": ([vp], vp),
        "lfasm_csr_symbolic": ([i, i, i, vp, vp, ct.POINTER(vp)], i64),
        "lfasm_free": ([vp], None),
        "lfasm_init_batch": ([vp, vp, i], i),
        "lfasm_batch_offsets": ([vp, vp, vp], i),
        "lfasm_variants_prepare": ([vp, i, vp], i),
        "lfasm_variants_assemble": ([vp, i, vp, i], i),
        "lfasm_variant_values": ([vp, i, ct.POINTER(vp)], i),
        "lfasm_download_variant": ([vp, i, vp, vp], i),
        "lfasm_set_incident": ([vp, vp], i),
        "lfasm_assemble_rhs": ([vp, i], i),
        "lfasm_download_rhs": ([vp, vp, vp], i),
        "lfasm_eps_sensitivity": ([vp, i, vp, vp, vp, i, vp], i),
        "lfasm_set_tensor_path": ([vp, i], i),
        "lfasm_set_element_concurrency": ([vp, i, i], i),
        "lfasm_set_variant_kernel": ([vp, i], i),
        "lfasm_set_quad_residency": ([vp, i], i),
        "lfasm_angles_assemble": ([vp, i, vp, i, i, i, i], i),
        "lfasm_angle_values": ([vp, i, ct.POINTER(vp)], i),
        "lfasm_download_angle": ([vp, i, vp, vp], i),
        "lfasm_step_capture": ([vp, i, i, i, i], i),
        "lfasm_step_launch": ([vp, i], i),
        "lfasm_build_global_device": ([vp, vp], i),
        "lfasm_init_vertices": ([vp, vp, vp], i),
        "lfasm_download_geometry": ([vp, vp, vp, vp, vp], i),
        "lfasm_download_contributors": ([vp, vp, vp], i),
        "lfasm_download_pattern": ([vp, vp, vp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def csr_symbolic(dof: np.ndarray, height: int):
    """Host multi-threaded CSR pattern build (lfasm_csr_symbolic)."""
    ne, nd = dof.shape
    dof = np.ascontiguousarray(dof, dtype=np.int32)
    I = np.zeros(height + 1, dtype=np.int32)
    pj = ct.c_void_p()
    nnz = lib().lfasm_csr_symbolic(ne, nd, height, _p(dof), _p(I), ct.byref(pj))
    if nnz < 0:
        raise AsmGpuError(int(nnz), "csr_symbolic failed")
    J = np.ctypeslib.as_array(ct.cast(pj, ct.POINTER(ct.c_int32)), shape=(max(nnz, 1),))[:nnz].copy()
    lib().lfasm_free(pj)
    return I, J


class AsmGpuGlobal:
    """One device assembly context (AsmGpu + AsmGpuGlobal of the reference)."""

    def __init__(self, device: int = 0):
        self._h = ct.c_void_p()
        rc = lib().lfasm_create(device, ct.byref(self._h))
        if rc != 0:
            msg = self.last_error()
            self.close()
            raise AsmGpuError(rc, msg)
        self._keep = []
        self.ne = self.nd = self.n = 0
        self.nnz = 0

    # -- lifetime
    def close(self):
        if self._h:
            lib().lfasm_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def last_error(self) -> str:
        s = lib().lfasm_last_error(self._h)
        return s.decode() if s else ""

    def _check(self, rc: int):
        if rc != 0:
            raise AsmGpuError(rc, self.last_error())

    # -- AsmGpu::Init
    def _mesh_struct(self, ad, problem, eps_elem, keep):
        def k(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        def rule(r):
            return _Rule(r.nq, _p(k(r.ip, np.float64)), _p(k(r.w, np.float64)),
                         _p(k(r.vshape, np.float64)), _p(k(r.curl, np.float64)))

        ee = None
        if eps_elem is not None:
            ee = k(np.asarray(eps_elem, dtype=np.complex128).view(np.float64), np.float64)
        return _Mesh(ad.ne, ad.nd, ad.p, ad.height,
                     _p(k(ad.dof, np.int32)), _p(k(ad.sign, np.float64)),
                     _p(k(ad.fo, np.int32)) if ad.p >= 2 else None,
                     _p(k(ad.attr, np.int32)), _p(k(ad.J, np.float64)),
                     _p(k(ad.Jinv, np.float64)), _p(k(ad.detJ, np.float64)),
                     _p(k(ad.v0, np.float64)), rule(ad.rule_cc), rule(ad.rule_mass),
                     len(problem.regions), _p(k(problem.eps_tab(), np.float64)), _p(ee),
                     _Pml(*problem.pml_params()), _p(k(ND_TINV, np.float64)))

    def init(self, ad, problem, eps_elem=None):
        keep = []
        m = self._mesh_struct(ad, problem, eps_elem, keep)
        self._check(lib().lfasm_init(self._h, ct.byref(m)))
        self.ne, self.nd, self.n = ad.ne, ad.nd, ad.height
        self.nmesh = 1

    def init_vertices(self, ad, problem, elem_xyz, eps_elem=None):
        """AsmGpu::Init with the geometry computed on the device from the
        element vertex coordinates (ne, 4, 3); ad's J, Jinv, detJ, v0 are not
        read."""
        keep = []
        m = self._mesh_struct(ad, problem, eps_elem, keep)
        x = np.ascontiguousarray(elem_xyz, dtype=np.float64)
        assert x.shape == (ad.ne, 4, 3)
        self._check(lib().lfasm_init_vertices(self._h, ct.byref(m), _p(x)))
        self.ne, self.nd, self.n = ad.ne, ad.nd, ad.height
        self.nmesh = 1

    def download_geometry(self):
        J = np.empty((self.ne, 9))
        Ji = np.empty((self.ne, 9))
        det = np.empty(self.ne)
        v0 = np.empty((self.ne, 3))
        self._check(lib().lfasm_download_geometry(self._h, _p(J), _p(Ji), _p(det), _p(v0)))
        return J, Ji, det, v0

    def download_contributors(self):
        ptr = np.empty(self.n + 1, dtype=np.int32)
        lst = np.empty(self.ne * self.nd, dtype=np.int32)
        self._check(lib().lfasm_download_contributors(self._h, _p(ptr), _p(lst)))
        return ptr, lst

    def init_batch(self, items):
        """items: sequence of (AsmData, Problem[, eps_elem]); the meshes are
        assembled as one block-diagonal system."""
        keep = []
        structs = [self._mesh_struct(it[0], it[1], it[2] if len(it) > 2 else None, keep)
                   for it in items]
        arr = (_Mesh * len(structs))(*structs)
        self._check(lib().lfasm_init_batch(self._h, arr, len(structs)))
        self.ne = sum(it[0].ne for it in items)
        self.nd = items[0][0].nd
        self.n = sum(it[0].height for it in items)
        self.nmesh = len(items)

    def batch_offsets(self):
        eo = np.zeros(self.nmesh + 1, np.int64)
        ro = np.zeros(self.nmesh + 1, np.int64)
        rc = lib().lfasm_batch_offsets(self._h, _p(eo), _p(ro))
        if rc < 0:
            self._check(rc)
        return eo, ro

    def set_eps_elem(self, eps_elem):
        if eps_elem is None:
            self._check(lib().lfasm_set_eps_elem(self._h, None))
            return
        a = np.ascontiguousarray(np.asarray(eps_elem, dtype=np.complex128).view(np.float64))
        self._check(lib().lfasm_set_eps_elem(self._h, _p(a)))

    # -- AsmGpuGlobal::BuildGlobal
    def build_global(self, I, Jcol, ess_list=None):
        I = np.ascontiguousarray(I, dtype=np.int32)
        Jcol = np.ascontiguousarray(Jcol, dtype=np.int32)
        ess = np.zeros(self.n, dtype=np.uint8)
        if ess_list is not None and len(ess_list):
            ess[np.asarray(ess_list)] = 1
        self._check(lib().lfasm_build_global(self._h, _p(I), _p(Jcol), len(Jcol), _p(ess)))
        self.nnz = len(Jcol)

    def build_global_device(self, ess_list=None):
        """CSR pattern built on the device (replaces the host
        build_csr_symbolic + build_global pair)."""
        ess = np.zeros(self.n, dtype=np.uint8)
        if ess_list is not None and len(ess_list):
            ess[np.asarray(ess_list)] = 1
        self._check(lib().lfasm_build_global_device(self._h, _p(ess)))
        nnz = np.zeros(1, dtype=np.int64)
        self._check(lib().lfasm_sizes(self._h, None, None, None, _p(nnz)))
        self.nnz = int(nnz[0])

    def download_pattern(self):
        I = np.empty(self.n + 1, dtype=np.int32)
        J = np.empty(self.nnz, dtype=np.int32)
        self._check(lib().lfasm_download_pattern(self._h, _p(I), _p(J)))
        return I, J

    # -- AsmGpu::ComputeLocal
    def compute_local(self, mask_cc=CURLCURL, mask_mass=MASS_ALL, apply_te=True,
                      kernel=KERNEL_AUTO):
        self._check(lib().lfasm_compute_local_kernel(self._h, mask_cc, mask_mass,
                                                     1 if apply_te else 0, kernel))

    def set_local_kernel(self, kernel: int):
        self._check(lib().lfasm_set_local_kernel(self._h, kernel))

    # -- AsmGpuGlobal::ScatterEliminate
    def scatter_eliminate(self, eliminate=True):
        self._check(lib().lfasm_scatter_eliminate(self._h, 1 if eliminate else 0))

    def download_local(self) -> np.ndarray:
        out = np.empty((self.ne, self.nd, self.nd), dtype=np.complex128)
        self._check(lib().lfasm_download_local(self._h, _p(out)))
        return out

    def download_values(self):
        re = np.empty(self.nnz)
        im = np.empty(self.nnz)
        self._check(lib().lfasm_download_values(self._h, _p(re), _p(im)))
        return re, im

    def device_csr(self):
        pi, pj, pv = ct.c_void_p(), ct.c_void_p(), ct.c_void_p()
        self._check(lib().lfasm_device_csr(self._h, ct.byref(pi), ct.byref(pj), ct.byref(pv)))
        return pi.value, pj.value, pv.value

    def residual(self, X2n, B2n) -> float:
        X2n = np.ascontiguousarray(X2n, dtype=np.float64)
        B2n = np.ascontiguousarray(B2n, dtype=np.float64)
        rel = ct.c_double()
        self._check(lib().lfasm_residual(self._h, _p(X2n), _p(B2n), ct.byref(rel)))
        return rel.value

    def residual_static(self, X2n, B2n) -> float:
        """lfasm_residual without device allocations (static workspace)."""
        X2n = np.ascontiguousarray(X2n, dtype=np.float64)
        B2n = np.ascontiguousarray(B2n, dtype=np.float64)
        assert X2n.shape == B2n.shape
        rel = ct.c_double()
        self._check(lib().lfasm_residual_static(self._h, X2n.size, _p(X2n), _p(B2n),
                                                ct.byref(rel)))
        return rel.value

    # -- eps variants (delta assembly)
    def variants_prepare(self, elems):
        el = np.ascontiguousarray(np.asarray(elems, dtype=np.int32))
        self._check(lib().lfasm_variants_prepare(self._h, len(el), _p(el)))
        self.nparam = len(el)

    def variants_assemble(self, deps, eliminate=True):
        d = np.ascontiguousarray(np.asarray(deps, dtype=np.complex128))
        assert d.ndim == 2 and d.shape[1] == self.nparam
        self._check(lib().lfasm_variants_assemble(self._h, d.shape[0], _p(d.view(np.float64)),
                                                  1 if eliminate else 0))
        self.nvar = d.shape[0]

    def download_variant(self, v):
        re = np.empty(self.nnz)
        im = np.empty(self.nnz)
        self._check(lib().lfasm_download_variant(self._h, v, _p(re), _p(im)))
        return re, im

    # -- angle sweeps
    def angles_assemble(self, kxy, mask_cc=CURLCURL, mask_mass=MASS_ALL, apply_te=True,
                        eliminate=True):
        """CSR values of every k_par = (kx, ky) in `kxy` (nang x 2) from the
        basis matrices of the k_par polynomial (lfasm_angles_assemble)."""
        k = np.ascontiguousarray(np.asarray(kxy, dtype=np.float64).reshape(-1, 2))
        self._check(lib().lfasm_angles_assemble(self._h, k.shape[0], _p(k), int(mask_cc),
                                                int(mask_mass), 1 if apply_te else 0,
                                                1 if eliminate else 0))
        self.nang = k.shape[0]

    def download_angle(self, j):
        re = np.empty(self.nnz)
        im = np.empty(self.nnz)
        self._check(lib().lfasm_download_angle(self._h, j, _p(re), _p(im)))
        return re, im

    # -- incident fields, RHS, sensitivity
    def set_incident(self, inc):
        """inc: lithofem_amd.incident.IncidentSet (or any object with
        nrhs, nslab, waves (…, 18), region_slab, region_deps, region_src)."""
        w = np.ascontiguousarray(inc.waves, dtype=np.float64)
        rs = np.ascontiguousarray(inc.region_slab, dtype=np.int32)
        rd = np.ascontiguousarray(np.asarray(inc.region_deps, dtype=np.complex128).view(np.float64))
        rsrc = np.ascontiguousarray(inc.region_src, dtype=np.int32)
        st = _Incident(inc.nrhs, inc.nslab, _p(w), _p(rs), _p(rd), _p(rsrc))
        self._check(lib().lfasm_set_incident(self._h, ct.byref(st)))
        self.nrhs = inc.nrhs

    def assemble_rhs(self, eliminate=True, download=True):
        """RHS of every incident field on the device; returns them as an
        (nrhs, n) complex array when `download`, else None (see
        download_rhs)."""
        self._check(lib().lfasm_assemble_rhs(self._h, 1 if eliminate else 0))
        return self.download_rhs() if download else None

    def download_rhs(self) -> np.ndarray:
        re = np.empty(self.nrhs * self.n)
        im = np.empty(self.nrhs * self.n)
        self._check(lib().lfasm_download_rhs(self._h, _p(re), _p(im)))
        return (re + 1j * im).reshape(self.nrhs, self.n)

    def eps_sensitivity(self, elems, lam, u, rhs_col=-1) -> np.ndarray:
        el = np.ascontiguousarray(np.asarray(elems, dtype=np.int32))
        lam = np.ascontiguousarray(np.asarray(lam, dtype=np.complex128))
        u = np.ascontiguousarray(np.asarray(u, dtype=np.complex128))
        out = np.empty(len(el), dtype=np.complex128)
        self._check(lib().lfasm_eps_sensitivity(self._h, len(el), _p(el), _p(lam.view(np.float64)),
                                                _p(u.view(np.float64)), rhs_col,
                                                _p(out.view(np.float64))))
        return out

    def timings(self) -> dict:
        ms = np.zeros(10)
        self._check(lib().lfasm_timings(self._h, _p(ms), 10))
        return dict(local_ms=ms[0], scatter_ms=ms[1], build_ms=ms[2], batch_ms=ms[3],
                    adjoint_ms=ms[4], local_tensor_ms=ms[5], local_quad_ms=ms[6],
                    step_ms=ms[7], pattern_ms=ms[8], geometry_ms=ms[9])

    # -- captured step (HIP graph)
    def step_capture(self, mask_cc=CURLCURL, mask_mass=MASS_ALL, apply_te=True, eliminate=True):
        self._check(lib().lfasm_step_capture(self._h, mask_cc, mask_mass, 1 if apply_te else 0,
                                             1 if eliminate else 0))

    def step_launch(self, sync=True):
        self._check(lib().lfasm_step_launch(self._h, 1 if sync else 0))

    TENSOR_OFF, TENSOR_AE, TENSOR_FUSED = 0, 1, 2

    def set_tensor_path(self, mode) -> int:
        """Element path of constant-coefficient elements: 0 (or False)
        quadrature kernels, 1 (or True, the default) reference-tensor kernel
        writing Ae, 2 tensor rows evaluated inside the CSR row kernels.
        Returns the number of elements on the tensor path."""
        mode = 1 if mode is True else 0 if mode is False else int(mode)
        rc = lib().lfasm_set_tensor_path(self._h, mode)
        if rc < 0:
            raise AsmGpuError(rc, self.last_error())
        return rc

    def set_element_concurrency(self, grid: int, scatter_overlap: int = 0) -> None:
        """0: the tensor-path kernel, then the quadrature kernel; grid > 0:
        the two side by side, the tensor kernel held to `grid` workgroups
        along its element chunks; scatter_overlap = 1: in the captured step,
        the rows formed only by tensor-path elements are scattered beside
        the quadrature kernel (lfasm_set_element_concurrency)."""
        self._check(lib().lfasm_set_element_concurrency(self._h, int(grid), int(scatter_overlap)))

    def set_quad_residency(self, wg_per_cu: int) -> None:
        """At most `wg_per_cu` quadrature workgroups per CU when the element
        kernels run side by side (0 = no cap; lfasm_set_quad_residency)."""
        self._check(lib().lfasm_set_quad_residency(self._h, int(wg_per_cu)))

    def set_variant_kernel(self, kernel: int) -> None:
        """Form of the eps-variant row kernel: 0 automatic, 1 accumulating
        (LDS atomics), 2 register (read-modify-write), 3 general
        (lfasm_set_variant_kernel)."""
        self._check(lib().lfasm_set_variant_kernel(self._h, int(kernel)))

    def stream(self) -> int:
        return lib().lfasm_stream(self._h) or 0
