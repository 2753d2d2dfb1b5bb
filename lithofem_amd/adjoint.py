"""Per-element eps gradients through the GPU assembly (adjoint method).

Forward: per-element complex eps on a set of parameter elements -> GPU
assembly of the eliminated system A(eps) u = b(eps) (matrix with
lfasm_compute_local + lfasm_scatter_eliminate, RHS of one incident field
with lfasm_assemble_rhs) -> host direct factorisation and solve.  The
reference factorises on the host with UMFPACK (solver/lithofem_solve.cpp:
366-400); SuiteSparse is not installed here, so scipy's SuperLU stands in,
with the same role: one factorisation per forward, reused by the adjoint.

Backward: A^T lam = conj(dL/du*) by a transpose solve with the forward
factors (essential entries of lam zeroed, matching the eliminated
system), then the hand-written kernel lfasm_eps_sensitivity gives
s_e = lam^T (dA/deps_e u - db/deps_e) for every parameter element.
PyTorch passes and expects complex gradients as 2 dL/d(conj z); with
c = conj(grad_u) = 2 (dL/du)^T the kernel's s_e is 2 lam_true^T (...), and
-conj(s_e) = 2 conj(dL/deps_e) is the gradient torch expects
(dL = 2 Re(dL/deps d eps) with dL/deps_e = -lam_true^T (dA/deps_e u -
db/deps_e)).

The module also provides the plane-sampling operator (field values of a
DOF vector at points, as a sparse matrix) and the objective used for the
inverse-design check: reflected +-1 diffraction-order efficiencies plus the
mean squared deviation of |E|^2 on a slice from a target.
"""

from __future__ import annotations

import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import gpu
from .fespace import AsmData
from .nedelec import ND_TINV, nd_tet


# ---------------------------------------------------------------- host solve

class HostLU:
    """Sparse LU of the eliminated complex CSR system (SuperLU)."""

    def __init__(self, I, J, vals, n, refine=1):
        self.A = sp.csr_matrix((vals, J, I), shape=(n, n))
        # the pattern is structurally symmetric: minimum degree on A^T + A
        # with diagonal pivots keeps the 3-D fill far below a column ordering
        # (threshold pivoting off the diagonal breaks the symmetric ordering
        # and multiplies the fill); iterative refinement against A restores
        # the accuracy static pivoting gives away
        self.lu = spla.splu(self.A.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                            options=dict(SymmetricMode=True))
        self.refine = refine

    def _solve(self, b, trans):
        b = np.asarray(b, dtype=np.complex128)
        A = self.A if trans == "N" else self.A.T
        x = self.lu.solve(b, trans=trans)
        for _ in range(self.refine):
            x = x + self.lu.solve(b - A @ x, trans=trans)
        return x

    def solve(self, b):
        return self._solve(b, "N")

    def solve_transpose(self, b):
        return self._solve(b, "T")


# ---------------------------------------------------------------- sampling

def locate(ad: AsmData, pts: np.ndarray, cand=None):
    """Element and reference coordinates of points (brute-force over the
    candidate elements, vectorised): the element whose barycentric
    coordinates are all >= -tol."""
    J = ad.J.reshape(-1, 3, 3)
    Ji = ad.Jinv.reshape(-1, 3, 3)
    elems = np.arange(ad.ne) if cand is None else np.asarray(cand)
    out_e = np.full(len(pts), -1, np.int64)
    out_xi = np.zeros((len(pts), 3))
    for s in range(0, len(pts), 256):
        P = pts[s:s + 256]
        d = P[:, None, :] - ad.v0[elems][None, :, :]                 # (np, ne, 3)
        xi = np.einsum("eij,pej->pei", Ji[elems], d)
        lam0 = 1.0 - xi.sum(-1)
        bmin = np.minimum(xi.min(-1), lam0)
        best = np.argmax(bmin, axis=1)
        out_e[s:s + 256] = elems[best]
        out_xi[s:s + 256] = xi[np.arange(len(P)), best]
    del J
    return out_e, out_xi


def sampling_matrix(ad: AsmData, pts: np.ndarray, periodic=None) -> sp.csr_matrix:
    """S (3 npts x n): S @ u = field components at the points.  `periodic`
    = (Lx, Ly) wraps points into the cell [0, Lx) x [0, Ly) first."""
    pts = np.array(pts, dtype=float)
    if periodic is not None:
        pts[:, 0] = np.mod(pts[:, 0], periodic[0])
        pts[:, 1] = np.mod(pts[:, 1], periodic[1])
    e, xi = locate(ad, pts)
    el = nd_tet(ad.p)
    nd, p = ad.nd, ad.p
    vs = el.vshape(xi)                                      # (np, nd, 3) reference
    Ji = ad.Jinv.reshape(-1, 3, 3)[e]                       # (np, 3, 3)
    v = np.einsum("pkj,pji->pki", vs, Ji)                   # physical covariant
    # field = sum_k l_k v_k with l = M (sign * g) on face pairs
    # (l_a = t0 g_a + t2 g_b, l_b = t1 g_a + t3 g_b): weights on g_a, g_b are
    # sign_a (t0 v_a + t1 v_b) and sign_b (t2 v_a + t3 v_b)
    W = v.copy()
    if p >= 2:
        ntd = p * (p - 1)
        Ti = ND_TINV.reshape(6, 4)
        for f in range(4):
            T = Ti[ad.fo[e, f]]                             # (np, 4)
            for i in range(ntd // 2):
                d0 = 6 * p + f * ntd + 2 * i
                va, vb = v[:, d0], v[:, d0 + 1]
                W[:, d0] = T[:, 0, None] * va + T[:, 1, None] * vb
                W[:, d0 + 1] = T[:, 2, None] * va + T[:, 3, None] * vb
    W = W * ad.sign[e][:, :, None]
    rows = (np.arange(len(pts))[:, None, None] * 3 + np.arange(3)[None, None, :])
    rows = np.broadcast_to(rows, W.shape)
    cols = np.broadcast_to(ad.dof[e][:, :, None], W.shape)
    S = sp.coo_matrix((W.ravel(), (rows.ravel(), cols.ravel())),
                      shape=(3 * len(pts), ad.height)).tocsr()
    S.sum_duplicates()
    return S


# the reference's plane sampler moves every sample point by a small fixed
# offset so that no point lies on an element face or edge and the element
# that evaluates it is deterministic (solver/lithofem_solve.cpp:961-964)
PLANE_NUDGE = (3.1e-7, 1.7e-7, -4.9e-7)


def plane_points(lx, ly, z, nx, ny, nudge=PLANE_NUDGE):
    """Cell-centred nx x ny grid on the plane z (x index first), shifted by
    `nudge` as the reference's plane sampler does."""
    x = (np.arange(nx) + 0.5) * lx / nx + nudge[0]
    y = (np.arange(ny) + 0.5) * ly / ny + nudge[1]
    X, Y = np.meshgrid(x, y, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, float(z) + nudge[2])], 1)


# ---------------------------------------------------------------- forward model

class EpsForwardModel:
    """GPU assembly + host solve as a function of per-element eps."""

    def __init__(self, ad: AsmData, problem, I, J, params, incident=None, rhs_col=0,
                 device=0, refine=1):
        self.ad, self.problem = ad, problem
        self.I, self.J = I, J
        self.params = np.asarray(params, dtype=np.int32)
        self.eps0 = np.array([problem.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
        self.rhs_col = rhs_col if incident is not None else -1
        self.g = gpu.AsmGpuGlobal(device)
        self.g.init(ad, problem)
        self.g.build_global(I, J, ad.ess)
        if incident is not None:
            self.g.set_incident(incident)
        self.refine = refine         # iterative-refinement steps of the host solves
        self.ess_mask = np.zeros(ad.height, bool)
        self.ess_mask[ad.ess] = True
        self.lu = None
        self.u = None
        self.times = {}          # seconds of the last solve / adjoint, by phase

    def close(self):
        self.g.close()

    def solve(self, eps_params: np.ndarray) -> np.ndarray:
        t = [time.perf_counter()]
        ee = self.eps0.copy()
        ee[self.params] = eps_params
        self.g.set_eps_elem(ee)
        self.g.compute_local()
        self.g.scatter_eliminate(True)
        vre, vim = self.g.download_values()       # waits for the assembly kernels
        t.append(time.perf_counter())
        if self.rhs_col >= 0:
            b = self.g.assemble_rhs(eliminate=True)[self.rhs_col]
        else:
            b = np.zeros(self.ad.height, dtype=np.complex128)
        t.append(time.perf_counter())
        self.lu = HostLU(self.I, self.J, vre + 1j * vim, self.ad.height, refine=self.refine)
        t.append(time.perf_counter())
        self.u = self.lu.solve(b)
        t.append(time.perf_counter())
        self.times = dict(zip(("assembly_and_download", "rhs", "factorisation", "solve"),
                              map(float, np.diff(t))))
        return self.u

    def adjoint_sensitivity(self, grad_u_conj: np.ndarray) -> np.ndarray:
        """s_e for every parameter element, given c = conj(dL/du*) (so that
        A^T lam = c)."""
        t0 = time.perf_counter()
        lam = self.lu.solve_transpose(grad_u_conj)
        lam[self.ess_mask] = 0.0
        t1 = time.perf_counter()
        s = self.g.eps_sensitivity(self.params, lam, self.u, self.rhs_col)
        t2 = time.perf_counter()
        self.times.update(transpose_solve=t1 - t0, sensitivity=t2 - t1)
        return s


def make_torch_function():
    import torch

    class EpsToField(torch.autograd.Function):
        """u = A(eps)^{-1} b(eps) with the GPU assembly; backward through the
        transpose solve and the eps-sensitivity kernel."""

        @staticmethod
        def forward(ctx, eps, model):
            u = model.solve(eps.detach().cpu().numpy())
            ctx.model = model
            return torch.from_numpy(u.copy())

        @staticmethod
        def backward(ctx, grad_u):
            c = np.conj(grad_u.detach().cpu().numpy())
            s = ctx.model.adjoint_sensitivity(c)
            return torch.from_numpy(-np.conj(s)), None

    return EpsToField


# ---------------------------------------------------------------- objective

def order_efficiencies(E, kpar, k0, lx, ly, orders, eps_top=1.0):
    """Reflected diffraction-order efficiencies from the scattered envelope
    E (nx, ny, 3) sampled on a plane above the structure: the Fourier
    coefficient of order (m, n) of the envelope is the amplitude of the
    plane wave with k_par + (2 pi m / lx, 2 pi n / ly); efficiency =
    |amplitude|^2 Re(kz_mn) / kz_inc for a unit incident amplitude."""
    import torch
    nx, ny = E.shape[0], E.shape[1]
    F = torch.fft.fft2(E, dim=(0, 1)) / (nx * ny)
    kz0 = np.sqrt(complex(k0 * k0 * eps_top - kpar[0] ** 2 - kpar[1] ** 2)).real
    out = []
    for (m, n) in orders:
        kx = kpar[0] + 2 * np.pi * m / lx
        ky = kpar[1] + 2 * np.pi * n / ly
        kz = np.sqrt(complex(k0 * k0 * eps_top - kx * kx - ky * ky)).real
        a = F[m % nx, n % ny]
        out.append((a.abs() ** 2).sum() * (kz / kz0))
    return out


class DesignObjective:
    """J(u) = sum of the reflected +-1 (x) order efficiencies + mean squared
    deviation of |E_total|^2 on a z slice from a target map; u is the
    scattered envelope, E_total = u + E_inc on the slice."""

    def __init__(self, ad, problem, lx, ly, z_top, z_slice, n_grid, target, e_inc_slice,
                 n_orders_grid=16):
        import torch
        pts_top = plane_points(lx, ly, z_top, n_orders_grid, n_orders_grid)
        pts_sl = plane_points(lx, ly, z_slice, n_grid, n_grid)
        self.S_top = _torch_sparse(sampling_matrix(ad, pts_top))
        self.S_sl = _torch_sparse(sampling_matrix(ad, pts_sl))
        self.n_o = n_orders_grid
        self.n_g = n_grid
        self.target = torch.as_tensor(target, dtype=torch.float64).reshape(-1)
        self.e_inc = torch.as_tensor(e_inc_slice, dtype=torch.complex128).reshape(-1, 3)
        self.kpar = (problem.kx, problem.ky)
        self.k0 = problem.k0
        self.lx, self.ly = lx, ly

    def __call__(self, u):
        import torch
        E_top = (self.S_top @ u.reshape(-1, 1)).reshape(self.n_o, self.n_o, 3)
        eff = order_efficiencies(E_top, self.kpar, self.k0, self.lx, self.ly, [(1, 0), (-1, 0)])
        E_sl = (self.S_sl @ u.reshape(-1, 1)).reshape(-1, 3) + self.e_inc
        I = (E_sl.abs() ** 2).sum(-1)
        return eff[0] + eff[1] + torch.mean((I - self.target) ** 2)


def _torch_sparse(S):
    import torch
    S = S.tocoo()
    idx = torch.as_tensor(np.vstack([S.row, S.col]), dtype=torch.int64)
    return torch.sparse_coo_tensor(idx, torch.as_tensor(S.data, dtype=torch.complex128),
                                   S.shape).coalesce()
