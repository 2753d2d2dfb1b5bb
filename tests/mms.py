"""Manufactured-solution check of the assembled operator (host side).

Solves  curl curl E - k0^2 eps E = f  on the unit cube with a
non-homogeneous PEC condition (the tangential trace of the exact field is
imposed on every boundary DOF) and measures the L2 error of the discrete
solution, as the reference's M4 verification does (solver/mms_test.cpp:1-12,
tests/test_mms.py:1-8: rate >= p - 0.2, complex symmetry < 1e-12, solve
residual < 1e-10).  The exact field is a divergence-free sine field plus a
gradient (which lies in the curl-curl kernel), so both operator terms and
the face-DOF transform are exercised.

The matrix comes from a caller-supplied assembler (the C oracle on the CPU,
the HIP library on a GPU); load vector, boundary values, solve and error
integration are plain numpy/scipy here.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from lithofem_amd.fespace import build_asm_data, csr_symbolic, interpolate
from lithofem_amd.mesh import structured_cell_mesh
from lithofem_amd.nedelec import FACES, ND_TINV, nd_tet
from lithofem_amd.problem import Problem, Region
from lithofem_amd.quadrature import tet_rule

K0 = 2.0
FREQ = np.array([[1.3, 2.1], [1.7, 0.9], [2.3, 1.1]])   # (a, b) of each component
AMP = np.array([1.0 + 0.5j, -0.7 + 0.2j, 0.4 - 0.9j])
GRAD = np.array([1.1, 0.8, 1.4])                       # phi = sin sin sin


def exact(X: np.ndarray) -> np.ndarray:
    """E = S + grad(phi): S_x = A_x sin(a y) sin(b z) (cyclic), div S = 0."""
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    c = (x, y, z)
    E = np.empty((len(X), 3), dtype=np.complex128)
    for d in range(3):
        u, v = c[(d + 1) % 3], c[(d + 2) % 3]
        a, b = FREQ[d]
        E[:, d] = AMP[d] * np.sin(a * u) * np.sin(b * v)
    g0, g1, g2 = GRAD
    E[:, 0] += g0 * np.cos(g0 * x) * np.sin(g1 * y) * np.sin(g2 * z)
    E[:, 1] += g1 * np.sin(g0 * x) * np.cos(g1 * y) * np.sin(g2 * z)
    E[:, 2] += g2 * np.sin(g0 * x) * np.sin(g1 * y) * np.cos(g2 * z)
    return E


def source(X: np.ndarray, eps: complex) -> np.ndarray:
    """f = curl curl E - k0^2 eps E; curl curl S = -Laplace S = (a^2 + b^2) S
    per component, curl curl grad(phi) = 0."""
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    c = (x, y, z)
    f = -K0 ** 2 * eps * exact(X)
    for d in range(3):
        u, v = c[(d + 1) % 3], c[(d + 2) % 3]
        a, b = FREQ[d]
        f[:, d] += (a * a + b * b) * AMP[d] * np.sin(a * u) * np.sin(b * v)
    return f


def cube_mesh(n: int, warp: float = 0.08):
    """n^3 Kuhn-split cubes on [0, 1]^3, interior vertices moved by a smooth
    map that fixes the boundary; every boundary face is PEC."""
    g = np.linspace(0.0, 1.0, n + 1)
    m = structured_cell_mesh(g, g.copy(), g.copy(), lambda x, y, z: np.ones_like(x, dtype=np.int32),
                             periodic=(False, False), warp_amp=warp, warp_z=(0.0, 1.0))
    fv = np.sort(m.elem_verts[:, np.array(FACES)].reshape(-1, 3), axis=1)
    key = (fv[:, 0] * m.nv + fv[:, 1]) * m.nv + fv[:, 2]
    uk, cnt = np.unique(key, return_counts=True)
    once = np.isin(key, uk[cnt == 1])
    m.bdr_faces = fv[once]
    return m


def problem(eps: complex) -> Problem:
    return Problem(k0=K0, kx=0.0, ky=0.0, z_min=-1e9, z_max=1e9, pml_thick=1.0,
                   regions=[Region(complex(eps))])


def _face_maps(ad):
    """Per element and face pair the 2x2 map from the element's global-DOF
    coordinates (sign applied) to reference-basis coefficients,
    M = [[t0, t2], [t1, t3]] of the orientation's TInv (the inverse of the
    map interpolate() undoes)."""
    Ti = ND_TINV.reshape(6, 4)
    return np.stack([np.stack([Ti[:, 0], Ti[:, 2]], -1), np.stack([Ti[:, 1], Ti[:, 3]], -1)], 1)


def _local_coeffs(ad, g):
    """Reference-basis coefficients of the discrete field, per element."""
    loc = g[ad.dof] * ad.sign
    if ad.p >= 2:
        M = _face_maps(ad)
        ntd = ad.p * (ad.p - 1)
        for f in range(4):
            Mf = M[ad.fo[:, f]]
            for i in range(ntd // 2):
                d0 = 6 * ad.p + f * ntd + 2 * i
                loc[:, d0:d0 + 2] = np.einsum("eij,ej->ei", Mf, loc[:, d0:d0 + 2].copy())
    return loc


def _points(ad, order):
    pts, w = tet_rule(order)
    ne = ad.ne
    J = ad.J.reshape(ne, 3, 3)
    X = ad.v0.reshape(ne, 1, 3) + np.einsum("eij,qj->eqi", J, pts)
    phi = nd_tet(ad.p).vshape(pts)                                    # (nq, nd, 3)
    Ji = ad.Jinv.reshape(ne, 3, 3)
    # covariant Piola: v = J^{-T} vhat
    V = np.einsum("eji,qkj->eqki", Ji, phi)                           # (ne, nq, nd, 3)
    return X, w, V


def load_vector(ad, eps, order):
    X, w, V = _points(ad, order)
    F = source(X.reshape(-1, 3), eps).reshape(ad.ne, len(w), 3)
    lref = np.einsum("q,e,eqd,eqkd->ek", w, ad.detJ, F, V)            # reference-basis loads
    loc = lref.copy()
    if ad.p >= 2:
        M = _face_maps(ad)
        ntd = ad.p * (ad.p - 1)
        for f in range(4):
            Mf = M[ad.fo[:, f]]
            for i in range(ntd // 2):
                d0 = 6 * ad.p + f * ntd + 2 * i
                loc[:, d0:d0 + 2] = np.einsum("eji,ej->ei", Mf, lref[:, d0:d0 + 2])
    b = np.zeros(ad.height, dtype=np.complex128)
    np.add.at(b, ad.dof.ravel(), (loc * ad.sign).ravel())
    return b


def l2_error(ad, g, order):
    X, w, V = _points(ad, order)
    c = _local_coeffs(ad, g)
    Eh = np.einsum("ek,eqkd->eqd", c, V)
    E = exact(X.reshape(-1, 3)).reshape(Eh.shape)
    d2 = np.einsum("q,e,eqd->", w, ad.detJ, np.abs(Eh - E) ** 2)
    return float(np.sqrt(d2))


def solve_level(n, p, eps, assemble):
    """One refinement level: returns (h, L2 error, symmetry, residual)."""
    mesh = cube_mesh(n)
    pb = problem(eps)
    ad = build_asm_data(mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    vre, vim = assemble(ad, pb, I, J)
    A = sp.csr_matrix((vre + 1j * vim, J, I), shape=(ad.height, ad.height))
    sym = spla.norm(A - A.T) / spla.norm(A)
    b = load_vector(ad, eps, 2 * p + 2)
    ess = ad.ess
    free = np.setdiff1d(np.arange(ad.height), ess)
    x = np.zeros(ad.height, dtype=np.complex128)
    x[ess] = interpolate(ad, exact)[ess]
    rhs = b - A @ x
    Aff = A[free][:, free].tocsc()
    x[free] = spla.spsolve(Aff, rhs[free])
    res = np.linalg.norm(Aff @ x[free] - rhs[free]) / np.linalg.norm(rhs[free])
    return 1.0 / n, l2_error(ad, x, 2 * p + 2), float(sym), float(res)


def convergence(p, eps, ns, assemble):
    out = [solve_level(n, p, eps, assemble) for n in ns]
    (h1, e1, _, _), (h2, e2, _, _) = out[-2], out[-1]
    rate = np.log(e1 / e2) / np.log(h1 / h2)
    return rate, out
