"""Nedelec (first kind) H(curl) reference element of order p on the tetrahedron.

The reference reads its basis from MFEM's ``ND_TetrahedronElement``
(``CalcVShape`` / ``CalcCurlShape`` at the rule points, solver/asm_data.inc:
47-73).  This module builds an equivalent nodal basis from scratch so the
standalone path has tables with the same layout and the same DOF structure:

* nd(p) = p (p+2) (p+3) / 2 DOFs (solver/asm_data.inc:80);
* edge DOFs first (6 edges x p), then face DOFs (4 faces x p(p-1), in pairs of
  two tangent directions at one face point), then interior DOFs
  (solver/asm_data.inc:342-353: ``foff = 6p``, ``d0 = foff + f p(p-1) + 2i``);
* edge table (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) and face table
  (1,2,3),(0,3,2),(0,1,3),(0,2,1).

DOF functionals are tangential components at points: edge DOF m of edge
(i,j) is u(x_m).(V_j - V_i) with x_m = V_i + (m+1)/(p+1) (V_j - V_i); the two
DOFs of a face point are u.(B-A) and u.(C-A) for the face vertex sequence
(A,B,C); interior DOFs are the three Cartesian components at interior lattice
points.  Under a change of a face's vertex sequence by the triangle
orientation o, the pair transforms with the 2x2 block T(o) of the
reference's ``asm_nd_T`` table (solver/asm_data.inc:17-25); the assembly
applies ``TInv(fo)`` to element matrices (solver/asm_data.inc:26-34,337-369).

The polynomial space is ND_p = P_{p-1}^3 + { x cross q : q homogeneous of
degree p-1 }; the basis is the dual basis of the functionals above.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

VERTS = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))
FACES = ((1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1))

# Triangle orientations: local vertex sequence = canonical[ORIENT[o]].
ORIENT = ((0, 1, 2), (1, 0, 2), (2, 0, 1), (2, 1, 0), (1, 2, 0), (0, 2, 1))

# 2x2 face-DOF blocks, column-major (m00, m10, m01, m11) per orientation.
# These are the values of solver/asm_data.inc:17-34 (the reference's copy of
# MFEM's ND_DofTransformation tables); the format has to match because the
# C-ABI takes the TInv table as an input exactly as the reference uploads it.
ND_T = np.array([
    1.0, 0.0, 0.0, 1.0,
    -1.0, -1.0, 0.0, 1.0,
    0.0, 1.0, -1.0, -1.0,
    1.0, 0.0, -1.0, -1.0,
    -1.0, -1.0, 1.0, 0.0,
    0.0, 1.0, 1.0, 0.0,
])
ND_TINV = np.array([
    1.0, 0.0, 0.0, 1.0,
    -1.0, -1.0, 0.0, 1.0,
    -1.0, -1.0, 1.0, 0.0,
    1.0, 0.0, -1.0, -1.0,
    0.0, 1.0, -1.0, -1.0,
    0.0, 1.0, 1.0, 0.0,
])


def nd_count(p: int) -> int:
    return p * (p + 2) * (p + 3) // 2


def face_points_bary(p: int) -> list[tuple[int, int, int]]:
    """Interior lattice points of a triangle, as integer barycentrics
    (i, j, k) with i+j+k = p+1, all >= 1, in a fixed enumeration order."""
    n = p + 1
    out = []
    for k in range(1, n):
        for j in range(1, n - k):
            i = n - j - k
            if i >= 1:
                out.append((i, j, k))
    return out


def interior_points_bary(p: int) -> list[tuple[int, int, int, int]]:
    n = p + 1
    out = []
    for l3 in range(1, n):
        for l2 in range(1, n - l3):
            for l1 in range(1, n - l3 - l2):
                l0 = n - l1 - l2 - l3
                if l0 >= 1:
                    out.append((l0, l1, l2, l3))
    return out


def dof_functionals(p: int) -> tuple[np.ndarray, np.ndarray]:
    """Points (nd x 3) and tangents (nd x 3) of the local DOF functionals."""
    pts, tans = [], []
    for (i, j) in EDGES:
        t = VERTS[j] - VERTS[i]
        for m in range(p):
            s = (m + 1) / (p + 1)
            pts.append(VERTS[i] + s * t)
            tans.append(t)
    for (a, b, c) in FACES:
        t1 = VERTS[b] - VERTS[a]
        t2 = VERTS[c] - VERTS[a]
        for (i, j, k) in face_points_bary(p):
            x = (i * VERTS[a] + j * VERTS[b] + k * VERTS[c]) / (p + 1)
            pts.append(x)
            tans.append(t1)
            pts.append(x)
            tans.append(t2)
    for (l0, l1, l2, l3) in interior_points_bary(p):
        x = (l0 * VERTS[0] + l1 * VERTS[1] + l2 * VERTS[2] + l3 * VERTS[3]) / (p + 1)
        for d in range(3):
            pts.append(x)
            tans.append(np.eye(3)[d])
    pts = np.array(pts)
    tans = np.array(tans)
    assert len(pts) == nd_count(p)
    return pts, tans


def _mono(p: int) -> list[tuple[int, int, int]]:
    return [(i, j, k) for s in range(p + 1) for i in range(s, -1, -1)
            for j in range(s - i, -1, -1) for k in [s - i - j]]


def _eval_mono(expo, x: np.ndarray) -> np.ndarray:
    """(n_points x n_mono) values."""
    e = np.array(expo)
    return np.prod(x[:, None, :] ** e[None, :, :], axis=2)


def _eval_mono_grad(expo, x: np.ndarray) -> np.ndarray:
    """(n_points x n_mono x 3) gradients."""
    e = np.array(expo, dtype=np.float64)
    out = np.zeros((len(x), len(expo), 3))
    for d in range(3):
        ed = e.copy()
        coef = ed[:, d].copy()
        ed[:, d] = np.maximum(ed[:, d] - 1, 0)
        out[:, :, d] = coef[None, :] * np.prod(x[:, None, :] ** ed[None, :, :], axis=2)
    return out


class NDTet:
    """Dual nodal basis of ND_p on the reference tetrahedron.

    ``coef`` has shape (nd, 3, n_mono): basis k, component d, monomial m.
    """

    def __init__(self, p: int):
        if p < 1 or p > 6:
            raise ValueError("order must be 1..6")
        self.p = p
        self.nd = nd_count(p)
        self.expo = _mono(p)
        nm = len(self.expo)
        idx = {e: m for m, e in enumerate(self.expo)}
        span = []
        for m, e in enumerate(self.expo):
            if sum(e) <= p - 1:
                for d in range(3):
                    v = np.zeros((3, nm))
                    v[d, m] = 1.0
                    span.append(v)
        for e in self.expo:
            if sum(e) != p - 1:
                continue
            # x cross (m e_d), with x = (x, y, z)
            sx = idx[(e[0] + 1, e[1], e[2])]
            sy = idx[(e[0], e[1] + 1, e[2])]
            sz = idx[(e[0], e[1], e[2] + 1)]
            for d in range(3):
                v = np.zeros((3, nm))
                if d == 0:      # (0, z m, -y m)
                    v[1, sz] = 1.0
                    v[2, sy] = -1.0
                elif d == 1:    # (-z m, 0, x m)
                    v[0, sz] = -1.0
                    v[2, sx] = 1.0
                else:           # (y m, -x m, 0)
                    v[0, sy] = 1.0
                    v[1, sx] = -1.0
                span.append(v)
        S = np.array([v.ravel() for v in span])
        _, sv, vt = np.linalg.svd(S, full_matrices=False)
        rank = int(np.sum(sv > 1e-10 * sv[0]))
        if rank != self.nd:
            raise RuntimeError(f"ND_{p} span rank {rank} != {self.nd}")
        basis = vt[:rank].reshape(rank, 3, nm)
        pts, tans = dof_functionals(p)
        vals = self._eval_coef(basis, pts)            # (nd_pts, rank, 3)
        V = np.einsum("kjd,kd->kj", vals, tans)       # DOF_k(b_j)
        self.vander_cond = float(np.linalg.cond(V))
        X = np.linalg.inv(V)
        self.coef = np.einsum("ji,jdm->idm", X, basis)
        self.dof_points, self.dof_tangents = pts, tans

    def _eval_coef(self, coef: np.ndarray, x: np.ndarray) -> np.ndarray:
        mv = _eval_mono(self.expo, np.atleast_2d(x))   # (np, nm)
        return np.einsum("pm,kdm->pkd", mv, coef)

    def vshape(self, x: np.ndarray) -> np.ndarray:
        """(n_points, nd, 3) basis values (reference element)."""
        return self._eval_coef(self.coef, x)

    def curl(self, x: np.ndarray) -> np.ndarray:
        """(n_points, nd, 3) reference curls."""
        g = _eval_mono_grad(self.expo, np.atleast_2d(x))  # (np, nm, 3)
        # du_d/dx_j = sum_m coef[k,d,m] g[p,m,j]
        du = np.einsum("kdm,pmj->pkdj", self.coef, g)
        c = np.empty(du.shape[:3])
        c[..., 0] = du[..., 2, 1] - du[..., 1, 2]
        c[..., 1] = du[..., 0, 2] - du[..., 2, 0]
        c[..., 2] = du[..., 1, 0] - du[..., 0, 1]
        return c


@lru_cache(maxsize=None)
def nd_tet(p: int) -> NDTet:
    return NDTet(p)


def ref_rule_tables(p: int, order: int):
    """Reference tables in the reference's layout (solver/asm_data.inc:38-45):
    ip[nq*3], w[nq], vshape[nq*nd*3], curl[nq*nd*3] with index
    ((q*nd + k)*3 + d)."""
    from .quadrature import tet_rule
    el = nd_tet(p)
    pts, w = tet_rule(max(0, order))
    vs = el.vshape(pts)
    cs = el.curl(pts)
    return (np.ascontiguousarray(pts.ravel()), np.ascontiguousarray(w),
            np.ascontiguousarray(vs.ravel()), np.ascontiguousarray(cs.ravel()))
