"""Nedelec DOF numbering and per-element assembly data (AsmData equivalent).

Produces, for a :class:`~lithofem_amd.mesh.TetMesh` and order p, the flat
element-major arrays the reference extracts from MFEM in ``extract_asm_data``
(solver/asm_data.inc:77-173): decoded dof indices, +/-1 signs, face
orientations, attributes, affine geometry (J, Jinv row-major, detJ, v0) and
the two reference rules, plus the essential (PEC) DOF list of the z caps
(``GetEssentialTrueDofs`` over the boundary attributes,
solver/lithofem_solve.cpp:670-674).

Global numbering: edge DOFs, then face DOFs, then interior DOFs.  A global
edge is oriented from the lower to the higher identified vertex id; a global
face uses its sorted vertex ids as canonical vertex sequence.  An element
edge that runs against the global orientation gets sign -1 and reversed point
order (the MFEM negative-index convention decoded at
solver/asm_data.inc:135-140); an element face whose vertex sequence is the
canonical one permuted by triangle orientation o gets its points permuted
and ``fo`` chosen so that TInv[fo] maps the global DOF pair to the local one.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import TetMesh
from .nedelec import (EDGES, FACES, ORIENT, face_points_bary, interior_points_bary,
                      nd_count, ref_rule_tables)

# T[o] == TInv[fo] for fo = FO_OF_ORIENT[o] (rotations 2 and 4 are mutual
# inverses, reflections are involutions).
FO_OF_ORIENT = np.array([0, 1, 4, 3, 2, 5], dtype=np.int32)


@dataclass
class RefRule:
    nq: int
    ip: np.ndarray
    w: np.ndarray
    vshape: np.ndarray
    curl: np.ndarray


@dataclass
class AsmData:
    ne: int
    nd: int
    p: int
    height: int
    dof: np.ndarray
    sign: np.ndarray
    fo: np.ndarray
    attr: np.ndarray
    J: np.ndarray
    Jinv: np.ndarray
    detJ: np.ndarray
    v0: np.ndarray
    rule_cc: RefRule
    rule_mass: RefRule
    ess: np.ndarray                   # sorted essential DOF list
    meta: dict = field(default_factory=dict)

    @property
    def has_kpar(self) -> bool:
        return bool(self.meta.get("has_kpar", False))


def make_rule(p: int, order: int) -> RefRule:
    ip, w, vs, cs = ref_rule_tables(p, order)
    return RefRule(nq=len(w), ip=ip, w=w, vshape=vs, curl=cs)


def _canon_point_table(p: int) -> np.ndarray:
    """cp[o, pt]: canonical face-point index of local face point pt when the
    local vertex sequence is canonical[ORIENT[o]]."""
    pts = face_points_bary(p)
    index = {b: m for m, b in enumerate(pts)}
    cp = np.zeros((6, max(1, len(pts))), dtype=np.int64)
    for o, perm in enumerate(ORIENT):
        for m, lam in enumerate(pts):
            can = [0, 0, 0]
            for i in range(3):
                can[perm[i]] = lam[i]
            cp[o, m] = index[tuple(can)]
    return cp


def build_asm_data(mesh: TetMesh, p: int) -> AsmData:
    ne = mesh.ne
    nd = nd_count(p)
    ev = mesh.elem_verts
    nv = mesh.nv

    # --- edges
    le = ev[:, np.array(EDGES)]                      # (ne, 6, 2)
    lo = np.minimum(le[..., 0], le[..., 1])
    hi = np.maximum(le[..., 0], le[..., 1])
    ekey = lo * nv + hi
    ukeys, edge_id = np.unique(ekey.ravel(), return_inverse=True)
    edge_id = edge_id.reshape(ne, 6)
    n_edge = len(ukeys)
    fwd = le[..., 0] < le[..., 1]

    # --- faces
    lf = ev[:, np.array(FACES)]                      # (ne, 4, 3)
    sf = np.sort(lf, axis=2)
    fkey = (sf[..., 0] * nv + sf[..., 1]) * nv + sf[..., 2]
    ufkeys, face_id = np.unique(fkey.ravel(), return_inverse=True)
    face_id = face_id.reshape(ne, 4)
    n_face = len(ufkeys)
    rank = np.argsort(np.argsort(lf, axis=2), axis=2)   # position in sorted
    code = rank[..., 0] * 9 + rank[..., 1] * 3 + rank[..., 2]
    o_of_code = np.full(27, -1, dtype=np.int64)
    for o, perm in enumerate(ORIENT):
        o_of_code[perm[0] * 9 + perm[1] * 3 + perm[2]] = o
    orient = o_of_code[code]
    assert np.all(orient >= 0)

    nfp = p * (p - 1) // 2
    nint = p * (p - 1) * (p - 2) // 2
    e_base = 0
    f_base = n_edge * p
    i_base = f_base + n_face * 2 * nfp
    height = i_base + ne * nint

    dof = np.empty((ne, nd), dtype=np.int64)
    sign = np.ones((ne, nd), dtype=np.float64)
    m = np.arange(p)
    for le_i in range(6):
        g = np.where(fwd[:, le_i, None], m[None, :], (p - 1 - m)[None, :])
        dof[:, le_i * p:(le_i + 1) * p] = e_base + edge_id[:, le_i, None] * p + g
        sign[:, le_i * p:(le_i + 1) * p] = np.where(fwd[:, le_i, None], 1.0, -1.0)
    if p >= 2:
        cp = _canon_point_table(p)
        foff = 6 * p
        for f in range(4):
            cpt = cp[orient[:, f]]                      # (ne, nfp)
            for pt in range(nfp):
                for c in range(2):
                    k = foff + f * 2 * nfp + 2 * pt + c
                    dof[:, k] = f_base + face_id[:, f] * 2 * nfp + 2 * cpt[:, pt] + c
    if nint:
        ioff = 6 * p + 8 * nfp
        dof[:, ioff:] = i_base + np.arange(ne)[:, None] * nint + np.arange(nint)[None, :]
    fo = FO_OF_ORIENT[orient].astype(np.int32) if p >= 2 else np.zeros((0, 4), np.int32)

    # --- geometry (affine): J[i][j] = d x_i / d xi_j
    x = mesh.elem_xyz
    J = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
    detJ = np.linalg.det(J)
    if np.any(detJ <= 0):
        raise ValueError("expected positively oriented tetrahedra")
    Jinv = np.linalg.inv(J)
    v0 = x[:, 0].copy()

    # --- essential DOFs: edges and faces lying in a PEC boundary layer
    ess_mask = np.zeros(height, dtype=bool)
    if mesh.bdr_faces is not None:
        # boundary triangles given explicitly: their element faces and the
        # three edges of each
        bkeys = np.sort(np.asarray(mesh.bdr_faces, dtype=np.int64), axis=1)
        bkeys = (bkeys[:, 0] * nv + bkeys[:, 1]) * nv + bkeys[:, 2]
        on_face = np.isin(fkey, bkeys)                      # (ne, 4)
        for f in range(4):
            sel = on_face[:, f]
            if not np.any(sel):
                continue
            fv = set(FACES[f])
            for le_i, (va, vb) in enumerate(EDGES):
                if va in fv and vb in fv:
                    ess_mask[dof[sel, le_i * p:(le_i + 1) * p].ravel()] = True
            if p >= 2:
                k0 = 6 * p + f * 2 * nfp
                ess_mask[dof[sel, k0:k0 + 2 * nfp].ravel()] = True
    vk = mesh.vert_k
    for layer in (mesh.bdr_layers if mesh.bdr_faces is None else ()):
        on_e = (vk[le[..., 0]] == layer) & (vk[le[..., 1]] == layer)
        for le_i in range(6):
            sel = on_e[:, le_i]
            if np.any(sel):
                ess_mask[dof[sel, le_i * p:(le_i + 1) * p].ravel()] = True
        if p >= 2:
            on_f = np.all(vk[lf] == layer, axis=2)
            for f in range(4):
                sel = on_f[:, f]
                if np.any(sel):
                    k0 = 6 * p + f * 2 * nfp
                    ess_mask[dof[sel, k0:k0 + 2 * nfp].ravel()] = True
    ess = np.nonzero(ess_mask)[0].astype(np.int32)

    ad = AsmData(
        ne=ne, nd=nd, p=p, height=int(height),
        dof=np.ascontiguousarray(dof.astype(np.int32)),
        sign=np.ascontiguousarray(sign),
        fo=np.ascontiguousarray(fo),
        attr=np.ascontiguousarray(mesh.attr.astype(np.int32)),
        J=np.ascontiguousarray(J.reshape(ne, 9)),
        Jinv=np.ascontiguousarray(Jinv.reshape(ne, 9)),
        detJ=np.ascontiguousarray(detJ),
        v0=np.ascontiguousarray(v0),
        rule_cc=make_rule(p, max(0, 2 * p - 2)),
        rule_mass=make_rule(p, 2 * p),
        ess=ess,
        meta=dict(n_edge=int(n_edge), n_face=int(n_face), orient=orient),
    )
    return ad


def csr_symbolic(dof: np.ndarray, height: int) -> tuple[np.ndarray, np.ndarray]:
    """Sorted-column CSR pattern of the union of element dof x dof blocks
    (numpy version of solver/asm_data.inc:178-210, for small meshes)."""
    ne, nd = dof.shape
    r = np.repeat(dof, nd, axis=1).ravel().astype(np.int64)
    c = np.tile(dof, (1, nd)).ravel().astype(np.int64)
    key = np.unique(r * height + c)
    rows = key // height
    cols = (key % height).astype(np.int32)
    I = np.zeros(height + 1, dtype=np.int32)
    np.add.at(I, rows + 1, 1)
    return np.cumsum(I).astype(np.int32), cols


def interpolate(ad: AsmData, field) -> np.ndarray:
    """Global DOF vector of a vector field F(x) (callable: (n,3) points ->
    (n,3) values, real or complex): every element evaluates its local
    functionals (tangential components at its DOF points, physical edge and
    face tangents), then maps them to the global numbering with the inverse
    of the element transform (sign, face-pair TInv block).  For a field in
    the discrete space the result is exact; fields with a tangential jump
    across elements get the value of the last element written."""
    from .nedelec import ND_TINV, dof_functionals
    pts, tans = dof_functionals(ad.p)
    ne, nd, p = ad.ne, ad.nd, ad.p
    J = ad.J.reshape(ne, 3, 3)
    X = ad.v0.reshape(ne, 1, 3) + np.einsum("eij,kj->eki", J, pts)
    T = np.einsum("eij,kj->eki", J, tans)
    F = np.asarray(field(X.reshape(-1, 3))).reshape(ne, nd, 3)
    loc = np.einsum("ekd,ekd->ek", F, T)
    if p >= 2:
        ntd = p * (p - 1)
        Ti = ND_TINV.reshape(6, 4)
        # l = M g with M = [[t0, t2], [t1, t3]]  ->  g = M^{-1} l
        M = np.stack([np.stack([Ti[:, 0], Ti[:, 2]], -1), np.stack([Ti[:, 1], Ti[:, 3]], -1)], 1)
        Minv = np.linalg.inv(M)                                  # (6, 2, 2)
        for f in range(4):
            Mi = Minv[ad.fo[:, f]]                               # (ne, 2, 2)
            for i in range(ntd // 2):
                d0 = 6 * p + f * ntd + 2 * i
                pair = loc[:, d0:d0 + 2].copy()
                loc[:, d0:d0 + 2] = np.einsum("eij,ej->ei", Mi, pair)
    g = np.zeros(ad.height, dtype=loc.dtype)
    g[ad.dof.ravel()] = (loc * ad.sign).ravel()
    return g
