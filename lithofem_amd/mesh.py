"""Structured conformal tetrahedral meshes of a Bloch-periodic unit cell.

The reference meshes its unit cell with Gmsh (src/lithofem/meshgen.py:365-513)
and identifies periodic vertices through a ``$Periodic`` section
(src/lithofem/meshgen.py:495-506).  Gmsh is not available here, so the
standalone path uses a deterministic structured mesher: a tensor grid in
(x, y, z) whose z breakpoints sit on every material interface and on both
PML interfaces, each hexahedron split into six tetrahedra along its main
diagonal (Kuhn split, conforming across hexahedra and across the periodic
boundary), and an optional lateral warp x' = x + a sin(2 pi x / Lx) w(z) that
keeps every z interface planar but gives each element its own affine map
(sloped "sidewalls").  Periodic x / y vertices are identified, exactly as the
vertex-identified .per.msh mesh does, while element geometry keeps the
unwrapped coordinates (the reference reads geometry from the element
transformation of the periodic mesh, solver/asm_data.inc:149-171).
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np


@dataclass
class TetMesh:
    elem_verts: np.ndarray      # (ne, 4) identified vertex ids, local order
    elem_xyz: np.ndarray        # (ne, 4, 3) unwrapped vertex coordinates
    attr: np.ndarray            # (ne,) 1-based region attribute
    nv: int                     # number of identified vertices
    vert_k: np.ndarray          # (nv,) z-layer index of each vertex
    nz: int                     # number of z cells (vertex layers 0..nz)
    bdr_layers: tuple = (0,)    # z-layer indices that carry PEC boundaries
    meta: dict = field(default_factory=dict)
    # PEC boundary triangles as identified vertex triples (meshes read from
    # files); when set they define the essential DOFs instead of bdr_layers
    bdr_faces: np.ndarray | None = None

    @property
    def ne(self) -> int:
        return len(self.attr)

    def volume(self) -> float:
        x = self.elem_xyz
        J = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
        return float(np.sum(np.linalg.det(J)) / 6.0)


# Kuhn split of the unit cube: for each axis permutation s the tetrahedron
# 000 -> e_s0 -> e_s0 + e_s1 -> 111.
def _kuhn_tets():
    tets = []
    for s in itertools.permutations(range(3)):
        c0 = (0, 0, 0)
        c1 = [0, 0, 0]
        c1[s[0]] = 1
        c2 = list(c1)
        c2[s[1]] = 1
        tets.append((c0, tuple(c1), tuple(c2), (1, 1, 1)))
    return tets


KUHN = _kuhn_tets()


def structured_cell_mesh(xs: np.ndarray, ys: np.ndarray, zs: np.ndarray,
                         region_of, periodic=(True, True),
                         warp_amp: float = 0.0, warp_z=(None, None)) -> TetMesh:
    """Kuhn-split tensor mesh.

    xs, ys, zs: grid lines (strictly increasing).  region_of(xc, yc, zc)
    maps element centroids (arrays) to 1-based attributes.  Periodic
    directions identify the last grid line with the first.
    """
    xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
    nx, ny, nz = len(xs) - 1, len(ys) - 1, len(zs) - 1
    px, py = periodic
    if (px and nx < 3) or (py and ny < 3):
        raise ValueError("periodic directions need at least 3 cells")
    mx = nx if px else nx + 1
    my = ny if py else ny + 1
    nzl = nz + 1
    lx = xs[-1] - xs[0]

    def vid(i, j, k):
        ii = i % nx if px else i
        jj = j % ny if py else j
        return (ii * my + jj) * nzl + k

    def coord(i, j, k):
        x = xs[i]
        y = ys[j]
        z = zs[k]
        if warp_amp != 0.0:
            z0, z1 = warp_z
            z0 = zs[0] if z0 is None else z0
            z1 = zs[-1] if z1 is None else z1
            wz = np.where((z > z0) & (z < z1), np.sin(np.pi * (z - z0) / (z1 - z0)), 0.0)
            x = x + warp_amp * np.sin(2 * np.pi * (x - xs[0]) / lx) * wz
        return np.stack(np.broadcast_arrays(x, y, z), axis=-1)

    I, Jg, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    I, Jg, K = I.ravel(), Jg.ravel(), K.ravel()
    ev, exyz = [], []
    for tet in KUHN:
        ids = [vid(I + c[0], Jg + c[1], K + c[2]) for c in tet]
        xyz = [coord(I + c[0], Jg + c[1], K + c[2]) for c in tet]
        ev.append(np.stack(ids, axis=1))
        exyz.append(np.stack(xyz, axis=1))
    # interleave so the six tets of one hexahedron are consecutive
    elem_verts = np.stack(ev, axis=1).reshape(-1, 4)
    elem_xyz = np.stack(exyz, axis=1).reshape(-1, 4, 3)
    # positive orientation: swap local vertices 2 and 3 where det < 0
    d = elem_xyz
    J = np.stack([d[:, 1] - d[:, 0], d[:, 2] - d[:, 0], d[:, 3] - d[:, 0]], axis=2)
    neg = np.linalg.det(J) < 0
    elem_verts[neg] = elem_verts[neg][:, [0, 1, 3, 2]]
    elem_xyz[neg] = elem_xyz[neg][:, [0, 1, 3, 2]]
    cen = elem_xyz.mean(axis=1)
    attr = np.asarray(region_of(cen[:, 0], cen[:, 1], cen[:, 2]), dtype=np.int32)
    nv = mx * my * nzl
    vert_k = np.arange(nv) % nzl
    return TetMesh(elem_verts=elem_verts.astype(np.int64), elem_xyz=elem_xyz,
                   attr=attr, nv=nv, vert_k=vert_k, nz=nz, bdr_layers=(0, nz),
                   meta=dict(nx=nx, ny=ny, nz=nz, periodic=periodic))


def graded_lines(breaks, h: float) -> np.ndarray:
    """Grid lines through every breakpoint, cells of size <= h."""
    out = [breaks[0]]
    for a, b in zip(breaks[:-1], breaks[1:]):
        n = max(1, int(np.ceil((b - a) / h - 1e-9)))
        out.extend(list(a + (b - a) * np.arange(1, n + 1) / n))
    return np.array(out)
