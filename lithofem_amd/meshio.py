"""Gmsh 2.2 ASCII meshes with a ``$Periodic`` section (the reference's
``.per.msh`` sibling, src/lithofem/meshgen.py:112-122,480-506).

The reference writes three siblings of each mesh; the one the periodic
solve reads is msh 2.2 with the lateral physical groups removed and the
node pairs (slave, master) of the periodic surfaces appended, which MFEM
turns into vertex identification.  ``read_periodic_msh`` gives the same
topology as a :class:`~lithofem_amd.mesh.TetMesh`: tetrahedra (type 4) with
their physical tag as the 1-based region attribute, vertex ids identified
through the periodic pairs (chains such as corner nodes resolved), each
element keeping its own (unwrapped) node coordinates, and the remaining
boundary triangles (type 2, the PEC z-caps) as the essential boundary.
Negatively oriented tetrahedra get local vertices 2 and 3 swapped.

``write_periodic_msh`` writes a TetMesh in the same layout (used to export
the synthetic meshes and by the round-trip tests).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from .mesh import TetMesh
from .nedelec import FACES


def _sections(text: str) -> dict:
    out = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("$") and not ln.startswith("$End"):
            name = ln[1:]
            j = i + 1
            while not lines[j].strip().startswith("$End" + name):
                j += 1
            out[name] = lines[i + 1:j]
            i = j
        i += 1
    return out


def read_periodic_msh(path) -> TetMesh:
    sec = _sections(Path(path).read_text())
    fmt = sec.get("MeshFormat", ["2.2 0 8"])[0].split()
    if not fmt or not fmt[0].startswith("2"):
        raise ValueError(f"{path}: expected msh 2.2 ASCII, got format {fmt}")
    nodes = sec["Nodes"]
    nn = int(nodes[0])
    nd = np.array([ln.split() for ln in nodes[1:1 + nn]], dtype=np.float64)
    node_id = nd[:, 0].astype(np.int64)
    xyz = nd[:, 1:4]
    row_of = {int(t): i for i, t in enumerate(node_id)}
    tets, tattr, tris = [], [], []
    elems = sec["Elements"]
    for ln in elems[1:1 + int(elems[0])]:
        v = [int(t) for t in ln.split()]
        etype, ntag = v[1], v[2]
        conn = v[3 + ntag:]
        if etype == 4:
            tets.append(conn[:4])
            tattr.append(v[3] if ntag > 0 else 1)
        elif etype == 2:
            tris.append(conn[:3])
    if not tets:
        raise ValueError(f"{path}: no tetrahedra")
    # periodic identification (union-find over node rows)
    parent = np.arange(nn)

    def find(a):
        root = a
        while parent[root] != root:
            root = parent[root]
        while parent[a] != root:
            parent[a], a = root, parent[a]
        return root

    if "Periodic" in sec:
        per = sec["Periodic"]
        nb = int(per[0])
        k = 1
        for _ in range(nb):
            k += 1                                   # "dim slave master"
            if per[k].strip().lower().startswith("affine"):
                k += 1
            npairs = int(per[k])
            k += 1
            for ln in per[k:k + npairs]:
                s, m = (row_of[int(t)] for t in ln.split()[:2])
                rs, rm = find(s), find(m)
                if rs != rm:
                    parent[max(rs, rm)] = min(rs, rm)
            k += npairs
    rep = np.array([find(a) for a in range(nn)])
    used_rows = np.array([[row_of[t] for t in tet] for tet in tets], dtype=np.int64)
    uniq, vid = np.unique(rep[used_rows], return_inverse=True)
    elem_verts = vid.reshape(-1, 4).astype(np.int64)
    elem_xyz = xyz[used_rows]
    d = elem_xyz
    J = np.stack([d[:, 1] - d[:, 0], d[:, 2] - d[:, 0], d[:, 3] - d[:, 0]], axis=2)
    neg = np.linalg.det(J) < 0
    elem_verts[neg] = elem_verts[neg][:, [0, 1, 3, 2]]
    elem_xyz[neg] = elem_xyz[neg][:, [0, 1, 3, 2]]
    bdr = None
    if tris:
        trows = np.array([[row_of[t] for t in tri] for tri in tris], dtype=np.int64)
        bdr = np.searchsorted(uniq, rep[trows]).astype(np.int64)
    nv = len(uniq)
    return TetMesh(elem_verts=elem_verts, elem_xyz=elem_xyz,
                   attr=np.asarray(tattr, dtype=np.int32), nv=nv,
                   vert_k=np.zeros(nv, dtype=np.int64), nz=0, bdr_layers=(),
                   meta=dict(source=str(path), flipped=int(neg.sum())), bdr_faces=bdr)


def write_periodic_msh(path, mesh: TetMesh) -> None:
    """Nodes = distinct (vertex, unwrapped position) pairs; periodic pairs
    map every further position of a vertex to its first one; boundary
    triangles = the element faces on the PEC layers (physical tags 1 and 2
    for the first and last layer) or the mesh's explicit boundary faces."""
    ev = mesh.elem_verts
    pos = np.round(mesh.elem_xyz.reshape(-1, 3), 9)
    key = np.concatenate([ev.reshape(-1, 1).astype(np.float64), pos], axis=1)
    uk, node_of = np.unique(key, axis=0, return_inverse=True)
    node_of = node_of.reshape(-1, 4)
    coords = uk[:, 1:]
    verts = uk[:, 0].astype(np.int64)
    first = {}
    pairs = []
    for nid, v in enumerate(verts):
        if v in first:
            pairs.append((nid + 1, first[v] + 1))
        else:
            first[int(v)] = nid
    # boundary triangles
    tri_nodes, tri_tag = [], []
    if mesh.bdr_faces is not None:
        bkeys = {tuple(sorted(map(int, f))) for f in mesh.bdr_faces}
    for e in range(mesh.ne):
        for f in FACES:
            vv = [int(ev[e, i]) for i in f]
            if mesh.bdr_faces is not None:
                hit = tuple(sorted(vv)) in bkeys
                tag = 1
            else:
                ks = [int(mesh.vert_k[v]) for v in vv]
                hit = ks[0] == ks[1] == ks[2] and ks[0] in mesh.bdr_layers
                tag = 1 if hit and ks[0] == min(mesh.bdr_layers) else 2
            if hit:
                tri_nodes.append([int(node_of[e, i]) + 1 for i in f])
                tri_tag.append(tag)
    with open(path, "w") as fh:
        fh.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        fh.write(f"$Nodes\n{len(coords)}\n")
        fh.write("".join("%d %.17g %.17g %.17g\n" % (i + 1, c[0], c[1], c[2])
                         for i, c in enumerate(coords)))
        fh.write("$EndNodes\n")
        ntot = len(tri_nodes) + mesh.ne
        fh.write(f"$Elements\n{ntot}\n")
        k = 1
        for t, tag in zip(tri_nodes, tri_tag):
            fh.write(f"{k} 2 2 {tag} {tag} {t[0]} {t[1]} {t[2]}\n")
            k += 1
        for e in range(mesh.ne):
            a = int(mesh.attr[e])
            n = node_of[e] + 1
            fh.write(f"{k} 4 2 {a} {a} {n[0]} {n[1]} {n[2]} {n[3]}\n")
            k += 1
        fh.write("$EndElements\n")
        if pairs:
            fh.write("$Periodic\n1\n2 1 1\n")
            fh.write(f"{len(pairs)}\n")
            fh.write("".join(f"{s} {m}\n" for s, m in pairs))
            fh.write("$EndPeriodic\n")
