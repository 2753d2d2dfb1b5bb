"""Benchmark workloads and the canonical work/traffic model of the assembly.

``canonical_flops`` is the algorithmic FP64 work per element of the
quadrature-form contraction (B^T C GEMMs re/im, forming C = w D B from D's
structural nonzeros, Piola maps, face transform):

    W(e) = sum_{r in {cc, mass}} nq_r [4 d_r nd^2 + 4 z_r nd + 6 d_r nd] + 48 p (p-1) nd
    d_cc = 3, z_cc = 3;  d_mass = 6, z_mass = 13 with k_par != 0, else 3 and 3.

``scatter_bytes`` is the compulsory HBM traffic of the materialised-Ae
scatter: read Ae (16 B) and its slot (1 B when every CSR row has at most
256 entries, else 2 B) per local entry, write each CSR value once (16 B),
plus the row pointers, contributor lists and signs.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .fespace import build_asm_data
from .problem import line_grating


def canonical_flops(p: int, nd: int, nq_cc: int, nq_m: int, kpar: bool) -> float:
    d_cc, z_cc = 3, 3
    d_m, z_m = (6, 13) if kpar else (3, 3)
    w = nq_cc * (4 * d_cc * nd * nd + 4 * z_cc * nd + 6 * d_cc * nd)
    w += nq_m * (4 * d_m * nd * nd + 4 * z_m * nd + 6 * d_m * nd)
    w += 48 * p * (p - 1) * nd
    return float(w)


def local_bytes(ne: int, nd: int) -> float:
    """Compulsory HBM bytes of the local-matrix kernel: per-element inputs
    (J, Jinv, detJ, v0, attr, fo) and the Ae write."""
    return float(ne) * (9 * 8 * 2 + 8 + 24 + 4 + 16 + 16.0 * nd * nd)


def tensor_bytes(ne: int, nd: int) -> float:
    """Compulsory HBM bytes of the reference-tensor kernel per launch: per
    element the list entry, J, Jinv, detJ, attr, fo and the Ae write (the
    reference tensors and region table stay in cache)."""
    return float(ne) * (4 + 9 * 8 * 2 + 8 + 4 + 16 + 16.0 * nd * nd)


def scatter_bytes(ne: int, nd: int, n: int, nnz: int, slot_bytes: int = 2) -> float:
    return float(ne) * nd * nd * (16 + slot_bytes) + 16.0 * nnz + 4.0 * (n + 1) + \
        float(ne) * nd * (4 + 4 + 8) + 4.0 * (n + 1)


@dataclass
class Workload:
    name: str
    cell: object
    ad: object
    p: int


def line_grating_p3(nx: int = 36, ny: int = 4, nz_scale: float = 1.0) -> Workload:
    """EUV line grating, cubic Nedelec, theta = 6 deg, p-polarised:
    72,576 tetrahedra, ~1.34M complex DOF at the default size."""
    cell = line_grating(nx=nx, ny=ny, nz_scale=nz_scale, theta=6.0)
    ad = build_asm_data(cell.mesh, 3)
    return Workload("euv_line_grating_p3", cell, ad, 3)
