"""Problem description the assembly reads, and the synthetic benchmark cells.

``Problem`` mirrors the subset of the reference's ``Problem``/``Region``
(solver/problem_common.inc:18-65) that the assembly consumes: k0, k_parallel,
the physical z range, PML thickness/order/target and the region table.
``pml_params`` reproduces the per-side smax precompute of ``AsmGpu::Init``
(solver/asm_gpu.inc:302-319): smax = -ln(target) (m+1) / (2 k0 n_med thick)
with n_med = |sqrt(eps)| of the side's PML region (1 if none).

The cell builders produce deterministic structured meshes for the workloads
the benchmark and tests use (sizes are stated in each builder's docstring).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import TetMesh, graded_lines, structured_cell_mesh

WAVELENGTH = 13.5
EPS_ABS = (0.95 + 0.031j) ** 2
EPS_CAP = (0.97 + 0.01j) ** 2


@dataclass
class Region:
    eps: complex
    is_pml: bool = False
    pml_top: bool = False
    name: str = ""
    slab: int = 0               # background slab of the region
    source: bool = False        # scattered-field source (frustum) region
    eps_bg: complex | None = None   # background eps of the slab (source regions)

    @property
    def deps(self) -> complex:
        return complex(self.eps) - complex(self.eps_bg) if self.source else 0j


@dataclass
class Problem:
    k0: float
    kx: float
    ky: float
    z_min: float
    z_max: float
    pml_thick: float
    regions: list
    pml_order: float = 2.0
    pml_target: float = 1e-8
    slabs: list = field(default_factory=list)   # [(z_lo, z_hi, eps_bg)] bottom to top
    meta: dict = field(default_factory=dict)

    @property
    def has_kpar(self) -> bool:
        return self.kx != 0.0 or self.ky != 0.0

    def n_med(self, top: bool) -> float:
        n = 1.0
        for r in self.regions:
            if r.is_pml and r.pml_top == top:
                n = abs(np.sqrt(complex(r.eps)))
        return float(n)

    def smax(self, top: bool) -> float:
        return (-math.log(self.pml_target) * (self.pml_order + 1)
                / (2.0 * self.k0 * self.n_med(top) * self.pml_thick))

    def pml_params(self) -> np.ndarray:
        """[z_min, z_max, thick, order, smax_bot, smax_top, k0, kx, ky]"""
        return np.array([self.z_min, self.z_max, self.pml_thick, self.pml_order,
                         self.smax(False), self.smax(True), self.k0, self.kx, self.ky],
                        dtype=np.float64)

    def eps_tab(self) -> np.ndarray:
        e = np.array([complex(r.eps) for r in self.regions], dtype=np.complex128)
        return np.ascontiguousarray(e.view(np.float64))

    def is_pml(self) -> np.ndarray:
        return np.array([1 if r.is_pml else 0 for r in self.regions], dtype=np.int32)

    def pml_top(self) -> np.ndarray:
        return np.array([1 if r.pml_top else 0 for r in self.regions], dtype=np.int32)


def k_parallel(theta_deg: float, phi_deg: float = 0.0, wavelength: float = WAVELENGTH):
    """k0 sin(theta) (cos phi, sin phi) (src/lithofem/config.py:472-473)."""
    k0 = 2.0 * math.pi / wavelength
    kk = k0 * math.sin(math.radians(theta_deg))
    return k0, kk * math.cos(math.radians(phi_deg)), kk * math.sin(math.radians(phi_deg))


@dataclass
class Cell:
    mesh: TetMesh
    problem: Problem
    name: str


def line_grating(nx=36, ny=4, nz_scale=1.0, theta=6.0, warp=0.3) -> Cell:
    """EUV line grating: period 64 x 8 nm; capping z in [0,40] (eps_cap),
    vacuum [40,60], absorber line [60,120] (CD 32 nm at its foot, 85 deg
    sidewalls), vacuum superstrate; z-PML one wavelength thick on both
    sides (bottom PML in the capping medium).  Default nx=36, ny=4 and
    84 z cells give 72,576 tetrahedra (1.34M DOF at p=3)."""
    lx, ly, lam = 64.0, 8.0, WAVELENGTH
    zb = [-lam, 0.0, 40.0, 60.0, 120.0, 120.0 + lam]
    nseg = np.array([8, 23, 11, 34, 8], dtype=float) * nz_scale
    nseg = np.maximum(1, np.round(nseg)).astype(int)
    zs = [zb[0]]
    for a, b, n in zip(zb[:-1], zb[1:], nseg):
        zs.extend(list(a + (b - a) * np.arange(1, n + 1) / n))
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    slope = math.tan(math.radians(5.0))

    def region_of(x, y, z):
        a = np.full(x.shape, 5, dtype=np.int32)
        a[z < 0] = 1
        a[(z >= 0) & (z < 40)] = 2
        a[(z >= 40) & (z < 60)] = 3
        half = 16.0 - (z - 60.0) * slope
        a[(z >= 60) & (z < 120) & (np.abs(x - 32.0) < half)] = 4
        a[z >= 120] = 6
        return a

    mesh = structured_cell_mesh(xs, ys, np.array(zs), region_of, periodic=(True, True),
                                warp_amp=warp, warp_z=(0.0, 120.0))
    k0, kx, ky = k_parallel(theta)
    regions = [Region(EPS_CAP, True, False, "pml_bottom", slab=0),
               Region(EPS_CAP, name="capping", slab=0),
               Region(1.0, name="vacuum", slab=1),
               Region(EPS_ABS, name="absorber", slab=2, source=True, eps_bg=1.0),
               Region(1.0, name="vacuum_bg", slab=2),
               Region(1.0, True, True, "pml_top", slab=2)]
    slabs = [(0.0, 40.0, EPS_CAP), (40.0, 60.0, 1.0), (60.0, 120.0, 1.0)]
    pb = Problem(k0=k0, kx=kx, ky=ky, z_min=0.0, z_max=120.0, pml_thick=lam, regions=regions,
                 slabs=slabs)
    return Cell(mesh, pb, "line_grating")


def contact_hole(n_xy=12, nz_scale=1.0, side=48.0, theta=6.0, phi=0.0, warp=0.4,
                 lx=96.0, z_abs=60.0, sidewall_deg=90.0) -> Cell:
    """Square contact hole through an absorber slab: cell lx x lx nm, absorber
    z in [0, z_abs] with an eps=1 hole of the given side centred in the
    cell (sidewall angle tapers the hole), vacuum above to z = 120, z-PML one
    wavelength on both sides (bottom PML in the absorber medium).  Default
    12 x 12 x 28 hexahedra = 24,192 tetrahedra (~1.5e5 DOF at p=2)."""
    lam = WAVELENGTH
    zb = [-lam, 0.0, z_abs, 120.0, 120.0 + lam]
    nseg = np.array([3, 12, 10, 3], dtype=float) * nz_scale
    nseg = np.maximum(1, np.round(nseg)).astype(int)
    zs = [zb[0]]
    for a, b, n in zip(zb[:-1], zb[1:], nseg):
        zs.extend(list(a + (b - a) * np.arange(1, n + 1) / n))
    xs = np.linspace(0.0, lx, n_xy + 1)
    c = lx / 2
    slope = 1.0 / math.tan(math.radians(sidewall_deg)) if sidewall_deg < 90 else 0.0

    def region_of(x, y, z):
        a = np.full(x.shape, 4, dtype=np.int32)
        a[z < 0] = 1
        inabs = (z >= 0) & (z < z_abs)
        a[inabs] = 2
        half = side / 2 + (z - z_abs / 2) * slope
        hole = inabs & (np.abs(x - c) < half) & (np.abs(y - c) < half)
        a[hole] = 3
        a[z >= 120] = 5
        return a

    mesh = structured_cell_mesh(xs, xs.copy(), np.array(zs), region_of, periodic=(True, True),
                                warp_amp=warp, warp_z=(0.0, 120.0))
    k0, kx, ky = k_parallel(theta, phi)
    regions = [Region(EPS_ABS, True, False, "pml_bottom", slab=0),
               Region(EPS_ABS, name="absorber", slab=0),
               Region(1.0, name="hole", slab=0, source=True, eps_bg=EPS_ABS),
               Region(1.0, name="vacuum", slab=1),
               Region(1.0, True, True, "pml_top", slab=1)]
    slabs = [(0.0, z_abs, EPS_ABS), (z_abs, 120.0, 1.0)]
    pb = Problem(k0=k0, kx=kx, ky=ky, z_min=0.0, z_max=120.0, pml_thick=lam, regions=regions,
                 slabs=slabs)
    return Cell(mesh, pb, "contact_hole")


def l_opening(n_xy=16, nz_scale=1.0, theta=6.0, warp=0.5) -> Cell:
    """L-shaped opening in a 70 nm absorber (80 deg sidewalls), 192 x 192 nm
    cell, arms 40 nm wide and 120 nm long, vacuum to z = 130, z-PML on both
    sides.  Default 16 x 16 x 48 hexahedra = 73,728 tetrahedra."""
    lam = WAVELENGTH
    lx = 192.0
    zb = [-lam, 0.0, 70.0, 130.0, 130.0 + lam]
    nseg = np.array([5, 25, 13, 5], dtype=float) * nz_scale
    nseg = np.maximum(1, np.round(nseg)).astype(int)
    zs = [zb[0]]
    for a, b, n in zip(zb[:-1], zb[1:], nseg):
        zs.extend(list(a + (b - a) * np.arange(1, n + 1) / n))
    xs = np.linspace(0.0, lx, n_xy + 1)
    slope = 1.0 / math.tan(math.radians(80.0))

    def region_of(x, y, z):
        a = np.full(x.shape, 4, dtype=np.int32)
        a[z < 0] = 1
        inabs = (z >= 0) & (z < 70.0)
        a[inabs] = 2
        grow = (z - 35.0) * slope
        arm1 = (x > 36 - grow) & (x < 76 + grow) & (y > 36 - grow) & (y < 156 + grow)
        arm2 = (x > 36 - grow) & (x < 156 + grow) & (y > 36 - grow) & (y < 76 + grow)
        a[inabs & (arm1 | arm2)] = 3
        a[z >= 130] = 5
        return a

    mesh = structured_cell_mesh(xs, xs.copy(), np.array(zs), region_of, periodic=(True, True),
                                warp_amp=warp, warp_z=(0.0, 130.0))
    k0, kx, ky = k_parallel(theta)
    regions = [Region(EPS_ABS, True, False, "pml_bottom", slab=0),
               Region(EPS_ABS, name="absorber", slab=0),
               Region(1.0, name="opening", slab=0, source=True, eps_bg=EPS_ABS),
               Region(1.0, name="vacuum", slab=1),
               Region(1.0, True, True, "pml_top", slab=1)]
    slabs = [(0.0, 70.0, EPS_ABS), (70.0, 130.0, 1.0)]
    pb = Problem(k0=k0, kx=kx, ky=ky, z_min=0.0, z_max=130.0, pml_thick=lam, regions=regions,
                 slabs=slabs)
    return Cell(mesh, pb, "l_opening")
