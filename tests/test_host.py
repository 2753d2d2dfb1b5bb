"""CPU tests of the C-ABI library's host side and of the input producers.

* liblfasm.so loads and exports every function include/lithofem_asm.h
  declares (no device call is made);
* the host-only entry points (lfasm_csr_symbolic, lfasm_pml_smax) against
  independent computations;
* the integration rules integrate every monomial of their order exactly;
* the Nedelec space built by lithofem_amd.fespace is H(curl)-conforming:
  a random global DOF vector has the same tangential field on both sides
  of every shared face, periodic faces included;
* the layered-stack incident field satisfies the interface conditions
  (tangential E and H, normal D) and the dispersion relation;
* bench.py refuses more than one GPU.
"""

from __future__ import annotations

import dataclasses
import json
import math
import re
import subprocess
import sys
from math import factorial
from pathlib import Path

import numpy as np
import pytest

from lithofem_amd import gpu
from lithofem_amd.fespace import build_asm_data, csr_symbolic
from lithofem_amd.incident import slab_waves
from lithofem_amd.nedelec import FACES, ND_TINV, nd_tet
from lithofem_amd.problem import contact_hole, line_grating
from lithofem_amd.quadrature import monomial_exponents, tet_rule

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(lfasm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not gpu.LIB_PATH.exists():
        pytest.fail("liblfasm.so is not built (run __graft_entry__.build())")
    return gpu.lib()


def test_library_exports_every_declared_function(lib):
    names = _declared_functions()
    assert len(names) >= 30, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("p", [1, 2, 3])
def test_library_csr_symbolic_equals_numpy(lib, p):
    c = contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=0.0)
    ad = build_asm_data(c.mesh, p)
    I0, J0 = csr_symbolic(ad.dof, ad.height)
    I1, J1 = gpu.csr_symbolic(ad.dof, ad.height)
    assert np.array_equal(I0, I1) and np.array_equal(J0, J1)
    # brute force: (i, j) is in the pattern iff some element holds both
    rows = np.repeat(ad.dof, ad.nd, axis=1).ravel()
    cols = np.tile(ad.dof, (1, ad.nd)).ravel()
    keys = np.unique(rows.astype(np.int64) * ad.height + cols)
    assert len(keys) == len(J1)
    got = np.repeat(np.arange(ad.height), np.diff(I1)).astype(np.int64) * ad.height + J1
    assert np.array_equal(np.sort(got), keys)


def test_library_rejects_repeated_dof_in_an_element(lib):
    """lfasm_init validates the mesh on the host before it needs the device:
    an element holding one global dof at two local positions is refused
    (the row scatter would lose one of its contributions), a valid mesh
    passes validation and then reports the missing device."""
    import ctypes as ct
    c = contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=0.0)
    ad = build_asm_data(c.mesh, 2)
    h = ct.c_void_p()
    lib.lfasm_create(0, ct.byref(h))          # no device here: the handle holds the error
    try:
        keep = []
        m = gpu.AsmGpuGlobal._mesh_struct(None, ad, c.problem, None, keep)
        assert lib.lfasm_init(h, ct.byref(m)) == gpu.E_STATE
        bad = ad.dof.copy()
        bad[5, 7] = bad[5, 2]
        ad_bad = dataclasses.replace(ad, dof=bad)
        m = gpu.AsmGpuGlobal._mesh_struct(None, ad_bad, c.problem, None, keep)
        assert lib.lfasm_init(h, ct.byref(m)) == gpu.E_ARG
        msg = lib.lfasm_last_error(h).decode()
        assert "element 5 holds global dof" in msg, msg
    finally:
        lib.lfasm_destroy(h)


def test_library_residual_static_limit(lib):
    """lfasm_residual_static checks the vector length against its static
    workspace before it needs the device; the limit is the header's."""
    import ctypes as ct
    hdr = (ROOT / "include" / "lithofem_asm.h").read_text()
    m = re.search(r"#define LFASM_RESIDUAL_STATIC_MAX_N (\d+)", hdr)
    assert m and int(m.group(1)) == gpu.RESIDUAL_STATIC_MAX_N
    h = ct.c_void_p()
    lib.lfasm_create(0, ct.byref(h))          # no device here
    try:
        x = np.zeros(4)
        rel = ct.c_double()
        big = 2 * (gpu.RESIDUAL_STATIC_MAX_N + 1)
        assert lib.lfasm_residual_static(h, big, x.ctypes.data, x.ctypes.data,
                                         ct.byref(rel)) == gpu.E_ARG
        assert "exceeds the static workspace" in lib.lfasm_last_error(h).decode()
        assert lib.lfasm_residual_static(h, 3, x.ctypes.data, x.ctypes.data,
                                         ct.byref(rel)) == gpu.E_ARG
        assert lib.lfasm_residual_static(h, 4, x.ctypes.data, x.ctypes.data,
                                         ct.byref(rel)) == gpu.E_STATE
    finally:
        lib.lfasm_destroy(h)


def test_library_pml_smax(lib):
    import ctypes as ct
    f = lib.lfasm_pml_smax
    f.restype = ct.c_double
    f.argtypes = [ct.c_double] * 5
    # -ln(1e-8) * 3 / (2 * 0.5 * 1.2 * 10) = 18.42068... * 3 / 12
    assert abs(f(1e-8, 2.0, 0.5, 1.2, 10.0) - 8.0 * math.log(10.0) * 3.0 / 12.0) < 1e-14
    pb = line_grating(nx=4, ny=3, nz_scale=0.15).problem
    for top in (False, True):
        ref = pb.smax(top)
        got = f(pb.pml_target, pb.pml_order, pb.k0, pb.n_med(top), pb.pml_thick)
        assert abs(got - ref) <= 1e-15 * abs(ref)


def test_tet_rule_point_counts():
    """The symmetric rules have the point counts of MFEM's tetrahedron rules
    for the orders the assembly asks for (2p - 2 and 2p, p = 1..4)."""
    assert [len(tet_rule(o)[1]) for o in (0, 2, 4, 6, 8)] == [1, 4, 11, 24, 45]
    # Keast's degree-8 rule carries one negative (centroid) weight
    w8 = tet_rule(8)[1]
    assert np.sum(w8 < 0) == 1 and abs(w8.sum() - 1.0 / 6.0) < 1e-16


@pytest.mark.parametrize("order", range(0, 9))
def test_tet_rule_is_exact_to_its_order(order):
    x, w = tet_rule(order)
    assert np.all(x >= -1e-14) and np.all(x.sum(1) <= 1 + 1e-14)
    for (i, j, k) in monomial_exponents(order):
        exact = factorial(i) * factorial(j) * factorial(k) / factorial(i + j + k + 3)
        got = float(np.sum(w * x[:, 0] ** i * x[:, 1] ** j * x[:, 2] ** k))
        assert abs(got - exact) < 1e-14, (order, i, j, k, got, exact)


def _field(ad, el, e, xi, g):
    """Physical field of global DOF vector g inside element e at reference
    point xi (the element transform: sign, then the face-pair block)."""
    vs = el.vshape(xi[None, :])[0]                       # (nd, 3)
    v = vs @ ad.Jinv.reshape(-1, 3, 3)[e]                # covariant map
    loc = g[ad.dof[e]] * ad.sign[e]
    if ad.p >= 2:
        ntd = ad.p * (ad.p - 1)
        Ti = ND_TINV.reshape(6, 4)
        for f in range(4):
            t = Ti[ad.fo[e, f]]
            for i in range(ntd // 2):
                d0 = 6 * ad.p + f * ntd + 2 * i
                a, b = loc[d0], loc[d0 + 1]
                loc[d0], loc[d0 + 1] = t[0] * a + t[2] * b, t[1] * a + t[3] * b
    return loc @ v


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("name", ["grating", "hole"])
def test_nedelec_space_is_tangentially_conforming(name, p):
    c = line_grating(nx=4, ny=3, nz_scale=0.15, theta=6.0) if name == "grating" else \
        contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=6.0)
    mesh = c.mesh
    ad = build_asm_data(mesh, p)
    el = nd_tet(p)
    rng = np.random.default_rng(11 + p)
    g = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
    owners = {}
    for e in range(ad.ne):
        for f, lv in enumerate(FACES):
            key = tuple(sorted(int(mesh.elem_verts[e, v]) for v in lv))
            owners.setdefault(key, []).append((e, lv))
    shared = [v for v in owners.values() if len(v) == 2]
    assert len(shared) > 100

    def centroid(e, lv):
        return mesh.elem_xyz[e][list(lv), :].mean(0)

    periodic = [np.linalg.norm(centroid(*a) - centroid(*b)) > 1e-9 for a, b in shared]
    per_idx = np.nonzero(periodic)[0]
    int_idx = np.nonzero(~np.array(periodic))[0]
    assert len(per_idx) > 0
    picks = np.concatenate([rng.choice(per_idx, min(60, len(per_idx)), replace=False),
                            rng.choice(int_idx, 60, replace=False)])
    worst_t, worst_n = 0.0, 0.0
    for s in picks:
        (e1, lv1), (e2, lv2) = shared[s]
        ids = sorted(int(mesh.elem_verts[e1, v]) for v in lv1)
        wts = rng.dirichlet(np.ones(3))
        fields = []
        for e, lv in ((e1, lv1), (e2, lv2)):
            lam = np.zeros(4)
            for v in lv:
                lam[v] = wts[ids.index(int(mesh.elem_verts[e, v]))]
            fields.append(_field(ad, el, e, lam[1:], g))
        xa, xb, xc = (mesh.elem_xyz[e1, v] for v in lv1)
        normal = np.cross(xb - xa, xc - xa)
        normal /= np.linalg.norm(normal)
        d = fields[0] - fields[1]
        ref = np.abs(fields[0]).max()
        worst_t = max(worst_t, np.abs(d - (d @ normal) * normal).max() / ref)
        worst_n = max(worst_n, abs(d @ normal) / ref)
    assert worst_t < 1e-11, worst_t
    # not the same element twice: the normal component does jump
    assert worst_n > 1e-3, worst_n


@pytest.mark.parametrize("phi", [0.0, 30.0])
@pytest.mark.parametrize("jones", ["s", "p"])
def test_incident_field_interface_conditions(phi, jones):
    c = contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=9.0, phi=phi)
    pb = c.problem
    W = slab_waves(pb, 9.0, phi, jones)
    slabs = pb.slabs
    eps = [complex(s[2]) for s in slabs]
    k0 = pb.k0
    kpar = k0 * np.sqrt(eps[-1]).real * math.sin(math.radians(9.0))
    kx, ky = kpar * math.cos(math.radians(phi)), kpar * math.sin(math.radians(phi))

    def fields(j, z):
        w = W[j]
        qA, qB = w[2] + 1j * w[3], w[4] + 1j * w[5]
        A = w[6:12:2] + 1j * w[7:12:2]
        B = w[12:18:2] + 1j * w[13:18:2]
        ea, eb = np.exp(1j * qA * (z - w[0])), np.exp(1j * qB * (z - w[1]))
        E = ea * A + eb * B
        dE = 1j * qA * ea * A + 1j * qB * eb * B
        # curl of E exp(i k_par . r), envelope part
        H = np.array([1j * ky * E[2] - dE[1], dE[0] - 1j * kx * E[2],
                      1j * kx * E[1] - 1j * ky * E[0]])
        div = 1j * kx * E[0] + 1j * ky * E[1] + dE[2]
        return E, H, div, qA, qB, A, B

    scale = np.abs(fields(len(slabs) - 1, slabs[-1][0])[0]).max()
    for j in range(len(slabs)):
        z = 0.5 * (slabs[j][0] + slabs[j][1])
        E, H, div, qA, qB, A, B = fields(j, z)
        assert abs(div) < 1e-12 * k0 * scale
        for q in (qA, qB):
            assert abs(q * q + kx * kx + ky * ky - k0 * k0 * eps[j]) < 1e-12 * k0 * k0
    for j in range(1, len(slabs)):
        z = slabs[j][0]
        Eb, Hb, *_ = fields(j - 1, z)
        Et, Ht, *_ = fields(j, z)
        assert np.abs(Eb[:2] - Et[:2]).max() < 1e-12 * scale
        assert abs(eps[j - 1] * Eb[2] - eps[j] * Et[2]) < 1e-12 * scale * abs(eps[j])
        assert np.abs(Hb[:2] - Ht[:2]).max() < 1e-12 * k0 * scale
    # nothing comes in from below: the bottom slab holds only the downward wave
    assert np.abs(W[0][6:12]).max() < 1e-14
    # unit incident amplitude from the top (|E_inc| = 1 for s, |H_inc| ~ n for p)
    E_in = W[-1][12:18:2] + 1j * W[-1][13:18:2]
    assert abs(np.linalg.norm(E_in) - 1.0) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3])
def test_sampling_matrix_reproduces_fields_in_the_space(p):
    """S @ g equals the field at the points for a field of the discrete
    space: on the (non-periodic) warped unit cube every polynomial of
    degree p - 1 is in it (a random complex linear field for p >= 2); on
    the periodic grating cell a constant field, with points outside the
    cell wrapped back by `periodic`."""
    sys.path.insert(0, str(ROOT / "tests"))
    import mms
    from lithofem_amd.adjoint import sampling_matrix
    from lithofem_amd.fespace import interpolate
    rng = np.random.default_rng(50 + p)
    a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    B = (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))) if p >= 2 else 0 * np.eye(3)

    def F(X):
        return a[None, :] + X @ np.asarray(B).T

    ad = build_asm_data(mms.cube_mesh(3), p)
    pts = rng.uniform(0.02, 0.98, (200, 3))
    S = sampling_matrix(ad, pts)
    got = (S @ interpolate(ad, F)).reshape(-1, 3)
    assert np.abs(got - F(pts)).max() < 1e-12 * np.abs(F(pts)).max()
    c = line_grating(nx=3, ny=3, nz_scale=0.1, theta=6.0)
    adg = build_asm_data(c.mesh, p)
    pts = np.stack([rng.uniform(-64, 128, 100), rng.uniform(-8, 16, 100),
                    rng.uniform(-10, 130, 100)], 1)
    S = sampling_matrix(adg, pts, periodic=(64.0, 8.0))
    got = (S @ interpolate(adg, lambda X: np.tile(a, (len(X), 1)))).reshape(-1, 3)
    assert np.abs(got - a[None, :]).max() < 1e-12 * np.abs(a).max()


def test_order_efficiencies_single_plane_wave():
    """A plane wave of order (m, n) sampled on the cell grid has one Fourier
    coefficient: its efficiency is |a|^2 kz_mn / kz_inc, every other order
    is 0; a sum of two orders gives both (the per-order flux of the
    reference's orders_from_plane / one_way_flux_z,
    src/lithofem/orders.py:36-92)."""
    import torch
    from lithofem_amd.adjoint import order_efficiencies, plane_points
    k0 = 2 * math.pi / 13.5
    lx, ly, nx, ny = 64.0, 8.0, 16, 4
    kpar = (k0 * math.sin(math.radians(6.0)), 0.0)
    pts = plane_points(lx, ly, 130.0, nx, ny, nudge=(0.0, 0.0, 0.0))
    x, y = pts[:, 0].reshape(nx, ny), pts[:, 1].reshape(nx, ny)
    a1 = np.array([0.2 + 0.1j, -0.05j, 0.3])
    a2 = np.array([0.1, 0.4 - 0.2j, -0.1j])

    def wave(m, n, amp):
        ph = np.exp(2j * math.pi * (m * x / lx + n * y / ly))
        return ph[..., None] * amp[None, None, :]

    E = torch.as_tensor(wave(1, 0, a1) + wave(-1, 0, a2))
    eff = order_efficiencies(E, kpar, k0, lx, ly, [(1, 0), (-1, 0), (0, 0), (2, 0)])

    def kz(m):
        kx = kpar[0] + 2 * math.pi * m / lx
        return math.sqrt(k0 * k0 - kx * kx)

    want = [np.sum(np.abs(a1) ** 2) * kz(1) / kz(0), np.sum(np.abs(a2) ** 2) * kz(-1) / kz(0),
            0.0, 0.0]
    for g, w in zip(eff, want):
        assert abs(float(g) - w) < 1e-14


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_bench_reads_the_committed_traffic_counters():
    """roofline.traffic of the default run comes from the counter summary
    under profiles/; its kernel names must be the ones bench.py reports."""
    sys.path.insert(0, str(ROOT))
    import bench
    t = bench._measured_traffic()
    assert set(t) == {"k_scatter_rows", "k_tensor_coef+k_local_tensor21"}
    # HBM bytes per launch near the algorithmic bytes (4.55 GB, 1.91 GB)
    assert 4.5e9 < t["k_scatter_rows"] < 6.0e9
    assert 1.9e9 < t["k_tensor_coef+k_local_tensor21"] < 2.5e9


def test_matrixmarket_round_trip(tmp_path):
    """The oracle's eliminated system and an RHS written in the reference's
    MatrixMarket format read back bit for bit (17 significant digits)."""
    from lithofem_amd import mmio
    from oracle import oracle as O
    c = contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=6.0)
    ad = build_asm_data(c.mesh, 2)
    I, J = csr_symbolic(ad.dof, ad.height)
    orc = O.Oracle(ad, c.problem, I, J)
    vre, vim = orc.assemble()
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    vre, vim = O.eliminate(I, J, is_ess, vre, vim)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(ad.height) * 1e-7 + 1j * rng.standard_normal(ad.height)
    prefix = str(tmp_path / "sys")
    mmio.export_system(prefix, I, J, vre, vim, b)
    head = open(prefix + ".mtx").read(200).splitlines()
    assert head[0] == "%%MatrixMarket matrix coordinate complex general"
    assert head[1] == f"{ad.height} {ad.height} {len(J)}"
    I2, J2, r2, i2 = mmio.read_system(prefix + ".mtx")
    assert np.array_equal(I2, I) and np.array_equal(J2, J)
    assert np.array_equal(r2, vre) and np.array_equal(i2, vim)
    assert np.array_equal(mmio.read_vector(prefix + ".rhs.mtx"), b)


def _stack(slabs, wavelength=13.5):
    from types import SimpleNamespace
    return SimpleNamespace(k0=2 * math.pi / wavelength, slabs=slabs)


def _fields(W, j, z):
    from lithofem_amd.incident import evaluate
    return evaluate(W, j, np.atleast_1d(z))[0]


@pytest.mark.parametrize("pol", ["s", "p"])
def test_incident_uniform_background_is_bare_wave(pol):
    """Vacuum everywhere: no reflected wave, |E| = 1, phase advance
    exp(-i qz dz) (the reference's tests/test_incident.py:78-97)."""
    pb = _stack([(0.0, 60.0, 1.0), (60.0, 120.0, 1.0)])
    W = slab_waves(pb, 20.0, 30.0, pol)
    k0 = pb.k0
    kp = k0 * math.sin(math.radians(20.0))
    qz = math.sqrt(k0 * k0 - kp * kp)
    z = np.linspace(1.0, 119.0, 57)
    e = np.array([_fields(W, 0 if zz < 60 else 1, zz) for zz in z])
    assert np.max(np.abs(np.linalg.norm(e, axis=1) - 1.0)) < 1e-12
    ok = np.abs(e[:-1]) > 1e-8
    ratio = np.where(ok, e[1:], 0.0) / np.where(ok, e[:-1], 1.0)
    assert np.max(np.abs(ratio[ok] - np.exp(-1j * qz * (z[1] - z[0])))) < 1e-10
    assert np.abs(W[1][6:12]).max() < 1e-14          # no upward wave on top


def test_incident_polarisation_conventions():
    """s at normal incidence: E = +y, amplitude 1 at z_max; oblique p: E at
    z_max = e_p = k_hat x e_s (tests/test_incident.py:117-140)."""
    pb = _stack([(0.0, 60.0, 1.0), (60.0, 120.0, 1.0)])
    e = _fields(slab_waves(pb, 0.0, 0.0, "s"), 1, 120.0)
    assert np.allclose(e, [0.0, 1.0, 0.0], atol=1e-13)
    th, ph = math.radians(30.0), math.radians(40.0)
    k_hat = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), -math.cos(th)])
    e_s = np.cross([0.0, 0.0, 1.0], k_hat)
    e_s /= np.linalg.norm(e_s)
    e_p = np.cross(k_hat, e_s)
    e = _fields(slab_waves(pb, 30.0, 40.0, "p"), 1, 120.0)
    assert np.linalg.norm(e - e_p) < 1e-12
    e = _fields(slab_waves(pb, 30.0, 40.0, "s"), 1, 120.0)
    assert np.linalg.norm(e - e_s) < 1e-12


def test_jones_vector_decomposition_closed_form():
    """A Jones vector is the transverse (Ex, Ey) of the incident field,
    decomposed onto (e_s, e_p) (src/lithofem/incident.py:220-225).  At
    phi = 0, e_s = y and e_p = (cos t, 0, sin t), so c_s = Ey and
    c_p = Ex / cos t: Jones (1, 0) is pure p, (0, 1) pure s, and
    (1, i)/sqrt 2 gives c_s = i/sqrt 2, c_p = 1/(sqrt 2 cos t)."""
    from lithofem_amd.incident import sp_coefficients
    r2 = 1.0 / math.sqrt(2.0)
    ct6 = math.cos(math.radians(6.0))
    cs, cp = sp_coefficients(6.0, 0.0, (r2, 1j * r2))
    assert abs(cs - 1j * r2) < 1e-15 and abs(cp - r2 / ct6) < 1e-15
    cs, cp = sp_coefficients(6.0, 0.0, (1.0, 0.0))
    assert abs(cs) < 1e-16 and abs(cp - 1.0 / ct6) < 1e-15
    cs, cp = sp_coefficients(6.0, 0.0, (0.0, 1.0))
    assert abs(cs - 1.0) < 1e-16 and abs(cp) < 1e-16
    assert sp_coefficients(6.0, 0.0, "s") == (1.0, 0.0)
    assert sp_coefficients(6.0, 0.0, "p") == (0.0, 1.0)


@pytest.mark.parametrize("theta,phi", [(6.0, 0.0), (14.0, 35.0), (0.0, 30.0)])
def test_jones_field_at_the_top(theta, phi):
    """In a uniform vacuum stack the field at z_max is the incident wave
    alone: its transverse part is amplitude x (Ex, Ey), and it is
    transverse to k_hat; at normal incidence s is along +y whatever phi
    (the reference's basis, src/lithofem/incident.py:82-97)."""
    from lithofem_amd.incident import sp_basis
    pb = _stack([(0.0, 60.0, 1.0), (60.0, 120.0, 1.0)])
    jones = (0.3 - 0.2j, 0.8 + 0.1j)
    amp = 0.7 * np.exp(0.4j)
    e = _fields(slab_waves(pb, theta, phi, jones, amp), 1, 120.0)
    assert np.abs(e[:2] - amp * np.array(jones)).max() < 1e-13
    _, _, k_hat = sp_basis(theta, phi)
    assert abs(e @ k_hat) < 1e-13
    if theta == 0.0:
        assert np.allclose(_fields(slab_waves(pb, 0.0, phi, "s"), 1, 120.0), [0, 1, 0],
                           atol=1e-14)
        assert np.allclose(_fields(slab_waves(pb, 0.0, phi, "p"), 1, 120.0), [1, 0, 0],
                           atol=1e-14)


def test_incident_tables_reference_layout():
    """groups[].incident of the reference's solve.json: one part per pure
    s / p component of a source, summed per source on reading.  A table
    written part by part in that layout reads back as the field of the
    Jones source, and a hand-made single-slab plane wave evaluates to its
    closed form."""
    from lithofem_amd.incident import (evaluate, sp_coefficients, tables_from_solve_json,
                                       tables_to_solve_json)
    c = contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=6.0, phi=20.0)
    pb = c.problem
    bounds = [pb.slabs[0][0]] + [sl[1] for sl in pb.slabs]
    jones = (1 / math.sqrt(2), 1j / math.sqrt(2))
    cs, cp = sp_coefficients(6.0, 20.0, jones)
    Ws, Wp = slab_waves(pb, 6.0, 20.0, "s", cs), slab_waves(pb, 6.0, 20.0, "p", cp)
    parts = (tables_to_solve_json({3: Ws}, bounds) + tables_to_solve_json({3: Wp}, bounds)
             + tables_to_solve_json({5: Ws}, bounds))
    parts = json.loads(json.dumps(parts))
    got = tables_from_solve_json(parts, len(pb.slabs))
    assert sorted(got) == [3, 5]
    want = slab_waves(pb, 6.0, 20.0, jones)
    z = np.linspace(0.5, 119.5, 41)
    for j, (lo, hi, _) in enumerate(pb.slabs):
        zz = z[(z >= lo) & (z < hi)]
        a, b = evaluate(got[3], j, zz), evaluate(want, j, zz)
        assert np.abs(a - b).max() < 1e-14 * np.abs(b).max()
    assert np.array_equal(got[5], Ws)
    # one vacuum slab, E = e_s exp(-i kz (z - 120)) written by hand
    kz = 0.3
    hand = [{"source_index": 0, "slabs": [{"z": [0.0, 120.0], "qA": [-kz, 0.0], "zA": 120.0,
                                           "A": [[0, 0], [1, 0], [0, 0]], "qB": [kz, 0.0],
                                           "zB": 0.0, "B": [[0, 0], [0, 0], [0, 0]]}]}]
    W = tables_from_solve_json(hand, 1)[0]
    e = evaluate(W, 0, z)
    assert np.abs(e[:, 1] - np.exp(-1j * kz * (z - 120.0))).max() < 1e-15
    assert np.abs(e[:, [0, 2]]).max() == 0.0
    # parts of one source with different propagation constants are refused
    bad = json.loads(json.dumps(hand))
    bad[0]["slabs"][0]["qA"] = [-0.31, 0.0]
    with pytest.raises(NotImplementedError):
        tables_from_solve_json(hand + bad, 1)


@pytest.mark.parametrize("pol", ["s", "p"])
@pytest.mark.parametrize("eps2", [2.13, (0.95 + 0.031j) ** 2])
def test_incident_fresnel_and_energy(pol, eps2):
    """Vacuum over a half space: reflection = the Fresnel coefficient, and
    the reflected plus transmitted flux at the interface equals the
    incident flux (R + T = 1, lossless or absorbing substrate)."""
    pb = _stack([(0.0, 60.0, eps2), (60.0, 120.0, 1.0)])
    theta = 15.0
    W = slab_waves(pb, theta, 0.0, pol)
    k0 = pb.k0
    kp = k0 * math.sin(math.radians(theta))
    q1 = np.sqrt(complex(k0 * k0 - kp * kp))
    q2 = np.sqrt(complex(k0 * k0 * eps2 - kp * kp))
    if q2.imag < 0:
        q2 = -q2
    top = W[1]
    A = top[6:12:2] + 1j * top[7:12:2]       # reflected (upward) wave at z_max
    B = top[12:18:2] + 1j * top[13:18:2]     # incident wave at z_max
    # both referenced at z_max = 120; the interface is at z = 60
    ph = np.exp(2j * q1 * (60.0 - 120.0))
    if pol == "s":
        r = (q1 - q2) / (q1 + q2)
        assert abs(A[1] / B[1] * ph - r) < 1e-12
    else:
        r = (eps2 * q1 - q2) / (eps2 * q1 + q2)     # H_s ratio of the TM wave
        # E_p of the waves; |H| ratio = |E| ratio in vacuum (|ph| = 1)
        assert abs(np.linalg.norm(A) - abs(r)) < 1e-12
    # Poynting flux S_z = Re(E x conj(H))_z / 2 with H = K x E / k0 (mu = 1)
    p_hat = np.array([1.0, 0.0, 0.0])
    z_hat = np.array([0.0, 0.0, 1.0])

    def flux(E, q):
        K = kp * p_hat + q * z_hat
        H = np.cross(K, E) / k0
        return 0.5 * np.real(np.cross(E, np.conj(H))[2])

    bot = W[0]
    Bt = bot[12:18:2] + 1j * bot[13:18:2]    # transmitted wave at the interface
    S_in = flux(B, -q1)                       # downward: negative
    R = -flux(A, q1) / S_in
    T = flux(Bt, -q2) / S_in
    # flux continuity at the interface (absorption happens below it)
    assert abs(R + T - 1.0) < 1e-12


def _z_field(p, rng):
    """A complex field in the Nedelec space of order p that is periodic in
    x and y: each component a polynomial of degree p-1 in z."""
    c = rng.standard_normal((3, p)) + 1j * rng.standard_normal((3, p))

    def F(X):
        z = X[:, 2] / 100.0
        return np.stack([sum(c[i, k] * z ** k for k in range(p)) for i in range(3)], axis=1)
    return F


@pytest.mark.parametrize("p", [1, 2, 3])
def test_periodic_msh_round_trip(tmp_path, p):
    """A mesh written as .per.msh (nodes per unwrapped position, $Periodic
    node pairs, PEC caps as boundary triangles) reads back as the same
    discrete space: same counts, same essential set size, and the same
    quadratic form u^T A u for an exactly interpolated periodic field."""
    from lithofem_amd.fespace import interpolate
    from lithofem_amd.meshio import read_periodic_msh, write_periodic_msh
    from oracle import oracle as O
    c = contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=6.0, phi=20.0)
    path = tmp_path / "cell.per.msh"
    write_periodic_msh(path, c.mesh)
    m2 = read_periodic_msh(path)
    assert m2.ne == c.mesh.ne and m2.nv == c.mesh.nv
    assert np.array_equal(m2.attr, c.mesh.attr)
    assert np.allclose(m2.elem_xyz, c.mesh.elem_xyz, atol=1e-12)
    forms = []
    for mesh in (c.mesh, m2):
        ad = build_asm_data(mesh, p)
        I, J = csr_symbolic(ad.dof, ad.height)
        vre, vim = O.Oracle(ad, c.problem, I, J).assemble()
        A = (vre + 1j * vim)
        u = interpolate(ad, _z_field(p, np.random.default_rng(5)))
        rows = np.repeat(np.arange(ad.height), np.diff(I))
        forms.append((ad.height, len(J), len(ad.ess), np.sum(u[rows] * A * u[J])))
    (n1, z1, e1, q1), (n2, z2, e2, q2) = forms
    assert (n1, z1, e1) == (n2, z2, e2)
    assert abs(q1 - q2) <= 1e-11 * abs(q1)


def test_solve_json_problem():
    """The reference's solve.json layout (regions slab-major, frustums after
    the slab background, PML regions last) gives the same matrix and RHS as
    the same cell described directly, with the attributes renumbered."""
    from lithofem_amd.incident import incident_set
    from lithofem_amd.solvejson import load_solve_json
    from oracle import oracle as O
    c = contact_hole(n_xy=3, nz_scale=0.3, side=30.0, theta=6.0, phi=0.0)
    eps_abs = c.problem.regions[1].eps
    cplx = lambda z: [complex(z).real, complex(z).imag]          # noqa: E731
    k0 = c.problem.k0
    d = {
        "schema_version": 1, "wavelength": 13.5, "k0": k0, "lateral_bc": "periodic",
        "domain": {"Lx": 96, "Ly": 96, "z_min": 0.0, "z_max": 120.0},
        "slabs": [0.0, 60.0, 120.0],
        "regions": [
            {"slab": 0, "z": [0, 60], "kind": "background", "epsilon": cplx(eps_abs)},
            {"slab": 0, "z": [0, 60], "kind": "frustum", "frustum": 0, "epsilon": cplx(1.0)},
            {"slab": 1, "z": [60, 120], "kind": "background", "epsilon": cplx(1.0)},
            {"slab": -1, "z": [-13.5, 0], "kind": "pml_bottom", "epsilon": cplx(eps_abs)},
            {"slab": 2, "z": [120, 133.5], "kind": "pml_top", "epsilon": cplx(1.0)},
        ],
        "sources": [
            {"type": "planewave", "amplitude": cplx(1.0), "theta": 6.0, "phi": 0.0,
             "from": "top", "polarization": "s", "jones": None,
             "kpar": [c.problem.kx, c.problem.ky]},
            {"type": "planewave", "amplitude": cplx(1.0), "theta": 6.0, "phi": 0.0,
             "from": "top", "polarization": "jones",
             "jones": [cplx(1 / np.sqrt(2)), cplx(1j / np.sqrt(2))],
             "kpar": [c.problem.kx, c.problem.ky]},
        ],
        "groups": [{"kpar": [c.problem.kx, c.problem.ky], "source_indices": [0, 1]}],
        "pml": {"thickness_wavelengths": 1.0, "order": 2.0, "target_reflection": 1e-8},
    }
    (g,) = load_solve_json(d)
    pb = g.problem
    for top in (False, True):
        assert abs(pb.smax(top) - c.problem.smax(top)) <= 1e-15 * c.problem.smax(top)
    # the direct cell numbers regions pml_bottom, absorber, hole, vacuum, pml_top
    remap = np.array([0, 4, 1, 2, 3, 5], dtype=np.int32)
    mesh2 = dataclasses.replace(c.mesh, attr=remap[c.mesh.attr])
    ad1 = build_asm_data(c.mesh, 2)
    ad2 = build_asm_data(mesh2, 2)
    I, J = csr_symbolic(ad1.dof, ad1.height)
    o1 = O.Oracle(ad1, c.problem, I, J)
    o2 = O.Oracle(ad2, pb, I, J)
    a1, a2 = o1.assemble(), o2.assemble()
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
    o1.set_incident(incident_set([c.problem], [g.fields]))
    o2.set_incident(incident_set([pb], [g.fields]))
    for r in range(2):
        assert np.array_equal(o1.rhs(r), o2.rhs(r))
    # Jones (1, i)/sqrt 2 read as the transverse (Ex, Ey) of the incident field
    assert g.fields[1][2] == (pytest.approx(1 / np.sqrt(2)), pytest.approx(1j / np.sqrt(2)))
    assert g.tables is None and g.incident_fields() is g.fields
    # the group's own incident tables (the reference driver's layout) are
    # read when present, and reproduce the recomputed RHS
    from lithofem_amd.incident import slab_waves as sw, tables_to_solve_json
    bounds = d["slabs"]
    d2 = json.loads(json.dumps(d))
    d2["groups"][0]["incident"] = tables_to_solve_json(
        {i: sw(pb, *f) for i, f in zip(g.source_indices, g.fields)}, bounds)
    (g2,) = load_solve_json(d2)
    assert g2.tables is not None and len(g2.tables) == 2
    o3 = O.Oracle(ad2, pb, I, J)
    o3.set_incident(incident_set([pb], [g2.incident_fields()]))
    for r in range(2):
        want = o2.rhs(r)
        assert np.abs(o3.rhs(r) - want).max() <= 1e-14 * np.abs(want).max()


@pytest.mark.parametrize("name,p", [("grating", 2), ("grating", 3), ("hole", 2), ("hole", 3)])
def test_parity_cases_cover_every_face_orientation(name, p):
    """The cases of the GPU parity tests hold every one of the six face
    orientations (bitmask 63), so each TInv block of the face transform is
    exercised; the reference's own check only observes {0, 1, 3, 5}
    (fo coverage 43, tests/test_asm_v25.py:134-140)."""
    from cases import case
    _, ad, _, _ = case(name, p, 6.0, "small")
    assert sum(1 << int(v) for v in np.unique(ad.fo)) == 63
