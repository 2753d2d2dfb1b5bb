"""Pins of the CPU oracle (oracle/asm_oracle.c) -- CPU only.

Each check is independent of the oracle's own code path, chosen so that a
plausible mistake (transposed Piola map, s and 1/s swapped, a dropped or
sign-flipped k_par term, a wrong face transform or dof sign, a scatter
that loses or duplicates contributions) fails at least one of them:

* reference-held golden data: the ND face transform tables;
* closed forms: PML stretch values; lowest-order (Whitney) element
  matrices of all five integrators from barycentric gradients, without and
  with a constant complex PML stretch;
* invariants: symmetry, CROSS1 = -CROSS0^T, curl-curl null space dimension,
  the eps-free mass is SPD, generalised eigenvalues of (curl-curl, mass)
  do not change under the face-DOF transform, sum of detJ / 6 = volume;
* global identities through face transform, signs and scatter: gradients
  of global polynomials are in the curl-curl kernel; closed-form quadratic
  forms of constant and linear fields;
* brute force on tiny meshes: CSR pattern, scatter, elimination, residual;
* finite differences for the eps sensitivity;
* end to end: manufactured solution on the unit cube (curl curl E -
  k0^2 eps E = f, non-homogeneous PEC), L2 convergence rate >= p - 0.2,
  complex symmetry and solve residual as in the reference's M4 test.
"""

from __future__ import annotations

import copy
import json
from pathlib import Path

import numpy as np
import pytest

from cases import case, maxnorm_err
from lithofem_amd.fespace import build_asm_data, csr_symbolic, interpolate
from lithofem_amd.mesh import structured_cell_mesh
from lithofem_amd.nedelec import EDGES, ND_T, ND_TINV
from lithofem_amd.problem import Problem, Region, line_grating
from oracle import oracle as O
import mms

GOLDEN = Path(__file__).with_name("golden")


# ---------------------------------------------------------------- golden data

def test_face_tables_match_reference_fixture():
    ref = json.loads((GOLDEN / "nd_face_tables.json").read_text())
    assert np.array_equal(ND_T, np.array(ref["asm_nd_T"]))
    assert np.array_equal(ND_TINV, np.array(ref["asm_nd_TInv"]))
    for o in range(6):
        t = ND_T[4 * o:4 * o + 4]
        ti = ND_TINV[4 * o:4 * o + 4]
        T = np.array([[t[0], t[2]], [t[1], t[3]]])
        Ti = np.array([[ti[0], ti[2]], [ti[1], ti[3]]])
        assert np.array_equal(T @ Ti, np.eye(2))


# ---------------------------------------------------------------- PML

def test_s_z_closed_form():
    c, ad, I, J = case("grating", 1, 6.0, "tiny")
    pb = c.problem
    orc = O.Oracle(ad, pb, I, J)
    k0 = 2 * np.pi / 13.5
    n_cap = abs(np.sqrt((0.97 + 0.01j) ** 2))
    smax_bot = -np.log(1e-8) * 3.0 / (2 * k0 * n_cap * 13.5)
    smax_top = -np.log(1e-8) * 3.0 / (2 * k0 * 1.0 * 13.5)
    assert abs(smax_top - 4.397613593276567) < 1e-12
    assert orc.s_z(60.0) == 1 + 0j
    assert orc.s_z(0.0) == 1 + 0j and orc.s_z(120.0) == 1 + 0j
    assert abs(orc.s_z(-6.75) - (1 + 0.25j * smax_bot)) < 1e-15
    assert abs(orc.s_z(120 + 13.5) - (1 + 1j * smax_top)) < 1e-15
    assert abs(orc.s_z(120 + 40.0) - (1 + 1j * smax_top)) < 1e-15   # d clamps at 1
    assert abs(pb.smax(False) - smax_bot) < 1e-14 and abs(pb.smax(True) - smax_top) < 1e-14


# ---------------------------------------------------------------- Whitney closed forms

def _whitney(X):
    """Barycentric gradients and the closed-form p=1 element tensors of the
    tet with vertices X (4x3): vol, int(phi_k) (6x3), curl (6x3) and
    T[k,l] = int phi_k phi_l^T (6x6x3x3)."""
    A = np.stack([X[1] - X[0], X[2] - X[0], X[3] - X[0]], axis=1)
    vol = np.linalg.det(A) / 6.0
    Gi = np.linalg.inv(A)            # rows: grad lambda_1..3
    G = np.vstack([-Gi.sum(0), Gi])
    M2 = vol * (np.ones((4, 4)) + np.eye(4)) / 20.0     # int lambda_m lambda_n
    curl = np.array([2 * np.cross(G[i], G[j]) for (i, j) in EDGES])
    mean = np.array([vol / 4 * (G[j] - G[i]) for (i, j) in EDGES])
    T = np.zeros((6, 6, 3, 3))
    for k, (i, j) in enumerate(EDGES):
        for l, (a, b) in enumerate(EDGES):
            T[k, l] = (M2[i, a] * np.outer(G[j], G[b]) - M2[i, b] * np.outer(G[j], G[a])
                       - M2[j, a] * np.outer(G[i], G[b]) + M2[j, b] * np.outer(G[i], G[a]))
    return vol, mean, curl, T


def test_whitney_curlcurl_linear_pml_pins_rule_point_z():
    """z at the rule points: with pml_order = 1 the stretch s(z) is linear
    in depth, and at p = 1 the curl-curl rule (order 0) is the centroid, so
    cc = vol curl diag(1/s, 1/s, s)(z_c) curl^T with z_c the mean vertex z.
    The tet is skewed so that the third row of J (z of the three edges,
    the right map z = v0_z + J[6..8] xi) and its third column (the three
    components of edge 3) differ: a transposed map would move z_c."""
    import dataclasses
    c = line_grating(nx=3, ny=3, nz_scale=0.1, theta=6.0)
    pb = copy.deepcopy(c.problem)
    pb.pml_order = 1.0
    ad = build_asm_data(c.mesh, 1)
    e = int(np.nonzero(ad.attr == 6)[0][0])              # a top-PML element
    X = np.array([[10.0, 3.0, 121.0], [14.0, 3.5, 124.0], [11.0, 8.0, 122.5],
                  [9.5, 4.0, 128.0]])
    A = (X[1:] - X[0]).T                                  # columns: edges 1..3
    assert np.linalg.det(A) > 0 and abs(A[2].sum() - A[:, 2].sum()) > 1
    Jn, Jin, dn, vn = ad.J.copy(), ad.Jinv.copy(), ad.detJ.copy(), ad.v0.copy()
    Jn.reshape(-1, 3, 3)[e] = A
    Jin.reshape(-1, 3, 3)[e] = np.linalg.inv(A)
    dn[e] = np.linalg.det(A)
    vn.reshape(-1, 3)[e] = X[0]
    ad = dataclasses.replace(ad, J=Jn, Jinv=Jin, detJ=dn, v0=vn)
    I, J = csr_symbolic(ad.dof, ad.height)
    orc = O.Oracle(ad, pb, I, J)
    vol, _, curl, _ = _whitney(X)
    zc = X[:, 2].mean()
    s = 1 + 1j * pb.smax(True) * (zc - pb.z_max) / pb.pml_thick
    assert abs(orc.s_z(zc) - s) < 1e-15
    ref = vol * np.einsum("ki,i,li->kl", curl, np.array([1 / s, 1 / s, s]), curl)
    got = orc.element(e, O.CURLCURL, False)
    assert maxnorm_err(got, ref) < 1e-13, maxnorm_err(got, ref)
    # the transposed map would evaluate s at another depth
    z_wrong = X[0, 2] + A[:, 2].sum() / 4
    s_w = 1 + 1j * pb.smax(True) * (z_wrong - pb.z_max) / pb.pml_thick
    wrong = vol * np.einsum("ki,i,li->kl", curl, np.array([1 / s_w, 1 / s_w, s_w]), curl)
    assert maxnorm_err(wrong, ref) > 1e-3


def _kmat(kx, ky):
    return np.array([[0, 0, ky], [0, 0, -kx], [-ky, kx, 0]], dtype=float)


@pytest.mark.parametrize("pml", [False, True], ids=["bulk", "pml_const_s"])
def test_whitney_closed_form_all_integrators(pml):
    c = line_grating(nx=3, ny=3, nz_scale=0.1, theta=6.0, warp=0.4)
    pb = copy.deepcopy(c.problem)
    if pml:
        pb.pml_order = 0.0          # s = 1 + i smax, constant inside the PML
    ad = build_asm_data(c.mesh, 1)
    I, J = csr_symbolic(ad.dof, ad.height)
    orc = O.Oracle(ad, pb, I, J)
    z_all = ad.v0[:, 2]
    if pml:
        e = int(np.nonzero(ad.attr == 6)[0][3])          # top PML
        s = 1 + 1j * pb.smax(True)
    else:
        e = int(np.nonzero((ad.attr == 4) & (z_all > 70))[0][0])   # absorber bulk
        s = 1.0 + 0j
    X = c.mesh.elem_xyz[e]
    vol, mean, curl, T = _whitney(X)
    li = np.array([1 / s, 1 / s, s])
    lam = np.array([s, s, 1 / s])
    eps = pb.regions[ad.attr[e] - 1].eps
    K = _kmat(pb.kx, pb.ky)
    k2 = pb.k0 ** 2
    cc = vol * np.einsum("ki,i,li->kl", curl, li, curl)
    mass = -k2 * eps * np.einsum("klij,ij->kl", T, np.diag(lam))
    cross0 = 1j * np.einsum("ki,ij,lj->kl", mean, K @ np.diag(li), curl)
    cross1 = 1j * np.einsum("ki,ij,lj->kl", curl, np.diag(li) @ K, mean)
    klk = -np.einsum("klij,ij->kl", T, K @ np.diag(li) @ K)
    for mask, ref in ((O.CURLCURL, cc), (O.MASS, mass), (O.CROSS0, cross0),
                      (O.CROSS1, cross1), (O.KLK, klk)):
        got = orc.element(e, mask, False)
        assert maxnorm_err(got, ref) < 1e-13, (mask, maxnorm_err(got, ref))
    assert abs(vol - ad.detJ[e] / 6) < 1e-14 * vol


# ---------------------------------------------------------------- invariants

@pytest.mark.parametrize("p", [2, 3, 4])
def test_element_invariants(p):
    c, ad, I, J = case("grating", p, 6.0, "tiny")
    orc = O.Oracle(ad, c.problem, I, J)
    bulk = int(np.nonzero((ad.attr == 2))[0][5])
    pml = int(np.nonzero((ad.attr == 1))[0][2])
    for e in (bulk, pml):
        cc = orc.element(e, O.CURLCURL, True)
        ms = orc.element(e, O.MASS, True)
        kl = orc.element(e, O.KLK, True)
        c0 = orc.element(e, O.CROSS0, True)
        c1 = orc.element(e, O.CROSS1, True)
        for m in (cc, ms, kl):
            assert maxnorm_err(m.T, m) < 1e-13
        assert maxnorm_err(-c0.T, c1) < 1e-13
        full = orc.element(e, O.ALL, True)
        assert maxnorm_err(full, cc + ms + kl + c0 + c1) < 1e-13
    # curl-curl kernel = gradients of degree-p H1 functions
    cc = orc.element(bulk, O.CURLCURL, True).real
    sv = np.linalg.svd(cc, compute_uv=False)
    null = int(np.sum(sv < 1e-10 * sv[0]))
    assert null == (p + 1) * (p + 2) * (p + 3) // 6 - 1


@pytest.mark.parametrize("p", [1, 2, 3])
def test_mass_spd_and_transform_invariance(p):
    """eps-free mass of a vacuum element is SPD; the face transform is a
    congruence, so the generalised eigenvalues of (curl-curl, mass) of a
    constant-coefficient element are the same with and without it."""
    import scipy.linalg as sl
    c, ad, I, J = case("grating", p, 0.0, "tiny")
    pb = c.problem
    orc = O.Oracle(ad, pb, I, J)
    vac = [e for e in range(ad.ne)
           if not pb.regions[ad.attr[e] - 1].is_pml and pb.regions[ad.attr[e] - 1].eps == 1.0]
    e = vac[len(vac) // 2]
    for te in (False, True):
        m0 = orc.element(e, O.MASS, te) / (-pb.k0 ** 2)
        assert np.max(np.abs(m0.imag)) == 0.0
        assert np.linalg.eigvalsh(m0.real).min() > 0.0
    lam = []
    for te in (False, True):
        cc = orc.element(e, O.CURLCURL, te).real
        m0 = (orc.element(e, O.MASS, te) / (-pb.k0 ** 2)).real
        lam.append(np.sort(sl.eigvalsh(cc, m0)))
    assert np.max(np.abs(lam[0] - lam[1])) < 1e-9 * np.max(np.abs(lam[0]))
    # and the transform does change the matrices themselves (p >= 2)
    if p >= 2:
        assert maxnorm_err(orc.element(e, O.CURLCURL, True), orc.element(e, O.CURLCURL, False)) > 1e-3


def test_geometry_volume():
    """Element Jacobians tile the unit cube: sum of detJ / 6 = 1 (interior
    vertices moved by a smooth map that fixes the boundary)."""
    ad = build_asm_data(mms.cube_mesh(5), 1)
    assert np.all(ad.detJ > 0)
    assert abs(np.sum(ad.detJ) / 6.0 - 1.0) < 1e-13


def _orc_assemble(ad, pb, I, J):
    orc = O.Oracle(ad, pb, I, J)
    return orc.scatter(orc.elements(0, ad.ne, O.CURLCURL | O.MASS, True))


@pytest.mark.parametrize("eps", [2.25, 2.25 + 0.75j], ids=["real", "complex"])
@pytest.mark.parametrize("p,ns", [(1, (4, 8)), (2, (2, 4)), (3, (2, 4))], ids=["p1", "p2", "p3"])
def test_mms_convergence(p, ns, eps):
    rate, levels = mms.convergence(p, eps, ns, _orc_assemble)
    print(f"p={p} eps={eps}: rate {rate:.3f}, L2 errors {[f'{e:.3e}' for _, e, _, _ in levels]}")
    assert rate >= p - 0.2, rate
    for h, err, sym, res in levels:
        assert sym < 1e-12, sym
        assert res < 1e-10, res


# ---------------------------------------------------------------- global identities

def _open_cell(p, kx=0.03, ky=-0.02, pml=True):
    xs = np.linspace(0.0, 6.0, 4)
    zs = np.linspace(-3.0, 9.0, 5)
    mesh = structured_cell_mesh(xs, xs.copy(), zs,
                                lambda x, y, z: np.where(z < 0, 1, np.where(z > 6, 3, 2)),
                                periodic=(False, False), warp_amp=0.3, warp_z=(0.0, 6.0))
    regions = [Region(1.0, True, False), Region(0.9 + 0.05j), Region(1.0, True, True)]
    pb = Problem(k0=0.46, kx=kx, ky=ky, z_min=0.0 if pml else -1e9, z_max=6.0 if pml else 1e9,
                 pml_thick=3.0, regions=regions)
    ad = build_asm_data(mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    return mesh, pb, ad, I, J


def _global(orc, ad, mask):
    Ae = orc.elements(0, ad.ne, mask, True)
    vre, vim = orc.scatter(Ae)
    return vre + 1j * vim


def _spmv(I, J, v, x):
    y = np.zeros(len(I) - 1, dtype=np.complex128)
    rows = np.repeat(np.arange(len(I) - 1), np.diff(I))
    np.add.at(y, rows, v * x[J])
    return y


@pytest.mark.parametrize("p", [1, 2, 3])
def test_gradients_in_global_curlcurl_kernel(p):
    mesh, pb, ad, I, J = _open_cell(p)
    orc = O.Oracle(ad, pb, I, J)
    A = _global(orc, ad, O.CURLCURL)
    # grad of phi = x^a y^b z^c with a+b+c = p (a global degree-p polynomial)
    a, b, cz = (p, 0, 0) if p == 1 else (1, p - 1, 0) if p == 2 else (1, 1, p - 2)

    def gradphi(X):
        x, y, z = X[:, 0], X[:, 1], X[:, 2]
        return np.stack([a * x ** max(a - 1, 0) * y ** b * z ** cz,
                         b * x ** a * y ** max(b - 1, 0) * z ** cz,
                         cz * x ** a * y ** b * z ** max(cz - 1, 0)], 1)

    g = interpolate(ad, gradphi)
    y = _spmv(I, J, A, g)
    scale = np.max(np.abs(A)) * np.max(np.abs(g))
    assert np.max(np.abs(y)) < 1e-11 * scale, np.max(np.abs(y)) / scale
    # the same vector with one face-pair transform dropped is NOT in the kernel
    if p >= 2:
        bad = copy.deepcopy(ad)
        bad.fo = np.zeros_like(ad.fo)
        orc_bad = O.Oracle(bad, pb, I, J)
        yb = _spmv(I, J, _global(orc_bad, bad, O.CURLCURL), g)
        assert np.max(np.abs(yb)) > 1e-6 * scale


@pytest.mark.parametrize("p", [1, 2, 3])
def test_quadratic_forms_closed_form(p):
    mesh, pb, ad, I, J = _open_cell(p, pml=False)
    orc = O.Oracle(ad, pb, I, J)
    V = float(np.sum(ad.detJ) / 6.0)
    K = _kmat(pb.kx, pb.ky)
    F0 = np.array([0.3, -1.1, 0.7])
    g0 = interpolate(ad, lambda X: np.tile(F0, (len(X), 1)))
    # mass: -k0^2 sum_e eps_e vol_e |F|^2
    eps_e = np.array([pb.regions[a - 1].eps for a in ad.attr])
    ref_mass = -pb.k0 ** 2 * np.sum(eps_e * ad.detJ / 6.0) * (F0 @ F0)
    M = _global(orc, ad, O.MASS)
    assert abs(g0 @ _spmv(I, J, M, g0) - ref_mass) < 1e-12 * abs(ref_mass)
    KL = _global(orc, ad, O.KLK)
    ref_klk = -V * (F0 @ K @ K @ F0)
    assert abs(g0 @ _spmv(I, J, KL, g0) - ref_klk) < 1e-12 * abs(ref_klk)
    # linear field F = a x r, curl F = 2a
    av = np.array([0.2, -0.5, 0.9])
    g1 = interpolate(ad, lambda X: np.cross(np.tile(av, (len(X), 1)), X))
    centroid = np.sum(ad.detJ[:, None] / 6.0 * (ad.v0 + ad.J.reshape(-1, 3, 3) @ np.full(3, 0.25)),
                      axis=0)   # int r dV
    C0 = _global(orc, ad, O.CROSS0)
    C1 = _global(orc, ad, O.CROSS1)
    ref_c0 = 2j * (np.cross(av, centroid) @ K @ av)
    ref_c1 = 2j * (av @ K @ np.cross(av, centroid))
    assert abs(g1 @ _spmv(I, J, C0, g1) - ref_c0) < 1e-11 * abs(ref_c0)
    assert abs(g1 @ _spmv(I, J, C1, g1) - ref_c1) < 1e-11 * abs(ref_c1)
    CC = _global(orc, ad, O.CURLCURL)
    ref_cc = V * 4 * (av @ av)
    assert abs(g1 @ _spmv(I, J, CC, g1) - ref_cc) < 1e-12 * ref_cc


# ---------------------------------------------------------------- brute force

@pytest.mark.parametrize("p", [1, 2])
def test_csr_scatter_elimination_brute_force(p):
    c, ad, I, J = case("hole", p, 6.0, "tiny")
    n = ad.height
    # CSR pattern: dense union of element blocks
    dense = np.zeros((n, n), dtype=bool)
    for e in range(ad.ne):
        dense[np.ix_(ad.dof[e], ad.dof[e])] = True
    Io, Jo = O.csr_symbolic(ad.dof, n)
    assert np.array_equal(Io, I) and np.array_equal(Jo, J)
    rows, cols = np.nonzero(dense)
    assert np.array_equal(np.diff(I), np.bincount(rows, minlength=n))
    assert np.array_equal(J, cols.astype(np.int32))
    # scatter: dense sum of sign-scaled element matrices
    orc = O.Oracle(ad, c.problem, I, J)
    Ae = orc.elements(0, ad.ne)
    Ad = np.zeros((n, n), dtype=np.complex128)
    for e in range(ad.ne):
        s = ad.sign[e]
        Ad[np.ix_(ad.dof[e], ad.dof[e])] += s[:, None] * Ae[e] * s[None, :]
    vre, vim = orc.assemble()
    assert maxnorm_err(vre + 1j * vim, Ad[rows, cols]) < 1e-14
    # elimination: essential rows/cols -> 0, unit diagonal
    is_ess = np.zeros(n, np.uint8)
    is_ess[ad.ess] = 1
    ere, eim = O.eliminate(I, J, is_ess, vre, vim)
    E = Ad.copy()
    E[ad.ess, :] = 0
    E[:, ad.ess] = 0
    E[ad.ess, ad.ess] = 1
    assert np.array_equal(ere + 1j * eim, E[rows, cols])
    # residual
    rng = np.random.default_rng(3)
    X = rng.standard_normal(2 * n)
    B = rng.standard_normal(2 * n)
    x = X[:n] + 1j * X[n:]
    b = B[:n] + 1j * B[n:]
    ref = np.linalg.norm(E @ x - b) / np.linalg.norm(b)
    assert abs(O.residual(I, J, ere, eim, X, B) - ref) < 1e-12 * ref


def _inc(pb, fields):
    from lithofem_amd.incident import incident_set
    return incident_set([pb], [fields])


def test_eps_sensitivity_matches_finite_differences():
    c, ad, I, J = case("grating", 2, 6.0, "tiny")
    pb = c.problem
    inc = _inc(pb, [(6.0, 0.0, "p")])
    rng = np.random.default_rng(11)
    eps0 = np.array([pb.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
    lam = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
    u = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
    o0 = O.Oracle(ad, pb, I, J, eps_elem=eps0)
    o0.set_incident(inc)
    g = o0.eps_sensitivity(lam, u, rhs_col=0)
    src = np.nonzero(ad.attr == 4)[0]            # absorber line: source region
    bulk = np.nonzero(ad.attr == 2)[0]           # capping: no source
    picks = list(rng.choice(src, 3, replace=False)) + list(rng.choice(bulk, 2, replace=False))
    for e in picks:
        for dirn in (1.0, 1j):
            h = 1e-3
            vals = []
            for sgn in (1, -1):
                ee = eps0.copy()
                ee[e] += sgn * h * dirn
                oe = O.Oracle(ad, pb, I, J, eps_elem=ee)
                oe.set_incident(inc)
                vre, vim = oe.assemble()
                vals.append(lam @ (_spmv(I, J, vre + 1j * vim, u) - oe.rhs(0)))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[e] * dirn) < 1e-8 * abs(g[e]), (e, dirn, fd, g[e])


@pytest.mark.parametrize("name,p", [("grating", 3), ("hole", 2)])
def test_per_nonzero_metric_and_scales(name, p):
    """The componentwise scale bounds every entry (|r_k| <= s_k), and two
    correct evaluation orders of the same formulas -- the oracle and its
    own -O3 FMA-contracted build -- agree to ~1e-16 of the scale at every
    nonzero.  Their plain relative error |g - r| / |r| over the entries at
    least 1e-8 of the largest is far larger (entries that cancel): that
    metric measures the rounding order of small entries, which is why the
    GPU parity tests assert it on well-conditioned nonzeros and the scaled
    metric on all of them."""
    from cases import pernz_err, scaled_err
    c, ad, I, J = case(name, p, 6.0, "small")
    o1 = O.Oracle(ad, c.problem, I, J)
    o2 = O.Oracle(ad, c.problem, I, J, fast=True)
    a1, a2 = o1.assemble(), o2.assemble()
    v1, v2 = a1[0] + 1j * a1[1], a2[0] + 1j * a2[1]
    assert not np.array_equal(v1, v2)
    sc = o1.scale_csr(4)
    assert np.all(np.abs(v1) <= sc * (1 + 1e-12))
    assert maxnorm_err(v2, v1) < 1e-14
    err_s = scaled_err(v2, v1, sc)
    err_g = pernz_err(v2, v1)
    print(f"{name} p={p}: scaled {err_s:.2e}, plain per-nonzero at 1e-8 gate {err_g:.2e}")
    assert err_s < 2e-15
    assert err_g > 50 * err_s
    E1, E2 = o1.elements(0, 150), o2.elements(0, 150)
    S = o1.scale_elements(np.arange(150))
    assert np.all(np.abs(E1) <= S * (1 + 1e-12))
    assert scaled_err(E2, E1, S) < 2e-15
    # the scale is the modulus sum: equal to |r| for a single real product
    # chain, e.g. the curl-curl entries of an element with constant s
    Sc = o1.scale_elements([0], O.CURLCURL, False)[0]
    Ec = o1.elements(0, 1, O.CURLCURL, False)[0]
    assert np.all(np.abs(np.diag(Ec)) <= np.diag(Sc) * (1 + 1e-12))


def test_geometry_from_vertices():
    """orc_geometry (the operation order of the device extraction): J columns
    are the element's edge vectors from vertex 0, J J^-1 = I, detJ = 6 x the
    signed volume from the scalar triple product, v0 = vertex 0, and the
    mesh producer's numpy geometry agrees to rounding."""
    c, ad, I, J = case("hole", 1, 6.0, "small")
    x = np.asarray(c.mesh.elem_xyz, dtype=np.float64)
    Jo, Jio, det, v0 = O.geometry(x)
    A = Jo.reshape(-1, 3, 3)
    for j in range(3):
        assert np.array_equal(A[:, :, j], x[:, j + 1] - x[:, 0])
    assert np.array_equal(v0, x[:, 0])
    eye = np.einsum("eij,ejk->eik", A, Jio.reshape(-1, 3, 3))
    assert np.abs(eye - np.eye(3)).max() < 1e-13
    e1, e2, e3 = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]
    triple = np.einsum("ei,ei->e", e1, np.cross(e2, e3))
    assert np.abs(det - triple).max() <= 1e-13 * np.abs(triple).max()
    assert np.abs(det - ad.detJ).max() <= 1e-13 * np.abs(ad.detJ).max()
    assert np.abs(Jio - ad.Jinv).max() <= 1e-12 * np.abs(ad.Jinv).max()


def test_threaded_oracle_equals_serial():
    """The host-thread forms used at full size (element blocks; disjoint CSR
    row ranges of the scatter) are bit-identical to the serial oracle."""
    c, ad, I, J = case("hole", 2, 6.0, "small")
    orc = O.Oracle(ad, c.problem, I, J)
    A1 = orc.elements(0, ad.ne)
    assert np.array_equal(orc.elements_threaded(5), A1)
    s1 = orc.scatter(A1)
    s2 = orc.scatter_threaded(A1, 7)
    assert np.array_equal(s1[0], s2[0]) and np.array_equal(s1[1], s2[1])
    a = orc.assemble()
    assert np.array_equal(a[0], s1[0]) and np.array_equal(a[1], s1[1])


def test_row_blocks_equal_full_reassembly():
    """The row-sample forms used on meshes too large for a full reassembly:
    values, scales and elimination of a few row blocks (only their
    contributor elements reassembled) equal the full oracle's bit for bit on
    those blocks."""
    c, ad, I, J = case("grating", 3, 6.0, "small")
    orc = O.Oracle(ad, c.problem, I, J)
    vre, vim = orc.assemble()
    s = orc.scale_csr(3)
    n = ad.height
    blocks = [(0, 37), (n // 2, n // 2 + 51), (n - 29, n)]
    bre, bim, bs = orc.row_blocks(blocks, 3)
    is_ess = np.zeros(n, np.uint8)
    is_ess[ad.ess] = 1
    ere, eim = O.eliminate(I, J, is_ess, vre, vim)
    fre, fim = O.eliminate_rows(I, J, is_ess, bre, bim, blocks)
    for r0, r1 in blocks:
        k = slice(I[r0], I[r1])
        assert np.array_equal(bre[k], vre[k]) and np.array_equal(bim[k], vim[k])
        assert np.array_equal(bs[k], s[k])
        assert np.array_equal(fre[k], ere[k]) and np.array_equal(fim[k], eim[k])
    assert np.any(is_ess[blocks[0][0]:blocks[0][1]]) or np.any(is_ess[blocks[2][0]:blocks[2][1]])


def test_sensitivity_element_list_equals_full():
    """The element-list form (host threads) of the sensitivity gives the full
    form's values bit for bit, in list order."""
    from lithofem_amd.incident import incident_set
    c, ad, I, J = case("hole", 2, 6.0, "small")
    orc = O.Oracle(ad, c.problem, I, J)
    orc.set_incident(incident_set([c.problem], [[(6.0, 30.0, "p")]]))
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
    u = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
    elems = rng.permutation(ad.ne)[: ad.ne // 3]
    for col in (-1, 0):
        full = orc.eps_sensitivity(lam, u, col)
        assert np.array_equal(orc.eps_sensitivity(lam, u, col, elems=elems, nthreads=4),
                              full[elems])


def test_rhs_constant_field_equals_mass_action():
    """For a constant incident field E0 and source strength deps, the RHS is
    b = -deps * M_src g(E0), with M_src the unit-eps mass matrix of the source
    elements and g(E0) the DOF vector of E0 (exact in the discrete space)."""
    from lithofem_amd.incident import IncidentSet
    c, ad, I, J = case("hole", 2, 6.0, "tiny")
    pb = c.problem
    E0 = np.array([0.3 - 0.2j, 1.1 + 0.4j, -0.7 + 0.1j])
    nslab = len(pb.slabs)
    w = np.zeros((nslab, 18))
    w[:, 6:12] = np.stack([E0.real, E0.imag], 1).ravel()      # A = E0, qA = 0
    inc = IncidentSet(nrhs=1, nslab=nslab, waves=w,
                      region_slab=np.array([r.slab for r in pb.regions], np.int32),
                      region_deps=np.array([r.deps for r in pb.regions]),
                      region_src=np.array([1 if r.source else 0 for r in pb.regions], np.int32))
    orc = O.Oracle(ad, pb, I, J)
    orc.set_incident(inc)
    b = orc.rhs(0)
    g = interpolate(ad, lambda X: np.tile(E0, (len(X), 1)))
    src = ad.attr == 3
    deps = pb.regions[2].deps
    Ae = orc.elements(0, ad.ne, O.MASS, True)
    Ae[~src] = 0.0
    eps_e = np.array([pb.regions[a - 1].eps for a in ad.attr])
    Ae[src] /= eps_e[src][:, None, None]            # unit-eps mass of the source elements
    vre, vim = orc.scatter(Ae)
    ref = -deps * _spmv(I, J, vre + 1j * vim, g)
    assert maxnorm_err(b, ref) < 1e-13
