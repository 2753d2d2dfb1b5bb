"""GPU tests of the per-element eps path and the batches built on it, each
against the oracle (not against another GPU result):

* element matrices and the eliminated CSR of a mesh with random per-element
  eps (+-5 % real and imaginary on the absorber, as the batched workload
  draws it) -- set at init and with lfasm_set_eps_elem -- equal the
  oracle's per-element reassembly;
* every eps variant of a delta-assembly batch equals the oracle's
  reassembly with that variant's eps;
* one RHS launch over several problems (incidence angles 0, 6, 14 deg,
  s and p each, as a block-diagonal batch) equals each problem's oracle
  RHS;
* with per-element eps, an element of a background region whose eps
  differs from the slab's is a scattered-field source: setting every
  absorber element of a contact hole to eps X gives the same matrix and RHS
  as the cell whose absorber is a source region of eps X over the same
  background.
"""

from __future__ import annotations

import copy
import math

import numpy as np
import pytest

from cases import assert_parity
from lithofem_amd import gpu
from lithofem_amd.fespace import build_asm_data, csr_symbolic
from lithofem_amd.incident import incident_set
from lithofem_amd.problem import EPS_CAP, Region, contact_hole, line_grating
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _perturbed(pb, ad, params, rng):
    """eps of every element: region eps, params perturbed by +-5 % of the
    real and imaginary parts independently."""
    eps0 = np.array([pb.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
    ee = eps0.copy()
    base = eps0[params]
    ee[params] = (base.real * (1 + 0.05 * rng.uniform(-1, 1, len(params)))
                  + 1j * base.imag * (1 + 0.05 * rng.uniform(-1, 1, len(params))))
    return eps0, ee


def _oracle_system(ad, pb, I, J, eps_elem):
    """Oracle, its eliminated CSR values and their componentwise scales."""
    orc = O.Oracle(ad, pb, I, J, eps_elem=eps_elem)
    rre, rim = orc.assemble()
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    xre, xim = O.eliminate(I, J, is_ess, rre, rim)
    return orc, xre + 1j * xim, O.eliminate_scale(I, J, is_ess, orc.scale_csr(4))


def _rhs_ref(orc, ad, r):
    want = orc.rhs(r)
    want[ad.ess] = 0.0
    sc = orc.rhs_scale(r)
    sc[ad.ess] = 0.0
    return want, sc


@pytest.mark.parametrize("tensor", [1, 0], ids=["tensor", "quadrature"])
@pytest.mark.parametrize("p", [2, 3])
def test_eps_elem_matches_oracle(p, tensor):
    c = line_grating(nx=4, ny=3, nz_scale=0.15, theta=6.0)
    pb = c.problem
    ad = build_asm_data(c.mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    params = np.nonzero(ad.attr == 4)[0]
    _, ee = _perturbed(pb, ad, params, np.random.default_rng(40 + p))
    print(f"eps_elem p={p} tensor={tensor}: ne={ad.ne} params={len(params)}", flush=True)
    orc, ref, sc = _oracle_system(ad, pb, I, J, ee)
    sample = sorted(set(params[:: max(1, len(params) // 24)].tolist()) |
                    set(np.nonzero(ad.attr != 4)[0][::97].tolist()))
    ref_el = np.stack([orc.element(e, O.ALL, True) for e in sample])
    sc_el = orc.scale_elements(sample, O.ALL, True)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb, eps_elem=ee)
        g.set_tensor_path(tensor)
        g.build_global(I, J, ad.ess)
        g.compute_local()
        assert_parity(g.download_local()[sample], ref_el, TOL, "element matrices", sc_el)
        g.scatter_eliminate(True)
        vre, vim = g.download_values()
        assert_parity(vre + 1j * vim, ref, TOL, "eliminated CSR", sc)
        # the same eps set after init gives the same values bit for bit
        g.set_eps_elem(None)
        g.compute_local()
        g.scatter_eliminate(True)
        pre, _ = g.download_values()
        assert not np.array_equal(pre, vre)
        g.set_eps_elem(ee)
        g.compute_local()
        g.scatter_eliminate(True)
        wre, wim = g.download_values()
        assert np.array_equal(wre, vre) and np.array_equal(wim, vim)
    finally:
        g.close()


@pytest.mark.parametrize("p", [2, 3])
def test_eps_variants_match_oracle(p):
    # at this size both variant kernels run: the row kernel and the range
    # copy of long runs of parameter-free rows
    c = line_grating(nx=4, ny=3, nz_scale=0.15, theta=6.0)
    pb = c.problem
    ad = build_asm_data(c.mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    params = np.nonzero(ad.attr == 4)[0].astype(np.int32)
    nvar = 6
    eps0 = np.array([pb.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
    deps = np.zeros((nvar, len(params)), dtype=np.complex128)
    for v in range(nvar):
        _, ee = _perturbed(pb, ad, params, np.random.default_rng(v))
        deps[v] = ee[params] - eps0[params]
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        g.compute_local()
        g.scatter_eliminate(True)
        g.variants_prepare(params)
        g.variants_assemble(deps, eliminate=True)
        var = [g.download_variant(v) for v in range(nvar)]
    finally:
        g.close()
    for v in range(nvar):
        ee = eps0.copy()
        ee[params] += deps[v]
        _, ref, sc = _oracle_system(ad, pb, I, J, ee)
        assert_parity(var[v][0] + 1j * var[v][1], ref, TOL, f"variant {v}", sc)


def _union_pattern(I, J, n, k):
    nnz = len(J)
    Iu = np.concatenate([I[:-1] + i * nnz for i in range(k)] + [np.array([k * nnz])])
    Ju = np.concatenate([J + i * n for i in range(k)])
    return Iu.astype(np.int32), Ju.astype(np.int32)


def test_batched_rhs_over_angles_matches_oracle():
    """Several problems in one RHS launch: the same contact-hole mesh at
    theta = 0, 6, 14 deg (k_par and so the matrix differ per angle), s and p
    fields each, assembled as one block-diagonal batch."""
    p = 2
    base = contact_hole(n_xy=4, nz_scale=0.34, side=44.0, theta=0.0, sidewall_deg=84.0)
    ad = build_asm_data(base.mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    thetas = [0.0, 6.0, 14.0]
    probs = []
    for th in thetas:
        pb = copy.deepcopy(base.problem)
        pb.kx = pb.k0 * math.sin(math.radians(th))
        pb.ky = 0.0
        probs.append(pb)
    fields = [[(th, 0.0, "s"), (th, 0.0, "p")] for th in thetas]
    inc = incident_set(probs, fields)
    Iu, Ju = _union_pattern(I, J, ad.height, len(probs))
    ess = np.concatenate([ad.ess + i * ad.height for i in range(len(probs))])
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init_batch([(ad, pb) for pb in probs])
        g.build_global(Iu, Ju, ess)
        g.set_incident(inc)
        b = g.assemble_rhs(eliminate=True)
        g.compute_local()
        g.scatter_eliminate(True)
        vre, vim = g.download_values()
    finally:
        g.close()
    assert b.shape == (2, len(probs) * ad.height)
    nnz = len(J)
    for i, (pb, fl) in enumerate(zip(probs, fields)):
        orc = O.Oracle(ad, pb, I, J)
        orc.set_incident(incident_set([pb], [fl]))
        for r in range(2):
            want, sc = _rhs_ref(orc, ad, r)
            got = b[r, i * ad.height:(i + 1) * ad.height]
            assert_parity(got, want, TOL, f"theta {thetas[i]} rhs {fl[r][2]}", sc)
        _, ref, sc = _oracle_system(ad, pb, I, J, None)
        sl = slice(i * nnz, (i + 1) * nnz)
        assert_parity(vre[sl] + 1j * vim[sl], ref, TOL, f"theta {thetas[i]} matrix", sc)


@pytest.mark.parametrize("p", [1, 2])
def test_eps_elem_source_equals_source_region(p):
    c = contact_hole(n_xy=4, nz_scale=0.34, side=44.0, theta=6.0, phi=30.0, sidewall_deg=84.0)
    pb = c.problem
    ad = build_asm_data(c.mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    absorber = [i for i, r in enumerate(pb.regions) if r.name == "absorber"][0]
    assert not pb.regions[absorber].source
    X = EPS_CAP
    ee = np.array([pb.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
    ee[ad.attr == absorber + 1] = X
    # the same cell with the absorber a source region of eps X over the
    # absorber background
    pb2 = copy.deepcopy(pb)
    r = pb2.regions[absorber]
    pb2.regions[absorber] = Region(X, name="absorber_src", slab=r.slab, source=True,
                                   eps_bg=pb.regions[absorber].eps)
    fields = [(6.0, 30.0, "s"), (6.0, 30.0, (0.6, 0.8j))]
    out = {}
    for key, prob, eps in (("elem", pb, ee), ("region", pb2, None)):
        inc = incident_set([prob], [fields])
        g = gpu.AsmGpuGlobal(0)
        try:
            g.init(ad, prob, eps_elem=eps)
            g.build_global(I, J, ad.ess)
            g.compute_local()
            g.scatter_eliminate(True)
            vals = g.download_values()
            g.set_incident(inc)
            out[key] = (vals, g.assemble_rhs(eliminate=True))
        finally:
            g.close()
        orc = O.Oracle(ad, prob, I, J, eps_elem=eps)
        orc.set_incident(inc)
        for k in range(len(fields)):
            want, sc = _rhs_ref(orc, ad, k)
            assert_parity(out[key][1][k], want, TOL, f"{key} rhs {k} vs oracle", sc)
        if key == "elem":
            _, _, sc_m = _oracle_system(ad, prob, I, J, eps)
            sc_r = [_rhs_ref(orc, ad, k)[1] for k in range(len(fields))]
    (ae, be), (ar, br) = out["elem"], out["region"]
    assert_parity(ae[0] + 1j * ae[1], ar[0] + 1j * ar[1], TOL, "matrix elem vs region", sc_m)
    for k in range(len(fields)):
        assert np.abs(br[k]).max() > 0
        assert_parity(be[k], br[k], TOL, f"rhs {k} elem vs region", sc_r[k])
