"""Parity of the other configurations at their full stated sizes.

* The CPU-reference contact hole (96 x 96 nm cell, square hole 24..72 nm,
  order 2, 24,192 tets, 154,656 DOF): the global CSR, the eliminated system
  and the s-polarised scattered-field RHS against the oracle's full
  reassembly, and the eps-sensitivity kernel of the adjoint configuration
  (same geometry) over every absorber element.
* The largest mesh of the batched contact-hole family (271,002 tets,
  order 2, 1.7M DOF): the eliminated system against the oracle's full
  reassembly; 64 eps variants (each absorber element's eps offset by a
  seeded uniform +-5 % real and imaginary) assembled on the GPU, the first
  and the last against the oracle on row blocks around absorber elements;
  s and p RHS.
* The order-4 L-opening (73,728 tets, 3.0M DOF, 437M nonzeros, Jones
  (1, i)/sqrt(2)): assembled in full on the GPU; the global CSR and the
  eliminated system on row blocks spread over the matrix (including rows
  of PML and absorber elements), reassembled by the oracle from their
  contributor elements only (a full reassembly takes minutes at order 4);
  the RHS of the Jones field and the eps-sensitivity kernel over every
  absorber element in full.

Metrics: the reference's max-normalised error, the per-nonzero error
against each entry's componentwise scale and the plain relative error on
well-conditioned nonzeros, each < 1e-13 (tests/cases.py); sensitivities
max-normalised < 1e-12.
"""

from __future__ import annotations

import os
import time

import numpy as np
import pytest

from cases import assert_parity, maxnorm_err
from lithofem_amd import gpu
from lithofem_amd.benchmarks import hole_family
from lithofem_amd.fespace import build_asm_data
from lithofem_amd.incident import incident_set
from lithofem_amd.problem import contact_hole, l_opening
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-13
ABSORBER = 2


def _threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _ess_mask(ad):
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    return is_ess


def _rhs_ref(orc, ad, r):
    ref = orc.rhs(r)
    ref[ad.ess] = 0.0
    sc = orc.rhs_scale(r)
    sc[ad.ess] = 0.0
    return ref, sc


def _field(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _blocks(I, starts, size, n):
    """Disjoint row blocks [r0, r0 + size) (clipped, merged when they overlap)."""
    out = []
    for r0 in sorted(int(min(max(s, 0), n - size)) for s in starts):
        if out and r0 <= out[-1][1]:
            out[-1] = (out[-1][0], min(n, max(out[-1][1], r0 + size)))
        else:
            out.append((r0, r0 + size))
    return out


def _on_blocks(a, I, blocks):
    return np.concatenate([a[I[r0]:I[r1]] for r0, r1 in blocks])


def test_contact_hole_full_size():
    t0 = time.perf_counter()
    c = contact_hole()
    pb = c.problem
    ad = build_asm_data(c.mesh, 2)
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    print(f"contact hole: ne={ad.ne} n={ad.height} nnz={len(J)} "
          f"(setup {time.perf_counter() - t0:.1f} s)", flush=True)
    inc = incident_set([pb], [[(6.0, 0.0, "s")]])
    params = np.nonzero(ad.attr == ABSORBER)[0].astype(np.int32)
    lam, u = _field(ad.height, 1), _field(ad.height, 2)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        g.set_incident(inc)
        g.compute_local()
        g.scatter_eliminate(False)
        vre, vim = g.download_values()
        g.scatter_eliminate(True)
        ere, eim = g.download_values()
        b = g.assemble_rhs(eliminate=True)
        sens = g.eps_sensitivity(params, lam, u, rhs_col=0)
    finally:
        g.close()
    nt = _threads()
    orc = O.Oracle(ad, pb, I, J)
    orc.set_incident(inc)
    rre, rim = orc.scatter_threaded(orc.elements_threaded(nt), nt)
    sc = orc.scale_csr(nt)
    assert_parity(vre + 1j * vim, rre + 1j * rim, TOL, "contact hole CSR (U5)", sc)
    is_ess = _ess_mask(ad)
    xre, xim = O.eliminate(I, J, is_ess, rre, rim)
    assert_parity(ere + 1j * eim, xre + 1j * xim, TOL, "contact hole eliminated (U6)",
                  O.eliminate_scale(I, J, is_ess, sc))
    ref, rsc = _rhs_ref(orc, ad, 0)
    assert_parity(b[0], ref, TOL, "contact hole RHS (s)", rsc)
    s_ref = orc.eps_sensitivity(lam, u, 0, elems=params, nthreads=nt)
    err = maxnorm_err(sens, s_ref)
    print(f"  sensitivity over {len(params)} absorber elements: maxnorm {err:.2e}", flush=True)
    assert err < 1e-12


def test_hole_family_largest_mesh_full_size():
    t0 = time.perf_counter()
    cells = hole_family()
    cell = cells[int(np.argmax([c.mesh.ne for c in cells]))]
    pb = cell.problem
    ad = build_asm_data(cell.mesh, 2)
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    print(f"hole family, largest mesh: ne={ad.ne} n={ad.height} nnz={len(J)} "
          f"(setup {time.perf_counter() - t0:.1f} s)", flush=True)
    assert ad.ne == 271002
    params = np.nonzero(ad.attr == ABSORBER)[0].astype(np.int32)
    eps0 = np.array([pb.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
    eps_abs = pb.regions[ABSORBER - 1].eps
    nvar = 64
    deps = np.empty((nvar, len(params)), np.complex128)
    for v in range(nvar):
        rng = np.random.default_rng(v)
        deps[v] = (eps_abs.real * 0.05 * rng.uniform(-1, 1, len(params))
                   + 1j * eps_abs.imag * 0.05 * rng.uniform(-1, 1, len(params)))
    inc = incident_set([pb], [[(6.0, 0.0, "s"), (6.0, 0.0, "p")]])
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        g.set_incident(inc)
        g.compute_local()
        g.scatter_eliminate(True)
        ere, eim = g.download_values()
        b = g.assemble_rhs(eliminate=True)
        g.variants_prepare(params)
        g.variants_assemble(deps, eliminate=True)
        var = {v: g.download_variant(v) for v in (0, nvar - 1)}
    finally:
        g.close()
    nt = _threads()
    is_ess = _ess_mask(ad)
    orc = O.Oracle(ad, pb, I, J)
    orc.set_incident(inc)
    rre, rim = orc.scatter_threaded(orc.elements_threaded(nt), nt)
    sc = orc.scale_csr(nt)
    xre, xim = O.eliminate(I, J, is_ess, rre, rim)
    del rre, rim
    assert_parity(ere + 1j * eim, xre + 1j * xim, TOL, "largest mesh eliminated (U6)",
                  O.eliminate_scale(I, J, is_ess, sc))
    del xre, xim, sc
    for r in range(2):
        ref, rsc = _rhs_ref(orc, ad, r)
        assert_parity(b[r], ref, TOL, f"largest mesh RHS {'sp'[r]}", rsc)
    # variants on row blocks around absorber elements (the rows they change)
    rng = np.random.default_rng(5)
    picks = rng.choice(params, 8, replace=False)
    blocks = _blocks(I, [int(ad.dof[e].min()) - 100 for e in picks], 400, ad.height)
    v0, v1 = (_on_blocks(var[v][1], I, blocks) for v in (0, nvar - 1))
    assert not np.array_equal(v0, v1)          # two distinct variants were downloaded
    for v, (gre, gim) in var.items():
        ee = eps0.copy()
        ee[params] += deps[v]
        ov = O.Oracle(ad, pb, I, J, eps_elem=ee)
        bre, bim, bs = ov.row_blocks(blocks, nt)
        fre, fim = O.eliminate_rows(I, J, is_ess, bre, bim, blocks)
        fs = O.eliminate_rows(I, J, is_ess, bs, np.zeros_like(bs), blocks)[0]
        assert_parity(_on_blocks(gre, I, blocks) + 1j * _on_blocks(gim, I, blocks),
                      _on_blocks(fre, I, blocks) + 1j * _on_blocks(fim, I, blocks), TOL,
                      f"variant {v} on {len(blocks)} row blocks", _on_blocks(fs, I, blocks))


def test_l_opening_order4_full_size_row_blocks():
    t0 = time.perf_counter()
    cell = l_opening()
    pb = cell.problem
    ad = build_asm_data(cell.mesh, 4)
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    n = ad.height
    print(f"L-opening p=4: ne={ad.ne} n={n} nnz={len(J)} "
          f"(setup {time.perf_counter() - t0:.1f} s)", flush=True)
    assert ad.ne == 73728 and n > 2_900_000
    r2 = 1.0 / np.sqrt(2.0)
    inc = incident_set([pb], [[(6.0, 0.0, (r2, 1j * r2))]])
    params = np.nonzero(ad.attr == ABSORBER)[0].astype(np.int32)
    lam, u = _field(n, 3), _field(n, 4)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        g.set_incident(inc)
        g.compute_local()
        g.scatter_eliminate(False)
        vre, vim = g.download_values()
        g.scatter_eliminate(True)
        ere, eim = g.download_values()
        b = g.assemble_rhs(eliminate=True)
        sens = g.eps_sensitivity(params, lam, u, rhs_col=0)
    finally:
        g.close()
    print(f"  GPU assembly downloaded ({time.perf_counter() - t0:.1f} s)", flush=True)
    pml = np.nonzero(np.asarray(pb.is_pml())[ad.attr - 1])[0]
    rng = np.random.default_rng(7)
    starts = [0, n // 4, n // 2, (3 * n) // 4, n,
              int(ad.dof[pml[len(pml) // 3]].min()), int(ad.dof[pml[-1]].max()) - 150,
              int(ad.dof[params[len(params) // 2]].min())] + list(rng.integers(0, n, 8))
    blocks = _blocks(I, starts, 300, n)
    nt = _threads()
    t1 = time.perf_counter()
    orc = O.Oracle(ad, pb, I, J)
    orc.set_incident(inc)
    bre, bim, bs = orc.row_blocks(blocks, nt)
    print(f"  oracle on {len(blocks)} row blocks ({sum(r1 - r0 for r0, r1 in blocks)} rows): "
          f"{time.perf_counter() - t1:.1f} s", flush=True)
    assert_parity(_on_blocks(vre, I, blocks) + 1j * _on_blocks(vim, I, blocks),
                  _on_blocks(bre, I, blocks) + 1j * _on_blocks(bim, I, blocks), TOL,
                  "L-opening CSR (U5) on row blocks", _on_blocks(bs, I, blocks))
    del vre, vim
    is_ess = _ess_mask(ad)
    fre, fim = O.eliminate_rows(I, J, is_ess, bre, bim, blocks)
    fs = O.eliminate_rows(I, J, is_ess, bs, np.zeros_like(bs), blocks)[0]
    assert_parity(_on_blocks(ere, I, blocks) + 1j * _on_blocks(eim, I, blocks),
                  _on_blocks(fre, I, blocks) + 1j * _on_blocks(fim, I, blocks), TOL,
                  "L-opening eliminated (U6) on row blocks", _on_blocks(fs, I, blocks))
    del ere, eim, bre, bim, bs, fre, fim, fs
    ref, rsc = _rhs_ref(orc, ad, 0)
    assert_parity(b[0], ref, TOL, "L-opening RHS, Jones (1, i)/sqrt 2", rsc)
    s_ref = orc.eps_sensitivity(lam, u, 0, elems=params, nthreads=nt)
    err = maxnorm_err(sens, s_ref)
    print(f"  sensitivity over {len(params)} absorber elements: maxnorm {err:.2e}", flush=True)
    assert err < 1e-12
