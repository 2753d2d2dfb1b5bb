"""GPU parity on the order-4 L-shaped opening (192 x 192 nm cell, 70 nm
absorber with 80 deg sidewalls, z-PML, theta = 6 deg) with the Jones
vector (1, i)/sqrt(2), against the oracle:

* element matrices (full sum + face transform) of a sample covering every
  region, including PML elements (quadrature kernel) and constant-
  coefficient elements (reference-tensor kernel);
* the CSR scatter bit-identical to the oracle's scatter of the same
  element matrices, and the eliminated system;
* the scattered-field RHS of (1, i)/sqrt(2), s and p;
* the eps-sensitivity kernel over every absorber element.
Order 4 (84 DOF per element) uses the 6-tile MFMA kernels; the fused
tensor-row path does not cover it (nd > 64) and must fall back.
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import assert_parity, maxnorm_err
from lithofem_amd import gpu
from lithofem_amd.fespace import build_asm_data, csr_symbolic
from lithofem_amd.incident import incident_set, sp_coefficients
from lithofem_amd.problem import l_opening
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-13


def test_l_opening_p4_jones():
    c = l_opening(n_xy=3, nz_scale=0.2)
    pb = c.problem
    ad = build_asm_data(c.mesh, 4)
    I, J = csr_symbolic(ad.dof, ad.height)
    print(f"l_opening p=4: ne={ad.ne} n={ad.height} nnz={len(J)}", flush=True)
    r2 = 1.0 / np.sqrt(2.0)
    # Jones (1, i)/sqrt(2): transverse (Ex, Ey) of the incident field
    fields = [(6.0, 0.0, (r2, 1j * r2)), (6.0, 0.0, "s"), (6.0, 0.0, "p")]
    inc = incident_set([pb], [fields])
    orc = O.Oracle(ad, pb, I, J)
    orc.set_incident(inc)
    sample = []
    for a in np.unique(ad.attr):
        idx = np.nonzero(ad.attr == a)[0]
        sample.extend(idx[np.linspace(0, len(idx) - 1, min(4, len(idx))).astype(int)])
    sample = sorted(set(int(e) for e in sample))
    ref = np.stack([orc.element(e, O.ALL, True) for e in sample])
    ref_sc = orc.scale_elements(sample, O.ALL, True)
    params = np.nonzero(ad.attr == 2)[0].astype(np.int32)
    rng = np.random.default_rng(8)
    lam = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
    u = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        nconst = g.set_tensor_path(1)
        assert 0 < nconst < ad.ne
        g.compute_local()
        Ae = g.download_local()
        assert_parity(Ae[sample], ref, TOL, f"element matrices ({len(sample)} sampled)", ref_sc)
        g.scatter_eliminate(False)
        vre, vim = g.download_values()
        sre, sim = orc.scatter(Ae)
        assert np.array_equal(vre, sre) and np.array_equal(vim, sim)
        g.scatter_eliminate(True)
        ere, eim = g.download_values()
        is_ess = np.zeros(ad.height, np.uint8)
        is_ess[ad.ess] = 1
        xre, xim = O.eliminate(I, J, is_ess, sre, sim)
        assert np.array_equal(ere, xre) and np.array_equal(eim, xim)
        # the fused path does not cover 84 DOF per element: same values
        g.set_tensor_path(2)
        g.compute_local()
        g.scatter_eliminate(True)
        fre, fim = g.download_values()
        assert np.array_equal(fre, ere) and np.array_equal(fim, eim)
        # RHS of the three fields
        g.set_incident(inc)
        b = g.assemble_rhs(eliminate=True)
        for r in range(len(fields)):
            want = orc.rhs(r)
            want[ad.ess] = 0.0
            sc = orc.rhs_scale(r)
            sc[ad.ess] = 0.0
            assert_parity(b[r], want, TOL, f"rhs {fields[r][2]}", sc)
        cs, cp = sp_coefficients(6.0, 0.0, (r2, 1j * r2))
        assert maxnorm_err(b[0], cs * b[1] + cp * b[2]) < TOL
        # eps sensitivity over the absorber elements
        s = g.eps_sensitivity(params, lam, u, rhs_col=0)
    finally:
        g.close()
    want = orc.eps_sensitivity(lam, u, 0)[params]
    e_s = maxnorm_err(s, want)
    print(f"  sensitivity over {len(params)} elements: {e_s:.2e}", flush=True)
    assert e_s < 1e-12, e_s
