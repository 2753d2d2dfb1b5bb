"""End-to-end parity: near fields and diffraction orders of a solve of the
HIP-assembled system against a solve of the oracle-assembled one.

The reference compares its GPU-assembled and MFEM-assembled solves on the
sampled planes (max-normalised < 1e-10, tests/test_asm_v25.py:199-219) and
on the order efficiencies (absolute < 1e-8, tests/test_asm_v25.py:274-290).
Here both eliminated systems (matrix and scattered-field RHS) are factorised
by the same host direct solver (SuperLU standing in for the reference's
UMFPACK), the solutions are sampled on cell-centred planes with the
reference sampler's nudge, and the reflected order efficiencies come from
the plane above the structure.
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import cell, maxnorm_err
from lithofem_amd import gpu
from lithofem_amd.adjoint import HostLU, order_efficiencies, plane_points, sampling_matrix
from lithofem_amd.fespace import build_asm_data, csr_symbolic
from lithofem_amd.incident import incident_set
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PLANE_TOL = 1e-10
ORDER_TOL = 1e-8

# (cell, order, field, cell size, z of the plane above the structure, z of a
# plane through it)
CASES = [("grating", 2, (6.0, 0.0, "p"), (64.0, 8.0), 119.0, 50.0),
         ("grating", 3, (6.0, 0.0, "p"), (64.0, 8.0), 119.0, 30.0),
         ("hole", 2, (6.0, 30.0, "s"), (96.0, 96.0), 100.0, 30.0),
         ("hole", 3, (6.0, 30.0, (0.6, 0.8j)), (96.0, 96.0), 100.0, -1.6)]


@pytest.mark.parametrize("name,p,field,lxy,z_top,z_mid", CASES,
                         ids=[f"{c[0]}-p{c[1]}" for c in CASES])
def test_e2e_planes_and_orders(name, p, field, lxy, z_top, z_mid):
    import torch
    c = cell(name, 6.0, "small")
    pb = c.problem
    ad = build_asm_data(c.mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    inc = incident_set([pb], [[field]])
    print(f"e2e {name} p={p}: ne={ad.ne} n={ad.height} nnz={len(J)}", flush=True)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        g.compute_local()
        g.scatter_eliminate(True)
        vre, vim = g.download_values()
        g.set_incident(inc)
        b_g = g.assemble_rhs(eliminate=True)[0]
    finally:
        g.close()
    orc = O.Oracle(ad, pb, I, J)
    rre, rim = orc.assemble()
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    xre, xim = O.eliminate(I, J, is_ess, rre, rim)
    orc.set_incident(inc)
    b_o = orc.rhs(0)
    b_o[ad.ess] = 0.0
    u_g = HostLU(I, J, vre + 1j * vim, ad.height).solve(b_g)
    u_o = HostLU(I, J, xre + 1j * xim, ad.height).solve(b_o)
    print(f"  solutions: maxnorm {maxnorm_err(u_g, u_o):.2e}", flush=True)
    lx, ly = lxy
    n_o = 16
    for z in (z_top, z_mid):
        S = sampling_matrix(ad, plane_points(lx, ly, z, n_o, n_o), periodic=lxy)
        eg, eo = S @ u_g, S @ u_o
        err = maxnorm_err(eg, eo)
        print(f"  plane z={z}: maxnorm {err:.2e}", flush=True)
        assert err < PLANE_TOL, (z, err)
    S = sampling_matrix(ad, plane_points(lx, ly, z_top, n_o, n_o), periodic=lxy)
    orders = [(m, n) for m in (-1, 0, 1) for n in (-1, 0, 1)]
    effs = []
    for u in (u_g, u_o):
        E = torch.as_tensor((S @ u).reshape(n_o, n_o, 3))
        effs.append(np.array([float(v) for v in
                              order_efficiencies(E, (pb.kx, pb.ky), pb.k0, lx, ly, orders)]))
    worst = float(np.abs(effs[0] - effs[1]).max())
    print(f"  orders {orders}: efficiencies {np.round(effs[1], 6).tolist()}, "
          f"max abs diff {worst:.2e}", flush=True)
    assert effs[1].max() > 1e-8
    assert worst < ORDER_TOL, worst
