"""Adjoint eps-gradient through the GPU assembly vs central finite differences.

Forward: per-element complex eps on every absorber and hole element of a
small contact-hole cell (p = 2, oblique incidence, s polarisation) ->
GPU matrix + RHS assembly -> host LU solve.  Objective: reflected +-1
order efficiencies + mean squared deviation of |E|^2 on the z = -1.6 nm
slice from a seeded random target.  The gradient comes from
torch.autograd through EpsToField (transpose solve + eps-sensitivity
kernel); central differences use a relative step of 1e-6 on 32 random
elements in the real and the imaginary direction, each within 1e-6 of the
adjoint gradient (the solves of the difference quotients take 3 steps of
iterative refinement, so their rounding stays far below the step).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from lithofem_amd.adjoint import DesignObjective, EpsForwardModel, make_torch_function
from lithofem_amd.fespace import build_asm_data, csr_symbolic
from lithofem_amd.incident import evaluate, incident_set
from lithofem_amd.problem import contact_hole

pytestmark = pytest.mark.gpu


def _setup(p=2):
    c = contact_hole(n_xy=4, nz_scale=0.3, side=44.0, theta=6.0, phi=0.0)
    pb = c.problem
    ad = build_asm_data(c.mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    inc = incident_set([pb], [[(6.0, 0.0, "s")]])
    params = np.nonzero((ad.attr == 2) | (ad.attr == 3))[0].astype(np.int32)
    model = EpsForwardModel(ad, pb, I, J, params, incident=inc, rhs_col=0, refine=3)
    rng = np.random.default_rng(3)
    ng = 8
    target = rng.uniform(0, 1, (ng, ng))
    from lithofem_amd.adjoint import plane_points
    pts = plane_points(96.0, 96.0, -1.6, ng, ng)
    e_inc = evaluate(inc.waves, 0, pts[:, 2])
    obj = DesignObjective(ad, pb, 96.0, 96.0, 115.0, -1.6, ng, target, e_inc, n_orders_grid=8)
    eps0 = model.eps0[params].copy()
    return model, obj, eps0, params, ad


def _J(model, obj, eps):
    u = torch.from_numpy(model.solve(eps))
    return float(obj(u))


def test_adjoint_gradient_matches_central_differences():
    model, obj, eps0, params, ad = _setup()
    try:
        F = make_torch_function()
        eps = torch.tensor(eps0, dtype=torch.complex128, requires_grad=True)
        u = F.apply(eps, model)
        J = obj(u)
        J.backward()
        grad = eps.grad.numpy()
        print(f"J = {float(J.detach()):.6e}; nparam = {len(params)}; |grad| = {np.abs(grad).max():.3e}",
              flush=True)
        rng = np.random.default_rng(4)
        picks = rng.choice(len(params), 32, replace=False)
        worst = 0.0
        for i in picks:
            for dirn in (1.0, 1j):
                h = 1e-6 * abs(eps0[i])
                ep = eps0.copy()
                ep[i] += h * dirn
                em = eps0.copy()
                em[i] -= h * dirn
                fd = (_J(model, obj, ep) - _J(model, obj, em)) / (2 * h)
                # torch's complex gradient is 2 dL/d(conj eps): dL = Re(conj(grad) d eps)
                an = (np.conj(grad[i]) * dirn).real
                rel = abs(fd - an) / max(abs(an), 1e-30)
                print(f"  elem {params[i]} dir {dirn}: adjoint {an:.9e} fd {fd:.9e} rel {rel:.2e}",
                      flush=True)
                worst = max(worst, rel)
        print(f"worst relative difference over {len(picks)} elements x 2 directions: {worst:.2e}",
              flush=True)
        assert worst < 1e-6, worst
    finally:
        model.close()


def test_region_sum_matches_two_solve_difference():
    """The reference has only region eps, so the quantity it could check is
    the region sum: sum of g_e over one region against the central
    difference of J when that whole region's eps moves (real and imaginary
    directions)."""
    model, obj, eps0, params, ad = _setup()
    try:
        F = make_torch_function()
        eps = torch.tensor(eps0, dtype=torch.complex128, requires_grad=True)
        J = obj(F.apply(eps, model))
        J.backward()
        grad = eps.grad.numpy()
        for attr in (2, 3):
            sel = ad.attr[params] == attr
            assert sel.any()
            for dirn in (1.0, 1j):
                h = 1e-6 * abs(eps0[sel]).max()
                d = np.where(sel, h * dirn, 0.0)
                fd = (_J(model, obj, eps0 + d) - _J(model, obj, eps0 - d)) / (2 * h)
                an = (np.conj(grad[sel]) * dirn).real.sum()
                rel = abs(fd - an) / max(abs(an), 1e-30)
                print(f"  region {attr} dir {dirn}: sum g_e {an:.9e} two-solve {fd:.9e} rel {rel:.2e}",
                      flush=True)
                assert rel < 1e-6, (attr, dirn, an, fd)
    finally:
        model.close()
