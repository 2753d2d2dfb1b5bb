"""The identity behind the angle sweeps (lfasm_angles_assemble), pinned on
the oracle's element matrices: for one mesh and wavelength the k_par terms
are polynomial in (kx, ky) -- CROSS0 + CROSS1 linear, KLK quadratic -- so

    A(kx, ky) = B0 + kx Bx + ky By + kx^2 Bxx + ky^2 Byy + kx ky (B11 - Bxx - Byy)

with B0 the curl-curl + eps-mass matrix, Bx / By the cross terms at
k_par = (1, 0) / (0, 1), Bxx / Byy / B11 the KLK term at (1, 0) / (0, 1) /
(1, 1).  Checked on PML and non-PML elements (Lambda != I and = I), with and
without the face transform, at several k_par including k_par = 0 (where the
reference switches the k_par integrators off).  A dropped KLK cross term, a
sign error in the cross terms or a wrong coefficient fails it."""

from __future__ import annotations

import copy

import numpy as np
import pytest

from lithofem_amd.fespace import build_asm_data
from lithofem_amd.problem import contact_hole
from oracle import oracle as O


def _at(pb, kx, ky):
    q = copy.deepcopy(pb)
    q.kx, q.ky = kx, ky
    return q


@pytest.mark.parametrize("p", [1, 2, 3])
def test_element_matrices_are_polynomial_in_kpar(p):
    c = contact_hole(n_xy=3, nz_scale=0.3, side=40.0, theta=6.0, phi=30.0, sidewall_deg=84.0)
    pb = c.problem
    ad = build_asm_data(c.mesh, p)
    # a few elements of every class: PML (both sides) and not
    pml = np.asarray(pb.is_pml(), dtype=bool)
    elems = []
    for want in (True, False):
        sel = [e for e in range(ad.ne) if bool(pml[ad.attr[e] - 1]) == want]
        elems += sel[:: max(1, len(sel) // 6)][:6]
    assert len(elems) >= 8
    cm = O.CURLCURL | O.MASS
    cr = O.CROSS0 | O.CROSS1
    k0 = pb.k0
    for te in (True, False):
        def mats(q, mask):
            orc = O.Oracle(ad, q)
            return [orc.element(e, mask, te) for e in elems]
        B0 = mats(_at(pb, 0.0, 0.0), cm)
        Bx = mats(_at(pb, 1.0, 0.0), cr)
        By = mats(_at(pb, 0.0, 1.0), cr)
        Bxx = mats(_at(pb, 1.0, 0.0), O.KLK)
        Byy = mats(_at(pb, 0.0, 1.0), O.KLK)
        B11 = mats(_at(pb, 1.0, 1.0), O.KLK)
        for kx, ky in ((0.0, 0.0), (0.07 * k0, 0.0), (0.1 * k0, 0.06 * k0), (-0.2 * k0, 0.15 * k0)):
            direct = mats(_at(pb, kx, ky), O.ALL)
            worst = 0.0
            for i in range(len(elems)):
                comb = (B0[i] + kx * Bx[i] + ky * By[i] + kx * kx * Bxx[i] + ky * ky * Byy[i]
                        + kx * ky * (B11[i] - Bxx[i] - Byy[i]))
                worst = max(worst, float(np.max(np.abs(comb - direct[i])) / np.max(np.abs(direct[i]))))
            print(f"p={p} te={te} k_par=({kx:.3f}, {ky:.3f}): max-normalised {worst:.2e}")
            assert worst < 1e-13, (kx, ky, worst)
