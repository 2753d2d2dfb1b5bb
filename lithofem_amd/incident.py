"""Incident (background) field of a layered stack, in the reference's
per-slab plane-wave form.

The reference builds the background field with a transfer-matrix solve of
the layer stack (src/lithofem/incident.py, src/lithofem/tmm.py) and hands
the solver, per slab, E(z) = A exp(i qA (z - zA)) + B exp(i qB (z - zB))
(solver/problem_common.inc:11-16,218-236).  This module computes the same
representation from the slab list of a :class:`~lithofem_amd.problem.Problem`:
homogeneous layers between z_min and z_max, the bottom and top layers
extending to -inf / +inf, illumination from the top with unit amplitude,
time convention exp(-i w t), fields as envelopes (exp(i k_par . r)
factored out).  Continuity of tangential E and H at every interface is
solved directly as one small linear system per polarisation.

Polarisation (the reference's convention, src/lithofem/incident.py:82-97):
s = E along e_s = (-sin phi, cos phi, 0); p = E along e_p = k_hat x e_s
(k_hat the downward incident direction), both with unit amplitude and zero
phase at z_max.  Jones vectors combine the two linearly.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def _kz(k0: float, eps: complex, kpar: float) -> complex:
    kz = np.sqrt(complex(k0 * k0 * eps - kpar * kpar))
    if kz.imag < 0 or (kz.imag == 0 and kz.real < 0):
        kz = -kz
    return kz


def _solve_stack(k0, kpar, eps, zs, pol):
    """Amplitudes (up, down) per layer, referenced at z_ref[j].  Layers j =
    0..L-1 bottom to top; interfaces zs[1..L-1]; TE (s): field = E_s,
    flux = kz E'; TM (p): field = H_s, flux = kz/eps H'."""
    L = len(eps)
    kz = [_kz(k0, e, kpar) for e in eps]
    # bottom layer referenced at its top, the top layer at z_max (where the
    # incident wave has unit amplitude and zero phase), the others at their
    # bottom interface
    zref = [zs[1]] + [zs[j] for j in range(1, L - 1)] + ([zs[L]] if L > 1 else [])
    M = np.zeros((2 * L, 2 * L), dtype=np.complex128)
    rhs = np.zeros(2 * L, dtype=np.complex128)
    row = 0
    for j in range(1, L):
        z = zs[j]
        for jj, sgn in ((j - 1, 1.0), (j, -1.0)):
            ep = np.exp(1j * kz[jj] * (z - zref[jj]))
            em = np.exp(-1j * kz[jj] * (z - zref[jj]))
            fac = kz[jj] if pol == "s" else kz[jj] / eps[jj]
            M[row, 2 * jj] += sgn * ep
            M[row, 2 * jj + 1] += sgn * em
            M[row + 1, 2 * jj] += sgn * fac * ep
            M[row + 1, 2 * jj + 1] -= sgn * fac * em
        row += 2
    M[row, 0] = 1.0                     # nothing incoming from below
    rhs[row] = 0.0
    M[row + 1, 2 * (L - 1) + 1] = 1.0   # unit incident amplitude from the top
    rhs[row + 1] = 1.0 if pol == "s" else np.sqrt(complex(eps[-1]))
    amp = np.linalg.solve(M, rhs).reshape(L, 2)
    return kz, zref, amp


def slab_waves(problem, theta_deg: float, phi_deg: float, jones=(1.0, 0.0),
               wavelength: float | None = None) -> np.ndarray:
    """(nslab, 18) table [zA, zB, qA, qB, A[3], B[3]] (complex interleaved)
    of the background field for one incidence and Jones vector (s, p)."""
    k0 = problem.k0
    slabs = problem.slabs
    eps = [complex(s[2]) for s in slabs]
    zs = [slabs[0][0]] + [s[1] for s in slabs[:-1]] + [slabs[-1][1]]
    # k_par = k0 sin(theta) whatever the top medium (the reference's
    # convention, src/lithofem/config.py:472-473)
    kpar = k0 * math.sin(math.radians(theta_deg))
    cphi, sphi = math.cos(math.radians(phi_deg)), math.sin(math.radians(phi_deg))
    s_hat = np.array([-sphi, cphi, 0.0])
    p_hat = np.array([cphi, sphi, 0.0])
    z_hat = np.array([0.0, 0.0, 1.0])
    out = np.zeros((len(slabs), 18))
    A = np.zeros((len(slabs), 3), dtype=np.complex128)
    B = np.zeros((len(slabs), 3), dtype=np.complex128)
    js, jp = complex(jones[0]), complex(jones[1])
    kz, zref, amp_s = _solve_stack(k0, kpar, eps, zs, "s")
    _, _, amp_p = _solve_stack(k0, kpar, eps, zs, "p")
    for j in range(len(slabs)):
        e = eps[j]
        # E of the p waves (H along -e_s for the downward wave), so that the
        # incident wave in vacuum is exactly e_p = cos(theta) p_hat + sin(theta) z_hat
        up_p = -(kz[j] / (k0 * e)) * p_hat + (kpar / (k0 * e)) * z_hat
        dn_p = (kz[j] / (k0 * e)) * p_hat + (kpar / (k0 * e)) * z_hat
        A[j] = js * amp_s[j, 0] * s_hat + jp * amp_p[j, 0] * up_p
        B[j] = js * amp_s[j, 1] * s_hat + jp * amp_p[j, 1] * dn_p
        out[j, 0] = zref[j]
        out[j, 1] = zref[j]
        out[j, 2], out[j, 3] = kz[j].real, kz[j].imag
        out[j, 4], out[j, 5] = -kz[j].real, -kz[j].imag
        out[j, 6:12] = np.stack([A[j].real, A[j].imag], 1).ravel()
        out[j, 12:18] = np.stack([B[j].real, B[j].imag], 1).ravel()
    return out


def evaluate(waves: np.ndarray, slab: int, z: np.ndarray) -> np.ndarray:
    """E(z) (n, 3) of one slab row of a waves table."""
    w = waves[slab]
    qA, qB = w[2] + 1j * w[3], w[4] + 1j * w[5]
    A = w[6:12:2] + 1j * w[7:12:2]
    B = w[12:18:2] + 1j * w[13:18:2]
    z = np.atleast_1d(z)
    return (np.exp(1j * qA * (z - w[0]))[:, None] * A[None, :]
            + np.exp(1j * qB * (z - w[1]))[:, None] * B[None, :])


@dataclass
class IncidentSet:
    """Input of lfasm_set_incident: waves [problem][rhs][slab][18] and the
    region tables (concatenated over problems like the eps tables)."""
    nrhs: int
    nslab: int
    waves: np.ndarray
    region_slab: np.ndarray
    region_deps: np.ndarray
    region_src: np.ndarray


def incident_set(problems, fields_per_problem) -> IncidentSet:
    """problems: list of Problem; fields_per_problem: list (per problem) of
    lists of (theta, phi, jones)."""
    nrhs = len(fields_per_problem[0])
    nslab = len(problems[0].slabs)
    waves, rs, rd, src = [], [], [], []
    for pb, flds in zip(problems, fields_per_problem):
        assert len(flds) == nrhs and len(pb.slabs) == nslab
        for (th, ph, jv) in flds:
            waves.append(slab_waves(pb, th, ph, jv))
        for r in pb.regions:
            rs.append(r.slab)
            rd.append(r.deps)
            src.append(1 if r.source else 0)
    return IncidentSet(nrhs=nrhs, nslab=nslab, waves=np.stack(waves).reshape(-1, 18),
                       region_slab=np.array(rs, np.int32),
                       region_deps=np.array(rd, np.complex128),
                       region_src=np.array(src, np.int32))
