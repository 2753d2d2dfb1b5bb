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

Polarisation (the reference's convention, src/lithofem/incident.py:82-97,
207-225): for a wave incident from the top, k_hat = (sin t cos f,
sin t sin f, -cos t), e_s = z_hat x k_hat normalised (e_s = y_hat at normal
incidence, whatever phi), e_p = k_hat x e_s; "s" and "p" are unit amplitude
along e_s / e_p with zero phase at z_max.  A Jones vector (Ex, Ey) gives
the transverse components of the incident E, and is decomposed onto
(e_s, e_p) by solving [[e_s.x, e_p.x], [e_s.y, e_p.y]] (c_s, c_p) = (Ex, Ey).

The reference's solve.json carries the tables it computed itself, per solve
group, as a list of parts (one per pure s / p component of each plane-wave
source; src/lithofem/driver.py:28-35, incident_to_json at
src/lithofem/incident.py:243-260).  :func:`tables_from_solve_json` reads
them (summing the parts of one source), :func:`tables_to_solve_json` writes
that layout.
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


def sp_basis(theta_deg: float, phi_deg: float):
    """(e_s, e_p, k_hat) of a plane wave incident from the top."""
    th, ph = math.radians(theta_deg), math.radians(phi_deg)
    k_hat = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph),
                      -math.cos(th)])
    if theta_deg == 0.0:
        e_s = np.array([0.0, 1.0, 0.0])
    else:
        e_s = np.cross([0.0, 0.0, 1.0], k_hat)
        e_s /= np.linalg.norm(e_s)
    return e_s, np.cross(k_hat, e_s), k_hat


def sp_coefficients(theta_deg: float, phi_deg: float, pol) -> tuple:
    """(c_s, c_p) of a polarisation: "s", "p", or a Jones vector (Ex, Ey) of
    the transverse incident field, decomposed onto (e_s, e_p)."""
    if isinstance(pol, str):
        if pol == "s":
            return 1.0 + 0j, 0j
        if pol == "p":
            return 0j, 1.0 + 0j
        raise ValueError(f"polarisation {pol!r}")
    ex, ey = complex(pol[0]), complex(pol[1])
    e_s, e_p, _ = sp_basis(theta_deg, phi_deg)
    det = e_s[0] * e_p[1] - e_p[0] * e_s[1]
    return (ex * e_p[1] - e_p[0] * ey) / det, (e_s[0] * ey - e_s[1] * ex) / det


def slab_waves(problem, theta_deg: float, phi_deg: float, pol="s",
               amplitude: complex = 1.0) -> np.ndarray:
    """(nslab, 18) table [zA, zB, qA, qB, A[3], B[3]] (complex interleaved)
    of the background field for one incidence and polarisation ("s", "p"
    or a Jones vector (Ex, Ey)), scaled by a complex amplitude."""
    k0 = problem.k0
    slabs = problem.slabs
    eps = [complex(s[2]) for s in slabs]
    zs = [slabs[0][0]] + [s[1] for s in slabs[:-1]] + [slabs[-1][1]]
    # k_par = k0 sin(theta) whatever the top medium (the reference's
    # convention, src/lithofem/config.py:472-473)
    kpar = k0 * math.sin(math.radians(theta_deg))
    s_hat, _, _ = sp_basis(theta_deg, phi_deg)
    p_hat = np.cross(s_hat, [0.0, 0.0, 1.0])     # in-plane direction of k_par
    z_hat = np.array([0.0, 0.0, 1.0])
    out = np.zeros((len(slabs), 18))
    A = np.zeros((len(slabs), 3), dtype=np.complex128)
    B = np.zeros((len(slabs), 3), dtype=np.complex128)
    js, jp = sp_coefficients(theta_deg, phi_deg, pol)
    js, jp = complex(amplitude) * js, complex(amplitude) * jp
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


def _row(zA, qA, Av, zB, qB, Bv) -> np.ndarray:
    r = np.zeros(18)
    r[0], r[1] = zA, zB
    r[2], r[3], r[4], r[5] = qA.real, qA.imag, qB.real, qB.imag
    r[6:12] = np.stack([np.real(Av), np.imag(Av)], 1).ravel()
    r[12:18] = np.stack([np.real(Bv), np.imag(Bv)], 1).ravel()
    return r


def tables_from_solve_json(parts, nslab: int) -> dict:
    """``groups[g]["incident"]`` of the reference's solve.json (a list of
    parts {"source_index", "slabs": [{"z", "qA", "zA", "A", "qB", "zB",
    "B"}]}) -> {source_index: (nslab, 18) table}.  The parts of one source
    (its s and p components) are summed; they share the propagation
    constants and reference planes of every slab (one layer stack), which
    is checked."""
    def c(v):
        return complex(float(v[0]), float(v[1]))

    out = {}
    for part in parts:
        si = int(part["source_index"])
        if len(part["slabs"]) != nslab:
            raise ValueError(f"incident part of source {si}: {len(part['slabs'])} slabs, "
                             f"expected {nslab}")
        rows = np.array([_row(float(sw["zA"]), c(sw["qA"]), [c(v) for v in sw["A"]],
                              float(sw["zB"]), c(sw["qB"]), [c(v) for v in sw["B"]])
                         for sw in part["slabs"]])
        if si not in out:
            out[si] = rows
            continue
        prev = out[si]
        same = np.allclose(prev[:, :6], rows[:, :6], rtol=1e-12, atol=1e-12)
        if not same:
            raise NotImplementedError(
                f"incident parts of source {si} have different propagation constants or "
                "reference planes; one table per source needs them equal")
        prev[:, 6:] += rows[:, 6:]
    return out


def tables_to_solve_json(tables: dict, slab_bounds) -> list:
    """{source_index: (nslab, 18) table} -> the reference's
    ``groups[g]["incident"]`` layout (one part per source)."""
    def c(v):
        return [float(np.real(v)), float(np.imag(v))]

    parts = []
    for si, W in tables.items():
        slabs = []
        for j, w in enumerate(W):
            slabs.append({"z": [float(slab_bounds[j]), float(slab_bounds[j + 1])],
                          "qA": [w[2], w[3]], "zA": float(w[0]),
                          "A": [c(w[6 + 2 * i] + 1j * w[7 + 2 * i]) for i in range(3)],
                          "qB": [w[4], w[5]], "zB": float(w[1]),
                          "B": [c(w[12 + 2 * i] + 1j * w[13 + 2 * i]) for i in range(3)]})
        parts.append({"source_index": int(si), "slabs": slabs})
    return parts


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
    lists of fields, each (theta, phi, pol[, amplitude]) with pol "s", "p"
    or a Jones vector (Ex, Ey), or a precomputed (nslab, 18) table (for
    instance one read from a solve.json)."""
    nrhs = len(fields_per_problem[0])
    nslab = len(problems[0].slabs)
    waves, rs, rd, src = [], [], [], []
    for pb, flds in zip(problems, fields_per_problem):
        assert len(flds) == nrhs and len(pb.slabs) == nslab
        for f in flds:
            if isinstance(f, np.ndarray):
                assert f.shape == (nslab, 18)
                waves.append(np.asarray(f, dtype=np.float64))
            else:
                waves.append(slab_waves(pb, *f))
        for r in pb.regions:
            rs.append(r.slab)
            rd.append(r.deps)
            src.append(1 if r.source else 0)
    return IncidentSet(nrhs=nrhs, nslab=nslab, waves=np.stack(waves).reshape(-1, 18),
                       region_slab=np.array(rs, np.int32),
                       region_deps=np.array(rd, np.complex128),
                       region_src=np.array(src, np.int32))
