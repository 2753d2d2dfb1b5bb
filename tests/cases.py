"""Small deterministic cases shared by the CPU and GPU tests."""

from __future__ import annotations

from functools import lru_cache

import numpy as np

from lithofem_amd.fespace import build_asm_data, csr_symbolic
from lithofem_amd.problem import contact_hole, line_grating


@lru_cache(maxsize=None)
def cell(name: str, theta: float, size: str = "small"):
    if name == "grating":
        if size == "tiny":
            return line_grating(nx=3, ny=3, nz_scale=0.1, theta=theta)
        return line_grating(nx=4, ny=3, nz_scale=0.15, theta=theta)
    if name == "hole":
        if size == "tiny":
            return contact_hole(n_xy=3, nz_scale=0.25, theta=theta, phi=30.0)
        return contact_hole(n_xy=4, nz_scale=0.34, theta=theta, phi=30.0)
    raise ValueError(name)


@lru_cache(maxsize=None)
def case(name: str, p: int, theta: float, size: str = "small"):
    c = cell(name, theta, size)
    ad = build_asm_data(c.mesh, p)
    I, J = csr_symbolic(ad.dof, ad.height)
    return c, ad, I, J


def maxnorm_err(got: np.ndarray, ref: np.ndarray) -> float:
    """max_k |g_k - r_k| / max_k |r_k| with complex modulus (the reference's
    metric, solver/asm_test.cpp:258-267)."""
    return float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))


# entries below this fraction of the largest |r| are left out of the plain
# per-nonzero metric (the gate the reference's discussion of the metric uses)
PERNZ_GATE = 1e-8
# an entry is well-conditioned when its terms cancel by less than this factor
# (scale / |r| below it)
WELL_COND = 10.0


def pernz_err(got: np.ndarray, ref: np.ndarray, gate: float = PERNZ_GATE) -> float:
    """max_k |g_k - r_k| / |r_k| over the entries with |r_k| >= gate max|r|
    (relative error per nonzero)."""
    a = np.abs(ref)
    m = a >= gate * a.max()
    return float(np.max(np.abs(got[m] - ref[m]) / a[m]))


def scaled_err(got: np.ndarray, ref: np.ndarray, scale: np.ndarray) -> float:
    """max_k |g_k - r_k| / s_k, s_k the componentwise error scale of entry k
    (the sum of the moduli of its terms, oracle.scale_*): the per-nonzero
    error relative to what rounding can change in that entry, whatever its
    cancellation.  Entries with s_k = 0 must agree exactly."""
    d = np.abs(np.asarray(got) - np.asarray(ref))
    z = scale == 0
    if np.any(d[z] != 0):
        return float("inf")
    return float(np.max(d[~z] / scale[~z])) if np.any(~z) else 0.0


def assert_parity(got, ref, tol: float = 1e-13, label: str = "", scale=None) -> dict:
    """The reference's max-normalised error < tol and, given the entries'
    componentwise scales, the per-nonzero error against them < tol at every
    entry, and the plain relative error |g - r| / |r| < tol at every
    well-conditioned nonzero (scale / |r| <= WELL_COND).  The plain relative
    error at the 1e-8 gate is printed: on entries that cancel it measures
    the rounding order, not the assembly."""
    got, ref = np.asarray(got), np.asarray(ref)
    out = dict(maxnorm=maxnorm_err(got, ref), pernz_gate=pernz_err(got, ref))
    msg = f"  {label}: maxnorm {out['maxnorm']:.2e} per-nonzero(|r|>=1e-8 max) {out['pernz_gate']:.2e}"
    assert out["maxnorm"] < tol, (label, "maxnorm", out["maxnorm"])
    if scale is not None:
        scale = np.asarray(scale).reshape(ref.shape)
        out["scaled"] = scaled_err(got, ref, scale)
        a = np.abs(ref)
        wc = (a > 0) & (scale <= WELL_COND * a)
        out["wellcond"] = float(np.max(np.abs(got[wc] - ref[wc]) / a[wc])) if wc.any() else 0.0
        out["wellcond_frac"] = float(wc.sum() / max(1, (a > 0).sum()))
        msg += (f" | per-nonzero vs scale {out['scaled']:.2e}; well-conditioned "
                f"({100 * out['wellcond_frac']:.1f}% of nonzeros) {out['wellcond']:.2e}")
    print(msg, flush=True)
    if scale is not None:
        assert out["scaled"] < tol, (label, "per-nonzero vs scale", out["scaled"])
        assert out["wellcond"] < tol, (label, "well-conditioned per-nonzero", out["wellcond"])
    return out
