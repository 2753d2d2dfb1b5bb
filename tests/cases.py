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
