"""lfasm_residual_static -- the residual entry point that allocates nothing --
against lfasm_residual and the oracle's residual, through the C-ABI.

Two calls in a row share the device's static workspace: first a larger
system, then a smaller one (the smaller call must not see the larger one's
staged vectors or partial sums), then the larger one again.  Each result
matches lfasm_residual to 1e-12 relative (the norms are reduced on the
device instead of the host) and the oracle's residual formula
(solver/asm_gpu.inc:533-565) to the same bar; a consistent right-hand side
(B = A X, formed by the oracle on the host) gives a residual at rounding
level through both entry points.
"""

from __future__ import annotations

import ctypes as ct

import numpy as np
import pytest

from cases import case
from lithofem_amd import gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _assembled(name, p, size):
    c, ad, I, J = case(name, p, 6.0, size)
    g = gpu.AsmGpuGlobal(0)
    g.init(ad, c.problem)
    g.build_global(I, J, ad.ess)
    g.compute_local()
    g.scatter_eliminate(True)
    vre, vim = g.download_values()
    return g, ad, I, J, vre, vim


def _spmv(I, J, vals, x):
    rows = np.repeat(np.arange(len(I) - 1), np.diff(I))
    y = np.zeros(len(I) - 1, dtype=np.complex128)
    np.add.at(y, rows, vals * x[J])
    return y


def _check(g, ad, I, J, vre, vim, seed):
    n = ad.height
    rng = np.random.default_rng(seed)
    X = rng.standard_normal(2 * n)
    B = rng.standard_normal(2 * n)
    r_ref = g.residual(X, B)
    r_st = g.residual_static(X, B)
    r_orc = O.residual(I, J, vre, vim, X, B)
    print(f"  n={n}: residual_static {r_st:.15e} residual {r_ref:.15e} oracle {r_orc:.15e}",
          flush=True)
    assert abs(r_st - r_ref) <= 1e-12 * r_ref
    assert abs(r_st - r_orc) <= 1e-12 * r_orc
    # consistent right-hand side: B = A X
    x = X[:n] + 1j * X[n:]
    b = _spmv(I, J, vre + 1j * vim, x)
    Bc = np.concatenate([b.real, b.imag])
    r0 = g.residual_static(X, Bc)
    r1 = g.residual(X, Bc)
    print(f"  consistent B: residual_static {r0:.2e} residual {r1:.2e}", flush=True)
    assert r0 < 1e-14 and r1 < 1e-14
    return r_st


def test_residual_static_matches_residual_two_calls():
    big = _assembled("grating", 3, "small")
    small = _assembled("grating", 2, "tiny")
    try:
        assert small[1].height < big[1].height
        print(f"larger n={big[1].height}, smaller n={small[1].height}", flush=True)
        r_big = _check(*big, seed=1)
        _check(*small, seed=2)
        # the larger system again after the smaller call: same value
        rng = np.random.default_rng(1)
        n = big[1].height
        X = rng.standard_normal(2 * n)
        B = rng.standard_normal(2 * n)
        assert big[0].residual_static(X, B) == r_big
    finally:
        big[0].close()
        small[0].close()


def test_residual_static_rejects_sizes():
    g, ad, I, J, vre, vim = _assembled("grating", 1, "tiny")
    try:
        L = gpu.lib()
        x = np.zeros(2 * ad.height + 2)
        rel = ct.c_double()
        # length that does not match the system
        rc = L.lfasm_residual_static(g._h, x.size, x.ctypes.data, x.ctypes.data, ct.byref(rel))
        assert rc == gpu.E_ARG and "does not match" in g.last_error()
        # more dofs than the static workspace holds (checked before any copy)
        big = 2 * (gpu.RESIDUAL_STATIC_MAX_N + 1)
        rc = L.lfasm_residual_static(g._h, big, x.ctypes.data, x.ctypes.data, ct.byref(rel))
        assert rc == gpu.E_ARG and "exceeds the static workspace" in g.last_error()
        # the handle still works
        X = np.ones(2 * ad.height)
        assert abs(g.residual_static(X, X) - g.residual(X, X)) <= 1e-12 * g.residual(X, X)
    finally:
        g.close()
