"""Captured assembly step (HIP graph) against the eager calls.

The graph holds the whole step (reference-tensor kernel, quadrature kernel
of the PML elements, scatter + elimination).  Every replay is checked bit
for bit against the eager result, with other eager work and host
synchronisation between replays;
a per-element eps change drops the graph and the next launch re-captures
it; asynchronous back-to-back launches end in the same values.
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import case
from lithofem_amd import gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [1, 2], ids=["tensor", "fused"])
@pytest.mark.parametrize("p", [2, 3])
def test_captured_step_replays_equal_eager(p, mode):
    c, ad, I, J = case("grating", p, 6.0, "small")
    print(f"graph p={p} tensor mode {mode} ne={ad.ne} nnz={len(J)}", flush=True)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, c.problem)
        g.build_global(I, J, ad.ess)
        g.set_tensor_path(mode)
        with pytest.raises(gpu.AsmGpuError):
            g.step_launch()                      # nothing captured yet
        g.compute_local()
        g.scatter_eliminate(True)
        ref = g.download_values()
        g.step_capture()
        for r in range(4):
            # different eager work between the replays
            g.compute_local(gpu.CURLCURL, gpu.MASS, False)
            g.scatter_eliminate(False)
            g.step_launch(sync=True)
            got = g.download_values()
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), r
        assert g.timings()["step_ms"] > 0
        # a per-element eps change drops the graph; the launch re-captures
        rng = np.random.default_rng(2)
        eps = np.array([c.problem.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
        eps *= 1.0 + 0.02 * rng.uniform(-1, 1, ad.ne)
        g.set_eps_elem(eps)
        g.step_launch(sync=True)
        got = g.download_values()
        g.compute_local()
        g.scatter_eliminate(True)
        eager = g.download_values()
        assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1])
        assert not np.array_equal(eager[0], ref[0])
        # asynchronous back-to-back launches
        for _ in range(3):
            g.step_launch(sync=False)
        got = g.download_values()
        assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1])
    finally:
        g.close()
