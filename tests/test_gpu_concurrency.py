"""Element kernels side by side (lfasm_set_element_concurrency): the
reference-tensor kernel on the handle's stream, held to a workgroup cap,
and the quadrature kernel on a second stream give exactly the values of
the two kernels in sequence, eager and as captured-step replays; so does
the captured step that scatters the rows formed only by tensor-path
elements beside the quadrature kernel, and the quadrature kernel held to a
few workgroups per CU (lfasm_set_quad_residency)."""

from __future__ import annotations

import numpy as np
import pytest

from cases import case
from lithofem_amd import gpu

pytestmark = pytest.mark.gpu


def _assemble(g, conc, graph, scat=0):
    g.set_element_concurrency(conc, scat)
    if graph:
        g.step_capture(gpu.CURLCURL, gpu.MASS_ALL, True, True)
        outs = []
        for _ in range(3):
            g.step_launch(sync=True)
            outs.append(g.download_values())
        for a, b in zip(outs, outs[1:]):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        return outs[-1]
    g.compute_local(gpu.CURLCURL, gpu.MASS_ALL, True)
    t = g.timings()
    assert t["local_tensor_ms"] > 0 and t["local_quad_ms"] > 0
    g.scatter_eliminate(True)
    return g.download_values()


@pytest.mark.parametrize("conc", [1, 7, 64])
def test_concurrent_element_kernels_match_sequence(conc):
    c, ad, I, J = case("grating", 3, 6.0, "small")
    print(f"concurrency cap {conc}: ne={ad.ne} nnz={len(J)}", flush=True)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, c.problem)
        g.build_global(I, J, ad.ess)
        assert g.set_tensor_path(1) > 0
        ref = _assemble(g, 0, False)
        for graph in (False, True):
            got = _assemble(g, conc, graph)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        # the captured step with the tensor-only rows scattered beside the
        # quadrature kernel (two row lists, every row formed once)
        for scat in (1,):
            got = _assemble(g, conc, True, scat)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        # the quadrature kernel held to 3, 2 and 1 workgroups per CU beside
        # the other kernels, eager and in both captured forms
        for res in (3, 2, 1):
            g.set_quad_residency(res)
            for graph, scat in ((False, 0), (True, 0), (True, 1)):
                got = _assemble(g, conc, graph, scat)
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        g.set_quad_residency(0)
        with pytest.raises(gpu.AsmGpuError):
            g.set_quad_residency(9)
        # back to the sequence, captured
        got = _assemble(g, 0, True)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    finally:
        g.close()


@pytest.mark.parametrize("p", [2, 4])
def test_quad_residency_general_kernels(p):
    # the general MFMA kernels (orders 2 and 4) under a residency cap beside
    # the tensor kernel give the values of the sequence
    c, ad, I, J = case("hole", p, 6.0, "small")
    print(f"residency, order {p}: ne={ad.ne} nnz={len(J)}", flush=True)
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, c.problem)
        g.build_global(I, J, ad.ess)
        assert g.set_tensor_path(1) > 0
        ref = _assemble(g, 0, False)
        for res in (0, 2):
            g.set_quad_residency(res)
            for graph, scat in ((False, 0), (True, 1)):
                got = _assemble(g, 16, graph, scat)
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    finally:
        g.close()
