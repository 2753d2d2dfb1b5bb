"""End-to-end check of the HIP assembly: manufactured solution on the unit
cube (curl curl E - k0^2 eps E = f, non-homogeneous PEC) solved with the
matrix the library assembles; L2 convergence rate >= p - 0.2, complex
symmetry < 1e-12 and solve residual < 1e-10, the criteria of the
reference's M4 test (tests/test_mms.py:1-8).  Both element paths of the
library are used: reference tensors (every element here has constant
coefficients) and the quadrature kernels.
"""

from __future__ import annotations

import pytest

import mms
from lithofem_amd import gpu

pytestmark = pytest.mark.gpu


def _assembler(mode):
    def run(ad, pb, I, J):
        g = gpu.AsmGpuGlobal(0)
        try:
            g.init(ad, pb)
            g.build_global(I, J, ad.ess)
            n = g.set_tensor_path(mode)
            assert (n == ad.ne) == (mode == 1)
            g.compute_local(gpu.CURLCURL, gpu.MASS, True)
            g.scatter_eliminate(False)
            return g.download_values()
        finally:
            g.close()
    return run


@pytest.mark.parametrize("mode", [1, 0], ids=["tensor", "quadrature"])
@pytest.mark.parametrize("p,ns", [(1, (4, 8)), (2, (2, 4)), (3, (2, 4)), (4, (3, 5))],
                         ids=["p1", "p2", "p3", "p4"])
def test_mms_convergence_gpu(p, ns, mode):
    print(f"mms p={p} levels {ns} tensor mode {mode}", flush=True)
    rate, levels = mms.convergence(p, 2.25 + 0.75j, ns, _assembler(mode))
    print(f"  rate {rate:.3f}, L2 errors {[f'{e:.3e}' for _, e, _, _ in levels]}", flush=True)
    assert rate >= p - 0.2, rate
    for h, err, sym, res in levels:
        assert sym < 1e-12, sym
        assert res < 1e-10, res
