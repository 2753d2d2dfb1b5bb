"""The launch autotuner's choices on a stand-in handle whose timings are a
known function of the launch settings: every candidate is timed, the fastest
is kept and left set on the handle (CPU only; the GPU timings themselves are
exercised by the benchmark)."""

from __future__ import annotations

from lithofem_amd import gpu
from lithofem_amd import tune


class FakeHandle:
    """Records the settings and reports synthetic timings for them."""

    def __init__(self, nd=45, tensor=True):
        self.nd = nd
        self.tensor = tensor
        self.kernel = 0
        self.conc = (0, 0)
        self.res = 0
        self.local_kernel = 0
        self.captures = 0
        self.last = {}

    # element-matrix launches
    def compute_local(self, mask_cc, mask_mass, te, kernel=0):
        k = kernel or self.local_kernel
        quad = 1.0 + 0.1 * abs(k - 7)          # kernel 7 fastest
        grid = self.conc[0]
        side = {0: 1.8, 64: 1.6, 96: 1.5, 128: 1.4, 160: 1.45, 224: 1.7}.get(grid, 1.9)
        self.last = {"local_tensor_ms": 0.5 if self.tensor else 0.0,
                     "local_quad_ms": quad,
                     "local_ms": (side if grid else quad + (0.5 if self.tensor else 0.0))}

    def timings(self):
        return dict(self.last)

    def set_local_kernel(self, k):
        self.local_kernel = k

    def set_element_concurrency(self, grid, scatter_overlap=0):
        self.conc = (grid, scatter_overlap)

    def set_quad_residency(self, k):
        self.res = k

    # captured step
    def step_capture(self, *a):
        self.captures += 1

    def step_launch(self, sync=True):
        grid, sg = self.conc
        t = 1.7 - 0.05 * sg - (0.04 if self.res == 2 else 0.02 if self.res == 3 else 0.0)
        t -= 0.01 if grid == 4 * 128 else 0.0
        self.last = {"step_ms": t}


def test_tune_local_picks_fastest_kernel_and_concurrency():
    g = FakeHandle()
    out = tune.tune_local(g, reps=2)
    assert out["best"] == gpu.kernel_mfma(5) == 7 and g.local_kernel == 7
    assert set(out["ms"]) == set(tune.candidates(45))
    assert out["concurrency"]["best"] == 128 and g.conc[0] == 128


def test_tune_local_without_tensor_path_skips_concurrency():
    g = FakeHandle(tensor=False)
    out = tune.tune_local(g, reps=1)
    assert "concurrency" not in out


def test_candidates_by_order():
    assert tune.candidates(84)[0] == gpu.kernel_mfma(0)       # no VALU kernel above nd 48
    assert tune.candidates(45)[0] == gpu.KERNEL_VALU
    assert len(tune.candidates(45)) == gpu.N_MFMA_VARIANTS + 1


def test_tune_step_times_every_triple_and_keeps_the_fastest():
    g = FakeHandle()
    out = tune.tune_step(g, 128, reps=2)
    grids = (128, 256, 512, 1024, 1 << 20)
    expect = {(128, 0, r) for r in tune.QUAD_RESIDENCIES} | \
             {(gr, 1, r) for gr in grids for r in tune.QUAD_RESIDENCIES}
    assert set(out["ms"]) == expect
    assert out["best"] == (512, 1, 2)
    assert g.conc == (512, 1) and g.res == 2
    assert g.captures == len(expect)


def test_tune_step_in_sequence_only():
    g = FakeHandle()
    out = tune.tune_step(g, 0, reps=1)
    assert set(out["ms"]) == {(0, 0, r) for r in tune.QUAD_RESIDENCIES} | \
        {(0, 1, r) for r in tune.QUAD_RESIDENCIES}
