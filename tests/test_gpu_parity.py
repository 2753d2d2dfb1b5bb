"""GPU parity of the HIP assembly against the CPU oracle, through the C-ABI.

Mirrors the reference's assembly harness (solver/asm_test.cpp, wrapped by
tests/test_asm_v25.py):
  U3  per-integrator element matrices (no face transform), p = 1..4, k_par on/off
  U4  full element matrix + dual face-DOF transform, every element
  U5  global CSR values (sorted-column pattern), no elimination
  U6  eliminated system (essential rows/cols, unit diagonal)
plus what the reference's atomics could not give: bitwise run-to-run
determinism, and a scatter that is bit-identical to the oracle's
element-order accumulation given the same element matrices.

Criterion (tests/cases.py:assert_parity): max-normalised complex error
< 1e-13 (tests/test_asm_v25.py:130-133,171-175); per nonzero, the error
against the entry's componentwise scale (the moduli of its terms, from the
oracle) < 1e-13 at every entry, and the plain relative error < 1e-13 at
every well-conditioned nonzero.  Every kernel variant (VALU and all MFMA launch variants)
is checked.
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import assert_parity, case
from lithofem_amd import gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-13
KERNELS = [gpu.KERNEL_AUTO, gpu.KERNEL_VALU] + [gpu.kernel_mfma(v) for v in range(gpu.N_MFMA_VARIANTS)]


def _sample(ad, n_per_attr=6):
    out = []
    for a in np.unique(ad.attr):
        idx = np.nonzero(ad.attr == a)[0]
        out.extend(idx[np.linspace(0, len(idx) - 1, min(n_per_attr, len(idx))).astype(int)])
    return sorted(set(int(e) for e in out))


def _ctx(c, ad, I=None, J=None, eps_elem=None):
    g = gpu.AsmGpuGlobal(0)
    g.init(ad, c.problem, eps_elem=eps_elem)
    if I is not None:
        g.build_global(I, J, ad.ess)
    return g


U3 = [("curlcurl", gpu.CURLCURL, 0, O.CURLCURL, False),
      ("mass", 0, gpu.MASS, O.MASS, False),
      ("cross0", 0, gpu.CROSS0, O.CROSS0, True),
      ("cross1", 0, gpu.CROSS1, O.CROSS1, True),
      ("cross01", 0, gpu.CROSS0 | gpu.CROSS1, O.CROSS0 | O.CROSS1, True),
      ("klk", 0, gpu.KLK, O.KLK, True)]

# (tensor path, quadrature kernels): with the tensor path on, constant-
# coefficient elements take the reference-tensor kernel and the PML
# elements the selected quadrature kernel; off, every element takes it
PATHS = [(True, [gpu.KERNEL_AUTO, gpu.KERNEL_VALU]), (False, KERNELS)]


@pytest.mark.parametrize("theta", [0.0, 6.0], ids=["t0", "t6"])
@pytest.mark.parametrize("p", [1, 2, 3, 4], ids=["p1", "p2", "p3", "p4"])
@pytest.mark.parametrize("name", ["grating", "hole"])
def test_u3_u4_local(name, p, theta):
    size = "tiny"
    c, ad, I, J = case(name, p, theta, size)
    print(f"case {name} p={p} theta={theta} ne={ad.ne}", flush=True)
    orc = O.Oracle(ad, c.problem, I, J)
    sample = _sample(ad, 3 if p == 4 else 6)
    kpar = c.problem.has_kpar
    refs = {nm: np.stack([orc.element(e, om, False) for e in sample])
            for nm, _, _, om, need_kpar in U3 if kpar or not need_kpar}
    full = O.ALL if kpar else (O.CURLCURL | O.MASS)
    ref_full = np.stack([orc.element(e, full, True) for e in sample])
    scales = {nm: orc.scale_elements(sample, om, False)
              for nm, _, _, om, need_kpar in U3 if kpar or not need_kpar}
    scale_full = orc.scale_elements(sample, full, True)
    g = _ctx(c, ad)
    try:
        for tensor, kernels in PATHS:
            nconst = g.set_tensor_path(tensor)
            assert (nconst > 0) == tensor
            for kern in kernels:
                # U3: each integrator alone, no face transform
                for nm, mcc, mm, om, need_kpar in U3:
                    if need_kpar and not kpar:
                        continue
                    g.compute_local(mcc, mm, False, kernel=kern)
                    Ag = g.download_local()
                    assert_parity(Ag[sample], refs[nm], TOL, f"tensor {tensor} kernel {kern} {nm}",
                                  scales[nm])
                # U4: full sum + face transform
                g.compute_local(gpu.CURLCURL, gpu.MASS_ALL, True, kernel=kern)
                Ag = g.download_local()
                assert_parity(Ag[sample], ref_full, TOL, f"tensor {tensor} kernel {kern} u4",
                              scale_full)
    finally:
        g.close()


@pytest.mark.parametrize("theta", [0.0, 6.0], ids=["t0", "t6"])
@pytest.mark.parametrize("p", [1, 2, 3], ids=["p1", "p2", "p3"])
@pytest.mark.parametrize("name", ["grating", "hole"])
def test_u5_u6_global(name, p, theta):
    c, ad, I, J = case(name, p, theta, "small")
    print(f"case {name} p={p} theta={theta} ne={ad.ne} n={ad.height} nnz={len(J)}", flush=True)
    orc = O.Oracle(ad, c.problem, I, J)
    g = _ctx(c, ad, I, J)
    try:
        g.compute_local()
        Ae = g.download_local()
        # U5: scatter without elimination
        g.scatter_eliminate(False)
        vre, vim = g.download_values()
        # bit-identical to the oracle's element-order scatter of the same Ae
        sre, sim = orc.scatter(Ae)
        assert np.array_equal(vre, sre) and np.array_equal(vim, sim)
        # against the full CPU reassembly
        rre, rim = orc.assemble()
        sc = orc.scale_csr(4)
        assert_parity(vre + 1j * vim, rre + 1j * rim, TOL, "u5", sc)
        # U6: essential elimination
        g.scatter_eliminate(True)
        ere, eim = g.download_values()
        is_ess = np.zeros(ad.height, np.uint8)
        is_ess[ad.ess] = 1
        xre, xim = O.eliminate(I, J, is_ess, rre, rim)
        sc6 = O.eliminate_scale(I, J, is_ess, sc)
        assert_parity(ere + 1j * eim, xre + 1j * xim, TOL, "u6", sc6)
        # bitwise equal to eliminating the un-eliminated GPU values
        yre, yim = O.eliminate(I, J, is_ess, vre, vim)
        assert np.array_equal(ere, yre) and np.array_equal(eim, yim)
        # determinism: a second full run is bitwise identical
        g.compute_local()
        g.scatter_eliminate(True)
        e2r, e2i = g.download_values()
        assert np.array_equal(ere, e2r) and np.array_equal(eim, e2i)
        # residual against the oracle's formula
        rng = np.random.default_rng(7)
        X = rng.standard_normal(2 * ad.height)
        B = rng.standard_normal(2 * ad.height)
        rg = g.residual(X, B)
        ro = O.residual(I, J, ere, eim, X, B)
        assert abs(rg - ro) <= 1e-12 * ro
        # the tensor / quadrature split against quadrature everywhere
        t = g.timings()
        assert t["local_tensor_ms"] > 0 and t["local_quad_ms"] > 0
        assert g.set_tensor_path(False) == 0
        g.compute_local()
        g.scatter_eliminate(True)
        qre, qim = g.download_values()
        assert g.timings()["local_tensor_ms"] == 0
        assert_parity(ere + 1j * eim, qre + 1j * qim, TOL, "tensor path vs quadrature only", sc6)
    finally:
        g.close()


@pytest.mark.parametrize("renumber", [False, True], ids=["native", "permuted"])
@pytest.mark.parametrize("p", [1, 2, 3], ids=["p1", "p2", "p3"])
def test_tensor_modes_global(p, renumber):
    """The three element paths give the same eliminated system.  With the
    global DOFs renumbered at random the face-row pairs of the fused path
    break up (single-row units) -- the path a caller's own numbering can
    take -- and the result still matches the oracle."""
    import dataclasses
    c, ad, I, J = case("hole", p, 6.0, "small")
    if renumber:
        rng = np.random.default_rng(17 + p)
        perm = rng.permutation(ad.height).astype(np.int32)
        ad = dataclasses.replace(ad, dof=np.ascontiguousarray(perm[ad.dof]),
                                 ess=np.sort(perm[ad.ess]).astype(np.int32))
        from lithofem_amd.fespace import csr_symbolic
        I, J = csr_symbolic(ad.dof, ad.height)
    orc = O.Oracle(ad, c.problem, I, J)
    rre, rim = orc.assemble()
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    xre, xim = O.eliminate(I, J, is_ess, rre, rim)
    sc6 = O.eliminate_scale(I, J, is_ess, orc.scale_csr(4))
    g = _ctx(c, ad, I, J)
    try:
        vals = {}
        for mode in (2, 1, 0):
            n = g.set_tensor_path(mode)
            assert (n > 0) == (mode > 0)
            g.compute_local()
            g.scatter_eliminate(True)
            vre, vim = g.download_values()
            vals[mode] = (vre, vim)
            assert_parity(vre + 1j * vim, xre + 1j * xim, TOL, f"mode {mode} vs oracle", sc6)
        same = np.array_equal(vals[2][0], vals[1][0]) and np.array_equal(vals[2][1], vals[1][1])
        print(f"  fused == materialised bitwise: {same}", flush=True)
        # the on-demand element matrices of the fused path
        g.set_tensor_path(2)
        g.compute_local()
        Af = g.download_local()
        # a scatter after the on-demand download still sees the fused path's
        # element coefficients
        g.scatter_eliminate(True)
        vre, vim = g.download_values()
        assert np.array_equal(vre, vals[2][0]) and np.array_equal(vim, vals[2][1])
        g.set_tensor_path(1)
        g.compute_local()
        Am = g.download_local()
        assert np.array_equal(Af, Am)
    finally:
        g.close()


def test_force_fail_env(monkeypatch):
    c, ad, I, J = case("grating", 2, 6.0, "tiny")
    monkeypatch.setenv("LITHOFEM_ASM_FORCE_FAIL", "1")
    g = gpu.AsmGpuGlobal(0)
    try:
        with pytest.raises(gpu.AsmGpuError) as ei:
            g.init(ad, c.problem)
        assert ei.value.code == gpu.E_FORCED
        assert "forced failure injected" in str(ei.value)
    finally:
        g.close()


def test_state_errors():
    c, ad, I, J = case("grating", 1, 6.0, "tiny")
    g = gpu.AsmGpuGlobal(0)
    try:
        with pytest.raises(gpu.AsmGpuError) as ei:
            g.compute_local()
        assert ei.value.code == gpu.E_STATE
        g.init(ad, c.problem)
        with pytest.raises(gpu.AsmGpuError) as ei:
            g.scatter_eliminate(True)
        assert ei.value.code == gpu.E_STATE
        # a pattern that misses element pairs is rejected before any scatter
        I2 = np.arange(ad.height + 1, dtype=np.int32)
        J2 = np.arange(ad.height, dtype=np.int32)
        with pytest.raises(gpu.AsmGpuError) as ei:
            g.build_global(I2, J2, ad.ess)
        assert ei.value.code == gpu.E_PATTERN
    finally:
        g.close()


@pytest.mark.parametrize("name,p,renumber", [("grating", 1, False), ("grating", 3, False),
                                             ("hole", 2, False), ("hole", 3, True)])
def test_device_csr_pattern(name, p, renumber):
    """The pattern built on the device equals the host symbolic phase bit for
    bit, and the assembled system through it equals the host-pattern one."""
    import dataclasses
    from lithofem_amd.fespace import csr_symbolic
    c, ad, I, J = case(name, p, 6.0, "small")
    if renumber:
        rng = np.random.default_rng(31)
        perm = rng.permutation(ad.height).astype(np.int32)
        ad = dataclasses.replace(ad, dof=np.ascontiguousarray(perm[ad.dof]),
                                 ess=np.sort(perm[ad.ess]).astype(np.int32))
        I, J = csr_symbolic(ad.dof, ad.height)
    g = _ctx(c, ad)
    try:
        g.build_global(I, J, ad.ess)
        g.compute_local()
        g.scatter_eliminate(True)
        ref = g.download_values()
        g.build_global_device(ad.ess)
        assert g.timings()["pattern_ms"] > 0
        Id, Jd = g.download_pattern()
        assert np.array_equal(Id, I) and np.array_equal(Jd, J)
        g.compute_local()
        g.scatter_eliminate(True)
        got = g.download_values()
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    finally:
        g.close()
