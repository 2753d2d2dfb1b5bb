"""Device-side extraction against the host path and the oracle.

* lfasm_init_vertices forms J, J^-1, detJ and v0 on the device from the
  element vertex coordinates bit for bit as the oracle's restatement of that
  step (orc_geometry), and close to the mesh producer's numpy geometry;
* the per-row contributor lists built on the device (both build_global
  calls) equal the host counting sort of the dof table (ascending element,
  then local index);
* the whole assembly through init_vertices + build_global_device equals the
  host-input path bit for bit and the oracle to the parity bar;
* an inverted element is refused with its index.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from cases import assert_parity, case
from lithofem_amd import gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _host_lists(ad):
    flat = ad.dof.ravel()
    order = np.argsort(flat, kind="stable").astype(np.int32)
    ptr = np.zeros(ad.height + 1, np.int32)
    np.add.at(ptr, flat + 1, 1)
    return np.cumsum(ptr).astype(np.int32), order


@pytest.mark.parametrize("name,p", [("grating", 2), ("hole", 3), ("grating", 3)])
def test_device_extraction(name, p):
    c, ad, I, J = case(name, p, 6.0, "small")
    xyz = np.ascontiguousarray(c.mesh.elem_xyz, dtype=np.float64)
    Jo, Jio, deto, v0o = O.geometry(xyz)
    # the oracle's geometry is the producer's up to rounding
    assert np.abs(Jo - ad.J).max() <= 1e-13 * np.abs(ad.J).max()
    assert np.abs(Jio - ad.Jinv).max() <= 1e-12 * np.abs(ad.Jinv).max()
    assert np.abs(deto - ad.detJ).max() <= 1e-13 * np.abs(ad.detJ).max()
    ado = dataclasses.replace(ad, J=Jo, Jinv=Jio, detJ=deto, v0=v0o)
    pb = c.problem
    g = gpu.AsmGpuGlobal(0)
    h = gpu.AsmGpuGlobal(0)
    try:
        g.init_vertices(ad, pb, xyz)
        Jg, Jig, detg, v0g = g.download_geometry()
        for a, b in ((Jg, Jo), (Jig, Jio), (detg, deto), (v0g, v0o)):
            assert np.array_equal(a, b)
        g.build_global_device(ad.ess)
        t = g.timings()
        print(f"  {name} p={p}: ne={ad.ne} geometry {t['geometry_ms']:.3f} ms, "
              f"pattern {t['pattern_ms']:.3f} ms", flush=True)
        ptr_d, lst_d = g.download_contributors()
        ptr_h, lst_h = _host_lists(ad)
        assert np.array_equal(ptr_d, ptr_h) and np.array_equal(lst_d, lst_h)
        Id, Jd = g.download_pattern()
        assert np.array_equal(Id, I) and np.array_equal(Jd, J)
        g.compute_local()
        g.scatter_eliminate(True)
        vre, vim = g.download_values()
        # the host-input path with the same geometry: same values bit for bit
        h.init(ado, pb)
        h.build_global(I, J, ad.ess)
        ptr2, lst2 = h.download_contributors()
        assert np.array_equal(ptr2, ptr_h) and np.array_equal(lst2, lst_h)
        h.compute_local()
        h.scatter_eliminate(True)
        wre, wim = h.download_values()
        assert np.array_equal(vre, wre) and np.array_equal(vim, wim)
    finally:
        g.close()
        h.close()
    orc = O.Oracle(ado, pb, I, J)
    rre, rim = orc.assemble()
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    xre, xim = O.eliminate(I, J, is_ess, rre, rim)
    sc = O.eliminate_scale(I, J, is_ess, orc.scale_csr(4))
    assert_parity(vre + 1j * vim, xre + 1j * xim, 1e-13, "vertices + device pattern vs oracle", sc)


def test_init_vertices_refuses_inverted_element():
    c, ad, I, J = case("grating", 2, 6.0, "tiny")
    xyz = np.array(c.mesh.elem_xyz, dtype=np.float64)
    xyz[7] = xyz[7][[0, 2, 1, 3]]            # swap two vertices: detJ < 0
    g = gpu.AsmGpuGlobal(0)
    try:
        with pytest.raises(gpu.AsmGpuError) as ei:
            g.init_vertices(ad, c.problem, xyz)
        assert ei.value.code == gpu.E_ARG and "element 7" in str(ei.value)
        # the handle is usable afterwards
        g.init_vertices(ad, c.problem, np.asarray(c.mesh.elem_xyz, dtype=np.float64))
        g.build_global_device(ad.ess)
        g.compute_local()
        g.scatter_eliminate(True)
    finally:
        g.close()
