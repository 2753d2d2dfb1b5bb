"""The drop-in header include/asm_hip.inc, compiled into a C++ driver program
(tests/shim/shim_driver.cpp) that runs the reference solve driver's
assembly sequence (solver/lithofem_solve.cpp:783-792) through it with
fixture types carrying the reference's member names.

CPU: the program builds warning-free against include/lithofem_asm.h and
liblfasm.so, and without a device takes the fallback path (the reference
driver then assembles with MFEM).  GPU: the same program assembles a small
grating; the values it hands to the host factorisation path equal the
oracle's eliminated system, the host operator it builds applies the
downloaded matrix, and its device residual of that product is at rounding
level.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cases import assert_parity, case
from lithofem_amd.nedelec import ND_TINV
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
SHIM = ROOT / "tests" / "shim"
EXE = SHIM / "shim_driver"


def write_case(d: Path, ad, pb, I, J):
    def w(name, a, dt):
        np.ascontiguousarray(a, dtype=dt).tofile(d / name)

    meta = dict(ne=ad.ne, nd=ad.nd, p=ad.p, height=ad.height, nq_cc=ad.rule_cc.nq,
                nq_mass=ad.rule_mass.nq, k0=pb.k0, kx=pb.kx, ky=pb.ky, z_min=pb.z_min,
                z_max=pb.z_max, pml_order=pb.pml_order, pml_target=pb.pml_target,
                pml_thick=pb.pml_thick)
    (d / "meta.txt").write_text("".join(f"{k} {v!r}\n" for k, v in meta.items()))
    w("dof.i32", ad.dof, np.int32)
    w("sign.f64", ad.sign, np.float64)
    if ad.p >= 2:
        w("fo.i32", ad.fo, np.int32)
    w("attr.i32", ad.attr, np.int32)
    for nm in ("J", "Jinv", "detJ", "v0"):
        w(f"{nm}.f64", getattr(ad, nm), np.float64)
    for pre, r in (("cc", ad.rule_cc), ("mass", ad.rule_mass)):
        w(f"{pre}_ip.f64", r.ip, np.float64)
        w(f"{pre}_w.f64", r.w, np.float64)
        w(f"{pre}_vs.f64", r.vshape, np.float64)
        w(f"{pre}_curl.f64", r.curl, np.float64)
    w("I.i32", I, np.int32)
    w("Jcol.i32", J, np.int32)
    w("ess.i32", ad.ess, np.int32)
    w("eps.f64", pb.eps_tab(), np.float64)
    w("is_pml.i32", pb.is_pml(), np.int32)
    w("pml_top.i32", pb.pml_top(), np.int32)
    w("tinv.f64", ND_TINV, np.float64)


def test_shim_builds_and_falls_back_without_device(tmp_path):
    r = subprocess.run(["make", "-s", "-B", "-C", str(SHIM)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert EXE.exists()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present: the GPU test runs the assembly")
    c, ad, I, J = case("grating", 2, 6.0, "tiny")
    write_case(tmp_path, ad, c.problem, I, J)
    r = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "asm_gpu: no HIP device" in r.stdout
    assert "falling back to MFEM cpu assembly" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("p", [2, 3])
def test_shim_assembles_through_the_drop_in(tmp_path, p):
    assert EXE.exists(), "tests/shim/shim_driver is built by __graft_entry__.build()"
    c, ad, I, J = case("grating", p, 6.0, "small")
    pb = c.problem
    write_case(tmp_path, ad, pb, I, J)
    r = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"asm_gpu: n {ad.height} nnz {len(J)}" in r.stdout
    assert "umfpack_factor_s" in r.stdout and "umfpack_solve_s" in r.stdout
    vre = np.fromfile(tmp_path / "vals_re.f64")
    vim = np.fromfile(tmp_path / "vals_im.f64")
    orc = O.Oracle(ad, pb, I, J)
    rre, rim = orc.assemble()
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    xre, xim = O.eliminate(I, J, is_ess, rre, rim)
    sc = O.eliminate_scale(I, J, is_ess, orc.scale_csr(4))
    assert_parity(vre + 1j * vim, xre + 1j * xim, 1e-13, "shim values", sc)
    # the host operator applies the downloaded matrix: y = A x
    n = ad.height
    x = 1.0 + 0.001 * (np.arange(2 * n) % 97)
    xc = x[:n] + 1j * x[n:]
    rows = np.repeat(np.arange(n), np.diff(I))
    want = np.zeros(n, dtype=np.complex128)
    np.add.at(want, rows, (vre + 1j * vim) * xc[J])
    y = np.fromfile(tmp_path / "y.f64")
    assert np.abs(y[:n] + 1j * y[n:] - want).max() <= 1e-13 * np.abs(want).max()
    rel = float(r.stdout.split("host product")[1].split()[0])
    assert rel < 1e-13
