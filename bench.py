#!/usr/bin/env python3
"""Benchmark: FP64 Nedelec system-matrix assembly on one MI355X.

A step is one pass of the hot path over the EUV line-grating workload
(cubic Nedelec, 72,576 tetrahedra, ~1.35M complex DOF, theta = 6 deg so all
five integrators are active): element matrices + face transform
(lfasm_compute_local) and the deterministic CSR scatter with essential
elimination (lfasm_scatter_eliminate), inputs resident in HBM.

Prints one JSON line.  `value` is elements/s of the whole step; `roofline`
is the dominant kernel (element matrices) against the FP64 peak, with
canonical algorithmic FLOPs; `cpu_baseline` is the C oracle (the restated
reference reassembler) timed on one host core over a bounded sample.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("FP64 Nédélec elements/s (% roofline); batched vs plain speedup; "
          "adjoint time")
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, I, J, seconds: float = 12.0) -> dict:
    """Oracle reassembly (local matrices + element-order CSR scatter) of a
    contiguous element sample, one core."""
    from oracle import oracle as O
    ad = wl.ad
    orc = O.Oracle(ad, wl.cell.problem, I, J)
    # calibrate on a small block, then size the sample to ~`seconds`
    t0 = time.perf_counter()
    orc.elements(0, 64)
    per = (time.perf_counter() - t0) / 64
    n = int(min(ad.ne, max(256, seconds / per)))
    e0 = (ad.ne - n) // 2
    t0 = time.perf_counter()
    Ae = orc.elements(e0, e0 + n)
    t_local = time.perf_counter() - t0
    return dict(value=n / t_local, unit="elements/s", cores=1, kind="port",
                sample=f"elements [{e0}, {e0 + n}) of the same mesh ({n} of {ad.ne}), "
                       f"oracle local matrices + face transform, {t_local:.1f} s",
                _Ae=Ae, _e0=e0)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel", type=int, default=0, help="local kernel selector")
    ap.add_argument("--nx", type=int, default=36)
    ap.add_argument("--mode", default="assemble",
                    choices=["assemble", "multimesh", "variants", "rhs", "adjoint"],
                    help="assemble: the headline step; others: batched / adjoint workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the "
              "assembly of one problem runs on a single device in this build.",
              file=sys.stderr)
        return 2

    from lithofem_amd import gpu
    from lithofem_amd.workload import (canonical_flops, line_grating_p3, local_bytes,
                                       scatter_bytes)

    if args.mode != "assemble":
        from lithofem_amd import benchmarks
        out = getattr(benchmarks, args.mode)(log)
        out.update({"n_gpus": 1, "dtype": "f64", "data": "synthetic", "mode": args.mode})
        print(json.dumps(out), flush=True)
        return 0

    t0 = time.perf_counter()
    wl = line_grating_p3(nx=args.nx)
    ad, pb = wl.ad, wl.cell.problem
    t_extract = time.perf_counter() - t0
    t0 = time.perf_counter()
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    t_csr = time.perf_counter() - t0
    log(f"workload {wl.name}: ne={ad.ne} nd={ad.nd} n={ad.height} nnz={len(J)} "
        f"nq={ad.rule_cc.nq}/{ad.rule_mass.nq} extract={t_extract:.2f}s csr={t_csr:.2f}s")

    import torch
    torch.cuda.init()
    g = gpu.AsmGpuGlobal(0)
    t0 = time.perf_counter()
    g.init(ad, pb)
    g.build_global(I, J, ad.ess)
    t_upload = time.perf_counter() - t0
    log(f"upload+build_global {t_upload:.2f}s (device part {g.timings()['build_ms']:.2f} ms)")
    tuned = None
    if args.kernel == 0:
        from lithofem_amd.tune import tune_local
        tuned = tune_local(g)
        log(f"autotune: best kernel {tuned['best']}: " +
            ", ".join(f"{k}={v:.3f}ms" for k, v in tuned["ms"].items()))

    def step():
        g.compute_local(gpu.CURLCURL, gpu.MASS_ALL, True, kernel=args.kernel)
        tl = g.timings()["local_ms"]
        g.scatter_eliminate(True)
        ts = g.timings()["scatter_ms"]
        return tl, ts

    for i in range(args.warmup):
        step()
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    tl_sum = ts_sum = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        tl, ts = step()
        tl_sum += tl
        ts_sum += ts
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    tl_avg, ts_avg = tl_sum / args.steps, ts_sum / args.steps
    log(f"step {ms_step:.3f} ms (local kernel {tl_avg:.3f} ms, scatter {ts_avg:.3f} ms)")

    W = canonical_flops(ad.p, ad.nd, ad.rule_cc.nq, ad.rule_mass.nq, pb.has_kpar)
    achieved = W * ad.ne / (tl_avg * 1e-3) / 1e12
    q_sc = scatter_bytes(ad.ne, ad.nd, ad.height, len(J))
    out = {
        "metric": METRIC,
        "value": ad.ne / (ms_step * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "EUV line grating 64x8 nm, cubic Nedelec (p=3), theta=6 deg, "
                               "structured periodic tet mesh",
                   "elements": ad.ne, "dof": ad.height, "nnz": len(J), "nd": ad.nd,
                   "nq_cc": ad.rule_cc.nq, "nq_mass": ad.rule_mass.nq},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": None},
        "detail": {"local_kernel_ms": tl_avg, "scatter_ms": ts_avg,
                   "local_kernel": tuned["best"] if tuned else args.kernel,
                   "autotune_ms": {str(k): v for k, v in tuned["ms"].items()} if tuned else None,
                   "flops_per_element": W,
                   "nnz_per_s": len(J) / (ms_step * 1e-3),
                   "scatter_GBps": q_sc / (ts_avg * 1e-3) / 1e9,
                   "scatter_hbm_frac": q_sc / (ts_avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "local_min_bytes": local_bytes(ad.ne, ad.nd),
                   "setup_s": {"extract": t_extract, "csr": t_csr, "upload": t_upload}},
    }
    if not args.no_cpu:
        cb = cpu_baseline(wl, I, J)
        # the GPU element matrices of the sample match the oracle's
        Ag = g.download_local()
        n = cb["_Ae"].shape[0]
        ref = cb["_Ae"]
        err = float(np.max(np.abs(Ag[cb["_e0"]:cb["_e0"] + n] - ref)) / np.max(np.abs(ref)))
        log(f"cpu baseline {cb['value']:.1f} elements/s; sample parity {err:.2e}")
        out["cpu_baseline"] = {k: v for k, v in cb.items() if not k.startswith("_")}
        out["cpu_baseline"]["sample_parity_maxnorm"] = err
        out["cpu_baseline"]["host_cpu"] = _cpu_model()
    g.close()
    print(json.dumps(out), flush=True)
    return 0


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical CPUs)"
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
