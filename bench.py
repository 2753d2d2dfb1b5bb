#!/usr/bin/env python3
"""Benchmark: FP64 Nedelec system-matrix assembly on one MI355X.

A step is one pass of the hot path over the EUV line-grating workload
(cubic Nedelec, 72,576 tetrahedra, ~1.35M complex DOF, theta = 6 deg so all
five integrators are active): element matrices + face transform
(lfasm_compute_local) and the deterministic CSR scatter with essential
elimination (lfasm_scatter_eliminate), inputs resident in HBM.

Element matrices of constant-coefficient elements (outside the PML) come
from the reference-tensor kernel (HBM-write bound), those of PML elements
from the FP64 quadrature kernel (MFMA), then the scatter (HBM bound).  The
autotuner picks the quadrature kernel and whether the two element kernels
run side by side (the tensor kernel held to a few workgroups per CU on one
stream, the quadrature kernel on a second stream); the per-kernel figures
come from eager steps with the kernels in sequence, `ms_per_step` from the
captured step with the tuned setting.

Prints one JSON line.  `value` is elements/s of the whole step; `roofline`
is the kernel with the largest share of the step against its own roofline
(HBM bytes or canonical FP64 FLOPs), every kernel's figures are under
`detail.kernels`; `cpu_baseline` is the C oracle (the restated reference
reassembler, gcc -O3 build) doing the step's work -- local matrices, face
transform, CSR scatter, elimination -- on one host core over a bounded
sample, with the GPU's element matrices of part of that sample checked
against the oracle's parity build; `detail.cpu_baseline_threads` is the same
work on the whole mesh over all usable host threads (at most 16).  Unless
--no-extras, `detail.variants_speedup` (64 eps variants in one delta
assembly against 64 sequential plain assemblies, order 2, 271k tets) and
`detail.adjoint_ratio` (forward + adjoint eps gradient against the forward
assembly + host solve, with one central-difference check) carry the other
two parts of the metric; `detail.angle_sweep_speedup` is the 8-angle x 2-
polarisation sweep (matrices from the k_par basis terms + one RHS launch)
against per-angle assembly on the same mesh.
"""

from __future__ import annotations

import argparse
import contextlib
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


@contextlib.contextmanager
def _stdout_to_stderr():
    """Route file descriptor 1 to stderr for the duration (the library's
    test-download call prints its marker line on stdout, which must carry
    only the result line)."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        yield
    finally:
        sys.stdout.flush()
        os.dup2(saved, 1)
        os.close(saved)


def _host_threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(wl, I, J, seconds: float = 12.0) -> dict:
    """The oracle (the restated reference reassembler, -O3 FMA build) doing
    the work of one GPU step on a contiguous element sample, one core: local
    matrices + face transform, their element-order CSR scatter, and the
    sample's share of the essential elimination (timed over the whole CSR,
    times sample / ne)."""
    from oracle import oracle as O
    ad = wl.ad
    orc = O.Oracle(ad, wl.cell.problem, I, J, fast=True)
    # calibrate on a small block, then size the sample to ~`seconds`
    t0 = time.perf_counter()
    orc.elements(0, 64)
    per = (time.perf_counter() - t0) / 64
    n = int(min(ad.ne, max(256, seconds / per)))
    e0 = (ad.ne - n) // 2
    nnz = len(J)
    vre, vim = np.zeros(nnz), np.zeros(nnz)
    t0 = time.perf_counter()
    Ae = orc.elements(e0, e0 + n)
    t1 = time.perf_counter()
    orc.scatter_elements_into(Ae, e0, vre, vim)
    t2 = time.perf_counter()
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    t3 = time.perf_counter()
    O.eliminate(I, J, is_ess, vre, vim)
    t_elim = time.perf_counter() - t3
    t_total = (t1 - t0) + (t2 - t1) + t_elim * n / ad.ne
    return dict(value=n / t_total, unit="elements/s", cores=1, kind="port",
                sample=(f"elements [{e0}, {e0 + n}) of the same mesh ({n} of {ad.ne}): oracle "
                        f"local matrices + face transform {t1 - t0:.1f} s, their CSR scatter "
                        f"{t2 - t1:.2f} s, elimination share {t_elim * n / ad.ne:.3f} s "
                        f"(gcc -O3 -march=x86-64-v3 -ffp-contract=fast)"),
                _Ae=Ae, _e0=e0, _per=t_total / n)


def cpu_baseline_threads(wl, I, J) -> dict:
    """The same work on the whole mesh over all host threads the process may
    use (at most 16): element blocks, then disjoint CSR row ranges of the
    scatter, then the elimination (one thread)."""
    from oracle import oracle as O
    ad = wl.ad
    nt = _host_threads()
    orc = O.Oracle(ad, wl.cell.problem, I, J, fast=True)
    is_ess = np.zeros(ad.height, np.uint8)
    is_ess[ad.ess] = 1
    t0 = time.perf_counter()
    Ae = orc.elements_threaded(nt)
    t1 = time.perf_counter()
    vre, vim = orc.scatter_threaded(Ae, nt)
    del Ae
    t2 = time.perf_counter()
    O.eliminate(I, J, is_ess, vre, vim)
    t3 = time.perf_counter()
    return dict(value=ad.ne / (t3 - t0), unit="elements/s", cores=nt, kind="port",
                sample=(f"all {ad.ne} elements: local matrices {t1 - t0:.1f} s ({nt} element "
                        f"blocks), CSR scatter {t2 - t1:.1f} s ({nt} row ranges), elimination "
                        f"{t3 - t2:.2f} s"))


def sample_parity(wl, I, J, Ag, e0: int, xyz, n: int = 192) -> dict:
    """GPU element matrices of `n` elements of the baseline sample against the
    parity build of the oracle (on the geometry the device formed from the
    vertex coordinates, orc_geometry): max-normalised error, error per entry
    against the entries' componentwise scales, and plain relative error per
    entry over the entries >= 1e-8 of the largest."""
    import dataclasses

    from oracle import oracle as O
    Jo, Jio, deto, v0o = O.geometry(xyz)
    ado = dataclasses.replace(wl.ad, J=Jo, Jinv=Jio, detJ=deto, v0=v0o)
    orc = O.Oracle(ado, wl.cell.problem, I, J)
    ref = orc.elements(e0, e0 + n)
    sc = orc.scale_elements(np.arange(e0, e0 + n))
    got = Ag[e0:e0 + n]
    d = np.abs(got - ref)
    a = np.abs(ref)
    gate = a >= 1e-8 * a.max()
    return dict(elements=f"[{e0}, {e0 + n})",
                maxnorm=float(d.max() / a.max()),
                per_entry_vs_scale=float(np.max(d[sc > 0] / sc[sc > 0])),
                per_entry_relative_gated=float(np.max(d[gate] / a[gate])))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the quadrature-only comparison launches (clean kernel traces)")
    ap.add_argument("--kernel", type=int, default=0, help="local kernel selector")
    ap.add_argument("--nx", type=int, default=36)
    ap.add_argument("--graph", default="on", choices=["on", "off"],
                    help="time the step as a captured HIP graph (on) or eager calls (off)")
    ap.add_argument("--tensor", default="on", choices=["on", "off", "fused"],
                    help="constant-coefficient elements: reference-tensor kernel (on), "
                         "quadrature kernels (off), tensor rows inside the scatter (fused)")
    ap.add_argument("--concurrency", type=int, default=None,
                    help="workgroup cap of the tensor kernel beside the quadrature kernel in "
                         "the captured step (0 = in sequence; default: the autotuner's choice)")
    ap.add_argument("--scatter-overlap", type=int, default=None, choices=[0, 1],
                    help="scatter the tensor-only rows beside the quadrature kernel in the "
                         "captured step (default: the autotuner's choice)")
    ap.add_argument("--quad-residency", type=int, default=None,
                    help="quadrature workgroups per CU beside the other kernels in the "
                         "captured step (0 = no cap; default: the autotuner's choice)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-variant and adjoint figures of the default run")
    ap.add_argument("--mode", default="assemble",
                    choices=["assemble", "multimesh", "variants", "rhs", "adjoint", "large",
                             "peaks"],
                    help="assemble: the headline step; others: batched / adjoint workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the "
              "assembly of one problem runs on a single device.",
              file=sys.stderr)
        return 2

    from lithofem_amd import gpu
    from lithofem_amd.workload import (canonical_flops, line_grating_p3, local_bytes,
                                       scatter_bytes, tensor_bytes)

    if args.mode == "peaks":
        out = measure_peaks()
        out.update({"n_gpus": 1, "mode": "peaks"})
        print(json.dumps(out), flush=True)
        return 0
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
    # host setup path: host geometry + host CSR symbolic + init + build_global
    gh = gpu.AsmGpuGlobal(0)
    t0 = time.perf_counter()
    gh.init(ad, pb)
    gh.build_global(I, J, ad.ess)
    t_upload = time.perf_counter() - t0
    gh.close()
    del gh
    # device setup path (the one the step runs on): geometry from the vertex
    # coordinates, contributor lists and CSR pattern all on the device
    xyz = np.ascontiguousarray(wl.cell.mesh.elem_xyz, dtype=np.float64)
    g = gpu.AsmGpuGlobal(0)
    t0 = time.perf_counter()
    g.init_vertices(ad, pb, xyz)
    t_init_v = time.perf_counter() - t0
    t0 = time.perf_counter()
    g.build_global_device(ad.ess)
    t_dev_build = time.perf_counter() - t0
    tim = g.timings()
    pattern_ms, geometry_ms = tim["pattern_ms"], tim["geometry_ms"]
    Id, Jd = g.download_pattern()
    same_pattern = bool(np.array_equal(Id, I) and np.array_equal(Jd, J))
    del Id, Jd
    log(f"device setup: init_vertices {t_init_v:.3f}s (geometry kernel {geometry_ms:.2f} ms) + "
        f"build_global_device {t_dev_build:.3f}s (pattern phase {pattern_ms:.1f} ms) vs host "
        f"extract {t_extract:.2f}s + csr_symbolic {t_csr:.3f}s + init+build_global "
        f"{t_upload:.3f}s; pattern identical: {same_pattern}")
    nconst = g.set_tensor_path({"on": 1, "off": 0, "fused": 2}[args.tensor])
    nquad = ad.ne - nconst
    log(f"element classes: {nconst} constant-coefficient (reference-tensor kernel), "
        f"{nquad} PML (quadrature kernel)")
    tuned = None
    conc = 0
    if args.kernel == 0:
        from lithofem_amd.tune import tune_local
        tuned = tune_local(g)
        log(f"autotune: best kernel {tuned['best']}: " +
            ", ".join(f"{k}={v:.3f}ms" for k, v in tuned["ms"].items()))
        if "concurrency" in tuned:
            conc = tuned["concurrency"]["best"]
            log(f"autotune: element kernels side by side with the tensor kernel capped at "
                f"{conc} workgroups (0 = in sequence): " +
                ", ".join(f"{k}={v:.3f}ms" for k, v in tuned["concurrency"]["ms"].items()))
    else:
        # the captured step launches the handle's default kernel
        g.set_local_kernel(args.kernel)
    if args.concurrency is not None:
        conc = args.concurrency
    scat = 0
    qres = 0
    tuned_step = None
    if args.scatter_overlap is not None:
        scat = args.scatter_overlap
    elif conc > 0 and args.graph == "on":
        from lithofem_amd.tune import tune_step
        tuned_step = tune_step(g, conc)
        conc, scat, qres = tuned_step["best"]
        log(f"autotune: captured step (tensor cap, tensor-only rows scattered beside the "
            f"quadrature kernel, quadrature workgroups per CU) = ({conc}, {scat}, {qres}): " +
            ", ".join(f"{k}={v:.3f}ms" for k, v in tuned_step["ms"].items()))
    if args.quad_residency is not None:
        qres = args.quad_residency
    g.set_quad_residency(qres)
    # the eager steps time every kernel on its own (element kernels in
    # sequence); the captured step runs with the chosen concurrency
    g.set_element_concurrency(0)

    def step():
        g.compute_local(gpu.CURLCURL, gpu.MASS_ALL, True, kernel=args.kernel)
        g.scatter_eliminate(True)
        t = g.timings()
        return np.array([t["local_ms"], t["scatter_ms"], t["local_tensor_ms"],
                         t["local_quad_ms"]])

    for i in range(args.warmup):
        step()
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    tsum = np.zeros(4)
    t0 = time.perf_counter()
    for i in range(args.steps):
        tsum += step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_eager = wall / args.steps * 1e3
    tl_avg, ts_avg, tt_avg, tq_avg = tsum / args.steps
    log(f"eager step {ms_eager:.3f} ms")
    ms_step = ms_eager
    ms_graph_seq = None
    if args.graph == "on":
        def graph_ms():
            # the same step as one captured HIP graph per launch
            g.step_capture(gpu.CURLCURL, gpu.MASS_ALL, True, True)
            for i in range(args.warmup):
                g.step_launch(sync=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(args.steps):
                g.step_launch(sync=False)
            torch.cuda.synchronize()       # device-wide: waits for the library's streams too
            return (time.perf_counter() - t0) / args.steps * 1e3
        if conc > 0:
            ms_graph_seq = graph_ms()
            log(f"graph step, element kernels in sequence {ms_graph_seq:.3f} ms")
            g.set_element_concurrency(conc, scat)
            g.set_quad_residency(qres)
        ms_step = graph_ms()
        log(f"graph step {ms_step:.3f} ms" +
            (f" (element kernels side by side, tensor kernel capped at {conc} workgroups; "
             f"tensor-only rows scattered beside the quadrature kernel: {scat}; quadrature "
             f"workgroups per CU: {qres or 'no cap'})"
             if conc > 0 else ""))
    log(f"step {ms_step:.3f} ms (element matrices {tl_avg:.3f} ms = tensor {tt_avg:.3f} + "
        f"quadrature {tq_avg:.3f}; scatter {ts_avg:.3f} ms)")

    # the same element matrices with the quadrature kernels on every element
    quad_only_ms = None
    if nconst > 0 and not args.no_compare:
        g.set_tensor_path(False)
        qs = []
        for _ in range(3):
            g.compute_local(gpu.CURLCURL, gpu.MASS_ALL, True, kernel=args.kernel)
            qs.append(g.timings()["local_ms"])
        quad_only_ms = min(qs)
        g.set_tensor_path({"on": 1, "off": 0, "fused": 2}[args.tensor])
        g.compute_local(gpu.CURLCURL, gpu.MASS_ALL, True, kernel=args.kernel)
        log(f"quadrature kernels on all elements: {quad_only_ms:.3f} ms")

    # the host-solver handoff (UMFPACK path): CSR values to host memory
    t0 = time.perf_counter()
    g.download_values()
    download_ms = (time.perf_counter() - t0) * 1e3
    log(f"download of the CSR values (pinned staging, re/im split on host threads): "
        f"{download_ms:.1f} ms")

    W = canonical_flops(ad.p, ad.nd, ad.rule_cc.nq, ad.rule_mass.nq, pb.has_kpar)
    # the library's slot entries are bytes when no CSR row exceeds 256 entries
    q_sc = scatter_bytes(ad.ne, ad.nd, ad.height, len(J),
                         1 if int(np.diff(I).max()) <= 256 else 2)
    kernels = {
        "k_scatter_rows": {"ms": ts_avg, "bound": "hbm", "unit": "GB/s",
                           "achieved": q_sc / (ts_avg * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                           "bytes": q_sc},
    }
    if nconst > 0:
        q_t = tensor_bytes(nconst, ad.nd)
        # the two launches of the tensor path: element coefficients, then
        # the entries (timed together)
        kernels["k_tensor_coef+k_local_tensor21"] = {
            "ms": tt_avg, "bound": "hbm", "unit": "GB/s", "achieved": q_t / (tt_avg * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS, "bytes": q_t, "elements": nconst}
    if nquad > 0:
        tq = tq_avg if nconst > 0 else tl_avg
        kernels["quadrature (MFMA/VALU)"] = {
            "ms": tq, "bound": "mfma", "unit": "TFLOP/s",
            "achieved": W * nquad / (tq * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
            "flops": W * nquad, "elements": nquad}
    for k in kernels.values():
        k["frac"] = k["achieved"] / k["peak"]
    dom_name = max(kernels, key=lambda k: kernels[k]["ms"])
    dom = kernels[dom_name]
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
    # this workload (default sizes and element split only)
    traffic = None
    if args.nx == 36 and args.tensor == "on":
        traffic = _measured_traffic().get(dom_name)
        for name, t in _measured_traffic().items():
            if name in kernels:
                kernels[name]["traffic_bytes"] = t
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
        "roofline": {"bound": dom["bound"], "achieved": dom["achieved"], "peak": dom["peak"],
                     "unit": dom["unit"], "frac": dom["frac"], "traffic": traffic,
                     "kernel": dom_name},
        "detail": {"element_matrices_ms": tl_avg, "scatter_ms": ts_avg,
                   "step_ms_eager": ms_eager, "step_launch": "hip graph" if args.graph == "on"
                   else "eager",
                   "element_concurrency_grid": conc,
                   "step_scatter_overlap": scat,
                   "step_quad_residency": qres,
                   "step_overlap_ms": ({",".join(map(str, k)): v for k, v in tuned_step["ms"].items()}
                                       if tuned_step else None),
                   "element_concurrency_ms": ({str(k): v for k, v in
                                               tuned["concurrency"]["ms"].items()}
                                              if tuned and "concurrency" in tuned else None),
                   "step_ms_graph_elements_in_sequence": ms_graph_seq,
                   "kernels": kernels,
                   "elements_tensor_path": nconst, "elements_quadrature": nquad,
                   "element_matrices_quadrature_only_ms": quad_only_ms,
                   "quadrature_kernel": tuned["best"] if tuned else args.kernel,
                   "quadrature_kernel_name": _quad_kernel_name(
                       tuned["best"] if tuned else args.kernel, ad.nd),
                   "autotune_ms": {str(k): v for k, v in tuned["ms"].items()} if tuned else None,
                   "canonical_flops_per_element": W,
                   "canonical_element_TFLOPs": W * ad.ne / (tl_avg * 1e-3) / 1e12,
                   "nnz_per_s": len(J) / (ms_step * 1e-3),
                   "download_values_ms": download_ms,
                   "elements_per_s_incl_download": ad.ne / ((ms_step + download_ms) * 1e-3),
                   "local_min_bytes": local_bytes(ad.ne, ad.nd),
                   "setup_s": {"host_path": {"extract_incl_geometry": t_extract,
                                             "csr_host_threads": t_csr,
                                             "init_and_build_global": t_upload},
                               "device_path": {"init_vertices": t_init_v,
                                               "geometry_kernel_ms": geometry_ms,
                                               "build_global_device": t_dev_build},
                               "device_pattern_phase_ms": pattern_ms,
                               "device_pattern_identical": same_pattern}},
    }
    if not args.no_cpu:
        cb = cpu_baseline(wl, I, J)
        with _stdout_to_stderr():
            Ag = g.download_local()
        par = sample_parity(wl, I, J, Ag, cb["_e0"], xyz)
        del Ag
        log(f"cpu baseline {cb['value']:.1f} elements/s ({cb['sample']}); sample parity {par}")
        out["cpu_baseline"] = {k: v for k, v in cb.items() if not k.startswith("_")}
        out["cpu_baseline"]["sample_parity"] = par
        out["cpu_baseline"]["host_cpu"] = _cpu_model()
    g.close()
    if not args.no_cpu:
        ct = cpu_baseline_threads(wl, I, J)
        log(f"cpu baseline, {ct['cores']} threads: {ct['value']:.1f} elements/s ({ct['sample']})")
        out["detail"]["cpu_baseline_threads"] = ct
    if not args.no_extras:
        from lithofem_amd import benchmarks
        v = benchmarks.variants(log)
        out["detail"]["variants_speedup"] = v["value"]
        out["detail"]["variants"] = v["detail"]
        a = benchmarks.adjoint(log, n_fd=1, region_check=False)
        out["detail"]["adjoint_ratio"] = a["value"]
        out["detail"]["adjoint"] = a["detail"]
        r = benchmarks.rhs(log)
        out["detail"]["angle_sweep_speedup"] = r["value"]
        out["detail"]["angle_sweep"] = r["detail"]
    print(json.dumps(out), flush=True)
    return 0


def measure_peaks(device: int = 0) -> dict:
    """Device ceilings from lithofem_amd/libpeaks.so (the probe library the
    assembly library does not contain): FP64 MFMA over chains per wave and
    waves per SIMD with the shader clock under load, FP64 VALU, HBM
    streams."""
    import ctypes as ct

    import torch  # noqa: F401  (one HIP runtime per process)
    L = ct.CDLL(str(ROOT / "lithofem_amd" / "libpeaks.so"))
    L.lfpeaks_measure.argtypes = [ct.c_int, ct.c_void_p, ct.c_int]
    L.lfpeaks_measure.restype = ct.c_int
    out = np.zeros(11 + 4 * 16 + 3)
    rows = L.lfpeaks_measure(device, out.ctypes.data, len(out))
    if rows < 0:
        raise RuntimeError(f"lfpeaks_measure failed ({rows})")
    sweep = [dict(chains=int(out[11 + 4 * i]), waves_per_simd=int(out[12 + 4 * i]),
                  tflops=out[13 + 4 * i], clock_ghz=out[14 + 4 * i]) for i in range(rows)]
    cus = int(out[9])
    return {"fp64_mfma_tflops": out[0], "fp64_mfma_clock_ghz": out[1],
            "fp64_mfma_best": {"chains": int(out[2]), "waves_per_simd": int(out[3])},
            "fp64_mfma_at_clock_peak_tflops": cus * 4 * 32 * out[1] * 1e9 / 1e12,
            "fp64_valu_tflops": out[4], "fp64_valu_clock_ghz": out[5],
            "fp64_valu_at_clock_peak_tflops": cus * 4 * 32 * out[5] * 1e9 / 1e12,
            "hbm_copy_gbps": out[6], "hbm_write_gbps": out[7], "hbm_read_gbps": out[8],
            "cus": cus, "reported_max_clock_ghz": out[10], "mfma_sweep": sweep,
            "mfma_valu_overlap_ms": {"mfma_alone": out[75], "valu_alone": out[76],
                                     "interleaved": out[77]}}


def _quad_kernel_name(selector: int, nd: int) -> str:
    """Kernel behind a selector of lfasm_compute_local_kernel for 32 < nd <= 48
    with all five integrators (the launch table of lfasm_api.hip)."""
    if not 32 < nd <= 48:
        return f"selector {selector}"
    # k_local_mfma_sym<staging buffers, waves per EU, priority mode, persistent>
    # (after its setup-record kernel k_sym_prep); priority 1: raised for
    # the MFMAs, 2: also for the epilogue
    return {1: "k_local_generic<4, 512> + k_te", 2: "k_local_mfma_sym<1, 3, 0, false>",
            3: "k_local_mfma_sym<2, 3, 1, false>", 4: "k_local_mfma_pipe<3, 1, 4>",
            5: "k_local_mfma<3, 1, 1, 2>", 6: "k_local_mfma_sym<1, 3, 2, false>",
            7: "k_local_mfma_sym<1, 3, 1, false>", 8: "k_local_mfma_sym<1, 3, 1, true>",
            9: "k_local_mfma_sym<1, 3, 2, true>"}.get(selector, f"selector {selector}")


def _measured_traffic() -> dict:
    """Per-kernel HBM bytes per launch of the default workload, from the
    committed counter summary (`traffic_bytes <kernel> <bytes>` lines)."""
    out = {}
    path = ROOT / "profiles" / "hbm_traffic_line_grating_p3.md"
    try:
        for line in path.read_text().splitlines():
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes":
                out[f[1]] = float(f[2])
    except OSError:
        pass
    return out


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
