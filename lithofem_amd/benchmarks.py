"""Secondary benchmark modes (bench.py --mode ...): batched assembly of many
meshes, eps variants and incident fields, and the adjoint gradient.

Each function returns a dict that bench.py prints as one JSON line.
Workloads (all FP64 complex, synthetic structured meshes):

* multimesh: 16 contact-hole cells, NE_i = round(5e4 * 8**u_i) tets
  (u_i ~ U(0,1), seed 2), hole side ~ U(20, 60) nm, sidewall ~ U(80, 90) deg,
  order 2, assembled as one block-diagonal batch vs one mesh at a time;
* variants: 64 eps variants on the largest of those meshes (every absorber
  element's eps perturbed by U(-5%, 5%) independently in real and imaginary
  part, seed = variant id): delta assembly vs 64 sequential plain
  assemblies with per-element eps;
* rhs: 8 incidence angles (0..14 deg) x 2 polarisations on that mesh: the 8
  matrices (one per k_par) as one batch, the 16 RHS in one launch;
* adjoint: contact-hole cell at order 2 (8 x 8 x 19 hexahedra, 7,296 tets,
  the host LU limits the size), gradient of the design objective with
  respect to every absorber/hole element's eps, vs the forward assembly +
  solve, with central differences on a few elements;
* large: L-shaped opening, 192 x 192 nm cell, order 4 (73,728 tets, about
  3M complex DOF): the assembly step, the RHS of the Jones vector
  (1, i)/sqrt(2) and of s and p, and the eps-sensitivity kernel over every
  absorber element (with synthetic adjoint and forward vectors -- the host
  factorisation of this size is out of reach), with the device memory the
  handle holds.
"""

from __future__ import annotations

import math
import time

import numpy as np

from . import gpu
from .fespace import build_asm_data
from .problem import contact_hole


def hole_cell(ne_target, side, sidewall, theta=6.0, phi=0.0):
    cubes = ne_target / 6.0
    h = (96.0 * 96.0 * 147.0 / cubes) ** (1.0 / 3.0)
    n_xy = max(3, int(round(96.0 / h)))
    nz = cubes / (n_xy * n_xy)
    return contact_hole(n_xy=n_xy, nz_scale=nz / 28.0, side=side, theta=theta, phi=phi,
                        sidewall_deg=sidewall)


def hole_family(n=16, seed=2):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, n)
    sides = rng.uniform(20, 60, n)
    angles = rng.uniform(80, 90, n)
    return [hole_cell(5e4 * 8 ** ui, s, a) for ui, s, a in zip(u, sides, angles)]


def _union_pattern(pats, heights):
    Is, Js, off, nnz = [], [], 0, 0
    for (I, J), n in zip(pats, heights):
        Is.append(I[:-1].astype(np.int64) + nnz)
        Js.append(J + off)
        off += n
        nnz += len(J)
    I = np.concatenate(Is + [np.array([nnz])])
    return I.astype(np.int32), np.concatenate(Js).astype(np.int32)


def _sync_time(fn, reps=3):
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
    return best


def multimesh(log) -> dict:
    cells = hole_family()
    ads = [build_asm_data(c.mesh, 2) for c in cells]
    pats = [gpu.csr_symbolic(ad.dof, ad.height) for ad in ads]
    ne = [ad.ne for ad in ads]
    log(f"multimesh: {len(ads)} meshes, ne {min(ne)}..{max(ne)}, total {sum(ne)}")
    # one mesh at a time (each its own handle, launches in sequence)
    singles = []
    for c, ad, (I, J) in zip(cells, ads, pats):
        g = gpu.AsmGpuGlobal(0)
        g.init(ad, c.problem)
        g.build_global(I, J, ad.ess)
        singles.append(g)

    def run_singles():
        for g in singles:
            g.compute_local()
            g.scatter_eliminate(True)

    run_singles()
    t_seq = _sync_time(run_singles)
    ref = [g.download_values() for g in singles]
    for g in singles:
        g.close()
    # one batch
    b = gpu.AsmGpuGlobal(0)
    b.init_batch([(ad, c.problem) for ad, c in zip(ads, cells)])
    I, J = _union_pattern(pats, [ad.height for ad in ads])
    offs = np.cumsum([0] + [ad.height for ad in ads[:-1]])
    ess = np.concatenate([ad.ess + o for ad, o in zip(ads, offs)])
    b.build_global(I, J, ess)

    def run_batch():
        b.compute_local()
        b.scatter_eliminate(True)

    run_batch()
    t_bat = _sync_time(run_batch)
    vre, vim = b.download_values()
    pos = 0
    same = True
    for (rre, rim) in ref:
        n = len(rre)
        same = same and np.array_equal(vre[pos:pos + n], rre) and np.array_equal(vim[pos:pos + n], rim)
        pos += n
    tm = b.timings()
    b.close()
    return {"metric": "batched multi-mesh assembly speedup (16 uneven contact-hole meshes, p=2)",
            "value": t_seq / t_bat, "unit": "x", "higher_is_better": True,
            "detail": {"elements": sum(ne), "elements_per_mesh": ne, "seq_ms": t_seq * 1e3,
                       "batch_ms": t_bat * 1e3, "batch_local_ms": tm["local_ms"],
                       "batch_scatter_ms": tm["scatter_ms"],
                       "elements_per_s": sum(ne) / t_bat, "bitwise_equal_to_single": same}}


def variants(log, nvar=64) -> dict:
    cells = hole_family()
    big = int(np.argmax([c.mesh.ne for c in cells]))
    cell = cells[big]
    pb = cell.problem
    ad = build_asm_data(cell.mesh, 2)
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    params = np.nonzero(ad.attr == 2)[0].astype(np.int32)      # absorber
    eps_abs = pb.regions[1].eps
    deps = np.empty((nvar, len(params)), np.complex128)
    for v in range(nvar):
        rng = np.random.default_rng(v)
        deps[v] = (eps_abs.real * 0.05 * rng.uniform(-1, 1, len(params))
                   + 1j * eps_abs.imag * 0.05 * rng.uniform(-1, 1, len(params)))
    log(f"variants: ne={ad.ne} n={ad.height} nnz={len(J)} params={len(params)} nvar={nvar}")
    g = gpu.AsmGpuGlobal(0)
    g.init(ad, pb)
    g.build_global(I, J, ad.ess)
    eps0 = np.array([pb.regions[a - 1].eps for a in ad.attr], dtype=np.complex128)
    ee = [eps0.copy() for _ in range(nvar)]
    for v in range(nvar):
        ee[v][params] += deps[v]

    def sequential():
        for v in range(nvar):
            g.set_eps_elem(ee[v])
            g.compute_local()
            g.scatter_eliminate(True)

    sequential()
    t_seq = _sync_time(sequential, reps=2)

    def batched():
        g.set_eps_elem(None)
        g.compute_local()
        g.scatter_eliminate(True)
        g.variants_prepare(params)
        g.variants_assemble(deps, eliminate=True)

    batched()
    t_bat = _sync_time(batched, reps=2)
    t_kernel = g.timings()["batch_ms"]
    # phases of one batched run (each call returns after its work is done)
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.set_eps_elem(None)
    g.compute_local()
    g.scatter_eliminate(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g.variants_prepare(params)
    t2 = time.perf_counter()
    g.variants_assemble(deps, eliminate=True)
    t3 = time.perf_counter()
    phases = {"base_assembly_ms": (t1 - t0) * 1e3, "prepare_ms": (t2 - t1) * 1e3,
              "upload_and_variant_kernel_ms": (t3 - t2) * 1e3}
    log("variants: " + ", ".join(f"{k} {v:.2f}" for k, v in phases.items()) +
        f"; kernel {t_kernel:.2f} ms; batched {t_bat * 1e3:.2f} ms vs sequential {t_seq * 1e3:.2f} ms")
    # spot check: variant 7 against a plain assembly
    vr, vi = g.download_variant(7)
    g.set_eps_elem(ee[7])
    g.compute_local()
    g.scatter_eliminate(True)
    pr, pi = g.download_values()
    err = float(np.max(np.abs((vr - pr) + 1j * (vi - pi))) / np.max(np.abs(pr + 1j * pi)))
    g.close()
    return {"metric": "64-variant eps batch speedup vs 64 sequential plain assemblies (p=2)",
            "value": t_seq / t_bat, "unit": "x", "higher_is_better": True,
            "detail": {"elements": ad.ne, "dof": ad.height, "nnz": len(J),
                       "params": len(params), "variants": nvar,
                       "sequential_ms": t_seq * 1e3, "batched_ms": t_bat * 1e3,
                       "variant_kernel_ms": t_kernel, "phases": phases,
                       "variant7_maxnorm_vs_plain": err,
                       "output_GB": nvar * len(J) * 16 / 1e9}}


def rhs(log) -> dict:
    from .incident import incident_set
    cells = hole_family()
    big = int(np.argmax([c.mesh.ne for c in cells]))
    base = cells[big]
    thetas = [2.0 * i for i in range(8)]
    ad = build_asm_data(base.mesh, 2)
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    import copy
    probs = []
    for th in thetas:
        pb = copy.deepcopy(base.problem)
        k0 = pb.k0
        pb.kx = k0 * math.sin(math.radians(th))
        pb.ky = 0.0
        probs.append(pb)
    fields = [[(th, 0.0, "s"), (th, 0.0, "p")] for th in thetas]
    inc = incident_set(probs, fields)
    log(f"rhs: ne={ad.ne} x {len(thetas)} angles x 2 polarisations")
    b = gpu.AsmGpuGlobal(0)
    b.init_batch([(ad, pb) for pb in probs])
    Iu, Ju = _union_pattern([(I, J)] * len(probs), [ad.height] * len(probs))
    offs = np.arange(len(probs)) * ad.height
    ess = np.concatenate([ad.ess + o for o in offs])
    b.build_global(Iu, Ju, ess)
    b.set_incident(inc)

    def batched():
        b.compute_local()
        b.scatter_eliminate(True)
        b.assemble_rhs(True, download=False)

    batched()
    t_bat = _sync_time(batched)
    rhs_kernel_ms = b.timings()["batch_ms"]
    t0 = time.perf_counter()
    rb = b.download_rhs()
    t_dl = time.perf_counter() - t0
    b.close()
    # per angle: own handle, matrix + 2 RHS
    singles = []
    for pb, fl in zip(probs, fields):
        g = gpu.AsmGpuGlobal(0)
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        g.set_incident(incident_set([pb], [fl]))
        singles.append(g)

    def seq():
        for g in singles:
            g.compute_local()
            g.scatter_eliminate(True)
            g.assemble_rhs(True, download=False)

    seq()
    t_seq = _sync_time(seq)
    # angle k's two columns of the batch are rows [k n, (k+1) n) of each
    # batched column
    same = True
    for k, g in enumerate(singles):
        r1 = g.download_rhs()
        same = same and np.array_equal(rb[:, k * ad.height:(k + 1) * ad.height], r1)
    # the same sweep on one single-mesh handle: every angle's matrix from
    # the basis matrices of the k_par polynomial (3 here: phi = 0), and the
    # 16 RHS columns in one launch
    a = gpu.AsmGpuGlobal(0)
    a.init(ad, probs[0])
    a.build_global(I, J, ad.ess)
    a.set_incident(incident_set([probs[0]], [sum(fields, [])]))
    kxy = np.array([[pb.kx, pb.ky] for pb in probs])

    def swept():
        a.angles_assemble(kxy, eliminate=True)
        a.assemble_rhs(True, download=False)

    swept()
    t_sw = _sync_time(swept)
    ra = a.download_rhs()
    err_m = 0.0
    for k, g in enumerate(singles):
        vr, vi = a.download_angle(k)
        pr, pi = g.download_values()
        err_m = max(err_m, float(np.max(np.abs((vr - pr) + 1j * (vi - pi))) /
                                 np.max(np.abs(pr + 1j * pi))))
    # column 2k + pol of the sweep is column pol of angle k's own RHS
    same_rhs = all(np.array_equal(ra[[2 * k, 2 * k + 1]], g.download_rhs())
                   for k, g in enumerate(singles))
    a.close()
    for g in singles:
        g.close()
    log(f"rhs: batched {t_bat * 1e3:.2f} ms (RHS kernels {rhs_kernel_ms:.2f} ms), per angle "
        f"{t_seq * 1e3:.2f} ms, download {t_dl * 1e3:.1f} ms, bitwise equal {same}; k_par "
        f"sweep {t_sw * 1e3:.2f} ms (matrices vs per angle {err_m:.1e}, RHS bitwise equal "
        f"{same_rhs})")
    return {"metric": "8 angles x 2 polarisations: matrices + RHS speedup vs per angle (p=2)",
            "value": t_seq / t_sw, "unit": "x", "higher_is_better": True,
            "detail": {"elements_per_angle": ad.ne, "kpar_sweep_ms": t_sw * 1e3,
                       "block_diagonal_batch_ms": t_bat * 1e3,
                       "block_diagonal_speedup": t_seq / t_bat,
                       "sequential_ms": t_seq * 1e3, "batched_rhs_kernels_ms": rhs_kernel_ms,
                       "rhs_download_ms": t_dl * 1e3, "rhs_columns": 2 * len(thetas),
                       "kpar_sweep_matrix_maxnorm_vs_per_angle": err_m,
                       "kpar_sweep_rhs_bitwise_equal": bool(same_rhs),
                       "bitwise_equal_to_per_angle": bool(same)}}


def adjoint(log, n_xy=8, nz_scale=0.67, n_fd=2, region_check=True) -> dict:
    import torch

    from .adjoint import DesignObjective, EpsForwardModel, make_torch_function, plane_points
    from .incident import evaluate, incident_set
    cell = contact_hole(n_xy=n_xy, nz_scale=nz_scale, side=48.0, theta=6.0)
    pb = cell.problem
    ad = build_asm_data(cell.mesh, 2)
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    inc = incident_set([pb], [[(6.0, 0.0, "s")]])
    params = np.nonzero((ad.attr == 2) | (ad.attr == 3))[0].astype(np.int32)
    log(f"adjoint: ne={ad.ne} n={ad.height} params={len(params)}")
    model = EpsForwardModel(ad, pb, I, J, params, incident=inc, rhs_col=0)
    rng = np.random.default_rng(3)
    target = rng.uniform(0, 1, (32, 32))
    pts = plane_points(96.0, 96.0, -1.6, 32, 32)
    obj = DesignObjective(ad, pb, 96.0, 96.0, 115.0, -1.6, 32, target,
                          evaluate(inc.waves, 0, pts[:, 2]), n_orders_grid=16)
    eps0 = model.eps0[params].copy()
    F = make_torch_function()
    eps = torch.tensor(eps0, dtype=torch.complex128, requires_grad=True)
    t0 = time.perf_counter()
    u = F.apply(eps, model)
    t_fwd = time.perf_counter() - t0
    fwd_phases = dict(model.times)
    log(f"adjoint: forward assembly + factorisation + solve {t_fwd:.2f} s "
        + " ".join(f"{k} {v:.3f}" for k, v in fwd_phases.items()))
    Jv = obj(u)
    t0 = time.perf_counter()
    Jv.backward()
    t_bwd = time.perf_counter() - t0
    bwd_phases = {k: model.times[k] for k in ("transpose_solve", "sensitivity")}
    log(f"adjoint: backward (transpose solve + sensitivity kernel) {t_bwd:.3f} s "
        + " ".join(f"{k} {v:.3f}" for k, v in bwd_phases.items()))
    grad = eps.grad.numpy()
    rels = []
    picks = np.random.default_rng(4).choice(len(params), n_fd, replace=False)
    for i in picks:
        h = 1e-6 * abs(eps0[i])
        ep, em = eps0.copy(), eps0.copy()
        ep[i] += h
        em[i] -= h
        jp = float(obj(torch.from_numpy(model.solve(ep))))
        jm = float(obj(torch.from_numpy(model.solve(em))))
        fd = (jp - jm) / (2 * h)
        an = float(np.conj(grad[i]).real)
        rels.append(abs(fd - an) / abs(an))
        log(f"adjoint: element {params[i]}: adjoint {an:.9e} central FD {fd:.9e} "
            f"rel {rels[-1]:.2e}")
    # the region sum (the only eps derivative the region-eps reference has):
    # sum of g_e over the absorber against a two-solve central difference
    region_rel = None
    if region_check:
        sel = ad.attr[params] == 2
        h = 1e-6 * abs(eps0[sel]).max()
        d = np.where(sel, h, 0.0)
        fd = (float(obj(torch.from_numpy(model.solve(eps0 + d))))
              - float(obj(torch.from_numpy(model.solve(eps0 - d))))) / (2 * h)
        an = float(np.conj(grad[sel]).real.sum())
        region_rel = abs(fd - an) / abs(an)
        log(f"adjoint: absorber region sum {an:.9e} two-solve difference {fd:.9e} "
            f"rel {region_rel:.2e}")
    adj_ms = model.g.timings()["adjoint_ms"]
    model.close()
    return {"metric": "forward + adjoint eps-gradient time / forward assembly+solve time (p=2)",
            "value": (t_fwd + t_bwd) / t_fwd, "unit": "x forward", "higher_is_better": False,
            "detail": {"elements": ad.ne, "dof": ad.height, "params": len(params),
                       "forward_s": t_fwd, "backward_s": t_bwd,
                       "forward_phases_s": fwd_phases, "backward_phases_s": bwd_phases,
                       "host_lu": "scipy SuperLU (splu, MMD_AT_PLUS_A) on the host",
                       "sensitivity_kernel_ms": adj_ms, "objective": float(Jv.detach()),
                       "fd_rel_errors": rels, "region_sum_rel_error": region_rel}}


def large(log, n_xy=16, nz_scale=1.0, steps=5) -> dict:
    import torch

    from .incident import incident_set, sp_coefficients
    from .problem import l_opening
    t0 = time.perf_counter()
    cell = l_opening(n_xy=n_xy, nz_scale=nz_scale)
    pb = cell.problem
    ad = build_asm_data(cell.mesh, 4)
    I, J = gpu.csr_symbolic(ad.dof, ad.height)
    log(f"large: ne={ad.ne} nd={ad.nd} n={ad.height} nnz={len(J)} "
        f"setup {time.perf_counter() - t0:.1f}s")
    free0, total = torch.cuda.mem_get_info()
    g = gpu.AsmGpuGlobal(0)
    try:
        g.init(ad, pb)
        g.build_global(I, J, ad.ess)
        free1, _ = torch.cuda.mem_get_info()
        r2 = 1.0 / math.sqrt(2.0)
        # Jones (1, i)/sqrt(2): transverse (Ex, Ey) of the incident field
        inc = incident_set([pb], [[(6.0, 0.0, (r2, 1j * r2)), (6.0, 0.0, "s"),
                                   (6.0, 0.0, "p")]])
        g.set_incident(inc)
        from .tune import tune_local, tune_step
        tuned = tune_local(g)
        log(f"large: autotune {tuned['best']}: " +
            ", ".join(f"{k}={v:.2f}ms" for k, v in tuned["ms"].items()))
        conc = tuned.get("concurrency", {}).get("best", 0)
        # eager steps with the element kernels in sequence (per-kernel times)
        g.set_element_concurrency(0)
        g.compute_local()
        g.scatter_eliminate(True)
        torch.cuda.synchronize()
        tl = ts = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            g.compute_local()
            g.scatter_eliminate(True)
            t = g.timings()
            tl += t["local_ms"]
            ts += t["scatter_ms"]
        torch.cuda.synchronize()
        eager_ms = (time.perf_counter() - t0) / steps * 1e3
        log(f"large: eager step {eager_ms:.2f} ms (element matrices {tl / steps:.2f}, "
            f"scatter {ts / steps:.2f})")
        # the captured step with the tuned overlap of the element kernels and
        # the scatters (or in sequence when the tuner keeps them apart)
        overlap = (0, 0, 0)
        if conc > 0:
            ts_ = tune_step(g, conc)
            overlap = ts_["best"]
            log(f"large: captured step (tensor cap, scatter overlap, quadrature residency) = "
                f"{overlap}: " + ", ".join(f"{k}={v:.2f}ms" for k, v in ts_["ms"].items()))
        g.set_element_concurrency(overlap[0], overlap[1])
        g.set_quad_residency(overlap[2])
        g.step_capture()
        g.step_launch(sync=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step_launch(sync=False)
        torch.cuda.synchronize()
        step_ms = (time.perf_counter() - t0) / steps * 1e3
        log(f"large: step {step_ms:.2f} ms (captured step, overlap {overlap})")
        b = g.assemble_rhs(eliminate=True)
        rhs_ms = g.timings()["batch_ms"]
        cs, cp = sp_coefficients(6.0, 0.0, (r2, 1j * r2))
        jones_lin = float(np.max(np.abs(b[0] - (cs * b[1] + cp * b[2]))) / np.max(np.abs(b[0])))
        params = np.nonzero(ad.attr == 2)[0].astype(np.int32)
        rng = np.random.default_rng(1)
        lam = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
        u = rng.standard_normal(ad.height) + 1j * rng.standard_normal(ad.height)
        g.eps_sensitivity(params, lam, u, rhs_col=0)
        sens = g.eps_sensitivity(params, lam, u, rhs_col=0)
        adj_ms = g.timings()["adjoint_ms"]
        log(f"large: rhs (3 fields) {rhs_ms:.2f} ms; sensitivity over {len(params)} "
            f"absorber elements {adj_ms:.2f} ms")
        used_gb = (free0 - free1) / 1e9
    finally:
        g.close()
    return {"metric": "order-4 assembly of the ~3M-DOF L-opening: elements/s",
            "value": ad.ne / (step_ms * 1e-3), "unit": "elements/s", "higher_is_better": True,
            "detail": {"elements": ad.ne, "dof": ad.height, "nnz": len(J),
                       "step_ms": step_ms, "step_ms_eager_in_sequence": eager_ms,
                       "step_overlap": list(overlap), "element_matrices_ms": tl / steps,
                       "scatter_ms": ts / steps, "nnz_per_s": len(J) / (step_ms * 1e-3),
                       "quadrature_kernel": tuned["best"],
                       "rhs_3_fields_ms": rhs_ms, "jones_linearity_maxnorm": jones_lin,
                       "sensitivity_elements": int(len(params)),
                       "sensitivity_kernel_ms": adj_ms,
                       "sensitivity_finite": bool(np.all(np.isfinite(sens))),
                       "device_memory_GB": used_gb, "device_memory_total_GB": total / 1e9}}
