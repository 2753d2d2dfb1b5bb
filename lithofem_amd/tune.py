"""Launch autotuner for the element-matrix kernels.

Times every launch variant of the quadrature element-matrix kernels (the
VALU kernel and the MFMA variants: elements per workgroup, waves per EU,
pipelined or not) on the handle's own mesh with HIP events, and makes the
fastest one the handle's default; then, when the reference-tensor kernel
and the quadrature kernel both run, the workgroup cap under which they run
side by side (0 = in sequence), and for the captured step the overlap of
the row scatter and the quadrature kernel's residency per CU.  Results are plain dicts so callers can log
them.
"""

from __future__ import annotations

from . import gpu


def candidates(nd: int) -> list[int]:
    ks = [gpu.kernel_mfma(v) for v in range(gpu.N_MFMA_VARIANTS)]
    if nd <= 48:
        ks = [gpu.KERNEL_VALU] + ks
    return ks


def tune_local(g: "gpu.AsmGpuGlobal", reps: int = 3, mask_cc: int = gpu.CURLCURL,
               mask_mass: int = gpu.MASS_ALL) -> dict:
    """Times the quadrature kernels on the elements they serve: the PML
    elements when the reference-tensor path covers the rest, else all."""
    ms = {}
    for k in candidates(g.nd):
        g.compute_local(mask_cc, mask_mass, True, kernel=k)   # warm-up launch
        best = float("inf")
        for _ in range(reps):
            g.compute_local(mask_cc, mask_mass, True, kernel=k)
            t = g.timings()
            best = min(best, t["local_quad_ms"] if t["local_tensor_ms"] > 0 else t["local_ms"])
        ms[k] = best
    choice = min(ms, key=ms.get)
    g.set_local_kernel(choice)
    out = {"best": choice, "ms": ms}
    # both element kernels in use (tensor path with PML elements): also
    # choose whether and how far they run side by side
    g.compute_local(mask_cc, mask_mass, True)
    t = g.timings()
    if t["local_tensor_ms"] > 0 and t["local_quad_ms"] > 0:
        out["concurrency"] = tune_concurrency(g, reps, mask_cc, mask_mass)
    return out


# workgroup caps of the tensor kernel tried beside the quadrature kernel
CONCURRENCY_GRIDS = (0, 64, 96, 128, 160, 224)


def tune_concurrency(g: "gpu.AsmGpuGlobal", reps: int = 3, mask_cc: int = gpu.CURLCURL,
                     mask_mass: int = gpu.MASS_ALL) -> dict:
    """Times the whole element-matrix phase with the tensor kernel and the
    quadrature kernel in sequence (0) and side by side with the tensor
    kernel held to each grid cap, and keeps the fastest."""
    ms = {}
    for grid in CONCURRENCY_GRIDS:
        g.set_element_concurrency(grid)
        g.compute_local(mask_cc, mask_mass, True)   # warm-up launch
        best = float("inf")
        for _ in range(reps):
            g.compute_local(mask_cc, mask_mass, True)
            best = min(best, g.timings()["local_ms"])
        ms[grid] = best
    choice = min(ms, key=ms.get)
    g.set_element_concurrency(choice)
    return {"best": choice, "ms": ms}




QUAD_RESIDENCIES = (0, 3, 2)   # quadrature workgroups per CU beside the other kernels


def tune_step(g: "gpu.AsmGpuGlobal", grid: int, reps: int = 5,
              mask_cc: int = gpu.CURLCURL, mask_mass: int = gpu.MASS_ALL) -> dict:
    """Times the captured step (element matrices, scatter, elimination) with
    the element kernels side by side under `grid`, with and without the
    rows formed by tensor-path elements only scattered beside the
    quadrature kernel -- the latter also with wider tensor caps, under which
    the quadrature kernel runs more beside that scatter than beside the
    tensor kernel -- each with the quadrature kernel's residency uncapped or
    held to 3 or 2 workgroups per CU, and keeps the fastest
    (grid, overlap, residency) triple."""
    ms = {}
    cands = [(grid, 0), (grid, 1)]
    if grid > 0:
        cands += [(2 * grid, 1), (4 * grid, 1), (8 * grid, 1), (1 << 20, 1)]
    for res in QUAD_RESIDENCIES:
        g.set_quad_residency(res)
        for gr, sg in cands:
            g.set_element_concurrency(gr, sg)
            g.step_capture(mask_cc, mask_mass, True, True)
            g.step_launch(sync=True)                    # warm-up replay
            best = float("inf")
            for _ in range(reps):
                g.step_launch(sync=True)
                best = min(best, g.timings()["step_ms"])
            ms[(gr, sg, res)] = best
    choice = min(ms, key=ms.get)
    g.set_element_concurrency(choice[0], choice[1])
    g.set_quad_residency(choice[2])
    return {"best": choice, "ms": ms}
