"""Launch autotuner for the element-matrix kernels.

Times every launch variant of the quadrature element-matrix kernels (the
VALU kernel and the MFMA variants: elements per workgroup, waves per EU,
pipelined or not) on the handle's own mesh with HIP events, and makes the
fastest one the handle's default.  Results are plain dicts so callers can log them.
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
    return {"best": choice, "ms": ms}
