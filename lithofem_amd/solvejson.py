"""The reference's ``solve.json`` (src/lithofem/config.py:689-815) as
problem data for the assembly: region table (attribute = 1-based index into
``regions``), PML parameters, background slabs, and per group the Bloch
vector and the plane-wave fields whose scattered-field RHS the group needs:
as descriptions (theta, phi, polarisation, amplitude) that
lithofem_amd.incident recomputes, and -- when the group carries them, as
the reference's driver writes it (src/lithofem/driver.py:28-35) -- as the
reference's own per-slab incident tables, one per source.

Only what the assembly path consumes is read.  Plane waves incident from
the bottom and point/line/sheet sources are outside this path and raise.
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

from .incident import tables_from_solve_json
from .problem import Problem, Region


def _cnum(v) -> complex:
    if isinstance(v, (list, tuple)):
        return complex(float(v[0]), float(v[1]))
    return complex(v)


@dataclass
class GroupCase:
    problem: Problem
    fields: list          # [(theta, phi, pol, amplitude)] for incident_set
    source_indices: list
    tables: list | None = None   # the group's own (nslab, 18) tables, per source

    def incident_fields(self) -> list:
        """The group's tables when the solve.json holds them, else the
        descriptions (both are inputs of incident_set)."""
        return self.tables if self.tables is not None else self.fields


def load_solve_json(src) -> list:
    """solve.json path or dict -> one GroupCase per group."""
    d = src if isinstance(src, dict) else json.loads(Path(src).read_text())
    if d.get("lateral_bc", "periodic") != "periodic":
        raise NotImplementedError("only Bloch-periodic lateral boundaries are on this path")
    k0 = float(d["k0"])
    wl = float(d["wavelength"])
    z_min, z_max = float(d["domain"]["z_min"]), float(d["domain"]["z_max"])
    pml = d["pml"]
    bounds = [float(z) for z in d["slabs"]]
    nslab = len(bounds) - 1
    bg = {}
    for r in d["regions"]:
        if r["kind"] == "background":
            bg[int(r["slab"])] = _cnum(r["epsilon"])
    slabs = [(bounds[i], bounds[i + 1], bg[i]) for i in range(nslab)]
    regions = []
    for r in d["regions"]:
        kind = r["kind"]
        eps = _cnum(r["epsilon"])
        slab = min(max(int(r["slab"]), 0), nslab - 1)
        if kind in ("pml_bottom", "pml_top"):
            regions.append(Region(eps, True, kind == "pml_top", kind, slab=slab))
        elif kind == "frustum":
            regions.append(Region(eps, name=f"frustum{r.get('frustum', '')}", slab=slab,
                                  source=True, eps_bg=bg[slab]))
        else:
            regions.append(Region(eps, name=kind, slab=slab))
    out = []
    for g in d["groups"]:
        kx, ky = (float(v) for v in g["kpar"])
        pb = Problem(k0=k0, kx=kx, ky=ky, z_min=z_min, z_max=z_max,
                     pml_thick=float(pml["thickness_wavelengths"]) * wl, regions=regions,
                     pml_order=float(pml["order"]), pml_target=float(pml["target_reflection"]),
                     slabs=slabs)
        fields = []
        for i in g["source_indices"]:
            s = d["sources"][i]
            if s["type"] != "planewave":
                raise NotImplementedError(f"source type {s['type']} is not on the assembly path")
            if s.get("from", "top") != "top":
                raise NotImplementedError("plane waves from the bottom are not on this path")
            amp = _cnum(s.get("amplitude", 1.0))
            pol = s.get("polarization", "s")
            if pol not in ("s", "p"):
                # a Jones vector (Ex, Ey) of the transverse incident field
                pol = tuple(_cnum(v) for v in s["jones"])
            fields.append((float(s["theta"]), float(s["phi"]), pol, amp))
        tables = None
        if g.get("incident"):
            by_src = tables_from_solve_json(g["incident"], nslab)
            tables = [by_src[i] for i in g["source_indices"]]
        out.append(GroupCase(pb, fields, list(g["source_indices"]), tables))
    return out
