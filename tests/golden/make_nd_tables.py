"""Extract the reference's ND face-DOF transform tables as a data fixture.

Reads solver/asm_data.inc (lines 17-34: asm_nd_T and asm_nd_TInv, the
reference's copy of MFEM's ND_DofTransformation tables) from the reference
tree as text and writes tests/golden/nd_face_tables.json.  Run in a
container that has the reference mounted:
    python tests/golden/make_nd_tables.py /root/reference
"""

import json
import re
import sys
from pathlib import Path


def main(ref_root: str) -> None:
    text = (Path(ref_root) / "solver" / "asm_data.inc").read_text()
    out = {}
    for name in ("asm_nd_T", "asm_nd_TInv"):
        m = re.search(name + r"\[24\]\s*=\s*\{([^}]*)\}", text)
        vals = [float(v) for v in re.findall(r"-?\d+\.\d+", m.group(1))]
        assert len(vals) == 24, (name, len(vals))
        out[name] = vals
    out["source"] = "solver/asm_data.inc:17-34 (column-major 2x2 blocks per face orientation 0..5)"
    dst = Path(__file__).with_name("nd_face_tables.json")
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
