"""Print the measured device ceilings (lfasm_measure_peaks) as JSON.

    python lithofem_amd/peaks.py [device]
"""

import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

from lithofem_amd import gpu  # noqa: E402

if __name__ == "__main__":
    dev = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    print(json.dumps(gpu.measure_peaks(dev)))
