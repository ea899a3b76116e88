"""Extract the LJ-8 NPT dead frames committed with the reference as a fixture.

Reads (as text only) tests/_assets/data/lj8_npt/output/lj8_npt.traj.extxyz of
the reference checkout and writes lj8_dead_frames.json next to this script:
for each frame the cell (rows = lattice vectors), positions, the logged
``ns_energy`` (= U_LJ + P V, float32 in the reference run) and the iteration.
The file's text is also copied beside it (``lj8_npt.traj.extxyz``), the
byte-level golden of the extxyz writer.  Run once in a container that holds
the reference; both outputs are committed.
"""

import json
import re
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else
           "/root/reference/tests/_assets/data/lj8_npt/output/lj8_npt.traj.extxyz")
OUT = Path(__file__).with_name("lj8_dead_frames.json")


def main():
    lines = SRC.read_text().splitlines()
    frames, i = [], 0
    while i < len(lines):
        n = int(lines[i].strip())
        hdr = lines[i + 1]
        lat = [float(x) for x in re.search(r'Lattice="([^"]+)"', hdr).group(1).split()]
        e = float(re.search(r"ns_energy=(\S+)", hdr).group(1))
        it = int(re.search(r"iter=(\d+)", hdr).group(1))
        pos = [[float(v) for v in lines[i + 2 + a].split()[1:4]] for a in range(n)]
        frames.append({"iter": it, "cell": [lat[0:3], lat[3:6], lat[6:9]],
                       "positions": pos, "ns_energy": e})
        i += 2 + n
    OUT.write_text(json.dumps({"source": "lj8_npt.traj.extxyz", "pressure": 0.1,
                               "epsilon": 1.0, "sigma": 1.0, "cutoff": 2.5,
                               "frames": frames}, indent=1))
    OUT.with_name("lj8_npt.traj.extxyz").write_text(SRC.read_text())
    print(f"wrote {len(frames)} frames to {OUT}")


if __name__ == "__main__":
    main()
