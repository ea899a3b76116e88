"""Copy the contour log of the reference's committed LJ-8 NPT run as a
fixture: tests/_assets/data/lj8_npt/output/lj8_npt.energies of the reference
checkout (200 rows ``iter H V`` after the header ``K n_cull n_dof 0.0 N``;
the run's settings are in lj8_npt.config.snapshot.yaml beside it) is read as
text and written to lj8_npt.energies next to this script.  Run once in a
container that holds the reference; the output is committed.
"""

import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else
           "/root/reference/tests/_assets/data/lj8_npt/output/lj8_npt.energies")
OUT = Path(__file__).with_name("lj8_npt.energies")


def main():
    text = SRC.read_text()
    rows = [ln for ln in text.splitlines()[1:] if ln.strip()]
    assert len(rows) == 200, len(rows)
    OUT.write_text(text)
    print(f"wrote {len(rows)} rows to {OUT}")


if __name__ == "__main__":
    main()
