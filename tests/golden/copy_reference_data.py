"""Copy the data files of the reference's own test suite into this directory.

Source: drons/nbody test/data/ (initial_state.txt, the 19 solver end states
used by test/solver/test_nbody_solver.cpp:172-266, the Zeno/G1 loader inputs).
They are fixtures (inputs and expected outputs), copied byte for byte; run this
script only where the reference checkout is present.
"""
import shutil
import sys
from pathlib import Path

SRC = Path("/root/reference/test/data")
DST = Path(__file__).resolve().parent / "reference_data"


def main() -> int:
    if not SRC.is_dir():
        print(f"{SRC} not found", file=sys.stderr)
        return 1
    DST.mkdir(exist_ok=True)
    for f in sorted(SRC.glob("*.txt")):
        shutil.copyfile(f, DST / f.name)
        print(f.name)
    return 0


if __name__ == "__main__":
    sys.exit(main())
