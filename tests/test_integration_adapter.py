"""The reference-side adapter (nbody_amd/integration/nbody_engine_mi355x.{h,cpp})
is compiled, syntax only, against the reference's own nbody_engine.h and
nbody_data.h with the Qt headers of this image: it overrides every pure
virtual of the interface it plugs into.  Skipped where either is absent (the
GPU box has neither); nothing of the reference is built or run."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference/nbody")
QT = Path("/opt/conda/include/qt")


@pytest.mark.skipif(not (REF / "nbody_engine.h").exists() or not (QT / "QtCore").exists()
                    or shutil.which("g++") is None, reason="reference headers or Qt headers not present")
def test_adapter_compiles_against_reference_interface(tmp_path):
    src = ROOT / "nbody_amd" / "integration" / "nbody_engine_mi355x.cpp"
    incs = [REF, ROOT / "include", QT] + [QT / m for m in ("QtCore", "QtGui", "QtWidgets", "QtOpenGL")]
    cmd = ["g++", "-std=gnu++11", "-fsyntax-only", "-fPIC", "-w", *[f"-I{p}" for p in incs], str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-3000:]
    # an abstract class would not be instantiable: the factory branch news both engines
    text = src.read_text()
    assert "new nbody_engine_mi355x()" in text and "new nbody_engine_mi355x_bh(" in text
