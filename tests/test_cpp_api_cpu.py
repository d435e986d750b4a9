"""CPU checks of the drop-in C++ API: the known-answer program (host path) and
the .field format round trip through ScalarField::write/readFromFile."""
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "phix_amd" / "_build" / "bin"


def test_fieldops_known_answers_cpu():
    r = subprocess.run([str(BIN / "fieldops"), "cpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
