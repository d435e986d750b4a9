"""Integrity of the data fixtures under tests/golden (no GPU): the PFHub-1a
MOOSE free-energy curve copied from the reference (develop/Spinodal
Decomposition/reference/MOOSE_IA_1a.csv) has its 442 rows, strictly
increasing times, and a free energy that does not rise after t = 1 beyond
1e-6 relative (the MOOSE solver's own noise at convergence is 2.6e-8 on the
last step; the t = 0 row holds 0, not an energy)."""
import csv
from pathlib import Path

import numpy as np


def test_pfhub1a_moose_curve_fixture():
    path = Path(__file__).resolve().parent / "golden" / "pfhub1a_moose_ia_energy.csv"
    rows = list(csv.DictReader(open(path)))
    assert len(rows) == 442 and set(rows[0]) == {"time", "TotalEnergy"}
    t = np.array([float(r["time"]) for r in rows])
    f = np.array([float(r["TotalEnergy"]) for r in rows])
    assert t[0] == 0.0 and f[0] == 0.0 and np.all(np.diff(t) > 0)
    late = f[t >= 1.0]
    assert np.all(np.diff(late) <= 1e-6 * late[1:]) and 318.0 < late[0] < 320.0
