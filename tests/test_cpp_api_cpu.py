"""CPU checks of the drop-in C++ API: the known-answer program (host path) and
the .field format round trip through ScalarField::write/readFromFile."""
import shutil
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


GOLDEN = ROOT / "tests" / "golden"


def test_config_io_cpu(tmp_path):
    """JSONC parser, ConfigFile on the reference's two sample settings files,
    buildBCs rules, OutputWriter files and cadence, .field/.dat round trips and
    restart resolution (tests/cpp/config_io.cpp), host code only."""
    r = subprocess.run([str(BIN / "config_io"), str(tmp_path / "scratch"), str(GOLDEN / "settings_ch_2d.jsonc"),
                        str(GOLDEN / "settings_ch_ac_2d.jsonc")], capture_output=True, text=True, timeout=120)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert " of " in r.stdout.splitlines()[-1]


def test_driver_usage_on_bad_settings(tmp_path):
    """The config-driven solvers exit 1 with the error and a usage line when the
    settings file is missing or malformed, before touching the device."""
    bad = tmp_path / "broken.jsonc"
    bad.write_text('{ "mesh": { "nx": 8, } }\n')
    for exe in ("ch_2d", "ch_ac_2d", "gfa_2d"):
        for arg in (str(tmp_path / "absent.jsonc"), str(bad)):
            r = subprocess.run([str(BIN / exe), arg], capture_output=True, text=True, timeout=120, cwd=tmp_path)
            assert r.returncode == 1, (exe, arg, r.stdout, r.stderr)
            assert "Usage: " in r.stderr and "[path/to/settings.jsonc]" in r.stderr, r.stderr
        r = subprocess.run([str(BIN / exe)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
        assert r.returncode == 1 and "settings/settings.jsonc" in r.stderr, r.stderr


def test_gfa_2d_reads_reference_settings(tmp_path):
    """gfa_2d parses the reference's FeB_ex1 settings file (JSONC with UTF-8
    comments), reads every constant it names, and stops at the missing initial
    field with the reference's error, before any device call."""
    (tmp_path / "settings").mkdir()
    shutil.copy(GOLDEN / "settings_gfa_2d.jsonc", tmp_path / "settings" / "settings.jsonc")
    r = subprocess.run([str(BIN / "gfa_2d")], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 1, (r.stdout, r.stderr)
    assert "x: n=600  d=0.0003" in r.stdout
    assert "IO::readField: cannot open file: settings/initial_field/c.field" in r.stderr, r.stderr


def test_composite_forms_host_path(tmp_path):
    """Composite DSL forms (grad/lap of expressions, div of a composite vector
    expression, Term*Term, VectorField divergence) on non-uniform fields, host
    path, against the numpy restatement; periodic and zero-flux BCs."""
    import sys
    sys.path.insert(0, str(ROOT / "tests"))
    from helpers import read_field
    from oracle import dsl
    for mode in ("periodic", "noflux"):
        r = subprocess.run([str(BIN / "composite"), "cpu", str(tmp_path), mode], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        ref = dsl.composite_forms(*dsl.composite_inputs(), 0.5, 0.25, mode)
        for name, want in ref.items():
            _, got = read_field(tmp_path / f"{name}_{mode}_cpu.field")
            err = np.abs(got[0, 0] - want).max()
            scale = max(1.0, np.abs(want).max())
            print(f"{mode:8s} {name:10s} max|host - numpy| = {err:.2e} (scale {scale:.1f})")
            assert err <= 1e-12 * scale, (mode, name, err)
