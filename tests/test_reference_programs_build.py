"""The drop-in boundary, pinned: the reference's own programs -- its tutorial,
its applications and its development drivers, as they lie in the reference
tree -- compile and link unchanged against include/phix and
libphix_amd.so. Build only (no GPU here); nothing is copied from the
reference. Skipped where the reference tree is absent."""
import concurrent.futures as cf
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
PROGRAMS = [
    "tutorials/quickstart/main.cu",
    "applications/solvers/Cahn-Hillard_double-well/2D/Cahn-Hillard_double-well.cu",
    "applications/solvers/Cahn-Hillard+Allen-Cahn_double-well/2D/CH+AC_double-well.cu",
    "applications/solvers/glass_formation/2D/GFA.cu",
    "develop/Ostwald_Ripening/AC-CH.cu",
    "develop/Spinodal Decomposition/Cahn-Hillard.cu",
    "develop/Spinodal Decomposition 2/Cahn-Hillard.cu",
    "develop/fieldOperation/test_fieldops.cu",
]

pytestmark = pytest.mark.skipif(not REF.is_dir(), reason="reference tree not present")


@pytest.mark.timeout(600)
def test_reference_programs_compile_and_link_unchanged(tmp_path):
    lib = ROOT / "phix_amd" / "libphix_amd.so"
    assert lib.exists(), "build the library first (python -m phix_amd.build)"

    def build(i_rel):
        i, rel = i_rel
        exe = tmp_path / f"prog{i}"
        cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "--offload-arch=gfx950", "-x", "hip", str(REF / rel),
               f"-I{ROOT / 'include' / 'phix'}", f"-L{lib.parent}", "-lphix_amd", f"-Wl,-rpath,{lib.parent}",
               "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        return rel, r.returncode, (r.stdout + r.stderr)[-2000:], exe.exists()

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        results = list(ex.map(build, enumerate(PROGRAMS)))
    failed = [(rel, log) for rel, rc, log, ok in results if rc != 0 or not ok]
    assert not failed, failed
