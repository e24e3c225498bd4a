"""The grid barrier's group arithmetic and protocol, on the CPU.

Builds tests/native/grid_sync_model.cpp against the library's own
grid_sync.h (host-compiled by hipcc) and runs it: member counts per group for
every grid size up to 8192 workgroups (2048 included: 8 four-wave workgroups
per CU on 256 CUs), and a thread model of the arrival/release protocol at
grid sizes up to 2048 that checks no workgroup leaves a barrier early and
every counter returns to zero.
"""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if pathlib.Path("/opt/rocm/bin/hipcc").exists() else None)


@pytest.mark.skipif(HIPCC is None, reason="no hipcc")
def test_grid_sync_arithmetic_and_protocol(tmp_path):
    exe = tmp_path / "grid_sync_model"
    build = subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O1", "-std=c++17", "-pthread",
                            f"-I{ROOT / 'include'}", f"-I{ROOT / 'perks_amd' / 'csrc'}",
                            str(ROOT / "tests" / "native" / "grid_sync_model.cpp"), "-o", str(exe)],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0 and run.stdout.strip().endswith("ok"), run.stdout + run.stderr
