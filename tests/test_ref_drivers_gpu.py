"""The reference's own drivers (j3d.driver.cpp, jacobi.driver.cpp), built
unmodified against the drop-in header and linked with libperks_amd.so
(oracle/_ref, built by `make -C oracle drivers` where the PERKS tree is
present), run on the GPU with their own --check: the driver computes the
reference's CPU gold on the same glibc-rand() input and fails (exit -1) when
the RMS error exceeds its TOLERANCE of 1e-5.  The drop-in runs the clamp
boundary with the reference's coefficients, so the error here is exactly 0."""
import pathlib
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

OUT = pathlib.Path(__file__).resolve().parent.parent / "oracle" / "_ref"


def run(exe, *args):
    path = OUT / exe
    if not path.exists():
        pytest.skip(f"{exe} is not built (make -C oracle drivers)")
    print(f"[ref-driver] {exe} {' '.join(args)}", flush=True)
    res = subprocess.run([str(path), *args], capture_output=True, text=True, timeout=120)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    return res


def rms(res):
    m = re.search(r"\[Test\] RMS Error : ([-+0-9.eE]+)", res.stdout)
    assert m, f"no RMS line in the driver's output: {res.stdout!r}"
    return float(m.group(1))


@pytest.mark.parametrize("prec", [[], ["--fp32"]])
@pytest.mark.parametrize("dims,iters", [(("64", "64", "64"), 8), (("37", "50", "130"), 5)])
def test_reference_3d_driver_checks_against_its_gold(gpu, dims, iters, prec):
    res = run("j3d_driver_3d7pt", *dims, f"--iter={iters}", "--check", *prec)
    assert res.returncode == 0, res.stdout
    assert rms(res) == 0.0


@pytest.mark.parametrize("prec", [[], ["--fp32"]])
@pytest.mark.parametrize("dims,iters", [(("256", "256"), 8), (("100", "333"), 7)])
def test_reference_2d_driver_checks_against_its_gold(gpu, dims, iters, prec):
    res = run("jacobi_driver_2d5pt", *dims, f"--iter={iters}", "--check", *prec)
    assert res.returncode == 0, res.stdout
    assert rms(res) == 0.0


def test_reference_3d_driver_small_domain_search(gpu):
    """--small: the driver asks j3d_iterative(..., getminHeight=true) for the
    height, then runs and checks that domain (j3d.driver.cpp:62-144)."""
    res = run("j3d_driver_3d7pt", "8", "64", "64", "--small", "--iter=2", "--check", "--fp32")
    assert res.returncode == 0, res.stdout
    assert rms(res) == 0.0


def test_reference_drivers_checkmindomain(gpu):
    res = run("j3d_driver_3d7pt", "64", "128", "256", "--checkmindomain")
    assert res.returncode == 0
    h, wy, wx = map(int, res.stdout.split()[:3])
    assert (wy, wx) == (128, 256) and h > 3
    res = run("jacobi_driver_2d5pt", "128", "512", "--checkmindomain")
    assert res.returncode == 0
    wx, wy = map(int, res.stdout.split()[:2])
    assert wx == 512 and wy > 0
