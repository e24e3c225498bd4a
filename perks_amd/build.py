"""Builds libperks_amd.so (HIP, gfx950) in-tree.

    python -m perks_amd.build            # compile every HIP source and link
    python -m perks_amd.build --changed  # recompile only sources whose object is
                                         # older than the source or a header it includes

Every source under perks_amd/csrc/ (*.hip: device + host, *.cpp: host only) is
compiled with hipcc for gfx950 and linked into perks_amd/libperks_amd.so.  Flags that matter for results:
  -ffp-contract=off      no FMA contraction: every * and + rounds on its own,
                         the condition for bit-identity with the CPU gold;
  -packed-fp32-ops off   no v_pk_*_f32 instructions are generated.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libperks_amd.so"

ARCH = os.environ.get("PERKS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and pathlib.Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; set HIPCC")


COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _up_to_date(src: pathlib.Path, obj: pathlib.Path) -> bool:
    """True when obj and its dependency list exist and nothing listed is newer."""
    dep = obj.with_suffix(".d")
    if not obj.exists() or not dep.exists():
        return False
    t = obj.stat().st_mtime
    text = dep.read_text().replace("\\\n", " ")
    files = [f for f in text.split(":", 1)[-1].split() if f]
    for f in [str(src), *files]:
        try:
            if pathlib.Path(f).stat().st_mtime > t:
                return False
        except OSError:
            return False
    return True


def _compile(src: pathlib.Path, changed_only: bool = False) -> pathlib.Path:
    obj = BUILD / (src.stem + ".o")
    if changed_only and _up_to_date(src, obj):
        return obj
    cmd = [hipcc(), *COMMON_FLAGS, "-MMD", "-MF", str(obj.with_suffix(".d")), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; drop that line.
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")
    return obj


def build(verbose: bool = True, changed_only: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(lambda src: _compile(src, changed_only), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    return LIB


if __name__ == "__main__":
    build(changed_only="--changed" in sys.argv[1:])
