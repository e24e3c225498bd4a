"""The C ABI as a library: loads without a GPU, exports every declared symbol,
struct layouts agree with the ctypes mirror, the C++ drop-in header compiles
against the reference's signatures, and the planner's refusals are pure host
logic."""
import ctypes as C
import pathlib
import re
import shutil
import subprocess

import pytest

import perks_amd as P
from perks_amd import _lib as L

ROOT = pathlib.Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "perks_amd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(perks_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(str(L.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_covers_the_header():
    assert set(declared_functions()) == set(L.SIGNATURES)


STRUCTS = {"perks_stencil_opts": L.StencilOpts, "perks_run_report": L.RunReport, "perks_grid2d": L.Grid2d,
           "perks_device_caps": L.DeviceCaps, "perks_plan_info": L.PlanInfo,
           "perks_kernel_config": L.KernelConfig, "perks_csr": L.Csr, "perks_dia": L.Dia,
           "perks_cg_opts": L.CgOpts, "perks_cg_report": L.CgReport}


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof and every field offset of the ctypes mirror against the C compiler's."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "perks_amd.h"', 'int main(void) {']
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    seen = 0
    for line in filter(None, out):
        cname, what, value = line.split()
        py = STRUCTS[cname]
        want = C.sizeof(py) if what == "sizeof" else getattr(py, what).offset
        assert int(value) == want, f"{cname}.{what}: C {value}, ctypes {want}"
        seen += 1
    assert seen == sum(len(p._fields_) + 1 for p in STRUCTS.values())


def test_error_strings_and_version():
    lib = P.load()
    assert b"gfx950" in lib.perks_version()
    for code in (0, -1, -2, -3, -4, -5, -10, -11, -12, -13, -14):
        assert lib.perks_error_string(code)


def test_shape_table():
    counts = {"3d7pt": 7, "3d13pt": 13, "3d17pt": 17, "3d27pt": 27, "poisson": 19, "2d5pt": 5, "2ds9pt": 9,
              "2d13pt": 13, "2d17pt": 17, "2d21pt": 21, "2ds25pt": 25, "2d9pt": 9, "2d25pt": 25}
    for ref, npts in counts.items():
        si = P.shape_info(ref)
        pts = [pt for g in si.offsets for pt in g]
        assert len(pts) == npts, ref
        assert len(set(pts)) == npts, ref
        assert max(max(abs(c) for c in pt) for pt in pts) == si.radius
        if si.dims == 2:
            assert all(pt[0] == 0 for pt in pts)


def caps(cu=256):
    c = L.DeviceCaps()
    c.cu_count, c.lds_per_cu, c.lds_per_block, c.vgprs_per_simd, c.simds_per_cu = cu, 163840, 163840, 512, 4
    c.warp_size, c.cooperative = 64, 1
    return c


def test_planner_is_host_only_and_refuses_like_the_reference():
    rc, info = P.plan(512, 512, 512, caps=caps(), occupancy=1, method="persistent")
    assert rc == 0 and info.grid[0] == 256 and info.block == 12 * 64  # FP64 default: 12 waves x 3 rows
    rc, _ = P.plan(512, 512, 512, caps=caps(), occupancy=0, method="persistent")
    assert rc == L.ERR_NOT_LAUNCHABLE                       # j3d.cu:570
    rc, info = P.plan(64, 64, 64, caps=caps(), occupancy=1, method="persistent", shape="3d27pt")
    assert rc == 0 and info.reserved[0] == 2 and info.tile_x == 64      # shape-generic kernel: 2 steps per pass
    rc, info = P.plan(64, 64, 64, caps=caps(), occupancy=1, method="persistent", shape="3d13pt")
    assert rc == 0 and info.reserved[0] == 1                            # radius 2: one step per pass
    rc, info = P.plan(512, 512, 512, caps=caps(), occupancy=2, method="plain")
    assert rc == 0 and info.grid[0] >= 256


def test_planner_balances_items_over_workgroups():
    # 512^3, 4 steps per pass, the clamp boundary: tiles of 64 x tile_y lanes
    # store 56 columns and tile_y - 6 rows (tile_y - 3 at the y faces); the z
    # split must not leave a lone round for a handful of items.
    rc, info = P.plan(512, 512, 512, caps=caps(), occupancy=1, method="persistent")
    rows = info.tile_y
    ntx = -(-512 // 56)
    nty = 2 + -(-(512 - 2 * (rows - 3)) // (rows - 6))
    nzs = info.reserved[4]
    items = ntx * nty * nzs
    rounds = -(-items // 256)
    assert items / (rounds * 256) > 0.9


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_cpp_dropin_header_compiles_and_links(tmp_path):
    src = tmp_path / "dropin.cpp"
    src.write_text(r'''
#include "perks_amd/jacobi_cuda.hpp"
#include <cstdio>
int main(int argc, char**) {
    if (argc < 0) {  // never taken: link check only, no device call
        double* p = nullptr;
        j3d_iterative<double>(p, 8, 8, 8, p, 256, 0, 3, false, false, 0, false);
        j3d_iterative<float>(nullptr, 8, 8, 8, nullptr, 256, 0, 3, true, false, 0, false, true);
        getMinWidthY<double>(64, 64, 256);
        jacobi_iterative<float>(nullptr, 8, 8, nullptr, 256, 0, 3, false, false, false, 0, false);
    }
    std::printf("%s\n", perks_version());
    return 0;
}
''')
    exe = tmp_path / "dropin"
    subprocess.run(["g++", "-std=c++14", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                    str(L.LIB_PATH), f"-Wl,-rpath,{L.LIB_PATH.parent}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert "gfx950" in out


def test_cg_driver_dropin_header_compiles_and_links(tmp_path):
    """The reference's conjugateGradient main calls myTest<ValueT,int,...>(devID,
    deviceProp, args); the drop-in header provides it over the C ABI."""
    src = tmp_path / "cgmain.cpp"
    src.write_text(r'''
#include "perks_amd/cg_driver.hpp"
#include <cstdio>
int main(int argc, char** argv) {
    CommandLineArgs args(argc, argv);
    perks_device_caps prop{};
    if (args.CheckCmdLineFlag("never-given")) {  // link check only, no device call
        myTest<double, int, true, false, false>(0, prop, args);
        myTest<float, int, false, true, true>(0, prop, args);
    }
    int iters = -1;
    args.GetCmdLineArgument("iters", iters);
    std::printf("%d %d\n", iters, (int)args.CheckCmdLineFlag("check"));
    return 0;
}
''')
    exe = tmp_path / "cgmain"
    subprocess.run(["g++", "-std=c++14", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
                    str(L.LIB_PATH), f"-Wl,-rpath,{L.LIB_PATH.parent}"], check=True)
    out = subprocess.run([str(exe), "--iters=37", "--check", "--mtx=x.mtx"], capture_output=True, text=True,
                         check=True).stdout
    assert out.split() == ["37", "1"]
