"""bench.py's roofline bookkeeping, on the CPU: the pass-structure model of
the algorithmic bytes and the counter-profile recorder (calibration on the
field-sized copies only, the source hash that ties a profile to the kernel
sources it was measured on)."""
import json
import pathlib
import sys
import types

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402

HEAD = ('"Correlation_Id","Dispatch_Id","Agent_Id","Queue_Id","Process_Id","Thread_Id","Grid_Size","Kernel_Id",'
        '"Kernel_Name","Workgroup_Size","LDS_Block_Size","Scratch_Size","VGPR_Count","Accum_VGPR_Count","SGPR_Count",'
        '"Counter_Name","Counter_Value","Start_Timestamp","End_Timestamp"')


def write_csv(d, counter, rows):
    d.mkdir(parents=True, exist_ok=True)
    lines = [HEAD]
    for disp, grid, name, value in rows:
        lines.append(f'{disp},{disp},"Agent 1",1,1,1,{grid},1,"{name}",256,0,0,8,0,32,"{counter}",{value},0,1')
    (d / "run_counter_collection.csv").write_text("\n".join(lines) + "\n")


def test_pass_roofline_prices_one_read_and_write_per_pass():
    r = bench.pass_roofline("k", 100, 10, 4, 8, 1e-6)
    assert r["algorithmic_bytes_per_launch"] == 2 * 8 * 100 * 3  # ceil(10 / 4) passes
    assert r["single_pass_gbs"] == round(2 * 8 * 100 * 10 / 1e-6 / 1e9, 1)
    assert r["frac"] == round(r["achieved"] / bench.HBM_PEAK_GBS, 4)
    b = bench.pass_roofline("k", [10, 20], [5, 8], 4, 4, 1.0)  # a batch: per-grid pass counts
    assert b["algorithmic_bytes_per_launch"] == 2 * 4 * (10 * 2 + 20 * 2)


def test_record_profile_calibrates_on_field_copies(tmp_path):
    n, field = 8, 8 ** 3 * 8
    kern = "void perks::j3dp::j3d_star1_persistent_kernel<0, 1>(...)"
    # FETCH_SIZE reads half the copy's bytes, WRITE_SIZE all of them; a small copy must not count
    write_csv(tmp_path / "f", "FETCH_SIZE", [(1, 4096, "__amd_rocclr_copyBuffer", field / 2048),
                                             (2, 512, "__amd_rocclr_copyBuffer", 3.0),
                                             (3, 131072, kern, 1000.0)])
    write_csv(tmp_path / "w", "WRITE_SIZE", [(1, 4096, "__amd_rocclr_copyBuffer", field / 1024),
                                             (2, 512, "__amd_rocclr_copyBuffer", 0.03),
                                             (3, 131072, kern, 500.0)])
    write_csv(tmp_path / "s", "SQ_INSTS_VALU", [(3, 131072, kern, 64.0)])
    rows = (tmp_path / "s" / "run_counter_collection.csv").read_text().splitlines()
    extra = [rows[1].replace("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU").replace(",64.0,", ",256.0,"),
             rows[1].replace("SQ_INSTS_VALU", "SQ_BUSY_CYCLES").replace(",64.0,", ",32.0,")]
    (tmp_path / "s" / "run_counter_collection.csv").write_text("\n".join(rows + extra) + "\n")
    out = tmp_path / "profile.json"
    args = types.SimpleNamespace(record_profile=[str(tmp_path / d) for d in "fws"], size=n, tsteps=10, dtype="f64")
    bench.record_profile(args, out_path=out)
    d = json.loads(out.read_text())
    assert d["fetch_factor"] == 2.0 and d["write_factor"] == 1.0
    assert d["hbm_bytes_per_launch"] == 1000 * 1024 * 2 + 500 * 1024
    assert d["source_sha256"] == bench.kernel_sources_sha()
    assert d["valu_lane_insts_per_cell"] == round(64 * 64 / (10 * n ** 3), 2)
    assert d["valu_busy"] == round(4 * 256 / (bench.SIMDS * 32 / bench.SHADER_ENGINES), 4)


def test_committed_profile_matches_the_kernel_sources():
    d = json.loads(bench.TRAFFIC_FILE.read_text())
    assert d["source_sha256"] == bench.kernel_sources_sha(), \
        "profiles/j3d_persistent_f64_512_traffic.json was recorded for other kernel sources: re-record it"


def test_committed_contracted_profile_matches_the_kernel_sources():
    d = json.loads(bench.TRAFFIC_FILE_FUSED.read_text())
    assert d["contracted"] is True
    assert d["source_sha256"] == bench.kernel_sources_sha(), \
        "profiles/j3d_persistent_fused_f64_512_traffic.json was recorded for other kernel sources: re-record it"


def test_record_profile_refuses_mixed_kernels(tmp_path):
    k0 = "void perks::j3dp::j3d_star1_persistent_kernel<0, 1, false>(...)"
    k1 = "void perks::j3dp::j3d_star1_persistent_kernel<0, 1, true>(...)"
    for d, c in (("f", "FETCH_SIZE"), ("w", "WRITE_SIZE")):
        write_csv(tmp_path / d, c, [(1, 4096, "__amd_rocclr_copyBuffer", 8.0), (2, 131072, k0, 1.0),
                                    (3, 131072, k1, 1.0)])
    args = types.SimpleNamespace(record_profile=[str(tmp_path / d) for d in "fw"], size=8, tsteps=10, dtype="f64")
    with pytest.raises(SystemExit):
        bench.record_profile(args, out_path=tmp_path / "p.json")
