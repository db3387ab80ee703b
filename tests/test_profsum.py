"""The rocprofv3 summarizer (pspaqr_public_amd.profsum) on small synthetic result files."""
import csv
import json

from pspaqr_public_amd import profsum

COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size", "Kernel_Id",
        "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count",
        "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f, quoting=csv.QUOTE_NONNUMERIC)
        w.writerow(COLS)
        for disp, name, counter, value in rows:
            w.writerow([disp, disp, "Agent 1", 1, 1, 1, 64, 1, name, 64, 0, 0, 8, 0, 16, counter, value, 0, 1])


def test_short_names_and_classes():
    assert profsum.short("void spaqr::k_vec_chain<float, 1024>(spaqr::VecChainTask<float> const*)") == \
        "k_vec_chain<float, 1024>"
    assert profsum.short("spaqr::(anonymous namespace)::k_minmax(int, int)") == "k_minmax"
    assert profsum.short("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
    assert profsum.kclass("void spaqr::k_convert_d2s<float>(int, double const*, float*)") == "precond_solve"
    assert profsum.kclass("void spaqr::k_tpqrt_coop<float>(int)") == "tpqrt_chain"
    assert profsum.kclass("void spaqr::k_panel_step<float>(int)") is None  # shared by two classes


def test_traffic_scales_and_sums_per_class(tmp_path, capsys):
    # calibration: a 4 MiB copy reads as 2048 KiB of FETCH_SIZE (x2) and 4096 KiB of WRITE_SIZE (x1)
    _write(tmp_path / "cf.csv", [(1, "fill", "FETCH_SIZE", 1.0), (2, "copy", "FETCH_SIZE", 1024.0),
                                 (2, "copy", "FETCH_SIZE", 1024.0)])  # two rows of one dispatch add up
    _write(tmp_path / "cw.csv", [(1, "copy", "WRITE_SIZE", 4096.0)])
    solve = "void spaqr::k_vec_q<float, 256>(int)"
    _write(tmp_path / "f.csv", [(1, solve, "FETCH_SIZE", 100.0), (2, solve, "FETCH_SIZE", 50.0),
                                (3, "void spaqr::k_panel_step<float>(int)", "FETCH_SIZE", 1000.0)])
    _write(tmp_path / "w.csv", [(1, solve, "WRITE_SIZE", 10.0), (2, solve, "WRITE_SIZE", 0.0),
                                (3, "void spaqr::k_panel_step<float>(int)", "WRITE_SIZE", 1000.0)])
    out = tmp_path / "t.json"
    profsum.main(["traffic", "--fetch-calib", str(tmp_path / "cf.csv"), "--write-calib", str(tmp_path / "cw.csv"),
                  "--calib-mib", "4", "--fetch", str(tmp_path / "f.csv"), "--write", str(tmp_path / "w.csv"),
                  "--steps", "2", "--json", str(out)])
    d = json.load(open(out))
    assert abs(d["fetch_scale"] - 2.0) < 1e-9 and abs(d["write_scale"] - 1.0) < 1e-9
    # (150 KiB x 2 + 10 KiB) over 2 steps; the shared panel kernel is in no class
    assert d["bytes_per_step"]["precond_solve"] == (150 * 2 + 10) * 1024 // 2
    assert d["bytes_per_step"]["tpqrt_chain"] == 0
    assert "precond_solve" in capsys.readouterr().out


def test_mfma_rates(tmp_path, capsys):
    k = "void spaqr::k_larfb_res<16>(int)"
    # 1e9 MOPS -> 512 GFLOP over GRBM 8 x 2.4e9 cycles = 1 s: 512 TFLOP/s would be over the FP32 peak; use 1e8
    _write(tmp_path / "m.csv", [(1, k, "SQ_INSTS_VALU_MFMA_MOPS_F32", 1e8), (1, k, "SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0),
                                (1, k, "SQ_VALU_MFMA_BUSY_CYCLES", 400.0), (1, k, "SQ_BUSY_CU_CYCLES", 1000.0),
                                (1, k, "GRBM_GUI_ACTIVE", 8 * 2.4e9)])
    profsum.main(["mfma", str(tmp_path / "m.csv")])
    line = [x for x in capsys.readouterr().out.splitlines() if "k_larfb_res<16>" in x][0]
    cells = [c.strip() for c in line.split("|")]
    assert cells[3] == "51.2" and cells[4] == "0.05" and cells[6] == "10.0%"
