"""CPU checks of bench.py's contract and byte accounting (no GPU): the
--gpus guard, the path-cell split, the roofline assembled from the timed
run, and the kernel-trace summary on a synthetic trace."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_path_cells_split():
    # counted: full and short as given, the rest skipped
    assert bench.path_cells(1000, 60, 100) == (60, 100)
    # short not counted: every other cell took the short path
    assert bench.path_cells(1000, 60, -1) == (60, 940)
    # full not counted: every cell took the full path
    assert bench.path_cells(1000, -1, -1) == (1000, 0)


def _kin(full, short, bpl):
    return {"full_path_share": full, "short_path_share": short, "bytes_per_launch": bpl}


def test_replay_roofline_uses_the_timed_run():
    kin = {"coef": _kin(0.03, 0.06, 1e8), "update": _kin(0.06, 0.13, 3e8),
           "adjoint": _kin(0.064, 0.041, 2.6e8)}
    # checkpoints only, K = 500, k = 25: the reverse sweep recomputes 24
    # steps of each of its 20 segments; the events time the recompute apart
    r = {"fwd_ms": 114.0, "adj_ms": 290.0, "rec_ms": 108.5, "stored_steps": 0}
    dom, rl = bench.replay_roofline(r, kin, 500, 25)
    assert rl["adjoint"]["avg_launch_ms"] == pytest.approx((290.0 - 108.5) / 500)
    assert rl["adjoint"]["gbs"] == pytest.approx(2.6e8 / (0.363e-3) / 1e9)
    assert rl["forward_step"]["avg_launch_ms"] == pytest.approx(0.228)
    assert rl["forward_step"]["bytes_per_launch"] == 3e8
    assert rl["recompute_step"]["launches_timed"] == 480
    assert rl["recompute_step"]["avg_launch_ms"] == pytest.approx(108.5 / 480)
    assert dom == "adjoint"  # 0.363 ms > 0.228 ms
    # the whole trajectory kept: nothing recomputed
    r = {"fwd_ms": 114.0, "adj_ms": 100.0, "rec_ms": 0.0, "stored_steps": 500}
    dom, rl = bench.replay_roofline(r, kin, 500, 25)
    assert dom == "forward_step" and "recompute_step" not in rl


def test_recomputed_steps():
    assert bench.recomputed_steps(20, 25, 0) == 19
    assert bench.recomputed_steps(500, 25, 0) == 480
    assert bench.recomputed_steps(500, 25, 441) == (59 - 3)   # segments of 25, 25, 9
    assert bench.recomputed_steps(500, 25, 500) == 0


def test_path_rates():
    kin = {"update": _kin(0.06, 0.14, 0), "adjoint": _kin(0.04, 0.06, 0)}
    r = bench.path_rates(kin, 10, 2.0, 1000)
    assert r["full_path_share"] == pytest.approx(0.05)
    assert r["full_path_cell_steps_per_s"] == pytest.approx(0.05 * 10000 / 2.0)
    assert r["visited_cell_steps_per_s"] == pytest.approx(0.15 * 10000 / 2.0)


def test_roofline_json_keys_and_fraction(tmp_path, monkeypatch):
    from discrete_adjoint_amd import _lib
    kin = {"coef": _kin(0.03, 0.06, 1e8), "update": _kin(0.06, 0.13, 3e8),
           "adjoint": _kin(0.064, 0.041, 2.6e8)}
    r = {"fwd_ms": 114.0, "adj_ms": 290.0, "rec_ms": 108.5, "stored_steps": 0}
    dom, rl = bench.replay_roofline(r, kin, 500, 25)
    j = bench.roofline_json(rl, dom, kin, _lib.TG_FP64, "no_such_workload",
                            run_config={"steps": 500})
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in j
    assert j["bound"] == "hbm" and j["unit"] == "GB/s" and j["peak"] == 8000.0
    assert j["frac"] == pytest.approx(j["achieved"] / 8000.0)
    assert j["traffic"] is None  # no committed summary for this workload
    assert j["algorithmic_bytes_per_cell"] == bench.algo_bytes(_lib.TG_FP64)["adjoint"]
    # a committed summary is used only for the run configuration it was taken on
    import json
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "traffic_adjoint_wl.json").write_text(json.dumps(
        {"hbm_bytes_per_launch": 3.9e8, "algorithmic_bytes_per_launch": 2.6e8,
         "config": {"steps": 500}}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    j = bench.roofline_json(rl, dom, kin, _lib.TG_FP64, "wl", run_config={"steps": 500})
    assert j["traffic"] == 3.9e8
    assert j["traffic_over_algorithmic"] == pytest.approx(1.5)
    j = bench.roofline_json(rl, dom, kin, _lib.TG_FP64, "wl", run_config={"steps": 20})
    assert j["traffic"] is None and "another configuration" in j["traffic_info"]["reason"]


def test_trace_summary_splits_the_last_sweep(tmp_path):
    # forward run 0, reverse 0, forward run 1, reverse 1: k_stats ends each forward run
    seq = (["k_coef<16>", "k_update<16>"] * 3 + ["k_stats"] + ["k_adjoint<16>"] * 2 +
           ["k_coef<16>", "k_update<16>"] * 2 + ["k_stats"] + ["k_adjoint<16>"] * 4)
    path = tmp_path / "trace.csv"
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        t = 1000
        for name in seq:
            w.writerow({"Kernel_Name": f"void tg::(anonymous namespace)::{name}(tg::Args)",
                        "Start_Timestamp": t, "End_Timestamp": t + 2000})
            t += 3000
    out = tmp_path / "summary.md"
    bench.trace_summary(str(path), str(out), "cmd")
    text = out.read_text()
    fwd, rev = text.split("### last reverse sweep")
    assert "5 dispatches" in fwd  # two coef + two update + k_stats
    assert "| `k_coef<16>` | 2 |" in fwd and "k_adjoint" not in fwd
    assert "| `k_adjoint<16>` | 4 | 2.0 |" in rev


def test_counter_summary_averages_the_last_dispatches(tmp_path):
    path = tmp_path / "c.csv"
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for v in (100, 1, 2, 3):  # the first dispatch is not among the last three
            w.writerow({"Kernel_Name": "void tg::k_adjoint<16>(tg::AdjArgs)",
                        "Counter_Name": "SQ_WAVES", "Counter_Value": v})
            w.writerow({"Kernel_Name": "void tg::k_update<16>(tg::StepArgs)",
                        "Counter_Name": "SQ_WAVES", "Counter_Value": 1000})
    out = tmp_path / "s.md"
    avg = bench.counter_summary([str(path)], "k_adjoint", 3, str(out), ["cmd"])
    assert avg == {"SQ_WAVES": 2.0}
    assert "| `SQ_WAVES` | 2 |" in out.read_text()
