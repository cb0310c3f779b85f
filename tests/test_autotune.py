"""CPU: the autotuner's table writing, with the GPU measurements replaced by
fixed timings: per kernel and grid the (chunk, schedule) of least launch time,
per solver/precision/grid the batch size of most probe-iterations per second,
and --k-only re-measuring the batch sizes while keeping every other entry."""
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
autotune = pytest.importorskip("sedi_amd.autotune")

# launch times (us) per (chunk, schedule), grid side and kernel
TIMES = {
    (8, 0): {"64": {"batch_upd_kernel": 70.0, "cg_update_kernel": 9.0, "batch_dir_kernel": 30.0}},
    (8, 1): {"64": {"batch_upd_kernel": 69.0, "cg_update_kernel": 12.0, "batch_dir_kernel": 30.0}},
    (16, 0): {"64": {"batch_upd_kernel": 60.0, "cg_update_kernel": 10.0, "batch_dir_kernel": 30.0}},
    (16, 1): {"64": {"batch_upd_kernel": 61.0, "cg_update_kernel": 11.0, "batch_dir_kernel": 30.0}},
    (32, 0): {"64": {"batch_upd_kernel": 65.0, "cg_update_kernel": 13.0, "batch_dir_kernel": 30.0}},
    (32, 1): {"64": {"batch_upd_kernel": 58.5, "cg_update_kernel": 14.0, "batch_dir_kernel": 30.0}},
}
for zc in (8, 16, 32):
    for sched in (3, 4):
        TIMES[(zc, sched)] = {"64": {"batch_upd_kernel": 80.0, "cg_update_kernel": 20.0,
                                     "batch_dir_kernel": 29.0 if (zc, sched) == (16, 3) else 31.0}}
BATCHES = [dict(side=64, kind=2, precision=1, k=32, rates={"8": 1.0, "32": 3.0, "64": 2.0}),
           dict(side=64, kind=2, precision=0, k=1, rates={"1": 5.0, "8": 4.0})]


def stencil_time(x2, zc, sched, pf, pair_wins=True):
    # fastest one-cell plan: chunk 32, one workgroup per tile, prefetch 4;
    # fastest pair plan: chunk 16, schedule 3, prefetch 2 (ahead of the
    # one-cell plans when pair_wins)
    if x2:
        return (95.0 if pair_wins else 105.0) + abs(zc - 16) + 3 * (sched != 3) + abs(pf - 2)
    return 100.0 + abs(zc - 32) + 3 * abs(sched - 2) + abs(pf - 4)


def fake_child(calls, pair_wins=True):
    def child(args, table_lines, mode):
        calls.append((mode, list(table_lines)))
        if mode == "batches":
            return BATCHES
        if mode == "stencil":
            x2, zc, sched, pf = (int(l.split()[-1]) for l in table_lines)
            return {"64": stencil_time(x2, zc, sched, pf, pair_wins)}
        zc = int(table_lines[0].split()[-1])
        sched = int(table_lines[1].split()[-1])
        return TIMES[(zc, sched)]
    return child


def entries(path):
    out = {}
    for line in open(path):
        line = line.split("#")[0].split()
        if line:
            out[(line[0], int(line[1]))] = int(line[4])
    return out


def test_table_takes_fastest_plan_and_best_batch(tmp_path, monkeypatch):
    calls = []
    monkeypatch.setattr(autotune, "child", fake_child(calls))
    out = tmp_path / "tuning.txt"
    monkeypatch.setattr(sys, "argv", ["autotune", "--sides", "64", "--out", str(out)])
    autotune.main()
    e = entries(out)
    assert e[("zc:batch_upd_kernel", 64)] == 32 and e[("sched:batch_upd_kernel", 64)] == 1
    assert e[("zc:cg_update_kernel", 64)] == 8 and e[("sched:cg_update_kernel", 64)] == 0
    assert e[("k:2:1", 64)] == 32 and e[("k:2:0", 64)] == 1
    # the chunked XCD order is a schedule choice like the others
    assert e[("zc:batch_dir_kernel", 64)] == 16 and e[("sched:batch_dir_kernel", 64)] == 3
    # the stencil's own plan: march, chunk, schedule and prefetch, from its
    # cold timings (here the pair march's 16/3/2)
    assert e[("x2:stencil_apply_kernel", 64)] == 1
    assert e[("zc:stencil_apply_kernel", 64)] == 16 and e[("sched:stencil_apply_kernel", 64)] == 3
    assert e[("pf:stencil_apply_kernel", 64)] == 2
    # twelve plans measured, then the batch sizes under the chosen plans, then
    # the 27 one-cell and 18 pair stencil plans
    assert [c[0] for c in calls] == ["chunks"] * 12 + ["batches"] + ["stencil"] * 45
    assert "zc:batch_upd_kernel 64 64 64 32" in [l.strip() for l in calls[12][1]]


def test_k_only_keeps_plans(tmp_path, monkeypatch):
    out = tmp_path / "tuning.txt"
    out.write_text("# header\nzc:batch_upd_kernel 64 64 64 16   # note\nsched:batch_upd_kernel 64 64 64 0\n"
                   "k:2:1 64 64 64 8   # old\n")
    calls = []
    monkeypatch.setattr(autotune, "child", fake_child(calls))
    monkeypatch.setattr(sys, "argv", ["autotune", "--sides", "64", "--k-only", "--out", str(out)])
    autotune.main()
    e = entries(out)
    assert e[("zc:batch_upd_kernel", 64)] == 16 and e[("sched:batch_upd_kernel", 64)] == 0
    assert e[("k:2:1", 64)] == 32 and e[("k:2:0", 64)] == 1
    assert [c[0] for c in calls] == ["batches"]
    text = out.read_text()
    assert text.startswith("# header\n") and "k entries re-measured by --k-only" in text
    assert text.count("k:2:1 ") == 1


@pytest.mark.parametrize("pair_wins", [True, False])
def test_stencil_only_keeps_everything_else(tmp_path, monkeypatch, pair_wins):
    out = tmp_path / "tuning.txt"
    out.write_text("# header\nzc:batch_upd_kernel 64 64 64 16   # note\nk:2:1 64 64 64 8\n"
                   "zc:stencil_apply_kernel 64 64 64 8\nsched:stencil_apply_kernel 64 64 64 0\n"
                   "x2:stencil_apply_kernel 64 64 64 1\n")
    calls = []
    monkeypatch.setattr(autotune, "child", fake_child(calls, pair_wins))
    monkeypatch.setattr(sys, "argv", ["autotune", "--sides", "64", "--stencil-only", "--out", str(out)])
    autotune.main()
    e = entries(out)
    assert e[("zc:batch_upd_kernel", 64)] == 16 and e[("k:2:1", 64)] == 8
    plan = (e[("x2:stencil_apply_kernel", 64)], e[("zc:stencil_apply_kernel", 64)],
            e[("sched:stencil_apply_kernel", 64)], e[("pf:stencil_apply_kernel", 64)])
    assert plan == ((1, 16, 3, 2) if pair_wins else (0, 32, 2, 4))
    assert [c[0] for c in calls] == ["stencil"] * 45
    text = out.read_text()
    assert text.count("zc:stencil_apply_kernel") == 1 and text.count("x2:stencil_apply_kernel") == 1
    # the note names the best one-cell plan beside the pair march
    assert "best one-cell plan 32/2/4:100.0" in text
