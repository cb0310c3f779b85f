"""GPU: every launch plan the autotuner can choose keeps parity.

The tuning table (sedi_amd/tuning.txt) sets per kernel and grid the planes per
tile (8, 16, 32) and the schedule (0: resident workgroups walk XCD-contiguous
runs of tiles; 1: tiles in launch order, more workgroups; 2: one workgroup
per tile where a kernel has no partial sums, else as 1; 3/4: launch order with
the tiles dealt to the XCDs in chunks of 4/8 tile rows), and the rows per
tile of the FP32 batched CG update pass (8, or 16 by default from 256 rows up).
The test grids have no entries of their own, so here a child process forces
one plan on every kernel (SEDI_TUNING_FILE, read once per process; two plans
also cap the plain kernels at one workgroup per CU, "wpc:* 1"; one plan
takes 16-row update tiles on the batched grid of 20 rows, a full tile and a
partial one) and the results are checked
against the reference fixtures with the bars of the parity suite: stencil
bit-exact, plain estimates within +-1 iteration and 1e-10 (1e-8 when the
counts differ; BiCGStab within the reference's own summation-order spread,
tests/golden/ref_spread.npz), fixed-iteration runs exact in counts
and within 1e-10, batched mixed runs converged on the true residual.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import ROOT, d_bar, iteration_bar, normwise

pytestmark = pytest.mark.gpu

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(tests)r)
import sedi_amd as sedi
import oracle_lib as O
out = {}
shape = (70, 40, 33)
x = np.random.default_rng(3).standard_normal(70 * 40 * 33)
out["stencil_exact"] = bool(np.array_equal(sedi.Stencil().apply(shape, x), O.stencil_apply(shape, x)))
for kind in (1, 2, 3):
    r = sedi.estimate((32, 32, 32), kind=kind, samples=8, maxiter=500, tol=1e-9, seed=1234)
    np.save(%(tmp)r + "/D%%d.npy" %% kind, r["D"])
    f = sedi.estimate((32, 32, 32), kind=kind, samples=8, maxiter=20, tol=1e-9, seed=1234, ignore_errors=True)
    np.save(%(tmp)r + "/F%%d.npy" %% kind, f["D"])
    b = sedi.estimate((24, 20, 18), kind=kind, samples=8, maxiter=2000, tol=1e-9, seed=5,
                      precision=sedi.FP32_MIXED, batch_k=8)
    np.save(%(tmp)r + "/B%%d.npy" %% kind, b["D"])
    out[str(kind)] = dict(iters=r["iters"].tolist(), codes=r["codes"].tolist(), status=int(r["status"]),
                          fiters=f["iters"].tolist(), fcodes=f["codes"].tolist(),
                          bcodes=b["codes"].tolist(), berrs=b["errs"].tolist())
print(json.dumps(out))
"""

# z chunk, schedule, update-pass rows (0: default); schedules 3 and 4 deal
# the tiles to the XCDs in chunks of 4 and 8 tile rows (2 is one workgroup per
# tile, which only the stencil takes; elsewhere it acts as 1)
PLANS = [(8, 1, 0, 0), (32, 0, 16, 0), (16, 1, 0, 0), (16, 3, 0, 0), (8, 4, 16, 0), (32, 2, 0, 0), (16, 3, 0, 1),
         (8, 0, 0, 1)]


@pytest.fixture(scope="module")
def batched_default():
    """The batched mixed estimates with the default plan (this process)."""
    import sedi_amd as sedi
    return {k: sedi.estimate((24, 20, 18), kind=k, samples=8, maxiter=2000, tol=1e-9, seed=5,
                             precision=sedi.FP32_MIXED, batch_k=8)["D"] for k in (1, 2, 3)}


@pytest.mark.parametrize("zc,sched,ty,wpc", PLANS)
def test_forced_launch_plan_keeps_parity(tmp_path, ref32, ref_spread, batched_default, zc, sched, ty, wpc):
    table = tmp_path / "tuning.txt"
    # the stencil's loads in flight vary with the plan too (2, 4 or 8 planes);
    # wpc 1 launches one workgroup per CU for every plain kernel (they loop
    # over the tiles)
    pf = {2: 8, 4: 4}.get(sched, 2)
    table.write_text("zc:* 0 0 0 %d\nsched:* 0 0 0 %d\npf:stencil_apply_kernel 0 0 0 %d\n" % (zc, sched, pf) +
                     ("ty:batch_upd_kernel 0 0 0 %d\n" % ty if ty else "") + ("wpc:* 0 0 0 %d\n" % wpc if wpc else ""))
    env = dict(os.environ, SEDI_TUNING_FILE=str(table))
    code = CHILD % dict(root=ROOT, tests=os.path.join(ROOT, "tests"), tmp=str(tmp_path))
    print("plan zc=%d sched=%d ty=%d wpc=%d" % (zc, sched, ty, wpc), flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["stencil_exact"]
    for kind in (1, 2, 3):
        o = out[str(kind)]
        D = np.load(tmp_path / ("D%d.npy" % kind))
        ref_it = ref32[f"s{kind}_iters"]
        assert o["status"] == 0 and all(c == 0 for c in o["codes"])
        it = np.array(o["iters"])
        assert np.all(np.abs(it - ref_it) <= iteration_bar(ref_spread, "c32", kind)), (it, ref_it)
        assert normwise(D, ref32[f"s{kind}_D"]) < d_bar(ref_spread, "c32", kind, np.array_equal(it, ref_it))
        assert o["fiters"] == ref32[f"s{kind}_fixed20_iters"].tolist()
        assert o["fcodes"] == ref32[f"s{kind}_fixed20_codes"].tolist()
        F = np.load(tmp_path / ("F%d.npy" % kind))
        assert normwise(F, ref32[f"s{kind}_fixed20_D"]) < 1e-10
        # batched mixed: converged on the true residual, D as with the default plan
        assert all(c == 0 for c in o["bcodes"]) and max(o["berrs"]) <= 1e-9
        B = np.load(tmp_path / ("B%d.npy" % kind))
        assert normwise(B, batched_default[kind]) < 1e-6
