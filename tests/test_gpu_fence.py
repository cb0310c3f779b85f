"""GPU: the cross-workgroup reduction's publication, with and without fences.

Every fused pass ends in grid_sum (sedi_amd/csrc/march.h): each workgroup
publishes its partial sums, the last one to arrive adds them all in a fixed
order.  The default build publishes with agent-scope atomic stores, waits for
their completion (s_waitcnt 0) and then increments the arrival counter; the
last workgroup reads the partials with agent-scope atomic loads -- no release
or acquire fence, because on gfx950 an agent-scope release writes back the
XCD's whole L2.  That ordering rests on the hardware (a completed agent-scope
store is visible to agent-scope loads of every XCD) rather than on the C++
memory model.  sedi_amd/variants/libsedi_amd_fence.so (built by build())
uses the memory model's own guarantee instead: plain stores, __threadfence()
(release) before the counter increment, __threadfence() (acquire) after the
last arrival.  Both publish the same partials in the same order, so every
result must be bitwise the same: plain and batched solves of every solver,
over many launches and workgroups.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

VARIANT = os.path.join(ROOT, "sedi_amd", "variants", "libsedi_amd_fence.so")

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, %(root)r)
import sedi_amd as sedi
out = {}
for kind in (1, 2, 3):
    r = sedi.estimate((48, 40, 36), kind=kind, samples=4, maxiter=400, tol=1e-9, seed=3, batch_k=1)
    np.save(%(tmp)r + "/plain%%d.npy" %% kind, r["D"])
    out["plain%%d" %% kind] = r["iters"].tolist()
    r = sedi.estimate((48, 40, 36), kind=kind, samples=8, maxiter=400, tol=1e-9, seed=3,
                      precision=sedi.FP32_MIXED, batch_k=8)
    np.save(%(tmp)r + "/mixed%%d.npy" %% kind, r["D"])
    out["mixed%%d" %% kind] = r["iters"].tolist()
print(json.dumps(out))
"""


def run_child(lib, tmp):
    # both children read the package's table (a variant's folder holds none:
    # it would run the untuned default plans, another summation order)
    env = dict(os.environ, SEDI_TUNING_FILE=os.path.join(ROOT, "sedi_amd", "tuning.txt"))
    if lib:
        env["SEDI_AMD_LIB"] = lib
    os.makedirs(tmp, exist_ok=True)
    r = subprocess.run([sys.executable, "-c", CHILD % dict(root=ROOT, tmp=tmp)], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_fenced_reduction_gives_bitwise_the_same_results(tmp_path):
    if not os.path.exists(VARIANT):
        pytest.skip("fence variant not built (sedi_amd.build.build_variant('fence'))")
    print("default build", flush=True)
    a = run_child(None, str(tmp_path / "default"))
    print("fenced build", flush=True)
    b = run_child(VARIANT, str(tmp_path / "fence"))
    assert a == b
    for name in a:
        Da = np.load(tmp_path / "default" / (name + ".npy"))
        Db = np.load(tmp_path / "fence" / (name + ".npy"))
        assert np.array_equal(Da, Db), name
