"""GPU: the f32compute build (FP32 arithmetic in the mixed pipelined CG and
BiCGStab) runs, and it is the variant, not the default, that it measures.

By default the mixed-precision pipelined CG and BiCGStab store their vectors
in FP32 and compute the stencil and vector updates in FP64 registers;
sedi_amd/variants/libsedi_amd_f32compute.so (built by build(),
-DSEDI_MIXED_COMPUTE_F64=0) computes them in FP32, and bench.py --mode arith
compares the two at 512^3.  Here, on a 48x40x36 grid (8 probes of seed 3, tol
1e-8) in a child process that loads the variant:
  * mixed pipelined CG converges on every probe and its D is within 1e-5
    (normwise) of the default build's mixed D; the two builds' D differ
    (FP32 arithmetic is a different computation);
  * mixed BiCGStab returns a status and a code for every probe (at this size it
    may converge or not: the 512^3 measurement is where it fails).
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, normwise

pytestmark = pytest.mark.gpu

VARIANT = os.path.join(ROOT, "sedi_amd", "variants", "libsedi_amd_f32compute.so")

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, %(root)r)
import sedi_amd as sedi
out = {"lib": sedi.lib_path}
for kind in (3, 1):
    r = sedi.estimate((48, 40, 36), kind=kind, samples=8, maxiter=2000, tol=1e-8, seed=3, ignore_errors=True,
                      precision=sedi.FP32_MIXED, batch_k=8)
    np.save(%(tmp)r + "/mixed%%d.npy" %% kind, r["D"])
    out["codes%%d" %% kind] = r["codes"].tolist()
    out["iters%%d" %% kind] = r["iters"].tolist()
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


def test_f32compute_build_runs_and_differs(tmp_path):
    if not os.path.exists(VARIANT):
        pytest.skip("f32compute variant not built (sedi_amd.build.build_variant('f32compute'))")
    print("default build", flush=True)
    a = run_child(None, str(tmp_path / "default"))
    print("f32compute build", flush=True)
    b = run_child(VARIANT, str(tmp_path / "f32"))
    assert os.path.realpath(b["lib"]) == os.path.realpath(VARIANT)
    assert all(c == 0 for c in a["codes3"]) and all(c == 0 for c in a["codes1"])
    assert all(c == 0 for c in b["codes3"]), b["codes3"]
    assert len(b["codes1"]) == 8 and len(b["iters1"]) == 8
    Da = np.load(tmp_path / "default" / "mixed3.npy")
    Db = np.load(tmp_path / "f32" / "mixed3.npy")
    print("pipelined CG: iterations %s (FP64 registers) vs %s (FP32), D normwise %.3g" % (
        a["iters3"], b["iters3"], normwise(Db, Da)), flush=True)
    assert not np.array_equal(Da, Db)
    assert normwise(Db, Da) < 1e-5
