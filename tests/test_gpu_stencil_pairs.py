"""GPU: the standalone stencil's two marches give bitwise the oracle's y = A x.

The pair march (two adjacent cells per lane, 16-byte accesses; the default
where n1 is even and x, y are 16-byte aligned) and the one-cell march are
forced in child processes (SEDI_TUNING_FILE, read once per process) on grids
with even n1 that exercise partial tiles along x (n1 = 2, 130, 258), y and z,
with each schedule the autotuner can pick for the pair march and with one and
two workgroups per CU.  In this process: device arrays that are only 8-byte
aligned, and an odd n1, take the one-cell march, still bitwise.
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from conftest import ROOT

sedi = pytest.importorskip("sedi_amd")

pytestmark = pytest.mark.gpu

SHAPES = [(2, 67, 70), (130, 9, 3), (64, 4, 3), (128, 17, 9), (258, 5, 40), (34, 33, 40), (256, 16, 33)]

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, %(root)r)
sys.path.insert(0, %(tests)r)
import sedi_amd as sedi
import oracle_lib as O
out = []
for k, shape in enumerate(%(shapes)r):
    n = int(np.prod(shape))
    x = np.random.default_rng(10 + k).standard_normal(n)
    for coef in ((6.0, 1.0, 1.0), (2.5, 0.25, 0.75)):
        y = sedi.Stencil(*coef).apply(tuple(shape), x)
        out.append(bool(np.array_equal(y, O.stencil_apply(tuple(shape), x, *coef))))
print(json.dumps(out))
"""

# (pair march, chunk, schedule, planes in flight, workgroups per CU)
PLANS = [(1, 16, 3, 2, 1), (1, 8, 0, 4, 1), (1, 32, 4, 2, 2), (1, 16, 2, 2, 1), (1, 8, 1, 4, 0),
         (0, 16, 3, 2, 0), (0, 32, 2, 8, 0)]


@pytest.mark.parametrize("x2,zc,sched,pf,wpc", PLANS)
def test_forced_stencil_march_bitwise(tmp_path, x2, zc, sched, pf, wpc):
    table = tmp_path / "tuning.txt"
    table.write_text("x2:stencil_apply_kernel 0 0 0 %d\nzc:stencil_apply_kernel 0 0 0 %d\n"
                     "sched:stencil_apply_kernel 0 0 0 %d\npf:stencil_apply_kernel 0 0 0 %d\n"
                     "wpc:stencil_apply_kernel 0 0 0 %d\n" % (x2, zc, sched, pf, wpc))
    env = dict(os.environ, SEDI_TUNING_FILE=str(table))
    code = CHILD % dict(root=ROOT, tests=os.path.join(ROOT, "tests"), shapes=SHAPES)
    print("stencil plan x2=%d zc=%d sched=%d pf=%d wpc=%d" % (x2, zc, sched, pf, wpc), flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    same = json.loads(r.stdout.strip().splitlines()[-1])
    assert all(same), same


def _apply_at(shape, x, offset_doubles):
    """sedi_stencil_apply on device arrays whose data starts `offset_doubles`
    doubles into the allocation (8-byte aligned only when odd)."""
    n = int(np.prod(shape))
    L = sedi.lib()
    g, _ = sedi._grid(shape)
    st = sedi.Stencil()._c()
    xd = sedi.DeviceArray(np.concatenate([np.zeros(offset_doubles), x]))
    yd = sedi.DeviceArray(n=n + offset_doubles)
    try:
        xp = ctypes.c_void_p(xd.ptr.value + 8 * offset_doubles)
        yp = ctypes.c_void_p(yd.ptr.value + 8 * offset_doubles)
        assert L.sedi_stencil_apply(ctypes.byref(g), ctypes.byref(st), xp, yp) == 0, L.sedi_last_error()
        return yd.get()[offset_doubles:]
    finally:
        xd.free()
        yd.free()


@pytest.mark.parametrize("offset", [0, 1, 2])
@pytest.mark.parametrize("shape", [(130, 9, 3), (256, 16, 33), (65, 5, 9)])
def test_stencil_device_arrays_any_alignment(shape, offset):
    x = np.random.default_rng(offset).standard_normal(int(np.prod(shape)))
    assert np.array_equal(_apply_at(shape, x, offset), O.stencil_apply(shape, x))
