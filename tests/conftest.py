import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsedi_amd.so)")
    config.addinivalue_line("markers", "slow: full-size (256^3 and above) GPU case")


@pytest.fixture(scope="session")
def rng_golden():
    with open(os.path.join(GOLDEN, "rng_golden.json")) as f:
        return json.load(f)


def load_npz(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated" % name)
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def ref32():
    return load_npz("ref_32cube.npz")


@pytest.fixture(scope="session")
def ref_odd():
    return load_npz("ref_odd.npz")


def normwise(a, b):
    """max|a-b| / max|b|: D estimates cross zero at some cells, so the per-cell
    relative error is not meaningful; this is the metric the parity bars use."""
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def sign_words(v):
    """Packs a +-1 vector into 32-bit words (bit c%32 of word c//32 = cell c is +1)."""
    bits = (np.asarray(v) > 0).astype(np.uint8)
    pad = (-len(bits)) % 32
    bits = np.concatenate([bits, np.zeros(pad, np.uint8)])
    b = np.packbits(bits.reshape(-1, 8), axis=1, bitorder="little").reshape(-1, 4)
    return b.view("<u4").reshape(-1).astype(np.uint64)


def wsum(words):
    i = np.arange(len(words), dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(words * (2 * i + 1), dtype=np.uint64))


@pytest.fixture(scope="session")
def ref_spread():
    return load_npz("ref_spread.npz")


def order_spread(spread, tag, kind):
    """How far the reference's own results move when only its summation order
    changes (tests/golden/ref_spread.npz: other -p decompositions and 4- and
    8-lane SIMD builds of the same solver code, against its -p 1x1x1 2-lane
    run): the largest |iteration count - the 1x1x1 run's| over every variant
    and probe, and the largest normwise distance of D.  CG and pipelined CG
    do not move (0 iterations, D within ~1e-12); BiCGStab does."""
    its = spread[f"{tag}_s{kind}_iters"]
    return int(np.max(np.abs(its[1:] - its[0]))), float(np.max(spread[f"{tag}_s{kind}_Ddist"]))


def iteration_bar(spread, tag, kind):
    """Allowed |GPU count - reference count| per probe: the reference's own
    spread under summation-order changes, and at least 1 (the GPU sums in yet
    another order)."""
    return max(1, order_spread(spread, tag, kind)[0])


def d_bar(spread, tag, kind, counts_equal):
    """Allowed normwise distance of D from the reference's: 1e-10 where every
    count agrees and the solver is order-stable (CG, pipelined CG); else twice
    the reference's own largest distance under summation-order changes."""
    dit, dD = order_spread(spread, tag, kind)
    if counts_equal and dit == 0:
        return 1e-10
    return max(1e-10, 2 * dD) if dit > 0 else 1e-8
