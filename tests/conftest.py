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
