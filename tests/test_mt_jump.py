"""CPU: the jump-ahead of the probe stream (sedi_mt_state_after, host code of
libsedi_amd.so) against sequential generation.

The state after J outputs, loaded into a textbook mt19937 (numpy's MT19937
with its position at the end of the block, so that its next output twists
the loaded words first), must continue the sequential stream at output J
exactly.  numpy's legacy-seeded MT19937 is pinned here to the oracle's
std::mt19937 restatement and to the standard's known answer (10000th output
of seed 5489 = 4123659995) before it serves as the sequential reference.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import ROOT

LIB = os.path.join(ROOT, "sedi_amd", "libsedi_amd.so")


def seq_mt(seed):
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    return bg


@pytest.fixture(scope="module")
def jump():
    if not os.path.exists(LIB):
        pytest.fail("libsedi_amd.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    L.sedi_mt_state_after.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p]
    L.sedi_mt_state_after.restype = ctypes.c_int

    def state_after(seed, draws):
        s = np.zeros(624, dtype=np.uint32)
        assert L.sedi_mt_state_after(seed, draws, s.ctypes.data) == 0
        return s
    return state_after


def test_sequential_reference_is_std_mt19937():
    bg = seq_mt(5489)
    assert int(bg.random_raw(10000)[-1]) == 4123659995
    for seed in (1234, 7, 23):
        m = O.MT19937(seed)
        assert [m.next() for _ in range(700)] == seq_mt(seed).random_raw(700).tolist()


# offsets: the start, within and across the first block, and probe starts of
# the 32^3 and 256^3 grids (two outputs per cell)
OFFSETS = [0, 1, 623, 624, 625, 1247, 100003, 2 * 32768 * 3, 2 * 256 ** 3]


@pytest.mark.parametrize("seed", [1234, 7])
@pytest.mark.parametrize("draws", OFFSETS)
def test_state_after_continues_the_stream(jump, seed, draws):
    s = jump(seed, draws)
    g = np.random.MT19937(0)
    g.state = {"bit_generator": "MT19937", "state": {"key": s, "pos": 624}}
    ref = seq_mt(seed)
    left = draws
    while left:
        step = min(left, 1 << 22)
        ref.random_raw(step)
        left -= step
    assert np.array_equal(g.random_raw(1500), ref.random_raw(1500))


def test_state_before_first_output_is_the_seeded_state(jump):
    # zero draws: std::mt19937's seeded words (Knuth's initialisation)
    s = jump(1234, 0)
    x = [1234]
    for i in range(1, 624):
        x.append((1812433253 * (x[-1] ^ (x[-1] >> 30)) + i) & 0xFFFFFFFF)
    assert s.tolist() == x
