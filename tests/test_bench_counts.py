"""The algorithmic counts bench.py reports (host only, no GPU)."""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from misc_amd import astro  # noqa: E402
from misc_amd.meshes import cantilever_block  # noqa: E402


def brute_flops(lt_range):
    R = np.asarray(lt_range, dtype=np.int64)
    rowlen = np.diff(R)
    s = np.arange(len(rowlen)) - rowlen + 1
    tot = 0
    for i in range(len(rowlen)):
        for j in range(s[i], i):
            tot += 2 * (j - max(s[i], s[j]))
    return tot


def test_factor_flops_closed_form_matches_brute_force():
    for dims, literal in [((6, 5, 3), False), ((4, 4, 2), True), ((7, 3, 2), False)]:
        p = astro.Prep(cantilever_block(*dims), astro.Options(literalMaxBW=literal))
        R = p.geti("elMatLTRange")
        assert bench.factor_flops(R) == brute_flops(R)


def test_factor_flops_with_decreasing_first_column():
    # rows 0..2: lengths 1, 2, 1 -> first columns 0, 0, 2; row 3 length 4 -> first column 0 (decreases)
    R = np.array([0, 1, 3, 4, 8])
    assert bench.factor_flops(R) == brute_flops(R) == 2
    rng = np.random.default_rng(4)
    s = np.array([max(0, i - int(rng.integers(0, 30))) for i in range(200)])
    R = np.concatenate([[0], np.cumsum(np.arange(200) - s + 1)])
    assert bench.factor_flops(R) == brute_flops(R)
