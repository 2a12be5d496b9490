"""The CPU oracle against the independent brute-force solver (naive.c).

The fixtures hold full-board DTZ codes (no king-slice symmetry) computed by
forward evaluation; the oracle's tables are looked up for every placement of
the pieces, through its own normalisation and index, for both sides to move.
"""
import glob
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "naive_*.npz")))
TYPE = {"K": 6, "Q": 5, "R": 4, "B": 3, "N": 2}

O.set_threads(min(8, os.cpu_count() or 1))


def full_board(material, table):
    """Oracle codes of every full-board placement, in naive.c's order
    (piece i of the material string on square (p >> 6i) & 63)."""
    toks, side = [], 0
    for ch in material:
        if ch == "v":
            side = 1
            continue
        toks.append(TYPE[ch] | (8 * side))
    k = len(toks)
    p = np.arange(64 ** k, dtype=np.int64)
    sqs = np.stack([(p >> (6 * i)) & 63 for i in range(k)], axis=1).astype(np.uint8)
    used, perm = [False] * k, []
    for pt in table.pt:
        for i in range(k):
            if not used[i] and toks[i] == pt:
                used[i] = True
                perm.append(i)
                break
    osq = np.ascontiguousarray(sqs[:, perm])
    out = []
    for stm in range(2):
        c = table.lookup_bulk(stm, osq).astype(np.int16)
        c[c == -1] = 0x7FFF  # two pieces on one square
        out.append(c)
    return out


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f)[6:-4] for f in FIXTURES])
def test_oracle_matches_brute_force(path):
    fx = np.load(path, allow_pickle=False)
    material = str(fx["material"])
    t = O.Table(material)
    arrs = full_board(material, t)
    for stm in range(2):
        a = arrs[stm]
        vals, cnts = np.unique(a, return_counts=True)
        assert np.array_equal(vals, fx[f"hist_codes_{stm}"]), (material, stm)
        assert np.array_equal(cnts, fx[f"hist_counts_{stm}"]), (material, stm)
        if f"codes_{stm}" in fx:
            assert np.array_equal(a, fx[f"codes_{stm}"]), (material, stm)
        else:
            idx = fx[f"sample_idx_{stm}"]
            assert np.array_equal(a[idx], fx[f"sample_code_{stm}"]), (material, stm)
        assert hashlib.sha256(a.tobytes()).hexdigest() == str(fx[f"sha256_{stm}"]), (material, stm)


def test_fixture_set_present():
    names = {os.path.basename(f) for f in FIXTURES}
    for m in ("KRvK", "KQvK", "KBvK", "KNvK", "KQvKR", "KBNvK", "KRRvK"):
        assert f"naive_{m}.npz" in names
