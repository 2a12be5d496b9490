"""Closed forms, invariants and known answers that pin the oracle."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

O.set_threads(min(8, os.cpu_count() or 1))

# The reference's KKIdx and InvTriangle tables (src/probe.c:441-522,
# 327-329) as data, every entry: the slice numbering, including the 21
# slices with both kings on a1-h8 (441..461).
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "kkidx_probe_c.json")) as _f:
    KKIDX = json.load(_f)


def test_kkidx_whole_table():
    for tri in range(10):
        for bk in range(64):
            assert O.kkidx(tri, bk) == KKIDX["kkidx"][tri][bk], (tri, bk)


def test_kkmap_agrees_with_kkidx():
    """KKMap (src/tb8gen.c:69-89) of each table entry's king pair is that
    entry; the pairs the table leaves out are equal, touching or, with the
    white king on a1-d4, have the black king above the diagonal."""
    for tri, wk in enumerate(KKIDX["inv_triangle"]):
        for bk in range(64):
            want = KKIDX["kkidx"][tri][bk]
            if want >= 0:
                assert O.kkmap(wk, bk) == want, (wk, bk)
            else:
                touching = max(abs((wk & 7) - (bk & 7)), abs((wk >> 3) - (bk >> 3))) <= 1
                above = (wk & 7) == (wk >> 3) and (bk & 7) < (bk >> 3)
                assert touching or above, (wk, bk)


def test_slices_partition_king_pairs():
    """462 classes of legal king pairs; 8-fold symmetric; diagonal slices."""
    seen = np.zeros(462, dtype=np.int64)
    for wk in range(64):
        for bk in range(64):
            s = O.kkmap(wk, bk)
            touching = max(abs((wk & 7) - (bk & 7)), abs((wk >> 3) - (bk >> 3))) <= 1
            if touching:
                assert s == -1
                continue
            seen[s] += 1
            f = lambda x: (x >> 3) * 8 + (7 - (x & 7))          # file mirror
            r = lambda x: ((7 - (x >> 3)) * 8) + (x & 7)        # rank mirror
            d = lambda x: ((x & 7) << 3) | (x >> 3)              # a1-h8 flip
            for g in (f, r, d, lambda x: f(r(x)), lambda x: d(f(x))):
                assert O.kkmap(g(wk), g(bk)) == s
    assert (seen > 0).all()
    assert seen.sum() == 64 * 63 - 420  # 3612 legal king pairs
    for s in range(462):
        wk, bk = O.kksquare(s)
        both_diag = (wk & 7) == (wk >> 3) and (bk & 7) == (bk >> 3)
        assert both_diag == (s >= 441)
        assert O.kkmap(wk, bk) == s


def test_index_is_a_bijection():
    t = O.Table("KRvKN")
    assert t.size == 62 * 61
    for s in (0, 170, 441):
        seen = set()
        for idx in range(t.size):
            sq = t.idx_to_sq(s, idx)
            assert len(set(sq)) == len(sq)
            assert t.sq_to_idx(sq) == idx
            seen.add(tuple(sq))
        assert len(seen) == t.size


def test_index_sizes_are_binomial_products():
    for material, size in [("KRRvK", math.comb(62, 2)), ("KQRvKR", 62 * 61 * 60),
                           ("KBNvKQ", 62 * 61 * 60)]:
        t = O.Table(material)
        assert t.size == size


def _max_n(table, stm, tag):
    c = table.codes(stm).ravel()
    sel = c[(c & 3) == tag]
    return int(sel.max() >> 2) if sel.size else 0


def test_known_answers_three_piece():
    """KQvK mate in 10 and KRvK mate in 16 (DTZ = DTM in plies here);
    KBvK and KNvK hold no win at all."""
    q, r = O.Table("KQvK"), O.Table("KRvK")
    assert _max_n(q, 0, 1) == 19 and _max_n(q, 1, 2) == 20
    assert _max_n(r, 0, 1) == 31 and _max_n(r, 1, 2) == 32
    for m in ("KBvK", "KNvK"):
        t = O.Table(m)
        for stm in range(2):
            c = t.codes(stm)
            assert not ((c & 3) == 1).any() and not ((c & 3) == 2).any()


def test_known_answer_kbnk():
    """KBNvK: the longest win is mate in 33 (65 plies)."""
    t = O.Table("KBNvK")
    assert _max_n(t, 0, 1) == 65


def test_count_lines_format_and_totals(tmp_path):
    log = tmp_path / "krvk.log"
    t = O.Table("KRvK", log_path=str(log)) if O.lib().tbo_find(b"KRvK") < 0 else None
    if t is None:
        O.free_all()
        t = O.Table("KRvK", log_path=str(log))
    lines = log.read_text().split("\n")
    assert lines[0].startswith("broken_w = ") and lines[3].startswith("l0_b = ")
    # every W_n / L_n line equals the number of codes with that value
    for stm, c in ((0, "w"), (1, "b")):
        codes = t.codes(stm).ravel()
        for line in lines[4:]:
            if not line:
                continue
            key, val = line.split(" = ")
            kind, rest = key[0], key[1:]
            n, side = rest.split("_")
            if side != c:
                continue
            tag = 2 if kind == "l" else 1
            assert int(val) == int(((codes >> 2) == int(n))[(codes & 3) == tag].sum())


def test_reference_king_capture_switch_is_inert_on_distinct_pieces():
    """The king-capture index variant only matters for duplicated sets."""
    O.free_all()
    base = [O.Table("KQvKR").codes(s).copy() for s in range(2)]
    O.free_all()
    O.set_ref_king_capture_index(True)
    try:
        alt = [O.Table("KQvKR").codes(s).copy() for s in range(2)]
    finally:
        O.set_ref_king_capture_index(False)
        O.free_all()
    for s in range(2):
        assert np.array_equal(base[s], alt[s])


def test_synthetic_slice_deterministic_and_seeded():
    a = O.Table("KQvKR", synth_slice=170, seed=7)
    b = O.Table("KQvKR", synth_slice=170, seed=7)
    c = O.Table("KQvKR", synth_slice=170, seed=8)
    for stm in range(2):
        assert np.array_equal(a.codes(stm), b.codes(stm))
    assert not np.array_equal(a.codes(0), c.codes(0))
    # the exits are Bernoulli(0.03) wins and losses
    h = np.array([O.synth_hash(7, 0, 0, 0, i) for i in range(20000)], dtype=np.uint64)
    win = (h < np.uint64(553402322211286548)).mean()
    assert 0.02 < win < 0.04


def test_synth_forward_check_pins_itself():
    """The sampled forward check used on the 8-piece slice (moves forward
    from each position) agrees with the oracle's retrograde generation
    (unmoves backward) on a 6-piece synthetic slice, and catches planted
    mistakes in the planes it is given: an off-by-one shift of one plane,
    and a plane emptied."""
    m, s, seed = "KQRvKRN", 170, 5
    o = O.Table(m, synth_slice=s, seed=seed, max_plies=3)
    words = ((o.size + 511) // 512) * 8
    pl = O.planes_from_codes(o.codes(0), o.codes(1), words)
    assert all(int(np.bitwise_count(p).sum()) > 0 for p in pl)
    bad, checked, first = O.synth_check(m, s, seed, pl, 20000, sample_seed=3)
    assert (bad, checked) == (0, 40000), first
    shifted = [p.copy() for p in pl]
    shifted[5] = (shifted[5] << np.uint64(1)) | np.concatenate(
        [[np.uint64(0)], shifted[5][:-1] >> np.uint64(63)])  # L_1(black), index + 1
    bad, _, first = O.synth_check(m, s, seed, shifted, 20000, sample_seed=3)
    assert bad > 0 and first is not None
    emptied = [p.copy() for p in pl]
    emptied[6][:] = 0  # W_2(white)
    bad, _, first = O.synth_check(m, s, seed, emptied, 20000, sample_seed=3)
    assert bad > 0 and first[:2] == (0, 3)
