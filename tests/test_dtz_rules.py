"""The oracle against the DTZ50 rules, checked position by position by an
independent checker (tests/golden/dtzcheck.c): the brute-force solver's move
generator, the codes only as data.  This pins the 50-move handling -- the
n = 1 and n = 101 capture cases, the cwin verification above ply 100, the
forced ply 101 and the cursed / blessed values -- independently of the
oracle's recurrences (src/tb8gen.c:964-970, 1010, 1046-1048, 1426-1433)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.golden import dtzcheck as D

O.set_threads(8)

FOUR = ["KQvK", "KRvK", "KQvKR", "KRvKN", "KBNvK", "KRRvK", "KBBvK", "KQvKQ"]
_tabs = {}


def _sub_materials(material):
    """Materials one capture away (colour orientation as the oracle makes
    them), excluding a bare pair of kings."""
    w, b = material.split("v")
    out = []
    for side, other, white in ((w, b, True), (b, w, False)):
        for i, ch in enumerate(side):
            if ch == "K":
                continue
            rest = side[:i] + side[i + 1:]
            m = f"{rest}v{other}" if white else f"{other}v{rest}"
            if len(m) - 1 > 2 and m not in out:
                out.append(m)
    return out


def _table(m):
    if m not in _tabs:
        _tabs[m] = O.Table(m)
    return _tabs[m]


def _codes(m):
    t = _table(m)
    return t.codes(0).ravel(), t.codes(1).ravel()


def _subs(m):
    out = []
    for sm in _sub_materials(m):
        # the oracle generates a sub-material in its own preferred colour
        # orientation; the checker accepts either
        w, b = sm.split("v")
        t = O.Table.find(sm) or O.Table.find(f"{b}v{w}")
        out.append((t.material, t.codes(0).ravel(), t.codes(1).ravel()))
    return out


@pytest.mark.parametrize("m", ["KQvKR", "KRRvK", "KBNvK"])
def test_checker_layout_matches_oracle(m):
    """The checker's own restatement of the slice layout unranks like the
    oracle's (which is pinned by its bijection tests)."""
    t = _table(m)
    assert D.slice_size(m) == t.size
    rng = np.random.default_rng(7)
    for _ in range(2000):
        s, idx = int(rng.integers(0, 462)), int(rng.integers(0, t.size))
        assert D.unrank(m, s, idx) == t.idx_to_sq(s, idx), (s, idx)


@pytest.mark.parametrize("m", FOUR)
def test_three_and_four_piece_tables_follow_the_rules(m):
    cw, cb = _codes(m)
    bad, st, first = D.check(m, cw, cb, _subs(m))
    assert bad == 0, first
    assert st["checked"] == 2 * cw.size


def test_kbnvkn_cursed_and_blessed_follow_the_rules():
    """KBNvKN reaches ply 153: white's cursed wins (W_n, n > 100) and
    black's blessed losses (L_n, n > 100) follow the rules at every 7th
    position."""
    m = "KBNvKN"
    cw, cb = _codes(m)
    bad, st, first = D.check(m, cw, cb, _subs(m), stride=7)
    assert bad == 0, first
    assert st["cursed_win"] > 10000 and st["blessed_loss"] > 5000, st


def test_checker_catches_planted_mistakes():
    """Mistakes planted in the oracle's codes (on the CPU) are found at the
    position planted: a win one ply too slow, a draw turned into a loss, a
    loss one ply too fast, and a cursed win turned into a plain win."""
    m = "KQvKR"
    cw, cb = (c.copy() for c in _codes(m))
    subs = _subs(m)

    def at(codes_w, codes_b, g):
        return D.check(m, codes_w, codes_b, subs, start=g, count=1)[0]

    for code_from, change in (
        (lambda c: (c & 3) == 1, lambda c: c + 4),             # W_n -> W_{n+1}
        (lambda c: c == 0, lambda c: (3 << 2) | 2),            # draw -> L_3
        (lambda c: ((c & 3) == 2) & ((c >> 2) > 1), lambda c: c - 4),  # L_n -> L_{n-1}
    ):
        hits = np.flatnonzero(code_from(cw))
        g = int(hits[len(hits) // 2])
        assert at(cw, cb, g) == 0
        planted = cw.copy()
        planted[g] = change(int(cw[g]))
        assert at(planted, cb, g) > 0
    m2 = "KBNvKN"
    w2, b2 = _codes(m2)
    subs2 = _subs(m2)
    g = int(np.flatnonzero(((w2 & 3) == 1) & ((w2 >> 2) > 100))[100])
    planted = w2.copy()
    planted[g] = (100 << 2) | 1  # a cursed win made a win within the 50-move rule
    assert D.check(m2, w2, b2, subs2, start=g, count=1)[0] == 0
    assert D.check(m2, planted, b2, subs2, start=g, count=1)[0] > 0
