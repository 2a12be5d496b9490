"""The engine's finished tables against the DTZ50 rules, position by
position (tests/golden/dtzcheck.c: the brute-force solver's move generator,
the engine's codes only as data; the check runs on the host cores).

- KBNvKN (cursed wins and blessed losses up to ply 153) and KQRvKR: every
  position of both sides.
- KBNNvKN and KBNvKNN, whose captures reach KBNvKN's cursed wins and
  blessed losses: the only tables here where the ply-101 uncapture
  (src/tb8gen.c:964-970, 1046-1048) has a non-empty source.  Every position
  the generator put into L_101 / W_101, a seeded sample of the rest, and a
  sample of the 5-piece tables they capture into that the digests do not
  pin.
"""
import numpy as np
import pytest

import tb8gen_amd as tbg
from tests.golden import dtzcheck as D

pytestmark = pytest.mark.gpu
THREADS = 16


@pytest.fixture(scope="module")
def reg():
    print("registry for the DTZ50 checks", flush=True)
    r = tbg.NativeRegistry(tbg.FORM_GRAPH)
    yield r
    r.close()


def _codes(t):
    return t.codes(0).ravel(), t.codes(1).ravel()


def _subs(reg, material):
    tabs = dict(reg.entries())
    out = []
    w, b = material.split("v")
    for side, other, white in ((w, b, True), (b, w, False)):
        for i, ch in enumerate(side):
            if ch == "K":
                continue
            rest = side[:i] + side[i + 1:]
            m = f"{rest}v{other}" if white else f"{other}v{rest}"
            if len(m) - 1 <= 2:
                continue
            mw, mb = m.split("v")
            for cand in (tbg.canonical(m), m, f"{mb}v{mw}"):
                if cand in tabs and all(cand != o[0] for o in out):
                    out.append((cand, *_codes(tabs[cand])))
                    break
    return out


@pytest.mark.parametrize("material", ["KBNvKN", "KQRvKR"])
def test_five_piece_tables_follow_dtz50_rules(reg, material):
    print(f"case dtz50 {material}", flush=True)
    t = reg.get(material)
    cw, cb = _codes(t)
    bad, st, first = D.check(material, cw, cb, _subs(reg, material), threads=THREADS)
    print(f"  {st}", flush=True)
    assert bad == 0, first
    assert st["checked"] == 2 * cw.size
    if material == "KBNvKN":
        assert st["cursed_win"] > 0 and st["blessed_loss"] > 0


@pytest.mark.parametrize("material,kind,stm", [("KBNNvKN", 2, 1), ("KBNvKNN", 1, 0)])
def test_six_piece_uncapture_at_ply_101(reg, material, kind, stm):
    """KBNNvKN: black captures a knight into KBNvKN positions that white
    wins only after more than 100 plies (sub/cwin without sub/win), the
    source of L_101 for black.  KBNvKNN: white captures a knight into
    positions black loses only after more than 100 plies (sub/bloss), the
    source of W_101 for white."""
    print(f"case dtz50 {material}", flush=True)
    t = reg.get(material)
    line = f"{'lw'[kind == 1]}101_{'wb'[stm]}"
    n101 = [int(l.split(" = ")[1]) for l in t.lines if l.startswith(line + " ")]
    print(f"  {line} = {n101}", flush=True)
    assert n101 and n101[0] > 0
    subs = _subs(reg, material)
    for m, sw, sb in subs:  # the 5-piece tables no digest pins, sampled
        if m in ("KBNvKN", "KQRvKR"):
            continue
        bad, st, first = D.check(m, sw, sb, _subs(reg, m), stride=7, threads=THREADS)
        print(f"  sub {m}: {st}", flush=True)
        assert bad == 0, (m, first)
    cw, cb = _codes(t)
    code = (101 << 2) | (2 if kind == 2 else 1)
    at101 = np.flatnonzero((cb if stm else cw) == code)
    assert at101.size == n101[0]
    bad, st, first = D.check(material, cw, cb, subs, positions=at101, threads=THREADS)
    print(f"  ply-101 positions: {st}", flush=True)
    assert bad == 0, first
    assert st["capture_into_cursed_or_blessed"] > 0
    bad, st, first = D.check(material, cw, cb, subs, start=12345, stride=509, threads=THREADS)
    print(f"  sample: {st}", flush=True)
    assert bad == 0, first
