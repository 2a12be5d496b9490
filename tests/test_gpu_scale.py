"""GPU parity at the sizes of the benchmark configurations.

- The 7-piece synthetic slice KQRBvKQR 170 (776,520,240 positions per
  side, 2^21 bitmap words): its codes against the oracle's full-fixpoint
  digest (tests/golden/synth_*.json).  Only slices of this size reach the
  large verification units, the two-pass verification and the wide
  predecessor kernel with multi-block LDS scatter.
- The 8-piece slice KQRBvKQRN 170 at real shape (44,261,653,680 positions
  per side, indices up to 2^35): plies 1 and 2 stopped, read back and
  checked position by position at seeded samples against the oracle's
  forward evaluation (tbo_synth_check).
- KQRvKR in the plain one-slice-per-launch form against its digest.
- 6-piece synthetic slices to the full fixpoint, code for code, in every
  scatter mode and under extreme launch geometries.
"""
import json
import os
import zlib

import numpy as np
import pytest

import tb8gen_amd as tbg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

O.set_threads(16)
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_synth = {}


def oracle_synth(material, s, seed):
    """The oracle's full-fixpoint synthetic slice: (codes per side, lines)."""
    key = (material, s, seed)
    if key not in _synth:
        log = f"/tmp/oracle_scale_{material}_{s}_{seed}.log"
        o = O.Table(material, synth_slice=s, seed=seed, log_path=log)
        with open(log) as f:
            lines = [l.strip() for l in f if l.strip()]
        _synth[key] = ([o.codes(0)[0], o.codes(1)[0]], lines)
    return _synth[key]


def assert_codes(material, t, ref_codes):
    for stm in range(2):
        g = t.codes(stm)[0]
        bad = np.flatnonzero(g != ref_codes[stm])
        assert len(bad) == 0, (f"{material} stm {stm}: {len(bad)} codes differ, first idx "
                               f"{bad[0]} gpu {g[bad[0]]} oracle {ref_codes[stm][bad[0]]}")


def test_seven_piece_slice_against_oracle_digest():
    """KQRBvKQR slice 170 to its fixpoint: count lines and the CRC32 of
    every run of 2^24 codes of each side equal the oracle's."""
    path = os.path.join(GOLDEN, "synth_KQRBvKQR_s170_20261203.json")
    with open(path) as f:
        ref = json.load(f)
    print(f"case synth digest {ref['material']} s{ref['slice']}", flush=True)
    t = tbg.Table(ref["material"], synth_slice=ref["slice"], seed=ref["seed"], keep_codes=True)
    assert t.size == ref["slice_size"]
    t.resolve([])
    lines = t.generate(tbg.FORM_BATCHED)
    assert lines == ref["lines"]
    chunk = ref["chunk"]
    for stm in range(2):
        c = t.codes(stm)[0]
        crc = [zlib.crc32(c[i:i + chunk].tobytes()) for i in range(0, c.size, chunk)]
        bad = [i for i in range(len(crc)) if crc[i] != ref["crc32"][str(stm)][i]]
        assert not bad, f"stm {stm}: chunks {bad[:10]} of {len(crc)} differ"
    t.close()


def test_eight_piece_slice_sampled_plies():
    """The 8-piece slice at real shape, stopped after plies 1 and 2: every
    sampled position's membership in wins, W_1, L_1, W_2 and L_2 (both
    sides; 10^5 samples per side, most indices above 2^32) equals the
    oracle's forward evaluation from the position's legal moves."""
    material, s, seed = "KQRBvKQRN", 170, 20261203
    print(f"case 8-piece sampled plies {material} s{s}", flush=True)
    t = tbg.Table(material, synth_slice=s, seed=seed)
    assert t.size == 44261653680
    t.resolve([])
    t.init_illegal_mate()
    lines1 = t.continue_(tbg.FORM_BATCHED, stop_at=2)
    print(f"  ply 1: {lines1}", flush=True)
    planes = {}
    for name, plane in (("wins1", tbg.PLANE_WINS), ("W1", tbg.PLANE_W), ("L1", tbg.PLANE_L)):
        planes[name] = [t.bitmap(plane, stm, s) for stm in range(2)]
    lines2 = t.continue_(tbg.FORM_BATCHED, stop_at=3)
    print(f"  ply 2: {lines2}", flush=True)
    for name, plane in (("W2", tbg.PLANE_W), ("L2", tbg.PLANE_L)):
        planes[name] = [t.bitmap(plane, stm, s) for stm in range(2)]
    t.close()
    # the planes agree with the count lines the engine printed
    counts = dict(l.split(" = ") for l in lines1 + lines2)
    for name, key in (("W1", "w1"), ("L1", "l1"), ("W2", "w2"), ("L2", "l2")):
        for stm in range(2):
            got = int(np.bitwise_count(planes[name][stm]).sum(dtype=np.int64))
            assert got == int(counts.get(f"{key}_{'wb'[stm]}", 0)), (name, stm)
    order = [planes[n][c] for n in ("wins1", "W1", "L1", "W2", "L2") for c in range(2)]
    bad, checked, first = O.synth_check(material, s, seed, order, 100000, sample_seed=20261204)
    print(f"  checked {checked} positions, {bad} mismatches", flush=True)
    assert checked == 200000
    assert bad == 0, f"first mismatch (stm, plane, index, engine bit) = {first}"


def test_plain_form_kqrvkr_against_digest():
    """KQRvKR in the plain form (one launch per slice and phase, manager
    order) equals the oracle's digest."""
    with open(os.path.join(GOLDEN, "oracle_KQRvKR.json")) as f:
        ref = json.load(f)
    print("case plain KQRvKR", flush=True)
    reg = tbg.NativeRegistry(tbg.FORM_GRAPH)
    subs = [reg.get(m) for m in ("KQvKR", "KRvKR", "KQRvK")]
    t = tbg.Table("KQRvKR")
    t.resolve(subs)
    assert t.generate(tbg.FORM_PLAIN) == ref["lines"]
    for stm in range(2):
        c = t.codes(stm)
        crc = [zlib.crc32(c[s].tobytes()) for s in range(c.shape[0])]
        assert crc == ref["crc32"][str(stm)], stm
    t.close()
    reg.close()


@pytest.mark.parametrize("mode", [tbg.SCATTER_BLOCK, tbg.SCATTER_GLOBAL,
                                  tbg.SCATTER_BLOCK | tbg.SCATTER_COMBINE,
                                  tbg.SCATTER_GLOBAL | tbg.SCATTER_COMBINE])
def test_six_piece_slice_full_fixpoint(mode):
    """A 6-piece slice spans many pred blocks (the multi-block LDS scatter,
    the wide pred kernel): every code equals the oracle's at the fixpoint,
    in every scatter mode."""
    material, s, seed = "KQRvKRN", 170, 5
    ref, olines = oracle_synth(material, s, seed)
    print(f"case synth 6-piece {material} scatter {mode}", flush=True)
    t = tbg.Table(material, synth_slice=s, seed=seed, keep_codes=True)
    t.set_scatter(mode)
    t.resolve([])
    assert t.generate(tbg.FORM_BATCHED) == olines
    assert_codes(material, t, ref)
    t.close()


@pytest.mark.parametrize("gi", range(3))
def test_six_piece_slice_geometries(gi):
    """Extreme unit sizes and grid caps on the 6-piece slice: the codes
    equal the oracle's (results never depend on the launch geometry)."""
    from tests.test_gpu_parity import _tune_candidates
    material, s, seed = "KQRvKRN", 170, 5
    ref, olines = oracle_synth(material, s, seed)
    geo = _tune_candidates()[gi]
    print(f"case synth 6-piece geometry {geo}", flush=True)
    t = tbg.Table(material, synth_slice=s, seed=seed, keep_codes=True)
    t.set_geometry(tbg.Geometry(**geo))
    t.resolve([])
    assert t.generate(tbg.FORM_BATCHED) == olines
    assert_codes(material, t, ref)
    t.close()


def test_six_piece_checkpoint_restore_against_oracle(tmp_path):
    """The 6-piece slice stopped at ply 4, checkpointed, restored into a
    fresh context and finished: codes and lines equal the oracle's."""
    material, s, seed = "KQRvKRN", 170, 5
    ref, olines = oracle_synth(material, s, seed)
    print(f"case checkpoint synth {material}", flush=True)
    t = tbg.Table(material, synth_slice=s, seed=seed, keep_codes=True)
    t.resolve([])
    t.init_illegal_mate()
    t.continue_(tbg.FORM_BATCHED, stop_at=4)
    t.checkpoint(tmp_path)
    t.close()
    t2 = tbg.Table(material, synth_slice=s, seed=seed, keep_codes=True)
    t2.restore(tmp_path)
    assert t2.continue_(tbg.FORM_BATCHED) == olines
    assert_codes(material, t2, ref)
    t2.close()
