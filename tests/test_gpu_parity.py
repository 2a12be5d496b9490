"""GPU parity: the HIP engine against the CPU oracle, bit-exact.

Every code of every (side, slice, index) and every count line of the
reference's format must agree, for materials whose captures exercise the
whole chain (3-piece tables probe KvK, 4-piece probe 3-piece, 5-piece probe
4-piece), for the batched and the plain (one slice per launch) forms.
"""
import glob
import json
import os
import zlib

import numpy as np
import pytest

import tb8gen_amd as tbg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

O.set_threads(16)
_oracles = {}


def oracle_table(material):
    if material not in _oracles:
        _oracles[material] = (O.Table(material, log_path=f"/tmp/oracle_{material}.log"))
    return _oracles[material]


def oracle_lines(material):
    oracle_table(material)
    with open(f"/tmp/oracle_{material}.log") as f:
        return [l.strip() for l in f if l.strip()]


@pytest.fixture(scope="module")
def registry():
    print("registry: batched form", flush=True)
    r = tbg.NativeRegistry(tbg.FORM_BATCHED)
    yield r
    r.close()


def assert_same_table(material, gpu):
    ora = oracle_table(material)
    for stm in range(2):
        g = gpu.codes(stm)
        o = ora.codes(stm)
        assert g.shape == o.shape
        bad = np.argwhere(g != o)
        assert len(bad) == 0, (
            f"{material} stm {stm}: {len(bad)} codes differ, first slice/idx {bad[0]} "
            f"gpu {g[tuple(bad[0])]} oracle {o[tuple(bad[0])]}")


@pytest.mark.parametrize("material", ["KRvK", "KQvK", "KBvK", "KNvK"])
def test_three_piece(registry, material):
    print(f"case {material}", flush=True)
    t = registry.get(material)
    assert t.lines == oracle_lines(material)
    assert_same_table(material, t)


@pytest.mark.parametrize("material", ["KQvKR", "KRvKN", "KBNvK", "KRRvK", "KBBvK", "KQvKQ"])
def test_four_piece(registry, material):
    print(f"case {material}", flush=True)
    t = registry.get(material)
    assert t.lines == oracle_lines(material)
    assert_same_table(material, t)


def test_sub_bitmaps(registry):
    """Capture-exit bitmaps (sub/loss, bloss, cwin, win) of KQvKR."""
    material = "KQvKR"
    t = registry.get(material)
    ora = oracle_table(material)
    for colour in range(2):
        for kind in range(4):
            for s in (0, 57, 170, 440, 441, 461):
                g = t.sub_bitmap(kind, colour, s)
                o = ora.sub_bitmap(kind, colour, s)
                assert np.array_equal(g, o), (colour, kind, s)


def test_plain_form_equals_batched(registry):
    """One slice per launch in manager order gives the same table."""
    material = "KRvKN"
    registry.get(material)
    print("case plain KRvKN", flush=True)
    t = tbg.Table(material)
    t.resolve(list(registry.tables.values()))
    lines = t.generate(tbg.FORM_PLAIN)
    assert lines == oracle_lines(material)
    assert_same_table(material, t)
    t.close()


_DIGESTS = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                         "oracle_*.json")))


@pytest.mark.parametrize("path", _DIGESTS, ids=[os.path.basename(p)[7:-5] for p in _DIGESTS])
def test_five_piece_against_oracle_digests(registry, path):
    """5-piece tables (KQRvKR 462 x 226,920 per side; KQQQvK with a triple
    set; KRRvKB, KBNvKN): count lines and per-slice CRC32 of the codes equal
    the oracle's (tests/golden/oracle_*.json)."""
    with open(path) as f:
        ref = json.load(f)
    material = ref["material"]
    print(f"case {material}", flush=True)
    t = registry.get(material)
    assert t.size == ref["slice_size"]
    assert t.lines == ref["lines"]
    for stm in range(2):
        c = t.codes(stm)
        crc = [zlib.crc32(c[s].tobytes()) for s in range(c.shape[0])]
        bad = [s for s in range(len(crc)) if crc[s] != ref["crc32"][str(stm)][s]]
        assert not bad, f"{material} stm {stm}: slices {bad[:10]} differ"


@pytest.mark.parametrize("material,slice_,seed", [("KQRvKR", 170, 7), ("KRBvKN", 57, 11),
                                                   ("KRRvKB", 300, 3)])
def test_synthetic_single_slice(material, slice_, seed):
    """Single-slice jobs with hashed capture and king-move exits (the form
    of the 8-piece workload): count lines and every code of the slice at the
    fixpoint equal the oracle's synthetic mode."""
    print(f"case synth {material} s{slice_} seed{seed}", flush=True)
    o = O.Table(material, synth_slice=slice_, seed=seed,
                log_path=f"/tmp/oracle_synth_{material}.log")
    with open(f"/tmp/oracle_synth_{material}.log") as f:
        olines = [l.strip() for l in f if l.strip()]
    t = tbg.Table(material, synth_slice=slice_, seed=seed, keep_codes=True)
    t.resolve([])
    lines = t.generate(tbg.FORM_BATCHED)
    assert lines == olines
    for stm in range(2):
        g, c = t.codes(stm)[0], o.codes(stm)[0]
        assert np.array_equal(g, c), (stm, int((g != c).sum()))
    t.close()


@pytest.mark.parametrize("mode", [tbg.SCATTER_GLOBAL, tbg.SCATTER_BLOCK,
                                  tbg.SCATTER_GLOBAL | tbg.SCATTER_COMBINE,
                                  tbg.SCATTER_BLOCK | tbg.SCATTER_COMBINE])
def test_scatter_modes_agree(registry, mode):
    """Global-atomic and LDS block-privatised scatter give the same table."""
    material = "KQvKR"
    registry.get(material)
    print(f"case scatter {mode} {material}", flush=True)
    t = tbg.Table(material)
    t.set_scatter(mode)
    t.resolve(list(registry.tables.values()))
    lines = t.generate(tbg.FORM_BATCHED)
    assert lines == oracle_lines(material)
    assert_same_table(material, t)
    t.close()


def _tune_candidates():
    # the autotuner's candidates span these unit sizes, grid caps and both
    # scatter modes; results must not depend on any of them
    return [
        dict(verify_unit_words=256, sweep_unit_words=512, pred_wg_per_cu=1, uncap_wg_per_cu=1,
             verify_wg_per_cu=1, sweep_wg_per_cu=1, scatter=tbg.SCATTER_BLOCK),
        dict(verify_unit_words=4096, sweep_unit_words=16384, pred_wg_per_cu=4, uncap_wg_per_cu=32,
             verify_wg_per_cu=8, sweep_wg_per_cu=8,
             scatter=tbg.SCATTER_BLOCK | tbg.SCATTER_COMBINE),
        dict(verify_unit_words=65536, sweep_unit_words=65536, pred_wg_per_cu=8, uncap_wg_per_cu=4,
             verify_wg_per_cu=64, sweep_wg_per_cu=64, scatter=tbg.SCATTER_GLOBAL),
        dict(verify_unit_words=768, sweep_unit_words=1536, pred_wg_per_cu=3, uncap_wg_per_cu=16,
             verify_wg_per_cu=16, sweep_wg_per_cu=16,
             scatter=tbg.SCATTER_GLOBAL | tbg.SCATTER_COMBINE),
    ]


@pytest.mark.parametrize("gi", range(4))
def test_geometry_does_not_change_results(registry, gi):
    """Unit sizes (including ones that do not divide a slice), grid caps and
    the scatter mode change speed only."""
    material = "KQvKR"
    registry.get(material)
    geo = _tune_candidates()[gi]
    print(f"case geometry {gi} {geo}", flush=True)
    t = tbg.Table(material)
    t.set_geometry(tbg.Geometry(**geo))
    assert t.geometry().as_dict() == geo
    t.resolve(list(registry.tables.values()))
    for form in (tbg.FORM_BATCHED, tbg.FORM_GRAPH):
        lines = t.generate(form)
        assert lines == oracle_lines(material)
        assert_same_table(material, t)
    t.close()


def test_autotune_picks_and_caches(registry, tmp_path):
    """The autotuner times its candidates, applies one, appends it to the
    cache file, and a second call is a cache hit; the table generated
    afterwards is exact."""
    material = "KQvKR"
    registry.get(material)
    cache = str(tmp_path / "geometry.txt")
    print("case autotune KQvKR", flush=True)
    t = tbg.Table(material)
    t.resolve(list(registry.tables.values()))
    g, ms = t.autotune(max_plies=0, cache_path=cache)
    assert ms > 0
    assert t.geometry().as_dict() == g.as_dict()
    with open(cache) as f:
        rows = [l for l in f if l.startswith("tbg-geometry KQvKR ")]
    assert len(rows) == 1
    g2, ms2 = t.autotune(max_plies=0, cache_path=cache)
    assert ms2 == -1 and g2.as_dict() == g.as_dict()
    lines = t.generate(tbg.FORM_BATCHED)
    assert lines == oracle_lines(material)
    assert_same_table(material, t)
    t.close()


def test_graph_form(registry):
    """Each ply captured once in a hipGraph and replayed, with the step flags
    and the ply counter kept on the device."""
    material = "KQvKR"
    registry.get(material)
    print("case graph KQvKR", flush=True)
    t = tbg.Table(material)
    t.resolve(list(registry.tables.values()))
    lines = t.generate(tbg.FORM_GRAPH)
    assert lines == oracle_lines(material)
    assert_same_table(material, t)
    t.close()


@pytest.mark.parametrize("form", [tbg.FORM_BATCHED, tbg.FORM_GRAPH])
def test_batch_of_materials_equals_each_alone(registry, form):
    """Several 4-piece materials in one batch (every launch covers all of
    them): each table and count line equals the oracle's, as alone."""
    mats = ["KQvKR", "KRvKN", "KBNvK", "KRRvK", "KBBvK", "KQvKQ"]
    for m in mats:
        registry.get(m)
    base = [registry.tables[m] for m in ("KQvK", "KRvK", "KBvK", "KNvK")]
    print(f"case batch form {form}", flush=True)
    batch = [tbg.Table(m) for m in mats]
    for t in batch:
        t.resolve(base)
    all_lines = tbg.generate_batch(batch, form)
    for m, t, lines in zip(mats, batch, all_lines):
        assert lines == oracle_lines(m), m
        assert_same_table(m, t)
        t.close()


def test_host_driven_steps_follow_main_loop(registry):
    """tbg_step_L / tbg_step_W / tbg_end_ply driven by the reference's own
    ply loop (tb8gen.c:1414-1444) give the same table and lines."""
    material = "KRvKN"
    registry.get(material)
    print("case manual KRvKN", flush=True)
    t = tbg.Table(material)
    t.resolve(list(registry.tables.values()))
    c = t.init_illegal_mate()
    lines = [f"broken_w = {c.broken[0]}", f"broken_b = {c.broken[1]}",
             f"l0_w = {c.mates[0]}", f"l0_b = {c.mates[1]}"]

    def L(stm, n, more):
        k = t.step("L", stm, n, more)
        lines.append(f"l{n}_{'wb'[stm]} = {k}")
        return k != 0

    def W(stm, n, more):
        k = t.step("W", stm, n, more)
        lines.append(f"w{n}_{'wb'[stm]} = {k}")
        return k != 0

    more_ww = more_wb = more_lw = more_lb = True
    n = 1
    while True:
        forced = n == 101
        if n > 101 and not (more_ww or more_wb or more_lw or more_lb):
            break
        wb_next = L(0, n, more_lw) if (forced or more_lw) else False
        ww_next = L(1, n, more_lb) if (forced or more_lb) else False
        more_lb = W(0, n, more_ww) if (forced or more_ww) else False
        more_lw = W(1, n, more_wb) if (forced or more_wb) else False
        more_wb, more_ww = wb_next, ww_next
        t.end_ply()
        n += 1
    assert lines == oracle_lines(material)
    # manual steps do not record codes for plies through tbg_generate's
    # history, but the codes array is written by the kernels themselves
    assert_same_table(material, t)
    t.close()


SERVED_JOBS = [("KQRvKR", 170, 7), ("KRBvKN", 57, 11), ("KRRvKB", 300, 3), ("KQNvKB", 12, 5),
               ("KQRvKR", 440, 9), ("KBNvKN", 200, 2), ("KRRvKB", 13, 4), ("KQvKRN", 99, 8),
               ("KQRvKRN", 170, 5), ("KRBvKRN", 57, 9)]


def _oracle_synth(material, s, seed):
    log = f"/tmp/oracle_srv_{material}_{s}_{seed}.log"
    o = O.Table(material, synth_slice=s, seed=seed, log_path=log)
    with open(log) as f:
        return o, [l.strip() for l in f if l.strip()]


def test_served_jobs_same_alone_and_in_the_mix():
    """Jobs arriving while others run join the running batch; each job's
    codes and count lines equal the oracle's for that job alone."""
    print("case served mix", flush=True)
    srv = tbg.Server(budget_bytes=8 << 30, max_active=6)
    # four jobs at once, the rest 1 ms later, while the first ones still run
    # (the last two are 6-piece slices of 13,388,280 positions)
    ids = [srv.submit(m, s, seed, arrival_ms=0.0 if i < 4 else 1.0)
           for i, (m, s, seed) in enumerate(SERVED_JOBS)]
    st = srv.run(tbg.FORM_BATCHED)
    assert st.jobs_done == len(SERVED_JOBS) and st.jobs_rejected == 0
    assert st.peak_active > 1  # jobs overlapped on the device
    for jid, (m, s, seed) in zip(ids, SERVED_JOBS):
        o, olines = _oracle_synth(m, s, seed)
        assert srv.lines(jid) == olines, (m, s)
        for stm in range(2):
            assert np.array_equal(srv.codes(jid, stm), o.codes(stm)[0]), (m, s, stm)
        r = srv.result(jid)
        assert r.arrival_ms <= r.admit_ms <= r.done_ms
    srv.close()


def test_served_admission_under_small_budget():
    """A budget that holds one job at a time serialises admission without
    changing results; a job larger than the whole budget is rejected."""
    print("case served budget", flush=True)
    probe = tbg.Table("KQRvKR", synth_slice=170, seed=7)
    one = probe.info.device_bytes
    probe.close()
    srv = tbg.Server(budget_bytes=int(one * 1.6) + (1 << 20), max_active=8)
    jobs = SERVED_JOBS[:3] + [("KQRBvKQR", 170, 1)]  # the 7-piece slice never fits
    ids = [srv.submit(m, s, seed, arrival_ms=0.0) for (m, s, seed) in jobs]
    st = srv.run(tbg.FORM_BATCHED)
    assert st.jobs_rejected == 1 and st.jobs_done == 3
    assert srv.result(ids[3]).status == -6  # TBG_E_BUDGET
    assert st.peak_active == 1
    for jid, (m, s, seed) in zip(ids[:3], jobs[:3]):
        o, olines = _oracle_synth(m, s, seed)
        assert srv.lines(jid) == olines
        assert np.array_equal(srv.codes(jid, 0), o.codes(0)[0])
    srv.close()


def _slice_file(s):
    wk, bk = O.kksquare(s)
    return f"{'abcdefgh'[wk & 7]}{wk // 8 + 1}{'abcdefgh'[bk & 7]}{bk // 8 + 1}"


@pytest.mark.parametrize("stop", [2, 5, 101])
def test_checkpoint_restore_continue(registry, tmp_path, stop):
    """Stop a generation at a ply boundary, write it in the reference's
    slice-file tree, restore it into a fresh context and continue: the table
    and the count lines equal the uninterrupted generation's."""
    material = "KRvKN"
    registry.get(material)
    subs = list(registry.tables.values())
    print(f"case checkpoint {material} stop {stop}", flush=True)
    t = tbg.Table(material)
    t.resolve(subs)
    t.init_illegal_mate()
    t.continue_(tbg.FORM_BATCHED, stop_at=stop)
    t.checkpoint(tmp_path)
    state = (tmp_path / "tbg_state").read_text()
    n = int([l for l in state.splitlines() if l.startswith("n ")][0].split()[1])
    assert n >= stop
    # the frontier files carry the reference's names and the plane's bytes
    w = t.bitmap(tbg.PLANE_W, 0, 100).view(np.uint8)
    got = tbg.file_read(tmp_path / str(n - 1) / "W" / "w" / _slice_file(100), w.nbytes)
    assert np.array_equal(got, w)
    lines_a = t.continue_(tbg.FORM_GRAPH)
    t2 = tbg.Table(material)
    t2.restore(tmp_path)
    lines_b = t2.continue_(tbg.FORM_BATCHED)
    assert lines_a == oracle_lines(material)
    assert lines_b == oracle_lines(material)
    assert_same_table(material, t)
    assert_same_table(material, t2)
    t.close()
    t2.close()


def test_export_plies_reference_files(registry, tmp_path):
    """Every executed half-step's W_n / L_n slice files, 0/L (mates) and
    0/wins of a finished table, in the reference's names and frame format,
    equal the positions the oracle's codes give."""
    material = "KRvK"
    t = registry.get(material)
    print(f"case export {material}", flush=True)
    t.export_plies(tmp_path)
    ora = oracle_table(material)
    words = t.words
    for stm in range(2):
        codes = ora.codes(stm)
        for s in (0, 37, 200, 440, 455):
            name = _slice_file(s)

            def bits(mask):  # bit idx & 7 of byte idx >> 3 (src/kslice.h:30-55)
                b = np.zeros(words * 64, dtype=bool)
                b[:codes.shape[1]] = mask
                return np.packbits(b, bitorder="little")

            c = codes[s]
            got = tbg.file_read(tmp_path / "0" / "L" / "wb"[stm] / name, words * 8)
            assert np.array_equal(got, bits(c == 2)), (stm, s, "L0")
            got = tbg.file_read(tmp_path / "0" / "wins" / "wb"[stm] / name, words * 8)
            assert np.array_equal(got, bits((c == 3) | ((c & 3) == 1))), (stm, s, "wins")
            for n in range(1, t.counts.nplies + 1):
                for kind, val in (("W", 1), ("L", 2)):
                    f = tmp_path / str(n) / kind / "wb"[stm] / name
                    if f.exists():
                        got = tbg.file_read(f, words * 8)
                        assert np.array_equal(got, bits(c == ((n << 2) | val))), (stm, s, n, kind)
    assert (tmp_path / "1" / "L" / "w" / _slice_file(0)).exists()


def test_native_registry_schedules_dependencies():
    """The library's material scheduler builds every smaller material a
    table's captures reach, one batch per piece count, and the result equals
    the oracle: 3/4-piece tables live, KQRvKR through its digest."""
    print("case native registry", flush=True)
    reg = tbg.NativeRegistry(tbg.FORM_GRAPH)
    t = reg.get("KQvKR")
    assert reg.count() == 3  # KQvK, KRvK (for KvKR), KQvKR
    assert t.lines == oracle_lines("KQvKR")
    assert_same_table("KQvKR", t)
    t = reg.get("KRRvK")
    assert reg.count() == 4
    assert t.lines == oracle_lines("KRRvK")
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_KQRvKR.json")) as f:
        ref = json.load(f)
    t = reg.get("KQRvKR")  # adds KRvKR and KQRvK
    assert reg.count() == 7
    assert t.lines == ref["lines"]
    for stm in range(2):
        c = t.codes(stm)
        crc = [zlib.crc32(c[s].tobytes()) for s in range(c.shape[0])]
        assert crc == ref["crc32"][str(stm)]
    assert reg.get("KQvKR").ptr.value is not None and reg.count() == 7  # resident, not rebuilt
    reg.close()


@pytest.mark.parametrize("form", ["graph", "plain"])
def test_c_driver_prints_the_reference_lines(form):
    """The C driver (the reference's command line over libtbg) prints the
    oracle's count lines for a 4-piece material, its smaller materials
    generated by the library's scheduler."""
    import subprocess
    drv = os.path.join(os.path.dirname(tbg.__file__), "tb8gen")
    print(f"case driver {form}", flush=True)
    r = subprocess.run([drv, "-f", form, "KRvKN"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = [l for l in r.stdout.splitlines() if l.strip()]
    assert out[0] == "number of threads = 1"  # tb8gen.c:1324
    assert out[1:] == oracle_lines("KRvKN")
