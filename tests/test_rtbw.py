"""Syzygy WDL files.

CPU: the oracle's restatement of the reference's table reader (src/probe.c)
against entries of the reference's encoding tables, index round trips, and
files written here by a small independent writer of the format.
GPU: the engine's final-table writer (tbg_write_rtbw), every legal position
of its files read back by that reader and compared with the engine's codes.
"""
import heapq
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

KNIGHT, BISHOP, ROOK, QUEEN, KING = 2, 3, 4, 5, 6
CODE = {"N": KNIGHT, "B": BISHOP, "R": ROOK, "Q": QUEEN, "K": KING}


def sq(name):
    return "abcdefgh".index(name[0]) + 8 * (int(name[1]) - 1)


# ------------------------------------------------------------ tables

def test_encoding_tables_match_reference_entries():
    """Spot entries of Triangle, OffDiag, Diag and Lower (src/probe.c:300-376)."""
    tri, off, dg, lo = O.rtbw_tables()
    want_tri = {"a1": 6, "b1": 0, "c1": 1, "d1": 2, "e1": 2, "h1": 6, "b2": 7, "c2": 3, "d2": 4,
                "d3": 5, "d4": 9, "e5": 9, "h8": 6, "a8": 6, "g7": 7, "c3": 8, "a2": 0}
    assert {k: tri[sq(k)] for k in want_tri} == want_tri
    want_off = {"a1": 0, "b1": -1, "a2": 1, "h8": 0, "h1": -1, "a8": 1, "e4": -1, "d5": 1}
    assert {k: off[sq(k)] for k in want_off} == want_off
    want_dg = {"a1": 0, "h1": 8, "b2": 1, "g2": 9, "d4": 3, "e4": 11, "a8": 15, "h8": 7, "c1": 0}
    assert {k: dg[sq(k)] for k in want_dg} == want_dg
    want_lo = {"a1": 28, "b1": 0, "h1": 6, "a2": 0, "b2": 29, "c2": 7, "h2": 12, "a3": 1, "b3": 7,
               "c3": 30, "d3": 13, "a8": 6, "g8": 27, "h8": 35, "f6": 33, "g6": 25}
    assert {k: lo[sq(k)] for k in want_lo} == want_lo


# ------------------------------------------------ a writer of the format

def header_layout(name, order=None):
    """Header pieces (kings, then a unique piece unless every other piece has
    a twin, then runs of like pieces), the lead size and group sizes."""
    w, b = name.split("v")
    codes = [KING, KING | 8]
    rest = [CODE[c] for c in w[1:]] + [CODE[c] | 8 for c in b[1:]]
    cnt = {c: (codes + rest).count(c) for c in set(codes + rest)}
    uniq = [c for c in rest if cnt[c] == 1]
    lead = 2
    if uniq:
        rest.remove(uniq[0])
        codes.append(uniq[0])
        lead = 3
    pieces = codes + rest
    groups, i, free = [], lead, 64 - lead
    while i < len(pieces):
        m = 1
        while i + m < len(pieces) and pieces[i + m] == pieces[i]:
            m += 1
        groups.append(math.comb(free, m))
        free -= m
        i += m
    size = (462 if lead == 2 else 31332) * math.prod(groups)
    if order is None:
        order = len(groups)
    return pieces, order, size


def huffman_lengths(counts):
    heap = [(c, v, (v,)) for v, c in counts.items()]
    heapq.heapify(heap)
    depth = {v: 0 for v in counts}
    tie = 100
    while len(heap) > 1:
        c1, _, a = heapq.heappop(heap)
        c2, _, b = heapq.heappop(heap)
        for v in a + b:
            depth[v] += 1
        heapq.heappush(heap, (c1 + c2, tie, a + b))
        tie += 1
    return depth


def pack_side(vals, block_log, idx_bits):
    vals = [int(v) for v in vals]
    counts = {}
    for v in vals:
        counts[v] = counts.get(v, 0) + 1
    if len(counts) == 1:
        return bytes([0x80, vals[0]]), b"", b"", b""
    ln = huffman_lengths(counts)
    minl, maxl = min(ln.values()), max(ln.values())
    syms = [v for l in range(maxl, minl - 1, -1) for v in sorted(counts) if ln[v] == l]
    h = maxl - minl + 1
    nlen = [sum(1 for v in syms if ln[v] == minl + i) for i in range(h)]
    offset, base = [0] * h, [0] * h
    for i in range(h - 2, -1, -1):
        offset[i] = offset[i + 1] + nlen[i + 1]
        base[i] = (base[i + 1] + nlen[i + 1]) // 2
    assert base[0] + nlen[0] == 1 << minl
    code = {v: base[ln[v] - minl] + syms.index(v) - offset[ln[v] - minl] for v in syms}
    cap = 8 << block_log
    starts, bits, cnt = [0], 0, 0
    for i, v in enumerate(vals):
        if bits + ln[v] > cap or cnt == 65536:
            starts.append(i)
            bits = cnt = 0
        bits += ln[v]
        cnt += 1
    nb = len(starts)
    starts.append(len(vals))
    data = bytearray(nb << block_log)
    for blk in range(nb):
        pos = blk << (block_log + 3)
        for v in vals[starts[blk]:starts[blk + 1]]:
            for t in range(ln[v] - 1, -1, -1):
                if (code[v] >> t) & 1:
                    data[pos >> 3] |= 0x80 >> (pos & 7)
                pos += 1
    sizes = b"".join((starts[b + 1] - starts[b] - 1).to_bytes(2, "little") for b in range(nb))
    sizes += (65535).to_bytes(2, "little")
    index = bytearray()
    n = len(vals)
    for m in range((n + (1 << idx_bits) - 1) >> idx_bits):
        P = (m << idx_bits) + (1 << (idx_bits - 1))
        b = nb
        if P < n:
            b = max(i for i in range(nb) if starts[i] <= P)
        index += b.to_bytes(4, "little") + (P - starts[b]).to_bytes(2, "little")
    head = bytearray([0, block_log, idx_bits, 1]) + nb.to_bytes(4, "little")
    head += bytes([maxl, minl]) + b"".join(o.to_bytes(2, "little") for o in offset)
    head += len(syms).to_bytes(2, "little")
    for v in syms:
        head += bytes([v, 0xF0, 0xFF])
    if len(syms) & 1:
        head += b"\0"
    return bytes(head), bytes(index), sizes, bytes(data)


def write_wdl(path, pieces, order, sides, block_log=6, idx_bits=10):
    """sides: one value array (0..4) per sub-table (1: symmetric material)."""
    split = len(sides) == 2
    f = bytearray((0x5D23E871).to_bytes(4, "little"))
    f.append(1 if split else 0)
    f.append(order | (order << 4 if split else 0))
    f += bytes(p | (p << 4 if split else 0) for p in pieces)
    if len(f) & 1:
        f.append(0)
    packed = [pack_side(v, block_log, idx_bits) for v in sides]
    for part in range(3):
        for pk in packed:
            f += pk[part]
    for pk in packed:
        f += bytes((-len(f)) % 64)
        f += pk[3]
    f += bytes(64)
    with open(path, "wb") as fh:
        fh.write(f)


# ---------------------------------------------------------------- tests

@pytest.mark.parametrize("name,order", [("KQvK", None), ("KRvKN", None), ("KRvKN", 0),
                                        ("KRRvK", None), ("KRRvK", 0), ("KQvKQ", None),
                                        ("KRNvKB", None), ("KRNvKB", 1)])
def test_index_round_trip(tmp_path, name, order):
    """decode then encode is the identity on every index (stride sampled for
    the larger spaces), whichever digit the leading group is."""
    pieces, order, size = header_layout(name, order)
    sym = name.split("v")[0][1:] == name.split("v")[1][1:]
    path = tmp_path / (name + ".rtbw")
    write_wdl(path, pieces, order, [[2]] * (1 if sym else 2))
    f = O.WdlFile(path, name)
    assert f.size(0) == size
    stride = max(1, size // 2_000_000)
    for side in range(1 if sym else 2):
        ok, bad = f.roundtrip(side, stride)
        assert bad == 0 and ok == len(range(0, size, stride))


@pytest.mark.parametrize("name,block_log,idx_bits", [("KRvK", 6, 10), ("KRvK", 5, 4),
                                                     ("KQvK", 6, 10)])
def test_huffman_file_probe(tmp_path, name, block_log, idx_bits):
    """Every position of a file with skewed random values probes back to its
    value: Huffman blocks, the size and index tables, and both sub-tables."""
    pieces, order, size = header_layout(name)
    sym = name.split("v")[0][1:] == name.split("v")[1][1:]
    rng = np.random.default_rng(7)
    sides = [rng.choice(5, size=size, p=[0.05, 0.1, 0.6, 0.15, 0.1]).astype(np.uint8)
             for _ in range(1 if sym else 2)]
    path = tmp_path / (name + ".rtbw")
    write_wdl(path, pieces, order, sides, block_log, idx_bits)
    f = O.WdlFile(path, name)
    for side, vals in enumerate(sides):
        for i in range(0, size, 7):
            pt, sqs = f.decode(side, i)
            n = len(pieces)
            assert f.probe(side, pt[:n], sqs[:n]) == int(vals[i]) - 2, (side, i)


def test_huffman_file_probe_kk(tmp_path):
    """Kings-led encoding (every non-king piece has a twin), many blocks."""
    name = "KRRvK"
    pieces, order, size = header_layout(name)
    vals = [(i * 7919 + (i >> 5)) % 5 for i in range(size)]
    vals = np.array(vals, dtype=np.uint8)
    vals[vals == 4] = 2  # four distinct values
    path = tmp_path / (name + ".rtbw")
    write_wdl(path, pieces, order, [vals, vals[::-1].copy()], 6, 8)
    f = O.WdlFile(path, name)
    for side in range(2):
        v = vals if side == 0 else vals[::-1]
        for i in range(0, size, 997):
            pt, sqs = f.decode(side, i)
            assert f.probe(side, pt[:4], sqs[:4]) == int(v[i]) - 2, (side, i)


def test_colour_flip_probe(tmp_path):
    """A position whose colours are the reverse of the file name's reads the
    other sub-table with the board mirrored (probe_table's flip)."""
    name = "KRvKN"
    pieces, order, size = header_layout(name)
    sides = [np.full(size, 4, np.uint8), np.full(size, 0, np.uint8)]
    sides[0][::3] = 1
    sides[1][::5] = 3
    path = tmp_path / (name + ".rtbw")
    write_wdl(path, pieces, order, sides)
    f = O.WdlFile(path, name)
    for i in range(0, size, 11):
        pt, sqs = f.decode(0, i)
        # the same position with colours swapped and the board mirrored,
        # black to move, reads sub-table 0 at index i
        pt2 = [p ^ 8 for p in pt[:4]]
        sq2 = [s ^ 0x38 for s in sqs[:4]]
        assert f.probe(1, pt2, sq2) == int(sides[0][i]) - 2
        pt, sqs = f.decode(1, i)
        assert f.probe(0, [p ^ 8 for p in pt[:4]], [s ^ 0x38 for s in sqs[:4]]) == \
            int(sides[1][i]) - 2


# ------------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def registry():
    import tb8gen_amd as tbg
    print("registry: batched form", flush=True)
    r = tbg.NativeRegistry(tbg.FORM_BATCHED)
    yield r
    r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("material,name", [("KQvK", "KQvK"), ("KRvK", "KRvK"), ("KBvK", "KBvK"),
                                           ("KQvKR", "KQvKR"), ("KRvKQ", "KQvKR"),
                                           ("KRRvK", "KRRvK"), ("KQvKQ", "KQvKQ"),
                                           ("KBNvK", "KBNvK"), ("KRvKN", "KRvKN")])
def test_engine_wdl_file(registry, tmp_path, material, name):
    """The engine's WDL file of a finished table holds the engine's value of
    every legal position, as the reference's reader decodes it."""
    print(f"case {material}", flush=True)
    t = registry.get(material)
    assert t.rtbw_name() == name
    path = t.write_rtbw(tmp_path)
    assert os.path.basename(path) == name + ".rtbw"
    f = O.WdlFile(path, name)
    c0, c1 = t.codes(0), t.codes(1)
    legal = int((c0 != 3).sum() + (c1 != 3).sum())
    n, bad, first = f.verify(material, c0, c1)
    assert bad == 0, f"first mismatch (stm, slice, idx, code, file value) = {first}"
    assert n == legal


@pytest.mark.gpu
@pytest.mark.parametrize("material,cursed", [("KQRvKR", False), ("KBNvKN", True)])
def test_engine_wdl_file_five_piece(registry, tmp_path, material, cursed):
    """Five-piece tables, every 13th position; KBNvKN has conversions after
    more than 100 plies (cursed wins, blessed losses: values 1 and -1)."""
    print(f"case {material}", flush=True)
    t = registry.get(material)
    path = t.write_rtbw(tmp_path)
    f = O.WdlFile(path, material)
    c0, c1 = t.codes(0), t.codes(1)
    n, bad, first = f.verify(material, c0, c1, stride=13)
    assert bad == 0, f"first mismatch (stm, slice, idx, code, file value) = {first}"
    assert n > 10_000_000
    w = np.concatenate([c0.ravel(), c1.ravel()])
    tag, ply = w & 3, w >> 2
    assert (((tag == 1) | (tag == 2)) & (ply > 100)).any() == cursed
