"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

DRAW, ILLEGAL = 0, 3


def build():
    """Compile the oracle library and CLI (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.tbo_generate.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.tbo_find.argtypes = [ctypes.c_char_p]
        L.tbo_slice_size.restype = ctypes.c_uint64
        L.tbo_codes.restype = ctypes.POINTER(ctypes.c_uint16)
        L.tbo_codes.argtypes = [ctypes.c_int, ctypes.c_int]
        L.tbo_count.restype = ctypes.c_int64
        L.tbo_broken.restype = ctypes.c_int64
        L.tbo_mates.restype = ctypes.c_int64
        L.tbo_sq_to_idx.restype = ctypes.c_uint64
        L.tbo_sq_to_idx.argtypes = [ctypes.c_int, u8p]
        L.tbo_idx_to_sq.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, u8p]
        L.tbo_capt_idx_to_sq.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, u8p]
        L.tbo_capt_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        L.tbo_sub_words.restype = ctypes.c_uint64
        L.tbo_sub_offset_words.restype = ctypes.c_uint64
        L.tbo_sub_bitmap.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint64)]
        L.tbo_lookup_bulk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, u8p,
                                      ctypes.POINTER(ctypes.c_uint16)]
        L.tbo_synth_slice.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_char_p, ctypes.c_int]
        L.tbo_synth_hash.restype = ctypes.c_uint64
        L.tbo_synth_hash.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_uint64]
        L.tbo_time_plies.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
        L.tbo_synth_init_time.restype = ctypes.c_double
        L.tbo_synth_init_time.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64,
                                          ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_int64)]
        L.tbo_synth_check.restype = ctypes.c_int64
        L.tbo_synth_check.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64,
                                      ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_int64)]
        L.tbo_rtbw_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.tbo_rtbw_close.argtypes = [ctypes.c_int]
        L.tbo_rtbw_size.restype = ctypes.c_uint64
        L.tbo_rtbw_size.argtypes = [ctypes.c_int, ctypes.c_int]
        L.tbo_rtbw_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u8p, u8p]
        L.tbo_rtbw_roundtrip.restype = ctypes.c_uint64
        L.tbo_rtbw_roundtrip.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64)]
        L.tbo_rtbw_decode.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, u8p, u8p]
        L.tbo_rtbw_tables.argtypes = [u8p, ctypes.POINTER(ctypes.c_int8), u8p, u8p]
        L.tbo_rtbw_verify.restype = ctypes.c_uint64
        L.tbo_rtbw_verify.argtypes = [ctypes.c_int, ctypes.c_char_p,
                                      ctypes.POINTER(ctypes.c_uint16),
                                      ctypes.POINTER(ctypes.c_uint16), ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_int64)]
        _lib = L
    return _lib


def set_threads(n):
    lib().tbo_set_threads(int(n))


def set_ref_king_capture_index(on):
    lib().tbo_set_ref_king_capture_index(1 if on else 0)


def set_keep_synth_codes(on):
    lib().tbo_set_keep_synth_codes(1 if on else 0)


def free_all():
    lib().tbo_free_all()


class Table:
    """One generated material (and, implicitly, its smaller materials)."""

    def __init__(self, material, log_path=None, synth_slice=None, seed=0, max_plies=0):
        L = lib()
        if synth_slice is None:
            h = L.tbo_generate(material.encode(), log_path.encode() if log_path else None)
        else:
            h = L.tbo_synth_slice(material.encode(), int(synth_slice), int(seed),
                                  log_path.encode() if log_path else None, int(max_plies))
        if h < 0:
            raise RuntimeError(f"oracle generation of {material} failed ({h})")
        self._wrap(h, material, synth_slice)

    @classmethod
    def find(cls, material):
        """A table the oracle already generated (a smaller material comes
        with the larger ones), or None."""
        h = lib().tbo_find(material.encode())
        if h < 0:
            return None
        t = cls.__new__(cls)
        t._wrap(h, material, None)
        return t

    def _wrap(self, h, material, synth_slice):
        L = lib()
        self.h = h
        self.material = material
        self.synth_slice = synth_slice
        self.size = L.tbo_slice_size(h)
        self.num = L.tbo_num_pieces(h)
        pt = (ctypes.c_uint8 * 8)()
        L.tbo_piece_types(h, pt)
        self.pt = list(pt)[: self.num]
        self.numsets = L.tbo_num_sets(h)

    def codes(self, stm):
        """uint16 codes [462, size] (synthetic mode: [1, size])."""
        n = self.size * (1 if self.synth_slice is not None else 462)
        p = lib().tbo_codes(self.h, stm)
        a = np.ctypeslib.as_array(p, shape=(n,)).copy()
        return a.reshape(-1, self.size)

    def count(self, kind, stm, n):
        """kind 'L' or 'W'; None if the step did not run."""
        c = lib().tbo_count(self.h, 0 if kind == "L" else 1, stm, n)
        return None if c < 0 else int(c)

    @property
    def nplies(self):
        return lib().tbo_num_plies(self.h)

    def broken(self, stm):
        return int(lib().tbo_broken(self.h, stm))

    def mates(self, stm):
        return int(lib().tbo_mates(self.h, stm))

    def capt_size(self, k):
        v = ctypes.c_uint64()
        lib().tbo_capt_size(self.h, k, ctypes.byref(v))
        return v.value

    def idx_to_sq(self, s, idx):
        out = (ctypes.c_uint8 * 8)()
        if lib().tbo_idx_to_sq(self.h, s, idx, out):
            raise ValueError("index out of range")
        return list(out)[: self.num]

    def capt_idx_to_sq(self, k, s, idx):
        out = (ctypes.c_uint8 * 8)()
        if lib().tbo_capt_idx_to_sq(self.h, k, s, idx, out):
            raise ValueError("index out of range")
        return list(out)[: self.num]

    def sq_to_idx(self, sq):
        a = (ctypes.c_uint8 * 8)(*sq)
        return int(lib().tbo_sq_to_idx(self.h, a))

    def sub_words(self, colour):
        return int(lib().tbo_sub_words(self.h, colour))

    def sub_bitmap(self, kind, colour, s):
        w = self.sub_words(colour)
        out = np.zeros(w, dtype=np.uint64)
        lib().tbo_sub_bitmap(self.h, kind, colour, s,
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return out

    def lookup_bulk(self, stm, squares):
        """squares: uint8 array [n, num] in slot order -> uint16 codes
        (0xffff for placements with two pieces on one square)."""
        sq = np.ascontiguousarray(squares, dtype=np.uint8)
        n = sq.shape[0]
        out = np.zeros(n, dtype=np.uint16)
        lib().tbo_lookup_bulk(self.h, stm, n, sq.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
        return out


def kkidx(tri, bk):
    return lib().tbo_kkidx(tri, bk)


def kkmap(wk, bk):
    return lib().tbo_kkmap(wk, bk)


def kksquare(s):
    out = (ctypes.c_uint8 * 2)()
    lib().tbo_kksquare(s, out)
    return out[0], out[1]


def synth_hash(seed, stm, kind, region, idx):
    return int(lib().tbo_synth_hash(seed, stm, kind, region, idx))


def planes_from_codes(codes_w, codes_b, words):
    """The ten planes tbo_synth_check reads (wins after ply 1, W_1, L_1,
    W_2, L_2, each [white, black]) derived from a synthetic slice's codes."""
    out = []
    cs = (np.asarray(codes_w).ravel(), np.asarray(codes_b).ravel())

    def plane(mask):
        b = np.zeros(words * 64, dtype=bool)
        b[:mask.size] = mask
        return np.packbits(b, bitorder="little").view(np.uint64)

    for which in ("wins1", "W1", "L1", "W2", "L2"):
        for c in cs:
            if which == "wins1":
                m = (c == 3) | (c == 5)
            else:
                n = int(which[1])
                m = c == ((n << 2) | (1 if which[0] == "W" else 2))
            out.append(plane(m))
    return out


def synth_check(material, s, seed, planes, nsamples, sample_seed=1):
    """Sampled forward check of plies 1 and 2 of a synthetic slice against
    the engine's planes (see tbo_synth_check); returns (mismatches, checked,
    first mismatch (stm, plane, index, engine bit) or None)."""
    arrs = [np.ascontiguousarray(p, dtype=np.uint64) for p in planes]
    ptrs = (ctypes.c_void_p * 10)(*[a.ctypes.data for a in arrs])
    chk = ctypes.c_uint64()
    fb = (ctypes.c_int64 * 4)()
    bad = lib().tbo_synth_check(material.encode(), int(s), int(seed), ptrs, int(nsamples),
                                int(sample_seed), ctypes.byref(chk), fb)
    if bad < 0:
        raise RuntimeError(f"synth_check of {material} failed ({bad})")
    return int(bad), chk.value, (tuple(fb) if bad else None)


def synth_init_time(material, s, seed, lo, n):
    """Seconds of the oracle's init (illegal, mates, synthetic exits) over
    indices [lo, lo + n) of a synthetic slice, and the four counts."""
    c = (ctypes.c_int64 * 4)()
    sec = lib().tbo_synth_init_time(material.encode(), int(s), int(seed), int(lo), int(n), c)
    if sec < 0:
        raise RuntimeError(f"oracle init sample of {material} failed")
    return sec, list(c)


def time_plies(material, max_plies):
    """Seconds spent in the first max_plies plies of the ply loop and the
    half-steps executed (the smaller materials are generated untimed)."""
    sec = ctypes.c_double()
    hs = ctypes.c_int64()
    r = lib().tbo_time_plies(material.encode(), int(max_plies), ctypes.byref(sec), ctypes.byref(hs))
    if r:
        raise RuntimeError(f"oracle timing of {material} failed ({r})")
    return sec.value, hs.value


class WdlFile:
    """A Syzygy WDL file read by the oracle's restatement of the probing
    code's reader (pawnless tables)."""

    def __init__(self, path, name):
        h = lib().tbo_rtbw_open(str(path).encode(), name.encode())
        if h < 0:
            raise RuntimeError(f"cannot read {path} as {name} ({h})")
        self.h = h
        self.name = name

    def close(self):
        if self.h >= 0:
            lib().tbo_rtbw_close(self.h)
            self.h = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self, side):
        return int(lib().tbo_rtbw_size(self.h, side))

    def probe(self, stm, pt, sq):
        n = len(pt)
        a = (ctypes.c_uint8 * 8)(*pt)
        b = (ctypes.c_uint8 * 8)(*sq)
        return lib().tbo_rtbw_probe(self.h, stm, n, a, b)

    def roundtrip(self, side, stride=1):
        """(indices that decode and encode back to themselves, others)"""
        bad = ctypes.c_uint64()
        ok = lib().tbo_rtbw_roundtrip(self.h, side, int(stride), ctypes.byref(bad))
        return int(ok), bad.value

    def decode(self, side, idx):
        pt = (ctypes.c_uint8 * 8)()
        sq = (ctypes.c_uint8 * 8)()
        if lib().tbo_rtbw_decode(self.h, side, int(idx), pt, sq):
            raise ValueError("index out of range")
        return list(pt), list(sq)

    def verify(self, material, codes_w, codes_b, stride=1):
        """Probes every stride-th legal position of a table given by its
        codes ([462, slice_size] per side, the material's slot order);
        returns (checked, mismatches, first mismatch or None)."""
        cw = np.ascontiguousarray(codes_w, dtype=np.uint16).ravel()
        cb = np.ascontiguousarray(codes_b, dtype=np.uint16).ravel()
        bad = ctypes.c_uint64()
        fb = (ctypes.c_int64 * 5)()
        u16p = ctypes.POINTER(ctypes.c_uint16)
        n = lib().tbo_rtbw_verify(self.h, material.encode(), cw.ctypes.data_as(u16p),
                                  cb.ctypes.data_as(u16p), int(stride), ctypes.byref(bad), fb)
        first = tuple(fb) if bad.value else None
        return int(n), bad.value, first


def rtbw_tables():
    """(Triangle, OffDiag, Diag, Lower) of the oracle's file reader, 64 each."""
    tri = (ctypes.c_uint8 * 64)()
    off = (ctypes.c_int8 * 64)()
    dg = (ctypes.c_uint8 * 64)()
    lo = (ctypes.c_uint8 * 64)()
    lib().tbo_rtbw_tables(tri, off, dg, lo)
    return list(tri), list(off), list(dg), list(lo)
