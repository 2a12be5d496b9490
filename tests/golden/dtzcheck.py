"""ctypes binding of the DTZ50 consistency checker (tests/golden/dtzcheck.c).

Test infrastructure only: the codes of a finished table (the engine's or the
oracle's) are checked position by position against the rules, with the
brute-force solver's move generator and no code of the oracle or the engine.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdtzcheck.so")
STATS = ["checked", "illegal", "win", "cursed_win", "loss", "blessed_loss", "draw",
         "capture_into_cursed_or_blessed"]
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB)
        u16p = ctypes.POINTER(ctypes.c_uint16)
        L.dc_check.restype = ctypes.c_int64
        L.dc_check.argtypes = [ctypes.c_char_p, u16p, u16p, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(u16p),
                               ctypes.POINTER(u16p), ctypes.c_uint64, ctypes.c_uint64,
                               ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                               ctypes.POINTER(ctypes.c_uint64)]
        L.dc_check_list.restype = ctypes.c_int64
        L.dc_check_list.argtypes = L.dc_check.argtypes[:7] + [ctypes.POINTER(ctypes.c_uint64)] + \
            L.dc_check.argtypes[7:]
        L.dc_slice_size.restype = ctypes.c_uint64
        L.dc_slice_size.argtypes = [ctypes.c_char_p]
        L.dc_unrank.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64,
                                ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def slice_size(material):
    return int(lib().dc_slice_size(material.encode()))


def unrank(material, s, idx):
    out = (ctypes.c_int * 8)()
    if lib().dc_unrank(material.encode(), int(s), int(idx), out):
        raise ValueError("index out of range")
    return list(out)[: sum(ch != "v" for ch in material)]


def _u16(a):
    a = np.ascontiguousarray(a, dtype=np.uint16).ravel()
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


def check(material, codes_w, codes_b, subs=(), start=0, stride=1, count=None, threads=8,
          positions=None):
    """Checks positions start, start + stride, ... (count of them; all by
    default) of the flattened [462 * slice_size] code arrays, or the
    flattened indices in `positions`.  subs: tables
    one capture smaller as (material, codes_w, codes_b).  Returns
    (mismatches, stats dict, first mismatch (stm, slice, index, code,
    rules' code) or None)."""
    keep = []
    cw, pw = _u16(codes_w)
    cb, pb = _u16(codes_b)
    keep += [cw, cb]
    n = len(subs)
    mats = (ctypes.c_char_p * max(1, n))()
    sw = (ctypes.POINTER(ctypes.c_uint16) * max(1, n))()
    sb = (ctypes.POINTER(ctypes.c_uint16) * max(1, n))()
    for i, (m, w, b) in enumerate(subs):
        mats[i] = m.encode()
        aw, qw = _u16(w)
        ab, qb = _u16(b)
        keep += [aw, ab]
        sw[i], sb[i] = qw, qb
    total = cw.size
    if count is None:
        count = (total - start + stride - 1) // stride
    fb = (ctypes.c_int64 * 5)()
    st = (ctypes.c_uint64 * 8)()
    if positions is not None:
        pos = np.ascontiguousarray(positions, dtype=np.uint64)
        keep.append(pos)
        bad = lib().dc_check_list(material.encode(), pw, pb, n, mats, sw, sb,
                                  pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 0, 1,
                                  pos.size, int(threads), fb, st)
    else:
        bad = lib().dc_check(material.encode(), pw, pb, n, mats, sw, sb, int(start),
                             int(stride), int(count), int(threads), fb, st)
    if bad < 0:
        raise RuntimeError(f"dtzcheck of {material} failed ({bad})")
    return int(bad), dict(zip(STATS, list(st))), (tuple(fb) if bad else None)
