"""Writes the golden fixtures under tests/golden/.

1. naive_<M>.npz -- from naive.c, the independent full-board brute-force
   solver (no king slices, no symmetry, forward evaluation).  3-piece
   materials keep the whole code array; 4-piece materials keep a SHA-256 of
   each side's array, the code histogram and a seeded sample of positions.
   These pin the oracle (tests/test_oracle_golden.py).
2. oracle_<M>.json -- digests of the CPU oracle's own tables for materials
   too large to regenerate inside a GPU test (per-slice CRC32 of the codes
   and the reference-format count lines).  These pin the HIP engine at that
   size (tests/test_gpu_parity.py).
3. synth_<M>_s<slice>_<seed>.json -- the same digest of one synthetic king
   slice (seeded hash exits) run by the oracle to its full fixpoint: the
   count lines and the CRC32 of every run of CHUNK positions of each side's
   codes.  The 7-piece KQRBvKQR slice 170 (776,520,240 positions per side)
   exercises the code that only slices above 2^22 bitmap words reach.

4. kkidx_probe_c.json -- the values of the reference's KKIdx table
   (src/probe.c:441-522, the king-slice numbering), read from the
   reference's source text as data: 10 rows of 64 entries.

Usage: python tests/golden/make_golden.py [naive|oracle|synth|kkidx] [MATERIAL ...]
       (synth: MATERIAL:SLICE:SEED)
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NAIVE_3 = ["KQvK", "KRvK", "KBvK", "KNvK"]
NAIVE_4 = ["KQvKR", "KRvKN", "KBNvK", "KRRvK", "KBBvK", "KQvKQ"]
ORACLE = ["KQRvKR"]
SYNTH = ["KQRBvKQR:170:20261203"]
SAMPLE = 4096
CHUNK = 1 << 24


def build_naive(outdir):
    exe = os.path.join(outdir, "naive")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-o", exe, os.path.join(HERE, "naive.c")], check=True)
    return exe


def run_naive(exe, material, outdir):
    path = os.path.join(outdir, f"{material}.bin")
    if not os.path.exists(path):
        subprocess.run([exe, material, path], check=True)
    k = sum(ch != "v" for ch in material)
    return np.fromfile(path, dtype=np.int16).reshape(2, 64 ** k)


def naive_fixture(material, arr):
    out = {"material": np.array(material)}
    for stm in range(2):
        a = arr[stm]
        out[f"sha256_{stm}"] = np.array(hashlib.sha256(a.tobytes()).hexdigest())
        vals, cnts = np.unique(a, return_counts=True)
        out[f"hist_codes_{stm}"] = vals.astype(np.int16)
        out[f"hist_counts_{stm}"] = cnts.astype(np.int64)
        if a.size <= 64 ** 3:
            out[f"codes_{stm}"] = a
        else:
            rng = np.random.default_rng(20261203 + stm)
            idx = rng.integers(0, a.size, SAMPLE)
            out[f"sample_idx_{stm}"] = idx.astype(np.int64)
            out[f"sample_code_{stm}"] = a[idx]
    np.savez_compressed(os.path.join(HERE, f"naive_{material}.npz"), **out)


def oracle_fixture(material):
    from oracle import oracle as O
    O.set_threads(os.cpu_count() or 8)
    log = os.path.join(tempfile.gettempdir(), f"golden_{material}.log")
    t = O.Table(material, log_path=log)
    with open(log) as f:
        lines = [l.strip() for l in f if l.strip()]
    crc = {}
    for stm in range(2):
        c = t.codes(stm)
        crc[str(stm)] = [zlib.crc32(c[s].tobytes()) for s in range(c.shape[0])]
    with open(os.path.join(HERE, f"oracle_{material}.json"), "w") as f:
        json.dump({"material": material, "slice_size": int(t.size), "lines": lines,
                   "crc32": crc}, f, separators=(",", ":"))


def synth_fixture(spec):
    from oracle import oracle as O
    material, s, seed = spec.split(":")
    s, seed = int(s), int(seed)
    O.set_threads(os.cpu_count() or 8)
    log = os.path.join(tempfile.gettempdir(), f"golden_synth_{material}.log")
    t = O.Table(material, synth_slice=s, seed=seed, log_path=log)
    with open(log) as f:
        lines = [l.strip() for l in f if l.strip()]
    crc = {}
    for stm in range(2):
        c = t.codes(stm).ravel()
        crc[str(stm)] = [zlib.crc32(c[i:i + CHUNK].tobytes()) for i in range(0, c.size, CHUNK)]
    name = f"synth_{material}_s{s}_{seed}.json"
    with open(os.path.join(HERE, name), "w") as f:
        json.dump({"material": material, "slice": s, "seed": seed, "slice_size": int(t.size),
                   "chunk": CHUNK, "lines": lines, "crc32": crc}, f, separators=(",", ":"))


def kkidx_fixture(probe_c="/root/reference/src/probe.c"):
    """The numbers of the KKIdx and InvTriangle initialisers, in order, read
    from the reference's text as data."""
    import re
    with open(probe_c) as f:
        text = f.read()

    def table(decl, n):
        start = text.index(decl)
        end = text.index("};", start)
        vals = [int(v) for v in re.findall(r"-?\d+", text[start + len(decl):end])]
        assert len(vals) == n, (decl, len(vals))
        return vals, f"src/probe.c:{text[:start].count(chr(10)) + 1}-{text[:end].count(chr(10)) + 1}"

    kk, kk_src = table("KKIdx[10][64] = {", 640)
    inv, inv_src = table("InvTriangle[10] = {", 10)
    with open(os.path.join(HERE, "kkidx_probe_c.json"), "w") as f:
        json.dump({"kkidx_source": kk_src, "kkidx": [kk[i * 64:(i + 1) * 64] for i in range(10)],
                   "inv_triangle_source": inv_src, "inv_triangle": inv}, f)


def main():
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    mats = sys.argv[2:]
    if what in ("all", "naive"):
        work = os.environ.get("NAIVE_WORKDIR", tempfile.gettempdir())
        exe = build_naive(work)
        for m in mats or NAIVE_3 + NAIVE_4:
            print("naive", m, flush=True)
            naive_fixture(m, run_naive(exe, m, work))
    if what in ("all", "oracle"):
        for m in mats or ORACLE:
            print("oracle", m, flush=True)
            oracle_fixture(m)
    if what in ("all", "kkidx"):
        kkidx_fixture()
    if what in ("all", "synth"):
        for spec in mats or SYNTH:
            print("synth", spec, flush=True)
            synth_fixture(spec)


if __name__ == "__main__":
    main()
