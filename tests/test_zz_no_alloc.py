"""The entry point that allocates nothing (tbg_generate_batch_noalloc)
against the existing batched entry point (tbg_generate_batch): every code
and count line of every context equal, on two calls in a row, the second
with a smaller batch.  The first batch mixes whole 4-piece tables with a
5-piece synthetic slice."""
import numpy as np
import pytest

import tb8gen_amd as tbg

pytestmark = pytest.mark.gpu

FIRST = ["KQvKR", "KRvKN", "KBNvK", "KRRvK", "KBBvK", "KQvKQ", ("KQRvKR", 170, 7)]
SECOND = ["KRvKN", "KQvKQ", ("KRBvKN", 57, 11)]


def _tables(specs, base):
    out = []
    for sp in specs:
        if isinstance(sp, tuple):
            t = tbg.Table(sp[0], synth_slice=sp[1], seed=sp[2], keep_codes=True)
            t.resolve([])
        else:
            t = tbg.Table(sp)
            t.resolve(base)
        out.append(t)
    return out


def test_noalloc_batch_equals_batched_entry_point():
    reg = tbg.NativeRegistry(tbg.FORM_GRAPH)
    base = [reg.get(m) for m in ("KQvK", "KRvK", "KBvK", "KNvK")]
    for call, specs in enumerate((FIRST, SECOND)):
        print(f"case noalloc call {call}: {len(specs)} contexts", flush=True)
        ref = _tables(specs, base)
        ref_lines = tbg.generate_batch(ref, tbg.FORM_BATCHED)
        got = _tables(specs, base)
        got_lines = tbg.generate_batch_noalloc(got)
        for sp, a, b, la, lb in zip(specs, ref, got, ref_lines, got_lines):
            assert lb == la, sp
            for stm in range(2):
                assert np.array_equal(b.codes(stm), a.codes(stm)), (sp, stm)
        for t in ref + got:
            t.close()
    reg.close()


def test_noalloc_rejects_oversized_batch():
    t = tbg.Table("KRvK")
    with pytest.raises(tbg.TbgError, match="at most 256"):
        tbg.generate_batch_noalloc([t] * (tbg.NOALLOC_MAX_BATCH + 1))
    t.close()
