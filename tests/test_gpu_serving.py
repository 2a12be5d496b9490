"""The served form: live submission while the serving loop runs, polling,
multi-slice jobs and preemption of admitted jobs to pinned host memory --
every job's codes and count lines equal the oracle's for that job alone."""
import threading
import time

import numpy as np
import pytest

import tb8gen_amd as tbg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

O.set_threads(16)
_oracle = {}


def oracle_slice(material, s, seed):
    key = (material, s, seed)
    if key not in _oracle:
        log = f"/tmp/oracle_serving_{material}_{s}_{seed}.log"
        o = O.Table(material, synth_slice=s, seed=seed, log_path=log)
        with open(log) as f:
            lines = [l.strip() for l in f if l.strip()]
        _oracle[key] = ([o.codes(0)[0], o.codes(1)[0]], lines)
    return _oracle[key]


def assert_job(srv, jid, material, slice_, seed, nslices=1):
    r = srv.result(jid)
    assert r.status == 0 and r.nslices == nslices
    for part in range(nslices):
        s = (slice_ + part) % 441
        codes, lines = oracle_slice(material, s, seed)
        assert srv.lines(jid, part) == lines, (material, s)
        for stm in range(2):
            assert np.array_equal(srv.codes(jid, stm, part), codes[stm]), (material, s, stm)


LIVE = [("KQRvKR", 170, 7), ("KRBvKN", 57, 11), ("KRRvKB", 300, 3), ("KQNvKB", 12, 5),
        ("KQRvKRN", 170, 5), ("KBNvKN", 200, 2), ("KRRvKB", 13, 4), ("KQvKRN", 99, 8)]


def test_live_submission_while_serving():
    """The serving loop runs on its own thread; jobs are submitted one by one
    while earlier ones are on the device, polled until done, and equal the
    oracle's."""
    print("case serving live", flush=True)
    srv = tbg.Server(budget_bytes=8 << 30, max_active=16)
    out = {}
    th = threading.Thread(target=lambda: out.setdefault("st", srv.serve(tbg.FORM_BATCHED)))
    th.start()
    ids = []
    for m, s, seed in LIVE:
        ids.append(srv.submit(m, s, seed, arrival_ms=-1.0))
        time.sleep(0.003)
    deadline = time.time() + 120
    while time.time() < deadline and any(srv.poll(j) is None for j in ids):
        time.sleep(0.005)
    srv.close_submissions()
    th.join(timeout=120)
    assert not th.is_alive()
    st = out["st"]
    assert st.jobs_done == len(LIVE)
    for jid, (m, s, seed) in zip(ids, LIVE):
        r = srv.poll(jid)
        assert r is not None and r.arrival_ms <= r.admit_ms <= r.done_ms
        assert_job(srv, jid, m, s, seed)
    with pytest.raises(tbg.TbgError):
        srv.submit("KRvK", 0, 1)  # closed to new jobs
    srv.close()


def test_multi_slice_jobs():
    """Jobs of several king slices (each with its own exits) in the batched
    and the plain server."""
    print("case serving multi-slice", flush=True)
    jobs = [("KQRvKR", 168, 7, 4), ("KRBvKN", 439, 3, 3), ("KQRvKRN", 60, 9, 2)]
    for form in (tbg.FORM_BATCHED, tbg.FORM_PLAIN):
        srv = tbg.Server(budget_bytes=8 << 30, max_active=16)
        ids = [srv.submit(m, s, seed, arrival_ms=0.0, nslices=k) for m, s, seed, k in jobs]
        st = srv.run(form)
        assert st.jobs_done == len(jobs)
        if form == tbg.FORM_PLAIN:
            assert st.peak_active == 1
        for jid, (m, s, seed, k) in zip(ids, jobs):
            assert_job(srv, jid, m, s, seed, k)
        srv.close()


def test_preemption_to_host_is_bit_exact():
    """A budget that holds the large two-slice job but not also the small
    one arriving after it: the large job is preempted to pinned host memory
    at a ply boundary, the small one runs, the large one is copied back and
    finishes -- both equal the oracle."""
    print("case serving preemption", flush=True)
    big = ("KQRvKRN", 170, 5, 2)
    small = ("KQRvKR", 170, 7, 1)
    probe = [tbg.Table(big[0], synth_slice=big[1], seed=big[2], keep_codes=True),
             tbg.Table(small[0], synth_slice=small[1], seed=small[2], keep_codes=True)]
    b_bytes = 2 * probe[0].info.device_bytes
    s_bytes = probe[1].info.device_bytes
    for t in probe:
        t.close()
    # admission estimates add 64 KiB per slice to what a context holds
    srv = tbg.Server(budget_bytes=b_bytes + (256 << 10) + s_bytes // 2, max_active=16)
    ib = srv.submit(*big[:3], arrival_ms=0.0, nslices=big[3])
    isml = srv.submit(*small[:3], arrival_ms=2.0, nslices=1)
    st = srv.run(tbg.FORM_BATCHED)
    print(f"  spills {st.spills} restores {st.restores} ({st.spill_ms:.2f} ms)", flush=True)
    assert st.jobs_done == 2 and st.spills >= 1 and st.restores >= 1
    rb, rs = srv.result(ib), srv.result(isml)
    assert rb.spills >= 1 and rs.spills == 0
    assert rs.done_ms < rb.done_ms
    assert_job(srv, ib, *big)
    assert_job(srv, isml, *small[:3])
    srv.close()
