"""Request server on the GPU: every job's result is the same alone and in a
mix of concurrent jobs; errors stay with their job; admission honours the
HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _box_job(n, seed, bound=False):
    from py_fof_amd.synth import clumpy_nfw
    L = float(n) ** (1.0 / 3.0)
    xp = clumpy_nfw(n=n, L=L, seed=seed)
    kw = dict(L_Box=L, linking_length=0.2, np_halo_min=10, return_ang_mom=True)
    vp = np.asfortranarray(np.random.default_rng(seed).normal(0.0, 0.5, xp.shape))
    if bound:
        kw.update(bound_check=True, ng_global=4096.0)
    return xp, vp, kw


def _halo_job(k, seed):
    from py_fof_amd.synth import isolated_halos
    xs, vs = isolated_halos(n_halos=k, seed=seed, nmax=20000)
    return xs, vs, dict(L_Box=512.0, linking_length=0.2, np_halo_min=10, bound_check=True,
                        return_ang_mom=True, ng_global=4096.0)


def _alone(xp, vp, kw):
    import py_fof_amd as pf
    return pf._fof_core(xp, vp, **kw)


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2])


def test_mixed_jobs_equal_their_standalone_results():
    from py_fof_amd.serve import Server
    jobs = [_box_job(20_000, 1), _box_job(300_000, 2), _box_job(50_000, 3, bound=True),
            _halo_job(40, 4), _box_job(1_000_000, 5), _halo_job(120, 6), _box_job(8_000, 7),
            _box_job(150_000, 8, bound=True)]
    alone = [_alone(*j) for j in jobs]
    with Server(slots=4) as srv:
        # two rounds: all at once, then interleaved submit / collect
        handles = [srv.submit(xp, vp, **kw) for xp, vp, kw in jobs]
        for h, ref in zip(handles, alone):
            _same(h.result(with_halo_id=True), ref)
            assert h.queue_ms >= 0.0 and h.run_ms > 0.0
        pending = []
        for i, (xp, vp, kw) in enumerate(jobs[::-1]):
            pending.append((srv.submit(xp, vp, **kw), alone[len(jobs) - 1 - i]))
            if i % 3 == 2:
                h, ref = pending.pop(0)
                _same(h.result(with_halo_id=True), ref)
        for h, ref in pending:
            _same(h.result(with_halo_id=True), ref)
        assert srv.status()["admitted_bytes"] == 0


def test_failing_job_keeps_its_error_and_others_succeed():
    from py_fof_amd.serve import Server
    import py_fof_amd as pf
    good = _box_job(40_000, 11)
    # a candidate of more than 1000 members spanning 25 cells > s_15 = 17.152
    x = np.vstack([np.linspace(10.0, 35.0, 1050), np.zeros(1050), np.zeros(1050)])
    bad = (x, np.zeros((3, 1050)), dict(L_Box=512.0, linking_length=2.0, np_halo_min=10,
                                        bound_check=True))
    with Server(slots=2) as srv:
        h1 = srv.submit(good[0], good[1], **good[2])
        h2 = srv.submit(bad[0], bad[1], **bad[2])
        h3 = srv.submit(good[0], good[1], **good[2])
        with pytest.raises(ValueError, match="Candidate halo span exceeds 15 Mpc limit"):
            h2.result()
        ref = _alone(*good)
        _same(h1.result(with_halo_id=True), ref)
        _same(h3.result(with_halo_id=True), ref)
    # the same error as the direct call
    with pytest.raises(ValueError, match="Candidate halo span exceeds 15 Mpc limit"):
        pf.fof(bad[0], bad[1], **bad[2])


def test_budget_admission():
    from py_fof_amd.serve import Server, ServerError
    xp, vp, kw = _box_job(60_000, 21)
    probe = Server(slots=1, hbm_budget=1 << 40)
    fp = probe.footprint(xp, vp, **kw)
    probe.close()
    assert fp > 60_000 * 160
    # a budget that admits one such job at a time still completes them all
    with Server(slots=4, hbm_budget=fp + 1) as srv:
        hs = [srv.submit(xp, vp, **kw) for _ in range(5)]
        ref = _alone(xp, vp, kw)
        for h in hs:
            _same(h.result(with_halo_id=True), ref)
        # a job larger than the whole budget is refused at submission
        big = _box_job(200_000, 22)
        with pytest.raises(ServerError, match="exceeds the server's HBM budget"):
            srv.submit(big[0], big[1], **big[2])


def test_served_jobs_match_the_oracle():
    """Each job of a concurrent mix against the oracle: plain FoF jobs (one of
    12,000,000 particles) bit-exact in in_halo and halo_id, angular momentum
    to rtol 1e-9; unbinding jobs exact on their direct-summation halos and
    within 1e-4 of the particles overall."""
    import os
    import sys
    from oracle import oracle_fof, oracle_fof_bound, oracle_set_link_threads
    from py_fof_amd.serve import Server
    jobs = [_box_job(12_000_000, 31), _box_job(200_000, 32, bound=True), _halo_job(150, 33),
            _box_job(400_000, 34), _box_job(60_000, 35, bound=True), _box_job(30_000, 36)]
    print("case: served mix of 6 jobs, largest 12,000,000 particles", flush=True)
    sys.stdout.flush()
    with Server(slots=4) as srv:
        handles = [srv.submit(xp, vp, **kw) for xp, vp, kw in jobs]
        results = [h.result(with_halo_id=True) for h in handles]
    oracle_set_link_threads(min(16, os.cpu_count() or 1))
    try:
        for (xp, vp, kw), (inh, hid, am) in zip(jobs, results):
            xs = xp if isinstance(xp, list) else [xp]
            vs = vp if isinstance(vp, list) else [vp]
            x = np.asfortranarray(np.hstack(xs))
            v = np.asfortranarray(np.hstack(vs))
            nps = [a.shape[1] for a in xs]
            if not kw.get("bound_check"):
                o_in, o_hid, o_am = oracle_fof(x, nps, kw["L_Box"], 0.2, 10, vp=v)
                assert np.array_equal(inh, o_in.astype(bool)) and np.array_equal(hid, o_hid)
                assert np.allclose(am, o_am, rtol=1e-9, atol=1e-12 * (np.abs(o_am).max() + 1e-300))
                continue
            s15 = 15.0 * 0.67 * kw["ng_global"] / 300.0
            r = oracle_fof_bound(x, nps, kw["L_Box"], 0.2, 10, v, s15=s15)
            assert r["ierr"] == 0
            ref = r["in_halo"].astype(bool)
            mism = int(np.sum(inh != ref))
            print(f"bound job of {x.shape[1]} particles: {mism} differ from the oracle", flush=True)
            assert mism <= 1e-4 * x.shape[1]
            if np.all(r["gid"] == 0):
                assert np.array_equal(inh, ref) and np.array_equal(hid, r["halo_id"])
    finally:
        oracle_set_link_threads(0)


def test_resident_workspaces_stay_within_budget():
    """Admission counts the workspace each worker keeps between jobs: after a
    large job, small concurrent jobs and the large one again, the bytes the
    server claims never exceed its budget, and neither does the device
    memory it holds (hipMemGetInfo, through torch)."""
    import torch
    from py_fof_amd.serve import Server
    big = _box_job(3_000_000, 41)
    small = [_box_job(150_000, 42 + k) for k in range(6)]
    probe = Server(slots=1, hbm_budget=1 << 40)
    fp_big = probe.footprint(big[0], big[1], **big[2])
    fp_small = probe.footprint(small[0][0], small[0][1], **small[0][2])
    probe.close()
    budget = fp_big + 2 * fp_small
    ref_big = _alone(*big)
    ref_small = [_alone(*j) for j in small]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    peak_used = 0
    with Server(slots=3, hbm_budget=budget) as srv:
        def check():
            nonlocal peak_used
            st = srv.status()
            assert st["claimed_bytes"] <= budget, st
            peak_used = max(peak_used, free0 - torch.cuda.mem_get_info()[0])
        _same(srv.submit(*big[:2], **big[2]).result(with_halo_id=True), ref_big)
        check()
        hs = [srv.submit(xp, vp, **kw) for xp, vp, kw in small]
        for h, ref in zip(hs, ref_small):
            _same(h.result(with_halo_id=True), ref)
            check()
        _same(srv.submit(*big[:2], **big[2]).result(with_halo_id=True), ref_big)
        check()
        assert srv.status()["resident_bytes"] > 0
    print(f"budget {budget / 2**20:.0f} MiB, peak device memory held {peak_used / 2**20:.0f} MiB",
          flush=True)
    # streams, events and kernel arguments take a few MiB beyond the workspaces
    assert peak_used <= budget + (64 << 20)
