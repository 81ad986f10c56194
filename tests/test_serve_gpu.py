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
    assert fp > 60_000 * 88
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
