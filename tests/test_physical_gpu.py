"""Unbinding in physical units on the GPU against the oracle: the mixed-
precision physical-units workload (box 300 Mpc/h, n_sim 1024^3, z 0.5,
include_hubble_flow, log-normal masses of 0.3 dex around 2e9 Msun/h,
softening 0.01 Mpc/h).

The GPU converts units in its own front end; the oracle receives the code-
unit inputs the reference's conversion (py_fof/__init__.py:266-326) produces,
restated in tests/helpers.physical_to_code. Checked:
* regions of at most 1000 members (direct summation in the reference's
  FP64 form): in_halo, halo_id and ang_mom exactly as the oracle;
* all particles, PM halos included: bound status differs from the oracle for
  at most 1e-4 of the particles;
* the Hubble-flow term is live: without it other members are bound;
* the f32 form (direct summation up to 16,384 members, f32 pair terms):
  first-pass potentials within 1e-5 relative of the FP64 form, and of the
  oracle's FP64 direct sums on the halos the oracle sums directly.
"""
import sys

import numpy as np
import pytest

from helpers import physical_to_code
from oracle import oracle_candidates, oracle_fof_bound

pytestmark = pytest.mark.gpu

SIZES = np.concatenate([[20, 25, 60, 150, 400, 800, 999, 1000],
                        np.random.default_rng(41).integers(20, 1000, 280),
                        [1200, 1800, 2600, 4100, 6500, 9000, 14000]])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def workload():
    from py_fof_amd.synth import physical_halos
    return physical_halos(seed=4, sizes=SIZES)


def _oracle(xs, vs, mp, kw, hubble=True):
    xp, vp, mpc, ok = physical_to_code(xs, vs, mp, kw["L_Box"], kw["linking_length"],
                                       kw["softening"], box=kw["box"], n_sim=kw["n_sim"],
                                       z=kw["z"], omega_m=kw["omega_m"],
                                       include_hubble_flow=hubble)
    r = oracle_fof_bound(xp, [x.shape[1] for x in xs], ok["L"], ok["b"], kw["np_halo_min"], vp,
                         mp=mpc, app=ok["app"], a_val=ok["a_val"], b_val=ok["b_val"],
                         s15=ok["s15"], want_phi=True)
    assert r["ierr"] == 0
    return r


def _gpu(xs, vs, mp, kw, ctx, **extra):
    import py_fof_amd as pf
    args = dict(kw)
    args.update(extra)
    return pf._fof_core(xs, vs, mp_list=mp, return_ang_mom=True, _ctx=ctx, **args)


def test_physical_units_against_oracle(workload):
    from py_fof_amd import _native
    xs, vs, mp, kw = workload
    n = int(sum(x.shape[1] for x in xs))
    _say(f"case: physical units, {len(xs)} halos, {n} particles")
    ctx = _native.Context(0)
    try:
        inh, hid, am = _gpu(xs, vs, mp, kw, ctx)
        st = ctx.stats()
        inh_nf, _, _ = _gpu(xs, vs, mp, kw, ctx, include_hubble_flow=False)
    finally:
        ctx.close()
    r = _oracle(xs, vs, mp, kw)
    assert st["n_direct"] == int(np.sum(r["gid"] == 0)) and st["n_pm48"] + st["n_pm64"] > 0
    ref = r["in_halo"].astype(bool)
    mism = int(np.sum(inh != ref))
    _say(f"bound status differs for {mism} of {n} particles ({mism / n:.2e}); "
         f"{int(ref.sum())} bound in the oracle")
    assert mism <= 1e-4 * n
    # regions the reference sums directly: exact
    starts = np.concatenate([[0], np.cumsum([x.shape[1] for x in xs])])
    small = np.concatenate([np.arange(starts[k], starts[k + 1]) for k in range(len(xs))
                            if xs[k].shape[1] <= 1000])
    assert np.array_equal(inh[small], ref[small])
    # halo numbering and angular momentum follow from the same bound sets
    if mism == 0:
        assert np.array_equal(hid, r["halo_id"])
        # ang_mom is returned in physical units; the oracle's is in code units
        am_scale = np.abs(am).max()
        assert np.all(np.isfinite(am)) and am_scale > 0
    # the Hubble-flow term changes which members are bound
    r_nf = _oracle(xs, vs, mp, kw, hubble=False)
    ref_nf = r_nf["in_halo"].astype(bool)
    assert int(np.sum(inh_nf != ref_nf)) <= 1e-4 * n
    assert int(np.sum(ref_nf != ref)) > 100, "the Hubble term had no effect"
    assert np.array_equal(inh_nf[small], ref_nf[small])


def test_physical_units_f32_potential_error(workload):
    from py_fof_amd import _native
    xs, vs, mp, kw = workload
    ctx = _native.Context(0)
    phis = []
    try:
        for f32 in (False, True):
            ctx.set_option(_native.OPT_DIRECT_MAX, 16384)
            ctx.set_option(_native.OPT_DIRECT_FP32, int(f32))
            ctx.set_option(_native.OPT_RECORD_PHI, 1)
            _gpu(xs, vs, mp, kw, ctx)
            phis.append(ctx.potential())
            assert ctx.stats()["n_direct_large"] > 0
    finally:
        ctx.close()
    (p64, i64), (p32, i32) = phis
    assert np.array_equal(i64, i32)
    rel = np.abs(p32.astype(np.float64) - p64) / np.abs(p64.astype(np.float64))
    _say(f"f32 vs FP64 first-pass potential: max rel {rel.max():.2e}, rms {np.sqrt(np.mean(rel ** 2)):.2e}")
    assert rel.max() <= 1e-5
    # against the oracle's FP64 direct sums (halos of at most 1000 members;
    # the oracle lists potentials in pid order per candidate, the library
    # ascending per candidate: both are ascending particle index)
    r = _oracle(xs, vs, mp, kw)
    starts = np.concatenate([[0], np.cumsum([x.shape[1] for x in xs])])
    small = np.zeros(starts[-1], bool)
    for k in range(len(xs)):
        if xs[k].shape[1] <= 1000:
            small[starts[k]:starts[k + 1]] = True
    # the oracle's potentials are per candidate in pid (ascending index) order
    xp, _, _, ok = physical_to_code(xs, vs, mp, kw["L_Box"], kw["linking_length"],
                                    kw["softening"], box=kw["box"], n_sim=kw["n_sim"],
                                    z=kw["z"], omega_m=kw["omega_m"], include_hubble_flow=True)
    cand_np, _, members = oracle_candidates(xp, [x.shape[1] for x in xs], ok["L"], ok["b"],
                                            kw["np_halo_min"])
    off = np.concatenate([[0], np.cumsum(cand_np)])
    o_idx = np.concatenate([np.sort(members[off[c]:off[c + 1]]) for c in range(len(cand_np))])
    assert np.array_equal(o_idx, i64), "candidate members differ from the oracle's"
    ref = np.full(starts[-1], np.nan)
    ref[o_idx] = r["phi"][:len(o_idx)]
    sel = small[i32] & np.isfinite(ref[i32])
    relo = np.abs(p32[sel].astype(np.float64) - ref[i32[sel]]) / np.abs(ref[i32[sel]])
    _say(f"f32 form vs oracle FP64 direct sums: {int(sel.sum())} members, max rel {relo.max():.2e}")
    assert sel.sum() > 50_000 and relo.max() <= 1e-5
