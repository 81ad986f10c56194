"""FoF at the benchmark's full size (134,217,728 particles, L = 512, b = 0.2).

Bit-exact against the oracle through the golden fixture
tests/golden/c2_nfw134M.json (written by tests/golden/make_c2_fixture.py,
which ran the oracle on this input for about 5 minutes on 8 cores): the
SHA-256 digests of the input, of the partition roots, of halo_id and of
in_halo, and the halo and member counts. The input digest is checked first,
so a generator that does not reproduce the same float64 positions on this
host fails loudly instead of comparing other data.

Size-independent properties are checked as well:

* the output is deterministic: a second call gives the same halo_id;
* the partition roots (per particle, the input index of its component's
  first CRS member) are consistent with in_halo and halo_id: a particle is in
  a halo exactly when its component has np_halo_min members or more, and
  halos are numbered 1..nh;
* closure on a sample: every particle within b of a sampled particle shares
  its component;
* no member of a component of two or more is isolated: each sampled such
  member has a partner within b.
The predicate works on f32 positions relative to the region's anchor (as the
reference's real(4) coordinates), up to 256 in magnitude where one ulp is
3.1e-5: a distance can move by up to 5.3e-5, 2.6e-4 of b. The float64
distances here therefore use b (1 -+ 1e-3) as the two bounds.
The partners come from a numpy cell grid of side 1 (> b), searched over the
27 neighbouring cells with periodic minimum images in float64.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 134_217_728
L = 512.0
B = 0.2
NP_MIN = 10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_nfw134M.json")


def _digest_input(xp):
    h = hashlib.sha256()
    flat = xp.reshape(-1, order="F")
    step = 1 << 24
    for s in range(0, flat.size, step):
        h.update(np.ascontiguousarray(flat[s:s + step]).tobytes())
    return h.hexdigest()


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()


def test_fof_134m_bit_exact_and_properties():
    import py_fof_amd as pf
    from py_fof_amd.synth import clumpy_nfw
    with open(FIXTURE) as f:
        gold = json.load(f)
    assert gold["n"] == N and gold["L_Box"] == L and gold["seed"] == 2
    _say("case: 134,217,728-particle NFW-clump set, seed 2")
    xp = clumpy_nfw(n=N, L=L, seed=2)
    assert _digest_input(xp) == gold["input_sha256"], \
        "the generator gave other positions on this host: the fixture does not apply"
    in_h, hid, _, ex = pf._fof_core(xp, L_Box=L, linking_length=B, np_halo_min=NP_MIN,
                                    _want_roots=True)
    roots = ex["roots"]
    assert int(hid.max()) == gold["nhalo"] and int(in_h.sum()) == gold["members"]
    assert _digest(roots) == gold["roots_sha256"], "partition differs from the oracle"
    assert _digest(hid) == gold["halo_id_sha256"], "halo_id differs from the oracle"
    assert _digest(in_h) == gold["in_halo_sha256"], "in_halo differs from the oracle"
    _say(f"bit-exact against the oracle: {gold['nhalo']} halos, {gold['members']} members")
    _, hid2, _ = pf._fof_core(xp, L_Box=L, linking_length=B, np_halo_min=NP_MIN)
    assert np.array_equal(hid, hid2), "second call differs"
    del hid2
    _say("deterministic")

    # roots against in_halo / halo_id
    assert roots.min() >= 0 and roots.max() < N
    assert np.all(roots[roots] == roots), "a root is not its own root"
    size = np.bincount(roots, minlength=N)
    big = size[roots] >= NP_MIN
    assert np.array_equal(in_h, big)
    assert np.array_equal(hid > 0, big)
    nh = int(hid.max())
    assert nh == int(np.count_nonzero(size >= NP_MIN)) > 100_000
    # one halo number per component
    r_of_h = np.zeros(nh + 1, dtype=np.int64)
    r_of_h[hid[big]] = roots[big]
    assert np.array_equal(r_of_h[hid[big]], roots[big])
    _say(f"consistent: {nh} halos, {int(big.sum())} particles in halos")

    # cell grid of side 1 for the neighbour searches
    ng = int(L)
    ijk = np.floor(xp).astype(np.int64) % ng
    cid = ijk[0] + ng * (ijk[1] + ng * ijk[2])
    order = np.argsort(cid, kind="stable")
    ptr = np.zeros(ng ** 3 + 1, dtype=np.int64)
    np.cumsum(np.bincount(cid, minlength=ng ** 3), out=ptr[1:])
    del ijk

    rng = np.random.default_rng(7)
    # half the sample from particles in halos, half from all particles
    sample = np.concatenate([rng.choice(np.flatnonzero(big), 1000, replace=False),
                             rng.choice(N, 1000, replace=False)])
    lo2, hi2 = (B * (1.0 - 1e-3)) ** 2, (B * (1.0 + 1e-3)) ** 2
    offs = np.array([(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    checked = 0
    for p in sample:
        c = np.floor(xp[:, p]).astype(np.int64) % ng
        cells = (c[None, :] + offs) % ng
        ids = cells[:, 0] + ng * (cells[:, 1] + ng * cells[:, 2])
        cand = np.concatenate([order[ptr[k]:ptr[k + 1]] for k in ids])
        cand = cand[cand != p]
        d = np.abs(xp[:, cand] - xp[:, p:p + 1])
        d = np.minimum(d, L - d)
        r2 = (d * d).sum(axis=0)
        close = cand[r2 < lo2]
        assert np.all(roots[close] == roots[p]), f"particle {p}: a partner within b is in another component"
        if size[roots[p]] >= 2:
            assert np.any(r2 <= hi2), f"particle {p}: isolated member of a component of {size[roots[p]]}"
        checked += len(close)
    _say(f"closure: {len(sample)} particles, {checked} partners within b")
