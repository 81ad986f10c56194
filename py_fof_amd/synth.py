"""Synthetic particle sets for the benchmark workloads.

All generators are seeded (numpy ``default_rng``) and return positions as a
(3, N) float64 array in Fortran order, the layout the reference's native entry
point takes (py_fof/__init__.py:210).

Workloads:
  * ``clumpy_plummer`` -- 262,144 particles in a 64^3 box: 60 % in 400 Plummer
    clumps (scale radius 0.3, Pareto(alpha=1, min 20) member counts), 40 %
    uniform background.
  * ``clumpy_nfw`` -- 134,217,728 particles in a 512^3 box: 50 % in 250,000
    NFW clumps (c = 8, Pareto(alpha=0.9) counts clipped to [20, 2e6],
    r_vir = 0.6 (n/100)^(1/3)), 50 % uniform background. Any size can be asked
    for; the clump count scales with it.
"""
import numpy as np

__all__ = [
    "pareto_counts",
    "plummer_radii",
    "nfw_radii",
    "isotropic",
    "clumpy_plummer",
    "clumpy_nfw",
    "halo_sizes",
    "isolated_halos",
    "physical_halos",
]


def _largest_remainder(weights, total, nmin):
    """Integer counts >= nmin summing to ``total``, proportional to weights."""
    w = np.asarray(weights, dtype=np.float64)
    k = len(w)
    if total < nmin * k:
        raise ValueError("total too small for the per-clump minimum")
    extra = total - nmin * k
    share = w / w.sum() * extra
    base = np.floor(share).astype(np.int64)
    rem = extra - int(base.sum())
    if rem > 0:
        order = np.argsort(-(share - base), kind="stable")
        base[order[:rem]] += 1
    return base + nmin


def pareto_counts(rng, k, alpha, total, nmin=20, nmax=None):
    """``k`` member counts from Pareto(alpha) with minimum ``nmin``, rescaled
    to sum to ``total`` and, when ``nmax`` is given, clipped to [nmin, nmax]."""
    raw = nmin * (1.0 + rng.pareto(alpha, k))
    if nmax is None:
        return _largest_remainder(raw - nmin + 1e-12, total, nmin)
    w = np.clip(raw, nmin, nmax)
    for _ in range(64):
        w = np.clip(w * (total / w.sum()), nmin, nmax)
        if abs(w.sum() - total) < 0.5:
            break
    c = np.floor(w).astype(np.int64)
    c = np.clip(c, nmin, nmax)
    diff = int(total - c.sum())
    if diff != 0:
        # put the leftover on the clumps with the most headroom
        room = (nmax - c) if diff > 0 else (c - nmin)
        order = np.argsort(-room, kind="stable")
        step = 1 if diff > 0 else -1
        i = 0
        while diff != 0:
            j = order[i % k]
            if nmin <= c[j] + step <= nmax:
                c[j] += step
                diff -= step
            i += 1
    return c


def isotropic(rng, n, dtype=np.float64):
    """Unit vectors, (3, n)."""
    mu = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - mu * mu)
    return np.stack([s * np.cos(phi), s * np.sin(phi), mu]).astype(dtype, copy=False)


def plummer_radii(rng, n, a):
    u = rng.uniform(1e-12, 1.0 - 1e-12, n)
    return a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)


_NFW_TAB = {}


def nfw_radii(rng, n, c, r_vir):
    """Radii inside r_vir of NFW profiles with concentration ``c`` (scalar or
    per-particle array) by inverting m(x) = ln(1+x) - x/(1+x) on [0, c]."""
    c = np.broadcast_to(np.asarray(c, dtype=np.float64), (n,))
    u = rng.uniform(0.0, 1.0, n)
    cu = np.unique(c)
    x = np.empty(n)
    for cv in cu:
        tab = _NFW_TAB.get(float(cv))
        if tab is None:
            xs = np.concatenate([[0.0], np.geomspace(1e-6, cv, 8192)])
            ms = np.log1p(xs) - xs / (1.0 + xs)
            tab = (xs, ms / ms[-1])
            _NFW_TAB[float(cv)] = tab
        sel = c == cv if len(cu) > 1 else slice(None)
        x[sel] = np.interp(u[sel], tab[1], tab[0])
    return x / c * r_vir


def _wrap(x, L):
    x = np.mod(x, L)
    x[x >= L] -= L
    return x


def clumpy_plummer(n=262144, L=64.0, n_clumps=400, frac=0.6, a=0.3, seed=0):
    """Plummer-clump + uniform background set (the 262,144-particle workload)."""
    rng = np.random.default_rng(seed)
    n_cl = int(round(frac * n))
    counts = pareto_counts(rng, n_clumps, 1.0, n_cl, nmin=20)
    centres = rng.uniform(0.0, L, (3, n_clumps))
    owner = np.repeat(np.arange(n_clumps), counts)
    r = plummer_radii(rng, n_cl, a)
    pos_cl = centres[:, owner] + r * isotropic(rng, n_cl)
    pos_bg = rng.uniform(0.0, L, (3, n - n_cl))
    xp = np.concatenate([pos_cl, pos_bg], axis=1)
    return np.asfortranarray(_wrap(xp, L))


def clumpy_nfw(n=134217728, L=512.0, n_clumps=None, frac=0.5, c=8.0, seed=2,
               nmax=2_000_000, alpha=0.9):
    """NFW-clump + uniform background set (the 134,217,728-particle workload
    at the defaults). ``n_clumps`` defaults to 250,000 * n / 2^27."""
    rng = np.random.default_rng(seed)
    if n_clumps is None:
        n_clumps = max(1, int(round(250_000 * n / 134_217_728)))
    n_cl = int(round(frac * n))
    nmax = min(nmax, max(20, n_cl))
    counts = pareto_counts(rng, n_clumps, alpha, n_cl, nmin=20, nmax=nmax)
    centres = rng.uniform(0.0, L, (3, n_clumps))
    r_vir = 0.6 * (counts / 100.0) ** (1.0 / 3.0)
    xp = np.empty((3, n), dtype=np.float64, order="F")
    # clump members, chunked to bound temporaries
    start = 0
    chunk = 1 << 22
    cum = np.concatenate([[0], np.cumsum(counts)])
    ci = 0
    while start < n_cl:
        stop = min(n_cl, start + chunk)
        # clumps overlapping [start, stop)
        c0 = np.searchsorted(cum, start, side="right") - 1
        c1 = np.searchsorted(cum, stop, side="left")
        lo = np.maximum(cum[c0:c1], start)
        hi = np.minimum(cum[c0 + 1:c1 + 1], stop)
        owner = np.repeat(np.arange(c0, c1), hi - lo)
        m = stop - start
        r = nfw_radii(rng, m, c, r_vir[owner])
        xp[:, start:stop] = centres[:, owner] + r * isotropic(rng, m)
        start = stop
        ci += 1
    xp[:, n_cl:] = rng.uniform(0.0, L, (3, n - n_cl))
    xp[:, :n_cl] = _wrap(xp[:, :n_cl], L)
    return xp


def halo_sizes(rng, k, alpha=1.9, nmin=20, nmax=1_000_000):
    """k sizes from dN/dn ~ n^-alpha on [nmin, nmax] (inverse-CDF sampling)."""
    u = rng.uniform(0.0, 1.0, k)
    e = 1.0 - alpha
    n = (nmin ** e + u * (nmax ** e - nmin ** e)) ** (1.0 / e)
    return np.maximum(nmin, np.floor(n).astype(np.int64))


def isolated_halos(n_halos=20_000, seed=3, alpha=1.9, nmin=20, nmax=1_000_000, L=512.0,
                   interlopers=0.1, a_val=1.0, sizes=None):
    """Isolated NFW halos for unbinding, one input region per halo.

    Each halo: NFW profile, c ~ U[4, 15], r_vir = 0.6 (n/100)^(1/3) cube
    units, isotropic Gaussian velocities with sigma^2 = G M / (2 r_vir) and
    G = a / (6 pi) (the cube-unit constant of the potential), unit masses.
    A fraction `interlopers` of the members are replaced by particles uniform
    in a sphere of radius 2 r_vir moving with dispersion 3 sigma.
    Returns (xp_list, vp_list) of (3, n) float64 Fortran-order arrays.
    """
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = halo_sizes(rng, n_halos, alpha, nmin, nmax)
    xs, vs = [], []
    for n in sizes:
        n = int(n)
        c = rng.uniform(4.0, 15.0)
        rvir = 0.6 * (n / 100.0) ** (1.0 / 3.0)
        sigma = np.sqrt(a_val / (6.0 * np.pi) * n / (2.0 * rvir))
        centre = rng.uniform(0.25 * L, 0.75 * L, 3)
        n_int = int(round(interlopers * n))
        n_h = n - n_int
        x = np.empty((3, n))
        v = np.empty((3, n))
        x[:, :n_h] = nfw_radii(rng, n_h, c, rvir) * isotropic(rng, n_h)
        v[:, :n_h] = sigma * rng.standard_normal((3, n_h))
        r_i = 2.0 * rvir * rng.uniform(0.0, 1.0, n_int) ** (1.0 / 3.0)
        x[:, n_h:] = r_i * isotropic(rng, n_int)
        v[:, n_h:] = 3.0 * sigma * rng.standard_normal((3, n_int))
        perm = rng.permutation(n)
        xs.append(np.asfortranarray(x[:, perm] + centre[:, None]))
        vs.append(np.asfortranarray(v[:, perm]))
    return xs, vs


def physical_halos(n_halos=20_000, seed=4, sizes=None, box=300.0, n_side=1024, z=0.5,
                   omega_m=0.3, log_mass=(2e9, 0.3)):
    """The isolated-halo set in physical units (Mpc/h, km/s, Msun/h).

    Positions of ``isolated_halos`` on a grid of ``n_side`` cells over ``box``
    (times dx = box / n_side), velocities from the virial sigma with the
    cube-unit G at a = 1 / (1 + z) converted to km/s, per-particle masses
    log-normal (``log_mass[1]`` dex) around ``log_mass[0]``. Returns
    (xp_list, vp_list, mp, kwargs) where kwargs are the fof() arguments of
    the set: b 0.2 cells, np_halo_min 10, bound_check, n_sim n_side^3, the
    Hubble flow, softening 0.01 Mpc/h.
    """
    a = 1.0 / (1.0 + z)
    ng = float(n_side)
    xs, vs = isolated_halos(n_halos=n_halos, seed=seed, L=ng, a_val=a, sizes=sizes)
    dx = box / ng
    v_scale = (a * ng) / (150.0 * np.sqrt(omega_m) * box)
    counts = [x.shape[1] for x in xs]
    rng = np.random.default_rng(seed + 100)
    mp = log_mass[0] * 10.0 ** rng.normal(0.0, log_mass[1], int(sum(counts)))
    xs = [np.asfortranarray(x * dx) for x in xs]
    vs = [np.asfortranarray(v / v_scale) for v in vs]
    kw = dict(L_Box=box, linking_length=0.2 * dx, np_halo_min=10, bound_check=True,
              units="physical", box=box, n_sim=int(n_side) ** 3, z=z, omega_m=omega_m,
              include_hubble_flow=True, softening=0.01)
    return xs, vs, mp, kw
