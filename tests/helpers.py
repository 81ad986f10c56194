"""Shared test inputs and independent checkers.

``chain_halo`` restates the constructive generator of the reference's tests
(test_fof.py:361-458): a chain in which each particle sits 0.5-0.95 b from
the previous one (so the chain must link), plus noise particles placed beyond
range + 5 b from the chain centre and more than b from each other, shuffled.
The expected membership is exact by construction.
"""
import numpy as np


def chain_halo(n_halo, b, n_noise, rng):
    xh = np.zeros((3, n_halo))
    for i in range(1, n_halo):
        r = rng.uniform(0.5 * b, 0.95 * b)
        th = rng.uniform(0, 2 * np.pi)
        ph = np.arccos(rng.uniform(-1, 1))
        xh[:, i] = xh[:, i - 1] + r * np.array(
            [np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)])
    lo, hi = xh.min(axis=1), xh.max(axis=1)
    centre = 0.5 * (lo + hi)
    dmin = (hi - lo).max() + 5 * b
    xn = np.zeros((3, n_noise))
    for i in range(n_noise):
        while True:
            th = rng.uniform(0, 2 * np.pi)
            ph = np.arccos(rng.uniform(-1, 1))
            r = rng.uniform(dmin, 2 * dmin)
            p = centre + r * np.array([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)])
            if i == 0 or np.all(np.sqrt(((xn[:, :i] - p[:, None]) ** 2).sum(0)) > b):
                xn[:, i] = p
                break
    xp = np.hstack([xh, xn])
    expect = np.zeros(n_halo + n_noise, bool)
    expect[:n_halo] = True
    perm = rng.permutation(n_halo + n_noise)
    return np.asfortranarray(xp[:, perm]), expect[perm]


def modulo(a, p):
    """gfortran real MODULO, elementwise (fmod then shift into sign of p)."""
    r = np.fmod(a, p)
    fix = (r != 0) & ((r < 0) != (p < 0))
    return np.where(fix, r + p, r)


def brute_force_labels(xp, L, b):
    """O(N^2) connected components of one region under the reference's
    predicate applied to EVERY pair (no cells), root = smallest input index.

    Valid as a checker when the reference's cells are at least b (+ f32
    rounding) wide, so that every pair within b sits in adjacent cells.
    Returns labels[i] = smallest input index in i's component.
    """
    n = xp.shape[1]
    Lh = 0.5 * L
    xr = modulo(xp - xp[:, :1] + Lh, L) - Lh
    rng_ = xr.max(axis=1) - xr.min(axis=1)
    periodic = bool(np.any(rng_ + 2 * b >= L))
    xc = xr.astype(np.float32)
    parent = np.arange(n)

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for i in range(n):
        d32 = xc[:, i + 1:] - xc[:, i:i + 1]  # float32 differences
        if periodic:
            d = np.abs(d32.astype(np.float64))
            d = np.where(d > Lh, L - d, d)
            r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
        else:
            s = d32[0] * d32[0]
            s = s + d32[1] * d32[1]
            s = s + d32[2] * d32[2]
            r2 = s.astype(np.float64)
        for j in np.nonzero(r2 <= b * b)[0] + i + 1:
            ri, rj = find(i), find(j)
            if ri != rj:
                parent[max(ri, rj)] = min(ri, rj)
    return np.array([find(i) for i in range(n)])


def labels_to_canonical(labels):
    """Map any labelling to 'smallest member index' labels."""
    labels = np.asarray(labels)
    n = len(labels)
    order = np.argsort(labels, kind="stable")
    sl = labels[order]
    starts = np.concatenate([[True], sl[1:] != sl[:-1]])
    first = order[starts]
    grp = np.cumsum(starts) - 1
    out = np.empty(n, dtype=np.int64)
    mins = np.minimum.reduceat(order, np.nonzero(starts)[0])
    out[order] = mins[grp]
    del first
    return out


def depth_kernel(r, h):
    """Softened depth D(r, h) of the direct sum (bound_fortran.f90:3080-3102),
    elementwise in f64; D(0, h) = 208 / (70 h). Used only to construct inputs."""
    r = np.asarray(r, np.float64)
    e = 2.0 * r / h
    with np.errstate(divide="ignore", invalid="ignore"):
        inner = (208 - e ** 2 * (112 + e ** 2 * (-56 + e * (14 + e * (8 - 3 * e))))) / (70 * h)
        outer = (128 - (e ** 2 * (448 + e * (-280 + e * (56 + e * (14 + e * (-8 + e)))))
                        - 224 * e - 12 / e)) / (70 * h)
        return np.where(r >= h, 1.0 / r, np.where(e <= 1.0, inner, outer))


def max_iter_halo(n_shells=24, n_core_pairs=60, step=0.1, app=0.1, seed=7):
    """A direct-summation halo built so that the removal loop deletes exactly
    one shell per potential refresh and so reaches the 20-pass cap
    (status 5, bound_fortran.f90:2857-2864).

    An inversion-symmetric core (the origin plus n_core_pairs +/- pairs inside
    radius 0.8 step) at rest, and shells k = 1..n_shells of six members at
    +/- k step on the three axes, each with a tangential speed (the six
    momenta cancel). Shell k's kinetic energy sits halfway between minus its
    potential with shell k + 1 present and without it, so it is bound
    (E = -delta/2) while shell k + 1 remains and unbound (E = +delta/2) once
    shell k + 1 is gone; the outermost shell starts unbound. Every pass
    therefore removes one shell and leaves the next one unbound for the next
    refresh: after 20 passes 20 shells are gone, n_shells - 20 remain with the
    core, and the candidate ends at MAX_ITER. Cube units, a = 1, masses 1,
    linking length 0.2 > step links every member. Returns (xp, vp, retained).
    """
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((3, n_core_pairs))
    u /= np.sqrt((u * u).sum(0))
    core_half = u * (0.8 * step * rng.uniform(0.2, 1.0, n_core_pairs) ** (1.0 / 3.0))
    core = np.hstack([np.zeros((3, 1)), core_half, -core_half])
    axes = np.eye(3)
    tang = np.roll(np.eye(3), 1, axis=0)  # x -> y, y -> z, z -> x
    shell_x, shell_dir = [], []
    for k in range(1, n_shells + 1):
        for ax in range(3):
            for sgn in (1.0, -1.0):
                shell_x.append(sgn * k * step * axes[:, ax])
                shell_dir.append(sgn * tang[:, ax])
    sx = np.array(shell_x).T
    sd = np.array(shell_dir).T
    x = np.hstack([core, sx])
    nc = core.shape[1]
    n = x.shape[1]
    G = 1.0 / (6.0 * np.pi)
    d = x[:, :, None] - x[:, None, :]
    D = depth_kernel(np.sqrt((d * d).sum(0)), app)
    np.fill_diagonal(D, 0.0)
    shell_of = np.concatenate([np.zeros(nc, int), np.repeat(np.arange(1, n_shells + 1), 6)])
    speed = np.zeros(n)
    for k in range(1, n_shells + 1):
        idx = np.nonzero(shell_of == k)[0]
        for i in idx:
            phi_with = -G * D[i, shell_of <= k + 1].sum()
            phi_without = -G * D[i, shell_of <= k].sum()
            if k == n_shells:
                ke = -phi_without * 1.5  # E = +phi/2 ... well above zero
            else:
                ke = -0.5 * (phi_with + phi_without)
            speed[i] = np.sqrt(2.0 * ke)
    v = np.zeros((3, n))
    v[:, nc:] = sd * speed[nc:]
    x = x + 256.0
    retained = nc + 6 * (n_shells - 20)
    return np.asfortranarray(x), np.asfortranarray(v), retained


def wrapped_clumps(L, per_clump, centres, n_bg, seed, sigma=0.3):
    """Gaussian clumps (sigma) at the given centres plus n_bg uniform
    particles, wrapped into [0, L) and shuffled; (3, N) f64 Fortran order.
    Clumps centred near a face straddle the periodic boundary."""
    rng = np.random.default_rng(seed)
    parts = [np.asarray(c, float)[:, None] + sigma * rng.standard_normal((3, per_clump))
             for c in centres]
    parts.append(rng.uniform(0, L, (3, n_bg)))
    x = np.mod(np.hstack(parts), L)
    return np.asfortranarray(x[:, rng.permutation(x.shape[1])])


# periodic regions with 1 or 2 reference cells per axis, where the -1 and +1
# neighbours of a cell are the same cell (fof_fortran.f90:729-755):
# (L, b, members per clump, clump centres, background, sigma, cells per axis)
TINY_GRID_CASES = [
    (10.0, 3.4, 50, [(0.1, 0.1, 0.1), (5.1, 0.1, 0.1), (0.1, 5.1, 5.1), (5.1, 5.1, 9.9)], 6, 0.3, 2),
    (10.0, 5.5, 60, [(0.1, 0.1, 9.95), (5.1, 5.1, 5.1)], 0, 0.3, 1),
    (10.0, 3.4, 600, [(0.1, 0.1, 0.1), (5.1, 0.1, 0.1), (0.1, 5.1, 5.1), (5.1, 5.1, 9.9)], 0, 0.15, 2),
    (10.0, 5.5, 900, [(0.1, 0.1, 9.95), (5.1, 5.1, 5.1)], 0, 0.15, 1),
]


def physical_to_code(xs, vs, mp, L_Box, linking_length, softening, box=300.0, n_sim=None,
                     z=0.0, h0=0.67, omega_m=0.3, omega_l=0.7, include_hubble_flow=False,
                     ng_global=None):
    """The reference's physical -> code unit conversion (py_fof/__init__.py:
    266-326) with its own expressions and operation order, for feeding the
    oracle the inputs the reference's native code would receive. Returns
    (xp, vp, mp, kw) with kw the oracle scalars L, b, app, a_val, b_val, s15."""
    a_val = 1.0 / (1.0 + float(z))
    if include_hubble_flow:
        t0 = 1.0 / (150.0 * float(h0) * np.sqrt(float(omega_m)))
        hubble = 100.0 * float(h0) * np.sqrt(float(omega_m) / (a_val ** 3) + float(omega_l))
        b_val = hubble * (a_val ** 2) * t0
    else:
        b_val = 0.0
    if n_sim is not None and float(n_sim) > 0:
        ng_eff = float(round(float(n_sim) ** (1.0 / 3.0)))
    else:
        ng_eff = float(ng_global) if ng_global is not None else 512.0
    box_mpc_h = float(L_Box)
    dx = box_mpc_h / ng_eff
    xp = np.asfortranarray(np.hstack(xs) / dx, dtype=np.float64)
    v_scale = (a_val * ng_eff) / (150.0 * np.sqrt(float(omega_m)) * box_mpc_h)
    vp = np.asfortranarray(np.hstack(vs) * v_scale, dtype=np.float64)
    m_grid = float(omega_m) * 2.77536627e11 * (box_mpc_h ** 3) / (ng_eff ** 3)
    mpc = np.ascontiguousarray(np.asarray(mp) / m_grid, dtype=np.float64)
    kw = dict(L=ng_eff, b=float(linking_length) / dx, app=float(softening) / dx, a_val=a_val,
              b_val=b_val, s15=15.0 * float(h0) * ng_eff / box_mpc_h)
    return xp, vp, mpc, kw
