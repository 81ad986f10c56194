"""TEST INFRASTRUCTURE ONLY -- ctypes front end of the C oracle (liboracle.so).

Mirrors the argument conventions of the reference's native entry point
(fof_fortran.f90:10-58): positions/velocities (3, N) in Fortran order, f64.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class OracleGeom(ctypes.Structure):
    _fields_ = [
        ("origin", ctypes.c_double * 3),
        ("cell", ctypes.c_double * 3),
        ("bmin", ctypes.c_double * 3),
        ("bmax", ctypes.c_double * 3),
        ("n", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32 * 3),
        ("is_periodic", ctypes.c_int32),
    ]


def build_oracle():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load_oracle():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build_oracle()
    lib = ctypes.CDLL(_LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.oracle_fof.restype = ctypes.c_int32
    lib.oracle_fof.argtypes = [dp, dp, dp, ip, ctypes.c_int32, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_int32, ip, ip, dp, ip]
    lib.oracle_fof_candidates.restype = ctypes.c_int32
    lib.oracle_fof_candidates.argtypes = [dp, ip, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_int32, ip, ip, ip, ip]
    lib.oracle_fof_bound.restype = ctypes.c_int32
    lib.oracle_fof_bound.argtypes = [dp, dp, dp, ip, ctypes.c_int32, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int64, ip, ip, dp, ctypes.c_int32, ip, ip, ip, dp,
                                     ctypes.POINTER(ctypes.c_float), ctypes.c_int64]
    lib.oracle_green.restype = None
    lib.oracle_green.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_float),
                                 ctypes.POINTER(ctypes.c_float)]
    lib.oracle_select.restype = ctypes.c_int32
    lib.oracle_select.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                                  ctypes.c_double, ip, ip, dp]
    lib.oracle_depth.restype = ctypes.c_double
    lib.oracle_depth.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.oracle_geometry_of.restype = ctypes.c_int32
    lib.oracle_geometry_of.argtypes = [dp, ctypes.c_int32, ctypes.c_double,
                                       ctypes.c_double, ctypes.POINTER(OracleGeom)]
    _lib = lib
    return lib


def oracle_available():
    try:
        load_oracle()
        return True
    except Exception:
        return False


def _f64_fortran(a):
    return np.asfortranarray(a, dtype=np.float64)


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def oracle_fof(xp, np_halo, L_Box, b, np_halo_min=10, vp=None, mp=None,
               n_threads=1, with_labels=False):
    """Plain FoF (bound_check=False) exactly as the reference computes it.

    xp: (3, N) f64; np_halo: region sizes. Returns (in_halo int32[N],
    halo_id int32[N], ang_mom (3, nhalo) f64[, labels int32[N]]).
    """
    lib = load_oracle()
    xp = _f64_fortran(xp)
    n = xp.shape[1]
    np_halo = np.ascontiguousarray(np_halo, dtype=np.int32)
    nh = len(np_halo)
    in_halo = np.zeros(n, np.int32)
    halo_id = np.zeros(n, np.int32)
    ang = np.zeros((3, max(nh, 1)), np.float64, order="F")
    labels = np.zeros(max(n, 1), np.int32) if with_labels else None
    vpa = _f64_fortran(vp) if vp is not None else None
    mpa = np.ascontiguousarray(mp, dtype=np.float64) if mp is not None else None
    dp = ctypes.c_double
    ierr = lib.oracle_fof(
        _ptr(xp, dp), _ptr(vpa, dp) if vpa is not None else None,
        _ptr(mpa, dp) if mpa is not None else None, _ptr(np_halo, ctypes.c_int32), nh,
        float(L_Box), float(b), int(np_halo_min), int(n_threads), n,
        _ptr(in_halo, ctypes.c_int32), _ptr(halo_id, ctypes.c_int32), _ptr(ang, dp),
        _ptr(labels, ctypes.c_int32) if labels is not None else None)
    if ierr != 0:
        raise RuntimeError(f"oracle_fof ierr={ierr}")
    ang = ang[:, :nh]
    if with_labels:
        return in_halo, halo_id, ang, labels[:n]
    return in_halo, halo_id, ang


def oracle_candidates(xp, np_halo, L_Box, b, np_halo_min=10):
    """Candidates in the reference's order with members in CRS order
    (fof_fortran.f90:263-308). Returns (cand_np, cand_region, members)."""
    lib = load_oracle()
    xp = _f64_fortran(xp)
    n = xp.shape[1]
    np_halo = np.ascontiguousarray(np_halo, dtype=np.int32)
    nm = ctypes.c_int32(0)
    cap_c = n
    cand_np = np.zeros(max(cap_c, 1), np.int32)
    cand_reg = np.zeros(max(cap_c, 1), np.int32)
    members = np.zeros(max(n, 1), np.int32)
    nc = lib.oracle_fof_candidates(_ptr(xp, ctypes.c_double), _ptr(np_halo, ctypes.c_int32),
                                   len(np_halo), float(L_Box), float(b), int(np_halo_min), n,
                                   cap_c, n, _ptr(cand_np, ctypes.c_int32),
                                   _ptr(cand_reg, ctypes.c_int32), _ptr(members, ctypes.c_int32),
                                   ctypes.byref(nm))
    if nc < 0:
        raise RuntimeError("oracle_fof_candidates failed")
    return cand_np[:nc].copy(), cand_reg[:nc].copy(), members[:nm.value].copy()


def oracle_geometry(xp, L_Box, b):
    """Reference cell geometry of one region (fof_fortran.f90:518-570)."""
    lib = load_oracle()
    xp = _f64_fortran(xp)
    g = OracleGeom()
    lib.oracle_geometry_of(_ptr(xp, ctypes.c_double), xp.shape[1], float(L_Box), float(b),
                           ctypes.byref(g))
    return {
        "origin": list(g.origin), "cell": list(g.cell), "n": list(g.n),
        "periodic": [bool(v) for v in g.periodic], "is_periodic": bool(g.is_periodic),
        "bmin": list(g.bmin), "bmax": list(g.bmax),
    }


def oracle_fof_bound(xp, np_halo, L_Box, b, np_halo_min, vp, mp=None, app=0.1, a_val=1.0,
                     b_val=0.0, s15=17.152, ng=256, want_phi=False):
    """FoF + unbinding as the reference's bound_check=True branch
    (fof_fortran.f90:253-356). Returns a dict with in_halo, halo_id,
    ang_mom, ierr and, per FoF candidate, status / np_bound / method gid /
    spacing (and the first-pass potential of every candidate member, pid
    order, when want_phi)."""
    lib = load_oracle()
    xp = _f64_fortran(xp)
    vp = _f64_fortran(vp)
    n = xp.shape[1]
    np_halo = np.ascontiguousarray(np_halo, dtype=np.int32)
    nh = len(np_halo)
    in_halo = np.zeros(max(n, 1), np.int32)
    halo_id = np.zeros(max(n, 1), np.int32)
    ang = np.zeros((3, max(nh, 1)), np.float64, order="F")
    cap = max(n, 1)
    st = np.full(cap, -1, np.int32)
    nb = np.zeros(cap, np.int32)
    gid = np.full(cap, -1, np.int32)
    sp = np.zeros(cap, np.float64)
    phi = np.zeros(max(n, 1), np.float32) if want_phi else None
    mpa = np.ascontiguousarray(mp, dtype=np.float64) if mp is not None else None
    dp = ctypes.c_double
    ierr = lib.oracle_fof_bound(
        _ptr(xp, dp), _ptr(vp, dp), _ptr(mpa, dp) if mpa is not None else None,
        _ptr(np_halo, ctypes.c_int32), nh, float(L_Box), float(b), int(np_halo_min), float(app),
        float(a_val), float(b_val), float(s15), int(ng), _ptr(in_halo, ctypes.c_int32),
        _ptr(halo_id, ctypes.c_int32), _ptr(ang, dp), cap, _ptr(st, ctypes.c_int32),
        _ptr(nb, ctypes.c_int32), _ptr(gid, ctypes.c_int32), _ptr(sp, dp),
        _ptr(phi, ctypes.c_float) if phi is not None else None, len(phi) if phi is not None else 0)
    ncand = int(np.sum(st >= 0))
    out = {"ierr": int(ierr), "in_halo": in_halo[:n], "halo_id": halo_id[:n],
           "ang_mom": ang[:, :nh], "status": st[:ncand], "np_bound": nb[:ncand],
           "gid": gid[:ncand], "spacing": sp[:ncand]}
    if want_phi:
        out["phi"] = phi[:n]
    return out


def oracle_green(n):
    """Half-spectrum Green function (n, n, n/2+1) float32 [z, y, x] and the
    5^3 self kernel [z, y, x] of an n^3 isolated grid."""
    lib = load_oracle()
    nh = n // 2 + 1
    G = np.zeros(nh * n * n, np.float32)
    self_ = np.zeros(125, np.float32)
    lib.oracle_green(int(n), _ptr(G, ctypes.c_float), _ptr(self_, ctypes.c_float))
    return G.reshape(n, n, nh), self_.reshape(5, 5, 5)


def oracle_select(np_, span, s15=17.152, ng=256, app=0.1):
    lib = load_oracle()
    g = ctypes.c_int32(0)
    l = ctypes.c_int32(0)
    sp = ctypes.c_double(0.0)
    st = lib.oracle_select(float(app), float(s15), int(ng), int(np_), float(span),
                           ctypes.byref(g), ctypes.byref(l), ctypes.byref(sp))
    return st, g.value, l.value, sp.value


def oracle_tl_info(q, spacing, origin, app):
    """Two-level block counts and ratios (512 each) of members q (3, n) f32
    relative positions on a 256-grid placement; returns (n_active, counts,
    ratios); n_active is -1 when a stencil leaves an inner grid."""
    lib = load_oracle()
    q = np.ascontiguousarray(np.asarray(q, np.float32).T)
    org = np.ascontiguousarray(origin, dtype=np.float32)
    counts = np.zeros(512, np.int32)
    ratios = np.zeros(512, np.int32)
    lib.oracle_tl_info.restype = ctypes.c_int32
    na = lib.oracle_tl_info(int(q.shape[0]), _ptr(q, ctypes.c_float), ctypes.c_float(spacing),
                            _ptr(org, ctypes.c_float), ctypes.c_float(app),
                            _ptr(counts, ctypes.c_int32), _ptr(ratios, ctypes.c_int32))
    return int(na), counts, ratios


def oracle_depth(r, h):
    return load_oracle().oracle_depth(float(r), float(h))


def oracle_set_force_full(on):
    """Checker switch: every later pass recomputes its potential from scratch
    over the kept members instead of the reference's incremental updates."""
    lib = load_oracle()
    lib.oracle_set_force_full.restype = None
    lib.oracle_set_force_full.argtypes = [ctypes.c_int32]
    lib.oracle_set_force_full(1 if on else 0)


def _incremental_check(name, q, w, keep0, removed, extra, phi_dtype):
    lib = load_oracle()
    q = np.ascontiguousarray(np.asarray(q, np.float32).T)  # (n, 3)
    n = q.shape[0]
    w = np.ascontiguousarray(w, dtype=np.float32)
    k0 = np.ascontiguousarray(keep0, dtype=np.int8)
    rem = np.ascontiguousarray(removed, dtype=np.int32)
    ct = ctypes.c_float if phi_dtype == np.float32 else ctypes.c_double
    inc = np.zeros(max(n, 1), phi_dtype)
    full = np.zeros(max(n, 1), phi_dtype)
    fn = getattr(lib, name)
    fn.restype = ctypes.c_int32
    rc = fn(int(n), _ptr(q, ctypes.c_float), _ptr(w, ctypes.c_float),
            k0.ctypes.data_as(ctypes.c_char_p), int(len(rem)), _ptr(rem, ctypes.c_int32), *extra,
            _ptr(inc, ct), _ptr(full, ct))
    if rc != 0:
        raise ValueError(f"{name}: invalid removal list")
    return inc[:n], full[:n]


def oracle_c2_check(q, w, keep0, removed, sp, app, a=1.0):
    """C2 residual after one removal batch: (incremental, full recompute),
    f32 (bound_fortran.f90:3403-3552)."""
    extra = (ctypes.c_float(sp), ctypes.c_float(app), ctypes.c_double(a))
    return _incremental_check("oracle_c2_check", q, w, keep0, removed, extra, np.float32)


def oracle_tl_pp_check(q, w, keep0, removed, sp, app, ratios, a=1.0):
    """Two-level short-range sums after one removal batch: (pp_update,
    pp_full), f32 (bound_fortran.f90:1088-1498)."""
    pr = np.ascontiguousarray(ratios, dtype=np.int32)
    extra = (ctypes.c_float(sp), ctypes.c_float(app), ctypes.c_double(a), _ptr(pr, ctypes.c_int32))
    return _incremental_check("oracle_tl_pp_check", q, w, keep0, removed, extra, np.float32)


def oracle_direct_check(q, w, keep0, removed, app, a=1.0):
    """Direct-sum potential after one removal batch: (add-back, fresh sum),
    f64 (bound_fortran.f90:2613-2647)."""
    extra = (ctypes.c_double(app), ctypes.c_double(a))
    return _incremental_check("oracle_direct_check", q, w, keep0, removed, extra, np.float64)


def oracle_set_link_threads(k):
    """Checker switch: link every region with k threads (0 restores the
    reference's rule). The partition does not depend on it."""
    lib = load_oracle()
    lib.oracle_set_link_threads.restype = None
    lib.oracle_set_link_threads.argtypes = [ctypes.c_int32]
    lib.oracle_set_link_threads(int(k))
