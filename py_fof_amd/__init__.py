"""MI355X-native Friends-of-Friends halo finder with gravitational unbinding.

Drop-in for ``py_fof``: ``fof()`` and ``fof_with_info()`` keep the reference's
signatures, defaults, unit handling, return types and error messages
(py_fof/__init__.py:168-518). The work runs in hand-written HIP kernels for
gfx950 in ``libpyfof_hip.so``, reached through the C-ABI in include/pyfof.h.
There is no CPU fallback: without the library every call raises.
"""
import numpy as np

from . import _native
from ._native import MEM_DEVICE, MEM_HOST, Context, NativeLibraryError  # noqa: F401
from . import _device
from ._units import unit_scalars

__all__ = ["fof", "fof_with_info", "Context", "NativeLibraryError"]
__version__ = "0.1.0"


def _pack(arrays):
    """(3, N) f64 Fortran-order concatenation (py_fof/__init__.py:210); a single
    array that already has that layout is used as is, without a copy."""
    if len(arrays) == 1:
        a = arrays[0]
        if a.dtype == np.float64 and a.flags.f_contiguous:
            return a
        return np.asfortranarray(a, dtype=np.float64)
    return np.asfortranarray(np.hstack(arrays), dtype=np.float64)


def _error_for(ierr):
    # py_fof/__init__.py:338-348
    if ierr == 1:
        return ValueError("Invalid input parameters or configuration in FOF/bound_test")
    if ierr == 2:
        return ValueError("Candidate halo span exceeds 15 Mpc limit (s_15)")
    if ierr == 4:
        return ValueError("Non-finite values (NaN/Inf) encountered during calculation")
    if ierr == 10:
        return ValueError("n_threads exceeds maximum supported slots (256)")
    return RuntimeError(f"FOF Fortran computation failed with error code ierr={ierr}")


class _Prepared:
    """Native arguments of one call plus what the result needs afterwards.
    `device` is None for host arrays, else the torch device of the inputs."""
    __slots__ = ("empty", "xp", "vp", "mp", "np_halo", "nhalo", "total_np", "L", "link",
                 "np_min", "bound_check", "verbose", "softening", "a", "b", "s15", "n_threads",
                 "ng", "ng_global", "need_am", "ang_mom_scale", "device")


def _prepare(xp_list, vp_list=None, L_Box=512.0, linking_length=0.2,
             np_halo_min=10, bound_check=False, verbose=0, return_ang_mom=False,
             n_threads=None, mp_list=None, softening=None, z=0.0,
             box=300.0, h0=0.67, omega_m=0.3, omega_l=0.7, n_sim=None,
             units='cube', include_hubble_flow=False, ng=None, ng_global=None):
    """Validation and unit conversion of py_fof/__init__.py:168-326, in the
    same order and with the same messages. Returns a _Prepared."""
    if units not in ('cube', 'physical'):
        raise ValueError(f"units must be 'cube' or 'physical', got '{units}'")

    if n_threads is None or n_threads <= 0:
        n_threads_val = 0
    else:
        if n_threads > 256:
            raise ValueError("n_threads cannot exceed 256")
        n_threads_val = int(n_threads)

    if _device.is_device_input(xp_list):
        P = _Prepared()
        U = _device.prepare(P, xp_list, vp_list, mp_list, units, bound_check, softening, L_Box,
                            linking_length, z, box, h0, omega_m, omega_l, n_sim, ng_global,
                            include_hubble_flow)
        if U is not None:
            _scalars_into(P, U, np_halo_min, bound_check, verbose, n_threads_val, ng, ng_global)
            P.need_am = bool(return_ang_mom) and P.vp is not None
        return P

    if isinstance(xp_list, np.ndarray) and xp_list.ndim == 2:
        xp_list_norm = [xp_list]
    elif isinstance(xp_list, (list, tuple)):
        xp_list_norm = list(xp_list)
    else:
        raise ValueError("xp_list must be a 2D numpy array (3, N) or a list of 2D arrays")

    nhalo = len(xp_list_norm)
    for i, xp in enumerate(xp_list_norm):
        if not isinstance(xp, np.ndarray) or xp.ndim != 2 or xp.shape[0] != 3:
            raise ValueError(f"xp_list[{i}] must have shape (3, N), got {getattr(xp, 'shape', type(xp))}")

    np_halo = np.array([xp.shape[1] for xp in xp_list_norm], dtype=np.int32)
    total_np = int(np.sum(np_halo))

    P = _Prepared()
    P.device = None
    P.nhalo = nhalo
    P.total_np = total_np
    P.empty = total_np == 0
    if P.empty:
        return P

    xp_halo = _pack(xp_list_norm)

    if bound_check and vp_list is None:
        raise ValueError("vp_list is required when bound_check=True")

    vp_halo = None
    if vp_list is not None:
        if isinstance(vp_list, np.ndarray) and vp_list.ndim == 2:
            vp_list_norm = [vp_list]
        elif isinstance(vp_list, (list, tuple)):
            vp_list_norm = list(vp_list)
        else:
            raise ValueError("vp_list must be a 2D numpy array (3, N) or a list of 2D arrays")
        if len(vp_list_norm) != nhalo:
            raise ValueError(f"vp_list length ({len(vp_list_norm)}) must match xp_list length ({nhalo})")
        for i, vp in enumerate(vp_list_norm):
            if not isinstance(vp, np.ndarray) or vp.ndim != 2 or vp.shape != xp_list_norm[i].shape:
                raise ValueError(f"vp_list[{i}] shape {getattr(vp, 'shape', None)} must match xp_list[{i}] shape {xp_list_norm[i].shape}")
        vp_halo = _pack(vp_list_norm)
    # a missing vp_list means zero velocities (py_fof/__init__.py:233); the
    # library treats a null vp that way, so no zero array is materialised.

    mp_raw = None
    if mp_list is not None:
        if np.isscalar(mp_list) or (isinstance(mp_list, np.ndarray) and mp_list.ndim == 0):
            val = float(mp_list)
            if val <= 0 or not np.isfinite(val):
                raise ValueError("mp_list scalar must be positive and finite")
            mp_raw = np.full(total_np, val, dtype=np.float64)
        elif isinstance(mp_list, np.ndarray):
            if mp_list.ndim == 1 and len(mp_list) == total_np:
                mp_raw = np.ascontiguousarray(mp_list, dtype=np.float64)
            elif nhalo > 1 and mp_list.ndim == 1 and len(mp_list) == nhalo:
                mp_raw = np.repeat(np.ascontiguousarray(mp_list, dtype=np.float64), np_halo)
            else:
                raise ValueError(f"mp_list 1D array length ({len(mp_list)}) does not match total particles ({total_np})")
        elif isinstance(mp_list, (list, tuple)):
            if len(mp_list) == nhalo and all(isinstance(m, np.ndarray) and m.ndim == 1 and len(m) == np_halo[i] for i, m in enumerate(mp_list)):
                mp_raw = np.concatenate(mp_list).astype(np.float64)
            elif len(mp_list) == nhalo and all(np.isscalar(m) for m in mp_list):
                scalars = np.array([float(m) for m in mp_list], dtype=np.float64)
                mp_raw = np.repeat(scalars, np_halo)
            elif len(mp_list) == total_np and all(np.isscalar(m) for m in mp_list):
                mp_raw = np.array(mp_list, dtype=np.float64)
            else:
                raise ValueError("mp_list structure does not match xp_list structure or particle counts")
        else:
            raise ValueError("Unsupported mp_list type")

        if not np.all(np.isfinite(mp_raw)) or not np.all(mp_raw > 0):
            raise ValueError("All particle masses in mp_list must be positive and finite")

    U = unit_scalars(units, bound_check, softening, L_Box, linking_length, z, box, h0, omega_m,
                     omega_l, n_sim, ng_global, include_hubble_flow, mp_raw is not None, total_np)
    if not U.physical:
        xp_code, vp_code = xp_halo, vp_halo
        mp_code = mp_raw  # None: unit masses inside the library
    else:
        xp_code = np.asfortranarray(xp_halo / U.dx, dtype=np.float64)
        vp_code = np.asfortranarray(vp_halo * U.v_scale, dtype=np.float64) if vp_halo is not None else None
        if mp_raw is not None:
            mp_code = np.ascontiguousarray(mp_raw / U.m_grid, dtype=np.float64)
        elif U.mp_fill is not None:
            mp_code = np.full(total_np, U.mp_fill, dtype=np.float64)
        else:
            mp_code = None

    P.xp, P.vp, P.mp, P.np_halo = xp_code, vp_code, mp_code, np_halo
    _scalars_into(P, U, np_halo_min, bound_check, verbose, n_threads_val, ng, ng_global)
    # the moment is only observable with return_ang_mom; with zero velocities it is 0
    P.need_am = bool(return_ang_mom) and vp_code is not None
    return P


def _scalars_into(P, U, np_halo_min, bound_check, verbose, n_threads_val, ng, ng_global):
    P.L, P.link, P.np_min = U.L, U.link, int(np_halo_min)
    P.bound_check, P.verbose = int(bool(bound_check)), int(verbose)
    P.softening, P.a, P.b, P.s15 = U.softening, U.a, U.b, U.s15
    P.n_threads = n_threads_val
    P.ng = int(ng) if ng is not None else 256
    P.ng_global = float(ng_global) if ng_global is not None else 512.0
    P.ang_mom_scale = U.ang_mom_scale


def _outputs(P):
    if P.device is not None:
        return _device.outputs(P)
    return (np.zeros(P.total_np, dtype=np.int32), np.zeros(P.total_np, dtype=np.int32),
            np.zeros((3, P.nhalo), dtype=np.float64, order='F'))


def _finish(P, ierr, in_halo_out, halo_id_out, ang_mom_out, message=""):
    """ierr -> exception (py_fof/__init__.py:338-348), then the scaled outputs."""
    if ierr != 0:
        if P.verbose:
            print(f"[py_fof_amd] ierr={ierr}: {message}")
        raise _error_for(ierr)
    if P.device is not None:
        return in_halo_out.bool(), halo_id_out, ang_mom_out * P.ang_mom_scale
    return in_halo_out.astype(bool), halo_id_out, ang_mom_out * P.ang_mom_scale


def _fof_core(xp_list, vp_list=None, L_Box=512.0, linking_length=0.2,
              np_halo_min=10, bound_check=False, verbose=0, return_ang_mom=False,
              n_threads=None, mp_list=None, softening=None, z=0.0,
              box=300.0, h0=0.67, omega_m=0.3, omega_l=0.7, n_sim=None,
              units='cube', include_hubble_flow=False,
              ng=None, ng_global=None, _ctx=None, _want_catalog=False, _want_roots=False):
    """py_fof/__init__.py:168-351: validation, units, the native call, errors.

    Returns (in_halo bool[N], halo_id int32[N], ang_mom (3, nhalo) f64) and,
    when asked, the device catalog and the partition roots.
    """
    P = _prepare(xp_list, vp_list=vp_list, L_Box=L_Box, linking_length=linking_length,
                 np_halo_min=np_halo_min, bound_check=bound_check, verbose=verbose,
                 return_ang_mom=return_ang_mom, n_threads=n_threads, mp_list=mp_list,
                 softening=softening, z=z, box=box, h0=h0, omega_m=omega_m, omega_l=omega_l,
                 n_sim=n_sim, units=units, include_hubble_flow=include_hubble_flow, ng=ng,
                 ng_global=ng_global)
    extras = {}
    if P.empty and P.device is not None:
        out = _device.empty_result(P, _want_catalog, _want_roots)
        return out if (_want_catalog or _want_roots) else out[:3]
    if P.empty:
        in_halo = np.zeros(0, dtype=bool)
        halo_id = np.zeros(0, dtype=np.int32)
        ang_mom = np.zeros((3, P.nhalo), dtype=np.float64)
        if _want_catalog:
            extras["catalog"] = (np.zeros(1, np.int64), np.zeros(0, np.int32))
        if _want_roots:
            extras["roots"] = np.zeros(0, np.int32)
        return (in_halo, halo_id, ang_mom, extras) if (_want_catalog or _want_roots) else (in_halo, halo_id, ang_mom)

    in_halo_out, halo_id_out, ang_mom_out = _outputs(P)
    if P.device is None:
        roots = np.zeros(P.total_np, dtype=np.int32) if _want_roots else None
        ctx = _ctx if _ctx is not None else _native.default_context()
        mem = MEM_HOST
    else:
        roots = _device.like_int32(P) if _want_roots else None
        ctx = _ctx if _ctx is not None else _native.default_context(P.device.index or 0)
        if (P.device.index or 0) != ctx.device:
            raise ValueError(f"inputs are on {P.device}, the context on device {ctx.device}")
        mem = MEM_DEVICE
    ctx.set_option(_native.OPT_CATALOG, 1 if _want_catalog else 0)
    ierr = ctx.fof(P.xp, P.vp, P.mp, P.np_halo, P.L, P.link, P.np_min, P.bound_check, P.verbose,
                   P.softening, P.a, P.b, P.s15, P.n_threads, P.total_np, P.ng, P.ng_global,
                   in_halo_out, halo_id_out, ang_mom_out if P.need_am else None, roots, mem)
    out = _finish(P, ierr, in_halo_out, halo_id_out, ang_mom_out,
                  _native.last_error() if ierr else "")
    if _want_catalog or _want_roots:
        if _want_catalog:
            extras["catalog"] = ctx.catalog() if P.device is None else _device.catalog(ctx, P.device)
        if _want_roots:
            extras["roots"] = roots
        return out + (extras,)
    return out


def fof(xp_list, vp_list=None, L_Box=512.0, linking_length=0.2,
        np_halo_min=10, bound_check=False, verbose=0, return_ang_mom=False,
        n_threads=None, mp_list=None, softening=None, z=0.0,
        box=300.0, h0=0.67, omega_m=0.3, omega_l=0.7, n_sim=None,
        units='cube', include_hubble_flow=False,
        ng=None, ng_global=None):
    """Friends-of-Friends halo finder (py_fof/__init__.py:354-447).

    Returns ``in_halo`` (bool, shape (total_np,)) and, with
    ``return_ang_mom=True``, also the (3, nhalo) per-region angular momentum.
    xp_list (and vp_list, mp_list) may also be torch tensors on the GPU; the
    results are then device tensors and nothing goes through host memory.
    """
    in_halo, _, ang_mom = _fof_core(
        xp_list=xp_list, vp_list=vp_list, L_Box=L_Box, linking_length=linking_length,
        np_halo_min=np_halo_min, bound_check=bound_check, verbose=verbose,
        return_ang_mom=return_ang_mom, n_threads=n_threads, mp_list=mp_list,
        softening=softening, z=z, box=box, h0=h0, omega_m=omega_m, omega_l=omega_l,
        n_sim=n_sim, units=units, include_hubble_flow=include_hubble_flow,
        ng=ng, ng_global=ng_global)
    if return_ang_mom:
        return in_halo, ang_mom
    return in_halo


def fof_with_info(xp_list, vp_list=None, L_Box=512.0, linking_length=0.2,
                  np_halo_min=10, bound_check=False, verbose=0, return_ang_mom=False,
                  n_threads=None, mp_list=None, softening=None, z=0.0,
                  box=300.0, h0=0.67, omega_m=0.3, omega_l=0.7, n_sim=None,
                  units='cube', include_hubble_flow=False,
                  ng=None, ng_global=None):
    """FoF with the halo catalog (py_fof/__init__.py:450-518).

    The catalog (members grouped by halo id, ascending indices) is built on
    the device by a stable radix sort instead of a host argsort. With device
    tensors as input, in_halo, ang_mom and each halo's indices are device
    tensors; counts and bound_status are host arrays, as in the reference.
    """
    in_halo, halo_id, ang_mom, extras = _fof_core(
        xp_list=xp_list, vp_list=vp_list, L_Box=L_Box, linking_length=linking_length,
        np_halo_min=np_halo_min, bound_check=bound_check, verbose=verbose,
        return_ang_mom=return_ang_mom, n_threads=n_threads, mp_list=mp_list,
        softening=softening, z=z, box=box, h0=h0, omega_m=omega_m, omega_l=omega_l,
        n_sim=n_sim, units=units, include_hubble_flow=include_hubble_flow,
        ng=ng, ng_global=ng_global, _want_catalog=True)

    off, idx = extras["catalog"]
    nh = len(off) - 1
    if nh <= 0:
        nhalo_found = 0
        np_per_halo = np.array([], dtype=int)
        halo_indices = []
    elif not isinstance(off, np.ndarray):
        nhalo_found, np_per_halo, halo_indices = _device.split_catalog(off, idx)
    else:
        counts = np.diff(off).astype(np.int64)
        idx64 = idx.astype(np.int64)
        halo_indices = [idx64[off[h]:off[h + 1]] for h in range(nh) if counts[h] > 0]
        np_per_halo = counts[counts > 0]
        nhalo_found = len(halo_indices)

    res = {
        'in_halo': in_halo,
        'nhalo_found': nhalo_found,
        'np_per_halo': np_per_halo,
        'halo_indices': halo_indices,
        'bound_status': np.zeros(nhalo_found, dtype=np.int32),
    }
    if return_ang_mom:
        res['ang_mom'] = ang_mom
    return res
