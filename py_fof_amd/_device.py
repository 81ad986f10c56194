"""Device-tensor front end: fof() / fof_with_info() on torch tensors in HBM.

The reference takes numpy arrays and packs, converts units and fills defaults
on the host (py_fof/__init__.py:188-326). Here positions, velocities and
masses that already sit on the GPU stay there: packing is one device copy
(none for a single (3, N) tensor whose columns are contiguous, the AoS layout
the kernels read), unit conversion is a device multiply, a missing vp_list is
a null pointer (no 3.2 GB zero array at 134M particles) and unit masses are
implied (no 1 GB array of ones). Outputs are device tensors: in_halo (bool),
halo_id (int32), ang_mom (f64, (3, nhalo)) and the catalog.

Accepted inputs, with the reference's validation order and messages:
* xp_list: a CUDA tensor of shape (3, N) or a list/tuple of them, float64 or
  float32 (converted to float64 on the device);
* vp_list: the same structure, or None;
* mp_list: None, a positive scalar, a 1-D CUDA tensor of total_np masses, or a
  list of per-region scalars or 1-D CUDA tensors.
Mixed host/device inputs raise ValueError.
"""
import numpy as np

from ._units import unit_scalars


def is_device_input(xp_list):
    """True when xp_list is a CUDA tensor or a non-empty list of them."""
    def dev(a):
        return type(a).__module__.split(".")[0] == "torch" and bool(getattr(a, "is_cuda", False))
    if isinstance(xp_list, (list, tuple)):
        return len(xp_list) > 0 and any(dev(a) for a in xp_list)
    return dev(xp_list)


def _torch():
    import torch
    return torch


def _norm_list(x, name):
    torch = _torch()
    if (isinstance(x, torch.Tensor) or isinstance(x, np.ndarray)) and x.ndim == 2:
        return [x]
    if isinstance(x, (list, tuple)):
        return list(x)
    raise ValueError(f"{name} must be a 2D numpy array (3, N) or a list of 2D arrays")


def _is_aos(a):
    return a.stride(0) == 1 and (a.shape[1] == 1 or a.stride(1) == 3)


def _aos(parts, device):
    """(3, N) tensors -> one float64 (3, N) view whose columns are contiguous
    (an (N, 3) row-major buffer), the layout the library reads; a single
    float64 tensor already laid out so is used as is."""
    torch = _torch()
    if len(parts) == 1:
        a = parts[0]
        if a.dtype == torch.float64 and _is_aos(a):
            return a
        return a.to(device=device, dtype=torch.float64).t().contiguous().t()
    return torch.cat([p.to(device=device, dtype=torch.float64).t() for p in parts], 0).contiguous().t()


def prepare(P, xp_list, vp_list, mp_list, units, bound_check, softening, L_Box, linking_length,
            z, box, h0, omega_m, omega_l, n_sim, ng_global, include_hubble_flow):
    """Fills P (a _Prepared) from device tensors and returns the unit
    scalars; returns None for an empty input (P.empty set)."""
    torch = _torch()
    xs = _norm_list(xp_list, "xp_list")
    for i, xp in enumerate(xs):
        if isinstance(xp, np.ndarray) or (isinstance(xp, torch.Tensor) and not xp.is_cuda):
            raise ValueError(f"xp_list[{i}] is a host array among device tensors")
        if not isinstance(xp, torch.Tensor) or xp.ndim != 2 or xp.shape[0] != 3:
            raise ValueError(f"xp_list[{i}] must have shape (3, N), got {getattr(xp, 'shape', type(xp))}")
        if not (xp.is_floating_point() and xp.dtype in (torch.float64, torch.float32)):
            raise ValueError(f"xp_list[{i}] must be float64 or float32, got {xp.dtype}")
    device = xs[0].device
    if any(x.device != device for x in xs):
        raise ValueError("xp_list tensors must be on one device")
    np_halo = np.array([int(x.shape[1]) for x in xs], dtype=np.int32)
    total_np = int(np_halo.sum())
    P.nhalo, P.total_np, P.empty = len(xs), total_np, total_np == 0
    P.device = device
    if P.empty:
        return None
    xp_halo = _aos(xs, device)

    if bound_check and vp_list is None:
        raise ValueError("vp_list is required when bound_check=True")
    vp_halo = None
    if vp_list is not None:
        vs = _norm_list(vp_list, "vp_list")
        if len(vs) != len(xs):
            raise ValueError(f"vp_list length ({len(vs)}) must match xp_list length ({len(xs)})")
        for i, vp in enumerate(vs):
            if not isinstance(vp, (torch.Tensor, np.ndarray)) or vp.ndim != 2 or tuple(vp.shape) != tuple(xs[i].shape):
                raise ValueError(f"vp_list[{i}] shape {getattr(vp, 'shape', None)} must match xp_list[{i}] shape {tuple(xs[i].shape)}")
            if isinstance(vp, np.ndarray) or not vp.is_cuda or vp.device != device:
                raise ValueError(f"vp_list[{i}] must be on the device of xp_list")
        vp_halo = _aos(vs, device)

    mp_raw = None
    if isinstance(mp_list, np.ndarray) and mp_list.ndim >= 1:
        mp_list = torch.as_tensor(mp_list, device=device)
    elif isinstance(mp_list, (list, tuple)):
        mp_list = [torch.as_tensor(m, device=device) if isinstance(m, np.ndarray) and m.ndim >= 1 else m
                   for m in mp_list]
    if mp_list is not None:
        if np.isscalar(mp_list) or (isinstance(mp_list, torch.Tensor) and mp_list.ndim == 0) or \
                (isinstance(mp_list, np.ndarray) and mp_list.ndim == 0):
            val = float(mp_list)
            if val <= 0 or not np.isfinite(val):
                raise ValueError("mp_list scalar must be positive and finite")
            mp_raw = torch.full((total_np,), val, dtype=torch.float64, device=device)
        elif isinstance(mp_list, torch.Tensor):
            if mp_list.ndim == 1 and mp_list.shape[0] == total_np:
                mp_raw = mp_list.to(device=device, dtype=torch.float64).contiguous()
            elif len(xs) > 1 and mp_list.ndim == 1 and mp_list.shape[0] == len(xs):
                counts = torch.as_tensor(np_halo.astype(np.int64), device=device)
                mp_raw = torch.repeat_interleave(mp_list.to(device=device, dtype=torch.float64), counts)
            else:
                raise ValueError(f"mp_list 1D array length ({mp_list.shape[0] if mp_list.ndim else 0}) does not match total particles ({total_np})")
        elif isinstance(mp_list, (list, tuple)):
            if len(mp_list) == len(xs) and all(isinstance(m, torch.Tensor) and m.ndim == 1 and
                                              m.shape[0] == np_halo[i] for i, m in enumerate(mp_list)):
                mp_raw = torch.cat([m.to(device=device, dtype=torch.float64) for m in mp_list])
            elif len(mp_list) == len(xs) and all(np.isscalar(m) for m in mp_list):
                vals = torch.tensor([float(m) for m in mp_list], dtype=torch.float64, device=device)
                counts = torch.as_tensor(np_halo.astype(np.int64), device=device)
                mp_raw = torch.repeat_interleave(vals, counts)
            else:
                raise ValueError("mp_list structure does not match xp_list structure or particle counts")
        else:
            raise ValueError("Unsupported mp_list type")
        if not bool(torch.isfinite(mp_raw).all()) or not bool((mp_raw > 0).all()):
            raise ValueError("All particle masses in mp_list must be positive and finite")

    U = unit_scalars(units, bound_check, softening, L_Box, linking_length, z, box, h0, omega_m,
                     omega_l, n_sim, ng_global, include_hubble_flow, mp_raw is not None, total_np)
    if U.physical:
        # torch divides a device tensor by a Python scalar as a product with
        # its reciprocal; a 0-d device divisor keeps the correctly rounded
        # quotient numpy (and the reference) compute
        def div(t, s):
            return t / torch.full((), s, dtype=torch.float64, device=device)
        xp_halo = _aos([div(xp_halo, U.dx)], device)
        if vp_halo is not None:
            vp_halo = _aos([vp_halo * U.v_scale], device)
        if mp_raw is not None:
            mp_raw = div(mp_raw, U.m_grid)
        elif U.mp_fill is not None:
            mp_raw = torch.full((total_np,), U.mp_fill, dtype=torch.float64, device=device)
    P.xp, P.vp, P.mp, P.np_halo = xp_halo, vp_halo, mp_raw, np_halo
    # the library runs on its own stream: the inputs must be complete
    torch.cuda.current_stream(device).synchronize()
    return U


def outputs(P):
    """in_halo, halo_id (int32[N]) and ang_mom ((3, nhalo) f64 with the
    library's per-halo xyz layout) on the input's device."""
    torch = _torch()
    dev = P.device
    am = torch.zeros((P.nhalo, 3), dtype=torch.float64, device=dev).t()
    return (torch.zeros(P.total_np, dtype=torch.int32, device=dev),
            torch.zeros(P.total_np, dtype=torch.int32, device=dev), am)


def catalog(ctx, device):
    """(offsets int64[nh+1], indices int32[total]) of the context's last call,
    as tensors on `device` (the indices are copied device to device)."""
    torch = _torch()
    from . import _native
    nh, nm = ctx.catalog_size()
    off = torch.zeros(nh + 1, dtype=torch.int64, device=device)
    idx = torch.zeros(max(nm, 1), dtype=torch.int32, device=device)
    ctx.catalog_copy(off, idx, _native.MEM_DEVICE)
    return off, idx[:nm]


def like_int32(P):
    torch = _torch()
    return torch.zeros(P.total_np, dtype=torch.int32, device=P.device)


def empty_result(P, want_catalog, want_roots):
    """(in_halo, halo_id, ang_mom, extras) of an input without particles."""
    torch = _torch()
    dev = P.device
    extras = {}
    if want_catalog:
        extras["catalog"] = (torch.zeros(1, dtype=torch.int64, device=dev),
                             torch.zeros(0, dtype=torch.int32, device=dev))
    if want_roots:
        extras["roots"] = torch.zeros(0, dtype=torch.int32, device=dev)
    return (torch.zeros(0, dtype=torch.bool, device=dev), torch.zeros(0, dtype=torch.int32, device=dev),
            torch.zeros((3, P.nhalo), dtype=torch.float64, device=dev), extras)


def split_catalog(off, idx):
    """fof_with_info's (nhalo_found, np_per_halo, halo_indices) from a device
    catalog: counts on the host (one small copy), indices as int64 views of
    one device tensor."""
    torch = _torch()
    counts = torch.diff(off).cpu().numpy().astype(np.int64)
    idx64 = idx.to(torch.int64)
    offs = off.cpu().numpy()
    halo_indices = [idx64[offs[h]:offs[h + 1]] for h in range(len(counts)) if counts[h] > 0]
    np_per_halo = counts[counts > 0]
    return len(halo_indices), np_per_halo, halo_indices
