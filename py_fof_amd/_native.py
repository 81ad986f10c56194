"""ctypes binding of libpyfof_hip.so (the C-ABI declared in include/pyfof.h).

There is no fallback: if the library is missing or cannot be loaded, every
call raises. Build it with ``python -m py_fof_amd.build``.
"""
import atexit
import ctypes
import os
import sys
import threading
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyfof_hip.so")

MEM_HOST = 0
MEM_DEVICE = 1
OPT_CATALOG = 1
OPT_DIRECT_MAX = 2
OPT_DIRECT_FP32 = 3
OPT_BATCHED = 4
OPT_RECORD_PHI = 5
OPT_POISON_ALLOC = 6  # process-wide: 0 off, 1 + b fills new device allocations with byte b
OPT_TIME_FFT = 7
OPT_FFT_LINES = 8  # value 1000 * n + lines

EXPORTS = [
    "pyfof_fof", "pyfof_fof_ctx", "pyfof_ctx_create", "pyfof_ctx_destroy",
    "pyfof_catalog_size", "pyfof_catalog_copy", "pyfof_ctx_timings", "pyfof_ctx_stats",
    "pyfof_ctx_stream", "pyfof_ctx_set_option", "pyfof_ctx_potential",
    "pyfof_last_error", "pyfof_version", "pyfof_release", "pyfof_fof_noalloc",
    # request server (py_fof_amd.serve)
    "pyfof_server_create", "pyfof_server_destroy", "pyfof_job_footprint", "pyfof_submit",
    "pyfof_poll", "pyfof_wait", "pyfof_server_status", "pyfof_server_last_error",
]

_lib = None
_lock = threading.Lock()


class NativeLibraryError(RuntimeError):
    pass


def _torch_first():
    """The torch wheel carries its own HIP runtime beside the system one this
    library links. Both share the device, but the runtime that initialises
    second finds no GPU when it is torch's. When torch is already imported,
    its runtime is initialised here, before this library's first HIP call."""
    t = sys.modules.get("torch")
    if t is None:
        return
    try:
        if t.cuda.device_count() > 0:
            t.cuda.init()
    except Exception:
        pass


def load(path=LIB_PATH):
    """Loads and types the library once; raises NativeLibraryError if absent."""
    global _lib
    with _lock:
        if _lib is None:
            _torch_first()
            _lib = _load_from(path)
        return _lib


def _load_from(path):
    if not os.path.exists(path):
        raise NativeLibraryError(
            f"{path} not found: build it with `python -m py_fof_amd.build`")
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise NativeLibraryError(f"cannot load {path}: {e}") from e
    d, i32, i64, vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    lib.pyfof_fof.restype = i32
    lib.pyfof_fof.argtypes = [vp, vp, vp, vp, i32, d, d, i32, i32, i32, d, d, d, d, i32, i32,
                              i64, d, vp, vp, vp, i32]
    lib.pyfof_fof_ctx.restype = i32
    lib.pyfof_fof_ctx.argtypes = [vp, vp, vp, vp, vp, i32, d, d, i32, i32, i32, d, d, d, d,
                                  i32, i32, i64, d, vp, vp, vp, vp, i32]
    lib.pyfof_ctx_create.restype = vp
    lib.pyfof_ctx_create.argtypes = [i32]
    lib.pyfof_ctx_destroy.restype = None
    lib.pyfof_ctx_destroy.argtypes = [vp]
    lib.pyfof_catalog_size.restype = i32
    lib.pyfof_catalog_size.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i64)]
    lib.pyfof_catalog_copy.restype = i32
    lib.pyfof_catalog_copy.argtypes = [vp, vp, vp, i32]
    lib.pyfof_ctx_timings.restype = i32
    lib.pyfof_ctx_timings.argtypes = [vp, ctypes.POINTER(d), i32]
    lib.pyfof_ctx_set_option.restype = i32
    lib.pyfof_ctx_set_option.argtypes = [vp, i32, i64]
    lib.pyfof_ctx_stats.restype = i32
    lib.pyfof_ctx_stats.argtypes = [vp, ctypes.POINTER(i64), i32]
    lib.pyfof_ctx_potential.restype = i64
    lib.pyfof_ctx_potential.argtypes = [vp, vp, vp, i64, i32]
    lib.pyfof_ctx_stream.restype = vp
    lib.pyfof_ctx_stream.argtypes = [vp]
    lib.pyfof_last_error.restype = ctypes.c_char_p
    lib.pyfof_last_error.argtypes = []
    lib.pyfof_version.restype = ctypes.c_char_p
    lib.pyfof_version.argtypes = []
    lib.pyfof_release.restype = None
    lib.pyfof_release.argtypes = []
    atexit.register(_release_all)
    return lib


_live = weakref.WeakSet()  # contexts still open
_servers = weakref.WeakSet()  # request servers still running


def register_for_exit(obj):
    """Closes obj (a request server) at interpreter exit, before the
    contexts and the default contexts are released."""
    _servers.add(obj)


def _release_all():
    """Runs at interpreter exit, before the HIP runtime and any profiler
    attached to it tear down: destroys every context still open (streams,
    HBM workspaces) and the library's per-device default contexts, so that
    nothing is freed after the runtime is gone."""
    for group in (_servers, _live):
        for c in list(group):
            try:
                c.close()
            except Exception:
                pass
    _default_ctx.clear()
    if _lib is not None:
        _lib.pyfof_release()


def last_error():
    return load().pyfof_last_error().decode(errors="replace")


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor on the device


class Context:
    """A device, its HIP stream and the reusable HBM workspace of the library."""

    def __init__(self, device=0):
        self.lib = load()
        self.h = self.lib.pyfof_ctx_create(int(device))
        if not self.h:
            raise NativeLibraryError("pyfof_ctx_create failed: " + last_error())
        self.device = device
        _live.add(self)

    def close(self):
        if self.h:
            self.lib.pyfof_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fof(self, xp, vp, mp, np_halo, L_Box, link, np_halo_min, bound_check, verbose,
            softening, a_val, b_val, s_15, n_threads, total_np, ng, ng_global, in_halo,
            halo_id, ang_mom, root_out=None, mem_kind=MEM_HOST):
        np_halo = np.ascontiguousarray(np_halo, dtype=np.int32)
        return self.lib.pyfof_fof_ctx(
            self.h, _ptr(xp), _ptr(vp), _ptr(mp), np_halo.ctypes.data, len(np_halo),
            float(L_Box), float(link), int(np_halo_min), int(bound_check), int(verbose),
            float(softening), float(a_val), float(b_val), float(s_15), int(n_threads),
            int(total_np), int(ng), float(ng_global), _ptr(in_halo), _ptr(halo_id),
            _ptr(ang_mom), _ptr(root_out), int(mem_kind))

    def set_option(self, option, value):
        err = self.lib.pyfof_ctx_set_option(self.h, int(option), int(value))
        if err:
            raise RuntimeError("pyfof_ctx_set_option failed: " + last_error())

    def catalog_size(self):
        """(halos, members) of the last call's catalog."""
        nh = ctypes.c_int32(0)
        nm = ctypes.c_int64(0)
        err = self.lib.pyfof_catalog_size(self.h, ctypes.byref(nh), ctypes.byref(nm))
        if err:
            raise RuntimeError("pyfof_catalog_size failed: " + last_error())
        return nh.value, nm.value

    def catalog_copy(self, offsets, indices, mem_kind):
        """Copies the catalog into caller buffers (numpy or device tensors)."""
        err = self.lib.pyfof_catalog_copy(self.h, _ptr(offsets), _ptr(indices), int(mem_kind))
        if err:
            raise RuntimeError("pyfof_catalog_copy failed: " + last_error())

    def catalog(self):
        """(offsets int64[nh+1], indices int32[total]) of the last call."""
        nh, nm = self.catalog_size()
        off = np.zeros(nh + 1, np.int64)
        idx = np.zeros(max(nm, 1), np.int32)
        self.catalog_copy(off, idx, MEM_HOST)
        return off, idx[:nm]

    def timings(self):
        buf = (ctypes.c_double * 8)()
        n = self.lib.pyfof_ctx_timings(self.h, buf, 8)
        return [buf[i] for i in range(n)]

    def stats(self):
        """dict of the last call's sizes (particles, cells, dense cells, halos, members)."""
        buf = (ctypes.c_int64 * 18)()
        n = self.lib.pyfof_ctx_stats(self.h, buf, 18)
        keys = ["n", "ncell", "ndense", "nhalo", "nmem", "nvalid", "n_ok", "n_too_few",
                "n_max_iter", "n_direct", "n_pm48", "n_pm64", "n_pm128", "n_pm256", "passes",
                "n_direct_large", "fft_slots", "fft_bytes"]
        return {keys[i]: int(buf[i]) for i in range(n)}

    def potential(self):
        """(phi f32, particle index int32) of the last bound_check call's
        candidate members, recorded with OPT_RECORD_PHI."""
        n = self.lib.pyfof_ctx_potential(self.h, None, None, 0, MEM_HOST)
        if n < 0:
            raise RuntimeError(last_error())
        phi = np.zeros(max(n, 1), np.float32)
        idx = np.zeros(max(n, 1), np.int32)
        if n > 0 and self.lib.pyfof_ctx_potential(self.h, phi.ctypes.data, idx.ctypes.data, n,
                                                  MEM_HOST) < 0:
            raise RuntimeError(last_error())
        return phi[:n], idx[:n]

    def stream(self):
        return self.lib.pyfof_ctx_stream(self.h)


_default_ctx = {}


def default_context(device=0):
    c = _default_ctx.get(device)
    if c is None:
        c = Context(device)
        _default_ctx[device] = c
    return c
