"""Request server: FoF / unbinding jobs submitted as they arrive.

``Server.submit`` takes the arguments of :func:`py_fof_amd.fof` (same
validation, units and error messages) and returns a :class:`Job` at once.
The library queues the job and starts it when one of the server's workers
is free and its estimated HBM footprint fits the budget next to the
running jobs. Workers have their own HIP streams and workspaces, so
admitted jobs overlap on the GPU. ``Job.result()`` blocks and returns what
``fof()`` returns for the same arguments, or raises the exception ``fof()``
would raise. A failing job does not affect the others.

The reference has no counterpart: its native call holds module state and
must run serially.
"""
import ctypes
import threading

import numpy as np

from . import _device, _finish, _native, _outputs, _prepare

__all__ = ["Server", "Job", "ServerError"]


class ServerError(RuntimeError):
    pass


class _JobDesc(ctypes.Structure):
    # pyfof_job_desc (include/pyfof.h)
    _fields_ = [
        ("xp", ctypes.c_void_p), ("vp", ctypes.c_void_p), ("mp", ctypes.c_void_p),
        ("np_halo", ctypes.c_void_p), ("in_halo", ctypes.c_void_p),
        ("halo_id", ctypes.c_void_p), ("ang_mom", ctypes.c_void_p),
        ("L_Box", ctypes.c_double), ("linking_length", ctypes.c_double),
        ("softening", ctypes.c_double), ("a_val", ctypes.c_double), ("b_val", ctypes.c_double),
        ("s_15_val", ctypes.c_double), ("ng_global_val", ctypes.c_double),
        ("ng_val", ctypes.c_int64),
        ("nhalo", ctypes.c_int32), ("np_halo_min", ctypes.c_int32),
        ("bound_check", ctypes.c_int32), ("verbose", ctypes.c_int32),
        ("n_threads", ctypes.c_int32), ("total_np", ctypes.c_int32),
        ("mem_kind", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


def _bind(lib):
    if getattr(lib, "_pyfof_server_bound", False):
        return lib
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.pyfof_server_create.restype = vp
    lib.pyfof_server_create.argtypes = [i32, i64, i32]
    lib.pyfof_server_destroy.restype = None
    lib.pyfof_server_destroy.argtypes = [vp]
    lib.pyfof_job_footprint.restype = i64
    lib.pyfof_job_footprint.argtypes = [ctypes.POINTER(_JobDesc)]
    lib.pyfof_submit.restype = i64
    lib.pyfof_submit.argtypes = [vp, ctypes.POINTER(_JobDesc)]
    for f in (lib.pyfof_poll, lib.pyfof_wait):
        f.restype = i32
        f.argtypes = [vp, i64, ctypes.POINTER(i32), ctypes.c_char_p, i32,
                      ctypes.POINTER(ctypes.c_double)]
    lib.pyfof_server_status.restype = i32
    lib.pyfof_server_status.argtypes = [vp, ctypes.POINTER(i64), i32]
    lib.pyfof_server_last_error.restype = ctypes.c_char_p
    lib.pyfof_server_last_error.argtypes = []
    lib._pyfof_server_bound = True
    return lib


class Job:
    """A submitted call. ``result()`` waits and returns fof()'s value."""

    def __init__(self, server, jid, P, outs, return_ang_mom, keep):
        self._server = server
        self.id = jid
        self._P = P
        self._outs = outs
        self._ret_am = return_ang_mom
        self._keep = keep  # arrays the device reads until the job is done
        self._done = False
        self._ierr = None
        self._msg = ""
        self.queue_ms = None
        self.run_ms = None

    def _collect(self, block):
        if self._done:
            return True
        lib = self._server._lib
        ierr = ctypes.c_int32(0)
        msg = ctypes.create_string_buffer(512)
        times = (ctypes.c_double * 2)()
        f = lib.pyfof_wait if block else lib.pyfof_poll
        r = f(self._server._h, self.id, ctypes.byref(ierr), msg, 512, times)
        if r == 0:
            return False
        if r < 0:
            raise ServerError(f"unknown job id {self.id}")
        self._done = True
        self._ierr = int(ierr.value)
        self._msg = msg.value.decode(errors="replace")
        self.queue_ms, self.run_ms = float(times[0]), float(times[1])
        self._keep = None
        return True

    def done(self):
        """True once the job has finished (never blocks)."""
        return self._collect(False)

    def result(self, with_halo_id=False):
        """fof()'s return value (``in_halo`` or ``(in_halo, ang_mom)``); with
        ``with_halo_id`` the tuple (in_halo, halo_id, ang_mom)."""
        self._collect(True)
        if self._P is None:  # empty input
            inh, hid, am = self._outs
        else:
            inh, hid, am = _finish(self._P, self._ierr, *self._outs, message=self._msg)
        if with_halo_id:
            return inh, hid, am
        return (inh, am) if self._ret_am else inh


class Server:
    """Continuous-batching FoF server on one GPU.

    device:      HIP device index.
    hbm_budget:  bytes of HBM the running jobs may claim together (by the
                 library's footprint estimate).
    slots:       workers (concurrent jobs), 1..16.
    """

    def __init__(self, device=0, hbm_budget=64 << 30, slots=4):
        self._lib = _bind(_native.load())
        self._h = self._lib.pyfof_server_create(int(device), int(hbm_budget), int(slots))
        if not self._h:
            raise ServerError(self._lib.pyfof_server_last_error().decode(errors="replace"))
        self.device = int(device)
        self._lock = threading.Lock()
        _native.register_for_exit(self)

    def close(self):
        """Waits for every queued and running job, then frees the workers."""
        if getattr(self, "_h", None):
            self._lib.pyfof_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def status(self):
        buf = (ctypes.c_int64 * 6)()
        n = self._lib.pyfof_server_status(self._h, buf, 6)
        keys = ["queued", "running", "admitted_bytes", "budget_bytes", "resident_bytes",
                "claimed_bytes"]
        return {keys[i]: int(buf[i]) for i in range(n)}

    def submit(self, xp_list, vp_list=None, **kwargs):
        """Queues fof(xp_list, vp_list, **kwargs); returns a Job. Argument
        errors raise here, exactly as fof() raises them."""
        ret_am = bool(kwargs.get("return_ang_mom", False))
        P = _prepare(xp_list, vp_list=vp_list, **kwargs)
        if P.empty:
            if P.device is not None:
                outs = _device.empty_result(P, False, False)[:3]
            else:
                outs = (np.zeros(0, dtype=bool), np.zeros(0, dtype=np.int32),
                        np.zeros((3, P.nhalo), dtype=np.float64))
            job = Job(self, -1, None, outs, ret_am, None)
            job._done = True
            return job
        if P.device is not None and (P.device.index or 0) != self.device:
            raise ValueError(f"inputs are on {P.device}, the server on device {self.device}")
        outs = _outputs(P)
        in_halo, halo_id, ang_mom = outs
        d = _JobDesc()
        d.xp = _native._ptr(P.xp)
        d.vp = _native._ptr(P.vp)
        d.mp = _native._ptr(P.mp)
        d.np_halo = P.np_halo.ctypes.data
        d.in_halo = _native._ptr(in_halo)
        d.halo_id = _native._ptr(halo_id)
        d.ang_mom = _native._ptr(ang_mom) if P.need_am else None
        d.L_Box, d.linking_length, d.softening = P.L, P.link, P.softening
        d.a_val, d.b_val, d.s_15_val, d.ng_global_val = P.a, P.b, P.s15, P.ng_global
        d.ng_val = P.ng
        d.nhalo, d.np_halo_min, d.bound_check = P.nhalo, P.np_min, P.bound_check
        d.verbose, d.n_threads, d.total_np = P.verbose, P.n_threads, P.total_np
        d.mem_kind = _native.MEM_HOST if P.device is None else _native.MEM_DEVICE
        keep = (P.xp, P.vp, P.mp, P.np_halo, outs)
        with self._lock:
            jid = self._lib.pyfof_submit(self._h, ctypes.byref(d))
        if jid < 0:
            raise ServerError(self._lib.pyfof_server_last_error().decode(errors="replace"))
        return Job(self, int(jid), P, outs, ret_am, keep)

    def footprint(self, xp_list, vp_list=None, **kwargs):
        """The library's HBM estimate for a job with these arguments."""
        P = _prepare(xp_list, vp_list=vp_list, **kwargs)
        if P.empty:
            return 0
        d = _JobDesc()
        d.vp = 1 if P.vp is not None else None
        d.mp = 1 if P.mp is not None else None
        d.total_np, d.bound_check = P.total_np, P.bound_check
        d.mem_kind = _native.MEM_HOST if P.device is None else _native.MEM_DEVICE
        return int(self._lib.pyfof_job_footprint(ctypes.byref(d)))
