"""The ctypes binding a py_fof maintainer puts in place of the f2py module
(py_fof/__init__.py:331-336), kept here word for word so that the tests call
pyfof_fof exactly as that binding does."""
# py_fof/_hip.py -- ctypes binding of the MI355X library
import ctypes, os
import numpy as np

_lib = ctypes.CDLL(os.environ.get("PYFOF_HIP_LIB", "libpyfof_hip.so"))
_d, _i32, _i64, _vp = ctypes.c_double, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
_lib.pyfof_fof.restype = _i32
_lib.pyfof_fof.argtypes = [_vp, _vp, _vp, _vp, _i32, _d, _d, _i32, _i32, _i32, _d, _d, _d,
                           _d, _i32, _i32, _i64, _d, _vp, _vp, _vp, _i32]
PYFOF_MEM_HOST = 0

def _p(a):
    return None if a is None else a.ctypes.data

def fof_fortran(xp, vp, mp, np_halo, nhalo, L_Box, link, np_halo_min, bound_check,
                verbose, softening, a_val, b_val, s_15, n_threads, total_np,
                ng_val=256, ng_global_val=512.0):
    """Same call and return tuple as the f2py wrapper."""
    xp = np.asfortranarray(xp, dtype=np.float64)
    vp = None if vp is None else np.asfortranarray(vp, dtype=np.float64)
    mp = None if mp is None else np.ascontiguousarray(mp, dtype=np.float64)
    np_halo = np.ascontiguousarray(np_halo, dtype=np.int32)
    in_halo = np.empty(total_np, np.int32)   # the library writes every entry,
    halo_id = np.empty(total_np, np.int32)   # zeros on an error or empty input
    ang_mom = np.empty((3, nhalo), np.float64, order="F")
    ierr = _lib.pyfof_fof(_p(xp), _p(vp), _p(mp), _p(np_halo), nhalo, L_Box, link,
                          np_halo_min, bound_check, verbose, softening, a_val, b_val, s_15,
                          n_threads, total_np, ng_val, ng_global_val, _p(in_halo),
                          _p(halo_id), _p(ang_mom), PYFOF_MEM_HOST)
    return in_halo, halo_id, ang_mom, ierr
