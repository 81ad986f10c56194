"""CPU-only checks of the boundary: the library loads and exports every entry
point include/pyfof.h declares; the Python layer validates like the reference
(py_fof/__init__.py:178-300) before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

import py_fof_amd
from py_fof_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pyfof.h")).read()
    return sorted(set(re.findall(r"\b(pyfof_[a-z_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = _declared()
    assert "pyfof_fof" in names
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/pyfof.h but not exported"
    assert set(names) <= set(_native.EXPORTS)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_native.NativeLibraryError):
        _native._load_from(str(tmp_path / "nope.so"))


@pytest.mark.parametrize("kwargs,msg", [
    (dict(units="bad"), "units must be"),
    (dict(n_threads=300), "n_threads cannot exceed 256"),
    (dict(bound_check=True), "vp_list is required"),
    (dict(z=-1.0), "1 + z must be positive"),
    (dict(mp_list=-1.0), "mp_list scalar must be positive"),
    (dict(mp_list=np.ones(3)), "does not match total particles"),
])
def test_validation_errors_match_reference(kwargs, msg):
    xp = np.zeros((3, 5))
    with pytest.raises(ValueError, match=re.escape(msg)):
        py_fof_amd.fof(xp, **kwargs)


def test_physical_units_need_softening():
    xp = np.zeros((3, 5))
    vp = np.zeros((3, 5))
    with pytest.raises(ValueError, match="softening must be explicitly provided"):
        py_fof_amd.fof(xp, vp_list=vp, units="physical", bound_check=True)
    with pytest.raises(ValueError, match="Either mp_list or n_sim"):
        py_fof_amd.fof(xp, vp_list=vp, units="physical", bound_check=True, softening=0.01)


def test_shape_errors():
    with pytest.raises(ValueError, match="must have shape"):
        py_fof_amd.fof(np.zeros((2, 5)))
    with pytest.raises(ValueError, match="xp_list must be"):
        py_fof_amd.fof("nope")
    with pytest.raises(ValueError, match="vp_list length"):
        py_fof_amd.fof([np.zeros((3, 2))], vp_list=[np.zeros((3, 2))] * 2)


def test_empty_input_needs_no_device():
    in_h, am = py_fof_amd.fof(np.zeros((3, 0)), return_ang_mom=True)
    assert in_h.shape == (0,) and in_h.dtype == bool and am.shape == (3, 1)


def test_job_descriptor_layout_matches_header():
    # pyfof_job_desc: 7 pointers, 7 doubles, 1 int64, 8 int32 -> 128 bytes
    from py_fof_amd.serve import _JobDesc
    assert ctypes.sizeof(_JobDesc) == 7 * 8 + 7 * 8 + 8 + 8 * 4
    assert _JobDesc.ng_val.offset == 112 and _JobDesc.nhalo.offset == 120


def test_server_argument_errors_raise_before_submission():
    # validation happens in the shared front end, before any device call
    from py_fof_amd import serve
    srv = serve.Server.__new__(serve.Server)  # no device: only the front end is used
    with pytest.raises(ValueError, match="units must be"):
        srv.submit(np.zeros((3, 4)), units="bogus")
    with pytest.raises(ValueError, match="n_threads cannot exceed 256"):
        srv.submit(np.zeros((3, 4)), n_threads=300)
