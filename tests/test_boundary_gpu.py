"""The drop-in boundary: pyfof_fof, called through the ctypes binding a
py_fof maintainer would put in place of the f2py module
(py_fof/__init__.py:331-336), against the oracle.

* the f2py return tuple (in_halo, halo_id, ang_mom, ierr) on the 262,144-
  particle clumped set: in_halo and halo_id bit-exact, ang_mom to rtol 1e-9;
* with bound_check on a direct-summation halo: bit-exact;
* outputs handed over uninitialised (np.empty) are fully written, and zeros
  come back on every path that returns no result: ierr 10 (n_threads > 256),
  ierr 1 (sizes that do not add up), empty input (fof_fortran.f90:110-123);
* the same entry point on device memory (PYFOF_MEM_DEVICE).
It runs on the library's per-device default context, which no other test
uses.
"""
import os

import numpy as np
import pytest
# imported before the library loads, so that torch's HIP runtime starts first
import torch  # noqa: F401

from oracle import oracle_fof, oracle_fof_bound

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "py_fof_amd", "libpyfof_hip.so")


@pytest.fixture(scope="module")
def stub():
    from py_fof_amd import _native
    _native.load()  # the device runtimes start in the supported order
    os.environ["PYFOF_HIP_LIB"] = LIB
    import integration_stub
    return integration_stub


def _c1(seed=0):
    from py_fof_amd.synth import clumpy_plummer
    xp = clumpy_plummer(n=262_144, L=64.0, n_clumps=400, seed=seed)
    vp = np.asfortranarray(np.random.default_rng(seed + 1).standard_normal(xp.shape))
    return xp, vp


def _args(xp, vp, L, b, np_min=10, bound=0, soft=0.1, nt=0, s15=17.152):
    n = xp.shape[1]
    return (xp, vp, None, np.array([n], np.int32), 1, L, b, np_min, bound, 0, soft, 1.0, 0.0,
            s15, nt, n)


def test_pyfof_fof_plain_matches_oracle(stub):
    xp, vp = _c1()
    n = xp.shape[1]
    in_h, hid, am, ierr = stub.fof_fortran(*_args(xp, vp, 64.0, 0.2))
    assert ierr == 0
    o_in, o_hid, o_am = oracle_fof(xp, [n], 64.0, 0.2, 10, vp=vp, n_threads=4)
    assert np.array_equal(in_h, o_in)
    assert np.array_equal(hid, o_hid)
    assert hid.max() > 100
    assert np.allclose(am, o_am, rtol=1e-9, atol=1e-12 * np.abs(o_am).max())
    # vp NULL: zero velocities, no angular momentum (all zeros)
    in2, hid2, am2, ierr = stub.fof_fortran(*_args(xp, None, 64.0, 0.2))
    assert ierr == 0 and np.array_equal(in2, in_h) and np.array_equal(hid2, hid)
    assert not am2.any()


def test_pyfof_fof_bound_matches_oracle(stub):
    from helpers import chain_halo
    rng = np.random.default_rng(888)
    xp, expect = chain_halo(200, 2.0, 400, rng)
    vp = rng.standard_normal(xp.shape) * 0.1
    hv = rng.choice(np.nonzero(expect)[0], 20, replace=False)
    vp[:, hv] *= 100.0
    vp = np.asfortranarray(vp)
    n = xp.shape[1]
    in_h, hid, am, ierr = stub.fof_fortran(*_args(xp, vp, 512.0, 2.0, bound=1))
    assert ierr == 0 and in_h.sum() == 180
    r = oracle_fof_bound(xp, [n], 512.0, 2.0, 10, vp, app=0.1)
    assert np.array_equal(in_h, r["in_halo"]) and np.array_equal(hid, r["halo_id"])
    assert np.allclose(am, r["ang_mom"], rtol=1e-9, atol=1e-12)


def test_pyfof_fof_error_paths_clear_outputs(stub):
    xp, vp = _c1(seed=3)
    n = xp.shape[1]
    args = list(_args(xp, vp, 64.0, 0.2))
    args[14] = 300  # n_threads > 256 -> ierr 10
    in_h, hid, am, ierr = stub.fof_fortran(*args)
    assert ierr == 10
    assert not in_h.any() and not hid.any() and not am.any()
    # region sizes that do not add up to total_np -> ierr 1
    args = list(_args(xp, vp, 64.0, 0.2))
    args[3] = np.array([n - 5], np.int32)
    in_h, hid, am, ierr = stub.fof_fortran(*args)
    assert ierr == 1 and not in_h.any() and not hid.any() and not am.any()
    # bound_check with too small a grid -> ierr 1 after the FoF ran
    args = list(_args(xp, vp, 64.0, 0.2, bound=1))
    in_h, hid, am, ierr = stub.fof_fortran(*args, ng_val=128)
    assert ierr == 1 and not in_h.any() and not hid.any() and not am.any()
    # empty input: ierr 0, nothing to write; empty regions: zero ang_mom
    e = np.zeros((3, 0))
    in_h, hid, am, ierr = stub.fof_fortran(e, None, None, np.zeros(0, np.int32), 0, 64.0, 0.2,
                                           10, 0, 0, 0.1, 1.0, 0.0, 17.152, 0, 0)
    assert ierr == 0 and in_h.size == 0 and hid.size == 0
    in_h, hid, am, ierr = stub.fof_fortran(e, None, None, np.zeros(2, np.int32), 2, 64.0, 0.2,
                                           10, 0, 0, 0.1, 1.0, 0.0, 17.152, 0, 0)
    assert ierr == 0 and am.shape == (3, 2) and not am.any()
    # the default context still works after the failed calls
    in_h, hid, _, ierr = stub.fof_fortran(*_args(xp, vp, 64.0, 0.2))
    o_in, o_hid, _ = oracle_fof(xp, [n], 64.0, 0.2, 10, n_threads=4)
    assert ierr == 0 and np.array_equal(in_h, o_in) and np.array_equal(hid, o_hid)


def test_pyfof_fof_device_memory(stub):
    xp, vp = _c1(seed=5)
    n = xp.shape[1]
    h_in, h_hid, h_am, ierr = stub.fof_fortran(*_args(xp, vp, 64.0, 0.2))
    assert ierr == 0
    dev = torch.device("cuda:0")
    x_d = torch.from_numpy(np.ascontiguousarray(xp.T)).to(dev)
    v_d = torch.from_numpy(np.ascontiguousarray(vp.T)).to(dev)
    in_d = torch.full((n,), -7, dtype=torch.int32, device=dev)
    id_d = torch.full((n,), -7, dtype=torch.int32, device=dev)
    am_d = torch.full((1, 3), -7.0, dtype=torch.float64, device=dev)
    nh = np.array([n], np.int32)
    torch.cuda.synchronize()

    def call(nt):
        return stub._lib.pyfof_fof(x_d.data_ptr(), v_d.data_ptr(), None, nh.ctypes.data, 1, 64.0,
                                   0.2, 10, 0, 0, 0.1, 1.0, 0.0, 17.152, nt, n, 256, 512.0,
                                   in_d.data_ptr(), id_d.data_ptr(), am_d.data_ptr(), 1)

    assert call(0) == 0
    assert np.array_equal(in_d.cpu().numpy(), h_in)
    assert np.array_equal(id_d.cpu().numpy(), h_hid)
    assert np.allclose(am_d.cpu().numpy().T, h_am, rtol=1e-12, atol=1e-12 * np.abs(h_am).max())
    # an error clears device outputs too
    assert call(300) == 10
    assert not in_d.any().item() and not id_d.any().item() and not am_d.any().item()
