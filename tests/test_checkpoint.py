"""Exact binary checkpoint (sph_checkpoint_write / sph_checkpoint_read).

CPU tests: the round trip of the host state (every field bit for bit, the
parameters and the step counter) into a handle that has read no input, and
the error paths - a missing, foreign or truncated file and the other
precision form - which leave the handle as it was.  GPU tests: a run resumed
from a checkpoint continues the uninterrupted run (the saved forces and dc/dt
drive its first half kick), to rounding, since the neighbour lists are
rebuilt at the first resumed step."""
import ctypes
import os

import numpy as np
import pytest

import helpers
import inputs
import sph_system
from helpers import comps, relerr

# every field of the state, ID order
P_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "mass", "dens", "press", "fx", "fy", "fz",
            "cxx", "cxy", "cyy", "cxz", "cyz", "czz", "dcdt_xx", "dcdt_xy", "dcdt_yy",
            "dcdt_xz", "dcdt_yz", "dcdt_zz"]
C_FIELDS = ["coll_x", "coll_y", "coll_z", "coll_vx", "coll_vy", "coll_vz", "coll_omegax",
            "coll_omegay", "coll_omegaz", "coll_theta", "coll_quat0", "coll_quatx",
            "coll_quaty", "coll_quatz", "fx_colloid", "fy_colloid", "fz_colloid",
            "tx_colloid", "ty_colloid", "tz_colloid"]
LISTS = ["type", "walls_list", "walls_start", "colloids_list", "colloids_start", "x_center",
         "y_center", "z_center"]


def _params_bytes(s):
    p = s.params
    return bytes(memoryview(p).cast("B"))


def _state(s):
    out = {n: s.get(n).copy() for n in P_FIELDS + LISTS}
    if s.params.N_colloids:
        out.update({n: s.get(n).copy() for n in C_FIELDS})
    return out


def _assert_same_state(a, b):
    assert a.keys() == b.keys()
    for n in a:
        assert a[n].dtype == b[n].dtype and a[n].shape == b[n].shape, n
        np.testing.assert_array_equal(a[n].view(np.uint8), b[n].view(np.uint8), err_msg=n)


@pytest.mark.parametrize("cfg", [helpers.cfg_colloids_3d(n=20), helpers.cfg_colloids_2d(n=32)],
                         ids=["3d_walls_colloids", "2d_walls_colloids"])
def test_round_trip_of_the_host_state(cfg, tmp_path):
    s = helpers.make_system(cfg, str(tmp_path))
    rng = np.random.default_rng(3)
    # values no input file could hold: every field filled with full-precision noise
    for n in P_FIELDS:
        s.set(n, rng.standard_normal(s.N))
    for n in C_FIELDS[:14]:
        s.set(n, rng.standard_normal(s.params.N_colloids))
    path = os.path.join(str(tmp_path), "state.ckpt")
    s.write_checkpoint(path)
    t = sph_system.System(sph_system.FP64)  # read no input: the checkpoint is the state
    t.read_checkpoint(path)
    assert _params_bytes(t) == _params_bytes(s)
    assert t.current_step() == s.current_step()
    _assert_same_state(_state(t), _state(s))


def test_error_paths_leave_the_handle_alone(tmp_path):
    cfg = helpers.cfg_colloids_2d(n=32)
    s = helpers.make_system(cfg, str(tmp_path))
    path = os.path.join(str(tmp_path), "state.ckpt")
    s.write_checkpoint(path)
    t = helpers.make_system(helpers.cfg_couette(n=16), str(tmp_path), name="other")
    before = _state(t)
    with pytest.raises(sph_system.SphError) as e:
        t.read_checkpoint(os.path.join(str(tmp_path), "missing.ckpt"))
    assert e.value.code == sph_system.ERR_IO
    foreign = os.path.join(str(tmp_path), "foreign.ckpt")
    open(foreign, "wb").write(b"not a checkpoint at all" * 10)
    with pytest.raises(sph_system.SphError) as e:
        t.read_checkpoint(foreign)
    assert e.value.code == sph_system.ERR_IO
    data = open(path, "rb").read()
    for cut in (4, 200, len(data) // 2, len(data) - 1):
        short = os.path.join(str(tmp_path), f"short{cut}.ckpt")
        open(short, "wb").write(data[:cut])
        with pytest.raises(sph_system.SphError) as e:
            t.read_checkpoint(short)
        assert e.value.code == sph_system.ERR_IO, cut
    f32 = sph_system.System(sph_system.FP32)
    with pytest.raises(sph_system.SphError) as e:
        f32.read_checkpoint(path)
    assert e.value.code == sph_system.ERR_ARG
    _assert_same_state(_state(t), before)
    assert t.params.N != s.params.N
    # and a good file still reads into the handle afterwards
    t.read_checkpoint(path)
    assert t.params.N == s.params.N


def test_write_before_initialize_is_refused(tmp_path):
    s = sph_system.System(sph_system.FP64)
    with pytest.raises(sph_system.SphError) as e:
        s.write_checkpoint(os.path.join(str(tmp_path), "x.ckpt"))
    assert e.value.code == sph_system.ERR_STATE


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [sph_system.FP64, sph_system.FP32])
@pytest.mark.parametrize("cfg", [helpers.cfg_colloids_3d(n=20), helpers.cfg_colloids_2d(n=40)],
                         ids=["3d_walls_colloids", "2d_walls_colloids"])
def test_resumed_run_continues_the_run(cfg, precision, tmp_path):
    s = helpers.make_system(cfg, str(tmp_path), precision)
    s.prepare()
    s.step(12)
    s.synchronize()
    path = os.path.join(str(tmp_path), "state.ckpt")
    s.write_checkpoint(path)
    saved = _state(s)
    s.step(15)                       # the uninterrupted run
    s.synchronize()
    t = sph_system.System(precision)
    t.read_checkpoint(path)
    assert t.current_step() == 12
    _assert_same_state(_state(t), saved)
    t.step(15)                       # the resumed run
    t.synchronize()
    assert t.current_step() == s.current_step() == 27
    d = s.params.dim
    tol = 1e-12 if precision == sph_system.FP64 else 1e-5
    for n in comps("v", d) + ["cxx", "cxy", "cyy"] + comps("f", d) + ["dens"]:
        assert relerr(t.get(n), s.get(n)) <= tol, n
    for n in comps("x", d):
        assert np.abs(t.get(n) - s.get(n)).max() <= (1e-13 if tol == 1e-12 else 1e-8), n
    if s.params.N_colloids:
        for n in ("coll_x", "coll_y", "coll_vx", "coll_omegaz"):
            assert np.abs(t.get(n) - s.get(n)).max() <= 1e3 * tol * max(
                1.0, np.abs(s.get(n)).max()), n
    ke_s, ke_t = s.kinetic_energy(), t.kinetic_energy()
    assert abs(ke_t - ke_s) <= tol * ke_s
