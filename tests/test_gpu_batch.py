"""Batches of independent simulations of uneven size (sph_batch_* in sph.h).

A batch concatenates the simulations and runs each phase as one launch over
all of them (replayed as a captured graph).  The kernels are the ones a single
simulation runs, on the same data in the same order, so every simulation of a
batch must reproduce its own stand-alone run bit for bit - in both precision
forms, with walls, colloids (counts 0 to 3) and under every launch shape.
"""
import numpy as np
import pytest

import helpers
import sph_system

pytestmark = pytest.mark.gpu


def _sims(kind):
    if kind == "2d":
        return [
            helpers.cfg_couette(n=24, tau=0.5, Vwall_bottom=-0.3, Vwall_top=0.3),
            helpers.cfg_colloids_2d(n=40),
            helpers.cfg_couette(n=40, tau=2.0, Vwall_bottom=-0.7, Vwall_top=0.7),
            helpers.cfg_colloids_2d(n=48, N_colloids=3, coll_R=0.06,
                                    coll_x=[[0.25, 0.3], [0.6, 0.45], [0.4, 0.75]],
                                    omega_magnetic=5.0),
        ]
    if kind == "3d_walls":
        return [helpers.cfg_colloids_3d(n=20),
                helpers.cfg_colloids_3d(n=24, n_colloids=2, tau=0.3,
                                        coll_x=[[0.3, 0.5, 0.5], [0.6, 0.5, 0.5]]),
                helpers.cfg_colloids_3d(n=18, n_colloids=0)]
    return [helpers.cfg_sine(n=12), helpers.cfg_sine(n=16, tau=2.0),
            helpers.cfg_sine(n=14, ext_force_type=1, ext_force=[0.5, 0.2, -0.1])]


def _systems(kind, tmpdir, precision, tag):
    out = []
    for k, cfg in enumerate(_sims(kind)):
        s = helpers.make_system(cfg, tmpdir, precision, noise=1e-2, seed=k + 1,
                                name=f"{tag}{k}")
        helpers.perturb(s, seed=20 + k)
        out.append(s)
    return out


def _names(p):
    d = p.dim
    n = helpers.comps("x", d) + helpers.comps("v", d) + helpers.comps("c", d)
    n += ["dens", "press"] + helpers.comps("f", d) + helpers.comps("dcdt", d)
    if p.N_colloids:
        n += ["coll_x", "coll_y", "coll_vx", "coll_omegaz", "fx_colloid", "fy_colloid",
              "tz_colloid"]
    if p.wall:
        n += ["fx_wall", "fy_wall"]
    return n


def _compare(batch, singles, when):
    for k, s in enumerate(singles):
        for n in _names(s.params):
            np.testing.assert_array_equal(batch.get(k, n), s.get(n),
                                          err_msg=f"sim {k} field {n} {when}")
    ke = batch.kinetic_energy()
    np.testing.assert_array_equal(ke, [s.kinetic_energy() for s in singles], err_msg=when)
    obs = batch.observables()
    for k, s in enumerate(singles):
        want = np.array(list(s.observables().values()))
        np.testing.assert_array_equal(obs[k], want, err_msg=f"observables sim {k} {when}")


@pytest.mark.parametrize("precision", [sph_system.FP64, sph_system.FP32])
@pytest.mark.parametrize("kind", ["2d", "3d_walls", "3d_periodic"])
def test_batch_reproduces_each_simulation(kind, precision, tmp_path):
    singles = _systems(kind, str(tmp_path), precision, "s")
    members = _systems(kind, str(tmp_path), precision, "b")
    b = sph_system.Batch(precision)
    for s in members:
        b.add(s)
    assert len(b) == len(singles)
    b.prepare()
    for s in singles:
        s.prepare()
    _compare(b, singles, "after prepare")
    b.step(3)
    for s in singles:
        s.step(3)
    _compare(b, singles, "after 3 steps")
    b.step(4)
    for s in singles:
        s.step(4)
    assert b.current_step() == 7
    _compare(b, singles, "after 7 steps")


def test_launch_shape_does_not_change_results(tmp_path):
    ref = sph_system.Batch(sph_system.FP32)
    for s in _systems("2d", str(tmp_path), sph_system.FP32, "r"):
        ref.add(s)
    ref.prepare()
    ref.step(4)
    for block, xcd in ((64, False), (128, True), (256, False)):
        b = sph_system.Batch(sph_system.FP32)
        for s in _systems("2d", str(tmp_path), sph_system.FP32, f"l{block}{int(xcd)}"):
            b.add(s)
        b.set_launch(block, xcd)
        b.prepare()
        b.step(2)
        b.set_launch(256 if block != 256 else 64, not xcd)   # change shape mid-run
        b.step(2)
        for k in range(len(b)):
            for n in ("x", "y", "vx", "vy", "cxy", "fx", "dens"):
                np.testing.assert_array_equal(b.get(k, n), ref.get(k, n),
                                              err_msg=f"{block} {xcd} sim {k} {n}")


def test_launch_shape_of_one_3d_simulation(tmp_path, monkeypatch):
    """A single 3D simulation (a batch of one) under every launch shape, the
    shape changed mid-run: its tile order (y bands of each XCD's share), list
    windows and rebuilds (skin 0.01 rcut) follow the shape, the trajectory is
    bitwise the same."""
    monkeypatch.setenv("SPH_VERLET_SKIN", "0.01")
    cfg = helpers.cfg_colloids_3d(n=24)

    def run(name, shapes):
        b = sph_system.Batch(sph_system.FP32)
        b.add(helpers.make_system(cfg, str(tmp_path), sph_system.FP32, noise=1e-2, name=name))
        b.set_launch(*shapes[0])
        b.prepare()
        b.step(4)
        b.set_launch(*shapes[1])
        b.step(4)
        return {n: b.get(0, n) for n in ("x", "y", "z", "vx", "vz", "cxy", "fx", "dens")}

    ref = run("r", ((256, True), (256, True)))
    for shapes in (((64, True), (128, False)), ((128, True), (64, True)),
                   ((256, False), (256, True))):
        out = run(f"l{shapes[0][0]}{shapes[1][0]}", shapes)
        for n in ref:
            np.testing.assert_array_equal(out[n], ref[n], err_msg=f"{shapes} {n}")


def test_autotune_keeps_the_trajectory(tmp_path):
    singles = _systems("2d", str(tmp_path), sph_system.FP64, "s")
    b = sph_system.Batch(sph_system.FP64)
    for s in _systems("2d", str(tmp_path), sph_system.FP64, "b"):
        b.add(s)
    b.prepare()
    block, xcd, ms = b.autotune(steps=1)
    assert block in (64, 128, 256) and ms > 0
    assert b.get_launch() == (block, xcd)
    n = b.current_step()
    assert n == 6 * 3
    for s in singles:
        s.prepare()
        s.step(n)
    _compare(b, singles, f"after autotune ({n} steps)")


def test_batch_of_one_equals_system(tmp_path):
    cfg = helpers.cfg_colloids_3d(n=20)
    s = helpers.make_system(cfg, str(tmp_path), sph_system.FP32, name="s")
    m = helpers.make_system(cfg, str(tmp_path), sph_system.FP32, name="m")
    b = sph_system.Batch(sph_system.FP32)
    b.add(m)
    b.prepare()
    s.prepare()
    b.step(5)
    s.step(5)
    _compare(b, [s], "after 5 steps")


def test_failed_simulation_is_flagged_and_isolated(tmp_path):
    """A particle moved beyond a wall band in one simulation flags that
    simulation only; the others reproduce their stand-alone runs bit for bit.
    The failed simulation is held: further steps leave its state as it was."""
    singles = _systems("2d", str(tmp_path), sph_system.FP64, "s")
    members = _systems("2d", str(tmp_path), sph_system.FP64, "b")
    y = members[1].get("y")
    y[5] = 3.0
    members[1].set("y", y)
    b = sph_system.Batch(sph_system.FP64)
    for s in members:
        b.add(s)
    b.prepare()
    b.step(3)
    flags = b.failed()
    assert flags[1] & 1 and not flags[0] and not flags[2] and not flags[3]
    for k in (0, 2, 3):
        s = singles[k]
        s.prepare()
        s.step(3)
        for n in _names(s.params):
            np.testing.assert_array_equal(b.get(k, n), s.get(n), err_msg=f"sim {k} {n}")
    held = {n: b.get(1, n) for n in ("x", "y", "vx", "vy", "cxy")}
    b.step(4)
    assert b.failed()[1] & 1
    for n, v in held.items():
        np.testing.assert_array_equal(b.get(1, n), v, err_msg=f"held sim {n}")
    for k in (0, 2, 3):
        s = singles[k]
        s.step(4)
        for n in ("x", "vx", "fx"):
            np.testing.assert_array_equal(b.get(k, n), s.get(n), err_msg=f"sim {k} {n} later")
