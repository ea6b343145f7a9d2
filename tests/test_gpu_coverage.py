"""GPU paths compared with the FP64 oracle beyond the basic parity cases.

  * every member of a batch against its own oracle run (each simulation of a
    batch is the reference's loop src/main.cu:358-459 on its own inputs):
    FP64 form, fields 1e-10 normwise and positions 1e-12 absolute after 7 steps,
    the tolerances of a stand-alone run;
  * the external force types 1, 2 and 3 (src/kernel_ext_force.cu:22-72) through
    prepare and 5 steps, and through the stand-alone calculate_external_forces;
  * simulations with more colloids than the pair pass stages in LDS (40 and
    the reference's maximum, 1000 > 32: the colloid states are read from
    global memory; all-pairs colloid repulsion and magnetic forces), FP64 and
    FP32 forms;
  * the cell-offset scan through k_scan_tile_sums + k_scan_add (forced at small
    grids by SPH_SCAN_SUM_TILES=1): the cell list bit-exact against the oracle
    and every field bitwise equal to the default path, single and batched;
  * calculate_vel_gradient after calculate_forces when every neighbour list
    overflows (capacity 8): the forces stay those of the force pass and the
    gradient is the oracle's, with the overflow queue emptied before each pass.
"""
import functools

import numpy as np
import pytest

import helpers
import inputs
import oracle
import sph_system
from helpers import comps, relerr

pytestmark = pytest.mark.gpu


def _check_state(s, o, tol=1e-10, xtol=1e-12, label=""):
    p = s.params
    d = p.dim
    for a, n in enumerate(comps("x", d)):
        diff = helpers.periodic_diff(s.get(n), o.get(n), p.L[a])
        assert np.max(np.abs(diff)) <= xtol, f"{label} {n}"
    for n in comps("v", d) + comps("c", d) + comps("f", d) + comps("dcdt", d) + ["dens", "press"]:
        e = relerr(s.get(n), o.get(n))
        assert e <= tol, f"{label} {n}: {e:.3e}"
    if p.N_colloids:
        for n in ("coll_x", "coll_y", "coll_vx", "coll_vy", "coll_omegaz", "fx_colloid",
                  "fy_colloid", "tz_colloid"):
            e = relerr(s.get(n), o.get(n))
            assert e <= 1e-9, f"{label} {n}: {e:.3e}"
    if p.wall:
        for n in ("fx_wall", "fy_wall"):
            e = relerr(s.get(n), o.get(n))
            assert e <= 1e-9, f"{label} {n}: {e:.3e}"


class _Member:
    """A batch member seen through the System-like get() the checks use."""

    def __init__(self, batch, k, params):
        self.b, self.k, self.params = batch, k, params

    def get(self, name):
        return self.b.get(self.k, name)


@pytest.mark.parametrize("kind", ["2d", "3d_walls", "3d_periodic"])
def test_batch_members_match_the_oracle(kind, tmp_path):
    import test_gpu_batch as tb

    members = tb._systems(kind, str(tmp_path), sph_system.FP64, "m")
    oracles = [oracle.Oracle.from_system(s) for s in members]
    b = sph_system.Batch(sph_system.FP64)
    for s in members:
        b.add(s)
    b.prepare()
    for o in oracles:
        o.prepare()
    for k, (s, o) in enumerate(zip(members, oracles)):
        _check_state(_Member(b, k, s.params), o, tol=1e-12, label=f"sim {k} prepare")
    b.step(7)
    for o in oracles:
        o.step(7)
    b.synchronize()
    ke = b.kinetic_energy()
    for k, (s, o) in enumerate(zip(members, oracles)):
        _check_state(_Member(b, k, s.params), o, label=f"sim {k} step 7")
        assert abs(ke[k] - o.kinetic_energy()) <= 1e-12 * abs(o.kinetic_energy())
    b.close()


EXT_CASES = {
    "type1_3d_periodic": lambda: helpers.cfg_sine(n=16, ext_force_type=1,
                                                  ext_force=[0.7, -0.3, 0.4]),
    "type1_2d_walls": lambda: helpers.cfg_couette(n=40, ext_force_type=1,
                                                  ext_force=[0.5, 0.2]),
    "type2_2d_walls": lambda: helpers.cfg_couette(n=40, ext_force_type=2,
                                                  ext_force=[0.8, 0.0]),
    "type3_2d_walls": lambda: helpers.cfg_couette(n=40, ext_force_type=3,
                                                  ext_force=[0.6, 0.0], Vwall_bottom=0.0,
                                                  Vwall_top=0.0),
    "type3_3d_walls": lambda: helpers.cfg_colloids_3d(n=20, n_colloids=0, ext_force_type=3,
                                                      ext_force=[0.6, 0.0, 0.0]),
}


@pytest.mark.parametrize("case", list(EXT_CASES))
def test_external_force_types(case, tmp_path):
    cfg = EXT_CASES[case]()
    s = helpers.make_system(cfg, str(tmp_path))
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    s.prepare()
    o.prepare()
    d = s.params.dim
    for n in comps("f", d):
        assert relerr(s.get(n), o.get(n)) <= 1e-12, n
    s.step(5)
    o.step(5)
    s.synchronize()
    _check_state(s, o, label=case)
    # the stand-alone method adds the body force once more
    f0 = s.get("fx")
    s.calculate_external_forces()
    o.calculate_external_forces()
    for n in comps("f", d):
        assert relerr(s.get(n), o.get(n)) <= 1e-12, n
    assert np.abs(s.get("fx") - f0).max() > 0


# (count, R, lattice n, dt, magnetic cut-off, F0_magnetic): 40 colloids, more than the
# pair pass stages in LDS; 1000, the reference's MAX_COLLOIDS (src/config.h:22),
# ~3 lattice spacings in radius on a 300 x 300 lattice (88,593 particles)
MANY_COLLOIDS = {40: (0.03, 96, 2e-5, 0.1, 1e-4), 1000: (0.0105, 300, 5e-6, 0.035, 1e-7)}


@functools.lru_cache(maxsize=None)
def _colloid_centres(count):
    R = MANY_COLLOIDS[count][0]
    return tuple(map(tuple, inputs.place_colloids(count, R, [1.0, 1.0], 2, True, seed=9,
                                                  wall_gap=3.0)))


def _many_colloids_cfg(count=40):
    R, n, dt, rmag, f0 = MANY_COLLOIDS[count]
    cfg = helpers.cfg_couette(n=n, dt=dt, Vwall_bottom=-0.2, Vwall_top=0.2)
    cfg.update(N_colloids=count, coll_R=R, coll_rho=1.0, coll_move=0,
               coll_x=[list(c) for c in _colloid_centres(count)],
               coll_repulsion_cuton=0.0, F0_repulsion=1.0, tau_repulsion=30.0,
               cutoff_magnetic=rmag, F0_magnetic=f0, omega_magnetic=3.0)
    return cfg


@pytest.mark.parametrize("count", [40, 1000])
@pytest.mark.parametrize("precision", [sph_system.FP64, sph_system.FP32])
def test_more_colloids_than_the_lds_stage(precision, count, tmp_path):
    s = helpers.make_system(_many_colloids_cfg(count), str(tmp_path), precision)
    assert s.params.N_colloids == count
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    s.prepare()
    o.prepare()
    d = s.params.dim
    if precision == sph_system.FP64:
        for n in ["dens"] + comps("gradv", d) + comps("f", d) + comps("dcdt", d):
            assert relerr(s.get(n), o.get(n)) <= 1e-12, n
        s.step(5)
        o.step(5)
        s.synchronize()
        _check_state(s, o, label=f"{count} colloids")
    else:
        assert relerr(s.get("dens"), o.get("dens")) <= 1e-6
        for n in comps("gradv", d) + comps("dcdt", d):
            assert relerr(s.get(n), o.get(n)) <= 1e-4, n
        for n in comps("f", d):
            assert relerr(s.get(n), o.get(n)) <= 1e-3, n
        for n, scale in helpers.sum_scales(o).items():
            if n.startswith(("fz", "tx", "ty")):
                continue
            assert np.all(np.abs(s.get(n) - o.get(n)) <= 1e-5 * scale + 1e-300), n


def _cells_equal(s, o):
    pc, pi, cs, ce = s.cell_list()
    opc, opi, ocs, oce = o.cell_list()
    np.testing.assert_array_equal(pc, opc)
    np.testing.assert_array_equal(pi, opi)
    occ = oce >= ocs
    np.testing.assert_array_equal(cs[occ], ocs[occ])
    np.testing.assert_array_equal(ce[occ], oce[occ])
    assert np.all(ce[~occ] < cs[~occ])


def test_multi_launch_scan_single(tmp_path, monkeypatch):
    """64^3 periodic flow: 21^3 = 9261 cells, two 8192-cell scan tiles."""
    outs = []
    for forced in (False, True):
        if forced:
            monkeypatch.setenv("SPH_SCAN_SUM_TILES", "1")
        s = helpers.make_system(helpers.cfg_sine(n=64), str(tmp_path), name=f"s{forced}")
        assert s.params.Ntotal_cells > 8192
        helpers.perturb(s)
        o = oracle.Oracle.from_system(s)
        s.neighbours_search()
        o.neighbours_search()
        _cells_equal(s, o)
        s.prepare()
        s.step(3)
        s.synchronize()
        outs.append({n: s.get(n) for n in ("x", "vx", "fx", "dens", "cxy")})
        s.close()
    for n in outs[0]:
        np.testing.assert_array_equal(outs[0][n], outs[1][n], err_msg=n)


def test_multi_launch_scan_batch(tmp_path, monkeypatch):
    """Twelve 2D wall-bounded flows of 100-160 sites per side: 24,142 cells,
    three scan tiles."""
    cfgs = [helpers.cfg_couette(n=n, tau=0.5 + 0.1 * k)
            for k, n in enumerate((100, 160, 128, 110, 150, 140, 120, 160, 105, 130, 155, 115))]
    outs = []
    for forced in (False, True):
        if forced:
            monkeypatch.setenv("SPH_SCAN_SUM_TILES", "1")
        b = sph_system.Batch(sph_system.FP32)
        systems = []
        for k, cfg in enumerate(cfgs):
            s = helpers.make_system(cfg, str(tmp_path), sph_system.FP32, seed=k + 1,
                                    name=f"b{forced}{k}")
            systems.append(s)
            b.add(s)
        assert sum(s.params.Ntotal_cells for s in systems) > 2 * 8192
        b.prepare()
        b.step(5)
        b.synchronize()
        outs.append([[b.get(k, n) for n in ("x", "y", "vx", "fy", "dens")]
                     for k in range(len(cfgs))])
        b.close()
        for s in systems:
            s.close()
    for k in range(len(cfgs)):
        for a, c in zip(outs[0][k], outs[1][k]):
            np.testing.assert_array_equal(a, c, err_msg=f"sim {k}")


def test_gradient_after_forces_with_overflowing_lists(tmp_path, monkeypatch):
    monkeypatch.setenv("SPH_NEIGHBOUR_CAPACITY", "8")
    s = helpers.make_system(helpers.cfg_colloids_3d(n=24), str(tmp_path))
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    d = s.params.dim
    for obj in (s, o):
        obj.neighbours_search()
        obj.set_step(2)
        obj.calculate_forces()
    f_forces = {n: s.get(n) for n in comps("f", d)}
    for n in comps("f", d) + comps("gradv", d):
        assert relerr(s.get(n), o.get(n)) <= 1e-12, n
    for rep in range(2):
        s.calculate_vel_gradient()
        o.calculate_vel_gradient()
        for n in comps("gradv", d):
            assert relerr(s.get(n), o.get(n)) <= 1e-12, f"pass {rep} {n}"
        for n in comps("f", d):
            np.testing.assert_array_equal(s.get(n), f_forces[n], err_msg=f"pass {rep} {n}")
