"""Parity of the HIP path (through the C-ABI) with the FP64 oracle.

States are "generic" (fluid positions jittered by 0.1 lattice spacing, random
symmetric perturbation of c, velocity noise) unless the test says "lattice".

Tolerances (normwise relative error ||gpu - oracle|| / ||oracle|| unless noted):
  cell list, sorted order, cell ranges           bit-exact (both precision forms)
  FP64 form, generic state: dens, gradv, press, F, dc/dt  1e-12
  wall and colloid force/torque sums: |gpu - oracle| <= 1e-13 (FP64) or 1e-5
      (FP32) times the sum of |terms| (these sums cancel strongly)
  FP64 form, positions after 5 steps              1e-12 absolute (box length 1)
  FP64 form, perfect-lattice start: F 1e-10 (the net force there is ~1e-3 of
      the individual pair terms, so summation-order rounding is amplified)
  FP32 form (pair arithmetic in float), generic state: dens 1e-6, gradv and
      dc/dt 1e-4, F 1e-3; positions 1e-9 absolute
  FP32 form, perfect-lattice start: F 1e-2
"""
import numpy as np
import pytest

import helpers
from helpers import comps, relerr
import oracle
import sph_system

pytestmark = pytest.mark.gpu

CASES = {
    "couette2d": lambda: helpers.cfg_couette(n=64),
    "sine3d": lambda: helpers.cfg_sine(n=24),
    "colloids3d": lambda: helpers.cfg_colloids_3d(n=30),
    "colloids2d": lambda: helpers.cfg_colloids_2d(n=48),
}


def _pair(case, tmpdir, precision=sph_system.FP64, generic=True):
    s = helpers.make_system(CASES[case](), tmpdir, precision)
    if generic:
        helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    return s, o


def _check_fields(s, o, tol, names):
    for n in names:
        e = relerr(s.get(n), o.get(n))
        assert e <= tol, f"{n}: rel err {e:.3e} > {tol:.1e}"


def _check_sums(s, o, rel):
    """Wall and colloid force/torque sums: |gpu - oracle| <= rel * sum|terms|."""
    for n, scale in helpers.sum_scales(o).items():
        if n.startswith(("fz", "tx", "ty")) and s.params.dim == 2:
            continue
        d = np.abs(s.get(n) - o.get(n))
        assert np.all(d <= rel * scale + 1e-300), f"{n}: {d} vs {rel} * {scale}"


def _check_cells(s, o):
    pc, pi, cs, ce = s.cell_list()
    opc, opi, ocs, oce = o.cell_list()
    np.testing.assert_array_equal(pc, opc)
    np.testing.assert_array_equal(pi, opi)
    occupied = oce >= ocs
    np.testing.assert_array_equal(cs[occupied], ocs[occupied])
    np.testing.assert_array_equal(ce[occupied], oce[occupied])
    assert np.all(ce[~occupied] < cs[~occupied])


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", [sph_system.FP64, sph_system.FP32])
def test_cell_list_bit_exact(case, precision, tmp_path):
    s, o = _pair(case, str(tmp_path), precision)
    p = s.params
    rng = np.random.default_rng(7)
    # Shuffle particles across cells, and put some exactly on cell faces and box faces.
    L = np.array(p.L)
    for a, n in enumerate(["x", "y", "z"][:p.dim]):
        v = s.get(n)
        t = s.get("type")
        move = t == 0
        v[move] = v[move] + rng.uniform(-0.3, 0.3, move.sum()) * p.rcut
        if not (a == 1 and p.wall):
            v = np.mod(v, L[a])
            k = rng.choice(np.flatnonzero(move), 8, replace=False)
            v[k[:3]] = 0.0
            v[k[3:5]] = L[a]
            v[k[5:]] = p.cell_size[a] * rng.integers(1, p.Ncells[a], 3)
        s.set(n, v)
        o.set(n, v)
    s.neighbours_search()
    o.neighbours_search()
    _check_cells(s, o)


@pytest.mark.parametrize("case", list(CASES))
def test_prepare_fp64(case, tmp_path):
    s, o = _pair(case, str(tmp_path))
    s.prepare()
    o.prepare()
    _check_cells(s, o)
    d = s.params.dim
    names = (["dens", "press"] + comps("gradv", d) + comps("f", d) + comps("dcdt", d))
    for n in names:
        e = relerr(s.get(n), o.get(n))
        assert e <= 1e-12, f"{case} {n}: {e:.3e}"
    _check_sums(s, o, 1e-13)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision,tol", [(sph_system.FP64, 1e-10), (sph_system.FP32, 1e-2)])
def test_lattice_start_forces(case, precision, tol, tmp_path):
    s, o = _pair(case, str(tmp_path), precision, generic=False)
    s.prepare()
    o.prepare()
    d = s.params.dim
    _check_fields(s, o, tol, comps("f", d))


@pytest.mark.parametrize("case", list(CASES))
def test_steps_fp64(case, tmp_path):
    s, o = _pair(case, str(tmp_path))
    s.prepare()
    o.prepare()
    nsteps = 5
    s.step(nsteps)
    o.step(nsteps)
    s.synchronize()
    p = s.params
    d = p.dim
    for a, n in enumerate(comps("x", d)):
        diff = helpers.periodic_diff(s.get(n), o.get(n), p.L[a])
        assert np.max(np.abs(diff)) <= 1e-12, n
    for grp in ("v", "c", "f", "dcdt"):
        for n in comps(grp, d):
            e = relerr(s.get(n), o.get(n))
            assert e <= 1e-10, f"{case} {n}: {e:.3e}"
    _check_fields(s, o, 1e-10, ["dens", "press"])
    ke, oke = s.kinetic_energy(), o.kinetic_energy()
    assert abs(ke - oke) <= 1e-12 * abs(oke)
    if p.N_colloids:
        _check_fields(s, o, 1e-9, ["coll_x", "coll_y", "coll_vx", "coll_vy", "coll_omegaz",
                                   "fx_colloid", "fy_colloid", "tz_colloid"])


@pytest.mark.parametrize("case", list(CASES))
def test_fp32_form(case, tmp_path):
    s, o = _pair(case, str(tmp_path), sph_system.FP32)
    s.prepare()
    o.prepare()
    _check_cells(s, o)
    d = s.params.dim
    _check_fields(s, o, 1e-6, ["dens"])
    _check_fields(s, o, 1e-4, comps("gradv", d) + comps("dcdt", d))
    _check_fields(s, o, 1e-3, comps("f", d))
    _check_sums(s, o, 1e-5)
    s.step(1)
    o.step(1)
    s.synchronize()
    p = s.params
    for a, n in enumerate(comps("x", d)):
        diff = helpers.periodic_diff(s.get(n), o.get(n), p.L[a])
        assert np.max(np.abs(diff)) <= 1e-9, n
    _check_fields(s, o, 1e-3, comps("f", d))


@pytest.mark.parametrize("precision", [sph_system.FP64, sph_system.FP32])
def test_bitwise_repeatable(precision, tmp_path):
    outs = []
    for rep in range(2):
        s = helpers.make_system(helpers.cfg_colloids_3d(n=24), str(tmp_path), precision,
                                name=f"in{rep}")
        s.prepare()
        s.step(3)
        s.synchronize()
        outs.append({n: s.get(n) for n in ["x", "vx", "fx", "dens", "fx_colloid", "fx_wall"]}
                    | {"ke": np.array([s.kinetic_energy()])})
    for n in outs[0]:
        np.testing.assert_array_equal(outs[0][n], outs[1][n], err_msg=n)


def test_method_sequence_matches_step(tmp_path):
    """VV1, neighbours_search, set_step, calculate_forces, VV2 called one by one
    against sph_step, whose steps replay a captured graph, use Verlet lists
    (same pairs, list order instead of cell order) and apply VV2's fluid kick
    inside the fused pair pass: FP64 agreement to 1e-12 after every replay, with
    host reads in between."""
    a = helpers.make_system(helpers.cfg_colloids_2d(), str(tmp_path), name="a")
    b = helpers.make_system(helpers.cfg_colloids_2d(), str(tmp_path), name="b")
    a.prepare()
    b.prepare()

    def same(step):
        for n in ["x", "y"]:
            assert np.abs(a.get(n) - b.get(n)).max() <= 1e-13, f"{n} at step {step}"
        for n in ["vx", "vy", "cxy", "fx", "dens", "coll_x", "tz_colloid"]:
            assert relerr(a.get(n), b.get(n)) <= 1e-12, f"{n} at step {step}"

    for step in (1, 2, 3, 4, 5):
        a.V_Verlet_step1()
        a.neighbours_search()
        a.set_step(step)
        a.calculate_forces()
        a.V_Verlet_step2()
        b.step(1)
        b.synchronize()
        same(step)
    for step in (6, 7, 8):   # several steps in one call
        a.V_Verlet_step1()
        a.neighbours_search()
        a.set_step(step)
        a.calculate_forces()
        a.V_Verlet_step2()
    b.step(3)
    same(8)


@pytest.mark.parametrize("skin", ["0.005", "0.1"])
@pytest.mark.parametrize("case", ["couette2d", "colloids3d"])
def test_verlet_steps_fp64(case, skin, tmp_path, monkeypatch):
    """40 steps with Verlet lists (a skin of 0.005 rcut rebuilds every few
    steps, 0.1 rcut rarely) against the oracle, which re-sorts every step."""
    monkeypatch.setenv("SPH_VERLET_SKIN", skin)
    cfg = {"couette2d": lambda: helpers.cfg_couette(n=48, Vwall_bottom=-2.0, Vwall_top=2.0),
           "colloids3d": lambda: helpers.cfg_colloids_3d(n=24)}[case]()
    s = helpers.make_system(cfg, str(tmp_path), noise=1e-2)
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    s.prepare()
    o.prepare()
    s.step(40)
    o.step(40)
    s.synchronize()
    p = s.params
    d = p.dim
    for a, n in enumerate(comps("x", d)):
        diff = helpers.periodic_diff(s.get(n), o.get(n), p.L[a])
        assert np.max(np.abs(diff)) <= 1e-12, n
    for n in comps("v", d) + comps("c", d) + comps("f", d) + ["dens"]:
        assert relerr(s.get(n), o.get(n)) <= 1e-10, n
    _check_cells(s, o)     # the cell list of the current positions


def test_verlet_list_overflow(tmp_path, monkeypatch):
    """Verlet lists of capacity 8 overflow everywhere: those particles fall back
    to the cell scan, and their simulation is rebuilt every step."""
    monkeypatch.setenv("SPH_VERLET_CAPACITY", "8")
    s, o = _pair("colloids3d", str(tmp_path))
    s.prepare()
    o.prepare()
    s.step(6)
    o.step(6)
    s.synchronize()
    d = s.params.dim
    for n in comps("v", d) + comps("f", d) + ["dens"]:
        assert relerr(s.get(n), o.get(n)) <= 1e-10, n


def test_neighbour_search_keeps_computed_fields(tmp_path):
    """A stand-alone neighbours_search after particles moved reorders the state;
    fields computed before it (forces, density, dc/dt) still belong to their
    particles (the reference never permutes its arrays)."""
    s, _ = _pair("sine3d", str(tmp_path))
    s.prepare()
    before = {n: s.get(n) for n in ("fx", "dens", "dcdt_xy", "press")}
    x = s.get("x")
    x = np.mod(x + 0.37, 1.0)          # move every particle across cells
    s.set("x", x)
    s.neighbours_search()
    for n, v in before.items():
        np.testing.assert_array_equal(s.get(n), v, err_msg=n)


def test_individual_methods_fp64(tmp_path):
    """The class_system methods one at a time, each against the oracle."""
    s, o = _pair("colloids3d", str(tmp_path))
    d = s.params.dim
    for obj in (s, o):
        obj.neighbours_search()
        obj.calculate_density()
    _check_fields(s, o, 1e-12, ["dens"])
    for obj in (s, o):
        obj.calculate_vel_gradient()
        obj.calculate_pressure()
    _check_fields(s, o, 1e-12, comps("gradv", d) + ["press"])
    for obj in (s, o):
        obj.set_step(3)
        obj.calculate_forces()
    _check_fields(s, o, 1e-12, comps("f", d) + comps("dcdt", d))
    _check_sums(s, o, 1e-13)
    for obj in (s, o):
        obj.V_Verlet_step2()
    _check_fields(s, o, 1e-12, comps("v", d) + comps("c", d))
    _check_fields(s, o, 1e-12, ["coll_vx", "coll_vy", "coll_vz", "coll_omegax", "coll_omegaz"])
    for obj in (s, o):
        obj.V_Verlet_step1()
    p = s.params
    for a, n in enumerate(comps("x", d)):
        diff = helpers.periodic_diff(s.get(n), o.get(n), p.L[a])
        assert np.max(np.abs(diff)) <= 1e-13, n
    _check_fields(s, o, 1e-12, ["coll_quat0", "coll_quatx", "coll_quaty", "coll_quatz"])


def test_external_force_method(tmp_path):
    s, o = _pair("sine3d", str(tmp_path))
    for obj in (s, o):
        obj.neighbours_search()
        obj.calculate_forces()
    f0 = s.get("fx")
    s.calculate_external_forces()   # adds the body force a second time
    o.calculate_external_forces()
    _check_fields(s, o, 1e-12, ["fx"])
    assert np.abs(s.get("fx") - f0).max() > 0


def test_out_of_grid_particle_is_reported(tmp_path):
    """A particle moved beyond a wall band raises SPH_ERR_OUT_OF_GRID (the
    reference's device assert), with every kernel still exiting normally."""
    s, _ = _pair("couette2d", str(tmp_path))
    s.prepare()
    y = s.get("y")
    y[5] = 3.0
    s.set("y", y)
    with pytest.raises(sph_system.SphError) as e:
        s.neighbours_search()
    assert e.value.code == 6


def test_small_neighbour_list_capacity_gives_identical_results(tmp_path, monkeypatch):
    """With a list capacity of 8 every particle overflows: the density pass adds
    the hits past the capacity during its scan and the pair pass re-scans the
    strips for them in the same order.  The pair sums are then bitwise those of
    the default capacity; the density differs only by the association of its
    sum (FP64: <= 1e-12 normwise after two steps, the parity bar)."""
    outs = []
    names = ("dens", "fx", "fy", "fz", "dcdt_xy", "vx", "tz_colloid")
    for cap in (None, "8"):
        if cap:
            monkeypatch.setenv("SPH_NEIGHBOUR_CAPACITY", cap)
        s = helpers.make_system(helpers.cfg_colloids_3d(n=24), str(tmp_path), name=f"i{cap}")
        helpers.perturb(s)
        s.prepare()
        s.step(2)
        s.synchronize()
        outs.append([s.get(n) for n in names])
        s.close()
    for n, a, b in zip(names, *outs):
        assert relerr(a, b) <= 1e-12, n


@pytest.mark.parametrize("precision", [sph_system.FP64, sph_system.FP32])
def test_colloid_states_from_lds_or_global_memory(precision, tmp_path, monkeypatch):
    """The pair pass reads the colloid states that the Morris velocities of
    fluid-colloid pairs need from an LDS copy (at most 32 colloids per
    simulation) or from global memory (more colloids; SPH_COLLOID_LDS=0 forces
    it): the same values either way, so the trajectories are bitwise equal."""
    outs = []
    names = ("x", "vx", "vy", "fx", "dcdt_xy", "fx_colloid", "tz_colloid")
    for lds in ("1", "0"):
        monkeypatch.setenv("SPH_COLLOID_LDS", lds)
        s = helpers.make_system(helpers.cfg_colloids_3d(n=24), str(tmp_path), precision,
                                name=f"l{lds}")
        helpers.perturb(s)
        s.prepare()
        s.step(3)
        s.synchronize()
        outs.append([s.get(n) for n in names])
        s.close()
    for n, a, b in zip(names, *outs):
        np.testing.assert_array_equal(a, b, err_msg=n)


@pytest.mark.parametrize("case", ["colloids3d", "colloids2d", "sine3d"])
def test_observables(case, tmp_path):
    s, o = _pair(case, str(tmp_path))
    s.prepare()
    o.prepare()
    s.step(3)
    o.step(3)
    a = np.array(list(s.observables().values()))
    b = o.observables()
    both_nan = np.isnan(a) & np.isnan(b)
    np.testing.assert_allclose(a[~both_nan], b[~both_nan], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("skin", ["0.01", "0.5"])
@pytest.mark.parametrize("precision", [sph_system.FP32, sph_system.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["colloids3d", "sine3d", "couette2d"])
def test_staged_density_walk_is_the_unstaged_one(case, precision, skin, tmp_path, monkeypatch):
    """The density walk reads neighbour positions from a per-tile LDS
    window (SPH_DENSITY_STAGE=2: in 3D, and in 2D for both forms; =0 reads
    them from memory; SPH_STAGE_CAP=400 leaves the tiles with larger windows
    to the unstaged walk).  The same positions in the same order: bitwise the
    same trajectory, with list rebuilds inside the run.  A skin of 0.01 rcut
    rebuilds every few steps; one of 0.5 rcut lists many entries from the
    second ring of cells, outside their tile's window (read from memory)."""
    monkeypatch.setenv("SPH_VERLET_SKIN", skin)
    cfg = {"couette2d": lambda: helpers.cfg_couette(n=48, Vwall_bottom=-2.0, Vwall_top=2.0),
           "colloids3d": lambda: helpers.cfg_colloids_3d(n=24),
           "sine3d": lambda: helpers.cfg_sine(n=24)}[case]()
    names = ("x", "y", "vx", "vy", "dens", "fx", "fy", "cxy")
    outs = []
    for mode in ({"SPH_DENSITY_STAGE": "2"}, {"SPH_DENSITY_STAGE": "0"},
                 {"SPH_DENSITY_STAGE": "2", "SPH_STAGE_CAP": "400"}):
        for k in ("SPH_DENSITY_STAGE", "SPH_STAGE_CAP"):
            monkeypatch.delenv(k, raising=False)
        for k, v in mode.items():
            monkeypatch.setenv(k, v)
        s = helpers.make_system(cfg, str(tmp_path), precision, noise=1e-2,
                                name=f"m{len(outs)}")
        helpers.perturb(s)
        s.prepare()
        s.step(12)
        s.synchronize()
        outs.append([s.get(n) for n in names])
        s.close()
    for ref in outs[1:]:
        for n, a, b in zip(names, outs[0], ref):
            np.testing.assert_array_equal(a, b, err_msg=n)


@pytest.mark.parametrize("precision", [sph_system.FP32, sph_system.FP64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("case", ["colloids3d", "sine3d"])
def test_tile_order_leaves_the_trajectory_alone(case, precision, tmp_path, monkeypatch):
    """A single 3D simulation walks each XCD's slab of tiles in y bands
    (SPH_TILE_ORDER=k bands, 0: layer after layer); tiles are independent, so
    every order gives bitwise the same trajectory, list rebuilds (a skin of
    0.01 rcut) included."""
    monkeypatch.setenv("SPH_VERLET_SKIN", "0.01")
    cfg = {"colloids3d": lambda: helpers.cfg_colloids_3d(n=24),
           "sine3d": lambda: helpers.cfg_sine(n=24)}[case]()
    names = ("x", "y", "z", "vx", "vy", "vz", "dens", "fx", "fy", "fz", "cxy", "czz")
    outs = []
    for bands in ("0", "4", "3"):
        monkeypatch.setenv("SPH_TILE_ORDER", bands)
        s = helpers.make_system(cfg, str(tmp_path), precision, noise=1e-2,
                                name=f"o{len(outs)}")
        helpers.perturb(s)
        s.prepare()
        s.step(12)
        s.synchronize()
        outs.append([s.get(n) for n in names])
        s.close()
    for ref in outs[1:]:
        for n, a, b in zip(names, outs[0], ref):
            np.testing.assert_array_equal(a, b, err_msg=n)


@pytest.mark.parametrize("precision", [sph_system.FP64, sph_system.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [9, 12, 16], ids=["3cells", "4cells", "5cells"])
def test_smallest_periodic_boxes(n, precision, tmp_path):
    """Fully periodic 3D boxes of 3, 4 and 5 cells per axis (for n = 9 the cell
    is exactly rcut wide).  Below 5 cells the steps search the cells (the cell
    scan visits every cell once even where both neighbours of a cell are the
    same image); 5 cells is the smallest box with Verlet lists, whose reach of
    two cells on each side then covers the axis exactly.  A cell scanned twice
    would count its pairs twice: density errors of a few per cent.  Generic
    state, 6 steps against the oracle: FP64 positions 1e-12 absolute, fields
    1e-10; FP32 density 1e-6 at the first step and 1e-5 after the steps,
    positions 1e-8 absolute."""
    s = helpers.make_system(helpers.cfg_sine(n=n), str(tmp_path), precision, noise=1e-2)
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    p = s.params
    assert list(p.Ncells) == [{9: 3, 12: 4, 16: 5}[n]] * 3
    s.prepare()
    o.prepare()
    fp64 = precision == sph_system.FP64
    _check_fields(s, o, 1e-12 if fp64 else 1e-6, ["dens"])
    s.step(6)
    o.step(6)
    s.synchronize()
    for a, c in enumerate(comps("x", 3)):
        diff = helpers.periodic_diff(s.get(c), o.get(c), p.L[a])
        assert np.max(np.abs(diff)) <= (1e-12 if fp64 else 1e-8), c
    _check_fields(s, o, 1e-10 if fp64 else 1e-5, ["dens"])
    _check_fields(s, o, 1e-10 if fp64 else 1e-3, comps("v", 3) + comps("f", 3))
    _check_cells(s, o)
