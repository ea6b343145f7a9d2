"""read_input / initialize / writers: closed forms and the reference's error paths
(src/system_read_input.cpp:271-507, src/system_initialize.cpp:128-261)."""
import math
import os

import numpy as np
import pytest

import helpers
import inputs
import sph_system
from sph_system import SphError


def _read(cfg, tmp_path, name="input"):
    path = os.path.join(str(tmp_path), name)
    inputs.write_input(path, cfg)
    s = sph_system.System("fp64")
    s.read_input(path)
    return s


def test_couette_2d_constants_and_lattice(tmp_path):
    s = _read(helpers.cfg_couette(n=64), tmp_path)
    s.initialize()
    p = s.params
    dx = 1.0 / 64
    assert (p.dim, p.wall, p.Nwalls) == (2, 1, 2)
    assert list(p.Nxyz[:2]) == [64, 70]          # + 2 ceil(overlap) wall layers
    assert p.N == p.Nmax == 64 * 70
    assert p.rcut == pytest.approx(3 * dx, rel=1e-15)
    assert p.wall_width == pytest.approx(3 * dx, rel=1e-15)
    assert p.cw == pytest.approx(63.0 / (478.0 * math.pi * p.rcut ** 2), rel=1e-15)
    assert p.c_gradw == pytest.approx(-15.0 * p.cw / p.rcut, rel=1e-15)
    assert (p.a, p.b) == pytest.approx((5.0 / 3.0, 4.0 / 3.0))   # 2D b = 4(zeta + eta/3)
    assert list(p.Ncells) == [21, 23, 1]
    assert p.cell_size[1] == pytest.approx((1 + 6 * dx) / 23, rel=1e-15)
    assert (p.y_bottom, p.y_top, p.beta_max) == (0.0, 1.0, 1.5)
    t = s.get("type")
    assert np.bincount(t).tolist() == [4096, 192, 192]
    x, y, m = s.get("x"), s.get("y"), s.get("mass")
    i, j = np.divmod(np.arange(p.N), 70)
    np.testing.assert_allclose(x, dx / 2 + i * dx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(y, dx / 2 + j * dx - 3 * dx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(m, (1 + 6 * dx) / (64 * 70), rtol=1e-15)
    assert np.all(t[y < 0] == 1) and np.all(t[y > 1] == 2) and np.all(t[(y > 0) & (y < 1)] == 0)
    np.testing.assert_array_equal(s.get("cxx"), 1.0)
    np.testing.assert_array_equal(s.get("cxy"), 0.0)
    wl, ws = s.get("walls_list"), s.get("walls_start")
    assert ws.tolist() == [0, 192]
    assert np.all(t[wl[:192]] == 1) and np.all(t[wl[192:]] == 2)
    assert np.all(np.diff(wl[:192]) > 0)


def test_sine_3d_constants(tmp_path):
    s = _read(helpers.cfg_sine(n=20), tmp_path)
    s.initialize()
    p = s.params
    assert p.N == 8000 and p.wall == 0
    assert p.b == pytest.approx(5.0 / 3.0)      # 3D b = 5(zeta + eta/3)
    assert p.cw == pytest.approx(27.0 / (120.0 * math.pi * p.rcut ** 3), rel=1e-15)
    assert p.ky == pytest.approx(2 * math.pi)
    # the reference's mass keeps 2*wall_width even without walls
    assert s.get("mass")[0] == pytest.approx((1 + 2 * 3 / 20) / 8000, rel=1e-15)


def test_colloid_shells_2d(tmp_path):
    cfg = helpers.cfg_colloids_2d(n=48)
    s = _read(cfg, tmp_path)
    s.initialize()
    p = s.params
    dx, R, ww = 1 / 48, cfg["coll_R"], p.wall_width
    x, y, t = s.get("x"), s.get("y"), s.get("type")
    for k, c in enumerate(cfg["coll_x"]):
        shell = t == 3 + k
        r = np.hypot(x[shell] - c[0], y[shell] - c[1])
        radii = [0.5 * dx + j * dx for j in range(round(R / dx)) if R - ww < 0.5 * dx + j * dx < R]
        expect = sum(round(2 * math.pi * rr / dx) for rr in radii)
        assert shell.sum() == expect
        assert np.all(np.min(np.abs(r[:, None] - np.array(radii)[None, :]), axis=1) < 1e-12)
        fluid = t == 0
        assert np.all(np.hypot(x[fluid] - c[0], y[fluid] - c[1]) > R)
    # colloid particles: mass, velocity, conformation are zero
    sh = t >= 3
    assert np.all(s.get("mass")[sh] == 0) and np.all(s.get("cxx")[sh] == 0)
    lst, st = s.get("colloids_list"), s.get("colloids_start")
    assert st.tolist() == [0, int((t == 3).sum())]
    xc, yc = s.get("x_center"), s.get("y_center")
    own = t[lst] - 3
    np.testing.assert_allclose(np.hypot(xc, yc), np.hypot(x[lst] - np.array(cfg["coll_x"])[own, 0],
                                                           y[lst] - np.array(cfg["coll_x"])[own, 1]),
                               atol=1e-14)


def test_colloid_shells_3d(tmp_path):
    cfg = helpers.cfg_colloids_3d(n=30)
    s = _read(cfg, tmp_path)
    s.initialize()
    p = s.params
    dx, R, ww = 1 / 30, cfg["coll_R"], p.wall_width
    x, y, z, t = s.get("x"), s.get("y"), s.get("z"), s.get("type")
    radii = np.array([0.5 * dx + j * dx for j in range(round(R / dx))
                      if R - ww < 0.5 * dx + j * dx < R])
    for k, c in enumerate(cfg["coll_x"]):
        sh = t == 3 + k
        r = np.sqrt((x[sh] - c[0]) ** 2 + (y[sh] - c[1]) ** 2 + (z[sh] - c[2]) ** 2)
        assert np.all(np.min(np.abs(r[:, None] - radii[None, :]), axis=1) < 1e-12)
        area = sum(4 * math.pi * rr ** 2 / dx ** 2 for rr in radii)
        assert abs(sh.sum() - area) < 0.15 * area
        # poles present on every shell
        for rr in radii:
            assert np.any(np.abs(z[sh] - (c[2] + rr)) < 1e-12)
            assert np.any(np.abs(z[sh] - (c[2] - rr)) < 1e-12)
    assert p.Nmax == int(30 * 36 * 30 * 1.01)
    assert p.coll_mass == pytest.approx(4 / 3 * math.pi * R ** 3, rel=1e-15)
    assert p.coll_I == pytest.approx(0.4 * p.coll_mass * R * R, rel=1e-15)


@pytest.mark.parametrize("key", inputs.KEY_ORDER)
def test_every_key_is_mandatory(key, tmp_path):
    if key == "coll_x":
        return
    cfg = helpers.cfg_couette(n=16)
    path = os.path.join(str(tmp_path), "input")
    inputs.write_input(path, cfg)
    lines = [l for l in open(path) if not l.startswith(key + " ")]
    open(path, "w").writelines(lines)
    s = sph_system.System("fp64")
    with pytest.raises(SphError) as e:
        s.read_input(path)
    assert e.value.code == 3 and f"{key} was not read" in e.value.message


@pytest.mark.parametrize("change,code,text", [
    (dict(N=[16, 16, 16]), 3, "dimensions of N"),
    (dict(L=[1.0, 1.0, 1.0]), 3, "dimensions of L"),
    (dict(ext_force_type=1, ext_force=[1.0, 0.0, 0.0]), 3, "dimensions of ext_force"),
    (dict(ext_force_type=4), 3, "not available"),
    (dict(N_colloids=2, coll_x=[[0.5, 0.5]]), 3, "coll_x"),
    (dict(N_colloids=2, coll_x=[[0.5, 0.5], [1.5, 0.5]], coll_R=0.05), 4, "out of the simulation box"),
    (dict(N_colloids=2, coll_x=[[0.5, 0.5], [0.55, 0.5]], coll_R=0.05), 4, "Overlapping"),
])
def test_read_input_errors(change, code, text, tmp_path):
    cfg = helpers.cfg_couette(n=16)
    cfg.update(change)
    path = os.path.join(str(tmp_path), "input")
    inputs.write_input(path, cfg)
    s = sph_system.System("fp64")
    with pytest.raises(SphError) as e:
        s.read_input(path)
    assert e.value.code == code and text in e.value.message


def test_3d_overlap_check_uses_z(tmp_path):
    # Same x and y, far apart in z: not overlapping (the reference's check used y
    # for z and would reject this input).
    cfg = helpers.cfg_sine(n=20)
    cfg.update(N_colloids=2, coll_R=0.1, coll_x=[[0.5, 0.5, 0.2], [0.5, 0.5, 0.7]],
               tau_repulsion=10.0, cutoff_magnetic=0.2)
    s = _read(cfg, tmp_path)
    s.initialize()
    assert s.params.N_colloids == 2


@pytest.mark.parametrize("change,text", [
    (dict(overlap=12.0), "neighbour cells are too big"),
    (dict(N_colloids=1, coll_x=[[0.5, 0.03]], coll_R=0.05), "bottom wall"),
    (dict(N_colloids=1, coll_x=[[0.5, 0.97]], coll_R=0.05), "top wall"),
    (dict(N_colloids=1, coll_x=[[0.5, 0.5]], coll_R=0.05, tau_repulsion=0.01),
     "colloids cells are too big"),
])
def test_initialize_errors(change, text, tmp_path):
    cfg = helpers.cfg_couette(n=32)
    cfg.update(change)
    s = _read(cfg, tmp_path)
    with pytest.raises(SphError) as e:
        s.initialize()
    assert e.value.code == 4 and text in e.value.message


def test_restart_roundtrip(tmp_path):
    cfg = helpers.cfg_colloids_2d(n=32)
    s = _read(cfg, tmp_path)
    s.initialize()
    inputs.seeded_velocity(s, 1e-2, 5)
    rpath = os.path.join(str(tmp_path), "restart.in")
    s.write_restart(rpath)
    cfg2 = dict(cfg, new_sim=1)
    s2 = _read(cfg2, tmp_path, "input2")
    s2.initialize(rpath)
    assert s2.params.N == s.params.N
    np.testing.assert_array_equal(s2.get("type"), s.get("type"))
    for n in ("x", "y", "vx", "vy", "mass", "cxx"):
        np.testing.assert_array_equal(s2.get(n), s.get(n).astype(np.float32).astype(np.float64))
    assert s2.get("colloids_list").tolist() == s.get("colloids_list").tolist()


def test_micro_file_format(tmp_path):
    cfg = helpers.cfg_couette(n=16)
    cfg.update(freq_micro=1, freq_walls=10**9, freq_macro=10**9)
    s = _read(cfg, tmp_path)
    s.initialize()
    out = str(tmp_path)
    # step 1: a micro step but not a macro step (macro output needs the GPU's
    # kinetic-energy reduction)
    s.print_output(1, out)
    lines = open(os.path.join(out, "micro-1.dat")).read().split("\n")
    assert len(lines) == s.N + 1 and lines[-1] == ""
    first = lines[0]
    assert first.endswith(" ")              # the 2D writer ends rows with " \n"
    cols = first.split()
    assert len(cols) == 12 and cols[0] == "0" and cols[8] == "1"
    assert cols[1] == "%.8f" % s.get("x")[0]


def test_sample_input_of_the_reference_format(tmp_path):
    # The reference's own sample layout: comments, blank lines, 2D colloids.
    text = """# Inputs for the SDPD program
  N        75  75
  L        25.0 25.0
  dim      2
  dt       0.0025
  t0       0.0
  Nsteps   1000000
  overlap  4.0
  rho      1.0
  c        1.0
  P0       0.9
  eta      0.1
  zeta     0.0
  #**** oldroyd ***
  np   10.0
  kT   0.001
  tau  10.0

  ext_force_type 0
  ext_force    1.0   0.0

  freq_micro    10000
  freq_macro    100
  freq_walls    10000
  freq_colloids 10000

  wall 0
  Vwall_bottom 0.01
  Vwall_top -0.01

  N_colloids 2
  coll_R  1.0
  coll_rho 1.0
  coll_move 0
  coll_x 12.5 12.5
  coll_x 14.6 12.5
  coll_repulsion_cuton  0.0
  F0_repulsion 0.02115
  tau_repulsion 1000.0

  cutoff_magnetic 3.0
  F0_magnetic 1.63
  omega_magnetic 0.003

  new_sim 0
"""
    path = os.path.join(str(tmp_path), "input")
    open(path, "w").write(text)
    s = sph_system.System("fp64")
    s.read_input(path)
    s.initialize()
    p = s.params
    assert (p.dim, p.N_colloids, p.wall) == (2, 2, 0)
    assert list(p.Ncells_colloids[:2]) == [8, 8]
    assert p.rcut == pytest.approx(25 / 75 * 4)
    assert s.get("coll_x").tolist() == [12.5, 14.6]
