"""Pins the FP64 oracle against mathematics that does not depend on it.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign, a transposed index, a wrong constant) fails it:
  kernel W, gradW      normalisation, W(0), continuity, W(rcut) = 0, gradW = dW/dr
  cell list            numpy stable argsort of the closed-form cell index
  density              O(N^2) brute force; analytic lattice sum on a perfect lattice
  velocity gradient    exact for a linear field: gradv = M G^T with the lattice
                       moment matrix M computed independently
  forces               total momentum conservation; zero on a lattice at rest
                       with c = I; pressure-only force vs an independent numpy sum
  dc/dt                steady simple shear of Oldroyd-B: c_xy = tau*g, c_xx = 1 + 2(tau*g)^2
  VV                   hand-computed single-particle updates, wall velocities
  colloids             two-body repulsion and magnetic forces in closed form;
                       torque of radial forces = 0; rotations preserve offsets
  kinetic energy       numpy sum
"""
import math

import numpy as np
import pytest
from scipy import integrate

import helpers
import oracle
from helpers import comps


def _sys(cfg, tmp_path, noise=0.0, seed=1):
    return helpers.make_system(cfg, str(tmp_path), noise=noise, seed=seed)


# ---------------------------------------------------------------- kernel W
@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_normalisation_and_shape(dim, tmp_path):
    cfg = helpers.cfg_couette(n=32) if dim == 2 else helpers.cfg_sine(n=16)
    p = _sys(cfg, tmp_path).params
    rc = p.rcut
    W = lambda r: oracle.kernel_W(p, r)
    gW = lambda r: oracle.kernel_gradW(p, r)
    if dim == 2:
        total = integrate.quad(lambda r: 2 * math.pi * r * W(r), 0, rc, points=[rc / 3, 2 * rc / 3],
                               epsabs=1e-14, epsrel=1e-13)[0]
    else:
        total = integrate.quad(lambda r: 4 * math.pi * r * r * W(r), 0, rc,
                               points=[rc / 3, 2 * rc / 3], epsabs=1e-14, epsrel=1e-13)[0]
    assert total == pytest.approx(1.0, abs=1e-11)
    assert W(0.0) == pytest.approx(66.0 * p.cw, rel=1e-15)
    assert abs(W(rc)) <= 1e-16 * p.cw
    assert gW(0.0) == pytest.approx(0.0, abs=1e-12 * abs(p.c_gradw))
    for q in (1.0, 2.0):   # continuity of W and gradW at q = 3r/rcut = 1, 2
        r = q * rc / 3
        assert W(r * (1 - 1e-12)) == pytest.approx(W(r * (1 + 1e-12)), rel=1e-9, abs=1e-12 * p.cw)
        assert gW(r * (1 - 1e-12)) == pytest.approx(gW(r * (1 + 1e-12)), rel=1e-9,
                                                   abs=1e-12 * abs(p.c_gradw))
    for r in np.linspace(0.05, 0.95, 19) * rc:
        h = 1e-6 * rc
        fd = (W(r + h) - W(r - h)) / (2 * h)
        assert gW(r) == pytest.approx(fd, rel=1e-7, abs=1e-9 * abs(p.c_gradw))
        assert gW(r) <= 0.0


# ---------------------------------------------------------------- cell list
@pytest.mark.parametrize("make", [lambda: helpers.cfg_couette(n=32), lambda: helpers.cfg_sine(n=16),
                                  lambda: helpers.cfg_colloids_3d(n=24)])
def test_cell_list_closed_form(make, tmp_path):
    s = _sys(make(), tmp_path)
    p = s.params
    rng = np.random.default_rng(3)
    pos = []
    for a, n in enumerate(["x", "y", "z"][:p.dim]):
        v = s.get(n)
        if not (a == 1 and p.wall):
            v = rng.uniform(0, p.L[a], v.shape)
            v[:4] = [0.0, p.L[a], p.cell_size[a], p.cell_size[a] * (p.Ncells[a] - 1)]
        s.set(n, v)
        pos.append(v)
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    pc, pi, cs, ce = o.cell_list()
    # closed form of kernel_cell_list (IEEE division then floor)
    cx = np.floor(pos[0] / p.cell_size[0]).astype(np.int64)
    cx[cx == p.Ncells[0]] -= 1
    if p.wall:
        cy = np.floor((pos[1] + p.wall_width) / p.cell_size[1]).astype(np.int64)
    else:
        cy = np.floor(pos[1] / p.cell_size[1]).astype(np.int64)
        cy[cy == p.Ncells[1]] -= 1
    cell = cy * p.Ncells[0] + cx
    if p.dim == 3:
        cz = np.floor(pos[2] / p.cell_size[2]).astype(np.int64)
        cz[cz == p.Ncells[2]] -= 1
        cell = cz * p.Ncells[0] * p.Ncells[1] + cell
    order = np.argsort(cell, kind="stable")
    np.testing.assert_array_equal(pi, order)
    np.testing.assert_array_equal(pc, cell[order])
    counts = np.bincount(cell, minlength=p.Ntotal_cells)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    occ = counts > 0
    np.testing.assert_array_equal(cs[occ], starts[occ])
    np.testing.assert_array_equal(ce[occ], starts[occ] + counts[occ] - 1)


# ---------------------------------------------------------------- density
@pytest.mark.parametrize("make", [lambda: helpers.cfg_couette(n=32),
                                  lambda: helpers.cfg_sine(n=14),
                                  lambda: helpers.cfg_colloids_2d(n=40)])
def test_density_equals_brute_force(make, tmp_path):
    s = _sys(make(), tmp_path)
    helpers.perturb(s, pos=0.3)
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_density()
    d = o.get("dens")
    bf = o.density_brute_force()
    np.testing.assert_allclose(d, bf, rtol=1e-13)


def test_density_on_perfect_lattice_equals_lattice_sum(tmp_path):
    s = _sys(helpers.cfg_sine(n=16), tmp_path)
    p = s.params
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_density()
    d = o.get("dens")
    dx = 1.0 / 16
    k = np.arange(-4, 5)
    off = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3) * dx
    r = np.linalg.norm(off, axis=1)
    r = r[r < p.rcut]
    q = 3 * r / p.rcut
    w = (3 - q) ** 5 - 6 * np.where(q < 2, (2 - q) ** 5, 0) + 15 * np.where(q < 1, (1 - q) ** 5, 0)
    expect = p.cw * w.sum()
    np.testing.assert_allclose(d, expect, rtol=1e-13)


# ---------------------------------------------------------------- velocity gradient
def test_gradv_of_linear_field(tmp_path):
    """v = G x on a perfect periodic lattice, checked on particles whose kernel
    support does not cross the periodic seam: gradv_ab = sum_c M_ac G_bc with
    M_ac = -(1/d) sum_j gradW(r) e_a r_c (closed form over lattice offsets)."""
    s = _sys(helpers.cfg_sine(n=16), tmp_path)
    p = s.params
    G = np.array([[0.3, -1.1, 0.7], [2.0, 0.5, -0.4], [-0.9, 1.3, 0.2]])
    x = np.stack([s.get("x"), s.get("y"), s.get("z")], 1)
    v = x @ G.T
    for a, n in enumerate(["vx", "vy", "vz"]):
        s.set(n, v[:, a])
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_density()
    o.calculate_vel_gradient()
    gv = np.stack([o.get(n) for n in comps("gradv", 3)], 1).reshape(-1, 3, 3)
    d = o.get("dens")
    dx = 1.0 / 16
    k = np.arange(-4, 5)
    off = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3) * dx
    r = np.linalg.norm(off, axis=1)
    keep = (r < p.rcut) & (r > 0)
    off, r = off[keep], r[keep]
    q = 3 * r / p.rcut
    gw = p.c_gradw * ((3 - q) ** 4 - 6 * np.where(q < 2, (2 - q) ** 4, 0)
                      + 15 * np.where(q < 1, (1 - q) ** 4, 0))
    e = off / r[:, None]
    interior = np.all((x > p.rcut + 1e-9) & (x < 1 - p.rcut - 1e-9), axis=1)
    assert interior.sum() > 50
    for i in np.flatnonzero(interior)[:200]:
        M = -(1.0 / d[i]) * np.einsum("j,ja,jc->ac", gw, e, off)
        np.testing.assert_allclose(gv[i], M @ G.T, rtol=1e-11, atol=1e-11 * np.abs(G).max())
    # M is isotropic on a cubic lattice, so gradv = m G^T; m -> 1 as overlap grows
    assert abs(M[0, 1]) < 1e-12 and M[0, 0] == pytest.approx(M[1, 1], rel=1e-12)


def test_uniform_velocity_has_zero_gradient(tmp_path):
    s = _sys(helpers.cfg_sine(n=14), tmp_path)
    helpers.perturb(s, pos=0.2)
    for n, val in (("vx", 0.3), ("vy", -0.2), ("vz", 0.1)):
        s.set(n, np.full(s.N, val))
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_density()
    o.calculate_vel_gradient()
    for n in comps("gradv", 3):
        assert np.max(np.abs(o.get(n))) <= 1e-10


# ---------------------------------------------------------------- forces
def test_momentum_conservation(tmp_path):
    s = _sys(helpers.cfg_sine(n=14, ext_force_type=0), tmp_path, noise=0.05)
    helpers.perturb(s, pos=0.25, conf=0.2)
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_forces()
    F = np.stack([o.get(n) for n in ("fx", "fy", "fz")], 1)
    assert np.abs(F).max() > 0
    np.testing.assert_allclose(F.sum(0), 0.0, atol=1e-12 * np.abs(F).sum())


def test_lattice_at_rest_has_zero_force(tmp_path):
    s = _sys(helpers.cfg_sine(n=14, ext_force_type=0), tmp_path)
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_forces()
    p = s.params
    pair_scale = abs(p.c_gradw) * 2 * o.get("press").max() / o.get("dens").min() ** 2
    for n in ("fx", "fy", "fz"):
        assert np.abs(o.get(n)).max() <= 1e-12 * pair_scale


def test_forces_match_independent_numpy_sum(tmp_path):
    """Periodic fluid: F_i = sum_j [-gradW (Pi_i/d_i^2 + Pi_j/d_j^2) e + gradW/(d_i d_j r)
    (a v_ij + b (e.v_ij) e)] evaluated with numpy over all pairs."""
    s = _sys(helpers.cfg_sine(n=10, ext_force_type=0), tmp_path, noise=0.05)
    helpers.perturb(s, pos=0.3, conf=0.3)
    p = s.params
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_forces()
    x = np.stack([o.get(n) for n in ("x", "y", "z")], 1)
    v = np.stack([o.get(n) for n in ("vx", "vy", "vz")], 1)
    c6 = [o.get(n) for n in comps("c", 3)]
    C = np.zeros((s.N, 3, 3))
    for (a, b), cc in zip([(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)], c6):
        C[:, a, b] = cc
        C[:, b, a] = cc
    d, pr = o.get("dens"), o.get("press")
    Pi = pr[:, None, None] * np.eye(3) - p.nkT * C
    Pd = Pi / (d ** 2)[:, None, None]
    r = x[:, None, :] - x[None, :, :]
    r -= np.round(r)               # L = 1 minimum image
    rr = np.linalg.norm(r, axis=2)
    mask = (rr < p.rcut) & (rr > 0)
    q = np.where(mask, 3 * rr / p.rcut, 3.0)
    gw = p.c_gradw * ((3 - q) ** 4 - 6 * np.where(q < 2, (2 - q) ** 4, 0)
                      + 15 * np.where(q < 1, (1 - q) ** 4, 0))
    gw = np.where(mask, gw, 0.0)
    e = np.where(mask[..., None], r / np.where(mask, rr, 1)[..., None], 0.0)
    vij = v[:, None, :] - v[None, :, :]
    rev = -gw[..., None] * np.einsum("ijab,ijb->ija", Pd[:, None] + Pd[None, :], e)
    aux = gw / (d[:, None] * d[None, :] * np.where(mask, rr, 1))
    irr = aux[..., None] * (p.a * vij + p.b * np.einsum("ijk,ijk->ij", e, vij)[..., None] * e)
    F = (rev + irr).sum(1)
    for a, n in enumerate(("fx", "fy", "fz")):
        np.testing.assert_allclose(o.get(n), F[:, a], rtol=1e-9, atol=1e-10 * np.abs(F).max())


# ---------------------------------------------------------------- dc/dt
def test_dcdt_steady_shear(tmp_path):
    s = _sys(helpers.cfg_sine(n=10), tmp_path)
    p = s.params
    g = 0.8                                   # v_x = g y  ->  dv_x/dy = gradv_yx = g
    tau = p.tau
    N = s.N
    o = oracle.Oracle.from_system(s)
    o.set("cxx", np.full(N, 1 + 2 * (tau * g) ** 2))
    o.set("cxy", np.full(N, tau * g))
    o.set("cyy", np.ones(N))
    o.set("czz", np.ones(N))
    for n in comps("gradv", 3):
        o.set(n, np.full(N, g if n == "gradvyx" else 0.0))
    o.calculate_dcdt()
    for n in comps("dcdt", 3):
        np.testing.assert_allclose(o.get(n), 0.0, atol=1e-14)
    # away from steady state: dcdt = c L^T + L c + (I - c)/tau with L = gradv^T
    rng = np.random.default_rng(2)
    A = rng.normal(size=(3, 3))
    C = np.eye(3) + 0.1 * (A + A.T)
    Gv = rng.normal(size=(3, 3))              # Gv[a, b] = dv_b/dx_a
    for (a, b), n in zip([(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)], comps("c", 3)):
        o.set(n, np.full(N, C[a, b]))
    for k, n in enumerate(comps("gradv", 3)):
        o.set(n, np.full(N, Gv[k // 3, k % 3]))
    o.calculate_dcdt()
    Lv = Gv.T
    D = C @ Lv.T + Lv @ C + (np.eye(3) - C) / tau
    for (a, b), n in zip([(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)], comps("dcdt", 3)):
        np.testing.assert_allclose(o.get(n), D[a, b], rtol=1e-13, atol=1e-14)


# ---------------------------------------------------------------- integrator
def test_velocity_verlet_single_particle_updates(tmp_path):
    s = _sys(helpers.cfg_couette(n=16), tmp_path)
    p = s.params
    o = oracle.Oracle.from_system(s)
    N = s.N
    rng = np.random.default_rng(4)
    F = rng.normal(size=(N, 2))
    dc = rng.normal(size=(N, 3))
    for a, n in enumerate(("fx", "fy")):
        o.set(n, F[:, a])
    for a, n in enumerate(("dcdt_xx", "dcdt_xy", "dcdt_yy")):
        o.set(n, dc[:, a])
    x0, y0, m = o.get("x"), o.get("y"), o.get("mass")
    c0 = o.get("cxy")
    t = o.get("type")
    o.V_Verlet_step1()
    vx = o.get("vx")
    fl = t == 0
    np.testing.assert_allclose(vx[fl], 0.5 * p.dt / m[fl] * F[fl, 0], rtol=1e-15)
    np.testing.assert_array_equal(vx[t == 1], p.V_bottom)
    np.testing.assert_array_equal(vx[t == 2], p.V_top)
    np.testing.assert_allclose(o.get("cxy")[fl], c0[fl] + 0.5 * p.dt * dc[fl, 1], rtol=1e-15)
    np.testing.assert_allclose(helpers.periodic_diff(o.get("x"), x0 + vx * p.dt, 1.0), 0,
                               atol=1e-15)
    np.testing.assert_allclose(o.get("y")[t > 0], y0[t > 0], atol=0)   # walls move in x only
    o.V_Verlet_step2()
    np.testing.assert_allclose(o.get("vx")[fl], p.dt / m[fl] * F[fl, 0], rtol=1e-14)
    np.testing.assert_array_equal(o.get("vx")[t == 1], p.V_bottom)


def test_kinetic_energy(tmp_path):
    s = _sys(helpers.cfg_colloids_3d(n=20), tmp_path, noise=0.01)
    o = oracle.Oracle.from_system(s)
    v2 = sum(o.get(n) ** 2 for n in ("vx", "vy", "vz"))
    assert o.kinetic_energy() == pytest.approx(0.5 * np.sum(o.get("mass") * v2), rel=1e-13)


# ---------------------------------------------------------------- colloids
def _two_colloids(tmp_path, dim, dist, **kw):
    if dim == 2:
        cfg = helpers.cfg_colloids_2d(n=48)
        cfg["coll_x"] = [[0.40, 0.50], [0.40 + dist, 0.50]]
    else:
        cfg = helpers.cfg_colloids_3d(n=24, n_colloids=2)
        cfg["coll_x"] = [[0.35, 0.50, 0.50], [0.35 + dist, 0.50, 0.50]]
    cfg.update(kw)
    s = _sys(cfg, tmp_path)
    o = oracle.Oracle.from_system(s)
    for n in ("fx", "fy", "fz"):
        o.set(n, np.zeros(s.N))
    return s, o


@pytest.mark.parametrize("dim", [2, 3])
def test_repulsion_and_magnetic_two_body(dim, tmp_path):
    s, o = _two_colloids(tmp_path, dim, 0.27 if dim == 3 else 0.23)
    p = s.params
    o.neighbours_search()
    step = 7
    o.set_step(step)
    o.calculate_force_colloids()
    X = np.stack([o.get(n) for n in ("coll_x", "coll_y", "coll_z")], 1)[:, :dim]
    r = X[0] - X[1]
    rr = np.linalg.norm(r)
    # repulsion: F0 tau e^{-tau s}/(1 - e^{-tau s}) rhat, s = (max(r, rcuton) - 2R)/R
    rs = max(rr, p.rcuton_coll_rep)
    Frep = 0.0
    if rs <= p.rcutoff_coll_rep:
        ex = math.exp(-p.tau_rep * (rs - 2 * p.coll_R) / p.coll_R)
        Frep = p.F0_rep * p.tau_rep * ex / (1 - ex)
    F = Frep * r / rs
    ang = p.omega_magnet * (p.t0 + p.dt * step)
    m = np.array([math.cos(ang), math.sin(ang), 0.0])[:dim]
    if rr * rr <= p.r0_magnet_sq:
        e = r / rr * (p.coll_R if dim == 3 else 1.0)      # 3D keeps the reference's R r/|r|
        me = m @ e
        F = F + p.F0_magnet / rr ** 4 * (2 * me * m - (5 * me * me - 1) * e)
    assert np.linalg.norm(F) > 0
    got = np.array([o.get(n)[0] for n in ("fx_colloid", "fy_colloid", "fz_colloid")[:dim]])
    np.testing.assert_allclose(got, F, rtol=1e-12)
    got1 = np.array([o.get(n)[1] for n in ("fx_colloid", "fy_colloid", "fz_colloid")[:dim]])
    np.testing.assert_allclose(got1, -F, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_radial_boundary_forces_give_force_without_torque(dim, tmp_path):
    s, o = _two_colloids(tmp_path, dim, 0.45)
    t = o.get("type")
    X = [o.get(n) for n in ("coll_x", "coll_y", "coll_z")]
    pos = [o.get(n) for n in ("x", "y", "z")]
    F = [np.zeros(s.N) for _ in range(3)]
    for k in range(2):
        sh = t == 3 + k
        for a in range(dim):
            F[a][sh] = 0.3 * (pos[a][sh] - X[a][k])
    for a, n in enumerate(("fx", "fy", "fz")):
        o.set(n, F[a])
    o.neighbours_search()
    o.set_step(0)
    o.calculate_force_colloids()
    for n in ("tz_colloid",) + (("tx_colloid", "ty_colloid") if dim == 3 else ()):
        np.testing.assert_allclose(o.get(n), 0.0, atol=1e-14)
    for k in range(2):
        sh = t == 3 + k
        assert o.get("fx_colloid")[k] == pytest.approx(F[0][sh].sum(), rel=1e-12, abs=1e-14)


def test_colloid_rotation_preserves_offsets(tmp_path):
    s, o = _two_colloids(tmp_path, 3, 0.45, coll_rho=0.5)
    t = o.get("type")
    o.set("tx_colloid", np.array([0.02, -0.01]))
    o.set("ty_colloid", np.array([0.01, 0.03]))
    o.set("tz_colloid", np.array([-0.02, 0.005]))
    before = []
    for k in range(2):
        sh = t == 3 + k
        X = [o.get(n)[k] for n in ("coll_x", "coll_y", "coll_z")]
        before.append(np.sqrt(sum((o.get(n)[sh] - X[a]) ** 2
                                  for a, n in enumerate(("x", "y", "z")))))
    for _ in range(5):
        o.V_Verlet_step1()
    q = np.stack([o.get(n) for n in ("coll_quat0", "coll_quatx", "coll_quaty", "coll_quatz")], 1)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-14)
    assert np.all(np.abs(q[:, 0]) < 1.0)      # they did rotate
    for k in range(2):
        sh = t == 3 + k
        X = [o.get(n)[k] for n in ("coll_x", "coll_y", "coll_z")]
        after = np.sqrt(sum((o.get(n)[sh] - X[a]) ** 2 for a, n in enumerate(("x", "y", "z"))))
        np.testing.assert_allclose(after, before[k], rtol=1e-12)


def test_colloid_rotation_2d_is_exact(tmp_path):
    s, o = _two_colloids(tmp_path, 2, 0.45)
    p = s.params
    t = o.get("type")
    o.set("tz_colloid", np.array([1e-3, 0.0]))
    sh = t == 3
    X0, Y0 = o.get("coll_x")[0], o.get("coll_y")[0]
    ang0 = np.arctan2(o.get("y")[sh] - Y0, o.get("x")[sh] - X0)
    o.V_Verlet_step1()
    om = 0.5 * p.dt / p.coll_I * 1e-3
    ang1 = np.arctan2(o.get("y")[sh] - Y0, o.get("x")[sh] - X0)
    np.testing.assert_allclose(np.angle(np.exp(1j * (ang1 - ang0))), om * p.dt, atol=1e-12)


def test_wall_force_is_sum_over_wall_particles(tmp_path):
    s = _sys(helpers.cfg_couette(n=32), tmp_path, noise=1e-2)
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_forces()
    t = o.get("type")
    for w, n in ((1, "fx"), (2, "fy")):
        pass
    fxw, fyw = o.get("fx_wall"), o.get("fy_wall")
    for w in (0, 1):
        sel = t == w + 1
        assert fxw[w] == pytest.approx(o.get("fx")[sel].sum(), rel=1e-12)
        assert fyw[w] == pytest.approx(o.get("fy")[sel].sum(), rel=1e-12)


# ---------------------------------------------------------------- observables
def test_observables_closed_forms(tmp_path):
    s, o = _two_colloids(tmp_path, 2, 0.25)
    p = s.params
    # two colloids 0.2 apart along the direction at 150 degrees
    phi = math.radians(150.0)
    X = np.array([0.5, 0.5 + 0.2 * math.cos(phi)])
    Y = np.array([0.5, 0.5 + 0.2 * math.sin(phi)])
    o.set("coll_x", X)
    o.set("coll_y", Y)
    t = o.get("type")
    rng = np.random.default_rng(9)
    cxx, cyy, cxy = (1 + rng.uniform(0, 1, s.N)), (1 + rng.uniform(0, 1, s.N)), rng.normal(size=s.N)
    for n, v in (("cxx", cxx), ("cyy", cyy), ("cxy", cxy)):
        o.set(n, v)
    o.set("fx", rng.normal(size=s.N))
    o.neighbours_search()
    o.calculate_force_walls()
    o.set_step(11)
    out = o.observables()
    fl = t == 0
    assert out[1] == pytest.approx(p.nkT * np.mean(cxx[fl] - cyy[fl]), rel=1e-12)
    assert out[2] == pytest.approx(p.nkT * np.mean(cxy[fl]), rel=1e-12)
    assert out[3] == pytest.approx(o.get("fx")[t == 1].sum() / p.L[0], rel=1e-12)
    assert out[5] == pytest.approx(phi - math.pi, abs=1e-12)     # 150 deg -> -30 deg mod pi
    fa = p.omega_magnet * (p.t0 + p.dt * 11)
    fa -= math.pi * math.ceil(fa / math.pi - 0.5)
    assert out[6] == pytest.approx(fa, abs=1e-12)
    lag = out[6] - out[5]
    lag -= math.pi * math.ceil(lag / math.pi - 0.5)
    assert out[7] == pytest.approx(lag, abs=1e-12)
    assert -math.pi / 2 < out[7] <= math.pi / 2
