"""Pins the FP64 oracle against mathematics that does not depend on it.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign, a transposed index, a wrong constant) fails it:
  derived constants    tests/test_derived_constants.py (the oracle derives them
                       itself from the raw inputs; EOS closed form, polymer modulus)
  kernel W, gradW      normalisation, W(0), continuity, W(rcut) = 0, gradW = dW/dr
  cell list            numpy stable argsort of the closed-form cell index
  density              O(N^2) brute force; analytic lattice sum on a perfect lattice
  velocity gradient    exact for a linear field: gradv = M G^T with the lattice
                       moment matrix M computed independently
  forces               total momentum conservation; zero on a lattice at rest
                       with c = I; full pair force vs an independent numpy sum
  external force       closed forms of the three force types
  Morris boundaries    walls: exact for a linear shear flow (gradient and full
                       force, with the fluid-partner density/pressure rule);
                       walls and colloids: the momentum balance of mirrored
                       pairs in closed form; colloids: zero gradient under a
                       common translation, spin torque opposing and linear in
                       the spin
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
    p = oracle.derive_params(_sys(cfg, tmp_path).params)   # the oracle's cw, c_gradw
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
    p = o.params   # the oracle's cell grid
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


# ---------------------------------------------------------------- external force
@pytest.mark.parametrize("kind", [1, 2, 3])
def test_external_force_closed_form(kind, tmp_path):
    """Fluid only: type 1 m f; type 2 m f_x sin(2 pi y / L1) along x; type 3
    +m f_x below mid-channel and -m f_x above it."""
    f = [0.7, -0.4, 0.25]
    s = _sys(helpers.cfg_colloids_3d(n=20, ext_force_type=kind, ext_force=f), tmp_path)
    helpers.perturb(s, pos=0.3)
    p = s.params
    o = oracle.Oracle.from_system(s)
    rng = np.random.default_rng(kind)
    F0 = [rng.normal(size=s.N) for _ in range(3)]
    for a, n in enumerate(("fx", "fy", "fz")):
        o.set(n, F0[a])
    o.calculate_external_forces()
    m, y, fl = o.get("mass"), o.get("y"), o.get("type") == 0
    want = np.zeros((3, s.N))
    if kind == 1:
        want[:, fl] = np.outer(f, m[fl])
    elif kind == 2:
        want[0, fl] = m[fl] * f[0] * np.sin(2 * np.pi / p.L[1] * y[fl])
    else:
        want[0, fl] = np.where(y[fl] < 0.5 * p.L[1], 1.0, -1.0) * m[fl] * f[0]
    for a, n in enumerate(("fx", "fy", "fz")):
        np.testing.assert_allclose(o.get(n) - F0[a], want[a], rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- Morris boundaries
def _pairs_within(x, L, periodic, rcut):
    """All ordered pairs (i, j), i != j, closer than rcut under the minimum image
    along the periodic axes: (i, j, r_ij = x_i - x_j, |r_ij|)."""
    r = x[:, None, :] - x[None, :, :]
    for a in range(x.shape[1]):
        if periodic[a]:
            r[..., a] -= L[a] * np.round(r[..., a] / L[a])
    rr = np.linalg.norm(r, axis=2)
    i, j = np.nonzero((rr < rcut) & (rr > 0))
    return i, j, r[i, j], rr[i, j]


def _gradW(p, r):
    q = 3 * r / p.rcut
    return p.c_gradw * ((3 - q) ** 4 - 6 * np.where(q < 2, (2 - q) ** 4, 0)
                        + 15 * np.where(q < 1, (1 - q) ** 4, 0))


@pytest.mark.parametrize("dim", [2, 3])
def test_wall_morris_is_exact_for_linear_shear(dim, tmp_path):
    """Wall particles squeezed to within 0.2 dx of their wall while the nearest
    fluid sits 0.5 dx from it keep beta = 1 + d_wall/d_fluid below its cap, so
    the Morris velocity beta (v_i - V_wall) equals v(y_i) - v(y_j) for a flow
    linear in y that takes the wall's velocity at the wall.  The oracle's
    velocity gradient and its complete pair force on every fluid particle must
    then equal plain SPH sums with those field differences, where a wall partner
    j enters the force with the fluid particle's density and pressure and its own
    conformation tensor (src/kernel_forces.cu:213-233)."""
    if dim == 2:
        cfg = helpers.cfg_couette(n=24, Vwall_bottom=-0.5, Vwall_top=0.3)
    else:
        cfg = helpers.cfg_colloids_3d(n=12, n_colloids=0, Vwall_bottom=-0.5, Vwall_top=0.3)
        cfg.update(N_colloids=0, coll_x=[])
    s = _sys(cfg, tmp_path)
    p = s.params
    rng = np.random.default_rng(5)
    dx = p.L[0] / p.Nxyz[0]
    t = s.get("type")
    fl = t == 0
    x = np.stack([s.get(n) for n in ("x", "y", "z")[:dim]], 1)
    for a in [0] + ([2] if dim == 3 else []):          # jitter along the periodic axes
        x[fl, a] = np.mod(x[fl, a] + rng.uniform(-0.3, 0.3, fl.sum()) * dx, p.L[a])
    nb, nt = (t == 1).sum(), (t == 2).sum()
    x[t == 1, 1] = p.y_bottom - rng.uniform(0.01, 0.2, nb) * dx
    x[t == 2, 1] = p.y_top + rng.uniform(0.01, 0.2, nt) * dx
    # Two linear fields: one for the lower half (bottom wall), one for the upper.
    gb, gt = np.array([1.3, 0.4, -0.6])[:dim], np.array([-0.8, -0.5, 0.9])[:dim]
    e0 = np.eye(dim)[0]
    low = x[:, 1] < 0.5 * p.L[1]
    field = np.where(low[:, None], p.V_bottom * e0 + (x[:, 1:2] - p.y_bottom) * gb,
                     p.V_top * e0 + (x[:, 1:2] - p.y_top) * gt)
    for a, n in enumerate(("x", "y", "z")[:dim]):
        s.set(n, x[:, a])
    for a, n in enumerate(("vx", "vy", "vz")[:dim]):
        v = np.where(fl, field[:, a], 0.0)
        s.set(n, v)
    for n in comps("c", dim):                           # every particle, walls included
        s.set(n, s.get(n) + rng.uniform(-0.3, 0.3, s.N))
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_density()
    o.calculate_pressure()
    o.calculate_vel_gradient()
    o.calculate_pair_forces()
    d, pr = o.get("dens"), o.get("press")
    i, j, r, rr = _pairs_within(x, p.L[:dim], [True, False, True][:dim], p.rcut)
    keep = fl[i]                                  # fluid i with any partner
    i, j, r, rr = i[keep], j[keep], r[keep], rr[keep]
    both_wall = ~fl[j]
    assert both_wall.any()
    beta = np.where(both_wall, 1 + np.abs(x[j, 1] - np.where(t[j] == 1, p.y_bottom, p.y_top))
                    / np.abs(x[i, 1] - np.where(t[j] == 1, p.y_bottom, p.y_top)), 0)
    assert beta.max() < p.beta_max
    vij = field[i] - field[j]                    # the linear field at both positions
    gw = _gradW(p, rr)
    e = r / rr[:, None]
    gv = np.zeros((s.N, dim, dim))
    np.add.at(gv, i, -(gw / d[i])[:, None, None] * e[:, :, None] * vij[:, None, :])
    got = np.stack([o.get(n) for n in comps("gradv", dim)], 1).reshape(-1, dim, dim)
    scale = np.abs(gv[fl]).max()
    np.testing.assert_allclose(got[fl], gv[fl], rtol=0, atol=1e-11 * scale)
    # Pair force with the fluid-partner rule for wall j.
    C = np.zeros((s.N, dim, dim))
    idx = [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]
    for (a, b), n in zip(idx, comps("c", dim)):
        C[:, a, b] = C[:, b, a] = o.get(n)
    dj = np.where(fl[j], d[j], d[i])
    pj = np.where(fl[j], pr[j], pr[i])
    Pi_i = pr[i][:, None, None] * np.eye(dim) - p.nkT * C[i]
    Pi_j = pj[:, None, None] * np.eye(dim) - p.nkT * C[j]
    P = Pi_i / (d[i] ** 2)[:, None, None] + Pi_j / (dj ** 2)[:, None, None]
    pair = -gw[:, None] * np.einsum("kab,kb->ka", P, e)
    pair += (gw / (d[i] * dj * rr))[:, None] * (p.a * vij + p.b * (e * vij).sum(1)[:, None] * e)
    F = np.zeros((s.N, dim))
    np.add.at(F, i, pair)
    got = np.stack([o.get(n) for n in comps("f", dim)], 1)
    np.testing.assert_allclose(got[fl], F[fl], rtol=0, atol=1e-10 * np.abs(F[fl]).max())


@pytest.mark.parametrize("case", ["couette", "colloids_2d", "colloids_3d"])
def test_momentum_balance_with_boundaries(case, tmp_path):
    """A pair of a fluid particle f and a wall or colloid particle w is evaluated
    from both sides with mirrored Morris velocities and f's density and pressure.
    The stress terms differ: from f's side w enters with Pi_w = p_f - nkT c_w
    (src/kernel_forces.cu:213-233), from w's side both particles take f's full
    Pi_f (src/kernel_forces.cu:319-335).  So the pair forces over all particles
    sum to  sum_(f,w) -gradW nkT (c_f - c_w) e_fw / d_f^2  exactly, which is zero
    when the boundary particles carry the fluid's c; checked with moving walls
    and translating, spinning colloids."""
    cfg = {"couette": lambda: helpers.cfg_couette(n=32),
           "colloids_2d": lambda: helpers.cfg_colloids_2d(n=40),
           "colloids_3d": lambda: helpers.cfg_colloids_3d(n=20)}[case]()
    s = _sys(cfg, tmp_path, noise=0.05)
    helpers.perturb(s, pos=0.3, conf=0.0)
    p = s.params
    dim = p.dim
    rng = np.random.default_rng(2)
    o = oracle.Oracle.from_system(s)
    if p.N_colloids:
        for n in ("coll_vx", "coll_vy", "coll_vz", "coll_omegax", "coll_omegay", "coll_omegaz"):
            o.set(n, rng.uniform(-0.5, 0.5, p.N_colloids))

    def total_force():
        o.neighbours_search()
        o.calculate_density()
        o.calculate_pressure()
        o.calculate_pair_forces()
        return np.stack([o.get(n) for n in comps("f", dim)], 1)

    for n in comps("c", dim):       # c = I everywhere (colloid shells start at c = 0)
        o.set(n, np.full(s.N, 1.0 if n in ("cxx", "cyy", "czz") else 0.0))
    F = total_force()
    t = o.get("type")
    assert np.abs(F[t != 0]).max() > 0
    np.testing.assert_allclose(F.sum(0), 0.0, atol=1e-13 * np.abs(F).sum())
    C = np.zeros((s.N, dim, dim))
    idx = [(0, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 2)]
    for (a, b), n in zip(idx, comps("c", dim)):
        c = o.get(n) + rng.uniform(-0.2, 0.2, s.N)      # fluid, walls and shells
        o.set(n, c)
        C[:, a, b] = C[:, b, a] = c
    F = total_force()
    x = np.stack([o.get(n) for n in comps("x", dim)], 1)
    i, j, r, rr = _pairs_within(x, p.L[:dim], [True, not p.wall, True][:dim], p.rcut)
    mixed = (t[i] == 0) & (t[j] != 0)
    i, j, r, rr = i[mixed], j[mixed], r[mixed], rr[mixed]
    e = r / rr[:, None]
    d = o.get("dens")
    coef = -_gradW(p, rr) * p.nkT / d[i] ** 2
    want = (coef[:, None] * np.einsum("kab,kb->ka", C[i] - C[j], e)).sum(0)
    assert np.abs(want).max() > 1e-6 * np.abs(F).sum() / s.N
    np.testing.assert_allclose(F.sum(0), want, rtol=1e-9, atol=1e-13 * np.abs(F).sum())


@pytest.mark.parametrize("dim", [2, 3])
def test_colloid_morris_galilean_and_spin(dim, tmp_path):
    """(1) Fluid, walls and a non-spinning colloid all moving with one velocity
    have no relative velocity: the velocity gradient of every fluid particle is
    zero.  (2) With fluid and walls at rest, spinning colloid 0 adds a torque
    that opposes the spin and is linear in it (the pressure part does not
    depend on velocity and drops out of the difference)."""
    V = 0.37
    if dim == 2:
        cfg = helpers.cfg_colloids_2d(n=40, Vwall_bottom=V, Vwall_top=V)
    else:
        cfg = helpers.cfg_colloids_3d(n=20, Vwall_bottom=V, Vwall_top=V)
    s = _sys(cfg, tmp_path)
    helpers.perturb(s, pos=0.3)
    p = s.params
    nc = p.N_colloids
    s.set("vx", np.where(s.get("type") == 0, V, 0.0))
    o = oracle.Oracle.from_system(s)
    o.set("coll_vx", np.full(nc, V))
    o.neighbours_search()
    o.calculate_density()
    o.calculate_vel_gradient()
    fl = o.get("type") == 0
    gscale = abs(p.c_gradw) * V / o.get("dens").min() * 100
    for n in comps("gradv", dim):
        assert np.abs(o.get(n)[fl]).max() <= 1e-13 * gscale, n

    def torque(omega):
        q = oracle.Oracle.from_system(s)
        q.set("vx", np.zeros(s.N))
        q.set("coll_vx", np.zeros(nc))
        w = np.zeros((3, nc))
        w[:, 0] = omega
        for a, n in enumerate(("coll_omegax", "coll_omegay", "coll_omegaz")):
            q.set(n, w[a])
        q.neighbours_search()
        q.calculate_density()
        q.calculate_pressure()
        q.calculate_pair_forces()
        q.set_step(0)
        q.calculate_force_colloids()
        return np.array([q.get(n)[0] for n in ("tx_colloid", "ty_colloid", "tz_colloid")])

    w = np.array([0.0, 0.0, 1.5]) if dim == 2 else np.array([0.6, -0.9, 1.2])
    T0, T1, T2 = torque(0 * w), torque(w), torque(2 * w)
    dT = T1 - T0
    assert dT @ w < 0 and np.linalg.norm(dT) > 0
    np.testing.assert_allclose(T2 - T0, 2 * dT, rtol=1e-9, atol=1e-12 * np.abs(dT).max())


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


@pytest.mark.parametrize("dim", [2, 3])
def test_colloid_rigid_body_step(dim, tmp_path):
    """M = pi R^2 rho_c (2D) or 4/3 pi R^3 rho_c, I = M R^2/2 or 2/5 M R^2.
    VV1: V += dt/(2M) F, omega += dt/(2I) T, X += V dt, and each shell particle
    is placed at X + (offset rotated by |omega| dt about omega) (Rodrigues'
    formula; the 3D quaternion path).  VV2: V += dt/(2M) F, omega += dt/(2I) T."""
    s, o = _two_colloids(tmp_path, dim, 0.45, coll_rho=0.8)
    p = s.params
    R = p.coll_R
    M = (math.pi * R * R if dim == 2 else 4.0 / 3.0 * math.pi * R ** 3) * 0.8
    assert p.coll_mass == pytest.approx(M, rel=1e-15)
    assert p.coll_I == pytest.approx((0.5 if dim == 2 else 0.4) * M * R * R, rel=1e-15)
    rng = np.random.default_rng(9)
    F = rng.normal(size=(3, 2)) * 1e-2
    T = rng.normal(size=(3, 2)) * 1e-3
    V = rng.normal(size=(3, 2)) * 0.3
    W = rng.normal(size=(3, 2)) * 2.0
    if dim == 2:
        F[2] = T[:2] = V[2] = W[:2] = 0.0
    for a, ax in enumerate("xyz"):
        o.set(f"f{ax}_colloid", F[a])
        o.set(f"t{ax}_colloid", T[a])
        o.set(f"coll_v{ax}", V[a])
        o.set(f"coll_omega{ax}", W[a])
    t = o.get("type")
    X0 = np.stack([o.get(f"coll_{ax}") for ax in "xyz"])
    pos0 = np.stack([o.get(ax) for ax in "xyz"])
    o.V_Verlet_step1()
    V1 = V + 0.5 * p.dt / M * F
    W1 = W + 0.5 * p.dt / p.coll_I * T
    for a, ax in enumerate("xyz"[:dim]):
        np.testing.assert_allclose(o.get(f"coll_v{ax}"), V1[a], rtol=1e-14, atol=1e-16)
        np.testing.assert_allclose(o.get(f"coll_{ax}"), X0[a] + V1[a] * p.dt, rtol=1e-14)
    for a, ax in enumerate("xyz"):
        if dim == 3 or ax == "z":
            np.testing.assert_allclose(o.get(f"coll_omega{ax}"), W1[a], rtol=1e-14, atol=1e-16)
    X1 = np.stack([o.get(f"coll_{ax}") for ax in "xyz"])
    pos1 = np.stack([o.get(ax) for ax in "xyz"])
    for k in range(2):
        sh = t == 3 + k
        off0 = pos0[:, sh] - X0[:, k:k + 1]
        w = W1[:, k]
        ang = np.linalg.norm(w) * p.dt
        u = w / np.linalg.norm(w)
        want = (off0 * math.cos(ang) + np.cross(u, off0.T).T * math.sin(ang)
                + np.outer(u, u @ off0) * (1 - math.cos(ang)))
        got = pos1[:, sh] - X1[:, k:k + 1]
        np.testing.assert_allclose(got[:dim], want[:dim], rtol=0, atol=1e-13)
    o.V_Verlet_step2()
    for a, ax in enumerate("xyz"[:dim]):
        np.testing.assert_allclose(o.get(f"coll_v{ax}"), V1[a] + 0.5 * p.dt / M * F[a],
                                   rtol=1e-14, atol=1e-16)


def test_wall_force_is_sum_over_wall_particles(tmp_path):
    s = _sys(helpers.cfg_couette(n=32), tmp_path, noise=1e-2)
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_forces()
    t = o.get("type")
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
