"""Derived constants and the equation of state, pinned without the host's arithmetic.

The oracle derives every constant of sph_params_t itself, from the raw input
values, by the reference's formulas (oracle.derive_params restating
src/system_initialize.cpp:47-262); Oracle.from_system no longer copies them from
the system under test.  These CPU tests pin that derivation and the pressure:

  * the host's derived constants equal the oracle's, field for field and bit
    for bit, on inputs where no term vanishes (P0 = 0.9, c = 3, np = 10,
    kT = 2e-3, zeta = 0.3, eta = 0.7, rho = 1.3, a non-integer overlap 2.7);
    every planted slip in the derivation breaks that equality;
  * raw values are the input file's (the parse);
  * the textbook facts the constants encode: colloid mass and moment of
    inertia of a disc / sphere of density rho_c, the wall band of ceil(overlap)
    lattice rows, cells at least rcut wide and fewer than L / rcut of them;
  * the oracle's pressure is p_i = c^2 (m_i d_i - P0 rho) (src/kernel_pressures.cu:21)
    evaluated from the input file's c, P0 and rho, and the check fails with
    c for c^2, rho0 = P0 + rho or the mass dropped (planted in the oracle);
  * the polymer modulus nkT = np kT enters the force exactly as the reference's
    stress Pi = p I - nkT c does: on a perfect periodic lattice at rest with
    c_yy linear in y, every particle away from the jump feels
    F_y = nkT alpha M_yy / d^2, M_yy = -sum_j gradW(r_ij) r_ij,y^2 / r_ij
    over the lattice offsets (the pressure and viscous parts cancel by the
    lattice symmetry); nkT = np + kT fails it.
"""
import math

import numpy as np
import pytest

import helpers
import inputs
import oracle

NONDEG = dict(P0=0.9, c=3.0, np=10.0, kT=2e-3, eta=0.7, zeta=0.3, rho=1.3, overlap=2.7)


def cfg_walls_colloids_3d():
    return inputs.walls_colloids_3d(n=20, n_colloids=2, R=0.08, seed=5, tau_repulsion=4.0,
                                    cutoff_magnetic=0.3, **NONDEG)


def cfg_colloids_2d():
    return helpers.cfg_colloids_2d(n=40, **NONDEG)


def cfg_sine_3d():
    return inputs.sine_3d(n=16, **NONDEG)


CFGS = {"walls_colloids_3d": cfg_walls_colloids_3d, "colloids_2d": cfg_colloids_2d,
        "sine_3d": cfg_sine_3d}


def _system(cfg, tmp_path, noise=0.0):
    return helpers.make_system(cfg, str(tmp_path), noise=noise)


def _field(p, name):
    v = getattr(p, name)
    return list(v) if hasattr(v, "_length_") else v


def _mismatches(host, derived):
    return [n for n in oracle.DERIVED_FIELDS if _field(host, n) != _field(derived, n)]


@pytest.mark.parametrize("case", list(CFGS))
def test_host_constants_equal_the_oracles_derivation(case, tmp_path):
    cfg = CFGS[case]()
    s = _system(cfg, tmp_path)
    host = s.params
    mine = oracle.derive_params(host)
    assert _mismatches(host, mine) == []
    # the raw values are the input file's
    for key, field in (("c", "c"), ("P0", "P0"), ("rho", "rho"), ("eta", "eta"),
                       ("zeta", "zeta"), ("np", "np"), ("kT", "kT"), ("tau", "tau"),
                       ("overlap", "overlap"), ("dt", "dt"), ("coll_R", "coll_R"),
                       ("coll_rho", "coll_rho"), ("tau_repulsion", "tau_rep"),
                       ("F0_repulsion", "F0_rep"), ("coll_repulsion_cuton", "coll_rep_cuton"),
                       ("cutoff_magnetic", "r0_magnet"), ("F0_magnetic", "F0_magnet"),
                       ("omega_magnetic", "omega_magnet"), ("Vwall_top", "V_top"),
                       ("Vwall_bottom", "V_bottom")):
        assert getattr(host, field) == cfg[key], key
    assert list(host.L)[:host.dim] == list(cfg["L"])
    assert host.N_colloids == cfg["N_colloids"]


def test_every_planted_slip_breaks_the_match(tmp_path):
    """Each plausible mistake in the derivation changes at least one derived
    field, so the equality above fails on it."""
    host = _system(cfg_walls_colloids_3d(), tmp_path).params
    for slip in oracle.SLIPS:
        bad = oracle.derive_params(host, slip=slip)
        assert _mismatches(host, bad), slip


@pytest.mark.parametrize("case", list(CFGS))
def test_constants_encode_the_textbook_facts(case, tmp_path):
    cfg = CFGS[case]()
    p = oracle.derive_params(_system(cfg, tmp_path).params)
    dim = p.dim
    n_raw = cfg["N"]
    dx = [cfg["L"][a] / n_raw[a] for a in range(dim)]
    assert p.csq == 9.0 and p.rho0 == pytest.approx(0.9 * 1.3, rel=1e-15)
    assert p.nkT == pytest.approx(10.0 * 2e-3, rel=1e-15)
    # rcut = overlap lattice spacings; the wall band is ceil(overlap) rows
    assert p.rcut == pytest.approx(2.7 * dx[0], rel=1e-15)
    assert p.wall_width == pytest.approx(math.ceil(2.7) * dx[1], rel=1e-15)
    if p.wall:
        assert p.Nxyz[1] == n_raw[1] + 2 * math.ceil(2.7)
    # cells: an integer number of them spans each axis, each at least rcut wide
    for a in range(dim):
        span = cfg["L"][a] + (2 * p.wall_width if a == 1 and p.wall else 0.0)
        assert p.cell_size[a] * p.Ncells[a] == pytest.approx(span, rel=1e-14)
        assert p.cell_size[a] >= p.rcut and (p.Ncells[a] + 1) * p.rcut > span
    # SDPD friction coefficients of the pair force a v_ij + b (e.v_ij) e: in
    # D = dim dimensions a + b / (D + 2) = 2 eta and b / (D + 2) = zeta + eta / 3
    # (shear and bulk viscosity; Espanol & Revenga 2003)
    assert p.a + p.b / (dim + 2) == pytest.approx(2 * 0.7, rel=1e-14)
    assert p.b / (dim + 2) == pytest.approx(0.3 + 0.7 / 3, rel=1e-14)
    # colloids: disc (2D) / sphere (3D) of density rho_c
    if p.N_colloids:
        R, rc = cfg["coll_R"], cfg["coll_rho"]
        vol = math.pi * R ** 2 if dim == 2 else 4.0 / 3.0 * math.pi * R ** 3
        assert p.coll_mass == pytest.approx(rc * vol, rel=1e-15)
        inertia = 0.5 * p.coll_mass * R ** 2 if dim == 2 else 0.4 * p.coll_mass * R ** 2
        assert p.coll_I == pytest.approx(inertia, rel=1e-15)
        # repulsion reaches 5 / tau_rep radii past contact; colloid cells hold
        # both ranges
        assert p.rcutoff_coll_rep - 2 * R == pytest.approx(5 * R / cfg["tau_repulsion"],
                                                           rel=1e-14)
        assert p.rcutoff_coll == max(p.rcutoff_coll_rep, cfg["cutoff_magnetic"])
        for a in range(dim):
            assert p.cell_colloids_size[a] >= p.rcutoff_coll


def _pressure_pin(o, cfg):
    """||p - c^2 (m d - P0 rho)|| / ||p|| with c, P0, rho from the input file."""
    m, d, pr = o.get("mass"), o.get("dens"), o.get("press")
    expect = cfg["c"] ** 2 * (m * d - cfg["P0"] * cfg["rho"])
    return np.linalg.norm(pr - expect) / np.linalg.norm(expect)


@pytest.mark.parametrize("case", list(CFGS))
def test_pressure_closed_form(case, tmp_path):
    cfg = CFGS[case]()
    s = _system(cfg, tmp_path, noise=1e-3)
    helpers.perturb(s, pos=0.2)
    o = oracle.Oracle.from_system(s)
    o.neighbours_search()
    o.calculate_density()
    o.calculate_pressure()
    assert _pressure_pin(o, cfg) <= 1e-15
    # the planted slips fail it
    for slip in ("csq_c", "rho0_sum"):
        q = oracle.Oracle.from_system(s, slip=slip)
        q.neighbours_search()
        q.calculate_density()
        q.calculate_pressure()
        assert _pressure_pin(q, cfg) > 1e-3, slip
    o.set_slip(oracle.ORC_SLIP_EOS_NO_MASS)
    o.calculate_pressure()
    assert _pressure_pin(o, cfg) > 1e-3
    o.set_slip(0)
    o.calculate_pressure()
    assert _pressure_pin(o, cfg) <= 1e-15


def _polymer_force_pin(s, cfg, slip=None):
    """Max relative deviation of F_y from nkT alpha M_yy / d^2 over the
    particles farther than rcut from the jump of c_yy, and max |F_x|, |F_z|
    relative to the expected magnitude."""
    alpha = 0.8
    p = oracle.derive_params(s.params)
    y = s.get("y")
    s.set("cyy", 1.0 + alpha * (y - 0.5))
    o = oracle.Oracle.from_system(s, slip=slip)
    o.neighbours_search()
    o.calculate_forces()
    d = o.get("dens")
    dx = cfg["L"][0] / cfg["N"][0]
    k = int(math.ceil(p.rcut / dx))
    off = np.array([(i, j, l) for i in range(-k, k + 1) for j in range(-k, k + 1)
                    for l in range(-k, k + 1)], float) * dx
    r = np.linalg.norm(off, axis=1)
    use = (r > 0) & (r < p.rcut)
    gw = np.array([oracle.kernel_gradW(p, rr) for rr in r[use]])
    M_yy = -np.sum(gw * off[use, 1] ** 2 / r[use])
    nkT = cfg["np"] * cfg["kT"]
    interior = np.abs(y - 0.5) < 0.5 - p.rcut - dx
    expect = nkT * alpha * M_yy / d[interior] ** 2
    fy = o.get("fy")[interior]
    dev = np.max(np.abs(fy - expect) / np.abs(expect))
    side = max(np.abs(o.get("fx")[interior]).max(), np.abs(o.get("fz")[interior]).max())
    return dev, side / np.abs(expect).max()


def test_polymer_modulus_on_a_lattice(tmp_path):
    cfg = inputs.sine_3d(n=16, ext_force_type=0, **NONDEG)
    s = _system(cfg, tmp_path)
    dev, side = _polymer_force_pin(s, cfg)
    assert dev <= 1e-9 and side <= 1e-9
    dev, _ = _polymer_force_pin(s, cfg, slip="nkT_sum")
    assert dev > 1.0
