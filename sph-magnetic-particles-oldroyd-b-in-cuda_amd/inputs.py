"""Input files in the reference's key/value format and the named workloads.

The grammar is the one read by class_system::read_input
(src/system_read_input.cpp:102-269): one `key value...` per line, every key
mandatory, `coll_x` once per colloid after `N_colloids`.

Workloads (physics as specified for this solver's benchmarks):
  couette_2d        64x64 2D Couette flow between walls moving at -/+0.5
  sine_3d           64^3 periodic 3D flow forced by f_x = sin(2 pi y / L)
  walls_colloids_3d 100^3 wall-bounded 3D flow with 16 rotating magnetic colloids
Smaller variants of the same physics are used by the parity tests.
"""
import math

import numpy as np

KEY_ORDER = [
    "N", "L", "dim", "dt", "t0", "Nsteps", "overlap", "rho", "c", "P0", "eta", "zeta",
    "np", "kT", "tau", "ext_force_type", "ext_force", "freq_micro", "freq_macro",
    "freq_walls", "freq_colloids", "wall", "Vwall_bottom", "Vwall_top", "N_colloids",
    "coll_R", "coll_rho", "coll_move", "coll_x", "coll_repulsion_cuton", "F0_repulsion",
    "tau_repulsion", "cutoff_magnetic", "F0_magnetic", "omega_magnetic", "new_sim",
]

BASE = dict(
    t0=0.0, overlap=3.0, rho=1.0, c=20.0, P0=0.0, eta=1.0, zeta=0.0, np=100.0, kT=1e-3,
    ext_force_type=0, freq_micro=10**9, freq_macro=10**9, freq_walls=10**9,
    freq_colloids=10**9, wall=0, Vwall_bottom=0.0, Vwall_top=0.0, N_colloids=0,
    coll_R=0.05, coll_rho=1.0, coll_move=0, coll_x=[], coll_repulsion_cuton=0.0,
    F0_repulsion=1.0, tau_repulsion=1.0, cutoff_magnetic=0.1, F0_magnetic=0.0,
    omega_magnetic=0.0, new_sim=0,
)


def _fmt(v):
    if isinstance(v, (list, tuple)):
        return " ".join(_fmt(x) for x in v)
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def write_input(path, cfg):
    """Writes `cfg` (dict of the reference's keys) as an input file."""
    lines = ["# SPH Oldroyd-B input"]
    for k in KEY_ORDER:
        if k == "coll_x":
            for c in cfg.get("coll_x", []):
                lines.append("coll_x " + _fmt(list(c)))
            continue
        if k not in cfg:
            raise KeyError(f"missing key {k}")
        lines.append(f"{k} {_fmt(cfg[k])}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def place_colloids(n, R, L, dim, wall, seed, min_sep=2.2, wall_gap=1.5, max_tries=100000):
    """Uniform centres from `seed`, rejected if closer than min_sep*R to another
    centre (minimum image in periodic directions) or within wall_gap*R of a wall."""
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("colloid placement failed")
        c = rng.uniform(0.0, 1.0, size=dim) * np.asarray(L[:dim])
        if wall and (c[1] < wall_gap * R or c[1] > L[1] - wall_gap * R):
            continue
        ok = True
        for o in out:
            d = c - o
            for a in range(dim):
                if a == 1 and wall:
                    continue
                d[a] -= L[a] * round(d[a] / L[a])
            if math.sqrt(float(np.dot(d, d))) < min_sep * R:
                ok = False
                break
        if ok:
            out.append(c)
    return [list(map(float, c)) for c in out]


def couette_2d(n=64, steps=4000, dt=5e-5, **kw):
    """2D Couette flow between walls moving at -/+0.5.  dt is 5e-5 instead of
    the nominal 1e-4, with twice the steps for the same physical time: at 1e-4
    the 64 x 64 flow (viscous number nu dt / dx^2 = 0.41 next to Morris walls)
    blows up after ~300 steps - in the FP64 oracle restating the reference as in
    the HIP path - while 5e-5 and 2.5e-5 stay stable and agree."""
    cfg = dict(BASE)
    cfg.update(N=[n, n], L=[1.0, 1.0], dim=2, dt=dt, Nsteps=steps, tau=1.0,
               ext_force=[0.0, 0.0], wall=1, Vwall_bottom=-0.5, Vwall_top=0.5)
    cfg.update(kw)
    return cfg


def sine_3d(n=64, steps=5000, dt=5e-5, **kw):
    cfg = dict(BASE)
    cfg.update(N=[n, n, n], L=[1.0, 1.0, 1.0], dim=3, dt=dt, Nsteps=steps, tau=0.5,
               ext_force_type=2, ext_force=[1.0, 0.0, 0.0])
    cfg.update(kw)
    return cfg


def walls_colloids_3d(n=100, n_colloids=16, R=0.05, seed=3, steps=2000, dt=2e-5,
                      tau_repulsion=1.25, F0_magnetic=1e-3, **kw):
    """The 1M-particle wall + magnetic colloid flow.  Two parameters differ from
    the nominal set, because the nominal values cannot run:
      * tau_repulsion = 1.25 instead of 0.01: with 0.01 the repulsion cut-off
        2R + 5R/tau is 25 box lengths and the colloid cell grid check
        (src/system_initialize.cpp:243-261) rejects the input; 1.25 gives a
        cut-off of 0.3 = cutoff_magnetic.
      * F0_magnetic = 1e-3 instead of 1: with 1 the dipole force F0/r^4 at the
        seeded spacing (>= 0.135) is ~150 on a colloid of mass 5e-4, which drives
        the colloids far past the sound speed c = 20 within ~20 steps (the
        dynamics diverge, in the reference as here); 1e-3 gives a Stokes
        terminal velocity ~0.16 (Mach ~0.01)."""
    cfg = dict(BASE)
    L = [1.0, 1.0, 1.0]
    cfg.update(N=[n, n, n], L=L, dim=3, dt=dt, Nsteps=steps, tau=1.0, ext_force=[0.0, 0.0, 0.0],
               wall=1, Vwall_bottom=-0.25, Vwall_top=0.25, N_colloids=n_colloids, coll_R=R,
               coll_rho=1.0, coll_move=0,
               coll_x=place_colloids(n_colloids, R, L, 3, True, seed),
               coll_repulsion_cuton=2 * R, F0_repulsion=10.0, tau_repulsion=tau_repulsion,
               cutoff_magnetic=0.3, F0_magnetic=F0_magnetic, omega_magnetic=2 * math.pi)
    cfg.update(kw)
    return cfg


def seeded_velocity(system, sigma, seed):
    """Adds N(0, sigma^2) noise to every velocity component of fluid particles."""
    rng = np.random.default_rng(seed)
    t = system.get("type")
    dim = system.params.dim
    names = ["vx", "vy", "vz"][:dim]
    for n in names:
        v = system.get(n)
        noise = rng.normal(0.0, sigma, size=v.shape)
        v = np.where(t == 0, v + noise, v)
        system.set(n, v)
