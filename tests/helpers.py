"""Shared set-up for the tests: workloads, state transfer, error norms."""
import os

import numpy as np

import inputs
import sph_system

# Small versions of the benchmark physics used by the parity tests.
def cfg_couette(n=64, **kw):
    return inputs.couette_2d(n=n, **kw)


def cfg_sine(n=24, **kw):
    return inputs.sine_3d(n=n, **kw)


def cfg_colloids_3d(n=30, **kw):
    """Walls + 3 colloids close enough for repulsion and magnetic forces."""
    base = dict(n=n, n_colloids=3, R=0.12, tau_repulsion=30.0, coll_repulsion_cuton=0.0,
                cutoff_magnetic=0.3, F0_repulsion=2.0, F0_magnetic=0.5,
                omega_magnetic=3.0)
    base.update(kw)
    cfg = inputs.walls_colloids_3d(**base)
    if "coll_x" not in kw:
        # Colloids 0 and 1 sit 0.27 apart: inside the repulsion and magnetic ranges.
        cfg["coll_x"] = [[0.30, 0.45, 0.40], [0.57, 0.50, 0.45], [0.70, 0.30, 0.80]][
            :cfg["N_colloids"]]
    return cfg


def cfg_colloids_2d(n=48, **kw):
    cfg = inputs.couette_2d(n=n, dt=5e-5)
    cfg.update(N_colloids=2, coll_R=0.1, coll_rho=1.0, coll_move=0,
               coll_x=[[0.40, 0.45], [0.62, 0.55]], coll_repulsion_cuton=0.0,
               F0_repulsion=1.0, tau_repulsion=20.0, cutoff_magnetic=0.3, F0_magnetic=0.2,
               omega_magnetic=2.0, Vwall_bottom=-0.2, Vwall_top=0.2)
    cfg.update(kw)
    return cfg


def make_system(cfg, tmpdir, precision=sph_system.FP64, noise=1e-3, seed=1, name="input"):
    path = os.path.join(tmpdir, name)
    inputs.write_input(path, cfg)
    s = sph_system.System(precision)
    s.read_input(path)
    s.initialize()
    if noise:
        inputs.seeded_velocity(s, noise, seed)
    return s


def perturb(s, seed=11, pos=0.1, conf=0.05):
    """Generic state: fluid positions jittered by U(-pos, pos) lattice spacings and
    a random symmetric perturbation of the conformation tensor (so every stress
    component, including the off-diagonal ones, enters the forces)."""
    rng = np.random.default_rng(seed)
    p = s.params
    t = s.get("type")
    fl = t == 0
    dx = p.L[0] / p.Nxyz[0]
    for a, n in enumerate(["x", "y", "z"][:p.dim]):
        v = s.get(n)
        v[fl] += rng.uniform(-pos, pos, fl.sum()) * dx
        if not (a == 1 and p.wall):
            v = np.mod(v, p.L[a])
        s.set(n, v)
    names = ["cxx", "cxy", "cyy"] + (["cxz", "cyz", "czz"] if p.dim == 3 else [])
    for n in names:
        c = s.get(n)
        c[fl] += rng.uniform(-conf, conf, fl.sum())
        s.set(n, c)


def sum_scales(o):
    """Sum of |terms| of every wall and colloid reduction, from the oracle's
    state: the scale against which a change of summation order is measured
    (these sums cancel strongly, e.g. the torque of near-radial forces)."""
    p = o.params
    out = {}
    F = [o.get(n) for n in ("fx", "fy", "fz")]
    t = o.get("type")
    if p.wall:
        for a, n in enumerate(("fx_wall", "fy_wall", "fz_wall")):
            out[n] = np.array([np.abs(F[a][t == w + 1]).sum() for w in range(p.Nwalls)])
    if p.N_colloids:
        X = [o.get(n) for n in ("x", "y", "z")]
        C = [o.get(n) for n in ("coll_x", "coll_y", "coll_z")]
        Fc = [o.get(n) for n in ("fx_colloid", "fy_colloid", "fz_colloid")]
        sc = {n: np.zeros(p.N_colloids) for n in
              ("fx_colloid", "fy_colloid", "fz_colloid", "tx_colloid", "ty_colloid", "tz_colloid")}
        for k in range(p.N_colloids):
            sh = t == 3 + k
            r = []
            for a in range(3):
                d = X[a][sh] - C[a][k]
                if not (a == 1 and p.wall):
                    d = d - p.L[a] * np.round(d / p.L[a])
                r.append(d if (a < p.dim) else 0 * d)
            f = [F[a][sh] for a in range(3)]
            for a, n in enumerate(("fx_colloid", "fy_colloid", "fz_colloid")):
                sc[n][k] = np.abs(f[a]).sum() + abs(Fc[a][k])
            sc["tx_colloid"][k] = np.abs(r[1] * f[2]).sum() + np.abs(r[2] * f[1]).sum()
            sc["ty_colloid"][k] = np.abs(r[2] * f[0]).sum() + np.abs(r[0] * f[2]).sum()
            sc["tz_colloid"][k] = np.abs(r[0] * f[1]).sum() + np.abs(r[1] * f[0]).sum()
        out.update(sc)
    return out


def relerr(a, b):
    """Normwise relative error ||a - b|| / ||b|| (0 when both vanish)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    na = np.linalg.norm(a - b)
    if nb == 0.0:
        return na
    return na / nb


def vec_relerr(getter_a, getter_b, names):
    a = np.concatenate([getter_a(n) for n in names])
    b = np.concatenate([getter_b(n) for n in names])
    return relerr(a, b)


def comps(prefix, dim):
    if prefix == "f":
        return ["fx", "fy", "fz"][:dim]
    if prefix == "v":
        return ["vx", "vy", "vz"][:dim]
    if prefix == "x":
        return ["x", "y", "z"][:dim]
    if prefix == "c":
        return ["cxx", "cxy", "cyy"] + (["cxz", "cyz", "czz"] if dim == 3 else [])
    if prefix == "dcdt":
        return ["dcdt_xx", "dcdt_xy", "dcdt_yy"] + (
            ["dcdt_xz", "dcdt_yz", "dcdt_zz"] if dim == 3 else [])
    if prefix == "gradv":
        ax = "xyz"[:dim]
        return [f"gradv{a}{b}" for a in ax for b in ax]
    raise KeyError(prefix)


def periodic_diff(a, b, L):
    d = np.asarray(a) - np.asarray(b)
    return d - L * np.round(d / L)
