"""Parameter sweeps of the 2D wall + magnetic-colloid flow, run as batches.

A sweep point is one independent simulation: a 2D box of n x n lattice sites
between walls moving at -V and +V, an Oldroyd-B fluid with relaxation time tau
(eta = 1, np = 100, kT = 1e-3), and k magnetic colloids of radius 0.06 placed by
seeded rejection sampling (at least 2.2 R apart and 4 R from the walls), driven
by a field rotating at omega with dipole strength F0.  Points of uneven size are grouped into batches (sph_batch_* in
sph.h); each point's observables are averaged over the second half of its run:

  wall_stress   shear stress on the walls, (F_x,bottom - F_x,top) / (2 L0)
  N1            first normal-stress difference of the polymer, nkT <cxx - cyy>
  phase_lag     lag of the colloid chain angle behind the field angle (mod pi;
                circular mean of 2*lag, halved); NaN when no pair of colloids
                is within the magnetic cut-off during the averaging window

The dimensionless inputs of a point (what the surrogate model maps from):
  Wi    Weissenberg number tau * gamma_dot, gamma_dot = 2 V / L1
  Mn    Mason number eta gamma_dot R (2R)^4 / F0: viscous over magnetic force
        at contact (the 2D dipole force is F0 / r^4)
  phi   colloid area fraction k pi R^2 / (L0 L1)
  V     wall speed
"""
import math
import os
import tempfile
import time

import numpy as np

import inputs
import sph_system

R_COLL = 0.06
TARGETS = ["wall_stress", "N1", "phase_lag"]
FEATURES = ["Wi", "Mn", "phi", "V"]


def sample_points(n_points, seed, n_range=(32, 160), v_range=(0.0, 1.0),
                  tau_range=(0.1, 10.0), colloid_range=(0, 12), omega_range=(0.5, 8.0),
                  f0_range=(1e-6, 1e-4)):
    """Seeded sweep points: side n uniform in n_range, V uniform, tau and F0
    log-uniform, colloid count uniform in colloid_range, omega uniform."""
    rng = np.random.default_rng(seed)
    pts = []
    for k in range(n_points):
        pts.append(dict(
            n=int(rng.integers(n_range[0], n_range[1] + 1)),
            V=float(rng.uniform(*v_range)),
            tau=float(math.exp(rng.uniform(math.log(tau_range[0]), math.log(tau_range[1])))),
            n_colloids=int(rng.integers(colloid_range[0], colloid_range[1] + 1)),
            omega=float(rng.uniform(*omega_range)),
            F0=float(math.exp(rng.uniform(math.log(f0_range[0]), math.log(f0_range[1])))),
            seed=int(rng.integers(0, 2**31 - 1)),
        ))
    return pts


DT = 1.6e-5
# Colloid centres start at least 4 R = 0.24 from a wall: the reference has no
# colloid-wall repulsion, and a rotating chain that reaches a wall pushes its
# shell particles through the wall band (the run then fails).
WALL_GAP = 4.0


def point_cfg(pt, steps, dt=DT):
    """Input of one sweep point.  Every point uses the same dt, so a run of
    `steps` steps covers the same physical time whatever its resolution.  The
    default 1.6e-5 gives the finest lattice (n = 160) the viscous number
    nu dt / dx^2 = 0.41 of the 64 x 64 Couette flow at dt = 1e-4 (twice that,
    dt = 3.2e-5, blows up within 10 steps at n = 160)."""
    n = pt["n"]
    cfg = inputs.couette_2d(n=n, steps=steps, dt=dt, tau=pt["tau"],
                            Vwall_bottom=-pt["V"], Vwall_top=pt["V"])
    k = pt["n_colloids"]
    cfg.update(N_colloids=k, coll_R=R_COLL, coll_rho=1.0, coll_move=0,
               coll_x=inputs.place_colloids(k, R_COLL, [1.0, 1.0, 1.0], 2, True, pt["seed"],
                                            wall_gap=WALL_GAP) if k else [],
               coll_repulsion_cuton=2 * R_COLL, F0_repulsion=1.0,
               tau_repulsion=5.0, cutoff_magnetic=0.3, F0_magnetic=pt["F0"],
               omega_magnetic=pt["omega"])
    return cfg


def features(pt):
    """(Wi, Mn, phi, V) of a point, with its colloid count as built."""
    gdot = 2.0 * pt["V"] / 1.0
    wi = pt["tau"] * gdot
    mn = 1.0 * gdot * R_COLL * (2 * R_COLL) ** 4 / pt["F0"]
    phi = pt["n_colloids"] * math.pi * R_COLL ** 2
    return [wi, mn, phi, pt["V"]]


def build_system(pt, steps, precision, tmpdir, noise_seed=None):
    """Initialized System of a point.  The reference sizes its particle arrays
    at 1.01 N when colloids are present (src/system_constructor.cpp:61-78), so a
    small box cannot hold many colloid shells: the colloid count is lowered
    until the set-up fits, and pt["n_colloids"] records what was built."""
    while True:
        cfg = point_cfg(pt, steps)
        path = os.path.join(tmpdir, f"sweep_{pt['seed']}_{pt['n']}_{pt['n_colloids']}")
        inputs.write_input(path, cfg)
        s = sph_system.System(precision)
        s.read_input(path)
        try:
            s.initialize()
        except sph_system.SphError as e:
            if "N > Nmax" in e.message and pt["n_colloids"] > 0:
                pt["n_colloids"] -= 1
                s.close()
                continue
            raise
        inputs.seeded_velocity(s, 1e-3, pt["seed"] if noise_seed is None else noise_seed)
        return s


def resolve(points, steps=100, precision=sph_system.FP32):
    """Fixes every point's colloid count to what its set-up can hold (see
    build_system), on the host only; returns the points."""
    with tempfile.TemporaryDirectory() as tmpdir:
        for pt in points:
            build_system(pt, steps, precision, tmpdir).close()
    return points


def _circ_half(angles):
    a = np.asarray(angles, np.float64)
    a = a[np.isfinite(a)]
    if a.size == 0:
        return float("nan")
    return 0.5 * math.atan2(np.sin(2 * a).mean(), np.cos(2 * a).mean())


def unsteady(wall, eta, dV, L1, factor=1.0):
    """True when the wall-stress samples `wall` of a run are not those of a
    steady sheared flow: their standard deviation exceeds `factor` times the
    Newtonian wall stress eta |dV| / L1 of the walls' relative speed dV (a
    colloid configuration gone unstable swings the stress to either sign by
    many times its scale; a steady run of the sweep fluctuates by a few per
    cent of it)."""
    return bool(np.std(np.asarray(wall, np.float64)) > factor * eta * abs(dV) / L1)


def _run_batch(b, g, steps, samples, autotune, targets, info, log, unstable_factor=1.0):
    n_part = sum(p.N for p in b.params)
    t0 = time.perf_counter()
    b.prepare()
    if autotune:
        b.autotune(steps=2)
    done = b.current_step()
    first = max(done, steps // 2)
    if first > done:
        b.step(first - done)
    every = max(1, (steps - first) // samples)
    obs = []
    k = first
    while k < steps:
        m = min(every, steps - k)
        b.step(m)
        k += m
        obs.append(b.observables())
    b.synchronize()
    failed = b.failed()
    dt = time.perf_counter() - t0
    info["batches"] += 1
    info["seconds"] += dt
    info["particle_steps"] += n_part * steps
    info["particles"].append(n_part)
    o = np.stack(obs)                  # [samples, sims, 8]
    for j, i in enumerate(g):
        if failed[j]:
            info["failed"] += 1
            continue
        wall = 0.5 * (o[:, j, 3] - o[:, j, 4])
        p = b.params[j]
        if unsteady(wall, p.eta, p.V_top - p.V_bottom, p.L[1], unstable_factor):
            info["failed"] += 1
            info["unstable"] += 1
            continue
        targets[i, 0] = float(np.mean(wall))
        # how settled the average is: the two halves of the window, relative
        h = len(wall) // 2
        if h:
            info["settle"][i] = float(abs(np.mean(wall[h:]) - np.mean(wall[:h]))
                                      / max(abs(targets[i, 0]), 1e-300))
        targets[i, 1] = float(np.mean(o[:, j, 1]))
        targets[i, 2] = _circ_half(o[:, j, 7])
    if log:
        log(f"batch of {len(g)} sims, {n_part} particles, {steps} steps: {dt:.2f} s")


def run_points(points, steps, precision=sph_system.FP32, max_particles=4_000_000,
               samples=20, autotune=False, log=None, unstable_factor=1.0):
    """Runs every point for `steps` steps in batches of at most `max_particles`
    particles and returns (targets [n, 3], info).  Observables are sampled
    `samples` times over the second half of the run and averaged.  A point
    whose simulation fails (a particle leaves the grid - e.g. a colloid driven
    into a wall - or its energy is not finite) gets NaN targets and is counted
    in info["failed"]; the rest of its batch is unaffected.  So does a point
    whose wall-stress samples are not those of a steady sheared flow: their
    standard deviation exceeds `unstable_factor` times the Newtonian wall
    stress eta (V_top - V_bottom) / L1 (counted in info["unstable"] as well;
    these are runs whose colloids went unstable - the stress swings to either
    sign by many times its scale - while no particle has left the grid yet).
    info["seconds"]
    is the device time of the batches (upload, steps, observables), without
    the host set-up of the inputs."""
    targets = np.full((len(points), 3), np.nan)
    info = dict(batches=0, particle_steps=0, seconds=0.0, particles=[], failed=0, unstable=0,
                settle=np.full(len(points), np.nan))
    order = sorted(range(len(points)), key=lambda i: -points[i]["n"])
    with tempfile.TemporaryDirectory() as tmpdir:
        b, g, total = None, [], 0
        for i in order:
            s = build_system(points[i], steps, precision, tmpdir)
            if g and total + s.N > max_particles:
                _run_batch(b, g, steps, samples, autotune, targets, info, log, unstable_factor)
                b.close()
                b, g, total = None, [], 0
            if b is None:
                b = sph_system.Batch(precision)
            b.add(s)
            g.append(i)
            total += s.N
            s.close()
        if g:
            _run_batch(b, g, steps, samples, autotune, targets, info, log, unstable_factor)
            b.close()
    return targets, info
