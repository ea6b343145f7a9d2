"""Parity with the FP64 oracle at the full size of the 1M-particle workload:
3D walls + 16 magnetic colloids on a 100^3 lattice (1,059,433 particles; the
bench's headline flow), generic state (positions jittered, c perturbed).

Same tolerances as the small cases (normwise relative error unless noted):
  cell list, sorted order, cell ranges      bit-exact
  FP32 form: dens 1e-6, gradv and dc/dt 1e-4, F 1e-3 (also after one step);
      wall and colloid sums 1e-5 of the sum of |terms|; positions after one
      step 1e-9 absolute
  FP64 form: dens, gradv, F, dc/dt 1e-12; sums 1e-13 of the sum of |terms|;
      after one step (Verlet-list summation order) positions 1e-12 absolute,
      F 1e-10

One step of the oracle at this size takes about a second on 16 threads.
"""
import numpy as np
import pytest

import helpers
import inputs
import oracle
import sph_system
from helpers import comps, relerr

pytestmark = pytest.mark.gpu


def _check_cells(s, o):
    pc, pi, cs, ce = s.cell_list()
    opc, opi, ocs, oce = o.cell_list()
    np.testing.assert_array_equal(pc, opc)
    np.testing.assert_array_equal(pi, opi)
    occupied = oce >= ocs
    np.testing.assert_array_equal(cs[occupied], ocs[occupied])
    np.testing.assert_array_equal(ce[occupied], oce[occupied])


def _check_sums(s, o, rel):
    for n, scale in helpers.sum_scales(o).items():
        d = np.abs(s.get(n) - o.get(n))
        assert np.all(d <= rel * scale + 1e-300), f"{n}: {d} vs {rel} * {scale}"


@pytest.mark.parametrize("precision,tol", [
    (sph_system.FP32, dict(dens=1e-6, gradv=1e-4, dcdt=1e-4, f=1e-3, sums=1e-5, x=1e-9,
                           f_step=1e-3)),
    (sph_system.FP64, dict(dens=1e-12, gradv=1e-12, dcdt=1e-12, f=1e-12, sums=1e-13, x=1e-12,
                           f_step=1e-10)),
], ids=["fp32", "fp64"])
def test_full_size_walls_colloids(precision, tol, tmp_path):
    cfg = inputs.walls_colloids_3d(n=100, n_colloids=16, seed=3)
    s = helpers.make_system(cfg, str(tmp_path), precision)
    assert s.N == 1059433
    helpers.perturb(s)
    o = oracle.Oracle.from_system(s)
    print(f"full size {precision}: N={s.N}", flush=True)
    s.prepare()
    o.prepare()
    _check_cells(s, o)
    for grp in ("dens", "gradv", "dcdt", "f"):
        names = ["dens"] if grp == "dens" else comps(grp, 3)
        for n in names:
            e = relerr(s.get(n), o.get(n))
            assert e <= tol[grp], f"prepare {n}: rel err {e:.3e} > {tol[grp]:.0e}"
    _check_sums(s, o, tol["sums"])
    s.step(1)
    o.step(1)
    s.synchronize()
    p = s.params
    for a, n in enumerate(comps("x", 3)):
        diff = helpers.periodic_diff(s.get(n), o.get(n), p.L[a])
        assert np.max(np.abs(diff)) <= tol["x"], n
    for n in comps("f", 3):
        e = relerr(s.get(n), o.get(n))
        assert e <= tol["f_step"], f"step {n}: rel err {e:.3e} > {tol['f_step']:.0e}"
    s.close()


def test_full_size_verlet_rebuilds_fp64(tmp_path, monkeypatch):
    """120 steps of the same flow, FP64 form, with Verlet lists of skin 0.005
    rcut (rebuilt every few steps near the walls, each rebuild reordering the
    rebuilt state in place) against a cell search every step (skin 0): the
    pairs are the same, only the summation order differs. Positions 1e-12
    absolute; v, c, F, density 1e-10 normwise."""
    cfg = inputs.walls_colloids_3d(n=100, n_colloids=16, seed=3)
    names = comps("x", 3) + comps("v", 3) + comps("c", 3) + comps("f", 3) + ["dens"]
    out = {}
    for skin in ("0.005", "0"):
        monkeypatch.setenv("SPH_VERLET_SKIN", skin)
        s = helpers.make_system(cfg, str(tmp_path), sph_system.FP64, noise=1e-2,
                                name=f"skin{skin}")
        helpers.perturb(s)
        print(f"full size verlet skin {skin}: N={s.N}", flush=True)
        s.prepare()
        x0 = [s.get(n) for n in comps("x", 3)]
        s.step(120)
        s.synchronize()
        out[skin] = {n: s.get(n) for n in names}
        L, rcut = s.params.L, s.params.rcut
        if skin != "0":
            # lists are rebuilt once a particle has moved skin / 2 from its
            # position at the last build: a largest displacement of four times
            # that means at least three rebuilds happened in these steps
            d2 = sum(helpers.periodic_diff(out[skin][n], x0[k], L[k]) ** 2
                     for k, n in enumerate(comps("x", 3)))
            assert np.sqrt(d2.max()) >= 4 * 0.5 * float(skin) * rcut
        s.close()
    a, b = out["0.005"], out["0"]
    for k, n in enumerate(comps("x", 3)):
        diff = helpers.periodic_diff(a[n], b[n], L[k])
        assert np.max(np.abs(diff)) <= 1e-12, n
    for n in names[3:]:
        e = relerr(a[n], b[n])
        assert e <= 1e-10, f"{n}: rel err {e:.3e}"


def test_full_size_sweep_batch_members_match_single_runs(tmp_path):
    """The batch of the sweep measurement at its full size: 256 2D wall +
    colloid simulations, lattice side 32-160 (seed 4), 2,839,441 particles, FP32
    form. After prepare and 20 steps (captured-graph replays with Verlet
    rebuilds), four members - the smallest, the largest and the two with the
    most colloids - equal their stand-alone runs bit for bit."""
    import sweep

    pts = sweep.sample_points(256, seed=4)
    systems = [sweep.build_system(p, 200, sph_system.FP32, str(tmp_path)) for p in pts]
    assert sum(s.N for s in systems) == 2839441
    b = sph_system.Batch(sph_system.FP32)
    for s in systems:
        b.add(s)
    sizes = [s.N for s in systems]
    ncoll = [s.params.N_colloids for s in systems]
    by_coll = sorted(range(len(systems)), key=lambda k: -ncoll[k])
    pick = list(dict.fromkeys([int(np.argmin(sizes)), int(np.argmax(sizes))] + by_coll[:2]))
    print(f"full size sweep batch: {len(systems)} sims, members {pick}", flush=True)
    b.prepare()
    b.step(20)
    b.synchronize()
    for k in pick:
        s = systems[k]
        s.prepare()
        s.step(20)
        s.synchronize()
        p = s.params
        names = (comps("x", 2) + comps("v", 2) + comps("c", 2) + comps("f", 2)
                 + ["dens", "fx_wall", "fy_wall"])
        if p.N_colloids:
            names += ["coll_x", "coll_vx", "coll_omegaz", "fx_colloid", "tz_colloid"]
        for n in names:
            np.testing.assert_array_equal(b.get(k, n), s.get(n), err_msg=f"sim {k} {n}")
    b.close()
    for s in systems:
        s.close()


def test_full_size_fp32_tracks_fp64_over_a_long_run(tmp_path):
    """The 1M-particle flow (walls at -/+0.25, 16 rotating magnetic colloids)
    for 600 steps in each form, from the same start; the FP64 form is the one
    pinned to the oracle at this size above.  Time averages over the second
    half, sampled every 50 steps: the kinetic energy, the shear stress on each
    wall and the colloids' mean spin <omega_z> within 1e-4 relative (measured:
    3.6e-6, 1.6e-6, 1.2e-6 and 2.4e-6; both forms repeat bitwise)."""
    cfg = inputs.walls_colloids_3d(n=100, n_colloids=16, seed=3)
    stats = {}
    for prec in (sph_system.FP64, sph_system.FP32):
        s = helpers.make_system(cfg, str(tmp_path), prec, name=f"long{prec}")
        s.prepare()
        ke, wb, wt, om = [], [], [], []
        for done in range(50, 601, 50):
            s.step(50)
            s.synchronize()
            if done > 300:
                o = s.observables()
                ke.append(o["kinetic_energy"])
                wb.append(o["wall_shear_bottom"])
                wt.append(o["wall_shear_top"])
                om.append(s.get("coll_omegaz").mean())
        stats[prec] = dict(ke=np.mean(ke), wb=np.mean(wb), wt=np.mean(wt), om=np.mean(om))
        s.close()
    r, g = stats[sph_system.FP64], stats[sph_system.FP32]
    print("FP32 vs FP64 time averages:",
          {k: (g[k], r[k], abs(g[k] - r[k]) / abs(r[k])) for k in r}, flush=True)
    for k in ("ke", "wb", "wt", "om"):
        assert abs(g[k] - r[k]) <= 1e-4 * abs(r[k]), (k, g[k], r[k])
    # the walls are dragged back by the fluid
    assert r["wb"] > 0 > r["wt"]
