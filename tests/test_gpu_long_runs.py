"""Long runs: statistics of the HIP path against the oracle over whole runs.

Over long runs trajectories separate (summation order, FP32 rounding), so what
is compared are time averages over the second half of the run, sampled every
50 steps:
  C1  64 x 64 Couette flow between walls at -/+0.5 (4,480 particles), 4000 steps
      of dt 5e-5 (the nominal 2000 x 1e-4 is unstable) from rest plus N(0, 1e-3)
      noise: kinetic energy and the shear stress on each wall, FP64 form within
      1e-8 and FP32 form within 2 %
  3D walls + 3 rotating magnetic colloids (20^3), 600 steps: kinetic energy,
      wall stress within 2 % and the colloids' mean spin <omega_z> within 5 %
      (FP32 form; FP64 within 1e-8)
The flows are also checked against what the physics fixes: the Couette walls
are dragged back by the fluid (stress opposite to their motion, equal and
opposite on the two walls once the start-up has passed the channel).
"""
import numpy as np
import pytest

import helpers
from helpers import relerr
import oracle
import sph_system

pytestmark = pytest.mark.gpu


def _run(obj, steps, every, is_gpu, n_colloids):
    ke, wb, wt, om = [], [], [], []
    done = 0
    while done < steps:
        obj.step(every)
        if is_gpu:
            obj.synchronize()      # reports a particle leaving the grid
        done += every
        if done <= steps // 2:
            continue
        if is_gpu:
            o = obj.observables()
            ke.append(o["kinetic_energy"])
            wb.append(o["wall_shear_bottom"])
            wt.append(o["wall_shear_top"])
        else:
            o = obj.observables()
            ke.append(o[0])
            wb.append(o[3])
            wt.append(o[4])
        if n_colloids:
            om.append(obj.get("coll_omegaz").mean())
    return dict(ke=np.mean(ke), wb=np.mean(wb), wt=np.mean(wt),
                om=np.mean(om) if om else 0.0)


def _long(cfg, tmp_path, precision, steps, name):
    s = helpers.make_system(cfg, str(tmp_path), precision, noise=1e-3, seed=1, name=name)
    o = oracle.Oracle.from_system(s)
    s.prepare()
    o.prepare()
    nc = s.params.N_colloids
    g = _run(s, steps, 50, True, nc)
    r = _run(o, steps, 50, False, nc)
    return g, r


@pytest.mark.parametrize("precision,tol", [(sph_system.FP64, 1e-8), (sph_system.FP32, 2e-2)])
def test_couette_long_run_statistics(precision, tol, tmp_path):
    g, r = _long(helpers.cfg_couette(n=64), tmp_path, precision, 4000, f"c{precision}")
    for k in ("ke", "wb", "wt"):
        assert abs(g[k] - r[k]) <= tol * abs(r[k]), (k, g[k], r[k])
    # walls at -0.5 (bottom) and +0.5 (top) are dragged back by the fluid
    assert r["wb"] > 0 > r["wt"]
    assert abs(r["wb"] + r["wt"]) < 0.05 * abs(r["wb"])


@pytest.mark.parametrize("precision,tol,tol_om",
                         [(sph_system.FP64, 1e-8, 1e-8), (sph_system.FP32, 2e-2, 5e-2)])
def test_colloid_long_run_statistics(precision, tol, tol_om, tmp_path):
    cfg = helpers.cfg_colloids_3d(n=20)
    g, r = _long(cfg, tmp_path, precision, 600, f"k{precision}")
    for k in ("ke", "wb", "wt"):
        assert abs(g[k] - r[k]) <= tol * abs(r[k]), (k, g[k], r[k])
    assert abs(g["om"] - r["om"]) <= tol_om * abs(r["om"]), (g["om"], r["om"])
    assert relerr([g["ke"]], [r["ke"]]) <= tol
