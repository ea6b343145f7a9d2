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
