"""sph_neighbours_search_noalloc: the allocation-free neighbour search.

Through the C-ABI, against the handle's sph_neighbours_search +
sph_get_cell_list on the same positions (bit-exact: particle_cell,
particle_index, cell_start, cell_end), and its colloid cell grid against the
oracle's (src/system_neighbours_search.cu:70-106; bit-exact).  Three calls in a
row on one GPU, each smaller than the one before (the 1M-particle walls + 16
colloids flow, a 3D walls + 3 colloids box, a 2D walls + 2 colloids box), so
that stale scratch contents from a larger call would show.  A particle outside
the grid is reported through the status word.  Sizes beyond the scratch array
are rejected before anything is launched (CPU test).

Device buffers for the test come from hipMalloc of the HIP runtime the library
itself loaded (the test allocates; the entry point does not).
"""
import ctypes

import numpy as np
import pytest

import helpers
import inputs
import oracle
import sph_system

SPH_ERR_ARG, SPH_ERR_OUT_OF_GRID, SPH_ERR_CAPACITY = 1, 6, 9


def test_rejects_sizes_beyond_the_scratch_array(tmp_path):
    lib = sph_system.library()
    s = helpers.make_system(helpers.cfg_couette(n=16), str(tmp_path), noise=0)
    p = s.params
    fake = ctypes.c_void_p(16)   # never dereferenced: the call returns first
    args = [fake] * 16
    big = sph_system.SphParams.from_buffer_copy(p)
    big.N = 1100001
    assert lib.sph_neighbours_search_noalloc(ctypes.byref(big), *args) == SPH_ERR_CAPACITY
    big = sph_system.SphParams.from_buffer_copy(p)
    big.Ntotal_cells = 65537
    assert lib.sph_neighbours_search_noalloc(ctypes.byref(big), *args) == SPH_ERR_CAPACITY
    big = sph_system.SphParams.from_buffer_copy(p)
    big.N_colloids, big.Ntotal_cells_colloids = 1001, 27
    assert lib.sph_neighbours_search_noalloc(ctypes.byref(big), *args) == SPH_ERR_CAPACITY
    assert lib.sph_neighbours_search_noalloc(ctypes.byref(p), None, *args[1:]) == SPH_ERR_ARG


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libsph_amd.so loaded."""

    H2D, D2H = 1, 2

    def __init__(self):
        sph_system.library()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, nbytes):
        ptr = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(max(int(nbytes), 4))) == 0
        self.bufs.append(ptr)
        return ptr

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        ptr = self.alloc(arr.nbytes)
        assert self.rt.hipMemcpy(ptr, arr.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_size_t(arr.nbytes), self.H2D) == 0
        return ptr

    def get(self, ptr, n, dtype=np.int32):
        out = np.empty(n, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ptr,
                                 ctypes.c_size_t(out.nbytes), self.D2H) == 0
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(hip, s, x=None):
    """One call of the entry point on the system's positions (x overrides them)."""
    lib = sph_system.library()
    p = s.params
    pos = [s.get(n) for n in ("x", "y", "z")[:p.dim]]
    if x is not None:
        pos = x
    dpos = [hip.put(a.astype(np.float64)) for a in pos] + ([None] if p.dim == 2 else [])
    N, NC = p.N, p.Ntotal_cells
    pc, pi = hip.alloc(4 * N), hip.alloc(4 * N)
    cs, ce = hip.alloc(4 * NC), hip.alloc(4 * NC)
    ncol, ncc = p.N_colloids, max(p.Ntotal_cells_colloids, 1)
    if ncol >= 2:
        dcol = [hip.put(s.get(n)) for n in ("coll_x", "coll_y", "coll_z")[:p.dim]] + (
            [None] if p.dim == 2 else [])
        ccell, cidx = hip.alloc(4 * ncol), hip.alloc(4 * ncol)
        ccs, cce = hip.alloc(4 * ncc), hip.alloc(4 * ncc)
    else:
        dcol, ccell, cidx, ccs, cce = [None] * 3, None, None, None, None
    status = hip.put(np.array([-1], np.int32))
    rc = lib.sph_neighbours_search_noalloc(ctypes.byref(p), *dpos, pc, pi, cs, ce, *dcol, ccell,
                                           cidx, ccs, cce, status, None)
    assert rc == 0
    hip.sync()
    out = dict(pc=hip.get(pc, N), pi=hip.get(pi, N), cs=hip.get(cs, NC), ce=hip.get(ce, NC),
               status=int(hip.get(status, 1)[0]))
    if ncol >= 2:
        out.update(ccell=hip.get(ccell, ncol), cidx=hip.get(cidx, ncol),
                   ccs=hip.get(ccs, p.Ntotal_cells_colloids),
                   cce=hip.get(cce, p.Ntotal_cells_colloids))
    hip.free()
    return out


CASES = [
    ("1M walls + 16 colloids", lambda: inputs.walls_colloids_3d(n=100, n_colloids=16, seed=3)),
    ("3D walls + 3 colloids", lambda: helpers.cfg_colloids_3d(n=24)),
    ("2D walls + 2 colloids", lambda: helpers.cfg_colloids_2d(n=32)),
]


@pytest.mark.gpu
def test_matches_the_handles_neighbour_search(tmp_path):
    hip = _Hip()
    sizes = []
    for k, (label, make) in enumerate(CASES):
        s = helpers.make_system(make(), str(tmp_path), name=f"in{k}")
        helpers.perturb(s, seed=30 + k)
        print(f"noalloc call {k}: {label}, N={s.N}", flush=True)
        sizes.append(s.N)
        got = _noalloc(hip, s)
        assert got["status"] == 0
        s.neighbours_search()
        pc, pi, cs, ce = s.cell_list()
        np.testing.assert_array_equal(got["pc"], pc, err_msg=label)
        np.testing.assert_array_equal(got["pi"], pi, err_msg=label)
        np.testing.assert_array_equal(got["cs"], cs, err_msg=label)
        np.testing.assert_array_equal(got["ce"], ce, err_msg=label)
        o = oracle.Oracle.from_system(s)
        o.neighbours_search()
        opc, opi, ocs, oce = o.cell_list()
        np.testing.assert_array_equal(got["pc"], opc)
        np.testing.assert_array_equal(got["pi"], opi)
        occ = oce >= ocs
        np.testing.assert_array_equal(got["cs"][occ], ocs[occ])
        np.testing.assert_array_equal(got["ce"][occ], oce[occ])
        assert np.all(got["ce"][~occ] < got["cs"][~occ])
        if s.params.N_colloids >= 2:
            occ_, oci, occs, occe = o.colloid_cell_list()
            np.testing.assert_array_equal(got["ccell"], occ_)
            np.testing.assert_array_equal(got["cidx"], oci)
            np.testing.assert_array_equal(got["ccs"], occs)
            np.testing.assert_array_equal(got["cce"], occe)
        s.close()
    assert sizes[0] > sizes[1] > sizes[2]


@pytest.mark.gpu
def test_reports_a_particle_outside_the_grid(tmp_path):
    hip = _Hip()
    s = helpers.make_system(helpers.cfg_couette(n=32), str(tmp_path))
    x, y = s.get("x"), s.get("y")
    y[7] = 3.0
    got = _noalloc(hip, s, x=[x, y])
    assert got["status"] == SPH_ERR_OUT_OF_GRID
    assert np.array_equal(np.sort(got["pi"]), np.arange(s.N))   # still a permutation
    ok = _noalloc(hip, s)
    assert ok["status"] == 0
