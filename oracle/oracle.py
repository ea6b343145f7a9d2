"""ctypes wrapper of liboracle.so - the FP64 CPU oracle (TEST INFRASTRUCTURE).

Only tests/, the smoke check and the bench's CPU-baseline leg use this module,
and only as the checker.  Field names and lengths are those of sph.h.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPH_ORACLE_LIB selects another build (e.g. liboracle_asan.so, make asan).
LIB_PATH = os.environ.get("SPH_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
LIB_V4_PATH = os.path.join(HERE, "liboracle_v4.so")   # AVX-512 build (CPU baseline)

_LIB = None
_LIBS = {}


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def host_has_avx512():
    """True when this host's CPU has the AVX-512 subsets x86-64-v4 requires."""
    try:
        flags = open("/proc/cpuinfo").read().split("flags", 2)[1].split("\n", 1)[0].split()
    except (OSError, IndexError):
        return False
    return all(f in flags for f in ("avx512f", "avx512dq", "avx512cd", "avx512bw", "avx512vl"))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def library(path=None):
    """The oracle library (default: the x86-64-v3 build the tests use)."""
    global _LIB
    path = path or LIB_PATH
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        build_oracle()
    from sph_system import SphParams  # the shared parameter-struct layout

    lib = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [ctypes.POINTER(SphParams)]
    lib.orc_destroy.restype = None
    lib.orc_destroy.argtypes = [vp]
    for fn in ("orc_set_field", "orc_get_field"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp, ctypes.c_char_p, vp, i64]
    for fn in ("orc_get_cell_list", "orc_get_colloid_cell_list"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp, vp, vp, vp, vp]
    for fn in ("orc_neighbours_search", "orc_calculate_density", "orc_calculate_vel_gradient",
               "orc_calculate_pressure", "orc_calculate_forces", "orc_V_Verlet_step1",
               "orc_V_Verlet_step2", "orc_prepare", "orc_calculate_pair_forces",
               "orc_calculate_external_forces", "orc_calculate_dcdt",
               "orc_calculate_force_walls", "orc_calculate_force_colloids"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp]
    lib.orc_set_step.restype = i32
    lib.orc_set_step.argtypes = [vp, i32]
    lib.orc_set_slip.restype = i32
    lib.orc_set_slip.argtypes = [vp, i32]
    lib.orc_step.restype = i32
    lib.orc_step.argtypes = [vp, i32]
    lib.orc_kinetic_energy.restype = ctypes.c_double
    lib.orc_kinetic_energy.argtypes = [vp]
    lib.orc_observables.restype = i32
    lib.orc_observables.argtypes = [vp, vp]
    lib.orc_num_threads.restype = i32
    lib.orc_num_threads.argtypes = []
    lib.orc_kernel_W.restype = ctypes.c_double
    lib.orc_kernel_W.argtypes = [ctypes.POINTER(SphParams), ctypes.c_double]
    lib.orc_kernel_gradW.restype = ctypes.c_double
    lib.orc_kernel_gradW.argtypes = [ctypes.POINTER(SphParams), ctypes.c_double]
    lib.orc_density_brute_force.restype = i32
    lib.orc_density_brute_force.argtypes = [vp, vp]
    _LIBS[path] = lib
    if path == LIB_PATH:
        _LIB = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


STATE_FIELDS = (
    ["x", "y", "z", "vx", "vy", "vz", "mass", "dens", "press", "fx", "fy", "fz",
     "cxx", "cxy", "cyy", "cxz", "cyz", "czz",
     "dcdt_xx", "dcdt_xy", "dcdt_yy", "dcdt_xz", "dcdt_yz", "dcdt_zz", "type"])
SETUP_FIELDS = ["walls_list", "walls_start", "colloids_list", "colloids_start",
                "x_center", "y_center", "z_center"]
COLLOID_FIELDS = [
    "coll_x", "coll_y", "coll_z", "coll_vx", "coll_vy", "coll_vz",
    "coll_omegax", "coll_omegay", "coll_omegaz", "coll_theta",
    "coll_quat0", "coll_quatx", "coll_quaty", "coll_quatz",
    "fx_colloid", "fy_colloid", "fz_colloid", "tx_colloid", "ty_colloid", "tz_colloid"]
ORC_SLIP_EOS_NO_MASS = 1

INT_FIELDS = {"type", "walls_list", "walls_start", "colloids_list", "colloids_start"}


# Raw parameters: the values read from the input file (src/system_read_input.cpp)
# plus the sizes of the arrays the set-up produced.  Everything else in
# sph_params_t is derived from these by derive_params.
RAW_FIELDS = (
    "dim", "N", "Nmax", "wall", "N_colloids", "Nlist_walls", "Nlist_colloids",
    "ext_force_type", "coll_move", "new_sim", "Nsteps",
    "freq_micro", "freq_macro", "freq_walls", "freq_colloids",
    "L", "dt", "t0", "overlap", "rho", "c", "P0", "eta", "zeta", "np", "kT", "tau",
    "ext_force", "V_top", "V_bottom", "coll_R", "coll_rho", "coll_rep_cuton", "F0_rep",
    "tau_rep", "r0_magnet", "F0_magnet", "omega_magnet")

# Planted mistakes for the tests that show the pins catch them (never used
# otherwise): each replaces one derived constant by a plausible wrong formula.
SLIPS = ("csq_c", "rho0_sum", "nkT_sum", "a_no_zeta", "b_2d_form", "cw_other_dim",
         "c_gradw_sign", "rcut_no_overlap", "coll_I_disc", "rep_cutoff_no_tau")


def derive_params(src, slip=None):
    """Every derived field of sph_params_t recomputed from the raw ones of `src`
    by the reference's formulas, in the reference's arithmetic order
    (src/system_initialize.cpp:47-262, src/system_constructor.cpp:61-78).
    The oracle runs on these, so the checker shares no host arithmetic with the
    system under test.  `src.Nxyz[1]` is taken as the system set it (lattice
    rows including the wall layers, src/system_initialize.cpp:76-80)."""
    from sph_system import SphParams

    p = SphParams()
    for name in RAW_FIELDS:
        v = getattr(src, name)
        if hasattr(v, "_length_"):
            for k in range(len(v)):
                getattr(p, name)[k] = v[k]
        else:
            setattr(p, name, v)
    dim, wall, overlap = p.dim, p.wall, p.overlap
    L = list(p.L)
    wall_layers = 2 * math.ceil(overlap) if wall == 1 else 0
    N0, N1, N2 = src.Nxyz[0], src.Nxyz[1] - wall_layers, src.Nxyz[2]
    p.Nwalls = 2 if wall else 0                                    # system_constructor.cpp
    p.nkT = p.np * p.kT                                            # :53
    if slip == "nkT_sum":
        p.nkT = p.np + p.kT
    p.rcut = L[0] / N0 * overlap                                   # :56
    if slip == "rcut_no_overlap":
        p.rcut = L[0] / N0
    p.rcutsq = p.rcut * p.rcut                                     # :57
    p.rcut_inv = 1.0 / p.rcut                                      # :58
    dx = [L[0] / N0, L[1] / N1, (L[2] / N2) if dim == 3 else 0.0]  # :61-63
    p.ky = 2.0 * math.pi / L[1]                                    # :66
    eps = 1.0e-10                                                  # :70-74
    if abs(overlap - math.floor(overlap)) < eps:
        p.wall_width = math.floor(overlap) * dx[1]
    else:
        p.wall_width = (math.floor(overlap) + 1) * dx[1]
    p.Nxyz[0], p.Nxyz[1], p.Nxyz[2] = N0, N1 + wall_layers, N2     # :76-80
    p.y_bottom, p.y_top = (0.0, L[1]) if wall == 1 else (0.0, 0.0)
    cw_2d = (dim == 2) != (slip == "cw_other_dim")
    if cw_2d:                                                      # :82-85
        p.cw = 63.0 / (478.0 * math.pi * p.rcut * p.rcut)
    else:
        p.cw = 27.0 / (120.0 * math.pi * p.rcut * p.rcut * p.rcut)
    p.c_gradw = -15.0 * p.cw / p.rcut                              # :86
    if slip == "c_gradw_sign":
        p.c_gradw = -p.c_gradw
    p.csq = p.c * p.c                                              # :91
    if slip == "csq_c":
        p.csq = p.c
    p.rho0 = p.P0 * p.rho                                          # :93
    if slip == "rho0_sum":
        p.rho0 = p.P0 + p.rho
    p.a = 5.0 * p.eta / 3.0 - p.zeta                               # :94-101
    if slip == "a_no_zeta":
        p.a = 5.0 * p.eta / 3.0
    b_factor = 5.0 if dim == 3 else 4.0
    if slip == "b_2d_form":
        b_factor = 4.0 if dim == 3 else 5.0
    p.b = b_factor * (p.zeta + p.eta / 3.0)
    p.beta_max = 1.5                                               # :105
    p.Ncells[0] = int(L[0] / p.rcut)                               # :108-125
    p.cell_size[0] = L[0] / p.Ncells[0]
    ly = L[1] + 2 * p.wall_width if wall == 1 else L[1]
    p.Ncells[1] = int(ly / p.rcut)
    p.cell_size[1] = ly / p.Ncells[1]
    if dim == 3:
        p.Ncells[2] = int(L[2] / p.rcut)
        p.cell_size[2] = L[2] / p.Ncells[2]
    else:
        p.Ncells[2] = 1
        p.cell_size[2] = 1
    p.Ntotal_cells = p.Ncells[0] * p.Ncells[1] * p.Ncells[2]
    R = p.coll_R
    p.coll_rep_cutoff = 5.0 * R / p.tau_rep                        # :161
    if slip == "rep_cutoff_no_tau":
        p.coll_rep_cutoff = 5.0 * R
    p.r0_magnet_sq = p.r0_magnet * p.r0_magnet                     # :163
    if dim == 2:                                                   # :167-174
        p.coll_mass = math.pi * R * R * p.coll_rho
        p.coll_I = 0.5 * p.coll_mass * R * R
    else:
        p.coll_mass = 4.0 / 3.0 * math.pi * R * R * R * p.coll_rho
        p.coll_I = 2.0 / 5.0 * p.coll_mass * R * R
        if slip == "coll_I_disc":
            p.coll_I = 0.5 * p.coll_mass * R * R
    if p.N_colloids > 0:                                           # :218-242
        p.rcuton_coll_rep = 2.0 * R + p.coll_rep_cuton
        p.rcuton_coll_rep_sq = p.rcuton_coll_rep * p.rcuton_coll_rep
        p.rcutoff_coll_rep = 2.0 * R + p.coll_rep_cutoff
        p.rcutoff_coll_rep_sq = p.rcutoff_coll_rep * p.rcutoff_coll_rep
        p.rcutoff_coll = p.rcutoff_coll_rep if p.rcutoff_coll_rep > p.r0_magnet else p.r0_magnet
        p.rcutoff_coll_sq = p.rcutoff_coll * p.rcutoff_coll
        for a in range(3):
            if a < dim:
                p.Ncells_colloids[a] = int(L[a] / p.rcutoff_coll)
                p.cell_colloids_size[a] = L[a] / p.Ncells_colloids[a]
            else:
                p.Ncells_colloids[a] = 1
                p.cell_colloids_size[a] = 1
        p.Ntotal_cells_colloids = (p.Ncells_colloids[0] * p.Ncells_colloids[1]
                                   * p.Ncells_colloids[2])
    return p


# Fields derive_params computes (compared with the system's in the tests).
DERIVED_FIELDS = (
    "Nwalls", "Nxyz", "Ncells", "Ntotal_cells", "Ncells_colloids", "Ntotal_cells_colloids",
    "cell_size", "cell_colloids_size", "csq", "rho0", "a", "b", "nkT", "rcut", "rcutsq",
    "rcut_inv", "cw", "c_gradw", "ky", "wall_width", "y_bottom", "y_top", "beta_max",
    "coll_mass", "coll_I", "coll_rep_cutoff", "rcuton_coll_rep", "rcuton_coll_rep_sq",
    "rcutoff_coll_rep", "rcutoff_coll_rep_sq", "rcutoff_coll", "rcutoff_coll_sq",
    "r0_magnet_sq")


class Oracle:
    def __init__(self, params, lib_path=None):
        self.lib = library(lib_path)
        self.params = params
        self.h = self.lib.orc_create(ctypes.byref(params))

    @classmethod
    def from_system(cls, system, slip=None, lib_path=None):
        """Oracle with the current state of a sph_system.System.  Its
        parameters are the system's raw input values with every derived
        constant recomputed here (derive_params), not copied."""
        o = cls(derive_params(system.params, slip=slip), lib_path=lib_path)
        for n in STATE_FIELDS + SETUP_FIELDS:
            o.set(n, system.get(n))
        if system.params.N_colloids:
            for n in COLLOID_FIELDS:
                o.set(n, system.get(n))
        return o

    def __del__(self):
        try:
            if self.h:
                self.lib.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def length(self, name):
        p = self.params
        if name in COLLOID_FIELDS or name == "colloids_start":
            return p.N_colloids
        if name in ("fx_wall", "fy_wall", "fz_wall", "walls_start"):
            return p.Nwalls
        if name == "walls_list":
            return p.Nlist_walls
        if name in ("colloids_list", "x_center", "y_center", "z_center"):
            return p.Nlist_colloids
        return p.N

    def set(self, name, values):
        n = self.length(name)
        arr = np.ascontiguousarray(values, dtype=np.int32 if name in INT_FIELDS else np.float64)
        assert arr.shape == (n,), (name, arr.shape, n)
        rc = self.lib.orc_set_field(self.h, name.encode(), _ptr(arr), n)
        if rc:
            raise KeyError(name)

    def get(self, name):
        n = self.length(name)
        out = np.zeros(n, dtype=np.int32 if name in INT_FIELDS else np.float64)
        rc = self.lib.orc_get_field(self.h, name.encode(), _ptr(out), n)
        if rc:
            raise KeyError(name)
        return out

    def cell_list(self):
        p = self.params
        pc = np.zeros(p.N, np.int32)
        pi = np.zeros(p.N, np.int32)
        cs = np.zeros(p.Ntotal_cells, np.int32)
        ce = np.zeros(p.Ntotal_cells, np.int32)
        self.lib.orc_get_cell_list(self.h, _ptr(pc), _ptr(pi), _ptr(cs), _ptr(ce))
        return pc, pi, cs, ce

    def colloid_cell_list(self):
        p = self.params
        nc, ncc = p.N_colloids, max(p.Ntotal_cells_colloids, 0)
        cc = np.zeros(nc, np.int32)
        ci = np.zeros(nc, np.int32)
        cs = np.zeros(ncc, np.int32)
        ce = np.zeros(ncc, np.int32)
        self.lib.orc_get_colloid_cell_list(self.h, _ptr(cc), _ptr(ci), _ptr(cs), _ptr(ce))
        return cc, ci, cs, ce

    def _call(self, fn):
        rc = getattr(self.lib, fn)(self.h)
        if rc:
            raise RuntimeError(f"{fn} returned {rc}")

    def neighbours_search(self):
        self._call("orc_neighbours_search")

    def calculate_density(self):
        self._call("orc_calculate_density")

    def calculate_vel_gradient(self):
        self._call("orc_calculate_vel_gradient")

    def calculate_pressure(self):
        self._call("orc_calculate_pressure")

    def calculate_pair_forces(self):
        self._call("orc_calculate_pair_forces")

    def calculate_external_forces(self):
        self._call("orc_calculate_external_forces")

    def calculate_dcdt(self):
        self._call("orc_calculate_dcdt")

    def calculate_force_walls(self):
        self._call("orc_calculate_force_walls")

    def calculate_force_colloids(self):
        self._call("orc_calculate_force_colloids")

    def calculate_forces(self):
        self._call("orc_calculate_forces")

    def V_Verlet_step1(self):
        self._call("orc_V_Verlet_step1")

    def V_Verlet_step2(self):
        self._call("orc_V_Verlet_step2")

    def set_step(self, step):
        self.lib.orc_set_step(self.h, int(step))

    def prepare(self):
        self._call("orc_prepare")

    def step(self, n=1):
        rc = self.lib.orc_step(self.h, int(n))
        if rc:
            raise RuntimeError(f"orc_step returned {rc}")

    def set_slip(self, code):
        """Plants a mistake in the oracle's arithmetic (ORC_SLIP_* of
        sph_oracle.h) for a test that shows a pin catches it."""
        self.lib.orc_set_slip(self.h, int(code))

    def kinetic_energy(self):
        return self.lib.orc_kinetic_energy(self.h)

    def observables(self):
        out = np.zeros(8, np.float64)
        self.lib.orc_observables(self.h, _ptr(out))
        return out

    def density_brute_force(self):
        out = np.zeros(self.params.N, np.float64)
        self.lib.orc_density_brute_force(self.h, _ptr(out))
        return out


def kernel_W(params, r):
    return library().orc_kernel_W(ctypes.byref(params), float(r))


def kernel_gradW(params, r):
    return library().orc_kernel_gradW(ctypes.byref(params), float(r))


def num_threads(lib_path=None):
    return library(lib_path).orc_num_threads()
