"""ctypes wrapper of liboracle.so - the FP64 CPU oracle (TEST INFRASTRUCTURE).

Only tests/, the smoke check and the bench's CPU-baseline leg use this module,
and only as the checker.  Field names and lengths are those of sph.h.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_LIB = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def library():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build_oracle()
    from sph_system import SphParams  # the shared parameter-struct layout

    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [ctypes.POINTER(SphParams)]
    lib.orc_destroy.restype = None
    lib.orc_destroy.argtypes = [vp]
    for fn in ("orc_set_field", "orc_get_field"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp, ctypes.c_char_p, vp, i64]
    lib.orc_get_cell_list.restype = i32
    lib.orc_get_cell_list.argtypes = [vp, vp, vp, vp, vp]
    for fn in ("orc_neighbours_search", "orc_calculate_density", "orc_calculate_vel_gradient",
               "orc_calculate_pressure", "orc_calculate_forces", "orc_V_Verlet_step1",
               "orc_V_Verlet_step2", "orc_prepare", "orc_calculate_pair_forces",
               "orc_calculate_external_forces", "orc_calculate_dcdt",
               "orc_calculate_force_walls", "orc_calculate_force_colloids"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp]
    lib.orc_set_step.restype = i32
    lib.orc_set_step.argtypes = [vp, i32]
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
INT_FIELDS = {"type", "walls_list", "walls_start", "colloids_list", "colloids_start"}


class Oracle:
    def __init__(self, params):
        self.lib = library()
        self.params = params
        self.h = self.lib.orc_create(ctypes.byref(params))

    @classmethod
    def from_system(cls, system):
        """Oracle with the parameters and current state of a sph_system.System."""
        o = cls(system.params)
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


def num_threads():
    return library().orc_num_threads()
