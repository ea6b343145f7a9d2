"""Python mirror of the reference's host class over the C-ABI of libsph_amd.so.

`System` has the methods of `class_system` (src/class_system.h:146-651) with the
same names and order of use - read_input, initialize, neighbours_search,
calculate_density, calculate_vel_gradient, calculate_pressure, calculate_forces,
calculate_external_forces, calculate_force_walls, calculate_force_colloids,
V_Verlet_step1, V_Verlet_step2, print_output - but the launch geometry and the
device pointers live inside the library.  Errors raise SphError carrying the
library's status code and message (the reference returns int 1 and prints).

The library must be built (build.py); there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPH_LIB selects another build of the library (e.g. the sanitizer build of
# build.py --sanitize).
LIB_PATH = os.environ.get("SPH_LIB") or os.path.join(HERE, "libsph_amd.so")

FP64, FP32 = 0, 1
STATUS = {
    0: "SPH_OK", 1: "SPH_ERR_ARG", 2: "SPH_ERR_IO", 3: "SPH_ERR_PARSE",
    4: "SPH_ERR_GEOMETRY", 5: "SPH_ERR_DEVICE", 6: "SPH_ERR_OUT_OF_GRID",
    7: "SPH_ERR_NAN", 8: "SPH_ERR_STATE", 9: "SPH_ERR_CAPACITY",
}
(OK, ERR_ARG, ERR_IO, ERR_PARSE, ERR_GEOMETRY, ERR_DEVICE, ERR_OUT_OF_GRID, ERR_NAN, ERR_STATE,
 ERR_CAPACITY) = range(10)

# The C-ABI entry points (include/sph.h).  Tests check that each is exported.
ENTRY_POINTS = [
    "sph_create", "sph_destroy", "sph_last_error", "sph_precision", "sph_read_input",
    "sph_initialize", "sph_get_params", "sph_get_field", "sph_set_field",
    "sph_get_cell_list", "sph_neighbours_search", "sph_calculate_density",
    "sph_calculate_vel_gradient", "sph_calculate_pressure",
    "sph_calculate_external_forces", "sph_calculate_force_walls",
    "sph_calculate_force_colloids", "sph_calculate_forces", "sph_V_Verlet_step1",
    "sph_V_Verlet_step2", "sph_set_step", "sph_prepare", "sph_step", "sph_synchronize",
    "sph_current_step", "sph_run", "sph_kinetic_energy", "sph_print_output",
    "sph_write_restart", "sph_checkpoint_write", "sph_checkpoint_read", "sph_set_timing",
    "sph_get_timing", "sph_get_counters",
    "sph_observables",
    "sph_batch_create", "sph_batch_destroy", "sph_batch_last_error", "sph_batch_add",
    "sph_batch_size", "sph_batch_prepare", "sph_batch_step", "sph_batch_synchronize",
    "sph_batch_current_step", "sph_batch_get_field", "sph_batch_kinetic_energy",
    "sph_batch_observables", "sph_batch_set_launch", "sph_batch_get_launch",
    "sph_batch_autotune", "sph_batch_failed", "sph_neighbours_search_noalloc",
]
# include/sph_surrogate.h (bound in surrogate.py)
SURROGATE_ENTRY_POINTS = [
    "sph_mlp_create", "sph_mlp_destroy", "sph_mlp_last_error", "sph_mlp_param_count",
    "sph_mlp_layout", "sph_mlp_set_params", "sph_mlp_get_params", "sph_mlp_forward",
    "sph_mlp_gradients", "sph_mlp_train_step", "sph_mlp_synchronize",
]
OBSERVABLES = ["kinetic_energy", "polymer_N1", "polymer_shear", "wall_shear_bottom",
               "wall_shear_top", "chain_angle", "field_angle", "phase_lag"]

PARTICLE_FIELDS = [
    "x", "y", "z", "vx", "vy", "vz", "mass", "dens", "press", "fx", "fy", "fz",
    "cxx", "cxy", "cyy", "cxz", "cyz", "czz",
    "dcdt_xx", "dcdt_xy", "dcdt_yy", "dcdt_xz", "dcdt_yz", "dcdt_zz",
    "gradvxx", "gradvxy", "gradvxz", "gradvyx", "gradvyy", "gradvyz",
    "gradvzx", "gradvzy", "gradvzz",
]
COLLOID_FIELDS = [
    "coll_x", "coll_y", "coll_z", "coll_vx", "coll_vy", "coll_vz",
    "coll_omegax", "coll_omegay", "coll_omegaz", "coll_theta",
    "coll_quat0", "coll_quatx", "coll_quaty", "coll_quatz",
    "fx_colloid", "fy_colloid", "fz_colloid", "tx_colloid", "ty_colloid", "tz_colloid",
]
WALL_FIELDS = ["fx_wall", "fy_wall", "fz_wall"]
INT_FIELDS = ["type", "walls_list", "walls_start", "colloids_list", "colloids_start"]
CENTER_FIELDS = ["x_center", "y_center", "z_center"]
PHASES = ["vv1", "neighbours", "density", "gradv_dcdt", "forces", "walls_colloids", "vv2",
          "other"]

_i32, _f64 = ctypes.c_int32, ctypes.c_double


class SphParams(ctypes.Structure):
    """Layout of sph_params_t (include/sph.h)."""
    _fields_ = (
        [(n, _i32) for n in ("dim", "N", "Nmax", "wall", "Nwalls")]
        + [("Nxyz", _i32 * 3), ("Ncells", _i32 * 3), ("Ntotal_cells", _i32),
           ("N_colloids", _i32), ("Ncells_colloids", _i32 * 3), ("Ntotal_cells_colloids", _i32),
           ("Nlist_walls", _i32), ("Nlist_colloids", _i32)]
        + [(n, _i32) for n in ("ext_force_type", "coll_move", "new_sim", "Nsteps",
                                "freq_micro", "freq_macro", "freq_walls", "freq_colloids",
                                "pad_")]
        + [("L", _f64 * 3), ("cell_size", _f64 * 3), ("cell_colloids_size", _f64 * 3)]
        + [(n, _f64) for n in ("dt", "t0", "overlap", "rho", "c", "csq", "P0", "rho0", "eta",
                                "zeta", "a", "b", "np", "kT", "nkT", "tau", "rcut", "rcutsq",
                                "rcut_inv", "cw", "c_gradw", "ky")]
        + [("ext_force", _f64 * 3)]
        + [(n, _f64) for n in ("wall_width", "y_bottom", "y_top", "V_top", "V_bottom",
                                "beta_max", "coll_R", "coll_rho", "coll_mass", "coll_I",
                                "coll_rep_cuton", "coll_rep_cutoff", "F0_rep", "tau_rep",
                                "rcuton_coll_rep", "rcuton_coll_rep_sq", "rcutoff_coll_rep",
                                "rcutoff_coll_rep_sq", "rcutoff_coll", "rcutoff_coll_sq",
                                "r0_magnet", "r0_magnet_sq", "F0_magnet", "omega_magnet")]
    )

    def as_dict(self):
        out = {}
        for name, typ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(typ, "_length_") else v
        return out


class SphError(RuntimeError):
    def __init__(self, code, message, call):
        super().__init__(f"{call}: {STATUS.get(code, code)}: {message}")
        self.code = code
        self.message = message


_LIB = None


def library(path=None):
    """Loads libsph_amd.so once; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with build.py")
    lib = ctypes.CDLL(p)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    cp = ctypes.c_char_p
    lib.sph_create.restype = vp
    lib.sph_create.argtypes = [i32]
    lib.sph_destroy.restype = None
    lib.sph_destroy.argtypes = [vp]
    lib.sph_last_error.restype = cp
    lib.sph_last_error.argtypes = [vp]
    for fn in ("sph_precision", "sph_synchronize", "sph_current_step", "sph_prepare",
               "sph_neighbours_search", "sph_calculate_density", "sph_calculate_vel_gradient",
               "sph_calculate_pressure", "sph_calculate_external_forces",
               "sph_calculate_force_walls", "sph_calculate_force_colloids",
               "sph_calculate_forces", "sph_V_Verlet_step1", "sph_V_Verlet_step2"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp]
    lib.sph_read_input.restype = i32
    lib.sph_read_input.argtypes = [vp, cp]
    lib.sph_initialize.restype = i32
    lib.sph_initialize.argtypes = [vp, cp]
    lib.sph_get_params.restype = i32
    lib.sph_get_params.argtypes = [vp, ctypes.POINTER(SphParams)]
    lib.sph_get_field.restype = i32
    lib.sph_get_field.argtypes = [vp, cp, vp, i64]
    lib.sph_set_field.restype = i32
    lib.sph_set_field.argtypes = [vp, cp, vp, i64]
    lib.sph_get_cell_list.restype = i32
    lib.sph_get_cell_list.argtypes = [vp, vp, vp, vp, vp]
    for fn in ("sph_set_step", "sph_step", "sph_set_timing"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp, i32]
    lib.sph_kinetic_energy.restype = i32
    lib.sph_kinetic_energy.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    lib.sph_print_output.restype = i32
    lib.sph_print_output.argtypes = [vp, i32, cp]
    lib.sph_run.restype = i32
    lib.sph_run.argtypes = [vp, cp]
    lib.sph_write_restart.restype = i32
    lib.sph_write_restart.argtypes = [vp, cp]
    lib.sph_checkpoint_write.restype = i32
    lib.sph_checkpoint_write.argtypes = [vp, cp]
    lib.sph_checkpoint_read.restype = i32
    lib.sph_checkpoint_read.argtypes = [vp, cp]
    lib.sph_observables.restype = i32
    lib.sph_observables.argtypes = [vp, vp]
    lib.sph_get_timing.restype = i32
    lib.sph_get_timing.argtypes = [vp, vp, vp]
    lib.sph_get_counters.restype = i32
    lib.sph_get_counters.argtypes = [vp, vp]
    lib.sph_batch_create.restype = vp
    lib.sph_batch_create.argtypes = [i32]
    lib.sph_batch_destroy.restype = None
    lib.sph_batch_destroy.argtypes = [vp]
    lib.sph_batch_last_error.restype = cp
    lib.sph_batch_last_error.argtypes = [vp]
    for fn in ("sph_batch_size", "sph_batch_prepare", "sph_batch_synchronize",
               "sph_batch_current_step"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp]
    lib.sph_batch_add.restype = i32
    lib.sph_batch_add.argtypes = [vp, vp, vp]
    lib.sph_batch_step.restype = i32
    lib.sph_batch_step.argtypes = [vp, i32]
    lib.sph_batch_get_field.restype = i32
    lib.sph_batch_get_field.argtypes = [vp, i32, cp, vp, i64]
    for fn in ("sph_batch_kinetic_energy", "sph_batch_observables", "sph_batch_failed"):
        getattr(lib, fn).restype = i32
        getattr(lib, fn).argtypes = [vp, vp]
    lib.sph_batch_set_launch.restype = i32
    lib.sph_batch_set_launch.argtypes = [vp, i32, i32]
    lib.sph_batch_get_launch.restype = i32
    lib.sph_batch_get_launch.argtypes = [vp, vp, vp]
    lib.sph_batch_autotune.restype = i32
    lib.sph_batch_autotune.argtypes = [vp, i32, vp, vp, vp]
    lib.sph_neighbours_search_noalloc.restype = i32
    lib.sph_neighbours_search_noalloc.argtypes = [ctypes.POINTER(SphParams)] + [vp] * 16
    if path is None:
        _LIB = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class System:
    """One simulation on one GPU; mirrors class_system's method sequence."""

    def __init__(self, precision=FP64, lib_path=None):
        if isinstance(precision, str):
            precision = {"fp64": FP64, "fp32": FP32}[precision.lower()]
        self.lib = library(lib_path)
        self.h = self.lib.sph_create(precision)
        if not self.h:
            raise ValueError("bad precision")
        self.precision = precision
        self._params = None

    def close(self):
        if self.h:
            self.lib.sph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- helpers ----
    def _check(self, rc, call):
        if rc != 0:
            msg = self.lib.sph_last_error(self.h).decode(errors="replace")
            raise SphError(rc, msg, call)

    def _call(self, name, *args):
        self._check(getattr(self.lib, name)(self.h, *args), name)

    # ---- set-up (class_system::read_input / initialize) ----
    def read_input(self, path):
        self._params = None
        self._check(self.lib.sph_read_input(self.h, os.fsencode(path)), "read_input")

    def initialize(self, restart_path=None):
        self._params = None
        arg = os.fsencode(restart_path) if restart_path else None
        self._check(self.lib.sph_initialize(self.h, arg), "initialize")

    @property
    def params(self):
        if self._params is None:
            p = SphParams()
            self._check(self.lib.sph_get_params(self.h, ctypes.byref(p)), "get_params")
            self._params = p
        return self._params

    @property
    def N(self):
        return self.params.N

    def field_length(self, name):
        p = self.params
        if name in PARTICLE_FIELDS or name == "type":
            return p.N
        if name in COLLOID_FIELDS or name == "colloids_start":
            return p.N_colloids
        if name in WALL_FIELDS or name == "walls_start":
            return p.Nwalls
        if name == "walls_list":
            return p.Nlist_walls
        if name in ("colloids_list",) or name in CENTER_FIELDS:
            return p.Nlist_colloids
        raise KeyError(name)

    def get(self, name):
        n = self.field_length(name)
        out = np.zeros(n, dtype=np.int32 if name in INT_FIELDS else np.float64)
        self._check(self.lib.sph_get_field(self.h, name.encode(), _ptr(out), n), f"get {name}")
        return out

    def set(self, name, values):
        n = self.field_length(name)
        arr = np.ascontiguousarray(values, dtype=np.int32 if name in INT_FIELDS else np.float64)
        if arr.shape != (n,):
            raise ValueError(f"{name}: expected {n} values, got {arr.shape}")
        self._check(self.lib.sph_set_field(self.h, name.encode(), _ptr(arr), n), f"set {name}")

    def state(self, names=None):
        names = names or PARTICLE_FIELDS + ["type"]
        return {n: self.get(n) for n in names}

    def cell_list(self):
        p = self.params
        pc = np.zeros(p.N, np.int32)
        pi = np.zeros(p.N, np.int32)
        cs = np.zeros(p.Ntotal_cells, np.int32)
        ce = np.zeros(p.Ntotal_cells, np.int32)
        self._check(self.lib.sph_get_cell_list(self.h, _ptr(pc), _ptr(pi), _ptr(cs), _ptr(ce)),
                    "get_cell_list")
        return pc, pi, cs, ce

    # ---- per-step methods ----
    def neighbours_search(self):
        self._call("sph_neighbours_search")

    def calculate_density(self):
        self._call("sph_calculate_density")

    def calculate_vel_gradient(self):
        self._call("sph_calculate_vel_gradient")

    def calculate_pressure(self):
        self._call("sph_calculate_pressure")

    def calculate_external_forces(self):
        self._call("sph_calculate_external_forces")

    def calculate_force_walls(self):
        self._call("sph_calculate_force_walls")

    def calculate_force_colloids(self):
        self._call("sph_calculate_force_colloids")

    def calculate_forces(self):
        self._call("sph_calculate_forces")

    def V_Verlet_step1(self):
        self._call("sph_V_Verlet_step1")

    def V_Verlet_step2(self):
        self._call("sph_V_Verlet_step2")

    def set_step(self, step):
        self._call("sph_set_step", int(step))

    # ---- drivers ----
    def prepare(self):
        self._call("sph_prepare")

    def step(self, nsteps=1):
        self._call("sph_step", int(nsteps))

    def synchronize(self):
        self._call("sph_synchronize")

    def current_step(self):
        return self.lib.sph_current_step(self.h)

    def run(self, out_dir):
        self._check(self.lib.sph_run(self.h, os.fsencode(out_dir)), "run")

    def kinetic_energy(self):
        ke = ctypes.c_double(0.0)
        self._check(self.lib.sph_kinetic_energy(self.h, ctypes.byref(ke)), "kinetic_energy")
        return ke.value

    def observables(self):
        """Dict of the sweep observables (see sph_observables in sph.h)."""
        out = np.zeros(8, np.float64)
        self._check(self.lib.sph_observables(self.h, _ptr(out)), "observables")
        return dict(zip(OBSERVABLES, out.tolist()))

    def print_output(self, step, out_dir):
        self._check(self.lib.sph_print_output(self.h, int(step), os.fsencode(out_dir)),
                    "print_output")

    def write_restart(self, path):
        self._check(self.lib.sph_write_restart(self.h, os.fsencode(path)), "write_restart")

    def write_checkpoint(self, path):
        """Exact binary checkpoint of the whole state and the step counter."""
        self._check(self.lib.sph_checkpoint_write(self.h, os.fsencode(path)),
                    "checkpoint_write")

    def read_checkpoint(self, path):
        """Replaces the whole state with a checkpoint (no read_input or
        initialize needed); the next step continues the saved run."""
        self._check(self.lib.sph_checkpoint_read(self.h, os.fsencode(path)), "checkpoint_read")
        self._params = None

    def set_timing(self, on=True):
        self._call("sph_set_timing", 1 if on else 0)

    def get_timing(self):
        ms = np.zeros(8, np.float64)
        calls = np.zeros(8, np.int64)
        self._check(self.lib.sph_get_timing(self.h, _ptr(ms), _ptr(calls)), "get_timing")
        return {PHASES[k]: (float(ms[k]), int(calls[k])) for k in range(8)}

    def counters(self):
        """sph_get_counters: steps run, Verlet rebuilds, neighbour pairs within
        rcut of the last density pass, Verlet-list entries of the last build."""
        out = np.zeros(4, np.int64)
        self._check(self.lib.sph_get_counters(self.h, _ptr(out)), "get_counters")
        return dict(steps=int(out[0]), rebuilds=int(out[1]), pairs=int(out[2]),
                    list_entries=int(out[3]))


class Batch:
    """Many independent simulations in one device state (sph_batch_* in sph.h).

    Add initialized `System`s, then prepare() and step(n): every phase of a
    step is one launch over all simulations, replayed as a captured graph.
    Each simulation follows exactly the numbers its own System.step would give.
    """

    def __init__(self, precision=FP64, lib_path=None):
        if isinstance(precision, str):
            precision = {"fp64": FP64, "fp32": FP32}[precision.lower()]
        self.lib = library(lib_path)
        self.h = self.lib.sph_batch_create(precision)
        if not self.h:
            raise ValueError("bad precision")
        self.precision = precision
        self.params = []

    def close(self):
        if self.h:
            self.lib.sph_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return self.lib.sph_batch_size(self.h)

    def _check(self, rc, call):
        if rc != 0:
            msg = self.lib.sph_batch_last_error(self.h).decode(errors="replace")
            raise SphError(rc, msg, call)

    def add(self, system):
        idx = ctypes.c_int(-1)
        self._check(self.lib.sph_batch_add(self.h, system.h, ctypes.byref(idx)), "batch_add")
        p = SphParams()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(system.params), ctypes.sizeof(p))
        self.params.append(p)
        return idx.value

    def prepare(self):
        self._check(self.lib.sph_batch_prepare(self.h), "batch_prepare")

    def step(self, nsteps=1):
        self._check(self.lib.sph_batch_step(self.h, int(nsteps)), "batch_step")

    def synchronize(self):
        self._check(self.lib.sph_batch_synchronize(self.h), "batch_synchronize")

    def current_step(self):
        return self.lib.sph_batch_current_step(self.h)

    def get(self, sim, name):
        p = self.params[sim]
        if name in PARTICLE_FIELDS or name == "type":
            n = p.N
        elif name in COLLOID_FIELDS:
            n = p.N_colloids
        elif name in WALL_FIELDS:
            n = p.Nwalls
        else:
            raise KeyError(name)
        out = np.zeros(n, dtype=np.int32 if name in INT_FIELDS else np.float64)
        self._check(self.lib.sph_batch_get_field(self.h, int(sim), name.encode(), _ptr(out), n),
                    f"batch get {name}")
        return out

    def kinetic_energy(self):
        out = np.zeros(len(self), np.float64)
        self._check(self.lib.sph_batch_kinetic_energy(self.h, _ptr(out)), "batch_kinetic_energy")
        return out

    def failed(self):
        """Per simulation: 0, or bit 1 (a particle left the grid) | bit 2
        (non-finite kinetic energy); sticky."""
        out = np.zeros(len(self), np.int32)
        self._check(self.lib.sph_batch_failed(self.h, _ptr(out)), "batch_failed")
        return out

    def observables(self):
        """Array [n_sims, 8] in the order of OBSERVABLES."""
        out = np.zeros((len(self), 8), np.float64)
        self._check(self.lib.sph_batch_observables(self.h, _ptr(out)), "batch_observables")
        return out

    def set_launch(self, block, xcd=True):
        self._check(self.lib.sph_batch_set_launch(self.h, int(block), 1 if xcd else 0),
                    "batch_set_launch")

    def get_launch(self):
        b, x = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.lib.sph_batch_get_launch(self.h, ctypes.byref(b), ctypes.byref(x)),
                    "batch_get_launch")
        return b.value, bool(x.value)

    def autotune(self, steps=5):
        """Times each launch shape over `steps` real steps, keeps the fastest;
        returns (block, xcd, ms_per_step)."""
        b, x, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0.0)
        self._check(self.lib.sph_batch_autotune(self.h, int(steps), ctypes.byref(b),
                                                ctypes.byref(x), ctypes.byref(ms)),
                    "batch_autotune")
        return b.value, bool(x.value), ms.value
