"""ctypes mirror of the gdon3d C API exported by lib/libgdon3d_hip.so.

The structs are field-for-field the reference's gdn_mesh
(src/mesh/mesh.h:15-45), gdn_model (src/models/model.h:41-62) and
gdn_simulation (src/simulation/simulation.h:17-38); the functions keep the
reference names (mesh_read_from_msh22_file, simulation_init, dg_solve, ...).
This module only marshals arguments: every computation runs in the C/HIP
library, which aborts with a message when no GPU is present.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GDN_LIB: another build of the same library (the host-sanitizer build,
# `make sanitize`: lib-san/libgdon3d_hip.so)
LIB_PATH = os.environ.get("GDN_LIB") or os.path.join(HERE, "lib", "libgdon3d_hip.so")

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)


class GdnMesh(C.Structure):
    _fields_ = [
        ("file_name", C.c_char * 1024),
        ("nbelems", C.c_int),
        ("nbnodes", C.c_int),
        ("nbqnodes_in", C.c_int),
        ("nbqnodes_out", C.c_int),
        ("nbfaces", C.c_int),
        ("nboundaryfaces", C.c_int),
        ("boundaryface", c_int_p),
        ("nmacrointerfaces", C.c_int),
        ("macrointerface", c_int_p),
        ("is_read", C.c_bool),
        ("is_built", C.c_bool),
        ("has_physical_groups", C.c_bool),
        ("physical_group_is_alloc", C.c_bool),
        ("phys_group_id_is_alloc", C.c_bool),
        ("nb_phys_groups", C.c_int),
        ("phys_group_id", c_int_p),
        ("physical_group", c_int_p),
        ("elem2node", c_int_p),
        ("elem2qnode", c_int_p),
        ("elem2elem", c_int_p),
        ("face2elem", c_int_p),
        ("fnode2enode", c_int_p),
        ("max_node2elem", C.c_int),
        ("node2elem", c_int_p),
        ("node", c_double_p),
        ("qnode", c_double_p),
        ("xmin", C.c_double * 3),
        ("xmax", C.c_double * 3),
    ]


class GdnModel(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("nb_v", C.c_int),
        ("nb_w", C.c_int),
        ("vi", c_double_p),
        ("lambda_", C.c_double),
        ("get_num_flux", C.c_void_p),
        ("get_num_flux_boundary", C.c_void_p),
        ("get_imposed_data", C.c_void_p),
        ("get_proj_pos", C.c_void_p),
        ("get_proj_neg", C.c_void_p),
        ("use_relax_scheme", C.c_bool),
        ("omega", C.c_double),
        ("get_w", C.c_void_p),
        ("get_feq", C.c_void_p),
        ("get_fR", C.c_void_p),
    ]


class GdnSimulation(C.Structure):
    _fields_ = [
        ("m", C.c_int),
        ("neq", C.c_int),
        ("wlen", C.c_int),
        ("itermax", C.c_int),
        ("use_kinetic_scheme", C.c_bool),
        ("t", C.c_double),
        ("tmax", C.c_double),
        ("dt", C.c_double),
        ("cfl", C.c_double),
        ("hmin", C.c_double),
        ("cpu_running_t", C.c_double),
        ("cpu_assembly_t", C.c_double),
        ("wn", c_double_p),
        ("mdl", C.POINTER(GdnModel)),
        ("tt", C.POINTER(GdnMesh)),
    ]


# signature of gdn_get_imposed_data, for callbacks written in Python
IMPOSED_CB = C.CFUNCTYPE(None, c_double_p, C.c_double, c_double_p)

_lib = None


def load_library(path=LIB_PATH):
    """Load the library (once). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (__graft_entry__.build())")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "mesh_read_from_msh22_file": (None, [P(GdnMesh), C.c_char_p]),
        "mesh_build_connectivity": (None, [P(GdnMesh)]),
        "mesh_kuhn_cube": (None, [P(GdnMesh), C.c_int, C.c_double, C.c_uint64]),
        "mesh_kuhn_cube_write_msh22": (C.c_int, [C.c_char_p, C.c_int, C.c_double, C.c_uint64]),
        "mesh_write_msh22_file": (C.c_int, [P(GdnMesh), C.c_char_p]),
        "mesh_get_t4_hmin": (C.c_double, [P(GdnMesh)]),
        "mesh_get_elem_det": (None, [P(GdnMesh), C.c_int, c_double_p]),
        "mesh_get_elem_codtau": (None, [P(GdnMesh), C.c_int, c_double_p]),
        "mesh_get_face_vn": (None, [P(GdnMesh), C.c_int, c_double_p]),
        "gdn_gdn_mesh_free": (None, [P(GdnMesh)]),
        "simulation_init": (None, [P(GdnMesh), P(GdnModel), P(GdnSimulation)]),
        "simulation_free": (None, [P(GdnSimulation)]),
        "simulation_set_time_parameters": (None, [P(GdnSimulation), C.c_double, C.c_double, C.c_double]),
        "simulation_error_l2": (C.c_double, [P(GdnSimulation)]),
        "dg_solve": (None, [P(GdnSimulation)]),
        "dg_solve_rk3": (None, [P(GdnSimulation)]),
        "dg_solve_relaxation": (None, [P(GdnSimulation)]),
        "dg_init_sol_macro": (None, [P(GdnSimulation)]),
        "dg_update_boundary": (None, [P(GdnSimulation), C.c_double]),
        "dg_apply_relaxation": (None, [P(GdnSimulation)]),
        "d3q4_set_model": (None, [C.c_void_p]),
        "d3q6_set_model": (None, [C.c_void_p]),
        "maxwell_plane_wave_set": (None, [c_double_p, c_double_p, C.c_double]),
        "maxwell_cavity_mode_set": (None, [C.c_int, C.c_int, C.c_int, c_double_p]),
        "maxwell_num_flux_upwind": (None, [c_double_p, c_double_p, c_double_p, c_double_p]),
        "maxwell_num_flux_boundary_metal": (None, [c_double_p, c_double_p, c_double_p, c_double_p]),
        "gdh_device_count": (C.c_int, []),
        "gdh_copy_calibration": (C.c_double, [C.c_size_t, C.c_int]),
        "gdn_upwind_classes": (C.POINTER(C.c_uint8), [P(GdnMesh), P(GdnModel)]),
        "dg_stepper_create": (C.c_void_p, [P(GdnSimulation)]),
        "dg_stepper_create_precision": (C.c_void_p, [P(GdnSimulation), C.c_int]),
        "dg_stepper_init": (None, [C.c_void_p]),
        "dg_stepper_set_permittivity": (None, [C.c_void_p, c_double_p]),
        "maxwell_plane_wave_set_eps": (None, [c_double_p, c_double_p, C.c_double, C.c_double]),
        "maxwell_gaussian_pulse_set": (None, [c_double_p, C.c_double, C.c_double]),
        "dg_stepper_advance": (None, [C.c_void_p, C.c_int]),
        "dg_stepper_time": (C.c_double, [C.c_void_p]),
        "dg_stepper_iterations": (C.c_int, [C.c_void_p]),
        "dg_stepper_levels": (C.c_int, [C.c_void_p]),
        "dg_stepper_synchronize": (None, [C.c_void_p]),
        "dg_stepper_error_l2": (C.c_double, [C.c_void_p]),
        "dg_stepper_energy": (C.c_double, [C.c_void_p]),
        "gdh_energy": (C.c_double, [C.c_void_p]),
        "gdh_energy_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
        "gdh_energy_dev_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p]),
        "gdh_upload_state": (None, [C.c_void_p, c_double_p]),
        "gdh_download_state": (None, [C.c_void_p, c_double_p]),
        "dg_stepper_finish": (None, [C.c_void_p]),
        "dg_stepper_profile": (None, [C.c_void_p, C.c_int, c_double_p]),
        "dg_stepper_free": (None, [C.c_void_p]),
        "dg_stepper_transport_stats": (None, [C.c_void_p, C.POINTER(C.c_int)]),
        "dg_batch_create": (C.c_void_p, [C.POINTER(C.POINTER(GdnSimulation)), C.c_int,
                                         c_double_p, c_double_p, C.c_int]),
        "dg_batch_tiles": (None, [C.c_void_p, C.POINTER(C.c_int)]),
        "dg_batch_advance": (None, [C.c_void_p, C.c_int]),
        "dg_batch_synchronize": (None, [C.c_void_p]),
        "dg_batch_iterations": (C.c_int, [C.c_void_p]),
        "dg_batch_errors": (None, [C.c_void_p, c_double_p]),
        "dg_batch_times": (None, [C.c_void_p, c_double_p]),
        "dg_batch_finish": (None, [C.c_void_p]),
        "dg_batch_free": (None, [C.c_void_p]),
        "dg_stepper_device_context": (C.c_void_p, [C.c_void_p]),
        "gdh_mlp_nparams": (C.c_int, []),
        "gdh_mlp_create": (C.c_void_p, [C.c_uint64]),
        "gdh_mlp_destroy": (None, [C.c_void_p]),
        "gdh_mlp_get_params": (None, [C.c_void_p, c_float_p]),
        "gdh_mlp_set_params": (None, [C.c_void_p, c_float_p]),
        "gdh_mlp_set_normalization": (None, [C.c_void_p, c_float_p]),
        "gdh_mlp_grad": (C.c_double, [C.c_void_p, c_float_p, c_float_p, C.c_int, c_float_p]),
        "gdh_mlp_train": (C.c_double, [C.c_void_p, c_float_p, c_float_p, C.c_int, C.c_int,
                                       C.c_int, C.c_float, C.c_uint64, c_double_p]),
        "gdh_mlp_predict": (None, [C.c_void_p, c_float_p, C.c_int, c_float_p]),
        "gdh_learned_data": (None, [C.c_void_p, C.c_int, C.POINTER(C.c_int), c_float_p, c_float_p]),
        "gdh_sweep_grid": (C.c_int, [C.c_void_p]),
        "gdh_relax_fused": (C.c_int, [C.c_void_p]),
        "gdh_sweep_info": (None, [C.c_void_p, C.POINTER(C.c_int)]),
        "gdh_rk3_info": (None, [C.c_void_p, C.POINTER(C.c_int)]),
        "gdh_learned_solve_stats": (None, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                           C.c_double, C.c_int, c_double_p, C.POINTER(C.c_int)]),
        "gdh_learned_solve_stats_ex": (None, [C.c_void_p, C.c_void_p, C.c_int,
                                              C.POINTER(C.c_int), C.c_double, C.c_int, C.c_int,
                                              c_double_p, C.POINTER(C.c_int)]),
        "gdh_learned_enable": (None, [C.c_void_p, C.c_void_p, C.c_double, c_double_p]),
        "t10_get_wpg_ref_entry": (C.c_double, [C.c_int]),
        "t10_get_mass_ref_entry": (C.c_double, [C.c_int, C.c_int]),
        "t10_get_inv_mass_ref_entry": (C.c_double, [C.c_int, C.c_int]),
        "t10_get_internal_ref_entry": (C.c_double, [C.c_int, C.c_int, C.c_int]),
        "t10_get_face_scale_entry": (C.c_double, [C.c_int, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def fnptr(name):
    """Address of an exported C function, for a gdn_model callback field."""
    lib = load_library()
    return C.cast(getattr(lib, name), C.c_void_p).value


c_float_p = C.POINTER(C.c_float)


def fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_float_p)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int))


def dptr(a):
    return a.ctypes.data_as(c_double_p)


class Mesh:
    """A built gdn_mesh owned by Python."""

    def __init__(self, path=None, kuhn=None):
        lib = load_library()
        self.s = GdnMesh()
        if path is not None:
            lib.mesh_read_from_msh22_file(C.byref(self.s), str(path).encode())
        else:
            n, jitter, seed = kuhn
            lib.mesh_kuhn_cube(C.byref(self.s), n, jitter, seed)
        lib.mesh_build_connectivity(C.byref(self.s))

    @property
    def nelem(self):
        return self.s.nbelems

    def arr(self, field, n, dtype=np.int32):
        p = getattr(self.s, field)
        ct = C.c_int if dtype == np.int32 else C.c_double
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).copy()

    def arrays(self):
        s = self.s
        nf, nb, ne = s.nbfaces, s.nboundaryfaces, s.nbelems
        return {
            "face2elem": self.arr("face2elem", 4 * nf),
            "fnode2enode": self.arr("fnode2enode", 12 * nf),
            "elem2qnode": self.arr("elem2qnode", 10 * ne + 6 * nb),
            "node": self.arr("node", 3 * s.nbqnodes_in, np.float64),
            "boundaryface": self.arr("boundaryface", nb),
            "elem2elem": self.arr("elem2elem", 4 * ne),
        }

    def free(self):
        if self.s.nbelems:
            load_library().gdn_gdn_mesh_free(C.byref(self.s))


class Model:
    """A Maxwell gdn_model using the library's callbacks."""

    def __init__(self, imposed="plane_wave", kinetic=None, boundary="upwind",
                 callback=None, lambda_=1.0, boundary_state="feq"):
        """boundary_state (kinetic): "feq", f_eq of the imposed data (the
        reference's get_fR), or "projector", the state of get_fR's
        use_feq = false branch (d3q4_get_fR_projector / d3q6_get_fR_projector)."""
        lib = load_library()
        self.s = GdnModel()
        self.s.nb_w = 6
        self.s.lambda_ = lambda_
        self.s.get_num_flux = fnptr("maxwell_num_flux_upwind")
        self.s.get_num_flux_boundary = fnptr(
            "maxwell_num_flux_upwind" if boundary == "upwind" else "maxwell_num_flux_boundary_metal")
        self._cb = None
        if callback is not None:
            self._cb = IMPOSED_CB(callback)
            self.s.get_imposed_data = C.cast(self._cb, C.c_void_p).value
        else:
            self.s.get_imposed_data = fnptr("maxwell_imposed_" + imposed)
        self.s.get_proj_pos = fnptr("maxwell_get_proj_pos_null_prod_w")
        self.s.get_proj_neg = fnptr("maxwell_get_proj_neg_prod_w")
        if kinetic:
            self.s.use_relax_scheme = True
            getattr(lib, kinetic + "_set_model")(C.byref(self.s))
            if boundary_state == "projector":
                self.s.get_fR = fnptr(kinetic + "_get_fR_projector")
        else:
            self.s.use_relax_scheme = False


def set_plane_wave(k, e0, phase=0.0):
    k = np.ascontiguousarray(k, dtype=np.float64)
    e0 = np.ascontiguousarray(e0, dtype=np.float64)
    load_library().maxwell_plane_wave_set(dptr(k), dptr(e0), phase)


def set_plane_wave_eps(k, e0, eps, phase=0.0):
    """plane wave of speed 1/sqrt(eps) (exact for a uniform permittivity)"""
    k = np.ascontiguousarray(k, dtype=np.float64)
    e0 = np.ascontiguousarray(e0, dtype=np.float64)
    load_library().maxwell_plane_wave_set_eps(dptr(k), dptr(e0), phase, eps)


def set_gaussian_pulse(center=(0.5, 0.5, 0.5), sigma=0.08, amp=1.0):
    c = np.ascontiguousarray(center, dtype=np.float64)
    load_library().maxwell_gaussian_pulse_set(dptr(c), sigma, amp)


def set_cavity_mode(l, m, n, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    load_library().maxwell_cavity_mode_set(l, m, n, dptr(a))


class Simulation:
    """gdn_simulation on a Mesh and a Model (simulation_init semantics)."""

    def __init__(self, mesh, model):
        self.mesh, self.model = mesh, model
        self.s = GdnSimulation()
        load_library().simulation_init(C.byref(mesh.s), C.byref(model.s), C.byref(self.s))

    def set_time(self, tmax, dt=0.0, cfl=-1.0):
        load_library().simulation_set_time_parameters(C.byref(self.s), tmax, dt, cfl)

    def solve(self):
        load_library().dg_solve(C.byref(self.s))

    def wn(self):
        return np.ctypeslib.as_array(self.s.wn, shape=(self.s.wlen,)).copy()

    def set_wn(self, values):
        dst = np.ctypeslib.as_array(self.s.wn, shape=(self.s.wlen,))
        dst[:] = values

    def error_l2(self):
        return load_library().simulation_error_l2(C.byref(self.s))

    def free(self):
        """simulation_free also frees the mesh arrays (reference semantics)."""
        load_library().simulation_free(C.byref(self.s))


class Stepper:
    """dg_stepper_* : set-up, initial condition, k-step advances, timing."""

    PRECISIONS = {"fp64": 0, "fp32": 1}

    def __init__(self, sim, precision="fp64", eps=None):
        """precision: "fp64" (the plain form) or "fp32" (binary32 fields,
        binary32 RK3 flux/volume arithmetic; see dg_stepper_create_precision).
        eps: permittivity per mesh element (dg_stepper_set_permittivity)."""
        self.lib = load_library()
        self.sim = sim
        self.h = self.lib.dg_stepper_create_precision(C.byref(sim.s),
                                                      self.PRECISIONS[precision])
        if eps is not None:
            self._eps = np.ascontiguousarray(eps, dtype=np.float64)
            assert self._eps.shape == (sim.mesh.nelem,)
            self.lib.dg_stepper_set_permittivity(self.h, dptr(self._eps))
        self.lib.dg_stepper_init(self.h)

    def advance(self, k):
        self.lib.dg_stepper_advance(self.h, k)

    def sync(self):
        self.lib.dg_stepper_synchronize(self.h)

    @property
    def t(self):
        return self.lib.dg_stepper_time(self.h)

    @property
    def levels(self):
        return self.lib.dg_stepper_levels(self.h)

    def error_l2(self):
        return self.lib.dg_stepper_error_l2(self.h)

    def energy(self):
        """1/2 w^T M w of the current state (binary64 device reduction)."""
        return self.lib.dg_stepper_energy(self.h)

    def upload(self, wn):
        """Replace the device state by wn (reference layout of simu->wn)."""
        wn = np.ascontiguousarray(wn, dtype=np.float64)
        self.lib.gdh_upload_state(self.lib.dg_stepper_device_context(self.h), dptr(wn))

    def download(self):
        """The device state in the reference layout (no ghost update)."""
        out = np.zeros(self.sim.s.wlen)
        self.lib.gdh_download_state(self.lib.dg_stepper_device_context(self.h), dptr(out))
        return out

    def profile(self, enable):
        out = np.zeros(4)
        self.lib.dg_stepper_profile(self.h, int(enable), dptr(out))
        return out

    def transport_stats(self):
        """(lagged inflow faces, sweeps of the last step, most sweeps in a step)"""
        out = (C.c_int * 3)()
        self.lib.dg_stepper_transport_stats(self.h, out)
        return tuple(out)

    def finish(self):
        self.lib.dg_stepper_finish(self.h)

    def free(self):
        if self.h:
            self.lib.dg_stepper_free(self.h)
            self.h = None


class Batch:
    """dg_batch_* : independent simulations advanced together (one device
    context over the union of their meshes, captured-graph steps,
    autotuned tiles). par: array (count, 16) of exact-solution parameter
    blocks; eps: permittivity per element (RK3), instance after instance."""

    def __init__(self, sims, par, eps=None, precision="fp32"):
        self.lib = load_library()
        self.sims = list(sims)
        arr = (C.POINTER(GdnSimulation) * len(self.sims))(
            *[C.pointer(s.s) for s in self.sims])
        self._par = np.ascontiguousarray(par, dtype=np.float64).reshape(-1)
        self._eps = None if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
        self.h = self.lib.dg_batch_create(
            arr, len(self.sims), dptr(self._par),
            dptr(self._eps) if self._eps is not None else None,
            Stepper.PRECISIONS[precision])

    def tiles(self):
        out = (C.c_int * 2)()
        self.lib.dg_batch_tiles(self.h, out)
        return tuple(out)

    def advance(self, k):
        self.lib.dg_batch_advance(self.h, k)

    def sync(self):
        self.lib.dg_batch_synchronize(self.h)

    def errors(self):
        out = np.zeros(len(self.sims))
        self.lib.dg_batch_errors(self.h, dptr(out))
        return out

    def times(self):
        out = np.zeros(len(self.sims))
        self.lib.dg_batch_times(self.h, dptr(out))
        return out

    def finish(self):
        self.lib.dg_batch_finish(self.h)

    def free(self):
        if self.h:
            self.lib.dg_batch_free(self.h)
            self.h = None


class BlockNet:
    """The learned block preconditioner's network (gdh_mlp_*)."""

    def __init__(self, seed=1):
        self.lib = load_library()
        self.h = self.lib.gdh_mlp_create(seed)
        self.n = self.lib.gdh_mlp_nparams()

    def params(self):
        out = np.zeros(self.n, np.float32)
        self.lib.gdh_mlp_get_params(self.h, fptr(out))
        return out

    def set_params(self, p):
        self.lib.gdh_mlp_set_params(self.h, fptr(np.ascontiguousarray(p, np.float32)))

    def set_normalization(self, norm):
        self.norm = np.ascontiguousarray(norm, np.float32)
        self.lib.gdh_mlp_set_normalization(self.h, fptr(self.norm))

    def grad(self, x, y):
        g = np.zeros(self.n, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = self.lib.gdh_mlp_grad(self.h, fptr(x), fptr(y), x.shape[0], fptr(g))
        return loss, g

    def train(self, x, y, steps, batch=1024, lr=1e-3, seed=7):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        hist = np.zeros(steps)
        self.lib.gdh_mlp_train(self.h, fptr(x), fptr(y), x.shape[0], steps, batch, lr, seed,
                               dptr(hist))
        return hist

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], 100), np.float32)
        self.lib.gdh_mlp_predict(self.h, fptr(x), x.shape[0], fptr(y))
        return y

    def free(self):
        if self.h:
            self.lib.gdh_mlp_destroy(self.h)
            self.h = None


def sweep_grid(stepper):
    """Workgroups of the one-launch transport sweep (measured residency)."""
    if not stepper.h:
        raise RuntimeError("stepper already freed")
    lib = load_library()
    return lib.gdh_sweep_grid(lib.dg_stepper_device_context(stepper.h))


SWEEP_FORMS = {0: "level", 1: "item", 2: "cluster"}


def sweep_info(stepper):
    """Form of the transport sweep: dict(form, clusters, cluster_levels,
    clustered_items, grid) (see gdh_sweep_info)."""
    if not stepper.h:
        raise RuntimeError("stepper already freed")
    lib = load_library()
    out = (C.c_int * 5)()
    lib.gdh_sweep_info(lib.dg_stepper_device_context(stepper.h), out)
    return dict(form=SWEEP_FORMS.get(out[0], str(out[0])), clusters=out[1],
                cluster_levels=out[2], clustered_items=out[3], grid=out[4])


RK3_FORMS = {0: "soa", 1: "tile", 2: "pair"}


def rk3_info(stepper):
    """Form of the RK3 stage: dict(kernel, k2_nontemporal, block)."""
    if not stepper.h:
        raise RuntimeError("stepper already freed")
    lib = load_library()
    out = (C.c_int * 3)()
    lib.gdh_rk3_info(lib.dg_stepper_device_context(stepper.h), out)
    return dict(kernel=RK3_FORMS.get(out[0], str(out[0])), k2_nontemporal=bool(out[1]),
                block=out[2])


def relax_fused(stepper):
    """1 if the sweep applies each step's relaxation as it reads its input."""
    if not stepper.h:
        raise RuntimeError("stepper already freed")
    lib = load_library()
    return lib.gdh_relax_fused(lib.dg_stepper_device_context(stepper.h))


def learned_data(stepper, idx):
    """Features x[n][7] and targets D^-1 - I [n][100] of sweep items idx."""
    lib = load_library()
    idx = np.ascontiguousarray(idx, np.int32)
    x = np.zeros((len(idx), 7), np.float32)
    y = np.zeros((len(idx), 100), np.float32)
    ctx = lib.dg_stepper_device_context(stepper.h)
    lib.gdh_learned_data(ctx, len(idx), iptr(idx), fptr(x), fptr(y))
    return x, y


def learned_solve_stats(stepper, net, idx, tol=1e-8, maxit=50, method="richardson"):
    """(mean iterations, max, fallback rate, max rel. L2 to exact, solves);
    method "richardson" or "gmres" (right-preconditioned, restart 10)"""
    lib = load_library()
    idx = np.ascontiguousarray(idx, np.int32)
    out = np.zeros(5)
    ctx = lib.dg_stepper_device_context(stepper.h)
    m = {"richardson": 0, "gmres": 1}[method]
    lib.gdh_learned_solve_stats_ex(ctx, net.h if net is not None else None, len(idx),
                                   iptr(idx), tol, maxit, m, dptr(out), None)
    return out


def learned_enable(stepper, net, tol=1e-8):
    """Put the sweep's block inverses on the network's refined predictions
    (gdh_learned_enable; net None: back on the stored exact inverses).
    Returns (items on the network, items on the stored inverse, largest
    relative deviation of a refined inverse from the stored one, median
    max|I - D X0| of the predictions)."""
    lib = load_library()
    out = np.zeros(4)
    lib.gdh_learned_enable(lib.dg_stepper_device_context(stepper.h),
                           net.h if net is not None else None, tol, dptr(out))
    return int(out[0]), int(out[1]), float(out[2]), float(out[3])


def normalization(x, y):
    """mu_x, sd_x, mu_y, sd_y of training data (sd floored)."""
    return np.concatenate([x.mean(0), np.maximum(x.std(0), 1e-6),
                           y.mean(0), np.maximum(y.std(0), 1e-6)]).astype(np.float32)
