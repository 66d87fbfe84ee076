"""ORACLE bindings -- test infrastructure only.

ctypes access to liboracle.so (the C99 restatement of the reference CPU
algorithm, oracle.c) and to _ref/libgdon3d_refsubset.so (the reference's own
SuiteSparse-free sources compiled from /root/reference by oracle/Makefile).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GDN_ORACLE_LIB: another build of oracle.c (`make -C oracle sanitize`)
ORACLE_LIB = os.environ.get("GDN_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libgdon3d_refsubset.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

# exact-solution kinds (same numbering as include/gdn_hip.h)
MACRO_COS, CAVITY_TM, STATIONARY, PLANE_WAVE, CAVITY_MODE, GAUSSIAN_PULSE = 1, 2, 3, 4, 5, 6
D3Q4, D3Q6 = 1, 2
UPWIND, METAL = 1, 2

_olib = None


def lib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run make -C oracle")
        L = C.CDLL(ORACLE_LIB)
        sig = {
            "orc_tables": (None, [dp, dp, dp, dp, dp]),
            "orc_mesh_load": (C.c_void_p, [C.c_char_p]),
            "orc_mesh_info": (None, [C.c_void_p, ip]),
            "orc_mesh_arrays": (None, [C.c_void_p, ip, ip, ip, dp, ip, ip]),
            "orc_mesh_hmin": (C.c_double, [C.c_void_p]),
            "orc_mesh_free": (None, [C.c_void_p]),
            "orc_sim_create": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int, dp, C.c_int, C.c_int]),
            "orc_sim_set_lambda": (None, [C.c_void_p, C.c_double]),
            "orc_sim_set_threads": (None, [C.c_void_p, C.c_int]),
            "orc_set_time": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
            "orc_sim_info": (None, [C.c_void_p, dp]),
            "orc_solve": (None, [C.c_void_p, C.c_int]),
            "orc_get_wn": (None, [C.c_void_p, dp]),
            "orc_set_wn": (None, [C.c_void_p, dp]),
            "orc_rk3_operator": (None, [C.c_void_p, dp, dp]),
            "orc_transport": (None, [C.c_void_p, C.c_int, dp, dp]),
            "orc_transport_residual": (C.c_double, [C.c_void_p, C.c_int, dp, dp]),
            "orc_block_stats": (None, [C.c_void_p, C.c_int, ip]),
            "orc_relax": (None, [C.c_void_p, C.c_int]),
            "orc_m1_dependencies": (None, [C.c_void_p, C.c_int, ip]),
            "orc_update_boundary": (None, [C.c_void_p, C.c_double]),
            "orc_init_solution": (None, [C.c_void_p]),
            "orc_error_l2": (C.c_double, [C.c_void_p]),
            "orc_energy": (C.c_double, [C.c_void_p]),
            "orc_sim_set_fr_projector": (None, [C.c_void_p, C.c_int]),
            "orc_sim_set_nofactor": (None, [C.c_void_p, C.c_int]),
            "orc_sim_set_eps": (None, [C.c_void_p, dp]),
            "orc_proj": (None, [dp, dp, C.c_int, dp]),
            "orc_fr_projector": (None, [C.c_int, dp, dp, dp, dp]),
            "orc_sim_free": (None, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _olib = L
    return _olib


def _d(a):
    return a.ctypes.data_as(dp)


def tables():
    wpg = np.zeros(10)
    mass = np.zeros((10, 10))
    imass = np.zeros((10, 10))
    internal = np.zeros((10, 10, 3))
    face = np.zeros((6, 6))
    lib().orc_tables(_d(wpg), _d(mass), _d(imass), _d(internal), _d(face))
    return dict(wpg=wpg, mass=mass, inv_mass=imass, internal=internal, face=face)


class OMesh:
    def __init__(self, path):
        self.h = lib().orc_mesh_load(str(path).encode())
        if not self.h:
            raise FileNotFoundError(path)
        info = (C.c_int * 5)()
        lib().orc_mesh_info(self.h, info)
        self.nelem, self.nnodes, self.nq, self.nfaces, self.nbf = list(info)

    def arrays(self):
        f2e = np.zeros(4 * self.nfaces, np.int32)
        fn = np.zeros(12 * self.nfaces, np.int32)
        e2q = np.zeros(10 * self.nelem + 6 * self.nbf, np.int32)
        xq = np.zeros(3 * self.nq)
        bf = np.zeros(max(self.nbf, 1), np.int32)
        e2e = np.zeros(4 * self.nelem, np.int32)
        P = lambda a: a.ctypes.data_as(ip)
        lib().orc_mesh_arrays(self.h, P(f2e), P(fn), P(e2q), _d(xq), P(bf), P(e2e))
        return dict(face2elem=f2e, fnode2enode=fn, elem2qnode=e2q, node=xq,
                    boundaryface=bf[: self.nbf], elem2elem=e2e)

    def hmin(self):
        return lib().orc_mesh_hmin(self.h)

    def free(self):
        if self.h:
            lib().orc_mesh_free(self.h)
            self.h = None


class OSim:
    """Assembled-matrix solve of the reference algorithm on an OMesh."""

    def __init__(self, mesh, scheme, imposed_kind, imposed_par=None, kin=0,
                 bflux=UPWIND, lam=None, threads=None, fr_projector=False, eps=None,
                 factor=True):
        """factor=False: the kinetic assembly leaves M1 unfactored (only
        transport_residual and the products are available)."""
        par = np.zeros(16)
        if imposed_par is not None:
            par[: len(imposed_par)] = imposed_par
        self.mesh = mesh
        self.h = lib().orc_sim_create(mesh.h, scheme, imposed_kind, _d(par), kin, bflux)
        if lam is not None:
            lib().orc_sim_set_lambda(self.h, lam)
        lib().orc_sim_set_threads(self.h, threads or os.cpu_count() or 1)
        lib().orc_sim_set_fr_projector(self.h, int(fr_projector))
        lib().orc_sim_set_nofactor(self.h, int(not factor))
        if eps is not None:
            self._eps = np.ascontiguousarray(eps, dtype=np.float64)
            assert self._eps.shape == (mesh.nelem,)
            lib().orc_sim_set_eps(self.h, _d(self._eps))
        self.info()

    def info(self):
        out = np.zeros(10)
        lib().orc_sim_info(self.h, _d(out))
        self.neq, self.m, self.itermax = int(out[0]), int(out[1]), int(out[2])
        self.tmax, self.dt, self.hmin, self.cfl, self.t = out[3], out[4], out[5], out[6], out[7]
        self.t_asm, self.t_run = out[8], out[9]
        return out

    def set_time(self, tmax, dt=0.0, cfl=-1.0):
        lib().orc_set_time(self.h, tmax, dt, cfl)
        self.info()

    def solve(self, nsteps=-1):
        lib().orc_solve(self.h, nsteps)
        self.info()

    def wn(self):
        out = np.zeros(self.m * self.neq)
        lib().orc_get_wn(self.h, _d(out))
        return out

    def set_wn(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        lib().orc_set_wn(self.h, _d(w))

    def rk3_operator(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.zeros_like(w)
        lib().orc_rk3_operator(self.h, _d(w), _d(out))
        return out

    def transport(self, iv, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(f)
        lib().orc_transport(self.h, iv, _d(f), _d(out))
        return out

    def transport_residual(self, iv, f, x):
        f = np.ascontiguousarray(f, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        return lib().orc_transport_residual(self.h, iv, _d(f), _d(x))

    def block_stats(self, iv):
        out = (C.c_int * 2)()
        lib().orc_block_stats(self.h, iv, out)
        return list(out)

    def m1_dependencies(self, iv):
        out = np.zeros(4 * self.mesh.nelem, np.int32)
        lib().orc_m1_dependencies(self.h, iv, out.ctypes.data_as(ip))
        return out.reshape(-1, 4)

    def relax(self, with_ghosts=True):
        lib().orc_relax(self.h, int(with_ghosts))

    def update_boundary(self, t):
        lib().orc_update_boundary(self.h, t)

    def init_solution(self):
        lib().orc_init_solution(self.h)

    def error_l2(self):
        return lib().orc_error_l2(self.h)

    def energy(self):
        """1/2 w^T M w of the current state (assembled P2 mass)."""
        return lib().orc_energy(self.h)

    def free(self):
        if self.h:
            lib().orc_sim_free(self.h)
            self.h = None


def proj(w, n, neg):
    """the oracle's Maxwell eigen-projector P^- (neg) or P^{+,0} applied to w"""
    w, n = np.ascontiguousarray(w, dtype=np.float64), np.ascontiguousarray(n, dtype=np.float64)
    out = np.zeros(6)
    lib().orc_proj(_d(w), _d(n), int(neg), _d(out))
    return out


def fr_projector(kind, fL, vn, wI):
    """the oracle's projector boundary state (kind D3Q4 or D3Q6)"""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (fL, vn, wI)]
    out = np.zeros(len(a[0]))
    lib().orc_fr_projector(kind, _d(a[0]), _d(a[1]), _d(a[2]), _d(out))
    return out


# ------------------------------------------------------------- reference
_rlib = None


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own mesh/model code (struct gdn_mesh layout)."""
    global _rlib
    if _rlib is None:
        _rlib = C.CDLL(REF_LIB)
    return _rlib
