"""Shared helpers of the test suite: synthetic meshes and comparisons."""
import ctypes as C
import os
import tempfile

import numpy as np

import gdon3d as G
import oracle as O

_TMP = tempfile.mkdtemp(prefix="gdon3d_tests_")


def kuhn_msh(n, jitter=0.0, seed=0):
    """gmsh v2.2 file of the Kuhn n^3 x 6 cube (written once per process)."""
    path = os.path.join(_TMP, f"kuhn_{n}_{jitter}_{seed}.msh")
    if not os.path.exists(path):
        rc = G.load_library().mesh_kuhn_cube_write_msh22(path.encode(), n, jitter, seed)
        assert rc == 0
    return path


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def plane_wave_par(k=(2 * np.pi, 0, 0), e0=(0, 1, 0), phase=0.0):
    """Parameter block of the plane-wave exact solution (as gdn_hip.h)."""
    k = np.asarray(k, dtype=np.float64)
    e0 = np.asarray(e0, dtype=np.float64)
    h0 = np.cross(k / np.linalg.norm(k), e0)
    par = np.zeros(16)
    par[0:3], par[3:6], par[6:9] = k, e0, h0
    par[9], par[10] = np.linalg.norm(k), phase
    return par


def cavity_par(lmn=(1, 1, 1), a=(1, -1, 0)):
    k = np.pi * np.asarray(lmn, dtype=np.float64)
    par = np.zeros(16)
    par[0:3], par[3:6] = k, a
    par[6] = np.pi * np.sqrt(float(np.dot(lmn, lmn)))
    return par


def twisted_msh(n, alpha):
    """The Kuhn n^3 x 6 cube twisted about the vertical axis through its
    centre by alpha * z radians (vertices only: the cells stay affine and
    positively oriented for the values used here). Strong twists make the
    upwind graphs of fixed directions cyclic, which jitter of a Kuhn mesh
    does not while its cells stay positively oriented."""
    src = kuhn_msh(n, 0.0, 0)
    path = os.path.join(_TMP, f"twist_{n}_{alpha}.msh")
    if not os.path.exists(path):
        lines = open(src).read().split("\n")
        i = lines.index("$Nodes")
        for k in range(i + 2, i + 2 + int(lines[i + 1])):
            t = lines[k].split()
            x, y, z = (float(c) for c in t[1:4])
            c, s = np.cos(alpha * z), np.sin(alpha * z)
            u, v = x - 0.5, y - 0.5
            lines[k] = f"{t[0]} {0.5 + c * u - s * v:.17g} {0.5 + s * u + c * v:.17g} {z:.17g}"
        with open(path, "w") as f:
            f.write("\n".join(lines))
    return path


class HipRuntime:
    """The HIP runtime the library itself loaded (same soname, so the same
    handle), for tests that hand device pointers to the C ABI: hipMalloc,
    hipMemcpy, hipFree."""

    H2D, D2H = 1, 2

    def __init__(self):
        G.load_library()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def check(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), max(int(nbytes), 16)), "hipMalloc")
        return p.value

    def free(self, p):
        self.check(self.h.hipFree(p), "hipFree")

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self.check(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy")
        return p

    def to_host(self, p, n, dtype=np.float64):
        out = np.zeros(n, dtype)
        self.check(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H), "hipMemcpy")
        return out


def plane_wave_eps_par(k, e0, eps, phase=0.0):
    """Parameter block of the plane wave in a uniform medium of permittivity
    eps: speed 1/sqrt(eps), H = sqrt(eps) (k/|k|) x e0 s."""
    par = plane_wave_par(k, e0, phase)
    par[6:9] *= np.sqrt(eps)
    par[9] /= np.sqrt(eps)
    return par


def gaussian_par(center=(0.5, 0.5, 0.5), sigma=0.08, amp=1.0):
    par = np.zeros(16)
    par[0:3], par[3], par[4] = center, sigma, amp
    return par


def block_eps(path, nblock=2, seed=3, lo=1.0, hi=4.0):
    """Piecewise-constant permittivity in [lo, hi] per cell of an nblock^3
    grid over the unit cube, by element centroid (numpy default_rng(seed))."""
    om = O.OMesh(path)
    a = om.arrays()
    X = a["node"].reshape(-1, 3)[a["elem2qnode"][: 10 * om.nelem]].reshape(om.nelem, 10, 3)
    c = X[:, :4, :].mean(axis=1)
    om.free()
    table = np.random.default_rng(seed).uniform(lo, hi, size=(nblock,) * 3)
    idx = np.minimum((c * nblock).astype(int), nblock - 1)
    return table[idx[:, 0], idx[:, 1], idx[:, 2]]
