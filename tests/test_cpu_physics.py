"""Maxwell fluxes, kinetic equilibria and exact solutions of the library
against the reference's own src/models (oracle/_ref) and against the
mathematics: conservation, consistency, upwind eigen-decomposition, moment
identities, Maxwell's equations for the exact fields."""
import ctypes as C

import numpy as np
import pytest

import gdon3d as G
import oracle as O

dp = C.POINTER(C.c_double)
rng = np.random.default_rng(1234)


def flux(lib, name, wl, wr, n):
    out = np.zeros(6)
    f = getattr(lib, name)
    f.argtypes = [dp] * 4
    f.restype = None
    f(G.dptr(np.ascontiguousarray(wl)), G.dptr(np.ascontiguousarray(wr)),
      G.dptr(np.ascontiguousarray(n)), G.dptr(out))
    return out


def A(n):
    """A(n) w = (-n x H, n x E)"""
    N = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    M = np.zeros((6, 6))
    M[:3, 3:] = -N
    M[3:, :3] = N
    return M


def test_upwind_flux_against_eigen_decomposition():
    lib = G.load_library()
    for _ in range(200):
        wl, wr, n = rng.normal(size=6), rng.normal(size=6), rng.normal(size=3)
        f = flux(lib, "maxwell_num_flux_upwind", wl, wr, n)
        An = A(n)
        lam, V = np.linalg.eigh(An)  # A(n) is symmetric
        absA = V @ np.diag(np.abs(lam)) @ V.T
        nn = np.linalg.norm(n)
        expect = An @ (wl + wr) / 2 - (nn / (nn + 1e-16)) * absA @ (wr - wl) / 2 / nn * nn
        assert np.allclose(f, expect, rtol=1e-13, atol=1e-13)
        # conservation and consistency
        g = flux(lib, "maxwell_num_flux_upwind", wr, wl, -n)
        assert np.array_equal(f, -g)
        assert np.allclose(flux(lib, "maxwell_num_flux_upwind", wl, wl, n), An @ wl,
                           rtol=1e-14, atol=1e-14)


def test_metal_flux():
    lib = G.load_library()
    wl, n = rng.normal(size=6), rng.normal(size=3)
    f = flux(lib, "maxwell_num_flux_boundary_metal", wl, np.zeros(6), n)
    assert np.allclose(f[:3], -np.cross(n, wl[3:]), atol=1e-15)
    assert np.all(f[3:] == 0)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_fluxes_match_reference_build():
    lib, R = G.load_library(), O.ref()
    for _ in range(200):
        wl, wr, n = rng.normal(size=6), rng.normal(size=6), rng.normal(size=3)
        for name in ("maxwell_num_flux_upwind", "maxwell_num_flux_boundary_metal"):
            a = flux(lib, name, wl, wr, n)
            b = flux(R, name, wl, wr, n)
            assert np.allclose(a, b, rtol=2e-15, atol=2e-15), name
    # projections on the eigenspaces of A(n/|n|)
    for name in ("maxwell_get_proj_neg_prod_w", "maxwell_get_proj_pos_prod_w",
                 "maxwell_get_proj_neg_null_prod_w", "maxwell_get_proj_pos_null_prod_w"):
        for _ in range(50):
            w, n = rng.normal(size=6), rng.normal(size=3)
            outs = []
            for L in (lib, R):
                f = getattr(L, name)
                f.argtypes = [dp] * 3
                o = np.zeros(6)
                f(G.dptr(w), G.dptr(n), G.dptr(o))
                outs.append(o)
            assert np.allclose(outs[0], outs[1], rtol=1e-13, atol=1e-13), name


def _model_struct(L, kin):
    m = G.GdnModel()
    m.nb_w = 6
    m.get_num_flux = C.cast(L.maxwell_num_flux_upwind, C.c_void_p).value
    getattr(L, kin + "_set_model").argtypes = [C.c_void_p]
    getattr(L, kin + "_set_model")(C.byref(m))
    return m


@pytest.mark.parametrize("kin,nv", [("d3q4", 4), ("d3q6", 6)])
def test_equilibrium_moments(kin, nv):
    lib = G.load_library()
    m = _model_struct(lib, kin)
    vi = np.ctypeslib.as_array(m.vi, shape=(3 * nv,)).reshape(nv, 3)
    feq_f = getattr(lib, kin + "_get_feq")
    feq_f.argtypes = [C.c_void_p, dp, dp]
    for _ in range(50):
        w = rng.normal(size=6)
        f = np.zeros(6 * nv)
        feq_f(C.byref(m), G.dptr(w), G.dptr(f))
        f = f.reshape(6, nv)
        assert np.allclose(f.sum(axis=1), w, atol=1e-14)  # sum_v f_eq = w
        q = np.stack([A(np.eye(3)[k]) @ w for k in range(3)], axis=1)  # q[iw, k]
        assert np.allclose(f @ vi, q, atol=1e-14)  # sum_v v f_eq = A(e_k) w


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("kin,nv", [("d3q4", 4), ("d3q6", 6)])
def test_kinetic_functions_match_reference_build(kin, nv):
    lib, R = G.load_library(), O.ref()
    ml, mr = _model_struct(lib, kin), _model_struct(R, kin)
    assert ml.lambda_ == mr.lambda_ and ml.omega == mr.omega
    assert np.array_equal(np.ctypeslib.as_array(ml.vi, shape=(3 * nv,)),
                          np.ctypeslib.as_array(mr.vi, shape=(3 * nv,)))
    for _ in range(50):
        w = rng.normal(size=6)
        fl, fr = np.zeros(6 * nv), np.zeros(6 * nv)
        for L, m, o in ((lib, ml, fl), (R, mr, fr)):
            f = getattr(L, kin + "_get_feq")
            f.argtypes = [C.c_void_p, dp, dp]
            f(C.byref(m), G.dptr(w), G.dptr(o))
        assert np.allclose(fl, fr, rtol=1e-15, atol=1e-15)
        wl, wr = np.zeros(6), np.zeros(6)
        for L, m, o in ((lib, ml, wl), (R, mr, wr)):
            f = getattr(L, kin + "_get_w")
            f.argtypes = [C.c_void_p, dp, dp]
            f(C.byref(m), G.dptr(fl), G.dptr(o))
        assert np.array_equal(wl, wr)


def imposed(L, name, x, t):
    f = getattr(L, name)
    f.argtypes = [dp, C.c_double, dp]
    w = np.zeros(6)
    f(G.dptr(np.ascontiguousarray(x, dtype=np.float64)), t, G.dptr(w))
    return w


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("name", ["maxwell_imposed_macro_cos", "maxwell_imposed_macro_cavity",
                                  "maxwell_imposed_stationary"])
def test_reference_exact_solutions_match_reference_build(name):
    lib, R = G.load_library(), O.ref()
    for _ in range(100):
        x, t = rng.uniform(0, 1, 3), rng.uniform(0, 2)
        assert np.allclose(imposed(lib, name, x, t), imposed(R, name, x, t),
                           rtol=1e-14, atol=1e-14)


def maxwell_residual(fun, x, t, h=1e-5):
    """d_t E - curl H and d_t H + curl E by central differences."""
    def d(axis):
        e = np.zeros(3)
        e[axis] = h
        return (fun(x + e, t) - fun(x - e, t)) / (2 * h)
    dt = (fun(x, t + h) - fun(x, t - h)) / (2 * h)
    g = [d(k) for k in range(3)]  # g[k][q] = d_k w_q
    curl = lambda o: np.array([g[1][o + 2] - g[2][o + 1], g[2][o] - g[0][o + 2],
                               g[0][o + 1] - g[1][o]])
    div_e = g[0][0] + g[1][1] + g[2][2]
    return np.concatenate([dt[:3] - curl(3), dt[3:] + curl(0)]), div_e


@pytest.mark.parametrize("which", ["plane", "cavity"])
def test_parameterised_solutions_solve_maxwell(which):
    lib = G.load_library()
    if which == "plane":
        k = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
        G.set_plane_wave(k, (0, 0, 1), 0.4)
        name = "maxwell_imposed_plane_wave"
    else:
        G.set_cavity_mode(1, 2, 1, (1.0, 0.0, -1.0))
        name = "maxwell_imposed_cavity_mode"
    fun = lambda x, t: imposed(lib, name, x, t)
    for _ in range(20):
        x, t = rng.uniform(0, 1, 3), rng.uniform(0, 1)
        r, div = maxwell_residual(fun, x, t)
        assert np.abs(r).max() < 1e-6 and abs(div) < 1e-6
    if which == "cavity":
        # perfect conductor: tangential E vanishes on every wall
        for _ in range(20):
            y = rng.uniform(0, 1, 3)
            for axis in range(3):
                for wall in (0.0, 1.0):
                    x = y.copy()
                    x[axis] = wall
                    E = fun(x, rng.uniform(0, 1))[:3]
                    tang = np.delete(E, axis)
                    assert np.abs(tang).max() < 1e-14
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
