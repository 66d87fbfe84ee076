"""The rest of the reference's public C API (src/models/d3q4.h:32-44,
src/models/d3q6.h, src/mesh/tetrageometry.h:30-32, src/mesh/t10.h:85): the
kinetic moment transforms Q / Q^-1, the projector boundary state that the
reference's get_fR computes with use_feq = false, the inverse geometric map
and the reference face mass, each against the reference's own sources built
in oracle/_ref and against the mathematics; and the reference's example
drivers compiled unchanged against include/ and linked with the library
(build and link only: no GPU here)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import gdon3d as G
import oracle as O

dp = C.POINTER(C.c_double)
rng = np.random.default_rng(77)
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def model(L, kin):
    """a gdn_model of library L with Maxwell fluxes and projectors"""
    m = G.GdnModel()
    m.nb_w = 6
    for field, name in (("get_num_flux", "maxwell_num_flux_upwind"),
                        ("get_proj_pos", "maxwell_get_proj_pos_null_prod_w"),
                        ("get_proj_neg", "maxwell_get_proj_neg_prod_w")):
        setattr(m, field, C.cast(getattr(L, name), C.c_void_p).value)
    f = getattr(L, kin + "_set_model")
    f.argtypes = [C.c_void_p]
    f(C.byref(m))
    return m


def call(L, name, m, *arrays):
    f = getattr(L, name)
    f.argtypes = [C.c_void_p] + [dp] * len(arrays)
    f.restype = None
    f(C.byref(m), *[G.dptr(a) for a in arrays])


NM = {"d3q4": (4, 3), "d3q6": (6, 5)}  # velocities, moments beyond w


@needs_ref
@pytest.mark.parametrize("kin", ["d3q4", "d3q6"])
def test_moment_transforms_match_reference_build(kin):
    lib, R = G.load_library(), O.ref()
    nv, nm = NM[kin]
    ml, mr = model(lib, kin), model(R, kin)
    for _ in range(50):
        f = rng.normal(size=6 * nv)
        outs = {}
        for tag, L, m in (("lib", lib, ml), ("ref", R, mr)):
            w, y = np.zeros(6), np.zeros(6 * nm)
            call(L, kin + "_apply_Q", m, f, w, y)
            back = np.zeros(6 * nv)
            call(L, kin + "_apply_Q_inv", m, w, y, back)
            outs[tag] = (w, y, back)
            if kin == "d3q6":
                q, qM, fm = np.zeros(6 * nm), np.zeros(6 * nm), np.zeros(6 * nv)
                call(L, "d3q6_apply_M", m, f, w, q)
                call(L, "d3q6_get_qM", m, w, qM)
                call(L, "d3q6_apply_M_inv", m, w, q, fm)
                outs[tag] += (q, qM, fm)
        for a, b in zip(outs["lib"], outs["ref"]):
            assert np.allclose(a, b, rtol=1e-14, atol=1e-14)
        # Q is invertible: Q^-1(Q(f)) = f
        assert np.allclose(outs["lib"][2], f, rtol=1e-13, atol=1e-13)


@needs_ref
def test_philippe_get_f_matches_reference_build():
    lib, R = G.load_library(), O.ref()
    ml, mr = model(lib, "d3q4"), model(R, "d3q4")
    for _ in range(20):
        w = rng.normal(size=6)
        a, b = np.zeros(24), np.zeros(24)
        call(lib, "test_philippe_d3q4_get_f", ml, w, a)
        call(R, "test_philippe_d3q4_get_f", mr, w, b)
        assert np.allclose(a, b, rtol=1e-14, atol=1e-14)


def proj(L, name, w, n):
    f = getattr(L, name)
    f.argtypes = [dp] * 3
    f.restype = None
    o = np.zeros(6)
    f(G.dptr(w), G.dptr(n), G.dptr(o))
    return o


@needs_ref
@pytest.mark.parametrize("kin", ["d3q4", "d3q6"])
def test_projector_boundary_state_composes_reference_functions(kin):
    """get_fR with use_feq = false is not reachable in the reference build
    (a const bool selects f_eq); its branch is composed here from the
    reference's own apply_Q, projectors and apply_Q_inv exactly as
    src/models/d3q4.c:316-331 / d3q6.c:362-378 compose them, and the
    library's d3qN_get_fR_projector must give the same."""
    lib, R = G.load_library(), O.ref()
    nv, nm = NM[kin]
    ml, mr = model(lib, kin), model(R, kin)
    for _ in range(50):
        fL, wI, vn = rng.normal(size=6 * nv), rng.normal(size=6), rng.normal(size=3)
        wL, yL = np.zeros(6), np.zeros(6 * nm)
        call(R, kin + "_apply_Q", mr, fL, wL, yL)
        pn = proj(R, "maxwell_get_proj_neg_prod_w", wI, vn)
        if kin == "d3q4":
            wR = pn  # the reference overwrites the sum with P^- wI
        else:
            wR = proj(R, "maxwell_get_proj_pos_null_prod_w", wL, vn) + pn
        expect = np.zeros(6 * nv)
        call(R, kin + "_apply_Q_inv", mr, wR, yL, expect)
        got = np.zeros(6 * nv)
        f = getattr(lib, kin + "_get_fR_projector")
        f.argtypes = [C.c_void_p, dp, dp, dp, dp]
        f(C.byref(ml), G.dptr(fL), G.dptr(vn), G.dptr(wI), G.dptr(got))
        assert np.allclose(got, expect, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("kin", ["d3q4", "d3q6"])
def test_projector_boundary_state_properties(kin):
    """What the branch fixes: the macro state of fR is P^- wI (D3Q4) or
    P^{+,0} wL + P^- wI (D3Q6), its non-equilibrium part is fL's; with
    wI = wL and fL at equilibrium the D3Q6 state reproduces fL."""
    lib = G.load_library()
    nv, nm = NM[kin]
    m = model(lib, kin)
    for _ in range(20):
        fL, wI, vn = rng.normal(size=6 * nv), rng.normal(size=6), rng.normal(size=3)
        fR = np.zeros(6 * nv)
        f = getattr(lib, kin + "_get_fR_projector")
        f.argtypes = [C.c_void_p, dp, dp, dp, dp]
        f(C.byref(m), G.dptr(fL), G.dptr(vn), G.dptr(wI), G.dptr(fR))
        wL, yL, wR, yR = np.zeros(6), np.zeros(6 * nm), np.zeros(6), np.zeros(6 * nm)
        call(lib, kin + "_apply_Q", m, fL, wL, yL)
        call(lib, kin + "_apply_Q", m, fR, wR, yR)
        pn = proj(lib, "maxwell_get_proj_neg_prod_w", wI, vn)
        expect = pn if kin == "d3q4" else proj(lib, "maxwell_get_proj_pos_null_prod_w", wL,
                                                   vn) + pn
        assert np.allclose(wR, expect, atol=1e-13) and np.allclose(yR, yL, atol=1e-12)
    if kin == "d3q6":
        w = rng.normal(size=6)
        feq = np.zeros(6 * nv)
        call(lib, "d3q6_get_feq", m, w, feq)
        fR = np.zeros(6 * nv)
        f(C.byref(m), G.dptr(feq), G.dptr(rng.normal(size=3)), G.dptr(w), G.dptr(fR))
        assert np.allclose(fR, feq, atol=1e-13)


def geo(L, name, pn, x):
    f = getattr(L, name)
    f.argtypes = [dp, dp, dp]
    f.restype = None
    out = np.zeros(3)
    f(G.dptr(np.ascontiguousarray(pn)), G.dptr(np.ascontiguousarray(x)), G.dptr(out))
    return out


def ref2phy(L, pn, xr):
    f = L.gdn_tetra_ref2phy
    f.argtypes = [dp, dp, dp, C.c_int, dp, dp, dp, dp, dp]
    f.restype = None
    out = np.zeros(3)
    f(G.dptr(np.ascontiguousarray(pn)), G.dptr(np.ascontiguousarray(xr)), None, -1,
      G.dptr(out), None, None, None, None)
    return out


def random_tet():
    while True:
        pn = rng.normal(size=(4, 3))
        if np.linalg.det((pn[1:] - pn[0]).T) > 0.1:
            return pn


def test_phy2ref_inverts_ref2phy():
    lib = G.load_library()
    for _ in range(100):
        pn, xr = random_tet(), rng.uniform(-0.5, 1.5, 3)
        x = ref2phy(lib, pn, xr)
        assert np.allclose(geo(lib, "gdn_tetra_phy2ref", pn, x), xr, atol=1e-12)
        # inside the unit cube of reference coordinates the homotopy reaches xphy
        xin = rng.uniform(0.05, 0.95, 3)
        assert np.allclose(geo(lib, "gdn_tetra_robust_phy2ref", pn, ref2phy(lib, pn, xin)),
                           xin, atol=1e-12)


@needs_ref
def test_phy2ref_matches_reference_build():
    lib, R = G.load_library(), O.ref()
    for _ in range(100):
        pn = random_tet()
        for xr in (rng.uniform(0.05, 0.95, 3), rng.uniform(-1.0, 2.0, 3)):
            x = ref2phy(R, pn, xr)
            for name in ("gdn_tetra_phy2ref", "gdn_tetra_robust_phy2ref"):
                a, b = geo(lib, name, pn, x), geo(R, name, pn, x)
                assert np.allclose(a, b, rtol=1e-11, atol=1e-11), (name, a, b)


def flux_ref(L):
    f = L.t10_get_flux_ref_entry
    f.argtypes = [C.c_int] * 3
    f.restype = C.c_double
    return np.array([[[f(i, j, k) for k in range(4)] for j in range(10)] for i in range(10)])


def test_flux_ref_is_the_face_mass():
    """int over face f of phi_i phi_j: symmetric, zero off the face, rows
    summing to int phi_i over the face, total = face area (1/2 on faces
    0..2, sqrt(3)/2 on face 3)."""
    T = flux_ref(G.load_library())
    f2n = [(0, 3, 2, 7, 9, 6), (0, 1, 3, 4, 8, 7), (0, 2, 1, 6, 5, 4), (1, 2, 3, 5, 9, 8)]
    for f in range(4):
        off = [i for i in range(10) if i not in f2n[f]]
        assert np.all(T[off, :, f] == 0) and np.all(T[:, off, f] == 0)
        assert np.array_equal(T[:, :, f], T[:, :, f].T)
        area = 0.5 if f < 3 else np.sqrt(3) / 2
        assert abs(T[:, :, f].sum() - area) < 1e-15
        # P2 on a triangle: vertex nodes integrate to 0, edge nodes to area / 3
        rows = T[:, :, f].sum(axis=1)
        assert np.allclose(rows[list(f2n[f][:3])], 0, atol=1e-16)
        assert np.allclose(rows[list(f2n[f][3:])], area / 3, atol=1e-16)


@needs_ref
def test_flux_ref_matches_reference_build():
    """The reference tabulates the face mass from a numerical integration:
    its entries carry rounding noise (1e-18 where the exact value is 0, two
    ulps on 0.15), the library's are the exact values rounded once."""
    a, b = flux_ref(G.load_library()), flux_ref(O.ref())
    assert np.abs(a - b).max() < 5e-17


REF_EXAMPLES = "/root/reference/examples"


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples absent")
@pytest.mark.parametrize("name", ["rk3_maxwell_plane_wave", "relax_maxwell_plane_wave",
                                  "relax_maxwell_stationary_wave"])
def test_reference_example_builds_against_include(name, tmp_path):
    """The reference's own drivers compile unchanged against include/ and
    link against libgdon3d_hip.so (-O2: rk3_maxwell_plane_wave.c:87 calls
    the never-defined simulation_display_info in unreachable code, which the
    reference's own Release build also relies on being removed)."""
    lib_dir = os.path.dirname(G.LIB_PATH)
    exe = tmp_path / name
    r = subprocess.run(["gcc", "-std=gnu99", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
                        os.path.join(REF_EXAMPLES, name + ".c"), "-L", lib_dir, "-lgdon3d_hip",
                        f"-Wl,-rpath,{lib_dir}", "-lm", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()


@needs_ref
@pytest.mark.parametrize("kin", ["d3q4", "d3q6"])
def test_oracle_projector_boundary_state_matches_reference(kin):
    """The oracle's restatement (6x6 eigen-projectors built from A(n) and
    its own moment transforms) against the reference's projectors and its
    apply_Q / apply_Q_inv composed as get_fR's use_feq = false branch."""
    R = O.ref()
    nv, nm = NM[kin]
    mr = model(R, kin)
    for _ in range(50):
        w, n = rng.normal(size=6), rng.normal(size=3)
        assert np.allclose(O.proj(w, n, True), proj(R, "maxwell_get_proj_neg_prod_w", w, n),
                           rtol=1e-13, atol=1e-13)
        assert np.allclose(O.proj(w, n, False),
                           proj(R, "maxwell_get_proj_pos_null_prod_w", w, n),
                           rtol=1e-13, atol=1e-13)
        fL, wI, vn = rng.normal(size=6 * nv), rng.normal(size=6), rng.normal(size=3)
        wL, yL = np.zeros(6), np.zeros(6 * nm)
        call(R, kin + "_apply_Q", mr, fL, wL, yL)
        pn = proj(R, "maxwell_get_proj_neg_prod_w", wI, vn)
        wR = pn if kin == "d3q4" else proj(R, "maxwell_get_proj_pos_null_prod_w", wL, vn) + pn
        expect = np.zeros(6 * nv)
        call(R, kin + "_apply_Q_inv", mr, wR, yL, expect)
        got = O.fr_projector(O.D3Q4 if kin == "d3q4" else O.D3Q6, fL, vn, wI)
        assert np.allclose(got, expect, rtol=1e-13, atol=1e-13)
