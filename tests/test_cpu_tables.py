"""P2 reference-element tables: the library's generated tables, the oracle's
independently integrated ones and the reference's own tables (src/mesh/t10.c,
through the reference build in oracle/_ref) against each other and against
closed-form identities."""
from fractions import Fraction as Fr

import numpy as np
import pytest

import gdon3d as G
import oracle as O

FACE2NODE = [(0, 3, 2, 7, 9, 6), (0, 1, 3, 4, 8, 7), (0, 2, 1, 6, 5, 4), (1, 2, 3, 5, 9, 8)]


def product_tables():
    lib = G.load_library()
    wpg = np.array([lib.t10_get_wpg_ref_entry(i) for i in range(10)])
    mass = np.array([[lib.t10_get_mass_ref_entry(i, j) for j in range(10)] for i in range(10)])
    imass = np.array([[lib.t10_get_inv_mass_ref_entry(i, j) for j in range(10)] for i in range(10)])
    internal = np.array([[[lib.t10_get_internal_ref_entry(i, j, k) for k in range(3)]
                          for j in range(10)] for i in range(10)])
    face = np.array([[lib.t10_get_face_scale_entry(a, b) for b in range(6)] for a in range(6)])
    return dict(wpg=wpg, mass=mass, inv_mass=imass, internal=internal, face=face)


def ref_tables():
    import ctypes as C
    R = O.ref()
    for name, n in [("t10_get_wpg_ref_entry", 1), ("t10_get_mass_ref_entry", 2),
                    ("t10_get_inv_mass_ref_entry", 2), ("t10_get_internal_ref_entry", 3),
                    ("t10_get_face_scale_entry", 2)]:
        getattr(R, name).restype = C.c_double
        getattr(R, name).argtypes = [C.c_int] * n
    return dict(
        wpg=np.array([R.t10_get_wpg_ref_entry(i) for i in range(10)]),
        mass=np.array([[R.t10_get_mass_ref_entry(i, j) for j in range(10)] for i in range(10)]),
        inv_mass=np.array([[R.t10_get_inv_mass_ref_entry(i, j) for j in range(10)]
                           for i in range(10)]),
        internal=np.array([[[R.t10_get_internal_ref_entry(i, j, k) for k in range(3)]
                            for j in range(10)] for i in range(10)]),
        face=np.array([[R.t10_get_face_scale_entry(a, b) for b in range(6)] for a in range(6)]))


def exact_face_mass():
    """P2 mass matrix of the area-1/2 triangle, (v0,v1,v2,e01,e12,e20)."""
    S = [[None] * 6 for _ in range(6)]
    for a in range(6):
        for b in range(6):
            if a < 3 and b < 3:
                S[a][b] = Fr(1, 60) if a == b else Fr(-1, 360)
            elif a >= 3 and b >= 3:
                S[a][b] = Fr(4, 45) if a == b else Fr(2, 45)
            else:
                v, e = (a, b - 3) if a < 3 else (b, a - 3)
                # edge e joins vertices e and e+1: vertex opposite to it couples
                S[a][b] = Fr(-1, 90) if v == (e + 2) % 3 else Fr(0)
    return S


def test_product_tables_match_oracle_integration():
    p, o = product_tables(), O.tables()
    for k in ("wpg", "mass", "internal"):
        assert np.allclose(p[k], o[k], rtol=0, atol=1e-15), k
    assert np.abs(p["inv_mass"] - o["inv_mass"]).max() < 5e-14 * np.abs(p["inv_mass"]).max()
    assert np.array_equal(p["face"], o["face"])


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_tables_match_reference_build():
    p, r = product_tables(), ref_tables()
    # the reference stores the face mass rounded to 10 digits: bit-identical
    assert np.array_equal(p["face"], r["face"])
    # exact rationals vs the reference's floating-point tables
    assert np.max(np.abs(p["wpg"] - r["wpg"])) <= 1e-17
    assert np.max(np.abs(p["mass"] - r["mass"])) <= 1e-17
    assert np.max(np.abs(p["inv_mass"] - r["inv_mass"])) <= 1e-12
    assert np.max(np.abs(p["internal"] - r["internal"])) <= 1e-17


def test_table_identities():
    p = product_tables()
    assert abs(p["wpg"].sum() - 1.0 / 6.0) < 1e-16
    assert np.allclose(p["mass"].sum(axis=1), p["wpg"], atol=1e-17)
    assert np.allclose(p["mass"] @ p["inv_mass"], np.eye(10), atol=1e-12)
    # partition of unity: sum_i d_k phi_i = 0
    assert np.abs(p["internal"].sum(axis=0)).max() < 1e-16


def test_internal_table_integration_by_parts():
    """I_ijk + I_jik = boundary integral of n_k phi_i phi_j, assembled from the
    exact face mass: faces 0-2 have normal -e_f and area 1/2, face 3 has
    normal (1,1,1)/sqrt 3 and area sqrt(3)/2 (so n_k * area factor = 1)."""
    I = product_tables()["internal"]
    S = np.array([[float(x) for x in row] for row in exact_face_mass()])
    B = np.zeros((10, 10, 3))
    for f in range(4):
        nodes = FACE2NODE[f]
        for a in range(6):
            for b in range(6):
                for k in range(3):
                    nk = (-1.0 if k == f else 0.0) if f < 3 else 1.0
                    B[nodes[a], nodes[b], k] += nk * S[a, b]
    assert np.abs(I + I.transpose(1, 0, 2) - B).max() < 1e-15


def test_face_mass_rounding():
    p = product_tables()["face"]
    S = exact_face_mass()
    for a in range(6):
        for b in range(6):
            ex = float(S[a][b])
            if ex == 0:
                assert p[a, b] == 0.0
            else:
                assert abs(p[a, b] - ex) / abs(ex) < 2.1e-10
                assert p[a, b] == float(f"{ex:.9e}")  # 10 significant digits
