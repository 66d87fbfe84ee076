#!/usr/bin/env python3
"""Generate t10_tables.h: the constant tables of the 10-node (P2) tetrahedron.

Every table is derived here from the P2 basis by exact rational integration
(``fractions.Fraction``) and emitted as correctly rounded binary64 hex-float
literals, so the C host code and the HIP kernels share one bit-exact copy.

The reference element and its numbering follow the reference's T10 element
(src/mesh/t10.c:23-42, basis src/mesh/t10.c:1337-1349):

* nodes: vertices 0..3 = (0,0,0),(1,0,0),(0,1,0),(0,0,1); edge midpoints
  4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(1,3) 9:(2,3);
* faces F0=(0,3,2|7,9,6) F1=(0,1,3|4,8,7) F2=(0,2,1|6,5,4) F3=(1,2,3|5,9,8);
* outward reference normals -x, -y, -z and (1,1,1)/sqrt(3).

Tables (all exact except ``FACE_SCAL``):

* ``WPG[i]   = int phi_i``                               (src/mesh/t10.c:44-55)
* ``MASS[i][j] = int phi_i phi_j``                       (src/mesh/t10.c:57-132)
* ``INV_MASS`` = MASS^-1                                 (src/mesh/t10.c:134-167)
* ``INTERNAL[i][j][k] = int d_k(phi_i) phi_j``           (src/mesh/t10.c:169-691)
* ``MINV_INTERNAL[k][m][j] = sum_l INV_MASS[m][l] INTERNAL[l][j][k]``: the
  inverse mass applied to the volume operator of direction k;
* ``MINV_LIFT[f][m][b] = sum_a INV_MASS[m][node(f, a)] FACE_SCAL[a][b]``:
  the inverse mass applied to the face-f lift of the (rounded) face mass;
* ``FLUX_REF[i][j][f] = int_{face f} phi_i phi_j dS``, the exact face mass of
  the reference tetrahedron (faces 0..2 of area 1/2, face 3 of area sqrt(3)/2,
  so its entries carry a factor sqrt(3), evaluated to 60 digits and rounded
  once); the reference tabulates it (src/mesh/t10.c:693-1315) but its
  operators use FACE_SCAL;
* ``FACE_SCAL[a][b]``: the P2 mass matrix of the area-1/2 reference triangle
  in face-node order (v0, v1, v2, e01, e12, e20), rounded to 10 significant
  decimal digits.  The reference stores exactly that rounding
  (src/mesh/t10.c:1317-1330); the operator must use the same rounded values to
  agree with it, so the rounding rule is reproduced here instead of the exact
  fractions.

Run: ``python gen_t10_tables.py > t10_tables.h``.
"""
from decimal import Decimal, ROUND_HALF_EVEN, getcontext
from fractions import Fraction as Fr
from math import factorial
import sys

# ---------------------------------------------------------------- polynomials
# A polynomial in (x, y, z) is a dict {(a, b, c): Fraction coefficient}.


def padd(*ps):
    out = {}
    for p in ps:
        for k, v in p.items():
            out[k] = out.get(k, 0) + v
    return {k: v for k, v in out.items() if v != 0}


def pmul(p, q):
    out = {}
    for (a1, b1, c1), v1 in p.items():
        for (a2, b2, c2), v2 in q.items():
            k = (a1 + a2, b1 + b2, c1 + c2)
            out[k] = out.get(k, 0) + v1 * v2
    return {k: v for k, v in out.items() if v != 0}


def pscale(p, s):
    return {k: v * s for k, v in p.items()}


def pderiv(p, d):
    out = {}
    for k, v in p.items():
        if k[d] > 0:
            kk = list(k)
            kk[d] -= 1
            out[tuple(kk)] = out.get(tuple(kk), 0) + v * k[d]
    return out


def integrate_tet(p):
    """Exact integral over {x,y,z >= 0, x+y+z <= 1}: a!b!c!/(a+b+c+3)!."""
    s = Fr(0)
    for (a, b, c), v in p.items():
        s += v * Fr(factorial(a) * factorial(b) * factorial(c),
                    factorial(a + b + c + 3))
    return s


def integrate_tri(p):
    """Exact integral over {x,y >= 0, x+y <= 1} (z-exponent must be 0)."""
    s = Fr(0)
    for (a, b, c), v in p.items():
        assert c == 0
        s += v * Fr(factorial(a) * factorial(b), factorial(a + b + 2))
    return s


ONE = {(0, 0, 0): Fr(1)}
X = {(1, 0, 0): Fr(1)}
Y = {(0, 1, 0): Fr(1)}
Z = {(0, 0, 1): Fr(1)}
L0 = padd(ONE, pscale(X, -1), pscale(Y, -1), pscale(Z, -1))  # 1 - x - y - z


def quad_vertex(l):
    """l (2 l - 1)"""
    return pmul(l, padd(pscale(l, 2), pscale(ONE, -1)))


def quad_edge(l1, l2):
    return pscale(pmul(l1, l2), 4)


BARY = [L0, X, Y, Z]
EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]  # nodes 4..9
PHI = [quad_vertex(l) for l in BARY] + [quad_edge(BARY[a], BARY[b]) for a, b in EDGES]

# 2D P2 basis on the reference triangle in face-node order (v0,v1,v2,e01,e12,e20)
T0 = padd(ONE, pscale(X, -1), pscale(Y, -1))
TB = [T0, X, Y]
TRI = [quad_vertex(l) for l in TB] + [quad_edge(TB[0], TB[1]),
                                      quad_edge(TB[1], TB[2]),
                                      quad_edge(TB[2], TB[0])]


def invert(mat):
    n = len(mat)
    a = [row[:] + [Fr(int(i == j)) for j in range(n)] for i, row in enumerate(mat)]
    for c in range(n):
        p = next(r for r in range(c, n) if a[r][c] != 0)
        a[c], a[p] = a[p], a[c]
        piv = a[c][c]
        a[c] = [v / piv for v in a[c]]
        for r in range(n):
            if r != c and a[r][c] != 0:
                f = a[r][c]
                a[r] = [vr - f * vc for vr, vc in zip(a[r], a[c])]
    return [row[n:] for row in a]


def round_sig(fr, digits):
    getcontext().prec = 60
    d = Decimal(fr.numerator) / Decimal(fr.denominator)
    if d == 0:
        return Fr(0)
    exp = d.adjusted() - (digits - 1)
    q = d.quantize(Decimal(1).scaleb(exp), rounding=ROUND_HALF_EVEN)
    return Fr(q)


def lit(fr):
    """C99 hex-float literal of the correctly rounded binary64 value."""
    v = float(fr)
    if v == 0.0:
        return "0.0"
    return v.hex()


def main(out=sys.stdout):
    wpg = [integrate_tet(p) for p in PHI]
    mass = [[integrate_tet(pmul(p, q)) for q in PHI] for p in PHI]
    inv_mass = invert(mass)
    internal = [[[integrate_tet(pmul(pderiv(PHI[i], k), PHI[j])) for k in range(3)]
                 for j in range(10)] for i in range(10)]
    face = [[round_sig(integrate_tri(pmul(p, q)), 10) for q in TRI] for p in TRI]

    w = out.write
    w("/* Generated by gen_t10_tables.py -- do not edit.\n"
      " * Constant tables of the P2 (T10) reference tetrahedron, exact rationals\n"
      " * rounded once to binary64 (FACE_SCAL: 10 significant digits, see the\n"
      " * generator). Each macro expands to a brace initializer. */\n")
    w("#ifndef GDN_T10_TABLES_H\n#define GDN_T10_TABLES_H\n\n")

    def emit1(name, vals, comment):
        w(f"/* {comment} */\n#define {name} {{ \\\n")
        w("".join(f"\t{lit(v)}, /* {v} */ \\\n" for v in vals))
        w("}\n\n")

    def emit2(name, rows, comment):
        w(f"/* {comment} */\n#define {name} {{ \\\n")
        for r in rows:
            w("\t{ " + ", ".join(lit(v) for v in r) + " }, \\\n")
        w("}\n\n")

    emit1("T10_WPG_INIT", wpg, "int phi_i over the reference tetrahedron")
    emit2("T10_MASS_INIT", mass, "int phi_i phi_j")
    emit2("T10_INV_MASS_INIT", inv_mass, "inverse of the mass matrix")
    w("/* int d_k(phi_i) phi_j, indexed [i][j][k] */\n#define T10_INTERNAL_INIT { \\\n")
    for i in range(10):
        w("\t{ " + ", ".join("{ " + ", ".join(lit(internal[i][j][k]) for k in range(3)) + " }"
                             for j in range(10)) + " }, \\\n")
    w("}\n\n")
    getcontext().prec = 60
    sqrt3 = Decimal(3).sqrt()
    tri_exact = [[integrate_tri(pmul(p, q)) for q in TRI] for p in TRI]
    f2n_ = [(0, 3, 2, 7, 9, 6), (0, 1, 3, 4, 8, 7), (0, 2, 1, 6, 5, 4), (1, 2, 3, 5, 9, 8)]
    flux = [[[Fr(0)] * 4 for _ in range(10)] for _ in range(10)]
    for f in range(4):
        for a in range(6):
            for b in range(6):
                v = tri_exact[a][b]
                if f == 3:
                    v = Fr(Decimal(v.numerator) / Decimal(v.denominator) * sqrt3)
                flux[f2n_[f][a]][f2n_[f][b]][f] = v
    w("/* int over face f of phi_i phi_j (exact face mass; face 3 of area sqrt(3)/2),\n"
      " * indexed [i][j][f] */\n#define T10_FLUX_REF_INIT { \\\n")
    for i in range(10):
        w("\t{ " + ", ".join("{ " + ", ".join(lit(flux[i][j][f]) for f in range(4)) + " }"
                             for j in range(10)) + " }, \\\n")
    w("}\n\n")
    emit2("T10_FACE_SCAL_INIT", face,
          "P2 face mass, area-1/2 triangle, (v0,v1,v2,e01,e12,e20), 10 sig. digits")
    minv_int = [[[sum(inv_mass[m][l] * internal[l][j][k] for l in range(10))
                  for j in range(10)] for m in range(10)] for k in range(3)]
    f2n = [(0, 3, 2, 7, 9, 6), (0, 1, 3, 4, 8, 7), (0, 2, 1, 6, 5, 4), (1, 2, 3, 5, 9, 8)]
    minv_lift = [[[sum(inv_mass[m][f2n[f][a]] * face[a][b] for a in range(6))
                   for b in range(6)] for m in range(10)] for f in range(4)]
    w("/* inverse mass times the volume operator of direction k, [k][m][j] */\n"
      "#define T10_MINV_INTERNAL_INIT { \\\n")
    for k in range(3):
        w("\t{ " + ", ".join("{ " + ", ".join(lit(minv_int[k][m][j]) for j in range(10)) + " }"
                             for m in range(10)) + " }, \\\n")
    w("}\n\n")
    w("/* inverse mass times the lift of the rounded face mass, [f][m][b] */\n"
      "#define T10_MINV_LIFT_INIT { \\\n")
    for f in range(4):
        w("\t{ " + ", ".join("{ " + ", ".join(lit(minv_lift[f][m][b]) for b in range(6)) + " }"
                             for m in range(10)) + " }, \\\n")
    w("}\n\n")
    w("#endif\n")


if __name__ == "__main__":
    main()
