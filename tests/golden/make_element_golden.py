"""Generate exact-rational golden vectors for the Tet10 element kernels.

Writes tests/golden/element_golden.json.  Everything here is computed from the
mathematics with sympy rationals, independently of the oracle and of the HIP
kernels:

  thermal   K_ij = k * integral(grad N_i . grad N_j) dV            (src/fem/element.rs:23-61)
  load      f_i  = Q * integral(N_i) dV                          (src/fem/element.rs:117-140)
  elastic   K    = integral(B^T D B) dV, Voigt [xx,yy,zz,xy,yz,zx] (src/mechanics/element.rs:33-65)
  gravity   f    = integral(N_i) dV * rho * g                    (src/mechanics/body_force.rs:49-84)

For affine (straight-sided) tets every integrand is a polynomial of degree
<= 2 in the barycentric coordinates, so the reference's 4-point rule is exact
and the floats derived from these rationals are what the FP64 kernels must
reproduce to rounding.

Cases: the reference tet with exact midpoints (the element every reference
unit test uses), and the six tets of the 6-tet (Kuhn) hex subdivision of
src/mesh_generator_improved.rs:105-118, after the same orientation fix
(:139-155), for two hex spacings: unit, and the 6x6x6 thermal benchmark's
(1000/6, 1000/6, 100).  Every element of the structured benchmark meshes is a
translate of one of these six shapes.

Run:  python tests/golden/make_element_golden.py
"""
import json
import os
from math import factorial

import sympy as S

HERE = os.path.dirname(os.path.abspath(__file__))
L1, L2, L3 = S.symbols("L1 L2 L3")
L0 = 1 - L1 - L2 - L3
LAM = [L0, L1, L2, L3]
EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]


def shape_funcs():
    N = [LAM[i] * (2 * LAM[i] - 1) for i in range(4)]
    N += [4 * LAM[a] * LAM[b] for a, b in EDGES]
    return [S.expand(n) for n in N]


def simplex_integral(expr):
    """Exact integral over the unit simplex {L1,L2,L3 >= 0, sum <= 1}:
    int L1^a L2^b L3^c = a! b! c! / (a+b+c+3)!."""
    poly = S.Poly(S.expand(expr), L1, L2, L3)
    tot = S.Integer(0)
    for (a, b, c), coef in poly.terms():
        tot += coef * S.Rational(factorial(a) * factorial(b) * factorial(c),
                                 factorial(a + b + c + 3))
    return tot


def tet10_nodes(v):
    v = [S.Matrix(p) for p in v]
    mids = [(v[a] + v[b]) / 2 for a, b in EDGES]
    return v + mids


def element_data(verts):
    verts = [[S.Rational(c) for c in p] for p in verts]
    v0 = S.Matrix(verts[0])
    J = S.Matrix.hstack(*[S.Matrix(verts[i]) - v0 for i in (1, 2, 3)])
    detJ = J.det()
    assert detJ != 0
    Jit = J.inv().T
    N = shape_funcs()
    grads = []
    for n in N:
        gref = S.Matrix([S.diff(n, s) for s in (L1, L2, L3)])
        grads.append(Jit * gref)
    absdet = abs(detJ)
    Kt = [[simplex_integral(grads[i].dot(grads[j])) * absdet for j in range(10)]
          for i in range(10)]
    intN = [simplex_integral(n) * absdet for n in N]
    nu = S.Rational(1, 4)
    fac = 1 / ((1 + nu) * (1 - 2 * nu))
    D = S.zeros(6, 6)
    for a in range(3):
        for b in range(3):
            D[a, b] = fac * (1 - nu if a == b else nu)
    for a in range(3, 6):
        D[a, a] = fac * (1 - 2 * nu) / 2
    B = S.zeros(6, 30)
    for i in range(10):
        dx, dy, dz = grads[i]
        c = 3 * i
        B[0, c] = dx
        B[1, c + 1] = dy
        B[2, c + 2] = dz
        B[3, c] = dy
        B[3, c + 1] = dx
        B[4, c + 1] = dz
        B[4, c + 2] = dy
        B[5, c + 2] = dx
        B[5, c] = dz
    BtDB = B.T * D * B
    Ke = [[simplex_integral(BtDB[a, b]) * absdet for b in range(30)]
          for a in range(30)]
    nodes = tet10_nodes(verts)
    return {
        "nodes": [[float(c) for c in p] for p in nodes],
        "volume": float(absdet / 6),
        "thermal_K_k1": [[float(x) for x in row] for row in Kt],
        "int_N": [float(x) for x in intN],
        "elastic_K_E1_nu025": [[float(x) for x in row] for row in Ke],
    }


def kuhn_tets(dx, dy, dz):
    """The six tets of one hex cell at the origin, vertex order as generated
    (src/mesh_generator_improved.rs:105-118) after the v2<->v3 swap when
    (v1-v0).((v2-v0)x(v3-v0)) < 0 (:148-155)."""
    P = {}
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                P[(i, j, k)] = [i * dx, j * dy, k * dz]
    tets = [((0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1)),
            ((0, 0, 0), (1, 0, 0), (1, 1, 1), (1, 0, 1)),
            ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 1, 1)),
            ((0, 0, 0), (0, 1, 0), (1, 1, 1), (0, 1, 1)),
            ((0, 0, 0), (0, 0, 1), (1, 1, 1), (1, 0, 1)),
            ((0, 0, 0), (0, 0, 1), (1, 1, 1), (0, 1, 1))]
    out = []
    for t in tets:
        v = [S.Matrix(P[c]) for c in t]
        det = (v[1] - v[0]).dot((v[2] - v[0]).cross(v[3] - v[0]))
        if det < 0:
            v[2], v[3] = v[3], v[2]
        out.append([list(p) for p in v])
    return out


def main():
    cases = {"reference_tet": element_data([[0, 0, 0], [1, 0, 0], [0, 1, 0],
                                            [0, 0, 1]])}
    spacings = {"unit": (1, 1, 1),
                "thermal_6x6x6": (S.Rational(1000, 6), S.Rational(1000, 6), 100)}
    for name, (dx, dy, dz) in spacings.items():
        for t, verts in enumerate(kuhn_tets(dx, dy, dz)):
            cases[f"kuhn_{name}_{t}"] = element_data(verts)
    path = os.path.join(HERE, "element_golden.json")
    with open(path, "w") as fh:
        json.dump({"generator": "tests/golden/make_element_golden.py",
                   "nu": 0.25, "cases": cases}, fh)
    print("wrote", path, len(cases), "cases")


if __name__ == "__main__":
    main()
