"""Generates tests/golden/tables_1d.json: exact 1D tables of the Q_k DG basis.

Independent of the oracle and of the GPU library: sympy builds the equidistant
Lagrange polynomials of R:dune/pdelab/finiteelement/qkdglagrange.hh:55-79 (nodes
j/k) symbolically, integrates the 1D mass / stiffness / convection matrices exactly
over [0,1] (rationals), and computes Gauss-Legendre nodes and weights on [0,1] as
roots of the Legendre polynomials to 40 digits (the rule family dune-geometry's
quadratureRule(geo, 2k) selects, R:dune/pdelab/common/quadraturerules.hh:110-120).

Run:  python tests/golden/make_tables_1d.py
"""
import json
import os

import sympy as sp

x = sp.Symbol("x")
DIGITS = 40


def lagrange(k):
    polys = []
    for i in range(k + 1):
        p = sp.Integer(1)
        for j in range(k + 1):
            if j != i:
                p *= (k * x - j) / sp.Integer(i - j)
        polys.append(sp.expand(p))
    return polys


def s(v):
    return str(sp.N(v, DIGITS))


def gauss(m):
    P = sp.legendre(m, 2 * x - 1)  # Legendre polynomial mapped to [0,1]
    roots = sorted(sp.Poly(P, x).nroots(n=DIGITS + 10))
    dP = sp.diff(sp.legendre(m, sp.Symbol("t")), sp.Symbol("t"))
    ws = []
    for r in roots:
        t = 2 * r - 1
        dv = dP.subs(sp.Symbol("t"), t)
        ws.append(sp.N(1 / ((1 - t * t) * dv * dv), DIGITS + 10))  # 2/((1-t^2)P'^2) / 2
    return roots, ws


def main():
    out = {"gauss": {}, "basis": {}}
    for m in range(1, 7):
        roots, ws = gauss(m)
        out["gauss"][str(m)] = {"x": [s(r) for r in roots], "w": [s(w) for w in ws]}
    for k in (1, 2, 3):
        ph = lagrange(k)
        dph = [sp.diff(p, x) for p in ph]
        n1 = k + 1
        M = [[s(sp.integrate(ph[i] * ph[j], (x, 0, 1))) for j in range(n1)] for i in range(n1)]
        K = [[s(sp.integrate(dph[i] * dph[j], (x, 0, 1))) for j in range(n1)] for i in range(n1)]
        G = [[s(sp.integrate(dph[i] * ph[j], (x, 0, 1))) for j in range(n1)] for i in range(n1)]
        ent = {
            "M": M, "K": K, "G": G,
            "e0": [s(p.subs(x, 0)) for p in ph], "e1": [s(p.subs(x, 1)) for p in ph],
            "d0": [s(p.subs(x, 0)) for p in dph], "d1": [s(p.subs(x, 1)) for p in dph],
            "int": [s(sp.integrate(p, (x, 0, 1))) for p in ph],
        }
        # basis values and derivatives at the Gauss nodes of the rule with k+1 points
        roots, _ = gauss(k + 1)
        ent["P_at_gauss"] = [[s(p.subs(x, r)) for p in ph] for r in roots]
        ent["D_at_gauss"] = [[s(p.subs(x, r)) for p in dph] for r in roots]
        out["basis"][str(k)] = ent
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables_1d.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
