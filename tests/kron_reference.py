"""Dense tensor-product ("Kronecker") form of the SIPG/NIPG/IIPG operator, built in
numpy from the exact sympy tables of tests/golden/tables_1d.json.

This is an independent second formulation of ConvectionDiffusionDG on an
axis-aligned grid (R:dune/pdelab/localoperator/convectiondiffusiondg.hh:106-188,
271-471, 684-879): per direction d the 1D DG operator A_d (volume S kappa K - Sc b G,
rank-one face terms in the traces e0/e1 and derivatives d0/d1) tensored with the
tangential 1D mass matrices. The oracle evaluates the same operator with
quadrature-point loops; agreement of the two to rounding pins both. Tiny grids only.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_tables(k):
    with open(os.path.join(HERE, "golden", "tables_1d.json")) as f:
        t = json.load(f)["basis"][str(k)]
    arr = lambda v: np.array(v, dtype=np.float64)  # noqa: E731
    return {key: arr([[float(s) for s in row] for row in t[key]]) if isinstance(t[key][0], list)
            else arr([float(s) for s in t[key]]) for key in t}


def face_coeffs(klo, khi, weights_on, alphak):
    if weights_on:
        den = khi + klo + 1e-20
        omega_s, omega_n = klo / den, khi / den
        return omega_n * klo, omega_s * khi, alphak * (2.0 * khi * klo / den)
    return 0.5 * klo, 0.5 * khi, alphak


def dense_operator(cfg, kappa=None, transpose=False):
    k = cfg.degree
    T = load_tables(k)
    M, K, G = T["M"], T["K"], T["G"]
    e0, e1, d0, d1 = T["e0"], T["e1"], T["d0"], T["d1"]
    n = [cfg.nx, cfg.ny, cfg.nz]
    L = [cfg.lx, cfg.ly, cfg.lz]
    h = [L[d] / n[d] for d in range(3)]
    V = h[0] * h[1] * h[2]
    S = [V / (h[d] * h[d]) for d in range(3)]
    Sc = [V / h[d] for d in range(3)]
    theta = {0: 1.0, 1: -1.0, 2: 0.0}[cfg.method]
    alphak = cfg.alpha * k * (k + 2)
    won = cfg.weights == 0
    b = [cfg.b[d] for d in range(3)]
    bc = [1 if cfg.bctype[s] == 0 else cfg.bctype[s] for s in range(6)]
    nc = n[0] * n[1] * n[2]
    n1 = k + 1
    nloc = n1 ** 3
    kap = np.full(nc, cfg.kappa0) if kappa is None else np.asarray(kappa, dtype=np.float64)
    A = np.zeros((nc * nloc, nc * nloc))
    o = np.outer
    stride = [1, n[0], n[0] * n[1]]

    def kron3(B, d):
        D = [M, M, M]
        D[d] = B
        return np.kron(D[2], np.kron(D[1], D[0]))

    def add(r, c, B):
        A[r * nloc:(r + 1) * nloc, c * nloc:(c + 1) * nloc] += B

    for c in range(nc):
        co = [c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1])]
        kc = kap[c]
        for d in range(3):
            Sd, Scd, bd = S[d], Sc[d], b[d]
            own = Sd * kc * K - Scd * bd * G
            if d == 0:
                own = own + V * cfg.c * M
            if co[d] > 0:
                cl = c - stride[d]
                wlo, whi, P = face_coeffs(kap[cl], kc, won, alphak)
                own = own + Sd * (whi * o(e0, d0) + P * o(e0, e0) - theta * whi * o(d0, e0))
                Bl = Sd * (wlo * o(e0, d1) - P * o(e0, e1) + theta * whi * o(d0, e1))
                if bd <= 0:
                    own = own + Scd * (-bd) * o(e0, e0)
                else:
                    Bl = Bl + Scd * (-bd) * o(e0, e1)
                add(c, cl, kron3(Bl, d))
            else:
                if bc[2 * d] == 1:
                    Pb = alphak * (kc if won else 1.0)
                    own = own + Sd * (kc * o(e0, d0) + Pb * o(e0, e0) - theta * kc * o(d0, e0))
                    if bd <= 0:
                        own = own + Scd * (-bd) * o(e0, e0)
                elif bc[2 * d] == -2:
                    own = own + Scd * (-bd) * o(e0, e0)
            if co[d] < n[d] - 1:
                cr = c + stride[d]
                wlo, whi, P = face_coeffs(kc, kap[cr], won, alphak)
                own = own + Sd * (-wlo * o(e1, d1) + P * o(e1, e1) + theta * wlo * o(d1, e1))
                Br = Sd * (-whi * o(e1, d0) - P * o(e1, e0) - theta * wlo * o(d1, e0))
                if bd > 0:
                    own = own + Scd * bd * o(e1, e1)
                else:
                    Br = Br + Scd * bd * o(e1, e0)
                add(c, cr, kron3(Br, d))
            else:
                if bc[2 * d + 1] == 1:
                    Pb = alphak * (kc if won else 1.0)
                    own = own + Sd * (-kc * o(e1, d1) + Pb * o(e1, e1) + theta * kc * o(d1, e1))
                    if bd >= 0:
                        own = own + Scd * bd * o(e1, e1)
                elif bc[2 * d + 1] == -2:
                    own = own + Scd * bd * o(e1, e1)
            add(c, c, kron3(own, d))
    return A.T.copy() if transpose else A
