"""Write the Gauss-Lobatto-Legendre fixtures `gll.json`.

Raven builds its 1-D operators in BigFloat at 256 bits and rounds them to the
cell's float type (src/lobattocells.jl:1-47, precision rule at :7).  This script
repeats that with mpmath at 256 bits from closed forms, independently of the
library and of the oracle:

* nodes: +-1 and the roots of P_N'(x) (Newton from Chebyshev-Lobatto guesses);
* weights: w_i = 2 / (N (N+1) P_N(x_i)^2);
* derivative: D[a][b] = l_b'(x_a) = P_N(x_a) / (P_N(x_b) (x_a - x_b)) for a != b,
  D[0][0] = -N(N+1)/4, D[N][N] = N(N+1)/4, other diagonal entries 0;
* tohalves: barycentric interpolation from the nodes to (x-1)/2 and (x+1)/2
  (lobattocells.jl:15-16).

Values are stored as float.hex strings rounded to nearest (float64) and, via
numpy, to float32.  Run:  python tests/golden/gen_gll.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.prec = 256


def legendre_and_deriv(n, x):
    p0, p1 = mp.mpf(1), x
    if n == 0:
        return p0, mp.mpf(0)
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    # P_n'(x) = n (x P_n - P_{n-1}) / (x^2 - 1)
    dp = n * (x * p1 - p0) / (x * x - 1) if abs(x) != 1 else None
    return p1, dp


def gll(nq):
    n = nq - 1
    xs = [mp.mpf(-1)]
    for i in range(1, n):
        x = -mp.cos(mp.pi * i / n)
        for _ in range(200):
            # f = P_n'(x); f' = P_n''(x) = (2x P_n' - n(n+1) P_n) / (1 - x^2)
            p, dp = legendre_and_deriv(n, x)
            ddp = (2 * x * dp - n * (n + 1) * p) / (1 - x * x)
            dx = dp / ddp
            x -= dx
            if abs(dx) < mp.mpf(2) ** (-250):
                break
        xs.append(x)
    xs.append(mp.mpf(1))
    # enforce the node symmetry x_i = -x_{N-i} (centre node exactly 0 for odd Nq)
    for i in range(nq // 2):
        xs[n - i] = -xs[i]
    if nq % 2 == 1:
        xs[n // 2] = mp.mpf(0)
    pn = [legendre_and_deriv(n, x)[0] for x in xs]
    ws = [mp.mpf(2) / (n * (n + 1) * p * p) for p in pn]
    D = [[mp.mpf(0)] * nq for _ in range(nq)]
    for a in range(nq):
        for b in range(nq):
            if a != b:
                D[a][b] = pn[a] / (pn[b] * (xs[a] - xs[b]))
    D[0][0] = -mp.mpf(n * (n + 1)) / 4
    D[n][n] = mp.mpf(n * (n + 1)) / 4
    return xs, ws, D


def interp_matrix(xs, ys):
    """Lagrange interpolation matrix I[y][b] = l_b(y)."""
    m = len(xs)
    out = []
    for y in ys:
        row = []
        for b in range(m):
            v = mp.mpf(1)
            for c in range(m):
                if c != b:
                    v *= (y - xs[c]) / (xs[b] - xs[c])
            row.append(v)
        out.append(row)
    return out


def h64(v):
    return float(v).hex()


def h32(v):
    # mpmath -> nearest float32 directly (not via float64) to avoid double rounding
    with mp.workprec(24):
        f = mp.mpf(v)
    return float(np.float32(float(f))).hex()


def main():
    out = {"source": "mpmath 256-bit closed forms; see gen_gll.py", "cells": {}}
    for nq in range(2, 17):
        xs, ws, D = gll(nq)
        lo = interp_matrix(xs, [(x - 1) / 2 for x in xs])
        hi = interp_matrix(xs, [(x + 1) / 2 for x in xs])
        out["cells"][str(nq)] = {
            "points_f64": [h64(x) for x in xs],
            "weights_f64": [h64(w) for w in ws],
            "D_f64": [[h64(v) for v in row] for row in D],
            "tolowerhalf_f64": [[h64(v) for v in row] for row in lo],
            "toupperhalf_f64": [[h64(v) for v in row] for row in hi],
            "points_f32": [h32(x) for x in xs],
            "weights_f32": [h32(w) for w in ws],
            "D_f32": [[h32(v) for v in row] for row in D],
        }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gll.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
