"""Writes tests/golden/lgf_golden.json: Laplace LGF values by scipy quadrature
of G(n) = -int_0^inf prod_i ive(n_i, 2t) dt (the definition in the
reference's table generator, include/iblgf/lgf/lgf_table_gen/lgf_integrator.py:7-25,
restated here with scipy as an independent library routine), and scaled Bessel
values ive(k, x) for the integrating-factor kernel.  Run from the repo root:
    python tests/golden/make_lgf_golden.py
"""
import json
import os

import numpy as np
from scipy import integrate, special

POINTS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 1, 1), (2, 1, 0), (3, 2, 1), (5, 0, 0), (7, 3, 2), (10, 10, 10),
          (20, 5, 3), (40, 0, 0), (63, 31, 7)]


def lgf(n):
    f = lambda t: np.prod([special.ive(k, 2.0 * t) for k in n])  # noqa: E731
    # split the range so quad resolves the peak and the t^-3/2 tail
    edges = [0.0, 1.0, 10.0, 100.0, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8]
    s = sum(integrate.quad(f, a, b, epsabs=1e-15, epsrel=1e-13, limit=400)[0] for a, b in zip(edges, edges[1:]))
    # tail beyond 1e8: (4 pi t)^{-3/2} (1 - sum(4k^2-1)/(16 t)) integrated
    T = edges[-1]
    m = sum(4 * k * k - 1 for k in n)
    s += (4 * np.pi) ** -1.5 * (2.0 / np.sqrt(T) - m / 24.0 * T ** -1.5)
    return -s


def main():
    out = {"lgf": [[list(p), lgf(p)] for p in POINTS],
           "ive": [[k, x, float(special.ive(k, x))] for k in (0, 1, 2, 5, 12, 25, 40) for x in (0.05, 0.2, 1.0, 3.4, 8.0)]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lgf_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
