"""Write the golden fixtures of tests/test_oracle.py.

The reference ships no fixtures or known-answer tests, so these vectors are
produced by the independent brute-force enumeration (oracle.brute_force, no
cell list) on small seeded inputs, with exact rational FMA emulation where the
coordinates are not dyadic.  Each fixture stores the inputs, edges, box and
mode with the expected uint64 counts and FP64 weighted sums.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import brute_force  # noqa: E402


def dyadic(rng, n, lo, hi, scale=64.0):
    return np.round(rng.uniform(lo, hi, size=(n, 3)) * scale) / scale


def case_survey(rng):
    n = 90
    p = np.ones((n, 4))
    p[:, :3] = dyadic(rng, n, 0.0, 30.0)
    p[:, 3] = np.round(rng.lognormal(0, 0.3, n) * 1024) / 1024
    edges = np.array([0.0, 1.5, 3.0, 4.5, 6.0, 7.5, 9.0])
    box = p[:, :3].max(axis=0) - p[:, :3].min(axis=0) + 1.0
    p[:, :3] -= p[:, :3].min(axis=0)
    c, w = brute_force(p, p, p, edges, None)
    return dict(p1=p, p2=p, p3=p, edges=edges, box=box, survey=1, counts=c, wsum=w)


def case_box(rng, m, nb, n, weighted):
    rmax = 8.0
    L = m * rmax
    p = np.ones((n, 4))
    p[:, :3] = dyadic(rng, n, 0.0, L)
    p[:, :3] = np.mod(p[:, :3], L)
    if weighted:
        p[:, 3] = np.round(rng.lognormal(0, 0.3, n) * 1024) / 1024
    edges = np.linspace(0.0, rmax, nb + 1)
    box = np.array([L, L, L])
    c, w = brute_force(p, p, p, edges, box)
    return dict(p1=p, p2=p, p3=p, edges=edges, box=box, survey=0, counts=c, wsum=w)


def case_three_catalogues(rng):
    rmax = 6.0
    L = 4 * rmax
    cats = []
    for n in (40, 50, 60):
        q = np.ones((n, 4))
        q[:, :3] = dyadic(rng, n, 0.0, L)
        cats.append(q)
    edges = np.array([0.5, 2.0, 3.5, 5.0, 6.0])
    box = np.array([L, L, L])
    c, w = brute_force(cats[0], cats[1], cats[2], edges, box)
    return dict(p1=cats[0], p2=cats[1], p3=cats[2], edges=edges, box=box, survey=0, counts=c, wsum=w)


def case_exact_fma(rng):
    """Non-dyadic FP64 coordinates: expected values use exact FMA emulation of
    fma(dz,dz,fma(dy,dy,dx*dx)) (slow; kept small)."""
    n = 28
    p = np.ones((n, 4))
    p[:, :3] = rng.uniform(0.0, 12.0, size=(n, 3))
    edges = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    box = p[:, :3].max(axis=0) + 1.0
    c, w = brute_force(p, p, p, edges, None, exact=True)
    return dict(p1=p, p2=p, p3=p, edges=edges, box=box, survey=1, counts=c, wsum=w)


def main():
    rng = np.random.default_rng(20261203)
    cases = {
        "survey_dyadic": case_survey(rng),
        "box_m3": case_box(rng, 3, 6, 120, weighted=False),
        "box_m4_weighted": case_box(rng, 4, 5, 150, weighted=True),
        "box_three_catalogues": case_three_catalogues(rng),
        "survey_exact_fma": case_exact_fma(rng),
    }
    for name, d in cases.items():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **d)
        print(name, int(d["counts"].sum()), "triplets ->", os.path.relpath(path))


if __name__ == "__main__":
    main()
