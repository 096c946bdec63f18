"""Seeded synthetic catalogues for the benchmark workloads.

The reference ships no data and its own random catalogue
(generateRandomCatalog, source/utils.cu:41-52) is seeded from time(0), so it
cannot be reproduced.  These generators use numpy's PCG64 with fixed seeds so
every run (here and on the GPU host) sees the same points.  All catalogues
are (n, 4) float64 arrays of x, y, z, w (the reference's double4).

Workloads (edges in r, nb = len(edges) - 1):
  box_uniform      20,000 uniform points in a 250 Mpc/h box, 10 bins over [2, 20]
  box_clustered    Neyman-Scott box, L=1000, 10,000 parents x 100 children,
                   sigma = 5, float32 coordinates, 15 bins over [0, 30]
  survey_cap       cap RA 0-60 deg, Dec 0-60 deg, comoving distance uniform in
                   volume over [1000, 2000] Mpc/h, lognormal(0, 0.3) float32
                   weights for data, unit weights for randoms
  batched_boxes    64 Neyman-Scott boxes with N log-uniform in [5e4, 2e6]
"""
from __future__ import annotations

import numpy as np

SEED = 20261203


def edges_linear(lo, hi, nb):
    """nb+1 equally spaced edges computed once in FP64: lo + i*(hi-lo)/nb."""
    step = (hi - lo) / nb
    return np.array([lo + i * step for i in range(nb + 1)], dtype=np.float64)


def logspace(start, end, nedge):
    """The reference's logspace helper (source/utils.cu:16-27)."""
    a, b = np.log10(start), np.log10(end)
    step = (b - a) / (nedge - 1)
    return np.array([10.0 ** (a + i * step) for i in range(nedge)], dtype=np.float64)


def uniform_box(n, L, seed, float32=False):
    rng = np.random.default_rng(seed)
    p = np.ones((n, 4), dtype=np.float64)
    xyz = rng.uniform(0.0, L, size=(n, 3))
    if float32:
        xyz = xyz.astype(np.float32).astype(np.float64)
    p[:, :3] = xyz
    return p


def neyman_scott_box(n_parents, n_children, sigma, L, seed):
    """Parents uniform in [0,L)^3, children with isotropic Gaussian offsets,
    wrapped periodically, rounded to float32 (then widened to float64)."""
    rng = np.random.default_rng(seed)
    parents = rng.uniform(0.0, L, size=(n_parents, 3))
    off = rng.normal(0.0, sigma, size=(n_parents, n_children, 3))
    xyz = np.mod(parents[:, None, :] + off, L).reshape(-1, 3)
    xyz = xyz.astype(np.float32)
    # float32 rounding can land exactly on L; the reference's clamp of
    # cell == numCells (cudaNcorr.cu:52-56) handles it, so keep it.
    p = np.ones((xyz.shape[0], 4), dtype=np.float64)
    p[:, :3] = xyz.astype(np.float64)
    return p


def _cap_xyz(rng, n, rmin=1000.0, rmax=2000.0, ra_max=60.0, dec_max=60.0):
    ra = np.deg2rad(rng.uniform(0.0, ra_max, n))
    sd = rng.uniform(0.0, np.sin(np.deg2rad(dec_max)), n)
    dec = np.arcsin(sd)
    r = np.cbrt(rmin ** 3 + rng.uniform(0.0, 1.0, n) * (rmax ** 3 - rmin ** 3))
    return np.stack([r * np.cos(dec) * np.cos(ra), r * np.cos(dec) * np.sin(ra), r * np.sin(dec)], -1)


def _in_cap(xyz, rmin=1000.0, rmax=2000.0, ra_max=60.0, dec_max=60.0):
    r = np.linalg.norm(xyz, axis=-1)
    ra = np.rad2deg(np.arctan2(xyz[:, 1], xyz[:, 0]))
    dec = np.rad2deg(np.arcsin(np.clip(xyz[:, 2] / np.maximum(r, 1e-300), -1, 1)))
    return (r >= rmin) & (r <= rmax) & (ra >= 0) & (ra <= ra_max) & (dec >= 0) & (dec <= dec_max)


def survey_cap(n_data, n_rand, seed, clustered=False, sigma=4.0, n_children=100):
    """Data and randoms in the cap/shell, Cartesian FP64.  Data weights are
    float32 lognormal(0, 0.3); random weights are 1."""
    rng = np.random.default_rng(seed)
    if clustered:
        n_par = max(1, n_data // n_children)
        par = _cap_xyz(rng, n_par)
        xyz = (par[:, None, :] + rng.normal(0.0, sigma, size=(n_par, n_children, 3))).reshape(-1, 3)
        bad = ~_in_cap(xyz)
        while bad.any():  # children outside the cap are redrawn around their parent
            idx = np.nonzero(bad)[0]
            xyz[idx] = par[idx // n_children] + rng.normal(0.0, sigma, size=(idx.size, 3))
            bad[idx] = ~_in_cap(xyz[idx])
        xyz = xyz[:n_data]
    else:
        xyz = _cap_xyz(rng, n_data)
    d = np.empty((xyz.shape[0], 4))
    d[:, :3] = xyz
    d[:, 3] = rng.lognormal(0.0, 0.3, size=xyz.shape[0]).astype(np.float32).astype(np.float64)
    r = np.ones((n_rand, 4))
    r[:, :3] = _cap_xyz(rng, n_rand)
    return d, r


def prep_cat_survey(*cats):
    """survey::prep_cat_survey (source/survey.cu:60-128): joint min/max, box =
    max - min + 1.0, ori = -min, every catalogue shifted by -min in place."""
    lo = np.min([c[:, :3].min(axis=0) for c in cats], axis=0)
    hi = np.max([c[:, :3].max(axis=0) for c in cats], axis=0)
    box = hi - lo + 1.0
    ori = -lo
    for c in cats:
        c[:, 0] -= lo[0]
        c[:, 1] -= lo[1]
        c[:, 2] -= lo[2]
    return box, ori


def batched_boxes(n_boxes=64, nmin=50_000, nmax=2_000_000, L=1000.0, sigma=5.0, n_children=100, seed=SEED + 3000):
    """Sizes N_b = round(exp(U(ln nmin, ln nmax))), each box a Neyman-Scott
    process with N_b / n_children parents.  Returns a list of catalogues."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_boxes))).astype(int)
    out = []
    for b, n in enumerate(sizes):
        npar = max(1, int(round(n / n_children)))
        out.append(neyman_scott_box(npar, n_children, sigma, L, seed + 17 * (b + 1)))
    return out


WORKLOADS = {
    "box_uniform": dict(L=250.0, edges=edges_linear(2.0, 20.0, 10)),
    "box_clustered": dict(L=1000.0, edges=edges_linear(0.0, 30.0, 15)),
    "survey_cap": dict(edges=edges_linear(0.0, 40.0, 20)),
    "batched_boxes": dict(L=1000.0, edges=edges_linear(0.0, 30.0, 15)),
    "survey_large": dict(edges=edges_linear(0.0, 50.0, 25)),
}
