"""Seeded synthetic structures for the five workloads of this path.

All coordinates are returned in bohr.  Constructions:

- ``octane``: all-anti n-octane C8H18 (C-C 1.53 A, C-H 1.09 A, tetrahedral
  angles) plus Gaussian jitter sigma = 0.02 A (seed 1001);
- ``water_cluster``: (H2O)_n, rigid monomers (O-H 0.9572 A, HOH 104.52 deg),
  O uniform in a cube at 1.0 g/cm^3 with rejection below O-O 2.5 A, Haar-random
  orientations (seed 1002);
- ``mixed_cluster``: monomers drawn from CH4, NH3, H2O, H2S, CH3Cl, C2H6 with
  probabilities .25/.15/.25/.10/.10/.15 in a cube at 0.9 g/cm^3, heavy atoms
  at least 2.8 A apart; monomers are drawn until the atom count reaches
  ``nat``, the last one is redrawn until the total is exactly ``nat`` (if no
  monomer fits, it is truncated to its heavy atom plus H) (seed 1003);
- ``random_molecule``: a random tree of H/C/N/O/F/S/Cl (fractions
  .50/.30/.06/.08/.02/.02/.02) grown from covalent radii sums +- 0.03 A and
  relaxed to a minimum non-bonded distance of 2.0 A (seeds 1004, 1005).
"""
from __future__ import annotations

import math

import numpy as np

from .params import AATOAU

AMU = {1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 9: 18.998, 16: 32.06, 17: 35.45}
NA = 6.02214076e23
# Pyykko single-bond covalent radii (Angstrom) used to grow random molecules
COV = {1: 0.32, 6: 0.75, 7: 0.71, 8: 0.63, 9: 0.64, 16: 1.03, 17: 0.99}
VALENCE = {1: 1, 6: 4, 7: 3, 8: 2, 9: 1, 16: 2, 17: 1}
TET = math.acos(-1.0 / 3.0)


def _rot(rng):
    """Haar-random rotation matrix."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def octane(seed: int = 1001):
    """n-octane, all-anti, with jitter.  Returns (numbers, xyz_bohr)."""
    rng = np.random.default_rng(seed)
    cc, ch = 1.53, 1.09
    half = TET / 2.0
    carbons = []
    for k in range(8):
        x = k * cc * math.sin(half)
        y = (cc * math.cos(half) / 2.0) * (1 if k % 2 == 0 else -1)
        carbons.append(np.array([x, y, 0.0]))
    nums, xyz = [], []
    for c in carbons:
        nums.append(6)
        xyz.append(c)
    for k, c in enumerate(carbons):
        nb = [carbons[j] for j in (k - 1, k + 1) if 0 <= j < 8]
        u = sum((c - n) / np.linalg.norm(c - n) for n in nb)
        u /= np.linalg.norm(u)
        if len(nb) == 2:
            # two H in the plane perpendicular to the C-C-C plane
            w = np.cross(nb[0] - c, nb[1] - c)
            w /= np.linalg.norm(w)
            for s in (1, -1):
                d = math.cos(half) * u + s * math.sin(half) * w
                nums.append(1)
                xyz.append(c + ch * d / np.linalg.norm(d))
        else:
            b = (c - nb[0]) / np.linalg.norm(c - nb[0])
            w = np.cross(b, np.array([0.0, 0.0, 1.0]))
            w /= np.linalg.norm(w)
            v = np.cross(b, w)
            for phi in (0.0, 2 * math.pi / 3, 4 * math.pi / 3):
                d = -math.cos(TET) * b + math.sin(TET) * (math.cos(phi) * v + math.sin(phi) * w)
                nums.append(1)
                xyz.append(c + ch * d / np.linalg.norm(d))
    xyz = np.array(xyz) + rng.normal(scale=0.02, size=(len(xyz), 3))
    return np.array(nums, dtype=np.int32), xyz * AATOAU


def _water_monomer():
    oh, ang = 0.9572, math.radians(104.52)
    return np.array([8, 1, 1]), np.array([[0.0, 0.0, 0.0],
                                           [oh * math.sin(ang / 2), 0.0, oh * math.cos(ang / 2)],
                                           [-oh * math.sin(ang / 2), 0.0, oh * math.cos(ang / 2)]])


def water_cluster(nmol: int = 512, seed: int = 1002, density: float = 1.0):
    rng = np.random.default_rng(seed)
    mass = nmol * 18.015 / NA
    side = (mass / density) ** (1.0 / 3.0) * 1e8  # cm -> Angstrom
    zs, mono = _water_monomer()
    centres = []
    while len(centres) < nmol:
        p = rng.uniform(0.0, side, 3)
        if centres:
            d = np.linalg.norm(np.array(centres) - p, axis=1)
            if d.min() < 2.5:
                continue
        centres.append(p)
    nums, xyz = [], []
    for p in centres:
        R = _rot(rng)
        nums.extend(zs)
        xyz.extend(p + mono @ R.T)
    return np.array(nums, dtype=np.int32), np.array(xyz) * AATOAU


def _monomers():
    s3 = 1.0 / math.sqrt(3.0)
    tet = np.array([[1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]]) * s3
    ch4 = ([6, 1, 1, 1, 1], np.vstack([[0, 0, 0], tet * 1.09]))
    nh3 = ([7, 1, 1, 1], np.vstack([[0, 0, 0], tet[:3] * 1.01]))
    h2o = ([8, 1, 1], _water_monomer()[1])
    h2s = ([16, 1, 1], np.array([[0, 0, 0], [1.336 * math.sin(0.806), 0, 1.336 * math.cos(0.806)],
                                 [-1.336 * math.sin(0.806), 0, 1.336 * math.cos(0.806)]]))
    ch3cl = ([6, 17, 1, 1, 1], np.vstack([[0, 0, 0], tet[0] * 1.78, tet[1:] * 1.09]))
    c2h6 = ([6, 6, 1, 1, 1, 1, 1, 1],
            np.vstack([[0, 0, 0], tet[0] * 1.53, tet[1:] * 1.09,
                       tet[0] * 1.53 - tet[1:] * 1.09]))
    return [ch4, nh3, h2o, h2s, ch3cl, c2h6]


def mixed_cluster(nat: int = 4000, seed: int = 1003, density: float = 0.9):
    rng = np.random.default_rng(seed)
    mons = _monomers()
    prob = np.array([0.25, 0.15, 0.25, 0.10, 0.10, 0.15])
    picks = []
    total = 0
    while total < nat:
        k = rng.choice(6, p=prob)
        if total + len(mons[k][0]) > nat:
            fits = [q for q in range(6) if total + len(mons[q][0]) == nat]
            if fits:
                k = fits[0]
            else:
                # truncate: heavy atom plus hydrogens up to the remaining count
                rest = nat - total
                zs, xyz = mons[k]
                picks.append((list(zs[:rest]), xyz[:rest]))
                total += rest
                break
        picks.append(mons[k])
        total += len(mons[k][0])
    mass = sum(sum(AMU[z] for z in zs) for zs, _ in picks) / NA
    side = (mass / density) ** (1.0 / 3.0) * 1e8
    centres = []
    nums, xyz = [], []
    for zs, x in picks:
        for _ in range(100000):
            p = rng.uniform(0.0, side, 3)
            if not centres or np.linalg.norm(np.array(centres) - p, axis=1).min() >= 2.8:
                break
        centres.append(p)
        R = _rot(rng)
        nums.extend(zs)
        xyz.extend(p + np.asarray(x) @ R.T)
    return np.array(nums, dtype=np.int32), np.array(xyz) * AATOAU


_EL = np.array([1, 6, 7, 8, 9, 16, 17])
_FR = np.array([0.50, 0.30, 0.06, 0.08, 0.02, 0.02, 0.02])


def random_molecule(nat: int, rng) -> tuple:
    """Random tree molecule of ``nat`` atoms (bohr).

    Heavy atoms are grown first as a branched chain, hydrogens are then
    attached to heavy atoms with free valence.  Each new atom sits at the
    covalent-radius sum (+- 0.03 A) from its parent, in the direction (of 64
    random trials) that keeps it furthest from the rest; trials that keep
    every non-bonded atom at >= 2.0 A (>= 1.4 A for atoms bonded to the same
    parent) are preferred.
    """
    zs = list(rng.choice(_EL, size=nat, p=_FR))
    heavy = [z for z in zs if z != 1]
    hyd = [z for z in zs if z == 1]
    if not heavy:
        heavy, hyd = [6], hyd[1:]
    zs = heavy + hyd
    pos = np.zeros((nat, 3))
    nbrs = [[] for _ in range(nat)]
    free = np.array([VALENCE[z] for z in zs])
    for k in range(1, nat):
        z = zs[k]
        cands = [a for a in range(k) if zs[a] != 1 and free[a] > 0]
        if not cands:
            cands = [a for a in range(k) if zs[a] != 1]
        if z != 1 and len(cands) > 1:
            # favour recent heavy atoms: chains with occasional branches
            w = np.exp(0.3 * (np.array(cands) - k))
            a = cands[rng.choice(len(cands), p=w / w.sum())]
        else:
            a = cands[rng.integers(len(cands))]
        d = COV[zs[a]] + COV[z] + rng.uniform(-0.03, 0.03)
        u = rng.normal(size=(64, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        trial = pos[a] + d * u
        others = np.array([q for q in range(k) if q != a], dtype=int)
        if len(others):
            r = np.linalg.norm(trial[:, None, :] - pos[others][None, :, :], axis=-1)
            thr = np.where(np.isin(others, nbrs[a]), 1.4, 2.0)
            best = int(np.argmax((r - thr[None, :]).min(axis=1)))
        else:
            best = 0
        pos[k] = trial[best]
        nbrs[a].append(k)
        nbrs[k].append(a)
        free[a] -= 1
        free[k] -= 1
    return np.array(zs, dtype=np.int32), pos * AATOAU


def _one_molecule(args):
    seed, nat_min, nat_max = args
    rng = np.random.default_rng(seed)
    nat = int(round(math.exp(rng.uniform(math.log(nat_min), math.log(nat_max)))))
    return random_molecule(nat, rng)


def molecule_batch(nmol: int = 2048, nat_min: int = 8, nat_max: int = 256, seed: int = 1004,
                   workers: int = 0):
    """``nmol`` random molecules with log-uniform atom counts in [nat_min, nat_max].

    Molecule k uses the k-th child of ``SeedSequence(seed)``, so the batch is
    the same whether it is built serially or with ``workers`` processes."""
    seeds = np.random.SeedSequence(seed).spawn(nmol)
    args = [(sq, nat_min, nat_max) for sq in seeds]
    if workers and workers > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=workers) as ex:
            return list(ex.map(_one_molecule, args, chunksize=16))
    return [_one_molecule(a) for a in args]
