"""get_lattice_points_cutoff (cutoff.f90:97-121): the oracle's restatement
pinned by what the geometry fixes, and the library's host implementation
(tbx_get_lattice_points) equal to it entry by entry, order included.

For lattice vectors a_k the reference repeats each a_k ceil(rthr / d_k)
times either way, d_k = V / |a_i x a_j| the distance between the lattice
planes spanned by the other two vectors; the points are then exactly
{n1 a1 + n2 a2 + n3 a3 : |n_k| <= rep_k}, the origin first, closed under
inversion."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import lib as L, structure


def _cells():
    rng = np.random.default_rng(4)
    yield np.diag([10.0, 12.0, 9.0]), 25.0
    yield np.array([[8.0, 0.0, 0.0], [4.0, 6.9282, 0.0], [0.0, 0.0, 13.0]]), 18.4  # hexagonal
    yield np.diag([7.0, 7.0, 7.0]) + rng.normal(scale=1.5, size=(3, 3)), 20.0  # triclinic
    yield np.diag([30.0, 30.0, 30.0]), 5.0


@pytest.mark.parametrize("k", range(4))
def test_oracle_lattice_points_closed_form(k):
    lat, rthr = list(_cells())[k]
    pts = O.get_lattice_points([True] * 3, lat, rthr)
    vol = abs(np.linalg.det(lat))
    rep = [int(np.ceil(rthr / (vol / np.linalg.norm(np.cross(lat[(i + 1) % 3], lat[(i + 2) % 3])))))
           for i in range(3)]
    assert len(pts) == np.prod([2 * r + 1 for r in rep])
    assert np.all(pts[0] == 0.0)
    brute = {tuple(np.round(np.array(n) @ lat, 9))
             for n in itertools.product(*[range(-r, r + 1) for r in rep])}
    got = {tuple(np.round(p, 9)) for p in pts}
    assert got == brute
    assert {tuple(np.round(-p, 9)) for p in pts} == got
    # every point within rthr of the origin is in the set (what the cutoff needs)
    assert all(tuple(np.round(p, 9)) in got for p in brute if np.linalg.norm(p) <= rthr)


@pytest.mark.parametrize("k", range(4))
def test_library_lattice_points_equal_oracle(k):
    lat, rthr = list(_cells())[k]
    ref = O.get_lattice_points([True, True, True], lat, rthr)
    got = structure.get_lattice_points([True, True, True], lat, rthr)
    assert got.shape == ref.shape and np.array_equal(got, ref)


def test_lattice_points_molecule_and_errors():
    assert np.array_equal(structure.get_lattice_points([False] * 3, np.eye(3), 25.0),
                          np.zeros((1, 3)))
    with pytest.raises(L.TbxError):  # a flat cell has no finite repetition count
        structure.get_lattice_points([True] * 3, np.diag([10.0, 10.0, 0.0]), 25.0)
