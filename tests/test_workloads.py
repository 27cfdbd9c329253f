"""Synthetic workload generators (host only)."""
import numpy as np

from mlip_hip import workloads as W


def test_random_structures_deterministic_and_well_formed():
    a = W.random_structures(64, 8, 512, seed=7)
    b = W.random_structures(64, 8, 512, seed=7)
    c = W.random_structures(64, 8, 512, seed=8)
    sizes = np.array([s.natoms for s in a])
    assert sizes.min() >= 8 and sizes.max() <= 512
    assert any(s1.natoms != s2.natoms for s1, s2 in zip(a, c))
    for s, t in zip(a, b):
        assert np.array_equal(s.positions, t.positions) and np.array_equal(s.cell, t.cell)
        assert np.array_equal(s.types, t.types)
    for s in a[:12]:
        vol = abs(np.linalg.det(s.cell))
        assert abs(s.natoms / vol - 0.06) < 1e-9
        assert set(np.unique(s.types)) <= {1, 2}
        # minimum distance over periodic images is at least rmin = 1.8
        nl = W.build_neighbor_list(s, 1.8, full=True)
        assert nl.numneigh.sum() == 0


def test_random_structure_sizes_only():
    s = W.random_structures(4096, 8, 512, seed=6)
    n = np.array([x.natoms for x in s])
    # log-uniform in [8, 512]: median near sqrt(8 * 512) = 64
    assert 40 < np.median(n) < 100
