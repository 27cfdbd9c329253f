"""Native periodic neighbor builder against the numpy reference builder:
identical pair sets (neighbor index and image shift per centre), bit-exact
counts, strict r < rc, self-images in small triclinic cells."""
import numpy as np
import pytest

from mlip_hip import periodic_full_list
from mlip_hip import workloads as W


def pairs_native(s, rc):
    off, idx, sh = periodic_full_list(s.cell, s.positions, rc)
    out = []
    for i in range(s.natoms):
        d = s.positions[i] - s.positions[idx[off[i]:off[i + 1]]] + sh[off[i]:off[i + 1]]
        out.append(sorted(zip(idx[off[i]:off[i + 1]].tolist(), np.round(d, 9).tolist())))
    return off, out


def pairs_numpy(s, rc):
    nl = W.build_neighbor_list(s, rc, full=True)
    out = []
    for i in range(s.natoms):
        js = nl.idx[nl.off[i]:nl.off[i + 1]]
        d = nl.x[i] - nl.x[js]
        r = np.linalg.norm(d, axis=1)
        keep = r < rc
        out.append(sorted(zip(nl.owner[js][keep].tolist(), np.round(d[keep], 9).tolist())))
    return out


@pytest.mark.parametrize("case", ["fcc", "tiny_triclinic", "reference"])
def test_native_list_matches_numpy(case):
    if case == "fcc":
        s, rc = W.fcc(3, 4.0, 0.08, 4, type_fraction=0.5, type_seed=4), 8.0
    elif case == "tiny_triclinic":
        s, rc = W.random_cell_structure(9, 0.06, np.random.default_rng(6)), 8.0
    else:
        from common import ref
        s, rc = W.read_lammps_data(ref("structure_00064_triclinic.lammps")), 8.0
    off, a = pairs_native(s, rc)
    b = pairs_numpy(s, rc)
    assert [len(v) for v in a] == [len(v) for v in b]
    for va, vb in zip(a, b):
        assert [j for j, _ in va] == [j for j, _ in vb]
        np.testing.assert_allclose(np.array([d for _, d in va]), np.array([d for _, d in vb]), atol=1e-8)
