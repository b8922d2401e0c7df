"""Cube-sphere generator and the reference's text input writers
(bem3_aw_qd2_amd/mesh.py)."""
import collections

import numpy as np
import pytest

from bem3_aw_qd2_amd import mesh


@pytest.mark.parametrize("n", [1, 2, 8])
def test_cube_sphere_is_closed_and_outward(n):
    P, Q = mesh.cube_sphere(n, 0.5e-3)
    assert len(Q) == 6 * n * n
    assert len(P) == 6 * n * n + 2          # Euler: V - E + F = 2 with E = 2F
    assert np.allclose(np.linalg.norm(P, axis=1), 0.5e-3)
    edges = collections.Counter()
    for q in Q:
        for a, b in zip(q, np.roll(q, -1)):
            edges[(a, b)] += 1
    # every directed edge appears once and its reverse once: closed, consistently oriented
    assert all(v == 1 for v in edges.values())
    assert all((b, a) in edges for (a, b) in edges)
    c = P[Q].mean(axis=1)
    nrm = np.cross(P[Q[:, 1]] - P[Q[:, 0]], P[Q[:, 3]] - P[Q[:, 0]])
    assert (np.einsum("ij,ij->i", nrm, c) > 0).all()


def test_split_triangles_keep_orientation():
    P, Q = mesh.cube_sphere(3, 1.0)
    T = mesh.split_triangles(P, Q)
    assert len(T) == 2 * len(Q)
    c = P[T].mean(axis=1)
    nrm = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    assert (np.einsum("ij,ij->i", nrm, c) > 0).all()


def test_msh_lists_each_element_twice(tmp_path):
    P, Q = mesh.cube_sphere(2, 1.0)
    path = str(tmp_path / "m.msh")
    mesh.write_msh(path, P, Q)
    lines = open(path).read().split("\n")
    i = lines.index("$Elements")
    ne = int(lines[i + 1])
    assert ne == 2 * len(Q)
    for k in range(len(Q)):
        a = lines[i + 2 + 2 * k].split()
        b = lines[i + 3 + 2 * k].split()
        assert a[3] == "1" and b[3] == "99"          # object side first, then the open domain
        na, nb = a[5:], b[5:]
        assert nb == [na[0], na[3], na[2], na[1]]   # reversed node order on the second listing


def test_workload_files(tmp_path):
    paths = mesh.write_config("sphere384", str(tmp_path))
    pw = open(paths[0]).read().split("\n")
    assert pw[2].split() == ["1000", "1500", "1000000"]
    per = open(paths[1]).read().split("\n")[2].split()
    assert per[0] == "1" and float(per[1]) == 2e-3 and float(per[4]) == 2e-3
    med = open(paths[2]).read().split("\n")
    assert med[2] == "1" and med[4].split() == ["1200", "2400"]
