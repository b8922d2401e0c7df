"""Cube-sphere meshes and the reference's text input files.

The reference reads gmsh 2.2 ASCII meshes in which every surface element is
listed twice (d3qd2_src/bem3_aw_qd2.c:531-828): first with the physical id of
the domain whose outward normal the node order gives, then with the other
domain's id and the node order reversed (n0, n3, n2, n1).  Physical id 99 is the
open lattice domain.  The shipped sphere_m1.msh follows this with the object's
listing first (normal pointing out of the sphere); the writer here does the same.

The cube-sphere is a cube with each face split into n x n equal squares, every
vertex projected radially onto the sphere.
"""
import math

import numpy as np


def cube_sphere(n, radius):
    """Nodes (N x 3) and quads (6 n^2 x 4, 0-based node ids) of a cube-sphere.
    Quads are ordered so that (p1-p0) x (p3-p0) points out of the sphere."""
    g = np.linspace(-1.0, 1.0, n + 1)
    # each face: fixed axis, sign, and the two in-plane axes ordered so that
    # e_u x e_v points along +sign*axis
    faces = []
    for axis in range(3):
        u, v = (axis + 1) % 3, (axis + 2) % 3
        faces.append((axis, +1.0, u, v))
        faces.append((axis, -1.0, v, u))
    index = {}
    nodes = []
    quads = []

    def node_id(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(nodes)
            nodes.append(p)
        return index[key]

    for axis, sgn, u, v in faces:
        ids = np.empty((n + 1, n + 1), dtype=np.int64)
        for i, a in enumerate(g):
            for j, b in enumerate(g):
                c = np.zeros(3)
                c[axis] = sgn
                c[u] = a
                c[v] = b
                ids[i, j] = node_id(c)
        for i in range(n):
            for j in range(n):
                quads.append([ids[i, j], ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]])
    P = np.array(nodes)
    P = radius * P / np.linalg.norm(P, axis=1, keepdims=True)
    Q = np.array(quads, dtype=np.int64)
    # check orientation: normal of each quad points away from the centre
    c = P[Q].mean(axis=1)
    w = np.cross(P[Q[:, 1]] - P[Q[:, 0]], P[Q[:, 3]] - P[Q[:, 0]])
    assert (np.einsum("ij,ij->i", w, c) > 0).all()
    return P, Q


def split_triangles(P, Q):
    """Each quad split along its shorter diagonal into two triangles with the
    quad's orientation."""
    tris = []
    for q in Q:
        d02 = np.linalg.norm(P[q[0]] - P[q[2]])
        d13 = np.linalg.norm(P[q[1]] - P[q[3]])
        if d02 <= d13:
            tris.append([q[0], q[1], q[2]])
            tris.append([q[0], q[2], q[3]])
        else:
            tris.append([q[0], q[1], q[3]])
            tris.append([q[1], q[2], q[3]])
    return np.array(tris, dtype=np.int64)


def write_msh(path, P, elems, inner_domain=1, outer_domain=99):
    """gmsh 2.2 ASCII with each element listed twice (object side first)."""
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(P))]
    for i, p in enumerate(P, start=1):
        lines.append("%d %.17g %.17g %.17g" % (i, p[0], p[1], p[2]))
    lines += ["$EndNodes", "$Elements", str(2 * len(elems))]
    eid = 1
    for k, e in enumerate(elems):
        e = [int(x) + 1 for x in e]
        etype = 3 if len(e) == 4 else 2
        rev = [e[0], e[3], e[2], e[1]] if etype == 3 else [e[0], e[2], e[1]]
        lines.append("%d %d 2 %d %d %s" % (eid, etype, inner_domain, k + 1, " ".join(map(str, e))))
        lines.append("%d %d 2 %d %d %s" % (eid + 1, etype, outer_domain, k + 1, " ".join(map(str, rev))))
        eid += 2
    lines.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_pw(path, rho0, c0, f, p=(1.0, 0.0), tv=(0.0, 0.0, 0.0), theta=0.0, psi=0.0):
    with open(path, "w") as fh:
        fh.write("## beam data for planewave (code aw_pw)\n# rho0[kg/m^3] c0[m/s] f[Hz]\n")
        fh.write("%.17g %.17g %.17g\n" % (rho0, c0, f))
        fh.write("# real(p)[Pa] imag(p)[Pa] tx[m] ty[m] tz[m] theta[rad] phi[rad]\n")
        fh.write("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (p[0], p[1], tv[0], tv[1], tv[2], theta, psi))


def write_periodicity(path, d1, d2, ps=1):
    with open(path, "w") as fh:
        fh.write("## periodicity datafile for bem3_aw_qd2\n# periodicity_setting d1_x d1_y d2_x d2_y\n")
        fh.write("%d %.17g %.17g %.17g %.17g\n" % (ps, d1[0], d1[1], d2[0], d2[1]))


def write_medium(path, media):
    with open(path, "w") as fh:
        fh.write("## medium data file for bem3_aw_qd2\n# number of medium\n%d\n# rho0[kg/m^3] c0[m/s]\n" % len(media))
        for rho, c in media:
            fh.write("%.17g %.17g\n" % (rho, c))


# ---------------------------------------------------------------------------
# The workloads of the benchmark configurations
# ---------------------------------------------------------------------------
CONFIGS = {
    # 0.5 mm sphere (1200 kg/m^3, 2400 m/s) in water on a 2.0 mm square lattice,
    # 1 MHz; normal incidence is rejected by the reference (kx = ky = 0), so a
    # tilt of 1e-6 rad at azimuth pi/4 is used.  6 x 8 x 8 quads.
    "sphere384": dict(n=8, radius=0.5e-3, d1=(2.0e-3, 0.0), d2=(0.0, 2.0e-3), f=1.0e6,
                      theta=1.0e-6, psi=math.pi / 4, host=(1000.0, 1500.0), sphere=(1200.0, 2400.0)),
    # same sphere on a triangular 1.8 mm lattice, theta 30 deg, phi 15 deg, 6 x 24 x 24 quads
    "sphere3456": dict(n=24, radius=0.5e-3, d1=(1.8e-3, 0.0), d2=(0.9e-3, 0.9e-3 * math.sqrt(3.0)), f=1.0e6,
                       theta=math.pi / 6, psi=math.pi / 12, host=(1000.0, 1500.0), sphere=(1200.0, 2400.0)),
    # the mixed-precision instance: 6 x 48 x 48 quads, ka = 6 in the host medium
    # (k = 12000 / m), square lattice with a period of 2.2 radii, incidence
    # tilted by 1e-6 rad at azimuth pi/4 as sphere384
    "sphere13824": dict(n=48, radius=0.5e-3, d1=(1.1e-3, 0.0), d2=(0.0, 1.1e-3), f=6.0 * 1500.0 / (2.0 * math.pi * 0.5e-3),
                        theta=1.0e-6, psi=math.pi / 4, host=(1000.0, 1500.0), sphere=(1200.0, 2400.0)),
    # the same physics on the 3456-quad mesh (a layout test that also fits the default layout)
    "sphere3456k6": dict(n=24, radius=0.5e-3, d1=(1.1e-3, 0.0), d2=(0.0, 1.1e-3), f=6.0 * 1500.0 / (2.0 * math.pi * 0.5e-3),
                         theta=1.0e-6, psi=math.pi / 4, host=(1000.0, 1500.0), sphere=(1200.0, 2400.0)),
    # small case for tests: 6 x 2 x 2 quads, same physics as sphere384
    "sphere24": dict(n=2, radius=0.5e-3, d1=(2.0e-3, 0.0), d2=(0.0, 2.0e-3), f=1.0e6,
                     theta=1.0e-6, psi=math.pi / 4, host=(1000.0, 1500.0), sphere=(1200.0, 2400.0)),
    "sphere96": dict(n=4, radius=0.5e-3, d1=(2.0e-3, 0.0), d2=(0.0, 2.0e-3), f=1.0e6,
                     theta=1.0e-6, psi=math.pi / 4, host=(1000.0, 1500.0), sphere=(1200.0, 2400.0)),
}


def write_config(name, directory, triangles=False, periodic=True, **over):
    """Write pw.txt, periodicity.txt, medium.txt and the mesh of a workload into
    directory; returns the four paths in read_aqd2 order.  periodic=False writes
    periodicity setting 0: the same sphere alone in free space.  Keyword
    arguments replace entries of the workload (n, radius, d1, d2, f, theta, psi,
    host, sphere) or give the plane wave's complex amplitude p = (re, im)."""
    import os
    c = dict(CONFIGS[name], **over)
    os.makedirs(directory, exist_ok=True)
    P, Q = cube_sphere(c["n"], c["radius"])
    elems = split_triangles(P, Q) if triangles else Q
    paths = [os.path.join(directory, x) for x in ("pw.txt", "periodicity.txt", "medium.txt", "mesh.msh")]
    write_pw(paths[0], c["host"][0], c["host"][1], c["f"], p=c.get("p", (1.0, 0.0)), theta=c["theta"], psi=c["psi"])
    write_periodicity(paths[1], c["d1"], c["d2"], ps=1 if periodic else 0)
    write_medium(paths[2], [c["sphere"]])
    write_msh(paths[3], P, elems)
    return paths
