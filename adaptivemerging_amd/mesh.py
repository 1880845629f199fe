"""Mesh bodies: the mass properties and bounding box that
XMLParser.createMesh (src/mergingBodies3D/XMLParser.java:449-538) computes
from a Wavefront OBJ file; the body's collision geometry is its .sph sphere
tree (scene.load_sph).

The mass properties are Mirtich's volume integrals (B. Mirtich, "Fast and
Accurate Computation of Polyhedral Mass Properties", journal of graphics
tools 1(2), 1996), the method of src/tools/moments/VolInt.java: by the
divergence theorem each volume integral becomes a sum of face integrals,
each face integral an integral over the face's projection onto the
coordinate plane where it is largest, and each projection integral a sum
over the projected edges.  As PolygonSoup.getPolyhedron does, each face
contributes the triangle of its first three vertices, and faces whose
normal has zero length are skipped.
"""
from __future__ import annotations

import math
import warnings

MIN_DOUBLE = 5e-324   # Double.MIN_VALUE (the smallest positive double)


def read_obj(path: str):
    """PolygonSoup (src/tools/moments/PolygonSoup.java:41-125): 'v x y z'
    lines give vertices, 'f a/b/c ...' lines faces (1-based vertex indices,
    texture and normal indices ignored).  Returns (vertices, faces) with
    0-based face indices."""
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("v "):
                t = line[2:].split()
                verts.append([float(t[0]), float(t[1]), float(t[2])])
            elif line.startswith("f "):
                faces.append([int(tok.split("/")[0]) - 1 for tok in line[2:].split()])
    return verts, faces


def _face_plane(p0, p1, p2):
    """Unit normal from the first edge crossed with the second, and the
    plane offset w = -n.p0; None for a degenerate triangle
    (Polyhedron.addFaceHelper, Polyhedron.java:216-238)."""
    dx1, dy1, dz1 = p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]
    dx2, dy2, dz2 = p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]
    nx = dy1 * dz2 - dy2 * dz1
    ny = dz1 * dx2 - dz2 * dx1
    nz = dx1 * dy2 - dx2 * dy1
    ln = math.sqrt(nx * nx + ny * ny + nz * nz)
    if ln == 0:
        return None
    n = [nx / ln, ny / ln, nz / ln]
    return n, -n[0] * p0[0] - n[1] * p0[1] - n[2] * p0[2]


def _projection_integrals(tri, A, B):
    """Integrals of 1, a, b, a^2, ab, b^2, a^3, a^2 b, a b^2, b^3 over the
    face's projection onto the (A, B) plane, as sums over its edges
    (Mirtich 1996, eq. 20-21 in Horner form)."""
    P1 = Pa = Pb = Paa = Pab = Pbb = Paaa = Paab = Pabb = Pbbb = 0.0
    for i in range(3):
        a0, b0 = tri[i][A], tri[i][B]
        a1, b1 = tri[(i + 1) % 3][A], tri[(i + 1) % 3][B]
        da, db = a1 - a0, b1 - b0
        a0_2 = a0 * a0
        a0_3 = a0_2 * a0
        a0_4 = a0_3 * a0
        b0_2 = b0 * b0
        b0_3 = b0_2 * b0
        b0_4 = b0_3 * b0
        a1_2 = a1 * a1
        a1_3 = a1_2 * a1
        b1_2 = b1 * b1
        b1_3 = b1_2 * b1
        C1 = a1 + a0
        Ca = a1 * C1 + a0_2
        Caa = a1 * Ca + a0_3
        Caaa = a1 * Caa + a0_4
        Cb = b1 * (b1 + b0) + b0_2
        Cbb = b1 * Cb + b0_3
        Cbbb = b1 * Cbb + b0_4
        Cab = 3 * a1_2 + 2 * a1 * a0 + a0_2
        Kab = a1_2 + 2 * a1 * a0 + 3 * a0_2
        Caab = a0 * Cab + 4 * a1_3
        Kaab = a1 * Kab + 4 * a0_3
        Cabb = 4 * b1_3 + 3 * b1_2 * b0 + 2 * b1 * b0_2 + b0_3
        Kabb = b1_3 + 2 * b1_2 * b0 + 3 * b1 * b0_2 + 4 * b0_3
        P1 += db * C1
        Pa += db * Ca
        Paa += db * Caa
        Paaa += db * Caaa
        Pb += da * Cb
        Pbb += da * Cbb
        Pbbb += da * Cbbb
        Pab += db * (b1 * Cab + b0 * Kab)
        Paab += db * (b1 * Caab + b0 * Kaab)
        Pabb += da * (a1 * Cabb + a0 * Kabb)
    return (P1 / 2.0, Pa / 6.0, Pb / -6.0, Paa / 12.0, Pab / 24.0, Pbb / -12.0, Paaa / 20.0, Paab / 60.0,
            Pabb / -60.0, Pbbb / -20.0)


def volume_integrals(verts, faces):
    """The ten volume integrals T0 (volume), T1 = (x, y, z), T2 = (x^2, y^2,
    z^2), TP = (xy, yz, zx) over the closed surface (Mirtich 1996, eq. 13
    and the face integrals of eq. 18-19)."""
    T0 = 0.0
    T1, T2, TP = [0.0] * 3, [0.0] * 3, [0.0] * 3
    for face in faces:
        tri = [verts[face[0]], verts[face[1]], verts[face[2]]]
        plane = _face_plane(*tri)
        if plane is None:
            continue
        n, w = plane
        nx, ny, nz = abs(n[0]), abs(n[1]), abs(n[2])
        if nx > ny and nx > nz:
            C = 0
        else:
            C = 1 if ny > nz else 2
        A = (C + 1) % 3
        B = (A + 1) % 3
        P1, Pa, Pb, Paa, Pab, Pbb, Paaa, Paab, Pabb, Pbbb = _projection_integrals(tri, A, B)
        nA, nB, nC = n[A], n[B], n[C]
        k1 = 1 / nC
        k2 = k1 * k1
        k3 = k2 * k1
        k4 = k3 * k1
        Fa = k1 * Pa
        Fb = k1 * Pb
        Fc = -k2 * (nA * Pa + nB * Pb + w * P1)
        Faa = k1 * Paa
        Fbb = k1 * Pbb
        Fcc = k3 * (nA * nA * Paa + 2 * nA * nB * Pab + nB * nB * Pbb + w * (2 * (nA * Pa + nB * Pb) + w * P1))
        Faaa = k1 * Paaa
        Fbbb = k1 * Pbbb
        Fccc = -k4 * (nA * nA * nA * Paaa + 3 * (nA * nA) * nB * Paab + 3 * nA * (nB * nB) * Pabb
                      + nB * nB * nB * Pbbb + 3 * w * (nA * nA * Paa + 2 * nA * nB * Pab + nB * nB * Pbb)
                      + w * w * (3 * (nA * Pa + nB * Pb) + w * P1))
        Faab = k1 * Paab
        Fbbc = -k2 * (nA * Pabb + nB * Pbbb + w * Pbb)
        Fcca = k3 * (nA * nA * Paaa + 2 * nA * nB * Paab + nB * nB * Pabb + w * (2 * (nA * Paa + nB * Pab) + w * Pa))
        T0 += n[0] * (Fa if A == 0 else (Fb if B == 0 else Fc))
        T1[A] += nA * Faa
        T1[B] += nB * Fbb
        T1[C] += nC * Fcc
        T2[A] += nA * Faaa
        T2[B] += nB * Fbbb
        T2[C] += nC * Fccc
        TP[A] += nA * Faab
        TP[B] += nB * Fbbc
        TP[C] += nC * Fcca
    T1 = [t / 2 for t in T1]
    T2 = [t / 3 for t in T2]
    TP = [t / 2 for t in TP]
    return T0, T1, T2, TP


def mass_properties(verts, faces, density: float):
    """(mass, inertia about the centre of mass (3x3 row-major list), centre
    of mass) of a uniform-density closed triangle mesh (Mirtich 1996, sec. 2;
    VolInt.computeMassProperties)."""
    T0, T1, T2, TP = volume_integrals(verts, faces)
    mass = density * T0
    r = [T1[0] / T0, T1[1] / T0, T1[2] / T0]
    J = [[0.0] * 3 for _ in range(3)]
    J[0][0] = density * (T2[1] + T2[2])
    J[1][1] = density * (T2[2] + T2[0])
    J[2][2] = density * (T2[0] + T2[1])
    J[0][1] = J[1][0] = -density * TP[0]
    J[1][2] = J[2][1] = -density * TP[1]
    J[2][0] = J[0][2] = -density * TP[2]
    # parallel-axis shift to the centre of mass
    J[0][0] -= mass * (r[1] * r[1] + r[2] * r[2])
    J[1][1] -= mass * (r[2] * r[2] + r[0] * r[0])
    J[2][2] -= mass * (r[0] * r[0] + r[1] * r[1])
    J[0][1] = J[1][0] = J[1][0] + mass * r[0] * r[1]
    J[1][2] = J[2][1] = J[2][1] + mass * r[1] * r[2]
    J[2][0] = J[0][2] = J[0][2] + mass * r[2] * r[0]
    return mass, [J[i][j] for i in range(3) for j in range(3)], r


def mesh_body(obj_path: str, scale: float, density: float):
    """What createMesh derives from the OBJ file: vertices scaled, mass
    properties (a negative mass, an inside-out mesh, is negated with a
    warning), vertices moved to the centre of mass, and the eight corners of
    the vertices' bounding box.  The box's upper corner starts from
    Double.MIN_VALUE, as in the reference, so an axis on which every vertex
    is negative gets 4.9e-324 there.  Returns dict(mass, inertia, com, bbox)."""
    verts, faces = read_obj(obj_path)
    verts = [[p[0] * scale, p[1] * scale, p[2] * scale] for p in verts]
    mass, inertia, com = mass_properties(verts, faces, density)
    if mass < 0:
        mass = -mass
        warnings.warn(f"mesh {obj_path} appears to be inside out")
    verts = [[p[0] - com[0], p[1] - com[1], p[2] - com[2]] for p in verts]
    ll = [1.7976931348623157e308] * 3
    ur = [MIN_DOUBLE] * 3
    for p in verts:
        for a in range(3):
            ll[a] = min(p[a], ll[a])
            ur[a] = max(p[a], ur[a])
    bbox = [[lx, ly, lz] for lx in (ll[0], ur[0]) for ly in (ll[1], ur[1]) for lz in (ll[2], ur[2])]
    return dict(mass=mass, inertia=inertia, com=com, bbox=bbox, n_vertices=len(verts), n_faces=len(faces))
