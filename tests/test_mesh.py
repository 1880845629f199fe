"""Pins the mesh mass properties (adaptivemerging_amd/mesh.py, Mirtich's
volume integrals as src/tools/moments/VolInt.java computes them) against
closed forms: boxes (m = rho abc, I = m (b^2 + c^2) / 12 ...), the regular
tetrahedron (I = m a^2 / 20 about its centre), rotated and translated copies
(centre moves with the mesh, I -> R I R^T), linearity in the density, and an
inside-out mesh (mass negated with a warning); plus the OBJ reader and the
createMesh path through the scene XML."""
import math
import os

import numpy as np
import pytest

from adaptivemerging_amd import mesh, xml_scene
from adaptivemerging_amd.scene import random_rotation


def box_mesh(a, b, c):
    V = [[x * a, y * b, z * c] for x in (0, 1) for y in (0, 1) for z in (0, 1)]
    i = lambda x, y, z: x * 4 + y * 2 + z
    F = []

    def quad(p, q, r, s):
        F.extend([[p, q, r], [p, r, s]])

    quad(i(0, 0, 0), i(0, 0, 1), i(0, 1, 1), i(0, 1, 0))
    quad(i(1, 0, 0), i(1, 1, 0), i(1, 1, 1), i(1, 0, 1))
    quad(i(0, 0, 0), i(1, 0, 0), i(1, 0, 1), i(0, 0, 1))
    quad(i(0, 1, 0), i(0, 1, 1), i(1, 1, 1), i(1, 1, 0))
    quad(i(0, 0, 0), i(0, 1, 0), i(1, 1, 0), i(1, 0, 0))
    quad(i(0, 0, 1), i(1, 0, 1), i(1, 1, 1), i(0, 1, 1))
    return V, F


def tetra_mesh(a):
    V = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) * (a / (2 * math.sqrt(2)))
    F = [[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]]
    return V.tolist(), F


def test_box_closed_form():
    a, b, c, rho = 0.7, 1.3, 0.4, 250.0
    m, I, com = mesh.mass_properties(*box_mesh(a, b, c), rho)
    M = rho * a * b * c
    assert m == pytest.approx(M, rel=1e-14)
    np.testing.assert_allclose(com, [a / 2, b / 2, c / 2], rtol=1e-14)
    want = np.diag([M * (b * b + c * c) / 12, M * (a * a + c * c) / 12, M * (a * a + b * b) / 12])
    np.testing.assert_allclose(np.reshape(I, (3, 3)), want, rtol=1e-12, atol=1e-12 * M)


def test_regular_tetrahedron_closed_form():
    a, rho = 1.7, 3.0
    m, I, com = mesh.mass_properties(*tetra_mesh(a), rho)
    M = rho * a ** 3 / (6 * math.sqrt(2))
    assert m == pytest.approx(M, rel=1e-13)
    np.testing.assert_allclose(com, 0.0, atol=1e-14)
    np.testing.assert_allclose(np.reshape(I, (3, 3)), np.eye(3) * M * a * a / 20, rtol=1e-12, atol=1e-13)


def test_rotation_translation_and_density():
    rng = np.random.default_rng(0)
    V, F = box_mesh(0.5, 0.9, 1.4)
    R = random_rotation(rng)
    t = rng.normal(size=3)
    V2 = (np.array(V) @ R.T + t).tolist()
    m1, I1, c1 = mesh.mass_properties(V, F, 2.0)
    m2, I2, c2 = mesh.mass_properties(V2, F, 2.0)
    assert m2 == pytest.approx(m1, rel=1e-13)
    np.testing.assert_allclose(c2, R @ np.array(c1) + t, atol=1e-13)
    np.testing.assert_allclose(np.reshape(I2, (3, 3)), R @ np.reshape(I1, (3, 3)) @ R.T, atol=1e-12)
    m3, I3, _ = mesh.mass_properties(V2, F, 6.0)
    assert m3 == pytest.approx(3 * m2, rel=1e-14)
    np.testing.assert_allclose(I3, 3 * np.array(I2), rtol=1e-13, atol=1e-14)


def test_inside_out_and_degenerate_faces(tmp_path):
    V, F = box_mesh(1, 1, 1)
    path = tmp_path / "inv.obj"
    lines = [f"v {x} {y} {z}" for x, y, z in V]
    lines += ["f " + " ".join(f"{k + 1}/{k + 1}/1" for k in reversed(f)) for f in F]
    lines += ["f 1 1 2"]   # zero area: skipped
    path.write_text("\n".join(lines) + "\n")
    with pytest.warns(UserWarning, match="inside out"):
        mb = mesh.mesh_body(str(path), 2.0, 10.0)
    assert mb["mass"] == pytest.approx(80.0, rel=1e-14)
    np.testing.assert_allclose(mb["com"], [1, 1, 1], rtol=1e-14)
    # vertices moved to the centre of mass: the box spans -1..1
    bb = np.reshape(mb["bbox"], (8, 3))
    np.testing.assert_allclose(bb[0], [-1, -1, -1]) ; np.testing.assert_allclose(bb[7], [1, 1, 1])


def write_mesh_files(tmp_path, a=0.6, b=0.3, c=0.4):
    V, F = box_mesh(a, b, c)
    (tmp_path / "block.obj").write_text("\n".join([f"v {x} {y} {z}" for x, y, z in V] +
                                                    ["f " + " ".join(str(k + 1) for k in f) for f in F]) + "\n")
    # root sphere around the block, eight leaves at its corners (in OBJ coordinates)
    leaves = [[x, y, z] for x in (0.1, a - 0.1) for y in (0.1, b - 0.1) for z in (0.1, c - 0.1)]
    lines = ["9", f"{a / 2} {b / 2} {c / 2} {math.sqrt(a * a + b * b + c * c) / 2}"]
    lines += [f"{x} {y} {z} 0.1" for x, y, z in leaves]
    lines += ["1 8 " + " ".join(str(k) for k in range(2, 10))]
    (tmp_path / "block.sph").write_text("\n".join(lines) + "\n")


def test_xml_mesh_body(tmp_path):
    write_mesh_files(tmp_path)
    text = f"""<root><body type="mesh" name="m" scale="2.0" density="100" obj="block.obj" st="block.sph">
        <x>0 1 0</x><R>0 0 1 0.3</R></body></root>"""
    scene, _ = xml_scene.parse_scene(text, data_root=str(tmp_path))
    b = scene.bodies[0]
    assert int(b["geom"]) == 1 and int(b["tree_count"]) == 9
    M = 100 * (1.2 * 0.6 * 0.8)
    assert float(b["mass"]) == pytest.approx(M, rel=1e-13)
    # the sphere tree is scaled and centred on the centre of mass
    root = scene.nodes[int(b["tree_first"])]
    np.testing.assert_allclose(root["c"], 0.0, atol=1e-15)
    assert float(root["r"]) == pytest.approx(math.sqrt(1.2 ** 2 + 0.6 ** 2 + 0.8 ** 2) / 2 * 1.0, rel=1e-14)
    np.testing.assert_allclose(np.reshape(b["jinv0"], (3, 3)) @ np.reshape(b["inertia0"], (3, 3)), np.eye(3),
                               atol=1e-12)


REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "data")), reason="the reference's data folder is not present")
def test_reference_meshes_have_valid_mass_properties():
    for obj in ("scaledtorso10.obj", "scaledhand10.obj", "scaledround-octa10.obj"):
        mb = mesh.mesh_body(os.path.join(REF, "data", obj), 1.0, 1.0)
        I = np.reshape(mb["inertia"], (3, 3))
        assert mb["mass"] > 0
        np.testing.assert_allclose(I, I.T, rtol=0, atol=0)
        ev = np.linalg.eigvalsh(I)
        assert ev.min() > 0 and ev[0] + ev[1] >= ev[2] * (1 - 1e-12)   # principal moments obey the triangle rule
