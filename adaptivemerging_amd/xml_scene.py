"""Scene XML ingestion: the reference's scene files drive this build unchanged.

Follows XMLParser (src/mergingBodies3D/XMLParser.java):
- parse (:41-66): the <system> nodes, then <collision>, then every <body>
  child of <root> in document order;
- parseCollision (:91-135): the solver settings common to all bodies;
- createBox (:360-396), createPlane (:404-413), createSphere (:420-447);
- setCommonAttributes (:542-626): x, R (axis-angle), v, omega, restitution,
  friction, pinned and the tags that do not touch the contact step.

Springs (:596-619, Spring.java) load with the attachment geometry their
constructors compute from the pose the body has when the <spring> tag is
read.  Mesh bodies (createMesh :449-538) load their mass properties from
the OBJ file and their sphere tree from the .sph file; composite bodies
(createComposite :180-353) their box and sphere parts.  Spinner vectors are
recorded in scene.spinners: only the app's keyboard applies them, by adding
them to a body's angular velocity.
Sphere-tree geometry for hand-built bodies comes from .sph files through
scene.load_sph (the format of :500-535).
"""
from __future__ import annotations

import os
import warnings
import xml.etree.ElementTree as ET

import numpy as np

from .scene import Scene, rot_axis_angle


class SceneFormatError(ValueError):
    """A scene file that this build cannot load faithfully."""


# <collision> attribute -> (CollisionProcessor field, parser); restitution and
# friction also switch their override on (XMLParser.java:112-122)
_COLLISION_ATTRS = {
    "iterations": ("iterations", int),
    "feedbackStiffness": ("feedbackStiffness", float),
    "restitution": ("restitution", float),
    "friction": ("friction", float),
    "enablePostStabilization": ("enablePostStabilization", None),
    "enableCompliance": ("enableCompliance", None),
}

# CollisionProcessor field -> am_params field
PARAM_NAMES = {
    "iterations": "iterations",
    "feedbackStiffness": "feedback_stiffness",
    "restitution": "restitution",
    "restitutionOverride": "restitution_override",
    "friction": "friction",
    "frictionOverride": "friction_override",
    "enablePostStabilization": "enable_post_stabilization",
    "enableCompliance": "enable_compliance",
}


def _java_bool(s: str) -> bool:
    """Boolean.parseBoolean: true iff the text is "true", ignoring case."""
    return s.strip().lower() == "true"


def _t3d(s: str) -> list[float]:
    v = s.split()
    if len(v) < 3:
        raise SceneFormatError(f"expected three numbers, got {s!r}")
    return [float(v[0]), float(v[1]), float(v[2])]


def _children(node, tag: str):
    return [c for c in node if c.tag.lower() == tag]


class _Common:
    """The per-body settings of setCommonAttributes (XMLParser.java:542-626);
    a later tag overrides an earlier one, as there."""

    def __init__(self, node, name: str, drop_springs: bool):
        self.x = None
        self.R = None
        self.v = None
        self.w = None
        self.friction = 0.8      # RigidBody.friction default (RigidBody.java:142)
        self.restitution = 0.0   # RigidBody.restitution default (RigidBody.java:145)
        self.pinned = False
        self.springs = []        # (attributes, pose (x, R) when the tag was read)
        self.spinner = None
        for c in node:
            tag = c.tag.lower()
            text = (c.text or "").strip()
            if tag == "x":
                self.x = _t3d(text)
            elif tag == "r":
                a = [float(t) for t in text.split()]
                if len(a) < 4:
                    raise SceneFormatError(f"body {name}: <R> needs an axis and an angle")
                n = np.linalg.norm(a[:3])
                self.R = np.eye(3) if n == 0.0 else rot_axis_angle(a[:3], a[3])
            elif tag == "v":
                self.v = _t3d(text)
            elif tag == "omega":
                self.w = _t3d(text)
            elif tag == "restitution":
                self.restitution = float(text.split()[0])
            elif tag == "friction":
                self.friction = float(text.split()[0])
            elif tag == "pinned":
                self.pinned = _java_bool(text)
            elif tag == "spinner":
                # RigidBody.spinner / canSpin: only the app's Q and W keys use
                # it, adding it to omega (LCPApp3D.java:918-935)
                self.spinner = _t3d(text)
            elif tag == "spring":
                if not drop_springs:
                    pose = (list(self.x) if self.x is not None else [0.0, 0.0, 0.0],
                            np.array(self.R) if self.R is not None else np.eye(3))
                    self.springs.append((dict(c.attrib), pose))
            # temporarilyPinned is ignored by the reference itself; magnetic
            # only matters once a magnet is switched on from the UI; factorypart
            # and col are for the factory feature and drawing
        self.pos = self.x if self.x is not None else [0.0, 0.0, 0.0]


def _add_springs(scene: Scene, body: int, com: "_Common", name: str, posed_planes: set) -> None:
    """The springs of one body, as setCommonAttributes builds them
    (XMLParser.java:596-619): pW gives a world spring, pB2 with body2 a
    body-body spring (body2 named among the bodies read before), otherwise a
    spring to the point's initial world position; k, d and ls override the
    defaults 100, 10 and 1 (Spring.java:27-31)."""
    for attrs, pose in com.springs:
        kw = {}
        for key, default in (("k", 100.0), ("d", 10.0), ("ls", 1.0)):
            kw[key] = float(attrs[key]) if attrs.get(key, "") != "" else default
        pB = _t3d(attrs.get("pB", "0 0 0"))
        if "pW" in attrs:
            scene.add_spring(body, pB, pW=_t3d(attrs["pW"]), pose1=pose, **kw)
        elif "pB2" in attrs and "body2" in attrs:
            target = attrs["body2"]
            others = [i for i, nm in enumerate(scene.names[:body]) if nm == target]
            if not others:
                # the reference prints "[ParseSpring] Something is wrong with the spring." and drops it
                warnings.warn(f"body {name}: spring to unknown body {target!r} dropped, as the reference does")
                continue
            if others[0] in posed_planes:
                raise SceneFormatError(f"body {name}: spring to plane {target!r}, which has a pose tag")
            scene.add_spring(body, pB, body2=others[0], pB2=_t3d(attrs["pB2"]), pose1=pose, **kw)
        else:
            scene.add_spring(body, pB, pose1=pose, **kw)


def _find_file(rel: str, search, name: str) -> str:
    """A data file named in the scene (relative to the directory the
    reference runs from): tried against each search directory in turn."""
    if not rel:
        raise SceneFormatError(f"body {name}: missing file name")
    for d in search:
        path = os.path.join(d, rel)
        if os.path.isfile(path):
            return path
    raise SceneFormatError(f"body {name}: file {rel!r} not found under {list(search)}")


def parse_scene(source: str, *, ignore_springs: bool = False, scene_index: int = 0, data_root=None):
    """Parses a scene (a file path or the XML text).  Returns (scene,
    collision), where collision maps CollisionProcessor field names to the
    values the file sets.  Mesh files (obj=, st=) are looked up under
    data_root, the current directory, and the scene file's directory and
    its parent (the reference's scenes name them relative to its root)."""
    text = source
    search = [d for d in (data_root, os.getcwd()) if d]
    if not source.lstrip().startswith("<"):
        with open(source) as f:
            text = f.read()
        here = os.path.dirname(os.path.abspath(source))
        search += [here, os.path.dirname(here)]
    try:
        root = ET.fromstring(text)
    except ET.ParseError as e:
        raise SceneFormatError(f"not an XML scene: {e}") from e
    if root.tag.lower() != "root":
        raise SceneFormatError(f"expected a <root> element, got <{root.tag}>")

    collision: dict = {}
    for node in _children(root, "collision"):
        for attr, (field, conv) in _COLLISION_ATTRS.items():
            if attr not in node.attrib:
                continue
            raw = node.attrib[attr]
            collision[field] = _java_bool(raw) if conv is None else conv(raw)
            if attr == "restitution":
                collision["restitutionOverride"] = True
            elif attr == "friction":
                collision["frictionOverride"] = True

    scene = Scene()
    # planes keep x = 0 and R = I here; the reference moves them with <x>/<R>,
    # which only matters to a spring attached to one
    posed_planes: set = set()
    for node in _children(root, "body"):
        kind = node.attrib.get("type", "").lower()
        name = node.attrib.get("name", "")
        com = _Common(node, name, ignore_springs)
        if kind == "box":
            s = np.array(_t3d(node.attrib["dim"]))
            density = float(node.attrib.get("density", 1.0))
            if "scale" in node.attrib:
                s = s * float(node.attrib["scale"])
            b = scene.add_box(s, com.pos, R=com.R, v=com.v, w=com.w, density=density, friction=com.friction,
                              restitution=com.restitution, pinned=com.pinned, scene=scene_index, name=name)
            _add_springs(scene, b, com, name, posed_planes)
            if com.spinner is not None:
                scene.spinners[b] = com.spinner
        elif kind == "plane":
            # PlaneRigidBody is always pinned (PlaneRigidBody.java:55-62)
            if com.springs:
                raise SceneFormatError(f"body {name}: springs on a plane are not supported")
            b = scene.add_plane(_t3d(node.attrib["p"]), _t3d(node.attrib["n"]), friction=com.friction,
                                restitution=com.restitution, scene=scene_index, name=name)
            if com.x is not None or com.R is not None:
                posed_planes.add(b)
        elif kind == "sphere":
            if com.pinned:
                raise SceneFormatError(f"body {name}: pinned spheres are not supported")
            density = float(node.attrib.get("density", 1.0))
            b = scene.add_sphere(float(node.attrib["r"]), com.pos, R=com.R, v=com.v, w=com.w, density=density,
                                 friction=com.friction, restitution=com.restitution, scene=scene_index, name=name)
            _add_springs(scene, b, com, name, posed_planes)
            if com.spinner is not None:
                scene.spinners[b] = com.spinner
        elif kind == "mesh":
            if "scale" not in node.attrib:
                raise SceneFormatError(f"body {name}: a mesh needs a scale")
            obj = _find_file(node.attrib.get("obj", ""), search, name)
            sph = _find_file(node.attrib.get("st", ""), search, name)
            b = scene.add_mesh(obj, sph, com.pos, scale=float(node.attrib["scale"]),
                               density=float(node.attrib.get("density", 1.0)), R=com.R, v=com.v, w=com.w,
                               friction=com.friction, restitution=com.restitution, pinned=com.pinned,
                               scene=scene_index, name=name)
            _add_springs(scene, b, com, name, posed_planes)
            if com.spinner is not None:
                scene.spinners[b] = com.spinner
        elif kind == "composite":
            parts = []
            for child in node:
                if child.tag.lower() != "body":
                    continue
                ckind = child.attrib.get("type", "").lower()
                cname = child.attrib.get("name", "")
                if ckind in ("plane", "mesh"):
                    # createComposite prints an error and leaves them out (XMLParser.java:218-226)
                    warnings.warn(f"composite {name}: {ckind} part {cname!r} left out, as the reference does")
                    continue
                part = _Common(child, cname, True)
                if part.springs or any(c.tag.lower() == "spring" for c in child):
                    warnings.warn(f"composite {name}: springs on part {cname!r} act on no integrated body; dropped")
                pd = dict(x=part.pos, R=part.R, density=float(child.attrib.get("density", 1.0)))
                if ckind == "box":
                    dim = np.array(_t3d(child.attrib["dim"]))
                    if "scale" in child.attrib:
                        dim = dim * float(child.attrib["scale"])
                    pd.update(kind="box", size=dim)
                elif ckind == "sphere":
                    pd.update(kind="sphere", r=float(child.attrib["r"]))
                else:
                    continue
                parts.append(pd)
            if not parts:
                raise SceneFormatError(f"body {name}: composite without box or sphere parts")
            b = scene.add_composite(parts, com.pos, R=com.R, v=com.v, w=com.w, friction=com.friction,
                                    restitution=com.restitution, pinned=com.pinned, scene=scene_index, name=name)
            _add_springs(scene, b, com, name, posed_planes)
            if com.spinner is not None:
                scene.spinners[b] = com.spinner
        else:
            raise SceneFormatError(f"body {name}: unknown type {kind!r}")
    return scene, collision


def apply_collision_settings(collision: dict, params) -> None:
    """Copies parsed <collision> settings onto an am_params structure."""
    for field, value in collision.items():
        setattr(params, PARAM_NAMES[field], int(value) if isinstance(value, bool) else value)


class XMLParser:
    """XMLParser.parse(system, filename) (XMLParser.java:41-66) for this build:
    builds the scene, then a RigidBodySystem on it with the file's collision
    settings."""

    def parse(self, filename: str, device: int = 0, ignore_springs: bool = False):
        from .system import RigidBodySystem

        scene, collision = parse_scene(filename, ignore_springs=ignore_springs)
        system = RigidBodySystem(scene, device)
        for field, value in collision.items():
            setattr(system.collision, field, value)
        system.name = os.path.basename(filename)
        return system
