"""Writes scene_<name>.npz: the arrays this package's XML loader builds from
some of the reference's scene files (scenes3D/*.xml), so that tests can step
those scenes where the reference is not present.  Each file holds the body
descriptions, sphere-tree nodes, composite parts and their transforms, the
initial state, the springs and the <collision> settings; no text of the
scene files is kept.  Usage: python make_scene_fixtures.py [reference root]"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from adaptivemerging_amd import xml_scene  # noqa: E402

SCENES = ("chariot", "stackOnSwing", "funnel", "seesaw", "tower", "venus")

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
for name in SCENES:
    path = os.path.join(ref, "scenes3D", name + ".xml")
    scene, coll = xml_scene.parse_scene(path, data_root=ref)
    b, nd, x, R, v, w = scene.arrays()
    subs, b2c = scene.sub_arrays()
    np.savez_compressed(os.path.join(HERE, f"scene_{name}.npz"), bodies=b, nodes=nd, x=x, R=R, v=v, w=w, subs=subs,
                        sub_b2c=b2c, springs=scene.spring_array(), names=np.array(scene.names),
                        collision=np.array(json.dumps(coll)))
    print(name, scene.n, "bodies,", len(scene.nodes), "nodes,", len(scene.subs), "parts,", len(scene.springs),
          "springs,", os.path.getsize(os.path.join(HERE, f"scene_{name}.npz")), "bytes")
