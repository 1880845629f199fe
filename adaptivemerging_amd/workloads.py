"""Seeded synthetic scenes for the measured configurations.

Each generator draws from numpy's PCG64 (np.random.default_rng(seed)) in a
fixed field order, so a seed reproduces a scene exactly.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import AM_GEOM_COMPOSITE, AM_GEOM_TREE, BODY_DTYPE
from .scene import Scene, random_rotation, rot_axis_angle


def jenga_tower(seed: int = 0, layers: int = 33, top: bool = True, size=(1.0, 0.25, 0.15), density: float = 1000.0,
                jitter: float = 1e-3, overlap: float = 0.0, friction: float = 0.6) -> Scene:
    """Jenga tower: `layers` layers of 3 blocks, alternating 90 degrees about y,
    plus one block on top, on a ground plane y = 0 (n = +y).  Block centres are
    jittered uniformly by +-jitter per axis.  `overlap` lowers every layer by
    that much per layer so neighbouring layers start interpenetrating."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.add_plane([0.0, 0.0, 0.0], [0.0, 1.0, 0.0], friction=friction, name="ground")
    L, W, H = size
    R90 = rot_axis_angle([0, 1, 0], math.pi / 2)
    n = 0
    for k in range(layers + (1 if top else 0)):
        y = H / 2 + k * (H - overlap)
        offsets = [0.0] if k == layers else [-W, 0.0, W]
        for o in offsets:
            j = rng.uniform(-jitter, jitter, size=3)
            if k % 2 == 0:
                x, R = [j[0], y + j[1], o + j[2]], np.eye(3)
            else:
                x, R = [o + j[0], y + j[1], j[2]], R90
            s.add_box([L, H, W], x, R, density=density, friction=friction, name=f"block{n}")
            n += 1
    return s


def box_pile(seed: int = 0, n: int = 20000, bin_size: float = 4.0, edge=(0.1, 0.5), density=(100.0, 10000.0),
             gap: float = 0.002, friction: float = 0.5, walls: bool = True, scene_id: int = 0,
             max_yaw: float = 0.2) -> Scene:
    """A pile of n boxes (edges U(edge), densities log-uniform over `density`,
    so mass ratios reach density[1]/density[0]) in a bin_size x bin_size bin
    (floor y = 0 plus four walls).  Boxes lie flat (shortest edge vertical),
    yawed by U(-max_yaw, max_yaw), and are shelf-packed in layers of similar
    height (tallest layers at the bottom) with `gap` clearance, so they settle
    under gravity within a few dozen steps."""
    rng = np.random.default_rng(seed)
    s = Scene()
    half = bin_size / 2
    s.add_plane([0, 0, 0], [0, 1, 0], friction=friction, scene=scene_id, name="floor")
    if walls:
        s.add_plane([-half, 0, 0], [1, 0, 0], friction=friction, scene=scene_id, name="wall-x")
        s.add_plane([half, 0, 0], [-1, 0, 0], friction=friction, scene=scene_id, name="wall+x")
        s.add_plane([0, 0, -half], [0, 0, 1], friction=friction, scene=scene_id, name="wall-z")
        s.add_plane([0, 0, half], [0, 0, -1], friction=friction, scene=scene_id, name="wall+z")
    raw = rng.uniform(edge[0], edge[1], size=(n, 3))
    dens = np.exp(rng.uniform(math.log(density[0]), math.log(density[1]), size=n))
    yaw = rng.uniform(-max_yaw, max_yaw, size=n)
    srt = np.sort(raw, axis=1)
    sizes = np.stack([srt[:, 2], srt[:, 0], srt[:, 1]], axis=1)  # long x, short y (vertical), mid z
    order = np.argsort(-sizes[:, 1], kind="stable")
    c, sn = np.abs(np.cos(yaw)), np.abs(np.sin(yaw))
    foot = np.stack([c * sizes[:, 0] + sn * sizes[:, 2], sn * sizes[:, 0] + c * sizes[:, 2]], axis=1)
    y0, q = 0.0, 0
    while q < n:
        row_z, layer_h = -half + gap, 0.0
        while q < n and row_z < half - gap:
            x, row_d = -half + gap, 0.0
            placed = False
            while q < n:
                i = order[q]
                fx, fz = foot[i]
                if x + fx > half - gap or row_z + fz > half - gap:
                    break
                cx, cz = x + fx / 2, row_z + fz / 2
                cy = y0 + sizes[i, 1] / 2 + gap
                s.add_box(sizes[i], [cx, cy, cz], rot_axis_angle([0, 1, 0], yaw[i]), density=dens[i],
                          friction=friction, scene=scene_id)
                x += fx + gap
                row_d = max(row_d, fz)
                layer_h = max(layer_h, sizes[i, 1])
                q += 1
                placed = True
            if not placed:
                break
            row_z += row_d + gap
        y0 += layer_h + gap
    return s


def domino_platforms(seed: int = 0, platforms: int = 10, rows_per_platform: int = 2, per_row: int = 100,
                     size=(0.05, 0.5, 0.25), spacing: float = 0.2, density: float = 1000.0, friction: float = 0.5,
                     amplitude: float = 0.02, frequency: float = 0.5, gap: float = 0.001) -> Scene:
    """Rows of dominoes (size: thickness along the row x height x width,
    spaced `spacing` apart along x) standing on kinematic platforms that
    oscillate vertically (amplitude, frequency in Hz, phase U[0, 2 pi) from
    the seed), over a ground plane; the shape of scenes3D/dominosPlatforms.xml
    with prescribed platform motion in place of its springs.  The first
    domino of every row is listed in scene.row_heads (for the push that
    starts the cascade, see domino_push)."""
    rng = np.random.default_rng(seed)
    phases = rng.uniform(0.0, 2.0 * math.pi, size=platforms)
    s = Scene()
    s.add_plane([0, 0, 0], [0, 1, 0], friction=friction, name="ground")
    T, H, W = size
    row_len = per_row * spacing
    row_gap = 1.5 * W
    plat_w = rows_per_platform * row_gap + W
    plat_h = 0.2
    y0 = 1.0
    omega = 2.0 * math.pi * frequency
    s.row_heads = []
    for k in range(platforms):
        zc = k * (plat_w + 1.0)
        pb = s.add_box([row_len + 1.0, plat_h, plat_w], [row_len / 2, y0, zc], density=density, friction=friction,
                       name=f"platform{k}")
        s.set_kinematic(pb, amplitude, omega, float(phases[k]), (0.0, 1.0, 0.0))
        top = y0 + amplitude * math.sin(float(phases[k])) + plat_h / 2
        for r in range(rows_per_platform):
            z = zc + (r - (rows_per_platform - 1) / 2) * row_gap
            for i in range(per_row):
                b = s.add_box([T, H, W], [0.5 + i * spacing, top + H / 2 + gap, z], density=density, friction=friction,
                              name=f"domino{k}_{r}_{i}")
                if i == 0:
                    s.row_heads.append(b)
    return s


def domino_push(scene: Scene, x_now, dt: float, speed: float = 5.0):
    """The push that starts every row: for each row head, (body, point, force)
    with force m * speed / dt along +x at its current centre (x_now: the
    bodies' positions), so one step of it adds `speed` to the head's
    velocity."""
    out = []
    for b in scene.row_heads:
        m = float(scene.bodies[b]["mass"])
        out.append((b, np.array(x_now[b], dtype=np.float64), np.array([m * speed / dt, 0.0, 0.0])))
    return out


def box_cluster(seed: int = 0, n: int = 64, spacing: float = 0.3, edge=(0.1, 0.5), with_plane: bool = True,
                spheres: int = 0) -> Scene:
    """Randomly oriented boxes on a loose grid, spaced so that many neighbours
    interpenetrate (every box-box and box-plane case of the narrow phase is
    exercised); optional single-sphere bodies mixed in."""
    rng = np.random.default_rng(seed)
    s = Scene()
    if with_plane:
        s.add_plane([0, 0, 0], [0, 1, 0])
    side = int(math.ceil(n ** (1 / 3)))
    k = 0
    for ix in range(side):
        for iy in range(side):
            for iz in range(side):
                if k >= n:
                    break
                c = np.array([ix * spacing, 0.1 + iy * spacing, iz * spacing]) + rng.uniform(-0.03, 0.03, 3)
                if k < spheres:
                    s.add_sphere(rng.uniform(0.05, 0.2), c, density=1000.0)
                else:
                    s.add_box(rng.uniform(edge[0], edge[1], 3), c, random_rotation(rng), density=1000.0)
                k += 1
    return s


def sphere_tree_body_shape(rng: np.random.Generator, depth: int = 3, branching: int = 8, leaf_mu=math.log(0.02),
                           leaf_sigma=0.3, cube=0.3):
    """A depth-`depth`, branching-`branching` sphere tree: leaf radii
    lognormal, leaf centres uniform in a cube, leaves grouped by recursive
    x/y/z median splits, internal spheres enclosing their children.  Returns
    (centres, radii, children, mass, inertia_diag) about the centre of mass."""
    n_leaf = branching ** depth
    leaves = rng.uniform(-cube / 2, cube / 2, size=(n_leaf, 3))
    lr = np.exp(rng.normal(leaf_mu, leaf_sigma, size=n_leaf))
    m_leaf = 1000.0 * 4.0 / 3.0 * math.pi * lr ** 3
    com = (leaves * m_leaf[:, None]).sum(0) / m_leaf.sum()
    leaves = leaves - com

    def split(idx, parts):
        # split idx into `parts` equal groups by alternating-axis median splits
        groups = [idx]
        axis = 0
        while len(groups) < parts:
            new = []
            for g in groups:
                order = g[np.argsort(leaves[g, axis], kind="stable")]
                h = len(order) // 2
                new += [order[:h], order[h:]]
            groups = new
            axis = (axis + 1) % 3
        return groups

    centres, radii, children = [], [], []

    def build(idx, level):
        node = len(centres)
        centres.append(None)
        radii.append(None)
        children.append([])
        if level == depth:
            leaf = int(idx[0])
            centres[node] = leaves[leaf].tolist()
            radii[node] = float(lr[leaf])
            return node
        kids = [build(g, level + 1) for g in split(idx, branching)]
        children[node] = kids
        cc = np.array([centres[k] for k in kids])
        rr = np.array([radii[k] for k in kids])
        c = cc.mean(0)
        centres[node] = c.tolist()
        radii[node] = float(np.max(np.linalg.norm(cc - c, axis=1) + rr))
        return node

    build(np.arange(n_leaf), 0)
    mass = float(m_leaf.sum())
    I = np.zeros(3)
    for p, r, m in zip(leaves, lr, m_leaf):
        I += 2.0 / 5.0 * m * r * r + m * (np.dot(p, p) - p * p)
    return centres, radii, children, mass, I


def sphere_tree_drop(seed: int = 0, n: int = 512, box=(2.0, 1.0, 2.0), heights=(1.0, 5.0), depth: int = 3,
                     branching: int = 8) -> Scene:
    """n sphere-tree bodies placed at random heights U(heights) in an open box
    of 5 static planes (floor + 4 walls), x/z extent box[0] x box[2], with no
    overlap check: the trees start interpenetrating, which gives detection
    tests many tree-tree leaf contacts at once (sphere_tree_pile is the
    contact-free drop)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    hx, hz = box[0] / 2, box[2] / 2
    s.add_plane([0, 0, 0], [0, 1, 0], friction=0.4)
    s.add_plane([-hx, 0, 0], [1, 0, 0], friction=0.4)
    s.add_plane([hx, 0, 0], [-1, 0, 0], friction=0.4)
    s.add_plane([0, 0, -hz], [0, 0, 1], friction=0.4)
    s.add_plane([0, 0, hz], [0, 0, -1], friction=0.4)
    for _ in range(n):
        centres, radii, children, mass, I = sphere_tree_body_shape(rng, depth, branching)
        x = [rng.uniform(-hx + 0.2, hx - 0.2), rng.uniform(*heights), rng.uniform(-hz + 0.2, hz - 0.2)]
        s.add_tree(centres, radii, children, x, random_rotation(rng), mass=mass, inertia=I, friction=0.4)
    return s


def sphere_tree_pile(seed: int = 0, n: int = 512, box=(2.0, 1.0, 2.0), bottom: float = 1.0, spacing: float = 0.4,
                     depth: int = 3, branching: int = 8, max_tries: int = 20) -> Scene:
    """The sphere-tree configuration: n tree bodies dropped into an open box
    of 5 static planes, spawned with no two leaf spheres overlapping, so the
    pile builds up from contact-free drops.  Bodies go to the slots of a
    lattice (`spacing` apart, lowest layer at `bottom`, filled layer by layer)
    with a random rotation and a jitter of +-3 cm, retried up to max_tries
    times while a leaf would overlap a leaf of a body already placed; a slot
    that stays blocked is skipped.  512 trees of 512 leaves fill up to about 19 m of
    the 2 x 2 m column this way (their leaves alone are about 13 m^3)."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    s = Scene()
    hx, hz = box[0] / 2, box[2] / 2
    s.add_plane([0, 0, 0], [0, 1, 0], friction=0.4)
    s.add_plane([-hx, 0, 0], [1, 0, 0], friction=0.4)
    s.add_plane([hx, 0, 0], [-1, 0, 0], friction=0.4)
    s.add_plane([0, 0, -hz], [0, 0, 1], friction=0.4)
    s.add_plane([0, 0, hz], [0, 0, -1], friction=0.4)
    # slots keep a body's leaves (within 0.3 m of its centre) off the walls
    nx = max(1, int((box[0] - 0.6) // spacing) + 1)
    nz = max(1, int((box[2] - 0.6) // spacing) + 1)
    xs = (np.arange(nx) - (nx - 1) / 2) * spacing
    zs = (np.arange(nz) - (nz - 1) / 2) * spacing
    cen = np.zeros((n, 3))
    rad = np.zeros(n)
    leaves = []  # per placed body: (leaf centres world, leaf radii, kd tree)
    slot = 0
    for q in range(n):
        centres, radii, children, mass, I = sphere_tree_body_shape(rng, depth, branching)
        leaf = [k for k in range(len(centres)) if not children[k]]
        lc = np.array([centres[k] for k in leaf])
        lr = np.array([radii[k] for k in leaf])
        rb = float(radii[0] + np.linalg.norm(centres[0]))
        while True:
            layer, rest = divmod(slot, nx * nz)
            slot += 1
            base = np.array([xs[rest % nx], bottom + layer * spacing, zs[rest // nx]])
            found = False
            for _ in range(max_tries):
                R = random_rotation(rng)
                x = base + rng.uniform(-0.03, 0.03, 3)
                w = lc @ R.T + x
                near = np.nonzero(np.linalg.norm(cen[:q] - x, axis=1) < rad[:q] + rb)[0]
                ok = True
                for j in near:
                    w2, lr2, tree = leaves[j]
                    for i, js in enumerate(tree.query_ball_point(w, lr + lr2.max())):
                        if js and np.any(np.linalg.norm(w2[js] - w[i], axis=1) < lr[i] + lr2[js]):
                            ok = False
                            break
                    if not ok:
                        break
                if ok:
                    found = True
                    break
            if found:
                break
        cen[q], rad[q] = x, rb
        leaves.append((w, lr, cKDTree(w)))
        s.add_tree(centres, radii, children, x, R, mass=mass, inertia=I, friction=0.4)
    return s


def batched_pile_scenes(seed: int = 0, n_scenes: int = 256, bodies=(50, 2000), box_frac: float = 0.7) -> list:
    """The scenes of batched_piles, each on its own: body counts log-uniform
    over `bodies`, each a random pile of boxes (edges U(0.1, 0.5)) and spheres
    (r U(0.05, 0.25)) above its own ground plane."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(bodies[0]), math.log(bodies[1]), size=n_scenes)).astype(int)
    out = []
    for sc in range(n_scenes):
        s = Scene()
        s.add_plane([0, 0, 0], [0, 1, 0], friction=0.5)
        nb = int(counts[sc])
        side = int(math.ceil(math.sqrt(nb / 4)))
        for k in range(nb):
            lx, lz, ly = k % side, (k // side) % side, k // (side * side)
            c = [lx * 0.55, 0.3 + ly * 0.55, lz * 0.55] + rng.uniform(-0.02, 0.02, 3)
            if rng.uniform() < box_frac:
                s.add_box(rng.uniform(0.1, 0.5, 3), c, random_rotation(rng), density=1000.0, friction=0.5)
            else:
                s.add_sphere(rng.uniform(0.05, 0.25), c, density=1000.0, friction=0.5)
        out.append(s)
    return out


def batched_piles(seed: int = 0, n_scenes: int = 256, bodies=(50, 2000), box_frac: float = 0.7) -> Scene:
    """n_scenes independent scenes (scene ids 0..n-1) of batched_pile_scenes
    in one context."""
    return batch(batched_pile_scenes(seed, n_scenes, bodies, box_frac))


def merge_stack(seed: int = 0, layers: int = 4, density: float = 500.0) -> Scene:
    """A short stack of boards and a lone cube on a ground plane: contacts
    that settle within a few steps, so body pairs merge (adaptive-merging
    scenes such as scenes3D/jenga.xml with merging on)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.add_plane([0, 0, 0], [0, 1, 0])
    for k in range(layers):
        jitter = rng.uniform(-1e-3, 1e-3, size=3)
        s.add_box([1.0, 0.2, 0.6], [0.02 * k + jitter[0], 0.1 + 0.2 * k, jitter[2]], density=density)
    s.add_box([0.3, 0.3, 0.3], [2.0, 0.15, 0.0], density=density)
    return s


def hung_platform(seed: int = 0, n_boxes: int = 16, k: float = 5000.0, d: float = 200.0) -> Scene:
    """A platform hung by four world springs at its corners with a few
    layers of boxes on it (one of them on a zero-length spring), a chain of
    spheres tied by body-body springs hanging from a world spring, over a
    ground plane (the spring scenes of scenes3D such as tower25platform.xml
    and dominosPlatforms.xml)."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.add_plane([0, -2.0, 0], [0, 1, 0])
    plat = s.add_box([3.0, 0.2, 3.0], [0.0, 0.0, 0.0], density=100.0)
    for cx in (1.5, -1.5):
        for cz in (1.5, -1.5):
            s.add_spring(plat, [cx, 0.1, cz], pW=[cx, 1.1, cz], k=k, d=d)
    side = max(1, int(math.ceil(math.sqrt(n_boxes / 2))))
    placed = 0
    layer = 0
    while placed < n_boxes:
        for i in range(side):
            for j in range(side):
                if placed == n_boxes:
                    break
                e = rng.uniform(0.25, 0.4)
                x = -0.9 + 1.8 * (i + 0.5) / side + rng.uniform(-0.02, 0.02)
                z = -0.9 + 1.8 * (j + 0.5) / side + rng.uniform(-0.02, 0.02)
                b = s.add_box([e, e, e], [x, 0.1 + 0.2 + 0.45 * layer + 0.01, z], density=rng.uniform(200.0, 800.0))
                if placed == 0:
                    # zero-length spring: its rest length is 0.5, so near its anchor it pushes the box down
                    s.add_spring(b, [0.0, 0.0, 0.0], k=0.02 * k, d=0.02 * d)
                placed += 1
        layer += 1
    prev = None
    for q in range(3):
        b = s.add_sphere(0.2, [2.6, 1.5 - 0.6 * q, 0.0], density=300.0)
        if prev is None:
            s.add_spring(b, [0.0, 0.2, 0.0], pW=[2.6, 2.2, 0.0], k=0.1 * k, d=0.05 * d)
        else:
            s.add_spring(b, [0.0, 0.2, 0.0], body2=prev, pB2=[0.0, -0.2, 0.0], k=0.1 * k, d=0.05 * d, ls=0.9)
        prev = b
    return s


def batch(scenes) -> Scene:
    """Independent scenes in one context (scene ids 0..n-1, bodies and sphere-
    tree nodes concatenated): bodies of different scenes never collide."""
    out = Scene()
    for k, sc in enumerate(scenes):
        node_off = len(out.nodes)
        sub_off = len(out.subs)
        for i, rec in enumerate(sc.bodies):
            r = np.array(rec, dtype=BODY_DTYPE).copy()
            r["scene"] = k
            if int(r["geom"]) == AM_GEOM_TREE:
                r["tree_first"] = int(r["tree_first"]) + node_off
            elif int(r["geom"]) == AM_GEOM_COMPOSITE:
                r["tree_first"] = int(r["tree_first"]) + sub_off
            out._append(r, sc.x[i], sc.R[i], sc.v[i], sc.w[i], f"s{k}_{sc.names[i]}")
        for rec, t in zip(sc.subs, sc.sub_b2c):
            r = np.array(rec, dtype=BODY_DTYPE).copy()
            r["scene"] = k
            if int(r["geom"]) == AM_GEOM_TREE:
                r["tree_first"] = int(r["tree_first"]) + node_off
            out.subs.append(r)
            out.sub_b2c.append(list(t))
        out.nodes.extend(sc.nodes)
    return out
