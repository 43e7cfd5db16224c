"""Deterministic synthetic scenes for the four workloads of this pipeline.

Every scene is a dict of numpy arrays in the layout of nphys_body_desc (include/nphys.h):
status, shape, dims(n,3), pos(n,3), rot(n,4 as i,j,k,w), linvel, angvel, density, margin,
linear_prediction, friction, restitution. Static colliders come first.

Randomness is SplitMix64 with a per-scene seed; a float is the top 24 bits of one output
scaled to [0, 1). Draw k of body b is element b*stride + k of one stream, so scenes can be
generated vectorised and reproduced in any language.

The scenes follow the reference's example scenes (examples3d/balls3.rs:17-94,
examples3d/boxes3.rs:17-94) scaled to the stated sizes:
  * balls(): 1,000 spheres r=0.5 on a 10^3 grid (spacing 1.1, jitter +-0.05) over a static
    cuboid ground of half extents (50, 0.5, 50); friction 0.5.
  * boxes(): 8,000 unit cuboids on a 20^3 grid (spacing 1.05), tilted by Rz*Ry*Rx with each
    angle in +-2 degrees; friction 0.6.
  * pile(): 262,144 bodies, even index a ball with r in [0.2, 0.6], odd a cuboid with half
    extents in [0.2, 0.5] and a uniform random orientation, on a 34 x 227 x 34 lattice filling
    x,z in [-30, 30], y in [45, 445], above an open five-wall box (60 x 40 x 60).
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

DEFAULT_MARGIN = 0.01
DEFAULT_LINEAR_PREDICTION = 0.001


def splitmix_uniform(seed, count):
    """count floats in [0,1): output k is SplitMix64 state seed + (k+1)*golden, top 24 bits."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)


def _empty(n):
    return {
        "status": np.zeros(n, np.int32),
        "shape": np.zeros(n, np.int32),
        "dims": np.zeros((n, 3), np.float32),
        "pos": np.zeros((n, 3), np.float32),
        "rot": np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1)),
        "linvel": np.zeros((n, 3), np.float32),
        "angvel": np.zeros((n, 3), np.float32),
        "density": np.ones(n, np.float32),
        "margin": np.full(n, DEFAULT_MARGIN, np.float32),
        "linear_prediction": np.full(n, DEFAULT_LINEAR_PREDICTION, np.float32),
        "friction": np.full(n, 0.5, np.float32),
        "restitution": np.zeros(n, np.float32),
    }


def concat(*scenes):
    out = {}
    for k in scenes[0]:
        out[k] = np.concatenate([s[k] for s in scenes], axis=0)
    return out


def with_sleep_threshold(scene, threshold):
    """Copy of `scene` whose bodies may sleep below `threshold` (a scalar or one value per body;
    negative: the body never sleeps). RigidBodyDesc's default is 0.01 (src/object/body.rs:72-74)."""
    out = dict(scene)
    n = scene["status"].shape[0]
    out["sleep_threshold"] = np.broadcast_to(np.asarray(threshold, np.float32), (n,)).copy()
    return out


def statics(half_extents, centres, friction=0.5):
    n = len(centres)
    s = _empty(n)
    s["status"][:] = 1
    s["shape"][:] = 1
    s["dims"][:] = np.asarray(half_extents, np.float32)
    s["pos"][:] = np.asarray(centres, np.float32)
    s["friction"][:] = friction
    return s


def ground(friction=0.5, half_extents=(50.0, 0.5, 50.0), centre=(0.0, -0.5, 0.0)):
    return statics([half_extents], [centre], friction)


def quat_axis(axis, angle):
    """(i, j, k, w) rotation about a unit coordinate axis, float64 arrays."""
    n = np.asarray(angle).shape[0]
    q = np.zeros((n, 4))
    q[:, axis] = np.sin(angle * 0.5)
    q[:, 3] = np.cos(angle * 0.5)
    return q


def quat_mul(a, b):
    ai, aj, ak, aw = a.T
    bi, bj, bk, bw = b.T
    return np.stack(
        [
            aw * bi + ai * bw + aj * bk - ak * bj,
            aw * bj - ai * bk + aj * bw + ak * bi,
            aw * bk + ai * bj - aj * bi + ak * bw,
            aw * bw - ai * bi - aj * bj - ak * bk,
        ],
        axis=1,
    )


def balls(seed=1, nx=10, ny=10, nz=10, radius=0.5, spacing=1.1, jitter=0.05, friction=0.5):
    n = nx * ny * nz
    u = splitmix_uniform(seed, 3 * n).reshape(n, 3)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c = np.stack([(i - (nx - 1) / 2) * spacing, 1.0 + j * spacing, (k - (nz - 1) / 2) * spacing], 1)
    c = c + (u * 2.0 - 1.0) * jitter
    s = _empty(n)
    s["shape"][:] = 0
    s["dims"][:, 0] = radius
    s["pos"][:] = c
    s["friction"][:] = friction
    return concat(ground(friction), s)


def boxes(seed=2, nx=20, ny=20, nz=20, half=0.5, spacing=1.05, max_tilt_deg=2.0, friction=0.6):
    n = nx * ny * nz
    u = splitmix_uniform(seed, 3 * n).reshape(n, 3)
    ang = np.deg2rad((u * 2.0 - 1.0) * max_tilt_deg)
    q = quat_mul(quat_axis(2, ang[:, 2]), quat_mul(quat_axis(1, ang[:, 1]), quat_axis(0, ang[:, 0])))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c = np.stack([(i - (nx - 1) / 2) * spacing, (half + 0.05) + j * spacing, (k - (nz - 1) / 2) * spacing], 1)
    s = _empty(n)
    s["shape"][:] = 1
    s["dims"][:] = half
    s["pos"][:] = c
    s["rot"][:] = q
    s["friction"][:] = friction
    return concat(ground(friction), s)


def pile_walls(friction=0.5, inner=30.0, height=20.0):
    he = [(inner + 1.0, 0.5, inner + 1.0), (0.5, height, inner + 1.0), (0.5, height, inner + 1.0),
          (inner, height, 0.5), (inner, height, 0.5)]
    cs = [(0.0, -0.5, 0.0), (inner + 0.5, height, 0.0), (-(inner + 0.5), height, 0.0),
          (0.0, height, inner + 0.5), (0.0, height, -(inner + 0.5))]
    return statics(he, cs, friction)


def pile(seed=3, n=262144, lattice=(34, 227, 34), x_range=(-30.0, 30.0), y_range=(45.0, 445.0),
         r_range=(0.2, 0.6), he_range=(0.2, 0.5), friction=0.5):
    lx, ly, lz = lattice
    assert n <= lx * ly * lz
    cx = (x_range[1] - x_range[0]) / lx
    cy = (y_range[1] - y_range[0]) / ly
    cz = cx
    b = np.arange(n)
    j = b // (lx * lz)
    i = (b // lz) % lx
    k = b % lz
    u = splitmix_uniform(seed, 9 * n).reshape(n, 9)
    is_ball = (b % 2) == 0
    s = _empty(n)
    s["shape"][:] = np.where(is_ball, 0, 1)
    r = r_range[0] + (r_range[1] - r_range[0]) * u[:, 0]
    he = he_range[0] + (he_range[1] - he_range[0]) * u[:, 0:3]
    dims = np.where(is_ball[:, None], np.stack([r, np.zeros(n), np.zeros(n)], 1), he)
    s["dims"][:] = dims
    # Shoemake uniform rotation from u[:, 3:6]
    u1, u2, u3 = u[:, 3], 2 * np.pi * u[:, 4], 2 * np.pi * u[:, 5]
    q = np.stack([np.sqrt(1 - u1) * np.sin(u2), np.sqrt(1 - u1) * np.cos(u2),
                  np.sqrt(u1) * np.sin(u3), np.sqrt(u1) * np.cos(u3)], 1)
    q[is_ball] = (0, 0, 0, 1)
    s["rot"][:] = q
    bound = np.where(is_ball, r, np.sqrt((he * he).sum(1)))
    cell = min(cx, cy)
    jit = np.maximum(0.0, cell / 2 - bound)[:, None] * (u[:, 6:9] * 2 - 1)
    c = np.stack([x_range[0] + (i + 0.5) * cx, y_range[0] + (j + 0.5) * cy, x_range[0] + (k + 0.5) * cz], 1)
    s["pos"][:] = c + jit
    s["friction"][:] = friction
    return concat(pile_walls(friction), s)


def small_pile(seed=7, n=512, lattice=(8, 16, 8), extent=8.0, y0=2.0, friction=0.5):
    """A scaled-down pile (same generator) that the CPU checker steps in seconds."""
    lx, ly, lz = lattice
    cell = 2.0 * extent / lx
    s = pile(seed=seed, n=n, lattice=lattice, x_range=(-extent, extent),
             y_range=(y0, y0 + cell * ly), friction=friction)
    # replace the 60 m box with one sized to the lattice
    w = pile_walls(friction, inner=extent, height=extent)
    for k in s:
        s[k] = np.concatenate([w[k], s[k][5:]], axis=0)
    return s


def ccd_room(num=4, rad=0.1, half_width=3.0, thickness=0.2, height=1.5, velocity=(-100.0, -10.0, 0.0),
             shape=1, seed=None):
    """The reference's CCD example (examples3d/ccd3.rs:18-124): a closed room of six static walls,
    cuboids of half extents (half_width, thickness, half_width) that are all colliders of one static
    body (static_body 0), and num^3 CCD-enabled bodies of half extent / radius `rad` (cuboids for
    shape 1, balls for shape 0; with a seed, alternating balls and cuboids with a random
    orientation) on a lattice of spacing 2 (rad + margin) at `height`, all moving at `velocity`
    (100 m/s crosses the room in 0.06 s, 3.6 steps). Density 1, no sleep thresholds (the example
    keeps RigidBodyDesc's default 0.01)."""
    hw = float(half_width)
    half = (hw, thickness, hw)
    zq = quat_axis(2, np.array([np.pi / 2]))[0]
    xq = quat_axis(0, np.array([np.pi / 2]))[0]
    ident = np.array([0.0, 0.0, 0.0, 1.0])
    walls = statics([half] * 6, [(0, -hw, 0), (0, hw, 0), (hw, 0, 0), (-hw, 0, 0), (0, 0, hw), (0, 0, -hw)])
    walls["rot"][:] = np.stack([ident, ident, zq, zq, xq, xq]).astype(np.float32)
    walls["static_body"] = np.zeros(6, np.int32)
    n = num ** 3
    b = _empty(n)
    shift = (rad + DEFAULT_MARGIN) * 2.0
    cx = shift * num / 2.0
    i, j, k = np.meshgrid(np.arange(num), np.arange(num), np.arange(num), indexing="ij")
    b["pos"][:] = np.stack([i.ravel() * shift - cx, j.ravel() * shift + shift / 2.0 + height, k.ravel() * shift - cx], 1)
    b["shape"][:] = shape
    b["dims"][:] = (rad, rad, rad) if shape == 1 else (rad, 0.0, 0.0)
    if seed is not None:
        v = splitmix_uniform(seed, 3 * n).reshape(n, 3)
        is_ball = (np.arange(n) % 2) == 0
        b["shape"][:] = np.where(is_ball, 0, 1)
        b["dims"][:] = np.where(is_ball[:, None], np.array([rad, 0.0, 0.0]), np.array([rad, rad, rad]))
        u1, u2, u3 = v[:, 0], 2 * np.pi * v[:, 1], 2 * np.pi * v[:, 2]
        q = np.stack([np.sqrt(1 - u1) * np.sin(u2), np.sqrt(1 - u1) * np.cos(u2),
                      np.sqrt(u1) * np.sin(u3), np.sqrt(u1) * np.cos(u3)], 1)
        q[is_ball] = (0, 0, 0, 1)
        b["rot"][:] = q
    b["linvel"][:] = np.asarray(velocity, np.float32)
    b["ccd_enabled"] = np.ones(n, np.int32)
    b["static_body"] = np.full(n, -1, np.int32)
    walls["ccd_enabled"] = np.zeros(6, np.int32)
    return concat(walls, b)


def collision_groups_example(num=8, layers=4, rad=0.1):
    """The reference's collision groups example (examples3d/collision_groups3.rs:18-132): one
    static body (static_body 0) with a main floor of half extents (3, 0.2, 3) whose top is at y = 0
    (default groups: it collides with everything), a green floor (1, 0.1, 1) at y = 1 in group 0
    whitelisting group 0, and a blue floor at y = 2 in group 1 whitelisting group 1; above them
    `layers` layers of num x num cuboids of half extent `rad` on a lattice of spacing 2 (rad +
    margin) from y = 2.5, the even layers green and the odd ones blue. Green and blue bodies never
    touch: the green layers fall through the blue floor onto the green one."""
    from nphysics_amd.abi import collision_groups
    green = collision_groups(membership=[0], whitelist=[0])
    blue = collision_groups(membership=[1], whitelist=[1])
    floors = statics([(3.0, 0.2, 3.0), (1.0, 0.1, 1.0), (1.0, 0.1, 1.0)], [(0.0, -0.2, 0.0), (0.0, 1.0, 0.0),
                                                                          (0.0, 2.0, 0.0)])
    floors["static_body"] = np.zeros(3, np.int32)
    floors["collision_groups"] = np.array([collision_groups(), green, blue], np.uint32)
    n = layers * num * num
    b = _empty(n)
    b["shape"][:] = 1
    b["dims"][:] = (rad, rad, rad)
    shift = (rad + DEFAULT_MARGIN) * 2.0
    c = shift * num / 2.0
    k, i, j = np.meshgrid(np.arange(layers), np.arange(num), np.arange(num), indexing="ij")
    b["pos"][:] = np.stack([i.ravel() * shift - c, k.ravel() * shift + 2.5, j.ravel() * shift - c], 1)
    b["static_body"] = np.full(n, -1, np.int32)
    b["collision_groups"] = np.where((k.ravel() % 2 == 0)[:, None], np.array(green, np.uint32),
                                     np.array(blue, np.uint32)).astype(np.uint32)
    return concat(floors, b)


def num_static(scene):
    return int((scene["status"] == 1).sum())


def batched(seed=4, worlds=4096, median=64, sigma=1.0, n_min=8, n_max=2048, cell=1.8, half_width=9.0,
            ground_half=(10.0, 0.5, 10.0), r_range=(0.2, 0.6), he_range=(0.2, 0.5), friction=0.5):
    """Independent worlds in one batch: world w has n_w = clip(round(exp(N(ln median, sigma^2))),
    n_min, n_max) bodies (even index in the world a ball, odd a cuboid, the pile's size ranges,
    uniform random orientation) on a jittered lattice of `cell` spacing over its own static
    ground of half extents ground_half; every body of world w has scene index w. Static grounds
    come first (one per world), then the bodies world by world."""
    u = splitmix_uniform(seed, 2 * worlds).reshape(worlds, 2)
    z = np.sqrt(-2.0 * np.log(np.maximum(u[:, 0], 1e-300))) * np.cos(2.0 * np.pi * u[:, 1])
    nw = np.clip(np.rint(np.exp(np.log(median) + sigma * z)), n_min, n_max).astype(np.int64)
    per_layer_side = int(2 * half_width / cell)
    per_layer = per_layer_side * per_layer_side
    total = int(nw.sum())
    world = np.repeat(np.arange(worlds), nw)
    first = np.concatenate([[0], np.cumsum(nw)[:-1]])
    k = np.arange(total) - first[world]          # index inside the world
    layer, rem = k // per_layer, k % per_layer
    ix, iz = rem // per_layer_side, rem % per_layer_side
    v = splitmix_uniform(seed + 1, 9 * total).reshape(total, 9)
    is_ball = (k % 2) == 0
    s = _empty(total)
    s["shape"][:] = np.where(is_ball, 0, 1)
    r = r_range[0] + (r_range[1] - r_range[0]) * v[:, 0]
    he = he_range[0] + (he_range[1] - he_range[0]) * v[:, 0:3]
    s["dims"][:] = np.where(is_ball[:, None], np.stack([r, np.zeros(total), np.zeros(total)], 1), he)
    u1, u2, u3 = v[:, 3], 2 * np.pi * v[:, 4], 2 * np.pi * v[:, 5]
    q = np.stack([np.sqrt(1 - u1) * np.sin(u2), np.sqrt(1 - u1) * np.cos(u2),
                  np.sqrt(u1) * np.sin(u3), np.sqrt(u1) * np.cos(u3)], 1)
    q[is_ball] = (0, 0, 0, 1)
    s["rot"][:] = q
    bound = np.where(is_ball, r, np.sqrt((he * he).sum(1)))
    jit = np.maximum(0.0, cell / 2 - bound)[:, None] * (v[:, 6:9] * 2 - 1)
    c = np.stack([-half_width + (ix + 0.5) * cell, cell * 0.5 + 0.1 + layer * cell, -half_width + (iz + 0.5) * cell], 1)
    s["pos"][:] = c + jit
    s["friction"][:] = friction
    s["scene"] = world.astype(np.int32)
    g = statics([ground_half] * worlds, [(0.0, -ground_half[1], 0.0)] * worlds, friction)
    g["scene"] = np.arange(worlds, dtype=np.int32)
    return concat(g, s)
