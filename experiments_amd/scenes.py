"""Seeded scene generators for the benchmark and parity workloads.

Scenes are described host-side (shape point clouds + body arrays) and turned
into a World by build_world().  Random streams are PCG32 (O'Neill's
pcg32_random_r) with seed 12345 and one stream per quantity: shapes=1,
positions=2, orientations=3, velocities=4, masses=5.
"""
import math
from dataclasses import dataclass, field

import numpy as np

SEED = 12345
MASK64 = (1 << 64) - 1


class PCG32:
    def __init__(self, seed=SEED, stream=1):
        self.state = 0
        self.inc = ((stream << 1) | 1) & MASK64
        self.next_u32()
        self.state = (self.state + seed) & MASK64
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, n=None):
        if n is None:
            return (self.next_u32() >> 8) * (1.0 / (1 << 24))
        return np.array([(self.next_u32() >> 8) * (1.0 / (1 << 24)) for _ in range(n)])

    def normal(self):
        u1 = max(self.uniform(), 1e-12)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def quaternion(self):
        """Uniform random unit quaternion (Shoemake), (x, y, z, w)."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1 - u1), math.sqrt(u1)
        return (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))


def cube_points(half=0.5):
    h = half
    return np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)], np.float32)


def box_points(hx, hy, hz):
    return np.array([[x, y, z] for x in (-hx, hx) for y in (-hy, hy) for z in (-hz, hz)],
                    np.float32)


def random_hull_points(rng, nverts):
    """nverts points uniform on a sphere whose radius is U[0.5, 1.0]."""
    r = 0.5 + 0.5 * rng.uniform()
    pts = []
    for _ in range(nverts):
        v = np.array([rng.normal(), rng.normal(), rng.normal()])
        v /= np.linalg.norm(v)
        pts.append(v * r)
    return np.array(pts, np.float32)


@dataclass
class Scene:
    name: str
    shapes: list = field(default_factory=list)
    mass: list = field(default_factory=list)
    pos: list = field(default_factory=list)
    quat: list = field(default_factory=list)
    shape: list = field(default_factory=list)
    scene: list = field(default_factory=list)
    linvel: list = field(default_factory=list)
    steps: int = 1
    iterations: int = 4
    dt: float = 1.0 / 60.0
    friction: float = 0.7

    def add(self, mass, pos, quat, shape, scene=0, linvel=(0.0, 0.0, 0.0)):
        self.mass.append(mass)
        self.pos.append(pos)
        self.quat.append(quat)
        self.shape.append(shape)
        self.scene.append(scene)
        self.linvel.append(linvel)

    @property
    def num_bodies(self):
        return len(self.mass)


def lattice_cubes(nx=16, ny=16, nz=16, spacing=1.05, gap=1.0, with_ground=True):
    """4096 unit cubes (half extent 0.5, mass 1) in a 16x16x16 lattice with
    spacing 1.05, the lowest layer `gap` above the top of a static ground box."""
    s = Scene("lattice_%dx%dx%d" % (nx, ny, nz), steps=300, iterations=8, friction=0.6)
    s.shapes = [cube_points(0.5)]
    ident = (0.0, 0.0, 0.0, 1.0)
    if with_ground:
        s.shapes.append(box_points(60.0, 1.0, 60.0))
        s.add(0.0, (0.0, -1.0, 0.0), ident, 1)
    for j in range(ny):
        for k in range(nz):
            for i in range(nx):
                x = (i - (nx - 1) / 2.0) * spacing
                z = (k - (nz - 1) / 2.0) * spacing
                y = gap + 0.5 + j * spacing
                s.add(1.0, (x, y, z), ident, 0)
    return s


def hull_pool(n8=1024, n32=1024, seed=SEED):
    rng = PCG32(seed, 1)
    p8 = [random_hull_points(rng, 8) for _ in range(n8)]
    p32 = [random_hull_points(rng, 32) for _ in range(n32)]
    return p8, p32


def random_hulls(n=65536, box=120.0, pool=(1024, 1024), seed=SEED, steps=500, iterations=10):
    """n hulls: a third unit cubes, a third random 8-vertex and a third random
    32-vertex hulls (round-robin over a pool of distinct shapes), uniform in a
    box^3 region above a static ground, uniform orientations, N(0,1) linear
    velocity per axis, mass U[1, 10]."""
    s = Scene("random_hulls_%d" % n, steps=steps, iterations=iterations)
    p8, p32 = hull_pool(pool[0], pool[1], seed)
    s.shapes = [cube_points(0.5)] + p8 + p32 + [box_points(box, 1.0, box)]
    ground_shape = len(s.shapes) - 1
    rp, rq, rv, rm = PCG32(seed, 2), PCG32(seed, 3), PCG32(seed, 4), PCG32(seed, 5)
    c8 = c32 = 0
    for i in range(n):
        kind = i % 3
        if kind == 0:
            sh = 0
        elif kind == 1:
            sh = 1 + (c8 % len(p8))
            c8 += 1
        else:
            sh = 1 + len(p8) + (c32 % len(p32))
            c32 += 1
        pos = ((rp.uniform() - 0.5) * box, 2.0 + rp.uniform() * box, (rp.uniform() - 0.5) * box)
        q = rq.quaternion()
        v = (rv.normal(), rv.normal(), rv.normal())
        m = 1.0 + 9.0 * rm.uniform()
        s.add(m, pos, q, sh, 0, v)
    s.add(0.0, (0.0, -1.0, 0.0), (0.0, 0.0, 0.0, 1.0), ground_shape)
    return s


def batched_piles(n_scenes=4096, lo=8, hi=512, seed=SEED, steps=1000, iterations=10, pool8=1024):
    """n_scenes independent piles; body count per scene log-uniform in
    [lo, hi]; cubes and 8-vertex hulls dropped on the scene's own ground."""
    s = Scene("batched_%d" % n_scenes, steps=steps, iterations=iterations)
    p8, _ = hull_pool(pool8, 0, seed)
    s.shapes = [cube_points(0.5)] + p8 + [box_points(30.0, 1.0, 30.0)]
    ground_shape = len(s.shapes) - 1
    rc, rq, rp = PCG32(seed, 6), PCG32(seed, 3), PCG32(seed, 2)
    c8 = 0
    for sc in range(n_scenes):
        cnt = int(math.exp(math.log(lo) + rc.uniform() * (math.log(hi) - math.log(lo))))
        cnt = max(lo, min(hi, cnt))
        s.add(0.0, (0.0, -1.0, 0.0), (0.0, 0.0, 0.0, 1.0), ground_shape, sc)
        m = min(8, int(math.ceil(math.sqrt(cnt))))
        for b in range(cnt):
            layer, cell = divmod(b, m * m)
            ix, iz = divmod(cell, m)
            x = (ix - (m - 1) / 2.0) * 2.2 + (rp.uniform() - 0.5) * 0.2
            z = (iz - (m - 1) / 2.0) * 2.2 + (rp.uniform() - 0.5) * 0.2
            y = 1.5 + layer * 2.2
            if b % 2 == 0:
                sh = 0
            else:
                sh = 1 + (c8 % len(p8))
                c8 += 1
            s.add(1.0, (x, y, z), rq.quaternion(), sh, sc)
    return s


def split_scenes(scene):
    """The scenes of a batched Scene, each as a Scene of its own holding only
    the shapes it uses."""
    by_scene = {}
    for i in range(scene.num_bodies):
        by_scene.setdefault(scene.scene[i], []).append(i)
    out = []
    for sc in sorted(by_scene):
        sub = Scene("%s_scene%d" % (scene.name, sc), steps=scene.steps,
                    iterations=scene.iterations, dt=scene.dt, friction=scene.friction)
        remap = {}
        for i in by_scene[sc]:
            sh = scene.shape[i]
            if sh not in remap:
                remap[sh] = len(sub.shapes)
                sub.shapes.append(scene.shapes[sh])
            sub.add(scene.mass[i], scene.pos[i], scene.quat[i], remap[sh], 0, scene.linvel[i])
        out.append(sub)
    return out


def build_world(scene, **cfg):
    """Register the scene's shapes and bodies in a World and upload it."""
    from .world import World, reference_config

    n = scene.num_bodies
    defaults = dict(max_bodies=max(n, 16), solver_iterations=scene.iterations,
                    max_pairs_per_body=64,
                    dt=float(scene.dt), solver_dt=float(scene.dt), friction=scene.friction,
                    max_pairs=max(1 << 16, 8 * n), max_contacts=max(1 << 16, 6 * n),
                    max_hull_vertices=max(1 << 16, 40 * len(scene.shapes)),
                    max_shapes=len(scene.shapes) + 1)
    defaults.update(cfg)
    w = World(reference_config(**defaults))
    for pts in scene.shapes:
        w.add_hull(pts)
    for i in range(n):
        w.add_body(scene.mass[i], scene.pos[i], scene.quat[i], scene.shape[i], scene.scene[i])
    w.upload()
    lv = np.zeros((n, 4), np.float32)
    lv[:, :3] = np.asarray(scene.linvel, np.float32)
    if np.any(lv):
        w.set_state(linvel=lv)
    return w


def broadphase_aabbs(n=1 << 20, box=400.0, rmin=0.25, rmax=2.0, seed=SEED):
    """Broadphase stress input: n spheres and boxes (half each), radius (box
    half extent) log-uniform in [rmin, rmax], centres uniform in a box^3
    domain, uniform box orientations, N(0, 1) velocity per axis.  Numpy's
    PCG64 seeded with `seed`.  Returns aabb_min, aabb_max (float32 [n, 4]),
    velocity [n, 3] and radius [n]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r = np.exp(rng.uniform(np.log(rmin), np.log(rmax), n))
    c = rng.uniform(-0.5 * box, 0.5 * box, (n, 3))
    is_box = rng.integers(0, 2, n).astype(bool)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                  np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                  np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    ext = np.where(is_box[:, None], np.abs(R).sum(2) * r[:, None], r[:, None])
    mn = np.zeros((n, 4), np.float32)
    mx = np.zeros((n, 4), np.float32)
    mn[:, :3] = c - ext
    mx[:, :3] = c + ext
    v = rng.normal(size=(n, 3)).astype(np.float32)
    return mn, mx, v, r.astype(np.float32)
