"""Host-side mirror of the reference's GPU physics API over the C ABI.

`World` is a thin Pythonic wrapper of mi_rb_* calls; `CLPhysicsDemo` keeps the
reference method names and argument meaning of
opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.h:48-80 so that code written
against the reference reads the same.
"""
import ctypes

import numpy as np

from . import _lib as L


def reference_config(**overrides):
    """mi_rb_config with the constants hard-coded in the reference kernels."""
    cfg = L.Config()
    L.lib().mi_rb_config_reference(ctypes.byref(cfg))
    for k, v in overrides.items():
        if k in ("gravity", "grid_dims"):
            arr = getattr(cfg, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


def _f32(a, n=None):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class World:
    """One simulation world on one GPU (mi_rb_create ... mi_rb_destroy)."""

    def __init__(self, cfg=None, **overrides):
        self.cfg = cfg if cfg is not None else reference_config(**overrides)
        h = ctypes.c_void_p()
        L.call("mi_rb_create", ctypes.byref(self.cfg), ctypes.byref(h))
        self.h = h
        self.num_bodies = 0
        self.num_shapes = 0

    def close(self):
        if self.h:
            L.lib().mi_rb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- registration ----
    def add_hull(self, verts, scale=(1.0, 1.0, 1.0)):
        v = _f32(verts).reshape(-1, 3).copy()
        s = _f32(scale)
        out = ctypes.c_int()
        L.call("mi_rb_register_convex_hull", self.h, _ptr(v), 12, len(v), _ptr(s),
               ctypes.byref(out))
        self.num_shapes += 1
        return out.value

    def add_polyhedron(self, verts, planes, face_index, indices, edges):
        """Polyhedral features as given (a btConvexUtility's: vertices xyz,
        face planes nx ny nz w, (first index, count) per face, face vertex
        indices, unique edge directions xyz), kept in their order."""
        v = _f32(verts).reshape(-1, 3).copy()
        pl = _f32(planes).reshape(-1, 4).copy()
        fi = np.ascontiguousarray(face_index, np.int32).reshape(-1, 2).copy()
        ind = np.ascontiguousarray(indices, np.int32).ravel().copy()
        e = _f32(edges).reshape(-1, 3).copy()
        out = ctypes.c_int()
        L.call("mi_rb_register_convex_polyhedron", self.h, _ptr(v), 12, len(v), _ptr(pl),
               _ptr(fi), len(fi), _ptr(ind), len(ind), _ptr(e), 12, len(e), ctypes.byref(out))
        self.num_shapes += 1
        return out.value

    def add_sphere(self, radius):
        out = ctypes.c_int()
        L.call("mi_rb_register_sphere", self.h, ctypes.c_float(radius), ctypes.byref(out))
        self.num_shapes += 1
        return out.value

    def add_plane(self, normal, constant):
        n = _f32(normal)[:3].copy()
        out = ctypes.c_int()
        L.call("mi_rb_register_plane", self.h, _ptr(n), ctypes.c_float(constant),
               ctypes.byref(out))
        self.num_shapes += 1
        return out.value

    def add_compound(self, child_shapes, child_pos, child_orn):
        """A rigid union of hull shapes: child i at child_pos[i] with
        orientation child_orn[i] (xyzw) in the compound's frame."""
        sh = np.ascontiguousarray(child_shapes, np.int32)
        cp = _f32(child_pos).reshape(-1, 3).copy()
        co = _f32(child_orn).reshape(-1, 4).copy()
        if not (len(sh) == len(cp) == len(co)):
            raise ValueError("one position and one orientation per child")
        out = ctypes.c_int()
        L.call("mi_rb_register_compound", self.h, len(sh), _ptr(sh), _ptr(cp), _ptr(co),
               ctypes.byref(out))
        self.num_shapes += 1
        return out.value

    def add_concave_mesh(self, vertices, indices, scaling=(1.0, 1.0, 1.0)):
        """A triangle mesh for static bodies: vertices (V, 3), triangles as
        index triples; each triangle takes a shape slot, the mesh one more."""
        v = _f32(vertices).reshape(-1, 3).copy()
        ix = np.ascontiguousarray(indices, np.int32).reshape(-1)
        if len(ix) % 3:
            raise ValueError("three indices per triangle")
        sc = _f32(scaling)[:3].copy()
        out = ctypes.c_int()
        L.call("mi_rb_register_concave_mesh", self.h, _ptr(v), 12, len(v), _ptr(ix), len(ix) // 3,
               _ptr(sc), ctypes.byref(out))
        self.num_shapes = out.value + 1
        return out.value

    def add_concave_mesh_obj(self, path, scaling=(1.0, 1.0, 1.0)):
        """A triangle mesh read from a Wavefront .obj file."""
        sc = _f32(scaling)[:3].copy()
        out = ctypes.c_int()
        L.call("mi_rb_register_concave_mesh_obj", self.h, str(path).encode(), _ptr(sc),
               ctypes.byref(out))
        self.num_shapes = out.value + 1
        return out.value

    def add_body(self, mass, pos, quat_xyzw, shape, scene=0):
        p = _f32(pos)
        q = _f32(quat_xyzw)
        out = ctypes.c_int()
        L.call("mi_rb_register_body", self.h, ctypes.c_float(mass), _ptr(p), _ptr(q), shape, scene,
               ctypes.byref(out))
        self.num_bodies += 1
        return out.value

    def upload(self):
        L.call("mi_rb_upload", self.h)

    # ---- state ----
    def set_state(self, pos=None, quat=None, linvel=None, angvel=None):
        arrs = []
        n = self.body_count()
        for a in (pos, quat, linvel, angvel):
            if a is None:
                arrs.append(None)
            else:
                a = _f32(a)
                if len(a) != n:
                    raise ValueError("state arrays need one row per body (%d)" % n)
                if a.shape[-1] == 3:
                    b = np.zeros((len(a), 4), np.float32)
                    b[:, :3] = a
                    a = b
                arrs.append(np.ascontiguousarray(a))
        L.call("mi_rb_set_body_state", self.h, *[None if a is None else _ptr(a) for a in arrs])

    def copy_transforms(self, dst, start_offset=0, stream=None):
        """Body transforms into a device float4 buffer `dst` (a torch tensor
        or a device pointer with `capacity` float4): (x, y, z, 1) of body i at
        start_offset/4 + i, its quaternion at start_offset/4 + n + i."""
        ptr, cap = (dst.data_ptr(), dst.numel() // 4) if hasattr(dst, "data_ptr") else dst
        L.call("mi_rb_copy_transforms", self.h, ctypes.c_void_p(ptr), int(start_offset), int(cap),
               ctypes.c_void_p(stream) if stream else None)

    def body_count(self):
        """Bodies registered in the library (however they were registered)."""
        return self.stats()["num_bodies"]

    def get_state(self):
        n = self.body_count()
        out = [np.zeros((n, 4), np.float32) for _ in range(4)]
        L.call("mi_rb_get_body_state", self.h, *[_ptr(a) for a in out])
        return dict(pos=out[0], quat=out[1], linvel=out[2], angvel=out[3])

    # ---- stepping ----
    def step(self, nsteps=1):
        L.call("mi_rb_step", self.h, nsteps)

    def step_partial(self, stage):
        L.call("mi_rb_step_partial", self.h, stage)

    def set_solve_timing(self, enable=True):
        """HIP events around the solver kernel's launches in every step
        (stats: solve_kernel_ms, solve_launches of the last step)."""
        L.call("mi_rb_set_solve_timing", self.h, 1 if enable else 0)

    def use_graph(self, enable=True):
        L.call("mi_rb_use_graph", self.h, 1 if enable else 0)

    def last_step_ms(self):
        ms = ctypes.c_float()
        L.call("mi_rb_last_step_ms", self.h, ctypes.byref(ms))
        return ms.value

    def stats(self):
        st = L.Stats()
        L.call("mi_rb_get_stats", self.h, ctypes.byref(st))
        return {k: (list(getattr(st, k)) if k == "stage_ms" else getattr(st, k))
                for k, _ in L.Stats._fields_}

    def params(self):
        pr = L.Params()
        L.call("mi_rb_get_params", self.h, ctypes.byref(pr))
        return {k: (list(getattr(pr, k)) if k == "dims" else getattr(pr, k))
                for k, _ in L.Params._fields_}

    def read(self, which, dtype, shape_tail=()):
        cnt = ctypes.c_int()
        L.call("mi_rb_read_buffer", self.h, which, None, 0, ctypes.byref(cnt))
        n = cnt.value
        arr = np.zeros((n,) + tuple(shape_tail), dtype)
        L.call("mi_rb_read_buffer", self.h, which, _ptr(arr) if arr.size else None, arr.nbytes,
               ctypes.byref(cnt))
        return arr

    def sorted_kv(self):
        return self.read(L.BUF_SORTED_KV, np.uint32, (2,))

    def aabbs(self):
        a = self.read(L.BUF_AABB, np.float32, (8,))
        n = len(a)
        flat = a.reshape(-1)
        return flat[:4 * n].reshape(n, 4), flat[4 * n:].reshape(n, 4)

    def pairs(self):
        return self.read(L.BUF_PAIRS, np.int32, (2,))

    def contacts(self, dtype):
        return self.read(L.BUF_CONTACTS, dtype)

    def set_solver(self, iterations, tolerance=0.0, warm_start=L.WARM_ZERO):
        L.call("mi_rb_set_solver", self.h, int(iterations), float(tolerance), int(warm_start))
        self.cfg.solver_iterations = int(iterations)
        self.cfg.residual_tolerance = float(tolerance)
        self.cfg.warm_start = int(warm_start)

    def set_tuning(self, solve_block=0, narrow_waves=0, sort_tile=0, sort_digit_bits=0):
        t = L.Tuning(solve_block, narrow_waves, sort_tile, sort_digit_bits)
        L.call("mi_rb_set_tuning", self.h, ctypes.byref(t))

    def get_tuning(self):
        t = L.Tuning()
        L.call("mi_rb_get_tuning", self.h, ctypes.byref(t))
        return {k: getattr(t, k) for k, _ in L.Tuning._fields_}

    def autotune(self, steps=5):
        """Time every launch shape from the current state (restored after
        each trial), apply the fastest; returns (best, {shape: ms/step})."""
        t = L.Tuning()
        ms = np.zeros(8, np.float32)
        L.call("mi_rb_autotune", self.h, int(steps), ctypes.byref(t), _ptr(ms))
        shapes = [(b, wv) for b in (256, 512) for wv in (2, 4, 6, 8)]
        return ({k: getattr(t, k) for k, _ in L.Tuning._fields_},
                {"%d/%d" % sh: float(m) for sh, m in zip(shapes, ms)})

    def set_warmstart_weights(self, w_bf16, bias):
        w = np.ascontiguousarray(w_bf16, np.uint16).reshape(-1)
        b = np.ascontiguousarray(bias, np.float32).reshape(-1)
        if w.size != L.MLP_NW or b.size != L.MLP_NB:
            raise ValueError("expected %d bf16 weights and %d biases" % (L.MLP_NW, L.MLP_NB))
        L.call("mi_rb_set_warmstart_weights", self.h, _ptr(w), _ptr(b))

    def copy_training_batch(self, x_dev_ptr, t_dev_ptr, capacity):
        """Copy the last step's features and solved impulses (network units)
        into device buffers; returns the number of contacts written."""
        cnt = ctypes.c_int()
        L.call("mi_rb_copy_training_batch", self.h, ctypes.c_void_p(x_dev_ptr),
               ctypes.c_void_p(t_dev_ptr), int(capacity), ctypes.byref(cnt))
        return cnt.value

    def load_bullet(self, path, keep_static=False):
        """Register a .bullet file's boxes, convex hulls and bodies (before
        upload), as the reference's btBulletDataExtractor does; returns the
        summary counts."""
        from .bullet import load_bullet
        summary = load_bullet(self, path, keep_static)
        self.num_bodies += summary["bodies_registered"] + summary["ground_bodies"]
        self.num_shapes += summary["shapes_registered"] + (1 if summary["ground_bodies"] else 0)
        return summary

    def save(self, path):
        L.call("mi_rb_save_state", self.h, path.encode())

    def load(self, path):
        L.call("mi_rb_load_state", self.h, path.encode())


class CLPhysicsDemo:
    """Reference method names over World (CLPhysicsDemo.h:48-80)."""

    def __init__(self, maxShapeBufferCapacityInBytes=0, maxNumObjectCapacity=1 << 17, cfg=None):
        self.world = World(cfg if cfg is not None else
                           reference_config(max_bodies=int(maxNumObjectCapacity)))
        self.m_numPhysicsInstances = 0
        self.m_numDynamicPhysicsInstances = 0

    def init(self, preferredDevice=0, preferredPlatform=0, useInterop=False):
        return None

    def registerConvexPolyhedron(self, vertices, strideInBytes, numVertices, scaling,
                                 noHeightField=True):
        v = np.frombuffer(np.ascontiguousarray(vertices).tobytes(), dtype=np.uint8)
        stride_f = strideInBytes // 4
        pts = np.frombuffer(v.tobytes(), np.float32).reshape(-1, stride_f)[:numVertices, :3]
        return self.world.add_hull(pts, scaling[:3])

    def registerSphereShape(self, radius):
        return self.world.add_sphere(radius)

    def registerPlaneShape(self, planeNormal, planeConstant):
        return self.world.add_plane(planeNormal, planeConstant)

    def registerConcaveMesh(self, vertices, indices, scaling=(1.0, 1.0, 1.0)):
        return self.world.add_concave_mesh(vertices, indices, scaling)

    def registerCompoundShape(self, childShapes):
        """childShapes: records with m_childPosition, m_childOrientation and
        m_shapeIndex (btGpuChildShape), or (shape, position, orientation)."""
        sh, cp, co = [], [], []
        for c in childShapes:
            if isinstance(c, (tuple, list)):
                s_, p_, o_ = c
            else:
                s_, p_, o_ = c.m_shapeIndex, c.m_childPosition, c.m_childOrientation
            sh.append(int(s_))
            cp.append(list(p_)[:3])
            co.append(list(o_)[:4])
        return self.world.add_compound(sh, cp, co)

    def registerPhysicsInstance(self, mass, position, orientation, collidableShape, userIndex):
        idx = self.world.add_body(mass, position[:3], orientation[:4], collidableShape, 0)
        self.m_numPhysicsInstances += 1
        if mass > 0:
            self.m_numDynamicPhysicsInstances += 1
        return idx

    def writeBodiesToGpu(self):
        self.world.upload()

    def createSceneFromBulletFile(self, fileName, keepStaticObjects=False):
        """createScene + btBulletDataExtractor::convertAllObjects
        (BulletDataExtractor.cpp:454-520, 594-700), then writeBodiesToGpu."""
        summary = self.world.load_bullet(fileName, keepStaticObjects)
        self.m_numPhysicsInstances += summary["bodies_registered"] + summary["ground_bodies"]
        self.writeBodiesToGpu()
        return summary

    def writeVelocitiesToGpu(self):
        return None

    def setObjectTransform(self, position, orientation, objectIndex):
        p = _f32(position[:3])
        q = _f32(orientation[:4])
        L.call("mi_rb_set_transform", self.world.h, objectIndex, _ptr(p), _ptr(q))

    def setObjectLinearVelocity(self, linVel, objectIndex):
        v = _f32(linVel[:3])
        L.call("mi_rb_set_linvel", self.world.h, objectIndex, _ptr(v))

    def stepSimulation(self):
        self.world.step(1)

    def readbackBodiesToCpu(self):
        L.call("mi_rb_readback", self.world.h)

    def getObjectTransformFromCpu(self, objectIndex):
        p = np.zeros(4, np.float32)
        q = np.zeros(4, np.float32)
        L.call("mi_rb_get_transform", self.world.h, objectIndex, _ptr(p), _ptr(q))
        return p, q

    def getBodiesGpu(self):
        v = L.DeviceView()
        L.call("mi_rb_device_ptrs", self.world.h, ctypes.byref(v))
        return v

    def getNumObjects(self):
        return self.m_numPhysicsInstances

    def cleanup(self):
        self.world.close()


def shape_arrays(world):
    """Registered hull arrays of a World as numpy (layouts of rb_oracle.h)."""
    import numpy as _np
    hull_dtype = _np.dtype([
        ("face_offset", "<i4"), ("num_faces", "<i4"), ("vertex_offset", "<i4"),
        ("num_vertices", "<i4"), ("edge_offset", "<i4"), ("num_edges", "<i4"),
        ("index_offset", "<i4"), ("num_indices", "<i4"),
        ("center", "<f4", 4), ("aabb_min", "<f4", 4), ("aabb_max", "<f4", 4)])
    return dict(
        hulls=world.read(L.BUF_HULLS, _np.uint8, (80,)).view(hull_dtype).reshape(-1),
        verts=world.read(L.BUF_VERTS, _np.float32, (4,)),
        planes=world.read(L.BUF_PLANES, _np.float32, (4,)),
        face_index=world.read(L.BUF_FACE_INDEX, _np.int32, (2,)),
        indices=world.read(L.BUF_INDICES, _np.int32),
        edges=world.read(L.BUF_EDGES, _np.float32, (4,)),
        inv_mass=world.read(L.BUF_INV_MASS, _np.float32),
        inv_inertia_local=world.read(L.BUF_INV_INERTIA_LOCAL, _np.float32, (4,)),
        scene=world.read(L.BUF_SCENE, _np.int32),
        shape=world.read(L.BUF_SHAPE, _np.int32))
