"""Python host mirror of the reference's MechanicalWorld surface over libnphys_hip.so.

The reference (nphysics 0.23, Rust) drives a step as
``MechanicalWorld::step(&mut GeometricalWorld, &mut bodies, &mut colliders, &mut joints,
&mut forces)`` (src/world/mechanical_world.rs:182-196); here one MechanicalWorld object owns the
geometrical world and the body/collider sets of balls and cuboids, on one GPU, and every call
goes through the C ABI of include/nphys.h. There is no CPU fallback: constructing a world
without the HIP library or without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnphys_hip.so")

_lib = None

EXPORTED = [
    "nphys_params_default", "nphys_world_create", "nphys_world_destroy", "nphys_set_params",
    "nphys_get_params", "nphys_add_bodies", "nphys_num_bodies", "nphys_set_state", "nphys_set_collision_groups",
    "nphys_get_state", "nphys_step", "nphys_collide", "nphys_contact_model_set",
    "nphys_num_pairs", "nphys_read_pairs", "nphys_num_contacts", "nphys_read_contacts",
    "nphys_num_manifolds", "nphys_read_manifolds", "nphys_num_solved_contacts", "nphys_read_impulses", "nphys_residual", "nphys_read_counters",
    "nphys_reset_counters", "nphys_counters_enable", "nphys_set_residual_target", "nphys_read_colours",
    "nphys_set_sweep_variant", "nphys_set_incremental_pairs", "nphys_read_activation", "nphys_warmstart_load",
    "nphys_capture_samples", "nphys_read_samples", "nphys_set_warmstart_judge",
    "nphys_mlp_create", "nphys_mlp_destroy", "nphys_mlp_num_params", "nphys_mlp_get_params",
    "nphys_mlp_set_params", "nphys_mlp_set_data", "nphys_mlp_train", "nphys_mlp_forward",
    "nphys_mlp_gradients", "nphys_mlp_gradients_noalloc", "nphys_last_error", "nphys_apply_force", "nphys_apply_force_at_point",
    "nphys_set_material_coefficient", "nphys_add_joints", "nphys_remove_joint", "nphys_num_joints",
    "nphys_read_joints",
]


class NphysError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Loads the HIP library and declares the C signatures; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NphysError(f"{path} is missing; build it with python -m nphysics_amd.build")
    L = C.CDLL(path)
    W = C.c_void_p
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    L.nphys_last_error.restype = C.c_char_p
    L.nphys_params_default.argtypes = [C.POINTER(abi.Params)]
    L.nphys_params_default.restype = None
    L.nphys_world_create.argtypes = [C.POINTER(abi.Params), C.POINTER(W)]
    L.nphys_world_destroy.argtypes = [W]
    L.nphys_set_params.argtypes = [W, C.POINTER(abi.Params)]
    L.nphys_get_params.argtypes = [W, C.POINTER(abi.Params)]
    L.nphys_add_bodies.argtypes = [W, C.POINTER(abi.BodyDesc), ip]
    L.nphys_num_bodies.argtypes = [W, ip]
    L.nphys_set_state.argtypes = [W, fp, fp, fp, fp]
    L.nphys_set_collision_groups.argtypes = [W, C.c_int32, ip, C.POINTER(C.c_uint32)]
    L.nphys_get_state.argtypes = [W, fp, fp, fp, fp]
    L.nphys_step.argtypes = [W, C.c_int32]
    L.nphys_collide.argtypes = [W]
    L.nphys_contact_model_set.argtypes = [W, C.c_int32]
    L.nphys_num_pairs.argtypes = [W, C.POINTER(C.c_int64)]
    L.nphys_read_pairs.argtypes = [W, ip, C.c_int64]
    L.nphys_num_contacts.argtypes = [W, C.POINTER(C.c_int64)]
    L.nphys_read_contacts.argtypes = [W, ip, C.POINTER(C.c_uint32), fp, fp, fp, fp, C.c_int64]
    L.nphys_num_manifolds.argtypes = [W, C.POINTER(C.c_int64)]
    L.nphys_read_manifolds.argtypes = [W, ip, ip, ip, C.c_int64]
    L.nphys_num_solved_contacts.argtypes = [W, C.POINTER(C.c_int64)]
    L.nphys_read_impulses.argtypes = [W, C.POINTER(C.c_uint64), fp, C.c_int64]
    L.nphys_residual.argtypes = [W, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.nphys_read_counters.argtypes = [W, C.POINTER(abi.Counters)]
    L.nphys_reset_counters.argtypes = [W]
    L.nphys_counters_enable.argtypes = [W, C.c_int32]
    L.nphys_set_residual_target.argtypes = [W, C.c_double, C.c_int32, C.c_int32, C.c_int32]
    L.nphys_warmstart_load.argtypes = [W, fp, C.c_int64]
    L.nphys_capture_samples.argtypes = [W, C.c_int32]
    L.nphys_read_samples.argtypes = [W, fp, fp, C.c_int64, C.POINTER(C.c_int64)]
    L.nphys_set_warmstart_judge.argtypes = [W, C.c_int32]
    L.nphys_set_sweep_variant.argtypes = [W, C.c_int32]
    L.nphys_set_incremental_pairs.argtypes = [W, C.c_int32]
    L.nphys_read_colours.argtypes = [W, ip, C.c_int64, ip]
    L.nphys_read_activation.argtypes = [W, fp, C.c_int64]
    L.nphys_apply_force.argtypes = [W, C.c_int32, ip, fp, C.c_int32, C.c_int32]
    L.nphys_apply_force_at_point.argtypes = [W, C.c_int32, ip, fp, fp, C.c_int32, C.c_int32]
    L.nphys_set_material_coefficient.argtypes = [W, C.c_int32, C.c_int32, C.c_int32, C.c_float]
    L.nphys_add_joints.argtypes = [W, C.POINTER(abi.JointDesc), ip]
    L.nphys_remove_joint.argtypes = [W, C.c_int32]
    L.nphys_num_joints.argtypes = [W, ip]
    L.nphys_read_joints.argtypes = [W, fp, fp, ip, C.c_int64]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise NphysError(f"nphys error {rc}: {_lib.nphys_last_error().decode()}")


def default_params(**kw):
    return abi.default_params(**kw)


class MechanicalWorld:
    """One world on the current HIP device (MechanicalWorld + GeometricalWorld + sets)."""

    def __init__(self, gravity=None, integration_parameters=None):
        """MechanicalWorld::new(gravity) (mechanical_world.rs:76-102). The gravity argument, when
        given, replaces the one carried by integration_parameters (default (0, -9.81, 0))."""
        self._L = load_library()
        p = abi.Params()
        C.pointer(p)[0] = integration_parameters if integration_parameters is not None else abi.default_params()
        if gravity is not None:
            p.gravity[0], p.gravity[1], p.gravity[2] = gravity
        self._h = C.c_void_p()
        _check(self._L.nphys_world_create(C.byref(p), C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.nphys_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters (integration_parameters / gravity pub fields, set_timestep)
    @property
    def integration_parameters(self):
        p = abi.Params()
        _check(self._L.nphys_get_params(self._h, C.byref(p)))
        return p

    @integration_parameters.setter
    def integration_parameters(self, p):
        _check(self._L.nphys_set_params(self._h, C.byref(p)))

    def timestep(self):
        return self.integration_parameters.dt

    def set_timestep(self, dt):
        p = self.integration_parameters
        p.dt = dt
        self.integration_parameters = p

    def set_contact_model(self, model):
        _check(self._L.nphys_contact_model_set(self._h, model))

    # -- scene
    def add_bodies(self, scene):
        desc, keep = abi.make_body_desc(scene)
        first = C.c_int32()
        _check(self._L.nphys_add_bodies(self._h, C.byref(desc), C.byref(first)))
        return first.value

    @property
    def num_bodies(self):
        n = C.c_int32()
        _check(self._L.nphys_num_bodies(self._h, C.byref(n)))
        return n.value

    def set_collision_groups(self, bodies, groups):
        """Collider::set_collision_groups for the given bodies; groups: (count, 3) masks."""
        b = np.ascontiguousarray(bodies, np.int32).reshape(-1)
        g = np.ascontiguousarray(groups, np.uint32).reshape(-1, 3)
        _check(self._L.nphys_set_collision_groups(self._h, int(b.shape[0]), abi.iptr(b),
                                                  g.ctypes.data_as(C.POINTER(C.c_uint32))))

    def set_state(self, pos=None, rot=None, linvel=None, angvel=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (pos, rot, linvel, angvel)]
        ptrs = [None if a is None else abi.fptr(a) for a in arrs]
        _check(self._L.nphys_set_state(self._h, *ptrs))

    def get_state(self):
        n = self.num_bodies
        pos = np.zeros((n, 3), np.float32)
        rot = np.zeros((n, 4), np.float32)
        lv = np.zeros((n, 3), np.float32)
        av = np.zeros((n, 3), np.float32)
        _check(self._L.nphys_get_state(self._h, abi.fptr(pos), abi.fptr(rot), abi.fptr(lv), abi.fptr(av)))
        return {"pos": pos, "rot": rot, "linvel": lv, "angvel": av}

    def apply_force(self, bodies, forces, force_type=abi.FORCE, auto_wake_up=True, points=None):
        """Body::apply_force per entry (forces (k, 6): linear, torque; world frame), or, with
        `points`, apply_force_at_point (forces (k, 3) linear at the world points (k, 3)).
        force_type: abi.FORCE, IMPULSE, ACCELERATION_CHANGE or VELOCITY_CHANGE."""
        b = np.ascontiguousarray(bodies, np.int32).reshape(-1)
        f = np.ascontiguousarray(forces, np.float32).reshape(-1)
        if points is None:
            _check(self._L.nphys_apply_force(self._h, b.shape[0], abi.iptr(b), abi.fptr(f), force_type,
                                             int(auto_wake_up)))
        else:
            p = np.ascontiguousarray(points, np.float32).reshape(-1)
            _check(self._L.nphys_apply_force_at_point(self._h, b.shape[0], abi.iptr(b), abi.fptr(f), abi.fptr(p),
                                                      force_type, int(auto_wake_up)))

    # -- joints (DefaultJointConstraintSet, src/joint/joint_constraint.rs:61-107)
    def add_joints(self, js):
        """Inserts the joints of an abi.joints() dict; returns the first joint handle."""
        desc, keep = abi.make_joint_desc(js)
        first = C.c_int32()
        _check(self._L.nphys_add_joints(self._h, C.byref(desc), C.byref(first)))
        return first.value

    def remove_joint(self, j):
        _check(self._L.nphys_remove_joint(self._h, j))

    def joints(self):
        """Cached impulses (lin (n, 3), ang (n, 3)) and broken flags of every joint."""
        n = C.c_int32()
        _check(self._L.nphys_num_joints(self._h, C.byref(n)))
        k = max(n.value, 1)
        lin = np.zeros((k, 3), np.float32)
        ang = np.zeros((k, 3), np.float32)
        br = np.zeros(k, np.int32)
        _check(self._L.nphys_read_joints(self._h, abi.fptr(lin), abi.fptr(ang), abi.iptr(br), k))
        return {"lin": lin[:n.value], "ang": ang[:n.value], "broken": br[:n.value].astype(bool)}

    def set_material_coefficient(self, which, id1, id2, coeff):
        """MaterialsCoefficientsTable: which 0 friction, 1 restitution, for material ids id1, id2."""
        _check(self._L.nphys_set_material_coefficient(self._h, which, id1, id2, coeff))

    # -- stepping
    def step(self, nsteps=1):
        _check(self._L.nphys_step(self._h, nsteps))

    def collide(self):
        _check(self._L.nphys_collide(self._h))

    # -- readback
    def pairs(self):
        n = C.c_int64()
        _check(self._L.nphys_num_pairs(self._h, C.byref(n)))
        out = np.zeros((n.value, 2), np.int32)
        _check(self._L.nphys_read_pairs(self._h, abi.iptr(out), n.value))
        return out

    def contacts(self):
        n = C.c_int64()
        _check(self._L.nphys_num_contacts(self._h, C.byref(n)))
        c = n.value
        d = {
            "bodies": np.zeros((c, 2), np.int32),
            "feature": np.zeros(c, np.uint32),
            "world1": np.zeros((c, 3), np.float32),
            "world2": np.zeros((c, 3), np.float32),
            "normal": np.zeros((c, 3), np.float32),
            "depth": np.zeros(c, np.float32),
        }
        _check(self._L.nphys_read_contacts(self._h, abi.iptr(d["bodies"]), abi.u32ptr(d["feature"]),
                                           abi.fptr(d["world1"]), abi.fptr(d["world2"]),
                                           abi.fptr(d["normal"]), abi.fptr(d["depth"]), c))
        return d

    def impulses(self):
        n = C.c_int64()
        _check(self._L.nphys_num_solved_contacts(self._h, C.byref(n)))
        keys = np.zeros(n.value, np.uint64)
        imp = np.zeros((n.value, 3), np.float32)
        _check(self._L.nphys_read_impulses(self._h, abi.u64ptr(keys), abi.fptr(imp), n.value))
        return keys, imp

    def residual(self):
        a, b = C.c_double(), C.c_double()
        _check(self._L.nphys_residual(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def counters(self):
        c = abi.Counters()
        _check(self._L.nphys_read_counters(self._h, C.byref(c)))
        return c.as_dict()

    def reset_counters(self):
        _check(self._L.nphys_reset_counters(self._h))

    def enable_counters(self, on=True):
        _check(self._L.nphys_counters_enable(self._h, 1 if on else 0))

    def set_residual_target(self, target=1e-4, max_sweeps=500, check_every=1, norm="max"):
        """Solve each step until the residual (norm "max" or "rms") is <= target (0 restores the
        fixed iteration count)."""
        _check(self._L.nphys_set_residual_target(self._h, target, max_sweeps, check_every,
                                                 0 if norm == "max" else 1))

    # -- learned warm start
    def load_warmstart(self, params):
        """Loads a warm-start network (nphysics_amd.learned.WarmStartNet.params) into the world."""
        p = np.ascontiguousarray(params, np.float32)
        _check(self._L.nphys_warmstart_load(self._h, abi.fptr(p), p.shape[0]))

    def capture_samples(self, on=True):
        _check(self._L.nphys_capture_samples(self._h, 1 if on else 0))

    def samples(self):
        """(features [n][24], targets [n][3]) recorded by the last step."""
        n = C.c_int64()
        _check(self._L.nphys_read_samples(self._h, None, None, 0, C.byref(n)))
        X = np.zeros((n.value, 24), np.float32)
        T = np.zeros((n.value, 3), np.float32)
        if n.value:
            _check(self._L.nphys_read_samples(self._h, abi.fptr(X), abi.fptr(T), n.value, C.byref(n)))
        return X, T

    def set_sweep_variant(self, variant=-1):
        """-1 autotune (default), 0 per-colour launches, 256/512/1024 scene-blocked (lanes per large scene)."""
        _check(self._L.nphys_set_sweep_variant(self._h, variant))

    def set_incremental_pairs(self, on=True):
        """Incremental pair-set updates (default) or a full tree build on every change."""
        _check(self._L.nphys_set_incremental_pairs(self._h, 1 if on else 0))

    def set_warmstart_judge(self, on=True):
        _check(self._L.nphys_set_warmstart_judge(self._h, 1 if on else 0))

    def manifolds(self):
        """Contact manifolds of the last collision detection (GeometricalWorld::contact_pairs):
        bodies (M, 2), index of the first contact in contacts() order, contact count."""
        n = C.c_int64()
        _check(self._L.nphys_num_manifolds(self._h, C.byref(n)))
        m = n.value
        b = np.zeros((max(m, 1), 2), np.int32)
        f = np.zeros(max(m, 1), np.int32)
        k = np.zeros(max(m, 1), np.int32)
        _check(self._L.nphys_read_manifolds(self._h, abi.iptr(b), abi.iptr(f), abi.iptr(k), max(m, 1)))
        return {"bodies": b[:m], "first": f[:m], "count": k[:m]}

    def activation(self):
        """Per-body activation energy (0: asleep; 0.04 for bodies that never sleep)."""
        n = self.num_bodies
        e = np.zeros(max(n, 1), np.float32)
        _check(self._L.nphys_read_activation(self._h, abi.fptr(e), e.shape[0]))
        return e[:n]

    def record_colours(self):
        """Starts recording the manifold colours of every following step."""
        nc = C.c_int32()
        _check(self._L.nphys_read_colours(self._h, None, 0, C.byref(nc)))

    def colours(self):
        """Colours of the manifolds solved by the last step (needs record_colours() before it)."""
        m = self.counters()["nmanifolds"]
        out = np.zeros(max(m, 1), np.int32)
        nc = C.c_int32()
        _check(self._L.nphys_read_colours(self._h, abi.iptr(out), out.shape[0], C.byref(nc)))
        return out[:m], nc.value
