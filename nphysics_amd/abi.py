"""ctypes mirrors of the structs declared in include/nphys.h.

The same structs are accepted by the product library (libnphys_hip.so) and by the CPU checker
used in tests, so scene descriptions are built once and handed to either.
"""
import ctypes as C

import numpy as np

NPHYS_OK = 0
NPHYS_ERR_INVALID = -1
NPHYS_ERR_HIP = -2
NPHYS_ERR_CAPACITY = -3
NPHYS_ERR_STATE = -4
NPHYS_ERR_NO_DEVICE = -5

STATUS_DYNAMIC = 0
STATUS_STATIC = 1
SHAPE_BALL = 0
SHAPE_CUBOID = 1
CONTACT_MODEL_SIGNORINI_COULOMB_PYRAMID = 0
CONTACT_MODEL_LEARNED_WARMSTART = 1
COMBINE_AVERAGE = 0
COMBINE_MIN = 1
COMBINE_MULTIPLY = 2
COMBINE_MAX = 3
FORCE = 0
IMPULSE = 1
ACCELERATION_CHANGE = 2
VELOCITY_CHANGE = 3

JOINT_BALL = 0
JOINT_FIXED = 1
JOINT_REVOLUTE = 2
JOINT_PRISMATIC = 3
JOINT_CYLINDRICAL = 4

# optional per-body arrays of nphys_body_desc: name -> (dtype, width)
OPTIONAL_BODY_FIELDS = {
    "scene": (np.int32, 1),
    "sleep_threshold": (np.float32, 1),
    "linear_damping": (np.float32, 1),
    "angular_damping": (np.float32, 1),
    "max_linear_velocity": (np.float32, 1),
    "max_angular_velocity": (np.float32, 1),
    "gravity_enabled": (np.int32, 1),
    "kinematic_dofs": (np.int32, 1),
    "friction_combine": (np.int32, 1),
    "restitution_combine": (np.int32, 1),
    "material_id": (np.int32, 1),
    "surface_velocity": (np.float32, 3),
    "ccd_enabled": (np.int32, 1),
    "linear_motion_interpolation": (np.int32, 1),
    "static_body": (np.int32, 1),
    "collision_groups": (np.uint32, 3),
}

# CollisionGroups masks (ncollide3d 0.31): groups 0-29
ALL_GROUPS = (1 << 30) - 1


def collision_groups(membership=None, whitelist=None, blacklist=()):
    """(membership, whitelist, blacklist) masks of CollisionGroups::new() followed by
    set_membership / set_whitelist / set_blacklist with these group lists (None keeps the
    default: all groups for the first two)."""
    def mask(groups, default):
        if groups is None:
            return default
        m = 0
        for g in groups:
            if not 0 <= g < 30:
                raise ValueError("collision group ids are 0-29")
            m |= 1 << g
        return m
    return (mask(membership, ALL_GROUPS), mask(whitelist, ALL_GROUPS), mask(blacklist, 0))


class Params(C.Structure):
    """IntegrationParameters (src/solver/integration_parameters.rs:5-76) + gravity."""

    _fields_ = [
        ("t", C.c_double),
        ("dt", C.c_double),
        ("erp", C.c_double),
        ("warmstart_coeff", C.c_double),
        ("restitution_velocity_threshold", C.c_double),
        ("allowed_linear_error", C.c_double),
        ("allowed_angular_error", C.c_double),
        ("max_linear_correction", C.c_double),
        ("max_angular_correction", C.c_double),
        ("max_stabilization_multiplier", C.c_double),
        ("max_velocity_iterations", C.c_int32),
        ("max_position_iterations", C.c_int32),
        ("max_ccd_position_iterations", C.c_int32),
        ("max_ccd_substeps", C.c_int32),
        ("return_after_ccd_substep", C.c_int32),
        ("multiple_ccd_substep_sensor_events_enabled", C.c_int32),
        ("ccd_on_penetration_enabled", C.c_int32),
        ("reserved0", C.c_int32),
        ("gravity", C.c_double * 3),
    ]


def default_params(**overrides):
    """IntegrationParameters::default() (integration_parameters.rs:169-189), gravity -9.81 y."""
    p = Params()
    p.t = 0.0
    p.dt = 1.0 / 60.0
    p.erp = 0.2
    p.warmstart_coeff = 1.0
    p.restitution_velocity_threshold = 1.0
    p.allowed_linear_error = 0.001
    p.allowed_angular_error = 0.001
    p.max_linear_correction = 0.2
    p.max_angular_correction = 0.2
    p.max_stabilization_multiplier = 0.2
    p.max_velocity_iterations = 8
    p.max_position_iterations = 3
    p.max_ccd_position_iterations = 10
    p.max_ccd_substeps = 1
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    for k, v in overrides.items():
        if k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        else:
            setattr(p, k, v)
    return p


class BodyDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("status", C.POINTER(C.c_int32)),
        ("shape", C.POINTER(C.c_int32)),
        ("dims", C.POINTER(C.c_float)),
        ("pos", C.POINTER(C.c_float)),
        ("rot", C.POINTER(C.c_float)),
        ("linvel", C.POINTER(C.c_float)),
        ("angvel", C.POINTER(C.c_float)),
        ("density", C.POINTER(C.c_float)),
        ("margin", C.POINTER(C.c_float)),
        ("linear_prediction", C.POINTER(C.c_float)),
        ("friction", C.POINTER(C.c_float)),
        ("restitution", C.POINTER(C.c_float)),
        ("scene", C.POINTER(C.c_int32)),
        ("sleep_threshold", C.POINTER(C.c_float)),
        ("linear_damping", C.POINTER(C.c_float)),
        ("angular_damping", C.POINTER(C.c_float)),
        ("max_linear_velocity", C.POINTER(C.c_float)),
        ("max_angular_velocity", C.POINTER(C.c_float)),
        ("gravity_enabled", C.POINTER(C.c_int32)),
        ("kinematic_dofs", C.POINTER(C.c_int32)),
        ("friction_combine", C.POINTER(C.c_int32)),
        ("restitution_combine", C.POINTER(C.c_int32)),
        ("material_id", C.POINTER(C.c_int32)),
        ("surface_velocity", C.POINTER(C.c_float)),
        ("ccd_enabled", C.POINTER(C.c_int32)),
        ("linear_motion_interpolation", C.POINTER(C.c_int32)),
        ("static_body", C.POINTER(C.c_int32)),
        ("collision_groups", C.POINTER(C.c_uint32)),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("step", C.c_double),
        ("collision_detection", C.c_double),
        ("broad_phase", C.c_double),
        ("narrow_phase", C.c_double),
        ("assembly", C.c_double),
        ("velocity_resolution", C.c_double),
        ("velocity_update", C.c_double),
        ("position_resolution", C.c_double),
        ("nsteps", C.c_int64),
        ("npairs", C.c_int64),
        ("ncontacts", C.c_int64),
        ("nmanifolds", C.c_int64),
        ("nconstraints", C.c_int64),
        ("ncolours", C.c_int64),
        ("colour_rounds", C.c_int64),
        ("sweep_kernel_time", C.c_double),
        ("sweep_kernel_launches", C.c_int64),
        ("sweep_kernel_contacts", C.c_int64),
        ("velocity_sweeps", C.c_int64),
        ("judge_plain_sweeps", C.c_int64),
        ("judge_learned_sweeps", C.c_int64),
        ("judge_kept", C.c_int64),
        ("sweep_variant", C.c_int64),
        ("loose_changes", C.c_int64),
        ("pair_rebuilds", C.c_int64),
        ("incremental_pair_updates", C.c_int64),
        ("active_bodies", C.c_int64),
        ("incremental_overflows", C.c_int64),
        ("warmstart_inference_time", C.c_double),
        ("warmstart_inference_rows", C.c_int64),
        ("ccd_time", C.c_double),
        ("ccd_substeps", C.c_int64),
        ("ccd_impacts", C.c_int64),
    ]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def u32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def u64ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


class JointDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("kind", C.POINTER(C.c_int32)),
        ("body1", C.POINTER(C.c_int32)),
        ("body2", C.POINTER(C.c_int32)),
        ("anchor1", C.POINTER(C.c_float)),
        ("anchor2", C.POINTER(C.c_float)),
        ("frame1", C.POINTER(C.c_float)),
        ("frame2", C.POINTER(C.c_float)),
        ("axis1", C.POINTER(C.c_float)),
        ("axis2", C.POINTER(C.c_float)),
        ("break_force", C.POINTER(C.c_float)),
        ("break_torque", C.POINTER(C.c_float)),
    ]


def joints(kind, body1, body2, anchor1, anchor2, frame1=None, frame2=None, axis1=None, axis2=None,
           break_force=None, break_torque=None):
    """A joint set as a dict of arrays (the fields of nphys_joint_desc); scalars broadcast."""
    body1 = np.atleast_1d(np.asarray(body1, np.int32))
    n = body1.shape[0]

    def arr(v, dtype, width):
        if v is None:
            return None
        a = np.asarray(v, dtype)
        return np.ascontiguousarray(np.broadcast_to(a, (n, width) if width > 1 else (n,)))

    out = {"kind": arr(kind, np.int32, 1), "body1": body1, "body2": arr(body2, np.int32, 1),
           "anchor1": arr(anchor1, np.float32, 3), "anchor2": arr(anchor2, np.float32, 3)}
    for name, v, w in (("frame1", frame1, 4), ("frame2", frame2, 4), ("axis1", axis1, 3), ("axis2", axis2, 3),
                       ("break_force", break_force, 1), ("break_torque", break_torque, 1)):
        a = arr(v, np.float32, w)
        if a is not None:
            out[name] = a
    return out


def make_joint_desc(js):
    """Builds a JointDesc over a joints() dict; returns (desc, keepalive)."""
    keep = {}
    d = JointDesc()
    d.n = int(js["body1"].shape[0])
    for name in ("kind", "body1", "body2"):
        keep[name] = np.ascontiguousarray(js[name], np.int32).reshape(-1)
        setattr(d, name, iptr(keep[name]))
    for name in ("anchor1", "anchor2", "frame1", "frame2", "axis1", "axis2", "break_force", "break_torque"):
        if js.get(name) is not None:
            keep[name] = np.ascontiguousarray(js[name], np.float32).reshape(-1)
            setattr(d, name, fptr(keep[name]))
    return d, keep


def make_body_desc(scene):
    """Builds a BodyDesc over the arrays of a scene dict; returns (desc, keepalive)."""
    keep = {}
    n = int(scene["status"].shape[0])

    def arr(name, dtype, width):
        a = np.ascontiguousarray(scene[name], dtype=dtype).reshape(n * width if width > 1 else n)
        keep[name] = a
        return a

    d = BodyDesc()
    d.n = n
    d.status = iptr(arr("status", np.int32, 1))
    d.shape = iptr(arr("shape", np.int32, 1))
    d.dims = fptr(arr("dims", np.float32, 3))
    d.pos = fptr(arr("pos", np.float32, 3))
    d.rot = fptr(arr("rot", np.float32, 4))
    d.linvel = fptr(arr("linvel", np.float32, 3))
    d.angvel = fptr(arr("angvel", np.float32, 3))
    d.density = fptr(arr("density", np.float32, 1))
    d.margin = fptr(arr("margin", np.float32, 1))
    d.linear_prediction = fptr(arr("linear_prediction", np.float32, 1))
    d.friction = fptr(arr("friction", np.float32, 1))
    d.restitution = fptr(arr("restitution", np.float32, 1))
    for name, (dtype, width) in OPTIONAL_BODY_FIELDS.items():
        if name in scene:
            a = arr(name, dtype, width)
            if dtype == np.int32:
                setattr(d, name, iptr(a))
            elif dtype == np.uint32:
                setattr(d, name, a.ctypes.data_as(C.POINTER(C.c_uint32)))
            else:
                setattr(d, name, fptr(a))
    return d, keep
