"""ctypes binding of the CPU restatement (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker and the timed CPU baseline. The product package
(nphysics_amd) never imports this module.

OracleWorld mirrors the product's MechanicalWorld wrapper call for call (add_bodies, step,
collide, pairs, contacts, impulses, residual, counters) with two extra switches: precision
(32 or 64; nphysics is generic over RealField) and the sweep order (the reference's row order,
or the GPU's colour order).
"""
import ctypes as C
import os

import numpy as np

from nphysics_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

ORDER_REFERENCE = 0
ORDER_COLOURED = 1

_lib = None


def build():
    import subprocess

    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_last_error.restype = C.c_char_p
        L.orc_world_create.argtypes = [C.POINTER(abi.Params), C.c_int32, C.POINTER(vp)]
        L.orc_world_destroy.argtypes = [vp]
        L.orc_set_params.argtypes = [vp, C.POINTER(abi.Params)]
        L.orc_set_order_mode.argtypes = [vp, C.c_int32]
        L.orc_set_brute_force_limit.argtypes = [vp, C.c_int32]
        L.orc_set_residual_target.argtypes = [vp, C.c_double, C.c_int32, C.c_int32, C.c_int32]
        L.orc_add_bodies.argtypes = [vp, C.POINTER(abi.BodyDesc), C.POINTER(C.c_int32)]
        L.orc_num_bodies.argtypes = [vp, C.POINTER(C.c_int32)]
        L.orc_add_joints.argtypes = [vp, C.POINTER(abi.JointDesc), C.POINTER(C.c_int32)]
        L.orc_remove_joint.argtypes = [vp, C.c_int32]
        L.orc_read_joints.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                      C.c_int64]
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        L.orc_set_state.argtypes = [vp, fp, fp, fp, fp]
        L.orc_set_collision_groups.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]
        L.orc_set_state_f64.argtypes = [vp, dp, dp, dp, dp]
        L.orc_get_state.argtypes = [vp, dp, dp, dp, dp]
        L.orc_step.argtypes = [vp, C.c_int32]
        L.orc_collide.argtypes = [vp]
        L.orc_num_pairs.argtypes = [vp, C.POINTER(C.c_int64)]
        L.orc_read_pairs.argtypes = [vp, C.POINTER(C.c_int32), C.c_int64]
        L.orc_num_contacts.argtypes = [vp, C.POINTER(C.c_int64)]
        ip = C.POINTER(C.c_int32)
        L.orc_read_contacts.argtypes = [vp, ip, C.POINTER(C.c_uint32), dp, dp, dp, dp, ip, dp, dp, dp, C.c_int64]
        L.orc_num_solved_contacts.argtypes = [vp, C.POINTER(C.c_int64)]
        L.orc_read_impulses.argtypes = [vp, C.POINTER(C.c_uint64), dp, C.c_int64]
        L.orc_residual.argtypes = [vp, dp, dp]
        L.orc_read_counters.argtypes = [vp, C.POINTER(abi.Counters)]
        L.orc_reset_counters.argtypes = [vp]
        L.orc_read_colours.argtypes = [vp, ip, C.c_int64, ip]
        L.orc_read_activation.argtypes = [vp, dp, C.c_int64]
        L.orc_num_manifolds.argtypes = [vp, C.POINTER(C.c_int64)]
        L.orc_read_manifolds.argtypes = [vp, ip, ip, ip, C.c_int64]
        L.orc_read_body_dynamics.argtypes = [vp, dp, dp, dp]
        L.orc_collide_shapes.argtypes = [C.c_int32, C.c_int32, dp, dp, C.c_int32, dp, dp, C.c_double,
                                         C.POINTER(C.c_uint32), dp, dp, dp, dp, ip]
        L.orc_colour_manifolds.argtypes = [C.c_int32, ip, ip, C.c_int32, ip, ip]
        L.orc_time_of_impact.argtypes = [C.c_int32, C.c_int32, dp, dp, C.c_int32, dp, dp, C.c_double, C.c_double,
                                         dp, ip]
        L.orc_apply_force.argtypes = [vp, C.c_int32, ip, fp, fp, C.c_int32, C.c_int32]
        L.orc_set_material_coefficient.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_float]
        L.orc_mlp_forward.argtypes = [dp, fp, C.c_int64, C.c_int32, dp]
        L.orc_mlp_gradients.argtypes = [dp, fp, fp, C.c_int64, C.c_int32, dp, dp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


class OracleWorld:
    def __init__(self, params=None, precision=32, order=ORDER_REFERENCE):
        self._L = lib()
        self._h = C.c_void_p()
        self.params = params if params is not None else abi.default_params()
        _check(self._L.orc_world_create(C.byref(self.params), precision, C.byref(self._h)))
        self.precision = precision
        self.set_order(order)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.orc_world_destroy(h)
            self._h = None

    def set_order(self, order):
        _check(self._L.orc_set_order_mode(self._h, order))

    def set_params(self, params):
        self.params = params
        _check(self._L.orc_set_params(self._h, C.byref(params)))

    def set_residual_target(self, target=1e-4, max_sweeps=500, check_every=1, norm="max"):
        _check(self._L.orc_set_residual_target(self._h, target, max_sweeps, check_every, 0 if norm == "max" else 1))

    def set_brute_force_limit(self, n):
        _check(self._L.orc_set_brute_force_limit(self._h, n))

    def add_bodies(self, scene):
        desc, keep = abi.make_body_desc(scene)
        first = C.c_int32()
        _check(self._L.orc_add_bodies(self._h, C.byref(desc), C.byref(first)))
        return first.value

    @property
    def num_bodies(self):
        n = C.c_int32()
        _check(self._L.orc_num_bodies(self._h, C.byref(n)))
        return n.value

    def add_joints(self, js):
        desc, keep = abi.make_joint_desc(js)
        first = C.c_int32()
        _check(self._L.orc_add_joints(self._h, C.byref(desc), C.byref(first)))
        self._njoints = getattr(self, "_njoints", 0) + desc.n
        return first.value

    def remove_joint(self, j):
        _check(self._L.orc_remove_joint(self._h, j))

    def joints(self):
        """Cached impulses (lin (n, 3), ang (n, 3)) and broken flags of every joint."""
        n = getattr(self, "_njoints", 0)
        lin = np.zeros((max(n, 1), 3))
        ang = np.zeros((max(n, 1), 3))
        br = np.zeros(max(n, 1), np.int32)
        _check(self._L.orc_read_joints(self._h, lin.ctypes.data_as(C.POINTER(C.c_double)),
                                       ang.ctypes.data_as(C.POINTER(C.c_double)),
                                       br.ctypes.data_as(C.POINTER(C.c_int32)), max(n, 1)))
        return {"lin": lin[:n], "ang": ang[:n], "broken": br[:n].astype(bool)}

    def set_collision_groups(self, bodies, groups):
        b = np.ascontiguousarray(bodies, np.int32).reshape(-1)
        g = np.ascontiguousarray(groups, np.uint32).reshape(-1, 3)
        _check(self._L.orc_set_collision_groups(self._h, int(b.shape[0]), abi.iptr(b),
                                                g.ctypes.data_as(C.POINTER(C.c_uint32))))

    def set_state(self, pos=None, rot=None, linvel=None, angvel=None):
        if all(a is None or np.asarray(a).dtype == np.float64 for a in (pos, rot, linvel, angvel)):
            arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (pos, rot, linvel, angvel)]
            ptrs = [None if a is None else abi.dptr(a) for a in arrs]
            _check(self._L.orc_set_state_f64(self._h, *ptrs))
        else:
            arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (pos, rot, linvel, angvel)]
            ptrs = [None if a is None else abi.fptr(a) for a in arrs]
            _check(self._L.orc_set_state(self._h, *ptrs))

    def get_state(self):
        n = self.num_bodies
        pos = np.zeros((n, 3))
        rot = np.zeros((n, 4))
        lv = np.zeros((n, 3))
        av = np.zeros((n, 3))
        _check(self._L.orc_get_state(self._h, abi.dptr(pos), abi.dptr(rot), abi.dptr(lv), abi.dptr(av)))
        return {"pos": pos, "rot": rot, "linvel": lv, "angvel": av}

    def step(self, n=1):
        _check(self._L.orc_step(self._h, n))

    def apply_force(self, bodies, forces, force_type=0, auto_wake_up=True, points=None):
        """Body::apply_force per entry (forces (k, 6): linear, torque), or apply_force_at_point
        with `points` (forces (k, 3) linear at world points (k, 3))."""
        b = np.ascontiguousarray(bodies, np.int32).reshape(-1)
        f = np.ascontiguousarray(forces, np.float32).reshape(-1)
        p = None if points is None else np.ascontiguousarray(points, np.float32).reshape(-1)
        _check(self._L.orc_apply_force(self._h, b.shape[0], abi.iptr(b), abi.fptr(f),
                                       None if p is None else abi.fptr(p), force_type, int(auto_wake_up)))

    def set_material_coefficient(self, which, id1, id2, coeff):
        _check(self._L.orc_set_material_coefficient(self._h, which, id1, id2, coeff))

    def collide(self):
        _check(self._L.orc_collide(self._h))

    def pairs(self):
        n = C.c_int64()
        _check(self._L.orc_num_pairs(self._h, C.byref(n)))
        out = np.zeros((n.value, 2), np.int32)
        _check(self._L.orc_read_pairs(self._h, abi.iptr(out), n.value))
        return out

    def contacts(self):
        n = C.c_int64()
        _check(self._L.orc_num_contacts(self._h, C.byref(n)))
        c = n.value
        d = {
            "bodies": np.zeros((c, 2), np.int32),
            "feature": np.zeros(c, np.uint32),
            "world1": np.zeros((c, 3)),
            "world2": np.zeros((c, 3)),
            "normal": np.zeros((c, 3)),
            "depth": np.zeros(c),
            "kind": np.zeros(c, np.int32),
            "lp1": np.zeros((c, 3)),
            "lp2": np.zeros((c, 3)),
            "ln": np.zeros((c, 3)),
        }
        _check(self._L.orc_read_contacts(
            self._h, abi.iptr(d["bodies"]), abi.u32ptr(d["feature"]), abi.dptr(d["world1"]),
            abi.dptr(d["world2"]), abi.dptr(d["normal"]), abi.dptr(d["depth"]), abi.iptr(d["kind"]),
            abi.dptr(d["lp1"]), abi.dptr(d["lp2"]), abi.dptr(d["ln"]), c))
        return d

    def impulses(self):
        n = C.c_int64()
        _check(self._L.orc_num_solved_contacts(self._h, C.byref(n)))
        keys = np.zeros(n.value, np.uint64)
        imp = np.zeros((n.value, 3))
        _check(self._L.orc_read_impulses(self._h, abi.u64ptr(keys), abi.dptr(imp), n.value))
        return keys, imp

    def residual(self):
        a, b = C.c_double(), C.c_double()
        _check(self._L.orc_residual(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def counters(self):
        c = abi.Counters()
        _check(self._L.orc_read_counters(self._h, C.byref(c)))
        return c.as_dict()

    def reset_counters(self):
        _check(self._L.orc_reset_counters(self._h))

    def manifolds(self):
        """Contact manifolds: bodies (M, 2), first contact index, contact count."""
        n = C.c_int64()
        _check(self._L.orc_num_manifolds(self._h, C.byref(n)))
        m = n.value
        b = np.zeros((max(m, 1), 2), np.int32)
        f = np.zeros(max(m, 1), np.int32)
        k = np.zeros(max(m, 1), np.int32)
        _check(self._L.orc_read_manifolds(self._h, abi.iptr(b), abi.iptr(f), abi.iptr(k), max(m, 1)))
        return {"bodies": b[:m], "first": f[:m], "count": k[:m]}

    def activation(self):
        """Per-body activation energy (0: asleep)."""
        e = np.zeros(max(1, self.num_bodies))
        _check(self._L.orc_read_activation(self._h, abi.dptr(e), e.shape[0]))
        return e[: self.num_bodies]

    def colours(self):
        c = self.counters()
        out = np.zeros(max(1, c["nmanifolds"]), np.int32)
        nc = C.c_int32()
        _check(self._L.orc_read_colours(self._h, abi.iptr(out), out.shape[0], C.byref(nc)))
        return out[: c["nmanifolds"]], nc.value

    def body_dynamics(self):
        n = self.num_bodies
        im = np.zeros(n)
        ia = np.zeros((n, 3, 3))
        bb = np.zeros((n, 6))
        _check(self._L.orc_read_body_dynamics(self._h, abi.dptr(im), abi.dptr(ia), abi.dptr(bb)))
        return im, ia, bb


def collide_shapes(kind1, dims1, pose1, kind2, dims2, pose2, prediction, precision=64):
    """Narrow phase of one pair; pose = (x, y, z, i, j, k, w). Returns a dict of arrays."""
    L = lib()
    d1 = np.ascontiguousarray(dims1, np.float64)
    d2 = np.ascontiguousarray(dims2, np.float64)
    p1 = np.ascontiguousarray(pose1, np.float64)
    p2 = np.ascontiguousarray(pose2, np.float64)
    feat = np.zeros(8, np.uint32)
    w1 = np.zeros((8, 3))
    w2 = np.zeros((8, 3))
    nn = np.zeros((8, 3))
    dep = np.zeros(8)
    kind = np.zeros(8, np.int32)
    nc = L.orc_collide_shapes(precision, kind1, abi.dptr(d1), abi.dptr(p1), kind2, abi.dptr(d2),
                              abi.dptr(p2), prediction, abi.u32ptr(feat), abi.dptr(w1), abi.dptr(w2),
                              abi.dptr(nn), abi.dptr(dep), abi.iptr(kind))
    if nc < 0:
        raise RuntimeError("collide_shapes failed")
    return {"n": nc, "feature": feat[:nc], "world1": w1[:nc], "world2": w2[:nc], "normal": nn[:nc],
            "depth": dep[:nc], "kind": kind[:nc]}


def time_of_impact(kind1, dims1, motion1, kind2, dims2, motion2, max_toi, target, precision=64):
    """Conservative-advancement time of impact of two shapes. motion = dict(kind=0 static / 1 linear /
    2 rotating, t0, pose=(x, y, z, i, j, k, w), v, w). Returns (toi, status) or None."""

    def mot(m):
        return np.array([m.get("kind", 2), m.get("t0", 0.0), *m["pose"], *m.get("v", (0, 0, 0)),
                         *m.get("w", (0, 0, 0))], np.float64)

    d1 = np.ascontiguousarray(dims1, np.float64)
    d2 = np.ascontiguousarray(dims2, np.float64)
    m1, m2 = mot(motion1), mot(motion2)
    toi = np.zeros(1)
    st = np.zeros(1, np.int32)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    found = lib().orc_time_of_impact(int(precision), int(kind1), d1.ctypes.data_as(dp), m1.ctypes.data_as(dp),
                                     int(kind2), d2.ctypes.data_as(dp), m2.ctypes.data_as(dp), float(max_toi),
                                     float(target), toi.ctypes.data_as(dp), st.ctypes.data_as(ip))
    return (float(toi[0]), int(st[0])) if found else None


def colour_manifolds(b1, b2, is_dynamic, initial=None):
    """Greedy colouring; entries >= 0 of `initial` are kept as given."""
    L = lib()
    b1 = np.ascontiguousarray(b1, np.int32)
    b2 = np.ascontiguousarray(b2, np.int32)
    dyn = np.ascontiguousarray(is_dynamic, np.int32)
    out = np.full(b1.shape[0], -1, np.int32) if initial is None else np.array(initial, np.int32)
    nc = L.orc_colour_manifolds(b1.shape[0], abi.iptr(b1), abi.iptr(b2), dyn.shape[0], abi.iptr(dyn), abi.iptr(out))
    return out, nc


MLP_NUM_TRAINED = 20099
MLP_NUM_PARAMS = 20150


def mlp_forward(params, X, emulate_bf16=True):
    """Network outputs (impulses) for rows X[n][24]; params in the C ABI layout."""
    L = lib()
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(X, np.float32)
    out = np.zeros((x.shape[0], 3))
    _check(L.orc_mlp_forward(abi.dptr(p), abi.fptr(x), x.shape[0], int(emulate_bf16), abi.dptr(out)))
    return out


def mlp_gradients(params, X, T, emulate_bf16=True):
    """(gradient over the 20,099 trained parameters, loss) of the batch X, T."""
    L = lib()
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(X, np.float32)
    t = np.ascontiguousarray(T, np.float32)
    g = np.zeros(MLP_NUM_TRAINED)
    loss = C.c_double()
    _check(L.orc_mlp_gradients(abi.dptr(p), abi.fptr(x), abi.fptr(t), x.shape[0], int(emulate_bf16), abi.dptr(g),
                               C.byref(loss)))
    return g, loss.value
