"""The learned warm start: a network that corrects each contact's cached initial impulses.

WarmStartNet wraps the C ABI's nphys_mlp (include/nphys.h): 24 contact features -> 128 -> 128 ->
3 impulse corrections (normal, tangent 1, tangent 2) in units of r_n |g| dt added to the cached
impulses (0 for a new contact), GELU, BF16 weights and
activations with FP32 accumulation on MFMA. It is trained on the GPU by the library's own
forward/backward kernels and Adam, on samples a MechanicalWorld records (its features before the
solve and its impulses after a solve to a tight residual), and loaded into a world with
MechanicalWorld.load_warmstart. The plain cache warm start of the reference
(signorini_coulomb_pyramid_model.rs:104-108) stays the fallback and the judge.
"""
import ctypes as C

import numpy as np

from . import abi
from .world import NphysError, load_library

NUM_FEATURES = 24
NUM_PARAMS = 20150
NUM_TRAINED = 20099

FEATURES = (
    "rhs_n/(g dt)", "rhs_t1/(g dt)", "rhs_t2/(g dt)",      # relative velocity in the contact frame
    "n_x", "n_y", "n_z",                                    # normal (body 1 -> body 2)
    "depth_eff/(g dt^2)",                                   # dilated penetration depth
    "m1^-1 r_n", "m2^-1 r_n",                               # inverse masses over the normal row's
    "ang1_n r_n", "ang1_t1 r_t1", "ang1_t2 r_t2",           # (r x d)^T I^-1 (r x d) per body and row,
    "ang2_n r_n", "ang2_t1 r_t1", "ang2_t2 r_t2",           #   each over the row's inverse effective mass
    "manifold size/4",
    "cache_n/ls", "cache_t1/ls", "cache_t2/ls",             # previous impulses, ls = r_n |g| dt
    "mu", "restitution", "|g| dt", "is_ground", "has_cache",
)


def _declare(L):
    if getattr(L, "_mlp_declared", False):
        return
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    L.nphys_mlp_create.argtypes = [C.c_uint64, C.POINTER(vp)]
    L.nphys_mlp_destroy.argtypes = [vp]
    L.nphys_mlp_num_params.argtypes = [C.POINTER(C.c_int64)]
    L.nphys_mlp_get_params.argtypes = [vp, fp, C.c_int64]
    L.nphys_mlp_set_params.argtypes = [vp, fp, C.c_int64]
    L.nphys_mlp_set_data.argtypes = [vp, fp, fp, C.c_int64, C.c_int32]
    L.nphys_mlp_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, fp]
    L.nphys_mlp_forward.argtypes = [vp, fp, C.c_int64, fp]
    L.nphys_mlp_gradients.argtypes = [vp, fp, fp, C.c_int64, fp, C.POINTER(C.c_double)]
    L._mlp_declared = True


def _check(L, rc):
    if rc != 0:
        raise NphysError(f"nphys error {rc}: {L.nphys_last_error().decode()}")


class WarmStartNet:
    def __init__(self, seed=1):
        self._L = load_library()
        _declare(self._L)
        h = C.c_void_p()
        _check(self._L, self._L.nphys_mlp_create(seed, C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            self._L.nphys_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.zeros(NUM_PARAMS, np.float32)
        _check(self._L, self._L.nphys_mlp_get_params(self._h, abi.fptr(out), out.shape[0]))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(self._L, self._L.nphys_mlp_set_params(self._h, abi.fptr(p), p.shape[0]))

    def set_data(self, features, targets, set_normalisation=True):
        x = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        if x.ndim != 2 or x.shape[1] != NUM_FEATURES or t.shape != (x.shape[0], 3):
            raise ValueError("features must be (n, 24) and targets (n, 3)")
        self._data = (x, t)  # the library copies; keep nothing else alive
        _check(self._L, self._L.nphys_mlp_set_data(self._h, abi.fptr(x), abi.fptr(t), x.shape[0],
                                                   1 if set_normalisation else 0))

    def train(self, epochs, batch=4096, lr=1e-3):
        losses = np.zeros(max(1, epochs), np.float32)
        _check(self._L, self._L.nphys_mlp_train(self._h, epochs, batch, lr, abi.fptr(losses)))
        return losses[:epochs]

    def forward(self, features):
        x = np.ascontiguousarray(features, np.float32)
        out = np.zeros((x.shape[0], 3), np.float32)
        _check(self._L, self._L.nphys_mlp_forward(self._h, abi.fptr(x), x.shape[0], abi.fptr(out)))
        return out

    def gradients(self, features, targets):
        x = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(NUM_TRAINED, np.float32)
        loss = C.c_double()
        _check(self._L, self._L.nphys_mlp_gradients(self._h, abi.fptr(x), abi.fptr(t), x.shape[0], abi.fptr(g),
                                                    C.byref(loss)))
        return g, loss.value


CACHE = slice(16, 19)  # features holding the cached impulses


def correction_targets(features, impulses):
    """Training targets of the network: solved impulses minus the cached ones (same unit)."""
    return impulses - features[:, CACHE]


def collect_samples(world, steps, target=1e-6, max_sweeps=500, check_every=4, norm="max"):
    """Steps `world` with its solve run to a tight residual target and returns the recorded
    (features, impulses) of every step: the impulses the plain solver converged to (or reached at
    max_sweeps). The world's residual mode is restored to the fixed count after."""
    world.set_residual_target(target, max_sweeps, check_every, norm)
    world.capture_samples(True)
    Xs, Ts = [], []
    for _ in range(steps):
        world.step(1)
        X, T = world.samples()
        Xs.append(X)
        Ts.append(T)
    world.capture_samples(False)
    world.set_residual_target(0.0)
    return np.concatenate(Xs), np.concatenate(Ts)


def judge(world, params, steps, target=1e-4, max_sweeps=200, check_every=2, norm="rms"):
    """Runs `steps` steps of `world` with the learned warm start under the judge (both solves from
    the same assembled system each step, the learned one kept only if it reached the target in
    fewer sweeps) and returns the judge's counters and the mean sweeps of each start."""
    world.load_warmstart(params)
    world.set_contact_model(abi.CONTACT_MODEL_LEARNED_WARMSTART)
    world.set_residual_target(target, max_sweeps, check_every, norm)
    world.set_warmstart_judge(True)
    world.reset_counters()
    world.step(steps)
    c = world.counters()
    world.set_warmstart_judge(False)
    world.set_contact_model(abi.CONTACT_MODEL_SIGNORINI_COULOMB_PYRAMID)
    world.set_residual_target(0.0)
    plain = c["judge_plain_sweeps"] / steps
    learned = c["judge_learned_sweeps"] / steps
    return {
        "steps": steps,
        "target": target,
        "norm": norm,
        "max_sweeps": max_sweeps,
        "check_every": check_every,
        "plain_sweeps_per_step": plain,
        "learned_sweeps_per_step": learned,
        "sweep_reduction": (plain - learned) / plain if plain > 0 else 0.0,
        "learned_kept_steps": c["judge_kept"],
    }
