"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the restatement still reproduces its frozen vectors (a regression guard on the checker).
GPU: the HIP pipeline reproduces the FP32 vectors bit for bit, with no CPU code run at test time,
and its error against an FP64 run of the same steps is reported and bounded.
"""
import os

import numpy as np
import pytest

from nphysics_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCENES = ("balls64", "boxes27", "pile64")
SCENE_KEYS = ("status", "shape", "dims", "pos", "rot", "linvel", "angvel", "density", "margin",
              "linear_prediction", "friction", "restitution")


def load(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))


def scene_of(f):
    return {k: f[f"in_{k}"] for k in SCENE_KEYS}


def params_of(f):
    return abi.default_params(max_velocity_iterations=int(f["params_velocity_iterations"]),
                              max_position_iterations=int(f["params_position_iterations"]))


@pytest.mark.parametrize("name", SCENES)
def test_oracle_reproduces_golden(oracle_mod, name):
    f = load(name)
    o = oracle_mod.OracleWorld(params_of(f), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(scene_of(f))
    o.step(int(f["steps"]))
    st = o.get_state()
    for k in ("pos", "rot", "linvel", "angvel"):
        assert np.array_equal(st[k].astype(np.float32), f[f"f32_{k}"]), k
    assert np.array_equal(o.pairs(), f["f32_pairs"])
    keys, imp = o.impulses()
    assert np.array_equal(keys, f["f32_impulse_keys"])
    assert np.array_equal(imp.astype(np.float32), f["f32_impulses"])


def test_oracle_shapes_golden(oracle_mod):
    f = load("shapes")
    for i in range(int(f["ncases"])):
        v = f[f"c{i}_in"]
        k1, d1, p1, k2, d2, p2 = int(v[0]), v[1:4], v[4:11], int(v[11]), v[12:15], v[15:22]
        r = oracle_mod.collide_shapes(k1, d1, p1, k2, d2, p2, prediction=0.022, precision=64)
        assert r["n"] == int(f[f"c{i}_n"]), i
        assert np.array_equal(r["feature"], f[f"c{i}_feature"]), i
        for k in ("world1", "world2", "normal", "depth"):
            np.testing.assert_allclose(r[k], f[f"c{i}_{k}"], rtol=0, atol=1e-12, err_msg=f"case {i} {k}")


# Absolute bounds on the FP32 error against the FP64 run in the same order after the fixture's
# steps (tens of steps of stacking and tumbling contact).
POS_TOL = 5e-3   # m
VEL_TOL = 5e-2   # m/s


def errors(f, pos, linvel, ref):
    return (float(np.abs(pos.astype(np.float64) - f[f"{ref}_pos"]).max()),
            float(np.abs(linvel.astype(np.float64) - f[f"{ref}_linvel"]).max()))


def test_golden_fixtures_are_nontrivial():
    """Every fixture exercises contacts; the FP32 error of the coloured order and of the
    reference's order against FP64 stay within the stated bounds."""
    for name in SCENES:
        f = load(name)
        assert f["f32_contact_depth"].shape[0] > 50, name
        assert f["f32_impulses"][:, 0].max() > 0, name
        ep, ev = errors(f, f["f32_pos"], f["f32_linvel"], "f64col")
        rp, rv = errors(f, f["f32ref_pos"], f["f32ref_linvel"], "f64ref")
        assert ep < POS_TOL and ev < VEL_TOL, (name, ep, ev)
        assert rp < POS_TOL and rv < VEL_TOL, (name, rp, rv)


@pytest.mark.gpu
@pytest.mark.parametrize("name", SCENES)
def test_gpu_reproduces_golden(nphys, name):
    f = load(name)
    w = nphys.MechanicalWorld(integration_parameters=params_of(f))
    w.add_bodies(scene_of(f))
    w.step(int(f["steps"]))
    st = w.get_state()
    for k in ("pos", "rot", "linvel", "angvel"):
        assert np.array_equal(st[k], f[f"f32_{k}"]), (name, k)
    assert np.array_equal(w.pairs(), f["f32_pairs"]), name
    c = w.contacts()
    assert np.array_equal(c["bodies"], f["f32_contact_bodies"]), name
    assert np.array_equal(c["feature"], f["f32_contact_feature"]), name
    assert np.array_equal(c["normal"], f["f32_contact_normal"]), name
    assert np.array_equal(c["depth"], f["f32_contact_depth"]), name
    keys, imp = w.impulses()
    assert np.array_equal(keys, f["f32_impulse_keys"]), name
    assert np.array_equal(imp, f["f32_impulses"]), name
    # FP32 error against FP64 at equal iteration count, beside the reference order's own
    ep, ev = errors(f, st["pos"], st["linvel"], "f64col")
    rp, rv = errors(f, f["f32ref_pos"], f["f32ref_linvel"], "f64ref")
    print(f"{name}: max |gpu - f64| pos {ep:.2e} m, linvel {ev:.2e} m/s; reference order in f32: "
          f"pos {rp:.2e} m, linvel {rv:.2e} m/s")
    assert ep < POS_TOL and ev < VEL_TOL, (ep, ev)


@pytest.mark.gpu
@pytest.mark.parametrize("name", SCENES)
def test_gpu_colouring_valid(nphys, oracle_mod, name):
    """No two manifolds of one colour share a dynamic body, and the colours equal the checker's."""
    f = load(name)
    w = nphys.MechanicalWorld(integration_parameters=params_of(f))
    w.add_bodies(scene_of(f))
    w.record_colours()
    w.step(int(f["steps"]))
    colours, ncol = w.colours()
    keys, _ = w.impulses()
    pair_of_key = np.stack([(keys >> np.uint64(40)).astype(np.int64),
                            ((keys >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.int64)], 1)
    _, first = np.unique(pair_of_key, axis=0, return_index=True)
    manifolds = pair_of_key[np.sort(first)]
    assert manifolds.shape[0] == colours.shape[0]
    assert colours.min() >= 0 and colours.max() < ncol
    dyn = f["in_status"] == 0
    for col in range(ncol):
        sel = manifolds[colours == col]
        bodies = np.concatenate([sel[dyn[sel[:, 0]], 0], sel[dyn[sel[:, 1]], 1]])
        assert np.unique(bodies).shape[0] == bodies.shape[0], (name, col)
    o = oracle_mod.OracleWorld(params_of(f), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(scene_of(f))
    o.step(int(f["steps"]))
    oc, onc = o.colours()
    assert onc == ncol and np.array_equal(oc, colours)
