"""GPU parity of the rigid-body and material attributes: damping and velocity clamps, the gravity
flag, the jacobian mask (kinematic dofs), material combine modes and the coefficient table,
surface velocities, and external forces of every ForceType applied between steps.

The HIP pipeline is compared bit for bit with the FP32 oracle run in the GPU's colour order on a
small pile whose bodies carry seeded random attributes (rigid_body.rs:467-505, 590-619, 672-722,
771-845; material.rs:72-86, 134-178). The oracle functions themselves are pinned by closed forms
in tests/test_oracle_pins.py.
"""
import numpy as np
import pytest

from nphysics_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32)


def attributed_pile(seed=11, n=216, masks=True, surface=True, table=True):
    s = scenes.small_pile(seed=seed, n=n, lattice=(6, 6, 6), extent=4.0)
    N = s["status"].shape[0]
    rng = np.random.default_rng(seed)
    dyn = s["status"] == 0
    s["linear_damping"] = np.where(rng.random(N) < 0.3, rng.uniform(0.0, 2.0, N), 0.0).astype(np.float32)
    s["angular_damping"] = np.where(rng.random(N) < 0.3, rng.uniform(0.0, 2.0, N), 0.0).astype(np.float32)
    fmax = np.finfo(np.float32).max
    s["max_linear_velocity"] = np.where(rng.random(N) < 0.2, rng.uniform(0.0, 4.0, N), fmax).astype(np.float32)
    s["max_angular_velocity"] = np.where(rng.random(N) < 0.2, rng.uniform(0.0, 3.0, N), fmax).astype(np.float32)
    s["gravity_enabled"] = (rng.random(N) > 0.05).astype(np.int32)
    if masks:
        choices = np.array([0, 0b000010, 0b111000, 0b001000, 0b100101, 0b000001], np.int32)
        s["kinematic_dofs"] = np.where(rng.random(N) < 0.3, choices[rng.integers(0, len(choices), N)], 0).astype(np.int32)
    s["friction"] = rng.uniform(0.1, 0.9, N).astype(np.float32)
    s["restitution"] = rng.uniform(0.0, 0.6, N).astype(np.float32)
    s["friction_combine"] = rng.integers(0, 4, N).astype(np.int32)
    s["restitution_combine"] = rng.integers(0, 4, N).astype(np.int32)
    if table:
        s["material_id"] = np.where(rng.random(N) < 0.6, rng.integers(0, 4, N), -1).astype(np.int32)
    if surface:
        sv = np.zeros((N, 3), np.float32)
        sv[~dyn] = rng.uniform(-1.0, 1.0, (int((~dyn).sum()), 3)).astype(np.float32)
        s["surface_velocity"] = sv
    return s


def _worlds(nphys, oracle_mod, s, table=True, **params):
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params(**params))
    g.add_bodies(s)
    o = oracle_mod.OracleWorld(abi.default_params(**params), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(s)
    if table:
        for which, a, b, v in ((0, 0, 1, 0.05), (0, 2, 2, 1.2), (1, 1, 3, 0.7), (0, 3, 0, 0.33), (1, 2, 0, 0.0)):
            g.set_material_coefficient(which, a, b, v)
            o.set_material_coefficient(which, a, b, v)
    return g, o


def _assert_same(g, o, what=""):
    gs, os_ = g.get_state(), o.get_state()
    for k in ("pos", "rot", "linvel", "angvel"):
        np.testing.assert_array_equal(gs[k], _f32(os_[k]), err_msg=f"{what} {k}")


@pytest.mark.parametrize("masks", [False, True])
def test_attributes_bitexact(nphys, oracle_mod, masks):
    s = attributed_pile(masks=masks)
    g, o = _worlds(nphys, oracle_mod, s)
    for k in range(40):
        g.step(1)
        o.step(1)
        print(f"masks={masks} step {k + 1}", flush=True)
        _assert_same(g, o, f"step {k + 1}")
    gk, gi = g.impulses()
    ok, oi = o.impulses()
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gi, _f32(oi))
    assert g.residual()[0] == pytest.approx(o.residual()[0], rel=1e-12, abs=0.0)


@pytest.mark.parametrize("variant", [0, 256, 1024])
def test_masked_sweep_variants_identical(nphys, oracle_mod, variant):
    """The masked kernels in every launch shape (per-colour, scene-blocked) give the same bits."""
    s = attributed_pile(seed=12)
    g, o = _worlds(nphys, oracle_mod, s)
    g.set_sweep_variant(variant)
    g.step(25)
    o.step(25)
    _assert_same(g, o, f"variant {variant}")


def test_forces_between_steps_bitexact(nphys, oracle_mod):
    """Forces, impulses, acceleration and velocity changes (plain and at points, repeated bodies,
    kinematic dofs), applied between steps, including after a set_state of the same interval."""
    s = attributed_pile(seed=13, surface=False)
    g, o = _worlds(nphys, oracle_mod, s)
    N = s["status"].shape[0]
    rng = np.random.default_rng(5)
    for k in range(30):
        if k % 3 == 0:
            for t in range(4):
                m = 20
                bodies = rng.integers(0, N, m).astype(np.int32)
                if t % 2 == 0:
                    f = rng.normal(0.0, 2.0, (m, 6)).astype(np.float32)
                    g.apply_force(bodies, f, t, auto_wake_up=bool(k % 2))
                    o.apply_force(bodies, f, t, auto_wake_up=bool(k % 2))
                else:
                    f = rng.normal(0.0, 2.0, (m, 3)).astype(np.float32)
                    pts = (s["pos"][bodies] + rng.normal(0.0, 0.3, (m, 3))).astype(np.float32)
                    g.apply_force(bodies, f, t, points=pts)
                    o.apply_force(bodies, f, t, points=pts)
        if k == 10:
            st = g.get_state()
            rot = st["rot"].copy()
            rot[5] = np.array([0.0, 0.0, np.sin(0.3), np.cos(0.3)], np.float32)
            g.set_state(rot=rot)
            o.set_state(rot=rot)
            f = np.array([[0.5, 0.2, -0.1, 0.3, 0.1, 0.4]], np.float32)
            g.apply_force([5], f, abi.IMPULSE)
            o.apply_force([5], f, abi.IMPULSE)
        g.step(1)
        o.step(1)
        print(f"forces step {k + 1}", flush=True)
        _assert_same(g, o, f"step {k + 1}")


def test_forces_wake_sleeping_bodies(nphys, oracle_mod):
    """Sleeping bodies woken by a force with auto_wake_up, and by an impulse without it."""
    # a pile where 10 bodies sleep after 400 steps (CPU restatement, colour order)
    s = scenes.with_sleep_threshold(scenes.small_pile(seed=5, n=64, lattice=(4, 4, 4), extent=4.0), 0.01)
    g = nphys.MechanicalWorld()
    g.add_bodies(s)
    o = oracle_mod.OracleWorld(abi.default_params(), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(s)
    g.step(400)
    o.step(400)
    _assert_same(g, o, "settled")
    asleep = np.flatnonzero((g.activation() == 0.0) & (s["status"] == 0))
    assert asleep.size > 0
    b = asleep[:2].astype(np.int32)
    f = np.array([[0.0, 3.0, 0.0, 0.0, 0.0, 0.0]] * len(b), np.float32)
    g.apply_force(b[:1], f[:1], abi.FORCE, auto_wake_up=True)
    o.apply_force(b[:1], f[:1], abi.FORCE, auto_wake_up=True)
    g.apply_force(b[1:], f[1:], abi.IMPULSE, auto_wake_up=False)
    o.apply_force(b[1:], f[1:], abi.IMPULSE, auto_wake_up=False)
    for k in range(20):
        g.step(1)
        o.step(1)
        _assert_same(g, o, f"after wake, step {k + 1}")
    np.testing.assert_array_equal(g.activation(), _f32(o.activation()))
