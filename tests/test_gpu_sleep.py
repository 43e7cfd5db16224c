"""GPU parity of sleeping / activation islands against the FP32 oracle in the GPU's colour order.

Energies, states, impulses of the solved contacts and the awake-body count must match bit for
bit while islands fall asleep, wake on contact and wake after the caller sets the state.
"""
import numpy as np
import pytest

from nphysics_amd import abi, scenes

from test_oracle_kat import single

pytestmark = pytest.mark.gpu


def _pair(nphys, oracle_mod, scene):
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params())
    g.add_bodies(scene)
    o = oracle_mod.OracleWorld(abi.default_params(), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(scene)
    return g, o


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32)


def _assert_same(g, o, what):
    gs, os_ = g.get_state(), o.get_state()
    for key in ("pos", "rot", "linvel", "angvel"):
        np.testing.assert_array_equal(gs[key], _f32(os_[key]), err_msg=f"{what}: {key}")
    np.testing.assert_array_equal(g.activation(), _f32(o.activation()), err_msg=f"{what}: energy")
    gc, oc = g.counters(), o.counters()
    assert gc["active_bodies"] == oc["active_bodies"], what
    assert gc["nmanifolds"] == oc["nmanifolds"] and gc["ncontacts"] == oc["ncontacts"], what
    gk, gi = g.impulses()
    ok, oi = o.impulses()
    order_g, order_o = np.argsort(gk, kind="stable"), np.argsort(ok, kind="stable")
    np.testing.assert_array_equal(gk[order_g], ok[order_o], err_msg=f"{what}: impulse keys")
    np.testing.assert_array_equal(gi[order_g], _f32(oi)[order_o], err_msg=f"{what}: impulses")


def _mixed_threshold(scene, seed, never=16):
    s = scenes.with_sleep_threshold(scene, 0.01)
    dyn = np.where(s["status"] == 0)[0]
    rng = np.random.default_rng(seed)
    s["sleep_threshold"][rng.choice(dyn, min(never, dyn.shape[0]), replace=False)] = -1.0
    return s


@pytest.mark.parametrize("case", ["pile", "batched"])
def test_sleeping_islands_bitexact(nphys, oracle_mod, case):
    """A settling pile and 32 independent worlds with the default threshold (a few bodies that
    never sleep): islands fall asleep, the solver only sees awake manifolds, sleeping contacts
    keep their cached impulses; every 25 steps the two trajectories are identical."""
    s = (_mixed_threshold(scenes.small_pile(seed=31, n=512), 5) if case == "pile"
         else _mixed_threshold(scenes.batched(worlds=32, seed=11), 6, never=40))
    g, o = _pair(nphys, oracle_mod, s)
    slept = 0
    for chunk in range(12):
        g.step(25)
        o.step(25)
        c = g.counters()
        slept = max(slept, int((s["status"] == 0).sum()) - c["active_bodies"])
        print(f"{case} step {25 * (chunk + 1)}: awake {c['active_bodies']}, solved manifolds {c['nmanifolds']}",
              flush=True)
        _assert_same(g, o, f"{case} step {25 * (chunk + 1)}")
    assert slept > (0 if case == "pile" else 500)


def test_wake_on_contact_and_on_set_state_bitexact(nphys, oracle_mod):
    r = 0.5
    rest = scenes.concat(single(0, 0, (r, 0, 0), (0, r + 0.005, 0)), single(0, 1, (0.4, 0.4, 0.4), (3, 0.405, 0)))
    s = scenes.with_sleep_threshold(scenes.concat(scenes.ground(), rest), 0.01)
    g, o = _pair(nphys, oracle_mod, s)
    g.step(250)
    o.step(250)
    _assert_same(g, o, "asleep")
    assert g.activation()[1] == 0.0 and g.activation()[2] == 0.0
    drop = scenes.with_sleep_threshold(single(0, 0, (r, 0, 0), (0.1, 3 * r + 0.6, 0)), 0.01)
    g.add_bodies(drop)
    o.add_bodies(drop)
    woke = None
    for k in range(90):
        g.step(1)
        o.step(1)
        _assert_same(g, o, f"drop step {k + 1}")
        if woke is None and g.activation()[1] != 0.0:
            woke = k + 1
    assert woke is not None
    assert g.activation()[2] == 0.0  # the box was never touched
    st = g.get_state()
    g.set_state(pos=st["pos"])
    o.set_state(pos=st["pos"])
    for k in range(5):
        g.step(1)
        o.step(1)
        _assert_same(g, o, f"after set_state step {k + 1}")
    assert g.activation()[2] != 0.0


def test_sleeping_long_run_identical_and_deterministic(nphys, oracle_mod):
    """800 steps of 48 worlds falling asleep and waking: two GPU worlds and the oracle stay
    identical. Regression test for the island flags: a body must store its own island root, never
    an older ancestor written by another body's path-halving walk (that made a few islands sleep
    or wake on one run and not on another, a few hundred steps in)."""
    s = _mixed_threshold(scenes.batched(worlds=48, seed=21), 3, never=30)
    g1, o = _pair(nphys, oracle_mod, s)
    g2 = nphys.MechanicalWorld(integration_parameters=abi.default_params())
    g2.add_bodies(s)
    for chunk in range(8):
        g1.step(100)
        g2.step(100)
        o.step(100)
        a, b = g1.get_state(), g2.get_state()
        for key in ("pos", "rot", "linvel", "angvel"):
            np.testing.assert_array_equal(a[key], b[key], err_msg=f"step {100 * (chunk + 1)}: {key}")
        _assert_same(g1, o, f"step {100 * (chunk + 1)}")
        print(f"step {100 * (chunk + 1)}: awake {g1.counters()['active_bodies']}, identical", flush=True)
