"""Joint constraints on the GPU against the CPU restatement in the colour order, bit for bit:
states, cached joint impulses and broken flags after many steps, in scenes that mix joints and
contacts (a pendulum, chains swinging into the ground, a fixed cantilever, a hinged plate, jointed
pairs in a pile, sleeping chains, joints that break, a removed joint, masked dofs)."""
import math

import numpy as np
import pytest

from nphysics_amd import abi, scenes
from tests.test_oracle_pins import body, merge

pytestmark = pytest.mark.gpu


def _pair(nphys, oracle_mod, s, js, **params):
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params(**params))
    g.add_bodies(s)
    o = oracle_mod.OracleWorld(abi.default_params(**params), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(s)
    if js is not None:
        g.add_joints(js)
        o.add_joints(js)
    return g, o


def _same(g, o, what):
    gs, os_ = g.get_state(), o.get_state()
    for k in ("pos", "rot", "linvel", "angvel"):
        a, b = gs[k], os_[k].astype(np.float32)
        if not np.array_equal(a, b):
            i = int(np.argmax(np.abs(a - b).max(axis=1)))
            raise AssertionError(f"{what}: {k} differs (body {i}: {a[i]} vs {b[i]})")
    gj, oj = g.joints(), o.joints()
    np.testing.assert_array_equal(gj["lin"], oj["lin"].astype(np.float32), err_msg=f"{what}: linear impulses")
    np.testing.assert_array_equal(gj["ang"], oj["ang"].astype(np.float32), err_msg=f"{what}: angular impulses")
    np.testing.assert_array_equal(gj["broken"], oj["broken"], err_msg=f"{what}: broken flags")


def _run(g, o, steps, what, every=10):
    for k in range(steps):
        g.step(1)
        o.step(1)
        if (k + 1) % every == 0 or k + 1 == steps:
            print(f"{what}: step {k + 1}", flush=True)
            _same(g, o, f"{what} step {k + 1}")


def chain_scene(n=6, he=(0.15, 0.3, 0.15), top=(0.0, 4.0, 0.0), tilt=0.6, ground=True):
    """Cuboid links hanging from a static anchor, released at `tilt` rad so they swing (and, with
    a ground, hit it)."""
    parts = [scenes.ground()] if ground else []
    parts.append(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=tuple(np.add(top, (0, 1.0, 0)))))
    d = np.array([math.sin(tilt), -math.cos(tilt), 0.0])
    link = 2 * he[1] + 0.02
    q = (0.0, 0.0, math.sin(tilt / 2), math.cos(tilt / 2))
    for k in range(n):
        parts.append(body(shape=1, dims=he, pos=tuple(np.array(top) + d * (k + 0.5) * link), rot=q))
    s = merge(*parts)
    a = 1 if ground else 0  # index of the static anchor
    b1 = [a] + [a + 1 + k for k in range(n - 1)]
    b2 = [a + 1 + k for k in range(n)]
    a1 = [(0.0, -1.0, 0.0)] + [(0.0, -he[1] - 0.01, 0.0)] * (n - 1)
    a2 = [(0.0, he[1] + 0.01, 0.0)] * n
    return s, b1, b2, a1, a2


@pytest.mark.parametrize("kind", ["ball", "fixed", "revolute", "prismatic", "cylindrical"])
def test_chain_swinging_into_the_ground(nphys, oracle_mod, kind):
    s, b1, b2, a1, a2 = chain_scene(n=6, top=(0.0, 2.5, 0.0))
    k = {"ball": abi.JOINT_BALL, "fixed": abi.JOINT_FIXED, "revolute": abi.JOINT_REVOLUTE,
         "prismatic": abi.JOINT_PRISMATIC, "cylindrical": abi.JOINT_CYLINDRICAL}[kind]
    js = abi.joints(k, b1, b2, a1, a2, axis1=(0, 0, 1), axis2=(0, 0, 1))
    g, o = _pair(nphys, oracle_mod, s, js)
    _run(g, o, 150, f"{kind} chain")
    assert g.counters()["ncontacts"] > 0


def test_pendulum_and_cantilever_and_hinge(nphys, oracle_mod):
    """Three independent mechanisms in one world (three joint kinds, one colour), no contacts."""
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, 8, 0)),
              body(shape=0, dims=(0.1, 0, 0), pos=(0.3, 4.05, 0.0)),
              body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(5, 5, 0)),
              body(shape=1, dims=(0.5, 0.05, 0.05), pos=(5.6, 5, 0)),
              body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(-5, 8, 0)),
              body(shape=1, dims=(0.02, 0.5, 0.3), pos=(-4.8, 4.55, 0.0), angvel=(0.3, 0.0, 1.0)))
    js = abi.joints([abi.JOINT_BALL, abi.JOINT_FIXED, abi.JOINT_REVOLUTE], [0, 2, 4], [1, 3, 5],
                    [(0, -3, 0), (0.1, 0, 0), (0, -3, 0)], [(-0.3, 0.95, 0), (-0.5, 0, 0), (-0.2, 0.45, 0)],
                    frame1=[(0, 0, 0, 1), (0, 0, math.sin(0.05), math.cos(0.05)), (0, 0, 0, 1)],
                    axis1=(0, 0, 1), axis2=(0, 0, 1))
    g, o = _pair(nphys, oracle_mod, s, js, max_velocity_iterations=12)
    _run(g, o, 200, "mechanisms", every=25)


def test_jointed_pairs_in_a_pile(nphys, oracle_mod):
    """A settling pile where every other pair of neighbouring bodies is ball-jointed: joints and
    contacts on the same bodies, joint colours before the contact colours in every sweep."""
    s = scenes.small_pile(seed=31, n=256, lattice=(8, 4, 8), extent=5.0)
    ns = scenes.num_static(s)
    bodies = np.arange(ns, s["status"].shape[0])
    pairs = bodies[: (len(bodies) // 2) * 2].reshape(-1, 2)[::2]
    mid = 0.5 * (s["pos"][pairs[:, 1]] - s["pos"][pairs[:, 0]])
    js = abi.joints(abi.JOINT_BALL, pairs[:, 0], pairs[:, 1], mid, -mid)
    g, o = _pair(nphys, oracle_mod, s, js)
    _run(g, o, 120, "jointed pile", every=20)


def test_break_force_and_removal(nphys, oracle_mod):
    s, b1, b2, a1, a2 = chain_scene(n=5, top=(0.0, 3.0, 0.0), tilt=0.3)
    m = 8 * 0.15 * 0.3 * 0.15
    # the top joint breaks (at 8 sweeps it carries about twice the weight of a link at release), the
    # others hold; joint 3 is removed after 20 steps
    bf = [1.5 * m * 9.81] + [-1.0] * 4
    js = abi.joints(abi.JOINT_BALL, b1, b2, a1, a2, break_force=bf)
    g, o = _pair(nphys, oracle_mod, s, js)
    _run(g, o, 20, "breaking chain", every=5)
    assert g.joints()["broken"][0]
    g.remove_joint(3)
    o.remove_joint(3)
    _run(g, o, 60, "after removal", every=20)


def test_sleeping_chain_on_the_ground(nphys, oracle_mod):
    """A chain lying on the ground with sleep thresholds: joints keep the chain one island; it
    falls asleep as a whole, and an impulse on one end wakes every link."""
    he = (0.3, 0.1, 0.1)
    parts = [scenes.ground()]
    for k in range(5):
        parts.append(body(shape=1, dims=he, pos=(k * 0.62, 0.1, 0.0)))
    s = scenes.with_sleep_threshold(merge(*parts), 0.01)
    js = abi.joints(abi.JOINT_BALL, [1, 2, 3, 4], [2, 3, 4, 5], (0.31, 0, 0), (-0.31, 0, 0))
    g, o = _pair(nphys, oracle_mod, s, js)
    _run(g, o, 300, "sleeping chain", every=50)
    e = g.activation()
    np.testing.assert_array_equal(e, o.activation().astype(np.float32))
    asleep = (e[1:] == 0.0)
    assert asleep.all() or not asleep.any()  # one island
    f = np.array([[0.0, 0.5, 0.0, 0.0, 0.0, 0.0]], np.float32)
    g.apply_force([1], f, abi.IMPULSE, auto_wake_up=True)
    o.apply_force([1], f, abi.IMPULSE, auto_wake_up=True)
    _run(g, o, 30, "woken chain", every=10)


def test_masked_dofs_with_joints(nphys, oracle_mod):
    s, b1, b2, a1, a2 = chain_scene(n=4, top=(0.0, 3.0, 0.0))
    s["kinematic_dofs"] = np.zeros(s["status"].shape[0], np.int32)
    s["kinematic_dofs"][3] = 8 | 16  # rotations about x and y kinematic on one link
    js = abi.joints(abi.JOINT_REVOLUTE, b1, b2, a1, a2, axis1=(0, 0, 1), axis2=(0, 0, 1))
    g, o = _pair(nphys, oracle_mod, s, js)
    _run(g, o, 80, "masked chain", every=20)


def test_residual_target_with_joints(nphys, oracle_mod):
    s, b1, b2, a1, a2 = chain_scene(n=6, top=(0.0, 2.5, 0.0))
    js = abi.joints(abi.JOINT_FIXED, b1, b2, a1, a2)
    g, o = _pair(nphys, oracle_mod, s, js)
    g.set_residual_target(1e-3, 200, 1, "max")
    o.set_residual_target(1e-3, 200, 1, "max")
    _run(g, o, 40, "residual target", every=10)
    assert g.counters()["velocity_sweeps"] == o.counters()["velocity_sweeps"]
    gi, _ = g.residual()
    oi, _ = o.residual()
    assert gi == pytest.approx(oi, rel=1e-6, abs=1e-9)


def test_many_joints_use_per_colour_launches(nphys, oracle_mod):
    """More joints than one workgroup handles (1,152 jointed pairs in a 2,304-body pile): the joint
    phases run one launch per joint colour over many workgroups; still bit for bit."""
    s = scenes.small_pile(seed=37, n=2304, lattice=(12, 16, 12), extent=9.0)
    ns = scenes.num_static(s)
    bodies = np.arange(ns, s["status"].shape[0])
    pairs = bodies.reshape(-1, 2)
    mid = 0.5 * (s["pos"][pairs[:, 1]] - s["pos"][pairs[:, 0]])
    js = abi.joints(abi.JOINT_BALL, pairs[:, 0], pairs[:, 1], mid, -mid)
    assert pairs.shape[0] > 1024
    g, o = _pair(nphys, oracle_mod, s, js)
    _run(g, o, 30, "many joints", every=10)


def test_prismatic_rails_into_the_ground(nphys, oracle_mod):
    """Cuboids on rails (a prismatic joint on an incline with the anchor off the centre of mass, a
    cylindrical one on a vertical axis) sliding into the ground, beside a revolute hinge: joint rows, the rotation and axis
    position constraints and the contacts, bit for bit."""
    s = merge(body(status=1, shape=1, dims=(20.0, 0.5, 20.0), pos=(0, -0.5, 0)),
              body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, 6, 0)),
              body(shape=1, dims=(0.3, 0.2, 0.25), pos=(1.0, 3.0, 0.0)),
              body(shape=1, dims=(0.25, 0.25, 0.25), pos=(-1.0, 3.5, 0.5)),
              body(shape=0, dims=(0.3, 0, 0), pos=(0.0, 2.0, 1.5), angvel=(0.0, 0.0, 2.0)))
    c = math.sqrt(0.5)
    js = abi.joints([abi.JOINT_PRISMATIC, abi.JOINT_CYLINDRICAL, abi.JOINT_REVOLUTE], [1, 1, 1], [2, 3, 4],
                    [(1.0, -3.0, 0.0), (-1.0, -2.5, 0.5), (0.0, -4.0, 1.5)],
                    [(0.1, 0.05, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)],
                    axis1=[(c, -c, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0)], axis2=(0, 0, 1))
    g, o = _pair(nphys, oracle_mod, s, js, max_velocity_iterations=10)
    _run(g, o, 120, "prismatic rails", every=20)
    assert g.counters()["ncontacts"] > 0
