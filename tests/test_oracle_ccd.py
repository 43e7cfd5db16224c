"""Pins of the oracle's continuous collision detection: the conservative-advancement time of impact
(ncollide's TOI queries are unvendored, so their iterates are parity unpinned) and the CCD phase of
the step (MechanicalWorld::solve_ccd, predict_next_impacts, step_ccd; mechanical_world.rs:400-1052,
moreau_jean_solver.rs:94-127).

Every expected value comes from the geometry or from the reference's stated rules evaluated in
closed form:

  * time of impact: head-on approaches give (gap - target) / speed; an oblique ball/ball approach
    and a spinning box falling onto a face are bracketed by the first times the exact distance
    reaches target + tolerance and target (found by dense sampling): every advancement step must be
    conservative (never after the target crossing) and stop within the tolerance; penetration at
    time 0, receding shapes and impacts beyond the bound;
  * the CCD phase: a ball fired at 200 m/s at a thin wall stops at the wall (without CCD it
    passes through), at the distance the CCD substep's ten position iterations leave
    (margins - 0.001 - 0.018 * 0.8^10); with two impacts in one step, a body frozen at the earlier
    impact is moved to its own impact time and then integrated for dt - (last impact time), so its
    tangential travel is v_t (t1 + dt - t2) (mechanical_world.rs:668-676); a window after the first
    (max_ccd_substeps 2) finds the same impact; linear motion interpolation keeps the rotation of
    the end of the step at the impact (rotation angle w (2 dt - t1) instead of w dt); CCD on
    penetration freezes an overlapping pair at time 0; the reference's CCD example keeps its 64
    boxes inside the room, which they leave without CCD.
"""
import math

import numpy as np
import pytest

from nphysics_amd import abi, scenes
from tests.test_oracle_pins import body, merge

TARGET = 0.001
TOL = 1e-4


def _pose(p=(0, 0, 0), q=(0, 0, 0, 1)):
    return tuple(p) + tuple(q)


def test_toi_head_on_closed_forms(oracle_mod):
    # ball / ball, ball / cuboid face, cuboid / cuboid faces: one advancement step, exact
    cases = [
        (0, (0.5,), 0, (0.5,), (3.0, 0, 0), 10.0, 2.0),
        (0, (0.5,), 1, (1.0, 1.0, 1.0), (-5.0, 0, 0), 20.0, 3.5),
        (1, (0.5, 0.5, 0.5), 1, (0.5, 0.5, 0.5), (4.0, 0, 0), 10.0, 3.0),
    ]
    for k1, d1, k2, d2, c2, speed, gap in cases:
        d1 = list(d1) + [0.0] * (3 - len(d1))
        d2 = list(d2) + [0.0] * (3 - len(d2))
        direction = np.sign(c2[0])
        m1 = dict(kind=1, pose=_pose(), v=(speed * direction, 0, 0))
        m2 = dict(kind=0, pose=_pose(c2))
        toi, status = oracle_mod.time_of_impact(k1, d1, m1, k2, d2, m2, 1.0, TARGET)
        assert status == 0
        assert toi == pytest.approx((gap - TARGET) / speed, abs=1e-12)
        # the same with the roles swapped (shape 2 moving towards a static shape 1)
        m1s = dict(kind=0, pose=_pose())
        m2s = dict(kind=1, pose=_pose(c2), v=(-speed * direction, 0, 0))
        toi2, _ = oracle_mod.time_of_impact(k1, d1, m1s, k2, d2, m2s, 1.0, TARGET)
        assert toi2 == pytest.approx(toi, abs=1e-12)


def _first_crossing(dist, level, t_max, step=1e-6):
    t = np.arange(0.0, t_max, step)
    d = dist(t)
    idx = np.nonzero(d <= level)[0]
    return t[idx[0]] if idx.size else None


def test_toi_oblique_balls_bracketed(oracle_mod):
    # ball 1 (r 0.5) at the origin moving +x at 10 m/s, static ball 2 (r 0.5) at (3, 0.8, 0)
    def dist(t):
        return np.sqrt((3.0 - 10.0 * t) ** 2 + 0.8 ** 2) - 1.0

    toi, status = oracle_mod.time_of_impact(0, (0.5, 0, 0), dict(kind=2, pose=_pose(), v=(10, 0, 0)), 0,
                                            (0.5, 0, 0), dict(kind=0, pose=_pose((3.0, 0.8, 0))), 1.0, TARGET)
    exact = (3.0 - math.sqrt((1.0 + TARGET) ** 2 - 0.64)) / 10.0
    lo = (3.0 - math.sqrt((1.0 + TARGET + TOL) ** 2 - 0.64)) / 10.0
    assert status in (0, 1)
    assert lo - 1e-12 <= toi <= exact + 1e-12
    assert dist(np.array([toi]))[0] <= TARGET + TOL + 1e-12


def _spin_z(theta):
    return (0.0, 0.0, math.sin(theta / 2), math.cos(theta / 2))


def test_toi_spinning_box_onto_face_bracketed(oracle_mod):
    # cuboid (0.5, 0.2, 0.3) falling at 20 m/s while spinning at 15 rad/s about z onto the top face
    # (y = 0) of a large static cuboid: distance 2 - 20 t - (0.5 |sin 15t| + 0.2 |cos 15t|)
    def dist(t):
        return 2.0 - 20.0 * t - (0.5 * np.abs(np.sin(15.0 * t)) + 0.2 * np.abs(np.cos(15.0 * t)))

    m1 = dict(kind=2, pose=_pose((0, 2.0, 0)), v=(0, -20, 0), w=(0, 0, 15.0))
    m2 = dict(kind=0, pose=_pose((0, -0.5, 0)))
    toi, status = oracle_mod.time_of_impact(1, (0.5, 0.2, 0.3), m1, 1, (10, 0.5, 10), m2, 1.0, TARGET)
    hi = _first_crossing(dist, TARGET, 0.2)
    lo = _first_crossing(dist, TARGET + TOL, 0.2)
    assert status in (0, 1)
    assert lo - 2e-6 <= toi <= hi + 1e-9
    # a motion started later (t0 < 0: the body has already moved for 0.01 s) is the same path
    m1b = dict(kind=2, t0=-0.01, pose=_pose((0, 2.2, 0), _spin_z(-0.15)), v=(0, -20, 0), w=(0, 0, 15.0))
    toi_b, _ = oracle_mod.time_of_impact(1, (0.5, 0.2, 0.3), m1b, 1, (10, 0.5, 10), m2, 1.0, TARGET)
    assert lo - 2e-6 <= toi_b <= hi + 1e-9
    # without rotation the same fall is a single step
    toi_c, _ = oracle_mod.time_of_impact(1, (0.5, 0.2, 0.3), dict(kind=1, pose=_pose((0, 2.0, 0)), v=(0, -20, 0)),
                                         1, (10, 0.5, 10), m2, 1.0, TARGET)
    assert toi_c == pytest.approx((1.8 - TARGET) / 20.0, abs=1e-12)


def test_toi_status_and_bounds(oracle_mod):
    ball = (0.5, 0, 0)
    # overlapping at time 0: penetrating
    r = oracle_mod.time_of_impact(0, ball, dict(kind=1, pose=_pose(), v=(1, 0, 0)), 0, ball,
                                  dict(kind=0, pose=_pose((0.9, 0, 0))), 1.0, TARGET)
    assert r == (0.0, 2)
    # closer than the target but apart: an impact at time 0
    r = oracle_mod.time_of_impact(0, ball, dict(kind=1, pose=_pose(), v=(1, 0, 0)), 0, ball,
                                  dict(kind=0, pose=_pose((1.0005, 0, 0))), 1.0, TARGET)
    assert r == (0.0, 0)
    # receding, or the impact after the bound
    assert oracle_mod.time_of_impact(0, ball, dict(kind=1, pose=_pose(), v=(-1, 0, 0)), 0, ball,
                                     dict(kind=0, pose=_pose((3, 0, 0))), 1.0, TARGET) is None
    assert oracle_mod.time_of_impact(0, ball, dict(kind=1, pose=_pose(), v=(1, 0, 0)), 0, ball,
                                     dict(kind=0, pose=_pose((3, 0, 0))), 1.0, TARGET) is None
    # a static motion ignores its velocity
    assert oracle_mod.time_of_impact(0, ball, dict(kind=0, pose=_pose(), v=(100, 0, 0)), 0, ball,
                                     dict(kind=0, pose=_pose((3, 0, 0))), 1.0, TARGET) is None


# ---------------------------------------------------------------- the CCD phase of the step
DT = 1.0 / 60.0


def _world(oracle_mod, scene, precision=64, **params):
    p = dict(gravity=(0.0, 0.0, 0.0))
    p.update(params)
    o = oracle_mod.OracleWorld(abi.default_params(**p), precision=precision, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(scene)
    return o


def _wall(x=0.0, z=0.0, friction=0.5, **extra):
    return body(status=1, shape=1, dims=(0.05, 2.0, 2.0), pos=(x, 0, z), friction=friction, **extra)


def _ball(x, vx, y=0.0, z=0.0, vy=0.0, ccd=1, friction=0.5, angvel=(0, 0, 0), **extra):
    return body(shape=0, dims=(0.1, 0, 0), pos=(x, y, z), linvel=(vx, vy, 0), angvel=angvel, friction=friction,
                ccd_enabled=ccd, **extra)


# rest distance of a ball stopped by CCD: the margins (0.02) less what ten position iterations at
# erp 0.2 leave of the correctable depth 0.019 - 0.001 (the scene's sizes are FP32 values)
_M2 = 2 * float(np.float32(0.01))
REST_GAP = _M2 - (0.001 + (_M2 - 0.002) * 0.8 ** 10)
_R, _H = float(np.float32(0.1)), float(np.float32(0.05))


@pytest.mark.parametrize("substeps", [1, 2])
def test_fast_ball_stops_at_thin_wall(oracle_mod, substeps):
    w = _world(oracle_mod, merge(_wall(), _ball(-1.0, 200.0)), max_ccd_substeps=substeps)
    w.step(1)
    st = w.get_state()
    assert st["pos"][1, 0] == pytest.approx(-_H - _R - REST_GAP, abs=1e-9)
    assert abs(st["linvel"][1, 0]) < 1e-9
    c = w.counters()
    assert c["ccd_substeps"] == 1 and c["ccd_impacts"] == 1
    w.step(5)
    assert w.get_state()["pos"][1, 0] < -0.15
    # a later impact, in the second window when there are two (min_overstep, mechanical_world.rs:612-624)
    w2 = _world(oracle_mod, merge(_wall(), _ball(-2.5, 200.0)), max_ccd_substeps=substeps)
    w2.step(1)
    assert w2.get_state()["pos"][1, 0] == pytest.approx(-_H - _R - REST_GAP, abs=1e-9)


def test_without_ccd_the_ball_tunnels(oracle_mod):
    w = _world(oracle_mod, merge(_wall(), _ball(-1.0, 200.0, ccd=0)))
    w.step(1)
    assert w.get_state()["pos"][1, 0] == pytest.approx(-1.0 + 200.0 * DT, abs=1e-9)
    assert w.counters()["ccd_substeps"] == 0


def test_frozen_body_moves_to_its_own_impact_time(oracle_mod):
    # two frictionless balls hit two walls at t1 < t2 in one step: the first is frozen at t1,
    # moved to t1 and integrated for dt - t2 (its tangential travel is 30 (t1 + dt - t2))
    scene = merge(_wall(friction=0.0), _wall(z=10.0, friction=0.0),
                  _ball(-1.0, 200.0, vy=30.0, friction=0.0), _ball(-2.0, 200.0, z=10.0, friction=0.0))
    w = _world(oracle_mod, scene)
    w.step(1)
    st = w.get_state()
    t1 = (1.0 - _H - _R - TARGET) / 200.0
    t2 = (2.0 - _H - _R - TARGET) / 200.0
    assert st["pos"][2, 1] == pytest.approx(30.0 * (t1 + DT - t2), abs=1e-9)
    assert st["pos"][2, 0] == pytest.approx(-_H - _R - REST_GAP, abs=1e-9)
    assert st["pos"][3, 0] == pytest.approx(-_H - _R - REST_GAP, abs=1e-9)
    assert st["linvel"][2, 1] == pytest.approx(30.0, abs=1e-9)
    assert w.counters()["ccd_impacts"] == 2


def _angle_x(q):
    return 2.0 * math.atan2(q[0], q[3])


@pytest.mark.parametrize("interp", [0, 1])
def test_linear_motion_interpolation_keeps_the_end_rotation(oracle_mod, interp):
    # a frictionless ball spinning about the wall normal: the contact exerts no torque
    ball = _ball(-1.0, 200.0, friction=0.0, angvel=(20.0, 0, 0), linear_motion_interpolation=interp)
    w = _world(oracle_mod, merge(_wall(friction=0.0), ball))
    w.step(1)
    st = w.get_state()
    t1 = (1.0 - _H - float(np.float32(0.1)) - TARGET) / 200.0
    expected = 20.0 * (2 * DT - t1) if interp else 20.0 * DT
    assert _angle_x(st["rot"][1]) == pytest.approx(expected, abs=1e-7)
    assert st["pos"][1, 0] < -0.15 + 1e-6


def test_freeze_resweeps_the_frozen_bodys_pairs(oracle_mod):
    # ball A (200 m/s) hits the wall at t_A; ball B follows at 150 m/s. The first pass, with A still
    # moving faster than B, finds no A/B impact; only the re-sweep of A's pairs after its freeze
    # (mechanical_world.rs:503-551) finds B reaching the frozen A at t_AB, before B would reach the
    # wall, so B stops on A instead of being frozen at the wall where A is
    scene = merge(_wall(friction=0.0), _ball(-1.0, 200.0, friction=0.0), _ball(-2.5, 150.0, friction=0.0))
    w = _world(oracle_mod, scene)
    w.step(1)
    assert w.counters()["ccd_impacts"] == 2
    x = w.get_state()["pos"][1:, 0]
    assert x[1] < x[0] - 0.19


def test_pair_of_two_frozen_bodies_counts_only_at_the_later_freeze(oracle_mod):
    # A is frozen by the wall at t_A, B by a plate behind A's start at t_B > t_A; the re-sweep after
    # A's freeze predicts B reaching A at t_AB > t_B (through the plate), an entry that finds both
    # frozen when it is popped and is kept only if t_AB equals the later freeze time
    # (mechanical_world.rs:476-489): two impacts, not three
    scene = merge(_wall(friction=0.0), _wall(x=-1.5, friction=0.0), _ball(-1.0, 200.0, friction=0.0),
                  _ball(-2.5, 150.0, friction=0.0))
    w = _world(oracle_mod, scene)
    w.step(1)
    assert w.counters()["ccd_impacts"] == 2
    assert w.get_state()["pos"][3, 0] < -1.55


def test_superseded_time_of_impact_is_skipped(oracle_mod):
    # A (200 m/s) and B (-200 m/s, 0.15 m higher) would meet at t0; a small static ball in A's path
    # only freezes A first at t_A < t0. The re-sweep replaces the A/B entry by the time B reaches
    # the frozen A (t1 > t0); the first-pass entry at t0 is stale and skipped (mechanical_world.rs:
    # 471-474), so B is frozen where it touches A, not at t0 half a metre short of it
    pebble = body(status=1, shape=0, dims=(0.02, 0, 0), pos=(-0.5, -0.05, 0), friction=0.0)
    scene = merge(pebble, _ball(-1.0, 200.0, friction=0.0), _ball(1.0, -200.0, y=0.15, friction=0.0))
    w = _world(oracle_mod, scene)
    w.step(1)
    assert w.counters()["ccd_impacts"] == 2
    x = w.get_state()["pos"][1:, 0]
    assert x[1] - x[0] < 0.2


@pytest.mark.parametrize("interp", [0, 1])
def test_body_still_impacting_is_clamped(oracle_mod, interp):
    # a frictionless, perfectly elastic ball spinning about x bounces off wall A in the substep and
    # is then predicted to hit wall B within the step: it is put back to the start of its motion
    # (the pose validated at the first impact, mechanical_world.rs:925-932, rigid_body.rs:530-537)
    # with that pose's rotation (w t1), or with linear motion interpolation with its current one
    # (w (2 dt - t1)); the velocity keeps the bounce
    ball = _ball(-0.3, 200.0, friction=0.0, angvel=(20.0, 0, 0), linear_motion_interpolation=interp,
                 restitution=1.0)
    scene = merge(_wall(friction=0.0, restitution=1.0), _wall(x=-0.6, friction=0.0, restitution=1.0), ball)
    w = _world(oracle_mod, scene)
    w.step(1)
    st = w.get_state()
    t1 = (0.3 - _H - _R - TARGET) / 200.0
    assert st["pos"][2, 0] == pytest.approx(-_H - _R - REST_GAP, abs=1e-9)
    assert st["linvel"][2, 0] == pytest.approx(-200.0, abs=1e-6)
    expected = 20.0 * (2 * DT - t1) if interp else 20.0 * t1
    assert _angle_x(st["rot"][2]) == pytest.approx(expected, abs=1e-7)


@pytest.mark.parametrize("on", [0, 1])
def test_ccd_on_penetration(oracle_mod, on):
    # two overlapping CCD balls fired together at the wall: with CCD on penetration their pair has
    # an impact at time 0, which freezes both where they start; without it each is frozen at its own
    # impact with the wall (the second one ignoring the first), and both stay on this side of it
    scene = merge(_wall(friction=0.0), _ball(-1.0, 200.0, friction=0.0), _ball(-1.15, 200.0, friction=0.0))
    w = _world(oracle_mod, scene, ccd_on_penetration_enabled=on)
    w.step(1)
    x = w.get_state()["pos"][1:, 0]
    if on:
        assert np.all(x < -0.9)
    else:
        assert np.all(x > -0.5) and np.all(x < 0.0)
        assert w.counters()["ccd_impacts"] == 2


def _inside_room(pos, hw=3.0, thickness=0.2):
    return np.all(np.abs(pos) < hw - thickness + 0.05, axis=1)


def test_reference_ccd_example_keeps_the_boxes_in(oracle_mod):
    s = scenes.ccd_room()
    ns = scenes.num_static(s)
    w = oracle_mod.OracleWorld(abi.default_params(), precision=32, order=oracle_mod.ORDER_COLOURED)
    w.add_bodies(s)
    w.step(30)
    assert np.all(_inside_room(w.get_state()["pos"][ns:]))
    assert w.counters()["ccd_substeps"] > 0
    off = dict(s)
    off["ccd_enabled"] = np.zeros_like(s["ccd_enabled"])
    w2 = oracle_mod.OracleWorld(abi.default_params(), precision=32, order=oracle_mod.ORDER_COLOURED)
    w2.add_bodies(off)
    w2.step(30)
    assert not np.all(_inside_room(w2.get_state()["pos"][ns:]))


def test_substep_leaves_joints_out(oracle_mod):
    # a pendulum beside a CCD impact: the substep solves contacts only (mechanical_world.rs:909),
    # and the joint holds again in the following steps
    s = merge(_wall(), _ball(-1.0, 200.0), body(shape=0, dims=(0.1, 0, 0), pos=(-0.5, 1.0, 1.0)),
              body(status=1, shape=0, dims=(0.05, 0, 0), pos=(-0.5, 1.5, 1.0)))
    w = _world(oracle_mod, s, precision=32, gravity=(0.0, -9.81, 0.0))
    w.add_joints(abi.joints(abi.JOINT_BALL, 3, 2, (0.0, 0.0, 0.0), (0.0, 0.5, 0.0)))
    w.step(10)
    assert w.counters()["ccd_substeps"] >= 1
    p = w.get_state()["pos"]
    assert np.linalg.norm(p[2] - p[3]) == pytest.approx(0.5, abs=5e-3)


def test_ccd_rejects_sleeping_and_scenes(oracle_mod):
    s = merge(_wall(), _ball(-1.0, 200.0, sleep_threshold=0.01))
    w = _world(oracle_mod, s)
    with pytest.raises(RuntimeError):
        w.step(1)
