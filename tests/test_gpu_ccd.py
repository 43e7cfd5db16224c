"""Continuous collision detection on the GPU against the CPU restatement in the colour order, bit for
bit: states after every step and the CCD counters, in the scenes that pin the oracle's CCD rules
(a fast ball at a thin wall, two impacts in one step, a re-sweep after a freeze, a superseded
entry, two frozen bodies, a body put back with and without linear motion interpolation, CCD on
penetration, two windows per step) and in the reference's CCD example (a room of six walls, one
static body, with 64 boxes at 100 m/s), also with balls and rotated boxes, joints, forces applied
after a CCD step, and the time step of the dynamics update a substep leaves for the next step."""
import numpy as np
import pytest

from nphysics_amd import abi, scenes
from tests.test_oracle_pins import body, merge

pytestmark = pytest.mark.gpu

ZERO_G = dict(gravity=(0.0, 0.0, 0.0))


def _pair(nphys, oracle_mod, s, **params):
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params(**params))
    g.add_bodies(s)
    o = oracle_mod.OracleWorld(abi.default_params(**params), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(s)
    return g, o


def _same(g, o, what):
    gs, os_ = g.get_state(), o.get_state()
    for k in ("pos", "rot", "linvel", "angvel"):
        a, b = gs[k], os_[k].astype(np.float32)
        if not np.array_equal(a, b):
            i = int(np.argmax(np.abs(a - b).max(axis=1)))
            raise AssertionError(f"{what}: {k} differs (body {i}: {a[i]} vs {b[i]})")
    gc, oc = g.counters(), o.counters()
    for k in ("ccd_substeps", "ccd_impacts"):
        assert gc[k] == oc[k], f"{what}: {k} {gc[k]} vs {oc[k]}"


def _run(g, o, steps, what, every=1):
    for k in range(steps):
        g.step(1)
        o.step(1)
        if (k + 1) % every == 0 or k + 1 == steps:
            print(f"{what}: step {k + 1}, substeps {g.counters()['ccd_substeps']}", flush=True)
            _same(g, o, f"{what} step {k + 1}")


def _wall(x=0.0, z=0.0, friction=0.5, **extra):
    return body(status=1, shape=1, dims=(0.05, 2.0, 2.0), pos=(x, 0, z), friction=friction, **extra)


def _ball(x, vx, y=0.0, z=0.0, vy=0.0, ccd=1, friction=0.5, angvel=(0, 0, 0), **extra):
    return body(shape=0, dims=(0.1, 0, 0), pos=(x, y, z), linvel=(vx, vy, 0), angvel=angvel, friction=friction,
                ccd_enabled=ccd, **extra)


SMALL = {
    "ball_at_wall": lambda: merge(_wall(), _ball(-1.0, 200.0)),
    "two_impacts": lambda: merge(_wall(friction=0.0), _wall(z=10.0, friction=0.0),
                                 _ball(-1.0, 200.0, vy=30.0, friction=0.0), _ball(-2.0, 200.0, z=10.0, friction=0.0)),
    "resweep": lambda: merge(_wall(friction=0.0), _ball(-1.0, 200.0, friction=0.0), _ball(-2.5, 150.0, friction=0.0)),
    "two_frozen": lambda: merge(_wall(friction=0.0), _wall(x=-1.5, friction=0.0), _ball(-1.0, 200.0, friction=0.0),
                                _ball(-2.5, 150.0, friction=0.0)),
    "superseded": lambda: merge(body(status=1, shape=0, dims=(0.02, 0, 0), pos=(-0.5, -0.05, 0), friction=0.0),
                                _ball(-1.0, 200.0, friction=0.0), _ball(1.0, -200.0, y=0.15, friction=0.0)),
    "clamped": lambda: merge(_wall(friction=0.0, restitution=1.0), _wall(x=-0.6, friction=0.0, restitution=1.0),
                             _ball(-0.3, 200.0, friction=0.0, angvel=(20.0, 0, 0), restitution=1.0)),
    "clamped_interp": lambda: merge(_wall(friction=0.0, restitution=1.0), _wall(x=-0.6, friction=0.0, restitution=1.0),
                                    _ball(-0.3, 200.0, friction=0.0, angvel=(20.0, 0, 0), restitution=1.0,
                                          linear_motion_interpolation=1)),
    "overlapping": lambda: merge(_wall(friction=0.0), _ball(-1.0, 200.0, friction=0.0),
                                 _ball(-1.15, 200.0, friction=0.0)),
    "spinning_box": lambda: merge(_wall(), body(shape=1, dims=(0.1, 0.2, 0.15), pos=(-1.0, 0.1, 0.2),
                                                linvel=(180.0, 5.0, -3.0), angvel=(3.0, 17.0, -8.0), ccd_enabled=1)),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_ccd_scenes_bitexact(nphys, oracle_mod, name):
    g, o = _pair(nphys, oracle_mod, SMALL[name](), **ZERO_G)
    _run(g, o, 6, name)
    assert g.counters()["ccd_substeps"] > 0


@pytest.mark.parametrize("opts", [dict(ccd_on_penetration_enabled=1), dict(max_ccd_substeps=2),
                                  dict(max_ccd_substeps=3, max_ccd_position_iterations=4)])
def test_ccd_parameters_bitexact(nphys, oracle_mod, opts):
    for name in ("overlapping", "two_impacts", "resweep"):
        g, o = _pair(nphys, oracle_mod, SMALL[name](), **ZERO_G, **opts)
        _run(g, o, 4, f"{name} {opts}")


def test_reference_ccd_example_bitexact(nphys, oracle_mod):
    s = scenes.ccd_room()
    g, o = _pair(nphys, oracle_mod, s)
    _run(g, o, 40, "ccd room", every=5)
    ns = scenes.num_static(s)
    pos = g.get_state()["pos"][ns:]
    assert np.all(np.abs(pos) < 3.0 - 0.2 + 0.05)
    assert g.counters()["ccd_substeps"] > 0


def test_ccd_room_4096_cascade_bitexact(nphys, oracle_mod):
    """The 4,096-cuboid room at full size: every body freezes in each step's cascade (about 3,000
    impacts per step), and the first pass leaves about 24,000 queue entries, so the queue runs with
    its deeper heap levels in HBM."""
    s = scenes.ccd_room(num=16, half_width=4.0, height=-1.76)
    g, o = _pair(nphys, oracle_mod, s)
    _run(g, o, 4, "ccd room 4096")
    assert g.counters()["ccd_impacts"] > 5000


def test_ccd_room_balls_and_rotated_boxes_bitexact(nphys, oracle_mod):
    s = scenes.ccd_room(num=6, rad=0.12, seed=11, velocity=(-60.0, -25.0, 40.0))
    g, o = _pair(nphys, oracle_mod, s)
    _run(g, o, 30, "mixed ccd room", every=5)


def test_ccd_with_joints_and_forces_bitexact(nphys, oracle_mod):
    # a pendulum next to the wall (joints are not solved in a substep), and impulses applied after
    # steps with CCD substeps (they act through the augmented mass of the substep's update)
    s = merge(_wall(), _ball(-1.0, 200.0), body(shape=0, dims=(0.1, 0, 0), pos=(-0.5, 1.0, 1.0)),
              body(status=1, shape=0, dims=(0.05, 0, 0), pos=(-0.5, 1.5, 1.0)))
    js = abi.joints(abi.JOINT_BALL, 3, 2, (0.0, 0.0, 0.0), (0.0, 0.5, 0.0))
    g, o = _pair(nphys, oracle_mod, s)
    g.add_joints(js)
    o.add_joints(js)
    for k in range(4):
        g.step(1)
        o.step(1)
        _same(g, o, f"joints step {k + 1}")
        imp = np.array([[0.0, 0.0, 0.0, 0.3, -0.2, 0.1]], np.float32)
        g.apply_force([1], imp, force_type=abi.IMPULSE)
        o.apply_force([1], imp, force_type=abi.IMPULSE)
        _same(g, o, f"after impulse {k + 1}")


def test_ccd_rejects_sleeping_bodies(nphys):
    s = merge(_wall(), _ball(-1.0, 200.0, sleep_threshold=0.01))
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params())
    g.add_bodies(s)
    with pytest.raises(Exception):
        g.step(1)
