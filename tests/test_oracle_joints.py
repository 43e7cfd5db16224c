"""Pins of the oracle's joint constraints (BallConstraint, FixedConstraint, RevoluteConstraint
without limits or motor; src/joint/, src/solver/helper.rs, sor_prox.rs, nonlinear_sor_prox.rs)
by closed forms of mechanics, each chosen so that a plausible slip fails it:

  * ball joint: the small-angle period of a physical pendulum 2 pi sqrt(I_pivot / (m g L)) (a
    jacobian sign or lever-arm slip changes the period or makes it diverge); the cached impulses of
    a hanging chain equal the weight below each joint times dt (a dropped warm start or a wrong
    impulse index fails it); one position iteration moves the free anchor by exactly
    (|error| - allowed_linear_error) erp (helper.rs:364-420, nonlinear_sor_prox.rs:85-119, 296-309);
  * fixed joint: a cantilever holds its weight: linear impulse m g dt, angular impulse m g l dt; one
    angular position iteration turns the body by (theta - allowed_angular_error) erp times
    (1/I) / (1/m + 1/I) - the reference adds the inverse mass to an angular row's inv_r
    (rigid_body.rs:707-708), and the pin fails without that term;
  * revolute joint: spin about the hinge axis is untouched, spin about the other axes is removed in
    one sweep (isotropic inertia: the two angular rows are decoupled), a hinged plate swings in the
    plane normal to the axis;
  * break force: a joint whose weight exceeds its break force breaks after one step, the body then
    falls freely;
  * the angle helpers (atan by half-angle reduction and series) against numpy.
"""
import math

import numpy as np
import pytest

from nphysics_amd import abi
from tests.test_oracle_pins import body, merge

G = 9.81
DT = 1.0 / 60.0


def world(oracle_mod, s, prec=64, order=None, **kw):
    o = oracle_mod.OracleWorld(abi.default_params(**kw), precision=prec,
                               order=oracle_mod.ORDER_REFERENCE if order is None else order)
    o.add_bodies(s)
    return o


def ball_mass(r, density=1.0):
    return density * 4.0 / 3.0 * math.pi * r ** 3


def test_angle_helpers_match_numpy(oracle_mod):
    """The rotation error of a fixed joint (2 atan2(|v|, |w|) of a random quaternion, by the oracle's
    series atan) through one full-strength angular position iteration: the body is turned back by
    10/11 of the angle (a ball: 1/I over 1/m + 1/I), against numpy's arctan2."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ang = rng.uniform(0.01, 3.1)
        q = np.concatenate([axis * math.sin(ang / 2), [math.cos(ang / 2)]])
        # a fixed joint between a static frame and a body rotated by q: the first position
        # iteration's rotation error is the scaled axis of q; the displacement is along -axis
        s = merge(body(status=1, shape=1, dims=(0.1, 0.1, 0.1), pos=(0, 0, 0)),
                  body(shape=0, dims=(0.5, 0, 0), pos=(0, 0, 0), rot=tuple(q)))
        o = world(oracle_mod, s, gravity=(0, 0, 0), max_velocity_iterations=0, max_position_iterations=1,
                  erp=1.0, max_angular_correction=10.0, allowed_angular_error=0.0)
        o.add_joints(abi.joints(abi.JOINT_FIXED, 0, 1, (0, 0, 0), (0, 0, 0)))
        o.step(1)
        # ball: I = 2/5 m r^2 = 0.1 m; angular row inv_r = 1/m + 1/I = 11/m; displacement = the
        # error times (1/I) / (1/m + 1/I) = 10/11 of the angle, about -axis
        q1 = o.get_state()["rot"][1]
        # remaining rotation angle: ang - 10/11 ang
        rem = 2 * math.atan2(np.linalg.norm(q1[:3]), abs(q1[3]))
        # tolerance: the displacement's quaternion exponential is a series truncated near 4e-10
        assert rem == pytest.approx(ang / 11.0, rel=1e-7, abs=1e-9)


@pytest.mark.parametrize("order", ["reference", "coloured"])
def test_ball_joint_pendulum_period(oracle_mod, order):
    L, r, theta = 1.0, 0.1, 0.1
    piv = np.array([0.0, 5.0, 0.0])
    c = piv + L * np.array([math.sin(theta), -math.cos(theta), 0.0])
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=tuple(piv)),
              body(shape=0, dims=(r, 0, 0), pos=tuple(c)))
    o = world(oracle_mod, s, order=getattr(oracle_mod, "ORDER_" + order.upper()), max_velocity_iterations=30)
    o.add_joints(abi.joints(abi.JOINT_BALL, 0, 1, (0, 0, 0), tuple(piv - c)))
    xs = []
    for _ in range(400):
        o.step(1)
        xs.append(o.get_state()["pos"][1, 0])
    xs = np.array(xs)
    down = [i for i in range(1, len(xs)) if xs[i - 1] > 0 >= xs[i]]
    period = (down[-1] - down[0]) / (len(down) - 1) * DT
    expected = 2 * math.pi * math.sqrt((0.4 * r * r + L * L) / (G * L))
    assert period == pytest.approx(expected, rel=0.005)
    # the bob stays on the sphere of radius L about the pivot
    p = o.get_state()["pos"][1]
    assert np.linalg.norm(p - piv) == pytest.approx(L, abs=2e-3)


@pytest.mark.parametrize("order", ["reference", "coloured"])
def test_hanging_chain_impulses(oracle_mod, order):
    """Five balls (r 0.1) hanging 0.3 m apart from a static anchor by ball joints: at rest joint k
    (0 at the top) carries the weight of the 5 - k balls below it, impulse (5 - k) m g dt."""
    n, gap, r = 5, 0.3, 0.1
    top = np.array([0.0, 5.0, 0.0])
    parts = [body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=tuple(top))]
    for k in range(n):
        parts.append(body(shape=0, dims=(r, 0, 0), pos=tuple(top - (k + 1) * gap * np.array([0, 1, 0]))))
    s = merge(*parts)
    o = world(oracle_mod, s, order=getattr(oracle_mod, "ORDER_" + order.upper()), max_velocity_iterations=50)
    a1 = [(0, 0, 0)] + [(0, -gap / 2, 0)] * (n - 1)
    a2 = [(0, gap, 0)] + [(0, gap / 2, 0)] * (n - 1)
    o.add_joints(abi.joints(abi.JOINT_BALL, list(range(n)), list(range(1, n + 1)), a1, a2))
    o.step(120)
    j = o.joints()
    w = ball_mass(r) * G * DT
    for k in range(n):
        assert abs(j["lin"][k, 1]) == pytest.approx((n - k) * w, rel=1e-3)
        assert abs(j["lin"][k, 0]) < 1e-3 * w and abs(j["lin"][k, 2]) < 1e-3 * w
    assert not j["broken"].any()
    st = o.get_state()
    np.testing.assert_allclose(st["pos"][1:, 1], top[1] - gap * np.arange(1, n + 1), atol=2e-3)


def test_ball_joint_position_iteration(oracle_mod):
    """Anchors 0.1 m apart, the free anchor at the centre of mass, no gravity and no velocity
    sweeps: each position iteration removes (error - 0.001) * 0.2 of the gap."""
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, 0, 0)),
              body(shape=0, dims=(0.25, 0, 0), pos=(1.125, 0, 0)))
    o = world(oracle_mod, s, gravity=(0, 0, 0), max_velocity_iterations=0, max_position_iterations=3)
    o.add_joints(abi.joints(abi.JOINT_BALL, 0, 1, (1.0, 0, 0), (0, 0, 0)))
    o.step(1)
    gap = 0.125
    for _ in range(3):
        gap -= (gap - 0.001) * 0.2
    assert o.get_state()["pos"][1, 0] - 1.0 == pytest.approx(gap, rel=1e-12)


def test_fixed_joint_cantilever(oracle_mod):
    """A 1 m bar (half extents 0.5, 0.05, 0.05) held at one end by a fixed joint: the joint carries
    its weight (linear impulse m g dt) and the moment about the anchor (angular m g 0.5 dt)."""
    he = (0.5, 0.05, 0.05)
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(-0.6, 5, 0)),
              body(shape=1, dims=he, pos=(0.0, 5, 0)))
    o = world(oracle_mod, s, max_velocity_iterations=40)
    o.add_joints(abi.joints(abi.JOINT_FIXED, 0, 1, (0.1, 0, 0), (-0.5, 0, 0)))
    o.step(60)
    m = 8 * he[0] * he[1] * he[2]
    j = o.joints()
    assert abs(j["lin"][0, 1]) == pytest.approx(m * G * DT, rel=1e-3)
    assert abs(j["ang"][0, 2]) == pytest.approx(m * G * 0.5 * DT, rel=1e-3)
    assert np.abs(j["ang"][0, :2]).max() < 1e-6
    st = o.get_state()
    assert abs(st["pos"][1, 1] - 5.0) < 2e-3 and abs(st["rot"][1, 2]) < 2e-3


def test_fixed_joint_rotation_iteration_has_inverse_mass_term(oracle_mod):
    """A ball of radius 0.5 turned by 0.3 rad about z from its fixed joint's frame, anchor at the
    centre: one angular position iteration turns it back by (0.3 - 0.001) 0.2 (1/I) / (1/m + 1/I)."""
    theta = 0.3
    q = (0.0, 0.0, math.sin(theta / 2), math.cos(theta / 2))
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, 0, 0)),
              body(shape=0, dims=(0.5, 0, 0), pos=(0, 0, 0), rot=q))
    o = world(oracle_mod, s, gravity=(0, 0, 0), max_velocity_iterations=0, max_position_iterations=1)
    o.add_joints(abi.joints(abi.JOINT_FIXED, 0, 1, (0, 0, 0), (0, 0, 0)))
    o.step(1)
    m = ball_mass(0.5)
    inertia = 0.4 * m * 0.25
    q32 = np.asarray(q, np.float32).astype(np.float64)
    theta = 2 * math.atan2(q32[2], q32[3])  # the angle of the FP32 input
    turn = (theta - 0.001) * 0.2 * (1 / inertia) / (1 / m + 1 / inertia)
    q1 = o.get_state()["rot"][1]
    # tolerance: the displacement's quaternion exponential is a series truncated near 4e-10
    assert 2 * math.atan2(q1[2], q1[3]) == pytest.approx(theta - turn, rel=1e-8)


@pytest.mark.parametrize("sweeps", [1, 2, 5])
@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (1 / 3, 2 / 3, 2 / 3)])
def test_revolute_keeps_axial_spin_and_removes_the_rest(oracle_mod, sweeps, axis):
    """A cube (isotropic inertia, anchor at its centre) hinged about an axis (given in the static
    body's frame, which is turned by 90 degrees about y) and spinning about all three axes: the spin
    about the world axis is untouched; each sweep scales the rest by (1/m) / (1/m + 1/I), the
    reference's angular row taking r = 1 / (1/m + 1/I)."""
    he = 0.3
    qs = (0.0, math.sin(math.pi / 4), 0.0, math.cos(math.pi / 4))
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, 3, 0), rot=qs),
              body(shape=1, dims=(he, he, he), pos=(0, 0, 0), angvel=(1.5, -0.7, 2.0)))
    o = world(oracle_mod, s, gravity=(0, 0, 0), max_velocity_iterations=sweeps, max_position_iterations=0)
    # the static body's frame maps local (x, y, z) to world (z, y, -x)
    anchor1 = (0.0, -3.0, 0.0)
    o.add_joints(abi.joints(abi.JOINT_REVOLUTE, 0, 1, anchor1, (0, 0, 0), axis1=axis, axis2=axis))
    o.step(1)
    m = 8 * he ** 3
    inertia = m * 8 * he * he / 12
    keep = (1 / m) / (1 / m + 1 / inertia)
    a = np.asarray(axis, np.float32).astype(np.float64)
    a_world = np.array([a[2], a[1], -a[0]])
    a_world /= np.linalg.norm(a_world)
    w = o.get_state()["angvel"][1]
    w0 = np.asarray((1.5, -0.7, 2.0), np.float32).astype(np.float64)
    along = w0 @ a_world
    np.testing.assert_allclose(w, along * a_world + (w0 - along * a_world) * keep ** sweeps, rtol=1e-5, atol=1e-6)


def test_revolute_plate_swings_in_its_plane(oracle_mod):
    """A plate hinged at its top edge about z, released at 0.4 rad: it swings in the xy plane (no z
    motion, no rotation about x or y) with the physical-pendulum period of a rod about its end,
    2 pi sqrt(2 h / (3 g)) for height 2 h (small amplitude; 0.4 rad lengthens it by about 1%)."""
    h = 0.5
    theta = 0.4
    piv = np.array([0.0, 5.0, 0.0])
    c = piv + h * np.array([math.sin(theta), -math.cos(theta), 0])
    q = (0.0, 0.0, math.sin(theta / 2), math.cos(theta / 2))
    # the static body sits away from the plate (jointed bodies still collide, as in the reference)
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=tuple(piv + (0, 3, 0))),
              body(shape=1, dims=(0.02, h, 0.3), pos=tuple(c), rot=q))
    o = world(oracle_mod, s, max_velocity_iterations=30)
    o.add_joints(abi.joints(abi.JOINT_REVOLUTE, 0, 1, (0, -3, 0), (0, h, 0), axis1=(0, 0, 1), axis2=(0, 0, 1)))
    xs = []
    for _ in range(300):
        o.step(1)
        st = o.get_state()
        xs.append(st["pos"][1, 0])
        assert abs(st["pos"][1, 2]) < 1e-9
        assert abs(st["angvel"][1, 0]) < 1e-6 and abs(st["angvel"][1, 1]) < 1e-6
    xs = np.array(xs)
    down = [i for i in range(1, len(xs)) if xs[i - 1] > 0 >= xs[i]]
    period = (down[-1] - down[0]) / (len(down) - 1) * DT
    # rod of length 2h about its end (the plate's thickness and depth do not change I about z much:
    # I_end = m ((2h)^2 + (0.04)^2) / 12 + m h^2)
    i_end = ((2 * h) ** 2 + 0.04 ** 2) / 12 + h * h
    small = 2 * math.pi * math.sqrt(i_end / (G * h))
    large = small * (1 + theta ** 2 / 16)
    assert period == pytest.approx(large, rel=0.01)


def test_break_force_then_free_fall(oracle_mod):
    r = 0.1
    m = ball_mass(r)
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, 5, 0)),
              body(shape=0, dims=(r, 0, 0), pos=(0, 4.5, 0)))
    for bf, should_break in ((0.5 * m * G, True), (2.0 * m * G, False)):
        o = world(oracle_mod, s, max_velocity_iterations=20)
        o.add_joints(abi.joints(abi.JOINT_BALL, 0, 1, (0, 0, 0), (0, 0.5, 0), break_force=bf))
        o.step(1)
        assert bool(o.joints()["broken"][0]) == should_break
        if should_break:
            v0 = o.get_state()["linvel"][1, 1]
            o.step(10)
            assert o.get_state()["linvel"][1, 1] == pytest.approx(v0 - 10 * G * DT, rel=1e-9)


def test_removed_joint_is_ignored(oracle_mod):
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, 5, 0)),
              body(shape=0, dims=(0.1, 0, 0), pos=(0, 4.5, 0)))
    o = world(oracle_mod, s)
    o.add_joints(abi.joints(abi.JOINT_BALL, 0, 1, (0, 0, 0), (0, 0.5, 0)))
    o.remove_joint(0)
    o.step(10)
    assert o.get_state()["linvel"][1, 1] == pytest.approx(-10 * G * DT, rel=1e-9)


def test_prismatic_slide_closed_form(oracle_mod):
    """A cuboid on a frictionless inclined rail (PrismaticConstraint, prismatic_constraint.rs:132-345;
    the anchor at its centre of mass, so the rows do not couple rotation): the two linear rows
    orthogonal to the axis are decoupled under the isotropic mass, and each step's solve keeps
    exactly the gravity component along the axis, v_k = k dt (g . a) a, with no rotation. Three
    linear rows (the fixed joint's) would hold it still; an axis row instead of an orthogonal one
    would let it fall freely."""
    c = math.sqrt(0.5)
    a = np.array([c, -c, 0.0])
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(0, -40, 0)),
              body(shape=1, dims=(0.2, 0.3, 0.25), pos=(0.0, 0.0, 0.0)))
    js = abi.joints(abi.JOINT_PRISMATIC, 0, 1, (0.0, 40.0, 0.0), (0.0, 0.0, 0.0), axis1=tuple(a))
    o = oracle_mod.OracleWorld(abi.default_params(), precision=64)
    o.add_bodies(s)
    o.add_joints(js)
    dt = 1.0 / 60.0
    g = np.array([0.0, -9.81, 0.0])
    for k in range(1, 31):
        o.step(1)
        st = o.get_state()
        np.testing.assert_allclose(st["linvel"][1], k * dt * g.dot(a) * a, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(st["angvel"][1], 0.0, atol=1e-12)
    # the anchor stays on the rail
    p = st["pos"][1]
    assert np.linalg.norm(p - a * p.dot(a)) < 1e-9


def test_cylindrical_slide_and_spin_closed_form(oracle_mod):
    """A cube on a vertical cylindrical joint (CylindricalConstraint, cylindrical_constraint.rs:125-292;
    the anchor at its centre of mass, isotropic inertia, so the two linear and the two angular rows
    are each decoupled): the first step's solve removes the velocity and spin across the axis
    exactly, and from then on the cube falls freely along the axis while its spin about it stays.
    The prismatic joint's third angular row would stop the spin, the revolute joint's third linear
    row the fall."""
    s = merge(body(status=1, shape=1, dims=(0.05, 0.05, 0.05), pos=(30, 0, 0)),
              body(shape=1, dims=(0.25, 0.25, 0.25), pos=(0.0, 0.0, 0.0), linvel=(1.0, 2.0, 3.0),
                   angvel=(0.4, 5.0, -0.7)))
    js = abi.joints(abi.JOINT_CYLINDRICAL, 0, 1, (-30.0, 0.0, 0.0), (0.0, 0.0, 0.0), axis1=(0, 1, 0),
                    axis2=(0, 1, 0))
    o = oracle_mod.OracleWorld(abi.default_params(), precision=64)
    o.add_bodies(s)
    o.add_joints(js)
    dt = 1.0 / 60.0
    for k in range(1, 21):
        o.step(1)
        st = o.get_state()
        np.testing.assert_allclose(st["linvel"][1], (0.0, 2.0 - 9.81 * dt * k, 0.0), rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(st["angvel"][1], (0.0, 5.0, 0.0), rtol=1e-9, atol=1e-10)
