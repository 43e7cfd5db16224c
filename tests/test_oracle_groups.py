"""Pins of the oracle's collision groups (ColliderDesc::collision_groups, collider.rs:444,506; the
rule is ncollide3d 0.31's CollisionGroups::can_interact_with_groups, unvendored): two colliders
interact only when each is a member of some group the other whitelists and of no group the other
blacklists.

  * the truth table below has a case for each of the rule's four conditions (dropping one of them,
    or making a whitelist test one-sided, turns a "no pair" case into a pair);
  * the reference's collision groups example (examples3d/collision_groups3.rs) fixes the outcome:
    the green layers fall through the blue floor and rest on the green floor, the blue layers rest
    on the blue floor, and no green body ever pairs with a blue one.
"""
import numpy as np
import pytest

from nphysics_amd import abi, scenes
from tests.test_oracle_pins import body, merge

cg = abi.collision_groups

# (groups of ball A, groups of ball B, whether the two overlapping balls form a pair)
TRUTH = {
    "defaults": (cg(), cg(), True),
    "disjoint_whitelists": (cg([0], [0]), cg([1], [1]), False),
    "one_side_default": (cg([0], [0]), cg(), True),
    "b_blacklists_a": (cg([0]), cg([1], blacklist=[0]), False),
    "a_blacklists_b": (cg([0], blacklist=[1]), cg([1]), False),
    "crossed_whitelists": (cg([0, 1], [2]), cg([2], [0]), True),
    "a_not_whitelisted_by_b": (cg([0], [0]), cg([0, 3], [1]), False),
    "b_not_whitelisted_by_a": (cg([0, 3], [1]), cg([0], [0]), False),
    "no_membership": (cg([]), cg(), False),
}


def pair_scene(ga, gb):
    return merge(body(shape=0, dims=(0.5, 0, 0), pos=(0.0, 2.0, 0.0), collision_groups=ga),
                 body(shape=0, dims=(0.5, 0, 0), pos=(0.6, 2.0, 0.0), collision_groups=gb))


@pytest.mark.parametrize("case", list(TRUTH))
def test_group_truth_table(oracle_mod, case):
    ga, gb, expect = TRUTH[case]
    o = oracle_mod.OracleWorld(abi.default_params(gravity=(0.0, 0.0, 0.0)), precision=32)
    o.add_bodies(pair_scene(ga, gb))
    o.collide()
    assert (len(o.pairs()) == 1) == expect
    # the rule is symmetric
    o2 = oracle_mod.OracleWorld(abi.default_params(gravity=(0.0, 0.0, 0.0)), precision=32)
    o2.add_bodies(pair_scene(gb, ga))
    o2.collide()
    assert (len(o2.pairs()) == 1) == expect


def test_groups_example_bodies_rest_on_their_floors(oracle_mod):
    s = scenes.collision_groups_example(num=4, layers=4)
    o = oracle_mod.OracleWorld(abi.default_params(), precision=32)
    o.add_bodies(s)
    ns = scenes.num_static(s)
    green = (s["collision_groups"][ns:, 0] == 1)
    member = s["collision_groups"][:, 0]
    for k in range(150):
        o.step(1)
        if k % 10 == 9:
            p = o.pairs()
            m1, m2 = member[p[:, 0]], member[p[:, 1]]
            assert not np.any(((m1 == 1) & (m2 == 2)) | ((m1 == 2) & (m2 == 1))), "a green body paired with a blue one"
    y = o.get_state()["pos"][ns:, 1]
    h = 0.1  # box half extent; floor tops at 1.1 (green) and 2.1 (blue)
    # the lower layer of each colour rests on its floor (within the margins' gap), the upper one on it
    assert np.all(y[green] > 1.1 + h - 0.005) and np.all(y[green] < 1.1 + 3 * h + 2 * 0.02 + 0.01)
    assert np.all(y[~green] > 2.1 + h - 0.005) and np.all(y[~green] < 2.1 + 3 * h + 2 * 0.02 + 0.01)
    assert np.abs(o.get_state()["linvel"][ns:]).max() < 0.05


def test_changed_groups_refilter_the_pairs(oracle_mod):
    """Collider::set_collision_groups between steps: a ball resting on a floor falls freely from
    the next step on once the floor stops accepting its group (v = -g t exactly), and lands again
    on a lower floor that accepts it."""
    s = merge(body(status=1, shape=1, dims=(2.0, 0.1, 2.0), pos=(0.0, -0.1, 0.0)),
              body(status=1, shape=1, dims=(2.0, 0.1, 2.0), pos=(0.0, -3.1, 0.0)),
              body(shape=0, dims=(0.25, 0, 0), pos=(0.0, 0.25, 0.0), collision_groups=cg([1])))
    o = oracle_mod.OracleWorld(abi.default_params(), precision=64)
    o.add_bodies(s)
    o.step(30)
    assert abs(o.get_state()["linvel"][2, 1]) < 1e-6
    o.set_collision_groups([0], [cg([0], [0])])  # the upper floor now accepts group 0 only
    dt = 1.0 / 60.0
    for k in range(1, 11):
        o.step(1)
        assert len(o.pairs()) == 0  # the lower floor is still out of reach
        np.testing.assert_allclose(o.get_state()["linvel"][2, 1], -9.81 * dt * k, rtol=1e-12)
    o.step(80)
    y = o.get_state()["pos"][2, 1]
    assert abs(y - (-3.0 + 0.25)) < 0.03  # resting on the lower floor
