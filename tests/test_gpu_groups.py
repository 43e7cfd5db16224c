"""Collision groups on the GPU against the CPU restatement: the pair rule on every case of the
truth table, and the reference's collision groups example (examples3d/collision_groups3.rs)
stepped bit for bit in the colour order, its pair sets equal at every checked step."""
import numpy as np
import pytest

from nphysics_amd import abi, scenes
from tests.test_oracle_groups import TRUTH, pair_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", list(TRUTH))
def test_group_pairs_match(nphys, oracle_mod, case):
    ga, gb, expect = TRUTH[case]
    s = pair_scene(ga, gb)
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params(gravity=(0.0, 0.0, 0.0)))
    g.add_bodies(s)
    o = oracle_mod.OracleWorld(abi.default_params(gravity=(0.0, 0.0, 0.0)), precision=32)
    o.add_bodies(s)
    g.collide()
    o.collide()
    np.testing.assert_array_equal(g.pairs(), o.pairs())
    assert (len(g.pairs()) == 1) == expect


def test_groups_example_bitexact(nphys, oracle_mod):
    s = scenes.collision_groups_example()
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params())
    g.add_bodies(s)
    o = oracle_mod.OracleWorld(abi.default_params(), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(s)
    for k in range(80):
        g.step(1)
        o.step(1)
        if k % 10 == 9:
            print(f"groups example: step {k + 1}, pairs {len(g.pairs())}", flush=True)
            np.testing.assert_array_equal(g.pairs(), o.pairs())
            gs, os_ = g.get_state(), o.get_state()
            for key in ("pos", "rot", "linvel", "angvel"):
                np.testing.assert_array_equal(gs[key], os_[key].astype(np.float32), err_msg=f"step {k + 1} {key}")
    ns = scenes.num_static(s)
    green = s["collision_groups"][ns:, 0] == 1
    y = g.get_state()["pos"][ns:, 1]
    assert np.all(y[green] < 1.6) and np.all(y[~green] > 2.15)


def test_changed_groups_bitexact(nphys, oracle_mod):
    """Collision groups changed between steps (the blue boxes of the example become green, then
    the green floor stops accepting them): pairs and states stay bit-exact."""
    s = scenes.collision_groups_example(num=6)
    g = nphys.MechanicalWorld(integration_parameters=abi.default_params())
    g.add_bodies(s)
    o = oracle_mod.OracleWorld(abi.default_params(), precision=32, order=oracle_mod.ORDER_COLOURED)
    o.add_bodies(s)
    ns = scenes.num_static(s)
    blue = ns + np.nonzero(s["collision_groups"][ns:, 0] == 2)[0]  # the blue boxes (not the floor)
    changes = {30: (blue, [abi.collision_groups([0], [0])] * len(blue)),
               60: ([1], [abi.collision_groups([2], [2])])}
    for k in range(90):
        if k in changes:
            bodies, groups = changes[k]
            g.set_collision_groups(bodies, groups)
            o.set_collision_groups(bodies, groups)
        g.step(1)
        o.step(1)
        if k % 10 == 9 or k in (30, 60):
            np.testing.assert_array_equal(g.pairs(), o.pairs())
            gs, os_ = g.get_state(), o.get_state()
            for key in ("pos", "rot", "linvel", "angvel"):
                np.testing.assert_array_equal(gs[key], os_[key].astype(np.float32), err_msg=f"step {k + 1} {key}")
    # no box pairs with the coloured floors any more: all fall to the main floor
    p = g.pairs()
    assert not np.isin(p, [1, 2]).any()
    assert np.all(g.get_state()["linvel"][ns:, 1] < 1.0)
