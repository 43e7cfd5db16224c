"""The benchmark workloads at their full sizes, checked through properties that do not depend on
the size (the CPU restatement checks them bit for bit at the sizes it finishes in seconds):
the dense pile (262,144 bodies) after most of it has landed, and the 4,096 batched worlds.

* pair lists: sorted, unique, no self pairs, every pair with a dynamic body, indices in range, and
  (batched) both bodies in the same scene;
* contacts: finite, unit normals (up to the drift of unrenormalised rotations);
* impulses: lambda_n >= 0 and |lambda_t| <= mu lambda_n (the projections of the solve);
* states: finite, every body of the pile inside its box;
* determinism: a second world stepped the same way holds the same bits."""
import numpy as np
import pytest

from nphysics_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _world(nphys, s, **params):
    w = nphys.MechanicalWorld(integration_parameters=abi.default_params(**params))
    w.add_bodies(s)
    return w


def _check_pairs(s, pairs):
    n = s["status"].shape[0]
    assert pairs.shape[0] > 0
    assert np.all((pairs >= 0) & (pairs < n))
    assert np.all(pairs[:, 0] != pairs[:, 1])
    key = pairs[:, 0].astype(np.int64) * n + pairs[:, 1]
    assert np.all(np.diff(key) > 0), "pairs sorted and unique"
    dyn = s["status"] == abi.STATUS_DYNAMIC
    assert np.all(dyn[pairs[:, 0]] | dyn[pairs[:, 1]])


def _check_contacts_and_impulses(w, mu):
    c = w.contacts()
    assert c["depth"].shape[0] > 0
    for k in ("world1", "world2", "normal", "depth"):
        assert np.all(np.isfinite(c[k])), k
    # unit up to the drift of the bodies' quaternions: box faces give the normals, and the reference
    # does not renormalise rotations (rigid_body.rs:309-310), so |q| drifts by about 1e-4 in 400 steps
    np.testing.assert_allclose(np.linalg.norm(c["normal"], axis=1), 1.0, atol=2e-3)
    _, imp = w.impulses()
    assert imp.shape[0] > 0
    assert np.all(imp[:, 0] >= 0.0)
    bound = mu * imp[:, 0] * (1 + 1e-6) + 1e-12
    assert np.all(np.abs(imp[:, 1]) <= bound) and np.all(np.abs(imp[:, 2]) <= bound)


def _same_bits(a, b):
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    np.testing.assert_array_equal(a.pairs(), b.pairs())


def test_dense_pile_full_size(nphys):
    s = scenes.pile()
    params = dict(max_velocity_iterations=16, max_position_iterations=3)
    a = _world(nphys, s, **params)
    b = _world(nphys, s, **params)
    for k in range(0, 400, 100):
        a.step(100)
        b.step(100)
        print(f"dense pile: step {k + 100}, contacts {a.counters()['ncontacts']}", flush=True)
    _same_bits(a, b)
    _check_pairs(s, a.pairs())
    _check_contacts_and_impulses(a, 0.5)
    st = a.get_state()
    for k in st:
        assert np.all(np.isfinite(st[k])), k
    dyn = s["status"] == abi.STATUS_DYNAMIC
    p = st["pos"][dyn]
    assert np.all(np.abs(p[:, 0]) < 31.0) and np.all(np.abs(p[:, 2]) < 31.0) and np.all(p[:, 1] > -1.0)
    assert a.counters()["ncontacts"] > 100000
    rinf, rrms = a.residual()
    assert np.isfinite(rinf) and np.isfinite(rrms)


def test_batched_4096_worlds_full_size(nphys):
    s = scenes.batched()
    params = dict(max_velocity_iterations=8, max_position_iterations=3)
    a = _world(nphys, s, **params)
    b = _world(nphys, s, **params)
    for k in range(0, 150, 50):
        a.step(50)
        b.step(50)
        print(f"batched: step {k + 50}, contacts {a.counters()['ncontacts']}", flush=True)
    _same_bits(a, b)
    pairs = a.pairs()
    _check_pairs(s, pairs)
    assert np.all(s["scene"][pairs[:, 0]] == s["scene"][pairs[:, 1]]), "pairs across scenes"
    _check_contacts_and_impulses(a, 0.5)
    st = a.get_state()
    for k in st:
        assert np.all(np.isfinite(st[k])), k
