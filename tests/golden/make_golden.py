"""Writes the golden fixtures of tests/golden/ from the CPU restatement (oracle/).

The reference (nphysics, Rust) cannot be built or run in this image and holds no vectors for
the contact pipeline, so these fixtures are regression vectors of the restatement, after the
restatement itself has been pinned by the known-answer and closed-form tests of
tests/test_oracle_kat.py. They freeze, for three small scenes:

  * the inputs (the scene arrays, the integration parameters);
  * after `steps` FP32 steps in the coloured order the GPU replays: body states, the sorted
    broad-phase pairs, the contacts (bodies, feature, normal, depth) and the solved impulses
    with their keys -- the GPU must reproduce these bit for bit;
  * positions and velocities of the same steps in FP64 in the same order (the error reference:
    nphysics is generic over RealField), in FP64 and in FP32 in the reference's own row order
    (what the reference's own FP32 error is at equal iteration count).

Also written: shape-pair cases of the narrow phase (inputs and FP64 outputs).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nphysics_amd import abi, scenes  # noqa: E402
from oracle import oracle  # noqa: E402

SCENE_KEYS = ("status", "shape", "dims", "pos", "rot", "linvel", "angvel", "density", "margin",
              "linear_prediction", "friction", "restitution")


def golden_scenes():
    """(name, scene, params, steps) of the fixtures."""
    out = []
    s = scenes.balls(seed=11, nx=4, ny=4, nz=4)
    out.append(("balls64", s, abi.default_params(max_velocity_iterations=8), 30))
    s = scenes.boxes(seed=12, nx=3, ny=3, nz=3)
    out.append(("boxes27", s, abi.default_params(max_velocity_iterations=10), 30))
    s = scenes.small_pile(seed=13, n=64, lattice=(4, 4, 4), extent=2.5, y0=0.3)
    out.append(("pile64", s, abi.default_params(max_velocity_iterations=16), 45))
    return out


def scene_fixture(name, scene, params, steps):
    f = {f"in_{k}": np.asarray(scene[k]) for k in SCENE_KEYS}
    f["params_dt"] = np.float64(params.dt)
    f["params_velocity_iterations"] = np.int32(params.max_velocity_iterations)
    f["params_position_iterations"] = np.int32(params.max_position_iterations)
    f["steps"] = np.int32(steps)

    o32 = oracle.OracleWorld(params, precision=32, order=oracle.ORDER_COLOURED)
    o32.add_bodies(scene)
    o32.step(steps)
    st = o32.get_state()
    for k, v in st.items():
        f[f"f32_{k}"] = v.astype(np.float32)
    f["f32_pairs"] = o32.pairs()
    c = o32.contacts()
    f["f32_contact_bodies"] = c["bodies"]
    f["f32_contact_feature"] = c["feature"]
    f["f32_contact_normal"] = c["normal"].astype(np.float32)
    f["f32_contact_depth"] = c["depth"].astype(np.float32)
    keys, imp = o32.impulses()
    f["f32_impulse_keys"] = keys
    f["f32_impulses"] = imp.astype(np.float32)
    f["f32_residual"] = np.array(o32.residual())

    # FP64 and FP32 runs for the error report: same colour order (the precision error alone) and
    # the reference's own row order
    for tag, prec, order in (("f64col", 64, oracle.ORDER_COLOURED), ("f64ref", 64, oracle.ORDER_REFERENCE),
                             ("f32ref", 32, oracle.ORDER_REFERENCE)):
        o = oracle.OracleWorld(params, precision=prec, order=order)
        o.add_bodies(scene)
        o.step(steps)
        for k in ("pos", "linvel"):
            f[f"{tag}_{k}"] = o.get_state()[k]
        f[f"{tag}_residual"] = np.array(o.residual())
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **f)
    return f


def shape_cases():
    """Narrow-phase cases: (kind1, dims1, pose1, kind2, dims2, pose2) with poses (x,y,z,i,j,k,w)."""
    q = lambda ax, a: np.concatenate([np.sin(a / 2) * np.asarray(ax, float) / np.linalg.norm(ax), [np.cos(a / 2)]])
    ident = np.array([0, 0, 0, 1.0])
    cases = [
        (0, [0.5, 0, 0], [0, 0, 0, *ident], 0, [0.5, 0, 0], [0.9, 0.1, 0, *ident]),
        (0, [0.3, 0, 0], [0, 0.75, 0, *ident], 1, [1, 0.5, 1], [0, 0, 0, *ident]),
        (1, [1, 0.5, 1], [0, 0, 0, *ident], 0, [0.3, 0, 0], [0.2, 0.7, -0.1, *ident]),
        (0, [0.3, 0, 0], [1.1, 0.6, 1.05, *ident], 1, [1, 0.5, 1], [0, 0, 0, *ident]),
        (1, [0.5, 0.5, 0.5], [0, 0.98, 0, *ident], 1, [2, 0.5, 2], [0, 0, 0, *ident]),
        (1, [0.5, 0.5, 0.5], [0.1, 0.95, -0.2, *q([0, 1, 0], 0.6)], 1, [2, 0.5, 2], [0, 0, 0, *ident]),
        (1, [0.4, 0.3, 0.2], [0.3, 0.62, 0.1, *q([1, 1, 0], 0.5)], 1, [0.5, 0.3, 0.5], [0, 0, 0, *q([0, 0, 1], 0.2)]),
        (1, [0.5, 0.5, 0.5], [0.0, 1.6, 0.0, *q([1, 0, 1], 0.9)], 1, [0.5, 0.5, 0.5], [0, 0, 0, *q([0, 1, 0], 0.785)]),
        (1, [0.5, 0.5, 0.5], [0, 1.2, 0, *ident], 1, [0.5, 0.5, 0.5], [0, 0, 0, *ident]),
    ]
    return cases


def shapes_fixture():
    f = {}
    cases = shape_cases()
    f["ncases"] = np.int32(len(cases))
    for i, (k1, d1, p1, k2, d2, p2) in enumerate(cases):
        r = oracle.collide_shapes(k1, d1, p1, k2, d2, p2, prediction=0.022, precision=64)
        f[f"c{i}_in"] = np.array([k1, *d1, *p1, k2, *d2, *p2], np.float64)
        f[f"c{i}_n"] = np.int32(r["n"])
        for k in ("feature", "world1", "world2", "normal", "depth", "kind"):
            f[f"c{i}_{k}"] = r[k]
    np.savez_compressed(os.path.join(HERE, "shapes.npz"), **f)


def main():
    oracle.build()
    for name, s, p, steps in golden_scenes():
        f = scene_fixture(name, s, p, steps)
        print(f"{name}: {s['status'].shape[0]} bodies, {steps} steps, {f['f32_contact_depth'].shape[0]} "
              f"contacts, {f['f32_pairs'].shape[0]} pairs")
    shapes_fixture()
    print("shapes: written")


if __name__ == "__main__":
    main()
