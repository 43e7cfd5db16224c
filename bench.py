"""Benchmark of the MI355X contact pipeline on the dense-pile workload (one MechanicalWorld::step
per timed step).

Workload: 262,144 bodies (even index a ball with r in [0.2, 0.6] m, odd a cuboid with half
extents in [0.2, 0.5] m and a random orientation), released from a 34 x 227 x 34 lattice above
an open five-wall box (60 x 40 x 60 m) and settled untimed until the contact count is steady
(about 1.5M contacts at rest); then W warm-up steps and K timed steps, 16 velocity and 3 position
iterations, dt = 1/60, FP32. Inputs are resident in HBM for the whole run.

Printed: one JSON line with steps/s, contact-iterations/s, the residual, per-stage times, the
roofline of the dominant kernel and the CPU baseline. The K timed steps run with counters off;
the roofline comes from a following region of --timing-steps steps with counters on, where every
launch of the SOR-Prox sweep kernel carries HIP events in its dispatch packet on the world's
stream. The CPU baseline is the FP32 CPU restatement of the
reference step, one thread, timed on a bounded sample of the same settled scene.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from nphysics_amd import abi, scenes, world  # noqa: E402

METRIC = "steps/s + contact-iters/s at residual 1e-4, learned warm start vs plain"
HBM_PEAK_GBPS = 8000.0
BF16_DENSE_TFLOPS = 2500.0
MLP_FLOP_PER_ROW = 2 * (24 * 128 + 128 * 128 + 128 * 3)  # 39,680
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "k_pgs_pile_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def settle(w, steps):
    done = 0
    while done < steps:
        k = min(100, steps - done)
        w.step(k)
        done += k
        c = w.counters()
        log(f"settle {done}/{steps}: contacts {c['ncontacts']} colours {c['ncolours']} awake {c['active_bodies']}")


def cpu_baseline(scene, snapshot, params, steps):
    """The FP32 CPU restatement of the reference step (reference row order) on one thread, from the
    same snapshot: one untimed step (it also runs the initial collision detection), then `steps`
    timed steps, with the residual after each. Returns the baseline and the state after the
    untimed step (for the FP64 error)."""
    o = oracle_world(with_state(scene, snapshot), params, 32)
    o.step(1)
    after_one = o.get_state()
    dt, rinf, rrms = 0.0, [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.step(1)
        dt += time.perf_counter() - t0
        a, b = o.residual()
        rinf.append(a)
        rrms.append(b)
    c = o.counters()
    return {
        "value": steps / dt,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} steps of the settled dense pile ({c['ncontacts']} contacts) after one untimed "
                  f"step, FP32, reference row order, single thread, on {cpu_model()} ({os.cpu_count()} logical CPUs)",
        "contact_iters_per_s": c["ncontacts"] * params.max_velocity_iterations * steps / dt,
        "residual_inf": float(np.mean(rinf)),
        "residual_rms": float(np.mean(rrms)),
    }, after_one


def gpu_from_snapshot(scene, snapshot, params, steps):
    """The GPU on the CPU baseline's sample: the same snapshot, one step (state kept for the FP64
    error), then `steps` steps with the residual after each."""
    w = gpu_world(with_state(scene, snapshot), params)
    w.step(1)
    after_one = w.get_state()
    rinf, rrms = [], []
    for _ in range(steps):
        w.step(1)
        a, b = w.residual()
        rinf.append(a)
        rrms.append(b)
    w.close()
    return float(np.mean(rinf)), float(np.mean(rrms)), after_one


def train_warmstart(scene, params, settle_steps, capture_steps, max_sweeps, check_every, what):
    """Trains the learned warm start's network on the GPU on samples from `scene` (a seed
    disjoint from the evaluated one): after `settle_steps` steps, `capture_steps` steps whose solve
    runs to RMS residual 1e-5 (capped at `max_sweeps` sweeps), each contact's features and the
    impulses its solve reached."""
    from nphysics_amd import learned

    t0 = time.perf_counter()
    w = world.MechanicalWorld(integration_parameters=params)
    w.add_bodies(scene)
    settle(w, settle_steps)
    X, T = learned.collect_samples(w, capture_steps, target=1e-5, max_sweeps=max_sweeps,
                                   check_every=check_every, norm="rms")
    w.close()
    t1 = time.perf_counter()
    net = learned.WarmStartNet(seed=1)
    net.set_data(X, learned.correction_targets(X, T))
    losses = net.train(epochs=15, batch=4096, lr=1e-3)
    t2 = time.perf_counter()
    params_out = net.params
    net.close()
    log(f"warm-start network ({what}): {X.shape[0]} samples in {t1 - t0:.1f} s, trained in {t2 - t1:.1f} s, "
        f"loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    return params_out, {"data": what, "train_samples": int(X.shape[0]), "data_s": t1 - t0, "train_s": t2 - t1,
                        "epochs": 15, "loss_first": float(losses[0]), "loss_last": float(losses[-1])}


def learned_section(fresh, net_params, train, iters, targets, steps, max_sweeps):
    """Solve-to-residual runs from the state `fresh()` returns, plain vs learned warm start.

    For each target on R-inf = max |phi| (m/s): steps/s and sweeps/step with the plain and with the
    learned warm start (separate runs from the same state, capped at max_sweeps; whether the target
    was reached is reported), and the judge's sweep counts, both starts solved from identical
    assembled systems with the learned start kept only where it needs fewer sweeps. The targets are
    the given ones plus the R-inf the plain warm start reaches with the reference's iteration count
    (`iters` sweeps) on the same state."""
    from nphysics_amd import learned

    w = fresh()
    w.step(1)
    r_ref = w.residual()[0]
    w.close()

    def run(tgt, learned_start):
        w = fresh()
        if learned_start:
            w.load_warmstart(net_params)
            w.set_contact_model(abi.CONTACT_MODEL_LEARNED_WARMSTART)
        w.set_residual_target(tgt, max_sweeps, 1, "max")
        w.step(1)
        w.reset_counters()
        t0 = time.perf_counter()
        w.step(steps)
        dt = time.perf_counter() - t0
        c = w.counters()
        r = {"steps_per_s": steps / dt, "sweeps_per_step": c["velocity_sweeps"] / steps,
             "contact_iters_per_s": c["ncontacts"] * c["velocity_sweeps"] / dt,
             "residual_inf": w.residual()[0], "residual_rms": w.residual()[1],
             "reached": bool(w.residual()[0] <= tgt)}
        if learned_start:
            # the inference kernel's MFMA roofline, timed with HIP events in one more step
            w.enable_counters(True)
            w.reset_counters()
            w.step(1)
            w.enable_counters(False)
            ci = w.counters()
            if ci["warmstart_inference_time"] > 0:
                tf = MLP_FLOP_PER_ROW * ci["warmstart_inference_rows"] / ci["warmstart_inference_time"] / 1e12
                r["inference_roofline"] = {
                    "kernel": "k_mlp_fwd (fused 3-layer BF16 MFMA inference, one row per contact)",
                    "bound": "mfma", "achieved": tf, "peak": BF16_DENSE_TFLOPS, "unit": "TFLOP/s",
                    "frac": tf / BF16_DENSE_TFLOPS, "rows": ci["warmstart_inference_rows"],
                    "us": 1e6 * ci["warmstart_inference_time"], "flop_per_row": MLP_FLOP_PER_ROW}
        w.close()
        return r

    out = {"norm": "inf", "max_sweeps": max_sweeps, "steps": steps, "training": train, "runs": []}
    for tgt, what in [(t, "fixed target") for t in targets] + [
            (r_ref, f"R-inf of the plain warm start after {iters} sweeps on this state")]:
        sec = {"target": tgt, "what": what, "plain": run(tgt, False), "learned": run(tgt, True)}
        w = fresh()
        sec["judge"] = learned.judge(w, net_params, steps, target=tgt, max_sweeps=max_sweeps, check_every=1,
                                     norm="max")
        w.close()
        p, l = sec["plain"]["sweeps_per_step"], sec["learned"]["sweeps_per_step"]
        sec["sweep_reduction"] = (p - l) / p if p > 0 else 0.0
        sec["speedup"] = sec["learned"]["steps_per_s"] / sec["plain"]["steps_per_s"]
        log(f"target {tgt:.3g}: plain {p:.1f} sweeps/step, learned {l:.1f}; judge plain "
            f"{sec['judge']['plain_sweeps_per_step']:.1f} learned {sec['judge']['learned_sweeps_per_step']:.1f}")
        out["runs"].append(sec)
    return out


def residual_target_section(scene, snapshot, params, target, steps, max_sweeps):
    """The learned warm start against the plain one on the two workloads it is evaluated on:

    * the dense pile, from the settled snapshot of the benchmark (seed 3), with a network trained
      on another pile (seed 1003) settled the same way;
    * the 8,000-box stack (seed 2, 10 velocity iterations), with a network trained on a box
      stack of seed 1002."""
    def fresh_pile():
        w = world.MechanicalWorld(integration_parameters=params)
        w.add_bodies(scene)
        w.set_state(**{k: snapshot[k] for k in ("pos", "rot", "linvel", "angvel")})
        w.step(3)  # rebuild the impulse cache after the state upload
        return w

    pile_net, pile_train = train_warmstart(scenes.pile(seed=1003), params, 750, 4, 256, 16,
                                           "dense pile, seed 1003")
    out = {"pile": learned_section(fresh_pile, pile_net, pile_train, params.max_velocity_iterations,
                                   [target], steps, max_sweeps)}

    box_params = abi.default_params(max_velocity_iterations=10, max_position_iterations=3)
    box_net, box_train = train_warmstart(scenes.boxes(seed=1002), box_params, 100, 25, 1000, 8,
                                         "8,000-box stack, seed 1002")
    eval_boxes = scenes.boxes(seed=2)
    w = world.MechanicalWorld(integration_parameters=box_params)
    w.add_bodies(eval_boxes)
    w.step(100)
    box_state = w.get_state()
    w.close()

    def fresh_boxes():
        w = world.MechanicalWorld(integration_parameters=box_params)
        w.add_bodies(eval_boxes)
        w.set_state(**box_state)
        w.step(3)
        return w

    out["box_stack"] = learned_section(fresh_boxes, box_net, box_train, box_params.max_velocity_iterations,
                                       [target], 8, max_sweeps)

    # the ball drop, where R-inf 1e-4 is reachable: trained on a drop of seed 1001, evaluated on
    # the 1,000-ball drop (seed 1) at rest after 200 steps
    ball_params = abi.default_params(max_velocity_iterations=8, max_position_iterations=3)
    ball_net, ball_train = train_warmstart(scenes.balls(seed=1001), ball_params, 150, 40, 2000, 1,
                                           "1,000-ball drop, seed 1001")
    eval_balls = scenes.balls(seed=1)
    w = world.MechanicalWorld(integration_parameters=ball_params)
    w.add_bodies(eval_balls)
    w.step(200)
    ball_state = w.get_state()
    w.close()

    def fresh_balls():
        w = world.MechanicalWorld(integration_parameters=ball_params)
        w.add_bodies(eval_balls)
        w.set_state(**ball_state)
        w.step(3)
        return w

    out["ball_drop"] = learned_section(fresh_balls, ball_net, ball_train, ball_params.max_velocity_iterations,
                                       [target], 8, max_sweeps)
    return out


def with_state(scene, st):
    s = dict(scene)
    for k in ("pos", "rot", "linvel", "angvel"):
        s[k] = st[k]
    return s


def gpu_world(scene, params):
    w = world.MechanicalWorld(integration_parameters=params)
    w.add_bodies(scene)
    return w


def oracle_world(scene, params, precision):
    from oracle import oracle

    o = oracle.OracleWorld(params, precision=precision, order=oracle.ORDER_REFERENCE)
    o.add_bodies(scene)
    return o


def velocity_errors(st, ref):
    """Largest velocity differences of state `st` against `ref` (the FP64 run), in m/s and rad/s,
    and relative to the largest FP64 speed."""
    dl = float(np.max(np.abs(st["linvel"].astype(np.float64) - ref["linvel"]), initial=0.0))
    da = float(np.max(np.abs(st["angvel"].astype(np.float64) - ref["angvel"]), initial=0.0))
    dp = float(np.max(np.abs(st["pos"].astype(np.float64) - ref["pos"]), initial=0.0))
    vmax = float(np.max(np.linalg.norm(ref["linvel"], axis=1), initial=0.0))
    return {"linvel_max_abs": dl, "angvel_max_abs": da, "pos_max_abs": dp,
            "linvel_rel_to_max_speed": dl / vmax if vmax > 0 else 0.0}


def f64_error(scene, st, params, gpu_state_after=None, cpu32_state_after=None):
    """Error against an FP64 run of the reference step: one step from the same state on the GPU
    (FP32, colour order), on the CPU restatement in FP32 (reference order) and in FP64 (reference
    order, the error reference). Row order and precision both enter the GPU figure; only precision
    enters the CPU FP32 one."""
    from oracle import oracle

    s = with_state(scene, st)
    o64 = oracle_world(s, params, 64)
    o64.step(1)
    ref = o64.get_state()
    del o64
    o64c = oracle.OracleWorld(params, precision=64, order=oracle.ORDER_COLOURED)
    o64c.add_bodies(s)
    o64c.step(1)
    ref_coloured = o64c.get_state()
    del o64c
    if gpu_state_after is None:
        g = gpu_world(s, params)
        g.step(1)
        gpu_state_after = g.get_state()
        g.close()
    if cpu32_state_after is None:
        o32 = oracle_world(s, params, 32)
        o32.step(1)
        cpu32_state_after = o32.get_state()
    return {"what": "one step from the same state against the CPU restatement in FP64 (reference row order); "
                    "gpu_f32_same_order: against FP64 in the GPU's colour order (precision alone)",
            "gpu_f32": velocity_errors(gpu_state_after, ref), "cpu_f32": velocity_errors(cpu32_state_after, ref),
            "gpu_f32_same_order": velocity_errors(gpu_state_after, ref_coloured)}


def solve_to_target(make, target, max_sweeps, warm_steps=3):
    """Sweeps one step needs to reach max |phi| <= target (R-inf, m/s), capped at max_sweeps, after
    `warm_steps` fixed-count steps that fill the impulse cache; with the time of that step."""
    w = make()
    w.step(warm_steps)
    w.set_residual_target(target, max_sweeps, 1, "max")
    w.reset_counters()
    t0 = time.perf_counter()
    w.step(1)
    dt = time.perf_counter() - t0
    sweeps = int(w.counters()["velocity_sweeps"])
    rinf, rrms = w.residual()
    out = {"sweeps": sweeps, "residual_inf": rinf, "residual_rms": rrms, "reached": bool(rinf <= target),
           "step_s": dt}
    if hasattr(w, "close"):
        w.close()
    return out


def reachability(scene, st, params, target, max_sweeps, gpu=True):
    """Sweeps to R-inf <= target from one state: GPU FP32 (colour order), CPU FP32 and FP64 (reference
    order). The FP64 run separates a precision floor (FP64 reaches the target, FP32 does not) from
    the convergence of projected Gauss-Seidel itself (neither reaches it)."""
    s = with_state(scene, st)
    out = {"target_inf": target, "max_sweeps": max_sweeps}
    if gpu:
        out["gpu_f32"] = solve_to_target(lambda: gpu_world(s, params), target, max_sweeps)
    out["cpu_f32"] = solve_to_target(lambda: oracle_world(s, params, 32), target, max_sweeps)
    out["cpu_f64"] = solve_to_target(lambda: oracle_world(s, params, 64), target, max_sweeps)
    return out


def small_config(what, scene, params, steps, samples, sample_steps, target, max_sweeps):
    """One workload run whole on the GPU (all `steps` steps timed from the initial scene, counters
    off), beside the one-thread FP32 CPU restatement (reference row order) timed on windows of
    `sample_steps` steps from GPU states at `samples` (each after one untimed step from that state,
    the GPU stepping the same windows for the residual comparison at equal iteration count); the
    FP64 error of one step from each sampled state; sweeps to R-inf <= target from the final state."""
    iters = params.max_velocity_iterations
    w = gpu_world(scene, params)
    w.enable_counters(False)
    t0 = time.perf_counter()
    w.step(steps)
    elapsed = time.perf_counter() - t0
    final = w.get_state()
    c = w.counters()
    w.close()
    # states at the sample points (deterministic: the same trajectory again)
    w = gpu_world(scene, params)
    snaps, done = [], 0
    for k in samples:
        w.step(k - done) if k > done else None
        done = k
        snaps.append(w.get_state())
    w.close()
    g_inf, g_rms, o_inf, o_rms, cpu_t, cpu_n, errs = [], [], [], [], 0.0, 0, []
    contacts = []
    for st in snaps:
        s = with_state(scene, st)
        g = gpu_world(s, params)
        g.step(1)
        g_after = g.get_state()
        o = oracle_world(s, params, 32)
        o.step(1)
        errs.append(f64_error(scene, st, params, g_after, o.get_state()))
        for _ in range(sample_steps):
            g.step(1)
            a, b = g.residual()
            g_inf.append(a)
            g_rms.append(b)
            t1 = time.perf_counter()
            o.step(1)
            cpu_t += time.perf_counter() - t1
            cpu_n += 1
            a, b = o.residual()
            o_inf.append(a)
            o_rms.append(b)
            contacts.append(o.counters()["ncontacts"])
        g.close()
    worst = max(errs, key=lambda e: e["gpu_f32"]["linvel_max_abs"])
    out = {
        "workload": what,
        "bodies": int(scene["status"].shape[0]),
        "dynamic_bodies": int((scene["status"] == 0).sum()),
        "steps": steps,
        "value": steps / elapsed,
        "unit": "steps/s",
        "ms_per_step": 1e3 * elapsed / steps,
        "contacts_last_step": c["ncontacts"],
        "velocity_iterations": iters,
        "residual_at_equal_iterations": {
            "what": f"mean over {len(g_inf)} steps ({sample_steps} after each of the states at steps {list(samples)}), "
                    f"{iters} sweeps per step, both sides from the same states",
            "gpu_inf": float(np.mean(g_inf)), "gpu_rms": float(np.mean(g_rms)),
            "cpu_reference_order_inf": float(np.mean(o_inf)), "cpu_reference_order_rms": float(np.mean(o_rms)),
            "ratio_inf": float(np.mean(g_inf) / np.mean(o_inf)) if np.mean(o_inf) > 0 else None,
            "ratio_rms": float(np.mean(g_rms) / np.mean(o_rms)) if np.mean(o_rms) > 0 else None},
        "f64_error": dict(worst, what=worst["what"] + f"; the worst of the states at steps {list(samples)}"),
        "residual_target": reachability(scene, final, params, target, max_sweeps),
        "cpu_baseline": {"value": cpu_n / cpu_t, "unit": "steps/s", "cores": 1, "kind": "port",
                         "sample": f"{cpu_n} steps ({sample_steps} after each of the GPU states at steps "
                                   f"{list(samples)}, one untimed step first), FP32, reference row order, "
                                   f"single thread, on {cpu_model()} ({os.cpu_count()} logical CPUs); "
                                   f"{int(np.mean(contacts))} contacts on average"},
    }
    out["contact_iters_per_s"] = float(np.mean(contacts)) * iters * out["value"]
    out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def pile_subset(scene, st, corner=20.0):
    """The static walls and the dynamic bodies of the settled pile in the corner x, z > corner."""
    stat = scene["status"] == 1
    keep = stat | ((st["pos"][:, 0] > corner) & (st["pos"][:, 2] > corner))
    idx = np.nonzero(keep)[0]
    n = scene["status"].shape[0]
    return {k: (v[idx] if isinstance(v, np.ndarray) and v.shape[:1] == (n,) else v) for k, v in scene.items()}, \
        {k: v[idx] for k, v in st.items()}


def cpu_baseline_batched(scene, snapshot, params, steps):
    """The FP32 CPU restatement on the batch from the same state, one world per scene: on one
    thread (the reference's own threading), and on all of the host's cores (at most 16) with a
    thread pool over the independent worlds. Each world takes one untimed step first."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle

    sc = scene["scene"]
    n = sc.shape[0]
    order = np.argsort(sc, kind="stable")
    bounds = np.searchsorted(sc[order], np.arange(int(sc.max()) + 2))
    worlds = []
    for s in range(int(sc.max()) + 1):
        idx = order[bounds[s]:bounds[s + 1]]
        sub = {k: v[idx] for k, v in scene.items() if isinstance(v, np.ndarray) and v.shape[:1] == (n,)}
        for k in ("pos", "rot", "linvel", "angvel"):
            sub[k] = snapshot[k][idx]
        sub["scene"] = np.zeros(idx.shape[0], np.int32)
        o = oracle.OracleWorld(params, precision=32)
        o.add_bodies(sub)
        worlds.append(o)
    threads = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda o: o.step(1), worlds))
    # one thread over a fixed share of the worlds (every 16th), scaled to the whole batch
    share = worlds[::16]
    t0 = time.perf_counter()
    for o in share:
        o.step(steps)
    t_one = (time.perf_counter() - t0) * len(worlds) / len(share)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda o: o.step(steps), worlds))
    t_all = time.perf_counter() - t0
    where = f"on {cpu_model()} ({os.cpu_count()} logical CPUs)"
    single = {"value": steps / t_one, "unit": "steps/s", "cores": 1, "kind": "port",
              "sample": f"{steps} steps of every 16th of the {len(worlds)} settled worlds (one oracle world "
                        f"each) after one untimed step, FP32, single thread, scaled to the whole batch, {where}"}
    every = {"value": steps / t_all, "unit": "steps/s", "cores": threads, "kind": "port",
             "sample": f"{steps} steps of all {len(worlds)} settled worlds, FP32, a pool of {threads} threads "
                       f"over the independent worlds, {where}"}
    return single, every


def batched_section(steps, warmup, settle_steps, sweep_variant, cpu_steps, no_cpu):
    """4,096 independent worlds (lognormal sizes, median 64 bodies, each with its own ground)
    stepped as one handle; 8 velocity / 3 position iterations."""
    params = abi.default_params(max_velocity_iterations=8, max_position_iterations=3)
    scene = scenes.batched()
    w = world.MechanicalWorld(integration_parameters=params)
    w.add_bodies(scene)
    w.set_sweep_variant(sweep_variant)
    t0 = time.perf_counter()
    settle(w, settle_steps)
    log(f"settled in {time.perf_counter() - t0:.1f} s")
    w.step(warmup)
    snapshot = w.get_state()
    w.reset_counters()
    t0 = time.perf_counter()
    w.step(steps)
    elapsed = time.perf_counter() - t0
    c = w.counters()
    rinf, rrms = w.residual()
    w.close()
    n_dyn = int((scene["status"] == 0).sum())
    out = {
        "metric": "steps/s of 4,096 independent worlds stepped as one batch",
        "value": steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched worlds: 4,096 scenes, lognormal sizes (median 64, clip 8-2048), "
                               "balls/cuboids, own ground each; 8 velocity + 3 position iterations, dt 1/60",
                   "bodies": int(scene["status"].shape[0]), "dynamic_bodies": n_dyn,
                   "settle_steps": settle_steps, "contacts": c["ncontacts"], "manifolds": c["nmanifolds"],
                   "colours": c["ncolours"], "sweep_variant": c["sweep_variant"]},
        "contact_iters_per_s": c["ncontacts"] * 8 * steps / elapsed,
        "world_steps_per_s": 4096 * steps / elapsed,
        "residual_inf": rinf,
        "residual_rms": rrms,
        "stages_ms_per_step": {k: 1e3 * c[k] / steps for k in
                               ("collision_detection", "broad_phase", "narrow_phase", "assembly",
                                "velocity_resolution", "velocity_update", "position_resolution")},
    }
    if not no_cpu:
        log("timing the CPU baselines ...")
        out["cpu_baseline"], out["cpu_baseline_all_cores"] = cpu_baseline_batched(scene, snapshot, params,
                                                                                  cpu_steps)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["speedup_vs_cpu_all_cores"] = out["value"] / out["cpu_baseline_all_cores"]["value"]
    else:
        out["cpu_baseline"] = None
    return out


def bench_batched(args):
    """The batched worlds as the bench's one JSON line."""
    print(json.dumps(batched_section(args.steps, args.warmup, min(args.settle, 200), args.sweep_variant,
                                     args.cpu_steps, args.no_cpu)), flush=True)


def bench_ccd(args):
    """The 4,096-cuboid CCD room (the reference's examples3d/ccd3.rs scaled), args.steps steps from
    its initial scene, with the CPU restatement on the same steps unless --no-cpu."""
    out = ccd_config("the reference's CCD example with 16^3 = 4,096 cuboids in a room of half width 4 m; "
                     f"{args.steps} steps from the start, dt 1/60",
                     scenes.ccd_room(num=16, half_width=4.0, height=-1.76), abi.default_params(), args.steps,
                     0 if args.no_cpu else args.steps)
    out = dict({"metric": "steps/s of the 4,096-cuboid CCD room", "n_gpus": 1, "warmup": 0,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                "data": "synthetic"}, **out)
    print(json.dumps(out), flush=True)


def bench_sleeping(args):
    """The batched worlds with the reference's default sleep threshold (0.01) on every body,
    settled with sleeping on: islands at rest leave the solve and the integration (a single dense
    pile is one island and stays awake while any body in it moves). Prints one JSON line."""
    pile = args.sleeping_scene == "pile"
    params = abi.default_params(max_velocity_iterations=16 if pile else 8, max_position_iterations=3)
    scene = scenes.with_sleep_threshold(scenes.pile() if pile else scenes.batched(), 0.01)
    w = world.MechanicalWorld(integration_parameters=params)
    w.add_bodies(scene)
    w.set_sweep_variant(args.sweep_variant)
    t0 = time.perf_counter()
    if args.load_state:
        # a settled state saved by an earlier run: setting it wakes every body (2 x threshold),
        # so the settle steps that follow let the energies decay again
        st = np.load(args.load_state)
        w.set_state(**{k: st[k] for k in ("pos", "rot", "linvel", "angvel")})
        log(f"loaded settled state from {args.load_state}")
    settle(w, args.settle)
    log(f"settled in {time.perf_counter() - t0:.1f} s")
    w.step(args.warmup)
    w.reset_counters()
    t0 = time.perf_counter()
    w.step(args.steps)
    elapsed = time.perf_counter() - t0
    c = w.counters()
    n_dyn = int((scene["status"] == 0).sum())
    out = {
        "metric": ("steps/s of the dense pile" if pile else "steps/s of 4,096 independent worlds")
                  + " with sleeping (threshold 0.01 on every body)",
        "value": args.steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("dense pile as the default workload (16 velocity + 3 position iterations)" if pile
                                else "batched worlds as --workload batched (8 velocity + 3 position iterations)")
                               + ", every body with sleep threshold 0.01, settled with sleeping on",
                   "bodies": int(scene["status"].shape[0]), "dynamic_bodies": n_dyn,
                   "settle_steps": args.settle, "awake_bodies": c["active_bodies"],
                   "pairs": c["npairs"], "solved_contacts": c["ncontacts"], "solved_manifolds": c["nmanifolds"],
                   "colours": c["ncolours"]},
        "stages_ms_per_step": {k: 1e3 * c[k] / args.steps for k in
                               ("collision_detection", "broad_phase", "narrow_phase", "assembly",
                                "velocity_resolution", "velocity_update", "position_resolution")},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--settle", type=int, default=750)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the residual-target / learned runs")
    ap.add_argument("--target", type=float, default=1e-4, help="R-inf target (m/s) of the residual-target runs")
    ap.add_argument("--subset-max-sweeps", type=int, default=1000)
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the ball drop, box stack, batched worlds, pile subset and FP64 error sections")
    ap.add_argument("--target-steps", type=int, default=4)
    ap.add_argument("--target-max-sweeps", type=int, default=1000)
    ap.add_argument("--save-state", default=None, help="write the settled pile's state (.npz)")
    ap.add_argument("--load-state", default=None, help="start from a state written by --save-state")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device copy 5 times (the reference kernel of the "
                         "FETCH_SIZE / WRITE_SIZE counter calibration) and exit")
    ap.add_argument("--sweep-variant", type=int, default=-1,
                    help="-1 autotune (default), 0 per-colour launches, 256/512/1024 scene-blocked")
    ap.add_argument("--workload", choices=("pile", "batched", "sleeping", "ccd"), default="pile",
                    help="pile: the dense pile (default); batched: 4,096 independent worlds in one handle; "
                         "sleeping: the batched worlds with sleep threshold 0.01; ccd: the reference's CCD "
                         "example scaled to 4,096 cuboids (--steps steps from the start)")
    ap.add_argument("--sleeping-scene", choices=("batched", "pile"), default="batched",
                    help="scene of --workload sleeping")
    args = ap.parse_args()

    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or ws != 1:
        print("bench.py runs on one GPU only: this build steps one world per GPU and has no "
              "multi-GPU path; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch

        a = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 512 << 20, "launches": 5}), flush=True)
        return
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "sleeping":
        return bench_sleeping(args)
    if args.workload == "ccd":
        return bench_ccd(args)
    params = abi.default_params(max_velocity_iterations=16, max_position_iterations=3)
    scene = scenes.pile()
    w = world.MechanicalWorld(integration_parameters=params)
    w.add_bodies(scene)
    w.set_sweep_variant(args.sweep_variant)
    if args.load_state:
        # a settled pile saved by an earlier run (--save-state); the impulse cache starts empty
        st = np.load(args.load_state)
        w.set_state(**{k: st[k] for k in ("pos", "rot", "linvel", "angvel")})
        log(f"loaded settled state from {args.load_state}")
    else:
        t0 = time.perf_counter()
        settle(w, args.settle)
        log(f"settled in {time.perf_counter() - t0:.1f} s")
    snapshot = w.get_state()
    if args.save_state:
        np.savez(args.save_state, **snapshot)
    w.step(args.warmup)

    # Timed region: K steps, counters off.
    w.reset_counters()
    w.enable_counters(False)
    t0 = time.perf_counter()
    w.step(args.steps)  # returns after the last step completed on the device
    elapsed = time.perf_counter() - t0
    c = w.counters()
    rinf, rrms = w.residual()

    # Kernel-timing region: the same workload continued for --timing-steps steps with counters
    # enabled, which launches the sweep kernel directly with HIP events carried by each dispatch.
    w.reset_counters()
    w.enable_counters(True)
    w.step(args.timing_steps)
    w.enable_counters(False)
    ct = w.counters()

    n_dyn = int((scene["status"] == 0).sum())
    C, M, ncol = c["ncontacts"], c["nmanifolds"], c["ncolours"]
    iters = params.max_velocity_iterations
    # Algorithmic bytes of one velocity sweep, 168 C + 76 N, attributed to the launches in
    # proportion to the contacts each processes; the timed launches are those of the large colour
    # classes (one launch per class), so bytes per launch = bytes per contact x their contacts.
    sweep_bytes = 168 * C + 76 * n_dyn
    bytes_per_contact = sweep_bytes / max(1, C)
    launches = max(1, ct["sweep_kernel_launches"])
    avg_launch_s = ct["sweep_kernel_time"] / launches
    bytes_per_launch = bytes_per_contact * ct["sweep_kernel_contacts"] / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            traffic = json.load(open(PROFILE_TRAFFIC)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    steps_per_s = args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "dense pile: 262,144 bodies (balls r 0.2-0.6 m / cuboids 0.2-0.5 m) settled in a "
                        "60x40x60 m open box; 16 velocity + 3 position iterations, dt 1/60",
            "bodies": int(scene["status"].shape[0]),
            "dynamic_bodies": n_dyn,
            "settle_steps": args.settle,
            "pairs": c["npairs"],
            "contacts": C,
            "manifolds": M,
            "colours": ncol,
            "velocity_iterations": iters,
            "position_iterations": params.max_position_iterations,
            "loose_changes_per_step": c["loose_changes"] / args.steps,
            "pair_rebuilds_per_step": c["pair_rebuilds"] / args.steps,
            "incremental_pair_updates_per_step": c["incremental_pair_updates"] / args.steps,
        },
        "contact_iters_per_s": C * iters * args.steps / elapsed,
        "residual_inf": rinf,
        "residual_rms": rrms,
        "stages_ms_per_step": {k: 1e3 * c[k] / args.steps for k in
                               ("collision_detection", "broad_phase", "narrow_phase", "assembly",
                                "velocity_resolution", "velocity_update", "position_resolution")},
        "stage_rooflines": stage_rooflines(c, args.steps, C, M, n_dyn, c["npairs"], iters,
                                           params.max_position_iterations),
        "roofline": {
            "kernel": "k_pgs (one colour class of one SOR-Prox sweep, classes of more than 512 manifolds; "
                      "average over every k_pgs launch of the timing region)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "avg_launch_us": avg_launch_s * 1e6,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "launches": ct["sweep_kernel_launches"],
            "timing_region_steps": args.timing_steps,
        },
    }
    w.close()
    if not args.no_learned:
        log("residual-target runs, plain vs learned warm start ...")
        out["residual_target"] = residual_target_section(scene, snapshot, params, args.target,
                                                         args.target_steps, args.target_max_sweeps)
        out["learned_sweep_reduction"] = {
            wl: {f"{r['target']:.3g}": r["sweep_reduction"] for r in sec["runs"]}
            for wl, sec in out["residual_target"].items()}
    if not args.no_cpu:
        log("timing the CPU baseline ...")
        out["cpu_baseline"], cpu_one = cpu_baseline(scene, snapshot, params, args.cpu_steps)
        out["speedup_vs_cpu"] = steps_per_s / out["cpu_baseline"]["value"]
        g_inf, g_rms, gpu_one = gpu_from_snapshot(scene, snapshot, params, args.cpu_steps)
        cb = out["cpu_baseline"]
        out["residual_at_equal_iterations"] = {
            "what": f"mean over the CPU baseline's {args.cpu_steps} steps from the settled snapshot (one step "
                    f"after it first), {iters} sweeps per step, both sides from the same state",
            "gpu_inf": g_inf, "gpu_rms": g_rms,
            "cpu_reference_order_inf": cb["residual_inf"], "cpu_reference_order_rms": cb["residual_rms"],
            "ratio_inf": g_inf / cb["residual_inf"], "ratio_rms": g_rms / cb["residual_rms"]}
        if not args.no_configs:
            log("FP64 error of one pile step ...")
            out["f64_error"] = f64_error(scene, snapshot, params, gpu_one, cpu_one)
    else:
        out["cpu_baseline"] = None
    if not args.no_configs:
        log("pile corner subset: sweeps to the R-inf target in FP32 and FP64 ...")
        sub, sub_st = pile_subset(scene, snapshot)
        r = reachability(sub, sub_st, params, args.target, args.subset_max_sweeps)
        r["what"] = (f"the settled pile's bodies in the corner x, z > 20 m ({int((sub['status'] == 0).sum())} "
                     f"dynamic bodies) with the walls; one step solved to the target after 3 fixed-count steps")
        out["residual_target_pile_subset"] = r
        out["configs"] = configs_section(args)
    print(json.dumps(out), flush=True)


def ccd_config(what, scene, params, steps, cpu_steps):
    """A CCD workload run from its initial scene on the GPU (all `steps` steps timed, counters off;
    a second run with counters gives the CCD phase's share and the impacts), beside the one-thread
    FP32 CPU restatement (reference row order) timed on the first `cpu_steps` steps of the same
    trajectory."""
    w = gpu_world(scene, params)
    w.enable_counters(False)
    t0 = time.perf_counter()
    w.step(steps)
    elapsed = time.perf_counter() - t0
    w.close()
    w = gpu_world(scene, params)
    w.enable_counters(True)
    w.step(steps)
    c = w.counters()
    w.close()
    if cpu_steps > 0:
        o = oracle_world(scene, params, 32)
        t1 = time.perf_counter()
        o.step(cpu_steps)
        cpu_t = time.perf_counter() - t1
        oc = o.counters()
    out = {
        "workload": what,
        "bodies": int(scene["status"].shape[0]),
        "dynamic_bodies": int((scene["status"] == 0).sum()),
        "steps": steps,
        "value": steps / elapsed,
        "unit": "steps/s",
        "ms_per_step": 1e3 * elapsed / steps,
        "ccd_ms_per_step": 1e3 * c["ccd_time"] / max(c["nsteps"], 1),
        "ccd_substeps": c["ccd_substeps"],
        "ccd_impacts": c["ccd_impacts"],
        "impacts_per_s": c["ccd_impacts"] / elapsed,
        "cpu_baseline": None,
    }
    if cpu_steps > 0:
        out["cpu_baseline"] = {
            "value": cpu_steps / cpu_t, "unit": "steps/s", "cores": 1, "kind": "port", "impacts": oc["ccd_impacts"],
            "sample": f"the first {cpu_steps} steps of the same run ({oc['ccd_impacts']} impacts), FP32, reference "
                      f"row order, single thread, on {cpu_model()} ({os.cpu_count()} logical CPUs)"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def stage_rooflines(c, steps, C, M, N, P, sweeps, pos_iters):
    """Each stage's algorithmic bytes per step (the per-kernel figures of the design notes) over
    its measured time per step (stage timers: HIP events on the world's stream), against the
    8 TB/s HBM peak. The velocity stage counts its sweeps only (the warm start, residual and
    impulse cache add to its time); the broad phase counts a full pass (the incremental update
    moves less)."""
    rows = {
        "velocity_resolution": (sweeps * (168 * C + 76 * N), "k_warmstart + k_pgs sweeps: (168 C + 76 N) per sweep"),
        "position_resolution": (pos_iters * (48 * C + 84 * N), "k_position: (48 C + 84 N) per iteration"),
        "assembly": (224 * C + 104 * N, "colouring + k_assemble: 224 C + 104 N"),
        "narrow_phase": (100 * P + 48 * C, "narrow phase: 100 P + 48 C"),
        "broad_phase": (124 * N + 8 * P, "broad phase: 124 N + 8 P (a full pass)"),
        "velocity_update": (152 * N, "k_integrate (with k_prestep): 152 N"),
    }
    out = {}
    for k, (b, what) in rows.items():
        t = c[k] / steps
        gbps = b / t / 1e9 if t > 0 else 0.0
        out[k] = {"bytes_per_step": int(b), "ms_per_step": 1e3 * t, "achieved": gbps, "unit": "GB/s",
                  "frac": gbps / HBM_PEAK_GBPS, "bytes": what}
    return out


def configs_section(args):
    """The other workloads, each run whole on the GPU beside the CPU restatement: the ball drop
    (1,000 balls, 8 + 3 iterations, 200 steps), the box stack (8,000 cuboids, 10 + 3 iterations,
    500 steps), the reference's CCD example and a 4,096-body version of it, and the batched worlds
    (4,096 scenes, 8 + 3 iterations)."""
    out = {}
    log("ball drop ...")
    out["ball_drop"] = small_config(
        "ball drop: 1,000 balls r 0.5 m on a 10^3 lattice (1.1 m, jitter 0.05 m) over a static ground; "
        "8 velocity + 3 position iterations, 200 steps, dt 1/60",
        scenes.balls(), abi.default_params(max_velocity_iterations=8, max_position_iterations=3),
        200, [0, 60, 120, 180], 5, args.target, 2000)
    log("box stack ...")
    out["box_stack"] = small_config(
        "box stack: 8,000 cuboids (half extents 0.5 m) on a 20^3 lattice (1.05 m, tilt 2 degrees) over a "
        "static ground; 10 velocity + 3 position iterations, 500 steps, dt 1/60",
        scenes.boxes(), abi.default_params(max_velocity_iterations=10, max_position_iterations=3),
        500, [0, 125, 250, 375, 490], 3, args.target, 1000)
    log("CCD example ...")
    out["ccd_example"] = ccd_config(
        "the reference's CCD example (examples3d/ccd3.rs): 64 CCD-enabled cuboids (half extent 0.1 m) fired at "
        "(-100, -10, 0) m/s in a closed room of six walls (one static body, half width 3 m); 8 velocity + 3 "
        "position iterations, 10 CCD position iterations, 200 steps, dt 1/60",
        scenes.ccd_room(), abi.default_params(), 200, 200)
    log("CCD room of 4,096 cuboids ...")
    out["ccd_room_4096"] = ccd_config(
        "the same example with 16^3 = 4,096 cuboids in a room of half width 4 m: every body freezes in each "
        "step's impact cascade; 12 steps, dt 1/60",
        scenes.ccd_room(num=16, half_width=4.0, height=-1.76), abi.default_params(), 12, 12)
    log("batched worlds ...")
    out["batched"] = batched_section(args.batched_steps, 10, 200, args.sweep_variant, args.cpu_steps,
                                     args.no_cpu)
    return out


if __name__ == "__main__":
    main()
