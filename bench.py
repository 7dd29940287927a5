"""Benchmark of the batched kinematics hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): Panda with its
59-sphere model, 65536 float32 configurations uniform in the joint limits, a
world of 256 spheres (centres uniform in the 2 m cube around the base, radii
U[0.02, 0.15] m).  One step = forward kinematics to all 13 link poses + the
minimum signed distance of each configuration to the world + the 66 self-pair
distances; two kernel launches (pk_fk_links_f32, pk_sphere_sweep_f32) on
inputs already resident in HBM.

Printed as one JSON line: whole-job configs/s, the roofline of the dominant
kernel (the sphere sweep, FP32 VALU-bound; timed with HIP events on the stream
it runs on), the FP64 CPU oracle timed on this host's cores for a bounded
sample of the same workload, and one object per further workload:
  ik        LM-IK, Panda, 16 targets x 64 seeds, 30 iterations, lambda = 1e-2
  grad      pose + hinge cost and its adjoint, 16384 Panda configurations
  ik_uneven 4096 targets over Panda/UR5/Fetch/Baxter, 16..256 seeds each, 50
            iterations, 256 world spheres: fused graph-captured solves/s vs
            one plain launch per target, at equal success
  trajopt   512 Panda trajectories (T in {32, 64, 128}), 1024 world spheres,
            100 L-BFGS iterations in one captured graph
  hjcd      HJCD-IK (Robot.inverse_kinematics' solver) on 1024 Panda targets
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LM-IK solves/s @≤1mm/0.01rad fused-batched vs plain; FK+coll configs/s"
B = 65536
M_WORLD = 256
PEAK_FP32_TFLOPS = 157.3
# The same peak without packed FP32 (one FMA per lane per cycle): this build
# compiles with packed FP32 instructions disabled.
PEAK_FP32_UNPACKED_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# Algorithmic work per configuration (flop convention: sphere-sphere distance
# including the min = 12; SE(3) sphere placement ~ 33 per sphere).
K_SPHERES = 59
SELF_SPHERE_PAIRS = 1230
# LM-IK: algorithmic flops per LM iteration per seed for Panda without
# collision (FK + Jacobian 1.6 k, normal equations 0.7 k, Cholesky 0.3 k, five
# line-search FK + residual 7 k); with collision, 33 per robot-sphere placement
# and 13 per executed sphere-pair test (distance, hinge and accumulate).
IK_FLOP_PER_SEED_ITER = 9600
IK_FLOP_PLACEMENT = 33
IK_FLOP_PAIR = 13
FLOP_SWEEP = K_SPHERES * M_WORLD * 12 + SELF_SPHERE_PAIRS * 12 + K_SPHERES * 33
BYTES_FK = 7 * 4 + 13 * 7 * 4  # cfg in + link poses out
BYTES_SWEEP = 13 * 7 * 4 + 4 + 66 * 4  # link poses in + min + self pairs out


def make_inputs(seed: int = 0):
    import pyroffi_amd as pk
    from pyroffi_amd.collision import RobotCollisionSpherized

    robot = pk.Robot.from_urdf(pk.bundled_urdf_path("panda"))
    coll = RobotCollisionSpherized.from_urdf(pk.bundled_urdf_path("panda"))
    rng = np.random.default_rng(seed)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    q = rng.uniform(lo, hi, (B, 7)).astype(np.float32)
    world = np.concatenate([rng.uniform(-1, 1, (M_WORLD, 3)), rng.uniform(0.02, 0.15, (M_WORLD, 1))], 1)
    return robot, coll, q, world.astype(np.float32)


def gpu_bench(steps: int, warmup: int):
    import torch

    from pyroffi_amd.kernels.ops import fk_links, sphere_sweep

    robot, coll, q_np, world_np = make_inputs()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    q = torch.as_tensor(q_np, device=dev)
    world = torch.as_tensor(world_np, device=dev)
    t = robot.device_tables(dev)
    ct = coll.device_tables(dev)
    P = coll.num_pairs

    def step():
        poses = fk_links(q, *t.robot_buffers(), t.link_parent_joint)
        _, wmin, selfd = sphere_sweep(poses, ct["spheres_local"], ct["link_sphere_start"], world,
                                      ct["self_entries"], P, world_min=True, self_pairs=True)
        return poses, wmin, selfd

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / steps

    # Per-kernel averages with HIP events on the launch stream.
    poses = out[0]
    stream = torch.cuda.current_stream()

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        stream.synchronize()
        # A spin kernel holds the stream while the host queues every launch, so
        # the events time back-to-back kernels rather than the host's launch rate
        # (a 15 us kernel is close to what the Python launch path sustains).
        torch.cuda._sleep(20_000_000)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps  # ms

    reps = max(20, steps)
    fk_ms = timed(lambda: fk_links(q, *t.robot_buffers(), t.link_parent_joint), reps)
    sweep_ms = timed(lambda: sphere_sweep(poses, ct["spheres_local"], ct["link_sphere_start"], world,
                                          ct["self_entries"], P, world_min=True, self_pairs=True), reps)
    return dict(ms_per_step=ms_per_step, fk_ms=fk_ms, sweep_ms=sweep_ms, results=out)


def ik_workload():
    """Config 1 IK: Panda, 16 FK-generated targets x 64 uniform seeds, 30 iterations, lambda 1e-2.
    Returns ``(robot, link, targets (16, 7), seeds, kw, solve)``."""
    import torch

    import pyroffi_amd as pk
    from pyroffi_amd.kernels.ops import ls_ik_select
    from pyroffi_amd.optimization_engines._ls_ik import ls_ik_run

    robot = pk.Robot.from_urdf(pk.bundled_urdf_path("panda"))
    link = robot.links.names.index("panda_hand")
    rng = np.random.default_rng(1)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    tgt_q = rng.uniform(lo, hi, (16, 7))
    tgt = robot.forward_kinematics(tgt_q.astype(np.float32))[:, link].astype(np.float64)
    seeds = torch.as_tensor(rng.uniform(lo, hi, (16, 64, 7)).astype(np.float32), device="cuda")
    tT = torch.as_tensor(tgt[:, None].astype(np.float32), device="cuda")
    prev = torch.zeros((16, 7), device="cuda")
    kw = dict(max_iter=30, lambda_init=1e-2)

    def solve():
        cfgs, errs = ls_ik_run(robot, seeds, tT, (link,), **kw)
        return ls_ik_select(cfgs, errs, prev, 0.0)[0], cfgs, errs

    return robot, link, tgt, seeds, kw, solve


def ik_bench(reps: int):
    """Config 1 IK (``ik_workload``): GPU time per call, success, FP64/FP32 oracle on the same seeds."""
    import torch

    import oracle
    from pyroffi_amd.optimization_engines._ls_ik import ee_tables

    robot, link, tgt, seeds, kw, solve = ik_workload()
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    solve()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(reps):
        best, cfgs, errs = solve()
    e1.record(stream)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    gpu_ms = e0.elapsed_time(e1) / reps

    def success(qs):
        lp = robot.forward_kinematics(np.asarray(qs, np.float32).reshape(-1, 7))[:, link].astype(np.float64)
        tt = np.repeat(tgt, qs.size // (16 * 7), axis=0) if qs.ndim == 3 else tgt
        from pyroffi_amd._se3 import quat_conj, quat_mul, so3_log

        pos = np.linalg.norm(lp[:, 4:] - tt[:, 4:], axis=1)
        rot = np.linalg.norm(so3_log(quat_mul(quat_conj(tt[:, :4]), lp[:, :4])), axis=1)
        return (pos <= 1e-3) & (rot <= 0.01), (pos <= 1e-3) & (rot <= 0.05)

    win01, win05 = success(best.cpu().numpy())
    seed01, _ = success(cfgs.cpu().numpy())
    # FP64 oracle on the same seeds (also the CPU baseline of this line).
    t_or = oracle.Tables.of(robot)
    tj, masks = ee_tables(robot, (link,))
    threads = cpu_threads()
    c0 = time.perf_counter()
    ocfgs, oerrs = oracle.ls_ik(t_or, seeds.cpu().numpy(), tj, masks, tgt[:, None], lo, hi, threads=threads, **kw)
    cpu_dt = time.perf_counter() - c0
    owin01, _ = success(ocfgs[np.arange(16), oerrs.argmin(1)])
    oseed01, _ = success(ocfgs)
    o32 = oracle.single_precision()
    c0 = time.perf_counter()
    o32.ls_ik(t_or, seeds.cpu().numpy(), tj, masks, tgt[:, None].astype(np.float32), lo.astype(np.float32),
              hi.astype(np.float32), threads=threads, **kw)
    cpu32_dt = time.perf_counter() - c0
    flop = IK_FLOP_PER_SEED_ITER * 16 * 64 * 30
    achieved = flop / (gpu_ms * 1e-3) / 1e12
    return {
        "workload": "panda_lm_ik_16targets_64seeds_30iters",
        "solves_per_s": 16 / dt,
        "ms_per_solve_call": dt * 1e3,
        "gpu_event_ms_per_call": gpu_ms,
        "roofline": {"bound": "valu", "kernel": "ls_ik_wave_kernel (pk_ls_ik_f32) + select", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "flop_per_seed_iteration": IK_FLOP_PER_SEED_ITER,
                     "note": "latency-bound: 1024 seeds, one wave each (line search on 8 lanes, the step spread over the wave), one wave per SIMD"},
        "success_1mm_0.01rad": float(win01.mean()),
        "success_1mm_0.05rad": float(win05.mean()),
        "seed_success_1mm_0.01rad": float(seed01.mean()),
        "oracle_success_1mm_0.01rad": float(owin01.mean()),
        "oracle_seed_success_1mm_0.01rad": float(oseed01.mean()),
        "cpu_baseline": {"value": 16 / cpu_dt, "unit": "solves/s", "cores": threads, "kind": "port",
                         "sample": "the same 16 problems x 64 seeds, FP64 oracle",
                         "fp32_solves_per_s": 16 / cpu32_dt, "cpu_model": cpu_model()},
    }


UNEVEN_EE = {"panda": ("panda_hand",), "ur5": ("ee_link",), "fetch": ("gripper_link",),
             "baxter": ("left_gripper_base", "right_gripper_base")}


def ik_uneven_bench(reps: int, per_robot: int = 1024, plain_per_robot: int = 32, cpu_per_robot: int = 2,
                    batch_per_robot: int = 256):
    """Config 4: 4096 targets over Panda/UR5/Fetch/Baxter (two EEs), seeds per target
    log-uniform in [16, 256], 50 LM iterations, collision penalty against 256 world
    spheres.  Fused = one autotuned persistent launch per robot over flat seeds,
    grid-culled penalty, all launches captured in one graph.  Plain = one
    single-problem launch per target against every world sphere (timed on the
    first ``plain_per_robot`` targets of each robot; per-call cost is additive).
    """
    import torch

    from pyroffi_amd.optimization_engines._ls_ik_batched import UnevenLsIkSolver

    groups, samples = uneven_workload(per_robot, plain_per_robot)
    n_seeds = int(sum(g.seeds.shape[0] for g in groups))
    solver = UnevenLsIkSolver(groups, max_iter=50)
    report = solver.autotune(reps=1, cache=False)
    launch_mode = solver.capture(keep_faster=True)
    solver.solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = solver.solve()
    torch.cuda.synchronize()
    fused_dt = (time.perf_counter() - t0) / reps
    winners = [w.cpu().numpy() for w, _ in out]
    work = solver.count_work()
    torch.cuda.synchronize()
    # The reference's batch call shape: one launch per robot, seeds padded to the largest count.
    solver.solve_batch_padded(n_problems=4)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    padded = [w.cpu().numpy() for w in solver.solve_batch_padded(n_problems=batch_per_robot)]
    padded_dt = time.perf_counter() - t0
    res = _uneven_report(groups, samples, winners, fused_dt, n_seeds, report, per_robot, plain_per_robot,
                         cpu_per_robot)
    res["launch_mode"] = launch_mode  # "graph": replay of the captured graph; "eager": direct launches (faster here)
    res.update(_uneven_roofline(groups, work, fused_dt, solver.kw["max_iter"], list(UNEVEN_EE)))
    res.update(_padded_report(groups, winners, padded, padded_dt, batch_per_robot, per_robot,
                              res["fused_solves_per_s"]))
    return res


def _uneven_roofline(groups, work, fused_dt, max_iter, names):
    """Work-based FLOP/s of the fused solve: per seed max_iter LM iterations at the
    Panda count scaled by DOF and end effectors, 1 + 6 max_iter penalty evaluations
    of robot-sphere placements, and the sphere-pair tests the kernel counted."""
    flop = 0.0
    per_robot = {}
    for name, g, w in zip(names, groups, work):
        S = g.seeds.shape[0]
        A = g.robot.joints.num_actuated_joints
        K = 0 if g.collision is None else len(g.collision[0])
        base = S * max_iter * IK_FLOP_PER_SEED_ITER * (A / 7.0) * len(g.target_link_indices)
        place = S * (1 + 6 * max_iter) * K * IK_FLOP_PLACEMENT
        pairs = w * IK_FLOP_PAIR
        flop += base + place + pairs
        per_robot[name] = {"seeds": S, "pair_tests": w, "pair_tests_per_seed_eval":
                                             w / max(1, S * (1 + 6 * max_iter) * max(K, 1))}
    achieved = flop / fused_dt / 1e12
    return {"roofline": {"bound": "valu", "kernel": "ls_ik_batched_kernel (pk_ls_ik_batched_f32), all robots",
                         "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / PEAK_FP32_TFLOPS, "flop_per_batch": flop,
                         "work": per_robot}}


def _padded_report(groups, winners, padded, padded_dt, batch_per_robot, per_robot, fused_rate):
    from pyroffi_amd._se3 import quat_conj, quat_mul, so3_log

    def ok(g, q):
        lp = g.robot.forward_kinematics(np.ascontiguousarray(q, np.float32)).astype(np.float64)
        r = np.ones(len(q), bool)
        for e, l in enumerate(g.target_link_indices):
            tt = g.targets[:len(q), e]
            pos = np.linalg.norm(lp[:, l, 4:] - tt[:, 4:], axis=1)
            rot = np.linalg.norm(so3_log(quat_mul(quat_conj(tt[:, :4]), lp[:, l, :4])), axis=1)
            r &= (pos <= 1e-3) & (rot <= 0.01)
        return r

    n = len(groups) * batch_per_robot
    rate = n / padded_dt
    return {"reference_batch_shape": {
        "solves_per_s": rate, "sample_targets": n, "seconds": padded_dt,
        "speedup_fused_vs_batch_shape": fused_rate / rate,
        "success_1mm_0.01rad": float(np.concatenate([ok(g, p) for g, p in zip(groups, padded)]).mean()),
        "fused_success_same_targets": float(np.concatenate([ok(g, w[:batch_per_robot])
                                                            for g, w in zip(groups, winners)]).mean()),
        "shape": "one pk_ls_ik_f32 launch per robot over its first targets, seeds padded to the largest "
                 "count, brute-force collision penalty (ls_ik_solve_cuda_batch's single seed count)"}}


def uneven_workload(per_robot: int = 1024, plain_per_robot: int = 32):
    """Config 4 problems: per robot, FK-generated targets, log-uniform seed counts, 256 world spheres."""
    import pyroffi_amd as pk
    from pyroffi_amd.collision import RobotCollisionSpherized
    from pyroffi_amd.optimization_engines._ls_ik_batched import RobotGroup, log_uniform_seed_counts, make_group

    rng = np.random.default_rng(4)
    world = np.concatenate([rng.uniform(-1, 1, (M_WORLD, 3)), rng.uniform(0.02, 0.15, (M_WORLD, 1))], 1)
    world = world.astype(np.float32)
    groups, samples = [], []
    for name, links in UNEVEN_EE.items():
        path = pk.bundled_urdf_path(name)
        robot = pk.Robot.from_urdf(path)
        li = [robot.links.names.index(n) for n in links]
        lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
        tq = rng.uniform(lo, hi, (per_robot, len(lo))).astype(np.float32)
        lp = robot.forward_kinematics(tq)
        tgt = np.stack([lp[:, l] for l in li], axis=1).astype(np.float64)
        counts = log_uniform_seed_counts(rng, per_robot, 16, 256)
        g = make_group(robot, li, tgt, counts, rng, collision_checker=RobotCollisionSpherized.from_urdf(path),
                       world=world)
        groups.append(g)
        n = plain_per_robot
        samples.append(RobotGroup(robot=robot, target_link_indices=g.target_link_indices, targets=g.targets[:n],
                                  seeds=g.seeds[:g.seed_start[n]], seed_start=g.seed_start[:n + 1],
                                  collision=g.collision))
    return groups, samples


def _uneven_report(groups, samples, winners, fused_dt, n_seeds, report, per_robot, plain_per_robot, cpu_per_robot):
    import torch

    import oracle
    from pyroffi_amd._se3 import quat_conj, quat_mul, so3_log
    from pyroffi_amd.optimization_engines._ls_ik import ee_tables
    from pyroffi_amd.optimization_engines._ls_ik_batched import UnevenLsIkSolver

    plain = UnevenLsIkSolver(samples, max_iter=50)
    plain.solve_per_call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_w = [q.cpu().numpy() for q in plain.solve_per_call()]
    plain_dt = time.perf_counter() - t0
    n_plain = sum(len(s.seed_start) - 1 for s in samples)

    def success(g, q, rows=None):
        lp = g.robot.forward_kinematics(np.ascontiguousarray(q, np.float32)).astype(np.float64)
        ok = np.ones(len(q), bool)
        rows = np.arange(len(q)) if rows is None else np.asarray(rows)
        for e, l in enumerate(g.target_link_indices):
            tt = g.targets[rows, e]
            pos = np.linalg.norm(lp[:, l, 4:] - tt[:, 4:], axis=1)
            rot = np.linalg.norm(so3_log(quat_mul(quat_conj(tt[:, :4]), lp[:, l, :4])), axis=1)
            ok &= (pos <= 1e-3) & (rot <= 0.01)
        return ok

    fused_ok = [success(g, w) for g, w in zip(groups, winners)]
    plain_ok = [success(s, w) for s, w in zip(samples, plain_w)]
    fused_on_sample = [ok[:plain_per_robot] for ok in fused_ok]

    # FP64 oracle on a bounded sample (the first targets of each robot, their own seeds).
    threads = cpu_threads()
    c0 = time.perf_counter()
    n_cpu = 0
    oracle_ok, fused_ok_cpu = [], []
    for g, ok_g in zip(groups, fused_ok):
        t_or = oracle.Tables.of(g.robot)
        tj, masks = ee_tables(g.robot, g.target_link_indices)
        rs, rj, ws = g.collision
        for p in range(cpu_per_robot):
            s0, s1 = g.seed_start[p], g.seed_start[p + 1]
            oc, oe = oracle.ls_ik(t_or, g.seeds[None, s0:s1], tj, masks, g.targets[p:p + 1],
                                  g.robot.joints.lower_limits, g.robot.joints.upper_limits, max_iter=50,
                                  robot_spheres=rs, sphere_joint=rj, world=ws, enable_collision=True,
                                  collision_weight=1e4, collision_margin=0.02, threads=threads)
            n_cpu += 1
            oracle_ok.append(bool(success(g, oc[0, int(np.argmin(oe[0]))][None], rows=[p])[0]))
            fused_ok_cpu.append(bool(ok_g[p]))
    cpu_dt = time.perf_counter() - c0
    names = list(UNEVEN_EE)
    return {
        "workload": "uneven_lm_ik_4robots_4096targets_seeds16to256_50iters_world256",
        "targets": per_robot * len(groups),
        "seeds_total": n_seeds,
        "fused_solves_per_s": per_robot * len(groups) / fused_dt,
        "fused_ms_per_batch": fused_dt * 1e3,
        "plain_solves_per_s": n_plain / plain_dt,
        "plain_sample_targets": n_plain,
        "speedup_fused_vs_plain": (per_robot * len(groups) / fused_dt) / (n_plain / plain_dt),
        "fused_success_1mm_0.01rad": {n: float(ok.mean()) for n, ok in zip(names, fused_ok)},
        "fused_success_on_plain_sample": float(np.concatenate(fused_on_sample).mean()),
        "plain_success_on_plain_sample": float(np.concatenate(plain_ok).mean()),
        "oracle_success_on_cpu_sample": float(np.mean(oracle_ok)),
        "fused_success_on_cpu_sample": float(np.mean(fused_ok_cpu)),
        "autotune": {n: list(g.tune) for n, g in zip(names, groups)},
        "autotune_ms": [v.get("ms") for v in report.values()],
        "cpu_baseline": {"value": n_cpu / cpu_dt, "unit": "solves/s", "cores": threads, "kind": "port",
                         "sample": f"first {cpu_per_robot} targets of each robot with their seeds, FP64 oracle",
                         "cpu_model": cpu_model()},
    }


def trajopt_bench(reps: int, cpu_sample: int = 16):
    """Config 5: 512 Panda trajectories with T mixed over {32, 64, 128}, random in-limit
    endpoints, straight-line initial paths, 1024 world spheres (centres in the 2 m cube,
    radii U[0.02, 0.1]); the direct L-BFGS of lbfgs_trajopt (default LbfgsTrajOptConfig:
    100 iterations, colldist activation with margin 0.01, collision-weight continuation)
    with the cost gradient from the hand-written adjoint; the 100 iterations and the
    final ranking run as one replay of a captured graph.
    """
    import torch

    import oracle

    robot, coll, world, trajs, lens, cfg, model, opt = trajopt_workload()
    opt.capture()
    opt.solve()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(reps):
        best, best_cost, rank = opt.solve()
    e1.record(stream)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    gpu_ms = e0.elapsed_time(e1) / reps
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    return _trajopt_report(robot, coll, world, trajs, lens, cfg, model, opt, best_cost, rank, dt, gpu_ms, lo, hi,
                           cpu_sample)


def chomp_bench(reps: int = 3, cpu_sample: int = 4):
    """CHOMP (chomp_trajopt's solver, default ChompTrajOptConfig: 100 iterations, covariant
    update, hinge^2 margin 0.01) on the 512-trajectory workload (T in {32, 64, 128}, 1024 world
    spheres): one launch of
    pk_traj_chomp_f32 for all 512 trajectories, then the ranking costs."""
    import torch

    import oracle
    from pyroffi_amd.optimization_engines._chomp_optimization import ChompTrajOptConfig, chomp_run
    from pyroffi_amd.optimization_engines._lbfgs_trajopt_optimization import TrajModel, flatten_trajs, traj_cost

    robot, coll, world, trajs, lens, _, _, _ = trajopt_workload()
    cfg = ChompTrajOptConfig()
    model = TrajModel(robot, coll, world, cfg.collision_margin, mode=0)
    x, start, max_len = flatten_trajs(trajs, model.device)
    chomp_run(model, x, start, max_len, cfg)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        out, cost, _ = chomp_run(model, x, start, max_len, cfg)
    e1.record(stream)
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1) / reps
    w = (cfg.w_smooth, 0.0, cfg.w_acc, cfg.w_jerk, cfg.w_limits, cfg.w_collision_max)
    c0, _ = traj_cost(model, x, start, max_len, w, grad=False)
    init = (c0 @ torch.tensor([cfg.w_smooth, cfg.w_limits, cfg.w_collision_max], device=c0.device)).cpu().numpy()
    final = cost.cpu().numpy()
    t = oracle.Tables.of(robot)
    om = oracle.TrajModel(t, oracle.padded_spheres(coll), coll.active_idx_i, coll.active_idx_j,
                          world.astype(np.float64), cfg.collision_margin, 0)
    kw = {k: getattr(cfg, k) for k in ("n_iters", "step_size", "w_smooth", "w_acc", "w_jerk", "w_collision",
                                        "w_collision_max", "collision_penalty_scale", "w_limits",
                                        "use_covariant_update", "smoothness_reg", "n_cg_iters", "cg_tol",
                                        "grad_clip_norm", "max_delta_per_step", "early_stop_patience",
                                        "min_cost_improve")}
    threads = cpu_threads()
    c_0 = time.perf_counter()
    _, ocost, _ = oracle.traj_chomp(om, [tr.astype(np.float64) for tr in trajs[:cpu_sample]],
                                    robot.joints.lower_limits, robot.joints.upper_limits, threads=threads, **kw)
    cpu_dt = time.perf_counter() - c_0
    return {
        "workload": "panda_chomp_512traj_T32_64_128_world1024_100iters",
        "trajectories": len(trajs),
        "ms_per_solve": gpu_ms,
        "trajectories_per_s": len(trajs) / (gpu_ms * 1e-3),
        "median_cost_ratio_final_over_initial": float(np.median(final / init)),
        "oracle_sample_median_cost_ratio_gpu_over_fp64": float(np.median(final[:cpu_sample] / ocost)),
        "cpu_baseline": {"value": cpu_sample / cpu_dt, "unit": "trajectories/s", "cores": threads, "kind": "port",
                         "sample": f"first {cpu_sample} trajectories, FP64 oracle, brute-force world",
                         "cpu_model": cpu_model()},
    }


def sco_bench(reps: int = 3, cpu_sample: int = 4):
    """SCO (sco_trajopt's solver, default ScoTrajOptConfig: 10 rounds of linearise + 30
    inner L-BFGS steps) on the 512-trajectory workload and its 1024-sphere world, one world group; the
    20 launches and the ranking cost replay from one captured graph."""
    import torch

    import oracle
    from pyroffi_amd.optimization_engines._sco_optimization import BatchedSco, ScoTrajOptConfig

    robot, coll, world, trajs, lens, _, _, _ = trajopt_workload()
    cfg = ScoTrajOptConfig()
    opt = BatchedSco(robot, coll, world, trajs, cfg)
    opt.capture()
    opt.solve()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        _, cost = opt.solve()
    e1.record(stream)
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1) / reps
    final = cost.cpu().numpy()
    t = oracle.Tables.of(robot)
    om = oracle.TrajModel(t, oracle.padded_spheres(coll), coll.active_idx_i, coll.active_idx_j,
                          world.astype(np.float64), cfg.collision_margin, 1)
    kw = {k: getattr(cfg, k) for k in ("n_outer_iters", "n_inner_iters", "m_lbfgs", "w_smooth", "w_vel", "w_acc",
                                        "w_jerk", "w_collision", "w_collision_max", "penalty_scale",
                                        "collision_margin", "w_trust", "w_limits", "smooth_min_temperature")}
    threads = cpu_threads()
    c_0 = time.perf_counter()
    _, ocost, _ = oracle.traj_sco(om, [tr.astype(np.float64) for tr in trajs[:cpu_sample]],
                                  robot.joints.lower_limits, robot.joints.upper_limits,
                                  np.array([0, len(world)], np.int32), threads=threads, **kw)
    cpu_dt = time.perf_counter() - c_0
    return {
        "workload": "panda_sco_512traj_T32_64_128_world1024_10x30iters",
        "trajectories": len(trajs),
        "ms_per_solve": gpu_ms,
        "trajectories_per_s": len(trajs) / (gpu_ms * 1e-3),
        "oracle_sample_median_cost_ratio_gpu_over_fp64": float(np.median(final[:cpu_sample] / ocost)),
        "cpu_baseline": {"value": cpu_sample / cpu_dt, "unit": "trajectories/s", "cores": threads, "kind": "port",
                         "sample": f"first {cpu_sample} trajectories, FP64 oracle",
                         "cpu_model": cpu_model()},
    }


def fk_small_bench(reps: int = 200):
    """Config 1 FK: 1,024 Panda configurations to all link poses (one launch, launch-latency bound)."""
    import torch

    import pyroffi_amd as pk
    from pyroffi_amd.kernels import ops

    robot = pk.Robot.from_urdf(pk.bundled_urdf_path("panda"))
    t = robot.device_tables()
    rng = np.random.default_rng(11)
    q = torch.as_tensor(rng.uniform(robot.joints.lower_limits, robot.joints.upper_limits, (1024, 7))
                        .astype(np.float32), device="cuda")
    bufs = t.robot_buffers()
    for _ in range(20):
        ops.fk_links(q, *bufs, t.link_parent_joint)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def per_call(queued):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        if queued:
            torch.cuda._sleep(20_000_000)  # the host queues every launch before the first runs
        e0.record(stream)
        for _ in range(reps):
            ops.fk_links(q, *bufs, t.link_parent_joint)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms = per_call(False)
    ms_kernel = per_call(True)
    return {"workload": "panda_fk_links_b1024", "configs": 1024, "ms_per_call": ms,
            "configs_per_s": 1024 / (ms * 1e-3), "ms_per_call_queued": ms_kernel,
            "note": "ms_per_call: back-to-back calls from Python (host launch path included); "
                    "ms_per_call_queued: the same launches queued behind a spin kernel (device time per call)"}


def trajopt_workload():
    """Config 5 inputs: 512 Panda straight-line trajectories (T in {32, 64, 128}), 1024 world spheres."""
    import pyroffi_amd as pk
    from pyroffi_amd.collision import RobotCollisionSpherized
    from pyroffi_amd.optimization_engines._lbfgs_trajopt_optimization import (BatchedLbfgsTrajOpt,
                                                                              LbfgsTrajOptConfig, TrajModel)

    robot = pk.Robot.from_urdf(pk.bundled_urdf_path("panda"))
    coll = RobotCollisionSpherized.from_urdf(pk.bundled_urdf_path("panda"))
    rng = np.random.default_rng(5)
    world = np.concatenate([rng.uniform(-1, 1, (1024, 3)), rng.uniform(0.02, 0.1, (1024, 1))], 1).astype(np.float32)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    lens = np.array([32, 64, 128])[rng.integers(0, 3, 512)]
    trajs = []
    for T in lens:
        q0, q1 = rng.uniform(lo, hi), rng.uniform(lo, hi)
        s = np.linspace(0, 1, T)[:, None]
        trajs.append(((1 - s) * q0 + s * q1).astype(np.float32))
    cfg = LbfgsTrajOptConfig()
    model = TrajModel(robot, coll, world, cfg.collision_margin, mode=1)
    opt = BatchedLbfgsTrajOpt(model, trajs, cfg)
    return robot, coll, world, trajs, lens, cfg, model, opt


def _trajopt_report(robot, coll, world, trajs, lens, cfg, model, opt, best_cost, rank, dt, gpu_ms, lo, hi,
                    cpu_sample):
    import oracle

    comps = opt.rank_comps.cpu().numpy()
    # Initial cost at the maximum collision weight, for the reduction ratio.
    from pyroffi_amd.optimization_engines._lbfgs_trajopt_optimization import traj_cost

    c0, _ = traj_cost(model, opt.x0, opt.start, opt.max_len,
                      (cfg.w_smooth, cfg.w_vel, cfg.w_acc, cfg.w_jerk, cfg.w_limits, cfg.w_collision_max), grad=False)
    c0 = c0.cpu().numpy()
    init_total = c0 @ [cfg.w_smooth, cfg.w_limits, cfg.w_collision_max]
    final_total = rank.cpu().numpy()
    # The reference tracks its "best" iterate against the initial cost at the
    # starting collision weight, so in a cluttered world it often keeps the
    # initial path; the last iterate shows what the optimiser did.
    cl, _ = traj_cost(model, opt.final_iterate(), opt.start, opt.max_len,
                      (cfg.w_smooth, cfg.w_vel, cfg.w_acc, cfg.w_jerk, cfg.w_limits, cfg.w_collision_max), grad=False)
    cl = cl.cpu().numpy()
    last_total = cl @ [cfg.w_smooth, cfg.w_limits, cfg.w_collision_max]
    waypoints = int(lens.sum())

    # FP64 oracle (brute force over the world) on the first trajectories, on the host cores.
    t = oracle.Tables.of(robot)
    om = oracle.TrajModel(t, oracle.padded_spheres(coll), coll.active_idx_i, coll.active_idx_j,
                          world.astype(np.float64), cfg.collision_margin, 1)
    threads = cpu_threads()
    c_start = time.perf_counter()
    _, ocost, _ = oracle.traj_lbfgs(om, [x.astype(np.float64) for x in trajs[:cpu_sample]], lo, hi, n_iters=cfg.n_iters,
                                    threads=threads)
    cpu_dt = time.perf_counter() - c_start
    gcost = best_cost.cpu().numpy()[:cpu_sample]
    return {
        "workload": "panda_lbfgs_trajopt_512traj_T32_64_128_world1024_100iters",
        "trajectories": len(trajs),
        "waypoints": waypoints,
        "iterations": cfg.n_iters,
        "trajectories_per_s": len(trajs) / dt,
        "ms_per_solve": dt * 1e3,
        "gpu_event_ms_per_solve": gpu_ms,
        "waypoint_iterations_per_s": waypoints * cfg.n_iters / dt,
        "collision_free_fraction": float(np.mean(comps[:, 2] == 0.0)),
        "initial_collision_free_fraction": float(np.mean(c0[:, 2] == 0.0)),
        "median_cost_ratio_best_over_initial": float(np.median(final_total / np.maximum(init_total, 1e-12))),
        "median_cost_ratio_last_iterate_over_initial": float(np.median(last_total / np.maximum(init_total, 1e-12))),
        "median_collision_ratio_last_iterate_over_initial": float(np.median(cl[:, 2] / np.maximum(c0[:, 2], 1e-12))),
        "last_iterate_collision_free_fraction": float(np.mean(cl[:, 2] == 0.0)),
        "oracle_sample_median_cost_ratio_gpu_over_fp64": float(np.median((gcost + 1e-3) / (ocost + 1e-3))),
        "cpu_baseline": {"value": cpu_sample / cpu_dt, "unit": "trajectories/s", "cores": threads, "kind": "port",
                         "sample": f"first {cpu_sample} trajectories (lengths {lens[:cpu_sample].tolist()}), "
                                   "FP64 oracle, brute-force world"},
    }


def hjcd_bench(reps: int, n_problems: int = 1024, cpu_sample: int = 8):
    """HJCD-IK (Robot.inverse_kinematics' solver) batched: 1024 FK-generated Panda targets,
    32 coarse seeds each (top 6 x 5 refinement seeds), 20 coarse + 40 LM iterations,
    continuity weight 1e-3 from the joint-range midpoint.  The FP64 oracle runs the same
    phases on the first targets with the GPU run's own seeds, noise and kick key."""
    import torch

    import oracle
    import pyroffi_amd as pk
    from pyroffi_amd.optimization_engines._hjcd_ik import hjcd_solve_batch
    from pyroffi_amd.optimization_engines._ls_ik import ee_tables

    robot = pk.Robot.from_urdf(pk.bundled_urdf_path("panda"))
    link = robot.links.names.index("panda_hand")
    rng = np.random.default_rng(7)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    tq = rng.uniform(lo, hi, (n_problems, 7)).astype(np.float32)
    tgt = robot.forward_kinematics(tq)[:, link].astype(np.float64)[:, None]
    prev = np.tile((lo + hi) / 2, (n_problems, 1))
    hjcd_solve_batch(robot, (link,), tgt, 0, prev, continuity_weight=1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        best, d = hjcd_solve_batch(robot, (link,), tgt, r, prev, continuity_weight=1e-3, return_details=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    best = best.cpu().numpy()
    t_or = oracle.Tables.of(robot)
    tj, _ = ee_tables(robot, (link,))

    def ok(q, rows):
        f = oracle.jres(t_or, q.astype(np.float64), tj, tgt[rows])
        rot = oracle.fk_links(t_or, q.astype(np.float64))[:, link]
        from scipy.spatial.transform import Rotation

        ang = np.linalg.norm((Rotation.from_quat(tgt[rows, 0][:, [1, 2, 3, 0]]).inv() *
                              Rotation.from_quat(rot[:, [1, 2, 3, 0]])).as_rotvec(), axis=1)
        return (np.linalg.norm(rot[:, 4:] - tgt[rows, 0, 4:], axis=1) <= 1e-3) & (ang <= 0.01) & np.isfinite(f[:, 0, 0])

    win = ok(best, np.arange(n_problems))
    # FP64 oracle: the same phases on the first targets, timed on the host cores.
    threads = cpu_threads()
    n = cpu_sample
    c0 = time.perf_counter()
    oc, oe = oracle.hjcd_coarse(t_or, d["seeds"][:n], tj, tgt[:n], d["ee_of_seed"], lo, hi, k_max=20,
                                threads=threads)
    top_k, repeats = d["top_k"], d["repeats"]
    R = top_k * repeats
    ref = np.empty((n, R, 7))
    for p in range(n):
        top = np.argsort(np.where(np.isnan(oe[p]), np.inf, oe[p]), kind="stable")[:top_k]
        base = np.tile(oc[p, top], (repeats, 1))
        ref[p] = np.where((np.arange(R) < top_k)[:, None], base,
                          np.clip(base + d["noise"][p] * 0.15 * (hi - lo), lo, hi))
    cfgs, errs = oracle.jlm(t_or, ref, tj, tgt[:n], lo, hi, max_iter=40, adaptive=True, prior_weight=1e-4,
                            kick_scale=0.05, kick_key=d["kick_key"], threads=threads)
    score = errs + 1e-3 * np.sum((cfgs - prev[:n, None]) ** 2, axis=-1)
    obest = cfgs[np.arange(n), np.argmin(score, axis=1)]
    cpu_dt = time.perf_counter() - c0
    return {
        "workload": "panda_hjcd_ik_1024targets_32seeds_20coarse_40lm",
        "solves_per_s": n_problems / dt,
        "ms_per_batch": dt * 1e3,
        "success_1mm_0.01rad": float(win.mean()),
        "success_on_oracle_sample": float(win[:n].mean()),
        "oracle_success_on_sample": float(ok(obest, np.arange(n)).mean()),
        "cpu_baseline": {"value": n / cpu_dt, "unit": "solves/s", "cores": threads, "kind": "port",
                         "sample": f"first {n} targets, FP64 oracle of the same phases", "cpu_model": cpu_model()},
    }


def grad_workload():
    """Config 3 inputs: Panda, 16384 configs, FK-generated targets, 512 world spheres.
    Returns (robot, coll, link, q_np, world_np, q, tgt, world, step, run): ``step`` is one
    cost + dcost/dq through the autograd Function, ``run`` one launch of the adjoint kernel
    on precomputed link poses."""
    import torch

    import pyroffi_amd as pk
    from pyroffi_amd.collision import RobotCollisionSpherized
    from pyroffi_amd.costs import pose_hinge_cost
    from pyroffi_amd.kernels.ops import fk_links, pose_hinge_cost as raw_cost

    nb, m = 16384, 512
    robot = pk.Robot.from_urdf(pk.bundled_urdf_path("panda"))
    coll = RobotCollisionSpherized.from_urdf(pk.bundled_urdf_path("panda"))
    link = robot.links.names.index("panda_hand")
    rng = np.random.default_rng(3)
    lo, hi = robot.joints.lower_limits, robot.joints.upper_limits
    q_np = rng.uniform(lo, hi, (nb, 7)).astype(np.float32)
    tq = rng.uniform(lo, hi, (nb, 7)).astype(np.float32)
    world_np = np.concatenate([rng.uniform(-1, 1, (m, 3)), rng.uniform(0.02, 0.15, (m, 1))], 1).astype(np.float32)
    q = torch.as_tensor(q_np, device="cuda")
    tgt = robot.forward_kinematics(torch.as_tensor(tq, device="cuda"))[:, link].contiguous()
    world = torch.as_tensor(world_np, device="cuda")

    def step():
        qq = q.detach().requires_grad_(True)
        c = pose_hinge_cost(robot, coll, qq, tgt, world, margin=0.05, ee_link=link)
        (g,) = torch.autograd.grad(c.sum(), qq)
        return c, g

    t = robot.device_tables()
    ct = coll.device_tables(q.device)
    poses = fk_links(q, *t.robot_buffers(), t.link_parent_joint)
    run = lambda: raw_cost(poses, ct["spheres_local"], ct["link_sphere_start"], world, 0.05, link, tgt, t.twists,
                           t.parent_idx, t.act_idx, t.mimic_act_idx, t.mimic_mul, t.topo_inv, t.link_parent_joint, 7)
    return robot, coll, link, q_np, world_np, q, tgt, world, step, run


def grad_bench(reps: int, cpu_sample: int):
    """Config 3: Panda, 16384 configs, pose + hinge (margin 0.05) against 512 world spheres,
    cost and dcost/dq through the hand-written adjoint (autograd Function)."""
    import torch

    import oracle

    robot, coll, link, q_np, world_np, q, tgt, world, step, run = grad_workload()
    nb, m = q_np.shape[0], world_np.shape[0]
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    # The adjoint kernel alone (HIP events on the launch stream).
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    run()
    e0.record(stream)
    for _ in range(reps):
        run()
    e1.record(stream)
    e1.synchronize()
    kms = e0.elapsed_time(e1) / reps
    flop = K_SPHERES * m * 13 * 3  # forward 13 flop per pair, backward ~2x forward
    # CPU: FP64 oracle cost + central-difference gradient (15 cost evaluations per config).
    t_or = oracle.Tables.of(robot)
    sp = oracle.padded_spheres(coll)
    tg = tgt.cpu().numpy().astype(np.float64)
    threads = cpu_threads()
    c0 = time.perf_counter()
    oracle.pose_hinge_cost(t_or, q_np[:cpu_sample], sp, world_np, 0.05, link, tg[:cpu_sample], threads=threads)
    oracle.pose_hinge_fd_grad(t_or, q_np[:cpu_sample], sp, world_np, 0.05, link, tg[:cpu_sample], threads=threads)
    cdt = time.perf_counter() - c0
    return {
        "workload": "panda_pose_hinge_cost_and_grad_b16384_world512",
        "configs_per_s": nb / dt,
        "ms_per_step": dt * 1e3,
        "adjoint_kernel_ms": kms,
        "adjoint_kernel_tflops": flop * nb / (kms * 1e-3) / 1e12,
        "flop_per_config": flop,
        "cpu_baseline": {"value": cpu_sample / cdt, "unit": "configs/s", "cores": threads, "kind": "port",
                         "sample": f"{cpu_sample} configs: FP64 cost + central-difference gradient (h=1e-4)"},
    }


def cpu_threads() -> int:
    """Every core this process may run on (its affinity set)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_sweep(o, t, sp, coll, q, world, threads):
    """FK + world min + self pairs with oracle module ``o`` (FP64 or its FP32 build)."""
    lp = o.fk_links(t, q, threads)
    _, mins = o.world_dist(lp, sp, world, want_full=False, threads=threads)
    selfd = o.self_dist(lp, sp, coll.active_idx_i, coll.active_idx_j, threads=threads)
    return lp, mins, selfd


def cpu_baseline(sample: int):
    """The oracle (C + OpenMP restatement) on `sample` configurations of the same workload:
    FP64 (the reference JAX path's arithmetic) on every core -- the reported value -- and
    on one core, and the FP32 build on every core and on one core, each the median of five
    timed runs after a warm-up.  One-core legs run on a sixteenth of the sample."""
    import oracle

    robot, coll, q, world = make_inputs()
    t = oracle.Tables.of(robot)
    threads = cpu_threads()
    legs = {}
    out = None
    for prec, o in (("fp64", oracle), ("fp32", oracle.single_precision())):
        R = o._R
        sp = o.padded_spheres(coll)
        w = world.astype(R)
        for cores in (threads, 1):
            n = sample if cores == threads else max(64, sample // 16)
            qq = q[:n].astype(R)
            _oracle_sweep(o, t, sp, coll, qq[:64], w, cores)  # warm-up (thread pool, page-in)
            times = []
            for _ in range(5):
                t0 = time.perf_counter()
                res = _oracle_sweep(o, t, sp, coll, qq, w, cores)
                times.append(time.perf_counter() - t0)
            legs[f"{prec}_{cores}_cores"] = n / float(np.median(times))
            if prec == "fp64" and cores == threads:
                out = res
    return {"value": legs[f"fp64_{threads}_cores"], "unit": "configs/s", "cores": threads, "kind": "port",
            "sample": f"first {sample} of the {B} configurations (one-core legs: the first {max(64, sample // 16)}), "
                      "oracle FK + world min + self pairs; value = FP64 on every core; median of 5 runs",
            "legs_configs_per_s": legs, "cpu_model": cpu_model()}, out


def load_traffic():
    """HBM bytes per sweep launch from the committed rocprofv3 counter summary (or None)."""
    path = os.path.join(ROOT, "profiles", "sphere_sweep_fk_panda_b65536_m256.md")
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":")[1])
    return None


def profile_workload(name: str):
    """Short runs for rocprofv3: 20 adjoint launches of the pose + hinge cost workload (16,384
    configurations, 512 world spheres; plus the copy calibration), or a few solves of the uneven
    LM-IK workload (4096 targets over four robots) or of a 512-trajectory workload."""
    import torch

    if name == "grad":
        *_, run = grad_workload()
        for _ in range(20):
            run()
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)  # reads 256 MiB, writes 256 MiB
    elif name == "ik":
        *_, solve = ik_workload()
        for _ in range(20):
            solve()
    elif name == "ik_uneven":
        from pyroffi_amd.optimization_engines._ls_ik_batched import UnevenLsIkSolver

        groups, _ = uneven_workload()
        solver = UnevenLsIkSolver(groups, max_iter=50)
        for g, tune in zip(groups, ((4096, 0), (2048, 0), (2048, 0), (1024, 0))):
            g.tune = tune
        for _ in range(3):
            solver.solve()
    elif name == "chomp":
        from pyroffi_amd.optimization_engines._chomp_optimization import ChompTrajOptConfig, chomp_run
        from pyroffi_amd.optimization_engines._lbfgs_trajopt_optimization import TrajModel, flatten_trajs

        robot, coll, world, trajs, *_ = trajopt_workload()
        cfg = ChompTrajOptConfig()
        model = TrajModel(robot, coll, world, cfg.collision_margin, mode=0)
        x, start, max_len = flatten_trajs(trajs, model.device)
        for _ in range(3):
            chomp_run(model, x, start, max_len, cfg)
    elif name == "sco":
        from pyroffi_amd.optimization_engines._sco_optimization import BatchedSco, ScoTrajOptConfig

        robot, coll, world, trajs, *_ = trajopt_workload()
        opt = BatchedSco(robot, coll, world, trajs, ScoTrajOptConfig())
        for _ in range(3):
            opt.solve()
    else:
        *_, opt = trajopt_workload()
        for _ in range(3):
            opt.solve()
    torch.cuda.synchronize()
    print(json.dumps({"profile": name}), flush=True)


def profile_run():
    import torch

    g = gpu_bench(steps=200, warmup=20)  # the default run's step counts
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)  # reads 256 MiB, writes 256 MiB
    torch.cuda.synchronize()
    print(json.dumps({"profile": True, "ms_per_step": g["ms_per_step"], "sweep_ms": g["sweep_ms"],
                      "fk_ms": g["fk_ms"], "copy_bytes_each_way": src.numel() * 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=B)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ik", action="store_true")
    ap.add_argument("--profile", nargs="?", const="sweep", choices=("sweep", "grad", "ik", "ik_uneven", "trajopt", "chomp", "sco"),
                    help="run for rocprofv3: 'sweep' = the headline workload's 20 + 200 steps and 200 timed "
                         "launches of each kernel, plus a 256 MiB device copy (the byte-count calibration for the HBM counters); 'grad' = 20 launches "
                         "of the pose + hinge adjoint kernel plus the same copy; 'ik' = 20 solve calls of 16 targets x 64 seeds; "
                         "'ik_uneven' / 'trajopt' / 'chomp' / 'sco' = 3 solves "
                         "of the uneven LM-IK / trajectory workload (L-BFGS, CHOMP, SCO)")
    args = ap.parse_args()
    if args.profile == "sweep":
        profile_run()
        return
    if args.profile:
        profile_workload(args.profile)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the runs here have one "
              "GPU (the workload shards by configuration with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)

    g = gpu_bench(args.steps, args.warmup)
    value = B / (g["ms_per_step"] * 1e-3)
    achieved = FLOP_SWEEP * B / (g["sweep_ms"] * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "configs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "panda_fk_sphere_collision_b65536_world256",
            "batch": B,
            "world_spheres": M_WORLD,
            "robot_spheres": K_SPHERES,
            "self_pairs": 66,
            "outputs": "link poses (B,13,7) + per-config world min (B) + self pair distances (B,66)",
            "parallelism": "single_gpu",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "sweep_kernel (pk_sphere_sweep_f32)",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": load_traffic(),
            "flop_per_config": FLOP_SWEEP,
            "avg_launch_ms": g["sweep_ms"],
            "peak_without_packed_fp32": PEAK_FP32_UNPACKED_TFLOPS,
            "frac_without_packed_fp32": achieved / PEAK_FP32_UNPACKED_TFLOPS,
        },
        "fk_kernel": {
            "bound": "hbm",
            "avg_launch_ms": g["fk_ms"],
            "achieved": BYTES_FK * B / (g["fk_ms"] * 1e-3) / 1e9,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
        },
    }
    _progress(f"sweep {g['ms_per_step']:.4f} ms/step")
    if not args.no_cpu_baseline:
        cb, _ = cpu_baseline(args.cpu_sample)
        line["cpu_baseline"] = cb
        _progress("cpu baseline done")
    if not args.no_ik:
        for key, fn in (("fk_small", lambda: fk_small_bench()), ("ik", lambda: ik_bench(reps=10)),
                        ("grad", lambda: grad_bench(reps=10, cpu_sample=256)),
                        ("ik_uneven", lambda: ik_uneven_bench(reps=5)), ("hjcd", lambda: hjcd_bench(reps=3)),
                        ("trajopt", lambda: trajopt_bench(reps=5)), ("chomp", lambda: chomp_bench()),
                        ("sco", lambda: sco_bench())):
            line[key] = fn()
            _progress(f"{key} done")
    print(json.dumps(line), flush=True)


def _progress(msg: str) -> None:
    """One line on stderr per finished leg (the JSON result stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
