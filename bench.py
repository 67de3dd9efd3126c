"""Benchmark of the GSO hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]

A step is one GSO step (pose transform + DFIRE scoring + luciferin, then
neighbour selection + movement) over every glowworm of every swarm of the
workload; inputs and state are resident in HBM before timing starts.  The
headline form is the batched FP32 guard-band form, its step captured once as
a hipGraph (before the timed region) and replayed --steps times; the plain
FP64 form is timed on the same workload for comparison, and the whole c0 run
on the CPU oracle and on the GPU.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from sparkledock_amd import ranking, sdgso, workloads  # noqa: E402

METRIC = "glowworm-pose evals/sec on 1 MI355X, plain FP64 vs batched FP32-guarded"
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def hip_sync():
    import torch
    torch.cuda.synchronize()


def timed_steps(run, n, graph):
    """Steps on torch's current stream, bracketed by device synchronisation."""
    hip_sync()
    t0 = time.perf_counter()
    run.steps(n, graph=graph)
    run.sync()
    hip_sync()
    return time.perf_counter() - t0


def precision_check(wl, n_per_complex=64, seed=5):
    """FP32-guard vs FP64 (the FP64 form is bit-exact to the CPU oracle in the
    parity tests) on a sample of poses of every complex: pair counts (bins)
    and energies relative to the sum of |T| over in-cutoff pairs."""
    rng = np.random.default_rng(seed)
    pos = wl["pos"].reshape(-1, wl["complexes"][0].P)
    G = wl["G"]
    start = 0
    bins_equal = 0
    n = 0
    worst_rel = 0.0
    worst_rel_score = 0.0
    rel_e = []
    tab_abs = np.abs(sdgso.synth_table(workloads.TABLE_SEED))
    for cx, S, inp in zip(wl["complexes"], wl["n_swarms"], wl["inputs"]):
        idx = start + rng.choice(S * G, size=min(n_per_complex, S * G), replace=False)
        poses = pos[idx]
        s64, _, p64 = cx.score(poses, sdgso.FP64)
        s32, _, p32 = cx.score(poses, sdgso.FP32_GUARD)
        cabs = sdgso.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"], tab_abs,
                             inp["lig_ref"], inp["r_modes"], inp["l_modes"])
        sabs, _, _ = cabs.score(poses, sdgso.FP64)
        cabs.close()
        denom = np.maximum((4.7 - sabs) / 0.0157, 1.0)
        e64, e32 = (4.7 - s64) / 0.0157, (4.7 - s32) / 0.0157
        worst_rel = max(worst_rel, float(np.max(np.abs(e32 - e64) / denom)))
        worst_rel_score = max(worst_rel_score,
                              float(np.max(np.abs(s32 - s64) / np.maximum(np.abs(s64), 1e-300))))
        rel_e.append(np.abs(e32 - e64) / np.maximum(np.abs(e64), 1e-300))
        bins_equal += int(np.sum(p64 == p32))
        n += len(idx)
        start += S * G
    rel_e = np.concatenate(rel_e)
    return dict(poses=n, pair_counts_equal=bins_equal, energy_rel_err_sum_abs=worst_rel,
                energy_rel_err_max=float(rel_e.max()), energy_rel_err_median=float(np.median(rel_e)),
                energy_rel_err_le_1e6=float(np.mean(rel_e <= 1e-6)), score_rel_err=worst_rel_score)


def ranking_overlap(stream, workload="c1", steps=30):
    """End-to-end: the same seeded run in both forms for `steps` steps; overlap
    of the top-100 glowworms ranked by final score and by luciferin
    (analysis/lgd_rank.py:311-336 ordering)."""
    wl = workloads.build(workload, steps=steps)
    out = {}
    finals = {}
    for form in (sdgso.FP32_GUARD, sdgso.FP64):
        r = sdgso.Run(wl["complexes"], wl["n_swarms"], wl["G"], wl["pos"], wl["stream"], steps,
                      form=form)
        r.set_stream(stream.cuda_stream)
        if form == sdgso.FP32_GUARD:
            r.tune()
        r.steps(steps, graph=True)
        finals[form] = r.read()
        r.close()
    a, b = finals[sdgso.FP32_GUARD], finals[sdgso.FP64]
    out["workload"] = workload
    out["steps"] = steps
    out["top100_overlap_by_scoring"] = ranking.top_overlap(a["score"], b["score"], 100)
    out["top100_overlap_by_luciferin"] = ranking.top_overlap(a["luciferin"], b["luciferin"], 100)
    out["selections_equal_last_step"] = float(np.mean(a["select"] == b["select"]))
    out["max_abs_pose_diff"] = float(np.max(np.abs(a["pos"] - b["pos"])))
    return out


def host_cpu():
    """The host's CPU model, its logical CPU count and the CPUs this process
    may run on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(cpu_model=model, nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)))


def cpu_baseline(wl, threads, target_s=12.0):
    """The CPU oracle (FP64 restatement of the reference CPU step) on a bounded
    sample: one full step (score, select, move) over `threads` swarms of g
    glowworms of every complex, with the workload's own poses and its random
    stream's first step; per-evaluation cost extrapolates to the whole
    workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    G = wl["G"]
    P = wl["complexes"][0].P
    pos = wl["pos"].reshape(-1, G, P)
    ustream = wl["stream"].reshape(-1, wl["steps"], G)
    # calibrate g so that the sample takes about target_s
    swarm0 = 0
    cost = 0.0
    samples = []
    for inp, S in zip(wl["inputs"], wl["n_swarms"]):
        cx = O.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"], inp["table"],
                       inp["lig_ref"], inp["r_modes"], inp["l_modes"], inp["r_mask"], inp["l_mask"])
        samples.append((cx, swarm0, S))
        cost += S * len(inp["r_xyz"]) * len(inp["l_xyz"])
        swarm0 += S
    # pairs per second of this host on all threads, from a small probe
    cx, s0, S = samples[0]
    ns = min(threads, S)
    st = O.State(ns, 2, P, O.default_params(), pos[s0:s0 + ns, :2].copy())
    u = ustream[s0:s0 + ns, :1, :2].copy()
    t0 = time.perf_counter()
    O.lib().or_gso_step(cx.c, O.default_params(), st.st, O._p(np.ascontiguousarray(u)), 1, 0, 1,
                        threads)
    rate = ns * 2 * cx.nr * cx.nl / max(time.perf_counter() - t0, 1e-6)
    # glowworms per sampled swarm so that the sample takes about target_s
    pairs_per_g = sum(min(threads, S2) * cx2.nr * cx2.nl for cx2, _, S2 in samples)
    g = int(max(1, min(G, target_s * rate / pairs_per_g)))
    t_total = 0.0
    evals = 0
    step_time = 0.0  # extrapolated time of one full workload step
    for cx, s0, S in samples:
        ns = min(threads, S)
        st = O.State(ns, g, P, O.default_params(), pos[s0:s0 + ns, :g].copy())
        u = ustream[s0:s0 + ns, :1, :g].copy()
        t0 = time.perf_counter()
        O.lib().or_gso_step(cx.c, O.default_params(), st.st, O._p(np.ascontiguousarray(u)), 1, 0,
                            1, threads)
        dt = time.perf_counter() - t0
        t_total += dt
        evals += ns * g
        step_time += dt / (ns * g) * S * G
    total_evals = sum(S for _, _, S in samples) * G
    out = dict(value=total_evals / step_time, unit="evals/s", cores=threads, kind="port",
               sample="one GSO step over %d swarms x %d glowworms of each of %d complexes "
                      "(%d evals, %.1f s; the workload's poses and its stream's first step); "
                      "rate extrapolated by per-complex evaluation cost"
                      % (min(threads, min(S for _, _, S in samples)), g, len(samples), evals,
                         t_total))
    out.update(host_cpu())
    return out


def cpu_c0_full(threads):
    """The whole c0 run (757/481 atoms, 130 x 200 glowworms, 7-value poses, 20
    steps, seed 0) on the CPU oracle with `threads` OpenMP threads, scoring
    only moved glowworms after the first step as the reference CPU path does
    (prepare_gso_tasks.cc:434-436); evals/s counts S*G*steps as every form
    here does."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    wl = workloads.build("c0", device=False)
    inp = wl["inputs"][0]
    cx = O.Complex(inp["r_xyz"], inp["r_type"], inp["l_xyz"], inp["l_type"], inp["table"],
                   inp["lig_ref"], inp["r_modes"], inp["l_modes"], inp["r_mask"], inp["l_mask"])
    S, G, steps = wl["n_swarms"][0], wl["G"], wl["steps"]
    st = O.State(S, G, cx.P, O.default_params(), wl["pos"])
    u = np.ascontiguousarray(wl["stream"])
    scored = 0
    t0 = time.perf_counter()
    for k in range(steps):
        scored += S * G if k == 0 else int(st.get()["moved"].sum())
        O.lib().or_gso_step(cx.c, O.default_params(), st.st, O._p(u), steps, k, 0, threads)
    dt = time.perf_counter() - t0
    out = dict(value=S * G * steps / dt, unit="evals/s", seconds=dt, cores=threads, kind="port",
               evaluations_performed=scored, evaluations_counted=S * G * steps,
               workload="c0 in full: 757/481 atoms, 130 x 200 glowworms, 20 steps, FP64")
    out.update(host_cpu())
    return out


def gpu_c0_full(stream):
    """The whole c0 run on the GPU in both forms (eager steps, as one run)."""
    out = {}
    for name, form in (("fp64", sdgso.FP64), ("fp32_guard", sdgso.FP32_GUARD)):
        wl = workloads.build("c0")
        r = sdgso.Run(wl["complexes"], wl["n_swarms"], wl["G"], wl["pos"], wl["stream"],
                      wl["steps"], form=form)
        r.set_stream(stream.cuda_stream)
        dt = timed_steps(r, wl["steps"], graph=False)
        out[name] = dict(value=wl["n_swarms"][0] * wl["G"] * wl["steps"] / dt, unit="evals/s",
                         seconds=dt)
        r.close()
    return out


def traffic_bytes(workload):
    """HBM bytes per scoring launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, calibrated), if any."""
    p = os.path.join(ROOT, "profiles", "score32_%s_traffic.md" % workload)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("bytes per launch:"):
            return int(line.split(":")[1].split()[0])
    return None


def profile_valu_insts(workload):
    """VALU instructions per scoring launch (SQ_INSTS_VALU) from the committed
    rocprofv3 counter summary of this workload, if any."""
    p = os.path.join(ROOT, "profiles", "score32_%s_counters.md" % workload)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("| SQ_INSTS_VALU |"):
            return float(line.split("|")[2].split()[0])
    return None


# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(workloads.CONFIGS))
    ap.add_argument("--swarm-scale", type=float, default=1.0)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-anm-probe", action="store_true",
                    help="skip timing the receptor displacement as an FP32 matrix-core GEMM "
                         "(sdgso_probe_anm_gemm) on the timed run's poses")
    ap.add_argument("--no-c0", action="store_true",
                    help="skip the full c0 runs (CPU oracle and both GPU forms)")
    ap.add_argument("--ranking-steps", type=int, default=30,
                    help="steps of the end-to-end FP32-vs-FP64 ranking comparison on c1 (0 = skip)")
    ap.add_argument("--variant", default=None,
                    help="FP32 scoring-kernel variant (name, e.g. bcast2/ring4/occ4, or index) "
                         "to use instead of the autotuner's choice")
    ap.add_argument("--xcd-chunk", type=int, default=None,
                    help="scoring work order (0 list order, K chunks of K items per XCD) instead "
                         "of the tuner's choice")
    ap.add_argument("--ligand-gemm", type=int, choices=(0, 1), default=None,
                    help="ligand displacement GEMM off/on instead of the tuner's choice")
    ap.add_argument("--profile-fp64-step", action="store_true",
                    help="only build the workload and run one plain FP64 step of each complex "
                         "(for rocprofv3 --pmc runs of the FP64 kernel)")
    ap.add_argument("--profile-step", action="store_true",
                    help="only build the workload and run one FP32 step (for rocprofv3 --pmc runs), "
                         "after a 1 GiB device copy used to calibrate FETCH_SIZE/WRITE_SIZE, "
                         "then the random-gather probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the GSO path has no multi-GPU form in this build "
              "(swarms would shard across GPUs with no exchange; see the design notes). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        sys.exit("no GPU visible")
    torch.cuda.set_device(0)
    # a stream of our own (default priority) as torch's current stream: the
    # runs launch on it and torch's synchronisation covers it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    # the library is loaded after torch has set up the device
    if args.variant is not None:
        names = sdgso.tune_variants()
        if args.variant.isdigit() and int(args.variant) < len(names):
            args.variant = int(args.variant)
        elif args.variant in names:
            args.variant = names.index(args.variant)
        else:
            sys.exit("unknown variant %r; compiled variants: %s" % (args.variant, ", ".join(names)))

    if args.profile_fp64_step:
        wl = workloads.build(args.workload, swarm_scale=args.swarm_scale, steps=1)
        for cx, S, k in zip(wl["complexes"], wl["n_swarms"], range(len(wl["complexes"]))):
            off = sum(wl["n_swarms"][:k]) * wl["G"] * cx.P
            r64 = sdgso.Run(cx, S, wl["G"], wl["pos"][off:off + S * wl["G"] * cx.P],
                            sdgso.stream(0, S, 1, wl["G"]), 1, form=sdgso.FP64)
            r64.steps(1)
            r64.sync()
            r64.close()
        log("profile FP64 step done: %d complexes, %d all-pairs evaluations"
            % (len(wl["complexes"]),
               sum(S * wl["G"] * cx.nr * cx.nl for cx, S in zip(wl["complexes"], wl["n_swarms"]))))
        return

    if args.profile_step:
        x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.mul(x, 2.0)  # 1 GiB read + 1 GiB write: the counter calibration kernel
        torch.cuda.synchronize()
        del x, y
        wl = workloads.build(args.workload, swarm_scale=args.swarm_scale, steps=2)
        run = sdgso.Run(wl["complexes"], wl["n_swarms"], wl["G"], wl["pos"], wl["stream"],
                        wl["steps"], form=sdgso.FP32_GUARD)
        if args.variant is not None:
            run.set_variant(args.variant)
        if args.xcd_chunk is not None:
            run.set_work_order(args.xcd_chunk)
        if args.ligand_gemm is not None:
            run.set_ligand_gemm(bool(args.ligand_gemm))
        run.steps(1)
        run.sync()
        log("profile step done: %d glowworms, %d in-cutoff pairs in the scoring launch"
            % (sum(wl["n_swarms"]) * wl["G"], int(run.read()["pairs"].sum())))
        # the random-gather ceiling kernel, for counters to compare against
        log("gather probe: %.3g gathers/s" % sdgso.probe_gather())
        return

    t0 = time.perf_counter()
    # stream steps: warm-up, the graphs' first (untimed) replays, the timed
    # replays, then replays of the one-graph step and eager steps to compare
    cmp_steps = max(1, min(args.steps, 5))
    wl = workloads.build(args.workload, swarm_scale=args.swarm_scale,
                         steps=args.warmup + 2 + args.steps + 2 * cmp_steps + 3)
    n_glow = sum(wl["n_swarms"]) * wl["G"]
    log("workload %s: %d complexes, %d glowworms, built in %.1f s" %
        (args.workload, len(wl["complexes"]), n_glow, time.perf_counter() - t0))

    # ---- batched FP32 guard-band form (headline)
    run = sdgso.Run(wl["complexes"], wl["n_swarms"], wl["G"], wl["pos"], wl["stream"], wl["steps"],
                    form=sdgso.FP32_GUARD)
    run.set_stream(stream.cuda_stream)
    run.steps(args.warmup)
    run.sync()
    if args.variant is None:
        chosen, tune_ms = run.tune()
        log("autotune: variants per complex %s" % chosen)
        log("autotune: ms per launch per complex %s" % tune_ms)
    else:
        run.set_variant(args.variant)
        chosen, tune_ms = run.variants(), None
    if args.xcd_chunk is not None:
        run.set_work_order(args.xcd_chunk)
    if args.ligand_gemm is not None:
        run.set_ligand_gemm(bool(args.ligand_gemm))
    # Timed region: args.steps replays of the captured step, with profiling
    # on, so that each replay is the scoring graph and the graph of the rest
    # of the step with HIP events on the run's stream between them (the
    # kernel's launch durations come from the timed region itself).  Both
    # graphs are captured, instantiated and replayed once before t0.
    run.profile(True)
    run.steps(1, graph=True)
    run.sync()
    run.profile(True)  # drop the untimed replay's events
    dt = timed_steps(run, args.steps, graph=True)
    prof = run.profile_read()
    guard = run.guard_stats()
    run.profile(False)
    value = n_glow * args.steps / dt
    st = run.read()
    pairs_per_launch = float(st["pairs"].sum())
    score_ms = prof["score_ms"] / max(1, prof["score_launches"])
    rest_ms = prof["move_ms"] / max(1, prof["move_launches"])
    lgemm_ms = prof["lgemm_ms"] / max(1, prof["lgemm_launches"])
    lgemm_on, lgemm_bytes = run.ligand_gemm()
    # the same step as one graph (no events) and eagerly, for comparison
    run.steps(1, graph=True)
    dt1 = timed_steps(run, cmp_steps, graph=True)
    dte = timed_steps(run, cmp_steps, graph=False)
    filt = run.filter_stats()
    anm = None
    if not args.no_anm_probe and max(cx.nmr for cx in wl["complexes"]) > 0:
        anm = {k: float(v) for k, v in run.probe_anm_gemm().items()}
        anm["fused_displacements"] = filt["near_atoms"]
        anm["gemm_TFLOPs"] = anm["flops"] / (anm["ms"] * 1e-3) / 1e12
        anm["peak_TFLOPs"] = FP32_PEAK_TFLOPS
        anm["frac"] = anm["gemm_TFLOPs"] / FP32_PEAK_TFLOPS
        anm["write_GBps"] = anm["bytes_written"] / (anm["ms"] * 1e-3) / 1e9
        # the whole (poses x modes) . (modes x 3N) products of one step, as the
        # reference materialises them (receptor and ligand)
        anm["full_gemm_flops_per_step"] = float(sum(
            2.0 * S * wl["G"] * (cx.nmr * 3 * cx.nr + cx.nml * 3 * cx.nl)
            for cx, S in zip(wl["complexes"], wl["n_swarms"])))
        anm["form"] = ("receptor displacement of the tiles each swarm reaches as "
                       "v_mfma_f32_32x32x2f32 products, written as FP32 coordinates; the "
                       "scoring kernel instead displaces the fused_displacements atoms that "
                       "pass its atom test")
        log("anm gemm probe:", anm)
    log("fp32 guard: %.2f ms/step (profiled graph replays: ligand GEMM %.2f ms + scoring %.2f "
        "ms + rest %.2f ms); one-graph replays %.2f ms/step, eager %.2f ms/step; pairs/launch "
        "%.3e, guard recomputes %d" % (1e3 * dt / args.steps, lgemm_ms, score_ms, rest_ms,
                                       1e3 * dt1 / cmp_steps, 1e3 * dte / cmp_steps,
                                       pairs_per_launch, guard["recomputed"]))
    # the ligand displacement GEMM, (poses x modes) . (modes x 3 * 64 slots) per
    # ligand block on the matrix cores with its FP32 output written to HBM: the
    # tuner keeps it when a scoring pass with it is faster; when it did not,
    # three more profiled replays with it on measure it (after the timed region)
    lgemm = None
    if max(cx.nml for cx in wl["complexes"]) > 0:
        t_with, t_without = run.ligand_gemm_tune_ms()
        lg = dict(ms=lgemm_ms, score_ms=score_ms, bytes=lgemm_bytes)
        if not lgemm_on:
            run.set_ligand_gemm(True)
            run.profile(True)
            run.steps(1, graph=True)
            run.sync()
            run.profile(True)
            run.steps(2, graph=True)
            p2 = run.profile_read()
            lg = dict(ms=p2["lgemm_ms"] / max(1, p2["lgemm_launches"]),
                      score_ms=p2["score_ms"] / max(1, p2["score_launches"]),
                      bytes=run.ligand_gemm()[1])
            run.profile(False)
            run.set_ligand_gemm(False)
        lg_flops = float(sum(2.0 * S * wl["G"] * cx.nml * 3 * 64 * -(-cx.nl // 64)
                             for cx, S in zip(wl["complexes"], wl["n_swarms"])))
        lgemm = {"on_in_timed_steps": lgemm_on,
                 "tuner_ms_per_scoring_pass": {"with": t_with, "without": t_without},
                 "ms": lg["ms"], "scoring_ms_with_it": lg["score_ms"],
                 "bytes_written": lg["bytes"],
                 "write_GBps": lg["bytes"] / (lg["ms"] * 1e-3) / 1e9,
                 "write_frac_of_hbm_peak": lg["bytes"] / (lg["ms"] * 1e-3) / 8e12,
                 "flops": lg_flops, "TFLOPs": lg_flops / (lg["ms"] * 1e-3) / 1e12,
                 "frac_of_fp32_mfma_peak": lg_flops / (lg["ms"] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                 "kernel": "ligand_disp_kernel (v_mfma_f32_32x32x2f32)",
                 "bound": "hbm write of its output"}
        log("ligand gemm:", lgemm)

    # ---- plain FP64 per-complex form on the same workload
    fp64 = None
    if not args.no_fp64 and args.fp64_steps > 0:
        t64 = 0.0
        for cx, S, k in zip(wl["complexes"], wl["n_swarms"], range(len(wl["complexes"]))):
            off = sum(wl["n_swarms"][:k]) * wl["G"] * cx.P
            pos_k = wl["pos"][off:off + S * wl["G"] * cx.P]
            # one untimed step first (first launches on this run's buffers)
            st_k = sdgso.stream(0, S, args.fp64_steps + 1, wl["G"])
            r64 = sdgso.Run(cx, S, wl["G"], pos_k, st_k, args.fp64_steps + 1, form=sdgso.FP64)
            r64.set_stream(stream.cuda_stream)
            r64.steps(1, graph=False)
            t64 += timed_steps(r64, args.fp64_steps, graph=False)
            r64.close()
        # the brute-force kernel evaluates the difference-form d2 of every pair
        all_pairs = sum(S * wl["G"] * cx.nr * cx.nl for cx, S in zip(wl["complexes"], wl["n_swarms"]))
        tf64 = 8.0 * all_pairs * args.fp64_steps / t64 / 1e12
        fp64 = dict(value=n_glow * args.fp64_steps / t64,
                    ms_per_step=1e3 * t64 / args.fp64_steps, steps=args.fp64_steps,
                    form="one run per complex, brute force over all pairs in FP64",
                    roofline={"bound": "fp64 valu", "achieved": tf64, "peak": FP64_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": tf64 / FP64_PEAK_TFLOPS,
                              "algorithmic": "8 FP64 flop per receptor-ligand pair, all Nr*Nl pairs",
                              # the reference rounds every product and sum (no FMA), so the
                              # 8 flop are 8 instructions, and the FMA-counted peak is twice
                              # what they can reach
                              "issue": {"achieved_Tops": tf64, "peak_Tops": FP64_PEAK_TFLOPS / 2,
                                        "frac": tf64 / (FP64_PEAK_TFLOPS / 2),
                                        "note": "non-FMA FP64 instructions at one per lane per "
                                                "cycle: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz"}})
        log("fp64 plain: %.1f ms/step" % fp64["ms_per_step"])

    prec = precision_check(wl)
    rank_cmp = None
    if args.ranking_steps > 0:
        rank_cmp = ranking_overlap(stream, "c1", args.ranking_steps)
        log("ranking:", rank_cmp)
    log("precision:", prec)

    cpu = c0cpu = c0gpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
        cpu = cpu_baseline(wl, threads)
        log("cpu baseline:", cpu)
        if not args.no_c0:
            c0cpu = cpu_c0_full(threads)
            log("cpu c0 in full:", c0cpu)
    if not args.no_c0:
        c0gpu = gpu_c0_full(stream)
        log("gpu c0 in full:", c0gpu)

    flops = 8.0 * pairs_per_launch
    achieved = flops / (score_ms * 1e-3) / 1e12
    # the table reads: 4 bytes per in-cutoff pair against the measured random-gather rate
    probe = sdgso.probe_gather()
    gather_gbs = 4.0 * pairs_per_launch / (score_ms * 1e-3) / 1e9
    probe_gbs = 4.0 * probe / 1e9
    gather_frac = gather_gbs / probe_gbs
    flop_frac = achieved / FP32_PEAK_TFLOPS
    # VALU issue, the co-bound: the committed counter profile's instructions
    # per launch over this run's launch time
    vi = profile_valu_insts(args.workload)
    valu_issue = None
    if vi:
        valu_issue = {"insts_per_launch": vi, "achieved_per_s": vi / (score_ms * 1e-3),
                      "peak_per_s": VALU_ISSUE_PEAK,
                      "frac": vi / (score_ms * 1e-3) / VALU_ISSUE_PEAK,
                      "source": "SQ_INSTS_VALU of profiles/score32_%s_counters.md over this "
                                "run's launch time; peak 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles"
                                % args.workload}
    names = sdgso.tune_variants()
    chosen_names = [names[v] for v in chosen]
    # per complex: in-cutoff pairs of the last scoring and the tuner's launch
    # time of its share with the chosen variant
    per_complex = []
    g0 = 0
    for k, (cx, S) in enumerate(zip(wl["complexes"], wl["n_swarms"])):
        npairs = float(st["pairs"][g0:g0 + S * wl["G"]].sum())
        g0 += S * wl["G"]
        ms = tune_ms[chosen_names[k]][k] if tune_ms else None
        per_complex.append({"atoms": [cx.nr, cx.nl], "swarms": S, "pairs": npairs,
                            "variant": chosen_names[k], "launch_ms": ms,
                            "gather_GBps": (4.0 * npairs / (ms * 1e-3) / 1e9) if ms else None})
    # the variant that scores the most work (waves) names the dominant kernel
    blocks = [S * wl["G"] * -(-cx.nl // 64) for cx, S in zip(wl["complexes"], wl["n_swarms"])]
    share = {}
    for v, w in zip(chosen_names, blocks):
        share[v] = share.get(v, 0) + w
    variant_name = max(share, key=share.get)
    kernel_name = "score32b_kernel"  # every FP32 variant is a form of it
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt / args.steps,
        "gso_steps_per_s": args.steps / dt,
        "timing": {"timed": "%d replays of the captured step (capture, instantiation and a "
                            "first replay before t0), each as the graphs of the ligand "
                            "displacement GEMM, the scoring launches and the rest of the step, "
                            "with HIP events between them"
                            % args.steps,
                   "ligand_gemm_ms": lgemm_ms if lgemm_on else None,
                   "scoring_ms": score_ms, "rest_ms": rest_ms,
                   "one_graph_ms_per_step": 1e3 * dt1 / cmp_steps,
                   "eager_ms_per_step": 1e3 * dte / cmp_steps, "compare_steps": cmp_steps},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 guard-band recompute, FP64 lane sums)",
        "data": "synthetic, seeded (random-walk globules, N(0,1)*0.5 table)",
        "config": {"workload": args.workload, "complexes": len(wl["complexes"]),
                   "swarms": int(sum(wl["n_swarms"])), "glowworms_per_swarm": wl["G"],
                   "modes": list(wl["spec"]["modes"]), "guard_band_A2": 1e-4,
                   "parallelism": "1 GPU, all swarms of all complexes in one run"},
        # The scoring kernel is bound by its data-dependent 4-byte table
        # gathers, served by L2 (the 2.26 MB table stays in each XCD's 4 MB
        # L2); there is no published L2 random-gather figure, so the peak is
        # the measured rate of a pure gather probe over a table of that size.
        # frac = max(gather fraction, FP32 flop fraction).
        "roofline": {"bound": "l2-gather", "achieved": gather_gbs, "peak": probe_gbs,
                     "unit": "GB/s", "frac": max(gather_frac, flop_frac),
                     "traffic": traffic_bytes(args.workload),
                     "kernel": kernel_name, "variant": variant_name, "launch_ms": score_ms,
                     "launch_ms_source": "HIP events on the run's stream around the scoring "
                                         "graph of each timed replay",
                     "algorithmic": "4 bytes of table per in-cutoff pair (0 < d2 <= 225), "
                                    "%.4e pairs per launch" % pairs_per_launch,
                     "peak_source": "sdgso_probe_gather: 4-byte gathers at hashed offsets over "
                                    "a 2.26 MB table, every CU busy",
                     "vs_hbm_peak": {"achieved_GBps": gather_gbs, "peak_GBps": 8000.0,
                                     "frac": gather_gbs / 8000.0},
                     "valu_issue": valu_issue,
                     "flop": {"achieved_TFLOPs": achieved, "peak_TFLOPs": FP32_PEAK_TFLOPS,
                              "frac": flop_frac,
                              "algorithmic": "8 FP32 flop per in-cutoff pair (difference-form "
                                             "d2)"}},
        "cpu_baseline": cpu,
        "c0_in_full": {"cpu": c0cpu, "gpu": c0gpu},
        "plain_fp64": fp64,
        "speedup_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "precision": prec,
        "ranking": rank_cmp,
        "guard_recomputes_per_step": guard["recomputed"] / max(1, args.steps),
        "pairs_per_eval": pairs_per_launch / n_glow,
        "evaluations": {"policy": "one per glowworm per step (swarms x glowworms x steps), as the "
                                  "reference GPU path scores; the reference CPU path re-scores only "
                                  "the glowworms that moved",
                        "moved_fraction_last_step": float(np.mean(st["moved"]))},
        "filter": filt,
        "per_complex": per_complex,
        "ligand_gemm": lgemm,
        "anm_gemm_probe": anm,
        "autotune": {"chosen_per_complex": chosen_names,
                     "work_order": {"xcd_chunk": run.work_order()[0],
                                    "tuner_ms_per_scoring_pass": dict(zip(
                                        ("list_order", "xcd_chunked"), run.work_order()[1]))},
                     "ms_per_launch": ({k: sum(v) for k, v in tune_ms.items()} if tune_ms else None)},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
