"""Benchmark of the docking hot path on one MI355X.

Headline (value): batched persistent docking throughput in ligands/s on the batched screening
workload -- 8,000-atom synthetic receptor (60 A cube, 10 A cavity), 30 A box at 0.375 A (81^3
points, all 17 FP32 maps resident in HBM), 1,024 synthetic ligands (10-60 heavy atoms,
0..min(16, atoms/3) torsions, seeds 0..1023), 256 Monte Carlo chains each with RILC-BFGS local
search.  One step = docking all 1,024 ligands.

Printed beside it on the same JSON line:
  per_ligand_plain  the plain per-ligand launch loop and the batched launch on the same ligand
                    subset, with their best scores (equal-accuracy check)
  grid              17-map precompute of the same receptor/box: kernel time and GB/s
  adjoint           forward + adjoint evaluations of 2^20 random conformations of a 40-atom,
                    12-torsion ligand (evals/s)
  roofline          the dominant kernel (the MC kernel) against HBM peak (8 TB/s, and beside it
                    the bandwidth a timed 1 GiB device-to-device copy reaches), algorithmic bytes; a
                    step's MC launches (one per lanes-per-chain class) run at the same time on
                    their own streams, so the unit is the step's set of MC launches: their
                    algorithmic bytes over their span (first start to last end, HIP events)
  post              the step after the search for the docked batch: pose lists, FP64 refinement of
                    every pose against the receptor atoms in one launch, final scores; beside it
                    the oracle's FP64 refinement of a sample of the same poses on host threads
  quickvina_w       the same screening step with the QuickVina-W policy (evaluations saved by its
                    gate, throughput, mean best score)
  large_box         large-box batched screening with the QuickVina2 policy (20,000-atom receptor,
                    188^3 maps, 4,096 ligands of 10-80 heavy atoms, 128 chains)
  cpu_baseline      the FP64 oracle's Monte Carlo (same schedule, same random stream) on host
                    cores, over a bounded sample of the same ligands
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vina-cuda_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_scene(vdock, ctx):
    xyz, xs = vdock.synth_receptor(8000, 60.0, 10.0, 42)
    ctx.set_receptor(xyz, xs)
    box = vdock.box_from_center([0, 0, 0], [30.0, 30.0, 30.0], 0.375)
    ms = ctx.grid_populate(box)
    return xyz, xs, box, ms


def screening_ligands(vdock, n=1024):
    rng = np.random.default_rng(2024)
    out = []
    for i in range(n):
        heavy = int(rng.integers(10, 61))
        tors = int(rng.integers(0, min(16, heavy // 3) + 1))
        out.append(vdock.synth_ligand(heavy, tors, i))
    return out


def dominant_kernel(launches):
    """The MC kernel instantiation with the most device time over the timed steps, with its
    launch count, average launch duration and algorithmic bytes per launch (per evaluation:
    32 B per heavy atom for the eight map corners + 16 B per intramolecular pair for one
    interpolation quad; the device counts evaluations per ligand)."""
    by = {}
    for rec in launches:
        d = by.setdefault(rec.kernel, dict(n=0, ms=0.0, bytes=0.0, evals=0))
        d["n"] += 1; d["ms"] += rec.kernel_ms; d["bytes"] += rec.eval_bytes; d["evals"] += rec.evals
    name = max(by, key=lambda k: by[k]["ms"])
    d = by[name]
    return name, d, by


def large_box_leg(vdock, log, n_ligands=4096, chains=128):
    """Large-box batched screening with the QuickVina2 policy: 20,000-atom synthetic receptor
    (90 A cube, 25 A cavity), 70 A box at 0.375 A (188^3 points, 17 FP32 maps = 452 MB), 4,096
    synthetic ligands (10-80 heavy atoms, 0-24 rotatable bonds drawn, capped by the bonds the tree
    has), 128 chains each, RILC-BFGS with the visited-history gate."""
    ctx = vdock.Context(0)
    xyz, xs = vdock.synth_receptor(20000, 90.0, 25.0, 4242)
    ctx.set_receptor(xyz, xs)
    box = vdock.box_from_center([0, 0, 0], [70.0, 70.0, 70.0], 0.375)
    ctx.grid_populate(box)
    g_ms = float(np.median([ctx.grid_populate(box) for _ in range(3)]))
    rng = np.random.default_rng(4096)
    ids = []
    for i in range(n_ligands):
        heavy = int(rng.integers(10, 81))
        tors = int(rng.integers(0, 25))
        ids.append(ctx.add_ligand(vdock.synth_ligand(heavy, tors, 100000 + i)))
    p = vdock.default_params(chains=chains, seed=11, policy=1)
    log("large box: warmup")
    ctx.dock_batch(ids, p)
    log("large box: timed")
    t0 = time.perf_counter()
    st = ctx.dock_batch(ids, p)
    wall = time.perf_counter() - t0
    npts = int(np.prod(box.dims))
    gbytes = 17 * npts * 4
    tors = [ctx.ligand_info(i).n_torsions for i in ids]
    return {"workload": "large-box batched screening, QuickVina2 policy: 20000-atom receptor (90 A cube, 25 A cavity), "
                        "70 A box at 0.375 A (188^3 x 17 FP32 maps), %d ligands (10-80 heavy atoms, 0-24 rotatable "
                        "bonds drawn), %d chains each" % (n_ligands, chains),
            "ligands_per_s": round(n_ligands / wall, 3), "ms": round(1000 * wall, 2), "kernel_ms": round(st.kernel_ms, 2),
            "evals_per_s": round(st.evals / (st.kernel_ms / 1000.0), 1), "mean_torsions": round(float(np.mean(tors)), 2),
            "mean_best_score": round(float(np.mean(ctx.batch_best())), 4),
            "grid": {"points": npts, "maps_MB": round(gbytes / 1e6, 1), "kernel_ms": round(g_ms, 3),
                     "GB_per_s": round(gbytes / (g_ms / 1000.0) / 1e9, 2)}}


def copy_bandwidth(ctx, nbytes=1 << 30, reps=5):
    """HBM bandwidth a plain device-to-device copy reaches on this device: bytes read plus bytes
    written over the best of `reps` timed copies (each call waits for its copy)."""
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    try:
        ctx.copy_device(dst, src, nbytes)
        best = min(_timed(lambda: ctx.copy_device(dst, src, nbytes)) for _ in range(reps))
    finally:
        ctx.free(src)
        ctx.free(dst)
    return 2 * nbytes / best / 1e9


def _timed(f):
    t0 = time.perf_counter()
    f()
    return time.perf_counter() - t0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def mc_step_spans(step_launches):
    """Per timed step: the span of its MC launches (first start to last end, from the HIP events
    recorded around each launch on its own stream), their summed algorithmic bytes and
    evaluations, and the share of the span in which both launches ran."""
    out = []
    for L in step_launches:
        t0 = min(r.start_ms for r in L)
        t1 = max(r.start_ms + r.kernel_ms for r in L)
        ends = sorted(r.start_ms + r.kernel_ms for r in L)
        starts = sorted(r.start_ms for r in L)
        both = max(0.0, ends[0] - starts[-1]) if len(L) > 1 else 0.0
        out.append(dict(span=t1 - t0, bytes=sum(r.eval_bytes for r in L), evals=sum(r.evals for r in L),
                        overlap=both / (t1 - t0)))
    return out


def traffic_from_profile(kernel, algorithmic_bytes):
    """Memory traffic (bytes) of one launch of `kernel`: the bytes moved per algorithmic byte that
    the committed counter runs (FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy) measured for
    the same kernel on this workload, times this launch's algorithmic bytes; None if the committed
    profile has no figure for this kernel."""
    path = os.path.join(ROOT, "profiles", "vd_mc_kernel_batched_screening.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("<!-- traffic:"):
            fields = dict(f.split("=", 1) for f in line.strip("<!-> \n").split()[1:] if "=" in f)
            if fields.get("kernel") == kernel.replace(" ", ""):
                return round(float(fields["bytes_per_algorithmic_byte"]) * algorithmic_bytes)
    return None


def run_cpu_baseline(vdock, texts, box, xyz, xs, chains, budget_s=15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    threads = max(1, min(16, os.cpu_count() or 1))
    lut = orc.Lut()
    b, e, n = box.as_tuple()
    t0 = time.time()
    grid = orc.Grid(lut, b, e, n)
    grid.set_receptor(xyz.astype(np.float64), xs)
    grid.populate(True)
    t_grid = time.time() - t0
    done, t_mc, evals = 0, 0.0, 0
    for k, text in enumerate(texts):
        lig = orc.Ligand(text)
        d, bf = vdock.heuristics(lig.n_atoms, lig.T, True)
        p = orc.mc_params(chains, d, bf, b, e, seed=7, stream=k, threads=threads)
        t1 = time.time()
        _, _, _, ev = orc.mc_run(grid, lig, p)
        t_mc += time.time() - t1
        evals += ev
        done += 1
        if t_mc >= budget_s:
            break
    return dict(value=done / t_mc, unit="ligands/s", cores=threads, kind="port", cpu_model=cpu_model(),
                sample=f"first {done} of the screening ligands, {chains} chains each, FP64 oracle MC "
                       f"(same schedule and heuristics) on {threads} host threads; maps built once in {t_grid:.1f} s",
                evals_per_s=evals / t_mc)


def post_leg(vdock, ctx, log, texts, box, xyz, xs, max_poses=20, cpu_budget_s=10.0):
    """The step after the search for every ligand of the last docking batch: each ligand's chains
    reduced to at most max_poses poses (add_to_output_container), every pose refined against the
    receptor atoms in one FP64 launch, final scores and mode selection; timed on its second call
    (the first, reported as first_call_wall_ms, also builds the per-box set-up).  Beside it, the oracle's
    FP64 refinement (the reference runs this step on the CPU) of a bounded sample of the same
    poses on host threads."""
    log("post: refinement + final scores of the docked batch")
    # the first call per receptor and box also builds the receptor cell lists of the direct
    # evaluation and the pinned staging buffers (once, like the maps); the timed call follows it
    t0 = time.perf_counter()
    ctx.postprocess_docked(max_poses)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    st = ctx.postprocess_docked(max_poses)
    wall = time.perf_counter() - t0
    n_lig = len(ctx._post_ids)
    modes = [ctx.post_results(k)[0] for k in range(n_lig)]
    best = [m[0].e for m in modes if m]
    out = {"ligands": n_lig, "poses": st.blocks, "kernel_ms": round(st.kernel_ms, 2), "wall_ms": round(1000 * wall, 2),
           "first_call_wall_ms": round(1000 * first, 2),
           "poses_per_s_kernel": round(st.blocks / (st.kernel_ms / 1000.0), 1),
           "ligands_per_s_wall": round(n_lig / wall, 2), "evals": int(st.evals),
           "evals_per_s_kernel": round(st.evals / (st.kernel_ms / 1000.0), 1),
           "mean_best_final_score": round(float(np.mean(best)), 4), "dtype": "f64"}
    # CPU: the oracle's refine_structure on the same poses, ctypes calls on host threads
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from concurrent.futures import ThreadPoolExecutor
    b, e, n = box.as_tuple()
    grid = orc.Grid(orc.Lut(), b, e, n)
    grid.set_receptor(xyz.astype(np.float64), xs)
    threads = max(1, min(16, os.cpu_count() or 1))
    jobs = []
    for k, lid in enumerate(ctx._post_ids[:64]):
        lig = orc.Ligand(texts[k])
        _, conf, _ = ctx.post_results(k)
        for c in conf:
            jobs.append((lig, c.astype(np.float64)))
    done, t_cpu = 0, 0.0
    with ThreadPoolExecutor(threads) as pool:
        t1 = time.perf_counter()
        for chunk in range(0, len(jobs), threads):
            list(pool.map(lambda j: orc.refine(grid, j[0], j[1], orc.refine_steps(j[0])), jobs[chunk:chunk + threads]))
            done = min(len(jobs), chunk + threads)
            if time.perf_counter() - t1 >= cpu_budget_s:
                break
        t_cpu = time.perf_counter() - t1
    out["cpu_refine"] = {"poses_per_s": round(done / t_cpu, 2), "cores": threads, "kind": "port",
                         "sample": f"{done} refined poses of the first ligands, FP64 oracle refine_structure"}
    out["speedup_vs_cpu_poses_per_s"] = round(out["poses_per_s_kernel"] / out["cpu_refine"]["poses_per_s"], 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ligands", type=int, default=1024)
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--plain-subset", type=int, default=0, help="0 = all ligands")
    ap.add_argument("--autotune", action="store_true",
                    help="time each size bucket's launch shape (vd_set_autotune) instead of the fixed table")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-large-box", action="store_true", help="skip the large-box QuickVina2 leg")
    ap.add_argument("--no-post", action="store_true", help="skip the refinement + final-score leg")
    ap.add_argument("--grid-only", action="store_true",
                    help="only the 17-map precompute: screening scene x5, large-box scene x3 (for a kernel profile)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, copy 256 MiB device to device twice (reference dispatch for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); ligands shard across GPUs "
              "with no exchange, so run one process per GPU.", file=sys.stderr)
        sys.exit(2)

    import vdock
    ctx = vdock.Context(0)
    ctx.set_autotune(args.autotune)
    if args.grid_only:
        xyz, xs, box, _ = build_scene(vdock, ctx)
        ms = [ctx.grid_populate(box) for _ in range(5)]
        x2, s2 = vdock.synth_receptor(20000, 90.0, 25.0, 4242)
        ctx.set_receptor(x2, s2)
        b2 = vdock.box_from_center([0, 0, 0], [70.0, 70.0, 70.0], 0.375)
        ms2 = [ctx.grid_populate(b2) for _ in range(3)]
        print(json.dumps({"grid_only": True, "screening_ms": ms, "large_box_ms": ms2}), flush=True)
        return
    log("scene: receptor + maps")
    xyz, xs, box, _ = build_scene(vdock, ctx)
    log("ligands")
    texts = screening_ligands(vdock, args.ligands)
    ids = [ctx.add_ligand(t) for t in texts]
    params = vdock.default_params(chains=args.chains, seed=7)

    for w in range(args.warmup):
        log(f"warmup {w}")
        ctx.dock_batch(ids, params)
    times, kernel_ms, evals, launches, per_step = [], 0.0, 0, [], []
    st = None
    for s in range(args.steps):
        log(f"step {s}")
        t0 = time.perf_counter()
        st = ctx.dock_batch(ids, params)
        times.append(time.perf_counter() - t0)
        evals += st.evals
        per_step.append(ctx.last_launches())
        launches += per_step[-1]
    spans = mc_step_spans(per_step)
    kernel_ms = sum(x["span"] for x in spans)
    wall = sum(times)
    value = args.ligands * args.steps / wall
    best = ctx.batch_best()
    post = None if args.no_post else post_leg(vdock, ctx, log, texts, box, xyz, xs)
    out = {
        "metric": "ligands/sec on 1 MI355X, batched vs per-ligand at equal score acc; grid GB/s",
        "value": round(value, 3),
        "unit": "ligands/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * wall / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched screening: 8000-atom receptor (60 A cube, 10 A cavity), 30 A box at 0.375 A "
                        "(81^3 x 17 FP32 maps), %d ligands (10-60 heavy atoms, 0..min(16,atoms/3) torsions), "
                        "%d MC chains each, RILC-BFGS, persistent work-queue launch" % (args.ligands, args.chains),
            "ligands": args.ligands, "chains": args.chains, "maps": "81x81x81x17 fp32",
            "launch_shapes": "timed per size bucket" if args.autotune else "fixed per size bucket",
        },
        "kernel_ms_per_step": round(kernel_ms / args.steps, 2),
        "evals_per_s": round(evals / (kernel_ms / 1000.0), 1),
        "mean_best_score": round(float(np.mean(best)), 4),
    }
    name, d, by_kernel = dominant_kernel(launches)
    ms_by_kernel = {k: round(v["ms"] / v["n"], 3) for k, v in by_kernel.items()}
    per_step_bytes = sum(x["bytes"] for x in spans) / len(spans)
    avg_span = sum(x["span"] for x in spans) / len(spans)
    achieved = per_step_bytes / (avg_span / 1000.0) / 1e9
    # traffic of the step: each instantiation's measured bytes per algorithmic byte times its bytes
    tr = [traffic_from_profile(k, v["bytes"] / args.steps) for k, v in by_kernel.items()]
    traffic = round(sum(tr)) if all(t is not None for t in tr) else None
    copy_gbs = copy_bandwidth(ctx)
    out["roofline"] = {"kernel": "vd_mc_kernel (the step's %d concurrent launches: %s)" % (
                           len(per_step[0]), ", ".join(r.kernel for r in per_step[0])),
                       "bound": "hbm", "achieved": round(achieved, 2),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                       "traffic": traffic, "measured_copy_GBs": round(copy_gbs, 1),
                       "frac_of_measured_copy": round(achieved / copy_gbs, 5),
                       "steps": len(spans), "avg_span_ms": round(avg_span, 3),
                       "both_running_share_of_span": round(sum(x["overlap"] for x in spans) / len(spans), 4),
                       "algorithmic_bytes_per_step": round(per_step_bytes),
                       "evals_per_step": sum(x["evals"] for x in spans) // len(spans),
                       "avg_launch_ms_by_instantiation": ms_by_kernel, "largest_share_instantiation": name,
                       "algorithmic_bytes_by_instantiation": {k: round(v["bytes"] / v["n"]) for k, v in by_kernel.items()},
                       "launches_by_instantiation": {k: v["n"] for k, v in by_kernel.items()}}

    if post is not None:
        out["post"] = post
    if not args.no_extras:
        # per-ligand plain launch loop (one kernel_monte-style launch per ligand) on the same
        # ligands and parameters as the headline step; equal accuracy = same best scores
        if args.plain_subset and args.plain_subset < len(ids):
            sub = ids[:: max(1, len(ids) // args.plain_subset)][: args.plain_subset]
            t0 = time.perf_counter()
            st_b = ctx.dock_batch(sub, params)
            t_batch = time.perf_counter() - t0
            best_batch = ctx.batch_best()
            batch_kernel = st_b.kernel_ms
        else:
            sub, t_batch, best_batch, batch_kernel = ids, wall / args.steps, best, kernel_ms / args.steps
        log(f"plain loop over {len(sub)} ligands")
        t0 = time.perf_counter()
        st_p = ctx.dock_batch(sub, params, plain=True)
        t_plain = time.perf_counter() - t0
        best_plain = ctx.batch_best()
        out["per_ligand_plain"] = {
            "ligands": len(sub), "plain_ligands_per_s": round(len(sub) / t_plain, 3),
            "batched_ligands_per_s": round(len(sub) / t_batch, 3), "speedup": round(t_plain / t_batch, 2),
            "mean_best_score_plain": round(float(np.mean(best_plain)), 4),
            "mean_best_score_batched": round(float(np.mean(best_batch)), 4),
            "mean_abs_best_score_diff": round(float(np.mean(np.abs(best_plain - best_batch))), 4),
            "plain_kernel_ms": round(st_p.kernel_ms, 2), "batched_kernel_ms": round(batch_kernel, 2)}
        # grid precompute
        ctx.set_receptor(xyz, xs)
        gms = [ctx.grid_populate(box) for _ in range(5)]
        g_ms = float(np.median(gms))
        npts = np.prod(box.dims)
        gbytes = 17 * npts * 4 + 16 * 8000 + 153 * 2051 * 4
        out["grid"] = {"points": int(npts), "kernel_ms": round(g_ms, 4), "algorithmic_bytes": int(gbytes),
                       "GB_per_s": round(gbytes / (g_ms / 1000.0) / 1e9, 2),
                       "frac_hbm_peak": round(gbytes / (g_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 5)}
        # adjoint microbenchmark: 2^20 random conformations of a 40-atom, 12-torsion ligand,
        # generated and kept in HBM; forward energy + adjoint gradient per conformation
        lid = ctx.add_ligand(vdock.synth_ligand(40, 12, 777))
        inf = ctx.ligand_info(lid)
        T = inf.n_torsions
        n = 1 << 20
        d_conf = ctx.alloc(n * (7 + T) * 4)
        d_e = ctx.alloc(n * 4)
        d_g = ctx.alloc(n * (6 + T) * 4)
        ctx.random_confs(lid, n, 1, d_conf)
        rates = {}
        for cpw in (1, 2, 4):
            if T + 7 > 64 // cpw:
                continue
            ctx.eval_device(lid, d_conf, d_e, d_g, n, chains_per_wave=cpw)
            ms = [ctx.eval_device(lid, d_conf, d_e, d_g, n, chains_per_wave=cpw) for _ in range(3)]
            rates[cpw] = float(np.median(ms))
        best_cpw = min(rates, key=rates.get)
        e_host = ctx.to_host(d_e, (n,))
        for ptr in (d_conf, d_e, d_g):
            ctx.free(ptr)
        out["adjoint"] = {"conformations": n, "atoms": inf.n_atoms, "torsions": T, "pairs": inf.n_pairs,
                          "chains_per_wave": best_cpw, "kernel_ms": round(rates[best_cpw], 3),
                          "evals_per_s": round(n / (rates[best_cpw] / 1000.0), 1),
                          "ms_by_chains_per_wave": {str(k): round(v, 3) for k, v in rates.items()},
                          "finite_fraction": float(np.isfinite(e_host).mean())}

    if not args.no_extras:
        # the QuickVina-W policy on the same ligands and schedule (global + per-chain buffers)
        log("QuickVina-W policy")
        pq = vdock.default_params(chains=args.chains, seed=7, policy=vdock.VD_MC_QVINAW)
        ctx.dock_batch(ids, pq)
        t0 = time.perf_counter()
        st_q = ctx.dock_batch(ids, pq)
        t_q = time.perf_counter() - t0
        best_q = ctx.batch_best()
        out["quickvina_w"] = {"ligands_per_s": round(len(ids) / t_q, 3), "kernel_ms": round(st_q.kernel_ms, 2),
                              "evals": int(st_q.evals), "evals_vs_vina_policy": round(st_q.evals / (evals / args.steps), 4),
                              "mean_best_score": round(float(np.mean(best_q)), 4)}

    if not args.no_extras and not args.no_large_box:
        out["large_box"] = large_box_leg(vdock, log)

    if args.copy_calibration:
        n = 256 << 20
        src, dst = ctx.alloc(n), ctx.alloc(n)
        for _ in range(2):
            ctx.copy_device(dst, src, n)
        ctx.free(src); ctx.free(dst)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = run_cpu_baseline(vdock, texts, box, xyz, xs, args.chains)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
