"""Benchmark of the docking hot path on one MI355X.

Headline (value): batched persistent docking throughput in ligands/s on the batched screening
workload -- 8,000-atom synthetic receptor (60 A cube, 10 A cavity), 30 A box at 0.375 A (81^3
points, all 17 FP32 maps resident in HBM), 1,024 synthetic ligands (10-60 heavy atoms,
0..min(16, atoms/3) torsions, seeds 0..1023), 256 Monte Carlo chains each with RILC-BFGS local
search.  One step = docking all 1,024 ligands.

Printed beside it on the same JSON line:
  roofline          the dominant kernel (the MC kernel): algorithmic bytes (map corners + pair
                    quads) over the span of the step's concurrent MC launches, against 8 TB/s as
                    the contract asks; beside it the HBM traffic those launches cause (derived from
                    the committed counter profile's bytes per algorithmic byte), its fraction of
                    8 TB/s, and the VALU busy share the counter profile measured -- the kernel is
                    bound by instruction issue and latency, not by bandwidth
  per_ligand_plain  the plain per-ligand launch loop and the batched launch on the same ligands;
                    equal accuracy over several seeds on a subset (best-score differences and the
                    RMSD between the two paths' top poses)
  grid              17-map precompute of the same receptor/box: kernel time, GB/s, and the
                    (grid point, receptor atom) terms per second it evaluates
  adjoint           forward + adjoint evaluations of 2^20 random conformations of a 40-atom,
                    12-torsion ligand (evals/s) and their error against the FP64 oracle
  single_ligand     one 32-atom, 8-torsion ligand with 8,192 chains through the plain launch
  small_scene       the CPU-runnable case: 2,000-atom receptor, 20 A box, one 20-atom 5-torsion
                    ligand, 512 chains -- on the device, and the oracle's Monte Carlo on host cores
  dense_bfgs        the screening step with dense BFGS (the reference's --rilc_bfgs 0)
  post              the step after the search for the docked batch: pose lists, FP64 refinement of
                    every pose against the receptor atoms in one launch, final scores; beside it
                    the oracle's FP64 refinement of a sample of the same poses on host threads
  quickvina_w       the same screening step with the QuickVina-W policy (evaluations saved by its
                    gate, throughput, mean best score)
  large_box         large-box batched screening with the QuickVina2 policy (20,000-atom receptor,
                    188^3 maps, 4,096 ligands of 10-80 heavy atoms, 128 chains)
  cpu_baseline      the FP64 oracle's Monte Carlo (same schedule, same random stream) on the host
                    cores this process may use, over a bounded sample of the same ligands
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
    from vdock import workloads as W
    xyz, xs, box = W.single_ligand_scene(vdock, ctx)
    return xyz, xs, box, vdock.load().vd_grid_last_ms(ctx.h)


def screening_ligands(vdock, n=1024):
    from vdock import workloads as W
    return W.texts(vdock, W.screening_specs(n))


def host_threads():
    """Worker threads for the CPU legs: the cores this process may run on (its affinity mask),
    capped by the cgroup's CPU quota and by OMP_NUM_THREADS when set (the share a batch host
    grants), with the counts behind the choice."""
    info = {"nproc": os.cpu_count()}
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    info["affinity"] = n
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            info["cgroup_quota_cpus"] = round(int(quota) / int(period), 2)
            n = min(n, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        info["OMP_NUM_THREADS"] = int(omp)
        n = min(n, int(omp))
    return max(1, n), info


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
    from vdock import workloads as W
    ctx = vdock.Context(0)
    _, _, box = W.large_box_scene(vdock, ctx)
    g_ms = float(np.median([ctx.grid_populate(box) for _ in range(3)]))
    ids = [ctx.add_ligand(t) for t in W.texts(vdock, W.large_box_specs(n_ligands))]
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
            "launches": [{"kernel": r.kernel, "ligands": r.ligands, "blocks": r.blocks, "start_ms": round(r.start_ms, 2),
                          "kernel_ms": round(r.kernel_ms, 2), "evals": int(r.evals), "algorithmic_bytes": round(r.eval_bytes)}
                         for r in ctx.last_launches()],
            "grid": {"points": npts, "maps_MB": round(gbytes / 1e6, 1), "kernel_ms": round(g_ms, 3),
                     "GB_per_s": round(gbytes / (g_ms / 1000.0) / 1e9, 2)}}


# The reference's own per-kernel timings on its batch example: QuickVina-W-CUDA's
# input_file_example/auto_test/log_1u1b.txt (NVIDIA GeForce RTX 3090, 8,192 chains per ligand --
# the logged ric buffer of 1,867,776 B = 8,192 x 228 B), receptor 1u1b, box of
# auto_test/1u1b_config.txt: kernel_grid over the 30 runs of the log (mean 10.0 ms) and
# kernel_monte per ligand, with the search depth and BFGS steps the log prints for it.
REFERENCE_LOG = {
    "kernel_grid_ms_mean": 10.0,
    "kernel_monte_s": {"drugbank1": 0.306398, "drugbank2": 1.406387, "drugbank3": 0.166095, "drugbank4": 1.146933,
                       "drugbank5": 0.070353, "drugbank6": 0.163469, "drugbank7": 0.264657, "drugbank8": 0.070946,
                       "drugbank9": 0.308865, "drugbank10": 0.364207},
    "depth_bfgs": {"drugbank1": (6, 16), "drugbank2": (19, 23), "drugbank3": (3, 15), "drugbank4": (15, 21),
                   "drugbank5": (1, 12), "drugbank6": (3, 15), "drugbank7": (5, 16), "drugbank8": (1, 12),
                   "drugbank9": (5, 15), "drugbank10": (6, 16)},
}


def reference_log_leg(vdock, log, chains=8192):
    """The inputs of the reference's own timing log (REFERENCE_LOG): the 1u1b receptor and its
    auto_test box (65 x 71 x 61 points), and drugbank1-10 docked with 8,192 chains and the
    heuristic search depth and BFGS steps, each through the plain per-ligand launch (the
    reference's one kernel_monte call per ligand) and all ten in one batched launch.  The
    reference ran QuickVina-W, whose gate as executed never skips a search, so the Vina policy
    (the same schedule and heuristics) is what is timed here.  Other hardware: for scale only."""
    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    ctx = vdock.Context(0)
    try:
        with open(os.path.join(root, "1u1b_protein.pdbqt")) as f:
            ctx.set_receptor_pdbqt(f.read())
        box = vdock.box_from_center([-30.661, -29.540, 17.010], [24.0, 26.25, 22.5], 0.375)
        g_ms = float(np.median([ctx.grid_populate(box) for _ in range(5)]))
        names = ["drugbank%d" % k for k in range(1, 11)]
        ids = []
        for n in names:
            with open(os.path.join(root, "qvw_auto", n + ".pdbqt")) as f:
                ids.append(ctx.add_ligand(f.read()))
        p = vdock.default_params(chains=chains, seed=1)
        per = {}
        for n, lid in zip(names, ids):
            inf = ctx.ligand_info(lid)
            ctx.dock(lid, p)
            log(f"reference log leg: {n}")
            t0 = time.perf_counter()
            _, _, _, st = ctx.dock(lid, p)
            wall = 1000.0 * (time.perf_counter() - t0)
            per[n] = {"atoms": inf.n_atoms, "torsions": inf.n_torsions,
                      "depth_bfgs": list(vdock.heuristics(inf.n_atoms, inf.n_torsions)),
                      "reference_depth_bfgs": list(REFERENCE_LOG["depth_bfgs"][n]),
                      "kernel_ms": round(st.kernel_ms, 3), "wall_ms": round(wall, 3),
                      "reference_kernel_monte_ms": round(1000.0 * REFERENCE_LOG["kernel_monte_s"][n], 3)}
        ctx.dock_batch(ids, p)
        t0 = time.perf_counter()
        st_b = ctx.dock_batch(ids, p)
        wall_b = 1000.0 * (time.perf_counter() - t0)
    finally:
        ctx.close()
    ref_sum = 1000.0 * sum(REFERENCE_LOG["kernel_monte_s"].values())
    plain_sum = sum(v["kernel_ms"] for v in per.values())
    return {"workload": "the reference's auto_test log inputs: 1u1b receptor, 24 x 26.25 x 22.5 A box at 0.375 A "
                        "(65 x 71 x 61 points), drugbank1-10, %d chains each, heuristic depth and BFGS steps, "
                        "Vina policy" % chains,
            "reference": "QuickVina-W-CUDA input_file_example/auto_test/log_1u1b.txt, NVIDIA GeForce RTX 3090 "
                         "(other hardware: for scale only)",
            "grid_ms": round(g_ms, 4), "reference_kernel_grid_ms": REFERENCE_LOG["kernel_grid_ms_mean"],
            "per_ligand": per,
            "plain_kernel_ms_sum": round(plain_sum, 3), "reference_kernel_monte_ms_sum": round(ref_sum, 3),
            "batched_kernel_ms": round(st_b.kernel_ms, 3), "batched_wall_ms": round(wall_b, 3),
            "plain_vs_reference": round(ref_sum / plain_sum, 2),
            "batched_vs_reference": round(ref_sum / st_b.kernel_ms, 2)}


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


def copy_calibration(ctx, n=256 << 20):
    """Two device-to-device copies of n bytes: the reference dispatch that scales the memory
    counters of a profile (bytes counted for a copy of known size)."""
    src, dst = ctx.alloc(n), ctx.alloc(n)
    for _ in range(2):
        ctx.copy_device(dst, src, n)
    ctx.free(src)
    ctx.free(dst)


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


def profile_figures(kernel):
    """What the committed counter profile of this workload measured for `kernel`, run alone: VALU
    busy share, resident waves per CU, TA and TD busy shares (the vector-memory path) and the L1
    and L2 hit rates (None when it has no figure for it)."""
    path = os.path.join(ROOT, "profiles", "vd_mc_kernel_batched_screening.md")
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        for tag in ("<!-- valu:", "<!-- vmem:"):
            if line.startswith(tag):
                fields = dict(f.split("=", 1) for f in line.strip("<!-> \n").split()[1:] if "=" in f)
                if fields.pop("kernel", None) == kernel.replace(" ", ""):
                    out.update({k: float(v) for k, v in fields.items()})
    return out or None


def grid_terms(xyz, xs, box):
    """(grid point, typed receptor atom) pairs within the 8 A cutoff: each is one table row of the
    17 map values, the unit of work of the precompute."""
    from scipy.spatial import cKDTree
    b, e, n = box.as_tuple()
    axes = [np.linspace(b[k], e[k], n[k] + 1) for k in range(3)]
    tree = cKDTree(np.asarray(xyz, dtype=np.float64)[np.asarray(xs) < 17])
    X, Y = np.meshgrid(axes[0], axes[1], indexing="ij")
    total = 0
    for z in axes[2]:
        pts = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], 1)
        total += int(tree.query_ball_point(pts, 8.0, return_length=True).sum())
    return total


def post_leg(ctx, log, max_poses=20):
    """The step after the search for every ligand of the last docking batch: each ligand's chains
    reduced to at most max_poses poses (add_to_output_container), every pose refined against the
    receptor atoms in one FP64 launch, final scores and mode selection; timed on its second call
    (the first, reported as first_call_wall_ms, also builds the per-box set-up)."""
    log("post: refinement + final scores of the docked batch")
    t0 = time.perf_counter()
    ctx.postprocess_docked(max_poses)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    st = ctx.postprocess_docked(max_poses)
    wall = time.perf_counter() - t0
    n_lig = len(ctx._post_ids)
    results = [ctx.post_results(k) for k in range(n_lig)]
    best = [m[0].e for m, _, _ in results if m]
    out = {"ligands": n_lig, "poses": st.blocks, "kernel_ms": round(st.kernel_ms, 2), "wall_ms": round(1000 * wall, 2),
           "first_call_wall_ms": round(1000 * first, 2),
           "poses_per_s_kernel": round(st.blocks / (st.kernel_ms / 1000.0), 1),
           "ligands_per_s_wall": round(n_lig / wall, 2), "evals": int(st.evals),
           "evals_per_s_kernel": round(st.evals / (st.kernel_ms / 1000.0), 1),
           "mean_best_final_score": round(float(np.mean(best)), 4), "dtype": "f64"}
    return out, [conf for _, conf, _ in results[:64]]


def plain_accuracy(vdock, ctx, log, ids, chains, seeds=(7, 8, 9, 10), every=8):
    """Equal accuracy of the batched launch and the plain per-ligand launches over several seeds,
    on every `every`-th ligand: per ligand and seed, the two paths' best scores and the RMSD
    (heavy atoms) between their top poses.  Both run the same chains; they differ only in the
    lanes per chain the plain launch takes (a launch of one ligand's chains fills the GPU with one
    chain per wavefront), so FP32 rounding, and with it the trajectories, part after a while."""
    sub = ids[::every]
    d_best, rms = [], []
    for seed in seeds:
        p = vdock.default_params(chains=chains, seed=seed)
        log(f"equal accuracy, seed {seed}: batched and plain over {len(sub)} ligands")
        runs = []
        for plain in (False, True):
            ctx.dock_batch(sub, p, plain=plain)
            runs.append([ctx.batch_results(k) for k in range(len(sub))])
        for (eb, _, hb), (ep, _, hp) in zip(*runs):
            ib, ip = int(np.argmin(eb)), int(np.argmin(ep))
            d_best.append(float(eb[ib] - ep[ip]))
            rms.append(float(np.sqrt(((hb[ib] - hp[ip]) ** 2).sum(-1).mean())))
    d_best, rms = np.array(d_best), np.array(rms)
    return {"ligands": len(sub), "seeds": list(seeds),
            "mean_best_score_diff_batched_minus_plain": round(float(d_best.mean()), 4),
            "mean_abs_best_score_diff": round(float(np.abs(d_best).mean()), 4),
            "top_pose_rmsd_median_A": round(float(np.median(rms)), 3),
            "top_pose_within_2A_fraction": round(float((rms <= 2.0).mean()), 4)}


def graph_leg(vdock, ctx, log, ids, chains, calls=3):
    """Batched docking as a captured hipGraph against the eager calls: the whole screening batch
    and its first 11 ligands (a small batch, where host work weighs most).  Call 0 of each mode is
    the eager warm call or the graph's capture; the next `calls` use new seeds (the graph's kernel
    nodes updated), so every call docks anew.  The results of a replay equal the eager path's bit
    for bit (tests/test_gpu_graph.py)."""
    out = {}
    for name, sub in (("screening", ids), ("first_11_ligands", ids[:11])):
        r = {"ligands": len(sub)}
        for mode in ("eager", "graph"):
            ctx.set_graphs(mode == "graph")
            walls, kms = [], []
            for k in range(calls + 1):
                log(f"graph leg: {name} {mode} call {k}")
                p = vdock.default_params(chains=chains, seed=100 + k)
                t0 = time.perf_counter()
                st = ctx.dock_batch(sub, p)
                walls.append(1000.0 * (time.perf_counter() - t0))
                kms.append(st.kernel_ms)
            r[mode] = {"first_call_ms": round(walls[0], 3), "ms_per_call": round(float(np.median(walls[1:])), 3),
                       "kernel_ms": round(float(np.median(kms[1:])), 3),
                       "host_ms": round(float(np.median(np.array(walls[1:]) - np.array(kms[1:]))), 3)}
        ctx.set_graphs(False)
        r["ligands_per_s_eager"] = round(len(sub) / (r["eager"]["ms_per_call"] / 1000.0), 2)
        r["ligands_per_s_graph"] = round(len(sub) / (r["graph"]["ms_per_call"] / 1000.0), 2)
        out[name] = r
    return out


def adjoint_leg(vdock, ctx, n=1 << 20):
    """The adjoint microbenchmark on the resident maps: 2^20 random conformations (translation
    uniform in the box, uniform orientation, torsions uniform) of a 40-atom, 12-torsion ligand,
    generated and kept in HBM; energy and adjoint gradient of each (vd_eval on device pointers),
    timed at 1, 2 and 4 chains per wavefront.  Algorithmic bytes per evaluation: the eight map
    corners of each heavy atom (32 B), one pair-table quad per pair (16 B), the conformation in
    (4 (7+T)), the energy and the gradient out (4 + 4 (6+T)).  Returns the record and a sample
    for the oracle check."""
    atext = vdock.synth_ligand(40, 12, 777)
    lid = ctx.add_ligand(atext)
    inf = ctx.ligand_info(lid)
    T = inf.n_torsions
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
    ctx.eval_device(lid, d_conf, d_e, d_g, n, chains_per_wave=best_cpw)
    e_host = ctx.to_host(d_e, (n,))
    sample = np.random.default_rng(5).choice(n, 256, replace=False)
    conf_host = ctx.to_host(d_conf, (n, 7 + T))[sample]
    g_host = ctx.to_host(d_g, (n, 6 + T))[sample]
    for ptr in (d_conf, d_e, d_g):
        ctx.free(ptr)
    per_eval = 32 * inf.n_heavy + 16 * inf.n_pairs + 4 * (7 + T) + 4 + 4 * (6 + T)
    ms = rates[best_cpw]
    rec = {"conformations": n, "atoms": inf.n_atoms, "heavy_atoms": inf.n_heavy, "torsions": T, "pairs": inf.n_pairs,
           "chains_per_wave": best_cpw, "kernel_ms": round(ms, 3), "evals_per_s": round(n / (ms / 1000.0), 1),
           "ms_by_chains_per_wave": {str(k): round(v, 3) for k, v in rates.items()},
           "algorithmic_bytes_per_eval": per_eval,
           "algorithmic_GB_per_s": round(per_eval * n / (ms / 1000.0) / 1e9, 1),
           "finite_fraction": float(np.isfinite(e_host).mean())}
    return rec, dict(text=atext, maps=ctx.grid_download(), conf=conf_host, e=e_host[sample], g=g_host)


def single_ligand_leg(vdock, ctx, log):
    """One 32-heavy-atom, 8-torsion ligand with 8,192 chains through the plain per-ligand launch
    (the drop-in for kernel_monte) on the screening scene's maps."""
    log("single ligand: 8192 chains")
    lid = ctx.add_ligand(vdock.synth_ligand(32, 8, 1))
    p = vdock.default_params(chains=8192, seed=11)
    ctx.dock(lid, p)
    t0 = time.perf_counter()
    e, _, _, st = ctx.dock(lid, p)
    wall = time.perf_counter() - t0
    inf = ctx.ligand_info(lid)
    return {"workload": "one ligand of 32 heavy atoms and 8 torsions, 8192 chains, plain launch, RILC-BFGS",
            "atoms": inf.n_atoms, "torsions": inf.n_torsions, "chains": 8192, "kernel_ms": round(st.kernel_ms, 2),
            "wall_ms": round(1000 * wall, 2), "ligands_per_s": round(1.0 / wall, 3),
            "evals_per_s": round(st.evals / (st.kernel_ms / 1000.0), 1), "best_score": round(float(e.min()), 4),
            "chains_per_wave": st.chains_per_wave}


def small_scene_leg(vdock, log):
    """The CPU-runnable case on the device: 2,000-atom receptor (36 A cube, 8 A cavity), 20 A box
    (55^3 points), one 20-heavy-atom ligand with 5 torsions, 512 chains, the heuristics' depth and
    steps, seed 42; maps and docking timed separately."""
    from vdock import workloads as W
    log("small scene: device")
    ctx = vdock.Context(0)
    try:
        xyz, xs, box = W.small_scene(vdock, ctx)
        g_ms = float(np.median([ctx.grid_populate(box) for _ in range(3)]))
        text = vdock.synth_ligand(20, 5, 1)
        lid = ctx.add_ligand(text)
        p = vdock.default_params(chains=512, seed=42)
        ctx.dock(lid, p)
        t0 = time.perf_counter()
        e, _, _, st = ctx.dock(lid, p)
        wall = time.perf_counter() - t0
        out = {"workload": "2000-atom receptor (36 A cube, 8 A cavity), 20 A box (55^3 points), one ligand of 20 "
                           "heavy atoms and 5 torsions, 512 chains, RILC-BFGS, seed 42",
               "grid_kernel_ms": round(g_ms, 4), "dock_kernel_ms": round(st.kernel_ms, 3),
               "dock_wall_ms": round(1000 * wall, 3), "best_score": round(float(e.min()), 4), "evals": int(st.evals)}
    finally:
        ctx.close()
    return out, dict(xyz=xyz, xs=xs, box=box, text=text, lid=lid)


def oracle_leg(vdock, texts, box, xyz, xs, chains, adjoint_sample, post_confs, small, threads, host_info,
               budget_s=15.0):
    """Everything that runs the FP64 oracle: the CPU baselines (the oracle's Monte Carlo, same
    schedule and random stream, on the screening ligands and on the CPU-runnable case; its
    refinement of the docked poses) and, as the checker, the adjoint workload's error against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from concurrent.futures import ThreadPoolExecutor
    out = {}
    lut = orc.Lut()
    b, e, n = box.as_tuple()
    t0 = time.time()
    grid = orc.Grid(lut, b, e, n)
    grid.set_receptor(xyz.astype(np.float64), xs)
    grid.populate(True)
    t_grid = time.time() - t0
    # screening ligands: the headline's CPU baseline
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
    out["cpu_baseline"] = dict(
        value=done / t_mc, unit="ligands/s", cores=threads, kind="port", cpu_model=cpu_model(), host=host_info,
        sample=f"first {done} of the screening ligands, {chains} chains each, FP64 oracle MC (same schedule and "
               f"heuristics) on {threads} host threads (the cores this process may use); maps built once in {t_grid:.1f} s",
        evals_per_s=evals / t_mc)
    # the CPU-runnable case, as specified: 512 chains, the heuristics' depth and steps, seed 42
    sb, se, sn = small["box"].as_tuple()
    t0 = time.time()
    sg = orc.Grid(lut, sb, se, sn)
    sg.set_receptor(small["xyz"].astype(np.float64), small["xs"])
    sg.populate(True)
    t_sgrid = time.time() - t0
    lig = orc.Ligand(small["text"])
    d, bf = vdock.heuristics(lig.n_atoms, lig.T, True)
    t0 = time.time()
    eo, _, _, ev = orc.mc_run(sg, lig, orc.mc_params(512, d, bf, sb, se, seed=42, stream=small["lid"], threads=threads))
    t_small = time.time() - t0
    # the CPU code's own defaults (dense BFGS, monte_carlo.h:71 and ssd.h:56: 2,500 MC steps, 300 BFGS
    # steps) on a sample: one chain per thread for a tenth of the steps, time scaled to 512 x 2,500
    k = threads
    t0 = time.time()
    orc.mc_run(sg, lig, orc.mc_params(k, 250, 300, sb, se, seed=42, stream=small["lid"], threads=threads, mode=0))
    t_def = (time.time() - t0) * (512 / k) * 10
    out["small_scene_cpu"] = {"grid_s": round(t_sgrid, 3), "matched_s": round(t_small, 3),
                              "matched_steps": [d, bf], "matched_best_score": round(float(eo.min()), 4),
                              "matched_evals": int(ev), "defaults_s_for_512_chains": round(t_def, 2),
                              "defaults_sample": f"{k} chains x 250 MC steps, dense BFGS of 300 steps, time scaled "
                                                 f"to 512 chains x 2500 steps",
                              "cores": threads, "kind": "port"}
    # the refinement after the search: the oracle's refine_structure on the docked poses
    jobs = []
    for kk, confs in enumerate(post_confs):
        lg = orc.Ligand(texts[kk])
        for c in confs:
            jobs.append((lg, c.astype(np.float64)))
    full = orc.Grid(lut, b, e, n)
    full.set_receptor(xyz.astype(np.float64), xs)
    done_r, t_cpu = 0, 0.0
    with ThreadPoolExecutor(threads) as pool:
        t1 = time.perf_counter()
        for chunk in range(0, len(jobs), threads):
            list(pool.map(lambda j: orc.refine(full, j[0], j[1], orc.refine_steps(j[0])), jobs[chunk:chunk + threads]))
            done_r = min(len(jobs), chunk + threads)
            if time.perf_counter() - t1 >= 10.0:
                break
        t_cpu = time.perf_counter() - t1
    out["cpu_refine"] = {"poses_per_s": round(done_r / t_cpu, 2), "cores": threads, "kind": "port",
                         "sample": f"{done_r} refined poses of the first ligands, FP64 oracle refine_structure"}
    # the adjoint workload's error: the oracle on the device's maps (the evaluation alone) and on
    # its own FP64 maps (end to end)
    shared = orc.Grid(lut, b, e, n)
    shared.set_maps(adjoint_sample["maps"].astype(np.float64))
    alig = orc.Ligand(adjoint_sample["text"])
    worst = {"shared_maps": [0.0, 0.0], "fp64_maps": [0.0, 0.0]}
    for c, ed, gd in zip(adjoint_sample["conf"], adjoint_sample["e"], adjoint_sample["g"]):
        for key, gr in (("shared_maps", shared), ("fp64_maps", grid)):
            er, gref, _, _ = orc.eval_deriv(gr, alig, c.astype(np.float64))
            worst[key][0] = max(worst[key][0], abs(float(ed) - er) / max(1.0, abs(er)))
            worst[key][1] = max(worst[key][1], float(np.abs(gd - gref).max()) / max(1.0, float(np.abs(gref).max())))
    out["adjoint_error"] = {"sample": len(adjoint_sample["conf"]),
                            "max_rel_energy_error_shared_maps": worst["shared_maps"][0],
                            "max_rel_gradient_error_shared_maps": worst["shared_maps"][1],
                            "max_rel_energy_error_fp64_oracle": worst["fp64_maps"][0],
                            "max_rel_gradient_error_fp64_oracle": worst["fp64_maps"][1],
                            "note": "relative to max(1, |e|) and max(1, max |g|); most conformations lie partly "
                                    "outside the box, where the penalty slope 1e6 dominates"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ligands", type=int, default=1024)
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--autotune", action="store_true",
                    help="time each size bucket's launch shape (vd_set_autotune) instead of the fixed table")
    ap.add_argument("--no-cpu", action="store_true", help="skip every leg that runs the oracle")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-large-box", action="store_true", help="skip the large-box QuickVina2 leg")
    ap.add_argument("--no-post", action="store_true", help="skip the refinement + final-score leg")
    ap.add_argument("--grid-only", action="store_true",
                    help="only the 17-map precompute: screening scene x5, large-box scene x3 (for a kernel profile)")
    ap.add_argument("--large-box-only", action="store_true",
                    help="run only the large-box QuickVina2 batch (for profiling it)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint microbenchmark on the screening scene's maps (for a kernel profile)")
    ap.add_argument("--classes-alone", action="store_true",
                    help="dock each lanes-per-chain class of the screening batch by itself (one launch each, "
                         "with the grid it has alone), print the kernel times and exit: the run that counter "
                         "profiles of one class at a time use")
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
        from vdock import workloads as W
        xyz, xs, box, _ = build_scene(vdock, ctx)
        ms = [ctx.grid_populate(box) for _ in range(5)]
        _, _, b2 = W.large_box_scene(vdock, ctx)
        ms2 = [ctx.grid_populate(b2) for _ in range(3)]
        if args.copy_calibration:
            copy_calibration(ctx)
        print(json.dumps({"grid_only": True, "screening_ms": ms, "large_box_ms": ms2}), flush=True)
        return
    if args.large_box_only:
        rec = large_box_leg(vdock, log)
        if args.copy_calibration:
            copy_calibration(ctx)
        print(json.dumps({"large_box_only": True, "large_box": rec}), flush=True)
        return
    log("scene: receptor + maps")
    xyz, xs, box, _ = build_scene(vdock, ctx)
    if args.adjoint_only:
        rec, _ = adjoint_leg(vdock, ctx)
        print(json.dumps({"adjoint_only": True, "adjoint": rec}), flush=True)
        return
    log("ligands")
    texts = screening_ligands(vdock, args.ligands)
    ids = [ctx.add_ligand(t) for t in texts]
    params = vdock.default_params(chains=args.chains, seed=7)

    if args.classes_alone:
        # the launch-shape table puts ligands under 80 atoms with at most 9 torsions in the 16-lane
        # class and the rest in the 64-lane class; each subset must run as one launch
        infos = [ctx.ligand_info(i) for i in ids]
        narrow = [i for i, f in zip(ids, infos) if f.n_atoms < 80 and f.n_torsions <= 9]
        wide = [i for i, f in zip(ids, infos) if not (f.n_atoms < 80 and f.n_torsions <= 9)]
        res = {}
        for sub in (narrow, wide):
            st = ctx.dock_batch(sub, params)
            L = ctx.last_launches()
            kinds = sorted({r.kernel for r in L})
            assert len(kinds) == 1, kinds
            res[kinds[0]] = {"ligands": len(sub), "kernel_ms": round(st.kernel_ms, 3), "evals": int(st.evals),
                             "blocks": [r.blocks for r in L]}
            log(f"{kinds[0]}: {res[kinds[0]]}")
        print(json.dumps({"classes_alone": res}), flush=True)
        return

    warm_ms = []
    for w in range(args.warmup):
        log(f"warmup {w}")
        t0 = time.perf_counter()
        ctx.dock_batch(ids, params)
        warm_ms.append(round(1000.0 * (time.perf_counter() - t0), 2))
    times, evals, launches, per_step = [], 0, [], []
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
    post, post_confs = (None, []) if args.no_post else post_leg(ctx, log)
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
        "warmup_ms": warm_ms,
        "warmup_note": "the first call of a batch also builds each ligand's device pair layout (an edge "
                       "colouring of its pair graph, on host threads) and the launch buffers",
    }
    name, d, by_kernel = dominant_kernel(launches)
    ms_by_kernel = {k: round(v["ms"] / v["n"], 3) for k, v in by_kernel.items()}
    per_step_bytes = sum(x["bytes"] for x in spans) / len(spans)
    avg_span = sum(x["span"] for x in spans) / len(spans)
    achieved = per_step_bytes / (avg_span / 1000.0) / 1e9
    # HBM traffic of the step: each instantiation's bytes per algorithmic byte (committed counter
    # profile of this workload) times its algorithmic bytes in this run -- derived, not counted here
    tr = [traffic_from_profile(k, v["bytes"] / args.steps) for k, v in by_kernel.items()]
    traffic = round(sum(tr)) if all(t is not None for t in tr) else None
    copy_gbs = copy_bandwidth(ctx)
    valu = {k: profile_figures(k) for k in by_kernel}
    out["roofline"] = {"kernel": "vd_mc_kernel (the step's %d concurrent launches: %s)" % (
                           len(per_step[0]), ", ".join(r.kernel for r in per_step[0])),
                       "bound": "hbm", "achieved": round(achieved, 2),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                       "traffic": traffic,
                       "achieved_is": "algorithmic bytes (8 map corners per heavy atom + 1 pair-table quad per "
                                      "pair, per evaluation) per second of the step's MC span; most are served "
                                      "from L2, so this is not HBM traffic",
                       "traffic_source": "derived: this run's algorithmic bytes x each instantiation's bytes per "
                                         "algorithmic byte from the committed FETCH_SIZE/WRITE_SIZE profile",
                       "hbm_traffic_GBs": round(traffic / (avg_span / 1000.0) / 1e9, 1) if traffic else None,
                       "hbm_traffic_frac": round(traffic / (avg_span / 1000.0) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                       "binding_resource": "the vector-memory data path (TA/TD) that serves the gathers of map "
                                           "corner records and pair-table quads through L1/L2; the 64-lane "
                                           "instantiation is also VALU-bound (no dense contraction, so no MFMA "
                                           "bound)",
                       "counters_by_instantiation_profiled": valu,
                       "measured_copy_GBs": round(copy_gbs, 1),
                       "frac_of_measured_copy": round(achieved / copy_gbs, 5),
                       "steps": len(spans), "avg_span_ms": round(avg_span, 3),
                       "both_running_share_of_span": round(sum(x["overlap"] for x in spans) / len(spans), 4),
                       "algorithmic_bytes_per_step": round(per_step_bytes),
                       "evals_per_step": sum(x["evals"] for x in spans) // len(spans),
                       "avg_launch_ms_by_instantiation": ms_by_kernel, "largest_share_instantiation": name,
                       "algorithmic_bytes_by_instantiation": {k: round(v["bytes"] / v["n"]) for k, v in by_kernel.items()},
                       "evals_by_instantiation": {k: round(v["evals"] / v["n"]) for k, v in by_kernel.items()},
                       "launches_by_instantiation": {k: v["n"] for k, v in by_kernel.items()}}
    if post is not None:
        out["post"] = post

    adjoint_sample = None
    small = None
    if not args.no_extras:
        # per-ligand plain launch loop (one kernel_monte-style launch per ligand) on the same
        # ligands and parameters as the headline step
        log(f"plain loop over {len(ids)} ligands")
        t0 = time.perf_counter()
        st_p = ctx.dock_batch(ids, params, plain=True)
        t_plain = time.perf_counter() - t0
        best_plain = ctx.batch_best()
        out["per_ligand_plain"] = {
            "ligands": len(ids), "plain_ligands_per_s": round(len(ids) / t_plain, 3),
            "batched_ligands_per_s": round(value, 3), "speedup": round(t_plain / (wall / args.steps), 2),
            "mean_best_score_plain": round(float(np.mean(best_plain)), 4),
            "mean_best_score_batched": round(float(np.mean(best)), 4),
            "mean_abs_best_score_diff": round(float(np.mean(np.abs(best_plain - best))), 4),
            "plain_kernel_ms": round(st_p.kernel_ms, 2), "batched_kernel_ms": round(kernel_ms / args.steps, 2),
            "accuracy_over_seeds": plain_accuracy(vdock, ctx, log, ids, args.chains)}
        # grid precompute
        ctx.set_receptor(xyz, xs)
        gms = [ctx.grid_populate(box) for _ in range(5)]
        g_ms = float(np.median(gms))
        npts = np.prod(box.dims)
        gbytes = 17 * npts * 4 + 16 * 8000 + 153 * 2051 * 4
        terms = grid_terms(xyz, xs, box)
        out["grid"] = {"points": int(npts), "kernel_ms": round(g_ms, 4), "algorithmic_bytes": int(gbytes),
                       "GB_per_s": round(gbytes / (g_ms / 1000.0) / 1e9, 2),
                       "frac_hbm_peak": round(gbytes / (g_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 5),
                       "terms": terms, "terms_per_s": round(terms / (g_ms / 1000.0), 1),
                       "terms_note": "(grid point, typed receptor atom) pairs within 8 A; each reads one 17-value "
                                     "table row and adds it to the point's 17 sums"}
        out["adjoint"], adjoint_sample = adjoint_leg(vdock, ctx)
        out["single_ligand"] = single_ligand_leg(vdock, ctx, log)
        out["graph"] = graph_leg(vdock, ctx, log, ids, args.chains)
        # the screening step with dense BFGS (--rilc_bfgs 0) and its own heuristics
        log("dense BFGS")
        pdn = vdock.default_params(chains=args.chains, seed=7, local_search=vdock.VD_LS_BFGS)
        ctx.dock_batch(ids, pdn)
        t0 = time.perf_counter()
        st_d = ctx.dock_batch(ids, pdn)
        t_d = time.perf_counter() - t0
        out["dense_bfgs"] = {"ligands_per_s": round(len(ids) / t_d, 3), "kernel_ms": round(st_d.kernel_ms, 2),
                             "evals": int(st_d.evals), "mean_best_score": round(float(np.mean(ctx.batch_best())), 4),
                             "launches": [r.kernel for r in ctx.last_launches()]}
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
        out["small_scene"], small = small_scene_leg(vdock, log)
        if not args.no_large_box:
            out["large_box"] = large_box_leg(vdock, log)
        out["reference_log"] = reference_log_leg(vdock, log)

    if args.copy_calibration:
        copy_calibration(ctx)
    if not args.no_cpu:
        log("cpu baseline and oracle checks")
        threads, host_info = host_threads()
        if small is None:
            _, small = small_scene_leg(vdock, log)
        if adjoint_sample is None:
            adjoint_sample = dict(text=vdock.synth_ligand(40, 12, 777), maps=ctx.grid_download(),
                                  conf=np.zeros((0, 19)), e=np.zeros(0), g=np.zeros((0, 18)))
        orc_out = oracle_leg(vdock, texts, box, xyz, xs, args.chains, adjoint_sample, post_confs, small, threads, host_info)
        out["cpu_baseline"] = orc_out["cpu_baseline"]
        if "small_scene" in out:
            out["small_scene"]["cpu"] = orc_out["small_scene_cpu"]
            out["small_scene"]["speedup_vs_cpu_matched"] = round(
                orc_out["small_scene_cpu"]["matched_s"] * 1000.0 / out["small_scene"]["dock_wall_ms"], 1)
        if "post" in out:
            out["post"]["cpu_refine"] = orc_out["cpu_refine"]
            out["post"]["speedup_vs_cpu_poses_per_s"] = round(out["post"]["poses_per_s_kernel"] /
                                                              orc_out["cpu_refine"]["poses_per_s"], 1)
        if "adjoint" in out:
            out["adjoint"]["error_vs_oracle"] = orc_out["adjoint_error"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
