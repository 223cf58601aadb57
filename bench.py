"""Benchmark: particle-substeps/s of the explicit MPM substep on one MI355X.

Workload: Drucker-Prager sand column collapse, 1,048,576 particles on a 128^3 grid, dt = 5e-5,
quadratic B-spline APIC, stress at F^n, box walls + sphere with Coulomb friction, exact return
mapping. Inputs are synthetic (seeded jittered lattice), resident in HBM before timing starts.

`value` is the exact-return-map path; the `surrogate` object reports the same workload with the
learned return map (trained on the GPU from pairs the exact path logs during the run), its
throughput and its position error against the exact run after the same substeps. The `batched`
object reports the batched many-scene workload (64 independent snow/sand scenes of 20k-500k
particles on 64^3-128^3 grids, one context, one launch per kernel per substep, autotuned launch
shapes) against the same scenes run one after another with the plain exact path, with and without
the learned return map. The `config3` object reports the cloth-sand coupling's throughput. The
`config4` object reports the learned return map's own workload (2^22 sand
particles with per-particle friction angles on a 256^3 grid): the exact path's throughput and the
surrogate against it after 1020 substeps.

Prints one JSON line. `roofline` is for the dominant kernel (its own algorithmic bytes over its
HIP-event-timed duration); `substep_roofline` is the whole substep against 208 B/particle +
56 B/active node. `cpu_baseline` is the CPU oracle (a single-threaded FP64 restatement of the
reference substep) timed on this host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-substeps/sec, surrogate vs exact return map at ≤1e-3 rel pos err"

# Algorithmic bytes of each stage (see the kernel descriptions in mpm_kernels.hip):
#   P2G  reads x, v, B, F, m, V (26 floats) per particle, writes m and momentum (16 B) per active node
#   GRID reads m, momentum (16 B), writes v and v_BF (24 B) per active node
#   G2P  reads x, F, plastic (13 floats), writes x, v, B, F, plastic (25 floats) per particle,
#        reads v and v_BF (24 B) per active node
#   SORT (counting sort) reads the key twice and the rank once, writes rank, scattered id, sorted key,
#        id and slot index: 36 B per particle, plus the cell counts scanned into the bins
STAGES = ["sort", "p2g", "grid", "g2p"]
ALG_PER_PARTICLE = {"sort": 36, "p2g": 104, "grid": 0, "g2p": 152}
ALG_PER_NODE = {"sort": 0, "p2g": 16, "grid": 40, "g2p": 24}
SUBSTEP_PER_PARTICLE, SUBSTEP_PER_NODE = 208, 56
KERNEL_NAMES = {"p2g": "k_p2g<2, P2G_MOM_STRESS>", "g2p": "k_g2p<2>", "grid": "k_grid<GRID_FUSED>",
                "sort": "k_cell_count + k_cell_scan + k_cell_scatter + k_cell_rank + k_claim x2"}


def load_traffic(kernel):
    """HBM traffic per launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_pmc_config2.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_baseline(sample_steps=3):
    """Time the FP64 oracle on this host: config 2 at full size, 1 warm-up + `sample_steps` substeps."""
    from anisotropicelastoplasticity_amd import scenes
    from oracle.bind import Oracle
    sc = scenes.config2_sand_column()
    o = Oracle(sc.params, sc.gmin, sc.gmax, sc.res, prims=sc.prims)
    o.load_scene(sc)
    o.substep(sc.dt)
    t0 = time.perf_counter()
    o.run(sc.dt, sample_steps)
    t = time.perf_counter() - t0
    return dict(value=sc.n * sample_steps / t, unit="particle-substeps/s", cores=1, kind="port",
                sample=f"config2 full size ({sc.n} particles, 128^3), 1 warm-up + {sample_steps} timed substeps, "
                       "FP64 oracle (C++ restatement of the reference substep, -O3, one thread)")


def _sync_torch():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def surrogate_leg(sc, warmup, steps, x_exact, wall_exact, gate_tol=1e-5, pair_chunks=10, pairs_per_chunk=200000,
                  epochs=8, batch=4096, lr=1e-3):
    """The learned return map against the exact one on the same workload: (trial, projected) pairs are
    logged by the exact path at `pair_chunks` substeps spread over the run, the network is trained on
    the GPU, then the run is repeated from the same initial state with the surrogate (gate `gate_tol`,
    exact fallback) and timed; positions are compared after warmup + steps substeps."""
    from anisotropicelastoplasticity_amd import Solver
    total = warmup + steps
    pairs = []
    with Solver(0) as s:
        sc.load_into(s)
        every = max(1, total // pair_chunks)
        done = 0
        while done < total:
            s.set_pair_logging(pairs_per_chunk)
            s.run(sc.dt, 1)
            pairs.append(s.get_pairs())
            s.set_pair_logging(0)
            k = min(every - 1, total - done - 1)
            if k > 0:
                s.run(sc.dt, k)
            done += 1 + max(k, 0)
        pairs = np.concatenate(pairs) if pairs else np.zeros((0, 8), np.float32)
        t0 = time.perf_counter()
        s.surrogate_init(1)
        losses = s.surrogate_train(pairs[:, :5], pairs[:, 5:8], epochs=epochs, batch=batch, lr=lr, seed=2)
        s.synchronize()
        train_s = time.perf_counter() - t0
        w = s.surrogate_get_weights()
    with Solver(0) as s:
        sc.load_into(s)
        s.surrogate_set_weights(w)
        s.set_surrogate(True, gate_tol)
        s.run(sc.dt, warmup)
        s.synchronize()
        _sync_torch()
        t0 = time.perf_counter()
        s.run(sc.dt, steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        x = s.get_particles()["x"]
        st = s.stats()
    dx = x.astype(np.float64) - x_exact.astype(np.float64)
    h = (sc.gmax[0] - sc.gmin[0]) / sc.res[0]
    cand = st["surrogate_candidates"]
    return {
        "value": sc.n * steps / wall, "ms_per_step": 1e3 * wall / steps, "speedup_vs_exact": wall_exact / wall,
        "rel_pos_err": float(np.linalg.norm(dx) / np.linalg.norm(x_exact.astype(np.float64))),
        "max_dx_over_h": float(np.max(np.abs(dx)) / h), "substeps_compared": total,
        "gate_tol": gate_tol, "candidates_per_substep": cand / total,
        "acceptance": (st["surrogate_accepted"] / cand) if cand else 0.0,
        "pairs_trained": int(len(pairs)), "train_epochs": epochs, "train_loss_first_last": [float(losses[0]), float(losses[-1])],
        "train_s": train_s, "status": st["status"],
    }, w


def config3_leg(steps, warmup):
    """Cloth-sand coupling: about 500k sand particles and a 256 x 256-vertex cloth (130,050 triangles)
    on 128^3, dt = 2e-5: throughput of the exact path over all moving points (particles, cloth
    vertices and elements)."""
    from anisotropicelastoplasticity_amd import Solver, scenes
    sc, mesh = scenes.config3_cloth_sand()
    kw = {k: mesh[k] for k in ("youngs_modulus", "poisson_ratio", "shear_stiffness", "stiffness", "friction_angle",
                               "density", "thickness")}
    with Solver(0) as s:
        sc.load_into(s)
        s.set_mesh(mesh["x"], mesh["faces"], **kw)
        s.run(sc.dt, warmup)
        s.synchronize()
        _sync_torch()
        t0 = time.perf_counter()
        s.run(sc.dt, steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        st = s.stats()
    nv, nf = mesh["x"].shape[1], len(mesh["faces"])
    npts = sc.n + nv + nf
    return {"workload": "config3_cloth_sand", "sand_particles": sc.n, "cloth_vertices": nv, "cloth_elements": nf,
            "grid": list(sc.res), "dt": sc.dt, "substeps": warmup + steps, "value": npts * steps / wall,
            "unit": "point-substeps/s (particles + cloth vertices + elements)", "ms_per_step": 1e3 * wall / steps,
            "status": st["status"]}


def config4_leg(surrogate, steps, warmup):
    """The surrogate's workload (2^22 sand particles with per-particle friction angles on 256^3): the
    exact path's throughput, and the learned return map against it after warmup + steps substeps."""
    from anisotropicelastoplasticity_amd import Solver, scenes
    sc = scenes.config4_sand_phi()
    with Solver(0) as s:
        sc.load_into(s)
        s.run(sc.dt, warmup)
        s.synchronize()
        _sync_torch()
        t0 = time.perf_counter()
        s.run(sc.dt, steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        x = s.get_particles()["x"]
        st = s.stats()
    out = {"workload": "config4_sand_phi", "particles": sc.n, "grid": list(sc.res), "dt": sc.dt,
           "substeps": warmup + steps, "value": sc.n * steps / wall, "ms_per_step": 1e3 * wall / steps,
           "fp64_apex_count": st["fp64_apex_count"], "status": st["status"]}
    if surrogate:
        out["surrogate"], _ = surrogate_leg(sc, warmup, steps, x, wall)
    return out


def batched_leg(surrogate, steps, warmup, seq_steps, gate_tol=1e-5, pair_chunks=10, pairs_per_chunk=400000,
                epochs=40, batch_size=4096, lr=1e-3):
    """64 scenes in one context against the same scenes one context at a time (exact path), and the
    batch with the learned return map (trained on the GPU from pairs the exact batch logs, exact
    fallback) against the exact batch after the same substeps."""
    from anisotropicelastoplasticity_amd import Solver, scenes
    batch = scenes.config5_batch()
    dt = batch[0].dt
    ntot = sum(sc.n for sc in batch)
    weights = None
    if surrogate:
        total = warmup + steps
        pairs = []
        with Solver(0) as s:
            s.load_batch(batch)
            every = max(1, total // pair_chunks)
            done = 0
            while done < total:
                s.set_pair_logging(pairs_per_chunk)
                s.run(dt, 1)
                pairs.append(s.get_pairs())
                s.set_pair_logging(0)
                k = min(every - 1, total - done - 1)
                if k > 0:
                    s.run(dt, k)
                done += 1 + max(k, 0)
            pairs = np.concatenate(pairs)
            s.surrogate_init(1)
            losses = s.surrogate_train(pairs[:, :5], pairs[:, 5:8], epochs=epochs, batch=batch_size, lr=lr, seed=2)
            weights = s.surrogate_get_weights()

    def run_batch(sur):
        with Solver(0) as s:
            s.load_batch(batch)
            tuning = s.autotune(dt, reps=3)
            if sur:
                s.surrogate_set_weights(weights)
                s.set_surrogate(True, gate_tol)
            s.run(dt, warmup)
            s.synchronize()
            t0 = time.perf_counter()
            s.run(dt, steps)
            s.synchronize()
            wall = time.perf_counter() - t0
            x = s.get_particles()["x"]
            st = s.stats()
        return wall, x, st, tuning

    wall_b, x_b, st_b, tuning = run_batch(False)
    seq_wall = 0.0
    for sc in batch:  # the plain path: one context per scene, default launch shapes
        with Solver(0) as s:
            sc.load_into(s)
            s.run(dt, warmup)
            s.synchronize()
            t0 = time.perf_counter()
            s.run(dt, seq_steps)
            s.synchronize()
            seq_wall += time.perf_counter() - t0
    seq_value = ntot * seq_steps / seq_wall
    out = {"workload": "config5 batched scenes", "scenes": len(batch), "particles": ntot, "dt": dt,
           "substeps": warmup + steps, "value": ntot * steps / wall_b, "ms_per_step": 1e3 * wall_b / steps,
           "sequential_value": seq_value, "sequential_substeps_per_scene": seq_steps,
           "speedup_vs_sequential": (ntot * steps / wall_b) / seq_value,
           "tuning": {k: tuning[k] for k in ("sort_mode", "p2g_ranks", "p2g_blocks", "grid_blocks")},
           "status": st_b["status"]}
    if weights is not None:
        wall_s, x_s, st_s, _ = run_batch(True)
        dx = x_s.astype(np.float64) - x_b.astype(np.float64)
        cand = st_s["surrogate_candidates"]
        out["surrogate"] = {"value": ntot * steps / wall_s, "ms_per_step": 1e3 * wall_s / steps,
                            "speedup_vs_sequential_exact": (ntot * steps / wall_s) / seq_value,
                            "speedup_vs_batched_exact": wall_b / wall_s,
                            "rel_pos_err": float(np.linalg.norm(dx) / np.linalg.norm(x_b.astype(np.float64))),
                            "acceptance": (st_s["surrogate_accepted"] / cand) if cand else 0.0,
                            "candidates": int(cand), "pairs_trained": int(len(pairs)), "train_epochs": epochs,
                            "train_loss_first_last": [float(losses[0]), float(losses[-1])],
                            "gate_tol": gate_tol, "status": st_s["status"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--stage-steps", type=int, default=50, help="substeps timed stage by stage with HIP events")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned return-map leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched many-scene leg")
    ap.add_argument("--no-config3", action="store_true", help="skip the cloth-sand leg")
    ap.add_argument("--config3-steps", type=int, default=200, help="timed substeps of the cloth-sand leg (after 20)")
    ap.add_argument("--no-config4", action="store_true", help="skip the 2^22-particle per-particle-angle leg")
    ap.add_argument("--config4-steps", type=int, default=1000, help="timed substeps of the config-4 leg (after 20)")
    ap.add_argument("--batch-steps", type=int, default=480, help="timed substeps of the batched leg (after 20 warm-up)")
    ap.add_argument("--seq-steps", type=int, default=40, help="timed substeps per scene of the one-at-a-time baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (reference point for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--only", choices=["config3", "config4", "batched"],
                    help="run this leg alone (exact path; for profiling) and print its JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)
    if args.only:
        leg = {"config3": lambda: config3_leg(args.config3_steps, 20),
               "config4": lambda: config4_leg(False, args.config4_steps, 20),
               "batched": lambda: batched_leg(False, args.batch_steps, 20, args.seq_steps)}[args.only]()
        print(json.dumps(leg))
        return

    from anisotropicelastoplasticity_amd import Solver, scenes
    sc = scenes.config2_sand_column()
    n = sc.n
    with Solver(0) as s:
        sc.load_into(s)
        s.run(sc.dt, args.warmup)
        s.synchronize()
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.synchronize()
        except Exception:
            pass
        t0 = time.perf_counter()
        s.run(sc.dt, args.steps)
        s.synchronize()
        wall = time.perf_counter() - t0
        x_exact = s.get_particles()["x"]
        # per-stage and per-main-kernel HIP-event timing on the context stream (eager launches)
        total_ms, stage_ms, kernel_ms = s.time_kernels(sc.dt, args.stage_steps)
        st = s.stats()
        copy = None
        if args.copy_calibration:
            nbytes, reps = 256 << 20, 3
            copy = {"bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes, "launches": reps,
                    "kernel": "k_copy", "avg_ms": s.bench_copy(nbytes, reps)}

    sur, weights = (None, None) if args.no_surrogate else surrogate_leg(sc, args.warmup, args.steps, x_exact, wall)
    bat = None if args.no_batch else batched_leg(not args.no_surrogate, args.batch_steps, 20, args.seq_steps)
    c3 = None if args.no_config3 else config3_leg(args.config3_steps, 20)
    c4 = None if args.no_config4 else config4_leg(not args.no_surrogate, args.config4_steps, 20)
    value = n * args.steps / wall
    ms_per_step = 1e3 * wall / args.steps
    n_nodes = st["n_active_nodes"]
    per_stage_ms = {k: stage_ms[i] / args.stage_steps for i, k in enumerate(STAGES)}
    per_kernel_ms = {k: kernel_ms[i] / args.stage_steps for i, k in enumerate(STAGES)}
    dom = max(("p2g", "grid", "g2p"), key=lambda k: per_kernel_ms[k])
    alg = ALG_PER_PARTICLE[dom] * n + ALG_PER_NODE[dom] * n_nodes
    achieved = alg / (per_kernel_ms[dom] * 1e-3) / 1e9
    traffic = load_traffic(dom)
    b_alg = SUBSTEP_PER_PARTICLE * n + SUBSTEP_PER_NODE * n_nodes
    sub_ach = b_alg / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": "config2_sand_column", "particles": n, "grid": list(sc.res), "dt": sc.dt,
                   "kernel": "quadratic B-spline APIC", "return_map": "exact (Drucker-Prager, FP32 + FP64 fallback)",
                   "stress": "at F^n", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": per_kernel_ms[dom]},
        "substep_roofline": {"achieved": sub_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": sub_ach / HBM_PEAK_GBS, "alg_bytes_per_substep": b_alg},
        "stage_ms": per_stage_ms,
        "kernel_ms": per_kernel_ms,
        "active_nodes": n_nodes,
        "status": st["status"],
        "fp64_svd_count": st["fp64_svd_count"],
        "fp64_apex_count": st["fp64_apex_count"],
    }
    if sur is not None:
        # the same workload with the learned return map (exact fallback behind the yield-residual
        # gate), compared with the exact run above after the same number of substeps
        out["surrogate"] = sur
    if bat is not None:
        out["batched"] = bat
    if c3 is not None:
        out["config3"] = c3
    if c4 is not None:
        out["config4"] = c4
    if copy is not None:
        out["copy_calibration"] = copy
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
