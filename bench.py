#!/usr/bin/env python3
"""One-GPU benchmark of the MI355X supervoxel VI-EM.

Default workload (the metric's single-frame configuration): one synthetic
frame of 10,000 Gaussians in a 2048x2048x400 volume (z anisotropy 5, radii
U[3,8], 15% close pairs), ~4.9M voxels in ~109k supervoxels, K = 5, 20 EM
iterations, FP32 E-step exponentials and log-likelihood voxel terms, FP64
moments, M-step statistics, Wishart inverse/log-det/digamma and likelihood
accumulation.

A step is one full EM solve of that frame with its inputs resident in HBM:
the initial Gaussians are restored device-to-device, then the 20 iterations
run as replays of one captured hipGraph. `value` is frames/s over the timed
steps. The JSON line also carries rnk-updates/s (S*K*iterations per solve),
p50/p99 per-frame latency, the accuracy against the FP64 oracle, the
roofline of the dominant kernel (hipEvent timing on the library's stream)
and the oracle timed on the host cores as the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--config 1|2|3|5]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = ("frames/s and rnk updates/s at ≤1e-5 rel err vs FP64; "
          "p99 frame latency")
HBM_PEAK = 8.0e12


def _frames_for(cfg):
    from tgmm_amd import synth
    if cfg == 1:
        f = synth.config1()
        return [f], 5, 10, ("config1: 200 Gaussians, 512x512x100 aZ=5, radii 3-6, "
                            "~40k voxels, ~900 supervoxels, K=5, 10 iterations")
    if cfg == 2:
        f = synth.config2()
        return [f], 5, 20, ("config2: 10,000 Gaussians, 2048x2048x400 aZ=5, radii 3-8, "
                            "15% close pairs, ~4.9M voxels, ~109k supervoxels, K=5, "
                            "20 iterations")
    if cfg == 3:
        fr = synth.config3_frames()
        return fr, 5, 15, ("config3: 64 frames, 100-20,000 Gaussians (log-uniform), "
                           "~25M voxels total, K=5, 15 iterations, one batch")
    if cfg == 5:
        f = synth.config5()
        return [f], 8, 25, ("config5: 60,000 Gaussians, 4096x4096x500 aZ=5, lognormal "
                            "brightness 0.8, ~30M voxels, ~600k supervoxels, K=8, "
                            "25 iterations")
    raise SystemExit("unknown config %r" % cfg)


def _percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-oracle", action="store_true",
                    help="skip the FP64 accuracy check and the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=2,
                    help="EM iterations of the oracle timed for the CPU baseline")
    ap.add_argument("--check-iters", type=int, default=0,
                    help="iterations of the accuracy solve (0 = the workload's count)")
    ap.add_argument("--plain-frames", type=int, default=64,
                    help="config 3: frames timed in the plain per-frame form")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the EM of a frame runs on one "
              "device and this build has no multi-GPU path (requested --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from tgmm_amd import em

    t0 = time.time()
    frames, K, iters, workload = _frames_for(args.config)
    gen_s = time.time() - t0
    print("# generated %d frame(s) in %.1fs: N=%d S=%d G=%d" % (
        len(frames), gen_s, sum(f.N for f in frames), sum(f.S for f in frames),
        sum(f.G for f in frames)), file=sys.stderr, flush=True)
    opts = em.make_opts(K=K, max_iter=iters, tol=-1.0)
    batch = em.Batch(frames, opts)
    S_tot = sum(f.S for f in frames)
    N_tot = sum(f.N for f in frames)

    # accuracy: one solve against the FP64 oracle (all frames when cheap)
    chk_iters = args.check_iters or iters
    acc = None
    if not args.no_oracle:
        import oracle_lib as O
        from helpers import record_errors, rel_err
        if chk_iters != iters:
            with em.Batch(frames, em.make_opts(K=K, max_iter=chk_iters, tol=-1.0)) as cb:
                cb.solve(em.SOLVE_GRAPH)
                outs = cb.fetch()
        else:
            batch.reset()
            batch.solve(em.SOLVE_GRAPH)
            outs = batch.fetch()
        errs = []
        knn_same = True
        idx = list(range(len(frames))) if args.config != 3 else list(range(0, len(frames), 8))
        sample = [frames[i] for i in idx]
        for i, f in zip(idx, sample):
            o = outs[i]
            gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(f, O.make_opts(K=K, max_iter=chk_iters))
            e = record_errors(o["gm"], gm_o)
            e["ll"] = rel_err(o["ll"], ll)
            errs.append(e)
            knn_same &= bool(np.array_equal(o["ind"], ind_o))
        acc = {k: max(e[k] for e in errs) for k in errs[0]}
        acc["knn_sets_identical"] = knn_same
        acc["frames_checked"] = len(sample)
        acc["iterations"] = chk_iters
        print("# accuracy vs FP64 oracle: %s" % json.dumps(acc), file=sys.stderr, flush=True)

    # warmup + timed steps
    for _ in range(args.warmup):
        batch.reset()
        batch.solve(em.SOLVE_GRAPH)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        batch.reset()
        batch.solve(em.SOLVE_GRAPH)
        lat.append(time.perf_counter() - t1)
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1000.0 * elapsed / args.steps
    frames_per_s = len(frames) * args.steps / elapsed
    rnk_updates = S_tot * K * iters * args.steps / elapsed
    vox_updates = N_tot * K * iters * args.steps / elapsed
    print("# timed %d steps: %.3f ms/step" % (args.steps, ms_per_step), file=sys.stderr,
          flush=True)

    # per-kernel device time with hipEvents on the library's stream
    kern = {}
    for name in ("loglik", "knn", "estep", "mstep"):
        ms, by = batch.time_kernel(name, 2 * iters)
        kern[name] = dict(avg_ms=ms, bytes=by, gbps=by / (ms * 1e-3) / 1e9 if ms > 0 else None)
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    achieved = kern[dom]["bytes"] / (kern[dom]["avg_ms"] * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s_config%d.json" % (dom, args.config))
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": traffic, "algorithmic_bytes": kern[dom]["bytes"],
                "avg_ms": kern[dom]["avg_ms"]}
    print("# kernels: %s" % json.dumps(kern), file=sys.stderr, flush=True)
    batch.close()

    # config 3: the plain per-frame GPU form over the same frames (one
    # workspace allocated per frame, host-synchronised iterations, no graph,
    # no batching), the denominator of the batched-over-plain ratio
    plain = None
    if args.config == 3 and args.plain_frames > 0:
        sub = frames[: args.plain_frames]
        t1 = time.perf_counter()
        for f in sub:
            with em.Batch([f], opts) as pb:
                pb.solve(em.SOLVE_HOSTSYNC)
                pb.fetch()
        dt = time.perf_counter() - t1
        vox_sub = sum(f.N for f in sub)
        plain = {"frames": len(sub), "seconds": dt, "frames_per_s": len(sub) / dt,
                 "voxels_per_s": vox_sub / dt,
                 "batched_over_plain_frames_per_s": frames_per_s / (len(sub) / dt)}
        print("# plain per-frame form: %s" % json.dumps(plain), file=sys.stderr, flush=True)

    cpu = None
    if not args.no_oracle:
        import oracle_lib as O
        f = frames[0]
        n_it = max(1, min(args.cpu_iters, iters))
        threads = O.oracle().orc_get_threads()
        t1 = time.perf_counter()
        O.em_run(f, O.make_opts(K=K, max_iter=n_it))
        dt = time.perf_counter() - t1
        per_frame = dt / n_it * iters
        cpu = {"value": (1.0 / per_frame) * (1 if args.config != 3 else 1),
               "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": "%d of %d EM iterations of frame 0 (N=%d, S=%d, G=%d) by the FP64 "
                         "oracle on %d host threads, scaled to one full solve"
                         % (n_it, iters, f.N, f.S, f.G, threads)}

    line = {
        "metric": METRIC,
        "value": frames_per_s,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator, tgmm_amd/csrc/synth.cpp)",
        "config": {"workload": workload, "frames_per_step": len(frames),
                   "voxels": N_tot, "supervoxels": S_tot,
                   "gaussians": sum(f.G for f in frames), "K": K, "iterations": iters,
                   "parallelism": "1 GPU"},
        "rnk_updates_per_s": rnk_updates,
        "voxel_candidate_updates_per_s": vox_updates,
        "latency_ms": {"p50": 1000 * _percentile(lat, 50), "p99": 1000 * _percentile(lat, 99)},
        "accuracy_vs_fp64": acc,
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if plain is not None:
        line["plain_per_frame_form"] = plain
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
