#!/usr/bin/env python3
"""One-GPU benchmark of the MI355X supervoxel VI-EM.

Default workload (the largest single-GPU configuration of the metric): one
synthetic frame of 60,000 Gaussians in a 4096x4096x500 volume (z anisotropy
5, radii U[3,8], log-normal brightness sigma 0.8), ~30M voxels in ~600k
supervoxels, K = 8, 25 EM iterations; FP32 E-step exponentials and
log-likelihood voxel terms, FP64 moments, M-step statistics, Wishart
inverse/log-det/digamma and likelihood accumulation.

A step is one full EM solve of that frame with its inputs resident in HBM:
the initial Gaussians are restored device-to-device, the 25 iterations run as
replays of one captured hipGraph, the last kNN search's slots are put in the
reference's layout, and the records (with the r > 0.5 supervoxel assignment
formed on the device) are copied to the host. `value` is frames/s over the
timed steps. The JSON line also carries rnk-updates/s (S*K*iterations per
solve), p50/p99 step latency, the accuracy of all 25 iterations against the
FP64 oracle, every kernel's HBM fraction with the roofline of the longest
one (hipEvent timing on the library's stream), the oracle timed on all host
cores and on one as the CPU baseline, and two keys measured on their own
workloads: config 3's batched-over-plain ratio (upload and fetch on both
sides) and config 4's served p50/p99 latency.

    python bench.py [--steps K] [--warmup W] [--config 1|2|3|5]

With --marshal it measures the input marshalling on the device (raw uint16
supervoxel voxels to the EM's float inputs and centroids) for the frame of
--config (default 2), from inputs resident in HBM and from pinned host memory,
beside the upload of the finished float arrays that the reference's host path
needs and the oracle's host restatement of the same work.

With --svknn it measures the supervoxel nearest neighbours (knnCUDA_) on the
frame's supervoxel centroids in raster order: the tracker's three searches
per frame (in space, K = 11; forward and backward against the next frame's
centroids, K = 10).

With --local it measures the local log-likelihood entry point
(calculateLocalLikelihood, the cell-division test's input) on the solved
frame of --config (default 2): the tracker's per-Gaussian supervoxel lists,
for 3% of the Gaussians (a division round) and for all of them.

With --config 4 it measures the continuous-batching server instead: four
synthetic time-lapse streams of 200 frames (2,000-15,000 cells, ~200 voxels
per cell, K = 5, 10 iterations), each stream's posterior kept on the device as
its next frame's priors. Two phases on fresh servers: a closed loop that keeps
every stream's next frame queued (throughput, `value`), then Poisson arrivals
at --qps (default half the closed-loop rate) for p50/p99 latency from a
frame's arrival to its results on the host. Input upload (host to device) is
inside both timed regions.
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


def _copy_calibration(nbytes):
    """One device-to-device hipMemcpy of nbytes (reads and writes nbytes of
    HBM): a dispatch of known traffic beside the profiled kernels."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError("hipMemcpy failed: %d" % rc)
    return nbytes


def _percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def _cpu_sample(fn, min_s=2.0, max_reps=1000):
    """Seconds per call of fn, repeated until at least min_s of wall time
    (a stable sample on the host cores); returns (seconds per call, calls)."""
    reps, t1 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        dt = time.perf_counter() - t1
        if dt >= min_s or reps >= max_reps:
            return dt / reps, reps


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _cpu_baseline(frames, K, iters, vox):
    """The FP64 oracle (the same update equations, OpenMP) on the host cores:
    a bounded sample of whole EM iterations of the workload on all threads
    available to the process and on one thread, scaled to full solves."""
    import oracle_lib as O
    L = O.oracle()
    n_all = L.orc_get_threads()
    idx = [0] if len(frames) == 1 else list(range(0, len(frames), 8))
    big = sum(frames[i].N for i in idx) > 2e7
    out = {}
    for threads, n_it, min_s in ((n_all, 2 if not big else 1, 4.0), (1, 1, 0.0)):
        L.orc_set_threads(threads)
        oo = O.make_opts(K=K, max_iter=n_it, voxel=vox)
        dt, reps = _cpu_sample(lambda: [O.em_run(frames[i], oo) for i in idx], min_s=min_s,
                               max_reps=50)
        per_step = dt / n_it * iters * len(frames) / len(idx)
        out[threads] = (len(frames) / per_step, n_it, reps, dt * reps)
    L.orc_set_threads(n_all)
    f = frames[0]
    what = ("frame 0 (N=%d, S=%d, G=%d)" % (f.N, f.S, f.G) if len(idx) == 1 else
            "frames %s of the batch (N=%d in total), scaled to all %d frames" % (
                ",".join(map(str, idx)), sum(frames[i].N for i in idx), len(frames)))
    v_all, it_all, r_all, s_all = out[n_all]
    v_one, it_one, r_one, s_one = out[1]
    return {"value": v_all, "unit": "frames/s", "cores": n_all, "kind": "port",
            "cpu": _cpu_model(),
            "single_thread": {"value": v_one, "unit": "frames/s", "cores": 1},
            "sample": "%d of %d EM iterations of %s by the FP64 oracle on %d host threads "
                      "(%d runs, %.1f s), and %d iteration on one thread (%.1f s), scaled to "
                      "full solves" % (it_all, iters, what, n_all, r_all, s_all, it_one, s_one)}


def _roofline_of(kern):
    """Every kernel's HBM fraction, and the roofline object of the kernel
    that takes the longest per iteration (whatever bounds it)."""
    for k, v in kern.items():
        v["hbm_frac"] = (v["bytes"] / (v["avg_ms"] * 1e-3) / HBM_PEAK) if v["avg_ms"] > 0 else None
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    achieved = kern[dom]["bytes"] / (kern[dom]["avg_ms"] * 1e-3)
    return dom, {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9,
                 "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                 "algorithmic_bytes": kern[dom]["bytes"], "avg_ms": kern[dom]["avg_ms"]}


def _plain_vs_batched(frames, opts, reps=2):
    """Frames/s of the batched form and of the plain per-frame form over the
    same frames, both from host arrays to host results: the batched form
    uploads all frames into one workspace, replays one captured iteration for
    all of them and fetches the records (rnk/ind included); the plain form
    gives each frame its own workspace, runs host-synchronised iterations
    without a graph and fetches the same outputs."""
    from tgmm_amd import em
    best_b = None
    for _ in range(reps):
        t1 = time.perf_counter()
        with em.Batch(frames, opts) as b:
            b.solve(em.SOLVE_GRAPH)
            b.fetch()
        dt = time.perf_counter() - t1
        best_b = dt if best_b is None else min(best_b, dt)
    t1 = time.perf_counter()
    for f in frames:
        with em.Batch([f], opts) as pb:
            pb.solve(em.SOLVE_HOSTSYNC)
            pb.fetch()
    dt_p = time.perf_counter() - t1
    return {"frames": len(frames), "batched_seconds": best_b, "plain_seconds": dt_p,
            "batched_frames_per_s": len(frames) / best_b, "plain_frames_per_s": len(frames) / dt_p,
            "batched_over_plain": dt_p / best_b,
            "both_sides": "upload from host arrays, workspace allocation, solve, fetch of "
                          "records, rnk and ind to host"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=None,
                    help="workload: 1, 2, 3, 5 (default 5), 4 (the server); the --marshal, "
                         "--svknn and --local modes default to 2")
    ap.add_argument("--no-oracle", action="store_true",
                    help="skip the FP64 accuracy check and the CPU baseline")
    ap.add_argument("--check-iters", type=int, default=0,
                    help="iterations of the accuracy solve (0 = the workload's count)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-over-plain (config 3) and served-latency (config 4) "
                         "keys of the default run")
    ap.add_argument("--plain-frames", type=int, default=64,
                    help="config 3: frames timed in the plain per-frame form")
    ap.add_argument("--stream-frames", type=int, default=200,
                    help="config 4: frames per time-lapse stream")
    ap.add_argument("--extra-stream-frames", type=int, default=50,
                    help="frames per stream of the served-latency key of the default run")
    ap.add_argument("--qps", type=float, default=0.0,
                    help="config 4: Poisson arrival rate, all streams (0 = half the "
                         "closed-loop throughput)")
    ap.add_argument("--hbm-budget-gb", type=float, default=24.0,
                    help="config 4: server workspace budget")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (for counter runs)")
    ap.add_argument("--tiles", type=str, default="",
                    help="launch shapes as JSON (the keys of em.DEFAULT_TILES) or as their "
                         "values in that order, comma-separated, instead of the autotuner's "
                         "choice (for profiler runs)")
    ap.add_argument("--voxel", action="store_true",
                    help="per-voxel EM (GMEMvariationalInferenceCUDA) on the same frames")
    ap.add_argument("--marshal", action="store_true",
                    help="measure the input marshalling on the device instead")
    ap.add_argument("--svknn", action="store_true",
                    help="measure the supervoxel nearest-neighbour search instead")
    ap.add_argument("--local", action="store_true",
                    help="measure the local log-likelihood entry point instead")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (hipMemcpy), "
                         "the known-size reference for HBM traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the EM of a frame runs on one "
              "device and this build has no multi-GPU path (requested --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from tgmm_amd import em
    from tgmm_amd.records import GMM_DTYPE

    if args.config is None:
        args.config = 2 if (args.local or args.marshal or args.svknn) else 5
    if args.config == 4:
        return serve_bench(args)
    if args.local:
        return local_bench(args)
    if args.marshal:
        return marshal_bench(args)
    if args.svknn:
        return svknn_bench(args)

    t0 = time.time()
    frames, K, iters, workload = _frames_for(args.config)
    gen_s = time.time() - t0
    print("# generated %d frame(s) in %.1fs: N=%d S=%d G=%d" % (
        len(frames), gen_s, sum(f.N for f in frames), sum(f.S for f in frames),
        sum(f.G for f in frames)), file=sys.stderr, flush=True)
    vox = bool(args.voxel)
    if vox:
        workload += "; voxel mode (every voxel its own kNN query)"
    opts = em.make_opts(K=K, max_iter=iters, tol=-1.0, voxel=vox)
    batch = em.Batch(frames, opts)
    S_tot = sum((f.N if vox else f.S) for f in frames)
    N_tot = sum(f.N for f in frames)

    # launch shapes: the autotuner times the candidates on this batch (not
    # part of the timed steps; results do not depend on the choice)
    if args.tiles:
        if args.tiles.lstrip().startswith("{"):
            want = json.loads(args.tiles)
        else:  # values in the order of em.DEFAULT_TILES' keys
            want = dict(zip(em.DEFAULT_TILES, (int(v) for v in args.tiles.split(","))))
        batch.set_tiles(want)
        tiles = batch.get_tiles()
    elif args.no_autotune:
        tiles = batch.get_tiles()
    else:
        tiles, tune_report = batch.autotune(reps=3, cache=False)
        print("# autotuned tiles %s\n%s" % (json.dumps(tiles), tune_report or ""),
              file=sys.stderr, flush=True)

    # a step: the initial Gaussians restored on the device, the solve (graph
    # replays, the last search's slot layout), and the results on the host:
    # the records with the r > 0.5 supervoxel assignment formed on the device
    # and copied back into page-locked host arrays kept across frames (rnk
    # and ind stay in HBM, as the reference leaves them)
    recs = [em.pinned(np.zeros(f.G, GMM_DTYPE)) for f in frames]

    def step():
        batch.reset()
        batch.solve(em.SOLVE_GRAPH)
        return batch.fetch(records_only=True, into=recs)
    for _ in range(args.warmup):
        step()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        step()
        lat.append(time.perf_counter() - t1)
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1000.0 * elapsed / args.steps
    frames_per_s = len(frames) * args.steps / elapsed
    rnk_updates = S_tot * K * iters * args.steps / elapsed
    vox_updates = N_tot * K * iters * args.steps / elapsed
    bstats = batch.stats()
    print("# timed %d steps: %.3f ms/step; %s" % (args.steps, ms_per_step, json.dumps(bstats)),
          file=sys.stderr, flush=True)

    # per-kernel device time with hipEvents on the library's stream
    kern = {}
    for name in ("loglik", "knn", "estep", "mstep"):
        ms, by = batch.time_kernel(name, 2 * iters)
        kern[name] = dict(avg_ms=ms, bytes=by, gbps=by / (ms * 1e-3) / 1e9 if ms > 0 else None)
    dom, roofline = _roofline_of(kern)
    tpath = os.path.join(ROOT, "profiles", "traffic_%s_config%d%s.json" % (
        dom, args.config, "_voxel" if vox else ""))
    roofline["traffic"] = None
    if os.path.exists(tpath):
        with open(tpath) as fh:
            roofline["traffic"] = json.load(fh).get("hbm_bytes_per_launch")
    print("# kernels: %s" % json.dumps(kern), file=sys.stderr, flush=True)

    # accuracy: the solve's outputs (all iterations) against the FP64 oracle
    chk_iters = args.check_iters or iters
    acc, cpu = None, None
    if not args.no_oracle:
        import oracle_lib as O
        from helpers import record_errors, rel_err
        if chk_iters != iters:
            with em.Batch(frames, em.make_opts(K=K, max_iter=chk_iters, tol=-1.0,
                                               voxel=vox)) as cb:
                cb.solve(em.SOLVE_GRAPH)
                outs = cb.fetch()
        else:
            batch.reset()
            batch.solve(em.SOLVE_GRAPH)
            outs = batch.fetch()
        st = batch.stats()
        errs = []
        same_slots = True
        idx = list(range(len(frames))) if len(frames) <= 4 else list(range(0, len(frames), 8))
        t1 = time.time()
        for i in idx:
            o = outs[i]
            gm_o, rnk_o, ind_o, hist, it, ll = O.em_run(
                frames[i], O.make_opts(K=K, max_iter=chk_iters, voxel=vox))
            e = record_errors(o["gm"], gm_o)
            e["ll"] = rel_err(o["ll"], ll)
            errs.append(e)
            same_slots &= bool(np.array_equal(o["ind"], ind_o))
        acc = {k: max(e[k] for e in errs) for k in errs[0]}
        acc["knn_slots_identical"] = same_slots
        acc["layout_mismatches"] = st["layout_mismatches"]
        acc["frames_checked"] = len(idx)
        acc["iterations"] = chk_iters
        acc["oracle_seconds"] = time.time() - t1
        print("# accuracy vs FP64 oracle: %s" % json.dumps(acc), file=sys.stderr, flush=True)
        cpu = _cpu_baseline(frames, K, iters, vox)
        print("# cpu baseline: %s" % json.dumps(cpu), file=sys.stderr, flush=True)
    batch.close()

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
                   "step": "reset + graph-replayed solve + slot layout + records with the "
                           "supervoxel assignment copied to page-locked host memory",
                   "parallelism": "1 GPU"},
        "rnk_updates_per_s": rnk_updates,
        "voxel_candidate_updates_per_s": vox_updates,
        "latency_ms": {"p50": 1000 * _percentile(lat, 50), "p99": 1000 * _percentile(lat, 99)},
        "accuracy_vs_fp64": acc,
        "kernels": kern,
        "dominant_kernel": dom,
        "knn_ties_last_iteration": bstats["knn_ties_last_iteration"],
        "tiles": tiles,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if args.copy_calibration_mb > 0:
        line["copy_calibration_bytes"] = _copy_calibration(args.copy_calibration_mb << 20)
    del frames
    if args.config == 5 and not args.no_extras and not vox:
        # the other headline numbers, on their own workloads
        try:
            fr3, K3, it3, wl3 = _frames_for(3)
            o3 = em.make_opts(K=K3, max_iter=it3, tol=-1.0)
            line["config3_batched_over_plain"] = dict(
                workload=wl3, **_plain_vs_batched(fr3[: args.plain_frames], o3))
            del fr3
            print("# config 3: %s" % json.dumps(line["config3_batched_over_plain"]),
                  file=sys.stderr, flush=True)
        except Exception as ex:  # reported, not fatal: the headline line still prints
            line["config3_batched_over_plain"] = {"error": repr(ex)}
        try:
            line["config4_served"] = serve_summary(args, args.extra_stream_frames)
            print("# config 4: %s" % json.dumps(line["config4_served"]), file=sys.stderr,
                  flush=True)
        except Exception as ex:
            line["config4_served"] = {"error": repr(ex)}
    print(json.dumps(line), flush=True)


def marshal_bench(args):
    """tgmm_marshal_frame on one frame: the frame's supervoxels as raw uint16
    voxels (8-byte linear index and 2-byte intensity per voxel, a threshold
    per supervoxel; intensities thr + round(img * 3000)). Timed per call
    through the C entry point (input copies, both kernels, the range
    read-back); outputs compared bitwise with the oracle."""
    import ctypes
    import numpy as np
    from tgmm_amd import em
    import oracle_lib as O
    from helpers import raw_from_frame
    frames, K, iters, workload = _frames_for(args.config)
    f = frames[0]
    dims = {1: (512, 512, 100), 2: (2048, 2048, 400), 5: (4096, 4096, 500)}[args.config]
    idx, raw, ptr, thr, d = raw_from_frame(f, dims, seed=1)
    N, S = f.N, f.S
    L = em.load()
    outs = (em.DeviceArray(nbytes=12 * N), em.DeviceArray(nbytes=4 * N),
            em.DeviceArray(nbytes=12 * S), em.DeviceArray(nbytes=8 * (S + 1)))
    sources = {
        "device": (em.DeviceArray(idx), em.DeviceArray(raw), em.DeviceArray(thr)),
        "pinned_host": (em.pinned(idx), em.pinned(raw), em.pinned(thr)),
    }
    res = {}
    for name, (a_idx, a_raw, a_thr) in sources.items():
        fr, keep = em.raw_frame(a_idx, a_raw, ptr, a_thr, d)

        def call():
            em._check(L.tgmm_marshal_frame(ctypes.byref(fr), outs[0].ptr, outs[1].ptr,
                                           outs[2].ptr, outs[3].ptr))
        for _ in range(max(1, args.warmup)):
            call()
        ts = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            call()
            ts.append(time.perf_counter() - t0)
        res[name] = {"ms_per_frame_p50": 1e3 * _percentile(ts, 50),
                     "ms_per_frame_min": 1e3 * min(ts)}
        print("# %s: %s" % (name, json.dumps(res[name])), file=sys.stderr, flush=True)
    # the reference's path uploads the finished float arrays instead
    qh, imh, ch = O.marshal_frame(idx, raw, ptr, thr, d)
    up = [em.pinned(qh), em.pinned(imh), em.pinned(ch)]
    dst = [em.DeviceArray(nbytes=a.nbytes) for a in up]
    H = em._hip_lib()
    ts = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        for a, b in zip(up, dst):
            H.hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, 1)
        ts.append(time.perf_counter() - t0)
    res["float_arrays_upload"] = {"ms_per_frame_p50": 1e3 * _percentile(ts, 50),
                                  "bytes": int(sum(a.nbytes for a in up))}
    got = [outs[0].numpy(np.float32)[: 3 * N], outs[1].numpy(np.float32)[:N],
           outs[2].numpy(np.float32)[: 3 * S]]
    exact = all(np.array_equal(g, w) for g, w in zip(got, (qh, imh, ch)))
    cpu, cpu_reps = _cpu_sample(lambda: O.marshal_frame(idx, raw, ptr, thr, d))
    t_dev = res["device"]["ms_per_frame_p50"] / 1e3
    line = {"metric": "marshalled voxels/s (raw uint16 frame to EM inputs)",
            "value": N / t_dev, "unit": "voxels/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * t_dev, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u16->f32", "data": "synthetic",
            "config": {"workload": workload + "; raw frame %dx%dx%d, %d voxels, %d supervoxels"
                       % (dims[0], dims[1], dims[2], N, S)},
            "bitwise_equal_to_oracle": bool(exact), "cases": res,
            "input_bytes": {"raw": int(10 * N + 10 * S + 8),
                            "float_arrays": int(16 * N + 12 * S)},
            "cpu_baseline": {"value": N / cpu, "unit": "voxels/s", "cores": 1, "kind": "port",
                             "sample": "the oracle's single-thread restatement of the whole "
                                       "frame, %d calls (at least 2 s)" % cpu_reps}}
    print(json.dumps(line), flush=True)


def svknn_bench(args):
    """tgmm_knnCUDA_ for one frame's three searches: centroids of the frame's
    supervoxels sorted by (z, y, x) (the raster order of a segmentation's
    labels), the next frame's centroids displaced by N(0, 1.5) voxels and
    re-sorted. Timed per frame (the three calls, host in and out); every
    slot of 2,000 sampled queries compared bitwise with the oracle; the
    oracle's time on those queries, scaled to the frame, is the CPU
    baseline."""
    import numpy as np
    from tgmm_amd import em
    import oracle_lib as O
    frames, K, iters, workload = _frames_for(args.config)
    f = frames[0]
    scale = np.array([1.0, 1.0, 5.0], np.float32)

    def raster(c):
        return c[:, np.lexsort((c[0], c[1], c[2]))]
    cur = raster(f.centroids.reshape(3, f.S).astype(np.float64))
    nxt = raster(cur + np.random.default_rng(2).normal(0, 1.5, cur.shape))
    a = np.ascontiguousarray(cur.astype(np.float32).reshape(-1))
    b = np.ascontiguousarray(nxt.astype(np.float32).reshape(-1))
    calls = [("space", a, a, 11), ("forward", a, b, 10), ("backward", b, a, 10)]

    def frame():
        return [em.knnCUDA_(q, r, k, scale) for _, q, r, k in calls]
    for _ in range(max(1, args.warmup)):
        outs = frame()
    ts = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        frame()
        ts.append(time.perf_counter() - t0)
    S = f.S
    rng = np.random.default_rng(3)
    sample = np.sort(rng.choice(S, min(S, 2000), replace=False))
    exact = True
    cpu_s = 0.0
    for (name, q, r, k), (gi, gd) in zip(calls, outs):
        qs = np.ascontiguousarray(q.reshape(3, S)[:, sample].reshape(-1))
        t0 = time.perf_counter()
        oi, od = O.sv_knn(qs, r, k, scale)
        cpu_s += time.perf_counter() - t0
        gi2 = gi.reshape(k, S)[:, sample].reshape(-1)
        gd2 = gd.reshape(k, S)[:, sample].reshape(-1)
        exact &= bool(np.array_equal(gi2, oi) and np.array_equal(gd2, od))
    cpu_frame = cpu_s * S / sample.size
    t = _percentile(ts, 50)
    line = {"metric": "supervoxel kNN frames/s (three searches per frame)", "value": 1.0 / t,
            "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * t, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": workload + "; %d supervoxel centroids, raster order, "
                       "searches in space (K=11), forward and backward (K=10)" % S},
            "bitwise_equal_to_oracle_on_sample": exact, "sampled_queries": int(sample.size),
            "cpu_baseline": {"value": 1.0 / cpu_frame, "unit": "frames/s",
                             "cores": O.oracle().orc_get_threads(), "kind": "port",
                             "sample": "the oracle's scan for %d sampled queries of each "
                                       "search, scaled to all %d" % (sample.size, S)}}
    print(json.dumps(line), flush=True)


def local_bench(args):
    """calculateLocalLikelihood on one solved frame. Lists as the division test
    builds them (mainTrackingGaussianMixtureModel.cpp:1312-1317): each chosen
    Gaussian's supervoxelIdx[0 .. supervoxelNum). Timed per call from the host
    through the C entry point with the lists already in CSR form (records and
    lists uploaded, results downloaded inside the call); the per-kernel device
    times come from a kernel trace of this run."""
    import numpy as np
    from tgmm_amd import em
    frames, K, iters, workload = _frames_for(args.config)
    if K != 5:
        raise SystemExit("the local log-likelihood has K = 5 slots; use --config 1, 2 or 3")
    f = frames[0]
    with em.Batch([f], em.make_opts(K=K, max_iter=iters, tol=-1.0)) as b:
        b.solve(em.SOLVE_GRAPH)
        res = b.fetch()[0]
    gm = np.ascontiguousarray(res["gm"])
    q, im = em.DeviceArray(f.query), em.DeviceArray(f.img)
    ind = em.DeviceArray(res["ind"].astype(np.int32))
    ptr = em.DeviceArray(f.sv_ptr.astype(np.int64))
    per_g = [list(gm["supervoxelIdx"][k][: gm["supervoxelNum"][k]]) for k in range(f.G)]
    rng = np.random.default_rng(1)
    div = sorted(rng.choice(f.G, max(1, f.G * 3 // 100), replace=False))
    cases = [("divisions_3pct", [per_g[k] for k in div]), ("all_gaussians", per_g)]
    sizes = np.diff(f.sv_ptr)
    out = {}
    L = em.load()
    for name, lists in cases:
        # the C entry point on prebuilt CSR arrays, as a C++ caller passes them
        lptr = np.zeros(len(lists) + 1, np.int64)
        lptr[1:] = np.cumsum([len(l) for l in lists])
        lidx = np.concatenate([np.asarray(l, np.int32) for l in lists]).astype(np.int32)
        res_ll = np.zeros(len(lists))

        def call():
            em._check(L.tgmm_calculateLocalLikelihood(
                res_ll.ctypes.data, lptr.ctypes.data, lidx.ctypes.data, len(lists), q.ptr,
                im.ptr, None, ind.ptr, gm.ctypes.data, ptr.ptr, f.N, f.G, f.S))
        for _ in range(max(1, args.warmup)):
            call()
        got = list(res_ll)
        assert got == em.calculateLocalLikelihood([], lists, q, im, None, ind, gm, ptr, f.N,
                                                  f.G, f.S)
        ts = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            call()
            ts.append(time.perf_counter() - t0)
        ent = sum(len(l) for l in lists)
        vox = int(sum(sizes[s] for l in lists for s in l))
        rec = {"lists": len(lists), "entries": ent, "voxels": vox,
               "ms_per_call_p50": 1e3 * _percentile(ts, 50),
               "ms_per_call_min": 1e3 * min(ts)}
        if not args.no_oracle:
            import oracle_lib as O
            t0 = time.perf_counter()
            want = O.local_loglik(f, gm, res["ind"].astype(np.int32), K, lists)
            cpu_s = time.perf_counter() - t0
            rec["max_rel_err_vs_oracle"] = float(max(abs(a - b) / max(1.0, abs(b))
                                                     for a, b in zip(got, want)))
            rec["cpu_oracle_ms"] = 1e3 * cpu_s
            rec["cpu_threads"] = O.oracle().orc_get_threads()
        out[name] = rec
        print("# %s: %s" % (name, json.dumps(rec)), file=sys.stderr, flush=True)
    d = out["divisions_3pct"]
    line = {"metric": "local log-likelihood calls/s (division lists)",
            "value": 1e3 / d["ms_per_call_p50"], "unit": "calls/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": d["ms_per_call_p50"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": workload + "; local log-likelihood "
                                            "of the per-Gaussian supervoxel lists"},
            "cases": out}
    print(json.dumps(line), flush=True)


def _serve(streams, opts, budget, order, arrivals=None, pct=(0.5, 0.5, 0.8), tiles=None):
    """Feed the frames of `streams` to a fresh server. Closed loop when
    `arrivals` is None (every stream keeps its next frame queued); otherwise
    arrivals[i] is the arrival time (s after the start) of each frame of
    stream i. Returns (elapsed seconds, per-frame records, outputs kept for
    `order`)."""
    from tgmm_amd import em
    ns = len(streams)
    G = [fr[0].G for fr in streams]
    srv = em.Server(opts, hbm_budget_bytes=budget, max_frames_in_flight=2 * ns,
                    beta_pct=pct[0], nu_pct=pct[1], alpha_pct=pct[2], borrow=True)
    if tiles is not None:
        srv.set_tiles(tiles)
    nxt = [0] * ns
    open_t = {}
    recs, keep = [], {}
    total = sum(len(fr) for fr in streams)
    depth = [0] * ns
    if arrivals is not None:
        events = sorted((arrivals[i][t], i, t) for i in range(ns) for t in range(len(streams[i])))
    ai = 0
    t0 = time.monotonic()
    while len(recs) < total:
        now = time.monotonic() - t0
        if arrivals is None:
            for i in range(ns):
                while depth[i] < 2 and nxt[i] < len(streams[i]):
                    t = nxt[i]
                    tk = srv.submit(streams[i][t], stream=i, with_gaussians=(t == 0))
                    open_t[tk] = (i, t, time.monotonic() - t0)
                    nxt[i] += 1
                    depth[i] += 1
        else:
            while ai < len(events) and events[ai][0] <= now:
                ta, i, t = events[ai]
                tk = srv.submit(streams[i][t], stream=i, with_gaussians=(t == 0))
                open_t[tk] = (i, t, ta)
                ai += 1
        if open_t:
            srv.step()
            for tk in list(open_t):
                if srv.query(tk) == 2:
                    i, t, ta = open_t.pop(tk)
                    o = srv.fetch(tk, G[i])
                    recs.append(dict(stream=i, t=t, arrival=ta, done=o["t_done"] - t0,
                                     admit=o["t_admit"] - t0))
                    depth[i] -= 1
                    if (i, t) in order:
                        keep[(i, t)] = o
        elif ai < len(events):
            time.sleep(max(0.0, min(0.002, events[ai][0] - now)))
    elapsed = time.monotonic() - t0
    stats = srv.stats()
    srv.close()
    return elapsed, recs, keep, stats


def _serve_phases(args, stream_frames):
    """The served time-lapse workload: closed loop (throughput) and Poisson
    arrivals (latency) on fresh servers; returns a summary and what the
    accuracy check needs."""
    import numpy as np
    from tgmm_amd import em, synth
    K, iters = 5, 10
    t0 = time.time()
    streams = synth.config4_streams(n_frames=stream_frames)
    n_frames = sum(len(fr) for fr in streams)
    n_vox = sum(f.N for fr in streams for f in fr)
    # the voxel streams in page-locked host memory, as a camera pipeline would
    # deliver them, so that uploads overlap the iterations
    t1 = time.time()
    for fr in streams:
        for f in fr:
            f.query = em.pinned(f.query)
            f.img = em.pinned(f.img)
    print("# page-locked %d frames in %.1fs" % (n_frames, time.time() - t1), file=sys.stderr,
          flush=True)
    cells = (min(f.meta["cells"] for fr in streams for f in fr),
             max(f.meta["cells"] for fr in streams for f in fr))
    print("# generated %d streams x %d frames in %.1fs: %d voxels, cells %d-%d per frame" % (
        len(streams), stream_frames, time.time() - t0, n_vox, cells[0], cells[1]),
        file=sys.stderr, flush=True)
    opts = em.make_opts(K=K, max_iter=iters, tol=-1.0)
    budget = int(args.hbm_budget_gb * (1 << 30))
    pct = (0.5, 0.5, 0.8)

    # launch shapes: autotuned on a batch of the streams' first frames
    if args.no_autotune:
        tiles = dict(em.DEFAULT_TILES)
    else:
        with em.Batch([fr[0] for fr in streams], opts) as tb:
            tiles, _ = tb.autotune(reps=3, cache=False)
    print("# autotuned tiles %s" % json.dumps(tiles), file=sys.stderr, flush=True)

    # warm-up: two frames per stream on a throw-away server
    _serve([fr[:2] for fr in streams], opts, budget, set(), tiles=tiles)

    # closed loop: throughput
    el, recs, keep, st1 = _serve(streams, opts, budget, {(0, 0), (0, 1)}, pct=pct, tiles=tiles)
    thr = n_frames / el
    print("# closed loop: %d frames in %.3fs = %.1f frames/s; host phases %s" % (
        n_frames, el, thr, json.dumps(st1)), file=sys.stderr, flush=True)

    # Poisson arrivals: latency
    qps = args.qps if args.qps > 0 else 0.5 * thr
    rng = np.random.default_rng(0xC0FFEE04)
    ns = len(streams)
    arrivals = [np.cumsum(rng.exponential(ns / qps, len(fr))) for fr in streams]
    el2, recs2, _, st2 = _serve(streams, opts, budget, set(), arrivals=arrivals, pct=pct,
                                tiles=tiles)
    lat = [r["done"] - r["arrival"] for r in recs2]
    qwait = [r["admit"] - r["arrival"] for r in recs2]
    lat_ms = {"p50": 1000 * _percentile(lat, 50), "p99": 1000 * _percentile(lat, 99),
              "max": 1000 * max(lat)}
    print("# Poisson at %.1f frames/s: served %.1f frames/s, latency %s" % (
        qps, n_frames / el2, json.dumps(lat_ms)), file=sys.stderr, flush=True)
    summary = {
        "workload": "config4: %d time-lapse streams x %d frames, %d-%d cells per frame, ~200 "
                    "voxels per cell, K=5, 10 iterations, continuous batching, priors from "
                    "each stream's posterior on the device" % (ns, stream_frames, cells[0],
                                                               cells[1]),
        "frames": n_frames, "voxels": n_vox,
        "closed_loop": {"seconds": el, "frames_per_s": thr, "host_phases": st1},
        "poisson": {"qps": qps, "seconds": el2, "frames_per_s": n_frames / el2,
                    "latency_ms": lat_ms,
                    "queue_wait_ms": {"p50": 1000 * _percentile(qwait, 50),
                                      "p99": 1000 * _percentile(qwait, 99)}},
        "tiles": tiles,
    }
    return summary, streams, keep, (K, iters, pct)


def serve_summary(args, stream_frames):
    summary, streams, keep, _ = _serve_phases(args, stream_frames)
    del streams, keep
    return summary


def serve_bench(args):
    import numpy as np
    from tgmm_amd import synth
    summary, streams, keep, (K, iters, pct) = _serve_phases(args, args.stream_frames)
    ns = len(streams)
    # accuracy: stream 0, frames 0 and 1 (device-propagated priors) against
    # the FP64 oracle chained through the host restatement of the prior update
    acc, cpu = None, None
    if not args.no_oracle:
        import oracle_lib as O
        from helpers import next_priors, record_errors, rel_err
        f0, f1 = streams[0][0], streams[0][1]
        oo = O.make_opts(K=K, max_iter=iters)
        g0, _, ind0, _, _, ll0 = O.em_run(f0, oo)
        dt, reps = _cpu_sample(lambda: O.em_run(f0, oo))
        pri = next_priors(g0, pct[0], pct[1], pct[2], float(g0["alpha_k"].sum()))
        f1g = synth.Frame(f1.query, f1.img, f1.centroids, f1.sv_ptr, pri)
        g1, _, ind1, _, _, ll1 = O.em_run(f1g, oo)
        errs = []
        for o, g, ll, ind in ((keep[(0, 0)], g0, ll0, ind0), (keep[(0, 1)], g1, ll1, ind1)):
            e = record_errors(o["gm"], g)
            e["ll"] = rel_err(o["ll"], ll)
            e["knn_slots_identical"] = bool(np.array_equal(o["ind"], ind))
            e["supervoxel_lists_identical"] = bool(
                np.array_equal(o["gm"]["supervoxelNum"], g["supervoxelNum"]) and
                np.array_equal(o["gm"]["supervoxelIdx"], g["supervoxelIdx"]))
            errs.append(e)
        acc = {k: (max(e[k] for e in errs) if not isinstance(errs[0][k], bool)
                   else all(e[k] for e in errs)) for k in errs[0]}
        acc["frames_checked"] = "stream 0, frames 0 and 1"
        print("# accuracy vs FP64 oracle: %s" % json.dumps(acc), file=sys.stderr, flush=True)
        threads = O.oracle().orc_get_threads()
        cpu = {"value": 1.0 / dt, "unit": "frames/s", "cores": threads, "kind": "port",
               "cpu": _cpu_model(),
               "sample": "full %d-iteration solves of stream 0 frame 0 (N=%d, S=%d, G=%d) "
                         "by the FP64 oracle on %d host threads, repeated %d times (%.1f s)" % (
                             iters, f0.N, f0.S, f0.G, threads, reps, dt * reps)}
    n_frames = summary["frames"]
    el = summary["closed_loop"]["seconds"]
    line = {
        "metric": METRIC,
        "value": summary["closed_loop"]["frames_per_s"],
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": n_frames,
        "warmup": 2 * ns,
        "ms_per_step": 1000.0 * el / n_frames,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator, tgmm_amd/csrc/synth.cpp)",
        "config": {"workload": summary["workload"], "voxels": summary["voxels"], "K": K,
                   "iterations": iters, "parallelism": "1 GPU"},
        "closed_loop": summary["closed_loop"],
        "poisson": summary["poisson"],
        "latency_ms": summary["poisson"]["latency_ms"],
        "accuracy_vs_fp64": acc,
        "tiles": summary["tiles"],
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
