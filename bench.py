#!/usr/bin/env python3
"""Benchmark: exact min-cut throughput on one MI355X.

Workload (the batch the headline metric is quoted on): 256 instances from the
ellipse-image generator, areas log-uniform on [128^2, 1024^2], aspect ratios
log-uniform on [1/2, 2], 16-bin colour data term with lambda = 50 computed on
the device (untimed setup; timed separately as `dataterm`).

A step is one cold solve of the whole batch: inputs resident in HBM ->
residual graphs built -> max-flow -> labels in HBM and int64 flows on the host.
`value` is batched Mpx/s (pixels = W*H of every instance, border included).
Reported beside it:
  * plain: the GC_* drop-in (GC_SetGraph + GC_Optimize, labels left on the
    device) looped over the same instances one at a time, and the ratio;
  * serve: continuous-batching server on a Poisson stream (sizes
    log-uniform on [256^2, 2048^2], 30% seed-stroke edits warm-started on
    device-resident residual graphs): p50/p99 latency, Mpx/s;
  * video: frames of a moving 1024x768 scene on one session, warm-started
    from the previous frame's residual graph, beside the same frames solved
    cold as new images;
  * roofline of the dominant kernel from an event-timed solve of the same batch;
  * cpu_baseline: the CPU restatement of the reference push-relabel schedule
    on a bounded sample of the same instances, with its results checked
    against the GPU's.
Every run checks that batched, plain and served flows agree bit-exactly.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "gc-fg_amd")]

import numpy as np  # noqa: E402

import gcfg  # noqa: E402

METRIC = "min-cut Mpixels/s (bit-exact) batched+served vs plain on 1 GPU; p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per node of one tile visit (halo rows and columns of a
# few hundred bytes per 1024-node tile left out).
PR_BYTES_PER_NODE = 49      # caps 16 + excess 4 + distance 4 read and written, mask 1 written
GR_INIT_BYTES_PER_NODE = 5  # initial BFS sweep from the kept mask: mask 1 read, distance 4 written
GR_RELAX_BYTES_PER_NODE = 9  # relaxation sweep: mask 1 + distance 4 read, distance 4 written
LABEL_BYTES_PER_NODE = 4    # label written by every sweep of an exact round and by FIN
COMPULSORY_BYTES_PER_PX = 28  # 6 int32 inputs read + int32 label written
CHILD_SOLVES = 3  # solves of the counter child run: --steps 1 --warmup 0, then two event-timed solves
BUILD_IN_BYTES_PER_PX = 24  # pos, neg and four n-link planes read by the build
BUILD_OUT_BYTES_PER_NODE = 9  # excess 4 + distance 4 + mask 1 written per arena node (capacities stay lazy)
BUILD_OUT_BYTES_PER_TILE = 640  # the tile's inflow slots cleared (160 int32)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Inst:
    def __init__(self, W, H, seed):
        self.W, self.H, self.seed = int(W), int(H), seed
        self.bufs = None
        self.d_label = None


def make_batch(n, seed):
    Ws, Hs = gcfg.gen_batch_shapes(n, seed, 128.0 ** 2, 1024.0 ** 2)
    insts = []
    dt_ms = 0.0
    for i in range(n):
        it = Inst(Ws[i], Hs[i], gcfg.mix_seed(seed, i))
        img = gcfg.gen_ellipse_image(it.W, it.H, it.seed)
        seeds = gcfg.gen_seed_map(it.W, it.H)
        dt = gcfg.DataTerm(it.W, it.H)
        dt.set_image(img)
        dt.set_seeds(seeds)
        t0 = time.perf_counter()
        dt.compute()
        dt_ms += (time.perf_counter() - t0) * 1e3
        nbytes = 4 * it.W * it.H
        it.bufs = []
        for p in dt.link_ptrs():
            b = gcfg.DeviceBuffer(nbytes)
            b.copy_from_device(p, nbytes)
            it.bufs.append(b)
        dt.close()
        it.d_label = gcfg.DeviceBuffer(nbytes)
        insts.append(it)
    return insts, dt_ms


def batch_items(insts):
    return [gcfg.BatchItem(it.W, it.H, it.bufs[0], it.bufs[1], edges=tuple(it.bufs[2:6])) for it in insts]


def host_links(it):
    n = it.W * it.H
    return [b.download(np.int32, n).reshape(it.H, it.W) for b in it.bufs]


def run_plain(insts, graphs):
    flows = []
    t0 = time.perf_counter()
    for it, g in zip(insts, graphs):
        g.build()
        st = g.optimize_device(it.d_label)
        flows.append(st.flow)
    gcfg.device_synchronize()
    return (time.perf_counter() - t0) * 1e3, np.array(flows, np.int64)


def serve(args, opts, qps, n):
    """Poisson stream against the continuous-batching server."""
    rng = np.random.default_rng(args.seed + 7)
    amin, amax = 256.0 ** 2, 2048.0 ** 2
    reqs = []
    sessions = []  # indices of image requests
    for i in range(n):
        if sessions and rng.random() < 0.3:
            j = int(rng.choice(sessions))
            W, H = reqs[j]["W"], reqs[j]["H"]
            k = int(rng.integers(1, 6))
            strokes = []
            for _ in range(k):
                x0 = int(rng.integers(0, max(1, W - 16)))
                y0 = int(rng.integers(0, max(1, H - 16)))
                strokes.append([x0, y0, 16, 16, int(rng.integers(1, 3))])
            reqs.append(dict(kind="edit", target=j, W=W, H=H, strokes=np.array(strokes, np.int32)))
        else:
            a = math.exp(math.log(amin) + rng.random() * (math.log(amax) - math.log(amin)))
            r = math.exp(math.log(0.5) + rng.random() * math.log(4.0))
            W, H = int(round(math.sqrt(a * r))), int(round(math.sqrt(a / r)))
            img = gcfg.gen_ellipse_image(W, H, gcfg.mix_seed(args.seed + 7, i))
            reqs.append(dict(kind="image", W=W, H=H, img=img, seeds=gcfg.gen_seed_map(W, H)))
            sessions.append(i)
    gaps = rng.exponential(1000.0 / qps, n)
    arrival = np.cumsum(gaps) - gaps[0]
    srv = gcfg.Server(args.serve_budget_gb << 30, opts)
    sess_of = {}
    rid_of = {}
    lat = {}
    flows = {}
    t0 = gcfg.now_ms()
    i = 0
    done = 0
    while done < n:
        now = gcfg.now_ms() - t0
        while i < n and arrival[i] <= now:
            q = reqs[i]
            if q["kind"] == "image":
                rid, sess, _ = srv.submit_image(q["img"], q["seeds"])
                sess_of[i] = sess
            else:
                rid, _ = srv.submit_edit(sess_of[q["target"]], q["strokes"], 1 << 16, shape=(q["H"], q["W"]))
            rid_of[rid] = i
            i += 1
        for r in srv.poll(wait=False):
            k = rid_of[r["request_id"]]
            if r["status"] != 0:
                raise RuntimeError(f"serving request {k} failed with status {r['status']}")
            lat[k] = r["done_ms"] - (t0 + arrival[k])
            flows[k] = r["flow"]
            done += 1
        inflight, queued, _ = srv.load()
        if inflight == 0 and queued == 0 and i < n:
            wait = arrival[i] - (gcfg.now_ms() - t0)
            if wait > 0:
                time.sleep(wait / 1e3)
    wall = gcfg.now_ms() - t0
    srv.close()
    px = sum(q["W"] * q["H"] for q in reqs)
    L = np.array([lat[k] for k in range(n)])
    return dict(requests=n, qps=qps, edits=sum(q["kind"] == "edit" for q in reqs),
                p50_ms=round(float(np.percentile(L, 50)), 3), p99_ms=round(float(np.percentile(L, 99)), 3),
                mean_ms=round(float(L.mean()), 3), mpx_s=round(px / (wall * 1e3), 2), wall_ms=round(wall, 1),
                hbm_budget_gb=args.serve_budget_gb), reqs, flows


def video(args, opts, W=1024, H=768, n=16):
    """Video loop on a serving session: a moving scene (the first image
    shifted by (3k, 2k) px per frame, 10 % of pixels re-drawn). Warm: every
    frame after the first is submit_frame on one session (reparametrised from
    the previous frame's residual graph). Cold: every frame is a new image
    session, built and solved from scratch. Latency is arrival -> labels on
    the host; warm flows are checked to repeat exactly across two runs (their
    exactness against cold solves is a GPU test)."""
    rng = np.random.default_rng(args.seed + 11)
    base = gcfg.gen_ellipse_image(W, H, args.seed + 11)
    frames = [base]
    for k in range(1, n):
        f = np.roll(base, (2 * k, 3 * k), axis=(0, 1)).copy()
        m = rng.random((H, W)) < 0.1
        f[m] = rng.integers(0, 256, (int(m.sum()), 3))
        frames.append(f.astype(np.uint8))
    seeds = gcfg.gen_seed_map(W, H)

    def run(warm):
        srv = gcfg.Server(args.serve_budget_gb << 30, opts)
        lat, flows = [], []
        sess = None
        for k, f in enumerate(frames):
            t0 = gcfg.now_ms()
            if warm and k > 0:
                rid, _ = srv.submit_frame(sess, f)
            else:
                rid, s, _ = srv.submit_image(f, seeds)
                sess = s if sess is None else sess
            r = None
            while r is None:
                for q in srv.poll(wait=True):
                    if q["request_id"] == rid:
                        r = q
            assert r["status"] == 0
            lat.append(r["done_ms"] - t0)
            flows.append(int(r["flow"]))
            if not warm and k > 0:
                srv.close_session(r["session"])
        srv.close()
        return np.array(lat), flows

    run(True)  # warm-up (first use of the server path)
    lw, fw = run(True)
    lc, _ = run(False)
    lw2, fw2 = run(True)
    assert fw == fw2, "video frames differ between runs"
    return dict(frames=n, shape=f"{W}x{H}", warm_frame_ms=round(float(np.mean(lw[1:])), 3),
                cold_frame_ms=round(float(np.mean(lc[1:])), 3),
                warm_p99_ms=round(float(np.percentile(lw[1:], 99)), 3),
                warm_mpx_s=round(W * H * (n - 1) / float(np.sum(lw[1:])) / 1e3, 1))


def cpu_baseline(insts, flows_gpu, labels_gpu_fn, budget_s):
    """Reference-schedule restatement on the host cores over a bounded sample."""
    import oracle

    threads = min(16, os.cpu_count() or 1)
    # Size-stratified order (every 8th instance by area, then the next
    # offset, ...), so that a sample cut short by the time budget still spans
    # the size range; with the default budget it covers most of the batch.
    by_size = sorted(range(len(insts)), key=lambda k: insts[k].W * insts[k].H)
    order = [by_size[j] for off in range(8) for j in range(off, len(by_size), 8)]
    px = 0
    secs = 0.0
    used = []
    match = True
    for k in order:
        it = insts[k]
        pos, neg, up, down, left, right = host_links(it)
        t0 = time.perf_counter()
        f, lab, iters, _ = oracle.refpr(it.W, it.H, pos, neg, edges=(up, down, left, right), threads=threads)
        secs += time.perf_counter() - t0
        px += it.W * it.H
        used.append(k)
        match &= (f == int(flows_gpu[k])) and bool(np.array_equal(lab, labels_gpu_fn(k)))
        if secs > budget_s:
            break
    out = dict(value=round(px / secs / 1e6, 3), unit="Mpx/s", cores=threads, kind="port",
               sample=f"{len(used)} of the {len(insts)} batch instances ({px / 1e6:.1f} Mpx), "
                      f"reference push-relabel schedule restated in C/OpenMP",
               seconds=round(secs, 2), matches_gpu=bool(match), cpu_model=cpu_model())
    # The single-image configurations: 320x240 (1 thread and all threads) and
    # 1024^2 (all threads), each checked against the GPU drop-in path.
    per = {}
    for kind, W, H, thread_list in (("c0", 320, 240, (1, threads)), ("c1", 1024, 1024, (threads,))):
        img = gcfg.gen_ellipse_image(W, H, 1)
        dt = gcfg.DataTerm(W, H)
        dt.set_image(img)
        dt.set_seeds(gcfg.gen_seed_map(W, H))
        dt.compute()
        pos, neg, up, down, left, right = dt.links()
        dt.close()
        gflow, glab, _ = gcfg.solve(W, H, pos, neg, edges=(up, down, left, right))
        row = {}
        for th in thread_list:
            t0 = time.perf_counter()
            f, lab, _, _ = oracle.refpr(W, H, pos, neg, edges=(up, down, left, right), threads=th)
            dt_s = time.perf_counter() - t0
            row[f"threads_{th}_mpx_s"] = round(W * H / dt_s / 1e6, 3)
            row["matches_gpu"] = bool(f == gflow and np.array_equal(lab, glab))
        per[kind] = row
    out["configs"] = per
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def single_instance(kind, seed):
    if kind == "c1":
        W = H = 1024
        img = gcfg.gen_ellipse_image(W, H, seed)
        seeds = gcfg.gen_seed_map(W, H)
    else:
        W = H = 4096
        img = gcfg.gen_perlin_image(W, H, seed)
        seeds = np.zeros((H, W), np.uint8)
    dt = gcfg.DataTerm(W, H)
    dt.set_image(img)
    dt.set_seeds(seeds)
    dt.compute()
    bufs = []
    for p in dt.link_ptrs():
        b = gcfg.DeviceBuffer(4 * W * H)
        b.copy_from_device(p, 4 * W * H)
        bufs.append(b)
    dt.close()
    if kind == "c2":
        pos, neg = gcfg.gen_stress_tlinks(W, H, seed + 1)
        bufs[0].upload(pos)
        bufs[1].upload(neg)
    return W, H, bufs


def extras(args):
    """Rows beyond the batch: the two data-term models and the 4- vs
    8-neighbourhood solver on the 1024^2 ellipse image (default options,
    penalty 50 for both neighbourhoods; medians of 8 runs after 2 warm-ups)."""
    def timeit(fn, reps=8):
        fn()
        fn()
        ts = []
        for _ in range(reps):
            gcfg.device_synchronize()
            t0 = time.perf_counter()
            fn()
            gcfg.device_synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return round(float(np.median(ts)), 3)

    W = H = 1024
    img = gcfg.gen_ellipse_image(W, H, args.seed)
    seeds = gcfg.gen_seed_map(W, H)
    out = dict(workload="extras", W=W, H=H)
    dt = gcfg.DataTerm(W, H)
    dt.set_image(img)
    dt.set_seeds(seeds)
    out["dataterm_histogram_ms"] = timeit(dt.compute)
    pos, neg = dt.links()[:2]
    dt.close()
    dk = gcfg.DataTerm(W, H)
    dk.set_model(gcfg.DT_MODEL_KDE)
    dk.set_sigma(30)
    dk.set_image(img)
    for k in range(8):
        dk.kde_select(100 + 100 * k, 500)
    out["dataterm_kde_seeds"] = dk.kde_count()
    out["dataterm_kde_ms"] = timeit(dk.compute)
    dk.close()
    for nb in (4, 8):
        g = gcfg.Graph(W, H, gcfg.default_options(), neighborhood=nb)
        g.set_dataterms(pos, neg)
        g.set_penalty(50)
        d_label = gcfg.DeviceBuffer(4 * W * H)

        def run():
            g.build()
            g.optimize_device(d_label)
        out[f"n{nb}_penalty_ms"] = timeit(run)
        st = g.optimize_device(d_label)
        out[f"n{nb}_steps"] = int(st.steps)
        out[f"n{nb}_flow"] = int(st.flow)
        g.close()
    print(json.dumps(out), flush=True)


def single(args, opts):
    """One plain GC_* drop-in solve at C1 or C2, with labels left on the device."""
    W, H, bufs = single_instance(args.single, args.seed)
    g = gcfg.Graph(W, H, opts)
    g.set_dataterms(bufs[0], bufs[1])
    g.set_edges(*bufs[2:6])
    d_label = gcfg.DeviceBuffer(4 * W * H)
    g.build()
    g.optimize_device(d_label)
    times, st = [], None
    for s in range(args.warmup + args.steps):
        gcfg.device_synchronize()
        t0 = time.perf_counter()
        g.build()
        st = g.optimize_device(d_label)
        gcfg.device_synchronize()
        if s >= args.warmup:
            times.append((time.perf_counter() - t0) * 1e3)
        log(f"{args.single} solve {s}: {(time.perf_counter() - t0) * 1e3:.2f} ms flow {st.flow} steps {st.steps} "
            f"gr_rounds {st.gr_rounds} pr_tiles {st.pr_tiles} gr_tiles {st.gr_tiles}")
    ms = float(np.mean(times))
    out = dict(workload=args.single, W=W, H=H, ms=round(ms, 3), mpx_s=round(W * H / ms / 1e3, 2), flow=int(st.flow),
               steps=int(st.steps), gr_rounds=int(st.gr_rounds), pr_tiles=int(st.pr_tiles),
               gr_tiles=int(st.gr_tiles), sweeps_per_gr=opts.sweeps_per_gr, inner_iters=opts.inner_iters)
    if not args.no_cpu:
        import oracle
        n = W * H
        host = [b.download(np.int32, n).reshape(H, W) for b in bufs]
        t0 = time.perf_counter()
        rf, rl = oracle.maxflow(W, H, host[0], host[1], edges=tuple(host[2:6]))
        out["oracle_s"] = round(time.perf_counter() - t0, 2)
        lab = d_label.download(np.int32, n).reshape(H, W)
        out["matches_oracle"] = bool(rf == st.flow and np.array_equal(rl, lab))
    g.close()
    print(json.dumps(out), flush=True)


def pmc_traffic(args, opts, kernel_prefix, launches_per_solve):
    """HBM bytes per launch of the dominant kernel from two rocprofv3 counter
    runs of this script (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each
    scaled by what the same counter reads for a 1 GiB device-to-device copy
    run in the same process. The child solves the batch CHILD_SOLVES times
    (one timed solve, two event-timed ones); a first solve may launch a few
    idle steps whose dispatches exit at once, so bytes per launch are the
    summed bytes over every dispatch / (solves x launches of one solve).
    Returns (traffic dict or None, note)."""
    import shutil
    import sqlite3
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    child = [sys.executable, os.path.abspath(__file__), "--no-cpu", "--no-serve", "--no-plain", "--no-pmc",
             "--steps", "1", "--warmup", "0", "--calibrate-copy", "--instances", str(args.instances),
             "--seed", str(args.seed), "--sweeps", str(opts.sweeps_per_gr), "--inner", str(opts.inner_iters),
             "--poll", str(opts.steps_per_poll), "--gr-block", f"{opts.gr_block_w}x{opts.gr_block_h}",
             "--gr-cap", str(opts.gr_cap)]
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="gcfg_pmc_")
        cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", counter, "-d", d, "-o", "run", "--"] + child
        log(f"counter run: {counter}")
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        db = os.path.join(d, "run_results.db")
        if r.returncode != 0 or not os.path.exists(db):
            shutil.rmtree(d, ignore_errors=True)
            return None, f"rocprofv3 --pmc {counter} failed (exit {r.returncode})"
        con = sqlite3.connect(db)
        rows = con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                           (counter,)).fetchall()
        con.close()
        shutil.rmtree(d, ignore_errors=True)
        calib = max((v for n, v in rows if "copyBuffer" in n), default=0.0)
        vals = [v for n, v in rows if kernel_prefix in n]
        if calib <= 0 or not vals:
            return None, f"no {counter} samples for the copy or {kernel_prefix}"
        scale = (1 << 20) / calib  # the copy moves 1 GiB = 2^20 KiB each way
        per_launch = sum(vals) / (CHILD_SOLVES * max(1, launches_per_solve))
        got[counter] = (per_launch * 1024 * scale, scale, len(vals))
    fetch, fscale, n = got["FETCH_SIZE"]
    write, wscale, _ = got["WRITE_SIZE"]
    return dict(bytes=int(fetch + write), fetch_bytes=int(fetch), write_bytes=int(write),
                fetch_scale=round(fscale, 4), write_scale=round(wscale, 4), dispatches=n), "ok"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--qps", type=float, default=40.0)
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-budget-gb", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="time budget of the CPU baseline's batch sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--sweeps", type=int, default=0)
    ap.add_argument("--inner", type=int, default=0)
    ap.add_argument("--poll", type=int, default=0)
    ap.add_argument("--gr-block", default=None, help="global-relabel block WxH in tiles, e.g. 1x4")
    ap.add_argument("--gr-cap", type=int, default=None, help="distance cap of intermediate global relabels")
    ap.add_argument("--gr-waves", type=int, default=None, help="relabel launch wavefronts per CU")
    ap.add_argument("--pr-waves", type=int, default=None, help="push/relabel launch wavefronts per CU")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter runs for roofline.traffic")
    ap.add_argument("--tune-cache", default=os.environ.get("GCFG_TUNE_CACHE",
                                                           os.path.expanduser("~/.cache/gcfg/tune.txt")))
    ap.add_argument("--serve-qps-high", type=float, default=400.0)
    ap.add_argument("--serve-requests-high", type=int, default=200)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="copy 1 GiB device-to-device first (a known-size kernel for profiler counters)")
    ap.add_argument("--extras", action="store_true",
                    help="time the two data-term models and the 4- vs 8-neighbourhood solver at 1024^2 instead")
    ap.add_argument("--single", choices=["c1", "c2"], default=None,
                    help="time one plain GC_* solve instead: c1 = 1024^2 ellipse image, "
                         "c2 = 4096^2 heavy-tailed t-links on Perlin n-links")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path is not sharded across GPUs",
              file=sys.stderr)
        sys.exit(2)
    if not gcfg.gpu_available():
        raise SystemExit("no GPU visible: bench.py needs an MI355X")

    opts = gcfg.default_options()
    if args.extras:
        extras(args)
        return
    if args.single:
        if args.sweeps:
            opts.sweeps_per_gr = args.sweeps
        if args.inner:
            opts.inner_iters = args.inner
        if args.gr_block:
            opts.gr_block_w, opts.gr_block_h = (int(v) for v in args.gr_block.split("x"))
        if args.gr_cap is not None:
            opts.gr_cap = args.gr_cap
        single(args, opts)
        return
    if args.sweeps:
        opts.sweeps_per_gr = args.sweeps
    if args.inner:
        opts.inner_iters = args.inner
    if args.poll:
        opts.steps_per_poll = args.poll
    if args.gr_block:
        opts.gr_block_w, opts.gr_block_h = (int(v) for v in args.gr_block.split("x"))
    if args.gr_cap is not None:
        opts.gr_cap = args.gr_cap
    if args.gr_waves:
        opts.gr_waves_per_cu = args.gr_waves
    if args.pr_waves:
        opts.pr_waves_per_cu = args.pr_waves

    log(f"generating {args.instances} instances + device data terms")
    insts, dt_ms = make_batch(args.instances, args.seed)
    total_px = sum(it.W * it.H for it in insts)
    log(f"batch: {total_px / 1e6:.2f} Mpx, data term {dt_ms:.1f} ms")

    if args.calibrate_copy:
        a, b = gcfg.DeviceBuffer(1 << 30), gcfg.DeviceBuffer(1 << 30)
        b.copy_from_device(a.ptr, 1 << 30)
        gcfg.device_synchronize()
        a.free()
        b.free()
    batch = gcfg.Batch(batch_items(insts), opts)
    tuned = None
    if not args.no_autotune and not (args.sweeps or args.inner or args.poll or args.gr_block or
                                     args.gr_cap is not None):
        key = f"batch:ellipse:n{args.instances}:seed{args.seed}"
        cache = args.tune_cache
        try:
            os.makedirs(os.path.dirname(cache) or ".", exist_ok=True)
            tuned = gcfg.tune_lookup(cache, key)
        except (OSError, gcfg.GcfgError):
            cache, tuned = None, None
        if tuned is None:
            t0 = time.perf_counter()
            tuned, tms = batch.autotune(trials=2)
            if cache:
                try:
                    gcfg.tune_store(cache, key, tuned)
                except (OSError, gcfg.GcfgError):
                    pass
            log(f"autotune: {gcfg.options_dict(tuned)} ({tms:.2f} ms) in {time.perf_counter() - t0:.1f} s")
        batch.close()
        opts = tuned
        batch = gcfg.Batch(batch_items(insts), opts)
    for _ in range(args.warmup):
        flows_b, _, _ = batch.solve(labels_to_host=False)
    gcfg.device_synchronize()
    times = []
    st = None
    for s in range(args.steps):
        gcfg.device_synchronize()
        t0 = time.perf_counter()
        flows_b, _, st = batch.solve(labels_to_host=False)
        gcfg.device_synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        log(f"batched step {s}: {times[-1]:.2f} ms, {total_px / times[-1] / 1e3:.1f} Mpx/s, steps {st.steps}")
    ms = float(np.mean(times))
    value = total_px / ms / 1e3

    # Event-timed solve of the same batch for the roofline.
    popts = gcfg.Options.from_buffer_copy(opts)
    popts.profile = 1
    pbatch = gcfg.Batch(batch_items(insts), popts)
    pbatch.solve(labels_to_host=False)  # the second solve launches exactly the steps the first needed
    flows_p, _, pst = pbatch.solve(labels_to_host=False)
    n_tiles = sum(((it.W + 63) // 64) * ((it.H + 15) // 16) for it in insts)
    pr_bytes = pst.pr_tiles * 1024 * PR_BYTES_PER_NODE
    per_item = 1024 * max(1, opts.gr_block_w) * max(1, opts.gr_block_h)  # nodes of one relabel item
    gr_bytes = per_item * (pst.gr_init_tiles * GR_INIT_BYTES_PER_NODE +
                       (pst.gr_tiles - pst.gr_init_tiles) * GR_RELAX_BYTES_PER_NODE +
                       (pst.gr_exact_tiles + pst.fin_items) * LABEL_BYTES_PER_NODE)
    wave = opts.kernels != 0
    # Build: the caller's six input planes read, residual graph written
    # (arena nodes incl. padding), labels too when the first sweep is exact.
    build_bytes = BUILD_IN_BYTES_PER_PX * total_px + BUILD_OUT_BYTES_PER_NODE * n_tiles * 1024 + \
        BUILD_OUT_BYTES_PER_TILE * n_tiles + \
        (LABEL_BYTES_PER_NODE * total_px if opts.gr_cap == 0 else 0)
    # name: (event-timed ms per solve, algorithmic bytes per solve, launches per solve)
    kern = {("k_buildw" if wave else "k_build"): (pst.build_kernel_ms, build_bytes, 1),
            ("k_prw" if wave else "k_pr"): (pst.pr_kernel_ms, pr_bytes, int(pst.launches)),
            ("k_grb" if wave else "k_gr"): (pst.gr_kernel_ms, gr_bytes, int(pst.launches))}
    per_kernel = {k: dict(ms=round(v[0], 4), gb_s=round(v[1] / (v[0] * 1e-3) / 1e9, 1) if v[0] > 0 else None,
                          frac=round(v[1] / (v[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if v[0] > 0 else None,
                          bytes=int(v[1]), launches=v[2]) for k, v in kern.items()}
    dom = max(kern, key=lambda k: kern[k][0])
    dms, dbytes, dlaunch = kern[dom]
    achieved = dbytes / (dms * 1e-3) / 1e9 if dms > 0 else 0.0
    solve_bytes = COMPULSORY_BYTES_PER_PX * total_px + pr_bytes + gr_bytes
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    avg_launch_us=round(dms * 1e3 / max(dlaunch, 1), 2),
                    bytes_per_launch=int(dbytes / max(dlaunch, 1)), launches=int(dlaunch),
                    kernels=per_kernel, steps=int(pst.launches),
                    pr_kernel_ms=round(pst.pr_kernel_ms, 3), gr_kernel_ms=round(pst.gr_kernel_ms, 3),
                    sched_kernel_ms=round(pst.sched_kernel_ms, 3), build_kernel_ms=round(pst.build_kernel_ms, 3),
                    solve_alg_frac=round(solve_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    compulsory_frac=round(COMPULSORY_BYTES_PER_PX * total_px / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                    pr_tiles=int(pst.pr_tiles), gr_tiles=int(pst.gr_tiles),
                    gr_init_tiles=int(pst.gr_init_tiles), gr_exact_tiles=int(pst.gr_exact_tiles),
                    fin_items=int(pst.fin_items), arena_tiles=int(n_tiles))
    assert np.array_equal(flows_p, flows_b), "profiled solve differs"
    pbatch.close()
    if not args.no_pmc:
        name = {"k_prw": "k_prw(", "k_pr": "k_pr(", "k_grb": "k_grb<", "k_gr": "k_gr(", "k_buildw": "k_buildw(",
                "k_build": "k_build("}[dom]
        traffic, note = pmc_traffic(args, opts, name, dlaunch)
        if traffic:
            roofline["traffic"] = traffic["bytes"]
            roofline["traffic_detail"] = traffic
            log(f"traffic of {dom}: {traffic}")
        else:
            roofline["traffic_note"] = note
            log(f"traffic not measured: {note}")

    extra = {}
    if not args.no_plain:
        # The plain path gets its own tuned options (on the median-size instance).
        popts = gcfg.Options.from_buffer_copy(opts)
        if tuned is not None:
            med = sorted(insts, key=lambda it: it.W * it.H)[len(insts) // 2]
            g = gcfg.Graph(med.W, med.H, gcfg.default_options())
            g.set_dataterms(med.bufs[0], med.bufs[1])
            g.set_edges(*med.bufs[2:6])
            g.build()
            popts, _ = g.autotune(trials=2)
            g.close()
            extra["plain_options"] = {k: getattr(popts, k) for k in ("sweeps_per_gr", "inner_iters", "gr_rounds",
                                                                   "steps_per_poll", "gr_block_w", "gr_block_h",
                                                                   "gr_cap")}
        graphs = []
        for it in insts:
            g = gcfg.Graph(it.W, it.H, popts)
            g.set_dataterms(it.bufs[0], it.bufs[1])
            g.set_edges(*it.bufs[2:6])
            g.build()
            g.optimize_device(it.d_label)
            graphs.append(g)
        plain_ms, flows_plain = run_plain(insts, graphs)
        log(f"plain pass: {plain_ms:.1f} ms, {total_px / plain_ms / 1e3:.2f} Mpx/s")
        assert np.array_equal(flows_plain, flows_b), "plain and batched flows differ"
        for k in range(0, len(insts), max(1, len(insts) // 8)):
            it = insts[k]
            a = it.d_label.download(np.int32, it.W * it.H)
            b = batch.device_labels(k).ravel()
            assert np.array_equal(a, b), f"plain and batched labels differ on instance {k}"
        extra["plain_mpx_s"] = round(total_px / plain_ms / 1e3, 3)
        extra["plain_ms"] = round(plain_ms, 2)
        extra["speedup_vs_plain"] = round(value / (total_px / plain_ms / 1e3), 2)
        for g in graphs:
            g.close()

    if not args.no_serve:
        log(f"serving {args.serve_requests} requests at {args.qps} qps")
        sv, reqs, sflows = serve(args, opts, args.qps, args.serve_requests)
        extra["serve"] = sv
        log(f"serve: {sv}")
        log(f"serving {args.serve_requests_high} requests at {args.serve_qps_high} qps")
        sv2, _, _ = serve(args, opts, args.serve_qps_high, args.serve_requests_high)
        extra["serve_high_load"] = sv2
        log(f"serve (high load): {sv2}")

    if not args.no_serve:
        log("video loop: warm frames on a session vs cold images")
        extra["video"] = video(args, opts)
        log(f"video: {extra['video']}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (reference schedule restated on the host)")
        cpu = cpu_baseline(insts, flows_b, lambda k: batch.device_labels(k), args.cpu_seconds)
        log(f"cpu: {cpu}")
    batch.close()

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"batch of {args.instances} ellipse images, area log-uniform 128^2..1024^2, "
                               f"aspect 1:2..2:1, 16-bin colour data term, lambda 50, cold solves",
                   "instances": args.instances, "total_mpx": round(total_px / 1e6, 3), "seed": args.seed,
                   "sweeps_per_gr": opts.sweeps_per_gr, "inner_iters": opts.inner_iters,
                   "steps_per_poll": opts.steps_per_poll, "gr_rounds": opts.gr_rounds,
                   "gr_block": f"{opts.gr_block_w}x{opts.gr_block_h}", "gr_cap": opts.gr_cap,
                   "autotuned": tuned is not None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "dataterm_mpx_s": round(total_px / dt_ms / 1e3, 2),
        "scheduler_steps": int(st.steps) if st is not None else None,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
