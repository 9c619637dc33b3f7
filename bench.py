#!/usr/bin/env python3
"""Benchmark: exact min-cut throughput on one MI355X.

Workload (the batch the headline metric is quoted on): 256 instances from the
ellipse-image generator, areas log-uniform on [128^2, 1024^2], aspect ratios
log-uniform on [1/2, 2], 16-bin colour data term with lambda = 50 computed on
the device by the batched data term (timed separately as `dataterm`).

A step is one cold solve of the whole batch: inputs resident in HBM ->
residual graphs built -> max-flow -> labels on the host (packed one bit per
pixel, one transfer for the batch) and int64 flows on the host. `value` is
that rate in Mpx/s (pixels = W*H of every instance, border included).
Reported beside it:
  * batched with the labels left in HBM, and with int32 labels copied to the
    host;
  * plain: the GC_* drop-in (GC_SetGraph + GC_Optimize with int32 labels to
    the host, as the reference returns them) looped over the same instances
    one at a time, and the ratios;
  * first_solve_unseen: one solve of a batch from another seed on a fresh
    handle (no step-count hint from an earlier solve), checked against the
    oracle;
  * serve: the continuous-batching server on Poisson streams of 1000 requests
    per point (sizes log-uniform on [256^2, 2048^2], 30% edits of 1-5 seed
    strokes warm-started on device-resident residual graphs), swept over
    arrival rates until it saturates: p50/p99 latency (arrival -> labels on
    the host), served Mpx/s; every result's flow and labels are compared with
    a cold solve of the same capacities (serve.matches_cold);
  * video: frames of a moving 1024x768 scene on one session, warm-started
    from the previous frame's residual graph, beside the same frames solved
    cold as new images;
  * roofline of the dominant kernel from an event-timed solve of the same batch;
  * cpu_baseline: the CPU restatement of the reference push-relabel schedule
    on a bounded sample of the same instances, with its results checked
    against the GPU's.
Every run checks that batched (packed and int32 labels), plain, served and
oracle results agree bit-exactly where it compares them.
"""
from __future__ import annotations

import argparse
import ctypes
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
DATATERM_BYTES_PER_PX = 30  # RGB read by two passes + six int32 planes written
BUILD_IN_BYTES_PER_PX = 24  # pos, neg and four n-link planes read by the build
BUILD_OUT_BYTES_PER_NODE = 9  # excess 4 + distance 4 + mask 1 written per arena node (capacities stay lazy)
BUILD_OUT_BYTES_PER_TILE = 640  # the tile's inflow slots cleared (160 int32)
_VP = ctypes.c_void_p


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Inst:
    def __init__(self, W, H, seed):
        self.W, self.H, self.seed = int(W), int(H), seed
        self.bufs = None
        self.d_label = None


def make_batch(n, seed):
    """Instances and their data terms (one batched data-term launch per
    pass). Returns (instances, data-term handle, data-term ms)."""
    Ws, Hs = gcfg.gen_batch_shapes(n, seed, 128.0 ** 2, 1024.0 ** 2)
    insts, images = [], []
    for i in range(n):
        it = Inst(Ws[i], Hs[i], gcfg.mix_seed(seed, i))
        images.append((gcfg.gen_ellipse_image(it.W, it.H, it.seed), gcfg.gen_seed_map(it.W, it.H)))
        insts.append(it)
    dtb = gcfg.DataTermBatch(images)
    dtb.compute()  # first launch of the kernels
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        dtb.compute()
        times.append((time.perf_counter() - t0) * 1e3)
    for i, it in enumerate(insts):
        it.bufs = dtb.planes[i]
        it.d_label = gcfg.DeviceBuffer(4 * it.W * it.H)
    return insts, dtb, float(np.median(times))


def batch_items(insts):
    return [gcfg.BatchItem(it.W, it.H, it.bufs[0], it.bufs[1], edges=tuple(it.bufs[2:6])) for it in insts]


def host_links(it):
    n = it.W * it.H
    return [b.download(np.int32, n).reshape(it.H, it.W) for b in it.bufs]


class BitsSolver:
    """gcfg_batch_solve_bits through preallocated ctypes arguments, so the
    timed region is the library call."""

    def __init__(self, batch):
        self.b = batch
        self.n = len(batch.items)
        self.flows = np.zeros(self.n, np.int64)
        self.ptrs = (ctypes.POINTER(ctypes.c_uint64) * self.n)()
        self.st = gcfg.Stats()

    def solve(self):
        rc = gcfg.lib().gcfg_batch_solve_bits(_VP(self.b.h), self.ptrs, _VP(self.flows.ctypes.data),
                                              ctypes.byref(self.st))
        if rc != 0:
            raise RuntimeError(f"gcfg_batch_solve_bits failed ({rc}): {gcfg.lib().gcfg_last_error().decode()}")
        return self.flows

    def words(self, i):
        it = self.b.items[i]
        return np.ctypeslib.as_array(self.ptrs[i], shape=(it.height, (it.width + 63) // 64))


class Int32Solver:
    """gcfg_batch_solve with int32 labels into preallocated host arrays."""

    def __init__(self, batch):
        self.b = batch
        n = len(batch.items)
        self.host = [np.empty((it.height, it.width), np.int32) for it in batch.items]
        self.ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in self.host])
        self.flows = np.zeros(n, np.int64)
        self.st = gcfg.Stats()

    def solve(self):
        rc = gcfg.lib().gcfg_batch_solve(_VP(self.b.h), self.ptrs, _VP(self.flows.ctypes.data), ctypes.byref(self.st))
        if rc != 0:
            raise RuntimeError(f"gcfg_batch_solve failed ({rc}): {gcfg.lib().gcfg_last_error().decode()}")
        return self.flows


def time_steps(fn, steps, warmup, label):
    for _ in range(warmup):
        fn()
    gcfg.device_synchronize()
    times = []
    for s in range(steps):
        gcfg.device_synchronize()
        t0 = time.perf_counter()
        fn()
        gcfg.device_synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        log(f"{label} step {s}: {times[-1]:.3f} ms")
    return float(np.mean(times)), times


def run_plain(insts, graphs, hosts):
    """GC_SetGraph + GC_Optimize per instance, int32 labels to the host."""
    flows = []
    st = gcfg.Stats()
    t0 = time.perf_counter()
    for it, g, h in zip(insts, graphs, hosts):
        g.build()
        rc = gcfg.lib().gcfg_graph_optimize(_VP(g.h), _VP(h.ctypes.data), ctypes.byref(st))
        if rc != 0:
            raise RuntimeError(f"gcfg_graph_optimize failed ({rc})")
        flows.append(st.flow)
    return (time.perf_counter() - t0) * 1e3, np.array(flows, np.int64)


# ---------------------------------------------------------------------------
# Serving
def serve_workload(args, n):
    """A deterministic request list: image requests drawn from a pool of
    distinct images (areas log-uniform on [256^2, 2048^2], aspect 1:2..2:1,
    ellipse generator, in pinned host memory), 30% edits of 1-5 random 16x16
    strokes on one of the 8 most recent sessions."""
    rng = np.random.default_rng(args.seed + 7)
    amin, amax = 256.0 ** 2, 2048.0 ** 2
    pool = []
    for k in range(args.serve_pool):
        a = math.exp(math.log(amin) + rng.random() * (math.log(amax) - math.log(amin)))
        r = math.exp(math.log(0.5) + rng.random() * math.log(4.0))
        W, H = int(round(math.sqrt(a * r))), int(round(math.sqrt(a / r)))
        img = gcfg.pinned_empty((H, W, 3), np.uint8)
        img[...] = gcfg.gen_ellipse_image(W, H, gcfg.mix_seed(args.seed + 7, k))
        seeds = gcfg.pinned_empty((H, W), np.uint8)
        seeds[...] = gcfg.gen_seed_map(W, H)
        pool.append((W, H, img, seeds))
    reqs, live = [], []
    for i in range(n):
        if live and rng.random() < 0.3:
            j = live[int(rng.integers(len(live)))]
            W, H = reqs[j]["W"], reqs[j]["H"]
            strokes = [[int(rng.integers(0, max(1, W - 16))), int(rng.integers(0, max(1, H - 16))), 16, 16,
                        int(rng.integers(1, 3))] for _ in range(int(rng.integers(1, 6)))]
            reqs.append(dict(kind="edit", target=j, W=W, H=H, strokes=np.array(strokes, np.int32)))
        else:
            p = int(rng.integers(len(pool)))
            reqs.append(dict(kind="image", pool=p, W=pool[p][0], H=pool[p][1]))
            live.append(i)
            if len(live) > 8:
                live.pop(0)
    last = {}
    for i, q in enumerate(reqs):
        last[i if q["kind"] == "image" else q["target"]] = i
    return pool, reqs, last


def serve_expected(pool, reqs):
    """Flow and packed labels of a cold plain solve of every request's
    capacities (an edit: its session's data term plus every stroke applied to
    the session so far)."""
    links = []
    for W, H, img, seeds in pool:
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
        links.append(bufs)
    host_pn = {}
    strokes_of = {}
    out = []
    for i, q in enumerate(reqs):
        W, H = q["W"], q["H"]
        if q["kind"] == "image":
            bufs = links[q["pool"]]
            strokes_of[i] = []
            pos_b, neg_b = bufs[0], bufs[1]
            tmp = None
        else:
            j = q["target"]
            strokes_of[j].extend(q["strokes"].tolist())
            bufs = links[reqs[j]["pool"]]
            if reqs[j]["pool"] not in host_pn:
                host_pn[reqs[j]["pool"]] = (bufs[0].download(np.int32, W * H).reshape(H, W),
                                            bufs[1].download(np.int32, W * H).reshape(H, W))
            pos, neg = (a.astype(np.int64) for a in host_pn[reqs[j]["pool"]])
            for x0, y0, w, h, cls in strokes_of[j]:
                (pos if cls == 1 else neg)[y0:y0 + h, x0:x0 + w] += 1 << 16
            pos_b = gcfg.DeviceBuffer.from_array(pos.astype(np.int32))
            neg_b = gcfg.DeviceBuffer.from_array(neg.astype(np.int32))
            tmp = (pos_b, neg_b)
        g = gcfg.Graph(W, H)
        g.set_dataterms(pos_b, neg_b)
        g.set_edges(*bufs[2:6])
        g.build()
        label, st = g.optimize()
        g.close()
        out.append((int(st.flow), gcfg.pack_bits(label)))
        del tmp
    return out


def serve_point(args, opts, pool, reqs, last, expected, qps):
    """One Poisson stream at `qps` against a fresh server (packed labels)."""
    n = len(reqs)
    rng = np.random.default_rng(args.seed + 13)
    gaps = rng.exponential(1000.0 / qps, n)
    arrival = np.cumsum(gaps) - gaps[0]
    srv = gcfg.Server(args.serve_budget_gb << 30, opts, label_format=gcfg.LABELS_BITS)
    sess_of, rid_of, lat = {}, {}, {}
    matches = True
    t0 = gcfg.now_ms()
    i = done = 0
    while done < n:
        now = gcfg.now_ms() - t0
        while i < n and arrival[i] <= now:
            q = reqs[i]
            # Labels are lent by the server (pinned memory, valid until the
            # next poll) instead of copied into a fresh caller array.
            if q["kind"] == "image":
                W, H, img, seeds = pool[q["pool"]]
                rid, sess, _ = srv.submit_image(img, seeds, borrow=True)
                sess_of[i] = sess
            else:
                rid, _ = srv.submit_edit(sess_of[q["target"]], q["strokes"], 1 << 16, shape=(q["H"], q["W"]),
                                         borrow=True)
            rid_of[rid] = i
            i += 1
        for r in srv.poll(wait=False):
            k = rid_of[r["request_id"]]
            if r["status"] != 0:
                raise RuntimeError(f"serving request {k} failed with status {r['status']}")
            lat[k] = r["done_ms"] - (t0 + arrival[k])
            ef, ew = expected[k]
            if r["flow"] != ef or not np.array_equal(r["labels"], ew):
                matches = False
                log(f"serve: request {k} differs from its cold solve (flow {r['flow']} vs {ef})")
            done += 1
            sess = k if reqs[k]["kind"] == "image" else reqs[k]["target"]
            if last[sess] == k:
                srv.close_session(sess_of[sess])
        inflight, queued, _ = srv.load()
        if inflight == 0 and queued == 0 and i < n:
            wait = arrival[i] - (gcfg.now_ms() - t0)
            if wait > 0:
                time.sleep(wait / 1e3)
    wall = gcfg.now_ms() - t0
    srv.close()
    px = sum(q["W"] * q["H"] for q in reqs)
    L = np.array([lat[k] for k in range(n)])
    offered = float(px / (arrival[-1] * 1e3)) if arrival[-1] > 0 else float("inf")
    return dict(qps=qps, requests=n, edits=sum(q["kind"] == "edit" for q in reqs),
                p50_ms=round(float(np.percentile(L, 50)), 3), p99_ms=round(float(np.percentile(L, 99)), 3),
                mean_ms=round(float(L.mean()), 3), served_mpx_s=round(px / (wall * 1e3), 1),
                offered_mpx_s=round(offered, 1), wall_ms=round(wall, 1), matches_cold=bool(matches))


def serve_sweep(args, opts):
    log(f"serving workload: {args.serve_requests} requests, pool of {args.serve_pool} images")
    pool, reqs, last = serve_workload(args, args.serve_requests)
    t0 = time.perf_counter()
    expected = serve_expected(pool, reqs)
    log(f"cold solves of every request: {time.perf_counter() - t0:.1f} s")
    points = []
    saturated = 0
    for qps in args.serve_qps:
        pt = serve_point(args, opts, pool, reqs, last, expected, qps)
        log(f"serve: {pt}")
        points.append(pt)
        if pt["served_mpx_s"] < 0.8 * pt["offered_mpx_s"]:
            saturated += 1
            if saturated >= 2:
                break
    for _, _, img, seeds in pool:
        gcfg.pinned_free(img)
        gcfg.pinned_free(seeds)
    px = sum(q["W"] * q["H"] for q in reqs)
    best = max(points, key=lambda p: p["served_mpx_s"])
    return dict(points=points, capacity_mpx_s=best["served_mpx_s"], capacity_qps=best["qps"],
                capacity_p99_ms=best["p99_ms"], matches_cold=all(p["matches_cold"] for p in points),
                requests_per_point=len(reqs), mean_request_mpx=round(px / len(reqs) / 1e6, 3),
                hbm_budget_gb=args.serve_budget_gb, label_format="packed bits")


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


# ---------------------------------------------------------------------------
# CPU baseline
def host_cores():
    """(threads used, nproc, OMP_NUM_THREADS): nproc as the affinity mask
    shows it, capped by OMP_NUM_THREADS where the environment sets the
    process's share of the host."""
    try:
        nproc = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        nproc = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    try:
        omp_n = int(omp) if omp else None
    except ValueError:
        omp_n = None
    return (min(nproc, omp_n) if omp_n else nproc), nproc, omp_n


def cpu_model():
    """The host's model-name string with its family/model/stepping numbers
    (a virtual machine may expose only a short name such as "AMD EPYC"; the
    numbers still identify the generation)."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if fields:
                        break  # the first processor's block is enough
                    continue
                k, v = (t.strip() for t in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "vendor_id") and k not in fields:
                    fields[k] = v
    except OSError:
        return None
    name = fields.get("model name")
    if not name:
        return None
    ids = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    return f"{name} ({', '.join(ids)})" if ids else name


def cpu_baseline(insts, flows_gpu, labels_gpu_fn, budget_s):
    """Reference-schedule restatement on the host cores over a bounded sample."""
    import oracle

    threads, nproc, omp_n = host_cores()
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
                      f"reference push-relabel schedule restated in C/OpenMP on {threads} threads",
               seconds=round(secs, 2), matches_gpu=bool(match), cpu_model=cpu_model(), nproc=nproc,
               omp_num_threads=omp_n, os_cpu_count=os.cpu_count())
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


# ---------------------------------------------------------------------------
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
    """One plain GC_* drop-in solve at C1 or C2, labels to the host (int32)."""
    if args.calibrate_copy:
        a, b = gcfg.DeviceBuffer(1 << 30), gcfg.DeviceBuffer(1 << 30)
        b.copy_from_device(a.ptr, 1 << 30)
        gcfg.device_synchronize()
        a.free()
        b.free()
    W, H, bufs = single_instance(args.single, args.seed)
    g = gcfg.Graph(W, H, opts)
    g.set_dataterms(bufs[0], bufs[1])
    g.set_edges(*bufs[2:6])
    d_label = gcfg.DeviceBuffer(4 * W * H)
    host = np.empty((H, W), np.int32)
    st = gcfg.Stats()
    g.build()
    g.optimize_device(d_label)
    times, dev_times = [], []
    for s in range(args.warmup + args.steps):
        gcfg.device_synchronize()
        t0 = time.perf_counter()
        g.build()
        rc = gcfg.lib().gcfg_graph_optimize(_VP(g.h), _VP(host.ctypes.data), ctypes.byref(st))
        if rc != 0:
            raise RuntimeError("gcfg_graph_optimize failed")
        t1 = time.perf_counter()
        g.build()
        std = g.optimize_device(d_label)
        gcfg.device_synchronize()
        t2 = time.perf_counter()
        if s >= args.warmup:
            times.append((t1 - t0) * 1e3)
            dev_times.append((t2 - t1) * 1e3)
        log(f"{args.single} solve {s}: {(t1 - t0) * 1e3:.2f} ms (labels to host), {(t2 - t1) * 1e3:.2f} ms "
            f"(labels in HBM), flow {st.flow} steps {st.steps} gr_rounds {st.gr_rounds}")
    ms, dms = float(np.mean(times)), float(np.mean(dev_times))
    out = dict(workload=args.single, W=W, H=H, ms=round(ms, 3), mpx_s=round(W * H / ms / 1e3, 2),
               device_resident_ms=round(dms, 3), device_resident_mpx_s=round(W * H / dms / 1e3, 2),
               flow=int(st.flow), steps=int(st.steps), gr_rounds=int(st.gr_rounds), pr_tiles=int(st.pr_tiles),
               gr_tiles=int(st.gr_tiles), sweeps_per_gr=opts.sweeps_per_gr, inner_iters=opts.inner_iters,
               gr_cap=opts.gr_cap, gr_cap_late=opts.gr_cap_late)
    # Event-timed solve (profile mode) for each kernel's time per solve and
    # its rate by the byte model of the batch roofline (relabel: REINIT and
    # relaxation visits and labels; push: tile passes).
    popts = gcfg.Options.from_buffer_copy(opts)
    popts.profile = 1
    gp = gcfg.Graph(W, H, popts)
    gp.set_dataterms(bufs[0], bufs[1])
    gp.set_edges(*bufs[2:6])
    for _ in range(2):  # the second solve launches exactly the steps the first needed
        gp.build()
        pst = gp.optimize_device(d_label)
    gp.close()
    gr_bytes = 1024 * max(1, opts.gr_block_w) * max(1, opts.gr_block_h) * (
        pst.gr_init_tiles * GR_INIT_BYTES_PER_NODE + (pst.gr_tiles - pst.gr_init_tiles) * GR_RELAX_BYTES_PER_NODE +
        (pst.gr_exact_tiles + pst.fin_items) * LABEL_BYTES_PER_NODE)
    pr_bytes = pst.pr_tiles * 1024 * PR_BYTES_PER_NODE

    def rate(ms_k, b):
        return dict(ms=round(ms_k, 4), bytes=int(b), gb_s=round(b / (ms_k * 1e-3) / 1e9, 1),
                    frac=round(b / (ms_k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)) if ms_k > 0 else None
    out["kernels"] = dict(k_grb=rate(pst.gr_kernel_ms, gr_bytes), k_prw=rate(pst.pr_kernel_ms, pr_bytes),
                          k_sched_ms=round(pst.sched_kernel_ms, 4), launches=int(pst.launches),
                          gr_items=int(pst.gr_tiles), gr_init_items=int(pst.gr_init_tiles),
                          gr_exact_items=int(pst.gr_exact_tiles), pr_items=int(pst.pr_tiles),
                          note="kernel time per solve includes the launches a step leaves empty")
    if not args.no_cpu:
        import oracle
        n = W * H
        h = [b.download(np.int32, n).reshape(H, W) for b in bufs]
        t0 = time.perf_counter()
        rf, rl = oracle.maxflow(W, H, h[0], h[1], edges=tuple(h[2:6]))
        out["oracle_s"] = round(time.perf_counter() - t0, 2)
        out["matches_oracle"] = bool(rf == st.flow and np.array_equal(rl, host) and std.flow == rf)
    g.close()
    print(json.dumps(out), flush=True)


def pmc_traffic(args, opts, kernel_prefix, launches_per_solve):
    """HBM bytes per launch of the dominant kernel from two rocprofv3 counter
    runs of this script (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each
    scaled by what the same counter reads for a 1 GiB device-to-device copy
    run in the same process. The child solves the batch several times (the
    timed solves of each label format and the event-timed ones); every solve
    arms its instances with one k_arm dispatch, so the child's solve count is
    the number of k_arm dispatches. A first solve may launch a few idle steps
    whose dispatches exit at once, so bytes per launch are the summed bytes
    over every dispatch / (solves x launches of one solve).
    Returns (traffic dict or None, note)."""
    import shutil
    import sqlite3
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    child = [sys.executable, os.path.abspath(__file__), "--no-cpu", "--no-serve", "--no-plain", "--no-pmc",
             "--no-unseen", "--steps", "1", "--warmup", "0", "--calibrate-copy", "--instances", str(args.instances),
             "--seed", str(args.seed), "--sweeps", str(opts.sweeps_per_gr), "--inner", str(opts.inner_iters),
             "--poll", str(opts.steps_per_poll), "--gr-block", f"{opts.gr_block_w}x{opts.gr_block_h}",
             "--gr-cap", str(opts.gr_cap), "--gr-cap-late", str(opts.gr_cap_late)]
    if opts.gr_waves_per_cu:
        child += ["--gr-waves", str(opts.gr_waves_per_cu)]
    if opts.pr_waves_per_cu:
        child += ["--pr-waves", str(opts.pr_waves_per_cu)]
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
        solves = sum(1 for n, _ in rows if "k_arm" in n)
        if calib <= 0 or not vals or not solves:
            return None, f"no {counter} samples for the copy, k_arm or {kernel_prefix}"
        scale = (1 << 20) / calib  # the copy moves 1 GiB = 2^20 KiB each way
        per_launch = sum(vals) / (solves * max(1, launches_per_solve))
        got[counter] = (per_launch * 1024 * scale, scale, len(vals), solves)
    fetch, fscale, n, solves = got["FETCH_SIZE"]
    write, wscale, _, _ = got["WRITE_SIZE"]
    return dict(bytes=int(fetch + write), fetch_bytes=int(fetch), write_bytes=int(write),
                fetch_scale=round(fscale, 4), write_scale=round(wscale, 4), dispatches=n, solves=solves), "ok"


def unseen_first_solve(args, opts):
    """A batch from another seed, solved once on a fresh handle (the first
    step chunk is the default length, no hint from an earlier solve; tuned
    options of this workload class), labels packed to the host; every flow
    and label map checked against the oracle."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor

    insts, dtb, _ = make_batch(args.instances, args.seed + 1)
    px = sum(it.W * it.H for it in insts)
    b = gcfg.Batch(batch_items(insts), opts)
    bs = BitsSolver(b)
    gcfg.device_synchronize()
    t0 = time.perf_counter()
    flows = bs.solve().copy()
    ms = (time.perf_counter() - t0) * 1e3
    steps = int(bs.st.steps)
    words = [bs.words(i).copy() for i in range(len(insts))]
    bs.solve()
    second_ms = bs.st.ms
    b.close()
    threads, _, _ = host_cores()

    def check(k):
        it = insts[k]
        pos, neg, up, down, left, right = host_links(it)
        rf, rl = oracle.maxflow(it.W, it.H, pos, neg, edges=(up, down, left, right))
        return rf == int(flows[k]) and np.array_equal(gcfg.pack_bits(rl), words[k])

    with ThreadPoolExecutor(max_workers=threads) as ex:
        ok = all(ex.map(check, range(len(insts))))
    dtb.close()
    return dict(seed=args.seed + 1, ms=round(ms, 3), mpx_s=round(px / ms / 1e3, 1), steps=steps,
                second_solve_ms=round(second_ms, 3), matches_oracle=bool(ok), total_mpx=round(px / 1e6, 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--serve-qps", type=float, nargs="+", default=[250, 500, 1000, 2000, 4000, 8000, 16000])
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-pool", type=int, default=96, help="distinct images the served requests draw from")
    ap.add_argument("--serve-budget-gb", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="time budget of the CPU baseline's batch sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-unseen", action="store_true")
    ap.add_argument("--sweeps", type=int, default=0)
    ap.add_argument("--inner", type=int, default=0)
    ap.add_argument("--poll", type=int, default=0)
    ap.add_argument("--gr-block", default=None, help="global-relabel block WxH in tiles, e.g. 1x4")
    ap.add_argument("--gr-cap", type=int, default=None, help="distance cap of the first global relabel")
    ap.add_argument("--gr-cap-late", type=int, default=None, help="distance cap of the next three global relabels")
    ap.add_argument("--gr-waves", type=int, default=None, help="relabel launch wavefronts per CU")
    ap.add_argument("--pr-waves", type=int, default=None, help="push/relabel launch wavefronts per CU")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter runs for roofline.traffic")
    ap.add_argument("--tune-cache", default=os.environ.get("GCFG_TUNE_CACHE",
                                                           os.path.expanduser("~/.cache/gcfg/tune.txt")))
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
    if args.gr_cap_late is not None:
        opts.gr_cap_late = args.gr_cap_late
    if args.gr_waves:
        opts.gr_waves_per_cu = args.gr_waves
    if args.pr_waves:
        opts.pr_waves_per_cu = args.pr_waves
    if args.extras:
        extras(args)
        return
    if args.single:
        single(args, opts)
        return

    log(f"generating {args.instances} instances + batched device data term")
    insts, dtb, dt_ms = make_batch(args.instances, args.seed)
    total_px = sum(it.W * it.H for it in insts)
    dt_gbs = DATATERM_BYTES_PER_PX * total_px / (dt_ms * 1e-3) / 1e9
    log(f"batch: {total_px / 1e6:.2f} Mpx, data term {dt_ms:.3f} ms ({dt_gbs:.0f} GB/s at 30 B/px)")

    if args.calibrate_copy:
        a, b = gcfg.DeviceBuffer(1 << 30), gcfg.DeviceBuffer(1 << 30)
        b.copy_from_device(a.ptr, 1 << 30)
        gcfg.device_synchronize()
        a.free()
        b.free()
    batch = gcfg.Batch(batch_items(insts), opts)
    tuned = None
    if not args.no_autotune and not (args.sweeps or args.inner or args.poll or args.gr_block or
                                     args.gr_cap is not None or args.gr_cap_late is not None):
        # One cache entry per workload class (instance count), not per seed.
        key = f"batch:ellipse:n{args.instances}"
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

    # Headline: labels on the host, packed one bit per pixel.
    bits = BitsSolver(batch)
    ms, _ = time_steps(bits.solve, args.steps, args.warmup, "batched (labels on host, packed)")
    value = total_px / ms / 1e3
    flows_b = bits.flows.copy()
    st = bits.st
    words_b = [bits.words(i).copy() for i in range(len(insts))]
    # The same solves with int32 labels to the host, and with labels left in HBM.
    i32 = Int32Solver(batch)
    ms_i32, _ = time_steps(i32.solve, max(3, args.steps // 4), 1, "batched (int32 labels to host)")
    assert np.array_equal(i32.flows, flows_b), "int32-label and packed-label solves differ"
    for k in range(len(insts)):
        assert np.array_equal(gcfg.pack_bits(i32.host[k]), words_b[k]), f"packed labels differ on instance {k}"
    dev_flows = {}

    def dev_solve():
        dev_flows["f"] = batch.solve(labels_to_host=False)[0]
    ms_dev, _ = time_steps(dev_solve, args.steps, 1, "batched (labels in HBM)")
    assert np.array_equal(dev_flows["f"], flows_b)

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
                    solve_alg_frac=round(solve_bytes / (ms_dev * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    compulsory_frac=round(COMPULSORY_BYTES_PER_PX * total_px / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                          5),
                    pr_tiles=int(pst.pr_tiles), gr_tiles=int(pst.gr_tiles),
                    gr_init_tiles=int(pst.gr_init_tiles), gr_exact_tiles=int(pst.gr_exact_tiles),
                    fin_items=int(pst.fin_items), arena_tiles=int(n_tiles))
    assert np.array_equal(flows_p, flows_b), "profiled solve differs"
    pbatch.close()
    if not args.no_pmc:
        name = {"k_prw": "k_prw<", "k_pr": "k_pr(", "k_grb": "k_grb<", "k_gr": "k_gr(", "k_buildw": "k_buildw<",
                "k_build": "k_build("}[dom]
        traffic, note = pmc_traffic(args, opts, name, dlaunch)
        if traffic:
            roofline["traffic"] = traffic["bytes"]
            roofline["traffic_detail"] = traffic
            log(f"traffic of {dom}: {traffic}")
        else:
            roofline["traffic_note"] = note
            log(f"traffic not measured: {note}")

    extra = dict(batched_labels_int32_to_host_mpx_s=round(total_px / ms_i32 / 1e3, 1),
                 batched_labels_int32_to_host_ms=round(ms_i32, 3),
                 batched_device_resident_mpx_s=round(total_px / ms_dev / 1e3, 1),
                 batched_device_resident_ms=round(ms_dev, 3))
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
                                                                   "gr_cap", "gr_cap_late")}
        graphs, hosts = [], []
        for it in insts:
            g = gcfg.Graph(it.W, it.H, popts)
            g.set_dataterms(it.bufs[0], it.bufs[1])
            g.set_edges(*it.bufs[2:6])
            g.build()
            g.optimize_device(it.d_label)
            graphs.append(g)
            hosts.append(np.empty((it.H, it.W), np.int32))
        run_plain(insts, graphs, hosts)  # warm-up pass
        plain_ms, flows_plain = run_plain(insts, graphs, hosts)
        log(f"plain pass (int32 labels to host): {plain_ms:.1f} ms, {total_px / plain_ms / 1e3:.2f} Mpx/s")
        assert np.array_equal(flows_plain, flows_b), "plain and batched flows differ"
        for k in range(len(insts)):
            assert np.array_equal(gcfg.pack_bits(hosts[k]), words_b[k]), f"plain and batched labels differ ({k})"
        t0 = time.perf_counter()
        for it, g in zip(insts, graphs):
            g.build()
            g.optimize_device(it.d_label)
        gcfg.device_synchronize()
        plain_dev_ms = (time.perf_counter() - t0) * 1e3
        extra["plain_mpx_s"] = round(total_px / plain_ms / 1e3, 3)
        extra["plain_ms"] = round(plain_ms, 2)
        extra["plain_device_resident_mpx_s"] = round(total_px / plain_dev_ms / 1e3, 3)
        extra["speedup_vs_plain"] = round(value / (total_px / plain_ms / 1e3), 2)
        extra["speedup_vs_plain_device_resident"] = round(ms_dev and (plain_dev_ms / ms_dev), 2)
        for g in graphs:
            g.close()

    if not args.no_unseen:
        log("first solve of an unseen batch")
        extra["first_solve_unseen"] = unseen_first_solve(args, opts)
        log(f"unseen: {extra['first_solve_unseen']}")

    if not args.no_serve:
        sv = serve_sweep(args, opts)
        extra["serve"] = sv
        if "plain_mpx_s" in extra:
            extra["served_speedup_vs_plain"] = round(sv["capacity_mpx_s"] / extra["plain_mpx_s"], 2)
        log("video loop: warm frames on a session vs cold images")
        extra["video"] = video(args, opts)
        log(f"video: {extra['video']}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (reference schedule restated on the host)")
        cpu = cpu_baseline(insts, flows_b, lambda k: gcfg.unpack_bits(words_b[k], insts[k].W, insts[k].H),
                           args.cpu_seconds)
        log(f"cpu: {cpu}")
    batch.close()

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpx/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"batch of {args.instances} ellipse images, area log-uniform 128^2..1024^2, "
                               f"aspect 1:2..2:1, 16-bin colour data term, lambda 50, cold solves, labels to the "
                               f"host packed one bit per pixel",
                   "instances": args.instances, "total_mpx": round(total_px / 1e6, 3), "seed": args.seed,
                   "sweeps_per_gr": opts.sweeps_per_gr, "inner_iters": opts.inner_iters,
                   "steps_per_poll": opts.steps_per_poll, "gr_rounds": opts.gr_rounds,
                   "gr_block": f"{opts.gr_block_w}x{opts.gr_block_h}", "gr_cap": opts.gr_cap,
                   "gr_cap_late": opts.gr_cap_late,
                   "gr_waves_per_cu": opts.gr_waves_per_cu, "autotuned": tuned is not None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "dataterm": dict(ms=round(dt_ms, 3), mpx_s=round(total_px / dt_ms / 1e3, 1), alg_gb_s=round(dt_gbs, 1),
                         frac=round(dt_gbs / HBM_PEAK_GBS, 4), bytes_per_px=DATATERM_BYTES_PER_PX,
                         note="gcfg_dt_batch_compute wall time (three launches and a synchronisation)"),
        "scheduler_steps": int(st.steps),
    }
    line.update(extra)
    dtb.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
