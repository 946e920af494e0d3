#!/usr/bin/env python3
"""Throughput of the HIP descale path on one MI355X.

Workload (one step): Debicubic(b=0, c=0.5) of 64 YUV420P16 frames
1920x1080 -> 1280x720 (chroma 960x540 -> 640x360, limited range, padding 3),
16-bit output, all three planes in one dsmvc_hip_inverse_2d_planes call.  In
the default solve mode that call takes the fused 2D route (one kernel:
X = M_v Y M_h^T with the planes' explicit operators, F32 accumulation, no
intermediate in memory); the two-pass route's per-axis kernels (u16 rows ->
F32 intermediate -> u16 columns) are timed beside it.
Frames are synthetic: a seeded 1280x720 U[0,65535] image blurred by a
Gaussian (sigma 1.5 px), upscaled by the forward bicubic operator, plus
N(0, 64^2) noise, clamped to 16 bits; inputs already in HBM, rotating through
more than 1 GiB of distinct buffers so the 256 MB Infinity Cache cannot serve
them.

Prints one JSON line: frames/s (whole job), the per-kernel roofline of the
dominant kernel (HIP events on the launch stream over the timed region) with
the frame-level fraction beside it, and the reference CPU engine's automatic
SIMD path on every host core (its scalar path on one core beside it) on a
bounded sample of the same workload on this host.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "vapoursynth-descalemvc_amd"))

METRIC = ("1080p→720p F32 frames/s + %HBM peak; mixed-size solves/s + p99 latency")
HBM_PEAK_GBPS = 8000.0
FRAMES = 64
SRC_W, SRC_H, DST_W, DST_H = 1920, 1080, 1280, 720
FRAME_BYTES_YUV16 = 2 * (SRC_W * SRC_H + 2 * (SRC_W // 2) * (SRC_H // 2)) \
    + 2 * (DST_W * DST_H + 2 * (DST_W // 2) * (DST_H // 2))          # 8,985,600
FRAME_BYTES_GRAY_F32 = 4 * (SRC_W * SRC_H + DST_W * DST_H)            # 11,980,800


def plane_requests(dh, sw, sh, dw, dh_, chroma, kernel, b, c):
    kw = dict(kernel=kernel, b=b, c=c, border=dh.SYMMETRIC, f64_mode=0)
    if not chroma:
        return (dh.AxisRequest(sw, dw, float(dw), 0.0, **kw),
                dh.AxisRequest(sh, dh_, float(dh_), 0.0, **kw))
    cw, ch = sw >> 1, sh >> 1
    # reference src/vs_plugin.cpp:656-705 (4:2:0 chroma requests)
    return (dh.AxisRequest(cw, dw >> 1, float(dw) * cw / sw, 0.25 - 0.25 * float(dw) / float(sw),
                           **kw),
            dh.AxisRequest(ch, dh_ >> 1, float(dh_) * ch / sh, 0.0, **kw))


def synthetic_frames(torch, dh, count, seed, device):
    """Config-2 frames as uint16 planes [count, H, W] (Y, U, V)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    planes = []
    for index, (sw, sh, dw, dh_) in enumerate([(SRC_W, SRC_H, DST_W, DST_H),
                                               (SRC_W // 2, SRC_H // 2, DST_W // 2, DST_H // 2),
                                               (SRC_W // 2, SRC_H // 2, DST_W // 2, DST_H // 2)]):
        low = torch.rand((count, 1, dh_, dw), generator=gen, device=device) * 65535.0
        radius = 5
        t = torch.arange(-radius, radius + 1, device=device, dtype=torch.float32)
        g = torch.exp(-0.5 * (t / 1.5) ** 2)
        g = g / g.sum()
        low = torch.nn.functional.conv2d(torch.nn.functional.pad(low, (radius, radius, 0, 0), mode="reflect"),
                                         g.view(1, 1, 1, -1))
        low = torch.nn.functional.conv2d(torch.nn.functional.pad(low, (0, 0, radius, radius), mode="reflect"),
                                         g.view(1, 1, -1, 1))[:, 0]
        hreq, vreq = plane_requests(dh, SRC_W, SRC_H, DST_W, DST_H, index > 0, dh.BICUBIC, 0.0, 0.5)
        A_h = forward_operator(torch, dh.Plan(hreq), device)      # [sw, dw]
        A_v = forward_operator(torch, dh.Plan(vreq), device)      # [sh, dh]
        up = torch.matmul(torch.matmul(A_v, low), A_h.t().contiguous())
        up = up + torch.randn(up.shape, generator=gen, device=device) * 64.0
        planes.append(up.round().clamp(0, 65535).to(torch.int32))
    return planes


def forward_operator(torch, plan, device):
    """Dense A (source x destination) from the plan's CSR of A^T."""
    a = plan.arrays()
    A = np.zeros((plan.source_size, plan.destination_size), np.float32)
    off, idx, w = a["offsets"], a["indices"], a["weights"]
    for i in range(plan.destination_size):
        A[idx[off[i]:off[i + 1]], i] = w[off[i]:off[i + 1]]
    return torch.from_numpy(A).to(device)


def host_cores():
    """Host threads this process may use, at most 16 (one GPU's share of the host)."""
    try:
        count = len(os.sched_getaffinity(0))
    except AttributeError:
        count = os.cpu_count() or 1
    return max(1, min(16, count))


def cpu_baseline(planes_host, seconds=12.0, scalar_seconds=4.0):
    """The reference engine (oracle/_ref, built from the reference's own
    sources) on this host: CpuExecutor with CpuPath::automatic (AVX-512 where
    the host has it, else AVX2; reference src/cpu_executor.cpp:787-799), one
    worker thread per host core, each taking whole YUV420P16 frames and
    calling inverse_2d_u16 per plane on one shared executor -- the way
    VapourSynth's fmParallel drives the plugin.  The scalar path (opt=1) on
    one thread is timed beside it.  Falls back to the oracle's C restatement
    (one thread, "port") when the reference build is absent."""
    import threading
    sys.path.insert(0, str(ROOT / "tests"))
    from support import Oracle, Reference, reference_available
    from workloads import chroma420_requests, integer_conversion, luma_requests
    lreq = luma_requests(SRC_W, SRC_H, DST_W, DST_H, 1, b=0.0, c=0.5)
    creq = chroma420_requests(SRC_W, SRC_H, DST_W, DST_H, 1, b=0.0, c=0.5)
    convs = [integer_conversion(16, False, False), integer_conversion(16, True, False)]
    count = planes_host[0].shape[0]

    def timed(run, threads, limit):
        stop = time.perf_counter() + limit
        done = [0] * threads

        def worker(t):
            k = t
            while time.perf_counter() < stop:
                f = k % count
                for p in range(3):
                    h, v = plans[0 if p == 0 else 1]
                    run(h, v, planes_host[p][f], convs[0 if p == 0 else 1])
                done[t] += 1
                k += threads
        start = time.perf_counter()
        pool = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
        for th in pool:
            th.start()
        for th in pool:
            th.join()
        elapsed = time.perf_counter() - start
        return sum(done), elapsed

    if not reference_available():
        eng = Oracle()
        plans = [(eng.build(lreq[0]), eng.build(lreq[1])), (eng.build(creq[0]), eng.build(creq[1]))]
        frames, elapsed = timed(lambda h, v, x, c: eng.inverse_2d_int(h, v, x, c), 1, seconds)
        return {"value": round(frames / elapsed, 3), "unit": "frames/s", "cores": 1,
                "kind": "port", "sample": f"{frames} YUV420P16 frames through the oracle's C "
                                          f"restatement, one thread, {elapsed:.1f} s"}
    eng = Reference()
    plans = [(eng.build(lreq[0]), eng.build(lreq[1])), (eng.build(creq[0]), eng.build(creq[1]))]
    threads = host_cores()
    path_name = eng.executor_name(Reference.CPU_AUTOMATIC)
    auto = lambda h, v, x, c: eng.inverse_2d_int(h, v, x, c, path=Reference.CPU_AUTOMATIC)  # noqa: E731
    scalar = lambda h, v, x, c: eng.inverse_2d_int(h, v, x, c, path=Reference.CPU_SCALAR)  # noqa: E731
    timed(auto, threads, 1.0)   # warm: plan packing, worker pool start
    frames, elapsed = timed(auto, threads, seconds)
    sframes, selapsed = timed(scalar, 1, scalar_seconds)
    return {"value": round(frames / elapsed, 3), "unit": "frames/s", "cores": threads,
            "kind": "reference",
            "sample": f"{frames} YUV420P16 1920x1080->1280x720 Debicubic frames in "
                      f"{elapsed:.1f} s: reference CpuExecutor CpuPath::automatic "
                      f"({path_name}), {threads} threads each taking whole frames, "
                      f"inverse_2d_u16 per plane",
            "scalar_one_thread": {"value": round(sframes / selapsed, 3), "unit": "frames/s",
                                  "cores": 1, "sample": f"{sframes} frames, CpuPath::scalar "
                                                        f"(opt=1), {selapsed:.1f} s"}}


def cpu_calibration(seconds=3.0):
    """CPU-only: the reference engine's scalar and automatic paths at 1..N
    threads on the GRAYS F32 1080p -> 720p workloads (Debilinear, and
    Debicubic) and the YUV420P16 Debicubic workload, plus the ratio of the
    oracle's C restatement to the reference scalar path on one thread.  Each
    worker takes whole frames; frames/s per entry."""
    import threading
    sys.path.insert(0, str(ROOT / "tests"))
    from support import BICUBIC, BILINEAR, Oracle, Reference, reference_available
    from workloads import chroma420_requests, integer_conversion, luma_requests
    if not reference_available():
        raise SystemExit("bench.py --cpu-calibration needs oracle/_ref (the reference build)")
    ref, orc = Reference(), Oracle()
    rng = np.random.default_rng(1)
    gray = rng.uniform(0, 1, (8, SRC_H, SRC_W)).astype(np.float32)
    yuv = [rng.integers(0, 65536, (4, SRC_H >> (p > 0), SRC_W >> (p > 0))).astype(np.uint16)
           for p in range(3)]
    convs = [integer_conversion(16, False, False), integer_conversion(16, True, False)]

    def rate(work, threads):
        stop = time.perf_counter() + seconds
        done = [0] * threads

        def worker(t):
            k = t
            while time.perf_counter() < stop:
                work(k)
                done[t] += 1
                k += threads
        start = time.perf_counter()
        pool = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
        for th in pool:
            th.start()
        for th in pool:
            th.join()
        return round(sum(done) / (time.perf_counter() - start), 2)

    out = {"host_threads": host_cores(), "seconds_per_entry": seconds,
           "automatic_resolves_to": ref.executor_name(Reference.CPU_AUTOMATIC)}
    counts = sorted({1, 2, 4, host_cores()})
    for name, kernel in (("grays_f32_debilinear", BILINEAR), ("grays_f32_debicubic", BICUBIC)):
        h, v = luma_requests(SRC_W, SRC_H, DST_W, DST_H, kernel, b=0.0, c=0.5)
        rh, rv = ref.build(h), ref.build(v)
        entry = {"scalar_1": rate(lambda k: ref.inverse_2d(rh, rv, gray[k % 8],
                                                           path=Reference.CPU_SCALAR), 1)}
        for n in counts:
            entry[f"automatic_{n}"] = rate(lambda k: ref.inverse_2d(
                rh, rv, gray[k % 8], path=Reference.CPU_AUTOMATIC), n)
        out[name] = entry
    lreq = luma_requests(SRC_W, SRC_H, DST_W, DST_H, BICUBIC, b=0.0, c=0.5)
    creq = chroma420_requests(SRC_W, SRC_H, DST_W, DST_H, BICUBIC, b=0.0, c=0.5)

    def yuv_frame(eng, plans, **kw):
        def work(k):
            for p in range(3):
                h, v = plans[0 if p == 0 else 1]
                eng.inverse_2d_int(h, v, yuv[p][k % 4], convs[0 if p == 0 else 1], **kw)
        return work
    rplans = [(ref.build(lreq[0]), ref.build(lreq[1])), (ref.build(creq[0]), ref.build(creq[1]))]
    oplans = [(orc.build(lreq[0]), orc.build(lreq[1])), (orc.build(creq[0]), orc.build(creq[1]))]
    entry = {"scalar_1": rate(yuv_frame(ref, rplans, path=Reference.CPU_SCALAR), 1)}
    for n in counts:
        entry[f"automatic_{n}"] = rate(yuv_frame(ref, rplans, path=Reference.CPU_AUTOMATIC), n)
    out["yuv420p16_debicubic"] = entry
    port = rate(yuv_frame(orc, oplans), 1)
    out["restatement_scalar_1_yuv420p16"] = port
    out["restatement_over_reference_scalar"] = round(port / entry["scalar_1"], 3)
    return out


def batched_scan(torch, dh, dev, frame, repeats):
    """Mixed-size solves: one 1920x1080 F32 frame descaled (Mitchell, b = c =
    1/3) to every height 500..1000 with width round(16h/9) -- 501 plan pairs,
    1002 axis solves per pass.  The multi-plan batch (one launch per
    direction and tap width over device job tables, replayed from a HIP
    graph) against one inverse_2d call per plan pair, both asynchronous on
    one stream and timed with HIP events."""
    heights = list(range(500, 1001))
    kw = dict(kernel=dh.BICUBIC, b=1.0 / 3.0, c=1.0 / 3.0, border=dh.SYMMETRIC, f64_mode=0)
    widths = [int(math.floor(h * 16.0 / 9.0 + 0.5)) for h in heights]
    requests = []
    for h, w in zip(heights, widths):
        requests += [dh.AxisRequest(SRC_W, w, float(w), 0.0, **kw),
                     dh.AxisRequest(SRC_H, h, float(h), 0.0, **kw)]
    # Host planning of the 1002 axis plans from an empty cache: one at a time,
    # then on the planner's host threads (dsmvc_hip_plan_build_many).
    dh.clear_planner_caches()
    t0 = time.perf_counter()
    serial = [dh.Plan(r) for r in requests]
    plan_serial_s = time.perf_counter() - t0
    del serial
    dh.clear_planner_caches()
    t0 = time.perf_counter()
    plans = dh.build_plans(requests)
    plan_s = time.perf_counter() - t0
    jobs, pair_bytes = [], 0
    for k, (h, w) in enumerate(zip(heights, widths)):
        jobs.append((plans[2 * k], plans[2 * k + 1], frame, torch.empty((h, w), device=dev)))
        pair_bytes += 4 * (SRC_W * SRC_H + w * h)
    ex = dh.Executor(0, "auto")
    t0 = time.perf_counter()
    batch = dh.Batch(ex, jobs, graph=True)
    setup_s = time.perf_counter() - t0
    info = batch.info()
    reference = [torch.empty_like(j[3]) for j in jobs]
    stream = torch.cuda.current_stream()

    def per_plan(outs):
        for (hp, vp, x, _), out in zip(jobs, outs):
            ex.inverse_2d(hp, vp, x, out=out, asynchronous=True)

    batch.run()
    per_plan(reference)
    torch.cuda.synchronize()
    diff = max(float((j[3] - r).abs().max()) for j, r in zip(jobs, reference))

    def timed(fn):
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(repeats):
            fn()
        stop.record(stream)
        stop.synchronize()
        return start.elapsed_time(stop) / 1e3 / repeats

    batch_s = timed(lambda: batch.run(stream, asynchronous=True))
    plain_s = timed(lambda: per_plan(reference))
    solves = 2 * len(heights)
    return {
        "workload": "Mitchell 1920x1080 F32 -> h x round(16h/9), h = 500..1000 (501 plan pairs)",
        "axis_solves_per_pass": solves,
        "batched": {"solves_per_s": round(solves / batch_s, 1),
                    "pairs_per_s": round(len(heights) / batch_s, 1),
                    "ms_per_pass": round(batch_s * 1e3, 3),
                    "achieved_GBps": round(pair_bytes / batch_s / 1e9, 1),
                    "frac_of_8TBps": round(pair_bytes / batch_s / 1e9 / HBM_PEAK_GBPS, 4),
                    "rate_note": "algorithmic rate: the one 1920x1080 source frame is counted "
                                 "once per pair, and it stays cache-resident across the pairs, "
                                 "so this is not HBM traffic",
                    "launches_per_pass": info["launches"], "graph": bool(info["graph"]),
                    "items_per_pass": info["items"], "fused_jobs": info["fused_jobs"]},
        "one_plan_per_launch": {"solves_per_s": round(solves / plain_s, 1),
                                "ms_per_pass": round(plain_s * 1e3, 3)},
        "speedup": round(plain_s / batch_s, 2),
        "max_abs_diff_batched_vs_per_plan": diff,
        "algorithmic_bytes_per_pass": pair_bytes,
        "host_plan_build_s": round(plan_s, 3),
        "host_plan_build_s_one_at_a_time": round(plan_serial_s, 3),
        "batch_setup_s": round(setup_s, 3),
        "repeats": repeats,
    }


SAT_REQUESTS = 400


def served_stream(torch, dh, dev, seconds, rate=400.0, seed=5):
    """Mixed-size serving: a seeded Poisson request stream (mean `rate`
    requests/s for `seconds`) of GRAYS / YUV420PS frames from sources
    3840x2160, 1920x1080 and 1280x720, target height uniform in
    [0.5, 0.95] x source height (16:9 width, even sizes), kernel drawn from
    {bilinear, bicubic(0, 0.5), lanczos3, spline36} and f64mode from {0, 1, 2},
    through the continuous-batching request loop.  Latency is submit ->
    completion observed by the loop; a client thread submits on schedule."""
    import threading
    rng = np.random.default_rng(seed)
    sizes = [(3840, 2160), (1920, 1080), (1280, 720)]
    kernels = [(dh.BILINEAR, 3, 0.0, 0.5), (dh.BICUBIC, 3, 0.0, 0.5), (dh.LANCZOS, 3, 0.0, 0.5),
               (dh.SPLINE36, 3, 0.0, 0.5)]
    # Source frames: smoothed seeded noise in [0, 1], two frames per size.
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    pool = {}
    for (w, h) in sizes:
        frames = []
        for _ in range(2):
            planes = []
            for (pw, ph) in [(w, h), (w // 2, h // 2), (w // 2, h // 2)]:
                x = torch.rand((1, 1, ph, pw), generator=gen, device=dev)
                x = torch.nn.functional.avg_pool2d(x, 3, stride=1, padding=1,
                                                   count_include_pad=False)[0, 0]
                planes.append(x.contiguous())
            frames.append(planes)
        pool[(w, h)] = frames
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    reqs = []
    for k, t in enumerate(arrivals):
        w, h = sizes[rng.integers(3)]
        dh_ = int(h * rng.uniform(0.5, 0.95)) & ~1
        dw = int(math.floor(dh_ * 16.0 / 9.0 + 0.5))
        dw += dw & 1
        yuv = bool(rng.integers(2))
        src = pool[(w, h)][k % 2]
        ins = src if yuv else src[:1]
        outs = [torch.empty((dh_, dw), device=dev)]
        if yuv:
            outs += [torch.empty((dh_ // 2, dw // 2), device=dev) for _ in range(2)]
        reqs.append((float(t), ins, outs, kernels[rng.integers(4)], int(rng.integers(3))))
    torch.cuda.synchronize()
    ex = dh.Executor(0, "auto")
    server = dh.Server(ex, memory_budget=8 << 30, slots=4)
    ids = [None] * len(reqs)
    lag = np.zeros(len(reqs))
    submit_s = np.zeros(len(reqs))
    start = time.perf_counter()

    def client():
        for k, (t, ins, outs, (kernel, taps, b, c), mode) in enumerate(reqs):
            delay = t - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            t_call = time.perf_counter()
            lag[k] = (t_call - start) - t
            ids[k] = server.submit(ins, outs, kernel=kernel, taps=taps, b=b, c=c, f64_mode=mode)
            submit_s[k] = time.perf_counter() - t_call

    thread = threading.Thread(target=client)
    thread.start()
    thread.join()
    lat = []
    for rid in ids:
        t0, t1 = server.wait(rid)
        lat.append(t1 - t0)
    elapsed = time.perf_counter() - start
    stats = server.stats()
    server.close()
    cache = ex.plan_cache_stats()
    lat = np.array(lat) * 1e3
    planes = sum(len(r[1]) for r in reqs)

    # Per-plane plans of every request, as the loop builds them (chroma per
    # the reference's 4:2:0 rules), to split latency by route and to run the
    # plain executor on the same work.
    def plane_plans(ins, outs, kernel, taps, b, c, mode):
        (sh, sw), (dh_, dw) = ins[0].shape, outs[0].shape
        kw = dict(kernel=kernel, taps=taps, b=b, c=c, blur=1.0, border=dh.SYMMETRIC, f64_mode=mode)
        luma = (dh.Plan(dh.AxisRequest(sw, dw, float(dw), 0.0, **kw)),
                dh.Plan(dh.AxisRequest(sh, dh_, float(dh_), 0.0, **kw)))
        if len(ins) == 1:
            return [luma]
        cw, ch = sw >> 1, sh >> 1
        chroma = (dh.Plan(dh.AxisRequest(cw, dw >> 1, dw * cw / sw, 0.25 - 0.25 * dw / sw, **kw)),
                  dh.Plan(dh.AxisRequest(ch, dh_ >> 1, dh_ * ch / sh, 0.0, **kw)))
        return [luma, chroma, chroma]

    plans = [plane_plans(ins, outs, *kern, mode) for (_, ins, outs, kern, mode) in reqs]
    f64 = np.array([any(p.requires_float64 for pair in pp for p in pair) for pp in plans])

    def pct(values):
        if len(values) == 0:
            return None
        return {"n": int(len(values)), "p50": round(float(np.percentile(values, 50)), 3),
                "p99": round(float(np.percentile(values, 99)), 3)}

    # Saturation: the first SAT requests of the stream (whose plans fit the
    # planner cache) submitted back to back, against the plain executor one
    # 2D plane per call on the same plans; both warmed once, best of two.
    sat = reqs[:SAT_REQUESTS]
    sat_plans = plans[:SAT_REQUESTS]
    sat_planes = sum(len(r[1]) for r in sat)

    def served_pass():
        # Operator images built at submit (the first, untimed-best pass pays
        # for them), as the plain executor builds them on its first call.
        server = dh.Server(ex, memory_budget=8 << 30, slots=4, fused="always")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sat_ids = [server.submit(ins, outs, kernel=k[0], taps=k[1], b=k[2], c=k[3], f64_mode=mode)
                   for (_, ins, outs, k, mode) in sat]
        t1 = time.perf_counter()
        for rid in sat_ids:
            server.wait(rid)
        t2 = time.perf_counter()
        sat_stats.update(server.stats())
        server.close()
        return t2 - t0, t1 - t0

    def plain_pass():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for (_, ins, outs, _, _), pp in zip(sat, sat_plans):
            for x, o, (hp, vp) in zip(ins, outs, pp):
                ex.inverse_2d(hp, vp, x, out=o)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    sat_stats = {}
    served_s, submit_s_sat = min(served_pass() for _ in range(3))
    plain_s = min(plain_pass() for _ in range(3))
    return {
        "workload": "Poisson %.0f req/s, GRAYS/YUV420PS from 3840x2160/1920x1080/1280x720, "
                    "h ~ U[0.5,0.95]H, 4 kernels, f64mode 0/1/2" % rate,
        "requests": len(reqs), "seconds": round(elapsed, 3),
        "requests_per_s": round(len(reqs) / elapsed, 1),
        "solves_per_s": round(2 * planes / elapsed, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
        "latency_ms_float32_requests": pct(lat[~f64]),
        "latency_ms_float64_requests": pct(lat[f64]),
        "client_submit_lag_ms_p99": round(float(np.percentile(lag, 99)) * 1e3, 3),
        "submit_call_ms": {"p50": round(float(np.percentile(submit_s, 50)) * 1e3, 3),
                           "p99": round(float(np.percentile(submit_s, 99)) * 1e3, 3)},
        "plan_cache": {k: int(cache[k]) for k in ("entries", "bytes", "hits", "misses",
                                                 "evictions")},
        "rounds": stats["rounds"], "largest_round": stats["largest_round"],
        "table_planes": stats["table_planes"], "fallback_planes": stats["fallback_planes"],
        "fused_planes": stats["fused_planes"],
        "memory_high_water_bytes": stats["memory_high_water"],
        "saturation": {"requests": len(sat), "server_fused": "always",
                       "fused_planes_per_pass": sat_stats.get("fused_planes"),
                       "table_planes_per_pass": sat_stats.get("table_planes"),
                       "served_solves_per_s": round(2 * sat_planes / served_s, 1),
                       "plain_one_plane_per_call_solves_per_s": round(2 * sat_planes / plain_s, 1),
                       "speedup": round(plain_s / served_s, 2),
                       "served_pass_ms": round(served_s * 1e3, 3),
                       "served_submit_ms": round(submit_s_sat * 1e3, 3),
                       "plain_pass_ms": round(plain_s * 1e3, 3),
                       "note": "the stream's first requests submitted at once (plans cached); "
                               "plain = dsmvc_hip_inverse_2d per plane, synchronous, same plans; "
                               "best of 3 warm passes each"},
    }


TRAFFIC_FILE = ROOT / "profiles" / "yuv420p16_debicubic_1080p_720p_traffic.json"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE
    profile, each counter scaled by its copy-kernel calibration factor."""
    try:
        data = json.loads(TRAFFIC_FILE.read_text())
        entry = data["kernels"][kernel]
        cal = data["calibration"]
    except (OSError, KeyError, ValueError):
        return None, None
    total = (entry["fetch_kib"] * cal["fetch_factor"] + entry["write_kib"] * cal["write_factor"]) * 1024
    return int(total), str(TRAFFIC_FILE.relative_to(ROOT))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--solve", default="auto", choices=["auto", "exact", "chunked", "windowed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gray", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the mixed-size multi-plan scan (501 plan pairs)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served Poisson request stream")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--cpu-calibration", action="store_true",
                    help="CPU only: time the reference engine's scalar and automatic paths at "
                         "1..N threads and the oracle restatement, print JSON, exit")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="run one 256 MiB device copy before the timed region (counter calibration)")
    args = ap.parse_args()
    if args.cpu_calibration:
        print(json.dumps(cpu_calibration()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the path shards by frame with no collectives)", file=sys.stderr)
        sys.exit(2)

    import torch
    import dsmvc_hip as dh
    from dsmvc_hip import _lib

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ex = dh.Executor(0, args.solve)
    lib = _lib.load()
    stream = torch.cuda.current_stream()
    sptr = ctypes.c_void_p(stream.cuda_stream)

    plans = []
    for chroma in (False, True):
        hreq, vreq = plane_requests(dh, SRC_W, SRC_H, DST_W, DST_H, chroma, dh.BICUBIC, 0.0, 0.5)
        hp, vp = dh.Plan(hreq), dh.Plan(vreq)
        ex.prepare(hp)
        ex.prepare(vp)
        plans.append((hp, vp))
    convs = [dh.integer_conversion(16, False, False).to_c(),
             dh.integer_conversion(16, True, False).to_c()]

    unique = synthetic_frames(torch, dh, 8, 20260, dev)
    batch_in_bytes = FRAMES * 2 * (SRC_W * SRC_H + 2 * (SRC_W // 2) * (SRC_H // 2))
    nbuf = max(2, math.ceil((1 << 30) / batch_in_bytes))
    inputs = []
    for b in range(nbuf):
        idx = torch.arange(FRAMES, device=dev) % unique[0].shape[0]
        inputs.append([p.index_select(0, (idx + b) % unique[0].shape[0]).to(torch.uint16)
                       .contiguous() for p in unique])
    outputs = [[torch.empty((FRAMES, DST_H >> (p > 0), DST_W >> (p > 0)), dtype=torch.uint16,
                            device=dev) for p in range(3)] for _ in range(nbuf)]
    inter = [torch.empty((FRAMES, SRC_H >> (p > 0), DST_W >> (p > 0)), dtype=torch.float32,
                         device=dev) for p in range(2)]

    plane_jobs = (_lib.PlaneJobC * 3)()
    conv_ptrs = [ctypes.cast(ctypes.pointer(c), ctypes.c_void_p) for c in convs]

    def frame_step(s):
        """All three planes of the 64-frame batch in one call."""
        ins, outs = inputs[s % nbuf], outputs[s % nbuf]
        for p in range(3):
            hp, vp = plans[0 if p == 0 else 1]
            x, y = ins[p], outs[p]
            plane_jobs[p] = _lib.PlaneJobC(hp.handle, vp.handle, ctypes.c_void_p(x.data_ptr()),
                                           x.stride(1), x.stride(0),
                                           ctypes.c_void_p(y.data_ptr()), y.stride(1),
                                           y.stride(0), FRAMES, 0, conv_ptrs[0 if p == 0 else 1])
        _lib.check(lib.dsmvc_hip_inverse_2d_planes(ex._handle, plane_jobs, 3, 2, sptr, _lib.ASYNC))

    def step(s, events=None):
        ins, outs = inputs[s % nbuf], outputs[s % nbuf]
        for p in range(3):
            hp, vp = plans[0 if p == 0 else 1]
            x, y, t = ins[p], outs[p], inter[0 if p == 0 else 1]
            conv = ctypes.byref(convs[0 if p == 0 else 1])
            if events is not None and p == 0:
                events[0].record(stream)
            _lib.check(lib.dsmvc_hip_inverse_rows_batch(
                ex._handle, hp.handle, 2, ctypes.c_void_p(x.data_ptr()), x.stride(1), x.stride(0),
                ctypes.c_void_p(t.data_ptr()), t.stride(1), t.stride(0), x.shape[1], FRAMES, conv,
                sptr, _lib.ASYNC))
            if events is not None and p == 0:
                events[1].record(stream)
            _lib.check(lib.dsmvc_hip_inverse_columns_batch(
                ex._handle, vp.handle, ctypes.c_void_p(t.data_ptr()), t.stride(1), t.stride(0), 2,
                ctypes.c_void_p(y.data_ptr()), y.stride(1), y.stride(0), t.shape[2], FRAMES, conv,
                sptr, _lib.ASYNC))
            if events is not None and p == 0:
                events[2].record(stream)

    # Warm-up: the per-plane calls first, so the autotuner picks each pass's
    # route on a launch that fills the GPU.
    for s in range(args.warmup):
        step(s)
        frame_step(s)
    torch.cuda.synchronize()
    ex.synchronize()
    if args.calibration_copy:
        # A plain 256 MiB device copy: the profile's FETCH_SIZE / WRITE_SIZE
        # for this kernel calibrate the counters against known bytes.
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        frame_step(s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ex.synchronize()
    fps = FRAMES * args.steps / elapsed

    # The frame call's own launches, bracketed by HIP events on its stream:
    # with the fused route that is one kernel for all three planes.
    rsteps = max(10, args.steps // 4)
    fevents = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(rsteps)]
    torch.cuda.synchronize()
    for s in range(rsteps):
        fevents[s][0].record(stream)
        frame_step(s)
        fevents[s][1].record(stream)
    torch.cuda.synchronize()
    ex.synchronize()
    frame_call_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in fevents]))
    fused = (args.solve == "auto" and os.environ.get("DSMVC_HIP_FUSED", "1") != "0"
             and all(p.operator()["rows_feasible"] for p in (plans[0][0], plans[1][0]))
             and all(p.operator()["columns_feasible"] for p in (plans[0][1], plans[1][1])))

    # The two-pass route's per-axis kernels: the luma plane's two passes on
    # their own, bracketed by HIP events on the launch stream.
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(rsteps)]
    torch.cuda.synchronize()
    for s in range(rsteps):
        step(s, events[s])
    torch.cuda.synchronize()
    ex.synchronize()
    h_ms = np.array([e[0].elapsed_time(e[1]) for e in events])
    v_ms = np.array([e[1].elapsed_time(e[2]) for e in events])
    per_plane_fps = None
    t1 = time.perf_counter()
    for s in range(rsteps):
        step(s)
    torch.cuda.synchronize()
    per_plane_fps = FRAMES * rsteps / (time.perf_counter() - t1)

    # Per-axis algorithmic bytes (source read once + destination written once),
    # luma plane, one launch = 64 frames.
    h_bytes = FRAMES * SRC_H * (2 * SRC_W + 4 * DST_W)
    v_bytes = FRAMES * DST_W * (4 * SRC_H + 2 * DST_H)
    h_avg, v_avg = float(np.mean(h_ms)) * 1e-3, float(np.mean(v_ms)) * 1e-3
    def family(rows):
        route = ex.tuned_route(plans[0][0 if rows else 1], rows, 2)
        name = "window" if route == -1 else ("chunked" if route > 0 else "seq")
        return (f"rows_{name}<u16> (luma horizontal pass, 64 frames)" if rows
                else f"columns_{name}<u16> (luma vertical pass, 64 frames)")
    if fused:
        dom = "operator2d<u16> (fused 2D, YUV420P16 frame call, 64 frames)"
        dom_bytes, dom_s = FRAMES * FRAME_BYTES_YUV16, frame_call_ms * 1e-3
    elif h_avg >= v_avg:
        dom, dom_bytes, dom_s = family(True), h_bytes, h_avg
    else:
        dom, dom_bytes, dom_s = family(False), v_bytes, v_avg
    achieved = dom_bytes / dom_s / 1e9
    traffic, traffic_source = measured_traffic(dom)
    frame_gbps = FRAME_BYTES_YUV16 * fps / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "traffic_source": traffic_source,
                "kernel": dom, "algorithmic_bytes_per_launch": dom_bytes,
                "avg_launch_ms": round(dom_s * 1e3, 4),
                # The frame call (events around it) and, beside it, the
                # two-pass route's luma passes launched on their own.
                "frame_call_ms": round(frame_call_ms, 4),
                "luma_rows_ms": round(h_avg * 1e3, 4), "luma_columns_ms": round(v_avg * 1e3, 4),
                # The whole frame (all passes of all planes, one call per step):
                # algorithmic frame bytes x frames/s.
                "frame_achieved": round(frame_gbps, 1),
                "frame_frac": round(frame_gbps / HBM_PEAK_GBPS, 4)}

    extra = {"frame_hbm": {"algorithmic_bytes_per_frame": FRAME_BYTES_YUV16,
                           "achieved_GBps": round(frame_gbps, 1),
                           "frac_of_8TBps": round(frame_gbps / HBM_PEAK_GBPS, 4)},
             "frame_call": "dsmvc_hip_inverse_2d_planes (3 planes per call)",
             "per_plane_calls_frames_per_s": round(per_plane_fps, 1),
             "routes": {"frame_call": "fused 2D (operator2d)" if fused else "two passes",
                        "per_plane_luma_rows": family(True), "per_plane_luma_columns": family(False)},
             "input_rotation_bytes": nbuf * batch_in_bytes, "solve": args.solve}

    if not args.no_gray:
        # GRAYS F32 1080p -> 720p through the 2D batch entry: Debicubic (the
        # config-2 kernel) and Debilinear (the config-1 kernel, padding 3).
        gin = [torch.rand((FRAMES, SRC_H, SRC_W), device=dev) for _ in range(nbuf)]
        gout = torch.empty((FRAMES, DST_H, DST_W), device=dev)
        bkw = dict(kernel=dh.BILINEAR, border=dh.SYMMETRIC, f64_mode=0)
        gray_plans = {"debicubic": plans[0],
                      "debilinear": (dh.Plan(dh.AxisRequest(SRC_W, DST_W, float(DST_W), 0.0, **bkw)),
                                     dh.Plan(dh.AxisRequest(SRC_H, DST_H, float(DST_H), 0.0, **bkw)))}
        for name, (hp, vp) in gray_plans.items():
            for s in range(5):
                ex.inverse_2d(hp, vp, gin[s % nbuf], out=gout, asynchronous=True)
            torch.cuda.synchronize()
            gsteps = max(20, args.steps // 4)
            t0 = time.perf_counter()
            for s in range(gsteps):
                ex.inverse_2d(hp, vp, gin[s % nbuf], out=gout, asynchronous=True)
            torch.cuda.synchronize()
            gel = time.perf_counter() - t0
            gfps = FRAMES * gsteps / gel
            extra[f"grays_f32_{name}_1080p_720p"] = {
                "frames_per_s": round(gfps, 1),
                "achieved_GBps": round(FRAME_BYTES_GRAY_F32 * gfps / 1e9, 1),
                "frac_of_8TBps": round(FRAME_BYTES_GRAY_F32 * gfps / 1e9 / HBM_PEAK_GBPS, 4)}
        del gin

    if not args.no_gray:
        # Cost per precision route: Delanczos taps=4 blur 0.9, GRAYS
        # F32 1920x1080 -> 1904x1071, 32 frames per call.  FP32 plain (plans
        # with f64mode 1), FP32 + one refinement step and the Float64 route
        # (f64mode 2 plans); their errors against the FP64 oracle are in
        # tests/test_gpu_precision.py.
        lk = dict(kernel=dh.LANCZOS, taps=4, b=0.0, c=0.5, blur=0.9, border=dh.SYMMETRIC)
        def lplans(mode):
            return (dh.Plan(dh.AxisRequest(SRC_W, 1904, 1904.0, 0.0, f64_mode=mode, **lk)),
                    dh.Plan(dh.AxisRequest(SRC_H, 1071, 1071.0, 0.0, f64_mode=mode, **lk)))
        refined = dh.Executor(0, "refined")
        routes = {"fp32": (ex, lplans(1)), "fp32_refine1": (refined, lplans(2)),
                  "float64": (ex, lplans(2))}
        pin = torch.rand((32, SRC_H, SRC_W), device=dev)
        pout = torch.empty((32, 1071, 1904), device=dev)
        prec = {}
        for name, (rex, (hp, vp)) in routes.items():
            rex.inverse_2d(hp, vp, pin, out=pout)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                rex.inverse_2d(hp, vp, pin, out=pout, asynchronous=True)
            rex.synchronize()
            prec[name] = {"frames_per_s": round(32 * 5 / (time.perf_counter() - t0), 1)}
        extra["precision_routes_lanczos4_1904x1071"] = prec
        del pin, pout, refined

    if not args.no_gray:
        # Host frames (PCIe-inclusive, never the headline): the same YUV420P16
        # workload from pageable host memory through dsmvc_hip_inverse_2d_host
        # per plane (the library stages each plane through pinned memory on an
        # execution slot: pack, upload, solve, download, unpack).  Four
        # threads, as fmParallel drives the plugin, 16 frames per call.
        import threading
        host_in = [torch.cat([p, p])[:16].cpu().numpy().astype(np.uint16) for p in unique]
        host_out = [[np.empty((16, DST_H >> (p > 0), DST_W >> (p > 0)), np.uint16)
                     for p in range(3)] for _ in range(4)]
        hconv = [dh.integer_conversion(16, False, False), dh.integer_conversion(16, True, False)]

        def host_frames(t, calls):
            for _ in range(calls):
                for p in range(3):
                    hp, vp = plans[0 if p == 0 else 1]
                    ex.inverse_2d_host(hp, vp, host_in[p], out=host_out[t][p],
                                       conversion=hconv[0 if p == 0 else 1])
        host_frames(0, 1)
        calls = 6
        t0 = time.perf_counter()
        pool = [threading.Thread(target=host_frames, args=(t, calls)) for t in range(4)]
        for th in pool:
            th.start()
        for th in pool:
            th.join()
        hel = time.perf_counter() - t0
        hfps = 4 * calls * 16 / hel
        extra["host_frames_yuv420p16_debicubic"] = {
            "frames_per_s": round(hfps, 1),
            "achieved_GBps": round(FRAME_BYTES_YUV16 * hfps / 1e9, 1),
            "note": "pageable host planes, 4 threads x 16-frame calls, per plane "
                    "dsmvc_hip_inverse_2d_host (PCIe-inclusive)"}
        del host_in, host_out

    if not args.no_batch:
        frame = (unique[0][0].to(torch.float32) / 65535.0).contiguous()
        extra["mixed_size_scan"] = batched_scan(torch, dh, dev, frame, repeats=5)

    if not args.no_serve:
        extra["served_stream"] = served_stream(torch, dh, dev, seconds=args.serve_seconds)

    cpu = None
    if not args.no_cpu_baseline:
        host = [p[:4].cpu().numpy().astype(np.uint16) for p in unique]
        cpu = cpu_baseline(host)

    line = {"metric": METRIC, "value": round(fps, 1), "unit": "frames/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "Debicubic(b=0,c=0.5) YUV420P16 1920x1080->1280x720, "
                                   "F32 internal, 16-bit output, padding=3, limited range",
                       "frames_per_step": FRAMES, "parallelism": "single GPU"},
            "roofline": roofline, "cpu_baseline": cpu, "extra": extra}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
