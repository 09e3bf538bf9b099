#!/usr/bin/env python3
"""Benchmark of the VBTPACK4 random-access decode path on one MI355X.

Workload (default): the single-field frame-cache build -- synthetic density
256x256x256 over 120 frames (24 advected Gaussian blobs, sigma 3-16, ~69%
empty leaves), VBTPACK4-encoded by this repository's encoder, resident in
HBM.  One step = decode every frame (t = 0..119) of every active leaf into the
packed compact active-leaf cache (meaningful controls only).  Also reported:
point-query throughput/latency of the batched kernel vs the plain
one-request-per-launch GPU decode (smoke256 and the scientific field), the
CPU-runnable smoke64 configuration, the serving loop (five more resident
assets: the three fire fields, the liquid level set and the scientific field;
Poisson and saturation traces against one-request-per-launch decode), and the
reference's own CPU decoders (compiled from its sources) and the oracle on
the host cores (bounded samples, core count and CPU model recorded).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(info, nt: int) -> dict:
    """Frame-cache build bytes per launch: compressed payload + offset table
    read once, nt frames of meaningful control floats written, page table."""
    L = info.header.leafCount
    written = 4.0 * info.controlsPerFrame * nt
    read = float(info.payloadBytes) + 4.0 * (L + 1)
    return {"read": read, "written": written, "page_table": 4.0 * L, "total": read + written + 4.0 * L}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of this frame-cache kernel (tile shape) on this
    workload, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (each
    scaled by what the counter reads for a plain copy of known size), or
    None when that kernel was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh).get(workload, {}).get(kernel)
    except (OSError, ValueError, AttributeError):
        return None
    return None if rec is None else rec.get("traffic_bytes")


def host_cpu() -> dict:
    """Host cores this run may use (the process's CPU affinity, capped by
    OMP_NUM_THREADS when set) and the CPU model."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(omp)) if omp.isdigit() and int(omp) > 0 else aff
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "affinity_cores": aff, "omp_num_threads": omp or None, "threads": threads,
            "model": model}


def _rate(fn, seconds: float, threads: int) -> tuple:
    """(calls, elapsed s) of fn(thread index, call index) run on `threads`
    host threads until `seconds` have passed (the decoders release the GIL)."""
    counter = {"n": 0}
    lock = threading.Lock()
    stop = time.perf_counter() + seconds

    def work(tid):
        i = tid
        while time.perf_counter() < stop:
            fn(tid, i)
            with lock:
                counter["n"] += 1
            i += threads

    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    t0 = time.perf_counter()
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    return counter["n"], time.perf_counter() - t0


def cpu_baseline(data: bytes, frames: int, budget_s: float, threads: int) -> dict:
    """The packed frame-cache decode of whole frames on the host cores, one
    frame per thread: the reference's own per-value decoders (compiled from
    its sources into oracle/_ref) when present, and the oracle (the C
    restatement, -O3) beside it; one thread (the reference's serial frame
    loop) and all threads (the reference's OpenMP-over-work-items pattern)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py

    of = oracle_py.OracleFile(data)
    offs = of.packed_offsets()
    per_frame_bytes = 4.0 * float(offs[-1])
    decoders = {}
    if oracle_py.have_ref():
        rf = oracle_py.RefFile(data)
        decoders["reference"] = lambda t, out: rf.cache_frame_into(t, offs, out)
    import ctypes

    O = oracle_py.oracle()

    def port(t, out):
        if O.vbto_frame_cache_packed(ctypes.byref(of.vf), t, offs.ctypes.data, out.ctypes.data) != 0:
            raise RuntimeError(oracle_py.last_error())

    decoders["port"] = port
    share = budget_s / len(decoders)
    res = {}
    for kind, dec in decoders.items():
        bufs = [np.empty(int(offs[-1]), np.float32) for _ in range(threads)]
        n1, dt1 = _rate(lambda k, i: dec(i % frames, bufs[k]), share / 3, 1)
        nn, dtn = _rate(lambda k, i: dec(i % frames, bufs[k]), share * 2 / 3, threads)
        res[kind] = {"gbs": round(nn * per_frame_bytes / dtn / 1e9, 4),
                     "single_thread_gbs": round(n1 * per_frame_bytes / dt1 / 1e9, 4),
                     "sample": f"{nn} whole-frame cache decodes ({dtn:.1f} s, {threads} threads) of the same asset; "
                               f"1 thread: {n1} frames in {dt1:.1f} s"}
    head = "reference" if "reference" in res else "port"
    return {"value": res[head]["gbs"], "unit": "GB/s", "cores": threads, "kind": head,
            "sample": res[head]["sample"], "single_thread_gbs": res[head]["single_thread_gbs"],
            "by_kind": res, "host": host_cpu()}


def random_queries(dims, n, seed=0x5EED0001) -> np.ndarray:
    W, H, D, F = dims
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, D, n),
                     rng.integers(0, F, n)], axis=1).astype(np.uint32)


def cpu_query_rates(data: bytes, dims, threads: int, seconds: float, scientific: bool) -> dict:
    """Point queries on the host cores over the same seeded random queries
    as the GPU figures: the oracle on 1 and all threads, and for scientific
    assets the reference's serial query loop (decodeScientificValueAtCpu, the
    reference's own CPU query harness) when oracle/_ref is present."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py

    of = oracle_py.OracleFile(data)
    q = random_queries(dims, 65536)
    out = {}
    for name, th in (("port_1_thread", 1), (f"port_{threads}_threads", threads)):
        n, dt = _rate(lambda k, i: of.query(q, threads=th, check=False), seconds / 3, 1)
        out[name] = round(n * len(q) / dt, 1)
    if scientific and oracle_py.have_ref():
        rf = oracle_py.RefFile(data)
        n, dt = _rate(lambda k, i: rf.sci_values(q[:16384]), seconds / 3, 1)
        out["reference_1_thread"] = round(n * 16384 / dt, 1)
    out["unit"] = "queries/s"
    out["queries"] = "65,536 uniform (x, y, z, t), seed 0x5EED0001"
    return out


def cfg1_bench(hip, cpu: bool) -> dict:
    """The CPU-runnable configuration: the reference-encoded 64^3 x 32 smoke
    asset (tests/golden/smoke64), decode of the full frame t = 17 plus 10,000
    seeded uniform (x, y, z, t) point queries, on the GPU (HIP events) and on
    one host thread (oracle; the reference's per-value decode of frame 17's
    controls beside it)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import golden_io
    import oracle_py

    data = golden_io.vbtp("smoke64")
    a = hip.Asset(data)
    s = a.stream
    W, H, D, F = a.dims
    q = random_queries(a.dims, 10000)
    dq = hip.DeviceArray.from_host(q, s)
    dout = hip.DeviceArray(len(q), np.float32, s)
    dense = hip.DeviceArray((D, H, W), np.float32, s)

    def gpu_step():
        a.dense_into(dense, 17)
        a.query_into(dq, len(q), dout)

    gpu_ms = timed(hip, s, gpu_step, 50, 3)
    res = {"workload": "smoke64 (reference-encoded) 64x64x64 x 32: dense frame 17 + 10,000 seeded point queries",
           "gpu_ms": round(gpu_ms, 4)}
    if not cpu:
        a.close()
        return res
    of = oracle_py.OracleFile(data)
    t0 = time.perf_counter()
    reps = 0
    while True:
        of.frame(17)
        of.query(q)
        reps += 1
        if time.perf_counter() - t0 > 1.0:
            break
    cpu_ms = (time.perf_counter() - t0) / reps * 1e3
    res.update({"cpu_port_1_thread_ms": round(cpu_ms, 3), "gpu_over_cpu": round(cpu_ms / gpu_ms, 1)})
    if oracle_py.have_ref():
        rf = oracle_py.RefFile(data)
        offs = of.packed_offsets()
        buf = np.empty(int(offs[-1]), np.float32)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 0.5:
            rf.cache_frame_into(17, offs, buf)
            k += 1
        res["cpu_reference_frame17_controls_1_thread_ms"] = round((time.perf_counter() - t0) / k * 1e3, 3)
    a.close()
    return res


def query_bench(asset, hip, n_batch: int, reps: int) -> dict:
    """Batched point queries (one launch over the batch, 8 lanes per query)
    vs the plain GPU decode (one thread per query); latency per batch."""
    q = random_queries(asset.dims, n_batch)
    s = asset.stream
    dq = hip.DeviceArray.from_host(q, s)
    out = hip.DeviceArray(n_batch, np.float32, s)
    res = {}
    for name, plain in (("batched", False), ("plain", True)):
        for _ in range(3):
            asset.query_into(dq, n_batch, out, plain=plain)
        s.sync()
        ev0, ev1 = hip.Event(), hip.Event()
        lat = []
        for _ in range(reps):
            ev0.record(s)
            asset.query_into(dq, n_batch, out, plain=plain)
            ev1.record(s)
            lat.append(ev0.elapsed_ms(ev1) * 1e3)
        lat = np.array(lat)
        res[name] = {"qps": round(n_batch / (np.median(lat) * 1e-6), 1), "p50_us": round(float(np.percentile(lat, 50)), 2),
                     "p99_us": round(float(np.percentile(lat, 99)), 2)}
    res["batch"] = n_batch
    return res


def query_credit_bytes(data: bytes, q: np.ndarray) -> float:
    """Credited bytes of a batch of point queries (render assets): 20 in, 4 out,
    12 of offset, header word and prefix, then per trilinear corner control
    (8 coarse, plus 8 fine for Fine6 leaves) a descriptor, a bin hint and two
    keyframe records; a Mode2 leaf adds its 80-byte occupancy section and one
    shell control.  Empty leaves cost the first 36 bytes only."""
    hdr = np.frombuffer(data, np.uint32, 11, 8)
    W, H, D, F, leaf_size, L, scales = (int(hdr[i]) for i in (1, 2, 3, 4, 5, 8, 10))
    base = 52 + 4 * scales
    offs = np.frombuffer(data, np.uint32, L + 1, base)
    words = np.frombuffer(data, np.uint32, offset=base + 4 * (L + 1))
    lcx, lcy, lcz = -(-W // leaf_size), -(-H // leaf_size), -(-D // leaf_size)
    bx = np.minimum(q[:, 0] // leaf_size, lcx - 1)
    by = np.minimum(q[:, 1] // leaf_size, lcy - 1)
    bz = np.minimum(q[:, 2] // leaf_size, lcz - 1)
    head = words[offs[(bz * lcy + by) * lcx + bx]]
    mode, cc, fc = head & 3, (head >> 2) & 3, (head >> 4) & 3
    kb = lambda c: np.where(c == 1, 8, 3)  # noqa: E731
    per = 36.0 + np.where(mode != 0, 8 * (4 + 2 + 2 * kb(cc)), 0)
    per = per + np.where(mode == 3, 8 * (4 + 2 + 2 * kb(fc)), 0) + np.where(mode == 2, 80 + 4 + 2 + 2 * kb(fc), 0)
    return float(per.sum())


def pattern_bench(hip, asset, data: bytes, reps: int = 50) -> dict:
    """Point queries in the reference's request shapes (render/src/query_patterns.cpp:
    random, same-t, same-xyz, coherent 8x8x8 tiles), as generated and in
    leaf-major order, 65,536 per batch: the batched kernel (8 lanes per
    query) against the plain one-thread-per-query kernel, HIP events around
    each launch; effective GB/s over the credited bytes."""
    s = asset.stream
    out = {}
    n = 65536
    dout = hip.DeviceArray(n, np.float32, s)
    for pattern in ("random", "same-t", "same-xyz", "coherent"):
        for order in ("generated", "leaf-sorted"):
            q = hip.generate_queries(asset.header, pattern, n, 0x5EED0001, sort_by_leaf=order == "leaf-sorted")
            credit = query_credit_bytes(data, q)
            dq = hip.DeviceArray.from_host(q, s)
            rec = {}
            for kind, plain in (("batched", False), ("plain", True)):
                for _ in range(3):
                    asset.query_into(dq, n, dout, plain=plain)
                s.sync()
                lat = []
                e0, e1 = hip.Event(), hip.Event()
                for _ in range(reps):
                    e0.record(s)
                    asset.query_into(dq, n, dout, plain=plain)
                    e1.record(s)
                    lat.append(e0.elapsed_ms(e1) * 1e3)
                med = float(np.median(lat))
                rec[kind] = {"qps": round(n / (med * 1e-6), 1), "p50_us": round(float(np.percentile(lat, 50)), 2),
                             "p99_us": round(float(np.percentile(lat, 99)), 2),
                             "effective_gbs": round(credit / (med * 1e-6) / 1e9, 1)}
            rec["credited_bytes_per_query"] = round(credit / n, 1)
            rec["batched_over_plain"] = round(rec["batched"]["qps"] / rec["plain"]["qps"], 2)
            out[f"{pattern}/{order}"] = rec
            dq.free()
    dout.free()
    print("[bench] query patterns: " + ", ".join(f"{k} {v['batched']['qps'] / 1e9:.2f} G/s ({v['batched_over_plain']}x)"
                                                 for k, v in out.items()), file=sys.stderr, flush=True)
    return out


def timed(hip, stream, fn, steps: int, warmup: int = 1) -> float:
    """Mean ms per call of fn() on `stream` (HIP events around `steps` calls)."""
    for _ in range(warmup):
        fn()
    stream.sync()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    stream.sync()
    return e0.elapsed_ms(e1) / steps


def frames_bench(hip, asset, workload: str, steps: int) -> dict:
    """Every frame of one asset into its packed cache in one launch, HIP
    events on the launch stream: the liquid level set (Mode2 shells, the
    sweep kernel with its autotuned tile) and the scientific field (the
    temporal-DCT frame sweep, one workgroup per leaf and run of frames)."""
    info = asset.info
    F = asset.dims[3]
    s = asset.stream
    tile = None
    if info.header.profileType != 2:
        tile = asset.autotune_table(F, reps=1)[0]
    out = hip.DeviceArray((F, info.controlsPerFrame), np.float32, s)
    ms = timed(hip, s, lambda: asset.frame_cache_into(out, 0, F), max(3, steps // 2), 1)
    out.free()
    byt = algorithmic_bytes(info, F)
    shapes = hip.Asset.tile_shapes()
    kname = (f"k_cache_render<0, {shapes[tile][0]}, {shapes[tile][1]}>" if tile is not None else "k_sci_frames<false>")
    achieved = byt["total"] / (ms * 1e-3) / 1e9
    res = {"workload": f"{workload} {'x'.join(map(str, asset.dims[:3]))} x {F} frames, all frames per launch",
           "kernel": kname, "ms": round(ms, 4), "frames": F, "slots_per_frame": info.controlsPerFrame,
           "bytes_per_launch": byt["total"], "written_per_launch": byt["written"],
           "decoded_gbs": round(byt["written"] / (ms * 1e-3) / 1e9, 1),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(workload, kname)}}
    print(f"[bench] all frames of {workload}: {ms:.4f} ms, {res['decoded_gbs']} GB/s decoded, "
          f"{100 * achieved / HBM_PEAK_GBS:.1f} % of HBM peak ({kname})", file=sys.stderr, flush=True)
    return res


def batch_bench(hip, assets: dict, names, steps: int, tune_reps: int = 2, fixed=None) -> dict:
    """All frames of each named asset into its packed cache: one replay of
    the batch graph (persistent work-queue kernel, shape autotuned) against
    one plain launch per asset (each with its own autotuned tile)."""
    s = hip.Stream()
    jobs = []
    byt = {"read": 0.0, "written": 0.0, "page_table": 0.0, "total": 0.0}
    for n in names:
        a = assets[n]
        F = a.dims[3]
        jobs.append((a, 0, F, hip.CACHE_PACKED, hip.DeviceArray((F, a.info.controlsPerFrame), np.float32, s)))
        for k, v in algorithmic_bytes(a.info, F).items():
            byt[k] += v
    tiles = {}
    for a, t0, nt, layout, out in jobs:
        tiles[a] = a.autotune_table(nt, reps=1)[0]

    def plain():
        for a, t0, nt, layout, out in jobs:
            a.frame_cache_into(out, t0, nt, layout, s)

    plain_ms = timed(hip, s, plain, steps)
    batch = hip.CacheBatch(jobs, stream=s)
    if fixed is None:
        best, tune_ms, trials = batch.autotune(reps=tune_reps)
    else:  # a given shape: persistent, tile, ranges, chunk, order, queues, heavyTile
        sh = hip.default_batch_shape()
        (sh.persistent, sh.tile, sh.ranges, sh.chunk, sh.order, sh.queues, sh.heavyTile) = fixed
        batch.shape = sh
        best, trials = sh.as_dict(), []
    items, wgs = batch.plan_info()
    graph_ms = timed(hip, s, batch.launch, steps)
    direct_ms = timed(hip, s, lambda: batch.launch(graph=False), steps)
    batch.close()
    for j in jobs:
        j[4].free()
    s.close()

    def rates(ms):
        return {"ms": round(ms, 4), "decoded_gbs": round(byt["written"] / (ms * 1e-3) / 1e9, 1),
                "roofline_frac": round(byt["total"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    return {"assets": list(names), "frames": [assets[n].dims[3] for n in names],
            "controls_per_frame": [assets[n].info.controlsPerFrame for n in names],
            "bytes_per_step": byt["total"], "written_per_step": byt["written"],
            "plain_per_asset": dict(rates(plain_ms), tiles=[tiles[assets[n]] for n in names]),
            "batched_graph": rates(graph_ms), "batched_direct": rates(direct_ms),
            "speedup_graph_over_plain": round(plain_ms / graph_ms, 3),
            "shape": best, "items": items, "workgroups": wgs,
            "autotune_ms": {"persistent/tile/ranges/chunk/order/queues/heavyTile": [
                [t["persistent"], t["tile"], t["ranges"], t["chunk"], t["order"], t["queues"], t["heavyTile"], ms]
                for t, ms in trials]}}


def serving_bench(hip, loadgen, assets: dict, requests: int, seed: int = 0x5EED0005) -> dict:
    """Serving loop over several resident assets: an open-loop trace of
    point-query batches (log-uniform 1..65536 queries, uniform x,y,z,t) and
    whole-frame cache builds (5% of requests, random t), replayed (a) into the
    device-resident server and (b) one request per launch (plain kernel, one
    thread per query), at saturation and at Poisson arrival rates set from the
    plain path's saturation rate.  Latency: arrival -> results in HBM."""
    names = list(assets)
    handles = [assets[n] for n in names]
    srv = hip.Server()  # 80 GiB HBM budget (default)
    ids = [srv.register(a) for a in handles]
    cf = 0.05

    def short(r):
        return {"qps": round(r["queriesPerSec"], 1), "req_per_s": round(r["requestsPerSec"], 1),
                "p50_us": round(r["p50Us"], 1), "p99_us": round(r["p99Us"], 1),
                "query_p50_us": round(r["qp50Us"], 1), "query_p99_us": round(r["qp99Us"], 1),
                "cache_p50_us": round(r["cp50Us"], 1), "cache_p99_us": round(r["cp99Us"], 1),
                "failed": r["failed"]}

    # both paths warmed up on a short trace of their own first (kernel loads,
    # the device loop's first launch, first touches of the outputs)
    loadgen.run(handles, loadgen.MODE_PLAIN, 0.0, 300, cf, 65536, seed + 7)
    loadgen.run(handles, loadgen.MODE_SERVER, 0.0, 300, cf, 65536, seed + 7, srv, ids)
    srv.reset_stats()
    plain_sat = loadgen.run(handles, loadgen.MODE_PLAIN, 0.0, requests, cf, 65536, seed)
    batched_sat = loadgen.run(handles, loadgen.MODE_PER_LAUNCH_BATCHED, 0.0, requests, cf, 65536, seed)
    server_sat = loadgen.run(handles, loadgen.MODE_SERVER, 0.0, requests, cf, 65536, seed, srv, ids)
    server_q = loadgen.run(handles, loadgen.MODE_SERVER, 0.0, requests, 0.0, 65536, seed + 1, srv, ids)
    plain_q = loadgen.run(handles, loadgen.MODE_PLAIN, 0.0, requests, 0.0, 65536, seed + 1)
    poisson = []
    for frac in (0.5, 0.9, 1.5):
        rate = frac * plain_sat["requestsPerSec"]
        sv = loadgen.run(handles, loadgen.MODE_SERVER, rate, 2 * requests, cf, 65536, seed + 2, srv, ids)
        pl = loadgen.run(handles, loadgen.MODE_PLAIN, rate, 2 * requests, cf, 65536, seed + 2)
        poisson.append({"offered_req_per_s": round(rate, 1), "of_plain_saturation": frac,
                        "server": short(sv), "plain": short(pl)})
    st = srv.stats()
    srv.close()
    return {
        "assets": {n: {"dims": list(a.dims), "active_leaves": a.info.activeLeaves,
                       "device_mb": round(a.info.deviceBytes / 1e6, 1)} for n, a in zip(names, handles)},
        "trace": {"requests": requests, "cache_fraction": cf, "query_batch": "log-uniform 1..65536",
                  "hbm_budget_gb": round(st["hbmBudget"] / 2**30, 1), "hbm_in_use_gb": round(st["hbmInUse"] / 2**30, 3)},
        "saturation": {"server": short(server_sat), "plain": short(plain_sat), "per_launch_batched": short(batched_sat),
                       "server_over_plain_qps": round(server_sat["queriesPerSec"] / plain_sat["queriesPerSec"], 2)},
        "saturation_queries_only": {"server": short(server_q), "plain": short(plain_q),
                                    "server_over_plain_qps": round(server_q["queriesPerSec"] / plain_q["queriesPerSec"], 2)},
        "poisson": poisson,
        "server_launches": st["launches"],
    }


def camera_rays(W, H, D, nx, ny, eye_scale, look_y, tmax):
    """Pinhole camera rays (origin, tmin 0, direction, tmax) looking at the
    volume from above and to the side, y up."""
    eye = np.array([W * eye_scale[0], H * eye_scale[1], D * eye_scale[2]])
    look = np.array([W / 2, H * look_y, D / 2])
    fwd = (look - eye) / np.linalg.norm(look - eye)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    u, v = np.meshgrid(np.linspace(-0.5, 0.5, nx), np.linspace(-0.5, 0.5, ny))
    d = fwd[None] + u.reshape(-1, 1) * right[None] + v.reshape(-1, 1) * up[None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = len(d)
    return np.concatenate([np.repeat(eye[None], n, 0), np.zeros((n, 1)), d, np.full((n, 1), tmax)],
                          1).astype(np.float32)


def step_points(rays, steps):
    """Density-mode sample positions of each ray (float64 estimate, for the
    fraction of samples inside the volume)."""
    dt = rays[:, 7:8] / max(steps, 1)
    t = rays[:, 3:4] + 0.5 * dt + dt * np.arange(steps)[None, :]
    return rays[:, None, :3] + rays[:, None, 4:7] * t[..., None]


def march_bench(hip, assets: dict, raws: dict, steps: int, cpu_threads: int, cpu_budget_s: float) -> dict:
    """The viewer's ray marcher on a one-frame stride-576 cache: a 1024x1024
    level-set image of liquid320 (frame 75, floor shadows) and a 512x512
    density image of smoke256 (frame 60, 256 samples per ray), HIP events
    around the march launches; the oracle's march on the host cores for a
    bounded sample of the same rays."""
    out = {}
    cases = []
    if "liquid320" in assets:
        cases.append(("liquid320_levelset_1024", assets["liquid320"], 75, 1024,
                      dict(renderMode=1, backgroundValue=4.0, stepCount=256, maxStepVoxels=4.0, surfaceModel=1,
                           floorHeight=0.5, shadowSteps=96, upAxis=1), (1.4, 1.1, -0.6), 0.35, 3.0))
    if "smoke256" in assets:
        # camera in front of the z = 0 face; rays span the volume's depth
        cases.append(("smoke256_density_512", assets["smoke256"], 60, 512,
                      dict(renderMode=0, stepCount=256, surfaceModel=1), (0.5, 0.5, -0.9), 0.5, 2.0))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    for name, a, t, npx, kw, eye, look_y, reach in cases:
        W, H, D, F = a.dims
        st = a.stream
        cache = hip.DeviceArray((1, a.info.activeLeaves * 576), np.float32, st)
        a.frame_cache_into(cache, t, 1, hip.CACHE_STRIDE576)
        rays = camera_rays(W, H, D, npx, npx, eye, look_y, reach * max(W, H, D))
        n = len(rays)
        d_rays = hip.DeviceArray.from_host(rays, st)
        d_out = hip.DeviceArray((n, 8), np.float32, st)
        p = hip.march_params(frameIndex=t, **kw)
        ms = timed(hip, st, lambda: a.march_into(cache, p, d_rays, n, d_out), max(3, steps // 4))
        res = d_out.host()
        entry = {"rays": n, "ms": round(ms, 3), "rays_per_s": round(n / (ms * 1e-3), 1)}
        if kw.get("renderMode", 0) == 0:
            entry["samples_per_s"] = round(n * kw["stepCount"] / (ms * 1e-3), 1)
            pts = step_points(rays[::97], kw["stepCount"])
            inside = (pts >= 0).all(-1) & (pts[..., 0] <= W - 1) & (pts[..., 1] <= H - 1) & (pts[..., 2] <= D - 1)
            entry["samples_inside_fraction"] = round(float(inside.mean()), 4)
        else:
            entry["hit_fraction"] = round(float((res[:, 3] >= 0).mean()), 4)
        if cpu_budget_s > 0:
            import oracle_py
            of = oracle_py.OracleFile(raws[name.split("_")[0]]) if name.split("_")[0] in raws else None
            if of is not None:
                hc = cache.host().reshape(-1, 576)
                m = min(n, 8192)
                sel = rays[np.random.default_rng(1).choice(n, m, replace=False)]
                t0 = time.perf_counter()
                k = 0
                while True:
                    of.march_rays(hc, p, sel, threads=cpu_threads)
                    k += 1
                    if time.perf_counter() - t0 > cpu_budget_s:
                        break
                dt = time.perf_counter() - t0
                entry["cpu_baseline"] = {"rays_per_s": round(k * m / dt, 1), "cores": cpu_threads, "kind": "port",
                                         "sample": f"{k} x {m} of the same rays in {dt:.1f} s"}
        out[name] = entry
        print(f"[bench] march {name}: {n} rays in {ms:.3f} ms = {entry['rays_per_s'] / 1e6:.1f} M rays/s"
              + (f" (CPU {entry['cpu_baseline']['rays_per_s'] / 1e6:.3f} M rays/s)" if "cpu_baseline" in entry else ""),
              file=sys.stderr, flush=True)
    return out


def single_frame_bench(hip, asset, info, F: int, s) -> dict:
    """One-frame packed-cache builds at 24 frames spread over the sequence,
    HIP events around each build, with the bracket search from the payload
    and then with the frame index (its size and build time reported)."""
    one = hip.DeviceArray((1, info.controlsPerFrame), np.float32, s)
    frames = [int(i * (F - 1) / 23) for i in range(24)]

    def run():
        for t in frames[:4]:
            asset.frame_cache_into(one, t, 1)
        return timed(hip, s, lambda: [asset.frame_cache_into(one, t, 1) for t in frames], 3, 0) / len(frames)

    exact_ms = run()
    t0 = time.perf_counter()
    nbytes = asset.build_frame_index()
    build_s = time.perf_counter() - t0
    index_ms = run()
    asset.drop_frame_index()
    one.free()
    out = {"frames": len(frames), "exact_us": round(exact_ms * 1e3, 2), "indexed_us": round(index_ms * 1e3, 2),
           "index_mb": round(nbytes / 1e6, 1), "index_build_ms": round(build_s * 1e3, 1),
           "out_mb_per_frame": round(4 * info.controlsPerFrame / 1e6, 2)}
    print(f"[bench] one-frame builds: {out['exact_us']} us (bracket search) vs {out['indexed_us']} us (frame index, "
          f"{out['index_mb']} MB, built in {out['index_build_ms']} ms)", file=sys.stderr, flush=True)
    return out


def cycles_bench(hip, assets: dict, raws: dict, steps: int, cpu_threads: int, cpu_budget_s: float) -> dict:
    """The Cycles resident blob and its samplers: the blob packed on the device
    for one frame, then 2^20 uniform points inside the volume sampled by the
    standalone form (8 lanes per sample) and the patched kernel's form (64
    lanes per sample), each beside its one-lane-per-sample form; HIP events
    around the launches; the oracle on the host cores for a bounded sample of
    the same points."""
    out = {}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    forms = {"standalone": 0, "standalone_plain": hip.CYCLES_PLAIN, "kernel": hip.CYCLES_KERNEL,
             "kernel_plain": hip.CYCLES_KERNEL | hip.CYCLES_PLAIN}
    for name, t in (("smoke256", 60), ("liquid320", 75)):
        if name not in assets:
            continue
        a = assets[name]
        W, H, D, F = a.dims
        st = a.stream
        blob = a.cycles_blob(t)
        pack_ms = timed(hip, st, lambda: hip.check(hip.lib().vbt_cycles_pack_blob(a.h, t, blob.ptr, blob.nbytes, st.h)),
                        max(3, steps // 4))
        n = 1 << 20
        pts = (np.random.default_rng(12).random((n, 3)) * [W - 1, H - 1, D - 1]).astype(np.float32)
        d_pts = hip.DeviceArray.from_host(pts, st)
        d_out = hip.DeviceArray(n, np.float32, st)
        entry = {"points": n, "frame": t, "blob_mb": round(blob.nbytes / 1e6, 2), "pack_ms": round(pack_ms, 3)}
        for form, flags in forms.items():
            ms = timed(hip, st, lambda: hip.cycles_sample_into(blob, d_pts, n, d_out, flags, st), max(3, steps // 4))
            entry[form] = {"ms": round(ms, 3), "samples_per_s": round(n / (ms * 1e-3), 1)}
        entry["nonzero_fraction"] = round(float((d_out.host() != 0).mean()), 4)
        if cpu_budget_s > 0 and name in raws:
            import oracle_py
            hb = blob.host()
            for form, flags in (("standalone", 0), ("kernel", hip.CYCLES_KERNEL)):
                m = 65536 if flags == 0 else 8192
                sel = pts[:m]
                t0 = time.perf_counter()
                k = 0
                while True:
                    oracle_py.cycles_sample(hb, sel, flags, threads=cpu_threads)
                    k += 1
                    if time.perf_counter() - t0 > cpu_budget_s:
                        break
                dt = time.perf_counter() - t0
                entry[form]["cpu_baseline"] = {"samples_per_s": round(k * m / dt, 1), "cores": cpu_threads,
                                               "kind": "port", "sample": f"{k} x {m} of the same points in {dt:.1f} s"}
        out[f"{name}_t{t}"] = entry
        print(f"[bench] cycles {name}: blob {entry['blob_mb']} MB packed in {pack_ms:.3f} ms; "
              + ", ".join(f"{f} {entry[f]['samples_per_s'] / 1e6:.1f} M/s" for f in forms)
              + "".join(f"; CPU {f} {entry[f]['cpu_baseline']['samples_per_s'] / 1e6:.2f} M/s"
                        for f in ("standalone", "kernel") if "cpu_baseline" in entry[f]),
              file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="smoke256")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-assets", default="fire_density,fire_flames,fire_temperature,liquid320,sci192",
                    help="assets resident beside the frame-cache workload in the serving run")
    ap.add_argument("--serve-requests", type=int, default=1500)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-queries", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-ceilings", action="store_true")
    ap.add_argument("--no-march", action="store_true")
    ap.add_argument("--no-cycles", action="store_true")
    ap.add_argument("--frames-assets", default="liquid320,sci192",
                    help="assets whose every frame is built in one launch (the Mode2 and scientific kernels)")
    ap.add_argument("--no-frames", action="store_true")
    ap.add_argument("--batch-tune-reps", type=int, default=2)
    ap.add_argument("--batch-shape", default="",
                    help="fixed batch shape 'persistent,tile,ranges,chunk,order,queues,heavyTile' (no autotune)")
    ap.add_argument("--batch-assets", default="fire_density,fire_flames,fire_temperature",
                    help="assets of the batched frame-cache build (all frames of each, one graph replay)")
    ap.add_argument("--tile", type=int, default=-1, help="frame-cache tile shape index (default: tuning table, "
                    "else autotune)")
    ap.add_argument("--tuning", default=os.path.join(ROOT, "vbt_amd", "tuning_gfx950.json"),
                    help="tuning table to start from (missing file: start empty)")
    ap.add_argument("--retune", action="store_true", help="start from an empty tuning table")
    ap.add_argument("--save-tuning", default="", help="write the tuning table here at the end")
    ap.add_argument("--single-frame-index", action="store_true",
                    help="build the frame index before the single-frame probe")
    ap.add_argument("--single-frame-probe", type=int, default=0,
                    help="also build this many one-frame packed caches (frames spread over the asset) for profiling")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one plain device copy of this many MiB (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the decode shards by leaf or "
              "frame ranges with replicated assets; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from vbt_amd import hip, loadgen, synth

    if hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if not args.retune and args.tuning and os.path.exists(args.tuning):
        hip.tuning_load(args.tuning)
    w = synth.WORKLOADS[args.workload]
    HT = host_cpu()["threads"]
    t0 = time.time()
    data = w.encode(threads=HT)
    enc_s = time.time() - t0
    print(f"[bench] encoded {args.workload} {w.dims}: {len(data) / 1e6:.1f} MB in {enc_s:.1f} s", file=sys.stderr,
          flush=True)
    s = hip.Stream()
    asset = hip.Asset(data, s)
    info = asset.info
    W, H, D, F = asset.dims
    nt = F
    shapes = hip.Asset.tile_shapes()
    tile_source = "fixed"
    if args.tile >= 0:
        asset.cache_tile = args.tile
        best, tile_ms = args.tile, [None] * len(shapes)
    else:
        best, measured, tile_ms = asset.autotune_table(nt, reps=2, times=True)
        tile_source = "measured" if measured else "table"
        tile_ms = tile_ms or [None] * len(shapes)
    tuning_key = asset.tuning_key(nt)
    print(f"[bench] tile {shapes[best]} ({tile_source}, key '{tuning_key}') ms per tile {tile_ms}", file=sys.stderr,
          flush=True)
    out = hip.DeviceArray((nt, info.controlsPerFrame), np.float32, s)
    for _ in range(args.warmup):
        asset.frame_cache_into(out, 0, nt)
    s.sync()
    ev0, ev1 = hip.Event(), hip.Event()
    per = []
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        e0, e1 = hip.Event(), hip.Event()
        e0.record(s)
        asset.frame_cache_into(out, 0, nt)
        e1.record(s)
        per.append((e0, e1))
    s.sync()
    wall = time.perf_counter() - t_wall
    ms = np.array([a.elapsed_ms(b) for a, b in per])
    ms_step = float(ms.mean())
    byt = algorithmic_bytes(info, nt)
    decoded_gbs = byt["written"] / (ms_step * 1e-3) / 1e9
    achieved = byt["total"] / (ms_step * 1e-3) / 1e9
    print(f"[bench] {args.steps} steps: kernel mean {ms_step:.4f} ms (min {ms.min():.4f}), wall/step "
          f"{wall / args.steps * 1e3:.4f} ms", file=sys.stderr, flush=True)

    # bandwidth ceilings of plain kernels on this device (1 GiB write-only fill,
    # 512 MiB -> 512 MiB copy): what streaming HBM traffic reaches here
    ceil = {}
    if not args.no_ceilings:
        nb = 1 << 30
        buf = hip.DeviceArray(nb // 4, np.float32, s)
        ms_fill = timed(hip, s, lambda: hip.device_fill(buf, 0x3F800000, nb, s), 10)
        half = nb // 2
        src = hip.DeviceArray(half // 4, np.float32, s)
        ms_copy = timed(hip, s, lambda: hip.device_copy(buf, src, half, s), 10)
        ceil = {"fill_gbs": round(nb / (ms_fill * 1e-3) / 1e9, 1),
                "copy_gbs": round(2 * half / (ms_copy * 1e-3) / 1e9, 1)}
        buf.free()
        src.free()
        print(f"[bench] ceilings: fill {ceil['fill_gbs']} GB/s, copy {ceil['copy_gbs']} GB/s", file=sys.stderr,
              flush=True)

    # the build's own memory traffic without its decode, same tiles and
    # stores: writes alone, and with each tile's keyframe records staged
    layout = None
    if not args.no_ceilings:
        ms_w = timed(hip, s, lambda: asset.cache_ceiling_into(out, nt, best, reads=False, stream=s), 10)
        ms_rw = timed(hip, s, lambda: asset.cache_ceiling_into(out, nt, best, reads=True, stream=s), 10)
        layout = {"write_only_ms": round(ms_w, 4), "stage_and_write_ms": round(ms_rw, 4),
                  "build_ms": round(ms_step, 4), "build_over_stage_and_write": round(ms_step / ms_rw, 3),
                  "stage_and_write_frac_of_peak": round(byt["total"] / (ms_rw * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        print(f"[bench] layout ceiling ({shapes[best][0]} x {shapes[best][1]} tiles, no decode): writes only "
              f"{ms_w:.4f} ms, records staged + writes {ms_rw:.4f} ms; the build takes {ms_step / ms_rw:.3f} x "
              f"the latter", file=sys.stderr, flush=True)

    if args.copy_probe_mb:
        nb = args.copy_probe_mb << 20
        src = hip.DeviceArray(nb // 4, np.float32, s)
        dst = hip.DeviceArray(nb // 4, np.float32, s)
        hip.device_copy(dst, src, nb, s)
        s.sync()
        print(f"[bench] copy probe: {nb} bytes read + {nb} bytes written by k_copy", file=sys.stderr, flush=True)
        src.free()
        dst.free()

    single = single_frame_bench(hip, asset, info, F, s)
    if args.single_frame_probe:
        if args.single_frame_index:
            asset.build_frame_index()
        one = hip.DeviceArray((1, info.controlsPerFrame), np.float32, s)
        k = args.single_frame_probe
        frames = [int(i * (F - 1) / max(k - 1, 1)) for i in range(k)]
        for t in frames:
            asset.frame_cache_into(one, t, 1)
        s.sync()
        print(f"[bench] single-frame probe: frames {frames}", file=sys.stderr, flush=True)
        one.free()
        asset.drop_frame_index()

    qres = small = None
    patterns = None
    if not args.no_queries:
        qres = query_bench(asset, hip, 65536, 50)
        small = query_bench(asset, hip, 256, 200)
        patterns = pattern_bench(hip, asset, data)

    kname = f"k_cache_render<0, {shapes[best][0]}, {shapes[best][1]}>"
    line = {
        "metric": "decoded GB/s + % HBM roofline; batched query p50/p99 us vs plain GPU decode",
        "value": round(decoded_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: frame-cache build {W}x{H}x{D} x {F} frames, all frames per step",
                   "active_leaves": info.activeLeaves, "leaves": info.header.leafCount,
                   "controls_per_frame": info.controlsPerFrame, "payload_mb": round(info.payloadBytes / 1e6, 2),
                   "fast_leaves": info.fast_leaves,
                   "tile": {"workgroup_slots": shapes[best][0], "lds_window_bytes": shapes[best][1],
                            "source": tile_source, "tuning_key": tuning_key,
                            "autotune_ms": dict(zip([f"{b}x{w}" for b, w in shapes], tile_ms))}},
        "hbm_roofline_pct": round(100.0 * achieved / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.workload, kname),
                     "bytes_per_launch": byt["total"],
                     "kernel": kname,
                     "device_ceilings_gbs": ceil, "layout_ceiling": layout},
        "queries": {"random_65536": qres, "random_256": small, "patterns_65536": patterns},
        "single_frame": single,
    }
    extra = {args.workload: asset}
    raws = {args.workload: data}
    want = ([] if args.no_batch else [x for x in args.batch_assets.split(",") if x]) + \
        ([] if args.no_serve else [x for x in args.serve_assets.split(",") if x]) + \
        ([] if args.no_frames else [x for x in args.frames_assets.split(",") if x]) + \
        ([] if args.no_march and args.no_cycles else ["liquid320"])
    for n in want:
        if n not in extra:
            t0 = time.time()
            raws[n] = synth.WORKLOADS[n].encode(threads=HT)
            extra[n] = hip.Asset(raws[n])
            print(f"[bench] asset {n}: encoded in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    if not args.no_queries:
        if not args.no_cpu:
            line["queries"]["cpu"] = cpu_query_rates(data, asset.dims, HT, 6.0, False)
        if "sci192" in extra:
            sa = extra["sci192"]
            sq = {"random_65536": query_bench(sa, hip, 65536, 50), "random_256": query_bench(sa, hip, 256, 200)}
            if not args.no_cpu:
                sq["cpu"] = cpu_query_rates(raws["sci192"], sa.dims, HT, 6.0, True)
            line["queries_sci192"] = sq
        line["cfg1_smoke64"] = cfg1_bench(hip, not args.no_cpu)
        c1 = line["cfg1_smoke64"]
        print(f"[bench] cfg1 smoke64: GPU {c1['gpu_ms']} ms, CPU {c1.get('cpu_port_1_thread_ms')} ms (1 thread)",
              file=sys.stderr, flush=True)
    if not args.no_frames:
        line["all_frames"] = {n: frames_bench(hip, extra[n], n, args.steps)
                              for n in [x for x in args.frames_assets.split(",") if x]}
    if not args.no_batch:
        fixed = tuple(int(x) for x in args.batch_shape.split(",")) if args.batch_shape else None
        line["batch_single"] = batch_bench(hip, extra, [args.workload], args.steps, args.batch_tune_reps, fixed)
        names = [x for x in args.batch_assets.split(",") if x]
        line["batch"] = batch_bench(hip, extra, names, args.steps, args.batch_tune_reps, fixed)
        bb = line["batch"]
        print(f"[bench] batch {names}: graph {bb['batched_graph']['ms']} ms vs plain "
              f"{bb['plain_per_asset']['ms']} ms ({bb['batched_graph']['decoded_gbs']} GB/s decoded), shape "
              f"{bb['shape']}; single {line['batch_single']['batched_graph']['ms']} ms vs one-shot "
              f"{line['batch_single']['plain_per_asset']['ms']} ms", file=sys.stderr, flush=True)
    if not args.no_march:
        line["march"] = march_bench(hip, extra, raws, args.steps, HT,
                                    0.0 if args.no_cpu else 3.0)
    if not args.no_cycles:
        line["cycles"] = cycles_bench(hip, extra, raws, args.steps, HT,
                                      0.0 if args.no_cpu else 2.0)
    if not args.no_serve:
        serve = {n: extra[n] for n in [args.workload] + [x for x in args.serve_assets.split(",") if x]}
        line["serving"] = serving_bench(hip, loadgen, serve, args.serve_requests)
        sat = line["serving"]["saturation"]
        print(f"[bench] serving: server {sat['server']['qps'] / 1e6:.1f} Mq/s vs plain {sat['plain']['qps'] / 1e6:.1f} "
              f"Mq/s at saturation", file=sys.stderr, flush=True)

    for a in extra.values():
        if a is not asset:
            a.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(data, F, args.cpu_seconds, HT)
    asset.close()
    if args.save_tuning:
        hip.tuning_save(args.save_tuning)
        print(f"[bench] tuning table: {hip.tuning_count()} entries written to {args.save_tuning}", file=sys.stderr,
              flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
