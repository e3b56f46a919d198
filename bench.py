#!/usr/bin/env python3
"""SSIMULACRA 2.1 throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1080p|4k|512]
                    [--planes f32|u16|f16|bf16] [--map default|reference|difference]
                    [--no-cpu]

A step scores one batch of synthetic pairs already resident in HBM: the 1080p
workload is 3 distortion strengths x 100 pairs (block step 4/12/24, noise
sigma 1/3/8), generated deterministically (vqm_amd/synth.py). Every pair of a
step goes through one prepare, one scale and one reduce launch of the batch
engine; the norms of all pairs come back to the host before the step ends.
The map form is the library default (the reference's operation order for
FP32 planes) unless --map names another.

Printed: one JSON line with pairs/s (`value`), the roofline of the dominant
kernel (algorithmic bytes per launch over its HIP-event duration on the batch
stream, against 8 TB/s, with the VALU issue rate that binds it), the CPU
baseline (the oracle's FP32 ref form on the host's cores, one pair per
thread, median of rounds), and extras: the other map form and plane form at
the same workload (`map_forms`, `plane_forms`), the 4K batch in FP32 and the
mixed form (`workload_4k`), streaming through Ssimulacra2Pool with 1 and 3
workers (`streaming_pool`), resident pairs through 1 and 3 pipelines by vqm's
own protocol (`resident_workers`), single-pair latency (`modes`), the
uneven-size batch (`ladder_batch`) and the serving loop (`serving`).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
RTX5070_1080P_RESIDENT = 3836.7  # 1080p resident, 3 workers, RTX 5070
WORKLOADS = {"1080p": (1920, 1080, 1), "4k": (3840, 2160, 2), "512": (512, 512, 0)}
PLANES = {"f32": 0, "f16": 1, "bf16": 2, "u16": 3}  # u16: the mixed form
MAPS = {"default": 0, "reference": 1, "difference": 2}
# vqm's published streaming rates with 3 workers (RTX 5070).
RTX5070_STREAMING = {"1920x1080": 1816.4, "3840x2160": 455.2}
# vqm's resident rates on an RTX 5070 with 1 and 3 workers (bench/README.md:32,162,169).
RTX5070_RESIDENT = {"1920x1080": {"workers_1": 3043.0, "workers_3": 3836.7},
                    "3840x2160": {"workers_1": 845.7, "workers_3": 915.0}}


def map_name(fmt: int, map_form: int) -> str:
    """The map form a configuration runs (VQM_MAP_DEFAULT resolved)."""
    if map_form == 0:
        return "difference" if fmt == 3 else "reference"
    return "reference" if map_form == 1 else "difference"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(kernel: str, planes: str, form: str, w: int, h: int, pairs: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by the
    plain copy kernel's reading of a known 1 GiB), or None when no summary
    covers it."""
    if (w, h, pairs) != (1920, 1080, 300):
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"ss2_kernels_1080p_batch300_{planes}_{form}.md")
    if not os.path.exists(path):
        return None
    row = f"| {kernel}_kernel<{planes.upper()}, {form}> |"
    derived = False
    for line in open(path):
        if line.startswith("| kernel | HBM read GB"):
            derived = True
        elif derived and line.startswith(row):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            return {"bytes": round((float(cells[1]) + float(cells[2])) * 1e9),
                    "valu_per_cu_cycle": float(cells[5]),
                    "source": os.path.relpath(path, os.path.dirname(os.path.abspath(__file__)))}
    return None


def plane_bytes(w: int, h: int, fmt: int) -> int:
    """Stored plane bytes of one image (pruned planes excluded, no padding):
    the library's algorithmic pair bytes are 6N input + 4 x these."""
    from vqm_amd import Size, algorithmic_bytes
    return (algorithmic_bytes(Size(w, h), fmt) - 6 * w * h) // 4


def make_pairs(w, h, cfg, per_strength):
    from vqm_amd import synth
    pairs = []
    for k, strength in enumerate(synth.STRENGTHS):
        for i in range(per_strength):
            pairs.append(synth.pair(cfg, k * per_strength + i, w, h, strength))
    return pairs


def make_pairs_cycled(w, h, cfg, count):
    """`count` pairs whose strengths cycle through the three."""
    from vqm_amd import synth
    return [synth.pair(cfg, i, w, h, synth.STRENGTHS[i % 3]) for i in range(count)]


def upload_pairs(pairs, w, h):
    """One device buffer holding every pair; returns it and the batch refs."""
    from vqm_amd import DeviceBuffer
    slot = (pairs[0][0].nbytes + 255) // 256 * 256
    dev = DeviceBuffer(2 * slot * len(pairs))
    refs = []
    for i, (a, b) in enumerate(pairs):
        dev.upload(a, 2 * i * slot)
        dev.upload(b, (2 * i + 1) * slot)
        refs.append((dev.ptr + 2 * i * slot, dev.ptr + (2 * i + 1) * slot, w, h))
    return dev, refs


class LadderPool:
    """Distinct seeded pairs of every ladder size and strength, resident on
    the device and, for host-input requests, in pinned host memory."""

    def __init__(self, cfg, variants, pinned):
        from vqm_amd import DeviceBuffer, PinnedBuffer, synth
        self.entries = {}
        self.keep = []
        t0 = time.perf_counter()
        for si, (w, h) in enumerate(synth.LADDER):
            for sk, strength in enumerate(synth.STRENGTHS):
                for v in range(variants):
                    idx = (si * len(synth.STRENGTHS) + sk) * variants + v
                    a, b = synth.pair(cfg, idx, w, h, strength)
                    da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
                    da.upload(a)
                    db.upload(b)
                    entry = {"dev": (da.ptr, db.ptr), "host": None}
                    self.keep += [da, db]
                    if pinned:
                        pa, pb = PinnedBuffer(a.nbytes), PinnedBuffer(b.nbytes)
                        pa.array[:] = a
                        pb.array[:] = b
                        entry["host"] = (pa.ptr, pb.ptr)
                        self.keep += [pa, pb]
                    self.entries[(w, h, sk, v)] = entry
        self.variants = variants
        log(f"ladder pool: {len(self.entries)} pairs ({variants} per size and strength) "
            f"in {time.perf_counter() - t0:.1f} s")

    def pick(self, cfg, index, weights):
        from vqm_amd import synth
        (w, h), strength = synth.ladder_pair_shape(cfg, index, weights)
        sk = synth.STRENGTHS.index(strength)
        v = min(int(synth.draw(cfg, index, 3) * self.variants), self.variants - 1)
        return w, h, self.entries[(w, h, sk, v)]


def h2d_gbps(nbytes: int = 256 << 20, reps: int = 5) -> float:
    """Pinned host-to-device copy rate (hipMemcpy from page-locked memory)."""
    from vqm_amd import DeviceBuffer, PinnedBuffer, memcpy_htod
    src, dst = PinnedBuffer(nbytes), DeviceBuffer(nbytes)
    src.array[:] = 1
    memcpy_htod(dst.ptr, src.array)
    t0 = time.perf_counter()
    for _ in range(reps):
        memcpy_htod(dst.ptr, src.array)
    dt = time.perf_counter() - t0
    src.free()
    dst.free()
    return nbytes * reps / dt / 1e9


def serving(args, pool):
    """Poisson arrivals over the five-size ladder (weights 0.1/0.2/0.35/0.2/
    0.15) replayed open-loop against the continuous-batching server: the
    capacity of a burst, then latency at fixed offered rates."""
    import math
    from vqm_amd import Ssimulacra2Server, synth
    n = args.serve_requests
    picks = [pool.pick(4, i, synth.SERVE_WEIGHTS) for i in range(n)]
    unit = [-math.log(1.0 - synth.draw(4, i, 2)) for i in range(n)]  # Exp(1) gaps
    budget = args.serve_budget_gib << 30

    def replay(server, host, qps):
        t, reqs = 0.0, []
        for (w, h, e), gap in zip(picks, unit):
            if qps:
                t += gap / qps
            a, b = e["host"] if host else e["dev"]
            reqs.append((a, b, w, h, host, int(t * 1e9)))
        out = server.Replay(reqs)
        lat = np.array([r.latency_ms for r in out])
        start = min(r.submit_ns for r in out)
        end = max(r.done_ns for r in out)
        return {"offered_qps": round(qps, 1) if qps else "burst",
                "pairs_per_s": round(n / ((end - start) * 1e-9), 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_batch": round(float(np.mean([r.batch_size for r in out])), 2)}

    report = {"requests": n, "hbm_budget_bytes": budget, "lanes": args.serve_lanes,
              "max_batch": args.serve_max_batch}
    # The streaming forms' roof: host inputs cross PCIe, 6 bytes per pixel of
    # each pair, at the pinned host-to-device copy rate measured here.
    h2d = h2d_gbps()
    mean_px = sum(w * h for w, h, _ in picks) / n
    report["h2d_gbps"] = round(h2d, 1)
    report["streaming_roof_pairs_per_s"] = round(h2d * 1e9 / (6 * mean_px), 1)
    # The plain FP32 single-pair form (one pair on the device at a time, no
    # batching) faces the same arrivals at the same offered rates.
    forms = (("f32", args.serve_lanes, args.serve_max_batch), ("u16", args.serve_lanes, args.serve_max_batch),
             ("single_pair_f32", 1, 1))
    for mode, host, fractions in (("resident", False, (0.25, 0.5, 0.75, 0.9)),
                                  ("streaming", True, (0.5, 0.9))):
        base_capacity = None
        for name, lanes, max_batch in forms:
            planes = "f32" if name.endswith("f32") else name
            server = Ssimulacra2Server(plane_format=PLANES[planes], lanes=lanes,
                                       max_batch=max_batch, hbm_budget=budget)
            replay(server, host, 0)  # warm the workspace pools
            burst = replay(server, host, 0)
            if base_capacity is None:
                base_capacity = burst["pairs_per_s"]
            points = [burst] + [replay(server, host, f * base_capacity) for f in fractions]
            stats = server.Stats()
            server.close()
            report[f"{mode}_{name}"] = {"capacity_pairs_per_s": burst["pairs_per_s"],
                                        "points": points,
                                        "peak_pool_bytes": stats["peak_pool_bytes"]}
            log(f"serving {mode} {name}: {points}")
    return report


def ladder_batch(args, pool):
    """256 pairs drawn uniformly from the ladder in one batched launch per
    kernel, against the same pairs scored one at a time through one
    single-pair pipeline per size."""
    from vqm_amd import DeviceImage, ImageFormat, Size, Ssimulacra2Batch, Ssimulacra2Config, Ssimulacra2Pipeline
    n = 256
    picks = [pool.pick(3, i, None) for i in range(n)]
    refs = [(e["dev"][0], e["dev"][1], w, h) for w, h, e in picks]
    out = {"pairs": n}
    for planes in ("f32", "u16"):
        batch = Ssimulacra2Batch(max_pairs=n, plane_format=PLANES[planes])
        for _ in range(2):
            batch.Finish(batch.Submit(refs), raw=True)
        t0 = time.perf_counter()
        steps = 5
        for _ in range(steps):
            batch.Finish(batch.Submit(refs), raw=True)
        dt = (time.perf_counter() - t0) / steps
        stages = batch.StageTimes()
        batch.close()
        out[f"batched_{planes}"] = {"pairs_per_s": round(n / dt, 1), "ms_per_launch": round(dt * 1e3, 3),
                                    "stage_ms": [round(s, 3) for s in stages]}
    pipes = {}
    for planes in ("f32",):
        for w, h, _ in picks:
            if (w, h) not in pipes:
                pipes[(w, h)] = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(w, h), plane_format=PLANES[planes]))
        for w, h, e in picks[:10]:
            pipes[(w, h)].Compute(DeviceImage(e["dev"][0], Size(w, h), ImageFormat.srgb8()), DeviceImage(e["dev"][1], Size(w, h), ImageFormat.srgb8()))
        t0 = time.perf_counter()
        for w, h, e in picks:
            pipes[(w, h)].Compute(DeviceImage(e["dev"][0], Size(w, h), ImageFormat.srgb8()), DeviceImage(e["dev"][1], Size(w, h), ImageFormat.srgb8()))
        dt = time.perf_counter() - t0
        out[f"single_pair_loop_{planes}"] = {"pairs_per_s": round(n / dt, 1)}
        for p in pipes.values():
            p.close()
    log(f"ladder batch: {out}")
    return out


def general_decode(pairs, w, h, map_form, count=60):
    """Pairs/s and per-stage times of one pipeline scoring resident device
    images one pair at a time, in 8-bit 4:2:0 limited-range BT.709 YUV (the
    general per-pixel decode of the prepare kernel) beside the same content
    as 8-bit sRGB RGB (the table decode). The YUV planes are the RGB pair's
    R plane as luma and its G and B planes subsampled 2x2 as chroma."""
    from vqm_amd import (ColorFamily, DeviceBuffer, DeviceImage, ImageFormat, MatrixCoefficients, PixelRange,
                         PixelType, Size, Ssimulacra2Config, Ssimulacra2Pipeline, TransferCharacteristics,
                         required_bytes)
    size = Size(w, h)
    yuv = ImageFormat(type=PixelType.Byte, family=ColorFamily.Yuv, range=PixelRange.Limited,
                      matrix=MatrixCoefficients.Bt709, transfer=TransferCharacteristics.Bt709,
                      subsample_w=1, subsample_h=1)
    plane = (w * h + 3) // 4 * 4

    def to_yuv(img):
        rgb = [img[k * plane:k * plane + w * h].reshape(h, w) for k in range(3)]
        out = np.zeros(required_bytes(yuv, size), np.uint8)
        out[:w * h] = rgb[0].ravel()
        cw, ch = (w + 1) // 2, (h + 1) // 2
        off = plane
        for k in (1, 2):
            out[off:off + cw * ch] = rgb[k][::2, ::2].ravel()
            off += (cw * ch + 3) // 4 * 4
        return out

    out = {"workload": f"{w}x{h}, {count} resident pairs scored one at a time by one pipeline"}
    for name, fmt, conv in (("yuv420p8_bt709_limited", yuv, to_yuv), ("srgb8", ImageFormat.srgb8(), None)):
        bufs = []
        for a, b in pairs[:count]:
            pair = []
            for img in (a, b):
                data = conv(img) if conv else img
                buf = DeviceBuffer(data.nbytes)
                buf.upload(data)
                pair.append(buf)
            bufs.append(pair)
        pipe = Ssimulacra2Pipeline(Ssimulacra2Config(size=size, map_form=map_form))
        for da, db in bufs[:3]:
            pipe.Compute(DeviceImage(da.ptr, size, fmt), DeviceImage(db.ptr, size, fmt))
        stages = []
        t0 = time.perf_counter()
        for da, db in bufs:
            pipe.Compute(DeviceImage(da.ptr, size, fmt), DeviceImage(db.ptr, size, fmt))
            stages.append(pipe.StageTimes())
        dt = time.perf_counter() - t0
        pipe.close()
        for pair in bufs:
            for buf in pair:
                buf.free()
        out[name] = {"pairs_per_s": round(len(bufs) / dt, 1),
                     "stage_ms": {k: round(statistics.mean(st[i] for st in stages), 4)
                                  for i, k in enumerate(("prepare", "scale", "reduce"))}}
    log(f"general decode: {out}")
    return out


def resident_batch(refs, w, h, fmt, steps, warmup, chunks=1, strip_rows=0, persistent=False, map_form=0,
                   strip_cols=0):
    """Pairs/s of the batch engine over resident pairs, with the per-stage
    HIP-event times and the scale kernel's HBM fraction."""
    from vqm_amd import Ssimulacra2Batch
    batch = Ssimulacra2Batch(max_pairs=len(refs), plane_format=fmt, strip_rows=strip_rows, chunks=chunks,
                             persistent=persistent, map_form=map_form, strip_cols=strip_cols)
    for _ in range(warmup):
        batch.Finish(batch.Submit(refs), raw=True)
    stages = []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, norms = batch.Finish(batch.Submit(refs), raw=True)
        stages.append(batch.StageTimes())
    total = time.perf_counter() - t0
    batch.close()
    if not np.all(np.isfinite(norms)):
        raise RuntimeError("non-finite norms")
    stage = [statistics.mean(s[k] for s in stages) for k in range(3)]
    planes = plane_bytes(w, h, fmt)
    prep_gbps = len(refs) * (6 * w * h + 2 * planes) / (stage[0] * 1e-3) / 1e9
    scale_gbps = len(refs) * 2 * planes / (stage[1] * 1e-3) / 1e9
    return {"pairs_per_s": round(len(refs) * steps / total, 1),
            "map_form": map_name(fmt, map_form),
            "ms_per_step": round(total / steps * 1e3, 3),
            "pairs_per_step": len(refs),
            "stage_ms": {"prepare": round(stage[0], 4), "scale": round(stage[1], 4),
                         "reduce": round(stage[2], 4)},
            "hbm_frac": {"prepare": round(prep_gbps / HBM_PEAK_GBPS, 4),
                         "scale": round(scale_gbps / HBM_PEAK_GBPS, 4)}}


def host_cores() -> int:
    """The cores this process may run on: its CPU affinity, capped by the
    job's CPU share when the environment states one (OMP_NUM_THREADS)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, h, pairs, rounds=5, warmups=3):
    """The oracle's FP32 ref form (the reference's operation order) on the
    host: one pair per thread on every available core, 3 warm-up rounds,
    then the median of `rounds` rounds of two pairs per thread; the oracle's
    device form beside it, and single-thread latency per pair."""
    import oracle
    threads = host_cores()
    per_round = 2 * threads
    chosen = [pairs[(i * 37) % len(pairs)] for i in range(per_round)]

    def rate(mode, n_rounds, n_warm):
        for _ in range(n_warm):
            oracle.compute_many(mode, w, h, chosen[:threads], threads)
        times = []
        for _ in range(n_rounds):
            t0 = time.perf_counter()
            oracle.compute_many(mode, w, h, chosen, threads)
            times.append(time.perf_counter() - t0)
        return per_round / statistics.median(times), times

    value, times = rate(oracle.F32_REF, rounds, warmups)
    dev_value, _ = rate(oracle.F32_DEV, 3, 1)
    lat = []
    for a, b in chosen[:3]:
        t0 = time.perf_counter()
        oracle.compute(oracle.F32_REF, w, h, a, b)
        lat.append((time.perf_counter() - t0) * 1e3)
    return {"value": round(value, 3), "unit": "pairs/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "single_thread_ms_per_pair": round(statistics.median(lat), 1),
            "device_form_pairs_per_s": round(dev_value, 3),
            "round_s": [round(t, 3) for t in times],
            "sample": f"{per_round} pairs of the {w}x{h} workload per round ({threads} threads, one pair "
                      f"per thread at a time), oracle FP32 ref form; {warmups} warm-up rounds, median "
                      f"of {rounds}; latency: median of 3 pairs on one thread"}


def streaming_pool(w, h, cfg, map_form, workers_list=(1, 3), count=60, rounds=3):
    """Host images scored through Ssimulacra2Pool (Submit(Image, Image):
    each call copies both images to the device), round-robin over 1 and 3
    pipelines as vqm's bench workers do (bench/driver/ssimulacra2.cc:175-285),
    median pairs/s of `rounds` passes over `count` distinct pairs."""
    from vqm_amd import Image, ImageFormat, Size, Ssimulacra2Config, Ssimulacra2Pool
    pairs = make_pairs_cycled(w, h, cfg, count)
    size = Size(w, h)
    imgs = [(Image(size, ImageFormat.srgb8(), a), Image(size, ImageFormat.srgb8(), b)) for a, b in pairs]
    out = {"pairs": count, "map_form": map_name(0, map_form)}
    for workers in workers_list:
        pool = Ssimulacra2Pool(Ssimulacra2Config(size=size, map_form=map_form), workers)
        pool.ComputeMany(imgs[:max(workers, 4)])  # warm
        rates = []
        for _ in range(rounds):
            t0 = time.perf_counter()
            res = pool.ComputeMany(imgs)
            rates.append(count / (time.perf_counter() - t0))
        if not all(np.isfinite(r.Score()) for r in res):
            raise RuntimeError("non-finite streaming score")
        pool.close()
        out[f"workers_{workers}"] = {"pairs_per_s": round(statistics.median(rates), 1)}
    key = f"{w}x{h}"
    if key in RTX5070_STREAMING:
        out["vqm_rtx5070_workers_3"] = RTX5070_STREAMING[key]
    return out


def resident_workers(w, h, cfg, map_form, workers_list=(1, 3), iterations=300, rounds=5, strip_rows=0,
                     strip_cols=0):
    """vqm's resident protocol (bench/driver/ssimulacra2.cc:175-285), one
    pipeline per worker: each uploads the original twice and the distorted
    image once and alternates the two originals, so no prepared original is
    reused; the workers' pairs overlap round-robin (submit on every worker,
    then finish one and submit it again). Median pairs/s of `rounds` rounds
    of `iterations` pairs."""
    from vqm_amd import Image, ImageFormat, Size, Ssimulacra2Config, Ssimulacra2Pipeline
    a, b = make_pairs_cycled(w, h, cfg, 1)[0]
    size = Size(w, h)
    ia, ib = Image(size, ImageFormat.srgb8(), a), Image(size, ImageFormat.srgb8(), b)
    out = {"map_form": map_name(0, map_form)}
    for workers in workers_list:
        pipes = [Ssimulacra2Pipeline(Ssimulacra2Config(size=size, map_form=map_form, gpu_resident_images=3,
                                                       strip_rows=strip_rows, strip_cols=strip_cols))
                 for _ in range(workers)]
        handles = [(p.Upload(ia), p.Upload(ia), p.Upload(ib)) for p in pipes]
        flip = [0] * workers

        def submit(k):
            original = handles[k][flip[k]]
            flip[k] ^= 1
            return pipes[k].Submit(original, handles[k][2])

        for _ in range(3):
            for k in range(workers):
                pipes[k].Finish(submit(k))
        rates, score = [], None
        for _ in range(rounds):
            t0 = time.perf_counter()
            tokens = [submit(k) for k in range(min(workers, iterations))]
            for i in range(iterations):
                k = i % len(tokens)
                score = pipes[k].Finish(tokens[k]).Score()
                if i + len(tokens) < iterations:
                    tokens[k] = submit(k)
            rates.append(iterations / (time.perf_counter() - t0))
        if not np.isfinite(score):
            raise RuntimeError("non-finite resident score")
        for p in pipes:
            p.close()
        out[f"workers_{workers}"] = {"pairs_per_s": round(statistics.median(rates), 1)}
    key = f"{w}x{h}"
    if key in RTX5070_RESIDENT:
        out["vqm_rtx5070"] = RTX5070_RESIDENT[key]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="1080p")
    ap.add_argument("--pairs-per-strength", type=int, default=100)
    ap.add_argument("--planes", choices=sorted(PLANES), default="f32")
    ap.add_argument("--map", choices=sorted(MAPS), default="default",
                    help="SSIM' map form (default: the library's, the reference order for FP32 planes)")
    ap.add_argument("--strip-rows", type=int, default=0,
                    help="output rows per scale work item of the batch (0 = the library default)")
    ap.add_argument("--strip-cols", type=int, default=0, choices=[0, 64, 128],
                    help="output columns per scale work item of the batch (0 = the library default)")
    ap.add_argument("--chunks", type=int, default=1,
                    help="runs of pairs per batch whose scale overlaps the next run's prepare")
    ap.add_argument("--persistent", action="store_true",
                    help="size the kernels' grids to the resident occupancy (blocks loop over the work)")
    ap.add_argument("--single-strip-rows", type=int, default=0,
                    help="output rows per scale work item of the single-pair modes (0 = the library default)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-ladder", action="store_true", help="skip the uneven-size batch workload")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--serve-max-batch", type=int, default=64)
    ap.add_argument("--serve-budget-gib", type=int, default=8)
    ap.add_argument("--ladder-variants", type=int, default=2,
                    help="distinct pairs per ladder size and strength")
    ap.add_argument("--cpu-rounds", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-streaming", action="store_true", help="skip the Ssimulacra2Pool streaming legs")
    ap.add_argument("--no-modes", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the 4K batch and the other plane forms at the main workload")
    ap.add_argument("--general-decode-only", action="store_true",
                    help="run only the general-decode leg (8-bit 4:2:0 YUV beside 8-bit sRGB) and print it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run the plain copy kernel over MIB MiB (traffic counter reference)")
    args = ap.parse_args()

    if args.gpus != 1:
        log(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (pairs are independent; "
            "several GPUs would each score their own share of the pairs)")
        return 2

    from vqm_amd import (DeviceBuffer, Image, ImageFormat, Size, Ssimulacra2Batch,
                         Ssimulacra2Config, Ssimulacra2Pipeline, device_info)

    w, h, cfg = WORKLOADS[args.workload]
    fmt = PLANES[args.planes]
    map_form = MAPS[args.map]
    form = map_name(fmt, map_form)
    info = device_info()
    log(f"device {info['name']} {info['arch']} cus={info['cus']}")

    if args.copy_calibration:
        import ctypes
        from vqm_amd import _lib
        nbytes = args.copy_calibration << 20
        src_buf, dst_buf = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        for _ in range(3):
            status = _lib.lib().vqm_device_copy(ctypes.c_void_p(dst_buf.ptr), ctypes.c_void_p(src_buf.ptr), nbytes)
            if status != 0:
                log("copy calibration failed")
                return 1
        log(f"copy calibration: 3 copies of {args.copy_calibration} MiB")
        src_buf.free()
        dst_buf.free()

    t0 = time.perf_counter()
    pairs = make_pairs(w, h, cfg, args.pairs_per_strength)
    log(f"generated {len(pairs)} pairs of {w}x{h} in {time.perf_counter() - t0:.1f} s")
    if args.general_decode_only:
        print(json.dumps({"general_decode": general_decode(pairs, w, h, map_form)}), flush=True)
        return 0
    dev, refs = upload_pairs(pairs, w, h)
    log(f"uploaded {2 * len(pairs)} images ({dev.nbytes / 1e9:.2f} GB)")

    batch = Ssimulacra2Batch(max_pairs=len(refs), plane_format=fmt, strip_rows=args.strip_rows,
                             chunks=args.chunks, persistent=args.persistent, map_form=map_form,
                             strip_cols=args.strip_cols)
    for _ in range(args.warmup):
        scales, norms = batch.Finish(batch.Submit(refs), raw=True)
    step_ms, stages = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        scales, norms = batch.Finish(batch.Submit(refs), raw=True)
        step_ms.append((time.perf_counter() - t) * 1e3)
        stages.append(batch.StageTimes())
    total = time.perf_counter() - t_start
    if not np.all(np.isfinite(norms)):
        log("non-finite norms")
        return 1
    value = len(refs) * args.steps / total
    log(f"batch ({form} map form): {value:.1f} pairs/s, {total / args.steps * 1e3:.2f} ms/step")

    # Roofline of the dominant kernel: algorithmic bytes per launch over the
    # average HIP-event duration of that launch on the batch stream.
    n_px = w * h
    planes_bytes = plane_bytes(w, h, fmt)
    launch_bytes = {
        "prepare": len(refs) * (2 * 3 * n_px + 2 * planes_bytes),
        "scale": len(refs) * (2 * planes_bytes),
    }
    avg = {"prepare": statistics.mean(s[0] for s in stages),
           "scale": statistics.mean(s[1] for s in stages),
           "reduce": statistics.mean(s[2] for s in stages)}
    dominant = max(("prepare", "scale"), key=lambda k: avg[k])
    achieved = launch_bytes[dominant] / (avg[dominant] * 1e-3) / 1e9
    pair_bytes = 2 * 3 * n_px + 4 * planes_bytes
    traffic = profiled_traffic(dominant, args.planes, form, w, h, len(refs))
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                "traffic": traffic["bytes"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                # The roof that binds: wave64 VALU instructions issued per CU
                # per cycle (non-packed FP32 issues at most one), from the
                # same counter summary.
                "binding": "valu_issue" if traffic and traffic["valu_per_cu_cycle"] > 0.8 else "hbm",
                "valu_issue": ({"achieved": traffic["valu_per_cu_cycle"], "peak": 1.0,
                                "unit": "wave64 VALU instructions per CU-cycle",
                                "frac": traffic["valu_per_cu_cycle"]} if traffic else None),
                "bytes_per_launch": launch_bytes[dominant],
                "stage_ms": {k: round(v, 4) for k, v in avg.items()},
                "pair_algorithmic_bytes": pair_bytes,
                "whole_path_gbps": round(value * pair_bytes / 1e9, 1)}

    out = {
        "metric": "SSIMULACRA2 pairs/s at 1080p & 4K (% HBM peak); p99 latency at fixed QPS",
        "value": round(value, 2),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / RTX5070_1080P_RESIDENT, 3) if args.workload == "1080p" else None,
        "dtype": {0: "f32", 1: "f32 (f16 planes)", 2: "f32 (bf16 planes)",
                  3: "f32 (u16 fixed-point planes)"}[fmt],
        "data": "synthetic (vqm_amd/synth.py: sinusoid field + block quantisation + Gaussian noise)",
        "config": {"workload": f"{w}x{h} 8-bit sRGB RGB, 6 scales, {len(refs)} pairs per step "
                               f"(3 strengths x {args.pairs_per_strength}), resident in HBM, "
                               f"{args.planes} planes, {form} map form, one batched launch per kernel",
                   "pairs_per_step": len(refs), "width": w, "height": h, "map_form": form,
                   "strip_cols": args.strip_cols or "library default", "parallelism": "single GPU"},
        "roofline": roofline,
        "vs_baseline_note": "value / 3836.7 pairs/s (vqm 1080p resident, 3 workers, RTX 5070, BASELINE.md)",
    }

    if not args.no_secondary:
        sec_steps = max(3, args.steps // 4)
        # The other map form and the other plane form on the same resident
        # pairs, then the 4K batch in FP32 and the mixed form.
        other = 2 if form == "reference" else 1
        out["map_forms"] = {map_name(fmt, other): resident_batch(refs, w, h, fmt, sec_steps, 1, args.chunks,
                                                                 persistent=args.persistent, map_form=other)}
        log(f"map forms: {out['map_forms']}")
        # Both scale strip widths (one or two output columns per lane) in
        # both map forms, at the batch's default strip height.
        out["strip_widths"] = {
            f"{map_name(fmt, m)}_{cols}": resident_batch(refs, w, h, fmt, sec_steps, 1, args.chunks,
                                                         persistent=args.persistent, map_form=m,
                                                         strip_cols=cols)
            for m in (1, 2) for cols in (64, 128)}
        log(f"strip widths: {out['strip_widths']}")
        forms = {}
        for name, f in sorted(PLANES.items()):
            if f != fmt and name in ("f32", "u16"):
                forms[name] = resident_batch(refs, w, h, f, sec_steps, 1, args.chunks, persistent=args.persistent)
        out["plane_forms"] = forms
        log(f"plane forms: {forms}")
        if args.workload == "1080p":
            w4, h4, cfg4 = WORKLOADS["4k"]
            pairs4 = make_pairs_cycled(w4, h4, cfg4, 50)
            dev4, refs4 = upload_pairs(pairs4, w4, h4)
            out["workload_4k"] = {"workload": f"{w4}x{h4} 8-bit sRGB RGB, {len(refs4)} resident pairs per step "
                                              f"(strengths cycled)"}
            for name in ("f32", "u16"):
                out["workload_4k"][name] = resident_batch(refs4, w4, h4, PLANES[name], sec_steps, 1, args.chunks,
                                                          persistent=args.persistent)
            dev4.free()
            del pairs4
            log(f"4k: {out['workload_4k']}")

    if not args.no_secondary:
        out["general_decode"] = general_decode(pairs, w, h, map_form)

    if not args.no_streaming:
        out["streaming_pool"] = {"1920x1080": streaming_pool(1920, 1080, 1, map_form, count=90),
                                 "3840x2160": streaming_pool(3840, 2160, 2, map_form, count=30)}
        log(f"streaming pool: {out['streaming_pool']}")

    if not args.no_streaming:
        out["resident_workers"] = {"1920x1080": resident_workers(1920, 1080, 1, map_form, iterations=300),
                                   "3840x2160": resident_workers(3840, 2160, 2, map_form, iterations=100)}
        log(f"resident workers: {out['resident_workers']}")

    if not args.no_modes:
        modes = {}
        from vqm_amd import DeviceImage
        # Single pair at a time, resident inputs: per-pair latency, with the
        # pair's launches replayed from one HIP graph and launched eagerly.
        for key, graph in (("resident_single_pair_graph", True), ("resident_single_pair", False)):
            pipe = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(w, h), plane_format=fmt, map_form=map_form,
                                                         strip_rows=args.single_strip_rows, graph=graph))
            lat = []
            for i in range(min(60, len(refs))):
                a, b = refs[i][0], refs[i][1]
                t = time.perf_counter()
                pipe.Compute(DeviceImage(a, Size(w, h), ImageFormat.srgb8()), DeviceImage(b, Size(w, h), ImageFormat.srgb8()))
                lat.append((time.perf_counter() - t) * 1e3)
            lat = lat[5:]
            modes[key] = {"pairs_per_s": round(1e3 / statistics.mean(lat), 1),
                          "p50_ms": round(float(np.percentile(lat, 50)), 4),
                          "p99_ms": round(float(np.percentile(lat, 99)), 4)}
            if graph:
                pipe.close()
        # Streaming: host images uploaded per call (PCIe-inclusive).
        imgs = [(Image(Size(w, h), ImageFormat.srgb8(), a), Image(Size(w, h), ImageFormat.srgb8(), b))
                for a, b in pairs[:30]]
        lat = []
        for a, b in imgs:
            t = time.perf_counter()
            pipe.Compute(a, b)
            lat.append((time.perf_counter() - t) * 1e3)
        lat = lat[5:]
        modes["streaming_single_pair"] = {"pairs_per_s": round(1e3 / statistics.mean(lat), 1),
                                          "p50_ms": round(float(np.percentile(lat, 50)), 4),
                                          "p99_ms": round(float(np.percentile(lat, 99)), 4)}
        pipe.close()
        # Small batches of resident pairs, eager and replayed from captured
        # graphs: where launch overhead is a visible share of a batch.
        from vqm_amd import Ssimulacra2Batch as _Batch
        small = {}
        for n in (1, 4, 16):
            for graph in (False, True):
                b = _Batch(max_pairs=n, plane_format=fmt, map_form=map_form, graph=graph)
                chunk = refs[:n]
                for _ in range(3):
                    b.Finish(b.Submit(chunk), raw=True)
                reps = max(10, 240 // n)
                t = time.perf_counter()
                for _ in range(reps):
                    b.Finish(b.Submit(chunk), raw=True)
                dt = time.perf_counter() - t
                b.close()
                small[f"batch{n}_{'graph' if graph else 'eager'}"] = {
                    "pairs_per_s": round(n * reps / dt, 1), "ms_per_batch": round(dt / reps * 1e3, 4)}
        modes["small_batches"] = small
        out["modes"] = modes
        log(f"modes: {modes}")

    if not (args.no_serve and args.no_ladder):
        pool = LadderPool(3, args.ladder_variants, pinned=not args.no_serve)
        if not args.no_ladder:
            out["ladder_batch"] = ladder_batch(args, pool)
        if not args.no_serve:
            out["serving"] = serving(args, pool)
        del pool

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, h, pairs, rounds=args.cpu_rounds)
        log(f"cpu baseline: {out['cpu_baseline']}")

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
