"""The request server: continuous batching over several HIP streams with
HBM-budget admission. Each result is the same alone and in the mix (bit for
bit: a pair's arithmetic and reduction order do not depend on its batch),
host and device inputs agree, and admission holds the budget."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import random_pair
import oracle
import vqm_amd
from vqm_amd import (DeviceBuffer, Image, ImageFormat, InvalidArgument, PinnedBuffer, Size,
                     Ssimulacra2Config, Ssimulacra2Pipeline, Ssimulacra2Server)

pytestmark = pytest.mark.gpu

SHAPES = [(64, 48), (200, 120), (37, 29), (320, 180), (129, 97), (256, 256), (8, 8), (640, 360)]


def make_requests(n, seed):
    reqs = []
    for i in range(n):
        w, h = SHAPES[(i * 5 + seed) % len(SHAPES)]
        a, b = random_pair(w, h, seed=seed * 1000 + i)
        da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        reqs.append({"w": w, "h": h, "a": a, "b": b, "da": da, "db": db})
    return reqs


def alone(r, fmt):
    """The request through a pipeline at the server's strip height (64 rows)."""
    w, h = r["w"], r["h"]
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(w, h), plane_format=fmt, strip_rows=64))
    res = p.Compute(Image(Size(w, h), ImageFormat.srgb8(), r["a"]),
                    Image(Size(w, h), ImageFormat.srgb8(), r["b"]))
    p.close()
    return res.norms()


@pytest.mark.parametrize("fmt", [vqm_amd.PLANES_F32, vqm_amd.PLANES_BF16])
def test_each_result_is_the_same_alone_and_in_the_mix(fmt):
    reqs = make_requests(40, seed=3)
    expect = [alone(r, fmt) for r in reqs]
    with Ssimulacra2Server(plane_format=fmt, lanes=3, max_batch=8) as server:
        # Two submitting threads, so arrivals interleave with batches in flight.
        tickets = [None] * len(reqs)

        def submit(idx):
            for i in idx:
                r = reqs[i]
                tickets[i] = server.SubmitDevice(r["da"].ptr, r["db"].ptr, r["w"], r["h"])

        threads = [threading.Thread(target=submit, args=(range(k, len(reqs), 2),)) for k in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        results = [server.Wait(t) for t in tickets]
        stats = server.Stats()
    for i, (r, res) in enumerate(zip(reqs, results)):
        assert res.status == 0
        assert res.scales == len(vqm_amd.plan_scales(Size(r["w"], r["h"])))
        assert np.array_equal(res.norms, expect[i]), f"request {i} {r['w']}x{r['h']}"
        assert res.submit_ns <= res.launch_ns <= res.done_ns
        assert 1 <= res.batch_size <= 8
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    assert stats["batched_requests"] == len(reqs)
    assert stats["max_batch"] > 1, "no request was batched with another"
    # And the mix is the metric: one pair against the plane format's oracle
    # form (the default map form).
    r = reqs[3]
    _, ref, _ = oracle.compute(oracle.form_for(fmt), r["w"], r["h"], r["a"], r["b"])
    np.testing.assert_allclose(results[3].norms, ref, rtol=1e-10, atol=1e-300)


def test_host_inputs_match_device_inputs():
    reqs = make_requests(12, seed=5)
    with Ssimulacra2Server(lanes=2, max_batch=4) as server:
        dev = [server.SubmitDevice(r["da"].ptr, r["db"].ptr, r["w"], r["h"]) for r in reqs]
        pinned = []
        host = []
        for r in reqs:
            pa, pb = PinnedBuffer(r["a"].nbytes), PinnedBuffer(r["b"].nbytes)
            pa.array[:] = r["a"]
            pb.array[:] = r["b"]
            pinned += [pa, pb]
            size = Size(r["w"], r["h"])
            host.append(server.SubmitHost(Image(size, ImageFormat.srgb8(), pa.array),
                                          Image(size, ImageFormat.srgb8(), pb.array)))
        # Pageable host memory works too (the copy is then synchronous).
        pageable = [server.SubmitHost(Image(Size(r["w"], r["h"]), ImageFormat.srgb8(), r["a"]),
                                      Image(Size(r["w"], r["h"]), ImageFormat.srgb8(), r["b"]))
                    for r in reqs[:4]]
        a = [server.Wait(t).norms for t in dev]
        b = [server.Wait(t).norms for t in host]
        c = [server.Wait(t).norms for t in pageable]
    for i in range(len(reqs)):
        assert np.array_equal(a[i], b[i]), f"request {i}"
    for i in range(4):
        assert np.array_equal(a[i], c[i])


def test_admission_holds_the_hbm_budget():
    reqs = make_requests(16, seed=7)
    probe = Ssimulacra2Server(lanes=1, max_batch=1)
    r0 = reqs[0]
    probe.Wait(probe.SubmitDevice(r0["da"].ptr, r0["db"].ptr, r0["w"], r0["h"]))
    probe.close()
    # Room for the largest workspace about twice: most requests wait for
    # workspaces held by batches in flight, and all complete.
    largest = max(reqs, key=lambda r: r["w"] * r["h"])
    one = Ssimulacra2Server(lanes=1, max_batch=1)
    one.Wait(one.SubmitDevice(largest["da"].ptr, largest["db"].ptr, largest["w"], largest["h"]))
    need = one.Stats()["peak_pool_bytes"]
    one.close()
    budget = 2 * need + need // 2
    expect = [alone(r, 0) for r in reqs]
    with Ssimulacra2Server(lanes=3, max_batch=4, hbm_budget=budget) as server:
        tickets = [server.SubmitDevice(r["da"].ptr, r["db"].ptr, r["w"], r["h"]) for r in reqs]
        results = [server.Wait(t) for t in tickets]
        stats = server.Stats()
    for i, res in enumerate(results):
        assert res.status == 0 and np.array_equal(res.norms, expect[i])
    assert 0 < stats["peak_pool_bytes"] <= budget
    # A request that cannot fit even alone fails by itself; the rest go on.
    with Ssimulacra2Server(lanes=2, max_batch=4, hbm_budget=need // 2) as server:
        small = min(reqs, key=lambda r: r["w"] * r["h"])
        t_big = server.SubmitDevice(largest["da"].ptr, largest["db"].ptr, largest["w"], largest["h"])
        t_small = server.SubmitDevice(small["da"].ptr, small["db"].ptr, small["w"], small["h"])
        with pytest.raises(InvalidArgument, match="HBM budget"):
            server.Wait(t_big)
        assert server.Wait(t_small).status == 0
        assert server.Stats()["failed"] == 1


def test_replay_and_misuse():
    reqs = make_requests(30, seed=9)
    expect = [alone(r, 0) for r in reqs]
    rng = np.random.default_rng(1)
    arrivals = np.cumsum(rng.exponential(1e9 / 2000.0, len(reqs)))  # 2000 requests/s
    with Ssimulacra2Server(lanes=3, max_batch=16) as server:
        out = server.Replay([(r["da"].ptr, r["db"].ptr, r["w"], r["h"], False, t)
                             for r, t in zip(reqs, arrivals)])
        for i, res in enumerate(out):
            assert res.status == 0 and np.array_equal(res.norms, expect[i])
            assert res.done_ns > res.submit_ns and res.latency_ms < 5000
        with pytest.raises(InvalidArgument):
            server.Wait(123456789)
        with pytest.raises(InvalidArgument, match="8x8"):
            server.SubmitDevice(reqs[0]["da"].ptr, reqs[0]["db"].ptr, 4, 8)
        with pytest.raises(InvalidArgument):
            server.SubmitDevice(0, reqs[0]["db"].ptr, 8, 8)
        t = server.SubmitDevice(reqs[0]["da"].ptr, reqs[0]["db"].ptr, reqs[0]["w"], reqs[0]["h"])
        server.Wait(t)
        with pytest.raises(InvalidArgument):
            server.Wait(t)  # a ticket is consumed once


def test_ladder_requests_are_the_same_alone_and_in_the_mix():
    """The serving workload's request mix: 60 requests whose sizes follow
    the ladder weights 0.1/0.2/0.35/0.2/0.15 (640x360 to 3840x2160, the
    synthetic serving generator), submitted from two threads to a server with
    three lanes; every result equals the same pair scored alone."""
    from vqm_amd import DeviceImage, synth
    reqs = []
    for i in range(60):
        (w, h), strength = synth.ladder_pair_shape(4, i, synth.SERVE_WEIGHTS)
        a, b = synth.pair(4, i, w, h, strength)
        da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        reqs.append({"w": w, "h": h, "da": da, "db": db})
    assert len({(r["w"], r["h"]) for r in reqs}) == 5
    print(f"ladder serving: {len(reqs)} requests", flush=True)
    with Ssimulacra2Server(lanes=3, max_batch=16, hbm_budget=8 << 30) as server:
        tickets = [None] * len(reqs)

        def submit(idx):
            for i in idx:
                r = reqs[i]
                tickets[i] = server.SubmitDevice(r["da"].ptr, r["db"].ptr, r["w"], r["h"])

        threads = [threading.Thread(target=submit, args=(range(k, len(reqs), 2),)) for k in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        results = [server.Wait(t) for t in tickets]
        stats = server.Stats()
    assert stats["completed"] == len(reqs) and stats["failed"] == 0 and stats["max_batch"] > 1
    pipes = {}
    for i, (r, res) in enumerate(zip(reqs, results)):
        size = (r["w"], r["h"])
        if size not in pipes:
            pipes[size] = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(*size), strip_rows=64))
        want = pipes[size].Compute(DeviceImage(r["da"].ptr, Size(*size), ImageFormat.srgb8()),
                                   DeviceImage(r["db"].ptr, Size(*size), ImageFormat.srgb8())).norms()
        assert res.status == 0
        assert np.array_equal(res.norms, want), f"request {i} {size}"
    for p in pipes.values():
        p.close()
    for r in reqs:
        r["da"].free()
        r["db"].free()
