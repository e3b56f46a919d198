"""Graph-captured single-pair pipelines: every replay equals the eager
launches bit for bit, with eager work of other pipelines and host
synchronisations between the replays, for every input kind (resident,
streaming, caller-owned device images with a stream to wait on)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import random_pair
import vqm_amd
from vqm_amd import (DeviceBuffer, DeviceImage, Image, ImageFormat, Size, Ssimulacra2Config,
                     Ssimulacra2Pipeline)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", [vqm_amd.PLANES_F32, vqm_amd.PLANES_F16])
@pytest.mark.parametrize("w,h", [(8, 8), (200, 130), (640, 360)])
def test_graph_replays_equal_eager_launches(w, h, fmt):
    size = Size(w, h)
    pairs = [random_pair(w, h, seed=40 + i) for i in range(6)]
    eager = Ssimulacra2Pipeline(Ssimulacra2Config(size=size, plane_format=fmt, graph=False))
    graph = Ssimulacra2Pipeline(Ssimulacra2Config(size=size, plane_format=fmt, graph=True))
    other = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(96, 64), graph=False))
    oa, ob = random_pair(96, 64, seed=99)
    img = lambda d: Image(size, ImageFormat.srgb8(), d)  # noqa: E731
    bufs = []
    for i, (a, b) in enumerate(pairs):
        want = eager.Compute(img(a), img(b)).norms()
        # Streaming input.
        assert np.array_equal(graph.Compute(img(a), img(b)).norms(), want), f"streaming pair {i}"
        other.Compute(Image(Size(96, 64), ImageFormat.srgb8(), oa), Image(Size(96, 64), ImageFormat.srgb8(), ob))
        vqm_amd.ssimulacra2.L.lib().vqm_device_synchronize()
        # Caller-owned device images at new addresses every time.
        da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        bufs += [da, db]
        got = graph.Compute(DeviceImage(da.ptr, size, ImageFormat.srgb8()), DeviceImage(db.ptr, size, ImageFormat.srgb8())).norms()
        assert np.array_equal(got, want), f"device pair {i}"
        # Resident slots.
        ua, ub = graph.Upload(img(a)), graph.Upload(img(b))
        assert np.array_equal(graph.Compute(ua, ub).norms(), want), f"resident pair {i}"
        graph.Release(ua)
        graph.Release(ub)
    # A second round over the same pairs in reverse order replays the same graph.
    for i, (a, b) in reversed(list(enumerate(pairs))):
        assert np.array_equal(graph.Compute(img(a), img(b)).norms(), eager.Compute(img(a), img(b)).norms())


def _device_pairs(shapes, seed):
    pairs, keep = [], []
    for i, (w, h) in enumerate(shapes):
        a, b = random_pair(w, h, seed=seed + i)
        da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        keep += [da, db]
        pairs.append((da.ptr, db.ptr, w, h))
    return pairs, keep


@pytest.mark.parametrize("map_form", [vqm_amd.MapForm.Reference, vqm_amd.MapForm.Difference])
def test_batch_graph_replays_equal_eager_batches(map_form):
    """A batch with graphs replays one captured graph per (geometry, pair
    count): every replay equals the eager batch bit for bit, with other eager
    batches and host synchronisations between the replays, over fresh
    workspaces and inputs at new addresses; a mixed batch and a chunked one
    run eagerly and still agree."""
    eager = vqm_amd.Ssimulacra2Batch(max_pairs=8, map_form=map_form)
    graph = vqm_amd.Ssimulacra2Batch(max_pairs=8, map_form=map_form, graph=True)
    other = vqm_amd.Ssimulacra2Batch(max_pairs=4)
    opairs, okeep = _device_pairs([(96, 64)] * 4, seed=7)
    for rnd in range(4):
        pairs, keep = _device_pairs([(320, 180)] * (5 + rnd % 2), seed=100 + 10 * rnd)
        _, want = eager.Finish(eager.Submit(pairs), raw=True)
        for rep in range(3):
            _, got = graph.Finish(graph.Submit(pairs), raw=True)
            assert np.array_equal(got, want), f"round {rnd} replay {rep}"
            assert graph.StageTimes()[0] > 0.0
            other.Finish(other.Submit(opairs), raw=True)
            vqm_amd.ssimulacra2.L.lib().vqm_device_synchronize()
    mixed, mkeep = _device_pairs([(320, 180), (64, 48), (200, 130)], seed=300)
    _, want = eager.Finish(eager.Submit(mixed), raw=True)
    _, got = graph.Finish(graph.Submit(mixed), raw=True)
    assert np.array_equal(got, want)
    chunked = vqm_amd.Ssimulacra2Batch(max_pairs=8, map_form=map_form, graph=True, chunks=2)
    pairs, keep = _device_pairs([(320, 180)] * 6, seed=400)
    _, want = eager.Finish(eager.Submit(pairs), raw=True)
    _, got = chunked.Finish(chunked.Submit(pairs), raw=True)
    assert np.array_equal(got, want)
    for b in (eager, graph, other, chunked):
        b.close()


def test_server_graph_results_equal_eager_server():
    """A server whose lanes replay graphs returns, for each request, the
    norms of the same request on an eager server."""
    shapes = [(320, 180)] * 12 + [(200, 130)] * 6
    pairs, keep = _device_pairs(shapes, seed=500)
    results = []
    for graph in (False, True):
        with vqm_amd.Ssimulacra2Server(lanes=2, max_batch=4, graph=graph) as server:
            got = []
            for _ in range(3):
                tickets = [server.SubmitDevice(a, b, w, h) for a, b, w, h in pairs]
                got.append([server.Wait(t).norms for t in tickets])
            results.append(got)
    for rnd in range(3):
        for i in range(len(pairs)):
            assert np.array_equal(results[1][rnd][i], results[0][0][i]), f"round {rnd} request {i}"
