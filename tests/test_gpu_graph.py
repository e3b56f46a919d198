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
