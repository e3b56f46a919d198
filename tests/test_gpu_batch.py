"""Batched scoring: pairs of uneven sizes in one launch give the same result
as each pair alone (the pool_test.cc:27-62 pattern, here bit for bit), and
the workspace pool respects its HBM budget."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import random_pair
import vqm_amd
from vqm_amd import (DeviceBuffer, Image, ImageFormat, InvalidArgument, LogicError, Size,
                     Ssimulacra2Batch, Ssimulacra2Config, Ssimulacra2Pipeline)

pytestmark = pytest.mark.gpu


def upload_pairs(shapes, seed=0):
    pairs, keep, host = [], [], []
    for i, (w, h) in enumerate(shapes):
        a, b = random_pair(w, h, seed=seed + i)
        da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        keep += [da, db]
        host.append((a, b))
        pairs.append((da.ptr, db.ptr, w, h))
    return pairs, keep, host


def alone(w, h, a, b, fmt=0, strip_rows=64, strip_cols=0):
    """The pair through a pipeline; at 64-row strips by default, the height
    the batches below are pinned to (a pipeline's own default follows the
    geometry)."""
    p = Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(w, h), plane_format=fmt, strip_rows=strip_rows,
                                              strip_cols=strip_cols))
    r = p.Compute(Image(Size(w, h), ImageFormat.srgb8(), a), Image(Size(w, h), ImageFormat.srgb8(), b))
    p.close()
    return r.norms()


@pytest.mark.parametrize("fmt", [vqm_amd.PLANES_F32, vqm_amd.PLANES_F16])
def test_mixed_batch_equals_each_pair_alone(fmt):
    shapes = [(64, 64), (200, 130), (33, 17), (640, 360), (64, 64), (129, 97), (8, 8)]
    pairs, keep, host = upload_pairs(shapes, seed=11)
    # Bit for bit with a pipeline at the same strip height (64 rows); the
    # batch default (chosen per geometry) groups the FP64 strip sums
    # differently.
    batch = Ssimulacra2Batch(max_pairs=16, plane_format=fmt, strip_rows=64)
    scales, norms = batch.Finish(batch.Submit(pairs), raw=True)
    tall = Ssimulacra2Batch(max_pairs=16, plane_format=fmt)
    _, tall_norms = tall.Finish(tall.Submit(pairs), raw=True)
    tall.close()
    np.testing.assert_allclose(tall_norms, norms, rtol=1e-12, atol=0)
    for i, ((w, h), (a, b)) in enumerate(zip(shapes, host)):
        assert np.array_equal(norms[i], alone(w, h, a, b, fmt)), f"pair {i} {w}x{h}"
        assert scales[i] == len(vqm_amd.plan_scales(Size(w, h)))
    # Again, in another order: workspaces are reused from the pool.
    before = batch.PoolBytes()
    order = [3, 0, 6, 2, 5, 1, 4]
    scales2, norms2 = batch.Finish(batch.Submit([pairs[i] for i in order]), raw=True)
    for k, i in enumerate(order):
        assert np.array_equal(norms2[k], norms[i])
    assert batch.PoolBytes() == before


def test_uniform_batch_equals_each_pair_alone():
    shapes = [(320, 180)] * 5
    pairs, keep, host = upload_pairs(shapes, seed=21)
    batch = Ssimulacra2Batch(max_pairs=8, strip_rows=64, graph=False)
    _, norms = batch.Finish(batch.Submit(pairs), raw=True)
    for i, (a, b) in enumerate(host):
        assert np.array_equal(norms[i], alone(320, 180, a, b))
    t = batch.StageTimes()
    assert all(v > 0 for v in t)
    # The default replays submissions of up to 8 pairs from a graph: the
    # same norms, and the whole graph timed as the prepare stage.
    auto = Ssimulacra2Batch(max_pairs=8, strip_rows=64)
    for _ in range(3):
        _, again = auto.Finish(auto.Submit(pairs), raw=True)
        assert np.array_equal(again, norms)
        t = auto.StageTimes()
        assert t[0] > 0 and t[1] == 0 and t[2] == 0
    auto.close()


def test_batch_budget_and_misuse():
    pairs, keep, _ = upload_pairs([(256, 256)] * 3, seed=31)
    batch = Ssimulacra2Batch(max_pairs=2, hbm_budget=1, strip_rows=64)
    with pytest.raises(InvalidArgument):
        batch.Submit(pairs[:1])           # one workspace exceeds a 1-byte budget
    big = Ssimulacra2Batch(max_pairs=2, strip_rows=64)
    with pytest.raises(InvalidArgument):
        big.Submit(pairs)                 # more than max_pairs
    token = big.Submit(pairs[:2])
    with pytest.raises(LogicError):
        big.Submit(pairs[:1])
    big.Finish(token)
    with pytest.raises(InvalidArgument):
        big.Finish(token)
    with pytest.raises(InvalidArgument):
        big.Submit([(pairs[0][0], pairs[0][1], 4, 4)])  # below 8x8


@pytest.mark.parametrize("chunks", [2, 3, 16])
def test_chunked_batch_equals_each_pair_alone(chunks):
    """Runs of pairs whose scale overlaps the next run's prepare on a second
    stream: uniform and mixed batches, more chunks than pairs included."""
    for shapes in ([(160, 90)] * 7, [(64, 64), (200, 130), (33, 17), (640, 360), (129, 97)]):
        pairs, keep, host = upload_pairs(shapes, seed=51)
        batch = Ssimulacra2Batch(max_pairs=8, chunks=chunks, strip_rows=64)
        _, norms = batch.Finish(batch.Submit(pairs), raw=True)
        for i, ((w, h), (a, b)) in enumerate(zip(shapes, host)):
            assert np.array_equal(norms[i], alone(w, h, a, b)), f"pair {i} {w}x{h}"
        assert all(v > 0 for v in batch.StageTimes())
        batch.close()
    with pytest.raises(InvalidArgument):
        Ssimulacra2Batch(max_pairs=2, chunks=17)


def test_autotune_picks_a_shape_and_keeps_results():
    """The tuner times every (strip height, strip width, chunks) shape and
    caches the fastest; a tuned shape keeps every pixel's arithmetic and FP32
    block sum, so the norms agree to the order of the FP64 sums over strips
    (relative 1e-10)."""
    first = vqm_amd.autotune(Size(320, 180), vqm_amd.PLANES_F32, pairs=6)
    assert first["strip_rows"] in (16, 32, 64, 128, 192, 256) and first["chunks"] in (1, 2, 4, 8)
    assert first["strip_cols"] in (64, 128)
    assert first["candidates"] == 6 * 2 * 4 and not first["cached"]
    assert 0.0 < first["ms_per_launch"] <= first["default_ms"] * 1.0001
    again = vqm_amd.autotune(Size(320, 180), vqm_amd.PLANES_F32, pairs=5)  # same size class
    assert again["cached"] and again["strip_rows"] == first["strip_rows"]
    assert again["strip_cols"] == first["strip_cols"]
    shapes = [(320, 180)] * 6 + [(64, 48), (200, 120)]
    pairs, keep, host = upload_pairs(shapes, seed=61)
    plain = Ssimulacra2Batch(max_pairs=8)
    tuned = Ssimulacra2Batch(max_pairs=8, autotune=True)
    _, a = plain.Finish(plain.Submit(pairs), raw=True)
    _, b = tuned.Finish(tuned.Submit(pairs), raw=True)
    np.testing.assert_allclose(b, a, rtol=1e-10, atol=0)
    with pytest.raises(InvalidArgument):
        vqm_amd.autotune(Size(4, 4))


@pytest.mark.parametrize("chunks", [1, 2])
def test_mixed_strip_widths_in_one_launch(chunks):
    """Geometries pinned to different strip shapes share one submission (one
    scale launch per strip width); each pair equals a pipeline of its shape
    bit for bit. A used geometry's shape cannot change."""
    shapes = [(320, 180), (200, 130), (320, 180), (64, 48), (200, 130)]
    pinned = {(320, 180): (32, 64), (200, 130): (64, 128), (64, 48): (16, 64)}
    pairs, keep, host = upload_pairs(shapes, seed=71)
    batch = Ssimulacra2Batch(max_pairs=8, chunks=chunks)
    for (w, h), (rows, cols) in pinned.items():
        batch.SetShape(Size(w, h), rows, cols)
    _, norms = batch.Finish(batch.Submit(pairs), raw=True)
    for i, ((w, h), (a, b)) in enumerate(zip(shapes, host)):
        rows, cols = pinned[(w, h)]
        assert np.array_equal(norms[i], alone(w, h, a, b, strip_rows=rows, strip_cols=cols)), f"pair {i}"
    with pytest.raises(LogicError):
        batch.SetShape(Size(320, 180), 64, 128)
    # A pinned shape takes precedence over the autotuner.
    tuned = Ssimulacra2Batch(max_pairs=8, chunks=chunks, autotune=True)
    tuned.SetShape(Size(200, 130), 16, 64)
    _, got = tuned.Finish(tuned.Submit(pairs), raw=True)
    for i, ((w, h), (a, b)) in enumerate(zip(shapes, host)):
        if (w, h) == (200, 130):
            assert np.array_equal(got[i], alone(w, h, a, b, strip_rows=16, strip_cols=64)), f"pinned pair {i}"
    tuned.close()
    with pytest.raises(InvalidArgument):
        batch.SetShape(Size(100, 100), 12, 128)
    with pytest.raises(InvalidArgument):
        batch.SetShape(Size(100, 100), 64, 32)
    batch.close()


@pytest.mark.parametrize("fmt", [vqm_amd.PLANES_F32, vqm_amd.PLANES_U16])
def test_persistent_grid_matches_one_block_per_item(fmt):
    """A persistent launch (blocks = resident occupancy, each looping over
    tiles and strips) gives the same norms bit for bit as one block per work
    item, for a uniform batch larger than the resident grid and for a mixed
    one; chunked persistent launches too."""
    for shapes in ([(1920, 1080)] * 6, [(64, 64), (1280, 720), (33, 17), (640, 360), (8, 8), (1000, 37)]):
        pairs, keep, host = upload_pairs(shapes, seed=31)
        plain = Ssimulacra2Batch(max_pairs=8, plane_format=fmt)
        _, want = plain.Finish(plain.Submit(pairs), raw=True)
        plain.close()
        for chunks in (1, 3):
            pers = Ssimulacra2Batch(max_pairs=8, plane_format=fmt, persistent=True, chunks=chunks)
            _, got = pers.Finish(pers.Submit(pairs), raw=True)
            _, again = pers.Finish(pers.Submit(pairs), raw=True)
            pers.close()
            assert np.array_equal(got, want), (shapes[0], chunks)
            assert np.array_equal(again, want)


def test_ladder_batch_at_workload_sizes(oracle_lib):
    """The uneven-size batch at its workload's sizes: 40 pairs drawn from the
    five-size ladder (640x360 to 3840x2160, the synthetic batch generator)
    in one launch per kernel. Each pair is bit-equal to the same pair alone
    through a pipeline at the batch's strip height, the default strip height
    changes only the FP64 summation grouping, and one pair of each size is
    bit-level with its oracle form (the FP32 ref form: the default map form)
    and within the contract of FP64."""
    from conftest import SCORE_ABS, assert_features_close
    from vqm_amd import DeviceImage, synth
    picks = [synth.ladder_pair_shape(3, i, None) for i in range(40)]
    for k, size in enumerate(synth.LADDER):  # every size at least once
        picks[k] = (size, synth.STRENGTHS[k % 3])
    host = [synth.pair(3, i, w, h, s) for i, ((w, h), s) in enumerate(picks)]
    keep, pairs = [], []
    for ((w, h), _), (a, b) in zip(picks, host):
        da, db = DeviceBuffer(a.nbytes), DeviceBuffer(b.nbytes)
        da.upload(a)
        db.upload(b)
        keep += [da, db]
        pairs.append((da.ptr, db.ptr, w, h))
    print(f"ladder batch: {len(pairs)} pairs", flush=True)
    batch = Ssimulacra2Batch(max_pairs=64, strip_rows=64)
    scales, norms = batch.Finish(batch.Submit(pairs), raw=True)
    batch.close()
    tall = Ssimulacra2Batch(max_pairs=64)
    _, tall_norms = tall.Finish(tall.Submit(pairs), raw=True)
    tall.close()
    np.testing.assert_allclose(tall_norms, norms, rtol=1e-12, atol=0)
    pipes = {size: Ssimulacra2Pipeline(Ssimulacra2Config(size=Size(*size), strip_rows=64)) for size in synth.LADDER}
    checked = set()
    for i, (((w, h), _), (a, b)) in enumerate(zip(picks, host)):
        alone_norms = pipes[(w, h)].Compute(DeviceImage(pairs[i][0], Size(w, h), ImageFormat.srgb8()),
                                           DeviceImage(pairs[i][1], Size(w, h), ImageFormat.srgb8())).norms()
        assert np.array_equal(norms[i], alone_norms), f"pair {i} {w}x{h}"
        assert scales[i] == 6
        if (w, h) in checked:
            continue
        checked.add((w, h))
        n, ref, _ = oracle_lib.compute(oracle_lib.F32_REF, w, h, a, b)
        np.testing.assert_allclose(norms[i], ref, rtol=1e-10, atol=0)
        _, f64, _ = oracle_lib.compute(oracle_lib.F64, w, h, a, b)
        assert abs(oracle_lib.score(n, norms[i]) - oracle_lib.score(n, f64)) <= SCORE_ABS
    for p in pipes.values():
        p.close()
    for d in keep:
        d.free()
