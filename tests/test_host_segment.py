"""Segmented estimate() on the host (utils/utils.m:592-1003, 1348-1540):
voxel-balanced slabs, halo ranges, cropping and reassembly."""
import numpy as np
import pytest

from gacelle_amd import segment


def test_balanced_boundaries_and_halo():
    mask = np.zeros((4, 4, 10), bool)
    mask[:, :, 2:8] = True                       # 6 slices of 16 voxels
    owned = segment.build_balanced_boundaries(mask, 32, 3)
    assert [list(o) for o in owned] == [[0, 1, 2, 3], [4, 5], [6, 7, 8, 9]]
    counts = [int(mask[:, :, o].sum()) for o in owned]
    assert counts == [32, 32, 32]
    seg = segment.expand_segments_with_halo(owned, 1, 10)
    assert list(seg[0]["fit"]) == [0, 1, 2, 3, 4] and list(seg[0]["local"]) == [0, 1, 2, 3]
    assert list(seg[1]["fit"]) == [3, 4, 5, 6] and list(seg[1]["local"]) == [1, 2]
    assert list(seg[2]["fit"]) == [5, 6, 7, 8, 9] and list(seg[2]["local"]) == [1, 2, 3, 4]


def test_find_segments_user_and_memory():
    mask = np.ones((8, 8, 12), bool)
    f = dict(NSegmentUser=3, memoryBudget=1 << 40)
    assert len(segment.find_optimal_segment_3D(mask, f, 100)) == 3
    assert len(segment.find_optimal_segment_3D(mask, dict(memoryBudget=1 << 40), 100)) == 1
    # a budget that holds about a third of the voxels
    bpv = 1000
    budget = segment.FIXED_BYTES + bpv * mask.sum() // 3 + 1
    segs = segment.find_optimal_segment_3D(mask, dict(memoryBudget=budget), bpv)
    assert len(segs) >= 3
    assert sorted(np.concatenate([s["owned"] for s in segs]).tolist()) == list(range(12))
    with pytest.raises(Exception):
        segment.find_optimal_segment_3D(mask, dict(memoryBudget=segment.FIXED_BYTES // 2), bpv)


def test_crop_and_restore_round_trip():
    rng = np.random.default_rng(0)
    full = rng.random((3, 2, 9, 4))
    mask = np.ones((3, 2, 9), bool)
    segs = segment.expand_segments_with_halo(segment.build_balanced_boundaries(mask, 18, 3), 1, 9)
    out = None
    for k, s in enumerate(segs):
        o = {"final": {"p": full[:, :, s["fit"]], "loss": float(k)}, "gd": np.ones((5, 2 * s["owned"].size))}
        out = segment.restore_segment_structure(out, segment.crop_segment_output(o, s), s["owned"], k, 9)
    np.testing.assert_array_equal(out["final"]["p"], full)
    assert out["final"]["loss"] == [0.0, 1.0, 2.0]
    assert out["gd"].shape == (5, 18)
