"""The launch-shape autotuner's candidate lists (host logic; the timing of the
candidates needs a GPU and runs in bench.py and the GPU tests).

A one-wave-strip filter launch is `strips x ceil(h / rows)` tiles, one
pass-A workgroup each, and a CU holds one such workgroup: the candidates
include, besides fixed row counts, the row counts that make the tiles k
whole rounds over the device's CUs (k = 1 .. 4), so that the last round is
not a few tiles on an otherwise idle chip.
"""
import pytest

from unblending_amd import autotune as A


def _tiles(w, h, r, block, rows):
    tw = block - 2 * r
    return -(-w // tw) * -(-h // rows)


@pytest.mark.parametrize("w,h,r,cus", [(4096, 4096, 32, 256), (1920, 1080, 16, 256), (2000, 3000, 60, 304),
                                       (640, 480, 8, 256)])
def test_balanced_rows_fill_whole_rounds(w, h, r, cus):
    for block in A.FILTER_WAVE_BLOCKS:
        if block - 2 * r < 64:
            continue
        rows_list = A.balanced_rows(w, h, r, block, cus)
        assert rows_list, (block, r)
        assert len(set(rows_list)) == len(rows_list)
        for rows in rows_list:
            assert rows >= 32
            # whole rounds: no more tiles than k rounds hold, for a k of 1 .. 4
            tiles = _tiles(w, h, r, block, rows)
            assert any(tiles <= k * cus for k in (1, 2, 3, 4)), (block, rows, tiles)
        # the one-round candidate leaves less than a row of strips idle
        # (guaranteed while the 32-row floor does not set it and a round is
        # at most 32 strip rows: then ceil(h / rows) is the floor(cus / strips)
        # strip rows asked for)
        rows1 = rows_list[0]
        strips = -(-w // (block - 2 * r))
        if rows1 > 32 and cus // strips <= 32:
            assert cus - strips < _tiles(w, h, r, block, rows1) <= cus, (block, rows1)


def test_balanced_rows_at_the_bench_shape():
    # 4096 x 4096, r = 32, 256-column strips (192 output columns: 22 strips)
    # on 256 CUs: 11 strip rows of 373 (242 tiles, one round), 23 of 179
    # (506 tiles, two rounds)
    rows = A.balanced_rows(4096, 4096, 32, 256, 256)
    assert rows[:2] == [373, 179]
    assert _tiles(4096, 4096, 32, 256, 373) == 242
    assert _tiles(4096, 4096, 32, 256, 179) == 506


def test_balanced_rows_degenerate_inputs():
    assert A.balanced_rows(4096, 4096, 32, 256, 0) == []
    assert A.balanced_rows(4096, 4096, 128, 256, 256) == []  # no output columns
    assert A.balanced_rows(100, 20, 4, 128, 256) == [32]      # a small image: the 32-row floor


@pytest.mark.parametrize("r", [1, 9, 32, 63, 96])
def test_wave_strip_candidates(r):
    cands = A.filter_candidates(r, 4096, 4096, 256)
    blocks = {b for b, _ in cands}
    assert blocks == {b for b in A.FILTER_WAVE_BLOCKS if b - 2 * r >= 64}
    assert len(set(cands)) == len(cands)
    for b in blocks:
        rows = [rw for bb, rw in cands if bb == b]
        assert set(A.FILTER_ROWS) <= set(rows)
        assert set(A.balanced_rows(4096, 4096, r, b, 256)) <= set(rows)


@pytest.mark.parametrize("r", [97, 150, 245, 480])
def test_workgroup_strip_candidates(r):
    cands = A.filter_candidates(r, 4096, 4096, 256)
    assert cands
    assert all(b in A.FILTER_BLOCKS for b, _ in cands)
    assert all(b - 2 * r >= 64 for b, _ in cands) or {b for b, _ in cands} == {1024}
    assert {rw for _, rw in cands} == set(A.FILTER_ROWS)
