"""CPU tests of the tiled kernel's host-side schedule (fslbm_tile_plan):
which tiles hold fluid (work items run on those only), the cells one step
moves, and the default chunk of the step-time model (rounds of one
workgroup per CU, 256 CUs in 8 XCDs, 3 planes of start-up per item,
chunks of at least 4 planes)."""
import numpy as np
import pytest

from flowcontrast_amd import inputs, tile_plan
from flowcontrast_amd._lib import FslbmError

PRECISIONS = ("f64", "f32", "f16s", "bf16s")


def rows_of(prec, periodic=False, blend=0.0, shape=None):
    """Tile rows: 12 for FP32, 4 for FP64 (on the pipelined kernel); the
    16-bit storage modes take 12 in a walled box at blend 0 and 8 otherwise.  An FP32 periodic
    box of `shape` takes 8 where the model rates the best 8-row schedule
    (at 0.79 of a 12-row tile-plane's time per tile-plane) faster."""
    if prec == "f32":
        if periodic and shape is not None:
            nz, ny, nx = shape

            def best(r):
                T = -(-nx // 64) * -(-ny // r)
                ln = model_chunk(nz, T)
                return model_cost(nz, T, -(-nz // ln))
            return 8 if best(8) * 0.79 < best(12) else 12
        return 12
    if prec in ("f16s", "bf16s") and not periodic and blend == 0.0:
        return 12
    if prec == "f64":
        return 4
    return 8


def brute_tiles(mask2d, rows):
    """Tiles of 64 columns x rows holding a fluid cell, and their cells."""
    ny, nx = mask2d.shape
    tiles = cells = 0
    for y0 in range(0, ny, rows):
        for x0 in range(0, nx, 64):
            blk = mask2d[y0:y0 + rows, x0:x0 + 64]
            if (blk == 0).any():
                tiles += 1
                cells += blk.size
    return tiles, cells


def model_cost(nz, ntile, nch, cus=256):
    items = ntile * nch
    per_xcd = -(-items // 8)
    rounds = -(-per_xcd // (cus // 8))
    return rounds * (-(-nz // nch) + 3)


def model_chunk(nz, ntile):
    """Cheapest chunk length; ties go to the longer chunk."""
    best = None
    for nch in range(1, nz // min(4, nz) + 1):
        ln = -(-nz // nch)
        if -(-nz // ln) != nch:
            continue
        key = (model_cost(nz, ntile, nch), -ln)
        if best is None or key < best[0]:
            best = (key, ln)
    return best[1]


def test_pipe_r120_cross_section():
    """The 256x256 cross-section of the R = 120 pipe: 78 of 88 tiles of
    64 x 12, 112 of 128 tiles of 64 x 8 and 224 of 256 tiles of 64 x 4
    hold fluid."""
    mask, _ = inputs.circle_mask(256, 256, 1, 120.0)
    nz = 1024
    for prec, blend, tiles in (("f32", 0.0, 78), ("f16s", 0.0, 78), ("f16s", 1.0, 112),
                               ("bf16s", 0.0, 78), ("bf16s", 1.0, 112), ("f64", 0.0, 224)):
        plan = tile_plan((nz, 256, 256), prec, mask, blend)
        rows = rows_of(prec, blend=blend)
        assert plan["tile_rows"] == rows
        assert plan["tiles"] == tiles
        assert plan["cells"] == tiles * 64 * rows * nz
        assert plan["chunk"] == model_chunk(nz, tiles)
    assert tile_plan((nz, 256, 256), "f32", mask)["chunk"] == 79
    assert tile_plan((nz, 256, 256), "f16s", mask)["chunk"] == 79
    assert tile_plan((nz, 256, 256), "bf16s", mask)["chunk"] == 79


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_masks_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    for nx, ny in ((70, 45), (130, 37), (200, 120), (9, 9)):
        m = np.full((ny, nx), 125, dtype=np.uint8)
        # a few fluid blobs away from the x/y faces
        for _ in range(rng.integers(1, 4)):
            cx, cy = rng.uniform(1, nx - 1), rng.uniform(1, ny - 1)
            r = rng.uniform(1.0, max(1.5, min(nx, ny) / 4))
            y, x = np.mgrid[0:ny, 0:nx]
            m[((x - cx) ** 2 + (y - cy) ** 2 <= r * r) & (x > 0) & (y > 0)
              & (x < nx - 1) & (y < ny - 1)] = 0
        nz = int(rng.integers(16, 80))
        for prec in PRECISIONS:
            for blend in (0.0, 0.5):
                plan = tile_plan((nz, ny, nx), prec, m, blend)
                rows = rows_of(prec, blend=blend)
                assert plan["tile_rows"] == rows
                tiles, cells = brute_tiles(m, rows)
                if tiles == 0:  # no fluid: one item, which stores nothing
                    tiles, cells = 1, min(64, nx) * min(rows, ny)
                assert plan["tiles"] == tiles, (nx, ny, prec)
                assert plan["cells"] == cells * nz, (nx, ny, prec)
                assert plan["chunk"] == model_chunk(nz, tiles), (nx, ny, nz, prec)


def test_periodic_box_steps_every_tile():
    for prec in PRECISIONS:
        plan = tile_plan((40, 88, 88), prec, None)
        rows = rows_of(prec, periodic=True, shape=(40, 88, 88))
        assert plan["tile_rows"] == rows
        assert plan["tiles"] == 2 * -(-88 // rows)
        assert plan["cells"] == 88 * 88 * 40


def test_fp32_periodic_box_rows_follow_the_model():
    """Measured FP32 Taylor-Green rates: 8-row tiles run the 256^3 box 7%
    faster (128 tiles in two 128-plane chunks fill one round of 256
    workgroups) and 192^3 10% slower; the model picks accordingly."""
    expect = {(256, 256, 256): 8, (192, 192, 192): 12, (128, 128, 128): 12,
              (96, 96, 96): 12, (1024, 256, 256): 12}
    for shape, rows in expect.items():
        assert rows_of("f32", periodic=True, shape=shape) == rows
        plan = tile_plan(shape, "f32", None)
        assert plan["tile_rows"] == rows, shape
        nz, ny, nx = shape
        assert plan["tiles"] == -(-nx // 64) * -(-ny // rows)
        assert plan["chunk"] == model_chunk(nz, plan["tiles"])


def test_all_wall_mask_keeps_one_item():
    m = np.full((30, 100), 125, dtype=np.uint8)
    plan = tile_plan((20, 30, 100), "f32", m)
    assert plan["tiles"] == 1
    assert plan["cells"] == 64 * 12 * 20


def test_chunk_model_shape():
    """Small domains get the shortest chunk (4 planes: their items run side
    by side); a domain that fills several rounds gets long chunks."""
    mask, _ = inputs.circle_mask(32, 32, 1, 14.0)
    assert tile_plan((64, 32, 32), "f64", mask)["chunk"] == 4
    assert tile_plan((64, 32, 32), "f32", mask)["chunk"] == 4
    big = tile_plan((1024, 256, 256), "f32", None)
    assert big["tiles"] == 88 and big["chunk"] == model_chunk(1024, 88) >= 16
    # nz below the 4-plane minimum: one chunk
    assert tile_plan((3, 32, 32), "f32", None)["chunk"] == 3


def test_invalid_arguments():
    with pytest.raises(FslbmError):
        tile_plan((0, 32, 32), "f32", None)
    with pytest.raises(FslbmError):
        tile_plan((8, 32, 32), "f32", None, float("nan"))
