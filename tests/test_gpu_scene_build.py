"""Scene construction on the device (init_chunk_from_grid,
src/Lustrine.cpp:676-722; the voxel model -> grid step of
init_grid_from_magika_voxel, src/VoxelLoader.cpp:47-100) against the oracle's
restatement: positions, colours and counts bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sub,mask_out", [(1, False), (2, False), (3, True), (1, True)])
def test_chunk_from_grid_matches_oracle(gpu_mod, oracle, sub, mask_out):
    rng = np.random.default_rng(sub * 10 + mask_out)
    X, Y, Z = 13, 7, 11
    cells = rng.integers(0, 4, X * Y * Z).astype(np.int32) * (rng.random(X * Y * Z) < 0.6)
    colors = rng.random((X * Y * Z, 4)).astype(np.float32)
    pos0 = (2.25, -1.5, 7.125)
    with gpu_mod.Gpu(0) as g:
        gp, gc, gk = g.chunk_from_grid(cells, X, Y, Z, pos0, 0.37, sub, mask_out, colors=colors)
    op, oc, ok = oracle.chunk_from_grid(cells, X, Y, Z, pos0, np.float32(0.37), sub, mask_out, colors=colors)
    assert gk == ok and np.array_equal(gp, op) and np.array_equal(gc, oc)


def test_chunk_from_grid_large_box(gpu_mod, oracle):
    # a 64 x 32 x 64 box, subdivision 2: 1,048,576 particles
    X, Y, Z = 64, 32, 64
    cells = np.ones(X * Y * Z, np.int32)
    with gpu_mod.Gpu(0) as g:
        gp, _, gk = g.chunk_from_grid(cells, X, Y, Z, (0.5, 0.5, 0.5), 1.0, 2)
    op, _, ok = oracle.chunk_from_grid(cells, X, Y, Z, (0.5, 0.5, 0.5), np.float32(1.0), 2)
    assert gk == ok == 1 << 20 and np.array_equal(gp, op)


def test_chunk_random_box_overcount_and_errors(gpu_mod, oracle):
    # init_grid_box_random counts every cell: the chunk's tail stays zero
    rng = np.random.default_rng(1)
    X, Y, Z = 6, 5, 4
    cells = (rng.random(X * Y * Z) <= 0.5).astype(np.int32)
    with gpu_mod.Gpu(0) as g:
        gp, _, gk = g.chunk_from_grid(cells, X, Y, Z, (0, 0, 0), 0.5, 2, num_occupied=X * Y * Z)
        op, _, ok = oracle.chunk_from_grid(cells, X, Y, Z, (0, 0, 0), np.float32(0.5), 2, num_occupied=X * Y * Z)
        assert gk == ok and gp.shape == op.shape == (X * Y * Z * 8, 3) and np.array_equal(gp, op)
        with pytest.raises(gpu_mod.LusError):  # fewer entries than points
            g.chunk_from_grid(cells, X, Y, Z, (0, 0, 0), 0.5, 2, num_occupied=1)


def test_grid_from_voxels_matches_oracle(gpu_mod, oracle):
    rng = np.random.default_rng(2)
    sx, sy, sz = 17, 9, 23
    vox = (rng.integers(1, 256, sx * sy * sz) * (rng.random(sx * sy * sz) < 0.3)).astype(np.uint8)
    pal = rng.integers(0, 256, (256, 4)).astype(np.uint8)
    with gpu_mod.Gpu(0) as g:
        gc, gcol, gocc = g.grid_from_voxels(vox, sx, sy, sz, pal)
    oc, ocol, oocc = oracle.grid_from_voxels(vox, sx, sy, sz, pal)
    assert gocc == oocc and np.array_equal(gc, oc) and np.array_equal(gcol, ocol)
