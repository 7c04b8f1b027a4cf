"""simulate_sand without the mid-step wait (lus_sand_set_list_capacity) and
replayed as hipGraphs (lus_sand_set_graph): the same bits as the waiting step
over steps with the attract and blow flags switched, for simulate_sand and
simulate_sand_v3; a capacity the lists exceed is reported at the next
synchronising call, with the step's kernels writing nothing past it."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu

FLAGS = [(False, False, None), (False, False, None), (True, False, (8.0, 6.0, 8.0)),
         (True, False, (8.0, 6.0, 8.0)), (False, True, (9.0, 5.0, 7.0)), (False, False, None),
         (False, False, None), (False, False, None), (False, False, None)]


def _scene(n_side=10, seed=5):
    pos = scenes.jittered_lattice(n_side, n_side, n_side, origin=(4.0, 3.0, 4.0), amplitude=0.2, seed=seed)
    sol = scenes.lattice(20, 2, 20, origin=(0.0, 0.0, 0.0))
    return pos, sol


def _run(gpu_mod, variant, capacity, graph, steps=len(FLAGS), box=(20, 20, 20), dt=0.016):
    pos, sol = _scene()
    n, ns = pos.shape[0], sol.shape[0]
    total = max(n + ns, box[0] * box[1] * box[2])
    solid_index = np.arange(total - 1, total - 1 - ns, -1, dtype=np.int32)
    sp = gpu_mod.sand_params(*box, kernel_radius_scale=3.1)
    sp.ptr_solid_start = total - ns
    with gpu_mod.Gpu(0) as g:
        g.sand_init(sp, pos, None, None, sol, solid_index)
        g.sand_set_variant(variant)
        if capacity:
            g.sand_set_list_capacity(capacity)
        if graph:
            g.sand_set_graph(True)
        for k in range(steps):
            a, b, pl = FLAGS[k % len(FLAGS)]
            g.sand_step(np.float32(dt), a, b, pl)
        g.sync()
        got = g.sand_download()
        off, idx = g.sand_neighbors()
    return got, off, idx


@pytest.mark.parametrize("variant", [0, 2])
def test_fixed_capacity_and_graphs_match_waiting_step(gpu_mod, variant):
    ref = _run(gpu_mod, variant, 0, False)
    cap = 4 * int(ref[1][-1]) + 64
    for graph in (False, True):
        got = _run(gpu_mod, variant, cap, graph)
        for a, b in zip(got[0], ref[0]):
            assert np.array_equal(a, b)
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])


def test_capacity_overflow_is_reported(gpu_mod):
    pos, sol = _scene()
    box = (20, 20, 20)
    n, ns = pos.shape[0], sol.shape[0]
    total = max(n + ns, box[0] * box[1] * box[2])
    sp = gpu_mod.sand_params(*box, kernel_radius_scale=3.1)
    sp.ptr_solid_start = total - ns
    solid_index = np.arange(total - 1, total - 1 - ns, -1, dtype=np.int32)
    with gpu_mod.Gpu(0) as g:
        g.sand_init(sp, pos, None, None, sol, solid_index)
        with pytest.raises(gpu_mod.LusError):
            g.sand_set_graph(True)  # needs a fixed capacity
        g.sand_set_list_capacity(100)  # ~30 entries per particle are needed
        g.sand_step(np.float32(0.016), False, False, None)
        with pytest.raises(gpu_mod.LusError, match="capacity"):
            g.sand_download()
        # reported once: the mark is cleared; a new scene steps normally
        g.sand_init(sp, pos, None, None, sol, solid_index)
        g.sand_step(np.float32(0.016), False, False, None)
        x = g.sand_download()[0]
        assert np.isfinite(x).all()


def test_no_wait_check_starts_clear_after_a_v3_error(gpu_mod):
    """Contexts made after one that reported a simulate_sand_v3 out-of-grid
    prediction (and freed its flag with the error still marked): their
    no-wait steps report nothing, whatever memory their flag reuses."""
    pos = scenes.jittered_lattice(2, 2, 2, origin=(6.0, 3.0, 6.0), amplitude=0.2, seed=1)
    sp = gpu_mod.sand_params(20, 20, 20, kernel_radius_scale=3.1)
    vel = np.zeros_like(pos)
    vel[0] = [0.0, -5000.0, 0.0]
    for _ in range(3):
        with gpu_mod.Gpu(0) as g:
            g.sand_init(sp, pos, vel, None, None, None)
            g.sand_set_variant(2)
            with pytest.raises(gpu_mod.LusError):
                g.sand_step(np.float32(0.016), False, False, None)
        for variant in (0, 2):
            with gpu_mod.Gpu(0) as g:
                g.sand_init(sp, pos, None, None, None, None)
                g.sand_set_variant(variant)
                g.sand_set_list_capacity(4096)
                g.sand_step(np.float32(0.016), False, False, None)
                g.sync()
                assert np.isfinite(g.sand_download()[0]).all()
