"""simulate_sand_credits and simulate_sand_v3 on the GPU (lus_sand_set_variant)
against the oracle's restatements (orc_simulate_sand_variant), through the C
ABI: permutation, neighbour lists, attracted flags, positions and velocities
bit for bit over several steps, with the attract / blow flags switched."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu


def _compare(gpu_mod, oracle, variant, pos, sol, att, steps, flags=None, box=(20, 20, 20), dt=0.016):
    s = oracle.init_constants(*box, 0.5, 1.0, 3.1)
    st = oracle.SandStream(s, pos, sol)
    n, ns = pos.shape[0], (0 if sol is None else sol.shape[0])
    st.att[:n] = att
    total = st.total
    solid_index = np.arange(total - 1, total - 1 - ns, -1, dtype=np.int32)
    sp = gpu_mod.sand_params(*box, kernel_radius_scale=3.1)
    sp.ptr_solid_start = total - ns if ns else total - 1
    with gpu_mod.Gpu(0) as g:
        g.sand_init(sp, pos, None, att, sol, solid_index)
        g.sand_set_variant(variant)
        for k in range(steps):
            fl = flags[k] if flags else (False, False, None)
            st.st.attract_flag, st.st.blow_flag = int(fl[0]), int(fl[1])
            if fl[2] is not None:
                st.st.player_position[:] = list(fl[2])
            st.step(np.float32(dt), variant=variant)
            g.sand_step(np.float32(dt), fl[0], fl[1], fl[2])
        got = g.sand_download()
        goff, gidx = g.sand_neighbors()
    want = (st.p[:n], st.ps[:n], st.v[:n], st.att[:n], st.perm())
    ooff, oidx = st.neighbors()
    assert np.array_equal(got[4], want[4])  # counting-sort permutation
    assert np.array_equal(goff, ooff) and np.array_equal(gidx, oidx)  # lists, in order
    assert np.array_equal(got[3], want[3])  # attracted
    for a, b, what in zip(got[:3], want[:3], ("x", "x*", "v")):
        assert np.array_equal(a, b), what
    return got, want


def _scene(n_side=10, seed=5):
    pos = scenes.jittered_lattice(n_side, n_side, n_side, origin=(4.0, 3.0, 4.0), amplitude=0.2, seed=seed)
    sol = scenes.lattice(20, 2, 20, origin=(0.0, 0.0, 0.0))
    return pos, sol


def test_credits_steps_match_reference(gpu_mod, oracle):
    pos, sol = _scene()
    att = np.where(np.arange(pos.shape[0]) % 3 == 0, 2, 0).astype(np.int32)
    got, want = _compare(gpu_mod, oracle, 1, pos, sol, att, steps=5)
    # some credits released at contact, the rest kept
    assert (want[3] == 2).any() or (att == 2).sum() > (want[3] == 2).sum()


def test_credits_attract_and_blow(gpu_mod, oracle):
    pos, sol = _scene(8, 9)
    att = np.where(np.arange(pos.shape[0]) % 2 == 0, 2, 0).astype(np.int32)
    player = np.array([8.0, 6.0, 8.0], np.float32)
    flags = [(True, False, player), (True, False, player), (False, True, player), (False, False, player)]
    _compare(gpu_mod, oracle, 1, pos, sol, att, steps=4, flags=flags)


def test_v3_steps_match_reference(gpu_mod, oracle):
    pos, sol = _scene(10, 7)
    _compare(gpu_mod, oracle, 2, pos, sol, np.zeros(pos.shape[0], np.int32), steps=5, dt=0.02)


def test_v3_walls_and_flags(gpu_mod, oracle):
    # a block against the x = 0 wall (the half-space projection acts), with
    # the unbounded attract / blow pushes of v3
    pos = scenes.jittered_lattice(6, 6, 6, origin=(0.3, 2.0, 6.0), amplitude=0.1, seed=3)
    sol = scenes.lattice(20, 1, 20, origin=(0.0, 0.0, 0.0))
    player = np.array([3.0, 4.0, 9.0], np.float32)
    flags = [(False, False, None), (True, False, player), (False, True, player)]
    _compare(gpu_mod, oracle, 2, pos, sol, np.zeros(pos.shape[0], np.int32), steps=3, flags=flags,
             dt=0.005)


@pytest.mark.parametrize("variant", [1, 2])
def test_variants_ragged(gpu_mod, oracle, variant):
    for lat in [(1, 1, 1), (1, 1, 2), (5, 13, 1)]:
        pos = scenes.jittered_lattice(*lat, origin=(6.0, 3.0, 6.0), amplitude=0.2, seed=sum(lat))
        sol = scenes.lattice(20, 2, 20, origin=(0.0, 0.0, 0.0))
        att = np.full(pos.shape[0], 2 if variant == 1 else 0, np.int32)
        _compare(gpu_mod, oracle, variant, pos, sol, att, steps=3)


def test_v3_prediction_outside_grid_is_reported(gpu_mod, oracle):
    pos = scenes.jittered_lattice(2, 2, 2, origin=(6.0, 3.0, 6.0), amplitude=0.2, seed=1)
    sp = gpu_mod.sand_params(20, 20, 20, kernel_radius_scale=3.1)
    vel = np.zeros_like(pos)
    vel[0] = [0.0, -5000.0, 0.0]
    with gpu_mod.Gpu(0) as g:
        g.sand_init(sp, pos, vel, None, None, None)
        g.sand_set_variant(2)
        with pytest.raises(gpu_mod.LusError):
            g.sand_step(np.float32(0.016), False, False, None)
        assert "outside the grid" in gpu_mod.last_error()
