"""Phase times in the reference's profiling slots (src/profiling/Profiling.hpp,
src/Simulate.cpp:669-802): which slots a profiled step fills, their
consistency, and that profiling changes no result."""
import numpy as np
import pytest

from lustrine_amd import scenes

pytestmark = pytest.mark.gpu

HZ = 2.7e9


def _check_slots(G):
    cyc, sec = G.profile_slots()
    assert sec[0] > 0.0 and sec[0] == sec[2]
    assert sec[1] == 0.0 and sec[5] == 0.0
    assert sec[3] > 0.0 and sec[4] > 0.0 and sec[6] > 0.0
    assert sec[3] + sec[4] + sec[6] <= sec[2] * (1 + 1e-6)
    for c, t in zip(cyc, sec):
        assert c == int(round(t * HZ))
    return sec


def test_sand_step_fills_slots(gpu_mod, oracle):
    pos = scenes.jittered_lattice(12, 12, 12, origin=(2.0, 3.0, 2.0), amplitude=0.25, seed=7)
    sp = gpu_mod.sand_params(20, 20, 20)
    out = []
    for prof in (False, True):
        with gpu_mod.Gpu(0) as g:
            g.sand_init(sp, pos)
            g.profiling(prof)
            for _ in range(3):
                g.sand_step(np.float32(0.016))
            out.append(g.sand_download())
            if prof:
                sec = _check_slots(gpu_mod)
                # an unprofiled context's step leaves the slots as they are
                with gpu_mod.Gpu(0) as h:
                    h.sand_init(sp, pos)
                    h.sand_step(np.float32(0.016))
                assert gpu_mod.profile_slots()[1] == sec
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_pbf_step_fills_slots(gpu_mod):
    cfg = scenes.cfg0()
    x0 = cfg.positions()
    out = []
    for prof in (False, True):
        with gpu_mod.Gpu(0) as g:
            g.pbf_init(gpu_mod.pbf_params(cfg), x0, None)
            g.pbf_set_graph(True)
            g.profiling(prof)
            g.pbf_step(np.float32(cfg.dt), 5)
            out.append(g.pbf_download(order=1)[:2])
            if prof:
                _check_slots(gpu_mod)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
