"""The shipped network (lustrine_amd/data/warm_start.npz) on the 1,048,576-particle
dam break (128 x 64 x 128 lattice in a 256 x 128 x 256 box, h = 2, dt = 1/60),
110 steps from the initial state, compared over steps 11-110 (the block has
fallen and spreads; over the first ~30 steps every solver's error is small and
the network gains nothing): what the network gains over the same step without
it, and the rerun bound.

Error measure: per step, e_t = max_i max(0, rho_i / rho0 - 1) at the step's
final positions, recorded on the GPU (lus_pbf_learned_stats).  The gain is
measured against the plain solver with the same number of Jacobi iterations
(the learned step minus the network) and against the plain 4-iteration solve.
The margins below leave room for a retrained network of the same quality; the
printed figures are the measured ones.
"""
import numpy as np
import pytest

from lustrine_amd import learned as L
from lustrine_amd import scenes

pytestmark = pytest.mark.gpu

STEPS = 110
WINDOW = slice(10, STEPS)
LEARNED_ITERS = 3   # plain iterations after the network's correction


def _run(g, gpu_mod, cfg, x0, iterations, learned_iters=None, bound=np.inf, exact=False):
    g.pbf_init(gpu_mod.pbf_params(cfg, iterations=iterations, exact=exact), x0)
    if learned_iters is None:
        g.pbf_set_mode(learned=False, check=True)
    else:
        g.pbf_set_mode(learned=True, learned_iterations=learned_iters, bound=bound)
    g.pbf_step(np.float32(cfg.dt), STEPS)
    n, reruns, errs, errm = g.pbf_learned_stats(with_mean=True)
    assert n == STEPS and errs.shape == (STEPS,)
    assert np.isfinite(errs).all()
    return reruns, errs[WINDOW].astype(np.float64), errm[WINDOW].astype(np.float64)


def test_shipped_network_gain_on_1m(gpu_mod):
    cfg = scenes.cfg2()
    x0 = cfg.positions()
    net, meta = L.load()
    with gpu_mod.Gpu(0) as g:
        g.pbf_set_network(net)
        _, e3, m3 = _run(g, gpu_mod, cfg, x0, LEARNED_ITERS)
        _, e4, m4 = _run(g, gpu_mod, cfg, x0, 4)
        reruns, el, ml = _run(g, gpu_mod, cfg, x0, 4, learned_iters=LEARNED_ITERS)
        # the same step without the network (zero correction, same
        # iterations) is the plain solver's, bit for bit in exact mode
        _, e3x, _ = _run(g, gpu_mod, cfg, x0, LEARNED_ITERS, exact=True)
        g.pbf_set_network(np.concatenate([net[:64], np.zeros(net.shape[0] - 64, np.float32)]))
        _, ez, _ = _run(g, gpu_mod, cfg, x0, 4, learned_iters=LEARNED_ITERS, exact=True)
    print(f"1m steps 11-{STEPS}: e mean/max  plain K=3 {e3.mean():.4f}/{e3.max():.4f}  plain K=4 "
          f"{e4.mean():.4f}/{e4.max():.4f}  learned {el.mean():.4f}/{el.max():.4f}; mean compression "
          f"K=3 {m3.mean():.5f} K=4 {m4.mean():.5f} learned {ml.mean():.5f} "
          f"(teacher {meta['teacher_iterations']} iterations)")
    assert reruns == 0
    assert np.array_equal(ez, e3x)
    # the network's gain over the same step without it
    assert el.mean() <= 0.6 * e3.mean()
    assert el.max() <= 0.6 * e3.max()
    assert ml.mean() <= 0.5 * m3.mean()
    # and over the plain 4-iteration solve, at 3 iterations + the network
    assert el.mean() < e4.mean() and el.max() < e4.max()


def test_shipped_network_rerun_bound_on_1m(gpu_mod):
    cfg = scenes.cfg2()
    x0 = cfg.positions()
    net, _ = L.load()
    with gpu_mod.Gpu(0) as g:
        g.pbf_set_network(net)
        _, e4, m4 = _run(g, gpu_mod, cfg, x0, 4, exact=True)
        # bound 1%: every step of this scene exceeds it, so every step reruns
        # the plain 4 iterations and the trajectory is the plain one bit for bit
        r1, e1, m1 = _run(g, gpu_mod, cfg, x0, 4, learned_iters=LEARNED_ITERS, bound=0.01, exact=True)
        assert r1 == STEPS
        assert np.array_equal(e1, e4) and np.array_equal(m1, m4)
        # a bound inside the learned step's error range: some steps rerun; a
        # step that kept the learned solution ended within the bound, so the
        # steps above the bound are at most the reruns
        _, el, _ = _run(g, gpu_mod, cfg, x0, 4, learned_iters=LEARNED_ITERS)
        bound = float(np.median(el))
        rb, eb, _ = _run(g, gpu_mod, cfg, x0, 16, learned_iters=LEARNED_ITERS, bound=bound)
    above = int((eb > bound).sum())
    print(f"1m x{STEPS}: bound 1% reruns {r1}/{STEPS}; bound {bound:.4f} (median learned e, steps 11-{STEPS}) "
          f"reruns {rb}/{STEPS}, steps above the bound {above}, e max {eb.max():.4f}")
    assert 0 < rb < STEPS
    assert above <= rb
