"""Learned warm start on the GPU: features, BF16-MFMA network kernels against
the CPU oracle (oracle/mlp.py), training, and the on-device fallback to the
plain step.  No reference counterpart (the reference's fluid solve is one
uncorrected pass, src/Simulate.cpp:83-154); the oracle is pinned by
tests/test_oracle_mlp.py."""
import numpy as np
import pytest

from lustrine_amd import scenes
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _scene(seed=5):
    cfg = scenes.small(10, 8, 10, box=(20, 16, 20), seed=seed)
    x0 = cfg.positions()
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(-0.8, 0.8, x0.shape).astype(np.float32)
    return cfg, x0, v0


def _predict_sorted(cfg, x, v):
    """x* and the cell order of the step, in the library's float32 operations."""
    dt = np.float32(cfg.dt)
    g = np.asarray(cfg.gravity, np.float32)
    r = np.float32(cfg.radius)
    lo = np.full(3, r, np.float32)
    hi = np.asarray(cfg.box, np.float32) - r
    vv = v + g * dt
    xp = x + vv * dt
    xp = np.minimum(np.maximum(xp, lo), hi).astype(np.float32)
    cs = np.float32(cfg.cell_size)
    gx, gy, gz = cfg.grid
    c = (xp / cs).astype(np.int32)
    key = c[:, 1] * gx * gz + c[:, 0] * gz + c[:, 2]
    order = np.argsort(key, kind="stable")
    return xp, order, lo, hi


def test_features_and_targets_match_oracle(gpu_mod, oracle):
    cfg, x0, v0 = _scene()
    p = gpu_mod.pbf_params(cfg, exact=True)
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(p, x0, v0)
        t = g.trainer(cfg.n)
        assert t.record_step(np.float32(cfg.dt), 1) == cfg.n
        feat, tgt = t.download()
        xn, _, ids = g.pbf_download(order=0)
    xp, order, lo, hi = _predict_sorted(cfg, x0, v0)
    assert np.array_equal(ids, order)
    want = M.features(xp[order], x0[order], cfg.h, cfg.kernel_factor, cfg.mass, cfg.rest_density(), lo, hi)
    assert np.abs(feat - want).max() < 2e-5
    # targets: the plain step's final position minus x*, bit for bit
    assert np.array_equal(tgt, xn - xp[order])
    # the recorded step is an ordinary plain step
    ox, ov = x0.copy(), v0.copy()
    oid = np.arange(cfg.n, dtype=np.int32)
    oracle.pbf_step(oracle.pbf_params(cfg), ox, ov, oid, np.float32(cfg.dt))
    assert np.array_equal(ox, xn)


def test_record_stride_selects_rotating_slots(gpu_mod):
    cfg, x0, v0 = _scene(6)
    p = gpu_mod.pbf_params(cfg, exact=True)
    dt = np.float32(cfg.dt)
    with gpu_mod.Gpu(0) as ga, gpu_mod.Gpu(0) as gb:
        ga.pbf_init(p, x0, v0)
        gb.pbf_init(p, x0, v0)
        ta, tb = ga.trainer(4 * cfg.n), gb.trainer(4 * cfg.n)
        full = []
        for k in range(3):
            ta.record_step(dt, 1)
            n_b = tb.record_step(dt, 3)
            off = k % 3
            p0 = (3 - off) % 3
            assert n_b == len(range(p0, cfg.n, 3))
        fa, ya = ta.download()
        fb, yb = tb.download()
    rows = []
    for k in range(3):
        p0 = (3 - k % 3) % 3
        rows.append(np.arange(p0, cfg.n, 3) + k * cfg.n)
    sel = np.concatenate(rows)
    assert np.array_equal(fb, fa[sel]) and np.array_equal(yb, ya[sel])


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


def test_inference_kernel_matches_oracle(gpu_mod):
    rng = np.random.default_rng(7)
    params = M.random_params(rng)
    net = np.concatenate([rng.normal(0, 0.3, 32), rng.uniform(0.5, 2.0, 32), params, [0.03]]).astype(np.float32)
    for rows in (1, 17, 1000, 4099):
        feat = rng.normal(size=(rows, 32)).astype(np.float32)
        with gpu_mod.Gpu(0) as g:
            out = g.mlp_infer(net, feat)
        emu = M.infer(net, feat, bf16_round=True)
        exact = M.infer(net, feat, bf16_round=False)
        assert _rel(out, emu) < 5e-3, rows
        assert _rel(out, exact) < 5e-2, rows


@pytest.mark.parametrize("rows", [300, 1000])
def test_training_gradient_matches_oracle(gpu_mod, rows):
    rng = np.random.default_rng(rows)
    params = M.random_params(rng)
    X = rng.normal(size=(rows, 32)).astype(np.float32)
    Y = rng.normal(size=(rows, 3)).astype(np.float32)
    with gpu_mod.Gpu(0) as g:
        loss, grad = g.mlp_loss_grad(params.astype(np.float32), X, Y, 0.5)
    # oracle with the GPU's parameter rounding to float32
    p32 = params.astype(np.float32).astype(np.float64)
    l_emu, g_emu = M.loss_grad(p32, X, Y, 0.5, bf16_round=True)
    l_64, g_64 = M.loss_grad(p32, X, Y, 0.5, bf16_round=False)
    assert loss == pytest.approx(l_emu, rel=1e-3)
    blocks = {"W1": (M.W1, M.B1), "b1": (M.B1, M.W2), "W2": (M.W2, M.B2), "b2": (M.B2, M.W3),
              "W3": (M.W3, M.B3), "b3": (M.B3, M.N_PARAMS)}
    for name, (a, b) in blocks.items():
        assert _rel(grad[a:b], g_emu[a:b]) < 1e-2, name
        assert _rel(grad[a:b], g_64[a:b]) < 6e-2, name


def test_fit_recovers_a_teacher_network(gpu_mod):
    rng = np.random.default_rng(11)
    teacher = M.random_params(rng)
    X = rng.normal(size=(24000, 32)).astype(np.float32)
    Y = (M.forward(teacher, X, bf16_round=False)[0] * 0.02).astype(np.float32)
    with gpu_mod.Gpu(0) as g:
        t = g.trainer(X.shape[0])
        t.add(X, Y)
        tl, vl = t.fit(epochs=60, batch=256, lr=3e-3, lr_final=1e-4, val_fraction=0.1, seed=2)
        net = t.network()
        Xt = rng.normal(size=(2000, 32)).astype(np.float32)
        pred = g.mlp_infer(net, Xt)
    assert np.isfinite(tl).all() and np.isfinite(vl).all()
    assert vl[-1] < 0.25 * vl[0] and tl[-1] < 0.25 * tl[0]
    want = M.forward(teacher, Xt, bf16_round=False)[0] * 0.02
    assert _rel(pred, want) < 0.5
    assert net[M.NET_OUT_SCALE] == pytest.approx(np.sqrt(np.mean(Y.astype(np.float64) ** 2)), rel=1e-3)


def _plain(gpu_mod, cfg, x0, steps, iterations=None):
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, iterations=iterations, exact=True), x0)
        g.pbf_step(np.float32(cfg.dt), steps)
        return g.pbf_download(order=1)


def _net(params, out_scale=1.0):
    return np.concatenate([np.zeros(32), np.ones(32), params, [out_scale]]).astype(np.float32)


@pytest.mark.parametrize("graph", [False, True])
def test_rerun_gives_plain_bits(gpu_mod, graph):
    cfg = scenes.small(10, 10, 10, box=(20, 20, 20), seed=8)
    x0 = cfg.positions()
    steps = 6
    px, pv, _ = _plain(gpu_mod, cfg, x0, steps)
    net = _net(M.random_params(np.random.default_rng(9)), 0.05)
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=True), x0)
        g.pbf_set_mode(learned=False, check=True)
        g.pbf_step(np.float32(cfg.dt), steps)
        _, _, perrs, perrm = g.pbf_learned_stats(with_mean=True)
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=True), x0)
        g.pbf_set_network(net)
        g.pbf_set_mode(learned=True, learned_iterations=1, bound=-1.0)
        g.pbf_set_graph(graph)
        g.pbf_step(np.float32(cfg.dt), steps)
        lx, lv, _ = g.pbf_download(order=1)
        n, reruns, errs, errm = g.pbf_learned_stats(with_mean=True)
    assert n == steps and reruns == steps and errs.shape == (steps,)
    assert np.array_equal(lx, px) and np.array_equal(lv, pv)
    # every step reran: the recorded errors are the plain step's
    assert np.array_equal(errs, perrs) and np.array_equal(errm, perrm)


def test_nan_network_falls_back(gpu_mod):
    cfg = scenes.small(8, 8, 8, box=(16, 16, 16), seed=10)
    x0 = cfg.positions()
    px, pv, _ = _plain(gpu_mod, cfg, x0, 4)
    net = _net(np.full(M.N_PARAMS, np.nan))
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=True), x0)
        g.pbf_set_network(net)
        g.pbf_set_mode(learned=True, learned_iterations=1, bound=0.01)
        g.pbf_step(np.float32(cfg.dt), 4)
        lx, lv, _ = g.pbf_download(order=1)
        n, reruns, _ = g.pbf_learned_stats()
    assert n == 4 and reruns == 4
    assert np.array_equal(lx, px) and np.array_equal(lv, pv)


def test_zero_network_with_k_iterations_is_the_plain_step(gpu_mod):
    # correction 0 and K learned-path iterations: every kernel sees the plain
    # step's inputs, so the learned path reproduces it bit for bit
    cfg = scenes.small(10, 10, 10, box=(20, 20, 20), seed=12)
    x0 = cfg.positions()
    px, pv, _ = _plain(gpu_mod, cfg, x0, 5)
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=True), x0)
        g.pbf_set_network(_net(np.zeros(M.N_PARAMS)))
        g.pbf_set_mode(learned=True, learned_iterations=cfg.iterations, bound=np.inf)
        g.pbf_set_graph(True)
        g.pbf_step(np.float32(cfg.dt), 5)
        lx, lv, _ = g.pbf_download(order=1)
        n, reruns, _ = g.pbf_learned_stats()
    assert n == 5 and reruns == 0
    assert np.array_equal(lx, px) and np.array_equal(lv, pv)


def test_density_check_matches_oracle(gpu_mod, oracle):
    cfg = scenes.small(10, 10, 10, box=(20, 20, 20), seed=13)
    x0 = cfg.positions()
    steps = 6
    x, v = x0.copy(), np.zeros_like(x0)
    ids = np.arange(cfg.n, dtype=np.int32)
    want, want_mean = [], []
    for _ in range(steps):
        _, e, em = oracle.pbf_step(oracle.pbf_params(cfg), x, v, ids, np.float32(cfg.dt), final_err=True)
        want.append(e)
        want_mean.append(em)
    with gpu_mod.Gpu(0) as g:
        g.pbf_init(gpu_mod.pbf_params(cfg, exact=True), x0)
        g.pbf_set_mode(learned=False, check=True)
        g.pbf_step(np.float32(cfg.dt), steps)
        n, reruns, errs, errm = g.pbf_learned_stats(with_mean=True)
    assert n == steps and reruns == 0
    assert np.array_equal(errs, np.array(want, np.float32))
    assert max(want_mean) > 0
    assert np.allclose(errm, want_mean, rtol=1e-5, atol=1e-9)
