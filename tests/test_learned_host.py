"""Host-side pieces of the learned warm start (no GPU): training scenes,
network files, argument checks of the C ABI entry points."""
import ctypes
import math

import numpy as np
import pytest

from lustrine_amd import gpu as G
from lustrine_amd import learned, scenes


def test_training_batch_aspects_and_sizes():
    b = scenes.training_batch()
    assert len(b.scenes) == 64 and b.steps == 200
    rng = np.random.default_rng(3)
    aspects = rng.uniform(0.5, 2.0, 64)
    for s, a in zip(b.scenes, aspects):
        nx, ny, nz = s.lattice
        assert abs(s.n - 8192) / 8192 < 0.06, s.name
        assert abs(math.log((nx / nz) / a)) < 0.12, (s.name, a)
        assert s.box == (2 * nx, 2 * ny, 2 * nz)
    # deterministic
    assert [s.lattice for s in scenes.training_batch().scenes] == [s.lattice for s in b.scenes]


def test_network_file_round_trip(tmp_path):
    net = np.arange(G.NET_FLOATS, dtype=np.float32) * 0.5
    p = tmp_path / "n.npz"
    learned.save(str(p), net, {"epochs": 3})
    got, meta = learned.load(str(p))
    assert np.array_equal(got, net) and meta == {"epochs": 3}
    learned.save(str(p), net[:10], {})
    with pytest.raises(ValueError):
        learned.load(str(p))


def test_layout_constants_agree():
    from oracle import mlp as M
    assert G.MLP_PARAMS == M.N_PARAMS == 32 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3
    assert G.NET_FLOATS == M.NET_FLOATS == 64 + G.MLP_PARAMS + 1
    assert G.NET_OUT_SCALE == M.NET_OUT_SCALE


def test_learned_entry_points_reject_bad_arguments():
    lib = G.load()
    assert lib.lus_pbf_set_network(None, None, 0) == 1
    assert lib.lus_pbf_set_mode(None, 1, 1, ctypes.c_float(0.01), 0) == 1
    out = ctypes.c_void_p()
    assert lib.lus_mlp_trainer_create(None, 10, ctypes.byref(out)) == 1
    assert lib.lus_mlp_fit(None, None, None, None) == 1
    assert lib.lus_mlp_loss_grad(None, None, 0, None, None, ctypes.c_float(1.0), None, None) == 1
    assert lib.lus_mlp_infer(None, None, 0, None, None) == 1
    assert lib.lus_pbf_record_step(None, None, ctypes.c_float(0.01), 1, None) == 1
