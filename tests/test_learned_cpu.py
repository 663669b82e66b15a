"""CPU checks of the learned preconditioner's host side: weight packing
layouts (forward and flipped/transposed backward-data weights), the
committed trained weights, and the training-data generator's statistics."""
import os

import numpy as np
import pytest
import torch

from nitorch_amd import learned as LN


def test_pack_weight_layout():
    rng = np.random.default_rng(0)
    co, ci = 5, 9
    w = torch.from_numpy(rng.standard_normal((co, 3, 3, 3, ci)).astype(np.float32))
    wk = LN.pack_weight(w, 16, 16).float()
    assert wk.shape == (16, LN._kp(16))
    # K = tap * CI_stored + ci, tap = 9 dx + 3 dy + dz
    for (o, i, j, k, c) in [(0, 0, 0, 0, 0), (4, 2, 1, 0, 8), (2, 1, 2, 2, 3)]:
        tap = 9 * i + 3 * j + k
        assert wk[o, tap * 16 + c] == w[o, i, j, k, c].to(torch.bfloat16).float()
    assert float(wk[co:].abs().sum()) == 0 and float(wk[:, 27 * 16:].abs().sum()) == 0
    wt = LN.pack_weight_t(w, 16, 16).float()
    # backward data: W'[ci][tap' * CO_stored + co] = W[co][26 - tap'][ci]
    for (o, i, j, k, c) in [(0, 0, 0, 0, 0), (4, 2, 1, 0, 8), (2, 1, 2, 2, 3)]:
        tap = 9 * i + 3 * j + k
        assert wt[c, (26 - tap) * 16 + o] == w[o, i, j, k, c].to(torch.bfloat16).float()


def test_committed_weights_load():
    path = LN.default_weights_path()
    if not os.path.exists(path):
        pytest.skip("no trained weights committed")
    with np.load(path, allow_pickle=False) as f:
        keys = set(f.files)
        for name, ci, ci_s, co, co_s, act in LN.LAYERS:
            assert f[name + ".w"].shape == (co, 3, 3, 3, ci)
            assert f[name + ".b"].shape == (co,)
            assert np.isfinite(f[name + ".w"]).all()
    assert keys == {n + s for n, *_ in LN.LAYERS for s in (".w", ".b")}


def test_make_system_statistics():
    g, h, pen = LN.make_system((16, 16, 16), seed=3, device="cpu")
    assert g.shape == (1, 3, 16, 16, 16) and h.shape == (1, 6, 16, 16, 16)
    assert abs(float(g.std()) - 1) < 1e-5
    assert float(h[:, :3].min()) >= 0.5 and float(h[:, :3].max()) <= 2.0
    assert float(h[:, 3:].abs().max()) <= 0.2
    for p, b in zip(pen, LN.BASE_PENALTY):
        assert 0.3 * b - 1e-12 <= p <= 3.0 * b + 1e-12
