"""The CPU restatement of the screener network: gradient against central
finite differences of the FP64 network (no BF16 rounding), and the BF16
rounding against hand-worked values."""
import numpy as np

import mlp as M


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.001953125, 1.005859375, -3.0e-5, 65504.0], dtype=np.float32)
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    got = M.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[3] == 1.0078125  # 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert abs(got[4] - (-3.0e-5)) < 3.0e-5 * 2 ** -8
    assert got[5] == 65536.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    nin, n = 6, 40
    p = rng.normal(0, 0.3, M.num_params(nin))
    X = rng.normal(size=(n, nin))
    t = rng.normal(size=n)
    w = rng.uniform(0.5, 2.0, n)
    loss, g = M.loss_grad(p, X, t, w, nin, round_bf16=False)
    idx = rng.choice(len(p), 60, replace=False)
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss_grad(pp, X, t, w, nin, False)[0] - M.loss_grad(pm, X, t, w, nin, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
