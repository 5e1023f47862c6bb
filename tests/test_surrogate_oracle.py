"""CPU checks of the surrogate's oracle (oracle/mlp_oracle.py).

The analytic gradient of the oracle is pinned by central finite
differences of its own loss with the BF16 roundings switched off (the
function is smooth then); the sample generator's inputs are pinned by
their ranges and log-uniform moments.  A planted error in the backward
formula (a dropped SiLU derivative, a transposed W2) fails the first test.
"""
import numpy as np

import mlp_oracle as mo
import oracle


def _weights(seed=0):
    rng = np.random.default_rng(seed)
    w = np.zeros(mo.NP)
    w[mo.OW1:mo.OB1] = rng.normal(0, 1 / np.sqrt(3), 192)
    w[mo.OB1:mo.OW2] = rng.normal(0, 0.1, 64)
    w[mo.OW2:mo.OB2] = rng.normal(0, 1 / 8, 4096)
    w[mo.OB2:mo.OW3] = rng.normal(0, 0.1, 64)
    w[mo.OW3:mo.OB3] = rng.normal(0, 1 / 8, 64)
    w[mo.OB3] = 0.05
    return w


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (40, 3))
    t = rng.uniform(-1, 1, 40)
    w = _weights()
    L, g = mo.loss_grad(w, X, t, bf=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False)
                          for a, b in ((mo.OW1, mo.OB1), (mo.OB1, mo.OW2), (mo.OW2, mo.OB2),
                                       (mo.OB2, mo.OW3), (mo.OW3, mo.OB3))] + [[mo.OB3]])
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (mo.loss_grad(wp, X, t, bf=False)[0] - mo.loss_grad(wm, X, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-3, 65504.0])
    b = mo.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: round to even (1.0)
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == 1.0078125
    assert np.all(np.abs(b - x) <= np.abs(x) * 2.0 ** -8)


def test_sample_inputs_ranges():
    N, T, r = mo.sample_inputs(200000, seed=7)
    for a, lo, hi in ((np.log10(N), -1, 4), (np.log10(T * 1e3), 1, 4), (np.log10(r), -3, 3)):
        assert a.min() >= lo and a.max() < hi
        assert abs(a.mean() - 0.5 * (lo + hi)) < 0.02 * (hi - lo)


def test_imc_solves_backward_euler_on_samples():
    # the training targets: IMC's solution satisfies the backward-Euler
    # equation to the accuracy of its bisection (a step of 1e-5 in P or 30
    # halvings of [1e-10, 1e10] in log10 P)
    N, T, r = mo.sample_inputs(400, seed=11)
    KB = 8.63359
    res = []
    for n_, T_, r_ in zip(N, T, r):
        rho, gi = 1.27 * n_, 2.0 / 3.0
        P0 = KB * n_ * T_
        L0 = oracle.fcool_ki(rho, T_)
        dt = r_ * P0 / (gi * abs(L0))
        P = oracle.imc(P0, n_, dt, rho, gi)
        res.append(abs(P + gi * oracle.fcool_ki(rho, P / (KB * n_)) * dt - P0) / P0)
    res = np.array(res)
    assert np.median(res) < 1e-5 and res.max() < 1e-2, (np.median(res), res.max())
