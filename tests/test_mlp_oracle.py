"""The surrogate's CPU oracle pinned by a library routine (torch's bfloat16
cast) and by finite differences of its own FP64 loss."""
import numpy as np
import pytest

import mlp_oracle as MO

torch = pytest.importorskip("torch")
RNG = np.random.default_rng(21)


def test_bf16_rounding_matches_torch():
    a = np.concatenate([RNG.standard_normal(20000).astype(np.float32) * 10.0 ** RNG.integers(-8, 8, 20000),
                        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0,
                                  3.3895314e38, 1e-40], np.float32)]).astype(np.float32)
    ours = MO.bf16(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    # ties go to even
    assert MO.bf16(np.float32(1.0 + 2.0 ** -8)) == np.float32(1.0)
    assert MO.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)


def _random_net():
    p = np.zeros(MO.NPAR, np.float32)
    p[MO.O_W1:MO.O_B1] = RNG.uniform(-2, 2, 64)
    p[MO.O_B1:MO.O_W2] = RNG.uniform(-0.5, 0.5, 64)
    p[MO.O_W2:MO.O_B2] = RNG.uniform(-0.3, 0.3, 4096)
    p[MO.O_B2:MO.O_W3] = RNG.uniform(-0.2, 0.2, 64)
    p[MO.O_W3:MO.O_B3] = RNG.uniform(-0.3, 0.3, 320)
    p[MO.O_B3:] = RNG.uniform(-0.1, 0.1, 5)
    norm = (0.3, 2.0, np.array([1.1, 1.0, 1.0, 1.0, 1.5]), np.array([0.2, 0.3, 0.1, 0.1, 0.8]))
    return p, norm


def test_backward_matches_finite_differences():
    """Without bf16 rounding the loss is smooth: every gradient group of the
    oracle's backward against central differences in float64."""
    p, norm = _random_net()
    h = RNG.uniform(-0.2, 0.8, 37)
    t = RNG.uniform(0.7, 2.0, (37, 5))
    L, g = MO.loss_grad(p, norm, h, t, rounding=False)
    p64 = p.astype(np.float64)

    def loss_at(q):
        y, _ = MO.forward(q, norm, h, rounding=False)
        e = (y - t) / t
        return (e * e).sum() / (5 * len(h))

    # forward() casts parameters to float32; evaluate the perturbed loss in float64 directly
    def loss64(q):
        w1, b1, W2, b2, W3, b3 = (q[MO.O_W1:MO.O_B1], q[MO.O_B1:MO.O_W2], q[MO.O_W2:MO.O_B2].reshape(64, 64),
                                  q[MO.O_B2:MO.O_W3], q[MO.O_W3:MO.O_B3].reshape(5, 64), q[MO.O_B3:])
        x = ((h - norm[0]) * norm[1]).astype(np.float32).astype(np.float64)
        a1 = np.tanh(np.outer(x, w1) + b1)
        a2 = np.tanh(a1 @ W2.T + b2)
        y = norm[2] + norm[3] * (a2 @ W3.T + b3)
        e = (y - t) / t
        return (e * e).sum() / (5 * len(h))

    assert abs(loss64(p64) - L) < 1e-12 * abs(L) + 1e-15
    idx = np.concatenate([RNG.choice(np.arange(lo, hi), min(12, hi - lo), replace=False)
                          for lo, hi in ((MO.O_W1, MO.O_B1), (MO.O_B1, MO.O_W2), (MO.O_W2, MO.O_B2),
                                         (MO.O_B2, MO.O_W3), (MO.O_W3, MO.O_B3), (MO.O_B3, MO.NPAR))])
    for i in idx:
        eps = 1e-6
        qp, qm = p64.copy(), p64.copy()
        qp[i] += eps
        qm[i] -= eps
        fd = (loss64(qp) - loss64(qm)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, fd, g[i])


def test_planted_sign_error_in_backward_is_caught():
    """A backward with the tanh derivative written as (1 + a^2) fails the
    finite-difference check above (the check can see a plausible slip)."""
    p, norm = _random_net()
    h = RNG.uniform(-0.2, 0.8, 19)
    t = RNG.uniform(0.7, 2.0, (19, 5))
    L, g = MO.loss_grad(p, norm, h, t, rounding=False)
    # recompute dW2 with the wrong derivative
    y, (x, a1, a2, o) = MO.forward(p, norm, h, rounding=False)
    dy = 2.0 * (y - t) / t / t / (5 * len(h)) * norm[3]
    W3 = p[MO.O_W3:MO.O_B3].reshape(5, 64).astype(np.float64)
    wrong = ((dy @ W3) * (1.0 + a2 * a2)).T @ a1
    assert np.abs(wrong.ravel() - g[MO.O_W2:MO.O_B2]).max() > 1e-3 * np.abs(g[MO.O_W2:MO.O_B2]).max()
