"""Pins the float64 restatement of the matching-order network (oracle/mlp.py) on the CPU:
its gradient against central finite differences of its own loss, the bf16 rounding against known
values, and Adam's first step against its closed form."""
import numpy as np

from oracle import mlp


def _batch(rows, rng):
    x = rng.normal(size=(rows, mlp.IN))
    t = np.stack([rng.normal(size=rows), rng.normal(size=rows),
                  rng.integers(0, 8, size=rows).astype(np.float64),
                  rng.uniform(0.5, 2.0, size=rows)], axis=1)
    return x, t


def _params(rng, scale=0.3):
    return rng.normal(scale=scale, size=mlp.NUM_PARAMS)


def test_num_params():
    assert mlp.NUM_PARAMS == 32 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 10 + 10 == 11082


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    x, t = _batch(7, rng)
    loss, g = mlp.loss_grad(p, x, t, depth_weight=0.7)
    # every bias and a spread of weights in every layer
    offsets, o = [], 0
    for _, s in mlp.SHAPES:
        n = int(np.prod(s))
        offsets += list(o + rng.choice(n, size=min(n, 12), replace=False))
        o += n
    h = 1e-6
    for k in offsets:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mlp.loss_grad(pp, x, t, 0.7)[0] - mlp.loss_grad(pm, x, t, 0.7)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_loss_terms():
    # zero network: outputs are zero, so the loss is the weighted mean of t0^2 + t1^2 + log 8
    p = np.zeros(mlp.NUM_PARAMS)
    x = np.ones((3, mlp.IN))
    t = np.array([[1.0, 2.0, 3, 1.0], [0.0, -1.0, 0, 3.0], [2.0, 0.0, 7, 0.0]])
    loss, g = mlp.loss_grad(p, x, t, depth_weight=2.0)
    expect = (1 * (1 + 4 + 2 * np.log(8)) + 3 * (0 + 1 + 2 * np.log(8))) / 4
    assert abs(loss - expect) < 1e-12
    # the output bias gradient of the two regression heads is 2 * weighted error
    b4 = mlp.unpack(g)["b4"]
    assert abs(b4[0] - 2 * (1 * -1 + 3 * 0) / 4) < 1e-12
    assert abs(b4[1] - 2 * (1 * -2 + 3 * 1) / 4) < 1e-12


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5,
                     np.uint32(0x7F7F7FFF).view(np.float32), 3.4e38], np.float32)
    got = mlp.round_bf16(vals)
    # ties go to the even mantissa (1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6); above a tie rounds
    # up; the largest bf16 is (2 - 2^-7) * 2^127: just below its upper tie stays there, 3.4e38
    # overflows to infinity
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, (2 - 2 ** -7) * 2.0 ** 127,
                     np.inf], np.float32)
    assert np.array_equal(got, want), (got, want)
    # a subnormal keeps its top 16 bits, rounded the same way
    u = int(np.float32(1e-40).view(np.uint32))
    assert int(mlp.round_bf16(np.float32(1e-40)).view(np.uint32)) == \
        (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000


def test_adam_first_step_closed_form():
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    opt = mlp.Adam(4, lr=0.1, eps=1e-8)
    p = opt.step(np.zeros(4), g)
    # first step: m_hat = g, v_hat = g^2, so the update is lr * g / (|g| + eps)
    assert np.allclose(p, -0.1 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)


def test_bf16_forward_close_to_f32():
    rng = np.random.default_rng(3)
    p = _params(rng, 0.15)
    x, _ = _batch(50, rng)
    y32, _ = mlp.forward(p, x)
    y16, _ = mlp.forward(p, x, bf16=True)
    scale = np.abs(y32).max()
    assert np.abs(y32 - y16).max() <= 0.05 * scale
    assert not np.array_equal(y32, y16)
