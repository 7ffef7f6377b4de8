"""The FP64 network oracle against finite differences and closed forms."""
import numpy as np

from oracle import mlp_oracle as M


def _params(rng):
    p = rng.standard_normal(M.NPARAM) * 0.3
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.standard_normal((9, M.NIN))
    t = rng.standard_normal(9)
    _, g = M.loss_and_grad(p, x, t)
    num = M.numeric_grad(p, x, t)
    assert np.max(np.abs(g - num)) < 1e-6 * max(1.0, np.max(np.abs(num)))


def test_forward_closed_form():
    # single hidden path: y = relu(relu(x.W1 + b1).W2 + b2).w3 + b3
    p = np.zeros(M.NPARAM)
    W1, b1, W2, b2, w3, _ = M.unpack(p)
    p[M.OW1 + 0 * M.NH + 5] = 2.0          # W1[0][5]
    p[M.OW2 + 5 * M.NH + 7] = 3.0          # W2[5][7]
    p[M.OW3 + 7] = 0.5                     # w3[7]
    p[M.OB3] = 1.0
    x = np.zeros((3, M.NIN))
    x[:, 0] = [1.0, -1.0, 0.25]
    y, _ = M.forward(p, x)
    assert np.allclose(y, [1 + 0.5 * 3 * 2 * 1.0, 1.0, 1 + 0.5 * 3 * 2 * 0.25])


def test_bf16_rounding():
    # bf16 keeps 8 significant bits: the spacing at 1 is 2^-7
    v = np.array([1.0, 1.00390625, 1.01171875, 3.14159265, -2.5e-3], dtype=np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0 and r[2] == 1.015625   # halfway cases go to the even neighbour
    assert abs(r[3] - 3.140625) < 1e-9
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)
