"""Pins for the CPU reference of the learned closure (oracle/les_mlp_oracle.c):
  * the FP64 gradient against central finite differences of the loss;
  * the [out][in] parameter layout against a hidden-unit permutation, which
    leaves the function unchanged only if rows and columns are indexed as
    stated (a transposed operand breaks it);
  * the bf16 rounding against torch's float32 -> bfloat16 conversion
    (round to nearest even) and the bf16 forward against the FP64 one."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def params(O, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    th = np.zeros(O.mlp_nparam())
    for (o, sh) in O.mlp_slices():
        fan_in = sh[1] if len(sh) == 2 else 4
        th[o:o + int(np.prod(sh))] = rng.standard_normal(int(np.prod(sh))) * scale / np.sqrt(fan_in)
    return th


def test_nparam(O):
    assert O.mlp_nparam() == 64 * 8 + 64 + 2 * (64 * 64 + 64) + 2 * 64 + 2


def test_gradient_matches_finite_differences(O):
    th = params(O, 1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((5, 8))
    t = rng.standard_normal((5, 2))
    loss, g = O.mlp_loss_grad(th, x, t, bf16=False)
    idx = np.concatenate([rng.choice(th.size, 600, replace=False),
                          np.arange(O.mlp_slices()[7][0], th.size)])   # plus the output biases
    h = 1e-6
    for j in idx:
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        fd = (O.mlp_loss_grad(tp, x, t, bf16=False)[0] - O.mlp_loss_grad(tm, x, t, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-7 * max(1.0, abs(g[j])) + 1e-9, (j, fd, g[j])
    # the loss itself: 1/(2P) sum (y - t)^2
    y = O.mlp_forward(th, x, bf16=False)
    assert abs(loss - 0.5 * np.sum((y - t) ** 2) / 5) < 1e-14


def test_hidden_permutation_invariance(O):
    th = params(O, 3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((7, 8))
    y0 = O.mlp_forward(th, x, bf16=False)
    sl = O.mlp_slices()
    W = [th[o:o + int(np.prod(sh))].reshape(sh).copy() for (o, sh) in sl]
    perm = rng.permutation(64)
    W[0], W[1] = W[0][perm], W[1][perm]          # layer-1 output units
    W[2] = W[2][:, perm]                         # layer-2 inputs
    perm2 = rng.permutation(64)
    W[2], W[3] = W[2][perm2], W[3][perm2]
    W[4] = W[4][:, perm2]
    th2 = np.concatenate([w.ravel() for w in W])
    assert np.max(np.abs(O.mlp_forward(th2, x, bf16=False) - y0)) < 1e-12
    # permuting the inputs of W1 without permuting x changes the function
    W[0] = W[0][:, ::-1]
    th3 = np.concatenate([w.ravel() for w in W])
    assert np.max(np.abs(O.mlp_forward(th3, x, bf16=False) - y0)) > 1e-3


def test_bf16_rounding_and_forward(O):
    torch = pytest.importorskip("torch")
    th = params(O, 5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((64, 8)) * 3
    # a network whose only rounding point is the input: W1 = e_i rows, no hidden effect
    yb = O.mlp_forward(th, x, bf16=True)
    yd = O.mlp_forward(th, x, bf16=False)
    rel = np.linalg.norm(yb - yd) / np.linalg.norm(yd)
    assert 1e-5 < rel < 3e-2, rel          # bf16 has 8 significant bits
    # rounding of one value through the layer-1 path: x exactly bf16 -> no change
    xr = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    th_id = np.zeros_like(th)
    (o1, _), (ob1, _), (o2, _), (ob2, _), (o3, _), (ob3, _), (o4, _), (ob4, _) = O.mlp_slices()
    # y0 = sum_i W4[0,i] silu(z3_i) with W3 = 0, b3 = c: exercises the bias path exactly
    th_id[ob3:ob3 + 64] = 1.0
    th_id[o4:o4 + 64] = 1.0 / 64
    y = O.mlp_forward(th_id, xr, bf16=True)
    silu1 = 1.0 / (1.0 + np.exp(-1.0))
    b = torch.tensor([silu1], dtype=torch.float32).to(torch.bfloat16).item()
    assert np.allclose(y[:, 0], b, rtol=0, atol=1e-15)
    assert np.all(y[:, 1] == 0.0)


def test_bf16_rounding_many_values(O):
    """W3 = 0 and a one-hot W4 row make y0 = bf16(silu(b3_i)) exactly: the
    oracle's rounding must equal torch's float32 -> bfloat16 conversion."""
    torch = pytest.importorskip("torch")
    sl = O.mlp_slices()
    (ob3, _), (o4, _) = sl[5], sl[6]
    rng = np.random.default_rng(8)
    b3 = rng.standard_normal(64) * 5
    x = np.zeros((1, 8))
    for i in range(64):
        th = np.zeros(O.mlp_nparam())
        th[ob3:ob3 + 64] = b3
        th[o4 + i] = 1.0
        y = O.mlp_forward(th, x, bf16=True)[0, 0]
        s = b3[i] / (1.0 + np.exp(-b3[i]))
        want = torch.tensor([s], dtype=torch.float32).to(torch.bfloat16).double().item()
        assert y == want, (i, y, want)
