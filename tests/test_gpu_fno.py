"""The spectral neural operator's device kernels against the FP64 oracle
(oracle/fno_oracle.py): the BF16-MFMA GEMM against numpy, the forward pass,
the loss and every parameter block's gradient, and a few Adam steps that
lower the loss.  BF16 operands (8-bit mantissa) with FP32 accumulation set
the tolerances: 1e-2 relative for products and gradients."""
import numpy as np
import pytest

from oracle.fno_oracle import FNOOracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("M,N,K", [(32, 128, 32), (200, 70, 45), (64, 64, 512), (32, 32, 9000), (513, 40, 3)])
def test_mfma_gemm_matches_numpy(M, N, K):
    from hawassi_c_amd.fno import gemm_check
    g = np.random.default_rng(M * 7 + N)
    print(f"gemm {M}x{N}x{K}", flush=True)
    # small integers are exact in BF16, their products and these sums exact in
    # FP32: the lane layout, the K tiling, the tails and split-K must give
    # the product exactly
    ai = g.integers(-8, 9, (M, K)).astype(np.float32)
    bi = g.integers(-8, 9, (K, N)).astype(np.float32)
    assert np.array_equal(gemm_check(ai, bi), (ai.astype(np.int64) @ bi.astype(np.int64)).astype(np.float32))
    # general data: BF16 rounding of both operands, ~2^-9 relative per operand
    a = g.standard_normal((M, K)).astype(np.float32)
    b = g.standard_normal((K, N)).astype(np.float32)
    exact = a.astype(np.float64) @ b.astype(np.float64)
    c = gemm_check(a, b)
    assert rel(c, exact) < 8e-3
    # split BF16 (hi + lo per operand, three products): 16 mantissa bits per
    # operand, ~2^-16 relative per product, FP32 accumulation
    cs = gemm_check(a, b, split=True)
    assert rel(cs, exact) < 5e-5
    assert rel(cs, exact) < 0.05 * rel(c, exact)
    # integer data stays exact (lo parts are zero)
    assert np.array_equal(gemm_check(ai, bi, split=True), (ai.astype(np.int64) @ bi.astype(np.int64)).astype(np.float32))


def _pair(n=16, width=8, modes=4, layers=3, hidden=16, batch=2, seed=5):
    from hawassi_c_amd.fno import FNO
    dev = FNO(n, width, modes, layers, hidden, max_batch=batch, seed=seed)
    orc = FNOOracle(n, width, modes, layers, hidden)
    assert dev.nparam == orc.nparam
    g = np.random.default_rng(seed)
    x = g.standard_normal((batch, 3, n, n)).astype(np.float32)
    t = g.standard_normal((batch, 2, n, n)).astype(np.float32)
    return dev, orc, x, t


@pytest.mark.parametrize("n,width,modes,layers,hidden", [(16, 8, 4, 3, 16), (32, 32, 8, 4, 64)])
def test_forward_and_gradients_match_the_oracle(n, width, modes, layers, hidden):
    dev, orc, x, t = _pair(n, width, modes, layers, hidden)
    p = dev.params().astype(np.float64)
    print(f"fno n={n} width={width} params={dev.nparam}", flush=True)
    y = dev.forward(x)
    yo, _ = orc.forward(p, x.astype(np.float64))
    assert rel(y, yo) < 1e-2
    loss, g = dev.loss_grad(x, t)
    lo, go = orc.loss_grad(p, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lo) < 1e-2 * lo
    for name, off, shp in orc.layout:
        k = int(np.prod(shp))
        e = rel(g[off:off + k], go[off:off + k])
        assert e < 2e-2, (name, e)


def test_split_precision_forward_is_smooth():
    """The split-BF16 forward follows the FP64 oracle to ~1e-5 and its
    response to a small input change follows the oracle's, where the BF16
    forward's response is dominated by rounding."""
    dev, orc, x, t = _pair(32, 32, 8, 4, 64)
    p = dev.params().astype(np.float64)
    dx = 3e-3 * np.random.default_rng(9).standard_normal(x.shape).astype(np.float32)
    yo, _ = orc.forward(p, x.astype(np.float64))
    yo2, _ = orc.forward(p, (x + dx).astype(np.float64))
    dev.set_precision(True)
    ys, ys2 = dev.forward(x), dev.forward(x + dx)
    dev.set_precision(False)
    yb, yb2 = dev.forward(x), dev.forward(x + dx)
    print("split", rel(ys, yo), rel(ys2 - ys, yo2 - yo), "bf16", rel(yb, yo), rel(yb2 - yb, yo2 - yo), flush=True)
    assert rel(ys, yo) < 1e-4
    assert rel(ys2 - ys, yo2 - yo) < 0.1
    assert rel(yb2 - yb, yo2 - yo) > 0.5


def test_adam_steps_lower_the_loss():
    from hawassi_c_amd.fno import FNO
    n, b = 32, 4
    dev = FNO(n, 16, 6, 2, 32, max_batch=b, seed=3)
    g = np.random.default_rng(2)
    # a learnable map: the target is a smoothed, shifted copy of the first input channel
    x = g.standard_normal((16, 3, n, n)).astype(np.float32)
    k = np.fft.fftfreq(n) * n
    lowpass = (np.abs(k)[None, :] < 6) & (np.abs(k)[:, None] < 6)
    xs = np.real(np.fft.ifft2(np.fft.fft2(x[:, 0]) * lowpass)).astype(np.float32)
    x[:, 0] = xs
    y = np.stack([np.roll(xs, 2, axis=-1), 0.5 * xs], axis=1).astype(np.float32)
    dev.set_dataset(x, y)
    first = [dev.train_step(np.arange(i, i + b) % 16, 3e-3) for i in range(0, 16, b)]
    for it in range(40):
        last = dev.train_step(np.arange((4 * it) % 16, (4 * it) % 16 + b) % 16, 3e-3)
    assert last < 0.5 * np.mean(first), (first, last)
