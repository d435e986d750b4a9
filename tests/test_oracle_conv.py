"""Pins of the FP64 conv oracle (oracle/conv_oracle.c) by mathematics alone.

* forward: the circular convolution theorem -- the oracle's direct sum equals
  a numpy FFT product with the kernel's transfer function built from the
  definition's offsets; identity and shift kernels reproduce x and np.roll(x);
* data gradient: the adjoint identity <conv(x), dy> = <x, dgrad(dy)>;
* weight and bias gradient: conv is linear in (w, b), so
  <conv(x; w, b), dy> = <w, wgrad(x, dy)> + <b, db(dy)>.
A transposed operand, a flipped tap or a wrong wrap direction fails at least
one of these (the shift-kernel case fixes the direction)."""
import numpy as np
import pytest

from oracle import oracle as O


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def fft_conv(x, w, b):
    """y[n,o] = b[o] + sum_i IFFT(FFT(x[n,i]) * K[o,i]), with
    K[o,i](k,l) = sum_{a,c in -1..1} w[o,i,a+1,c+1] exp(2 pi i (k a / H + l c / W))
    (y[h] gathers x[h + a]: a shift by -a)."""
    N, Ci, H, W = x.shape
    Co = w.shape[0]
    kh = np.fft.fftfreq(H) * H
    kw = np.fft.fftfreq(W) * W
    X = np.fft.fft2(x)
    y = np.zeros((N, Co, H, W))
    for o in range(Co):
        acc = np.zeros((N, H, W), dtype=complex)
        for i in range(Ci):
            K = np.zeros((H, W), dtype=complex)
            for a in (-1, 0, 1):
                for c in (-1, 0, 1):
                    K += w[o, i, a + 1, c + 1] * np.exp(2j * np.pi * (kh[:, None] * a / H + kw[None, :] * c / W))
            acc += X[:, i] * K
        y[:, o] = np.fft.ifft2(acc).real + (b[o] if b is not None else 0.0)
    return y


@pytest.mark.parametrize("shape", [(2, 3, 4, 8, 6), (1, 1, 2, 5, 7), (3, 2, 1, 4, 4)])
def test_forward_equals_fft_convolution(shape):
    N, Ci, Co, H, W = shape
    x, w, b = rnd((N, Ci, H, W), 1), rnd((Co, Ci, 3, 3), 2), rnd(Co, 3)
    y = O.conv3x3_fwd(x, w, b)
    ref = fft_conv(x, w, b)
    assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_identity_and_shift_kernels():
    x = rnd((2, 2, 6, 5), 4)
    w = np.zeros((2, 2, 3, 3))
    w[0, 0, 1, 1] = w[1, 1, 1, 1] = 1.0
    b = np.array([0.25, -0.5])
    y = O.conv3x3_fwd(x, w, b)
    assert np.array_equal(y, x + b[None, :, None, None])
    # tap (dh, dw) = (0, 2): y[h, q] = x[h - 1, q + 1]
    w = np.zeros((1, 1, 3, 3))
    w[0, 0, 0, 2] = 1.0
    y = O.conv3x3_fwd(x[:, :1], w)
    assert np.array_equal(y[:, 0], np.roll(np.roll(x[:, 0], 1, axis=1), -1, axis=2))


@pytest.mark.parametrize("shape", [(2, 3, 4, 8, 6), (1, 32, 32, 6, 6), (2, 32, 1, 5, 9)])
def test_dgrad_is_the_adjoint(shape):
    N, Ci, Co, H, W = shape
    x, w, dy = rnd((N, Ci, H, W), 5), rnd((Co, Ci, 3, 3), 6), rnd((N, Co, H, W), 7)
    lhs = np.sum(O.conv3x3_fwd(x, w) * dy)
    dx = O.conv3x3_dgrad(dy, w)
    assert dx.shape == x.shape
    rhs = np.sum(x * dx)
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


@pytest.mark.parametrize("shape", [(2, 3, 4, 8, 6), (1, 32, 32, 6, 6), (3, 1, 32, 5, 4)])
def test_wgrad_is_the_gradient_of_a_linear_map(shape):
    N, Ci, Co, H, W = shape
    x, w, b, dy = rnd((N, Ci, H, W), 8), rnd((Co, Ci, 3, 3), 9), rnd(Co, 10), rnd((N, Co, H, W), 11)
    lhs = np.sum(O.conv3x3_fwd(x, w, b) * dy)
    dw, db = O.conv3x3_wgrad(x, dy)
    assert dw.shape == w.shape and db.shape == b.shape
    rhs = np.sum(w * dw) + np.sum(b * db)
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
    # and per tap: a one-hot weight picks that tap's entry
    e = np.zeros_like(w)
    e[Co - 1, 0, 2, 0] = 1.0
    assert abs(np.sum(O.conv3x3_fwd(x, e) * dy) - dw[Co - 1, 0, 2, 0]) <= 1e-12 * max(1.0, abs(dw).max())
