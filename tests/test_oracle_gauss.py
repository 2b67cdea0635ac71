"""Pins of the screen-space depth smoothing (GaussRenderPass.cpp:8-48,
gauss.frag): the kernel table against the Gaussian in FP64 and its window
normalisation, the library's host-side table equal to the oracle's, and the
oracle's window sum against an independent FP64 convolution with edge-clamped
borders (numpy), on constants, impulses, steps at the border and noise.  No
GPU."""
import math

import numpy as np
import pytest

import oracle as O
from pmarch import _lib


@pytest.mark.parametrize("N", [0, 1, 2, 8, 31])
def test_kernel_table(N):
    K = O.gauss_kernel(N).astype(np.float64)
    i, j = np.meshgrid(np.arange(N + 1), np.arange(N + 1), indexing="ij")
    g = np.exp(-0.5 * (i * i + j * j))
    mult = np.where((i == 0) & (j == 0), 1, np.where((i != 0) & (j != 0), 4, 2))
    # float exponent -(r r)/2 with r = sqrtf(i^2 + j^2) and E rounded to float: the relative
    # error grows with r^2
    want = g / (g * mult).sum()
    # (and entries below float's normal range, r^2 > 170, are denormal or 0)
    assert np.all(np.abs(K - want) <= want * (1e-6 + 2e-7 * (i * i + j * j)) + 1e-40)
    assert abs((K * mult).sum() - 1.0) < 1e-5
    assert np.array_equal(_lib.gauss_kernel(N), O.gauss_kernel(N))    # library host code == oracle


def _ref(depth, N):
    K = O.gauss_kernel(N).astype(np.float64)
    full = np.zeros((2 * N + 1, 2 * N + 1))
    for a in range(-N, N + 1):
        for b in range(-N, N + 1):
            full[a + N, b + N] = K[abs(a), abs(b)]        # [i (x), j (y)]
    P = np.pad(depth.astype(np.float64), N, mode="edge")
    H, W = depth.shape
    out = np.zeros((H, W))
    for a in range(-N, N + 1):
        for b in range(-N, N + 1):
            out += full[a + N, b + N] * P[N + b: N + b + H, N + a: N + a + W]
    return out


@pytest.mark.parametrize("N", [0, 1, 3, 8])
def test_window_sum_against_fp64_convolution(N):
    rng = np.random.default_rng(N)
    H, W = 37, 53
    imgs = [np.full((H, W), 0.75, np.float32), rng.uniform(0, 1, (H, W)).astype(np.float32)]
    step = np.zeros((H, W), np.float32); step[:, :3] = 1.0
    imp = np.zeros((H, W), np.float32); imp[18, 26] = 1.0
    imgs += [step, imp]
    for img in imgs:
        got = O.depth_smooth(img, N).astype(np.float64)
        assert np.abs(got - _ref(img, N)).max() <= 2e-6 * max(1.0, np.abs(img).max())
    # the impulse response is the kernel itself (x horizontal, y vertical)
    got = O.depth_smooth(imp, N)
    K = O.gauss_kernel(N)
    if N:
        assert got[18, 26 + 1] == K[1, 0] and got[18 + 1, 26] == K[0, 1] and got[18, 26] == K[0, 0]
    assert np.allclose(O.depth_smooth(imgs[0], N), 0.75, rtol=1e-5)
