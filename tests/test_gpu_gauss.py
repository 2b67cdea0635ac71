"""GPU parity of the screen-space depth smoothing (GaussRenderPass,
gauss.frag): the HIP kernels' images equal the CPU oracle's bit for bit on
ragged image sizes (images smaller than one tile included) and window sizes
0, 1, 7, 8 (the reference's default, which has a kernel of its own) and 31
(its largest), with the border clamped."""
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H,N", [(203, 77, 8), (64, 48, 0), (64, 48, 1), (97, 131, 31), (640, 480, 8),
                                   (5, 3, 8), (1, 1, 8), (33, 40, 7), (1920, 1080, 8)])
def test_depth_smooth_bitexact(marcher, W, H, N):
    rng = np.random.default_rng(W * 7 + N)
    depth = rng.uniform(0.2, 1.0, (H, W)).astype(np.float32)
    depth[: H // 3, : W // 4] = 1.0                     # clear value where no sprite is drawn
    print(f"depth smooth {W}x{H} N={N}", flush=True)
    got = marcher.depth_smooth(depth, N)
    assert np.array_equal(got, O.depth_smooth(depth, N))


def test_depth_smooth_1080p_timing(marcher):
    rng = np.random.default_rng(1)
    depth = rng.uniform(0.2, 1.0, (1080, 1920)).astype(np.float32)
    marcher.depth_smooth(depth, 8)
    t0 = time.perf_counter()
    for _ in range(5):
        out = marcher.depth_smooth(depth, 8)
    dt = (time.perf_counter() - t0) / 5
    rows = slice(0, 1080, 97)
    ref = O.depth_smooth(depth[:40], 8)
    assert np.array_equal(out[:20], ref[:20])           # rows whose window stays inside the slab
    print(f"1080p N=8 smoothing incl. copies: {dt * 1e3:.2f} ms", flush=True)
