"""The learned step proposal on the GPU: trained here on traces of other seeded
scenes, then compared with the plain march on a held-out view.  Every jump is
certified (the samples it passes are proven below iso), so the image is the
plain march's bit for bit; the network's gain is the samples it skips."""
import numpy as np
import pytest

import mlp_oracle as M
from pmarch import scenes, train

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(marcher):
    marcher.dataset_clear(600_000)
    n = 0
    for i, s in enumerate([scenes.dam_break(31 + i, nx=60, ny=120, nz=60, width=640, height=360) for i in range(3)]):
        rng = np.random.default_rng(50 + i)
        cam = scenes.random_views(rng, scenes.COLUMN_CENTRE, 1, 640, 640, 360, 360, radius=(1.1, 1.8))[0]
        marcher.set_frame(30, s.xyz, s.h)
        n = marcher.trace_append(30, s.settings(), cam.inv_pv, cam.eye, cam.width, cam.height, 2, 2, seed=i)
    assert n > 50_000
    marcher.mlp_set(train.he_init(0))
    losses = marcher.mlp_train(6, batch=2048, lr=2e-3, w_over=train.W_OVER, seed=1)
    print("rows", n, "loss per epoch", losses, flush=True)
    assert losses[-1] < 0.6 * losses[0]
    return marcher.mlp_get()


@pytest.mark.parametrize("skip_mode", [0, 4])
def test_learned_matches_plain(marcher, trained, skip_mode):
    """Held-out view, without and with the exact bound skip: the learned march's
    images equal the plain march's bit for bit, its certificates skip samples
    the plain march evaluates, and its march samples plus certified skips cover
    at least the plain march's samples (nothing is lost, nothing invented)."""
    sc = scenes.dam_break(0, nx=60, ny=120, nz=60, width=960, height=540)
    c = sc.cameras[0]
    print(f"learned vs plain 960x540 skip_mode={skip_mode}", flush=True)
    marcher.set_frame(31, sc.xyz, sc.h)
    marcher.mlp_set(trained)
    plain = marcher.march(31, sc.settings(skip_mode=skip_mode), c.inv_pv, c.eye, c.width, c.height)
    lrn = marcher.march(31, sc.settings(learned=1, skip_mode=skip_mode), c.inv_pv, c.eye, c.width, c.height)
    sp, sl = plain[3], lrn[3]
    print("plain", sp, "\nlearned", sl, flush=True)
    for a, b in zip(plain[:3], lrn[:3]):
        assert np.array_equal(a, b)
    assert sl["hits"] == sp["hits"] and sl["rays_in_aabb"] == sp["rays_in_aabb"]
    assert sl["mlp_evals"] > 0 and sl["cert_walks"] > 0 and sl["cert_skipped"] > 0
    assert sl["march_samples"] < sp["march_samples"]
    # the finish is the same: every sample after the crossing is the plain one's
    assert sl["samples"] - sl["march_samples"] == sp["samples"] - sp["march_samples"]


def test_learned_without_network_is_rejected(marcher):
    import ctypes
    import pmarch
    m2 = pmarch.Marcher(0)
    sc = scenes.lattice_cube(0, width=64, height=48)
    m2.set_frame(0, sc.xyz, sc.h)
    with pytest.raises(pmarch.PmError):
        m2.march(0, sc.settings(learned=1), sc.cameras[0].inv_pv, sc.cameras[0].eye, 64, 48)
    m2.close()
    assert ctypes is not None


def test_mlp_init_matches_oracle_layout():
    p = train.he_init(0)
    assert p.size == M.NPARAM
    assert np.all(p[M.OFF_W1:M.OFF_B1].reshape(M.H, M.K1)[:, M.NIN:] == 0)
