"""Training-set plumbing of the denoiser measurement (CPU): the fixed crop
set lies inside the renders, every crop is used once per epoch, and a flip
mirrors the normal's x channel."""

import numpy as np

from sundog_amd import denoise


class _Recorder(object):
    def __init__(self):
        self.batches = []

    def train_step(self, cx, cy, lr):
        self.batches.append((cx.copy(), cy.copy()))
        return 0.0


def test_crop_list_inside_renders():
    c = denoise.crop_list(5000, 7, 100, 120, 32, seed=1)
    assert c.shape == (5000, 3)
    assert c[:, 0].min() >= 0 and c[:, 0].max() <= 6
    assert c[:, 1].min() >= 0 and c[:, 1].max() <= 100 - 32
    assert c[:, 2].min() >= 0 and c[:, 2].max() <= 120 - 32


def test_fixed_crops_walk_each_epoch_once():
    n_img, h, w, crop = 3, 20, 24, 8
    xs = np.zeros((n_img, h, w, 9), np.float32)
    ys = np.zeros((n_img, h, w, 3), np.float32)
    # every pixel of every image carries a unique id in channel 0 (x) and 0 (y)
    ids = np.arange(n_img * h * w, dtype=np.float32).reshape(n_img, h, w)
    xs[..., 0] = ids
    ys[..., 0] = ids
    crops = denoise.crop_list(10, n_img, h, w, crop, seed=2)
    rec = _Recorder()
    denoise.train(rec, xs, ys, steps=5, batch=4, crop=crop, crops=crops)
    assert len(rec.batches) == 5
    # the crop each batch entry came from, identified by its pixel ids
    used = []
    for cx, cy in rec.batches:
        for k in range(cx.shape[0]):
            corner = min(cy[k, 0, 0, 0], cy[k, 0, -1, 0])   # a flip mirrors columns
            b, rem = divmod(int(corner), h * w)
            used.append((b, rem // w, rem % w))
            assert np.array_equal(cx[k, ..., 0], cy[k, ..., 0])
    want = [tuple(int(v) for v in c) for c in crops]
    # 20 picks = two full epochs over the 10 crops
    assert sorted(used) == sorted(want + want)


def test_flip_mirrors_normal_x():
    xs = np.zeros((1, 4, 6, 9), np.float32)
    ys = np.zeros((1, 4, 6, 3), np.float32)
    xs[..., 6] = 1.0
    rng = np.random.default_rng(0)
    picks = np.array([[0, 0, 0]] * 16)
    cx, _ = denoise.random_crops(xs, ys, rng, 16, 4, picks)
    signs = cx[:, 0, 0, 6]
    assert set(np.unique(signs)) == {-1.0, 1.0}
