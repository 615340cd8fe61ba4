"""Batched fitting (ds_batch) on the GPU: a batch of uneven instances spread
over several streams must fit each instance exactly as fitting it alone does
(same steps from the same start; the only differences come from the order of
float atomics), and the autotuner must leave a working batch behind."""
from __future__ import annotations

import numpy as np
import pytest

import diffsplat_amd as ds
from helpers import close

pytestmark = pytest.mark.gpu

LR = {"lr_pos": 1e-2, "lr_sh_dc": 1e-2, "lr_scale": 5e-3, "lr_rot": 6e-4, "lr_opacity": 6e-3}
SIZES = [(96, 64, 700, 11), (160, 120, 2500, 12), (64, 64, 300, 13), (200, 112, 4000, 14),
         (128, 128, 1500, 15), (80, 48, 500, 16)]


def make(w, h, n, seed):
    view, proj, cam = ds.fitter_camera(w, h)
    fit = ds.ImageFitter(w, h, n, view=view, proj=proj, cam_pos=cam, lr=LR,
                         isect_cap=1 << 20)
    fit.set_target(ds.target_image(seed, w, h, 16))
    fit.upload_params(ds.frustum_splats(seed, n, aspect=w / h))
    fit.set_graph(True)
    return fit


def test_batch_matches_individual_fits():
    alone = [make(*s) for s in SIZES]
    for f in alone:
        f.step(5)
        f.synchronize()
    batched = [make(*s) for s in SIZES]
    batch = ds.BatchFitter(batched)
    batch.set_streams(3)
    batch.step(5)
    batch.synchronize()
    for (w, h, n, seed), a, b in zip(SIZES, alone, batched):
        ok, mabs, mrel = close(b.params(), a.params(), atol=1e-5, rtol=1e-3)
        print(f"[parity] batch_vs_alone_{w}x{h}_{n}: max_abs={mabs:.3e}")
        assert ok, (w, h, n, mabs, mrel)
        a.forward()
        b.forward()
        assert abs(a.loss() - b.loss()) <= 1e-4 * a.loss()
    batch.close()
    for f in alone + batched:
        f.close()


def test_batch_autotune_and_progress():
    fits = [make(*s) for s in SIZES]
    for f in fits:
        f.forward()
    loss0 = [f.loss() for f in fits]
    batch = ds.BatchFitter(fits)
    tuned = batch.autotune(reps=2)
    assert tuned["streams"] in (1, 2, 4, 8) and tuned["ms_per_batch_step"] > 0
    batch.step(40)
    batch.synchronize()
    for f, l0 in zip(fits, loss0):
        f.forward()
        assert np.isfinite(f.loss()) and f.loss() < l0
    batch.close()
    # the fitters keep working on their own streams after the batch is gone
    fits[0].step(1)
    fits[0].synchronize()
    for f in fits:
        f.close()
