"""Forward-only renderer (SplatRenderer) on the clustered SH-degree-3 scene of
the viewer workload: parity against the oracle at a reduced size, and the
full 1,000,000-splat 1920x1080 orbit through size-independent invariants."""
from __future__ import annotations

import numpy as np
import pytest

import diffsplat_amd as ds
import oracle
from helpers import close

pytestmark = pytest.mark.gpu


def render_parity(n, W, H, cam_index):
    params = ds.cluster_scene(2000, n, 3, 64)
    view, proj, eye = ds.orbit_cameras(120, W, H)[cam_index]
    r = ds.SplatRenderer(W, H, n, 3)
    r.upload_params(params)
    r.render(view, proj, eye)
    img = r.image()
    st = r.raster_state()
    pv = ds.mat_mul(proj, view)
    s3 = oracle.act_fwd(params, n, 3, eye)
    s2 = oracle.proj_fwd(s3, n, pv)
    return img, st, s2, pv, eye, params


@pytest.mark.parametrize("cam_index", [0, 37])
def test_viewer_render_matches_oracle(cam_index):
    n, W, H = 150_000, 480, 270
    img, st, s2_64, pv, eye, params = render_parity(n, W, H, cam_index)
    # the renderer's Splat2D is internal: re-derive it with the C-ABI layer kernels
    d_p = ds.DeviceArray.from_numpy(params)
    d_s2 = ds.DeviceArray(4 * 10 * n)
    d_s2.zero()
    g = ds.gauss_view(d_p, n, 3)
    s2v = ds.splat2d_view(d_s2, n)
    ds.check(ds.lib().ds_activ_project_fwd(g, s2v, 3, ds.fptr(eye), ds.fptr(pv), None), "fwd")
    s2 = d_s2.download(np.float32, 10 * n)
    live = np.repeat(s2_64[2 * n:3 * n] < 1e30, 10)
    ok, mabs, _ = close(s2[live], s2_64[live])
    assert ok, mabs
    ref = oracle.tile_assign_sort(s2, n, W, H)
    assert st["n_isects"] == ref["n_isects"] > 100_000
    assert np.array_equal(st["flatten_ids"], ref["ids"])
    assert np.array_equal(st["tile_ranges"], ref["ranges"])
    comp = oracle.comp_fwd(s2, n, ref["ids"], ref["ranges"], W, H)
    good = np.repeat(comp["margin"] > 5e-5, 3)
    assert (~good).sum() / 3 <= W * H // 500
    ok, mabs, mrel = close(img[good], comp["rgb"][good])
    print(f"[parity] viewer_cam{cam_index}: max_abs={mabs:.3e} max_rel={mrel:.3e}")
    assert ok


def test_renderer_from_ply_matches_uploaded_scene(tmp_path):
    """SplatRenderer::loadPLY: a scene saved as PLY (un-normalised coordinates)
    renders bit-identically to the same file loaded, normalised on the host
    and uploaded; the render of an uploaded slab is pinned against the oracle
    by the parity test above."""
    n, W, H = 20000, 320, 180
    params = ds.cluster_scene(77, n, 3, 16)
    params[0:3 * n] = params[0:3 * n] * np.float32(5.0) + np.float32(2.0)
    path = tmp_path / "scene.ply"
    ds.save_ply(path, params, n, 3)
    r = ds.SplatRenderer.from_ply(path, W, H, scene_scale=1.0)
    assert (r.count, r.sh_degree) == (n, 3)
    slab, _, _ = ds.load_ply(path)
    ds.normalize_scene(slab, n, 1.0)
    u = ds.SplatRenderer(W, H, n, 3)
    u.upload_params(slab)
    for k in (5, 60):
        view, proj, eye = ds.orbit_cameras(120, W, H)[k]
        r.render(view, proj, eye)
        u.render(view, proj, eye)
        a, b = r.image(), u.image()
        assert a.max() > 0.05 and np.array_equal(a, b)
    r.close()
    u.close()


def test_fitter_save_ply_round_trips(tmp_path):
    W, H, n = 96, 64, 800
    view, proj, cam = ds.fitter_camera(W, H)
    fit = ds.ImageFitter(W, H, n, view=view, proj=proj, cam_pos=cam)
    fit.set_target(ds.target_image(3, W, H, 8))
    fit.upload_params(ds.frustum_splats(3, n, aspect=W / H))
    fit.step(5)
    path = tmp_path / "fit.ply"
    fit.save_ply(path)
    slab, count, deg = ds.load_ply(path)
    assert (count, deg) == (n, 0)
    np.testing.assert_array_equal(slab, fit.params())
    fit.close()


def test_viewer_full_size_orbit_invariants():
    """1M splats, SH 3, 1920x1080: colours lie in [0, 1 - T] for every pixel
    of several orbit frames, and re-rendering a frame is bit-identical."""
    n, W, H = 1_000_000, 1920, 1080
    params = ds.cluster_scene(2000, n, 3, 64)
    r = ds.SplatRenderer(W, H, n, 3)
    r.upload_params(params)
    cams = ds.orbit_cameras(120, W, H)
    first = None
    for k in (0, 30, 60, 90, 0):
        view, proj, eye = cams[k]
        r.render(view, proj, eye)
        img = r.image().reshape(-1, 3)
        st = r.raster_state()
        T = st["T"]
        assert st["n_isects"] > 1_000_000 and st["wanted"] == st["n_isects"]
        assert np.all(np.isfinite(img)) and img.min() >= 0.0
        assert np.all((T >= 0) & (T <= 1))
        assert np.all(img <= (1.0 - T)[:, None] + 1e-5)
        if k == 0:
            if first is None:
                first = img.copy()
            else:
                assert np.array_equal(first, img)
