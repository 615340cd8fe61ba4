"""Interactive cameras (host code, CPU tests).

ds_camera_* is compared frame by frame with the float64 restatement in
oracle/camera.py on random input sequences. Both are pinned by what the
geometry fixes: GLM's matrices for the start pose, closed-form orbit / pan /
zoom / fly / roll results for single frames, the reference's order inside a
frame (zoom along the pre-orbit direction), clamps (mouse, pitch, zoom
distance, ortho size) and invariants (orbit keeps the distance, pan moves eye
and target together inside the view plane, view rotations stay orthonormal)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import diffsplat_amd as ds
from oracle import camera as ref


class Native:
    """The library's camera behind the oracle's interface (4x4 row-major)."""

    def __init__(self, kind, aspect, fov_deg=40.0, near=0.1, far=100.0):
        cls = ds.ArcballCamera if kind == "arcball" else ds.FlyCamera
        self.c = cls(aspect, fov_deg, near, far)

    def update(self, inp, dt):
        return self.c.update(ds.Input(inp["mouse_delta"], inp["scroll_delta"],
                                      inp["mouse_left_held"], inp["keys"]), dt)

    def set_aspect(self, a):
        self.c.set_aspect(a)

    def set_ortho_mode(self, e):
        self.c.set_ortho_mode(e)

    @property
    def view(self):
        return self.c.view_matrix.reshape(4, 4).T.astype(np.float64)

    @property
    def proj(self):
        return self.c.projection_matrix.reshape(4, 4).T.astype(np.float64)

    @property
    def position(self):
        return self.c.position.astype(np.float64)


def make(impl, kind, *a, **k):
    if impl == "native":
        return Native(kind, *a, **k)
    return (ref.Arcball if kind == "arcball" else ref.Fly)(*a, **k)


def inp(dx=0.0, dy=0.0, scroll=0.0, left=False, keys=()):
    return {"mouse_delta": (dx, dy), "scroll_delta": scroll, "mouse_left_held": left,
            "keys": set(keys)}


IMPLS = ["oracle", "native"]
TOL = dict(rtol=1e-5, atol=2e-5)


def view_space(cam, p):
    return (cam.view @ np.append(p, 1.0))[:3]


# ------------------------------------------------------------- start pose
@pytest.mark.parametrize("impl", IMPLS)
@pytest.mark.parametrize("kind", ["arcball", "fly"])
def test_start_pose_is_glm_look_at_and_perspective(impl, kind):
    cam = make(impl, kind, 1.5, 40.0, 0.1, 100.0)
    want_v = np.eye(4)
    want_v[2, 3] = -3.0  # eye (0,0,3) looking down -z
    np.testing.assert_allclose(cam.view, want_v, atol=1e-7)
    t = math.tan(math.radians(20.0))
    want_p = np.zeros((4, 4))
    want_p[0, 0], want_p[1, 1] = 1 / (1.5 * t), 1 / t
    want_p[2, 2], want_p[2, 3], want_p[3, 2] = -100.1 / 99.9, -20.0 / 99.9, -1.0
    np.testing.assert_allclose(cam.proj, want_p, rtol=1e-6)
    np.testing.assert_allclose(cam.position, [0, 0, 3], atol=0)


# ------------------------------------------------------------- arcball
@pytest.mark.parametrize("impl", IMPLS)
def test_orbit_yaw_and_pitch_closed_forms(impl):
    cam = make(impl, "arcball", 1.0)
    assert cam.update(inp(dx=10, left=True), 1.0)  # yaw 0 -> -2 rad
    np.testing.assert_allclose(cam.position, 3 * np.array([math.sin(-2), 0, math.cos(-2)]), **TOL)
    np.testing.assert_allclose(view_space(cam, np.zeros(3)), [0, 0, -3], **TOL)
    cam = make(impl, "arcball", 1.0)
    cam.update(inp(dy=10, left=True), 0.1)  # dragging down raises the eye by 0.2 rad
    np.testing.assert_allclose(cam.position, 3 * np.array([0, math.sin(0.2), math.cos(0.2)]),
                               **TOL)


@pytest.mark.parametrize("impl", IMPLS)
def test_orbit_clamps_mouse_and_pitch_and_keeps_distance(impl):
    a, b = make(impl, "arcball", 1.0), make(impl, "arcball", 1.0)
    a.update(inp(dx=500, dy=-7, left=True), 0.05)
    b.update(inp(dx=50, dy=-7, left=True), 0.05)
    np.testing.assert_allclose(a.position, b.position, atol=0)
    for _ in range(40):  # keep pulling up: pitch stops 0.01 rad short of the pole
        a.update(inp(dy=50, left=True), 0.1)
        assert np.linalg.norm(a.position) == pytest.approx(3.0, rel=1e-5)
    assert a.position[1] == pytest.approx(3 * math.sin(math.pi / 2 - 0.01), rel=1e-5)
    r = a.view[:3, :3]
    np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-5)


@pytest.mark.parametrize("impl", IMPLS)
def test_pan_moves_eye_and_target_in_the_view_plane(impl):
    cam = make(impl, "arcball", 1.0)
    cam.update(inp(dx=10, left=True, keys={"shift"}), 1.0)  # right * (-10 * 3 * 0.005)
    np.testing.assert_allclose(cam.position, [-0.15, 0, 3], **TOL)
    np.testing.assert_allclose(view_space(cam, np.array([-0.15, 0, 0])), [0, 0, -3], **TOL)
    cam = make(impl, "arcball", 1.0)
    cam.update(inp(dx=13, left=True), 0.05)  # orbit somewhere first
    e0, v0 = cam.position.copy(), cam.view.copy()
    cam.update(inp(dx=-4, dy=9, left=True, keys={"shift"}), 1.0)
    t = cam.position - e0
    fwd = -v0[2, :3]
    assert abs(t @ fwd) < 1e-5
    assert np.linalg.norm(t) == pytest.approx(3 * 0.005 * math.hypot(4, 9), rel=1e-5)
    np.testing.assert_allclose(cam.view[:3, :3], v0[:3, :3], atol=1e-5)  # pure translation


@pytest.mark.parametrize("impl", IMPLS)
def test_zoom_dolly_and_clamps(impl):
    cam = make(impl, "arcball", 1.0)
    cam.update(inp(scroll=1), 0.016)  # dt does not scale zoom
    np.testing.assert_allclose(cam.position, [0, 0, 2.7], **TOL)
    cam.update(inp(scroll=9.5), 0.016)  # 2.565 forward leaves 0.135 < 0.5
    np.testing.assert_allclose(cam.position, [0, 0, 0.5], **TOL)
    cam.update(inp(scroll=-1e4), 0.016)
    np.testing.assert_allclose(cam.position, [0, 0, 50], rtol=1e-6)
    # A dolly through the target lands behind it unclamped when the new
    # distance is in range (the reference checks only the distance).
    cam = make(impl, "arcball", 1.0)
    cam.update(inp(scroll=10), 0.016)  # 3 forward: at the target, clamped back
    np.testing.assert_allclose(cam.position, [0, 0, 0.5], **TOL)
    cam.update(inp(scroll=100), 0.016)  # 5 forward from 0.5
    np.testing.assert_allclose(cam.position, [0, 0, -4.5], **TOL)
    np.testing.assert_allclose(view_space(cam, np.zeros(3)), [0, 0, -4.5], **TOL)


@pytest.mark.parametrize("impl", IMPLS)
def test_orbit_and_zoom_in_one_frame_zoom_along_the_old_direction(impl):
    cam = make(impl, "arcball", 1.0)
    cam.update(inp(dx=10, scroll=1, left=True), 1.0)
    want = 3 * np.array([math.sin(-2), 0, math.cos(-2)]) + np.array([0, 0, -1]) * 0.3
    np.testing.assert_allclose(cam.position, want, **TOL)


@pytest.mark.parametrize("impl", IMPLS)
def test_ortho_mode_matches_the_perspective_frustum_at_the_target(impl):
    cam = make(impl, "arcball", 2.0, 40.0, 0.1, 100.0)
    cam.update(inp(scroll=2), 0.1)  # distance 2.4
    cam.set_ortho_mode(True)
    h = math.tan(math.radians(20)) * 2.4
    p = cam.proj
    assert p[0, 0] == pytest.approx(1 / (2 * h), rel=1e-5)
    assert p[1, 1] == pytest.approx(1 / h, rel=1e-5)
    assert p[2, 2] == pytest.approx(-2 / 99.9, rel=1e-6) and p[3, 3] == 1
    e0 = cam.position.copy()
    cam.update(inp(scroll=2), 0.1)  # ortho zoom scales the frustum, not the eye
    np.testing.assert_allclose(cam.position, e0, atol=0)
    assert cam.proj[1, 1] == pytest.approx(1 / (0.8 * h), rel=1e-5)
    cam.update(inp(scroll=50), 0.1)
    assert cam.proj[1, 1] == pytest.approx(1 / 0.01, rel=1e-5)
    cam.set_ortho_mode(True)  # already on: the size is kept
    assert cam.proj[1, 1] == pytest.approx(1 / 0.01, rel=1e-5)
    cam.set_aspect(0.5)
    assert cam.proj[0, 0] == pytest.approx(1 / (0.5 * 0.01), rel=1e-5)
    cam.set_ortho_mode(False)
    assert cam.proj[3, 2] == -1 and cam.proj[0, 0] == pytest.approx(
        1 / (0.5 * math.tan(math.radians(20))), rel=1e-5)


@pytest.mark.parametrize("impl", IMPLS)
def test_arcball_idle_frames_change_nothing(impl):
    cam = make(impl, "arcball", 1.0)
    assert not cam.update(inp(), 0.1)
    assert not cam.update(inp(left=True), 0.1)  # held without motion
    assert not cam.update(inp(dx=5, dy=5), 0.1)  # motion without the button
    assert not cam.update(inp(keys={"w", "shift"}), 0.1)  # fly keys mean nothing here


# ------------------------------------------------------------- fly
@pytest.mark.parametrize("impl", IMPLS)
def test_fly_moves_along_local_axes_with_speed_multipliers(impl):
    for keys, want in ((("w",), [0, 0, 2.5]), (("s",), [0, 0, 3.5]), (("a",), [-0.5, 0, 3]),
                       (("d",), [0.5, 0, 3]), (("r",), [0, 0.5, 3]), (("f",), [0, -0.5, 3]),
                       (("w", "shift"), [0, 0, 1.0]), (("w", "ctrl"), [0, 0, 2.9]),
                       (("w", "shift", "ctrl"), [0, 0, 2.6]), (("w", "s"), [0, 0, 3])):
        cam = make(impl, "fly", 1.0)
        assert cam.update(inp(keys=keys), 0.1)
        np.testing.assert_allclose(cam.position, want, **TOL)


@pytest.mark.parametrize("impl", IMPLS)
def test_fly_yaw_pitch_roll_closed_forms(impl):
    cam = make(impl, "fly", 1.0)
    cam.update(inp(dx=50, left=True), math.pi / 20)  # -50 * 0.2 * dt = -pi/2 about local up
    cam.update(inp(keys={"w"}), 0.1)  # now forward is +x
    np.testing.assert_allclose(cam.position, [0.5, 0, 3], **TOL)
    np.testing.assert_allclose(view_space(cam, np.array([1.5, 0, 3])), [0, 0, -1], **TOL)

    cam = make(impl, "fly", 1.0)
    cam.update(inp(dy=-10, left=True), 0.1)  # mouse up: pitch +0.2
    np.testing.assert_allclose(cam.view[2, :3], -np.array([0, math.sin(0.2), -math.cos(0.2)]),
                               **TOL)
    cam.update(inp(keys={"w"}), 0.1)  # movement stays in the unpitched plane
    np.testing.assert_allclose(cam.position, [0, 0, 2.5], **TOL)
    for _ in range(20):
        cam.update(inp(dy=-50, left=True), 0.2)
    np.testing.assert_allclose(cam.view[2, 1], -math.sin(math.pi / 2 - 0.01), rtol=1e-5)

    cam = make(impl, "fly", 1.0)
    cam.update(inp(keys={"q"}), 0.5)  # about -z by -0.1: right -> (cos .1, sin .1, 0)
    np.testing.assert_allclose(cam.view[0, :3], [math.cos(0.1), math.sin(0.1), 0], **TOL)
    cam.update(inp(keys={"e"}), 0.5)
    np.testing.assert_allclose(cam.view[:3, :3], np.eye(3), atol=2e-6)


@pytest.mark.parametrize("impl", IMPLS)
def test_fly_idle_and_ortho_is_a_no_op(impl):
    cam = make(impl, "fly", 1.0)
    assert not cam.update(inp(), 0.1)
    assert not cam.update(inp(dx=4, dy=4), 0.1)  # look needs the button
    assert not cam.update(inp(scroll=3), 0.1)
    p0 = cam.proj.copy()
    cam.set_ortho_mode(True)
    np.testing.assert_allclose(cam.proj, p0, atol=0)
    cam.set_aspect(2.0)
    assert cam.proj[0, 0] == pytest.approx(p0[0, 0] / 2, rel=1e-6)


# ------------------------------------------------------------- native vs oracle
@pytest.mark.parametrize("kind", ["arcball", "fly"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_native_follows_the_oracle_on_random_sessions(kind, seed):
    rng = np.random.default_rng(seed)
    a, b = make("oracle", kind, 1.7), make("native", kind, 1.7)
    names = ["shift", "ctrl", "w", "a", "s", "d", "r", "f", "q", "e"]
    for frame in range(80):
        keys = {k for k in names if rng.random() < 0.15}
        i = inp(dx=rng.normal(0, 25), dy=rng.normal(0, 25), left=rng.random() < 0.7,
                scroll=rng.choice([0.0, 0.0, rng.normal(0, 1.5)]), keys=keys)
        dt = float(rng.uniform(0.005, 0.05))
        assert a.update(i, dt) == b.update(i, dt), frame
        if kind == "arcball" and frame in (30, 60):
            a.set_ortho_mode(frame == 30)
            b.set_ortho_mode(frame == 30)
        np.testing.assert_allclose(b.position, a.position, rtol=1e-4, atol=1e-4)
        np.testing.assert_allclose(b.view, a.view, rtol=1e-4, atol=1e-4)
        np.testing.assert_allclose(b.proj, a.proj, rtol=1e-4, atol=1e-6)


def test_speeds_defaults_and_setters():
    cam = ds.ArcballCamera(1.0)
    assert (cam.speed_rotate, cam.speed_zoom, cam.zoom_max_dist) == pytest.approx((0.2, 0.1, 50))
    cam.speed_zoom = 0.5
    cam.update(ds.Input(scroll_delta=1.0), 0.1)
    np.testing.assert_allclose(cam.position, [0, 0, 1.5], rtol=1e-6)
    fly = ds.FlyCamera(1.0)
    assert (fly.speed_move, fly.speed_mult_shift) == pytest.approx((5, 4))
    fly.speed_move = 1.0
    fly.update(ds.Input(keys={"d"}), 0.5)
    np.testing.assert_allclose(fly.position, [0.5, 0, 3], rtol=1e-6)
    with pytest.raises(AttributeError):
        _ = fly.speed_rotate


def test_invalid_arguments():
    for args in ((0.0,), (1.0, 180.0), (1.0, 40.0, 0.0), (1.0, 40.0, 5.0, 5.0)):
        with pytest.raises(ds.DiffsplatError, match="ds_camera_create"):
            ds.ArcballCamera(*args)
    cam = ds.ArcballCamera(1.0)
    with pytest.raises(ds.DiffsplatError):
        cam.set_aspect(-1.0)
    with pytest.raises(ds.DiffsplatError, match="zoom_min_dist"):
        cam.zoom_min_dist = 0.0
