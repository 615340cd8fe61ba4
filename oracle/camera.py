"""Interactive cameras, restated in float64 numpy -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of ds_camera_* (the engine
never calls it). It follows the reference's ArcballCamera
(src/camera/arcball_camera.cpp:11-144, constants arcball_camera.h:40-71) and
FlyCamera (src/camera/fly_camera.cpp:13-125, constants fly_camera.h:44-76)
step by step; GLM's lookAt / perspective / ortho are written out from their
published definitions (right-handed, clip z in [-1, 1]).

Matrices are returned as 4x4 row-major arrays M acting on column vectors
(M @ [x, y, z, 1]); the library's 16-float column-major output is M.T.ravel().

Input: dict-like with mouse_delta (dx, dy), scroll_delta, mouse_left_held,
keys (set of "shift", "ctrl", "w", "a", "s", "d", "r", "f", "q", "e").
"""
from __future__ import annotations

import math

import numpy as np

MAX_PITCH = math.pi / 2 - 0.01  # arcball_camera.h:70-71, fly_camera.h:75-76


def normalize(v):
    return v / np.linalg.norm(v)


def look_at(eye, center, up):
    """glm::lookAt (right-handed): rows s, u, -f and the eye moved to the origin."""
    f = normalize(center - eye)
    s = normalize(np.cross(f, up))
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def perspective(fovy, aspect, n, f):
    """glm::perspective (right-handed, z in [-1, 1])."""
    t = math.tan(fovy / 2)
    m = np.zeros((4, 4))
    m[0, 0], m[1, 1] = 1 / (aspect * t), 1 / t
    m[2, 2], m[2, 3] = -(f + n) / (f - n), -2 * f * n / (f - n)
    m[3, 2] = -1
    return m


def ortho(l, r, b, t, n, f):
    """glm::ortho (right-handed, z in [-1, 1])."""
    m = np.eye(4)
    m[0, 0], m[1, 1], m[2, 2] = 2 / (r - l), 2 / (t - b), -2 / (f - n)
    m[0, 3], m[1, 3], m[2, 3] = -(r + l) / (r - l), -(t + b) / (t - b), -(f + n) / (f - n)
    return m


def qmul(a, b):
    """Hamilton product of (w, x, y, z) quaternions."""
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def angle_axis(angle, axis):
    return np.concatenate([[math.cos(angle / 2)], np.asarray(axis, float) * math.sin(angle / 2)])


def qrotate(q, v):
    """q v q^-1 for a unit quaternion, via the rotation matrix."""
    w, x, y, z = q
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return r @ np.asarray(v, float)


def _mouse(inp):
    dx, dy = inp["mouse_delta"]
    return float(np.clip(dx, -50, 50)), -float(np.clip(dy, -50, 50))


class Arcball:
    """ArcballCamera (arcball_camera.cpp:11-27 constructor)."""

    def __init__(self, aspect, fov_deg=40.0, near=0.1, far=100.0):
        self.aspect, self.fov, self.near, self.far = aspect, math.radians(fov_deg), near, far
        self.rotate, self.pan, self.zoom, self.zmin, self.zmax = 0.2, 0.005, 0.1, 0.5, 50.0
        self.position, self.target = np.array([0.0, 0.0, 3.0]), np.zeros(3)
        self.ortho, self.half_h = False, 1.0
        self._refresh()

    def _refresh(self):  # updateViewSpaceVectors + updateMatrices (:130-144)
        fwd = normalize(self.target - self.position)
        self.right = normalize(np.cross(fwd, [0.0, 1.0, 0.0]))
        self.up = normalize(np.cross(self.right, fwd))
        self.view = look_at(self.position, self.target, self.up)
        self._proj()

    def _proj(self):
        h, a = self.half_h, self.aspect
        self.proj = (ortho(-h * a, h * a, -h, h, self.near, self.far) if self.ortho
                     else perspective(self.fov, a, self.near, self.far))

    def update(self, inp, dt):  # :31-101
        mx, my = _mouse(inp)
        offset = self.position - self.target
        view_dir = normalize(-offset)
        dist = np.linalg.norm(offset)
        shift = "shift" in inp["keys"]
        drag = inp["mouse_left_held"] and (mx != 0 or my != 0)
        dirty = False
        if not shift and drag:
            d = normalize(offset)
            pitch = math.asin(np.clip(d[1], -1, 1)) - my * self.rotate * dt
            yaw = math.atan2(d[0], d[2]) - mx * self.rotate * dt
            pitch = float(np.clip(pitch, -MAX_PITCH, MAX_PITCH))
            yaw = yaw - 2 * math.pi * math.floor(yaw / (2 * math.pi))
            nd = np.array([math.sin(yaw) * math.cos(pitch), math.sin(pitch),
                           math.cos(yaw) * math.cos(pitch)])
            self.position = self.target + nd * dist
            dirty = True
        if shift and drag:
            t = self.right * (-mx * dist * self.pan) + self.up * (-my * dist * self.pan)
            self.position, self.target = self.position + t, self.target + t
            dirty = True
        if inp["scroll_delta"] != 0:
            s = inp["scroll_delta"]
            if self.ortho:
                self.half_h = max(self.half_h * (1 - s * self.zoom), 0.01)
            else:
                self.position = self.position + view_dir * (s * dist * self.zoom)
                nd = np.linalg.norm(self.position - self.target)
                if nd < self.zmin:
                    self.position = self.target - view_dir * self.zmin
                if nd > self.zmax:
                    self.position = self.target - view_dir * self.zmax
            dirty = True
        if dirty:
            self._refresh()
        return dirty

    def set_aspect(self, aspect):  # :105-113
        self.aspect = aspect
        self._proj()

    def set_ortho_mode(self, enable):  # :115-121
        if enable and not self.ortho:
            self.half_h = math.tan(self.fov / 2) * np.linalg.norm(self.position - self.target)
        self.ortho = bool(enable)
        self._refresh()


class Fly:
    """FlyCamera (fly_camera.cpp:13-32 constructor)."""

    def __init__(self, aspect, fov_deg=40.0, near=0.1, far=100.0):
        self.aspect, self.fov, self.near, self.far = aspect, math.radians(fov_deg), near, far
        self.move, self.look, self.roll, self.mshift, self.mctrl = 5.0, 0.2, 0.2, 4.0, 0.2
        self.position, self.q, self.pitch = np.array([0.0, 0.0, 3.0]), np.array([1.0, 0, 0, 0]), 0.0
        self._refresh()

    def _refresh(self):  # updateViewSpaceVectors + updateMatrices (:106-125)
        self.q = self.q / np.linalg.norm(self.q)
        self.fwd = qrotate(self.q, [0, 0, -1])
        self.right = qrotate(self.q, [1, 0, 0])
        self.up = qrotate(self.q, [0, 1, 0])
        d = normalize(self.fwd * math.cos(self.pitch) + self.up * math.sin(self.pitch))
        vup = normalize(np.cross(self.right, d))
        self.view = look_at(self.position, self.position + d, vup)
        self.proj = perspective(self.fov, self.aspect, self.near, self.far)

    def update(self, inp, dt):  # :36-93
        mx, my = _mouse(inp)
        keys = inp["keys"]
        mult = (self.mshift if "shift" in keys else 1.0) * (self.mctrl if "ctrl" in keys else 1.0)
        move, roll, look = self.move * mult * dt, self.roll * mult * dt, self.look * mult * dt
        dirty = False
        for k, axis, sign in (("w", self.fwd, 1), ("s", self.fwd, -1), ("a", self.right, -1),
                              ("d", self.right, 1), ("r", self.up, 1), ("f", self.up, -1)):
            if k in keys:
                self.position = self.position + sign * axis * move
                dirty = True
        if "q" in keys:
            self.q, dirty = qmul(self.q, angle_axis(-roll, [0, 0, -1])), True
        if "e" in keys:
            self.q, dirty = qmul(self.q, angle_axis(roll, [0, 0, -1])), True
        if inp["mouse_left_held"]:
            if mx != 0:
                self.q, dirty = qmul(self.q, angle_axis(-mx * look, [0, 1, 0])), True
            if my != 0:
                self.pitch = float(np.clip(self.pitch + my * look, -MAX_PITCH, MAX_PITCH))
                dirty = True
        if dirty:
            self._refresh()
        return dirty

    def set_aspect(self, aspect):  # :97-101
        self.aspect = aspect
        self.proj = perspective(self.fov, aspect, self.near, self.far)

    def set_ortho_mode(self, enable):  # Camera::setOrthoMode default (camera.h:12)
        pass
