// Interactive cameras that produce the renderer's view/projection matrices
// (OpenGL conventions, column-major, clip z in [-1, 1]).
//
//   DS_CAMERA_ARCBALL   ArcballCamera (src/camera/arcball_camera.cpp:11-144,
//                       arcball_camera.h:40-71): orbit around a target with
//                       yaw/pitch, shift-drag pan, scroll zoom, optional
//                       orthographic projection matched to the perspective view.
//   DS_CAMERA_FLY       FlyCamera (src/camera/fly_camera.cpp:13-125,
//                       fly_camera.h:44-76): WASD/RF translation in the local
//                       frame, Q/E roll, drag yaw about the local up axis and a
//                       clamped pitch kept outside the orientation quaternion.
//
// Each update follows the reference's order of operations (mouse deltas
// clamped to +-50 px, screen y negated; the arcball's zoom uses the view
// direction and distance taken before that frame's orbit), in float, and
// reports whether the matrices changed. The input is a plain struct instead of
// a GLFW-fed object: keys are a bitmask of the ones the cameras read.
#include <cmath>
#include <new>

#include "../../../include/ds_api.h"
#include "../ds_internal.hpp"

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kMaxPitch = kPi * 0.5f - 0.01f;  // both cameras stop 0.01 rad short of the pole
constexpr float kMouseClamp = 50.f;

struct V3 {
    float x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(V3 a) { return std::sqrt(dot(a, a)); }
V3 normalize(V3 a) { return a * (1.f / length(a)); }
float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Unit quaternion w + xi + yj + zk; q * p is the Hamilton product, so
// q * axis_angle(a, n) turns about n expressed in q's local frame.
struct Quat {
    float w, x, y, z;
};
Quat operator*(Quat a, Quat b) {
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
            a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
Quat axis_angle(float angle, V3 n) {
    const float s = std::sin(angle * 0.5f);
    return {std::cos(angle * 0.5f), n.x * s, n.y * s, n.z * s};
}
Quat normalize(Quat q) {
    const float inv = 1.f / std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    return {q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}
// v' = v + 2w (u x v) + 2 u x (u x v), u = (x, y, z): q v q* for a unit q.
V3 rotate(Quat q, V3 v) {
    const V3 u{q.x, q.y, q.z};
    const V3 t = cross(u, v) * 2.f;
    return v + t * q.w + cross(u, t);
}

bool key(const ds_camera_input &in, uint32_t k) { return (in.keys & k) != 0; }

}  // namespace

struct ds_camera {
    int kind;
    float aspect, fovy, znear, zfar;
    ds_camera_speeds sp;
    float view[16], proj[16];
    V3 position{0.f, 0.f, 3.f};
    // arcball
    V3 target{0.f, 0.f, 0.f};
    V3 right{1.f, 0.f, 0.f}, up{0.f, 1.f, 0.f};
    bool ortho = false;
    float ortho_half_h = 1.f;
    // fly
    Quat orientation{1.f, 0.f, 0.f, 0.f};
    float pitch = 0.f;
    V3 plane_forward{0.f, 0.f, -1.f}, plane_right{1.f, 0.f, 0.f}, plane_up{0.f, 1.f, 0.f};

    void projection() {
        if (ortho)
            ds_mat_ortho(-ortho_half_h * aspect, ortho_half_h * aspect, -ortho_half_h,
                         ortho_half_h, znear, zfar, proj);
        else
            ds_mat_perspective(fovy, aspect, znear, zfar, proj);
    }

    // Arcball frame (arcball_camera.cpp:130-135) and matrices (:137-144).
    void arcball_refresh() {
        const V3 forward = normalize(target - position);
        right = normalize(cross(forward, V3{0.f, 1.f, 0.f}));
        up = normalize(cross(right, forward));
        const float e[3] = {position.x, position.y, position.z};
        const float c[3] = {target.x, target.y, target.z};
        const float u[3] = {up.x, up.y, up.z};
        ds_mat_look_at(e, c, u, view);
        projection();
    }

    // Fly frame (fly_camera.cpp:106-113) and matrices (:115-125).
    void fly_refresh() {
        orientation = normalize(orientation);
        plane_forward = rotate(orientation, V3{0.f, 0.f, -1.f});
        plane_right = rotate(orientation, V3{1.f, 0.f, 0.f});
        plane_up = rotate(orientation, V3{0.f, 1.f, 0.f});
        const V3 dir = normalize(plane_forward * std::cos(pitch) + plane_up * std::sin(pitch));
        const V3 vup = normalize(cross(plane_right, dir));
        const V3 ctr = position + dir;
        const float e[3] = {position.x, position.y, position.z};
        const float c[3] = {ctr.x, ctr.y, ctr.z};
        const float u[3] = {vup.x, vup.y, vup.z};
        ds_mat_look_at(e, c, u, view);
        ds_mat_perspective(fovy, aspect, znear, zfar, proj);
    }

    void refresh() { kind == DS_CAMERA_ARCBALL ? arcball_refresh() : fly_refresh(); }

    // arcball_camera.cpp:31-101
    bool arcball_update(const ds_camera_input &in, float dt) {
        const float mx = clampf(in.mouse_dx, -kMouseClamp, kMouseClamp);
        const float my = -clampf(in.mouse_dy, -kMouseClamp, kMouseClamp);
        const V3 offset = position - target;
        const V3 view_dir = normalize(offset * -1.f);
        const float distance = length(offset);
        const bool shift = key(in, DS_KEY_SHIFT);
        const bool drag = in.mouse_left && (mx != 0.f || my != 0.f);
        bool dirty = false;
        if (!shift && drag) {  // orbit
            const V3 dir = normalize(offset);
            float p = std::asin(clampf(dir.y, -1.f, 1.f));
            float y = std::atan2(dir.x, dir.z);
            p -= my * sp.rotate * dt;
            y -= mx * sp.rotate * dt;
            p = clampf(p, -kMaxPitch, kMaxPitch);
            const float two_pi = 2.f * kPi;
            y = y - two_pi * std::floor(y / two_pi);
            const V3 nd{std::sin(y) * std::cos(p), std::sin(p), std::cos(y) * std::cos(p)};
            position = target + nd * distance;
            dirty = true;
        }
        if (shift && drag) {  // pan: eye and target move together in the view plane
            const V3 t = right * (-mx * distance * sp.pan) + up * (-my * distance * sp.pan);
            position = position + t;
            target = target + t;
            dirty = true;
        }
        if (in.scroll != 0.f) {
            if (ortho) {
                ortho_half_h *= (1.f - in.scroll * sp.zoom);
                ortho_half_h = ortho_half_h > 0.01f ? ortho_half_h : 0.01f;
            } else {
                position = position + view_dir * (in.scroll * distance * sp.zoom);
                const float nd = length(position - target);
                if (nd < sp.zoom_min_dist) position = target + view_dir * -sp.zoom_min_dist;
                if (nd > sp.zoom_max_dist) position = target + view_dir * -sp.zoom_max_dist;
            }
            dirty = true;
        }
        if (dirty) arcball_refresh();
        return dirty;
    }

    // fly_camera.cpp:36-93
    bool fly_update(const ds_camera_input &in, float dt) {
        const float mx = clampf(in.mouse_dx, -kMouseClamp, kMouseClamp);
        const float my = -clampf(in.mouse_dy, -kMouseClamp, kMouseClamp);
        float mult = 1.f;
        if (key(in, DS_KEY_SHIFT)) mult *= sp.mult_shift;
        if (key(in, DS_KEY_CTRL)) mult *= sp.mult_ctrl;
        const float move = sp.move * mult * dt;
        const float roll = sp.roll * mult * dt;
        const float look = sp.look * mult * dt;
        bool dirty = false;
        if (key(in, DS_KEY_W)) { position = position + plane_forward * move; dirty = true; }
        if (key(in, DS_KEY_S)) { position = position - plane_forward * move; dirty = true; }
        if (key(in, DS_KEY_A)) { position = position - plane_right * move; dirty = true; }
        if (key(in, DS_KEY_D)) { position = position + plane_right * move; dirty = true; }
        if (key(in, DS_KEY_R)) { position = position + plane_up * move; dirty = true; }
        if (key(in, DS_KEY_F)) { position = position - plane_up * move; dirty = true; }
        const V3 back_axis{0.f, 0.f, -1.f};
        if (key(in, DS_KEY_Q)) { orientation = orientation * axis_angle(-roll, back_axis); dirty = true; }
        if (key(in, DS_KEY_E)) { orientation = orientation * axis_angle(roll, back_axis); dirty = true; }
        if (in.mouse_left) {
            if (mx != 0.f) {
                orientation = orientation * axis_angle(-mx * look, V3{0.f, 1.f, 0.f});
                dirty = true;
            }
            if (my != 0.f) {
                pitch = clampf(pitch + my * look, -kMaxPitch, kMaxPitch);
                dirty = true;
            }
        }
        if (dirty) fly_refresh();
        return dirty;
    }
};

extern "C" {

void ds_camera_default_speeds(ds_camera_speeds *s) {
    if (!s) return;
    // arcball_camera.h:40-44, fly_camera.h:44-48
    s->rotate = 0.2f;
    s->pan = 0.005f;
    s->zoom = 0.1f;
    s->zoom_min_dist = 0.5f;
    s->zoom_max_dist = 50.f;
    s->move = 5.f;
    s->look = 0.2f;
    s->roll = 0.2f;
    s->mult_shift = 4.f;
    s->mult_ctrl = 0.2f;
}

ds_status ds_camera_create(int kind, float aspect, float fovy_deg, float znear, float zfar,
                           ds_camera **out) {
    if (!out) return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_create: null out");
    *out = nullptr;
    if (kind != DS_CAMERA_ARCBALL && kind != DS_CAMERA_FLY)
        return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_create: unknown camera kind");
    if (!(aspect > 0.f) || !(fovy_deg > 0.f && fovy_deg < 180.f) || !(znear > 0.f) ||
        !(zfar > znear))
        return ds::set_error(DS_ERR_INVALID_ARG,
                             "ds_camera_create: need aspect > 0, 0 < fovy < 180, 0 < near < far");
    ds_camera *c = new (std::nothrow) ds_camera();
    if (!c) return ds::set_error(DS_ERR_OUT_OF_MEMORY, "ds_camera_create: out of host memory");
    c->kind = kind;
    c->aspect = aspect;
    c->fovy = fovy_deg * (kPi / 180.f);
    c->znear = znear;
    c->zfar = zfar;
    ds_camera_default_speeds(&c->sp);
    c->refresh();
    *out = c;
    return DS_OK;
}

void ds_camera_destroy(ds_camera *c) { delete c; }

ds_status ds_camera_update(ds_camera *c, const ds_camera_input *in, float dt, int *changed) {
    if (!c || !in) return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_update: null argument");
    const bool dirty = c->kind == DS_CAMERA_ARCBALL ? c->arcball_update(*in, dt)
                                                    : c->fly_update(*in, dt);
    if (changed) *changed = dirty ? 1 : 0;
    return DS_OK;
}

ds_status ds_camera_set_aspect(ds_camera *c, float aspect) {
    if (!c || !(aspect > 0.f))
        return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_set_aspect: need a camera, aspect > 0");
    c->aspect = aspect;  // arcball_camera.cpp:105-113, fly_camera.cpp:97-101
    c->projection();
    return DS_OK;
}

ds_status ds_camera_set_ortho(ds_camera *c, int enable) {
    if (!c) return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_set_ortho: null camera");
    if (c->kind != DS_CAMERA_ARCBALL) return DS_OK;  // Camera::setOrthoMode is a no-op (camera.h:12)
    // arcball_camera.cpp:115-121: the ortho frustum matches the perspective
    // one's height at the target's distance.
    if (enable && !c->ortho)
        c->ortho_half_h = std::tan(c->fovy * 0.5f) * length(c->position - c->target);
    c->ortho = enable != 0;
    c->arcball_refresh();
    return DS_OK;
}

ds_status ds_camera_get_speeds(const ds_camera *c, ds_camera_speeds *s) {
    if (!c || !s) return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_get_speeds: null argument");
    *s = c->sp;
    return DS_OK;
}

ds_status ds_camera_set_speeds(ds_camera *c, const ds_camera_speeds *s) {
    if (!c || !s) return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_set_speeds: null argument");
    if (!(s->zoom_min_dist > 0.f) || !(s->zoom_max_dist >= s->zoom_min_dist))
        return ds::set_error(DS_ERR_INVALID_ARG,
                             "ds_camera_set_speeds: need 0 < zoom_min_dist <= zoom_max_dist");
    c->sp = *s;
    return DS_OK;
}

ds_status ds_camera_matrices(const ds_camera *c, float view[16], float proj[16], float eye[3]) {
    if (!c) return ds::set_error(DS_ERR_INVALID_ARG, "ds_camera_matrices: null camera");
    for (int i = 0; i < 16; ++i) {
        if (view) view[i] = c->view[i];
        if (proj) proj[i] = c->proj[i];
    }
    if (eye) {
        eye[0] = c->position.x;
        eye[1] = c->position.y;
        eye[2] = c->position.z;
    }
    return DS_OK;
}

}  // extern "C"
