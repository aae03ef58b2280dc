// Host-side camera math for primary rays (the only libm-dependent part of
// the render path: cosf/sinf/tanf stay on the host).
//
// Orbit camera as the reference viewer sets it up (src/Main.cpp:212,244-250):
// MODEL = rotXYZ(pitch,0,0) * rotXYZ(0,yaw,0), VIEW = translate(0,0,-radius),
// tform = pseudoInvert(MODEL) * VIEW (MatrixStack::get(INV_MODELVIEW_STACK),
// src/math/MatrixStack.cpp:65-66), ray origin = tform*0 + center + 1
// (src/Main.cpp:150-152). Matrix products accumulate left to right in single
// precision (src/math/Mat4.cpp:67-86), which fixes the float results.
#include <cmath>
#include <cstring>

#include "svo_internal.hpp"

namespace svo {
namespace {

struct M4 {
    float a[16];
};

M4 identity() {
    M4 m{};
    m.a[0] = m.a[5] = m.a[10] = m.a[15] = 1.0f;
    return m;
}

M4 mul(const M4 &x, const M4 &y) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) {
            float s = x.a[i * 4 + 0] * y.a[k];
            s = s + x.a[i * 4 + 1] * y.a[4 + k];
            s = s + x.a[i * 4 + 2] * y.a[8 + k];
            s = s + x.a[i * 4 + 3] * y.a[12 + k];
            r.a[i * 4 + k] = s;
        }
    return r;
}

M4 translate(float x, float y, float z) {
    M4 m = identity();
    m.a[3] = x;
    m.a[7] = y;
    m.a[11] = z;
    return m;
}

// Mat4::rotXYZ, src/math/Mat4.cpp:114-125 (angles in degrees)
M4 rot_xyz(float rx, float ry, float rz) {
    const float pi = float(3.14159265358979323846);
    const float r[3] = {rx * pi / 180.0f, ry * pi / 180.0f, rz * pi / 180.0f};
    float c[3], s[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = std::cos(r[k]);
        s[k] = std::sin(r[k]);
    }
    M4 m;
    m.a[0] = c[1] * c[2];
    m.a[1] = -c[0] * s[2] + s[0] * s[1] * c[2];
    m.a[2] = s[0] * s[2] + c[0] * s[1] * c[2];
    m.a[3] = 0.0f;
    m.a[4] = c[1] * s[2];
    m.a[5] = c[0] * c[2] + s[0] * s[1] * s[2];
    m.a[6] = -s[0] * c[2] + c[0] * s[1] * s[2];
    m.a[7] = 0.0f;
    m.a[8] = -s[1];
    m.a[9] = s[0] * c[1];
    m.a[10] = c[0] * c[1];
    m.a[11] = 0.0f;
    m.a[12] = m.a[13] = m.a[14] = 0.0f;
    m.a[15] = 1.0f;
    return m;
}

// Mat4::pseudoInvert, src/math/Mat4.cpp:59-65: transpose the rotation, undo
// the translation.
M4 pseudo_invert(const M4 &x) {
    M4 t = translate(-x.a[3], -x.a[7], -x.a[11]);
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) r.a[i * 4 + k] = x.a[k * 4 + i];
    r.a[12] = r.a[13] = r.a[14] = 0.0f;
    return mul(r, t);
}

}  // namespace

void orbit_camera(float pitch, float yaw, float radius, const float center[3], svo_camera &cam) {
    M4 model = mul(rot_xyz(pitch, 0.0f, 0.0f), rot_xyz(0.0f, yaw, 0.0f));
    M4 tf = mul(pseudo_invert(model), translate(0.0f, 0.0f, -radius));
    std::memcpy(cam.tform, tf.a, sizeof(tf.a));
    for (int r = 0; r < 3; ++r) {
        float p = tf.a[r * 4 + 0] * 0.0f;
        p = p + tf.a[r * 4 + 1] * 0.0f;
        p = p + tf.a[r * 4 + 2] * 0.0f;
        p = p + tf.a[r * 4 + 3];
        cam.pos[r] = (p + center[r]) + 1.0f;
    }
}

bool camera_ok(const svo_camera &cam) {
    for (float v : cam.tform)
        if (!(std::fabs(v) <= 1e6f)) return false;
    for (float v : cam.pos)
        if (!(std::fabs(v) <= 1e6f)) return false;
    return true;
}

}  // namespace svo
