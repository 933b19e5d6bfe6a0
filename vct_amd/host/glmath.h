// glmath.h -- GLM-convention matrices for the host application layer.
//
// Column-major, right-handed, clip depth [-1, 1], as glm::ortho / lookAt /
// perspective / inverse (GLM 0.9.8, which the reference README links) are
// called at src/Application.cpp:200-210, 689-692, 804.  normalize is
// v / sqrt(dot(v, v)) (the GLSL definition).  Operation order is fixed and
// the file is compiled without FP contraction so the matrices are
// reproducible bit for bit.
#pragma once

#include <cmath>
#include <cstring>

namespace vcthost {

inline void identity(float m[16]) {
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

inline void mul4(const float a[16], const float b[16], float out[16]) {
    float t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float s = a[0 * 4 + r] * b[c * 4 + 0];
            s = s + a[1 * 4 + r] * b[c * 4 + 1];
            s = s + a[2 * 4 + r] * b[c * 4 + 2];
            s = s + a[3 * 4 + r] * b[c * 4 + 3];
            t[c * 4 + r] = s;
        }
    std::memcpy(out, t, sizeof(t));
}

inline void nrm3(const float v[3], float o[3]) {
    float l = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    o[0] = v[0] / l;
    o[1] = v[1] / l;
    o[2] = v[2] / l;
}
inline void cross3(const float a[3], const float b[3], float o[3]) {
    float t0 = a[1] * b[2] - a[2] * b[1], t1 = a[2] * b[0] - a[0] * b[2], t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
}
inline float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

inline void ortho(float l, float r, float b, float t, float n, float f, float out[16]) {
    identity(out);
    out[0] = 2.0f / (r - l);
    out[5] = 2.0f / (t - b);
    out[10] = -2.0f / (f - n);
    out[12] = -(r + l) / (r - l);
    out[13] = -(t + b) / (t - b);
    out[14] = -(f + n) / (f - n);
}

inline void lookat(const float eye[3], const float center[3], const float up[3], float out[16]) {
    float f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    nrm3(f, f);
    float s[3];
    cross3(f, up, s);
    nrm3(s, s);
    float u[3];
    cross3(s, f, u);
    identity(out);
    out[0] = s[0];
    out[4] = s[1];
    out[8] = s[2];
    out[1] = u[0];
    out[5] = u[1];
    out[9] = u[2];
    out[2] = -f[0];
    out[6] = -f[1];
    out[10] = -f[2];
    out[12] = -dot3(s, eye);
    out[13] = -dot3(u, eye);
    out[14] = dot3(f, eye);
}

inline void perspective(float fovy, float aspect, float n, float f, float out[16]) {
    float th = std::tan(fovy / 2.0f);
    std::memset(out, 0, 16 * sizeof(float));
    out[0] = 1.0f / (aspect * th);
    out[5] = 1.0f / th;
    out[10] = -(f + n) / (f - n);
    out[11] = -1.0f;
    out[14] = -(2.0f * f * n) / (f - n);
}

// cofactor expansion; returns false for a singular matrix
inline bool inverse4(const float m[16], float out[16]) {
#define A(r, c) m[(c)*4 + (r)]
    float s0 = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1);
    float s1 = A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2);
    float s2 = A(0, 0) * A(1, 3) - A(1, 0) * A(0, 3);
    float s3 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    float s4 = A(0, 1) * A(1, 3) - A(1, 1) * A(0, 3);
    float s5 = A(0, 2) * A(1, 3) - A(1, 2) * A(0, 3);
    float c5 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3);
    float c4 = A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3);
    float c3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2);
    float c2 = A(2, 0) * A(3, 3) - A(3, 0) * A(2, 3);
    float c1 = A(2, 0) * A(3, 2) - A(3, 0) * A(2, 2);
    float c0 = A(2, 0) * A(3, 1) - A(3, 0) * A(2, 1);
    float det = ((((s0 * c5 - s1 * c4) + s2 * c3) + s3 * c2) - s4 * c1) + s5 * c0;
    if (det == 0.0f) return false;
    float inv[4][4];
    inv[0][0] = (A(1, 1) * c5 - A(1, 2) * c4) + A(1, 3) * c3;
    inv[0][1] = (-A(0, 1) * c5 + A(0, 2) * c4) - A(0, 3) * c3;
    inv[0][2] = (A(3, 1) * s5 - A(3, 2) * s4) + A(3, 3) * s3;
    inv[0][3] = (-A(2, 1) * s5 + A(2, 2) * s4) - A(2, 3) * s3;
    inv[1][0] = (-A(1, 0) * c5 + A(1, 2) * c2) - A(1, 3) * c1;
    inv[1][1] = (A(0, 0) * c5 - A(0, 2) * c2) + A(0, 3) * c1;
    inv[1][2] = (-A(3, 0) * s5 + A(3, 2) * s2) - A(3, 3) * s1;
    inv[1][3] = (A(2, 0) * s5 - A(2, 2) * s2) + A(2, 3) * s1;
    inv[2][0] = (A(1, 0) * c4 - A(1, 1) * c2) + A(1, 3) * c0;
    inv[2][1] = (-A(0, 0) * c4 + A(0, 1) * c2) - A(0, 3) * c0;
    inv[2][2] = (A(3, 0) * s4 - A(3, 1) * s2) + A(3, 3) * s0;
    inv[2][3] = (-A(2, 0) * s4 + A(2, 1) * s2) - A(2, 3) * s0;
    inv[3][0] = (-A(1, 0) * c3 + A(1, 1) * c1) - A(1, 2) * c0;
    inv[3][1] = (A(0, 0) * c3 - A(0, 1) * c1) + A(0, 2) * c0;
    inv[3][2] = (-A(3, 0) * s3 + A(3, 1) * s1) - A(3, 2) * s0;
    inv[3][3] = (A(2, 0) * s3 - A(2, 1) * s1) + A(2, 2) * s0;
#undef A
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c * 4 + r] = inv[r][c] / det;
    return true;
}

// mvp_x/y/z of src/Application.cpp:689-692
inline void voxel_projections(const float c[3], const float mn[3], const float mx[3], float mvp_x[16],
                              float mvp_y[16], float mvp_z[16]) {
    float P[16], V[16];
    ortho(mn[0], mx[0], mn[1], mx[1], 0.0f, mx[2] - mn[2], P);
    const float ex[3] = {mx[0] + c[0], 0.0f + c[1], 0.0f + c[2]};
    const float ey[3] = {0.0f + c[0], mx[1] + c[1], 0.0f + c[2]};
    const float ez[3] = {0.0f + c[0], 0.0f + c[1], mx[2] + c[2]};
    const float upy[3] = {0, 1, 0}, upnz[3] = {0, 0, -1};
    lookat(ex, c, upy, V);
    mul4(P, V, mvp_x);
    lookat(ey, c, upnz, V);
    mul4(P, V, mvp_y);
    lookat(ez, c, upy, V);
    mul4(P, V, mvp_z);
}

}  // namespace vcthost
