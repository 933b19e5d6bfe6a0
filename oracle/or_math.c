/*
 * or_math.c -- GLM-convention matrices and storage-format conversions for the
 * oracle.  TEST INFRASTRUCTURE ONLY (see vct_oracle.h).
 *
 * Matrices follow GLM 0.9.8 (the version the reference README links,
 * README.md:18): column-major, right-handed, clip depth [-1, 1].  The
 * reference calls glm::ortho / glm::lookAt / glm::perspective / glm::inverse at
 * src/Application.cpp:200-210, 689-692, 804.  GLM's normalize uses
 * inversesqrt; here normalize is v / sqrt(dot(v, v)), the GLSL definition,
 * which the HIP host code uses as well.
 */
#include "vct_oracle.h"

#include <math.h>
#include <string.h>

void or_identity(float m[16]) {
    memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

/* out = a * b (column-major; out[c*4+r] = sum_k a[k*4+r] * b[c*4+k]) */
void or_mul4(const float a[16], const float b[16], float out[16]) {
    float t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float s = a[0 * 4 + r] * b[c * 4 + 0];
            s = s + a[1 * 4 + r] * b[c * 4 + 1];
            s = s + a[2 * 4 + r] * b[c * 4 + 2];
            s = s + a[3 * 4 + r] * b[c * 4 + 3];
            t[c * 4 + r] = s;
        }
    memcpy(out, t, sizeof(t));
}

void or_mat4_vec4(const float m[16], const float v[4], float out[4]) {
    float t[4];
    for (int r = 0; r < 4; ++r) {
        float s = m[0 * 4 + r] * v[0];
        s = s + m[1 * 4 + r] * v[1];
        s = s + m[2 * 4 + r] * v[2];
        s = s + m[3 * 4 + r] * v[3];
        t[r] = s;
    }
    memcpy(out, t, sizeof(t));
}

/* glm::ortho(left, right, bottom, top, zNear, zFar), RH, NO */
void or_ortho(float l, float r, float b, float t, float n, float f, float out[16]) {
    or_identity(out);
    out[0] = 2.0f / (r - l);
    out[5] = 2.0f / (t - b);
    out[10] = -2.0f / (f - n);
    out[12] = -(r + l) / (r - l);
    out[13] = -(t + b) / (t - b);
    out[14] = -(f + n) / (f - n);
}

static void nrm3(const float v[3], float o[3]) {
    float d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    float l = sqrtf(d);
    o[0] = v[0] / l;
    o[1] = v[1] / l;
    o[2] = v[2] / l;
}

static void cross3(const float a[3], const float b[3], float o[3]) {
    float t0 = a[1] * b[2] - a[2] * b[1];
    float t1 = a[2] * b[0] - a[0] * b[2];
    float t2 = a[0] * b[1] - a[1] * b[0];
    o[0] = t0;
    o[1] = t1;
    o[2] = t2;
}

static float dot3(const float a[3], const float b[3]) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* glm::lookAt(eye, center, up), RH */
void or_lookat(const float eye[3], const float center[3], const float up[3], float out[16]) {
    float f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    nrm3(f, f);
    float s[3];
    cross3(f, up, s);
    nrm3(s, s);
    float u[3];
    cross3(s, f, u);
    or_identity(out);
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

/* glm::perspective(fovy, aspect, zNear, zFar), RH, NO */
void or_perspective(float fovy, float aspect, float n, float f, float out[16]) {
    float th = tanf(fovy / 2.0f);
    memset(out, 0, 16 * sizeof(float));
    out[0] = 1.0f / (aspect * th);
    out[5] = 1.0f / th;
    out[10] = -(f + n) / (f - n);
    out[11] = -1.0f;
    out[14] = -(2.0f * f * n) / (f - n);
}

/* 4x4 inverse by cofactor expansion (adjugate / determinant).  Returns 0 for a
 * singular matrix (out untouched). */
int or_inverse4(const float m[16], float out[16]) {
    /* a[r][c] = m[c*4+r] */
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
    if (det == 0.0f) return 0;
    float inv[4][4]; /* inv[r][c] */
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
    return 1;
}

/* mat3(transpose(inverse(model))) (voxelize.vert:18, phong.vert:37) computed
 * as cofactor(A) / det(A) of the upper-left 3x3 block A, column-major out. */
void or_normal_matrix(const float m[16], float out[9]) {
#define A(r, c) m[(c)*4 + (r)]
    float c00 = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
    float c01 = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
    float c02 = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
    float c10 = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
    float c11 = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
    float c12 = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
    float c20 = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
    float c21 = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
    float c22 = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
    float det = (A(0, 0) * c00 + A(0, 1) * c01) + A(0, 2) * c02;
#undef A
    /* N[r][c] = cof[r][c] / det ; out[c*3+r] */
    out[0] = c00 / det;
    out[1] = c10 / det;
    out[2] = c20 / det;
    out[3] = c01 / det;
    out[4] = c11 / det;
    out[5] = c21 / det;
    out[6] = c02 / det;
    out[7] = c12 / det;
    out[8] = c22 / det;
}

/* mvp_x/y/z (src/Application.cpp:689-692):
 *   P = ortho(min.x, max.x, min.y, max.y, 0, max.z - min.z)
 *   mvp_x = P * lookAt(vec3(max.x,0,0) + center, center, (0,1,0))
 *   mvp_y = P * lookAt(vec3(0,max.y,0) + center, center, (0,0,-1))
 *   mvp_z = P * lookAt(vec3(0,0,max.z) + center, center, (0,1,0)) */
void or_voxel_axis_projections(const float c[3], const float mn[3], const float mx[3],
                               float mvp_x[16], float mvp_y[16], float mvp_z[16]) {
    float P[16], V[16];
    or_ortho(mn[0], mx[0], mn[1], mx[1], 0.0f, mx[2] - mn[2], P);
    float ex[3] = {mx[0] + c[0], 0.0f + c[1], 0.0f + c[2]};
    float ey[3] = {0.0f + c[0], mx[1] + c[1], 0.0f + c[2]};
    float ez[3] = {0.0f + c[0], 0.0f + c[1], mx[2] + c[2]};
    const float upy[3] = {0, 1, 0}, upnz[3] = {0, 0, -1};
    or_lookat(ex, c, upy, V);
    or_mul4(P, V, mvp_x);
    or_lookat(ey, c, upnz, V);
    or_mul4(P, V, mvp_y);
    or_lookat(ez, c, upy, V);
    or_mul4(P, V, mvp_z);
}

/* ---------------- IEEE binary16 ------------------------------------------ */
uint16_t or_float_to_half(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) { /* inf or nan */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    }
    if (ax >= 0x477ff000u) { /* >= 65520 rounds to inf */
        return (uint16_t)(sign | 0x7c00u);
    }
    if (ax < 0x38800000u) { /* below 2^-14: half subnormal or zero */
        /* value = ax_float; half subnormal step 2^-24 */
        float a;
        memcpy(&a, &ax, 4);
        float q = rintf(a * 16777216.0f); /* a / 2^-24, exact scale, RNE */
        return (uint16_t)(sign | (uint32_t)q);
    }
    /* normal: round mantissa from 23 to 10 bits, RNE */
    uint32_t e = ((ax >> 23) - 127 + 15);
    uint32_t mant = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1; /* may carry into exponent */
    return (uint16_t)(sign | h);
}

float or_half_to_float(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    float v;
    if (e == 0) {
        v = ldexpf((float)m, -24);
        uint32_t b;
        memcpy(&b, &v, 4);
        b |= sign;
        memcpy(&v, &b, 4);
        return v;
    }
    uint32_t bits;
    if (e == 31) bits = sign | 0x7f800000u | (m << 13);
    else bits = sign | ((e - 15 + 127) << 23) | (m << 13);
    memcpy(&v, &bits, 4);
    return v;
}

/* ---------------- OCP FP8 E4M3 (bias 7, max 448, 0x7f/0xff = NaN) ---------- */
uint8_t or_float_to_fp8(float f) {
    if (isnan(f)) return 0x7f;
    uint8_t sign = signbit(f) ? 0x80 : 0x00;
    float a = fabsf(f);
    if (a >= 448.0f) return (uint8_t)(sign | 0x7e); /* saturate */
    if (a < 0.015625f) {                             /* < 2^-6: subnormal step 2^-9 */
        float q = rintf(a * 512.0f);
        return (uint8_t)(sign | (uint8_t)q); /* q == 8 is the min normal code 0x08 */
    }
    int E;
    float mfr = frexpf(a, &E); /* a = mfr * 2^E, mfr in [0.5, 1) */
    int e = E - 1;             /* a = (2*mfr) * 2^e                 */
    float mant = mfr * 2.0f;   /* [1, 2)                            */
    float q = rintf((mant - 1.0f) * 8.0f);
    if (q >= 8.0f) {
        q = 0.0f;
        e += 1;
    }
    int code = ((e + 7) << 3) | (int)q;
    if (code > 0x7e) code = 0x7e;
    return (uint8_t)(sign | code);
}

float or_fp8_to_float(uint8_t v) {
    int e = (v >> 3) & 15, m = v & 7;
    float s = (v & 0x80) ? -1.0f : 1.0f;
    if ((v & 0x7f) == 0x7f) return NAN;
    if (e == 0) return s * ldexpf((float)m, -9);
    return s * ldexpf(1.0f + (float)m / 8.0f, e - 7);
}

/* packUnorm4x8: round(clamp(c, 0, 1) * 255.0), R in the low byte.  GLSL
 * leaves the direction of a .5 tie to the implementation; ties go to even. */
uint32_t or_pack_unorm4x8(const float c[4]) {
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i) {
        float v = c[i];
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        uint32_t q = (uint32_t)rintf(v * 255.0f);
        r |= (q & 0xffu) << (8 * i);
    }
    return r;
}

void or_unpack_unorm4x8(uint32_t v, float out[4]) {
    for (int i = 0; i < 4; ++i) out[i] = (float)((v >> (8 * i)) & 0xffu) / 255.0f;
}
