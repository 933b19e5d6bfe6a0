/*
 * or_label.c -- FP64 reference of the diffuse gather used as training labels
 * and as the error yardstick of the learned gather.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference's 6-cone table (shaders/phong.frag:470-478) spreads its cones
 * around the tangent-space y axis, i.e. around TBN * (0,1,0).  The 64-cone
 * reference integrates the same hemisphere finely:
 *   - 64 directions, a Fibonacci set uniform in cos(theta) around local +y
 *     (y_i = 1 - (i + 0.5)/64, phi_i = i * golden angle), mapped through the
 *     pixel's TBN like the 6-cone table;
 *   - weights w_i = y_i / 32 (cosine-weighted; they sum to 1 like the table's
 *     0.25 + 5 * 0.15);
 *   - 20-degree apertures (64 cones of 2*pi/64 sr each);
 *   - the traceCone recurrence of phong.frag:145-168 with half the step
 *     (h += r/2, i.e. 0.25 of the footprint diameter per step): each sample
 *     then stands for half a segment, so its opacity is corrected to
 *     1 - (1 - a)^(1/2) and its colour weighted by 1/2;
 *   - up to 256 steps, stopping at alpha >= 0.999 or when leaving [0,1]^3;
 *   - everything in double precision; the pyramid holds the stored values.
 * The result is reported like the debugIndirect view (phong.frag:487-489):
 * rgb * occlusion with occlusion = 1 - clamp(sum_i w_i alpha_i, 0, 1).
 */
#include <math.h>
#include <string.h>

#include "vct_oracle.h"

#define LABEL_CONES 64

void or_label_dirs(double dirs[LABEL_CONES][3], double w[LABEL_CONES]) {
    const double ga = M_PI * (3.0 - sqrt(5.0));
    for (int i = 0; i < LABEL_CONES; ++i) {
        double y = 1.0 - (i + 0.5) / LABEL_CONES;
        double r = sqrt(1.0 - y * y);
        double phi = ga * i;
        dirs[i][0] = r * cos(phi);
        dirs[i][1] = y;
        dirs[i][2] = r * sin(phi);
        w[i] = y / (LABEL_CONES / 2.0);
    }
}

static void lin_d(const float *vol, int s, const double p[3], double out[4]) {
    double u[3], f[3];
    int i0[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = p[k] * s - 0.5;
        double fl = floor(u[k]);
        i0[k] = (int)fl;
        f[k] = u[k] - fl;
    }
    for (int q = 0; q < 4; ++q) out[q] = 0;
    for (int c = 0; c < 8; ++c) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        int x = i0[0] + dx, y = i0[1] + dy, z = i0[2] + dz;
        if (x < 0 || y < 0 || z < 0 || x >= s || y >= s || z >= s) continue;
        double wgt = (dx ? f[0] : 1 - f[0]) * (dy ? f[1] : 1 - f[1]) * (dz ? f[2] : 1 - f[2]);
        const float *t = vol + (((size_t)z * s + y) * s + x) * 4;
        for (int q = 0; q < 4; ++q) out[q] += wgt * (double)t[q];
    }
}

static void sample_d(float *const *levels, int dim, int nlevels, const double p[3], double lam, double out[4]) {
    int q = nlevels - 1;
    if (!(lam < q)) {
        int s = dim >> q;
        lin_d(levels[q], s < 1 ? 1 : s, p, out);
        return;
    }
    if (lam < 0) lam = 0; /* lod + 0.5 > 0 always here */
    int d1 = (int)floor(lam);
    double fr = lam - d1, a[4], b[4];
    int s1 = dim >> d1, s2 = dim >> (d1 + 1);
    lin_d(levels[d1], s1 < 1 ? 1 : s1, p, a);
    lin_d(levels[d1 + 1], s2 < 1 ? 1 : s2, p, b);
    for (int c = 0; c < 4; ++c) out[c] = (1 - fr) * a[c] + fr * b[c];
}

void or_label_cone64(float *const *levels, int dim, int nlevels, const double pos[3], const double nrm[3],
                     const double dir_in[3], double tan_half, double out[4]) {
    double l = sqrt(dir_in[0] * dir_in[0] + dir_in[1] * dir_in[1] + dir_in[2] * dir_in[2]);
    double dir[3] = {dir_in[0] / l, dir_in[1] / l, dir_in[2] / l};
    double c[3] = {0, 0, 0}, a = 0, scale = 1.0 / dim, h = 1.0;
    double start[3];
    for (int k = 0; k < 3; ++k) start[k] = pos[k] + 1.0 * nrm[k] * scale; /* bias 1 voxel along N */
    for (int i = 0; i < 256 && a < 0.999; ++i) {
        double r = h * tan_half;
        double lod = log2(fmax(1.0, 2.0 * r)) + 0.5;
        double sp[3];
        int out_of = 0;
        for (int k = 0; k < 3; ++k) {
            sp[k] = start[k] + h * dir[k] * scale;
            if (sp[k] < 0 || sp[k] > 1) out_of = 1;
        }
        if (out_of) break;
        double s[4];
        sample_d(levels, dim, nlevels, sp, lod, s);
        double sa = s[3] < 0 ? 0 : (s[3] > 1 ? 1 : s[3]);
        double ah = 1.0 - sqrt(1.0 - sa);
        double tr = 1.0 - a;
        for (int k = 0; k < 3; ++k) c[k] += 0.5 * s[k] * tr;
        a += tr * ah;
        h += 0.5 * r;
    }
    out[0] = c[0];
    out[1] = c[1];
    out[2] = c[2];
    out[3] = a;
}

/* One G-buffer texel -> rgb * occlusion of the 64-cone gather (out[3] = occlusion). */
void or_label64(const or_gbuf_texel *g, int npix, float *const *levels, int dim, int nlevels,
                const float center[3], const float vmin[3], const float vmax[3], double *out) {
    double dirs[LABEL_CONES][3], w[LABEL_CONES];
    or_label_dirs(dirs, w);
    const double tan_half = tan(10.0 * M_PI / 180.0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < npix; ++i) {
        double *o = out + (size_t)i * 4;
        o[0] = o[1] = o[2] = o[3] = 0;
        const or_gbuf_texel *t = &g[i];
        if (!t->covered) continue;
        double pos[3], N[3], n2 = 0;
        for (int k = 0; k < 3; ++k) {
            pos[k] = ((double)t->position[k] - center[k] - vmin[k]) / ((double)vmax[k] - vmin[k]);
            N[k] = t->normal[k];
            n2 += N[k] * N[k];
        }
        n2 = sqrt(n2);
        for (int k = 0; k < 3; ++k) N[k] /= n2;
        double acc[4] = {0, 0, 0, 0};
        for (int c = 0; c < LABEL_CONES; ++c) {
            double d[3];
            for (int k = 0; k < 3; ++k)
                d[k] = t->tangent[k] * dirs[c][0] + t->bitangent[k] * dirs[c][1] + t->normal[k] * dirs[c][2];
            double r[4];
            or_label_cone64(levels, dim, nlevels, pos, N, d, tan_half, r);
            for (int q = 0; q < 4; ++q) acc[q] += w[c] * r[q];
        }
        double occ = 1.0 - (acc[3] < 0 ? 0 : (acc[3] > 1 ? 1 : acc[3]));
        o[0] = acc[0] * occ;
        o[1] = acc[1] * occ;
        o[2] = acc[2] * occ;
        o[3] = occ;
    }
}
