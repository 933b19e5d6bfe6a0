/*
 * or_trace.c -- FP32 and FP64 instantiations of the phong.frag cone gather.
 * TEST INFRASTRUCTURE ONLY (see vct_oracle.h).
 */
#include "vct_oracle.h"

#include <math.h>
#include <string.h>

#define REAL float
#define SUFFIX _f
#define SQRT sqrtf
#define FLOOR floorf
#define LOG2 log2f
#define TAN tanf
#include "or_trace_impl.h"
#undef REAL
#undef SUFFIX
#undef SQRT
#undef FLOOR
#undef LOG2
#undef TAN

#define REAL double
#define SUFFIX _d
#define SQRT sqrt
#define FLOOR floor
#define LOG2 log2
#define TAN tan
#include "or_trace_impl.h"
#undef REAL
#undef SUFFIX
#undef SQRT
#undef FLOOR
#undef LOG2
#undef TAN

void or_trace(const or_gbuf_texel *g, int npix, float *const *levels, float *const *const *aniso,
              const or_trace_params *p, float *out_indirect, float *out_specular) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < npix; ++i) {
        float ind[4], spec[4];
        trace_pixel_f(&g[i], levels, aniso, p, ind, spec);
        memcpy(out_indirect + (size_t)i * 4, ind, 16);
        if (out_specular) memcpy(out_specular + (size_t)i * 4, spec, 16);
    }
}

void or_trace64(const or_gbuf_texel *g, int npix, float *const *levels, float *const *const *aniso,
                const or_trace_params *p, double *out_indirect, double *out_specular) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < npix; ++i) {
        double ind[4], spec[4];
        trace_pixel_d(&g[i], levels, aniso, p, ind, spec);
        memcpy(out_indirect + (size_t)i * 4, ind, 32);
        if (out_specular) memcpy(out_specular + (size_t)i * 4, spec, 32);
    }
}

void or_trace_cone(float *const *levels, float *const *const *aniso, int dim, int nlevels, const float pos[3],
                   const float nrm[3], const float dir[3], const or_cone *c, float out[4]) {
    trace_cone_f(levels, aniso, dim, nlevels, pos, nrm, dir, c->steps, c->bias, tanf(c->angle / 2.0f), c->h0,
                 c->lod_offset, NULL, out);
}

void or_sample(float *const *levels, int dim, int nlevels, const float p[3], float lambda, float out[4]) {
    const float dir[3] = {0, 0, 1};
    sample_lod_f(levels, NULL, dim, nlevels, p, dir, lambda, out);
}
