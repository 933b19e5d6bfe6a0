/*
 * or_trace_impl.h -- cone gather of phong.frag, instantiated twice by
 * or_trace.c: REAL = float (the as-GL FP32 restatement) and REAL = double (the
 * FP64 oracle of the cone integral).  TEST INFRASTRUCTURE ONLY.
 *
 * Follows shaders/phong.frag:135-180 (traceCone) and :455-512 (the 6-cone
 * diffuse gather, occlusion, the specular cone) with the sampler state of
 * src/Application.h:144-146 and src/Application.cpp:1094-1098:
 * LINEAR_MIPMAP_LINEAR minification, NEAREST magnification, CLAMP_TO_BORDER
 * with border (0,0,0,0), levels 0..L-1.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

static REAL FN(dot3)(const REAL a[3], const REAL b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static void FN(nrm3)(const REAL v[3], REAL o[3]) {
    REAL l = SQRT(FN(dot3)(v, v));
    o[0] = v[0] / l;
    o[1] = v[1] / l;
    o[2] = v[2] / l;
}

/* GL LINEAR filter inside one level of a float4 volume of edge s. */
static void FN(linear)(const float *vol, int s, const REAL p[3], REAL out[4]) {
    REAL u[3], f[3];
    int i0[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = p[k] * (REAL)s - (REAL)0.5;
        REAL fl = FLOOR(u[k]);
        i0[k] = (int)fl;
        f[k] = u[k] - fl;
    }
    for (int q = 0; q < 4; ++q) out[q] = 0;
    /* GL spec order: i0j0k0, i1j0k0, i0j1k0, i1j1k0, i0j0k1, i1j0k1, i0j1k1, i1j1k1 */
    for (int c = 0; c < 8; ++c) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
        int x = i0[0] + dx, y = i0[1] + dy, z = i0[2] + dz;
        REAL w = ((dx ? f[0] : (REAL)1 - f[0]) * (dy ? f[1] : (REAL)1 - f[1])) * (dz ? f[2] : (REAL)1 - f[2]);
        if (x < 0 || y < 0 || z < 0 || x >= s || y >= s || z >= s) continue; /* border 0 */
        const float *t = vol + (((size_t)z * s + y) * s + x) * 4;
        for (int q = 0; q < 4; ++q) out[q] = out[q] + w * (REAL)t[q];
    }
}

/* One mip level: isotropic, or the direction-weighted mix of three faces. */
static void FN(level_sample)(float *const *levels, float *const *const *aniso, int dim, int l,
                              const REAL p[3], const REAL dir[3], REAL out[4]) {
    int s = dim >> l;
    if (s < 1) s = 1;
    if (!aniso || l == 0) {
        FN(linear)(levels[l], s, p, out);
        return;
    }
    REAL acc[4] = {0, 0, 0, 0};
    for (int ax = 0; ax < 3; ++ax) {
        REAL w = dir[ax] * dir[ax];
        int face = ax * 2 + (dir[ax] >= 0 ? 0 : 1);
        REAL t[4];
        FN(linear)(aniso[l][face], s, p, t);
        for (int q = 0; q < 4; ++q) acc[q] = acc[q] + w * t[q];
    }
    for (int q = 0; q < 4; ++q) out[q] = acc[q];
}

/* textureLod with LINEAR_MIPMAP_LINEAR and the lambda clamp to [0, q]. */
static void FN(sample_lod)(float *const *levels, float *const *const *aniso, int dim, int nlevels,
                            const REAL p[3], const REAL dir[3], REAL lambda, REAL out[4]) {
    int q = nlevels - 1;
    if (lambda <= 0) { /* magnification: NEAREST on level 0 */
        int s = dim;
        int i[3];
        for (int k = 0; k < 3; ++k) i[k] = (int)FLOOR(p[k] * (REAL)s);
        for (int c = 0; c < 4; ++c) out[c] = 0;
        if (i[0] < 0 || i[1] < 0 || i[2] < 0 || i[0] >= s || i[1] >= s || i[2] >= s) return;
        const float *t = levels[0] + (((size_t)i[2] * s + i[1]) * s + i[0]) * 4;
        for (int c = 0; c < 4; ++c) out[c] = (REAL)t[c];
        return;
    }
    if (!(lambda < (REAL)q)) { /* lambda >= q (or NaN): the last level only */
        FN(level_sample)(levels, aniso, dim, q, p, dir, out);
        return;
    }
    REAL fl = FLOOR(lambda);
    int d1 = (int)fl;
    d1 = d1 < 0 ? 0 : (d1 > q - 1 ? q - 1 : d1);
    REAL fr = lambda - fl;
    REAL t1[4], t2[4];
    FN(level_sample)(levels, aniso, dim, d1, p, dir, t1);
    FN(level_sample)(levels, aniso, dim, d1 + 1, p, dir, t2);
    for (int c = 0; c < 4; ++c) out[c] = ((REAL)1 - fr) * t1[c] + fr * t2[c];
}

/* texture(warpmap, p).xyz on the 32^3 warp texture: LINEAR, CLAMP_TO_EDGE */
static void FN(warp_sample)(const float *warpmap, const REAL tc[3], REAL out[3]) {
    int i0[3], i1[3];
    REAL f[3];
    for (int k = 0; k < 3; ++k) {
        REAL u = tc[k] * (REAL)OR_WARP_DIM - (REAL)0.5;
        REAL fl = FLOOR(u);
        f[k] = u - fl;
        int a = (int)fl, b = a + 1;
        i0[k] = a < 0 ? 0 : (a > OR_WARP_DIM - 1 ? OR_WARP_DIM - 1 : a);
        i1[k] = b < 0 ? 0 : (b > OR_WARP_DIM - 1 ? OR_WARP_DIM - 1 : b);
    }
    REAL acc[3] = {0, 0, 0};
    for (int c = 0; c < 8; ++c) {
        int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        REAL w = ((dx ? f[0] : (REAL)1 - f[0]) * (dy ? f[1] : (REAL)1 - f[1])) * (dz ? f[2] : (REAL)1 - f[2]);
        int x = dx ? i1[0] : i0[0], y = dy ? i1[1] : i0[1], z = dz ? i1[2] : i0[2];
        const float *t = warpmap + (((size_t)z * OR_WARP_DIM + y) * OR_WARP_DIM + x) * 4;
        for (int k = 0; k < 3; ++k) acc[k] = acc[k] + w * (REAL)t[k];
    }
    for (int k = 0; k < 3; ++k) out[k] = acc[k];
}

/* traceCone (phong.frag:135-180).  tan_half = tan(coneAngle / 2).  warpmap:
 * warpTexture (:150-152), every sample position goes through the warp texture. */
static void FN(trace_cone)(float *const *levels, float *const *const *aniso, int dim, int nlevels,
                            const REAL pos[3], const REAL nrm[3], const REAL dir_in[3], int steps,
                            REAL bias, REAL tan_half, REAL h0, REAL lod_offset, const float *warpmap,
                            REAL out[4]) {
    REAL dir[3];
    FN(nrm3)(dir_in, dir);
    REAL c[3] = {0, 0, 0}, a = 0;
    REAL scale = (REAL)1 / (REAL)dim;
    REAL start[3];
    for (int k = 0; k < 3; ++k) start[k] = pos[k] + (bias * nrm[k]) * scale;
    REAL h = h0;
    for (int i = 0; i < steps && a < (REAL)0.95; ++i) {
        REAL r = h * tan_half;
        REAL m = (REAL)2 * r;
        REAL lod = LOG2(m > (REAL)1 ? m : (REAL)1);
        REAL sp[3];
        int outside = 0;
        for (int k = 0; k < 3; ++k) {
            sp[k] = start[k] + (h * dir[k]) * scale;
            if (sp[k] < 0 || sp[k] > 1) outside = 1;
        }
        if (outside) break;
        if (warpmap) FN(warp_sample)(warpmap, sp, sp);
        REAL s[4];
        FN(sample_lod)(levels, aniso, dim, nlevels, sp, dir, lod + lod_offset, s);
        REAL aa = (REAL)1 - a;
        for (int k = 0; k < 3; ++k) c[k] = c[k] + s[k] * aa;
        a = a + aa * s[3];
        h = h + r;
    }
    out[0] = c[0];
    out[1] = c[1];
    out[2] = c[2];
    out[3] = a;
}

/* phong.frag:470-478: tangent-space cone table (the fifth entry keeps the
 * reference's -0.5909037 x component) and weights. */
static const float FN(CONE_DIRS)[6][3] = {{0.0f, 1.0f, 0.0f},
                                          {0.0f, 0.5f, 0.866025f},
                                          {0.823639f, 0.5f, 0.267617f},
                                          {0.509037f, 0.5f, -0.700629f},
                                          {-0.5909037f, 0.5f, -0.700629f},
                                          {-0.823639f, 0.5f, 0.267617f}};
static const float FN(CONE_W)[6] = {0.25f, 0.15f, 0.15f, 0.15f, 0.15f, 0.15f};

static void FN(trace_pixel)(const or_gbuf_texel *g, float *const *levels, float *const *const *aniso,
                             const or_trace_params *p, REAL out_ind[4], REAL out_spec[4]) {
    for (int q = 0; q < 4; ++q) out_ind[q] = out_spec[q] = 0;
    if (!g->covered) return;
    REAL pos[3], nraw[3], N[3], T[3], B[3];
    for (int k = 0; k < 3; ++k) {
        pos[k] = (((REAL)g->position[k] - (REAL)p->center[k]) - (REAL)p->vmin[k]) /
                 ((REAL)p->vmax[k] - (REAL)p->vmin[k]);
        nraw[k] = (REAL)g->normal[k];
        T[k] = (REAL)g->tangent[k];
        B[k] = (REAL)g->bitangent[k];
    }
    FN(nrm3)(nraw, N);
    REAL ind[4] = {0, 0, 0, 0};
    REAL tan_d = (REAL)TAN((REAL)p->diffuse.angle / (REAL)2);
    for (int i = 0; i < 6; ++i) {
        const float *cd = FN(CONE_DIRS)[i];
        REAL d[3];
        /* TBN * coneDir with TBN = mat3(T, B, N_interpolated) */
        for (int k = 0; k < 3; ++k) d[k] = (T[k] * (REAL)cd[0] + B[k] * (REAL)cd[1]) + nraw[k] * (REAL)cd[2];
        REAL dn[3];
        FN(nrm3)(d, dn);
        REAL r[4];
        FN(trace_cone)(levels, aniso, p->dim, p->nlevels, pos, N, dn, p->diffuse.steps, (REAL)p->diffuse.bias,
                       tan_d, (REAL)p->diffuse.h0, (REAL)p->diffuse.lod_offset, p->warpmap, r);
        for (int q = 0; q < 4; ++q) ind[q] = ind[q] + (REAL)FN(CONE_W)[i] * r[q];
    }
    REAL ia = ind[3] < 0 ? 0 : (ind[3] > 1 ? 1 : ind[3]);
    REAL occ = (REAL)1 - ia;
    REAL m = p->draw_occlusion ? occ : (REAL)1;
    out_ind[0] = ind[0] * m;
    out_ind[1] = ind[1] * m;
    out_ind[2] = ind[2] * m;
    out_ind[3] = occ;
    if (p->enable_reflections) {
        REAL angle = (REAL)p->specular.angle;
        if (p->angle_from_roughness && p->has_roughness)
            angle = ((REAL)g->roughness * (REAL)3.1415982f) * (REAL)0.1f; /* PI as in common.glsl:1 */
        REAL I[3];
        for (int k = 0; k < 3; ++k) I[k] = (REAL)g->position[k] - (REAL)p->eye[k];
        REAL dni = FN(dot3)(N, I);
        REAL R[3];
        for (int k = 0; k < 3; ++k) R[k] = I[k] - ((REAL)2 * dni) * N[k];
        REAL r[4];
        FN(trace_cone)(levels, aniso, p->dim, p->nlevels, pos, N, R, p->specular.steps, (REAL)p->specular.bias,
                       (REAL)TAN(angle / (REAL)2), (REAL)p->specular.h0, (REAL)p->specular.lod_offset, p->warpmap,
                       r);
        for (int q = 0; q < 4; ++q) out_spec[q] = r[q];
    }
}

#undef FN
#undef CAT
#undef CAT2
