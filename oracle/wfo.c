/*
 * wfo.c -- CPU oracle: restatement of the reference engine's layout tables,
 * Yee/CPML block kernels, Z-slab stepping, sources and probes.
 *
 * TEST INFRASTRUCTURE ONLY (see wfo.h).  Compile with -ffp-contract=off and
 * without -ffast-math: every f32 operation below is a separately rounded
 * IEEE operation in the reference's order.
 */
#define _GNU_SOURCE
#include "wfo.h"

#include <math.h>
#include <stdio.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#define BD WFO_BLOCK_DIM
#define VPB WFO_VOX_PER_BLOCK

/* src/layout.rs:29-31 */
static const float EPS0 = 8.854188e-12f;
static const float MU0 = 1.2566371e-6f;
/* std::f32::consts::PI */
static const float PI_F = 3.14159265358979323846f;

/* ===================================================================== */
/* Q16.16 and materials                                                   */
/* ===================================================================== */

/* Rust `f32 as i32`: saturating, NaN -> 0. */
static int32_t sat_i32(float r)
{
    if (isnan(r)) return 0;
    if (r >= 2147483648.0f) return INT32_MAX;
    if (r <= -2147483648.0f) return INT32_MIN;
    return (int32_t)r;
}

/* Rust `f32 as usize`: saturating, NaN -> 0. */
static size_t sat_usize(float r)
{
    if (isnan(r) || r <= 0.0f) return 0;
    if (r >= 18446744073709551616.0f) return SIZE_MAX;
    return (size_t)r;
}

/* src/layout.rs:114-116: (v * 65536).round() as i32 */
int32_t wfo_q16_from_f32(float v) { return sat_i32(roundf(v * 65536.0f)); }

/* src/layout.rs:123-125: (q as f32) * (1/65536) */
float wfo_q16_to_f32(int32_t q) { return (float)q * (1.0f / 65536.0f); }

/* src/layout.rs:194-208 */
void wfo_material_coeffs(float eps_r, float mu_r, float sigma, float dt, float d,
                         int32_t out[4])
{
    float eps = eps_r * EPS0;
    float mu = mu_r * MU0;
    float loss = sigma * dt / (2.0f * eps);
    float ca = (1.0f - loss) / (1.0f + loss);
    float cb = (dt / (eps * d)) / (1.0f + loss);
    float da = 1.0f;
    float db = dt / (mu * d);
    out[0] = wfo_q16_from_f32(ca);
    out[1] = wfo_q16_from_f32(cb);
    out[2] = wfo_q16_from_f32(da);
    out[3] = wfo_q16_from_f32(db);
}

/* src/layout.rs:525-533 */
size_t wfo_effective_pml_thickness_blocks(size_t nx, size_t ny, size_t nz,
                                          size_t requested_voxels)
{
    if (requested_voxels == 0) return 0;
    size_t bx = nx / BD, by = ny / BD, bz = nz / BD;
    size_t req = requested_voxels / BD + (requested_voxels % BD != 0); /* usize::div_ceil */
    if (req < 1) req = 1;
    size_t mn = bx < by ? bx : by;
    mn = mn < bz ? mn : bz;
    size_t mx = mn / 2;
    if (mx < 1) mx = 1;
    return req < mx ? req : mx;
}

/* src/layout.rs:471-513 (build + coeffs_from_physical) */
void wfo_pml_profile(size_t n, float m, float kappa_max, float alpha_max,
                     float target_reflection, float dt, float dx,
                     size_t thickness, float *b, float *a, float *ik)
{
    for (size_t i = 0; i < n; i++) {
        b[i] = 1.0f;
        a[i] = 0.0f;
        ik[i] = 1.0f;
    }
    if (thickness == 0) return;

    float depth = (float)thickness * dx;
    float eta0 = sqrtf(MU0 / EPS0);
    float sigma_max = -(m + 1.0f) * logf(target_reflection) / (2.0f * eta0 * depth);

    for (size_t g = 0; g < thickness; g++) {
        float rho = (float)(thickness - g) / (float)thickness;
        float sigma = sigma_max * powf(rho, m);
        float kappa = 1.0f + (kappa_max - 1.0f) * powf(rho, m);
        float alpha = alpha_max * (1.0f - rho);

        float bb = expf(-(sigma / kappa + alpha) * dt / EPS0);
        float denom = kappa * (sigma + kappa * alpha);
        float aa = fabsf(denom) > 1.0e-30f ? sigma * (bb - 1.0f) / denom : 0.0f;
        float kk = 1.0f / kappa;

        b[g] = bb;
        a[g] = aa;
        ik[g] = kk;
        b[n - 1 - g] = bb;
        a[n - 1 - g] = aa;
        ik[n - 1 - g] = kk;
    }
}

/* src/source.rs:46-59 */
float wfo_waveform(int kind, float p0, float p1, float t)
{
    switch (kind) {
    case 0: {
        float arg = (t - p0) / p1;
        return expf(-0.5f * arg * arg);
    }
    case 1:
        return sinf(2.0f * PI_F * p0 * t);
    case 2: {
        float arg = PI_F * p0 * (t - p1);
        float sq = arg * arg;
        return (1.0f - 2.0f * sq) * expf(-sq);
    }
    default:
        return 0.0f;
    }
}

/* ===================================================================== */
/* Simulation state                                                       */
/* ===================================================================== */

typedef struct {
    float ex[VPB], ey[VPB], ez[VPB], hx[VPB], hy[VPB], hz[VPB];
} __attribute__((aligned(64))) blk_t;

typedef struct {
    int32_t ca, cb, da, db;
} coef_t;

/* PmlAux field order, src/layout.rs:545-558 */
enum {
    PSI_EZ_DX, PSI_EY_DX, PSI_EZ_DY, PSI_EX_DY, PSI_EY_DZ, PSI_EX_DZ,
    PSI_HZ_DX, PSI_HY_DX, PSI_HX_DY, PSI_HZ_DY, PSI_HY_DZ, PSI_HX_DZ,
    PSI_COUNT
};

typedef struct {
    float p[PSI_COUNT][VPB];
} __attribute__((aligned(64))) aux_t;

typedef struct {
    float b, a, ik;
} pmlc_t;

typedef struct {
    size_t x, y, z;
    int comp, hard, kind;
    float amp, p0, p1;
} src_t;

typedef struct {
    size_t x, y, z;
    int comp;
    size_t nfreq;
    float *freqs;
    float start;
    double *sum_cos, *sum_sin;
    uint64_t n;
} probe_t;

struct wfo_sim {
    size_t nx, ny, nz, nbx, nby, nbz, nblocks;
    blk_t *blocks;
    coef_t *coefs; /* nblocks * VPB, block-major like FieldGrid */
    int pml;
    pmlc_t *prof[3];
    size_t shell[3];
    long *aux_idx;
    aux_t *aux;
    size_t naux;
    unsigned periodic;
    src_t *src;
    size_t nsrc;
    probe_t *probes;
    size_t nprobes;
};

static const blk_t ZERO_BLOCK; /* OUTER_ZERO_BLOCK, src/engine.rs:83 */

static inline float q2f(int32_t q) { return (float)q * (1.0f / 65536.0f); }

static inline size_t lidx(size_t lx, size_t ly, size_t lz) { return (lz * BD + ly) * BD + lx; }

static inline size_t bidx(const wfo_sim *s, size_t bx, size_t by, size_t bz)
{
    return (bz * s->nby + by) * s->nbx + bx;
}

wfo_sim *wfo_sim_create(size_t nx, size_t ny, size_t nz, const uint8_t *ids,
                        const int32_t *table, const size_t shell_blocks[3],
                        const float *profiles, unsigned periodic_mask)
{
    if (nx == 0 || ny == 0 || nz == 0 || nx % BD || ny % BD || nz % BD) return NULL;
    wfo_sim *s = calloc(1, sizeof(*s));
    if (!s) return NULL;
    s->nx = nx;
    s->ny = ny;
    s->nz = nz;
    s->nbx = nx / BD;
    s->nby = ny / BD;
    s->nbz = nz / BD;
    s->nblocks = s->nbx * s->nby * s->nbz;
    s->periodic = periodic_mask;

    s->blocks = aligned_alloc(64, s->nblocks * sizeof(blk_t));
    s->coefs = aligned_alloc(64, s->nblocks * VPB * sizeof(coef_t));
    s->aux_idx = malloc(s->nblocks * sizeof(long));
    if (!s->blocks || !s->coefs || !s->aux_idx) {
        wfo_sim_destroy(s);
        return NULL;
    }
    memset(s->blocks, 0, s->nblocks * sizeof(blk_t));

    /* CoeffGrid::build, src/layout.rs:230-260 */
    for (size_t bz = 0; bz < s->nbz; bz++)
        for (size_t by = 0; by < s->nby; by++)
            for (size_t bx = 0; bx < s->nbx; bx++) {
                coef_t *bc = s->coefs + bidx(s, bx, by, bz) * VPB;
                for (size_t lz = 0; lz < BD; lz++)
                    for (size_t ly = 0; ly < BD; ly++)
                        for (size_t lx = 0; lx < BD; lx++) {
                            size_t x = bx * BD + lx, y = by * BD + ly, z = bz * BD + lz;
                            unsigned id = ids ? ids[(z * ny + y) * nx + x] : 0u;
                            const int32_t *t = table + 4 * id;
                            coef_t c = {t[0], t[1], t[2], t[3]};
                            bc[lidx(lx, ly, lz)] = c;
                        }
            }

    /* PmlContext::build + PmlAuxGrid::build, src/layout.rs:594-656 */
    s->pml = profiles != NULL;
    for (int ax = 0; ax < 3; ax++) s->shell[ax] = shell_blocks ? shell_blocks[ax] : 0;
    if (s->pml) {
        size_t n[3] = {nx, ny, nz};
        const float *p = profiles;
        for (int ax = 0; ax < 3; ax++) {
            s->prof[ax] = malloc(n[ax] * sizeof(pmlc_t));
            if (!s->prof[ax]) {
                wfo_sim_destroy(s);
                return NULL;
            }
            for (size_t i = 0; i < n[ax]; i++) {
                s->prof[ax][i].b = p[i];
                s->prof[ax][i].a = p[n[ax] + i];
                s->prof[ax][i].ik = p[2 * n[ax] + i];
            }
            p += 3 * n[ax];
        }
    }
    size_t nb[3] = {s->nbx, s->nby, s->nbz};
    s->naux = 0;
    for (size_t bz = 0; bz < s->nbz; bz++)
        for (size_t by = 0; by < s->nby; by++)
            for (size_t bx = 0; bx < s->nbx; bx++) {
                size_t c[3] = {bx, by, bz};
                int in_shell = 0;
                for (int ax = 0; ax < 3; ax++) {
                    size_t t = s->shell[ax];
                    if (t > 0 && (c[ax] < t || c[ax] >= nb[ax] - t)) in_shell = 1;
                }
                s->aux_idx[bidx(s, bx, by, bz)] = in_shell ? (long)s->naux++ : -1;
            }
    if (s->naux) {
        s->aux = aligned_alloc(64, s->naux * sizeof(aux_t));
        if (!s->aux) {
            wfo_sim_destroy(s);
            return NULL;
        }
        memset(s->aux, 0, s->naux * sizeof(aux_t));
    }
    return s;
}

void wfo_sim_destroy(wfo_sim *s)
{
    if (!s) return;
    free(s->blocks);
    free(s->coefs);
    free(s->aux_idx);
    free(s->aux);
    for (int ax = 0; ax < 3; ax++) free(s->prof[ax]);
    free(s->src);
    for (size_t p = 0; p < s->nprobes; p++) {
        free(s->probes[p].freqs);
        free(s->probes[p].sum_cos);
        free(s->probes[p].sum_sin);
    }
    free(s->probes);
    free(s);
}

int wfo_sim_add_source(wfo_sim *s, size_t x, size_t y, size_t z, int comp, int hard,
                       float amplitude, int kind, float p0, float p1)
{
    if (x >= s->nx || y >= s->ny || z >= s->nz || comp < 0 || comp > 2) return -1;
    src_t *n = realloc(s->src, (s->nsrc + 1) * sizeof(src_t));
    if (!n) return -1;
    s->src = n;
    src_t v = {x, y, z, comp, hard, kind, amplitude, p0, p1};
    s->src[s->nsrc++] = v;
    return 0;
}

int wfo_sim_add_probe(wfo_sim *s, size_t x, size_t y, size_t z, int comp, size_t nfreq,
                      const float *freqs, float start)
{
    if (x >= s->nx || y >= s->ny || z >= s->nz || comp < 0 || comp > 5) return -1;
    probe_t *n = realloc(s->probes, (s->nprobes + 1) * sizeof(probe_t));
    if (!n) return -1;
    s->probes = n;
    probe_t *p = &s->probes[s->nprobes++];
    memset(p, 0, sizeof(*p));
    p->x = x;
    p->y = y;
    p->z = z;
    p->comp = comp;
    p->nfreq = nfreq;
    p->start = start;
    p->freqs = calloc(nfreq ? nfreq : 1, sizeof(float));
    p->sum_cos = calloc(nfreq ? nfreq : 1, sizeof(double));
    p->sum_sin = calloc(nfreq ? nfreq : 1, sizeof(double));
    if (nfreq) memcpy(p->freqs, freqs, nfreq * sizeof(float));
    return 0;
}

/* ===================================================================== */
/* Kernels: src/fdtd.rs                                                   */
/* ===================================================================== */

/* pml_correct_x / pml_correct_scalar, src/fdtd.rs:136-163 */
static inline float pml_correct(float raw, pmlc_t c, float *psi)
{
    float psi_new = c.b * *psi + c.a * raw;
    *psi = psi_new;
    return raw * c.ik + psi_new;
}

/* update_h_field, src/fdtd.rs:194-253 */
static void update_h(blk_t *c, const blk_t *px, const blk_t *py, const blk_t *pz,
                     const coef_t *co)
{
    for (size_t lz = 0; lz < BD; lz++)
        for (size_t ly = 0; ly < BD; ly++) {
            size_t rb = lidx(0, ly, lz);
            const float *ez_py, *ex_py, *ey_pz, *ex_pz;
            if (ly + 1 < BD) {
                ez_py = &c->ez[lidx(0, ly + 1, lz)];
                ex_py = &c->ex[lidx(0, ly + 1, lz)];
            } else {
                ez_py = &py->ez[lidx(0, 0, lz)];
                ex_py = &py->ex[lidx(0, 0, lz)];
            }
            if (lz + 1 < BD) {
                ey_pz = &c->ey[lidx(0, ly, lz + 1)];
                ex_pz = &c->ex[lidx(0, ly, lz + 1)];
            } else {
                ey_pz = &pz->ey[lidx(0, ly, 0)];
                ex_pz = &pz->ex[lidx(0, ly, 0)];
            }
            /* shifted_row_plus, src/fdtd.rs:40-45 */
            float ez_px[BD], ey_px[BD];
            for (size_t l = 0; l + 1 < BD; l++) {
                ez_px[l] = c->ez[rb + l + 1];
                ey_px[l] = c->ey[rb + l + 1];
            }
            ez_px[BD - 1] = px->ez[rb];
            ey_px[BD - 1] = px->ey[rb];

            for (size_t l = 0; l < BD; l++) {
                size_t i = rb + l;
                float ex = c->ex[i], ey = c->ey[i], ez = c->ez[i];
                float chx = (ez_py[l] - ez) - (ey_pz[l] - ey);
                float chy = (ex_pz[l] - ex) - (ez_px[l] - ez);
                float chz = (ey_px[l] - ey) - (ex_py[l] - ex);
                float da = q2f(co[i].da), db = q2f(co[i].db);
                float hx = c->hx[i], hy = c->hy[i], hz = c->hz[i];
                c->hx[i] = da * hx - db * chx;
                c->hy[i] = da * hy - db * chy;
                c->hz[i] = da * hz - db * chz;
            }
        }
}

/* update_h_field_pml, src/fdtd.rs:273-345 */
static void update_h_pml(blk_t *c, const blk_t *px, const blk_t *py, const blk_t *pz,
                         const coef_t *co, const pmlc_t *wx, const pmlc_t *wy,
                         const pmlc_t *wz, aux_t *aux)
{
    for (size_t lz = 0; lz < BD; lz++) {
        pmlc_t pzc = wz[lz];
        for (size_t ly = 0; ly < BD; ly++) {
            pmlc_t pyc = wy[ly];
            size_t rb = lidx(0, ly, lz);
            const float *ez_py, *ex_py, *ey_pz, *ex_pz;
            if (ly + 1 < BD) {
                ez_py = &c->ez[lidx(0, ly + 1, lz)];
                ex_py = &c->ex[lidx(0, ly + 1, lz)];
            } else {
                ez_py = &py->ez[lidx(0, 0, lz)];
                ex_py = &py->ex[lidx(0, 0, lz)];
            }
            if (lz + 1 < BD) {
                ey_pz = &c->ey[lidx(0, ly, lz + 1)];
                ex_pz = &c->ex[lidx(0, ly, lz + 1)];
            } else {
                ey_pz = &pz->ey[lidx(0, ly, 0)];
                ex_pz = &pz->ex[lidx(0, ly, 0)];
            }
            float ez_px[BD], ey_px[BD];
            for (size_t l = 0; l + 1 < BD; l++) {
                ez_px[l] = c->ez[rb + l + 1];
                ey_px[l] = c->ey[rb + l + 1];
            }
            ez_px[BD - 1] = px->ez[rb];
            ey_px[BD - 1] = px->ey[rb];

            for (size_t l = 0; l < BD; l++) {
                size_t i = rb + l;
                float ex = c->ex[i], ey = c->ey[i], ez = c->ez[i];
                float d_ez_dy = pml_correct(ez_py[l] - ez, pyc, &aux->p[PSI_EZ_DY][i]);
                float d_ex_dy = pml_correct(ex_py[l] - ex, pyc, &aux->p[PSI_EX_DY][i]);
                float d_ey_dz = pml_correct(ey_pz[l] - ey, pzc, &aux->p[PSI_EY_DZ][i]);
                float d_ex_dz = pml_correct(ex_pz[l] - ex, pzc, &aux->p[PSI_EX_DZ][i]);
                float d_ez_dx = pml_correct(ez_px[l] - ez, wx[l], &aux->p[PSI_EZ_DX][i]);
                float d_ey_dx = pml_correct(ey_px[l] - ey, wx[l], &aux->p[PSI_EY_DX][i]);
                float chx = d_ez_dy - d_ey_dz;
                float chy = d_ex_dz - d_ez_dx;
                float chz = d_ey_dx - d_ex_dy;
                float da = q2f(co[i].da), db = q2f(co[i].db);
                float hx = c->hx[i], hy = c->hy[i], hz = c->hz[i];
                c->hx[i] = da * hx - db * chx;
                c->hy[i] = da * hy - db * chy;
                c->hz[i] = da * hz - db * chz;
            }
        }
    }
}

/* update_e_field, src/fdtd.rs:369-428 */
static void update_e(blk_t *c, const blk_t *mx, const blk_t *my, const blk_t *mz,
                     const coef_t *co)
{
    for (size_t lz = 0; lz < BD; lz++)
        for (size_t ly = 0; ly < BD; ly++) {
            size_t rb = lidx(0, ly, lz);
            const float *hz_my, *hx_my, *hy_mz, *hx_mz;
            if (ly > 0) {
                hz_my = &c->hz[lidx(0, ly - 1, lz)];
                hx_my = &c->hx[lidx(0, ly - 1, lz)];
            } else {
                hz_my = &my->hz[lidx(0, BD - 1, lz)];
                hx_my = &my->hx[lidx(0, BD - 1, lz)];
            }
            if (lz > 0) {
                hy_mz = &c->hy[lidx(0, ly, lz - 1)];
                hx_mz = &c->hx[lidx(0, ly, lz - 1)];
            } else {
                hy_mz = &mz->hy[lidx(0, ly, BD - 1)];
                hx_mz = &mz->hx[lidx(0, ly, BD - 1)];
            }
            /* shifted_row_minus, src/fdtd.rs:52-57 */
            float hz_mx[BD], hy_mx[BD];
            hz_mx[0] = mx->hz[lidx(BD - 1, ly, lz)];
            hy_mx[0] = mx->hy[lidx(BD - 1, ly, lz)];
            for (size_t l = 1; l < BD; l++) {
                hz_mx[l] = c->hz[rb + l - 1];
                hy_mx[l] = c->hy[rb + l - 1];
            }

            for (size_t l = 0; l < BD; l++) {
                size_t i = rb + l;
                float hx = c->hx[i], hy = c->hy[i], hz = c->hz[i];
                float cex = (hz - hz_my[l]) - (hy - hy_mz[l]);
                float cey = (hx - hx_mz[l]) - (hz - hz_mx[l]);
                float cez = (hy - hy_mx[l]) - (hx - hx_my[l]);
                float ca = q2f(co[i].ca), cb = q2f(co[i].cb);
                float ex = c->ex[i], ey = c->ey[i], ez = c->ez[i];
                c->ex[i] = ca * ex + cb * cex;
                c->ey[i] = ca * ey + cb * cey;
                c->ez[i] = ca * ez + cb * cez;
            }
        }
}

/* update_e_field_pml, src/fdtd.rs:437-509 */
static void update_e_pml(blk_t *c, const blk_t *mx, const blk_t *my, const blk_t *mz,
                         const coef_t *co, const pmlc_t *wx, const pmlc_t *wy,
                         const pmlc_t *wz, aux_t *aux)
{
    for (size_t lz = 0; lz < BD; lz++) {
        pmlc_t pzc = wz[lz];
        for (size_t ly = 0; ly < BD; ly++) {
            pmlc_t pyc = wy[ly];
            size_t rb = lidx(0, ly, lz);
            const float *hz_my, *hx_my, *hy_mz, *hx_mz;
            if (ly > 0) {
                hz_my = &c->hz[lidx(0, ly - 1, lz)];
                hx_my = &c->hx[lidx(0, ly - 1, lz)];
            } else {
                hz_my = &my->hz[lidx(0, BD - 1, lz)];
                hx_my = &my->hx[lidx(0, BD - 1, lz)];
            }
            if (lz > 0) {
                hy_mz = &c->hy[lidx(0, ly, lz - 1)];
                hx_mz = &c->hx[lidx(0, ly, lz - 1)];
            } else {
                hy_mz = &mz->hy[lidx(0, ly, BD - 1)];
                hx_mz = &mz->hx[lidx(0, ly, BD - 1)];
            }
            float hz_mx[BD], hy_mx[BD];
            hz_mx[0] = mx->hz[lidx(BD - 1, ly, lz)];
            hy_mx[0] = mx->hy[lidx(BD - 1, ly, lz)];
            for (size_t l = 1; l < BD; l++) {
                hz_mx[l] = c->hz[rb + l - 1];
                hy_mx[l] = c->hy[rb + l - 1];
            }

            for (size_t l = 0; l < BD; l++) {
                size_t i = rb + l;
                float hx = c->hx[i], hy = c->hy[i], hz = c->hz[i];
                float d_hz_dy = pml_correct(hz - hz_my[l], pyc, &aux->p[PSI_HZ_DY][i]);
                float d_hx_dy = pml_correct(hx - hx_my[l], pyc, &aux->p[PSI_HX_DY][i]);
                float d_hy_dz = pml_correct(hy - hy_mz[l], pzc, &aux->p[PSI_HY_DZ][i]);
                float d_hx_dz = pml_correct(hx - hx_mz[l], pzc, &aux->p[PSI_HX_DZ][i]);
                float d_hz_dx = pml_correct(hz - hz_mx[l], wx[l], &aux->p[PSI_HZ_DX][i]);
                float d_hy_dx = pml_correct(hy - hy_mx[l], wx[l], &aux->p[PSI_HY_DX][i]);
                float cex = d_hz_dy - d_hy_dz;
                float cey = d_hx_dz - d_hz_dx;
                float cez = d_hy_dx - d_hx_dy;
                float ca = q2f(co[i].ca), cb = q2f(co[i].cb);
                float ex = c->ex[i], ey = c->ey[i], ez = c->ez[i];
                c->ex[i] = ca * ex + cb * cex;
                c->ey[i] = ca * ey + cb * cey;
                c->ez[i] = ca * ez + cb * cez;
            }
        }
    }
}

/* ===================================================================== */
/* Slab traversal: src/engine.rs:130-275                                  */
/* ===================================================================== */

static const blk_t *nbr_plus(const wfo_sim *s, size_t bx, size_t by, size_t bz, int ax)
{
    size_t c[3] = {bx, by, bz}, nb[3] = {s->nbx, s->nby, s->nbz};
    if (c[ax] + 1 < nb[ax]) {
        c[ax] += 1;
    } else if (s->periodic & (1u << ax)) {
        c[ax] = 0;
    } else {
        return &ZERO_BLOCK;
    }
    return &s->blocks[bidx(s, c[0], c[1], c[2])];
}

static const blk_t *nbr_minus(const wfo_sim *s, size_t bx, size_t by, size_t bz, int ax)
{
    size_t c[3] = {bx, by, bz}, nb[3] = {s->nbx, s->nby, s->nbz};
    if (c[ax] > 0) {
        c[ax] -= 1;
    } else if (s->periodic & (1u << ax)) {
        c[ax] = nb[ax] - 1;
    } else {
        return &ZERO_BLOCK;
    }
    return &s->blocks[bidx(s, c[0], c[1], c[2])];
}

static void slab_h(wfo_sim *s, size_t bz0, size_t bz1)
{
    for (size_t bz = bz0; bz < bz1; bz++)
        for (size_t by = 0; by < s->nby; by++)
            for (size_t bx = 0; bx < s->nbx; bx++) {
                size_t idx = bidx(s, bx, by, bz);
                blk_t *c = &s->blocks[idx];
                const blk_t *px = nbr_plus(s, bx, by, bz, 0);
                const blk_t *py = nbr_plus(s, bx, by, bz, 1);
                const blk_t *pz = nbr_plus(s, bx, by, bz, 2);
                const coef_t *co = s->coefs + idx * VPB;
                long a = s->aux_idx[idx];
                if (s->pml && a >= 0) {
                    update_h_pml(c, px, py, pz, co, s->prof[0] + bx * BD,
                                 s->prof[1] + by * BD, s->prof[2] + bz * BD, &s->aux[a]);
                } else {
                    update_h(c, px, py, pz, co);
                }
            }
}

static void slab_e(wfo_sim *s, size_t bz0, size_t bz1)
{
    for (size_t bz = bz0; bz < bz1; bz++)
        for (size_t by = 0; by < s->nby; by++)
            for (size_t bx = 0; bx < s->nbx; bx++) {
                size_t idx = bidx(s, bx, by, bz);
                blk_t *c = &s->blocks[idx];
                const blk_t *mx = nbr_minus(s, bx, by, bz, 0);
                const blk_t *my = nbr_minus(s, bx, by, bz, 1);
                const blk_t *mz = nbr_minus(s, bx, by, bz, 2);
                const coef_t *co = s->coefs + idx * VPB;
                long a = s->aux_idx[idx];
                if (s->pml && a >= 0) {
                    update_e_pml(c, mx, my, mz, co, s->prof[0] + bx * BD,
                                 s->prof[1] + by * BD, s->prof[2] + bz * BD, &s->aux[a]);
                } else {
                    update_e(c, mx, my, mz, co);
                }
            }
}

/* ===================================================================== */
/* Sources and probes: src/source.rs:79-95, src/probe.rs:101-138          */
/* ===================================================================== */

static float *comp_ptr(wfo_sim *s, size_t x, size_t y, size_t z, int comp)
{
    blk_t *b = &s->blocks[bidx(s, x / BD, y / BD, z / BD)];
    size_t l = lidx(x % BD, y % BD, z % BD);
    float *arr[6] = {b->ex, b->ey, b->ez, b->hx, b->hy, b->hz};
    return &arr[comp][l];
}

static void inject_sources(wfo_sim *s, float t)
{
    for (size_t i = 0; i < s->nsrc; i++) {
        const src_t *src = &s->src[i];
        float value = src->amp * wfo_waveform(src->kind, src->p0, src->p1, t);
        float *f = comp_ptr(s, src->x, src->y, src->z, src->comp);
        if (src->hard)
            *f = value;
        else
            *f += value;
    }
}

static void accumulate_probes(wfo_sim *s, float t, float *trace_row)
{
    for (size_t p = 0; p < s->nprobes; p++) {
        probe_t *pr = &s->probes[p];
        float v32 = *comp_ptr(s, pr->x, pr->y, pr->z, pr->comp);
        if (trace_row) trace_row[p] = v32;
        if (t < pr->start) continue;
        double value = (double)v32;
        for (size_t i = 0; i < pr->nfreq; i++) {
            double omega = 2.0 * M_PI * (double)pr->freqs[i];
            double phase = omega * (double)t;
            pr->sum_cos[i] += value * cos(phase);
            pr->sum_sin[i] += value * sin(phase);
        }
        pr->n += 1;
    }
}

/* Probe::accumulate(&grid, t) for every probe, on the fields as they are
 * (src/probe.rs:126-138): the probes' known-answer tests drive a voxel by
 * writing the fields and call this once per step. */
void wfo_sim_accumulate_probes(wfo_sim *s, float t) { accumulate_probes(s, t, NULL); }

/* ===================================================================== */
/* SPMD runner: Z-slabs (src/engine.rs:666-669) with a barrier per phase   */
/* ===================================================================== */

typedef struct {
    wfo_sim *s;
    size_t first, nsteps;
    float dt;
    float *trace;
    int nthreads;
    size_t rows_per_slab;
    pthread_barrier_t *bar;
    int fields_only;
} run_ctx;

typedef struct {
    run_ctx *rc;
    int id;
} worker_arg;

static void *worker(void *p)
{
    worker_arg *wa = p;
    run_ctx *rc = wa->rc;
    wfo_sim *s = rc->s;
    size_t bz0 = (size_t)wa->id * rc->rows_per_slab;
    size_t bz1 = bz0 + rc->rows_per_slab;
    if (bz0 > s->nbz) bz0 = s->nbz;
    if (bz1 > s->nbz) bz1 = s->nbz;
    for (size_t k = 0; k < rc->nsteps; k++) {
        size_t step = rc->first + k;
        slab_h(s, bz0, bz1);
        if (rc->nthreads > 1) pthread_barrier_wait(rc->bar);
        slab_e(s, bz0, bz1);
        if (rc->nthreads > 1) pthread_barrier_wait(rc->bar);
        if (wa->id == 0 && !rc->fields_only) {
            float t = ((float)step + 1.0f) * rc->dt;
            inject_sources(s, t);
            accumulate_probes(s, t, rc->trace ? rc->trace + k * s->nprobes : NULL);
        }
        if (rc->nthreads > 1) pthread_barrier_wait(rc->bar);
    }
    return NULL;
}

static void run_impl(wfo_sim *s, size_t first, size_t nsteps, float dt, int threads,
                     float *trace, int fields_only)
{
    if (threads < 1) threads = 1;
    size_t rows = (s->nbz + (size_t)threads - 1) / (size_t)threads;
    if (rows < 1) rows = 1;
    size_t nslabs = (s->nbz + rows - 1) / rows;
    run_ctx rc = {s, first, nsteps, dt, trace, (int)nslabs, rows, NULL, fields_only};
    if (nslabs <= 1) {
        worker_arg wa = {&rc, 0};
        worker(&wa);
        return;
    }
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)nslabs);
    rc.bar = &bar;
    pthread_t *th = malloc(nslabs * sizeof(pthread_t));
    worker_arg *wa = malloc(nslabs * sizeof(worker_arg));
    for (size_t i = 0; i < nslabs; i++) {
        wa[i].rc = &rc;
        wa[i].id = (int)i;
        if (i > 0) pthread_create(&th[i], NULL, worker, &wa[i]);
    }
    worker(&wa[0]);
    for (size_t i = 1; i < nslabs; i++) pthread_join(th[i], NULL);
    pthread_barrier_destroy(&bar);
    free(th);
    free(wa);
}

void wfo_sim_run(wfo_sim *s, size_t first_step, size_t nsteps, float dt, int threads,
                 float *trace)
{
    run_impl(s, first_step, nsteps, dt, threads, trace, 0);
}

void wfo_sim_step_fields(wfo_sim *s, int threads) { run_impl(s, 0, 1, 0.0f, threads, NULL, 1); }

size_t wfo_sim_bytes(const wfo_sim *s) { return s->nblocks * sizeof(blk_t); }

/* serialize_snapshot, src/engine.rs:622-640 */
void wfo_sim_read_aosoa(const wfo_sim *s, void *dst)
{
    memcpy(dst, s->blocks, s->nblocks * sizeof(blk_t));
}

void wfo_sim_write_aosoa(wfo_sim *s, const void *src)
{
    memcpy(s->blocks, src, s->nblocks * sizeof(blk_t));
}

size_t wfo_sim_shell_block_count(const wfo_sim *s) { return s->naux; }

void wfo_sim_read_psi(const wfo_sim *s, float *dst)
{
    if (s->naux) memcpy(dst, s->aux, s->naux * sizeof(aux_t));
}

size_t wfo_sim_probe_count(const wfo_sim *s) { return s->nprobes; }

size_t wfo_sim_probe_nfreq(const wfo_sim *s, size_t p)
{
    return p < s->nprobes ? s->probes[p].nfreq : 0;
}

void wfo_sim_probe_sums(const wfo_sim *s, size_t p, double *sum_cos, double *sum_sin,
                        uint64_t *n)
{
    const probe_t *pr = &s->probes[p];
    memcpy(sum_cos, pr->sum_cos, pr->nfreq * sizeof(double));
    memcpy(sum_sin, pr->sum_sin, pr->nfreq * sizeof(double));
    *n = pr->n;
}

/* src/probe.rs:143-167 */
void wfo_probe_spectrum(size_t nfreq, const double *sum_cos, const double *sum_sin,
                        uint64_t n, double *amplitude, double *phase)
{
    double nn = (double)(n > 0 ? n : 1);
    for (size_t i = 0; i < nfreq; i++) {
        double a = 2.0 * sum_cos[i] / nn;
        double b = 2.0 * sum_sin[i] / nn;
        amplitude[i] = sqrt(a * a + b * b);
        phase[i] = atan2(a, b);
    }
}

/* ===================================================================== */
/* Scenes: src/scene.rs:81-230, src/main.rs:403-425                        */
/* ===================================================================== */

void wfo_voxelize_demo_sphere(size_t nx, size_t ny, size_t nz, uint8_t *ids)
{
    float cx = (float)nx / 2.0f, cy = (float)ny / 2.0f, cz = (float)nz / 2.0f;
    size_t mn = nx < ny ? nx : ny;
    mn = mn < nz ? mn : nz;
    float radius = (float)mn / 4.0f;
    float r2 = radius * radius;
    for (size_t z = 0; z < nz; z++)
        for (size_t y = 0; y < ny; y++)
            for (size_t x = 0; x < nx; x++) {
                float dx = (float)x + 0.5f - cx;
                float dy = (float)y + 0.5f - cy;
                float dz = (float)z + 0.5f - cz;
                if (dx * dx + dy * dy + dz * dz <= r2) ids[(z * ny + y) * nx + x] = 1;
            }
}

/* Voxels [lo, hi) along one axis whose own term can pass the sphere
 * predicate, fl(d * d) <= r2 with d = fl(fl(v + 0.5) - c): the predicate of
 * src/scene.rs:197-209 compares a rounded sum of three non-negative rounded
 * squares, which is never below any of them, so a voxel failing on one axis
 * fails whatever the others are.  Every v of the axis is tested, so the box
 * holds every voxel the reference's full scan sets, for any centre and
 * radius. */
static void sphere_range(float c, float r2, size_t n, size_t *lo, size_t *hi)
{
    size_t a = n, b = n;
    for (size_t v = 0; v < n; v++) {
        float d = (float)v + 0.5f - c;
        if (d * d <= r2) {
            if (a == n) a = v;
            b = v + 1;
        }
    }
    *lo = a == n ? 0 : a;
    *hi = a == n ? 0 : b;
}

typedef struct {
    int is_box;
    float v[9];
    int nv;
} prim_t;

/* src/scene.rs:96 parses each number with Rust's f32 FromStr, whose token
 * grammar is [+-]? (inf | infinity | nan | decimal), letters in any case, the
 * decimal D+ | D+ '.' D* | D* '.' D+ with an optional exponent [eE][+-]?D+.
 * strtof takes more (hex floats, nan(...)); the reference rejects those. */
static int rust_f32_token(const char *s)
{
    if (*s == '+' || *s == '-') s++;
    if (!strcasecmp(s, "inf") || !strcasecmp(s, "infinity") || !strcasecmp(s, "nan")) return 1;
    int nd = 0;
    while (*s >= '0' && *s <= '9') s++, nd++;
    if (*s == '.')
        for (s++; *s >= '0' && *s <= '9'; s++) nd++;
    if (!nd) return 0;
    if (*s == 'e' || *s == 'E') {
        s++;
        if (*s == '+' || *s == '-') s++;
        if (!(*s >= '0' && *s <= '9')) return 0;
        while (*s >= '0' && *s <= '9') s++;
    }
    return *s == 0;
}

static int parse_line(const char *line, size_t len, prim_t *out, int *has, char *msg,
                      size_t msg_len, size_t line_no)
{
    /* split_whitespace */
    char buf[1024];
    if (len >= sizeof(buf)) len = sizeof(buf) - 1;
    memcpy(buf, line, len);
    buf[len] = 0;
    char *save = NULL;
    char *tok = strtok_r(buf, " \t\r\v\f", &save);
    *has = 0;
    if (!tok || tok[0] == '#') return 0;
    int is_sphere = strcmp(tok, "sphere") == 0, is_box = strcmp(tok, "box") == 0;
    if (!is_sphere && !is_box) {
        snprintf(msg, msg_len, "scene line %zu: unknown primitive '%s'", line_no, tok);
        return -1;
    }
    int nv = 0;
    float v[16];
    while ((tok = strtok_r(NULL, " \t\r\v\f", &save)) != NULL) {
        char *end = NULL;
        float f = strtof(tok, &end);
        if (!rust_f32_token(tok) || end == tok || *end != 0) {
            snprintf(msg, msg_len, "scene line %zu: invalid number '%s'", line_no, tok);
            return -1;
        }
        if (nv < 16) v[nv] = f;
        nv++;
    }
    int want = is_sphere ? 7 : 9;
    if (nv != want) {
        snprintf(msg, msg_len, "scene line %zu: expected %d values, got %d", line_no, want,
                 nv);
        return -1;
    }
    out->is_box = is_box;
    out->nv = nv;
    memcpy(out->v, v, (size_t)nv * sizeof(float));
    *has = 1;
    return 0;
}

int wfo_scene_voxelize(const char *text, size_t nx, size_t ny, size_t nz, float dt,
                       float dx, uint8_t *ids, int32_t *table, char *msg, size_t msg_len)
{
    uint32_t keys[256][3];
    int nkeys = 0;
    size_t line_no = 0;
    const char *p = text;
    /* parse everything first (Scene::parse), then voxelize (Scene::voxelize) */
    prim_t *prims = NULL;
    size_t nprims = 0;
    while (*p) {
        const char *e = strchr(p, '\n');
        size_t len = e ? (size_t)(e - p) : strlen(p);
        line_no++;
        prim_t pr;
        int has = 0;
        if (parse_line(p, len, &pr, &has, msg, msg_len, line_no) != 0) {
            free(prims);
            return -1;
        }
        if (has) {
            prim_t *np = realloc(prims, (nprims + 1) * sizeof(prim_t));
            if (!np) {
                free(prims);
                return -1;
            }
            prims = np;
            prims[nprims++] = pr;
        }
        if (!e) break;
        p = e + 1;
    }
    for (size_t k = 0; k < nprims; k++) {
        const prim_t *pr = &prims[k];
        uint32_t key[3];
        memcpy(key, pr->v, 3 * sizeof(float));
        int id = -1;
        for (int i = 0; i < nkeys; i++)
            if (keys[i][0] == key[0] && keys[i][1] == key[1] && keys[i][2] == key[2]) id = i + 1;
        if (id < 0) {
            if (nkeys + 1 > 255) {
                snprintf(msg, msg_len, "scene uses more than 255 distinct materials");
                free(prims);
                return -1;
            }
            memcpy(keys[nkeys], key, sizeof(key));
            nkeys++;
            id = nkeys;
            wfo_material_coeffs(pr->v[0], pr->v[1], pr->v[2], dt, dx, table + 4 * id);
        }
        if (!pr->is_box) {
            float cx = pr->v[3], cy = pr->v[4], cz = pr->v[5], r = pr->v[6];
            float r2 = r * r;
            size_t x0, x1, y0, y1, z0, z1;
            sphere_range(cx, r2, nx, &x0, &x1);
            sphere_range(cy, r2, ny, &y0, &y1);
            sphere_range(cz, r2, nz, &z0, &z1);
            for (size_t z = z0; z < z1; z++)
                for (size_t y = y0; y < y1; y++)
                    for (size_t x = x0; x < x1; x++) {
                        float ddx = (float)x + 0.5f - cx;
                        float ddy = (float)y + 0.5f - cy;
                        float ddz = (float)z + 0.5f - cz;
                        if (ddx * ddx + ddy * ddy + ddz * ddz <= r2)
                            ids[(z * ny + y) * nx + x] = (uint8_t)id;
                    }
        } else {
            const float *v = pr->v + 3;
            size_t lo[3], hi[3], n[3] = {nx, ny, nz};
            for (int ax = 0; ax < 3; ax++) {
                float a = v[ax], b = v[3 + ax];
                lo[ax] = sat_usize(floorf(fmaxf(fminf(a, b), 0.0f)));
                size_t h = sat_usize(fmaxf(ceilf(fmaxf(a, b)), 0.0f));
                hi[ax] = h < n[ax] ? h : n[ax];
            }
            for (size_t z = lo[2]; z < hi[2]; z++)
                for (size_t y = lo[1]; y < hi[1]; y++)
                    for (size_t x = lo[0]; x < hi[0]; x++) ids[(z * ny + y) * nx + x] = (uint8_t)id;
        }
    }
    free(prims);
    return nkeys;
}
