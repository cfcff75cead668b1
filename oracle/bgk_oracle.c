/*
 * bgk_oracle.c -- CPU restatement of BGK2D's time-step loop (scalar C).
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library; the product never does.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile).
 * -ffp-contract=off keeps every multiply and add separately rounded, as the
 * reference's x86-64 build does (no FMA in the x86-64 baseline ISA).
 *
 * The per-precision routines live in orc_body.inc, included once for each
 * (storage kind, compute kind) pair of SRC/precision.F90:13-47.
 */
#define _GNU_SOURCE
#include "bgk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_vtab orc_vtab;

struct orc {
    orc_params p;
    int l, m, ld;
    size_t plane;
    int esize;
    void *a[9], *b[9];
    int32_t *obs;
    int imin, imax, jmin, jmax, nobs;
    int stepx, stepy;
    double omega, fgrad, u00, force;
    double umax;  /* OPEN inflow peak in the compute kind */
    double scale, inv_scale;
    const orc_vtab *vt;
};

struct orc_vtab {
    void (*setup)(orc *);
    void (*boundaries)(orc *);
    void (*movef)(orc *);
    void (*col)(orc *);
    void (*colmc)(orc *);
    void (*diagno)(const orc *, int, double *);
    void (*varm)(const orc *, double *, double *, double *);
    void (*prof_i)(const orc *, int, double *, double *, double *);
    void (*prof_j)(const orc *, int, double *, double *, double *);
    void (*probe)(const orc *, int, int, double *);
    void (*decode_plane)(const orc *, const void *, double *);
    void (*draglift)(const orc *, int, int, int, int, double *);
};

/* ---------------------------------------------------- storage rounding */

/* IEEE binary16 from binary32, round to nearest even (own implementation
 * of the IEEE 754 rounding rule; checked against numpy in the tests).   */
uint16_t orc_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) /* inf / nan */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u));
    if (ax >= 0x477ff000u) /* >= 65520 rounds to inf */
        return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) { /* below 2^-14: subnormal half (or zero) */
        float a;
        memcpy(&a, &ax, 4);
        return (uint16_t)(sign | (uint32_t)lrintf(a * 16777216.0f));
    }
    uint32_t e = (ax >> 23) - 127u + 15u;
    uint32_t mant = ax & 0x7fffffu;
    uint32_t hb = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (hb & 1u))) hb++;
    return (uint16_t)(sign | hb);
}

float orc_f16_to_f32(uint16_t hv)
{
    uint32_t sign = ((uint32_t)hv & 0x8000u) << 16;
    uint32_t e = (hv >> 10) & 0x1fu, mant = hv & 0x3ffu;
    float r;
    if (e == 0) {
        r = (float)mant * (1.0f / 16777216.0f);
        uint32_t bits;
        memcpy(&bits, &r, 4);
        bits |= sign;
        memcpy(&r, &bits, 4);
        return r;
    }
    uint32_t bits;
    if (e == 31)
        bits = sign | 0x7f800000u | (mant << 13);
    else
        bits = sign | ((e - 15u + 127u) << 23) | (mant << 13);
    memcpy(&r, &bits, 4);
    return r;
}

/* bfloat16 from binary32, round to nearest even */
uint16_t orc_f32_to_bf16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}

float orc_bf16_to_f32(uint16_t v)
{
    uint32_t x = (uint32_t)v << 16;
    float r;
    memcpy(&r, &x, 4);
    return r;
}

/* Stochastic rounding of half/bf16 stores: a random word is added below
 * the kept bits of the magnitude before truncation, the rule of gfx950's
 * v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 that the HIP codec uses (restated
 * from the hardware's observed behaviour, pinned against the probe table
 * tests/golden/sr_convert.npz).  The word of population k at node (i, j):
 * the low 32 bits of (low 24 bits of (the value's binary32 pattern xor the
 * node's salt)) x multiplier_k; salt = i x 0x9E3779 + j x 0x85EBCA (24-bit
 * factors) modulo 2^32.                                                  */
static uint32_t salt_of(int i, int j)
{
    return ((uint32_t)i & 0xffffffu) * 0x9E3779u + ((uint32_t)j & 0xffffffu) * 0x85EBCAu;
}
static const uint32_t SR_MUL[9] = {0xB5297Bu, 0x9E3779u, 0xC2B2AFu, 0x85EBCBu, 0xD35A2Du,
                                   0xA3B195u, 0xE6546Bu, 0x8B1E5Fu, 0xF1357Bu};
static uint32_t sr_word(uint32_t bits, int i, int j, int k)
{
    return ((bits ^ salt_of(i, j)) & 0xffffffu) * SR_MUL[k];
}
/* binary32 bits b to binary16 with random word r: the top 13 bits of r
 * are added to the 24-bit significand below the 11 kept bits (for a half
 * subnormal result, aligned to the top of the wider dropped field), then
 * truncated; NaN keeps its top payload bits and turns quiet, a carry past
 * 65504 gives inf, binary32 zeros and subnormals give signed zero.       */
uint16_t orc_sr_f16(uint32_t b, uint32_t r)
{
    uint32_t sign = (b >> 16) & 0x8000u, a = b & 0x7fffffffu;
    if (a >= 0x7f800000u)
        return (uint16_t)(sign | 0x7c00u | (a > 0x7f800000u ? (0x200u | ((a >> 13) & 0x3ffu)) : 0u));
    if (a < 0x00800000u) return (uint16_t)sign;
    int e = (int)(a >> 23) - 127;
    uint64_t full = (a & 0x7fffffu) | 0x800000u;
    int d = e >= -14 ? 13 : 13 + (-14 - e); /* dropped significand bits */
    if (d >= 37) return (uint16_t)sign;    /* full + offset < 2^d */
    uint32_t v = (uint32_t)((full + ((uint64_t)(r >> 19) << (d - 13))) >> d);
    if (e < -14) return (uint16_t)(sign | v); /* subnormal (0x400: carried into 2^-14) */
    if (v >= 0x800u) {
        v >>= 1;
        e++;
    }
    if (e > 15) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | ((uint32_t)(e + 15) << 10) | (v & 0x3ffu));
}
/* binary32 bits b to bfloat16 with random word r: the top 16 bits of r are
 * added to the magnitude, then truncated; NaN turns quiet, a carry into
 * the exponent's all-ones pattern gives inf.                             */
uint16_t orc_sr_bf16(uint32_t b, uint32_t r)
{
    uint32_t a = b & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
    if (a == 0x7f800000u) return (uint16_t)(b >> 16);
    uint32_t a2 = a + (r >> 16);
    if (a2 >= 0x7f800000u) return (uint16_t)(((b >> 16) & 0x8000u) | 0x7f80u);
    return (uint16_t)(((b & 0x80000000u) | a2) >> 16);
}
static uint16_t enc_f16(float v, float sc, int sr, int i, int j, int k)
{
    float x = v * sc;
    if (sr) {
        uint32_t b;
        memcpy(&b, &x, 4);
        return orc_sr_f16(b, sr_word(b, i, j, k));
    }
    return orc_f32_to_f16(x);
}
static uint16_t enc_bf16(float v, float sc, int sr, int i, int j, int k)
{
    float x = v * sc;
    if (sr) {
        uint32_t b;
        memcpy(&b, &x, 4);
        return orc_sr_bf16(b, sr_word(b, i, j, k));
    }
    return orc_f32_to_bf16(x);
}

/* ------------------------------------------------- precision instances */

/* SINGLE: sp storage, sp compute */
#define CT float
#define ST float
#define LDS(v) ((CT)(v))
#define ENC(v, i, j, k) ((ST)(v))
#define HUGEFILL ((float)2147483647)
#define SINCOS sincosf
#define SFX f32
#include "orc_body.inc"
#undef CT
#undef ST
#undef LDS
#undef ENC
#undef HUGEFILL
#undef SINCOS
#undef SFX

/* DOUBLE: dp storage, dp compute */
#define CT double
#define ST double
#define LDS(v) ((CT)(v))
#define ENC(v, i, j, k) ((ST)(v))
#define HUGEFILL ((double)2147483647)
#define SINCOS sincos
#define SFX f64
#include "orc_body.inc"
#undef CT
#undef ST
#undef LDS
#undef ENC
#undef HUGEFILL
#undef SINCOS
#undef SFX

/* MIXED2: sp storage, dp compute */
#define CT double
#define ST float
#define LDS(v) ((CT)(v))
#define ENC(v, i, j, k) ((ST)(v))
#define HUGEFILL ((float)2147483647)
#define SINCOS sincos
#define SFX mx2
#include "orc_body.inc"
#undef CT
#undef ST
#undef LDS
#undef ENC
#undef HUGEFILL
#undef SINCOS
#undef SFX

/* MIXED1: IEEE half storage holding scale*(f - w), sp compute */
#define CT float
#define ST uint16_t
#define LDS(v) (orc_f16_to_f32(v) * (float)h->inv_scale)
#define ENC(v, i, j, k) enc_f16((float)(v), (float)h->scale, h->p.rounding, (i), (j), (k))
#define HUGEFILL ((uint16_t)0x7c00u)
#define SINCOS sincosf
#define SFX h16
#include "orc_body.inc"
#undef CT
#undef ST
#undef LDS
#undef ENC
#undef HUGEFILL
#undef SINCOS
#undef SFX

/* BF16: bfloat16 storage holding scale*(f - w), sp compute */
#define CT float
#define ST uint16_t
#define LDS(v) (orc_bf16_to_f32(v) * (float)h->inv_scale)
#define ENC(v, i, j, k) enc_bf16((float)(v), (float)h->scale, h->p.rounding, (i), (j), (k))
#define HUGEFILL ((uint16_t)0x4f00u)
#define SINCOS sincosf
#define SFX b16
#include "orc_body.inc"
#undef CT
#undef ST
#undef LDS
#undef ENC
#undef HUGEFILL
#undef SINCOS
#undef SFX

/* ------------------------------------------------------------- the API */

orc *orc_create(const orc_params *p)
{
    if (!p || p->lx < 2 || p->ly < 2) return NULL;
    orc *h = (orc *)calloc(1, sizeof(orc));
    if (!h) return NULL;
    h->p = *p;
    h->l = p->lx;
    h->m = p->ly;
    h->ld = p->lx + 2;
    h->plane = (size_t)(p->lx + 2) * (size_t)(p->ly + 2);
    h->stepx = p->lx / 20;  /* SRC/input.F90:74-75 */
    h->stepy = p->ly / 2;
    h->scale = (p->scale > 0) ? p->scale : 1.0;
    h->inv_scale = 1.0 / h->scale;
    switch (p->prec) {
    case ORC_DOUBLE: h->vt = &vtab_f64; h->esize = 8; break;
    case ORC_MIXED1: h->vt = &vtab_h16; h->esize = 2; break;
    case ORC_MIXED2: h->vt = &vtab_mx2; h->esize = 4; break;
    case ORC_BF16: h->vt = &vtab_b16; h->esize = 2; break;
    default: h->vt = &vtab_f32; h->esize = 4; break;
    }
    for (int k = 0; k < 9; k++) {
        h->a[k] = malloc(h->plane * (size_t)h->esize);
        h->b[k] = malloc(h->plane * (size_t)h->esize);
        if (!h->a[k] || !h->b[k]) { orc_destroy(h); return NULL; }
    }
    h->obs = (int32_t *)calloc((size_t)p->lx * (size_t)p->ly, sizeof(int32_t));
    if (!h->obs) { orc_destroy(h); return NULL; }
    return h;
}

void orc_destroy(orc *h)
{
    if (!h) return;
    for (int k = 0; k < 9; k++) { free(h->a[k]); free(h->b[k]); }
    free(h->obs);
    free(h);
}

int orc_storage_bytes(const orc *h) { return h->esize; }

int orc_setup(orc *h)
{
    if (h->p.init_v < 0 || h->p.init_v > 5) return -1; /* SRC/init.F90:68-71 */
    /* NOSHIFT stores whole populations: the deviation form and the OPEN
     * channel's inlet are defined on deviations only                     */
    if (h->p.noshift && (h->p.form || h->p.bc == ORC_BC_OPEN)) return -1;
    h->vt->setup(h);
    return 0;
}

void orc_boundaries(orc *h) { h->vt->boundaries(h); }

void orc_propagation(orc *h)
{
    if (h->p.impl == ORC_ORIGINAL) h->vt->movef(h);
}

void orc_collision(orc *h)
{
    if (h->p.impl == ORC_FUSED) h->vt->colmc(h);
    else h->vt->col(h);
}

/* one pass of the loop body of SRC/bgk2d.F90:92-113 without diagnostics */
void orc_step(orc *h, int32_t nsteps)
{
    for (int32_t s = 0; s < nsteps; s++) {
        orc_boundaries(h);
        orc_propagation(h);
        orc_collision(h);
    }
}

void orc_constants(const orc *h, double out[4])
{
    out[0] = h->omega;
    out[1] = h->fgrad;
    out[2] = h->u00;
    out[3] = h->force;
}

void orc_get_pop(const orc *h, int32_t k, double *out) { h->vt->decode_plane(h, h->a[k], out); }

void orc_get_raw(const orc *h, int32_t k, void *out) { memcpy(out, h->a[k], h->plane * (size_t)h->esize); }

void orc_set_raw(orc *h, int32_t k, const void *in) { memcpy(h->a[k], in, h->plane * (size_t)h->esize); }

void orc_get_raw_b(const orc *h, int32_t k, void *out) { memcpy(out, h->b[k], h->plane * (size_t)h->esize); }

void orc_get_obs(const orc *h, int32_t *out, int32_t bbox[4])
{
    memcpy(out, h->obs, sizeof(int32_t) * (size_t)h->l * (size_t)h->m);
    bbox[0] = h->imin; bbox[1] = h->imax; bbox[2] = h->jmin; bbox[3] = h->jmax;
}

void orc_diagno(const orc *h, int32_t mode, double out[4]) { h->vt->diagno(h, mode, out); }

void orc_set_obs(orc *h, const int32_t *mask)
{
    const int l = h->l, m = h->m;
    if (!h->obs) h->obs = (int32_t *)calloc((size_t)l * m, sizeof(int32_t));
    h->imin = l; h->jmin = m; h->imax = 0; h->jmax = 0; h->nobs = 0;
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= l; i++) {
            size_t q = (size_t)(i - 1) + (size_t)(j - 1) * l;
            h->obs[q] = mask[q] == 1 ? 1 : 0;
            if (h->obs[q]) {
                h->nobs++;
                if (i < h->imin) h->imin = i;
                if (i > h->imax) h->imax = i;
                if (j < h->jmin) h->jmin = j;
                if (j > h->jmax) h->jmax = j;
            }
        }
    h->p.obstacle = h->nobs > 0;
}

void orc_draglift(const orc *h, int32_t i0, int32_t i1, int32_t j0, int32_t j1, double out[2])
{
    h->vt->draglift(h, i0 < 1 ? 1 : i0, i1 > h->l ? h->l : i1, j0 < 1 ? 1 : j0, j1 > h->m ? h->m : j1, out);
}

void orc_varm(const orc *h, double *u, double *v, double *rho) { h->vt->varm(h, u, v, rho); }

void orc_prof_i(const orc *h, int32_t jc, double *u, double *v, double *rho) { h->vt->prof_i(h, jc, u, v, rho); }

void orc_prof_j(const orc *h, int32_t ic, double *u, double *v, double *rho) { h->vt->prof_j(h, ic, u, v, rho); }

void orc_probe(const orc *h, int32_t i0, int32_t j0, double out[3]) { h->vt->probe(h, i0, j0, out); }
