// bgk_host.hpp -- host runtime of libbgk2d_hip.so shared by its translation
// units: the context and batch state, the launchers of the kernels of
// bgk_kernels.hpp and bgk_serve.hpp, and the step loop.  The heavy
// templates (the step loop of each storage kind, the batch kernels) are
// instantiated in their own translation units (bgk_kind_*.hip), declared
// `extern template` below, so the library builds in parallel.
#pragma once
#include "../../include/bgk2d.h"
#include "bgk_kernels.hpp"
#include "bgk_serve.hpp"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <map>
#include <string>
#include <vector>

using namespace bgk;

struct bgk_ctx {
    bgk_params_t p;
    int l, m;
    int esize, A, off, pitch;
    size_t plane_elems;  // elements per plane, incl. alignment padding
    int dev;
    hipStream_t s;
    hipStream_t s2;        // the obstacle's strip blocks of the two-step kernel
    hipEvent_t evf, evj;   // fork to s2 / join back to s
    void *mem;
    void *plane[18];  // [0..7] set 0, [8..15] set 1, [16], [17] pop 19
    int cur;          // which set is `a`
    int rcur;         // which pop-19 plane is current (16 + rcur)
    int tb_h;         // strip height of the two-step kernel (0: not sized yet)
    bool tb_fill;     // the lattice fills the device with strips of >= TB_HMIN rows
    uint8_t *d_obs;
    int imin, imax, jmin, jmax, nobs;
    std::vector<int32_t> h_obs;
    double *d_rows, *d_red, *h_red;
    void *d_stage;
    float *d_fields;  // bgk_fields output, allocated on first use
    size_t stage_bytes;
    // constants in the compute kind, widened
    double omega, fgrad, u00, force, scale;
    double inflow[3];
    double umax;  // OPEN: peak inflow speed in the compute kind
    bool inited;
    std::string err;
    hipEvent_t ev0, ev1;
    double last_ms;
    std::map<long long, hipGraphExec_t> graphs;
};

namespace bgk_host {


inline int fail(bgk_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail((c), BGK_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

inline bool is_fp16(int prec) { return prec == BGK_PREC_FP16; }
inline bool is_bf16(int prec) { return prec == BGK_PREC_BF16; }

inline int esize_of(int prec)
{
    switch (prec) {
    case BGK_PREC_DOUBLE: return 8;
    case BGK_PREC_FP16:
    case BGK_PREC_BF16: return 2;
    default: return 4;
    }
}

// hencol + build_bcond in the compute kind C (SRC/hencol.F90:30-50,
// SRC/build.bcond.F90:36-45), and the stored inflow equilibria of
// SRC/bcond.inflow.F90:101-113 with u_inflow = 0.1 (a default-real literal).
template <class C>
void derive(bgk_ctx *c)
{
    const bgk_params_t &p = c->p;
    C u0 = (C)p.u0, u00;
    c->umax = p.bc == BGK_BC_OPEN ? (double)u0 : 0.0;
    if (p.bc == BGK_BC_CHANNEL) {
        u00 = (C)0.0;
    } else if (p.bc == BGK_BC_OPEN) {  // no lid and no body force: u0 is the inflow peak
        u00 = (C)0.0;
        u0 = (C)0.0;
    } else {
        u00 = u0;
        u0 = (C)0.0;
    }
    C svisc = (C)p.svisc;
    C omega = ((C)1.0 + (C)1.0) / (((C)6.0 * svisc) + (C)1.0);
    if (is_fp16(p.prec) && p.mixed1_quirk) u0 = (C)0.0;
    C fgrad = (C)4.0 * u0 * svisc / ((C)3.0 * (C)c->m * (C)c->m);
    c->omega = (double)omega;
    c->fgrad = (double)fgrad;
    c->u00 = (double)u00;
    c->force = (double)(u00 / (C)6.0);
    const C xj = (C)0.1f, yj = (C)0.0;
    const C cvsq = xj * xj + yj * yj;
    const C cxq[3] = {eq_d<C>(xj - yj, cvsq), eq_d<C>(xj + yj, cvsq), eq_d<C>(xj, cvsq)};
    const C wq[3] = {K<C>::p2, K<C>::p2, K<C>::p1};
    // crho*p*(cte1+cx) in the compute kind (cte1 = 1 for whole populations,
    // SRC/bcond.inflow.F90:48-52); each stored copy is rounded to the
    // storage kind where it is stored (x = 0, row j)
    const C cte1 = p.noshift ? (C)1.0 : (C)0.0;
    for (int q = 0; q < 3; q++) c->inflow[q] = (double)((C)1.0 * wq[q] * (cte1 + cxq[q]));
}

template <class C>
Phys<C> phys(const bgk_ctx *c)
{
    Phys<C> P;
    P.omega = (C)c->omega;
    P.fgrad = (C)c->fgrad;
    P.force = (C)c->force;
    P.sc = (C)c->scale;
    P.inv = (C)(1.0 / c->scale);
    for (int q = 0; q < 3; q++) P.inflow[q] = (C)c->inflow[q];
    P.umax = (C)c->umax;
    P.channel_force = c->p.bc == BGK_BC_CHANNEL ? (c->p.forcing_y ? 2 : 1) : 0;
    P.form = c->p.form;
    P.les = c->p.les;
    P.trick1 = c->p.trick1;
    P.ctes = (C)c->p.cte_s;
    P.rounding = (is_fp16(c->p.prec) || is_bf16(c->p.prec)) ? c->p.rounding : 0;
    // LES constants with the operations of SRC/col_MC.F90:200-206 in the
    // compute kind (IEEE division and products, so the device would round
    // them the same way)
    const C omega1 = P.omega, cs = P.ctes;
    const C inv1 = (C)1.0 / omega1;
    P.les_h = (C)1.0 / ((C)2.0 * omega1);
    P.les_i2 = inv1 * inv1;
    P.les_c = (C)18.0 * (cs * cs);
    P.usc = (C)1.0;
    P.dsc = (C)1.0;
    P.noshift = c->p.noshift;
    return P;
}

// Kernel instance flags (Cfg in bgk_kernels.hpp): 0 for the reference's
// form with nearest-even stores, 1 for the deviation form with stochastic
// rounding of 16-bit stores, 2 (fp32/fp64 storage) for whole populations
// (NOSHIFT, which needs form 0), -1 (16-bit kinds only) for the other mixes.
inline int fl_of(const bgk_ctx *c)
{
    const bool narrow = is_fp16(c->p.prec) || is_bf16(c->p.prec);
    const int form = c->p.form, rnd = narrow ? c->p.rounding : form;
    if (c->p.noshift) return narrow ? -1 : 2;
    if (form == 0 && rnd == 0) return 0;
    if (form == 1 && rnd == 1) return 1;
    return -1;
}

// calls f(std::integral_constant<int, FL>) for the instance flags of c;
// the generic instance exists for the 16-bit kinds only
template <class S, class F>
void with_fl(const bgk_ctx *c, F &&f)
{
    const int fl = fl_of(c);
    if (fl == 1) {
        f(std::integral_constant<int, 1>{});
    } else if (fl == 0) {
        f(std::integral_constant<int, 0>{});
    } else if (fl == 2) {
        if constexpr (sizeof(S) != 2) f(std::integral_constant<int, 2>{});
    } else {
        if constexpr (sizeof(S) == 2) f(std::integral_constant<int, -1>{});
    }
}

inline Lat lat(const bgk_ctx *c)
{
    Lat L;
    const int a = c->cur, b = 1 - c->cur;
    for (int k = 0; k < 8; k++) {
        L.a[k] = c->plane[a * 8 + k];
        L.b[k] = c->plane[b * 8 + k];
    }
    L.r = c->plane[16 + c->rcur];
    L.rn = c->plane[17 - c->rcur];
    L.obs = c->d_obs;
    L.l = c->l;
    L.m = c->m;
    L.pitch = c->pitch;
    L.off = c->off;
    L.rs = c->pitch;
    L.imin = c->imin;
    L.imax = c->imax;
    L.jmin = c->jmin;
    L.jmax = c->jmax;
    L.stepx = c->l / 20;  // SRC/input.F90:74-75
    L.stepy = c->m / 2;
    return L;
}

inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

inline const dim3 B2(64, 4);

template <class S, class C>
int launch_boundaries(bgk_ctx *c)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    const int l = c->l, m = c->m;
    auto pass = [&](int ps, int t0, int n, int lid) {
        if (n <= 0) return;
        hipLaunchKernelGGL((k_bc<S, C>), dim3((n + 255) / 256), dim3(256), 0, c->s, L, P, ps, t0, n, lid);
    };
    switch (c->p.bc) {
    case BGK_BC_PERIODIC:
        pass(BP_XPER, 0, m + 2, 0);
        pass(BP_YPER, 0, l + 2, 0);
        break;
    case BGK_BC_CHANNEL:
        pass(BP_XPER, 0, m + 2, 0);
        pass(BP_WALL_Y, 0, l + 2, 0);
        break;
    case BGK_BC_INFLOW:
        pass(BP_INFLOW_X, 0, m + 2, 0);
        pass(BP_YPER, 0, l + 2, 0);
        break;
    case BGK_BC_COUETTE:
        pass(BP_XPER, 1, m, 0);
        pass(BP_WALL_Y, 1, l, 1);
        break;
    case BGK_BC_OPEN:
        pass(BP_OPEN_X, 1, m, 0);
        pass(BP_FREESLIP_Y, 0, l + 2, 0);
        break;
    case BGK_BC_STEP:
        pass(BP_STEP_OUT, 0, m + 2, 0);
        pass(BP_STEP_IN, L.stepy, m + 2 - L.stepy, 0);
        pass(BP_WALL_Y, 1, l, 0);
        pass(BP_STEP_TOP, 1, L.stepx, 0);
        pass(BP_STEP_FACE, 1, L.stepy, 0);
        break;
    default:
        if (c->p.trick1) {  // SRC/bcond.driven.F90:43-81: side walls after the lid
            pass(BP_WALL_Y, 1, l, 1);
            pass(BP_DRV_X, 1, m, 0);
        } else {
            pass(BP_DRV_X, 1, m, 0);
            pass(BP_WALL_Y, 1, l, 1);
        }
        break;
    }
    if (c->p.obstacle && c->nobs > 0) {
        int nx = c->imax - c->imin + 1, ny = c->jmax - c->jmin + 1;
        hipLaunchKernelGGL((k_bc_obs<S>), grid2(nx, ny, dim3(16, 16)), dim3(16, 16), 0, c->s, L);
    }
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S, class C>
int launch_propagation(bgk_ctx *c)
{
    if (c->p.impl != BGK_IMPL_ORIGINAL) return BGK_OK;
    Lat L = lat(c);
    hipLaunchKernelGGL((k_movef<S>), grid2(c->l, c->m, B2), B2, 0, c->s, L);
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S, class C>
int launch_collision(bgk_ctx *c)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    if (c->p.impl == BGK_IMPL_ORIGINAL) {
        hipLaunchKernelGGL((k_col<S, C>), grid2(c->l, c->m, B2), B2, 0, c->s, L, P);
    } else {
        hipLaunchKernelGGL((k_colmc<S, C>), grid2(c->l, c->m, B2), B2, 0, c->s, L, P);
        c->cur = 1 - c->cur;  // pointer swap of SRC/col_MC.F90:311-341
    }
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

// the boundary folded into the hot kernels: FUSED, all variants but STEP
inline bool foldable(const bgk_ctx *c)
{
    return c->p.impl == BGK_IMPL_FUSED && c->p.bc != BGK_BC_STEP;
}

template <class S, class C, int BC, bool OBS>
void launch_fused_t(bgk_ctx *c)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    with_fl<S>(c, [&](auto flc) {
        constexpr int FL = decltype(flc)::value;
        if constexpr (sizeof(S) == 2)  // 16-bit storage: two nodes per thread
            hipLaunchKernelGGL((k_fused2<S, C, BC, OBS, FL>), grid2((c->l + 1) / 2, c->m, B2), B2, 0, c->s, L, P);
        else
            hipLaunchKernelGGL((k_fused<S, C, BC, OBS, FL>), grid2(c->l, c->m, B2), B2, 0, c->s, L, P);
    });
    c->cur = 1 - c->cur;
    c->rcur = 1 - c->rcur;  // pop 19 written to the second plane
}

template <class S, class C, bool OBS>
void launch_fused_o(bgk_ctx *c)
{
    switch (c->p.bc) {
    case BGK_BC_PERIODIC: launch_fused_t<S, C, BK_PERIODIC, OBS>(c); break;
    case BGK_BC_CHANNEL: launch_fused_t<S, C, BK_CHANNEL, OBS>(c); break;
    case BGK_BC_INFLOW: launch_fused_t<S, C, BK_INFLOW, OBS>(c); break;
    case BGK_BC_COUETTE: launch_fused_t<S, C, BK_COUETTE, OBS>(c); break;
    case BGK_BC_OPEN: launch_fused_t<S, C, BK_OPEN, OBS>(c); break;
    default: launch_fused_t<S, C, BK_DRIVEN, OBS>(c); break;
    }
}

template <class S, class C>
int launch_step(bgk_ctx *c)
{
    if (foldable(c)) {
        if (c->p.obstacle && c->nobs > 0)
            launch_fused_o<S, C, true>(c);
        else
            launch_fused_o<S, C, false>(c);
        HIPCHK(c, hipGetLastError());
        return BGK_OK;
    }
    int r = launch_boundaries<S, C>(c);
    if (r == BGK_OK) r = launch_propagation<S, C>(c);
    if (r == BGK_OK) r = launch_collision<S, C>(c);
    return r;
}

// Two FUSED steps per pass over HBM (k_fused_tb): the default on lattices
// large enough to fill the device with strips of at least TB_HMIN rows,
// where it beats one step per pass (Taylor-Green 8192^2: FP32 122 k against
// 75 k MLUPS, FP16 157 k against 145 k).  Needs size_tb first.
constexpr int TB_HMIN = 8;
// Wavefronts per CU that the two-step grids are sized for, at most (a
// grid is one round of the blocks the device holds): 16-bit blocks of 256
// columns, which the occupancy query allows seven of per CU, ran slower
// sized for seven than for six (Taylor-Green 8192^2 FP16: 158 k against
// 161 k MLUPS); the fp32 blocks hold six anyway (their LDS ring), the
// 16-bit 512-column blocks three (their registers).
constexpr int TB_MAX_WAVES_PER_CU = 24;
constexpr int TB_MAX_PER_CU = TB_MAX_WAVES_PER_CU / (TB_TW / 64);  // blocks of the batch kernel
inline bool use_tb(const bgk_ctx *c)
{
    if (!foldable(c)) return false;
    if (c->p.tblock == 2) return true;
    return c->p.tblock == 0 && c->tb_fill;
}

// Strip blocks of height H that must apply the obstacle rule: those whose
// updated nodes reach the obstacle's bounding box grown by one node, where
// fix_pulls applies it.  Strip bx computes step t+1 on columns bx*TW ..
// bx*TW+TW+1 (its strip and rim columns), so it reaches the grown columns
// imin-1..imax+1 exactly when (imin-3)/TW <= bx <= (imax+1)/TW (imin >= 2);
// block row by computes rows by*H .. by*H+H+1, likewise.  Along a
// periodic direction the first and last blocks also update wrapped nodes
// of the far edge (their rim or halo row): when the grown box reaches an
// edge (the obstacle touches the second or second-to-last row or column),
// every block along that direction is taken.
// (first strip, last strip, first row, last row) of blocks; x > y: none.
inline int4 tb_box(const bgk_ctx *c, int H, int TW, bool xper, bool yper)
{
    if (!(c->p.obstacle && c->nobs > 0)) return make_int4(0, -1, 0, -1);
    const int strips = (c->l + TW - 1) / TW, rows = (c->m + H - 1) / H;
    int4 b = make_int4(max(0, (c->imin - 3) / TW), min(strips - 1, (c->imax + 1) / TW),
                       max(0, (c->jmin - 3) / H), min(rows - 1, (c->jmax + 1) / H));
    if (xper && (c->imin - 1 <= 1 || c->imax + 1 >= c->l)) {
        b.x = 0;
        b.y = strips - 1;
    }
    if (yper && (c->jmin - 1 <= 1 || c->jmax + 1 >= c->m)) {
        b.z = 0;
        b.w = rows - 1;
    }
    return b;
}

// Strip height of k_fused_tb: as many strips of rows as make the grid one
// round of the blocks the device holds at once, so no CU idles in a
// partial last round; at least TB_HMIN rows, since each strip also
// computes two rows of its neighbours.
template <class S, class C, int BC, bool OBS>
int size_tb(bgk_ctx *c)
{
    constexpr int TW = tb_tw<S>();
    int per_cu = 0, cus = 0;
    // with an obstacle the grid is still that of the obstacle-free kernel:
    // the obstacle's few strip blocks run beside it (launch_tb_t); leaving
    // them room in the round measured slower (91 k against 94 k MLUPS)
    hipError_t e = hipSuccess;
    with_fl<S>(c, [&](auto flc) {
        constexpr int FL = decltype(flc)::value;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_tb<S, C, BC, false, FL>, TW, 0);
    });
    if (e != hipSuccess) return fail(c, BGK_E_HIP, "hipOccupancyMaxActiveBlocksPerMultiprocessor: %s", hipGetErrorString(e));
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->dev));
    if (per_cu > TB_MAX_WAVES_PER_CU / (TW / 64)) per_cu = TB_MAX_WAVES_PER_CU / (TW / 64);
    if (per_cu < 1) per_cu = 1;
    if (cus < 1) cus = 1;
    const long long slots = (long long)per_cu * cus;
    const long long strips = (c->l + TW - 1) / TW;
    long long per_strip = slots / strips;
    if (per_strip < 1) per_strip = 1;
    const int h = (int)((c->m + per_strip - 1) / per_strip);
    c->tb_fill = h >= TB_HMIN;
    c->tb_h = h >= TB_HMIN ? h : TB_HMIN;
    return BGK_OK;
}

template <class S, class C, bool OBS>
int size_tb_o(bgk_ctx *c)
{
    switch (c->p.bc) {
    case BGK_BC_PERIODIC: return size_tb<S, C, BK_PERIODIC, OBS>(c);
    case BGK_BC_CHANNEL: return size_tb<S, C, BK_CHANNEL, OBS>(c);
    case BGK_BC_INFLOW: return size_tb<S, C, BK_INFLOW, OBS>(c);
    case BGK_BC_COUETTE: return size_tb<S, C, BK_COUETTE, OBS>(c);
    case BGK_BC_OPEN: return size_tb<S, C, BK_OPEN, OBS>(c);
    default: return size_tb<S, C, BK_DRIVEN, OBS>(c);
    }
}

// the obstacle's strip blocks march H / TB_BOX_DIV rows each (C4 cylinder
// channel, FP32 / FP16 MLUPS: H 79.9 k / 85.3 k, H/2 98.3 k / 102.6 k,
// H/4 98.1 k / 102.0 k, H/8 97.3 k / 101.7 k)
#ifndef BGK_TB_BOX_DIV
#define BGK_TB_BOX_DIV 4
#endif
constexpr int TB_BOX_DIV = BGK_TB_BOX_DIV;

template <class S, class C, int BC, bool OBS>
int launch_tb_t(bgk_ctx *c)
{
    if (c->tb_h == 0) {
        int r = size_tb<S, C, BC, OBS>(c);
        if (r != BGK_OK) return r;
    }
    constexpr int TW = tb_tw<S>();
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    const int H = c->tb_h;
    const int strips = (c->l + TW - 1) / TW, rows = (c->m + H - 1) / H;
    const int rows8 = (rows + NXCD - 1) / NXCD * NXCD;
    // strip blocks with the obstacle rule
    const int4 box = OBS ? tb_box(c, H, TW, Per<BC>::x, Per<BC>::y) : make_int4(0, -1, 0, -1);
    if (OBS && box.y >= box.x) {
        // the obstacle's blocks on a second stream, beside the others:
        // the rows of the skipped blocks, in blocks of a quarter of H
        const int nbx = box.y - box.x + 1, r0 = box.z * H + 1, r1 = min((box.w + 1) * H, c->m);
        const int hb = max(TB_HMIN, H / TB_BOX_DIV), nsub = (r1 - r0 + hb) / hb;
        HIPCHK(c, hipEventRecord(c->evf, c->s));
        HIPCHK(c, hipStreamWaitEvent(c->s2, c->evf, 0));
        with_fl<S>(c, [&](auto flc) {
            constexpr int FL = decltype(flc)::value;
            hipLaunchKernelGGL((k_fused_tb_box<S, C, BC, FL>), dim3(nbx * nsub), dim3(TW), 0, c->s2, L, P, box.x,
                               nbx, r0, r1, hb);
        });
        HIPCHK(c, hipEventRecord(c->evj, c->s2));
    }
    with_fl<S>(c, [&](auto flc) {
        constexpr int FL = decltype(flc)::value;
        hipLaunchKernelGGL((k_fused_tb<S, C, BC, false, FL>), dim3(strips * rows8), dim3(TW), 0, c->s, L, P, H,
                           strips, rows, box);
    });
    if (OBS && box.y >= box.x) HIPCHK(c, hipStreamWaitEvent(c->s, c->evj, 0));
    c->cur = 1 - c->cur;
    c->rcur = 1 - c->rcur;
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S, class C, bool OBS>
int launch_tb_o(bgk_ctx *c)
{
    switch (c->p.bc) {
    case BGK_BC_PERIODIC: return launch_tb_t<S, C, BK_PERIODIC, OBS>(c);
    case BGK_BC_CHANNEL: return launch_tb_t<S, C, BK_CHANNEL, OBS>(c);
    case BGK_BC_INFLOW: return launch_tb_t<S, C, BK_INFLOW, OBS>(c);
    case BGK_BC_COUETTE: return launch_tb_t<S, C, BK_COUETTE, OBS>(c);
    case BGK_BC_OPEN: return launch_tb_t<S, C, BK_OPEN, OBS>(c);
    default: return launch_tb_t<S, C, BK_DRIVEN, OBS>(c);
    }
}

// n complete steps: pairs through k_fused_tb where it is used, the rest one
// step at a time
template <class S, class C>
int launch_steps(bgk_ctx *c, int n)
{
    int t = 0;
    if (use_tb(c)) {
        const bool obs = c->p.obstacle && c->nobs > 0;
        for (; t + 2 <= n; t += 2) {
            int r = obs ? launch_tb_o<S, C, true>(c) : launch_tb_o<S, C, false>(c);
            if (r != BGK_OK) return r;
        }
    }
    for (; t < n; t++) {
        int r = launch_step<S, C>(c);
        if (r != BGK_OK) return r;
    }
    return BGK_OK;
}

// Steps are replayed from HIP graphs of GRAPH_STEPS steps each (an even
// count, so the FUSED set parities are the same before and after a
// replay); one graph per starting parity, built on first use.
constexpr int GRAPH_STEPS = 32;

template <class S, class C>
int graph_for(bgk_ctx *c, hipGraphExec_t *out)
{
    const long long key = c->cur + 2 * c->rcur;
    auto it = c->graphs.find(key);
    if (it != c->graphs.end()) {
        *out = it->second;
        return BGK_OK;
    }
    const int cur0 = c->cur, rcur0 = c->rcur;
    HIPCHK(c, hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    // the capture ends whatever launch_steps returns
    int r = launch_steps<S, C>(c, GRAPH_STEPS);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->s, &g);
    c->cur = cur0;
    c->rcur = rcur0;
    if (r != BGK_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (r != BGK_OK) return r;
        return fail(c, BGK_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, BGK_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    c->graphs[key] = x;
    *out = x;
    return BGK_OK;
}

// enqueue nsteps steps on the context's stream, without waiting
template <class S, class C>
int enqueue_t(bgk_ctx *c, int nsteps)
{
    if (foldable(c) && c->tb_h == 0) {
        // size the two-step kernel (which decides whether it is used)
        // before any capture
        int r = c->p.obstacle && c->nobs > 0 ? size_tb_o<S, C, true>(c) : size_tb_o<S, C, false>(c);
        if (r != BGK_OK) return r;
    }
    int t = 0;
    if (nsteps >= GRAPH_STEPS) {
        hipGraphExec_t g;
        int r = graph_for<S, C>(c, &g);
        if (r != BGK_OK) return r;
        for (; t + GRAPH_STEPS <= nsteps; t += GRAPH_STEPS) HIPCHK(c, hipGraphLaunch(g, c->s));
    }
    return launch_steps<S, C>(c, nsteps - t);
}

template <class S, class C>
int run_t(bgk_ctx *c, int nsteps)
{
    HIPCHK(c, hipEventRecord(c->ev0, c->s));
    int r = enqueue_t<S, C>(c, nsteps);
    if (r != BGK_OK) return r;
    HIPCHK(c, hipEventRecord(c->ev1, c->s));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = nsteps > 0 ? (double)ms / nsteps : 0.0;
    return BGK_OK;
}

// sin and cos of one argument through the C library's sincos: the
// reference's build (and gcc's) fuse the sin/cos pair of SRC/init.F90:109-110
// into one sincos call, whose last bit can differ from separate sin and cos
template <class C>
void host_sincos(C arg, C *s, C *c);
template <>
inline void host_sincos<float>(float arg, float *s, float *c) { sincosf(arg, s, c); }
template <>
inline void host_sincos<double>(double arg, double *s, double *c) { sincos(arg, s, c); }

// the trig tables of init: sin and cos of 2 pi x over nodes 0..l+1, then
// of 2 pi y over nodes 0..m+1 (2 (l + m + 4) elements of the compute kind)
inline size_t init_table_bytes(const bgk_ctx *c)
{
    const size_t ck = (c->p.prec == BGK_PREC_DOUBLE || c->p.prec == BGK_PREC_MIXED2) ? 8 : 4;
    return 2 * ((size_t)c->l + (size_t)c->m + 4) * ck;
}

template <class C>
void init_tables(const bgk_ctx *c, C *h)
{
    const int nx = c->l + 2, ny = c->m + 2;
    // SRC/init.F90:88-111: pi is a default-real literal in kind mykind,
    // x = (real(i)-0.5d0)/real(l) evaluated in double, sin/cos in mykind
    const C pi = (C)(float)3.141592653589793238462643383279;
    for (int i = 0; i < nx; i++) {
        C x = (C)(((double)(C)i - 0.5) / (double)(C)c->l);
        host_sincos<C>((C)2 * pi * x, &h[i], &h[nx + i]);
    }
    for (int j = 0; j < ny; j++) {
        C y = (C)(((double)(C)j - 0.5) / (double)(C)c->m);
        host_sincos<C>((C)2 * pi * y, &h[2 * nx + j], &h[2 * nx + ny + j]);
    }
}

// uploads the tables h into the device buffer d, launches k_init and
// waits for it (h and d are free again on return, also on failure)
template <class S, class C>
int init_launch(bgk_ctx *c, double amp, int init_v, const C *h, C *d)
{
    const bool tgv = c->p.bc == BGK_BC_PERIODIC;
    const int nx = c->l + 2, ny = c->m + 2;
    hipError_t e = hipMemcpyAsync(d, h, init_table_bytes(c), hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) {
        Lat L = lat(c);
        Phys<C> P = phys<C>(c);
        hipLaunchKernelGGL((k_init<S, C>), grid2(nx, ny, B2), B2, 0, c->s, L, P, d, d + nx, d + 2 * nx,
                           d + 2 * nx + ny, amp, tgv ? (init_v == 1 ? 2 : 1) : 0);
        e = hipGetLastError();
    }
    const hipError_t w = hipStreamSynchronize(c->s);
    if (e == hipSuccess) e = w;
    if (e != hipSuccess) return fail(c, BGK_E_HIP, "init: %s", hipGetErrorString(e));
    return BGK_OK;
}

template <class S, class C>
int init_t(bgk_ctx *c, double amp, int init_v)
{
    std::vector<C> h(init_table_bytes(c) / sizeof(C));
    init_tables<C>(c, h.data());
    C *d = nullptr;
    HIPCHK(c, hipMalloc(&d, init_table_bytes(c)));
    const int r = init_launch<S, C>(c, amp, init_v, h.data(), d);
    const hipError_t e = hipFree(d);
    if (r == BGK_OK && e != hipSuccess) return fail(c, BGK_E_HIP, "init: %s", hipGetErrorString(e));
    return r;
}

// init with caller-provided tables: hstage host memory and dscratch device
// memory of init_table_bytes(c) bytes each (no allocation)
template <class S, class C>
int init_static_t(bgk_ctx *c, double amp, int init_v, void *dscratch, void *hstage)
{
    init_tables<C>(c, (C *)hstage);
    return init_launch<S, C>(c, amp, init_v, (const C *)hstage, (C *)dscratch);
}

template <class S, class C>
int rows_t(bgk_ctx *c, bool final_)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    hipLaunchKernelGGL((k_rows<S, C>), dim3(c->m), dim3(ROWS_T), 0, c->s, L, P, c->d_rows);
    if (final_) hipLaunchKernelGGL(k_rows_final<>, dim3(1), dim3(ROWS_T), 0, c->s, c->d_rows, c->m, c->d_red);
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S, class C>
int line_t(bgk_ctx *c, int mode, int c0, int c1, int n)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    hipLaunchKernelGGL((k_line<S, C>), dim3((n + 255) / 256), dim3(256), 0, c->s, L, P, mode, c0, c1, c->d_red);
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S, class C>
int draglift_t(bgk_ctx *c, int i0, int i1, int j0, int j1)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    hipLaunchKernelGGL((k_draglift<S, C>), dim3(1), dim3(DL_T), 0, c->s, L, P, i0, i1, j0, j1, c->d_red);
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S, class C>
int fields_t(bgk_ctx *c, float *d)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    const size_t n = (size_t)c->l * c->m;
    hipLaunchKernelGGL((k_fields<S, C>), grid2(c->l, c->m, B2), B2, 0, c->s, L, P, d, d + n, d + 2 * n);
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S>
int fill_t(bgk_ctx *c, S v)
{
    long long n = (long long)c->plane_elems * 18;
    hipLaunchKernelGGL((k_fill<S>), dim3(2048), dim3(256), 0, c->s, c->mem, n, v);
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

template <class S>
int pack_t(bgk_ctx *c, int k, bool up)
{
    Lat L = lat(c);
    void *plane = k == 8 ? c->plane[16 + c->rcur] : c->plane[c->cur * 8 + k];
    if (up)
        hipLaunchKernelGGL((k_unpack<S>), grid2(c->l + 2, c->m + 2, B2), B2, 0, c->s, L, plane, (const S *)c->d_stage);
    else
        hipLaunchKernelGGL((k_pack<S>), grid2(c->l + 2, c->m + 2, B2), B2, 0, c->s, L, (const void *)plane,
                           (S *)c->d_stage);
    HIPCHK(c, hipGetLastError());
    return BGK_OK;
}

// precision dispatch: storage kind S and compute kind C of each mode
#define DISPATCH(c, FN, ...)                                                              \
    [&]() -> int {                                                                        \
        switch ((c)->p.prec) {                                                            \
        case BGK_PREC_DOUBLE: return FN<double, double>((c), ##__VA_ARGS__);              \
        case BGK_PREC_FP16: return FN<_Float16, float>((c), ##__VA_ARGS__);               \
        case BGK_PREC_MIXED2: return FN<float, double>((c), ##__VA_ARGS__);               \
        case BGK_PREC_BF16: return FN<bgk::bf16_t, float>((c), ##__VA_ARGS__);            \
        default: return FN<float, float>((c), ##__VA_ARGS__);                             \
        }                                                                                 \
    }()

// pitched plane layout: a row holds nodes 0..l+1 from element off = A-1
// (A elements = 128 bytes, so node 1 starts a line), plus two elements of
// slack for the paired 16-bit accesses at the end of a row
inline void layout_of(const bgk_params_t *p, int *esize, int *A, int *off, int *pitch, size_t *plane_elems)
{
    *esize = esize_of(p->prec);
    *A = 128 / *esize;
    *off = *A - 1;
    *pitch = ((*A + p->lx + 3) + *A - 1) / *A * *A;
    size_t pe = (size_t)(p->ly + 2) * (size_t)*pitch;
    const size_t align = 256 / *esize;
    *plane_elems = (pe + align - 1) / align * align;
}

// device bytes one context allocates (populations, mask, reduction scratch)
inline size_t footprint_of(const bgk_params_t *p)
{
    int es, A, off, pitch;
    size_t pe;
    layout_of(p, &es, &A, &off, &pitch, &pe);
    size_t obs_bytes = ((size_t)p->lx * p->ly + 255) / 256 * 256;
    size_t rows_bytes = sizeof(double) * ROW_VALS * (size_t)p->ly;
    size_t red_bytes = sizeof(double) * 3 * (size_t)(p->lx > p->ly ? p->lx : p->ly) + 64;
    return pe * (size_t)es * 18 + obs_bytes + rows_bytes + red_bytes;
}

inline int check_ctx(bgk_ctx *c)
{
    if (!c) return BGK_E_INVALID;
    hipError_t e = hipSetDevice(c->dev);
    if (e != hipSuccess) return fail(c, BGK_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    return BGK_OK;
}

inline int fetch_red(bgk_ctx *c, int n)
{
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->d_red, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return BGK_OK;
}

// Everything a context holds, whatever part of bgk_create succeeded:
// handles and buffers not yet made are null (the context is value-initialised).
inline void ctx_release(bgk_ctx *c)
{
    if (c->s) (void)hipStreamSynchronize(c->s);
    if (c->s2) (void)hipStreamSynchronize(c->s2);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
    if (c->mem) (void)hipFree(c->mem);
    if (c->d_obs) (void)hipFree(c->d_obs);
    if (c->d_rows) (void)hipFree(c->d_rows);
    if (c->d_red) (void)hipFree(c->d_red);
    if (c->h_red) (void)hipHostFree(c->h_red);
    if (c->d_stage) (void)hipFree(c->d_stage);
    if (c->d_fields) (void)hipFree(c->d_fields);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evf) (void)hipEventDestroy(c->evf);
    if (c->evj) (void)hipEventDestroy(c->evj);
    if (c->s) (void)hipStreamDestroy(c->s);
    if (c->s2) (void)hipStreamDestroy(c->s2);
    delete c;
}

// Message of the last context-less failure (bgk_create, bgk_footprint) on
// this thread; bgk_last_error(NULL) returns it.
inline thread_local std::string t_param_err;

inline int param_fail(const char *why)
{
    t_param_err = why;
    return BGK_E_INVALID;
}

// Largest plane the hot kernels address: they form each node's byte offset
// within a plane in 32 bits (one VGPR per node), so a plane of (m+2) rows
// of `pitch` elements must stay below 4 GiB.
constexpr uint64_t PLANE_LIMIT = (uint64_t)1 << 32;

// nullptr if the parameters describe a solve this library runs, else why not
inline const char *param_error(const bgk_params_t *p)
{
    if (p->lx < 2 || p->ly < 2 || p->lx > (1 << 20) || p->ly > (1 << 20)) return "lx, ly must be in 2..2^20";
    if (p->bc < 0 || p->bc > 6) return "unknown boundary variant";
    if (p->prec < 0 || p->prec > 4) return "unknown precision";
    if (p->impl != 0 && p->impl != 1) return "unknown implementation";
    if (p->bc == BGK_BC_STEP && p->lx < 20) return "STEP needs lx >= 20 (stepx = l/20 >= 1)";
    if (!(p->svisc > 0.0)) return "svisc must be positive";
    if (p->form != 0 && p->form != 1) return "form must be 0 or 1";
    if (p->rounding != 0 && p->rounding != 1) return "rounding must be 0 or 1";
    if (p->tblock < 0 || p->tblock > 2) return "tblock must be 0, 1 or 2";
    if ((p->les != 0 && p->les != 1) || !(p->cte_s >= 0.0)) return "les must be 0 or 1, cte_s >= 0";
    // TRICK1 is the driven cavity's and, as in the reference, square only
    if (p->trick1 != 0 && (p->trick1 != 1 || p->bc != BGK_BC_DRIVEN || p->lx != p->ly))
        return "trick1 needs the square driven cavity";
    // FORCING_Y changes the CHANNEL build's forcing term only
    if (p->forcing_y != 0 && (p->forcing_y != 1 || p->bc != BGK_BC_CHANNEL)) return "forcing_y needs the channel";
    if (p->obstacle < 0 || p->obstacle > 4) return "obstacle must be 0..4";
    if (p->noshift != 0 && (p->noshift != 1 || p->form != 0 || p->bc == BGK_BC_OPEN))
        return "noshift needs form 0 and a reference boundary variant";
    if (p->prec == BGK_PREC_FP16 || p->prec == BGK_PREC_BF16) {
        int ex;
        const double sc = p->scale > 0.0 ? p->scale : 1.0;
        if (frexp(sc, &ex) != 0.5) return "scale must be a power of two (exact scaling)";
    }
    int es, A, off, pitch;
    size_t pe;
    layout_of(p, &es, &A, &off, &pitch, &pe);
    if ((uint64_t)(p->ly + 2) * (uint64_t)pitch * (uint64_t)es >= PLANE_LIMIT)
        return "lattice too large: a population plane ((ly+2) rows x pitch x element size) must stay below 4 GiB";
    return nullptr;
}


}  // namespace bgk_host

struct bgk_batch {
    int dev;
    hipStream_t s;
    size_t budget, used;
    struct Slot {
        bgk_ctx *ctx;
        long long remaining, done;
        size_t bytes;
        int tw_log2, th;
        bool held;  // skipped by bgk_batch_run
    };
    std::vector<Slot> slots;
    bgk::Member *d_members, *h_members;
    int2 *d_tiles, *h_tiles;  // task table: (member, first tile)
    int *d_ntiles, *h_ntiles;  // per group: task counts, then first tasks
    size_t tiles_cap;
    std::map<long long, hipGraphExec_t> graphs;  // key: (chunk, group, full grid)
    std::map<long long, int> uses;
    // The groups' chains run on NQ streams made one after another at
    // creation; each chunk assigns the groups to them by work.  With the
    // batch's own stream they take the four hardware queues a process gets
    // by default, one each: on the mixed stream 3 gave 68.6 k MLUPS run
    // after run, 2 gave 63 k, 1 gave 49 k, and 4 (two streams then share a
    // queue) 61-72 k depending on which.
    static constexpr int NQ = 3;
    hipStream_t qs[NQ];
    hipEvent_t qev[NQ];
    int cus, grid_tb[3][2][6];
    int grid;
    hipEvent_t ev0, ev1;
    double last_ms;
    std::string err;
};

namespace bgk_host {


inline int bfail(bgk_batch_t *b, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

#define BHIPCHK(b, expr)                                                                           \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return bfail((b), BGK_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// tile codes: 1 = 64x4, 2 = 128x2, 3 = 256x1, 4 = 32x8
// tiles of a member: (1 << tw_log2) columns -- node pairs for the 16-bit
// kinds -- by th rows
inline int member_tiles_x(const bgk_ctx *c, int tw_log2)
{
    const int w = (1 << tw_log2) * (c->esize == 2 ? 2 : 1);
    return (c->l + w - 1) / w;
}
inline size_t member_tiles(const bgk_ctx *c, int tw_log2, int th)
{
    return (size_t)member_tiles_x(c, tw_log2) * (size_t)((c->m + th - 1) / th);
}
// A member of at least SOLO_NODES nodes (2048^2) keeps the chip busy by
// itself for a whole step: it advances with bgk_run's own kernels (the
// two-step kernel where that is faster: 122 k against 68 k MLUPS at 4096^2
// FP32) instead of through the task table.  Smaller members run faster
// together in the batch kernels.  Both paths give the same bits.
constexpr size_t SOLO_NODES = (size_t)1 << 22;
// tasks up to which a group's chain launches SMALL_GROUP blocks
constexpr size_t SMALL_GROUP = 512;
inline bool member_solo(const bgk_ctx *c) { return (size_t)c->l * (size_t)c->m >= SOLO_NODES; }

// A member advances two steps per launch (k_batch_tb) in a chunk of even
// length that it runs to the end, with bgk_run's rule for the kernel:
// tblock 0 or 2, every storage kind; tblock 1 never; otherwise one step per
// launch.  (The full serving stream with its 16-bit members in two-step
// groups: 111.8 k MLUPS; in one-step groups: 108.5 k.)
inline bool member_tb(const bgk_ctx *c, long long remaining, int chunk)
{
    const bool two = c->p.tblock != 1;
    return two && (chunk & 1) == 0 && remaining >= chunk;
}
// two-step strip blocks of a member, h rows high
inline int member_strips(const bgk_ctx *c) { return (c->l + bgk::TB_TW - 1) / bgk::TB_TW; }
inline size_t member_tb_tiles(const bgk_ctx *c, int h)
{
    return (size_t)member_strips(c) * (size_t)((c->m + h - 1) / h);
}
// entries of the batch task table for a member: one per strip block for
// the two-step kernel, one per TASK_TILES tiles otherwise
inline size_t member_tasks(const bgk_ctx *c, int tw_log2, int th, bool tb, int h)
{
    if (member_solo(c)) return 0;
    if (tb) return member_tb_tiles(c, h);
    return (member_tiles(c, tw_log2, th) + bgk::TASK_TILES - 1) / bgk::TASK_TILES;
}
// Strip height of the two-step members of a chunk: the strip blocks of all
// of them make whole rounds of the blocks the device holds (as k_fused_tb's
// height does for one lattice), in as few rounds as strips of at most
// TB_BATCH_H rows allow, and at least TB_HMIN rows.
inline int batch_tb_h(size_t strip_rows, int slots)
{
    if (slots < 1) slots = 1;
    const size_t per_round = (size_t)slots * bgk::TB_BATCH_H;
    const size_t rounds = strip_rows > per_round ? (strip_rows + per_round - 1) / per_round : 1;
    size_t h = (strip_rows + rounds * slots - 1) / (rounds * slots);
    if (h < (size_t)TB_HMIN) h = TB_HMIN;
    if (h > (size_t)bgk::TB_BATCH_H) h = bgk::TB_BATCH_H;
    return (int)h;
}
inline int member_kind(const bgk_ctx *c)
{
    return c->p.prec == BGK_PREC_FP16 ? bgk::MK_F16 : (c->p.prec == BGK_PREC_BF16 ? bgk::MK_BF16 : bgk::MK_F32);
}
inline int member_group(const bgk_ctx *c, bool tb)
{
    int v = 0;
    while (v < 5 && bgk::BATCH_BCS[v] != c->p.bc) v++;
    return bgk::group_of((tb ? bgk::TWO_STEP : 0) + member_kind(c), fl_of(c), v);
}

// one step of a group's members
template <class S, int FL>
void launch_batch_sf(bgk_batch_t *b, int g, int nb, int t, hipStream_t s)
{
    const dim3 grid(nb), blk(256);
    const bgk::Member *dm = b->d_members;
    const int2 *dt = b->d_tiles;
    const int *meta = b->d_ntiles;
    switch (bgk::BATCH_BCS[bgk::group_variant(g)]) {
    case bgk::BK_PERIODIC: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_PERIODIC, FL>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_CHANNEL: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_CHANNEL, FL>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_INFLOW: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_INFLOW, FL>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_COUETTE: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_COUETTE, FL>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_OPEN: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_OPEN, FL>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    default: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_DRIVEN, FL>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    }
}
template <class S>
void launch_batch_s(bgk_batch_t *b, int g, int nb, int t, hipStream_t s)
{
    if (bgk::group_fl(g))
        launch_batch_sf<S, 1>(b, g, nb, t, s);
    else
        launch_batch_sf<S, 0>(b, g, nb, t, s);
}
template <class S, int FL>
void launch_batch_tbf(bgk_batch_t *b, int g, int nb, int t2, hipStream_t s)
{
    const dim3 grid(nb), blk(bgk::TB_TW);
    const bgk::Member *dm = b->d_members;
    const int2 *dt = b->d_tiles;
    const int *meta = b->d_ntiles;
    switch (bgk::BATCH_BCS[bgk::group_variant(g)]) {
    case bgk::BK_PERIODIC: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_PERIODIC, FL>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_CHANNEL: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_CHANNEL, FL>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_INFLOW: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_INFLOW, FL>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_COUETTE: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_COUETTE, FL>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_OPEN: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_OPEN, FL>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    default: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_DRIVEN, FL>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    }
}
template <class S>
void launch_batch_tb(bgk_batch_t *b, int g, int nb, int t2, hipStream_t s)
{
    if (bgk::group_fl(g))
        launch_batch_tbf<S, 1>(b, g, nb, t2, s);
    else
        launch_batch_tbf<S, 0>(b, g, nb, t2, s);
}
// launch t of a group's chain (a two-step group's chain has chunk / 2)
inline void launch_batch(bgk_batch_t *b, int g, int nb, int t, hipStream_t s)
{
    switch (bgk::group_slot(g)) {
    case bgk::MK_F16: launch_batch_s<_Float16>(b, g, nb, t, s); break;
    case bgk::MK_BF16: launch_batch_s<bgk::bf16_t>(b, g, nb, t, s); break;
    case bgk::TWO_STEP + bgk::MK_F32: launch_batch_tb<float>(b, g, nb, t, s); break;
    case bgk::TWO_STEP + bgk::MK_F16: launch_batch_tb<_Float16>(b, g, nb, t, s); break;
    case bgk::TWO_STEP + bgk::MK_BF16: launch_batch_tb<bgk::bf16_t>(b, g, nb, t, s); break;
    default: launch_batch_s<float>(b, g, nb, t, s); break;
    }
}
// blocks of the two-step batch kernel the device holds at once: its grid,
// since a persistent block that starts late would run its whole share late
template <class S, int FL>
int occupancy_tbf(int v, int *per_cu)
{
    switch (bgk::BATCH_BCS[v]) {
    case bgk::BK_PERIODIC: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_PERIODIC, FL>, bgk::TB_TW, 0);
    case bgk::BK_CHANNEL: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_CHANNEL, FL>, bgk::TB_TW, 0);
    case bgk::BK_INFLOW: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_INFLOW, FL>, bgk::TB_TW, 0);
    case bgk::BK_COUETTE: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_COUETTE, FL>, bgk::TB_TW, 0);
    case bgk::BK_OPEN: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_OPEN, FL>, bgk::TB_TW, 0);
    default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_DRIVEN, FL>, bgk::TB_TW, 0);
    }
}
template <class S>
int occupancy_tb(int v, int fl, int *per_cu)
{
    return fl ? occupancy_tbf<S, 1>(v, per_cu) : occupancy_tbf<S, 0>(v, per_cu);
}
// blocks of the two-step batch kernel of storage kind `kind`, variant v,
// that the device holds at once: its grid, since a persistent block that
// starts late would run its whole share late
inline int grid_tb(bgk_batch_t *b, int kind, int fl, int v)
{
    if (b->grid_tb[kind][fl][v] > 0) return b->grid_tb[kind][fl][v];
    int per_cu = 0;
    const int e = kind == bgk::MK_F16 ? occupancy_tb<_Float16>(v, fl, &per_cu)
                  : kind == bgk::MK_BF16 ? occupancy_tb<bgk::bf16_t>(v, fl, &per_cu)
                                         : occupancy_tb<float>(v, fl, &per_cu);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    if (per_cu > TB_MAX_PER_CU) per_cu = TB_MAX_PER_CU;
    b->grid_tb[kind][fl][v] = per_cu * b->cus;
    return b->grid_tb[kind][fl][v];
}

inline bool tile_shape(int code, int *tw_log2, int *th)
{
    switch (code) {
    case 0:
    case 1: *tw_log2 = 6; *th = 4; return true;
    case 2: *tw_log2 = 7; *th = 2; return true;
    case 3: *tw_log2 = 8; *th = 1; return true;
    case 4: *tw_log2 = 5; *th = 8; return true;
    default: return false;
    }
}

inline void drop_graphs(bgk_batch_t *b)
{
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

inline int ensure_tiles(bgk_batch_t *b, size_t n)
{
    if (n <= b->tiles_cap) return BGK_OK;
    size_t cap = b->tiles_cap ? b->tiles_cap : 4096;
    while (cap < n) cap *= 2;
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->h_tiles) (void)hipHostFree(b->h_tiles);
    b->d_tiles = nullptr;
    b->h_tiles = nullptr;
    b->tiles_cap = 0;
    BHIPCHK(b, hipMalloc(&b->d_tiles, sizeof(int2) * cap));
    BHIPCHK(b, hipHostMalloc(&b->h_tiles, sizeof(int2) * cap));
    b->tiles_cap = cap;
    drop_graphs(b);  // the graphs hold the old table address
    return BGK_OK;
}

// Everything a batch holds, whatever part of bgk_batch_create succeeded:
// handles and buffers not yet made are null (the batch is value-initialised).
inline void batch_release(bgk_batch_t *b)
{
    if (b->s) (void)hipStreamSynchronize(b->s);
    for (int q = 0; q < bgk_batch_t::NQ; q++)
        if (b->qs[q]) (void)hipStreamSynchronize(b->qs[q]);
    drop_graphs(b);
    for (auto &sl : b->slots)
        if (sl.ctx) bgk_destroy(sl.ctx);
    if (b->d_members) (void)hipFree(b->d_members);
    if (b->h_members) (void)hipHostFree(b->h_members);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->h_tiles) (void)hipHostFree(b->h_tiles);
    if (b->d_ntiles) (void)hipFree(b->d_ntiles);
    if (b->h_ntiles) (void)hipHostFree(b->h_ntiles);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (int q = 0; q < bgk_batch_t::NQ; q++) {
        if (b->qev[q]) (void)hipEventDestroy(b->qev[q]);
        if (b->qs[q]) (void)hipStreamDestroy(b->qs[q]);
    }
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
}


}  // namespace bgk_host

// instantiated in bgk_kind_*.hip
namespace bgk_host {
extern template int run_t<float, float>(bgk_ctx *, int);
extern template int enqueue_t<float, float>(bgk_ctx *, int);
extern template int run_t<double, double>(bgk_ctx *, int);
extern template int enqueue_t<double, double>(bgk_ctx *, int);
extern template int run_t<_Float16, float>(bgk_ctx *, int);
extern template int enqueue_t<_Float16, float>(bgk_ctx *, int);
extern template int run_t<float, double>(bgk_ctx *, int);
extern template int enqueue_t<float, double>(bgk_ctx *, int);
extern template int run_t<bgk::bf16_t, float>(bgk_ctx *, int);
extern template int enqueue_t<bgk::bf16_t, float>(bgk_ctx *, int);
extern template void launch_batch_s<float>(bgk_batch_t *, int, int, int, hipStream_t);
extern template void launch_batch_tb<float>(bgk_batch_t *, int, int, int, hipStream_t);
extern template int occupancy_tb<float>(int, int, int *);
extern template void launch_batch_s<_Float16>(bgk_batch_t *, int, int, int, hipStream_t);
extern template void launch_batch_tb<_Float16>(bgk_batch_t *, int, int, int, hipStream_t);
extern template int occupancy_tb<_Float16>(int, int, int *);
extern template void launch_batch_s<bgk::bf16_t>(bgk_batch_t *, int, int, int, hipStream_t);
extern template void launch_batch_tb<bgk::bf16_t>(bgk_batch_t *, int, int, int, hipStream_t);
extern template int occupancy_tb<bgk::bf16_t>(int, int, int *);
}  // namespace bgk_host
