// bgk_api.hip -- C ABI and runtime of libbgk2d_hip.so (see include/bgk2d.h).
//
// A context owns one HBM allocation (18 pitched population planes plus
// the obstacle mask and reduction scratch), one HIP stream and the HIP
// graphs it replays.  Host code derives the reference's constants in the
// compute kind (hencol, build_bcond) and launches the kernels of
// bgk_kernels.hpp in the reference's phase order.
#include "../../include/bgk2d.h"
#include "bgk_kernels.hpp"
#include "bgk_serve.hpp"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
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

namespace {

int fail(bgk_ctx *c, int code, const char *fmt, ...)
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

bool is_fp16(int prec) { return prec == BGK_PREC_FP16; }
bool is_bf16(int prec) { return prec == BGK_PREC_BF16; }

int esize_of(int prec)
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
    // crho*p*(cte1+cx) in the compute kind; each stored copy is rounded to
    // the storage kind where it is stored (x = 0, row j)
    for (int q = 0; q < 3; q++) c->inflow[q] = (double)((C)1.0 * wq[q] * ((C)0.0 + cxq[q]));
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
    return P;
}

Lat lat(const bgk_ctx *c)
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

dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

const dim3 B2(64, 4);

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

bool foldable(const bgk_ctx *c) { return c->p.impl == BGK_IMPL_FUSED && c->p.bc != BGK_BC_STEP; }

// the settings of the FAST kernel instances (16-bit storage only)
template <class C>
bool fast_flags(const Phys<C> &P) { return P.form == 1 && P.rounding == 1; }

template <class S, class C, int BC, bool OBS>
void launch_fused_t(bgk_ctx *c)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    if constexpr (sizeof(S) == 2) {
        // 16-bit storage: two nodes per thread
        if (fast_flags(P))
            hipLaunchKernelGGL((k_fused2<S, C, BC, OBS, true>), grid2((c->l + 1) / 2, c->m, B2), B2, 0, c->s, L, P);
        else
            hipLaunchKernelGGL((k_fused2<S, C, BC, OBS>), grid2((c->l + 1) / 2, c->m, B2), B2, 0, c->s, L, P);
    } else {
        hipLaunchKernelGGL((k_fused<S, C, BC, OBS>), grid2(c->l, c->m, B2), B2, 0, c->s, L, P);
    }
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

// Two FUSED steps per pass over HBM (k_fused_tb): the default for fp32 and
// fp64 storage on lattices large enough to fill the device with strips of
// at least TB_HMIN rows, where it beats one step per pass; for FP16/BF16
// storage the paired one-step kernel is as fast, so it stays the default
// there.  Needs size_tb first.
constexpr int TB_HMIN = 8;
bool use_tb(const bgk_ctx *c)
{
    if (!foldable(c)) return false;
    if (c->p.tblock == 2) return true;
    return c->p.tblock == 0 && c->esize >= 4 && c->tb_fill;
}

// Strip blocks of height H that must apply the obstacle rule: those whose
// updated nodes (their strip, its rim columns, one row above and below)
// reach the obstacle's bounding box grown by one node, where fix_pulls
// applies it -- with one more block on each side as a margin.  Along a
// periodic direction the first and last blocks also update wrapped nodes
// of the far edge (their rim or halo row): when the grown box reaches an
// edge (the obstacle touches the second or second-to-last row or column),
// every block along that direction is taken.
// (first strip, last strip, first row, last row) of blocks; x > y: none.
int4 tb_box(const bgk_ctx *c, int H, bool xper, bool yper)
{
    if (!(c->p.obstacle && c->nobs > 0)) return make_int4(0, -1, 0, -1);
    const int strips = (c->l + TB_TW - 1) / TB_TW, rows = (c->m + H - 1) / H;
    int4 b = make_int4(max(0, (c->imin - 3) / TB_TW - 1), min(strips - 1, (c->imax + 1) / TB_TW + 1),
                       max(0, (c->jmin - 3) / H - 1), min(rows - 1, (c->jmax + 1) / H + 1));
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
    int per_cu = 0, cus = 0;
    // with an obstacle the grid is still that of the obstacle-free kernel:
    // the obstacle's few strip blocks run beside it (launch_tb_t); leaving
    // them room in the round measured slower (91 k against 94 k MLUPS)
    if (sizeof(S) == 2 && fast_flags(phys<C>(c)))
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_tb<S, C, BC, false, sizeof(S) == 2>, 256, 0));
    else
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_tb<S, C, BC, false>, 256, 0));
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->dev));
    if (per_cu < 1) per_cu = 1;
    if (cus < 1) cus = 1;
    const long long slots = (long long)per_cu * cus;
    const long long strips = (c->l + TB_TW - 1) / TB_TW;
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

template <class S, class C, int BC, bool OBS>
int launch_tb_t(bgk_ctx *c)
{
    if (c->tb_h == 0) {
        int r = size_tb<S, C, BC, OBS>(c);
        if (r != BGK_OK) return r;
    }
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    const bool fast = sizeof(S) == 2 && fast_flags(P);
    const int H = c->tb_h;
    const int strips = (c->l + TB_TW - 1) / TB_TW, rows = (c->m + H - 1) / H;
    const int rows8 = (rows + NXCD - 1) / NXCD * NXCD;
    // strip blocks with the obstacle rule
    const int4 box = OBS ? tb_box(c, H, Per<BC>::x, Per<BC>::y) : make_int4(0, -1, 0, -1);
    if (OBS && box.y >= box.x) {
        // the obstacle's blocks on a second stream, beside the others
        // the rows of the skipped blocks, in blocks of a quarter of H
        const int nbx = box.y - box.x + 1, r0 = box.z * H + 1, r1 = min((box.w + 1) * H, c->m);
        const int hb = max(TB_HMIN, H / 4), nsub = (r1 - r0 + hb) / hb;
        HIPCHK(c, hipEventRecord(c->evf, c->s));
        HIPCHK(c, hipStreamWaitEvent(c->s2, c->evf, 0));
        if (fast)
            hipLaunchKernelGGL((k_fused_tb_box<S, C, BC, sizeof(S) == 2>), dim3(nbx * nsub), dim3(256), 0, c->s2, L, P,
                               box.x, nbx, r0, r1, hb);
        else
            hipLaunchKernelGGL((k_fused_tb_box<S, C, BC>), dim3(nbx * nsub), dim3(256), 0, c->s2, L, P, box.x, nbx, r0,
                               r1, hb);
        HIPCHK(c, hipEventRecord(c->evj, c->s2));
    }
    if (fast)
        hipLaunchKernelGGL((k_fused_tb<S, C, BC, false, sizeof(S) == 2>), dim3(strips * rows8), dim3(256), 0, c->s, L, P,
                           H, strips, rows, box);
    else
        hipLaunchKernelGGL((k_fused_tb<S, C, BC, false>), dim3(strips * rows8), dim3(256), 0, c->s, L, P, H, strips,
                           rows, box);
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
    int r = launch_steps<S, C>(c, GRAPH_STEPS);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->s, &g);
    c->cur = cur0;
    c->rcur = rcur0;
    if (r != BGK_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(c, BGK_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
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
void host_sincos<float>(float arg, float *s, float *c) { sincosf(arg, s, c); }
template <>
void host_sincos<double>(double arg, double *s, double *c) { sincos(arg, s, c); }

template <class S, class C>
int init_t(bgk_ctx *c, double amp, int init_v)
{
    const bool tgv = c->p.bc == BGK_BC_PERIODIC;
    const int nx = c->l + 2, ny = c->m + 2;
    std::vector<C> tsx(nx), tcx(nx), tsy(ny), tcy(ny);
    // SRC/init.F90:88-111: pi is a default-real literal in kind mykind,
    // x = (real(i)-0.5d0)/real(l) evaluated in double, sin/cos in mykind
    const C pi = (C)(float)3.141592653589793238462643383279;
    for (int i = 0; i < nx; i++) {
        C x = (C)(((double)(C)i - 0.5) / (double)(C)c->l);
        host_sincos<C>((C)2 * pi * x, &tsx[i], &tcx[i]);
    }
    for (int j = 0; j < ny; j++) {
        C y = (C)(((double)(C)j - 0.5) / (double)(C)c->m);
        host_sincos<C>((C)2 * pi * y, &tsy[j], &tcy[j]);
    }
    C *d = nullptr;
    HIPCHK(c, hipMalloc(&d, sizeof(C) * 2 * (nx + ny)));
    HIPCHK(c, hipMemcpyAsync(d, tsx.data(), sizeof(C) * nx, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(d + nx, tcx.data(), sizeof(C) * nx, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(d + 2 * nx, tsy.data(), sizeof(C) * ny, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(d + 2 * nx + ny, tcy.data(), sizeof(C) * ny, hipMemcpyHostToDevice, c->s));
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    hipLaunchKernelGGL((k_init<S, C>), grid2(nx, ny, B2), B2, 0, c->s, L, P, d, d + nx, d + 2 * nx,
                       d + 2 * nx + ny, amp, tgv ? (init_v == 1 ? 2 : 1) : 0);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->s));
    HIPCHK(c, hipFree(d));
    return BGK_OK;
}

template <class S, class C>
int rows_t(bgk_ctx *c, bool final_)
{
    Lat L = lat(c);
    Phys<C> P = phys<C>(c);
    hipLaunchKernelGGL((k_rows<S, C>), dim3(c->m), dim3(ROWS_T), 0, c->s, L, P, c->d_rows);
    if (final_) hipLaunchKernelGGL(k_rows_final, dim3(1), dim3(ROWS_T), 0, c->s, c->d_rows, c->m, c->d_red);
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
void layout_of(const bgk_params_t *p, int *esize, int *A, int *off, int *pitch, size_t *plane_elems)
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
size_t footprint_of(const bgk_params_t *p)
{
    int es, A, off, pitch;
    size_t pe;
    layout_of(p, &es, &A, &off, &pitch, &pe);
    size_t obs_bytes = ((size_t)p->lx * p->ly + 255) / 256 * 256;
    size_t rows_bytes = sizeof(double) * ROW_VALS * (size_t)p->ly;
    size_t red_bytes = sizeof(double) * 3 * (size_t)(p->lx > p->ly ? p->lx : p->ly) + 64;
    return pe * (size_t)es * 18 + obs_bytes + rows_bytes + red_bytes;
}

int check_ctx(bgk_ctx *c)
{
    if (!c) return BGK_E_INVALID;
    hipError_t e = hipSetDevice(c->dev);
    if (e != hipSuccess) return fail(c, BGK_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    return BGK_OK;
}

int fetch_red(bgk_ctx *c, int n)
{
    HIPCHK(c, hipMemcpyAsync(c->h_red, c->d_red, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return BGK_OK;
}

// Everything a context holds, whatever part of bgk_create succeeded:
// handles and buffers not yet made are null (the context is value-initialised).
void ctx_release(bgk_ctx *c)
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
thread_local std::string t_param_err;

int param_fail(const char *why)
{
    t_param_err = why;
    return BGK_E_INVALID;
}

// Largest plane the hot kernels address: they form each node's byte offset
// within a plane in 32 bits (one VGPR per node), so a plane of (m+2) rows
// of `pitch` elements must stay below 4 GiB.
constexpr uint64_t PLANE_LIMIT = (uint64_t)1 << 32;

// nullptr if the parameters describe a solve this library runs, else why not
const char *param_error(const bgk_params_t *p)
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
    if (p->obstacle < 0 || p->obstacle > 3) return "obstacle must be 0..3";
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

}  // namespace

// =====================================================================
extern "C" {

int bgk_version(void) { return BGK_ABI_VERSION; }

const char *bgk_last_error(const bgk_ctx_t *c) { return c ? c->err.c_str() : t_param_err.c_str(); }

int bgk_create(const bgk_params_t *p, bgk_ctx_t **out)
{
    if (!p || !out) return param_fail("null argument");
    *out = nullptr;
    if (const char *why = param_error(p)) return param_fail(why);
    bgk_ctx *c = new bgk_ctx();
    c->p = *p;
    c->l = p->lx;
    c->m = p->ly;
    c->dev = p->device;
    c->scale = (p->scale > 0.0) ? p->scale : 1.0;
    if (!is_fp16(p->prec) && !is_bf16(p->prec)) c->scale = 1.0;
    layout_of(p, &c->esize, &c->A, &c->off, &c->pitch, &c->plane_elems);
    c->imin = c->l;
    c->jmin = c->m;
    c->imax = 0;
    c->jmax = 0;
    c->nobs = 0;
    c->cur = 0;
    c->rcur = 0;
    c->tb_h = 0;
    c->tb_fill = false;
    c->last_ms = 0.0;
    c->inited = false;
    c->stage_bytes = 0;
    c->d_stage = nullptr;
    c->d_fields = nullptr;
    if (hipSetDevice(c->dev) != hipSuccess) {
        delete c;
        return param_fail("hipSetDevice failed");
    }
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
        hipEventCreate(&c->ev1) != hipSuccess || hipEventCreateWithFlags(&c->evf, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evj, hipEventDisableTiming) != hipSuccess) {
        ctx_release(c);
        return param_fail("stream or event creation failed");
    }
    size_t bytes = c->plane_elems * (size_t)c->esize * 18;
    size_t obs_bytes = ((size_t)c->l * c->m + 255) / 256 * 256;
    size_t rows_bytes = sizeof(double) * ROW_VALS * (size_t)c->m;
    size_t red_bytes = sizeof(double) * 3 * (size_t)(c->l > c->m ? c->l : c->m) + 64;
    if (hipMalloc(&c->mem, bytes) != hipSuccess || hipMalloc(&c->d_obs, obs_bytes) != hipSuccess ||
        hipMalloc(&c->d_rows, rows_bytes) != hipSuccess || hipMalloc(&c->d_red, red_bytes) != hipSuccess ||
        hipHostMalloc(&c->h_red, red_bytes) != hipSuccess) {
        ctx_release(c);
        t_param_err = "device or pinned host allocation failed";
        return BGK_E_NOMEM;
    }
    for (int q = 0; q < 18; q++) c->plane[q] = (char *)c->mem + (size_t)q * c->plane_elems * c->esize;
    hipMemsetAsync(c->d_obs, 0, obs_bytes, c->s);
    c->h_obs.assign((size_t)c->l * c->m, 0);
    // alloca: huge(mykind) = 2147483647 converted to the storage kind
    int r;
    switch (p->prec) {
    case BGK_PREC_DOUBLE: r = fill_t<double>(c, 2147483647.0); break;
    case BGK_PREC_FP16: r = fill_t<_Float16>(c, (_Float16)2147483647.0f); break;
    case BGK_PREC_BF16: {
        bgk::bf16_t v;
        v.bits = 0x4f00;  // 2^31
        r = fill_t<bgk::bf16_t>(c, v);
    } break;
    default: r = fill_t<float>(c, (float)2147483647); break;
    }
    if (r != BGK_OK) {
        bgk_destroy(c);
        return r;
    }
    if (p->prec == BGK_PREC_DOUBLE || p->prec == BGK_PREC_MIXED2)
        derive<double>(c);
    else
        derive<float>(c);
    hipStreamSynchronize(c->s);
    *out = c;
    return BGK_OK;
}

int bgk_destroy(bgk_ctx_t *c)
{
    if (!c) return BGK_E_INVALID;
    (void)hipSetDevice(c->dev);
    ctx_release(c);
    return BGK_OK;
}

int bgk_constants(const bgk_ctx_t *c, double out[4])
{
    if (!c || !out) return BGK_E_INVALID;
    out[0] = c->omega;
    out[1] = c->fgrad;
    out[2] = c->u00;
    out[3] = c->force;
    return BGK_OK;
}

static int set_mask(bgk_ctx *c, const int32_t *obs)
{
    int imin = c->l, jmin = c->m, imax = 0, jmax = 0, n = 0;
    std::vector<uint8_t> h((size_t)c->l * c->m);
    for (int j = 1; j <= c->m; j++)
        for (int i = 1; i <= c->l; i++) {
            size_t q = (size_t)(i - 1) + (size_t)(j - 1) * c->l;
            h[q] = obs[q] == 1 ? 1 : 0;
            if (h[q]) {
                n++;
                if (i < imin) imin = i;
                if (j < jmin) jmin = j;
                if (i > imax) imax = i;
                if (j > jmax) jmax = j;
            }
        }
    if (n > 0 && (imin < 2 || imax > c->l - 1 || jmin < 2 || jmax > c->m - 1))
        return fail(c, BGK_E_INVALID, "obstacle must not touch the first/last interior row or column");
    c->imin = imin;
    c->imax = imax;
    c->jmin = jmin;
    c->jmax = jmax;
    c->nobs = n;
    if (n == 0) c->p.obstacle = 0;             // a mask turns bcond_obs on, as the
    else if (c->p.obstacle == 0) c->p.obstacle = 1;  // OBSTACLE build; keeps a shape code
    c->tb_h = 0;                    // the two-step kernel is sized per instance
    c->h_obs.assign(obs, obs + (size_t)c->l * c->m);
    HIPCHK(c, hipMemcpyAsync(c->d_obs, h.data(), h.size(), hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    for (auto &g : c->graphs) hipGraphExecDestroy(g.second);
    c->graphs.clear();
    return BGK_OK;
}

int bgk_build_obs(bgk_ctx_t *c)
{
    int r = check_ctx(c);
    if (r) return r;
    // SRC/build.obs.F90:49-205, centred at (2l/5, m/2), comparisons in
    // mykind: obstacle 1 (and 0) the cylinder (icoord-i)^2+(jcoord-j)^2 <=
    // R*R (default branch), 2 the square of half side radius (SQUARE), 3 the
    // cylinder's rows m/2..m (HALF_CIRCLE)
    const int l = c->l, m = c->m, ic = 2 * l / 5, jc = m / 2, shape = c->p.obstacle;
    std::vector<int32_t> obs((size_t)l * m, 0);
    const bool dbl = c->p.prec == BGK_PREC_DOUBLE || c->p.prec == BGK_PREC_MIXED2;
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= l; i++) {
            long long d2i = (long long)(ic - i) * (ic - i) + (long long)(jc - j) * (jc - j);
            bool in;
            if (shape == 2) {
                // integer bounds against a real half size: exact for these sizes
                const double h = (double)c->p.radius;
                in = j >= jc - h && j <= jc + h && i >= ic - h && i <= ic + h;
            } else if (dbl) {
                double R = (double)c->p.radius * 1.0;
                in = (double)(int)d2i <= R * R;
            } else {
                float R = (float)c->p.radius * 1.0f;
                in = (float)(int)d2i <= R * R;
            }
            if (shape == 3 && j < m / 2) in = false;
            obs[(size_t)(i - 1) + (size_t)(j - 1) * l] = in ? 1 : 0;
        }
    return set_mask(c, obs.data());
}

int bgk_set_obstacle(bgk_ctx_t *c, const int32_t *obs)
{
    int r = check_ctx(c);
    if (r) return r;
    if (!obs) return fail(c, BGK_E_INVALID, "null mask");
    return set_mask(c, obs);
}

int bgk_get_obstacle(bgk_ctx_t *c, int32_t *obs, int32_t bbox[4])
{
    if (!c) return BGK_E_INVALID;
    if (obs) memcpy(obs, c->h_obs.data(), sizeof(int32_t) * c->h_obs.size());
    if (bbox) {
        bbox[0] = c->imin;
        bbox[1] = c->imax;
        bbox[2] = c->jmin;
        bbox[3] = c->jmax;
    }
    return BGK_OK;
}

int bgk_init(bgk_ctx_t *c, int32_t init_v, double amplitude)
{
    int r = check_ctx(c);
    if (r) return r;
    if (init_v < 0 || init_v > 5) return fail(c, BGK_E_INVALID, "Initial condition out of range[0,5] %d", init_v);
    r = DISPATCH(c, init_t, amplitude, (int)init_v);
    if (r == BGK_OK) c->inited = true;
    return r;
}

int bgk_boundaries(bgk_ctx_t *c)
{
    int r = check_ctx(c);
    if (r) return r;
    if (!c->inited) return fail(c, BGK_E_STATE, "bgk_init not called");
    return DISPATCH(c, launch_boundaries);
}

int bgk_propagation(bgk_ctx_t *c)
{
    int r = check_ctx(c);
    if (r) return r;
    if (!c->inited) return fail(c, BGK_E_STATE, "bgk_init not called");
    return DISPATCH(c, launch_propagation);
}

int bgk_collision(bgk_ctx_t *c, int32_t itime)
{
    (void)itime;  // only the LES/debug builds of the reference use it
    int r = check_ctx(c);
    if (r) return r;
    if (!c->inited) return fail(c, BGK_E_STATE, "bgk_init not called");
    return DISPATCH(c, launch_collision);
}

int bgk_run(bgk_ctx_t *c, int32_t nsteps)
{
    int r = check_ctx(c);
    if (r) return r;
    if (!c->inited) return fail(c, BGK_E_STATE, "bgk_init not called");
    if (nsteps < 0) return fail(c, BGK_E_INVALID, "negative step count");
    return DISPATCH(c, run_t, nsteps);
}

int bgk_sync(bgk_ctx_t *c)
{
    int r = check_ctx(c);
    if (r) return r;
    HIPCHK(c, hipStreamSynchronize(c->s));
    return BGK_OK;
}

int bgk_diagno(bgk_ctx_t *c, double out[4])
{
    int r = check_ctx(c);
    if (r) return r;
    if (!out) return BGK_E_INVALID;
    r = DISPATCH(c, rows_t, true);
    if (r == BGK_OK) r = fetch_red(c, 6);
    if (r) return r;
    const double *h = c->h_red;
    out[0] = (h[1] / c->l) / c->m;
    out[1] = (h[2] / c->l) / c->m;
    out[2] = (h[0] / c->l) / c->m;
    out[3] = (h[3] / c->l) / c->m;
    return BGK_OK;
}

int bgk_check(bgk_ctx_t *c, double umax)
{
    int r = check_ctx(c);
    if (r) return r;
    r = DISPATCH(c, rows_t, true);
    if (r == BGK_OK) r = fetch_red(c, 6);
    if (r) return r;
    if (c->h_red[5] > 0.0) return fail(c, BGK_E_DIVERGED, "non-finite populations");
    if (c->h_red[4] > umax * umax) return fail(c, BGK_E_DIVERGED, "|u| = %g > %g", sqrt(c->h_red[4]), umax);
    return BGK_OK;
}

int bgk_varm(bgk_ctx_t *c, double *u, double *v, double *rho)
{
    int r = check_ctx(c);
    if (r) return r;
    r = DISPATCH(c, rows_t, false);
    if (r) return r;
    std::vector<double> rows((size_t)ROW_VALS * c->m);
    HIPCHK(c, hipMemcpyAsync(rows.data(), c->d_rows, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    for (int j = 0; j < c->m; j++) {
        const double *q = &rows[(size_t)j * ROW_VALS];
        if (rho) rho[j] = q[4] / c->l;
        if (u) u[j] = q[5] / c->l;
        if (v) v[j] = q[6] / c->l;
    }
    return BGK_OK;
}

static int line_common(bgk_ctx *c, int mode, int c0, int c1, int n, double *u, double *v, double *rho)
{
    int r = check_ctx(c);
    if (r) return r;
    r = DISPATCH(c, line_t, mode, c0, c1, n);
    if (r == BGK_OK) r = fetch_red(c, 3 * n);
    if (r) return r;
    for (int t = 0; t < n; t++) {
        if (u) u[t] = c->h_red[t];
        if (v) v[t] = c->h_red[n + t];
        if (rho) rho[t] = c->h_red[2 * n + t];
    }
    return BGK_OK;
}

int bgk_prof_i(bgk_ctx_t *c, int32_t jc, double *u, double *v, double *rho)
{
    if (!c || jc < 0 || jc > c->m + 1) return BGK_E_INVALID;
    return line_common(c, 0, jc, 0, c->l, u, v, rho);
}

int bgk_prof_j(bgk_ctx_t *c, int32_t ic, double *u, double *v, double *rho)
{
    if (!c || ic < 0 || ic > c->l + 1) return BGK_E_INVALID;
    return line_common(c, 1, ic, 0, c->m, u, v, rho);
}

int bgk_probe(bgk_ctx_t *c, int32_t i0, int32_t j0, double out[3])
{
    if (!c || !out || i0 < 0 || i0 > c->l + 1 || j0 < 0 || j0 > c->m + 1) return BGK_E_INVALID;
    return line_common(c, 2, i0, j0, 1, &out[0], &out[1], &out[2]);
}

int bgk_draglift(bgk_ctx_t *c, int32_t i0, int32_t i1, int32_t j0, int32_t j1, double out[2])
{
    if (!c || !out) return BGK_E_INVALID;
    int r = check_ctx(c);
    if (r) return r;
    // the box is clipped to the interior; outside it no node is solid
    i0 = i0 < 1 ? 1 : i0;
    j0 = j0 < 1 ? 1 : j0;
    i1 = i1 > c->l ? c->l : i1;
    j1 = j1 > c->m ? c->m : j1;
    if (i0 > i1 || j0 > j1) return fail(c, BGK_E_INVALID, "draglift: empty box");
    r = DISPATCH(c, draglift_t, i0, i1, j0, j1);
    if (r == BGK_OK) r = fetch_red(c, 2);
    if (r) return r;
    out[0] = c->h_red[0];
    out[1] = c->h_red[1];
    return BGK_OK;
}

int bgk_fields(bgk_ctx_t *c, float *u, float *v, float *rho)
{
    if (!c || !u || !v || !rho) return BGK_E_INVALID;
    int r = check_ctx(c);
    if (r) return r;
    const size_t n = (size_t)c->l * c->m;
    if (!c->d_fields) HIPCHK(c, hipMalloc((void **)&c->d_fields, 3 * n * sizeof(float)));
    float *d = c->d_fields;
    r = DISPATCH(c, fields_t, d);
    if (r == BGK_OK) {
        hipError_t e = hipMemcpyAsync(u, d, n * sizeof(float), hipMemcpyDeviceToHost, c->s);
        if (e == hipSuccess) e = hipMemcpyAsync(v, d + n, n * sizeof(float), hipMemcpyDeviceToHost, c->s);
        if (e == hipSuccess) e = hipMemcpyAsync(rho, d + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost, c->s);
        if (e != hipSuccess) r = fail(c, BGK_E_HIP, "bgk_fields copy: %s", hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(c->s);
    if (r == BGK_OK && e != hipSuccess) r = fail(c, BGK_E_HIP, "bgk_fields: %s", hipGetErrorString(e));
    return r;
}

int bgk_storage_bytes(const bgk_ctx_t *c) { return c ? c->esize : BGK_E_INVALID; }

static int ensure_stage(bgk_ctx *c)
{
    size_t need = (size_t)(c->l + 2) * (c->m + 2) * c->esize;
    if (c->stage_bytes >= need) return BGK_OK;
    if (c->d_stage) hipFree(c->d_stage);
    c->d_stage = nullptr;
    c->stage_bytes = 0;
    HIPCHK(c, hipMalloc(&c->d_stage, need));
    c->stage_bytes = need;
    return BGK_OK;
}

int bgk_download(bgk_ctx_t *c, int32_t k, void *host)
{
    int r = check_ctx(c);
    if (r) return r;
    if (k < 0 || k > 8 || !host) return fail(c, BGK_E_INVALID, "bad population index %d", k);
    if ((r = ensure_stage(c))) return r;
    switch (c->esize) {
    case 8: r = pack_t<double>(c, k, false); break;
    case 2: r = pack_t<uint16_t>(c, k, false); break;
    default: r = pack_t<float>(c, k, false); break;
    }
    if (r) return r;
    size_t n = (size_t)(c->l + 2) * (c->m + 2) * c->esize;
    HIPCHK(c, hipMemcpyAsync(host, c->d_stage, n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return BGK_OK;
}

int bgk_upload(bgk_ctx_t *c, int32_t k, const void *host)
{
    int r = check_ctx(c);
    if (r) return r;
    if (k < 0 || k > 8 || !host) return fail(c, BGK_E_INVALID, "bad population index %d", k);
    if ((r = ensure_stage(c))) return r;
    size_t n = (size_t)(c->l + 2) * (c->m + 2) * c->esize;
    HIPCHK(c, hipMemcpyAsync(c->d_stage, host, n, hipMemcpyHostToDevice, c->s));
    switch (c->esize) {
    case 8: r = pack_t<double>(c, k, true); break;
    case 2: r = pack_t<uint16_t>(c, k, true); break;
    default: r = pack_t<float>(c, k, true); break;
    }
    if (r) return r;
    HIPCHK(c, hipStreamSynchronize(c->s));
    c->inited = true;
    return BGK_OK;
}

void *bgk_stream(bgk_ctx_t *c) { return c ? (void *)c->s : nullptr; }

// ===================================================================== batch

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
    int cus, grid_tb[3][6];
    int grid;
    hipEvent_t ev0, ev1;
    double last_ms;
    std::string err;
};

namespace {

int bfail(bgk_batch_t *b, int code, const char *fmt, ...)
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
int member_tiles_x(const bgk_ctx *c, int tw_log2)
{
    const int w = (1 << tw_log2) * (c->esize == 2 ? 2 : 1);
    return (c->l + w - 1) / w;
}
size_t member_tiles(const bgk_ctx *c, int tw_log2, int th)
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
bool member_solo(const bgk_ctx *c) { return (size_t)c->l * (size_t)c->m >= SOLO_NODES; }

// A member advances two steps per launch (k_batch_tb) in a chunk of even
// length that it runs to the end, with bgk_run's rule for the kernel:
// tblock = 2 always, tblock = 0 for fp32 storage (for the 16-bit kinds the
// one-step groups, which overlap, served a mixed stream 5-10 % faster),
// tblock = 1 never; otherwise one step per launch.
bool member_tb(const bgk_ctx *c, long long remaining, int chunk)
{
    const bool two = c->p.tblock == 2 || (c->p.tblock == 0 && c->p.prec == BGK_PREC_SINGLE);
    return two && (chunk & 1) == 0 && remaining >= chunk;
}
// two-step strip blocks of a member, h rows high
int member_strips(const bgk_ctx *c) { return (c->l + bgk::TB_TW - 1) / bgk::TB_TW; }
size_t member_tb_tiles(const bgk_ctx *c, int h)
{
    return (size_t)member_strips(c) * (size_t)((c->m + h - 1) / h);
}
// entries of the batch task table for a member: one per strip block for
// the two-step kernel, one per TASK_TILES tiles otherwise
size_t member_tasks(const bgk_ctx *c, int tw_log2, int th, bool tb, int h)
{
    if (member_solo(c)) return 0;
    if (tb) return member_tb_tiles(c, h);
    return (member_tiles(c, tw_log2, th) + bgk::TASK_TILES - 1) / bgk::TASK_TILES;
}
// Strip height of the two-step members of a chunk: the strip blocks of all
// of them make whole rounds of the blocks the device holds (as k_fused_tb's
// height does for one lattice), in as few rounds as strips of at most
// TB_BATCH_H rows allow, and at least TB_HMIN rows.
int batch_tb_h(size_t strip_rows, int slots)
{
    if (slots < 1) slots = 1;
    const size_t per_round = (size_t)slots * bgk::TB_BATCH_H;
    const size_t rounds = strip_rows > per_round ? (strip_rows + per_round - 1) / per_round : 1;
    size_t h = (strip_rows + rounds * slots - 1) / (rounds * slots);
    if (h < (size_t)TB_HMIN) h = TB_HMIN;
    if (h > (size_t)bgk::TB_BATCH_H) h = bgk::TB_BATCH_H;
    return (int)h;
}
int member_kind(const bgk_ctx *c)
{
    return c->p.prec == BGK_PREC_FP16 ? bgk::MK_F16 : (c->p.prec == BGK_PREC_BF16 ? bgk::MK_BF16 : bgk::MK_F32);
}
int member_group(const bgk_ctx *c, bool tb)
{
    int v = 0;
    while (v < 5 && bgk::BATCH_BCS[v] != c->p.bc) v++;
    return ((tb ? bgk::TWO_STEP : 0) + member_kind(c)) * 6 + v;
}

// one step of a group's members
extern "C++" {
template <class S>
void launch_batch_s(bgk_batch_t *b, int g, int nb, int t, hipStream_t s)
{
    const dim3 grid(nb), blk(256);
    const bgk::Member *dm = b->d_members;
    const int2 *dt = b->d_tiles;
    const int *meta = b->d_ntiles;
    switch (bgk::BATCH_BCS[g % 6]) {
    case bgk::BK_PERIODIC: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_PERIODIC>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_CHANNEL: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_CHANNEL>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_INFLOW: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_INFLOW>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_COUETTE: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_COUETTE>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    case bgk::BK_OPEN: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_OPEN>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    default: hipLaunchKernelGGL((bgk::k_batch<S, bgk::BK_DRIVEN>), grid, blk, 0, s, dm, dt, meta, g, t); break;
    }
}
template <class S>
void launch_batch_tb(bgk_batch_t *b, int g, int nb, int t2, hipStream_t s)
{
    const dim3 grid(nb), blk(256);
    const bgk::Member *dm = b->d_members;
    const int2 *dt = b->d_tiles;
    const int *meta = b->d_ntiles;
    switch (bgk::BATCH_BCS[g % 6]) {
    case bgk::BK_PERIODIC: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_PERIODIC>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_CHANNEL: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_CHANNEL>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_INFLOW: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_INFLOW>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_COUETTE: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_COUETTE>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    case bgk::BK_OPEN: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_OPEN>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    default: hipLaunchKernelGGL((bgk::k_batch_tb<S, bgk::BK_DRIVEN>), grid, blk, 0, s, dm, dt, meta, g, t2); break;
    }
}
// launch t of a group's chain (a two-step group's chain has chunk / 2)
void launch_batch(bgk_batch_t *b, int g, int nb, int t, hipStream_t s)
{
    switch (g / 6) {
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
template <class S>
int occupancy_tb(int v, int *per_cu)
{
    switch (bgk::BATCH_BCS[v]) {
    case bgk::BK_PERIODIC: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_PERIODIC>, 256, 0);
    case bgk::BK_CHANNEL: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_CHANNEL>, 256, 0);
    case bgk::BK_INFLOW: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_INFLOW>, 256, 0);
    case bgk::BK_COUETTE: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_COUETTE>, 256, 0);
    case bgk::BK_OPEN: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_OPEN>, 256, 0);
    default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, bgk::k_batch_tb<S, bgk::BK_DRIVEN>, 256, 0);
    }
}
// blocks of the two-step batch kernel of storage kind `kind`, variant v,
// that the device holds at once: its grid, since a persistent block that
// starts late would run its whole share late
int grid_tb(bgk_batch_t *b, int kind, int v)
{
    if (b->grid_tb[kind][v] > 0) return b->grid_tb[kind][v];
    int per_cu = 0;
    const int e = kind == bgk::MK_F16 ? occupancy_tb<_Float16>(v, &per_cu)
                  : kind == bgk::MK_BF16 ? occupancy_tb<bgk::bf16_t>(v, &per_cu)
                                         : occupancy_tb<float>(v, &per_cu);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    b->grid_tb[kind][v] = per_cu * b->cus;
    return b->grid_tb[kind][v];
}
}  // extern "C++"

bool tile_shape(int code, int *tw_log2, int *th)
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

void drop_graphs(bgk_batch_t *b)
{
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
}

int ensure_tiles(bgk_batch_t *b, size_t n)
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
void batch_release(bgk_batch_t *b)
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

}  // namespace

int bgk_batch_create(int32_t device, double budget_bytes, int32_t max_members, bgk_batch_t **out)
{
    if (!out || max_members < 1 || max_members > 65536 || !(budget_bytes > 0)) return BGK_E_INVALID;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return BGK_E_HIP;
    bgk_batch_t *b = new bgk_batch_t();
    b->dev = device;
    b->budget = (size_t)budget_bytes;
    b->used = 0;
    b->slots.assign((size_t)max_members, bgk_batch_t::Slot{nullptr, 0, 0, 0, 6, 4, false});
    b->tiles_cap = 0;
    b->d_tiles = nullptr;
    b->h_tiles = nullptr;
    b->last_ms = 0.0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        cus = 256;
    b->grid = cus * 8;
    b->cus = cus;
    for (int k = 0; k < 3; k++)
        for (int v = 0; v < 6; v++) b->grid_tb[k][v] = 0;
    if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess) {
        batch_release(b);
        return BGK_E_HIP;
    }
    for (int q = 0; q < bgk_batch_t::NQ; q++) {
        if (hipStreamCreateWithFlags(&b->qs[q], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->qev[q], hipEventDisableTiming) != hipSuccess) {
            batch_release(b);
            return BGK_E_HIP;
        }
    }
    if (hipMalloc(&b->d_members, sizeof(bgk::Member) * max_members) != hipSuccess ||
        hipHostMalloc(&b->h_members, sizeof(bgk::Member) * max_members) != hipSuccess ||
        hipMalloc(&b->d_ntiles, sizeof(int) * 2 * bgk::NGROUPS) != hipSuccess ||
        hipHostMalloc(&b->h_ntiles, sizeof(int) * 2 * bgk::NGROUPS) != hipSuccess) {
        batch_release(b);
        return BGK_E_NOMEM;
    }
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        batch_release(b);
        return BGK_E_HIP;
    }
    *out = b;
    return BGK_OK;
}

int bgk_batch_destroy(bgk_batch_t *b)
{
    if (!b) return BGK_E_INVALID;
    (void)hipSetDevice(b->dev);
    batch_release(b);
    return BGK_OK;
}

const char *bgk_batch_last_error(const bgk_batch_t *b) { return b ? b->err.c_str() : "null batch"; }

int bgk_footprint(const bgk_params_t *p, double *bytes)
{
    if (!p || !bytes) return param_fail("null argument");
    if (const char *why = param_error(p)) return param_fail(why);
    *bytes = (double)footprint_of(p);
    return BGK_OK;
}

int bgk_batch_add(bgk_batch_t *b, const bgk_params_t *p, int32_t init_v, double amplitude, int64_t nsteps,
                  int32_t tile, int32_t *slot)
{
    if (!b || !p || !slot || nsteps < 0) return BGK_E_INVALID;
    *slot = -1;
    if (p->impl != BGK_IMPL_FUSED || p->bc == BGK_BC_STEP || p->obstacle ||
        !(p->prec == BGK_PREC_SINGLE || p->prec == BGK_PREC_FP16 || p->prec == BGK_PREC_BF16))
        return bfail(b, BGK_E_INVALID, "batch members are FUSED, without obstacle or step, fp32/fp16/bf16");
    if (p->device != b->dev) return bfail(b, BGK_E_INVALID, "member on device %d, batch on %d", p->device, b->dev);
    if (const char *why = param_error(p)) return bfail(b, BGK_E_INVALID, "%s", why);
    int tw_log2, th;
    if (!tile_shape(tile, &tw_log2, &th)) return bfail(b, BGK_E_INVALID, "unknown tile code %d", tile);
    const size_t bytes = footprint_of(p);
    if (b->used + bytes > b->budget) return bfail(b, BGK_E_NOMEM, "HBM budget: %zu + %zu > %zu", b->used, bytes, b->budget);
    int free_slot = -1;
    for (size_t q = 0; q < b->slots.size(); q++)
        if (!b->slots[q].ctx) {
            free_slot = (int)q;
            break;
        }
    if (free_slot < 0) return bfail(b, BGK_E_NOMEM, "all %zu slots in use", b->slots.size());
    bgk_ctx_t *c = nullptr;
    int r = bgk_create(p, &c);
    if (r != BGK_OK) return bfail(b, r, "bgk_create failed (%d)", r);
    r = bgk_init(c, init_v, amplitude);
    if (r != BGK_OK) {
        bfail(b, r, "bgk_init: %s", c->err.c_str());
        bgk_destroy(c);
        return r;
    }
    b->slots[free_slot] = bgk_batch_t::Slot{c, (long long)nsteps, 0, bytes, tw_log2, th, false};
    b->used += bytes;
    *slot = free_slot;
    return BGK_OK;
}

int bgk_batch_ctx(bgk_batch_t *b, int32_t slot, bgk_ctx_t **ctx)
{
    if (!b || !ctx || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].ctx) return BGK_E_INVALID;
    *ctx = b->slots[slot].ctx;
    return BGK_OK;
}

int bgk_batch_remaining(bgk_batch_t *b, int32_t slot, int64_t *remaining)
{
    if (!b || !remaining || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].ctx) return BGK_E_INVALID;
    *remaining = b->slots[slot].remaining;
    return BGK_OK;
}

int bgk_batch_remove(bgk_batch_t *b, int32_t slot)
{
    if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].ctx) return BGK_E_INVALID;
    (void)hipStreamSynchronize(b->s);
    bgk_destroy(b->slots[slot].ctx);
    b->used -= b->slots[slot].bytes;
    b->slots[slot] = bgk_batch_t::Slot{nullptr, 0, 0, 0, 6, 4, false};
    return BGK_OK;
}

int bgk_batch_hold(bgk_batch_t *b, int32_t slot, int32_t held)
{
    if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].ctx || held < 0 || held > 1)
        return BGK_E_INVALID;
    b->slots[slot].held = held != 0;
    return BGK_OK;
}

int bgk_batch_used(const bgk_batch_t *b, double *bytes)
{
    if (!b || !bytes) return BGK_E_INVALID;
    *bytes = (double)b->used;
    return BGK_OK;
}

// Advance every member with steps left by up to `chunk` steps: one graph of
// `chunk` k_batch launches.  *finished = members that reached zero.
int bgk_batch_run(bgk_batch_t *b, int32_t chunk, int32_t *finished)
{
    if (!b || chunk < 1 || chunk > 4096) return BGK_E_INVALID;
    BHIPCHK(b, hipSetDevice(b->dev));
    int nm = 0;
    size_t nt = 0;
    std::vector<int> members_slot;
    std::vector<char> tbm;  // member advances two steps per launch in this chunk
    constexpr int G = bgk::NGROUPS;
    size_t strip_rows[G] = {};  // per two-step group
    for (size_t q = 0; q < b->slots.size(); q++) {
        const auto &sl = b->slots[q];
        if (!sl.ctx || sl.remaining <= 0 || sl.held) continue;
        const bgk_ctx *c = sl.ctx;
        const bool tb = member_tb(c, sl.remaining, chunk) && !member_solo(c);
        if (tb) strip_rows[member_group(c, true)] += (size_t)member_strips(c) * (size_t)c->m;
        members_slot.push_back((int)q);
        tbm.push_back(tb);
        nm++;
    }
    // Two-step groups run one after another, each sized to fill the chip in
    // whole rounds of strips of at least TB_HMIN rows.  A group too small
    // for one such round would leave most of the chip idle while it runs:
    // its members take the one-step kernels, whose groups overlap.
    int tbh[G];
    for (int g = 0; g < G; g++) {
        tbh[g] = bgk::TB_BATCH_H;
        if (!strip_rows[g]) continue;
        const int slots = grid_tb(b, g / 6 - bgk::TWO_STEP, g % 6);
        if (strip_rows[g] < (size_t)TB_HMIN * (size_t)slots) {
            for (int mi = 0; mi < nm; mi++)
                if (tbm[mi] && member_group(b->slots[members_slot[mi]].ctx, true) == g) tbm[mi] = 0;
            strip_rows[g] = 0;
            continue;
        }
        tbh[g] = batch_tb_h(strip_rows[g], slots);
    }
    auto height = [&](int mi) { return tbm[mi] ? tbh[member_group(b->slots[members_slot[mi]].ctx, true)] : 0; };
    for (int mi = 0; mi < nm; mi++) {
        const auto &sl = b->slots[members_slot[mi]];
        nt += member_tasks(sl.ctx, sl.tw_log2, sl.th, tbm[mi], height(mi));
    }
    if (finished) *finished = 0;
    if (nm == 0) {
        b->last_ms = 0.0;
        return BGK_OK;
    }
    int r = ensure_tiles(b, nt);
    if (r) return r;
    // the task table holds the groups one after another
    size_t cnt[G] = {}, pos[G];
    double gnodes[G] = {};  // nodes a group updates per step
    for (int mi = 0; mi < nm; mi++) {
        const auto &sl = b->slots[members_slot[mi]];
        cnt[member_group(sl.ctx, tbm[mi])] += member_tasks(sl.ctx, sl.tw_log2, sl.th, tbm[mi], height(mi));
    }
    for (int g = 0, acc = 0; g < G; g++) {
        pos[g] = (size_t)acc;
        b->h_ntiles[g] = (int)cnt[g];
        b->h_ntiles[G + g] = acc;
        acc += (int)cnt[g];
    }
    for (int mi = 0; mi < nm; mi++) {
        const auto &sl = b->slots[members_slot[mi]];
        const bgk_ctx *c = sl.ctx;
        bgk::Member &d = b->h_members[mi];
        // offsets from the table's device address (see bgk::Member)
        const auto rel = [b](void *q) { return (long long)((uintptr_t)q - (uintptr_t)b->d_members); };
        for (int k = 0; k < 8; k++) {
            d.set[0][k] = rel(c->plane[k]);
            d.set[1][k] = rel(c->plane[8 + k]);
        }
        d.r[0] = rel(c->plane[16]);
        d.r[1] = rel(c->plane[17]);
        d.rparity0 = c->rcur;
        d.l = c->l;
        d.m = c->m;
        d.off = c->off;
        d.rs = c->pitch;
        d.tw_log2 = sl.tw_log2;
        d.parity0 = c->cur;
        d.remaining = (int)(sl.remaining < chunk ? sl.remaining : chunk);
        d.P = phys<float>(c);
        if (tbm[mi]) {
            d.tiles_x = member_strips(c);
            d.th = height(mi);
            d.ntiles = (int)member_tb_tiles(c, d.th);
        } else {
            d.tiles_x = member_tiles_x(c, sl.tw_log2);
            d.th = sl.th;
            d.ntiles = (int)member_tiles(c, sl.tw_log2, sl.th);
        }
        if (member_solo(c)) continue;
        gnodes[member_group(c, tbm[mi])] += (double)c->l * (double)c->m;
        size_t &t = pos[member_group(c, tbm[mi])];
        const int per = tbm[mi] ? 1 : bgk::TASK_TILES;
        for (int q = 0; q < d.ntiles; q += per) b->h_tiles[t++] = make_int2(mi, q);
    }
    BHIPCHK(b, hipMemcpyAsync(b->d_members, b->h_members, sizeof(bgk::Member) * nm, hipMemcpyHostToDevice, b->s));
    BHIPCHK(b, hipMemcpyAsync(b->d_tiles, b->h_tiles, sizeof(int2) * nt, hipMemcpyHostToDevice, b->s));
    BHIPCHK(b, hipMemcpyAsync(b->d_ntiles, b->h_ntiles, sizeof(int) * 2 * G, hipMemcpyHostToDevice, b->s));
    BHIPCHK(b, hipEventRecord(b->ev0, b->s));
    // The large members run one after another: their kernels are sized to
    // fill the chip in one round, and sharing it would leave a second,
    // half-empty round.
    hipEvent_t prev = b->ev0;
    for (int mi = 0; mi < nm; mi++) {
        bgk_ctx *c = b->slots[members_slot[mi]].ctx;
        if (!member_solo(c)) continue;
        BHIPCHK(b, hipStreamWaitEvent(c->s, prev, 0));
        const int rs = DISPATCH(c, enqueue_t, b->h_members[mi].remaining);  // flips c->cur, c->rcur
        if (rs != BGK_OK) return bfail(b, rs, "member %d: %s", members_slot[mi], bgk_last_error(c));
        BHIPCHK(b, hipEventRecord(c->ev1, c->s));
        prev = c->ev1;
    }
    BHIPCHK(b, hipStreamWaitEvent(b->s, prev, 0));
    // Each group's chain of `chunk` steps replays from its own graph, after
    // the tables are in place.  Two-step groups, each sized to fill the chip
    // in whole rounds, run one after another on the first stream (sharing
    // the chip would leave a partial last round in each); one-step groups
    // go to the stream with the least work so far, largest first, and
    // overlap with the other streams.
    double load[bgk_batch_t::NQ] = {};
    int order[G], ng = 0, where[G];
    for (int g = 0; g < G; g++) {
        if (!cnt[g]) continue;
        if (g / 6 >= bgk::TWO_STEP) {
            where[g] = 0;
            load[0] += 0.5 * gnodes[g];  // two steps per pass: about half the time per update
        } else {
            order[ng++] = g;
        }
    }
    std::sort(order, order + ng, [&](int x, int y) { return gnodes[x] > gnodes[y]; });
    for (int k = 0; k < ng; k++) {
        int q = 0;
        for (int u = 1; u < bgk_batch_t::NQ; u++)
            if (load[u] < load[q]) q = u;
        where[order[k]] = q;
        load[q] += gnodes[order[k]];
    }
    bool used[bgk_batch_t::NQ] = {};
    for (int g = 0; g < G; g++) {
        if (!cnt[g]) continue;
        hipStream_t qs = b->qs[where[g]];
        if (!used[where[g]]) {
            BHIPCHK(b, hipStreamWaitEvent(qs, prev, 0));
            used[where[g]] = true;
        }
        // a group with few tasks launches fewer blocks, leaving the rest of
        // the chip to the other groups' chains
        // (two-step groups: always the full grid, whose surplus blocks exit
        // at once; their strip height already sizes the work to it)
        const bool two = g / 6 >= bgk::TWO_STEP;
        const int full = two ? grid_tb(b, g / 6 - bgk::TWO_STEP, g % 6) : b->grid;
        const int nb = !two && cnt[g] <= SMALL_GROUP ? (int)SMALL_GROUP : full;
        const int nl = two ? chunk / 2 : chunk;  // launches of the chain
        const long long key = ((long long)chunk * G + g) * 2 + (nb == full);
        // a chain seen for the first time is launched directly; its graph
        // is captured when it comes again (capture and instantiation cost
        // more than the launches of one chunk)
        if (++b->uses[key] < 2) {
            for (int tt = 0; tt < nl; tt++) launch_batch(b, g, nb, tt, qs);
            BHIPCHK(b, hipGetLastError());
        } else {
            hipGraphExec_t ge;
            auto it = b->graphs.find(key);
            if (it == b->graphs.end()) {
                BHIPCHK(b, hipStreamBeginCapture(qs, hipStreamCaptureModeThreadLocal));
                for (int tt = 0; tt < nl; tt++) launch_batch(b, g, nb, tt, qs);
                hipGraph_t gr = nullptr;
                BHIPCHK(b, hipStreamEndCapture(qs, &gr));
                hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) return bfail(b, BGK_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
                b->graphs[key] = ge;
            } else {
                ge = it->second;
            }
            BHIPCHK(b, hipGraphLaunch(ge, qs));
        }
    }
    for (int q = 0; q < bgk_batch_t::NQ; q++) {
        if (!used[q]) continue;
        BHIPCHK(b, hipEventRecord(b->qev[q], b->qs[q]));
        BHIPCHK(b, hipStreamWaitEvent(b->s, b->qev[q], 0));
    }
    BHIPCHK(b, hipEventRecord(b->ev1, b->s));
    BHIPCHK(b, hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BHIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    int nfin = 0;
    for (int mi = 0; mi < nm; mi++) {
        auto &sl = b->slots[members_slot[mi]];
        const int steps = b->h_members[mi].remaining;
        if (!member_solo(sl.ctx)) {
            // each launch swaps the sets and the pop-19 planes once
            const int launches = tbm[mi] ? steps / 2 : steps;
            sl.ctx->cur = (sl.ctx->cur + launches) & 1;
            sl.ctx->rcur = (sl.ctx->rcur + launches) & 1;
        }
        sl.remaining -= steps;
        sl.done += steps;
        if (sl.remaining == 0) nfin++;
    }
    if (finished) *finished = nfin;
    return BGK_OK;
}

int bgk_batch_last_ms(const bgk_batch_t *b, double *ms)
{
    if (!b || !ms) return BGK_E_INVALID;
    *ms = b->last_ms;
    return BGK_OK;
}

int bgk_last_run_ms(const bgk_ctx_t *c, double *ms)
{
    if (!c || !ms) return BGK_E_INVALID;
    *ms = c->last_ms;
    return BGK_OK;
}

}  // extern "C"
