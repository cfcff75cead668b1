// bgk_api.hip -- C ABI and runtime of libbgk2d_hip.so (see include/bgk2d.h).
//
// A context owns one HBM allocation (18 pitched population planes plus
// the obstacle mask and reduction scratch), one HIP stream and the HIP
// graphs it replays.  Host code derives the reference's constants in the
// compute kind (hencol, build_bcond) and launches the kernels of
// bgk_kernels.hpp in the reference's phase order.
#include "bgk_host.hpp"

#include <mutex>

using namespace bgk_host;

// ---------------------------------------------------------------------
// Memory of bgk_init_noalloc: the trig tables of init in `scratch`, one
// copy per GPU (held until the process ends), staged from h_stage.  The
// array's name is a host-side stand-in: host code uses only the device
// address hipGetSymbolAddress returns, kept per device in scratch_ptr.
// Calls serialize on scratch_mu, which is held until the kernel that
// reads the tables has ended.
static __device__ __align__(256) unsigned char scratch[BGK_INIT_SCRATCH_BYTES];
static unsigned char h_stage[BGK_INIT_SCRATCH_BYTES];
static std::mutex scratch_mu;
constexpr int SCRATCH_DEVICES = 64;
static void *scratch_ptr[SCRATCH_DEVICES];

// device address of `scratch` on the current device (c's), looked up once
// per device; call with scratch_mu held
static int scratch_on(bgk_ctx *c, void **out)
{
    if (c->dev < 0 || c->dev >= SCRATCH_DEVICES)
        return fail(c, BGK_E_INVALID, "bgk_init_noalloc: device %d beyond %d", c->dev, SCRATCH_DEVICES - 1);
    if (!scratch_ptr[c->dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(c, BGK_E_HIP, "bgk_init_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        scratch_ptr[c->dev] = ptr;
    }
    *out = scratch_ptr[c->dev];
    return BGK_OK;
}


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
    // mykind: obstacle 4 the porous layout (below), 1 (and 0) the cylinder
    // (icoord-i)^2+(jcoord-j)^2 <=
    // R*R (default branch), 2 the square of half side radius (SQUARE), 3 the
    // cylinder's rows m/2..m (HALF_CIRCLE)
    const int l = c->l, m = c->m, ic = 2 * l / 5, jc = m / 2, shape = c->p.obstacle;
    std::vector<int32_t> obs((size_t)l * m, 0);
    const bool dbl = c->p.prec == BGK_PREC_DOUBLE || c->p.prec == BGK_PREC_MIXED2;
    if (shape == 4) {
        // POROUS (SRC/build.obs.F90:117-147): a plus of five solid nodes
        // centred at (i, j), j = 5, 10, .. <= m-4, i = l/4, l/4+5, .. <=
        // 3l/4, wherever the next uniform draw exceeds 0.4.  The reference
        // draws from its compiler's RAND extension; here the draws come
        // from the Park-Miller minimal standard generator (x <- 16807 x mod
        // 2^31-1, seed 1, draw x / (2^31-1), the generator gfortran
        // documents for RAND), compared in the compute kind.
        if (l < 12 || m < 10) return fail(c, BGK_E_INVALID, "POROUS needs l >= 12, m >= 10");
        uint64_t x = 1;
        for (int j = 5; j <= m - 4; j += 5)
            for (int i = l / 4; i <= 3 * l / 4; i += 5) {
                x = x * 16807u % 2147483647u;
                const bool hit = dbl ? (double)x / 2147483647.0 > (double)0.4f
                                     : (float)((double)x / 2147483647.0) > 0.4f;
                if (!hit) continue;
                const int di[5] = {1, 0, 0, -1, 0}, dj[5] = {0, 1, 0, 0, -1};
                for (int q = 0; q < 5; q++) obs[(size_t)(i + di[q] - 1) + (size_t)(j + dj[q] - 1) * l] = 1;
            }
        return set_mask(c, obs.data());
    }
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

int bgk_init_noalloc(bgk_ctx_t *c, int32_t init_v, double amplitude)
{
    int r = check_ctx(c);
    if (r) return r;
    if (init_v < 0 || init_v > 5) return fail(c, BGK_E_INVALID, "Initial condition out of range[0,5] %d", init_v);
    const size_t need = init_table_bytes(c), per_node = need / ((size_t)c->l + c->m + 4);
    if (need > (size_t)BGK_INIT_SCRATCH_BYTES)
        return fail(c, BGK_E_INVALID,
                    "bgk_init_noalloc: trig tables of %zu bytes exceed the %d-byte scratch array "
                    "(lx + ly + 4 <= %zu for this compute kind); use bgk_init",
                    need, BGK_INIT_SCRATCH_BYTES, (size_t)BGK_INIT_SCRATCH_BYTES / per_node);
    std::lock_guard<std::mutex> hold(scratch_mu);
    void *d = nullptr;
    if ((r = scratch_on(c, &d))) return r;
    r = DISPATCH(c, init_static_t, amplitude, (int)init_v, d, (void *)h_stage);
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
        for (int f = 0; f < 2; f++)
            for (int v = 0; v < 6; v++) b->grid_tb[k][f][v] = 0;
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
    if (p->impl != BGK_IMPL_FUSED || p->bc == BGK_BC_STEP || p->obstacle || p->noshift ||
        !(p->prec == BGK_PREC_SINGLE || p->prec == BGK_PREC_FP16 || p->prec == BGK_PREC_BF16))
        return bfail(b, BGK_E_INVALID, "batch members are FUSED, shifted, without obstacle or step, fp32/fp16/bf16");
    if (p->device != b->dev) return bfail(b, BGK_E_INVALID, "member on device %d, batch on %d", p->device, b->dev);
    if (const char *why = param_error(p)) return bfail(b, BGK_E_INVALID, "%s", why);
    {
        const bool narrow = p->prec == BGK_PREC_FP16 || p->prec == BGK_PREC_BF16;
        if (narrow && p->form != p->rounding)
            return bfail(b, BGK_E_INVALID, "16-bit batch members: form 0 with nearest-even stores or form 1 with "
                                           "stochastic rounding");
    }
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
        const int slots = grid_tb(b, bgk::group_slot(g) - bgk::TWO_STEP, bgk::group_fl(g), bgk::group_variant(g));
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
        if (bgk::group_slot(g) >= bgk::TWO_STEP) {
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
        const bool two = bgk::group_slot(g) >= bgk::TWO_STEP;
        const int full =
            two ? grid_tb(b, bgk::group_slot(g) - bgk::TWO_STEP, bgk::group_fl(g), bgk::group_variant(g)) : b->grid;
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
                // nothing between the begin and the end of the capture
                // returns early, so no error leaves qs capturing
                BHIPCHK(b, hipStreamBeginCapture(qs, hipStreamCaptureModeThreadLocal));
                for (int tt = 0; tt < nl; tt++) launch_batch(b, g, nb, tt, qs);
                hipGraph_t gr = nullptr;
                hipError_t e = hipStreamEndCapture(qs, &gr);
                if (e != hipSuccess) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return bfail(b, BGK_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
                }
                e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
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
