// bgk_kernels.hpp -- gfx950 kernels of the BGK2D time step.
//
// Kernel inventory (reference loop each one replaces):
//   k_fill        alloca's huge() fill            SRC/alloca.F90:133-151
//   k_init        init                            SRC/init.F90:83-137
//   k_bc          the 1-D halo passes of every bcond_* variant
//   k_bc_obs      bcond_obs                       SRC/bcond.obs.F90:50-59
//   k_movef       movef (ORIGINAL stream)         SRC/movef.F90:46-53
//   k_col         col (ORIGINAL collide)          SRC/col.F90:90-237
//   k_colmc       col_MC (FUSED pull + collide)   SRC/col_MC.F90:114-293
//   k_fused       col_MC with the boundary folded into the pull (no halo
//                 pass): the hot kernel of bgk_run
//   k_rows/k_rows_final  diagno + varm + divergence check, FP64
//   k_line        prof_i / prof_j / probe
//   k_pack/k_unpack      population <-> dense host layout
#pragma once
#include "bgk_device.hpp"

#include <type_traits>

namespace bgk {

template <class S, class C>
using CodecT = Codec<S, C>;

template <class S>
__device__ __forceinline__ S &el(void *p, long long idx) { return ((S *)p)[idx]; }
template <class S>
__device__ __forceinline__ S elc(const void *p, long long idx) { return ((const S *)p)[idx]; }

// decoded read / encoded write of population k (0..7 in a set, 8 = rest)
template <class S, class C>
__device__ __forceinline__ C get(const Lat &L, const Phys<C> &P, void *const *set, int k, int i, int j)
{
    const void *p = (k == 8) ? L.r : set[k];
    return CodecT<S, C>::dec(elc<S>(p, L.at(i, j)), P.inv);
}
template <class S, class C>
__device__ __forceinline__ typename CodecT<S, C>::raw encode(const Phys<C> &P, C v, int i, int j, int k)
{
    return CodecT<S, C>::enc(v, P.sc, P.rounding, P.rounding ? salt_of(i, j) : 0u, k);
}
// the same with the node's salt already formed
template <class S, class C>
__device__ __forceinline__ typename CodecT<S, C>::raw encode_s(const Phys<C> &P, C v, uint32_t salt, int k)
{
    return CodecT<S, C>::enc(v, P.sc, P.rounding, salt, k);
}
// A fresh copy of a 32-bit value for the compiler: an address formed from
// it after the kernel's branches stays a scalar base plus this 32-bit vector
// offset (the saddr form of global loads and stores) instead of a 64-bit
// vector address per access.  An empty asm statement; no instruction.
__device__ __forceinline__ uint32_t fresh(uint32_t x)
{
    asm volatile("" : "+v"(x));
    return x;
}
template <class S, class C>
__device__ __forceinline__ void put(const Lat &L, const Phys<C> &P, void *const *set, int k, int i, int j, C v)
{
    void *p = (k == 8) ? L.r : set[k];
    el<S>(p, L.at(i, j)) = encode<S, C>(P, v, i, j, k);
}
// streaming store of a result population: written once per step and not
// read again by this launch, so it bypasses cache retention
__device__ __forceinline__ void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(_Float16 *p, _Float16 v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(bf16_t *p, bf16_t v) { __builtin_nontemporal_store(v.bits, &p->bits); }
template <class S, class C>
__device__ __forceinline__ void put_nt(const Lat &L, const Phys<C> &P, void *const *set, int k, int i, int j, C v)
{
    void *p = (k == 8) ? L.r : set[k];
    st_nt(&el<S>(p, L.at(i, j)), encode<S, C>(P, v, i, j, k));
}

// raw copy of a stored element (the reference's `a(x) = a(y)` assignments)
template <class S>
__device__ __forceinline__ void cpy(const Lat &L, void *dst, int di, int dj, const void *src, int si, int sj)
{
    el<S>(dst, L.at(di, dj)) = elc<S>(src, L.at(si, sj));
}
// storage round trip of a compute-kind value stored at halo node (i, j) as
// population k by a boundary and read back by the pull
template <class S, class C>
__device__ __forceinline__ C rt(const Phys<C> &P, C v, int i, int j, int k)
{
    return CodecT<S, C>::dec(encode<S, C>(P, v, i, j, k), P.inv);
}

// ------------------------------------------------------------ fill/init
template <class S>
__global__ void k_fill(void *base, long long n, S v)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) ((S *)base)[t] = v;
}

// init (SRC/init.F90:83-137): tx = {sin, cos}(2 pi x_i) per column and
// ty = {sin, cos}(2 pi y_j) per row, evaluated on the host with the C
// library's sin/cos in the compute kind as the reference's build does.
template <class S, class C>
__global__ void k_init(Lat L, Phys<C> P, const C *tsx, const C *tcx, const C *tsy, const C *tcy, double amp,
                       int tgv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > L.l + 1 || j > L.m + 1) return;
    C xj = (C)0.0, yj = (C)0.0;
    if (tgv) {
        xj = (C)(amp * (double)tsx[i] * (double)tcy[j]);
        yj = (C)(-(amp * (double)tcx[i] * (double)tsy[j]));
    }
    if (tgv == 2) {
        // Taylor-Green with the density of its pressure field,
        // rho = 1 + (3 A^2/4)(cos 4 pi x + cos 4 pi y), so that no acoustic
        // transient is launched (an extension; the reference starts at rho = 1)
        const double c2x = (double)tcx[i] * (double)tcx[i] - (double)tsx[i] * (double)tsx[i];
        const double c2y = (double)tcy[j] * (double)tcy[j] - (double)tsy[j] * (double)tsy[j];
        const double drho = 0.75 * amp * amp * (c2x + c2y);
        const double u = (double)xj, v = (double)yj, usq = u * u + v * v;
        const double w[9] = {1.0 / 36, 1.0 / 36, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 9, 1.0 / 9, 4.0 / 9};
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double cu = cx_of(k) * u + cy_of(k) * v;
            const double f = w[k] * drho + w[k] * (1.0 + drho) * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq) +
                             (P.noshift ? w[k] : 0.0);
            put<S, C>(L, P, L.a, k, i, j, (C)f);
        }
        return;
    }
    const C rf = K<C>::rf, qf = K<C>::qf, p0 = K<C>::p0, p1 = K<C>::p1, p2 = K<C>::p2;
    // NOSHIFT stores whole populations (SRC/init.F90:73-77)
    const C crho = (C)1.0, cte1 = P.noshift ? (C)1.0 : (C)0.0;
    C cvsq = xj * xj + yj * yj;
    C c01 = eq_d<C>(xj - yj, cvsq);
    C c03 = eq_d<C>(xj + yj, cvsq);
    C c05 = eq_d<C>(xj, cvsq);
    C c08 = eq_d<C>(yj, cvsq);
    C c10 = (C)((double)(rf * (-xj - yj)) + (double)qf * (3.0 * (double)(xj + yj) * (double)(xj + yj) - (double)cvsq));
    C c12 = (C)((double)(rf * (-xj + yj)) + (double)qf * (3.0 * (double)(xj - yj) * (double)(xj - yj) - (double)cvsq));
    C c14 = (C)((double)(rf * (-xj)) + (double)qf * (3.0 * (double)xj * (double)xj - (double)cvsq));
    C c17 = (C)((double)(rf * (-yj)) + (double)qf * (3.0 * (double)yj * (double)yj - (double)cvsq));
    C c19 = (C)((double)rf * 0.0 + (double)qf * (3.0 * 0.0 * 0.0 - (double)cvsq));
    put<S, C>(L, P, L.a, 0, i, j, crho * p2 * (cte1 + c01));
    put<S, C>(L, P, L.a, 1, i, j, crho * p2 * (cte1 + c03));
    put<S, C>(L, P, L.a, 2, i, j, crho * p1 * (cte1 + c05));
    put<S, C>(L, P, L.a, 3, i, j, crho * p1 * (cte1 + c08));
    put<S, C>(L, P, L.a, 4, i, j, crho * p2 * (cte1 + c10));
    put<S, C>(L, P, L.a, 5, i, j, crho * p2 * (cte1 + c12));
    put<S, C>(L, P, L.a, 6, i, j, crho * p1 * (cte1 + c14));
    put<S, C>(L, P, L.a, 7, i, j, crho * p1 * (cte1 + c17));
    put<S, C>(L, P, L.a, 8, i, j, crho * p0 * (cte1 + c19));
}

// ------------------------------------------------------------ boundaries
// One 1-D halo pass; the passes of a variant are launched in the order the
// reference's loops run, so corner priorities are the reference's.
enum BcPass {
    BP_DRV_X = 0,    // bcond_driven x-walls, j = 1..m  SRC/bcond.driven.F90:96-109
    BP_WALL_Y = 1,   // bottom wall + top wall (lid: +-u00/6), i = t0..t0+n-1
    BP_XPER = 2,     // x periodic copy, j = t0..t0+n-1  SRC/bcond.periodic.F90:60-68
    BP_YPER = 3,     // y periodic copy, i = 0..l+1      SRC/bcond.periodic.F90:99-107
    BP_INFLOW_X = 4, // outflow + uniform inflow, j = 0..m+1  SRC/bcond.inflow.F90:80-113
    BP_STEP_OUT = 5, // outflow only, j = 0..m+1        SRC/bcond.step.F90:66-102
    BP_STEP_IN = 6,  // parabolic inflow, j = stepy..m+1 SRC/bcond.step.F90:116-151
    BP_STEP_TOP = 7, // step top face, i = 1..stepx     SRC/bcond.step.F90:205-223
    BP_STEP_FACE = 8, // step vertical face, j = 1..stepy SRC/bcond.step.F90:237-255
    BP_OPEN_X = 9,    // OPEN: parabolic inflow + zero-gradient outflow, j = 1..m
    BP_FREESLIP_Y = 10 // OPEN: specular (free-slip) walls, i = 0..l+1
};

// OPEN inflow profile: u(j) = 4 umax y (1 - y), y = (j - 1/2) / m, the
// parabola through the two halfway walls
template <class C>
__device__ __host__ __forceinline__ C open_u(const Phys<C> &P, int j, int m)
{
    const C y = ((C)j - (C)0.5) / (C)m;
    return (C)4.0 * P.umax * y * ((C)1.0 - y);
}
// stored value of inflow population k (0..2: 01,03,05) at x = 0, row j:
// the shifted equilibrium of u = (open_u(j), 0) at the density 1 + drho of
// the first fluid node (1, j), rho w_k eq_k + w_k drho -- a velocity inlet
// whose density follows the interior
template <class C>
__device__ __forceinline__ C open_in(const Phys<C> &P, int k, int j, int m, C drho)
{
    const C rho = (C)1.0 + drho, yj = (C)0.0, xj = open_u<C>(P, j, m);
    const C cvsq = xj * xj + yj * yj;
    if (k == 0) return rho * K<C>::p2 * eq_d<C>(xj - yj, cvsq) + K<C>::p2 * drho;
    if (k == 1) return rho * K<C>::p2 * eq_d<C>(xj + yj, cvsq) + K<C>::p2 * drho;
    return rho * K<C>::p1 * eq_d<C>(xj, cvsq) + K<C>::p1 * drho;
}
// density deviation of a node from its nine populations, in the collision's
// summation order
template <class C, class A>
__device__ __forceinline__ C drho_of(const A &a, int i, int j)
{
    return ((a(0, i, j) + a(1, i, j) + a(4, i, j) + a(5, i, j)) + (a(2, i, j) + a(3, i, j) + a(6, i, j) + a(7, i, j))) +
           a(8, i, j);
}
// specular image across a horizontal wall: 01<->03, 08<->17, 10<->12
__device__ __forceinline__ constexpr int mirror_y(int k)
{
    return k == 0 ? 1 : k == 1 ? 0 : k == 3 ? 7 : k == 7 ? 3 : k == 4 ? 5 : k == 5 ? 4 : k;
}

template <class S, class C>
__device__ void outflow_col(const Lat &L, const Phys<C> &P, int j)
{
    const int l = L.l, l1 = l + 1;
    const C rf = K<C>::rf, qf = K<C>::qf, p1 = K<C>::p1, p2 = K<C>::p2;
    const C rhoinv = (C)1.0, crho = (C)1.0, cte1 = P.noshift ? (C)1.0 : (C)0.0;
    C x01 = get<S, C>(L, P, L.a, 0, l, j), x03 = get<S, C>(L, P, L.a, 1, l, j);
    C x05 = get<S, C>(L, P, L.a, 2, l, j), x08 = get<S, C>(L, P, L.a, 3, l, j);
    C x10 = get<S, C>(L, P, L.a, 4, l, j), x12 = get<S, C>(L, P, L.a, 5, l, j);
    C x14 = get<S, C>(L, P, L.a, 6, l, j), x17 = get<S, C>(L, P, L.a, 7, l, j);
    C xj = ((x03 - x12) + (x01 - x10) + (x05 - x14)) * rhoinv;
    C yj = ((x03 - x01) + (x12 - x10) + (x08 - x17)) * rhoinv;
    C cvsq = xj * xj + yj * yj;
    C cx10 = rf * (-xj - yj) + qf * ((C)3.0 * (xj + yj) * (xj + yj) - cvsq);
    C cx12 = rf * (-xj + yj) + qf * ((C)3.0 * (xj - yj) * (xj - yj) - cvsq);
    C cx14 = rf * (-xj) + qf * ((C)3.0 * (xj) * (xj)-cvsq);
    put<S, C>(L, P, L.a, 4, l1, j, crho * p2 * (cte1 + cx10));
    put<S, C>(L, P, L.a, 5, l1, j, crho * p2 * (cte1 + cx12));
    put<S, C>(L, P, L.a, 6, l1, j, crho * p1 * (cte1 + cx14));
}

template <class S, class C>
__device__ void inflow_col(const Lat &L, const Phys<C> &P, int j, C xj)
{
    const C p1 = K<C>::p1, p2 = K<C>::p2;
    const C crho = (C)1.0, cte1 = P.noshift ? (C)1.0 : (C)0.0, yj = (C)0.0;
    C cvsq = xj * xj + yj * yj;
    C c01 = eq_d<C>(xj - yj, cvsq), c03 = eq_d<C>(xj + yj, cvsq), c05 = eq_d<C>(xj, cvsq);
    put<S, C>(L, P, L.a, 0, 0, j, crho * p2 * (cte1 + c01));
    put<S, C>(L, P, L.a, 1, 0, j, crho * p2 * (cte1 + c03));
    put<S, C>(L, P, L.a, 2, 0, j, crho * p1 * (cte1 + c05));
}

template <class S, class C>
__global__ void k_bc(Lat L, Phys<C> P, int pass, int t0, int n, int lid)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    t += t0;
    const int l = L.l, m = L.m, l1 = l + 1, m1 = m + 1;
    void *const *a = L.a;
    switch (pass) {
    case BP_DRV_X: {
        int j = t;
        cpy<S>(L, a[5], l1, j - 1, a[0], l, j);  // a12(l1,j-1) = a01(l,j)
        cpy<S>(L, a[4], l1, j + 1, a[1], l, j);  // a10(l1,j+1) = a03(l,j)
        cpy<S>(L, a[6], l1, j, a[2], l, j);      // a14(l1,j)   = a05(l,j)
        cpy<S>(L, a[1], 0, j - 1, a[4], 1, j);   // a03(0,j-1)  = a10(1,j)
        cpy<S>(L, a[0], 0, j + 1, a[5], 1, j);   // a01(0,j+1)  = a12(1,j)
        cpy<S>(L, a[2], 0, j, a[6], 1, j);       // a05(0,j)    = a14(1,j)
    } break;
    case BP_WALL_Y: {
        int i = t;
        cpy<S>(L, a[3], i, 0, a[7], i, 1);                       // a08(i,0)   = a17(i,1)
        if (i + 1 <= l1) cpy<S>(L, a[5], i + 1, 0, a[0], i, 1);  // a12(i+1,0) = a01(i,1)
        if (i - 1 >= 0) cpy<S>(L, a[1], i - 1, 0, a[4], i, 1);   // a03(i-1,0) = a10(i,1)
        if (!lid) {  // channel / step walls: plain copies
            if (i + 1 <= l1) cpy<S>(L, a[4], i + 1, m1, a[1], i, m);  // a10(i+1,m1) = a03(i,m)
            if (i - 1 >= 0) cpy<S>(L, a[0], i - 1, m1, a[5], i, m);   // a01(i-1,m1) = a12(i,m)
        } else {      // driven / couette lid: a03(i,m) - force, a12(i,m) + force
            if (i + 1 <= l1) put<S, C>(L, P, a, 4, i + 1, m1, get<S, C>(L, P, a, 1, i, m) - P.force);
            if (i - 1 >= 0) put<S, C>(L, P, a, 0, i - 1, m1, get<S, C>(L, P, a, 5, i, m) + P.force);
        }
        cpy<S>(L, a[7], i, m1, a[3], i, m);  // a17(i,m1) = a08(i,m)
    } break;
    case BP_XPER: {
        int j = t;
        cpy<S>(L, a[0], 0, j, a[0], l, j);
        cpy<S>(L, a[1], 0, j, a[1], l, j);
        cpy<S>(L, a[2], 0, j, a[2], l, j);
        cpy<S>(L, a[4], l1, j, a[4], 1, j);
        cpy<S>(L, a[5], l1, j, a[5], 1, j);
        cpy<S>(L, a[6], l1, j, a[6], 1, j);
    } break;
    case BP_YPER: {
        int i = t;
        cpy<S>(L, a[7], i, m1, a[7], i, 1);
        cpy<S>(L, a[0], i, m1, a[0], i, 1);
        cpy<S>(L, a[4], i, m1, a[4], i, 1);
        cpy<S>(L, a[1], i, 0, a[1], i, m);
        cpy<S>(L, a[3], i, 0, a[3], i, m);
        cpy<S>(L, a[5], i, 0, a[5], i, m);
    } break;
    case BP_INFLOW_X: {
        int j = t;
        outflow_col<S, C>(L, P, j);
        inflow_col<S, C>(L, P, j, (C)0.1f);
    } break;
    case BP_STEP_OUT: outflow_col<S, C>(L, P, t); break;
    case BP_STEP_IN: {
        int j = t;
        const int stepy = L.stepy;
        C xj = (C)0.1f * (C)(float)(j - stepy) * (C)(float)(m - j) / (C)(float)((m - stepy) * (m - stepy) / 4);
        inflow_col<S, C>(L, P, j, xj);
    } break;
    case BP_STEP_TOP: {
        int i = t;
        const int sy = L.stepy;
        cpy<S>(L, a[3], i, sy - 1, a[7], i, sy);
        cpy<S>(L, a[5], i + 1, sy - 1, a[0], i, sy);
        cpy<S>(L, a[1], i - 1, sy - 1, a[4], i, sy);
    } break;
    case BP_OPEN_X: {
        int j = t;
        auto g = [&](int k, int si, int sj) { return get<S, C>(L, P, a, k, si, sj); };
        const C drho = drho_of<C>(g, 1, j);
#pragma unroll
        for (int k = 0; k < 3; k++) put<S, C>(L, P, a, k, 0, j, open_in<C>(P, k, j, m, drho));
        cpy<S>(L, a[4], l1, j, a[4], l, j);  // zero gradient: column l+1 = column l
        cpy<S>(L, a[5], l1, j, a[5], l, j);
        cpy<S>(L, a[6], l1, j, a[6], l, j);
    } break;
    case BP_FREESLIP_Y: {
        int i = t;
        cpy<S>(L, a[1], i, 0, a[0], i, 1);   // a03(i,0)  = a01(i,1)
        cpy<S>(L, a[3], i, 0, a[7], i, 1);   // a08(i,0)  = a17(i,1)
        cpy<S>(L, a[5], i, 0, a[4], i, 1);   // a12(i,0)  = a10(i,1)
        cpy<S>(L, a[0], i, m1, a[1], i, m);  // a01(i,m1) = a03(i,m)
        cpy<S>(L, a[7], i, m1, a[3], i, m);  // a17(i,m1) = a08(i,m)
        cpy<S>(L, a[4], i, m1, a[5], i, m);  // a10(i,m1) = a12(i,m)
    } break;
    case BP_STEP_FACE: {
        int j = t;
        const int sx = L.stepx;
        cpy<S>(L, a[1], sx - 1, j - 1, a[4], sx, j);
        cpy<S>(L, a[0], sx - 1, j + 1, a[5], sx, j);
        cpy<S>(L, a[2], sx - 1, j, a[6], sx, j);
    } break;
    default: break;
    }
}

// bcond_obs without the serial loop: for a solid node s and population k
// the reference sets a_k(s) = a_opp(k)(s + c_k) in j-outer/i-inner order.
// When the neighbour n = s + c_k comes earlier in that order (c_k with
// cy = -1, or cy = 0 and cx = -1: pops 01,10,14,17) and is itself solid,
// it already copied a_k(s) into a_opp(k)(n), so a_k(s) keeps its value;
// otherwise a_k(s) takes n's value, which no thread writes.  Threads that
// would store an unchanged value skip the store, which makes every read
// race-free and the result equal to the serial loop.
__device__ __forceinline__ bool before_pop(int k) { return k == 0 || k == 4 || k == 6 || k == 7; }

__device__ __forceinline__ bool solid(const Lat &L, int i, int j)
{
    if (i < L.imin || i > L.imax || j < L.jmin || j > L.jmax) return false;
    return L.obs[(long long)(i - 1) + (long long)(j - 1) * L.l] != 0;
}

template <class S>
__global__ void k_bc_obs(Lat L)
{
    int i = L.imin + blockIdx.x * blockDim.x + threadIdx.x;
    int j = L.jmin + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > L.imax || j > L.jmax || !solid(L, i, j)) return;
    S v[8];
    bool w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int ni = i + cx_of(k), nj = j + cy_of(k);
        w[k] = !(before_pop(k) && solid(L, ni, nj));
        if (w[k]) v[k] = elc<S>(L.a[opp_of(k)], L.at(ni, nj));
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (w[k]) el<S>(L.a[k], L.at(i, j)) = v[k];
}

// ------------------------------------------------------- stream/collide
template <class S>
__global__ void k_movef(Lat L)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > L.l || j > L.m) return;
#pragma unroll
    for (int k = 0; k < 8; k++) cpy<S>(L, L.b[k], i, j, L.a[k], i - cx_of(k), j - cy_of(k));
}

template <class S, class C>
__global__ void k_col(Lat L, Phys<C> P)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > L.l || j > L.m) return;
    C x[9], o[9];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = get<S, C>(L, P, L.b, k, i, j);
    x[8] = get<S, C>(L, P, L.a, 8, i, j);
    relax<C, true, -1, false, false, true>(P, x, o);
#pragma unroll
    for (int k = 0; k < 9; k++) put_nt<S, C>(L, P, L.a, k, i, j, o[k]);
}

template <class S, class C>
__global__ void k_colmc(Lat L, Phys<C> P)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > L.l || j > L.m) return;
    C x[9], o[9];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = get<S, C>(L, P, L.a, k, i - cx_of(k), j - cy_of(k));
    x[8] = get<S, C>(L, P, L.a, 8, i, j);
    relax<C, true, -1, false, false, true>(P, x, o);
#pragma unroll
    for (int k = 0; k < 8; k++) put_nt<S, C>(L, P, L.b, k, i, j, o[k]);
    put_nt<S, C>(L, P, L.a, 8, i, j, o[8]);
}

// ------------------------------------------------ fused, boundary folded
// Every boundary variant but STEP writes each halo value that the pull
// reads as a function of post-collision interior values of the same step
// (own reversed population +- lid term, periodic wrap, or the equilibrium
// of a column-l node); the pull evaluates that function directly, so one
// kernel per step replaces bcond_* + bcond_obs + col_MC and no halo is
// ever written.  Rules, in the reference's corner priority:
//   DRIVEN  : source row m+1 -> own opposite +/- u00/6 (lid wins corners);
//             any other halo -> own opposite.
//   PERIODIC: wrap both coordinates.
//   CHANNEL : source row out -> own opposite; else wrap x.
//   COUETTE : source row m+1 -> lid; row 0 -> own opposite; else wrap x.
//   INFLOW  : wrap y first; column 0 -> uniform inflow equilibrium,
//             column l+1 -> equilibrium of the momentum of (l, row).
//   OPEN    : row 0 / m+1 -> the mirror population (01<->03, 08<->17,
//             10<->12) of the same column in row 1 / m, that column's value
//             after the x pass; column 0 -> parabolic inflow equilibrium at
//             the density of (1, row); column l+1 -> the same population
//             of (l, row).
// BK_* shares the numbering of the foldable BGK_BC_* variants (5 is STEP).
enum BcKind { BK_DRIVEN = 0, BK_PERIODIC = 1, BK_CHANNEL = 2, BK_INFLOW = 3, BK_COUETTE = 4, BK_OPEN = 6 };

// Periodic directions of each foldable variant
template <int BC>
struct Per {
    static constexpr bool x = BC == BK_PERIODIC || BC == BK_CHANNEL || BC == BK_COUETTE;
    static constexpr bool y = BC == BK_PERIODIC || BC == BK_INFLOW;
};

// Post-collision value of population k at node (si, sj) of the current set,
// read from HBM; coordinates one step outside the interior along a
// periodic direction wrap to the opposite edge.
template <class S, class C, int BC>
struct HbmAcc {
    const Lat &L;
    const Phys<C> &P;
    __device__ __forceinline__ C operator()(int k, int si, int sj) const
    {
        if (Per<BC>::x) si = si < 1 ? si + L.l : (si > L.l ? si - L.l : si);
        if (Per<BC>::y) sj = sj < 1 ? sj + L.m : (sj > L.m ? sj - L.m : sj);
        return get<S, C>(L, P, L.a, k, si, sj);
    }
};

// Value that the boundary pass would have stored at halo node (si, sj) for
// population k, as read by interior node (i, j); `a` returns current
// post-collision interior values (and wraps periodic directions).
template <class S, class C, int BC, class A>
__device__ C folded_a(const A &a, const Phys<C> &P, int l, int m, int k, int i, int j, int si, int sj)
{
    if (BC == BK_PERIODIC) return a(k, si, sj);
    if (BC == BK_OPEN) {
        if (sj < 1 || sj > m) {
            // the y pass copies the mirror population of row 1 / m, which
            // is the node's own row (sj < 1 only for j = 1, sj > m only for
            // j = m); kept apart from the x rule below so that both
            // population indices stay compile-time constants
            const int kk = mirror_y(k);
            if (si < 1) return rt<S, C>(P, open_in<C>(P, kk, j, m, drho_of<C>(a, 1, j)), 0, j, kk);
            if (si > l) return a(kk, l, j);
            return a(kk, si, j);
        }
        if (si < 1) return rt<S, C>(P, open_in<C>(P, k, sj, m, drho_of<C>(a, 1, sj)), 0, sj, k);
        if (si > l) return a(k, l, sj);
        return a(k, si, sj);
    }
    if (BC == BK_INFLOW) {
        const int sjw = sj < 1 ? m : (sj > m ? 1 : sj);  // the y pass runs after the x pass
        if (si < 1) return rt<S, C>(P, P.inflow[k], 0, sjw, k);  // pops 01,03,05 from x = 0
        if (si > l) {
            // outflow equilibrium from the momentum of (l, sj); whole
            // populations in the NOSHIFT build (SRC/bcond.inflow.F90:48-52)
            const C rf = K<C>::rf, qf = K<C>::qf, p1 = K<C>::p1, p2 = K<C>::p2;
            const C rhoinv = (C)1.0, crho = (C)1.0, cte1 = P.noshift ? (C)1.0 : (C)0.0;
            C x01 = a(0, l, sj), x03 = a(1, l, sj);
            C x05 = a(2, l, sj), x08 = a(3, l, sj);
            C x10 = a(4, l, sj), x12 = a(5, l, sj);
            C x14 = a(6, l, sj), x17 = a(7, l, sj);
            C xj = ((x03 - x12) + (x01 - x10) + (x05 - x14)) * rhoinv;
            C yj = ((x03 - x01) + (x12 - x10) + (x08 - x17)) * rhoinv;
            C cvsq = xj * xj + yj * yj;
            C v;
            if (k == 4)
                v = crho * p2 * (cte1 + (rf * (-xj - yj) + qf * ((C)3.0 * (xj + yj) * (xj + yj) - cvsq)));
            else if (k == 5)
                v = crho * p2 * (cte1 + (rf * (-xj + yj) + qf * ((C)3.0 * (xj - yj) * (xj - yj) - cvsq)));
            else
                v = crho * p1 * (cte1 + (rf * (-xj) + qf * ((C)3.0 * (xj) * (xj)-cvsq)));
            return rt<S, C>(P, v, l + 1, sjw, k);
        }
        return a(k, si, sj);
    }
    // wall bounded in y: DRIVEN, CHANNEL, COUETTE
    const int o = opp_of(k);
    // TRICK1 corner order (SRC/bcond.driven.F90:43-81): the side walls are
    // written after the lid, so the two top corners hold plain copies
    if (BC == BK_DRIVEN && sj > m && P.trick1 && (si < 1 || si > l)) return a(o, i, j);
    if (sj > m && (BC == BK_DRIVEN || BC == BK_COUETTE)) {
        C own = a(o, i, j);
        const int c = cx_of(k);
        if (c == 0) return own;
        return rt<S, C>(P, c > 0 ? own + P.force : own - P.force, si, sj, k);
    }
    if (sj < 1 || sj > m) return a(o, i, j);
    if (BC == BK_DRIVEN) return a(o, i, j);
    return a(k, si, sj);  // CHANNEL / COUETTE: x periodic
}

template <class S, class C, int BC>
__device__ C folded(const Lat &L, const Phys<C> &P, int k, int i, int j, int si, int sj)
{
    return folded_a<S, C, BC>(HbmAcc<S, C, BC>{L, P}, P, L.l, L.m, k, i, j, si, sj);
}

// Replace the pulled values whose source is a halo node (folded boundary)
// or a solid node (bcond_obs), in that order, as the reference applies the
// wall boundary before the obstacle one.  With WRAPPED the caller's pulls
// already hold the wrapped values along periodic directions, so only the
// sources outside a wall-bounded direction are replaced.
template <class S, class C, int BC, bool OBS, bool WRAPPED, class A>
__device__ __forceinline__ void fix_body(const A &a, const Lat &L, const Phys<C> &P, int i, int j, C *x)
{
    constexpr bool fx = !(WRAPPED && Per<BC>::x), fy = !(WRAPPED && Per<BC>::y);
    if ((fx && (i == 1 || i == L.l)) || (fy && (j == 1 || j == L.m))) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            int si = i - cx_of(k), sj = j - cy_of(k);
            if ((fx && (si < 1 || si > L.l)) || (fy && (sj < 1 || sj > L.m)))
                x[k] = folded_a<S, C, BC>(a, P, L.l, L.m, k, i, j, si, sj);
        }
    }
    if (OBS && i >= L.imin - 1 && i <= L.imax + 1 && j >= L.jmin - 1 && j <= L.jmax + 1) {
        const bool me = solid(L, i, j);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            int si = i - cx_of(k), sj = j - cy_of(k);
            // the pulled value was set by bcond_obs: keep it only when this
            // node is solid and precedes the source in the serial order
            if (solid(L, si, sj) && !(me && before_pop(k))) x[k] = a(opp_of(k), i, j);
        }
    }
}

template <class S, class C, int BC, bool OBS, bool WRAPPED = false, class A>
__device__ __forceinline__ void fix_pulls_a(const A &a, const Lat &L, const Phys<C> &P, int i, int j, C x[9])
{
    constexpr bool fx = !(WRAPPED && Per<BC>::x), fy = !(WRAPPED && Per<BC>::y);
    const bool edge = (fx && (i == 1 || i == L.l)) || (fy && (j == 1 || j == L.m));
    const bool near = OBS && i >= L.imin - 1 && i <= L.imax + 1 && j >= L.jmin - 1 && j <= L.jmax + 1;
    if (edge || near) fix_body<S, C, BC, OBS, WRAPPED>(a, L, P, i, j, x);
}

template <class S, class C, int BC, bool OBS>
__device__ __forceinline__ void fix_pulls(const Lat &L, const Phys<C> &P, int i, int j, C x[9])
{
    const bool edge = i == 1 || i == L.l || j == 1 || j == L.m;
    const bool near = OBS && i >= L.imin - 1 && i <= L.imax + 1 && j >= L.jmin - 1 && j <= L.jmax + 1;
    if (edge || near) {
        fix_body<S, C, BC, OBS, false>(HbmAcc<S, C, BC>{L, P}, L, P, i, j, x);
        // the folded values were loaded into the pulls' registers: wait for
        // them here, so that no path out of this branch leaves a load into
        // those registers in flight, which would make the wait for the
        // pulls themselves a wait for every memory operation issued
        __builtin_amdgcn_s_waitcnt(0);
    }
}

// Compile-time configuration of a hot-kernel instance.  FL: -1 the generic
// instance (form, rounding and the NOSHIFT switch read from Phys); 0 the
// reference's form with nearest-even stores; 1 the deviation form with
// stochastic rounding of 16-bit stores; 2 the reference's form on whole
// populations (the NOSHIFT build, SRC/col_MC.F90:70-74).  The host picks 0 or 1 when the parameters are those, so
// the values are those of the generic instance; fixing them lets the
// compiler drop the other form's density terms and store path.  With FL = 1
// and 16-bit storage, the variants whose folded boundary rules are copies
// and the lid term (not the INFLOW/OPEN equilibria) compute on the
// populations as stored (SCALED: no multiply per load or store).
template <class S, class C, int BC, int FL>
struct Cfg {
    static constexpr bool FORCE = BC == BK_CHANNEL;
    static constexpr bool SCALED = FL == 1 && sizeof(S) == 2 &&
                                   (BC == BK_DRIVEN || BC == BK_PERIODIC || BC == BK_CHANNEL || BC == BK_COUETTE);
    static constexpr bool Z0 = std::is_same<S, C>::value;
};

template <class S, class C, int BC, int FL>
__device__ __forceinline__ Phys<C> fixed_flags(const Phys<C> &P)
{
    Phys<C> Q = P;
    if constexpr (FL >= 0) {
        Q.form = FL == 1 ? 1 : 0;
        Q.rounding = FL == 1 ? 1 : 0;  // only the 16-bit codecs read it
        Q.noshift = FL == 2 ? 1 : 0;
    }
    if constexpr (Cfg<S, C, BC, FL>::SCALED) {
        // populations stay scale*(f - w) in registers: codecs without the
        // scale, the lid term scaled once (exact: a power of two)
        Q.usc = P.sc;
        Q.dsc = P.inv;
        Q.sc = (C)1.0;
        Q.inv = (C)1.0;
        Q.force = P.force * P.sc;
    }
    return Q;
}

template <class S, class C, int BC, int FL>
__device__ __forceinline__ void relax_k(const Phys<C> &P, const C x[9], C o[9])
{
    using K = Cfg<S, C, BC, FL>;
    relax<C, K::FORCE, (FL == 2 ? 0 : FL), K::SCALED, K::Z0, (FL == 2 || FL < 0)>(P, x, o);
}

// Addressing of the hot kernel: one 32-bit byte offset per node, and per
// population a uniform base pointer that already includes the population's
// pull shift, so every load and store is scalar base + vector offset with
// no per-access address arithmetic (planes are smaller than 4 GiB, which
// bgk_create enforces).
template <class S>
__device__ __forceinline__ const char *pull_base(const Lat &L, const void *plane, int k)
{
    return (const char *)plane + ((long long)L.off - (long long)cy_of(k) * L.rs - cx_of(k)) * (long long)sizeof(S);
}

// one FUSED step of interior node (i, j): pull (boundary folded), relax,
// store into the other set.  The rest population goes to the second pop-19
// plane rather than in place: the OPEN inflow reads the rest population of
// the neighbouring rows' first column, which must still be the old one.
template <class S, class C, int BC, bool OBS, int FL>
__device__ __forceinline__ void fused_node(const Lat &L, const Phys<C> &P, int i, int j)
{
    using R = typename CodecT<S, C>::raw;
    const uint32_t bo = fresh(((uint32_t)j * (uint32_t)L.rs + (uint32_t)i) * (uint32_t)sizeof(S));
    C x[9], o[9];
    // plain pulls; the halo ring is allocated, so edge nodes read it too
    // and then replace those values below
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = CodecT<S, C>::dec(*(const R *)(pull_base<S>(L, L.a[k], k) + bo), P.inv);
    x[8] = CodecT<S, C>::dec(*(const R *)(pull_base<S>(L, L.r, 8) + bo), P.inv);
    fix_pulls<S, C, BC, OBS>(L, P, i, j, x);
    relax_k<S, C, BC, FL>(P, x, o);
    const uint32_t bs = fresh(bo);
    const uint32_t salt = P.rounding ? salt_of(i, j) : 0u;
#pragma unroll
    for (int k = 0; k < 8; k++) st_nt((R *)(pull_base<S>(L, L.b[k], 8) + bs), encode_s<S, C>(P, o[k], salt, k));
    st_nt((R *)(pull_base<S>(L, L.rn, 8) + bs), encode_s<S, C>(P, o[8], salt, 8));
}

template <class S, class C, int BC, bool OBS, int FL = -1>
__global__ __launch_bounds__(256) void k_fused(Lat L, Phys<C> P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > L.l || j > L.m) return;
    fused_node<S, C, BC, OBS, FL>(L, fixed_flags<S, C, BC, FL>(P), i, j);
}

// The same step for 16-bit storage, two nodes (i0, i0+1) per thread with
// i0 odd: each population is moved with 32-bit accesses, halving the
// memory instructions of the 16-bit kinds.  Node i0 starts a dword (the
// first interior node is 128-byte aligned); the x-shifted pulls start at
// an odd element and are assembled from the two aligned dwords around them.
// The results of node i0 are converted into the low halves of the stored
// dwords, then node i0+1's into the high halves in place.
template <class S, class C, int BC, bool OBS, int FL>
__device__ __forceinline__ void fused_pair(const Lat &L, const Phys<C> &P, int i0, int j)
{
    static_assert(sizeof(S) == 2, "pairs of 16-bit elements");
    // With the column-l / inlet equilibria of INFLOW and OPEN, the boundary
    // code inlined twice below would raise the register count of every pair
    // (to ~150): there, pairs that touch an edge or the obstacle take the
    // one-node path instead.  The other variants keep it inline, which is
    // faster for them (the obstacle rule included).  A last node without a
    // partner (odd l) takes the one-node path in every variant.
    constexpr bool split = BC == BK_INFLOW || BC == BK_OPEN;
    if (i0 == L.l || (split && (i0 == 1 || i0 + 1 >= L.l || j == 1 || j == L.m ||
                                (OBS && i0 + 1 >= L.imin - 1 && i0 <= L.imax + 1 && j >= L.jmin - 1 &&
                                 j <= L.jmax + 1)))) {
        fused_node<S, C, BC, OBS, FL>(L, P, i0, j);
        if (i0 + 1 <= L.l) fused_node<S, C, BC, OBS, FL>(L, P, i0 + 1, j);
        return;
    }
    using R = typename CodecT<S, C>::raw;
    using Cd = CodecT<S, C>;
    const uint32_t bo = fresh(((uint32_t)j * (uint32_t)L.rs + (uint32_t)i0) * 2u);
    C x[2][9], o[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int cx = k == 8 ? 0 : cx_of(k);
        const char *base = pull_base<S>(L, k == 8 ? L.r : L.a[k], k);
        uint32_t w;
        if (cx == 0) {
            w = *(const uint32_t *)(base + bo);
        } else {
            const uint32_t lo = *(const uint32_t *)(base + bo - 2), hi = *(const uint32_t *)(base + bo + 2);
            w = (lo >> 16) | (hi << 16);
        }
        x[0][k] = Cd::dec(__builtin_bit_cast(R, (uint16_t)(w & 0xffffu)), P.inv);
        x[1][k] = Cd::dec(__builtin_bit_cast(R, (uint16_t)(w >> 16)), P.inv);
    }
    uint32_t ow[9];
    const uint32_t s0 = P.rounding ? salt_of(i0, j) : 0u;
    const uint32_t s1 = s0 + 0x9E3779u;  // salt_of(i0 + 1, j)
    if (!split) fix_pulls<S, C, BC, OBS>(L, P, i0, j, x[0]);
    relax_k<S, C, BC, FL>(P, x[0], o);
#pragma unroll
    for (int k = 0; k < 9; k++) ow[k] = Cd::enc_lo(o[k], P.sc, P.rounding, s0, k);
    if (!split) fix_pulls<S, C, BC, OBS>(L, P, i0 + 1, j, x[1]);
    relax_k<S, C, BC, FL>(P, x[1], o);
#pragma unroll
    for (int k = 0; k < 9; k++) ow[k] = Cd::enc_hi(ow[k], o[k], P.sc, P.rounding, s1, k);
    const uint32_t bs = fresh(bo);
#pragma unroll
    for (int k = 0; k < 9; k++)
        __builtin_nontemporal_store(ow[k], (uint32_t *)((char *)pull_base<S>(L, k == 8 ? L.rn : L.b[k], 8) + bs));
}

template <class S, class C, int BC, bool OBS, int FL = -1>
__global__ __launch_bounds__(256) void k_fused2(Lat L, Phys<C> P)
{
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 2 + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i0 > L.l || j > L.m) return;
    fused_pair<S, C, BC, OBS, FL>(L, fixed_flags<S, C, BC, FL>(P), i0, j);
}

// ------------------------------------------------ two steps per HBM pass
// k_fused_tb advances the lattice by two FUSED steps in one pass over HBM.
// A block owns a strip of TB_TW columns and H rows and marches up the strip
// one row r at a time, thread t owning column i0+t in both phases:
//   phase 1: step t+1 of row r, pulled from HBM (boundary folded),
//            collided and rounded to the storage kind;
//   phase 2: step t+2 of row r-1, pulled from the step t+1 values, stored
//            to HBM.
// The six populations that move along x (01,03,05,10,12,14) pass between
// threads through an LDS ring of four rows, which also holds the rim
// columns i0-1 (pops 01,03,05) and i0+TW (pops 10,12,14) that the strip's
// edge nodes pull; 08, 17 and 19 stay in the owning thread's registers as
// a three-row history.  The rim columns' step t+1 values are computed
// TB_RB rows at a time (a few wavefront passes instead of one per row) and
// copied into the ring row by row.  The raw pulls of row r+1 are issued
// before phase 2 of row r-1, so their latency overlaps it; phase 1 runs at
// raised wave priority (s_setprio 1) until they are issued, so among the
// CU's waves those about to issue loads get the VALU first (FP16
// Taylor-Green 8192^2: 184 k to 189 k MLUPS; fp32 cavity 4096^2: +1 %).
// Along periodic directions the rim columns and the rows above/below the
// domain hold the wrapped nodes, so phase 2 pulls them directly; along
// wall-bounded directions phase 2 folds the boundary from step t+1 values
// of its own column (own node; the column-l outflow of INFLOW).
// Every value equals what two k_fused launches would store: each node is
// the same function of the same rounded inputs (rim nodes are computed by
// both blocks that read them).  Pop 19 moves from the plane L.r to the
// plane L.rn, because neighbouring blocks still read the old value of a
// node when its owner writes the new one.
// Ring hazard: phase 2 of row r-1 reads ring rows r-2..r; phase 1 of row
// r+1 overwrites ring row r-3, read last by phase 2 of row r-2, which every
// thread has finished when it passes the barrier of row r.
// strip width of the two-step kernels: one column per thread, one block
// per strip.  TB_TW is the width of the batch kernel and of bgk_run's fp32
// and fp64 kernels; bgk_run's 16-bit kernels use strips twice as wide (FP16
// Taylor-Green 8192^2: 165 k against 161 k MLUPS; for fp32 512 columns gave
// 128.5 k against 128.2 k, and the fp64 ring would not fit; 64, 128 and
// 1024 columns were slower for every kind).
#ifndef BGK_TB_TW
#define BGK_TB_TW 256
#endif
constexpr int TB_TW = BGK_TB_TW;
template <class S>
__host__ __device__ constexpr int tb_tw()
{
    return sizeof(S) == 2 ? 2 * TB_TW : TB_TW;
}
constexpr int TB_RB = 16;   // rows of rim columns computed per batch (power of two; 16 beat 8, 32, 64)

// Two-step instances that march with no branch around their loads and
// stores (tb_block), under a bound of six wavefronts per SIMD (80 VGPRs):
// 16-bit storage with the boundary variants where that measured faster.
// MLUPS, branched march -> branch-free (FP16, 8192^2 Taylor-Green,
// 4096^2 cavity, 8192x2048 channels): periodic 165 k -> 186 k, driven
// 152 k -> 171 k, channel 141 k -> 156 k, inflow 91 k -> 110 k; open
// (120 k -> 104 k) and couette (127 k -> 113 k) lost occupancy to the
// extra registers and keep the branched march, as do fp32 (127.5 k ->
// 127.0 k) and fp64.
template <class S, int BC>
__host__ __device__ constexpr bool tb_branch_free()
{
    return sizeof(S) == 2 && (BC == BK_DRIVEN || BC == BK_PERIODIC || BC == BK_CHANNEL || BC == BK_INFLOW);
}

// ring slot of population k (0..7, not 08/17): 01,03,05,10,12,14 -> 0..5
__device__ __forceinline__ constexpr int tb_slot(int k) { return k < 3 ? k : k - 1; }

template <class S, class C, int E>
struct TbAcc {
    using R = typename CodecT<S, C>::raw;
    const R (*ring)[6][E];
    int e0, r0;  // node column of ring column 0; node row of ring row 0
    int q;       // row of the node being updated (its own column is e0+e)
    R h08[3], h17[3], h19[3];  // rows q-1, q, q+1 of the own column
    C inv;
    __device__ __forceinline__ C operator()(int k, int si, int sj) const
    {
        const int d = sj - q + 1;
        if (k == 3) return CodecT<S, C>::dec(d == 0 ? h08[0] : (d == 1 ? h08[1] : h08[2]), inv);
        if (k == 7) return CodecT<S, C>::dec(d == 0 ? h17[0] : (d == 1 ? h17[1] : h17[2]), inv);
        if (k == 8) return CodecT<S, C>::dec(d == 0 ? h19[0] : (d == 1 ? h19[1] : h19[2]), inv);
        return CodecT<S, C>::dec(ring[(sj - r0) & 3][tb_slot(k)][si - e0], inv);
    }
};

// step t+1 of interior node (i, j) from raw pulls, all nine populations
template <class S, class C, int BC, bool OBS, int FL>
__device__ __forceinline__ void tb_node(const Lat &L, const Phys<C> &P, int i, int j,
                                        const typename CodecT<S, C>::raw xr[9], C o[9])
{
    C x[9];
#pragma unroll
    for (int k = 0; k < 9; k++) x[k] = CodecT<S, C>::dec(xr[k], P.inv);
    fix_pulls<S, C, BC, OBS>(L, P, i, j, x);
    relax_k<S, C, BC, FL>(P, x, o);
}

template <class S, class R>
__device__ __forceinline__ void tb_load(const Lat &L, int i, int j, R xr[9])
{
    const uint32_t bo = fresh(((uint32_t)j * (uint32_t)L.rs + (uint32_t)i) * (uint32_t)sizeof(S));
#pragma unroll
    for (int k = 0; k < 8; k++) xr[k] = *(const R *)(pull_base<S>(L, L.a[k], k) + bo);
    xr[8] = *(const R *)(pull_base<S>(L, L.r, 8) + bo);
}

// the strip block of strip bx (columns bx*TW+1..) and rows j0..jend, run
// by TW threads
template <class S, class C, int BC, bool OBS, int FL, int TW = TB_TW>
__device__ __forceinline__ void tb_block(const Lat &L, const Phys<C> &P, int bx, int j0, int jend)
{
    using R = typename CodecT<S, C>::raw;
    constexpr int E = TW + 2;
    constexpr bool XP = Per<BC>::x, YP = Per<BC>::y;
    __shared__ R ring[4][6][E];
    constexpr int RB = TB_RB;
    __shared__ R rim[2][RB][3];
    const int l = L.l, m = L.m;
    const int i0 = bx * TW + 1;
    const int ncols = min(TW, l - i0 + 1);
    const int tid = threadIdx.x;
    // node row / column that a ring position holds, -1 for a halo of a
    // wall-bounded direction
    auto wrow = [&](int r) { return (r >= 1 && r <= m) ? r : (YP ? (r < 1 ? r + m : r - m) : -1); };
    auto wcol = [&](int c) { return (c >= 1 && c <= l) ? c : (XP ? (c < 1 ? c + l : c - l) : -1); };
    TbAcc<S, C, E> acc;
    acc.ring = ring;
    acc.e0 = i0 - 1;
    acc.r0 = j0 - 1;
    acc.inv = P.inv;
    acc.q = j0;
#pragma unroll
    for (int d = 0; d < 3; d++) acc.h08[d] = acc.h17[d] = acc.h19[d] = R{};
    const int e = tid + 1, i = i0 + tid;
    const bool own = tid < ncols;
    // one row r of the march: the pulls of row r are in xa (node row wa);
    // after phase 1 the buffer receives those of row r+1
    auto row = [&](int r, R(&xa)[9], int &wa) {
        const int slot = (r - j0 + 1) & 3;
        if (((r - j0 + 1) & (RB - 1)) == 0) {
            // rim columns, step t+1 of the next RB rows: column i0-1 keeps
            // pops 01,03,05 (pulled by column i0), column i0+ncols pops
            // 10,12,14.  The previous segment's rim rows were all copied
            // into the ring before the barrier of their own row.
            const int nr = min(RB, jend + 2 - r);
            for (int t = tid; t < 2 * nr; t += TW) {
                const int side = t >= nr ? 1 : 0, rr = t - side * nr;
                const int wi = wcol(side ? i0 + ncols : i0 - 1), wj = wrow(r + rr);
                if (wi < 0 || wj < 0) continue;
                R yr[9];
                C o[9];
                tb_load<S>(L, wi, wj, yr);
                tb_node<S, C, BC, OBS, FL>(L, P, wi, wj, yr, o);
                const uint32_t salt = P.rounding ? salt_of(wi, wj) : 0u;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int kk = side ? k + 4 : k;
                    rim[side][rr][k] = encode_s<S, C>(P, o[kk], salt, kk);
                }
            }
            __syncthreads();
        }
        // phase 1: step t+1 of row r, at raised wave priority until the
        // pulls of row r+1 are issued (waves about to issue loads go first)
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int d = 0; d < 2; d++) {
            acc.h08[d] = acc.h08[d + 1];
            acc.h17[d] = acc.h17[d + 1];
            acc.h19[d] = acc.h19[d + 1];
        }
        if (own && wa >= 0) {
            C o[9];
            tb_node<S, C, BC, OBS, FL>(L, P, i, wa, xa, o);
            const uint32_t salt = P.rounding ? salt_of(i, wa) : 0u;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (cx_of(k) != 0) ring[slot][tb_slot(k)][e] = encode_s<S, C>(P, o[k], salt, k);
            acc.h08[2] = encode_s<S, C>(P, o[3], salt, 3);
            acc.h17[2] = encode_s<S, C>(P, o[7], salt, 7);
            acc.h19[2] = encode_s<S, C>(P, o[8], salt, 8);
        }
        if (tid < 6) {
            const int side = tid >= 3 ? 1 : 0, k = tid - side * 3;
            ring[slot][side ? k + 3 : k][side ? ncols + 1 : 0] = rim[side][(r - j0 + 1) & (RB - 1)][k];
        }
        // prefetch the raw pulls of row r+1
        wa = r + 1 <= jend + 1 ? wrow(r + 1) : -1;
        if (own && wa >= 0) tb_load<S>(L, i, wa, xa);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        // phase 2: step t+2 of row q = r-1 (history rows r-2, r-1, r)
        const int q = r - 1;
        if (q < j0 || !own) return;
        acc.q = q;
        C x[9], o[9];
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (cx_of(k) != 0) x[k] = CodecT<S, C>::dec(ring[(q - cy_of(k) - j0 + 1) & 3][tb_slot(k)][e - cx_of(k)], P.inv);
        x[3] = CodecT<S, C>::dec(acc.h08[0], P.inv);  // 08 from (i, q-1)
        x[7] = CodecT<S, C>::dec(acc.h17[2], P.inv);  // 17 from (i, q+1)
        x[8] = CodecT<S, C>::dec(acc.h19[1], P.inv);
        fix_pulls_a<S, C, BC, OBS, true>(acc, L, P, i, q, x);
        relax_k<S, C, BC, FL>(P, x, o);
        const uint32_t bo = fresh(((uint32_t)q * (uint32_t)L.rs + (uint32_t)i) * (uint32_t)sizeof(S));
        const uint32_t salt = P.rounding ? salt_of(i, q) : 0u;
#pragma unroll
        for (int k = 0; k < 8; k++) st_nt((R *)(pull_base<S>(L, L.b[k], 8) + bo), encode_s<S, C>(P, o[k], salt, k));
        st_nt((R *)(pull_base<S>(L, L.rn, 8) + bo), encode_s<S, C>(P, o[8], salt, 8));
    };
    // The same march with no branch around its loads or stores
    // (tb_branch_free): every thread pulls each row (those past the strip
    // from column 1) and stores each row (those past the strip into one
    // slot of the destination set's bottom halo row, which no folded step
    // reads: edge pulls from it are replaced).  vmcnt counts loads and
    // stores in issue order, so the wait for a row's pulls then leaves the
    // previous row's stores in flight; behind a branch that may skip them,
    // the compiler's wait covers those stores too.
    auto safe = [&](int r) { return r < 1 ? 1 : (r > m ? m : r); };
    const int li = own ? i : 1;
    // one slot per block (the lanes past the strip coalesce into one write)
    const uint32_t fill = (uint32_t)((j0 * 61 + i0) % (l + 2)) * (uint32_t)sizeof(S);
    auto row_u = [&](auto p2c, int r, const R(&x)[9], R(&y)[9], int rn) {
        const int slot = (r - j0 + 1) & 3;
        if (((r - j0 + 1) & (RB - 1)) == 0) {
            const int nr = min(RB, jend + 2 - r);
            for (int t = tid; t < 2 * nr; t += TW) {
                const int side = t >= nr ? 1 : 0, rr = t - side * nr;
                const int wi = wcol(side ? i0 + ncols : i0 - 1), wj = wrow(r + rr);
                if (wi < 0 || wj < 0) continue;
                R yr[9];
                C o[9];
                tb_load<S>(L, wi, wj, yr);
                tb_node<S, C, BC, OBS, FL>(L, P, wi, wj, yr, o);
                const uint32_t salt = P.rounding ? salt_of(wi, wj) : 0u;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int kk = side ? k + 4 : k;
                    rim[side][rr][k] = encode_s<S, C>(P, o[kk], salt, kk);
                }
            }
            __syncthreads();
        }
        __builtin_amdgcn_s_setprio(1);  // as in row()
#pragma unroll
        for (int d = 0; d < 2; d++) {
            acc.h08[d] = acc.h08[d + 1];
            acc.h17[d] = acc.h17[d + 1];
            acc.h19[d] = acc.h19[d + 1];
        }
        // the pulls of row r are decoded first, so row rn's are issued
        // before phase 1's collision and stay in flight through both
        // phases (FP16 Taylor-Green 8192^2: 190 k to 199 k MLUPS; in the
        // branched march the same order was slower for the wall and
        // inflow variants, whose edge-fold loads then wait for it too)
        const int wa = wrow(r);
        C xc[9];
#pragma unroll
        for (int k = 0; k < 9; k++) xc[k] = CodecT<S, C>::dec(x[k], P.inv);
        {
            const int wn = wrow(rn);
            tb_load<S>(L, li, wn >= 0 ? wn : safe(rn), y);
        }
        if (own && wa >= 0) {
            C o[9];
            fix_pulls<S, C, BC, OBS>(L, P, i, wa, xc);
            relax_k<S, C, BC, FL>(P, xc, o);
            const uint32_t salt = P.rounding ? salt_of(i, wa) : 0u;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (cx_of(k) != 0) ring[slot][tb_slot(k)][e] = encode_s<S, C>(P, o[k], salt, k);
            acc.h08[2] = encode_s<S, C>(P, o[3], salt, 3);
            acc.h17[2] = encode_s<S, C>(P, o[7], salt, 7);
            acc.h19[2] = encode_s<S, C>(P, o[8], salt, 8);
        }
        if (tid < 6) {
            const int side = tid >= 3 ? 1 : 0, k = tid - side * 3;
            ring[slot][side ? k + 3 : k][side ? ncols + 1 : 0] = rim[side][(r - j0 + 1) & (RB - 1)][k];
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if constexpr (decltype(p2c)::value) {
            const int q = r - 1;
            acc.q = q;
            C xx[9], o[9];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (cx_of(k) != 0)
                    xx[k] = CodecT<S, C>::dec(ring[(q - cy_of(k) - j0 + 1) & 3][tb_slot(k)][e - cx_of(k)], P.inv);
            xx[3] = CodecT<S, C>::dec(acc.h08[0], P.inv);
            xx[7] = CodecT<S, C>::dec(acc.h17[2], P.inv);
            xx[8] = CodecT<S, C>::dec(acc.h19[1], P.inv);
            fix_pulls_a<S, C, BC, OBS, true>(acc, L, P, i, q, xx);
            relax_k<S, C, BC, FL>(P, xx, o);
            const uint32_t bo =
                own ? fresh(((uint32_t)q * (uint32_t)L.rs + (uint32_t)i) * (uint32_t)sizeof(S)) : fill;
            const uint32_t salt = P.rounding ? salt_of(i, q) : 0u;
#pragma unroll
            for (int k = 0; k < 8; k++)
                st_nt((R *)(pull_base<S>(L, L.b[k], 8) + bo), encode_s<S, C>(P, o[k], salt, k));
            st_nt((R *)(pull_base<S>(L, L.rn, 8) + bo), encode_s<S, C>(P, o[8], salt, 8));
        }
    };
    using NoP2 = std::integral_constant<bool, false>;
    using P2 = std::integral_constant<bool, true>;
    R xa[9];
    if constexpr (tb_branch_free<S, BC>()) {
        (void)row;
        {
            const int w0 = wrow(j0 - 1);
            tb_load<S>(L, li, w0 >= 0 ? w0 : safe(j0 - 1), xa);
        }
        row_u(NoP2{}, j0 - 1, xa, xa, j0);
        row_u(NoP2{}, j0, xa, xa, j0 + 1);
        __builtin_amdgcn_s_waitcnt(0);  // no path enters the loop with pulls last
        for (int r = j0 + 1; r <= jend + 1; r++) row_u(P2{}, r, xa, xa, r + 1);
    } else {
        (void)row_u;
        int wa = wrow(j0 - 1);
        if (own && wa >= 0) tb_load<S>(L, i, wa, xa);
        for (int r = j0 - 1; r <= jend + 1; r++) row(r, xa, wa);
    }
}

// Blocks are dispatched to the eight XCDs (each with its own L2) in turn:
// block b runs on XCD b % 8.  tb_xcd_block puts every strip of a row of
// strip blocks on one XCD, so the lines that neighbouring strips both read
// (the rim columns, and the x-shifted pulls across the strip edge) meet in
// one L2.  The grid covers rows rounded up to a multiple of eight; blocks
// past the last row return at once.
constexpr int NXCD = 8;
__device__ __forceinline__ bool tb_xcd_block(int b, int strips, int rows, int &bx, int &by)
{
    const int x = b % NXCD, k = b / NXCD;
    bx = k % strips;
    by = NXCD * (k / strips) + x;
    return by < rows;
}

// box = (first, last strip; first, last row of strip blocks): blocks left
// to k_fused_tb_box (an empty range, x > y, skips none)
template <class S, class C, int BC, bool OBS, int FL = -1>
__global__ __launch_bounds__(tb_tw<S>(), (tb_branch_free<S, BC>() ? 6 : 1)) void k_fused_tb(Lat L, Phys<C> P, int H,
                                                                                           int strips, int rows, int4 box)
{
    int bx, by;
    if (!tb_xcd_block((int)blockIdx.x, strips, rows, bx, by)) return;
    if (bx >= box.x && bx <= box.y && by >= box.z && by <= box.w) return;
    tb_block<S, C, BC, OBS, FL, tb_tw<S>()>(L, fixed_flags<S, C, BC, FL>(P), bx, by * H + 1, min(by * H + H, L.m));
}

// The strip blocks that reach the obstacle, with the obstacle rule: strips
// bx0.. (nbx of them) over rows r0..r1, cut into blocks of hb rows.
// Launched beside k_fused_tb on a second stream: the rule's code raises
// the register count (to ~100) and the time per node, so only these blocks
// carry it, shorter than the others so that they end with them, and the
// grid of the others keeps its occupancy.
template <class S, class C, int BC, int FL = -1>
__global__ __launch_bounds__(tb_tw<S>()) void k_fused_tb_box(Lat L, Phys<C> P, int bx0, int nbx, int r0, int r1, int hb)
{
    const int j0 = r0 + ((int)blockIdx.x / nbx) * hb;
    tb_block<S, C, BC, true, FL, tb_tw<S>()>(L, fixed_flags<S, C, BC, FL>(P), bx0 + (int)blockIdx.x % nbx, j0,
                                             min(j0 + hb - 1, r1));
}

// ------------------------------------------------------------ diagnostics
// One block per row j; each thread accumulates a fixed strided subset of
// the row in FP64 and the block reduces in a fixed order (wavefront
// shuffles, then the wavefronts' partials), so results depend only on
// (l, m).  Per row: diagno sums (rho, u, v, u2+v2 with diagno's
// per-node formulas), varm sums (rho, u, v with varm's formulas), the max
// of u2+v2 and a non-finite counter.
constexpr int ROWS_T = 256;
constexpr int ROW_VALS = 9;

// Fixed-order block reduction of N FP64 values per thread (value MX by max,
// the others by sum): each 64-lane wavefront folds its lanes with shuffles
// (offsets 32, 16, .., 1), lane 0 of each wavefront leaves its N partials in
// sh[N][ROWS_T / 64], and thread q < N combines value q over the wavefronts
// in order into acc[0].
template <int N, int MX>
__device__ __forceinline__ void block_reduce(double (&acc)[N], double *sh)
{
    constexpr int NW = ROWS_T / 64;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int q = 0; q < N; q++) {
            const double o = __shfl_down(acc[q], off, 64);
            acc[q] = q == MX ? fmax(acc[q], o) : acc[q] + o;
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) sh[q * NW + w] = acc[q];
    }
    __syncthreads();
    if ((int)threadIdx.x < N) {
        const int q = threadIdx.x;
        double t = sh[q * NW];
        for (int k = 1; k < NW; k++) t = q == MX ? fmax(t, sh[q * NW + k]) : t + sh[q * NW + k];
        acc[0] = t;
    }
}

template <class S, class C>
__global__ __launch_bounds__(ROWS_T) void k_rows(Lat L, Phys<C> P, double *out)
{
    const int j = blockIdx.x + 1;
    const C cte1 = P.noshift ? (C)0.0 : (C)1.0;  // NOSHIFT: rho = sum
    double acc[ROW_VALS] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x + 1; i <= L.l; i += ROWS_T) {
        C a01 = get<S, C>(L, P, L.a, 0, i, j), a03 = get<S, C>(L, P, L.a, 1, i, j);
        C a05 = get<S, C>(L, P, L.a, 2, i, j), a08 = get<S, C>(L, P, L.a, 3, i, j);
        C a10 = get<S, C>(L, P, L.a, 4, i, j), a12 = get<S, C>(L, P, L.a, 5, i, j);
        C a14 = get<S, C>(L, P, L.a, 6, i, j), a17 = get<S, C>(L, P, L.a, 7, i, j);
        C a19 = get<S, C>(L, P, L.a, 8, i, j);
        // diagno (SRC/diagno.F90:63-96)
        C rho = (((+a01 + a03 + a10 + a12) + (+a05 + a08 + a14 + a17)) + a19) + cte1;
        C rhoinv = (C)1.0 / rho;
        C xj = ((a03 - a12) + (a01 - a10) + (a05 - a14)) * rhoinv;
        C yj = ((a03 - a01) + (a12 - a10) + (a08 - a17)) * rhoinv;
        C s = xj * xj + yj * yj;
        acc[0] += (double)rho;
        acc[1] += (double)xj;
        acc[2] += (double)yj;
        acc[3] += (double)s;
        // varm (SRC/varm.F90:54-71)
        C rv = (+a01 + a03 + a05 + a08 + a10 + a12 + a14 + a17 + a19) + cte1;
        C rvi = (C)1.0 / rv;
        acc[4] += (double)rv;
        acc[5] += (double)((a01 + a03 + a05 - a10 - a12 - a14) * rvi);
        acc[6] += (double)((a03 + a08 + a12 - a01 - a10 - a17) * rvi);
        double sd = (double)s;
        acc[7] = sd > acc[7] ? sd : acc[7];
        if (!(sd <= 1.0e300) || !((double)rho <= 1.0e300) || !((double)rho >= -1.0e300)) acc[8] += 1.0;
    }
    __shared__ double sh[ROW_VALS][ROWS_T / 64];
    block_reduce<ROW_VALS, 7>(acc, &sh[0][0]);
    if (threadIdx.x < ROW_VALS) out[(long long)(j - 1) * ROW_VALS + threadIdx.x] = acc[0];
}

// sums of the per-row values in a fixed order: out = {Srho, Su, Sv, Ss,
// max s, nonfinite}
template <int = 0>
__global__ __launch_bounds__(ROWS_T) void k_rows_final(const double *rows, int m, double *out)
{
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = threadIdx.x; j < m; j += ROWS_T) {
        const double *r = rows + (long long)j * ROW_VALS;
        acc[0] += r[0];
        acc[1] += r[1];
        acc[2] += r[2];
        acc[3] += r[3];
        acc[4] = fmax(acc[4], r[7]);
        acc[5] += r[8];
    }
    __shared__ double sh[6][ROWS_T / 64];
    block_reduce<6, 4>(acc, &sh[0][0]);
    if (threadIdx.x < 6) out[threadIdx.x] = acc[0];
}

// prof_i (mode 0, row jc, n = l), prof_j (mode 1, column ic, n = m),
// probe (mode 2, one node); out = u[n], v[n], rho[n]
template <class S, class C>
__global__ void k_line(Lat L, Phys<C> P, int mode, int c0, int c1, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = mode == 0 ? L.l : (mode == 1 ? L.m : 1);
    if (t >= n) return;
    int i, j;
    if (mode == 0) { i = t + 1; j = c0; }
    else if (mode == 1) { i = c0; j = t + 1; }
    else { i = c0; j = c1; }
    const C cte1 = P.noshift ? (C)0.0 : (C)1.0;  // NOSHIFT: rho = sum
    C a01 = get<S, C>(L, P, L.a, 0, i, j), a03 = get<S, C>(L, P, L.a, 1, i, j);
    C a05 = get<S, C>(L, P, L.a, 2, i, j), a08 = get<S, C>(L, P, L.a, 3, i, j);
    C a10 = get<S, C>(L, P, L.a, 4, i, j), a12 = get<S, C>(L, P, L.a, 5, i, j);
    C a14 = get<S, C>(L, P, L.a, 6, i, j), a17 = get<S, C>(L, P, L.a, 7, i, j);
    C a19 = get<S, C>(L, P, L.a, 8, i, j);
    C u, v, den;
    if (mode == 1) {  // SRC/prof_j.F90:47-67
        den = ((+a01 + a03 + a05 + a08 + a10 + a12 + a14 + a17) + a19) + cte1;
        u = ((a01 - a10) + (a03 - a12) + (a05 - a14)) / den;
        v = ((a03 - a01) + (a08 - a17) + (a12 - a10)) / den;
    } else {  // SRC/prof_i.F90:47-63 and SRC/probe.F90:46-62
        den = (+a01 + a03 + a05 + a08 + a10 + a12 + a14 + a17 + a19) + cte1;
        u = (a01 + a03 + a05 - a10 - a12 - a14) / den;
        v = (a03 + a08 + a12 - a01 - a10 - a17) / den;
    }
    out[t] = (double)u;
    out[n + t] = (double)v;
    out[2 * n + t] = (double)den;
}

// population k (0..8) <-> dense (l+2)*(m+2) array, x fastest
// Momentum-exchange force on the obstacle (SRC/draglift.F90:58-98): over
// the fluid nodes of the box [i0,i1] x [j0,j1], 2 c_k a_k(i,j) for every
// population k whose neighbour (i,j)+c_k is solid; x and y components,
// per-link terms in the compute kind, summed in FP64 by one block in a
// fixed order.
constexpr int DL_T = 256;
template <class S, class C>
__global__ __launch_bounds__(DL_T) void k_draglift(Lat L, Phys<C> P, int i0, int i1, int j0, int j1, double *out)
{
    __shared__ double sx[DL_T], sy[DL_T];
    const int w = i1 - i0 + 1, n = w * (j1 - j0 + 1);
    double fx = 0.0, fy = 0.0;
    for (int t = threadIdx.x; t < n; t += DL_T) {
        const int i = i0 + t % w, j = j0 + t / w;
        if (solid(L, i, j)) continue;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (!solid(L, i + cx_of(k), j + cy_of(k))) continue;
            const C v = get<S, C>(L, P, L.a, k, i, j);
            fx += (double)((C)2.0 * (C)cx_of(k) * v);
            fy += (double)((C)2.0 * (C)cy_of(k) * v);
        }
    }
    sx[threadIdx.x] = fx;
    sy[threadIdx.x] = fy;
    __syncthreads();
    for (int h = DL_T / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sx[threadIdx.x] += sx[threadIdx.x + h];
            sy[threadIdx.x] += sy[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sx[0];
        out[1] = sy[0];
    }
}

// Macroscopic fields of the VTK writers (SRC/vtk_xy.binary.F90:60-140):
// per interior node, the population sums in the storage kind (FP32 for the
// 16-bit kinds) rounded to FP32, then u/den, w/den and den in FP32; dense
// l*m arrays, x fastest.
template <class S>
struct SumKind {
    using T = float;
};
template <>
struct SumKind<double> {
    using T = double;
};
template <class S, class C>
__global__ void k_fields(Lat L, Phys<C> P, float *u, float *v, float *rho)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > L.l || j > L.m) return;
    using T = typename SumKind<S>::T;
    T a[9];
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = (T)get<S, C>(L, P, L.a, k, i, j);
    const float us = (float)(a[0] + a[1] + a[2] - a[4] - a[5] - a[6]);
    const float ws = (float)(a[1] + a[3] + a[5] - a[0] - a[4] - a[7]);
    const float den = (float)((a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7] + a[8]) +
                              (P.noshift ? (T)0.0 : (T)1.0));
    const size_t q = (size_t)(i - 1) + (size_t)(j - 1) * L.l;
    u[q] = us / den;
    v[q] = ws / den;
    rho[q] = den;
}

template <class S>
__global__ void k_pack(Lat L, const void *plane, S *dense)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > L.l + 1 || j > L.m + 1) return;
    dense[(long long)j * (L.l + 2) + i] = elc<S>(plane, L.at(i, j));
}
template <class S>
__global__ void k_unpack(Lat L, void *plane, const S *dense)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > L.l + 1 || j > L.m + 1) return;
    el<S>(plane, L.at(i, j)) = dense[(long long)j * (L.l + 2) + i];
}

}  // namespace bgk
