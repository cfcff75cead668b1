// bgk_device.hpp -- device-side building blocks of the gfx950 BGK2D path.
//
// Layout in HBM (one allocation per context):
//   17 planes: pops 01,03,05,08,10,12,14,17 of set 0, the same of set 1,
//   and the rest population 19 (a single plane updated in place, as
//   SRC/col_MC.F90:293 does).  A plane holds (m+2) rows of `pitch`
//   elements; node (i,j), 0 <= i <= l+1, 0 <= j <= m+1, sits at element
//   off + j*pitch + i with off = (128 B / element size) - 1, so that the
//   first interior node i = 1 of every row starts a 128-byte line and a
//   64-lane wavefront reading 64 consecutive nodes is fully coalesced.
//
// All arithmetic follows the reference's operation order and is compiled
// with -ffp-contract=off, so every value is bit-identical to the reference
// CPU build on the same inputs (x86-64 flang makes no FMA either).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bgk {

// population index k = 0..8  <->  01 03 05 08 10 12 14 17 19
__host__ __device__ constexpr int cx_of(int k) { return k < 3 ? 1 : (k == 3 ? 0 : (k < 7 ? -1 : 0)); }
__host__ __device__ constexpr int cy_of(int k)
{
    return k == 0 ? -1 : k == 1 ? 1 : k == 2 ? 0 : k == 3 ? 1 : k == 4 ? -1 : k == 5 ? 1 : k == 6 ? 0 : k == 7 ? -1 : 0;
}
__host__ __device__ constexpr int opp_of(int k)
{
    return k == 0 ? 5 : k == 1 ? 4 : k == 2 ? 6 : k == 3 ? 7 : k == 4 ? 1 : k == 5 ? 0 : k == 6 ? 2 : k == 7 ? 3 : 8;
}

// ------------------------------------------------------------ storage
// A codec maps one stored element to the compute kind and back.  FP16 and
// BF16 hold scale*(f - w) with a power-of-two scale; scaling by a power
// of two is exact, so scale = 1 reproduces the reference's MIXED1 storage.
struct bf16_t {
    uint16_t bits;
};

// Stochastic rounding of FP16/BF16 stores (opt-in): the bits below the
// target precision get a pseudo-random offset before truncation, which makes
// the rounding unbiased; with round-to-nearest, slowly varying fields lose
// their sub-ulp per-step changes systematically.  The offset is a hash of
// the value's own bits and its position (i, j, population), so results are
// deterministic and independent of launch grouping.  The conversion itself
// is gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32: it adds the top 13
// (FP16) or 16 (BF16) bits of its random operand to the magnitude below the
// kept bits (aligned to the top of the dropped bits for half subnormals)
// and truncates; NaN stays NaN, overflow gives inf.  One instruction in
// place of the range checks, add, mask and conversion.
// position salt: 24-bit multiplies (full rate on CDNA) and a Weyl step per
// population; the per-store hash below does the mixing
__host__ __device__ inline uint32_t mul24(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umul24(a, b);
#else
    return (a & 0xffffffu) * (b & 0xffffffu);
#endif
}
__host__ __device__ inline uint32_t salt_of(int i, int j, int k)
{
    return mul24((uint32_t)i, 0x9E3779u) + mul24((uint32_t)j, 0x85EBCAu) + (uint32_t)k * 0x9E3779B9u;
}
// per-store random word: one 24-bit multiply (full rate on CDNA) of the
// low bits of value^salt, which change from step to step; the conversion
// takes its top bits, the best-mixed bits of the product
__host__ __device__ inline uint32_t sr_word(uint32_t bits, uint32_t salt)
{
    return mul24(bits ^ salt, 0xB5297Au);
}

typedef __fp16 sr_half2 __attribute__((ext_vector_type(2)));
typedef short sr_short2 __attribute__((ext_vector_type(2)));

template <class S, class C>
struct Codec;

template <>
struct Codec<float, float> {
    using raw = float;
    __device__ __forceinline__ static float dec(float v, float) { return v; }
    __device__ __forceinline__ static float enc(float v, float, int, uint32_t) { return v; }
};
template <>
struct Codec<double, double> {
    using raw = double;
    __device__ __forceinline__ static double dec(double v, double) { return v; }
    __device__ __forceinline__ static double enc(double v, double, int, uint32_t) { return v; }
};
template <>
struct Codec<float, double> {  // MIXED2
    using raw = float;
    __device__ __forceinline__ static double dec(float v, double) { return (double)v; }
    __device__ __forceinline__ static float enc(double v, double, int, uint32_t) { return (float)v; }
};
template <>
struct Codec<_Float16, float> {
    using raw = _Float16;
    // dec/enc take the scale factor that applies: inv_scale / scale
    __device__ __forceinline__ static float dec(_Float16 v, float inv) { return (float)v * inv; }
    __device__ __forceinline__ static _Float16 enc(float v, float sc, int sr, uint32_t salt)
    {
        const float x = v * sc;
        if (sr) {
            const sr_half2 h = __builtin_amdgcn_cvt_sr_f16_f32(sr_half2{}, x, sr_word(__float_as_uint(x), salt), false);
            return (_Float16)h[0];
        }
        return (_Float16)x;
    }
};
template <>
struct Codec<bf16_t, float> {
    using raw = bf16_t;
    __device__ __forceinline__ static float dec(bf16_t v, float inv)
    {
        return __uint_as_float((uint32_t)v.bits << 16) * inv;
    }
    __device__ __forceinline__ static bf16_t enc(float v, float sc, int sr, uint32_t salt)
    {
        const float f = v * sc;
        uint32_t x = __float_as_uint(f);
        bf16_t r;
        if (sr) {
            const sr_short2 h = __builtin_amdgcn_cvt_sr_bf16_f32(sr_short2{}, f, sr_word(x, salt), false);
            r.bits = (uint16_t)h[0];
        } else if ((x & 0x7fffffffu) > 0x7f800000u) {
            r.bits = (uint16_t)((x >> 16) | 0x40u);
        } else {
            x += 0x7fffu + ((x >> 16) & 1u);
            r.bits = (uint16_t)(x >> 16);
        }
        return r;
    }
};

// ---------------------------------------------------------- descriptors
struct Lat {
    void *a[8];  // current set, pops 01..17
    void *b[8];  // other set (FUSED target / ORIGINAL streamed copy)
    void *r;     // pop 19 (current)
    void *rn;    // second pop-19 plane: target of the two-step kernel
    const uint8_t *obs;  // l*m solid mask (x fastest) or null
    int l, m;
    int pitch, off;
    long long rs;                // elements from row j to row j+1 of one population
    int imin, imax, jmin, jmax;  // obstacle bounding box (1-based)
    int stepx, stepy;
    __device__ __forceinline__ long long at(int i, int j) const { return (long long)off + (long long)j * rs + i; }
};

template <class C>
struct Phys {
    C omega, fgrad, force;  // hencol / bcond_driven constants
    C sc, inv;              // storage scale and its inverse (FP16/BF16)
    C inflow[3];            // stored-and-decoded inflow equilibria 01,03,05
    C umax;                 // OPEN: peak of the parabolic inflow profile
    int channel_force;      // CHANNEL body force in collision: 1 along x,
                            // 2 along y (the FORCING_Y build)
    int form;               // 0: reference operation order; 1: deviation form
    int rounding;           // FP16/BF16 stores: 0 nearest-even, 1 stochastic
    int les;                // 1: Smagorinsky LES relaxation rate per node
    int trick1;             // 1: DRIVEN top corners take the side walls' copy
    C ctes;                 // LES: Smagorinsky constant cteS
};

// ---------------------------------------------------------- arithmetic
// Reference constants (SRC/storage.F90:44-73,130-138): evaluated in double
// precision, then converted to the compute kind.
template <class C>
struct K {
    static constexpr C rf = (C)3.0;
    static constexpr C qf = (C)1.5;
    static constexpr C p0 = (C)(4.0 / 9.0);
    static constexpr C p1 = (C)(1.0 / 9.0);
    static constexpr C p2 = (C)(1.0 / (3.0 * 3.0 * (1.0 + 3.0)));
};

// BGK relaxation of one node in the shifted form, operation for operation
// as SRC/col_MC.F90:125-293 (and SRC/col.F90:104-237).  FORCE: the kernel
// may serve CHANNEL, whose body force P.channel_force selects; elsewhere
// both force terms are the compile-time zeros of the reference's
// non-CHANNEL builds.
template <class C, bool FORCE = true>
__device__ __forceinline__ void relax(const Phys<C> &P, const C x[9], C o[9])
{
    const C cte1 = (C)1.0, cte0 = (C)1.0 - cte1;
    const C rf = K<C>::rf, qf = K<C>::qf, p0 = K<C>::p0, p1 = K<C>::p1, p2 = K<C>::p2;
    const C x01 = x[0], x03 = x[1], x05 = x[2], x08 = x[3], x10 = x[4];
    const C x12 = x[5], x14 = x[6], x17 = x[7], x19 = x[8];
    // deviation form: drho = sum of the shifted populations is kept apart
    // from the 1, and the equilibria use w*drho instead of (rho*w - w), so
    // the O(u^2) density deviation is not rounded against 1
    const C drho = ((x01 + x03 + x10 + x12) + (x05 + x08 + x14 + x17)) + x19;
    C rho = P.form ? (C)1.0 + drho : drho + cte1;
    C rhoinv = (C)1.0 / rho;
    C vx = ((x03 - x12) + (x01 - x10) + (x05 - x14)) * rhoinv;
    C vy = ((x03 - x01) + (x12 - x10) + (x08 - x17)) * rhoinv;
    C vx2 = vx * vx, vy2 = vy * vy;
    C vsq = vx2 + vy2;
    C vxpy = vx + vy, vxmy = vx - vy;
    C qxpy = cte0 + qf * ((C)3.0 * vxpy * vxpy - vsq);
    C qxmy = cte0 + qf * ((C)3.0 * vxmy * vxmy - vsq);
    C qx = cte0 + qf * ((C)3.0 * vx2 - vsq);
    C qy = cte0 + qf * ((C)3.0 * vy2 - vsq);
    C q0 = cte0 + qf * (-vsq);
    vx = rf * vx;
    vy = rf * vy;
    vxpy = rf * vxpy;
    vxmy = rf * vxmy;
    C rp0 = rho * p0, rp1 = rho * p1, rp2 = rho * p2;
    // the density term cte1*(rp - p) of the reference, or p*drho
    const C d0 = P.form ? p0 * drho : cte1 * (rp0 - p0);
    const C d1 = P.form ? p1 * drho : cte1 * (rp1 - p1);
    const C d2 = P.form ? p2 * drho : cte1 * (rp2 - p2);
    C e01 = rp2 * (+vxmy + qxmy) + d2;
    C e03 = rp2 * (+vxpy + qxpy) + d2;
    C e05 = rp1 * (+vx + qx) + d1;
    C e08 = rp1 * (+vy + qy) + d1;
    C e10 = rp2 * (-vxpy + qxpy) + d2;
    C e12 = rp2 * (-vxmy + qxmy) + d2;
    C e14 = rp1 * (-vx + qx) + d1;
    C e17 = rp1 * (-vy + qy) + d1;
    C e19 = rp0 * (+q0) + d0;
    C om = P.omega;
    if (P.les) {
        // Smagorinsky LES (SRC/col_MC.F90:178-230, SRC/col.F90:154-205):
        // the relaxation rate from the non-equilibrium stress; omega1 is the
        // molecular rate of hencol (= P.omega)
        const C n01 = x01 - e01, n03 = x03 - e03, n05 = x05 - e05, n08 = x08 - e08;
        const C n10 = x10 - e10, n12 = x12 - e12, n14 = x14 - e14, n17 = x17 - e17;
        const C pxx = n01 + n03 + n05 + n10 + n12 + n14;
        const C pyy = n01 + n03 + n08 + n10 + n12 + n17;
        const C pxy = -n01 + n03 + n10 - n12;
        const C pyx = pxy;
        const C ptot = sqrt(pxx * pxx + (C)2.0 * pxy * pyx + pyy * pyy);
        const C omega1 = P.omega, cs = P.ctes;
        const C inv1 = (C)1.0 / omega1;
        const C ts = (C)1.0 / ((C)2.0 * omega1) + sqrt((C)18.0 * (cs * cs) * ptot + inv1 * inv1) / (C)2.0;
        om = (C)1.0 / ts;
    }
    // forcing term (SRC/col_MC.F90:266-274, SRC/col.F90:210-218): fgrad*rho
    // on x, or on y in the FORCING_Y build
    C forcex = (C)0.0, forcey = (C)0.0;
    if (FORCE && P.channel_force == 1) forcex = P.fgrad * rho;
    if (FORCE && P.channel_force == 2) forcey = P.fgrad * rho;
    o[0] = x01 - om * (x01 - e01) + (+forcex - forcey);
    o[1] = x03 - om * (x03 - e03) + (+forcex + forcey);
    o[2] = x05 - om * (x05 - e05) + (+forcex);
    o[3] = x08 - om * (x08 - e08) + (+forcey);
    o[4] = x10 - om * (x10 - e10) + (-forcex - forcey);
    o[5] = x12 - om * (x12 - e12) + (-forcex + forcey);
    o[6] = x14 - om * (x14 - e14) + (-forcex);
    o[7] = x17 - om * (x17 - e17) + (-forcey);
    o[8] = x19 - om * (x19 - e19);
}

// Shifted equilibrium at rho = 1 with the quadratic part in double
// precision (the 3.d0 literal of SRC/init.F90:114-122 and
// SRC/bcond.inflow.F90:104-106).
template <class C>
__host__ __device__ inline C eq_d(C cu, C cvsq)
{
    double q = 3.0 * (double)cu * (double)cu - (double)cvsq;
    return (C)((double)(K<C>::rf * cu) + (double)K<C>::qf * q);
}

}  // namespace bgk
