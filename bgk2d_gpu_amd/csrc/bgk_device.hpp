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
// their sub-ulp per-step changes systematically.  The offset is derived from
// the value's own bits and its position (i, j) and population k, so results
// are deterministic and independent of launch grouping.  The conversion
// itself is gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32: it adds the top 13
// (FP16) or 16 (BF16) bits of its random operand to the magnitude below the
// kept bits (aligned to the top of the dropped bits for half subnormals) and
// truncates; NaN stays NaN, overflow gives inf.
__host__ __device__ inline uint32_t mul24(uint32_t a, uint32_t b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __umul24(a, b);
#else
    return (a & 0xffffffu) * (b & 0xffffffu);
#endif
}
// position salt of node (i, j): two 24-bit multiplies (full rate on CDNA),
// computed once per node
__host__ __device__ inline uint32_t salt_of(int i, int j)
{
    return mul24((uint32_t)i, 0x9E3779u) + mul24((uint32_t)j, 0x85EBCAu);
}
// odd 24-bit multiplier of population k's random word
__host__ __device__ constexpr uint32_t sr_mul(int k)
{
    return k == 0 ? 0xB5297Bu : k == 1 ? 0x9E3779u : k == 2 ? 0xC2B2AFu : k == 3 ? 0x85EBCBu : k == 4 ? 0xD35A2Du
         : k == 5 ? 0xA3B195u : k == 6 ? 0xE6546Bu : k == 7 ? 0x8B1E5Fu : 0xF1357Bu;
}
// per-store random word: one 24-bit multiply-add (v_mad_u32_u24) of the low
// bits of the stored value, which change from step to step, and the node's
// salt; the conversion takes its top bits
__host__ __device__ inline uint32_t sr_word(uint32_t bits, uint32_t salt, int k)
{
    return mul24(bits, sr_mul(k)) + salt;
}

typedef __fp16 sr_half2 __attribute__((ext_vector_type(2)));
typedef short sr_short2 __attribute__((ext_vector_type(2)));

template <class S, class C>
struct Codec;

template <>
struct Codec<float, float> {
    using raw = float;
    __device__ __forceinline__ static float dec(float v, float) { return v; }
    __device__ __forceinline__ static float enc(float v, float, int, uint32_t, int) { return v; }
};
template <>
struct Codec<double, double> {
    using raw = double;
    __device__ __forceinline__ static double dec(double v, double) { return v; }
    __device__ __forceinline__ static double enc(double v, double, int, uint32_t, int) { return v; }
};
template <>
struct Codec<float, double> {  // MIXED2
    using raw = float;
    __device__ __forceinline__ static double dec(float v, double) { return (double)v; }
    __device__ __forceinline__ static float enc(double v, double, int, uint32_t, int) { return (float)v; }
};
template <>
struct Codec<_Float16, float> {
    using raw = _Float16;
    // dec/enc take the scale factor that applies: inv_scale / scale
    __device__ __forceinline__ static float dec(_Float16 v, float inv) { return (float)v * inv; }
    __device__ __forceinline__ static _Float16 enc(float v, float sc, int sr, uint32_t salt, int k)
    {
        const float x = v * sc;
        if (sr) {
            const sr_half2 h = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(sr_half2, x), x,
                                                               sr_word(__float_as_uint(x), salt, k), false);
            return (_Float16)h[0];
        }
        return (_Float16)x;
    }
    // two values stored as the halves of one dword: enc_lo converts the
    // first into the low half (the high half is left undefined with
    // stochastic rounding, zero otherwise), enc_hi the second into the high
    // half of that word in place
    __device__ __forceinline__ static uint32_t enc_lo(float v, float sc, int sr, uint32_t salt, int k)
    {
        const float x = v * sc;
        if (sr)
            return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(sr_half2, x), x,
                                                                                sr_word(__float_as_uint(x), salt, k),
                                                                                false));
        return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
    }
    __device__ __forceinline__ static uint32_t enc_hi(uint32_t w, float v, float sc, int sr, uint32_t salt, int k)
    {
        const float x = v * sc;
        if (sr)
            return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(sr_half2, w), x,
                                                                                sr_word(__float_as_uint(x), salt, k),
                                                                                true));
        return w | ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x) << 16);
    }
};
template <>
struct Codec<bf16_t, float> {
    using raw = bf16_t;
    __device__ __forceinline__ static float dec(bf16_t v, float inv)
    {
        return __uint_as_float((uint32_t)v.bits << 16) * inv;
    }
    __device__ __forceinline__ static uint16_t rne(float f)
    {
        uint32_t x = __float_as_uint(f);
        if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
        x += 0x7fffu + ((x >> 16) & 1u);
        return (uint16_t)(x >> 16);
    }
    __device__ __forceinline__ static bf16_t enc(float v, float sc, int sr, uint32_t salt, int k)
    {
        const float f = v * sc;
        bf16_t r;
        if (sr) {
            const sr_short2 h = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(sr_short2, f), f,
                                                                 sr_word(__float_as_uint(f), salt, k), false);
            r.bits = (uint16_t)h[0];
        } else {
            r.bits = rne(f);
        }
        return r;
    }
    __device__ __forceinline__ static uint32_t enc_lo(float v, float sc, int sr, uint32_t salt, int k)
    {
        const float x = v * sc;
        if (sr)
            return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(sr_short2, x), x,
                                                                                 sr_word(__float_as_uint(x), salt, k),
                                                                                 false));
        return (uint32_t)rne(x);
    }
    __device__ __forceinline__ static uint32_t enc_hi(uint32_t w, float v, float sc, int sr, uint32_t salt, int k)
    {
        const float x = v * sc;
        if (sr)
            return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(sr_short2, w), x,
                                                                                 sr_word(__float_as_uint(x), salt, k),
                                                                                 true));
        return w | ((uint32_t)rne(x) << 16);
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
    // LES terms that depend on omega and cteS only, evaluated once on the
    // host in the compute kind with the reference's operations:
    // 1/(2 omega), (1/omega)^2, 18 cteS^2
    C les_h, les_i2, les_c;
    // kernels that compute on scaled populations (scale*(f - w) as stored,
    // see fixed_flags): the storage scale and its inverse; 1 otherwise
    C usc, dsc;
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
// as SRC/col_MC.F90:125-293 (and SRC/col.F90:104-237).
//   FORCE : the kernel may serve CHANNEL, whose body force P.channel_force
//           selects; without it the force terms of the reference's
//           non-CHANNEL builds are zeros.
//   FL    : -1 the form is P.form; 0 or 1 that form, fixed at compile time.
//   SCALED: x and o are the populations as stored, scale*(f - w) (P.usc =
//           scale, P.dsc = 1/scale; form 1 only).  Scaling by a power of two
//           commutes with every rounding here, so each result is exactly
//           scale times the unscaled one: the store needs no multiply.
//   Z0    : storage kind = compute kind.  No stored population is then ever
//           -0 (a difference a - b is -0 only for a = -0), so adding the
//           +0 force of a non-CHANNEL build changes nothing and is left out.
// The cte0 + q terms of the four non-rest equilibria are left out too: q is
// qf*(3a*a - vsq) with 3a*a >= +0, never -0, so 0 + q = q.  The rest
// population's q0 = cte0 + qf*(-vsq) keeps it (-vsq is -0 at rest).
// Form 1 (deviation form, an extension) has no force terms at all outside
// CHANNEL.
template <class C, bool FORCE = true, int FL = -1, bool SCALED = false, bool Z0 = false>
__device__ __forceinline__ void relax(const Phys<C> &P, const C x[9], C o[9])
{
    static_assert(!SCALED || FL == 1, "scaled populations need the deviation form");
    const int form = FL >= 0 ? FL : P.form;
    const C cte1 = (C)1.0, cte0 = (C)1.0 - cte1;
    const C rf = K<C>::rf, qf = K<C>::qf, p0 = K<C>::p0, p1 = K<C>::p1, p2 = K<C>::p2;
    const C x01 = x[0], x03 = x[1], x05 = x[2], x08 = x[3], x10 = x[4];
    const C x12 = x[5], x14 = x[6], x17 = x[7], x19 = x[8];
    // deviation form: drho = sum of the shifted populations is kept apart
    // from the 1, and the equilibria use w*drho instead of (rho*w - w), so
    // the O(u^2) density deviation is not rounded against 1
    const C drs = ((x01 + x03 + x10 + x12) + (x05 + x08 + x14 + x17)) + x19;  // scaled when SCALED
    const C drho = SCALED ? drs * P.dsc : drs;
    C rho = form ? (C)1.0 + drho : drho + cte1;
    C rhoinv = (C)1.0 / rho;
    const C rinv = SCALED ? rhoinv * P.dsc : rhoinv;
    C vx = ((x03 - x12) + (x01 - x10) + (x05 - x14)) * rinv;
    C vy = ((x03 - x01) + (x12 - x10) + (x08 - x17)) * rinv;
    C vx2 = vx * vx, vy2 = vy * vy;
    C vsq = vx2 + vy2;
    C vxpy = vx + vy, vxmy = vx - vy;
    C qxpy = qf * ((C)3.0 * vxpy * vxpy - vsq);
    C qxmy = qf * ((C)3.0 * vxmy * vxmy - vsq);
    C qx = qf * ((C)3.0 * vx2 - vsq);
    C qy = qf * ((C)3.0 * vy2 - vsq);
    C q0 = cte0 + qf * (-vsq);
    vx = rf * vx;
    vy = rf * vy;
    vxpy = rf * vxpy;
    vxmy = rf * vxmy;
    const C rhoe = SCALED ? rho * P.usc : rho;
    C rp0 = rhoe * p0, rp1 = rhoe * p1, rp2 = rhoe * p2;
    // the density term cte1*(rp - p) of the reference, or p*drho
    const C d0 = form ? p0 * drs : cte1 * (rp0 - p0);
    const C d1 = form ? p1 * drs : cte1 * (rp1 - p1);
    const C d2 = form ? p2 * drs : cte1 * (rp2 - p2);
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
        C ptot = sqrt(pxx * pxx + (C)2.0 * pxy * pyx + pyy * pyy);
        if (SCALED) ptot = ptot * P.dsc;
        // Ts = 1/(2 omega1) + sqrt(18 cteS^2 P + (1/omega1)^2) / 2
        const C ts = P.les_h + sqrt(P.les_c * ptot + P.les_i2) / (C)2.0;
        om = (C)1.0 / ts;
    }
    C t[9];
    t[0] = x01 - om * (x01 - e01);
    t[1] = x03 - om * (x03 - e03);
    t[2] = x05 - om * (x05 - e05);
    t[3] = x08 - om * (x08 - e08);
    t[4] = x10 - om * (x10 - e10);
    t[5] = x12 - om * (x12 - e12);
    t[6] = x14 - om * (x14 - e14);
    t[7] = x17 - om * (x17 - e17);
    o[8] = x19 - om * (x19 - e19);
    const bool fon = FORCE && P.channel_force != 0;
    if (fon || (form == 0 && !Z0)) {
        // forcing term (SRC/col_MC.F90:266-274, SRC/col.F90:210-218):
        // fgrad*rho on x, or on y in the FORCING_Y build
        C forcex = (C)0.0, forcey = (C)0.0;
        if (FORCE && P.channel_force == 1) forcex = P.fgrad * rhoe;
        if (FORCE && P.channel_force == 2) forcey = P.fgrad * rhoe;
        o[0] = t[0] + (+forcex - forcey);
        o[1] = t[1] + (+forcex + forcey);
        o[2] = t[2] + (+forcex);
        o[3] = t[3] + (+forcey);
        o[4] = t[4] + (-forcex - forcey);
        o[5] = t[5] + (-forcex + forcey);
        o[6] = t[6] + (-forcex);
        o[7] = t[7] + (-forcey);
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) o[k] = t[k];
    }
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
