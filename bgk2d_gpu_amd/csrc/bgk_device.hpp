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
// per-store random word: the low bits of the stored value, which change
// from step to step, mixed with the node's salt, times population k's
// multiplier (a xor and a full-rate 24-bit multiply); the conversion takes
// its top bits.  Mixing the salt in before the multiply makes the word a
// non-linear function of the dropped bits (adding it after the multiply
// left the word linear in them and the BF16 error 3.6 times larger).
__host__ __device__ inline uint32_t sr_word(uint32_t bits, uint32_t salt, int k)
{
    return mul24(bits ^ salt, sr_mul(k));
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
    int noshift;  // 1: whole populations f (the NOSHIFT build)
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

// Smagorinsky LES rate (SRC/col_MC.F90:178-230, SRC/col.F90:154-205): the
// relaxation rate from the non-equilibrium stress of populations x with
// equilibria e; omega1 is the molecular rate of hencol.  With SCALED inputs
// the stress is rescaled once (exact, a power of two).
template <class C, bool SCALED>
__device__ __forceinline__ C les_rate(const Phys<C> &P, const C x[9], const C e[9])
{
    const C n01 = x[0] - e[0], n03 = x[1] - e[1], n05 = x[2] - e[2], n08 = x[3] - e[3];
    const C n10 = x[4] - e[4], n12 = x[5] - e[5], n14 = x[6] - e[6], n17 = x[7] - e[7];
    const C pxx = n01 + n03 + n05 + n10 + n12 + n14;
    const C pyy = n01 + n03 + n08 + n10 + n12 + n17;
    const C pxy = -n01 + n03 + n10 - n12;
    const C pyx = pxy;
    C ptot = sqrt(pxx * pxx + (C)2.0 * pxy * pyx + pyy * pyy);
    if (SCALED) ptot = ptot * P.dsc;
    // Ts = 1/(2 omega1) + sqrt(18 cteS^2 P + (1/omega1)^2) / 2
    const C ts = P.les_h + sqrt(P.les_c * ptot + P.les_i2) / (C)2.0;
    return (C)1.0 / ts;
}

// BGK relaxation of one node, form 0: the shifted form operation for
// operation as SRC/col_MC.F90:125-293 (and SRC/col.F90:104-237), bit-exact
// with the reference's builds.
//   FORCE : the kernel may serve CHANNEL, whose body force P.channel_force
//           selects; without it the force terms of the reference's
//           non-CHANNEL builds are zeros.
//   Z0    : storage kind = compute kind.  No stored population is then ever
//           -0 (a difference a - b is -0 only for a = -0), so adding the
//           +0 force of a non-CHANNEL build changes nothing and is left out.
// The cte0 + q terms of the four non-rest equilibria are left out too: q is
// qf*(3a*a - vsq) with 3a*a >= +0, never -0, so 0 + q = q.  The rest
// population's q0 = cte0 + qf*(-vsq) keeps it (-vsq is -0 at rest).
// NS: the kernel may serve the NOSHIFT build (P.noshift): cte1 = 0 and
// cte0 = 1 (SRC/col_MC.F90:70-74), so the cte0 terms are added.
template <class C, bool FORCE, bool Z0, bool NS = false>
__device__ __forceinline__ void relax_ref(const Phys<C> &P, const C x[9], C o[9])
{
    const C cte1 = (NS && P.noshift) ? (C)0.0 : (C)1.0, cte0 = (C)1.0 - cte1;
    const C rf = K<C>::rf, qf = K<C>::qf, p0 = K<C>::p0, p1 = K<C>::p1, p2 = K<C>::p2;
    const C x01 = x[0], x03 = x[1], x05 = x[2], x08 = x[3], x10 = x[4];
    const C x12 = x[5], x14 = x[6], x17 = x[7], x19 = x[8];
    C rho = (((x01 + x03 + x10 + x12) + (x05 + x08 + x14 + x17)) + x19) + cte1;
    C rhoinv = (C)1.0 / rho;
    C vx = ((x03 - x12) + (x01 - x10) + (x05 - x14)) * rhoinv;
    C vy = ((x03 - x01) + (x12 - x10) + (x08 - x17)) * rhoinv;
    C vx2 = vx * vx, vy2 = vy * vy;
    C vsq = vx2 + vy2;
    C vxpy = vx + vy, vxmy = vx - vy;
    C qxpy = qf * ((C)3.0 * vxpy * vxpy - vsq);
    C qxmy = qf * ((C)3.0 * vxmy * vxmy - vsq);
    C qx = qf * ((C)3.0 * vx2 - vsq);
    C qy = qf * ((C)3.0 * vy2 - vsq);
    if (NS) {
        qxpy = cte0 + qxpy;
        qxmy = cte0 + qxmy;
        qx = cte0 + qx;
        qy = cte0 + qy;
    }
    C q0 = cte0 + qf * (-vsq);
    vx = rf * vx;
    vy = rf * vy;
    vxpy = rf * vxpy;
    vxmy = rf * vxmy;
    C rp0 = rho * p0, rp1 = rho * p1, rp2 = rho * p2;
    const C d0 = cte1 * (rp0 - p0), d1 = cte1 * (rp1 - p1), d2 = cte1 * (rp2 - p2);
    C e[9];
    e[0] = rp2 * (+vxmy + qxmy) + d2;
    e[1] = rp2 * (+vxpy + qxpy) + d2;
    e[2] = rp1 * (+vx + qx) + d1;
    e[3] = rp1 * (+vy + qy) + d1;
    e[4] = rp2 * (-vxpy + qxpy) + d2;
    e[5] = rp2 * (-vxmy + qxmy) + d2;
    e[6] = rp1 * (-vx + qx) + d1;
    e[7] = rp1 * (-vy + qy) + d1;
    e[8] = rp0 * (+q0) + d0;
    const C om = P.les ? les_rate<C, false>(P, x, e) : P.omega;
    C t[9];
#pragma unroll
    for (int k = 0; k < 9; k++) t[k] = x[k] - om * (x[k] - e[k]);
    o[8] = t[8];
    const bool fon = FORCE && P.channel_force != 0;
    if (fon || !Z0) {
        // forcing term (SRC/col_MC.F90:266-274, SRC/col.F90:210-218):
        // fgrad*rho on x, or on y in the FORCING_Y build
        C forcex = (C)0.0, forcey = (C)0.0;
        if (FORCE && P.channel_force == 1) forcex = P.fgrad * rho;
        if (FORCE && P.channel_force == 2) forcey = P.fgrad * rho;
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

// Form 1, the deviation form (an extension of the reference): the density
// deviation drho = sum of the shifted populations is never added to 1 where
// it multiplies a weight (the w*drho terms), so the O(u^2) deviation is not
// rounded against 1, and every multiply-add is a fused one:
//   u = j / rho, usq = ux^2 + uy^2, m = -1.5 usq,
//   a_w = fma(w rho, m, w drho), b_w = 4.5 w rho, c_w = 3 w rho,
//   e_k, e_k' = fma(b_w, cu^2, a_w) +- c_w cu for each opposite pair,
//   e_19 = a_0,   o_k = fma(omega, e_k - x_k, x_k),
// plus the CHANNEL body force.  SCALED: x and o are the populations as
// stored, scale*(f - w) (P.usc = scale, P.dsc = 1/scale); scaling by a
// power of two commutes with every rounding here, so each result is
// exactly scale times the unscaled one and the store needs no multiply.
template <class C, bool FORCE, bool SCALED>
__device__ __forceinline__ void relax_dev(const Phys<C> &P, const C x[9], C o[9])
{
    const C p0 = K<C>::p0, p1 = K<C>::p1, p2 = K<C>::p2;
    const C x01 = x[0], x03 = x[1], x05 = x[2], x08 = x[3], x10 = x[4];
    const C x12 = x[5], x14 = x[6], x17 = x[7], x19 = x[8];
    const C drs = ((x01 + x03 + x10 + x12) + (x05 + x08 + x14 + x17)) + x19;  // scaled when SCALED
    const C drho = SCALED ? drs * P.dsc : drs;
    const C rho = (C)1.0 + drho;
    const C rhoinv = (C)1.0 / rho;
    const C rinv = SCALED ? rhoinv * P.dsc : rhoinv;
    const C ux = ((x03 - x12) + (x01 - x10) + (x05 - x14)) * rinv;
    const C uy = ((x03 - x01) + (x12 - x10) + (x08 - x17)) * rinv;
    const C ux2 = ux * ux, uy2 = uy * uy;
    const C m = (C)(-1.5) * (ux2 + uy2);
    const C rw = SCALED ? rho * P.usc : rho;
    const C r0 = rw * p0, r1 = rw * p1, r2 = rw * p2;
    const C a0 = fma(r0, m, p0 * drs), a1 = fma(r1, m, p1 * drs), a2 = fma(r2, m, p2 * drs);
    const C b1 = (C)4.5 * r1, b2 = (C)4.5 * r2, c1 = (C)3.0 * r1, c2 = (C)3.0 * r2;
    C e[9];
    {
        const C sx = fma(b1, ux2, a1), tx = c1 * ux;
        e[2] = sx + tx;  // 05
        e[6] = sx - tx;  // 14
        const C sy = fma(b1, uy2, a1), ty = c1 * uy;
        e[3] = sy + ty;  // 08
        e[7] = sy - ty;  // 17
        const C up = ux + uy, sp = fma(b2, up * up, a2), tp = c2 * up;
        e[1] = sp + tp;  // 03
        e[4] = sp - tp;  // 10
        const C um = ux - uy, sm = fma(b2, um * um, a2), tm = c2 * um;
        e[0] = sm + tm;  // 01
        e[5] = sm - tm;  // 12
        e[8] = a0;       // 19
    }
    const C om = P.les ? les_rate<C, SCALED>(P, x, e) : P.omega;
#pragma unroll
    for (int k = 0; k < 9; k++) o[k] = fma(om, e[k] - x[k], x[k]);
    if (FORCE && P.channel_force != 0) {
        C forcex = (C)0.0, forcey = (C)0.0;
        if (P.channel_force == 1) forcex = P.fgrad * rw;
        if (P.channel_force == 2) forcey = P.fgrad * rw;
        o[0] = o[0] + (+forcex - forcey);
        o[1] = o[1] + (+forcex + forcey);
        o[2] = o[2] + (+forcex);
        o[3] = o[3] + (+forcey);
        o[4] = o[4] + (-forcex - forcey);
        o[5] = o[5] + (-forcex + forcey);
        o[6] = o[6] + (-forcex);
        o[7] = o[7] + (-forcey);
    }
}

// FL: -1 the form is P.form; 0 or 1 that form, fixed at compile time.
// NS: see relax_ref (the per-phase kernels).
template <class C, bool FORCE = true, int FL = -1, bool SCALED = false, bool Z0 = false, bool NS = false>
__device__ __forceinline__ void relax(const Phys<C> &P, const C x[9], C o[9])
{
    static_assert(!SCALED || FL == 1, "scaled populations need the deviation form");
    if (FL == 1 || (FL < 0 && P.form))
        relax_dev<C, FORCE, SCALED>(P, x, o);
    else
        relax_ref<C, FORCE, Z0, NS>(P, x, o);
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
