/*
 * bgk_oracle.h -- CPU restatement of BGK2D's time-step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (libbgk2d_hip.so) never links
 * or calls it.
 *
 * It restates, scalar and in the reference's own operation order, the
 * Fortran routines named beside each entry point, for every precision mode
 * (SINGLE, DOUBLE, MIXED1 = half storage/fp32 compute, MIXED2 = fp32
 * storage/fp64 compute, plus BF16 storage/fp32 compute which the reference
 * does not have), both implementations (ORIGINAL two-pass, FUSED), and
 * every boundary variant.
 *
 * Population order used throughout (index k = 0..8):
 *     k    :  0    1    2    3    4    5    6    7    8
 *     name :  01   03   05   08   10   12   14   17   19
 *     c    : (1,-1)(1,1)(1,0)(0,1)(-1,-1)(-1,1)(-1,0)(0,-1)(0,0)
 * (SRC/alloca.F90:30-108, SRC/bgk2d.F90:29-38).
 */
#ifndef BGK_ORACLE_H
#define BGK_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_DRIVEN = 0, ORC_BC_PERIODIC = 1, ORC_BC_CHANNEL = 2,
       ORC_BC_INFLOW = 3, ORC_BC_COUETTE = 4, ORC_BC_STEP = 5,
       ORC_BC_OPEN = 6 /* parabolic inflow of peak u0, zero-gradient outflow,
                          free-slip walls: no reference counterpart      */ };
enum { ORC_SINGLE = 0, ORC_DOUBLE = 1, ORC_MIXED1 = 2, ORC_MIXED2 = 3,
       ORC_BF16 = 4 };
enum { ORC_ORIGINAL = 0, ORC_FUSED = 1 };

typedef struct {
    int32_t lx, ly;        /* interior size l, m                          */
    int32_t bc;            /* ORC_BC_*                                    */
    int32_t obstacle;      /* build_obs shape + bcond_obs: 1 cylinder (VKS),
                              2 square (SQUARE), 3 half cylinder
                              (HALF_CIRCLE)                               */
    int32_t impl;          /* ORC_ORIGINAL / ORC_FUSED                    */
    int32_t prec;          /* ORC_SINGLE ...                              */
    int32_t radius;        /* cylinder radius (namelist `radius`)         */
    int32_t init_v;        /* namelist init_v, range-checked only         */
    int32_t mixed1_quirk;  /* 1: hencol zeroes u0 for MIXED1 (reference)  */
    int32_t tgv;           /* 1: Taylor-Green initial field (PERIODIC)    */
    int32_t form;          /* 0: reference operation order; 1: deviation  */
    int32_t rounding;      /* half/bf16 stores: 0 nearest, 1 stochastic   */
    int32_t les;           /* 1: Smagorinsky LES collision (-DLES build)  */
    int32_t trick1;        /* 1: driven corner order of the TRICK1 build  */
    int32_t forcing_y;     /* 1: CHANNEL force along y (FORCING_Y build)  */
    int32_t noshift;       /* 1: whole populations f, the NOSHIFT build   */
    double svisc, u0;      /* namelist values                             */
    double amplitude;      /* TGV amplitude (reference hard-codes 0.1)    */
    double scale;          /* storage scale for half/bf16 (power of two)  */
    double cte_s;          /* Smagorinsky constant cteS, taken into the
                              compute kind (reference default 0.1, a
                              default-real literal)                       */
} orc_params;

typedef struct orc orc;

orc   *orc_create(const orc_params *p);
void   orc_destroy(orc *h);
int    orc_storage_bytes(const orc *h);
/* alloca fill + build_bcond + build_obs + init + hencol */
int    orc_setup(orc *h);
void   orc_boundaries(orc *h);
void   orc_propagation(orc *h);
void   orc_collision(orc *h);
void   orc_step(orc *h, int32_t nsteps);
/* constants in compute kind, widened: omega, fgrad, u00, force(u00/6) */
void   orc_constants(const orc *h, double out[4]);
/* decoded population k of the current `a` set, (l+2)*(m+2), x fastest   */
void   orc_get_pop(const orc *h, int32_t k, double *out);
/* raw storage bytes of population k of the current `a` set             */
void   orc_get_raw(const orc *h, int32_t k, void *out);
void   orc_set_raw(orc *h, int32_t k, const void *in);
/* raw storage bytes of the `b` set (FUSED: the next-step target)        */
void   orc_get_raw_b(const orc *h, int32_t k, void *out);
void   orc_get_obs(const orc *h, int32_t *out, int32_t bbox[4]);
/* caller mask obs(1:l,1:m), x fastest (1 = solid); turns bcond_obs on   */
void   orc_set_obs(orc *h, const int32_t *mask);
/* draglift (SRC/draglift.F90:58-98) over the box, out = {Fx, Fy}: per-
 * link terms 2 c_k a_k in compute kind, summed in long double           */
void   orc_draglift(const orc *h, int32_t i0, int32_t i1, int32_t j0, int32_t j1, double out[2]);
/* diagno: <u>,<v>,<rho>,<u2+v2>; per-node values in compute kind,
 * sums in long double (mode 0) or serially in compute kind exactly as
 * SRC/diagno.F90:63-101 (mode 1)                                        */
void   orc_diagno(const orc *h, int32_t mode, double out[4]);
void   orc_varm(const orc *h, double *u, double *v, double *rho);
void   orc_prof_i(const orc *h, int32_t jcoord, double *u, double *v, double *rho);
void   orc_prof_j(const orc *h, int32_t icoord, double *u, double *v, double *rho);
void   orc_probe(const orc *h, int32_t i0, int32_t j0, double out[3]);
/* storage conversions (for tests of the rounding helpers)               */
uint16_t orc_f32_to_f16(float f);
float    orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);
float    orc_bf16_to_f32(uint16_t h);
/* stochastic-rounding conversions of binary32 bits b with random word r  */
uint16_t orc_sr_f16(uint32_t b, uint32_t r);
uint16_t orc_sr_bf16(uint32_t b, uint32_t r);

#ifdef __cplusplus
}
#endif
#endif
