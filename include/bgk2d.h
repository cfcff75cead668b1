/*
 * bgk2d.h -- C ABI of libbgk2d_hip.so, the MI355X (gfx950) implementation
 * of BGK2D's time-step loop: D2Q9 BGK collision, ORIGINAL (two-pass) and
 * FUSED (single-pass) streaming, every boundary variant, and the FP64
 * diagnostics.
 *
 * Plain C types only.  Every entry point returns 0 on success or a
 * negative BGK_E* code; bgk_last_error() gives the message.  One context is
 * driven by one host thread; distinct contexts are independent and may be
 * used from different threads.  All work of a context runs in order on its
 * own HIP stream.  Coordinates are the reference's: 1-based interior
 * (1..l, 1..m), halo ring at 0 and l+1 / m+1, x fastest.
 *
 * Reference interfaces replaced (the reference has no plugin registry; its
 * "operators" are the parameterless subroutines below, all state living in
 * module `storage`, SRC/storage.F90:25-140):
 *   bgk_create          alloca + build_bcond + hencol  (SRC/alloca.F90:110-151,
 *                       SRC/build.bcond.F90:36-45, SRC/hencol.F90:30-50)
 *   bgk_build_obs       build_obs          (SRC/build.obs.F90:170-205)
 *   bgk_set_obstacle    obs(:,:) from a file/caller (SRC/r.obs.F90)
 *   bgk_init            init(opt)          (SRC/init.F90:83-137)
 *   bgk_boundaries      boundaries         (SRC/boundaries.F90:23-64)
 *   bgk_propagation     propagation/movef  (SRC/propagation.F90:31-45)
 *   bgk_collision       collision(itime)   (SRC/collision.F90:37-41)
 *   bgk_run             the loop body of SRC/bgk2d.F90:92-113 for n steps
 *   bgk_diagno          diagno             (SRC/diagno.F90:63-101)
 *   bgk_varm            varm               (SRC/varm.F90:48-79)
 *   bgk_prof_i/_j       prof_i / prof_j    (SRC/prof_i.F90:47-63, SRC/prof_j.F90:47-67)
 *   bgk_probe           probe              (SRC/probe.F90:46-62)
 *   bgk_download/upload `!$omp target update` of a01..a19 and the records
 *                       of save_raw/restore_raw (SRC/save_raw.F90:37-57,
 *                       SRC/restore_raw.F90:41-51)
 */
#ifndef BGK2D_H
#define BGK2D_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define BGK_ABI_VERSION 4

/* boundary variant: the compile-time -DLDC/-DPERIODIC/-DCHANNEL/-DINFLOW/
 * -DCOUETTE/-DSTEP choice of SRC/boundaries.F90:31-56, made at run time  */
enum bgk_bc {
    BGK_BC_DRIVEN = 0,   /* bcond_driven   SRC/bcond.driven.F90:96-144   */
    BGK_BC_PERIODIC = 1, /* bcond_periodic SRC/bcond.periodic.F90:60-107 */
    BGK_BC_CHANNEL = 2,  /* bcond_channel  SRC/bcond.channel.F90:72-123  */
    BGK_BC_INFLOW = 3,   /* bcond_inflow   SRC/bcond.inflow.F90:59-153   */
    BGK_BC_COUETTE = 4,  /* bcond_couette  SRC/bcond.couette.F90:41-103  */
    BGK_BC_STEP = 5,     /* bcond_step     SRC/bcond.step.F90:55-254     */
    BGK_BC_OPEN = 6      /* no reference counterpart: channel with a
                            parabolic inflow of peak u0 at x = 0 (shifted
                            equilibrium at the density of the first fluid
                            node), zero-gradient outflow at x = l+1 and
                            free-slip (specular) walls at y = 0, m+1     */
};

/* precision: SRC/precision.F90:13-47 plus bf16 storage */
enum bgk_prec {
    BGK_PREC_SINGLE = 0, /* fp32 storage, fp32 compute                    */
    BGK_PREC_DOUBLE = 1, /* fp64 storage, fp64 compute                    */
    BGK_PREC_FP16 = 2,   /* IEEE half storage of scale*(f-w), fp32 compute
                            (MIXED1 when scale = 1)                        */
    BGK_PREC_MIXED2 = 3, /* fp32 storage, fp64 compute                    */
    BGK_PREC_BF16 = 4    /* bfloat16 storage of scale*(f-w), fp32 compute */
};

enum bgk_impl { BGK_IMPL_ORIGINAL = 0, BGK_IMPL_FUSED = 1 };

enum bgk_err {
    BGK_OK = 0,
    BGK_E_INVALID = -1,   /* bad argument / unsupported combination */
    BGK_E_NOMEM = -2,     /* device or host allocation failed        */
    BGK_E_HIP = -3,       /* HIP runtime error                       */
    BGK_E_DIVERGED = -4,  /* NaN/Inf or |u| > 0.5 found by a check   */
    BGK_E_STATE = -5      /* call out of order (e.g. run before init) */
};

typedef struct bgk_params {
    int32_t lx, ly;       /* interior size (namelist lx, ly)                */
    int32_t ipad, jpad;   /* namelist padding; accepted, layout is pitched  */
    int32_t bc;           /* enum bgk_bc                                    */
    int32_t obstacle;     /* bcond_obs with the build_obs mask: 1 cylinder
                             (VKS build), 2 square (SQUARE), 3 half
                             cylinder (HALF_CIRCLE), 4 porous (POROUS:
                             five-node plus shapes on a 5-spaced lattice
                             where a Park-Miller minimal standard draw,
                             seed 1, exceeds 0.4); a bgk_set_obstacle mask
                             turns a 0 into 1                               */
    int32_t impl;         /* enum bgk_impl                                  */
    int32_t prec;         /* enum bgk_prec                                  */
    int32_t radius;       /* cylinder radius (namelist radius)              */
    int32_t device;       /* HIP device ordinal                             */
    int32_t mixed1_quirk; /* 1: zero u0 in hencol for FP16 as the MIXED1
                             reference build does (SRC/hencol.F90:33-46)    */
    int32_t form;         /* 0: the reference's operation order (bit-exact
                             with its builds); 1: deviation form, the
                             density deviation kept apart from 1 and fused
                             multiply-adds (an extension)                  */
    int32_t rounding;     /* FP16/BF16 stores: 0 round to nearest even (the
                             reference's MIXED1); 1 deterministic stochastic
                             rounding (random word from the value bits and
                             the position)                                 */
    int32_t tblock;       /* FUSED steps per pass over HBM in bgk_run: 0 the
                             library's choice (2 where the lattice fills the
                             device with strips of >= 8 rows, else 1), 1 one
                             step per kernel, 2 two steps per kernel;
                             results are identical                         */
    int32_t les;          /* 1: Smagorinsky LES collision, the -DLES build
                             (SRC/col_MC.F90:178-230)                       */
    int32_t trick1;       /* 1: driven-cavity corner order of the TRICK1 build
                             (SRC/bcond.driven.F90:43-81; lx == ly)         */
    int32_t forcing_y;    /* 1: CHANNEL body force along y, the FORCING_Y
                             build (SRC/col_MC.F90:266-274)                 */
    int32_t noshift;      /* 1: whole populations f stored instead of the
                             deviations f - w, the NOSHIFT build
                             (SRC/col_MC.F90:70-74, SRC/init.F90:73-77);
                             form 0 only, not OPEN; bgk_run folds the
                             boundary into the one- and two-step kernels
                             as for shifted populations (SRC/bcond.inflow
                             .F90:48-52 for the inflow equilibria)          */
    double svisc;         /* namelist svisc (kinematic viscosity)           */
    double u0;            /* namelist u0 (lid speed, or channel force scale)*/
    double scale;         /* FP16/BF16 storage scale (power of two)         */
    double cte_s;         /* LES: namelist cteS (SRC/input.F90:55 defaults it
                             to 0.1, a default-real literal)                */
} bgk_params_t;

typedef struct bgk_ctx bgk_ctx_t;

/* Context: allocates the two population sets in HBM (18 pitched planes:
 * eight moving populations per set and two planes of the rest population),
 * fills them with huge(mykind) as alloca does, and derives omega, fgrad
 * and the lid speed in the compute kind exactly as hencol/build_bcond.
 * Size limit: one plane, (ly+2) rows of pitch = roundup(lx+3+A, A)
 * elements (A = 128 bytes / element size), must stay below 4 GiB, because
 * the hot kernels form per-node byte offsets in 32 bits: about 32,700^2
 * nodes in FP32, 23,100^2 in FP64, 46,300^2 in FP16/BF16.  Larger
 * requests fail with BGK_E_INVALID before any device call.              */
int bgk_create(const bgk_params_t *p, bgk_ctx_t **out);
int bgk_destroy(bgk_ctx_t *ctx);
/* message of the context's last failure; with ctx == NULL, of the last
 * failed bgk_create or bgk_footprint on the calling thread               */
const char *bgk_last_error(const bgk_ctx_t *ctx);
int bgk_version(void);

/* omega, fgrad, u00, force (= u00/6) in the compute kind, widened */
int bgk_constants(const bgk_ctx_t *ctx, double out[4]);

/* Mask of SRC/build.obs.F90:49-205 for params.obstacle (cylinder, square,
 * half cylinder: centre 2l/5, m/2; porous: l/4..3l/4 x 5..m-4).  The mask
 * must stay at least two nodes from every wall (else BGK_E_INVALID).    */
int bgk_build_obs(bgk_ctx_t *ctx);
/* Caller-supplied mask obs(1:l,1:m) as l*m int32, x fastest (1 = solid). */
int bgk_set_obstacle(bgk_ctx_t *ctx, const int32_t *obs);
int bgk_get_obstacle(bgk_ctx_t *ctx, int32_t *obs, int32_t bbox[4]);

/* init(opt): shifted rest equilibrium, or for PERIODIC the Taylor-Green
 * field u = A sin(2 pi x) cos(2 pi y), v = -A cos(2 pi x) sin(2 pi y).
 * The reference hard-codes A = 0.1 (SRC/init.F90:109-110) and only
 * range-checks opt (0..5); here opt = 1 with PERIODIC also sets the
 * density of the vortex's pressure field, rho = 1 + (3A^2/4)(cos 4 pi x +
 * cos 4 pi y), which launches no acoustic transient.                    */
int bgk_init(bgk_ctx_t *ctx, int32_t init_v, double amplitude);

/* bgk_init that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory at the point of the
 * call (code under HIP graph capture, serving loops that recycle a
 * context for the next request, harnesses that forbid hipMalloc).
 * bgk_init builds its trig tables (2 (lx + ly + 4) elements of the compute
 * kind) in a hipMalloc'd buffer; this entry point uses a static device
 * array of BGK_INIT_SCRATCH_BYTES per GPU instead (held until the process
 * ends) and gives the same populations bit for bit.  Lattices whose
 * tables exceed it (lx + ly + 4 > 131072 for fp32 compute, > 65536 for
 * fp64) are refused with BGK_E_INVALID.  Calls on every GPU serialize on
 * one host lock, held until the kernel that reads the array has ended, so
 * calls that overlap wait rather than race; returns after the kernel. */
#define BGK_INIT_SCRATCH_BYTES (1 << 20)
int bgk_init_noalloc(bgk_ctx_t *ctx, int32_t init_v, double amplitude);

/* The three per-step phases of SRC/bgk2d.F90:100-102, each a kernel
 * launch on the context stream with the reference's halo semantics.     */
int bgk_boundaries(bgk_ctx_t *ctx);
int bgk_propagation(bgk_ctx_t *ctx);
int bgk_collision(bgk_ctx_t *ctx, int32_t itime);

/* nsteps complete steps.  FUSED with a foldable boundary (all but STEP)
 * runs the boundary folded into the pull, one or two steps per kernel
 * (params.tblock); the rest replay boundaries/propagation/collision.
 * Steps are captured into HIP graphs in chunks.  Interior results are
 * identical to the per-phase entry points; the folded path neither reads
 * nor maintains the halo ring (the two-step 16-bit kernels may leave
 * arbitrary values in the bottom halo row), which bgk_boundaries rewrites
 * before any per-phase step reads it.                                    */
int bgk_run(bgk_ctx_t *ctx, int32_t nsteps);
int bgk_sync(bgk_ctx_t *ctx);

/* FP64 diagnostics, deterministic fixed-order reductions on the device.
 * diagno: out = { <u>, <v>, <rho>, <u^2+v^2> } over 1..l x 1..m.        */
int bgk_diagno(bgk_ctx_t *ctx, double out[4]);
int bgk_varm(bgk_ctx_t *ctx, double *u, double *v, double *rho);           /* m each */
int bgk_prof_i(bgk_ctx_t *ctx, int32_t jcoord, double *u, double *v, double *rho); /* l each */
int bgk_prof_j(bgk_ctx_t *ctx, int32_t icoord, double *u, double *v, double *rho); /* m each */
int bgk_probe(bgk_ctx_t *ctx, int32_t i0, int32_t j0, double out[3]);
/* draglift (SRC/draglift.F90:58-98): momentum-exchange force on the
 * obstacle, out = { Fx, Fy } = sum over the fluid nodes of the box
 * [i0,i1] x [j0,j1] (clipped to the interior) of 2 c_k a_k for each
 * population whose neighbour is solid, in FP64.  The reference's box is
 * the cylinder centre +- (radius + 3); Cd = Fx / (U^2 radius).            */
int bgk_draglift(bgk_ctx_t *ctx, int32_t i0, int32_t i1, int32_t j0, int32_t j1, double out[2]);
/* Fields of the VTK writers vtk_xy_bin / vtk_om_bin (SRC/vtk_xy.binary.F90:
 * 60-140, SRC/vtk_om.binary.F90): u/rho, v/rho and rho per interior node in
 * FP32 (population sums in the storage kind), l*m each, x fastest.        */
int bgk_fields(bgk_ctx_t *ctx, float *u, float *v, float *rho);
/* 0 if every node is finite with |u| <= umax, else BGK_E_DIVERGED */
int bgk_check(bgk_ctx_t *ctx, double umax);

/* Population k (0..8 = 01,03,05,08,10,12,14,17,19) of the current set,
 * (l+2)*(m+2) elements of the storage kind (fp32, fp64, or the raw 16-bit
 * pattern for FP16/BF16), x fastest, halos included.                    */
int bgk_download(bgk_ctx_t *ctx, int32_t k, void *host);
int bgk_upload(bgk_ctx_t *ctx, int32_t k, const void *host);
int bgk_storage_bytes(const bgk_ctx_t *ctx);

/* Stream and timing helpers for benchmarks: the context stream as an
 * opaque handle, and the average time per step (ms) of the last bgk_run,
 * measured with HIP events on that stream around all of its launches.    */
void *bgk_stream(bgk_ctx_t *ctx);
int bgk_last_run_ms(const bgk_ctx_t *ctx, double *ms);

/* ------------------------------------------------------------------------
 * Batched solving (no counterpart in the reference, which runs one solve
 * per process): many independent FUSED solves of any size, boundary
 * (driven, periodic, channel, inflow, couette, open; no obstacle) and
 * storage kind (fp32, fp16, bf16) stay resident in HBM and advance
 * together: per step one persistent-kernel launch per group of members of
 * one storage kind and boundary variant (two steps per launch where
 * bgk_run would use the two-step kernel), chunks of steps replayed from
 * HIP graphs.  A member's results are bit-identical to the same solve run
 * alone with bgk_run.  Members are admitted while their device footprint
 * fits the byte budget given at creation.                                */
typedef struct bgk_batch bgk_batch_t;

int bgk_batch_create(int32_t device, double budget_bytes, int32_t max_members, bgk_batch_t **out);
int bgk_batch_destroy(bgk_batch_t *b);
const char *bgk_batch_last_error(const bgk_batch_t *b);
/* device bytes a context with these parameters allocates; BGK_E_INVALID
 * for parameters bgk_create would refuse (the same checks)               */
int bgk_footprint(const bgk_params_t *p, double *bytes);
/* create + init a member that will run nsteps steps; tile: 0 default,
 * 1 = 64x4, 2 = 128x2, 3 = 256x1, 4 = 32x8 nodes per block.  Returns
 * BGK_E_NOMEM (nothing allocated) when the budget or the slots are full. */
int bgk_batch_add(bgk_batch_t *b, const bgk_params_t *p, int32_t init_v, double amplitude, int64_t nsteps,
                  int32_t tile, int32_t *slot);
/* the member's context, for bgk_diagno/bgk_check/bgk_download...        */
int bgk_batch_ctx(bgk_batch_t *b, int32_t slot, bgk_ctx_t **ctx);
int bgk_batch_remaining(bgk_batch_t *b, int32_t slot, int64_t *remaining);
int bgk_batch_remove(bgk_batch_t *b, int32_t slot);
/* held = 1: bgk_batch_run skips the member (its state and steps left stay
 * as they are) until it is released with held = 0; for a request loop's
 * scheduling policy                                                      */
int bgk_batch_hold(bgk_batch_t *b, int32_t slot, int32_t held);
int bgk_batch_used(const bgk_batch_t *b, double *bytes);
/* advance every member with steps left by min(chunk, remaining) steps   */
int bgk_batch_run(bgk_batch_t *b, int32_t chunk, int32_t *finished);
int bgk_batch_last_ms(const bgk_batch_t *b, double *ms);

#ifdef __cplusplus
}
#endif
#endif
