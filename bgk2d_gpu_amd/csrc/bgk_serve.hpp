// bgk_serve.hpp -- persistent multi-request kernel of the batched solver.
//
// Many independent solves ("members", each a bgk_ctx with its own size,
// boundary variant and storage kind) stay resident in HBM and advance
// together.  Members fall into groups by (storage kind, boundary variant);
// one launch of k_batch<S, BC> performs one FUSED step of every active
// member of its group, so each kernel carries only its own group's code and
// registers (one kernel for every group needed up to 142 VGPRs and ran the
// large members 15-25 % slower than the one-step kernels).  Blocks
// grid-stride over the group's part of a device-resident task table; a task
// names a member and a run of up to TASK_TILES consecutive tiles of it (a
// tile is 256 nodes, or 256 node pairs for the 16-bit kinds), so the
// member's description is fetched once per task and the tile loop issues no
// dependent scalar loads.  The block runs exactly the per-node code of the
// one-step kernels (fused_node; fused_pair for 16-bit storage), so a
// member's populations are bit-identical whether it runs alone or in any
// mix.  Each launch takes the index t of the step inside a chunk; a member
// whose remaining step count is <= t sits the launch out, and its FUSED
// parity is (parity0 + t) & 1 (the same for the two pop-19 planes from
// rparity0).  The host rewrites the small member/task tables between chunks
// (admission and retirement), never during one, so chunks replay from HIP
// graphs, one chain per group on its own stream.
#pragma once
#include "bgk_kernels.hpp"

namespace bgk {

// storage kinds the batch kernels serve
enum MemberKind {
    MK_F32 = 0,   // fp32 storage + compute
    MK_F16 = 1,   // IEEE half storage, fp32 compute
    MK_BF16 = 2,  // bfloat16 storage, fp32 compute
};

// boundary variants the batch kernels serve, in group order
constexpr int BATCH_BCS[6] = {BK_DRIVEN, BK_PERIODIC, BK_CHANNEL, BK_INFLOW, BK_COUETTE, BK_OPEN};
// group = (kind slot * 2 + fl) * 6 + index of the variant in BATCH_BCS:
// kind slots TWO_STEP + MK_* hold the members of that storage kind that
// advance two steps per launch in this chunk; fl is the kernel instance's
// form/rounding flag (Cfg in bgk_kernels.hpp): 0 the reference's form
// with nearest-even stores, 1 the deviation form with stochastic rounding
// of 16-bit stores (the serving settings)
constexpr int TWO_STEP = 3;
constexpr int NGROUPS = 6 * 2 * 6;
__host__ __device__ constexpr int group_of(int slot, int fl, int v) { return (slot * 2 + fl) * 6 + v; }
__host__ __device__ constexpr int group_slot(int g) { return g / 12; }
__host__ __device__ constexpr int group_fl(int g) { return (g / 6) % 2; }
__host__ __device__ constexpr int group_variant(int g) { return g % 6; }
// largest strip height of the two-step batch kernel
constexpr int TB_BATCH_H = 64;

// tiles per task: enough to hide the member lookup, few enough that a
// 128^2 member still spreads over several blocks (measured: 8 as fast as 4
// for fp32 and 3-20 % slower for fp16, 16 slower for both; 1, 2, 32, 64
// and strided tasks slower than 8)
#ifndef BGK_TASK_TILES
#define BGK_TASK_TILES 4
#endif
constexpr int TASK_TILES = BGK_TASK_TILES;

// The planes are held as byte offsets from the member table's own device
// address: a pointer read from memory is generic to the compiler and its
// accesses become flat instructions (which also wait on the LDS counter),
// while an offset from a kernel argument keeps them global loads and stores.
struct Member {
    long long set[2][8];  // the two population sets
    long long r[2];       // the two pop-19 planes
    int l, m, off;
    long long rs;
    int tiles_x;             // tiles per row of tiles (strips for the two-step kernel)
    int ntiles;              // tiles of the member
    int tw_log2, th;         // tile shape: (1 << tw_log2) x th = 256 nodes (two-step: th = strip height)
    int parity0, remaining;  // set holding the state at t = 0; steps left
    int rparity0;            // pop-19 plane holding the state at t = 0
    Phys<float> P;
};

// meta: [0, NGROUPS) task counts of the groups; [NGROUPS, 2 NGROUPS) their
// first task in the table.  A task (member, first tile) covers the tiles
// first .. min(first + TASK_TILES, ntiles) - 1.
// No occupancy bound: under a bound of 8 waves per SIMD (64 VGPRs) the
// compiler spilled 100-270 bytes per lane and a mixed stream of 128..1024
// lattices ran at 20 GLUPS instead of 34.
template <class S, int BC, int FL>
__global__ __launch_bounds__(256) void k_batch(const Member *__restrict__ M, const int2 *__restrict__ tasks,
                                               const int *__restrict__ meta, int group, int t)
{
    const int n = meta[group];
    const int2 *__restrict__ tk = tasks + meta[NGROUPS + group];
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        const int2 tl = tk[q];
        const Member &d = M[tl.x];
        if (t >= d.remaining) continue;
        const int par = (d.parity0 + t) & 1;
        const int rpar = (d.rparity0 + t) & 1;
        char *const base = (char *)M;
        Lat L;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            L.a[k] = base + d.set[par][k];
            L.b[k] = base + d.set[par ^ 1][k];
        }
        L.r = base + d.r[rpar];
        L.rn = base + d.r[rpar ^ 1];
        L.obs = nullptr;
        L.l = d.l;
        L.m = d.m;
        L.pitch = 0;
        L.off = d.off;
        L.rs = d.rs;
        L.imin = d.l;
        L.imax = 0;
        L.jmin = d.m;
        L.jmax = 0;
        L.stepx = 0;
        L.stepy = 0;
        const Phys<float> P = fixed_flags<S, float, BC, FL>(d.P);
        const int tiles_x = d.tiles_x, tw_log2 = d.tw_log2, th = d.th;
        const int tend = min(tl.y + TASK_TILES, d.ntiles);
        const int lane_x = (int)threadIdx.x & ((1 << tw_log2) - 1);
        const int lane_y = (int)threadIdx.x >> tw_log2;
        for (int tile = tl.y; tile < tend; tile++) {
            const int tx = tile % tiles_x, ty = tile / tiles_x;
            // column of this thread's node (16-bit kinds: first node of its pair)
            const int col = (tx << tw_log2) + lane_x;
            const int i = sizeof(S) == 2 ? 2 * col + 1 : col + 1;
            const int j = ty * th + lane_y + 1;
            if (i > L.l || j > L.m) continue;
            // one node (fp32) or the pair (i, i+1) (16-bit kinds), as bgk_run's kernels
            if constexpr (sizeof(S) == 2)
                fused_pair<S, float, BC, false, FL>(L, P, i, j);
            else
                fused_node<S, float, BC, false, FL>(L, P, i, j);
        }
    }
}

// Two FUSED steps of every member of a two-step group per launch: a task
// is one strip block of k_fused_tb (TB_TW columns by th rows), run by the
// same device code, so the values are those of two one-step launches.
// Launch t2 performs steps 2 t2 and 2 t2 + 1 of the chunk; a member in the
// group has at least 2 (t2 + 1) steps left.  Every task begins with a
// barrier: the previous task's last reads of the LDS ring must be done
// before this one refills it.
// The tables are read through the constant address space: they do not
// change during a launch, and so their loads stay scalar after the
// barriers (which would otherwise make them vector loads, the members'
// fields then held in VGPRs).  Field by field, so that only what is used is
// loaded and nothing goes through private memory.
template <class T>
__device__ __forceinline__ T cld(const T *p)
{
    return *(const __attribute__((address_space(4))) T *)p;
}
__device__ __forceinline__ Phys<float> cld_phys(const Phys<float> *p)
{
    static_assert(sizeof(Phys<float>) == 21 * 4, "cld_phys copies every field of Phys");
    Phys<float> r;
    r.omega = cld(&p->omega);
    r.fgrad = cld(&p->fgrad);
    r.force = cld(&p->force);
    r.sc = cld(&p->sc);
    r.inv = cld(&p->inv);
#pragma unroll
    for (int k = 0; k < 3; k++) r.inflow[k] = cld(&p->inflow[k]);
    r.umax = cld(&p->umax);
    r.channel_force = cld(&p->channel_force);
    r.form = cld(&p->form);
    r.rounding = cld(&p->rounding);
    r.les = cld(&p->les);
    r.trick1 = cld(&p->trick1);
    r.ctes = cld(&p->ctes);
    r.les_h = cld(&p->les_h);
    r.les_i2 = cld(&p->les_i2);
    r.les_c = cld(&p->les_c);
    r.usc = cld(&p->usc);
    r.dsc = cld(&p->dsc);
    r.noshift = cld(&p->noshift);
    return r;
}

template <class S, int BC, int FL>
__global__ __launch_bounds__(TB_TW) void k_batch_tb(const Member *__restrict__ M, const int2 *__restrict__ tasks,
                                                  const int *__restrict__ meta, int group, int t2)
{
    const int n = cld(meta + group);
    const int *tk = (const int *)tasks + 2 * cld(meta + NGROUPS + group);
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        __syncthreads();
        const int mi = cld(tk + 2 * q), tile = cld(tk + 2 * q + 1);
        const Member *d = M + mi;
        if (2 * t2 + 2 > cld(&d->remaining)) continue;
        const int par = (cld(&d->parity0) + t2) & 1;
        const int rpar = (cld(&d->rparity0) + t2) & 1;
        char *const base = (char *)M;
        Lat L;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            L.a[k] = base + cld(&d->set[par][k]);
            L.b[k] = base + cld(&d->set[par ^ 1][k]);
        }
        L.r = base + cld(&d->r[rpar]);
        L.rn = base + cld(&d->r[rpar ^ 1]);
        L.obs = nullptr;
        L.l = cld(&d->l);
        L.m = cld(&d->m);
        L.pitch = 0;
        L.off = cld(&d->off);
        L.rs = cld(&d->rs);
        L.imin = L.l;
        L.imax = 0;
        L.jmin = L.m;
        L.jmax = 0;
        L.stepx = 0;
        L.stepy = 0;
        const Phys<float> P = fixed_flags<S, float, BC, FL>(cld_phys(&d->P));
        const int tiles_x = cld(&d->tiles_x);
        const int th = cld(&d->th), j0 = (tile / tiles_x) * th + 1;
        tb_block<S, float, BC, false, FL>(L, P, tile % tiles_x, j0, min(j0 + th - 1, L.m));
    }
}

}  // namespace bgk
