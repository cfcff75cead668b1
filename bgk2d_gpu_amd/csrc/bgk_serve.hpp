// bgk_serve.hpp -- persistent multi-request kernel of the batched solver.
//
// Many independent solves ("members", each a bgk_ctx with its own size,
// boundary variant and storage kind) stay resident in HBM and advance
// together: one launch of k_batch performs one FUSED step of every active
// member.  Blocks grid-stride over a device-resident tile table; a tile
// names a member and a 256-node tile of it, and the block runs exactly the
// per-node code of k_fused (fused_node), so a member's populations are
// bit-identical whether it runs alone or in any mix.  Each launch takes the
// index t of the step inside a chunk; a member whose remaining step count is
// <= t sits the launch out, and its FUSED parity is (parity0 + t) & 1 (the
// same for the two pop-19 planes from rparity0).  The
// host rewrites the small member/tile tables between chunks (admission and
// retirement), never during one, so chunks replay from HIP graphs.
#pragma once
#include "bgk_kernels.hpp"

namespace bgk {

// (storage kind, boundary) pairs the batch kernel serves
enum MemberKind {
    MK_F32 = 0,   // fp32 storage + compute
    MK_F16 = 1,   // IEEE half storage, fp32 compute
    MK_BF16 = 2,  // bfloat16 storage, fp32 compute
};

struct Member {
    void *set[2][8];  // the two population sets
    void *r[2];       // the two pop-19 planes
    int l, m, off, store;
    long long rs;
    int bc;                  // BcKind
    int tiles_x;             // tiles per row of tiles
    int tw_log2, th;         // tile shape: (1 << tw_log2) x th = 256 nodes
    int parity0, remaining;  // set holding the state at t = 0; steps left
    int rparity0;            // pop-19 plane holding the state at t = 0
    Phys<float> P;
};

template <class S, int BC>
__device__ __forceinline__ void member_node(const Lat &L, const Phys<float> &P, int i, int j)
{
    fused_node<S, float, BC, false>(L, P, i, j);
}

template <class S>
__device__ __forceinline__ void member_bc(const Lat &L, const Phys<float> &P, int bc, int i, int j)
{
    switch (bc) {
    case BK_PERIODIC: member_node<S, BK_PERIODIC>(L, P, i, j); break;
    case BK_CHANNEL: member_node<S, BK_CHANNEL>(L, P, i, j); break;
    case BK_INFLOW: member_node<S, BK_INFLOW>(L, P, i, j); break;
    case BK_COUETTE: member_node<S, BK_COUETTE>(L, P, i, j); break;
    case BK_OPEN: member_node<S, BK_OPEN>(L, P, i, j); break;
    default: member_node<S, BK_DRIVEN>(L, P, i, j); break;
    }
}

__global__ __launch_bounds__(256) void k_batch(const Member *__restrict__ M, const int2 *__restrict__ tiles,
                                               const int *__restrict__ ntiles, int t)
{
    const int n = *ntiles;
    for (int q = blockIdx.x; q < n; q += gridDim.x) {
        const int2 tl = tiles[q];
        const Member &d = M[tl.x];
        if (t >= d.remaining) continue;
        const int par = (d.parity0 + t) & 1;
        const int tx = tl.y % d.tiles_x, ty = tl.y / d.tiles_x;
        const int tw = 1 << d.tw_log2;
        const int i = (tx << d.tw_log2) + ((int)threadIdx.x & (tw - 1)) + 1;
        const int j = ty * d.th + ((int)threadIdx.x >> d.tw_log2) + 1;
        if (i > d.l || j > d.m) continue;
        Lat L;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            L.a[k] = d.set[par][k];
            L.b[k] = d.set[par ^ 1][k];
        }
        const int rpar = (d.rparity0 + t) & 1;
        L.r = d.r[rpar];
        L.rn = d.r[rpar ^ 1];
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
        switch (d.store) {
        case MK_F16: member_bc<_Float16>(L, d.P, d.bc, i, j); break;
        case MK_BF16: member_bc<bf16_t>(L, d.P, d.bc, i, j); break;
        default: member_bc<float>(L, d.P, d.bc, i, j); break;
        }
    }
}

}  // namespace bgk
