// rd_kernels.h -- building blocks of the BD step kernels on gfx950.
//
// Data layout in HBM (all replicas of a handle; a plain system is one):
//   ps[N]   float4, sorted by cell: x, y, z, w = (local index << 6) | type
//   pn[N]   float4, positions after the update, in the slot order of ps
//   kr[N]   uint2: next global cell of slot s and its rank inside that cell
//   count[ncell], start[ncell+1]: counting-sort histogram and offsets over the
//   concatenated cell ranges of all replicas; tsum[]: per-2048-cell tile sums
// One step = force/update (gather forces, Euler-Maruyama, wrap, next cell key,
// histogram) -> scan -> scatter.  Replaces orderOne, warpOrderTwo and update
// (CudaSimulation.cu:2309-2476) and the linked-list cell maintenance with
// spin locks (CudaSimulation.cu:1860-1914).
#pragma once
#include "rd_device.h"

#define RD_BLOCK 256
#define RD_SCAN_ITEMS 8
#define RD_SCAN_TILE (RD_BLOCK * RD_SCAN_ITEMS)
#define RD_SCAN_TILE_LOG2 11
static_assert(RD_SCAN_TILE == (1 << RD_SCAN_TILE_LOG2), "scan tile is a power of two");

struct RdStepState {
    unsigned long long step;        // global step counter (noise counter)
    unsigned long long noise_step0; // first step held by the noise array
    int noise_nsteps;               // 0: generator
    unsigned int err;               // bit 0: non-finite position, bit 1: noise array exhausted,
                                    // bit 2: position still outside the box after the periodic
                                    // correction (moved more than one box length in a step),
                                    // bit 3: a pair or bond gradient term beyond the fixed-point range
};

#define RD_ERR_NONFINITE 1u
#define RD_ERR_NOISE 2u
#define RD_ERR_OUTSIDE 4u
#define RD_ERR_FIXRANGE 8u

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of slots so that neighbour cells share one L2.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// number of set bits of a wave ballot below this lane
__device__ __forceinline__ int lanes_below(uint64_t b)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

// Histogram entry of cell k and of its scan tile (the per-tile sums the scan
// of the cell counts starts from).  Lanes of a wave are in slot order, so
// their cells fall in one or two tiles: one atomic per distinct tile, issued
// by the lowest lane of the tile's group.
__device__ __forceinline__ uint32_t cell_histogram(uint32_t* __restrict__ count, uint32_t* __restrict__ tsum,
                                                   uint32_t k)
{
    const uint32_t t = k >> RD_SCAN_TILE_LOG2;
    uint64_t todo = __ballot(1);
    while (todo) {
        const int lead = __builtin_ctzll(todo);
        const uint32_t t0 = (uint32_t)__shfl((int)t, lead, 64);
        const uint64_t m = __ballot(t == t0) & todo;
        if ((int)(threadIdx.x & 63) == lead) atomicAdd(&tsum[t0], (uint32_t)__popcll(m));
        todo &= ~m;
    }
    return atomicAdd(&count[k], 1u);
}

// orders a wave's LDS writes before its following LDS reads by other lanes
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x-range of a stencil row as at most two runs of contiguous cells
__device__ __forceinline__ int x_runs(int cx, int S, int L, int* a0, int* b0, int* a1, int* b1)
{
    if (2 * S + 1 >= L) { *a0 = 0; *b0 = L - 1; return 1; }
    int lo = cx - S, hi = cx + S;
    if (lo >= 0 && hi <= L - 1) { *a0 = lo; *b0 = hi; return 1; }
    if (lo < 0) { *a0 = 0; *b0 = hi; *a1 = L + lo; *b1 = L - 1; return 2; }
    *a0 = lo; *b0 = L - 1; *a1 = 0; *b1 = hi - L; return 2;
}

__device__ __forceinline__ void cell_coords(const RdRep& R, float x, float y, float z, int& cx, int& cy, int& cz)
{
    cx = cell_coord(x, R.lo[0], R.cell, R.L[0]);
    cy = cell_coord(y, R.lo[1], R.cell, R.L[1]);
    cz = cell_coord(z, R.lo[2], R.cell, R.L[2]);
}

// global cell of a position: the replica's first cell + the reference field
// index (CudaSimulation.cu:586-588)
__device__ __forceinline__ uint32_t cell_key(const RdRep& R, float x, float y, float z)
{
    int cx, cy, cz;
    cell_coords(R, x, y, z, cx, cy, cz);
    return (uint32_t)(R.coff + (cz * R.L[1] + cy) * R.L[0] + cx);
}

// Visit every slot j of the cells within S of (cx,cy,cz).  Rows along x are
// contiguous in the sorted arrays, so each row is one or two slot ranges.
template <class F>
__device__ __forceinline__ void for_stencil(const RdRep& R, const uint32_t* __restrict__ start, int cx, int cy,
                                            int cz, int S, F&& f)
{
    const int Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    int ylo, yhi, zlo, zhi;
    bool fy = 2 * S + 1 >= Ly, fz = 2 * S + 1 >= Lz;
    if (fy) { ylo = 0; yhi = Ly - 1; } else { ylo = -S; yhi = S; }
    if (fz) { zlo = 0; zhi = Lz - 1; } else { zlo = -S; zhi = S; }
    int a0, b0, a1, b1;
    int nr = x_runs(cx, S, Lx, &a0, &b0, &a1, &b1);
    for (int oz = zlo; oz <= zhi; ++oz) {
        int zz = fz ? oz : cz + oz;
        if (zz < 0) zz += Lz; else if (zz >= Lz) zz -= Lz;
        for (int oy = ylo; oy <= yhi; ++oy) {
            int yy = fy ? oy : cy + oy;
            if (yy < 0) yy += Ly; else if (yy >= Ly) yy -= Ly;
            int row = R.coff + (zz * Ly + yy) * Lx;
            uint32_t j0 = start[row + a0], j1 = start[row + b0 + 1];
            for (uint32_t j = j0; j < j1; ++j) f(j);
            if (nr == 2) {
                j0 = start[row + a1]; j1 = start[row + b1 + 1];
                for (uint32_t j = j0; j < j1; ++j) f(j);
            }
        }
    }
}

// for_stencil with each slot's record loaded one slot ahead of its use, so
// that a load is in flight while the previous candidate is examined;
// f(j, q) with q = ps[j]
template <class F>
__device__ __forceinline__ void for_stencil_q(const RdRep& R, const uint32_t* __restrict__ start,
                                              const float4* __restrict__ ps, int cx, int cy, int cz, int S, F&& f)
{
    auto run = [&](uint32_t j, uint32_t j1) {
        if (j >= j1) return;
        float4 q = ps[j];
        for (; j + 1 < j1; ++j) {
            const float4 n = ps[j + 1];
            f(j, q);
            q = n;
        }
        f(j, q);
    };
    const int Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    int ylo, yhi, zlo, zhi;
    bool fy = 2 * S + 1 >= Ly, fz = 2 * S + 1 >= Lz;
    if (fy) { ylo = 0; yhi = Ly - 1; } else { ylo = -S; yhi = S; }
    if (fz) { zlo = 0; zhi = Lz - 1; } else { zlo = -S; zhi = S; }
    int a0, b0, a1, b1;
    int nr = x_runs(cx, S, Lx, &a0, &b0, &a1, &b1);
    for (int oz = zlo; oz <= zhi; ++oz) {
        int zz = fz ? oz : cz + oz;
        if (zz < 0) zz += Lz; else if (zz >= Lz) zz -= Lz;
        for (int oy = ylo; oy <= yhi; ++oy) {
            int yy = fy ? oy : cy + oy;
            if (yy < 0) yy += Ly; else if (yy >= Ly) yy -= Ly;
            int row = R.coff + (zz * Ly + yy) * Lx;
            run(start[row + a0], start[row + b0 + 1]);
            if (nr == 2) run(start[row + a1], start[row + b1 + 1]);
        }
    }
}

// for_stencil_q over the half of the stencil that holds each unordered pair
// once: the rows (dy, dz) with dz > 0, or dz == 0 and dy > 0, in full, and of
// the home row the slots after s.  The mirror image of every other row is
// left out, so a pair is met from exactly one of its two particles.  Needs
// 2S+1 < L in every dimension (no stencil row is its own mirror image).
template <class F>
__device__ __forceinline__ void for_stencil_half_q(const RdRep& R, const uint32_t* __restrict__ start,
                                                   const float4* __restrict__ ps, uint32_t s, int cx, int cy, int cz,
                                                   int S, F&& f)
{
    auto run = [&](uint32_t j, uint32_t j1) {
        if (j >= j1) return;
        float4 q = ps[j];
        for (; j + 1 < j1; ++j) {
            const float4 n = ps[j + 1];
            f(j, q);
            q = n;
        }
        f(j, q);
    };
    const int Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    int a0, b0, a1, b1;
    const int nr = x_runs(cx, S, Lx, &a0, &b0, &a1, &b1);
    for (int oz = 0; oz <= S; ++oz) {
        int zz = cz + oz;
        if (zz >= Lz) zz -= Lz;
        for (int oy = oz == 0 ? 0 : -S; oy <= S; ++oy) {
            int yy = cy + oy;
            if (yy < 0) yy += Ly; else if (yy >= Ly) yy -= Ly;
            const int row = R.coff + (zz * Ly + yy) * Lx;
            const uint32_t lo = (oz == 0 && oy == 0) ? s + 1u : 0u;
            run(max(start[row + a0], lo), start[row + b0 + 1]);
            if (nr == 2) run(max(start[row + a1], lo), start[row + b1 + 1]);
        }
    }
}

// Pair gradients are summed in 64-bit fixed point (2^-32 resolution): integer
// addition is exact, so the sum does not depend on the order of particles
// inside a cell, which the atomic counting sort leaves unspecified.  The step
// is therefore bitwise reproducible.
#define RD_FIX_SCALE 4294967296.0f
#define RD_FIX_INV 2.3283064365386963e-10
#define RD_FIX_MAX 5.0e8f        // |term| limit; a larger term is clamped and reported

// Per-particle accumulator: fixed-point gradient sums, FP64 energy, and
// whether any term was clamped to the fixed-point range (reported by the
// step as an error instead of being summed silently).
struct RdAcc {
    long long x, y, z;
    double e;
    bool sat;
};

// s += floor(v * 2^32), clamped to the range: the integer part and the exact
// fraction from two 32-bit conversions (instead of a float -> int64
// conversion sequence), added as the low word and, with its carry, the high
// word
__device__ __forceinline__ void fix_add(long long& s, float v)
{
    v = fminf(fmaxf(v, -RD_FIX_MAX), RD_FIX_MAX);
    const float fl = floorf(v);
    const uint32_t hi = (uint32_t)(int)fl;
    const uint32_t lo = (uint32_t)((v - fl) * RD_FIX_SCALE);
    unsigned int c;
    const uint32_t sl = __builtin_addc((uint32_t)s, lo, 0u, &c);
    const uint32_t sh = __builtin_addc((uint32_t)((unsigned long long)s >> 32), hi, c, &c);
    s = (long long)(((unsigned long long)sh << 32) | sl);
}

__device__ __forceinline__ void acc_add(RdAcc& a, float pre, float dx, float dy, float dz)
{
    const float fx = pre * dx, fy = pre * dy, fz = pre * dz;
    a.sat = a.sat || !(fabsf(fx) < RD_FIX_MAX && fabsf(fy) < RD_FIX_MAX && fabsf(fz) < RD_FIX_MAX);
    fix_add(a.x, fx); fix_add(a.y, fy); fix_add(a.z, fz);
}

// Per-block LDS copy of the type-pair tables and EM coefficients: the inner
// loop then reads them with LDS latency instead of a global round trip.
struct RdTab {
    float rc2[RD_MAXT * RD_MAXT];
    float r0[RD_MAXT * RD_MAXT];
    uint32_t mask[RD_MAXT * RD_MAXT];
    float a[RD_MAXT], b[RD_MAXT], prad[RD_MAXT];
    RdPot2Dev pot2[RD_MAXP2];
    int npot;
};

// all threads of the block must call this (it ends with a barrier).  BLOCK
// is the workgroup size: the loop strides are constants (blockDim.x would be
// read from the dispatch packet, a dependent load before each loop), and
// every thread issues all its loads before its LDS stores, so that the table
// costs one global latency at the start of a workgroup instead of several.
template <int BLOCK>
__device__ __forceinline__ void load_tab(RdTab& T, const RdDev* __restrict__ P)
{
    constexpr int NP = (RD_MAXT * RD_MAXT + BLOCK - 1) / BLOCK;
    constexpr int NW = (int)(sizeof(RdPot2Dev) / 4) * RD_MAXP2;
    constexpr int NQ = (NW + BLOCK - 1) / BLOCK;
    const int t = threadIdx.x;
    const int n2 = P->nt * P->nt;
    float rc2[NP], r0[NP];
    uint32_t mk[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int k = t + i * BLOCK;
        if (k < n2) { rc2[i] = P->rc2[k]; r0[i] = P->r0[k]; mk[i] = P->p2mask[k]; }
    }
    float a = 0.0f, b = 0.0f, prad = 0.0f;
    if (t < RD_MAXT) { a = P->a[t]; b = P->b[t]; prad = P->prad[t]; }
    uint32_t pw[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int k = t + i * BLOCK;
        if (k < NW) pw[i] = reinterpret_cast<const uint32_t*>(P->pot2)[k];
    }
    const int np = P->n_pot2;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int k = t + i * BLOCK;
        if (k < n2) { T.rc2[k] = rc2[i]; T.r0[k] = r0[i]; T.mask[k] = mk[i]; }
    }
    if (t < RD_MAXT) { T.a[t] = a; T.b[t] = b; T.prad[t] = prad; }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int k = t + i * BLOCK;
        if (k < NW) reinterpret_cast<uint32_t*>(T.pot2)[k] = pw[i];
    }
    if (t == 0) T.npot = np;
    __syncthreads();
}

// Gradient contribution of q on p for the separation d = p - q (|d|^2 = d2),
// added to the fixed-point sums.  Every force kernel evaluates a pair through
// this one function, so all of them produce the same bits.  LISTED: the pair
// already passed the same range test (0 < d2 < rc2) on the same d2.
template <bool ENERGY, bool LISTED = false>
__device__ __forceinline__ void pair_add(const RdTab& T, int rowbase, float dx, float dy, float dz, float d2, int tj,
                                         RdAcc& acc)
{
    if (LISTED || (d2 < T.rc2[rowbase + tj] && d2 > 0.0f)) {
        const float rinv = __builtin_amdgcn_rsqf(d2);
        const float r = d2 * rinv;
        const float r0 = T.r0[rowbase + tj];
        const uint32_t m = T.mask[rowbase + tj];
        float ep = 0.0f;
        const float pre = pair_pre_wave<ENERGY>(T.pot2, T.npot, m, r0, r, rinv, &ep);
        acc_add(acc, pre, dx, dy, dz);
        if (ENERGY) acc.e += (double)(0.5f * ep);
    }
}

__device__ __forceinline__ float dist2(float dx, float dy, float dz)
{
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}

// the pair (p, q) with the minimum-image separation
template <bool ENERGY>
__device__ __forceinline__ void pair_accum(const RdTab& T, int rowbase, float4 p, float4 q, const RdRep& R,
                                           RdAcc& acc)
{
    float dx = min_image(p.x - q.x, R.Lb[0], R.half[0]);
    float dy = min_image(p.y - q.y, R.Lb[1], R.half[1]);
    float dz = min_image(p.z - q.z, R.Lb[2], R.half[2]);
    pair_add<ENERGY>(T, rowbase, dx, dy, dz, dist2(dx, dy, dz), w_type(q.w), acc);
}

__device__ __forceinline__ bool interior_cell(const RdRep& R, int cx, int cy, int cz)
{
    return cx >= 1 && cx <= R.L[0] - 2 && cy >= 1 && cy <= R.L[1] - 2 && cz >= 1 && cz <= R.L[2] - 2;
}

// external potentials, added to the fixed-point pair sum; a gradient computed
// by a separate external pass (composite geometry, surrogate) is added first
template <bool ENERGY>
__device__ __forceinline__ float3 finish_grad(const RdDev* __restrict__ P, const RdTab& T,
                                              const RdPot1Dev* __restrict__ pot1, uint32_t s, float4 p,
                                              const RdAcc& acc, double* e)
{
    float gx = (float)((double)acc.x * RD_FIX_INV);
    float gy = (float)((double)acc.y * RD_FIX_INV);
    float gz = (float)((double)acc.z * RD_FIX_INV);
    const float4* __restrict__ gext = P->gext;
    const int ti = w_type(p.w);
    if (gext && ((P->ext_mask >> ti) & 1u)) {
        const float4 q = gext[s];
        gx = gx + q.x; gy = gy + q.y; gz = gz + q.z;
        if (ENERGY) *e += (double)q.w;
    }
    const int np1 = P->n_pot1;
    const float prad = T.prad[ti];
    for (int k = 0; k < np1; ++k) {
        const RdPot1Dev& q = pot1[k];
        if (!((q.mask >> ti) & 1u)) continue;
        float ek = 0.0f;
        ext_grad<ENERGY>(q, prad, p.x, p.y, p.z, gx, gy, gz, &ek);
        if (ENERGY) *e += (double)ek;
    }
    return make_float3(gx, gy, gz);
}

// bonded partners of the particle in slot s (group potentials,
// CudaSimulation.cu:1829-1836): minimum-image distance, no cutoff, added to
// the fixed-point pair sum
template <bool ENERGY>
__device__ __forceinline__ void bond_accum(const RdDev* __restrict__ P, const RdRep& R, const RdTab& T,
                                           const float4* __restrict__ ps, float4 p, RdAcc& acc)
{
    const uint32_t gi = (uint32_t)R.off + w_index(p.w);
    const int b1 = P->bstart[gi + 1];
    const int rowbase = w_type(p.w) * P->nt;
    for (int b = P->bstart[gi]; b < b1; ++b) {
        const int2 bd = P->bonds[b];
        const float4 q = ps[P->slot_of[bd.x]];
        const float dx = min_image(p.x - q.x, R.Lb[0], R.half[0]);
        const float dy = min_image(p.y - q.y, R.Lb[1], R.half[1]);
        const float dz = min_image(p.z - q.z, R.Lb[2], R.half[2]);
        const float d2 = dist2(dx, dy, dz);
        if (!(d2 > 0.0f)) continue;
        const float rinv = __builtin_amdgcn_rsqf(d2);
        const float r = d2 * rinv;
        const float fr0 = P->bpot_r0[bd.y];
        const float r0 = fr0 >= 0.0f ? fr0 : T.r0[rowbase + w_type(q.w)];
        float ep = 0.0f;
        const float pre = pair_pre_fast<ENERGY>(P->bpot[bd.y], r0, r, rinv, &ep);
        acc_add(acc, pre, dx, dy, dz);
        if (ENERGY) acc.e += (double)(0.5f * ep);
    }
}

// per-wave list of candidate pairs that passed the distance test (slot of
// the partner), transposed so that a wave's lanes write one LDS row
#define RD_HITS 16
struct RdHitList {
    uint32_t j[RD_HITS][64];
};

// gradient of U (pair + external) for the particle in slot s, from global
// memory (the per-particle gather)
template <bool ENERGY>
__device__ __forceinline__ float3 gather_grad(const RdDev* __restrict__ P, const RdRep& R, const RdTab& T,
                                              const RdPot1Dev* __restrict__ pot1,
                                              const float4* __restrict__ ps, const uint32_t* __restrict__ start,
                                              uint32_t s, float4 p, double* energy, bool* sat,
                                              RdHitList* __restrict__ il)
{
    const int ti = w_type(p.w);
    const int rowbase = ti * P->nt;
    RdAcc acc;             // pair, bond and external energies accumulate in FP64
    acc.x = 0; acc.y = 0; acc.z = 0; acc.e = 0.0; acc.sat = false;
    const int Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    int cx, cy, cz;
    cell_coords(R, p.x, p.y, p.z, cx, cy, cz);
    if (P->n_pot2 > 0) {
        if (P->S == 1 && Lx >= 3 && Ly >= 3 && Lz >= 3) {
            const bool in = interior_cell(R, cx, cy, cz);
            const int lane = threadIdx.x & 63;
            // The walks address the particle array from a copy of its address
            // in vector registers (an identity permutation across the lanes
            // hides that it is uniform): held in scalar registers, it was
            // evicted by the values the walks keep there and reloaded from
            // the kernel arguments, with a wait, on every iteration.
            typedef float rd_f4 __attribute__((ext_vector_type(4)));
            typedef const __attribute__((address_space(1))) rd_f4* gptr;
            gptr pva;
            {
                const uint64_t a = (uint64_t)(uintptr_t)ps;
                const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(lane << 2, (int)(uint32_t)a);
                const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(lane << 2, (int)(uint32_t)(a >> 32));
                pva = (gptr)(((uint64_t)hi << 32) | lo);
            }
            auto pv = [&](uint32_t j) {
                const rd_f4 v = pva[j];
                return make_float4(v.x, v.y, v.z, v.w);
            };
            // Phase 1: every stencil candidate gets the cheap distance test
            // against the range of its type pair (one LDS table read); hits
            // are listed in LDS.
            // Phase 2 evaluates the listed pairs only, so the interaction
            // code runs (max hits over the wave) times, not once per
            // candidate iteration in which some lane has a hit.
            // The 27-cell stencil is 9 rows of at most two runs of
            // contiguous slots (two only for homes at the x edge of the
            // lattice): runs 0..8 are the first runs of rows 0..8, runs
            // 9..17 the second runs.  Runs are walked in groups of three,
            // whose six bounds are loaded together (independent loads in
            // flight at once).
            // Dense stencils: a lane whose list fills notes its first
            // unlisted hit (run, slot) and stops listing; before its next
            // run the wave leaves the listing-only walk for a second one
            // that goes on from there, listing while a lane's list has room
            // and evaluating further hits as they are met.  Keeping the
            // interaction code out of the listing-only loops keeps their
            // registers free (inlined there, it made the compiler reload the
            // particle array's address from the kernel arguments on every
            // iteration).
            int nhit = 0;                   // hits met; more than RD_HITS: the list is full
            int cur_run = 0, res_run = 0;   // (res_run, res_j): the first unlisted hit
            uint32_t res_j = 0u;
            auto delta = [&](float4 q, float& dx, float& dy, float& dz) {
                dx = p.x - q.x; dy = p.y - q.y; dz = p.z - q.z;
                if (!in) {
                    dx = min_image(dx, R.Lb[0], R.half[0]);
                    dy = min_image(dy, R.Lb[1], R.half[1]);
                    dz = min_image(dz, R.Lb[2], R.half[2]);
                }
            };
            auto test = [&](uint32_t j, float4 q) {           // listing only
                float dx, dy, dz;
                delta(q, dx, dy, dz);
                const float d2 = dist2(dx, dy, dz);
                if (d2 < T.rc2[rowbase + w_type(q.w)] && d2 > 0.0f) {
                    if (nhit < RD_HITS) il->j[nhit][lane] = j;
                    const bool first = nhit == RD_HITS;
                    res_j = first ? j : res_j;
                    res_run = first ? cur_run : res_run;
                    ++nhit;
                }
            };
            auto test_dense = [&](uint32_t j, float4 q) {     // listing, then evaluating
                float dx, dy, dz;
                delta(q, dx, dy, dz);
                const float d2 = dist2(dx, dy, dz);
                if (d2 < T.rc2[rowbase + w_type(q.w)] && d2 > 0.0f) {
                    if (nhit < RD_HITS) {
                        il->j[nhit][lane] = j;
                        ++nhit;
                    } else {
                        pair_add<ENERGY, true>(T, rowbase, dx, dy, dz, d2, w_type(q.w), acc);
                    }
                }
            };
            int a0, b0, a1, b1;
            const int nr = x_runs(cx, 1, Lx, &a0, &b0, &a1, &b1);
            const bool two = __any(nr == 2);
            const int nruns = two ? 18 : 9;
            auto row_of = [&](int r) {
                int yy = cy + (r % 3) - 1, zz = cz + (r / 3) - 1;
                yy += yy < 0 ? Ly : 0; yy -= yy >= Ly ? Ly : 0;
                zz += zz < 0 ? Lz : 0; zz -= zz >= Lz ? Lz : 0;
                return R.coff + (zz * Ly + yy) * Lx;
            };
            auto bounds = [&](int run, uint32_t& j0, uint32_t& j1) {
                j0 = 0u; j1 = 0u;
                if (run < 9 || nr == 2) {
                    // both bounds from one 16-byte load: a run is at most
                    // three cells, so its end is one of the next three
                    // starts (the start array has 16 entries of padding)
                    const int row = row_of(run < 9 ? run : run - 9);
                    const int a = run < 9 ? a0 : a1, len = (run < 9 ? b0 : b1) + 1 - a;
                    uint4 v;
                    __builtin_memcpy(&v, start + row + a, sizeof(v));
                    j0 = v.x;
                    j1 = len == 3 ? v.w : (len == 2 ? v.z : v.y);
                }
            };
            // slots j .. j1-1 two at a time, the next two loaded while the
            // current two are tested (the gather is bound by load latency;
            // the same in the dense walk below measured slower on the 1M
            // workload, whose register count it raises from 54 to 60)
            // (the loads step a pointer, so that each pair of candidates costs
            // one address update instead of two index-to-address conversions)
            auto walk = [&](uint32_t j, uint32_t j1) {
                if (j + 1 < j1) {
                    gptr a = pva + j;
                    rd_f4 q0 = a[0], q1 = a[1];
                    for (; j + 3 < j1; j += 2) {
                        a += 2;
                        const rd_f4 n0 = a[0], n1 = a[1];
                        test(j, make_float4(q0.x, q0.y, q0.z, q0.w));
                        test(j + 1, make_float4(q1.x, q1.y, q1.z, q1.w));
                        q0 = n0;
                        q1 = n1;
                    }
                    test(j, make_float4(q0.x, q0.y, q0.z, q0.w));
                    test(j + 1, make_float4(q1.x, q1.y, q1.z, q1.w));
                    j += 2;
                }
                if (j < j1) test(j, pv(j));
            };
            int rnext = nruns;              // first run the listing-only walk did not reach
            bool stop = false;
#pragma unroll
            for (int g = 0; g < 18; g += 3) {
                if (g < nruns && !stop) {
                    uint32_t sb[3], se[3];
#pragma unroll
                    for (int r = 0; r < 3; ++r) bounds(g + r, sb[r], se[r]);
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        if (!stop) {
                            if (__any(nhit > RD_HITS)) {
                                stop = true;
                                rnext = g + k;
                            } else {
                                cur_run = g + k;
                                walk(sb[k], se[k]);
                            }
                        }
                    }
                }
            }
            const bool over = nhit > RD_HITS;
            nhit = min(nhit, RD_HITS);
            if (__any(over)) {
                // go on from the run in which lists filled (rnext - 1, the
                // last one walked): lanes whose list filled from their first
                // unlisted hit, the others from run rnext
#pragma unroll 1
                for (int r0 = rnext - 1; r0 < nruns; r0 += 3) {
                    uint32_t sb[3], se[3];
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        const int run = r0 + r;
                        sb[r] = 0u;
                        se[r] = 0u;
                        if (run < nruns && run >= (over ? res_run : rnext)) {
                            bounds(run, sb[r], se[r]);
                            if (over && run == res_run) sb[r] = res_j;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        uint32_t j = sb[k];
                        const uint32_t j1 = se[k];
                        for (; j + 1 < j1; j += 2) {
                            const float4 q0 = pv(j), q1 = pv(j + 1);
                            test_dense(j, q0);
                            test_dense(j + 1, q1);
                        }
                        if (j < j1) test_dense(j, pv(j));
                    }
                }
            }
            // Phase 2, the next listed partner loaded while one is evaluated
            if (__any(nhit > 0)) {
                float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
                if (nhit > 0) q = pv(il->j[0][lane]);
                for (int k = 0; __any(k < nhit); ++k) {
                    float4 n = q;
                    if (k + 1 < nhit) n = pv(il->j[k + 1][lane]);
                    if (k < nhit) {
                        float dx, dy, dz;
                        delta(q, dx, dy, dz);
                        pair_add<ENERGY, true>(T, rowbase, dx, dy, dz, dist2(dx, dy, dz), w_type(q.w), acc);
                    }
                    q = n;
                }
            }
        } else {
            for_stencil(R, start, cx, cy, cz, P->S, [&](uint32_t j) {
                if (j != s) pair_accum<ENERGY>(T, rowbase, p, ps[j], R, acc);
            });
        }
    }
    if (P->bstart) bond_accum<ENERGY>(P, R, T, ps, p, acc);
    double e = acc.e;
    float3 g = finish_grad<ENERGY>(P, T, pot1, s, p, acc, &e);
    if (ENERGY) *energy = e;
    *sat = acc.sat;
    return g;
}

__device__ __forceinline__ float wrap_coord(float x, float lo, float hi, float Lb)
{
    if (x > hi) x -= Lb;
    if (x < lo) x += Lb;
    return x;
}

// Euler-Maruyama for slot s (CudaSimulation.cu:1850-1858): noise, x += -a*dU
// + b*xi, one periodic correction, the next cell key and its histogram entry.
// The reference repeats the correction until the particle is inside
// (:1857-1858); a particle that is still outside after one correction moved
// more than a box length in one step, which is reported as an error
// (RD_ERR_OUTSIDE) rather than wrapped.  The noise counter is (local index,
// step, replica stream); a caller-supplied noise array is indexed by the
// global particle index.
__device__ __forceinline__ void em_update_store(const RdDev* __restrict__ P, const RdRep& R, const RdTab& T,
                                                RdStepState* __restrict__ st, const float* __restrict__ noise,
                                                int ntot, uint32_t s, float4 p, float3 g, bool sat,
                                                float4* __restrict__ pn,
                                                uint2* __restrict__ kr,
                                                uint32_t* __restrict__ count, uint32_t* __restrict__ tsum)
{
    const int t = w_type(p.w);
    const uint32_t idx = w_index(p.w);
    const unsigned long long step = st->step;
    float3 z;
    if (st->noise_nsteps == 0) {
        z = rd_normal3(P->seed, R.stream, step, idx);
    } else {
        long long ls = (long long)(step - st->noise_step0);
        if (ls >= 0 && ls < st->noise_nsteps) {
            size_t base = (size_t)ls * 3 * (size_t)ntot + (size_t)R.off + idx;
            z = make_float3(noise[base], noise[base + ntot], noise[base + 2 * (size_t)ntot]);
        } else {
            z = make_float3(0.0f, 0.0f, 0.0f);
            atomicOr(&st->err, RD_ERR_NOISE);
        }
    }
    const float a = T.a[t], bb = T.b[t];
    float x = (p.x - a * g.x) + bb * z.x;
    float y = (p.y - a * g.y) + bb * z.y;
    float zc = (p.z - a * g.z) + bb * z.z;
    x = wrap_coord(x, R.lo[0], R.hi[0], R.Lb[0]);
    y = wrap_coord(y, R.lo[1], R.hi[1], R.Lb[1]);
    zc = wrap_coord(zc, R.lo[2], R.hi[2], R.Lb[2]);
    const bool inside = x >= R.lo[0] && x <= R.hi[0] && y >= R.lo[1] && y <= R.hi[1] && zc >= R.lo[2] &&
                        zc <= R.hi[2];
    if (!inside || sat) {
        const bool fin = fabsf(x) < 3.0e38f && fabsf(y) < 3.0e38f && fabsf(zc) < 3.0e38f;
        atomicOr(&st->err, (fin ? (inside ? 0u : RD_ERR_OUTSIDE) : RD_ERR_NONFINITE) | (sat ? RD_ERR_FIXRANGE : 0u));
    }
    pn[s] = make_float4(x, y, zc, p.w);
    if (P->gid_n) P->gid_n[s] = (uint32_t)R.off + idx;
    uint32_t k = cell_key(R, x, y, zc);
    kr[s] = make_uint2(k, cell_histogram(count, tsum, k));
}

// the wave's work item: replica and slot range (waves never span replicas)
__device__ __forceinline__ RdWave wave_item(const RdWave* __restrict__ waves, int nwaves, int w)
{
    RdWave it;
    if (w < nwaves) it = waves[w];
    else { it.rep = 0; it.s0 = 0; it.s1 = 0; it.pad = 0; }
    return it;
}
