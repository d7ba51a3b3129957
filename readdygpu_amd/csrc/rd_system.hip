// rd_system.hip -- the BD step behind the rd_* C ABI.
//
// Host side mirrors CudaSimulation (CudaSimulation.cuh:39-116): set-up, step,
// transfers, RDF.  A handle holds one system or a batch of independent
// replicas; all replicas advance together through one set of kernels.
#include "rd_host.h"
#include <hip/hip_ext.h>
#include "rd_kernels.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

// ===========================================================================
// Kernels
// ===========================================================================

// Pack unsorted host input into slot form and histogram the cells.
__global__ __launch_bounds__(RD_BLOCK) void k_pack(const float* __restrict__ xyz, const int* __restrict__ type,
                                                   const int* __restrict__ rep_of, int n,
                                                   const RdRep* __restrict__ reps, float4* __restrict__ pn,
                                                   uint2* __restrict__ kr,
                                                   uint32_t* __restrict__ count, uint32_t* __restrict__ tsum,
                                                   uint32_t* __restrict__ gid_n)
{
    int i = blockIdx.x * RD_BLOCK + threadIdx.x;
    if (i >= n) return;
    const RdRep R = reps[rep_of[i]];
    float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    pn[i] = make_float4(x, y, z, w_pack((uint32_t)(i - R.off), type[i]));
    if (gid_n) gid_n[i] = (uint32_t)i;
    uint32_t k = cell_key(R, x, y, z);
    kr[i] = make_uint2(k, cell_histogram(count, tsum, k));
}

// wave index and lane of this thread; blocks of RD_BLOCK = 4 waves
__device__ __forceinline__ int wave_id() { return xcd_block(blockIdx.x, gridDim.x) * (RD_BLOCK / 64) + (threadIdx.x >> 6); }

// One BD step for the particle in slot s with the per-particle global gather
// (CudaSimulation.cu:2314-2465 fused): gradient, Euler-Maruyama
// (CudaSimulation.cu:1852), periodic wrap (:1857-1858), next cell key.
// BLOCK threads per workgroup and PPT particles per lane (the autotuner's
// choices; a wave takes PPT consecutive 64-slot work items one after another);
// the type-pair table is staged in LDS once per workgroup.
#define RD_FORCE_ARGS                                                                                   \
    const float4 *__restrict__ ps, float4 *__restrict__ pn, uint2 *__restrict__ kr,                     \
        uint32_t *__restrict__ count, uint32_t *__restrict__ tsum, const uint32_t *__restrict__ start, \
        const RdDev *__restrict__ P, const RdRep *__restrict__ reps, const RdWave *__restrict__ waves, \
        int nwaves, const RdPot1Dev *__restrict__ pot1, RdStepState *__restrict__ st,                   \
        const float *__restrict__ noise, int ntot
template <int BLOCK, int PPT>
__device__ __forceinline__ void force_update(RD_FORCE_ARGS)
{
    __shared__ RdTab T;
    __shared__ RdHitList il[BLOCK / 64];
    const int w = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / 64) + (threadIdx.x >> 6)) * PPT;
    if constexpr (PPT == 1) {
        const RdWave it = wave_item(waves, nwaves, __builtin_amdgcn_readfirstlane(w));
        const uint32_t s = (uint32_t)(it.s0 + (threadIdx.x & 63));
        const bool act = s < (uint32_t)it.s1;
        // the particle's record is loaded before the table, so that the two
        // loads' latencies overlap instead of following each other
        float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (act) p = ps[s];
        load_tab<BLOCK>(T, P);
        if (!act) return;
        const RdRep R = reps[it.rep];
        bool sat;
        float3 g = gather_grad<false>(P, R, T, pot1, ps, start, s, p, nullptr, &sat, &il[threadIdx.x >> 6]);
        em_update_store(P, R, T, st, noise, ntot, s, p, g, sat, pn, kr, count, tsum);
    } else {
        load_tab<BLOCK>(T, P);
        for (int k = 0; k < PPT; ++k) {
            const RdWave it = wave_item(waves, nwaves, __builtin_amdgcn_readfirstlane(w + k));
            const uint32_t s = (uint32_t)(it.s0 + (threadIdx.x & 63));
            if (s < (uint32_t)it.s1) {
                const RdRep R = reps[it.rep];
                float4 p = ps[s];
                bool sat;
                float3 g = gather_grad<false>(P, R, T, pot1, ps, start, s, p, nullptr, &sat, &il[threadIdx.x >> 6]);
                em_update_store(P, R, T, st, noise, ntot, s, p, g, sat, pn, kr, count, tsum);
            }
        }
    }
}

// One particle per lane: the register budget is held to 64 VGPRs (8 waves per
// SIMD where the LDS allows it): the gather is bound by the latency of its
// dependent loads, and waves in flight hide it.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_force_update(RD_FORCE_ARGS)
{
    force_update<BLOCK, 1>(ps, pn, kr, count, tsum, start, P, reps, waves, nwaves, pot1, st, noise, ntot);
}

// PPT > 1 particles per lane (the compiler's own register budget)
template <int BLOCK, int PPT>
__global__ __launch_bounds__(BLOCK) void k_force_update_ppt(RD_FORCE_ARGS)
{
    force_update<BLOCK, PPT>(ps, pn, kr, count, tsum, start, P, reps, waves, nwaves, pot1, st, noise, ntot);
}

// Resident-grid form: occupancy x CUs workgroups, each owning a contiguous
// share of the 64-slot work items (shares in XCD order, as xcd_block maps
// the per-item form, so one XCD's waves work on neighbouring items and share
// its L2).  The waves of a workgroup take the share's items from a ticket in
// LDS until none are left: a wave that finishes early takes the next item, as
// the dispatcher would refill its slot, while the type-pair table is loaded
// once per workgroup instead of once per 4 items.  (Tickets in global memory,
// one per XCD, serialized ~5,000 atomics per address per step: 3.8x slower.)
// Every wave leaves the loop when the ticket passes the share's end.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_force_update_queue(RD_FORCE_ARGS)
{
    __shared__ RdTab T;
    __shared__ RdHitList il[BLOCK / 64];
    __shared__ unsigned int ticket;
    if (threadIdx.x == 0) ticket = 0u;
    load_tab<BLOCK>(T, P);                  // ends with a workgroup barrier
    const int lane = threadIdx.x & 63;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int per = (nwaves + (int)gridDim.x - 1) / (int)gridDim.x;
    const int lo = b * per, hi = min(nwaves, lo + per);
    for (;;) {
        int w = 0;
        if (lane == 0) w = lo + (int)atomicAdd(&ticket, 1u);
        w = __builtin_amdgcn_readfirstlane(__shfl(w, 0, 64));
        if (w >= hi) break;
        const RdWave it = waves[w];
        const uint32_t s = (uint32_t)(it.s0 + lane);
        if (s < (uint32_t)it.s1) {
            const RdRep R = reps[it.rep];
            const float4 p = ps[s];
            bool sat;
            const float3 g = gather_grad<false>(P, R, T, pot1, ps, start, s, p, nullptr, &sat, &il[threadIdx.x >> 6]);
            em_update_store(P, R, T, st, noise, ntot, s, p, g, sat, pn, kr, count, tsum);
        }
    }
}

// gradient + energy (FP64 accumulation) at the current positions
__global__ __launch_bounds__(RD_BLOCK) void k_gradients(const float4* __restrict__ ps,
                                                        const uint32_t* __restrict__ start,
                                                        const RdDev* __restrict__ P, const RdRep* __restrict__ reps,
                                                        const RdWave* __restrict__ waves, int nwaves,
                                                        const RdPot1Dev* __restrict__ pot1,
                                                        float* __restrict__ grad, double* __restrict__ etot,
                                                        RdStepState* __restrict__ st)
{
    __shared__ double red[RD_BLOCK / 64];
    __shared__ RdTab T;
    __shared__ RdHitList il[RD_BLOCK / 64];
    load_tab<RD_BLOCK>(T, P);
    const RdWave it = wave_item(waves, nwaves, __builtin_amdgcn_readfirstlane(wave_id()));
    const uint32_t s = (uint32_t)(it.s0 + (threadIdx.x & 63));
    double e = 0.0;
    if (s < (uint32_t)it.s1) {
        const RdRep R = reps[it.rep];
        float4 p = ps[s];
        double ef = 0.0;
        bool sat;
        float3 g = gather_grad<true>(P, R, T, pot1, ps, start, s, p, &ef, &sat, &il[threadIdx.x >> 6]);
        if (sat) atomicOr(&st->err, RD_ERR_FIXRANGE);
        size_t gi = (size_t)R.off + w_index(p.w);
        grad[3 * gi] = g.x; grad[3 * gi + 1] = g.y; grad[3 * gi + 2] = g.z;
        e = ef;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < RD_BLOCK / 64; ++w) t += red[w];
        atomicAdd(etot, t);
    }
}

// ---- exclusive scan of count[ncell] -> start[ncell+1], zeroing count ----
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// block-wide exclusive scan of one value per thread (NT threads); returns
// the block total
template <int NT = RD_BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* excl, uint32_t* lds)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        uint32_t t = lds[w];
        if (w < wid) off += t;
        tot += t;
    }
    *excl = off + inc - v;
    __syncthreads();
    return tot;
}

__global__ __launch_bounds__(RD_BLOCK) void k_scan_reduce(const uint32_t* __restrict__ count, int ncell,
                                                          uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t lds[RD_BLOCK / 64];
    size_t base = (size_t)blockIdx.x * RD_SCAN_TILE + (size_t)threadIdx.x * RD_SCAN_ITEMS;
    uint32_t v = 0;
    if (base + RD_SCAN_ITEMS <= (size_t)ncell) {
        const uint4* c4 = reinterpret_cast<const uint4*>(count + base);
#pragma unroll
        for (int k = 0; k < RD_SCAN_ITEMS / 4; ++k) {
            uint4 q = c4[k];
            v += q.x + q.y + q.z + q.w;
        }
    } else {
        for (int k = 0; k < RD_SCAN_ITEMS; ++k)
            if (base + k < (size_t)ncell) v += count[base + k];
    }
    uint32_t ex;
    uint32_t tot = block_excl_scan(v, &ex, lds);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Each block (NT threads, one scan tile) adds up the tile sums before it (a
// few KB, L2-resident), scans its own tile, writes start and zeroes count;
// the last block writes start[ncell] = total.  512 threads (4 cells each)
// halve each thread's share of that prefix against 256 (8 cells each).
#define RD_SCAN_APPLY_THREADS 512
template <int NT>
__global__ __launch_bounds__(NT) void k_scan_apply(uint32_t* __restrict__ count, int ncell,
                                                   const uint32_t* __restrict__ bsum,
                                                   uint32_t* __restrict__ start)
{
    constexpr int ITEMS = RD_SCAN_TILE / NT;
    static_assert(ITEMS % 4 == 0 && ITEMS * NT == RD_SCAN_TILE, "whole 16-byte groups per thread");
    __shared__ uint32_t lds[NT / 64];
    __shared__ uint32_t pre_sh;
    // this tile's counts are loaded first so that their latency overlaps the
    // sum over the earlier tiles
    size_t base = (size_t)blockIdx.x * RD_SCAN_TILE + (size_t)threadIdx.x * ITEMS;
    uint32_t c[ITEMS];
    bool full = base + ITEMS <= (size_t)ncell;
    if (full) {
        const uint4* c4 = reinterpret_cast<const uint4*>(count + base);
#pragma unroll
        for (int k = 0; k < ITEMS / 4; ++k) {
            uint4 q = c4[k];
            c[4 * k] = q.x; c[4 * k + 1] = q.y; c[4 * k + 2] = q.z; c[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) c[k] = (base + k < (size_t)ncell) ? count[base + k] : 0u;
    }
    {
        uint32_t acc = 0;
        for (int i = threadIdx.x; i < (int)blockIdx.x; i += NT) acc += bsum[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < NT / 64; ++w) t += lds[w];
            pre_sh = t;
            if (blockIdx.x == gridDim.x - 1) start[ncell] = t + bsum[blockIdx.x];
        }
        __syncthreads();
    }
    const uint32_t pre = pre_sh;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) v += c[k];
    uint32_t ex;
    block_excl_scan<NT>(v, &ex, lds);
    uint32_t run = pre + ex;
    if (full) {
        uint4* s4 = reinterpret_cast<uint4*>(start + base);
        uint4* z4 = reinterpret_cast<uint4*>(count + base);
#pragma unroll
        for (int k = 0; k < ITEMS / 4; ++k) {
            uint4 o;
            o.x = run; run += c[4 * k];
            o.y = run; run += c[4 * k + 1];
            o.z = run; run += c[4 * k + 2];
            o.w = run; run += c[4 * k + 3];
            s4[k] = o;
            z4[k] = make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
        for (int k = 0; k < ITEMS; ++k)
            if (base + k < (size_t)ncell) { start[base + k] = run; run += c[k]; count[base + k] = 0u; }
    }
}

// In-place exclusive scan of per-particle pair counts v[0..n) into offsets,
// v[n] = the total; bsum holds the per-tile sums written by k_scan_reduce.
// Each thread reads its own items before the block's barriers and writes
// only them, so counts and offsets can share one array.
__global__ __launch_bounds__(RD_BLOCK) void k_scan_offsets(uint32_t* v, int n, const uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t lds[RD_BLOCK / 64];
    __shared__ uint32_t pre_sh;
    const size_t base = (size_t)blockIdx.x * RD_SCAN_TILE + (size_t)threadIdx.x * RD_SCAN_ITEMS;
    uint32_t c[RD_SCAN_ITEMS];
#pragma unroll
    for (int k = 0; k < RD_SCAN_ITEMS; ++k) c[k] = (base + k < (size_t)n) ? v[base + k] : 0u;
    {
        uint32_t acc = 0;
        for (int i = threadIdx.x; i < (int)blockIdx.x; i += RD_BLOCK) acc += bsum[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < RD_BLOCK / 64; ++w) t += lds[w];
            pre_sh = t;
        }
        __syncthreads();
    }
    const uint32_t pre = pre_sh;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < RD_SCAN_ITEMS; ++k) sum += c[k];
    uint32_t ex;
    block_excl_scan(sum, &ex, lds);
    uint32_t run = pre + ex;
#pragma unroll
    for (int k = 0; k < RD_SCAN_ITEMS; ++k) {
        if (base + k < (size_t)n) v[base + k] = run;
        run += c[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) v[n] = pre + bsum[blockIdx.x];
}

// counting-sort scatter; advances the step counter when `advance` is set.
// Each thread moves RD_SCATTER_PPL records RD_BLOCK slots apart (every load
// and store instruction covers contiguous slots), with all of its key and
// record loads, then all of its start lookups, in flight at once.
#define RD_SCATTER_PPL 4
__global__ __launch_bounds__(RD_BLOCK) void k_scatter(const float4* __restrict__ pn, const uint2* __restrict__ kr,
                                                      const uint32_t* __restrict__ start, float4* __restrict__ ps,
                                                      int n, RdStepState* __restrict__ st, int advance,
                                                      const uint32_t* __restrict__ gid_n, uint32_t* __restrict__ slot_of,
                                                      uint32_t* __restrict__ tsum, int ntile)
{
    const int tid = blockIdx.x * RD_BLOCK + threadIdx.x;
    for (int t = tid; t < ntile; t += gridDim.x * RD_BLOCK) tsum[t] = 0u;
    if (advance && tid == 0) st->step += 1ull;
    if (n <= 0) return;
    // slots past the end are read and moved as slot n - 1 (the same record
    // to the same place, a duplicate of that slot's own store), so that no
    // load or store waits on a branch
    const int s0 = blockIdx.x * (RD_BLOCK * RD_SCATTER_PPL) + threadIdx.x;
    uint2 q[RD_SCATTER_PPL];
    float4 v[RD_SCATTER_PPL];
#pragma unroll
    for (int u = 0; u < RD_SCATTER_PPL; ++u) {
        const int s = min(s0 + u * RD_BLOCK, n - 1);
        q[u] = kr[s];
        v[u] = pn[s];
    }
    uint32_t d[RD_SCATTER_PPL];
#pragma unroll
    for (int u = 0; u < RD_SCATTER_PPL; ++u) d[u] = start[q[u].x] + q[u].y;
#pragma unroll
    for (int u = 0; u < RD_SCATTER_PPL; ++u) ps[d[u]] = v[u];
    if (slot_of) {
#pragma unroll
        for (int u = 0; u < RD_SCATTER_PPL; ++u) slot_of[gid_n[min(s0 + u * RD_BLOCK, n - 1)]] = d[u];
    }
}

// replica of each particle index (replica r owns indices [off_r, off_r + n_r),
// the same range as its slots)
__global__ __launch_bounds__(RD_BLOCK) void k_rep_of(const RdWave* __restrict__ waves, int nwaves,
                                                     int* __restrict__ rep_of)
{
    const RdWave it = wave_item(waves, nwaves, wave_id());
    const int s = it.s0 + (threadIdx.x & 63);
    if (s < it.s1) rep_of[s] = it.rep;
}

// positions back to particle-index order
__global__ __launch_bounds__(RD_BLOCK) void k_unsort(const float4* __restrict__ ps, const RdRep* __restrict__ reps,
                                                     const RdWave* __restrict__ waves, int nwaves,
                                                     float* __restrict__ xyz, int* __restrict__ type)
{
    const RdWave it = wave_item(waves, nwaves, wave_id());
    const int s = it.s0 + (threadIdx.x & 63);
    if (s >= it.s1) return;
    float4 p = ps[s];
    size_t i = (size_t)reps[it.rep].off + w_index(p.w);
    xyz[3 * i] = p.x; xyz[3 * i + 1] = p.y; xyz[3 * i + 2] = p.z;
    type[i] = w_type(p.w);
}

// Reaction compaction (the cReactions protocol, lib/ReaDDyGPU.cpp:303-463,
// on the device): survivors keep their order, so the new index of particle i
// is i minus the number of removed indices below it (binary search in the
// sorted removal list) plus the particles created in earlier replicas;
// removed particles are dropped.
__global__ __launch_bounds__(RD_BLOCK) void k_react_compact(const float4* __restrict__ ps,
                                                            const RdRep* __restrict__ reps,
                                                            const RdWave* __restrict__ waves, int nwaves,
                                                            const int* __restrict__ del, int ndel,
                                                            const int* __restrict__ shift, int nt,
                                                            uint32_t* __restrict__ delcnt,
                                                            float* __restrict__ xyz, int* __restrict__ type)
{
    const RdWave it = wave_item(waves, nwaves, wave_id());
    const int s = it.s0 + (threadIdx.x & 63);
    if (s >= it.s1) return;
    const float4 p = ps[s];
    const int gi = reps[it.rep].off + (int)w_index(p.w);
    int lo = 0, hi = ndel;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (del[mid] < gi) lo = mid + 1;
        else hi = mid;
    }
    if (lo < ndel && del[lo] == gi) {
        atomicAdd(&delcnt[it.rep * nt + w_type(p.w)], 1u);       // removed particles per (replica, type)
        return;
    }
    const size_t ni = (size_t)(gi - lo + shift[it.rep]);
    xyz[3 * ni] = p.x; xyz[3 * ni + 1] = p.y; xyz[3 * ni + 2] = p.z;
    type[ni] = w_type(p.w);
}

// new particles of a reaction list to their places (after the survivors of
// their replica)
__global__ __launch_bounds__(RD_BLOCK) void k_react_place(const float* __restrict__ bxyz, const int* __restrict__ bty,
                                                          const int* __restrict__ bdst, int nb,
                                                          float* __restrict__ xyz, int* __restrict__ type)
{
    const int k = blockIdx.x * RD_BLOCK + threadIdx.x;
    if (k >= nb) return;
    const size_t d = (size_t)bdst[k];
    xyz[3 * d] = bxyz[3 * k]; xyz[3 * d + 1] = bxyz[3 * k + 1]; xyz[3 * d + 2] = bxyz[3 * k + 2];
    type[d] = bty[k];
}

// reference cell index (replica-local) per particle
__global__ __launch_bounds__(RD_BLOCK) void k_cell_ids(const float4* __restrict__ ps, const RdRep* __restrict__ reps,
                                                       const RdWave* __restrict__ waves, int nwaves,
                                                       int* __restrict__ field)
{
    const RdWave it = wave_item(waves, nwaves, wave_id());
    const int s = it.s0 + (threadIdx.x & 63);
    if (s >= it.s1) return;
    const RdRep R = reps[it.rep];
    float4 p = ps[s];
    field[(size_t)R.off + w_index(p.w)] = (int)cell_key(R, p.x, p.y, p.z) - R.coff;
}

// noise arrays out[(s*3+d)*ntot + i], counter (local index, step, replica stream)
__global__ __launch_bounds__(RD_BLOCK) void k_noise_fill(const RdDev* __restrict__ P, const RdRep* __restrict__ reps,
                                                         int nrep, unsigned long long step0, int nsteps, int ntot,
                                                         float* __restrict__ out)
{
    size_t g = (size_t)blockIdx.x * RD_BLOCK + threadIdx.x;
    size_t total = (size_t)nsteps * (size_t)ntot;
    if (g >= total) return;
    int s = (int)(g / (size_t)ntot);
    int i = (int)(g - (size_t)s * ntot);
    int lo = 0, hi = nrep - 1;          // last replica with off <= i
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (reps[mid].off <= i) lo = mid; else hi = mid - 1;
    }
    const RdRep& R = reps[lo];
    float3 z = rd_normal3(P->seed, R.stream, step0 + (unsigned long long)s, (uint32_t)(i - R.off));
    size_t base = (size_t)s * 3 * (size_t)ntot + i;
    out[base] = z.x; out[base + ntot] = z.y; out[base + 2 * (size_t)ntot] = z.z;
}

// RDF histogram (CudaSimulation.cu:1784-1827 counts every ordered pair once
// per side; here each slot counts its own partners).
// counts[((rep*nt + ti)*nt + tj)*nbins + bin]
// Candidates with d2 >= rcut2 are skipped before the square root: rcut2 is
// rmax^2 rounded up by 1e-6 relative, and sqrtf is correctly rounded, so every
// skipped candidate has sqrtf(d2) >= rmax and the counted set is unchanged.
// `wrap` false: the home cell's stencil neither wraps nor touches cell 0
// (which also holds the partial slab at the top of the box), and it spans
// less than half the box, so every raw difference is already its minimum
// image (min_image would return it unchanged).
// bin of the pair (p, q), or -1 beyond rmax
__device__ __forceinline__ int rdf_bin(const RdRep& R, float4 p, float4 q, float rmax, float rcut2, int nbins,
                                       bool wrap)
{
    float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    if (wrap) {
        dx = min_image(dx, R.Lb[0], R.half[0]);
        dy = min_image(dy, R.Lb[1], R.half[1]);
        dz = min_image(dz, R.Lb[2], R.half[2]);
    }
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    if (!(d2 < rcut2)) return -1;
    float r = sqrtf(d2);
    if (!(r < rmax)) return -1;
    int bin = (int)floorf((r * (float)nbins) / rmax);
    if (bin > nbins - 1) bin = nbins - 1;
    return bin;
}

__device__ __forceinline__ bool stencil_wraps(const RdRep& R, int cx, int cy, int cz, int S)
{
    const int L[3] = { R.L[0], R.L[1], R.L[2] }, c[3] = { cx, cy, cz };
    for (int d = 0; d < 3; ++d)
        if (c[d] < S + 1 || c[d] > L[d] - 1 - S || (float)(S + 2) * R.cell >= R.half[d]) return true;
    return false;
}

// Visits every pair of the particle in slot s with the RDF histogram index
// of each direction in which the full-stencil count (every particle counting
// its partners) meets it: add(h, c) with c = 1, or 2 where both directions
// share the index (equal types).  Where the lattice allows, the half stencil
// meets each pair once and adds both directions, half the candidates.
template <class F>
__device__ __forceinline__ void rdf_pairs(const RdRep& R, const uint32_t* __restrict__ start,
                                          const float4* __restrict__ ps, int s, float4 p, int nt, int S, float rmax,
                                          float rcut2, int nbins, F&& add)
{
    const int ti = w_type(p.w);
    int cx, cy, cz;
    cell_coords(R, p.x, p.y, p.z, cx, cy, cz);
    const bool wrap = stencil_wraps(R, cx, cy, cz, S);
    const bool half = 2 * S + 1 < R.L[0] && 2 * S + 1 < R.L[1] && 2 * S + 1 < R.L[2];
    if (half) {
        for_stencil_half_q(R, start, ps, (uint32_t)s, cx, cy, cz, S, [&](uint32_t j, float4 q) {
            const int b = rdf_bin(R, p, q, rmax, rcut2, nbins, wrap);
            if (b < 0) return;
            const int tj = w_type(q.w);
            const int h1 = (ti * nt + tj) * nbins + b, h2 = (tj * nt + ti) * nbins + b;
            if (h1 == h2) {
                add(h1, 2u);
            } else {
                add(h1, 1u);
                add(h2, 1u);
            }
        });
    } else {
        for_stencil_q(R, start, ps, cx, cy, cz, S, [&](uint32_t j, float4 q) {
            if ((int)j == s) return;
            const int b = rdf_bin(R, p, q, rmax, rcut2, nbins, wrap);
            if (b >= 0) add((ti * nt + w_type(q.w)) * nbins + b, 1u);
        });
    }
}

// One LDS histogram per workgroup, shared by its waves when their work items
// belong to one replica (a workgroup straddling two replicas adds to the
// global counts directly).  Workgroup b adds its histogram to copy b % ncopy
// of the counts; k_rdf_fold sums the copies, so the 64-bit atomics of
// thousands of workgroups do not queue on the same few thousand addresses.
__global__ __launch_bounds__(RD_BLOCK) void k_rdf_lds(const float4* __restrict__ ps, const uint32_t* __restrict__ start,
                                                      const RdDev* __restrict__ P, const RdRep* __restrict__ reps,
                                                      const RdWave* __restrict__ waves, int nwaves, int S, float rmax,
                                                      float rcut2, int nbins, unsigned long long* __restrict__ part,
                                                      long part_stride, int ncopy,
                                                      unsigned long long* __restrict__ counts)
{
    extern __shared__ uint32_t hist[];
    __shared__ int rep_sh[RD_BLOCK / 64];
    const int nt = P->nt;
    const int nh = nt * nt * nbins;
    const RdWave it = wave_item(waves, nwaves, __builtin_amdgcn_readfirstlane(wave_id()));
    if ((threadIdx.x & 63) == 0) rep_sh[threadIdx.x >> 6] = it.s1 > it.s0 ? it.rep : -1;
    for (int k = threadIdx.x; k < nh; k += RD_BLOCK) hist[k] = 0u;
    __syncthreads();
    int rep0 = -1;
    bool one = true;
    for (int v = 0; v < RD_BLOCK / 64; ++v) {
        const int r = rep_sh[v];
        if (r < 0) continue;
        if (rep0 < 0) rep0 = r;
        else if (r != rep0) one = false;
    }
    const int s = it.s0 + (threadIdx.x & 63);
    if (s < it.s1) {
        const RdRep R = reps[it.rep];
        unsigned long long* cnt = counts + (size_t)it.rep * nh;
        const float4 p = ps[s];
        rdf_pairs(R, start, ps, s, p, nt, S, rmax, rcut2, nbins, [&](int h, uint32_t c) {
            if (one) atomicAdd(&hist[h], c);
            else atomicAdd(&cnt[h], (unsigned long long)c);
        });
    }
    __syncthreads();
    if (one && rep0 >= 0) {
        unsigned long long* cnt = part + (size_t)(blockIdx.x % ncopy) * (size_t)part_stride + (size_t)rep0 * nh;
        for (int k = threadIdx.x; k < nh; k += RD_BLOCK) {
            const uint32_t v = hist[k];
            if (v) atomicAdd(&cnt[k], (unsigned long long)v);
        }
    }
}

// counts += the sum of the copies; the copies are cleared for the next call
__global__ __launch_bounds__(RD_BLOCK) void k_rdf_fold(unsigned long long* __restrict__ part, long n, int ncopy,
                                                       unsigned long long* __restrict__ counts)
{
    const long i = (long)blockIdx.x * RD_BLOCK + threadIdx.x;
    if (i >= n) return;
    unsigned long long v = 0;
    for (int c = 0; c < ncopy; ++c) {
        v += part[(size_t)c * n + i];
        part[(size_t)c * n + i] = 0ull;
    }
    counts[i] += v;
}

// histogram too large for LDS: direct 64-bit atomics on the global counts
__global__ __launch_bounds__(RD_BLOCK) void k_rdf_global(const float4* __restrict__ ps,
                                                         const uint32_t* __restrict__ start,
                                                         const RdDev* __restrict__ P, const RdRep* __restrict__ reps,
                                                         const RdWave* __restrict__ waves, int nwaves, int S,
                                                         float rmax, float rcut2, int nbins,
                                                         unsigned long long* __restrict__ counts)
{
    const int nt = P->nt;
    const int nh = nt * nt * nbins;
    const RdWave it = wave_item(waves, nwaves, wave_id());
    const int s = it.s0 + (threadIdx.x & 63);
    if (s >= it.s1) return;
    const RdRep R = reps[it.rep];
    unsigned long long* cnt = counts + (size_t)it.rep * nh;
    const float4 p = ps[s];
    rdf_pairs(R, start, ps, s, p, nt, S, rmax, rcut2, nbins,
              [&](int h, uint32_t c) { atomicAdd(&cnt[h], (unsigned long long)c); });
}

// pairs (i<j by index) with d2 < thr2[ti][tj]; global particle indices
// Pair sets (neighbour pairs d^2 < rc^2, reaction candidates d^2 < R^2) in
// global-index order without a global sort: a counting pass writes the number
// of partners j > i of every particle i, an exclusive scan turns the counts
// into offsets, and an emitting pass writes each particle's pairs into its own
// segment and sorts that (short) segment by j.
// pair (i, j) of the emitted list: uint2 in the library's own buffer,
// longlong2 where the list is copied straight into the caller's int64 array
__device__ __forceinline__ void set_pair(uint2& o, uint32_t i, uint32_t j) { o = make_uint2(i, j); }
__device__ __forceinline__ void set_pair(longlong2& o, uint32_t i, uint32_t j)
{
    o = make_longlong2((long long)i, (long long)j);
}

template <bool EMIT, class PairT>
__global__ __launch_bounds__(RD_BLOCK) void k_pairs(const float4* __restrict__ ps, const uint32_t* __restrict__ start,
                                                    const RdDev* __restrict__ P, const RdRep* __restrict__ reps,
                                                    const RdWave* __restrict__ waves, int nwaves,
                                                    const float* __restrict__ thr2, int S,
                                                    uint32_t* __restrict__ cnt_off, PairT* __restrict__ out)
{
    const RdWave it = wave_item(waves, nwaves, wave_id());
    const int s = it.s0 + (threadIdx.x & 63);
    if (s >= it.s1) return;
    const RdRep R = reps[it.rep];
    float4 p = ps[s];
    const int nt = P->nt;
    int ti = w_type(p.w);
    uint32_t ii = w_index(p.w);
    const uint32_t gi = (uint32_t)R.off + ii;
    // a type with no partner type (every threshold 0, as for the species
    // that take part in no reaction) has no pairs: no walk
    float tmax = 0.0f;
    for (int k = 0; k < nt; ++k) tmax = fmaxf(tmax, thr2[ti * nt + k]);
    if (!(tmax > 0.0f)) {
        if (!EMIT) cnt_off[gi] = 0u;
        return;
    }
    int cx, cy, cz;
    cell_coords(R, p.x, p.y, p.z, cx, cy, cz);
    const uint32_t o = EMIT ? cnt_off[gi] : 0u;
    const bool wrap = stencil_wraps(R, cx, cy, cz, S);
    uint32_t c = 0;
    for_stencil_q(R, start, ps, cx, cy, cz, S, [&](uint32_t j, float4 q) {
        uint32_t jj = w_index(q.w);
        if (jj <= ii) return;
        float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        if (wrap) {
            dx = min_image(dx, R.Lb[0], R.half[0]);
            dy = min_image(dy, R.Lb[1], R.half[1]);
            dz = min_image(dz, R.Lb[2], R.half[2]);
        }
        float d2 = dx * dx;
        d2 = d2 + dy * dy;
        d2 = d2 + dz * dz;
        if (d2 < thr2[ti * nt + w_type(q.w)]) {
            if (EMIT) {
                // insertion into the sorted segment
                const uint32_t gj = (uint32_t)R.off + jj;
                uint32_t k = c;
                while (k > 0 && (uint32_t)out[o + k - 1].y > gj) {
                    out[o + k] = out[o + k - 1];
                    --k;
                }
                set_pair(out[o + k], gi, gj);
            }
            ++c;
        }
    });
    if (!EMIT) cnt_off[gi] = c;
}

// ===========================================================================
// Host runtime
// ===========================================================================

static int set_err(rd_system* h, int code, const std::string& m)
{
    if (h) h->err = m;
    return code;
}

#define HIPCK(h, x)                                                                            \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            return set_err((h), RD_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

// error bits the step kernels raise (RdStepState::err) as an RD_ERANGE message
static int step_error(rd_system* h, unsigned int err)
{
    if (err & RD_ERR_FIXRANGE)
        return set_err(h, RD_ERANGE, "a pair or bond gradient term exceeds the fixed-point range (|term| >= 5e8)");
    if (err & RD_ERR_NONFINITE) return set_err(h, RD_ERANGE, "non-finite particle position");
    if (err & RD_ERR_OUTSIDE)
        return set_err(h, RD_ERANGE, "a particle moved more than one box length in one step (outside the box after "
                                     "the periodic correction)");
    if (err & RD_ERR_NOISE) return set_err(h, RD_ERANGE, "noise array exhausted");
    return RD_OK;
}

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

static void drop_graph(rd_system* h)
{
    for (auto& kv : h->graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    h->graphs.clear();
}

static void free_bonds(rd_system* h)
{
    void* ptrs[] = { h->d_bstart, h->d_bonds, h->d_slot, h->d_gid };
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h->d_bstart = nullptr; h->d_bonds = nullptr; h->d_slot = nullptr; h->d_gid = nullptr;
}

static void free_frames(rd_system* h)
{
    if (h->d_snap_xyz) (void)hipFree(h->d_snap_xyz);
    if (h->d_snap_type) (void)hipFree(h->d_snap_type);
    if (h->hp_xyz) (void)hipHostFree(h->hp_xyz);
    if (h->hp_type) (void)hipHostFree(h->hp_type);
    if (h->hp_st) (void)hipHostFree(h->hp_st);
    if (h->copy_stream) (void)hipStreamSynchronize(h->copy_stream);
    if (h->ev_frame) (void)hipEventDestroy(h->ev_frame);
    if (h->ev_snap) (void)hipEventDestroy(h->ev_snap);
    if (h->copy_stream) (void)hipStreamDestroy(h->copy_stream);
    h->d_snap_xyz = nullptr; h->d_snap_type = nullptr; h->hp_xyz = nullptr; h->hp_type = nullptr;
    h->hp_st = nullptr; h->ev_frame = nullptr; h->ev_snap = nullptr; h->copy_stream = nullptr;
    h->snap_cap = 0; h->snap_valid = false;
}

static void free_dev(rd_system* h)
{
    free_bonds(h);
    free_frames(h);
    void* ptrs[] = { h->d_ps, h->d_pn, h->d_kr, h->d_count, h->d_start, h->d_bsum,
                     h->d_P, h->d_pot1, h->d_st, h->d_noise, h->d_rdf, h->d_rdf_part, h->d_pairs,
                     h->d_poff, h->d_pbs,
                     h->d_xyz_io, h->d_type_io, h->d_etot, h->d_reps, h->d_waves, h->d_rep_of };
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h->d_ps = h->d_pn = nullptr;
    h->d_kr = nullptr;
    h->d_count = h->d_start = h->d_bsum = nullptr;
    h->d_P = nullptr; h->d_pot1 = nullptr; h->d_st = nullptr; h->d_noise = nullptr;
    h->d_rdf = nullptr; h->d_rdf_part = nullptr; h->d_pairs = nullptr;
    h->d_poff = nullptr; h->d_pbs = nullptr; h->poff_cap = 0;
    h->d_xyz_io = nullptr; h->d_type_io = nullptr; h->d_etot = nullptr;
    h->d_reps = nullptr; h->d_waves = nullptr; h->d_rep_of = nullptr;
    h->cap = 0; h->ncell_alloc = 0; h->pairs_cap = 0; h->noise_cap = 0; h->nrep_alloc = 0; h->nwave_alloc = 0;
}

extern "C" int rd_create(const rd_params* prm, rd_system** out)
{
    if (!prm || !out) return RD_EINVAL;
    *out = nullptr;
    for (int d = 0; d < 3; ++d)
        if (!(prm->lattice_bounds[2 * d + 1] > prm->lattice_bounds[2 * d])) return RD_EINVAL;
    if (!(prm->dt > 0) || !(prm->boltzmann > 0) || !(prm->temperature > 0) || !(prm->max_cutoff > 0))
        return RD_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return RD_EHIP;
    rd_system* h = new rd_system();
    h->prm = *prm;
    h->dev = (prm->device >= 0 && prm->device < ndev) ? prm->device : 0;
    if (hipSetDevice(h->dev) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return RD_EHIP;
    }
    if (hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, h->dev) != hipSuccess || h->ncu < 1)
        h->ncu = 1;
    h->step = 0;
    *out = h;
    return RD_OK;
}

extern "C" void rd_destroy(rd_system* h)
{
    if (!h) return;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    free_dev(h);
    rd_host_free_surrogate(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" const char* rd_last_error(const rd_system* h)
{
    return h ? h->err.c_str() : "null handle";
}

extern "C" int rd_set_particle_types(rd_system* h, int nt, const double* D, const double* rad, const double* radii)
{
    if (!h) return RD_EINVAL;
    if (nt < 1 || nt > RD_MAX_TYPES || !D || !rad) return set_err(h, RD_EINVAL, "bad particle types");
    h->nt = nt;
    h->D.assign(D, D + nt);
    h->rad.assign(rad, rad + nt);
    h->radii.assign((size_t)nt * nt, 0.0);
    if (radii) h->radii.assign(radii, radii + (size_t)nt * nt);
    h->rxn.assign((size_t)nt * nt, 0.0);
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_add_order_one_potential(rd_system* h, const rd_pot1* p)
{
    if (!h || !p) return RD_EINVAL;
    if (p->type < RD_POT1_DISK || p->type > RD_POT1_BOX) return set_err(h, RD_EINVAL, "unknown order-one potential");
    h->pot1.push_back(*p);
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_add_order_two_potential(rd_system* h, const rd_pot2* p, const uint8_t* mask)
{
    if (!h || !p || !mask) return RD_EINVAL;
    if (h->nt == 0) return set_err(h, RD_ESTATE, "set particle types first");
    if ((int)h->pot2.size() >= RD_MAX_POT2) return set_err(h, RD_EINVAL, "too many order-two potentials");
    if (p->type < RD_POT2_HARMONIC || p->type > RD_POT2_PIECEWISE_WELL)
        return set_err(h, RD_EINVAL, "unknown order-two potential");
    if (p->type == RD_POT2_PIECEWISE_WELL && !(p->length > 0)) return set_err(h, RD_EINVAL, "well length must be > 0");
    h->pot2.push_back(*p);
    std::vector<uint8_t> m((size_t)h->nt * h->nt, 0);
    for (int i = 0; i < h->nt; ++i)
        for (int j = 0; j < h->nt; ++j)
            if (mask[i * h->nt + j]) { m[i * h->nt + j] = 1; m[j * h->nt + i] = 1; }
    h->pot2mask.push_back(m);
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_set_reaction_radii(rd_system* h, const double* R)
{
    if (!h || !R) return RD_EINVAL;
    if (h->nt == 0) return set_err(h, RD_ESTATE, "set particle types first");
    h->rxn.assign(R, R + (size_t)h->nt * h->nt);
    for (int i = 0; i < h->nt; ++i)
        for (int j = 0; j < h->nt; ++j) {
            double v = std::max(h->rxn[i * h->nt + j], h->rxn[j * h->nt + i]);
            h->rxn[i * h->nt + j] = h->rxn[j * h->nt + i] = v;
        }
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_set_bonds(rd_system* h, int n_bonds, const int* bonds, const int* pot_index, int n_pots,
                            const rd_pot2* pots, const double* r0)
{
    if (!h || n_bonds < 0 || n_pots < 0) return RD_EINVAL;
    if (n_bonds > 0 && (!bonds || !pot_index || n_pots == 0 || !pots)) return RD_EINVAL;
    static_assert(RD_MAXBP == RD_MAX_BOND_POTS, "bond potential limits differ");
    if (n_pots > RD_MAXBP) return set_err(h, RD_EINVAL, "too many bond potentials (at most " + std::to_string(RD_MAXBP) + ")");
    for (int p = 0; p < n_pots; ++p) {
        const int t = pots[p].type;
        if (t != RD_POT2_HARMONIC && t != RD_POT2_WEAK_HARMONIC && t != RD_POT2_PIECEWISE_WELL)
            return set_err(h, RD_EINVAL, "unknown bond potential type");
    }
    for (int b = 0; b < n_bonds; ++b) {
        const int a = bonds[2 * b], c = bonds[2 * b + 1];
        if (a < 0 || c < 0 || a == c) return set_err(h, RD_EINVAL, "a bond needs two distinct particle indices");
        if (pot_index[b] < 0 || pot_index[b] >= n_pots) return set_err(h, RD_EINVAL, "bond potential index out of range");
    }
    h->bond_pairs.assign(bonds, bonds + 2 * (size_t)n_bonds);
    h->bond_pot.assign(pot_index, pot_index + n_bonds);
    h->bpots.assign(pots, pots + n_pots);
    h->bpot_r0.assign((size_t)n_pots, -1.0);
    if (r0)
        for (int p = 0; p < n_pots; ++p) h->bpot_r0[p] = r0[p];
    h->dirty = true;
    return RD_OK;
}

// Range of the evaluated interaction for a type pair: the largest natural
// range of the potentials that apply, capped at maxCutoff.
static double pair_range(const rd_system* h, int i, int j, double r0)
{
    double m = 0.0;
    for (size_t p = 0; p < h->pot2.size(); ++p) {
        if (!h->pot2mask[p][i * h->nt + j]) continue;
        const rd_pot2& q = h->pot2[p];
        double r = 0.0;
        if (q.type == RD_POT2_HARMONIC) r = (q.subtype == RD_SUB_REPULSIVE) ? r0 : h->prm.max_cutoff;
        else r = r0 + q.length;
        m = std::max(m, std::min(r, h->prm.max_cutoff));
    }
    return m;
}

double rd_host_r0(const rd_system* h, int i, int j)
{
    int nt = h->nt;
    double ri = h->radii[i * nt + j] == 0.0 ? h->rad[i] : h->radii[i * nt + j];
    double rj = h->radii[j * nt + i] == 0.0 ? h->rad[j] : h->radii[j * nt + i];
    return ri + rj;
}

// Build the device constant block (CudaSimulation.cu:237-391 tables) and the
// replica geometry: per replica L_d = (int)((hi-lo)/cell) in FP32
// (CudaSimulation.cu:524-528), cells of all replicas concatenated.
static int build_tables(rd_system* h)
{
    // everything is computed into locals and stored only once every replica
    // passes the lattice checks, so a refused size leaves the handle as it was
    RdDev P;
    memset(&P, 0, sizeof(P));
    const rd_params& g = h->prm;
    int nt = h->nt;
    P.nt = nt;
    double maxr = 0.0, maxrx = 0.0;
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j) {
            double r0 = rd_host_r0(h, i, j);
            double rc = pair_range(h, i, j, r0);
            P.r0[i * nt + j] = (float)r0;
            P.rc2[i * nt + j] = (float)(rc * rc);
            P.rr2[i * nt + j] = (float)(h->rxn[i * nt + j] * h->rxn[i * nt + j]);
            maxr = std::max(maxr, rc);
            maxrx = std::max(maxrx, h->rxn[i * nt + j]);
            uint32_t m = 0;
            for (size_t p = 0; p < h->pot2.size(); ++p)
                if (h->pot2mask[p][i * nt + j]) m |= 1u << p;
            P.p2mask[i * nt + j] = m;
        }
    double mr = std::max(maxr, maxrx);
    double cell = g.cell_size > 0 ? g.cell_size : (mr > 0 ? 1.01 * mr : g.max_cutoff);
    P.S = std::max(1, (int)std::ceil(maxr * 1.0001 / cell));
    long coff = 0;
    int off = 0;
    std::vector<RdRep> rep_dev(h->reps.size());
    for (size_t r = 0; r < h->reps.size(); ++r) {
        const rd_system::Rep& A = h->reps[r];
        RdRep& R = rep_dev[r];
        memset(&R, 0, sizeof(R));
        R.cell = (float)cell;
        long nc = 1;
        for (int d = 0; d < 3; ++d) {
            R.lo[d] = (float)A.lo[d];
            R.hi[d] = (float)A.hi[d];
            R.Lb[d] = R.hi[d] - R.lo[d];
            R.half[d] = R.Lb[d] / 2.0f;
            const float q = R.Lb[d] / R.cell;
            if (!(q < 268435456.0f)) return set_err(h, RD_ERANGE, "too many cells; raise cell_size");
            int l = (int)q;
            R.L[d] = l < 1 ? 1 : l;
            nc *= R.L[d];
            if (coff + nc > (1L << 28)) return set_err(h, RD_ERANGE, "too many cells; raise cell_size");
        }
        R.coff = (int)coff;
        R.off = off;
        R.n = A.n;
        R.stream = A.stream;
        coff += nc;
        off += A.n;
    }
    P.n_pot2 = (int)h->pot2.size();
    for (int p = 0; p < P.n_pot2; ++p) {
        const rd_pot2& q = h->pot2[p];
        RdPot2Dev& d = P.pot2[p];
        d.kind = q.type; d.subtype = q.subtype; d.k = q.force_const;
        d.length = q.length; d.depth = q.depth; d.cutoff = (float)g.max_cutoff;
        d.wc = q.type == RD_POT2_PIECEWISE_WELL ? (float)(q.depth / (0.25 * q.length * q.length)) : 0.0f;
    }
    for (int t = 0; t < nt; ++t) {
        P.a[t] = (float)(g.dt * h->D[t] / (g.boltzmann * g.temperature));
        P.b[t] = (float)std::sqrt(2.0 * h->D[t] * g.dt);
        P.prad[t] = (float)h->rad[t];
    }
    P.seed = g.seed;
    P.n_pot1 = (int)h->pot1.size();
    std::vector<RdPot1Dev> pot1dev(h->pot1.size());
    for (size_t k = 0; k < h->pot1.size(); ++k) {
        const rd_pot1& q = h->pot1[k];
        RdPot1Dev& d = pot1dev[k];
        memset(&d, 0, sizeof(d));
        d.kind = q.type; d.subtype = q.subtype; d.k = q.force_const;
        double nn = std::sqrt((double)q.normal[0] * q.normal[0] + (double)q.normal[1] * q.normal[1] +
                              (double)q.normal[2] * q.normal[2]);
        for (int a = 0; a < 3; ++a) {
            d.o[a] = q.origin[a];
            d.ext[a] = q.extension[a];
            d.n[a] = nn > 0 ? (float)(q.normal[a] / nn) : (a == 2 ? 1.0f : 0.0f);
        }
        d.hh = 0.5f * q.height;
        d.R = q.radius;
        d.mask = q.type_mask;
    }
    // waves: 64 consecutive slots of one replica each
    std::vector<RdWave> wave_tab;
    for (size_t r = 0; r < rep_dev.size(); ++r) {
        const RdRep& R = rep_dev[r];
        for (int s = 0; s < R.n; s += 64) {
            RdWave w;
            w.rep = (int)r; w.s0 = R.off + s; w.s1 = R.off + std::min(R.n, s + 64); w.pad = 0;
            wave_tab.push_back(w);
        }
    }
    h->P = P;
    h->max_range = maxr;
    h->max_rxn = maxrx;
    h->cell = cell;
    h->rep_dev.swap(rep_dev);
    h->ncell = (int)std::max(coff, 1L);
    h->pot1dev.swap(pot1dev);
    h->wave_tab.swap(wave_tab);
    h->nwaves = (int)h->wave_tab.size();
    return RD_OK;
}

static int validate_particles(rd_system* h, int n, const float* xyz, const int* type, const double lo[3],
                              const double hi[3], int base)
{
    for (int i = 0; i < n; ++i) {
        if (type[i] < 0 || type[i] >= h->nt) return set_err(h, RD_EINVAL, "particle type out of range");
        for (int d = 0; d < 3; ++d) {
            float x = xyz[3 * i + d];
            if (!(x >= (float)lo[d] && x <= (float)hi[d]))
                return set_err(h, RD_ERANGE, "particle is out of the box: " + std::to_string(base + i));
        }
    }
    return RD_OK;
}

extern "C" int rd_set_particles(rd_system* h, int n, const float* xyz, const int* type)
{
    if (!h) return RD_EINVAL;
    if (n < 0 || n > RD_MAX_PARTICLES || (n > 0 && (!xyz || !type))) return set_err(h, RD_EINVAL, "bad particle arrays");
    if (h->nt == 0) return set_err(h, RD_ESTATE, "set particle types first");
    rd_system::Rep A;
    for (int d = 0; d < 3; ++d) {
        A.lo[d] = h->prm.lattice_bounds[2 * d];
        A.hi[d] = h->prm.lattice_bounds[2 * d + 1];
    }
    A.stream = h->prm.stream;
    A.n = n;
    int rc = validate_particles(h, n, xyz, type, A.lo, A.hi, 0);
    if (rc) return rc;
    h->reps.assign(1, A);
    h->n = n;
    h->hxyz.assign(xyz, xyz + 3 * (size_t)n);
    h->htype.assign(type, type + n);
    h->tcount.assign((size_t)h->nt, 0);
    for (int i = 0; i < n; ++i) h->tcount[type[i]]++;
    h->snap_valid = false;
    h->particles_dirty = true;
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_set_replicas(rd_system* h, int nrep, const double* bounds, const uint32_t* streams,
                               const int* counts, const float* xyz, const int* type)
{
    if (!h || nrep < 1 || !bounds || !counts) return RD_EINVAL;
    if (h->nt == 0) return set_err(h, RD_ESTATE, "set particle types first");
    long total = 0;
    for (int r = 0; r < nrep; ++r) {
        if (counts[r] < 0) return set_err(h, RD_EINVAL, "negative replica size");
        total += counts[r];
    }
    if (total > RD_MAX_PARTICLES) return set_err(h, RD_EINVAL, "too many particles");
    if (total > 0 && (!xyz || !type)) return set_err(h, RD_EINVAL, "bad particle arrays");
    std::vector<rd_system::Rep> reps(nrep);
    long base = 0;
    for (int r = 0; r < nrep; ++r) {
        rd_system::Rep& A = reps[r];
        for (int d = 0; d < 3; ++d) {
            A.lo[d] = bounds[6 * r + 2 * d];
            A.hi[d] = bounds[6 * r + 2 * d + 1];
            if (!(A.hi[d] > A.lo[d])) return set_err(h, RD_EINVAL, "bad replica box");
        }
        A.stream = streams ? streams[r] : (uint32_t)r;
        A.n = counts[r];
        int rc = validate_particles(h, A.n, xyz + 3 * base, type + base, A.lo, A.hi, (int)base);
        if (rc) return rc;
        base += A.n;
    }
    h->reps = reps;
    h->n = (int)total;
    h->hxyz.assign(xyz, xyz + 3 * (size_t)total);
    h->htype.assign(type, type + total);
    h->tcount.assign((size_t)nrep * h->nt, 0);
    base = 0;
    for (int r = 0; r < nrep; ++r)
        for (int i = 0; i < counts[r]; ++i, ++base) h->tcount[(size_t)r * h->nt + type[base]]++;
    h->snap_valid = false;
    h->particles_dirty = true;
    h->dirty = true;
    return RD_OK;
}

// per-particle arrays for at least `need` particles; with keep_ps the sorted
// state (h->n slots) survives the move
static int grow_particles(rd_system* h, int need, bool keep_ps)
{
    drop_graph(h);
    int cap = std::max(need, h->cap + h->cap / 2);
    float4* nps = nullptr;
    HIPCK(h, hipMalloc(&nps, sizeof(float4) * cap));
    if (keep_ps && h->d_ps && h->n > 0) {
        HIPCK(h, hipMemcpyAsync(nps, h->d_ps, sizeof(float4) * (size_t)h->n, hipMemcpyDeviceToDevice, h->stream));
        HIPCK(h, hipStreamSynchronize(h->stream));
    }
    {
        void* ps[] = { h->d_ps, h->d_pn, h->d_kr, h->d_xyz_io, h->d_type_io, h->d_rep_of };
        for (void* p : ps)
            if (p) (void)hipFree(p);
        h->d_pn = nullptr; h->d_kr = nullptr; h->d_xyz_io = nullptr; h->d_type_io = nullptr;
        h->d_rep_of = nullptr;
        h->d_ps = nps;
        HIPCK(h, hipMalloc(&h->d_pn, sizeof(float4) * cap));
        HIPCK(h, hipMalloc(&h->d_kr, sizeof(uint2) * cap));
        HIPCK(h, hipMalloc(&h->d_xyz_io, sizeof(float) * 3 * cap));
        HIPCK(h, hipMalloc(&h->d_type_io, sizeof(int) * cap));
        HIPCK(h, hipMalloc(&h->d_rep_of, sizeof(int) * cap));
        h->cap = cap;
    }
    return RD_OK;
}

static int ensure_capacity(rd_system* h)
{
    int n = std::max(h->n, 1);
    if (n > h->cap) {
        int rc = grow_particles(h, n, false);
        if (rc) return rc;
    }
    if (h->ncell != h->ncell_alloc) {
        drop_graph(h);
        if (h->d_count) (void)hipFree(h->d_count);
        if (h->d_start) (void)hipFree(h->d_start);
        if (h->d_bsum) (void)hipFree(h->d_bsum);
        size_t nc = (size_t)h->ncell;
        HIPCK(h, hipMalloc(&h->d_count, sizeof(uint32_t) * (nc + 16)));
        HIPCK(h, hipMalloc(&h->d_start, sizeof(uint32_t) * (nc + 16)));
        h->nscan = nblk((long)nc, RD_SCAN_TILE);
        HIPCK(h, hipMalloc(&h->d_bsum, sizeof(uint32_t) * (h->nscan + 16)));
        HIPCK(h, hipMemsetAsync(h->d_count, 0, sizeof(uint32_t) * (nc + 16), h->stream));
        HIPCK(h, hipMemsetAsync(h->d_bsum, 0, sizeof(uint32_t) * (h->nscan + 16), h->stream));
        h->ncell_alloc = h->ncell;
    }
    if ((int)h->rep_dev.size() > h->nrep_alloc) {
        drop_graph(h);
        if (h->d_reps) (void)hipFree(h->d_reps);
        HIPCK(h, hipMalloc(&h->d_reps, sizeof(RdRep) * h->rep_dev.size()));
        h->nrep_alloc = (int)h->rep_dev.size();
    }
    if (std::max(h->nwaves, 1) > h->nwave_alloc) {
        drop_graph(h);
        if (h->d_waves) (void)hipFree(h->d_waves);
        HIPCK(h, hipMalloc(&h->d_waves, sizeof(RdWave) * std::max(h->nwaves, 1)));
        h->nwave_alloc = std::max(h->nwaves, 1);
    }
    if (!h->d_P) HIPCK(h, hipMalloc(&h->d_P, sizeof(RdDev)));
    if (!h->d_st) {
        HIPCK(h, hipMalloc(&h->d_st, sizeof(RdStepState)));
        RdStepState s0;
        memset(&s0, 0, sizeof(s0));
        s0.step = h->step;
        HIPCK(h, hipMemcpyAsync(h->d_st, &s0, sizeof(s0), hipMemcpyHostToDevice, h->stream));
    }
    if (!h->d_etot) HIPCK(h, hipMalloc(&h->d_etot, sizeof(double)));
    return RD_OK;
}

// scan of count -> start (the per-tile sums were accumulated with the
// histogram), then scatter pn -> ps, clearing the tile sums for the next step
static void launch_sort(rd_system* h, int advance)
{
    int nc = h->ncell;
    hipLaunchKernelGGL(k_scan_apply<RD_SCAN_APPLY_THREADS>, dim3(h->nscan), dim3(RD_SCAN_APPLY_THREADS), 0, h->stream,
                       h->d_count, nc, h->d_bsum, h->d_start);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(std::max(h->n, 1), RD_BLOCK * RD_SCATTER_PPL)), dim3(RD_BLOCK), 0,
                       h->stream, h->d_pn,
                       h->d_kr, h->d_start, h->d_ps, h->n, h->d_st, advance,
                       (const uint32_t*)h->d_gid, h->d_slot, h->d_bsum, h->nscan);
}

static int wave_blocks(const rd_system* h) { return std::max(1, nblk(h->nwaves, RD_BLOCK / 64)); }

// ev0/ev1 (measurement only): events the dispatch itself timestamps at the
// kernel's start and end (hipExtLaunchKernelGGL), so their interval is the
// kernel's own duration
template <int B, int PPT>
static void launch_force(rd_system* h, hipEvent_t ev0, hipEvent_t ev1)
{
    const dim3 grid(std::max(1, nblk(h->nwaves, B / 64 * PPT)));
    if constexpr (PPT == 1) {
        if (ev0)
            hipExtLaunchKernelGGL((k_force_update<B>), grid, dim3(B), 0, h->stream, ev0, ev1, 0,
                                  (const float4*)h->d_ps, h->d_pn, h->d_kr, h->d_count, h->d_bsum,
                                  (const uint32_t*)h->d_start, (const RdDev*)h->d_P, (const RdRep*)h->d_reps,
                                  (const RdWave*)h->d_waves, h->nwaves, (const RdPot1Dev*)h->d_pot1, h->d_st,
                                  (const float*)h->d_noise, h->n);
        else
            hipLaunchKernelGGL((k_force_update<B>), grid, dim3(B), 0, h->stream, h->d_ps, h->d_pn, h->d_kr,
                               h->d_count, h->d_bsum, h->d_start, h->d_P, h->d_reps, h->d_waves, h->nwaves, h->d_pot1,
                               h->d_st, h->d_noise, h->n);
    } else {
        if (ev0)
            hipExtLaunchKernelGGL((k_force_update_ppt<B, PPT>), grid, dim3(B), 0, h->stream, ev0, ev1, 0,
                                  (const float4*)h->d_ps, h->d_pn, h->d_kr, h->d_count, h->d_bsum,
                                  (const uint32_t*)h->d_start, (const RdDev*)h->d_P, (const RdRep*)h->d_reps,
                                  (const RdWave*)h->d_waves, h->nwaves, (const RdPot1Dev*)h->d_pot1, h->d_st,
                                  (const float*)h->d_noise, h->n);
        else
            hipLaunchKernelGGL((k_force_update_ppt<B, PPT>), grid, dim3(B), 0, h->stream, h->d_ps, h->d_pn, h->d_kr,
                               h->d_count, h->d_bsum, h->d_start, h->d_P, h->d_reps, h->d_waves, h->nwaves,
                               h->d_pot1, h->d_st, h->d_noise, h->n);
    }
}

// resident workgroups of the work-queue form, at most one per work-item group
template <int B>
static int queue_grid(rd_system* h)
{
    if (h->qgrid_block != B) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_force_update_queue<B>, B, 0) !=
                hipSuccess || per_cu < 1)
            per_cu = 1;
        h->qgrid = per_cu * h->ncu;
        h->qgrid_block = B;
    }
    return std::max(1, std::min(h->qgrid, nblk(h->nwaves, B / 64)));
}

template <int B>
static void launch_force_queue(rd_system* h, hipEvent_t ev0, hipEvent_t ev1)
{
    const dim3 grid(queue_grid<B>(h));
    if (ev0)
        hipExtLaunchKernelGGL((k_force_update_queue<B>), grid, dim3(B), 0, h->stream, ev0, ev1, 0,
                              (const float4*)h->d_ps, h->d_pn, h->d_kr, h->d_count, h->d_bsum,
                              (const uint32_t*)h->d_start, (const RdDev*)h->d_P, (const RdRep*)h->d_reps,
                              (const RdWave*)h->d_waves, h->nwaves, (const RdPot1Dev*)h->d_pot1, h->d_st,
                              (const float*)h->d_noise, h->n);
    else
        hipLaunchKernelGGL((k_force_update_queue<B>), grid, dim3(B), 0, h->stream, h->d_ps, h->d_pn, h->d_kr,
                           h->d_count, h->d_bsum, h->d_start, h->d_P, h->d_reps, h->d_waves, h->nwaves, h->d_pot1,
                           h->d_st, h->d_noise, h->n);
}

template <int B>
static void launch_force_ppt(rd_system* h, hipEvent_t ev0, hipEvent_t ev1)
{
    if (h->kern == RD_KERNEL_QUEUE) {
        launch_force_queue<B>(h, ev0, ev1);
        return;
    }
    switch (h->ppt) {
    case 2: launch_force<B, 2>(h, ev0, ev1); break;
    case 4: launch_force<B, 4>(h, ev0, ev1); break;
    default: launch_force<B, 1>(h, ev0, ev1); break;
    }
}

void rd_host_launch_force(rd_system* h, hipEvent_t ev0, hipEvent_t ev1)
{
    if (h->n <= 0) return;
    switch (h->block) {
    case 64: launch_force_ppt<64>(h, ev0, ev1); break;
    case 128: launch_force_ppt<128>(h, ev0, ev1); break;
    case 512: launch_force_ppt<512>(h, ev0, ev1); break;
    default: launch_force_ppt<256>(h, ev0, ev1); break;
    }
}

static void launch_step(rd_system* h)
{
    if (h->ext_active()) rd_host_launch_ext(h);
    rd_host_launch_force(h);
    launch_sort(h, 1);
}

// bonds as per-particle lists (both directions) by global index, the bond
// potential table, and the slot map the force kernel reads partners through
static int upload_bonds(rd_system* h)
{
    RdDev& P = h->P;
    free_bonds(h);
    P.bstart = nullptr; P.bonds = nullptr; P.slot_of = nullptr; P.gid_n = nullptr;
    P.n_bpot = (int)h->bpots.size();
    for (int p = 0; p < P.n_bpot; ++p) {
        const rd_pot2& q = h->bpots[p];
        RdPot2Dev& d = P.bpot[p];
        memset(&d, 0, sizeof(d));
        d.kind = q.type; d.subtype = q.subtype; d.k = q.force_const;
        d.length = q.length; d.depth = q.depth; d.cutoff = INFINITY;
        d.wc = q.type == RD_POT2_PIECEWISE_WELL ? (float)(q.depth / (0.25 * q.length * q.length)) : 0.0f;
        P.bpot_r0[p] = (float)h->bpot_r0[p];
    }
    const size_t nb = h->bond_pot.size();
    if (nb == 0 || h->n <= 0) return RD_OK;
    // replica of a global index: replicas hold consecutive index ranges
    auto rep_of = [&](int i) {
        size_t r = 0;
        while (r + 1 < h->rep_dev.size() && h->rep_dev[r + 1].off <= i) ++r;
        return r;
    };
    std::vector<int> start((size_t)h->n + 1, 0);
    for (size_t b = 0; b < nb; ++b) {
        const int a = h->bond_pairs[2 * b], c = h->bond_pairs[2 * b + 1];
        if (a >= h->n || c >= h->n) return set_err(h, RD_EINVAL, "a bond names a particle index beyond the particle count");
        if (h->rep_dev.size() > 1 && rep_of(a) != rep_of(c)) return set_err(h, RD_EINVAL, "a bond joins two replicas");
        start[a + 1]++; start[c + 1]++;
    }
    for (int i = 0; i < h->n; ++i) start[i + 1] += start[i];
    std::vector<int> fill(start.begin(), start.end() - 1);
    std::vector<int2> ent(2 * nb);
    for (size_t b = 0; b < nb; ++b) {
        const int a = h->bond_pairs[2 * b], c = h->bond_pairs[2 * b + 1], p = h->bond_pot[b];
        ent[fill[a]++] = make_int2(c, p);
        ent[fill[c]++] = make_int2(a, p);
    }
    HIPCK(h, hipMalloc(&h->d_bstart, sizeof(int) * start.size()));
    HIPCK(h, hipMalloc(&h->d_bonds, sizeof(int2) * ent.size()));
    HIPCK(h, hipMalloc(&h->d_slot, sizeof(uint32_t) * (size_t)h->n));
    HIPCK(h, hipMalloc(&h->d_gid, sizeof(uint32_t) * (size_t)h->n));
    HIPCK(h, hipMemcpyAsync(h->d_bstart, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipMemcpyAsync(h->d_bonds, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    P.bstart = h->d_bstart; P.bonds = h->d_bonds; P.slot_of = h->d_slot; P.gid_n = h->d_gid;
    return RD_OK;
}

// upload tables and particles, build the cell-sorted arrays
static int upload(rd_system* h)
{
    int rc = build_tables(h);
    if (rc) return rc;
    rc = ensure_capacity(h);
    if (rc) return rc;
    rc = rd_host_ext_capacity(h);
    if (rc) return rc;
    rc = upload_bonds(h);
    if (rc) return rc;
    h->P.gext = h->ext_active() ? h->d_gext : nullptr;
    h->P.ext_mask = h->comp_mask;
    h->kern = h->lc.kernel == RD_KERNEL_QUEUE ? RD_KERNEL_QUEUE : RD_KERNEL_GATHER;
    h->block = (h->lc.block == 64 || h->lc.block == 128 || h->lc.block == 512) ? h->lc.block : 256;
    h->ppt = (h->lc.ppt == 2 || h->lc.ppt == 4) ? h->lc.ppt : 1;
    HIPCK(h, hipMemcpyAsync(h->d_P, &h->P, sizeof(RdDev), hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipMemcpyAsync(h->d_reps, h->rep_dev.data(), sizeof(RdRep) * h->rep_dev.size(), hipMemcpyHostToDevice,
                            h->stream));
    if (h->nwaves > 0)
        HIPCK(h, hipMemcpyAsync(h->d_waves, h->wave_tab.data(), sizeof(RdWave) * h->nwaves, hipMemcpyHostToDevice,
                                h->stream));
    if (h->noise_n >= 0 && h->noise_n != h->n) {
        // the noise array was laid out for another particle count: back to the generator
        RdStepState st;
        HIPCK(h, hipMemcpyAsync(&st, h->d_st, sizeof(st), hipMemcpyDeviceToHost, h->stream));
        HIPCK(h, hipStreamSynchronize(h->stream));
        st.noise_nsteps = 0;
        HIPCK(h, hipMemcpyAsync(h->d_st, &st, sizeof(st), hipMemcpyHostToDevice, h->stream));
        h->noise_n = -1;
    }
    if (h->d_pot1) { (void)hipFree(h->d_pot1); h->d_pot1 = nullptr; }
    if (!h->pot1dev.empty()) {
        HIPCK(h, hipMalloc(&h->d_pot1, sizeof(RdPot1Dev) * h->pot1dev.size()));
        HIPCK(h, hipMemcpyAsync(h->d_pot1, h->pot1dev.data(), sizeof(RdPot1Dev) * h->pot1dev.size(),
                                hipMemcpyHostToDevice, h->stream));
    }
    if (h->n > 0) {
        // replica of each particle index: the wave table covers the same ranges
        hipLaunchKernelGGL(k_rep_of, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_waves, h->nwaves,
                           h->d_rep_of);
        if (!h->io_on_device) {
            HIPCK(h, hipMemcpyAsync(h->d_xyz_io, h->hxyz.data(), sizeof(float) * 3 * h->n, hipMemcpyHostToDevice,
                                    h->stream));
            HIPCK(h, hipMemcpyAsync(h->d_type_io, h->htype.data(), sizeof(int) * h->n, hipMemcpyHostToDevice,
                                    h->stream));
        }
        hipLaunchKernelGGL(k_pack, dim3(nblk(h->n, RD_BLOCK)), dim3(RD_BLOCK), 0, h->stream, h->d_xyz_io, h->d_type_io,
                           h->d_rep_of, h->n, h->d_reps, h->d_pn, h->d_kr, h->d_count, h->d_bsum, h->d_gid);
    }
    launch_sort(h, 0);
    HIPCK(h, hipGetLastError());
    HIPCK(h, hipStreamSynchronize(h->stream));
    h->dirty = false;
    h->particles_dirty = false;
    drop_graph(h);
    return RD_OK;
}

static int scratch_setup(rd_system* h);

extern "C" int rd_initialize(rd_system* h)
{
    if (!h) return RD_EINVAL;
    if (h->nt == 0) return set_err(h, RD_ESTATE, "no particle types");
    if (h->reps.empty()) return set_err(h, RD_ESTATE, "no particles set");
    HIPCK(h, hipSetDevice(h->dev));
    int rc = scratch_setup(h);
    if (rc) return rc;
    rc = upload(h);
    if (rc) return rc;
    h->initialized = true;
    return RD_OK;
}

static int refresh(rd_system* h)
{
    if (!h->initialized) return set_err(h, RD_ESTATE, "rd_initialize has not been called");
    HIPCK(h, hipSetDevice(h->dev));
    if (h->dirty || h->particles_dirty) {
        int rc = rd_copy_pos_from_device(h, nullptr, nullptr);   // keep moved positions unless replaced
        (void)rc;
        return upload(h);
    }
    return RD_OK;
}

// One hipGraph per step count: K steps captured once, launched as one graph.
static int step_graph(rd_system* h, int k, hipGraphExec_t* exec)
{
    auto it = h->graphs.find(k);
    if (it != h->graphs.end()) { *exec = it->second.exec; return RD_OK; }
    rd_system::Graph gr;
    HIPCK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < k; ++i) launch_step(h);
    hipError_t e = hipStreamEndCapture(h->stream, &gr.graph);
    if (e != hipSuccess) return set_err(h, RD_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(gr.graph);
        return set_err(h, RD_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    h->graphs[k] = gr;
    *exec = gr.exec;
    return RD_OK;
}

// launch nsteps steps as graphs of up to RD_GRAPH_STEPS steps
static int launch_steps(rd_system* h, int nsteps)
{
    const int chunk = 256;
    while (nsteps > 0) {
        int k = nsteps >= chunk ? chunk : (nsteps >= 16 ? 16 : 1);
        hipGraphExec_t ex;
        int rc = step_graph(h, k, &ex);
        if (rc) return rc;
        HIPCK(h, hipGraphLaunch(ex, h->stream));
        nsteps -= k;
        h->step += (uint64_t)k;
    }
    return RD_OK;
}

extern "C" int rd_simulate(rd_system* h, int nsteps)
{
    if (!h || nsteps < 0) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    if (nsteps == 0) return RD_OK;
    rc = launch_steps(h, nsteps);
    if (rc) return rc;
    RdStepState st;
    HIPCK(h, hipMemcpyAsync(&st, h->d_st, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    return step_error(h, st.err);
}

// pieces of the step for the measurement entry points (rd_bench.hip)
int rd_host_refresh(rd_system* h) { return refresh(h); }
void rd_host_launch_sort(rd_system* h, int advance) { launch_sort(h, advance); }
int rd_host_step_graph(rd_system* h, int k, hipGraphExec_t* exec) { return step_graph(h, k, exec); }

// reset the global step counter (noise counter), e.g. after autotuning
int rd_host_set_step(rd_system* h, uint64_t step)
{
    RdStepState st;
    HIPCK(h, hipMemcpyAsync(&st, h->d_st, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    st.step = step;
    st.err &= ~2u;
    HIPCK(h, hipMemcpyAsync(h->d_st, &st, sizeof(st), hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    h->step = step;
    return RD_OK;
}

// replace positions and types, keeping the replica layout (autotuner)
int rd_host_set_state(rd_system* h, const std::vector<float>& xyz, const std::vector<int>& type)
{
    if (xyz.size() != 3 * (size_t)h->n || type.size() != (size_t)h->n) return set_err(h, RD_EINVAL, "state size");
    h->hxyz = xyz;
    h->htype = type;
    h->particles_dirty = true;
    return RD_OK;
}

extern "C" int rd_set_launch_config(rd_system* h, const rd_launch_config* cfg)
{
    if (!h || !cfg) return RD_EINVAL;
    if (cfg->kernel < RD_KERNEL_AUTO || cfg->kernel > RD_KERNEL_QUEUE) return set_err(h, RD_EINVAL, "bad kernel");
    if (cfg->block != 0 && cfg->block != 64 && cfg->block != 128 && cfg->block != 256 && cfg->block != 512)
        return set_err(h, RD_EINVAL, "workgroup size must be 64, 128, 256 or 512");
    if (cfg->ppt != 0 && cfg->ppt != 1 && cfg->ppt != 2 && cfg->ppt != 4)
        return set_err(h, RD_EINVAL, "particles per lane must be 1, 2 or 4");
    h->lc = *cfg;
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_get_launch_config(rd_system* h, rd_launch_config* cfg)
{
    if (!h || !cfg) return RD_EINVAL;
    *cfg = h->lc;
    cfg->used_kernel = h->kern;
    cfg->used_block = h->block;
    cfg->used_ppt = h->ppt;
    return RD_OK;
}

extern "C" int rd_copy_pos_from_device(rd_system* h, float* xyz, int* type)
{
    if (!h) return RD_EINVAL;
    if (!h->initialized) return set_err(h, RD_ESTATE, "not initialized");
    if (h->particles_dirty) {
        if (xyz) memcpy(xyz, h->hxyz.data(), sizeof(float) * 3 * h->n);
        if (type) memcpy(type, h->htype.data(), sizeof(int) * h->n);
        return RD_OK;
    }
    HIPCK(h, hipSetDevice(h->dev));
    if (h->n > 0) {
        hipLaunchKernelGGL(k_unsort, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps, h->d_reps,
                           h->d_waves, h->nwaves, h->d_xyz_io, h->d_type_io);
        HIPCK(h, hipMemcpyAsync(h->hxyz.data(), h->d_xyz_io, sizeof(float) * 3 * h->n, hipMemcpyDeviceToHost, h->stream));
        HIPCK(h, hipMemcpyAsync(h->htype.data(), h->d_type_io, sizeof(int) * h->n, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCK(h, hipStreamSynchronize(h->stream));
    if (xyz) memcpy(xyz, h->hxyz.data(), sizeof(float) * 3 * h->n);
    if (type) memcpy(type, h->htype.data(), sizeof(int) * h->n);
    return RD_OK;
}

extern "C" int rd_cell_ids(rd_system* h, int* field)
{
    if (!h || !field) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    if (h->n == 0) return RD_OK;
    hipLaunchKernelGGL(k_cell_ids, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps, h->d_reps, h->d_waves,
                       h->nwaves, h->d_type_io);
    HIPCK(h, hipGetLastError());
    HIPCK(h, hipMemcpyAsync(field, h->d_type_io, sizeof(int) * h->n, hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    return RD_OK;
}

extern "C" int rd_gradients(rd_system* h, float* grad, double* energy)
{
    if (!h) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    if (h->n == 0) { if (energy) *energy = 0.0; return RD_OK; }
    HIPCK(h, hipMemsetAsync(h->d_etot, 0, sizeof(double), h->stream));
    if (h->ext_active()) rd_host_launch_ext(h);
    hipLaunchKernelGGL(k_gradients, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps, h->d_start, h->d_P,
                           h->d_reps, h->d_waves, h->nwaves, h->d_pot1, h->d_xyz_io, h->d_etot, h->d_st);
    HIPCK(h, hipGetLastError());
    std::vector<float> tmp(3 * (size_t)h->n);
    double e = 0.0;
    RdStepState st;
    HIPCK(h, hipMemcpyAsync(tmp.data(), h->d_xyz_io, sizeof(float) * 3 * h->n, hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipMemcpyAsync(&e, h->d_etot, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipMemcpyAsync(&st, h->d_st, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    if (grad) memcpy(grad, tmp.data(), sizeof(float) * 3 * h->n);
    if (energy) *energy = e;
    return (st.err & RD_ERR_FIXRANGE) ? step_error(h, RD_ERR_FIXRANGE) : RD_OK;
}

extern "C" int rd_noise_fill(rd_system* h, uint64_t step0, int nsteps, float* out)
{
    if (!h || nsteps < 0 || !out) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    size_t total = (size_t)nsteps * (size_t)h->n;
    if (total == 0) return RD_OK;
    float* d = nullptr;
    HIPCK(h, hipMalloc(&d, sizeof(float) * 3 * total));
    hipLaunchKernelGGL(k_noise_fill, dim3((unsigned)((total + RD_BLOCK - 1) / RD_BLOCK)), dim3(RD_BLOCK), 0, h->stream,
                       h->d_P, h->d_reps, (int)h->rep_dev.size(), (unsigned long long)step0, nsteps, h->n, d);
    hipError_t e1 = hipMemcpyAsync(out, d, sizeof(float) * 3 * total, hipMemcpyDeviceToHost, h->stream);
    hipError_t e2 = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e1 != hipSuccess || e2 != hipSuccess) return set_err(h, RD_EHIP, "noise fill failed");
    return RD_OK;
}

extern "C" int rd_set_noise_array(rd_system* h, const float* noise, int nsteps)
{
    if (!h || nsteps < 0 || (nsteps > 0 && !noise)) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    RdStepState st;
    HIPCK(h, hipMemcpyAsync(&st, h->d_st, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    size_t total = (size_t)nsteps * 3 * (size_t)h->n;
    if (total > h->noise_cap) {
        drop_graph(h);
        if (h->d_noise) (void)hipFree(h->d_noise);
        h->d_noise = nullptr;
        HIPCK(h, hipMalloc(&h->d_noise, sizeof(float) * total));
        h->noise_cap = total;
    }
    if (total) HIPCK(h, hipMemcpyAsync(h->d_noise, noise, sizeof(float) * total, hipMemcpyHostToDevice, h->stream));
    st.noise_step0 = st.step;
    st.noise_nsteps = nsteps;
    h->noise_n = nsteps > 0 ? h->n : -1;
    st.err &= ~2u;
    HIPCK(h, hipMemcpyAsync(h->d_st, &st, sizeof(st), hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    drop_graph(h);   // the captured kernels hold the noise pointer
    return RD_OK;
}

// ---- RDF (per replica) ----
extern "C" int rd_rdf_configure(rd_system* h, double rmax, int nbins)
{
    if (!h || !(rmax > 0) || nbins < 1) return RD_EINVAL;
    if (h->nt == 0 || h->reps.empty()) return set_err(h, RD_ESTATE, "set particle types and particles first");
    for (const auto& A : h->reps)
        for (int d = 0; d < 3; ++d)
            if (rmax > 0.5 * (A.hi[d] - A.lo[d])) return set_err(h, RD_EINVAL, "RDF range exceeds half a box");
    h->rdf_rmax = rmax;
    h->rdf_bins = nbins;
    h->rdf_nrep = (int)h->reps.size();
    size_t nh = (size_t)h->rdf_nrep * h->nt * h->nt * nbins;
    if (h->d_rdf) (void)hipFree(h->d_rdf);
    if (h->d_rdf_part) (void)hipFree(h->d_rdf_part);
    h->d_rdf = nullptr;
    h->d_rdf_part = nullptr;
    HIPCK(h, hipMalloc(&h->d_rdf, sizeof(unsigned long long) * nh));
    HIPCK(h, hipMemsetAsync(h->d_rdf, 0, sizeof(unsigned long long) * nh, h->stream));
    // copies of the counts the LDS histograms are flushed into (k_rdf_lds),
    // 16 while they stay within 64 MB
    h->rdf_ncopy = 16;
    while (h->rdf_ncopy > 1 && sizeof(unsigned long long) * nh * (size_t)h->rdf_ncopy > ((size_t)64 << 20))
        h->rdf_ncopy >>= 1;
    HIPCK(h, hipMalloc(&h->d_rdf_part, sizeof(unsigned long long) * nh * (size_t)h->rdf_ncopy));
    HIPCK(h, hipMemsetAsync(h->d_rdf_part, 0, sizeof(unsigned long long) * nh * (size_t)h->rdf_ncopy, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    h->rdf_frames = 0;
    h->rdf_pairs_sum.assign((size_t)h->rdf_nrep * h->nt * h->nt, 0.0);
    h->rdf_vol.assign(h->rdf_nrep, 1.0);
    for (int r = 0; r < h->rdf_nrep; ++r)
        for (int d = 0; d < 3; ++d) h->rdf_vol[r] *= h->reps[r].hi[d] - h->reps[r].lo[d];
    return RD_OK;
}

extern "C" int rd_rdf_reset(rd_system* h)
{
    if (!h || !h->d_rdf) return RD_ESTATE;
    return rd_rdf_configure(h, h->rdf_rmax, h->rdf_bins);
}

extern "C" int rd_rdf_calculate(rd_system* h)
{
    if (!h) return RD_EINVAL;
    if (!h->d_rdf) return set_err(h, RD_ESTATE, "rd_rdf_configure first");
    int rc = refresh(h);
    if (rc) return rc;
    if ((int)h->rep_dev.size() != h->rdf_nrep) return set_err(h, RD_ESTATE, "replicas changed: reconfigure the RDF");
    int S = std::max(1, (int)std::ceil(h->rdf_rmax * 1.0001 / h->cell));
    size_t nh = (size_t)h->nt * h->nt * h->rdf_bins;
    // one histogram per workgroup of up to 160 KB
    size_t lds = nh * sizeof(uint32_t);
    const float rmax = (float)h->rdf_rmax;
    const float rcut2 = (float)((double)rmax * (double)rmax * (1.0 + 1e-6));
    if (h->n > 0) {
        if (lds + 64 <= 160 * 1024 && h->d_rdf_part) {          // + the static replica table
            if (lds > 64 * 1024)
                HIPCK(h, hipFuncSetAttribute((const void*)k_rdf_lds, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            const long total = (long)h->rdf_nrep * (long)nh;
            hipLaunchKernelGGL(k_rdf_lds, dim3(wave_blocks(h)), dim3(RD_BLOCK), lds, h->stream, h->d_ps, h->d_start,
                               h->d_P, h->d_reps, h->d_waves, h->nwaves, S, rmax, rcut2, h->rdf_bins, h->d_rdf_part,
                               total, h->rdf_ncopy, h->d_rdf);
            hipLaunchKernelGGL(k_rdf_fold, dim3(std::max(1, nblk(total, RD_BLOCK))), dim3(RD_BLOCK), 0, h->stream,
                               h->d_rdf_part, total, h->rdf_ncopy, h->d_rdf);
        } else {
            hipLaunchKernelGGL(k_rdf_global, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps, h->d_start,
                               h->d_P, h->d_reps, h->d_waves, h->nwaves, S, rmax, rcut2, h->rdf_bins, h->d_rdf);
        }
        HIPCK(h, hipGetLastError());
    }
    // per-frame normalisation constants (particle numbers may change by reactions)
    const int nt = h->nt;
    for (int r = 0; r < h->rdf_nrep; ++r) {
        const long* npt = &h->tcount[(size_t)r * nt];
        for (int a = 0; a < nt; ++a)
            for (int b = 0; b < nt; ++b)
                h->rdf_pairs_sum[((size_t)r * nt + a) * nt + b] += (double)npt[a] * (double)(npt[b] - (a == b ? 1 : 0));
    }
    h->rdf_frames++;
    HIPCK(h, hipStreamSynchronize(h->stream));
    return RD_OK;
}

extern "C" int rd_rdf_copy(rd_system* h, uint64_t* counts, long* frames)
{
    if (!h || !counts) return RD_EINVAL;
    if (!h->d_rdf) return set_err(h, RD_ESTATE, "rd_rdf_configure first");
    size_t nh = (size_t)h->rdf_nrep * h->nt * h->nt * h->rdf_bins;
    HIPCK(h, hipMemcpyAsync(counts, h->d_rdf, sizeof(uint64_t) * nh, hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    if (frames) *frames = h->rdf_frames;
    return RD_OK;
}

// g_ab(r_k) = C_ab(k) / (sum_frames N_a (N_b - [a==b]) / V * V_shell(k)), FP64, per replica.
extern "C" int rd_rdf_normalize(rd_system* h, double* g)
{
    if (!h || !g) return RD_EINVAL;
    if (!h->d_rdf) return set_err(h, RD_ESTATE, "rd_rdf_configure first");
    int nt = h->nt, nb = h->rdf_bins;
    std::vector<uint64_t> c((size_t)h->rdf_nrep * nt * nt * nb);
    int rc = rd_rdf_copy(h, c.data(), nullptr);
    if (rc) return rc;
    const double pi = 3.14159265358979323846;
    double dr = h->rdf_rmax / nb;
    for (int r = 0; r < h->rdf_nrep; ++r)
        for (int a = 0; a < nt; ++a)
            for (int b = 0; b < nt; ++b)
                for (int k = 0; k < nb; ++k) {
                    double r1 = k * dr, r2 = (k + 1) * dr;
                    double vs = 4.0 / 3.0 * pi * (r2 * r2 * r2 - r1 * r1 * r1);
                    double den = h->rdf_pairs_sum[((size_t)r * nt + a) * nt + b] / h->rdf_vol[r] * vs;
                    size_t idx = (((size_t)r * nt + a) * nt + b) * nb + k;
                    g[idx] = den > 0 ? (double)c[idx] / den : 0.0;
                }
    return RD_OK;
}

// ---- pair sets ----
// per-particle partner counts, then their exclusive scan in place: off[i] is
// the first slot of particle i's pairs, off[n] their total
static void launch_pair_counts(rd_system* h, const float* dtable, int S, uint32_t* off, uint32_t* bsum)
{
    const int nsc = nblk(h->n, RD_SCAN_TILE);
    hipLaunchKernelGGL((k_pairs<false, uint2>), dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps,
                       h->d_start, h->d_P, h->d_reps, h->d_waves, h->nwaves, dtable, S, off, (uint2*)nullptr);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nsc), dim3(RD_BLOCK), 0, h->stream, off, h->n, bsum);
    hipLaunchKernelGGL(k_scan_offsets, dim3(nsc), dim3(RD_BLOCK), 0, h->stream, off, h->n, bsum);
}

static int pairs_common(rd_system* h, bool reactions, int64_t* out, long cap, long* count)
{
    if (!h || cap < 0 || (cap > 0 && !out)) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    const float* table = reactions ? h->P.rr2 : h->P.rc2;
    const float* dtable = (const float*)((const char*)h->d_P + (reactions ? offsetof(RdDev, rr2) : offsetof(RdDev, rc2)));
    float m2 = 0.0f;
    for (int a = 0; a < h->nt * h->nt; ++a) m2 = std::max(m2, table[a]);
    if (h->n == 0 || m2 <= 0.0f) { if (count) *count = 0; return RD_OK; }
    int S = std::max(1, (int)std::ceil(std::sqrt((double)m2) * 1.0001 / h->cell));
    // per-particle counts -> offsets (the counting-sort scan over n entries),
    // in buffers the handle keeps between calls
    const int n = h->n;
    const int nsc = nblk(n, RD_SCAN_TILE);
    hipError_t e = hipSuccess;
    if ((size_t)n + 16 > h->poff_cap) {
        for (void* p : { (void*)h->d_poff, (void*)h->d_pbs })
            if (p) (void)hipFree(p);
        h->d_poff = nullptr;
        h->d_pbs = nullptr;
        h->poff_cap = 0;
        e = hipMalloc(&h->d_poff, sizeof(uint32_t) * ((size_t)n + 16));
        if (e == hipSuccess) e = hipMalloc(&h->d_pbs, sizeof(uint32_t) * ((size_t)nblk((long)n + 16, RD_SCAN_TILE) + 16));
        if (e == hipSuccess) h->poff_cap = (size_t)n + 16;
    }
    uint32_t* d_off = h->d_poff;
    uint32_t* d_bs = h->d_pbs;
    uint32_t total = 0;
    if (e == hipSuccess) launch_pair_counts(h, dtable, S, d_off, d_bs);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&total, d_off + n, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && total > h->pairs_cap) {
        if (h->d_pairs) (void)hipFree(h->d_pairs);
        h->d_pairs = nullptr;
        h->pairs_cap = 0;
        e = hipMalloc(&h->d_pairs, sizeof(uint2) * total);
        if (e == hipSuccess) h->pairs_cap = total;
    }
    const long nout = std::min<long>((long)total, cap);
    std::vector<uint2> raw((size_t)nout);
    if (e == hipSuccess && total > 0) {
        hipLaunchKernelGGL((k_pairs<true, uint2>), dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps,
                           h->d_start, h->d_P, h->d_reps, h->d_waves, h->nwaves, dtable, S, d_off, h->d_pairs);
        e = hipGetLastError();
        if (e == hipSuccess && nout > 0)
            e = hipMemcpyAsync(raw.data(), h->d_pairs, sizeof(uint2) * (size_t)nout, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) return set_err(h, RD_EHIP, std::string("pair emission: ") + hipGetErrorString(e));
    for (long k = 0; k < nout; ++k) { out[2 * k] = (int64_t)raw[k].x; out[2 * k + 1] = (int64_t)raw[k].y; }
    if (count) *count = (long)total;
    return RD_OK;
}

// ---- reaction candidates without allocation ----
// rd_reaction_candidates_noalloc runs the kernels of rd_reaction_candidates
// with their scratch memory (per-particle offsets, scan tile sums, the pair
// list) in this file-scope device array instead of buffers allocated per
// call.  Each GPU holds its own copy for the life of the process; its address
// is taken once per device (rd_initialize) and kept in scratch_dev.  Every
// kernel writes the scratch it later reads, so nothing needs clearing between
// calls.  Calls that overlap on one GPU share the array and race.
#define RD_NA_MAX_N (1 << 23)               // particles
#define RD_NA_MAX_PAIRS (1 << 20)           // candidate pairs
#define RD_NA_OFF_WORDS (RD_NA_MAX_N + 64)
#define RD_NA_BSUM_WORDS (RD_NA_MAX_N / RD_SCAN_TILE + 64)
#define RD_NA_WORDS (RD_NA_OFF_WORDS + RD_NA_BSUM_WORDS + 4 * RD_NA_MAX_PAIRS)
static_assert((RD_NA_OFF_WORDS + RD_NA_BSUM_WORDS) % 4 == 0, "pair list is 16-byte aligned");
static __device__ __align__(256) uint32_t scratch[RD_NA_WORDS];

#define RD_NA_MAX_DEVICES 64
static std::mutex scratch_mu;
static std::atomic<void*> scratch_dev[RD_NA_MAX_DEVICES];

// address of this device's scratch array (the device must be current)
static int scratch_setup(rd_system* h)
{
    if (h->dev < 0 || h->dev >= RD_NA_MAX_DEVICES) return set_err(h, RD_EINVAL, "device index beyond 64");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_dev[h->dev].load()) return RD_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_err(h, RD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_dev[h->dev].store(ptr);
    return RD_OK;
}

extern "C" int rd_reaction_candidates_noalloc(rd_system* h, int64_t* out, long cap, long* count)
{
    if (!h || cap < 0 || (cap > 0 && !out)) return RD_EINVAL;
    if (!h->initialized) return set_err(h, RD_ESTATE, "rd_initialize has not been called");
    if (h->dirty || h->particles_dirty)
        return set_err(h, RD_ESTATE, "the model or particles changed since the last upload: call rd_simulate "
                                     "(or rd_reaction_candidates) first");
    void* base = (h->dev >= 0 && h->dev < RD_NA_MAX_DEVICES) ? scratch_dev[h->dev].load() : nullptr;
    if (!base) return set_err(h, RD_ESTATE, "no scratch address for this device");
    float m2 = 0.0f;
    for (int a = 0; a < h->nt * h->nt; ++a) m2 = std::max(m2, h->P.rr2[a]);
    if (h->n == 0 || m2 <= 0.0f) { if (count) *count = 0; return RD_OK; }
    if (h->n > RD_NA_MAX_N)
        return set_err(h, RD_ERANGE, "more than 2^23 particles: the allocation-free scratch is too small; use "
                                     "rd_reaction_candidates");
    HIPCK(h, hipSetDevice(h->dev));
    const int S = std::max(1, (int)std::ceil(std::sqrt((double)m2) * 1.0001 / h->cell));
    const float* dtable = (const float*)((const char*)h->d_P + offsetof(RdDev, rr2));
    uint32_t* off = (uint32_t*)base;
    uint32_t* bsum = off + RD_NA_OFF_WORDS;
    longlong2* pairs = reinterpret_cast<longlong2*>(bsum + RD_NA_BSUM_WORDS);
    launch_pair_counts(h, dtable, S, off, bsum);
    HIPCK(h, hipGetLastError());
    uint32_t total = 0;
    HIPCK(h, hipMemcpyAsync(&total, off + h->n, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    if (total > (uint32_t)RD_NA_MAX_PAIRS)
        return set_err(h, RD_ERANGE, "more than 2^20 candidate pairs: the allocation-free scratch is too small; "
                                     "use rd_reaction_candidates");
    if (total > 0) {
        hipLaunchKernelGGL((k_pairs<true, longlong2>), dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps,
                           h->d_start, h->d_P, h->d_reps, h->d_waves, h->nwaves, dtable, S, off, pairs);
        HIPCK(h, hipGetLastError());
        const long nout = std::min<long>((long)total, cap);
        if (nout > 0)
            HIPCK(h, hipMemcpyAsync(out, pairs, sizeof(longlong2) * (size_t)nout, hipMemcpyDeviceToHost, h->stream));
        HIPCK(h, hipStreamSynchronize(h->stream));
    }
    if (count) *count = (long)total;
    return RD_OK;
}

extern "C" int rd_neighbor_pairs(rd_system* h, int64_t* out, long cap, long* count)
{
    return pairs_common(h, false, out, cap, count);
}

extern "C" int rd_reaction_candidates(rd_system* h, int64_t* out, long cap, long* count)
{
    return pairs_common(h, true, out, cap, count);
}

// ---- reactions: lib/ReaDDyGPU.cpp:303-463 (single system) ----
// cReactions = [count, (readdyId, newType, x, y, z, index) x count]
// (lib/ReaDDyGPU.cpp:303-365): index != -1 removes that particle, newType != -1
// appends a particle of that type at (x, y, z); both is a type change (the
// particle is removed and its successor appended).  Survivors keep their
// order, new particles follow in entry order, per replica.  Applied on the
// device: the sorted state is compacted into index order, new particles are
// placed, and the cell list is rebuilt; no particle data crosses PCIe.
static int apply_reactions(rd_system* h, const float* cr, const int* creation_replica)
{
    if (!h->initialized) return set_err(h, RD_ESTATE, "not initialized");
    int nr = (int)cr[0];
    if (nr <= 0) return RD_OK;
    int rc = refresh(h);
    if (rc) return rc;
    const int n = h->n;
    const int R = (int)h->reps.size();
    const int nt = h->nt;
    // replica of a particle index (replicas hold consecutive index ranges)
    auto rep_of = [&](int i) {
        int lo = 0, hi = R - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (h->rep_dev[mid].off <= i) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    std::vector<int> del;
    struct Born { int rep, type; float x, y, z; };
    std::vector<Born> born;
    for (int r = 0; r < nr; ++r) {
        const float* e = &cr[1 + 6 * r];
        int newType = (int)e[1];
        int index = (int)e[5];
        if (index != -1) {
            if (index < 0 || index >= n) return set_err(h, RD_EINVAL, "reaction deletes a nonexistent particle");
            del.push_back(index);
        }
        if (newType != -1) {
            if (newType < 0 || newType >= h->nt) return set_err(h, RD_EINVAL, "reaction creates an unknown type");
            int rep = 0;
            if (index != -1) rep = rep_of(index);
            else if (R > 1) {
                if (!creation_replica) return set_err(h, RD_EINVAL, "a new particle needs its replica");
                rep = creation_replica[r];
                if (rep < 0 || rep >= R) return set_err(h, RD_EINVAL, "replica index out of range");
            }
            born.push_back({ rep, newType, e[2], e[3], e[4] });
        }
    }
    std::sort(del.begin(), del.end());
    if (std::adjacent_find(del.begin(), del.end()) != del.end())
        return set_err(h, RD_EINVAL, "reaction deletes a nonexistent particle");
    for (const Born& b : born) {
        float x[3] = { b.x, b.y, b.z };
        rc = validate_particles(h, 1, x, &b.type, h->reps[b.rep].lo, h->reps[b.rep].hi, 0);
        if (rc) return rc;
    }
    // per replica: removed, created; new offsets; shift of the survivors
    std::vector<int> ndel(R, 0), ncre(R, 0), shift(R, 0), noff(R + 1, 0);
    for (int i : del) ndel[rep_of(i)]++;
    for (const Born& b : born) ncre[b.rep]++;
    for (int r = 0; r < R; ++r) {
        shift[r] = (r ? shift[r - 1] + ncre[r - 1] : 0);
        noff[r + 1] = noff[r] + h->reps[r].n - ndel[r] + ncre[r];
    }
    const int n_new = noff[R];
    if (n_new > RD_MAX_PARTICLES) return set_err(h, RD_EINVAL, "too many particles");
    if (n_new > h->cap) {
        rc = grow_particles(h, n_new, true);
        if (rc) return rc;
    }
    // destinations of the new particles: after the survivors of their replica
    std::vector<int> fill(R, 0), bdst(born.size()), bty(born.size());
    std::vector<float> bxyz(3 * born.size());
    for (size_t k = 0; k < born.size(); ++k) {
        const Born& b = born[k];
        bdst[k] = noff[b.rep] + h->reps[b.rep].n - ndel[b.rep] + fill[b.rep]++;
        bty[k] = b.type;
        bxyz[3 * k] = b.x; bxyz[3 * k + 1] = b.y; bxyz[3 * k + 2] = b.z;
    }
    int *d_del = nullptr, *d_shift = nullptr, *d_bdst = nullptr, *d_bty = nullptr;
    float* d_bxyz = nullptr;
    uint32_t* d_delcnt = nullptr;
    std::vector<uint32_t> delcnt((size_t)R * nt, 0);
    hipError_t e = hipSuccess;
    auto up = [&](auto** dst, const auto* src, size_t cnt) {
        if (e != hipSuccess || cnt == 0) return;
        e = hipMalloc((void**)dst, sizeof(**dst) * cnt);
        if (e == hipSuccess) e = hipMemcpyAsync(*dst, src, sizeof(**dst) * cnt, hipMemcpyHostToDevice, h->stream);
    };
    up(&d_del, del.data(), del.size());
    up(&d_shift, shift.data(), (size_t)R);
    up(&d_bdst, bdst.data(), bdst.size());
    up(&d_bty, bty.data(), bty.size());
    up(&d_bxyz, bxyz.data(), bxyz.size());
    up(&d_delcnt, delcnt.data(), delcnt.size());
    if (e == hipSuccess && n > 0)
        hipLaunchKernelGGL(k_react_compact, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps, h->d_reps,
                           h->d_waves, h->nwaves, d_del, (int)del.size(), d_shift, nt, d_delcnt, h->d_xyz_io,
                           h->d_type_io);
    if (e == hipSuccess && !born.empty())
        hipLaunchKernelGGL(k_react_place, dim3(nblk((long)born.size(), RD_BLOCK)), dim3(RD_BLOCK), 0, h->stream,
                           d_bxyz, d_bty, d_bdst, (int)born.size(), h->d_xyz_io, h->d_type_io);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && d_delcnt)
        e = hipMemcpyAsync(delcnt.data(), d_delcnt, sizeof(uint32_t) * delcnt.size(), hipMemcpyDeviceToHost, h->stream);
    hipError_t se = hipStreamSynchronize(h->stream);
    for (void* p : { (void*)d_del, (void*)d_shift, (void*)d_bdst, (void*)d_bty, (void*)d_bxyz, (void*)d_delcnt })
        if (p) (void)hipFree(p);
    if (e != hipSuccess || se != hipSuccess)
        return set_err(h, RD_EHIP, std::string("reaction compaction: ") + hipGetErrorString(e != hipSuccess ? e : se));
    // particle counts per (replica, type); the host position/type mirrors are
    // refreshed from the device when next read
    for (size_t k = 0; k < delcnt.size(); ++k) h->tcount[k] -= (long)delcnt[k];
    for (const Born& b : born) h->tcount[(size_t)b.rep * nt + b.type]++;
    h->hxyz.resize(3 * (size_t)n_new);
    h->htype.resize((size_t)n_new);
    auto new_index = [&](int i) {
        auto it = std::lower_bound(del.begin(), del.end(), i);
        return (it != del.end() && *it == i) ? -1 : i - (int)(it - del.begin()) + shift[rep_of(i)];
    };
    // bonds follow their particles; a bond of a removed particle is dropped
    if (!h->bond_pot.empty()) {
        std::vector<int> bp, bo;
        for (size_t b = 0; b < h->bond_pot.size(); ++b) {
            const int a0 = h->bond_pairs[2 * b], c0 = h->bond_pairs[2 * b + 1];
            const int a = a0 < n ? new_index(a0) : -1, c = c0 < n ? new_index(c0) : -1;
            if (a < 0 || c < 0) continue;
            bp.push_back(a); bp.push_back(c); bo.push_back(h->bond_pot[b]);
        }
        h->bond_pairs.swap(bp);
        h->bond_pot.swap(bo);
    }
    for (int r = 0; r < R; ++r) h->reps[r].n = noff[r + 1] - noff[r];
    h->n = n_new;
    h->snap_valid = false;
    h->io_on_device = true;
    rc = upload(h);
    h->io_on_device = false;
    return rc;
}

extern "C" int rd_apply_reactions(rd_system* h, const float* cr)
{
    if (!h || !cr) return RD_EINVAL;
    return apply_reactions(h, cr, nullptr);
}

extern "C" int rd_apply_reactions_replicas(rd_system* h, const float* cr, const int* creation_replica)
{
    if (!h || !cr) return RD_EINVAL;
    return apply_reactions(h, cr, creation_replica);
}

// ---- frame hand-off (the frame loop of lib/ReaDDyGPU.cpp:235-301) ----
extern "C" int rd_simulate_async(rd_system* h, int nsteps)
{
    if (!h || nsteps < 0) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    return nsteps > 0 ? launch_steps(h, nsteps) : RD_OK;
}

extern "C" int rd_frame_copy(rd_system* h)
{
    if (!h) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    const int n = std::max(h->n, 1);
    if (n > h->snap_cap) {
        const int cap = std::max(n, h->snap_cap + h->snap_cap / 2);
        free_frames(h);
        HIPCK(h, hipMalloc(&h->d_snap_xyz, sizeof(float) * 3 * (size_t)cap));
        HIPCK(h, hipMalloc(&h->d_snap_type, sizeof(int) * (size_t)cap));
        HIPCK(h, hipHostMalloc((void**)&h->hp_xyz, sizeof(float) * 3 * (size_t)cap, hipHostMallocDefault));
        HIPCK(h, hipHostMalloc((void**)&h->hp_type, sizeof(int) * (size_t)cap, hipHostMallocDefault));
        HIPCK(h, hipHostMalloc((void**)&h->hp_st, sizeof(RdStepState), hipHostMallocDefault));
        HIPCK(h, hipEventCreateWithFlags(&h->ev_frame, hipEventDisableTiming));
        HIPCK(h, hipEventCreateWithFlags(&h->ev_snap, hipEventDisableTiming));
        HIPCK(h, hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
        h->snap_cap = cap;
    } else if (h->snap_valid) {
        // the previous copy to the host must be done before the snapshot is overwritten
        HIPCK(h, hipStreamWaitEvent(h->stream, h->ev_frame, 0));
    }
    // snapshot on the step stream; the copy to the host runs on its own stream
    // so the steps enqueued next do not wait for it
    if (h->n > 0) {
        hipLaunchKernelGGL(k_unsort, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps, h->d_reps,
                           h->d_waves, h->nwaves, h->d_snap_xyz, h->d_snap_type);
        HIPCK(h, hipGetLastError());
    }
    HIPCK(h, hipMemcpyAsync(h->hp_st, h->d_st, sizeof(RdStepState), hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipEventRecord(h->ev_snap, h->stream));
    HIPCK(h, hipStreamWaitEvent(h->copy_stream, h->ev_snap, 0));
    if (h->n > 0) {
        HIPCK(h, hipMemcpyAsync(h->hp_xyz, h->d_snap_xyz, sizeof(float) * 3 * (size_t)h->n, hipMemcpyDeviceToHost,
                                h->copy_stream));
        HIPCK(h, hipMemcpyAsync(h->hp_type, h->d_snap_type, sizeof(int) * (size_t)h->n, hipMemcpyDeviceToHost,
                                h->copy_stream));
    }
    HIPCK(h, hipEventRecord(h->ev_frame, h->copy_stream));
    h->snap_n = h->n;
    h->snap_step = h->step;
    h->snap_valid = true;
    return RD_OK;
}

extern "C" int rd_frame_wait(rd_system* h, const float** xyz, const int** type, int* n)
{
    if (!h) return RD_EINVAL;
    if (!h->snap_valid) return set_err(h, RD_ESTATE, "no frame copy in flight");
    HIPCK(h, hipEventSynchronize(h->ev_frame));
    const int rc = step_error(h, h->hp_st->err);
    if (rc) return rc;
    if (xyz) *xyz = h->hp_xyz;
    if (type) *type = h->hp_type;
    if (n) *n = h->snap_n;
    return RD_OK;
}

extern "C" int rd_frame_rewind(rd_system* h)
{
    if (!h) return RD_EINVAL;
    if (!h->snap_valid || h->snap_n != h->n) return set_err(h, RD_ESTATE, "no frame copy of the current particle set");
    HIPCK(h, hipStreamSynchronize(h->stream));
    if (h->n > 0) {
        HIPCK(h, hipMemcpyAsync(h->d_xyz_io, h->d_snap_xyz, sizeof(float) * 3 * (size_t)h->n,
                                hipMemcpyDeviceToDevice, h->stream));
        HIPCK(h, hipMemcpyAsync(h->d_type_io, h->d_snap_type, sizeof(int) * (size_t)h->n, hipMemcpyDeviceToDevice,
                                h->stream));
    }
    int rc = rd_host_set_step(h, h->snap_step);
    if (rc) return rc;
    // same particle set and tables: pack the snapshot and sort it again (the
    // captured step graphs stay valid)
    if (h->n > 0)
        hipLaunchKernelGGL(k_pack, dim3(nblk(h->n, RD_BLOCK)), dim3(RD_BLOCK), 0, h->stream, h->d_xyz_io,
                           h->d_type_io, h->d_rep_of, h->n, h->d_reps, h->d_pn, h->d_kr, h->d_count,
                           h->d_bsum, h->d_gid);
    launch_sort(h, 0);
    HIPCK(h, hipGetLastError());
    HIPCK(h, hipStreamSynchronize(h->stream));
    return RD_OK;
}

extern "C" int rd_get_info(rd_system* h, rd_info* info)
{
    if (!h || !info) return RD_EINVAL;
    memset(info, 0, sizeof(*info));
    info->n = h->n;
    info->n_types = h->nt;
    if (!h->rep_dev.empty())
        for (int d = 0; d < 3; ++d) info->lattice[d] = h->rep_dev[0].L[d];
    info->cell_size = h->cell;
    info->stencil = h->P.S;
    info->step = h->step;
    info->rdf_frames = h->rdf_frames;
    info->n_replicas = (int)h->reps.size();
    info->n_cells = h->ncell;
    return RD_OK;
}

// ---- checkpoint / resume ----
// No counterpart in the reference, whose de-facto restart is the XML
// trajectory (outputHandler.cpp:135-148, read back by CoordReader.cpp:27-48):
// that format drops the step counter, so the noise after a restart would not
// continue the run.  The snapshot holds the particle state in index order, the
// replica layout, the bonds, the seed and the step (noise) counter.
static const char RD_CKPT_MAGIC[8] = { 'R', 'D', 'C', 'K', 'P', 'T', '0', '1' };

extern "C" int rd_set_step(rd_system* h, uint64_t step)
{
    if (!h) return RD_EINVAL;
    if (!h->d_st) {                  // not initialized yet: used when the state is created
        h->step = step;
        return RD_OK;
    }
    HIPCK(h, hipSetDevice(h->dev));
    HIPCK(h, hipStreamSynchronize(h->stream));
    return rd_host_set_step(h, step);
}

extern "C" int rd_checkpoint_write(rd_system* h, const char* path)
{
    if (!h || !path) return RD_EINVAL;
    int rc = rd_copy_pos_from_device(h, nullptr, nullptr);
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return set_err(h, RD_EINVAL, std::string("cannot write ") + path);
    const uint32_t version = 1, nt = (uint32_t)h->nt, nrep = (uint32_t)h->reps.size(), n = (uint32_t)h->n;
    const uint32_t nb = (uint32_t)h->bond_pot.size();
    const uint64_t seed = h->prm.seed, step = h->step;
    bool ok = fwrite(RD_CKPT_MAGIC, 1, 8, f) == 8;
    ok = ok && fwrite(&version, 4, 1, f) == 1 && fwrite(&seed, 8, 1, f) == 1 && fwrite(&step, 8, 1, f) == 1;
    ok = ok && fwrite(&nt, 4, 1, f) == 1 && fwrite(&nrep, 4, 1, f) == 1 && fwrite(&n, 4, 1, f) == 1;
    for (const rd_system::Rep& A : h->reps) {
        const uint32_t cnt = (uint32_t)A.n;
        ok = ok && fwrite(A.lo, 8, 3, f) == 3 && fwrite(A.hi, 8, 3, f) == 3;
        ok = ok && fwrite(&A.stream, 4, 1, f) == 1 && fwrite(&cnt, 4, 1, f) == 1;
    }
    ok = ok && fwrite(h->hxyz.data(), 4, 3 * (size_t)n, f) == 3 * (size_t)n;
    ok = ok && fwrite(h->htype.data(), 4, (size_t)n, f) == (size_t)n;
    ok = ok && fwrite(&nb, 4, 1, f) == 1;
    ok = ok && fwrite(h->bond_pairs.data(), 4, 2 * (size_t)nb, f) == 2 * (size_t)nb;
    ok = ok && fwrite(h->bond_pot.data(), 4, (size_t)nb, f) == (size_t)nb;
    if (fclose(f) != 0) ok = false;
    return ok ? RD_OK : set_err(h, RD_EINVAL, std::string("write failed: ") + path);
}

extern "C" int rd_checkpoint_read(rd_system* h, const char* path)
{
    if (!h || !path) return RD_EINVAL;
    FILE* f = fopen(path, "rb");
    if (!f) return set_err(h, RD_EINVAL, std::string("cannot read ") + path);
    char magic[8];
    uint32_t version = 0, nt = 0, nrep = 0, n = 0, nb = 0;
    uint64_t seed = 0, step = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, RD_CKPT_MAGIC, 8) == 0;
    ok = ok && fread(&version, 4, 1, f) == 1 && version == 1;
    ok = ok && fread(&seed, 8, 1, f) == 1 && fread(&step, 8, 1, f) == 1;
    ok = ok && fread(&nt, 4, 1, f) == 1 && fread(&nrep, 4, 1, f) == 1 && fread(&n, 4, 1, f) == 1;
    ok = ok && nrep >= 1 && nrep <= (1u << 20) && n <= (uint32_t)RD_MAX_PARTICLES;
    std::vector<double> bounds;
    std::vector<uint32_t> streams;
    std::vector<int> counts;
    uint64_t csum = 0;
    for (uint32_t r = 0; ok && r < nrep; ++r) {
        double lo[3], hi[3];
        uint32_t st = 0, cnt = 0;
        ok = fread(lo, 8, 3, f) == 3 && fread(hi, 8, 3, f) == 3 && fread(&st, 4, 1, f) == 1 && fread(&cnt, 4, 1, f) == 1;
        for (int d = 0; d < 3; ++d) { bounds.push_back(lo[d]); bounds.push_back(hi[d]); }
        streams.push_back(st);
        // the replica sizes must add up to the particle count the arrays hold
        ok = ok && cnt <= (uint32_t)RD_MAX_PARTICLES;
        counts.push_back(ok ? (int)cnt : 0);
        csum += cnt;
    }
    ok = ok && csum == (uint64_t)n;
    std::vector<float> xyz(ok ? 3 * (size_t)n : 0);
    std::vector<int> type(ok ? (size_t)n : 0);
    ok = ok && fread(xyz.data(), 4, xyz.size(), f) == xyz.size() && fread(type.data(), 4, type.size(), f) == type.size();
    ok = ok && fread(&nb, 4, 1, f) == 1 && nb <= (1u << 28);
    std::vector<int> bp(ok ? 2 * (size_t)nb : 0), bo(ok ? (size_t)nb : 0);
    ok = ok && fread(bp.data(), 4, bp.size(), f) == bp.size() && fread(bo.data(), 4, bo.size(), f) == bo.size();
    fclose(f);
    if (!ok) return set_err(h, RD_EINVAL, std::string("not a snapshot of this format: ") + path);
    if ((int)nt != h->nt) return set_err(h, RD_EINVAL, "snapshot has another number of particle types");
    if (seed != h->prm.seed) return set_err(h, RD_EINVAL, "snapshot has another noise seed");
    int rc;
    if (nrep == 1 && streams[0] == h->prm.stream) {
        bool same_box = true;
        for (int d = 0; d < 3; ++d)
            same_box = same_box && bounds[2 * d] == h->prm.lattice_bounds[2 * d] &&
                       bounds[2 * d + 1] == h->prm.lattice_bounds[2 * d + 1];
        rc = same_box ? rd_set_particles(h, (int)n, xyz.data(), type.data())
                      : rd_set_replicas(h, 1, bounds.data(), streams.data(), counts.data(), xyz.data(), type.data());
    } else {
        rc = rd_set_replicas(h, (int)nrep, bounds.data(), streams.data(), counts.data(), xyz.data(), type.data());
    }
    if (rc) return rc;
    if (nb > 0) {
        if (h->bpots.empty()) return set_err(h, RD_ESTATE, "snapshot has bonds: set the bond potentials first");
        rc = rd_set_bonds(h, (int)nb, bp.data(), bo.data(), (int)h->bpots.size(), h->bpots.data(),
                          h->bpot_r0.data());
    } else {
        h->bond_pairs.clear();
        h->bond_pot.clear();
        h->dirty = true;
    }
    if (rc) return rc;
    if (!h->initialized) {
        h->step = step;
        return RD_OK;
    }
    rc = refresh(h);
    if (rc) return rc;
    return rd_set_step(h, step);
}

// ---- invariant check of the sorted state (debug aid; no counterpart in the
// reference, whose only integrity checks are the host list walkers of
// CudaSimulation.cu:628-726) ----
// Each slot lies inside the slot range of the cell its position maps to; each
// particle index occurs once; the cell starts rise and end at the particle
// count.  Violations are counted.
__global__ __launch_bounds__(RD_BLOCK) void k_check_slots(const float4* __restrict__ ps,
                                                          const uint32_t* __restrict__ start,
                                                          const RdRep* __restrict__ reps,
                                                          const RdWave* __restrict__ waves, int nwaves,
                                                          uint32_t* __restrict__ seen, uint32_t* __restrict__ bad)
{
    const RdWave it = wave_item(waves, nwaves, wave_id());
    const int s = it.s0 + (threadIdx.x & 63);
    if (s >= it.s1) return;
    const RdRep R = reps[it.rep];
    const float4 p = ps[s];
    const uint32_t k = cell_key(R, p.x, p.y, p.z);
    if (!((uint32_t)s >= start[k] && (uint32_t)s < start[k + 1])) atomicAdd(bad, 1u);
    const uint32_t li = w_index(p.w);
    if (li >= (uint32_t)R.n) { atomicAdd(bad, 1u); return; }
    atomicAdd(&seen[R.off + li], 1u);
}

__global__ __launch_bounds__(RD_BLOCK) void k_check_counts(const uint32_t* __restrict__ seen, int n,
                                                           const uint32_t* __restrict__ start, int ncell,
                                                           const uint32_t* __restrict__ count,
                                                           const uint32_t* __restrict__ tsum, int ntile,
                                                           uint32_t* __restrict__ bad)
{
    const long i = (long)blockIdx.x * RD_BLOCK + threadIdx.x;
    if (i < n && seen[i] != 1u) atomicAdd(bad, 1u);
    if (i < ncell && start[i] > start[i + 1]) atomicAdd(bad, 1u);
    if (i == 0 && start[ncell] != (uint32_t)n) atomicAdd(bad, 1u);
    // between steps the histogram and its tile sums are clear
    if (i < ncell && count[i] != 0u) atomicAdd(bad, 1u);
    if (i < ntile && tsum[i] != 0u) atomicAdd(bad, 1u);
}

extern "C" int rd_check_invariants(rd_system* h, long* violations)
{
    if (!h || !violations) return RD_EINVAL;
    int rc = refresh(h);
    if (rc) return rc;
    uint32_t *d_seen = nullptr, *d_bad = nullptr;
    hipError_t e = hipMalloc(&d_seen, sizeof(uint32_t) * ((size_t)h->n + 1));
    if (e == hipSuccess) e = hipMalloc(&d_bad, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(d_seen, 0, sizeof(uint32_t) * ((size_t)h->n + 1), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, sizeof(uint32_t), h->stream);
    uint32_t bad = 0;
    if (e == hipSuccess) {
        if (h->n > 0)
            hipLaunchKernelGGL(k_check_slots, dim3(wave_blocks(h)), dim3(RD_BLOCK), 0, h->stream, h->d_ps, h->d_start,
                               h->d_reps, h->d_waves, h->nwaves, d_seen, d_bad);
        const long m = std::max<long>(h->n, h->ncell);
        hipLaunchKernelGGL(k_check_counts, dim3(std::max(1, nblk(m, RD_BLOCK))), dim3(RD_BLOCK), 0, h->stream, d_seen,
                           h->n, h->d_start, h->ncell, h->d_count, h->d_bsum, h->nscan, d_bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (d_seen) (void)hipFree(d_seen);
    if (d_bad) (void)hipFree(d_bad);
    if (e != hipSuccess) return set_err(h, RD_EHIP, std::string("invariant check: ") + hipGetErrorString(e));
    *violations = (long)bad;
    return RD_OK;
}
