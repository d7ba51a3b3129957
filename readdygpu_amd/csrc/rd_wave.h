// rd_wave.h -- wave-tiled force gather (gfx950).
//
// Each wave owns up to 64 consecutive sorted slots of one replica.  Their
// cells form one range [ca, cb] of the replica's linear cell order; the
// 27-cell stencil of that range lies inside 9 slot ranges, one per (dy, dz)
// row offset d:
//     [start[ca + d - 1], start[cb + d + 2])
// The wave stages those ranges into LDS (contiguous loads, all in flight),
// about 9 x 64 particles whatever the density, then each lane scans its own 9
// row runs in LDS.  Runs that fall outside the staged ranges (x/y/z wrap,
// ranges beyond the LDS budget) are read from global memory, so every
// configuration is exact.  Pair gradients are summed in fixed point, so the
// result does not depend on which path a pair took.
#pragma once
#include "rd_kernels.h"

#define RD_WAVE_BUDGET 512          // staged particles per wave
#define RD_WAVE_WAVES 4             // waves per workgroup

struct RdWaveRange {
    int g0;       // first staged global slot
    int g1;       // one past the last staged slot (g0 when not staged)
    int lb;       // LDS offset
    int pad;
};

enum { RD_MODE_UPDATE = 0, RD_MODE_GRAD = 1 };

template <int MODE>
__global__ __launch_bounds__(64 * RD_WAVE_WAVES) void k_force_wave(
    const float4* __restrict__ ps, float4* __restrict__ pn, uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
    uint32_t* __restrict__ count, const uint32_t* __restrict__ start, const RdDev* __restrict__ P,
    const RdRep* __restrict__ reps, const RdWave* __restrict__ waves, int nwaves,
    const RdPot1Dev* __restrict__ pot1, RdStepState* __restrict__ st, const float* __restrict__ noise, int ntot,
    float* __restrict__ grad, double* __restrict__ etot)
{
    __shared__ RdTab T;
    __shared__ float4 buf[RD_WAVE_WAVES][RD_WAVE_BUDGET];
    __shared__ RdWaveRange rg[RD_WAVE_WAVES][9];
    __shared__ int tot_sh[RD_WAVE_WAVES];
    load_tab(T, P);   // ends with a barrier

    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int wid = xcd_block(blockIdx.x, gridDim.x) * RD_WAVE_WAVES + wv;
    const RdWave it = wave_item(waves, nwaves, wid);
    const RdRep R = reps[it.rep];
    const int s = it.s0 + lane;
    const bool valid = s < it.s1;
    const int nvalid = it.s1 - it.s0;                  // 0..64
    const int Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    const int LxLy = Lx * Ly;
    const int ncell = LxLy * Lz;

    float4 p = valid ? ps[s] : make_float4(R.lo[0], R.lo[1], R.lo[2], 0.f);
    int cx, cy, cz;
    cell_coords(R, p.x, p.y, p.z, cx, cy, cz);
    const int c = (cz * Ly + cy) * Lx + cx;            // replica-local cell
    const int last = nvalid > 0 ? nvalid - 1 : 0;
    const int ca = __shfl(c, 0, 64);
    const int cb = __shfl(c, last, 64);

    // --- the 9 shifted ranges (lanes 0..8), staged within the budget ---
    if (lane < 9) {
        int g0 = 0, len = 0;
        if (nvalid > 0 && P->n_pot2 > 0) {
            int dz = lane / 3 - 1, dy = lane % 3 - 1;
            int d = dz * LxLy + dy * Lx;
            int lo = ca + d - 1, hi = cb + d + 1;
            lo = lo < 0 ? 0 : lo;
            hi = hi > ncell - 1 ? ncell - 1 : hi;
            if (lo <= hi) {
                g0 = (int)start[R.coff + lo];
                len = (int)start[R.coff + hi + 1] - g0;
            }
        }
        rg[wv][lane].g0 = g0;
        rg[wv][lane].g1 = len;     // length for now
    }
    __syncthreads();
    if (lane == 0) {
        int acc = 0;
        for (int k = 0; k < 9; ++k) {
            int len = rg[wv][k].g1;
            bool fit = acc + len <= RD_WAVE_BUDGET;
            rg[wv][k].lb = acc;
            rg[wv][k].g1 = rg[wv][k].g0 + (fit ? len : 0);
            acc += fit ? len : 0;
        }
        tot_sh[wv] = acc;
    }
    __syncthreads();

    // --- stage (all loads in flight) ---
    const int total = tot_sh[wv];
    for (int t = lane; t < total; t += 64) {
        // staged ranges are laid out in order: t belongs to the last
        // non-empty range that starts at or before it
        int k = 0;
#pragma unroll
        for (int q = 1; q < 9; ++q)
            if (rg[wv][q].lb <= t && rg[wv][q].g1 > rg[wv][q].g0) k = q;
        const RdWaveRange& W = rg[wv][k];
        buf[wv][t] = ps[W.g0 + (t - W.lb)];
    }

    // --- own 9 row runs: global slot ranges ---
    int a0 = 0, b0 = 0, a1 = 0, b1 = 0;
    const int nr = x_runs(cx, 1, Lx, &a0, &b0, &a1, &b1);
    int rb[9], re[9], rb2[9], re2[9];
    const bool fast = valid && P->n_pot2 > 0 && P->S == 1 && Lx >= 3 && Ly >= 3 && Lz >= 3;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        rb[r] = re[r] = rb2[r] = re2[r] = 0;
        if (fast) {
            int yy = cy + (r % 3) - 1, zz = cz + (r / 3) - 1;
            yy += yy < 0 ? Ly : 0; yy -= yy >= Ly ? Ly : 0;
            zz += zz < 0 ? Lz : 0; zz -= zz >= Lz ? Lz : 0;
            const int row = R.coff + (zz * Ly + yy) * Lx;
            rb[r] = (int)start[row + a0];
            re[r] = (int)start[row + b0 + 1];
            if (nr == 2) { rb2[r] = (int)start[row + a1]; re2[r] = (int)start[row + b1 + 1]; }
        }
    }
    __syncthreads();

    long long ax = 0, ay = 0, az = 0;
    float e = 0.0f;
    const int rowbase = w_type(p.w) * P->nt;
    const uint32_t pw = __float_as_uint(p.w);
    auto scan = [&](int k, int a, int b) {
        if (a >= b) return;
        const RdWaveRange& W = rg[wv][k];
        if (a >= W.g0 && b <= W.g1) {
            for (int j = W.lb + (a - W.g0), je = W.lb + (b - W.g0); j < je; ++j) {
                float4 q = buf[wv][j];
                if (__float_as_uint(q.w) != pw) pair_accum<MODE == RD_MODE_GRAD>(T, rowbase, p, q, R, ax, ay, az, e);
            }
        } else {
            int j = a;
            for (; j + 1 < b; j += 2) {
                float4 q0 = ps[j], q1 = ps[j + 1];
                if (__float_as_uint(q0.w) != pw) pair_accum<MODE == RD_MODE_GRAD>(T, rowbase, p, q0, R, ax, ay, az, e);
                if (__float_as_uint(q1.w) != pw) pair_accum<MODE == RD_MODE_GRAD>(T, rowbase, p, q1, R, ax, ay, az, e);
            }
            if (j < b) {
                float4 q = ps[j];
                if (__float_as_uint(q.w) != pw) pair_accum<MODE == RD_MODE_GRAD>(T, rowbase, p, q, R, ax, ay, az, e);
            }
        }
    };
    if (fast) {
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            scan(r, rb[r], re[r]);
            scan(r, rb2[r], re2[r]);
        }
    } else if (valid && P->n_pot2 > 0) {
        for_stencil(R, start, cx, cy, cz, P->S, [&](uint32_t j) {
            float4 q = ps[j];
            if (__float_as_uint(q.w) != pw) pair_accum<MODE == RD_MODE_GRAD>(T, rowbase, p, q, R, ax, ay, az, e);
        });
    }

    double eblk = 0.0;
    if (valid) {
        float3 g = finish_grad<MODE == RD_MODE_GRAD>(P, T, pot1, (uint32_t)s, p, ax, ay, az, &e);
        if (MODE == RD_MODE_GRAD) {
            const uint32_t gi = (uint32_t)R.off + w_index(p.w);
            grad[3 * (size_t)gi] = g.x; grad[3 * (size_t)gi + 1] = g.y; grad[3 * (size_t)gi + 2] = g.z;
            eblk = (double)e;
        } else {
            em_update_store(P, R, T, st, noise, ntot, (uint32_t)s, p, g, pn, key, rank, count);
        }
    }
    if (MODE == RD_MODE_GRAD) {
        __shared__ double red[RD_WAVE_WAVES];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) eblk += __shfl_down(eblk, o, 64);
        if (lane == 0) red[wv] = eblk;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < RD_WAVE_WAVES; ++w) t += red[w];
            atomicAdd(etot, t);
        }
    }
}
