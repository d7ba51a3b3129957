// rd_tiled.h -- LDS-tiled force gather for the 27-cell stencil (gfx950).
//
// A workgroup owns a tile of cells: TY consecutive rows (y) of one plane (z),
// and along x either the whole row (nseg == 1) or a segment of TW cells.  It
// stages into LDS every particle of the 3 x (TY+2) neighbour rows restricted
// to the tile's x-range plus a one-cell halo, together with the LDS offset of
// every staged cell.  Staging loads are contiguous slot ranges, issued by all
// lanes at once; the stencil scan then reads LDS only.
//
// Replaces warpOrderTwo (CudaSimulation.cu:1402-1782), which walks per-cell
// linked lists with 32-lane blocks and global float atomics.
#pragma once
#include "rd_wave.h"

struct RdTileGeom {
    int TY;      // rows per tile
    int TW;      // cells per segment (== Lx when nseg == 1)
    int W;       // staged cells per row (Lx when nseg == 1, else TW + 2)
    int nseg;    // segments per row
    int nyt;     // tiles along y
    int cap;     // staged particle capacity
    int ntiles;
};

// LDS layout (dynamic): RdTab | rows info | cell offsets | particles
struct RdTileRow {
    int row;      // linear index of the row's first cell
    int g0;       // global slot of the first staged particle
    int nA;       // staged particles before the x wrap
    int g1;       // global slot continuing after the wrap (start[row])
    int lbase;    // LDS offset of the row's first staged particle
    int n;        // staged particles of the row
};

__host__ __device__ inline size_t rd_tile_lds_bytes(const RdTileGeom& G)
{
    int nr = 3 * (G.TY + 2);
    size_t b = sizeof(RdTab);
    b += sizeof(RdTileRow) * (size_t)nr + 16;
    b += sizeof(int) * (size_t)nr * (G.W + 1);
    b = (b + 15) & ~(size_t)15;
    b += sizeof(float4) * (size_t)G.cap;
    return b;
}


template <int BLOCK, int MODE>
__global__ __launch_bounds__(BLOCK) void k_force_tiled(const float4* __restrict__ ps, float4* __restrict__ pn,
                                                       uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
                                                       uint32_t* __restrict__ count,
                                                       const uint32_t* __restrict__ start,
                                                       const RdDev* __restrict__ P,
                                                       const RdRep* __restrict__ reps,
                                                       const RdPot1Dev* __restrict__ pot1,
                                                       RdStepState* __restrict__ st,
                                                       const float* __restrict__ noise, int n, RdTileGeom G,
                                                       float* __restrict__ grad, double* __restrict__ etot)
{
    extern __shared__ __align__(16) unsigned char smem[];
    RdTab& T = *reinterpret_cast<RdTab*>(smem);
    const int TY = G.TY, W = G.W;
    const int NR = 3 * (TY + 2);
    RdTileRow* rows = reinterpret_cast<RdTileRow*>(smem + sizeof(RdTab));
    int* lsx = reinterpret_cast<int*>(rows + NR) + 4;          // [NR][W+1]
    size_t off = sizeof(RdTab) + sizeof(RdTileRow) * (size_t)NR + 16 + sizeof(int) * (size_t)NR * (W + 1);
    off = (off + 15) & ~(size_t)15;
    float4* buf = reinterpret_cast<float4*>(smem + off);
    int* total_sh = reinterpret_cast<int*>(rows + NR);          // [0] staged total

    load_tab(T, P);   // ends with a barrier

    const RdRep RG = reps[0];   // the tiled kernel runs single-replica systems
    const int Lx = RG.L[0], Ly = RG.L[1], Lz = RG.L[2];
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int seg = tile % G.nseg;
    const int yt = (tile / G.nseg) % G.nyt;
    const int z = tile / (G.nseg * G.nyt);
    const int y0 = yt * TY;
    const int x0 = seg * G.TW;
    const int xs = G.nseg == 1 ? 0 : (x0 == 0 ? Lx - 1 : x0 - 1);   // first staged cell

    // --- rows: global ranges and counts ---
    if (threadIdx.x < NR) {
        int k = threadIdx.x;
        int dzi = k / (TY + 2), yi = k % (TY + 2);
        int yy = y0 - 1 + yi, zz = z - 1 + dzi;
        yy = (yy % Ly + Ly) % Ly;
        zz = (zz % Lz + Lz) % Lz;
        int row = RG.coff + (zz * Ly + yy) * Lx;
        int g0 = (int)start[row + xs];
        int nA, g1 = 0, ntot;
        if (xs + W <= Lx) {
            nA = (int)start[row + xs + W] - g0;
            ntot = nA;
        } else {
            nA = (int)start[row + Lx] - g0;
            g1 = (int)start[row];
            ntot = nA + (int)start[row + xs + W - Lx] - g1;
        }
        rows[k].row = row; rows[k].g0 = g0; rows[k].nA = nA; rows[k].g1 = g1; rows[k].n = ntot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < NR; ++k) { rows[k].lbase = acc; acc += rows[k].n; }
        total_sh[0] = acc;
    }
    __syncthreads();
    const int total = total_sh[0];
    const bool staged = total <= G.cap;

    // --- cell offsets in LDS coordinates ---
    if (staged) {
        for (int t = threadIdx.x; t < NR * (W + 1); t += BLOCK) {
            int k = t / (W + 1), i = t - k * (W + 1);
            const RdTileRow& R = rows[k];
            int rel;
            if (i == W) rel = R.n;
            else {
                int x = xs + i;
                if (x < Lx) rel = (int)start[R.row + x] - R.g0;
                else rel = R.nA + (int)start[R.row + x - Lx] - R.g1;
            }
            lsx[t] = R.lbase + rel;
        }
        // --- particles ---
        for (int t = threadIdx.x; t < total; t += BLOCK) {
            int k = 0;
            while (k + 1 < NR && rows[k + 1].lbase <= t) ++k;
            const RdTileRow& R = rows[k];
            int r = t - R.lbase;
            int gsl = r < R.nA ? R.g0 + r : R.g1 + (r - R.nA);
            buf[t] = ps[gsl];
        }
    }
    __syncthreads();

    // --- home particles: rows yi = 1..TY of the middle plane, cells x0..x0+TW-1 ---
    const int ci0 = G.nseg == 1 ? 0 : 1;
    const int twe = min(G.TW, Lx - x0);
    const int hy = min(TY, Ly - y0);
    // home ranges per home row (LDS coordinates)
    int hcount = 0;
    for (int yi = 1; yi <= hy; ++yi) {
        int k = (TY + 2) + yi;
        hcount += staged ? lsx[k * (W + 1) + ci0 + twe] - lsx[k * (W + 1) + ci0]
                         : 0;
    }
    double eblk = 0.0;

    if (staged) {
        for (int h = threadIdx.x; h < hcount; h += BLOCK) {
            // locate home row
            int yi = 1, acc = 0, k = TY + 2 + 1, lh0 = 0, lh1 = 0;
            for (;; ++yi) {
                k = (TY + 2) + yi;
                lh0 = lsx[k * (W + 1) + ci0];
                lh1 = lsx[k * (W + 1) + ci0 + twe];
                if (h < acc + (lh1 - lh0) || yi == hy) break;
                acc += lh1 - lh0;
            }
            const int l = lh0 + (h - acc);
            const RdTileRow& R = rows[k];
            const int rr = l - R.lbase;
            const uint32_t s = (uint32_t)(rr < R.nA ? R.g0 + rr : R.g1 + (rr - R.nA));
            const float4 p = buf[l];
            const int ti = w_type(p.w);
            const int rowbase = ti * P->nt;
            const int cx = cell_coord(p.x, RG.lo[0], RG.cell, Lx);
            int ci = cx - xs;
            if (ci < 0) ci += Lx;
            long long ax = 0, ay = 0, az = 0;
            float e = 0.0f;
            auto scan = [&](int a, int b) {
                for (int j = a; j < b; ++j) {
                    float4 q = buf[j];
                    if (__float_as_uint(q.w) != __float_as_uint(p.w))
                        pair_accum<MODE == RD_MODE_GRAD>(T, rowbase, p, q, RG, ax, ay, az, e);
                }
            };
            if (P->n_pot2 > 0) {
                for (int dzi = 0; dzi < 3; ++dzi) {
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int* L = &lsx[(dzi * (TY + 2) + yi + dy) * (W + 1)];
                        if (G.nseg > 1 || (ci >= 1 && ci <= Lx - 2)) {
                            scan(L[ci - 1], L[ci + 2]);
                        } else if (ci == 0) {
                            scan(L[0], L[2]);
                            scan(L[Lx - 1], L[Lx]);
                        } else {
                            scan(L[Lx - 2], L[Lx]);
                            scan(L[0], L[1]);
                        }
                    }
                }
            }
            float gx = (float)((double)ax * RD_FIX_INV);
            float gy = (float)((double)ay * RD_FIX_INV);
            float gz = (float)((double)az * RD_FIX_INV);
            const int np1 = P->n_pot1;
            const float prad = T.prad[ti];
            for (int kk = 0; kk < np1; ++kk) {
                const RdPot1Dev& q = pot1[kk];
                if (!((q.mask >> ti) & 1u)) continue;
                ext_grad<MODE == RD_MODE_GRAD>(q, prad, p.x, p.y, p.z, gx, gy, gz, &e);
            }
            if (MODE == RD_MODE_GRAD) {
                const uint32_t idx = (uint32_t)RG.off + w_index(p.w);
                grad[3 * idx] = gx; grad[3 * idx + 1] = gy; grad[3 * idx + 2] = gz;
                eblk += (double)e;
            } else {
                em_update_store(P, RG, T, st, noise, n, s, p, make_float3(gx, gy, gz), pn, key, rank, count);
            }
        }
    } else {
        // the tile overflowed its LDS budget: gather from global memory
        for (int yi = 1; yi <= hy; ++yi) {
            int row = RG.coff + ((z * Ly) + (y0 + yi - 1)) * Lx;
            uint32_t s0 = start[row + x0], s1 = start[row + x0 + twe];
            for (uint32_t s = s0 + threadIdx.x; s < s1; s += BLOCK) {
                float4 p = ps[s];
                float e = 0.0f;
                float3 g = gather_grad<MODE == RD_MODE_GRAD>(P, RG, T, pot1, ps, start, s, p, &e);
                if (MODE == RD_MODE_GRAD) {
                    const uint32_t idx = (uint32_t)RG.off + w_index(p.w);
                    grad[3 * idx] = g.x; grad[3 * idx + 1] = g.y; grad[3 * idx + 2] = g.z;
                    eblk += (double)e;
                } else {
                    em_update_store(P, RG, T, st, noise, n, s, p, g, pn, key, rank, count);
                }
            }
        }
    }
    if (MODE == RD_MODE_GRAD) {
        __shared__ double red[BLOCK / 64];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) eblk += __shfl_down(eblk, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = eblk;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
            atomicAdd(etot, t);
        }
    }
}
