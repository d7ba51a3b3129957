// rd_kernels.h -- kernels of the plain BD step on gfx950.
//
// Data layout in HBM (one system):
//   ps[N]   float4, sorted by cell: x, y, z, w = (index << 6) | type
//   pn[N]   float4, positions after the update, in the slot order of ps
//   key[N]  new cell id of slot s;  rank[N]: rank of slot s inside its cell
//   count[ncell], start[ncell+1]: counting-sort histogram and offsets
// One step = force_update (gather forces, Euler-Maruyama, wrap, new cell key,
// histogram) -> scan (3 phases) -> scatter.  Replaces orderOne, warpOrderTwo
// and update (CudaSimulation.cu:2309-2476) and the linked-list cell
// maintenance with spin locks (CudaSimulation.cu:1860-1914).
#pragma once
#include "rd_device.h"

#define RD_BLOCK 256
#define RD_SCAN_ITEMS 16
#define RD_SCAN_TILE (RD_BLOCK * RD_SCAN_ITEMS)

struct RdStepState {
    unsigned long long step;        // global step counter (noise counter)
    unsigned long long noise_step0; // first step held by the noise array
    int noise_nsteps;               // 0: generator
    unsigned int err;               // bit 0: non-finite position
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of slots so that neighbour cells share one L2.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
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

// Visit every slot j of the cells within S of (cx,cy,cz).  Rows along x are
// contiguous in the sorted arrays, so each row is one or two slot ranges.
template <class F>
__device__ __forceinline__ void for_stencil(const RdDev* __restrict__ P, const uint32_t* __restrict__ start,
                                            int cx, int cy, int cz, int S, F&& f)
{
    const int Lx = P->L[0], Ly = P->L[1], Lz = P->L[2];
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
            int row = (zz * Ly + yy) * Lx;
            uint32_t j0 = start[row + a0], j1 = start[row + b0 + 1];
            for (uint32_t j = j0; j < j1; ++j) f(j);
            if (nr == 2) {
                j0 = start[row + a1]; j1 = start[row + b1 + 1];
                for (uint32_t j = j0; j < j1; ++j) f(j);
            }
        }
    }
}

// Pair gradients are summed in 64-bit fixed point (2^-32 resolution): integer
// addition is exact, so the sum does not depend on the order of particles
// inside a cell, which the atomic counting sort leaves unspecified.  The step
// is therefore bitwise reproducible.
#define RD_FIX_SCALE 4294967296.0f
#define RD_FIX_INV 2.3283064365386963e-10
__device__ __forceinline__ long long to_fix(float v)
{
    v = fminf(fmaxf(v, -1.0e9f), 1.0e9f);
    return (long long)(v * RD_FIX_SCALE);
}

// gradient of U (pair + external) for the particle in slot s
template <bool ENERGY>
__device__ __forceinline__ float3 gather_grad(const RdDev* __restrict__ P, const RdPot1Dev* __restrict__ pot1,
                                              const float4* __restrict__ ps, const uint32_t* __restrict__ start,
                                              uint32_t s, float4 p, float* energy)
{
    const int ti = w_type(p.w);
    const int nt = P->nt;
    long long ax = 0, ay = 0, az = 0;
    float e = 0.0f;
    int cx = cell_coord(p.x, P->lo[0], P->cell, P->L[0]);
    int cy = cell_coord(p.y, P->lo[1], P->cell, P->L[1]);
    int cz = cell_coord(p.z, P->lo[2], P->cell, P->L[2]);
    const float Lbx = P->Lb[0], Lby = P->Lb[1], Lbz = P->Lb[2];
    const float hx = P->half[0], hy = P->half[1], hz = P->half[2];
    const float* rc2row = &P->rc2[ti * nt];
    const float* r0row = &P->r0[ti * nt];
    const uint32_t* mrow = &P->p2mask[ti * nt];
    if (P->n_pot2 > 0) {
        for_stencil(P, start, cx, cy, cz, P->S, [&](uint32_t j) {
            if (j == s) return;
            float4 q = ps[j];
            float dx = min_image(p.x - q.x, Lbx, hx);
            float dy = min_image(p.y - q.y, Lby, hy);
            float dz = min_image(p.z - q.z, Lbz, hz);
            float d2 = dx * dx;
            d2 = d2 + dy * dy;
            d2 = d2 + dz * dz;
            int tj = w_type(q.w);
            if (d2 < rc2row[tj] && d2 > 0.0f) {
                float r = sqrtf(d2);
                float r0 = r0row[tj];
                uint32_t m = mrow[tj];
                float pre = 0.0f, ep = 0.0f;
                while (m) {
                    int pi = __builtin_ctz(m);
                    m &= m - 1;
                    pre += pair_pre<ENERGY>(P->pot2[pi], r0, r, &ep);
                }
                ax += to_fix(pre * dx); ay += to_fix(pre * dy); az += to_fix(pre * dz);
                if (ENERGY) e += 0.5f * ep;
            }
        });
    }
    float gx = (float)((double)ax * RD_FIX_INV);
    float gy = (float)((double)ay * RD_FIX_INV);
    float gz = (float)((double)az * RD_FIX_INV);
    const int np1 = P->n_pot1;
    const float prad = P->prad[ti];
    for (int k = 0; k < np1; ++k) {
        const RdPot1Dev& q = pot1[k];
        if (!((q.mask >> ti) & 1u)) continue;
        ext_grad<ENERGY>(q, prad, p.x, p.y, p.z, gx, gy, gz, &e);
    }
    if (ENERGY) *energy = e;
    return make_float3(gx, gy, gz);
}

__device__ __forceinline__ float wrap_coord(float x, float lo, float hi, float Lb)
{
    if (x > hi) x -= Lb;
    if (x < lo) x += Lb;
    return x;
}

__device__ __forceinline__ uint32_t cell_key(const RdDev* __restrict__ P, float x, float y, float z)
{
    int cx = cell_coord(x, P->lo[0], P->cell, P->L[0]);
    int cy = cell_coord(y, P->lo[1], P->cell, P->L[1]);
    int cz = cell_coord(z, P->lo[2], P->cell, P->L[2]);
    return (uint32_t)((cz * P->L[1] + cy) * P->L[0] + cx);
}
