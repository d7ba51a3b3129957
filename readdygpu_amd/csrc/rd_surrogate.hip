// rd_surrogate.hip -- composite external potential and its BF16 MLP surrogate.
//
// The composite is a geometry of order-one pieces (attractive disks and
// cylinders, the forms of CudaSimulation.cu:1049-1148 with any normal); a
// particle feels the piece nearest to it in energy, or the sum of all pieces
// (rd_set_composite_combine).  Evaluating it exactly
// costs one pass over every piece per particle (the reference's orderOne loop
// shape, CudaSimulation.cu:1049-1236), VALU-bound.  The surrogate is an MLP
// 3 -> 64 -> 64 -> 64 -> 3 trained on the GPU, by the kernels below, on
// gradients that the exact evaluator generates; in the step it replaces the
// exact pass.  Both passes write one float4 per sorted slot (gradient,
// energy) that the force kernel adds to the particle's gradient.
#include "rd_host.h"
#include "rd_kernels.h"
#include "rd_mlp.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

// ===========================================================================
// Exact composite
// ===========================================================================
struct RdPiece {
    float o[3];
    float n[3];        // unit axis
    float R;           // radius
    float hh;          // half height (0 for a disk: its energy along the axis is s^2)
    float k;
    int zaxis;         // axis is +z: cheaper projection
};                     // 40 bytes: 512 pieces and a 2 KB queue leave room for 7 workgroups per CU

__device__ __forceinline__ float piece_energy(const RdPiece& p, float x, float y, float z)
{
    float vx = x - p.o[0], vy = y - p.o[1], vz = z - p.o[2];
    float s, r2;
    if (p.zaxis) {
        s = vz;
        r2 = vx * vx + vy * vy;
    } else {
        s = vx * p.n[0] + vy * p.n[1] + vz * p.n[2];
        float qx = vx - s * p.n[0], qy = vy - s * p.n[1], qz = vz - s * p.n[2];
        r2 = qx * qx + qy * qy + qz * qz;
    }
    float dn = fmaxf(fabsf(s) - p.hh, 0.0f);
    float r = __builtin_amdgcn_sqrtf(r2);
    float dr = fmaxf(r - p.R, 0.0f);
    return 0.5f * p.k * (dn * dn + dr * dr);
}

// gradient of one piece: attractive disk / cylinder (ext_grad forms)
__device__ __forceinline__ float3 piece_grad(const RdPiece& p, float x, float y, float z)
{
    float vx = x - p.o[0], vy = y - p.o[1], vz = z - p.o[2];
    float s = vx * p.n[0] + vy * p.n[1] + vz * p.n[2];
    float qx = vx - s * p.n[0], qy = vy - s * p.n[1], qz = vz - s * p.n[2];
    float r, rinv;
    fast_len(qx * qx + qy * qy + qz * qz, &r, &rinv);
    float as = fabsf(s), sg = s > 0.0f ? 1.0f : (s < 0.0f ? -1.0f : 0.0f);
    float gn = as > p.hh ? p.k * (as - p.hh) * sg : 0.0f;
    float gx = gn * p.n[0], gy = gn * p.n[1], gz = gn * p.n[2];
    if (r > p.R) {
        float pre = p.k * (r - p.R) * rinv;
        gx += pre * qx; gy += pre * qy; gz += pre * qz;
    }
    return make_float3(gx, gy, gz);
}

#define RD_PIECES_MAX 2048

// composite at one position: energy and gradient.  SUM: the sum over every
// piece, as orderOne adds every order-one potential (CudaSimulation.cu:
// 1049-1236); otherwise the nearest piece in energy (first argmin).
template <bool SUM>
__device__ __forceinline__ float4 composite_at(const RdPiece* __restrict__ pc, int np, float x, float y, float z)
{
    if constexpr (SUM) {
        float gx = 0.f, gy = 0.f, gz = 0.f, e = 0.f;
        for (int k = 0; k < np; ++k) {
            const float3 g = piece_grad(pc[k], x, y, z);
            gx += g.x; gy += g.y; gz += g.z;
            e += piece_energy(pc[k], x, y, z);
        }
        return make_float4(gx, gy, gz, e);
    }
    float best = INFINITY;
    int kb = 0;
    for (int k = 0; k < np; ++k) {
        float e = piece_energy(pc[k], x, y, z);
        if (e < best) { best = e; kb = k; }
    }
    if (np == 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    float3 g = piece_grad(pc[kb], x, y, z);
    return make_float4(g.x, g.y, g.z, best);
}

// exact pass over sorted slots: gext[s] for the particles whose type is in the
// mask (other slots are left alone: the force kernel reads gext only for
// those types).  One round of resident workgroups takes chunks of 256 slots
// in turn (selected particles gather in some cells, e.g. on a membrane, so
// contiguous shares would be unbalanced), and compacts the selected ones
// into an LDS queue; whenever 256 are queued, every thread evaluates one, so
// no lane idles on an unselected particle.
// Piece table in dynamic LDS: np * sizeof(RdPiece) bytes.
template <bool SUM>
__global__ __launch_bounds__(256) void k_composite_slots(const float4* __restrict__ ps, int n,
                                                         const RdPiece* __restrict__ pieces, int np,
                                                         uint32_t type_mask, float4* __restrict__ gext)
{
    extern __shared__ RdPiece pc[];
    __shared__ uint32_t q[512];
    __shared__ int wcnt[4];
    for (int k = threadIdx.x; k < np; k += 256) pc[k] = pieces[k];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    auto eval = [&](uint32_t s) {
        float4 p = ps[s];
        gext[s] = composite_at<SUM>(pc, np, p.x, p.y, p.z);
    };
    int qn = 0;                       // queued entries (same in every thread)
    for (long base = (long)xcd_block(blockIdx.x, gridDim.x) * 256; base < n; base += (long)gridDim.x * 256) {
        const long s = base + tid;
        const bool on = s < n && ((type_mask >> w_type(ps[s].w)) & 1u);
        const uint64_t b = __ballot(on);
        if (lane == 0) wcnt[w] = __popcll(b);
        __syncthreads();
        int off = qn, tot = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            off += k < w ? wcnt[k] : 0;
            tot += wcnt[k];
        }
        if (on) q[off + lanes_below(b)] = (uint32_t)s;      // qn < 256, tot <= 256
        qn += tot;
        __syncthreads();
        if (qn >= 256) {
            eval(q[tid]);
            __syncthreads();
            if (tid < qn - 256) q[tid] = q[256 + tid];
            qn -= 256;
            __syncthreads();
        }
    }
    if (tid < qn) eval(q[tid]);
}

// exact pass when the composite acts on every type: one slot per thread
template <bool SUM>
__global__ __launch_bounds__(256) void k_composite_all(const float4* __restrict__ ps, int n,
                                                       const RdPiece* __restrict__ pieces, int np,
                                                       float4* __restrict__ gext)
{
    extern __shared__ RdPiece pc[];
    for (int k = threadIdx.x; k < np; k += 256) pc[k] = pieces[k];
    __syncthreads();
    const long s = (long)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (s >= n) return;
    const float4 p = ps[s];
    gext[s] = composite_at<SUM>(pc, np, p.x, p.y, p.z);
}

// exact evaluation at given points (float4 input, .w ignored)
template <bool SUM>
__global__ __launch_bounds__(256) void k_composite_points(const float4* __restrict__ x, long m,
                                                          const RdPiece* __restrict__ pieces, int np,
                                                          float4* __restrict__ out)
{
    extern __shared__ RdPiece pc[];
    for (int k = threadIdx.x; k < np; k += 256) pc[k] = pieces[k];
    __syncthreads();
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    float4 p = x[i];
    out[i] = composite_at<SUM>(pc, np, p.x, p.y, p.z);
}

// ===========================================================================
// Training data: uniform positions in the domain, exact gradients
// ===========================================================================
__device__ __forceinline__ float u01(uint32_t b) { return (float)(b >> 8) * (1.0f / 16777216.0f); }

__device__ __forceinline__ float norm_coord(float x, float lo, float sc) { return (x - lo) * sc - 1.0f; }

template <bool SUM>
__global__ __launch_bounds__(256) void k_sur_data(long m, uint64_t seed, RdMlpNorm nm, float span0, float span1,
                                                  float span2, const RdPiece* __restrict__ pieces, int np,
                                                  float4* __restrict__ xn, float4* __restrict__ tg,
                                                  double* __restrict__ bsum)
{
    extern __shared__ RdPiece pc[];
    __shared__ double red[4];
    for (int k = threadIdx.x; k < np; k += 256) pc[k] = pieces[k];
    __syncthreads();
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    double sq = 0.0;
    if (i < m) {
        uint32_t c[4] = { (uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0x44415441u, 0x53555247u };
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        float x = nm.lo[0] + u01(c[0]) * span0;
        float y = nm.lo[1] + u01(c[1]) * span1;
        float z = nm.lo[2] + u01(c[2]) * span2;
        float4 g = composite_at<SUM>(pc, np, x, y, z);
        xn[i] = make_float4(norm_coord(x, nm.lo[0], nm.sc[0]), norm_coord(y, nm.lo[1], nm.sc[1]),
                            norm_coord(z, nm.lo[2], nm.sc[2]), 0.0f);
        tg[i] = make_float4(g.x, g.y, g.z, 0.0f);
        sq = (double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_down(sq, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// gscale = sqrt(sum / (3 m)) in a fixed order; targets divided by it
__global__ __launch_bounds__(256) void k_sur_scale_sum(const double* __restrict__ bsum, int nb, long m,
                                                       float* __restrict__ gscale)
{
    __shared__ double red[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) a += bsum[i];
    red[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < 256; ++i) t += red[i];
        double gs = sqrt(t / (3.0 * (double)m));
        *gscale = gs > 0.0 ? (float)gs : 1.0f;
    }
}

__global__ __launch_bounds__(256) void k_sur_normalise(float4* __restrict__ tg, long m, const float* __restrict__ gscale)
{
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    float inv = 1.0f / *gscale;
    float4 g = tg[i];
    tg[i] = make_float4(g.x * inv, g.y * inv, g.z * inv, 0.0f);
}

// ===========================================================================
// Training: forward + backward per wave of 64 samples -> partial gradients
// ===========================================================================
struct RdTrainCtl {
    unsigned long long step;   // optimiser steps taken
    float loss;                // loss of the last step
    float pad;
};

#define RD_TLD 72   // row stride (bf16) of the per-wave staging matrices [64][RD_TLD]

// activations (C layout) -> staging matrix S[feature][sample] in BF16
__device__ __forceinline__ void stage_act(__bf16* S, const float (&a)[4][4][4], int lane)
{
    const int j = lane & 15, g = lane >> 4;
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int r = 0; r < 4; ++r) S[(ti * 16 + 4 * g + r) * RD_TLD + tj * 16 + j] = (__bf16)a[tj][ti][r];
}

__device__ __forceinline__ rd_bf16x8 lds_frag(const __bf16* S, int row, int col)
{
    return *reinterpret_cast<const rd_bf16x8*>(S + row * RD_TLD + col);
}

// dW[i][k] = sum_samples D[i][s] H[k][s] for 64x64, written to out (row-major [64][64])
__device__ __forceinline__ void wgrad64(const __bf16* SD, const __bf16* SH, float* __restrict__ out, int lane)
{
    const int j = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
        rd_bf16x8 a0 = lds_frag(SD, ti * 16 + j, 8 * g), a1 = lds_frag(SD, ti * 16 + j, 32 + 8 * g);
#pragma unroll
        for (int tk = 0; tk < 4; ++tk) {
            rd_f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
            acc = RD_MFMA(a0, lds_frag(SH, tk * 16 + j, 8 * g), acc);
            acc = RD_MFMA(a1, lds_frag(SH, tk * 16 + j, 32 + 8 * g), acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) out[(ti * 16 + 4 * g + r) * 64 + tk * 16 + j] = acc[r];
        }
    }
}

// dH = W^T dZ with W^T fragments wt[to][kb][lane]; dZ in the C layout
template <class WP>
__device__ __forceinline__ void back64(const WP& wt, const float (&dz)[4][4][4], float (&dh)[4][4][4], int lane)
{
    rd_bf16x8 bop[4][2];
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
        bop[tj][0] = rd_mlp_bop(dz, tj, 0);
        bop[tj][1] = rd_mlp_bop(dz, tj, 1);
    }
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        const rd_bf16x8 a0 = wt[to][0][lane], a1 = wt[to][1][lane];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
            rd_f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
            acc = RD_MFMA(a0, bop[tj][0], acc);
            acc = RD_MFMA(a1, bop[tj][1], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) dh[tj][to][r] = acc[r];
        }
    }
}

// bias gradient: sum over the 64 samples of d[.][ti][r]; lanes j == 0 write
__device__ __forceinline__ void bgrad(const float (&d)[4][4][4], float* __restrict__ out, int lane)
{
    const int j = lane & 15, g = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float v = d[0][ti][r] + d[1][ti][r] + d[2][ti][r] + d[3][ti][r];
            v = rd_sum16(v);
            if (j == 0) out[ti * 16 + 4 * g + r] = v;
        }
}

// Samples: idx = sequential ? c : Philox(seed, step, c) mod m.  Samples at
// c >= batch (padding of the last wave) contribute nothing.
__global__ __launch_bounds__(256) void k_mlp_train(const float4* __restrict__ xn_all, const float4* __restrict__ tn_all,
                                                   long m, int batch, int sequential, uint64_t seed,
                                                   const RdMlpPack* __restrict__ pk_g, RdTrainCtl* __restrict__ ctl,
                                                   float* __restrict__ partial)
{
    extern __shared__ __align__(16) unsigned char smem[];
    RdMlpPack& pk = *reinterpret_cast<RdMlpPack*>(smem);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __bf16* SD = reinterpret_cast<__bf16*>(smem + sizeof(RdMlpPack)) + (size_t)wv * 2 * 64 * RD_TLD;
    __bf16* SH = SD + 64 * RD_TLD;
    {
        const uint4* src = reinterpret_cast<const uint4*>(pk_g);
        uint4* dst = reinterpret_cast<uint4*>(smem);
        for (int i = threadIdx.x; i < (int)(sizeof(RdMlpPack) / 16); i += 256) dst[i] = src[i];
    }
    __syncthreads();
    const unsigned long long step = ctl->step;
    const int w = blockIdx.x * 4 + wv;
    const int j = lane & 15, g = lane >> 4;
    float* out = partial + (size_t)w * RD_MLP_ROW;

    float xn[4][3], tn[4][3];
    bool live[4];
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
        const int c = w * 64 + tj * 16 + j;
        live[tj] = c < batch;
        long idx = c;
        if (!sequential) {
            uint32_t q[4] = { (uint32_t)c, (uint32_t)step, (uint32_t)(step >> 32), 0x5452414eu };
            philox4x32_10(q, (uint32_t)seed, (uint32_t)(seed >> 32));
            idx = (long)(((uint64_t)q[1] << 32 | q[0]) % (uint64_t)m);
        }
        if (!live[tj]) idx = 0;
        float4 a = xn_all[idx], b = tn_all[idx];
        xn[tj][0] = a.x; xn[tj][1] = a.y; xn[tj][2] = a.z;
        tn[tj][0] = b.x; tn[tj][1] = b.y; tn[tj][2] = b.z;
    }

    // ---- forward ----
    float z1[4][4][4], z2[4][4][4], z3[4][4][4], h[4][4][4];
    rd_mlp_layer1(pk.w1, pk.b1, xn, z1, lane);
    rd_mlp_silu_all(z1, h);
    rd_mlp_layer64([&](int to, int kb) { return pk.w2[to][kb][lane]; }, pk.b2, h, z2, lane);
    rd_mlp_silu_all(z2, h);
    rd_mlp_layer64([&](int to, int kb) { return pk.w3[to][kb][lane]; }, pk.b3, h, z3, lane);
    rd_mlp_silu_all(z3, h);                  // h = h3
    float y[4][3];
    rd_mlp_layer4([&](int kb) { return pk.w4[kb][lane]; }, pk.b4, h, y, lane);

    // ---- loss and dy (lanes of group 0 hold the outputs) ----
    const float sc2 = 2.0f / (float)batch;
    float dy[4][3];
    float lsum = 0.0f;
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float r = (g == 0 && live[tj]) ? y[tj][c] - tn[tj][c] : 0.0f;
            lsum += r * r;
            dy[tj][c] = sc2 * r;
        }

    // dW4 = dy h3^T: stage dy (rows 0..2) and h3
    stage_act(SH, h, lane);
    if (g == 0) {
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
            for (int c = 0; c < 3; ++c) SD[c * RD_TLD + tj * 16 + j] = (__bf16)dy[tj][c];
    }
    __syncthreads();
    {
        rd_bf16x8 a0, a1;
        if (j < 3) { a0 = lds_frag(SD, j, 8 * g); a1 = lds_frag(SD, j, 32 + 8 * g); }
        else {
#pragma unroll
            for (int t = 0; t < 8; ++t) { a0[t] = (__bf16)0.0f; a1[t] = (__bf16)0.0f; }
        }
#pragma unroll
        for (int tk = 0; tk < 4; ++tk) {
            rd_f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
            acc = RD_MFMA(a0, lds_frag(SH, tk * 16 + j, 8 * g), acc);
            acc = RD_MFMA(a1, lds_frag(SH, tk * 16 + j, 32 + 8 * g), acc);
            if (g == 0) {
#pragma unroll
                for (int r = 0; r < 3; ++r) out[RD_MLP_W4 + r * 64 + tk * 16 + j] = acc[r];
            }
        }
    }
    // db4 and the loss
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = rd_sum16(dy[0][c] + dy[1][c] + dy[2][c] + dy[3][c]);
        if (lane == 0) out[RD_MLP_B4 + c] = v;
    }
    {
        float l = lsum;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
        if (lane == 0) out[RD_MLP_NPARAM] = l / (float)batch;
    }

    // dh3 = W4^T dy: B operand = dy in lanes of group 0 (k = 0..2)
    float dz[4][4][4], dh[4][4][4];
    {
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
            rd_bf16x8 b;
#pragma unroll
            for (int t = 0; t < 8; ++t) b[t] = (__bf16)((g == 0 && t < 3) ? dy[tj][t < 3 ? t : 0] : 0.0f);
#pragma unroll
            for (int to = 0; to < 4; ++to) {
                rd_f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
                acc = RD_MFMA(pk.w4t[to][lane], b, acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) dh[tj][to][r] = acc[r];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) dz[a][b][r] = dh[a][b][r] * rd_dsilu(z3[a][b][r]);
    bgrad(dz, out + RD_MLP_B3, lane);

    // dW3 = dz3 h2^T
    __syncthreads();
    rd_mlp_silu_all(z2, h);                  // h = h2
    stage_act(SD, dz, lane);
    stage_act(SH, h, lane);
    __syncthreads();
    wgrad64(SD, SH, out + RD_MLP_W3, lane);

    // dh2 = W3^T dz3, dz2
    back64(pk.w3t, dz, dh, lane);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) dz[a][b][r] = dh[a][b][r] * rd_dsilu(z2[a][b][r]);
    bgrad(dz, out + RD_MLP_B2, lane);

    // dW2 = dz2 h1^T
    __syncthreads();
    rd_mlp_silu_all(z1, h);                  // h = h1
    stage_act(SD, dz, lane);
    stage_act(SH, h, lane);
    __syncthreads();
    wgrad64(SD, SH, out + RD_MLP_W2, lane);

    // dh1 = W2^T dz2, dz1 (FP32), dW1 = dz1 xn^T on the VALU
    back64(pk.w2t, dz, dh, lane);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) dz[a][b][r] = dh[a][b][r] * rd_dsilu(z1[a][b][r]);
    bgrad(dz, out + RD_MLP_B1, lane);
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float v = dz[0][ti][r] * xn[0][c];
                v = v + dz[1][ti][r] * xn[1][c];
                v = v + dz[2][ti][r] * xn[2][c];
                v = v + dz[3][ti][r] * xn[3][c];
                v = rd_sum16(v);
                if (j == 0) out[RD_MLP_W1 + 3 * (ti * 16 + 4 * g + r) + c] = v;
            }
}

// write the packed BF16 / FP32 copies of master parameter p
__device__ __forceinline__ void pack_param(RdMlpPack* __restrict__ pk, int p, float w)
{
    const __bf16 wb = (__bf16)w;
    auto put = [&](rd_bf16x8* frag, int t) { reinterpret_cast<__bf16*>(frag)[t] = wb; };
    auto kpos = [](int b, int& kb, int& g, int& t) {
        kb = (b >> 4) >> 1;
        t = 4 * ((b >> 4) & 1) + (b & 3);
        g = (b & 15) >> 2;
    };
    if (p < RD_MLP_B1) {
        pk->w1[p] = (float)wb;
        pk->w1s[p] = (float)wb * RD_MLP_NLOG2E;
        return;
    }
    if (p < RD_MLP_W2) {
        pk->b1[p - RD_MLP_B1] = w;
        pk->b1s[p - RD_MLP_B1] = w * RD_MLP_NLOG2E;
        return;
    }
    if (p < RD_MLP_B2 || (p >= RD_MLP_W3 && p < RD_MLP_B3)) {
        const bool l3 = p >= RD_MLP_W3;
        const int q = p - (l3 ? RD_MLP_W3 : RD_MLP_W2);
        const int a = q >> 6, b = q & 63;       // W[a][b]: output a, input b
        int kb, g, t;
        kpos(b, kb, g, t);
        put(l3 ? &pk->w3[a >> 4][kb][g * 16 + (a & 15)] : &pk->w2[a >> 4][kb][g * 16 + (a & 15)], t);
        kpos(a, kb, g, t);
        put(l3 ? &pk->w3t[b >> 4][kb][g * 16 + (b & 15)] : &pk->w2t[b >> 4][kb][g * 16 + (b & 15)], t);
        return;
    }
    if (p < RD_MLP_W3) {
        pk->b2[p - RD_MLP_B2] = w;
        pk->b2s[p - RD_MLP_B2] = w * RD_MLP_NLOG2E;
        return;
    }
    if (p < RD_MLP_W4) {
        pk->b3[p - RD_MLP_B3] = w;
        pk->b3s[p - RD_MLP_B3] = w * RD_MLP_NLOG2E;
        return;
    }
    if (p < RD_MLP_B4) {
        const int q = p - RD_MLP_W4, o = q >> 6, b = q & 63;
        int kb, g, t;
        kpos(b, kb, g, t);
        put(&pk->w4[kb][g * 16 + o], t);
        put(&pk->w4t[b >> 4][b & 15], o);
        reinterpret_cast<__bf16*>(&pk->w4s[kb][g * 16 + o])[t] = (__bf16)(w * RD_MLP_NLN2);
        return;
    }
    pk->b4[p - RD_MLP_B4] = w;
}

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ master, RdMlpPack* __restrict__ pk)
{
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p < RD_MLP_NPARAM) pack_param(pk, p, master[p]);
}

// Sum the partial gradients in wave order (deterministic), Adam step on the
// FP32 master weights, refresh the packed copies.  Cosine learning rate from
// lr0 to lr1 over total steps.
__global__ __launch_bounds__(256) void k_mlp_adam(const float* __restrict__ partial, int nw,
                                                  float* __restrict__ master, float* __restrict__ mom,
                                                  float* __restrict__ vel, RdMlpPack* __restrict__ pk,
                                                  RdTrainCtl* __restrict__ ctl, float lr0, float lr1,
                                                  long long total, float b1, float b2, float eps,
                                                  float* __restrict__ grad_out, int apply,
                                                  float* __restrict__ loss_hist)
{
    // 32 parameters per workgroup, the per-wave partial rows split over 8
    // slices, the slices added in a fixed order: deterministic, and enough
    // workgroups to keep the loads in flight
    __shared__ float red[8][32];
    const int pl = threadIdx.x & 31, slice = threadIdx.x >> 5;
    const int p = blockIdx.x * 32 + pl;
    float part = 0.0f;
    if (p <= RD_MLP_NPARAM)
        for (int w = slice; w < nw; w += 8) part += partial[(size_t)w * RD_MLP_ROW + p];
    red[slice][pl] = part;
    __syncthreads();
    if (slice != 0 || p > RD_MLP_NPARAM) return;
    float gsum = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) gsum += red[k][pl];
    if (p == RD_MLP_NPARAM) {        // loss column
        ctl->loss = gsum;
        if (loss_hist) loss_hist[ctl->step] = gsum;
        return;
    }
    if (grad_out) grad_out[p] = gsum;
    if (!apply) return;
    const unsigned long long t = ctl->step + 1;
    const double ph = total > 0 ? fmin(1.0, (double)(t - 1) / (double)total) : 0.0;
    const float lr = lr1 + 0.5f * (lr0 - lr1) * (float)(1.0 + cos(3.14159265358979323846 * ph));
    float m = b1 * mom[p] + (1.0f - b1) * gsum;
    float v = b2 * vel[p] + (1.0f - b2) * gsum * gsum;
    mom[p] = m;
    vel[p] = v;
    const float mh = m / (1.0f - (float)pow((double)b1, (double)t));
    const float vh = v / (1.0f - (float)pow((double)b2, (double)t));
    const float w = master[p] - lr * mh / (sqrtf(vh) + eps);
    master[p] = w;
    pack_param(pk, p, w);
}

__global__ void k_mlp_step_inc(RdTrainCtl* ctl) { ctl->step += 1; }

// ===========================================================================
// Inference
// ===========================================================================
// Persistent waves: the forward weights (A fragments, 18 KB) and layer-1
// weights and biases are staged in LDS per workgroup; the compiler keeps the
// fragments a lane uses in registers across tiles (about 210 VGPRs: 2 waves
// per SIMD, which overlap one wave's SiLU work on the VALU with the other's
// MFMAs; the MFMAs write VGPRs directly, built with
// -mllvm -amdgpu-mfma-vgpr-form, so no accumulator-register reads follow
// them).  Samples go through the network one sample tile (16 samples) at a
// time, so only one tile's activations are live, as packed BF16 B operands.
// Lane l holds sample l % 16 of the tile; outputs land in lanes 0..15.
//
// The staged copies are the pack's inference copies: layer 1 and the biases
// carry -log2(e) and W4 carries -ln 2, so every layer yields zs = -log2(e) z
// and the activation zs / (1 + 2^zs) = -log2(e) SiLU(z) feeds the next layer
// unchanged (W2 and W3 see the factor in both their input and their output);
// W4's -ln 2 removes it.  SiLU then costs 2^x, a reciprocal, an add and one
// multiply.  The inference is VALU-bound (two transcendentals per
// activation), so each VALU instruction saved counts.
struct RdMlpFwdLds {
    rd_bf16x8 w2[4][2][64], w3[4][2][64], w4[2][64];
    float w1[192], b1[64], b2[64], b3[64], b4[4];
};

// BLOCK threads stage the weights with compile-time strides (blockDim.x
// would be a load from the dispatch packet ahead of each loop), loads issued
// in two batches ahead of their LDS stores (two global latencies instead of
// one per loop iteration; one batch of everything spills registers)
template <int BLOCK>
__device__ __forceinline__ void mlp_stage(RdMlpFwdLds& L, const RdMlpPack* __restrict__ pk_g)
{
    const uint4* src = reinterpret_cast<const uint4*>(pk_g);      // w2, w3 lead the pack
    uint4* dst = reinterpret_cast<uint4*>(&L);
    constexpr int N23 = (int)(sizeof(rd_bf16x8) * (2 * 4 * 2 * 64) / 16);
    static_assert(N23 % (2 * BLOCK) == 0 && BLOCK >= 192, "staging layout");
    const int t = threadIdx.x;
    const uint4* s4 = reinterpret_cast<const uint4*>(pk_g->w4s);
    uint4* d4 = reinterpret_cast<uint4*>(L.w4);
    // batch 1: the small arrays and half of w2, w3
    uint4 v4 = make_uint4(0u, 0u, 0u, 0u);
    float v1 = 0.0f, vb1 = 0.0f, vb2 = 0.0f, vb3 = 0.0f, vb4 = 0.0f;
    if (t < 2 * 64) v4 = s4[t];
    if (t < 192) v1 = pk_g->w1s[t];
    if (t < 64) { vb1 = pk_g->b1s[t]; vb2 = pk_g->b2s[t]; vb3 = pk_g->b3s[t]; }
    if (t < 4) vb4 = pk_g->b4[t];
    constexpr int KH = N23 / (2 * BLOCK);
    for (int h = 0; h < 2; ++h) {
        uint4 v[KH];
#pragma unroll
        for (int k = 0; k < KH; ++k) v[k] = src[t + (h * KH + k) * BLOCK];
#pragma unroll
        for (int k = 0; k < KH; ++k) dst[t + (h * KH + k) * BLOCK] = v[k];
    }
    if (t < 2 * 64) d4[t] = v4;
    if (t < 192) L.w1[t] = v1;
    if (t < 64) { L.b1[t] = vb1; L.b2[t] = vb2; L.b3[t] = vb3; }
    if (t < 4) L.b4[t] = vb4;
}

// -log2(e) SiLU(z) from zs = -log2(e) z
__device__ __forceinline__ float rd_silu_l2(float zs)
{
    return zs * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(zs));
}

// network output for this lane's sample of the tile (valid in lanes 0..15)
__device__ __forceinline__ float3 mlp_tile(const RdMlpFwdLds& L, const float (&xn)[3])
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    // layer 1 (FP32 FMAs, VALU) -> SiLU -> B operands
    rd_f32x4 z[4];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int f = ti * 16 + 4 * g + r;
            float v = __builtin_fmaf(L.w1[3 * f], xn[0], L.b1[f]);
            v = __builtin_fmaf(L.w1[3 * f + 1], xn[1], v);
            z[ti][r] = __builtin_fmaf(L.w1[3 * f + 2], xn[2], v);
        }
    rd_bf16x8 hb[2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int t = 0; t < 8; ++t) hb[kb][t] = (__bf16)rd_silu_l2(z[2 * kb + t / 4][t & 3]);
    // layers 2 and 3 (MFMA) -> SiLU
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const float* bb = layer == 0 ? L.b2 : L.b3;
#pragma unroll
        for (int to = 0; to < 4; ++to) {
            rd_f32x4 acc;
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = bb[to * 16 + 4 * g + r];
            acc = RD_MFMA(layer == 0 ? L.w2[to][0][lane] : L.w3[to][0][lane], hb[0], acc);
            acc = RD_MFMA(layer == 0 ? L.w2[to][1][lane] : L.w3[to][1][lane], hb[1], acc);
            z[to] = acc;
        }
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int t = 0; t < 8; ++t) hb[kb][t] = (__bf16)rd_silu_l2(z[2 * kb + t / 4][t & 3]);
    }
    // layer 4 (MFMA, rows 0..2 of 16)
    rd_f32x4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = (g == 0 && r < 3) ? L.b4[r] : 0.0f;
    acc = RD_MFMA(L.w4[0][lane], hb[0], acc);
    acc = RD_MFMA(L.w4[1][lane], hb[1], acc);
    return make_float3(acc[0], acc[1], acc[2]);
}

// surrogate pass over sorted slots: gext[s] = gscale * mlp(normalised position)
// for the particles whose type is in the mask (other slots are not written).
// One round of resident waves takes chunks of 64 slots in turn; a wave
// appends the selected slots of its chunk to a ring queue in LDS and runs a
// tile whenever 16 are queued; the last partial tile runs at the end.
#define RD_MLP_RING 128
__global__ __launch_bounds__(256) void k_mlp_slots(const RdMlpPack* __restrict__ pk, const float4* __restrict__ ps,
                                                   int n, RdMlpNorm nm, uint32_t type_mask,
                                                   float4* __restrict__ gext)
{
    __shared__ RdMlpFwdLds L;
    __shared__ uint32_t ring[4][RD_MLP_RING];
    mlp_stage<256>(L, pk);
    __syncthreads();
    const int lane = threadIdx.x & 63, j = lane & 15;
    uint32_t* q = ring[threadIdx.x >> 6];
    auto tile = [&](uint32_t head, int cnt) {            // cnt in 1..16 entries from head
        const bool ok = j < cnt;
        const uint32_t s = ok ? q[(head + j) & (RD_MLP_RING - 1)] : 0u;
        const float4 p = ps[s];
        float xn[3] = { norm_coord(p.x, nm.lo[0], nm.sc[0]), norm_coord(p.y, nm.lo[1], nm.sc[1]),
                        norm_coord(p.z, nm.lo[2], nm.sc[2]) };
        const float3 y = mlp_tile(L, xn);
        if (lane < 16 && ok)
            gext[s] = make_float4(y.x * nm.gscale, y.y * nm.gscale, y.z * nm.gscale, 0.f);
    };
    const long nchunks = ((long)n + 63) / 64;
    const long wpb = blockDim.x >> 6;
    const long stride = (long)gridDim.x * wpb;
    uint32_t head = 0, tail = 0;                         // tail - head <= 15 between chunks
    for (long c = (long)xcd_block(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6); c < nchunks; c += stride) {
        const long s = c * 64 + lane;
        const bool on = s < n && ((type_mask >> w_type(ps[s].w)) & 1u);
        const uint64_t b = __ballot(on);
        if (on) q[(tail + lanes_below(b)) & (RD_MLP_RING - 1)] = (uint32_t)s;
        tail += (uint32_t)__popcll(b);
        wave_lds_sync();
#pragma unroll 1
        while (tail - head >= 16) {
            tile(head, 16);
            head += 16;
        }
        wave_lds_sync();
    }
    if (tail != head) tile(head, (int)(tail - head));
}

// inference at normalised points (float4 xn) -> raw network output (float4)
__global__ __launch_bounds__(256) void k_mlp_points(const RdMlpPack* __restrict__ pk, const float4* __restrict__ xin,
                                                    long m, float4* __restrict__ out)
{
    __shared__ RdMlpFwdLds L;
    mlp_stage<256>(L, pk);
    __syncthreads();
    const int lane = threadIdx.x & 63, j = lane & 15;
    const long ntiles = (m + 15) / 16;
    const long wpb = blockDim.x >> 6;
    const long stride = (long)gridDim.x * wpb;
    for (long t = (long)xcd_block(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6); t < ntiles; t += stride) {
        const long i = t * 16 + j;
        const float4 p = i < m ? xin[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        float xn[3] = { p.x, p.y, p.z };
        const float3 y = mlp_tile(L, xn);
        if (lane < 16 && i < m) out[i] = make_float4(y.x, y.y, y.z, 0.f);
    }
}

// ===========================================================================
// Host side
// ===========================================================================
static int sset_err(rd_system* h, int code, const std::string& m)
{
    h->err = m;
    return code;
}

#define SHIPCK(h, x)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            return sset_err((h), RD_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)

static inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }
static inline size_t pieces_lds(const rd_system* h) { return sizeof(RdPiece) * (size_t)(h->npieces > 0 ? h->npieces : 1); }

void rd_host_free_surrogate(rd_system* h)
{
    void* ptrs[] = { h->d_pieces, h->d_gext, h->d_master, h->d_mom, h->d_vel, h->d_pack, h->d_ctl, h->d_partial,
                     h->d_xn, h->d_tg, h->d_gscale, h->d_sbsum, h->d_loss_hist };
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h->d_pieces = nullptr; h->d_gext = nullptr; h->d_master = h->d_mom = h->d_vel = nullptr;
    h->d_pack = nullptr; h->d_ctl = nullptr; h->d_partial = nullptr; h->d_xn = h->d_tg = nullptr;
    h->d_gscale = nullptr; h->d_sbsum = nullptr; h->d_loss_hist = nullptr;
    h->gext_cap = 0; h->partial_rows = 0; h->nsamples = 0; h->loss_hist_cap = 0;
}

// persistent grid of k_mlp_points: one tile of 16 points per wave and round
static inline unsigned mlp_point_blocks(long m) { return (unsigned)std::min(((m + 15) / 16 + 3) / 4, 2048L); }

// external pass of the step (sorted slots -> gext)
void rd_host_launch_ext(rd_system* h)
{
    if (h->n <= 0) return;
    const uint32_t all_types = h->P.nt >= 32 ? 0xffffffffu : (1u << h->P.nt) - 1u;
    if (h->ext_mode == RD_EXT_SURROGATE) {
        unsigned blocks = std::min(h->ext_blocks[1] > 0 ? (unsigned)h->ext_blocks[1] : 2048u, nb256(h->n));
        hipLaunchKernelGGL(k_mlp_slots, dim3(blocks), dim3(256), 0, h->stream, (const RdMlpPack*)h->d_pack,
                           (const float4*)h->d_ps, h->n, h->mlp_norm, h->comp_mask, h->d_gext);
    } else if ((h->comp_mask & all_types) == all_types) {
        hipLaunchKernelGGL(h->comp_sum ? k_composite_all<true> : k_composite_all<false>, dim3(nb256(h->n)), dim3(256),
                           pieces_lds(h), h->stream, h->d_ps, h->n, (const RdPiece*)h->d_pieces, h->npieces,
                           h->d_gext);
    } else {
        unsigned blocks = std::min(h->ext_blocks[0] > 0 ? (unsigned)h->ext_blocks[0] : 2048u, nb256(h->n));
        hipLaunchKernelGGL(h->comp_sum ? k_composite_slots<true> : k_composite_slots<false>, dim3(blocks), dim3(256),
                           pieces_lds(h), h->stream, h->d_ps, h->n, (const RdPiece*)h->d_pieces, h->npieces,
                           h->comp_mask, h->d_gext);
    }
}

// gext buffer sized for the current particle count; resident workgroups of
// the slot passes for the current piece count
int rd_host_ext_capacity(rd_system* h)
{
    if (!h->ext_active()) return RD_OK;
    int per = 0;
    SHIPCK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_composite_slots<false>, 256,
                                                           pieces_lds(h)));
    h->ext_blocks[0] = std::max(per, 1) * h->ncu;
    SHIPCK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_mlp_slots, 256, 0));
    h->ext_blocks[1] = std::max(per, 1) * h->ncu;
    int need = h->n > 0 ? h->n : 1;
    if (need > h->gext_cap) {
        if (h->d_gext) (void)hipFree(h->d_gext);
        h->d_gext = nullptr;
        SHIPCK(h, hipMalloc(&h->d_gext, sizeof(float4) * need));
        h->gext_cap = need;
    }
    return RD_OK;
}

static void to_piece(const rd_pot1& q, RdPiece& p)
{
    memset(&p, 0, sizeof(p));
    double nn = std::sqrt((double)q.normal[0] * q.normal[0] + (double)q.normal[1] * q.normal[1] +
                          (double)q.normal[2] * q.normal[2]);
    for (int a = 0; a < 3; ++a) {
        p.o[a] = q.origin[a];
        p.n[a] = nn > 0 ? (float)(q.normal[a] / nn) : (a == 2 ? 1.0f : 0.0f);
    }
    p.R = q.radius;
    p.hh = q.type == RD_POT1_CYLINDER ? 0.5f * q.height : 0.0f;
    p.k = q.force_const;
    p.zaxis = (p.n[0] == 0.0f && p.n[1] == 0.0f && p.n[2] == 1.0f) ? 1 : 0;
}

extern "C" int rd_set_composite(rd_system* h, int n_pieces, const rd_pot1* pieces, uint32_t type_mask)
{
    if (!h || n_pieces < 0 || (n_pieces > 0 && !pieces)) return RD_EINVAL;
    if (n_pieces > RD_PIECES_MAX) return sset_err(h, RD_EINVAL, "too many composite pieces");
    static bool attr = false;
    if (!attr) {
        const int bytes = (int)(sizeof(RdPiece) * RD_PIECES_MAX);
        const void* fns[] = { (const void*)k_composite_slots<false>, (const void*)k_composite_slots<true>,
                              (const void*)k_composite_all<false>, (const void*)k_composite_all<true>,
                              (const void*)k_composite_points<false>, (const void*)k_composite_points<true>,
                              (const void*)k_sur_data<false>, (const void*)k_sur_data<true> };
        for (const void* f : fns) SHIPCK(h, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        attr = true;
    }
    for (int i = 0; i < n_pieces; ++i) {
        const rd_pot1& q = pieces[i];
        if ((q.type != RD_POT1_DISK && q.type != RD_POT1_CYLINDER) || q.subtype != RD_SUB_ATTRACTIVE)
            return sset_err(h, RD_EINVAL, "composite pieces are attractive disks or cylinders");
    }
    SHIPCK(h, hipSetDevice(h->dev));
    std::vector<RdPiece> pc(n_pieces);
    for (int i = 0; i < n_pieces; ++i) to_piece(pieces[i], pc[i]);
    if (h->d_pieces) (void)hipFree(h->d_pieces);
    h->d_pieces = nullptr;
    if (n_pieces > 0) {
        SHIPCK(h, hipMalloc(&h->d_pieces, sizeof(RdPiece) * n_pieces));
        SHIPCK(h, hipMemcpy(h->d_pieces, pc.data(), sizeof(RdPiece) * n_pieces, hipMemcpyHostToDevice));
    }
    h->npieces = n_pieces;
    h->comp_mask = type_mask;
    h->comp.assign(pieces, pieces + n_pieces);
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_set_composite_combine(rd_system* h, int combine)
{
    if (!h || (combine != RD_COMBINE_NEAREST && combine != RD_COMBINE_SUM)) return RD_EINVAL;
    h->comp_sum = combine == RD_COMBINE_SUM;
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_set_external_mode(rd_system* h, int mode)
{
    if (!h || mode < RD_EXT_EXACT || mode > RD_EXT_SURROGATE) return RD_EINVAL;
    if (mode == RD_EXT_SURROGATE && !h->sur_ready) return sset_err(h, RD_ESTATE, "the surrogate has no weights");
    h->ext_mode = mode;
    h->dirty = true;
    return RD_OK;
}

// exact composite at host points (test and data checks)
extern "C" int rd_composite_eval(rd_system* h, long m, const float* xyz, float* grad, float* energy)
{
    if (!h || m < 0 || (m > 0 && !xyz)) return RD_EINVAL;
    if (m == 0) return RD_OK;
    SHIPCK(h, hipSetDevice(h->dev));
    std::vector<float4> x(m), o(m);
    for (long i = 0; i < m; ++i) x[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.f);
    float4 *dx = nullptr, *dout = nullptr;
    SHIPCK(h, hipMalloc(&dx, sizeof(float4) * m));
    hipError_t e = hipMalloc(&dout, sizeof(float4) * m);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x.data(), sizeof(float4) * m, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(h->comp_sum ? k_composite_points<true> : k_composite_points<false>, dim3(nb256(m)),
                           dim3(256), pieces_lds(h), h->stream, dx, m, (const RdPiece*)h->d_pieces, h->npieces,
                           dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(o.data(), dout, sizeof(float4) * m, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dx);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return sset_err(h, RD_EHIP, std::string("composite eval: ") + hipGetErrorString(e));
    for (long i = 0; i < m; ++i) {
        if (grad) { grad[3 * i] = o[i].x; grad[3 * i + 1] = o[i].y; grad[3 * i + 2] = o[i].z; }
        if (energy) energy[i] = o[i].w;
    }
    return RD_OK;
}

static int sur_alloc(rd_system* h)
{
    if (h->d_master) return RD_OK;
    SHIPCK(h, hipSetDevice(h->dev));
    SHIPCK(h, hipMalloc(&h->d_master, sizeof(float) * RD_MLP_NPARAM));
    SHIPCK(h, hipMalloc(&h->d_mom, sizeof(float) * RD_MLP_NPARAM));
    SHIPCK(h, hipMalloc(&h->d_vel, sizeof(float) * RD_MLP_NPARAM));
    SHIPCK(h, hipMalloc(&h->d_pack, sizeof(RdMlpPack)));
    SHIPCK(h, hipMalloc(&h->d_ctl, sizeof(RdTrainCtl)));
    SHIPCK(h, hipMalloc(&h->d_gscale, sizeof(float)));
    SHIPCK(h, hipMemset(h->d_pack, 0, sizeof(RdMlpPack)));    // zero padding rows of W4 / W4^T
    SHIPCK(h, hipMemset(h->d_ctl, 0, sizeof(RdTrainCtl)));
    if (!h->norm_set) {
        // default domain: the (first) box
        const double* lo = h->reps.empty() ? nullptr : h->reps[0].lo;
        for (int d = 0; d < 3; ++d) {
            double l = lo ? h->reps[0].lo[d] : h->prm.lattice_bounds[2 * d];
            double u = lo ? h->reps[0].hi[d] : h->prm.lattice_bounds[2 * d + 1];
            h->mlp_norm.lo[d] = (float)l;
            h->mlp_norm.sc[d] = (float)(2.0 / (u - l));
            h->norm_hi[d] = u;
        }
        h->mlp_norm.gscale = 1.0f;
    }
    return RD_OK;
}

static int sur_upload_master(rd_system* h, const std::vector<float>& w)
{
    SHIPCK(h, hipMemcpy(h->d_master, w.data(), sizeof(float) * RD_MLP_NPARAM, hipMemcpyHostToDevice));
    SHIPCK(h, hipMemset(h->d_mom, 0, sizeof(float) * RD_MLP_NPARAM));
    SHIPCK(h, hipMemset(h->d_vel, 0, sizeof(float) * RD_MLP_NPARAM));
    SHIPCK(h, hipMemset(h->d_ctl, 0, sizeof(RdTrainCtl)));
    hipLaunchKernelGGL(k_mlp_pack, dim3(nb256(RD_MLP_NPARAM)), dim3(256), 0, h->stream, h->d_master,
                       (RdMlpPack*)h->d_pack);
    SHIPCK(h, hipGetLastError());
    SHIPCK(h, hipStreamSynchronize(h->stream));
    h->sur_ready = true;
    return RD_OK;
}

// weights: uniform(-a, a) with a = sqrt(6 / fan_in) (He-uniform for SiLU), biases 0;
// host generator: splitmix64 of the seed
extern "C" int rd_surrogate_init(rd_system* h, uint64_t seed)
{
    if (!h) return RD_EINVAL;
    int rc = sur_alloc(h);
    if (rc) return rc;
    std::vector<float> w(RD_MLP_NPARAM, 0.0f);
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int count, int fan_in) {
        double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; ++i) w[off + i] = (float)((2.0 * next() - 1.0) * a);
    };
    fill(RD_MLP_W1, 192, 3);
    fill(RD_MLP_W2, 4096, 64);
    fill(RD_MLP_W3, 4096, 64);
    fill(RD_MLP_W4, 192, 64);
    return sur_upload_master(h, w);
}

extern "C" int rd_surrogate_set_weights(rd_system* h, const float* w)
{
    if (!h || !w) return RD_EINVAL;
    int rc = sur_alloc(h);
    if (rc) return rc;
    std::vector<float> v(w, w + RD_MLP_NPARAM);
    return sur_upload_master(h, v);
}

extern "C" int rd_surrogate_get_weights(rd_system* h, float* w)
{
    if (!h || !w) return RD_EINVAL;
    if (!h->d_master) return sset_err(h, RD_ESTATE, "no surrogate");
    SHIPCK(h, hipMemcpy(w, h->d_master, sizeof(float) * RD_MLP_NPARAM, hipMemcpyDeviceToHost));
    return RD_OK;
}

extern "C" int rd_surrogate_set_norm(rd_system* h, const double* lo, const double* hi, double gscale)
{
    if (!h || !lo || !hi || !(gscale > 0)) return RD_EINVAL;
    for (int d = 0; d < 3; ++d) {
        if (!(hi[d] > lo[d])) return RD_EINVAL;
        h->mlp_norm.lo[d] = (float)lo[d];
        h->mlp_norm.sc[d] = (float)(2.0 / (hi[d] - lo[d]));
        h->norm_hi[d] = hi[d];
    }
    h->mlp_norm.gscale = (float)gscale;
    h->norm_set = true;
    h->dirty = true;
    return RD_OK;
}

extern "C" int rd_surrogate_get_norm(rd_system* h, double* lo, double* hi, double* gscale)
{
    if (!h) return RD_EINVAL;
    for (int d = 0; d < 3; ++d) {
        if (lo) lo[d] = h->mlp_norm.lo[d];
        if (hi) hi[d] = h->norm_hi[d];
    }
    if (gscale) *gscale = h->mlp_norm.gscale;
    return RD_OK;
}

// training set: m uniform positions in the domain, exact composite gradients
// normalised by their RMS component (which becomes the output scale)
extern "C" int rd_surrogate_make_data(rd_system* h, long m, uint64_t seed, double* gscale)
{
    if (!h || m < 64) return RD_EINVAL;
    if (h->npieces <= 0) return sset_err(h, RD_ESTATE, "set the composite first");
    int rc = sur_alloc(h);
    if (rc) return rc;
    if (m > h->nsamples) {
        if (h->d_xn) (void)hipFree(h->d_xn);
        if (h->d_tg) (void)hipFree(h->d_tg);
        if (h->d_sbsum) (void)hipFree(h->d_sbsum);
        h->d_xn = h->d_tg = nullptr; h->d_sbsum = nullptr;
        SHIPCK(h, hipMalloc(&h->d_xn, sizeof(float4) * m));
        SHIPCK(h, hipMalloc(&h->d_tg, sizeof(float4) * m));
        SHIPCK(h, hipMalloc(&h->d_sbsum, sizeof(double) * nb256(m)));
    }
    h->nsamples = m;
    const RdMlpNorm& nm = h->mlp_norm;
    float span[3];
    for (int d = 0; d < 3; ++d) span[d] = (float)(h->norm_hi[d] - nm.lo[d]);
    hipLaunchKernelGGL(h->comp_sum ? k_sur_data<true> : k_sur_data<false>, dim3(nb256(m)), dim3(256), pieces_lds(h),
                       h->stream, m, seed, nm, span[0], span[1], span[2], (const RdPiece*)h->d_pieces, h->npieces,
                       h->d_xn, h->d_tg, h->d_sbsum);
    hipLaunchKernelGGL(k_sur_scale_sum, dim3(1), dim3(256), 0, h->stream, h->d_sbsum, (int)nb256(m), m, h->d_gscale);
    hipLaunchKernelGGL(k_sur_normalise, dim3(nb256(m)), dim3(256), 0, h->stream, h->d_tg, m, h->d_gscale);
    SHIPCK(h, hipGetLastError());
    float gs = 1.0f;
    SHIPCK(h, hipMemcpyAsync(&gs, h->d_gscale, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    SHIPCK(h, hipStreamSynchronize(h->stream));
    h->mlp_norm.gscale = gs;
    if (gscale) *gscale = gs;
    return RD_OK;
}

static size_t train_lds() { return sizeof(RdMlpPack) + (size_t)4 * 2 * 64 * RD_TLD * sizeof(__bf16); }

static int ensure_partial(rd_system* h, int batch)
{
    size_t rows = (size_t)((batch + 255) / 256) * 4;
    if (rows > h->partial_rows) {
        if (h->d_partial) (void)hipFree(h->d_partial);
        h->d_partial = nullptr;
        SHIPCK(h, hipMalloc(&h->d_partial, sizeof(float) * RD_MLP_ROW * rows));
        h->partial_rows = rows;
    }
    static bool attr = false;
    if (!attr) {
        SHIPCK(h, hipFuncSetAttribute((const void*)k_mlp_train, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)train_lds()));
        attr = true;
    }
    return RD_OK;
}

static void launch_train_step(rd_system* h, const float4* xn, const float4* tn, long m, int batch, int sequential,
                              uint64_t seed, const rd_train_params& tp, float* grad_out, int apply, float* loss_hist)
{
    const int blocks = (batch + 255) / 256;
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), train_lds(), h->stream, xn, tn, m, batch, sequential,
                       seed, (const RdMlpPack*)h->d_pack, (RdTrainCtl*)h->d_ctl, h->d_partial);
    hipLaunchKernelGGL(k_mlp_adam, dim3((RD_MLP_NPARAM + 1 + 31) / 32), dim3(256), 0, h->stream, h->d_partial, blocks * 4,
                       h->d_master, h->d_mom, h->d_vel, (RdMlpPack*)h->d_pack, (RdTrainCtl*)h->d_ctl, tp.lr, tp.lr_min,
                       (long long)tp.steps, tp.beta1, tp.beta2, tp.eps, grad_out, apply, loss_hist);
    if (apply) hipLaunchKernelGGL(k_mlp_step_inc, dim3(1), dim3(1), 0, h->stream, (RdTrainCtl*)h->d_ctl);
}

// loss and gradient of one batch (no update): the gradient check entry
extern "C" int rd_surrogate_loss_grad(rd_system* h, long m, const float* xn, const float* tn, double* loss, float* grad)
{
    if (!h || m < 1 || !xn || !tn) return RD_EINVAL;
    if (!h->sur_ready) return sset_err(h, RD_ESTATE, "the surrogate has no weights");
    if (m > (1 << 22)) return sset_err(h, RD_EINVAL, "batch too large");
    int rc = ensure_partial(h, (int)m);
    if (rc) return rc;
    std::vector<float4> hx(m), ht(m);
    for (long i = 0; i < m; ++i) {
        hx[i] = make_float4(xn[3 * i], xn[3 * i + 1], xn[3 * i + 2], 0.f);
        ht[i] = make_float4(tn[3 * i], tn[3 * i + 1], tn[3 * i + 2], 0.f);
    }
    float4 *dx = nullptr, *dt = nullptr;
    float* dg = nullptr;
    SHIPCK(h, hipMalloc(&dx, sizeof(float4) * m));
    SHIPCK(h, hipMalloc(&dt, sizeof(float4) * m));
    SHIPCK(h, hipMalloc(&dg, sizeof(float) * RD_MLP_NPARAM));
    SHIPCK(h, hipMemcpyAsync(dx, hx.data(), sizeof(float4) * m, hipMemcpyHostToDevice, h->stream));
    SHIPCK(h, hipMemcpyAsync(dt, ht.data(), sizeof(float4) * m, hipMemcpyHostToDevice, h->stream));
    rd_train_params tp{};
    tp.steps = 1;
    launch_train_step(h, dx, dt, m, (int)m, 1, 0, tp, dg, 0, nullptr);
    hipError_t e = hipGetLastError();
    RdTrainCtl ctl{};
    if (e == hipSuccess && grad) e = hipMemcpyAsync(grad, dg, sizeof(float) * RD_MLP_NPARAM, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&ctl, h->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(dg);
    if (e != hipSuccess) return sset_err(h, RD_EHIP, std::string("loss_grad: ") + hipGetErrorString(e));
    if (loss) *loss = ctl.loss;
    return RD_OK;
}

// Adam on the training set: tp.steps steps of tp.batch samples drawn with
// replacement (Philox of seed and step); loss per step into loss_curve
extern "C" int rd_surrogate_train(rd_system* h, const rd_train_params* tp, float* loss_curve)
{
    if (!h || !tp || tp->steps < 1 || tp->batch < 64 || tp->batch > (1 << 20)) return RD_EINVAL;
    if (!h->sur_ready) return sset_err(h, RD_ESTATE, "the surrogate has no weights");
    if (h->nsamples < 64) return sset_err(h, RD_ESTATE, "no training data");
    int rc = ensure_partial(h, tp->batch);
    if (rc) return rc;
    if ((size_t)tp->steps > h->loss_hist_cap) {
        if (h->d_loss_hist) (void)hipFree(h->d_loss_hist);
        h->d_loss_hist = nullptr;
        SHIPCK(h, hipMalloc(&h->d_loss_hist, sizeof(float) * tp->steps));
        h->loss_hist_cap = tp->steps;
    }
    SHIPCK(h, hipMemset(h->d_ctl, 0, sizeof(RdTrainCtl)));
    SHIPCK(h, hipMemset(h->d_mom, 0, sizeof(float) * RD_MLP_NPARAM));
    SHIPCK(h, hipMemset(h->d_vel, 0, sizeof(float) * RD_MLP_NPARAM));
    // capture a block of steps once and replay it
    const int chunk = tp->steps >= 64 ? 64 : tp->steps;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    SHIPCK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk; ++i)
        launch_train_step(h, h->d_xn, h->d_tg, h->nsamples, tp->batch, 0, tp->seed, *tp, nullptr, 1, h->d_loss_hist);
    SHIPCK(h, hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    int done = 0;
    while (e == hipSuccess && done + chunk <= tp->steps) {
        e = hipGraphLaunch(ge, h->stream);
        done += chunk;
    }
    for (; e == hipSuccess && done < tp->steps; ++done) {
        launch_train_step(h, h->d_xn, h->d_tg, h->nsamples, tp->batch, 0, tp->seed, *tp, nullptr, 1, h->d_loss_hist);
        e = hipGetLastError();
    }
    if (e == hipSuccess && loss_curve)
        e = hipMemcpyAsync(loss_curve, h->d_loss_hist, sizeof(float) * tp->steps, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) return sset_err(h, RD_EHIP, std::string("train: ") + hipGetErrorString(e));
    h->dirty = true;     // new packed weights: the captured step graphs stay valid (same buffers)
    return RD_OK;
}

// raw network outputs at normalised inputs
extern "C" int rd_surrogate_predict(rd_system* h, long m, const float* xn, float* y)
{
    if (!h || m < 0 || (m > 0 && (!xn || !y))) return RD_EINVAL;
    if (!h->sur_ready) return sset_err(h, RD_ESTATE, "the surrogate has no weights");
    if (m == 0) return RD_OK;
    std::vector<float4> hx(m), ho(m);
    for (long i = 0; i < m; ++i) hx[i] = make_float4(xn[3 * i], xn[3 * i + 1], xn[3 * i + 2], 0.f);
    float4 *dx = nullptr, *dout = nullptr;
    SHIPCK(h, hipMalloc(&dx, sizeof(float4) * m));
    hipError_t e = hipMalloc(&dout, sizeof(float4) * m);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, hx.data(), sizeof(float4) * m, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_points, dim3(mlp_point_blocks(m)), dim3(256), 0, h->stream,
                           (const RdMlpPack*)h->d_pack, (const float4*)dx, m, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ho.data(), dout, sizeof(float4) * m, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dx);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return sset_err(h, RD_EHIP, std::string("predict: ") + hipGetErrorString(e));
    for (long i = 0; i < m; ++i) { y[3 * i] = ho[i].x; y[3 * i + 1] = ho[i].y; y[3 * i + 2] = ho[i].z; }
    return RD_OK;
}

// relative RMSE of the surrogate gradient against the exact composite on m
// fresh uniform samples: sqrt(sum |g_mlp - g|^2 / sum |g|^2)
extern "C" int rd_surrogate_validate(rd_system* h, long m, uint64_t seed, double* rel_rmse)
{
    if (!h || m < 64 || !rel_rmse) return RD_EINVAL;
    if (!h->sur_ready || h->npieces <= 0) return sset_err(h, RD_ESTATE, "composite and surrogate required");
    float4 *dx = nullptr, *dt = nullptr, *dy = nullptr;
    double* bs = nullptr;
    float* gs = nullptr;
    SHIPCK(h, hipMalloc(&dx, sizeof(float4) * m));
    SHIPCK(h, hipMalloc(&dt, sizeof(float4) * m));
    SHIPCK(h, hipMalloc(&dy, sizeof(float4) * m));
    SHIPCK(h, hipMalloc(&bs, sizeof(double) * nb256(m)));
    const RdMlpNorm& nm = h->mlp_norm;
    float span[3];
    for (int d = 0; d < 3; ++d) span[d] = (float)(h->norm_hi[d] - nm.lo[d]);
    hipLaunchKernelGGL(h->comp_sum ? k_sur_data<true> : k_sur_data<false>, dim3(nb256(m)), dim3(256), pieces_lds(h),
                       h->stream, m, seed, nm, span[0], span[1], span[2], (const RdPiece*)h->d_pieces, h->npieces, dx,
                       dt, bs);
    hipLaunchKernelGGL(k_mlp_points, dim3(mlp_point_blocks(m)), dim3(256), 0, h->stream, (const RdMlpPack*)h->d_pack,
                       (const float4*)dx, m, dy);
    hipError_t e = hipGetLastError();
    std::vector<float4> t(m), y(m);
    if (e == hipSuccess) e = hipMemcpyAsync(t.data(), dt, sizeof(float4) * m, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(y.data(), dy, sizeof(float4) * m, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dx); (void)hipFree(dt); (void)hipFree(dy); (void)hipFree(bs);
    (void)gs;
    if (e != hipSuccess) return sset_err(h, RD_EHIP, std::string("validate: ") + hipGetErrorString(e));
    double num = 0.0, den = 0.0;
    const double sc = nm.gscale;
    for (long i = 0; i < m; ++i) {
        double ex = t[i].x, ey = t[i].y, ez = t[i].z;
        double dxv = y[i].x * sc - ex, dyv = y[i].y * sc - ey, dzv = y[i].z * sc - ez;
        num += dxv * dxv + dyv * dyv + dzv * dzv;
        den += ex * ex + ey * ey + ez * ez;
    }
    *rel_rmse = den > 0 ? std::sqrt(num / den) : 0.0;
    return RD_OK;
}
