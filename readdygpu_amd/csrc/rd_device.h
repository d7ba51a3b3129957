// rd_device.h -- device-side building blocks of the BD step (gfx950).
//
// Everything here is compiled with -ffp-contract=off: FP32 expressions are
// evaluated exactly in source order, so cell ids, pair tests, RDF bins and the
// noise stream are reproducible bit for bit on any IEEE-754 machine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RD_MAXT 16
#define RD_MAXP2 8
#define RD_MAXBP 8      // bond (group) potentials

// order-two potential, device form (CudaOrderTwoPotential, CudaSimulation.cuh:30-37)
struct RdPot2Dev {
    int kind, subtype;
    float k, length, depth, cutoff;
    float wc;      // piecewise well: depth / (length/2)^2
    float pad;
};

// order-one potential, device form (CudaOrderOnePotential, CudaSimulation.cuh:18-28)
struct RdPot1Dev {
    int kind, subtype;
    float k;
    float o[3];
    float n[3];     // unit normal
    float ext[3];
    float hh;       // half height
    float R;
    uint32_t mask;
    float pad;
};

// Geometry of one replica (an independent periodic box).  A plain system is
// one replica.  Replicas occupy consecutive slot ranges [off, off+n) and
// consecutive cell ranges [coff, coff + Lx*Ly*Lz) of the shared arrays, so
// one counting sort over all cells keeps every replica in its own range.
struct RdRep {
    float lo[3], hi[3], Lb[3], half[3];
    float cell;
    int L[3];
    int coff;           // first global cell
    int off;            // first global slot / particle index
    int n;              // particles
    uint32_t stream;    // noise stream (replica id)
};

// one wave of work: 64 consecutive slots of one replica
struct RdWave {
    int rep, s0, s1, pad;
};

// per-system constants, read through one pointer (scalar loads)
struct RdDev {
    int S;              // pair stencil radius in cells
    int nt, n_pot2, n_pot1;
    uint64_t seed;
    float a[RD_MAXT];   // dt*D/(kB T)
    float b[RD_MAXT];   // sqrt(2 D dt)
    float prad[RD_MAXT];
    float rc2[RD_MAXT * RD_MAXT];   // squared pair range (evaluation threshold)
    float r0[RD_MAXT * RD_MAXT];    // collision radius sums
    float rr2[RD_MAXT * RD_MAXT];   // squared reaction radius (0: none)
    uint32_t p2mask[RD_MAXT * RD_MAXT];
    RdPot2Dev pot2[RD_MAXP2];
    const float4* gext;   // external gradient per slot (x,y,z, energy) from a separate pass, or null
    uint32_t ext_mask;    // types that gext applies to (other slots of gext are not written)
    // bonded pairs of group potentials, listed per particle (both directions)
    // by global index; bstart is null when there are none
    int n_bpot;
    RdPot2Dev bpot[RD_MAXBP];
    float bpot_r0[RD_MAXBP];      // fixed rest length, or < 0: collision radius sum of the pair
    const int* bstart;            // [n+1]
    const int2* bonds;            // (partner global index, potential)
    const uint32_t* slot_of;      // global index -> slot of the current sorted array
    uint32_t* gid_n;              // global index of each slot of the next-step array
};

// packed particle word: low 6 bits type, high 26 bits particle index
__device__ __forceinline__ int w_type(float w) { return (int)(__float_as_uint(w) & 63u); }
__device__ __forceinline__ uint32_t w_index(float w) { return __float_as_uint(w) >> 6; }
__device__ __forceinline__ float w_pack(uint32_t index, int type)
{
    return __uint_as_float((index << 6) | (uint32_t)type);
}

// ---------------------------------------------------------------------------
// Lattice: L_d = (int)((hi-lo)/cell); coordinate floor((x-lo)/cell) % L_d
// (CudaSimulation.cu:524-528, 586-588); the partial last cell wraps to 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_coord(float x, float lo, float cell, int L)
{
    float t = (x - lo) / cell;
    if (!(t >= 0.0f)) return 0;
    if (t > 1.0e9f) t = 1.0e9f;
    int i = (int)floorf(t);
    return i < L ? i : i % L;
}

__device__ __forceinline__ float min_image(float d, float Lb, float h)
{
    if (d > h) d -= Lb;
    if (d < -h) d += Lb;
    return d;
}

// ---------------------------------------------------------------------------
// Noise: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller with an in-house
// log and sin/cos built from IEEE-exact operations.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32x32->64 multiply per word gives both halves
        const uint64_t m0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t m1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t lo0 = (uint32_t)m0, hi0 = (uint32_t)(m0 >> 32);
        const uint32_t lo1 = (uint32_t)m1, hi1 = (uint32_t)(m1 >> 32);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

__device__ __forceinline__ float rd_log_unit(float x)
{
    uint32_t b = __float_as_uint(x);
    int e = (int)((b >> 23) & 0xffu) - 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 0.0909090909f;
    p = p * s2 + 0.111111111f;
    p = p * s2 + 0.142857143f;
    p = p * s2 + 0.2f;
    p = p * s2 + 0.333333333f;
    p = p * s2 + 1.0f;
    float lnm = (s + s) * p;
    float fe = (float)e;
    return fe * 0.693145752f + (fe * 1.42860677e-06f + lnm);
}

__device__ __forceinline__ void rd_sincos_turns(float t, float* sn, float* cs)
{
    float t4 = t * 4.0f;
    int q = (int)t4;
    float f = t4 - (float)q;
    bool swap = false;
    if (f > 0.5f) { f = 1.0f - f; swap = true; }
    float ph = f * 1.57079637f;
    float p2 = ph * ph;
    float sp = -2.50521084e-08f;
    sp = sp * p2 + 2.75573192e-06f;
    sp = sp * p2 - 1.98412698e-04f;
    sp = sp * p2 + 8.33333333e-03f;
    sp = sp * p2 - 1.66666667e-01f;
    sp = sp * p2;
    float s0 = ph + ph * sp;
    float cp = 2.08767570e-09f;
    cp = cp * p2 - 2.75573192e-07f;
    cp = cp * p2 + 2.48015873e-05f;
    cp = cp * p2 - 1.38888889e-03f;
    cp = cp * p2 + 4.16666667e-02f;
    cp = cp * p2 - 0.5f;
    float c0 = 1.0f + p2 * cp;
    if (swap) { float tmp = s0; s0 = c0; c0 = tmp; }
    float s, c;
    switch (q & 3) {
    case 0: s = s0; c = c0; break;
    case 1: s = c0; c = -s0; break;
    case 2: s = -s0; c = -c0; break;
    default: s = -c0; c = s0; break;
    }
    *sn = s; *cs = c;
}

// three N(0,1) values for (seed, stream, step, index)
__device__ __forceinline__ float3 rd_normal3(uint64_t seed, uint32_t stream, uint64_t step, uint32_t index)
{
    uint32_t c[4] = { index, (uint32_t)step, (uint32_t)(step >> 32) ^ (stream << 16), 0x52444E53u };
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    float a1 = (float)((c[0] >> 8) + 1u) * 5.96046448e-08f;
    float a2 = (float)(c[1] >> 8) * 5.96046448e-08f;
    float b1 = (float)((c[2] >> 8) + 1u) * 5.96046448e-08f;
    float b2 = (float)(c[3] >> 8) * 5.96046448e-08f;
    float ra = sqrtf(-2.0f * rd_log_unit(a1));
    float rb = sqrtf(-2.0f * rd_log_unit(b1));
    float sa, ca, sb, cb;
    rd_sincos_turns(a2, &sa, &ca);
    rd_sincos_turns(b2, &sb, &cb);
    return make_float3(ra * ca, ra * sa, rb * cb);
}

// ---------------------------------------------------------------------------
// Pair potentials (FP32).  d = x_i - x_j (minimum image).  Returns pre such
// that grad_i U += pre * d; *e receives the pair energy.
// HARMONIC        CudaSimulation.cu:1007-1019
// WEAK_HARMONIC   CudaSimulation.cu:1021-1031
// PIECEWISE_WELL  P2_WeakInteractionPiecewiseHarmonic.java:114-200
// ---------------------------------------------------------------------------
// rinv = 1/r from v_rsq_f32 (forces need no IEEE division)
template <bool ENERGY>
__device__ __forceinline__ float pair_pre_fast(const RdPot2Dev& p, float r0, float r, float rinv, float* e)
{
    float pre = 0.0f;
    if (r > p.cutoff) return 0.0f;
    if (p.kind == 1) {
        bool on = (p.subtype == 1 && r > r0) || (p.subtype == 2 && r < r0) || (p.subtype == 3);
        if (on) {
            pre = p.k * (r - r0) * rinv;
            if (ENERGY) *e += 0.5f * p.k * (r - r0) * (r - r0);
        }
    } else if (p.kind == 2) {
        float ri = p.length + r0;
        if (r < ri && r > r0) pre = p.k * (r - r0) * (ri - r);
        if (ENERGY && r < ri) {
            float a = r0 + ri, b = r0 * ri;
            float s = r > r0 ? r : r0;
            float gs = -s * s * s * s * 0.25f + a * s * s * s * 0.333333343f - b * s * s * 0.5f;
            float gi = -ri * ri * ri * ri * 0.25f + a * ri * ri * ri * 0.333333343f - b * ri * ri * 0.5f;
            *e += -p.k * (gi - gs);
        }
    } else if (p.kind == 3) {
        float l = p.length;
        float hl = 0.5f * l;
        if (r < r0) {
            pre = p.k * (r - r0) * rinv;
            if (ENERGY) *e += 0.5f * p.k * (r - r0) * (r - r0) - p.depth;
        } else if (r < r0 + hl) {
            pre = p.wc * (r - r0) * rinv;
            if (ENERGY) *e += 0.5f * p.wc * (r - r0) * (r - r0) - p.depth;
        } else if (r < r0 + l) {
            pre = -p.wc * (r - r0 - l) * rinv;
            if (ENERGY) *e += -0.5f * p.wc * (r - r0 - l) * (r - r0 - l);
        }
    }
    return pre;
}

// The sum of pair_pre_fast over the potentials of bit mask m (bit p:
// potential p), in potential order, with the same FP32 operations on every
// branch, so the same bits.  The potential index runs uniformly across the
// wave (potentials no lane needs are skipped), so kind and subtype select
// uniform branches, and a lane's conditions become selects instead of a
// per-lane loop that branches on values read per lane.
__device__ __forceinline__ float rd_uniform(float x)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

template <bool ENERGY>
__device__ __forceinline__ float pair_pre_wave(const RdPot2Dev* pot2, int npot, uint32_t m, float r0, float r,
                                               float rinv, float* e)
{
    float pre = 0.0f;
    const float u = r - r0;
    npot = __builtin_amdgcn_readfirstlane(npot);
    for (int pi = 0; pi < npot; ++pi) {
        const bool has = (m >> pi) & 1u;
        if (!__any(has)) continue;
        // the record is the same for every lane: read whole, held in scalar registers
        const RdPot2Dev p = pot2[pi];
        const int kind = __builtin_amdgcn_readfirstlane(p.kind);
        const float k = rd_uniform(p.k), cut = rd_uniform(p.cutoff);
        const bool in = has & !(r > cut);
        float v = 0.0f;
        bool on = false;
        if (kind == 1) {
            const int sub = __builtin_amdgcn_readfirstlane(p.subtype);
            on = in & ((sub == 1 & r > r0) | (sub == 2 & r < r0) | sub == 3);
            v = k * u * rinv;
        } else if (kind == 2) {
            const float ri = rd_uniform(p.length) + r0;
            on = in & (r < ri) & (r > r0);
            v = k * u * (ri - r);
        } else if (kind == 3) {
            const float l = rd_uniform(p.length), wc = rd_uniform(p.wc);
            const float hl = 0.5f * l;
            const bool a = r < r0, b = r < r0 + hl;
            on = in & (r < r0 + l);
            const float c = a ? k : (b ? wc : -wc);
            const float w = (a | b) ? u : u - l;
            v = c * w * rinv;
        }
        pre += on ? v : 0.0f;
        if (ENERGY && has) (void)pair_pre_fast<true>(p, r0, r, rinv, e);
    }
    return pre;
}

// |v| and 1/|v| from v_rsq_f32; (0, 0) for v = 0
__device__ __forceinline__ void fast_len(float v2, float* len, float* inv)
{
    if (v2 > 0.0f) {
        float ri = __builtin_amdgcn_rsqf(v2);
        *inv = ri;
        *len = v2 * ri;
    } else {
        *inv = 0.0f;
        *len = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// External potentials (FP32), gradient of U at x.
// DISK/CYLINDER  CudaSimulation.cu:1049-1172 (any normal; cylinder repulsive
//                pushes particles out, as P1_Cylinder.java:250-270)
// SPHERE         CudaSimulation.cu:1173-1190
// BOX            CudaSimulation.cu:1191-1228 (considerParticleRadius = true)
// ---------------------------------------------------------------------------
template <bool ENERGY>
__device__ __forceinline__ void ext_grad(const RdPot1Dev& p, float prad, float x, float y, float z,
                                         float& gx, float& gy, float& gz, float* e)
{
    if (p.kind == 1 || p.kind == 2) {
        float vx = x - p.o[0], vy = y - p.o[1], vz = z - p.o[2];
        float s = vx * p.n[0] + vy * p.n[1] + vz * p.n[2];
        float qx = vx - s * p.n[0], qy = vy - s * p.n[1], qz = vz - s * p.n[2];
        float r, rinv;
        fast_len(qx * qx + qy * qy + qz * qz, &r, &rinv);
        float as = fabsf(s);
        float sg = s > 0.0f ? 1.0f : (s < 0.0f ? -1.0f : 0.0f);
        if (p.subtype == 1) {
            float gn;
            if (p.kind == 1) {
                gn = p.k * s;
                if (ENERGY) *e += 0.5f * p.k * s * s;
            } else {
                gn = as > p.hh ? p.k * (as - p.hh) * sg : 0.0f;
                if (ENERGY && as > p.hh) *e += 0.5f * p.k * (as - p.hh) * (as - p.hh);
            }
            gx += gn * p.n[0]; gy += gn * p.n[1]; gz += gn * p.n[2];
            if (r > p.R) {
                float pre = p.k * (r - p.R) * rinv;
                gx += pre * qx; gy += pre * qy; gz += pre * qz;
                if (ENERGY) *e += 0.5f * p.k * (r - p.R) * (r - p.R);
            }
        } else if (p.subtype == 2 && p.kind == 2) {
            if (as < p.hh) {
                float gn = p.k * (as - p.hh) * sg;
                gx += gn * p.n[0]; gy += gn * p.n[1]; gz += gn * p.n[2];
                if (ENERGY) *e += 0.5f * p.k * (as - p.hh) * (as - p.hh);
            }
            if (r < p.R && r > 0.0f) {
                float pre = p.k * (r - p.R) * rinv;
                gx += pre * qx; gy += pre * qy; gz += pre * qz;
                if (ENERGY) *e += 0.5f * p.k * (r - p.R) * (r - p.R);
            }
        }
    } else if (p.kind == 3) {
        float vx = x - p.o[0], vy = y - p.o[1], vz = z - p.o[2];
        float d, dinv;
        fast_len(vx * vx + vy * vy + vz * vz, &d, &dinv);
        bool on = (p.subtype == 1 && d > p.R) || (p.subtype == 2 && d < p.R && d > 0.0f);
        if (on) {
            float pre = p.k * (d - p.R) * dinv;
            gx += pre * vx; gy += pre * vy; gz += pre * vz;
            if (ENERGY) *e += 0.5f * p.k * (d - p.R) * (d - p.R);
        }
    } else if (p.kind == 4) {
        float xs[3] = { x, y, z };
        float gs[3] = { 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (p.subtype == 1) {
                float ub = p.o[d] + p.ext[d] - prad;
                float lb = p.o[d] + prad;
                if (xs[d] > ub) { gs[d] = p.k * (xs[d] - ub); if (ENERGY) *e += 0.5f * p.k * (xs[d] - ub) * (xs[d] - ub); }
                else if (xs[d] < lb) { gs[d] = p.k * (xs[d] - lb); if (ENERGY) *e += 0.5f * p.k * (xs[d] - lb) * (xs[d] - lb); }
            } else if (p.subtype == 2) {
                float ub = p.o[d] + p.ext[d] + prad;
                float lb = p.o[d] - prad;
                if (xs[d] > ub) { gs[d] = p.k * (ub - xs[d]); if (ENERGY) *e -= 0.5f * p.k * (xs[d] - ub) * (xs[d] - ub); }
                else if (xs[d] < lb) { gs[d] = p.k * (lb - xs[d]); if (ENERGY) *e -= 0.5f * p.k * (xs[d] - lb) * (xs[d] - lb); }
            }
        }
        gx += gs[0]; gy += gs[1]; gz += gs[2];
    }
}
