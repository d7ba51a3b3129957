/*
 * rd_oracle.c -- CPU oracle for the Brownian-dynamics step (TEST INFRASTRUCTURE).
 *
 * Only tests, the smoke check and the benchmark's CPU-baseline leg load this
 * library, as the checker.  The product path never calls it.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Which reference lines each function restates is given above each function.
 */
#include "rd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================
 * Noise: Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11) + Box-Muller (Box & Muller 1958).
 * ====================================================================== */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline float as_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t as_u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* Natural log for x in (0, 1]: x = m * 2^e with m in (sqrt(1/2), sqrt(2)],
 * ln m = 2 atanh(s), s = (m-1)/(m+1), atanh series to s^11. */
float orc_log_f32(float x)
{
    uint32_t b = as_u(x);
    int e = (int)((b >> 23) & 0xffu) - 127;
    float m = as_f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 0.0909090909f;                 /* 1/11 */
    p = p * s2 + 0.111111111f;               /* 1/9 */
    p = p * s2 + 0.142857143f;               /* 1/7 */
    p = p * s2 + 0.2f;                       /* 1/5 */
    p = p * s2 + 0.333333333f;               /* 1/3 */
    p = p * s2 + 1.0f;
    float lnm = (s + s) * p;
    float fe = (float)e;
    return fe * 0.693145752f + (fe * 1.42860677e-06f + lnm);   /* ln2 = hi + lo */
}

/* sin and cos of 2*pi*t for t in [0,1): quadrant split, then Taylor
 * polynomials on [0, pi/4]. */
void orc_sincos_turns_f32(float t, float* sn, float* cs)
{
    float t4 = t * 4.0f;
    int q = (int)t4;
    float f = t4 - (float)q;                /* [0,1) quarter turn fraction */
    int swap = 0;
    if (f > 0.5f) { f = 1.0f - f; swap = 1; }
    float ph = f * 1.57079637f;             /* pi/2 */
    float p2 = ph * ph;
    float sp = -2.50521084e-08f;           /* -1/11! */
    sp = sp * p2 + 2.75573192e-06f;         /* 1/9! */
    sp = sp * p2 - 1.98412698e-04f;         /* -1/7! */
    sp = sp * p2 + 8.33333333e-03f;         /* 1/5! */
    sp = sp * p2 - 1.66666667e-01f;         /* -1/3! */
    sp = sp * p2;
    float s0 = ph + ph * sp;
    float cp = 2.08767570e-09f;             /* 1/12! */
    cp = cp * p2 - 2.75573192e-07f;         /* -1/10! */
    cp = cp * p2 + 2.48015873e-05f;         /* 1/8! */
    cp = cp * p2 - 1.38888889e-03f;         /* -1/6! */
    cp = cp * p2 + 4.16666667e-02f;         /* 1/4! */
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

/* Three N(0,1) values for (seed, stream, step, index).
 * counter = {index, step_lo, step_hi ^ (stream << 16), 0x52444E53}, key = seed. */
void orc_normal3(uint64_t seed, uint32_t stream, uint64_t step, uint32_t index, float out[3])
{
    uint32_t ctr[4] = { index, (uint32_t)step, (uint32_t)(step >> 32) ^ (stream << 16), 0x52444E53u };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t u[4];
    orc_philox4x32_10(ctr, key, u);
    float a1 = (float)((u[0] >> 8) + 1u) * 5.96046448e-08f;   /* (0,1] */
    float a2 = (float)(u[1] >> 8) * 5.96046448e-08f;          /* [0,1) */
    float b1 = (float)((u[2] >> 8) + 1u) * 5.96046448e-08f;
    float b2 = (float)(u[3] >> 8) * 5.96046448e-08f;
    float ra = sqrtf(-2.0f * orc_log_f32(a1));
    float rb = sqrtf(-2.0f * orc_log_f32(b1));
    float sa, ca, sb, cb;
    orc_sincos_turns_f32(a2, &sa, &ca);
    orc_sincos_turns_f32(b2, &sb, &cb);
    (void)sb;
    out[0] = ra * ca;
    out[1] = ra * sa;
    out[2] = rb * cb;
}

void orc_noise_fill(uint64_t seed, uint32_t stream, uint64_t step0, int nsteps, int n, float* out)
{
    for (int s = 0; s < nsteps; ++s)
        for (int i = 0; i < n; ++i) {
            float z[3];
            orc_normal3(seed, stream, step0 + (uint64_t)s, (uint32_t)i, z);
            for (int d = 0; d < 3; ++d) out[((size_t)s * 3 + d) * n + i] = z[d];
        }
}

/* ======================================================================
 * Lattice: CudaSimulation.cu:522-528 (L_d = (int)((hi-lo)/maxCutoff), FP32)
 * and :586-588 (field index with % wrap of the partial last cell).
 * ====================================================================== */
void orc_lattice_dims(const orc_system* s, int L[3])
{
    for (int d = 0; d < 3; ++d) {
        float ext = (float)s->hi[d] - (float)s->lo[d];
        int l = (int)(ext / (float)s->cell);
        L[d] = l < 1 ? 1 : l;
    }
}

static inline int cell_coord_f32(float x, float lo, float cell, int L)
{
    float t = (x - lo) / cell;
    if (!(t >= 0.0f)) return 0;            /* below the box or NaN */
    if (t > 1.0e9f) t = 1.0e9f;
    int i = (int)floorf(t);
    return i % L;
}

void orc_cell_index(const orc_system* s, int n, const float* xyz, int* field, int* cxyz)
{
    int L[3];
    orc_lattice_dims(s, L);
    for (int i = 0; i < n; ++i) {
        int c[3];
        for (int d = 0; d < 3; ++d)
            c[d] = cell_coord_f32(xyz[3 * i + d], (float)s->lo[d], (float)s->cell, L[d]);
        if (field) field[i] = c[2] * L[0] * L[1] + c[1] * L[0] + c[0];
        if (cxyz) { cxyz[3 * i] = c[0]; cxyz[3 * i + 1] = c[1]; cxyz[3 * i + 2] = c[2]; }
    }
}

/* ======================================================================
 * Pair potentials (FP64).  d = x_i - x_j (minimum image, decision Q1).
 * HARMONIC             CudaSimulation.cu:1007-1019, P2_Harmonic.java:144-234
 * WEAK_HARMONIC        CudaSimulation.cu:1021-1031 (depth unused; energy is
 *                      the integral of the force so that U(r0+length) = 0)
 * PIECEWISE_WELL       P2_WeakInteractionPiecewiseHarmonic.java:114-200
 * All return 0 beyond `cutoff` (the reference skips r > maxCutoff, .cu:1002).
 * ====================================================================== */
double orc_pair_eval(const orc_pot2* p, double r0, const double d[3], double g[3])
{
    g[0] = g[1] = g[2] = 0.0;
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r == 0.0 || r > p->cutoff) return 0.0;
    double pre = 0.0, e = 0.0;
    if (p->kind == ORC_P2_HARMONIC) {
        int on = (p->subtype == ORC_SUB_ATTRACTIVE && r > r0) ||
                 (p->subtype == ORC_SUB_REPULSIVE && r < r0) ||
                 (p->subtype == ORC_SUB_SPRING);
        if (on) { pre = p->k * (r - r0) / r; e = 0.5 * p->k * (r - r0) * (r - r0); }
    } else if (p->kind == ORC_P2_WEAK_HARMONIC) {
        double ri = p->length + r0;
        double a = r0 + ri, b = r0 * ri;
#define GW(s) (-(s) * (s) * (s) * (s) / 4.0 + a * (s) * (s) * (s) / 3.0 - b * (s) * (s) / 2.0)
        if (r < ri && r > r0) { pre = p->k * (r - r0) * (ri - r); e = -p->k * (GW(ri) - GW(r)); }
        else if (r <= r0) { e = -p->k * (GW(ri) - GW(r0)); }
#undef GW
    } else if (p->kind == ORC_P2_PIECEWISE_WELL) {
        double l = p->length, dd = p->depth, hl = 0.5 * l;
        double c = dd / (hl * hl);
        if (r < r0) { pre = p->k * (r - r0) / r; e = 0.5 * p->k * (r - r0) * (r - r0) - dd; }
        else if (r < r0 + hl) { pre = c * (r - r0) / r; e = 0.5 * c * (r - r0) * (r - r0) - dd; }
        else if (r < r0 + l) { pre = -c * (r - r0 - l) / r; e = -0.5 * c * (r - r0 - l) * (r - r0 - l); }
    }
    g[0] = pre * d[0]; g[1] = pre * d[1]; g[2] = pre * d[2];
    return e;
}

/* ======================================================================
 * External potentials (FP64), gradient of U at x.
 * DISK      CudaSimulation.cu:1049-1116 generalised to any normal (the GPU
 *           reference only handles +x,+y,z normals); repulsive = no-op
 *           without a particle radius, as in the reference.
 * CYLINDER  CudaSimulation.cu:1117-1172; repulsive uses the sign of
 *           P1_Cylinder.java:250-270 (particles pushed out).
 * SPHERE    CudaSimulation.cu:1173-1190
 * BOX/CUBE  CudaSimulation.cu:1191-1228 with considerParticleRadius = true.
 * ====================================================================== */
static void unit3(const double* v, double* u)
{
    double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n == 0.0) { u[0] = 0; u[1] = 0; u[2] = 1; return; }
    u[0] = v[0] / n; u[1] = v[1] / n; u[2] = v[2] / n;
}

double orc_ext_eval(const orc_pot1* p, double prad, const double x[3], double g[3])
{
    double e = 0.0;
    g[0] = g[1] = g[2] = 0.0;
    if (p->kind == ORC_P1_DISK || p->kind == ORC_P1_CYLINDER) {
        double n[3];
        unit3(p->normal, n);
        double v[3] = { x[0] - p->origin[0], x[1] - p->origin[1], x[2] - p->origin[2] };
        double s = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
        double q[3] = { v[0] - s * n[0], v[1] - s * n[1], v[2] - s * n[2] };
        double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        double as = fabs(s), sg = (s > 0) - (s < 0);
        if (p->kind == ORC_P1_DISK) {
            if (p->subtype == ORC_SUB_ATTRACTIVE) {
                for (int d = 0; d < 3; ++d) g[d] += p->k * s * n[d];
                e += 0.5 * p->k * s * s;
                if (r > p->radius) {
                    double pre = p->k * (r - p->radius) / r;
                    for (int d = 0; d < 3; ++d) g[d] += pre * q[d];
                    e += 0.5 * p->k * (r - p->radius) * (r - p->radius);
                }
            }
        } else {
            double hh = 0.5 * p->height;
            if (p->subtype == ORC_SUB_ATTRACTIVE) {
                if (as > hh) {
                    for (int d = 0; d < 3; ++d) g[d] += p->k * (as - hh) * sg * n[d];
                    e += 0.5 * p->k * (as - hh) * (as - hh);
                }
                if (r > p->radius) {
                    double pre = p->k * (r - p->radius) / r;
                    for (int d = 0; d < 3; ++d) g[d] += pre * q[d];
                    e += 0.5 * p->k * (r - p->radius) * (r - p->radius);
                }
            } else if (p->subtype == ORC_SUB_REPULSIVE) {
                if (as < hh) {
                    for (int d = 0; d < 3; ++d) g[d] += p->k * (as - hh) * sg * n[d];
                    e += 0.5 * p->k * (as - hh) * (as - hh);
                }
                if (r < p->radius && r > 0.0) {
                    double pre = p->k * (r - p->radius) / r;
                    for (int d = 0; d < 3; ++d) g[d] += pre * q[d];
                    e += 0.5 * p->k * (r - p->radius) * (r - p->radius);
                }
            }
        }
    } else if (p->kind == ORC_P1_SPHERE) {
        double v[3] = { x[0] - p->origin[0], x[1] - p->origin[1], x[2] - p->origin[2] };
        double dd = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        int on = (p->subtype == ORC_SUB_ATTRACTIVE && dd > p->radius) ||
                 (p->subtype == ORC_SUB_REPULSIVE && dd < p->radius && dd > 0.0);
        if (on) {
            double pre = p->k * (dd - p->radius) / dd;
            for (int d = 0; d < 3; ++d) g[d] += pre * v[d];
            e += 0.5 * p->k * (dd - p->radius) * (dd - p->radius);
        }
    } else if (p->kind == ORC_P1_BOX) {
        for (int d = 0; d < 3; ++d) {
            if (p->subtype == ORC_SUB_ATTRACTIVE) {
                double ub = p->origin[d] + p->extension[d] - prad;
                double lb = p->origin[d] + prad;
                if (x[d] > ub) { g[d] += p->k * (x[d] - ub); e += 0.5 * p->k * (x[d] - ub) * (x[d] - ub); }
                else if (x[d] < lb) { g[d] += p->k * (x[d] - lb); e += 0.5 * p->k * (x[d] - lb) * (x[d] - lb); }
            } else if (p->subtype == ORC_SUB_REPULSIVE) {
                double ub = p->origin[d] + p->extension[d] + prad;
                double lb = p->origin[d] - prad;
                if (x[d] > ub) { g[d] += p->k * (ub - x[d]); e -= 0.5 * p->k * (x[d] - ub) * (x[d] - ub); }
                else if (x[d] < lb) { g[d] += p->k * (lb - x[d]); e -= 0.5 * p->k * (x[d] - lb) * (x[d] - lb); }
            }
        }
    }
    return e;
}

/* P1_Disk.java:112-275 verbatim in method (force convention: returns the
 * Java "gradient", which points along the force), particle radius included. */
double orc_disk_java(const orc_pot1* p, double prad, const double x[3], double f[3])
{
    double n[3];
    unit3(p->normal, n);
    double dco = p->origin[0] * n[0] + p->origin[1] * n[1] + p->origin[2] * n[2];
    double fac = (x[0] * n[0] + x[1] * n[1] + x[2] * n[2]) - dco;
    double pp[3] = { x[0] - n[0] * fac, x[1] - n[1] * fac, x[2] - n[2] * fac };
    double dpl = fabs(fac);
    double cv[3] = { pp[0] - p->origin[0], pp[1] - p->origin[1], pp[2] - p->origin[2] };
    double din = sqrt(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]);
    double e = 0.0;
    f[0] = f[1] = f[2] = 0.0;
    if (p->subtype == ORC_SUB_ATTRACTIVE) {
        if (dpl > 0.0) {
            double pre = p->k * dpl / dpl;
            for (int d = 0; d < 3; ++d) f[d] += pre * (pp[d] - x[d]);
            e += 0.5 * p->k * dpl * dpl;
        }
        double r = din + prad;
        if (r > p->radius) {
            double pre = p->k * (r - p->radius) / r;
            for (int d = 0; d < 3; ++d) f[d] += pre * (p->origin[d] - pp[d]);
            e += 0.5 * p->k * (r - p->radius) * (r - p->radius);
        }
    } else {
        double r1 = din - prad;
        if (dpl < prad && r1 < p->radius) {
            double pre = p->k * (dpl - prad) / dpl;
            for (int d = 0; d < 3; ++d) f[d] += pre * (pp[d] - x[d]);
            e += 0.5 * p->k * (dpl - prad) * (dpl - prad);
        }
    }
    return e;
}

/* ======================================================================
 * Gradients.
 * ====================================================================== */
static inline double mi64(double d, double Lb)
{
    double h = 0.5 * Lb;
    if (d > h) d -= Lb;
    if (d < -h) d += Lb;
    return d;
}

static double ext_all(const orc_system* s, int t, const double* x, double* g)
{
    double e = 0.0;
    for (int p = 0; p < s->n_pot1; ++p) {
        if (!((s->pot1[p].type_mask >> t) & 1u)) continue;
        double gg[3];
        e += orc_ext_eval(&s->pot1[p], s->prad[t], x, gg);
        g[0] += gg[0]; g[1] += gg[1]; g[2] += gg[2];
    }
    return e;
}

/* energy of the pair counted once; gradient added to particle i only */
static double pair_all(const orc_system* s, int ti, int tj, const double* d, double* gi)
{
    int nt = s->n_types;
    double e = 0.0;
    for (int p = 0; p < s->n_pot2; ++p) {
        if (!s->pot2[p].mask[ti * nt + tj]) continue;
        double gg[3];
        e += orc_pair_eval(&s->pot2[p], s->r0[ti * nt + tj], d, gg);
        gi[0] += gg[0]; gi[1] += gg[1]; gi[2] += gg[2];
    }
    return e;
}

/* Bonded pairs (CudaSimulation.cu:1829-1836 with each group template's own
 * potential and the pair's radius sum, P2_Harmonic.java:122-126): the
 * minimum-image separation, no cutoff; gradient added to both particles,
 * energy counted once.  Returns the bond energy. */
static double bonds_all(const orc_system* s, const double* xyz, const int* type, double* grad)
{
    double Lb[3], e = 0.0;
    for (int d = 0; d < 3; ++d) Lb[d] = s->hi[d] - s->lo[d];
    int nt = s->n_types;
    for (int b = 0; b < s->n_bonds; ++b) {
        int i = s->bonds[2 * b], j = s->bonds[2 * b + 1];
        const orc_pot2* p = &s->bpot[s->bond_pot[b]];
        orc_pot2 q = *p;
        q.cutoff = INFINITY;
        double r0 = s->bond_r0[s->bond_pot[b]];
        if (r0 < 0) r0 = s->r0[type[i] * nt + type[j]];
        double d[3], g[3];
        for (int k = 0; k < 3; ++k) d[k] = mi64(xyz[3 * i + k] - xyz[3 * j + k], Lb[k]);
        e += orc_pair_eval(&q, r0, d, g);
        for (int k = 0; k < 3; ++k) {
            grad[3 * i + k] += g[k];
            grad[3 * j + k] -= g[k];
        }
    }
    return e;
}

double orc_gradients_bruteforce(const orc_system* s, int n, const double* xyz, const int* type,
                                double* grad)
{
    double Lb[3];
    for (int d = 0; d < 3; ++d) Lb[d] = s->hi[d] - s->lo[d];
    double etot = 0.0;
    for (int i = 0; i < n; ++i) {
        double g[3] = { 0, 0, 0 };
        etot += ext_all(s, type[i], &xyz[3 * i], g);
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = mi64(xyz[3 * i + k] - xyz[3 * j + k], Lb[k]);
            double e = pair_all(s, type[i], type[j], d, g);
            if (j > i) etot += e;
        }
        grad[3 * i] = g[0]; grad[3 * i + 1] = g[1]; grad[3 * i + 2] = g[2];
    }
    return etot + bonds_all(s, xyz, type, grad);
}

/* ---- cell list shared by the cell-based routines ---- */
typedef struct {
    int L[3];
    int ncell;
    int* start;   /* [ncell+1] */
    int* order;   /* particle indices sorted by cell */
    int* cc;      /* [3n] cell coordinates */
} clist;

static void clist_build(const orc_system* s, int n, const float* xyzf, clist* cl)
{
    orc_lattice_dims(s, cl->L);
    cl->ncell = cl->L[0] * cl->L[1] * cl->L[2];
    cl->start = (int*)calloc((size_t)cl->ncell + 1, sizeof(int));
    cl->order = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    cl->cc = (int*)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    int* field = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    orc_cell_index(s, n, xyzf, field, cl->cc);
    for (int i = 0; i < n; ++i) cl->start[field[i] + 1]++;
    for (int c = 0; c < cl->ncell; ++c) cl->start[c + 1] += cl->start[c];
    int* fill = (int*)malloc(sizeof(int) * (size_t)cl->ncell);
    memcpy(fill, cl->start, sizeof(int) * (size_t)cl->ncell);
    for (int i = 0; i < n; ++i) cl->order[fill[field[i]]++] = i;
    free(fill);
    free(field);
}

static void clist_free(clist* cl)
{
    free(cl->start); free(cl->order); free(cl->cc);
}

/* stencil extent per dimension: offsets -S..S, or the whole dimension once
   when 2S+1 >= L (no cell visited twice) */
static void stencil_range(int S, int L, int* lo, int* hi)
{
    if (2 * S + 1 >= L) { *lo = 0; *hi = L - 1; }
    else { *lo = -S; *hi = S; }
}

static int stencil_s(double range, double cell)
{
    int S = (int)ceil(range * 1.0001 / cell);
    return S < 1 ? 1 : S;
}

/* visit every particle j in cells within S of particle i's cell */
#define FOR_STENCIL(cl, S, ci, BODY)                                                   \
    do {                                                                               \
        int lx_, hx_, ly_, hy_, lz_, hz_;                                              \
        stencil_range(S, (cl)->L[0], &lx_, &hx_);                                      \
        stencil_range(S, (cl)->L[1], &ly_, &hy_);                                      \
        stencil_range(S, (cl)->L[2], &lz_, &hz_);                                      \
        int full_x_ = (2 * S + 1 >= (cl)->L[0]), full_y_ = (2 * S + 1 >= (cl)->L[1]);  \
        int full_z_ = (2 * S + 1 >= (cl)->L[2]);                                       \
        for (int oz_ = lz_; oz_ <= hz_; ++oz_)                                         \
            for (int oy_ = ly_; oy_ <= hy_; ++oy_)                                     \
                for (int ox_ = lx_; ox_ <= hx_; ++ox_) {                               \
                    int cx_ = full_x_ ? ox_ : ((ci)[0] + ox_ + (cl)->L[0]) % (cl)->L[0]; \
                    int cy_ = full_y_ ? oy_ : ((ci)[1] + oy_ + (cl)->L[1]) % (cl)->L[1]; \
                    int cz_ = full_z_ ? oz_ : ((ci)[2] + oz_ + (cl)->L[2]) % (cl)->L[2]; \
                    int c_ = (cz_ * (cl)->L[1] + cy_) * (cl)->L[0] + cx_;              \
                    for (int q_ = (cl)->start[c_]; q_ < (cl)->start[c_ + 1]; ++q_) {   \
                        int j = (cl)->order[q_];                                       \
                        BODY                                                           \
                    }                                                                  \
                }                                                                      \
    } while (0)

static double max_pair_range(const orc_system* s)
{
    double m = 0.0;
    int nt = s->n_types;
    for (int a = 0; a < nt * nt; ++a) {
        double r = sqrt((double)s->rc2[a]);
        if (r > m) m = r;
    }
    return m;
}

double orc_gradients_cells(const orc_system* s, int n, const double* xyz, const int* type,
                           double* grad, int nthreads)
{
    float* xf = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < 3 * n; ++i) xf[i] = (float)xyz[i];
    clist cl;
    clist_build(s, n, xf, &cl);
    int S = stencil_s(max_pair_range(s), s->cell);
    double Lb[3];
    for (int d = 0; d < 3; ++d) Lb[d] = s->hi[d] - s->lo[d];
    double etot = 0.0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256) reduction(+ : etot)
#else
    (void)nthreads;
#endif
    for (int i = 0; i < n; ++i) {
        double g[3] = { 0, 0, 0 };
        double ei = ext_all(s, type[i], &xyz[3 * i], g);
        const int* ci = &cl.cc[3 * i];
        FOR_STENCIL(&cl, S, ci, {
            if (j != i) {
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = mi64(xyz[3 * i + k] - xyz[3 * j + k], Lb[k]);
                double e = pair_all(s, type[i], type[j], d, g);
                if (j > i) ei += e;
            }
        });
        grad[3 * i] = g[0]; grad[3 * i + 1] = g[1]; grad[3 * i + 2] = g[2];
        etot += ei;
    }
    clist_free(&cl);
    free(xf);
    return etot + bonds_all(s, xyz, type, grad);
}

/* ======================================================================
 * Euler-Maruyama, CudaSimulation.cu:1850-1859:
 *   x += -dt*D*dU/(kB T) + sqrt(2 D dt) * xi,  then one periodic correction.
 * ====================================================================== */
int orc_run(const orc_system* s, int n, double* xyz, const int* type, uint64_t step0, int nsteps,
            const float* noise, int nthreads)
{
    double* grad = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    double a[ORC_MAX_TYPES], b[ORC_MAX_TYPES];
    for (int t = 0; t < s->n_types; ++t) {
        a[t] = s->dt * s->D[t] / s->kbt;
        b[t] = sqrt(2.0 * s->D[t] * s->dt);
    }
    int bad = 0;
    for (int st = 0; st < nsteps; ++st) {
        orc_gradients_cells(s, n, xyz, type, grad, nthreads);
#ifdef _OPENMP
        int nth = nthreads <= 0 ? omp_get_max_threads() : nthreads;
#pragma omp parallel for num_threads(nth) reduction(| : bad)
#endif
        for (int i = 0; i < n; ++i) {
            float z[3];
            if (noise) {
                for (int d = 0; d < 3; ++d) z[d] = noise[((size_t)st * 3 + d) * n + i];
            } else {
                orc_normal3(s->seed, s->stream, step0 + (uint64_t)st, (uint32_t)i, z);
            }
            int t = type[i];
            for (int d = 0; d < 3; ++d) {
                double x = xyz[3 * i + d] - a[t] * grad[3 * i + d] + b[t] * (double)z[d];
                double Lb = s->hi[d] - s->lo[d];
                if (x > s->hi[d]) x -= Lb;
                if (x < s->lo[d]) x += Lb;
                if (!(x >= s->lo[d] && x <= s->hi[d])) bad |= 1;
                xyz[3 * i + d] = x;
            }
        }
    }
    free(grad);
    return bad;
}

/* ======================================================================
 * Set-valued outputs in FP32, canonical expressions:
 *   dx = xi - xj; if (dx > h) dx -= Lb; if (dx < -h) dx += Lb;   (.cu:996-1000)
 *   d2 = (dx*dx + dy*dy) + dz*dz;  test d2 < thr2.
 * ====================================================================== */
static inline float mi32(float d, float Lb, float h)
{
    if (d > h) d -= Lb;
    if (d < -h) d += Lb;
    return d;
}

static inline float d2_f32(const float* a, const float* b, const float* Lb, const float* h)
{
    float dx = mi32(a[0] - b[0], Lb[0], h[0]);
    float dy = mi32(a[1] - b[1], Lb[1], h[1]);
    float dz = mi32(a[2] - b[2], Lb[2], h[2]);
    float r = dx * dx;
    r = r + dy * dy;
    r = r + dz * dz;
    return r;
}

static void box_f32(const orc_system* s, float* Lb, float* h)
{
    for (int d = 0; d < 3; ++d) {
        Lb[d] = (float)s->hi[d] - (float)s->lo[d];
        h[d] = Lb[d] / 2.0f;
    }
}

long orc_pairs_bruteforce(const orc_system* s, int n, const float* xyz, const int* type,
                          const float* thr2, int64_t* out, long cap)
{
    const float* T = thr2 ? thr2 : s->rc2;
    int nt = s->n_types;
    float Lb[3], h[3];
    box_f32(s, Lb, h);
    long cnt = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            float t2 = T[type[i] * nt + type[j]];
            if (d2_f32(&xyz[3 * i], &xyz[3 * j], Lb, h) < t2) {
                if (cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                cnt++;
            }
        }
    return cnt;
}

long orc_pairs_cells(const orc_system* s, int n, const float* xyz, const int* type,
                     const float* thr2, int64_t* out, long cap)
{
    const float* T = thr2 ? thr2 : s->rc2;
    int nt = s->n_types;
    float Lb[3], h[3];
    box_f32(s, Lb, h);
    float m2 = 0.0f;
    for (int a = 0; a < nt * nt; ++a) if (T[a] > m2) m2 = T[a];
    clist cl;
    clist_build(s, n, xyz, &cl);
    int S = stencil_s(sqrt((double)m2), s->cell);
    long cnt = 0;
    for (int i = 0; i < n; ++i) {
        const int* ci = &cl.cc[3 * i];
        FOR_STENCIL(&cl, S, ci, {
            if (j > i) {
                float t2 = T[type[i] * nt + type[j]];
                if (d2_f32(&xyz[3 * i], &xyz[3 * j], Lb, h) < t2) {
                    if (cnt < cap) { out[2 * cnt] = i; out[2 * cnt + 1] = j; }
                    cnt++;
                }
            }
        });
    }
    clist_free(&cl);
    return cnt;
}

/* RDF: CudaSimulation.cu:1805-1819, r = sqrtf(d2); r < rmax;
   bin = (int)floorf(r*nbins/rmax), clamped to nbins-1 */
void orc_rdf_counts(const orc_system* s, int n, const float* xyz, const int* type, float rmax,
                    int nbins, uint64_t* counts, int use_cells)
{
    int nt = s->n_types;
    float Lb[3], h[3];
    box_f32(s, Lb, h);
    float fnb = (float)nbins;
#define RDF_ACC(i, j)                                                          \
    do {                                                                       \
        float r = sqrtf(d2_f32(&xyz[3 * (i)], &xyz[3 * (j)], Lb, h));          \
        if (r < rmax) {                                                        \
            int bin = (int)floorf((r * fnb) / rmax);                           \
            if (bin > nbins - 1) bin = nbins - 1;                              \
            counts[((size_t)type[i] * nt + type[j]) * nbins + bin] += 1;       \
        }                                                                      \
    } while (0)
    if (!use_cells) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (j != i) RDF_ACC(i, j);
    } else {
        clist cl;
        clist_build(s, n, xyz, &cl);
        int S = stencil_s((double)rmax, s->cell);
        for (int i = 0; i < n; ++i) {
            const int* ci = &cl.cc[3 * i];
            FOR_STENCIL(&cl, S, ci, { if (j != i) RDF_ACC(i, j); });
        }
        clist_free(&cl);
    }
#undef RDF_ACC
}

/* g_ab(r_k) = C_ab(k) / (frames * N_a * (N_b - [a==b]) / V * V_shell(k)),
   V_shell = 4/3 pi (r_{k+1}^3 - r_k^3).  Replaces Simulation.cpp:40-62 (Q7). */
void orc_rdf_normalize(int nt, int nbins, double rmax, double volume, const int* npt, long frames,
                       const uint64_t* counts, double* g)
{
    const double pi = 3.14159265358979323846;
    double dr = rmax / nbins;
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
            double pairs = (double)npt[a] * (double)(npt[b] - (a == b ? 1 : 0));
            for (int k = 0; k < nbins; ++k) {
                double r1 = k * dr, r2 = (k + 1) * dr;
                double vs = 4.0 / 3.0 * pi * (r2 * r2 * r2 - r1 * r1 * r1);
                double den = (double)frames * pairs / volume * vs;
                size_t idx = ((size_t)a * nt + b) * nbins + k;
                g[idx] = den > 0.0 ? (double)counts[idx] / den : 0.0;
            }
        }
}
