/*
 * rd_oracle_mlp.c -- CPU oracle for the composite external potential and the
 * MLP surrogate (TEST INFRASTRUCTURE, see rd_oracle_mlp.h).
 */
#include "rd_oracle_mlp.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

double orc_piece_eval(const orc_pot1* p, const double x[3], double g[3])
{
    /* attractive disk / cylinder forms of CudaSimulation.cu:1071-1148 with a
     * general normal (orc_ext_eval); the particle radius does not enter */
    return orc_ext_eval(p, 0.0, x, g);
}

void orc_composite_eval(const orc_pot1* pieces, int npieces, long m, const double* xyz, double* grad,
                        double* energy, int* which, double* gap, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long i = 0; i < m; ++i) {
        const double* x = xyz + 3 * i;
        double best = INFINITY, second = INFINITY, gb[3] = { 0, 0, 0 };
        int kb = -1;
        for (int k = 0; k < npieces; ++k) {
            double g[3];
            double e = orc_piece_eval(&pieces[k], x, g);
            if (e < best) {
                second = best;
                best = e;
                kb = k;
                gb[0] = g[0]; gb[1] = g[1]; gb[2] = g[2];
            } else if (e < second) {
                second = e;
            }
        }
        if (grad) { grad[3 * i] = gb[0]; grad[3 * i + 1] = gb[1]; grad[3 * i + 2] = gb[2]; }
        if (energy) energy[i] = kb >= 0 ? best : 0.0;
        if (which) which[i] = kb;
        if (gap) gap[i] = second - best;
    }
}

/* Sum of the pieces (CudaSimulation.cu:1049-1236 adds every order-one
 * potential of a particle's type): energy and gradient per position. */
void orc_composite_sum(const orc_pot1* pieces, int npieces, long m, const double* xyz, double* grad,
                       double* energy, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long i = 0; i < m; ++i) {
        const double* x = xyz + 3 * i;
        double e = 0.0, gs[3] = { 0, 0, 0 };
        for (int k = 0; k < npieces; ++k) {
            double g[3];
            e += orc_piece_eval(&pieces[k], x, g);
            gs[0] += g[0]; gs[1] += g[1]; gs[2] += g[2];
        }
        if (grad) { grad[3 * i] = gs[0]; grad[3 * i + 1] = gs[1]; grad[3 * i + 2] = gs[2]; }
        if (energy) energy[i] = e;
    }
}

float orc_bf16(float v)
{
    uint32_t u;
    memcpy(&u, &v, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v;   /* inf / nan unchanged */
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static double rnd(double v, int bf) { return bf ? (double)orc_bf16((float)v) : v; }
static double sigm(double z) { return 1.0 / (1.0 + exp(-z)); }

/* weights, rounded to BF16 where the kernels store them as BF16 */
typedef struct {
    double W1[64][3], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[3][64], b4[3];
} mlp_w;

static void load_w(const double* w, int bf, mlp_w* M)
{
    for (int i = 0; i < 64; ++i)
        for (int c = 0; c < 3; ++c) M->W1[i][c] = rnd(w[ORC_MLP_W1 + 3 * i + c], bf);
    for (int i = 0; i < 64; ++i) {
        M->b1[i] = w[ORC_MLP_B1 + i];
        M->b2[i] = w[ORC_MLP_B2 + i];
        M->b3[i] = w[ORC_MLP_B3 + i];
        for (int k = 0; k < 64; ++k) {
            M->W2[i][k] = rnd(w[ORC_MLP_W2 + 64 * i + k], bf);
            M->W3[i][k] = rnd(w[ORC_MLP_W3 + 64 * i + k], bf);
        }
    }
    for (int o = 0; o < 3; ++o) {
        M->b4[o] = w[ORC_MLP_B4 + o];
        for (int k = 0; k < 64; ++k) M->W4[o][k] = rnd(w[ORC_MLP_W4 + 64 * o + k], bf);
    }
}

typedef struct {
    double z1[64], h1[64], z2[64], h2[64], z3[64], h3[64], y[3];
} mlp_act;

static void fwd(const mlp_w* M, const float* xn, int bf, mlp_act* A)
{
    for (int i = 0; i < 64; ++i) {
        double z = M->b1[i];
        for (int c = 0; c < 3; ++c) z += M->W1[i][c] * (double)xn[c];
        A->z1[i] = z;
        A->h1[i] = rnd(z * sigm(z), bf);
    }
    for (int i = 0; i < 64; ++i) {
        double z = M->b2[i];
        for (int k = 0; k < 64; ++k) z += M->W2[i][k] * A->h1[k];
        A->z2[i] = z;
        A->h2[i] = rnd(z * sigm(z), bf);
    }
    for (int i = 0; i < 64; ++i) {
        double z = M->b3[i];
        for (int k = 0; k < 64; ++k) z += M->W3[i][k] * A->h2[k];
        A->z3[i] = z;
        A->h3[i] = rnd(z * sigm(z), bf);
    }
    for (int o = 0; o < 3; ++o) {
        double y = M->b4[o];
        for (int k = 0; k < 64; ++k) y += M->W4[o][k] * A->h3[k];
        A->y[o] = y;
    }
}

void orc_mlp_forward(const double* w, long m, const float* xn, int bf16, double* y)
{
    mlp_w* M = (mlp_w*)malloc(sizeof(mlp_w));
    load_w(w, bf16, M);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long i = 0; i < m; ++i) {
        mlp_act A;
        fwd(M, xn + 3 * i, bf16, &A);
        y[3 * i] = A.y[0]; y[3 * i + 1] = A.y[1]; y[3 * i + 2] = A.y[2];
    }
    free(M);
}

/* inference form (rd_surrogate.hip: mlp_tile, pack_param's w1s/b1s/b2s/b3s/w4s) */
void orc_mlp_forward_infer(const double* w, long m, const float* xn, int bf16, double* y)
{
    const double nl2e = -1.0 / log(2.0), nln2 = -log(2.0);
    mlp_w* M = (mlp_w*)malloc(sizeof(mlp_w));
    load_w(w, bf16, M);                 /* W2, W3 as the kernels hold them */
    for (int i = 0; i < 64; ++i) {
        for (int c = 0; c < 3; ++c) {
            const double v = w[ORC_MLP_W1 + 3 * i + c];
            M->W1[i][c] = bf16 ? (double)((float)orc_bf16((float)v) * -1.44269504f) : v * nl2e;
        }
        M->b1[i] = bf16 ? (double)((float)w[ORC_MLP_B1 + i] * -1.44269504f) : w[ORC_MLP_B1 + i] * nl2e;
        M->b2[i] = bf16 ? (double)((float)w[ORC_MLP_B2 + i] * -1.44269504f) : w[ORC_MLP_B2 + i] * nl2e;
        M->b3[i] = bf16 ? (double)((float)w[ORC_MLP_B3 + i] * -1.44269504f) : w[ORC_MLP_B3 + i] * nl2e;
    }
    for (int o = 0; o < 3; ++o)
        for (int k = 0; k < 64; ++k) {
            const double v = w[ORC_MLP_W4 + 64 * o + k];
            M->W4[o][k] = bf16 ? (double)orc_bf16((float)v * -0.693147181f) : v * nln2;
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long s = 0; s < m; ++s) {
        double h[64], hn[64];
        for (int i = 0; i < 64; ++i) {
            double z = M->b1[i];
            for (int c = 0; c < 3; ++c) z += M->W1[i][c] * (double)xn[3 * s + c];
            h[i] = rnd(z / (1.0 + exp2(z)), bf16);
        }
        for (int l = 0; l < 2; ++l) {
            for (int i = 0; i < 64; ++i) {
                double z = l == 0 ? M->b2[i] : M->b3[i];
                for (int k = 0; k < 64; ++k) z += (l == 0 ? M->W2[i][k] : M->W3[i][k]) * h[k];
                hn[i] = rnd(z / (1.0 + exp2(z)), bf16);
            }
            memcpy(h, hn, sizeof(h));
        }
        for (int o = 0; o < 3; ++o) {
            double v = M->b4[o];
            for (int k = 0; k < 64; ++k) v += M->W4[o][k] * h[k];
            y[3 * s + o] = v;
        }
    }
    free(M);
}

static double dsilu(double z)
{
    double s = sigm(z);
    return s * (1.0 + z * (1.0 - s));
}

double orc_mlp_loss_grad(const double* w, long m, const float* xn, const float* t, int bf16, double* grad)
{
    mlp_w* M = (mlp_w*)malloc(sizeof(mlp_w));
    load_w(w, bf16, M);
    memset(grad, 0, sizeof(double) * ORC_MLP_NPARAM);
    double loss = 0.0;
    const double inv = 1.0 / (double)m;
    for (long s = 0; s < m; ++s) {
        mlp_act A;
        fwd(M, xn + 3 * s, bf16, &A);
        double dy[3], dyb[3], dh[64], dz3[64], dz3b[64], dz2[64], dz2b[64], dz1[64];
        for (int o = 0; o < 3; ++o) {
            double r = A.y[o] - (double)t[3 * s + o];
            loss += r * r * inv;
            dy[o] = 2.0 * inv * r;
            dyb[o] = rnd(dy[o], bf16);
        }
        for (int o = 0; o < 3; ++o) {
            grad[ORC_MLP_B4 + o] += dy[o];
            for (int k = 0; k < 64; ++k) grad[ORC_MLP_W4 + 64 * o + k] += dyb[o] * A.h3[k];
        }
        for (int k = 0; k < 64; ++k) {
            double a = 0.0;
            for (int o = 0; o < 3; ++o) a += M->W4[o][k] * dyb[o];
            dh[k] = a;
        }
        for (int i = 0; i < 64; ++i) {
            dz3[i] = dh[i] * dsilu(A.z3[i]);
            dz3b[i] = rnd(dz3[i], bf16);
            grad[ORC_MLP_B3 + i] += dz3[i];
            for (int k = 0; k < 64; ++k) grad[ORC_MLP_W3 + 64 * i + k] += dz3b[i] * A.h2[k];
        }
        for (int k = 0; k < 64; ++k) {
            double a = 0.0;
            for (int i = 0; i < 64; ++i) a += M->W3[i][k] * dz3b[i];
            dh[k] = a;
        }
        for (int i = 0; i < 64; ++i) {
            dz2[i] = dh[i] * dsilu(A.z2[i]);
            dz2b[i] = rnd(dz2[i], bf16);
            grad[ORC_MLP_B2 + i] += dz2[i];
            for (int k = 0; k < 64; ++k) grad[ORC_MLP_W2 + 64 * i + k] += dz2b[i] * A.h1[k];
        }
        for (int k = 0; k < 64; ++k) {
            double a = 0.0;
            for (int i = 0; i < 64; ++i) a += M->W2[i][k] * dz2b[i];
            dh[k] = a;
        }
        for (int i = 0; i < 64; ++i) {
            dz1[i] = dh[i] * dsilu(A.z1[i]);
            grad[ORC_MLP_B1 + i] += dz1[i];
            for (int c = 0; c < 3; ++c) grad[ORC_MLP_W1 + 3 * i + c] += dz1[i] * (double)xn[3 * s + c];
        }
    }
    free(M);
    return loss;
}
