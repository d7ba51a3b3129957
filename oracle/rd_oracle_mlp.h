/*
 * rd_oracle_mlp.h -- CPU oracle for the composite external potential and its
 * MLP surrogate (TEST INFRASTRUCTURE: only tests, the smoke check and the
 * benchmark's CPU leg load it).
 *
 * Composite potential: a geometry built from order-one pieces (disks and
 * cylinders with the attractive forms of CudaSimulation.cu:1049-1172,
 * generalised to any normal as orc_ext_eval).  A particle feels the piece
 * nearest to it in energy:
 *     U(x) = min_i U_i(x),   grad U(x) = grad U_k(x),  k = first argmin.
 * (Harmonic in the distance to the union of the pieces.)  This combination
 * has no counterpart in the reference, which sums every order-one potential.
 *
 * MLP: 3 -> 64 -> 64 -> 64 -> 3, SiLU after the three hidden layers.  Weight
 * vector layout (8771 values):
 *     W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
 * Input: the normalised position xn (FP32).  Output: the normalised gradient.
 * With bf16 != 0 the oracle rounds (round-to-nearest-even) at exactly the
 * points where the GPU kernels store BF16: every weight, the hidden
 * activations h1..h3, and in the backward pass dy, dz3, dz2 (the MFMA
 * operands); everything else is FP64.
 */
#ifndef RD_ORACLE_MLP_H
#define RD_ORACLE_MLP_H
#include "rd_oracle.h"

#define ORC_MLP_NPARAM 8771
#define ORC_MLP_H 64

/* offsets into the weight vector */
#define ORC_MLP_W1 0
#define ORC_MLP_B1 192
#define ORC_MLP_W2 256
#define ORC_MLP_B2 4352
#define ORC_MLP_W3 4416
#define ORC_MLP_B3 8512
#define ORC_MLP_W4 8576
#define ORC_MLP_B4 8768

/* energy of one piece (attractive disk or cylinder) and its gradient */
double orc_piece_eval(const orc_pot1* p, const double x[3], double g[3]);

/* nearest-piece composite for m positions (OpenMP); which[i] = argmin piece,
 * gap[i] = second-smallest minus smallest piece energy (NULL to skip) */
void orc_composite_eval(const orc_pot1* pieces, int npieces, long m, const double* xyz, double* grad,
                        double* energy, int* which, double* gap, int nthreads);
/* sum of the pieces: energy[m] and gradient[3m] (either may be NULL) */
void orc_composite_sum(const orc_pot1* pieces, int npieces, long m, const double* xyz, double* grad,
                       double* energy, int nthreads);

/* round an FP32 value to BF16 (nearest, ties to even), returned as FP32 */
float orc_bf16(float v);

/* forward pass: y[m][3] */
void orc_mlp_forward(const double* w, long m, const float* xn, int bf16, double* y);
/* The same network in the inference form of the kernels: layer 1 and the
 * biases times -log2(e), W4 times -ln 2, activations zs / (1 + 2^zs) =
 * -log2(e) SiLU(z).  bf16 = 0: exact scale factors in FP64 (the network
 * itself, to rounding); bf16 != 0: the kernels' FP32 products and BF16
 * roundings (W1 rounded to BF16 before its FP32 scaling, W4's product
 * rounded to BF16, activations rounded to BF16). */
void orc_mlp_forward_infer(const double* w, long m, const float* xn, int bf16, double* y);

/* loss = (1/m) sum_i |y_i - t_i|^2 and its gradient with respect to w */
double orc_mlp_loss_grad(const double* w, long m, const float* xn, const float* t, int bf16, double* grad);

#endif
