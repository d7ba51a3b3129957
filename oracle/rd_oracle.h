/*
 * rd_oracle.h -- CPU oracle for the Brownian-dynamics step (TEST INFRASTRUCTURE).
 *
 * This library is the checker, never the product: only the test suite, the
 * smoke check and the benchmark's CPU-baseline leg load it.  It restates the
 * reference method in plain C:
 *
 *   cell lattice and cell index  CudaSimulation.cu:522-528, 586-588
 *   pair potentials              CudaSimulation.cu:992-1032,
 *                                P2_Harmonic.java:144-234,
 *                                P2_WeakInteractionPiecewiseHarmonic.java:114-200
 *   external potentials          CudaSimulation.cu:1034-1236, P1_Disk.java:112-275,
 *                                P1_Sphere.java:222-341, P1_Cylinder.java:224-280
 *   Euler-Maruyama update        CudaSimulation.cu:1838-1919 (x += -dt*D*dU/kT + sqrt(2Ddt)*N)
 *   RDF histogram                CudaSimulation.cu:1784-1827
 *   reaction candidates          ReactionObserver.java:101-130 (dist < R strictly)
 *   bonded groups                CudaSimulation.cu:1829-1836 (groups kernel), pairs and
 *                                potentials from the group templates (Simulation.cpp:254-264)
 *
 * Trajectories and gradients are FP64.  Set-valued outputs (cell ids, pair
 * sets, RDF bins) are evaluated with the canonical FP32 expressions spelled
 * out in rd_oracle.c so that they can be compared bit for bit.
 *
 * The noise generator is Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
 * followed by a Box-Muller transform written with IEEE-exact operations only
 * (+ - * / sqrt, no fused multiply-add), so the same FP32 N(0,1) values come
 * out of any IEEE-754 machine that evaluates the operations in order.
 */
#ifndef RD_ORACLE_H
#define RD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_TYPES 64

/* potential kinds: numbering follows Simulation.cpp:100-199 (typeID) */
enum { ORC_P1_DISK = 1, ORC_P1_CYLINDER = 2, ORC_P1_SPHERE = 3, ORC_P1_BOX = 4 };
enum { ORC_P2_HARMONIC = 1, ORC_P2_WEAK_HARMONIC = 2, ORC_P2_PIECEWISE_WELL = 3 };
enum { ORC_SUB_ATTRACTIVE = 1, ORC_SUB_REPULSIVE = 2, ORC_SUB_SPRING = 3 };

typedef struct {
    int kind, subtype;
    double k;
    double origin[3];     /* centre (disk, cylinder, sphere) or corner (box) */
    double normal[3];     /* disk / cylinder axis (normalised by the oracle) */
    double extension[3];  /* box */
    double height;        /* cylinder */
    double radius;        /* disk, cylinder, sphere */
    uint64_t type_mask;   /* bit t set: applies to particle type t */
} orc_pot1;

typedef struct {
    int kind, subtype;
    double k;
    double length;        /* weak harmonic: width; piecewise well: l */
    double depth;         /* piecewise well: d */
    double cutoff;        /* range cap for attractive / spring harmonic */
    const uint8_t* mask;  /* [nt*nt] 1 where the potential applies */
} orc_pot2;

typedef struct {
    int n_types;
    double lo[3], hi[3];       /* periodic box = latticeBounds */
    double cell;               /* lattice cell edge (maxCutoff) */
    double kbt;                /* kB*T */
    double dt;
    const double* D;           /* [nt] diffusion constants */
    const double* prad;        /* [nt] particle radius (box potential) */
    const double* r0;          /* [nt*nt] collision radius sums */
    const float* rc2;          /* [nt*nt] FP32 squared interaction range */
    int n_pot1;
    const orc_pot1* pot1;
    int n_pot2;
    const orc_pot2* pot2;
    uint64_t seed;
    uint32_t stream;           /* replica id for batched runs */
    int n_bonds;               /* bonded pairs of group potentials */
    const int* bonds;          /* [2*n_bonds] particle indices */
    const int* bond_pot;       /* [n_bonds] index into bpot */
    const orc_pot2* bpot;      /* bond potentials (mask unused, cutoff ignored) */
    const double* bond_r0;     /* per bond potential: rest length, < 0: r0 of the type pair */
} orc_system;

/* ---- noise ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_log_f32(float x);
void orc_sincos_turns_f32(float t, float* s, float* c);
void orc_normal3(uint64_t seed, uint32_t stream, uint64_t step, uint32_t index, float out[3]);
/* out[(s*3+d)*n + i] for s in [0,nsteps), i in [0,n) (index i) */
void orc_noise_fill(uint64_t seed, uint32_t stream, uint64_t step0, int nsteps, int n, float* out);

/* ---- lattice ---- */
void orc_lattice_dims(const orc_system* s, int L[3]);
/* reference field index from FP32 positions; also per-dimension coordinates */
void orc_cell_index(const orc_system* s, int n, const float* xyz, int* field, int* cxyz);

/* ---- potentials (single evaluation; used by the known-answer tests) ---- */
/* gradient of U with respect to x_i for the pair (i,j); d = x_i - x_j (minimum image) */
double orc_pair_eval(const orc_pot2* p, double r0, const double d[3], double grad_i[3]);
/* gradient of U at x; prad = particle radius (box only, and the Java disk form) */
double orc_ext_eval(const orc_pot1* p, double prad, const double x[3], double grad[3]);
/* Java P1_Disk form with particle radius (P1_Disk.java:112-275), for the golden vectors */
double orc_disk_java(const orc_pot1* p, double prad, const double x[3], double force[3]);

/* ---- gradients / energy / step ---- */
/* brute force O(N^2) FP64 gradient and total energy (minimum image pairs) */
double orc_gradients_bruteforce(const orc_system* s, int n, const double* xyz, const int* type,
                                double* grad);
/* cell-list FP64 gradient: same result as brute force, O(N); nthreads<=0 -> all */
double orc_gradients_cells(const orc_system* s, int n, const double* xyz, const int* type,
                           double* grad, int nthreads);
/* nsteps EM steps in place (FP64) with the seeded noise; step0 = global step counter.
   noise != NULL: take noise[(s*3+d)*n+i] instead of the generator. */
int orc_run(const orc_system* s, int n, double* xyz, const int* type, uint64_t step0, int nsteps,
            const float* noise, int nthreads);

/* ---- set-valued outputs (FP32 canonical) ---- */
/* pairs (i<j) with d2 < rc2[ti][tj] (or < thr2[ti][tj] when thr2 != NULL).
   Writes up to cap pairs to out[2*k], returns total count. */
long orc_pairs_bruteforce(const orc_system* s, int n, const float* xyz, const int* type,
                          const float* thr2, int64_t* out, long cap);
long orc_pairs_cells(const orc_system* s, int n, const float* xyz, const int* type,
                     const float* thr2, int64_t* out, long cap);
/* RDF: counts[(ti*nt+tj)*nbins+bin] += 1 for every ordered pair i!=j with r < rmax */
void orc_rdf_counts(const orc_system* s, int n, const float* xyz, const int* type, float rmax,
                    int nbins, uint64_t* counts, int use_cells);
/* textbook g(r) in FP64: counts / (frames * N_i * rho_j * V_shell) */
void orc_rdf_normalize(int nt, int nbins, double rmax, double volume, const int* n_per_type,
                       long frames, const uint64_t* counts, double* g);

#ifdef __cplusplus
}
#endif
#endif
