/*
 * rdgpu.h -- C ABI of libCReaDDyGPU.so, the MI355X (gfx950) Brownian-dynamics
 * step of ReaDDyGPU.
 *
 * Every entry point returns 0 on success and a nonzero RD_E* code on error,
 * as the reference's CudaSimulation methods return 0/1
 * (CudaSimulation.cuh:40-71).  rd_last_error() gives the message.
 * Handles are not thread safe; one handle drives one GPU stream.
 *
 * Units are whatever the caller uses consistently (the reference GPU input
 * uses nm, s and kJ/mol; only D*dt and kB*T matter).
 */
#ifndef RDGPU_H
#define RDGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RD_MAX_TYPES 16        /* particle types per system */
#define RD_MAX_POT2 8          /* order-two potentials per system */
#define RD_MAX_BOND_POTS 8     /* group (bond) potentials per system */
#define RD_MAX_PARTICLES (1 << 26)

enum {
    RD_OK = 0,
    RD_EINVAL = 1,      /* bad argument */
    RD_ESTATE = 2,      /* call out of order (e.g. step before initialize) */
    RD_EHIP = 3,        /* HIP runtime error */
    RD_ERANGE = 4,      /* particle outside the box, non-finite position, capacity */
    RD_ENOMEM = 5
};

/* order-one potential kinds: typeID of Simulation.cpp:100-199 */
enum { RD_POT1_DISK = 1, RD_POT1_CYLINDER = 2, RD_POT1_SPHERE = 3, RD_POT1_BOX = 4 };
/* order-two potential kinds: HARMONIC, WEAK_INTERACTION_HARMONIC
 * (Simulation.cpp:172-204) and WEAK_INTERACTION_PIECEWISE_HARMONIC
 * (P2_WeakInteractionPiecewiseHarmonic.java) */
enum { RD_POT2_HARMONIC = 1, RD_POT2_WEAK_HARMONIC = 2, RD_POT2_PIECEWISE_WELL = 3 };
/* subtypeID of Simulation.cpp:105-110, 176-181 */
enum { RD_SUB_ATTRACTIVE = 1, RD_SUB_REPULSIVE = 2, RD_SUB_SPRING = 3 };

typedef struct rd_system rd_system;

/* Global parameters: Simulation fields read by GlobalParametersHandler.cpp:73-143
 * and hard-coded in lib/ReaDDyGPU.cpp:96-123. */
typedef struct {
    double lattice_bounds[6]; /* latticeBounds: xlo,xhi,ylo,yhi,zlo,zhi (periodic box) */
    double dt;                /* BD step size (dtOpenMM / stepSizeInPs) */
    double boltzmann;         /* kB */
    double temperature;       /* T */
    double max_cutoff;        /* maxCutoff: pairs with r > max_cutoff are skipped
                                 (CudaSimulation.cu:1002); reference value 10 */
    double cell_size;         /* lattice edge; <= 0 chooses 1.01 x the largest
                                 pair range.  The reference uses maxCutoff. */
    uint64_t seed;            /* noise seed (Philox key) */
    uint32_t stream;          /* noise stream (replica id) */
    int device;               /* cudaDeviceIndex (CudaSimulation.cu:189-200) */
} rd_params;

/* CudaOrderOnePotential (CudaSimulation.cuh:18-28) plus its row of the
 * type mask M1[pot][type] (CudaSimulation.cu:286-306). */
typedef struct {
    int type;              /* RD_POT1_* */
    int subtype;           /* RD_SUB_ATTRACTIVE / RD_SUB_REPULSIVE */
    float force_const;
    float origin[3];       /* centre (disk, cylinder, sphere) or corner (box) */
    float normal[3];       /* disk / cylinder axis, any direction */
    float extension[3];    /* box */
    float height;          /* cylinder */
    float radius;          /* disk, cylinder, sphere */
    uint32_t type_mask;    /* bit t: applies to particle type t */
} rd_pot1;

/* CudaOrderTwoPotential (CudaSimulation.cuh:30-37). */
typedef struct {
    int type;              /* RD_POT2_* */
    int subtype;           /* RD_SUB_* (harmonic) */
    float force_const;     /* k */
    float length;          /* weak harmonic: width; piecewise well: l */
    float depth;           /* piecewise well: d */
} rd_pot2;

typedef struct {
    int n;                 /* live particles */
    int n_types;
    int lattice[3];        /* cells per dimension */
    double cell_size;
    int stencil;           /* pair stencil radius in cells */
    uint64_t step;         /* global step counter */
    long rdf_frames;
    int n_replicas;        /* independent systems held by the handle (1: plain) */
    int n_cells;           /* cells of all replicas */
} rd_info;

/* ---- set-up: Simulation::parse (Simulation.cpp:206-280) and
 *      CudaSimulation::initialize (CudaSimulation.cu:138-455) ---- */
int rd_create(const rd_params* params, rd_system** out);
void rd_destroy(rd_system* h);
const char* rd_last_error(const rd_system* h);

/* particle types: D (ParticleType.D), default radius, and the collision radius
 * matrix radii[i*nt+j] (0 -> default radius of i).  r0(i,j) = r_i(j) + r_j(i)
 * as CudaSimulation.cu:357-391. */
int rd_set_particle_types(rd_system* h, int n_types, const double* D, const double* default_radius,
                          const double* radii_matrix);
/* toCudaOrderOnePotential (CudaSimulation.cu:79-117) */
int rd_add_order_one_potential(rd_system* h, const rd_pot1* pot);
/* toCudaOrderTwoPotential (CudaSimulation.cu:119-136); pair_mask[i*nt+j]
 * = M2[i][j][pot] (CudaSimulation.cu:311-353), symmetrised */
int rd_add_order_two_potential(rd_system* h, const rd_pot2* pot, const uint8_t* pair_mask);
/* reaction radius per type pair (0: no reaction); candidates are pairs with
 * dist < R strictly (ReactionObserver.java:101-130) */
int rd_set_reaction_radii(rd_system* h, const double* radius_matrix);
/* Bonded groups: the groups kernel (CudaSimulation.cu:1829-1836) over the
 * GroupPart list and group potential table built in CudaSimulation.cu:394-435
 * from the group templates (Simulation.cpp:254-264, GroupParametersHandler.cpp)
 * and instances (GroupReader.cpp:13-40).  n_bonds pairs of global particle
 * indices bonds[2b], bonds[2b+1]; bond b uses pots[pot_index[b]] (HARMONIC,
 * WEAK_HARMONIC or PIECEWISE_WELL) with rest length r0[p] (NULL or < 0: the
 * collision radius sum of the two particles' types, as P2_Harmonic.java:122-126),
 * minimum-image distance and no cutoff.  Both particles of a bond must lie in
 * the same replica.  Reactions renumber bonds with their particles and drop
 * those of removed particles.  n_bonds = 0 removes every bond. */
int rd_set_bonds(rd_system* h, int n_bonds, const int* bonds, const int* pot_index, int n_pots, const rd_pot2* pots,
                 const double* r0);
/* CoordReader (CoordReader.cpp:16-65) + copyPosToDevice/copyTypesToDevice
 * (CudaSimulation.cu:905-970).  Particle index = position in these arrays. */
int rd_set_particles(rd_system* h, int n, const float* xyz, const int* type);
/* Batched replicas (no counterpart in the reference, which runs one system per
 * process: the Simulation object of Simulation.hpp:138-181 holds one box).
 * n_replicas independent periodic boxes share the particle types and
 * potentials of the handle; replica r has box bounds[6r..6r+5] (lo,hi per
 * axis), noise stream streams[r] (NULL: r) and counts[r] particles, stored
 * consecutively in xyz/type.  Global particle index = offset of the replica
 * + local index.  One step advances every replica; the noise of a particle
 * depends only on (seed, stream, step, local index), so a replica evolves
 * bit for bit as the same system run alone with params.stream = its stream.
 * Reactions need a single-system handle. */
int rd_set_replicas(rd_system* h, int n_replicas, const double* bounds, const uint32_t* streams,
                    const int* counts, const float* xyz, const int* type);
/* CudaSimulation::initialize: upload tables, build the first cell list */
int rd_initialize(rd_system* h);

/* ---- the step: Simulation::simulate(int) (Simulation.cpp:349-355) ->
 *      CudaSimulation::simulate (CudaSimulation.cu:2309-2476) ---- */
int rd_simulate(rd_system* h, int nsteps);

/* copyPosFromDevice (CudaSimulation.cu:972-984): positions in index order */
int rd_copy_pos_from_device(rd_system* h, float* xyz, int* type);
/* reference cell index of every particle, index order: the field formula of
 * CudaSimulation.cu:586-588 evaluated on the current FP32 positions */
int rd_cell_ids(rd_system* h, int* field);
/* gradient of U at the current positions (what orderOne + warpOrderTwo leave
 * in cudaForces), index order, and the total energy accumulated in FP64 */
int rd_gradients(rd_system* h, float* grad, double* energy);

/* ---- noise: seeded Philox4x32-10 + Box-Muller, FP32 N(0,1).
 *      Replaces the curand states of CudaSimulation.cu:749-885, 1921-1926. ---- */
/* out[(s*3+d)*n + i] = noise used for particle index i, dimension d, step step0+s */
int rd_noise_fill(rd_system* h, uint64_t step0, int nsteps, float* out);
/* take noise from the given array (same layout, steps step..step+nsteps-1)
 * instead of generating it; nsteps = 0 returns to the generator */
int rd_set_noise_array(rd_system* h, const float* noise, int nsteps);

/* ---- RDF: callRDFCalculation + copyRDFMatrix (CudaSimulation.cu:457-520),
 *      normalizeRDF (Simulation.cpp:40-62) ---- */
int rd_rdf_configure(rd_system* h, double r_max, int n_bins);
int rd_rdf_calculate(rd_system* h);
/* counts[((r*nt+ti)*nt+tj)*n_bins+bin] over frames so far (ordered pairs),
 * r = replica (0 only for a plain system) */
int rd_rdf_copy(rd_system* h, uint64_t* counts, long* frames);
/* g(r) in FP64 with the shell volume and the box volume */
int rd_rdf_normalize(rd_system* h, double* g);
int rd_rdf_reset(rd_system* h);

/* ---- pair sets, index pairs (i<j) sorted lexicographically ---- */
/* pairs with d^2 < rc^2(ti,tj): the pairs whose interaction is evaluated */
int rd_neighbor_pairs(rd_system* h, int64_t* out, long cap, long* count);
/* pairs with d^2 < R^2(ti,tj), R from rd_set_reaction_radii */
int rd_reaction_candidates(rd_system* h, int64_t* out, long cap, long* count);
/* The same candidate set as rd_reaction_candidates, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): no counterpart in the reference.  Its
 * scratch (per-particle offsets, scan sums, the pair list) is a fixed device
 * array of the library, about 48 MB per GPU, held until the process ends; the
 * call makes no device allocation.  Limits: at most 2^23 particles and 2^20
 * candidate pairs (RD_ERANGE beyond them); the handle must be up to date
 * (after rd_initialize or a step; RD_ESTATE otherwise).  Calls that overlap on
 * one GPU share the scratch and race. */
int rd_reaction_candidates_noalloc(rd_system* h, int64_t* out, long cap, long* count);

/* ---- reactions: the cReactions protocol of lib/ReaDDyGPU.cpp:303-463:
 *      creactions = [count, (readdyId, newType, x, y, z, index) x count];
 *      index != -1 deletes particle `index`, newType != -1 appends one.
 *      Survivors keep their order; new particles follow.  The compaction runs
 *      on the device (the reference rebuilds every array on the host and
 *      uploads it again, CudaSimulation.cu:936-970). ---- */
int rd_apply_reactions(rd_system* h, const float* creactions);
/* The same in a handle of replicas (no counterpart in the reference, which
 * holds one system): entry r that only creates (index == -1) puts its particle
 * into replica creation_replica[r]; a type change stays in the replica of its
 * index.  Replica-local order: survivors, then new particles in entry order. */
int rd_apply_reactions_replicas(rd_system* h, const float* creactions, const int* creation_replica);

/* ---- output files of the reference's stand-alone driver
 *      (outputHandler.cpp:17-148, opened in lib/ReaDDyGPU_Cpp/ReaDDyGPU.cpp:68-85):
 *      the same text, printf format for printf format; the XML trajectory is
 *      closed with "</traj>" (the reference writes "<" TAB "raj>").  No GPU
 *      needed. ---- */
enum { RD_TRAJ_PDB = 1, RD_TRAJ_XML = 2 };
typedef struct rd_traj rd_traj;
int rd_traj_open(const char* path, int format, rd_traj** out);
/* one MODEL / tplgy_coords block: particle a is written as id a+1 */
int rd_traj_write_frame(rd_traj* t, int frame, int n, const float* xyz, const int* type);
int rd_traj_close(rd_traj* t);
/* RDF table: header "d" + ",\tRDF(a,b)" per type pair, then per bin k the
 * distance k*range/nbins ("%8.5f") and g[(i*nt+j)*nbins+k] (", %8.3f") */
int rd_rdf_write(const char* path, int n_types, const char* const* type_names, int nbins, double range,
                 const double* g);

/* ---- frame hand-off: the frame loop of lib/ReaDDyGPU.cpp:235-301
 *      (simulate, copyPosFromDevice, frameCallback) without its synchronous
 *      copies, so the next frame's steps can run during the hand-off ----
 * rd_simulate_async: enqueue steps and return (errors surface at the next
 *   rd_frame_wait or rd_simulate).
 * rd_frame_copy: enqueue a copy of the current positions and types (index
 *   order) into a device snapshot and into pinned host memory.
 * rd_frame_wait: wait for that copy; *xyz [3n] and *type [n] stay valid until
 *   the next rd_frame_copy.
 * rd_frame_rewind: discard the steps enqueued after the last rd_frame_copy and
 *   return to its state (positions, types, step counter, so the same noise
 *   follows); the particle set must be unchanged since the copy. */
int rd_simulate_async(rd_system* h, int nsteps);
int rd_frame_copy(rd_system* h);
int rd_frame_wait(rd_system* h, const float** xyz, const int** type, int* n);
int rd_frame_rewind(rd_system* h);

int rd_get_info(rd_system* h, rd_info* info);

/* ---- checkpoint / resume.  No counterpart in the reference: its de-facto
 *      restart is the XML trajectory (outputHandler.cpp:135-148, read by
 *      CoordReader.cpp:27-48), which drops the step counter, so the noise of a
 *      restarted run does not continue the original one. ----
 * rd_checkpoint_write: binary snapshot of the particle state (index order),
 *   replica boxes, streams and counts, bonds, seed and step counter.
 * rd_checkpoint_read: load it into a handle built with the same particle types,
 *   potentials, bond potentials and seed; the run then continues bit for bit
 *   as the uninterrupted one.
 * rd_set_step: set the step (noise) counter. */
int rd_checkpoint_write(rd_system* h, const char* path);
int rd_checkpoint_read(rd_system* h, const char* path);
int rd_set_step(rd_system* h, uint64_t step);

/* Debug invariant check of the cell-sorted state (no counterpart in the
 * reference, whose only integrity checks are the host list walkers of
 * CudaSimulation.cu:628-726): every particle lies in the slot range of its
 * cell, every particle index occurs once, cell starts rise to the particle
 * count.  *violations = number of failed checks (0: consistent). */
int rd_check_invariants(rd_system* h, long* violations);

/* ---- composite external potential and its learned surrogate ----
 * No counterpart in the reference.  The composite is a geometry of order-one
 * pieces: attractive disks and cylinders (the forms of orderOne,
 * CudaSimulation.cu:1049-1148, with any normal).  A particle whose type is in
 * type_mask feels the piece nearest to it in energy, U = min_i U_i, with the
 * gradient of that piece (first index on ties).  It is evaluated exactly by a
 * pass over all pieces, or by a BF16 MLP 3-64-64-64-3 (SiLU) trained on the
 * GPU on gradients of the exact composite.  Both add to the particle's
 * gradient before the Euler-Maruyama update. */
enum { RD_EXT_EXACT = 0, RD_EXT_SURROGATE = 1 };
int rd_set_composite(rd_system* h, int n_pieces, const rd_pot1* pieces, uint32_t type_mask);
/* How the pieces combine: RD_COMBINE_NEAREST (default) as above, or
 * RD_COMBINE_SUM: U = sum_i U_i and the gradient is the sum of the pieces'
 * gradients, as the reference's orderOne adds every order-one potential
 * (CudaSimulation.cu:1049-1236).  Applies to the exact pass, rd_composite_eval
 * and the surrogate's training data. */
enum { RD_COMBINE_NEAREST = 0, RD_COMBINE_SUM = 1 };
int rd_set_composite_combine(rd_system* h, int combine);
int rd_set_external_mode(rd_system* h, int mode);
/* exact composite at m host positions: grad[3m], energy[m] (either may be NULL) */
int rd_composite_eval(rd_system* h, long m, const float* xyz, float* grad, float* energy);

/* Surrogate parameters: FP32 master vector of 8771 values,
 * W1[64][3] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3].
 * Input xn = (x - lo) * (2 / (hi - lo)) - 1 per axis (FP32); output times
 * gscale = dU/dx. */
typedef struct {
    int batch;             /* samples per optimiser step (multiple of 64 is best) */
    int steps;             /* optimiser steps */
    float lr, lr_min;      /* cosine schedule from lr to lr_min */
    float beta1, beta2, eps;
    uint64_t seed;         /* minibatch sampling */
} rd_train_params;
int rd_surrogate_init(rd_system* h, uint64_t seed);
int rd_surrogate_set_weights(rd_system* h, const float* w);
int rd_surrogate_get_weights(rd_system* h, float* w);
int rd_surrogate_set_norm(rd_system* h, const double* lo3, const double* hi3, double gscale);
int rd_surrogate_get_norm(rd_system* h, double* lo3, double* hi3, double* gscale);
/* m training samples: uniform positions in the domain, exact gradients; sets gscale */
int rd_surrogate_make_data(rd_system* h, long m, uint64_t seed, double* gscale);
/* Adam over the training samples; loss of every step into loss_curve[steps] (may be NULL) */
int rd_surrogate_train(rd_system* h, const rd_train_params* params, float* loss_curve);
/* loss (1/m) sum |y - t|^2 and its gradient for one batch of normalised
 * inputs xn[3m] and targets t[3m], without an update */
int rd_surrogate_loss_grad(rd_system* h, long m, const float* xn, const float* t, double* loss, float* grad);
/* raw network outputs y[3m] at normalised inputs xn[3m] */
int rd_surrogate_predict(rd_system* h, long m, const float* xn, float* y);
/* relative RMSE of the surrogate gradient against the exact composite on m
 * fresh uniform samples */
int rd_surrogate_validate(rd_system* h, long m, uint64_t seed, double* rel_rmse);

/* ---- input files: the reference's XML schema (Xerces-free reader) ----
 * rd_model is the parsed Simulation of Simulation::readXML / readGlobalXML /
 * parse (Simulation.cpp:206-338): param_global (GlobalParametersHandler.cpp:
 * 73-143), param_particles (ParticleParametersHandler.cpp:48-162),
 * tplgy_potentials (PotentialParametersHandler.cpp:38-83), tplgy_coordinates
 * (CoordReader.cpp:16-65), param_groups (GroupParametersHandler.cpp) and
 * tplgy_groups (GroupReader.cpp:13-40).  Parsing needs no GPU. */
typedef struct rd_model rd_model;
typedef struct {
    int n_particles;
    int n_types;
    int n_pot1, n_pot2;
    int n_groups;            /* group instances (tplgy_groups <g>) */
    int n_bonds;             /* bonded pairs they define */
    int n_bond_pots;         /* group potentials of all templates, in file order */
    int frames;              /* nSimulationSteps */
    int steps_per_frame;     /* round(dt / dtOpenMM) (XMLReader.cpp:67) */
    double dt_frame;         /* dt */
    double dt_step;          /* dtOpenMM / dt_dynamic / dt_dyn */
    double temperature, boltzmann;
    double lattice_bounds[6];
    int device;              /* cudaDeviceIndex / cudaDevice */
    int testmode;
} rd_model_info;
int rd_model_read_xml(const char* param_particles, const char* tplgy_potentials, const char* param_groups,
                      const char* tplgy_coordinates, const char* tplgy_groups, const char* param_global,
                      rd_model** out, char* err, int err_len);
void rd_model_free(rd_model* m);
int rd_model_get_info(const rd_model* m, rd_model_info* info);
/* D, default radius [nt]; radii matrix [nt*nt] (0 = default) */
int rd_model_types(const rd_model* m, double* D, double* radius, double* radii_matrix);
/* positions [3n], types [n], ReaDDy particle ids [n] in file order */
int rd_model_particles(const rd_model* m, float* xyz, int* type, int* readdy_id);
int rd_model_pot1(const rd_model* m, int i, rd_pot1* pot);
int rd_model_pot2(const rd_model* m, int i, rd_pot2* pot, uint8_t* pair_mask);
/* bonds [2*n_bonds] (particle indices in file order), pot_index [n_bonds],
 * pots [n_bond_pots], r0 [n_bond_pots] (< 0: collision radius sum); any
 * output may be NULL */
int rd_model_bonds(const rd_model* m, int* bonds, int* pot_index, rd_pot2* pots, double* r0);
/* Simulation + CudaSimulation::initialize from a parsed model (maxCutoff as
 * given; the reference JNI path hard-codes 10, lib/ReaDDyGPU.cpp:101) */
int rd_create_from_model(const rd_model* m, double max_cutoff, uint64_t seed, rd_system** out);

/* ---- launch configuration of the force/update kernel.  No counterpart in
 *      the reference (its launch shape is fixed: 128 threads per particle
 *      block, 32 per cell, CudaSimulation.cu:193, 2347).  Any configuration
 *      gives bitwise the same trajectory: pair gradients are summed in fixed
 *      point, so only speed changes. ---- */
enum { RD_KERNEL_AUTO = 0, RD_KERNEL_GATHER = 1, RD_KERNEL_QUEUE = 2 };
typedef struct {
    int kernel;            /* RD_KERNEL_*: GATHER one lane per particle walks its stencil rows,
                              testing distances first and evaluating the listed hits after, one
                              workgroup per group of 64-slot work items.  QUEUE: the same per-lane
                              work on a resident grid (occupancy x compute units) whose workgroups
                              each own a contiguous share of the work items, taken by their waves
                              from a ticket in LDS until none are left (ppt is ignored).
                              AUTO = GATHER */
    int used_kernel;       /* out: kernel in use */
    int block;             /* threads per workgroup of the force kernel: 64, 128, 256 or 512; 0 = 256 */
    int used_block;        /* out: workgroup size in use */
    int ppt;               /* particles per lane of the force kernel: 1, 2 or 4 (a wave takes that many
                              consecutive 64-slot work items in turn); 0 = 1 */
    int used_ppt;          /* out: particles per lane in use */
} rd_launch_config;
int rd_set_launch_config(rd_system* h, const rd_launch_config* cfg);
int rd_get_launch_config(rd_system* h, rd_launch_config* cfg);

/* Autotuner: times `trial_steps` steps for each candidate cell size (factors
 * of the interaction range) and force-kernel workgroup size, then particles
 * per lane at the winning pair; keeps the fastest (the default shape, 256
 * threads and one particle per lane, unless another is more than 2% faster),
 * and restores the particle state and step counter, so tuning does not change
 * the run.  Values fixed by rd_set_launch_config (kernel, block, ppt) or the
 * cell size are kept; the work-queue form is tuned only when selected there. */
int rd_autotune(rd_system* h, int trial_steps, double* best_ms_per_step);

/* Lattice cell edge of a running system (<= 0: 1.01 x the largest interaction
 * range); the particles keep their positions and the step counter is kept.
 * RD_ERANGE (previous lattice kept) when the lattice would be too large.  No
 * counterpart in the reference (its cell edge is fixed at creation,
 * CudaSimulation.cu:522-560); used to apply a cached autotuner choice. */
int rd_set_cell_size(rd_system* h, double cell);

/* "<name> <arch> <N>CU" of a HIP device, the key of cached autotuner choices.
 * No counterpart in the reference. */
int rd_device_name(int device, char* buf, int cap);

#ifdef __cplusplus
}
#endif
#endif
