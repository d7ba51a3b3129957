// rd_host.h -- host-side state of one system (the CudaSimulation object of
// CudaSimulation.cuh:39-116, re-laid out for the sorted SoA design).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/rdgpu.h"
#include "rd_device.h"
#include "rd_mlp.h"

struct RdStepState;

struct rd_system {
    rd_params prm{};
    std::string err;
    int dev = 0;
    int ncu = 1;                     // compute units of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // model (Simulation fields, Simulation.hpp:138-181)
    int nt = 0;
    std::vector<double> D, rad, radii, rxn;
    std::vector<rd_pot1> pot1;
    std::vector<rd_pot2> pot2;
    std::vector<std::vector<uint8_t>> pot2mask;
    std::vector<RdPot1Dev> pot1dev;
    RdDev P{};
    double cell = 0.0, max_range = 0.0, max_rxn = 0.0;
    int ncell = 0;                // cells of all replicas
    int nscan = 0;

    // replicas: independent systems of one handle, particles concatenated
    struct Rep {
        double lo[3], hi[3];
        uint32_t stream;
        int n;
    };
    std::vector<Rep> reps;
    std::vector<RdRep> rep_dev;
    std::vector<RdWave> wave_tab;
    std::vector<long> tcount;        // particles per (replica, type), [R * nt]
    int nwaves = 0;

    // particles in index order (host mirror)
    int n = 0;
    std::vector<float> hxyz;
    std::vector<int> htype;

    bool initialized = false;
    bool dirty = true;            // tables changed
    bool particles_dirty = true;  // host arrays newer than device
    uint64_t step = 0;

    // device buffers
    float4* d_ps = nullptr;
    float4* d_pn = nullptr;
    uint2* d_kr = nullptr;        // (next cell, rank inside it) of each unsorted slot
    uint32_t* d_count = nullptr;
    uint32_t* d_start = nullptr;
    uint32_t* d_bsum = nullptr;
    RdDev* d_P = nullptr;
    RdPot1Dev* d_pot1 = nullptr;
    RdStepState* d_st = nullptr;
    float* d_noise = nullptr;
    float* d_xyz_io = nullptr;
    int* d_type_io = nullptr;
    double* d_etot = nullptr;
    RdRep* d_reps = nullptr;
    RdWave* d_waves = nullptr;
    int* d_rep_of = nullptr;
    int cap = 0;
    int ncell_alloc = 0;
    int nrep_alloc = 0;
    int nwave_alloc = 0;
    size_t noise_cap = 0;
    int noise_n = -1;             // particle count the noise array was laid out for
    size_t io_cap = 0;

    // bonded pairs of group potentials (global particle indices)
    std::vector<int> bond_pairs;      // 2 per bond
    std::vector<int> bond_pot;        // potential per bond
    std::vector<rd_pot2> bpots;
    std::vector<double> bpot_r0;      // < 0: collision radius sum
    int* d_bstart = nullptr;
    int2* d_bonds = nullptr;
    uint32_t* d_slot = nullptr;       // global index -> slot (null without bonds)
    uint32_t* d_gid = nullptr;        // slot of the next-step array -> global index (null without bonds)

    // device-side reactions: the io arrays hold the new particle set until upload
    bool io_on_device = false;

    // frame hand-off: the state at the last rd_frame_copy, on the device (for
    // rd_frame_rewind) and in pinned host memory
    float* d_snap_xyz = nullptr;
    int* d_snap_type = nullptr;
    float* hp_xyz = nullptr;
    int* hp_type = nullptr;
    RdStepState* hp_st = nullptr;
    int snap_cap = 0, snap_n = 0;
    uint64_t snap_step = 0;
    bool snap_valid = false;
    hipEvent_t ev_frame = nullptr;   // copy to the host done (on copy_stream)
    hipEvent_t ev_snap = nullptr;    // snapshot written (on stream)
    hipStream_t copy_stream = nullptr;

    // RDF
    unsigned long long* d_rdf = nullptr;
    unsigned long long* d_rdf_part = nullptr;   // rdf_ncopy copies of d_rdf (cleared between calls)
    int rdf_ncopy = 0;
    double rdf_rmax = 0.0;
    int rdf_bins = 0;
    long rdf_frames = 0;
    int rdf_nrep = 0;
    std::vector<double> rdf_vol, rdf_pairs_sum;

    // pair emission
    uint2* d_pairs = nullptr;
    size_t pairs_cap = 0;
    uint32_t* d_poff = nullptr;   // per-particle pair offsets and their scan tile sums
    uint32_t* d_pbs = nullptr;
    size_t poff_cap = 0;

    // captured steps: one graph per step count
    struct Graph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::map<int, Graph> graphs;

    // force-kernel launch configuration (autotuned)
    rd_launch_config lc{};        // requested
    int kern = RD_KERNEL_GATHER;  // kernel in use (RD_KERNEL_*)
    int block = 256;              // force-kernel workgroup size in use
    int ppt = 1;                  // force-kernel particles per lane in use
    int qgrid = 0, qgrid_block = 0; // resident workgroups of the work-queue form (for qgrid_block threads)

    // composite external potential (nearest piece) and its MLP surrogate
    std::vector<rd_pot1> comp;
    uint32_t comp_mask = 0;
    bool comp_sum = false;           // RD_COMBINE_SUM: sum of the pieces (else nearest piece)
    int npieces = 0;
    void* d_pieces = nullptr;
    int ext_mode = RD_EXT_EXACT;
    int ext_blocks[2] = { 0, 0 };    // resident workgroups of the exact / surrogate slot passes (0: not yet known)
    float4* d_gext = nullptr;        // external gradient per sorted slot
    int gext_cap = 0;
    float *d_master = nullptr, *d_mom = nullptr, *d_vel = nullptr;
    void* d_pack = nullptr;          // RdMlpPack
    void* d_ctl = nullptr;           // training control block
    float* d_partial = nullptr;
    size_t partial_rows = 0;
    float4 *d_xn = nullptr, *d_tg = nullptr;
    long nsamples = 0;
    float* d_gscale = nullptr;
    double* d_sbsum = nullptr;       // per-block sums of the training-data scale
    float* d_loss_hist = nullptr;
    size_t loss_hist_cap = 0;
    RdMlpNorm mlp_norm{};
    double norm_hi[3] = { 0, 0, 0 };
    bool norm_set = false;
    bool sur_ready = false;
    bool ext_active() const { return npieces > 0 || (ext_mode == RD_EXT_SURROGATE && sur_ready); }
};

double rd_host_r0(const rd_system* h, int i, int j);
int rd_host_refresh(rd_system* h);
int rd_host_set_step(rd_system* h, uint64_t step);
int rd_host_set_state(rd_system* h, const std::vector<float>& xyz, const std::vector<int>& type);
int rd_host_step_graph(rd_system* h, int k, hipGraphExec_t* exec);
void rd_host_launch_force(rd_system* h, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void rd_host_launch_sort(rd_system* h, int advance);
void rd_host_launch_ext(rd_system* h);
int rd_host_ext_capacity(rd_system* h);
void rd_host_free_surrogate(rd_system* h);
