// Host-side batch object: problems resident in HBM and the solve loop.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "amie_gpu.h"
#include "device.hpp"

struct amie_gpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    size_t total_mem = 0;
    std::atomic<int> children{0};   // live batches and servers (created and destroyed on any host thread)
    bool closing = false;    // amie_gpu_destroy called while children live
    int32_t *voxel_scratch = nullptr;   // this device's copy of the voxel ownership array (voxels.hip)
};

namespace amie {

void set_error(const char *fmt, ...);
bool hip_ok(hipError_t e, const char *what);
#define AMIE_HIP(call)                                      \
    do {                                                    \
        if (!::amie::hip_ok((call), #call)) return AMIE_ERR_HIP; \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    bool alloc(size_t bytes);
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Element types: 1 P1, 2 P2, 3 TET4
struct EType {
    int et, nper, dim, b, s;
};
EType etype_info(int et);

struct Kern {
    void (*k1)(SolveArgs);          // SpMV + dots
    void (*red)(SolveArgs);         // per-problem scalars and decisions
    void (*k2)(SolveArgs, int);     // vector update (arg: checkpoint slot follows)
    void (*cx)(SolveArgs);          // refinement: x64 += S y
    void (*cr)(SolveArgs);          // refinement: FP64 residual, p
    void (*norms)(SolveArgs);
    void (*norms_final)(SolveArgs);
    void (*kp)(SolveArgs, PersistArgs) = nullptr;   // the persistent solve of these kernels, where built
};

// State of one asynchronous solve (begin / advance / poll / end).
// The SpMV kernel the autotuner picked for one precision / preconditioner.
struct K1Choice {
    void (*k1)(SolveArgs) = nullptr;
    int steps = 0;           // blocks per step of the row loop
    int format = 0;          // 0: the precision's value planes, 1: the value-indexed FIX48 matrix
    int tile_order = 0;      // SolveArgs::tile_order
    double us = 0.;          // its time per launch with every tile active
};

struct SolveRun {
    amie_solve_opts o{};
    int mode = 0;            // precision of the current pass
    bool mixed = false, bj = false;
    std::vector<int64_t> iters_done;   // iterations of earlier passes, per problem
    std::vector<int64_t> nckpt_done;   // FP64 refinements of earlier passes, per problem
    int ce = 4;              // iterations between checkpoint opportunities
    int per_check = 8;       // checkpoint periods per host status check
    Kern K{};
    SolveArgs a{};
    int gt = 1, gp = 1;      // grids: active tiles, active problems
    hipGraphExec_t exec = nullptr;
    std::vector<double> pbytes;       // algorithmic bytes of one SpMV launch, per problem of the pass
    std::vector<double> pbytes_k64;   // the same with the FP64 matrix (after a precision hand-over)
    std::vector<hipEvent_t> tev;
    int nt = 0;
    std::vector<char> was_running;
    int64_t periods = 0, max_periods = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr, ev_status = nullptr;
    bool started = false, in_flight = false, finished = false;
    int pgrid = 0;                   // > 0: this pass runs as one persistent launch of pgrid workgroups
    bool persist_launched = false;
    amie_solve_stats stats{};
};

struct ProbHost {
    int64_t nrows = 0;                // DOF nodes (block rows)
    int64_t nelem = 0, nnodes = 0;
    int32_t slice0 = 0, nslices = 0, tile0 = 0, ntiles = 0;
    int64_t row0 = 0;                 // first internal global row
    int64_t elem0 = 0, node0 = 0;     // offsets of this problem's elements / geometry nodes
    int32_t cmat0 = 0;
    std::vector<uint32_t> row_size, column_index;   // reference layout, original numbering
    std::vector<int64_t> ref_slot;                  // reference block -> global SELL slot
    std::vector<int32_t> iperm;                     // original row -> local internal row
    std::vector<int64_t> elem_ref;                  // element (i,j) -> reference block index
    int64_t positions = 0;                          // block positions of the slices (one per 64 slots)
    int64_t offset_blocks = 0;                      // real blocks at positions with a shared column offset
    bool has_elements = false;
};

struct Batch {
    amie_gpu_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;    // every launch of this batch goes here
    SolveRun run;
    EType et{};
    int nrhs = 1;
    int nprob = 0;
    std::vector<ProbHost> ph;
    int64_t nslices = 0, ntiles = 0, nrows_pad = 0, nslots = 0, nelem = 0, nnodes_geo = 0;
    bool has_elements = false, has_macro = false, has_damage_state = false;
    bool fix48_ok = false;            // the last assembly's scaled values fit the fixed-point range
    std::vector<int32_t> colour_ptr;  // colour c elements: colour_elems[colour_ptr[c]..colour_ptr[c+1])
    std::map<int, K1Choice> k1_tuned; // by precision | (block Jacobi ? 8 : 0)

    // SELL structure
    DevBuf slice_ptr, slice_width, slice_rows, col, dslot, row_len;
    DevBuf tile_desc, slice_desc;     // packed {slice0, nslices, prob} / {slot ptr lo, hi, width, rows}
    DevBuf slot_off;                  // per 64 slots: column - row shared by the position's 64 lanes, or NO_OFFSET
    DevBuf dict_ids, dict_table;      // value-indexed FIX48 matrix: 16-bit index per slot, distinct blocks (FP64)
    bool dict_wanted = false;         // rebuild the dictionary at every assembly (set by the autotuner)
    bool dict_ok = false;             // the dictionary matches the current FIX48 planes
    int64_t dict_n = 0;               // its distinct blocks
    DevBuf v64, v32, vsp, vfx;   // FP64 unscaled K; FP32, split FP32+bf16 and 48-bit fixed-point scaled Ks
    DevBuf vs64;                 // FP64 scaled Ks, built on the first FP64 solve after an assembly
    bool pre64_ok = false;       // vs64 matches the current assembly
    DevBuf fix48_bad;            // int flag: a scaled value outside the fixed-point range
    DevBuf tile_slice0, tile_nslices, tile_prob, active, n_active, active_prob, n_active_prob;
    DevBuf prob, part1, part2, partc, partn;
    // vectors: one plane of nrows_pad * b entries per right-hand side; 8 bytes
    // per entry (the iteration vectors y, r, p, q hold floats in mixed mode)
    DevBuf x64, b64, fext, S, y, r, p, q, minv32, minv64;
    DevBuf p2, p3, p4;           // p's other plane sets (p_k in set k % YGROUP; set 0 is p)
    DevBuf bcmask, bcval;
    // elements
    DevBuf coords, econn, erow, eslot, ephase, edamage, ekappa, ed0, eprob, cmat, cmat_off, macro, colour_elems,
        prob_e0, prob_ne, grown;
    // gather assembly: per slot its contributions (e*nper + i)*nper + j, per
    // internal row its (element, local node) pairs e*nper + i, in the
    // reference's summation order
    DevBuf slot_cptr, slot_contrib, row_cptr, row_contrib;
    int asm_method = 0;               // AMIE_ASSEMBLE_GATHER | AMIE_ASSEMBLE_COLOURED
    // non-local strain smoothing (amie_gpu_batch_set_nonlocal)
    DevBuf egeo, ngrid, ngdim, cell_start, cell_elem;
    bool nonlocal_on = false;
    int nl_kernel = 0;
    double nl_radius = 0., nl_threshold = 1.;
    bool nl_wave = false, ev_wave = false;   // a wavefront per element in k_smooth / k_fiber_decide
    // event-driven damage (amie_gpu_batch_event_step): an element with
    // damage >= frac_at keeps frac_residual of its stiffness
    double frac_at = HUGE_VAL, frac_residual = 0.;
    DevBuf ev_tmp, ev_sbar, ev_score, ev_dprev, ev_flags, ev_crit, ev_prm, ev_count;
    DevBuf ev_inc, ev_tmax;   // IsotropicLinearDamage: increments of the pass, largest score per problem
    DevBuf small_desc, small_out, small_part;   // k_cg_small: per problem {slice0, nslices}, results; grid sums
    std::vector<int32_t> h_active, h_active_prob;
    std::vector<ProbDev> h_prob;
    ProbDev *h_prob_pinned = nullptr;
    int32_t *h_nactive_pinned = nullptr;   // [0] tiles, [1] problems
    DevBuf persist;                        // persistent solve: barrier count (8 bytes), error flag, periods
    int32_t *h_persist_pinned = nullptr;   // [0] error flag, [1] periods run

    // captured iteration chunks, keyed by solve variant
    struct GraphEntry {
        int key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;

    ~Batch();
    size_t bytes() const;
    SolveArgs args(int mode) const;
    int grid(int64_t tiles = -1) const;
    int assemble();
    int ensure_pre64(int mode);
    int solve(const amie_solve_opts &o, amie_solve_stats *st);
    int solve_begin(const amie_solve_opts &o);
    int start_pass(int mode, const std::vector<char> &subset);
    int solve_advance();
    int solve_poll(bool wait, bool *finished);
    int solve_end(amie_solve_stats *st);
    int set_active_all();
    int set_active(const std::vector<char> &probs);
    int autotune(int mode, bool bj, int iters);
    int damage_update(const amie_damage_params &prm, int64_t *n_grown);
    int event_step(const amie_event_params &prm, const amie_solve_opts &o, amie_event_stats *st);
    void apply_tuned(int mode, bool bj, Kern &K, SolveArgs &a) const;
    int k1_format(int mode, bool bj) const;
    int64_t vplane() const { return nrows_pad * et.b; }
};

// Builds the value-indexed FIX48 matrix of the current assembly (dict_ok
// tells whether it exists: at most DICT_MAX distinct blocks).
int build_dict(Batch &B);
int build_dict_or_drop(Batch &B);   // a build out of device memory leaves the batch without one

// FP64 plane vector of problem p <-> original DOF order, one column of
// nrows*b entries per right-hand side.
// gather assembly: whether the batch's element blocks fit 32-bit list
// offsets, and the device build of the slot and row lists
bool gather_fits(int64_t nelem, int nper, int64_t nslots);
int build_gather_lists(Batch &B);
int download_vec(Batch &B, int p, const DevBuf &buf, double *out);
int upload_vec(Batch &B, int p, DevBuf &buf, const double *in);

int batch_create(amie_gpu_ctx *ctx, int nprob, const amie_problem_desc *probs, Batch **out);
// Device address of the voxel ownership array on the current device (set up
// once per device; the device must be current).
int voxel_scratch_address(int device, int32_t **ptr);
void ctx_retain(amie_gpu_ctx *ctx);
void ctx_release(amie_gpu_ctx *ctx);
int batch_from_bcsr(amie_gpu_ctx *ctx, int stride, int64_t nrows, const uint32_t *row_size,
                    const uint32_t *column_index, const double *array, const double *rhs, Batch **out);
int upload_bcsr_values(Batch &B, const double *array, const double *rhs);

} // namespace amie

struct amie_gpu_batch {
    amie::Batch *b;
};
