// Device-side data structures and kernels of the AMIE FE-solve hot path on
// MI355X (gfx950, wave64).
//
// Matrix storage ("SELL-64 block"): the node-block CSR of the reference
// (sparse/sparse_matrix.h:129-137) is re-laid out in slices of 64 block rows,
// one slice per wavefront.  Within a slice the k-th block of every row is
// stored contiguously over the 64 lanes, so each load instruction of a
// wavefront reads one contiguous run (a 2x2 FP32 block is one 16-byte
// float4 per lane: 1 KiB per wavefront).  Rows shorter than the slice width
// are padded with zero blocks, whose column is the row plus the column offset
// most common at that block position (the row itself where that falls
// outside the problem).  Block entries
// keep the reference's column-major order (entry (r,c) at c*b + r), without
// the reference's 4x3 padding of 3x3 blocks.  Blocks are split into 16-byte
// "planes" so every plane load is a dwordx4 (or dword/dwordx2 remainder):
//   FP32 2x2: 1 x float4            FP32 3x3: 2 x float4 + 1 float
//   FP64 2x2: 2 x double2           FP64 3x3: 4 x double2 + 1 double
//
// Vectors hold one plane per right-hand side: entry (k, row, i) sits at
// (k * nrows_pad + row) * b + i.  A wavefront's access to one right-hand
// side of its 64 rows, and its gather of neighbour rows (whose columns mostly
// run along the rows), is then one contiguous run of 64 b-vectors.
//
// Work is distributed in tiles of up to 4 slices (256 rows, one 256-thread
// workgroup) that never straddle two problems of a batch; the iteration
// kernels run one workgroup per active tile.
//
// One CG iteration on the scaled system Ks = S K S (S = diag(K)^-1/2) with
// preconditioner M (identity, or the inverses of the scaled diagonal blocks):
//   K1 (k_spmv)   q = Ks p; partial sums of p.q, q.z and q.Mq (z = M r)
//   R  (k_reduce) per problem: alpha = rho / p.q, the next
//                 rho' = (r - alpha q).M(r - alpha q) = rho - 2 alpha q.z +
//                 alpha^2 q.Mq, beta = rho' / rho; convergence, refinement
//                 and stagnation decisions
//   K2 (k_update) y += alpha p; r -= alpha q; z = M r; p = z + beta p;
//                 partial sums of the direct rho = r.z and r.r for the next R
// so K1 gathers a single vector and K2 is a pure stream.  A refinement (the
// mixed-precision outer step, and the FP64 confirmation of convergence) is
// C1 (k_ckpt_x) x64 += S y, y = 0 and C2 (k_ckpt_r) r64 = b - K x64 in FP64,
// r = S r64, p = M r + beta p.
//
// Reductions never cross workgroups inside a kernel: on CDNA the L2 of each
// XCD is not coherent with the others, so making one workgroup's writes
// visible to another inside a kernel needs a device-scope release (a write
// back of the whole L2).  Every kernel writes per-tile partial sums instead,
// and R sums a problem's tiles in tile order.  All per-problem solver state
// is written by R (thread 0 of the problem's workgroup), except the flag
// C1 raises for R; kernel boundaries publish it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace amie {

constexpr int WAVE = 64;
constexpr int TILE_SLICES = 4;
constexpr int WG = WAVE * TILE_SLICES;
constexpr int MAXRHS = 4;
constexpr int YGROUP = 4;   // CG iterations per y += alpha p group (a power of 2, at most 4)
constexpr int MAXQ = 3 * MAXRHS;
constexpr int NXCD = 8;

enum Status : int32_t { ST_IDLE = 0, ST_RUNNING = 1, ST_CONVERGED = 2, ST_MAXIT = 3, ST_NAN = 4, ST_STALLED = 5 };

// Per-problem constants (host) and CG state (device), FP64, resident in HBM.
struct ProbDev {
    int32_t tile0, ntiles;
    int32_t slice0, nslices;
    int64_t row0;
    int32_t nrows;
    int32_t refine;          // 0 replace the residual, keep p; 1 restart at each refinement
    double rtol, drop;
    double ref_eps;          // > 0: converged when ||S r64|| <= ref_eps (absolute, scaled)
    int64_t maxit;
    int64_t ckpt_interval;   // refinement at least every this many iterations
    int64_t stall_iters;     // > 0: ST_STALLED at the accuracy floor, or when the FP64 residual has not halved in that many iterations
    // state, written by k_reduce (src_ckpt: raised by k_ckpt_x)
    int32_t status;          // a finished status only ever stops work
    int32_t ckpt_pending;    // a refinement is due at the next checkpoint slot
    int32_t src_ckpt;        // the latest residual partials are a refinement's
    int32_t first;           // the next refinement starts the recurrence (p = M r)
    int32_t done[MAXRHS];
    double alpha[MAXRHS], beta[MAXRHS];
    double rho_pred[MAXRHS]; // the recurrence's rho for the next residual
    double relres[MAXRHS];
    int64_t iters, nckpt, last_ckpt;
    int32_t ladder;          // the split matrix may hand over to the FP64 matrix
    int32_t kmode;           // 1: the iteration's SpMV uses the FP64 matrix (S K S on the fly)
    int64_t iters_k64;       // iterations run with kmode = 1
    double rr[MAXRHS], rr_ckpt[MAXRHS], rr_target[MAXRHS], bb[MAXRHS], bb64[MAXRHS], rel_best[MAXRHS];
    int64_t best_it[MAXRHS]; // iteration of the last halving of rel_best
    // The update adds alpha p to y every YGROUP-th iteration, YGROUP terms
    // at once: p_k lives in plane set k % YGROUP (k = iters when the SpMV
    // runs), so p_{k-3} .. p_{k-1} are still there when the update of p_k
    // adds them.
    int32_t ypend;           // terms the previous updates left out of y (0 .. YGROUP - 1)
    int32_t pad_y;
    double alpha_ring[MAXRHS][YGROUP]; // alpha_k at [k % YGROUP]
    // results for the host
    double relres_out[MAXRHS];
    int32_t done_out[MAXRHS];
};

// Arguments shared by the solve kernels.
struct SolveArgs {
    const int64_t *slice_ptr;
    const int32_t *slice_width;
    const int32_t *slice_rows;
    const uint32_t *col;
    int64_t nslots;
    const double *v64;       // FP64 planes (unscaled K)
    const void *vw;          // planes of the iteration's SpMV (scaled FP32 S K S, or v64)
    const int32_t *tile_slice0, *tile_nslices, *tile_prob;
    const int4 *tile_desc;   // per tile {first slice, slices, problem, 0}
    const int4 *slice_desc;  // per slice {slot pointer lo, hi, width, rows}
    const int32_t *slot_off; // per 64 slots (one block position of a slice): column - row shared by all lanes, or NO_OFFSET
    const double *dict_table; // value-indexed FIX48 matrix: the distinct blocks, decoded (b*b doubles each)
    int32_t ntiles_all;
    int32_t tile_order;      // 0: XCD-contiguous tile ranges (xcd_index); C > 0: chunks of C tiles dealt
                             // round-robin over the XCDs (1: launch order)
    const int32_t *active;   // active tiles
    const int32_t *n_active;
    const int32_t *active_prob;
    const int32_t *n_active_prob;
    ProbDev *prob;
    double *part1;           // [tile][MAXQ] p.q, q.z, q.Mq
    double *part2;           // [tile][MAXQ] r.z, r.r after the update
    double *partc;           // [tile][MAXQ] r.z, r.r, |b - K x|^2, |S (b - K x)|^2 after a refinement
    double *partn;           // [tile][MAXQ] |S b|^2, |b|^2
    int64_t vplane;          // entries per right-hand-side plane (nrows_pad * b)
    double *x64;
    const double *b64;
    const double *S;         // symmetric Jacobi scaling, per dof
    void *y, *r, *p, *q;     // iteration vectors (float or double)
    void *pset[YGROUP];      // p's plane sets: p_k in pset[k % YGROUP] (pset[0] = p)
    const void *minv;        // block-Jacobi inverses of the scaled diagonal blocks, per row (b*b, row-major)
    int32_t *fix48_bad;      // set when a scaled value does not fit the fixed-point range
};

// ------------------------------------------------------------ block planes
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <typename W, int B> struct Blk;

// Blocks are held in registers in their stored form (Raw) and expanded to
// values only where they are used.
template <typename M, int B> struct BlkRaw {
    typedef typename Blk<M, B>::T T;
    struct Raw {
        T v[B * B];
    };
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, Raw &r) { Blk<M, B>::load(base, ns, s, r.v); }
    __device__ static inline T at(const Raw &r, int k) { return r.v[k]; }
};

template <> struct Blk<float, 2> {
    typedef float T;
    static constexpr int N = 4;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, float *a)
    {
        f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(base) + s);
        a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
        (void)ns;
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const float *a)
    {
        reinterpret_cast<float4 *>(base)[s] = make_float4(a[0], a[1], a[2], a[3]);
        (void)ns;
    }
    static constexpr size_t bytes_per_slot = 16;
};
template <> struct Blk<float, 3> {
    typedef float T;
    static constexpr int N = 9;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, float *a)
    {
        const f4v *p = reinterpret_cast<const f4v *>(base);
        f4v v0 = __builtin_nontemporal_load(p + s);
        f4v v1 = __builtin_nontemporal_load(p + ns + s);
        float v2 = __builtin_nontemporal_load(reinterpret_cast<const float *>(p + 2 * ns) + s);
        a[0] = v0.x; a[1] = v0.y; a[2] = v0.z; a[3] = v0.w;
        a[4] = v1.x; a[5] = v1.y; a[6] = v1.z; a[7] = v1.w; a[8] = v2;
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const float *a)
    {
        float4 *p = reinterpret_cast<float4 *>(base);
        p[s] = make_float4(a[0], a[1], a[2], a[3]);
        p[ns + s] = make_float4(a[4], a[5], a[6], a[7]);
        reinterpret_cast<float *>(p + 2 * ns)[s] = a[8];
    }
    static constexpr size_t bytes_per_slot = 36;
};
template <> struct Blk<double, 2> {
    typedef double T;
    static constexpr int N = 4;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, double *a)
    {
        const d2v *p = reinterpret_cast<const d2v *>(base);
        d2v v0 = __builtin_nontemporal_load(p + s);
        d2v v1 = __builtin_nontemporal_load(p + ns + s);
        a[0] = v0.x; a[1] = v0.y; a[2] = v1.x; a[3] = v1.y;
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const double *a)
    {
        double2 *p = reinterpret_cast<double2 *>(base);
        p[s] = make_double2(a[0], a[1]);
        p[ns + s] = make_double2(a[2], a[3]);
    }
    static constexpr size_t bytes_per_slot = 32;
};
template <> struct Blk<double, 3> {
    typedef double T;
    static constexpr int N = 9;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, double *a)
    {
        const d2v *p = reinterpret_cast<const d2v *>(base);
        d2v v0 = __builtin_nontemporal_load(p + s);
        d2v v1 = __builtin_nontemporal_load(p + ns + s);
        d2v v2 = __builtin_nontemporal_load(p + 2 * ns + s);
        d2v v3 = __builtin_nontemporal_load(p + 3 * ns + s);
        double v4 = __builtin_nontemporal_load(reinterpret_cast<const double *>(p + 4 * ns) + s);
        a[0] = v0.x; a[1] = v0.y; a[2] = v1.x; a[3] = v1.y; a[4] = v2.x;
        a[5] = v2.y; a[6] = v3.x; a[7] = v3.y; a[8] = v4;
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const double *a)
    {
        double2 *p = reinterpret_cast<double2 *>(base);
        p[s] = make_double2(a[0], a[1]);
        p[ns + s] = make_double2(a[2], a[3]);
        p[2 * ns + s] = make_double2(a[4], a[5]);
        p[3 * ns + s] = make_double2(a[6], a[7]);
        reinterpret_cast<double *>(p + 4 * ns)[s] = a[8];
    }
    static constexpr size_t bytes_per_slot = 72;
};


// Split FP32 + bf16 values: hi = fl32(v), lo = bf16(v - hi), so hi + lo
// carries v to a relative error below 2^-32 in 6 bytes.  The hi planes come
// first (laid out as the FP32 planes), then the lo planes:
//   2x2: 4 bf16 (8 bytes) per slot          3x3: 8 bf16 (16 bytes) + 1 bf16
struct Split {};

__host__ __device__ inline uint16_t bf16_rne(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_value(uint32_t h)
{
    const uint32_t u = h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

typedef uint32_t u2v __attribute__((ext_vector_type(2)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));

template <> struct Blk<Split, 2> {
    typedef double T;
    static constexpr int N = 4;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, double *a)
    {
        const f4v h = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(base) + s);
        const u2v l = __builtin_nontemporal_load(reinterpret_cast<const u2v *>(reinterpret_cast<const char *>(base) + ns * 16) + s);
        a[0] = (double)h.x + (double)bf16_value(l.x & 0xFFFFu);
        a[1] = (double)h.y + (double)bf16_value(l.x >> 16);
        a[2] = (double)h.z + (double)bf16_value(l.y & 0xFFFFu);
        a[3] = (double)h.w + (double)bf16_value(l.y >> 16);
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const double *a)
    {
        float h[4];
        uint32_t lo[4];
        for (int i = 0; i < 4; i++) {
            h[i] = (float)a[i];
            lo[i] = bf16_rne((float)(a[i] - (double)h[i]));
        }
        reinterpret_cast<float4 *>(base)[s] = make_float4(h[0], h[1], h[2], h[3]);
        reinterpret_cast<uint2 *>(reinterpret_cast<char *>(base) + ns * 16)[s] = make_uint2(lo[0] | lo[1] << 16, lo[2] | lo[3] << 16);
    }
    static constexpr size_t bytes_per_slot = 24;
};
template <> struct Blk<Split, 3> {
    typedef double T;
    static constexpr int N = 9;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, double *a)
    {
        const f4v *p = reinterpret_cast<const f4v *>(base);
        const f4v h0 = __builtin_nontemporal_load(p + s);
        const f4v h1 = __builtin_nontemporal_load(p + ns + s);
        const float h2 = __builtin_nontemporal_load(reinterpret_cast<const float *>(p + 2 * ns) + s);
        const char *lb = reinterpret_cast<const char *>(base) + ns * 36;
        const u4v l0 = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(lb) + s);
        const uint16_t l1 = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(lb + ns * 16) + s);
        const float hv[9] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w, h2};
        const uint32_t lw[4] = {l0.x, l0.y, l0.z, l0.w};
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = (double)hv[i] + (double)bf16_value((lw[i / 2] >> (16 * (i % 2))) & 0xFFFFu);
        a[8] = (double)hv[8] + (double)bf16_value(l1);
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const double *a)
    {
        float h[9];
        uint32_t lo[9];
        for (int i = 0; i < 9; i++) {
            h[i] = (float)a[i];
            lo[i] = bf16_rne((float)(a[i] - (double)h[i]));
        }
        float4 *p = reinterpret_cast<float4 *>(base);
        p[s] = make_float4(h[0], h[1], h[2], h[3]);
        p[ns + s] = make_float4(h[4], h[5], h[6], h[7]);
        reinterpret_cast<float *>(p + 2 * ns)[s] = h[8];
        char *lb = reinterpret_cast<char *>(base) + ns * 36;
        reinterpret_cast<uint4 *>(lb)[s] = make_uint4(lo[0] | lo[1] << 16, lo[2] | lo[3] << 16, lo[4] | lo[5] << 16, lo[6] | lo[7] << 16);
        reinterpret_cast<uint16_t *>(lb + ns * 16)[s] = (uint16_t)lo[8];
    }
    static constexpr size_t bytes_per_slot = 54;
};

template <> struct BlkRaw<Split, 2> {
    typedef double T;
    struct Raw {
        f4v h;
        u2v l;
    };
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, Raw &r)
    {
        r.h = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(base) + s);
        r.l = __builtin_nontemporal_load(reinterpret_cast<const u2v *>(reinterpret_cast<const char *>(base) + ns * 16) + s);
    }
    __device__ static inline double at(const Raw &r, int k)
    {
        const float h = k == 0 ? r.h.x : k == 1 ? r.h.y : k == 2 ? r.h.z : r.h.w;
        const uint32_t w = k < 2 ? r.l.x : r.l.y;
        return (double)h + (double)bf16_value((w >> (16 * (k & 1))) & 0xFFFFu);
    }
};
template <> struct BlkRaw<Split, 3> {
    typedef double T;
    struct Raw {
        f4v h0, h1;
        float h2;
        u4v l0;
        uint32_t l1;
    };
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, Raw &r)
    {
        const f4v *p = reinterpret_cast<const f4v *>(base);
        r.h0 = __builtin_nontemporal_load(p + s);
        r.h1 = __builtin_nontemporal_load(p + ns + s);
        r.h2 = __builtin_nontemporal_load(reinterpret_cast<const float *>(p + 2 * ns) + s);
        const char *lb = reinterpret_cast<const char *>(base) + ns * 36;
        r.l0 = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(lb) + s);
        r.l1 = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(lb + ns * 16) + s);
    }
    __device__ static inline double at(const Raw &r, int k)
    {
        float h;
        uint32_t w;
        switch (k) {
        case 0: h = r.h0.x; w = r.l0.x; break;
        case 1: h = r.h0.y; w = r.l0.x >> 16; break;
        case 2: h = r.h0.z; w = r.l0.y; break;
        case 3: h = r.h0.w; w = r.l0.y >> 16; break;
        case 4: h = r.h1.x; w = r.l0.z; break;
        case 5: h = r.h1.y; w = r.l0.z >> 16; break;
        case 6: h = r.h1.z; w = r.l0.w; break;
        case 7: h = r.h1.w; w = r.l0.w >> 16; break;
        default: h = r.h2; w = r.l1; break;
        }
        return (double)h + (double)bf16_value(w & 0xFFFFu);
    }
};

// 48-bit fixed point: every entry of the scaled matrix Ks = S K S of an SPD
// K lies in (-1, 1] (unit diagonal, |Ks_ij| <= sqrt(Ks_ii Ks_jj)), so one
// global scale 2^-46 holds each value with an absolute error below 2^-47
// (7e-15) in 6 bytes: the low 32 bits in the first planes (laid out as the
// FP32 planes), the high 16 bits (signed) after them:
//   2x2: 4 x uint32 + 4 x int16          3x3: 8 + 1 x uint32, 8 + 1 x int16
struct Fix48 {};
constexpr double FIX48_SCALE = 70368744177664.0;        // 2^46
constexpr double FIX48_ULP = 1.0 / 70368744177664.0;    // 2^-46
constexpr double FIX48_LIMIT = 2.0;                     // |value| < 2 fits 48 bits

__host__ __device__ inline double fix48_value(uint32_t lo, uint32_t hi16)
{
    const double h = (double)(int32_t)(int16_t)(uint16_t)hi16;
    return (h * 4294967296.0 + (double)lo) * FIX48_ULP;
}
__host__ __device__ inline void fix48_encode(double v, uint32_t &lo, uint32_t &hi16)
{
    const int64_t q = (int64_t)llrint(v * FIX48_SCALE);
    lo = (uint32_t)(uint64_t)q;
    hi16 = (uint32_t)(uint16_t)(int16_t)(q >> 32);
}

template <> struct Blk<Fix48, 2> {
    typedef double T;
    static constexpr int N = 4;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, double *a)
    {
        const u4v l = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(base) + s);
        const u2v h = __builtin_nontemporal_load(reinterpret_cast<const u2v *>(reinterpret_cast<const char *>(base) + ns * 16) + s);
        a[0] = fix48_value(l.x, h.x & 0xFFFFu);
        a[1] = fix48_value(l.y, h.x >> 16);
        a[2] = fix48_value(l.z, h.y & 0xFFFFu);
        a[3] = fix48_value(l.w, h.y >> 16);
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const double *a)
    {
        uint32_t lo[4], hi[4];
        for (int i = 0; i < 4; i++) fix48_encode(a[i], lo[i], hi[i]);
        reinterpret_cast<uint4 *>(base)[s] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
        reinterpret_cast<uint2 *>(reinterpret_cast<char *>(base) + ns * 16)[s] = make_uint2(hi[0] | hi[1] << 16, hi[2] | hi[3] << 16);
    }
    static constexpr size_t bytes_per_slot = 24;
};
template <> struct Blk<Fix48, 3> {
    typedef double T;
    static constexpr int N = 9;
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, double *a)
    {
        const u4v *p = reinterpret_cast<const u4v *>(base);
        const u4v l0 = __builtin_nontemporal_load(p + s);
        const u4v l1 = __builtin_nontemporal_load(p + ns + s);
        const uint32_t l2 = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p + 2 * ns) + s);
        const char *hb = reinterpret_cast<const char *>(base) + ns * 36;
        const u4v h0 = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(hb) + s);
        const uint16_t h1 = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(hb + ns * 16) + s);
        const uint32_t lv[9] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w, l2};
        const uint32_t hw[4] = {h0.x, h0.y, h0.z, h0.w};
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = fix48_value(lv[i], (hw[i / 2] >> (16 * (i % 2))) & 0xFFFFu);
        a[8] = fix48_value(lv[8], h1);
    }
    __device__ static inline void store(void *base, int64_t ns, int64_t s, const double *a)
    {
        uint32_t lo[9], hi[9];
        for (int i = 0; i < 9; i++) fix48_encode(a[i], lo[i], hi[i]);
        uint4 *p = reinterpret_cast<uint4 *>(base);
        p[s] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
        p[ns + s] = make_uint4(lo[4], lo[5], lo[6], lo[7]);
        reinterpret_cast<uint32_t *>(p + 2 * ns)[s] = lo[8];
        char *hb = reinterpret_cast<char *>(base) + ns * 36;
        reinterpret_cast<uint4 *>(hb)[s] = make_uint4(hi[0] | hi[1] << 16, hi[2] | hi[3] << 16, hi[4] | hi[5] << 16, hi[6] | hi[7] << 16);
        reinterpret_cast<uint16_t *>(hb + ns * 16)[s] = (uint16_t)hi[8];
    }
    static constexpr size_t bytes_per_slot = 54;
};
template <> struct BlkRaw<Fix48, 2> {
    typedef double T;
    struct Raw {
        u4v l;
        u2v h;
    };
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, Raw &r)
    {
        r.l = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(base) + s);
        r.h = __builtin_nontemporal_load(reinterpret_cast<const u2v *>(reinterpret_cast<const char *>(base) + ns * 16) + s);
    }
    __device__ static inline double at(const Raw &r, int k)
    {
        const uint32_t lo = k == 0 ? r.l.x : k == 1 ? r.l.y : k == 2 ? r.l.z : r.l.w;
        const uint32_t w = k < 2 ? r.h.x : r.h.y;
        return fix48_value(lo, (w >> (16 * (k & 1))) & 0xFFFFu);
    }
};
template <> struct BlkRaw<Fix48, 3> {
    typedef double T;
    struct Raw {
        u4v l0, l1;
        uint32_t l2;
        u4v h0;
        uint32_t h1;
    };
    __device__ static inline void load(const void *base, int64_t ns, int64_t s, Raw &r)
    {
        const u4v *p = reinterpret_cast<const u4v *>(base);
        r.l0 = __builtin_nontemporal_load(p + s);
        r.l1 = __builtin_nontemporal_load(p + ns + s);
        r.l2 = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p + 2 * ns) + s);
        const char *hb = reinterpret_cast<const char *>(base) + ns * 36;
        r.h0 = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(hb) + s);
        r.h1 = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(hb + ns * 16) + s);
    }
    __device__ static inline double at(const Raw &r, int k)
    {
        uint32_t lo, w;
        switch (k) {
        case 0: lo = r.l0.x; w = r.h0.x; break;
        case 1: lo = r.l0.y; w = r.h0.x >> 16; break;
        case 2: lo = r.l0.z; w = r.h0.y; break;
        case 3: lo = r.l0.w; w = r.h0.y >> 16; break;
        case 4: lo = r.l1.x; w = r.h0.z; break;
        case 5: lo = r.l1.y; w = r.h0.z >> 16; break;
        case 6: lo = r.l1.z; w = r.h0.w; break;
        case 7: lo = r.l1.w; w = r.h0.w >> 16; break;
        default: lo = r.l2; w = r.h1; break;
        }
        return fix48_value(lo, w & 0xFFFFu);
    }
};

// Matrix value kinds: FP64 planes hold the unscaled K (the scaling is applied
// on the fly); FP32 and split planes hold the scaled Ks.  FP64 and split
// products accumulate in FP64.
template <typename M> struct MatKind {
    static constexpr bool onfly = false, acc64 = false;
};
template <> struct MatKind<double> {
    static constexpr bool onfly = true, acc64 = true;
};

// FP64 planes of the scaled Ks = S K S, the FP64 CG's matrix.  Scaling on
// the fly gathers S beside p for every block column (32 bytes a 2x2 column
// instead of 16) and that second gather is what the P2 meshes, whose
// columns spread widely, re-read from HBM; the copy costs one more FP64
// matrix in HBM, built once per assembly when an FP64 solve asks for it.
struct Pre64 {};
template <int B> struct Blk<Pre64, B> : Blk<double, B> {};
template <> struct MatKind<Pre64> {
    static constexpr bool onfly = false, acc64 = true;
};
template <> struct MatKind<Split> {
    static constexpr bool onfly = false, acc64 = true;
};
template <> struct MatKind<Fix48> {
    static constexpr bool onfly = false, acc64 = true;
};

// Value-indexed FIX48 matrix: structured meshes hold few distinct blocks
// (the 512-RVE batch: about 10^3 among 9 x 10^7), so each slot stores a
// 16-bit index into a table of the distinct FIX48 blocks, decoded to FP64.
// The iteration then reads 2 bytes a block from HBM instead of 24 (2x2) or
// 54 (3x3); the table stays in the caches.  The values are bitwise those of
// the FIX48 planes, so are the results.
struct Dict {};
template <int B> struct Blk<Dict, B> {
    typedef double T;
    static constexpr int N = B * B;
    static constexpr size_t bytes_per_slot = 2;
};
template <> struct MatKind<Dict> {
    static constexpr bool onfly = false, acc64 = true;
};
constexpr int DICT_MAX = 65536;   // table entries addressable by a 16-bit index
// doubles per table entry: b*b, padded to a 16-byte multiple
template <int B> struct DictEntry {
    static constexpr int N = B == 2 ? 4 : 10;
};

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Sums v[0..NQ) over the workgroup; every thread gets the totals.
template <int NQ> __device__ inline void wg_sum_all(double *v, double (*lds)[MAXQ])
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double s = wave_sum(v[q]);
        if (lane == 0) lds[wave][q] = s;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = (lds[0][q] + lds[1][q]) + (lds[2][q] + lds[3][q]);
}

// Writes this tile's partial sums (from thread 0 after a workgroup sum).
template <int NQ> __device__ inline void write_partials(double *v, double *part, int tile, double (*lds)[MAXQ])
{
    wg_sum_all<NQ>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NQ; q++) part[(size_t)tile * MAXQ + q] = v[q];
}

// Sums the tile partials of problem P in tile order: the same bits on every
// run, whatever the launch shape.
template <int NQ> __device__ inline void problem_sums(const double *part, const ProbDev &P, double *out,
                                                      double (*lds)[MAXQ])
{
#pragma unroll
    for (int q = 0; q < NQ; q++) out[q] = 0.;
    for (int t = threadIdx.x; t < P.ntiles; t += WG)
#pragma unroll
        for (int q = 0; q < NQ; q++) out[q] += part[(size_t)(P.tile0 + t) * MAXQ + q];
    wg_sum_all<NQ>(out, lds);
}

// problem_sums for two partial arrays at once (NA values from pa, NB from
// pb), with the loads of four tiles in flight per thread: the same additions
// in the same order as problem_sums, one round of barriers instead of two.
template <int NA, int NB>
__device__ inline void problem_sums2(const double *pa, const double *pb, const ProbDev &P, double *oa, double *ob,
                                     double (*lds)[2 * MAXQ])
{
    constexpr int NQ = NA + NB;
    double out[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) out[q] = 0.;
    const int nt = P.ntiles;
    const size_t t0 = (size_t)P.tile0 * MAXQ;
    int t = threadIdx.x;
    for (; t + 3 * WG < nt; t += 4 * WG) {
        double v[4][NQ];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const size_t o = t0 + (size_t)(t + u * WG) * MAXQ;
#pragma unroll
            for (int q = 0; q < NA; q++) v[u][q] = pa[o + q];
#pragma unroll
            for (int q = 0; q < NB; q++) v[u][NA + q] = pb[o + q];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int q = 0; q < NQ; q++) out[q] += v[u][q];
    }
    for (; t < nt; t += WG) {
        const size_t o = t0 + (size_t)t * MAXQ;
#pragma unroll
        for (int q = 0; q < NA; q++) out[q] += pa[o + q];
#pragma unroll
        for (int q = 0; q < NB; q++) out[NA + q] += pb[o + q];
    }
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double sw = wave_sum(out[q]);
        if (lane == 0) lds[wave][q] = sw;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) out[q] = (lds[0][q] + lds[1][q]) + (lds[2][q] + lds[3][q]);
#pragma unroll
    for (int q = 0; q < NA; q++) oa[q] = out[q];
#pragma unroll
    for (int q = 0; q < NB; q++) ob[q] = out[NA + q];
}

// ------------------------------------------------------- tile scheduling
// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8).  In a
// one-workgroup-per-tile launch each XCD takes a contiguous range of the
// active tiles, so the neighbour rows a tile gathers were mostly brought
// into the same L2 by nearby tiles.  A bijection of [0, gridDim.x).
__device__ inline int xcd_index()
{
    const int G = gridDim.x, b = blockIdx.x, x = b % NXCD, per = G / NXCD, rem = G % NXCD;
    return x * per + (x < rem ? x : rem) + b / NXCD;
}

// The wavefront index as a scalar (uniform) value.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE); }

// ------------------------------------------------------------ vector nodes
// The b components of one right-hand side of one row: a float2 / double2
// (b = 2) or three scalars (b = 3).  NT: streaming (non-temporal) access.
template <typename W, int B, bool NT = false> struct Node {
    __device__ static inline void load(const W *v, W *o)
    {
        if constexpr (B == 2) {
            typedef W v2 __attribute__((ext_vector_type(2)));
            const v2 t = NT ? __builtin_nontemporal_load(reinterpret_cast<const v2 *>(v)) : *reinterpret_cast<const v2 *>(v);
            o[0] = t.x;
            o[1] = t.y;
        } else {
#pragma unroll
            for (int j = 0; j < B; j++) o[j] = NT ? __builtin_nontemporal_load(v + j) : v[j];
        }
    }
    __device__ static inline void store(W *v, const W *o)
    {
        if constexpr (B == 2) {
            typedef W v2 __attribute__((ext_vector_type(2)));
            v2 t;
            t.x = o[0];
            t.y = o[1];
            if (NT) __builtin_nontemporal_store(t, reinterpret_cast<v2 *>(v));
            else *reinterpret_cast<v2 *>(v) = t;
        } else {
#pragma unroll
            for (int j = 0; j < B; j++) {
                if (NT) __builtin_nontemporal_store(o[j], v + j);
                else v[j] = o[j];
            }
        }
    }
};

template <typename W> __device__ inline W *vw(void *p) { return reinterpret_cast<W *>(p); }
template <typename W> __device__ inline const W *cvw(const void *p) { return reinterpret_cast<const W *>(p); }
// the plane set of p_k
__device__ inline void *p_planes(const SolveArgs &a, int64_t k) { return a.pset[k & (YGROUP - 1)]; }

// z = M v with M the row's inverse diagonal block (BJ) or the identity.
template <typename W, int B, bool BJ> __device__ inline void precond(const W *mi, const W *v, W *z)
{
#pragma unroll
    for (int i = 0; i < B; i++) {
        if (BJ) {
            W s = 0;
#pragma unroll
            for (int j = 0; j < B; j++) s += mi[i * B + j] * v[j];
            z[i] = s;
        } else {
            z[i] = v[i];
        }
    }
}

template <typename W, int B, bool BJ> __device__ inline void load_minv(const SolveArgs &a, int64_t row, W *mi)
{
    if (BJ) {
        const W *m = cvw<W>(a.minv) + (size_t)row * B * B;
#pragma unroll
        for (int k = 0; k < B * B; k++) mi[k] = __builtin_nontemporal_load(m + k);
    }
}

// ------------------------------------------------------------ SpMV core
// Blocks requested per step: their column indices and values are all loaded
// before the dependent gathers, so a wavefront keeps many independent loads
// in flight (a P1 row of 7 blocks is one step with 2x2 FP32 blocks).
#ifndef AMIE_UNROLL_16
#define AMIE_UNROLL_16 8      // blocks per step for slots of <= 16 bytes
#endif
#ifndef AMIE_UNROLL_36
#define AMIE_UNROLL_36 4      // ... of <= 36 bytes (split 2x2, FP32 3x3, FP64 2x2)
#endif
template <typename M, int B> struct Unroll {
    static constexpr size_t bytes = Blk<M, B>::bytes_per_slot;
    static constexpr int U = bytes <= 16 ? AMIE_UNROLL_16 : (bytes <= 36 ? AMIE_UNROLL_36 : 2);
};

// acc[k][i] += sum over the row's blocks of A(row, c)(i, j) * x_k(c)_j for
// lane `lane` of slice s; gather(c, xv) supplies x_k(c) for every
// right-hand side k.  The slice width is uniform over the wavefront, so the
// step tails are scalar branches.
template <typename A, typename M, int B, int NR, typename G>
__device__ inline void spmv_row(const SolveArgs &a, const void *vals, int s, int lane, A (*acc)[B], G gather)
{
    constexpr int U = Unroll<M, B>::U;
    const int w = a.slice_width[s];
    const int64_t sp = a.slice_ptr[s] + lane;
    for (int k = 0; k < w; k += U) {
        typedef BlkRaw<M, B> BR;
        uint32_t c[U];
        typename BR::Raw blk[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) c[u] = __builtin_nontemporal_load(a.col + sp + (int64_t)(k + u) * WAVE);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) BR::load(vals, a.nslots, sp + (int64_t)(k + u) * WAVE, blk[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) {
                A xv[NR][B];
                gather(c[u], xv);
#pragma unroll
                for (int j = 0; j < B; j++)
#pragma unroll
                    for (int i = 0; i < B; i++) {
                        const A m = (A)BR::at(blk[u], j * B + i);
#pragma unroll
                        for (int r = 0; r < NR; r++) acc[r][i] += m * xv[r][j];
                    }
            }
    }
}

// Chunked order: the XCD of workgroup b (b mod 8) takes chunks of C
// consecutive tiles, chunk after chunk round-robin over the XCDs; workgroups
// past the last full round keep launch order.  A bijection of
// [0, gridDim.x); C = 1 is launch order.
__device__ inline int chunk_index(int C)
{
    const int G = gridDim.x, b = blockIdx.x, span = NXCD * C, M = G / span * span;
    if (C == 1 || b >= M) return b;
    const int x = b % NXCD, j = b / NXCD;
    return (j / C * NXCD + x) * C + j % C;
}

// The tile of a one-workgroup-per-tile launch, or -1.  With every tile
// active the list lookup is skipped.  tile_order (picked by the autotuner)
// chooses the XCD-contiguous or a chunked order.
__device__ inline int my_tile(const SolveArgs &a)
{
    const int it = a.tile_order ? chunk_index(a.tile_order) : xcd_index();
    const int na = *a.n_active;
    if (it >= na) return -1;
    return na == a.ntiles_all ? it : a.active[it];
}

// Row loop of one slice from its packed descriptor (same arithmetic as
// spmv_row).  slot_off holds, per block position of the slice, the column
// offset that every one of its 64 lanes has (column = row + offset), or
// NO_OFFSET.  A group of positions that all have one gathers at once,
// without loading the columns; any other group loads its columns first.  On
// structured meshes most positions have an offset, which saves the column
// bytes and takes the column load off the gather's critical path.
constexpr int32_t NO_OFFSET = INT32_MIN;
template <typename A, typename M, int B, int NR, int U, typename G>
__device__ inline void spmv_slice(const SolveArgs &a, const void *vals, int64_t sp0, int w, int lane, int64_t row,
                                  A (*acc)[B], G gather)
{
    constexpr bool IS_DICT = std::is_same<M, Dict>::value;
    typedef BlkRaw<typename std::conditional<IS_DICT, double, M>::type, B> BR;
    const int64_t sp = sp0 + lane;
    const int32_t *offs = a.slot_off + (sp0 >> 6);
    for (int k = 0; k < w; k += U) {
        int32_t o[U];
        bool direct = true;
#pragma unroll
        for (int u = 0; u < U; u++) {
            o[u] = k + u < w ? offs[k + u] : 0;
            direct &= o[u] != NO_OFFSET;
        }
        uint32_t c[U], id[U];
        typename BR::Raw blk[U];
        if (direct) {
#pragma unroll
            for (int u = 0; u < U; u++) c[u] = (uint32_t)(row + o[u]);
        } else {
#pragma unroll
            for (int u = 0; u < U; u++)
                if (k + u < w) c[u] = __builtin_nontemporal_load(a.col + sp + (int64_t)(k + u) * WAVE);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) {
                if constexpr (IS_DICT)
                    id[u] = __builtin_nontemporal_load(reinterpret_cast<const uint16_t *>(vals) + sp + (int64_t)(k + u) * WAVE);
                else
                    BR::load(vals, a.nslots, sp + (int64_t)(k + u) * WAVE, blk[u]);
            }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) {
                A xv[NR][B];
                gather(c[u], xv);
                double dv[IS_DICT ? B * B : 1];
                if constexpr (IS_DICT) {
                    const d2v *e = reinterpret_cast<const d2v *>(a.dict_table + (size_t)id[u] * DictEntry<B>::N);
#pragma unroll
                    for (int h = 0; h < B * B / 2; h++) {
                        const d2v t = e[h];
                        dv[2 * h] = t.x;
                        dv[2 * h + 1] = t.y;
                    }
                    if constexpr (B * B % 2) dv[B * B - 1] = reinterpret_cast<const double *>(e)[B * B - 1];
                }
#pragma unroll
                for (int j = 0; j < B; j++)
#pragma unroll
                    for (int i = 0; i < B; i++) {
                        A m;
                        if constexpr (IS_DICT) m = (A)dv[j * B + i];
                        else m = (A)BR::at(blk[u], j * B + i);
#pragma unroll
                        for (int r = 0; r < NR; r++) acc[r][i] += m * xv[r][j];
                    }
            }
    }
}

// ------------------------------------------------------------ K1: SpMV
// q = Ks p with partial sums of p.q, q.z and q.Mq (z = M r).  With the FP64
// matrix the unscaled K is used and the scaling applied on the fly:
// q = S K (S p).  With AMIE_EARLY_OWN the wave's own p and r are requested
// with its first matrix loads (the row loop's latency covers them, at the
// cost of registers); otherwise after the row loop.
#ifndef AMIE_EARLY_OWN
#define AMIE_EARLY_OWN 0
#endif
// Body of K1 for one tile with matrix kind M (values at `vals`), U blocks
// per step.
template <typename W, typename M, int B, int NR, bool BJ, int U>
__device__ inline void spmv_tile(const SolveArgs &a, const void *vals, int t, const int4 &td, int wave, int lane)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    constexpr bool ONFLY = MatKind<M>::onfly;
    // products accumulate in FP64 with the FP64 and split matrices, else in W
    typedef typename std::conditional<MatKind<M>::acc64, double, W>::type A;
    const bool has = wave < td.y;
    const int4 sd = a.slice_desc[has ? td.x + wave : td.x];
    const W *p = cvw<W>(p_planes(a, a.prob[td.z].iters)), *r = cvw<W>(a.r);
    W *q = vw<W>(a.q);
    double part[3 * NR];
#pragma unroll
    for (int k = 0; k < 3 * NR; k++) part[k] = 0.;
    if (has) {
        const int s = td.x + wave;
        const int64_t row = (int64_t)s * WAVE + lane;
        const bool own = lane < sd.w;
        W pv[NR][B], rv[NR][B];
        auto load_own = [&]() {
#pragma unroll
            for (int k = 0; k < NR; k++) {
                const size_t idx = k * a.vplane + (size_t)row * B;
                Node<W, B>::load(p + idx, pv[k]);
                Node<W, B, true>::load(r + idx, rv[k]);
            }
        };
        if (AMIE_EARLY_OWN && own) load_own();
        A acc[NR][B];
#pragma unroll
        for (int k = 0; k < NR; k++)
#pragma unroll
            for (int i = 0; i < B; i++) acc[k][i] = 0;
        auto gather = [&](uint32_t c, A (*xv)[B]) {
            double sc[B];
            if (ONFLY) Node<double, B>::load(a.S + (size_t)c * B, sc);
#pragma unroll
            for (int k = 0; k < NR; k++) {
                W tv[B];
                Node<W, B>::load(p + k * a.vplane + (size_t)c * B, tv);
#pragma unroll
                for (int j = 0; j < B; j++) xv[k][j] = ONFLY ? (A)tv[j] * (A)sc[j] : (A)tv[j];
            }
        };
        const int64_t sp0 = (int64_t)(uint32_t)sd.x | ((int64_t)sd.y << 32);
        spmv_slice<A, M, B, NR, U>(a, vals, sp0, sd.z, lane, row, acc, gather);
        if (own) {
            if (!AMIE_EARLY_OWN) load_own();
            W mi[B * B];
            load_minv<W, B, BJ>(a, row, mi);
            double sr[B];
            if (ONFLY) Node<double, B>::load(a.S + (size_t)row * B, sr);
#pragma unroll
            for (int k = 0; k < NR; k++) {
                const size_t idx = k * a.vplane + (size_t)row * B;
                W qv[B], zv[B], mq[B];
#pragma unroll
                for (int i = 0; i < B; i++) qv[i] = ONFLY ? (W)(acc[k][i] * (A)sr[i]) : (W)acc[k][i];
                Node<W, B, true>::store(q + idx, qv);
                precond<W, B, BJ>(mi, rv[k], zv);
                precond<W, B, BJ>(mi, qv, mq);
#pragma unroll
                for (int i = 0; i < B; i++) {
                    part[k] += (double)pv[k][i] * (double)qv[i];
                    part[NR + k] += (double)qv[i] * (double)zv[i];
                    part[2 * NR + k] += (double)qv[i] * (double)mq[i];
                }
            }
        }
    }
    write_partials<3 * NR>(part, a.part1, t, lds);
}

// ------------------------------------------------------------ K1: SpMV
// q = Ks p with partial sums of p.q, q.z and q.Mq (z = M r).  With the FP64
// matrix the unscaled K is used and the scaling applied on the fly:
// q = S K (S p).  A problem on the split matrix whose reduce kernel found its
// accuracy floor (kmode) runs the same product from the FP64 matrix.
template <typename W, typename M, int B, int NR, bool BJ, int U = Unroll<M, B>::U>
__device__ inline void k1_body(const SolveArgs &a, int t)
{
    const int4 td = a.tile_desc[t];
    const ProbDev &P = a.prob[td.z];
    if (P.status != ST_RUNNING) return;
    const int wave = wave_id(), lane = threadIdx.x % WAVE;
    if (std::is_same<M, Split>::value && P.kmode)
        spmv_tile<W, double, B, NR, BJ, Unroll<double, B>::U>(a, a.v64, t, td, wave, lane);
    else spmv_tile<W, M, B, NR, BJ, U>(a, a.vw, t, td, wave, lane);
}

template <typename W, typename M, int B, int NR, bool BJ, int U = Unroll<M, B>::U>
__global__ __launch_bounds__(WG) void k_spmv(SolveArgs a)
{
    const int t = my_tile(a);
    if (t < 0) return;
    k1_body<W, M, B, NR, BJ, U>(a, t);
}

// ------------------------------------------- R: per-problem CG scalars
// One workgroup per active problem.  Consumes the residual partials of the
// previous update (or refinement) and this iteration's SpMV partials.
template <int NR>
__device__ inline void reduce_body(const SolveArgs &a, int pi)
{
    __shared__ double lds2[TILE_SLICES][2 * MAXQ];
    ProbDev &P = a.prob[a.active_prob[pi]];
    if (P.status != ST_RUNNING) return;   // read by every thread before thread 0 writes (barriers below)
    const bool ck = P.src_ckpt != 0;
    double s1[3 * NR], s0[4 * NR];
    if (ck) problem_sums2<3 * NR, 4 * NR>(a.part1, a.partc, P, s1, s0, lds2);
    else problem_sums2<3 * NR, 2 * NR>(a.part1, a.part2, P, s1, s0, lds2);
    if (threadIdx.x != 0) return;
    bool all = true, bad = false, stalled = false, need = false, ladder_up = false;
    for (int k = 0; k < NR; k++) {
        if (P.done[k]) {
            P.alpha[k] = P.beta[k] = 0.;
            continue;
        }
        const double rho = s0[k], rr = s0[NR + k];
        if (ck) {
            // accuracy floor: the true (replaced) rho well above the
            // recurrence's prediction for the same residual means the working
            // precision can no longer follow the iteration.  The split matrix
            // hands over to the FP64 matrix; otherwise the problem is stalled
            // (and continues in FP64 when escalation is on).
            if (P.rho_pred[k] > 0. && s0[k] > 4. * P.rho_pred[k]) {
                if (P.ladder && !P.kmode) ladder_up = true;
                else if (P.stall_iters > 0) stalled = true;
            }
            // the refinement's FP64 residual decides convergence
            const double rr64 = s0[2 * NR + k];
            if (!(rr64 == rr64)) bad = true;
            const double rel = P.bb64[k] > 0. ? sqrt(rr64 / P.bb64[k]) : sqrt(rr64);
            P.relres[k] = P.relres_out[k] = rel;
            P.rr_ckpt[k] = rr;
            // the stop measure and its bound: relative FP64 residual, or the
            // reference's absolute sqrt(r . D^-1 r) = ||S r64||
            const double meas = P.ref_eps > 0. ? sqrt(s0[3 * NR + k]) : rel;
            const double bound = P.ref_eps > 0. ? fmax(1e-12, P.ref_eps) : P.rtol;
            if (meas <= bound) {
                P.done[k] = P.done_out[k] = 1;
                P.alpha[k] = P.beta[k] = 0.;
                continue;
            }
            // expected working-precision r.r at convergence, from this ratio
            P.rr_target[k] = rr * (bound / meas) * (bound / meas);
            if (rel < 0.5 * P.rel_best[k]) {
                P.rel_best[k] = rel;
                P.best_it[k] = P.iters;
            } else if (P.stall_iters > 0 && P.iters - P.best_it[k] >= P.stall_iters) {
                stalled = true;
            }
        }
        all = false;
        const double pq = s1[k], qz = s1[NR + k], qmq = s1[2 * NR + k];
        if (!(pq == pq) || !(rho == rho) || !(qz == qz) || !(qmq == qmq)) bad = true;
        const double alpha = pq != 0. ? rho / pq : 0.;
        const double rho1 = rho - 2. * alpha * qz + alpha * alpha * qmq;
        P.alpha[k] = alpha;
        P.alpha_ring[k][P.iters & (YGROUP - 1)] = alpha;
        P.beta[k] = rho != 0. ? rho1 / rho : 0.;
        P.rho_pred[k] = rho1;
        P.rr[k] = rr;
        if (rr < P.drop * P.drop * P.rr_ckpt[k]) need = true;   // FP32 residual dropped: refine
        if (rr <= P.rr_target[k]) need = true;                   // may have converged: confirm in FP64
    }
    if (ladder_up) P.kmode = 1;
    if (P.kmode) P.iters_k64 += 1;
    if (ck) {
        P.nckpt += 1;
        P.last_ckpt = P.iters;
        P.ckpt_pending = 0;
        P.src_ckpt = 0;
        P.first = 0;
    }
    if (bad) { P.status = ST_NAN; return; }
    if (all) { P.status = ST_CONVERGED; return; }
    if (ck && P.iters >= P.maxit) { P.status = ST_MAXIT; return; }
    if (stalled) { P.status = ST_STALLED; return; }
    // the previous update deferred its y term (one more pending) or added
    // every pending one; a refinement added everything
    P.ypend = (ck || P.ypend == YGROUP - 1) ? 0 : P.ypend + 1;
    P.iters += 1;
    if (P.iters >= P.maxit || P.iters - P.last_ckpt >= P.ckpt_interval) need = true;
    if (need) P.ckpt_pending = 1;
}

template <int NR>
__global__ __launch_bounds__(WG) void k_reduce(SolveArgs a)
{
    if ((int)blockIdx.x >= *a.n_active_prob) return;
    reduce_body<NR>(a, (int)blockIdx.x);
}

// ------------------------------------ K2: y, r, p update + dots (streaming)
// slot != 0: a checkpoint follows this iteration; a problem with a pending
// refinement then only completes its y update (C2 replaces r and forms p).
template <typename W, int B, int NR, bool BJ>
__device__ inline void k2_body(const SolveArgs &a, int t, int slot)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    const ProbDev &P = a.prob[a.tile_prob[t]];
    if (P.status != ST_RUNNING) return;
    const bool ck = slot && P.ckpt_pending;
    // k, the iteration being completed: the reduce kernel has counted it.
    // p_k is read from its plane set and p_{k+1} written to the set of
    // p_{k+1-YGROUP}.  y += alpha p is grouped: an update with fewer than
    // YGROUP - 1 pending terms leaves y alone; the next one (or a
    // refinement's slot update) adds alpha_j p_j for the pending j in order
    // and then alpha_k p_k: the same FP64 operations in the same order as one
    // add per iteration.
    const int64_t k_it = P.iters - 1;
    const int pend = P.ypend;
    const bool flush = pend == YGROUP - 1 || ck;
    const int wave = wave_id(), lane = threadIdx.x % WAVE;
    const W *pk = cvw<W>(p_planes(a, k_it));
    W *pn = vw<W>(p_planes(a, k_it + 1));
    W *r = vw<W>(a.r);
    double *y = vw<double>(a.y);   // the iterate since the last refinement: FP64 in every mode
    const W *q = cvw<W>(a.q);
    double part[2 * NR];
#pragma unroll
    for (int k = 0; k < 2 * NR; k++) part[k] = 0.;
    if (wave < a.tile_nslices[t]) {
        const int s = a.tile_slice0[t] + wave;
        const int64_t row = (int64_t)s * WAVE + lane;
        if (lane < a.slice_rows[s]) {
            W mi[B * B];
            if (!ck) load_minv<W, B, BJ>(a, row, mi);
#pragma unroll
            for (int k = 0; k < NR; k++) {
                if (P.done[k]) continue;
                const W al = (W)P.alpha[k], be = (W)P.beta[k];
                const double al64 = P.alpha[k];
                const size_t idx = k * a.vplane + (size_t)row * B;
                W pv[B];
                Node<W, B, true>::load(pk + idx, pv);
                if (flush) {
                    double yv[B];
                    Node<double, B, true>::load(y + idx, yv);
                    for (int j = pend; j >= 1; j--) {
                        W pp[B];
                        Node<W, B, true>::load(cvw<W>(p_planes(a, k_it - j)) + idx, pp);
                        const double alp = P.alpha_ring[k][(k_it - j) & (YGROUP - 1)];
#pragma unroll
                        for (int i = 0; i < B; i++) yv[i] += alp * (double)pp[i];
                    }
#pragma unroll
                    for (int i = 0; i < B; i++) yv[i] += al64 * (double)pv[i];
                    Node<double, B, true>::store(y + idx, yv);
                }
                if (ck) continue;
                W qv[B], rv[B], zv[B];
                Node<W, B, true>::load(q + idx, qv);
                Node<W, B, true>::load(r + idx, rv);
#pragma unroll
                for (int i = 0; i < B; i++) rv[i] -= al * qv[i];
                Node<W, B, true>::store(r + idx, rv);
                precond<W, B, BJ>(mi, rv, zv);
#pragma unroll
                for (int i = 0; i < B; i++) {
                    pv[i] = zv[i] + be * pv[i];
                    part[k] += (double)rv[i] * (double)zv[i];
                    part[NR + k] += (double)rv[i] * (double)rv[i];
                }
                Node<W, B, true>::store(pn + idx, pv);
            }
        }
    }
    if (!ck) write_partials<2 * NR>(part, a.part2, t, lds);
}

template <typename W, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_update(SolveArgs a, int slot)
{
    const int t = my_tile(a);
    if (t < 0) return;
    k2_body<W, B, NR, BJ>(a, t, slot);
}

// ------------------------------------------- refinement 1: x64 += S y, y = 0
template <typename W, int B, int NR>
__device__ inline void cx_body(const SolveArgs &a, int t)
{
    const int pi = a.tile_prob[t];
    const ProbDev &P = a.prob[pi];
    if (P.status != ST_RUNNING || !P.ckpt_pending) return;
    if (t == P.tile0 && threadIdx.x == 0) a.prob[pi].src_ckpt = 1;
    const int wave = wave_id(), lane = threadIdx.x % WAVE;
    if (wave >= a.tile_nslices[t]) return;
    const int s = a.tile_slice0[t] + wave;
    if (lane >= a.slice_rows[s]) return;
    const int64_t row = (int64_t)s * WAVE + lane;
    double *y = vw<double>(a.y);
    double sr[B];
    Node<double, B>::load(a.S + (size_t)row * B, sr);
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const size_t idx = k * a.vplane + (size_t)row * B;
        double yv[B], xv[B];
        Node<double, B>::load(y + idx, yv);
        Node<double, B>::load(a.x64 + idx, xv);
#pragma unroll
        for (int i = 0; i < B; i++) {
            xv[i] += sr[i] * yv[i];
            yv[i] = 0;
        }
        Node<double, B>::store(a.x64 + idx, xv);
        Node<double, B>::store(y + idx, yv);
    }
}

template <typename W, int B, int NR>
__global__ __launch_bounds__(WG) void k_ckpt_x(SolveArgs a)
{
    const int t = my_tile(a);
    if (t < 0) return;
    cx_body<W, B, NR>(a, t);
}

// ------------- refinement 2: r64 = b - K x64 (FP64), r = S r64, p = M r + beta p
template <typename W, int B, int NR, bool BJ>
__device__ inline void cr_body(const SolveArgs &a, int t)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    const ProbDev &P = a.prob[a.tile_prob[t]];
    if (P.status != ST_RUNNING || !P.ckpt_pending) return;
    const int wave = wave_id(), lane = threadIdx.x % WAVE;
    // p_{k+1} = M r + beta p_k: from p_k's plane set to the next one (k + 1 =
    // iters, counted by the reduce kernel; at a pass's opening refinement
    // iters = 0 and both sets hold zeros)
    W *r = vw<W>(a.r);
    const W *pk = cvw<W>(p_planes(a, P.iters - 1));
    W *pn = vw<W>(p_planes(a, P.iters));
    double part[4 * NR];
#pragma unroll
    for (int k = 0; k < 4 * NR; k++) part[k] = 0.;
    if (wave < a.tile_nslices[t]) {
        const int s = a.tile_slice0[t] + wave;
        const int64_t row = (int64_t)s * WAVE + lane;
        double acc[NR][B];
#pragma unroll
        for (int k = 0; k < NR; k++)
#pragma unroll
            for (int i = 0; i < B; i++) acc[k][i] = 0.;
        auto gather = [&](uint32_t c, double (*xv)[B]) {
#pragma unroll
            for (int k = 0; k < NR; k++) Node<double, B>::load(a.x64 + k * a.vplane + (size_t)c * B, xv[k]);
        };
        spmv_row<double, double, B, NR>(a, a.v64, s, lane, acc, gather);
        if (lane < a.slice_rows[s]) {
            W mi[B * B];
            load_minv<W, B, BJ>(a, row, mi);
            double sr[B];
            Node<double, B>::load(a.S + (size_t)row * B, sr);
            const bool restart = P.first || P.refine == 1;
#pragma unroll
            for (int k = 0; k < NR; k++) {
                const size_t idx = k * a.vplane + (size_t)row * B;
                const W be = restart ? (W)0 : (W)P.beta[k];
                double bv[B];
                Node<double, B>::load(a.b64 + idx, bv);
                W rv[B], zv[B], pv[B];
#pragma unroll
                for (int i = 0; i < B; i++) {
                    const double r64 = bv[i] - acc[k][i];
                    const double rs = sr[i] * r64;
                    part[2 * NR + k] += r64 * r64;
                    part[3 * NR + k] += rs * rs;
                    rv[i] = (W)rs;
                }
                Node<W, B>::store(r + idx, rv);
                precond<W, B, BJ>(mi, rv, zv);
                Node<W, B>::load(pk + idx, pv);
#pragma unroll
                for (int i = 0; i < B; i++) {
                    pv[i] = zv[i] + be * pv[i];
                    part[k] += (double)rv[i] * (double)zv[i];
                    part[NR + k] += (double)rv[i] * (double)rv[i];
                }
                Node<W, B>::store(pn + idx, pv);
            }
        }
    }
    write_partials<4 * NR>(part, a.partc, t, lds);
}

template <typename W, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_ckpt_r(SolveArgs a)
{
    const int t = my_tile(a);
    if (t < 0) return;
    cr_body<W, B, NR, BJ>(a, t);
}

// ------------------------------------------------------ persistent solve
// One launch runs whole checkpoint periods: the phases of queue_period (K1,
// R, K2 per iteration, then C1 and C2) as loops over the same per-tile and
// per-problem bodies, separated by device-wide barriers instead of kernel
// boundaries.  The results are those of the launch path bit for bit: the
// same bodies in the same order, the same sums.  Meant for small groups:
// a barrier among G workgroups costs about 2-3 us at G = 64 and grows with
// G (one arrival counter), while a kernel boundary costs about the same at
// any size.
struct PersistArgs {
    unsigned long long *bar;   // barrier arrivals (only grows), zeroed before the launch
    int *err;                  // set when a barrier wait passes its bound; every workgroup then returns
    int *periods;              // checkpoint periods run (workgroup 0)
    int ce;                    // iterations per checkpoint period
    int max_periods;
};

// bound of a barrier wait, in polls (each an L2 round trip plus a short
// sleep): seconds, far beyond any phase of a group small enough for the
// persistent path; the wait then gives up instead of hanging the device
constexpr unsigned PERSIST_SPIN_MAX = 1u << 22;

// Device-wide barrier of the persistent solve: a release (the workgroup's
// writes visible device-wide: L2 write-back on CDNA, whose XCD L2s are not
// coherent), one arrival on the counter, a bounded wait until all gridDim.x
// workgroups of this round arrived, an acquire.  Returns false (every
// thread) when the wait gave up or another workgroup's did.
__device__ inline bool persist_barrier(const PersistArgs &pa, unsigned long long &round)
{
    __shared__ int ok_s;
    round++;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(pa.bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long target = round * gridDim.x;
        unsigned spins = 0;
        while (__hip_atomic_load(pa.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > PERSIST_SPIN_MAX || __hip_atomic_load(pa.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(pa.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_s = ok;
    }
    __syncthreads();
    return ok_s != 0;
}

// ------------------------------------------- one-workgroup CG (small meshes)
// The whole FP64 CG of one problem in one workgroup of WG1 threads: the
// phases of an iteration are separated by workgroup barriers instead of
// kernel boundaries or device-wide barriers, which is what the solves of a
// few thousand block rows are bound by (the passes of the event-driven
// damage).  The scaled system S K S (symmetric Jacobi, point Jacobi after
// scaling) on the unscaled FP64 planes, right-hand side 0; x64 in place
// (warm start when use_x0).  The recurrence residual decides when to check,
// the FP64 relative residual ||b - K x|| / ||b|| <= rtol decides
// convergence, and the checked residual replaces the recurrence's, as the
// batch solve does.  Every sum runs in a fixed order.
//
// A thread keeps its rows' x, r, q and p in registers (slices wave,
// wave + 16, ... up to MAXJ of them): the only global stores of an
// iteration are p's, which the next product gathers, so the barriers of the
// two sums wait for no store.  x goes to memory for a residual check and at
// the end.
//
// coop: one problem over the whole grid, launched cooperatively (every
// workgroup resident): global wavefront g = block * 16 + wave takes slices
// g, g + 16 G, ...; a sum adds the workgroups' sums (written to part, read
// by all in workgroup order after a device-wide barrier, persist_barrier),
// and the barriers that order p's and x's stores are device-wide.  A wait
// that gives up (pa.err) ends every workgroup with ST_NAN; the host reports
// it as an error.  Otherwise one workgroup per problem.
constexpr int WG1 = 1024;
struct SmallOut {
    int64_t iters;
    double relres;
    int32_t status;   // ST_CONVERGED, ST_MAXIT or ST_NAN
    int32_t pad;
};

template <int B, int U, typename G>
__device__ inline void row_product64(const SolveArgs &a, int s, int lane, double *acc, G gather)
{
    // U blocks in flight (all of a row's 7-9 at once spills registers at
    // 1 024 threads and is slower)
    const int w = a.slice_width[s];
    const int64_t sp = a.slice_ptr[s] + lane, ns = a.nslots;
    const d2v *pl = reinterpret_cast<const d2v *>(a.v64);
    for (int k = 0; k < w; k += U) {
        uint32_t c[U];
        double m[U][B * B];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) c[u] = a.col[sp + (int64_t)(k + u) * WAVE];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) {
                const int64_t sl = sp + (int64_t)(k + u) * WAVE;
                const d2v v0 = pl[sl], v1 = pl[ns + sl];
                m[u][0] = v0.x; m[u][1] = v0.y; m[u][2] = v1.x; m[u][3] = v1.y;
                if constexpr (B == 3) {
                    const d2v v2 = pl[2 * ns + sl], v3 = pl[3 * ns + sl];
                    m[u][4] = v2.x; m[u][5] = v2.y; m[u][6] = v3.x; m[u][7] = v3.y;
                    m[u][8] = reinterpret_cast<const double *>(pl + 4 * ns)[sl];
                }
            }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < w) {
                double xv[B];
                gather(c[u], xv);
#pragma unroll
                for (int j = 0; j < B; j++)
#pragma unroll
                    for (int i = 0; i < B; i++) acc[i] += m[u][j * B + i] * xv[j];
            }
    }
}

template <int B, int MAXJ>
__global__ __launch_bounds__(WG1) void k_cg_small(SolveArgs a, const int2 *desc, double rtol, int64_t maxit,
                                                   int use_x0, SmallOut *out, int coop, PersistArgs pa, double *part)
{
    __shared__ double red[2][WG1 / WAVE];
    constexpr int NW = WG1 / WAVE;
    // blocks in flight per row: fewer the more rows a lane keeps (128
    // registers a lane at 1 024 threads; the compiler spills 2 of them at
    // B = 3, MAXJ = 2, none below, dozens beyond)
    constexpr int U = (B == 2 && MAXJ == 1) ? 4 : (B * MAXJ <= 4 ? 2 : 1);
    static_assert(B * MAXJ <= 6, "more rows a lane spill most of them");
    int rb = 0;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    // the workgroup's sum, the same bits in every thread (butterfly within
    // the wavefront, then the wavefront sums in order; two buffers, so a
    // sum's reads finish before the buffer is written again two sums later)
    auto wg_sum = [&](double v) -> double {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
        if (lane == 0) red[rb][wave] = v;
        __syncthreads();
        double t = 0.;
        for (int i = 0; i < NW; i++) t += red[rb][i];
        rb ^= 1;
        return t;
    };
    // coop: the grid's barrier and sum (part: two buffers of G sums, as the
    // LDS ones); ok turns false for every workgroup once a wait gave up
    const int G = coop ? (int)gridDim.x : 1;
    unsigned long long round = 0;
    int pb = 0;
    bool ok = true;
    auto barrier = [&]() {
        if (coop) ok = persist_barrier(pa, round) && ok;
        else __syncthreads();
    };
    auto sum = [&](double v) -> double {
        double t = wg_sum(v);
        if (!coop) return t;
        if (threadIdx.x == 0) part[pb * G + blockIdx.x] = t;
        barrier();
        t = 0.;
        for (int i = 0; i < G; i++) t += part[pb * G + i];
        pb ^= 1;
        return t;
    };
    const int2 d = desc[coop ? 0 : blockIdx.x];   // first slice, slices (at most MAXJ * NW * G)
    const int g0 = coop ? (int)blockIdx.x * NW : 0, gstride = G * NW;
    double *xg = a.x64, *pg = vw<double>(a.p);
    const double *bv = a.b64, *S = a.S;
    bool own[MAXJ];
    int sl[MAXJ];
    size_t rw[MAXJ];   // first vector entry of the row
    double x[MAXJ][B], r[MAXJ][B], p[MAXJ][B], q[MAXJ][B];
#pragma unroll
    for (int j = 0; j < MAXJ; j++) {
        sl[j] = d.x + g0 + wave + j * gstride;
        own[j] = sl[j] < d.x + d.y && lane < a.slice_rows[sl[j]];
        rw[j] = ((size_t)sl[j] * WAVE + lane) * B;
#pragma unroll
        for (int i = 0; i < B; i++) x[j][i] = r[j][i] = p[j][i] = q[j][i] = 0.;
    }
    // r = S (b - K x) from x in memory; partials |b - K x|^2 and |r|^2
    auto residual = [&](double &rr64, double &rho) {
        rr64 = rho = 0.;
#pragma unroll
        for (int j = 0; j < MAXJ; j++)
            if (own[j]) {
                double acc[B];
#pragma unroll
                for (int i = 0; i < B; i++) acc[i] = 0.;
                row_product64<B, U>(a, sl[j], lane, acc, [&](uint32_t c, double *xv) {
#pragma unroll
                    for (int i = 0; i < B; i++) xv[i] = xg[(size_t)c * B + i];
                });
#pragma unroll
                for (int i = 0; i < B; i++) {
                    const double r64 = bv[rw[j] + i] - acc[i];
                    const double rs = S[rw[j] + i] * r64;
                    rr64 += r64 * r64;
                    rho += rs * rs;
                    r[j][i] = rs;
                }
            }
    };
    double bb = 0.;
#pragma unroll
    for (int j = 0; j < MAXJ; j++)
        if (own[j])
#pragma unroll
            for (int i = 0; i < B; i++) {
                const double b = bv[rw[j] + i];
                bb += b * b;
                if (use_x0) x[j][i] = xg[rw[j] + i];
                else xg[rw[j] + i] = 0.;
            }
    bb = sum(bb);   // its barrier also orders the zeroed x before the residual's gathers
    double rr64, rho;
    residual(rr64, rho);
#pragma unroll
    for (int j = 0; j < MAXJ; j++)
        if (own[j])
#pragma unroll
            for (int i = 0; i < B; i++) {
                p[j][i] = r[j][i];
                pg[rw[j] + i] = p[j][i];
            }
    rr64 = sum(rr64);   // p in memory past this barrier
    rho = sum(rho);
    double rel = bb > 0. ? sqrt(rr64 / bb) : sqrt(rr64);
    int64_t it = 0;
    int status = rel <= rtol ? ST_CONVERGED : ST_RUNNING;
    double target = rel > 0. ? rho * (rtol / rel) * (rtol / rel) : 0.;
    while (status == ST_RUNNING) {
        // q = S K S p, partial p.q
        double pq = 0.;
#pragma unroll
        for (int j = 0; j < MAXJ; j++)
            if (own[j]) {
                double acc[B];
#pragma unroll
                for (int i = 0; i < B; i++) acc[i] = 0.;
                row_product64<B, U>(a, sl[j], lane, acc, [&](uint32_t c, double *xv) {
#pragma unroll
                    for (int i = 0; i < B; i++) xv[i] = S[(size_t)c * B + i] * pg[(size_t)c * B + i];
                });
#pragma unroll
                for (int i = 0; i < B; i++) {
                    q[j][i] = S[rw[j] + i] * acc[i];
                    pq += p[j][i] * q[j][i];
                }
            }
        pq = sum(pq);   // every gather of p is done past this barrier
        const double alpha = pq != 0. ? rho / pq : 0.;
        double rho1 = 0.;
#pragma unroll
        for (int j = 0; j < MAXJ; j++)
            if (own[j])
#pragma unroll
                for (int i = 0; i < B; i++) {
                    x[j][i] += S[rw[j] + i] * (alpha * p[j][i]);
                    r[j][i] -= alpha * q[j][i];
                    rho1 += r[j][i] * r[j][i];
                }
        rho1 = sum(rho1);
        it++;
        if (!ok || !(rho1 == rho1) || !(pq == pq)) {
            status = ST_NAN;
            break;
        }
        if (rho1 <= target || it >= maxit) {
            // the checked residual replaces the recurrence's
#pragma unroll
            for (int j = 0; j < MAXJ; j++)
                if (own[j])
#pragma unroll
                    for (int i = 0; i < B; i++) xg[rw[j] + i] = x[j][i];
            barrier();   // x in memory before the residual's gathers
            residual(rr64, rho1);
            rr64 = sum(rr64);
            rho1 = sum(rho1);
            rel = bb > 0. ? sqrt(rr64 / bb) : sqrt(rr64);
            if (!(rel == rel)) { status = ST_NAN; break; }
            if (rel <= rtol) { status = ST_CONVERGED; break; }
            if (it >= maxit) { status = ST_MAXIT; break; }
            target = rho1 * (rtol / rel) * (rtol / rel);
        }
        const double beta = rho != 0. ? rho1 / rho : 0.;
#pragma unroll
        for (int j = 0; j < MAXJ; j++)
            if (own[j])
#pragma unroll
                for (int i = 0; i < B; i++) {
                    p[j][i] = r[j][i] + beta * p[j][i];
                    pg[rw[j] + i] = p[j][i];
                }
        rho = rho1;
        barrier();   // p complete before the next product's gathers
        if (!ok) {
            status = ST_NAN;
            break;
        }
    }
#pragma unroll
    for (int j = 0; j < MAXJ; j++)
        if (own[j])
#pragma unroll
            for (int i = 0; i < B; i++) xg[rw[j] + i] = x[j][i];
    if (threadIdx.x == 0 && (!coop || blockIdx.x == 0)) {
        SmallOut &o = out[coop ? 0 : blockIdx.x];
        o.iters = it;
        o.relres = rel;
        o.status = ok ? status : ST_NAN;
    }
}

// Workgroup b takes active tiles b, b + G, ... and active problems b, b + G,
// ...; every workgroup leaves together once no problem runs (the statuses
// read after the period's last barrier are the same for all) or after
// max_periods.  Launched cooperatively (all workgroups resident).
template <typename W, typename M, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_persistent(SolveArgs a, PersistArgs pa)
{
    unsigned long long round = 0;
    const int G = gridDim.x, b = blockIdx.x;
    const int nt = *a.n_active, np = *a.n_active_prob;
    int period = 0;
    while (period < pa.max_periods) {
        for (int i = 0; i < pa.ce; i++) {
            for (int it = b; it < nt; it += G) k1_body<W, M, B, NR, BJ>(a, a.active[it]);
            if (!persist_barrier(pa, round)) return;
            for (int pi = b; pi < np; pi += G) reduce_body<NR>(a, pi);
            if (!persist_barrier(pa, round)) return;
            for (int it = b; it < nt; it += G) k2_body<W, B, NR, BJ>(a, a.active[it], i == pa.ce - 1 ? 1 : 0);
            if (!persist_barrier(pa, round)) return;
        }
        for (int it = b; it < nt; it += G) cx_body<W, B, NR>(a, a.active[it]);
        if (!persist_barrier(pa, round)) return;
        for (int it = b; it < nt; it += G) cr_body<W, B, NR, BJ>(a, a.active[it]);
        if (!persist_barrier(pa, round)) return;
        period++;
        bool any = false;
        for (int pi = 0; pi < np; pi++) any |= a.prob[a.active_prob[pi]].status == ST_RUNNING;
        if (!any) break;
    }
    if (b == 0 && threadIdx.x == 0) *pa.periods = period;
}

// ------------------------------------------ norms of the right-hand sides
template <int B, int NR>
__global__ __launch_bounds__(WG) void k_rhs_norms(SolveArgs a)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    const int t = my_tile(a);
    if (t < 0) return;
    const int wave = wave_id(), lane = threadIdx.x % WAVE;
    double part[2 * NR];
#pragma unroll
    for (int k = 0; k < 2 * NR; k++) part[k] = 0.;
    if (wave < a.tile_nslices[t]) {
        const int s = a.tile_slice0[t] + wave;
        const int64_t row = (int64_t)s * WAVE + lane;
        if (lane < a.slice_rows[s]) {
#pragma unroll
            for (int k = 0; k < NR; k++)
#pragma unroll
                for (int i = 0; i < B; i++) {
                    const double bv = a.b64[k * a.vplane + (size_t)row * B + i];
                    const double sb = a.S[(size_t)row * B + i] * bv;
                    part[k] += sb * sb;
                    part[NR + k] += bv * bv;
                }
        }
    }
    write_partials<2 * NR>(part, a.partn, t, lds);
}

// One workgroup per active problem: |S b|^2 and |b|^2 from the norm partials.
template <int NR>
__global__ __launch_bounds__(WG) void k_norms_final(SolveArgs a)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    if ((int)blockIdx.x >= *a.n_active_prob) return;
    ProbDev &P = a.prob[a.active_prob[blockIdx.x]];
    if (P.status != ST_RUNNING) return;
    double sums[2 * NR];
    problem_sums<2 * NR>(a.partn, P, sums, lds);
    if (threadIdx.x != 0) return;
    for (int k = 0; k < NR; k++) {
        P.bb[k] = sums[k];
        P.bb64[k] = sums[NR + k];
        P.rr_target[k] = P.ref_eps > 0. ? fmax(1e-24, P.ref_eps * P.ref_eps) : P.rtol * P.rtol * sums[k];
        // a homogeneous right-hand side is solved by x = 0
        // (solvers/conjugategradient.cpp:74-78)
        if (sums[NR + k] == 0.) {
            P.done[k] = P.done_out[k] = 1;
            P.relres[k] = P.relres_out[k] = 0.;
        }
    }
}

// ------------------------------------------------- plain SpMV (test hook)
// y = K x with the FP64 planes (mode 0), or S^-1 (Ks (S^-1 x)) with the
// scaled FP32 planes (mode 1, FP32 accumulation) or the split planes (mode 2,
// FP64 accumulation), a.vw pointing at them.  x, y: FP64 planes.
template <typename A, typename M, int B, int NR>
__device__ inline void spmv_plain_row(const SolveArgs &a, const void *vals, bool scaled, const double *x, double *yout,
                                      int s, int lane)
{
    const int64_t row = (int64_t)s * WAVE + lane;
    A acc[NR][B];
#pragma unroll
    for (int k = 0; k < NR; k++)
#pragma unroll
        for (int i = 0; i < B; i++) acc[k][i] = 0;
    auto gather = [&](uint32_t c, A (*xv)[B]) {
#pragma unroll
        for (int k = 0; k < NR; k++)
#pragma unroll
            for (int j = 0; j < B; j++) {
                const double v = x[k * a.vplane + (size_t)c * B + j];
                xv[k][j] = (A)(scaled ? v / a.S[(size_t)c * B + j] : v);
            }
    };
    spmv_row<A, M, B, NR>(a, vals, s, lane, acc, gather);
    if (lane < a.slice_rows[s])
        for (int k = 0; k < NR; k++)
            for (int i = 0; i < B; i++)
                yout[k * a.vplane + (size_t)row * B + i] = scaled ? (double)acc[k][i] / a.S[(size_t)row * B + i] : (double)acc[k][i];
}

template <int B, int NR>
__global__ __launch_bounds__(WG) void k_spmv_plain(SolveArgs a, int mode, const double *x, double *yout, int ntiles)
{
    const int wave = wave_id(), lane = threadIdx.x % WAVE;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (wave >= a.tile_nslices[t]) continue;
        const int s = a.tile_slice0[t] + wave;
        if (mode == 0) spmv_plain_row<double, double, B, NR>(a, a.v64, false, x, yout, s, lane);
        else if (mode == 1) spmv_plain_row<float, float, B, NR>(a, a.vw, true, x, yout, s, lane);
        else if (mode == 2) spmv_plain_row<double, Split, B, NR>(a, a.vw, true, x, yout, s, lane);
        else spmv_plain_row<double, Fix48, B, NR>(a, a.vw, true, x, yout, s, lane);
    }
}

} // namespace amie
