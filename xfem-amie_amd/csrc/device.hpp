// Device-side data structures and kernels of the AMIE FE-solve hot path on
// MI355X (gfx950, wave64).
//
// Matrix storage ("SELL-64 block"): the node-block CSR of the reference
// (sparse/sparse_matrix.h:129-137) is re-laid out in slices of 64 block rows,
// one slice per wavefront.  Within a slice the k-th block of every row is
// stored contiguously over the 64 lanes, so each load instruction of a
// wavefront reads one contiguous run (a 2x2 FP32 block is one 16-byte
// float4 per lane: 1 KiB per wavefront).  Rows shorter than the slice width
// are padded with zero blocks whose column is the row itself.  Block entries
// keep the reference's column-major order (entry (r,c) at c*b + r), without
// the reference's 4x3 padding of 3x3 blocks.  Blocks are split into 16-byte
// "planes" so every plane load is a dwordx4 (or dword/dwordx2 remainder):
//   FP32 2x2: 1 x float4            FP32 3x3: 2 x float4 + 1 float
//   FP64 2x2: 2 x double2           FP64 3x3: 4 x double2 + 1 double
//
// Vectors are stored per internal block row as [row][rhs][axis].
//
// Work is distributed in tiles of up to 4 slices (256 rows, one 256-thread
// workgroup) that never straddle two problems of a batch.
//
// Reductions never cross workgroups inside a kernel: on CDNA the L2 of each
// XCD is not coherent with the others, so making one workgroup's writes
// visible to another inside a kernel needs a device-scope release (a write
// back of the whole L2).  Instead every kernel writes per-tile partial sums
// and the workgroups of the NEXT kernel sum their problem's partials (in tile
// order, so every workgroup gets the same bits) and derive the CG scalars
// from them.  Per-problem state that later kernels need is written by one
// designated workgroup (the problem's first tile) into fields that no other
// workgroup of the same kernel reads; kernel boundaries publish it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amie {

constexpr int WAVE = 64;
constexpr int TILE_SLICES = 4;
constexpr int WG = WAVE * TILE_SLICES;
constexpr int MAXRHS = 4;
constexpr int MAXQ = 3 * MAXRHS;

enum Status : int32_t { ST_IDLE = 0, ST_RUNNING = 1, ST_CONVERGED = 2, ST_MAXIT = 3, ST_NAN = 4, ST_STALLED = 5 };

// Per-problem constants (host) and CG state (device), FP64, resident in HBM.
struct ProbDev {
    int32_t tile0, ntiles;
    int32_t slice0, nslices;
    int64_t row0;
    int32_t nrows;
    int32_t refine;          // 0 replace residual, keep p; 1 restart at each refinement
    double rtol, drop;
    int64_t maxit;
    int64_t ckpt_interval;   // FP64 refinement at least every this many iterations
    int32_t stall_limit;     // > 0: ST_STALLED after that many refinements without halving
    // State written by the designated workgroup and read by later kernels.
    // Fields the p-update kernel reads AND its designated workgroup writes
    // are kept per iteration parity: the kernel of parity par reads slot
    // par ^ 1 and writes slot par, so no workgroup sees a value of its own
    // kernel; the refinement kernels write slot 1, which the next p update
    // (parity 0) reads.
    int32_t status;          // a finished status only ever stops work
    int32_t ckpt_pending;    // a refinement is due at the next checkpoint slot
    int32_t src[2];          // 1: the latest rho/rr partials are a refinement's
    int32_t first[2];        // 1: the next p update starts a new recurrence (beta = 0)
    int32_t done[2][MAXRHS];
    double relres[2][MAXRHS];
    double rho_hist[2][MAXRHS];   // rho of the p update
    int64_t iters, nckpt, last_ckpt;
    double rr[MAXRHS], rr_ckpt[MAXRHS], bb[MAXRHS], bb64[MAXRHS], rel_best[MAXRHS];
    int32_t stall[MAXRHS];
    // results for the host (never read by kernels)
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
    const void *vw;          // planes used by the p update's SpMV (scaled FP32 S K S, or v64)
    const int32_t *tile_slice0, *tile_nslices, *tile_prob;
    const int32_t *active;
    const int32_t *n_active;
    ProbDev *prob;
    double *part1;           // [tile][MAXQ] p.q
    double *part2;           // [tile][MAXQ] r.z, r.r after the update
    double *partc;           // [tile][MAXQ] r.z, r.r, |b - K x|^2 after a refinement
    double *partn;           // [tile][MAXQ] |S b|^2, |b|^2
    double *x64;
    const double *b64;
    const double *S;         // symmetric Jacobi scaling, per dof
    void *y, *r, *z, *p0, *p1, *q;
    const void *minv;        // block-Jacobi inverses of the scaled diagonal blocks, per row (b*b)
};

// ------------------------------------------------------------ block planes
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

template <typename W, int B> struct Blk;

template <> struct Blk<float, 2> {
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

// Status of a problem as ONE value for the whole workgroup.  The p-update
// kernel's designated workgroup may finish a problem while other workgroups
// of that kernel start tiles of it; without the broadcast, waves of one
// workgroup could read different values and disagree about entering the
// workgroup barriers below.
__device__ inline int32_t wg_status(const ProbDev &P, int32_t *sh)
{
    __syncthreads();
    if (threadIdx.x == 0) *sh = __hip_atomic_load(&P.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return *sh;
}

// Sums the partials a previous kernel wrote for problem P, in a fixed order:
// every workgroup of the problem obtains the same totals.
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

// ------------------------------------------------------- tile scheduling
// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8); each
// XCD gets a contiguous range of the active tiles, so rows whose vector
// entries neighbouring rows gather are processed under the same L2.
constexpr int NXCD = 8;
template <typename F> __device__ inline void for_tiles(int na, F body)
{
    if (gridDim.x >= NXCD) {
        const int xcd = blockIdx.x % NXCD, lw = blockIdx.x / NXCD;
        const int nw = (gridDim.x - xcd + NXCD - 1) / NXCD;
        const int per = (na + NXCD - 1) / NXCD;
        for (int j = lw; j < per; j += nw) {
            const int it = xcd * per + j;
            if (it >= na) break;
            body(it);
        }
    } else {
        for (int it = blockIdx.x; it < na; it += gridDim.x) body(it);
    }
}

// ------------------------------------------------------------ SpMV core
// Accumulates acc[k][i] += sum_blocks A(i,j) * src(c, k, j) for one lane
// (row) of slice s, reading the block planes `vals` of precision W.
// src(c, k, j) is supplied by a functor (gathers of neighbouring rows).
template <typename W, typename M, int B, int NR, typename Src>
__device__ inline void sell_row(const SolveArgs &a, const void *vals, int s, int lane, W (*acc)[B], Src src)
{
    // U blocks per step: their column indices and values are all requested
    // before the dependent gathers, so each wavefront keeps 2U independent
    // loads in flight instead of one load -> gather chain per block.
    constexpr int U = 4;
    const int w = a.slice_width[s];
    const int64_t sp = a.slice_ptr[s] + lane;
    int k = 0;
    for (; k + U <= w; k += U) {
        uint32_t c[U];
        M blk[U][B * B];
#pragma unroll
        for (int u = 0; u < U; u++) c[u] = __builtin_nontemporal_load(a.col + sp + (int64_t)(k + u) * WAVE);
#pragma unroll
        for (int u = 0; u < U; u++) Blk<M, B>::load(vals, a.nslots, sp + (int64_t)(k + u) * WAVE, blk[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int r = 0; r < NR; r++) {
                W xv[B];
                src(c[u], r, xv);
#pragma unroll
                for (int j = 0; j < B; j++)
#pragma unroll
                    for (int i = 0; i < B; i++) acc[r][i] += (W)blk[u][j * B + i] * xv[j];
            }
    }
    for (; k < w; k++) {
        const int64_t slot = sp + (int64_t)k * WAVE;
        const uint32_t c = __builtin_nontemporal_load(a.col + slot);
        M blk[B * B];
        Blk<M, B>::load(vals, a.nslots, slot, blk);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            W xv[B];
            src(c, r, xv);
#pragma unroll
            for (int j = 0; j < B; j++)
#pragma unroll
                for (int i = 0; i < B; i++) acc[r][i] += (W)blk[j * B + i] * xv[j];
        }
    }
}


template <typename W> __device__ inline W *vw(void *p) { return reinterpret_cast<W *>(p); }
template <typename W> __device__ inline const W *cvw(const void *p) { return reinterpret_cast<const W *>(p); }

// The B components of one right-hand side of node c, as one vector load for B = 2.
template <typename W, int B> __device__ inline void load_node(const W *v, size_t idx, W *o)
{
    if constexpr (B == 2) {
        typedef W v2 __attribute__((ext_vector_type(2)));
        const v2 t = *reinterpret_cast<const v2 *>(v + idx);
        o[0] = t.x;
        o[1] = t.y;
    } else {
#pragma unroll
        for (int j = 0; j < B; j++) o[j] = v[idx + j];
    }
}

// ---------------------------------------------------- K1: p update + SpMV
// beta from the partials of the previous update (or refinement), then
// p_new = z + beta p_old (own rows stored, neighbours recomputed on the fly),
// q = A p_new and partial p_new . q.  With the FP64 matrix the unscaled K is
// used with the scaling applied on the fly: q = S K (S p).
// The designated workgroup records rho, consumes a refinement's residual
// (convergence, stagnation) and decides when the next refinement is due.
template <typename W, typename M, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_spmv_p(SolveArgs a, int parity)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    __shared__ int32_t sh_status;
    constexpr bool ONFLY = sizeof(M) == 8;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const W *zs = BJ ? cvw<W>(a.z) : cvw<W>(a.r);
    const W *pold = cvw<W>(parity ? a.p1 : a.p0);
    W *pnew = vw<W>(parity ? a.p0 : a.p1);
    W *q = vw<W>(a.q);
    const int na = *a.n_active;
    for_tiles(na, [&](int it) {
        const int t = a.active[it];
        const int p = a.tile_prob[t];
        const ProbDev &P = a.prob[p];
        if (wg_status(P, &sh_status) != ST_RUNNING) return;
        const int rd = parity ^ 1;   // slot written before this kernel
        const bool ck = P.src[rd] != 0;
        double sums[3 * NR];
        if (ck) problem_sums<3 * NR>(a.partc, P, sums, lds);
        else problem_sums<2 * NR>(a.part2, P, sums, lds);
        bool done[NR], all = true, bad = false;
        double rel[NR];
        W beta[NR];
#pragma unroll
        for (int k = 0; k < NR; k++) {
            bool dk = P.done[rd][k] != 0;
            rel[k] = P.relres[rd][k];
            if (ck && !dk) {
                const double rr64 = sums[2 * NR + k];
                if (!(rr64 == rr64)) bad = true;
                rel[k] = P.bb64[k] > 0. ? sqrt(rr64 / P.bb64[k]) : sqrt(rr64);
                dk = rel[k] <= P.rtol;
            }
            if (!(sums[k] == sums[k])) bad = true;
            done[k] = dk;
            all = all && dk;
            const double rho_old = P.rho_hist[rd][k];
            const bool restart = P.first[rd] || (ck && P.refine == 1) || rho_old == 0.;
            beta[k] = (dk || restart) ? (W)0 : (W)(sums[k] / rho_old);
        }
        const bool designated = t == P.tile0 && threadIdx.x == 0;
        if (all || bad) {
            if (designated) {
                ProbDev &Pw = a.prob[p];
                for (int k = 0; k < NR; k++) {
                    Pw.done[parity][k] = Pw.done_out[k] = 1;
                    Pw.relres[parity][k] = Pw.relres_out[k] = rel[k];
                }
                Pw.status = bad ? ST_NAN : ST_CONVERGED;
            }
            return;
        }
        double part[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) part[r] = 0.;
        if (wave < a.tile_nslices[t]) {
            const int s = a.tile_slice0[t] + wave;
            const int64_t row = (int64_t)s * WAVE + lane;
            W acc[NR][B];
#pragma unroll
            for (int r = 0; r < NR; r++)
#pragma unroll
                for (int i = 0; i < B; i++) acc[r][i] = 0;
            auto src = [&](uint32_t c, int r, W *xv) {
                W zv[B], pv[B];
                const size_t idx = ((size_t)c * NR + r) * B;
                load_node<W, B>(zs, idx, zv);
                load_node<W, B>(pold, idx, pv);
#pragma unroll
                for (int j = 0; j < B; j++) {
                    W v = zv[j] + beta[r] * pv[j];
                    if (ONFLY) v *= (W)a.S[(size_t)c * B + j];
                    xv[j] = v;
                }
            };
            sell_row<W, M, B, NR>(a, a.vw, s, lane, acc, src);
            if (lane < a.slice_rows[s]) {
#pragma unroll
                for (int r = 0; r < NR; r++)
#pragma unroll
                    for (int i = 0; i < B; i++) {
                        const size_t idx = ((size_t)row * NR + r) * B + i;
                        W pn = zs[idx] + beta[r] * pold[idx];
                        W qv = acc[r][i];
                        if (ONFLY) qv *= (W)a.S[(size_t)row * B + i];
                        pnew[idx] = pn;
                        q[idx] = qv;
                        part[r] += (double)pn * (double)qv;
                    }
            }
        }
        write_partials<NR>(part, a.part1, t, lds);
        if (designated) {
            ProbDev &Pw = a.prob[p];
            bool stalled = false;
            for (int k = 0; k < NR; k++) Pw.rho_hist[parity][k] = sums[k];
            for (int k = 0; k < NR; k++) {
                Pw.relres[parity][k] = Pw.relres_out[k] = rel[k];
                Pw.done[parity][k] = Pw.done_out[k] = done[k];
            }
            if (ck) {
                // a refinement's FP64 residual: record, watch for stagnation
                for (int k = 0; k < NR; k++) {
                    Pw.rr_ckpt[k] = sums[NR + k];
                    if (done[k]) continue;
                    if (rel[k] < 0.5 * Pw.rel_best[k]) {
                        Pw.rel_best[k] = rel[k];
                        Pw.stall[k] = 0;
                    } else if (++Pw.stall[k] >= Pw.stall_limit && Pw.stall_limit > 0) {
                        stalled = true;
                    }
                }
                Pw.nckpt += 1;
                Pw.last_ckpt = Pw.iters;
                Pw.ckpt_pending = 0;
            }
            Pw.src[parity] = 0;
            Pw.first[parity] = 0;
            if (ck && Pw.iters >= Pw.maxit) { Pw.status = ST_MAXIT; return; }
            if (stalled) { Pw.status = ST_STALLED; return; }
            Pw.iters += 1;
            bool need = Pw.iters >= Pw.maxit || Pw.iters - Pw.last_ckpt >= Pw.ckpt_interval;
            for (int k = 0; k < NR; k++) {
                if (done[k]) continue;
                const double rr = sums[NR + k];
                Pw.rr[k] = rr;
                if (rr < Pw.drop * Pw.drop * Pw.rr_ckpt[k]) need = true;
                if (rr <= Pw.rtol * Pw.rtol * Pw.bb[k]) need = true;
            }
            if (need) Pw.ckpt_pending = 1;
        }
    });
}

// ------------------------------------ K2: y, r update + preconditioner + dots
template <typename W, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_update(SolveArgs a, int parity)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    __shared__ int32_t sh_status;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const W *pnew = cvw<W>(parity ? a.p0 : a.p1);
    const W *q = cvw<W>(a.q);
    W *y = vw<W>(a.y), *r = vw<W>(a.r), *z = vw<W>(a.z);
    const W *minv = cvw<W>(a.minv);
    const int na = *a.n_active;
    for_tiles(na, [&](int it) {
        const int t = a.active[it];
        const int p = a.tile_prob[t];
        const ProbDev &P = a.prob[p];
        if (wg_status(P, &sh_status) != ST_RUNNING) return;
        double pq[NR];
        problem_sums<NR>(a.part1, P, pq, lds);
        W alpha[NR];
#pragma unroll
        for (int k = 0; k < NR; k++)
            alpha[k] = (P.done[parity][k] || pq[k] == 0.) ? (W)0 : (W)(P.rho_hist[parity][k] / pq[k]);
        double part[2 * NR];
#pragma unroll
        for (int k = 0; k < 2 * NR; k++) part[k] = 0.;
        if (wave < a.tile_nslices[t]) {
            const int s = a.tile_slice0[t] + wave;
            const int64_t row = (int64_t)s * WAVE + lane;
            if (lane < a.slice_rows[s]) {
                W mi[B * B];
                if (BJ) {
#pragma unroll
                    for (int k = 0; k < B * B; k++) mi[k] = minv[(size_t)row * B * B + k];
                }
#pragma unroll
                for (int k = 0; k < NR; k++) {
                    W rv[B];
#pragma unroll
                    for (int i = 0; i < B; i++) {
                        const size_t idx = ((size_t)row * NR + k) * B + i;
                        y[idx] = y[idx] + alpha[k] * pnew[idx];
                        rv[i] = r[idx] - alpha[k] * q[idx];
                        r[idx] = rv[i];
                    }
#pragma unroll
                    for (int i = 0; i < B; i++) {
                        W zv = rv[i];
                        if (BJ) {
                            zv = 0;
#pragma unroll
                            for (int j = 0; j < B; j++) zv += mi[i * B + j] * rv[j];
                            z[((size_t)row * NR + k) * B + i] = zv;
                        }
                        part[k] += (double)rv[i] * (double)zv;
                        part[NR + k] += (double)rv[i] * (double)rv[i];
                    }
                }
            }
        }
        write_partials<2 * NR>(part, a.part2, t, lds);
    });
}

// ------------------------------------------- refinement 1: x64 += S y, y = 0
template <typename W, int B, int NR>
__global__ __launch_bounds__(WG) void k_ckpt_x(SolveArgs a)
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    W *y = vw<W>(a.y);
    const int na = *a.n_active;
    for_tiles(na, [&](int it) {
        const int t = a.active[it];
        const ProbDev &P = a.prob[a.tile_prob[t]];
        if (P.status != ST_RUNNING || !P.ckpt_pending) return;
        if (t == P.tile0 && threadIdx.x == 0) a.prob[a.tile_prob[t]].src[1] = 1;
        if (wave >= a.tile_nslices[t]) return;
        const int s = a.tile_slice0[t] + wave;
        if (lane >= a.slice_rows[s]) return;
        const int64_t row = (int64_t)s * WAVE + lane;
#pragma unroll
        for (int k = 0; k < NR; k++)
#pragma unroll
            for (int i = 0; i < B; i++) {
                const size_t idx = ((size_t)row * NR + k) * B + i;
                a.x64[idx] += a.S[(size_t)row * B + i] * (double)y[idx];
                y[idx] = 0;
            }
    });
}

// -------------------- refinement 2: r64 = b - K x64 (FP64), r = S r64 (W), z
template <typename W, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_ckpt_r(SolveArgs a)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    __shared__ int32_t sh_status;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    W *r = vw<W>(a.r), *z = vw<W>(a.z);
    const W *minv = cvw<W>(a.minv);
    const int na = *a.n_active;
    for_tiles(na, [&](int it) {
        const int t = a.active[it];
        const int p = a.tile_prob[t];
        const ProbDev &P = a.prob[p];
        if (P.status != ST_RUNNING || !P.ckpt_pending) return;
        double part[3 * NR];
#pragma unroll
        for (int k = 0; k < 3 * NR; k++) part[k] = 0.;
        if (wave < a.tile_nslices[t]) {
            const int s = a.tile_slice0[t] + wave;
            const int64_t row = (int64_t)s * WAVE + lane;
            double acc[NR][B];
#pragma unroll
            for (int k = 0; k < NR; k++)
#pragma unroll
                for (int i = 0; i < B; i++) acc[k][i] = 0.;
            auto src = [&](uint32_t c, int k, double *xv) { load_node<double, B>(a.x64, ((size_t)c * NR + k) * B, xv); };
            sell_row<double, double, B, NR>(a, a.v64, s, lane, acc, src);
            if (lane < a.slice_rows[s]) {
                W mi[B * B];
                if (BJ) {
#pragma unroll
                    for (int k = 0; k < B * B; k++) mi[k] = minv[(size_t)row * B * B + k];
                }
#pragma unroll
                for (int k = 0; k < NR; k++) {
                    W rv[B];
#pragma unroll
                    for (int i = 0; i < B; i++) {
                        const size_t idx = ((size_t)row * NR + k) * B + i;
                        double r64 = a.b64[idx] - acc[k][i];
                        part[2 * NR + k] += r64 * r64;
                        rv[i] = (W)(a.S[(size_t)row * B + i] * r64);
                        r[idx] = rv[i];
                    }
#pragma unroll
                    for (int i = 0; i < B; i++) {
                        W zv = rv[i];
                        if (BJ) {
                            zv = 0;
#pragma unroll
                            for (int j = 0; j < B; j++) zv += mi[i * B + j] * rv[j];
                            z[((size_t)row * NR + k) * B + i] = zv;
                        }
                        part[k] += (double)rv[i] * (double)zv;
                        part[NR + k] += (double)rv[i] * (double)rv[i];
                    }
                }
            }
        }
        write_partials<3 * NR>(part, a.partc, t, lds);
    });
}

// ------------------------------------------ norms of the right-hand sides
template <int B, int NR>
__global__ __launch_bounds__(WG) void k_rhs_norms(SolveArgs a)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int na = *a.n_active;
    for_tiles(na, [&](int it) {
        const int t = a.active[it];
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
                        double bv = a.b64[((size_t)row * NR + k) * B + i];
                        double sb = a.S[(size_t)row * B + i] * bv;
                        part[k] += sb * sb;
                        part[NR + k] += bv * bv;
                    }
            }
        }
        write_partials<2 * NR>(part, a.partn, t, lds);
    });
}

// One workgroup per problem: |S b|^2 and |b|^2 from the norm partials.
template <int NR>
__global__ __launch_bounds__(WG) void k_norms_final(SolveArgs a, int nprob)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    for (int p = blockIdx.x; p < nprob; p += gridDim.x) {
        const ProbDev &P = a.prob[p];
        if (P.status != ST_RUNNING) continue;
        double sums[2 * NR];
        problem_sums<2 * NR>(a.partn, P, sums, lds);
        if (threadIdx.x == 0) {
            ProbDev &Pw = a.prob[p];
            for (int k = 0; k < NR; k++) {
                Pw.bb[k] = sums[k];
                Pw.bb64[k] = sums[NR + k];
                // a homogeneous right-hand side is solved by x = 0
                // (solvers/conjugategradient.cpp:74-78)
                if (sums[NR + k] == 0.) {
                    Pw.done[0][k] = Pw.done[1][k] = Pw.done_out[k] = 1;
                    Pw.relres[0][k] = Pw.relres[1][k] = Pw.relres_out[k] = 0.;
                }
            }
        }
    }
}

// ------------------------------------------------- plain SpMV (test hook)
// y = K x with the FP64 planes (mode 0) or S^-1 (Ks (S^-1 x)) with the
// scaled working planes Ks = S K S (mode 1).  One lane per row, every tile.
template <typename W, int B, int NR>
__global__ __launch_bounds__(WG) void k_spmv_plain(SolveArgs a, int mode, const double *x, double *yout, int ntiles)
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (wave >= a.tile_nslices[t]) continue;
        const int s = a.tile_slice0[t] + wave;
        const int64_t row = (int64_t)s * WAVE + lane;
        if (mode == 0) {
            double acc[NR][B];
#pragma unroll
            for (int k = 0; k < NR; k++)
#pragma unroll
                for (int i = 0; i < B; i++) acc[k][i] = 0.;
            auto src = [&](uint32_t c, int k, double *xv) {
#pragma unroll
                for (int j = 0; j < B; j++) xv[j] = x[((size_t)c * NR + k) * B + j];
            };
            sell_row<double, double, B, NR>(a, a.v64, s, lane, acc, src);
            if (lane < a.slice_rows[s])
                for (int k = 0; k < NR; k++)
                    for (int i = 0; i < B; i++) yout[((size_t)row * NR + k) * B + i] = acc[k][i];
        } else {
            W acc[NR][B];
#pragma unroll
            for (int k = 0; k < NR; k++)
#pragma unroll
                for (int i = 0; i < B; i++) acc[k][i] = 0;
            auto src = [&](uint32_t c, int k, W *xv) {
#pragma unroll
                for (int j = 0; j < B; j++) xv[j] = (W)(x[((size_t)c * NR + k) * B + j] / a.S[(size_t)c * B + j]);
            };
            sell_row<W, W, B, NR>(a, a.vw, s, lane, acc, src);
            if (lane < a.slice_rows[s])
                for (int k = 0; k < NR; k++)
                    for (int i = 0; i < B; i++)
                        yout[((size_t)row * NR + k) * B + i] = (double)acc[k][i] / a.S[(size_t)row * B + i];
        }
    }
}

} // namespace amie
