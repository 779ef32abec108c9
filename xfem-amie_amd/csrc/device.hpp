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
// workgroup) that never straddle two problems of a batch; kernels loop over
// the active-tile list with a grid stride.  Every reduction is deterministic:
// each tile writes its partial sums, and the last tile of a problem to finish
// (ticket counter) sums that problem's partials in tile order and updates
// the problem's scalars.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amie {

constexpr int WAVE = 64;
constexpr int TILE_SLICES = 4;
constexpr int WG = WAVE * TILE_SLICES;
constexpr int MAXRHS = 4;
constexpr int MAXQ = 3 * MAXRHS;

enum Status : int32_t { ST_IDLE = 0, ST_RUNNING = 1, ST_CONVERGED = 2, ST_MAXIT = 3, ST_NAN = 4 };

// Per-problem constants and Krylov scalars (FP64), resident in HBM.
struct ProbDev {
    int32_t tile0, ntiles;
    int32_t slice0, nslices;
    int64_t row0;
    int32_t nrows;
    int32_t refine;      // 0 replace residual, keep p; 1 restart
    double rtol, drop;
    int64_t maxit;
    int32_t status, ckpt_pending;
    int64_t iters, nckpt;
    double rho[MAXRHS], rho_prev[MAXRHS], alpha[MAXRHS], beta[MAXRHS];
    double rr[MAXRHS], rr_ckpt[MAXRHS], bb[MAXRHS], bb64[MAXRHS], relres[MAXRHS], rel_prev[MAXRHS];
    int32_t done[MAXRHS];
};

// Arguments shared by the solve kernels.
struct SolveArgs {
    const int64_t *slice_ptr;
    const int32_t *slice_width;
    const int32_t *slice_rows;
    const uint32_t *col;
    int64_t nslots;
    const double *v64;       // FP64 planes (unscaled K)
    const void *vw;          // working planes (scaled FP32 K in mixed mode; == v64 in FP64 mode)
    const int32_t *tile_slice0, *tile_nslices, *tile_prob;
    const int32_t *active;
    const int32_t *n_active;
    ProbDev *prob;
    unsigned *ticket;
    double *partial;         // [tile][MAXQ]
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

// Sums v[0..NQ) over the workgroup; result valid in thread 0.
template <int NQ> __device__ inline void wg_sum(double *v, double (*lds)[MAXQ])
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double s = wave_sum(v[q]);
        if (lane == 0) lds[wave][q] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) v[q] = lds[0][q] + lds[1][q] + lds[2][q] + lds[3][q];
    }
    __syncthreads();
}

__device__ inline double load_coherent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Writes this tile's partials; returns true (in every thread) when this
// workgroup holds the last tile of problem `p`, in which case tot[] (thread 0)
// holds the problem totals summed in tile order.
template <int NQ>
__device__ inline bool tile_reduce(double *v, int tile, int p, const ProbDev &P, unsigned *ticket, double *partial,
                                   double (*lds)[MAXQ], int *flag)
{
    wg_sum<NQ>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++)
            __hip_atomic_store(partial + (size_t)tile * MAXQ + q, v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        unsigned prev = atomicAdd(ticket + p, 1u);
        *flag = (prev == (unsigned)(P.ntiles - 1));
    }
    __syncthreads();
    bool last = *flag;
    if (last) {
        __threadfence();
        double acc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[q] = 0.;
        for (int t = threadIdx.x; t < P.ntiles; t += WG) {
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[q] += load_coherent(partial + (size_t)(P.tile0 + t) * MAXQ + q);
        }
        wg_sum<NQ>(acc, lds);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < NQ; q++) v[q] = acc[q];
            ticket[p] = 0;
        }
    }
    __syncthreads();
    return last;
}

// ------------------------------------------------------------ SpMV core
// Accumulates acc[k][i] += sum_blocks A(i,j) * src(c, k, j) for one lane
// (row) of slice s, reading the block planes `vals` of precision W.
// src(c, k, j) is supplied by a functor (gathers of neighbouring rows).
template <typename W, int B, int NR, typename Src>
__device__ inline void sell_row(const SolveArgs &a, const void *vals, int s, int lane, W (*acc)[B], Src src)
{
    const int w = a.slice_width[s];
    const int64_t sp = a.slice_ptr[s] + lane;
    for (int k = 0; k < w; k++) {
        const int64_t slot = sp + (int64_t)k * WAVE;
        const uint32_t c = __builtin_nontemporal_load(a.col + slot);
        W blk[B * B];
        Blk<W, B>::load(vals, a.nslots, slot, blk);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            W xv[B];
            src(c, r, xv);
#pragma unroll
            for (int j = 0; j < B; j++)
#pragma unroll
                for (int i = 0; i < B; i++) acc[r][i] += blk[j * B + i] * xv[j];
        }
    }
}

template <typename W> __device__ inline W *vw(void *p) { return reinterpret_cast<W *>(p); }
template <typename W> __device__ inline const W *cvw(const void *p) { return reinterpret_cast<const W *>(p); }

// ---------------------------------------------------- K1: p update + SpMV
// p_new = z + beta p_old (own rows stored; neighbours recomputed on the fly),
// q = A p_new, and pq = p_new . q per right-hand side.  In FP64 mode the
// unscaled K is used with the scaling applied on the fly: q = S K (S p).
template <typename W, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_spmv_p(SolveArgs a, int parity)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    __shared__ int flag;
    constexpr bool ONFLY = sizeof(W) == 8;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const W *zs = BJ ? cvw<W>(a.z) : cvw<W>(a.r);
    const W *pold = cvw<W>(parity ? a.p1 : a.p0);
    W *pnew = vw<W>(parity ? a.p0 : a.p1);
    W *q = vw<W>(a.q);
    const int na = *a.n_active;
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const int t = a.active[it];
        const int p = a.tile_prob[t];
        const ProbDev &P = a.prob[p];
        if (P.status != ST_RUNNING) continue;
        W beta[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) beta[r] = (W)P.beta[r];
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
#pragma unroll
                for (int j = 0; j < B; j++) {
                    const size_t idx = ((size_t)c * NR + r) * B + j;
                    W v = zs[idx] + beta[r] * pold[idx];
                    if (ONFLY) v *= (W)a.S[(size_t)c * B + j];
                    xv[j] = v;
                }
            };
            sell_row<W, B, NR>(a, a.vw, s, lane, acc, src);
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
        if (tile_reduce<NR>(part, t, p, P, a.ticket, a.partial, lds, &flag) && threadIdx.x == 0) {
            ProbDev &Pw = a.prob[p];
            for (int r = 0; r < NR; r++) {
                double pq = part[r];
                Pw.alpha[r] = (Pw.done[r] || pq == 0.) ? 0. : Pw.rho[r] / pq;
            }
        }
    }
}

// ------------------------------------ K2: y, r update + preconditioner + dots
template <typename W, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_update(SolveArgs a, int parity)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    __shared__ int flag;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const W *pnew = cvw<W>(parity ? a.p0 : a.p1);
    const W *q = cvw<W>(a.q);
    W *y = vw<W>(a.y), *r = vw<W>(a.r), *z = vw<W>(a.z);
    const W *minv = cvw<W>(a.minv);
    const int na = *a.n_active;
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const int t = a.active[it];
        const int p = a.tile_prob[t];
        const ProbDev &P = a.prob[p];
        if (P.status != ST_RUNNING) continue;
        W alpha[NR];
#pragma unroll
        for (int k = 0; k < NR; k++) alpha[k] = (W)P.alpha[k];
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
        if (tile_reduce<2 * NR>(part, t, p, P, a.ticket, a.partial, lds, &flag) && threadIdx.x == 0) {
            ProbDev &Pw = a.prob[p];
            Pw.iters += 1;
            bool need = false, bad = false;
            for (int k = 0; k < NR; k++) {
                if (Pw.done[k]) { Pw.alpha[k] = 0.; Pw.beta[k] = 0.; continue; }
                double rho = part[k], rr = part[NR + k];
                if (!(rho == rho) || !(rr == rr)) bad = true;
                Pw.beta[k] = (Pw.rho[k] != 0.) ? rho / Pw.rho[k] : 0.;
                Pw.rho_prev[k] = Pw.rho[k];
                Pw.rho[k] = rho;
                Pw.rr[k] = rr;
                if (rr < Pw.drop * Pw.drop * Pw.rr_ckpt[k]) need = true;
                if (rr <= Pw.rtol * Pw.rtol * Pw.bb[k]) need = true;
            }
            if (Pw.iters >= Pw.maxit) need = true;
            if (bad) Pw.status = ST_NAN;
            if (need) Pw.ckpt_pending = 1;
        }
    }
}

// ------------------------------------------- checkpoint 1: x64 += S y, y = 0
template <typename W, int B, int NR>
__global__ __launch_bounds__(WG) void k_ckpt_x(SolveArgs a)
{
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    W *y = vw<W>(a.y);
    const int na = *a.n_active;
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const int t = a.active[it];
        const ProbDev &P = a.prob[a.tile_prob[t]];
        if (P.status != ST_RUNNING || !P.ckpt_pending) continue;
        if (wave >= a.tile_nslices[t]) continue;
        const int s = a.tile_slice0[t] + wave;
        if (lane >= a.slice_rows[s]) continue;
        const int64_t row = (int64_t)s * WAVE + lane;
#pragma unroll
        for (int k = 0; k < NR; k++)
#pragma unroll
            for (int i = 0; i < B; i++) {
                const size_t idx = ((size_t)row * NR + k) * B + i;
                a.x64[idx] += a.S[(size_t)row * B + i] * (double)y[idx];
                y[idx] = 0;
            }
    }
}

// ----------------------- checkpoint 2: r64 = b - K x64 (FP64), r = S r64 (W)
template <typename W, int B, int NR, bool BJ>
__global__ __launch_bounds__(WG) void k_ckpt_r(SolveArgs a)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    __shared__ int flag;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    W *r = vw<W>(a.r), *z = vw<W>(a.z);
    const W *minv = cvw<W>(a.minv);
    const int na = *a.n_active;
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const int t = a.active[it];
        const int p = a.tile_prob[t];
        const ProbDev &P = a.prob[p];
        if (P.status != ST_RUNNING || !P.ckpt_pending) continue;
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
            auto src = [&](uint32_t c, int k, double *xv) {
#pragma unroll
                for (int j = 0; j < B; j++) xv[j] = a.x64[((size_t)c * NR + k) * B + j];
            };
            sell_row<double, B, NR>(a, a.v64, s, lane, acc, src);
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
        if (tile_reduce<3 * NR>(part, t, p, P, a.ticket, a.partial, lds, &flag) && threadIdx.x == 0) {
            ProbDev &Pw = a.prob[p];
            bool all = true, bad = false;
            for (int k = 0; k < NR; k++) {
                double rho = part[k], rr = part[NR + k], rr64 = part[2 * NR + k];
                if (!(rr64 == rr64)) bad = true;
                double rel = Pw.bb64[k] > 0. ? sqrt(rr64 / Pw.bb64[k]) : sqrt(rr64);
                Pw.relres[k] = rel;
                if (Pw.done[k] || rel <= Pw.rtol) {
                    Pw.done[k] = 1;
                    Pw.alpha[k] = 0.;
                    Pw.beta[k] = 0.;
                    continue;
                }
                all = false;
                // restart the recurrences (beta = 0) when the FP64 residual did
                // not halve since the previous checkpoint: the search directions
                // kept across a residual replacement have lost conjugacy
                const bool stalled = Pw.nckpt > 0 && rel > 0.5 * Pw.rel_prev[k];
                Pw.rel_prev[k] = rel;
                bool restart = Pw.refine == 1 || Pw.nckpt == 0 || stalled;
                Pw.beta[k] = (restart || Pw.rho_prev[k] == 0.) ? 0. : rho / Pw.rho_prev[k];
                Pw.rho[k] = rho;
                Pw.rr[k] = rr;
                Pw.rr_ckpt[k] = rr;
            }
            Pw.nckpt += 1;
            Pw.ckpt_pending = 0;
            if (bad) Pw.status = ST_NAN;
            else if (all) Pw.status = ST_CONVERGED;
            else if (Pw.iters >= Pw.maxit) Pw.status = ST_MAXIT;
        }
    }
}

// ------------------------------------------ norms of the right-hand sides
template <int B, int NR>
__global__ __launch_bounds__(WG) void k_rhs_norms(SolveArgs a)
{
    __shared__ double lds[TILE_SLICES][MAXQ];
    __shared__ int flag;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int na = *a.n_active;
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const int t = a.active[it];
        const int p = a.tile_prob[t];
        const ProbDev &P = a.prob[p];
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
        if (tile_reduce<2 * NR>(part, t, p, P, a.ticket, a.partial, lds, &flag) && threadIdx.x == 0) {
            ProbDev &Pw = a.prob[p];
            for (int k = 0; k < NR; k++) {
                Pw.bb[k] = part[k];
                Pw.bb64[k] = part[NR + k];
                // a homogeneous right-hand side is solved by x = 0
                // (solvers/conjugategradient.cpp:74-78)
                if (part[NR + k] == 0.) Pw.done[k] = 1;
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
            sell_row<double, B, NR>(a, a.v64, s, lane, acc, src);
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
            sell_row<W, B, NR>(a, a.vw, s, lane, acc, src);
            if (lane < a.slice_rows[s])
                for (int k = 0; k < NR; k++)
                    for (int i = 0; i < B; i++)
                        yout[((size_t)row * NR + k) * B + i] = (double)acc[k][i] / a.S[(size_t)row * B + i];
        }
    }
}

} // namespace amie
