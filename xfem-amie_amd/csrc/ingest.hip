// Mesh input on the device: the reference's node numbering and its P2
// mid-edge nodes, for meshes handed over as element lists.
//
// FeatureTree::renumber (features/features.cpp:603-729) walks the elements in
// order and gives each node the next id the first time it is met.  In
// parallel: every position q = e*nper + j of the connectivity proposes
// itself to its node (atomic minimum, order-free), the positions that won
// are the first touches, and an exclusive scan over the win flags numbers
// them in position order, which is the serial loop's order.
//
// DelaunayTree::addSharedNodes(1, 1, ...) (mesher/delaunay.cpp:260-355) gives
// side s of triangle t the mid-edge node an earlier triangle made on the same
// edge, else a new one at a*0.5 + b*0.5.  In parallel: the 3*ntri sides are
// radix-sorted (stable) by edge, the first side of each run (the smallest
// position) makes the node, and a scan over "makes a node" flags numbers the
// nodes in position order, which is the order the serial loop makes them.
//
// Both are integer passes over the connectivity (HBM-bound); results are
// bit-exact and do not depend on the launch.
#include <hipcub/hipcub.hpp>

#include "batch.hpp"

namespace amie {

namespace {

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// first[node] = smallest position naming the node
__global__ void k_first_touch(const int32_t *conn, int64_t n, uint32_t *first)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) atomicMin(&first[conn[q]], (uint32_t)q);
}

__global__ void k_touch_flags(const int32_t *conn, int64_t n, const uint32_t *first, uint32_t *flag)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) flag[q] = first[conn[q]] == (uint32_t)q;
}

__global__ void k_assign_ids(const int32_t *conn, int64_t n, const uint32_t *flag, const uint32_t *rank, int32_t *ids)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n && flag[q]) ids[conn[q]] = (int32_t)rank[q];
}

// side q = 3t + s: edge key lo*nnodes + hi, value q
__global__ void k_side_keys(const int32_t *conn3, int64_t nsides, int64_t nnodes, uint64_t *key, uint32_t *val)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsides) return;
    const int64_t t = q / 3;
    const int s = (int)(q - t * 3);
    const uint32_t a = (uint32_t)conn3[q], b = (uint32_t)conn3[t * 3 + (s + 1) % 3];
    const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
    key[q] = lo * (uint64_t)nnodes + hi;
    val[q] = (uint32_t)q;
}

// after the stable sort: a run of one edge starts where the key changes;
// head[i] = i there, 0 elsewhere, and a running maximum gives every sorted
// side the start of its run
__global__ void k_run_heads(const uint64_t *skey, int64_t nsides, uint32_t *head)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsides) head[i] = (i == 0 || skey[i] != skey[i - 1]) ? (uint32_t)i : 0u;
}

// every side learns the first side of its edge (runs hold ascending
// positions, so the run head is the smallest)
__global__ void k_side_first(const uint32_t *sval, const uint32_t *hmax, int64_t nsides, uint32_t *first, uint32_t *makes)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsides) return;
    const uint32_t h = hmax[i];
    first[sval[i]] = sval[h];
    makes[sval[i]] = h == (uint32_t)i;
}

struct MaxOp {
    __device__ __host__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

__global__ void k_p2_conn(const int32_t *conn3, int64_t nsides, const uint32_t *first, const uint32_t *rank,
                          int64_t nnodes, const double *xy, int32_t *conn6, double *mid)
{
#pragma clang fp contract(off)
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsides) return;
    const int64_t t = q / 3;
    const int s = (int)(q - t * 3);
    const int32_t a = conn3[q], b = conn3[t * 3 + (s + 1) % 3];
    const uint32_t f = first[q];
    conn6[t * 6 + 2 * s] = a;
    conn6[t * 6 + 2 * s + 1] = (int32_t)(nnodes + rank[f]);
    if (f == (uint32_t)q) {
        const int64_t m = rank[q];
        for (int d = 0; d < 2; d++) mid[m * 2 + d] = xy[(int64_t)a * 2 + d] * (1. - .5) + xy[(int64_t)b * 2 + d] * .5;
    }
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

struct Timer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Timer()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// the connectivity names nodes in [0, nnodes) only: checked on the host
// before any kernel indexes by it
bool conn_in_range(const int32_t *conn, int64_t n, int64_t nnodes)
{
    int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
    for (int64_t q = 0; q < n; q++) bad |= conn[q] < 0 || conn[q] >= nnodes;
    return !bad;
}

} // namespace

} // namespace amie

using namespace amie;

int amie_gpu_renumber(amie_gpu_ctx *ctx, int64_t nnodes, int64_t nelem, int32_t nper, const int32_t *conn, int32_t *ids,
                      int64_t *count, double *device_ms)
{
    if (!ctx || !conn || !ids || nnodes <= 0 || nnodes > INT32_MAX || nelem <= 0 || nper <= 0) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    const int64_t n = nelem * nper;
    if (n >= INT32_MAX) { set_error("more than 2^31 - 1 connectivity entries"); return AMIE_ERR_ARG; }
    if (!conn_in_range(conn, n, nnodes)) { set_error("connectivity names a node outside [0, nnodes)"); return AMIE_ERR_ARG; }
    AMIE_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf dconn, dfirst, dflag, drank, dids, tmp;
    if (!dconn.alloc(n * 4) || !dfirst.alloc(nnodes * 4) || !dflag.alloc(n * 4) || !drank.alloc(n * 4) ||
        !dids.alloc(nnodes * 4))
        return AMIE_ERR_NOMEM;
    size_t need = 0;
    AMIE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, dflag.as<uint32_t>(), drank.as<uint32_t>(), (int)n, st));
    if (!tmp.alloc(need ? need : 4)) return AMIE_ERR_NOMEM;
    AMIE_HIP(hipMemcpyAsync(dconn.p, conn, n * 4, hipMemcpyHostToDevice, st));
    Timer tm;
    if (device_ms) {
        AMIE_HIP(hipEventCreate(&tm.e0));
        AMIE_HIP(hipEventCreate(&tm.e1));
        AMIE_HIP(hipEventRecord(tm.e0, st));
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(nnodes)), dim3(256), 0, st, dfirst.as<uint32_t>(), nnodes, 0xffffffffu);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(nnodes)), dim3(256), 0, st, dids.as<int32_t>(), nnodes, -1);
    hipLaunchKernelGGL(k_first_touch, dim3(blocks(n)), dim3(256), 0, st, dconn.as<int32_t>(), n, dfirst.as<uint32_t>());
    hipLaunchKernelGGL(k_touch_flags, dim3(blocks(n)), dim3(256), 0, st, dconn.as<int32_t>(), n, dfirst.as<uint32_t>(),
                       dflag.as<uint32_t>());
    AMIE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, need, dflag.as<uint32_t>(), drank.as<uint32_t>(), (int)n, st));
    hipLaunchKernelGGL(k_assign_ids, dim3(blocks(n)), dim3(256), 0, st, dconn.as<int32_t>(), n, dflag.as<uint32_t>(),
                       drank.as<uint32_t>(), dids.as<int32_t>());
    AMIE_HIP(hipGetLastError());
    if (device_ms) AMIE_HIP(hipEventRecord(tm.e1, st));
    uint32_t last_flag = 0, last_rank = 0;
    AMIE_HIP(hipMemcpyAsync(&last_flag, dflag.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipMemcpyAsync(&last_rank, drank.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipMemcpyAsync(ids, dids.p, nnodes * 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipStreamSynchronize(st));
    if (count) *count = (int64_t)last_rank + last_flag;
    if (device_ms) {
        float ms = 0.f;
        AMIE_HIP(hipEventElapsedTime(&ms, tm.e0, tm.e1));
        *device_ms = ms;
    }
    return AMIE_OK;
}

int amie_gpu_add_midpoints(amie_gpu_ctx *ctx, int64_t nnodes, const double *xy, int64_t ntri, const int32_t *conn3,
                           int32_t *conn6, double *mid_xy, int64_t *nmid, double *device_ms)
{
    if (!ctx || !xy || !conn3 || !conn6 || !mid_xy || nnodes <= 0 || ntri <= 0) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    const int64_t ns = ntri * 3;
    if (ns >= INT32_MAX || nnodes + ns > INT32_MAX) { set_error("too many triangles or nodes for 32-bit ids"); return AMIE_ERR_ARG; }
    if (!conn_in_range(conn3, ns, nnodes)) { set_error("connectivity names a node outside [0, nnodes)"); return AMIE_ERR_ARG; }
    AMIE_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf dconn, dxy, key, key2, val, val2, first, makes, rank, dconn6, dmid, tmp, head, hmax;
    if (!dconn.alloc(ns * 4) || !dxy.alloc(nnodes * 16) || !key.alloc(ns * 8) || !key2.alloc(ns * 8) ||
        !val.alloc(ns * 4) || !val2.alloc(ns * 4) || !first.alloc(ns * 4) || !makes.alloc(ns * 4) ||
        !rank.alloc(ns * 4) || !dconn6.alloc(ntri * 24) || !dmid.alloc(ns * 16) || !head.alloc(ns * 4) ||
        !hmax.alloc(ns * 4))
        return AMIE_ERR_NOMEM;
    // edge keys < nnodes^2: sort only the bits that can be set
    const uint64_t kmax = (uint64_t)nnodes * (uint64_t)nnodes - 1;
    int end_bit = 1;
    while (end_bit < 64 && (kmax >> end_bit)) end_bit++;
    size_t need_sort = 0, need_scan = 0, need_max = 0;
    AMIE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need_sort, key.as<uint64_t>(), key2.as<uint64_t>(),
                                                val.as<uint32_t>(), val2.as<uint32_t>(), (int)ns, 0, end_bit, st));
    AMIE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need_scan, makes.as<uint32_t>(), rank.as<uint32_t>(), (int)ns, st));
    AMIE_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, need_max, head.as<uint32_t>(), hmax.as<uint32_t>(), MaxOp(),
                                               (int)ns, st));
    const size_t need = std::max(std::max(need_sort, need_scan), need_max);
    if (!tmp.alloc(need ? need : 4)) return AMIE_ERR_NOMEM;
    AMIE_HIP(hipMemcpyAsync(dconn.p, conn3, ns * 4, hipMemcpyHostToDevice, st));
    AMIE_HIP(hipMemcpyAsync(dxy.p, xy, nnodes * 16, hipMemcpyHostToDevice, st));
    Timer tm;
    if (device_ms) {
        AMIE_HIP(hipEventCreate(&tm.e0));
        AMIE_HIP(hipEventCreate(&tm.e1));
        AMIE_HIP(hipEventRecord(tm.e0, st));
    }
    hipLaunchKernelGGL(k_side_keys, dim3(blocks(ns)), dim3(256), 0, st, dconn.as<int32_t>(), ns, nnodes,
                       key.as<uint64_t>(), val.as<uint32_t>());
    size_t have = need;
    AMIE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, have, key.as<uint64_t>(), key2.as<uint64_t>(), val.as<uint32_t>(),
                                                val2.as<uint32_t>(), (int)ns, 0, end_bit, st));
    hipLaunchKernelGGL(k_run_heads, dim3(blocks(ns)), dim3(256), 0, st, key2.as<uint64_t>(), ns, head.as<uint32_t>());
    have = need;
    AMIE_HIP(hipcub::DeviceScan::InclusiveScan(tmp.p, have, head.as<uint32_t>(), hmax.as<uint32_t>(), MaxOp(), (int)ns, st));
    hipLaunchKernelGGL(k_side_first, dim3(blocks(ns)), dim3(256), 0, st, val2.as<uint32_t>(), hmax.as<uint32_t>(), ns,
                       first.as<uint32_t>(), makes.as<uint32_t>());
    have = need;
    AMIE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, have, makes.as<uint32_t>(), rank.as<uint32_t>(), (int)ns, st));
    hipLaunchKernelGGL(k_p2_conn, dim3(blocks(ns)), dim3(256), 0, st, dconn.as<int32_t>(), ns, first.as<uint32_t>(),
                       rank.as<uint32_t>(), nnodes, dxy.as<double>(), dconn6.as<int32_t>(), dmid.as<double>());
    AMIE_HIP(hipGetLastError());
    if (device_ms) AMIE_HIP(hipEventRecord(tm.e1, st));
    uint32_t last_make = 0, last_rank = 0;
    AMIE_HIP(hipMemcpyAsync(&last_make, makes.as<uint32_t>() + ns - 1, 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipMemcpyAsync(&last_rank, rank.as<uint32_t>() + ns - 1, 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipMemcpyAsync(conn6, dconn6.p, ntri * 24, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipStreamSynchronize(st));
    const int64_t nm = (int64_t)last_rank + last_make;
    AMIE_HIP(hipMemcpy(mid_xy, dmid.p, nm * 16, hipMemcpyDeviceToHost));
    if (nmid) *nmid = nm;
    if (device_ms) {
        float ms = 0.f;
        AMIE_HIP(hipEventElapsedTime(&ms, tm.e0, tm.e1));
        *device_ms = ms;
    }
    return AMIE_OK;
}
