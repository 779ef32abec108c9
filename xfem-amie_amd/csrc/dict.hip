// Value-indexed FIX48 matrix: finds the distinct blocks of the FIX48 planes
// on the device (hash every slot's block, radix-sort the hashes, number the
// runs), stores a 16-bit index per slot and a table of the distinct blocks
// decoded to FP64, then checks every slot against its table entry bit for
// bit (a hash collision makes the dictionary unavailable, never wrong).
#include <hipcub/hipcub.hpp>

#include "batch.hpp"

namespace amie {

namespace {

// 64-bit finaliser of MurmurHash3 (Austin Appleby, public domain)
__device__ inline uint64_t fmix64(uint64_t h)
{
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

__device__ inline uint64_t bits(double v)
{
    uint64_t u;
    __builtin_memcpy(&u, &v, 8);
    return u;
}

template <int B>
__global__ __launch_bounds__(256) void k_dict_hash(const void *vfx, int64_t ns, uint64_t *keys, uint32_t *idx)
{
    for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < ns; s += (int64_t)gridDim.x * 256) {
        double v[B * B];
        Blk<Fix48, B>::load(vfx, ns, s, v);
        uint64_t h = 0x9E3779B97F4A7C15ULL;
#pragma unroll
        for (int i = 0; i < B * B; i++) h = fmix64(h ^ bits(v[i]));
        keys[s] = h;
        idx[s] = (uint32_t)s;
    }
}

__global__ __launch_bounds__(256) void k_dict_head(const uint64_t *keys, int64_t n, uint32_t *head)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

template <int B>
__global__ __launch_bounds__(256) void k_dict_assign(const uint32_t *idx, const uint32_t *uid, const uint32_t *head,
                                                     int64_t n, const void *vfx, int64_t ns, uint16_t *ids,
                                                     double *table)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const uint32_t id = uid[i] - 1u;
        const int64_t s = idx[i];
        ids[s] = (uint16_t)id;
        if (head[i]) {
            double v[B * B];
            Blk<Fix48, B>::load(vfx, ns, s, v);
            double *e = table + (size_t)id * DictEntry<B>::N;
#pragma unroll
            for (int k = 0; k < DictEntry<B>::N; k++) e[k] = k < B * B ? v[k] : 0.;
        }
    }
}

template <int B>
__global__ __launch_bounds__(256) void k_dict_verify(const void *vfx, int64_t ns, const uint16_t *ids,
                                                     const double *table, int32_t *bad)
{
    for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < ns; s += (int64_t)gridDim.x * 256) {
        double v[B * B];
        Blk<Fix48, B>::load(vfx, ns, s, v);
        const double *e = table + (size_t)ids[s] * DictEntry<B>::N;
        bool diff = false;
#pragma unroll
        for (int k = 0; k < B * B; k++) diff |= bits(v[k]) != bits(e[k]);
        if (diff) *bad = 1;   // every writer stores the same value
    }
}

} // namespace

int build_dict(Batch &B)
{
    B.dict_ok = false;
    B.dict_n = 0;
    const int64_t n = B.nslots;
    if (n <= 0 || n > INT32_MAX || !B.fix48_ok) return AMIE_OK;
    hipStream_t st = B.stream;
    const int b = B.et.b;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    DevBuf keys, keys2, idx, idx2, head, uid, tmp, bad;
    if (!keys.alloc(n * 8) || !keys2.alloc(n * 8) || !idx.alloc(n * 4) || !idx2.alloc(n * 4) || !head.alloc(n * 4) ||
        !uid.alloc(n * 4) || !bad.alloc(4))
        return AMIE_ERR_NOMEM;
    if (b == 2) hipLaunchKernelGGL(k_dict_hash<2>, dim3(grid), dim3(256), 0, st, B.vfx.p, n, keys.as<uint64_t>(), idx.as<uint32_t>());
    else hipLaunchKernelGGL(k_dict_hash<3>, dim3(grid), dim3(256), 0, st, B.vfx.p, n, keys.as<uint64_t>(), idx.as<uint32_t>());
    size_t need_sort = 0, need_scan = 0;
    AMIE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need_sort, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                                idx.as<uint32_t>(), idx2.as<uint32_t>(), (int)n, 0, 64, st));
    AMIE_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, need_scan, head.as<uint32_t>(), uid.as<uint32_t>(), (int)n, st));
    if (!tmp.alloc(std::max<size_t>(std::max(need_sort, need_scan), 256))) return AMIE_ERR_NOMEM;
    size_t have = tmp.n;
    AMIE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, have, keys.as<uint64_t>(), keys2.as<uint64_t>(),
                                                idx.as<uint32_t>(), idx2.as<uint32_t>(), (int)n, 0, 64, st));
    hipLaunchKernelGGL(k_dict_head, dim3(grid), dim3(256), 0, st, keys2.as<uint64_t>(), n, head.as<uint32_t>());
    have = tmp.n;
    AMIE_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, have, head.as<uint32_t>(), uid.as<uint32_t>(), (int)n, st));
    uint32_t nuniq = 0;
    AMIE_HIP(hipMemcpyAsync(&nuniq, uid.as<uint32_t>() + (n - 1), 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipStreamSynchronize(st));
    if (nuniq == 0 || nuniq > (uint32_t)DICT_MAX) {
        B.dict_wanted = false;   // too many distinct blocks (a damaged mesh, say): stop trying
        return AMIE_OK;
    }
    const size_t tbytes = (size_t)nuniq * (b == 2 ? DictEntry<2>::N : DictEntry<3>::N) * 8;
    if (B.dict_ids.n != (size_t)n * 2 && !B.dict_ids.alloc(n * 2)) return AMIE_ERR_NOMEM;
    if (B.dict_table.n < tbytes && !B.dict_table.alloc(tbytes)) return AMIE_ERR_NOMEM;
    if (b == 2)
        hipLaunchKernelGGL(k_dict_assign<2>, dim3(grid), dim3(256), 0, st, idx2.as<uint32_t>(), uid.as<uint32_t>(),
                           head.as<uint32_t>(), n, B.vfx.p, n, B.dict_ids.as<uint16_t>(), B.dict_table.as<double>());
    else
        hipLaunchKernelGGL(k_dict_assign<3>, dim3(grid), dim3(256), 0, st, idx2.as<uint32_t>(), uid.as<uint32_t>(),
                           head.as<uint32_t>(), n, B.vfx.p, n, B.dict_ids.as<uint16_t>(), B.dict_table.as<double>());
    AMIE_HIP(hipMemsetAsync(bad.p, 0, 4, st));
    if (b == 2)
        hipLaunchKernelGGL(k_dict_verify<2>, dim3(grid), dim3(256), 0, st, B.vfx.p, n, B.dict_ids.as<uint16_t>(),
                           B.dict_table.as<double>(), bad.as<int32_t>());
    else
        hipLaunchKernelGGL(k_dict_verify<3>, dim3(grid), dim3(256), 0, st, B.vfx.p, n, B.dict_ids.as<uint16_t>(),
                           B.dict_table.as<double>(), bad.as<int32_t>());
    AMIE_HIP(hipGetLastError());
    int32_t hbad = 1;
    AMIE_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipStreamSynchronize(st));
    B.dict_ok = hbad == 0;
    B.dict_n = B.dict_ok ? nuniq : 0;
    return AMIE_OK;
}

} // namespace amie
