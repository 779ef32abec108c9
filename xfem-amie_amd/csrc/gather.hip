// Contribution lists of the gather assembly, built on the device at batch
// creation.  Every (element, block) pair becomes a (slot, code) pair, code =
// (e*nper + i)*nper + j, laid out in the reference's summation order
// (solvers/assembly.cpp:657-735: elements in order, inside one element
// (j,j), then (j,k), (k,j) for k > j); a stable radix sort by slot keeps that
// order inside each slot's list, and a count + scan gives the list offsets.
// The macro-load lists (element, node) per internal row are built the same
// way from the element -> row map.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "batch.hpp"

namespace amie {

namespace {

// the m-th block pair of an element in the reference's order
__device__ inline void pair_of(int m, int nper, int &a, int &b)
{
    int j = 0;
    while (m >= 1 + 2 * (nper - 1 - j)) {
        m -= 1 + 2 * (nper - 1 - j);
        j++;
    }
    if (m == 0) { a = j; b = j; return; }
    const int k = j + 1 + (m - 1) / 2;
    if ((m - 1) % 2 == 0) { a = j; b = k; }
    else { a = k; b = j; }
}

__global__ void k_pair_keys(const uint32_t *eslot, int64_t n, int nper, uint32_t *key, uint32_t *val, uint32_t *cnt)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t e = q / (nper * nper);
    int a, b;
    pair_of((int)(q - e * nper * nper), nper, a, b);
    const int64_t pos = (e * nper + a) * nper + b;
    key[q] = eslot[pos];
    val[q] = (uint32_t)pos;
    atomicAdd(&cnt[eslot[pos]], 1u);
}

__global__ void k_row_keys(const int32_t *erow, int64_t n, uint32_t *key, uint32_t *val, uint32_t *cnt)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    key[q] = (uint32_t)erow[q];
    val[q] = (uint32_t)q;
    atomicAdd(&cnt[erow[q]], 1u);
}

__global__ void k_zero(uint32_t *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

int bits_for(int64_t nkeys)
{
    int b = 1;
    while (b < 32 && ((uint64_t)(nkeys - 1) >> b)) b++;
    return b;
}

// keys in [0, nkeys): list offsets ptr[nkeys + 1] and values sorted by key
// (stable), from n (key, value) pairs made by `make`
template <typename Make>
int sorted_lists(Batch &B, int64_t n, int64_t nkeys, DevBuf &ptr, DevBuf &sorted, Make make)
{
    hipStream_t st = B.stream;
    DevBuf key, key2, val, cnt, tmp;
    if (!key.alloc(n * 4) || !key2.alloc(n * 4) || !val.alloc(n * 4) || !cnt.alloc((nkeys + 1) * 4) ||
        !ptr.alloc((nkeys + 1) * 4) || !sorted.alloc(n * 4))
        return AMIE_ERR_NOMEM;
    hipLaunchKernelGGL(k_zero, dim3(blocks(nkeys + 1)), dim3(256), 0, st, cnt.as<uint32_t>(), nkeys + 1);
    make(key.as<uint32_t>(), val.as<uint32_t>(), cnt.as<uint32_t>());
    const int end_bit = bits_for(nkeys);
    size_t need_sort = 0, need_scan = 0;
    AMIE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need_sort, key.as<uint32_t>(), key2.as<uint32_t>(),
                                                val.as<uint32_t>(), sorted.as<uint32_t>(), (int)n, 0, end_bit, st));
    AMIE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need_scan, cnt.as<uint32_t>(), ptr.as<uint32_t>(),
                                              (int)(nkeys + 1), st));
    const size_t need = std::max(need_sort, need_scan);
    if (!tmp.alloc(need ? need : 4)) return AMIE_ERR_NOMEM;
    size_t have = need;
    AMIE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, have, key.as<uint32_t>(), key2.as<uint32_t>(),
                                                val.as<uint32_t>(), sorted.as<uint32_t>(), (int)n, 0, end_bit, st));
    have = need;
    AMIE_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, have, cnt.as<uint32_t>(), ptr.as<uint32_t>(), (int)(nkeys + 1), st));
    AMIE_HIP(hipGetLastError());
    AMIE_HIP(hipStreamSynchronize(st));
    return AMIE_OK;
}

} // namespace

bool gather_fits(int64_t nelem, int nper, int64_t nslots)
{
    return nelem * nper * nper < INT32_MAX && nslots < INT32_MAX;
}

int build_gather_lists(Batch &B)
{
    const int nper = B.et.nper;
    const int64_t n = B.nelem * nper * nper;
    int rc = sorted_lists(B, n, B.nslots, B.slot_cptr, B.slot_contrib, [&](uint32_t *key, uint32_t *val, uint32_t *cnt) {
        hipLaunchKernelGGL(k_pair_keys, dim3(blocks(n)), dim3(256), 0, B.stream, B.eslot.as<uint32_t>(), n, nper, key, val,
                           cnt);
    });
    if (rc || !B.has_macro) return rc;
    const int64_t nr = B.nelem * nper;
    return sorted_lists(B, nr, B.nrows_pad, B.row_cptr, B.row_contrib, [&](uint32_t *key, uint32_t *val, uint32_t *cnt) {
        hipLaunchKernelGGL(k_row_keys, dim3(blocks(nr)), dim3(256), 0, B.stream, B.erow.as<int32_t>(), nr, key, val, cnt);
    });
}

} // namespace amie
