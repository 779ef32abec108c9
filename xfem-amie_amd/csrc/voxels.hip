// VoxelWriter sampling on the device (utilities/writer/voxel_writer.cpp:
// 88-272).  The reference walks the tetrahedra in order and lets every
// tetrahedron overwrite the voxels of its bounding-box index range that it
// contains, so a voxel ends with the last containing tetrahedron.  Here one
// thread per tetrahedron visits the same index range and raises the voxel's
// owner to its index with an atomic maximum (order-free, the same winner);
// a second pass gathers the owner's values.
#include <mutex>

#include "batch.hpp"

namespace amie {

// Ownership array of amie_gpu_voxelize_noalloc: one int32 per voxel, a
// static device array of AMIE_VOXEL_SCRATCH_MAX entries, so that call
// allocates nothing.  Each GPU holds its own copy at its own address until
// the process ends; host code uses the address hipGetSymbolAddress gives on
// that device (voxel_scratch_address), never the array's name.
static __device__ __align__(256) int32_t scratch[AMIE_VOXEL_SCRATCH_MAX];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
int32_t *scratch_ptr[MAX_DEVICES] = {};
} // namespace

int voxel_scratch_address(int device, int32_t **ptr)
{
    if (device < 0 || device >= MAX_DEVICES) { set_error("device %d: at most %d devices", device, MAX_DEVICES); return AMIE_ERR_ARG; }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress of the voxel scratch array: %s", hipGetErrorString(e));
            return AMIE_ERR_HIP;
        }
        scratch_ptr[device] = static_cast<int32_t *>(p);
    }
    *ptr = scratch_ptr[device];
    return AMIE_OK;
}

namespace {

__device__ inline double det3(const double *a, const double *b, const double *c)
{
#pragma clang fp contract(off)
    return a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

// barycentric weights of vertices 2, 3, 0 (the reference's column order,
// geometry_3D.cpp:451-533), by volumes; tolerance 1e-12 per weight, 3e-12 on
// their sum
__device__ inline bool in_tet(const double (*P)[3], const double *p)
{
#pragma clang fp contract(off)
    double e[3][3], q[3][3], f[3];
    for (int d = 0; d < 3; d++) {
        e[0][d] = P[3][d] - P[2][d];
        e[1][d] = P[0][d] - P[2][d];
        e[2][d] = P[1][d] - P[2][d];
    }
    const double V = det3(e[0], e[1], e[2]);
    if (V == 0.) return false;
    for (int d = 0; d < 3; d++) {
        q[0][d] = P[3][d] - p[d];
        q[1][d] = P[0][d] - p[d];
        q[2][d] = P[1][d] - p[d];
        f[d] = p[d] - P[2][d];
    }
    const double l2 = det3(q[0], q[1], q[2]) / V;
    const double l3 = det3(f, e[1], e[2]) / V;
    const double l0 = det3(e[0], f, e[2]) / V;
    if (l2 < -1e-12 || l3 < -1e-12 || l0 < -1e-12) return false;
    return !(l2 + l3 + l0 > 1 + 3. * 1e-12);
}

struct Grid {
    double bl[3], tr[3];
    int n[3];
};

__global__ void k_fill_owner(int32_t *owner, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) owner[i] = -1;
}

// First index of a voxel range that starts at s (the reference truncates s
// to an int; indices below 0 are skipped there): clamped to [0, n] before
// the conversion, so a tetrahedron far outside the grid loops over nothing.
__device__ inline int first_index(double s, int n)
{
    return !(s > 0.) ? 0 : (s >= (double)n ? n : (int)s);
}

// Tetrahedra that name a node outside [0, nnodes) own nothing (the host
// entry point refuses them before the launch; the device-array entry point
// cannot read the connectivity on the host).
__global__ void k_voxel_owner(const double *xyz, const int32_t *conn, int64_t ntet, int64_t nnodes, Grid g,
                              int32_t *owner)
{
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntet) return;
    double P[4][3], lo[3], hi[3];
    for (int a = 0; a < 4; a++)
        if (conn[t * 4 + a] < 0 || conn[t * 4 + a] >= nnodes) return;
    for (int a = 0; a < 4; a++)
        for (int d = 0; d < 3; d++) P[a][d] = xyz[(int64_t)conn[t * 4 + a] * 3 + d];
    for (int d = 0; d < 3; d++) {
        lo[d] = hi[d] = P[0][d];
        for (int a = 1; a < 4; a++) {
            lo[d] = fmin(lo[d], P[a][d]);
            hi[d] = fmax(hi[d], P[a][d]);
        }
    }
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const double ei = nx * (hi[0] - g.bl[0]) / (g.tr[0] - g.bl[0]);
    const double ej = ny * (hi[1] - g.bl[1]) / (g.tr[1] - g.bl[1]);
    const double ek = nz * (hi[2] - g.bl[2]) / (g.tr[2] - g.bl[2]);
    const int i0 = first_index(nx * (lo[0] - g.bl[0]) / (g.tr[0] - g.bl[0]), nx);
    const int j0 = first_index(ny * (lo[1] - g.bl[1]) / (g.tr[1] - g.bl[1]), ny);
    const int k0 = first_index(nz * (lo[2] - g.bl[2]) / (g.tr[2] - g.bl[2]), nz);
    for (int i = i0; i < nx && i < ei; i++) {
        for (int j = j0; j < ny && j < ej; j++) {
            for (int k = k0; k < nz && k < ek; k++) {
                double p[3] = {g.bl[0], g.bl[1], g.bl[2]};
                p[0] += (g.tr[0] - g.bl[0]) * ((double)i) / (double)(nx - 1);
                p[1] += (g.tr[1] - g.bl[1]) * ((double)j) / (double)(ny - 1);
                p[2] += (g.tr[2] - g.bl[2]) * ((double)k) / (double)(nz - 1);
                bool in = in_tet(P, p);
                for (int m = 0; m < 8 && !in; m++) {
                    const double q[3] = {p[0] + ((m & 4) ? 16. : -16.) * 1e-12, p[1] + ((m & 2) ? 16. : -16.) * 1e-12,
                                         p[2] + ((m & 1) ? 16. : -16.) * 1e-12};
                    in = in_tet(P, q);
                }
                if (in) atomicMax(&owner[k + (int64_t)ny * j + (int64_t)ny * nx * i], (int32_t)t);
            }
        }
    }
}

__global__ void k_voxel_values(const int32_t *owner, int64_t nvox, int ncomp, const double *field, double *values)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int32_t t = owner[v];
    for (int c = 0; c < ncomp; c++) values[(int64_t)c * nvox + v] = t >= 0 ? field[(int64_t)t * ncomp + c] : -1e9;
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// The grid of voxel_writer.cpp:88-104: corners c -/+ 0.5 s (n-1)/n.
int make_grid(const double *box, int nx, int ny, int nz, Grid &g)
{
    g.n[0] = nx; g.n[1] = ny; g.n[2] = nz;
    for (int d = 0; d < 3; d++) {
        if (!(box[3 + d] > 0)) { set_error("box size must be positive"); return AMIE_ERR_ARG; }
        const double v = box[3 + d] * (g.n[d] - 1) / (g.n[d]) * 0.5;
        g.bl[d] = box[d] - v;
        g.tr[d] = box[d] + v;
    }
    return AMIE_OK;
}

} // namespace

} // namespace amie

using namespace amie;

int amie_gpu_voxelize(amie_gpu_ctx *ctx, int64_t nnodes, const double *xyz, int64_t ntet, const int32_t *conn,
                      int32_t ncomp, const double *field, const double *box, int32_t nx, int32_t ny, int32_t nz,
                      double *values, double *device_ms)
{
    if (!ctx || !xyz || !conn || !field || !box || !values || nnodes <= 0 || ntet <= 0 || ntet >= INT32_MAX || ncomp <= 0 ||
        nx <= 1 || ny <= 1 || nz <= 1 || (int64_t)nx * ny * nz >= INT32_MAX) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    // the reference's voxel index k + ny*j + ny*nx*i numbers a cubic grid
    // only; other shapes would run past the end
    if (nx != ny || ny != nz) { set_error("the voxel grid must be cubic (nx = ny = nz)"); return AMIE_ERR_ARG; }
    for (int64_t q = 0; q < ntet * 4; q++)
        if (conn[q] < 0 || conn[q] >= nnodes) { set_error("connectivity names a node outside [0, nnodes)"); return AMIE_ERR_ARG; }
    Grid g;
    if (int rc = make_grid(box, nx, ny, nz, g)) return rc;
    AMIE_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t nvox = (int64_t)nx * ny * nz;
    DevBuf dxyz, dconn, dfield, owner, dval;
    if (!dxyz.alloc(nnodes * 24) || !dconn.alloc(ntet * 16) || !dfield.alloc(ntet * ncomp * 8) || !owner.alloc(nvox * 4) ||
        !dval.alloc(nvox * ncomp * 8))
        return AMIE_ERR_NOMEM;
    AMIE_HIP(hipMemcpyAsync(dxyz.p, xyz, nnodes * 24, hipMemcpyHostToDevice, st));
    AMIE_HIP(hipMemcpyAsync(dconn.p, conn, ntet * 16, hipMemcpyHostToDevice, st));
    AMIE_HIP(hipMemcpyAsync(dfield.p, field, ntet * ncomp * 8, hipMemcpyHostToDevice, st));
    struct Events {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    hipEvent_t &e0 = ev.e0, &e1 = ev.e1;
    if (device_ms) {
        AMIE_HIP(hipEventCreate(&e0));
        AMIE_HIP(hipEventCreate(&e1));
        AMIE_HIP(hipEventRecord(e0, st));
    }
    hipLaunchKernelGGL(k_fill_owner, dim3(blocks(nvox)), dim3(256), 0, st, owner.as<int32_t>(), nvox);
    hipLaunchKernelGGL(k_voxel_owner, dim3(blocks(ntet)), dim3(256), 0, st, dxyz.as<double>(), dconn.as<int32_t>(), ntet,
                       nnodes, g, owner.as<int32_t>());
    hipLaunchKernelGGL(k_voxel_values, dim3(blocks(nvox)), dim3(256), 0, st, owner.as<int32_t>(), nvox, (int)ncomp,
                       dfield.as<double>(), dval.as<double>());
    AMIE_HIP(hipGetLastError());
    if (device_ms) AMIE_HIP(hipEventRecord(e1, st));
    AMIE_HIP(hipMemcpyAsync(values, dval.p, nvox * ncomp * 8, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipStreamSynchronize(st));
    if (device_ms) {
        float ms = 0.f;
        AMIE_HIP(hipEventElapsedTime(&ms, e0, e1));
        *device_ms = ms;
    }
    return AMIE_OK;
}

int amie_gpu_voxelize_noalloc(amie_gpu_ctx *ctx, void *stream, int64_t nnodes, const double *xyz, int64_t ntet,
                              const int32_t *conn, int32_t ncomp, const double *field, const double *box, int32_t nx,
                              int32_t ny, int32_t nz, double *values)
{
    if (!ctx || !xyz || !conn || !field || !box || !values || nnodes <= 0 || ntet <= 0 || ntet >= INT32_MAX ||
        ncomp <= 0 || nx <= 1 || ny <= 1 || nz <= 1) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    if (nx != ny || ny != nz) { set_error("the voxel grid must be cubic (nx = ny = nz)"); return AMIE_ERR_ARG; }
    const int64_t nvox = (int64_t)nx * ny * nz;
    if (nvox > AMIE_VOXEL_SCRATCH_MAX) {
        set_error("%lld voxels: amie_gpu_voxelize_noalloc takes at most %lld (AMIE_VOXEL_SCRATCH_MAX)", (long long)nvox,
                  (long long)AMIE_VOXEL_SCRATCH_MAX);
        return AMIE_ERR_ARG;
    }
    if (!ctx->voxel_scratch) { set_error("the context has no voxel scratch array"); return AMIE_ERR_STATE; }
    Grid g;
    if (int rc = make_grid(box, nx, ny, nz, g)) return rc;
    AMIE_HIP(hipSetDevice(ctx->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    int32_t *owner = ctx->voxel_scratch;
    hipLaunchKernelGGL(k_fill_owner, dim3(blocks(nvox)), dim3(256), 0, st, owner, nvox);
    hipLaunchKernelGGL(k_voxel_owner, dim3(blocks(ntet)), dim3(256), 0, st, xyz, conn, ntet, nnodes, g, owner);
    hipLaunchKernelGGL(k_voxel_values, dim3(blocks(nvox)), dim3(256), 0, st, owner, nvox, (int)ncomp, field, values);
    AMIE_HIP(hipGetLastError());
    return AMIE_OK;
}
