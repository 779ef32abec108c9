// Result files in the reference's TriangleWriter text format, for its
// viewers.  Host code: the fields are computed on the device and copied out
// first (amie_gpu_batch_get_solution, amie_gpu_batch_element_strain, ...).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "amie_gpu.h"

namespace amie {
void set_error(const char *fmt, ...);
}

// utilities/writer/triangle_writer.cpp: the header (:1461-1475) counts the
// columns after the six coordinates in threes; the constructor (:31-86)
// collects the coordinates and the displacements, each field adds its
// columns (:578-608) component by component, vertex 0, 1, 2 within a
// component (:1024-1075, reversed at :589), and write() (:407-445) prints
// every column of a triangle as a C++ stream prints a double, each followed
// by a space, one triangle a line.
int amie_write_triangles(const char *path, int64_t ntri, int32_t nper, const int32_t *conn, int64_t nnodes,
                         const double *xy, const double *uv, int32_t nfield, const int32_t *sizes,
                         const double *const *fields)
{
    if (!path || ntri <= 0 || nper <= 0 || nper % 3 || !conn || nnodes <= 0 || !xy || !uv || nfield < 0 ||
        (nfield && (!sizes || !fields))) {
        amie::set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    for (int64_t q = 0; q < ntri * nper; q++)
        if (conn[q] < 0 || conn[q] >= nnodes) { amie::set_error("connectivity names a node outside [0, nnodes)"); return AMIE_ERR_ARG; }
    int64_t cols = 0;
    for (int f = 0; f < nfield; f++) {
        if (sizes[f] <= 0 || !fields[f]) { amie::set_error("field %d: bad size or data", f); return AMIE_ERR_ARG; }
        cols += sizes[f];
    }
    std::ofstream out(path, std::ios::out);
    if (!out) { amie::set_error("cannot open %s", path); return AMIE_ERR_ARG; }
    out << "TRIANGLES" << std::endl;
    out << (int)ntri << std::endl;
    out << 3 << std::endl;
    out << (int)(2 + cols) << std::endl;
    const int n = nper / 3;  // vertices at local 0, n, 2n
    for (int64_t t = 0; t < ntri; t++) {
        const int32_t v[3] = {conn[t * nper], conn[t * nper + n], conn[t * nper + 2 * n]};
        for (int k = 0; k < 3; k++) out << xy[(int64_t)v[k] * 2] << " " << xy[(int64_t)v[k] * 2 + 1] << " ";
        for (int c = 0; c < 2; c++)
            for (int k = 0; k < 3; k++) out << uv[t * 6 + k * 2 + c] << " ";
        for (int f = 0; f < nfield; f++) {
            const int s = sizes[f];
            const double *d = fields[f] + t * 3 * s;
            for (int c = 0; c < s; c++)
                for (int k = 0; k < 3; k++) out << d[k * s + c] << " ";
        }
        out << std::endl;
    }
    out.close();
    if (!out) { amie::set_error("write to %s failed", path); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

// VoxelWriter output (utilities/writer/voxel_writer.cpp:52-86, 433-442,
// 548-571): each component to <base>_<nx>_<c>.vox, a text header ("VOXELS",
// 1, nx, ny, nz, one a line) and one byte per voxel in index order.  The
// byte is the value mapped from [vmin, vmax] to [1, 255] and rounded, where
// vmax and vmin are the values at the 99 % and 1 % positions of the sorted
// component, widened to the symmetric range of the larger magnitude
// (normalizeArray, min 1, max 255); voxels no tetrahedron owns keep the
// sampling's -1e9 and take part in the percentiles as the reference's do.
int amie_write_voxels(const char *base, int32_t nx, int32_t ny, int32_t nz, int32_t ncomp, const double *values)
{
    if (!base || nx <= 1 || ny <= 1 || nz <= 1 || ncomp <= 0 || !values) {
        amie::set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    const size_t n = (size_t)nx * ny * nz;
    std::vector<double> sorted(n);
    std::vector<unsigned char> bytes(n);
    const double lo = 1., hi = 255.;
    for (int c = 0; c < ncomp; c++) {
        const double *val = values + (size_t)c * n;
        sorted.assign(val, val + n);
        std::sort(sorted.begin(), sorted.end());
        double vmax = sorted[std::min(n * 1 - 1, (size_t)(n * .99))];
        double vmin = sorted[(size_t)(n * .01)];
        if (std::abs(vmax) > std::abs(vmin)) vmin = -std::abs(vmax);
        else vmax = std::abs(vmin);
        for (size_t i = 0; i < n; i++)
            bytes[i] = (unsigned char)std::min(std::max(std::round(lo + (hi - lo) * ((val[i] - vmin) / (vmax - vmin))), lo), hi);
        std::ostringstream name;
        name << base << "_" << nx << "_" << c << ".vox";
        std::ofstream out(name.str(), std::ios::out | std::ios::binary);
        if (!out) { amie::set_error("cannot open %s", name.str().c_str()); return AMIE_ERR_ARG; }
        out << "VOXELS" << std::endl << 1 << std::endl << nx << std::endl << ny << std::endl << nz << std::endl;
        out.write(reinterpret_cast<const char *>(bytes.data()), (std::streamsize)n);
        out.close();
        if (!out) { amie::set_error("write to %s failed", name.str().c_str()); return AMIE_ERR_ARG; }
    }
    return AMIE_OK;
}

