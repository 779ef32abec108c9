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

namespace {

// normalizeArray (utilities/writer/voxel_writer.cpp:548-571, min 1, max
// 255, no voids): the values at the 99 % and 1 % positions of the sorted
// array, widened to the symmetric range of the larger magnitude, map the
// values to [1, 255], rounded
void normalize(const double *val, size_t n, size_t stride, std::vector<double> &sorted, unsigned char *out)
{
    sorted.resize(n);
    for (size_t i = 0; i < n; i++) sorted[i] = val[i * stride];
    std::sort(sorted.begin(), sorted.end());
    double vmax = sorted[std::min(n * 1 - 1, (size_t)(n * .99))];
    double vmin = sorted[(size_t)(n * .01)];
    if (std::abs(vmax) > std::abs(vmin)) vmin = -std::abs(vmax);
    else vmax = std::abs(vmin);
    const double lo = 1., hi = 255.;
    for (size_t i = 0; i < n; i++)
        out[i] = (unsigned char)std::min(std::max(std::round(lo + (hi - lo) * ((val[i * stride] - vmin) / (vmax - vmin))), lo), hi);
}

// the columns of a TriangleWriter line: coordinates (6), displacements (6),
// then each field component over vertices 0, 1, 2
struct Columns {
    int64_t ntri;
    int32_t nper;
    const int32_t *conn;
    const double *xy, *uv;
    int32_t nfield;
    const int32_t *sizes;
    const double *const *fields;
    int64_t count() const
    {
        int64_t c = 12;
        for (int f = 0; f < nfield; f++) c += 3 * sizes[f];
        return c;
    }
    // column `col` of triangle t
    double at(int64_t t, int64_t col) const
    {
        const int n = nper / 3;
        if (col < 6) return xy[(int64_t)conn[t * nper + (col / 2) * n] * 2 + col % 2];
        if (col < 12) return uv[t * 6 + ((col - 6) % 3) * 2 + (col - 6) / 3];
        col -= 12;
        for (int f = 0; f < nfield; f++) {
            const int s = sizes[f];
            if (col < 3 * s) return fields[f][t * 3 * s + (col % 3) * s + col / 3];
            col -= 3 * s;
        }
        return 0.;
    }
};

int check_args(const char *path, int64_t ntri, int32_t nper, const int32_t *conn, int64_t nnodes, const double *xy,
               const double *uv, int32_t nfield, const int32_t *sizes, const double *const *fields)
{
    if (!path || ntri <= 0 || nper <= 0 || nper % 3 || !conn || nnodes <= 0 || !xy || !uv || nfield < 0 ||
        (nfield && (!sizes || !fields))) {
        amie::set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    for (int64_t q = 0; q < ntri * nper; q++)
        if (conn[q] < 0 || conn[q] >= nnodes) { amie::set_error("connectivity names a node outside [0, nnodes)"); return AMIE_ERR_ARG; }
    for (int f = 0; f < nfield; f++)
        if (sizes[f] <= 0 || !fields[f]) { amie::set_error("field %d: bad size or data", f); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

} // namespace

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
    int rc = check_args(path, ntri, nper, conn, nnodes, xy, uv, nfield, sizes, fields);
    if (rc) return rc;
    const Columns C{ntri, nper, conn, xy, uv, nfield, sizes, fields};
    const int64_t cols = C.count();
    std::ofstream out(path, std::ios::out);
    if (!out) { amie::set_error("cannot open %s", path); return AMIE_ERR_ARG; }
    out << "TRIANGLES" << std::endl;
    out << (int)ntri << std::endl;
    out << 3 << std::endl;
    out << (int)((cols - 6) / 3) << std::endl;
    for (int64_t t = 0; t < ntri; t++) {
        for (int64_t c = 0; c < cols; c++) out << C.at(t, c) << " ";
        out << std::endl;
    }
    out.close();
    if (!out) { amie::set_error("write to %s failed", path); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

// MultiTriangleWriter (triangle_writer.cpp:73-86, 538-573, 1493-1500): the
// index file starts with "MULTI"; every append writes <base>_<step> in the
// TriangleWriter format (step counting 1, 2, ... from the first append) and
// adds that name to the index file.
int amie_multi_triangles_begin(const char *head)
{
    if (!head) { amie::set_error("bad arguments"); return AMIE_ERR_ARG; }
    std::ofstream out(head, std::ios::out);
    if (!out) { amie::set_error("cannot open %s", head); return AMIE_ERR_ARG; }
    out << "MULTI" << std::endl;
    out.close();
    if (!out) { amie::set_error("write to %s failed", head); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

int amie_multi_triangles_append(const char *head, const char *base, int32_t step, int64_t ntri, int32_t nper,
                                const int32_t *conn, int64_t nnodes, const double *xy, const double *uv,
                                int32_t nfield, const int32_t *sizes, const double *const *fields)
{
    if (!head || !base || step < 1) { amie::set_error("bad arguments"); return AMIE_ERR_ARG; }
    const std::string name = std::string(base) + "_" + std::to_string(step);
    int rc = amie_write_triangles(name.c_str(), ntri, nper, conn, nnodes, xy, uv, nfield, sizes, fields);
    if (rc) return rc;
    std::ofstream out(head, std::ios::out | std::ios::app);
    if (!out) { amie::set_error("cannot open %s", head); return AMIE_ERR_ARG; }
    out << name << std::endl;
    out.close();
    if (!out) { amie::set_error("write to %s failed", head); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

// BinaryTriangleWriter (triangle_writer.cpp:447-478, 1477-1491): the header
// "BIN_TRIANGLES", the column count, 3, and the reference's (layers - 6) / 3
// for one layer (-1); per triangle the six coordinates as doubles, then
// every later column normalised over the triangles to 1..255 (as the voxel
// files) and written as an integer, each value followed by a space.
int amie_write_triangles_bin(const char *path, int64_t ntri, int32_t nper, const int32_t *conn, int64_t nnodes,
                             const double *xy, const double *uv, int32_t nfield, const int32_t *sizes,
                             const double *const *fields)
{
    int rc = check_args(path, ntri, nper, conn, nnodes, xy, uv, nfield, sizes, fields);
    if (rc) return rc;
    const Columns C{ntri, nper, conn, xy, uv, nfield, sizes, fields};
    const int64_t cols = C.count();
    std::vector<double> col((size_t)ntri), sorted;
    std::vector<unsigned char> norm((size_t)(cols - 6) * ntri);
    for (int64_t c = 6; c < cols; c++) {
        for (int64_t t = 0; t < ntri; t++) col[t] = C.at(t, c);
        normalize(col.data(), (size_t)ntri, 1, sorted, norm.data() + (c - 6) * ntri);
    }
    std::ofstream out(path, std::ios::out);
    if (!out) { amie::set_error("cannot open %s", path); return AMIE_ERR_ARG; }
    out << "BIN_TRIANGLES" << std::endl;
    out << (int)cols << std::endl;
    out << 3 << std::endl;
    out << ((int)1 - 6) / 3 << std::endl;
    for (int64_t t = 0; t < ntri; t++) {
        for (int64_t c = 0; c < 6; c++) out << C.at(t, c) << " ";
        for (int64_t c = 6; c < cols; c++) out << (unsigned short)norm[(c - 6) * ntri + t] << " ";
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
    std::vector<double> sorted;
    std::vector<unsigned char> bytes(n);
    for (int c = 0; c < ncomp; c++) {
        normalize(values + (size_t)c * n, n, 1, sorted, bytes.data());
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

