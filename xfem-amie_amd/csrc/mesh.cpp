// Host-side mesh helpers: orientation rule, structured meshes of the
// benchmark configurations, periodic DOF map, inclusion placement, phases,
// and the initial-damage random field.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "amie_gpu.h"

namespace amie {
void set_error(const char *fmt, ...);
}

namespace {

void orient_one(int et, const double *coords, int32_t *c)
{
    if (et == AMIE_P1 || et == AMIE_P2) {
        // geometry/geometry_2D.cpp:476-497 with ConvexPolygon::isTrigoOriented
        // (geometry/geometry_base.cpp:3292-3307): swap vertices 1 and 2 of a
        // clockwise triangle (mid-nodes follow their edges).
        const int v0 = 0, v1 = et == AMIE_P2 ? 2 : 1, v2 = et == AMIE_P2 ? 4 : 2;
        const int v[3] = {v0, v1, v2};
        bool ccw = true;
        for (int i = 0; i < 2; i++) {
            const double *p0 = coords + 2 * (size_t)c[v[i]], *p1 = coords + 2 * (size_t)c[v[(i + 1) % 3]],
                         *p2 = coords + 2 * (size_t)c[v[(i + 2) % 3]];
            double ax = p1[0] - p0[0], ay = p1[1] - p0[1], bx = p2[0] - p1[0], by = p2[1] - p1[1];
            if (ax * by - ay * bx < 0) ccw = false;
        }
        if (!ccw) {
            if (et == AMIE_P1) std::swap(c[1], c[2]);
            else { std::swap(c[2], c[4]); std::swap(c[1], c[5]); }
        }
    } else {
        // geometry/geometry_3D.cpp:254-268 with Tetrahedron::computeVolume (:571-576)
        const double *p0 = coords + 3 * (size_t)c[0], *p1 = coords + 3 * (size_t)c[1], *p2 = coords + 3 * (size_t)c[2],
                     *p3 = coords + 3 * (size_t)c[3];
        double u[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
        double v[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        double w[3] = {p1[0] - p3[0], p1[1] - p3[1], p1[2] - p3[2]};
        double x[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        if ((x[0] * w[0] + x[1] * w[1] + x[2] * w[2]) / 6. < 0) std::swap(c[0], c[1]);
    }
}

struct CircleGrid {
    double cell;
    int nc;
    std::vector<std::vector<int64_t>> cells;
    const double *xyr;
    CircleGrid(double cell_, const double *xyr_, int64_t n) : cell(cell_), xyr(xyr_)
    {
        nc = std::max(1, (int)std::ceil(1. / cell));
        cells.resize((size_t)nc * nc);
        for (int64_t i = 0; i < n; i++) add(i);
    }
    int idx(double v) const { return std::min(nc - 1, std::max(0, (int)std::floor(v / cell))); }
    void add(int64_t i)
    {
        const double x = xyr[3 * i], y = xyr[3 * i + 1], r = xyr[3 * i + 2];
        for (int a = idx(x - r); a <= idx(x + r); a++)
            for (int b = idx(y - r); b <= idx(y + r); b++) cells[(size_t)a * nc + b].push_back(i);
    }
    const std::vector<int64_t> &at(double x, double y) const { return cells[(size_t)idx(x) * nc + idx(y)]; }
};

} // namespace

extern "C" {

void amie_cauchy_green(double E, double nu, int32_t dim, double *C)
{
    // utilities/tensor.cpp:1845-1931, YOUNG_POISSON; 2D = plane stress.
    // The shear entry is E/(1+nu) = 2G, used with engineering shear strain.
    double Ciiii, Ciijj, Cijij;
    if (dim == 3) {
        double c = E / ((1. + nu) * (1. - 2. * nu));
        Ciiii = c * (1. - nu); Ciijj = c * nu; Cijij = c * (1. - 2. * nu);
    } else {
        double c = E / (1. - nu * nu);
        Ciiii = c; Ciijj = c * nu; Cijij = c * (1. - nu);
    }
    const int s = dim == 3 ? 6 : 3;
    for (int i = 0; i < s * s; i++) C[i] = 0.;
    if (dim == 3) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) C[i * 6 + j] = i == j ? Ciiii : Ciijj;
        C[21] = C[28] = C[35] = Cijij;
    } else {
        C[0] = C[4] = Ciiii; C[1] = C[3] = Ciijj; C[8] = Cijij;
    }
}

int amie_mesh_orient(int32_t et, int64_t nelem, int32_t *conn, const double *coords)
{
    if (!conn || !coords || (et < AMIE_P1 || et > AMIE_TET4)) { amie::set_error("bad arguments"); return AMIE_ERR_ARG; }
    const int nper = et == AMIE_P1 ? 3 : (et == AMIE_P2 ? 6 : 4);
    for (int64_t e = 0; e < nelem; e++) orient_one(et, coords, conn + e * nper);
    return AMIE_OK;
}

int amie_mesh_structured_size(int32_t et, int32_t n, int64_t *nnodes, int64_t *nelem)
{
    if (n <= 0) { amie::set_error("n must be positive"); return AMIE_ERR_ARG; }
    if (et == AMIE_P1) { *nnodes = (int64_t)(n + 1) * (n + 1); *nelem = 2LL * n * n; }
    else if (et == AMIE_P2) { *nnodes = (int64_t)(2 * n + 1) * (2 * n + 1); *nelem = 2LL * n * n; }
    else if (et == AMIE_TET4) { *nnodes = (int64_t)(n + 1) * (n + 1) * (n + 1); *nelem = 6LL * n * n * n; }
    else { amie::set_error("bad element type"); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

int amie_mesh_structured(int32_t et, int32_t n, double *coords, int32_t *conn)
{
    int64_t nn, ne;
    int rc = amie_mesh_structured_size(et, n, &nn, &ne);
    if (rc) return rc;
    if (et == AMIE_P1) {
        // node (i,j) -> i*(n+1)+j at (i/n, j/n); two triangles per cell as
        // mesher/structuredmesh.cpp:26-31 (without its random jitter)
        const int m = n + 1;
        for (int i = 0; i <= n; i++)
            for (int j = 0; j <= n; j++) {
                coords[2 * ((size_t)i * m + j)] = (double)i / n;
                coords[2 * ((size_t)i * m + j) + 1] = (double)j / n;
            }
        int64_t e = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int32_t a = i * m + j, b = i * m + j + 1, c = (i + 1) * m + j + 1, d = (i + 1) * m + j;
                int32_t t0[3] = {a, b, c}, t1[3] = {a, d, c};
                std::memcpy(conn + 3 * e++, t0, 12);
                std::memcpy(conn + 3 * e++, t1, 12);
            }
    } else if (et == AMIE_P2) {
        // (2n+1)^2 grid of vertex and mid-edge nodes, node (I,J) -> I*(2n+1)+J
        const int m = 2 * n + 1;
        for (int I = 0; I < m; I++)
            for (int J = 0; J < m; J++) {
                coords[2 * ((size_t)I * m + J)] = (double)I / (2 * n);
                coords[2 * ((size_t)I * m + J) + 1] = (double)J / (2 * n);
            }
        auto id = [&](int I, int J) { return (int32_t)(I * m + J); };
        int64_t e = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const int I = 2 * i, J = 2 * j;
                // vertex triangles as P1: (a,b,c) and (a,d,c)
                const int tv[2][3][2] = {{{I, J}, {I, J + 2}, {I + 2, J + 2}}, {{I, J}, {I + 2, J}, {I + 2, J + 2}}};
                for (int h = 0; h < 2; h++) {
                    const int(*v)[2] = tv[h];
                    int32_t c6[6] = {id(v[0][0], v[0][1]), id((v[0][0] + v[1][0]) / 2, (v[0][1] + v[1][1]) / 2),
                                     id(v[1][0], v[1][1]), id((v[1][0] + v[2][0]) / 2, (v[1][1] + v[2][1]) / 2),
                                     id(v[2][0], v[2][1]), id((v[2][0] + v[0][0]) / 2, (v[2][1] + v[0][1]) / 2)};
                    std::memcpy(conn + 6 * e++, c6, 24);
                }
            }
    } else {
        // voxel 6-tet split of filters/voxelfilter.cpp:352-405, node
        // (i,j,k) -> i*(n+1)^2 + j*(n+1) + k at (i/n, j/n, k/n)
        const int m = n + 1;
        for (int i = 0; i <= n; i++)
            for (int j = 0; j <= n; j++)
                for (int k = 0; k <= n; k++) {
                    const size_t q = ((size_t)i * m + j) * m + k;
                    coords[3 * q] = (double)i / n;
                    coords[3 * q + 1] = (double)j / n;
                    coords[3 * q + 2] = (double)k / n;
                }
        auto id = [&](int i, int j, int k) { return (int32_t)(((int64_t)i * m + j) * m + k); };
        static const int split[6][4] = {{0, 1, 2, 4}, {3, 5, 7, 6}, {2, 3, 4, 6}, {1, 2, 3, 4}, {1, 3, 4, 5}, {3, 4, 6, 5}};
        int64_t e = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                for (int k = 0; k < n; k++) {
                    const int32_t c[8] = {id(i, j + 1, k), id(i + 1, j + 1, k), id(i, j + 1, k + 1), id(i + 1, j + 1, k + 1),
                                          id(i, j, k),     id(i + 1, j, k),     id(i, j, k + 1),     id(i + 1, j, k + 1)};
                    for (int s = 0; s < 6; s++) {
                        for (int q = 0; q < 4; q++) conn[4 * e + q] = c[split[s][q]];
                        e++;
                    }
                }
    }
    return amie_mesh_orient(et, ne, conn, coords);
}

int amie_mesh_periodic_dofmap(int32_t n, int32_t *dofmap)
{
    if (n <= 0 || !dofmap) { amie::set_error("bad arguments"); return AMIE_ERR_ARG; }
    for (int i = 0; i <= n; i++)
        for (int j = 0; j <= n; j++) dofmap[i * (n + 1) + j] = (i % n) * n + (j % n);
    return AMIE_OK;
}

int amie_inclusions_2d(uint64_t seed, double target, double rmin, double rmax, double gap, int64_t max_circles,
                       double *xyr, int64_t *ncircles, double *fraction)
{
    if (!xyr || !ncircles || rmin <= 0 || rmax < rmin || max_circles <= 0) { amie::set_error("bad arguments"); return AMIE_ERR_ARG; }
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0., 1.);
    const double cell = 2. * rmax + gap;
    const int nc = std::max(1, (int)std::ceil(1. / cell));
    std::vector<std::vector<int64_t>> grid((size_t)nc * nc);
    auto cid = [&](double v) { return std::min(nc - 1, std::max(0, (int)std::floor(v / cell))); };
    int64_t n = 0;
    double area = 0.;
    int64_t fails = 0;
    const int64_t max_fails = 200000;
    while (area < target && n < max_circles && fails < max_fails) {
        const double r = rmin + (rmax - rmin) * U(rng);
        const double x = r + (1. - 2. * r) * U(rng), y = r + (1. - 2. * r) * U(rng);
        bool ok = true;
        for (int a = std::max(0, cid(x) - 1); ok && a <= std::min(nc - 1, cid(x) + 1); a++)
            for (int b = std::max(0, cid(y) - 1); ok && b <= std::min(nc - 1, cid(y) + 1); b++)
                for (int64_t q : grid[(size_t)a * nc + b]) {
                    const double dx = x - xyr[3 * q], dy = y - xyr[3 * q + 1];
                    const double md = r + xyr[3 * q + 2] + gap;
                    if (dx * dx + dy * dy < md * md) { ok = false; break; }
                }
        if (!ok) { fails++; continue; }
        fails = 0;
        xyr[3 * n] = x; xyr[3 * n + 1] = y; xyr[3 * n + 2] = r;
        grid[(size_t)cid(x) * nc + cid(y)].push_back(n);
        n++;
        area += M_PI * r * r;
    }
    *ncircles = n;
    if (fraction) *fraction = area;
    return AMIE_OK;
}

int amie_phase_by_centroid(int32_t et, int64_t nelem, const int32_t *conn, const double *coords, int64_t ncircles,
                           const double *xyr, int32_t *phase)
{
    if (!conn || !coords || !phase || (et != AMIE_P1 && et != AMIE_P2)) { amie::set_error("bad arguments"); return AMIE_ERR_ARG; }
    double rmax = 0.;
    for (int64_t i = 0; i < ncircles; i++) rmax = std::max(rmax, xyr[3 * i + 2]);
    CircleGrid G(std::max(0.01, 2. * rmax), xyr, ncircles);
    const int nper = et == AMIE_P1 ? 3 : 6;
    const int v[3] = {0, et == AMIE_P1 ? 1 : 2, et == AMIE_P1 ? 2 : 4};
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nelem; e++) {
        double cx = 0., cy = 0.;
        for (int i = 0; i < 3; i++) {
            cx += coords[2 * (size_t)conn[e * nper + v[i]]];
            cy += coords[2 * (size_t)conn[e * nper + v[i]] + 1];
        }
        cx /= 3.; cy /= 3.;
        int ph = 0;
        for (int64_t q : G.at(cx, cy)) {
            const double dx = cx - xyr[3 * q], dy = cy - xyr[3 * q + 1];
            if (dx * dx + dy * dy <= xyr[3 * q + 2] * xyr[3 * q + 2]) { ph = 1; break; }
        }
        phase[e] = ph;
    }
    return AMIE_OK;
}

int amie_damage_field_3d(uint64_t seed, int32_t n, double ell, double lo, double hi, double *d0)
{
    if (n <= 0 || !d0 || ell <= 0) { amie::set_error("bad arguments"); return AMIE_ERR_ARG; }
    const size_t nv = (size_t)n * n * n;
    std::vector<double> f(nv), g(nv);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> N(0., 1.);
    for (auto &v : f) v = N(rng);
    // separable Gaussian filter (std. deviation ell in voxel units of the
    // unit cube), truncated at 4 sigma, clamped at the boundary
    const double sig = ell * n;
    const int R = std::max(1, (int)std::ceil(4. * sig));
    std::vector<double> w(2 * R + 1);
    double ws = 0.;
    for (int k = -R; k <= R; k++) ws += (w[k + R] = std::exp(-0.5 * k * k / (sig * sig)));
    for (auto &v : w) v /= ws;
    auto idx = [&](int i, int j, int k) { return ((size_t)i * n + j) * n + k; };
    for (int axis = 0; axis < 3; axis++) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                for (int k = 0; k < n; k++) {
                    double s = 0.;
                    for (int t = -R; t <= R; t++) {
                        int a = i, b = j, c = k;
                        int &m = axis == 0 ? a : (axis == 1 ? b : c);
                        m = std::min(n - 1, std::max(0, m + t));
                        s += w[t + R] * f[idx(a, b, c)];
                    }
                    g[idx(i, j, k)] = s;
                }
        f.swap(g);
    }
    double mn = *std::min_element(f.begin(), f.end()), mx = *std::max_element(f.begin(), f.end());
    const double span = mx > mn ? mx - mn : 1.;
    for (size_t v = 0; v < nv; v++) {
        const double val = lo + (hi - lo) * (f[v] - mn) / span;
        for (int t = 0; t < 6; t++) d0[6 * v + t] = val;
    }
    return AMIE_OK;
}

// placement2D (utilities/placement.cpp:87-162) for circular inclusions in an
// axis-aligned rectangle, without pre-placed inclusions or exclusion zones.
// Restated step by step:
//  * the generator is std::default_random_engine(seed) with uniform
//    distributions over the rectangle's corner coordinates, x from the top
//    left to the bottom right corner's x, y from the top left's y to the
//    bottom right's (a decreasing range), and a rotation in
//    [-orientation, orientation] that circles ignore (geometry_base.cpp:
//    588-660); one try draws the rotation first, then y, then x, the order
//    in which the reference's g++ build evaluates the arguments of
//    transform2D(inc, x(), y(), r());
//  * with a gap, the radius is scaled by (r + gap)/r for the placement and
//    back by 1/scale once placed (Circle::setRadius, geometry_2D.cpp:1256);
//  * a candidate must lie in the rectangle with its bounding box
//    (Rectangle::in, geometry_2D.cpp:993-1008, 1e-12 tolerance), then neither
//    overlap a placed circle (Geometry::intersects, circle branch,
//    geometry_base.cpp:1430-1445) nor have its centre inside one
//    (Circle::in, geometry_2D.cpp:1472-1484); placed circles keep their
//    unscaled radii.  The reference finds the placed circles to test through
//    a pixel grid (utilities/grid.cpp:959-991, 1090-1105); every circle the
//    grid returns is tested the same way and a circle it does not return
//    cannot overlap, so testing all placed circles decides the same;
//  * one try budget for the whole list: an inclusion still searching when it
//    runs out is not placed and placement stops.
// Outputs the placed inclusions' indices, centres and final radii in order.
int amie_place_inclusions_2d(double width, double height, double cx, double cy, int64_t n, const double *radii,
                             double min_dist, int64_t tries_max, double orientation, uint64_t seed,
                             int64_t *placed_index, double *placed_xyr, int64_t *nplaced)
{
#pragma clang fp contract(off)
    if (n < 0 || (n && (!radii || !placed_index || !placed_xyr)) || !nplaced || width <= 0 || height <= 0) {
        amie::set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    for (int64_t i = 0; i < n; i++)
        if (!(radii[i] > 0)) { amie::set_error("radius %lld is not positive", (long long)i); return AMIE_ERR_ARG; }
    const double tol = 1e-12;
    const double left = cx - 0.5 * width, right = cx + 0.5 * width, top = cy + 0.5 * height, bottom = cy - 0.5 * height;
    auto box_in = [&](double x, double y) {
        if (x < cx - 0.5 * width - tol) return false;
        if (x > cx + 0.5 * width + tol) return false;
        if (y > cy + 0.5 * height + tol) return false;
        if (y < cy - 0.5 * height - tol) return false;
        return true;
    };
    std::default_random_engine gen(seed);
    std::uniform_real_distribution<double> xd(left, right), yd(top, bottom), rd(-orientation, orientation);
    struct Circ { double x, y, r, r2; };
    std::vector<Circ> placed;
    double x = 0., y = 0.;
    auto draw = [&]() {
        (void)rd(gen);
        y = yd(gen);
        x = xd(gen);
    };
    auto fits_box = [&](double r) {
        return box_in(x, y) && box_in(x + (-r), y + r) && box_in(x + r, y + r) && box_in(x + r, y + (-r)) &&
               box_in(x + (-r), y + (-r));
    };
    auto free_of_others = [&](double r) {
        for (const Circ &t : placed) {
            const double birad = r + t.r;
            bool inter = true;
            if (x + birad < t.x) inter = false;
            else if (x - birad > t.x) inter = false;
            else if (y + birad < t.y) inter = false;
            else if (y - birad > t.y) inter = false;
            if (inter) {
                const double dx = x - t.x, dy = y - t.y;
                if (dx * dx + dy * dy < birad * birad) return false;
            }
            if (!(x < t.x - t.r * 1.01) && !(x > t.x + t.r * 1.01) && !(y < t.y - t.r * 1.01) && !(y > t.y + t.r * 1.01)) {
                const double dx = x - t.x, dy = y - t.y;
                if (dx * dx + dy * dy < t.r2) return false;
            }
        }
        return true;
    };
    int64_t tries = 0, np = 0;
    for (int64_t i = 0; i < n && tries < tries_max; i++) {
        tries++;
        double r = radii[i], scale = 1.;
        if (min_dist > tol) {
            scale = (radii[i] + min_dist) / radii[i];
            r = r * scale;
        }
        draw();
        while (!fits_box(r) && tries < tries_max) {
            tries++;
            draw();
        }
        bool added;
        while (!(added = free_of_others(r)) && tries < tries_max) {
            tries++;
            draw();
            while (!fits_box(r) && tries < tries_max) {
                tries++;
                draw();
            }
        }
        if (added && tries < tries_max) {
            if (scale > 1.) r = r * (1. / scale);
            placed.push_back({x, y, r, r * r});
            placed_index[np] = i;
            placed_xyr[np * 3] = x;
            placed_xyr[np * 3 + 1] = y;
            placed_xyr[np * 3 + 2] = r;
            np++;
        } else if (added) {
            // added to the reference's grid on the last try, not returned
            if (scale > 1.) r = r * (1. / scale);
            placed.push_back({x, y, r, r * r});
        }
    }
    *nplaced = np;
    return AMIE_OK;
}

} // extern "C"
