// Host-side layout builder: sparsity pattern in the reference order, the
// internal SELL-64 block layout, element -> block maps and the element
// colouring used by the coloured scatter.
//
// Pattern: the reference inserts every (ids[j], ids[k]) node pair of every
// element into a std::set and fills missing diagonals
// (solvers/assembly.cpp:448-524, 832-909), giving row_size and column_index
// sorted by (row, column).  Here the same set is produced row by row from the
// node -> element adjacency (sorted, unique neighbour lists), which is
// identical by construction and parallel over rows.
#include "layout.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace amie {

void build_pattern(int nper, int64_t nelem, const int32_t *edof, int64_t nrows, std::vector<uint32_t> &row_size,
                   std::vector<uint32_t> &column_index)
{
    // node -> elements adjacency
    std::vector<int64_t> cnt(nrows + 1, 0);
    for (int64_t e = 0; e < nelem; e++)
        for (int i = 0; i < nper; i++) cnt[edof[e * nper + i] + 1]++;
    for (int64_t r = 0; r < nrows; r++) cnt[r + 1] += cnt[r];
    std::vector<int64_t> adj(cnt[nrows]);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < nelem; e++)
        for (int i = 0; i < nper; i++) adj[fill[edof[e * nper + i]]++] = e;
    row_size.assign(nrows, 0);
    std::vector<std::vector<uint32_t>> rows(nrows);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < nrows; r++) {
        std::vector<uint32_t> &v = rows[r];
        v.push_back((uint32_t)r); // diagonal always present (gap filling)
        for (int64_t q = cnt[r]; q < cnt[r + 1]; q++) {
            const int64_t e = adj[q];
            for (int i = 0; i < nper; i++) v.push_back((uint32_t)edof[e * nper + i]);
        }
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        row_size[r] = (uint32_t)v.size();
    }
    size_t nnzb = 0;
    for (int64_t r = 0; r < nrows; r++) nnzb += row_size[r];
    column_index.resize(nnzb);
    size_t off = 0;
    for (int64_t r = 0; r < nrows; r++) {
        std::memcpy(column_index.data() + off, rows[r].data(), rows[r].size() * sizeof(uint32_t));
        off += rows[r].size();
    }
}

void choose_permutation(const std::vector<uint32_t> &row_size, const std::vector<int64_t> &row_start,
                        const std::vector<uint32_t> &column_index, std::vector<int32_t> &perm)
{
    // Candidate order, per window of 64 slices: rows by decreasing length,
    // and within a length, when the window holds at most 32 stencils (the
    // column - row offsets of a row; structured meshes), the rows of one
    // stencil together in order of first appearance; otherwise in their
    // own order.  Consecutive rows of a stencil group have their neighbours
    // at fixed offsets, which the SpMV's column offsets turn into gathers
    // without column loads.  The candidate is taken when it saves more than
    // 5 % of the padded slots (P2 meshes, whose vertex and mid-edge rows
    // alternate: 54 -> 40 us per SpMV on P2 400^2 with 64-slice windows,
    // against 47 us with 16-slice ones).  Meshes whose rows all have one
    // length keep their order: grouping their boundary and periodic rows
    // apart measured no faster (P1 RVE batch) or slower (Tet4 64^3).
    const int64_t n = (int64_t)row_size.size();
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    auto padded = [&](const std::vector<int32_t> &pm) {
        int64_t tot = 0;
        for (int64_t s = 0; s < n; s += 64) {
            uint32_t w = 0;
            for (int64_t i = s; i < std::min(n, s + 64); i++) w = std::max(w, row_size[pm[i]]);
            tot += (int64_t)w * 64;
        }
        return tot;
    };
    const int64_t base = padded(perm);
    // stencil signature: FNV-1a over the row's column - row offsets
    std::vector<uint64_t> sig(n);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) {
        uint64_t h = 1469598103934665603ull;
        for (int64_t q = row_start[r]; q < row_start[r + 1]; q++) {
            const uint64_t d = (uint64_t)((int64_t)column_index[q] - r);
            for (int byte = 0; byte < 8; byte++) h = (h ^ ((d >> (8 * byte)) & 0xFF)) * 1099511628211ull;
        }
        sig[r] = h;
    }
    std::vector<int32_t> cand(perm);
    const int64_t win = 64 * 64;
    for (int64_t s = 0; s < n; s += win) {
        const int64_t e = std::min(n, s + win);
        std::vector<uint64_t> seen;
        std::vector<int32_t> group(e - s, 0);
        for (int64_t r = s; r < e && seen.size() <= 32; r++) {
            auto it = std::find(seen.begin(), seen.end(), sig[r]);
            if (it == seen.end()) {
                seen.push_back(sig[r]);
                it = seen.end() - 1;
            }
            group[r - s] = (int32_t)(it - seen.begin());
        }
        const bool by_stencil = seen.size() <= 32;
        std::stable_sort(cand.begin() + s, cand.begin() + e, [&](int32_t x, int32_t y) {
            if (row_size[x] != row_size[y]) return row_size[x] > row_size[y];
            return by_stencil && group[x - s] < group[y - s];
        });
    }
    if (padded(cand) < base * 0.95) perm.swap(cand);
}

void colour_elements(int nper, int64_t nelem, const int32_t *edof, int64_t nrows, std::vector<int32_t> &colour,
                     int &ncolours)
{
    // Greedy colouring in element order: an element takes the lowest colour
    // that none of the elements sharing one of its DOF nodes already holds.
    std::vector<uint64_t> used(nrows * 2, 0);
    colour.resize(nelem);
    ncolours = 0;
    for (int64_t e = 0; e < nelem; e++) {
        uint64_t m0 = 0, m1 = 0;
        for (int i = 0; i < nper; i++) {
            const int64_t r = edof[e * nper + i];
            m0 |= used[2 * r];
            m1 |= used[2 * r + 1];
        }
        int c;
        if (~m0) c = __builtin_ctzll(~m0);
        else if (~m1) c = 64 + __builtin_ctzll(~m1);
        else c = -1;
        colour[e] = c;
        if (c < 0) { ncolours = -1; return; }
        for (int i = 0; i < nper; i++) {
            const int64_t r = edof[e * nper + i];
            if (c < 64) used[2 * r] |= (1ull << c);
            else used[2 * r + 1] |= (1ull << (c - 64));
        }
        ncolours = std::max(ncolours, c + 1);
    }
}

int64_t find_block(const std::vector<uint32_t> &column_index, const std::vector<int64_t> &row_start, int64_t r,
                   uint32_t c)
{
    const uint32_t *b = column_index.data() + row_start[r], *e = column_index.data() + row_start[r + 1];
    const uint32_t *p = std::lower_bound(b, e, c);
    if (p == e || *p != c) return -1;
    return p - column_index.data();
}

} // namespace amie
