#pragma once
#include <cstdint>
#include <vector>

namespace amie {

// Reference-order node-block pattern (row_size, sorted column_index) of the
// elements' DOF nodes edof[nelem*nper] over nrows block rows.
void build_pattern(int nper, int64_t nelem, const int32_t *edof, int64_t nrows, std::vector<uint32_t> &row_size,
                   std::vector<uint32_t> &column_index);
// perm[internal] = original row (row_start: prefix sums of row_size)
void choose_permutation(const std::vector<uint32_t> &row_size, const std::vector<int64_t> &row_start,
                        const std::vector<uint32_t> &column_index, std::vector<int32_t> &perm);
// colour[e]; ncolours = -1 when more than 128 colours would be needed
void colour_elements(int nper, int64_t nelem, const int32_t *edof, int64_t nrows, std::vector<int32_t> &colour,
                     int &ncolours);
int64_t find_block(const std::vector<uint32_t> &column_index, const std::vector<int64_t> &row_start, int64_t r,
                   uint32_t c);

} // namespace amie
