// Internal interface of the matrix-statistics kernel (kernels_features.hip).
#pragma once

#include "cghip_kernels.h"

namespace cghip {

constexpr int kStatQuantities = 5;  // log d, Gershgorin radius, log strongest, log weakest, anisotropy
constexpr int kStatHistBins = 16;   // row lengths 1..15 and >= 16
// per block: (count, sum, sum of squares, min, max) per quantity, the
// histogram, the number of rows without a positive diagonal
constexpr int kStatWidth = 5 * kStatQuantities + kStatHistBins + 1;

int matrix_stats_grid(int n);
hipError_t matrix_stats(const Crs<double> &A, double *partials, hipStream_t s);

}  // namespace cghip
