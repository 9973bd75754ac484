// (R, G) pairs that get a compiled fit kernel for every model: the default
// group size for every R in [3, 32] plus tuning alternatives for the ROI sizes
// of the benchmark workloads (7, 9, 10, 11, 13, 16, 20, 25).  The list is
// split in four parts so the build compiles them in parallel.
#pragma once

#define PPA_FIT_SIZES_0(X) X(3, 1) X(4, 1) X(5, 2) X(6, 2) X(7, 1) X(7, 2) X(7, 4) X(8, 2) X(9, 2) X(9, 4) X(10, 2)
#define PPA_FIT_SIZES_1(X) X(10, 4) X(10, 8) X(10, 16) X(11, 2) X(11, 4) X(12, 4) X(13, 4) X(13, 8) X(14, 4) X(15, 4) X(16, 4)
#define PPA_FIT_SIZES_2(X) X(16, 8) X(17, 8) X(18, 8) X(19, 8) X(20, 8) X(20, 16) X(21, 8) X(22, 8) X(23, 8) X(24, 8)
#define PPA_FIT_SIZES_3(X) X(25, 16) X(25, 8) X(26, 16) X(27, 16) X(28, 16) X(29, 16) X(30, 16) X(31, 16) X(32, 16)

// Grid-form spline: default_group_grid(R) for every R plus alternatives at the
// benchmark workloads' sizes.
#define PPA_GRID_SIZES_0(X) X(3, 4) X(4, 4) X(5, 4) X(6, 4) X(7, 8) X(7, 4) X(8, 8) X(9, 8) X(9, 16) X(10, 8)
#define PPA_GRID_SIZES_1(X) X(11, 8) X(11, 16) X(12, 8) X(13, 16) X(13, 8) X(14, 16) X(15, 16) X(16, 16) X(16, 32)
#define PPA_GRID_SIZES_2(X) X(17, 16) X(18, 32) X(19, 32) X(20, 16) X(20, 32) X(20, 64) X(21, 32) X(22, 32) X(23, 64)
#define PPA_GRID_SIZES_3(X) X(24, 64) X(25, 16) X(25, 32) X(25, 64) X(26, 64) X(27, 64) X(28, 64) X(29, 64) X(30, 64) X(31, 64) X(32, 64)

// The reference's other Gaussian variants and the intensity/background model:
// default_group(R) for every R.
#define PPA_VAR_SIZES_0(X) X(3, 1) X(4, 1) X(5, 2) X(6, 2) X(7, 2) X(8, 2) X(9, 2) X(10, 2)
#define PPA_VAR_SIZES_1(X) X(11, 2) X(12, 4) X(13, 4) X(14, 4) X(15, 4) X(16, 4)
#define PPA_VAR_SIZES_2(X) X(17, 8) X(18, 8) X(19, 8) X(20, 8) X(21, 8) X(22, 8) X(23, 8) X(24, 8)
#define PPA_VAR_SIZES_3(X) X(25, 16) X(26, 16) X(27, 16) X(28, 16) X(29, 16) X(30, 16) X(31, 16) X(32, 16)

// Grid-form spline variants: default_group_grid(R) for every R.
#define PPA_GRID_VAR_SIZES_0(X) X(3, 4) X(4, 4) X(5, 4) X(6, 4) X(7, 8) X(8, 8) X(9, 8) X(10, 8)
#define PPA_GRID_VAR_SIZES_1(X) X(11, 8) X(12, 8) X(13, 16) X(14, 16) X(15, 16) X(16, 16)
#define PPA_GRID_VAR_SIZES_2(X) X(17, 16) X(18, 32) X(19, 32) X(20, 32) X(21, 32) X(22, 32) X(23, 64) X(24, 64)
#define PPA_GRID_VAR_SIZES_3(X) X(25, 64) X(26, 64) X(27, 64) X(28, 64) X(29, 64) X(30, 64) X(31, 64) X(32, 64)

#define PPA_FIT_CASE(R_, G_) \
    if (R == R_ && G == G_) return launch_fit_t<Model, R_, G_>(m, a, events, st);
