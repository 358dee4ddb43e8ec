// blas1.hpp -- launch helpers shared by blas1.hip and the CG drivers.
#pragma once
#include "common.hpp"

namespace numamd {

int elementwise_grid(int64_t n);
int dot_grid(int64_t n);
// out[0] = sum x_i y_i (sqrt if do_sqrt); partial needs dot_grid(n) doubles.
int launch_dot(int64_t n, const double* x, const double* y, double* out, double* partial,
               int do_sqrt, hipStream_t stream);

// Resolve the device address of the static scratch of num_dot_noalloc on
// the current device (idempotent; called by num_init).
int noalloc_setup();

__global__ void sum_partials_kernel(const double* __restrict__ partial, int np, double* out,
                                    int do_sqrt);

}  // namespace numamd
