// Internal interface of the reduction kernels (nbx_blas.hip).
#pragma once
#include "nbx_reduce.h"

namespace nbx {

// Device scratch of a deterministic reduction: one partial per tile and value
// and a small result slot.
struct ReduceWork {
  double* partials = nullptr;
  double* result = nullptr;
  unsigned int capacity = 0;  // tiles
  int init(int64_t maxN, int nvals);
  void release();
};

// out_dev[0] = <x,y>_w, and out_dev[1] = <x,x>_w when xx (device pointers)
int launch_wdot(int32_t N, int32_t offset, int Nf, const double* w, const double* x, const double* y, bool xx,
                ReduceWork& ws, double* out_dev, hipStream_t s);

// q -= (unweighted sum of the E-vector over Nf fields) / count, on the whole
// Nf*offset range (BPZeroMean, nekBone/solve.cpp:115-134); device only
int launch_zero_mean(int32_t N, int32_t offset, int Nf, double count, double* q, ReduceWork& ws, hipStream_t s);

}  // namespace nbx
