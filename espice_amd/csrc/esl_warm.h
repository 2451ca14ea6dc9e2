// The warm-start network as the lane engine uses it (mlp_host.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/espice_lanes.h"

struct EslMlp;

namespace eslw {
// x [rows][n_out + 1] on the device from live values [rows][n_in] on the
// device, on stream s.
esl_status predict_device(EslMlp &n, const double *d_inputs, uint32_t rows, double *d_x, hipStream_t s);
EslMlp *lookup_mlp(esl_mlp *h);
uint32_t mlp_n_in(const EslMlp &n);
uint32_t mlp_n_out(const EslMlp &n);
}  // namespace eslw
