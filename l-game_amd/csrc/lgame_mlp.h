// lgame_mlp.h -- device-side entry points of the radiation-surrogate MLP used by the step.
#pragma once

#include <hip/hip_runtime.h>

struct lgame_mlp;

namespace lgame {

// x (ncol x n_in, FP32, device) -> y (ncol x n_out, FP32, device), enqueued on `stream`.
int mlp_infer_device(lgame_mlp *m, const float *x, long long ncol, float *y, hipStream_t stream);
int mlp_dims(const lgame_mlp *m, int *n_in, int *n_out);
// device arrays of the output normalisation: y_norm = (y - mean) / scale, scale the training
// targets' standard deviation
void mlp_out_norm(const lgame_mlp *m, const float **mean, const float **scale);

}  // namespace lgame
