// lgame_mlp.h -- device-side entry points of the radiation-surrogate MLP used by the step.
#pragma once

#include <hip/hip_runtime.h>

struct lgame_mlp;

namespace lgame {

// Inference buffers of one user of a network (a domain's radiation update): the normalised
// input, the activations of every hidden layer and the outputs of up to `rows` rows per pass.
struct MlpWork;
int mlp_work_create(const lgame_mlp *m, long long rows, MlpWork **out);
void mlp_work_destroy(MlpWork *w);
// x (ncol x n_in, FP32, device) -> y (ncol x n_out, FP32, device), enqueued on `stream`, in
// passes through the workspace; allocates nothing.
int mlp_infer_device(const lgame_mlp *m, MlpWork *w, const float *x, long long ncol, float *y, hipStream_t stream);
// whether the caller set the input and output normalisation (lgame_mlp_set_norm)
bool mlp_has_norm(const lgame_mlp *m);
int mlp_dims(const lgame_mlp *m, int *n_in, int *n_out);
// device arrays of the output normalisation: y_norm = (y - mean) / scale, scale the training
// targets' standard deviation
void mlp_out_norm(const lgame_mlp *m, const float **mean, const float **scale);

}  // namespace lgame
