// Hooks through which the dynamical core runs a surrogate network on its own
// stream (defined in mlp.hip).
#pragma once
#include "../../include/thor_mlp.h"

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_layer.h"

namespace thor {
namespace mlp {
// Makes room for rows input rows; input() is then the BF16 [rows][n_in]
// buffer the caller fills, output() the FP32 [rows][n_out] result of
// forward_on().
hipError_t prepare(thor_mlp *m, long rows);
uint16_t *input(thor_mlp *m);
const float *output(thor_mlp *m);
void forward_on(thor_mlp *m, long rows, hipStream_t s);
void dims(const thor_mlp *m, int *nin, int *nout);
// A dynamical core's stream that runs the network: the network's own work
// that rewrites weights or buffers waits for it (attach), and order_on makes
// s wait for the network's queued work (training, new parameters).
void attach(thor_mlp *m, hipStream_t s);
void detach(thor_mlp *m, hipStream_t s);
hipError_t order_on(thor_mlp *m, hipStream_t s);
// Weights of the fused inference kernel (false if the network's shape has none).
bool fused_weights(const thor_mlp *m, InferWeights *w);
} // namespace mlp
} // namespace thor
