// hps_surrogate.hpp -- the field surrogate network as the engine uses it.
#pragma once
#include <hip/hip_runtime.h>

namespace hps {

struct NetImpl;
// forward pass of B samples [B][H][W][4] on stream s; the prediction is
// net_output() [B][H][W][net_out_channels()], channels 0 and 1 used
int net_forward(NetImpl* n, const float* in, int B, hipStream_t s);
const float* net_output(const NetImpl* n);
int net_out_channels();

}  // namespace hps
