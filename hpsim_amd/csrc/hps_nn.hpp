// hps_nn.hpp -- launchers of the surrogate network kernels (hps_nn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hps {
namespace nn {

constexpr int kMaxC = 64;   // most channels of a layer

// one convolution: B samples of H x W positions (NHWC), Cin -> Cout channels,
// 3x3 taps with dilation dil (or 1x1)
struct ConvGeom {
  int B, H, W, Cin, Cout, dil;
};

// out = [resid +] act(conv(in) + bias); act_out (optional) keeps the
// activation before the residual add.  Cin % 4 == 0, Cout % 16 == 0.
hipError_t conv_fwd(const float* in, const uint16_t* wt, const float* bias, const float* resid,
                    float* act_out, float* out, const ConvGeom& g, int taps, int relu,
                    hipStream_t s);
// din = [dres +] conv^T(dZ), dZ = dout masked by ReLU'(act) when relu.  Cin % 16 == 0.
hipError_t conv_bwd_data(const float* dout, const float* act, const uint16_t* wb, const float* dres,
                         float* din, const ConvGeom& g, int taps, int relu, hipStream_t s);
// dw += dZ^T im2col(in), db += column sums of dZ (FP32 atomics)
hipError_t conv_bwd_w(const float* dout, const float* act, const float* in, float* dw, float* db,
                      const ConvGeom& g, int taps, int relu, hipStream_t s);
// per-sample relative L2 loss over the first 2 channels of pred ([B][HW][pc])
// against tgt ([B][HW][2]); dpred (optional) its gradient, zero in the other
// channels
hipError_t rel_l2_loss(const float* pred, int pc, const float* tgt, float* dpred, float* loss,
                       int HW, int B, hipStream_t s);
hipError_t adam(float* w, const float* g, float* m, float* v, int n, float lr_t, float b1,
                float b2, float eps, hipStream_t s);
// BF16 copies of w[Cout][taps][Cin]: wt (same layout) and wb[Cin][taps][Cout]
hipError_t weights_bf16(const float* w, uint16_t* wt, uint16_t* wb, int Cout, int taps, int Cin,
                        hipStream_t s);
// dst[i][..] = src[idx[i]][..] for rows of `row` floats
hipError_t gather_rows(const float* src, const int32_t* idx, float* dst, int n, int row,
                       hipStream_t s);

}  // namespace nn
}  // namespace hps
