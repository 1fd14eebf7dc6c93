// hps_unet.hpp -- kernels of the SCHEFF field surrogate network
// (hps_unet.hip) and the network object the batch engine runs it through
// (hps_surrogate.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hps {
namespace unet {

// Layer modes: source position zi of output position zo for tap t.
enum : int {
  kSame = 0,       // 3 taps: zi = zo + (t - 1) * dil
  kPoint = 1,      // 1 tap:  zi = zo
  kDown = 2,       // 2 taps: zi = 2 zo + t            (Zout = Zin / 2)
  kUp = 3,         // 2 taps: zi = zo / 2 if zo % 2 == t (Zout = 2 Zin)
  kSameFlip = 4    // 3 taps: zi = zo + (1 - t) * dil  (backward data of kSame)
};
constexpr int kMaxCout = 144;        // most output channels of a layer
constexpr int kWgradChunk = 256;     // positions per weight-gradient task

struct ConvArgs {
  const float* in;          // [B][Zin][Cin]
  const uint16_t* w;        // BF16 [Cout][taps][Cin]
  float* pre;               // optional: the product before scale and skip
  float* out;               // [B][Zout][Cout]
  const float* scale;       // optional [B][Cout]
  const float* skip;        // optional [B][Zout][Cout], added last
  const uint32_t* dyn_b;    // optional device sample count (<= B)
  int B, Zin, Zout, Cin, Cout, taps, dil, mode;
};

struct WgradArgs {
  const float* dpre;        // [B][Zout][Cout]
  const float* in;          // [B][Zin][Cin]
  float* dw;                // [Cout][taps][Cin], accumulated
  int B, Zin, Zout, Cin, Cout, taps, dil, mode;
};

struct FilmLayer {
  int cout;
  int pad_;
  size_t a_off;             // FiLM coefficients A[cout][3] in the parameters
  size_t s_off;             // this layer's scales [max_b][cout] in the scale buffer
};

hipError_t conv(const ConvArgs& a, hipStream_t s);
hipError_t wgrad(const WgradArgs& a, hipStream_t s);
hipError_t film(const float* params, const FilmLayer* layers, int nlayers, const float* u,
                const uint32_t* dyn_b, int B, float* scale, hipStream_t s);
hipError_t film_bwd(const float* dy, const float* pre, const float* scale, const float* u,
                    float* dpre, float* dA, int B, int Z, int C, hipStream_t s);
// per-sample relative squared error of y against tgt * inv_scale; dy the
// gradient of its mean (optional)
hipError_t loss(const float* y, const float* tgt, float* dy, float* l, int Z0, int Cpad, int zn,
                int cout, float inv_scale, int B, hipStream_t s);
hipError_t adam(float* w, const float* g, float* m, float* v, int n, float lr_t, float b1,
                float b2, float eps, hipStream_t s);
hipError_t bf16_copies(const float* w, uint16_t* wt, uint16_t* wb, int Cout, int taps, int Cin,
                       hipStream_t s);
hipError_t gather(const float* src, const uint32_t* idx, float* dst, int n, int row,
                  hipStream_t s);

}  // namespace unet

// The network as the batch engine uses it (hps_surrogate.hip).
struct NetImpl;
struct NetView {
  int nr, nz, Z0, cin, cout, cpad, zn;
  float out_scale, u_mu, u_sigma;
};
NetView net_view(const NetImpl* n);
// forward pass of up to max_b samples (count in device memory): x
// [max_b][Z0][cin], u [max_b] (normalised conditioning) -> y [max_b][Z0][cpad]
int net_infer(NetImpl* n, const float* x, const float* u, const uint32_t* dyn_b, int max_b,
              float** y, hipStream_t s);
// inference buffers for up to B samples (allocates; call before capture)
int net_reserve_infer(NetImpl* n, int B);
// device dataset the batch appends (charge, field) pairs to
struct NetData {
  float *x, *y, *u;          // [cap][nz][cin], [cap][zn][cout], [cap] (log aspect)
  uint32_t* count;           // device counter
  uint32_t cap;
};
int net_data(NetImpl* n, NetData* out);

}  // namespace hps

struct hps_nn;
namespace hps {
NetImpl* net_of(hps_nn* h);

}  // namespace hps
