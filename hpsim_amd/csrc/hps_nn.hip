// hps_nn.hip -- the learned SCHEFF field surrogate: a small dilated
// convolutional network on the (nz+1) x (nr+1) field mesh, with BF16 weights
// and FP32 accumulation on the MFMA units (v_mfma_f32_16x16x16_bf16).
//
// Layout: activations NHWC FP32 ([position][channel], position = (b*H + y)*W
// + x, y along z, x along r).  Convolutions are implicit GEMMs: for the
// forward pass M = positions, N = output channels, K = taps x input channels
// (tap-major).  One wavefront computes 16 positions x all output channels;
// fragments are loaded straight from global memory (the tensors are small
// and L2-resident), converted to BF16 by round-to-nearest-even.
//
// MFMA 16x16x16 operand layout (wave64): lane l holds A[m][k] for
// m = l % 16, k = 4 (l / 16) + j and B[k][n] for n = l % 16,
// k = 4 (l / 16) + j (j = 0..3); the result D[m][n] for n = l % 16,
// m = 4 (l / 16) + i (i = 0..3).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hps_nn.hpp"

namespace hps {
namespace nn {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;   // wavefronts per workgroup

// float -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ s4 pack(float a, float b, float c, float d)
{
  s4 r;
  r.x = (short)f2bf(a); r.y = (short)f2bf(b); r.z = (short)f2bf(c); r.w = (short)f2bf(d);
  return r;
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// tap offset (dy, dx) of tap t
template <int TAPS>
__device__ __forceinline__ void tap(int t, int dil, int& dy, int& dx)
{
  if (TAPS == 1) { dy = 0; dx = 0; }
  else { dy = (t / 3 - 1) * dil; dx = (t % 3 - 1) * dil; }
}

// Forward: out = [resid +] act(conv(in) + bias), act = ReLU if relu; the
// activation (before the residual add) is kept in act_out for the backward
// pass when given.  Cout <= kMaxC, multiple of 16; Cin multiple of 4.
template <int TAPS>
__global__ __launch_bounds__(64 * kWaves) void k_conv_fwd(const float* __restrict__ in,
                                                        const uint16_t* __restrict__ wt,
                                                        const float* __restrict__ bias,
                                                        const float* __restrict__ resid,
                                                        float* __restrict__ act_out,
                                                        float* __restrict__ out, ConvGeom g,
                                                        int relu)
{
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int P = g.B * g.H * g.W;
  const int p0 = (blockIdx.x * kWaves + wave) * 16;
  if (p0 >= P) return;
  const int K = TAPS * g.Cin;
  const int NT = g.Cout / 16;
  f4 acc[kMaxC / 16];
#pragma unroll
  for (int n = 0; n < kMaxC / 16; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
  // this lane's A row: position pa
  const int pa = p0 + (lane & 15);
  const int xa = pa % g.W, ya = (pa / g.W) % g.H, ba = pa / (g.W * g.H);
  const int kq = 4 * (lane >> 4);
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int kk = k0 + kq;
    s4 a = s4{0, 0, 0, 0};
    if (kk < K && pa < P) {
      const int t = kk / g.Cin, ci = kk % g.Cin;
      int dy, dx;
      tap<TAPS>(t, g.dil, dy, dx);
      const int yy = ya + dy, xx = xa + dx;
      if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W) {
        const float4 v = *reinterpret_cast<const float4*>(
            in + ((size_t)(ba * g.H + yy) * g.W + xx) * g.Cin + ci);
        a = pack(v.x, v.y, v.z, v.w);
      }
    }
#pragma unroll
    for (int n = 0; n < kMaxC / 16; ++n) {
      if (n < NT) {
        s4 b = s4{0, 0, 0, 0};
        if (kk < K) {
          const uint16_t* w = wt + (size_t)(n * 16 + (lane & 15)) * K + kk;
          b.x = (short)w[0]; b.y = (short)w[1]; b.z = (short)w[2]; b.w = (short)w[3];
        }
        acc[n] = mfma(a, b, acc[n]);
      }
    }
  }
#pragma unroll
  for (int n = 0; n < kMaxC / 16; ++n) {
    if (n >= NT) continue;
    const int co = n * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = p0 + 4 * (lane >> 4) + i;
      if (p >= P) continue;
      float v = acc[n][i] + bias[co];
      if (relu) v = v > 0.f ? v : 0.f;
      const size_t o = (size_t)p * g.Cout + co;
      if (act_out) act_out[o] = v;
      if (resid) v += resid[o];
      out[o] = v;
    }
  }
}

// dZ = dOut, masked by ReLU'(act) when relu
__device__ __forceinline__ float dz_at(const float* dout, const float* act, size_t o, int relu)
{
  const float d = dout[o];
  return relu ? (act[o] > 0.f ? d : 0.f) : d;
}

// Backward data: din[q][ci] = [dres[q][ci] +] sum_{t,co} W[co][t][ci] dZ[q - off_t][co]
// (wb = W as [Cin][TAPS][Cout] bf16).  Cin multiple of 16, Cout multiple of 4.
template <int TAPS>
__global__ __launch_bounds__(64 * kWaves) void k_conv_bwd_data(const float* __restrict__ dout,
                                                             const float* __restrict__ act,
                                                             const uint16_t* __restrict__ wb,
                                                             const float* __restrict__ dres,
                                                             float* __restrict__ din,
                                                             ConvGeom g, int relu)
{
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int P = g.B * g.H * g.W;
  const int q0 = (blockIdx.x * kWaves + wave) * 16;
  if (q0 >= P) return;
  const int K = TAPS * g.Cout;
  const int NT = g.Cin / 16;
  f4 acc[kMaxC / 16];
#pragma unroll
  for (int n = 0; n < kMaxC / 16; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
  const int qa = q0 + (lane & 15);
  const int xa = qa % g.W, ya = (qa / g.W) % g.H, ba = qa / (g.W * g.H);
  const int kq = 4 * (lane >> 4);
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int kk = k0 + kq;
    s4 a = s4{0, 0, 0, 0};
    if (kk < K && qa < P) {
      const int t = kk / g.Cout, co = kk % g.Cout;
      int dy, dx;
      tap<TAPS>(t, g.dil, dy, dx);
      const int yy = ya - dy, xx = xa - dx;    // the position whose tap t reads q
      if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W) {
        const size_t o = ((size_t)(ba * g.H + yy) * g.W + xx) * g.Cout + co;
        a = pack(dz_at(dout, act, o, relu), dz_at(dout, act, o + 1, relu),
                 dz_at(dout, act, o + 2, relu), dz_at(dout, act, o + 3, relu));
      }
    }
#pragma unroll
    for (int n = 0; n < kMaxC / 16; ++n) {
      if (n < NT) {
        s4 b = s4{0, 0, 0, 0};
        if (kk < K) {
          const uint16_t* w = wb + (size_t)(n * 16 + (lane & 15)) * K + kk;
          b.x = (short)w[0]; b.y = (short)w[1]; b.z = (short)w[2]; b.w = (short)w[3];
        }
        acc[n] = mfma(a, b, acc[n]);
      }
    }
  }
#pragma unroll
  for (int n = 0; n < kMaxC / 16; ++n) {
    if (n >= NT) continue;
    const int ci = n * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int q = q0 + 4 * (lane >> 4) + i;
      if (q >= P) continue;
      const size_t o = (size_t)q * g.Cin + ci;
      din[o] = acc[n][i] + (dres ? dres[o] : 0.f);
    }
  }
}

// Backward weights: dw[co][t][ci] += sum_p dZ[p][co] in[p + off_t][ci] over a
// chunk of positions per workgroup (split K, FP32 atomics: the order of the
// chunk sums is not fixed, which training tolerates).  One wavefront per
// (co tile, (t,ci) tile) pair.
template <int TAPS>
__global__ __launch_bounds__(64 * kWaves) void k_conv_bwd_w(const float* __restrict__ dout,
                                                          const float* __restrict__ act,
                                                          const float* __restrict__ in,
                                                          float* __restrict__ dw, ConvGeom g,
                                                          int relu, int chunk)
{
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int P = g.B * g.H * g.W;
  const int N = TAPS * g.Cin;
  const int MT = g.Cout / 16, NTl = (N + 15) / 16;
  const int pair = blockIdx.y * kWaves + wave;
  if (pair >= MT * NTl) return;
  const int mt = pair / NTl, nt = pair % NTl;
  const int pbeg = blockIdx.x * chunk;
  const int pend = pbeg + chunk < P ? pbeg + chunk : P;
  const int co = mt * 16 + (lane & 15);          // A row
  const int nb = nt * 16 + (lane & 15);          // B column
  const bool nok = nb < N;
  const int t = nok ? nb / g.Cin : 0, ci = nok ? nb % g.Cin : 0;
  int dy, dx;
  tap<TAPS>(t, g.dil, dy, dx);
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = pbeg; k0 < pend; k0 += 16) {
    float av[4], bv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = k0 + 4 * (lane >> 4) + j;
      av[j] = 0.f;
      bv[j] = 0.f;
      if (p < pend) {
        av[j] = dz_at(dout, act, (size_t)p * g.Cout + co, relu);
        if (nok) {
          const int x = p % g.W, y = (p / g.W) % g.H, b = p / (g.W * g.H);
          const int yy = y + dy, xx = x + dx;
          if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W)
            bv[j] = in[((size_t)(b * g.H + yy) * g.W + xx) * g.Cin + ci];
        }
      }
    }
    acc = mfma(pack(av[0], av[1], av[2], av[3]), pack(bv[0], bv[1], bv[2], bv[3]), acc);
  }
  const int n = nt * 16 + (lane & 15);
  if (n < N) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = mt * 16 + 4 * (lane >> 4) + i;
      atomicAdd(&dw[(size_t)m * N + n], acc[i]);
    }
  }
}

// dbias[co] += sum over a chunk of positions of dZ[p][co]
__global__ void k_bias_grad(const float* __restrict__ dout, const float* __restrict__ act,
                            float* __restrict__ db, int P, int C, int relu, int chunk)
{
  const int c = threadIdx.x;
  if (c >= C) return;
  const int pbeg = blockIdx.x * chunk;
  const int pend = pbeg + chunk < P ? pbeg + chunk : P;
  float s = 0.f;
  for (int p = pbeg; p < pend; ++p) s += dz_at(dout, act, (size_t)p * C + c, relu);
  atomicAdd(&db[c], s);
}

// Relative-L2 loss per sample over the first 2 channels of pred ([B][HW][pc])
// against tgt ([B][HW][2]): loss_b = sum (p - t)^2 / (sum t^2 + eps);
// dpred = 2 (p - t) / (sum t^2 + eps) / B (0 in channels >= 2).  One
// workgroup per sample.
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ pred, int pc,
                                              const float* __restrict__ tgt,
                                              float* __restrict__ dpred,
                                              float* __restrict__ loss, int HW, int B)
{
  __shared__ float rn[256], rd[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const size_t pb = (size_t)b * HW * pc, tb = (size_t)b * HW * 2;
  float num = 0.f, den = 0.f;
  for (int i = t; i < HW * 2; i += 256) {
    const float e = pred[pb + (size_t)(i >> 1) * pc + (i & 1)] - tgt[tb + i];
    num += e * e;
    den += tgt[tb + i] * tgt[tb + i];
  }
  rn[t] = num; rd[t] = den;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (t < o) { rn[t] += rn[t + o]; rd[t] += rd[t + o]; }
    __syncthreads();
  }
  const float inv = 1.f / (rd[0] + 1e-30f);
  if (t == 0) loss[b] = rn[0] * inv;
  if (dpred)
    for (int i = t; i < HW * pc; i += 256) {
      const int pos = i / pc, c = i % pc;
      dpred[pb + i] = c < 2 ? 2.f * (pred[pb + i] - tgt[tb + 2 * (size_t)pos + c]) * inv / (float)B
                            : 0.f;
    }
}

__global__ void k_gather_rows(const float* __restrict__ src, const int32_t* __restrict__ idx,
                              float* __restrict__ dst, int n, int row)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)n * row) return;
  const size_t r = i / row, c = i % row;
  dst[i] = src[(size_t)idx[r] * row + c];
}

// Adam step on the FP32 master parameters (decoupled bias correction in lr_t)
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr_t, float b1, float b2, float eps)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float g = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// BF16 weight copies for the two GEMM directions: wt[co][t][ci] (forward B
// operand) and wb[ci][t][co] (backward-data B operand) from w[co][t][ci].
__global__ void k_weights_bf16(const float* __restrict__ w, uint16_t* __restrict__ wt,
                               uint16_t* __restrict__ wb, int Cout, int taps, int Cin)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int n = Cout * taps * Cin;
  if (i >= n) return;
  const int ci = i % Cin, t = (i / Cin) % taps, co = i / (Cin * taps);
  const uint16_t h = f2bf(w[i]);
  wt[i] = h;
  wb[((size_t)ci * taps + t) * Cout + co] = h;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

hipError_t conv_fwd(const float* in, const uint16_t* wt, const float* bias, const float* resid,
                    float* act_out, float* out, const ConvGeom& g, int taps, int relu,
                    hipStream_t s)
{
  const int P = g.B * g.H * g.W;
  const dim3 grid(cdiv(P, 16 * kWaves)), block(64 * kWaves);
  if (taps == 9)
    hipLaunchKernelGGL(k_conv_fwd<9>, grid, block, 0, s, in, wt, bias, resid, act_out, out, g, relu);
  else
    hipLaunchKernelGGL(k_conv_fwd<1>, grid, block, 0, s, in, wt, bias, resid, act_out, out, g, relu);
  return hipGetLastError();
}

hipError_t conv_bwd_data(const float* dout, const float* act, const uint16_t* wb, const float* dres,
                         float* din, const ConvGeom& g, int taps, int relu, hipStream_t s)
{
  const int P = g.B * g.H * g.W;
  const dim3 grid(cdiv(P, 16 * kWaves)), block(64 * kWaves);
  if (taps == 9)
    hipLaunchKernelGGL(k_conv_bwd_data<9>, grid, block, 0, s, dout, act, wb, dres, din, g, relu);
  else
    hipLaunchKernelGGL(k_conv_bwd_data<1>, grid, block, 0, s, dout, act, wb, dres, din, g, relu);
  return hipGetLastError();
}

hipError_t conv_bwd_w(const float* dout, const float* act, const float* in, float* dw, float* db,
                      const ConvGeom& g, int taps, int relu, hipStream_t s)
{
  const int P = g.B * g.H * g.W;
  const int N = taps * g.Cin;
  const int pairs = (g.Cout / 16) * cdiv(N, 16);
  const int chunk = 1024;
  const dim3 grid(cdiv(P, chunk), cdiv(pairs, kWaves)), block(64 * kWaves);
  if (taps == 9)
    hipLaunchKernelGGL(k_conv_bwd_w<9>, grid, block, 0, s, dout, act, in, dw, g, relu, chunk);
  else
    hipLaunchKernelGGL(k_conv_bwd_w<1>, grid, block, 0, s, dout, act, in, dw, g, relu, chunk);
  hipLaunchKernelGGL(k_bias_grad, dim3(cdiv(P, chunk)), dim3(kMaxC), 0, s, dout, act, db, P,
                     g.Cout, relu, chunk);
  return hipGetLastError();
}

hipError_t rel_l2_loss(const float* pred, int pc, const float* tgt, float* dpred, float* loss,
                       int HW, int B, hipStream_t s)
{
  hipLaunchKernelGGL(k_loss, dim3(B), dim3(256), 0, s, pred, pc, tgt, dpred, loss, HW, B);
  return hipGetLastError();
}

hipError_t gather_rows(const float* src, const int32_t* idx, float* dst, int n, int row,
                       hipStream_t s)
{
  const size_t total = (size_t)n * row;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src,
                     idx, dst, n, row);
  return hipGetLastError();
}

hipError_t adam(float* w, const float* g, float* m, float* v, int n, float lr_t, float b1,
                float b2, float eps, hipStream_t s)
{
  hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, w, g, m, v, n, lr_t, b1, b2,
                     eps);
  return hipGetLastError();
}

hipError_t weights_bf16(const float* w, uint16_t* wt, uint16_t* wb, int Cout, int taps, int Cin,
                        hipStream_t s)
{
  const int n = Cout * taps * Cin;
  hipLaunchKernelGGL(k_weights_bf16, dim3(cdiv(n, 256)), dim3(256), 0, s, w, wt, wb, Cout, taps,
                     Cin);
  return hipGetLastError();
}

}  // namespace nn
}  // namespace hps
