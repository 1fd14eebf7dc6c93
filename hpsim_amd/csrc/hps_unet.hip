// hps_unet.hip -- kernels of the learned SCHEFF field surrogate: a
// one-dimensional convolutional U-Net along the bunch axis z with the source
// rings as channels (network object and C ABI in hps_surrogate.hip): the
// convolutions forward and backward on the BF16 matrix cores, FiLM, loss,
// Adam.
//
// Why this shape: the SCHEFF field is linear in the charge histogram and
// invariant under shifts along z (the Green's table depends on zp - zs only,
// space_charge_kernel.cu:38-100), but not under shifts in r.  So the network
// convolves along z, mixes rings densely in its channels, has no biases and
// no activations (superposition holds exactly), and takes the mesh aspect --
// the one parameter of the normalised Green's function -- through FiLM
// scales 1 + sum_j A[l][c][j] u^j (j = 1..3) on every layer's channels.
// The U-Net levels (z, z/2, z/4, z/8, and dilated convolutions at the
// coarsest level) reach the whole bunch, as the Coulomb field does.
//
// Input x[b][z][rs] = n(zs = z, rs) * nr * nz / Q for z < nz and 0 for the
// padding up to Z0 = nz + 8 (n particles per bin, Q their sum); output
// y[b][z][2 rp + comp] for the field nodes z <= nz; the field (before the
// c1 / dr^2 factor) is y * Q * out_scale.
//
// Precision: weights BF16 (FP32 master copies for Adam), accumulation FP32
// on v_mfma_f32_16x16x16_bf16; activations stay FP32 in memory and enter
// the products as two BF16 terms (hi + lo), so the only rounding of an
// activation is that of FP32.
//
// MFMA 16x16x16 bf16 operand layout (wave64): lane l supplies A[m][k] for
// m = l % 16, k = 4 (l / 16) + j and B[k][n] for n = l % 16, k = 4 (l / 16)
// + j (j = 0..3); the result D[m][n] is held for n = l % 16, m = 4 (l / 16)
// + i in element i.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "hpsim_amd.h"
#include "hps_unet.hpp"

namespace hps {

int fail_msg(const std::string& m);   // hps_engine.hip

namespace unet {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;
constexpr int kMaxFrag = kMaxCout / 16;

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// float -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// v = hi + lo with hi = bf16(v), lo = bf16(v - hi)
__device__ __forceinline__ void split4(const float4& v, s4& hi, s4& lo)
{
  const float a[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint16_t h = f2bf(a[j]);
    hi[j] = (short)h;
    lo[j] = (short)f2bf(a[j] - bf2f(h));
  }
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// source position of output position zo for tap t (-1: none)
__device__ __forceinline__ int src_of(int mode, int zo, int t, int dil, int zin)
{
  int zi;
  switch (mode) {
    case kSame: zi = zo + (t - 1) * dil; break;
    case kSameFlip: zi = zo + (1 - t) * dil; break;
    case kPoint: zi = zo; break;
    case kDown: zi = 2 * zo + t; break;
    default: zi = (zo & 1) == t ? zo >> 1 : -1; break;   // kUp
  }
  return (zi >= 0 && zi < zin) ? zi : -1;
}

// pre = conv(in); out = pre * scale[b][co] (+ skip).  One wave: 16 output
// positions x all Cout.  Cin % 16 == 0, Cout % 16 == 0, Cout <= kMaxCout.
__global__ __launch_bounds__(64 * kWaves) void k_conv(ConvArgs a)
{
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int B = a.dyn_b ? (int)a.dyn_b[0] : a.B;
  const int P = B * a.Zout;
  const int p0 = (blockIdx.x * kWaves + wave) * 16;
  if (p0 >= P) return;
  const int NF = a.Cout / 16;
  f4 acc[kMaxFrag];
#pragma unroll
  for (int n = 0; n < kMaxFrag; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
  const int pa = p0 + (lane & 15);
  const int ba = pa / a.Zout, za = pa % a.Zout;
  const int kq = 4 * (lane >> 4);
  for (int t = 0; t < a.taps; ++t) {
    const int zi = pa < P ? src_of(a.mode, za, t, a.dil, a.Zin) : -1;
    const float* row = a.in + ((size_t)ba * a.Zin + (zi < 0 ? 0 : zi)) * a.Cin;
    for (int c0 = 0; c0 < a.Cin; c0 += 16) {
      s4 hi = s4{0, 0, 0, 0}, lo = s4{0, 0, 0, 0};
      if (zi >= 0) split4(*reinterpret_cast<const float4*>(row + c0 + kq), hi, lo);
#pragma unroll
      for (int n = 0; n < kMaxFrag; ++n) {
        if (n >= NF) continue;
        const uint16_t* w = a.w + ((size_t)(n * 16 + (lane & 15)) * a.taps + t) * a.Cin + c0 + kq;
        const uint2 wv = *reinterpret_cast<const uint2*>(w);
        s4 b;
        b[0] = (short)(wv.x & 0xffffu); b[1] = (short)(wv.x >> 16);
        b[2] = (short)(wv.y & 0xffffu); b[3] = (short)(wv.y >> 16);
        acc[n] = mfma(hi, b, acc[n]);
        acc[n] = mfma(lo, b, acc[n]);
      }
    }
  }
#pragma unroll
  for (int n = 0; n < kMaxFrag; ++n) {
    if (n >= NF) continue;
    const int co = n * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p = p0 + 4 * (lane >> 4) + i;
      if (p >= P) continue;
      const size_t o = (size_t)p * a.Cout + co;
      float v = acc[n][i];
      if (a.pre) a.pre[o] = v;
      if (a.scale) v *= a.scale[(size_t)(p / a.Zout) * a.Cout + co];
      if (a.skip) v += a.skip[o];
      a.out[o] = v;
    }
  }
}

// dw[co][t][ci] += sum over positions of dpre[p][co] * in[src(p, t)][ci]:
// one wave per (tap, 16 co, 16 ci, chunk of 256 positions); FP32 atomics.
__global__ __launch_bounds__(64 * kWaves) void k_wgrad(WgradArgs a)
{
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int P = a.B * a.Zout;
  const int nco = a.Cout / 16, nci = a.Cin / 16;
  const int nchunk = cdiv(P, kWgradChunk);
  const int task = blockIdx.x * kWaves + wave;
  if (task >= a.taps * nco * nci * nchunk) return;
  const int chunk = task % nchunk;
  int r = task / nchunk;
  const int cit = r % nci; r /= nci;
  const int cot = r % nco;
  const int t = r / nco;
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  const int kq = 4 * (lane >> 4);
  const int co = cot * 16 + (lane & 15), ci = cit * 16 + (lane & 15);
  const int pend = min(P, (chunk + 1) * kWgradChunk);
  for (int q0 = chunk * kWgradChunk; q0 < pend; q0 += 16) {
    float dv[4], xv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = q0 + kq + j;
      dv[j] = 0.f;
      xv[j] = 0.f;
      if (p < pend) {
        const int b = p / a.Zout, zo = p % a.Zout;
        const int zi = src_of(a.mode, zo, t, a.dil, a.Zin);
        if (zi >= 0) {
          dv[j] = a.dpre[(size_t)p * a.Cout + co];
          xv[j] = a.in[((size_t)b * a.Zin + zi) * a.Cin + ci];
        }
      }
    }
    s4 da, xh, xl;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      da[j] = (short)f2bf(dv[j]);
      const uint16_t h = f2bf(xv[j]);
      xh[j] = (short)h;
      xl[j] = (short)f2bf(xv[j] - bf2f(h));
    }
    acc = mfma(da, xh, acc);
    acc = mfma(da, xl, acc);
  }
  // D[m = co][n = ci]: this lane holds ci = cit*16 + lane%16, co rows 4(lane/16)+i
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int cor = cot * 16 + 4 * (lane >> 4) + i;
    atomicAdd(&a.dw[((size_t)cor * a.taps + t) * a.Cin + ci], acc[i]);
  }
}

// Bounded conditioning v = 2 tanh(u / 2) (the FiLM polynomial stays within
// the range it was trained on)
__device__ __forceinline__ float film_v(float u) { return 2.f * tanhf(0.5f * u); }

// FiLM scales of every layer: scale[l][b][co] = 1 + sum_j A[l][co][j] v_b^j
__global__ void k_film(const float* __restrict__ params, const FilmLayer* __restrict__ layers,
                       int nlayers, const float* __restrict__ u, const uint32_t* dyn_b, int B,
                       float* __restrict__ scale, int max_b)
{
  const int nb = dyn_b ? (int)dyn_b[0] : B;
  const int l = blockIdx.y;
  if (l >= nlayers) return;
  const FilmLayer L = layers[l];
  for (int e = blockIdx.x * 256 + threadIdx.x; e < nb * L.cout; e += gridDim.x * 256) {
    const int b = e / L.cout, co = e % L.cout;
    const float x = film_v(u[b]);
    const float* A = params + L.a_off + (size_t)co * 3;
    scale[L.s_off + (size_t)b * L.cout + co] = 1.f + A[0] * x + A[1] * x * x + A[2] * x * x * x;
  }
}

// dpre = dy * scale; g[b][co] = sum_z dy * pre -> dA[co][j] += g v_b^j
__global__ void k_film_bwd(const float* __restrict__ dy, const float* __restrict__ pre,
                           const float* __restrict__ scale, const float* __restrict__ u,
                           float* __restrict__ dpre, float* __restrict__ dA, int B, int Z, int C)
{
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= B * C) return;
  const int b = e / C, co = e % C;
  const float s = scale[(size_t)b * C + co];
  float g = 0.f;
  for (int z = 0; z < Z; ++z) {
    const size_t o = ((size_t)b * Z + z) * C + co;
    const float d = dy[o];
    g += d * pre[o];
    dpre[o] = d * s;
  }
  const float x = film_v(u[b]);
  atomicAdd(&dA[(size_t)co * 3 + 0], g * x);
  atomicAdd(&dA[(size_t)co * 3 + 1], g * x * x);
  atomicAdd(&dA[(size_t)co * 3 + 2], g * x * x * x);
}

// loss_b = |y - t / out_scale|^2 / |t / out_scale|^2 over z < zn, c < cout
// (y [B][Z0][Cpad], t [B][zn][cout]); dy (optional) = d loss_mean / d y
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ tgt,
                                              float* __restrict__ dy, float* __restrict__ loss,
                                              int Z0, int Cpad, int zn, int cout, float inv_scale,
                                              int B)
{
  __shared__ float red[2][256];
  const int b = blockIdx.x, t = threadIdx.x;
  float nt = 0.f, nd = 0.f;
  for (int e = t; e < zn * cout; e += 256) {
    const int z = e / cout, c = e % cout;
    const float tv = tgt[((size_t)b * zn + z) * cout + c] * inv_scale;
    const float d = y[((size_t)b * Z0 + z) * Cpad + c] - tv;
    nt += tv * tv;
    nd += d * d;
  }
  red[0][t] = nt;
  red[1][t] = nd;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) { red[0][t] += red[0][t + o]; red[1][t] += red[1][t + o]; }
    __syncthreads();
  }
  const float ntot = red[0][0] > 0.f ? red[0][0] : 1.f;
  if (t == 0) loss[b] = red[1][0] / ntot;
  if (!dy) return;
  const float k = 2.f / (ntot * (float)B);
  for (int e = t; e < Z0 * Cpad; e += 256) {
    const int z = e / Cpad, c = e % Cpad;
    float v = 0.f;
    if (z < zn && c < cout)
      v = k * (y[((size_t)b * Z0 + z) * Cpad + c] - tgt[((size_t)b * zn + z) * cout + c] * inv_scale);
    dy[((size_t)b * Z0) * Cpad + e] = v;
  }
}

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

// BF16 copies of W[Cout][taps][Cin]: wt (same layout), wb[Cin][taps][Cout]
__global__ void k_bf16(const float* __restrict__ w, uint16_t* __restrict__ wt,
                       uint16_t* __restrict__ wb, int Cout, int taps, int Cin)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= Cout * taps * Cin) return;
  const int ci = i % Cin, t = (i / Cin) % taps, co = i / (Cin * taps);
  const uint16_t h = f2bf(w[i]);
  wt[i] = h;
  wb[((size_t)ci * taps + t) * Cout + co] = h;
}

// rows idx[i] of src (row floats each) -> dst[i]
__global__ void k_gather(const float* __restrict__ src, const uint32_t* __restrict__ idx,
                         float* __restrict__ dst, int n, int row)
{
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)n * row) return;
  const size_t i = e / row, k = e % row;
  dst[e] = src[(size_t)idx[i] * row + k];
}

}  // namespace

hipError_t conv(const ConvArgs& a, hipStream_t s)
{
  const int P = a.B * a.Zout;   // dyn_b <= B: the grid covers the largest count
  hipLaunchKernelGGL(k_conv, dim3(cdiv(P, 16 * kWaves)), dim3(64 * kWaves), 0, s, a);
  return hipGetLastError();
}

hipError_t wgrad(const WgradArgs& a, hipStream_t s)
{
  const int tasks = a.taps * (a.Cout / 16) * (a.Cin / 16) * cdiv(a.B * a.Zout, kWgradChunk);
  hipLaunchKernelGGL(k_wgrad, dim3(cdiv(tasks, kWaves)), dim3(64 * kWaves), 0, s, a);
  return hipGetLastError();
}

hipError_t film(const float* params, const FilmLayer* layers, int nlayers, const float* u,
                const uint32_t* dyn_b, int B, float* scale, hipStream_t s)
{
  hipLaunchKernelGGL(k_film, dim3(64, nlayers), dim3(256), 0, s, params, layers, nlayers, u, dyn_b,
                     B, scale, B);
  return hipGetLastError();
}

hipError_t film_bwd(const float* dy, const float* pre, const float* scale, const float* u,
                    float* dpre, float* dA, int B, int Z, int C, hipStream_t s)
{
  hipLaunchKernelGGL(k_film_bwd, dim3(cdiv(B * C, 256)), dim3(256), 0, s, dy, pre, scale, u, dpre,
                     dA, B, Z, C);
  return hipGetLastError();
}

hipError_t loss(const float* y, const float* tgt, float* dy, float* l, int Z0, int Cpad, int zn,
                int cout, float inv_scale, int B, hipStream_t s)
{
  hipLaunchKernelGGL(k_loss, dim3(B), dim3(256), 0, s, y, tgt, dy, l, Z0, Cpad, zn, cout,
                     inv_scale, B);
  return hipGetLastError();
}

hipError_t adam(float* w, const float* g, float* m, float* v, int n, float lr_t, float b1,
                float b2, float eps, hipStream_t s)
{
  hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, w, g, m, v, n, lr_t, b1, b2,
                     eps);
  return hipGetLastError();
}

hipError_t bf16_copies(const float* w, uint16_t* wt, uint16_t* wb, int Cout, int taps, int Cin,
                       hipStream_t s)
{
  hipLaunchKernelGGL(k_bf16, dim3(cdiv(Cout * taps * Cin, 256)), dim3(256), 0, s, w, wt, wb, Cout,
                     taps, Cin);
  return hipGetLastError();
}

hipError_t gather(const float* src, const uint32_t* idx, float* dst, int n, int row, hipStream_t s)
{
  const size_t total = (size_t)n * row;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, idx,
                     dst, n, row);
  return hipGetLastError();
}

}  // namespace unet
}  // namespace hps
