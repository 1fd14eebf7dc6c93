// hps_field_fft.hip -- the grouped field product of a batch in the frequency
// domain along z.
//
// The grouped product (hps_field_gemm.hip) evaluates, for every remeshing
// member's column p and its aspect class c,
//
//   W[p][(rp, comp)][zp] = sum over rs, zs of S_c[rp][rs][comp][zp - zs] q[p][rs][zs]
//
// for zp = 0 .. nz: per pair of rings a convolution along z with the signed
// offset table S (offsets 1 - nz .. nz, the Green's entries of
// UpdateFldTbl1Kernel, space_charge_kernel.cu:38-100).  With L = 2 nz the
// circular convolution of the zero-padded charges with S laid out by offset
// modulo L has no wrap-around for these zp, so
//
//   W[p][row][.] = IDFT_L( sum over rs of S^_c[row][rs][f] q^_p[rs][f] )
//
// with ^ the DFT of length L (row = 2 rp + comp).  Per frequency f the sum
// over rs is a complex matrix product (2 (nr + 1) rows x nr rings x the
// columns of the class), run here on the FP32 matrix cores as a real product
// of twice the size; the transforms are radix-2 FFTs in LDS.  Per column
// this is about 2 (L/2 + 1) x 2(nr+1) x 2 nr x 2 flops for the product plus
// the transforms, against 2 x 2(nr+1)(nz+1) x nr nz for the direct product:
// 40 times fewer at 64 x 256.
//
// Launch sequence (one "spectral" shape of launch_gemm_product; it writes
// the same w[p][m] layout as the direct product with one K split, so the
// store, audit and capture kernels read it unchanged):
//   k_fft_sched    slots in use whose spectrum is not that of their class
//   k_fft_tables   forward transforms of those slots' offset tables
//   k_fft_charges  forward transforms of every column's charges
//   k_fft_product  per frequency and column tile: the complex product (MFMA)
//   k_fft_inverse  inverse transforms, W[p][m] for zp = 0 .. nz
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "hps_internal.hpp"

namespace hps {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kFftThreads = 256;
constexpr int kFftLdsComplex = 4096;   // complex entries of a workgroup's transform rows (32 KB)
constexpr int kProdNT = 32;            // columns per product tile
constexpr int kProdMaxK = 128;         // 2 nr: the real K of the product (nr <= 64)

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ f4 mfma(float a, float b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t ilog2(uint32_t v) { return 31u - __builtin_clz(v); }

// In-place forward DFT (exponent -2 pi i f k / L) of `cnt` rows of length L
// (a power of two) in LDS, x[r L + k]: bit-reversal permutation, then
// log2(L) radix-2 stages; tw[k] = exp(-2 pi i k / L), k < L / 2.  All threads
// of the workgroup take part; the rows are in order on return.
__device__ void fft_rows(float2* x, const float2* tw, uint32_t L, uint32_t cnt)
{
  const uint32_t lg = ilog2(L);
  for (uint32_t e = threadIdx.x; e < cnt * L; e += blockDim.x) {
    const uint32_t r = e >> lg, k = e & (L - 1);
    const uint32_t j = __brev(k) >> (32 - lg);
    if (j > k) {
      const float2 t = x[r * L + k];
      x[r * L + k] = x[r * L + j];
      x[r * L + j] = t;
    }
  }
  __syncthreads();
  for (uint32_t half = 1; half < L; half <<= 1) {
    const uint32_t step = L / (2 * half);
    for (uint32_t e = threadIdx.x; e < cnt * (L / 2); e += blockDim.x) {
      const uint32_t r = e / (L / 2), j = e % (L / 2);
      const uint32_t g = j / half, k = j % half;
      const uint32_t i0 = r * L + g * 2 * half + k, i1 = i0 + half;
      const float2 b = cmul(x[i1], tw[k * step]);
      const float2 a = x[i0];
      x[i0] = make_float2(a.x + b.x, a.y + b.y);
      x[i1] = make_float2(a.x - b.x, a.y - b.y);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void load_twiddles(float2* tw_s, const float2* tw, uint32_t L)
{
  for (uint32_t k = threadIdx.x; k < L / 2; k += blockDim.x) tw_s[k] = tw[k];
}

// Spectra of two real rows transformed together as x = a + i b: for
// f = 0 .. L/2, A[f] = (X[f] + conj X[-f]) / 2 and B[f] = (X[f] - conj X[-f]) / (2 i).
__device__ __forceinline__ void split_real_pair(const float2* X, uint32_t L, uint32_t f,
                                                float2& A, float2& B)
{
  const float2 u = X[f], v = X[(L - f) & (L - 1)];
  A = make_float2(0.5f * (u.x + v.x), 0.5f * (u.y - v.y));
  B = make_float2(0.5f * (u.y + v.y), -0.5f * (u.x - v.x));
}

}  // namespace

// ---- spectral class tables ----------------------------------------------------

// Slots the product uses at this kick point (slot_count > 0) whose stored
// spectrum is not that of the class they hold: listed for k_fft_tables and
// marked as done.  One thread.
__global__ void k_fft_sched(const GemmState* __restrict__ gs, int32_t* spec_class,
                            uint32_t* list)
{
  if (threadIdx.x != 0) return;
  uint32_t n = 0;
  for (uint32_t s = 0; s < kGemmSlots; ++s)
    if (gs->slot_count[s] > 0 && spec_class[s] != gs->slot_class[s]) {
      list[1 + n++] = s;
      spec_class[s] = gs->slot_class[s];
    }
  list[0] = n;
}

// S^[slot][f][row][rs] for the listed slots: the offset rows of (rp, rs),
// both components transformed together (x = E_r + i E_z), laid out by offset
// modulo L (k <= nz: offset k at o = k + nz - 1; k > nz: offset k - L at
// o = k - nz - 1).  Grid-stride over (listed slot, group of ring pairs).
__global__ __launch_bounds__(kFftThreads) void k_fft_tables(const uint32_t* __restrict__ list,
                                                            const float* __restrict__ tables,
                                                            float2* __restrict__ shat,
                                                            const float2* __restrict__ tw,
                                                            uint32_t nr, uint32_t nz)
{
  __shared__ float2 x[kFftLdsComplex];
  __shared__ float2 tw_s[kFftLdsComplex / 2];
  const uint32_t L = 2 * nz, F = nz + 1, R = 2 * (nr + 1);
  const uint32_t rows = kFftLdsComplex / L;              // transforms per workgroup
  const uint32_t pairs = (nr + 1) * nr;                   // (rp, rs)
  const uint32_t groups = (pairs + rows - 1) / rows;
  const uint32_t units = list[0] * groups;
  const size_t slot_sz = (size_t)(nr + 1) * nr * 4 * nz;
  load_twiddles(tw_s, tw, L);
  for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint32_t slot = list[1 + u / groups];
    const uint32_t pr0 = (u % groups) * rows;
    const uint32_t cnt = min(rows, pairs - pr0);
    __syncthreads();   // previous unit done with x; twiddles loaded
    for (uint32_t e = threadIdx.x; e < cnt * L; e += blockDim.x) {
      const uint32_t r = e / L, k = e % L;
      const float* row = tables + slot * slot_sz + (size_t)(pr0 + r) * 4 * nz;
      const uint32_t o = k <= nz ? k + nz - 1 : k - nz - 1;
      x[e] = make_float2(row[o], row[2 * nz + o]);
    }
    __syncthreads();
    fft_rows(x, tw_s, L, cnt);
    for (uint32_t e = threadIdx.x; e < cnt * F; e += blockDim.x) {
      const uint32_t r = e / F, f = e % F;
      const uint32_t pr = pr0 + r, rp = pr / nr, rs = pr % nr;
      float2 a, b;
      split_real_pair(x + r * L, L, f, a, b);
      float2* dst = shat + ((size_t)slot * F + f) * R * nr;
      dst[(rp * 2 + 0) * nr + rs] = a;
      dst[(rp * 2 + 1) * nr + rs] = b;
    }
  }
}

// ---- charge spectra -------------------------------------------------------------

// q^[p][f][rs] of every product column p < nrem: rings rs and rs + 1
// transformed together, zero-padded from nz to L.  Grid-stride over (column,
// group of ring pairs).
__global__ __launch_bounds__(kFftThreads) void k_fft_charges(const GemmState* __restrict__ gs,
                                                             const float* __restrict__ q,
                                                             float2* __restrict__ qhat,
                                                             const float2* __restrict__ tw,
                                                             uint32_t nr, uint32_t nz)
{
  __shared__ float2 x[kFftLdsComplex];
  __shared__ float2 tw_s[kFftLdsComplex / 2];
  const uint32_t L = 2 * nz, F = nz + 1, K = nr * nz;
  const uint32_t rows = kFftLdsComplex / L;
  const uint32_t pairs = nr / 2;
  const uint32_t groups = (pairs + rows - 1) / rows;
  const uint32_t units = gs->nrem * groups;
  load_twiddles(tw_s, tw, L);
  for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint32_t p = u / groups, pr0 = (u % groups) * rows;
    const uint32_t cnt = min(rows, pairs - pr0);
    __syncthreads();
    const float* qp = q + (size_t)p * K;
    for (uint32_t e = threadIdx.x; e < cnt * L; e += blockDim.x) {
      const uint32_t r = e / L, k = e % L, rs = 2 * (pr0 + r);
      x[e] = k < nz ? make_float2(qp[rs * nz + k], qp[(rs + 1) * nz + k]) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_rows(x, tw_s, L, cnt);
    for (uint32_t e = threadIdx.x; e < cnt * F; e += blockDim.x) {
      const uint32_t r = e / F, f = e % F, rs = 2 * (pr0 + r);
      float2 a, b;
      split_real_pair(x + r * L, L, f, a, b);
      float2* dst = qhat + ((size_t)p * F + f) * nr;
      dst[rs] = a;
      dst[rs + 1] = b;
    }
  }
}

// ---- the product per frequency --------------------------------------------------

// Unit u = tile * F + f: the columns of one beam tile (all of one slot, slot
// order as the direct product's tiles) at frequency f.  The complex product
// W^[p][f][row] = sum over rs of S^[slot][f][row][rs] q^[p][f][rs] as the
// real product of M' = 2 R rows (m' = 2 row + part: part 0 the real, 1 the
// imaginary part) by K' = 2 nr (k' = 2 rs + c: c 0 the real, 1 the imaginary
// part of q^) with A'[2 row][2 rs] = Re S^, A'[2 row][2 rs + 1] = -Im S^,
// A'[2 row + 1][2 rs] = Im S^, A'[2 row + 1][2 rs + 1] = Re S^.  The tile's
// charge spectra are staged in LDS; each wave takes 16-row tiles of M' in
// turn and reads S^ through the caches.
__global__ __launch_bounds__(kFftThreads) void k_fft_product(const GemmState* __restrict__ gs,
                                                             const float2* __restrict__ shat,
                                                             const float2* __restrict__ qhat,
                                                             float2* __restrict__ what,
                                                             uint32_t nr, uint32_t nz)
{
  __shared__ uint32_t s_start[kGemmSlots], s_count[kGemmSlots];
  __shared__ float b_lds[kProdNT * (kProdMaxK + 1)];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  if (t < (int)kGemmSlots) { s_start[t] = gs->slot_start[t]; s_count[t] = gs->slot_count[t]; }
  __syncthreads();
  uint32_t ntiles = 0;
  for (uint32_t s = 0; s < kGemmSlots; ++s) ntiles += (s_count[s] + kProdNT - 1) / kProdNT;
  const uint32_t F = nz + 1, R = 2 * (nr + 1), Kp = 2 * nr, Mp = 2 * R;
  const uint32_t pitch = Kp + 1;
  const uint32_t mtiles = (Mp + 15) / 16;
  const uint32_t units = ntiles * F;
  for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint32_t f = u % F;
    uint32_t tile = u / F, slot = 0, p0 = 0, nb = 0;
    for (uint32_t s = 0; s < kGemmSlots; ++s) {
      const uint32_t nt = (s_count[s] + kProdNT - 1) / kProdNT;
      if (tile < nt) {
        slot = s;
        p0 = s_start[s] + tile * kProdNT;
        nb = min((uint32_t)kProdNT, s_count[s] - tile * kProdNT);
        break;
      }
      tile -= nt;
    }
    __syncthreads();   // the previous unit is done with b_lds
    for (uint32_t e = t; e < kProdNT * Kp; e += kFftThreads) {
      const uint32_t j = e / Kp, kk = e % Kp;
      float v = 0.f;
      if (j < nb) {
        const float2 c = qhat[((size_t)(p0 + j) * F + f) * nr + kk / 2];
        v = (kk & 1) ? c.y : c.x;
      }
      b_lds[j * pitch + kk] = v;
    }
    __syncthreads();
    const float2* A = shat + ((size_t)slot * F + f) * R * nr;
    for (uint32_t mt = wave; mt < mtiles; mt += kFftThreads / 64) {
      f4 acc[kProdNT / 16];
#pragma unroll
      for (int g = 0; g < kProdNT / 16; ++g) acc[g] = f4{0.f, 0.f, 0.f, 0.f};
      const uint32_t m = mt * 16 + (lane & 15);     // this lane's row of A'
      const uint32_t row = m / 2, part = m & 1;
      const bool mrow = m < Mp;
      for (uint32_t k4 = 0; k4 < Kp; k4 += 4) {
        const uint32_t kk = k4 + (lane >> 4);
        float a = 0.f;
        if (mrow) {
          const float2 s = A[(size_t)row * nr + kk / 2];
          // part 0: (Re, -Im); part 1: (Im, Re) for c = 0, 1
          a = part == 0 ? ((kk & 1) ? -s.y : s.x) : ((kk & 1) ? s.x : s.y);
        }
#pragma unroll
        for (int g = 0; g < kProdNT / 16; ++g)
          acc[g] = mfma(a, b_lds[(g * 16 + (lane & 15)) * pitch + kk], acc[g]);
      }
      // D[m][n]: n = lane % 16, m = 4 (lane / 16) + i
#pragma unroll
      for (int g = 0; g < kProdNT / 16; ++g) {
        const uint32_t j = g * 16 + (lane & 15);
        if (j >= nb) continue;
        float* out = (float*)(what + ((size_t)(p0 + j) * F + f) * R);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t mo = mt * 16 + 4 * (lane >> 4) + i;
          if (mo < Mp) out[mo] = acc[g][i];   // m' = 2 row + part: interleaved complex
        }
      }
    }
  }
}

// ---- inverse transforms -----------------------------------------------------------

// w[p][(2 rp + comp) (nz + 1) + zp] for zp = 0 .. nz: the two components of
// rp transformed back together, z = IDFT(E_r^ + i E_z^) = conj(DFT(conj(.))) / L
// with the spectra extended by Hermitian symmetry.  Grid-stride over (column,
// group of radii).
__global__ __launch_bounds__(kFftThreads) void k_fft_inverse(const GemmState* __restrict__ gs,
                                                             const float2* __restrict__ what,
                                                             float* __restrict__ w,
                                                             const float2* __restrict__ tw,
                                                             uint32_t nr, uint32_t nz)
{
  __shared__ float2 x[kFftLdsComplex];
  __shared__ float2 tw_s[kFftLdsComplex / 2];
  const uint32_t L = 2 * nz, F = nz + 1, R = 2 * (nr + 1);
  const uint32_t M = R * (nz + 1);
  const uint32_t rows = kFftLdsComplex / L;
  const uint32_t groups = (nr + 1 + rows - 1) / rows;
  const uint32_t units = gs->nrem * groups;
  const float inv_l = 1.0f / (float)L;
  load_twiddles(tw_s, tw, L);
  for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint32_t p = u / groups, rp0 = (u % groups) * rows;
    const uint32_t cnt = min(rows, nr + 1 - rp0);
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < cnt * L; e += blockDim.x) {
      const uint32_t r = e / L, k = e % L, rp = rp0 + r;
      const uint32_t fk = k <= nz ? k : L - k;
      const float2* src = what + ((size_t)p * F + fk) * R;
      float2 a = src[rp * 2 + 0], b = src[rp * 2 + 1];
      if (k > nz) { a.y = -a.y; b.y = -b.y; }        // X[L - f] = conj X[f]
      // Z = A + i B, stored conjugated for the forward transform
      x[e] = make_float2(a.x - b.y, -(a.y + b.x));
    }
    __syncthreads();
    fft_rows(x, tw_s, L, cnt);
    for (uint32_t e = threadIdx.x; e < cnt * (nz + 1); e += blockDim.x) {
      const uint32_t r = e / (nz + 1), zp = e % (nz + 1), rp = rp0 + r;
      const float2 z = x[r * L + zp];                  // conj of the inverse, times L
      float* out = w + (size_t)p * M;
      out[(rp * 2 + 0) * (nz + 1) + zp] = z.x * inv_l;
      out[(rp * 2 + 1) * (nz + 1) + zp] = -z.y * inv_l;
    }
  }
}

// ---- launcher -----------------------------------------------------------------------

bool fft_product_supported(uint32_t nr, uint32_t nz)
{
  const uint32_t L = 2 * nz;
  return nr >= 2 && nr % 2 == 0 && (L & (L - 1)) == 0 && L >= 8 &&
         L <= (uint32_t)kFftLdsComplex && 2 * nr <= (uint32_t)kProdMaxK;
}

hipError_t launch_fft_product(const GemmState* gs, const float* tables, const float* q, float* w,
                              const FftBufs& fb, uint32_t nr, uint32_t nz, uint32_t cus,
                              hipStream_t s)
{
  hipLaunchKernelGGL(k_fft_sched, dim3(1), dim3(64), 0, s, gs, fb.spec_class, fb.list);
  hipLaunchKernelGGL(k_fft_tables, dim3(cus * 4), dim3(kFftThreads), 0, s, fb.list, tables,
                     fb.shat, fb.tw, nr, nz);
  hipLaunchKernelGGL(k_fft_charges, dim3(cus * 4), dim3(kFftThreads), 0, s, gs, q, fb.qhat,
                     fb.tw, nr, nz);
  hipLaunchKernelGGL(k_fft_product, dim3(cus * 4), dim3(kFftThreads), 0, s, gs, fb.shat,
                     fb.qhat, fb.what, nr, nz);
  hipLaunchKernelGGL(k_fft_inverse, dim3(cus * 4), dim3(kFftThreads), 0, s, gs, fb.what, w,
                     fb.tw, nr, nz);
  return hipGetLastError();
}

}  // namespace hps
