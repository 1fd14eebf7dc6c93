// hps_kernels.hip -- gfx950 kernels of the tracking engine.
//
// Kernel families (one HBM sweep each):
//   k_segment  SCHEFF kick of the previous kick point + every element map up
//              to the next kick point, fused in registers, + per-chunk beam
//              moments the next kick needs (reference: SimulatePartialDrift/
//              HalfQuad/RFGap kernels, ApplySpchKickKernel, and the
//              reductions of Beam::UpdateStatForSpaceCharge up to the x/y
//              moments); k_seg_final finishes them in one workgroup.
//   k_phase    per-chunk relative-phase moments; k_phase_final finishes them
//              and runs the Scheff::Start control logic on the device (no
//              host round trip per kick).
//   k_sc_build / k_field   the SCHEFF rebuild (Green's table and charge
//              deposit in one launch, then the field sum), run only when the
//              control step asks for a remesh (flag in device memory).
// Reductions follow one canonical order (see kChunk): lane sums of 4
// particles strided by 256, halving tree over 256 lanes, halving tree over
// the chunk partials zero-padded to a power of two.  The grid kernels write
// one partial per chunk; a one-workgroup kernel launched after them performs
// the cross-chunk tree (kernel order makes the partials visible, with no
// device-wide fence per workgroup).
#include "hps_internal.hpp"

namespace hps {

// hps_push_fast.hip builds this file a second time with HPS_FAST_PUSH
// defined and reciprocal division; that build's kernels and launchers live
// in hps::fast.
#ifdef HPS_FAST_PUSH
namespace fast {
#endif

namespace {

constexpr int kBlock = 256;

// Loads and stores through global-address-space pointers.  Pointers that a
// kernel reads from device memory (the per-simulation records of the
// batched launches) are generic to the compiler, which then emits flat
// accesses (64-bit VGPR addresses, waits on both memory counters); every
// buffer of the engine is global memory.
template <class T>
__device__ __forceinline__ T gld(const T* p, size_t i)
{
  return ((const __attribute__((address_space(1))) T*)p)[i];
}

template <class T>
__device__ __forceinline__ void gst(T* p, size_t i, T v)
{
  ((__attribute__((address_space(1))) T*)p)[i] = v;
}

// --- canonical reductions -------------------------------------------------

// Canonical halving tree over the 256 lane values of Q sums (levels 128 and
// 64 through LDS, 32..1 with wavefront shuffles: lane i adds lane i + o at
// every level, the same pairing as the serial definition), plus an integer
// count and a max carried alongside.  The results are valid in thread 0;
// red must hold Q*256 doubles, ured 256 words.
template <int Q>
__device__ __forceinline__ void block_reduce(double (&v)[Q], uint32_t& cnt, double& mx, double* red,
                                             uint32_t* ured, double* mred, int t)
{
#pragma unroll
  for (int q = 0; q < Q; ++q) red[q * kBlock + t] = v[q];
  ured[t] = cnt;
  mred[t] = mx;
  __syncthreads();
  if (t < 128) {
#pragma unroll
    for (int q = 0; q < Q; ++q) red[q * kBlock + t] = red[q * kBlock + t] + red[q * kBlock + t + 128];
    ured[t] += ured[t + 128];
    mred[t] = ::fmax(mred[t], mred[t + 128]);
  }
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = red[q * kBlock + t] + red[q * kBlock + t + 64];
    cnt = ured[t] + ured[t + 64];
    mx = ::fmax(mred[t], mred[t + 64]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
      for (int q = 0; q < Q; ++q) v[q] = v[q] + __shfl_down(v[q], o, 64);
      cnt += __shfl_down(cnt, o, 64);
      mx = ::fmax(mx, __shfl_down(mx, o, 64));
    }
  }
  __syncthreads();
}

// Largest of the 256 lane values (valid in thread 0).  The comparison
// a < v keeps a when v is NaN; a maximum does not depend on the order.
// buf holds kBlock / 64 doubles.
__device__ __forceinline__ double block_max(double v, double* buf, int t)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double u = __shfl_down(v, o, 64);
    v = v < u ? u : v;
  }
  if ((t & 63) == 0) buf[t >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (t == 0) {
    r = buf[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) r = r < buf[w] ? buf[w] : r;
  }
  __syncthreads();
  return r;
}

// Thread t's entries c = t + 256 k, k < K, of the cross-chunk trees below
// (mpad = 256 K, or K = 1 and mpad <= 256): the Q partial sums folded by the
// upper levels of the halving tree, the integer counts C[c] of c < nch (C
// may be null), and, when M is given, the largest M[c] of c < nch (-inf if
// none; a NaN entry is never taken).  Every load is issued before the first
// addition, so the K * (Q + 2) loads share one round trip.
template <int Q, int K>
__device__ __forceinline__ void fold_entries(const double* P, const uint32_t* C, const double* M,
                                             uint32_t mpad, uint32_t nch, int t, double (&v)[Q],
                                             uint32_t& cnt, double& mx)
{
  const bool in = K > 1 || (uint32_t)t < mpad;
  double r[Q][K];
  uint32_t n[K];
  double m[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t c = (uint32_t)t + kBlock * k;
#pragma unroll
    for (int q = 0; q < Q; ++q) r[q][k] = in ? gld(P, (size_t)q * mpad + c) : 0.0;
    n[k] = (C && in) ? gld(C, c) : 0u;
    m[k] = (M && in) ? gld(M, c) : -__builtin_huge_val();
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
#pragma unroll
    for (int h = K / 2; h >= 1; h /= 2) {
#pragma unroll
      for (int k = 0; k < h; ++k) r[q][k] = r[q][k] + r[q][k + h];
    }
    v[q] = r[q][0];
  }
  cnt = 0;
  mx = -__builtin_huge_val();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool live = (uint32_t)t + kBlock * k < nch;
    cnt += live ? n[k] : 0u;
    if (live && m[k] > mx) mx = m[k];
  }
}

// Cross-chunk trees of Q quantities stored P[q * mpad + c] (zero padded to
// mpad, a power of two) and the sum of the integer counts C[0, nch).  Thread
// t first folds its entries t + 256k in registers (the upper levels of the
// halving tree), then one block tree finishes.  Every thread gets the
// results; with M given, *own_max receives the largest M[c] over the
// thread's own entries c < nch.  mpad <= 4096 (16 entries per thread).
template <int Q>
__device__ void final_reduce(const double* P, const uint32_t* C, uint32_t mpad, uint32_t nch,
                             double* red, uint32_t* ured, double* mred, int t, double (&out)[Q],
                             uint32_t& cnt_out, const double* M = nullptr,
                             double* own_max = nullptr)
{
  double v[Q];
  uint32_t cnt;
  double own;
  switch (mpad > (uint32_t)kBlock ? mpad / kBlock : 1u) {
    case 1: fold_entries<Q, 1>(P, C, M, mpad, nch, t, v, cnt, own); break;
    case 2: fold_entries<Q, 2>(P, C, M, mpad, nch, t, v, cnt, own); break;
    case 4: fold_entries<Q, 4>(P, C, M, mpad, nch, t, v, cnt, own); break;
    case 8: fold_entries<Q, 8>(P, C, M, mpad, nch, t, v, cnt, own); break;
    default: fold_entries<Q, 16>(P, C, M, mpad, nch, t, v, cnt, own); break;
  }
  if (own_max) *own_max = own;
  double mx = 0.0;
  block_reduce<Q>(v, cnt, mx, red, ured, mred, t);
  __shared__ double bcast[kMaxQ];
  __shared__ uint32_t bcnt;
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) bcast[q] = v[q];
    bcnt = cnt;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) out[q] = bcast[q];
  cnt_out = bcnt;
  __syncthreads();
}

__device__ __forceinline__ bool is_good(uint32_t loss, uint32_t lloss)
{
  return (uint32_t)(loss + lloss) == 0u;
}

// --- SCHEFF kick (ApplySpchKickKernel, space_charge_kernel.cu:572-706) -----

struct KickConst {
  double phimc, xbar, ybar, epsq, gmg, dr, dzg, hl, c2, c3, xfac, yfac;
  double er_ratio, ez_ratio, cbgcoef, r_beta, phi_ref, pl;
  uint32_t nx, nz, adj;
};

// The kick constants of ApplySpchKickKernel (space_charge_kernel.cu:578-640):
// the parts fixed by the control step come from ScDev (kick_prepare in
// k_phase_final); the element length, the on-mesh count (k_deposit) and the
// phase references (which a host frequency step may change) are read here.
__device__ __forceinline__ KickConst kick_const(const ScDev* scg, const hps_beam_stats* stg,
                                                double r_length)
{
  // read through the constant address space (scalar loads: neither record
  // changes while k_segment runs)
  typedef const __attribute__((address_space(4))) ScDev CScDev;
  typedef const __attribute__((address_space(4))) hps_beam_stats CStats;
  CScDev* sc = (CScDev*)scg;
  CStats* st = (CStats*)stg;
  KickConst k;
  k.phimc = st->phi_avg_r;
  k.xbar = sc->k_xbar;
  k.ybar = sc->k_ybar;
  k.epsq = sc->k_epsq;
  k.gmg = sc->k_gmg;
  k.dr = sc->k_dr;
  k.dzg = sc->k_dzg;
  k.nx = sc->st.nx;
  k.nz = sc->st.nz;
  k.hl = sc->k_hl;
  k.c2 = sc->k_c2;
  k.c3 = r_length / sc->k_bgmg;
  k.xfac = sc->k_xfac;
  k.yfac = sc->k_yfac;
  k.er_ratio = sc->k_er_ratio;
  k.ez_ratio = sc->k_ez_ratio;
  k.cbgcoef = sc->k_c1 * k.c3 * PI * 0.5 * sc->st.num_on_mesh;
  k.r_beta = sc->beta;
  k.phi_ref = st->phi_avg_good;
  k.pl = sc->k_pl;
  k.adj = sc->adj_eff;
  return k;
}

// The control-step part of the kick constants, in the expressions of
// ApplySpchKickKernel; r_num_part, r_mass: the beam's size and mass.
__device__ __forceinline__ void kick_prepare(ScDev* sc, double r_beta, double r_freq,
    double r_current, double x_avg_good, double y_avg_good, double x_sig_good,
    double y_sig_good, double dr, double dz, uint32_t nz, double ratio_r, double ratio_z,
    double ratio_q, double ratio_g, uint32_t r_num_part, double r_mass)
{
  const double epsq = x_sig_good / y_sig_good;
  const double gmg = 1.0 / ::sqrt(1.0 - r_beta * r_beta);
  const double dzg = dz * gmg * ratio_z;
  const double wave_len = CLIGHT / r_freq;
  const double xfac = 2.0 / (epsq + 1.0);
  sc->k_xbar = x_avg_good;
  sc->k_ybar = y_avg_good;
  sc->k_epsq = epsq;
  sc->k_gmg = gmg;
  sc->k_dr = dr * ratio_r;
  sc->k_dzg = dzg;
  sc->k_hl = nz * 0.5 * dzg;
  sc->k_c1 = 5721.67 * r_current / (r_freq * 1e6 * (double)r_num_part * r_mass);
  sc->k_c2 = r_beta * gmg * wave_len / TWOPI;
  sc->k_pl = r_beta * gmg * wave_len;
  sc->k_bgmg = r_beta * gmg;
  sc->k_xfac = xfac;
  sc->k_yfac = epsq * xfac;
  sc->k_er_ratio = ratio_q / (ratio_r * ratio_z * ratio_g * ratio_g);
  sc->k_ez_ratio = ratio_q / (ratio_z * ratio_z);
}

__device__ __forceinline__ void kick(Particle& p, const KickConst& k, const double* tbl,
                                     double r_mass)
{
  const double phi = relative_phi(p.phi, k.phi_ref);   // phi_r of UpdateRelativePhi(true)
  const double gm = p.w / r_mass + 1.0;
  const double bgz = psqrt((gm - 1.0) * (gm + 1.0));
  double bgx = bgz * p.xp;
  double bgy = bgz * p.yp;
  double bgzstar = k.gmg * (bgz - k.r_beta * gm);
  const double xxbar = p.x - k.xbar;
  const double yybar = p.y - k.ybar;
  double r = psqrt(xxbar * xxbar / k.epsq + yybar * yybar * k.epsq);
  if (r == 0.0) r = 1.0e-8;
  const double xovr = xxbar * k.xfac / r;
  const double yovr = yybar * k.yfac / r;
  double ddz = phi - k.phimc;
  for (int guard = 0; (ddz > PI || ddz < -PI) && guard < 4096; ++guard) {
    if (ddz > 0.0) ddz -= TWOPI;
    else ddz += TWOPI;
  }
  double z = -k.c2 * ddz;
  double cbgr, cbgzs;
  if (r > k.nx * k.dr || __builtin_fabs(z) > k.hl) {
    z = -k.c2 * (phi - k.phimc);
    const double d = psqrt(z * z + r * r);
    const double invd3 = 1.0 / (d * d * d);
    double rod3 = r * invd3;
    double zod3 = z * invd3;
    // point charges of the neighbouring bunches (space_charge_kernel.cu:650-662)
    for (uint32_t nb = 1; nb <= k.adj; ++nb) {
      int xi = (int)nb;
      for (int lr = 0; lr < 2; ++lr) {
        const double sz = z + xi * k.pl;
        const double dd = psqrt(sz * sz + r * r);
        const double iv = 1.0 / (dd * dd * dd);
        rod3 += r * iv;
        zod3 += sz * iv;
        xi = -xi;
      }
    }
    cbgr = k.cbgcoef * rod3;
    cbgzs = k.cbgcoef * zod3;
  } else {
    const uint32_t i = (uint32_t)(1 + to_int(r / k.dr));
    const double a = r / k.dr - (i - 1.0);
    const double b = 1.0 - a;
    const uint32_t j = (uint32_t)(1 + to_int((z + k.hl) / k.dzg));
    const double c = (z + k.hl) / k.dzg - (j - 1.0);
    const double d = 1.0 - c;
    const uint32_t tbl_sz = (k.nx + 1) * (k.nz + 1);
    uint32_t l = (j - 1) * (k.nx + 1) + i;
    uint32_t m = l + (k.nx + 1);
    if (m >= tbl_sz) m = tbl_sz - 1;
    if (l >= tbl_sz) l = tbl_sz - 1;
    const double tl0 = gld(tbl, 2 * l), tl1 = gld(tbl, 2 * l + 1);
    const double tk0 = gld(tbl, 2 * (l - 1)), tk1 = gld(tbl, 2 * (l - 1) + 1);
    const double tm0 = gld(tbl, 2 * m), tm1 = gld(tbl, 2 * m + 1);
    const double tn0 = gld(tbl, 2 * (m - 1)), tn1 = gld(tbl, 2 * (m - 1) + 1);
    cbgr = k.er_ratio * k.c3 * (d * (a * tl0 + b * tk0) + c * (a * tm0 + b * tn0));
    cbgzs = k.ez_ratio * k.c3 * (d * (a * tl1 + b * tk1) + c * (a * tm1 + b * tn1));
  }
  bgx += cbgr * xovr;
  bgy += cbgr * yovr;
  bgzstar += cbgzs;
  const double gstar = 1.0 + 0.5 * bgzstar * bgzstar;
  const double bgzf = k.gmg * (bgzstar + k.r_beta * gstar);
  if (bgzf <= 0.0) {
    p.loss = 9999;
  } else {
    p.xp = bgx / bgzf;
    p.yp = bgy / bgzf;
    p.w *= 1.0 + (gm + 1.0) * (bgzf - bgz) / (gm * bgz);
  }
}

// --- k_segment ---------------------------------------------------------------

// Per-chunk moments of the SCHEFF statistics (quantity order in partials):
//   0 phi, 1 x, 2 x^2, 3 y, 4 y^2, 5 w   over (loss + lloss) == 0
// plus the good count and the largest phi among particles not yet flagged
// as longitudinally lost (the chunks a new lloss flag can appear in).
constexpr int kScQ = 6;

struct RedSmem {
  double red[kMaxQ * kBlock];
  double mred[kBlock];
  uint32_t ured[kBlock];
};

// Sets the longitudinal-loss flag of every particle of chunk c beyond the
// threshold thr and recomputes the chunk's statistics partials with the new
// flags, in one pass over the chunk, on one wavefront: lane L stands for the
// canonical lanes L, L+64, L+128 and L+192 of k_segment's 256-lane order
// (each the same sequential sum over j), then folds them as levels 128 and 64
// of the halving tree before the shuffle levels 32..1 -- the same tree as
// block_reduce, so the partials are identical.  The waves of a workgroup
// rescan different chunks at the same time.
__device__ void chunk_flag_stats_wave(const BeamArrays& b, uint32_t c, double thr, int lane,
                                      const Partials& pr)
{
  constexpr int kSlots = kBlock / 64;
  double acc[kSlots][kScQ];
  uint32_t cnt = 0;
  double mx = -__builtin_huge_val();
#pragma unroll
  for (int sl = 0; sl < kSlots; ++sl) {
#pragma unroll
    for (int q = 0; q < kScQ; ++q) acc[sl][q] = 0.0;
#pragma unroll
    for (int j = 0; j < (int)kPerLane; ++j) {
      const uint32_t i = c * kChunk + j * kBlock + sl * 64 + lane;
      double phi = 0.0, x = 0.0, y = 0.0, w = 0.0;
      uint32_t loss = 1, lloss = 0;
      if (i < b.n) {
        phi = b.phi[i]; x = b.x[i]; y = b.y[i]; w = b.w[i]; loss = b.loss[i]; lloss = b.lloss[i];
        if (phi > thr) {
          if (lloss != 1) b.lloss[i] = 1;
          lloss = 1;
        }
        if (lloss == 0) mx = ::fmax(mx, phi);
      }
      const bool g = (i < b.n) && is_good(loss, lloss);
      const double vphi = g ? phi : 0.0, vx = g ? x : 0.0, vy = g ? y : 0.0, vw = g ? w : 0.0;
      acc[sl][0] = acc[sl][0] + vphi;
      acc[sl][1] = acc[sl][1] + vx;
      acc[sl][2] = acc[sl][2] + vx * vx;
      acc[sl][3] = acc[sl][3] + vy;
      acc[sl][4] = acc[sl][4] + vy * vy;
      acc[sl][5] = acc[sl][5] + vw;
      cnt += g ? 1u : 0u;
    }
  }
  static_assert(kSlots == 4, "the fold below is written for 256 canonical lanes");
  double v[kScQ];
#pragma unroll
  for (int q = 0; q < kScQ; ++q) {
    const double l0 = acc[0][q] + acc[2][q];   // level 128: lane L += lane L+128
    const double l1 = acc[1][q] + acc[3][q];   //            lane L+64 += lane L+192
    v[q] = l0 + l1;                            // level 64
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
    for (int q = 0; q < kScQ; ++q) v[q] = v[q] + __shfl_down(v[q], o, 64);
    cnt += __shfl_down(cnt, o, 64);
    mx = ::fmax(mx, __shfl_down(mx, o, 64));
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kScQ; ++q) pr.sums[q * pr.mpad + c] = v[q];
    pr.cnt[c] = cnt;
    pr.pmax[c] = mx;
  }
}

// Particles of one lane in flight at a time in k_segment, per map class.
#ifndef HPS_SEG_GROUP_GAP
#define HPS_SEG_GROUP_GAP 1
#endif
#ifndef HPS_SEG_GROUP_QUAD
#define HPS_SEG_GROUP_QUAD 1
#endif
#ifndef HPS_SEG_GROUP_LIGHT
#define HPS_SEG_GROUP_LIGHT 1
#endif
constexpr int kSegGroupGap = HPS_SEG_GROUP_GAP;
constexpr int kSegGroupQuad = HPS_SEG_GROUP_QUAD;
constexpr int kSegGroupLight = HPS_SEG_GROUP_LIGHT;

// Particle i of the beam, or an excluded placeholder (loss = 1) past the end.
__device__ __forceinline__ Particle load_particle(const BeamArrays& b, uint32_t i)
{
  Particle p;
  p.loss = 1; p.lloss = 0;
  p.x = p.xp = p.y = p.yp = p.phi = p.w = 0.0;
  if (i < b.n) {
    p.loss = gld(b.loss, i);
    p.lloss = gld(b.lloss, i);
    p.x = gld(b.x, i); p.xp = gld(b.xp, i); p.y = gld(b.y, i); p.yp = gld(b.yp, i);
    p.phi = gld(b.phi, i); p.w = gld(b.w, i);
  }
  return p;
}

// Writes back a pushed particle that was alive on entry (the map kernels write
// only those); a new loss code is written with it.
__device__ __forceinline__ void store_particle(const BeamArrays& b, uint32_t i, const Particle& q)
{
  gst(b.x, i, q.x); gst(b.xp, i, q.xp); gst(b.y, i, q.y); gst(b.yp, i, q.yp);
  gst(b.phi, i, q.phi); gst(b.w, i, q.w);
  if (q.loss != 0) gst(b.loss, i, q.loss);
}

// One workgroup per 1024-particle chunk; each lane carries its four
// particles (i = chunk*1024 + j*256 + lane) through the kick, the maps and
// the store in groups of kGroup, and accumulates the chunk statistics in the
// canonical lane order (j = 0..3, from +0.0):
//   STAT_SC        0 phi, 1 x, 2 x^2, 3 y, 4 y^2, 5 w over (loss + lloss) == 0,
//                  the good count and the largest phi among particles not yet
//                  flagged as longitudinally lost;
//   STAT_APERTURE  x and y over loss == 0 and the count of loss != 0
//                  (UpdateLoss + UpdateAvgXY, simulation_engine.cu:311-312).
// The group size trades registers (waves per SIMD) against independent
// chains per wave.
template <int kClass, int kGroup>
__device__ __forceinline__ void segment_body(const SegArgs& a, uint32_t c, RedSmem& sm)
{
  static_assert(kPerLane % kGroup == 0, "group must divide the particles per lane");
  const int t = threadIdx.x;
  const BeamArrays& b = a.b;
  const double xa = a.st->x_avg, ya = a.st->y_avg;
  // per-lane accumulators in LDS rows (lane t owns column t): 0..5 sums,
  // 6 largest r^2, ured count, mred max phi.  Held in LDS rather than
  // registers across the particle loop, they leave the registers to the
  // maps (14 fewer VGPRs in the RF-gap class).
  double* const acc = sm.red + t;
#pragma unroll
  for (int q = 0; q < kScQ + 1; ++q) acc[q * kBlock] = 0.0;
  sm.ured[t] = 0;
  sm.mred[t] = -__builtin_huge_val();
#pragma unroll 1
  for (int j0 = 0; j0 < (int)kPerLane; j0 += kGroup) {
    Particle p[kGroup];
    bool alive_in[kGroup];
#pragma unroll
    for (int g = 0; g < kGroup; ++g) {
      const uint32_t i = c * kChunk + (j0 + g) * kBlock + t;
      p[g] = load_particle(b, i);
      alive_in[g] = i < b.n && p[g].loss == 0;
    }
    if (a.kick) {
      // the kick constants are read again for every particle group (scalar
      // loads from the scalar cache) rather than held in SGPRs across the maps
      const ScDev* scp = a.sc;
      const hps_beam_stats* stp = a.st;
      asm volatile("" : "+s"(scp), "+s"(stp));
      const KickConst k = kick_const(scp, stp, a.kick_len);
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        const uint32_t i = c * kChunk + (j0 + g) * kBlock + t;
        if (alive_in[g] && p[g].lloss == 0 && (i & a.kick_mask) != 0)
          kick(p[g], k, a.fld2, b.mass);
      }
    }
    for (int o = a.op_begin; o < a.op_end; ++o) {
      const OpConst& op = ((const OpConst*)a.ops)[o];
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        const uint32_t i = c * kChunk + (j0 + g) * kBlock + t;
        if (p[g].loss == 0) apply_op<kClass>(op, p[g], i, b.mass, b.charge, xa, ya);
      }
    }
#pragma unroll
    for (int g = 0; g < kGroup; ++g) {
      const uint32_t i = c * kChunk + (j0 + g) * kBlock + t;
      const bool in = i < b.n;
      const Particle& q = p[g];
      if (alive_in[g]) store_particle(b, i, q);
      if (a.stats == STAT_SC) {
        if (in && q.lloss == 0) sm.mred[t] = ::fmax(sm.mred[t], q.phi);
        const bool gd = in && is_good(q.loss, q.lloss);
        const double vphi = gd ? q.phi : 0.0, vx = gd ? q.x : 0.0, vy = gd ? q.y : 0.0,
                     vw = gd ? q.w : 0.0;
        acc[0] = acc[0] + vphi;
        acc[kBlock] = acc[kBlock] + vx;
        acc[2 * kBlock] = acc[2 * kBlock] + vx * vx;
        acc[3 * kBlock] = acc[3 * kBlock] + vy;
        acc[4 * kBlock] = acc[4 * kBlock] + vy * vy;
        acc[5 * kBlock] = acc[5 * kBlock] + vw;
        sm.ured[t] += gd ? 1u : 0u;
        if (a.rmax && in && q.loss == 0) {
          const double tx = q.x - xa, ty = q.y - ya;
          const double r2 = tx * tx + ty * ty;
          const double r2m = acc[6 * kBlock];
          acc[6 * kBlock] = r2m < r2 ? r2 : r2m;
        }
      } else if (a.stats == STAT_APERTURE) {
        const bool ok = in && q.loss == 0;
        acc[0] = acc[0] + (ok ? q.x : 0.0);
        acc[kBlock] = acc[kBlock] + (ok ? q.y : 0.0);
        sm.ured[t] += (in && q.loss != 0) ? 1u : 0u;
      }
    }
  }
  const Partials& pr = a.pr;
  if (a.stats == STAT_SC) {
    double v[kScQ];
#pragma unroll
    for (int q = 0; q < kScQ; ++q) v[q] = acc[q * kBlock];
    const double r2m = acc[6 * kBlock];
    uint32_t cnt = sm.ured[t];
    double mx = sm.mred[t];
    block_reduce<kScQ>(v, cnt, mx, sm.red, sm.ured, sm.mred, t);
    if (t == 0) {
#pragma unroll
      for (int q = 0; q < kScQ; ++q) pr.sums[q * pr.mpad + c] = v[q];
      pr.cnt[c] = cnt;
      pr.pmax[c] = mx;
    }
    if (a.rmax) {
      const double m = block_max(r2m, sm.mred, t);
      if (t == 0) pr.rmax2[c] = m;
    }
  } else if (a.stats == STAT_APERTURE) {
    double ap[2] = {acc[0], acc[kBlock]};
    uint32_t cnt = sm.ured[t];
    double m0 = 0.0;
    block_reduce<2>(ap, cnt, m0, sm.red, sm.ured, sm.mred, t);
    if (t == 0) { pr.sums[c] = ap[0]; pr.sums[pr.mpad + c] = ap[1]; pr.cnt[c] = cnt; }
  }
}

#ifdef HPS_FAST_PUSH
// The batched push of the grouped solve (this file as built by
// hps_push_fast.hip): the same push, moments and chunk partials as
// segment_body, but only over the chunk's live particles.  Particles lost
// before the segment are never pushed.  Scattered at random among the live
// ones, they would otherwise keep their lanes of every wavefront busy (on
// the 64-beam stress batch about half the particles are lost, on average,
// over the lattice).  The workgroup first reads the loss flags of its 1,024
// particles.  A lost particle adds to the loss count (aperture statistics)
// or, when not flagged longitudinally, offers its unchanged phase to the
// largest-phase test.  The live ones are listed in particle order in LDS
// (ballots per 64-particle group, an exclusive scan over the 16 groups),
// and thread t pushes list entries t, t + 256, ...  Each sum then adds its
// terms in another order than the canonical one; that is within this
// build's error bounds, and the counts and maxima are the same.
struct LiveSmem {
  uint16_t list[kChunk];     // chunk offsets (j * 256 + lane) of the live particles
  uint32_t gcnt[kPerLane * (kBlock / 64)];
};

template <int kClass>
__device__ __forceinline__ void segment_body_live(const SegArgs& a, uint32_t c, RedSmem& sm,
                                                  LiveSmem& ls)
{
  static_assert(kBlock == 256 && kPerLane == 4, "16 groups of 64 particles per chunk");
  const int t = threadIdx.x, wave = t >> 6;
  const BeamArrays& b = a.b;
  const double xa = a.st->x_avg, ya = a.st->y_avg;
  double* const acc = sm.red + t;
#pragma unroll
  for (int q = 0; q < kScQ + 1; ++q) acc[q * kBlock] = 0.0;
  // flags of the lane's four particles; lost ones' contributions
  uint32_t nlost = 0;
  double mlost = -__builtin_huge_val();
  bool live[kPerLane];
#pragma unroll
  for (int j = 0; j < (int)kPerLane; ++j) {
    const uint32_t i = c * kChunk + j * kBlock + t;
    live[j] = false;
    if (i < b.n) {
      if (gld(b.loss, i) == 0) {
        live[j] = true;
      } else {
        ++nlost;
        if (a.stats == STAT_SC && gld(b.lloss, i) == 0) mlost = ::fmax(mlost, gld(b.phi, i));
      }
    }
  }
  sm.ured[t] = a.stats == STAT_APERTURE ? nlost : 0u;
  sm.mred[t] = mlost;
  uint64_t bal[kPerLane];
#pragma unroll
  for (int j = 0; j < (int)kPerLane; ++j) {
    bal[j] = __ballot(live[j]);
    if ((t & 63) == 0) ls.gcnt[j * 4 + wave] = (uint32_t)__popcll(bal[j]);
  }
  __syncthreads();
  uint32_t nlive = 0, base[kPerLane] = {0, 0, 0, 0};
#pragma unroll
  for (int g = 0; g < (int)(kPerLane * 4); ++g) {
    const uint32_t n = ls.gcnt[g];
#pragma unroll
    for (int j = 0; j < (int)kPerLane; ++j)
      if (g < j * 4 + wave) base[j] += n;
    nlive += n;
  }
#pragma unroll
  for (int j = 0; j < (int)kPerLane; ++j) {
    if (live[j]) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(bal[j] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal[j], 0u));
      ls.list[base[j] + below] = (uint16_t)(j * kBlock + t);
    }
  }
  __syncthreads();
#pragma unroll 1
  for (uint32_t sl = (uint32_t)t; sl < nlive; sl += kBlock) {
    const uint32_t i = c * kChunk + ls.list[sl];
    Particle p = load_particle(b, i);
    if (a.kick) {
      const ScDev* scp = a.sc;
      const hps_beam_stats* stp = a.st;
      asm volatile("" : "+s"(scp), "+s"(stp));
      const KickConst k = kick_const(scp, stp, a.kick_len);
      if (p.lloss == 0 && (i & a.kick_mask) != 0) kick(p, k, a.fld2, b.mass);
    }
    for (int o = a.op_begin; o < a.op_end; ++o) {
      const OpConst& op = ((const OpConst*)a.ops)[o];
      if (p.loss == 0) apply_op<kClass>(op, p, i, b.mass, b.charge, xa, ya);
    }
    store_particle(b, i, p);
    if (a.stats == STAT_SC) {
      if (p.lloss == 0) sm.mred[t] = ::fmax(sm.mred[t], p.phi);
      const bool gd = is_good(p.loss, p.lloss);
      const double vphi = gd ? p.phi : 0.0, vx = gd ? p.x : 0.0, vy = gd ? p.y : 0.0,
                   vw = gd ? p.w : 0.0;
      acc[0] = acc[0] + vphi;
      acc[kBlock] = acc[kBlock] + vx;
      acc[2 * kBlock] = acc[2 * kBlock] + vx * vx;
      acc[3 * kBlock] = acc[3 * kBlock] + vy;
      acc[4 * kBlock] = acc[4 * kBlock] + vy * vy;
      acc[5 * kBlock] = acc[5 * kBlock] + vw;
      sm.ured[t] += gd ? 1u : 0u;
      if (a.rmax && p.loss == 0) {
        const double tx = p.x - xa, ty = p.y - ya;
        const double r2 = tx * tx + ty * ty;
        const double r2m = acc[6 * kBlock];
        acc[6 * kBlock] = r2m < r2 ? r2 : r2m;
      }
    } else if (a.stats == STAT_APERTURE) {
      const bool ok = p.loss == 0;
      acc[0] = acc[0] + (ok ? p.x : 0.0);
      acc[kBlock] = acc[kBlock] + (ok ? p.y : 0.0);
      sm.ured[t] += ok ? 0u : 1u;
    }
  }
  const Partials& pr = a.pr;
  if (a.stats == STAT_SC) {
    double v[kScQ];
#pragma unroll
    for (int q = 0; q < kScQ; ++q) v[q] = acc[q * kBlock];
    const double r2m = acc[6 * kBlock];
    uint32_t cnt = sm.ured[t];
    double mx = sm.mred[t];
    block_reduce<kScQ>(v, cnt, mx, sm.red, sm.ured, sm.mred, t);
    if (t == 0) {
#pragma unroll
      for (int q = 0; q < kScQ; ++q) pr.sums[q * pr.mpad + c] = v[q];
      pr.cnt[c] = cnt;
      pr.pmax[c] = mx;
    }
    if (a.rmax) {
      const double m = block_max(r2m, sm.mred, t);
      if (t == 0) pr.rmax2[c] = m;
    }
  } else if (a.stats == STAT_APERTURE) {
    double ap[2] = {acc[0], acc[kBlock]};
    uint32_t cnt = sm.ured[t];
    double m0 = 0.0;
    block_reduce<2>(ap, cnt, m0, sm.red, sm.ured, sm.mred, t);
    if (t == 0) { pr.sums[c] = ap[0]; pr.sums[pr.mpad + c] = ap[1]; pr.cnt[c] = cnt; }
  }
}
#endif

template <int kClass, int kGroup>
__global__ __launch_bounds__(kBlock) void k_segment(SegArgs a)
{
  __shared__ RedSmem sm;
  segment_body<kClass, kGroup>(a, blockIdx.x, sm);
}

// The same over the chunks of several simulations: map[block] packs the
// simulation (high bits) and its chunk (low kBatchLocalBits bits).
// A batched launch finds its records at recs + *roff bytes: 0 for an eager
// launch, the replayed lattice period's offset inside a period graph.
template <class T>
__device__ __forceinline__ const T* period_recs(const T* recs, const uint64_t* roff)
{
  return (const T*)((const char*)recs + roff[0]);
}

template <int kClass, int kGroup>
__global__ __launch_bounds__(kBlock) void k_segment_batch(const SegArgs* __restrict__ recs,
                                                          const uint32_t* __restrict__ map,
                                                          const uint64_t* __restrict__ roff)
{
  __shared__ RedSmem sm;
  const uint32_t e = map[blockIdx.x];
  const SegArgs& a = period_recs(recs, roff)[e >> kBatchLocalBits];
#ifdef HPS_FAST_PUSH
  __shared__ LiveSmem ls;
  segment_body_live<kClass>(a, e & kBatchLocalMask, sm, ls);
#else
  segment_body<kClass, kGroup>(a, e & kBatchLocalMask, sm);
#endif
}

// The same push for beams too small to fill the GPU with k_segment's 256-thread
// workgroups: one chunk per 1024-thread workgroup, one particle per thread.
// Thread j*256 + l pushes particle i = chunk*1024 + j*256 + l (the particle
// canonical lane l takes as its j-th) and stores its statistics
// contributions in LDS; lane l (< 256) then adds its four in the canonical
// order j = 0..3 from +0.0, and the 256 lane sums go through block_reduce's
// halving tree -- the chunk partials are k_segment's bit for bit.
constexpr uint32_t kWideBlock = kChunk;
// at most this many chunks (k_segment would run at most two wavefronts per
// SIMD; measured faster up to 512 chunks, slower at 1024)
constexpr uint32_t kWideMaxChunks = 512;
// Batches of at least this many chunks run the phase pass one chunk per
// wavefront (MI355X, 262,144 chunks of the 64-beam stress batch: phase pass
// 189 -> 164 ms per step); a lone beam of 1,024 chunks in that form has one
// wave per SIMD and runs slower (7.3 -> 13.5 ms per step), so lone beams
// keep a workgroup per chunk.
constexpr uint32_t kPhaseWaveMinChunks = 8192;
struct WideSmem {
  double q[kScQ][kWideBlock];   // contributions, then lane sums in [q][0, 256)
  double mx[kWideBlock];        // phi of particles not longitudinally flagged, else -inf
  double r2[kWideBlock];        // r^2 about (x_avg, y_avg), loss == 0
  uint32_t cnt[kWideBlock];     // good (STAT_SC) or lost (STAT_APERTURE) flags
};

template <int kClass>
__device__ __forceinline__ void segment_body_wide(const SegArgs& a, uint32_t c, WideSmem& w)
{
  const int t = threadIdx.x;
  const BeamArrays& b = a.b;
  const uint32_t i = c * kChunk + (uint32_t)t;
  Particle p = load_particle(b, i);
  const bool alive_in = i < b.n && p.loss == 0;
  if (a.kick && alive_in && p.lloss == 0 && (i & a.kick_mask) != 0) {
    const KickConst k = kick_const(a.sc, a.st, a.kick_len);
    kick(p, k, a.fld2, b.mass);
  }
  const double xa = a.st->x_avg, ya = a.st->y_avg;
  for (int o = a.op_begin; o < a.op_end; ++o) {
    const OpConst& op = ((const OpConst*)a.ops)[o];
    if (p.loss == 0) apply_op<kClass>(op, p, i, b.mass, b.charge, xa, ya);
  }
  const bool in = i < b.n;
  if (alive_in) store_particle(b, i, p);
  if (a.stats == STAT_NONE) return;
  const Partials& pr = a.pr;
  if (a.stats == STAT_SC) {
    const bool gd = in && is_good(p.loss, p.lloss);
    const double vx = gd ? p.x : 0.0, vy = gd ? p.y : 0.0;
    w.q[0][t] = gd ? p.phi : 0.0;
    w.q[1][t] = vx;
    w.q[2][t] = vx * vx;
    w.q[3][t] = vy;
    w.q[4][t] = vy * vy;
    w.q[5][t] = gd ? p.w : 0.0;
    w.cnt[t] = gd ? 1u : 0u;
    w.mx[t] = (in && p.lloss == 0) ? p.phi : -__builtin_huge_val();
    double r2 = 0.0;
    if (a.rmax && in && p.loss == 0) {
      // MaxR2PhiReduceKernel's first pass (beam_kernel.cu:826-843)
      const double tx = p.x - xa, ty = p.y - ya;
      r2 = tx * tx + ty * ty;
    }
    w.r2[t] = r2;
  } else {
    const bool ok = in && p.loss == 0;
    w.q[0][t] = ok ? p.x : 0.0;
    w.q[1][t] = ok ? p.y : 0.0;
    w.cnt[t] = (in && p.loss != 0) ? 1u : 0u;
    w.mx[t] = -__builtin_huge_val();
    w.r2[t] = 0.0;
  }
  __syncthreads();
  const int nq = a.stats == STAT_SC ? kScQ : 2;
  // lane sums in the canonical order over j, kept in [q][0, 256)
  if (t < (int)kBlock) {
    for (int q = 0; q < nq; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < (int)kPerLane; ++j) acc = acc + w.q[q][j * kBlock + t];
      w.q[q][t] = acc;
    }
    uint32_t n = 0;
    double m = -__builtin_huge_val(), r = 0.0;
#pragma unroll
    for (int j = 0; j < (int)kPerLane; ++j) {
      n += w.cnt[j * kBlock + t];
      m = ::fmax(m, w.mx[j * kBlock + t]);
      const double u = w.r2[j * kBlock + t];
      r = r < u ? u : r;
    }
    w.cnt[t] = n;
    w.mx[t] = m;
    w.r2[t] = r;
  }
  __syncthreads();
  // block_reduce's levels over the 256 lane sums
  if (t < 128) {
    for (int q = 0; q < nq; ++q) w.q[q][t] = w.q[q][t] + w.q[q][t + 128];
    w.cnt[t] += w.cnt[t + 128];
    w.mx[t] = ::fmax(w.mx[t], w.mx[t + 128]);
    const double u = w.r2[t + 128];
    w.r2[t] = w.r2[t] < u ? u : w.r2[t];
  }
  __syncthreads();
  if (t < 64) {
    double v[kScQ];
    for (int q = 0; q < kScQ; ++q) v[q] = q < nq ? w.q[q][t] + w.q[q][t + 64] : 0.0;
    uint32_t n = w.cnt[t] + w.cnt[t + 64];
    double m = ::fmax(w.mx[t], w.mx[t + 64]);
    double r = w.r2[t];
    {
      const double u = w.r2[t + 64];
      r = r < u ? u : r;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
      for (int q = 0; q < kScQ; ++q) v[q] = v[q] + __shfl_down(v[q], o, 64);
      n += __shfl_down(n, o, 64);
      m = ::fmax(m, __shfl_down(m, o, 64));
      const double u = __shfl_down(r, o, 64);
      r = r < u ? u : r;
    }
    if (t == 0) {
      if (a.stats == STAT_SC) {
#pragma unroll
        for (int q = 0; q < kScQ; ++q) pr.sums[q * pr.mpad + c] = v[q];
        pr.cnt[c] = n;
        pr.pmax[c] = m;
        if (a.rmax) pr.rmax2[c] = r;
      } else {
        pr.sums[c] = v[0];
        pr.sums[pr.mpad + c] = v[1];
        pr.cnt[c] = n;
      }
    }
  }
}

template <int kClass>
__global__ __launch_bounds__(kWideBlock) void k_segment_wide(SegArgs a)
{
  extern __shared__ double wide_lds[];
  segment_body_wide<kClass>(a, blockIdx.x, *reinterpret_cast<WideSmem*>(wide_lds));
}

// Cross-chunk step of the k_segment statistics in one workgroup, launched
// after k_segment so that kernel order, not a device-wide fence per
// workgroup, makes the partials visible: the aperture loss count and
// centroid (UpdateLoss + UpdateAvgXY), or the longitudinal-loss update and
// good moments of the space-charge statistics.
__device__ __forceinline__ void seg_final_body(const SegArgs& a)
{
  __shared__ RedSmem sm;
  __shared__ uint32_t rescan_list[64];
  __shared__ uint32_t rescan_n;
  const int t = threadIdx.x;
  const BeamArrays& b = a.b;
  const Partials& pr = a.pr;
  if (a.stats == STAT_APERTURE) {
    double s[2];
    uint32_t nloss;
    final_reduce<2>(pr.sums, pr.cnt, pr.mpad, pr.nch, sm.red, sm.ured, sm.mred, t, s, nloss);
    if (t == 0) {
      hps_beam_stats* st = a.st;
      st->num_loss = nloss;
      const double surv = (double)(uint32_t)(b.n - nloss);
      st->x_avg = s[0] / surv;
      st->y_avg = s[1] / surv;
    }
    return;
  }

  // STAT_SC: the longitudinal-loss update of Beam::UpdateLongitudinalLoss
  // (beam.cpp:776-782), then the good moments of beam.cpp:1008-1014.
  hps_beam_stats* st = a.st;
  const uint32_t num_loss = st->num_loss;
  // the chunk maxima are loaded together with the partials (one round trip)
  double s[kScQ];
  uint32_t g1;
  double pm;
  final_reduce<kScQ>(pr.sums, pr.cnt, pr.mpad, pr.nch, sm.red, sm.ured, sm.mred, t, s, g1,
                     pr.pmax, &pm);
  const double avg1 = g1 > 0 ? s[0] / (double)g1 : -99999;
  const double thr = avg1 + TWOPI + PI;
  // Chunks where a not-yet-flagged particle is beyond the threshold get their
  // lloss flags set and their partials recomputed (rare: once per particle).
  if (t == 0) rescan_n = 0;
  __syncthreads();
  if (pm > thr) atomicAdd(&rescan_n, 1u);
  __syncthreads();
  bool any = false;
  for (uint32_t it = 0; rescan_n != 0 && it < pr.nch / 64 + 2; ++it) {
    __syncthreads();
    if (t == 0) rescan_n = 0;
    __syncthreads();
    for (uint32_t cc = t; cc < pr.nch; cc += kBlock) {
      if (pr.pmax[cc] > thr) {
        const uint32_t slot = atomicAdd(&rescan_n, 1u);
        if (slot < 64) rescan_list[slot] = cc;
      }
    }
    __syncthreads();
    const uint32_t nfound = rescan_n;
    if (nfound == 0) break;
    any = true;
    const uint32_t take = nfound < 64 ? nfound : 64;
    for (uint32_t r = (uint32_t)t / 64; r < take; r += kBlock / 64)
      chunk_flag_stats_wave(b, rescan_list[r], thr, t & 63, pr);
    __syncthreads();
  }
  uint32_t g2 = g1;
  if (any)
    final_reduce<kScQ>(pr.sums, pr.cnt, pr.mpad, pr.nch, sm.red, sm.ured, sm.mred, t, s, g2);
  if (t == 0) {
    st->num_good = g2;
    st->phi_avg_good = g2 > 0 ? s[0] / (double)g2 : -99999;
    const double surv = (double)(uint32_t)(b.n - num_loss);
    const double xavg = s[1] / surv, x2avg = s[2] / surv, yavg = s[3] / surv, y2avg = s[4] / surv;
    st->x_avg_good = xavg;
    st->x_sig_good = ::sqrt(x2avg - xavg * xavg);
    st->y_avg_good = yavg;
    st->y_sig_good = ::sqrt(y2avg - yavg * yavg);
    st->w_avg_good = g2 > 0 ? s[5] / (double)g2 : -99999;
  }
}


__global__ __launch_bounds__(kBlock) void k_seg_final(SegArgs a) { seg_final_body(a); }

// one workgroup per simulation of a batch
__global__ __launch_bounds__(kBlock) void k_seg_final_batch(const SegArgs* recs,
                                                            const uint64_t* roff)
{
  const SegArgs a = period_recs(recs, roff)[blockIdx.x];
  seg_final_body(a);
}

// --- k_phase -------------------------------------------------------------------

// Per-chunk sums of the relative phase (UpdateRelativePhi(true) then the
// phi_r moments of Beam::UpdateStatForSpaceCharge, beam.cpp:1003-1016):
//   0 phi_r over loss == 0, 1 phi_r and 2 phi_r^2 over good particles.
__device__ __forceinline__ void phase_body(const PhaseArgs& a, uint32_t c)
{
  __shared__ RedSmem sm;
  const int t = threadIdx.x;
  const BeamArrays& b = a.b;
  const Partials& pr = a.pr;
  const double ref = a.st->phi_avg_good;
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < (int)kPerLane; ++j) {
    const uint32_t i = c * kChunk + j * kBlock + t;
    double pr_ = 0.0;
    uint32_t loss = 1, lloss = 0;
    if (i < b.n) {
      loss = b.loss[i];
      lloss = b.lloss[i];
      // the relative phase enters the sums only for loss == 0; a lost
      // particle keeps its last phase, which falls ever further behind the
      // reference, so its exact remainder would take the longest
      if (loss == 0) pr_ = relative_phi(b.phi[i], ref);
    }
    const bool tg = (i < b.n) && loss == 0;
    const bool g = (i < b.n) && is_good(loss, lloss);
    acc[0] = acc[0] + (tg ? pr_ : 0.0);
    const double vg = g ? pr_ : 0.0;
    acc[1] = acc[1] + vg;
    acc[2] = acc[2] + vg * vg;
  }
  uint32_t cnt = 0;
  double mx = 0.0;
  block_reduce<3>(acc, cnt, mx, sm.red, sm.ured, sm.mred, t);
  if (t == 0) {
    pr.sums[c] = acc[0];
    pr.sums[pr.mpad + c] = acc[1];
    pr.sums[2 * pr.mpad + c] = acc[2];
  }
}

__global__ __launch_bounds__(kBlock) void k_phase(PhaseArgs a) { phase_body(a, blockIdx.x); }

// The same for beams of at most kWideMaxChunks chunks, one particle per
// thread (as k_segment_wide): contributions in LDS, lane l (< 256) adds its
// four in the canonical order, then block_reduce's halving tree.
__global__ __launch_bounds__(kWideBlock) void k_phase_wide(PhaseArgs a)
{
  __shared__ double w[3][kWideBlock];
  const int t = threadIdx.x;
  const BeamArrays& b = a.b;
  const Partials& pr = a.pr;
  const uint32_t c = blockIdx.x;
  const double ref = a.st->phi_avg_good;
  const uint32_t i = c * kChunk + (uint32_t)t;
  double pr_ = 0.0;
  uint32_t loss = 1, lloss = 0;
  if (i < b.n) {
    loss = b.loss[i];
    lloss = b.lloss[i];
    if (loss == 0) pr_ = relative_phi(b.phi[i], ref);   // used only for loss == 0
  }
  const bool tg = (i < b.n) && loss == 0;
  const bool g = (i < b.n) && is_good(loss, lloss);
  const double vg = g ? pr_ : 0.0;
  w[0][t] = tg ? pr_ : 0.0;
  w[1][t] = vg;
  w[2][t] = vg * vg;
  __syncthreads();
  if (t < (int)kBlock) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < (int)kPerLane; ++j) acc = acc + w[q][j * kBlock + t];
      w[q][t] = acc;
    }
  }
  __syncthreads();
  if (t < 128) {
#pragma unroll
    for (int q = 0; q < 3; ++q) w[q][t] = w[q][t] + w[q][t + 128];
  }
  __syncthreads();
  if (t < 64) {
    double v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] = w[q][t] + w[q][t + 64];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
      for (int q = 0; q < 3; ++q) v[q] = v[q] + __shfl_down(v[q], o, 64);
    }
    if (t == 0) {
      pr.sums[c] = v[0];
      pr.sums[pr.mpad + c] = v[1];
      pr.sums[2 * pr.mpad + c] = v[2];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_phase_batch(const PhaseArgs* recs, const uint32_t* map,
                                                        const uint64_t* roff)
{
  const uint32_t e = map[blockIdx.x];
  const PhaseArgs a = period_recs(recs, roff)[e >> kBatchLocalBits];
  phase_body(a, e & kBatchLocalMask);
}

// The same sums with one chunk per wavefront (batches of at least
// kPhaseWaveMinChunks chunks): lane l stands for the canonical
// lanes l, l + 64, l + 128 and l + 192 of phase_body, sums each one's four
// particles in its order and folds them as levels 128 and 64 of
// block_reduce's tree, then the same shuffle levels: the same bits, without
// LDS or barriers, and all sixteen particles' loads issued before the first
// relative phase.
__device__ __forceinline__ void phase_wave_body(const PhaseArgs& a, uint32_t c, int l)
{
  const BeamArrays& b = a.b;
  const Partials& pr = a.pr;
  const double ref = a.st->phi_avg_good;
  double ph[4][kPerLane];
  uint32_t lo[4][kPerLane], ll[4][kPerLane];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < (int)kPerLane; ++j) {
      const uint32_t i = c * kChunk + j * kBlock + q * 64 + l;
      const bool in = i < b.n;
      ph[q][j] = in ? gld(b.phi, i) : 0.0;
      lo[q][j] = in ? gld(b.loss, i) : 1u;
      ll[q][j] = in ? gld(b.lloss, i) : 0u;
    }
  double v[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < (int)kPerLane; ++j) {
      // used only for loss == 0 (beyond n, loss reads 1)
      const double pr_ = lo[q][j] == 0 ? relative_phi(ph[q][j], ref) : 0.0;
      const bool tg = lo[q][j] == 0;
      const bool g = is_good(lo[q][j], ll[q][j]);
      acc[0] = acc[0] + (tg ? pr_ : 0.0);
      const double vg = g ? pr_ : 0.0;
      acc[1] = acc[1] + vg;
      acc[2] = acc[2] + vg * vg;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) v[q][k] = acc[k];
  }
  double s[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) s[k] = (v[0][k] + v[2][k]) + (v[1][k] + v[3][k]);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = s[k] + __shfl_down(s[k], o, 64);
  }
  if (l == 0) {
    gst(pr.sums, c, s[0]);
    gst(pr.sums, pr.mpad + c, s[1]);
    gst(pr.sums, 2 * pr.mpad + c, s[2]);
  }
}

__global__ __launch_bounds__(kBlock) void k_phase_wave_batch(const PhaseArgs* recs,
                                                             const uint32_t* map, uint32_t nch,
                                                             const uint64_t* roff)
{
  // the chunk map deals each member's chunks to one XCD's positions (8 j + x
  // for XCD x); workgroup b runs on XCD b mod 8, so its four waves take
  // positions 32 (b / 8) + 8 w + b mod 8, w < 4: every position once, each
  // on its XCD
  static_assert(kBlock == 256, "four waves per workgroup");
  const uint32_t b = blockIdx.x, w = threadIdx.x >> 6;
  const uint32_t k = __builtin_amdgcn_readfirstlane(32u * (b >> 3) + 8u * w + (b & 7u));
  if (k >= nch) return;
  const uint32_t e = map[k];
  const PhaseArgs a = period_recs(recs, roff)[e >> kBatchLocalBits];
  phase_wave_body(a, e & kBatchLocalMask, threadIdx.x & 63);
}

// One workgroup after k_phase: the phi_r moments, then the Scheff::Start
// control logic of space_charge.cpp:84-138 (mesh_size_cutoff_w = 0) on the
// device, and the cleared charge histogram
// when it decides to remesh.  Thread 0 reads every input into registers
// before it writes any output.
__device__ __forceinline__ void phase_final_body(const PhaseArgs& a)
{
  __shared__ RedSmem sm;
  __shared__ int remesh;
  __shared__ uint32_t nbins;
  const int t = threadIdx.x;
  const BeamArrays& b = a.b;
  hps_beam_stats* st = a.st;
  ScDev* sc = a.sc;
  // thread 0's inputs, loaded before the reduction so that both round trips
  // overlap
  uint32_t good_cnt = 0, num_loss = 0;
  double x_sig_good = 0.0, y_sig_good = 0.0, ke = 0.0, phi_sig_good = 0.0;
  double x_avg_good = 0.0, y_avg_good = 0.0;
  hps_spch_state ss{};
  if (t == 0) {
    good_cnt = st->num_good;
    num_loss = st->num_loss;
    x_sig_good = st->x_sig_good;
    y_sig_good = st->y_sig_good;
    x_avg_good = st->x_avg_good;
    y_avg_good = st->y_avg_good;
    ke = st->w_avg_good;
    phi_sig_good = st->phi_sig_good;
    ss = sc->st;
  }
  double s[3] = {0.0, 0.0, 0.0};
  if (a.do_sums) {
    uint32_t dummy;
    final_reduce<3>(a.pr.sums, nullptr, a.pr.mpad, a.pr.nch, sm.red, sm.ured, sm.mred, t, s,
                    dummy);
  }
  // Beam::UpdateMaxRPhi (beam.cu:829-851): the largest r^2 about the last
  // all-particle centroid over loss == 0, for the adjacent-bunch mesh; from
  // the segment's chunk maxima, or directly for beams of at most two
  // particles (no segment statistics).
  double r2max = 0.0;
  if (a.rmax == 1) {
    double v = 0.0;
    for (uint32_t c = t; c < a.pr.nch; c += kBlock) {
      const double u = a.pr.rmax2[c];
      v = v < u ? u : v;
    }
    r2max = block_max(v, sm.mred, t);
  } else if (a.rmax == 2 && t == 0) {
    const double xa = st->x_avg, ya = st->y_avg;
    for (uint32_t i = 0; i < b.n; ++i) {
      if (b.loss[i] != 0) continue;
      const double tx = b.x[i] - xa, ty = b.y[i] - ya;
      const double r2 = tx * tx + ty * ty;
      r2max = r2max < r2 ? r2 : r2max;
    }
  }
  if (t == 0) {
    if (a.do_sums) {
      const double gd = (double)good_cnt;
      const double avg = s[1] / gd, avg2 = s[2] / gd;
      phi_sig_good = ::sqrt(avg2 - avg * avg);
      st->phi_avg_r = s[0] / (double)(uint32_t)(b.n - num_loss);
      st->phi_sig_good = phi_sig_good;
    }
    const double sigphi = b.n < 2 ? 0.0 : phi_sig_good;
    const double sigx = b.n < 2 ? 0.0 : x_sig_good;
    const double sigy = b.n < 2 ? 0.0 : y_sig_good;
    const double gamma = (double)ke / b.mass + 1.0;
    const double beta = ::sqrt(1.0 - 1.0 / (gamma * gamma));
    const double lmbd = 299.792458 / a.freq;
    const double sigz = sigphi * 0.5 / 3.14159265 * beta * lmbd;
    const double sigr = ::sqrt(sigx * sigx + sigy * sigy);
    const double cratior = sigr * sigz;
    const double cratioz = sigz * sigz * gamma * gamma;
    const double cratior_old = ss.sigr_old * ss.sigz_old;
    const double cratioz_old = ss.sigz_old * ss.sigz_old * ss.gm_old * ss.gm_old;
    // adjacent bunches below the cutoff energy only (space_charge.cpp:102-105)
    const uint32_t adj = (ss.adj_bunch_w == 0.0 || ke < ss.adj_bunch_w) ? ss.adj_bunch : 0u;
    double ratio_r = 1.0, ratio_z = 1.0, ratio_q = 1.0, ratio_g = 1.0;
    int rm = 0;
    ss.kick_count += 1;
    if (cratior_old == 0.0 || cratioz_old == 0.0 ||
        __builtin_fabs(cratior - cratior_old) / cratior_old > ss.remesh_val ||
        __builtin_fabs(cratioz - cratioz_old) / cratioz_old > ss.remesh_val) {
      ss.sigr_old = sigr;
      ss.sigz_old = sigz;
      ss.gm_old = gamma;
      ss.good_cnt_old = good_cnt;
      ss.remesh_count += 1;
      // UpdateScheffMeshSizeKernel via UpdateMeshKernelCall: the reference
      // compares the device pointers x_sig_good > y_sig_good (space_charge.cu:131);
      // x_sig_good is allocated first (beam.cu:51,54), so sigma_y is used.
      const double bl = beta * CLIGHT / a.freq;
      if (adj != 0) {
        // with adjacent bunches the mesh spans the largest radius and one
        // RF period (UpdateScheffMeshSizeKernel with r_coef = 1)
        ss.dr = 1.0 * ::sqrt(r2max) / ss.nx;
        ss.dz = bl / ss.nz;
      } else {
        ss.dr = 3.0 * y_sig_good / ss.nx;
        ss.dz = 3.0 * bl * phi_sig_good / (ss.nz * PI);
      }
      // the aspect the grouped field solve of a batch interpolates its
      // shared Green's tables in (gm as the Green's table computes it)
      sc->log_aspect = log_aspect(ss.dr, ss.dz, 1.0 / ::sqrt(1.0 - beta * beta));
      ss.num_on_mesh = 0;
      rm = 1;
    } else {
      ratio_r = sigr / ss.sigr_old;
      ratio_z = sigz / ss.sigz_old;
      // uint / uint, space_charge.cpp:137-138
      ratio_q = ss.good_cnt_old ? (double)(good_cnt / ss.good_cnt_old) : 0.0;
      ratio_g = gamma / ss.gm_old;
    }
    sc->st = ss;
    sc->freq = a.freq;
    sc->beta = beta;
    sc->frac = a.frac;
    sc->current = a.current;
    sc->ratio_r = ratio_r;
    sc->ratio_z = ratio_z;
    sc->ratio_q = ratio_q;
    sc->ratio_g = ratio_g;
    sc->remesh = rm;
    sc->adj_eff = adj;
    if (rm) sc->valid = 1;
    kick_prepare(sc, beta, a.freq, a.current * a.frac, x_avg_good, y_avg_good, x_sig_good,
                 y_sig_good, ss.dr, ss.dz, ss.nz, ratio_r, ratio_z, ratio_q, ratio_g, b.n,
                 b.mass);
    remesh = rm;
    nbins = ss.nx * ss.nz;
  }
  __syncthreads();
  if (remesh)
    for (uint32_t i = t; i < nbins; i += kBlock) a.bins[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_phase_final(PhaseArgs a) { phase_final_body(a); }

__global__ __launch_bounds__(kBlock) void k_phase_final_batch(const PhaseArgs* recs,
                                                              const uint64_t* roff)
{
  const PhaseArgs a = period_recs(recs, roff)[blockIdx.x];
  phase_final_body(a);
}

// --- SCHEFF rebuild ----------------------------------------------------------------

// One entry of the ring Green's table, UpdateFldTbl1Kernel
// (space_charge_kernel.cu:38-100), stored [rp][dz][rs] so that the field sum
// reads consecutive rs.
__device__ __forceinline__ void greens_entry(const ScDev* sc, double* fld1, uint32_t nr,
                                             uint32_t nz, uint32_t r_num_part, double mass,
                                             uint32_t idx)
{
  const uint32_t total = (nr + 1) * nz * nr;
  if (idx >= total) return;
  const uint32_t k = idx % nr;            // source ring index rs
  const uint32_t t = (idx / nr) % nz;     // dz
  const uint32_t rpi = idx / (nr * nz);
  const double r_beta = sc->beta, r_freq = sc->freq;
  const double r_current_ov_mc2 = sc->current * sc->frac / mass;
  const double gm = 1.0 / ::sqrt(1.0 - r_beta * r_beta);
  const double dr = sc->st.dr;
  const double dzg = sc->st.dz * gm;
  const double rs = ::sqrt(0.5 * (k * k + (k + 1) * (k + 1))) * dr;
  const double zs = 0.5 * dzg;
  const double zp = (t + 1) * dzg;
  const double rp = rpi * dr;
  const double d = zp - zs;
  double er1, ez1;
  ring_green(rs, rp, d, er1, ez1);
  // image rings of the neighbouring bunches, one RF period apart in the
  // bunch frame (space_charge_kernel.cu:74-94)
  const uint32_t adj = sc->adj_eff;
  if (adj != 0) {
    const double wave_len = CLIGHT / r_freq;
    const double pl = r_beta * gm * wave_len;
    for (uint32_t i = 1; i <= adj; ++i) {
      double xi = (double)i;
      for (int j = 0; j < 2; ++j) {
        double er, ez;
        ring_green(rs, rp, d - xi * pl, er, ez);
        if (rp != 0.0) er1 += er;
        ez1 += ez;
        xi = -xi;
      }
    }
  }
  const double c1 = 5721.67 * r_current_ov_mc2 / (r_freq * 1e6 * (double)r_num_part);
  fld1[2 * (size_t)idx] = c1 * er1;
  fld1[2 * (size_t)idx + 1] = c1 * ez1;
}

// Charge deposit, DistributeParticleKernel (space_charge_kernel.cu:122-224),
// of workgroup `blk` of `nblk` into an LDS-private int64 fixed-point
// histogram, merged into bins with 64-bit integer atomics: integer sums make
// the result independent of the order.
// Constants of one rebuild's deposit (DistributeParticleKernel's
// prologue, space_charge_kernel.cu:122-150).
struct DepConst {
  uint32_t nx, nz;
  double phimc, xbar, ybar, epsq, ref, dr, dzg, hl, c2, drdr;
};

__device__ __forceinline__ DepConst dep_const(const hps_beam_stats* st, const ScDev* sc)
{
  DepConst d;
  d.nx = sc->st.nx;
  d.nz = sc->st.nz;
  const double r_beta = sc->beta, r_freq = sc->freq;
  d.phimc = st->phi_avg_r;
  d.xbar = st->x_avg_good;
  d.ybar = st->y_avg_good;
  d.epsq = st->x_sig_good / st->y_sig_good;
  d.ref = st->phi_avg_good;
  const double gm = 1.0 / ::sqrt(1.0 - r_beta * r_beta);
  d.dr = sc->st.dr;
  d.dzg = sc->st.dz * gm;
  const double wave_len = CLIGHT / r_freq;
  d.hl = d.nz * 0.5 * d.dzg;
  d.c2 = r_beta * gm * wave_len / TWOPI;
  d.drdr = d.dr * d.dr;
  return d;
}

// One good particle's charge into the LDS histogram: the area weights in r,
// the linear weights in z, as 2^-36 fixed point (DistributeParticleKernel's
// body, space_charge_kernel.cu:151-218); returns whether it landed on the
// mesh.
__device__ __forceinline__ bool deposit_one(double x, double y, double phi, const DepConst& d,
                                            unsigned long long* hist)
{
  const uint32_t nx = d.nx, nz = d.nz;
  const double dr = d.dr, dzg = d.dzg, drdr = d.drdr, hl = d.hl;
  double z = relative_phi(phi, d.ref);
  const double rsq_sh = (x - d.xbar) * (x - d.xbar) / d.epsq + (y - d.ybar) * (y - d.ybar) * d.epsq;
  const int i = to_int(psqrt(rsq_sh) / dr + 1);
  if (i < 1 || i > (int)nx) return false;
  double ddz = z - d.phimc;
  for (int guard = 0; (ddz > PI || ddz < -PI) && guard < 4096; ++guard) {
    if (ddz > 0.0) ddz -= TWOPI;
    else ddz += TWOPI;
  }
  z = -d.c2 * ddz;
  if (!(__builtin_fabs(z) < hl)) return false;
  const double zz = z + hl;
  const int jm1 = to_int(zz / dzg + 1);
  int i1 = i + 1;
  if (rsq_sh < 0.5 * drdr * (double)(i * i + (i - 1) * (i - 1))) i1 = i - 1;
  if (i1 < 1) i1 = 1;
  if (i1 > (int)nx) i1 = (int)nx;
  int j1 = jm1 + 1;
  if (zz < (jm1 - 1 + 0.5) * dzg) j1 = jm1 - 1;
  if (j1 < 1) j1 = 1;
  if (j1 > (int)nz) j1 = (int)nz;
  double a;
  if (i1 == i) {
    a = 1.0;
  } else {
    const double sqrtmp = psqrt(4.0 * rsq_sh / drdr - 1.0);
    const double rminsq = 0.25 * drdr * (sqrtmp - 1.0) * (sqrtmp - 1.0);
    const double rmaxsq = rminsq + sqrtmp * drdr;
    if (i1 < i) a = (rmaxsq - (i - 1) * (i - 1) * drdr) / (rmaxsq - rminsq);
    else a = ((i1 - 1) * (i1 - 1) * drdr - rminsq) / (rmaxsq - rminsq);
  }
  double cc = 1.0;
  if (j1 != jm1) cc = (zz - (j1 - 1 + 0.5) * dzg) / ((jm1 - j1) * dzg);
  const double bb = 1.0 - a;
  const double dd = 1.0 - cc;
  if (jm1 < 1 || jm1 > (int)nz) return false;
  int k = (jm1 - 1) * (int)nx + i;
  atomicAdd(&hist[k - 1], (unsigned long long)(int64_t)(a * cc * 0x1p36));
  k += i1 - i;
  atomicAdd(&hist[k - 1], (unsigned long long)(int64_t)(bb * cc * 0x1p36));
  k = (j1 - 1) * (int)nx + i;
  atomicAdd(&hist[k - 1], (unsigned long long)(int64_t)(a * dd * 0x1p36));
  k += i1 - i;
  atomicAdd(&hist[k - 1], (unsigned long long)(int64_t)(bb * dd * 0x1p36));
  return true;
}

// The deposit of a workgroup's share of the particles (strided by the
// deposit grid) into its LDS histogram, flushed into the global one with
// 64-bit integer atomics.  Each thread loads its next particle (flags and
// coordinates, unconditionally) before depositing the current one; the
// histogram is the same (integer sums), only the load order differs (the
// 1,048,576-particle beam: 14.7 -> 14.2 ms per 800-element step; 64 beams
// of 4,194,304 particles on 64 x 256: 226 -> 185 ms per 160-element step).
__device__ __forceinline__ void deposit_block(const BeamArrays& b, const hps_beam_stats* st,
                                              ScDev* sc, int64_t* bins, uint32_t blk,
                                              uint32_t nblk, unsigned long long* hist,
                                              uint32_t* ured)
{
  const int t = threadIdx.x;
  const DepConst d = dep_const(st, sc);
  const uint32_t nb = d.nx * d.nz;
  for (uint32_t i = t; i < nb; i += kDepositBlock) hist[i] = 0ull;
  __syncthreads();
  uint32_t count = 0;
  const uint32_t stride = nblk * kDepositBlock;
  uint32_t np = blk * kDepositBlock + t;
  bool have = np < b.n;
  uint32_t lo = 1, ll = 0;
  double x = 0.0, y = 0.0, ph = 0.0;
  if (have) {
    lo = gld(b.loss, np); ll = gld(b.lloss, np);
    x = gld(b.x, np); y = gld(b.y, np); ph = gld(b.phi, np);
  }
  while (have) {
    const uint32_t clo = lo, cll = ll;
    const double cx = x, cy = y, cph = ph;
    np += stride;
    have = np < b.n;
    if (have) {
      lo = gld(b.loss, np); ll = gld(b.lloss, np);
      x = gld(b.x, np); y = gld(b.y, np); ph = gld(b.phi, np);
    }
    if (clo == 0 && cll == 0 && deposit_one(cx, cy, cph, d, hist)) ++count;
  }
  __syncthreads();
  for (uint32_t i = t; i < nb; i += kDepositBlock) {
    const unsigned long long v = hist[i];
    if (v != 0ull) atomicAdd((unsigned long long*)&bins[i], v);
  }
  ured[t] = count;
  __syncthreads();
  for (int o = kDepositBlock / 2; o >= 1; o >>= 1) {
    if (t < o) ured[t] += ured[t + o];
    __syncthreads();
  }
  if (t == 0 && ured[0]) atomicAdd(&sc->st.num_on_mesh, ured[0]);
}

// The SCHEFF rebuild after a remesh decision, one launch of
// sc_build_blocks(a) workgroups per simulation.  When the deposit takes at
// least as many workgroups as the Green's table, there are dep_blocks of
// them: the first greens_blocks fill a Green's entry per thread and all
// then deposit (fewer, fuller workgroups than side by side: the
// 1,048,576-particle step measured 65.1 -> 63.7 ms).  Otherwise workgroups
// [0, greens_blocks) fill the table and the next dep_blocks deposit
// (65,536 particles: sc_build 1.55 ms per step side by side, 1.82
// combined).  Both roles read only the control step's results; all exit
// when there is no remesh.

__device__ __forceinline__ void sc_build_body(const BuildArgs& a, uint32_t nr, uint32_t nz,
                                              uint32_t blk, unsigned long long* hist,
                                              uint32_t* ured)
{
  if (!a.sc->remesh) return;
  if (a.dep_blocks >= a.greens_blocks) {
    if (blk < a.greens_blocks)
      greens_entry(a.sc, a.fld1, nr, nz, a.b.n, a.b.mass, blk * kDepositBlock + threadIdx.x);
    deposit_block(a.b, a.st, a.sc, a.bins, blk, a.dep_blocks, hist, ured);
    return;
  }
  if (blk < a.greens_blocks) {
    greens_entry(a.sc, a.fld1, nr, nz, a.b.n, a.b.mass, blk * kDepositBlock + threadIdx.x);
    return;
  }
  deposit_block(a.b, a.st, a.sc, a.bins, blk - a.greens_blocks, a.dep_blocks, hist, ured);
}

__global__ __launch_bounds__(kDepositBlock) void k_sc_build(BuildArgs a, uint32_t nr, uint32_t nz)
{
  extern __shared__ unsigned long long hist[];
  __shared__ uint32_t ured[kDepositBlock];
  sc_build_body(a, nr, nz, blockIdx.x, hist, ured);
}

__global__ __launch_bounds__(kDepositBlock) void k_sc_build_batch(const BuildArgs* recs,
                                                                  const uint32_t* map,
                                                                  uint32_t nr, uint32_t nz)
{
  extern __shared__ unsigned long long hist[];
  __shared__ uint32_t ured[kDepositBlock];
  const uint32_t e = map[blockIdx.x];
  const BuildArgs a = recs[e >> kBatchLocalBits];
  sc_build_body(a, nr, nz, e & kBatchLocalMask, hist, ured);
}

// Field table E(zp, rp) = sum over bins of charge * Green's value,
// UpdateFldTbl2Kernel (space_charge_kernel.cu:230-342), in the canonical
// order: P = min(NextPow2(nr), 64) partial sums, partial p takes the source
// rings rs = p, p + P, ... in ascending order and for each the rows
// zs = 0 .. nz-1 in ascending order (fma), then a halving tree over the P
// partials.  Field point (zp, rp) reads the Green's entry at dz = zs - zp
// (zs >= zp, E_z sign -1) or zp - zs - 1 (zs < zp).
//
// The P partials may be split by residue over S workgroup sets (S a power of
// two dividing P): set s holds the partials p = s, s + S, ...; its halving
// tree runs the levels o >= S (which pair partials of the same residue) and
// stores the level-S value; k_field_combine runs the levels o < S.
//
// A group of P / S lanes computes K consecutive rows zp0 .. zp0 + K - 1 of
// one rp: lane j holds partial s + j S of all K points.  Along zs the K
// points need the Green's entries of K consecutive offsets d = zs - zp, so a
// K-deep window of (E_r, signed E_z) values stays in registers and each step
// loads one new entry and one charge.  The window slot of offset d is d mod K
// (zp0 is a multiple of K), so with the zs loop unrolled by K every slot
// index is a compile-time constant.  kLds: the workgroup's columns (rings of
// its residue) of the rp slice of the table and of the bins are staged in
// LDS first.
struct FieldGeom {
  uint32_t P, S, groups_per_wg, wgs_per_set;
};

template <int K, bool kLds>
__device__ __forceinline__ void field_win_body(const FieldArgs& f, uint32_t nr, uint32_t nz,
                                               const FieldGeom& fg, uint32_t blk, double* lds)
{
  const ScDev* sc = f.sc;
  if (!sc->remesh) return;
  const uint32_t S = fg.S, G = fg.P / S;
  const uint32_t per_rp = S * fg.wgs_per_set;
  const uint32_t rp = blk / per_rp;
  const uint32_t set = (blk % per_rp) / fg.wgs_per_set;
  const uint32_t wg = blk % fg.wgs_per_set;
  const uint32_t fld_sz = nr * nz;
  // columns of this set: rings rs = set + c S, c < nc
  const uint32_t nc = set < nr ? (nr - set + S - 1) / S : 0;
  const double2* gsrc = reinterpret_cast<const double2*>(f.fld1) + (size_t)rp * fld_sz;
  const double2* g = gsrc;
  const double* lb = nullptr;
  uint32_t gstride = nr;
  if (kLds) {
    double2* lg = reinterpret_cast<double2*>(lds);   // [nz][nc] (E_r, E_z)
    double* lq = lds + 2 * (size_t)nz * nc;           // [nz][nc] charge
    // kStage entries per thread are loaded before any is stored: one memory
    // round trip per kStage entries instead of one per entry.  Indexes past
    // the slice are clamped to its last entry, so every load and store is
    // unconditional (the clamped stores write that entry's own values).
    constexpr uint32_t kStage = 8;
    const uint32_t total = nz * nc;
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += kStage * blockDim.x) {
      double2 gv[kStage];
      int64_t qv[kStage];
#pragma unroll
      for (uint32_t u = 0; u < kStage; ++u) {
        const uint32_t i = ::min(i0 + u * blockDim.x, total - 1);
        const uint32_t src = (i / nc) * nr + set + (i % nc) * S;
        const double* e = reinterpret_cast<const double*>(gsrc) + 2 * (size_t)src;
        gv[u] = make_double2(gld(e, 0), gld(e, 1));
        qv[u] = gld(f.bins, src);
      }
#pragma unroll
      for (uint32_t u = 0; u < kStage; ++u) {
        const uint32_t i = ::min(i0 + u * blockDim.x, total - 1);
        lg[i] = gv[u];
        lq[i] = (double)qv[u] * 0x1p-36;
      }
    }
    __syncthreads();
    g = lg;
    lb = lq;
    gstride = nc;
  }
  const uint32_t j = threadIdx.x % G;
  const uint32_t block = wg * fg.groups_per_wg + threadIdx.x / G;
  const int zp0 = (int)(block * K);
  if (zp0 > (int)nz) return;
  const int inz = (int)nz;
  // column of ring rs in the (staged or global) table
  auto col = [&](uint32_t rs) -> uint32_t { return kLds ? (rs - set) / S : rs; };
  // window entry for offset d of ring rs; offsets of the zs loop always
  // index the table (|d| < nz for rows zp <= nz), only the prefill of rows
  // past nz needs the bound
  auto entry = [&](int d, uint32_t c) -> double2 {
    const int t = d >= 0 ? d : -d - 1;
    double2 v = g[(uint32_t)t * gstride + c];
    if (d >= 0) v.y = -v.y;
    return v;
  };
  auto entry_bounded = [&](int d, uint32_t c) -> double2 {
    const int t = d >= 0 ? d : -d - 1;
    return t < inz ? entry(d, c) : make_double2(0.0, 0.0);
  };
  auto charge = [&](int zs, uint32_t c) -> double {
    const uint32_t bi = (uint32_t)zs * gstride + c;
    return kLds ? lb[bi] : (double)f.bins[bi] * 0x1p-36;
  };
  double tx[K], ty[K];
#pragma unroll
  for (int k = 0; k < K; ++k) { tx[k] = 0.0; ty[k] = 0.0; }
  for (uint32_t rs = set + j * S; rs < nr; rs += fg.P) {
    const uint32_t c = col(rs);
    double2 w[K];
    w[0] = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 1; k < K; ++k) w[K - k] = entry_bounded(-zp0 - k, c);
    int zs0 = 0;
    for (; zs0 + K <= inz; zs0 += K) {
#pragma unroll
      for (int u = 0; u < K; ++u) {
        w[u] = entry(zs0 + u - zp0, c);
        const double q = charge(zs0 + u, c);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double2 v = w[(u - k + K) % K];
          tx[k] = ::fma(q, v.x, tx[k]);
          ty[k] = ::fma(q, v.y, ty[k]);
        }
      }
    }
    if (zs0 < inz) {
#pragma unroll
      for (int u = 0; u < K; ++u) {
        if (zs0 + u < inz) {
          w[u] = entry(zs0 + u - zp0, c);
          const double q = charge(zs0 + u, c);
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const double2 v = w[(u - k + K) % K];
            tx[k] = ::fma(q, v.x, tx[k]);
            ty[k] = ::fma(q, v.y, ty[k]);
          }
        }
      }
    }
  }
  // levels o = P/2 .. S of the tree: lane offsets G/2 .. 1 inside the group
  for (uint32_t o = G / 2; o >= 1; o >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double ox = __shfl_down(tx[k], o, (int)G);
      const double oy = __shfl_down(ty[k], o, (int)G);
      tx[k] = tx[k] + ox;
      ty[k] = ty[k] + oy;
    }
  }
  if (j == 0) {
    const uint32_t nout = (nz + 1) * (nr + 1);
    double* dst = S == 1 ? f.fld2 : f.part + 2 * (size_t)set * nout;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int zp = zp0 + k;
      if (zp <= inz) {
        const size_t o = (size_t)zp * (nr + 1) + rp;
        dst[2 * o] = tx[k];
        dst[2 * o + 1] = ty[k];
      }
    }
  }
}

template <int K, bool kLds>
__global__ __launch_bounds__(1024) void k_field_win(FieldArgs f, uint32_t nr, uint32_t nz,
                                                    FieldGeom fg)
{
  extern __shared__ double lds[];
  field_win_body<K, kLds>(f, nr, nz, fg, blockIdx.x, lds);
}

// per simulation of a batch: blocks [m * per, (m + 1) * per)
template <int K, bool kLds>
__global__ __launch_bounds__(1024) void k_field_win_batch(const FieldArgs* recs, uint32_t per,
                                                          uint32_t nr, uint32_t nz, FieldGeom fg)
{
  extern __shared__ double lds[];
  const FieldArgs f = recs[blockIdx.x / per];
  field_win_body<K, kLds>(f, nr, nz, fg, blockIdx.x % per, lds);
}

// Levels o = S/2 .. 1 of the field tree over the split results.
__global__ __launch_bounds__(kBlock) void k_field_combine(FieldArgs f, uint32_t nout, uint32_t S)
{
  if (!f.sc->remesh) return;
  const uint32_t o = blockIdx.x * kBlock + threadIdx.x;
  if (o >= nout) return;
  double ax[kFieldMaxSplit], ay[kFieldMaxSplit];
#pragma unroll
  for (uint32_t l = 0; l < kFieldMaxSplit; ++l) {
    ax[l] = l < S ? f.part[2 * ((size_t)l * nout + o)] : 0.0;
    ay[l] = l < S ? f.part[2 * ((size_t)l * nout + o) + 1] : 0.0;
  }
#pragma unroll
  for (uint32_t h = kFieldMaxSplit / 2; h >= 1; h >>= 1) {
    if (h < S) {
#pragma unroll
      for (uint32_t l = 0; l < h; ++l) {
        ax[l] = ax[l] + ax[l + h];
        ay[l] = ay[l] + ay[l + h];
      }
    }
  }
  f.fld2[2 * (size_t)o] = ax[0];
  f.fld2[2 * (size_t)o + 1] = ay[0];
}

// --- getter reductions ---------------------------------------------------------

// With rel_ref given, the first variable is the relative phase of pa
// against rel_ref[0] (what k_relative_phi stores), computed in registers.
__global__ __launch_bounds__(kBlock) void k_sums(BeamArrays b, const double* pa, const double* pb,
                                                 int filter, Partials pr, const double* rel_ref)
{
  __shared__ RedSmem sm;
  const int t = threadIdx.x;
  const uint32_t c = blockIdx.x;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  uint32_t nloss = 0, ngood = 0, nlloss = 0;
#pragma unroll
  for (int j = 0; j < (int)kPerLane; ++j) {
    const uint32_t i = c * kChunk + j * kBlock + t;
    if (i >= b.n) continue;
    const uint32_t loss = b.loss[i], lloss = b.lloss[i];
    nloss += loss != 0;
    nlloss += lloss != 0;
    ngood += is_good(loss, lloss);
    const bool inc = filter == 0 ? loss == 0 : (filter == 1 ? is_good(loss, lloss) : true);
    double va = (inc && pa) ? pa[i] : 0.0;
    if (inc && pa && rel_ref) va = relative_phi(va, rel_ref[0]);
    const double vb = (inc && pb) ? pb[i] : 0.0;
    acc[0] = acc[0] + va;
    acc[1] = acc[1] + va * va;
    acc[2] = acc[2] + vb;
    acc[3] = acc[3] + vb * vb;
    acc[4] = acc[4] + va * vb;
  }
  // counts are packed 3 x 11 bits (at most 1024 per chunk each)
  uint32_t packed = nloss | (ngood << 11) | (nlloss << 22);
  double mx = 0.0;
  block_reduce<5>(acc, packed, mx, sm.red, sm.ured, sm.mred, t);
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < 5; ++q) pr.sums[q * pr.mpad + c] = acc[q];
    pr.cnt[c] = packed & 0x7ffu;
    pr.cnt[pr.mpad + c] = (packed >> 11) & 0x7ffu;
    pr.cnt[2 * pr.mpad + c] = packed >> 22;
  }
}

// Cross-chunk trees of k_sums in one workgroup.
__global__ __launch_bounds__(kBlock) void k_sums_final(Partials pr)
{
  __shared__ RedSmem sm;
  const int t = threadIdx.x;
  double s[5];
  uint32_t c0, c1, c2;
  final_reduce<5>(pr.sums, pr.cnt, pr.mpad, pr.nch, sm.red, sm.ured, sm.mred, t, s, c0);
  double d1[1], d2[1];
  final_reduce<1>(pr.sums, pr.cnt + pr.mpad, pr.mpad, pr.nch, sm.red, sm.ured, sm.mred, t, d1, c1);
  final_reduce<1>(pr.sums, pr.cnt + 2 * pr.mpad, pr.mpad, pr.nch, sm.red, sm.ured, sm.mred, t, d2, c2);
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < 5; ++q) pr.out[q] = s[q];
    pr.out_cnt[0] = c0;
    pr.out_cnt[1] = c1;
    pr.out_cnt[2] = c2;
  }
}

__global__ void k_relative_phi(BeamArrays b, const double* ref)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < b.n) b.phi_r[i] = relative_phi(b.phi[i], ref[0]);
}

// UpdateLongitudinalLossCoordinateKernel, beam_kernel.cu:1549-1559
__global__ void k_lloss_flag(BeamArrays b, const double* avg)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < b.n && b.phi[i] > avg[0] + TWOPI + PI) b.lloss[i] = 1;
}

// ShiftVariableKernel, beam_kernel.cu:1489-1496 (particle 0 untouched)
__global__ void k_shift(double* a, uint32_t n, double v)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i > 0 && i < n) a[i] += v;
}

// ChangeFrequencyKernel, beam_kernel.cu:1531-1537
__global__ void k_scale(double* a, uint32_t n, double v)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] *= v;
}

// CutBeamKernel, beam_kernel.cu:1464-1480
__global__ void k_cut(const double* a, uint32_t* loss, uint32_t n, double lo, double hi)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i > 0 && i < n && loss[i] == 0) {
    const double x = a[i];
    if (x < hi && x > lo) loss[i] = 99999;
  }
}

__global__ void k_probe(int fn, const double* x, const double* y, double* out, uint32_t n)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  double r = 0.0, s, c;
  switch (fn) {
    case 0: r = hps::exp(v); break;
    case 1: r = hps::log(v); break;
    case 2: hps::sincos(v, &s, &c); r = s; break;
    case 3: hps::sincos(v, &s, &c); r = c; break;
    case 4: r = hps::sinh(v); break;
    case 5: r = hps::cosh(v); break;
    case 6: r = ::sqrt(v); break;
    case 7: r = hps::fmod_pos(v, TWOPI); break;
    case 8: r = v / y[i]; break;
    case 9: r = bessi0c(v); break;
    case 10: r = bessi1c(v); break;
    case 11: r = bessi1p(v); break;
    case 12: r = psqrt(v); break;
    case 13: psincos(v, &s, &c); r = s; break;
    case 14: psincos(v, &s, &c); r = c; break;
    case 15: psinhcosh(v, &s, &c); r = s; break;
    case 16: psinhcosh(v, &s, &c); r = c; break;
    default: break;
  }
  out[i] = r;
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_segment(const SegArgs& a, hipStream_t s)
{
  const uint32_t nch = (a.b.n + kChunk - 1) / kChunk;
  if (nch <= kWideMaxChunks) {
    const size_t lds = sizeof(WideSmem);
    if (a.op_class >= OPC_GAP)
      hipLaunchKernelGGL((k_segment_wide<OPC_GAP>), dim3(nch), dim3(kWideBlock), lds, s, a);
    else if (a.op_class == OPC_QUAD)
      hipLaunchKernelGGL((k_segment_wide<OPC_QUAD>), dim3(nch), dim3(kWideBlock), lds, s, a);
    else
      hipLaunchKernelGGL((k_segment_wide<OPC_LIGHT>), dim3(nch), dim3(kWideBlock), lds, s, a);
    return hipGetLastError();
  }
  if (a.op_class >= OPC_GAP)
    hipLaunchKernelGGL((k_segment<OPC_GAP, kSegGroupGap>), dim3(nch), dim3(kBlock), 0, s, a);
  else if (a.op_class == OPC_QUAD)
    hipLaunchKernelGGL((k_segment<OPC_QUAD, kSegGroupQuad>), dim3(nch), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_segment<OPC_LIGHT, kSegGroupLight>), dim3(nch), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_segment_batch(const SegArgs* recs, const uint32_t* map, uint32_t nblocks,
                                int op_class, const uint64_t* roff, hipStream_t s)
{
  if (op_class >= OPC_GAP)
    hipLaunchKernelGGL((k_segment_batch<OPC_GAP, kSegGroupGap>), dim3(nblocks), dim3(kBlock), 0,
                       s, recs, map, roff);
  else if (op_class == OPC_QUAD)
    hipLaunchKernelGGL((k_segment_batch<OPC_QUAD, kSegGroupQuad>), dim3(nblocks), dim3(kBlock),
                       0, s, recs, map, roff);
  else
    hipLaunchKernelGGL((k_segment_batch<OPC_LIGHT, kSegGroupLight>), dim3(nblocks),
                       dim3(kBlock), 0, s, recs, map, roff);
  return hipGetLastError();
}

hipError_t launch_seg_final(const SegArgs& a, hipStream_t s)
{
  hipLaunchKernelGGL(k_seg_final, dim3(1), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_seg_final_batch(const SegArgs* recs, uint32_t nmem, const uint64_t* roff,
                                  hipStream_t s)
{
  hipLaunchKernelGGL(k_seg_final_batch, dim3(nmem), dim3(kBlock), 0, s, recs, roff);
  return hipGetLastError();
}

hipError_t launch_phase(const PhaseArgs& a, hipStream_t s)
{
  const uint32_t nch = (a.b.n + kChunk - 1) / kChunk;
  if (nch <= kWideMaxChunks)
    hipLaunchKernelGGL(k_phase_wide, dim3(nch), dim3(kWideBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_phase, dim3(nch), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_phase_final(const PhaseArgs& a, hipStream_t s)
{
  hipLaunchKernelGGL(k_phase_final, dim3(1), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_phase_batch(const PhaseArgs* recs, const uint32_t* map, uint32_t nblocks,
                              const uint64_t* roff, hipStream_t s)
{
  // nblocks: one per chunk of the batch.  Batches of at least
  // kPhaseWaveMinChunks chunks run a wavefront per chunk (four chunks per
  // workgroup); smaller ones a workgroup per chunk, which keeps more waves in
  // flight when there are few chunks.
  if (nblocks >= kPhaseWaveMinChunks)
    hipLaunchKernelGGL(k_phase_wave_batch, dim3((nblocks + 31) / 32 * 8), dim3(kBlock), 0, s,
                       recs, map, nblocks, roff);
  else
    hipLaunchKernelGGL(k_phase_batch, dim3(nblocks), dim3(kBlock), 0, s, recs, map, roff);
  return hipGetLastError();
}

hipError_t launch_phase_final_batch(const PhaseArgs* recs, uint32_t nmem, const uint64_t* roff,
                                    hipStream_t s)
{
  hipLaunchKernelGGL(k_phase_final_batch, dim3(nmem), dim3(kBlock), 0, s, recs, roff);
  return hipGetLastError();
}

// roff[0] = v (set) or roff[0] += v (advance): the record offset of the
// next replay of a period graph, as a kernel node (no memset node).
__global__ void k_roff(uint64_t* roff, uint64_t v, int add)
{
  if (threadIdx.x == 0) roff[0] = add ? roff[0] + v : v;
}

hipError_t launch_roff(uint64_t* roff, uint64_t v, bool add, hipStream_t s)
{
  hipLaunchKernelGGL(k_roff, dim3(1), dim3(64), 0, s, roff, v, add ? 1 : 0);
  return hipGetLastError();
}

uint32_t deposit_blocks(uint32_t n, uint32_t per_thread)
{
  uint32_t grid = (n + per_thread * kDepositBlock - 1) / (per_thread * kDepositBlock);
  if (grid < 1) grid = 1;
  if (grid > kDepositMaxBlocks) grid = kDepositMaxBlocks;
  return grid;
}

// Deposit workgroups of a batch member whose rebuild is the deposit alone
// (grouped field solve, no per-beam Green's table): four particles per
// thread, at most kDepositMaxBlocksBatch workgroups, or kDepositMaxBlocksBig
// when the histogram takes more than 64 KB of LDS (one workgroup per CU).
// Fewer, fuller workgroups spread the zeroing and flush of each workgroup's
// histogram over more particles: 64 members of 4,194,304 particles on the
// 64 x 256 mesh, deposit per 160-element step 360 / 294 / 261 / 241 / 233 /
// 227 ms at 1,024 / 512 / 256 / 128 / 64 / 32 workgroups per member; the
// 256-beam scan on 32 x 128 takes 84 ms at 128 and 91 ms at 32.
uint32_t deposit_blocks_batch(uint32_t n, uint32_t nr, uint32_t nz)
{
  const uint32_t cap = (size_t)nr * nz * sizeof(unsigned long long) > 65536
                           ? kDepositMaxBlocksBig
                           : kDepositMaxBlocksBatch;
  uint32_t grid = (n + 4 * kDepositBlock - 1) / (4 * kDepositBlock);
  if (grid < 1) grid = 1;
  if (grid > cap) grid = cap;
  return grid;
}

uint32_t greens_blocks(uint32_t nr, uint32_t nz)
{
  return ((nr + 1) * nz * nr + kDepositBlock - 1) / kDepositBlock;
}

hipError_t launch_sc_build(ScDev* sc, double* fld1, uint32_t nr, uint32_t nz, BeamArrays b,
                           const hps_beam_stats* st, int64_t* bins, bool deposit,
                           hipStream_t s)
{
  BuildArgs a;
  a.sc = sc; a.fld1 = fld1; a.b = b; a.st = st; a.bins = bins;
  a.greens_blocks = greens_blocks(nr, nz);
  // a lone beam spreads its deposit over up to one particle per thread (65,536
  // particles: 64 workgroups, 9.9 -> 9.0 ms per 160-element step); from
  // 262,144 particles on the 256-workgroup cap holds
  a.dep_blocks = deposit ? deposit_blocks(b.n, 1) : 0;
  const size_t lds = (size_t)nr * nz * sizeof(unsigned long long);
  hipLaunchKernelGGL(k_sc_build, dim3(sc_build_blocks(a)), dim3(kDepositBlock), lds, s, a, nr,
                     nz);
  return hipGetLastError();
}

hipError_t launch_sc_build_batch(const BuildArgs* recs, const uint32_t* map, uint32_t nblocks,
                                 uint32_t nr, uint32_t nz, hipStream_t s)
{
  const size_t lds = (size_t)nr * nz * sizeof(unsigned long long);
  hipLaunchKernelGGL(k_sc_build_batch, dim3(nblocks), dim3(kDepositBlock), lds, s, recs, map, nr,
                     nz);
  return hipGetLastError();
}

namespace {

// Launch shape of the field sum.
struct FieldShape {
  bool lds;
  int K;
  FieldGeom g;
  uint32_t threads, blocks;
  size_t lds_bytes;
};

FieldShape field_shape_k(uint32_t nr, uint32_t nz, size_t max_lds, int K, uint32_t S,
                         uint32_t max_threads)
{
  FieldShape f{};
  uint32_t P = 1;
  while (P < nr && P < 64) P <<= 1;
  if (S > P) S = P;
  f.g.P = P;
  f.g.S = S;
  f.K = K;
  const uint32_t G = P / S;
  const uint32_t nblk = (nz + 1 + K - 1) / K;
  uint32_t threads = nblk * G;
  threads = (threads + 63) / 64 * 64;
  if (threads > max_threads) threads = max_threads;
  f.threads = threads;
  f.g.groups_per_wg = threads / G;
  f.g.wgs_per_set = (nblk + f.g.groups_per_wg - 1) / f.g.groups_per_wg;
  f.blocks = (nr + 1) * S * f.g.wgs_per_set;
  const uint32_t nc = (nr + S - 1) / S;   // columns of the widest set
  f.lds_bytes = (size_t)24 * nz * nc;
  f.lds = f.lds_bytes <= max_lds;
  if (!f.lds) f.lds_bytes = 0;
  return f;
}

// A batch has workgroups enough with the deepest window and all partials in
// one workgroup.  A lone simulation uses small (256-thread) workgroups with
// all partials, so that the rows spread over more CUs and no combine step is
// needed, when its table slice fits in LDS (32 x 128: 11.7 ms/step of field
// time against 13.2 for the 8-set split plus combine); otherwise it splits
// its partials over sets (each set stages only its rings).
FieldShape field_shape(uint32_t nr, uint32_t nz, size_t max_lds, uint32_t nsim)
{
  if (nsim * (nr + 1) >= 1024) return field_shape_k(nr, nz, max_lds, 8, 1, 1024);
  const FieldShape f = field_shape_k(nr, nz, max_lds, 4, 1, 256);
  if (f.lds) return f;
  return field_shape_k(nr, nz, max_lds, 4, kFieldMaxSplit, 1024);
}

template <bool kBatch>
hipError_t launch_field_shape(const FieldShape& f, const FieldArgs* recs, const FieldArgs& one,
                              uint32_t nsim, uint32_t nr, uint32_t nz, hipStream_t s)
{
  const dim3 grid(f.blocks * nsim), block(f.threads);
#define HPS_FIELD_LAUNCH(KK, LL)                                                              \
  if (kBatch)                                                                                 \
    hipLaunchKernelGGL((k_field_win_batch<KK, LL>), grid, block, f.lds_bytes, s, recs,        \
                       f.blocks, nr, nz, f.g);                                                \
  else                                                                                        \
    hipLaunchKernelGGL((k_field_win<KK, LL>), grid, block, f.lds_bytes, s, one, nr, nz, f.g);
  if (f.K == 8) {
    if (f.lds) { HPS_FIELD_LAUNCH(8, true) } else { HPS_FIELD_LAUNCH(8, false) }
  } else if (f.K == 4) {
    if (f.lds) { HPS_FIELD_LAUNCH(4, true) } else { HPS_FIELD_LAUNCH(4, false) }
  } else if (f.K == 2) {
    if (f.lds) { HPS_FIELD_LAUNCH(2, true) } else { HPS_FIELD_LAUNCH(2, false) }
  } else {
    if (f.lds) { HPS_FIELD_LAUNCH(1, true) } else { HPS_FIELD_LAUNCH(1, false) }
  }
#undef HPS_FIELD_LAUNCH
  return hipGetLastError();
}

}  // namespace

hipError_t launch_field(const FieldArgs& a, uint32_t nr, uint32_t nz, size_t max_lds,
                        hipStream_t s)
{
  const FieldShape f = field_shape(nr, nz, max_lds, 1);
  hipError_t e = launch_field_shape<false>(f, nullptr, a, 1, nr, nz, s);
  if (e != hipSuccess || f.g.S == 1) return e;
  const uint32_t nout = (nz + 1) * (nr + 1);
  hipLaunchKernelGGL(k_field_combine, dim3((nout + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a,
                     nout, f.g.S);
  return hipGetLastError();
}

namespace {

// Field-sum shapes a batch can use (all partials in one workgroup set): window
// depth x workgroup size.  Every shape computes the same canonical sums.
constexpr int kFieldBatchK[] = {8, 4, 2};
constexpr uint32_t kFieldBatchThreads[] = {1024, 512, 256};
constexpr uint32_t kFieldBatchShapes = 9;

FieldShape field_batch_shape(uint32_t nr, uint32_t nz, size_t max_lds, uint32_t idx)
{
  return field_shape_k(nr, nz, max_lds, kFieldBatchK[idx / 3], 1, kFieldBatchThreads[idx % 3]);
}

}  // namespace

uint32_t field_batch_shapes() { return kFieldBatchShapes; }

uint32_t field_batch_default_shape(uint32_t nr, uint32_t nmem)
{
  return nmem * (nr + 1) >= 1024 ? 0u : 3u;   // K 8 or 4, 1024 threads
}

hipError_t launch_field_batch(const FieldArgs* recs, uint32_t nmem, uint32_t nr, uint32_t nz,
                              size_t max_lds, uint32_t shape, hipStream_t s)
{
  FieldArgs none{};
  if (shape >= kFieldBatchShapes) shape = field_batch_default_shape(nr, nmem);
  return launch_field_shape<true>(field_batch_shape(nr, nz, max_lds, shape), recs, none, nmem,
                                  nr, nz, s);
}

hipError_t launch_sums(BeamArrays b, const double* a, const double* bb, int filter, Partials pr,
                       hipStream_t s, const double* rel_ref)
{
  const uint32_t nch = (b.n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(k_sums, dim3(nch), dim3(kBlock), 0, s, b, a, bb, filter, pr, rel_ref);
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(kBlock), 0, s, pr);
  return hipGetLastError();
}

hipError_t launch_relative_phi(BeamArrays b, const double* ref, hipStream_t s)
{
  hipLaunchKernelGGL(k_relative_phi, dim3(blocks_for(b.n)), dim3(kBlock), 0, s, b, ref);
  return hipGetLastError();
}

hipError_t launch_lloss_flag(BeamArrays b, const double* avg, hipStream_t s)
{
  hipLaunchKernelGGL(k_lloss_flag, dim3(blocks_for(b.n)), dim3(kBlock), 0, s, b, avg);
  return hipGetLastError();
}

hipError_t launch_shift(double* a, uint32_t n, double v, hipStream_t s)
{
  hipLaunchKernelGGL(k_shift, dim3(blocks_for(n)), dim3(kBlock), 0, s, a, n, v);
  return hipGetLastError();
}

hipError_t launch_scale(double* a, uint32_t n, double v, hipStream_t s)
{
  hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(kBlock), 0, s, a, n, v);
  return hipGetLastError();
}

hipError_t launch_cut(const double* a, uint32_t* loss, uint32_t n, double lo, double hi,
                      hipStream_t s)
{
  hipLaunchKernelGGL(k_cut, dim3(blocks_for(n)), dim3(kBlock), 0, s, a, loss, n, lo, hi);
  return hipGetLastError();
}

// --- synthetic beams ---------------------------------------------------------------

// Philox4x32-10 block (Salmon et al., SC'11): counter c, key (k0, k1).
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

// Coordinate c of particle i: center + sigma * z, z a standard normal
// truncated to |z| <= 3 by rejection with the attempt number in the Philox
// counter (i, c, attempt), Box-Muller from two 53-bit uniforms -- the
// sampler of hpsim_amd/configs.py (truncated_normal) with the engine's own
// log and cos.  Particle 0 is the reference particle (on axis, centred).
__global__ void k_sample_gaussian(BeamArrays b, uint32_t seed, GaussSpec g)
{
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= b.n) return;
  double* out[6] = {b.x, b.xp, b.y, b.yp, b.phi, b.w};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double z = 0.0;
    for (uint32_t attempt = 0; attempt <= 64; ++attempt) {
      uint32_t r[4] = {i, 0u, (uint32_t)c, attempt};
      philox(r, seed, 0x48505331u);
      const double u1 = 1.0 - u53(r[0], r[1]);
      const double u2 = u53(r[2], r[3]);
      double sn, cs;
      hps::sincos(2.0 * PI * u2, &sn, &cs);
      z = ::sqrt(-2.0 * hps::log(u1)) * cs;
      if (__builtin_fabs(z) <= 3.0) break;
    }
    double v = g.center[c] + g.sigma[c] * z;
    if (i == 0) v = c < 4 ? 0.0 : g.center[c];
    out[c][i] = v;
  }
  b.loss[i] = 0;
  b.lloss[i] = 0;
}

hipError_t launch_sample_gaussian(BeamArrays b, uint32_t seed, const GaussSpec& g, hipStream_t s)
{
  hipLaunchKernelGGL(k_sample_gaussian, dim3(blocks_for(b.n)), dim3(kBlock), 0, s, b, seed, g);
  return hipGetLastError();
}

hipError_t launch_probe(int fn, const double* x, const double* y, double* out, uint32_t n,
                        hipStream_t s)
{
  hipLaunchKernelGGL(k_probe, dim3(blocks_for(n)), dim3(kBlock), 0, s, fn, x, y, out, n);
  return hipGetLastError();
}

#ifdef HPS_FAST_PUSH
}  // namespace fast
#endif

}  // namespace hps
