// hps_field_gemm.hip -- grouped SCHEFF field solve of a batch on the FP32
// matrix cores.
//
// The exact (FP64) batch evaluates one Green's table per beam and one table
// x charge sum per beam.  Here the Green's table is normalised to dr = 1
// (it then depends on the bunch-frame cell aspect kappa = dz*gamma/dr
// alone and scales as 1/dr^2, space_charge_kernel.cu:38-100) and kept for
// the aspect classes kappa_c = (1+s)^c of a grid; a beam whose aspect is
// nearest kappa_c gets the field interpolated quadratically in log(kappa)
// between the fields of classes c - 1, c, c + 1, on its own mesh (the
// interpolation error is third order in the grid step: below 5e-6 of the
// field at s = 0.05).  All beams that remesh at a kick point share the
// class tables, and the fields of all of them are one grouped product
//
//   W[(rp, comp, zp)][col] = sum over (rs, zs) of A_c[(rp, comp, zp)][(rs, zs)] * n_beam[(rs, zs)]
//
// (n = particles per bin; three columns per beam), then E_beam = (sum of
// the Lagrange weights times its columns) * c1_beam / dr_beam^2.  The product
// replaces UpdateFldTbl2Kernel (space_charge_kernel.cu:230-342) for every
// beam that remeshes at a kick point, as one grouped GEMM with FP32 operands
// and FP32 accumulation on v_mfma_f32_16x16x4f32.
//
// A_c is never stored dense: for fixed (rp, rs, comp) its (zp, zs) block is
// Toeplitz, A = S[zp - zs], with S the Green's entries by signed offset
// (zs >= zp: dz = zs - zp, E_z negated; zs < zp: dz = zp - zs - 1), so a
// class table holds S[rp][rs][comp][2 nz] floats (2.2 MB at 32 x 128) and a
// 128 x 32 operand tile is built in LDS from 2 x 95 entries.
//
// Launch sequence per kick point (after the control step and the deposit):
//   k_gemm_schedule  one workgroup: remeshing beams grouped by class slot
//                    (direct-mapped cache of class tables), slots to build
//   k_gemm_tables    the normalised Green's entries of the slots to build
//   k_gemm_pack      fixed-point histograms -> FP32 particle counts
//   k_gemm_field     persistent grid over (K split, beam tile, row tile)
//   k_gemm_store     sum of the K splits, c1/dr^2 scale, field table of
//                    each remeshing beam in the [zp][rp] double2 layout
//
// MFMA 16x16x4 f32 operand layout (wave64): lane l supplies A[m][k] for
// m = l % 16, k = l / 16, and B[k][n] for n = l % 16, k = l / 16; the result
// D[m][n] is held for n = l % 16, m = 4 (l / 16) + i in element i.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "hps_internal.hpp"
#include "hps_unet.hpp"

namespace hps {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;       // 4 wavefronts
constexpr int kTileM = 128;         // rows per tile
constexpr int kTileK = 32;          // zs per K stage
constexpr int kZBlock = 64;         // zp per Toeplitz row tile (x 2 components)
constexpr int kWin = 96;            // Toeplitz window per component (95 used)
constexpr int kLdsK = kTileK + 1;   // padded row pitch of the staged tiles

__device__ __forceinline__ f4 mfma(float a, float b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__host__ __device__ __forceinline__ uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

// ---- schedule --------------------------------------------------------------

// One workgroup.  Thread 0 walks the members in order.  A remeshing member
// gets the exact product when no surrogate runs, or at an audit kick
// point, or when its last audit failed: three product columns, for its
// classes c - 1, c, c + 1 (class table of class k in slot k mod kGemmSlots; a slot
// holding another class is rebuilt, unless another class already claimed it
// at this kick point -- then the column uses the slot's class, counted in
// conflicts).  Columns are grouped by slot (counting sort, member order
// kept) in perm; with a surrogate, every remeshing member also gets a
// position in the network's batch.
__global__ __launch_bounds__(kThreads) void k_gemm_schedule(const GemmMember* __restrict__ mem,
                                                            uint32_t nmem, GemmState* gs,
                                                            uint32_t* perm, MemberMap mm,
                                                            SurCtl* ctl, double lnr)
{
  __shared__ uint32_t count[kGemmSlots];
  __shared__ uint32_t claimed[kGemmSlots];
  static_assert(kGemmSlots <= 127, "slot numbers are kept in int8");
  __shared__ int8_t mslot[kGemmCols * kGemmMaxMembers];
  __shared__ int32_t mclass[kGemmMaxMembers];   // first class c - 1 of a remeshing member
  __shared__ uint8_t mrem[kGemmMaxMembers];     // the member remeshes at this kick point
  __shared__ uint32_t any_rem;
  const int t = threadIdx.x;
  if (t < (int)kGemmSlots) { count[t] = 0; claimed[t] = 0; }
  if (t == 0) any_rem = 0;
  __syncthreads();
  // every member's remesh flag and aspect class, loaded by all threads (one
  // dependent pair of loads per member); thread 0 then walks them in order
  for (uint32_t m = t; m < nmem; m += kThreads) {
    const ScDev* sc = mem[m].sc;
    const uint32_t rem = sc->remesh;
    mrem[m] = rem ? 1 : 0;
    if (rem) {
      double wl[3];
      mclass[m] = aspect_lagrange(sc->log_aspect, lnr, wl) - 1;
      any_rem = 1;
    }
  }
  __syncthreads();
  if (t != 0) return;
  bool audit = false;
  if (ctl) {
    const uint32_t any = any_rem;
    if (ctl->audit_every && ctl->kick % ctl->audit_every == 0) ctl->pending = 1;
    ctl->kick += 1;
    audit = ctl->pending && any;
    if (audit) ctl->pending = 0;
    ctl->audit_now = audit ? 1 : 0;
  }
  uint32_t nbuild = 0, conflicts = gs->conflicts, nn = 0, ncol = 0;
  for (uint32_t m = 0; m < nmem; ++m) {
    mm.pos_nn[m] = -1;
    for (uint32_t h = 0; h < kGemmCols; ++h) mslot[kGemmCols * m + h] = -1;
    if (!mrem[m]) continue;
    if (ctl) mm.pos_nn[m] = (int32_t)nn++;
    if (ctl && !audit && !mm.fallback[m]) continue;
    const int32_t c0 = mclass[m];
    for (uint32_t h = 0; h < kGemmCols; ++h) {
      const int32_t c = c0 + (int32_t)h;
      const int32_t slot = (int32_t)((uint32_t)c & (kGemmSlots - 1));
      if (gs->slot_class[slot] != c) {
        if (!claimed[slot]) {
          gs->slot_class[slot] = c;
          gs->build[nbuild++] = (uint32_t)slot;
        } else {
          ++conflicts;
        }
      }
      claimed[slot] = 1;
      count[slot] += 1;
      mslot[kGemmCols * m + h] = (int8_t)slot;
      ++ncol;
    }
  }
  gs->nbuild = nbuild;
  gs->builds_total += nbuild;
  gs->conflicts = conflicts;
  uint32_t start = 0;
  for (uint32_t s = 0; s < kGemmSlots; ++s) {
    gs->slot_start[s] = start;
    gs->slot_count[s] = count[s];
    start += count[s];
    count[s] = gs->slot_start[s];   // running insertion point
  }
  gs->nrem = start;
  gs->exact_total += start;
  for (uint32_t e = 0; e < kGemmCols * nmem; ++e) {
    const int32_t slot = mslot[e];
    mm.pos_exact[e] = -1;
    if (slot >= 0) {
      mm.pos_exact[e] = (int32_t)count[slot];
      perm[count[slot]++] = e / kGemmCols;
    }
  }
  if (ctl) {
    ctl->nn_n = nn;
    ctl->nn_member_kicks += nn;
    ctl->exact_member_kicks += ncol / kGemmCols;
  }
}

// ---- class tables ------------------------------------------------------------

// S[slot][rp][rs][comp][o], o = delta + nz - 1 for delta = zp - zs in
// [1 - nz, nz]: the ring Green's entries of UpdateFldTbl1Kernel with dr = 1,
// dzg = (1+s)^c, adjacent bunches off, no c1 factor.  One thread per
// (rp, rs, dz) entry writes its two signed offsets.
__global__ __launch_bounds__(kThreads) void k_gemm_tables(const GemmState* __restrict__ gs,
                                                          float* __restrict__ tables,
                                                          uint32_t nr, uint32_t nz, double lnr)
{
  const uint32_t per = (nr + 1) * nr * nz;
  const uint32_t total = gs->nbuild * per;
  const size_t slot_sz = (size_t)(nr + 1) * nr * 4 * nz;
  for (uint32_t idx = blockIdx.x * kThreads + threadIdx.x; idx < total;
       idx += gridDim.x * kThreads) {
    const uint32_t slot = gs->build[idx / per];
    const uint32_t e = idx % per;
    const uint32_t dzi = e % nz;
    const uint32_t rs_i = (e / nz) % nr;
    const uint32_t rp_i = e / (nz * nr);
    const double dzg = hps::exp((double)gs->slot_class[slot] * lnr);
    const double rs = ::sqrt(0.5 * (rs_i * rs_i + (rs_i + 1) * (rs_i + 1)));
    const double d = (dzi + 1) * dzg - 0.5 * dzg;
    double er, ez;
    ring_green(rs, (double)rp_i, d, er, ez);
    float* row = tables + slot * slot_sz + ((size_t)rp_i * nr + rs_i) * 4 * nz;
    // delta = -dz (zs >= zp, E_z negated) and delta = dz + 1 (zs < zp)
    row[nz - 1 - dzi] = (float)er;
    row[nz + dzi] = (float)er;
    row[2 * nz + nz - 1 - dzi] = (float)(-ez);
    row[2 * nz + nz + dzi] = (float)ez;
  }
}

// ---- charges -----------------------------------------------------------------

// q[p][rs * nz + zs] = particles in bin (zs, rs) of the p-th remeshing
// member; the histogram is stored [zs][rs] (DistributeParticleKernel's bin
// index).  Grid: nmem x nr*nz/256; workgroups past the remeshing count exit.
__global__ __launch_bounds__(kThreads) void k_gemm_pack(const GemmMember* __restrict__ mem,
                                                        const GemmState* __restrict__ gs,
                                                        const uint32_t* __restrict__ perm,
                                                        float* __restrict__ q, uint32_t nr,
                                                        uint32_t nz)
{
  const uint32_t p = blockIdx.x;
  if (p >= gs->nrem) return;
  const int64_t* bins = mem[perm[p]].bins;
  const uint32_t k = blockIdx.y * kThreads + threadIdx.x;
  if (k >= nr * nz) return;
  const uint32_t rs = k / nz, zs = k % nz;
  q[(size_t)p * nr * nz + k] = (float)((double)bins[zs * nr + rs] * 0x1p-36);
}

// ---- the grouped product -----------------------------------------------------

// Beam tile `tile` (in slot order) for NT beams per tile: its slot, first
// remeshing position and beam count.  Returns false past the last tile.
template <int NT>
__device__ __forceinline__ bool beam_tile(const uint32_t* start, const uint32_t* count,
                                          uint32_t tile, uint32_t& slot, uint32_t& p0,
                                          uint32_t& nb)
{
  for (uint32_t s = 0; s < kGemmSlots; ++s) {
    const uint32_t nt = ceil_div(count[s], NT);
    if (tile < nt) {
      slot = s;
      p0 = start[s] + tile * NT;
      nb = min((uint32_t)NT, count[s] - tile * NT);
      return true;
    }
    tile -= nt;
  }
  return false;
}

// Persistent grid over work units u = ((split * tiles) + tile) * mtiles + mt:
// row tile mt of the field rows, beam tile `tile`, K split `split` (source
// rings [split * nr / S, (split + 1) * nr / S)).  Row tiles 0 .. (nr+1)*nz/64
// - 1 are Toeplitz tiles (one rp, 64 consecutive zp, both components); the
// rest hold the last row zp = nz for 64 radii x 2 components each.  A wave
// computes 32 rows x NT beams (2 x NT/16 MFMA tiles).  Partial sums go to
// w[split][p][m], m = (rp * 2 + comp) * (nz + 1) + zp.
template <int NT>
__global__ __launch_bounds__(kThreads) void k_gemm_field(const GemmState* __restrict__ gs,
                                                         const float* __restrict__ tables,
                                                         const float* __restrict__ q,
                                                         float* __restrict__ w, uint32_t nr,
                                                         uint32_t nz, uint32_t splits)
{
  __shared__ uint32_t s_start[kGemmSlots], s_count[kGemmSlots];
  __shared__ float a_lds[kTileM * kLdsK];   // Toeplitz window (2 x kWin) or tail tile
  __shared__ float b_lds[NT * kLdsK];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  if (t < (int)kGemmSlots) { s_start[t] = gs->slot_start[t]; s_count[t] = gs->slot_count[t]; }
  __syncthreads();
  uint32_t ntiles = 0;
  for (uint32_t s = 0; s < kGemmSlots; ++s) ntiles += ceil_div(s_count[s], NT);
  const uint32_t zblocks = nz / kZBlock;
  const uint32_t toep = (nr + 1) * zblocks;
  const uint32_t mtiles = toep + ceil_div(2 * (nr + 1), kTileM);
  const uint32_t units = splits * ntiles * mtiles;
  const uint32_t M = 2 * (nr + 1) * (nz + 1);
  const uint32_t K = nr * nz;
  const size_t slot_sz = (size_t)(nr + 1) * nr * 4 * nz;
  const uint32_t nrem = gs->nrem;

  for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint32_t mt = u % mtiles;
    const uint32_t tile = (u / mtiles) % ntiles;
    const uint32_t split = u / (mtiles * ntiles);
    uint32_t slot = 0, p0 = 0, nb = 0;
    beam_tile<NT>(s_start, s_count, tile, slot, p0, nb);
    const float* tab = tables + slot * slot_sz;
    const bool is_toep = mt < toep;
    const uint32_t rp = is_toep ? mt / zblocks : 0;
    const uint32_t zp0 = is_toep ? (mt % zblocks) * kZBlock : 0;
    const uint32_t tt = is_toep ? 0 : mt - toep;
    const uint32_t rs0 = split * nr / splits, rs1 = (split + 1) * nr / splits;

    f4 acc[2][NT / 16];
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int g = 0; g < NT / 16; ++g) acc[f][g] = f4{0.f, 0.f, 0.f, 0.f};

    // K stages (source ring rs, 32 rows zs0..); for the Toeplitz tiles the
    // operands of the next stage are loaded into registers while this
    // stage's products run (tail tiles, at most two per product, stage
    // theirs directly)
    const uint32_t zst = nz / kTileK;
    const uint32_t nst = (rs1 - rs0) * zst;
    constexpr int kBPer = NT * kTileK / kThreads;           // charge entries per thread
    float ra = 0.f;
    float rb[kBPer];
    auto fetch = [&](uint32_t st) {
      const uint32_t rs = rs0 + st / zst, zs0 = (st % zst) * kTileK;
      if (is_toep && t < 2 * kWin) {
        // window o = zp0 - zs0 - 31 + j + nz - 1, j < 95, per component
        const float* src = tab + ((size_t)rp * nr + rs) * 4 * nz;
        const int base = (int)zp0 - (int)zs0 - 31 + (int)nz - 1;
        const int comp = t / kWin, j = t % kWin;
        ra = j < kWin - 1 ? src[comp * 2 * nz + base + j] : 0.f;
      }
#pragma unroll
      for (int i = 0; i < kBPer; ++i) {
        const int e = t + i * kThreads;
        const int j = e / kTileK, kk = e % kTileK;
        rb[i] = (uint32_t)j < nb ? q[(size_t)(p0 + j) * K + rs * nz + zs0 + kk] : 0.f;
      }
    };
    fetch(0);
    for (uint32_t st = 0; st < nst; ++st) {
      __syncthreads();   // the previous stage's products are done with the LDS tiles
      if (is_toep) {
        if (t < 2 * kWin) a_lds[t] = ra;
      } else {
        // tail rows: r = 2 (rp - 64 tt) + comp at zp = nz, o = 2 nz - 1 - zs
        const uint32_t rs = rs0 + st / zst, zs0 = (st % zst) * kTileK;
        for (int e = t; e < kTileM * kTileK; e += kThreads) {
          const int r = e / kTileK, kk = e % kTileK;
          const uint32_t rpt = tt * (kTileM / 2) + r / 2;
          a_lds[r * kLdsK + kk] =
              rpt <= nr ? tab[((size_t)rpt * nr + rs) * 4 * nz + (r & 1) * 2 * nz + 2 * nz - 1 -
                              zs0 - kk]
                        : 0.f;
        }
      }
#pragma unroll
      for (int i = 0; i < kBPer; ++i) {
        const int e = t + i * kThreads;
        b_lds[(e / kTileK) * kLdsK + e % kTileK] = rb[i];
      }
      __syncthreads();
      if (st + 1 < nst) fetch(st + 1);
      // 8 MFMA K steps of 4
#pragma unroll
      for (int k4 = 0; k4 < kTileK; k4 += 4) {
        const int kk = k4 + (lane >> 4);
        float af[2];
#pragma unroll
        for (int f = 0; f < 2; ++f) {
          const int row = wave * 32 + f * 16 + (lane & 15);
          if (is_toep) {
            const int comp = row / kZBlock, zl = row % kZBlock;
            af[f] = a_lds[comp * kWin + zl - kk + 31];
          } else {
            af[f] = a_lds[row * kLdsK + kk];
          }
        }
#pragma unroll
        for (int g = 0; g < NT / 16; ++g) {
          const float bf = b_lds[(g * 16 + (lane & 15)) * kLdsK + kk];
#pragma unroll
          for (int f = 0; f < 2; ++f) acc[f][g] = mfma(af[f], bf, acc[f][g]);
        }
      }
    }
    // store the partial sums of this unit
#pragma unroll
    for (int g = 0; g < NT / 16; ++g) {
      const uint32_t j = g * 16 + (lane & 15);
      if (j >= nb || p0 + j >= nrem) continue;
      float* out = w + ((size_t)split * nrem + p0 + j) * M;
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        const int row0 = wave * 32 + f * 16 + 4 * (lane >> 4);
        if (is_toep) {
          const int comp = row0 / kZBlock, zl = row0 % kZBlock;
          const uint32_t m = (rp * 2 + comp) * (nz + 1) + zp0 + zl;
#pragma unroll
          for (int i = 0; i < 4; ++i) out[m + i] = acc[f][g][i];
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int r = row0 + i;
            const uint32_t rpt = tt * (kTileM / 2) + r / 2;
            if (rpt <= nr) out[(rpt * 2 + (r & 1)) * (nz + 1) + nz] = acc[f][g][i];
          }
        }
      }
    }
  }
}

// ---- field tables --------------------------------------------------------------

// The exact field of member m at product row e before the c1 / dr^2
// factor: its three class columns col[0..2], each summed over the K splits
// in split order, with the Lagrange weights of its aspect.
__device__ __forceinline__ float exact_field(const float* __restrict__ w, uint32_t splits,
                                             uint32_t ncol, uint32_t M, const int32_t* col,
                                             const double (&wl)[3], uint32_t e)
{
  float f = 0.f;
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    float a = 0.f;
    for (uint32_t k = 0; k < splits; ++k) a += w[((size_t)k * ncol + col[h]) * M + e];
    f += (float)wl[h] * a;
  }
  return f;
}

// Field table of every remeshing member m: fld2[(zp (nr+1) + rp) 2 + comp]
// = c1 / dr^2 x (the exact field when the member had product columns, else
// the network's y[pos_nn[m]][zp][2 rp + comp] x Q x out_scale); c1 =
// 5721.67 I f_sc / (m f 1e6 N) as in UpdateFldTbl1Kernel.  Grid: nmem x
// points / 256.
__global__ __launch_bounds__(kThreads) void k_gemm_store(const GemmMember* __restrict__ mem,
                                                         const GemmState* __restrict__ gs,
                                                         MemberMap mm, const float* __restrict__ w,
                                                         const float* __restrict__ y,
                                                         const SurCtl* __restrict__ ctl,
                                                         uint32_t nr, uint32_t nz, uint32_t splits,
                                                         double lnr)
{
  const uint32_t m = blockIdx.x;
  const int32_t col[3] = {mm.pos_exact[3 * m], mm.pos_exact[3 * m + 1], mm.pos_exact[3 * m + 2]};
  const int32_t pn = mm.pos_nn[m];
  if (col[0] < 0 && pn < 0) return;
  const uint32_t e = blockIdx.y * kThreads + threadIdx.x;   // zp-fastest point index
  const uint32_t npts = (nr + 1) * (nz + 1);
  if (e >= npts) return;
  const uint32_t zp = e % (nz + 1), rp = e / (nz + 1);
  const GemmMember& mb = mem[m];
  const ScDev* sc = mb.sc;
  const double c1 = 5721.67 * (sc->current * sc->frac / mb.mass) /
                    (sc->freq * 1e6 * (double)mb.n);
  const double dr = sc->st.dr;
  const double scale = c1 / (dr * dr);
  const uint32_t M = 2 * (nr + 1) * (nz + 1);
  double wl[3];
  (void)aspect_lagrange(sc->log_aspect, lnr, wl);
  double v[2];
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
    const uint32_t row = (rp * 2 + comp) * (nz + 1) + zp;
    float s;
    if (col[0] >= 0) {
      s = exact_field(w, splits, gs->nrem, M, col, wl, row);
    } else {
      const uint32_t cpad = (2 * (nr + 1) + 15) / 16 * 16;
      s = y[((size_t)pn * (nz + kZPad) + zp) * cpad + rp * 2 + comp] * mm.qtot[m] * ctl->out_scale;
    }
    v[comp] = scale * (double)s;
  }
  double* out = mb.fld2 + 2 * ((size_t)zp * (nr + 1) + rp);
  out[0] = v[0];
  out[1] = v[1];
}

// ---- surrogate input, audit and training capture -------------------------------

// Total charge of member m's histogram (one workgroup, result in thread 0).
__device__ float member_charge(const int64_t* bins, uint32_t nbins, float* red)
{
  const int t = threadIdx.x;
  double q = 0.0;
  for (uint32_t k = t; k < nbins; k += kThreads) q += (double)bins[k] * 0x1p-36;
  red[t] = (float)q;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  return red[0];
}

// Network input of every member with a network position: x[pos][z][rs] =
// n(z, rs) nr nz / Q (the histogram is stored [zs][rs]), the padding rows
// z >= nz left zero; u = (class log(1+step) - mu) / sigma; qtot[m] = Q.
__global__ __launch_bounds__(kThreads) void k_sur_input(const GemmMember* __restrict__ mem,
                                                        MemberMap mm, const SurCtl* __restrict__ ctl,
                                                        float* __restrict__ x, float* __restrict__ u,
                                                        uint32_t nr, uint32_t nz)
{
  __shared__ float red[kThreads];
  const uint32_t m = blockIdx.x;
  const int32_t pn = mm.pos_nn[m];
  if (pn < 0) return;
  const int64_t* bins = mem[m].bins;
  const uint32_t nb = nr * nz;
  const float q = member_charge(bins, nb, red);
  const float k = q > 0.f ? (float)nb / q : 0.f;
  float* row = x + (size_t)pn * (nz + kZPad) * nr;
  for (uint32_t e = threadIdx.x; e < nb; e += kThreads)
    row[e] = (float)((double)bins[e] * 0x1p-36) * k;
  if (threadIdx.x == 0) {
    mm.qtot[m] = q;
    u[pn] = ((float)mem[m].sc->log_aspect - ctl->u_mu) * ctl->inv_u_sigma;
  }
}

// Audit of the network against the exact product (members that have both):
// relative L2 error of the field; above the bound the member falls back to
// the exact product until an audit passes.
__global__ __launch_bounds__(kThreads) void k_sur_audit(const GemmMember* __restrict__ mem,
                                                        const GemmState* __restrict__ gs,
                                                        MemberMap mm, const float* __restrict__ w,
                                                        const float* __restrict__ y, SurCtl* ctl,
                                                        uint32_t nr, uint32_t nz, uint32_t splits)
{
  __shared__ float red[2][kThreads];
  if (!ctl->audit_now) return;
  const uint32_t m = blockIdx.x;
  const int32_t col[3] = {mm.pos_exact[3 * m], mm.pos_exact[3 * m + 1], mm.pos_exact[3 * m + 2]};
  const int32_t pn = mm.pos_nn[m];
  if (col[0] < 0 || pn < 0) return;
  const int t = threadIdx.x;
  const uint32_t M = 2 * (nr + 1) * (nz + 1), ncol = gs->nrem;
  const uint32_t cpad = (2 * (nr + 1) + 15) / 16 * 16;
  const float qs = mm.qtot[m] * ctl->out_scale;
  double wl[3];
  (void)aspect_lagrange(mem[m].sc->log_aspect, ctl->lnr, wl);
  float ne = 0.f, nd = 0.f;
  for (uint32_t e = t; e < M; e += kThreads) {
    const uint32_t zp = e % (nz + 1), rc = e / (nz + 1);   // rc = 2 rp + comp
    const float ex = exact_field(w, splits, ncol, M, col, wl, e);
    const float nv = y[((size_t)pn * (nz + kZPad) + zp) * cpad + rc] * qs;
    ne += ex * ex;
    nd += (nv - ex) * (nv - ex);
  }
  red[0][t] = ne;
  red[1][t] = nd;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (t < o) { red[0][t] += red[0][t + o]; red[1][t] += red[1][t + o]; }
    __syncthreads();
  }
  if (t == 0) {
    const float err = red[0][0] > 0.f ? sqrtf(red[1][0] / red[0][0]) : 0.f;
    const uint32_t fail = err > ctl->err_bound ? 1u : 0u;
    mm.fallback[m] = fail;
    atomicAdd(&ctl->audited, 1u);
    if (fail) atomicAdd(&ctl->audit_fail, 1u);
    atomicMax((uint32_t*)&ctl->max_err, __float_as_uint(err));   // err >= 0: bits order as floats
  }
}

// Training pairs from the exact products of this kick point: for member m
// with product columns, x = n(z, rs) nr nz / Q ([nz][nr]), y = exact field /
// Q ([nz+1][2 (nr+1)], before c1 / dr^2), u = log aspect, appended at the
// device counter (while the capacity lasts); one workgroup per member.
__global__ __launch_bounds__(kThreads) void k_capture(const GemmMember* __restrict__ mem,
                                                      const GemmState* __restrict__ gs,
                                                      MemberMap mm, const float* __restrict__ w,
                                                      NetData d, double lnr, uint32_t nr,
                                                      uint32_t nz, uint32_t splits)
{
  __shared__ float red[kThreads];
  __shared__ uint32_t slot;
  const uint32_t m = blockIdx.x;
  const int32_t col[3] = {mm.pos_exact[3 * m], mm.pos_exact[3 * m + 1], mm.pos_exact[3 * m + 2]};
  if (col[0] < 0) return;
  if (threadIdx.x == 0) slot = atomicAdd(d.count, 1u);
  __syncthreads();
  if (slot >= d.cap) return;
  const int64_t* bins = mem[m].bins;
  const uint32_t nb = nr * nz, M = 2 * (nr + 1) * (nz + 1), cout = 2 * (nr + 1);
  const float q = member_charge(bins, nb, red);
  const float k = q > 0.f ? (float)nb / q : 0.f, iq = q > 0.f ? 1.f / q : 0.f;
  double wl[3];
  (void)aspect_lagrange(mem[m].sc->log_aspect, lnr, wl);
  float* xr = d.x + (size_t)slot * nb;
  for (uint32_t e = threadIdx.x; e < nb; e += kThreads)
    xr[e] = (float)((double)bins[e] * 0x1p-36) * k;
  float* yr = d.y + (size_t)slot * M;
  for (uint32_t e = threadIdx.x; e < M; e += kThreads) {
    const uint32_t zp = e / cout, rc = e % cout;
    yr[e] = exact_field(w, splits, gs->nrem, M, col, wl, rc * (nz + 1) + zp) * iq;
  }
  if (threadIdx.x == 0) d.u[slot] = (float)mem[m].sc->log_aspect;
}

__global__ void k_capture_clamp(NetData d)
{
  if (threadIdx.x == 0 && d.count[0] > d.cap) d.count[0] = d.cap;
}

// ---- launchers -------------------------------------------------------------------

uint32_t gemm_shapes() { return kGemmShapes; }

GemmShape gemm_shape(uint32_t i)
{
  // (beams per tile, K splits, workgroups per CU of the persistent grid,
  // spectral): eight shapes of the direct product, then the product in the
  // frequency domain along z
  static const GemmShape shapes[kGemmShapes] = {
      {64, 1, 2, 0}, {64, 2, 2, 0}, {64, 4, 2, 0}, {32, 1, 2, 0}, {32, 2, 2, 0}, {32, 4, 2, 0},
      {64, 2, 4, 0}, {32, 4, 4, 0}, {32, 1, 4, 1}};
  return shapes[i < kGemmShapes ? i : 0];
}

hipError_t launch_gemm_prepare(const GemmMember* mem, uint32_t nmem, GemmState* gs,
                               uint32_t* perm, MemberMap mm, SurCtl* ctl, float* tables, float* q,
                               uint32_t nr, uint32_t nz, double lnr, uint32_t cus, hipStream_t s)
{
  hipLaunchKernelGGL(k_gemm_schedule, dim3(1), dim3(kThreads), 0, s, mem, nmem, gs, perm, mm,
                     ctl, lnr);
  hipLaunchKernelGGL(k_gemm_tables, dim3(cus), dim3(kThreads), 0, s, gs, tables, nr, nz, lnr);
  hipLaunchKernelGGL(k_gemm_pack, dim3(kGemmCols * nmem, ceil_div(nr * nz, kThreads)),
                     dim3(kThreads), 0, s, mem, gs, perm, q, nr, nz);
  return hipGetLastError();
}

hipError_t launch_gemm_store(const GemmMember* mem, uint32_t nmem, const GemmState* gs,
                             MemberMap mm, const float* w, const float* y, const SurCtl* ctl,
                             uint32_t nr, uint32_t nz, double lnr, uint32_t shape, hipStream_t s)
{
  hipLaunchKernelGGL(k_gemm_store, dim3(nmem, ceil_div((nr + 1) * (nz + 1), kThreads)),
                     dim3(kThreads), 0, s, mem, gs, mm, w, y, ctl, nr, nz,
                     gemm_shape(shape).splits, lnr);
  return hipGetLastError();
}

hipError_t launch_sur_input(const GemmMember* mem, uint32_t nmem, MemberMap mm, const SurCtl* ctl,
                            float* x, float* u, uint32_t nr, uint32_t nz, hipStream_t s)
{
  hipLaunchKernelGGL(k_sur_input, dim3(nmem), dim3(kThreads), 0, s, mem, mm, ctl, x, u, nr, nz);
  return hipGetLastError();
}

hipError_t launch_sur_audit(const GemmMember* mem, uint32_t nmem, const GemmState* gs,
                            MemberMap mm, const float* w, const float* y, SurCtl* ctl,
                            uint32_t nr, uint32_t nz, uint32_t shape, hipStream_t s)
{
  hipLaunchKernelGGL(k_sur_audit, dim3(nmem), dim3(kThreads), 0, s, mem, gs, mm, w, y, ctl, nr,
                     nz, gemm_shape(shape).splits);
  return hipGetLastError();
}

hipError_t launch_capture(const GemmMember* mem, uint32_t nmem, const GemmState* gs, MemberMap mm,
                          const float* w, const NetData& d, double lnr, uint32_t nr, uint32_t nz,
                          uint32_t shape, hipStream_t s)
{
  hipLaunchKernelGGL(k_capture, dim3(nmem), dim3(kThreads), 0, s, mem, gs, mm, w, d, lnr, nr, nz,
                     gemm_shape(shape).splits);
  hipLaunchKernelGGL(k_capture_clamp, dim3(1), dim3(64), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_gemm_product(const GemmState* gs, const float* tables, const float* q, float* w,
                               uint32_t nr, uint32_t nz, uint32_t shape, uint32_t cus,
                               const FftBufs* fb, hipStream_t s)
{
  const GemmShape sh = gemm_shape(shape);
  if (sh.spectral) {
    if (!fb) return hipErrorInvalidValue;
    return launch_fft_product(gs, tables, q, w, *fb, nr, nz, cus, s);
  }
  const dim3 grid(cus * sh.per_cu), block(kThreads);
  if (sh.nt == 64)
    hipLaunchKernelGGL(k_gemm_field<64>, grid, block, 0, s, gs, tables, q, w, nr, nz, sh.splits);
  else
    hipLaunchKernelGGL(k_gemm_field<32>, grid, block, 0, s, gs, tables, q, w, nr, nz, sh.splits);
  return hipGetLastError();
}

}  // namespace hps
