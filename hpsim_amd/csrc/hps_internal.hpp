// hps_internal.hpp -- device data layout and kernel launch interface shared by
// the kernels (hps_kernels.hip) and the host engine (hps_engine.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "hpsim_amd.h"
#include "hps_device.hpp"

namespace hps {

// Structure-of-arrays particle state in HBM (reference beam.h:156-174).
struct BeamArrays {
  double *x, *xp, *y, *yp, *phi, *phi_r, *w;
  uint32_t *loss, *lloss;
  uint32_t n;
  double mass, charge;
};

// Device-resident SCHEFF state: the SpaceCharge members the reference keeps
// on the host (space_charge.h:83-124) plus the values of the last control
// step (Scheff::Start, space_charge.cpp:82-139).
struct ScDev {
  hps_spch_state st;
  int32_t remesh;       // last control step rebuilt the tables
  int32_t valid;        // tables hold a field
  double beta, freq, frac, current;
  double ratio_r, ratio_z, ratio_q, ratio_g;
  // kick constants of ApplySpchKickKernel (space_charge_kernel.cu:578-640)
  // that the control step fixes, computed once per kick point
  double k_xbar, k_ybar, k_epsq, k_gmg, k_dr, k_dzg, k_hl, k_c1, k_c2, k_bgmg;
  double k_xfac, k_yfac, k_er_ratio, k_ez_ratio;
  double k_pl;          // bunch spacing beta*gamma*lambda (adjacent-bunch images)
  uint32_t adj_eff;     // adjacent bunches each side at this kick point
  uint32_t pad_;
  double log_aspect;    // log(dz * gamma / dr) of the last remesh (grouped field solve)
};

// Scratch for the canonical reductions: per-chunk partial sums, padded to a
// power of two with zeros.
struct Partials {
  double* sums;        // [kMaxQ][mpad]
  uint32_t* cnt;       // [3][mpad]
  double* pmax;        // [mpad]
  double* rmax2;       // [mpad] largest r^2 about (x_avg, y_avg), loss == 0
  double* out;         // [kMaxQ] final sums for host-side getters
  uint32_t* out_cnt;   // [1]
  uint32_t mpad;       // NextPow2(nchunks)
  uint32_t nch;        // chunks
};
constexpr int kMaxQ = 8;

enum StatKind : int32_t { STAT_NONE = 0, STAT_SC = 1, STAT_APERTURE = 2 };

struct SegArgs {
  BeamArrays b;
  hps_beam_stats* st;
  Partials pr;
  const Op* ops;
  int32_t op_begin, op_end;
  int32_t kick;            // apply the pending SCHEFF kick first
  uint32_t kick_mask;      // particles i with (i & mask) == 0 are not kicked (stride - 1)
  double kick_len;
  const ScDev* sc;
  const double* fld2;      // (nz+1)*(nr+1) double2, [zp][rp]
  int32_t stats;           // StatKind computed after the maps
  int32_t op_class;        // OpClass: heaviest map in [op_begin, op_end)
  int32_t rmax;            // STAT_SC also keeps the per-chunk largest r^2
};

struct PhaseArgs {
  BeamArrays b;
  hps_beam_stats* st;
  Partials pr;
  ScDev* sc;
  int64_t* bins;
  int32_t do_sums;         // n > 2 (UpdateStatForSpaceCharge guard)
  int32_t rmax;            // the segment kept per-chunk r^2 maxima (adj_bunch > 0)
  double frac;             // current fraction at this kick point
  double current;          // beam current (A)
  double freq;             // beam frequency (MHz)
};

// Arguments of the SCHEFF rebuild and field launches.
struct BuildArgs {
  ScDev* sc;
  double* fld1;
  BeamArrays b;
  const hps_beam_stats* st;
  int64_t* bins;
  uint32_t greens_blocks;   // leading workgroups that fill the Green's table
  uint32_t dep_blocks;      // deposit workgroups: after them, or the first dep_blocks
                            // when a lone beam has dep_blocks >= greens_blocks
                            // (0: empty deposit)
};

// Workgroups of one simulation's rebuild grid (k_sc_build): dep_blocks when
// the deposit covers the Green's table (its first workgroups fill the table),
// else the table workgroups followed by the deposit workgroups.  The host
// builds the batched block map from it and the kernel splits the roles by it.
__host__ __device__ __forceinline__ uint32_t sc_build_blocks(const BuildArgs& a)
{
  return a.dep_blocks >= a.greens_blocks ? a.dep_blocks : a.greens_blocks + a.dep_blocks;
}

struct FieldArgs {
  const ScDev* sc;
  const int64_t* bins;
  const double* fld1;
  double* fld2;
  double* part;     // kFieldMaxSplit * (nz+1)*(nr+1) double2: split partial sums
};
// The field sum may split its partial sums over up to this many workgroup
// sets (a lone simulation, to use more CUs); part holds their results.
constexpr uint32_t kFieldMaxSplit = 8;

// Batched launches cover the blocks of several simulations in one grid; a
// block map entry packs the simulation (high bits) and its local block.
constexpr uint32_t kBatchLocalBits = 20;
constexpr uint32_t kBatchLocalMask = (1u << kBatchLocalBits) - 1;

// Launchers (hps_kernels.hip).
hipError_t launch_segment(const SegArgs& a, hipStream_t s);
hipError_t launch_seg_final(const SegArgs& a, hipStream_t s);
hipError_t launch_phase(const PhaseArgs& a, hipStream_t s);
hipError_t launch_phase_final(const PhaseArgs& a, hipStream_t s);
// k_sc_build workgroup size (deposit: four particles per thread at full
// grid), the most deposit workgroups, and the LDS a deposit workgroup uses
// besides its histogram.
constexpr uint32_t kDepositBlock = 1024;
constexpr uint32_t kDepositMaxBlocks = 256;
constexpr uint32_t kDepositMaxBlocksBatch = 128;
constexpr uint32_t kDepositMaxBlocksBig = 32;
constexpr size_t kBlockLdsReserve = 4 * kDepositBlock;
// deposit workgroups for n particles, per_thread particles per thread (at most
// kDepositMaxBlocks workgroups)
uint32_t deposit_blocks(uint32_t n, uint32_t per_thread);
uint32_t deposit_blocks_batch(uint32_t n, uint32_t nr, uint32_t nz);
// Green's table + charge deposit (deposit = false for beams whose reference
// deposit launch is empty).
hipError_t launch_sc_build(ScDev* sc, double* fld1, uint32_t nr, uint32_t nz, BeamArrays b,
                           const hps_beam_stats* st, int64_t* bins, bool deposit,
                           hipStream_t s);
hipError_t launch_field(const FieldArgs& a, uint32_t nr, uint32_t nz, size_t max_lds,
                        hipStream_t s);
uint32_t greens_blocks(uint32_t nr, uint32_t nz);

// Batched forms: recs holds one argument record per simulation (the segment
// and phase kernels also a block map over their chunks).
// The kernels read their records at recs + roff[0] bytes (a device word:
// zero for eager launches, the lattice period's offset in a period graph).
hipError_t launch_segment_batch(const SegArgs* recs, const uint32_t* map, uint32_t nblocks,
                                int op_class, const uint64_t* roff, hipStream_t s);
hipError_t launch_seg_final_batch(const SegArgs* recs, uint32_t nmem, const uint64_t* roff,
                                  hipStream_t s);
hipError_t launch_phase_batch(const PhaseArgs* recs, const uint32_t* map, uint32_t nblocks,
                              const uint64_t* roff, hipStream_t s);
hipError_t launch_phase_final_batch(const PhaseArgs* recs, uint32_t nmem, const uint64_t* roff,
                                    hipStream_t s);
// roff[0] = v, or roff[0] += v when add
hipError_t launch_roff(uint64_t* roff, uint64_t v, bool add, hipStream_t s);
hipError_t launch_sc_build_batch(const BuildArgs* recs, const uint32_t* map, uint32_t nblocks,
                                 uint32_t nr, uint32_t nz, hipStream_t s);
// The same batched push and rebuild with reciprocal FP64 division
// (hps_push_fast.hip), for batches in the grouped field solve (field mode 1,
// with or without a surrogate), whose results are held to error bounds rather
// than to the oracle's bits.
namespace fast {
hipError_t launch_segment_batch(const SegArgs* recs, const uint32_t* map, uint32_t nblocks,
                                int op_class, const uint64_t* roff, hipStream_t s);
hipError_t launch_sc_build_batch(const BuildArgs* recs, const uint32_t* map, uint32_t nblocks,
                                 uint32_t nr, uint32_t nz, hipStream_t s);
hipError_t launch_probe(int fn, const double* x, const double* y, double* out, uint32_t n,
                        hipStream_t s);
}  // namespace fast
// Field sum of a batch in launch shape `shape` < field_batch_shapes() (any
// larger value: the default for the batch size); every shape gives the same
// bits, so a caller may time them and keep the fastest.
uint32_t field_batch_shapes();
uint32_t field_batch_default_shape(uint32_t nr, uint32_t nmem);
hipError_t launch_field_batch(const FieldArgs* recs, uint32_t nmem, uint32_t nr, uint32_t nz,
                              size_t max_lds, uint32_t shape, hipStream_t s);

// Grouped field solve of a batch (hps_field_gemm.hip): per-member record,
// device state, launch shapes.
struct GemmMember {
  const ScDev* sc;
  const int64_t* bins;     // fixed-point histogram [zs][rs]
  double* fld2;            // field table [zp][rp] double2
  uint32_t n;              // beam size (the c1 denominator)
  uint32_t pad_;
  double mass;
};
// class tables: direct-mapped cache, class c in slot c mod kGemmSlots
constexpr uint32_t kGemmSlots = 64;
constexpr uint32_t kGemmMaxMembers = 4096;
constexpr uint32_t kGemmCols = 3;   // product columns per member: classes c - 1, c, c + 1
struct GemmState {
  int32_t slot_class[kGemmSlots];   // class whose table the slot holds (INT32_MIN: none)
  uint32_t slot_start[kGemmSlots];  // first remeshing position of the slot's members
  uint32_t slot_count[kGemmSlots];
  uint32_t build[kGemmSlots];       // slots to build at this kick point
  uint32_t nbuild;
  uint32_t nrem;                    // product columns at this kick point (3 per member)
  uint32_t conflicts;               // members that used another class's table (cumulative)
  uint32_t builds_total;            // class tables built (cumulative)
  uint64_t exact_total;             // product columns so far (cumulative)
};
struct GemmShape {
  uint32_t nt;        // beams per tile (32 or 64)
  uint32_t splits;    // K splits (source-ring ranges)
  uint32_t per_cu;    // persistent workgroups per CU
  uint32_t spectral;  // 1: the product in the frequency domain along z (hps_field_fft.hip)
};
constexpr uint32_t kGemmShapes = 9;
uint32_t gemm_shapes();
GemmShape gemm_shape(uint32_t i);
constexpr uint32_t kGemmMaxSplits = 4;
// zero rows after the nz bins in the surrogate network's input and output
constexpr uint32_t kZPad = 8;
// Per-member positions of one kick point: its product columns (classes
// c - 1, c, c + 1: pos_exact[3 m + h]) and its place in the network's batch
// (-1: none), the audit verdict, the total charge.
struct MemberMap {
  int32_t* pos_exact;
  int32_t* pos_nn;
  uint32_t* fallback;   // last audit failed: exact product until one passes
  float* qtot;
};
// Device control of the surrogate inside a batch (null: no surrogate):
// configuration (uploaded at every simulate), then running counters.
struct SurCtl {
  uint32_t audit_every;          // audit at every audit_every-th kick point (0: never)...
  float err_bound;               // ... with this relative L2 field error bound
  float out_scale, u_mu, inv_u_sigma;
  uint32_t pad_;
  double lnr;                    // log(1 + aspect step)
  uint32_t kick;                 // kick points so far
  uint32_t pending;              // an audit is due at the next kick point with remeshes
  uint32_t audit_now;
  uint32_t nn_n;                 // members the network serves at this kick point
  uint32_t audited, audit_fail;  // member audits, failed member audits
  uint32_t nn_member_kicks;      // member remeshes served by the network...
  uint32_t exact_member_kicks;   // ... and by the exact product
  float max_err;                 // largest audited relative L2 field error
  uint32_t pad2_;
};
constexpr size_t kSurCtlConfigBytes = 32;   // audit_every .. lnr
static_assert(offsetof(SurCtl, kick) == kSurCtlConfigBytes, "SurCtl configuration block");
// one kick point: prepare (schedule, class tables to build, charges), the
// product in launch shape `shape` (all shapes give the same field up to the
// FP32 summation order of their K splits), store (field tables; y: the
// network's output for members without an exact product)
hipError_t launch_gemm_prepare(const GemmMember* mem, uint32_t nmem, GemmState* gs,
                               uint32_t* perm, MemberMap mm, SurCtl* ctl, float* tables, float* q,
                               uint32_t nr, uint32_t nz, double lnr, uint32_t cus, hipStream_t s);
// Buffers of the spectral product (one K split): twiddles exp(-2 pi i k / L)
// for k < L / 2 (L = 2 nz), the slots' table spectra [slot][f][row][rs]
// (f = 0 .. nz, row = 2 rp + comp), the class each slot's spectrum belongs
// to, a worklist of slots to transform, and per product column the charge
// spectra [p][f][rs] and field spectra [p][f][row].
struct FftBufs {
  float2* tw;
  float2* shat;
  int32_t* spec_class;
  uint32_t* list;
  float2* qhat;
  float2* what;
};
bool fft_product_supported(uint32_t nr, uint32_t nz);
hipError_t launch_fft_product(const GemmState* gs, const float* tables, const float* q, float* w,
                              const FftBufs& fb, uint32_t nr, uint32_t nz, uint32_t cus,
                              hipStream_t s);
// fb: null when the spectral shape is not available (then only the direct
// shapes may be asked for)
hipError_t launch_gemm_product(const GemmState* gs, const float* tables, const float* q, float* w,
                               uint32_t nr, uint32_t nz, uint32_t shape, uint32_t cus,
                               const FftBufs* fb, hipStream_t s);
hipError_t launch_gemm_store(const GemmMember* mem, uint32_t nmem, const GemmState* gs,
                             MemberMap mm, const float* w, const float* y, const SurCtl* ctl,
                             uint32_t nr, uint32_t nz, double lnr, uint32_t shape, hipStream_t s);
// surrogate: network inputs of this kick point, audit against the exact product
hipError_t launch_sur_input(const GemmMember* mem, uint32_t nmem, MemberMap mm, const SurCtl* ctl,
                            float* x, float* u, uint32_t nr, uint32_t nz, hipStream_t s);
hipError_t launch_sur_audit(const GemmMember* mem, uint32_t nmem, const GemmState* gs,
                            MemberMap mm, const float* w, const float* y, SurCtl* ctl,
                            uint32_t nr, uint32_t nz, uint32_t shape, hipStream_t s);
struct NetData;
// training pairs (charge, exact field) of this kick point into a dataset
hipError_t launch_capture(const GemmMember* mem, uint32_t nmem, const GemmState* gs, MemberMap mm,
                          const float* w, const NetData& d, double lnr, uint32_t nr, uint32_t nz,
                          uint32_t shape, hipStream_t s);

// Generic canonical sums for the getters: over the filtered particles
// (filter 0: loss == 0, 1: loss + lloss == 0, 2: all) of a, a*a, b, b*b, a*b,
// and the count of loss != 0, of (loss+lloss) == 0 and of lloss != 0.
// Results land in pr.out[0..4] and pr.out_cnt[0..2].  With rel_ref, a is
// replaced by its relative phase against rel_ref[0] (k_relative_phi's value).
hipError_t launch_sums(BeamArrays b, const double* a, const double* bb, int filter,
                       Partials pr, hipStream_t s, const double* rel_ref = nullptr);
hipError_t launch_relative_phi(BeamArrays b, const double* ref, hipStream_t s);
hipError_t launch_lloss_flag(BeamArrays b, const double* avg, hipStream_t s);
hipError_t launch_shift(double* a, uint32_t n, double v, hipStream_t s);
hipError_t launch_scale(double* a, uint32_t n, double v, hipStream_t s);
hipError_t launch_cut(const double* a, uint32_t* loss, uint32_t n, double lo, double hi,
                      hipStream_t s);
// Truncated-Gaussian synthetic beam (Philox4x32-10, Box-Muller, |z| <= 3)
struct GaussSpec {
  double center[6];   // x, xp, y, yp, phi, w
  double sigma[6];
};
hipError_t launch_sample_gaussian(BeamArrays b, uint32_t seed, const GaussSpec& g, hipStream_t s);
hipError_t launch_probe(int fn, const double* x, const double* y, double* out, uint32_t n,
                        hipStream_t s);

}  // namespace hps
