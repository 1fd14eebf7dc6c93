// hps_push_fast.hip -- the tracking kernels of hps_kernels.hip built a
// second time, with reciprocal division and square roots from v_rsq_f64,
// for batches in the grouped field
// solve (field mode 1, whose field comes from the FP32 product or from the
// surrogate attached to it).
//
// Built with -fapprox-func and -freciprocal-math, an FP64 quotient x / y
// becomes v_rcp_f64(y), two Newton steps, x times the refined reciprocal and
// one FMA correction (7 instructions instead of 11).  It skips the
// v_div_scale / v_div_fmas / v_div_fixup steps of the correctly rounded
// sequence, and quotients by one denominator share its reciprocal (x times
// the rounded 1 / y).  With HPS_FAST_PUSH defined, the square roots of the
// maps, the kick and the deposit (psqrt, hps_math.hpp) come from v_rsq_f64
// with one Goldschmidt step and one Newton correction (12 instructions
// instead of 21).  Measured on the device, both gave the correctly rounded
// result for every one of 400,000 random quotients and roots between
// 10^-200 and 10^200.  A zero or infinite denominator gives NaN instead of
// an infinity or zero.  The push guards the denominators that can be zero:
// r and the gap radius on the axis, and bgzf <= 0 is a loss.  Every other
// denominator is a mass, a mesh spacing, or a velocity or energy that is
// positive for any particle still alive.  Static VALU instruction counts of
// the batched push drop by 12 % (quad class) and 15 % (RF-gap class), and
// by 10 % in the batched deposit.
//
// The batched push of this build visits only the live particles of each
// chunk (segment_body_live in hps_kernels.hip).  It adds each chunk's sums
// in another order than the canonical one.
//
// These modes already differ from the FP64 solve: each member's field comes
// from an FP32 product, within about 1e-6 of the FP64 field.  A last-bit
// change in a division is ten orders of magnitude below that.  Their results
// are held to error bounds against the FP64 oracle, not to its bits.  The
// FP64 field mode and single beams, which are bit-exact against the oracle,
// launch the kernels of hps_kernels.hip as built there.  Index arithmetic
// that takes a quotient (mesh cell, table row) keeps its range checks and
// clamps, so a quotient one unit off at a cell boundary selects a
// neighbouring cell, never an address outside the tables.
#define HPS_FAST_PUSH 1
#include "hps_kernels.hip"
