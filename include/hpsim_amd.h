/*
 * hpsim_amd.h -- C ABI of the MI355X HPSim tracking engine.
 *
 * The reference exposes its engine to Python through a hand-written CPython
 * extension module "HPSim" (src/pywrapper/src/hpsim_module.cpp:398-416) whose
 * methods call C++ objects Beam, BeamLine, SpaceCharge and SimulationEngine.
 * This header is the language-neutral boundary that replaces those calls:
 * plain handles, pointers and sizes, an int status on every call (0 = ok,
 * negative = error, message from hps_last_error()).  hpsim_amd/HPSim.py binds
 * it with ctypes and keeps the reference's Python types and method names.
 *
 * Units are the engine units of the reference: m, rad, MeV, MHz, A.
 * All arrays are host pointers; the library owns device memory.
 */
#ifndef HPSIM_AMD_H
#define HPSIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hps_beam hps_beam;
typedef struct hps_beamline hps_beamline;
typedef struct hps_spch hps_spch;
typedef struct hps_sim hps_sim;

/* Element kinds: the reference element type strings
 * (src/beamline/inc/beamline_element.h:46-263). */
enum hps_element_kind {
  HPS_DRIFT = 1,        /* "Drift"       */
  HPS_QUAD = 2,         /* "Quad"        */
  HPS_GAP_DTL = 3,      /* "RFGap-DTL"   */
  HPS_GAP_CCL = 4,      /* "RFGap-CCL"   */
  HPS_APERTURE_C = 5,   /* "ApertureC"   */
  HPS_APERTURE_R = 6,   /* "ApertureR"   */
  HPS_SPCH_COMP = 7,    /* "SpchComp"    */
  HPS_DIAGNOSTICS = 8,  /* "Diagnostics" */
  HPS_ROTATION = 9,     /* "Rotation"    */
  HPS_STEERER = 10,     /* "Steerer"     */
  HPS_BUNCHER = 11,     /* "Buncher"     */
  HPS_DIPOLE = 12       /* "Dipole"      */
};

/* RF gap parameters, field for field RFGapParameter
 * (src/beamline/inc/rfgap_parameter.h:4-32). */
typedef struct {
  double length, frequency, cell_length_over_beta_lambda, amplitude;
  double phase_ref, phase_shift, energy_out, beta_center, dg, t, tp, sp;
  double fit_beta_center, fit_beta_min;
  double fit_t0, fit_t1, fit_t2, fit_t3, fit_t4, fit_t5;
  double fit_s1, fit_s2, fit_s3, fit_s4, fit_s5;
} hps_gap_param;

/* One beamline element: the union of the element classes' host fields
 * (beamline_element.h); a kind reads only its own fields. */
typedef struct {
  int32_t kind;            /* enum hps_element_kind */
  int32_t aperture_in;     /* ApertureC/ApertureR::IsIn() */
  double length;           /* m */
  double aperture;         /* m, 0 = none */
  double gradient;         /* Quad, T/m */
  double angle;            /* Rotation, rad */
  double bl_h, bl_v;       /* Steerer integrated fields, T*m */
  double aper_xl, aper_xr, aper_yt, aper_yb;  /* ApertureR, m */
  double fraction;         /* SpchComp */
  hps_gap_param gap;       /* RFGap-DTL / RFGap-CCL */
  /* Buncher (beamline_element.h:83-107): voltage MV, frequency MHz, phase rad */
  double voltage, frequency, phase;
  /* Dipole (dipole_parameter in beamline_element.h:115-150); the bend angle
   * is `angle` above */
  double rho, edge_angle_in, edge_angle_out, half_gap, k1, k2, field_index,
         kinetic_energy;
  int32_t on, pad_;        /* Buncher::IsOn() */
} hps_element;

/* Scalar beam statistics: the device scalars of the reference Beam
 * (src/beam/inc/beam.h:175-263). */
typedef struct {
  double x_avg, x_avg_good, xp_avg, y_avg, y_avg_good, yp_avg;
  double phi_avg, phi_avg_r, phi_avg_good, w_avg, w_avg_good;
  double x_sig, x_sig_good, xp_sig, y_sig, y_sig_good, yp_sig;
  double phi_sig, phi_sig_r, phi_sig_good, w_sig, w_sig_good;
  double x_emit, y_emit, z_emit;
  uint32_t num_loss, num_lloss, num_good, pad_;
} hps_beam_stats;

/* Space-charge state: SpaceCharge members (src/simulation/inc/space_charge.h:83-124)
 * plus the SCHEFF mesh spacing (space_charge_kernel.cu:12-13). */
typedef struct {
  uint32_t nx, nz, nx_init, nz_init, adj_bunch, good_cnt_old;
  double fraction, interval, interval_init, adj_bunch_w, mesh_w;
  double sigr_old, sigz_old, gm_old, remesh_val;
  double dr, dz;
  uint32_t num_on_mesh, remesh_count, kick_count, pad_;
} hps_spch_state;

const char* hps_last_error(void);

/* Device selection; replaces set_gpu / SetGPU (src/beamline/src/init.cpp:12-18). */
int hps_set_device(int device);

/* ---- Beam (src/pywrapper/src/wrap_beam.cpp) ---- */
/* Beam(mass=, charge=, current=, num=), wrap_beam.cpp:26-50 / beam.cpp:124-134 */
int hps_beam_create(hps_beam** out, uint32_t num, double mass, double charge, double current);
int hps_beam_destroy(hps_beam* b);
/* set_distribution(x, xp, y, yp, phi, w[, loss[, lloss]]), wrap_beam.cpp:61-118,
 * beam.cpp:317-368.  loss / lloss may be NULL (zeros). */
int hps_beam_set_distribution(hps_beam* b, const double* x, const double* xp,
    const double* y, const double* yp, const double* phi, const double* w,
    const uint32_t* loss, const uint32_t* lloss);
/* set_waterbag(ax, bx, ex, ay, by, ey, az, bz, ez, phi, w, freq[, seed]),
 * wrap_beam.cpp:125-144 / Beam::InitWaterbagBeam beam.cpp:391-463: PARMILA
 * Twiss input (beta in cm, emittance in cm*rad), 6D waterbag drawn with the C
 * library's srand(seed)/rand (seed 0 = time), sets the beam frequency. */
int hps_beam_set_waterbag(hps_beam* b, double ax, double bx, double ex, double ay, double by,
    double ey, double az, double bz, double ez, double sync_phi, double sync_w, double freq,
    uint32_t seed);
/* set_dc(ax, bx, ex, ay, by, ey, dphi, sync_phi, sync_w[, seed]),
 * wrap_beam.cpp:146-165 / Beam::InitDCBeam beam.cpp:480-537: 4D transverse
 * waterbag, phase a linear ramp over [-dphi, dphi], w = sync_w. */
int hps_beam_set_dc(hps_beam* b, double ax, double bx, double ex, double ay, double by,
    double ey, double dphi, double sync_phi, double sync_w, uint32_t seed);
/* A synthetic beam drawn on the device: coordinate c (x, xp, y, yp, phi, w)
 * of particle i = center[c] + sigma[c] * z, z a standard normal truncated to
 * |z| <= 3 by rejection, from Philox4x32-10 with counter (i, c, attempt) and
 * key (seed, 0x48505331), Box-Muller on two 53-bit uniforms; particle 0 is
 * the reference particle (on axis, at the centre); no losses.  The sampler
 * of hpsim_amd/configs.py, computed in place of a host upload (the
 * distribution set_distribution would receive from it, up to the last bit
 * of log and cos).  No reference counterpart (the reference samples on the
 * host, beam.cpp:391-537). */
int hps_beam_sample_gaussian(hps_beam* b, uint32_t seed, const double* center, const double* sigma);
/* Copy coordinates to host; any pointer may be NULL.  Replaces the array
 * part of get_x/get_xp/.../get_losses, wrap_beam.cpp:440-912. */
int hps_beam_download(hps_beam* b, double* x, double* xp, double* y, double* yp,
    double* phi, double* w, uint32_t* loss, uint32_t* lloss);
/* Recompute the relative phase (reference particle, or average of good
 * particles when use_good != 0) and copy it out; beam.cpp:1063-1066. */
int hps_beam_relative_phi(hps_beam* b, int use_good, double* out);
uint32_t hps_beam_size(const hps_beam* b);
/* set_ref_w / set_ref_phi / set_frequency, wrap_beam.cpp:271-341 */
int hps_beam_set_ref_w(hps_beam* b, double w);
int hps_beam_set_ref_phi(hps_beam* b, double phi);
int hps_beam_set_frequency(hps_beam* b, double freq);
/* get_mass/get_charge/get_frequency/get_ref_w/get_ref_phi/get_current and the
 * design energy (beam.h:140-151); wrap_beam.cpp:343-422,1286-1298.  which:
 * 0 mass, 1 charge, 2 frequency, 3 ref_w (w[0]), 4 ref_phi (phi[0]),
 * 5 current, 6 design_w, 7 current scaled by transmission. */
int hps_beam_get_scalar(hps_beam* b, int which, double* out);
/* get_loss_num('t'|'l'|'all'), wrap_beam.cpp:1245-1284 (which 0|1|2). */
int hps_beam_get_loss_num(hps_beam* b, int which, uint32_t* out);
/* get_avg_{x,y,phi,w}, wrap_beam.cpp:914-1049.  var: 0 x, 1 y, 2 phi,
 * 3 w, 4 relative phi; good: 0 "transverse_good"/"absolute", 1 "longitudinal_good". */
int hps_beam_get_avg(hps_beam* b, int var, int good, double* out);
/* get_sig_{x,y,phi,w}, wrap_beam.cpp:1051-1186, same var/good codes. */
int hps_beam_get_sig(hps_beam* b, int var, int good, double* out);
/* get_emittance_{x,y,z}, wrap_beam.cpp:1188-1243 (axis 0|1|2). */
int hps_beam_get_emittance(hps_beam* b, int axis, double* out);
/* Full scalar state (for diagnostics and tests). */
int hps_beam_get_stats(hps_beam* b, hps_beam_stats* out);
/* save/restore_{initial,intermediate}_beam, wrap_beam.cpp:163-233 (slot 0|1). */
int hps_beam_save(hps_beam* b, int slot);
int hps_beam_restore(hps_beam* b, int slot);
/* translate(axis, val), wrap_beam.cpp:1317-1351; axis 0 x,1 xp,2 y,3 yp,4 phi,5 w. */
int hps_beam_translate(hps_beam* b, int axis, double val);
/* apply_cut(coord, min, max), wrap_beam.cpp:1353-1371; coord 'x','y','w','p'. */
int hps_beam_apply_cut(hps_beam* b, char coord, double lo, double hi);

/* ---- BeamLine (src/beamline/inc/beamline.h:11-34) ---- */
int hps_beamline_create(hps_beamline** out);
int hps_beamline_destroy(hps_beamline* bl);
/* BeamLine::AddElement, beamline.h:53-57 (the reference fills it from the
 * database, src/beamline/src/init.cpp:20-429). */
int hps_beamline_add(hps_beamline* bl, const char* name, const hps_element* e);
/* Replace element i (model-parameter update, the PV path of the reference). */
int hps_beamline_set(hps_beamline* bl, uint32_t index, const hps_element* e);
uint32_t hps_beamline_size(const hps_beamline* bl);
/* BeamLine::GetElementModelIndex (case-insensitive), beamline.cpp:70-77. */
int hps_beamline_index(const hps_beamline* bl, const char* name, uint32_t* out);

/* ---- SpaceCharge (src/pywrapper/src/wrap_spacecharge.cpp) ---- */
/* SpaceCharge(nr, nz, interval, adj_bunch), wrap_spacecharge.cpp:31-56 */
int hps_spch_create(hps_spch** out, uint32_t nr, uint32_t nz, double interval, uint32_t adj_bunch);
int hps_spch_destroy(hps_spch* s);
/* set_{interval,remesh_threshold,adj_bunch,adj_bunch_cutoff_w,mesh_size_cutoff_w}
 * (what 0..4), space_charge.h:60-109. */
int hps_spch_set(hps_spch* s, int what, double value);
/* set_mesh_size(nr, nz), space_charge.cpp:61-73 */
int hps_spch_set_mesh_size(hps_spch* s, uint32_t nr, uint32_t nz);
int hps_spch_get_state(hps_spch* s, hps_spch_state* out);
/* Tables of the last remesh: Green's table in reference layout
 * [rp][rs][dz] ((nr+1)*nr*nz double2), field table [zp][rp]
 * ((nz+1)*(nr+1) double2), fixed-point charge histogram (nr*nz int64, units
 * of 2^-36 particle).  Any pointer may be NULL. */
int hps_spch_get_tables(hps_spch* s, double* fld1, double* fld2, int64_t* bins);

/* ---- Simulator (src/pywrapper/src/wrap_simulator.cpp) ---- */
/* Simulator(beam, beamline, spch=None), wrap_simulator.cpp:27-54 */
int hps_sim_create(hps_sim** out, hps_beam* b, hps_beamline* bl, hps_spch* s);
int hps_sim_destroy(hps_sim* m);
/* simulate() / simulate(start, end): inclusive model indices, -1 = first /
 * last element; SimulationEngine::Simulate, simulation_engine.cpp:73-95. */
int hps_sim_simulate(hps_sim* m, int32_t start, int32_t end);
/* reset(), wrap_simulator.cpp:67-72 */
int hps_sim_reset(hps_sim* m);
/* set_space_charge('on'|'off'), wrap_simulator.cpp:101-118 */
int hps_sim_set_space_charge(hps_sim* m, int on);
/* Per-kernel-family device timing with HIP events around every launch of
 * later simulate() calls (families: 0 fused element maps + kick + per-chunk
 * moments, 1 per-chunk relative-phase moments, 2 Green's table + charge
 * deposit, 3 field sum, 4 cross-chunk moments and longitudinal loss, 5
 * relative-phase moments + SCHEFF control, 6 the matrix-core product of a
 * batch's grouped field solve, also counted in family 3).  ms and launches
 * hold HPS_PROFILE_FAMILIES entries.  Enabling resets the counters.  No reference
 * counterpart: the reference only times whole simulate() calls
 * (timer.cpp:4-19). */
#define HPS_PROFILE_FAMILIES 7
int hps_sim_set_profiling(hps_sim* m, int on);
int hps_sim_get_profile(hps_sim* m, double* ms, uint32_t* launches);
/* Launch the kernel sequence of simulate() as a HIP graph (default on),
 * re-captured whenever any kernel argument or map differs from the cached
 * one; profiling and beam-frequency changes use eager launches.  No
 * reference counterpart (the reference launches every kernel from the host
 * loop, simulation_engine.cu:125-373). */
int hps_sim_set_graph(hps_sim* m, int on);
int hps_sim_graph_replays(hps_sim* m, uint32_t* out);
/* Device time of the last simulate() in ms (the reference prints it,
 * simulation_engine.cpp:84-92). */
int hps_sim_last_time_ms(hps_sim* m, double* out);

/* ---- Batch: several simulations in one launch sequence ---- */
/* The PV-scan loop of the reference, pytest/sim-lbeg-scan-pv.py:140-160
 * (restore_initial_beam + set_db_epics + simulate per setting, one after the
 * other), run as one program: every simulator of the batch (its own beam,
 * beamline settings and SpaceCharge) is compiled as by hps_sim_simulate,
 * and each kernel launch covers all of them.  Results are identical to
 * separate hps_sim_simulate calls.  The programs must have the same
 * structure (same element layout and kinds; element values and beam sizes
 * may differ, beams of at least 3 particles), the same SCHEFF mesh size,
 * no mesh_size_cutoff_w.  A simulator belongs to the batch by reference
 * (destroy the batch first). */
typedef struct hps_batch hps_batch;
int hps_batch_create(hps_batch** out);
int hps_batch_destroy(hps_batch* bt);
int hps_batch_add(hps_batch* bt, hps_sim* m);
uint32_t hps_batch_size(const hps_batch* bt);
/* simulate(start, end) of every simulator (same index rules as hps_sim_simulate) */
int hps_batch_simulate(hps_batch* bt, int32_t start, int32_t end);
int hps_batch_last_time_ms(hps_batch* bt, double* out);
/* Field-sum launch shape autotuning (default on): the first rebuild of a
 * batch (per mesh size and batch size) times every launch shape of the
 * field sum on the live tables and keeps the fastest; all shapes give the
 * same bits.  get_tuning: the chosen shape (UINT32_MAX: none yet) and up to
 * n candidate times in ms.  No reference counterpart. */
int hps_batch_set_autotune(hps_batch* bt, int on);
int hps_batch_get_tuning(hps_batch* bt, uint32_t* shape, float* ms, uint32_t n);
/* Lattice periods as one HIP graph (default on): when the launch sequence of
 * a simulate() holds n >= 2 repetitions of one lattice period's launches
 * (same kernels, records advancing by a fixed stride), that period is
 * captured once and replayed n times, each replay reading the next period's
 * records through a device offset; the graph is kept while the program's
 * structure and buffers are unchanged.  Launches outside the repeated run,
 * profiling runs, launch-shape tuning and programs with a beam frequency
 * change go eagerly.  No reference counterpart (the reference launches every
 * kernel from its element loop, simulation_engine.cu:125-373). */
int hps_batch_set_graph(hps_batch* bt, int on);
/* period graph replays so far and launches per replayed period (0: none) */
int hps_batch_graph_info(hps_batch* bt, uint64_t* replays, uint32_t* period_launches);
/* Field solve of the batch.  mode 0 (default): each simulation's own FP64
 * Green's table and field sum, results identical to hps_sim_simulate.  mode
 * 1: the grouped solve -- the Green's table normalised to dr = 1 depends on
 * the bunch-frame cell aspect kappa = dz*gamma/dr alone; it is kept for the
 * grid kappa_c = (1 + aspect_step)^c, each simulation's field is
 * interpolated quadratically in log(kappa) between its three nearest classes
 * (on its own, unchanged mesh), and the fields of all simulations that remesh at a
 * kick point are one grouped matrix product on the FP32 matrix cores (FP32
 * operands and accumulation), scaled per simulation by its c1 / dr^2.
 * Replaces UpdateFldTbl1Kernel + UpdateFldTbl2Kernel
 * (space_charge_kernel.cu:38-342) per beam; needs adj_bunch = 0 and nz a
 * multiple of 64.  Results are approximate: the interpolation error is
 * third order in the step (below 5e-6 of the field at 0.05) plus FP32
 * rounding.  The push and rebuild of such a batch divide by refined
 * reciprocals and take square roots from the reciprocal square root
 * (within one unit in the last place) and shorter sin/cos/sinh/cosh
 * polynomials (within 2.5e-14); the push visits only live particles and
 * adds each chunk's moments in its own order.  No reference counterpart. */
int hps_batch_set_field_mode(hps_batch* bt, int mode, double aspect_step);
/* Launch shape of the grouped product (index < 9, or UINT32_MAX: autotune
 * at the first rebuild, the default).  Shapes 0-7 run the direct product
 * (columns per tile, K splits, workgroups per CU); shape 8 runs it in the
 * frequency domain along z (radix-2 transforms of length 2 nz, the
 * per-frequency complex product on the matrix cores; needs an even nr and
 * a power-of-two nz, else simulate fails when it is asked for). */
int hps_batch_set_gemm_shape(hps_batch* bt, uint32_t shape);
/* info[0] chosen shape, [1] class tables built, [2] product columns that
 * used another class's table (cache conflicts), [3] product columns of the
 * last kick point (three per simulation), [4] product columns so far;
 * ms: the autotuner's time per shape. */
int hps_batch_get_gemm_info(hps_batch* bt, uint64_t* info, float* ms, uint32_t n);
/* per-kernel-family device time of the batched launches (families as
 * hps_sim_set_profiling) */
int hps_batch_set_profiling(hps_batch* bt, int on);
int hps_batch_get_profile(hps_batch* bt, double* ms, uint32_t* launches);

/* ---- SCHEFF field surrogate network ---- */
/* A one-dimensional convolutional U-Net along the bunch axis z with the
 * source rings as channels, for an nr x nz SCHEFF mesh (nr a multiple of
 * 16, nz of 64): input x[z][rs] = charge of bin (z, rs) * nr * nz / Q for
 * the nz bins (Q the total charge), conditioning the log cell aspect
 * log(dz*gamma/dr); output y[zp][2 rp + comp] for the (nz+1) x (nr+1) field
 * nodes = the SCHEFF field before the c1 / dr^2 factor, divided by Q.  Four
 * levels of widths widths[0..3] (z, z/2, z/4, z/8), nbot dilated layers at
 * the coarsest, FiLM scales from the conditioning on every layer, no biases
 * or activations (the field is linear in the charge).  BF16 weights, FP32
 * accumulation on the matrix cores, Adam training by the library's own
 * forward and backward kernels.  No reference counterpart: the reference
 * always solves SCHEFF (space_charge.cu:140-210). */
typedef struct hps_nn hps_nn;
int hps_nn_create(hps_nn** out, uint32_t nr, uint32_t nz, const uint32_t* widths, uint32_t nbot,
                  uint32_t max_batch, uint32_t seed);
int hps_nn_destroy(hps_nn* net);
uint32_t hps_nn_num_params(const hps_nn* net);
int hps_nn_get_params(hps_nn* net, float* host);
int hps_nn_set_params(hps_nn* net, const float* host);
/* layer rows {mode (0 same, 1 point, 2 down, 3 up), Cin, Cout, taps,
 * dilation, skip layer (-1 none), weight offset, FiLM offset} */
int hps_nn_layers(const hps_nn* net, int32_t* table, uint32_t max_layers, uint32_t* nlayers);
/* {Z0 (positions incl. padding), Cin, Cout, Cout padded, field rows, nr, nz} */
int hps_nn_dims(const hps_nn* net, uint32_t* out);
/* normalisation: the network predicts y / out_scale from the conditioning
 * (log aspect - u_mu) / u_sigma */
int hps_nn_set_norm(hps_nn* net, float out_scale, float u_mu, float u_sigma);
int hps_nn_get_norm(const hps_nn* net, float* out);
/* device dataset of (x [nz][nr], y [nz+1][2(nr+1)], log aspect) samples */
int hps_nn_data_reserve(hps_nn* net, uint32_t capacity);
int hps_nn_data_append(hps_nn* net, const float* x, const float* y, const float* log_aspect,
                       uint32_t count);
int hps_nn_data_size(hps_nn* net, uint32_t* count);
/* random order of the samples (so that any range is a random subset) */
int hps_nn_data_shuffle(hps_nn* net, uint32_t seed);
int hps_nn_data_get(hps_nn* net, uint32_t first, uint32_t count, float* x, float* y,
                    float* log_aspect);
/* Adam on random batches of dataset samples [first, first + count), the
 * learning rate decaying from lr to lr_end (cosine); losses[k] (optional)
 * the mean relative squared error of step k's batch */
int hps_nn_train(hps_nn* net, uint32_t first, uint32_t count, uint32_t steps, uint32_t batch,
                 float lr, float lr_end, uint32_t seed, float* losses);
/* per-sample relative squared error of dataset samples */
int hps_nn_eval(hps_nn* net, uint32_t first, uint32_t count, float* loss);
/* fields of host samples: y [count][nz+1][2(nr+1)] (times out_scale) */
int hps_nn_predict(hps_nn* net, const float* x, const float* log_aspect, uint32_t count, float* y);
/* per-sample losses and the gradient of their mean (for gradient checks) */
int hps_nn_loss_grad(hps_nn* net, const float* x, const float* log_aspect, const float* y,
                     uint32_t count, float* loss, float* grad);

/* Surrogate in a batch (grouped field solve only).  set_capture: every kick
 * point appends the (charge, exact field) pairs of its exact products to
 * the network's device dataset (while its capacity lasts): training data
 * from the engine's own SCHEFF solve.  set_surrogate: the network serves the
 * field of every remeshing simulation; at every audit_every-th kick point
 * (or the next one with remeshes) the exact product runs too and a
 * simulation whose relative L2 field error exceeds max_err falls back to the
 * exact product until a later audit passes; audited kick points use the
 * exact field; audit_every 0: no audits (every field from the network, for
 * ablations).  net NULL switches either off.  The network must be made for
 * the batch's mesh size.  No reference counterpart. */
int hps_batch_set_capture(hps_batch* bt, hps_nn* net);
int hps_batch_set_surrogate(hps_batch* bt, hps_nn* net, uint32_t audit_every, double max_err);
/* info: kick points, member audits, failed member audits, largest audited
 * error, member remeshes served by the network, by the exact product */
int hps_batch_get_surrogate_info(hps_batch* bt, double* info);

/* ---- Emittances without allocation ---- */
/* The three emittances of hps_beam_get_emittance (axis 0, 1, 2 -> out[0..2],
 * the same sums, order and expressions, so the same bits) over particle
 * arrays the caller holds in device memory on the current device: x, xp, y,
 * yp, phi, w (n doubles each) and loss, lloss (n words each).  No
 * counterpart in the reference: this entry point is for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc);
 * it calls no allocator.  Its reduction partials live in a static device
 * array of 311,616 bytes that the library holds on every device it runs on
 * until the process ends, so n is at most 4,194,304 (larger n is an error).
 * Work runs on the device's null stream and the call returns when out is
 * written.  Calls that overlap on one device race on that array. */
int hps_emittance_noalloc(const double* x, const double* xp, const double* y, const double* yp,
                          const double* phi, const double* w, const uint32_t* loss,
                          const uint32_t* lloss, uint32_t n, double mass, double* out);

/* ---- elementary-function probes (host-callable, run on the device) ---- */
/* Evaluates fn over n inputs on the GPU: 0 exp, 1 log, 2 sin, 3 cos, 4 sinh,
 * 5 cosh, 6 sqrt, 7 fmod(x, 2pi), 8 x/y, 9 Bessel I0, 10 Bessel I1,
 * 11 Bessel I1', and the per-particle maps' 12 sqrt, 13 sin, 14 cos, 15 sinh,
 * 16 cosh.  32 + k runs routine k as the reciprocal-division build of the
 * grouped-solve push computes it (division and square root within one unit
 * in the last place of the correctly rounded ones, the maps' sin and cos
 * within 2.5e-14, sinh and cosh within 4e-15 relative). */
int hps_probe_math(int fn, const double* x, const double* y, double* out, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
