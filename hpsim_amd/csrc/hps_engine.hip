// hps_engine.hip -- host side of the tracking engine and its C ABI.
//
// Beam, BeamLine, SpaceCharge and Simulator mirror the reference C++ classes
// (src/beam, src/beamline, src/simulation) behind the C ABI of hpsim_amd.h.
// simulate() compiles the element range into fused segments (all maps between
// two space-charge kicks) and launches them back to back on one stream; the
// Scheff::Start decisions run on the device, so the host never waits inside a
// simulate() call except for the rare beam frequency change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>
#include <vector>

#include "hpsim_amd.h"
#include "hps_internal.hpp"
#include "hps_unet.hpp"

using namespace hps;

namespace {

thread_local std::string g_error;

int fail(const std::string& m)
{
  g_error = m;
  return -1;
}

#define HPS_CHECK(expr)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(std::string(#expr) + ": " + hipGetErrorString(e_));                     \
  } while (0)

uint32_t next_pow2(uint32_t x)   // utility.cpp:7-16
{
  --x;
  x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16;
  return ++x;
}

hipStream_t g_stream[16];

}  // namespace

namespace hps {
// the error text of hps_last_error(), for the other translation units
int fail_msg(const std::string& m) { return fail(m); }
}  // namespace hps

namespace {

int stream_for(int dev, hipStream_t* out)
{
  if (dev < 0 || dev >= 16) return fail("device index out of range");
  if (!g_stream[dev]) {
    HPS_CHECK(hipSetDevice(dev));
    HPS_CHECK(hipStreamCreate(&g_stream[dev]));
  }
  *out = g_stream[dev];
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Beam
// ---------------------------------------------------------------------------
struct hps_beam {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n = 0;
  double mass = 0, charge = 0, current = 0, freq = 0, design_w = 0;
  uint32_t grid_size = 1, blck_size = 1;
  BeamArrays d{};
  hps_beam_stats* st = nullptr;     // device
  Partials pr{};
  double* snap[2][6] = {{nullptr}};
  uint32_t* snap_l[2][2] = {{nullptr}};
  bool saved[2] = {false, false};
  double snap_dw[2] = {0, 0}, snap_f[2] = {0, 0}, snap_c[2] = {0, 0};
};

namespace {

// Canonical-order sums over the beam (k_sums), copied to host.
struct Sums {
  double s[5];
  uint32_t nloss, ngood, nlloss;
};

int run_sums(hps_beam* b, const double* a, const double* bb, int filter, Sums* out)
{
  HPS_CHECK(launch_sums(b->d, a, bb, filter, b->pr, b->stream));
  HPS_CHECK(hipMemcpyAsync(out->s, b->pr.out, 5 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  uint32_t c[3];
  HPS_CHECK(hipMemcpyAsync(c, b->pr.out_cnt, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
  HPS_CHECK(hipStreamSynchronize(b->stream));
  out->nloss = c[0];
  out->ngood = c[1];
  out->nlloss = c[2];
  return 0;
}

int pull(hps_beam* b, hps_beam_stats* s)
{
  HPS_CHECK(hipMemcpyAsync(s, b->st, sizeof(*s), hipMemcpyDeviceToHost, b->stream));
  HPS_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

int push(hps_beam* b, const hps_beam_stats* s)
{
  HPS_CHECK(hipMemcpyAsync(b->st, s, sizeof(*s), hipMemcpyHostToDevice, b->stream));
  HPS_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

double surv(const hps_beam* b, const hps_beam_stats& s) { return (double)(uint32_t)(b->n - s.num_loss); }

// UpdateLossCountKernelCall, beam.cu:493-516
int update_loss(hps_beam* b, hps_beam_stats& s)
{
  Sums r;
  if (run_sums(b, nullptr, nullptr, 2, &r)) return -1;
  s.num_loss = r.nloss;
  return 0;
}

// UpdateGoodParticleCountKernelCall, beam.cu:1233-1255
int update_good(hps_beam* b, hps_beam_stats& s)
{
  Sums r;
  if (run_sums(b, nullptr, nullptr, 2, &r)) return -1;
  s.num_good = r.ngood;
  return 0;
}

// UpdateAvgOfOneVariableKernelCall + finalise kernels, beam.cu:411-436,
// beam_kernel.cu:1396-1410
int avg_of(hps_beam* b, hps_beam_stats& s, const double* a, bool good, double* out)
{
  Sums r;
  if (run_sums(b, a, nullptr, good ? 1 : 0, &r)) return -1;
  if (!good) *out = r.s[0] / surv(b, s);
  else *out = s.num_good > 0 ? r.s[0] / (double)s.num_good : -99999;
  return 0;
}

// UpdateSigmaOfOneVariableKernelCall, beam.cu:651-677
int sig_of(hps_beam* b, hps_beam_stats& s, const double* a, bool good, double* out)
{
  Sums r;
  if (run_sums(b, a, nullptr, good ? 1 : 0, &r)) return -1;
  const double d = good ? (double)s.num_good : surv(b, s);
  const double avg = r.s[0] / d, avg2 = r.s[1] / d;
  *out = std::sqrt(avg2 - avg * avg);
  return 0;
}

// UpdateRelativePhiKernelCall, beam.cu:1122-1131 (writes phi_r)
int relative_phi(hps_beam* b, bool use_good)
{
  const double* ref = use_good ? &b->st->phi_avg_good : b->d.phi;
  HPS_CHECK(launch_relative_phi(b->d, ref, b->stream));
  return 0;
}

// Beam::UpdateLongitudinalLoss, beam.cpp:776-782
int update_longitudinal_loss(hps_beam* b, hps_beam_stats& s)
{
  if (update_good(b, s)) return -1;
  if (avg_of(b, s, b->d.phi, true, &s.phi_avg_good)) return -1;
  if (push(b, &s)) return -1;
  HPS_CHECK(launch_lloss_flag(b->d, &b->st->phi_avg_good, b->stream));
  Sums r;
  if (run_sums(b, nullptr, nullptr, 2, &r)) return -1;
  s.num_lloss = r.nlloss;
  return 0;
}

// Beam::InitPhiAvgGood, beam.cpp:655-661
int init_phi_avg_good(hps_beam* b)
{
  hps_beam_stats s;
  if (pull(b, &s) || update_loss(b, s) || avg_of(b, s, b->d.phi, false, &s.phi_avg)) return -1;
  s.phi_avg_good = s.phi_avg;
  return push(b, &s);
}

// Beam::ChangeFrequency, beam.cpp:1654-1668
int change_frequency(hps_beam* b, double f)
{
  const double ratio = f / b->freq;
  b->current *= ratio;
  HPS_CHECK(launch_scale(b->d.phi, b->n, ratio, b->stream));
  b->freq = f;
  hps_beam_stats s;
  if (pull(b, &s) || avg_of(b, s, b->d.phi, false, &s.phi_avg) || push(b, &s)) return -1;
  if (relative_phi(b, false)) return -1;
  if (avg_of(b, s, b->d.phi_r, false, &s.phi_avg_r)) return -1;
  return push(b, &s);
}

int emittance(hps_beam* b, hps_beam_stats& s)   // Beam::UpdateEmittance, beam.cpp:985-994
{
  if (b->n <= 2) return 0;
  double w0;
  HPS_CHECK(hipMemcpyAsync(&w0, b->d.w, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HPS_CHECK(hipStreamSynchronize(b->stream));
  for (int axis = 0; axis < 2; ++axis) {   // beam.cu:1007-1064, beam_kernel.cu:1333-1354
    Sums r;
    if (run_sums(b, axis == 0 ? b->d.x : b->d.y, axis == 0 ? b->d.xp : b->d.yp, 0, &r)) return -1;
    const double sv = surv(b, s);
    const double x = r.s[0] / sv, x2 = r.s[1] / sv, xp = r.s[2] / sv, xp2 = r.s[3] / sv,
                 xxp = r.s[4] / sv;
    const double gamma1 = w0 / b->mass;
    const double betagamma = std::sqrt(gamma1 * (gamma1 + 2.0));
    const double sig = std::sqrt(x2 - x * x), sigp = std::sqrt(xp2 - xp * xp);
    const double emit = std::sqrt((x2 - x * x) * (xp2 - xp * xp) - (xxp - x * xp) * (xxp - x * xp)) *
                        betagamma;
    if (axis == 0) { s.x_avg = x; s.x_sig = sig; s.xp_avg = xp; s.xp_sig = sigp; s.x_emit = emit; }
    else { s.y_avg = x; s.y_sig = sig; s.yp_avg = xp; s.yp_sig = sigp; s.y_emit = emit; }
  }
  if (relative_phi(b, false)) return -1;
  Sums r;   // beam.cu:1069-1092, beam_kernel.cu:1356-1373
  if (run_sums(b, b->d.phi_r, b->d.w, 0, &r)) return -1;
  const double sv = surv(b, s);
  const double phi = r.s[0] / sv, phi2 = r.s[1] / sv, w = r.s[2] / sv, w2 = r.s[3] / sv,
               phi_w = r.s[4] / sv;
  const double phi_var = phi2 - phi * phi;
  const double w_var = w2 - w * w;
  const double cov = phi_w - phi * w;
  s.phi_avg_r = phi;
  s.w_avg = w;
  s.phi_sig_r = std::sqrt(phi_var);
  s.w_sig = std::sqrt(w_var);
  s.z_emit = std::sqrt(phi_var * w_var - cov * cov);
  return 0;
}

void free_beam(hps_beam* b)
{
  double* arrs[7] = {b->d.x, b->d.xp, b->d.y, b->d.yp, b->d.phi, b->d.phi_r, b->d.w};
  for (double* p : arrs) (void)hipFree(p);
  (void)hipFree(b->d.loss);
  (void)hipFree(b->d.lloss);
  (void)hipFree(b->st);
  (void)hipFree(b->pr.sums);
  (void)hipFree(b->pr.cnt);
  (void)hipFree(b->pr.pmax);
  (void)hipFree(b->pr.rmax2);
  (void)hipFree(b->pr.out);
  (void)hipFree(b->pr.out_cnt);
  for (int k = 0; k < 2; ++k) {
    for (int j = 0; j < 6; ++j) (void)hipFree(b->snap[k][j]);
    (void)hipFree(b->snap_l[k][0]);
    (void)hipFree(b->snap_l[k][1]);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// BeamLine / SpaceCharge / Simulator
// ---------------------------------------------------------------------------
struct hps_beamline {
  std::vector<std::string> names;
  std::vector<hps_element> el;
};

struct hps_spch {
  int device = 0;
  hipStream_t stream = nullptr;
  hps_spch_state cfg{};   // host copy of the configuration
  ScDev* dev = nullptr;
  double* fld1 = nullptr;
  double* fld2 = nullptr;
  int64_t* bins = nullptr;
  double* part = nullptr;    // split partial sums of the field (kFieldMaxSplit tables)
  size_t max_lds = 0;        // LDS bytes per workgroup on this device
};

struct hps_sim {
  hps_beam* b = nullptr;
  hps_beamline* bl = nullptr;
  hps_spch* sc = nullptr;
  bool sc_on = false;
  uint32_t ccl_cell_num = 0;
  double design_w = 0, wavelen = 0;
  int32_t prev_end = -1;      // uint max stored in int, simulation_engine.cpp:52
  Op* d_ops = nullptr;
  size_t ops_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  // per-kernel-family device time (HIP events around each launch)
  bool profile = false;
  std::vector<hipEvent_t> evpool;
  double prof_ms[HPS_PROFILE_FAMILIES] = {};
  uint32_t prof_n[HPS_PROFILE_FAMILIES] = {};
  // HIP graph of the last launch sequence, reused while the sequence
  // (every kernel argument and the map table) is byte-identical
  bool use_graph = true;
  std::vector<uint8_t> graph_sig;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint32_t graph_replays = 0;
};

extern "C" {

const char* hps_last_error(void) { return g_error.c_str(); }

int hps_set_device(int device)
{
  HPS_CHECK(hipSetDevice(device));
  return 0;
}

int hps_beam_create(hps_beam** out, uint32_t num, double mass, double charge, double current)
{
  if (!out) return fail("null output handle");
  if (num == 0) return fail("Beam needs num > 0 particles");
  // the cross-chunk reductions fold at most 4096 chunks of 1024 particles
  if (num > 4096u * kChunk) return fail("a beam holds at most 4,194,304 particles");
  hps_beam* b = new hps_beam();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { delete b; return fail("no HIP device"); }
  b->device = dev;
  if (stream_for(dev, &b->stream)) { delete b; return -1; }
  b->n = num; b->mass = mass; b->charge = charge; b->current = current;
  // Beam::SetNumThreadsPerBlock(256), beam.cpp:148-168
  const uint32_t nt = next_pow2(num);
  if (nt > 256) { b->grid_size = nt / 256; b->blck_size = 256; }
  else { b->grid_size = 1; b->blck_size = nt; }
  const size_t nd = (size_t)num * sizeof(double), nu = (size_t)num * sizeof(uint32_t);
  double** arrs[7] = {&b->d.x, &b->d.xp, &b->d.y, &b->d.yp, &b->d.phi, &b->d.phi_r, &b->d.w};
  bool ok = true;
  for (double** p : arrs) ok = ok && hipMalloc(p, nd) == hipSuccess && hipMemset(*p, 0, nd) == hipSuccess;
  ok = ok && hipMalloc(&b->d.loss, nu) == hipSuccess && hipMemset(b->d.loss, 0, nu) == hipSuccess;
  ok = ok && hipMalloc(&b->d.lloss, nu) == hipSuccess && hipMemset(b->d.lloss, 0, nu) == hipSuccess;
  ok = ok && hipMalloc(&b->st, sizeof(hps_beam_stats)) == hipSuccess &&
       hipMemset(b->st, 0, sizeof(hps_beam_stats)) == hipSuccess;
  b->pr.nch = (num + kChunk - 1) / kChunk;
  b->pr.mpad = next_pow2(b->pr.nch);
  const size_t nsum = (size_t)kMaxQ * b->pr.mpad + b->pr.mpad / 2 + 256;
  ok = ok && hipMalloc(&b->pr.sums, nsum * sizeof(double)) == hipSuccess &&
       hipMemset(b->pr.sums, 0, nsum * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc(&b->pr.cnt, 3 * (size_t)b->pr.mpad * 4) == hipSuccess &&
       hipMemset(b->pr.cnt, 0, 3 * (size_t)b->pr.mpad * 4) == hipSuccess;
  ok = ok && hipMalloc(&b->pr.pmax, (size_t)b->pr.mpad * 8) == hipSuccess &&
       hipMemset(b->pr.pmax, 0, (size_t)b->pr.mpad * 8) == hipSuccess;
  ok = ok && hipMalloc(&b->pr.rmax2, (size_t)b->pr.mpad * 8) == hipSuccess &&
       hipMemset(b->pr.rmax2, 0, (size_t)b->pr.mpad * 8) == hipSuccess;
  ok = ok && hipMalloc(&b->pr.out, kMaxQ * sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc(&b->pr.out_cnt, 4 * sizeof(uint32_t)) == hipSuccess;
  if (!ok) { free_beam(b); delete b; return fail("device allocation failed for the beam"); }
  b->d.n = num; b->d.mass = mass; b->d.charge = charge;
  *out = b;
  return 0;
}

int hps_beam_destroy(hps_beam* b)
{
  if (!b) return 0;
  (void)hipStreamSynchronize(b->stream);
  free_beam(b);
  delete b;
  return 0;
}

uint32_t hps_beam_size(const hps_beam* b) { return b ? b->n : 0; }

int hps_beam_set_distribution(hps_beam* b, const double* x, const double* xp, const double* y,
    const double* yp, const double* phi, const double* w, const uint32_t* loss, const uint32_t* lloss)
{
  if (!b || !x || !xp || !y || !yp || !phi || !w) return fail("set_distribution needs six coordinate arrays");
  const size_t nd = (size_t)b->n * sizeof(double), nu = (size_t)b->n * 4;
  HPS_CHECK(hipMemcpy(b->d.x, x, nd, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.xp, xp, nd, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.y, y, nd, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.yp, yp, nd, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.phi, phi, nd, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.w, w, nd, hipMemcpyHostToDevice));
  // the reference particle sits on axis, beam.cpp:344-345
  const double zero = 0.0;
  HPS_CHECK(hipMemcpy(b->d.x, &zero, 8, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.xp, &zero, 8, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.y, &zero, 8, hipMemcpyHostToDevice));
  HPS_CHECK(hipMemcpy(b->d.yp, &zero, 8, hipMemcpyHostToDevice));
  const uint32_t uz = 0;
  if (loss) {
    HPS_CHECK(hipMemcpy(b->d.loss, loss, nu, hipMemcpyHostToDevice));
    HPS_CHECK(hipMemcpy(b->d.loss, &uz, 4, hipMemcpyHostToDevice));
  } else {
    HPS_CHECK(hipMemset(b->d.loss, 0, nu));
  }
  if (lloss) {
    HPS_CHECK(hipMemcpy(b->d.lloss, lloss, nu, hipMemcpyHostToDevice));
    HPS_CHECK(hipMemcpy(b->d.lloss, &uz, 4, hipMemcpyHostToDevice));
  } else {
    HPS_CHECK(hipMemset(b->d.lloss, 0, nu));
  }
  HPS_CHECK(hipDeviceSynchronize());
  return init_phi_avg_good(b);
}

// Initial-beam samplers of the reference Beam (beam.cpp:391-537).  They run
// on the host like the reference's and draw from the C library's
// srand/rand, so one seed gives the reference's particles on the same libc.
namespace {

double rand_pm1() { return 2.0 * std::rand() / RAND_MAX - 1.0; }   // beam.cpp:424-425

// PARMILA ellipse half-widths from Twiss (a, b, e): g = (1 + a^2)/b,
// umax = sqrt(e/g), upmax = sqrt(e*g), shear du = -a/g (beam.cpp:409-418).
struct Ellipse {
  double umax, upmax, du;
  Ellipse(double a, double bt, double e)
  {
    const double g = (1.0 + a * a) / bt;
    umax = std::sqrt(e / g);
    upmax = std::sqrt(e * g);
    du = -a / g;
  }
};

int upload_sampled(hps_beam* b, std::vector<double>& v)
{
  const size_t n = b->n;
  return hps_beam_set_distribution(b, v.data(), v.data() + n, v.data() + 2 * n,
                                    v.data() + 3 * n, v.data() + 4 * n, v.data() + 5 * n,
                                    nullptr, nullptr);
}

}  // namespace

int hps_beam_set_waterbag(hps_beam* b, double ax, double bx, double ex, double ay, double by,
    double ey, double az, double bz, double ez, double sync_phi, double sync_w, double freq,
    uint32_t seed)
{
  if (!b) return fail("null beam");
  if (!(bx > 0 && by > 0 && bz > 0 && freq > 0 && sync_w > 0))
    return fail("set_waterbag needs positive beta functions, frequency and energy");
  std::srand(seed == 0 ? (unsigned)std::time(nullptr) : seed);
  const Ellipse ex_(ax, bx, ex), ey_(ay, by, ey), ez_(az, bz, ez);
  const double zpmx = ez / ez_.umax;               // beam.cpp:418
  const double wave_len = 29979.2458 / freq;       // cm
  const double gm = sync_w / b->mass + 1.0;
  const double beta = std::sqrt(1.0 - 1.0 / (gm * gm));
  const size_t n = b->n;
  std::vector<double> v(6 * n);
  for (size_t i = 0; i < n; ++i) {
    double r1, r2, r3, r4, r5, r6;
    // a point of the unit 6-ball: three discs by rejection, then the joint radius
    do {
      do { r1 = rand_pm1(); r2 = rand_pm1(); } while (r1 * r1 + r2 * r2 > 1.0);
      do { r3 = rand_pm1(); r4 = rand_pm1(); } while (r3 * r3 + r4 * r4 > 1.0);
      do { r5 = rand_pm1(); r6 = rand_pm1(); } while (r5 * r5 + r6 * r6 > 1.0);
    } while (r1 * r1 + r2 * r2 + r3 * r3 + r4 * r4 + r5 * r5 + r6 * r6 > 1.0);
    v[i] = 0.01 * (r1 * ex_.umax + ex_.du * r2 * ex_.upmax);
    v[n + i] = r2 * ex_.upmax;
    v[2 * n + i] = 0.01 * (r3 * ey_.umax + ey_.du * r4 * ey_.upmax);
    v[3 * n + i] = r4 * ey_.upmax;
    v[4 * n + i] = -(r5 * ez_.umax + ez_.du * zpmx * r6) * 2.0 * hps::PI / (beta * wave_len) + sync_phi;
    v[5 * n + i] = 2.0 * sync_w * zpmx * r6 + sync_w;
  }
  b->freq = freq;                                  // beam.cpp:421
  return upload_sampled(b, v);
}

int hps_beam_set_dc(hps_beam* b, double ax, double bx, double ex, double ay, double by,
    double ey, double dphi, double sync_phi, double sync_w, uint32_t seed)
{
  if (!b) return fail("null beam");
  if (!(bx > 0 && by > 0)) return fail("set_dc needs positive beta functions");
  std::srand(seed == 0 ? (unsigned)std::time(nullptr) : seed);
  const Ellipse ex_(ax, bx, ex), ey_(ay, by, ey);
  const size_t n = b->n;
  std::vector<double> v(6 * n);
  for (size_t i = 0; i < n; ++i) {
    double r1, r2, r3, r4;
    // a point of the unit 4-ball
    do {
      do { r1 = rand_pm1(); r2 = rand_pm1(); } while (r1 * r1 + r2 * r2 > 1.0);
      do { r3 = rand_pm1(); r4 = rand_pm1(); } while (r3 * r3 + r4 * r4 > 1.0);
    } while (r1 * r1 + r2 * r2 + r3 * r3 + r4 * r4 > 1.0);
    const double r5 = 2.0 * ((double)i + 1.0) / (double)n - 1.0;   // phase ramp by index
    v[i] = 0.01 * (r1 * ex_.umax + ex_.du * r2 * ex_.upmax);
    v[n + i] = r2 * ex_.upmax;
    v[2 * n + i] = 0.01 * (r3 * ey_.umax + ey_.du * r4 * ey_.upmax);
    v[3 * n + i] = r4 * ey_.upmax;
    v[4 * n + i] = r5 * dphi + sync_phi;
    v[5 * n + i] = sync_w;
  }
  return upload_sampled(b, v);
}

int hps_beam_sample_gaussian(hps_beam* b, uint32_t seed, const double* center, const double* sigma)
{
  if (!b || !center || !sigma) return fail("null argument");
  GaussSpec g{};
  for (int c = 0; c < 6; ++c) { g.center[c] = center[c]; g.sigma[c] = sigma[c]; }
  HPS_CHECK(launch_sample_gaussian(b->d, seed, g, b->stream));
  HPS_CHECK(hipStreamSynchronize(b->stream));
  return init_phi_avg_good(b);
}

int hps_beam_download(hps_beam* b, double* x, double* xp, double* y, double* yp, double* phi,
                      double* w, uint32_t* loss, uint32_t* lloss)
{
  if (!b) return fail("null beam");
  HPS_CHECK(hipStreamSynchronize(b->stream));
  const size_t nd = (size_t)b->n * sizeof(double), nu = (size_t)b->n * 4;
  if (x) HPS_CHECK(hipMemcpy(x, b->d.x, nd, hipMemcpyDeviceToHost));
  if (xp) HPS_CHECK(hipMemcpy(xp, b->d.xp, nd, hipMemcpyDeviceToHost));
  if (y) HPS_CHECK(hipMemcpy(y, b->d.y, nd, hipMemcpyDeviceToHost));
  if (yp) HPS_CHECK(hipMemcpy(yp, b->d.yp, nd, hipMemcpyDeviceToHost));
  if (phi) HPS_CHECK(hipMemcpy(phi, b->d.phi, nd, hipMemcpyDeviceToHost));
  if (w) HPS_CHECK(hipMemcpy(w, b->d.w, nd, hipMemcpyDeviceToHost));
  if (loss) HPS_CHECK(hipMemcpy(loss, b->d.loss, nu, hipMemcpyDeviceToHost));
  if (lloss) HPS_CHECK(hipMemcpy(lloss, b->d.lloss, nu, hipMemcpyDeviceToHost));
  return 0;
}

int hps_beam_relative_phi(hps_beam* b, int use_good, double* out)
{
  if (!b) return fail("null beam");
  if (relative_phi(b, use_good != 0)) return -1;
  HPS_CHECK(hipStreamSynchronize(b->stream));
  if (out) HPS_CHECK(hipMemcpy(out, b->d.phi_r, (size_t)b->n * 8, hipMemcpyDeviceToHost));
  return 0;
}

int hps_beam_set_ref_w(hps_beam* b, double w)   // beam.cpp:1074-1078
{
  if (!b) return fail("null beam");
  HPS_CHECK(hipMemcpy(b->d.w, &w, 8, hipMemcpyHostToDevice));
  b->design_w = w;
  return 0;
}

int hps_beam_set_ref_phi(hps_beam* b, double phi)   // beam.cpp:1086-1089
{
  if (!b) return fail("null beam");
  HPS_CHECK(hipMemcpy(b->d.phi, &phi, 8, hipMemcpyHostToDevice));
  return 0;
}

int hps_beam_set_frequency(hps_beam* b, double f)   // wrap_beam.cpp:315-341
{
  if (!b) return fail("null beam");
  if (b->freq == 0) { b->freq = f; return 0; }
  return change_frequency(b, f);
}

int hps_beam_get_scalar(hps_beam* b, int which, double* out)
{
  if (!b || !out) return fail("null argument");
  double v = 0;
  switch (which) {
    case 0: v = b->mass; break;
    case 1: v = b->charge; break;
    case 2: v = b->freq; break;
    case 3: HPS_CHECK(hipStreamSynchronize(b->stream)); HPS_CHECK(hipMemcpy(&v, b->d.w, 8, hipMemcpyDeviceToHost)); break;
    case 4: HPS_CHECK(hipStreamSynchronize(b->stream)); HPS_CHECK(hipMemcpy(&v, b->d.phi, 8, hipMemcpyDeviceToHost)); break;
    case 5: v = b->current; break;
    case 6: v = b->design_w; break;
    case 7: {   // BeamGetCurrent, wrap_beam.cpp:1286-1298
      hps_beam_stats s;
      if (pull(b, &s) || update_loss(b, s) || push(b, &s)) return -1;
      v = b->current * (1.0 - ((double)s.num_loss) / ((double)b->n));
      break;
    }
    default: return fail("unknown scalar");
  }
  *out = v;
  return 0;
}

int hps_beam_get_loss_num(hps_beam* b, int which, uint32_t* out)   // wrap_beam.cpp:1245-1284
{
  if (!b || !out) return fail("null argument");
  hps_beam_stats s;
  if (pull(b, &s)) return -1;
  if (which == 0) {
    if (update_loss(b, s)) return -1;
    *out = s.num_loss;
  } else if (which == 1) {
    if (update_longitudinal_loss(b, s)) return -1;
    *out = s.num_lloss;
  } else {
    if (update_loss(b, s) || update_longitudinal_loss(b, s) || update_good(b, s)) return -1;
    *out = b->n - s.num_good;
  }
  return push(b, &s);
}

int hps_beam_get_avg(hps_beam* b, int var, int good, double* out)   // wrap_beam.cpp:914-1049
{
  if (!b || !out) return fail("null argument");
  hps_beam_stats s;
  if (pull(b, &s) || update_loss(b, s)) return -1;
  double v = 0;
  int rc = 0;
  switch (var) {
    case 0: rc = avg_of(b, s, b->d.x, good != 0, good ? &s.x_avg_good : &s.x_avg); v = good ? s.x_avg_good : s.x_avg; break;
    case 1: rc = avg_of(b, s, b->d.y, good != 0, good ? &s.y_avg_good : &s.y_avg); v = good ? s.y_avg_good : s.y_avg; break;
    case 2:
      if (!good) { rc = avg_of(b, s, b->d.phi, false, &s.phi_avg); v = s.phi_avg; }
      else { rc = update_good(b, s) || avg_of(b, s, b->d.phi, true, &s.phi_avg_good); v = s.phi_avg_good; }
      break;
    case 3: rc = avg_of(b, s, b->d.w, good != 0, good ? &s.w_avg_good : &s.w_avg); v = good ? s.w_avg_good : s.w_avg; break;
    default:
      rc = relative_phi(b, false) || avg_of(b, s, b->d.phi_r, good != 0, &s.phi_avg_r);
      v = s.phi_avg_r;
      break;
  }
  if (rc || push(b, &s)) return -1;
  *out = v;
  return 0;
}

int hps_beam_get_sig(hps_beam* b, int var, int good, double* out)   // wrap_beam.cpp:1051-1186
{
  if (!b || !out) return fail("null argument");
  hps_beam_stats s;
  if (pull(b, &s) || update_loss(b, s)) return -1;
  if (b->n < 2) { *out = 0.0; return push(b, &s); }
  const bool upd = b->n > 2;
  double v = 0;
  int rc = 0;
  switch (var) {
    case 0:
      if (upd) rc = sig_of(b, s, b->d.x, good != 0, good ? &s.x_sig_good : &s.x_sig);
      v = good ? s.x_sig_good : s.x_sig;
      break;
    case 1:
      if (upd) rc = sig_of(b, s, b->d.y, good != 0, good ? &s.y_sig_good : &s.y_sig);
      v = good ? s.y_sig_good : s.y_sig;
      break;
    case 2:
      if (!good) { if (upd) rc = sig_of(b, s, b->d.phi, false, &s.phi_sig); v = s.phi_sig; break; }
      rc = update_good(b, s) || avg_of(b, s, b->d.phi, true, &s.phi_avg_good) || push(b, &s) ||
           relative_phi(b, true);
      if (!rc && upd) rc = sig_of(b, s, b->d.phi_r, true, &s.phi_sig_good);
      v = s.phi_sig_good;
      break;
    case 3:
      if (upd) rc = sig_of(b, s, b->d.w, good != 0, good ? &s.w_sig_good : &s.w_sig);
      v = good ? s.w_sig_good : s.w_sig;
      break;
    default:
      rc = relative_phi(b, false);
      if (!rc && upd) rc = sig_of(b, s, b->d.phi_r, false, &s.phi_sig_r);
      v = s.phi_sig_r;
      break;
  }
  if (rc || push(b, &s)) return -1;
  *out = v;
  return 0;
}

int hps_beam_get_emittance(hps_beam* b, int axis, double* out)   // wrap_beam.cpp:1188-1243
{
  if (!b || !out) return fail("null argument");
  hps_beam_stats s;
  if (pull(b, &s) || update_loss(b, s) || emittance(b, s) || push(b, &s)) return -1;
  if (b->n < 2) { *out = 0.0; return 0; }
  *out = axis == 0 ? s.x_emit : (axis == 1 ? s.y_emit : s.z_emit);
  return 0;
}

int hps_beam_get_stats(hps_beam* b, hps_beam_stats* out)
{
  if (!b || !out) return fail("null argument");
  return pull(b, out);
}

int hps_beam_save(hps_beam* b, int slot)   // beam.cpp:547-595
{
  if (!b || slot < 0 || slot > 1) return fail("bad save slot");
  const size_t nd = (size_t)b->n * 8, nu = (size_t)b->n * 4;
  double* src[6] = {b->d.x, b->d.xp, b->d.y, b->d.yp, b->d.phi, b->d.w};
  for (int j = 0; j < 6; ++j) {
    if (!b->snap[slot][j]) HPS_CHECK(hipMalloc(&b->snap[slot][j], nd));
    HPS_CHECK(hipMemcpyAsync(b->snap[slot][j], src[j], nd, hipMemcpyDeviceToDevice, b->stream));
  }
  for (int j = 0; j < 2; ++j) {
    if (!b->snap_l[slot][j]) HPS_CHECK(hipMalloc(&b->snap_l[slot][j], nu));
    HPS_CHECK(hipMemcpyAsync(b->snap_l[slot][j], j ? b->d.lloss : b->d.loss, nu,
                             hipMemcpyDeviceToDevice, b->stream));
  }
  HPS_CHECK(hipStreamSynchronize(b->stream));
  b->saved[slot] = true;
  b->snap_dw[slot] = b->design_w; b->snap_f[slot] = b->freq; b->snap_c[slot] = b->current;
  return 0;
}

int hps_beam_restore(hps_beam* b, int slot)   // beam.cpp:606-647
{
  if (!b || slot < 0 || slot > 1) return fail("bad restore slot");
  if (!b->saved[slot]) return fail("Beam can't be restored because no previous beam has been saved.");
  const size_t nd = (size_t)b->n * 8, nu = (size_t)b->n * 4;
  double* dst[6] = {b->d.x, b->d.xp, b->d.y, b->d.yp, b->d.phi, b->d.w};
  for (int j = 0; j < 6; ++j)
    HPS_CHECK(hipMemcpyAsync(dst[j], b->snap[slot][j], nd, hipMemcpyDeviceToDevice, b->stream));
  HPS_CHECK(hipMemcpyAsync(b->d.loss, b->snap_l[slot][0], nu, hipMemcpyDeviceToDevice, b->stream));
  HPS_CHECK(hipMemcpyAsync(b->d.lloss, b->snap_l[slot][1], nu, hipMemcpyDeviceToDevice, b->stream));
  b->design_w = b->snap_dw[slot]; b->freq = b->snap_f[slot]; b->current = b->snap_c[slot];
  hps_beam_stats s;
  if (pull(b, &s) || avg_of(b, s, b->d.phi, true, &s.phi_avg_good) || push(b, &s)) return -1;
  if (relative_phi(b, true)) return -1;
  HPS_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

int hps_beam_translate(hps_beam* b, int axis, double val)   // beam.cpp:1612-1652
{
  if (!b || axis < 0 || axis > 5) return fail("bad translate axis");
  double* arr[6] = {b->d.x, b->d.xp, b->d.y, b->d.yp, b->d.phi, b->d.w};
  HPS_CHECK(launch_shift(arr[axis], b->n, val, b->stream));
  if (axis == 4) return init_phi_avg_good(b);
  HPS_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

int hps_beam_apply_cut(hps_beam* b, char coord, double lo, double hi)   // beam.cpp:1595-1610
{
  if (!b) return fail("null beam");
  const double* cd = nullptr;
  switch (coord) {
    case 'x': case 'X': cd = b->d.x; break;
    case 'y': case 'Y': cd = b->d.y; break;
    case 'w': case 'W': cd = b->d.w; break;
    case 'p': case 'P': cd = b->d.phi; break;
    default: return fail("apply_cut coordinate must be one of x, y, w, p");
  }
  HPS_CHECK(launch_cut(cd, b->d.loss, b->n, lo, hi, b->stream));
  HPS_CHECK(hipStreamSynchronize(b->stream));
  return 0;
}

// ---- BeamLine ----
int hps_beamline_create(hps_beamline** out)
{
  if (!out) return fail("null output handle");
  *out = new hps_beamline();
  return 0;
}

int hps_beamline_destroy(hps_beamline* bl)
{
  delete bl;
  return 0;
}

int hps_beamline_add(hps_beamline* bl, const char* name, const hps_element* e)
{
  if (!bl || !e) return fail("null argument");
  if (e->kind < HPS_DRIFT || e->kind > HPS_DIPOLE) return fail("unknown element kind");
  bl->names.push_back(name ? name : "");
  bl->el.push_back(*e);
  return 0;
}

int hps_beamline_set(hps_beamline* bl, uint32_t index, const hps_element* e)
{
  if (!bl || !e || index >= bl->el.size()) return fail("element index out of range");
  bl->el[index] = *e;
  return 0;
}

uint32_t hps_beamline_size(const hps_beamline* bl) { return bl ? (uint32_t)bl->el.size() : 0; }

int hps_beamline_index(const hps_beamline* bl, const char* name, uint32_t* out)
{
  if (!bl || !name || !out) return fail("null argument");
  std::string key(name);
  std::transform(key.begin(), key.end(), key.begin(), ::tolower);
  for (size_t i = 0; i < bl->names.size(); ++i) {
    std::string n = bl->names[i];
    std::transform(n.begin(), n.end(), n.begin(), ::tolower);
    if (n == key) { *out = (uint32_t)i; return 0; }
  }
  return fail(std::string("BeamLineElementNotFoundException: ") + name);
}

// ---- SpaceCharge ----
// Bound on adjacent bunches (each image is one more ring per Green's-table
// entry and one more point charge per off-mesh kick).
constexpr uint32_t kMaxAdjBunch = 1024;

static int spch_alloc_tables(hps_spch* s)
{
  (void)hipFree(s->fld1); (void)hipFree(s->fld2); (void)hipFree(s->bins); (void)hipFree(s->part);
  s->fld1 = nullptr; s->fld2 = nullptr; s->bins = nullptr; s->part = nullptr;
  const uint32_t nr = s->cfg.nx, nz = s->cfg.nz;
  const size_t n1 = (size_t)(nr + 1) * nr * nz * 2, n2 = (size_t)(nz + 1) * (nr + 1) * 2;
  HPS_CHECK(hipMalloc(&s->fld1, n1 * 8));
  HPS_CHECK(hipMalloc(&s->fld2, n2 * 8));
  HPS_CHECK(hipMalloc(&s->part, kFieldMaxSplit * n2 * 8));
  HPS_CHECK(hipMalloc(&s->bins, (size_t)nr * nz * 8));
  HPS_CHECK(hipMemset(s->fld1, 0, n1 * 8));
  HPS_CHECK(hipMemset(s->fld2, 0, n2 * 8));
  HPS_CHECK(hipMemset(s->bins, 0, (size_t)nr * nz * 8));
  return 0;
}

static int spch_push_cfg(hps_spch* s)
{
  ScDev d;
  HPS_CHECK(hipMemcpy(&d, s->dev, sizeof d, hipMemcpyDeviceToHost));
  hps_spch_state& st = d.st;
  st.nx = s->cfg.nx; st.nz = s->cfg.nz; st.nx_init = s->cfg.nx_init; st.nz_init = s->cfg.nz_init;
  st.adj_bunch = s->cfg.adj_bunch; st.fraction = s->cfg.fraction; st.interval = s->cfg.interval;
  st.interval_init = s->cfg.interval_init; st.adj_bunch_w = s->cfg.adj_bunch_w;
  st.mesh_w = s->cfg.mesh_w; st.remesh_val = s->cfg.remesh_val;
  HPS_CHECK(hipMemcpy(s->dev, &d, sizeof d, hipMemcpyHostToDevice));
  return 0;
}

int hps_spch_create(hps_spch** out, uint32_t nr, uint32_t nz, double interval, uint32_t adj_bunch)
{
  if (!out) return fail("null output handle");
  if (nr == 0 || nz == 0) return fail("SpaceCharge needs nr > 0 and nz > 0");
  if (adj_bunch > kMaxAdjBunch) return fail("adj_bunch is limited to 1024 bunches each side");
  hps_spch* s = new hps_spch();
  if (hipGetDevice(&s->device) != hipSuccess || stream_for(s->device, &s->stream)) {
    delete s;
    return fail("no HIP device");
  }
  // SpaceCharge / Scheff constructors, space_charge.cpp:13-40; SetInterval
  s->cfg.nx = nr; s->cfg.nz = nz; s->cfg.nx_init = nr; s->cfg.nz_init = nz;
  s->cfg.adj_bunch = adj_bunch; s->cfg.fraction = 1.0;
  s->cfg.interval = interval; s->cfg.interval_init = interval;
  s->cfg.remesh_val = 0.05;
  int lds = 0;
  s->max_lds = 65536;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) == hipSuccess &&
      lds > 0)
    s->max_lds = (size_t)lds;
  if (hipMalloc(&s->dev, sizeof(ScDev)) != hipSuccess ||
      hipMemset(s->dev, 0, sizeof(ScDev)) != hipSuccess) {
    (void)hipFree(s->dev);
    delete s;
    return fail("device allocation failed for the SpaceCharge");
  }
  if (spch_alloc_tables(s) || spch_push_cfg(s)) {   // these set the error message
    (void)hipFree(s->dev); (void)hipFree(s->fld1); (void)hipFree(s->fld2); (void)hipFree(s->bins);
    (void)hipFree(s->part);
    delete s;
    return -1;
  }
  *out = s;
  return 0;
}

int hps_spch_destroy(hps_spch* s)
{
  if (!s) return 0;
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(s->dev); (void)hipFree(s->fld1); (void)hipFree(s->fld2); (void)hipFree(s->bins);
  (void)hipFree(s->part);
  delete s;
  return 0;
}

int hps_spch_set(hps_spch* s, int what, double v)
{
  if (!s) return fail("null space charge");
  switch (what) {
    case 0: s->cfg.interval = v; s->cfg.interval_init = v; break;
    case 1: s->cfg.remesh_val = v; break;
    case 2:
      if (!(v >= 0.0 && v <= (double)kMaxAdjBunch))
        return fail("adj_bunch must lie in [0, 1024]");
      s->cfg.adj_bunch = (uint32_t)v;
      break;
    case 3: s->cfg.adj_bunch_w = v; break;
    case 4: s->cfg.mesh_w = v; break;
    default: return fail("unknown space-charge setting");
  }
  HPS_CHECK(hipStreamSynchronize(s->stream));
  return spch_push_cfg(s);
}

int hps_spch_set_mesh_size(hps_spch* s, uint32_t nr, uint32_t nz)   // space_charge.cpp:61-73
{
  if (!s || nr == 0 || nz == 0) return fail("bad mesh size");
  if (nr == s->cfg.nx && nz == s->cfg.nz) return 0;
  HPS_CHECK(hipStreamSynchronize(s->stream));
  s->cfg.nx = nr; s->cfg.nz = nz;
  if (spch_alloc_tables(s)) return -1;
  return spch_push_cfg(s);
}

int hps_spch_get_state(hps_spch* s, hps_spch_state* out)
{
  if (!s || !out) return fail("null argument");
  HPS_CHECK(hipStreamSynchronize(s->stream));
  ScDev d;
  HPS_CHECK(hipMemcpy(&d, s->dev, sizeof d, hipMemcpyDeviceToHost));
  *out = d.st;
  return 0;
}

int hps_spch_get_tables(hps_spch* s, double* fld1, double* fld2, int64_t* bins)
{
  if (!s) return fail("null space charge");
  HPS_CHECK(hipStreamSynchronize(s->stream));
  const uint32_t nr = s->cfg.nx, nz = s->cfg.nz;
  if (fld1) {   // device layout [rp][dz][rs] -> reference layout [rp][rs][dz]
    std::vector<double> tmp((size_t)(nr + 1) * nr * nz * 2);
    HPS_CHECK(hipMemcpy(tmp.data(), s->fld1, tmp.size() * 8, hipMemcpyDeviceToHost));
    for (uint32_t rp = 0; rp <= nr; ++rp)
      for (uint32_t dz = 0; dz < nz; ++dz)
        for (uint32_t rs = 0; rs < nr; ++rs) {
          const size_t src = ((size_t)rp * nz + dz) * nr + rs;
          const size_t dst = (size_t)rp * nr * nz + (size_t)rs * nz + dz;
          fld1[2 * dst] = tmp[2 * src];
          fld1[2 * dst + 1] = tmp[2 * src + 1];
        }
  }
  if (fld2) HPS_CHECK(hipMemcpy(fld2, s->fld2, (size_t)(nz + 1) * (nr + 1) * 16, hipMemcpyDeviceToHost));
  if (bins) HPS_CHECK(hipMemcpy(bins, s->bins, (size_t)nr * nz * 8, hipMemcpyDeviceToHost));
  return 0;
}

// ---- Simulator ----
int hps_sim_reset(hps_sim* m)   // Reset, simulation_engine.cu:101-111
{
  if (!m) return fail("null simulator");
  m->ccl_cell_num = 0;
  m->design_w = m->b->design_w;
  m->wavelen = CLIGHT / m->b->freq;
  return 0;
}

int hps_sim_create(hps_sim** out, hps_beam* b, hps_beamline* bl, hps_spch* s)
{
  if (!out || !b || !bl) return fail("Simulator needs a beam and a beamline");
  hps_sim* m = new hps_sim();
  m->b = b; m->bl = bl; m->sc = s; m->sc_on = s != nullptr;
  if (hipEventCreate(&m->ev0) != hipSuccess) {
    delete m;
    return fail("event creation failed");
  }
  if (hipEventCreate(&m->ev1) != hipSuccess) {
    (void)hipEventDestroy(m->ev0);
    delete m;
    return fail("event creation failed");
  }
  hps_sim_reset(m);
  m->prev_end = -1;
  *out = m;
  return 0;
}

int hps_sim_destroy(hps_sim* m)
{
  if (!m) return 0;
  (void)hipStreamSynchronize(m->b->stream);
  for (hipEvent_t e : m->evpool) (void)hipEventDestroy(e);
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  (void)hipFree(m->d_ops);
  (void)hipEventDestroy(m->ev0);
  (void)hipEventDestroy(m->ev1);
  delete m;
  return 0;
}

int hps_sim_set_space_charge(hps_sim* m, int on)
{
  if (!m) return fail("null simulator");
  m->sc_on = on != 0;
  return 0;
}

int hps_sim_set_profiling(hps_sim* m, int on)
{
  if (!m) return fail("null simulator");
  m->profile = on != 0;
  for (int k = 0; k < HPS_PROFILE_FAMILIES; ++k) { m->prof_ms[k] = 0; m->prof_n[k] = 0; }
  return 0;
}

int hps_sim_set_graph(hps_sim* m, int on)
{
  if (!m) return fail("null simulator");
  m->use_graph = on != 0;
  return 0;
}

int hps_sim_graph_replays(hps_sim* m, uint32_t* out)
{
  if (!m || !out) return fail("null argument");
  *out = m->graph_replays;
  return 0;
}

int hps_sim_get_profile(hps_sim* m, double* ms, uint32_t* launches)
{
  if (!m || !ms || !launches) return fail("null argument");
  for (int k = 0; k < HPS_PROFILE_FAMILIES; ++k) { ms[k] = m->prof_ms[k]; launches[k] = m->prof_n[k]; }
  return 0;
}

int hps_sim_last_time_ms(hps_sim* m, double* out)
{
  if (!m || !out) return fail("null argument");
  *out = m->last_ms;
  return 0;
}

}  // extern "C"

namespace {

// One launch unit of a simulate() call.
struct Step {
  int kind;            // 0 segment, 1 frequency change
  int32_t op_begin, op_end;
  int32_t kick;        // segment starts with the pending SCHEFF kick
  double kick_len;
  int32_t stats;       // StatKind computed by the segment kernel
  int32_t kick_point;  // Scheff::Start control + rebuild follow the segment
  double frac;         // current fraction of that kick point
  double new_freq;     // kind 1
};

struct Compiler {
  hps_sim* m;
  std::vector<Op> ops;
  std::vector<Step> steps;
  int32_t seg_begin = 0;
  bool pending_kick = false;
  double pending_len = 0;
  double frac = 1.0;
  double freq = 0;           // beam frequency as the program runs
  bool quad_gap_run = true;  // launches with blck/2 threads are empty for N = 1
  bool dipole_run = true;    // launches with blck/4 threads are empty for N <= 2
  bool sc_stats = true;      // UpdateStatForSpaceCharge runs only for N > 2

  void close(int32_t stats, bool kick_point, double kick_len_next)
  {
    Step s{};
    s.kind = 0;
    s.op_begin = seg_begin;
    s.op_end = (int32_t)ops.size();
    s.kick = pending_kick ? 1 : 0;
    s.kick_len = pending_len;
    s.stats = stats;
    s.kick_point = kick_point ? 1 : 0;
    s.frac = frac;
    steps.push_back(s);
    seg_begin = (int32_t)ops.size();
    pending_kick = kick_point;
    pending_len = kick_len_next;
  }

  // Scheff::Start at this point with kick length len.
  void kick_point(double len) { close(sc_stats ? STAT_SC : STAT_NONE, true, len); }
};

Op make_op(int32_t kind, uint32_t elem)
{
  Op o{};
  o.kind = kind;
  o.elem = elem;
  return o;
}

bool is_kind(const hps_beamline* bl, int64_t i, int kind)
{
  return i >= 0 && (uint64_t)i < bl->el.size() && bl->el[i].kind == kind;
}

int compile_element(Compiler& c, uint32_t id)
{
  hps_sim* m = c.m;
  const hps_element& e = m->bl->el[id];
  const bool sc = m->sc_on;
  switch (e.kind) {
    case HPS_DRIFT: {   // SimulateDrift, simulation_engine.cu:125-145
      int nk = 1;
      if (sc) {
        if (!m->sc) return fail("space charge is on but no SpaceCharge object was given");
        nk = (int)std::ceil(e.length / m->sc->cfg.interval);
      }
      const double h = 0.5 * e.length / nk;
      for (int k = 0; k < nk; ++k) {
        Op o = make_op(OP_DRIFT, id);
        o.a[0] = h; o.a[1] = e.aperture; o.a[2] = m->wavelen;
        c.ops.push_back(o);
        if (sc) c.kick_point(2.0 * h);
        c.ops.push_back(o);
      }
      return 0;
    }
    case HPS_QUAD: {   // SimulateQuad, simulation_engine.cu:147-161
      Op o = make_op(OP_QUAD, id);
      o.a[0] = e.length; o.a[1] = e.aperture; o.a[2] = e.gradient; o.a[3] = m->wavelen;
      if (c.quad_gap_run) c.ops.push_back(o);
      if (sc) {
        if (!m->sc) return fail("space charge is on but no SpaceCharge object was given");
        c.kick_point(e.length);
      }
      if (c.quad_gap_run) c.ops.push_back(o);
      return 0;
    }
    case HPS_GAP_DTL:
    case HPS_GAP_CCL: {
      const bool ccl = e.kind == HPS_GAP_CCL;
      // PrepSimulateRFGap, simulation_engine.cu:163-175
      if (c.freq != e.gap.frequency) {
        Step s{};
        s.kind = 1;
        s.new_freq = e.gap.frequency;
        c.close(STAT_NONE, false, 0);
        c.steps.push_back(s);
        c.freq = e.gap.frequency;
        m->wavelen = CLIGHT / e.gap.frequency;
      }
      const hps_beamline* bl = m->bl;
      const int64_t i = id;
      const double ref = e.gap.phase_ref;
      double phi_in = 0, phi_out = 0, q1 = 0, q2 = 0;
      uint32_t cell = 0;
      if (!ccl) {   // SimulateDTLRFGap, simulation_engine.cu:182-241
        if (i == 1 || (i > 1 && !is_kind(bl, i - 2, HPS_GAP_DTL))) {
          if (is_kind(bl, i + 2, HPS_GAP_DTL))
            phi_in = ref - 0.5 * (bl->el[i + 2].gap.phase_ref - ref);
          else
            phi_in = ref;   // only gap in tank (also for a final gap)
        } else if (i > 1) {
          phi_in = 0.5 * (bl->el[i - 2].gap.phase_ref + ref);
        } else {
          return fail("A quad is missing before the DTL gap!");
        }
        q1 = is_kind(bl, i - 1, HPS_QUAD) ? bl->el[i - 1].length : 0.0;
        q2 = is_kind(bl, i + 1, HPS_QUAD) ? bl->el[i + 1].length : 0.0;
        phi_out = phi_in;
      } else {      // SimulateCCLRFGap, simulation_engine.cu:243-293
        if (i == 0 || !is_kind(bl, i - 1, HPS_GAP_CCL)) m->ccl_cell_num = 0;
        const auto is_gap = [&](int64_t k) { return is_kind(bl, k, HPS_GAP_DTL) || is_kind(bl, k, HPS_GAP_CCL); };
        if (m->ccl_cell_num == 0) {
          if (!is_gap(i + 1)) return fail("a CCL tank needs a gap after its first gap");
          const double next = bl->el[i + 1].gap.phase_ref;
          phi_in = ref - 0.5 * (next - ref);
          phi_out = 0.5 * (next + ref);
        } else if (!is_kind(bl, i + 1, HPS_GAP_CCL)) {
          const double prev = bl->el[i - 1].gap.phase_ref;
          phi_in = 0.5 * (ref + prev);
          phi_out = ref + 0.5 * (ref - prev);
        } else {
          phi_in = 0.5 * (ref + bl->el[i - 1].gap.phase_ref);
          phi_out = 0.5 * (ref + bl->el[i + 1].gap.phase_ref);
        }
        cell = m->ccl_cell_num++;
      }
      const bool run = ccl || c.quad_gap_run;
      const GapUniform gu = gap_uniform(e.gap, m->design_w, phi_in, ccl ? phi_out : phi_in,
                                        e.length, q1, q2, cell, ccl, m->b->mass,
                                        m->b->charge);
      Op o1 = make_op(OP_GAP1, id);
      o1.ccl = ccl; o1.g = e.gap; o1.gu = gu;
      o1.a[0] = m->design_w; o1.a[1] = phi_in; o1.a[2] = e.length; o1.a[3] = q1; o1.a[4] = q2;
      if (run) {
        c.ops.push_back(o1);
        m->wavelen = CLIGHT / e.gap.frequency;   // every thread writes d_wavelen
      }
      if (sc) {
        if (!m->sc) return fail("space charge is on but no SpaceCharge object was given");
        c.kick_point(e.length);
      }
      Op o2 = make_op(OP_GAP2, id);
      o2.ccl = ccl; o2.ccl_cell_num = cell; o2.g = e.gap; o2.gu = gu;
      o2.a[0] = m->design_w; o2.a[1] = ccl ? phi_out : phi_in; o2.a[2] = e.length;
      o2.a[3] = q1; o2.a[4] = q2; o2.a[5] = m->wavelen;
      if (run) c.ops.push_back(o2);
      m->design_w = e.gap.energy_out;
      m->b->design_w = m->design_w;
      return 0;
    }
    case HPS_APERTURE_C:
    case HPS_APERTURE_R: {   // simulation_engine.cu:307-329
      if (!e.aperture_in) return 0;
      c.close(STAT_APERTURE, false, 0);
      Op o = make_op(e.kind == HPS_APERTURE_C ? OP_APER_C : OP_APER_R, id);
      if (e.kind == HPS_APERTURE_C) o.a[0] = e.aperture;
      else { o.a[0] = e.aper_xl; o.a[1] = e.aper_xr; o.a[2] = e.aper_yt; o.a[3] = e.aper_yb; }
      c.ops.push_back(o);
      return 0;
    }
    case HPS_SPCH_COMP:   // simulation_engine.cu:301-305
      if (sc && m->sc) { c.frac = e.fraction; m->sc->cfg.fraction = e.fraction; }
      return 0;
    case HPS_ROTATION: {
      Op o = make_op(OP_ROT, id);
      o.a[0] = e.angle;
      c.ops.push_back(o);
      return 0;
    }
    case HPS_STEERER: {   // simulation_engine.cu:337-344
      if (e.bl_h > 1e-10 || e.bl_v > 1e-10) {
        Op o = make_op(OP_STEER, id);
        o.a[0] = e.bl_h; o.a[1] = e.bl_v;
        c.ops.push_back(o);
      }
      return 0;
    }
    case HPS_DIAGNOSTICS: return 0;
    case HPS_BUNCHER: {   // SimulateBuncher, simulation_engine.cu:346-358
      if (!e.on) return 0;
      m->wavelen = CLIGHT / e.frequency;   // UpdateWaveLengthKernel(bfreq)
      Op o = make_op(OP_BUNCH, id);
      o.a[0] = m->design_w; o.a[1] = e.frequency; o.a[2] = e.voltage; o.a[3] = e.phase;
      o.a[4] = e.aperture; o.a[5] = e.frequency / c.freq; o.a[6] = m->wavelen;
      c.ops.push_back(o);
      return 0;
    }
    case HPS_DIPOLE: {    // SimulateDipole, simulation_engine.cu:360-374
      const bool run = c.dipole_run;   // launches <<<grid, blck/4>>>
      Op o = make_op(OP_DIP1, id);
      o.a[0] = e.rho; o.a[1] = e.angle; o.a[2] = e.edge_angle_in; o.a[3] = e.edge_angle_out;
      o.a[4] = e.half_gap; o.a[5] = e.k1; o.a[6] = e.k2; o.a[7] = e.field_index;
      o.a[8] = e.kinetic_energy; o.a[9] = m->wavelen;
      if (run) c.ops.push_back(o);
      if (sc) {
        if (!m->sc) return fail("space charge is on but no SpaceCharge object was given");
        c.kick_point(e.rho * e.angle);
      }
      o.kind = OP_DIP2;
      if (run) c.ops.push_back(o);
      return 0;
    }
    default: return fail("unknown element kind");
  }
}

}  // namespace

// Scheff::Start's mesh-size cutoff (space_charge.cpp:107-117), decided on
// the host from the beam's good mean energy of this kick point: above
// mesh_w half the mesh in each direction (tables reallocated zeroed, as
// SetMeshSize does) and four times the kick interval.
static int mesh_cutoff_step(hps_sim* m)
{
  hps_spch* sc = m->sc;
  hps_beam* b = m->b;
  HPS_CHECK(hipStreamSynchronize(b->stream));
  double ke = 0.0;
  HPS_CHECK(hipMemcpy(&ke, (const char*)b->st + offsetof(hps_beam_stats, w_avg_good), sizeof ke,
                      hipMemcpyDeviceToHost));
  hps_spch_state& c = sc->cfg;
  const double mw = c.mesh_w;
  if (ke > mw && (c.nx != c.nx_init / 2 || c.nz != c.nz_init / 2)) {
    if (hps_spch_set_mesh_size(sc, c.nx_init / 2, c.nz_init / 2)) return -1;
  } else if (ke <= mw && (c.nx != c.nx_init || c.nz != c.nz_init)) {
    if (hps_spch_set_mesh_size(sc, c.nx_init, c.nz_init)) return -1;
  }
  bool push = false;
  if (ke > mw && c.interval != c.interval_init * 4) {
    c.interval = c.interval_init * 4;
    push = true;
  } else if (ke <= mw && c.interval != c.interval_init) {
    c.interval = c.interval_init;
    push = true;
  }
  if (push) return spch_push_cfg(sc);
  return 0;
}

extern "C" int hps_sim_simulate(hps_sim* m, int32_t start, int32_t end)
{
  if (!m) return fail("null simulator");
  hps_beam* b = m->b;
  const uint32_t nel = (uint32_t)m->bl->el.size();
  if (nel == 0) return 0;
  const int32_t start_index = start < 0 ? 0 : start;
  const int32_t end_index = end < 0 ? (int32_t)nel - 1 : end;
  if (end_index >= (int32_t)nel || start_index >= (int32_t)nel) return fail("element index out of range");
  // with a mesh-size cutoff energy the kick interval and the mesh follow the
  // beam energy kick by kick: the program is compiled and issued kick point
  // by kick point, with one host read of the mean energy at each
  const bool stepped = m->sc_on && m->sc && m->sc->cfg.mesh_w != 0.0;
  if (m->sc_on && m->sc) {
    const hps_spch_state& cf = m->sc->cfg;
    if (stepped && (cf.nx_init < 2 || cf.nz_init < 2))
      return fail("mesh_size_cutoff_w needs a mesh of at least 2 x 2");
    // the deposit keeps one int64 histogram of nr * nz bins per workgroup in LDS
    const size_t bins = std::max((size_t)cf.nx * cf.nz, stepped ? (size_t)cf.nx_init * cf.nz_init : 0);
    if (bins * 8 + kBlockLdsReserve > m->sc->max_lds)
      return fail("the SCHEFF mesh has more bins than one workgroup's LDS histogram holds");
  }
  // SimulationEngine::Simulate, simulation_engine.cpp:73-95
  if (start_index < m->prev_end) hps_sim_reset(m);
  m->prev_end = end_index;
  Compiler c;
  c.m = m;
  c.frac = m->sc ? m->sc->cfg.fraction : 1.0;
  c.freq = b->freq;
  c.quad_gap_run = b->blck_size / 2 != 0;
  c.dipole_run = b->blck_size / 4 != 0;
  c.sc_stats = b->n > 2;

  hipStream_t s = b->stream;
  const uint32_t kstride = b->grid_size * (b->blck_size / 4);
  const bool kick_ok = kstride != 0;
  const bool deposit_ok = b->blck_size / 4 != 0;

  // The launch list of this call.
  // kinds 0..5 are the profiling families of the header; L_HOST is a host step
  enum { L_SEG = 0, L_PHASE, L_BUILD, L_FIELD, L_SEG_FINAL, L_PHASE_FINAL, L_HOST };
  struct Launch {
    int kind;
    SegArgs seg;
    PhaseArgs ph;
    double new_freq;
  };
  std::vector<Launch> launches;
  bool host_steps = false;
  auto add_step = [&](const Step& st) {
    if (st.kind == 1) {
      Launch l{};
      l.kind = L_HOST;
      l.new_freq = st.new_freq;
      launches.push_back(l);
      host_steps = true;
      return;
    }
    if (st.op_end > st.op_begin || st.kick || st.stats != STAT_NONE) {
      Launch l{};
      l.kind = L_SEG;
      SegArgs& a = l.seg;
      a.b = b->d;
      a.st = b->st;
      a.pr = b->pr;
      a.ops = m->d_ops;
      a.op_begin = st.op_begin;
      a.op_end = st.op_end;
      a.kick = (st.kick && kick_ok && m->sc) ? 1 : 0;
      a.kick_mask = kstride ? kstride - 1 : 0;
      a.kick_len = st.kick_len;
      a.sc = m->sc ? m->sc->dev : nullptr;
      a.fld2 = m->sc ? m->sc->fld2 : nullptr;
      a.stats = st.stats;
      a.rmax = (st.stats == STAT_SC && m->sc && m->sc->cfg.adj_bunch != 0) ? 1 : 0;
      a.op_class = OPC_LIGHT;
      for (int32_t o = st.op_begin; o < st.op_end; ++o)
        a.op_class = std::max(a.op_class, op_class_of(c.ops[o].kind));
      launches.push_back(l);
      if (a.stats != STAT_NONE) {
        l.kind = L_SEG_FINAL;
        launches.push_back(l);
      }
    }
    if (st.kick_point) {
      hps_spch* sc = m->sc;
      Launch l{};
      l.kind = L_PHASE;
      PhaseArgs& p = l.ph;
      p.b = b->d; p.st = b->st; p.pr = b->pr; p.sc = sc->dev; p.bins = sc->bins;
      p.do_sums = b->n > 2 ? 1 : 0;
      p.rmax = sc->cfg.adj_bunch == 0 ? 0 : (c.sc_stats ? 1 : 2);
      p.frac = st.frac; p.current = b->current; p.freq = b->freq;
      if (p.do_sums) launches.push_back(l);
      Launch g = l;
      g.kind = L_PHASE_FINAL;
      launches.push_back(g);
      g.kind = L_BUILD;
      launches.push_back(g);
      g.kind = L_FIELD;
      launches.push_back(g);
    }
  };
  auto issue = [&](const Launch& l, hipStream_t st) -> hipError_t {
    hps_spch* sc = m->sc;
    switch (l.kind) {
      case L_SEG: return launch_segment(l.seg, st);
      case L_SEG_FINAL: return launch_seg_final(l.seg, st);
      case L_PHASE: return launch_phase(l.ph, st);
      case L_PHASE_FINAL: return launch_phase_final(l.ph, st);
      case L_BUILD:
        return launch_sc_build(sc->dev, sc->fld1, sc->cfg.nx, sc->cfg.nz, b->d, b->st, sc->bins,
                               deposit_ok, st);
      default: {
        FieldArgs f;
        f.sc = sc->dev; f.bins = sc->bins; f.fld1 = sc->fld1; f.fld2 = sc->fld2; f.part = sc->part;
        return launch_field(f, sc->cfg.nx, sc->cfg.nz, sc->max_lds, st);
      }
    }
  };

  // profiling: HIP events around every launch (eager launches only)
  std::vector<std::pair<int, size_t>> marks;
  size_t nev = 0;
  // one eager launch; host steps run in order with the launches
  auto issue_eager = [&](const Launch& l) -> int {
    if (l.kind == L_HOST) {
      // the phases of a frequency change are not known at compile time of
      // later launches' arguments: rebuild them after the host step
      return change_frequency(b, l.new_freq);
    }
    if (m->profile) {
      while (m->evpool.size() < nev + 2) {
        hipEvent_t e;
        HPS_CHECK(hipEventCreate(&e));
        m->evpool.push_back(e);
      }
      marks.push_back({l.kind, nev});
      HPS_CHECK(hipEventRecord(m->evpool[nev], s));
    }
    Launch lc = l;
    if (l.kind == L_PHASE || l.kind == L_PHASE_FINAL) {
      lc.ph.current = b->current;
      lc.ph.freq = b->freq;
      lc.ph.bins = m->sc->bins;   // a mesh-size step may have reallocated it
    }
    HPS_CHECK(issue(lc, s));
    if (m->profile) {
      HPS_CHECK(hipEventRecord(m->evpool[nev + 1], s));
      nev += 2;
    }
    return 0;
  };

  HPS_CHECK(hipEventRecord(m->ev0, s));
  if (stepped) {
    // compile one element, then issue every closed step; the map table grows
    // on the device as the program does
    size_t done = 0;
    auto flush = [&]() -> int {
      for (; done < c.steps.size(); ++done) {
        const Step& st = c.steps[done];
        if (st.kind == 0 && st.op_end > st.op_begin) {
          if ((size_t)st.op_end > m->ops_cap) {
            HPS_CHECK(hipStreamSynchronize(s));
            (void)hipFree(m->d_ops);
            m->d_ops = nullptr;
            m->ops_cap = 0;
            HPS_CHECK(hipMalloc(&m->d_ops, (2 * (size_t)st.op_end + 64) * sizeof(Op)));
            m->ops_cap = 2 * (size_t)st.op_end + 64;
          }
          HPS_CHECK(hipMemcpyAsync(m->d_ops + st.op_begin, c.ops.data() + st.op_begin,
                                   (size_t)(st.op_end - st.op_begin) * sizeof(Op),
                                   hipMemcpyHostToDevice, s));
          // the host vector may move when the program grows
          HPS_CHECK(hipStreamSynchronize(s));
        }
        launches.clear();
        add_step(st);
        for (const Launch& l : launches) {
          if (l.kind == L_PHASE_FINAL && mesh_cutoff_step(m)) return -1;
          if (issue_eager(l)) return -1;
        }
      }
      return 0;
    };
    for (uint32_t i = 0; (int32_t)i <= end_index; ++i)
      if ((int32_t)i >= start_index || m->bl->el[i].kind == HPS_SPCH_COMP) {
        if (compile_element(c, i) || flush()) return -1;
      }
    c.close(STAT_NONE, false, 0);
    if (flush()) return -1;
  } else {
    for (uint32_t i = 0; (int32_t)i <= end_index; ++i)
      if ((int32_t)i >= start_index || m->bl->el[i].kind == HPS_SPCH_COMP)
        if (compile_element(c, i)) return -1;
    c.close(STAT_NONE, false, 0);

    if (c.ops.size() > m->ops_cap) {
      (void)hipFree(m->d_ops);
      m->d_ops = nullptr;
      m->ops_cap = 0;
      HPS_CHECK(hipMalloc(&m->d_ops, (c.ops.size() + 64) * sizeof(Op)));
      m->ops_cap = c.ops.size() + 64;
    }
    if (!c.ops.empty())
      HPS_CHECK(hipMemcpyAsync(m->d_ops, c.ops.data(), c.ops.size() * sizeof(Op),
                               hipMemcpyHostToDevice, s));
    for (const Step& st : c.steps) add_step(st);

    const bool eager = m->profile || host_steps || !m->use_graph;
    if (eager) {
      for (const Launch& l : launches)
        if (issue_eager(l)) return -1;
    } else {
      std::vector<uint8_t> sig;
      const auto put = [&](const void* p, size_t nb) {
        const uint8_t* q = (const uint8_t*)p;
        sig.insert(sig.end(), q, q + nb);
      };
      for (const Launch& l : launches) {
        put(&l.kind, sizeof l.kind);
        if (l.kind == L_SEG || l.kind == L_SEG_FINAL) put(&l.seg, sizeof l.seg);
        else put(&l.ph, sizeof l.ph);
      }
      put(c.ops.data(), c.ops.size() * sizeof(Op));
      if (m->sc) {
        const hps_spch* sc = m->sc;
        const void* ptrs[4] = {sc->dev, sc->fld1, sc->fld2, sc->bins};
        put(ptrs, sizeof ptrs);
        put(&sc->cfg.nx, sizeof sc->cfg.nx);
        put(&sc->cfg.nz, sizeof sc->cfg.nz);
        put(&sc->max_lds, sizeof sc->max_lds);
      }
      if (!m->graph_exec || sig != m->graph_sig) {
        if (m->graph_exec) { (void)hipGraphExecDestroy(m->graph_exec); m->graph_exec = nullptr; }
        if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
        HPS_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
        hipError_t err = hipSuccess;
        for (const Launch& l : launches) {
          err = issue(l, s);
          if (err != hipSuccess) break;
        }
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (err != hipSuccess) return fail(std::string("graph capture: ") + hipGetErrorString(err));
        HPS_CHECK(e2);
        m->graph = g;
        HPS_CHECK(hipGraphInstantiate(&m->graph_exec, g, nullptr, nullptr, 0));
        m->graph_sig.swap(sig);
      } else {
        m->graph_replays += 1;
      }
      HPS_CHECK(hipGraphLaunch(m->graph_exec, s));
    }
  }
  HPS_CHECK(hipEventRecord(m->ev1, s));
  HPS_CHECK(hipEventSynchronize(m->ev1));
  float ms = 0;
  HPS_CHECK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  m->last_ms = ms;
  for (const auto& mk : marks) {
    float t = 0;
    HPS_CHECK(hipEventElapsedTime(&t, m->evpool[mk.second], m->evpool[mk.second + 1]));
    m->prof_ms[mk.first] += t;
    m->prof_n[mk.first] += 1;
  }
  HPS_CHECK(hipGetLastError());
  return 0;
}

// ---- Batch: several simulations, one grid per launch ----

struct hps_batch {
  std::vector<hps_sim*> sims;
  // field-sum launch shape (UINT32_MAX: not chosen yet); autotune times the
  // candidates at the first rebuild and keeps the fastest
  bool autotune = true;
  uint32_t field_shape = UINT32_MAX;
  uint32_t field_key[3] = {0, 0, 0};   // nr, nz, batch size the shape was chosen for
  std::vector<float> field_times;      // ms per candidate of the last tuning
  std::vector<uint8_t> host_rec;      // launch records + block maps of the last call
  uint8_t* d_rec = nullptr;
  size_t rec_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  bool profile = false;
  std::vector<hipEvent_t> evpool;
  double prof_ms[HPS_PROFILE_FAMILIES] = {};
  uint32_t prof_n[HPS_PROFILE_FAMILIES] = {};
  // grouped field solve (field_mode 1): class tables, charges, partial sums
  int field_mode = 0;
  double aspect_step = 0.05, aspect_lnr = 0.0;
  uint32_t gemm_key[3] = {0, 0, 0};    // nr, nz, members the buffers were sized for
  double gemm_key_lnr = 0.0;           // aspect grid the class tables hold
  GemmMember* d_mem = nullptr;
  GemmState* d_gs = nullptr;
  uint32_t* d_perm = nullptr;
  float *d_tables = nullptr, *d_q = nullptr, *d_w = nullptr;
  MemberMap mm{};
  // surrogate field network (grouped mode): inputs, device control
  hps_nn* sur = nullptr;
  hps_nn* capture = nullptr;
  SurCtl sur_cfg{};
  SurCtl* d_sur = nullptr;
  float *d_xinf = nullptr, *d_uinf = nullptr;
  FftBufs fft{};                       // spectral product buffers (tw null: unavailable)
  uint32_t gemm_shape = UINT32_MAX;    // chosen product shape (UINT32_MAX: tune)
  uint32_t gemm_shape_key[3] = {0, 0, 0};
  std::vector<float> gemm_times;
  uint32_t cus = 0;
  // lattice-period graph: [0] record offset of the replayed period, [1] zero
  uint64_t* d_ctl = nullptr;
  bool use_graph = true;
  hipGraphExec_t pgraph_exec = nullptr;
  std::vector<uint8_t> pgraph_sig;
  uint64_t period_replays = 0;
  uint32_t period_launches = 0;
};

// One launch of a batch program for the period search: its kind and map
// class, whether it reads per-step records, and their offset.
struct PeriodKey {
  int kind, op_class;
  bool carries;
  size_t rec;
};
// A run of n >= 2 repetitions of the p launches starting at a, whose record
// offsets advance by stride bytes per repetition.
struct PeriodRun {
  size_t a, p;
  uint32_t n;
  uint64_t stride;
};

// The longest periodic run (most launches covered, then the shortest
// period) with a period of at most 128 launches starting within the first
// two periods.
static PeriodRun find_period_run(const std::vector<PeriodKey>& k)
{
  PeriodRun best{0, 0, 0, 0};
  const size_t L = k.size();
  for (size_t p = 1; p <= 128 && 2 * p <= L; ++p) {
    for (size_t a = 0; a < 2 * p && a + 2 * p <= L; ++a) {
      size_t j0 = a;
      while (j0 < a + p && !k[j0].carries) ++j0;
      if (j0 == a + p) continue;
      const uint64_t stride = (uint64_t)(k[j0 + p].rec - k[j0].rec);
      uint32_t n = 1;
      for (;;) {
        const size_t b = a + (size_t)n * p;
        if (b + p > L) break;
        bool same = true;
        for (size_t j = b; j < b + p && same; ++j) {
          const PeriodKey &x = k[j], &y = k[j - p];
          same = x.kind == y.kind && x.op_class == y.op_class && x.carries == y.carries &&
                 (!x.carries || x.rec - y.rec == stride) && (x.carries || x.rec == y.rec);
        }
        if (!same) break;
        ++n;
      }
      if (n >= 2 && (size_t)n * p > (size_t)best.n * best.p) best = {a, p, n, stride};
    }
  }
  return best;
}

// Chunk map of a batch's segment and phase launches.  The workgroups of a
// launch go to the part's eight XCDs round robin by index (workgroup b on
// XCD b mod 8) and each XCD has its own L2, so the members are dealt to
// eight groups with balanced chunk totals (largest member first, each to the
// lightest group) and block 8 j + x takes the j-th chunk of group x: all
// chunks of a member run on one XCD, whose L2 then holds the field tables
// the kicks gather from for its members only.  Any order gives the same
// results (every chunk writes its own partial sums).
static std::vector<uint32_t> xcd_chunk_map(const std::vector<uint32_t>& nch)
{
  constexpr uint32_t kXcds = 8;
  const size_t B = nch.size();
  std::vector<size_t> order(B);
  for (size_t k = 0; k < B; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return nch[a] > nch[b]; });
  std::vector<std::vector<uint32_t>> grp(kXcds);
  std::vector<uint64_t> load(kXcds, 0);
  size_t total = 0;
  for (size_t k : order) {
    uint32_t x = 0;
    for (uint32_t y = 1; y < kXcds; ++y)
      if (load[y] < load[x]) x = y;
    for (uint32_t c = 0; c < nch[k]; ++c) grp[x].push_back(((uint32_t)k << kBatchLocalBits) | c);
    load[x] += nch[k];
    total += nch[k];
  }
  std::vector<uint32_t> map;
  map.reserve(total);
  for (size_t j = 0; map.size() < total; ++j)
    for (uint32_t x = 0; x < kXcds; ++x)
      if (j < grp[x].size()) map.push_back(grp[x][j]);
  return map;
}

static void free_gemm(hps_batch* bt)
{
  (void)hipFree(bt->d_mem); (void)hipFree(bt->d_gs); (void)hipFree(bt->d_perm);
  (void)hipFree(bt->d_tables); (void)hipFree(bt->d_q); (void)hipFree(bt->d_w);
  (void)hipFree(bt->mm.pos_exact); (void)hipFree(bt->mm.pos_nn); (void)hipFree(bt->mm.fallback);
  (void)hipFree(bt->mm.qtot); (void)hipFree(bt->d_xinf); (void)hipFree(bt->d_uinf);
  (void)hipFree(bt->fft.tw); (void)hipFree(bt->fft.shat); (void)hipFree(bt->fft.spec_class);
  (void)hipFree(bt->fft.list); (void)hipFree(bt->fft.qhat); (void)hipFree(bt->fft.what);
  bt->fft = FftBufs{};
  bt->mm = MemberMap{};
  bt->d_xinf = nullptr; bt->d_uinf = nullptr;
  bt->d_mem = nullptr; bt->d_gs = nullptr; bt->d_perm = nullptr;
  bt->d_tables = nullptr; bt->d_q = nullptr; bt->d_w = nullptr;
  bt->gemm_key[0] = bt->gemm_key[1] = bt->gemm_key[2] = 0;
}

// Buffers of the grouped field solve for B members on an nr x nz mesh; the
// class-table cache starts empty (a new mesh size or aspect grid discards it).
static int alloc_gemm(hps_batch* bt, uint32_t nr, uint32_t nz, uint32_t B, hipStream_t s)
{
  if (bt->d_gs && bt->gemm_key[0] == nr && bt->gemm_key[1] == nz && bt->gemm_key[2] >= B &&
      bt->gemm_key_lnr == bt->aspect_lnr)
    return 0;
  HPS_CHECK(hipStreamSynchronize(s));
  free_gemm(bt);
  const size_t slot_sz = (size_t)(nr + 1) * nr * 4 * nz;
  const size_t M = 2 * (size_t)(nr + 1) * (nz + 1), K = (size_t)nr * nz;
  HPS_CHECK(hipMalloc(&bt->d_mem, B * sizeof(GemmMember)));
  HPS_CHECK(hipMalloc(&bt->d_gs, sizeof(GemmState)));
  // kGemmCols product columns per member (its aspect classes)
  HPS_CHECK(hipMalloc(&bt->d_perm, kGemmCols * B * sizeof(uint32_t)));
  HPS_CHECK(hipMalloc(&bt->d_tables, kGemmSlots * slot_sz * sizeof(float)));
  HPS_CHECK(hipMalloc(&bt->d_q, kGemmCols * B * K * sizeof(float)));
  HPS_CHECK(hipMalloc(&bt->d_w, kGemmMaxSplits * kGemmCols * B * M * sizeof(float)));
  HPS_CHECK(hipMalloc(&bt->mm.pos_exact, kGemmCols * B * sizeof(int32_t)));
  HPS_CHECK(hipMalloc(&bt->mm.pos_nn, B * sizeof(int32_t)));
  HPS_CHECK(hipMalloc(&bt->mm.fallback, B * sizeof(uint32_t)));
  HPS_CHECK(hipMemset(bt->mm.fallback, 0, B * sizeof(uint32_t)));
  HPS_CHECK(hipMalloc(&bt->mm.qtot, B * sizeof(float)));
  HPS_CHECK(hipMalloc(&bt->d_xinf, B * (size_t)(nz + kZPad) * nr * sizeof(float)));
  HPS_CHECK(hipMemset(bt->d_xinf, 0, B * (size_t)(nz + kZPad) * nr * sizeof(float)));
  HPS_CHECK(hipMalloc(&bt->d_uinf, B * sizeof(float)));
  GemmState g{};
  for (uint32_t k = 0; k < kGemmSlots; ++k) g.slot_class[k] = INT32_MIN;
  HPS_CHECK(hipMemcpy(bt->d_gs, &g, sizeof g, hipMemcpyHostToDevice));
  if (fft_product_supported(nr, nz)) {
    // spectral product: twiddles in FP64 rounded to FP32, empty spectrum cache
    const size_t L = 2 * (size_t)nz, F = nz + 1, R = 2 * (size_t)(nr + 1);
    std::vector<float2> tw(L / 2);
    for (size_t k = 0; k < L / 2; ++k) {
      const double a = -2.0 * M_PI * (double)k / (double)L;
      tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    std::vector<int32_t> none(kGemmSlots, INT32_MIN);
    HPS_CHECK(hipMalloc(&bt->fft.tw, tw.size() * sizeof(float2)));
    HPS_CHECK(hipMemcpy(bt->fft.tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    HPS_CHECK(hipMalloc(&bt->fft.shat, kGemmSlots * F * R * nr * sizeof(float2)));
    HPS_CHECK(hipMalloc(&bt->fft.spec_class, kGemmSlots * sizeof(int32_t)));
    HPS_CHECK(hipMemcpy(bt->fft.spec_class, none.data(), kGemmSlots * sizeof(int32_t),
                        hipMemcpyHostToDevice));
    HPS_CHECK(hipMalloc(&bt->fft.list, (kGemmSlots + 1) * sizeof(uint32_t)));
    HPS_CHECK(hipMalloc(&bt->fft.qhat, kGemmCols * (size_t)B * F * nr * sizeof(float2)));
    HPS_CHECK(hipMalloc(&bt->fft.what, kGemmCols * (size_t)B * F * R * sizeof(float2)));
  }
  bt->gemm_key[0] = nr; bt->gemm_key[1] = nz; bt->gemm_key[2] = B;
  bt->gemm_key_lnr = bt->aspect_lnr;
  return 0;
}

// Times every product shape once on the live operands of a kick point whose
// schedule, class tables and charges are in place, and keeps the fastest;
// the caller then runs the chosen shape for the results.
static int tune_gemm(hps_batch* bt, uint32_t nr, uint32_t nz, uint32_t B, hipStream_t s)
{
  const uint32_t n = gemm_shapes();
  const FftBufs* fb = bt->fft.tw ? &bt->fft : nullptr;
  std::vector<hipEvent_t> ev(2 * n);
  for (auto& e : ev) HPS_CHECK(hipEventCreate(&e));
  // untimed first launches: the direct product, and the spectral one, whose
  // first launch also transforms the class tables in use (kept for later)
  HPS_CHECK(launch_gemm_product(bt->d_gs, bt->d_tables, bt->d_q, bt->d_w, nr, nz, 0, bt->cus,
                                fb, s));
  for (uint32_t k = 0; k < n; ++k)
    if (gemm_shape(k).spectral && fb)
      HPS_CHECK(launch_gemm_product(bt->d_gs, bt->d_tables, bt->d_q, bt->d_w, nr, nz, k,
                                    bt->cus, fb, s));
  for (uint32_t k = 0; k < n; ++k) {
    HPS_CHECK(hipEventRecord(ev[2 * k], s));
    if (!gemm_shape(k).spectral || fb)
      HPS_CHECK(launch_gemm_product(bt->d_gs, bt->d_tables, bt->d_q, bt->d_w, nr, nz, k,
                                    bt->cus, fb, s));
    HPS_CHECK(hipEventRecord(ev[2 * k + 1], s));
  }
  HPS_CHECK(hipEventSynchronize(ev[2 * n - 1]));
  bt->gemm_times.assign(n, 0.0f);
  uint32_t best = 0;
  for (uint32_t k = 0; k < n; ++k) {
    HPS_CHECK(hipEventElapsedTime(&bt->gemm_times[k], ev[2 * k], ev[2 * k + 1]));
    if (gemm_shape(k).spectral && !fb) continue;   // not available for this mesh
    if (bt->gemm_times[k] < bt->gemm_times[best]) best = k;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  bt->gemm_shape = best;
  bt->gemm_shape_key[0] = nr; bt->gemm_shape_key[1] = nz; bt->gemm_shape_key[2] = B;
  return 0;
}

// Times every field-sum launch shape once on the batch's live tables (the
// rebuild of a kick point the control step marked for remesh, so the kernels
// do the full work) and keeps the fastest in bt->field_shape.  Needs eager
// launches; each candidate overwrites the field tables with the same values.
static int tune_field(hps_batch* bt, const FieldArgs* d_field, uint32_t nmem, uint32_t nr,
                      uint32_t nz, size_t max_lds, hipStream_t s)
{
  const uint32_t n = field_batch_shapes();
  std::vector<hipEvent_t> ev(2 * n);
  for (auto& e : ev) HPS_CHECK(hipEventCreate(&e));
  // one untimed launch first (instruction caches, table residency)
  HPS_CHECK(launch_field_batch(d_field, nmem, nr, nz, max_lds, field_batch_default_shape(nr, nmem), s));
  for (uint32_t k = 0; k < n; ++k) {
    HPS_CHECK(hipEventRecord(ev[2 * k], s));
    HPS_CHECK(launch_field_batch(d_field, nmem, nr, nz, max_lds, k, s));
    HPS_CHECK(hipEventRecord(ev[2 * k + 1], s));
  }
  HPS_CHECK(hipEventSynchronize(ev[2 * n - 1]));
  bt->field_times.assign(n, 0.0f);
  uint32_t best = 0;
  for (uint32_t k = 0; k < n; ++k) {
    HPS_CHECK(hipEventElapsedTime(&bt->field_times[k], ev[2 * k], ev[2 * k + 1]));
    if (bt->field_times[k] < bt->field_times[best]) best = k;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  bt->field_shape = best;
  bt->field_key[0] = nr; bt->field_key[1] = nz; bt->field_key[2] = nmem;
  return 0;
}

extern "C" int hps_batch_create(hps_batch** out)
{
  if (!out) return fail("null output handle");
  hps_batch* bt = new hps_batch();
  if (hipEventCreate(&bt->ev0) != hipSuccess || hipEventCreate(&bt->ev1) != hipSuccess) {
    delete bt;
    return fail("event creation failed");
  }
  *out = bt;
  return 0;
}

extern "C" int hps_batch_destroy(hps_batch* bt)
{
  if (!bt) return 0;
  if (!bt->sims.empty()) (void)hipStreamSynchronize(bt->sims[0]->b->stream);
  for (hipEvent_t e : bt->evpool) (void)hipEventDestroy(e);
  free_gemm(bt);
  if (bt->pgraph_exec) (void)hipGraphExecDestroy(bt->pgraph_exec);
  (void)hipFree(bt->d_ctl);
  (void)hipFree(bt->d_sur);
  (void)hipFree(bt->d_rec);
  (void)hipEventDestroy(bt->ev0);
  (void)hipEventDestroy(bt->ev1);
  delete bt;
  return 0;
}

extern "C" int hps_batch_add(hps_batch* bt, hps_sim* m)
{
  if (!bt || !m) return fail("null argument");
  if (bt->sims.size() >= (1u << (32 - kBatchLocalBits)) - 1) return fail("batch is full");
  for (hps_sim* o : bt->sims) {
    if (o == m) return fail("simulator is already in the batch");
    if (o->b == m->b) return fail("two simulators of a batch share one beam");
    if (m->sc && o->sc == m->sc) return fail("two simulators of a batch share one SpaceCharge");
  }
  if (!bt->sims.empty()) {
    const hps_beam* b0 = bt->sims[0]->b;
    if (m->b->device != b0->device) return fail("the simulators of a batch must share one device");
  }
  bt->sims.push_back(m);
  return 0;
}

extern "C" uint32_t hps_batch_size(const hps_batch* bt) { return bt ? (uint32_t)bt->sims.size() : 0; }

extern "C" int hps_batch_set_profiling(hps_batch* bt, int on)
{
  if (!bt) return fail("null batch");
  bt->profile = on != 0;
  for (int k = 0; k < HPS_PROFILE_FAMILIES; ++k) { bt->prof_ms[k] = 0; bt->prof_n[k] = 0; }
  return 0;
}

extern "C" int hps_batch_get_profile(hps_batch* bt, double* ms, uint32_t* launches)
{
  if (!bt || !ms || !launches) return fail("null argument");
  for (int k = 0; k < HPS_PROFILE_FAMILIES; ++k) { ms[k] = bt->prof_ms[k]; launches[k] = bt->prof_n[k]; }
  return 0;
}

extern "C" int hps_batch_set_autotune(hps_batch* bt, int on)
{
  if (!bt) return fail("null batch");
  bt->autotune = on != 0;
  bt->field_shape = UINT32_MAX;
  bt->field_key[0] = bt->field_key[1] = bt->field_key[2] = 0;
  return 0;
}

extern "C" int hps_batch_get_tuning(hps_batch* bt, uint32_t* shape, float* ms, uint32_t n)
{
  if (!bt || !shape) return fail("null argument");
  *shape = bt->field_shape;
  for (uint32_t k = 0; ms && k < n && k < bt->field_times.size(); ++k) ms[k] = bt->field_times[k];
  return 0;
}

extern "C" int hps_batch_set_capture(hps_batch* bt, hps_nn* net)
{
  if (!bt) return fail("null batch");
  bt->capture = net;
  return 0;
}

extern "C" int hps_batch_set_surrogate(hps_batch* bt, hps_nn* net, uint32_t audit_every,
                                       double max_err)
{
  if (!bt) return fail("null batch");
  if (net && !(max_err > 0.0)) return fail("the surrogate needs a positive error bound");
  bt->sur = net;
  if (!net) return 0;
  if (!bt->d_sur) HPS_CHECK(hipMalloc(&bt->d_sur, sizeof(SurCtl)));
  SurCtl c{};
  c.audit_every = audit_every;
  c.err_bound = (float)max_err;
  bt->sur_cfg = c;
  HPS_CHECK(hipMemcpy(bt->d_sur, &c, sizeof c, hipMemcpyHostToDevice));
  if (bt->mm.fallback && !bt->sims.empty())
    HPS_CHECK(hipMemset(bt->mm.fallback, 0, bt->gemm_key[2] * sizeof(uint32_t)));
  return 0;
}

extern "C" int hps_batch_get_surrogate_info(hps_batch* bt, double* info)
{
  if (!bt || !info) return fail("null argument");
  SurCtl c{};
  if (bt->d_sur) {
    if (!bt->sims.empty()) HPS_CHECK(hipStreamSynchronize(bt->sims[0]->b->stream));
    HPS_CHECK(hipMemcpy(&c, bt->d_sur, sizeof c, hipMemcpyDeviceToHost));
  }
  info[0] = c.kick;
  info[1] = c.audited;
  info[2] = c.audit_fail;
  info[3] = c.max_err;
  info[4] = c.nn_member_kicks;
  info[5] = c.exact_member_kicks;
  return 0;
}

extern "C" int hps_batch_set_graph(hps_batch* bt, int on)
{
  if (!bt) return fail("null batch");
  bt->use_graph = on != 0;
  return 0;
}

extern "C" int hps_batch_graph_info(hps_batch* bt, uint64_t* replays, uint32_t* period_launches)
{
  if (!bt || !replays || !period_launches) return fail("null argument");
  *replays = bt->period_replays;
  *period_launches = bt->period_launches;
  return 0;
}

extern "C" int hps_batch_set_field_mode(hps_batch* bt, int mode, double aspect_step)
{
  if (!bt) return fail("null batch");
  if (mode != 0 && mode != 1) return fail("field mode must be 0 (FP64 per beam) or 1 (grouped FP32 GEMM)");
  if (mode == 1 && !(aspect_step > 0.0 && aspect_step <= 0.5))
    return fail("the aspect grid step must be in (0, 0.5]");
  bt->field_mode = mode;
  if (mode == 1) {
    bt->aspect_step = aspect_step;
    bt->aspect_lnr = hps::log(1.0 + aspect_step);
  } else {
    bt->aspect_lnr = 0.0;
  }
  return 0;
}

extern "C" int hps_batch_set_gemm_shape(hps_batch* bt, uint32_t shape)
{
  if (!bt) return fail("null batch");
  if (shape != UINT32_MAX && shape >= gemm_shapes()) return fail("no such GEMM shape");
  bt->gemm_shape = shape;
  bt->gemm_shape_key[0] = bt->gemm_shape_key[1] = bt->gemm_shape_key[2] = 0;
  if (shape != UINT32_MAX) bt->gemm_shape_key[2] = UINT32_MAX;   // fixed for every batch shape
  return 0;
}

extern "C" int hps_batch_get_gemm_info(hps_batch* bt, uint64_t* info, float* ms, uint32_t n)
{
  if (!bt || !info) return fail("null argument");
  GemmState g{};
  if (bt->d_gs && !bt->sims.empty()) {
    HPS_CHECK(hipStreamSynchronize(bt->sims[0]->b->stream));
    HPS_CHECK(hipMemcpy(&g, bt->d_gs, sizeof g, hipMemcpyDeviceToHost));
  }
  info[0] = bt->gemm_shape;
  info[1] = g.builds_total;
  info[2] = g.conflicts;
  info[3] = g.nrem;
  info[4] = g.exact_total;
  for (uint32_t k = 0; ms && k < n && k < bt->gemm_times.size(); ++k) ms[k] = bt->gemm_times[k];
  return 0;
}

extern "C" int hps_batch_last_time_ms(hps_batch* bt, double* out)
{
  if (!bt || !out) return fail("null argument");
  *out = bt->last_ms;
  return 0;
}

// simulate(start, end) of every simulation of the batch, with results
// identical to separate hps_sim_simulate calls.  Every simulation is compiled
// on its own; the programs must have the same structure (the same maps in
// the same segments and kick points: the same lattice layout, element
// values may differ), the same SCHEFF mesh and no mesh-size cutoff.  Each
// launch then covers all simulations: the segment and phase kernels over
// the concatenated chunks, one workgroup per simulation for the
// cross-chunk steps, the rebuild and field grids side by side.
extern "C" int hps_batch_simulate(hps_batch* bt, int32_t start, int32_t end)
{
  if (!bt) return fail("null batch");
  const size_t B = bt->sims.size();
  if (B == 0) return 0;
  hps_sim* m0 = bt->sims[0];
  const uint32_t nel = (uint32_t)m0->bl->el.size();
  if (nel == 0) return 0;
  const int32_t start_index = start < 0 ? 0 : start;
  const int32_t end_index = end < 0 ? (int32_t)nel - 1 : end;
  if (end_index >= (int32_t)nel || start_index >= (int32_t)nel) return fail("element index out of range");
  const bool sc_on = m0->sc_on && m0->sc;
  for (hps_sim* m : bt->sims) {
    if (m->bl->el.size() != nel) return fail("the beamlines of a batch must have the same elements");
    if ((m->sc_on && m->sc) != sc_on) return fail("space charge must be on for all or none of a batch");
    if (m->b->stream != m0->b->stream) return fail("the simulators of a batch must share one device");
    if (sc_on) {
      const hps_spch_state& cf = m->sc->cfg;
      if (cf.mesh_w != 0.0) return fail("mesh_size_cutoff_w is not supported in a batch");
      if (cf.nx != m0->sc->cfg.nx || cf.nz != m0->sc->cfg.nz)
        return fail("the SCHEFF meshes of a batch must have the same size");
      if ((size_t)cf.nx * cf.nz * 8 + kBlockLdsReserve > m->sc->max_lds)
        return fail("the SCHEFF mesh has more bins than one workgroup's LDS histogram holds");
      if (bt->field_mode == 1) {
        if (cf.adj_bunch != 0)
          return fail("the grouped field solve needs adj_bunch = 0 (no image rings)");
        if (cf.nz % 64 != 0)
          return fail("the grouped field solve needs nz to be a multiple of 64");
      }
    }
  }
  if (sc_on && bt->field_mode == 1 && B > kGemmMaxMembers)
    return fail("the grouped field solve takes at most 4096 simulations");
  if (sc_on && (bt->sur || bt->capture)) {
    if (bt->field_mode != 1)
      return fail("the surrogate and the training capture need the grouped field solve");
    for (hps_nn* h : {bt->sur, bt->capture}) {
      if (!h) continue;
      const NetView v = net_view(net_of(h));
      if ((uint32_t)v.nr != m0->sc->cfg.nx || (uint32_t)v.nz != m0->sc->cfg.nz)
        return fail("the surrogate network was made for another mesh size");
    }
  }
  // compile every simulation (restoring the simulators on a structure error)
  struct Saved { int32_t prev_end; uint32_t ccl; double dw, wl; };
  std::vector<Saved> saved(B);
  for (size_t k = 0; k < B; ++k) {
    hps_sim* m = bt->sims[k];
    saved[k] = {m->prev_end, m->ccl_cell_num, m->design_w, m->wavelen};
  }
  auto restore = [&]() {
    for (size_t k = 0; k < B; ++k) {
      hps_sim* m = bt->sims[k];
      m->prev_end = saved[k].prev_end; m->ccl_cell_num = saved[k].ccl;
      m->design_w = saved[k].dw; m->wavelen = saved[k].wl;
    }
  };
  std::vector<Compiler> cs(B);
  for (size_t k = 0; k < B; ++k) {
    hps_sim* m = bt->sims[k];
    hps_beam* b = m->b;
    if (start_index < m->prev_end) hps_sim_reset(m);
    m->prev_end = end_index;
    Compiler& c = cs[k];
    c.m = m;
    c.frac = m->sc ? m->sc->cfg.fraction : 1.0;
    c.freq = b->freq;
    c.quad_gap_run = b->blck_size / 2 != 0;
    c.dipole_run = b->blck_size / 4 != 0;
    c.sc_stats = b->n > 2;
    for (uint32_t i = 0; (int32_t)i <= end_index; ++i)
      if ((int32_t)i >= start_index || m->bl->el[i].kind == HPS_SPCH_COMP)
        if (compile_element(c, i)) { restore(); return -1; }
    c.close(STAT_NONE, false, 0);
  }
  const Compiler& c0 = cs[0];
  for (size_t k = 1; k < B; ++k) {
    const Compiler& c = cs[k];
    bool same = c.steps.size() == c0.steps.size() && c.ops.size() == c0.ops.size() &&
                c.sc_stats == c0.sc_stats && c.quad_gap_run == c0.quad_gap_run &&
                c.dipole_run == c0.dipole_run;
    for (size_t i = 0; same && i < c.steps.size(); ++i) {
      const Step &a = c.steps[i], &o = c0.steps[i];
      same = a.kind == o.kind && a.op_begin == o.op_begin && a.op_end == o.op_end &&
             a.kick == o.kick && a.stats == o.stats && a.kick_point == o.kick_point;
    }
    for (size_t i = 0; same && i < c.ops.size(); ++i) same = c.ops[i].kind == c0.ops[i].kind;
    if (!same) {
      restore();
      return fail("the simulations of a batch compile to different programs (lattice layout, "
                  "element kinds or beam sizes differ)");
    }
  }
  hipStream_t s = m0->b->stream;
  for (size_t k = 0; k < B; ++k) {
    hps_sim* m = bt->sims[k];
    const Compiler& c = cs[k];
    if (c.ops.size() > m->ops_cap) {
      HPS_CHECK(hipStreamSynchronize(s));
      (void)hipFree(m->d_ops);
      m->d_ops = nullptr;
      m->ops_cap = 0;
      HPS_CHECK(hipMalloc(&m->d_ops, (c.ops.size() + 64) * sizeof(Op)));
      m->ops_cap = c.ops.size() + 64;
    }
    if (!c.ops.empty())
      HPS_CHECK(hipMemcpyAsync(m->d_ops, c.ops.data(), c.ops.size() * sizeof(Op),
                               hipMemcpyHostToDevice, s));
  }

  // record layout: chunk map, rebuild map, rebuild and field records, then
  // per launch the records of its step
  const uint32_t nr = sc_on ? m0->sc->cfg.nx : 0, nz = sc_on ? m0->sc->cfg.nz : 0;
  if (bt->field_key[0] != nr || bt->field_key[1] != nz || bt->field_key[2] != (uint32_t)B)
    bt->field_shape = bt->autotune ? UINT32_MAX : field_batch_default_shape(nr, (uint32_t)B);
  std::vector<uint32_t> build_map;
  std::vector<uint32_t> nchs(B);
  for (size_t k = 0; k < B; ++k) nchs[k] = bt->sims[k]->b->pr.nch;
  const std::vector<uint32_t> chunk_map = xcd_chunk_map(nchs);
  std::vector<BuildArgs> build(B);
  std::vector<FieldArgs> field(B);
  if (sc_on) {
    const uint32_t gb = greens_blocks(nr, nz);
    for (size_t k = 0; k < B; ++k) {
      hps_sim* m = bt->sims[k];
      BuildArgs& a = build[k];
      a.sc = m->sc->dev; a.fld1 = m->sc->fld1; a.b = m->b->d; a.st = m->b->st; a.bins = m->sc->bins;
      // the grouped field solve has no per-beam Green's table
      a.greens_blocks = bt->field_mode == 1 ? 0 : gb;
      // deposit-only rebuilds (grouped solve) take fuller deposit workgroups
      a.dep_blocks = m->b->blck_size / 4 == 0 ? 0
                     : bt->field_mode == 1 ? deposit_blocks_batch(m->b->n, nr, nz)
                                           : deposit_blocks(m->b->n, 4);
      const uint32_t nb = sc_build_blocks(a);
      for (uint32_t c = 0; c < nb; ++c)
        build_map.push_back(((uint32_t)k << kBatchLocalBits) | c);
      FieldArgs& f = field[k];
      f.sc = m->sc->dev; f.bins = m->sc->bins; f.fld1 = m->sc->fld1; f.fld2 = m->sc->fld2;
      f.part = m->sc->part;
    }
  }
  const bool gemm = sc_on && bt->field_mode == 1;
  if (gemm) {
    if (alloc_gemm(bt, nr, nz, (uint32_t)B, s)) { restore(); return -1; }
    if (!bt->cus) {
      int dev = 0, n = 0;
      HPS_CHECK(hipGetDevice(&dev));
      HPS_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
      bt->cus = (uint32_t)n;
    }
    std::vector<GemmMember> mem(B);
    for (size_t k = 0; k < B; ++k) {
      hps_sim* m = bt->sims[k];
      mem[k].sc = m->sc->dev; mem[k].bins = m->sc->bins; mem[k].fld2 = m->sc->fld2;
      mem[k].n = m->b->n; mem[k].mass = m->b->mass;
    }
    HPS_CHECK(hipMemcpyAsync(bt->d_mem, mem.data(), B * sizeof(GemmMember),
                             hipMemcpyHostToDevice, s));
    if (bt->gemm_shape_key[2] != UINT32_MAX &&
        (bt->gemm_shape_key[0] != nr || bt->gemm_shape_key[1] != nz ||
         bt->gemm_shape_key[2] != (uint32_t)B))
      bt->gemm_shape = UINT32_MAX;
    if (bt->sur) {
      NetImpl* net = net_of(bt->sur);
      if (net_reserve_infer(net, (int)B)) { restore(); return -1; }
      const NetView v = net_view(net);
      SurCtl& c = bt->sur_cfg;
      c.out_scale = v.out_scale; c.u_mu = v.u_mu; c.inv_u_sigma = 1.0f / v.u_sigma;
      c.lnr = bt->aspect_lnr;
      HPS_CHECK(hipMemcpyAsync(bt->d_sur, &c, kSurCtlConfigBytes, hipMemcpyHostToDevice, s));
    }
  }
  NetData cap_data{};
  if (gemm && bt->capture && net_data(net_of(bt->capture), &cap_data)) { restore(); return -1; }
  if (gemm && bt->capture && !cap_data.x) { restore(); return fail("the capture dataset has no capacity (hps_nn_data_reserve)"); }
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t off_chunk = off; off = align(off + chunk_map.size() * 4);
  const size_t off_bmap = off; off = align(off + build_map.size() * 4);
  const size_t off_build = off; off = align(off + B * sizeof(BuildArgs));
  const size_t off_field = off; off = align(off + B * sizeof(FieldArgs));
  size_t nseg = 0, nkp = 0;
  for (const Step& st : c0.steps) {
    if (st.kind != 0) continue;
    if (st.op_end > st.op_begin || st.kick || st.stats != STAT_NONE) ++nseg;
    if (st.kick_point) ++nkp;
  }
  const size_t off_steps = off;
  off = align(off + nseg * B * sizeof(SegArgs) + nkp * B * sizeof(PhaseArgs) + 256);
  if (off > bt->rec_cap) {
    HPS_CHECK(hipStreamSynchronize(s));
    (void)hipFree(bt->d_rec);
    bt->d_rec = nullptr;
    bt->rec_cap = 0;
    HPS_CHECK(hipMalloc(&bt->d_rec, off + off / 4));
    bt->rec_cap = off + off / 4;
  }
  std::vector<uint8_t>& H = bt->host_rec;
  H.assign(off, 0);
  std::memcpy(H.data() + off_chunk, chunk_map.data(), chunk_map.size() * 4);
  std::memcpy(H.data() + off_bmap, build_map.data(), build_map.size() * 4);
  std::memcpy(H.data() + off_build, build.data(), B * sizeof(BuildArgs));
  std::memcpy(H.data() + off_field, field.data(), B * sizeof(FieldArgs));
  const uint32_t* d_chunk_map = (const uint32_t*)(bt->d_rec + off_chunk);
  const uint32_t* d_build_map = (const uint32_t*)(bt->d_rec + off_bmap);
  const BuildArgs* d_build = (const BuildArgs*)(bt->d_rec + off_build);
  const FieldArgs* d_field = (const FieldArgs*)(bt->d_rec + off_field);
  const uint32_t total_chunks = (uint32_t)chunk_map.size();
  const uint32_t build_blocks = (uint32_t)build_map.size();

  enum { L_SEG = 0, L_PHASE, L_BUILD, L_FIELD, L_SEG_FINAL, L_PHASE_FINAL };
  struct BLaunch {
    int kind;
    size_t rec;        // byte offset of the records
    int op_class;
  };
  std::vector<BLaunch> launches;
  size_t cursor = off_steps;
  size_t uploaded = 0;         // bytes of H already on the device
  std::vector<std::pair<int, size_t>> marks;
  size_t nev = 0;
  if (!bt->d_ctl) {
    HPS_CHECK(hipMalloc(&bt->d_ctl, 2 * sizeof(uint64_t)));
    HPS_CHECK(hipMemset(bt->d_ctl, 0, 2 * sizeof(uint64_t)));
  }
  uint64_t* const d_roff = bt->d_ctl;           // record offset of a replayed period
  const uint64_t* const d_zero = bt->d_ctl + 1;  // eager launches
  // one launch of the program (roff: where the batched kernels add the
  // replayed period's record offset)
  auto issue = [&](const BLaunch& l, const uint64_t* roff) -> int {
    const void* rec = bt->d_rec + l.rec;
    switch (l.kind) {
      case L_SEG:
        if (gemm)
          HPS_CHECK(fast::launch_segment_batch((const SegArgs*)rec, d_chunk_map, total_chunks,
                                               l.op_class, roff, s));
        else
          HPS_CHECK(launch_segment_batch((const SegArgs*)rec, d_chunk_map, total_chunks,
                                         l.op_class, roff, s));
        break;
      case L_SEG_FINAL:
        HPS_CHECK(launch_seg_final_batch((const SegArgs*)rec, (uint32_t)B, roff, s));
        break;
      case L_PHASE:
        HPS_CHECK(launch_phase_batch((const PhaseArgs*)rec, d_chunk_map, total_chunks, roff, s));
        break;
      case L_PHASE_FINAL:
        HPS_CHECK(launch_phase_final_batch((const PhaseArgs*)rec, (uint32_t)B, roff, s));
        break;
      case L_BUILD:
        if (gemm)
          HPS_CHECK(fast::launch_sc_build_batch(d_build, d_build_map, build_blocks, nr, nz, s));
        else
          HPS_CHECK(launch_sc_build_batch(d_build, d_build_map, build_blocks, nr, nz, s));
        break;
      default: {
        if (gemm) {
          SurCtl* ctl = bt->sur ? bt->d_sur : nullptr;
          HPS_CHECK(launch_gemm_prepare(bt->d_mem, (uint32_t)B, bt->d_gs, bt->d_perm, bt->mm, ctl,
                                        bt->d_tables, bt->d_q, nr, nz, bt->aspect_lnr, bt->cus,
                                        s));
          if (bt->gemm_shape == UINT32_MAX && tune_gemm(bt, nr, nz, (uint32_t)B, s)) return -1;
          if (gemm_shape(bt->gemm_shape).spectral && !bt->fft.tw)
            return fail("the spectral product needs an even nr and a power-of-two nz");
          if (bt->profile) {   // the product alone (family 6), inside the field family
            while (bt->evpool.size() < nev + 2) {
              hipEvent_t e;
              HPS_CHECK(hipEventCreate(&e));
              bt->evpool.push_back(e);
            }
            marks.push_back({6, nev});
            HPS_CHECK(hipEventRecord(bt->evpool[nev], s));
          }
          HPS_CHECK(launch_gemm_product(bt->d_gs, bt->d_tables, bt->d_q, bt->d_w, nr, nz,
                                        bt->gemm_shape, bt->cus,
                                        bt->fft.tw ? &bt->fft : nullptr, s));
          if (bt->profile) {
            HPS_CHECK(hipEventRecord(bt->evpool[nev + 1], s));
            nev += 2;
          }
          float* y = nullptr;
          if (ctl) {
            // the network serves every remeshing member; the audit compares it
            // with the exact product where both ran
            HPS_CHECK(launch_sur_input(bt->d_mem, (uint32_t)B, bt->mm, ctl, bt->d_xinf,
                                       bt->d_uinf, nr, nz, s));
            if (net_infer(net_of(bt->sur), bt->d_xinf, bt->d_uinf, &ctl->nn_n, (int)B, &y, s))
              return -1;
            HPS_CHECK(launch_sur_audit(bt->d_mem, (uint32_t)B, bt->d_gs, bt->mm, bt->d_w, y, ctl,
                                       nr, nz, bt->gemm_shape, s));
          }
          if (bt->capture)
            HPS_CHECK(launch_capture(bt->d_mem, (uint32_t)B, bt->d_gs, bt->mm, bt->d_w, cap_data,
                                     bt->aspect_lnr, nr, nz, bt->gemm_shape, s));
          HPS_CHECK(launch_gemm_store(bt->d_mem, (uint32_t)B, bt->d_gs, bt->mm, bt->d_w, y, ctl,
                                      nr, nz, bt->aspect_lnr, bt->gemm_shape, s));
          break;
        }
        if (bt->autotune && bt->field_shape == UINT32_MAX) {
          // first rebuild of this batch shape: time every field launch shape
          // on the live tables (each gives the same bits) and keep the fastest
          if (tune_field(bt, d_field, (uint32_t)B, nr, nz, m0->sc->max_lds, s)) return -1;
        }
        HPS_CHECK(launch_field_batch(d_field, (uint32_t)B, nr, nz, m0->sc->max_lds,
                                     bt->field_shape, s));
        break;
      }
    }
    return 0;
  };
  auto upload = [&]() -> int {
    if (cursor > uploaded) {
      HPS_CHECK(hipMemcpyAsync(bt->d_rec + uploaded, H.data() + uploaded, cursor - uploaded,
                               hipMemcpyHostToDevice, s));
      uploaded = cursor;
    }
    return 0;
  };
  auto eager = [&](size_t from, size_t to) -> int {
    for (size_t j = from; j < to; ++j) {
      const BLaunch& l = launches[j];
      size_t ev = 0;   // this launch's event pair (events inside it come after)
      if (bt->profile) {
        while (bt->evpool.size() < nev + 2) {
          hipEvent_t e;
          HPS_CHECK(hipEventCreate(&e));
          bt->evpool.push_back(e);
        }
        ev = nev;
        nev += 2;
        marks.push_back({l.kind, ev});
        HPS_CHECK(hipEventRecord(bt->evpool[ev], s));
      }
      if (issue(l, d_zero)) return -1;
      if (bt->profile) HPS_CHECK(hipEventRecord(bt->evpool[ev + 1], s));
    }
    return 0;
  };
  auto flush = [&]() -> int {
    if (upload() || eager(0, launches.size())) return -1;
    launches.clear();
    return 0;
  };
  bool host_steps = false;

  HPS_CHECK(hipEventRecord(bt->ev0, s));
  // the records of the host beam state (current, frequency) are built after
  // every host step that changes it
  for (size_t i = 0; i < c0.steps.size(); ++i) {
    const Step& st0 = c0.steps[i];
    if (st0.kind == 1) {
      if (flush()) return -1;
      host_steps = true;
      for (size_t k = 0; k < B; ++k)
        if (change_frequency(bt->sims[k]->b, cs[k].steps[i].new_freq)) return -1;
      continue;
    }
    if (st0.op_end > st0.op_begin || st0.kick || st0.stats != STAT_NONE) {
      SegArgs* recs = (SegArgs*)(H.data() + cursor);
      for (size_t k = 0; k < B; ++k) {
        hps_sim* m = bt->sims[k];
        hps_beam* b = m->b;
        const Step& st = cs[k].steps[i];
        const uint32_t kstride = b->grid_size * (b->blck_size / 4);
        SegArgs& a = recs[k];
        a.b = b->d;
        a.st = b->st;
        a.pr = b->pr;
        a.ops = m->d_ops;
        a.op_begin = st.op_begin;
        a.op_end = st.op_end;
        a.kick = (st.kick && kstride != 0 && m->sc) ? 1 : 0;
        a.kick_mask = kstride ? kstride - 1 : 0;
        a.kick_len = st.kick_len;
        a.sc = m->sc ? m->sc->dev : nullptr;
        a.fld2 = m->sc ? m->sc->fld2 : nullptr;
        a.stats = st.stats;
        a.rmax = (st.stats == STAT_SC && m->sc && m->sc->cfg.adj_bunch != 0) ? 1 : 0;
        a.op_class = OPC_LIGHT;
        for (int32_t o = st.op_begin; o < st.op_end; ++o)
          a.op_class = std::max(a.op_class, op_class_of(cs[k].ops[o].kind));
      }
      launches.push_back({L_SEG, cursor, recs[0].op_class});
      if (st0.stats != STAT_NONE) launches.push_back({L_SEG_FINAL, cursor, 0});
      cursor += B * sizeof(SegArgs);
    }
    if (st0.kick_point) {
      PhaseArgs* recs = (PhaseArgs*)(H.data() + cursor);
      for (size_t k = 0; k < B; ++k) {
        hps_sim* m = bt->sims[k];
        hps_beam* b = m->b;
        PhaseArgs& p = recs[k];
        p.b = b->d; p.st = b->st; p.pr = b->pr; p.sc = m->sc->dev; p.bins = m->sc->bins;
        p.do_sums = b->n > 2 ? 1 : 0;
        p.rmax = m->sc->cfg.adj_bunch == 0 ? 0 : (cs[k].sc_stats ? 1 : 2);
        p.frac = cs[k].steps[i].frac; p.current = b->current; p.freq = b->freq;
      }
      if (c0.sc_stats) launches.push_back({L_PHASE, cursor, 0});
      launches.push_back({L_PHASE_FINAL, cursor, 0});
      launches.push_back({L_BUILD, 0, 0});
      launches.push_back({L_FIELD, 0, 0});
      cursor += B * sizeof(PhaseArgs);
    }
  }
  // Lattice periods as one replayed HIP graph: a run of n repetitions of the
  // same p launches whose records advance by a fixed stride is captured once
  // (the batched kernels read their records at the base plus a device offset
  // that the graph's last node advances); the launches before and after the
  // run go eagerly.  Needs every launch shape chosen (tuning is eager) and no
  // host step (beam frequency change) inside the program.
  const bool tuned = !sc_on || (gemm ? bt->gemm_shape != UINT32_MAX
                                     : (!bt->autotune || bt->field_shape != UINT32_MAX));
  PeriodRun run{};
  if (bt->use_graph && !bt->profile && !host_steps && tuned) {
    std::vector<PeriodKey> keys(launches.size());
    for (size_t j = 0; j < launches.size(); ++j) {
      const BLaunch& l = launches[j];
      const bool carries = l.kind != L_BUILD && l.kind != L_FIELD;
      keys[j] = {l.kind, l.op_class, carries, l.rec};
    }
    run = find_period_run(keys);
  }
  if (run.n >= 2) {
    if (upload() || eager(0, run.a)) return -1;
    std::vector<uint8_t> sig;
    const auto put = [&](const void* p, size_t nb) {
      const uint8_t* q = (const uint8_t*)p;
      sig.insert(sig.end(), q, q + nb);
    };
    put(&run, sizeof run);
    for (size_t j = run.a; j < run.a + run.p; ++j) put(&launches[j], sizeof(BLaunch));
    const void* ptrs[] = {bt->d_rec, d_chunk_map, d_build, d_build_map, d_field, bt->d_mem,
                          bt->d_gs, bt->d_perm, bt->d_tables, bt->d_q, bt->d_w, bt->d_ctl,
                          bt->sur, bt->capture, bt->d_sur, bt->d_xinf, cap_data.x,
                          bt->mm.pos_exact, bt->fft.shat, bt->fft.qhat, bt->fft.what};
    put(ptrs, sizeof ptrs);
    const uint32_t dims[] = {(uint32_t)B, nr, nz, total_chunks, build_blocks, bt->field_shape,
                             bt->gemm_shape, bt->cus, gemm ? 1u : 0u};
    put(dims, sizeof dims);
    put(&bt->aspect_lnr, sizeof bt->aspect_lnr);
    const size_t max_lds = sc_on ? m0->sc->max_lds : 0;
    put(&max_lds, sizeof max_lds);
    if (!bt->pgraph_exec || sig != bt->pgraph_sig) {
      if (bt->pgraph_exec) { (void)hipGraphExecDestroy(bt->pgraph_exec); bt->pgraph_exec = nullptr; }
      HPS_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
      int rc = 0;
      for (size_t j = run.a; j < run.a + run.p && rc == 0; ++j) rc = issue(launches[j], d_roff);
      hipError_t e1 = rc == 0 ? launch_roff(d_roff, run.stride, true, s) : hipSuccess;
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(s, &g);
      if (rc != 0) { if (g) (void)hipGraphDestroy(g); return -1; }
      if (e1 != hipSuccess) { if (g) (void)hipGraphDestroy(g); HPS_CHECK(e1); }
      HPS_CHECK(e2);
      hipError_t e3 = hipGraphInstantiate(&bt->pgraph_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HPS_CHECK(e3);
      bt->pgraph_sig.swap(sig);
    }
    HPS_CHECK(launch_roff(d_roff, 0, false, s));
    for (uint32_t r = 0; r < run.n; ++r) HPS_CHECK(hipGraphLaunch(bt->pgraph_exec, s));
    bt->period_replays += run.n;
    bt->period_launches = (uint32_t)run.p;
    if (eager(run.a + (size_t)run.n * run.p, launches.size())) return -1;
    launches.clear();
  } else if (flush()) {
    return -1;
  }
  HPS_CHECK(hipEventRecord(bt->ev1, s));
  HPS_CHECK(hipEventSynchronize(bt->ev1));
  float ms = 0;
  HPS_CHECK(hipEventElapsedTime(&ms, bt->ev0, bt->ev1));
  bt->last_ms = ms;
  for (hps_sim* m : bt->sims) m->last_ms = ms;
  for (const auto& mk : marks) {
    float t = 0;
    HPS_CHECK(hipEventElapsedTime(&t, bt->evpool[mk.second], bt->evpool[mk.second + 1]));
    bt->prof_ms[mk.first] += t;
    bt->prof_n[mk.first] += 1;
  }
  HPS_CHECK(hipGetLastError());
  return 0;
}

extern "C" int hps_probe_math(int fn, const double* x, const double* y, double* out, uint32_t n)
{
  if (!x || !out) return fail("null argument");
  if (n == 0) return 0;
  int dev = 0;
  HPS_CHECK(hipGetDevice(&dev));
  hipStream_t s;
  if (stream_for(dev, &s)) return -1;
  double *dx = nullptr, *dy = nullptr, *dout = nullptr;
  const size_t nb = (size_t)n * 8;
  HPS_CHECK(hipMalloc(&dx, nb));
  HPS_CHECK(hipMalloc(&dy, nb));
  HPS_CHECK(hipMalloc(&dout, nb));
  HPS_CHECK(hipMemcpy(dx, x, nb, hipMemcpyHostToDevice));
  if (y) HPS_CHECK(hipMemcpy(dy, y, nb, hipMemcpyHostToDevice));
  else HPS_CHECK(hipMemset(dy, 0, nb));
  if (fn >= 32)
    HPS_CHECK(fast::launch_probe(fn - 32, dx, dy, dout, n, s));
  else
    HPS_CHECK(launch_probe(fn, dx, dy, dout, n, s));
  HPS_CHECK(hipStreamSynchronize(s));
  HPS_CHECK(hipMemcpy(out, dout, nb, hipMemcpyDeviceToHost));
  (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dout);
  return 0;
}

// ---- Emittances over caller-owned device arrays, allocating nothing ----

namespace {

// The per-chunk partials of the canonical sums live in this array instead of
// a Beam's allocation: 8 quantities and 3 counts for every chunk of the
// largest beam, then the final sums.  Every device holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host
// stand-in there).
constexpr uint32_t kNoAllocMaxParticles = 4194304;
constexpr uint32_t kNoAllocChunks = kNoAllocMaxParticles / kChunk;             // 4096
constexpr size_t kNoAllocSumsOff = 0;
constexpr size_t kNoAllocCntOff = kNoAllocSumsOff + (size_t)kMaxQ * kNoAllocChunks * 8;
constexpr size_t kNoAllocOutOff = kNoAllocCntOff + 3 * (size_t)kNoAllocChunks * 4;
constexpr size_t kNoAllocOutCntOff = kNoAllocOutOff + kMaxQ * 8;
constexpr size_t kNoAllocBytes = kNoAllocOutCntOff + 256;                       // 311,616 B

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[16] = {};

// The device address of `scratch` on device dev (looked up once per device,
// with that device current).
int scratch_for(int dev, unsigned char** out)
{
  if (dev < 0 || dev >= 16) return fail("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HPS_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int hps_emittance_noalloc(const double* x, const double* xp, const double* y,
                                     const double* yp, const double* phi, const double* w,
                                     const uint32_t* loss, const uint32_t* lloss, uint32_t n,
                                     double mass, double* out)
{
  if (!x || !xp || !y || !yp || !phi || !w || !loss || !lloss || !out)
    return fail("null argument");
  if (n > kNoAllocMaxParticles)
    return fail("hps_emittance_noalloc: " + std::to_string(n) + " particles exceed its limit of " +
                std::to_string(kNoAllocMaxParticles));
  if (n <= 2) {   // Beam::UpdateEmittance computes nothing for such beams (beam.cpp:985-994)
    out[0] = out[1] = out[2] = 0.0;
    return 0;
  }
  int dev = 0;
  HPS_CHECK(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  if (scratch_for(dev, &base)) return -1;
  hipStream_t s = nullptr;   // the device's null stream: nothing is created
  Partials pr{};
  pr.sums = (double*)(base + kNoAllocSumsOff);
  pr.cnt = (uint32_t*)(base + kNoAllocCntOff);
  pr.out = (double*)(base + kNoAllocOutOff);
  pr.out_cnt = (uint32_t*)(base + kNoAllocOutCntOff);
  pr.nch = (n + kChunk - 1) / kChunk;
  pr.mpad = next_pow2(pr.nch);
  // the trees read the zero padding past the last chunk (left over by an
  // earlier, larger call otherwise)
  if (pr.mpad > pr.nch) {
    for (int q = 0; q < kMaxQ; ++q)
      HPS_CHECK(hipMemsetAsync(pr.sums + (size_t)q * pr.mpad + pr.nch, 0,
                               (size_t)(pr.mpad - pr.nch) * 8, s));
    for (int q = 0; q < 3; ++q)
      HPS_CHECK(hipMemsetAsync(pr.cnt + (size_t)q * pr.mpad + pr.nch, 0,
                               (size_t)(pr.mpad - pr.nch) * 4, s));
  }
  BeamArrays a{};
  a.x = const_cast<double*>(x); a.xp = const_cast<double*>(xp);
  a.y = const_cast<double*>(y); a.yp = const_cast<double*>(yp);
  a.phi = const_cast<double*>(phi); a.w = const_cast<double*>(w);
  a.loss = const_cast<uint32_t*>(loss); a.lloss = const_cast<uint32_t*>(lloss);
  a.n = n;
  a.mass = mass;
  double w0 = 0;
  HPS_CHECK(hipMemcpyAsync(&w0, w, sizeof(double), hipMemcpyDeviceToHost, s));
  // the same sums and expressions as hps_beam_get_emittance (UpdateLoss, then
  // Beam::UpdateEmittance, beam.cpp:985-994); num_loss is the loss count of
  // the same pass
  auto sums = [&](const double* pa, const double* pb, const double* rel, double* r5,
                  uint32_t* nloss) -> int {
    HPS_CHECK(launch_sums(a, pa, pb, 0, pr, s, rel));
    uint32_t c[3];
    HPS_CHECK(hipMemcpyAsync(r5, pr.out, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    HPS_CHECK(hipMemcpyAsync(c, pr.out_cnt, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HPS_CHECK(hipStreamSynchronize(s));
    *nloss = c[0];
    return 0;
  };
  for (int axis = 0; axis < 2; ++axis) {
    double r[5];
    uint32_t nloss = 0;
    if (sums(axis == 0 ? x : y, axis == 0 ? xp : yp, nullptr, r, &nloss)) return -1;
    const double sv = (double)(uint32_t)(n - nloss);
    const double m1 = r[0] / sv, m2 = r[1] / sv, p1 = r[2] / sv, p2 = r[3] / sv, mp = r[4] / sv;
    const double gamma1 = w0 / mass;
    const double betagamma = std::sqrt(gamma1 * (gamma1 + 2.0));
    out[axis] = std::sqrt((m2 - m1 * m1) * (p2 - p1 * p1) - (mp - m1 * p1) * (mp - m1 * p1)) *
                betagamma;
  }
  double r[5];
  uint32_t nloss = 0;
  if (sums(phi, w, phi, r, &nloss)) return -1;
  const double sv = (double)(uint32_t)(n - nloss);
  const double p1 = r[0] / sv, p2 = r[1] / sv, w1 = r[2] / sv, w2 = r[3] / sv, pw = r[4] / sv;
  const double cov = pw - p1 * w1;
  out[2] = std::sqrt((p2 - p1 * p1) * (w2 - w1 * w1) - cov * cov);
  return 0;
}
