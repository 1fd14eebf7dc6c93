/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial FP64 CPU restatement of the reference HPSim hot path, used as the
 * parity checker for the HIP product and as the timed CPU baseline.  Nothing
 * in the product links or calls this file.  Every function cites the
 * reference file:line it restates (paths relative to the reference root).
 *
 * The particle, chunk and table loops carry OpenMP pragmas: every iteration
 * writes only its own index (the deposit adds integers atomically), so the
 * OpenMP build liboracle_omp.so gives bit-identical results to the serial
 * liboracle.so, which ignores the pragmas.
 *
 * Deliberate, documented differences from the reference (applied identically
 * in the GPU product so that parity is bit-exact):
 *  - Reductions use one canonical order (ORC_CANON_*: each of 256 lanes sums
 *    4 particles strided by 256 inside a 1024-particle chunk, halving tree
 *    over the 256 lanes, halving tree over the zero-padded chunk partials)
 *    instead of the launch-shape-dependent tree of beam_kernel.cu:84-166.
 *  - The SCHEFF charge deposit is an integer fixed-point histogram (weights
 *    truncated to multiples of 2^-36, int64 sums) instead of float atomics
 *    (space_charge_kernel.cu:200-218): order independent, deterministic.
 *  - The field sum of UpdateFldTbl2Kernel (space_charge_kernel.cu:230-342)
 *    uses fused multiply-adds in a canonical order: with P = min(NextPow2(nr),
 *    64) partial sums, partial p takes the source rings rs = p, p+P, ... in
 *    ascending order and, for each, the source rows zs = 0 .. nz-1 in
 *    ascending order; then a halving tree over the P partials (partials
 *    with no ring stay +0.0).
 *  - Elementary functions come from orc_math.h (project-owned, see there).
 *  - rsqrt(x) is 1.0/sqrt(x); double->int conversions saturate (orc_to_int).
 *  - Lattice conventions for configurations the reference cannot run
 *    (simulation_engine.cu:200-206,224-227 exit or dereference null): a DTL
 *    gap whose neighbour is not a Quad uses quad length 0, and a DTL gap that
 *    is the last element is treated as the only gap in its tank.
 */
#include "oracle.h"
#include "orc_math.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CLIGHT 299.792458                       /* constant.h:7 */
#define PI 3.14159265358979323846              /* constant.h:8 */
#define TWOPI (3.14159265358979323846 * 2.0)   /* constant.h:9 */

static char g_err[512];
const char* orc_last_error(void) { return g_err; }
static int set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); return -1; }

uint32_t orc_t_nextpow2(uint32_t x)   /* utility.cpp:7-16 */
{
  --x;
  x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16;
  return ++x;
}

/* ------------------------------------------------------------------ */
/* Canonical reduction                                                 */
/* ------------------------------------------------------------------ */

/* v[i] already holds the filtered value (excluded particles: 0.0). */
static double canon_sum(const double* v, uint32_t n)
{
  uint32_t nch = (n + ORC_CANON_CHUNK - 1) / ORC_CANON_CHUNK;
  if (nch == 0) nch = 1;
  uint32_t m = orc_t_nextpow2(nch);
  double* part = (double*)calloc(m, sizeof(double));
  #pragma omp parallel for schedule(static)
  for (uint32_t c = 0; c < nch; ++c) {
    double lane[ORC_CANON_LANES];
    for (uint32_t t = 0; t < ORC_CANON_LANES; ++t) {
      double acc = 0.0;
      for (uint32_t j = 0; j < ORC_CANON_PER_LANE; ++j) {
        uint64_t i = (uint64_t)c * ORC_CANON_CHUNK + (uint64_t)j * ORC_CANON_LANES + t;
        acc = acc + (i < n ? v[i] : 0.0);
      }
      lane[t] = acc;
    }
    for (uint32_t o = ORC_CANON_LANES / 2; o >= 1; o >>= 1)
      for (uint32_t t = 0; t < o; ++t) lane[t] = lane[t] + lane[t + o];
    part[c] = lane[0];
  }
  for (uint32_t o = m / 2; o >= 1; o >>= 1)
    for (uint32_t c = 0; c < o; ++c) part[c] = part[c] + part[c + o];
  double r = part[0];
  free(part);
  return r;
}
double orc_t_canon_sum(const double* v, uint32_t n) { return canon_sum(v, n); }

/* Threads the loops of this build use (1 for the serial build). */
int orc_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Beam                                                                */
/* ------------------------------------------------------------------ */

struct orc_beam {
  uint32_t n;
  double mass, charge, current, freq, design_w;
  uint32_t grid_size, blck_size;
  double *x, *xp, *y, *yp, *phi, *phi_r, *w;
  uint32_t *loss, *lloss;
  double* tmp;   /* scratch for filtered values */
  double* tmp2;
  orc_beam_stats st;
  /* saved copies, beam.h:279-311 */
  int saved0, saved1;
  double *s0[6], *s1[6];
  uint32_t *s0l, *s0ll, *s1l, *s1ll;
  double dw0, f0, c0, dw1, f1, c1;
};

/* Beam::SetNumThreadsPerBlock(256), beam.cpp:148-168 */
static void set_launch_geometry(orc_beam* b)
{
  uint32_t num_thread = orc_t_nextpow2(b->n);
  if (num_thread > 256) { b->grid_size = num_thread / 256; b->blck_size = 256; }
  else { b->grid_size = 1; b->blck_size = num_thread; }
}

orc_beam* orc_beam_create(uint32_t num, double mass, double charge, double current)
{
  if (num == 0) { set_err("beam needs at least one particle"); return NULL; }
  orc_beam* b = (orc_beam*)calloc(1, sizeof(orc_beam));
  b->n = num; b->mass = mass; b->charge = charge; b->current = current;
  b->freq = 0.0; b->design_w = 0.0;
  double** arrs[7] = {&b->x, &b->xp, &b->y, &b->yp, &b->phi, &b->phi_r, &b->w};
  for (int k = 0; k < 7; ++k) *arrs[k] = (double*)calloc(num, sizeof(double));
  b->loss = (uint32_t*)calloc(num, sizeof(uint32_t));
  b->lloss = (uint32_t*)calloc(num, sizeof(uint32_t));
  b->tmp = (double*)calloc(num, sizeof(double));
  b->tmp2 = (double*)calloc(num, sizeof(double));
  set_launch_geometry(b);
  return b;
}

void orc_beam_destroy(orc_beam* b)
{
  if (!b) return;
  free(b->x); free(b->xp); free(b->y); free(b->yp); free(b->phi); free(b->phi_r); free(b->w);
  free(b->loss); free(b->lloss); free(b->tmp); free(b->tmp2);
  for (int k = 0; k < 6; ++k) { free(b->s0[k]); free(b->s1[k]); }
  free(b->s0l); free(b->s0ll); free(b->s1l); free(b->s1ll);
  free(b);
}

uint32_t orc_beam_size(const orc_beam* b) { return b->n; }
double orc_beam_get_frequency(const orc_beam* b) { return b->freq; }
double orc_beam_get_design_w(const orc_beam* b) { return b->design_w; }
void orc_beam_get_stats(const orc_beam* b, orc_beam_stats* s) { *s = b->st; }
void orc_beam_set_num_loss(orc_beam* b, uint32_t n) { b->st.num_loss = n; }

void orc_beam_get_arrays(const orc_beam* b, double* x, double* xp, double* y,
    double* yp, double* phi, double* phi_r, double* w, uint32_t* loss, uint32_t* lloss)
{
  size_t nb = (size_t)b->n * sizeof(double);
  if (x) memcpy(x, b->x, nb);
  if (xp) memcpy(xp, b->xp, nb);
  if (y) memcpy(y, b->y, nb);
  if (yp) memcpy(yp, b->yp, nb);
  if (phi) memcpy(phi, b->phi, nb);
  if (phi_r) memcpy(phi_r, b->phi_r, nb);
  if (w) memcpy(w, b->w, nb);
  if (loss) memcpy(loss, b->loss, (size_t)b->n * 4);
  if (lloss) memcpy(lloss, b->lloss, (size_t)b->n * 4);
}

/* filters of the reduce kernels, beam_kernel.cu:105-111 */
static inline int is_tgood(const orc_beam* b, uint32_t i) { return b->loss[i] == 0; }
static inline int is_good(const orc_beam* b, uint32_t i) { return (uint32_t)(b->loss[i] + b->lloss[i]) == 0; }

static double sum_filtered(orc_beam* b, const double* a, int good, int square)
{
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    int inc = good ? is_good(b, i) : is_tgood(b, i);
    double v = inc ? a[i] : 0.0;
    b->tmp[i] = square ? v * v : v;
  }
  return canon_sum(b->tmp, b->n);
}

static double sum_product(orc_beam* b, const double* a1, const double* a2)
{
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i)
    b->tmp[i] = is_tgood(b, i) ? a1[i] * a2[i] : 0.0;
  return canon_sum(b->tmp, b->n);
}

/* UpdateLossCountKernelCall, beam.cu:493-516 */
void orc_beam_update_loss(orc_beam* b)
{
  uint32_t c = 0;
  #pragma omp parallel for reduction(+:c) schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) c += (b->loss[i] != 0);
  b->st.num_loss = c;
}

static void update_lloss_count(orc_beam* b)
{
  uint32_t c = 0;
  #pragma omp parallel for reduction(+:c) schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) c += (b->lloss[i] != 0);
  b->st.num_lloss = c;
}

/* UpdateGoodParticleCountKernelCall, beam.cu:1233-1255 */
static void update_good_count(orc_beam* b)
{
  uint32_t c = 0;
  #pragma omp parallel for reduction(+:c) schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) c += is_good(b, i);
  b->st.num_good = c;
}

/* UpdateAvgOfOneVariableKernelCall + UpdateVariableAvg(WithLloss)Kernel,
 * beam.cu:411-436, beam_kernel.cu:1396-1410 */
static double avg_of(orc_beam* b, const double* a, int good)
{
  double s = sum_filtered(b, a, good, 0);
  if (!good) {
    double surv = (double)(uint32_t)(b->n - b->st.num_loss);
    return s / surv;
  }
  if (b->st.num_good > 0) return s / (double)b->st.num_good;
  return -99999;
}

/* UpdateSigmaOfOneVariableKernelCall, beam.cu:651-677; kernels
 * beam_kernel.cu:1412-1429 */
static double sig_of(orc_beam* b, const double* a, int good)
{
  double s1 = sum_filtered(b, a, good, 0);
  double s2 = sum_filtered(b, a, good, 1);
  double d = good ? (double)b->st.num_good : (double)(uint32_t)(b->n - b->st.num_loss);
  double avg = s1 / d;
  double avg2 = s2 / d;
  return sqrt(avg2 - avg * avg);
}

/* UpdateRelativePhiKernel, beam_kernel.cu:1506-1522 */
static void update_relative_phi(orc_beam* b, int use_good)
{
  double ref = use_good ? b->st.phi_avg_good : b->phi[0];
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    double tmp = b->phi[i] - ref + PI;
    double sign = 1.0;
    if (tmp < 0) { sign = -1.0; tmp = -tmp; }
    b->phi_r[i] = sign * (orc_fmod_pos(tmp, TWOPI) - PI);
  }
}

/* Beam::UpdateLongitudinalLoss, beam.cpp:776-782;
 * UpdateLongitudinalLossCoordinateKernel beam_kernel.cu:1549-1559 */
static void update_longitudinal_loss(orc_beam* b)
{
  update_good_count(b);
  b->st.phi_avg_good = avg_of(b, b->phi, 1);
  double thr = b->st.phi_avg_good + TWOPI + PI;
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i)
    if (b->phi[i] > thr) b->lloss[i] = 1;
  update_lloss_count(b);
}

/* UpdateAvgSigXYKernelCall + UpdateHorizontalAvgSigKernel,
 * beam.cu:919-938, beam_kernel.cu:1375-1390 */
static void update_avg_sig_xy(orc_beam* b)
{
  double sx = sum_filtered(b, b->x, 1, 0);
  double sx2 = sum_filtered(b, b->x, 1, 1);
  double sy = sum_filtered(b, b->y, 1, 0);
  double sy2 = sum_filtered(b, b->y, 1, 1);
  double surv = (double)(uint32_t)(b->n - b->st.num_loss);
  double xavg = sx / surv, x2avg = sx2 / surv, yavg = sy / surv, y2avg = sy2 / surv;
  b->st.x_avg_good = xavg;
  b->st.x_sig_good = sqrt(x2avg - xavg * xavg);
  b->st.y_avg_good = yavg;
  b->st.y_sig_good = sqrt(y2avg - yavg * yavg);
}

/* UpdateAvgXYKernelCall + UpdateHorizontalAvgXYKernel, beam.cu:739-760,
 * beam_kernel.cu:1364-1373 */
static void update_avg_xy(orc_beam* b)
{
  double sx = sum_filtered(b, b->x, 0, 0);
  double sy = sum_filtered(b, b->y, 0, 0);
  double surv = (double)(uint32_t)(b->n - b->st.num_loss);
  b->st.x_avg = sx / surv;
  b->st.y_avg = sy / surv;
}

/* Beam::UpdateStatForSpaceCharge, beam.cpp:1003-1016 */
void orc_beam_update_stat_for_space_charge(orc_beam* b)
{
  if (b->n > 2) {
    update_longitudinal_loss(b);
    update_good_count(b);
    b->st.phi_avg_good = avg_of(b, b->phi, 1);
    update_relative_phi(b, 1);
    b->st.phi_avg_r = avg_of(b, b->phi_r, 0);
    b->st.phi_sig_good = sig_of(b, b->phi_r, 1);
    update_avg_sig_xy(b);
    b->st.w_avg_good = avg_of(b, b->w, 1);
  }
}

/* Beam::InitPhiAvgGood, beam.cpp:655-661 */
static void init_phi_avg_good(orc_beam* b)
{
  orc_beam_update_loss(b);
  b->st.phi_avg = avg_of(b, b->phi, 0);
  b->st.phi_avg_good = b->st.phi_avg;
}

/* Beam::InitBeamFromDistribution, beam.cpp:317-368 */
int orc_beam_set_distribution(orc_beam* b, const double* x, const double* xp,
    const double* y, const double* yp, const double* phi, const double* w,
    const uint32_t* loss, const uint32_t* lloss)
{
  size_t nb = (size_t)b->n * sizeof(double);
  memcpy(b->x, x, nb); memcpy(b->xp, xp, nb); memcpy(b->y, y, nb);
  memcpy(b->yp, yp, nb); memcpy(b->phi, phi, nb); memcpy(b->w, w, nb);
  b->x[0] = 0.0; b->xp[0] = 0.0; b->y[0] = 0.0; b->yp[0] = 0.0;
  if (loss) { memcpy(b->loss, loss, (size_t)b->n * 4); b->loss[0] = 0; }
  else memset(b->loss, 0, (size_t)b->n * 4);
  if (lloss) { memcpy(b->lloss, lloss, (size_t)b->n * 4); b->lloss[0] = 0; }
  else memset(b->lloss, 0, (size_t)b->n * 4);
  init_phi_avg_good(b);
  return 0;
}

/* Uniform deviate on [-1, 1] from the C library generator, as the
 * reference draws it (beam.cpp:424-425). */
static double rand_pm1(void) { return 2.0 * rand() / RAND_MAX - 1.0; }

/* Twiss half-widths of a PARMILA phase-space ellipse (beam.cpp:409-418):
 * g = (1 + a^2)/b, umax = sqrt(e/g), upmax = sqrt(e*g), du = -a/g. */
static void twiss_widths(double a, double bt, double e, double* umax, double* upmax,
                         double* du)
{
  double g = (1.0 + a * a) / bt;
  *umax = sqrt(e / g);
  *upmax = sqrt(e * g);
  *du = -a / g;
}

/* Beam::InitWaterbagBeam, beam.cpp:391-463.  Points uniform in the 6-ball
 * (three disc draws by rejection, then rejection on the joint radius),
 * mapped onto the three Twiss ellipses; x, y in m from cm, phase from the
 * longitudinal ellipse in cm at beta*lambda of sync_w. */
int orc_beam_set_waterbag(orc_beam* b, double ax, double bx, double ex, double ay,
    double by, double ey, double az, double bz, double ez, double sync_phi,
    double sync_w, double freq, uint32_t seed)
{
  srand(seed == 0 ? (unsigned)time(NULL) : seed);
  double xmx, xpmx, dx, ymx, ypmx, dy, zmx, zpmx_unused, dz;
  twiss_widths(ax, bx, ex, &xmx, &xpmx, &dx);
  twiss_widths(ay, by, ey, &ymx, &ypmx, &dy);
  twiss_widths(az, bz, ez, &zmx, &zpmx_unused, &dz);
  double zpmx = ez / zmx;   /* beam.cpp:418 uses e/zmax, not sqrt(e*g) */
  double wave_len = 29979.2458 / freq;   /* cm */
  b->freq = freq;
  double gm = sync_w / b->mass + 1.0;
  double beta = sqrt(1.0 - 1.0 / (gm * gm));
  uint32_t n = b->n;
  double* v = (double*)malloc(sizeof(double) * 6 * (size_t)n);
  if (!v) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    double r1, r2, r3, r4, r5, r6;
    do {
      do { r1 = rand_pm1(); r2 = rand_pm1(); } while (r1 * r1 + r2 * r2 > 1.0);
      do { r3 = rand_pm1(); r4 = rand_pm1(); } while (r3 * r3 + r4 * r4 > 1.0);
      do { r5 = rand_pm1(); r6 = rand_pm1(); } while (r5 * r5 + r6 * r6 > 1.0);
    } while (r1 * r1 + r2 * r2 + r3 * r3 + r4 * r4 + r5 * r5 + r6 * r6 > 1.0);
    v[i] = 0.01 * (r1 * xmx + dx * r2 * xpmx);
    v[n + i] = r2 * xpmx;
    v[2 * (size_t)n + i] = 0.01 * (r3 * ymx + dy * r4 * ypmx);
    v[3 * (size_t)n + i] = r4 * ypmx;
    v[4 * (size_t)n + i] = -(r5 * zmx + dz * zpmx * r6) * 2.0 * PI / (beta * wave_len) + sync_phi;
    v[5 * (size_t)n + i] = 2.0 * sync_w * zpmx * r6 + sync_w;
  }
  int rc = orc_beam_set_distribution(b, v, v + n, v + 2 * (size_t)n, v + 3 * (size_t)n,
                                     v + 4 * (size_t)n, v + 5 * (size_t)n, NULL, NULL);
  free(v);
  return rc;
}

/* Beam::InitDCBeam, beam.cpp:480-537: transverse points uniform in the
 * 4-ball, phase a uniform ramp over [-dphi, dphi) by index, w = sync_w. */
int orc_beam_set_dc(orc_beam* b, double ax, double bx, double ex, double ay,
    double by, double ey, double dphi, double sync_phi, double sync_w, uint32_t seed)
{
  srand(seed == 0 ? (unsigned)time(NULL) : seed);
  double xmx, xpmx, dx, ymx, ypmx, dy;
  twiss_widths(ax, bx, ex, &xmx, &xpmx, &dx);
  twiss_widths(ay, by, ey, &ymx, &ypmx, &dy);
  uint32_t n = b->n;
  double* v = (double*)malloc(sizeof(double) * 6 * (size_t)n);
  if (!v) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    double r1, r2, r3, r4;
    do {
      do { r1 = rand_pm1(); r2 = rand_pm1(); } while (r1 * r1 + r2 * r2 > 1.0);
      do { r3 = rand_pm1(); r4 = rand_pm1(); } while (r3 * r3 + r4 * r4 > 1.0);
    } while (r1 * r1 + r2 * r2 + r3 * r3 + r4 * r4 > 1.0);
    double r5 = 2.0 * ((double)i + 1.0) / (double)n - 1.0;
    v[i] = 0.01 * (r1 * xmx + dx * r2 * xpmx);
    v[n + i] = r2 * xpmx;
    v[2 * (size_t)n + i] = 0.01 * (r3 * ymx + dy * r4 * ypmx);
    v[3 * (size_t)n + i] = r4 * ypmx;
    v[4 * (size_t)n + i] = r5 * dphi + sync_phi;
    v[5 * (size_t)n + i] = sync_w;
  }
  int rc = orc_beam_set_distribution(b, v, v + n, v + 2 * (size_t)n, v + 3 * (size_t)n,
                                     v + 4 * (size_t)n, v + 5 * (size_t)n, NULL, NULL);
  free(v);
  return rc;
}

/* Beam::SetRefEnergy / SetRefPhase, beam.cpp:1074-1089 */
void orc_beam_set_ref_w(orc_beam* b, double w) { b->w[0] = w; b->design_w = w; }
void orc_beam_set_ref_phi(orc_beam* b, double phi) { b->phi[0] = phi; }

/* Beam::ChangeFrequency, beam.cpp:1654-1668 (ChangeFrequencyKernel
 * beam_kernel.cu:1531-1537) */
static void change_frequency(orc_beam* b, double f)
{
  double ratio = f / b->freq;
  b->current *= ratio;
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) b->phi[i] *= ratio;
  b->freq = f;
  b->st.phi_avg = avg_of(b, b->phi, 0);
  update_relative_phi(b, 0);
  b->st.phi_avg_r = avg_of(b, b->phi_r, 0);
}

/* BeamSetFrequency, wrap_beam.cpp:315-341 */
void orc_beam_set_frequency(orc_beam* b, double f)
{
  if (b->freq == 0) b->freq = f;
  else change_frequency(b, f);
}

/* Transverse emittance, beam.cu:1007-1064 + beam_kernel.cu:1333-1354 */
static void update_transverse_emittance(orc_beam* b, int axis)
{
  const double* a = axis == 0 ? b->x : b->y;
  const double* ap = axis == 0 ? b->xp : b->yp;
  double s_x = sum_filtered(b, a, 0, 0);
  double s_x2 = sum_filtered(b, a, 0, 1);
  double s_xp = sum_filtered(b, ap, 0, 0);
  double s_xp2 = sum_filtered(b, ap, 0, 1);
  double s_xxp = sum_product(b, a, ap);
  double surv = (double)(uint32_t)(b->n - b->st.num_loss);
  double x = s_x / surv, x2 = s_x2 / surv, xp = s_xp / surv, xp2 = s_xp2 / surv, xxp = s_xxp / surv;
  double gamma1 = b->w[0] / b->mass;
  double betagamma = sqrt(gamma1 * (gamma1 + 2.0));
  double sig = sqrt(x2 - x * x), sigp = sqrt(xp2 - xp * xp);
  double emit = sqrt((x2 - x * x) * (xp2 - xp * xp) - (xxp - x * xp) * (xxp - x * xp)) * betagamma;
  if (axis == 0) { b->st.x_avg = x; b->st.x_sig = sig; b->st.xp_avg = xp; b->st.xp_sig = sigp; b->st.x_emit = emit; }
  else { b->st.y_avg = x; b->st.y_sig = sig; b->st.yp_avg = xp; b->st.yp_sig = sigp; b->st.y_emit = emit; }
}

/* Longitudinal emittance, beam.cu:1069-1092 + beam_kernel.cu:1356-1373 */
static void update_longitudinal_emittance(orc_beam* b)
{
  double s_p = sum_filtered(b, b->phi_r, 0, 0);
  double s_p2 = sum_filtered(b, b->phi_r, 0, 1);
  double s_w = sum_filtered(b, b->w, 0, 0);
  double s_w2 = sum_filtered(b, b->w, 0, 1);
  double s_pw = sum_product(b, b->phi_r, b->w);
  double surv = (double)(uint32_t)(b->n - b->st.num_loss);
  double phi = s_p / surv, phi2 = s_p2 / surv, w = s_w / surv, w2 = s_w2 / surv, phi_w = s_pw / surv;
  double phi_var = phi2 - phi * phi;
  double w_var = w2 - w * w;
  double cov = phi_w - phi * w;
  b->st.phi_avg_r = phi;
  b->st.w_avg = w;
  b->st.phi_sig_r = sqrt(phi_var);
  b->st.w_sig = sqrt(w_var);
  b->st.z_emit = sqrt(phi_var * w_var - cov * cov);
}

/* Beam::UpdateEmittance, beam.cpp:985-994 */
static void update_emittance(orc_beam* b)
{
  if (b->n > 2) {
    update_transverse_emittance(b, 0);
    update_transverse_emittance(b, 1);
    update_relative_phi(b, 0);
    update_longitudinal_emittance(b);
  }
}

/* BeamGetEmitX/Y/Z, wrap_beam.cpp:1188-1243 */
double orc_beam_emittance(orc_beam* b, int axis)
{
  orc_beam_update_loss(b);
  update_emittance(b);
  if (b->n < 2) return 0.0;
  return axis == 0 ? b->st.x_emit : (axis == 1 ? b->st.y_emit : b->st.z_emit);
}

/* BeamGetLossNum, wrap_beam.cpp:1245-1284 */
uint32_t orc_beam_loss_num(orc_beam* b, int which)
{
  if (which == 0) { orc_beam_update_loss(b); return b->st.num_loss; }
  if (which == 1) { update_longitudinal_loss(b); return b->st.num_lloss; }
  orc_beam_update_loss(b);
  update_longitudinal_loss(b);
  update_good_count(b);
  return b->n - b->st.num_good;
}

/* BeamGetAvg{X,Y,Phi,W}, wrap_beam.cpp:914-1049 */
double orc_beam_avg(orc_beam* b, int var, int good)
{
  orc_beam_update_loss(b);
  switch (var) {
    case 0:
      if (!good) return b->st.x_avg = avg_of(b, b->x, 0);
      return b->st.x_avg_good = avg_of(b, b->x, 1);
    case 1:
      if (!good) return b->st.y_avg = avg_of(b, b->y, 0);
      return b->st.y_avg_good = avg_of(b, b->y, 1);
    case 2:
      if (!good) return b->st.phi_avg = avg_of(b, b->phi, 0);
      update_good_count(b);
      return b->st.phi_avg_good = avg_of(b, b->phi, 1);
    case 3:
      if (!good) return b->st.w_avg = avg_of(b, b->w, 0);
      return b->st.w_avg_good = avg_of(b, b->w, 1);
    default:
      update_relative_phi(b, 0);
      return b->st.phi_avg_r = avg_of(b, b->phi_r, good);
  }
}

/* BeamGetSig{X,Y,Phi,W}, wrap_beam.cpp:1051-1186 */
double orc_beam_sig(orc_beam* b, int var, int good)
{
  orc_beam_update_loss(b);
  if (b->n < 2) return 0.0;
  int upd = b->n > 2;
  switch (var) {
    case 0:
      if (upd) { if (good) b->st.x_sig_good = sig_of(b, b->x, 1); else b->st.x_sig = sig_of(b, b->x, 0); }
      return good ? b->st.x_sig_good : b->st.x_sig;
    case 1:
      if (upd) { if (good) b->st.y_sig_good = sig_of(b, b->y, 1); else b->st.y_sig = sig_of(b, b->y, 0); }
      return good ? b->st.y_sig_good : b->st.y_sig;
    case 2:
      if (!good) { if (upd) b->st.phi_sig = sig_of(b, b->phi, 0); return b->st.phi_sig; }
      update_good_count(b);
      b->st.phi_avg_good = avg_of(b, b->phi, 1);
      update_relative_phi(b, 1);
      if (upd) b->st.phi_sig_good = sig_of(b, b->phi_r, 1);
      return b->st.phi_sig_good;
    case 3:
      if (upd) { if (good) b->st.w_sig_good = sig_of(b, b->w, 1); else b->st.w_sig = sig_of(b, b->w, 0); }
      return good ? b->st.w_sig_good : b->st.w_sig;
    default:
      update_relative_phi(b, 0);
      if (upd) b->st.phi_sig_r = sig_of(b, b->phi_r, 0);
      return b->st.phi_sig_r;
  }
}

/* Save / restore, beam.cpp:547-647 */
static void save_beam(orc_beam* b, double** s, uint32_t** l, uint32_t** ll)
{
  double* src[6] = {b->x, b->xp, b->y, b->yp, b->phi, b->w};
  for (int k = 0; k < 6; ++k) {
    if (!s[k]) s[k] = (double*)malloc((size_t)b->n * sizeof(double));
    memcpy(s[k], src[k], (size_t)b->n * sizeof(double));
  }
  if (!*l) *l = (uint32_t*)malloc((size_t)b->n * 4);
  if (!*ll) *ll = (uint32_t*)malloc((size_t)b->n * 4);
  memcpy(*l, b->loss, (size_t)b->n * 4);
  memcpy(*ll, b->lloss, (size_t)b->n * 4);
}
static void restore_beam(orc_beam* b, double** s, uint32_t* l, uint32_t* ll)
{
  double* dst[6] = {b->x, b->xp, b->y, b->yp, b->phi, b->w};
  for (int k = 0; k < 6; ++k) memcpy(dst[k], s[k], (size_t)b->n * sizeof(double));
  memcpy(b->loss, l, (size_t)b->n * 4);
  memcpy(b->lloss, ll, (size_t)b->n * 4);
}
void orc_beam_save_initial(orc_beam* b)
{
  save_beam(b, b->s0, &b->s0l, &b->s0ll);
  b->saved0 = 1; b->dw0 = b->design_w; b->f0 = b->freq; b->c0 = b->current;
}
void orc_beam_save_intermediate(orc_beam* b)
{
  save_beam(b, b->s1, &b->s1l, &b->s1ll);
  b->saved1 = 1; b->dw1 = b->design_w; b->f1 = b->freq; b->c1 = b->current;
}
void orc_beam_restore_initial(orc_beam* b)
{
  if (!b->saved0) return;
  restore_beam(b, b->s0, b->s0l, b->s0ll);
  b->design_w = b->dw0; b->freq = b->f0; b->current = b->c0;
  b->st.phi_avg_good = avg_of(b, b->phi, 1);
  update_relative_phi(b, 1);
}
void orc_beam_restore_intermediate(orc_beam* b)
{
  if (!b->saved1) return;
  restore_beam(b, b->s1, b->s1l, b->s1ll);
  b->design_w = b->dw1; b->freq = b->f1; b->current = b->c1;
  b->st.phi_avg_good = avg_of(b, b->phi, 1);
  update_relative_phi(b, 1);
}

/* ShiftVariableKernel beam_kernel.cu:1489-1496 (skips particle 0);
 * Beam::ShiftPhi re-runs InitPhiAvgGood, beam.cpp:1637-1641 */
void orc_beam_translate(orc_beam* b, int axis, double val)
{
  double* arr[6] = {b->x, b->xp, b->y, b->yp, b->phi, b->w};
  if (axis < 0 || axis > 5) return;
  for (uint32_t i = 1; i < b->n; ++i) arr[axis][i] += val;
  if (axis == 4) init_phi_avg_good(b);
}

/* Beam::ApplyCut + CutBeamKernel, beam.cpp:1595-1610,
 * beam_kernel.cu:1464-1480: particles strictly inside (lo, hi) are marked
 * 99999; particle 0 is skipped. */
void orc_beam_apply_cut(orc_beam* b, char coord, double lo, double hi)
{
  const double* cd = NULL;
  if (coord == 'x' || coord == 'X') cd = b->x;
  if (coord == 'y' || coord == 'Y') cd = b->y;
  if (coord == 'w' || coord == 'W') cd = b->w;
  if (coord == 'p' || coord == 'P') cd = b->phi;
  if (!cd) return;
  for (uint32_t i = 1; i < b->n; ++i)
    if (b->loss[i] == 0 && cd[i] < hi && cd[i] > lo) b->loss[i] = 99999;
}

/* ------------------------------------------------------------------ */
/* Element maps (simulation_kernel.cu)                                 */
/* ------------------------------------------------------------------ */

/* Bessi0c, simulation_kernel.cu:15-36 */
static double bessi0c(double r_x)
{
  double ans, y, ax = fabs(r_x);
  if (ax < 3.75) {
    y = r_x / 3.75;
    y = y * y;
    ans = 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
          y * (0.360768e-1 + y * 0.45813e-2)))));
  } else {
    y = 3.75 / ax;
    ans = (orc_exp(ax) * (1.0 / sqrt(ax))) * (0.39894228 + y * (0.1328592e-1 +
          y * (0.225319e-2 + y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 +
          y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
  }
  return ans;
}

/* Bessi1c, simulation_kernel.cu:38-64 */
static double bessi1c(double r_x)
{
  double y, ans, ax = fabs(r_x);
  if (ax < 3.75) {
    y = r_x / 3.75;
    y = y * y;
    ans = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 +
          y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    y = 3.75 / ax;
    ans = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    ans = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 + y *
          (-0.1031555e-1 + y * ans))));
    ans = ans * (orc_exp(ax) * (1.0 / sqrt(ax)));
  }
  return r_x < 0.0 ? -ans : ans;
}

/* Bessi1p, simulation_kernel.cu:66-73 */
static double bessi1p(double r_x)
{
  if (r_x != 0.0) return bessi0c(r_x) - bessi1c(r_x) / r_x;
  return 0.5;
}

/* GetTransitTimeFactors, simulation_kernel.cu:290-308 */
static void ttf(double r_beta, const orc_gap_param* g, double* t, double* tp, double* sp)
{
  double beta = r_beta;
  if (r_beta > 1.1 * g->fit_beta_center) beta = g->fit_beta_center;
  else if (r_beta < g->fit_beta_min) beta = g->fit_beta_min;
  *t = g->fit_t0 + beta * (g->fit_t1 + beta * (g->fit_t2 + beta * (g->fit_t3 + beta *
       (g->fit_t4 + beta * g->fit_t5))));
  double coef = beta * beta / (TWOPI * g->fit_beta_center);
  *tp = coef * (g->fit_t1 + beta * (2 * g->fit_t2 + beta * (3 * g->fit_t3 + beta *
        (4 * g->fit_t4 + 5 * beta * g->fit_t5))));
  *sp = -coef * (g->fit_s1 + beta * (2 * g->fit_s2 + beta * (3 * g->fit_s3 + beta *
        (4 * g->fit_s4 + 5 * beta * g->fit_s5))));
}

double orc_t_exp(double x) { return orc_exp(x); }
double orc_t_log(double x) { return orc_log(x); }
void orc_t_sincos(double x, double* s, double* c) { orc_sincos(x, s, c); }
double orc_t_sinh(double x) { return orc_sinh(x); }
double orc_t_cosh(double x) { return orc_cosh(x); }
double orc_t_fmod_pos(double x, double y) { return orc_fmod_pos(x, y); }
double orc_t_tan(double x) { return orc_tan(x); }
double orc_t_bessi0(double x) { return bessi0c(x); }
double orc_t_bessi1(double x) { return bessi1c(x); }
double orc_t_bessi1p(double x) { return bessi1p(x); }
void orc_t_ttf(double beta, const orc_gap_param* g, double* t, double* tp, double* sp) { ttf(beta, g, t, tp, sp); }

/* SimulatePartialDriftKernel, simulation_kernel.cu:86-118 */
static void map_partial_drift(orc_beam* b, double wavelen, double h, double aper, uint32_t idx)
{
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double aper2 = aper * aper;
    double x = b->x[i];
    double y = b->y[i];
    x += b->xp[i] * h;
    y += b->yp[i] * h;
    if (aper2 != 0.0 && x * x + y * y > aper2)
      b->loss[i] = idx;
    else {
      b->x[i] = x;
      b->y[i] = y;
      double gm1 = b->w[i] / b->mass;
      double bt = sqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
      b->phi[i] += TWOPI * h / (wavelen * bt);
    }
  }
}

/* SimulateHalfQuadKernel, simulation_kernel.cu:142-218 */
static void map_half_quad(orc_beam* b, double wavelen, double length, double aper,
    double gradient, uint32_t idx)
{
  double aper2 = aper;
  aper2 *= aper2;
  double qg = gradient * b->charge;
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double x = b->x[i], y = b->y[i], xp = b->xp[i], yp = b->yp[i];
    double w = b->w[i];
    if (qg == 0) {
      b->x[i] = x + xp * length * 0.5;
      b->y[i] = y + yp * length * 0.5;
    } else {
      double gamma1 = w / b->mass;
      double bg = sqrt(gamma1 * (gamma1 + 2.0) * (1.0 + xp * xp + yp * yp));
      double absqg = qg > 0.0 ? qg : -qg;
      double k = sqrt(absqg / (bg * b->mass) * CLIGHT);
      double inv_k = 1.0 / k;
      double kds = length * 0.25 * k;
      double s, c, sh = orc_sinh(kds), ch = orc_cosh(kds);
      orc_sincos(kds, &s, &c);
      double xx, yy, sovk = s * inv_k, shovk = sh * inv_k, sk = s * k, shk = sh * k;
      for (int rep = 0; rep < 2; ++rep) {
        if (qg > 0.0) {
          xx = c * x + sovk * xp; xp = -sk * x + c * xp; x = xx;
          yy = ch * y + shovk * yp; yp = shk * y + ch * yp; y = yy;
        } else {
          xx = ch * x + shovk * xp; xp = shk * x + ch * xp; x = xx;
          yy = c * y + sovk * yp; yp = -sk * y + c * yp; y = yy;
        }
      }
      b->x[i] = x; b->y[i] = y; b->xp[i] = xp; b->yp[i] = yp;
    }
    if (aper2 != 0 && x * x + y * y > aper2)
      b->loss[i] = idx;
    else {
      double gm1 = w / b->mass;
      double bt = sqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
      b->phi[i] += TWOPI * length * 0.5 / (wavelen * bt);
    }
  }
}

/* SimulateRFGapFirstHalfKernel, simulation_kernel.cu:226-288 */
static void map_gap_first_half(orc_beam* b, double design_w_in, double phi_in,
    const orc_gap_param* gap, double length, double q1, double q2, int ccl, double* wavelen)
{
  double wave_len = CLIGHT / gap->frequency;
  *wavelen = wave_len;
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double x = b->x[i], y = b->y[i], xp = b->xp[i], yp = b->yp[i];
    double phi = b->phi[i], w = b->w[i];
    double cell_len = length + q1 * 0.5 + q2 * 0.5;
    double dd1 = 0.5 * cell_len - gap->dg - 0.5 * q1;
    b->x[i] = x + dd1 * xp;
    b->y[i] = y + dd1 * yp;
    if (gap->amplitude == 0.0) {
      double gm1 = w / b->mass;
      double bt = sqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
      b->phi[i] += TWOPI * dd1 / (wave_len * bt);
    } else {
      double gam = design_w_in / b->mass;
      double beta_in = sqrt((gam + 2.0) * gam) / (gam + 1.0);
      gam = w / b->mass;
      double beta = sqrt((gam + 2.0) * gam) / (gam + 1.0);
      if (!ccl) {
        double dps = gap->phase_ref - phi_in;
        double d_cell_len = cell_len - gap->beta_center * wave_len;
        double dphi_qlen1 = TWOPI * 0.5 * q1 / (beta * wave_len);
        double dphi_clen = 0.0;
        if (d_cell_len > 1e-15 || d_cell_len < -1e-15)
          dphi_clen = PI * d_cell_len / (beta * wave_len);
        b->phi[i] = phi + (1.0 - (beta - beta_in) / beta) *
          (PI * gap->cell_length_over_beta_lambda + dps) + dphi_clen - dphi_qlen1;
      } else {
        double dps = gap->phase_ref - phi_in;
        double betag = cell_len / (gap->cell_length_over_beta_lambda * wave_len);
        b->phi[i] = phi + (1.0 - (beta - beta_in) * betag / (beta * beta_in)) *
          (PI * gap->cell_length_over_beta_lambda + dps);
      }
    }
  }
}

/* SimulateRFGapSecondHalfKernel, simulation_kernel.cu:310-497 */
static void map_gap_second_half(orc_beam* b, double wavelen, double design_w_in,
    double phi_in, const orc_gap_param* gap, double length, uint32_t ccl_cell_num,
    double q1, double q2, int ccl)
{
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double cell_len = length + q1 * 0.5 + q2 * 0.5;
    double dd2 = 0.5 * cell_len + gap->dg - q2 * 0.5;
    double x = b->x[i], y = b->y[i], xp = b->xp[i], yp = b->yp[i];
    double inv_mass = 1.0 / b->mass;
    double phi = b->phi[i];
    double w = b->w[i];
    double gm_in = design_w_in * inv_mass + 1.0;
    double beta_in = sqrt(1.0 - 1.0 / (gm_in * gm_in));
    double gam = w * inv_mass;
    double bg1 = sqrt(gam * (gam + 2.0));
    double beta = bg1 / (gam + 1.0);
    double t_fac, tp_fac, sp_fac;
    ttf(beta, gap, &t_fac, &tp_fac, &sp_fac);
    if (gap->amplitude == 0.0 || (ccl && beta < gap->fit_beta_min)) {
      b->x[i] = x + dd2 * xp;
      b->y[i] = y + dd2 * yp;
      double gm1 = b->w[i] / b->mass;
      double bt = sqrt(gm1 * (gm1 + 2.0)) / (gm1 + 1.0);
      b->phi[i] += TWOPI * dd2 / (wavelen * bt);
      continue;
    }
    double wave_len = CLIGHT / gap->frequency;
    double cl = cell_len / gap->cell_length_over_beta_lambda;
    double betag_ccl = cl / wave_len;
    double betag = ccl ? gap->beta_center : betag_ccl;
    double betag2 = betag * betag;
    double gmg = 1.0 / sqrt(1.0 - betag2);
    double inv_gmg2 = 1.0 / (gmg * gmg);
    double q = b->charge;
    q = q > 0.0 ? q : -q;
    double gm_out = gap->energy_out * inv_mass + 1.0;
    double beta_out = sqrt(1.0 - 1.0 / (gm_out * gm_out));
    double efl = gap->amplitude * cell_len;
    double concomm = q * efl / (b->mass * betag2);
    double concomm1 = concomm * inv_gmg2 / gmg;
    double con1 = PI * concomm1;
    double con2nex = concomm * inv_gmg2;
    double con5nex, tw, inv_fkis, inv_gmgcl;
    if (ccl) {
      inv_gmgcl = 1.0 / (gmg * betag * wave_len);
      inv_fkis = betag * wave_len / TWOPI;
      con5nex = concomm * (t_fac * inv_fkis - betag_ccl * wave_len * tp_fac * inv_gmg2);
      tw = t_fac - TWOPI * tp_fac * (beta_in / beta - 1.0) * betag_ccl / beta_in;
    } else {
      inv_gmgcl = 1.0 / (gmg * cl);
      inv_fkis = cl / TWOPI;
      con5nex = concomm * (t_fac * inv_fkis - beta_in * wave_len * tp_fac * inv_gmg2);
      tw = t_fac;
    }
    double con5nex1 = concomm1 * t_fac;
    double con5nex2 = -con2nex * t_fac * inv_fkis;
    double fkr = TWOPI * inv_gmgcl;
    double sinps, cosps;
    orc_sincos(gap->phase_ref, &sinps, &cosps);
    double phi1, sinp1, cosp1;
    if (ccl)
      phi1 = phi + gap->phase_shift - TWOPI * gap->dg / (beta * wave_len) + ccl_cell_num * PI;
    else
      phi1 = phi + gap->phase_shift;
    orc_sincos(phi1, &sinp1, &cosp1);
    double prome1 = -con1 * (tp_fac * (sinp1 - sinps) + sp_fac * (cosp1 - cosps));
    orc_sincos(phi1 + prome1, &sinp1, &cosp1);
    double rs_p = x * x + y * y;
    double rrp = x * xp + y * yp;
    double ris, fkrris, rpris = 0.0, bes1covr;
    ris = sqrt(rs_p);
    fkrris = fkr * ris;
    double b0c = bessi0c(fkrris);
    double b1c = bessi1c(fkrris);
    double b1p = bessi1p(fkrris);
    if (ris != 0.0) bes1covr = b1c / ris;
    else bes1covr = fkr * 0.5;
    double delte = q * efl * (tw * b0c * cosp1 + ((tw - TWOPI * tp_fac) / gmg * bes1covr +
                   tw / gmg * fkr * b1p - tw * gmg * bes1covr) * rrp * sinp1);
    w += delte;
    gam = w * inv_mass;
    double bg = sqrt(gam * (gam + 2.0));
    beta = bg / (gam + 1.0);
    double bgr = bg1 / bg;
    double onema = 1.0 - (con5nex * bes1covr + con5nex1 * b1p + con5nex2 * bes1covr) * cosp1;
    double dovbl = con2nex * bes1covr * tw * sinp1;
    double bgrovonema = bgr / onema;
    xp = -dovbl * x + bgrovonema * xp;
    x = x * onema + dd2 * xp;
    yp = -dovbl * y + bgrovonema * yp;
    y = y * onema + dd2 * yp;
    if (rs_p > 0.0) rpris = rrp / ris;
    double prome = 2.0 * con1 * (-tp_fac * b0c + ris * tw * b1c * inv_gmgcl) * (sinp1 - sinps) -
                   efl * inv_mass * tw * b1c * rpris * (cosp1 - cosps);
    double phi_new;
    if (ccl) {
      double dps = phi_in - gap->phase_ref;
      phi_new = phi + (1.0 - (beta - beta_out) * betag_ccl / (beta * beta_out)) *
                (PI * gap->cell_length_over_beta_lambda + dps) + prome;
    } else {
      double d_cell_len = cell_len - gap->beta_center * wave_len;
      double dphi_qlen2 = TWOPI * 0.5 * q2 / (beta * wave_len);
      double dphi_clen = 0.0;
      if (d_cell_len > 1e-15 || d_cell_len < -1e-15)
        dphi_clen = PI * d_cell_len / (beta * wave_len);
      double dps = gap->phase_ref - phi_in;
      phi_new = phi + (1.0 - (beta - beta_out) / beta) * (PI * gap->cell_length_over_beta_lambda +
                dps) + prome + dphi_clen - dphi_qlen2;
    }
    if (w <= 0.0 || x != x || y != y || xp != xp || yp != yp || w != w || phi_new != phi_new) {
      b->loss[i] = 33333333;
    } else {
      b->x[i] = x; b->y[i] = y; b->xp[i] = xp; b->yp[i] = yp; b->w[i] = w; b->phi[i] = phi_new;
    }
  }
}

/* SimulateRotationKernel, simulation_kernel.cu:511-529 (no grid stride:
 * indices >= grid*blck are untouched, which never happens since that
 * product is NextPow2(N) >= N) */
static void map_rotation(orc_beam* b, double angle)
{
  double sina, cosa;
  orc_sincos(angle, &sina, &cosa);
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double x = b->x[i], y = b->y[i], xp = b->xp[i], yp = b->yp[i];
    b->x[i] = x * cosa + y * sina;
    b->y[i] = y * cosa - x * sina;
    b->xp[i] = xp * cosa + yp * sina;
    b->yp[i] = yp * cosa - xp * sina;
  }
}

/* SimulateSteererKernel, simulation_kernel.cu:120-140 */
static void map_steerer(orc_beam* b, double blh, double blv)
{
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double gm1 = b->w[i] / b->mass;
    double btgm = sqrt(gm1 * (gm1 + 2.0));
    double p = btgm * b->mass / CLIGHT;
    if (blh > 1e-10) b->xp[i] += blh * b->charge / p;
    if (blv > 1e-10) b->yp[i] += blv * b->charge / p;
  }
}

/* SimulateCircularApertureKernel, simulation_kernel.cu:543-556 */
static void map_circular_aperture(orc_beam* b, double aper, uint32_t idx)
{
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 1; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double aper2 = aper * aper;
    double x = b->x[i] - b->st.x_avg;
    double y = b->y[i] - b->st.y_avg;
    if (x * x + y * y > aper2) b->loss[i] = idx;
  }
}

/* SimulateRectangularApertureKernel, simulation_kernel.cu:558-575 */
static void map_rect_aperture(orc_beam* b, double xl, double xr, double yt, double yb, uint32_t idx)
{
  double xmin = b->st.x_avg - xl, xmax = b->st.x_avg + xr;
  double ymin = b->st.y_avg - yb, ymax = b->st.y_avg + yt;
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 1; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double x = b->x[i], y = b->y[i];
    if (x < xmin || x > xmax || y < ymin || y > ymax) b->loss[i] = idx;
  }
}

/* SimulateBuncherKernel, simulation_kernel.cu:807-874 (launch
 * <<<grid, blck>>>, every particle).  wavelen = CLIGHT / buncher frequency
 * (UpdateWaveLengthKernel, simulation_engine.cu:352); the phase argument
 * uses phi scaled by freq_ratio = buncher / beam frequency. */
static void map_buncher(orc_beam* b, double w_ref, double freq, double voltage, double phase,
    double aper, double freq_ratio, double wavelen, uint32_t idx)
{
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double x = b->x[i], y = b->y[i], xp = b->xp[i], yp = b->yp[i];
    double phi = b->phi[i] * freq_ratio;
    double w = b->w[i];
    double r2 = x * x + y * y;
    double aper2 = aper * aper;
    if (aper2 != 0.0 && r2 > aper2) { b->loss[i] = idx; continue; }
    double h = freq * wavelen / CLIGHT;
    double gm1 = w_ref / b->mass;
    double gm = gm1 + 1.0;
    double btgm = sqrt(gm1 * (gm1 + 2.0));
    double beta = btgm / gm;
    double k = h * TWOPI / (btgm * wavelen);
    double k2 = k * k;
    double q = b->charge;
    q = q > 0.0 ? q : -q;
    double con1 = TWOPI * h * voltage * q / (beta * wavelen * b->mass);
    double arg = k2 * r2 * 0.25;
    double ph = h * phi + phase;
    double sph, cph;
    orc_sincos(ph, &sph, &cph);
    double dw = q * voltage * cph * (1.0 + arg * (1.0 + arg * (0.25 + arg / 36.0)));
    double wb_par = w + 0.5 * dw;
    double wf = w + dw;
    if (wb_par <= 0.0 || wf <= 0.0) { b->loss[i] = idx; continue; }
    double gmtmp = wb_par / b->mass;
    double bav = sqrt(gmtmp * (gmtmp + 2.0)) / (gmtmp + 1.0);
    gmtmp = w / b->mass;
    double bg = sqrt(gmtmp * (gmtmp + 2.0));
    gmtmp = wf / b->mass;
    double rbgf = 1.0 / sqrt(gmtmp * (gmtmp + 2.0));
    double del = -con1 * sph * (1.0 - bav * beta) * (0.5 + arg * (0.25 + arg / 24.0)) / bav;
    b->xp[i] = (xp * bg + del * x) * rbgf;
    b->yp[i] = (yp * bg + del * y) * rbgf;
    b->w[i] = wf;
  }
}

/* Transfer coefficients of half a sector dipole (arc dl = rho * angle / 2)
 * with field index n, simulation_kernel.cu:625-683 / 725-780. */
typedef struct {
  double r11, r12, r16, r21, r22, r26, r33, r34, r43, r44, r51, r52, r56, beta, dl;
} orc_dipole_half;

static orc_dipole_half dipole_half(const orc_element* e, double mass)
{
  orc_dipole_half o;
  double h = 1.0 / e->rho;
  double h2 = h * h;
  double dl = 0.5 * e->rho * e->angle;
  double yk2 = e->field_index * h2;
  double yk = sqrt(fabs(yk2));
  double inv_yk = 1.0 / yk;
  double ykl = yk * dl;
  double xk2 = h2 - yk2;
  double inv_xk2 = 1.0 / xk2;
  double xk = sqrt(fabs(xk2));
  double inv_xk = 1.0 / xk;
  double xkl = xk * dl;
  double cxkl = 1.0, sxokx = dl, s, c;
  if (xk2 > 0.0) { orc_sincos(xkl, &s, &c); cxkl = c; sxokx = s * inv_xk; }
  else if (xk2 < 0.0) { cxkl = orc_cosh(xkl); sxokx = orc_sinh(xkl) * inv_xk; }
  double cykl = 1.0, syoky = dl;
  if (yk2 > 0.0) { orc_sincos(ykl, &s, &c); cykl = c; syoky = s * inv_yk; }
  else if (yk2 < 0.0) { cykl = orc_cosh(ykl); syoky = orc_sinh(ykl) * inv_yk; }
  o.r11 = cxkl; o.r12 = sxokx; o.r21 = -xk2 * sxokx; o.r22 = cxkl;
  o.r33 = cykl; o.r34 = syoky; o.r43 = -yk2 * syoky; o.r44 = cykl;
  o.r26 = h * sxokx;
  o.r16 = h * (1.0 - cxkl) * inv_xk2;
  o.r51 = -o.r26;
  o.r52 = -o.r16;
  double gamma = e->kinetic_energy / mass + 1.0;
  double beta2 = 1.0 - 1.0 / (gamma * gamma);
  o.beta = sqrt(beta2);
  o.r56 = -h2 * (dl * beta2 - sxokx) * inv_xk2 + dl * (1.0 - h2 * inv_xk2) / (gamma * gamma);
  o.dl = dl;
  return o;
}

/* Edge kick of a dipole face at angle b with fringe-field correction psi,
 * simulation_kernel.cu:628-641 / 788-796: returns r21 and r43. */
static void dipole_edge(const orc_element* e, double b, double* r21, double* r43)
{
  double h = 1.0 / e->rho;
  double c1 = 2.0 * e->half_gap * e->k1 * h;
  double sinb, cosb;
  orc_sincos(b, &sinb, &cosb);
  double inv_cosb = 1.0 / cosb;
  double psi = c1 * (1.0 + sinb * sinb) * inv_cosb * (1.0 - c1 * e->k2 * sinb * inv_cosb);
  *r21 = orc_tan(b) * h;
  *r43 = orc_tan(b - psi) * h;
}

/* SimulateFirstHalfDipoleKernel, simulation_kernel.cu:603-705: entrance
 * edge, then the first half sector. */
static void map_dipole_first_half(orc_beam* b, const orc_element* e, double wavelen)
{
  double er21, er43;
  dipole_edge(e, e->edge_angle_in, &er21, &er43);
  orc_dipole_half d = dipole_half(e, b->mass);
  double design_w = e->kinetic_energy;
  double gamma = design_w / b->mass + 1.0;
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double x = b->x[i], xp = b->xp[i], y = b->y[i], yp = b->yp[i];
    xp += er21 * x;
    yp -= er43 * y;
    double dpop = gamma / (design_w * (gamma + 1.0)) * (b->w[i] - design_w);
    b->x[i] = d.r11 * x + d.r12 * xp + d.r16 * dpop;
    b->xp[i] = d.r21 * x + d.r22 * xp + d.r26 * dpop;
    b->y[i] = d.r33 * y + d.r34 * yp;
    b->yp[i] = d.r43 * y + d.r44 * yp;
    b->phi[i] += TWOPI * (-(d.r51 * x + d.r52 * xp + d.r56 * dpop) + d.dl) / (d.beta * wavelen);
  }
}

/* SimulateSecondHalfDipoleKernel, simulation_kernel.cu:706-805: second half
 * sector, then the exit edge. */
static void map_dipole_second_half(orc_beam* b, const orc_element* e, double wavelen)
{
  double er21, er43;
  dipole_edge(e, e->edge_angle_out, &er21, &er43);
  orc_dipole_half d = dipole_half(e, b->mass);
  double design_w = e->kinetic_energy;
  double gamma = design_w / b->mass + 1.0;
  #pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double x = b->x[i], xp = b->xp[i], y = b->y[i], yp = b->yp[i];
    double dpop = gamma / (design_w * (gamma + 1.0)) * (b->w[i] - design_w);
    b->phi[i] += TWOPI * (-(d.r51 * x + d.r52 * xp + d.r56 * dpop) + d.dl) / (d.beta * wavelen);
    double x_new = d.r11 * x + d.r12 * xp + d.r16 * dpop;
    xp = d.r21 * x + d.r22 * xp + d.r26 * dpop;
    double y_new = d.r33 * y + d.r34 * yp;
    yp = d.r43 * y + d.r44 * yp;
    b->x[i] = x_new;
    b->y[i] = y_new;
    b->xp[i] = xp + er21 * x_new;
    b->yp[i] = yp - er43 * y_new;
  }
}

/* ------------------------------------------------------------------ */
/* SCHEFF space charge                                                 */
/* ------------------------------------------------------------------ */

struct orc_spch {
  orc_spch_state s;
  double* fld1;     /* (nr+1)*nr*nz double2 */
  double* fld2;     /* (nz+1)*(nr+1) double2 */
  int64_t* bins;    /* nr*nz */
  /* SetParameters state, space_charge.cu:77-86 */
  double beta, freq, frac;
  uint32_t adj_eff;
  /* field table stored in single precision, as the reference's default
   * build keeps it (DOUBLE_PRECISION off, make_root.mk:22-23,
   * space_charge.h:147-161): the precision floor the engine's approximate
   * field modes are compared with; 0 = FP64 (the checker) */
  int fp32_table;
  /* > 0: the field interpolated linearly in log(dz*gm/dr) between the
   * aspect classes (1+step)^c and (1+step)^(c+1) of the engine's grouped
   * batch solve, in FP64 (log(1+step)); 0 = the reference field */
  double interp_lnr;
  int interp_order;   /* 1 linear (two classes), 2 quadratic (three) */
};

static void spch_alloc(orc_spch* s)
{
  free(s->fld1); free(s->fld2); free(s->bins);
  uint32_t nr = s->s.nx, nz = s->s.nz;
  s->fld1 = (double*)calloc((size_t)(nr + 1) * nr * nz * 2, sizeof(double));
  s->fld2 = (double*)calloc((size_t)(nz + 1) * (nr + 1) * 2, sizeof(double));
  s->bins = (int64_t*)calloc((size_t)nr * nz, sizeof(int64_t));
}

/* SpaceCharge / Scheff constructors, space_charge.cpp:13-40 */
orc_spch* orc_spch_create(uint32_t nr, uint32_t nz, int adj_bunch)
{
  if (nr == 0 || nz == 0) { set_err("mesh sizes must be positive"); return NULL; }
  orc_spch* s = (orc_spch*)calloc(1, sizeof(orc_spch));
  s->s.nx = nr; s->s.nz = nz; s->s.nx_init = nr; s->s.nz_init = nz;
  s->s.adj_bunch = (uint32_t)adj_bunch;
  s->s.fraction = 1.0; s->s.interval = 0.01; s->s.interval_init = 0.01;
  s->s.adj_bunch_w = 0.0; s->s.mesh_w = 0.0;
  s->s.sigr_old = 0.0; s->s.sigz_old = 0.0; s->s.gm_old = 0.0; s->s.good_cnt_old = 0;
  s->s.remesh_val = 0.05;
  spch_alloc(s);
  return s;
}
void orc_spch_destroy(orc_spch* s)
{
  if (!s) return;
  free(s->fld1); free(s->fld2); free(s->bins); free(s);
}
void orc_spch_set_interval(orc_spch* s, double v) { s->s.interval = v; s->s.interval_init = v; }
void orc_spch_set_remesh_threshold(orc_spch* s, double v) { s->s.remesh_val = v; }
void orc_spch_set_adj_bunch(orc_spch* s, uint32_t v) { s->s.adj_bunch = v; }
void orc_spch_set_adj_bunch_cutoff_w(orc_spch* s, double v) { s->s.adj_bunch_w = v; }
void orc_spch_set_mesh_size_cutoff_w(orc_spch* s, double v) { s->s.mesh_w = v; }
void orc_spch_set_fp32_table(orc_spch* s, int on) { s->fp32_table = on != 0; }
void orc_spch_set_aspect_interp(orc_spch* s, double step, int order)
{
  s->interp_lnr = step > 0.0 ? orc_log(1.0 + step) : 0.0;
  s->interp_order = order;
}
/* Scheff::SetMeshSize, space_charge.cpp:61-73 */
void orc_spch_set_mesh_size(orc_spch* s, uint32_t nr, uint32_t nz)
{
  if (nr != s->s.nx || nz != s->s.nz) { s->s.nx = nr; s->s.nz = nz; spch_alloc(s); }
}
void orc_spch_get_state(const orc_spch* s, orc_spch_state* st) { *st = s->s; }
void orc_spch_get_tables(const orc_spch* s, double* fld1, double* fld2, int64_t* bins)
{
  uint32_t nr = s->s.nx, nz = s->s.nz;
  if (fld1) memcpy(fld1, s->fld1, (size_t)(nr + 1) * nr * nz * 2 * sizeof(double));
  if (fld2) memcpy(fld2, s->fld2, (size_t)(nz + 1) * (nr + 1) * 2 * sizeof(double));
  if (bins) memcpy(bins, s->bins, (size_t)nr * nz * sizeof(int64_t));
}

/* UpdateFldTbl1Kernel, space_charge_kernel.cu:38-100, including the image
 * rings of r_adj_bunch neighbouring bunches on each side, one RF period
 * (beta*gamma*lambda in the bunch frame) apart (:74-94). */
void orc_t_green_table(uint32_t nr, uint32_t nz, double dr_, double dz_, double r_freq,
    double r_beta, uint32_t r_num_part, double r_current_ov_mc2, uint32_t r_adj_bunch,
    double* fld1)
{
  double gm = 1.0 / sqrt(1.0 - r_beta * r_beta);
  double dr = dr_;
  double dzg = dz_ * gm;
  double c1 = 5721.67 * r_current_ov_mc2 / (r_freq * 1e6 * (double)r_num_part);
  #pragma omp parallel for collapse(3) schedule(static)
  for (uint32_t rpi = 0; rpi <= nr; ++rpi)
    for (uint32_t k = 0; k < nr; ++k)
      for (uint32_t t = 0; t < nz; ++t) {
        double rs = sqrt(0.5 * (k * k + (k + 1) * (k + 1))) * dr;
        double zs = 0.5 * dzg;
        double zp = (t + 1) * dzg;
        double rp = rpi * dr;
        size_t cnt = (size_t)rpi * nr * nz + (size_t)k * nz + t;
        double d = zp - zs;
        double d2 = d * d;
        double c = (rp - rs) * (rp - rs);
        double b = (rp + rs) * (rp + rs);
        double rsrp = rs * rp;
        double a = 4.0 * rsrp / (b + d2);
        double g = 1.0 - a;
        double h = orc_log(g);
        double ee = 1.0 + g * (0.4630106 - 0.2452740 * h + g * (0.1077857 - 0.04125321 * h));
        double ek = 1.38629436 - 0.5 * h + g * (0.1119697 - 0.1213486 * h + g *
                    (0.07253230 - 0.02887472 * h));
        double er1 = 0.0;
        a = sqrt(b + d2);
        double rs2mrp2 = (rs * rs - rp * rp);
        if (rp != 0.0) er1 = (ek - (rs2mrp2 + d2) * ee / (c + d2)) / (2.0 * rp * a);
        double ez1 = d * ee / (a * (c + d2));
        double wave_len = CLIGHT / r_freq;
        double pl = r_beta * gm * wave_len;
        for (int i = 1; i <= (int)r_adj_bunch; ++i) {
          double xi = i;
          for (int j = 0; j < 2; ++j) {
            double d1 = d - xi * pl;
            d2 = d1 * d1;
            a = 4.0 * rsrp / (b + d2);
            g = 1.0 - a;
            h = orc_log(g);
            ee = 1.0 + g * (0.4630106 - 0.2452740 * h + g * (0.1077857 - 0.04125321 * h));
            ek = 1.38629436 - 0.5 * h + g * (0.1119697 - 0.1213486 * h + g *
                 (0.07253230 - 0.02887472 * h));
            a = sqrt(b + d2);
            if (rp != 0.0) er1 += (ek - (rs2mrp2 + d2) * ee / (c + d2)) / (2.0 * rp * a);
            ez1 += d1 * ee / (a * (c + d2));
            xi = -xi;
          }
        }
        fld1[2 * cnt] = c1 * er1;
        fld1[2 * cnt + 1] = c1 * ez1;
      }
}

/* Field sum of UpdateFldTbl2Kernel(Double), space_charge_kernel.cu:230-568,
 * in the canonical order described at the top of this file.  The field
 * point (zp, rp) sees source bin (zs, rs) through the Green's entry at
 * dz = zs - zp (zs >= zp, E_z sign -1) or zp - zs - 1 (zs < zp). */
void orc_t_field(uint32_t nr, uint32_t nz, const int64_t* bins, const double* fld1, double* fld2)
{
  uint32_t fld_sz = nr * nz;
  uint32_t P = orc_t_nextpow2(nr);
  if (P > 64) P = 64;
  #pragma omp parallel for collapse(2) schedule(static)
  for (uint32_t zp = 0; zp <= nz; ++zp)
    for (uint32_t rp = 0; rp <= nr; ++rp) {
      double ax[64], ay[64];
      for (uint32_t p = 0; p < P; ++p) {
        double tx = 0.0, ty = 0.0;
        for (uint32_t rs = p; rs < nr; rs += P)
          for (uint32_t zs = 0; zs < nz; ++zs) {
            double q = (double)bins[zs * nr + rs] * 0x1p-36;
            uint32_t dz;
            double sign;
            if (zs >= zp) { dz = zs - zp; sign = -1.0; }
            else { dz = zp - zs - 1; sign = 1.0; }
            size_t fi = (size_t)dz + (size_t)rs * nz + (size_t)rp * fld_sz;
            tx = fma(q, fld1[2 * fi], tx);
            ty = fma(q, sign * fld1[2 * fi + 1], ty);
          }
        ax[p] = tx; ay[p] = ty;
      }
      for (uint32_t o = P / 2; o >= 1; o >>= 1)
        for (uint32_t l = 0; l < o; ++l) { ax[l] = ax[l] + ax[l + o]; ay[l] = ay[l] + ay[l + o]; }
      size_t f2 = (size_t)zp * (nr + 1) + rp;
      fld2[2 * f2] = ax[0];
      fld2[2 * f2 + 1] = ay[0];
    }
}

/* DistributeParticleKernel, space_charge_kernel.cu:122-224 (fixed-point) */
static void deposit(orc_spch* s, orc_beam* b)
{
  uint32_t nx = s->s.nx, nz = s->s.nz;
  memset(s->bins, 0, (size_t)nx * nz * sizeof(int64_t));
  s->s.num_on_mesh = 0;
  if (b->blck_size / 4 == 0) return;   /* launch <<<grid, blck/4>>> is empty */
  double r_beta = s->beta, r_freq = s->freq;
  uint32_t on_mesh = 0;
  #pragma omp parallel for reduction(+:on_mesh) schedule(static)
  for (uint32_t np = 0; np < b->n; ++np) {
    if (!(b->loss[np] == 0 && b->lloss[np] == 0)) continue;
    double x = b->x[np], y = b->y[np], z = b->phi_r[np];
    double phimc = b->st.phi_avg_r;
    double xbar = b->st.x_avg_good, ybar = b->st.y_avg_good;
    double epsq = b->st.x_sig_good / b->st.y_sig_good;
    double gm = 1.0 / sqrt(1.0 - r_beta * r_beta);
    double dr = s->s.dr;
    double dzg = s->s.dz * gm;
    double wave_len = CLIGHT / r_freq;
    double hl = nz * 0.5 * dzg;
    double c2 = r_beta * gm * wave_len / TWOPI;
    double zz, a, cc, rminsq, rmaxsq, bb, d;
    int i, jm1, i1, j1, k;
    double rsq_sh = (x - xbar) * (x - xbar) / epsq + (y - ybar) * (y - ybar) * epsq;
    i = orc_to_int(sqrt(rsq_sh) / dr + 1);
    if (i < 1 || i > (int)nx) continue;
    double ddz = z - phimc;
    for (int guard = 0; (ddz > PI || ddz < -PI) && guard < 4096; ++guard) {
      if (ddz > 0.0) ddz -= TWOPI;
      else ddz += TWOPI;
    }
    z = -c2 * ddz;
    if (!(fabs(z) < hl)) continue;
    zz = z + hl;
    jm1 = orc_to_int(zz / dzg + 1);
    i1 = i + 1;
    double drdr = dr * dr;
    if (rsq_sh < 0.5 * drdr * (double)(i * i + (i - 1) * (i - 1))) i1 = i - 1;
    if (i1 < 1) i1 = 1;
    if (i1 > (int)nx) i1 = (int)nx;
    j1 = jm1 + 1;
    if (zz < (jm1 - 1 + 0.5) * dzg) j1 = jm1 - 1;
    if (j1 < 1) j1 = 1;
    if (j1 > (int)nz) j1 = (int)nz;
    if (i1 == i) a = 1.0;
    else {
      double sqrtmp = sqrt(4.0 * rsq_sh / drdr - 1.0);
      rminsq = 0.25 * drdr * (sqrtmp - 1.0) * (sqrtmp - 1.0);
      rmaxsq = rminsq + sqrtmp * drdr;
      if (i1 < i) a = (rmaxsq - (i - 1) * (i - 1) * drdr) / (rmaxsq - rminsq);
      else a = ((i1 - 1) * (i1 - 1) * drdr - rminsq) / (rmaxsq - rminsq);
    }
    cc = 1.0;
    if (j1 != jm1) cc = (zz - (j1 - 1 + 0.5) * dzg) / ((jm1 - j1) * dzg);
    bb = 1.0 - a;
    d = 1.0 - cc;
    if (jm1 < 1 || jm1 > (int)nz) continue;
    /* integer adds: the same totals in any order */
    int64_t* bins = s->bins;
    k = (jm1 - 1) * (int)nx + i;
    #pragma omp atomic
    bins[k - 1] += (int64_t)(a * cc * 0x1p36);
    k += i1 - i;
    #pragma omp atomic
    bins[k - 1] += (int64_t)(bb * cc * 0x1p36);
    k = (j1 - 1) * (int)nx + i;
    #pragma omp atomic
    bins[k - 1] += (int64_t)(a * d * 0x1p36);
    k += i1 - i;
    #pragma omp atomic
    bins[k - 1] += (int64_t)(bb * d * 0x1p36);
    on_mesh += 1;
  }
  s->s.num_on_mesh = on_mesh;
}

/* Field factors r/d^3 and z/d^3 of the bunch as a point charge at
 * distance (r, z), plus its r_adj_bunch neighbours on each side at z +- k*pl
 * (the off-mesh branch of ApplySpchKickKernel, space_charge_kernel.cu:642-662). */
void orc_t_offmesh_field(double z, double r, double pl, uint32_t r_adj_bunch, double* rod3,
                         double* zod3)
{
  double d = sqrt(z * z + r * r);
  double invd3 = 1.0 / (d * d * d);
  double ro = r * invd3, zo = z * invd3;
  for (int nb = 1; nb <= (int)r_adj_bunch; ++nb) {
    int xi = nb;
    for (int lr = 1; lr <= 2; ++lr) {
      double sz = z + xi * pl;
      d = sqrt(sz * sz + r * r);
      invd3 = 1.0 / (d * d * d);
      ro += r * invd3;
      zo += sz * invd3;
      xi = -xi;
    }
  }
  *rod3 = ro;
  *zod3 = zo;
}

/* ApplySpchKickKernel, space_charge_kernel.cu:572-706; launch geometry of
 * KickBeamKernelCall space_charge.cu:230-238 (Q3: indices that are
 * multiples of grid*(blck/4) are never kicked) */
static void kick(orc_spch* s, orc_beam* b, double r_length, double r_ratio_r,
    double r_ratio_z, double r_ratio_q, double r_ratio_gm)
{
  uint32_t stride = b->grid_size * (b->blck_size / 4);
  if (stride == 0) return;
  uint32_t nx = s->s.nx, nz = s->s.nz;
  double r_mass = b->mass, r_current = b->current * s->frac, r_freq = s->freq, r_beta = s->beta;
  uint32_t r_num_part = b->n;
  #pragma omp parallel for schedule(static)
  for (uint32_t indx = 0; indx < b->n; ++indx) {
    if (indx % stride == 0) continue;
    double x = b->x[indx], y = b->y[indx], xp = b->xp[indx], yp = b->yp[indx];
    double phi = b->phi_r[indx];
    double w = b->w[indx];
    if (!(b->loss[indx] == 0 && b->lloss[indx] == 0)) continue;
    double phimc = b->st.phi_avg_r;
    double xbar = b->st.x_avg_good, ybar = b->st.y_avg_good;
    double epsq = b->st.x_sig_good / b->st.y_sig_good;
    double gmg = 1.0 / sqrt(1.0 - r_beta * r_beta);
    double dr = s->s.dr * r_ratio_r;
    double dzg = s->s.dz * gmg * r_ratio_z;
    double wave_len = CLIGHT / r_freq;
    double hl = nz * 0.5 * dzg;
    double c1 = 5721.67 * r_current / (r_freq * 1e6 * (double)r_num_part * r_mass);
    double c2 = r_beta * gmg * wave_len / TWOPI;
    double c3 = r_length / (r_beta * gmg);
    double xfac = 2.0 / (epsq + 1.0);
    double yfac = epsq * xfac;
    double bgx, bgy, bgzstar, gstar, xxbar, yybar, r, xovr, yovr, ddz, z, d, rod3, zod3,
           cbgr, cbgzs, a, bb, c, bgzf, er_ratio, ez_ratio;
    uint32_t i, j, l, m;
    double gm = w / r_mass + 1.0;
    double bgz = sqrt((gm - 1.0) * (gm + 1.0));
    bgx = bgz * xp;
    bgy = bgz * yp;
    bgzstar = gmg * (bgz - r_beta * gm);
    gstar = gmg * (gm - r_beta * bgz);
    (void)gstar;
    xxbar = x - xbar;
    yybar = y - ybar;
    r = sqrt(xxbar * xxbar / epsq + yybar * yybar * epsq);
    if (r == 0.0) r = 1.0e-8;
    xovr = xxbar * xfac / r;
    yovr = yybar * yfac / r;
    ddz = phi - phimc;
    for (int guard = 0; (ddz > PI || ddz < -PI) && guard < 4096; ++guard) {
      if (ddz > 0.0) ddz -= TWOPI;
      else ddz += TWOPI;
    }
    z = -c2 * ddz;
    if (r > nx * dr || fabs(z) > hl) {
      z = -c2 * (phi - phimc);
      orc_t_offmesh_field(z, r, r_beta * gmg * wave_len, s->adj_eff, &rod3, &zod3);
      double cbgcoef = c1 * c3 * PI * 0.5 * s->s.num_on_mesh;
      cbgr = cbgcoef * rod3;
      cbgzs = cbgcoef * zod3;
    } else {
      i = (uint32_t)(1 + orc_to_int(r / dr));
      a = r / dr - (i - 1.0);
      bb = 1.0 - a;
      j = (uint32_t)(1 + orc_to_int((z + hl) / dzg));
      c = (z + hl) / dzg - (j - 1.0);
      d = 1.0 - c;
      uint32_t tbl_sz = (nx + 1) * (nz + 1);
      l = (j - 1) * (nx + 1) + i;
      m = l + (nx + 1);
      if (m >= tbl_sz) m = tbl_sz - 1;
      if (l >= tbl_sz) l = tbl_sz - 1;
      er_ratio = r_ratio_q / (r_ratio_r * r_ratio_z * r_ratio_gm * r_ratio_gm);
      ez_ratio = r_ratio_q / (r_ratio_z * r_ratio_z);
      const double* t = s->fld2;
      cbgr = er_ratio * c3 * (d * (a * t[2 * l] + bb * t[2 * (l - 1)]) +
             c * (a * t[2 * m] + bb * t[2 * (m - 1)]));
      cbgzs = ez_ratio * c3 * (d * (a * t[2 * l + 1] + bb * t[2 * (l - 1) + 1]) +
              c * (a * t[2 * m + 1] + bb * t[2 * (m - 1) + 1]));
    }
    bgx += cbgr * xovr;
    bgy += cbgr * yovr;
    bgzstar += cbgzs;
    gstar = 1.0 + 0.5 * bgzstar * bgzstar;
    bgzf = gmg * (bgzstar + r_beta * gstar);
    if (bgzf <= 0.0) b->loss[indx] = 9999;
    else {
      b->xp[indx] = bgx / bgzf;
      b->yp[indx] = bgy / bgzf;
      b->w[indx] *= 1.0 + (gm + 1.0) * (bgzf - bgz) / (gm * bgz);
    }
  }
}

/* Test hook: the SCHEFF kick alone (ApplySpchKickKernel,
 * space_charge_kernel.cu:572-706) with a caller's field table fld2
 * ((nz+1)*(nr+1) double2, [zp][rp]) on a mesh of spacing dr, dz, in the
 * frame of velocity beta at frequency freq, all remesh ratios 1: the beam's
 * SCHEFF statistics and relative phases are recomputed first
 * (UpdateStatForSpaceCharge), then every particle is kicked. */
void orc_t_kick(orc_spch* s, orc_beam* b, const double* fld2, double dr, double dz, double beta,
                double freq, double length, uint32_t num_on_mesh)
{
  orc_beam_update_stat_for_space_charge(b);
  s->s.dr = dr;
  s->s.dz = dz;
  s->beta = beta;
  s->freq = freq;
  s->frac = 1.0;
  s->adj_eff = 0;
  s->s.num_on_mesh = num_on_mesh;
  memcpy(s->fld2, fld2, (size_t)(s->s.nz + 1) * (s->s.nx + 1) * 2 * sizeof(double));
  kick(s, b, length, 1.0, 1.0, 1.0, 1.0);
}

/* Largest (x - x_avg)^2 + (y - y_avg)^2 over particles with loss == 0
 * (MaxR2PhiReduceKernel, beam_kernel.cu:793-930): the comparison a < v
 * keeps a when v is NaN, as fmax does; a maximum is order-free. */
static double max_r2(const orc_beam* b)
{
  double r2_max = 0.0;
  double xa = b->st.x_avg, ya = b->st.y_avg;
  #pragma omp parallel for reduction(max:r2_max) schedule(static)
  for (uint32_t i = 0; i < b->n; ++i) {
    if (b->loss[i] != 0) continue;
    double tx = b->x[i] - xa, ty = b->y[i] - ya;
    double r2 = tx * tx + ty * ty;
    r2_max = r2_max < r2 ? r2 : r2_max;
  }
  return r2_max;
}

/* Scheff::Start, space_charge.cpp:82-139 */
int orc_spch_start(orc_spch* s, orc_beam* b, double r_length)
{
  orc_beam_update_stat_for_space_charge(b);
  uint32_t good_cnt = b->st.num_good;
  double sigphi = b->n < 2 ? 0.0 : b->st.phi_sig_good;
  double sigx = b->n < 2 ? 0.0 : b->st.x_sig_good;
  double sigy = b->n < 2 ? 0.0 : b->st.y_sig_good;
  double ke = b->st.w_avg_good;
  double gamma = (double)ke / b->mass + 1.0;
  double beta = sqrt(1.0 - 1.0 / (gamma * gamma));
  double lmbd = 299.792458 / b->freq;
  double sigz = sigphi * 0.5 / 3.14159265 * beta * lmbd;
  double sigr = sqrt(sigx * sigx + sigy * sigy);
  double cratior = sigr * sigz;
  double cratioz = sigz * sigz * gamma * gamma;
  double cratior_old = s->s.sigr_old * s->s.sigz_old;
  double cratioz_old = s->s.sigz_old * s->s.sigz_old * s->s.gm_old * s->s.gm_old;
  uint32_t adj_bunch;
  if (s->s.adj_bunch_w == 0.0 || ke < s->s.adj_bunch_w) adj_bunch = s->s.adj_bunch;
  else adj_bunch = 0;
  /* above the mesh cutoff energy: half the mesh in each direction and four
   * times the kick interval (space_charge.cpp:107-117); SetMeshSize
   * reallocates the tables zeroed (space_charge.cpp:61-73) */
  if (s->s.mesh_w != 0.0) {
    if (s->s.nx_init < 2 || s->s.nz_init < 2)
      return set_err("mesh_size_cutoff_w needs a mesh of at least 2 x 2");
    if (ke > s->s.mesh_w && (s->s.nx != s->s.nx_init / 2 || s->s.nz != s->s.nz_init / 2))
      orc_spch_set_mesh_size(s, s->s.nx_init / 2, s->s.nz_init / 2);
    else if (ke <= s->s.mesh_w && (s->s.nx != s->s.nx_init || s->s.nz != s->s.nz_init))
      orc_spch_set_mesh_size(s, s->s.nx_init, s->s.nz_init);
    if (ke > s->s.mesh_w && s->s.interval != s->s.interval_init * 4)
      s->s.interval = s->s.interval_init * 4;
    else if (ke <= s->s.mesh_w && s->s.interval != s->s.interval_init)
      s->s.interval = s->s.interval_init;
  }
  /* SetParameters, space_charge.cu:77-86 */
  s->freq = b->freq; s->beta = beta; s->frac = s->s.fraction; s->adj_eff = adj_bunch;
  s->s.kick_count += 1;
  if (cratior_old == 0.0 || cratioz_old == 0.0 ||
      fabs(cratior - cratior_old) / cratior_old > s->s.remesh_val ||
      fabs(cratioz - cratioz_old) / cratioz_old > s->s.remesh_val) {
    s->s.sigr_old = sigr;
    s->s.sigz_old = sigz;
    s->s.gm_old = gamma;
    s->s.good_cnt_old = good_cnt;
    s->s.remesh_count += 1;
    memset(s->fld2, 0, (size_t)(s->s.nz + 1) * (s->s.nx + 1) * 2 * sizeof(double));
    /* UpdateMeshKernelCall + UpdateScheffMeshSizeKernel, space_charge.cu:122-138,
     * space_charge_kernel.cu:18-32.  The reference compares the device
     * pointers x_sig_good > y_sig_good; x_sig_good is allocated first
     * (beam.cu:51,54), so the y branch is the one taken. */
    double bl = beta * CLIGHT / s->freq;
    if (adj_bunch != 0) {
      /* UpdateMaxRPhi (beam.cu:829-851, beam_kernel.cu:793-930,1378-1383):
       * largest r^2 about the beam's last all-particle centroid (x_avg,
       * y_avg) over loss == 0; the mesh spans it radially and one RF
       * period longitudinally (space_charge_kernel.cu:18-32 with r_coef 1) */
      s->s.dr = 1.0 * sqrt(max_r2(b)) / s->s.nx;
      s->s.dz = bl / s->s.nz;
    } else {
      s->s.dr = 3.0 * b->st.y_sig_good / s->s.nx;
      s->s.dz = 3.0 * bl * b->st.phi_sig_good / (s->s.nz * PI);
    }
    orc_t_green_table(s->s.nx, s->s.nz, s->s.dr, s->s.dz, s->freq, s->beta, b->n,
                      b->current * s->frac / b->mass, adj_bunch, s->fld1);
    deposit(s, b);
    orc_t_field(s->s.nx, s->s.nz, s->bins, s->fld1, s->fld2);
    if (s->interp_lnr > 0.0 && adj_bunch == 0) {
      const uint32_t nx = s->s.nx, nzz = s->s.nz;
      const size_t n1 = (size_t)(nx + 1) * nx * nzz * 2, n2 = (size_t)(nzz + 1) * (nx + 1) * 2;
      double* t1 = (double*)malloc(n1 * sizeof(double));
      double* f0 = (double*)malloc(n2 * sizeof(double));
      double* f1 = (double*)malloc(n2 * sizeof(double));
      const double gm = 1.0 / sqrt(1.0 - s->beta * s->beta);
      const double x = orc_log(s->s.dz * gm / s->s.dr) / s->interp_lnr;
      double* f2 = (double*)malloc(n2 * sizeof(double));
      if (s->interp_order == 2) {
        /* three classes around the nearest, Lagrange weights */
        const double c = floor(x + 0.5), t = x - c;
        const double w[3] = {0.5 * t * (t - 1.0), (1.0 - t) * (1.0 + t), 0.5 * t * (t + 1.0)};
        double* fs[3] = {f0, f1, f2};
        for (int h = 0; h < 3; ++h) {
          const double dzc = orc_exp((c + h - 1) * s->interp_lnr) * s->s.dr / gm;
          orc_t_green_table(nx, nzz, s->s.dr, dzc, s->freq, s->beta, b->n,
                            b->current * s->frac / b->mass, 0, t1);
          orc_t_field(nx, nzz, s->bins, t1, fs[h]);
        }
        for (size_t i = 0; i < n2; ++i) s->fld2[i] = w[0] * f0[i] + w[1] * f1[i] + w[2] * f2[i];
      } else {
        const double c = floor(x), t = x - c;
        for (int h = 0; h < 2; ++h) {
          const double dzc = orc_exp((c + h) * s->interp_lnr) * s->s.dr / gm;
          orc_t_green_table(nx, nzz, s->s.dr, dzc, s->freq, s->beta, b->n,
                            b->current * s->frac / b->mass, 0, t1);
          orc_t_field(nx, nzz, s->bins, t1, h ? f1 : f0);
        }
        for (size_t i = 0; i < n2; ++i) s->fld2[i] = (1.0 - t) * f0[i] + t * f1[i];
      }
      free(t1); free(f0); free(f1); free(f2);
    }
    if (s->fp32_table)
      for (size_t i = 0; i < (size_t)(s->s.nz + 1) * (s->s.nx + 1) * 2; ++i)
        s->fld2[i] = (double)(float)s->fld2[i];
    kick(s, b, r_length, 1.0, 1.0, 1.0, 1.0);
  } else {
    double rq = s->s.good_cnt_old ? (double)(good_cnt / s->s.good_cnt_old) : 0.0;
    kick(s, b, r_length, sigr / s->s.sigr_old, sigz / s->s.sigz_old, rq, gamma / s->s.gm_old);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Simulation engine (simulation_engine.cpp / simulation_engine.cu)    */
/* ------------------------------------------------------------------ */

struct orc_sim {
  orc_beam* b;
  orc_element* el;
  uint32_t n;
  orc_spch* sc;
  int sc_on;
  uint32_t cur_id, ccl_cell_num;
  double design_w, wavelen;
  int32_t prev_end;
};

/* Reset, simulation_engine.cu:101-111 */
void orc_sim_reset(orc_sim* m)
{
  m->cur_id = 0;
  m->ccl_cell_num = 0;
  m->design_w = m->b->design_w;
  m->wavelen = CLIGHT / m->b->freq;
}

orc_sim* orc_sim_create(orc_beam* b, const orc_element* elems, uint32_t n, orc_spch* s)
{
  orc_sim* m = (orc_sim*)calloc(1, sizeof(orc_sim));
  m->b = b;
  m->el = (orc_element*)malloc((n ? n : 1) * sizeof(orc_element));
  memcpy(m->el, elems, n * sizeof(orc_element));
  m->n = n;
  m->sc = s;
  m->sc_on = s != NULL;
  orc_sim_reset(m);
  m->prev_end = -1;   /* uint max stored in int, simulation_engine.cpp:52 */
  return m;
}
void orc_sim_destroy(orc_sim* m) { if (m) { free(m->el); free(m); } }
void orc_sim_set_space_charge(orc_sim* m, int on) { m->sc_on = on; }

static int spch_start(orc_sim* m, double len)
{
  if (!m->sc) return set_err("space charge is on but no SpaceCharge object was given");
  return orc_spch_start(m->sc, m->b, len);
}

static int is_type(const orc_sim* m, int64_t i, int type)
{
  return i >= 0 && (uint64_t)i < m->n && m->el[i].type == type;
}
static int is_gap(const orc_sim* m, int64_t i)
{
  return is_type(m, i, ORC_EL_GAP_DTL) || is_type(m, i, ORC_EL_GAP_CCL);
}

/* SimulateDrift, simulation_engine.cu:125-145 */
static int sim_drift(orc_sim* m, const orc_element* e)
{
  int nk = 1;
  double length = e->length;
  if (m->sc_on) {
    if (!m->sc) return set_err("space charge is on but no SpaceCharge object was given");
    nk = (int)ceil(length / m->sc->s.interval);
  }
  double h = 0.5 * length / nk;
  for (int k = 0; k < nk; ++k) {
    map_partial_drift(m->b, m->wavelen, h, e->aperture, m->cur_id);
    if (m->sc_on && spch_start(m, 2.0 * h)) return -1;
    map_partial_drift(m->b, m->wavelen, h, e->aperture, m->cur_id);
  }
  return 0;
}

/* SimulateQuad, simulation_engine.cu:147-161 (launch <<<grid, blck/2>>>) */
static int sim_quad(orc_sim* m, const orc_element* e)
{
  int run = m->b->blck_size / 2 != 0;
  if (run) map_half_quad(m->b, m->wavelen, e->length, e->aperture, e->gradient, m->cur_id);
  if (m->sc_on && spch_start(m, e->length)) return -1;
  if (run) map_half_quad(m->b, m->wavelen, e->length, e->aperture, e->gradient, m->cur_id);
  return 0;
}

/* PrepSimulateRFGap, simulation_engine.cu:163-175 */
static void prep_gap(orc_sim* m, const orc_element* e)
{
  if (m->b->freq != e->gap.frequency) {
    m->wavelen = CLIGHT / e->gap.frequency;
    change_frequency(m->b, e->gap.frequency);
  }
}

/* SimulateDTLRFGap, simulation_engine.cu:182-241 */
static int sim_dtl_gap(orc_sim* m, const orc_element* e)
{
  prep_gap(m, e);
  double phi_in = 0.0;
  int64_t id = m->cur_id;
  double ref = e->gap.phase_ref;
  if (id == 1 || (id > 1 && !is_type(m, id - 2, ORC_EL_GAP_DTL))) {
    if (is_type(m, id + 2, ORC_EL_GAP_DTL))
      phi_in = ref - 0.5 * (m->el[id + 2].gap.phase_ref - ref);
    else
      phi_in = ref;   /* only gap in tank (also used for a final gap) */
  } else if (id > 1) {
    phi_in = 0.5 * (m->el[id - 2].gap.phase_ref + ref);
  } else {
    return set_err("A quad is missing before the DTL gap");
  }
  double q1 = is_type(m, id - 1, ORC_EL_QUAD) ? m->el[id - 1].length : 0.0;
  double q2 = is_type(m, id + 1, ORC_EL_QUAD) ? m->el[id + 1].length : 0.0;
  int run = m->b->blck_size / 2 != 0;
  if (run) map_gap_first_half(m->b, m->design_w, phi_in, &e->gap, e->length, q1, q2, 0, &m->wavelen);
  if (m->sc_on && spch_start(m, e->length)) return -1;
  if (run) map_gap_second_half(m->b, m->wavelen, m->design_w, phi_in, &e->gap, e->length, 0, q1, q2, 0);
  m->design_w = e->gap.energy_out;
  m->b->design_w = m->design_w;
  return 0;
}

/* SimulateCCLRFGap, simulation_engine.cu:243-293 */
static int sim_ccl_gap(orc_sim* m, const orc_element* e)
{
  prep_gap(m, e);
  int64_t id = m->cur_id;
  if (id == 0 || !is_type(m, id - 1, ORC_EL_GAP_CCL)) m->ccl_cell_num = 0;
  double phi_in, phi_out, cur = e->gap.phase_ref;
  if (m->ccl_cell_num == 0) {
    if (!is_gap(m, id + 1)) return set_err("a CCL tank needs a gap after its first gap");
    double next = m->el[id + 1].gap.phase_ref;
    phi_in = cur - 0.5 * (next - cur);
    phi_out = 0.5 * (next + cur);
  } else if (!is_type(m, id + 1, ORC_EL_GAP_CCL)) {
    double prev = m->el[id - 1].gap.phase_ref;
    phi_in = 0.5 * (cur + prev);
    phi_out = cur + 0.5 * (cur - prev);
  } else {
    phi_in = 0.5 * (cur + m->el[id - 1].gap.phase_ref);
    phi_out = 0.5 * (cur + m->el[id + 1].gap.phase_ref);
  }
  map_gap_first_half(m->b, m->design_w, phi_in, &e->gap, e->length, 0.0, 0.0, 1, &m->wavelen);
  if (m->sc_on && spch_start(m, e->length)) return -1;
  map_gap_second_half(m->b, m->wavelen, m->design_w, phi_out, &e->gap, e->length,
                      m->ccl_cell_num++, 0.0, 0.0, 1);
  m->design_w = e->gap.energy_out;
  m->b->design_w = m->design_w;
  return 0;
}

static int visit(orc_sim* m, const orc_element* e)
{
  switch (e->type) {
    case ORC_EL_DRIFT: return sim_drift(m, e);
    case ORC_EL_QUAD: return sim_quad(m, e);
    case ORC_EL_GAP_DTL: return sim_dtl_gap(m, e);
    case ORC_EL_GAP_CCL: return sim_ccl_gap(m, e);
    case ORC_EL_APERTURE_C:   /* SimulateApertureCircular, simulation_engine.cu:307-316 */
      if (!e->aperture_in) return 0;
      orc_beam_update_loss(m->b);
      update_avg_xy(m->b);
      map_circular_aperture(m->b, e->aperture, m->cur_id);
      return 0;
    case ORC_EL_APERTURE_R:   /* simulation_engine.cu:318-329 */
      if (!e->aperture_in) return 0;
      orc_beam_update_loss(m->b);
      update_avg_xy(m->b);
      map_rect_aperture(m->b, e->aper_xl, e->aper_xr, e->aper_yt, e->aper_yb, m->cur_id);
      return 0;
    case ORC_EL_SPCH_COMP:    /* simulation_engine.cu:301-305 */
      if (m->sc_on && m->sc) m->sc->s.fraction = e->fraction;
      return 0;
    case ORC_EL_ROTATION: map_rotation(m->b, e->angle); return 0;
    case ORC_EL_STEERER:      /* simulation_engine.cu:337-344 */
      if (e->bl_h > 1e-10 || e->bl_v > 1e-10) map_steerer(m->b, e->bl_h, e->bl_v);
      return 0;
    case ORC_EL_DIAGNOSTICS: return 0;
    case ORC_EL_BUNCHER:      /* SimulateBuncher, simulation_engine.cu:346-358 */
      if (!e->on) return 0;
      m->wavelen = CLIGHT / e->frequency;
      map_buncher(m->b, m->design_w, e->frequency, e->voltage, e->phase, e->aperture,
                  e->frequency / m->b->freq, m->wavelen, m->cur_id);
      return 0;
    case ORC_EL_DIPOLE: {     /* SimulateDipole, simulation_engine.cu:360-374 */
      int run = m->b->blck_size / 4 != 0;   /* launches <<<grid, blck/4>>> */
      if (run) map_dipole_first_half(m->b, e, m->wavelen);
      if (m->sc_on && spch_start(m, e->rho * e->angle)) return -1;
      if (run) map_dipole_second_half(m->b, e, m->wavelen);
      return 0;
    }
    default: return set_err("unknown element type");
  }
}

/* SimulationEngine::Simulate, simulation_engine.cpp:73-95 */
int orc_sim_simulate(orc_sim* m, int32_t start, int32_t end)
{
  if (m->n == 0) return 0;
  int32_t start_index = start < 0 ? 0 : start;
  int32_t end_index = end < 0 ? (int32_t)m->n - 1 : end;
  if (end_index >= (int32_t)m->n) return set_err("end element out of range");
  if (start_index < m->prev_end) orc_sim_reset(m);
  m->prev_end = end_index;
  for (uint32_t i = 0; (int32_t)i <= end_index; ++i) {
    if ((int32_t)i >= start_index || m->el[i].type == ORC_EL_SPCH_COMP) {
      m->cur_id = i;
      if (visit(m, &m->el[i])) return -1;
    }
  }
  return 0;
}
