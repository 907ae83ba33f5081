// cup_capi.cpp -- the extern "C" surface of include/cup.h: context lifetime, field I/O,
// and the operator entry points, each launching the hand-written kernels on the
// context's stream. No exception crosses this boundary.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "cup_internal.h"

#include <algorithm>

namespace cup {

int set_error(Ctx& c, int code, const std::string& msg)
{
  c.err = msg;
  return code;
}

int check_hip(Ctx& c, hipError_t e, const char* what)
{
  if (e == hipSuccess) return CUP_OK;
  c.err = std::string(what) + ": " + hipGetErrorString(e);
  return CUP_E_DEVICE;
}

int ensure_partials(Ctx& c, size_t ndoubles)
{
  if (ndoubles <= c.partials_cap) return CUP_OK;
  hipStreamSynchronize(c.stream);
  if (c.partials) hipFree(c.partials);
  c.partials = nullptr;
  c.partials_cap = 0;
  if (check_hip(c, hipMalloc(&c.partials, ndoubles * sizeof(double)), "hipMalloc partials")) return CUP_E_NOMEM;
  c.partials_cap = ndoubles;
  return CUP_OK;
}

static void timing_release(Ctx& c);

static void release(Ctx& c)
{
  if (c.stream) hipStreamSynchronize(c.stream);
  timing_release(c);
  bodies_free(c);
  for (auto& p : c.f)
    if (p) { hipFree(p); p = nullptr; }
  if (c.ana) hipFree(c.ana);
  c.ana = nullptr;
  void** bufs[] = {&c.spec, &c.spec2, &c.spec3, &c.rbuf, &c.xface, &c.tw[0], &c.tw[1], &c.tw[2]};
  for (void** b : bufs)
    if (*b) { hipFree(*b); *b = nullptr; }
  for (auto& p : c.cos4)
    if (p) { hipFree(p); p = nullptr; }
  for (auto& row : c.r2r)
    for (auto& p : row)
      if (p) { hipFree(p); p = nullptr; }
  for (auto& p : c.mden)
    if (p) { hipFree(p); p = nullptr; }
  unbounded_free(c);
  if (c.partials) hipFree(c.partials);
  if (c.ke_part) hipFree(c.ke_part);
  c.ke_part = nullptr;
  if (c.low_part) hipFree(c.low_part);
  if (c.low_red) hipFree(c.low_red);
  if (c.low_coef) hipFree(c.low_coef);
  if (c.low_counter) hipFree(c.low_counter);
  if (c.low_dev) hipFree(c.low_dev);
  if (c.low_rows) hipFree(c.low_rows);
  if (c.adjf) hipFree(c.adjf);
  c.adjf = nullptr;
  if (c.body_ft) hipFree(c.body_ft);
  c.body_ft = nullptr;
  c.body_ft_n = 0;
  c.low_dev = nullptr;
  c.low_rows = nullptr;
  c.low_part = nullptr; c.low_red = nullptr; c.low_counter = nullptr;
  if (c.sc) hipFree(c.sc);
  if (c.sc_host) hipHostFree(c.sc_host);
  if (c.host_red) hipHostFree(c.host_red);
  c.partials = nullptr; c.low_coef = nullptr; c.sc = nullptr; c.sc_host = nullptr;
  c.host_red = nullptr;
  if (c.stream) hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

// Reads DevScalars back (synchronizes the stream).
static int read_scalars(Ctx& c)
{
  if (check_hip(c, hipMemcpyAsync(c.sc_host, c.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream),
                "read scalars"))
    return CUP_E_DEVICE;
  return check_hip(c, hipStreamSynchronize(c.stream), "synchronize");
}

// roctx range names of the fused step's stages, after the reference's profiler regions
// (AdvectionDiffusion.cpp:406, PressureRHS.cpp:285, PoissonSolverPeriodic.cpp:131-143,
// PressureProjection.cpp:115, Penalization.cpp:133)
static const char* const kStageName[] = {"AdvDiff Kernel", "PresRHS Kernel + FFTW r2c x", "Forcing reduce",
                                         "FFTW r2c y",     "FFTW solve z",                "FFTW c2r y",
                                         "FFTW c2r x",     "GradP",                       "Penalization",
                                         "Step finalize"};

void stage_mark(Ctx& c, int stage, bool end)
{
  if (end)
    roctxRangePop();
  else
    roctxRangePushA(kStageName[stage]);
  if (stage != c.timing_stage || c.timing_skip || c.timing_count >= c.timing_cap) return;
  hipEventRecord(c.tev[2 * c.timing_count + (end ? 1 : 0)], c.stream);
  if (end) c.timing_count++;
}

static void timing_release(Ctx& c)
{
  for (auto e : c.tev) hipEventDestroy(e);
  c.tev.clear();
  c.timing_stage = -1;
  c.timing_cap = c.timing_count = 0;
}

}  // namespace cup

using namespace cup;

#define CUP_GUARD(ctx)                                 \
  if (!(ctx)) return CUP_E_INVALID;                    \
  Ctx& c = (ctx)->c;                                   \
  if (hipSetDevice(c.dev) != hipSuccess) return set_error(c, CUP_E_DEVICE, "hipSetDevice");

// Message of this thread's last failed call that has no context to hold it (cup_create,
// cup_diagnostics_noalloc).
static thread_local std::string g_create_error;

static int create_error(int code, const std::string& msg)
{
  g_create_error = msg;
  return code;
}

int cup::free_error(int code, const std::string& msg) { return create_error(code, msg); }

extern "C" {

int cup_create(const cup_grid_desc* d, cup_ctx** out)
{
  if (!d || !out) return create_error(CUP_E_INVALID, "cup_create: null argument");
  *out = nullptr;
  for (int a = 0; a < 3; ++a)
    if (d->bpd[a] < 1) return create_error(CUP_E_INVALID, "cup_create: blocks per dimension must be >= 1");
  if (d->precision != 4 && d->precision != 8)
    return create_error(CUP_E_INVALID, "cup_create: precision must be 4 (FP32) or 8 (FP64)");
  if (d->decomposition != CUP_PLAIN && d->decomposition != CUP_FUSED)
    return create_error(CUP_E_INVALID, "cup_create: decomposition must be CUP_PLAIN or CUP_FUSED");
  if (!(d->maxextent > 0)) return create_error(CUP_E_INVALID, "cup_create: maxextent must be > 0");
  bool periodic = true;
  int nfree = 0;
  for (int a = 0; a < 3; ++a) {
    if (d->bc[a] != CUP_BC_PERIODIC && d->bc[a] != CUP_BC_OPEN && d->bc[a] != CUP_BC_WALL &&
        d->bc[a] != CUP_BC_FREESPACE)
      return create_error(CUP_E_INVALID,
                          "cup_create: boundary conditions must be CUP_BC_PERIODIC, _OPEN, _WALL or _FREESPACE");
    periodic = periodic && d->bc[a] == CUP_BC_PERIODIC;
    nfree += d->bc[a] == CUP_BC_FREESPACE;
  }
  if (nfree != 0 && nfree != 3)  // SimulationData.cpp:136-143
    return create_error(CUP_E_INVALID, "cup_create: either all or no boundary condition can be CUP_BC_FREESPACE");
  if (!periodic && d->decomposition != CUP_PLAIN)
    return create_error(CUP_E_INVALID, "cup_create: non-periodic boundaries take the CUP_PLAIN decomposition");
  cup_ctx* h = new (std::nothrow) cup_ctx();
  if (!h) return create_error(CUP_E_NOMEM, "cup_create: out of host memory");
  Ctx& c = h->c;
  c.dev = d->device;
  // freespace faces are open faces for every stencil (Definitions.h:355-373); they select
  // PoissonSolverUnbounded
  for (int a = 0; a < 3; ++a) c.bc[a] = d->bc[a] == CUP_BC_FREESPACE ? CUP_BC_OPEN : d->bc[a];
  c.unbounded = nfree == 3;
  c.precision = d->precision;
  c.esz = (size_t)d->precision;
  c.decomposition = d->decomposition;
  int maxN = 0;
  for (int a = 0; a < 3; ++a) {
    c.bpd[a] = d->bpd[a];
    c.N[a] = d->bpd[a] * BS;
    maxN = c.N[a] > maxN ? c.N[a] : maxN;
  }
  c.nb = c.bpd[0] * c.bpd[1] * c.bpd[2];
  c.ncell = (size_t)c.nb * BS3;
  c.maxextent = d->maxextent;
  c.h = d->maxextent / maxN;
  for (int a = 0; a < 3; ++a) c.extent[a] = ((double)c.N[a] / maxN) * d->maxextent;
  int rc = CUP_OK;
  do {
    if ((rc = check_hip(c, hipSetDevice(c.dev), "hipSetDevice"))) break;
    if ((rc = check_hip(c, hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate"))) break;
    for (int k = 0; k < 8 && !rc; ++k) {
      if ((rc = check_hip(c, hipMalloc(&c.f[k], c.ncell * c.esz), "hipMalloc field"))) break;
      rc = check_hip(c, hipMemsetAsync(c.f[k], 0, c.ncell * c.esz, c.stream), "hipMemset field");
    }
    if (rc) break;
    if ((rc = check_hip(c, hipMalloc(&c.spec, c.ncell * c.esz), "hipMalloc spectrum"))) break;
    if ((rc = check_hip(c, hipMalloc(&c.rbuf, c.ncell * c.esz), "hipMalloc rhs"))) break;
    if ((rc = check_hip(c, hipMalloc(&c.xface, (size_t)c.nb * 2 * 64 * c.esz), "hipMalloc x faces"))) break;
    if ((rc = check_hip(c, hipMalloc(&c.sc, sizeof(DevScalars)), "hipMalloc scalars"))) break;
    if ((rc = check_hip(c, hipMemsetAsync(c.sc, 0, sizeof(DevScalars), c.stream), "hipMemset scalars"))) break;
    if ((rc = check_hip(c, hipHostMalloc((void**)&c.sc_host, sizeof(DevScalars), hipHostMallocDefault),
                        "hipHostMalloc")))
      break;
    if ((rc = check_hip(c, hipHostMalloc((void**)&c.host_red, 64 * sizeof(double), hipHostMallocDefault),
                        "hipHostMalloc")))
      break;
    if ((rc = ensure_partials(c, 20 * (size_t)c.nb + 1024))) break;
    if ((rc = check_hip(c, hipMalloc(&c.ke_part, 4 * (size_t)c.nb * sizeof(double)), "hipMalloc ke"))) break;
    if ((rc = check_hip(c, hipMemsetAsync(c.ke_part, 0, 4 * (size_t)c.nb * sizeof(double), c.stream), "memset ke")))
      break;
    if ((rc = fft_setup(c))) break;
    if ((rc = spectral_setup(c))) break;
    rc = check_hip(c, hipStreamSynchronize(c.stream), "synchronize");
  } while (0);
  if (rc) {
    const std::string msg = c.err;
    release(c);
    delete h;
    return create_error(rc, msg);
  }
  *out = h;
  return CUP_OK;
}

int cup_destroy(cup_ctx* ctx)
{
  if (!ctx) return CUP_E_INVALID;
  hipSetDevice(ctx->c.dev);
  release(ctx->c);
  delete ctx;
  return CUP_OK;
}

const char* cup_last_error(const cup_ctx* ctx)
{
  if (ctx) return ctx->c.err.c_str();
  return g_create_error.empty() ? "null context" : g_create_error.c_str();
}

int cup_get_geometry(const cup_ctx* ctx, int cells[3], double* h, double extent[3])
{
  if (!ctx) return CUP_E_INVALID;
  const Ctx& c = ctx->c;
  for (int a = 0; a < 3; ++a) {
    if (cells) cells[a] = c.N[a];
    if (extent) extent[a] = c.extent[a];
  }
  if (h) *h = c.h;
  return CUP_OK;
}

int cup_synchronize(cup_ctx* ctx)
{
  CUP_GUARD(ctx);
  return check_hip(c, hipStreamSynchronize(c.stream), "synchronize");
}

int cup_stream(cup_ctx* ctx, void** s)
{
  CUP_GUARD(ctx);
  if (!s) return CUP_E_INVALID;
  *s = (void*)c.stream;
  return CUP_OK;
}

// Valid field id; the adjoint fields are allocated (zeroed) on first use.
static int field_ready(Ctx& c, int field, const char* what)
{
  if (field < 0 || field > CUP_ADJ_W) return set_error(c, CUP_E_INVALID, std::string(what) + ": bad field");
  if (c.f[field]) return CUP_OK;
  if (check_hip(c, hipMalloc(&c.f[field], c.ncell * c.esz), "hipMalloc adjoint field")) return CUP_E_NOMEM;
  return check_hip(c, hipMemsetAsync(c.f[field], 0, c.ncell * c.esz, c.stream), "hipMemset adjoint field");
}

int cup_set_field(cup_ctx* ctx, int field, const double* host)
{
  CUP_GUARD(ctx);
  if (!host) return set_error(c, CUP_E_INVALID, "set_field: null pointer");
  if (int rc = field_ready(c, field, "set_field")) return rc;
  hipStreamSynchronize(c.stream);
  hipError_t e;
  if (c.precision == 4) {
    std::vector<float> tmp(c.ncell);
    for (size_t i = 0; i < c.ncell; ++i) tmp[i] = (float)host[i];
    e = hipMemcpy(c.f[field], tmp.data(), c.ncell * sizeof(float), hipMemcpyHostToDevice);
  } else {
    e = hipMemcpy(c.f[field], host, c.ncell * sizeof(double), hipMemcpyHostToDevice);
  }
  return check_hip(c, e, "set_field");
}

int cup_get_field(cup_ctx* ctx, int field, double* host)
{
  CUP_GUARD(ctx);
  if (!host) return set_error(c, CUP_E_INVALID, "get_field: null pointer");
  if (int rc = field_ready(c, field, "get_field")) return rc;
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "synchronize")) return rc;
  hipError_t e;
  if (c.precision == 4) {
    std::vector<float> tmp(c.ncell);
    e = hipMemcpy(tmp.data(), c.f[field], c.ncell * sizeof(float), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < c.ncell; ++i) host[i] = tmp[i];
  } else {
    e = hipMemcpy(host, c.f[field], c.ncell * sizeof(double), hipMemcpyDeviceToHost);
  }
  return check_hip(c, e, "get_field");
}

int cup_field_device_ptr(cup_ctx* ctx, int field, void** dptr)
{
  CUP_GUARD(ctx);
  if (!dptr) return set_error(c, CUP_E_INVALID, "field_device_ptr: null pointer");
  if (int rc = field_ready(c, field, "field_device_ptr")) return rc;
  *dptr = c.f[field];
  return CUP_OK;
}

int cup_set_field_device(cup_ctx* ctx, int field, const void* src)
{
  CUP_GUARD(ctx);
  if (!src) return set_error(c, CUP_E_INVALID, "set_field_device: null pointer");
  if (int rc = field_ready(c, field, "set_field_device")) return rc;
  return check_hip(c, hipMemcpyAsync(c.f[field], src, c.ncell * c.esz, hipMemcpyDeviceToDevice, c.stream),
                   "set_field_device");
}

int cup_get_field_device(cup_ctx* ctx, int field, void* dst)
{
  CUP_GUARD(ctx);
  if (!dst) return set_error(c, CUP_E_INVALID, "get_field_device: null pointer");
  if (int rc = field_ready(c, field, "get_field_device")) return rc;
  return check_hip(c, hipMemcpyAsync(dst, c.f[field], c.ncell * c.esz, hipMemcpyDeviceToDevice, c.stream),
                   "get_field_device");
}

int cup_lab_shift(cup_ctx* ctx, int g, int dx, int dy, int dz)
{
  CUP_GUARD(ctx);
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "lab_shift: needs a fully periodic box");
  if (int rc = launch_lab_shift(c, g, dx, dy, dz)) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "lab_shift");
}

int cup_max_velocity(cup_ctx* ctx, const double uinf[3], double* umax)
{
  CUP_GUARD(ctx);
  if (!uinf || !umax) return set_error(c, CUP_E_INVALID, "max_velocity: null argument");
  if (int rc = launch_max_u(c, uinf)) return rc;
  if (int rc = launch_reduce(c, c.nb, 2, 1u, c.sc->red)) return rc;
  if (int rc = read_scalars(c)) return rc;
  *umax = c.sc_host->red[0];
  return CUP_OK;
}

int cup_calc_max_timestep(cup_ctx* ctx, double CFL, double nu, const double uinf[3], int step, int rampup,
                          double* dt_out, double* umax_out)
{
  CUP_GUARD(ctx);
  if (!(CFL > 0) || nu < 0) return set_error(c, CUP_E_INVALID, "calc_max_timestep: CFL must be > 0 and nu >= 0");
  double umax = 0;
  if (int rc = cup_max_velocity(ctx, uinf, &umax)) return rc;
  c.umax_measured = umax;  // sim.uMax_measured (Simulation.cpp:407)
  const double hMin = c.h;
  const double dtDif = nu > 0 ? hMin * hMin / nu : INFINITY;
  const double dtAdv = hMin / (umax + 1e-8);
  const double m = std::fmin(dtDif, dtAdv);
  double dt = CFL * m;
  if (step < rampup) {
    const double x = (step + 1.0) / rampup;
    const double rampCFL = std::exp(std::log(1e-3) * (1 - x) + std::log(CFL) * x);
    dt = rampCFL * m;
  }
  if (dt_out) *dt_out = dt;
  if (umax_out) *umax_out = umax;
  return launch_set_dt(c, dt);
}

int cup_advect_diffuse(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme)
{
  CUP_GUARD(ctx);
  if (!uinf || !(dt > 0)) return set_error(c, CUP_E_INVALID, "advect_diffuse: dt must be > 0");
  if (int rc = launch_set_dt(c, dt)) return rc;
  if (int rc = launch_advdiff(c, nu, uinf, scheme)) return rc;
  if (c.decomposition == CUP_PLAIN) {  // u = tmp = u*, as after the reference's copy
    if (int rc = c.ustar == CUP_U ? launch_copy_u_to_tmp(c) : launch_copy_tmp_to_u(c)) return rc;
    if (!c.periodic())
      if (int rc = launch_inflow_correction(c)) return rc;  // on u, AdvectionDiffusion.cpp:209-296
  }
  return CUP_OK;
}

int cup_add_uniform_velocity(cup_ctx* ctx, const double du[3])
{
  CUP_GUARD(ctx);
  if (!du) return set_error(c, CUP_E_INVALID, "add_uniform_velocity: null pointer");
  return launch_add_uniform(c, du);
}

int cup_pressure_rhs(cup_ctx* ctx, double dt)
{
  CUP_GUARD(ctx);
  if (!(dt > 0)) return set_error(c, CUP_E_INVALID, "pressure_rhs: dt must be > 0");
  if (int rc = launch_set_dt(c, dt)) return rc;
  if (c.decomposition == CUP_PLAIN) {
    if (int rc = launch_fill(c, c.rbuf, c.ncell, 0.0)) return rc;  // PoissonSolver::reset
    if (int rc = launch_rhs_plain(c)) return rc;
    if (c.body.present)
      if (int rc = launch_rhs_obstacle(c, c.rbuf)) return rc;
  } else if (c.body.present) {
    if (int rc = launch_rhs_obstacle(c, nullptr)) return rc;
  }
  return CUP_OK;
}

int cup_poisson_periodic(cup_ctx* ctx)
{
  CUP_GUARD(ctx);
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "poisson_periodic: non-periodic boundaries take cup_poisson_mixed");
  if (c.decomposition == CUP_PLAIN) {
    if (int rc = launch_poisson_plain(c)) return rc;
    return launch_copy(c, c.f[CUP_P], c.rbuf, c.ncell);  // PoissonSolver::_fftw2cub
  }
  return launch_poisson_fused(c);
}

int cup_poisson_mixed(cup_ctx* ctx)
{
  CUP_GUARD(ctx);
  if (c.decomposition != CUP_PLAIN)
    return set_error(c, CUP_E_STATE, "poisson_mixed: takes the CUP_PLAIN decomposition's RHS buffer");
  if (int rc = launch_poisson_mixed(c)) return rc;
  return launch_copy(c, c.f[CUP_P], c.rbuf, c.ncell);  // PoissonSolver::_fftw2cub
}

int cup_poisson_unbounded(cup_ctx* ctx)
{
  CUP_GUARD(ctx);
  if (c.decomposition != CUP_PLAIN)
    return set_error(c, CUP_E_STATE, "poisson_unbounded: takes the CUP_PLAIN decomposition's RHS buffer");
  return launch_poisson_unbounded(c);
}

int cup_pressure_project(cup_ctx* ctx, double dt)
{
  CUP_GUARD(ctx);
  if (!(dt > 0)) return set_error(c, CUP_E_INVALID, "pressure_project: dt must be > 0");
  if (int rc = launch_set_dt(c, dt)) return rc;
  if (c.decomposition == CUP_PLAIN) return launch_gradp_plain(c);
  const double zero[3] = {0, 0, 0};
  return launch_gradp_fused(c, zero, nullptr);
}

int cup_diagnostics(cup_ctx* ctx, double nu, double out[20])
{
  CUP_GUARD(ctx);
  if (!out) return set_error(c, CUP_E_INVALID, "diagnostics: null output");
  int nparts = 0;
  if (int rc = launch_diagnostics(c, nu, &nparts)) return rc;
  if (int rc = launch_reduce(c, nparts, 20, 0u, c.sc->red)) return rc;
  if (int rc = read_scalars(c)) return rc;
  for (int i = 0; i < 20; ++i) out[i] = c.sc_host->red[i];
  return CUP_OK;
}

int cup_check_nan(cup_ctx* ctx, long long counts[4])
{
  CUP_GUARD(ctx);
  if (!counts) return set_error(c, CUP_E_INVALID, "check_nan: null output");
  return launch_count_nan(c, counts);
}

int cup_hit_analysis(cup_ctx* ctx, double nu, double out[9], double* spectrum, int nbin)
{
  CUP_GUARD(ctx);
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "hit_analysis: needs a fully periodic box");
  const int maxN = std::max(c.N[0], std::max(c.N[1], c.N[2]));
  const int nb = maxN / 2 - 1;
  if (!out || !(nu > 0)) return set_error(c, CUP_E_INVALID, "hit_analysis: bad argument");
  if (nbin < nb || !spectrum)
    return set_error(c, CUP_E_INVALID, "hit_analysis: the spectrum needs max(N)/2 - 1 entries");
  double sums[3], grad[2];
  std::vector<double> E(nb);
  if (int rc = launch_hit_analysis(c, sums, E.data(), nb, grad)) return rc;
  // SpectralManipFFTW.cpp:191-206 (normalization by the FFT size) and Analysis.cpp:171-187
  const double norm = (double)c.N[0] * c.N[1] * c.N[2];
  for (int b = 0; b < nb; ++b) spectrum[b] = E[b] / (norm * norm);
  const double tke = sums[0] / (norm * norm);
  const double eps = sums[1] * 2 * nu / ((norm * 2 * c.h) * (norm * 2 * c.h));
  const double uprime = std::sqrt(2 * tke / 3);
  const double lambda = std::sqrt(15 * nu / eps) * uprime;
  const double gmean = grad[0] / norm;
  out[0] = tke;
  out[1] = eps;
  out[2] = uprime;
  out[3] = lambda;
  out[4] = uprime * lambda / nu;
  out[5] = sums[2] / sums[0];
  out[6] = out[5] / uprime;
  out[7] = gmean;
  out[8] = std::sqrt(grad[1] / norm - gmean * gmean);
  return CUP_OK;
}

int cup_ic_taylor_green(cup_ctx* ctx, double uMax)
{
  CUP_GUARD(ctx);
  if (int rc = launch_ic_taylor_green(c, uMax)) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "ic_taylor_green");
}

int cup_ic_hit(cup_ctx* ctx, double k0, double u_rms, uint64_t seed)
{
  CUP_GUARD(ctx);
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "ic_hit: needs a fully periodic box");
  if (!(k0 > 0) || !(u_rms > 0)) return set_error(c, CUP_E_INVALID, "ic_hit: k0 and u_rms must be > 0");
  if (int rc = launch_ic_hit(c, k0, nullptr, nullptr, 0, u_rms, seed)) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "ic_hit");
}

int cup_ic_hit_spectrum(cup_ctx* ctx, const double* K, const double* E, int n, double u_rms, uint64_t seed)
{
  CUP_GUARD(ctx);
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "ic_hit_spectrum: needs a fully periodic box");
  if (!K || !E || n < 2) return set_error(c, CUP_E_INVALID, "ic_hit_spectrum: the table needs two entries or more");
  for (int i = 1; i < n; ++i)
    if (!(K[i] > K[i - 1])) return set_error(c, CUP_E_INVALID, "ic_hit_spectrum: K must increase");
  if (int rc = launch_ic_hit(c, 0.0, K, E, n, u_rms, seed)) return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "ic_hit_spectrum");
}

int cup_spectral_forcing(cup_ctx* ctx, double dt, double sim_dt, double nu, double tke_target,
                         double injection_rate, double stats_out[5])
{
  CUP_GUARD(ctx);
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "spectral_forcing: needs a fully periodic box");
  if (!(dt > 0)) return set_error(c, CUP_E_INVALID, "spectral_forcing: dt must be > 0");
  double st[5];
  if (int rc = forcing_plain(c, dt, sim_dt, nu, tke_target, injection_rate, st)) return rc;
  if (stats_out)
    for (int i = 0; i < 5; ++i) stats_out[i] = st[i];
  return CUP_OK;
}

// The bodies other than body 0 are dropped (the single-body entry points replace every body).
static void drop_other_bodies(Ctx& c)
{
  for (size_t j = 0; j < c.more.size(); ++j) {
    BodyScope scope(c, (int)j + 1);
    body_free(c);
  }
  c.more.clear();
}

// A call that replaces every body by one and fails leaves no body: the buffers and merged tables
// are freed and the grid chi cleared, so no later call sees a body half replaced.
static int replace_failed(Ctx& c, int rc)
{
  const std::string msg = c.err;
  bodies_free(c);
  launch_fill(c, c.f[CUP_CHI], c.ncell, 0.0);
  hipStreamSynchronize(c.stream);
  c.err = msg;
  return rc;
}

int cup_chi_from_sphere(cup_ctx* ctx, const double center[3], double radius, double com_mass_out[4])
{
  CUP_GUARD(ctx);
  if (!center || !(radius > 0)) return set_error(c, CUP_E_INVALID, "chi_from_sphere: radius must be > 0");
  drop_other_bodies(c);
  if (int rc = body_create(c, center, radius, 0.0)) return replace_failed(c, rc);
  if (com_mass_out) {
    for (int a = 0; a < 3; ++a) com_mass_out[a] = c.body.com[a];
    com_mass_out[3] = c.body.mass;
  }
  return CUP_OK;
}

int cup_chi_from_hemisphere(cup_ctx* ctx, const double center[3], double radius, double com_mass_out[4])
{
  CUP_GUARD(ctx);
  if (!center || !(radius > 0)) return set_error(c, CUP_E_INVALID, "chi_from_hemisphere: radius must be > 0");
  drop_other_bodies(c);
  if (int rc = body_create(c, center, radius, 0.0, true)) return replace_failed(c, rc);
  if (com_mass_out) {
    for (int a = 0; a < 3; ++a) com_mass_out[a] = c.body.com[a];
    com_mass_out[3] = c.body.mass;
  }
  return CUP_OK;
}

int cup_chi_from_sphere_mollified(cup_ctx* ctx, const double center[3], double radius, double width,
                                  double com_mass_out[4])
{
  CUP_GUARD(ctx);
  if (!center || !(radius > 0)) return set_error(c, CUP_E_INVALID, "chi_from_sphere_mollified: radius must be > 0");
  if (!(width > 0 && width <= 2))
    return set_error(c, CUP_E_INVALID, "chi_from_sphere_mollified: the half-width must lie in (0, 2] cells");
  drop_other_bodies(c);
  if (int rc = body_create(c, center, radius, width)) return replace_failed(c, rc);
  if (com_mass_out) {
    for (int a = 0; a < 3; ++a) com_mass_out[a] = c.body.com[a];
    com_mass_out[3] = c.body.mass;
  }
  return CUP_OK;
}

int cup_body_count(cup_ctx* ctx, int* n)
{
  CUP_GUARD(ctx);
  if (!n) return set_error(c, CUP_E_INVALID, "body_count: null pointer");
  *n = body_count(c);
  return CUP_OK;
}

int cup_add_sphere(cup_ctx* ctx, const double center[3], double radius, int hemi, double com_mass_out[4], int* id)
{
  CUP_GUARD(ctx);
  if (!center || !(radius > 0)) return set_error(c, CUP_E_INVALID, "add_sphere: radius must be > 0");
  int k = 0;
  if (c.body.present) {
    c.more.emplace_back();
    k = (int)c.more.size();
  }
  int rc;
  {
    BodyScope scope(c, k);
    rc = body_create(c, center, radius, 0.0, hemi != 0);
    if (!rc && com_mass_out) {
      for (int a = 0; a < 3; ++a) com_mass_out[a] = c.body.com[a];
      com_mass_out[3] = c.body.mass;
    }
  }
  if (rc) {
    if (k > 0) c.more.pop_back();
    return rc;
  }
  if (id) *id = k;
  return CUP_OK;
}

static int body_arg(Ctx& c, int id, const char* what)
{
  if (id < 0 || id >= body_count(c)) return set_error(c, CUP_E_INVALID, std::string(what) + ": no such body");
  return CUP_OK;
}

int cup_move_body(cup_ctx* ctx, int id, const double center[3], double com_mass_out[4])
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "move_body")) return rc;
  if (!center) return set_error(c, CUP_E_INVALID, "move_body: null pointer");
  BodyScope scope(c, id);
  if (c.body.moll > 0) return set_error(c, CUP_E_STATE, "move_body: the mollified sphere is built by its own call");
  if (int rc = body_create(c, center, c.body.radius, 0.0, c.body.hemi)) return rc;
  if (com_mass_out) {
    for (int a = 0; a < 3; ++a) com_mass_out[a] = c.body.com[a];
    com_mass_out[3] = c.body.mass;
  }
  return CUP_OK;
}

int cup_body_obstacle_blocks(cup_ctx* ctx, int id, int* ids_out, int cap, int* n_out)
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "body_obstacle_blocks")) return rc;
  if (!n_out) return set_error(c, CUP_E_INVALID, "body_obstacle_blocks: null pointer");
  BodyScope scope(c, id);
  const int n = c.body.nob;
  *n_out = n;
  if (ids_out) {
    if (cap < n) return set_error(c, CUP_E_INVALID, "body_obstacle_blocks: cap below the number of obstacle blocks");
    std::copy(c.body.ids_host.begin(), c.body.ids_host.end(), ids_out);
  }
  return CUP_OK;
}

int cup_body_set_udef(cup_ctx* ctx, int id, const double* udef, int n)
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "body_set_udef")) return rc;
  if (!udef) return set_error(c, CUP_E_INVALID, "body_set_udef: null pointer");
  BodyScope scope(c, id);
  Body& B = c.body;
  if (n != B.nob) return set_error(c, CUP_E_INVALID, "body_set_udef: n must be the number of obstacle blocks");
  const size_t bytes = (size_t)n * 3 * BS3 * sizeof(double);
  if (!B.ob_udef && check_hip(c, hipMalloc(&B.ob_udef, bytes), "hipMalloc udef")) return CUP_E_NOMEM;
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "body_set_udef")) return rc;
  return check_hip(c, hipMemcpy(B.ob_udef, udef, bytes, hipMemcpyHostToDevice), "body_set_udef");
}

int cup_body_get_udef(cup_ctx* ctx, int id, double* udef, int n)
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "body_get_udef")) return rc;
  if (!udef) return set_error(c, CUP_E_INVALID, "body_get_udef: null pointer");
  BodyScope scope(c, id);
  Body& B = c.body;
  if (n != B.nob) return set_error(c, CUP_E_INVALID, "body_get_udef: n must be the number of obstacle blocks");
  const size_t cnt = (size_t)n * 3 * BS3;
  if (!B.ob_udef) {
    std::fill(udef, udef + cnt, 0.0);
    return CUP_OK;
  }
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "body_get_udef")) return rc;
  return check_hip(c, hipMemcpy(udef, B.ob_udef, cnt * sizeof(double), hipMemcpyDeviceToHost), "body_get_udef");
}

int cup_body_remove_udef_momenta(cup_ctx* ctx, int id, double out[13])
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "body_remove_udef_momenta")) return rc;
  if (!out) return set_error(c, CUP_E_INVALID, "body_remove_udef_momenta: null pointer");
  BodyScope scope(c, id);
  return launch_udef_remove(c, out);
}

int cup_penalize_body(cup_ctx* ctx, int id, double dt, double lambda, int implicit, const double trans_vel[3],
                      const double ang_vel[3], double ft_out[6])
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "penalize_body")) return rc;
  BodyScope scope(c, id);
  return cup_penalize(ctx, dt, lambda, implicit, trans_vel, ang_vel, ft_out);
}

int cup_surface_forces_body(cup_ctx* ctx, int id, double nu, const double trans_vel[3], const double ang_vel[3],
                            double out[22])
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "surface_forces_body")) return rc;
  BodyScope scope(c, id);
  return cup_surface_forces(ctx, nu, trans_vel, ang_vel, out);
}

int cup_obstacle_momenta_body(cup_ctx* ctx, int id, double dt, double lambda, int implicit, double out[29])
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "obstacle_momenta_body")) return rc;
  BodyScope scope(c, id);
  return cup_obstacle_momenta(ctx, dt, lambda, implicit, out);
}

int cup_penalize(cup_ctx* ctx, double dt, double lambda, int implicit, const double trans_vel[3],
                 const double ang_vel[3], double ft_out[6])
{
  CUP_GUARD(ctx);
  if (!c.body.present) return set_error(c, CUP_E_STATE, "penalize: no obstacle (call cup_chi_from_sphere)");
  if (!(dt > 0) || !trans_vel || !ang_vel) return set_error(c, CUP_E_INVALID, "penalize: bad argument");
  if (int rc = launch_set_dt(c, dt)) return rc;
  if (int rc = launch_penalize(c, lambda, implicit, trans_vel, ang_vel)) return rc;
  if (int rc = read_scalars(c)) return rc;
  if (ft_out)
    for (int i = 0; i < 6; ++i) ft_out[i] = c.sc_host->red[i];
  return CUP_OK;
}

int cup_obstacle_momenta(cup_ctx* ctx, double dt, double lambda, int implicit, double out[29])
{
  CUP_GUARD(ctx);
  if (!c.body.present) return set_error(c, CUP_E_STATE, "obstacle_momenta: no obstacle");
  if (!out || !(dt > 0) || !(lambda > 0)) return set_error(c, CUP_E_INVALID, "obstacle_momenta: bad argument");
  if (int rc = launch_obstacle_momenta(c, dt, lambda, implicit ? 1 : 0)) return rc;
  if (int rc = read_scalars(c)) return rc;
  for (int i = 0; i < 29; ++i) out[i] = c.sc_host->red[i];
  return CUP_OK;
}

int cup_surface_forces(cup_ctx* ctx, double nu, const double trans_vel[3], const double ang_vel[3], double out[22])
{
  CUP_GUARD(ctx);
  if (!c.body.present) return set_error(c, CUP_E_STATE, "surface_forces: no obstacle");
  if (!trans_vel || !ang_vel) return set_error(c, CUP_E_INVALID, "surface_forces: bad argument");
  if (int rc = launch_surface_forces(c, nu, trans_vel, ang_vel)) return rc;
  if (int rc = read_scalars(c)) return rc;
  if (out)
    for (int i = 0; i < 22; ++i) out[i] = c.sc_host->red[i];
  return CUP_OK;
}

// One step of the PLAIN decomposition, host-driven like Simulation::run (the reference's
// operator order, dt from the host each step).
static int plain_step(cup_ctx* ctx, const cup_step_params* p)
{
  Ctx& c = ctx->c;
  double dt = p->fixed_dt, umax = 0;
  if (!(dt > 0)) {
    if (int rc = cup_calc_max_timestep(ctx, p->CFL, p->nu, p->uinf, (int)c.sc_host->step, p->rampup, &dt, &umax))
      return rc;
  }
  if (int rc = cup_advect_diffuse(ctx, dt, p->nu, p->uinf, p->scheme)) return rc;
  if (p->spectral_forcing)
    if (int rc = cup_spectral_forcing(ctx, dt, dt, p->nu, p->tke_target, p->injection_rate, nullptr)) return rc;
  if (int rc = cup_pressure_rhs(ctx, dt)) return rc;
  if (int rc = c.periodic() ? cup_poisson_periodic(ctx) : cup_poisson_mixed(ctx)) return rc;
  if (int rc = cup_pressure_project(ctx, dt)) return rc;
  if (p->has_body && c.body.present) {
    if (int rc = cup_penalize(ctx, dt, p->lambda, p->implicit_penalization, p->body_vel, p->body_angvel, nullptr))
      return rc;
    for (int k = 1; k < body_count(c) && k < c.body_ft_n; ++k) {  // dt on the device from body 0's call
      BodyScope scope(c, k);
      if (int rc = launch_penalize(c, p->lambda, p->implicit_penalization, c.body.vel, c.body.vel + 3,
                                   c.body_ft + 6 * k))
        return rc;
    }
  }
  c.sc_host->step += 1;
  c.sc_host->time += dt;
  c.sc_host->dt = dt;
  c.sc_host->last_dt = dt;
  return check_hip(c, hipMemcpyAsync(c.sc, c.sc_host, sizeof(DevScalars), hipMemcpyHostToDevice, c.stream), "scalars");
}

int cup_run_steps(cup_ctx* ctx, const cup_step_params* p, int nsteps)
{
  CUP_GUARD(ctx);
  if (!p || nsteps < 0) return set_error(c, CUP_E_INVALID, "run_steps: bad argument");
  if (!(p->CFL > 0) && !(p->fixed_dt > 0)) return set_error(c, CUP_E_INVALID, "run_steps: need CFL > 0 or fixed_dt");
  if (p->has_body && !c.body.present) return set_error(c, CUP_E_STATE, "run_steps: has_body without obstacle");
  const int nbody = p->has_body ? body_count(c) : 0;
  if (nbody > 1 && c.body_ft_n < nbody) {  // the other bodies' forces of the last step
    if (c.body_ft) hipFree(c.body_ft);
    c.body_ft = nullptr;
    c.body_ft_n = 0;
    if (check_hip(c, hipMalloc(&c.body_ft, 6 * (size_t)nbody * sizeof(double)), "hipMalloc body forces"))
      return CUP_E_NOMEM;
    c.body_ft_n = nbody;
  }
  if (c.decomposition == CUP_PLAIN) {
    if (int rc = read_scalars(c)) return rc;
    for (int s = 0; s < nsteps; ++s)
      if (int rc = plain_step(ctx, p)) return rc;
    return check_hip(c, hipStreamSynchronize(c.stream), "run_steps");
  }
  // FUSED: measure the current velocity and set dt for the coming step on the device.
  if (int rc = launch_reset_umax(c)) return rc;
  c.umax_bank = 0;
  if (int rc = launch_max_u(c, p->uinf, true)) return rc;
  if (int rc = launch_finalize_step(c, p->CFL, p->nu, p->rampup, p->fixed_dt, 0)) return rc;
  // With the central scheme each step's bookkeeping (k_finalize_step) runs inside the next
  // step's advection; the ramp needs the device step counter, read once here.
  const bool fold_ok = p->scheme == CUP_CENTRAL_EULER;
  long long step0 = 0;
  if (fold_ok && p->rampup > 0) {
    if (int rc = read_scalars(c)) return rc;
    step0 = c.sc_host->step;
  }
  const bool body = c.body.present && p->has_body;
  // The band-limited forcing stays on the device when its low-mode set fits the fused
  // kernels; otherwise the reference decomposition of SpectralForcing runs (host round trip).
  // With a body the forced u* must be in place before the RHS masking reads it (the fused
  // forcing adds its modes after), so the reference decomposition runs then as well.
  const bool force_fused = p->spectral_forcing && c.low_fused_ok && !body;
  struct SkipReset {
    Ctx& c;
    ~SkipReset() { c.timing_skip = false; }
  } skip_reset{c};
  for (int s = 0; s < nsteps; ++s) {
    // The first step of the call is left out of the stage timing: its events follow an idle
    // stream (they would hold the host's launch latency) and it runs without the folded
    // bookkeeping of a previous step.
    c.timing_skip = s == 0;
    const bool fold = fold_ok && s > 0;
    const StepFold sf{umax_slots_dev(c, c.umax_bank), step0 + s, p->CFL, p->rampup, p->fixed_dt};
    stage_mark(c, ST_ADVDIFF, false);
    if (int rc = launch_advdiff(c, p->nu, p->uinf, p->scheme, fold ? &sf : nullptr)) return rc;
    stage_mark(c, ST_ADVDIFF, true);
    if (fold) c.umax_bank ^= 1;  // this step's gradient fills the other bank and clears the folded one
    if (p->spectral_forcing && !force_fused) {
      double st[5];
      if (int rc = forcing_plain(c, -1.0, -1.0, p->nu, p->tke_target, p->injection_rate, st)) return rc;
    }
    if (body)
      if (int rc = launch_rhs_obstacle(c, nullptr)) return rc;
    if (int rc = launch_poisson_fused(c, force_fused, p->tke_target, p->injection_rate, true)) return rc;
    stage_mark(c, ST_GRADP, false);
    if (int rc = launch_gradp_fused(c, p->uinf, body ? c.ob_slot_all : nullptr, force_fused, true)) return rc;
    stage_mark(c, ST_GRADP, true);
    if (body)
      if (int rc = launch_penalize_step(c, p->lambda, p->implicit_penalization, p->body_vel, p->body_angvel, p->uinf))
        return rc;
    for (int k = 1; k < nbody; ++k) {  // the other bodies, each with its cup_body_velocity
      BodyScope scope(c, k);
      if (int rc = launch_penalize_step(c, p->lambda, p->implicit_penalization, c.body.vel, c.body.vel + 3, p->uinf,
                                        c.body_ft + 6 * k))
        return rc;
    }
    if (!fold_ok || s == nsteps - 1)
      if (int rc = launch_finalize_step(c, p->CFL, p->nu, p->rampup, p->fixed_dt, 1)) return rc;
  }
  return CUP_OK;
}

int cup_stage_timing_begin(cup_ctx* ctx, int stage, int max_steps)
{
  CUP_GUARD(ctx);
  if (stage < 0 || stage > ST_FINAL || max_steps < 1) return set_error(c, CUP_E_INVALID, "stage timing: bad argument");
  hipStreamSynchronize(c.stream);
  timing_release(c);
  c.tev.resize(2 * (size_t)max_steps);
  for (auto& e : c.tev)
    if (int rc = check_hip(c, hipEventCreate(&e), "hipEventCreate")) return rc;
  c.timing_stage = stage;
  c.timing_cap = max_steps;
  c.timing_count = 0;
  c.timing_skip = false;
  return CUP_OK;
}

int cup_stage_timing_end(cup_ctx* ctx, double* avg_ms, int* count)
{
  CUP_GUARD(ctx);
  if (int rc = check_hip(c, hipStreamSynchronize(c.stream), "synchronize")) return rc;
  double tot = 0;
  for (int i = 0; i < c.timing_count; ++i) {
    float ms = 0;
    if (int rc = check_hip(c, hipEventElapsedTime(&ms, c.tev[2 * i], c.tev[2 * i + 1]), "hipEventElapsedTime"))
      return rc;
    tot += ms;
  }
  if (count) *count = c.timing_count;
  if (avg_ms) *avg_ms = c.timing_count ? tot / c.timing_count : 0.0;
  timing_release(c);
  return CUP_OK;
}

int cup_step_state(cup_ctx* ctx, int* step, double* time, double* last_dt, double* umax, double* tke)
{
  CUP_GUARD(ctx);
  if (c.decomposition == CUP_FUSED) {
    if (int rc = launch_reduce(c, 4 * c.nb, 1, 0u, &c.sc->ke, c.ke_part)) return rc;
  }
  if (int rc = read_scalars(c)) return rc;
  if (c.decomposition == CUP_FUSED) c.sc_host->ke *= 0.5 * c.h * c.h * c.h;
  if (step) *step = (int)c.sc_host->step;
  if (time) *time = c.sc_host->time;
  if (last_dt) *last_dt = c.sc_host->last_dt;
  if (umax) *umax = c.sc_host->umax;
  if (tke) *tke = c.sc_host->ke;
  return CUP_OK;
}

int cup_set_step_state(cup_ctx* ctx, int step, double time)
{
  CUP_GUARD(ctx);
  if (int rc = read_scalars(c)) return rc;
  c.sc_host->step = step;
  c.sc_host->time = time;
  if (int rc = check_hip(c, hipMemcpy(c.sc, c.sc_host, sizeof(DevScalars), hipMemcpyHostToDevice), "scalars"))
    return rc;
  return CUP_OK;
}

// u_n (the velocity fields) kept aside in the forcing adjoint's last three fields while the
// midpoint form computes u* in their place, and put back.
static int keep_velocity(Ctx& c, bool back)
{
  const size_t fb = c.ncell * c.esz;
  char* keep = (char*)c.adjf + 6 * fb;
  for (int q = 0; q < 3; ++q) {
    void* dst = back ? c.f[CUP_U + q] : (void*)(keep + q * fb);
    const void* src = back ? (const void*)(keep + q * fb) : c.f[CUP_U + q];
    if (int e = check_hip(c, hipMemcpyAsync(dst, src, fb, hipMemcpyDeviceToDevice, c.stream), "keep velocity"))
      return e;
  }
  return CUP_OK;
}

static int alloc_adjf(Ctx& c)
{
  if (!c.adjf && check_hip(c, hipMalloc(&c.adjf, 9 * c.ncell * c.esz), "hipMalloc forcing adjoint"))
    return CUP_E_NOMEM;
  return CUP_OK;
}

// The adjoint fields += the linear forcing's transpose at u* = A(u_n), the scheme's advection of
// the velocity fields, which are left as they were.
static int adjoint_forcing_at_ustar(Ctx& c, double dt, double nu, const double uinf[3], int scheme,
                                    double tke_target, double injection_rate)
{
  if (int e = alloc_adjf(c)) return e;
  const bool midpoint = scheme == CUP_UPWIND3_RK2;
  if (midpoint)
    if (int e = keep_velocity(c, false)) return e;
  const int saved = c.ustar;
  int e = launch_advdiff(c, nu, uinf, scheme);
  const int us0 = c.ustar;
  c.ustar = saved;
  if (!e) e = launch_adjoint_forcing(c, us0, dt, tke_target, injection_rate);
  if (midpoint) {  // u_n back in place, also after a failure
    const int r = keep_velocity(c, true);
    if (!e) e = r;
  }
  return e;
}

// lam_n = A'(u_n)^T lam* for the scheme's advection, into tmp; then tmp and the adjoint fields
// trade buffers.
static int adjoint_advection(Ctx& c, double nu, const double uinf[3], int scheme)
{
  if (scheme == CUP_CENTRAL_EULER) {
    if (int e = launch_adjoint_advdiff(c, nu, uinf)) return e;
  } else if (scheme == CUP_UPWIND3_EULER) {
    // lam_n = lam* + dt L'(u_n)^T lam*
    if (int e = launch_adjoint_up3(c, CUP_U, CUP_ADJ_U, CUP_ADJ_U, -1, CUP_TMPU, nu, uinf, 1.0)) return e;
  } else {
    // midpoint: u_h = u_n + dt/2 L(u_n); mu = dt L'(u_h)^T lam*; lam_n = lam* + mu + dt/2 L'(u_n)^T mu
    if (int e = launch_up3_stage(c, CUP_U, CUP_U, CUP_TMPU, nu, uinf, 0.5)) return e;
    if (int e = launch_adjoint_up3(c, CUP_TMPU, CUP_ADJ_U, -1, -1, F_ADJ2, nu, uinf, 1.0)) return e;
    if (int e = launch_adjoint_up3(c, CUP_U, F_ADJ2, CUP_ADJ_U, F_ADJ2, CUP_TMPU, nu, uinf, 0.5)) return e;
  }
  for (int q = 0; q < 3; ++q) std::swap(c.f[CUP_TMPU + q], c.f[CUP_ADJ_U + q]);
  return CUP_OK;
}

// Argument checks shared by the adjoint steps; allocates the midpoint form's scratch fields.
static int adjoint_scheme_ready(Ctx& c, int scheme, double tke_target, double injection_rate)
{
  if (tke_target < 0 || injection_rate < 0)
    return set_error(c, CUP_E_INVALID, "adjoint_step: tke_target and injection_rate must be >= 0");
  if (scheme != CUP_CENTRAL_EULER && scheme != CUP_UPWIND3_EULER && scheme != CUP_UPWIND3_RK2)
    return set_error(c, CUP_E_INVALID, "adjoint_step: unknown scheme");
  if (scheme == CUP_UPWIND3_RK2)
    for (int q = 0; q < 3; ++q)
      if (!c.f[F_ADJ2 + q] &&
          check_hip(c, hipMalloc(&c.f[F_ADJ2 + q], c.ncell * c.esz), "hipMalloc adjoint scratch"))
        return CUP_E_NOMEM;
  return CUP_OK;
}

int cup_adjoint_step(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme)
{
  return cup_adjoint_step_forcing(ctx, dt, nu, uinf, scheme, 0.0, 0.0);
}

int cup_adjoint_step_forcing(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme,
                             double tke_target, double injection_rate)
{
  CUP_GUARD(ctx);
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "adjoint_step: needs a fully periodic box");
  if (!uinf || !(dt > 0)) return set_error(c, CUP_E_INVALID, "adjoint_step: dt must be > 0");
  if (c.body.present) return set_error(c, CUP_E_STATE, "adjoint_step: obstacles are not supported");
  if (int rc = adjoint_scheme_ready(c, scheme, tke_target, injection_rate)) return rc;
  for (int f = CUP_ADJ_U; f <= CUP_ADJ_W; ++f)
    if (int rc = field_ready(c, f, "adjoint_step")) return rc;
  if (int rc = launch_set_dt(c, dt)) return rc;
  // lam* = P lam: the forward projection on the adjoint fields (P is symmetric)
  const int saved = c.ustar;
  c.ustar = CUP_ADJ_U;
  int rc = launch_poisson_fused(c, false, 0.0, 0.0);
  c.ustar = saved;
  if (rc) return rc;
  if (int e = launch_gradp_inplace(c, CUP_ADJ_U)) return e;
  if (tke_target > 0 || injection_rate > 0)
    if (int e = adjoint_forcing_at_ustar(c, dt, nu, uinf, scheme, tke_target, injection_rate)) return e;
  return adjoint_advection(c, nu, uinf, scheme);
}

int cup_adjoint_step_body(cup_ctx* ctx, double dt, double nu, const double uinf[3], double lambda, int implicit,
                          const double body_vel[3], double wdrag, double dJdV[3], double dJdc[3])
{
  const double zero[3] = {0, 0, 0};
  return cup_adjoint_step_body_full(ctx, dt, nu, uinf, CUP_CENTRAL_EULER, 0.0, 0.0, lambda, implicit, body_vel, zero,
                                    wdrag, dJdV, nullptr, dJdc);
}

// The adjoint step with bodies 0..nbody-1 (nbody == body_count, their obstacle blocks apart, so
// each body's transposes act on its own blocks): body k forced at vel[3k..], its drag weighted by
// wdrag[k]; dJdV[3k..] and dJdW[3k..] (may be NULL) set. dJdc: the centre of body 0 (nbody == 1).
static int adjoint_bodies(Ctx& c, double dt, double nu, const double uinf[3], int scheme, double tke_target,
                          double injection_rate, double lambda, int implicit, int nbody, const double* vel,
                          const double* wdrag, double* dJdV, double* dJdW, double* dJdc)
{
  for (int f = CUP_ADJ_U; f <= CUP_ADJ_W; ++f)
    if (int rc = field_ready(c, f, "adjoint_step_body")) return rc;
  auto args = [&](int k) {
    AdjBodyArgs a;
    a.dt = dt;
    a.lambda = lambda;
    a.implicit = implicit;
    a.wdrag = wdrag[k];
    for (int d = 0; d < 3; ++d) a.vel[d] = vel[3 * k + d];
    a.centre = dJdc != nullptr;
    return a;
  };
  const bool forced = tke_target > 0 || injection_rate > 0, midpoint = scheme == CUP_UPWIND3_RK2;
  if (int rc = launch_set_dt(c, dt)) return rc;
  // forward state at this step: u* = F(A(u_n)), the scheme's advection (the midpoint form writes
  // it over u_n, which is kept aside) and the forcing
  if (midpoint) {
    if (int rc = alloc_adjf(c)) return rc;
    if (int rc = keep_velocity(c, false)) return rc;
  }
  // from the forward state to lam* = P_M^T Pen^T lam; u_n is put back afterwards, also after a failure
  auto transposes = [&]() -> int {
    if (int rc = launch_advdiff(c, nu, uinf, scheme)) return rc;
    if (forced) {
      double st[5];
      if (int rc = forcing_plain(c, dt, dt, nu, tke_target, injection_rate, st, c.ustar)) return rc;
    }
    if (dJdc) {
      // the penalization's input P_M(u*) needs the forward pressure: masked and corrected RHS,
      // Poisson solve into P; then the chi sensitivity of the penalization and its drag
      const AdjBodyArgs a = args(0);
      if (int rc = launch_adjoint_body(c, ADJB_FWD_MASK, a)) return rc;
      if (int rc = launch_adjoint_body(c, ADJB_FWD_CORRECT, a)) return rc;
      if (int rc = launch_poisson_fused(c, false, 0.0, 0.0)) return rc;
      if (int rc = launch_adjoint_body(c, ADJB_CHI_PEN, a)) return rc;
    }
    // each body's penalization and drag weight; its dJ/dV and dJ/dW of this step
    for (int k = 0; k < nbody; ++k) {
      BodyScope scope(c, k);
      if (int rc = launch_adjoint_body(c, ADJB_PENALIZE, args(k))) return rc;
      if (int rc = read_scalars(c)) return rc;
      for (int d = 0; d < 3; ++d) {
        dJdV[3 * k + d] = c.sc_host->red[8 + d];
        if (dJdW) dJdW[3 * k + d] = c.sc_host->red[17 + d];
      }
    }
    // transposed projection: y = L^{-1} R lam' (no obstacle RHS), y <- M'^T y body by body (each
    // body's masked RHS of u* on its blocks, the linearisation point, formed first), lam* = lam' -
    // (dt/2h) grad y
    const int us0 = c.ustar;
    c.ustar = CUP_ADJ_U;
    c.plain_rhs = true;
    int rc = launch_poisson_fused(c, false, 0.0, 0.0);
    c.plain_rhs = false;
    c.ustar = us0;
    if (rc) return rc;
    for (int k = 0; k < nbody; ++k) {
      BodyScope scope(c, k);
      if (int e = launch_adjoint_body(c, ADJB_FWD_MASK, args(k))) return e;
      if (int e = launch_adjoint_body(c, ADJB_MASK, args(k))) return e;
    }
    if (int e = launch_gradp_inplace(c, CUP_ADJ_U)) return e;
    return CUP_OK;
  };
  int rc = transposes();
  if (midpoint) {
    const int r = keep_velocity(c, true);
    if (!rc) rc = r;
  }
  if (rc) return rc;
  // the forcing's transpose at A(u_n), then the advection's
  if (forced)
    if (int e = adjoint_forcing_at_ustar(c, dt, nu, uinf, scheme, tke_target, injection_rate)) return e;
  if (int e = adjoint_advection(c, nu, uinf, scheme)) return e;
  if (dJdc) {
    if (int e = read_scalars(c)) return e;
    for (int d = 0; d < 3; ++d) dJdc[d] = c.sc_host->red[11 + d] + c.sc_host->red[14 + d];
  }
  return CUP_OK;
}

// Checks shared by the adjoint steps with bodies.
static int adjoint_body_ready(Ctx& c, double dt, const double uinf[3], int scheme, double tke_target,
                              double injection_rate)
{
  if (!c.periodic()) return set_error(c, CUP_E_INVALID, "adjoint_step_body: needs a fully periodic box");
  if (!uinf || !(dt > 0)) return set_error(c, CUP_E_INVALID, "adjoint_step_body: bad argument");
  if (int rc = adjoint_scheme_ready(c, scheme, tke_target, injection_rate)) return rc;
  if (!c.body.present) return set_error(c, CUP_E_STATE, "adjoint_step_body: no obstacle");
  for (int k = 0; k < body_count(c); ++k) {
    BodyScope scope(c, k);
    if (c.body.ob_udef) return set_error(c, CUP_E_STATE, "adjoint_step_body: a body's udef is not differentiated");
  }
  return CUP_OK;
}

int cup_adjoint_step_body_full(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme,
                               double tke_target, double injection_rate, double lambda, int implicit,
                               const double body_vel[3], const double body_angvel[3], double wdrag, double dJdV[3],
                               double dJdW[3], double dJdc[3])
{
  CUP_GUARD(ctx);
  if (!body_vel || !body_angvel || !dJdV) return set_error(c, CUP_E_INVALID, "adjoint_step_body: bad argument");
  if (int rc = adjoint_body_ready(c, dt, uinf, scheme, tke_target, injection_rate)) return rc;
  if (dJdc && (body_angvel[0] != 0 || body_angvel[1] != 0 || body_angvel[2] != 0))
    return set_error(c, CUP_E_INVALID,
                     "adjoint_step_body: the centre gradient is formed for a body without rotation only");
  if (body_count(c) > 1)
    return set_error(c, CUP_E_STATE, "adjoint_step_body: one body only (cup_adjoint_step_bodies takes several)");
  if (dJdc && !(c.body.moll > 0))
    return set_error(c, CUP_E_STATE,
                     "adjoint_step_body: the centre gradient needs the mollified chi (cup_chi_from_sphere_mollified)");
  return adjoint_bodies(c, dt, nu, uinf, scheme, tke_target, injection_rate, lambda, implicit, 1, body_vel, &wdrag,
                        dJdV, dJdW, dJdc);
}

int cup_adjoint_step_bodies(cup_ctx* ctx, double dt, double nu, const double uinf[3], int scheme, double tke_target,
                            double injection_rate, double lambda, int implicit, int nbody, const double* body_vel,
                            const double* body_angvel, const double* wdrag, double* dJdV, double* dJdW)
{
  CUP_GUARD(ctx);
  if (!body_vel || !body_angvel || !wdrag || !dJdV)
    return set_error(c, CUP_E_INVALID, "adjoint_step_bodies: bad argument");
  if (int rc = adjoint_body_ready(c, dt, uinf, scheme, tke_target, injection_rate)) return rc;
  if (nbody != body_count(c))
    return set_error(c, CUP_E_INVALID, "adjoint_step_bodies: nbody must be the number of bodies");
  return adjoint_bodies(c, dt, nu, uinf, scheme, tke_target, injection_rate, lambda, implicit, nbody, body_vel, wdrag,
                        dJdV, dJdW, nullptr);
}

int cup_body_velocity(cup_ctx* ctx, int id, const double vel[3], const double angvel[3])
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "body_velocity")) return rc;
  if (!vel || !angvel) return set_error(c, CUP_E_INVALID, "body_velocity: null pointer");
  BodyScope scope(c, id);
  for (int d = 0; d < 3; ++d) {
    c.body.vel[d] = vel[d];
    c.body.vel[3 + d] = angvel[d];
  }
  return CUP_OK;
}

int cup_last_body_force_body(cup_ctx* ctx, int id, double out[6])
{
  CUP_GUARD(ctx);
  if (int rc = body_arg(c, id, "last_body_force_body")) return rc;
  if (!out) return set_error(c, CUP_E_INVALID, "last_body_force_body: null pointer");
  if (id == 0) return cup_last_body_force(ctx, out);
  if (id >= c.body_ft_n) return set_error(c, CUP_E_STATE, "last_body_force_body: no loop step with this body yet");
  if (int rc = check_hip(c, hipMemcpyAsync(out, c.body_ft + 6 * id, 6 * sizeof(double), hipMemcpyDeviceToHost,
                                           c.stream), "last_body_force_body"))
    return rc;
  return check_hip(c, hipStreamSynchronize(c.stream), "last_body_force_body");
}

int cup_last_body_force(cup_ctx* ctx, double out[6])
{
  CUP_GUARD(ctx);
  if (!out) return set_error(c, CUP_E_INVALID, "last_body_force: null pointer");
  if (!c.body.present) return set_error(c, CUP_E_STATE, "last_body_force: no obstacle");
  if (int rc = read_scalars(c)) return rc;
  for (int i = 0; i < 6; ++i) out[i] = c.sc_host->red[i];
  return CUP_OK;
}

}  // extern "C"
