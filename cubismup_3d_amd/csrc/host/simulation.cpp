// simulation.cpp -- the reference's Simulation / SimulationData / Operator classes for the
// periodic path, in host C++ over the C-ABI of include/cup.h, plus the cup_sim_* C entry
// points the Python bindings load.
#include "simulation.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <limits>
#include <cstring>
#include <sys/stat.h>

#include <rocprofiler-sdk-roctx/roctx.h>

namespace cubismup3d {

BCflag string2BCflag(const std::string& s)
{
  if (s == "periodic") return periodic;
  if (s == "dirichlet") return dirichlet;
  if (s == "wall") return wall;
  if (s == "freespace" || s == "unbounded") return freespace;
  throw std::invalid_argument("BC not recognized " + s);
}

// ------------------------------------------------------------------ arguments

static std::string strip_dashes(const std::string& k)
{
  size_t i = 0;
  while (i < k.size() && k[i] == '-') ++i;
  return k.substr(i);
}

ArgumentParser::ArgumentParser(int argc, const char* const* argv)
{
  for (int i = 0; i < argc; ++i) {
    const std::string a = argv[i];
    if (a.size() < 2 || a[0] != '-' || (a[1] >= '0' && a[1] <= '9') || a[1] == '.') continue;
    std::string v = "true";
    if (i + 1 < argc) {
      const std::string n = argv[i + 1];
      const bool nextIsKey = n.size() >= 2 && n[0] == '-' && !(n[1] >= '0' && n[1] <= '9') && n[1] != '.';
      if (!nextIsKey) {
        v = n;
        ++i;
      }
    }
    args_[strip_dashes(a)] = v;
  }
}

ArgumentParser::Value ArgumentParser::operator()(const std::string& key) const
{
  auto it = args_.find(strip_dashes(key));
  if (it == args_.end()) return Value(false, "");
  return Value(true, it->second);
}

int ArgumentParser::Value::asInt(int def) const { return present_ ? (int)std::stod(s_) : def; }
int ArgumentParser::Value::asInt() const
{
  if (!present_) throw std::invalid_argument("mandatory integer argument missing");
  return (int)std::stod(s_);
}
double ArgumentParser::Value::asDouble(double def) const { return present_ ? std::stod(s_) : def; }
double ArgumentParser::Value::asDouble() const
{
  if (!present_) throw std::invalid_argument("mandatory argument missing");
  return std::stod(s_);
}
std::string ArgumentParser::Value::asString(const std::string& def) const { return present_ ? s_ : def; }
bool ArgumentParser::Value::asBool(bool def) const
{
  if (!present_) return def;
  if (s_ == "true" || s_ == "1" || s_ == "yes") return true;
  if (s_ == "false" || s_ == "0" || s_ == "no") return false;
  return std::stod(s_) != 0;
}

// ---------------------------------------------------------------------- logger

std::stringstream& BufferedLogger::get_stream(const std::string& filename)
{
  auto it = files_.find(filename);
  if (it != files_.end()) {
    if (++it->second.requests_since_last_flush == AUTO_FLUSH_COUNT) flush_one(it->first, it->second);
    return it->second.stream;
  }
  return files_[filename].stream;
}

void BufferedLogger::flush_one(const std::string& name, Stream& s)
{
  std::ofstream f(dir_ + "/" + name, std::ios::app | std::ios::out);
  f << s.stream.rdbuf();
  s.stream.str("");
  s.stream.clear();
  s.requests_since_last_flush = 0;
}

void BufferedLogger::flush()
{
  for (auto& p : files_) flush_one(p.first, p.second);
}

// ------------------------------------------------------------- SimulationData

SimulationData::SimulationData() : logger(std::make_shared<BufferedLogger>()), profiler(std::make_shared<Profiler>()) {}

SimulationData::SimulationData(const SimulationData& o) { *this = o; }

SimulationData& SimulationData::operator=(const SimulationData& o)
{
  if (this == &o) return *this;
  step = o.step; nsteps = o.nsteps; time = o.time; endTime = o.endTime; dt = o.dt;
  bpdx = o.bpdx; bpdy = o.bpdy; bpdz = o.bpdz; maxextent = o.maxextent; extent = o.extent;
  bImplicitPenalization = o.bImplicitPenalization; hmin = o.hmin; hmax = o.hmax; hmean = o.hmean;
  uinf = o.uinf; nu = o.nu; CFL = o.CFL; lambda = o.lambda; DLM = o.DLM;
  initCond = o.initCond; spectralIC = o.spectralIC; k0 = o.k0; tke0 = o.tke0; icSeed = o.icSeed; icUrms = o.icUrms;
  uMax_forced = o.uMax_forced; uMax_measured = o.uMax_measured; spectralForcing = o.spectralForcing;
  bChannelFixedMassFlux = o.bChannelFixedMassFlux;
  turbKinEn_target = o.turbKinEn_target; enInjectionRate = o.enInjectionRate;
  freqDiagnostics = o.freqDiagnostics; rampup = o.rampup; verbose = o.verbose; muteAll = o.muteAll;
  saveFreq = o.saveFreq; saveTime = o.saveTime; nextSaveTime = o.nextSaveTime; bDump = o.bDump;
  b3Ddump = o.b3Ddump; bRestart = o.bRestart;
  analysis = o.analysis; timeAnalysis = o.timeAnalysis; nextAnalysisTime = o.nextAnalysisTime;
  freqAnalysis = o.freqAnalysis; grad_mean = o.grad_mean; grad_std = o.grad_std; nu_sgs = o.nu_sgs;
  cs2_avg = o.cs2_avg;
  path4serialization = o.path4serialization;
  BCx_flag = o.BCx_flag; BCy_flag = o.BCy_flag; BCz_flag = o.BCz_flag; bUseFourierBC = o.bUseFourierBC;
  bUseUnboundedBC = o.bUseUnboundedBC;
  precision = o.precision; decomposition = o.decomposition; scheme = o.scheme; device = o.device;
  obstacles.clear();
  for (const auto& s : o.obstacles) obstacles.push_back(std::make_shared<Sphere>(*s));
  bCheckNaN = o.bCheckNaN;
  logger = std::make_shared<BufferedLogger>(o.path4serialization);
  profiler = std::make_shared<Profiler>();
  ctx = nullptr;  // a copy never shares the device context
  return *this;
}

SimulationData::~SimulationData()
{
  pipeline.clear();
  if (ctx) cup_destroy(ctx);
  ctx = nullptr;
}

// SimulationData(MPI_Comm, ArgumentParser&) (SimulationData.cpp:32-172), periodic subset.
SimulationData::SimulationData(int argc, const char* const* argv) : SimulationData()
{
  ArgumentParser parser(argc, argv);
  bpdx = parser("-bpdx").asInt();
  bpdy = parser("-bpdy").asInt();
  bpdz = parser("-bpdz").asInt();
  extent[0] = parser("extentx").asDouble(1);
  extent[1] = parser("extenty").asDouble(0);
  extent[2] = parser("extentz").asDouble(0);
  nu = parser("-nu").asDouble();
  initCond = parser("-initCond").asString("zero");
  spectralIC = parser("-spectralIC").asString("");
  k0 = parser("-k0").asDouble(10.0);
  tke0 = parser("-tke0").asDouble(1.0);
  icSeed = (uint64_t)parser("-icSeed").asDouble(20260101);
  icUrms = parser("-icUrms").asDouble(0.0);  // 0: the target spectrum's energy
  turbKinEn_target = parser("-turbKinEn_target").asDouble(0);
  enInjectionRate = parser("-energyInjectionRate").asDouble(0);
  const bool hint = turbKinEn_target > 0 || enInjectionRate > 0;
  spectralForcing = parser("-spectralForcing").asBool(hint);
  if (turbKinEn_target > 0 && enInjectionRate > 0)
    throw std::invalid_argument("either constant energy injection rate or forcing to fixed energy target");
  freqDiagnostics = parser("-freqDiagnostics").asInt(100);
  b3Ddump = parser("-dump3D").asBool(true);
  {
    const int dumpFreq = (int)parser("-fdump").asDouble(0);
    const double dumpTime = parser("-tdump").asDouble(0.0);
    saveFreq = parser("-fsave").asInt(0);
    saveTime = parser("-tsave").asDouble(0.0);
    if (saveFreq <= 0 && dumpFreq > 0) saveFreq = dumpFreq;
    if (saveTime <= 0 && dumpTime > 0) saveTime = dumpTime;
  }
  bRestart = parser("-restart").asBool(false);
  analysis = parser("-analysis").asString("");
  timeAnalysis = parser("-tAnalysis").asDouble(0.0);
  freqAnalysis = parser("-fAnalysis").asInt(0);
  if (parser("-iterativePenalization").asBool(false))
    throw std::invalid_argument("-iterativePenalization is not supported");
  bImplicitPenalization = parser("-implicitPenalization").asBool(false);
  uMax_forced = parser("-uMax_forced").asDouble(0.0);
  bChannelFixedMassFlux = parser("-channelFixedMassFlux").asBool(false);
  lambda = parser("-lambda").asDouble(1e6);
  DLM = parser("-use-dlm").asDouble(0);
  CFL = parser("-CFL").asDouble(.1);
  uinf[0] = parser("-uinfx").asDouble(0.0);
  uinf[1] = parser("-uinfy").asDouble(0.0);
  uinf[2] = parser("-uinfz").asDouble(0.0);
  verbose = parser("-verbose").asBool(false);
  muteAll = parser("-muteAll").asBool(false);
  bCheckNaN = parser("-checkNaN").asBool(false);
  rampup = parser("-rampup").asInt(100);
  nsteps = parser("-nsteps").asInt(0);
  endTime = parser("-tend").asDouble(0);
  path4serialization = parser("-serialization").asString("./");
  const std::string BC_x = parser("-BC_x").asString("dirichlet");
  const std::string BC_y = parser("-BC_y").asString("dirichlet");
  const std::string BC_z = parser("-BC_z").asString("dirichlet");
  BCx_flag = string2BCflag(BC_x);
  BCy_flag = string2BCflag(BC_y);
  BCz_flag = string2BCflag(BC_z);
  bUseFourierBC = BCx_flag == periodic && BCy_flag == periodic && BCz_flag == periodic;
  const std::string prec = parser("-precision").asString("double");
  precision = (prec == "float" || prec == "single" || prec == "4" || prec == "fp32") ? 4 : 8;
  // FUSED (the device-resident step) takes periodic boxes; non-periodic ones default to PLAIN
  const std::string dec = parser("-decomposition").asString(precision == 4 && bUseFourierBC ? "fused" : "plain");
  decomposition = dec == "fused" ? CUP_FUSED : CUP_PLAIN;
  device = parser("-device").asInt(0);
  logger->set_dir(path4serialization);
  // Sphere from the factory idiom of launch/settingsSphere.sh:19 as flags.
  if (parser("-sphereL").exists()) {
    auto s = std::make_shared<Sphere>();
    s->radius = 0.5 * parser("-sphereL").asDouble();
    s->position[0] = parser("-xpos").asDouble();
    s->position[1] = parser("-ypos").asDouble(-1);
    s->position[2] = parser("-zpos").asDouble(-1);
    const bool forced = parser("-bForcedInSimFrame").asBool(false);
    const bool fix = parser("-bFixFrameOfRef").asBool(false);
    for (int d = 0; d < 3; ++d) {
      s->bForcedInSimFrame[d] = forced;
      s->bFixFrameOfRef[d] = fix;
    }
    // enforcedVelocity = -(xvel, yvel, zvel) (Obstacle.cpp:64)
    s->setVelocities({-parser("-xvel").asDouble(0.0), -parser("-yvel").asDouble(0.0), -parser("-zvel").asDouble(0.0)});
    // the ramped x velocity (Sphere.cpp:90-98): umax = xvel, tmax = T
    s->accel_decel = parser("-accel").asBool(false);
    if (s->accel_decel) {
      s->umax = parser("-xvel").asDouble(0.0);
      s->tmax = parser("-T").asDouble(1.0);
    }
    s->bHemi = parser("-hemisphere").asBool(false);
    obstacles.push_back(s);
  }
}

// SimulationData::_preprocessArguments (SimulationData.cpp:174-247), one rank.
void SimulationData::_preprocessArguments()
{
  if (bpdx < 1 || bpdy < 1 || bpdz < 1) throw std::invalid_argument("Invalid bpd");
  const double NFE[3] = {(double)bpdx * 8, (double)bpdy * 8, (double)bpdz * 8};
  const double maxbpd = std::max({NFE[0], NFE[1], NFE[2]});
  maxextent = std::max({extent[0], extent[1], extent[2]});
  if (extent[0] <= 0 || extent[1] <= 0 || extent[2] <= 0) {
    for (int d = 0; d < 3; ++d) extent[d] = (NFE[d] / maxbpd) * maxextent;
  } else {
    // all three extents given: the reference switches to a stretched grid, which the
    // periodic FFT path does not support unless the extents match the uniform spacing
    for (int d = 0; d < 3; ++d)
      if (std::fabs(extent[d] - (NFE[d] / maxbpd) * maxextent) > 1e-12 * maxextent)
        throw std::invalid_argument("stretched grids are not supported");
  }
  if (!(nu >= 0)) throw std::invalid_argument("nu must be >= 0");
  if (!(CFL > 0)) throw std::invalid_argument("CFL must be > 0");
  if (analysis == "channel") throw std::invalid_argument("channel analysis needs walls; this build has periodic boxes");
  if (bChannelFixedMassFlux)
    throw std::invalid_argument("the fixed mass flux of channel flows needs walls; this build has periodic boxes");
  if (analysis == "HIT" && (freqAnalysis > 0 || timeAnalysis > 0) && !(nu > 0))
    throw std::invalid_argument("HIT analysis needs nu > 0");
  // SimulationData.cpp:136-143: freespace faces select PoissonSolverUnbounded, on every axis
  if (BCx_flag == freespace || BCy_flag == freespace || BCz_flag == freespace) {
    if (!(BCx_flag == freespace && BCy_flag == freespace && BCz_flag == freespace))
      throw std::invalid_argument("either all or no BC can be freespace/unbounded");
    bUseUnboundedBC = true;
  }
  if (!bUseFourierBC) {  // the non-periodic path (PoissonSolverMixed)
    if (decomposition == CUP_FUSED)
      throw std::invalid_argument("non-periodic boundaries take -decomposition plain");
    if (spectralForcing || initCond == "HITurbulence" || analysis == "HIT")
      throw std::invalid_argument("spectral forcing, the HIT initial condition and the HIT analysis need a periodic box");
    if (scheme != CUP_CENTRAL_EULER)
      throw std::invalid_argument("the upwind schemes need a periodic box");
  }
  hmin = hmax = hmean = maxextent / maxbpd;
  for (auto& s : obstacles) {
    if (s->position[1] < 0) s->position[1] = extent[1] / 2;  // ObstacleArguments defaults
    if (s->position[2] < 0) s->position[2] = extent[2] / 2;
  }
}

void SimulationData::setCells(int nx, int ny, int nz)
{
  if (nx <= 0 || ny <= 0 || nz <= 0 || nx % 8 != 0 || ny % 8 != 0 || nz % 8 != 0)
    throw std::invalid_argument("Number of cells must be multiple of block size * number of processes.");
  bpdx = nx / 8;
  bpdy = ny / 8;
  bpdz = nz / 8;
}

// ------------------------------------------------------------------ operators

void Operator::check(int rc, const char* what)
{
  if (rc == CUP_OK) return;
  const std::string m = std::string(what) + ": " + cup_last_error(sim.ctx);
  if (rc == CUP_E_INVALID) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

// Operator::check's scan on the device when the run asks for it.
static void checkNaN(const SimulationData& sim, const std::string& infoText)
{
  if (!sim.bCheckNaN) return;
  long long n[4];
  if (cup_check_nan(sim.ctx, n) != CUP_OK) throw std::runtime_error(std::string("check: ") + cup_last_error(sim.ctx));
  if (n[0] || n[1] || n[2] || n[3]) {
    const std::string m = "GenericCoordinator::check isnan " + infoText;
    fflush(stderr);
    printf("%s\n", m.c_str());
    fflush(stdout);
    throw std::runtime_error(m);
  }
}

void Operator::check(const std::string& infoText) { checkNaN(sim, infoText); }

// ------------------------------------------------------------------ profiler

void Profiler::push_start(const std::string& name, cup_ctx* ctx)
{
  roctxRangePushA(name.c_str());
  const bool timed = timing && ctx;
  if (timed) cup_synchronize(ctx);
  stack_.push_back({timed ? name : std::string(), std::chrono::steady_clock::now()});
}

void Profiler::pop_stop(cup_ctx* ctx)
{
  if (stack_.empty()) return;
  const auto& top = stack_.back();
  if (!top.first.empty()) {  // a timed region
    if (ctx) cup_synchronize(ctx);
    Agent& a = agents_[top.first];
    a.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - top.second).count();
    a.count++;
  }
  stack_.pop_back();
  roctxRangePop();
}

void Profiler::printSummary(FILE* out) const
{
  double total = 0;
  for (const auto& kv : agents_) total += kv.second.seconds;
  std::vector<std::pair<std::string, Agent>> rows(agents_.begin(), agents_.end());
  std::sort(rows.begin(), rows.end(), [](const auto& a, const auto& b) { return a.second.seconds > b.second.seconds; });
  fprintf(out, "%-30s %12s %12s %7s %8s\n", "region", "total[ms]", "mean[ms]", "share", "count");
  for (const auto& r : rows)
    fprintf(out, "%-30s %12.3f %12.4f %6.1f%% %8ld\n", r.first.c_str(), 1e3 * r.second.seconds,
            1e3 * r.second.seconds / std::max(1L, r.second.count), total > 0 ? 100 * r.second.seconds / total : 0.0,
            r.second.count);
  fflush(out);
}

void Profiler::reset() { agents_.clear(); }

void SimulationData::startProfiler(const std::string& name) const
{
  profiler->timing = verbose;
  profiler->push_start(name, ctx);
}

void SimulationData::stopProfiler() const { profiler->pop_stop(ctx); }

void SimulationData::printResetProfiler()
{
  profiler->printSummary();
  profiler->reset();
}

namespace {
// One profiler region for the lifetime of the object, closed also when an operator throws.
struct Region {
  const SimulationData& sim;
  Region(const SimulationData& s, const char* name) : sim(s) { sim.startProfiler(name); }
  ~Region() { sim.stopProfiler(); }
};
}  // namespace

// InitialConditions::operator() (InitialConditions.cpp:164-230).
// SpectralIcGenerator::_generateTarget (SpectralIcGenerator.cpp:21-98): the target spectrum
// table of -spectralIC. "art" (and no value): E = 32 tke0/(3 k0) sqrt(2/pi) (k/k0)^4
// exp(-2 (k/k0)^2) at bin centres; "cbc": the Comte-Bellot and Corrsin measurements in the
// reference's units; "fromFit": HITstatistics::getTargetSpectrumFit (SpectralManip.cpp:26-46)
// for the injection rate and nu. The table values are data and fit constants of the reference.
void spectralTarget(const SimulationData& sim, std::vector<double>& K, std::vector<double>& E)
{
  const int maxN = 8 * std::max({sim.bpdx, sim.bpdy, sim.bpdz});
  const double maxL = sim.maxextent;
  const int nBins = (int)std::ceil(std::sqrt(3.0) * maxN / 2.0) + 1;
  if (sim.spectralIC == "" || sim.spectralIC == "art") {
    const double binSize = M_PI * std::sqrt(3.0) * maxN / (nBins * maxL);
    const double coef = 32 * sim.tke0 / (3 * sim.k0) * std::sqrt(2.0 / M_PI);
    K.resize(nBins);
    E.resize(nBins);
    for (int i = 0; i < nBins; ++i) {
      K[i] = (i + 0.5) * binSize;
      E[i] = coef * std::pow(K[i] / sim.k0, 4) * std::exp(-2 * std::pow(K[i] / sim.k0, 2));
    }
  } else if (sim.spectralIC == "cbc") {
    // wavenumbers in 1/cm and energies in cm^3/s^2, grid mesh M = 5.08 cm
    K = {0.20, 0.25, 0.30, 0.40, 0.50, 0.70, 1.00, 1.50, 2.00, 2.50, 3.00, 4.00, 6.00, 8.00, 10.00, 12.50, 15.00,
         17.50, 20.00};
    E = {129.0, 230.0, 322.0, 435.0, 457.0, 380.0, 270.0, 168.0, 120.0, 89.0, 70.3, 47.0, 24.7, 12.6, 7.42, 3.96, 2.33,
         1.34, 0.8};
    const double M = 5.08, u_ref = std::sqrt(1.5) * 22.2, L_ref = 10.8 * M;
    for (size_t i = 0; i < K.size(); ++i) {
      K[i] *= L_ref * 0.01 * 2 * M_PI;
      E[i] /= u_ref * u_ref * L_ref;
    }
  } else if (sim.spectralIC == "fromFit") {
    const double eps = sim.enInjectionRate, nu = sim.nu;
    if (!(eps > 0) || !(nu > 0)) throw std::invalid_argument("spectralIC fromFit needs -energyInjectionRate and -nu");
    const double Lint = 0.74885397 * std::pow(eps, -0.0233311) * std::pow(nu, 0.07192009);
    const double Lk = std::pow(eps, -0.25) * std::pow(nu, 0.75);
    const double C = 5.7, CI = 1e-3, CE = 0, BE = 4.9, P0 = 600;
    K.resize(nBins);
    E.resize(nBins);
    for (int i = 0; i < nBins; ++i) {
      K[i] = (i + 0.5) * 2 * M_PI / maxL;
      const double KI = K[i] * Lint, KE4 = std::pow(K[i] * Lk, 4);
      const double FL = std::pow(KI / (KI + CI), 5 / 3.0 + P0);
      const double FE = std::exp(-BE * (std::pow(KE4 + std::pow(CE, 4), 0.25) - CE));
      E[i] = C * std::pow(eps, 2 / 3.0) * std::pow(K[i], -5 / 3.0) * FL * FE;
    }
  } else {
    throw std::invalid_argument("spectralIC not supported: " + sim.spectralIC + " (art, cbc, fromFit)");
  }
}

void InitialConditions::operator()(double)
{
  if (sim.initCond == "zero") {
    for (int f = 0; f < 8; ++f) {
      std::vector<double> z((size_t)sim.bpdx * sim.bpdy * sim.bpdz * 512, 0.0);
      check(cup_set_field(sim.ctx, f, z.data()), "IC zero");
    }
  } else if (sim.initCond == "taylorGreen") {
    check(cup_ic_taylor_green(sim.ctx, sim.uMax_forced), "IC taylorGreen");
  } else if (sim.initCond == "HITurbulence") {
    // SpectralIcGenerator::run (SpectralIcGenerator.cpp:118-135) on the target of -spectralIC;
    // the field keeps the target's energy unless -icUrms asks for a u_rms
    std::vector<double> K, E;
    spectralTarget(sim, K, E);
    check(cup_ic_hit_spectrum(sim.ctx, K.data(), E.data(), (int)K.size(), sim.icUrms, sim.icSeed),
          "IC HITurbulence");
    // SpectralIcGenerator::run (SpectralIcGenerator.cpp:126-130): target = IC energy
    if (sim.spectralForcing && sim.turbKinEn_target <= 0 && sim.enInjectionRate <= 0) {
      double d[20];
      check(cup_diagnostics(sim.ctx, sim.nu, d), "IC energy");
      const double vol = sim.extent[0] * sim.extent[1] * sim.extent[2];
      sim.turbKinEn_target = d[18] / vol;
    }
  } else {
    throw std::invalid_argument("initCond not supported: " + sim.initCond);
  }
}

// ------------------------------------------------------------------- Sphere

void Sphere::setVelocities(const std::array<double, 3>& imposed)
{
  bool any = false;
  for (int d = 0; d < 3; ++d) {
    transVel_imposed[d] = bForcedInSimFrame[d] ? imposed[d] : 0.0;
    transVel[d] = transVel_imposed[d];
    any = any || bForcedInSimFrame[d];
  }
  bBlockRotation = {any, any, any};
  angVel = {0, 0, 0};
}

bool Sphere::constantVelocity() const
{
  if (accel_decel) return false;
  for (int d = 0; d < 3; ++d)
    if (!bForcedInSimFrame[d]) return false;
  return true;
}

bool Sphere::isStatic() const
{
  if (!constantVelocity()) return false;
  for (int d = 0; d < 3; ++d)
    if (!(bFixFrameOfRef[d] || transVel_imposed[d] == 0)) return false;
  return true;
}

void Sphere::update(double dt, const std::array<double, 3>& uinf)
{
  for (int d = 0; d < 3; ++d) {
    position[d] += dt * (transVel[d] + uinf[d]);
    absPos[d] += dt * transVel[d];
  }
  // the quaternion moves along the unit sphere by the rotation of dt * dq/dt
  const auto& Q = quaternion;
  const double w0 = angVel[0], w1 = angVel[1], w2 = angVel[2];
  const double DQ[4] = {0.5 * dt * (-w0 * Q[1] - w1 * Q[2] - w2 * Q[3]), 0.5 * dt * (w0 * Q[0] + w1 * Q[3] - w2 * Q[2]),
                        0.5 * dt * (-w0 * Q[3] + w1 * Q[0] + w2 * Q[1]), 0.5 * dt * (w0 * Q[2] - w1 * Q[1] + w2 * Q[0])};
  const double n = std::sqrt(DQ[0] * DQ[0] + DQ[1] * DQ[1] + DQ[2] * DQ[2] + DQ[3] * DQ[3]);
  if (n > std::numeric_limits<double>::epsilon()) {
    const double f = std::tan(n) / n;
    double D[4], m = 0;
    for (int i = 0; i < 4; ++i) {
      D[i] = Q[i] + f * DQ[i];
      m += D[i] * D[i];
    }
    m = 1 / std::sqrt(m);
    for (int i = 0; i < 4; ++i) quaternion[i] = D[i] * m;
  }
  _2Dangle = 2 * std::atan2(quaternion[3], quaternion[0]);
}

void Sphere::save(const std::string& filename, double time) const
{
  std::ofstream f(filename + ".txt");
  if (!f) throw IOError("Couldn't open " + filename + ".txt");
  f.setf(std::ios::scientific);
  f.precision(std::numeric_limits<double>::digits10 + 1);
  f << time << "\n";
  f << position[0] << "\t" << position[1] << "\t" << position[2] << "\n";
  f << absPos[0] << "\t" << absPos[1] << "\t" << absPos[2] << "\n";
  f << quaternion[0] << "\t" << quaternion[1] << "\t" << quaternion[2] << "\t" << quaternion[3] << "\n";
  f << transVel[0] << "\t" << transVel[1] << "\t" << transVel[2] << "\n";
  f << angVel[0] << "\t" << angVel[1] << "\t" << angVel[2] << "\n";
  f << _2Dangle << "\n";
}

bool Sphere::restart(const std::string& filename)
{
  std::ifstream f(filename + ".txt");
  if (!f.good()) {
    printf("Could not restart from file\n");
    return false;
  }
  double t = 0;
  f >> t;
  f >> position[0] >> position[1] >> position[2];
  f >> absPos[0] >> absPos[1] >> absPos[2];
  f >> quaternion[0] >> quaternion[1] >> quaternion[2] >> quaternion[3];
  f >> transVel[0] >> transVel[1] >> transVel[2];
  f >> angVel[0] >> angVel[1] >> angVel[2];
  f >> _2Dangle;
  if (!f) throw IOError("Error reading " + filename + ".txt");
  return true;
}

// Gaussian elimination with partial pivoting on a 6x6 system (the LU solve of the reference).
static void solve6(double A[6][6], double b[6], double x[6])
{
  for (int k = 0; k < 6; ++k) {
    int p = k;
    for (int i = k + 1; i < 6; ++i)
      if (std::fabs(A[i][k]) > std::fabs(A[p][k])) p = i;
    if (!(std::fabs(A[p][k]) > 0)) throw std::runtime_error("computeVelocities: singular momentum system");
    if (p != k) {
      for (int j = 0; j < 6; ++j) std::swap(A[k][j], A[p][j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < 6; ++i) {
      const double f = A[i][k] / A[k][k];
      for (int j = k; j < 6; ++j) A[i][j] -= f * A[k][j];
      b[i] -= f * b[k];
    }
  }
  for (int i = 5; i >= 0; --i) {
    double r = b[i];
    for (int j = i + 1; j < 6; ++j) r -= A[i][j] * x[j];
    x[i] = r / A[i][i];
  }
}

void Sphere::computeVelocities(const double M[29], bool implicit, double time)
{
  if (accel_decel) {  // Sphere::computeVelocities (Sphere.cpp:147-159)
    if (time < tmax) transVel_imposed[0] = umax * time / tmax;
    else if (time < 2 * tmax) transVel_imposed[0] = umax * (2 * tmax - time) / tmax;
    else transVel_imposed[0] = 0;
  }
  // the penalized momenta (KernelFinalizeObstacleVel, ObstaclesUpdate.cpp:172-187)
  const double pm = implicit ? M[13] : M[0];
  const double cm[3] = {implicit ? M[14] : 0.0, implicit ? M[15] : 0.0, implicit ? M[16] : 0.0};
  const double* J = implicit ? M + 17 : M + 7;
  const double* L = implicit ? M + 23 : M + 1;
  const double* Am = implicit ? M + 26 : M + 4;
  double A[6][6] = {{pm, 0, 0, 0, cm[2], -cm[1]},       {0, pm, 0, -cm[2], 0, cm[0]},
                    {0, 0, pm, cm[1], -cm[0], 0},       {0, -cm[2], cm[1], J[0], J[3], J[4]},
                    {cm[2], 0, -cm[0], J[3], J[1], J[5]}, {-cm[1], cm[0], 0, J[4], J[5], J[2]}};
  double b[6] = {L[0], L[1], L[2], Am[0], Am[1], Am[2]};
  penalM = pm;
  for (int d = 0; d < 3; ++d) {
    penalCM[d] = cm[d];
    penalLmom[d] = L[d];
    penalAmom[d] = Am[d];
  }
  for (int i = 0; i < 6; ++i) {
    penalJ[i] = J[i];
    this->J[i] = M[7 + i];
  }
  // a forced axis or a blocked rotation keeps only its diagonal (Obstacle.cpp:172-200)
  for (int d = 0; d < 3; ++d) {
    if (bForcedInSimFrame[d]) {
      for (int j = 0; j < 6; ++j)
        if (j != d) A[d][j] = 0;
      b[d] = pm * transVel_imposed[d];
    }
    if (bBlockRotation[d]) {
      for (int j = 0; j < 6; ++j)
        if (j != 3 + d) A[3 + d][j] = 0;
      b[3 + d] = 0;
    }
  }
  double x[6];
  solve6(A, b, x);
  for (int d = 0; d < 3; ++d) {
    transVel_computed[d] = x[d];
    angVel_computed[d] = x[3 + d];
    transVel[d] = bForcedInSimFrame[d] ? transVel_imposed[d] : x[d];
    angVel[d] = bBlockRotation[d] ? 0.0 : x[3 + d];
  }
}

// CreateObstacles (ObstaclesCreate.cpp:444-540) for the spheres: the frame velocity
// (ObstacleVector::updateUinf, ObstacleVector.h:60-74: minus the mean velocity of the bodies
// fixed to the frame, per axis), each body's update with the step's dt (Obstacle::update), then
// chi, surface points and centre of mass where a body moved. Body k of the device is
// sim.obstacles[k].
// ObstacleVector::updateUinf (ObstacleVector.h:60-74): per axis, minus the mean velocity of the
// bodies fixed to the frame on that axis (0 when none is).
static std::array<double, 3> frameVelocity(const std::vector<std::shared_ptr<Sphere>>& obstacles)
{
  std::array<double, 3> u = {0, 0, 0};
  int nfix[3] = {0, 0, 0};
  for (const auto& sp : obstacles)
    for (int d = 0; d < 3; ++d)
      if (sp->bFixFrameOfRef[d]) {
        ++nfix[d];
        u[d] -= sp->transVel[d];
      }
  for (int d = 0; d < 3; ++d)
    if (nfix[d]) u[d] /= nfix[d];
  return u;
}

void CreateObstacles::operator()(double dt)
{
  if (sim.obstacles.empty()) return;
  sim.uinf = frameVelocity(sim.obstacles);
  velHeader.resize(sim.obstacles.size(), 0);
  std::vector<int> moved;
  for (size_t k = 0; k < sim.obstacles.size(); ++k) {
    auto& s = *sim.obstacles[k];
    const auto before = s.position;
    s.update(dt, sim.uinf);
    if (s.position != before) moved.push_back((int)k);
    if (dt > 0 && !sim.muteAll) {  // Obstacle::_writeComputedVelToFile (Obstacle.cpp:491-515)
      std::stringstream& f = sim.logger->get_stream("computedVelocity_" + std::to_string(s.obstacleID) + ".dat");
      const std::string tab("\t");
      if (sim.step == 0 && !velHeader[k]) {
        velHeader[k] = 1;
        f << "step" << tab << "time" << tab << "CMx" << tab << "CMy" << tab << "CMz" << tab << "quat_0" << tab
          << "quat_1" << tab << "quat_2" << tab << "quat_3" << tab << "vel_x" << tab << "vel_y" << tab << "vel_z"
          << tab << "angvel_x" << tab << "angvel_y" << tab << "angvel_z" << tab << "mass" << tab << "J0" << tab
          << "J1" << tab << "J2" << tab << "J3" << tab << "J4" << tab << "J5" << std::endl;
      }
      f << sim.step << tab;
      f.setf(std::ios::scientific);
      f.precision(std::numeric_limits<float>::digits10 + 1);
      f << sim.time;
      for (double v : s.position) f << tab << v;
      for (double v : s.quaternion) f << tab << v;
      for (double v : s.transVel) f << tab << v;
      for (double v : s.angVel) f << tab << v;
      f << tab << s.mass;
      for (double v : s.J) f << tab << v;
      f << std::endl;
    }
  }
  if (!created) {
    build();
    return;
  }
  Region r(sim, "Obst CHI");  // ObstaclesCreate.cpp:466-493: the SDF, chi and centre of mass
  for (int k : moved) {
    auto& s = *sim.obstacles[k];
    double cm[4];
    check(cup_move_body(sim.ctx, k, s.position.data(), cm), "CreateObstacles");
    s.centerOfMass = {cm[0], cm[1], cm[2]};
    s.mass = cm[3];
  }
}

// chi, surface points and centre of mass of every body where it is (set-up, and after an IC)
void CreateObstacles::build()
{
  Region r(sim, "Obst CHI");
  for (size_t k = 0; k < sim.obstacles.size(); ++k) {
    auto& s = *sim.obstacles[k];
    double cm[4];
    if (k == 0) {  // replaces every body of the device
      check(s.bHemi ? cup_chi_from_hemisphere(sim.ctx, s.position.data(), s.radius, cm)
                    : cup_chi_from_sphere(sim.ctx, s.position.data(), s.radius, cm),
            "CreateObstacles");
    } else {
      int id = -1;
      check(cup_add_sphere(sim.ctx, s.position.data(), s.radius, s.bHemi ? 1 : 0, cm, &id), "CreateObstacles");
    }
    s.centerOfMass = {cm[0], cm[1], cm[2]};
    s.mass = cm[3];
  }
  created = true;
}

void ExternalForcing::operator()(double dt)
{
  const int dir = sim.BCy_flag == wall ? 1 : 2;
  const double H = sim.extent[dir];
  const double du[3] = {8 * sim.uMax_forced * sim.nu / H / H * dt, 0, 0};
  Region r(sim, "Forcing Kernel");
  check(cup_add_uniform_velocity(sim.ctx, du), "ExternalForcing");
}

// The kernel and the copy back to u (AdvectionDiffusion.cpp:406-413) are one call here.
void AdvectionDiffusion::operator()(double dt)
{
  Region r(sim, "AdvDiff Kernel");
  check(cup_advect_diffuse(sim.ctx, dt, sim.nu, sim.uinf.data(), sim.scheme), "AdvectionDiffusion");
}

// SpectralForcing::operator() (SpectralForcing.cpp:23-82) with its forcingData.dat rows.
void SpectralForcing::operator()(double dt)
{
  double st[5];
  Region r(sim, "SpectralForcing");
  check(cup_spectral_forcing(sim.ctx, dt, sim.dt, sim.nu, sim.turbKinEn_target, sim.enInjectionRate, st),
        "SpectralForcing");
  const double totalKinEn = st[0], viscousDissip = st[1], largeModesKinEn = st[2];
  sim.dissipationRate = (totalKinEnPrev - totalKinEn) / dt;
  sim.actualInjectionRate = 0;
  if (sim.turbKinEn_target > 0) sim.actualInjectionRate = (sim.turbKinEn_target - totalKinEn) / dt;
  else if (sim.enInjectionRate > 0) sim.actualInjectionRate = sim.enInjectionRate;
  if (!sim.muteAll) {
    std::stringstream& ssF = sim.logger->get_stream("forcingData.dat");
    const std::string tab("\t");
    if (sim.step == 0)
      ssF << "step \t time \t dt \t totalKinEn \t largeModesKinEn \t viscousDissip \t totalDissipRate \t "
             "injectionRate \t lIntegral\n";
    ssF << sim.step << tab;
    ssF.setf(std::ios::scientific);
    ssF.precision(std::numeric_limits<float>::digits10 + 1);
    ssF << sim.time << tab << sim.dt << tab << totalKinEn << tab << largeModesKinEn << tab << viscousDissip << tab
        << sim.dissipationRate << tab << sim.actualInjectionRate << tab << st[3] << "\n";
  }
  totalKinEnPrev = st[4] > 0 ? totalKinEn + dt * sim.actualInjectionRate : totalKinEn;
}

// The RHS and its obstacle correction (PressureRHS.cpp:285-345) are one call here.
void PressureRHS::operator()(double dt)
{
  {
    Region r(sim, "PresRHS Kernel");
    check(cup_pressure_rhs(sim.ctx, dt), "PressureRHS");
  }
  check("PressureRHS");
}

// PressureProjection::operator() (PressureProjection.cpp:107-129).
void PressureProjection::operator()(double dt)
{
  // PressureProjection's solver (PressureProjection.cpp:86-105): the FFT solver when every face
  // is periodic (bUseFourierBC), PoissonSolverUnbounded with freespace faces (bUseUnboundedBC),
  // PoissonSolverMixed otherwise
  {
    Region r(sim, sim.bUseFourierBC ? "FFTW solve" : (sim.bUseUnboundedBC ? "UFFTW solve" : "MFFTW solve"));
    check(sim.bUseFourierBC ? cup_poisson_periodic(sim.ctx)
                            : (sim.bUseUnboundedBC ? cup_poisson_unbounded(sim.ctx) : cup_poisson_mixed(sim.ctx)),
          "PressureProjection solve");
  }
  {
    Region r(sim, "GradP");
    check(cup_pressure_project(sim.ctx, dt), "PressureProjection gradient");
  }
  check("PressureProjection");
}

// UpdateObstacles (ObstaclesUpdate.cpp:202-236): the body's momentum integrals on the
// projected velocity, then Obstacle::computeVelocities. A static body keeps its velocities
// (every axis forced, rotation blocked), so without log files its step skips the integrals.
void UpdateObstacles::operator()(double dt)
{
  for (size_t k = 0; k < sim.obstacles.size(); ++k) {
    auto& s = *sim.obstacles[k];
    if (s.isStatic() && sim.muteAll) continue;
    double M[29];
    {
      Region r(sim, "Obst Int Vel");
      check(cup_obstacle_momenta_body(sim.ctx, (int)k, dt, sim.lambda, sim.bImplicitPenalization ? 1 : 0, M),
            "UpdateObstacles");
    }
    Region r(sim, "Obst Upd Vel");
    s.computeVelocities(M, sim.bImplicitPenalization, sim.time);
  }
}

// Penalization::operator() (Penalization.cpp:129-156), body by body.
void Penalization::operator()(double dt)
{
  Region r(sim, "Penalization");
  for (size_t k = 0; k < sim.obstacles.size(); ++k) {
    auto& s = *sim.obstacles[k];
    double ft[6];
    check(cup_penalize_body(sim.ctx, (int)k, dt, sim.lambda, sim.bImplicitPenalization ? 1 : 0, s.transVel.data(),
                            s.angVel.data(), ft),
          "Penalization");
    s.force = {ft[0], ft[1], ft[2]};
    s.torque = {ft[3], ft[4], ft[5]};
  }
}

// ComputeForces (FluidSolidForces.cpp:205-223) and Obstacle::_writeSurfForcesToFile
// (Obstacle.cpp:518-545).
void ComputeForces::operator()(double)
{
  Region r(sim, "Obst. Forces");
  for (size_t k = 0; k < sim.obstacles.size(); ++k) write((int)k);
}

void ComputeForces::write(int k)
{
  auto& s = *sim.obstacles[k];
  double q[22];
  check(cup_surface_forces_body(sim.ctx, k, sim.nu, s.transVel.data(), s.angVel.data(), q), "ComputeForces");
  for (int i = 0; i < 22; ++i) s.surfaceQoI[i] = q[i];
  if (sim.muteAll) return;
  std::stringstream& ssF = sim.logger->get_stream("forceValues_surface_" + std::to_string(s.obstacleID) + ".dat");
  const std::string tab("\t");
  if (sim.step == 0)
    ssF << "step" << tab << "time" << tab << "mass" << tab << "force_x" << tab << "force_y" << tab << "force_z" << tab
        << "torque_x" << tab << "torque_y" << tab << "torque_z" << tab << "presF_x" << tab << "presF_y" << tab
        << "presF_z" << tab << "viscF_x" << tab << "viscF_y" << tab << "viscF_z" << tab << "gamma_x" << tab
        << "gamma_y" << tab << "gamma_z" << tab << "drag" << tab << "thrust" << std::endl;
  ssF << sim.step << tab;
  ssF.setf(std::ios::scientific);
  ssF.precision(std::numeric_limits<float>::digits10 + 1);
  ssF << sim.time << tab << s.mass << tab << q[0] << tab << q[1] << tab << q[2] << tab << q[9] << tab << q[10] << tab
      << q[11] << tab << q[3] << tab << q[4] << tab << q[5] << tab << q[6] << tab << q[7] << tab << q[8] << tab << q[12]
      << tab << q[13] << tab << q[14] << tab << q[15] << tab << q[16] << std::endl;

  // powerValues_surface_<id>.dat: the derived powers of Obstacle::computeForces (:288-296)
  const double EPS = std::numeric_limits<double>::epsilon();
  const double vn = std::sqrt(s.transVel[0] * s.transVel[0] + s.transVel[1] * s.transVel[1] + s.transVel[2] * s.transVel[2]);
  const double Pthrust = q[16] * vn, Pdrag = q[15] * vn, Pout = q[17], PoutBnd = q[18], defPower = q[19],
               defPowerBnd = q[20], pLocom = q[21];
  const double EffPDef = Pthrust / (Pthrust - std::min(defPower, 0.0) + EPS);
  const double EffPDefBnd = Pthrust / (Pthrust - defPowerBnd + EPS);
  std::stringstream& ssP = sim.logger->get_stream("powerValues_surface_" + std::to_string(s.obstacleID) + ".dat");
  if (sim.step == 0)
    ssP << "step" << tab << "time" << tab << "Pthrust" << tab << "Pdrag" << tab << "Pout" << tab << "pDef" << tab
        << "etaPDef" << tab << "pLocom" << tab << "PoutBnd" << tab << "defPowerBnd" << tab << "etaPDefBnd" << std::endl;
  ssP << sim.step << tab;
  ssP.setf(std::ios::scientific);
  ssP.precision(std::numeric_limits<float>::digits10 + 1);
  ssP << sim.time << tab << Pthrust << tab << Pdrag << tab << Pout << tab << -defPower << tab << EffPDef << tab << pLocom
      << tab << PoutBnd << tab << -defPowerBnd << tab << EffPDefBnd << std::endl;

  // forceValues_<id>.dat: the penalization force and the momentum solve (Obstacle.cpp:557-589)
  std::stringstream& ssD = sim.logger->get_stream("forceValues_" + std::to_string(s.obstacleID) + ".dat");
  if (sim.step == 0)
    ssD << "step" << tab << "time" << tab << "mass" << tab << "force_x" << tab << "force_y" << tab << "force_z" << tab
        << "torque_x" << tab << "torque_y" << tab << "torque_z" << tab << "penalLmom_x" << tab << "penalLmom_y" << tab
        << "penalLmom_z" << tab << "penalAmom_x" << tab << "penalAmom_y" << tab << "penalAmom_z" << tab << "penalCM_x"
        << tab << "penalCM_y" << tab << "penalCM_z" << tab << "linVel_comp_x" << tab << "linVel_comp_y" << tab
        << "linVel_comp_z" << tab << "angVel_comp_x" << tab << "angVel_comp_y" << tab << "angVel_comp_z" << tab
        << "penalM" << tab << "penalJ0" << tab << "penalJ1" << tab << "penalJ2" << tab << "penalJ3" << tab << "penalJ4"
        << tab << "penalJ5" << std::endl;
  ssD << sim.step << tab;
  ssD.setf(std::ios::scientific);
  ssD.precision(std::numeric_limits<float>::digits10 + 1);
  ssD << sim.time << tab << s.mass;
  for (double v : s.force) ssD << tab << v;
  for (double v : s.torque) ssD << tab << v;
  for (double v : s.penalLmom) ssD << tab << v;
  for (double v : s.penalAmom) ssD << tab << v;
  for (double v : s.penalCM) ssD << tab << v;
  for (double v : s.transVel_computed) ssD << tab << v;
  for (double v : s.angVel_computed) ssD << tab << v;
  ssD << tab << s.penalM;
  for (double v : s.penalJ) ssD << tab << v;
  ssD << std::endl;
}

// ComputeDissipation::operator() (ComputeDissipation.cpp:119-178) and diagnostics.dat.
void ComputeDissipation::operator()(double)
{
  if (sim.freqDiagnostics == 0 || sim.step % sim.freqDiagnostics) return;
  double R[20];
  {
    Region r(sim, "Dissip Kernel");
    check(cup_diagnostics(sim.ctx, sim.nu, R), "ComputeDissipation");
  }
  check("ComputeDissipation");
  for (int i = 0; i < 20; ++i) last[i] = R[i];
  if (sim.muteAll) return;
  std::stringstream& f = sim.logger->get_stream("diagnostics.dat");
  if (sim.step == 0)
    f << "step_id time circ_x circ_y circ_y linImp_x linImp_y linImp_z linMom_x linMom_y linMom_z angImp_x angImp_y "
         "angImp_z angMom_x angMom_y angMom_z presPow viscPow helicity kineticEn enstrophy"
      << std::endl;
  f << sim.step << " " << sim.time;
  for (int i = 0; i < 20; ++i) f << " " << R[i];
  f << std::endl;
}

// Analysis::operator() (Analysis.cpp:96-112 trigger and file number, :159-196 HIT branch)
// and SpectralAnalysis::dump2File (SpectralAnalysis.cpp:124-205): the same rows, names and
// column widths, so the reference's post-processing reads these files. The directory is
// made when a file is due (the reference makes it at step 0 of every run).
void Analysis::operator()(double dt)
{
  if (sim.analysis != "HIT") return;
  const bool bFreq = sim.freqAnalysis > 0 && (sim.step + 1) % sim.freqAnalysis == 0;
  const bool bTime = sim.timeAnalysis > 0 && sim.time + dt > sim.nextAnalysisTime;
  if (!bFreq && !bTime) return;
  int nFile = sim.step;
  if (bTime) {
    sim.nextAnalysisTime += sim.timeAnalysis;
    nFile = sim.time <= 0 ? 0 : 1 + (int)(sim.time / sim.timeAnalysis);
  }
  const int nb = std::max({sim.bpdx, sim.bpdy, sim.bpdz}) * 4 - 1;  // max(N)/2 - 1 shells
  spectrum.assign(nb, 0.0);
  {
    Region r(sim, "HIT Analysis");
    check(cup_hit_analysis(sim.ctx, sim.nu, last.data(), spectrum.data(), nb), "Analysis");
  }
  sim.grad_mean = last[7];
  sim.grad_std = last[8];
  ++files;
  const double nu = sim.nu, eps = last[1];
  if (sim.verbose)
    printf("step:%d time:%e totalKinEn:%e viscousDissip:%e totalDissipRate:%e injectionRate:%e lIntegral:%e\n",
           sim.step, sim.time, last[0], eps, sim.dissipationRate, sim.actualInjectionRate, last[5]);
  struct Row {
    const char* name;
    double value;
    const char* what;
  };
  const Row rows[] = {{"time", sim.time, "simulation time"},
                      {"lBox", sim.maxextent, "simulation box length"},
                      {"tke", last[0], "turbulent kinetic energy"},
                      {"eps", eps, "Viscous dissipation rate"},
                      {"eps_f", sim.dissipationRate, "Total dissipation rate"},
                      {"lambda", last[3], "Taylor microscale"},
                      {"Re_lambda", last[4], "Turbulent Reynolds based on lambda"},
                      {"eta", std::pow(nu * nu * nu / eps, 0.25), "Kolmogorov length scale"},
                      {"tau_integral", last[6], "Integral time scale"},
                      {"l_integral", last[5], "Integral length scale"},
                      {"tau_eta", std::sqrt(nu / eps), "Kolmogorov time scale"},
                      {"nu_sgs", sim.nu_sgs, "Average SGS viscosity"},
                      {"cs2_avg", sim.cs2_avg, "Average Cs2 if dynamic model"},
                      {"mean_grad", sim.grad_mean, "Average gradient magnitude"},
                      {"std_grad", sim.grad_std, "Stdev gradient magnitude"}};
  mkdir("analysis", 0775);  // EEXIST when it is there already
  char name[64];
  snprintf(name, sizeof name, "analysis/spectralAnalysis_%09d", nFile);
  std::ofstream f(name);
  if (!f) throw IOError(std::string("Analysis: cannot write ") + name);
  f << std::left << "Spectral Analysis :" << "\n";
  for (const Row& r : rows) f << std::setw(15) << r.name << std::setw(15) << r.value << " #" << r.what << "\n";
  f << "\n" << std::setw(15) << "k * (lBox/2pi)" << std::setw(15) << "E_k" << "\n";
  for (int i = 0; i < nb; ++i)
    f << std::setw(15) << (i + 1) * 2 * M_PI / sim.maxextent << std::setw(15) << spectrum[i] << "\n";
}

// ------------------------------------------------------------- field files

// The saves' field container (the reference writes HDF5 through the Cubism library, whose
// layout is not available here): a 64-byte header -- the 8 bytes "CUPFIELD", int32 version
// (1), int32 ncomp, int32 cells x, y, z, int32 step, int32 0, FP64 time, 20 bytes of zeros --
// then ncomp FP64 arrays in [comp][z][y][x] order, little-endian.
// The XDMF 2 descriptor beside a field file, so that ParaView and VisIt open the fields as the
// reference's HDF5 dumps open (Simulation.cpp:475-505, Cubism's DumpHDF5_MPI writes <name>.h5
// and <name>.xmf): a 3DCoRectMesh of the cell corners with origin 0 and spacing h, and
// cell-centred attributes read from the raw FP64 arrays after the 64-byte header (component q
// of n cells at byte 64 + 8 n q). A vector field gets its three components and the vector
// joined from them. The binary file is named relative to the descriptor.
static void write_xdmf(const std::string& path, const std::string& binary, const SimulationData& sim,
                       const char* const* names, int ncomp, const char* vector_name)
{
  const int N[3] = {sim.bpdx * 8, sim.bpdy * 8, sim.bpdz * 8};
  const int maxN = std::max(N[0], std::max(N[1], N[2]));
  const double h = sim.maxextent / maxN;
  const size_t n = (size_t)N[0] * N[1] * N[2];
  FILE* f = fopen(path.c_str(), "w");
  if (!f) throw IOError("cannot write " + path);
  const std::string base = binary.substr(binary.find_last_of('/') + 1);
  fprintf(f, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n");
  fprintf(f, " <Domain>\n  <Grid Name=\"cells\" GridType=\"Uniform\">\n");
  fprintf(f, "   <Time Value=\"%.17e\"/>\n", sim.time);
  fprintf(f, "   <Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%d %d %d\"/>\n", N[2] + 1, N[1] + 1, N[0] + 1);
  fprintf(f, "   <Geometry GeometryType=\"ORIGIN_DXDYDZ\">\n");
  fprintf(f, "    <DataItem Name=\"Origin\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" Format=\"XML\">0 0 0</DataItem>\n");
  fprintf(f, "    <DataItem Name=\"Spacing\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" Format=\"XML\">%.17e %.17e %.17e</DataItem>\n", h, h, h);
  fprintf(f, "   </Geometry>\n");
  auto item = [&](int q, const char* indent) {
    fprintf(f, "%s<DataItem Dimensions=\"%d %d %d\" NumberType=\"Float\" Precision=\"8\" Format=\"Binary\" Endian=\"Little\" Seek=\"%zu\">%s</DataItem>\n",
            indent, N[2], N[1], N[0], 64 + 8 * n * (size_t)q, base.c_str());
  };
  for (int q = 0; q < ncomp; ++q) {
    fprintf(f, "   <Attribute Name=\"%s\" AttributeType=\"Scalar\" Center=\"Cell\">\n", names[q]);
    item(q, "    ");
    fprintf(f, "   </Attribute>\n");
  }
  if (vector_name) {
    fprintf(f, "   <Attribute Name=\"%s\" AttributeType=\"Vector\" Center=\"Cell\">\n", vector_name);
    fprintf(f, "    <DataItem ItemType=\"Function\" Function=\"JOIN($0, $1, $2)\" Dimensions=\"%d %d %d 3\">\n", N[2], N[1], N[0]);
    for (int q = 0; q < 3; ++q) item(q, "     ");
    fprintf(f, "    </DataItem>\n   </Attribute>\n");
  }
  fprintf(f, "  </Grid>\n </Domain>\n</Xdmf>\n");
  if (fclose(f) != 0) throw IOError("cannot write " + path);
}

static void write_field_file(const std::string& path, const SimulationData& sim, const int* fields, int ncomp)
{
  const int N[3] = {sim.bpdx * 8, sim.bpdy * 8, sim.bpdz * 8};
  const size_t n = (size_t)N[0] * N[1] * N[2];
  std::vector<double> blocks(n), grid(n);
  std::ofstream f(path, std::ios::binary);
  if (!f) throw IOError("cannot write " + path);
  char head[64] = {};
  std::memcpy(head, "CUPFIELD", 8);
  const int32_t ints[7] = {1, ncomp, N[0], N[1], N[2], sim.step, 0};
  std::memcpy(head + 8, ints, sizeof ints);
  std::memcpy(head + 36, &sim.time, sizeof(double));
  f.write(head, sizeof head);
  for (int q = 0; q < ncomp; ++q) {
    if (cup_get_field(sim.ctx, fields[q], blocks.data()) != CUP_OK)
      throw std::runtime_error(std::string("save: ") + cup_last_error(sim.ctx));
    for (size_t b = 0; b < n / 512; ++b) {
      const int bx = (int)(b % sim.bpdx), by = (int)((b / sim.bpdx) % sim.bpdy), bz = (int)(b / ((size_t)sim.bpdx * sim.bpdy));
      for (int c = 0; c < 512; ++c) {
        const size_t x = bx * 8 + (c & 7), y = by * 8 + ((c >> 3) & 7), z = bz * 8 + (c >> 6);
        grid[(z * N[1] + y) * N[0] + x] = blocks[b * 512 + c];
      }
    }
    f.write(reinterpret_cast<const char*>(grid.data()), n * sizeof(double));
  }
  if (!f) throw IOError("cannot write " + path);
}

static void read_field_file(const std::string& path, const SimulationData& sim, const int* fields, int ncomp)
{
  const int N[3] = {sim.bpdx * 8, sim.bpdy * 8, sim.bpdz * 8};
  const size_t n = (size_t)N[0] * N[1] * N[2];
  std::ifstream f(path, std::ios::binary);
  if (!f) throw IOError("Unable to restart: cannot read " + path);
  char head[64];
  f.read(head, sizeof head);
  int32_t ints[7];
  std::memcpy(ints, head + 8, sizeof ints);
  if (!f || std::memcmp(head, "CUPFIELD", 8) != 0 || ints[0] != 1 || ints[1] != ncomp || ints[2] != N[0] ||
      ints[3] != N[1] || ints[4] != N[2])
    throw IOError("Unable to restart: " + path + " does not hold this grid's fields");
  std::vector<double> blocks(n), grid(n);
  for (int q = 0; q < ncomp; ++q) {
    f.read(reinterpret_cast<char*>(grid.data()), n * sizeof(double));
    if (!f) throw IOError("Unable to restart: " + path + " is truncated");
    for (size_t b = 0; b < n / 512; ++b) {
      const int bx = (int)(b % sim.bpdx), by = (int)((b / sim.bpdx) % sim.bpdy), bz = (int)(b / ((size_t)sim.bpdx * sim.bpdy));
      for (int c = 0; c < 512; ++c) {
        const size_t x = bx * 8 + (c & 7), y = by * 8 + ((c >> 3) & 7), z = bz * 8 + (c >> 6);
        blocks[b * 512 + c] = grid[(z * N[1] + y) * N[0] + x];
      }
    }
    if (cup_set_field(sim.ctx, fields[q], blocks.data()) != CUP_OK)
      throw std::runtime_error(std::string("restart: ") + cup_last_error(sim.ctx));
  }
}

// ----------------------------------------------------------------- Simulation

// Simulation::_serialize (Simulation.cpp:426-532): <path>/<name>.status, the obstacles'
// <name>_<i>.txt, the fields, then <path>/restart.status; name = restart_ or the given prefix,
// followed by the step in 9 digits.
void Simulation::_serialize(const std::string& append)
{
  char num[16];
  snprintf(num, sizeof num, "%09d", sim.step);
  const std::string name = (append.empty() ? std::string("restart_") : append) + num;
  mkdir(sim.path4serialization.c_str(), 0775);  // EEXIST when it is there already
  const std::string fpath = sim.path4serialization + "/" + name;
  for (size_t i = 0; i < sim.obstacles.size(); ++i) sim.obstacles[i]->save(fpath + "_" + std::to_string(i), sim.time);
  auto status = [&](const std::string& file) {
    FILE* f = fopen(file.c_str(), "w");
    if (!f) throw IOError("cannot write " + file);
    fprintf(f, "time: %20.20e\n", sim.time);
    fprintf(f, "stepid: %d\n", sim.step);
    fprintf(f, "uinfx: %20.20e\n", sim.uinf[0]);
    fprintf(f, "uinfy: %20.20e\n", sim.uinf[1]);
    fprintf(f, "uinfz: %20.20e\n", sim.uinf[2]);
    fclose(f);
  };
  status(fpath + ".status");
  if (sim.b3Ddump) {
    const int vel[3] = {CUP_U, CUP_V, CUP_W}, pres[1] = {CUP_P}, chi[1] = {CUP_CHI};
    const char* const vel_names[3] = {"u", "v", "w"};
    const char* const pres_names[1] = {"pressure"};
    const char* const chi_names[1] = {"chi"};
    const std::string root = sim.path4serialization + "/";
    write_field_file(root + "vel_" + name + ".cupf", sim, vel, 3);
    write_xdmf(root + "vel_" + name + ".xmf", root + "vel_" + name + ".cupf", sim, vel_names, 3, "velocity");
    write_field_file(root + "pres_" + name + ".cupf", sim, pres, 1);
    write_xdmf(root + "pres_" + name + ".xmf", root + "pres_" + name + ".cupf", sim, pres_names, 1, nullptr);
    write_field_file(root + "chi_" + name + ".cupf", sim, chi, 1);
    write_xdmf(root + "chi_" + name + ".xmf", root + "chi_" + name + ".cupf", sim, chi_names, 1, nullptr);
  }
  status(sim.path4serialization + "/restart.status");
}

// Simulation::_deserialize (Simulation.cpp:534-580): restart.status, then the velocity of
// restart_<step> and the bodies' state; chi is built again where the bodies are. Without a
// restart.status the run starts afresh from the IC (the reference starts from the fields as
// allocated).
void Simulation::_deserialize()
{
  const std::string file = sim.path4serialization + "/restart.status";
  FILE* f = fopen(file.c_str(), "r");
  if (!f) {
    printf("Could not restart... starting a new sim.\n");
    _ic();
    return;
  }
  double t = -1, u[3] = {0, 0, 0};
  int step = -1;
  bool ok = fscanf(f, "time: %le\n", &t) == 1;
  ok = ok && fscanf(f, "stepid: %d\n", &step) == 1;
  ok = ok && fscanf(f, "uinfx: %le\n", &u[0]) == 1;
  ok = ok && fscanf(f, "uinfy: %le\n", &u[1]) == 1;
  ok = ok && fscanf(f, "uinfz: %le\n", &u[2]) == 1;
  fclose(f);
  if (!ok || step < 0 || t < 0) throw IOError("Error reading restart file " + file);
  sim.time = t;
  sim.step = step;
  sim.uinf = {u[0], u[1], u[2]};
  char num[16];
  snprintf(num, sizeof num, "%09d", step);
  const std::string name = std::string("restart_") + num;
  const int vel[3] = {CUP_U, CUP_V, CUP_W};
  read_field_file(sim.path4serialization + "/vel_" + name + ".cupf", sim, vel, 3);
  for (size_t i = 0; i < sim.obstacles.size(); ++i)
    sim.obstacles[i]->restart(sim.path4serialization + "/" + name + "_" + std::to_string(i));
  for (auto& op : sim.pipeline)
    if (auto* c = dynamic_cast<CreateObstacles*>(op.get())) c->build();
  if (cup_set_step_state(sim.ctx, sim.step, sim.time) != CUP_OK)
    throw std::runtime_error(std::string("restart: ") + cup_last_error(sim.ctx));
  sim.nextSaveTime = sim.time + sim.saveTime;
  printf("DESERIALIZATION: time is %f and step id is %d\n", sim.time, sim.step);
}

Simulation::Simulation(const SimulationData& sd) : sim(sd)
{
  sim._preprocessArguments();
  setupGrid();
  setupOperators();
  _init();
}

Simulation::Simulation(int argc, const char* const* argv) : sim(argc, argv)
{
  sim._preprocessArguments();
  setupGrid();
  setupOperators();
  _init();
}

// Simulation::_init (Simulation.cpp:154-165): restart or IC, then the "init" save. The init
// save is written only when saves are asked for (-fsave/-tsave, or -fdump/-tdump).
void Simulation::_init()
{
  if (sim.bRestart) _deserialize();
  else _ic();
  if (sim.saveFreq > 0 || sim.saveTime > 0) _serialize("init");
}

Simulation::~Simulation()
{
  sim.pipeline.clear();
  if (sim.logger) sim.logger->flush();
}

void Simulation::setupGrid()
{
  cup_grid_desc d;
  d.bpd[0] = sim.bpdx;
  d.bpd[1] = sim.bpdy;
  d.bpd[2] = sim.bpdz;
  d.maxextent = sim.maxextent;
  d.precision = sim.precision;
  d.decomposition = sim.decomposition;
  d.device = sim.device;
  const BCflag f[3] = {sim.BCx_flag, sim.BCy_flag, sim.BCz_flag};
  for (int a = 0; a < 3; ++a)
    d.bc[a] = f[a] == periodic ? CUP_BC_PERIODIC
                               : (f[a] == wall ? CUP_BC_WALL : (f[a] == freespace ? CUP_BC_FREESPACE : CUP_BC_OPEN));
  const int rc = cup_create(&d, &sim.ctx);
  if (rc != CUP_OK) throw std::runtime_error("cup_create failed (" + std::to_string(rc) + ")");
}

void Simulation::addObstacle(std::shared_ptr<Sphere> s)
{
  s->obstacleID = (int)sim.obstacles.size();
  sim.obstacles.push_back(std::move(s));
  try {
    setupOperators();
  } catch (...) {  // e.g. obstacle blocks overlapping another body's: the body is not added
    sim.obstacles.pop_back();
    setupOperators();
    throw;
  }
}

// Simulation::setupOperators (Simulation.cpp:307-401), periodic, non-iterative pipeline.
void Simulation::setupOperators()
{
  sim.pipeline.clear();
  auto* create = new CreateObstacles(sim);
  sim.pipeline.emplace_back(create);
  sim.pipeline.emplace_back(new AdvectionDiffusion(sim));
  // Simulation.cpp:336-339: the uniform forcing unless the TGV (which also reads uMax_forced)
  if (sim.uMax_forced > 0 && sim.initCond != "taylorGreen" && !sim.bChannelFixedMassFlux)
    sim.pipeline.emplace_back(new ExternalForcing(sim));
  if (sim.spectralForcing) sim.pipeline.emplace_back(new SpectralForcing(sim));
  sim.pipeline.emplace_back(new PressureRHS(sim));
  sim.pipeline.emplace_back(new PressureProjection(sim));
  sim.pipeline.emplace_back(new UpdateObstacles(sim));
  sim.pipeline.emplace_back(new Penalization(sim));
  sim.pipeline.emplace_back(new ComputeForces(sim));
  sim.pipeline.emplace_back(new ComputeDissipation(sim));
  sim.pipeline.emplace_back(new Analysis(sim));
  create->build();
}

void Simulation::_ic()
{
  InitialConditions ic(sim);
  {
    Region r(sim, ic.getName().c_str());
    ic(0);
  }
  // The zero and Taylor-Green ICs clear every field, chi included. A body that does not move
  // keeps its chi (the reference rebuilds it in every step's CreateObstacles), so it is built
  // again here.
  for (auto& op : sim.pipeline)
    if (auto* c = dynamic_cast<CreateObstacles*>(op.get())) c->build();
}

void Simulation::reset()
{
  _ic();
  sim.step = 0;
  sim.time = 0;
  sim.uinf = {0, 0, 0};
  cup_set_step_state(sim.ctx, 0, 0.0);
}

// Simulation::calcMaxTimestep (Simulation.cpp:403-424).
double Simulation::calcMaxTimestep()
{
  Region r(sim, "DT");
  double dt = 0, umax = 0;
  const int rc =
      cup_calc_max_timestep(sim.ctx, sim.CFL, sim.nu, sim.uinf.data(), sim.step, sim.rampup, &dt, &umax);
  if (rc != CUP_OK) throw std::runtime_error(std::string("calcMaxTimestep: ") + cup_last_error(sim.ctx));
  sim.uMax_measured = umax;
  sim.dt = dt;
  if (sim.DLM > 0) sim.lambda = sim.DLM / sim.dt;
  return sim.dt;
}

// Simulation::timestep (Simulation.cpp:593-623).
bool Simulation::timestep(double dt)
{
  const bool bDumpFreq = sim.saveFreq > 0 && (sim.step + 1) % sim.saveFreq == 0;
  const bool bDumpTime = sim.saveTime > 0 && sim.time + dt > sim.nextSaveTime;
  if (bDumpTime) sim.nextSaveTime += sim.saveTime;
  sim.bDump = bDumpFreq || bDumpTime;
  for (auto& op : sim.pipeline) (*op)(dt);
  sim.step++;
  sim.time += dt;
  cup_set_step_state(sim.ctx, sim.step, sim.time);
  if (sim.verbose) printf("%d : %f uInf {%f %f %f}\n", sim.step, sim.time, sim.uinf[0], sim.uinf[1], sim.uinf[2]);
  {
    Region r(sim, "Save");
    if (sim.bDump) _serialize();
  }
  if (sim.step % 50 == 0 && sim.verbose) sim.printResetProfiler();
  if ((sim.endTime > 0 && sim.time > sim.endTime) || (sim.nsteps != 0 && sim.step >= sim.nsteps)) return true;
  return false;
}

void Simulation::syncStepState()
{
  int step = 0;
  double time = 0, last = 0, umax = 0, ke = 0;
  if (cup_step_state(sim.ctx, &step, &time, &last, &umax, &ke) != CUP_OK)
    throw std::runtime_error(std::string("step state: ") + cup_last_error(sim.ctx));
  sim.step = step;
  sim.time = time;
  sim.dt = last;
  sim.uMax_measured = umax;
}

bool Simulation::advance(int n)
{
  // The device-resident loop neither moves a body nor builds its chi again, and has no uniform
  // forcing: it is taken only when every body keeps its imposed velocity and is at rest in the
  // frame, transVel + uinf = 0 on every axis with uinf the bodies' frame velocity (a body at rest
  // beside a towed one that fixes the frame moves, and so do frame-fixing bodies of different
  // velocities); other steps take the pipeline, which moves the bodies.
  const std::array<double, 3> uframe = frameVelocity(sim.obstacles);
  bool moving = false;
  for (const auto& s : sim.obstacles) {
    moving = moving || !s->constantVelocity();
    for (int d = 0; d < 3; ++d) moving = moving || s->transVel[d] + uframe[d] != 0;
  }
  const bool external = sim.uMax_forced > 0 && sim.initCond != "taylorGreen";
  if (sim.decomposition != CUP_FUSED || moving || external) {
    for (int i = 0; i < n; ++i)
      if (timestep(calcMaxTimestep())) return true;
    return false;
  }
  // The frame velocity is set by the first step's CreateObstacles, after that step's dt
  // was computed with the previous one: that step goes through the pipeline when they differ.
  if (!sim.obstacles.empty() && n > 0) {
    if (uframe != sim.uinf) {
      if (timestep(calcMaxTimestep())) return true;
      --n;
    }
  }
  cup_step_params p{};
  p.CFL = sim.CFL;
  p.nu = sim.nu;
  for (int d = 0; d < 3; ++d) p.uinf[d] = sim.uinf[d];
  p.rampup = sim.rampup;
  p.scheme = sim.scheme;
  p.spectral_forcing = sim.spectralForcing ? 1 : 0;
  p.tke_target = sim.turbKinEn_target;
  p.injection_rate = sim.enInjectionRate;
  p.fixed_dt = 0;
  if (!sim.obstacles.empty()) {
    auto& s = *sim.obstacles[0];
    p.has_body = 1;
    p.lambda = sim.lambda;
    p.implicit_penalization = sim.bImplicitPenalization ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
      p.body_vel[d] = s.transVel[d];
      p.body_angvel[d] = s.angVel[d];
    }
  }
  if (n <= 0) return (sim.endTime > 0 && sim.time > sim.endTime) || (sim.nsteps != 0 && sim.step >= sim.nsteps);
  for (size_t k = 1; k < sim.obstacles.size(); ++k)
    if (cup_body_velocity(sim.ctx, (int)k, sim.obstacles[k]->transVel.data(), sim.obstacles[k]->angVel.data()) !=
        CUP_OK)
      throw std::runtime_error(std::string("advance: ") + cup_last_error(sim.ctx));
  {
    // the device-resident loop: its stages are roctx ranges of their own
    Region r(sim, "Fused steps");
    if (cup_run_steps(sim.ctx, &p, n) != CUP_OK)
      throw std::runtime_error(std::string("advance: ") + cup_last_error(sim.ctx));
  }
  checkNaN(sim, "Fused steps");
  syncStepState();
  for (size_t k = 0; k < sim.obstacles.size(); ++k) {  // each body's last penalization force and surface QoIs
    auto& s = *sim.obstacles[k];
    double ft[6], q[22];
    if (cup_last_body_force_body(sim.ctx, (int)k, ft) != CUP_OK ||
        cup_surface_forces_body(sim.ctx, (int)k, sim.nu, s.transVel.data(), s.angVel.data(), q) != CUP_OK)
      throw std::runtime_error(std::string("advance: ") + cup_last_error(sim.ctx));
    s.force = {ft[0], ft[1], ft[2]};
    s.torque = {ft[3], ft[4], ft[5]};
    for (int i = 0; i < 22; ++i) s.surfaceQoI[i] = q[i];
  }
  return (sim.endTime > 0 && sim.time > sim.endTime) || (sim.nsteps != 0 && sim.step >= sim.nsteps);
}

// Simulation::run (Simulation.cpp:582-591).
void Simulation::run()
{
  const bool analysis = sim.analysis == "HIT" && (sim.freqAnalysis > 0 || sim.timeAnalysis > 0);
  const bool external = sim.uMax_forced > 0 && sim.initCond != "taylorGreen";  // not in the device loop
  // the device-resident loop writes no per-step files: taken when none is asked for
  const bool logs = (!sim.muteAll && (!sim.obstacles.empty() || sim.spectralForcing)) || sim.saveFreq > 0 ||
                    sim.saveTime > 0;
  if (sim.decomposition == CUP_FUSED && sim.nsteps > 0 && sim.endTime <= 0 && sim.freqDiagnostics == 0 && !analysis &&
      !logs && !external) {
    advance(sim.nsteps - sim.step);
    return;
  }
  for (;;) {
    const double dt = calcMaxTimestep();
    if (timestep(dt)) break;
  }
}

}  // namespace cubismup3d

// ------------------------------------------------------------ C entry points

using namespace cubismup3d;

struct cup_sim {
  std::unique_ptr<Simulation> S;
  std::unique_ptr<SimulationData> pending;  // before the Simulation is built
  std::string err;
};

template <typename F>
static int guarded(cup_sim* h, F&& f)
{
  if (!h) return CUP_E_INVALID;
  try {
    f();
    return CUP_OK;
  } catch (const std::invalid_argument& e) {
    h->err = e.what();
    return CUP_E_INVALID;
  } catch (const IOError& e) {
    h->err = e.what();
    return CUP_E_IO;
  } catch (const std::exception& e) {
    h->err = e.what();
    return CUP_E_DEVICE;
  }
}

extern "C" {

int cup_sim_create(int argc, const char* const* argv, cup_sim** out)
{
  if (!out) return CUP_E_INVALID;
  *out = new cup_sim();
  return guarded(*out, [&] { (*out)->pending.reset(new SimulationData(argc, argv)); });
}

int cup_sim_add_sphere(cup_sim* h, double radius, const double position[3], const double enforced_vel[3],
                       int forced_mask, int fix_mask, int accel_decel, double umax, double tmax, int hemi)
{
  return guarded(h, [&] {
    if (!(radius > 0)) throw std::invalid_argument("radius must be > 0");
    if (!position || !enforced_vel) throw std::invalid_argument("null argument");
    if (accel_decel && !(tmax > 0)) throw std::invalid_argument("tmax must be > 0");
    auto s = std::make_shared<Sphere>();
    s->radius = radius;
    for (int d = 0; d < 3; ++d) {
      s->position[d] = position[d];
      s->bForcedInSimFrame[d] = (forced_mask >> d) & 1;
      s->bFixFrameOfRef[d] = (fix_mask >> d) & 1;
    }
    s->setVelocities({enforced_vel[0], enforced_vel[1], enforced_vel[2]});
    s->accel_decel = accel_decel != 0;
    s->umax = umax;
    s->tmax = tmax;
    s->bHemi = hemi != 0;
    if (h->S) h->S->addObstacle(s);
    else if (h->pending) h->pending->obstacles.push_back(s);
    else throw std::invalid_argument("no simulation");
  });
}

int cup_sim_init(cup_sim* h)
{
  return guarded(h, [&] {
    if (h->S) return;
    if (!h->pending) throw std::invalid_argument("no simulation data");
    h->S.reset(new Simulation(*h->pending));
    h->pending.reset();
  });
}

int cup_sim_destroy(cup_sim* h)
{
  if (!h) return CUP_E_INVALID;
  delete h;
  return CUP_OK;
}

const char* cup_sim_last_error(const cup_sim* h) { return h ? h->err.c_str() : "null simulation"; }

int cup_sim_run(cup_sim* h)
{
  return guarded(h, [&] {
    if (!h->S) throw std::invalid_argument("simulation not initialised");
    h->S->run();
  });
}

int cup_sim_calc_max_timestep(cup_sim* h, double* dt)
{
  return guarded(h, [&] {
    if (!h->S) throw std::invalid_argument("simulation not initialised");
    const double v = h->S->calcMaxTimestep();
    if (dt) *dt = v;
  });
}

int cup_sim_timestep(cup_sim* h, double dt, int* finished)
{
  return guarded(h, [&] {
    if (!h->S) throw std::invalid_argument("simulation not initialised");
    h->S->sim.dt = dt;
    const bool f = h->S->timestep(dt);
    if (finished) *finished = f ? 1 : 0;
  });
}

int cup_sim_advance(cup_sim* h, int nsteps, int* finished)
{
  return guarded(h, [&] {
    if (!h->S) throw std::invalid_argument("simulation not initialised");
    const bool f = h->S->advance(nsteps);
    if (finished) *finished = f ? 1 : 0;
  });
}

int cup_sim_context(cup_sim* h, cup_ctx** ctx)
{
  return guarded(h, [&] {
    if (!h->S) throw std::invalid_argument("simulation not initialised");
    *ctx = h->S->sim.ctx;
  });
}

// Scalar members of SimulationData by name.
static double* sim_field(SimulationData& s, const std::string& n)
{
  static double tmp;
  if (n == "time") return &s.time;
  if (n == "dt") return &s.dt;
  if (n == "endTime") return &s.endTime;
  if (n == "nu") return &s.nu;
  if (n == "CFL") return &s.CFL;
  if (n == "lambda") return &s.lambda;
  if (n == "uinf_x") return &s.uinf[0];
  if (n == "uinf_y") return &s.uinf[1];
  if (n == "uinf_z") return &s.uinf[2];
  if (n == "extent_x") return &s.extent[0];
  if (n == "extent_y") return &s.extent[1];
  if (n == "extent_z") return &s.extent[2];
  if (n == "h") return &s.hmean;
  if (n == "uMax_measured") return &s.uMax_measured;
  if (n == "turbKinEn_target") return &s.turbKinEn_target;
  if (n == "dissipationRate") return &s.dissipationRate;
  if (n == "actualInjectionRate") return &s.actualInjectionRate;
  if (n == "timeAnalysis") return &s.timeAnalysis;
  if (n == "nextAnalysisTime") return &s.nextAnalysisTime;
  if (n == "grad_mean") return &s.grad_mean;
  if (n == "grad_std") return &s.grad_std;
  (void)tmp;
  return nullptr;
}

static int* sim_ifield(SimulationData& s, const std::string& n)
{
  if (n == "step") return &s.step;
  if (n == "nsteps") return &s.nsteps;
  if (n == "rampup") return &s.rampup;
  if (n == "freqDiagnostics") return &s.freqDiagnostics;
  if (n == "freqAnalysis") return &s.freqAnalysis;
  if (n == "bpdx") return &s.bpdx;
  if (n == "bpdy") return &s.bpdy;
  if (n == "bpdz") return &s.bpdz;
  if (n == "precision") return &s.precision;
  if (n == "decomposition") return &s.decomposition;
  if (n == "scheme") return &s.scheme;
  return nullptr;
}

static bool* sim_bfield(SimulationData& s, const std::string& n)
{
  if (n == "verbose") return &s.verbose;
  if (n == "muteAll") return &s.muteAll;
  if (n == "bCheckNaN") return &s.bCheckNaN;
  if (n == "bImplicitPenalization") return &s.bImplicitPenalization;
  if (n == "spectralForcing") return &s.spectralForcing;
  return nullptr;
}

int cup_sim_get(cup_sim* h, const char* name, double* v)
{
  return guarded(h, [&] {
    SimulationData* s = h->S ? &h->S->sim : h->pending.get();
    if (!s || !name || !v) throw std::invalid_argument("bad argument");
    if (double* p = sim_field(*s, name)) *v = *p;
    else if (int* q = sim_ifield(*s, name)) *v = *q;
    else if (bool* b = sim_bfield(*s, name)) *v = *b ? 1 : 0;
    else throw std::invalid_argument(std::string("unknown member ") + name);
  });
}

int cup_sim_set(cup_sim* h, const char* name, double v)
{
  return guarded(h, [&] {
    SimulationData* s = h->S ? &h->S->sim : h->pending.get();
    if (!s || !name) throw std::invalid_argument("bad argument");
    const std::string n = name;
    if (h->S && (n == "bpdx" || n == "bpdy" || n == "bpdz" || n == "precision" || n == "decomposition"))
      throw std::invalid_argument(n + " is fixed once the simulation exists");
    if (double* p = sim_field(*s, n)) *p = v;
    else if (int* q = sim_ifield(*s, n)) *q = (int)v;
    else if (bool* b = sim_bfield(*s, n)) *b = v != 0;
    else throw std::invalid_argument("unknown member " + n);
    if (h->S && n == "step") cup_set_step_state(s->ctx, s->step, s->time);
    if (h->S && n == "time") cup_set_step_state(s->ctx, s->step, s->time);
  });
}

int cup_sim_set_cells(cup_sim* h, int nx, int ny, int nz)
{
  return guarded(h, [&] {
    if (!h->pending) throw std::invalid_argument("cells are fixed once the simulation exists");
    h->pending->setCells(nx, ny, nz);
  });
}

// Obstacle state: out = CoM[3], mass, force[3], torque[3], surface QoIs[22] (34 doubles).
int cup_sim_obstacle_state(cup_sim* h, int idx, double out[13])
{
  return guarded(h, [&] {
    if (!h->S || idx < 0 || idx >= (int)h->S->sim.obstacles.size()) throw std::invalid_argument("no such obstacle");
    const Sphere& s = *h->S->sim.obstacles[idx];
    for (int d = 0; d < 3; ++d) out[d] = s.position[d];
    for (int i = 0; i < 4; ++i) out[3 + i] = s.quaternion[i];
    for (int d = 0; d < 3; ++d) {
      out[7 + d] = s.transVel[d];
      out[10 + d] = s.angVel[d];
    }
  });
}

int cup_sim_obstacle(cup_sim* h, int idx, double out[34])
{
  return guarded(h, [&] {
    if (!h->S || idx < 0 || idx >= (int)h->S->sim.obstacles.size()) throw std::invalid_argument("no such obstacle");
    const Sphere& s = *h->S->sim.obstacles[idx];
    for (int d = 0; d < 3; ++d) out[d] = s.centerOfMass[d];
    out[3] = s.mass;
    for (int d = 0; d < 3; ++d) { out[4 + d] = s.force[d]; out[7 + d] = s.torque[d]; }
    for (int i = 0; i < 22; ++i) out[10 + i] = s.surfaceQoI[i];
  });
}

int cup_sim_diagnostics(cup_sim* h, double out[20])
{
  return guarded(h, [&] {
    if (!h->S) throw std::invalid_argument("simulation not initialised");
    if (cup_diagnostics(h->S->sim.ctx, h->S->sim.nu, out) != CUP_OK)
      throw std::runtime_error(cup_last_error(h->S->sim.ctx));
  });
}

int cup_sim_flush_logs(cup_sim* h)
{
  return guarded(h, [&] {
    if (h->S && h->S->sim.logger) h->S->sim.logger->flush();
  });
}

}  // extern "C"
