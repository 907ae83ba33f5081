// simulation.h -- host C++ layer of cubismup_3d_amd: the reference's
// Simulation / SimulationData / Operator API (source/Simulation.h:24-71,
// source/SimulationData.h:46-179, source/operators/Operator.h:16-177) for the periodic
// per-timestep path, each operator calling the C-ABI of include/cup.h, which launches the
// HIP kernels. One Simulation owns one device context; calls are not re-entrant.
#pragma once

#include <array>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/cup.h"

namespace cubismup3d {

using Real = double;

enum BCflag { dirichlet, periodic, wall, freespace };  // Definitions.h:74
BCflag string2BCflag(const std::string& s);

// "-name value" command-line arguments (the subset of cubism::ArgumentParser used by
// SimulationData's constructor, SimulationData.cpp:32-172). Keys are matched with or
// without their leading dash.
// A file that could not be read or written (CUP_E_IO at the C ABI).
struct IOError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class ArgumentParser
{
 public:
  class Value
  {
   public:
    Value(bool present, std::string s) : present_(present), s_(std::move(s)) {}
    int asInt(int def) const;
    int asInt() const;
    double asDouble(double def) const;
    double asDouble() const;
    std::string asString(const std::string& def) const;
    bool asBool(bool def) const;
    bool exists() const { return present_; }

   private:
    bool present_;
    std::string s_;
  };
  ArgumentParser(int argc, const char* const* argv);
  Value operator()(const std::string& key) const;

 private:
  std::map<std::string, std::string> args_;
};

// cubism::Profiler as SimulationData drives it (SimulationData.cpp:293-305): nested named
// regions opened by push_start and closed by pop_stop. Every region is a roctx range, so a
// rocprofv3 run with --marker-trace files each kernel under the region that launched it. With
// timing on, a region also adds its wall-clock duration to its name's total; both ends wait
// for the context's stream first, so the time is that of the device work the region issued
// (the reference's regions time synchronous host code the same way).
class Profiler
{
 public:
  struct Agent {
    double seconds = 0;
    long count = 0;
  };
  bool timing = false;
  void push_start(const std::string& name, cup_ctx* ctx);
  void pop_stop(cup_ctx* ctx);
  // a table of the regions by total time: name, total and mean milliseconds, share, count
  void printSummary(FILE* out = stdout) const;
  void reset();
  const std::map<std::string, Agent>& agents() const { return agents_; }

 private:
  std::map<std::string, Agent> agents_;
  std::vector<std::pair<std::string, std::chrono::steady_clock::time_point>> stack_;
};

// BufferedLogger (utils/BufferedLogger.cpp:18-76): per-file string streams appended to
// disk every AUTO_FLUSH_COUNT requests and at destruction.
class BufferedLogger
{
 public:
  static constexpr int AUTO_FLUSH_COUNT = 100;
  explicit BufferedLogger(std::string dir = ".") : dir_(std::move(dir)) {}
  ~BufferedLogger() { flush(); }
  std::stringstream& get_stream(const std::string& filename);
  void flush();
  void set_dir(const std::string& d) { dir_ = d; }

 private:
  struct Stream {
    std::stringstream stream;
    int requests_since_last_flush = 0;
  };
  std::string dir_;
  std::map<std::string, Stream> files_;
  void flush_one(const std::string& name, Stream& s);
};

// A sphere (obstacles/Sphere.h, obstacles/Obstacle.cpp:24-151 arguments). Only forced
// motion is supported: velocities stay at their imposed values
// (Obstacle::computeVelocities with every degree of freedom forced, Obstacle.cpp:153-264).
struct Sphere {
  double radius = 0;
  std::array<double, 3> position = {0, 0, 0}, absPos = {0, 0, 0};
  std::array<double, 4> quaternion = {1, 0, 0, 0};
  double _2Dangle = 0;
  std::array<bool, 3> bForcedInSimFrame = {false, false, false};
  std::array<bool, 3> bFixFrameOfRef = {false, false, false};
  std::array<bool, 3> bBlockRotation = {false, false, false};
  std::array<double, 3> transVel_imposed = {0, 0, 0};
  std::array<double, 3> transVel = {0, 0, 0};
  std::array<double, 3> angVel = {0, 0, 0};
  std::array<double, 3> centerOfMass = {0, 0, 0};
  double mass = 0;
  std::array<double, 3> force = {0, 0, 0}, torque = {0, 0, 0};  // penalization integrals
  std::array<double, 22> surfaceQoI = {};                        // ObstacleBlock::sumQoI order
  // UpdateObstacles' penalized momenta and the velocities they solve to (Obstacle.h:120-140)
  double penalM = 0;
  std::array<double, 3> penalLmom = {0, 0, 0}, penalAmom = {0, 0, 0}, penalCM = {0, 0, 0};
  std::array<double, 6> penalJ = {0, 0, 0, 0, 0, 0}, J = {0, 0, 0, 0, 0, 0};
  std::array<double, 3> transVel_computed = {0, 0, 0}, angVel_computed = {0, 0, 0};
  int obstacleID = 0;
  // SphereArguments (Sphere.h:17-29): the ramped x velocity, up to umax at tmax, 0 at 2 tmax
  bool accel_decel = false, bHemi = false;
  double umax = 0, tmax = 1;

  // The Obstacle constructor's velocity set-up (Obstacle.cpp:119-137): forced axes take the
  // imposed velocity, the others start at rest; any forced axis blocks the rotation.
  void setVelocities(const std::array<double, 3>& imposed);
  // Every axis forced, no ramp: the velocity stays at its imposed value and the rotation at 0.
  bool constantVelocity() const;
  // constantVelocity, and at rest in the frame when it is the only body: each axis is fixed to
  // the frame or imposed at 0 (UpdateObstacles skips its momenta when logs are muted).
  bool isStatic() const;
  // Obstacle::update (Obstacle.cpp:314-345): position, absolute position and quaternion.
  void update(double dt, const std::array<double, 3>& uinf);
  // Obstacle::save / restart (Obstacle.cpp:402-450): <filename>.txt with time, position,
  // absPos, quaternion, transVel, angVel and the 2-D angle.
  void save(const std::string& filename, double time) const;
  bool restart(const std::string& filename);
  // Sphere::computeVelocities then Obstacle::computeVelocities (Sphere.cpp:147-159,
  // Obstacle.cpp:153-264) on UpdateObstacles' 29 momentum sums at simulation time `time`.
  void computeVelocities(const double M[29], bool implicit, double time);
};

class Operator;

struct SimulationData {
  // simulation status (SimulationData.h:76-79)
  int step = 0, nsteps = 0;
  double time = 0, endTime = 0, dt = 0;
  // grid (SimulationData.h:86-94)
  int bpdx = -1, bpdy = -1, bpdz = -1;
  Real maxextent = 1;
  std::array<Real, 3> extent = {{1, 0, 0}};
  bool bImplicitPenalization = false;
  Real hmin = 0, hmax = 0, hmean = 0;
  // flow (SimulationData.h:97-98)
  std::array<Real, 3> uinf = {{0, 0, 0}};
  double nu = 0, CFL = 0, lambda = -1, DLM = 1;
  // initial conditions (SimulationData.h:101-106)
  std::string initCond = "zero";
  std::string spectralIC = "";
  double k0 = 0, tke0 = 0;
  uint64_t icSeed = 20260101;
  double icUrms = 0.0;  // > 0: the HIT IC scaled to this u_rms; 0: the target spectrum's energy
  // forcing (SimulationData.h:109-115)
  Real uMax_forced = 0, uMax_measured = 0;
  bool bChannelFixedMassFlux = false;
  bool spectralForcing = false;
  double turbKinEn_target = 0, enInjectionRate = 0, dissipationRate = 0, actualInjectionRate = 0;
  // settings (SimulationData.h:141-147)
  int freqDiagnostics = 0;
  int rampup = 100;
  // serialization (SimulationData.h:148-153, SimulationData.cpp:96-115)
  int saveFreq = 0;
  double saveTime = 0, nextSaveTime = 0;
  bool bDump = false, b3Ddump = true, bRestart = false;
  // analysis (SimulationData.h:122-131); nu_sgs and cs2_avg stay 0 (no SGS model here)
  std::string analysis = "";
  double timeAnalysis = 0, nextAnalysisTime = 0;
  int freqAnalysis = 0;
  double grad_mean = 0, grad_std = 0, nu_sgs = 0, cs2_avg = 0;
  bool verbose = false, muteAll = false;
  // Operator::check (Operator.h:23-50) after PressureRHS, PressureProjection and
  // ComputeDissipation; the reference compiles it into debug builds, here -checkNaN 1 turns it on
  bool bCheckNaN = false;
  std::string path4serialization = "./";
  BCflag BCx_flag = dirichlet, BCy_flag = dirichlet, BCz_flag = dirichlet;
  bool bUseFourierBC = false;
  bool bUseUnboundedBC = false;  // SimulationData.h:157: every face freespace
  // device-specific: context precision (4/8), step decomposition, advection scheme
  int precision = 8;
  int decomposition = CUP_PLAIN;
  int scheme = CUP_CENTRAL_EULER;
  int device = 0;

  // runtime objects
  cup_ctx* ctx = nullptr;
  std::vector<std::shared_ptr<Sphere>> obstacles;
  std::vector<std::unique_ptr<Operator>> pipeline;
  std::shared_ptr<BufferedLogger> logger;
  std::shared_ptr<Profiler> profiler;  // timing on for verbose runs

  SimulationData();
  SimulationData(int argc, const char* const* argv);
  SimulationData(const SimulationData&);
  SimulationData& operator=(const SimulationData&);
  ~SimulationData();
  void _preprocessArguments();
  void setCells(int nx, int ny, int nz);
  Real uniformH() const { return hmean; }
  // SimulationData::startProfiler / stopProfiler / printResetProfiler (SimulationData.cpp:293-305)
  void startProfiler(const std::string& name) const;
  void stopProfiler() const;
  void printResetProfiler();
};

class Operator
{
 protected:
  SimulationData& sim;
  void check(int rc, const char* what);
  // Operator::check (Operator.h:23-50) when sim.bCheckNaN: a NaN in u, v, w or p prints
  // "GenericCoordinator::check isnan <infoText>" and throws std::runtime_error with that text
  // (the reference aborts the MPI job).
  void check(const std::string& infoText);

 public:
  explicit Operator(SimulationData& s) : sim(s) {}
  virtual ~Operator() = default;
  virtual void operator()(double dt) = 0;
  virtual std::string getName() = 0;
};

class InitialConditions : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "IC"; }
};

class CreateObstacles : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  void build();
  std::string getName() override { return "CreateObstacles"; }
  bool created = false;
  std::vector<char> velHeader;  // per body: computedVelocity's header written
};

class AdvectionDiffusion : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "AdvectionDiffusion"; }
};

// ExternalForcing (ExternalForcing.cpp:35-48): the uniform pressure gradient that drives a
// mean flow of centreline velocity uMax_forced, u += 8 uMax_forced nu dt / H^2.
class ExternalForcing : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "ExternalForcing"; }
};

class SpectralForcing : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "SpectralForcing"; }
  double totalKinEnPrev = 0;
};

class PressureRHS : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "PressureRHS"; }
};

class PressureProjection : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "PressureProjection"; }
};

class UpdateObstacles : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "UpdateObstacles"; }
};

class Penalization : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "Penalization"; }
};

class ComputeForces : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  void write(int k);  // body k's surface forces and its files
  std::string getName() override { return "ComputeForces"; }
};

class ComputeDissipation : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "ComputeDissipation"; }
  std::array<double, 20> last = {};
};

// Analysis (Analysis.cpp:96-196), HIT branch: gradient statistics and the spectral
// analysis every fAnalysis steps or tAnalysis time units, written to
// analysis/spectralAnalysis_<nFile, 9 digits> (SpectralAnalysis.cpp:124-205).
class Analysis : public Operator
{
 public:
  using Operator::Operator;
  void operator()(double dt) override;
  std::string getName() override { return "Analysis"; }
  std::array<double, 9> last = {};  // tke eps uprime lambda Re_lambda l_integral tau_integral mean std
  std::vector<double> spectrum;
  int files = 0;
};

// SpectralIcGenerator::_generateTarget: the target spectrum table (K, E) of -spectralIC.
void spectralTarget(const SimulationData& sim, std::vector<double>& K, std::vector<double>& E);

class Simulation
{
 public:
  SimulationData sim;

  explicit Simulation(const SimulationData& sd);
  Simulation(int argc, const char* const* argv);
  virtual ~Simulation();

  void addObstacle(std::shared_ptr<Sphere> s);
  void setupOperators();
  virtual void run();
  double calcMaxTimestep();
  bool timestep(double dt);
  // The FUSED decomposition's loop: dt, step and time stay on the device; returns true
  // when the run is finished.
  bool advance(int nsteps);
  void reset();

 private:
  void setupGrid();
  void _init();
  void _ic();
  void syncStepState();

 public:
  // Simulation::_serialize / _deserialize (Simulation.cpp:426-580): <name>.status and
  // restart.status as the reference writes them, the obstacles' .txt files, and the fields in
  // vel_/pres_/chi_<name>.cupf (write_field_file); a restart reads the velocity and the bodies.
  void _serialize(const std::string& append = "");
  void _deserialize();
};

}  // namespace cubismup3d
