// cosmo_host.h -- C++ host classes over the C ABI (include/cosmo_gpu.h)
// with the reference's class and method names, so that a CosmoGRaPH
// simulation driver reads the same: ConfigParser (utils/ConfigParser.h),
// IOData (IO/IOData.h), BSSN (components/bssn/bssn.h:53-170), Lambda
// (components/Lambda/lambda.h), and the sims CosmoSim / StaticSim / DustSim /
// VacuumSim (sims/sim.h, sims/static.h, sims/dust.h).
//
// The grid lives on the GPU; host arrays are copies taken on request.  The
// matter sources (Static::addBSSNSrc, Dust::addBSSNSrc, Lambda) are formed
// inside the stage kernel, so clearSrc and the addBSSNSrc calls of the
// reference's step sequence have nothing left to do here.  Errors follow
// the reference's conventions: `throw -1` for configuration/IC errors,
// `throw 2` for an unknown simulation type, `throw 10` for NaN.
#pragma once

#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/cosmo_gpu.h"

namespace cosmo {

// utils/ConfigParser.h: whitespace-separated "param = value" triples
class ConfigParser {
 public:
  ConfigParser() = default;
  explicit ConfigParser(const std::string& fname) { parse(fname); }
  void parse(const std::string& fname);
  std::string getFileName() const { return fileName; }
  std::string operator[](const std::string& param) const;                      // required
  std::string operator()(const std::string& param, const std::string& dflt) const;  // optional
  bool has(const std::string& param) const { return config.count(param) != 0; }
  // parameter names beginning with `prefix`, in sorted order
  std::vector<std::string> keys(const std::string& prefix) const;

 private:
  std::map<std::string, std::string> config;
  std::string fileName;
};

// IO/IOData.h: output directory (renamed "<dir>.<n>" if it holds a run)
// and log file
class IOData {
 public:
  explicit IOData(const std::string& output_dir, bool quiet = false);
  ~IOData();
  std::string dir() const { return output_dir; }
  void log(const std::string& msg);

 private:
  std::string output_dir;
  FILE* logfile = nullptr;
  bool quiet;
};

// Grid and physics parameters from the config keys the reference reads
// (cosmo.cc:41-42, cosmo_macros.h:12-40, bssn.cc:15-22,
// BSSNGaugeHandler.h:175), plus the keys of this build: N (or NX/NY/NZ),
// stencil_order (the reference's compile-time STENCIL_ORDER, default 8),
// H_LEN_FRAC (default 0.5), gpu_precision (fp64 | fp32_split), gpu_device.
cg_params params_from_config(const ConfigParser& cfg);

// components/bssn/bssn.h:53-170 on the GPU
class BSSN {
 public:
  explicit BSSN(const ConfigParser& cfg);
  explicit BSSN(const cg_params& p);
  ~BSSN();
  BSSN(const BSSN&) = delete;
  BSSN& operator=(const BSSN&) = delete;

  void stepInit();
  void RKEvolve();
  void K1Finalize();
  void K2Finalize();
  void K3Finalize();
  void K4Finalize();
  void clearSrc() {}  // sources are formed inside the stage kernel
  // one or more whole RK4 steps in the StaticSim/DustSim sequence, fused
  void step(int nsteps = 1);
  void setConstraintCalcs(double H[8], double M[8], double G[7], double A[7], double S[7]);
  // bssn.cc:903-938: H and M_axis at the N_axis points (x, n1, n2)
  void set1DConstraintOutput(double* H, double* M, int axis, long n1, long n2);
  // the values io_dump_strip writes (IO/io.cc:719-771); false if the field
  // map has no such array here
  bool strip(const std::string& name, int axis, long n1, long n2, std::vector<double>& out);
  // conformal_average, conformal_standard_deviation of a field and the
  // volume_average of DIFFphi_a (IO/io.cc:8-14, utils/math.h:1382-1470)
  void fieldStats(const std::string& name, double out[3]);
  // Fourier::powerDump's bin averages (utils/Fourier.h:147-220)
  std::vector<double> powerSpectrum(const std::string& name);
  // refreshes K_min, K_avg, avg_vol, rho_avg (bssn.cc:165-170)
  void setExtraFieldData();
  double K_min = 0, K_avg = 0, avg_vol = 1, rho_avg = 0;

  // the `fields` map: copies to and from the device
  std::vector<double> field(const std::string& name) const;
  void setField(const std::string& name, const std::vector<double>& v);
  long numNaNs(const std::string& name) const;

  // FRW<real_t> owned by BSSN (utils/FRW.h)
  void frwSet(double phi, double K);
  void frwAddFluid(double rho, double w);
  void frwGet(double out_phi_K_rho_S[4]) const;

  long points() const { return (long)p.nx * p.ny * p.nz; }
  const cg_params& params() const { return p; }
  cg_ctx* context() { return ctx; }

 private:
  void check(int rc) const;
  cg_params p{};
  cg_ctx* ctx = nullptr;
};

// components/Lambda/lambda.h
class Lambda {
 public:
  void setLambda(double l) { lambda = l; }
  double getLambda() const { return lambda; }
  void addBSSNSource(BSSN* bssn) const;

 private:
  double lambda = 0.0;
};

// sims/sim.h: the step loop, NaN guard and output of every simulation type
class CosmoSim {
 public:
  explicit CosmoSim(const ConfigParser& cfg);
  virtual ~CosmoSim();
  virtual void init();
  virtual void setICs();
  virtual void runStep();
  void run();
  void runCommonStepTasks();
  void outputStep();
  long steps_run() const { return rk_steps; }
  BSSN* bssn() { return bssnSim; }

 protected:
  const ConfigParser& cfg;
  IOData* iodata = nullptr;
  BSSN* bssnSim = nullptr;
  Lambda* lambda = nullptr;
  long step = 0, num_steps = 0, rk_steps = 0;
  double t = 0.0, dt = 0.0;
  double t_rk = 0.0, t_out = 0.0;
};

// sims/static.h: BSSN + static dust (+ Lambda); the lapse must be Static or
// ConformalFLRW (sims/static.cc:8-20)
class StaticSim : public CosmoSim {
 public:
  explicit StaticSim(const ConfigParser& cfg);
};

// sims/dust.h: BSSN + the dust fluid (+ Lambda); no shift (sims/dust.cc:8-20)
class DustSim : public CosmoSim {
 public:
  explicit DustSim(const ConfigParser& cfg);
};

// sims/vacuum.h: BSSN without matter
class VacuumSim : public CosmoSim {
 public:
  explicit VacuumSim(const ConfigParser& cfg) : CosmoSim(cfg) {}
};

// cosmo.cc: the simulation named by `simulation_type`; throws 2 otherwise
CosmoSim* make_sim(const ConfigParser& cfg);

}  // namespace cosmo
