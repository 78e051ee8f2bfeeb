// Command-line driver with the reference mains' interface (xyz.cpp,
// kitaev.cpp, potts.cpp): `dmrgx <model> l=.. m=.. n=.. s=.. <params>` runs
// the reference's schedule on the GPU and writes energy<suffix> and
// singular<suffix> in the working directory.  Extra words select the
// arithmetic (prec=fp64|fp32|bf16x3), the schedule (schedule=dmrg|dmrg2; the
// mains call dmrg), the sign rule (signs=physical|reference) and periodic
// edge terms (pbc=1, the reference's PBC build).
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "dmrgx.h"
#include "dmrgx_driver.h"

namespace {

struct Model {
  const char* name;
  int id;
  std::vector<const char*> params;
};

const Model kModels[] = {
    {"xyz", DMRGX_MODEL_XYZ, {"jx", "jy", "jz"}},           // xyz.cpp:25
    {"kitaev", DMRGX_MODEL_KITAEV, {"t", "p", "u"}},        // kitaev.cpp:25
    {"potts3", DMRGX_MODEL_POTTS3, {"f", "j", "p", "t"}},   // potts.cpp:27
    {"potts2", DMRGX_MODEL_POTTS2, {"h", "j"}},             // H_Potts2(h, j), src/operators.cpp
    {"4e", DMRGX_MODEL_4E, {"t1", "t3", "f0", "f1", "frac"}},  // 4e.cpp:39
};

// the labels of the 4e main's correlation lines (4e.cpp:98-196)
const char* label_4e(int i, int n_tensors) {
  static const char* ends[13] = {"L1 R1: ", "L1 R2: ", "L2 R1: ", "L2 R2: ", "up: L1 R1: ", "up: L1 R2: ",
                                 "up: L2 R1: ", "up: L2 R2: ", "down: L1 R1: ", "down: L1 R2: ", "down: L2 R1: ",
                                 "down: L2 R2: ", "Normalization: "};
  if (i < 13) return ends[i];
  return i < 13 + 2 * (n_tensors - 2) ? "Cooper-pair: " : "pair-pair: ";
}

std::string word(int argc, char** argv, const char* key, const char* def) {
  const size_t n = std::strlen(key);
  for (int i = 0; i < argc; ++i)
    if (!std::strncmp(argv[i], key, n) && argv[i][n] == '=') return argv[i] + n + 1;
  return def;
}

int fail(const char* what) {
  std::fprintf(stderr, "%s failed: %s\n", what, dmrgx_last_error());
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s xyz|kitaev|potts3|potts2|4e l=<sites> m=<cutoff> n=<sweeps> s=<sector> "
                 "<model parameters> [prec=fp64|fp32|bf16x3] [schedule=dmrg|dmrg2] [signs=physical|reference] [warm=0|1] [pbc=0|1]\n",
                 argv[0]);
    return 2;
  }
  const Model* model = nullptr;
  for (const auto& m : kModels)
    if (!std::strcmp(argv[1], m.name)) model = &m;
  if (!model) {
    std::fprintf(stderr, "unknown model %s\n", argv[1]);
    return 2;
  }
  std::time_t start = std::time(nullptr);
  dmrgx_run_config cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  double p[DMRGX_NPARAM] = {};
  char suffix[4096];
  if (dmrgx_parse_args(argc, argv, (int)model->params.size(), model->params.data(), &cfg.L, &cfg.m, &cfg.sweeps,
                       &cfg.sector, p, suffix, sizeof(suffix)))
    return fail("dmrgx_parse_args");
  // src/functions.cpp:68-72
  std::printf("\n ***** The parameters for this running ***** \n\n  sites(l) = %d\t cutoff(m) = %d\t sweep(n) = %d"
              "\t symmetry_section(s) = %d\n", cfg.L, cfg.m, cfg.sweeps, cfg.sector);
  for (size_t i = 0; i < model->params.size(); ++i) std::printf("\t %s = %g\n", model->params[i], p[i]);
  // the reference's PBC build (xyz.cpp:39-44), here a run-time switch
  cfg.pbc = word(argc, argv, "pbc", "0") == "1" ? 1 : 0;
  std::printf(cfg.pbc ? "\nUsing periodic boundary condition\n\n" : "\nUsing open boundary condition\n\n");
  cfg.model = model->id;
  for (int i = 0; i < DMRGX_NPARAM; ++i) cfg.p[i] = p[i];
  const std::string prec = word(argc, argv, "prec", "fp64");
  cfg.prec = prec == "bf16x3" ? DMRGX_BF16X3 : prec == "fp32" ? DMRGX_FP32 : DMRGX_FP64;
  cfg.schedule = word(argc, argv, "schedule", "dmrg") == "dmrg2" ? DMRGX_SCHEDULE_DMRG2 : DMRGX_SCHEDULE_DMRG;
  cfg.sign_rule = word(argc, argv, "signs", "physical") == "reference" ? DMRGX_SIGNS_REFERENCE
                                                                       : DMRGX_SIGNS_PHYSICAL;
  cfg.ladder = 1;
  cfg.seed = 1234;
  // FP64 residual tolerance: ARPACK's 1e-10 for an FP64 run; the FP64
  // refinement of a reduced-precision run goes half a digit further, which
  // keeps the block SVD's Jacobi sweeps at the FP64 run's count
  cfg.tol = cfg.prec == DMRGX_FP64 ? 1e-10 : 5e-11;
  cfg.tol_lo = 3e-5;  // the reduced-precision Lanczos solve ahead of the FP64 refinement
  cfg.refine_k = 8;
  cfg.output = suffix;
  cfg.warm_start = word(argc, argv, "warm", "0") == "1" ? 1 : 0;
  if (dmrgx_init(0)) return fail("dmrgx_init");
  dmrgx_run_stats st;
  int n_steps = 0;
  dmrgx_state* state = nullptr;
  const bool keep = model->id == DMRGX_MODEL_KITAEV || model->id == DMRGX_MODEL_4E;
  if (dmrgx_run_dmrg_state(&cfg, &st, nullptr, 0, &n_steps, keep ? &state : nullptr))
    return fail("dmrgx_run_dmrg");
  std::printf("Energy:\n%.10f\n", st.energy);
  if (state && model->id == DMRGX_MODEL_4E) {
    // 4e.cpp:71-196: the correlation functions, printed and written to data<suffix>
    std::printf("\n********** Calculate expectation values **********\n\n");
    std::vector<double> rows(3 * 4096);
    int n = 0;
    if (dmrgx_state_correlations_4e(state, rows.data(), 4096, &n)) return fail("dmrgx_state_correlations_4e");
    FILE* f = std::fopen((std::string("data") + suffix).c_str(), "w");
    if (!f) return fail("opening the data file");
    const int tensors = cfg.L / 2;
    for (int i = 0; i < n; ++i) {
      const double* r = &rows[3 * i];
      if (i < 13) std::printf("%s(%.10g,%.10g)\n", label_4e(i, tensors), r[1], r[2]);
      else std::printf("%s%.10g \t (%.10g,%.10g)\n", label_4e(i, tensors), r[0], r[1], r[2]);
      std::fprintf(f, "%.10g\t%.10g\t%.10g\n", r[0], r[1], r[2]);
    }
    std::fclose(f);
    dmrgx_state_destroy(state);
  } else if (state) {  // kitaev.cpp:50-52
    double nrm[2];
    if (dmrgx_state_norm(state, nrm)) return fail("dmrgx_state_norm");
    std::printf("Normalization: %.10f\n", nrm[0]);
    dmrgx_state_destroy(state);
  }
  std::printf("Time used: %.0f\n", std::difftime(std::time(nullptr), start));
  return 0;
}
