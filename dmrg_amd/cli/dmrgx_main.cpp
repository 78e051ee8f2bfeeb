// Command-line driver with the reference mains' interface (xyz.cpp,
// kitaev.cpp, potts.cpp): `dmrgx <model> l=.. m=.. n=.. s=.. <params>` runs
// the reference's schedule on the GPU and writes energy<suffix> and
// singular<suffix> in the working directory.  Extra words select the
// arithmetic (prec=fp64|fp32|bf16x3), the schedule (schedule=dmrg|dmrg2; the
// mains call dmrg) and the sign rule (signs=physical|reference).
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
};

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
    std::fprintf(stderr, "usage: %s xyz|kitaev|potts3|potts2 l=<sites> m=<cutoff> n=<sweeps> s=<sector> "
                 "<model parameters> [prec=fp64|fp32|bf16x3] [schedule=dmrg|dmrg2] [signs=physical|reference] [warm=0|1]\n",
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
  double p[4] = {0, 0, 0, 0};
  char suffix[4096];
  if (dmrgx_parse_args(argc, argv, (int)model->params.size(), model->params.data(), &cfg.L, &cfg.m, &cfg.sweeps,
                       &cfg.sector, p, suffix, sizeof(suffix)))
    return fail("dmrgx_parse_args");
  // src/functions.cpp:68-72
  std::printf("\n ***** The parameters for this running ***** \n\n  sites(l) = %d\t cutoff(m) = %d\t sweep(n) = %d"
              "\t symmetry_section(s) = %d\n", cfg.L, cfg.m, cfg.sweeps, cfg.sector);
  for (size_t i = 0; i < model->params.size(); ++i) std::printf("\t %s = %g\n", model->params[i], p[i]);
  std::printf("\nUsing open boundary condition\n\n");
  cfg.model = model->id;
  for (int i = 0; i < 4; ++i) cfg.p[i] = p[i];
  const std::string prec = word(argc, argv, "prec", "fp64");
  cfg.prec = prec == "bf16x3" ? DMRGX_BF16X3 : prec == "fp32" ? DMRGX_FP32 : DMRGX_FP64;
  cfg.schedule = word(argc, argv, "schedule", "dmrg") == "dmrg2" ? DMRGX_SCHEDULE_DMRG2 : DMRGX_SCHEDULE_DMRG;
  cfg.sign_rule = word(argc, argv, "signs", "physical") == "reference" ? DMRGX_SIGNS_REFERENCE
                                                                       : DMRGX_SIGNS_PHYSICAL;
  cfg.ladder = 1;
  cfg.seed = 1234;
  cfg.tol = 1e-10;
  cfg.tol_lo = 1e-4;
  cfg.refine_k = 8;
  cfg.output = suffix;
  cfg.warm_start = word(argc, argv, "warm", "0") == "1" ? 1 : 0;
  if (dmrgx_init(0)) return fail("dmrgx_init");
  dmrgx_run_stats st;
  int n_steps = 0;
  dmrgx_state* state = nullptr;
  if (dmrgx_run_dmrg_state(&cfg, &st, nullptr, 0, &n_steps, model->id == DMRGX_MODEL_KITAEV ? &state : nullptr))
    return fail("dmrgx_run_dmrg");
  std::printf("Energy:\n%.10f\n", st.energy);
  if (state) {  // kitaev.cpp:50-52
    double nrm[2];
    if (dmrgx_state_norm(state, nrm)) return fail("dmrgx_state_norm");
    std::printf("Normalization: %.10f\n", nrm[0]);
    dmrgx_state_destroy(state);
  }
  std::printf("Time used: %.0f\n", std::difftime(std::time(nullptr), start));
  return 0;
}
