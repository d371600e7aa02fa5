// C ABI of libinductioneq_hip: program build from the reference settings string,
// kernel-list execution by reference kernel name, buffer helpers, fused stepping.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/inductioneq.h"
#include "ieq_adjoint.hpp"
#include "ieq_cg.hpp"
#include "ieq_device.hpp"
#include "ieq_diag.hpp"
#include "ieq_fused.hpp"
#include "ieq_generic.hpp"
#include "ieq_tiled.hpp"

using ieq::Params;

namespace {

thread_local std::string g_build_error;

struct OpMeta {
  const char* name;
  int order, nb, w, w_hod;
  int lap_nb[3];
};

const OpMeta kOps[6] = {
    {"2ndOrder", IEQ_2ND_ORDER, IEQ_2ND_NB, IEQ_2ND_W, IEQ_2ND_W_HOD, {IEQ_2ND_LAP_WS_D0_NB, IEQ_2ND_LAP_NS_D0_NB, IEQ_2ND_LAP_WS_LN_NB}},
    {"4thOrder", IEQ_4TH_ORDER, IEQ_4TH_NB, IEQ_4TH_W, IEQ_4TH_W_HOD, {IEQ_4TH_LAP_WS_D0_NB, IEQ_4TH_LAP_NS_D0_NB, IEQ_4TH_LAP_WS_LN_NB}},
    {"6thOrder", IEQ_6TH_ORDER, IEQ_6TH_NB, IEQ_6TH_W, IEQ_6TH_W_HOD, {IEQ_6TH_LAP_WS_D0_NB, IEQ_6TH_LAP_NS_D0_NB, IEQ_6TH_LAP_WS_LN_NB}},
    {"2ndOrderExtended", IEQ_2ND_EXT_ORDER, IEQ_2ND_EXT_NB, IEQ_2ND_EXT_W, IEQ_2ND_EXT_W_HOD,
     {IEQ_2ND_EXT_LAP_WS_D0_NB, IEQ_2ND_EXT_LAP_NS_D0_NB, IEQ_2ND_EXT_LAP_WS_LN_NB}},
    {"4thOrderExtended", IEQ_4TH_EXT_ORDER, IEQ_4TH_EXT_NB, IEQ_4TH_EXT_W, IEQ_4TH_EXT_W_HOD,
     {IEQ_4TH_EXT_LAP_WS_D0_NB, IEQ_4TH_EXT_LAP_NS_D0_NB, IEQ_4TH_EXT_LAP_WS_LN_NB}},
    {"6thOrderExtended", IEQ_6TH_EXT_ORDER, IEQ_6TH_EXT_NB, IEQ_6TH_EXT_W, IEQ_6TH_EXT_W_HOD,
     {IEQ_6TH_EXT_LAP_WS_D0_NB, IEQ_6TH_EXT_LAP_NS_D0_NB, IEQ_6TH_EXT_LAP_WS_LN_NB}},
};

const char* kTestcases[8] = {"rotation_2D",        "rotation_3D",      "hall_periodic",        "hall_travelling_wave",
                             "confined_domain",    "alfven_periodic_2D", "gaussian_rotation_3D", "uniform_flow"};

// One reference stage-family kernel (kernel_time_integrator.cl).
enum Kind { K_STAGE, K_2R, K_2N, K_S06, K_TIME, K_CALC_TIME, K_U, K_CURL };
struct StageDesc {
  const char* name;
  int nf;     // 2 or 3 field registers in the argument list
  Kind kind;
  int out, P;  // K_STAGE: out = cP*P + cQ*Q + cD*DT*DB(Y);  K_2R: X=P, O=out, D=Q;  K_2N: X=out, D=P
  double cP;
  int Q;
  double cQ;
  int Y;
  double cD;   // K_STAGE: multiple of DT;  K_2R/K_2N: first/second update coefficient in cP/cQ
  double tc;   // time[1] = time[0] + tc*DT, < 0: none
};

#define CK(n, d) ((double)n / (double)d)
const StageDesc kStages[] = {
    {"SSPRK33_1", 3, K_STAGE, 1, 0, 1.0, -1, 0, 0, 1.0, -1},
    {"SSPRK33_2a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 1.0},
    {"SSPRK33_2b", 3, K_STAGE, 2, 0, 0.75, 1, 0.25, 1, 0.25, -1},
    {"SSPRK33_3a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 0.5},
    {"SSPRK33_3b", 3, K_STAGE, 0, 0, 1.0 / 3.0, 2, 2.0 / 3.0, 2, 2.0 / 3.0, -1},
    {"SSPRK104_01", 3, K_STAGE, 1, 0, 1.0, -1, 0, 0, 1.0 / 6.0, -1},
    {"SSPRK104_02a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 1.0 / 6.0},
    {"SSPRK104_02b", 3, K_STAGE, 2, 1, 1.0, -1, 0, 1, 1.0 / 6.0, -1},
    {"SSPRK104_03a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 1.0 / 3.0},
    {"SSPRK104_03b", 3, K_STAGE, 1, 2, 1.0, -1, 0, 2, 1.0 / 6.0, -1},
    {"SSPRK104_04a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 0.5},
    {"SSPRK104_04b", 3, K_STAGE, 2, 1, 1.0, -1, 0, 1, 1.0 / 6.0, -1},
    {"SSPRK104_05a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 2.0 / 3.0},
    {"SSPRK104_05b", 3, K_STAGE, 1, 2, 1.0, -1, 0, 2, 1.0 / 6.0, -1},
    {"SSPRK104_06", 3, K_S06, -1, -1, 0, -1, 0, -1, 0, 1.0 / 3.0},
    {"SSPRK104_07", 3, K_STAGE, 0, 1, 1.0, -1, 0, 1, 1.0 / 6.0, -1},
    {"SSPRK104_08a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 0.5},
    {"SSPRK104_08b", 3, K_STAGE, 1, 0, 1.0, -1, 0, 0, 1.0 / 6.0, -1},
    {"SSPRK104_09a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 2.0 / 3.0},
    {"SSPRK104_09b", 3, K_STAGE, 0, 1, 1.0, -1, 0, 1, 1.0 / 6.0, -1},
    {"SSPRK104_10a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 5.0 / 6.0},
    {"SSPRK104_10b", 3, K_STAGE, 1, 0, 1.0, -1, 0, 0, 1.0 / 6.0, -1},
    {"SSPRK104_11a", 3, K_TIME, -1, -1, 0, -1, 0, -1, 0, 1.0},
    {"SSPRK104_11b", 3, K_STAGE, 0, 2, 1.0, 1, 0.6, 1, 0.6 / 6.0, -1},
    {"KennedyCarpenterLewis2R54C_1a", 3, K_STAGE, 2, -1, 0, -1, 0, 0, 1.0, -1},
    {"KennedyCarpenterLewis2R54C_1b", 3, K_2R, 1, 0, 0.22502245872571303, 2, -0.17379315208537388, -1, 0, 0.22502245872571303},
    {"KennedyCarpenterLewis2R54C_2a", 3, K_STAGE, 2, -1, 0, -1, 0, 0, 1.0, -1},
    {"KennedyCarpenterLewis2R54C_2b", 3, K_2R, 0, 1, 0.5440433129514047, 2, -0.1630884872250029, -1, 0, 0.5952726195917439},
    {"KennedyCarpenterLewis2R54C_3a", 3, K_STAGE, 2, -1, 0, -1, 0, 1, 1.0, -1},
    {"KennedyCarpenterLewis2R54C_3b", 3, K_2R, 1, 0, 0.14456824349399464, 2, -0.5179208398863779, -1, 0, -1},
    {"KennedyCarpenterLewis2R54C_4a", 3, K_STAGE, 2, -1, 0, -1, 0, 0, 1.0, -1},
    {"KennedyCarpenterLewis2R54C_4b", 3, K_2R, 0, 1, 0.7866643421983568, 2, -0.19416305717199442, -1, 0, 0.8454958781727144},
    {"KennedyCarpenterLewis2R54C_5", 3, K_STAGE, 0, 0, 1.0, -1, 0, 1, 0.34866717899927996, -1},
    {"CalvoFrancoRandez2R64_1a", 3, K_STAGE, 2, -1, 0, -1, 0, 0, 1.0, -1},
    {"CalvoFrancoRandez2R64_1b", 3, K_2R, 1, 0, 0.10893125722541, 2, 0.17985400977138, -1, 0, 0.28878526699679},
    {"CalvoFrancoRandez2R64_2a", 3, K_STAGE, 2, -1, 0, -1, 0, 1, 1.0, -1},
    {"CalvoFrancoRandez2R64_2b", 3, K_2R, 1, 0, 0.13201701492152, 2, 0.14081893152111, -1, 0, 0.38176720366804},
    {"CalvoFrancoRandez2R64_3a", 3, K_STAGE, 2, -1, 0, -1, 0, 1, 1.0, -1},
    {"CalvoFrancoRandez2R64_3b", 3, K_2R, 1, 0, 0.38911623225517, 2, 0.08255631629428, -1, 0, 0.71262082069639},
    {"CalvoFrancoRandez2R64_4a", 3, K_STAGE, 2, -1, 0, -1, 0, 1, 1.0, -1},
    {"CalvoFrancoRandez2R64_4b", 3, K_2R, 1, 0, -0.59203884581148, 2, 0.65804425034331, -1, 0, 0.69606990893393},
    {"CalvoFrancoRandez2R64_5a", 3, K_STAGE, 2, -1, 0, -1, 0, 1, 1.0, -1},
    {"CalvoFrancoRandez2R64_5b", 3, K_2R, 1, 0, 0.47385028714844, 2, 0.31862993413251, -1, 0, 0.83050587987157},
    {"CalvoFrancoRandez2R64_6", 3, K_STAGE, 0, 0, 1.0, -1, 0, 1, 0.48812405426094, -1},
    {"CarpenterKennedy2N54_1a", 2, K_STAGE, 1, -1, 0, -1, 0, 0, 1.0, -1},
    {"CarpenterKennedy2N54_1b", 2, K_2N, 0, 1, 0, -1, CK(1432997174477.0, 9575080441755.0), -1, 0, CK(1432997174477.0, 9575080441755.0)},
    {"CarpenterKennedy2N54_2a", 2, K_STAGE, 1, 1, CK(-567301805773.0, 1357537059087.0), -1, 0, 0, 1.0, -1},
    {"CarpenterKennedy2N54_2b", 2, K_2N, 0, 1, 0, -1, CK(5161836677717.0, 13612068292357.0), -1, 0, CK(2526269341429.0, 6820363962896.0)},
    {"CarpenterKennedy2N54_3a", 2, K_STAGE, 1, 1, CK(-2404267990393.0, 2016746695238.0), -1, 0, 0, 1.0, -1},
    {"CarpenterKennedy2N54_3b", 2, K_2N, 0, 1, 0, -1, CK(1720146321549.0, 2090206949498.0), -1, 0, CK(2006345519317.0, 3224310063776.0)},
    {"CarpenterKennedy2N54_4a", 2, K_STAGE, 1, 1, CK(-3550918686646.0, 2091501179385.0), -1, 0, 0, 1.0, -1},
    {"CarpenterKennedy2N54_4b", 2, K_2N, 0, 1, 0, -1, CK(3134564353537.0, 4481467310338.0), -1, 0, CK(2802321613138.0, 2924317926251.0)},
    {"CarpenterKennedy2N54_5a", 2, K_STAGE, 1, 1, CK(-1275806237668.0, 842570457699.0), -1, 0, 0, 1.0, -1},
    {"CarpenterKennedy2N54_5b", 2, K_2N, 0, 1, 0, -1, CK(2277821191437.0, 14882151754819.0), -1, 0, -1},
    {"ToulorgeDesmet2N84F_1a", 2, K_STAGE, 1, -1, 0, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_1b", 2, K_2N, 0, 1, 0, -1, 0.08037936882736950, -1, 0, 0.08037936882736950},
    {"ToulorgeDesmet2N84F_2a", 2, K_STAGE, 1, 1, -0.5534431294501569, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_2b", 2, K_2N, 0, 1, 0, -1, 0.5388497458569843, -1, 0, 0.3210064250338430},
    {"ToulorgeDesmet2N84F_3a", 2, K_STAGE, 1, 1, 0.01065987570203490, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_3b", 2, K_2N, 0, 1, 0, -1, 0.01974974409031960, -1, 0, 0.3408501826604660},
    {"ToulorgeDesmet2N84F_4a", 2, K_STAGE, 1, 1, -0.5515812888932000, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_4b", 2, K_2N, 0, 1, 0, -1, 0.09911841297339970, -1, 0, 0.3850364824285470},
    {"ToulorgeDesmet2N84F_5a", 2, K_STAGE, 1, 1, -1.885790377558741, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_5b", 2, K_2N, 0, 1, 0, -1, 0.7466920411064123, -1, 0, 0.5040052477534100},
    {"ToulorgeDesmet2N84F_6a", 2, K_STAGE, 1, 1, -5.701295742793264, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_6b", 2, K_2N, 0, 1, 0, -1, 1.679584245618894, -1, 0, 0.6578977561168540},
    {"ToulorgeDesmet2N84F_7a", 2, K_STAGE, 1, 1, 2.113903965664793, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_7b", 2, K_2N, 0, 1, 0, -1, 0.2433728067008188, -1, 0, 0.9484087623348481},
    {"ToulorgeDesmet2N84F_8a", 2, K_STAGE, 1, 1, -0.5339578826675280, -1, 0, 0, 1.0, -1},
    {"ToulorgeDesmet2N84F_8b", 2, K_2N, 0, 1, 0, -1, 0.1422730459001373, -1, 0, -1},
    {"calc_time_2_args", 2, K_CALC_TIME, -1, -1, 0, -1, 0, -1, 0, -1},
    {"calc_time_3_args", 3, K_CALC_TIME, -1, -1, 0, -1, 0, -1, 0, -1},
    {"calc_u_2_args", 2, K_U, -1, -1, 0, -1, 0, -1, 0, -1},
    {"calc_u_3_args", 3, K_U, -1, -1, 0, -1, 0, -1, 0, -1},
    {"calc_curlB_rho_1_2args", 2, K_CURL, -1, -1, 0, -1, 0, 0, 0, -1},
    {"calc_curlB_rho_2_2args", 2, K_CURL, -1, -1, 0, -1, 0, 1, 0, -1},
    {"calc_curlB_rho_1_3args", 3, K_CURL, -1, -1, 0, -1, 0, 0, 0, -1},
    {"calc_curlB_rho_2_3args", 3, K_CURL, -1, -1, 0, -1, 0, 1, 0, -1},
    {"calc_curlB_rho_3_3args", 3, K_CURL, -1, -1, 0, -1, 0, 2, 0, -1},
};
#undef CK

const StageDesc* find_stage(const char* name) {
  for (const auto& s : kStages)
    if (!std::strcmp(s.name, name)) return &s;
  return nullptr;
}

std::string module_base(const char* m) {
  std::string s(m);
  size_t sl = s.find_last_of('/');
  if (sl != std::string::npos) s = s.substr(sl + 1);
  size_t dot = s.find_last_of('.');
  if (dot != std::string::npos) s = s.substr(0, dot);
  return s;
}

}  // namespace

struct ieq_program {
  Params p{};
  ieq_program_info info{};
  std::set<std::string> modules;
  hipStream_t stream = nullptr;
  std::string err;
  ieq::FusedWork* fused = nullptr;
  ieq::CgWork* cg = nullptr;
  int last_path = -1;  // path of the last ieq_run / ieq_step: 0 reference kernels, 1 fused kernels, 2 reference
                       // sequence with the tiled stage kernel
  bool tiled_used = false;  // the current ieq_run / ieq_step launched a tiled stage
  double* diag = nullptr;  // ieq_diagnostics: 3 * kDiagBlocks partials + the 3 results

  bool has(const char* m) const { return modules.count(m) != 0; }
  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
};

namespace {

int hip_fail(ieq_program* prog, hipError_t e, const char* what) {
  return prog->fail(IEQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(prog, call)                                  \
  do {                                                       \
    hipError_t _e = (call);                                  \
    if (_e != hipSuccess) return hip_fail((prog), _e, #call); \
  } while (0)

// Scratch device memory and timing events owned by one call: released on every return path
// (hipFree waits for the work still using the buffer).
struct ScopedDev {
  void* p = nullptr;
  ~ScopedDev() {
    if (p) (void)hipFree(p);
  }
};
struct ScopedEvents {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~ScopedEvents() {
    for (hipEvent_t v : e)
      if (v) (void)hipEventDestroy(v);
  }
};

int parse_build(ieq_program* pr, const char* settings, const char* const* modules, int nmod) {
  std::map<std::string, std::string> defs;
  if (settings) {
    const char* s = settings;
    while (*s) {
      while (*s == ' ' || *s == '\t' || *s == '\n') ++s;
      const char* e = s;
      while (*e && *e != ' ' && *e != '\t' && *e != '\n') ++e;
      std::string tok(s, e - s);
      s = e;
      if (tok.rfind("-D", 0) == 0) {
        std::string body = tok.substr(2);
        size_t eq = body.find('=');
        if (eq == std::string::npos) defs[body] = "1";
        else defs[body.substr(0, eq)] = body.substr(eq + 1);
      }
    }
  }
  for (int i = 0; i < nmod; ++i)
    if (modules[i]) pr->modules.insert(module_base(modules[i]));

  auto has_def = [&](const char* k) { return defs.count(k) != 0; };
  auto num = [&](const char* k, double dflt, bool* ok) -> double {
    auto it = defs.find(k);
    if (it == defs.end()) {
      if (ok) *ok = false;
      return dflt;
    }
    char* end = nullptr;
    double v = std::strtod(it->second.c_str(), &end);
    if (ok) *ok = end && *end == 0;
    return v;
  };

  // REAL type
  std::string real = defs.count("REAL") ? defs["REAL"] : "";
  if (real != "double" && real != "float") return pr->fail(IEQ_ERR_SETTINGS, "REAL must be defined as double or float");
  ieq_program_info& I = pr->info;
  I.real_bytes = real == "double" ? 8 : 4;

  // operator set
  I.opset = -1;
  for (int k = 0; k < 6; ++k)
    if (pr->has(kOps[k].name)) {
      if (I.opset >= 0) return pr->fail(IEQ_ERR_SETTINGS, "more than one operator header in the module list");
      I.opset = k;
    }
  if (I.opset < 0) return pr->fail(IEQ_ERR_SETTINGS, "module list names no operator header (2ndOrder ... 6thOrderExtended)");
  const OpMeta& op = kOps[I.opset];
  I.order = op.order;
  I.num_bounds = op.nb;
  I.stencil_width = op.w;

  bool ok = true, okk;
  I.nx = (int)num("NODES_X", 0, &okk); ok &= okk;
  I.ny = (int)num("NODES_Y", 0, &okk); ok &= okk;
  I.nz = (int)num("NODES_Z", 0, &okk); ok &= okk;
  if (!ok || I.nx <= 0 || I.ny <= 0 || I.nz <= 0) return pr->fail(IEQ_ERR_SETTINGS, "NODES_X/Y/Z must be positive integers");
  I.dx = num("DX", 0, &okk); ok &= okk;
  I.dy = num("DY", 0, &okk); ok &= okk;
  I.dz = num("DZ", 0, &okk); ok &= okk;
  if (!ok || !(I.dx > 0) || !(I.dy > 0) || !(I.dz > 0)) return pr->fail(IEQ_ERR_SETTINGS, "DX/DY/DZ must be positive numbers");
  I.xmin = num("XMIN", 0, nullptr);
  I.ymin = num("YMIN", 0, nullptr);
  I.zmin = num("ZMIN", 0, nullptr);
  I.dt = num("DT", 0, nullptr);
  I.periodic = has_def("USE_PERIODIC");
  I.hall = has_def("USE_HALL");
  I.w_size = (int)num("W_SIZE", 256, nullptr);
  if (I.w_size <= 0) return pr->fail(IEQ_ERR_SETTINGS, "W_SIZE must be positive");
  I.ho_factor = num("HO_DISSIPATION_FACTOR", 1.0, nullptr);
  I.mp2min = num("MP2MIN", 0, nullptr);
  I.mp2max = num("MP2MAX", 0, nullptr);
  I.cmin = num("CMIN", 0, nullptr);
  I.cmax = num("CMAX", 0, nullptr);
  I.ux = num("UX", 0, nullptr);
  I.uy = num("UY", 0, nullptr);
  I.uz = num("UZ", 0, nullptr);

  // volume-term forms (induction_Eq_volume.cl:13-142, #error when absent)
  const bool volume = pr->has("induction_Eq_volume");
  I.form_uibj = has_def("USE_UIBJ_CENTRAL") ? 0 : has_def("USE_UIBJ_SPLIT") ? 1 : has_def("USE_UIBJ_PRODUCT") ? 2 : -1;
  I.form_source = has_def("USE_SOURCE_ZERO") ? 0 : has_def("USE_SOURCE_CENTRAL") ? 1 : has_def("USE_SOURCE_SPLIT") ? 2 : -1;
  I.form_ujbi = has_def("USE_UJBI_CENTRAL") ? 0 : has_def("USE_UJBI_SPLIT") ? 1 : has_def("USE_UJBI_PRODUCT") ? 2 : -1;
  if (volume) {
    if (I.form_uibj < 0) return pr->fail(IEQ_ERR_SETTINGS, "Error in induction_Eq_volume.cl: No discretization of 'D_j(u_i B_j)' specified!");
    if (I.form_source < 0) return pr->fail(IEQ_ERR_SETTINGS, "Error in induction_Eq_volume.cl: No discretization of 'u_i D_j B_j' specified!");
    if (I.form_ujbi < 0) return pr->fail(IEQ_ERR_SETTINGS, "Error in induction_Eq_volume.cl: No discretization of 'D_j(u_j B_i)' specified!");
  }

  // artificial dissipation (artificial_dissipation.cl:164-177)
  I.dissipation = 0;
  if (has_def("USE_ARTIFICIAL_DISSIPATION")) {
    if (!pr->has("artificial_dissipation"))
      return pr->fail(IEQ_ERR_SETTINGS, "USE_ARTIFICIAL_DISSIPATION requires the artificial_dissipation module");
    I.dissipation = has_def("USE_HIGH_ORDER_DISSIPATION") ? 1 : has_def("USE_FIRST_ORDER_DISSIPATION") ? 2
                    : has_def("USE_ADAPTIVE_DISSIPATION") ? 3 : 0;
    if (I.dissipation == 3 && !(has_def("MP2MIN") && has_def("MP2MAX") && has_def("CMIN") && has_def("CMAX")))
      return pr->fail(IEQ_ERR_SETTINGS, "adaptive dissipation needs MP2MIN, MP2MAX, CMIN and CMAX");
  }

  // testcase
  I.testcase = -1;
  for (int k = 0; k < 8; ++k)
    if (pr->has(kTestcases[k])) I.testcase = k;
  if (I.testcase < 0 && (pr->has("kernel_time_integrator") || pr->has("kernel_init")))
    return pr->fail(IEQ_ERR_SETTINGS, "module list names no testcase (needed by kernel_init / kernel_time_integrator)");
  if (I.testcase < 0) I.testcase = 7;

  // divergence cleaning (kernel_div_cleaning.cl:45-83, *Extended.h NS_D0 #error)
  I.laplace = -1;
  I.bnodes = (int)num("BNODES", 0, nullptr);
  if (pr->has("kernel_div_cleaning")) {
    I.laplace = has_def("USE_LAPLACE_WIDE_STENCIL_DIRICHLET") ? 0 : has_def("USE_LAPLACE_WIDE_STENCIL_LNS") ? 2
                : has_def("USE_LAPLACE_NARROW_STENCIL_DIRICHLET") ? 1 : -1;
    if (I.laplace < 0) return pr->fail(IEQ_ERR_SETTINGS, "Define a discretization of the Laplace operator for divergence cleaning!");
    if (I.laplace == 1 && op.lap_nb[1] == 0)
      return pr->fail(IEQ_ERR_SETTINGS, std::string("Error in include/") + op.name +
                                            ".h: Coefficients of the narrow stencil operator using homogeneous Dirichlet "
                                            "boundary conditions have not been derived.");
  }
  if (I.bnodes < 0 || 2 * I.bnodes >= std::min(I.nx, std::min(I.ny, I.nz)))
    return pr->fail(IEQ_ERR_SETTINGS, "BNODES out of range");

  // The periodic neighbour rule wraps once (utils.h:91-98); a grid narrower than the
  // widest stencil half-width would index outside the field.
  if (I.periodic) {
    int hw = (op.w - 1) / 2;
    int nbl = 0, wl = 0;
    if (I.laplace >= 0 && ieq::host_lap_table(I.opset, I.laplace, &nbl, &wl)) hw = std::max(hw, (wl - 1) / 2);
    const int nmin = std::min(I.nx, std::min(I.ny, I.nz));
    if (nmin < hw)
      return pr->fail(IEQ_ERR_SETTINGS, "periodic grid has fewer nodes per direction than the stencil half-width");
  }

  I.num_nodes = (long long)I.nx * I.ny * I.nz;
  I.num_nodes_pad = (I.num_nodes + I.w_size - 1) / I.w_size * I.w_size;
  I.num_groups = I.num_nodes_pad / I.w_size;
  I.stride = (I.num_nodes_pad + 63) / 64 * 64;

  Params& p = pr->p;
  p.nx = I.nx; p.ny = I.ny; p.nz = I.nz;
  p.periodic = I.periodic;
  p.stride = I.stride;
  p.dx = I.dx; p.dy = I.dy; p.dz = I.dz;
  p.xmin = I.xmin; p.ymin = I.ymin; p.zmin = I.zmin;
  p.dt = I.dt;
  p.hall = I.hall;
  p.diss = I.dissipation;
  p.form_uibj = I.form_uibj < 0 ? 2 : I.form_uibj;
  p.form_source = I.form_source < 0 ? 1 : I.form_source;
  p.form_ujbi = I.form_ujbi < 0 ? 0 : I.form_ujbi;
  p.testcase = I.testcase;
  p.laplace = I.laplace < 0 ? 0 : I.laplace;
  p.bnodes = I.bnodes;
  p.ho_factor = I.ho_factor;
  p.mp2min = I.mp2min; p.mp2max = I.mp2max; p.cmin = I.cmin; p.cmax = I.cmax;
  p.ux = I.ux; p.uy = I.uy; p.uz = I.uz;
  return IEQ_OK;
}

bool range_ok(const ieq_program* pr, const uint32_t g[3], int off) {
  return g[0] + (uint32_t)off <= (uint32_t)pr->info.nx && g[1] + (uint32_t)off <= (uint32_t)pr->info.ny &&
         g[2] + (uint32_t)off <= (uint32_t)pr->info.nz;
}

// Does the NDRange g cover the program's whole grid?
bool full_grid(const ieq_program* pr, const uint32_t g[3]) {
  return g[0] == (uint32_t)pr->info.nx && g[1] == (uint32_t)pr->info.ny && g[2] == (uint32_t)pr->info.nz;
}

template <typename R>
int run_one(ieq_program* pr, const char* name, const uint32_t g[3], const uint32_t l[3], void* const* args, int nargs) {
  const Params& p = pr->p;
  const int ops = pr->info.opset;
  hipStream_t s = pr->stream;
  auto A = [&](int i) { return (R*)args[i]; };
  for (int i = 0; i < nargs; ++i)
    if (!args[i]) return pr->fail(IEQ_ERR_ARGS, std::string(name) + ": argument " + std::to_string(i) + " is NULL");

  if (const StageDesc* st = find_stage(name)) {
    if (!pr->has("kernel_time_integrator")) return pr->fail(IEQ_ERR_KERNEL, std::string("kernel not in program: ") + name);
    if (st->kind == K_STAGE && !(pr->has("induction_Eq_volume") && pr->has("induction_Eq_surface")))
      return pr->fail(IEQ_ERR_KERNEL, std::string(name) + " needs the induction_Eq_volume and induction_Eq_surface modules");
    if (nargs != st->nf + 4)
      return pr->fail(IEQ_ERR_ARGS, std::string(name) + " takes " + std::to_string(st->nf + 4) + " arguments");
    if (!range_ok(pr, g, 0)) return pr->fail(IEQ_ERR_ARGS, std::string(name) + ": NDRange exceeds the grid");
    R* F[3] = {A(0), A(1), st->nf == 3 ? A(2) : nullptr};
    R *J = A(st->nf), *u = A(st->nf + 1), *rho = A(st->nf + 2), *time = A(st->nf + 3);
    switch (st->kind) {
      case K_STAGE: {
        R* P = st->P >= 0 ? F[st->P] : nullptr;
        R* Q = st->Q >= 0 ? F[st->Q] : nullptr;
        // the whole grid of a program the tiled kernel covers: z-marching tiles; otherwise (or for
        // a part of the grid) one work-item per node
        if (full_grid(pr, g) && ieq::tiled_stage_ok(p, ops, (int)sizeof(R))) {
          hipError_t e = ieq::tiled_stage<R>(p, ops, s, F[st->out], P, (R)st->cP, Q, (R)st->cQ, F[st->Y], (R)(st->cD * p.dt), u, time);
          if (e != hipSuccess) return hip_fail(pr, e, name);
          pr->tiled_used = true;
        } else {
          ieq::gen_stage<R>(p, ops, s, g, F[st->out], P, (R)st->cP, Q, (R)st->cQ, F[st->Y], (R)(st->cD * p.dt), J, u, time);
        }
        break;
      }
      case K_2R:
        ieq::gen_update_2r<R>(p, s, g, F[st->P], F[st->out], F[st->Q], (R)st->cP, (R)st->cQ, time, st->tc);
        break;
      case K_2N:
        ieq::gen_update_2n<R>(p, s, g, F[st->out], F[st->P], (R)st->cQ, time, st->tc);
        break;
      case K_S06:
        ieq::gen_ssprk104_06<R>(p, s, g, F[0], F[1], F[2], time);
        break;
      case K_TIME:
        ieq::gen_time<R>(s, time, st->tc, p.dt, 0);
        break;
      case K_CALC_TIME:
        ieq::gen_time<R>(s, time, 0, p.dt, 1);
        break;
      case K_U:
        ieq::gen_calc_u<R>(p, s, g, u, time);
        break;
      case K_CURL:
        ieq::gen_curl_rho<R>(p, ops, s, g, F[st->Y], J, rho);
        break;
    }
    return IEQ_OK;
  }

  std::string n(name);
  auto need = [&](const char* mod, int na) -> int {
    if (!pr->has(mod)) return pr->fail(IEQ_ERR_KERNEL, std::string("kernel not in program: ") + name + " (module " + mod + ")");
    if (nargs != na) return pr->fail(IEQ_ERR_ARGS, n + " takes " + std::to_string(na) + " arguments");
    return IEQ_OK;
  };
  int rc;
  if (n == "init_b" || n == "init_u" || n == "init_rho" || n == "analytical_b") {
    int which = n == "init_b" ? 0 : n == "init_u" ? 1 : n == "init_rho" ? 2 : 3;
    if ((rc = need("kernel_init", which == 3 ? 2 : 1))) return rc;
    if (!range_ok(pr, g, 0)) return pr->fail(IEQ_ERR_ARGS, n + ": NDRange exceeds the grid");
    ieq::gen_init<R>(p, s, g, A(0), which, which == 3 ? A(1) : nullptr);
    return IEQ_OK;
  }
  if (n == "norm2" || n == "norm2_diff" || n == "norm2_S" || n == "dot_product") {
    int which = n == "norm2" ? 0 : n == "norm2_diff" ? 1 : n == "norm2_S" ? 2 : 3;
    if ((rc = need(which == 3 ? "kernel_dot_product" : "kernel_operator", (which == 1 || which == 3) ? 3 : 2))) return rc;
    long long gsz = g[0], lsz = l ? l[0] : pr->info.w_size;
    if (lsz <= 0 || gsz % lsz || gsz > pr->info.stride)
      return pr->fail(IEQ_ERR_ARGS, n + ": global size must be a multiple of the local size and fit the field");
    const R* a = A(0);
    const R* b = (which == 1 || which == 3) ? A(1) : nullptr;
    R* out = (which == 1 || which == 3) ? A(2) : A(1);
    ieq::gen_group_reduce<R>(p, ops, s, a, b, out, which, lsz, gsz / lsz, std::min<long long>(gsz, pr->info.num_nodes));
    return IEQ_OK;
  }
  if (n == "calc_div" || n == "calc_curl" || n == "calc_grad") {
    if ((rc = need("kernel_operator", 2))) return rc;
    if (!range_ok(pr, g, 0)) return pr->fail(IEQ_ERR_ARGS, n + ": NDRange exceeds the grid");
    if (n == "calc_div") ieq::gen_div<R>(p, ops, s, g, A(0), A(1), 0);
    else if (n == "calc_curl") ieq::gen_curl<R>(p, ops, s, g, A(0), A(1));
    else ieq::gen_grad<R>(p, ops, s, g, A(0), A(1));
    return IEQ_OK;
  }
  if (n == "calc_div_divcleaning" || n == "calc_laplace_divcleaning" || n == "projector_divcleaning") {
    if ((rc = need("kernel_div_cleaning", 2))) return rc;
    int off = n == "projector_divcleaning" ? 0 : p.bnodes;
    if (!range_ok(pr, g, off)) return pr->fail(IEQ_ERR_ARGS, n + ": NDRange exceeds the grid");
    if (n == "calc_div_divcleaning") ieq::gen_div<R>(p, ops, s, g, A(0), A(1), off);
    else if (n == "calc_laplace_divcleaning") ieq::gen_lap<R>(p, ops, s, g, A(0), A(1));
    else ieq::gen_proj<R>(p, ops, s, g, A(0), A(1));
    return IEQ_OK;
  }
  if (n == "calc_axpy" || n == "calc_xpay") {
    if ((rc = need("kernel_CG", 3))) return rc;
    if (!range_ok(pr, g, p.bnodes)) return pr->fail(IEQ_ERR_ARGS, n + ": NDRange exceeds the grid");
    ieq::gen_axpy<R>(p, s, g, A(0), A(1), A(2), n == "calc_xpay");
    return IEQ_OK;
  }
  return pr->fail(IEQ_ERR_KERNEL, std::string("kernel not in program: ") + name);
}

int run_list(ieq_program* pr, const char* const* names, int n, const uint32_t g[3], const uint32_t l[3], void* const* args,
             int nargs) {
  for (int i = 0; i < n; ++i) {
    if (!names[i]) return pr->fail(IEQ_ERR_ARGS, "NULL kernel name");
    int rc = pr->info.real_bytes == 8 ? run_one<double>(pr, names[i], g, l, args, nargs)
                                      : run_one<float>(pr, names[i], g, l, args, nargs);
    if (rc) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(pr, e, names[i]);
  }
  return IEQ_OK;
}

// Kernel time of one ABI call (hip events on the program's stream); destroys its events on
// every path.
struct KernelTimer {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int begin(ieq_program* pr, bool on) {
    if (!on) return IEQ_OK;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, pr->stream);
    if (e != hipSuccess) {
      release();
      return hip_fail(pr, e, "kernel timer");
    }
    return IEQ_OK;
  }
  void end(ieq_program* pr, int rc, double* seconds) {
    if (!seconds || !e0) return;
    float ms = 0;
    if (hipEventRecord(e1, pr->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      ms = 0;
    *seconds = rc ? 0.0 : ms * 1e-3;
    release();
  }
  void release() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    e0 = e1 = nullptr;
  }
  ~KernelTimer() { release(); }
};

// IEQ_NO_FUSED=1 in the environment keeps every call on the reference kernel sequence.
bool fused_allowed() {
  const char* v = std::getenv("IEQ_NO_FUSED");
  return !(v && *v && std::strcmp(v, "0") != 0);
}


// Is names[0..n) the RK_Step list of one integrator (compute_numerical_solution_setup.m:140-371, without
// calc_u) repeated k >= 1 times, with that integrator's argument count?
bool match_rk_list(const ieq_program* pr, const char* const* names, int n, int nargs, const char** integ, long long* k) {
  static const char* kIntegrators[6] = {"SSPRK33", "SSPRK104", "KennedyCarpenterLewis2R54C", "CalvoFrancoRandez2R64",
                                        "CarpenterKennedy2N54", "ToulorgeDesmet2N84F"};
  if (n <= 0 || !names[0]) return false;
  for (const char* I : kIntegrators) {
    if (std::strncmp(names[0], "calc_curlB_rho", 14) && std::strncmp(names[0], I, std::strlen(I))) continue;
    std::vector<std::string> sched;
    int nf = 0;
    if (ieq::schedule(I, pr->info.hall, 0, sched, nf) || nargs != nf + 4) continue;
    const int m = (int)sched.size();
    if (n % m) continue;
    bool ok = true;
    for (int i = 0; i < n && ok; ++i) ok = names[i] && sched[i % m] == names[i];
    if (ok) {
      *integ = I;
      *k = n / m;
      return true;
    }
  }
  return false;
}

}  // namespace

// ------------------------------------------------------------------------------------
extern "C" {

int ieq_program_build(ieq_program** prog, const char* settings, const char* const* modules, int nmod) {
  if (!prog) return IEQ_ERR_ARGS;
  *prog = nullptr;
  ieq_program* pr = new ieq_program();
  int rc = parse_build(pr, settings, modules, nmod);
  if (rc) {
    g_build_error = pr->err;
    delete pr;
    return rc;
  }
  g_build_error.clear();
  *prog = pr;
  return IEQ_OK;
}

void ieq_program_release(ieq_program* prog) {
  if (!prog) return;
  ieq::fused_release(prog->fused);
  ieq::cg_release(prog->cg);
  if (prog->diag) (void)hipFree(prog->diag);
  delete prog;
}

const char* ieq_last_error(const ieq_program* prog) { return prog ? prog->err.c_str() : g_build_error.c_str(); }

int ieq_program_get_info(const ieq_program* prog, ieq_program_info* info) {
  if (!prog || !info) return IEQ_ERR_ARGS;
  *info = prog->info;
  return IEQ_OK;
}

int ieq_set_stream(ieq_program* prog, void* hip_stream) {
  if (!prog) return IEQ_ERR_ARGS;
  prog->stream = (hipStream_t)hip_stream;
  return IEQ_OK;
}

int ieq_synchronize(ieq_program* prog) {
  if (!prog) return IEQ_ERR_ARGS;
  HIP_TRY(prog, hipStreamSynchronize(prog->stream));
  return IEQ_OK;
}

size_t ieq_vector_field_bytes(const ieq_program* prog) {
  return prog ? (size_t)(3 * prog->info.stride * prog->info.real_bytes) : 0;
}
size_t ieq_scalar_field_bytes(const ieq_program* prog) { return prog ? (size_t)(prog->info.stride * prog->info.real_bytes) : 0; }

int ieq_alloc(ieq_program* prog, size_t bytes, void** dptr) {
  if (!prog || !dptr) return IEQ_ERR_ARGS;
  *dptr = nullptr;
  HIP_TRY(prog, hipMalloc(dptr, bytes ? bytes : 1));
  HIP_TRY(prog, hipMemset(*dptr, 0, bytes ? bytes : 1));
  return IEQ_OK;
}

int ieq_free(ieq_program* prog, void* dptr) {
  if (!prog) return IEQ_ERR_ARGS;
  if (dptr) HIP_TRY(prog, hipFree(dptr));
  return IEQ_OK;
}

int ieq_upload_real4(ieq_program* prog, void* dst, const void* host_aos, size_t n) {
  if (!prog || !dst || !host_aos) return IEQ_ERR_ARGS;
  if ((long long)n > prog->info.stride) return prog->fail(IEQ_ERR_ARGS, "upload larger than the field");
  const size_t rb = prog->info.real_bytes;
  ScopedDev buf;
  HIP_TRY(prog, hipMalloc(&buf.p, 4 * n * rb + 1));
  void* tmp = buf.p;
  HIP_TRY(prog, hipMemcpyAsync(tmp, host_aos, 4 * n * rb, hipMemcpyHostToDevice, prog->stream));
  if (rb == 8) ieq::gen_aos_to_soa<double>(prog->stream, (const double*)tmp, (double*)dst, (long long)n, prog->info.stride);
  else ieq::gen_aos_to_soa<float>(prog->stream, (const float*)tmp, (float*)dst, (long long)n, prog->info.stride);
  HIP_TRY(prog, hipStreamSynchronize(prog->stream));
  return IEQ_OK;
}

int ieq_download_real4(ieq_program* prog, void* host_aos, const void* src, size_t n) {
  if (!prog || !src || !host_aos) return IEQ_ERR_ARGS;
  if ((long long)n > prog->info.stride) return prog->fail(IEQ_ERR_ARGS, "download larger than the field");
  const size_t rb = prog->info.real_bytes;
  ScopedDev buf;
  HIP_TRY(prog, hipMalloc(&buf.p, 4 * n * rb + 1));
  void* tmp = buf.p;
  if (rb == 8) ieq::gen_soa_to_aos<double>(prog->stream, (const double*)src, (double*)tmp, (long long)n, prog->info.stride);
  else ieq::gen_soa_to_aos<float>(prog->stream, (const float*)src, (float*)tmp, (long long)n, prog->info.stride);
  HIP_TRY(prog, hipMemcpyAsync(host_aos, tmp, 4 * n * rb, hipMemcpyDeviceToHost, prog->stream));
  HIP_TRY(prog, hipStreamSynchronize(prog->stream));
  return IEQ_OK;
}

int ieq_upload_real(ieq_program* prog, void* dst, const void* host, size_t n) {
  if (!prog || !dst || !host) return IEQ_ERR_ARGS;
  if ((long long)n > prog->info.stride) return prog->fail(IEQ_ERR_ARGS, "upload larger than the field");
  HIP_TRY(prog, hipMemcpyAsync(dst, host, n * prog->info.real_bytes, hipMemcpyHostToDevice, prog->stream));
  HIP_TRY(prog, hipStreamSynchronize(prog->stream));
  return IEQ_OK;
}

int ieq_download_real(ieq_program* prog, void* host, const void* src, size_t n) {
  if (!prog || !src || !host) return IEQ_ERR_ARGS;
  if ((long long)n > prog->info.stride) return prog->fail(IEQ_ERR_ARGS, "download larger than the field");
  HIP_TRY(prog, hipMemcpyAsync(host, src, n * prog->info.real_bytes, hipMemcpyDeviceToHost, prog->stream));
  HIP_TRY(prog, hipStreamSynchronize(prog->stream));
  return IEQ_OK;
}

int ieq_run(ieq_program* prog, const char* const* names, int n, const uint32_t g[3], const uint32_t l[3], void* const* args,
            int nargs, double* kernel_seconds) {
  if (!prog || (!names && n > 0) || !g || (!args && nargs > 0) || n < 0) return IEQ_ERR_ARGS;
  KernelTimer tm;
  int rc = tm.begin(prog, kernel_seconds != nullptr);
  if (rc) return rc;
  int used = 0;
  const char* integ = nullptr;
  long long k = 0;
  // The reference's time loop runs repmat(RK_Step, 1, k) in one call (compute_numerical_solution.m:91-108):
  // such a list over the whole grid takes the fused kernels when the program qualifies.
  if (fused_allowed() && full_grid(prog, g) && match_rk_list(prog, names, n, nargs, &integ, &k)) {
    ieq::FusedOpts opt;
    opt.final_curl = true;  // every argument ends as after the listed kernels
    rc = ieq::fused_step(prog->p, prog->info, prog->stream, &prog->fused, integ, k, 0, args, opt, &used, prog->err);
  }
  prog->tiled_used = false;
  if (rc == IEQ_OK && !used) rc = run_list(prog, names, n, g, l, args, nargs);
  prog->last_path = rc == IEQ_OK ? (used ? 1 : prog->tiled_used ? 2 : 0) : -1;
  tm.end(prog, rc, kernel_seconds);
  return rc;
}

int ieq_run_host(ieq_program* prog, const char* const* names, int n, const uint32_t g[3], const uint32_t l[3],
                 const ieq_host_arg* args, int nargs, double* kernel_seconds) {
  if (!prog || !args || nargs <= 0) return IEQ_ERR_ARGS;
  for (int i = 0; i < nargs; ++i) {
    if (!args[i].host) return prog->fail(IEQ_ERR_ARGS, "host argument " + std::to_string(i) + " is NULL");
    if ((long long)args[i].n > prog->info.stride)
      return prog->fail(IEQ_ERR_ARGS, "host argument " + std::to_string(i) + " is larger than a field");
  }
  const size_t rb = prog->info.real_bytes;
  std::vector<void*> dev(nargs, nullptr);
  int rc = IEQ_OK;
  for (int i = 0; i < nargs && rc == IEQ_OK; ++i) {  // copy in
    const ieq_host_arg& h = args[i];
    rc = ieq_alloc(prog, h.vector ? ieq_vector_field_bytes(prog) : std::max(ieq_scalar_field_bytes(prog), h.n * rb), &dev[i]);
    if (rc == IEQ_OK) rc = h.vector ? ieq_upload_real4(prog, dev[i], h.host, h.n) : ieq_upload_real(prog, dev[i], h.host, h.n);
  }
  if (rc == IEQ_OK) rc = ieq_run(prog, names, n, g, l, dev.data(), nargs, kernel_seconds);
  for (int i = 0; i < nargs && rc == IEQ_OK; ++i)  // write back in place
    rc = args[i].vector ? ieq_download_real4(prog, args[i].host, dev[i], args[i].n)
                        : ieq_download_real(prog, args[i].host, dev[i], args[i].n);
  const std::string err = prog->err;
  for (void* d : dev)
    if (d) (void)ieq_free(prog, d);
  prog->err = err;  // the first failure's message
  return rc;
}

static int step_impl(ieq_program* prog, const char* integrator, long long nsteps, int variable_u, void* const* args,
                     int nargs, const double* uniform, double* kernel_seconds) {
  if (!prog || !integrator || nsteps < 0 || !args) return IEQ_ERR_ARGS;
  std::vector<std::string> sched;
  int nf = 0;
  int rc = ieq::schedule(integrator, prog->info.hall, variable_u, sched, nf);
  if (rc) return prog->fail(IEQ_ERR_ARGS, std::string("Unkown time integrator ") + integrator + "!");
  if (nargs != nf + 4) return prog->fail(IEQ_ERR_ARGS, std::string(integrator) + " steps take " + std::to_string(nf + 4) + " arguments");
  KernelTimer tm;
  if ((rc = tm.begin(prog, kernel_seconds != nullptr))) return rc;
  int used_fused = 0;
  if (fused_allowed()) {
    ieq::FusedOpts opt;
    opt.uniform = uniform;
    rc = ieq::fused_step(prog->p, prog->info, prog->stream, &prog->fused, integrator, nsteps, variable_u, args, opt,
                         &used_fused, prog->err);
  }
  prog->tiled_used = false;
  if (rc == IEQ_OK && !used_fused) {
    std::vector<const char*> names;
    for (auto& s : sched) names.push_back(s.c_str());
    uint32_t g[3] = {(uint32_t)prog->info.nx, (uint32_t)prog->info.ny, (uint32_t)prog->info.nz};
    uint32_t l[3] = {(uint32_t)prog->info.w_size, 1, 1};
    for (long long k = 0; k < nsteps && rc == IEQ_OK; ++k) rc = run_list(prog, names.data(), (int)names.size(), g, l, args, nargs);
  }
  prog->last_path = rc == IEQ_OK ? (used_fused ? 1 : prog->tiled_used ? 2 : 0) : -1;
  tm.end(prog, rc, kernel_seconds);
  return rc;
}

int ieq_step(ieq_program* prog, const char* integrator, long long nsteps, int variable_u, void* const* args, int nargs,
             double* kernel_seconds) {
  return step_impl(prog, integrator, nsteps, variable_u, args, nargs, nullptr, kernel_seconds);
}

int ieq_step_uniform(ieq_program* prog, const char* integrator, long long nsteps, const double u[3], double rho,
                     void* const* args, int nargs, double* kernel_seconds) {
  if (!u) return IEQ_ERR_ARGS;
  const double uni[4] = {u[0], u[1], u[2], rho};
  for (double v : uni)
    if (!std::isfinite(v)) return prog ? prog->fail(IEQ_ERR_ARGS, "uniform u and rho must be finite") : IEQ_ERR_ARGS;
  return step_impl(prog, integrator, nsteps, 0, args, nargs, uni, kernel_seconds);
}

int ieq_last_step_path(const ieq_program* prog) { return prog ? prog->last_path : -1; }

int ieq_diagnostics(ieq_program* prog, const void* b, double time, double out[3]) {
  if (!prog || !b || !out) return IEQ_ERR_ARGS;
  if (!prog->diag) HIP_TRY(prog, hipMalloc(&prog->diag, (3 * ieq::kDiagBlocks + 3) * sizeof(double)));
  double* res = prog->diag + 3 * ieq::kDiagBlocks;
  const int analytic = prog->info.testcase != 7;  // uniform_flow has no analytic field
  hipError_t e = prog->info.real_bytes == 8
                     ? ieq::diag_launch<double>(prog->p, prog->info.opset, prog->stream, (const double*)b, time, analytic,
                                                prog->diag, res)
                     : ieq::diag_launch<float>(prog->p, prog->info.opset, prog->stream, (const float*)b, time, analytic,
                                               prog->diag, res);
  if (e != hipSuccess) return hip_fail(prog, e, "diagnostics launch");
  HIP_TRY(prog, hipMemcpyAsync(out, res, 3 * sizeof(double), hipMemcpyDeviceToHost, prog->stream));
  HIP_TRY(prog, hipStreamSynchronize(prog->stream));
  return IEQ_OK;
}

int ieq_diagnostics_noalloc(ieq_program* prog, const void* b, double time, double* out_dev) {
  if (!prog || !b || !out_dev) return IEQ_ERR_ARGS;
  if (ieq::diag_blocks(prog->p) > ieq::kDiagBlocks)  // the partial pass never uses more blocks than the scratch holds
    return prog->fail(IEQ_ERR_ARGS, "diagnostics need more partial sums than the static scratch holds");
  double* part = nullptr;
  hipError_t e = ieq::diag_scratch(&part);
  if (e != hipSuccess) return hip_fail(prog, e, "diagnostics scratch lookup");
  const int analytic = prog->info.testcase != 7;
  e = prog->info.real_bytes == 8
          ? ieq::diag_launch<double>(prog->p, prog->info.opset, prog->stream, (const double*)b, time, analytic, part, out_dev)
          : ieq::diag_launch<float>(prog->p, prog->info.opset, prog->stream, (const float*)b, time, analytic, part, out_dev);
  if (e != hipSuccess) return hip_fail(prog, e, "diagnostics launch");
  return IEQ_OK;
}

}  // extern "C"

namespace {

// Host sum of the group partials of a reduction over the whole padded field.
template <typename R>
int reduce_host(ieq_program* pr, int which, const R* a, const R* b, R* partials, double* out) {
  const ieq_program_info& I = pr->info;
  ieq::gen_group_reduce<R>(pr->p, I.opset, pr->stream, a, b, partials, which, I.w_size, I.num_groups, I.num_nodes);
  std::vector<R> h((size_t)I.num_groups);
  HIP_TRY(pr, hipMemcpyAsync(h.data(), partials, h.size() * sizeof(R), hipMemcpyDeviceToHost, pr->stream));
  HIP_TRY(pr, hipStreamSynchronize(pr->stream));
  double acc = 0.0;
  for (R v : h) acc += (double)v;
  *out = acc;
  return IEQ_OK;
}

template <typename R>
int clean_div_t(ieq_program* pr, R* b, R* phi, double thr, double rel, int max_it, ieq_cg_result* res) {
  const Params& p = pr->p;
  const ieq_program_info& I = pr->info;
  int nbl = 0, wl = 0;
  const double* tab = ieq::host_lap_table(I.opset, I.laplace, &nbl, &wl);
  if (!tab) return pr->fail(IEQ_ERR_SETTINGS, "no -Laplacian table for this operator set");
  std::memset(res, 0, sizeof(*res));
  const size_t sbytes = (size_t)I.stride * sizeof(R);
  ScopedDev buf;  // scalar work fields + reduction partials
  HIP_TRY(pr, hipMalloc(&buf.p, 4 * sbytes + (size_t)I.num_groups * sizeof(R) + 64));
  R* tmp = (R*)buf.p;
  R *r = tmp, *u = tmp + I.stride, *q = tmp + 2 * I.stride, *dv = tmp + 3 * I.stride, *part = tmp + 4 * I.stride;
  int rc = ieq::cg_run_t<R>(p, I, tab, nbl, wl, pr->stream, &pr->cg, b, phi, thr, rel, max_it, 16, res, pr->err);
  if (rc < 0) return IEQ_ERR_HIP;
  const uint32_t g_dc[3] = {(uint32_t)(p.nx - 2 * p.bnodes), (uint32_t)(p.ny - 2 * p.bnodes), (uint32_t)(p.nz - 2 * p.bnodes)};
  if (rc == 1) {  // reference kernel sequence with host-side scalars (clean_div_CG.m:31-76)
    HIP_TRY(pr, hipMemsetAsync(tmp, 0, 3 * sbytes, pr->stream));
    HIP_TRY(pr, hipMemsetAsync(phi, 0, sbytes, pr->stream));
    R* coef = part + I.num_groups;  // 1-element coefficient buffer
    ScopedEvents ev;
    HIP_TRY(pr, hipEventCreate(&ev.e[0]));
    HIP_TRY(pr, hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    (void)hipEventRecord(e0, pr->stream);
    ieq::gen_div<R>(p, I.opset, pr->stream, g_dc, b, r, p.bnodes);
    double rr = 0, old = 1.0, uq = 0;
    if ((rc = reduce_host<R>(pr, 3, r, r, part, &rr))) return rc;
    const double rr0 = rr;
    res->divergence_norm_initial = std::sqrt(rr);
    int it = 0;
    auto go = [&]() { return rr > thr * thr && (rel <= 0 || rr > rel * rel * rr0) && it < max_it; };
    while (go()) {
      R c = (R)(rr / old);
      HIP_TRY(pr, hipMemcpyAsync(coef, &c, sizeof(R), hipMemcpyHostToDevice, pr->stream));
      ieq::gen_axpy<R>(p, pr->stream, g_dc, u, r, coef, 0);
      ieq::gen_lap<R>(p, I.opset, pr->stream, g_dc, u, q);
      if ((rc = reduce_host<R>(pr, 3, u, q, part, &uq))) return rc;
      const double alpha = rr / uq;
      c = (R)alpha;
      HIP_TRY(pr, hipMemcpyAsync(coef, &c, sizeof(R), hipMemcpyHostToDevice, pr->stream));
      ieq::gen_axpy<R>(p, pr->stream, g_dc, phi, u, coef, 1);
      c = (R)(-alpha);
      HIP_TRY(pr, hipMemcpyAsync(coef, &c, sizeof(R), hipMemcpyHostToDevice, pr->stream));
      ieq::gen_axpy<R>(p, pr->stream, g_dc, r, q, coef, 1);
      old = rr;
      if ((rc = reduce_host<R>(pr, 3, r, r, part, &rr))) return rc;
      ++it;
    }
    (void)hipEventRecord(e1, pr->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    res->iterations = it;
    res->residual_norm = std::sqrt(rr);
    res->seconds = ms * 1e-3;
  }
  // projection over the whole grid and the final divergence norm (clean_div_CG.m:78-86)
  const uint32_t g_full[3] = {(uint32_t)p.nx, (uint32_t)p.ny, (uint32_t)p.nz};
  ieq::gen_proj<R>(p, I.opset, pr->stream, g_full, b, phi);
  HIP_TRY(pr, hipMemsetAsync(dv, 0, sbytes, pr->stream));
  ieq::gen_div<R>(p, I.opset, pr->stream, g_full, b, dv, 0);
  double dn = 0;
  if ((rc = reduce_host<R>(pr, 2, dv, nullptr, part, &dn))) return rc;
  res->divergence_norm_final = std::sqrt(std::max(dn, 0.0));
  HIP_TRY(pr, hipStreamSynchronize(pr->stream));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(pr, e, "clean_div_cg");
  return IEQ_OK;
}

}  // namespace

extern "C" {

int ieq_clean_div_cg(ieq_program* prog, void* b, void* phi, double abs_threshold, double rel_tolerance, int max_iterations,
                     ieq_cg_result* result) {
  if (!prog || !b || !phi || !result || max_iterations < 0) return IEQ_ERR_ARGS;
  if (!prog->has("kernel_div_cleaning") || prog->info.laplace < 0)
    return prog->fail(IEQ_ERR_KERNEL, "divergence cleaning needs the kernel_div_cleaning module and a USE_LAPLACE_* define");
  if (prog->info.real_bytes == 8)
    return clean_div_t<double>(prog, (double*)b, (double*)phi, abs_threshold, rel_tolerance, max_iterations, result);
  return clean_div_t<float>(prog, (float*)b, (float*)phi, abs_threshold, rel_tolerance, max_iterations, result);
}

struct ieq_ensemble {
  ieq_program* prog;
  ieq::Ensemble* e;
};

long long ieq_member_stride(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  return ((long long)nx * ny * nz + 63) / 64 * 64;
}

int ieq_ensemble_create(ieq_ensemble** ens, ieq_program* prog, const ieq_member* members, int nmembers, const char* integrator) {
  if (!ens || !prog || !members || nmembers < 1 || !integrator) return IEQ_ERR_ARGS;
  *ens = nullptr;
  ieq::Ensemble* e = nullptr;
  int rc = ieq::ensemble_create(prog->p, prog->info, members, nmembers, integrator, &e, prog->err);
  if (rc) return rc;
  *ens = new ieq_ensemble{prog, e};
  return IEQ_OK;
}

int ieq_ensemble_step(ieq_ensemble* ens, long long nsteps, double* kernel_seconds) {
  if (!ens || nsteps < 0) return IEQ_ERR_ARGS;
  // order after the work already enqueued on the program's stream
  HIP_TRY(ens->prog, hipStreamSynchronize(ens->prog->stream));
  return ieq::ensemble_step(ens->e, nsteps, kernel_seconds, ens->prog->err);
}

int ieq_ensemble_config(const ieq_ensemble* ens, int* tile_shape, int* z_chunk, int* order) {
  if (!ens || !tile_shape || !z_chunk || !order) return IEQ_ERR_ARGS;
  ieq::ensemble_config(ens->e, tile_shape, z_chunk, order);
  return IEQ_OK;
}

void ieq_ensemble_release(ieq_ensemble* ens) {
  if (!ens) return;
  ieq::ensemble_release(ens->e);
  delete ens;
}

struct ieq_adjoint {
  ieq_program* prog;
  ieq::Adjoint* a;
};

int ieq_adjoint_create(ieq_adjoint** adj, ieq_program* prog, long long nsteps, int snapshots) {
  if (!adj || !prog) return IEQ_ERR_ARGS;
  *adj = nullptr;
  ieq::Adjoint* a = nullptr;
  int rc = ieq::adjoint_create(prog->p, prog->info, nsteps, snapshots, &a, prog->err);
  if (rc) return rc;
  *adj = new ieq_adjoint{prog, a};
  return IEQ_OK;
}

int ieq_adjoint_forward(ieq_adjoint* adj, const void* b0, void* bT, double c_hall, const double u[3], double* kernel_seconds) {
  if (!adj || !b0 || !bT) return IEQ_ERR_ARGS;
  return ieq::adjoint_forward(adj->a, adj->prog->stream, b0, bT, c_hall, u, kernel_seconds, adj->prog->err);
}

int ieq_adjoint_backward(ieq_adjoint* adj, const void* lambda_T, void* grad_b0, double grads[4], double* kernel_seconds) {
  if (!adj || !lambda_T || !grad_b0 || !grads) return IEQ_ERR_ARGS;
  return ieq::adjoint_backward(adj->a, adj->prog->stream, lambda_T, grad_b0, grads, kernel_seconds, adj->prog->err);
}

int ieq_adjoint_counts(const ieq_adjoint* adj, long long* forward_steps, long long* adjoint_steps) {
  if (!adj || !forward_steps || !adjoint_steps) return IEQ_ERR_ARGS;
  ieq::adjoint_counts(adj->a, forward_steps, adjoint_steps);
  return IEQ_OK;
}

void ieq_adjoint_release(ieq_adjoint* adj) {
  if (!adj) return;
  ieq::adjoint_release(adj->a);
  delete adj;
}

}  // extern "C"
