// Run-time composition of a user pair term into the force kernels: the
// PairForceFragment / composer path of the reference
// (src/forces/jit_composed.rs:112-151 fragment, :697-720 NVRTC compile,
// :1162-1643 composed kernel), with hiprtc.  The fragment is a functor
// struct, inserted ahead of the library's own pair-kernel source (embedded as
// text at build time); the pair kernel calls it for every listed pair, with
// the built-in LJ / Coulomb terms, the same cutoff and exclusion masking, and
// the same fixed-point accumulation.  The result is a second code object
// whose six pair-kernel variants replace the built-in ones for that force
// field.
#include <hip/hiprtc.h>

#include <cctype>
#include <string>
#include <vector>

#include "forcefield.h"

extern "C" const char hmd_jit_src_params[];
extern "C" const char hmd_jit_src_device[];
extern "C" const char hmd_jit_src_pair[];

namespace hmd {

static void check_rtc(hiprtcResult r, const char* what) {
  if (r != HIPRTC_SUCCESS)
    throw Error(HMD_ERR_GPU, std::string(what) + ": " + hiprtcGetErrorString(r));
}

static bool is_identifier(const std::string& s) {
  if (s.empty() || !(std::isalpha((unsigned char)s[0]) || s[0] == '_')) return false;
  for (char ch : s)
    if (!(std::isalnum((unsigned char)ch) || ch == '_')) return false;
  return true;
}

std::vector<char> compile_pair_fragment(const std::string& name, const std::string& source,
                                        std::string* log_out) {
  if (!is_identifier(name)) throw Error(HMD_ERR_INVALID, "fragment name must be an identifier");
  const std::string src = "#define HMD_USER_FUNCTOR " + name + "\n" +
                          "#include \"hmd_device.h\"\n" + source + "\n" + hmd_jit_src_pair;
  const char* headers[] = {hmd_jit_src_params, hmd_jit_src_device};
  const char* names[] = {"hmd_params.h", "hmd_device.h"};
  hiprtcProgram prog;
  check_rtc(hiprtcCreateProgram(&prog, src.c_str(), "hmd_pair_composed.hip", 2, headers, names),
            "hiprtcCreateProgram");
  // the flags of the built-in code object: no packed FP32, IEEE f32 divide
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-fhip-fp32-correctly-rounded-divide-sqrt"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  size_t log_size = 0;
  (void)hiprtcGetProgramLogSize(prog, &log_size);
  std::string log(log_size, '\0');
  if (log_size) (void)hiprtcGetProgramLog(prog, &log[0]);
  if (log_out) *log_out = log;
  if (rc != HIPRTC_SUCCESS) {
    (void)hiprtcDestroyProgram(&prog);
    throw Error(HMD_ERR_INVALID, "pair fragment does not compile:\n" + log.substr(0, 4000));
  }
  size_t code_size = 0;
  check_rtc(hiprtcGetCodeSize(prog, &code_size), "hiprtcGetCodeSize");
  std::vector<char> code(code_size);
  check_rtc(hiprtcGetCode(prog, code.data()), "hiprtcGetCode");
  (void)hiprtcDestroyProgram(&prog);
  return code;
}

JitPairModule::~JitPairModule() {
  if (module) (void)hipModuleUnload(module);
}

void ForceField::set_pair_fragment(const std::string& name, const std::string& source,
                                   const float* params, uint32_t n_params) {
  std::vector<char> code = compile_pair_fragment(name, source, nullptr);
  auto m = std::make_unique<JitPairModule>();
  check(hipModuleLoadData(&m->module, code.data()), "hipModuleLoadData (composed pair kernels)");
  const KernelId ks[7] = {K_hmd_pair_force_f, K_hmd_pair_force_fev, K_hmd_pair_force_lj_f,
                          K_hmd_pair_force_lj_fev, K_hmd_pair_force_coul_f,
                          K_hmd_pair_force_coul_fev, K_hmd_pair_scaled};
  for (int k = 0; k < 7; ++k) {
    m->kernel[k] = ks[k];
    check(hipModuleGetFunction(&m->fn[k], m->module, kKernelNames[ks[k]]),
          kKernelNames[ks[k]]);
  }
  user_params.alloc(std::max<uint32_t>(1u, n_params));
  if (n_params) {
    user_params.upload(params, n_params, c.stream);
    check(hipStreamSynchronize(c.stream), "sync");
  }
  jit = std::move(m);
  graph_invalid = true;
}

hipFunction_t ForceField::pair_function(KernelId k) const {
  if (jit)
    for (int i = 0; i < 7; ++i)
      if (jit->kernel[i] == k) return jit->fn[i];
  return c.fn[k];
}

}  // namespace hmd

using namespace hmd;

extern "C" {

hmd_status hmd_forcefield_set_pair_fragment(hmd_forcefield* f, const char* functor_name,
                                            const char* functor_source, const float* params,
                                            uint32_t n_params) {
  HMD_GUARD({
    if (!f || !functor_name || !functor_source || (n_params && !params))
      throw Error(HMD_ERR_INVALID, "null argument");
    f->ff->set_pair_fragment(functor_name, functor_source, params, n_params);
  });
}

hmd_status hmd_pair_fragment_compile(const char* functor_name, const char* functor_source,
                                     char* log, int32_t log_len) {
  std::string text;
  hmd_status st = HMD_OK;
  try {
    if (!functor_name || !functor_source) throw Error(HMD_ERR_INVALID, "null argument");
    (void)compile_pair_fragment(functor_name, functor_source, &text);
  } catch (const Error& e) {
    hmd_set_error(e.what());
    st = e.code;
    if (text.empty()) text = e.what();
  }
  if (log && log_len > 0) std::snprintf(log, (size_t)log_len, "%s", text.c_str());
  return st;
}

}  // extern "C"
