// gsx_runtime.hip -- fields registered at run time (gsx_field_register).
//
// A registered field is HIP C++ source for a `struct Field` with the same members as the
// built-in fields of gsx_core.h (D, P, autonomous, f and vjp templated on the scalar type;
// jac and jac_vjp when stiff). The library wraps the source in namespace gsx_user, appends it
// to the kernel templates of gsx_tsit5_kernels.h / gsx_rb23_kernels.h (embedded in the library
// at build time, so the library needs no header files at run time), and compiles it for gfx950
// with hiprtc, one kernel at a time on its first launch:
//
//   family         namespace    kernels
//   kTsit5         gsx          tsit5 forward (recording and plain), replay, adjoint
//                               (pair kernel for C = 2, segment kernels for C <= 1/4/8/16/32)
//   kTsit5Strict   gsx_strict   the same with contraction off and the library's pow/division
//   kTsit5F32      gsx          the FP32 forward, replay and adjoint (autonomous fields)
//   kRb23          gsx          Rosenbrock23 forward, replay, adjoint (stiff fields; pair kernel
//                               for C = 2)
//   kRb23Strict    gsx_strict   the same, strict arithmetic
//   kRb23F32       gsx          the FP32 Rosenbrock23 forward, replay and adjoint (autonomous)
//   kTsit5Wide     gsx          the wide Tsit5 forward, replay and adjoint (gsx_wide.h): the
//                               only family of a field with D above 64
//
// The recording forward of the primary family (kTsit5, or kRb23 for a stiff field) is
// compiled at registration so that errors in the source surface there; every other kernel on
// its first launch. Code objects are loaded into each device's context on first launch there. Launches pass the same argument blocks as the
// built-in kernels (FwdArgs, ReplayArgs, AdjArgs by value).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gsx_internal.h"
#include "gsx_rt_headers.inc"  // kRtHeaderNames, kRtHeaderSources, kRtHeaderCount (generated)

namespace gsx {
namespace rt {

namespace {

struct FamilySpec {
  const char* ns;                 // namespace of the kernel templates
  const char* include;            // kernel header
  std::vector<const char*> opts;  // extra compile options
  std::vector<std::pair<const char*, const char*>> kernels;  // key -> template (with %s = F)
};

const FamilySpec& spec(Family f) {
  static const std::vector<const char*> kStrict = {"-DGSX_NS=gsx_strict", "-DGSX_STRICT_ARITH",
                                                   "-ffp-contract=off"};
  static const FamilySpec specs[kNumFamilies] = {
      {"gsx", "gsx_tsit5_kernels.h", {},
       {{"fwd_rec", "tsit5_fwd_kernel<%s, true>"}, {"fwd", "tsit5_fwd_kernel<%s, false>"},
        {"replay", "tsit5_replay_kernel<%s>"}, {"adj1", "tsit5_adj_kernel<%s, 1>"},
        {"adj_pair", "tsit5_adj_pair_kernel<%s, true>"},
        {"adj_pair_plain", "tsit5_adj_pair_kernel<%s, false>"}, {"adj4", "tsit5_adj_kernel<%s, 4>"},
        {"adj8", "tsit5_adj_kernel<%s, 8>"}, {"adj16", "tsit5_adj_kernel<%s, 16>"},
        {"adj32", "tsit5_adj_kernel<%s, 32>"}, {"saveat", "tsit5_saveat_kernel<%s>"},
        {"saveat_vjp", "tsit5_saveat_vjp_kernel<%s>"}}},
      {"gsx_strict", "gsx_tsit5_kernels.h", kStrict,
       {{"fwd_rec", "tsit5_fwd_kernel<%s, true>"}, {"fwd", "tsit5_fwd_kernel<%s, false>"},
        {"replay", "tsit5_replay_kernel<%s>"}, {"adj1", "tsit5_adj_kernel<%s, 1>"},
        {"adj_pair", "tsit5_adj_pair_kernel<%s, true>"},
        {"adj_pair_plain", "tsit5_adj_pair_kernel<%s, false>"}, {"adj4", "tsit5_adj_kernel<%s, 4>"},
        {"adj8", "tsit5_adj_kernel<%s, 8>"}, {"adj16", "tsit5_adj_kernel<%s, 16>"},
        {"adj32", "tsit5_adj_kernel<%s, 32>"}, {"saveat", "tsit5_saveat_kernel<%s>"},
        {"saveat_vjp", "tsit5_saveat_vjp_kernel<%s>"}}},
      {"gsx", "gsx_tsit5_kernels.h", {},
       {{"fwd", "tsit5_fwd_f32_kernel<%s>"}, {"replay", "tsit5_replay_f32_kernel<%s>"},
        {"adj1", "tsit5_adj_f32_kernel<%s, 1>"}, {"adj4", "tsit5_adj_f32_kernel<%s, 4>"},
        {"adj8", "tsit5_adj_f32_kernel<%s, 8>"}, {"adj16", "tsit5_adj_f32_kernel<%s, 16>"},
        {"adj32", "tsit5_adj_f32_kernel<%s, 32>"}}},
      {"gsx", "gsx_rb23_kernels.h", {},
       {{"fwd", "rb23_fwd_kernel<%s>"}, {"replay", "rb23_replay_kernel<%s>"},
        {"adj1", "rb23_adj_kernel<%s, 1>"}, {"adj_pair", "rb23_adj_pair_kernel<%s, true>"},
        {"adj_pair_plain", "rb23_adj_pair_kernel<%s, false>"}, {"adj4", "rb23_adj_kernel<%s, 4>"},
        {"adj8", "rb23_adj_kernel<%s, 8>"}, {"adj16", "rb23_adj_kernel<%s, 16>"},
        {"adj32", "rb23_adj_kernel<%s, 32>"}, {"saveat", "rb23_saveat_kernel<%s>"},
        {"saveat_vjp", "rb23_saveat_vjp_kernel<%s>"}}},
      {"gsx_strict", "gsx_rb23_kernels.h", kStrict,
       {{"fwd", "rb23_fwd_kernel<%s>"}, {"replay", "rb23_replay_kernel<%s>"},
        {"adj1", "rb23_adj_kernel<%s, 1>"}, {"adj_pair", "rb23_adj_pair_kernel<%s, true>"},
        {"adj_pair_plain", "rb23_adj_pair_kernel<%s, false>"}, {"adj4", "rb23_adj_kernel<%s, 4>"},
        {"adj8", "rb23_adj_kernel<%s, 8>"}, {"adj16", "rb23_adj_kernel<%s, 16>"},
        {"adj32", "rb23_adj_kernel<%s, 32>"}, {"saveat", "rb23_saveat_kernel<%s>"},
        {"saveat_vjp", "rb23_saveat_vjp_kernel<%s>"}}},
      {"gsx", "gsx_rodas5p_kernels.h", {},
       {{"fwd", "r5_fwd_kernel<%s>"}, {"replay", "r5_replay_kernel<%s>"},
        {"adj1", "r5_adj_kernel<%s, 1>"}, {"adj4", "r5_adj_kernel<%s, 4>"},
        {"adj8", "r5_adj_kernel<%s, 8>"}, {"adj16", "r5_adj_kernel<%s, 16>"},
        {"adj32", "r5_adj_kernel<%s, 32>"}, {"saveat", "r5_saveat_kernel<%s>"},
        {"saveat_vjp", "r5_saveat_vjp_kernel<%s>"},
        {"saveat_dense", "r5_saveat_dense_kernel<%s>"}}},
      {"gsx_strict", "gsx_rodas5p_kernels.h", kStrict,
       {{"fwd", "r5_fwd_kernel<%s>"}, {"replay", "r5_replay_kernel<%s>"},
        {"adj1", "r5_adj_kernel<%s, 1>"}, {"adj4", "r5_adj_kernel<%s, 4>"},
        {"adj8", "r5_adj_kernel<%s, 8>"}, {"adj16", "r5_adj_kernel<%s, 16>"},
        {"adj32", "r5_adj_kernel<%s, 32>"}, {"saveat", "r5_saveat_kernel<%s>"},
        {"saveat_vjp", "r5_saveat_vjp_kernel<%s>"},
        {"saveat_dense", "r5_saveat_dense_kernel<%s>"}}},
      {"gsx", "gsx_rb23_kernels.h", {},
       {{"fwd", "rb23_fwd_f32_kernel<%s>"}, {"replay", "rb23_replay_f32_kernel<%s>"},
        {"adj1", "rb23_adj_f32_kernel<%s, 1>"}, {"adj4", "rb23_adj_f32_kernel<%s, 4>"},
        {"adj8", "rb23_adj_f32_kernel<%s, 8>"}, {"adj16", "rb23_adj_f32_kernel<%s, 16>"},
        {"adj32", "rb23_adj_f32_kernel<%s, 32>"}}},
      {"gsx", "gsx_wide.h", {},
       {{"fwd", "tsit5w_fwd_kernel<%s>"}, {"replay", "tsit5w_replay_kernel<%s>"},
        {"adj", "tsit5w_adj_kernel<%s>"}}},
  };
  return specs[f];
}

const char* family_name(Family f) {
  static const char* names[kNumFamilies] = {"tsit5", "tsit5 (strict)", "tsit5 (FP32)",
                                            "rosenbrock23", "rosenbrock23 (strict)",
                                            "rodas5p", "rodas5p (strict)",
                                            "rosenbrock23 (FP32)", "tsit5 (wide)"};
  return names[f];
}

// One hiprtc program per kernel, compiled on its first launch (a wide field's kernels take
// seconds each to compile, and a run needs only a few of them).
struct Module {
  bool tried = false;
  int rc = GSX_OK;
  std::string error;
  std::string code;                  // gfx950 code object
  std::string lowered;               // mangled kernel name
  std::map<int, hipModule_t> loaded;  // device -> module
  std::map<int, hipFunction_t> fn;    // device -> kernel
};

struct RtField {
  int D = 0, P = 0;
  bool stiff = false, autonomous = true, released = false;
  std::string source;
  std::map<std::pair<int, std::string>, Module> kernels;  // (family, key) -> module
};

std::mutex g_mu;
std::vector<std::unique_ptr<RtField>> g_fields;

RtField* lookup(int field) {
  const int idx = field - GSX_FIELD_RUNTIME_BASE;
  if (idx < 0 || idx >= static_cast<int>(g_fields.size())) return nullptr;
  RtField* f = g_fields[idx].get();
  return f->released ? nullptr : f;
}

std::string kernel_expr(const FamilySpec& sp, const char* tmpl) {
  char buf[256];
  std::snprintf(buf, sizeof buf, tmpl, "gsx_user::Field");
  return std::string(sp.ns) + "::" + buf;
}

const char* kernel_template(const FamilySpec& sp, const char* key) {
  for (const auto& k : sp.kernels)
    if (std::strcmp(k.first, key) == 0) return k.second;
  return nullptr;
}

// Segment adjoints keep CMAX checkpoint rows per lane in LDS: CMAX * rows * 64 lanes * 8 B,
// at most 64 KiB per workgroup. A checkpoint interval that would need more is rejected for a
// wide field (the pair adjoint needs no LDS).
constexpr int kSegLdsBytes = 65536;

int seg_cmax(const char* key) {
  if (std::strncmp(key, "adj", 3) != 0 || std::strncmp(key, "adj_pair", 8) == 0) return 0;
  return std::atoi(key + 3);
}

int ckpt_rows_of(const RtField& f, Family fam) {
  return f.D + ((f.autonomous || fam == kTsit5F32 || fam == kRb23F32) ? 0 : 1);
}

std::string program_source(const RtField& f, const FamilySpec& sp) {
  std::string s = "#include \"";
  s += sp.include;
  s += "\"\nnamespace gsx_user {\nusing namespace ::GSX_NS;\n";
  s += f.source;
  s += "\n}  // namespace gsx_user\n";
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "static_assert(gsx_user::Field::D == %d, \"Field::D differs from the registered "
                "D=%d\");\nstatic_assert(gsx_user::Field::P == %d, \"Field::P differs from the "
                "registered P=%d\");\nstatic_assert(gsx_user::Field::autonomous == %s, "
                "\"Field::autonomous differs from the registration flags\");\n",
                f.D, f.D, f.P, f.P, f.autonomous ? "true" : "false");
  s += buf;
  return s;
}

// Compiles kernel `key` of family `fam` for field `f` (caller holds g_mu). Returns GSX_OK or
// an error code with the message in the module's error.
int compile(RtField& f, Family fam, const char* key, Module** out) {
  Module& m = f.kernels[{static_cast<int>(fam), key}];
  *out = &m;
  if (m.tried) return m.rc;
  m.tried = true;
  if (f.D > kMaxRegisterD && fam != kTsit5Wide) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "D=%d is above the %d components a lane's registers hold: "
                  "a field this wide runs on the wide Tsit5 lane only (gsx_tsit5_wide_*)", f.D,
                  kMaxRegisterD);
    m.error = buf;
    return m.rc = GSX_E_UNSUPPORTED;
  }
  const FamilySpec& sp = spec(fam);
  const char* tmpl = kernel_template(sp, key);
  if (!tmpl) {
    m.error = std::string("no ") + family_name(fam) + " kernel '" + key + "'";
    return m.rc = GSX_E_INVALID;
  }
  const int cmax = seg_cmax(key);
  if (cmax > 0 && cmax * ckpt_rows_of(f, fam) * 64 * 8 > kSegLdsBytes) {
    char buf[320];
    std::snprintf(buf, sizeof buf,
                  "a checkpoint interval of up to %d needs %d B of LDS per wavefront for the "
                  "field's %d checkpoint rows (limit %d B); use a smaller ckpt_interval", cmax,
                  cmax * ckpt_rows_of(f, fam) * 64 * 8, ckpt_rows_of(f, fam), kSegLdsBytes);
    m.error = buf;
    return m.rc = GSX_E_INVALID;
  }
  const std::string src = program_source(f, sp);
  const std::string expr = kernel_expr(sp, tmpl);
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "gsx_user_field.hip",
                                       kRtHeaderCount, kRtHeaderSources, kRtHeaderNames);
  if (r != HIPRTC_SUCCESS) {
    m.error = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
    return m.rc = GSX_E_HIP;
  }
  hiprtcAddNameExpression(prog, expr.c_str());
  std::vector<const char*> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                                   "-target-feature", "-Xclang", "-packed-fp32-ops"};
  for (const char* o : sp.opts) opts.push_back(o);
  r = hiprtcCompileProgram(prog, static_cast<int>(opts.size()), opts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    m.error = std::string("compiling ") + expr + " for the field: " + log;
    hiprtcDestroyProgram(&prog);
    return m.rc = GSX_E_INVALID;
  }
  const char* low = nullptr;
  if (hiprtcGetLoweredName(prog, expr.c_str(), &low) != HIPRTC_SUCCESS || !low) {
    m.error = "hiprtcGetLoweredName failed for " + expr;
    hiprtcDestroyProgram(&prog);
    return m.rc = GSX_E_HIP;
  }
  m.lowered = low;
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  m.code.assign(cs, '\0');
  if (cs) hiprtcGetCode(prog, &m.code[0]);
  hiprtcDestroyProgram(&prog);
  return m.rc = GSX_OK;
}

int fail(const Module& m, int field) {
  return set_error(m.rc, "field %d: %s", field, m.error.c_str());
}

// The kernel `key` of family `fam` for the current device, compiling and loading as needed.
int function(int field, Family fam, const char* key, hipFunction_t* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  RtField* f = lookup(field);
  if (!f) return set_error(GSX_E_INVALID, "field %d is not a registered field", field);
  Module* m = nullptr;
  if (compile(*f, fam, key, &m) != GSX_OK) return fail(*m, field);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(GSX_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  auto it = m->fn.find(dev);
  if (it != m->fn.end()) {
    *out = it->second;
    return GSX_OK;
  }
  hipModule_t mod;
  e = hipModuleLoadData(&mod, m->code.data());
  if (e != hipSuccess)
    return set_error(GSX_E_HIP, "hipModuleLoadData (field %d, %s kernel '%s'): %s", field,
                     family_name(fam), key, hipGetErrorString(e));
  m->loaded[dev] = mod;
  hipFunction_t fn;
  e = hipModuleGetFunction(&fn, mod, m->lowered.c_str());
  if (e != hipSuccess)
    return set_error(GSX_E_HIP, "hipModuleGetFunction %s: %s", m->lowered.c_str(),
                     hipGetErrorString(e));
  m->fn[dev] = fn;
  *out = fn;
  return GSX_OK;
}

template <class Args>
int launch(int field, Family fam, const char* key, int64_t n, int block, const Args& a,
           hipStream_t s) {
  if (n == 0) return GSX_OK;
  hipFunction_t fn;
  const int rc = function(field, fam, key, &fn);
  if (rc != GSX_OK) return rc;
  const unsigned grid = static_cast<unsigned>((n + block - 1) / block);
  Args copy = a;
  void* params[] = {&copy};
  const hipError_t e =
      hipModuleLaunchKernel(fn, grid, 1, 1, static_cast<unsigned>(block), 1, 1, 0, s, params,
                            nullptr);
  if (e != hipSuccess)
    return set_error(GSX_E_HIP, "launching %s kernel '%s' of field %d: %s", family_name(fam),
                     key, field, hipGetErrorString(e));
  return GSX_OK;
}

const char* adj_key(int C, bool pair_ok) {
  if (C <= 1) return "adj1";
  if (C == 2 && pair_ok) return "adj_pair";
  if (C <= 4) return "adj4";
  if (C <= 8) return "adj8";
  if (C <= 16) return "adj16";
  if (C <= 32) return "adj32";
  return nullptr;
}

}  // namespace

bool is_runtime(int field) { return field >= GSX_FIELD_RUNTIME_BASE; }

bool info(int field, FieldInfo* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  RtField* f = lookup(field);
  if (!f) return false;
  *out = {f->D, f->P, f->autonomous, f->stiff};
  return true;
}

int tsit5_fwd(int field, Family fam, const FwdArgs& a, hipStream_t s) {
  const char* key = (fam != kTsit5F32 && a.tape && a.ckpt) ? "fwd_rec" : "fwd";
  return launch(field, fam, key, a.n, kFwdBlock, a, s);
}

int tsit5_replay(int field, Family fam, const ReplayArgs& a, hipStream_t s) {
  return launch(field, fam, "replay", a.n, kFwdBlock, a, s);
}

int tsit5_adj(int field, Family fam, const AdjArgs& a, hipStream_t s) {
  const bool pair = fam != kTsit5F32;
  const char* key = adj_key(a.ckpt_interval, pair);
  if (!key)
    return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d",
                     a.ckpt_interval, kMaxCkptInterval);
  const bool is_pair = std::strcmp(key, "adj_pair") == 0;
  if (is_pair && a.nsave == 0 && !a.pbar_in) key = "adj_pair_plain";
  return launch(field, fam, key, a.n, is_pair ? kFwdBlock : kAdjBlock, a, s);
}

int tsit5_saveat(int field, Family fam, const SaveArgs& a, hipStream_t s) {
  return launch(field, fam, "saveat", a.n, kFwdBlock, a, s);
}

int tsit5_saveat_vjp(int field, Family fam, const SaveVjpArgs& a, hipStream_t s) {
  return launch(field, fam, "saveat_vjp", a.n, kFwdBlock, a, s);
}

int rb23_fwd(int field, Family fam, const FwdArgs& a, hipStream_t s) {
  return launch(field, fam, "fwd", a.n, kFwdBlock, a, s);
}

int rb23_replay(int field, Family fam, const ReplayArgs& a, hipStream_t s) {
  return launch(field, fam, "replay", a.n, kFwdBlock, a, s);
}

int rb23_adj(int field, Family fam, const AdjArgs& a, hipStream_t s) {
  const char* key = adj_key(a.ckpt_interval, fam != kRb23F32);
  if (!key)
    return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d",
                     a.ckpt_interval, kMaxCkptInterval);
  const bool is_pair = std::strcmp(key, "adj_pair") == 0;
  if (is_pair && a.nsave == 0 && !a.pbar_in) key = "adj_pair_plain";
  return launch(field, fam, key, a.n, is_pair ? kFwdBlock : kAdjBlock, a, s);
}

int r5_launch(int field, Family fam, const char* key, int64_t n, const FwdArgs& a,
              hipStream_t s) {
  return launch(field, fam, key, n, kFwdBlock, a, s);
}

int r5_launch(int field, Family fam, const char* key, int64_t n, const ReplayArgs& a,
              hipStream_t s) {
  return launch(field, fam, key, n, kFwdBlock, a, s);
}

int r5_launch(int field, Family fam, const char* key, int64_t n, const SaveArgs& a,
              hipStream_t s) {
  return launch(field, fam, key, n, kFwdBlock, a, s);
}

int r5_launch(int field, Family fam, const char* key, int64_t n, const SaveVjpArgs& a,
              hipStream_t s) {
  return launch(field, fam, key, n, kFwdBlock, a, s);
}

int r5_adj(int field, Family fam, const AdjArgs& a, hipStream_t s) {
  const char* key = adj_key(a.ckpt_interval, false);
  if (!key)
    return set_error(GSX_E_INVALID, "ckpt_interval %d exceeds the adjoint's maximum of %d",
                     a.ckpt_interval, kMaxCkptInterval);
  return launch(field, fam, key, a.n, kAdjBlock, a, s);
}

int rb23_saveat(int field, Family fam, const SaveArgs& a, hipStream_t s) {
  return launch(field, fam, "saveat", a.n, kFwdBlock, a, s);
}

int rb23_saveat_vjp(int field, Family fam, const SaveVjpArgs& a, hipStream_t s) {
  return launch(field, fam, "saveat_vjp", a.n, kFwdBlock, a, s);
}

int tsit5w_fwd(int field, const WideFwdArgs& w, hipStream_t s) {
  return launch(field, kTsit5Wide, "fwd", w.f.n, kFwdBlock, w, s);
}

int tsit5w_replay(int field, const WideReplayArgs& w, hipStream_t s) {
  return launch(field, kTsit5Wide, "replay", w.r.n, kFwdBlock, w, s);
}

int tsit5w_adj(int field, const WideAdjArgs& w, hipStream_t s) {
  return launch(field, kTsit5Wide, "adj", w.a.n, kFwdBlock, w, s);
}

}  // namespace rt
}  // namespace gsx

using namespace gsx;

extern "C" {

int gsx_field_register(const char* source, int D, int P, int flags, int* field_id) {
  if (!source || !field_id) return set_error(GSX_E_INVALID, "field_register: null argument");
  if (D < 1 || D > rt::kMaxWideD)
    return set_error(GSX_E_INVALID, "field_register: D=%d outside 1..%d", D, rt::kMaxWideD);
  if (P < 0 || P > 64) return set_error(GSX_E_INVALID, "field_register: P=%d outside 0..64", P);
  if (flags & ~(GSX_FIELD_STIFF | GSX_FIELD_TIME_DEPENDENT))
    return set_error(GSX_E_INVALID, "field_register: unknown flags 0x%x", flags);
  const bool stiff = (flags & GSX_FIELD_STIFF) != 0;
  const bool autonomous = (flags & GSX_FIELD_TIME_DEPENDENT) == 0;
  // The stiff lanes hold a D x D W and its LU per lane (gradsolve/warp/warp_rosenbrock.py:67,
  // 324-340 refuses its fused stiff kernel above the same 12).
  if (stiff && D > rt::kMaxStiffD)
    return set_error(GSX_E_INVALID, "field_register: a stiff field of D=%d is above the register "
                     "limit of the fused stiff lanes (D <= %d: each lane factorises a D x D "
                     "W in registers)", D, rt::kMaxStiffD);
  std::lock_guard<std::mutex> lk(rt::g_mu);
  auto f = std::make_unique<rt::RtField>();
  f->D = D;
  f->P = P;
  f->stiff = stiff;
  f->autonomous = autonomous;
  f->source = source;
  // The recording forward of the primary family is compiled now, so that errors in the
  // source surface at registration with the compiler's log; a stiff field that reads t also
  // compiles the Rodas5P forward, the one user of its dfdt.
  const int id = GSX_FIELD_RUNTIME_BASE + static_cast<int>(rt::g_fields.size());
  rt::Module* m = nullptr;
  int rc = stiff ? rt::compile(*f, rt::kRb23, "fwd", &m)
                 : D > rt::kMaxRegisterD ? rt::compile(*f, rt::kTsit5Wide, "fwd", &m)
                                         : rt::compile(*f, rt::kTsit5, "fwd_rec", &m);
  if (rc != GSX_OK) return rt::fail(*m, id);
  if (stiff && !autonomous) {
    rc = rt::compile(*f, rt::kRodas5p, "fwd", &m);
    if (rc != GSX_OK) return rt::fail(*m, id);
  }
  rt::g_fields.push_back(std::move(f));
  *field_id = id;
  return GSX_OK;
}

int gsx_field_release(int field_id) {
  std::lock_guard<std::mutex> lk(rt::g_mu);
  rt::RtField* f = rt::lookup(field_id);
  if (!f) return set_error(GSX_E_INVALID, "field %d is not a registered field", field_id);
  int dev0 = 0;
  const bool restore = hipGetDevice(&dev0) == hipSuccess;
  for (auto& km : f->kernels) {
    rt::Module& m = km.second;
    for (auto& kv : m.loaded) {
      if (hipSetDevice(kv.first) == hipSuccess) (void)hipModuleUnload(kv.second);
    }
    m.loaded.clear();
    m.fn.clear();
  }
  if (restore) (void)hipSetDevice(dev0);
  f->released = true;
  return GSX_OK;
}

}  // extern "C"
