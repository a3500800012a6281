// gsx_core.h -- per-trajectory (per-lane) building blocks of the ensemble ODE kernels.
//
// Everything here is written once for one trajectory and instantiated by the HIP kernels
// with one trajectory per lane. The functions are __host__ __device__ so that the same
// source also compiles as plain C++ for a CPU test harness that checks indexing and
// arithmetic against the oracle without a GPU; the product library only ever runs them
// on the device.
//
// Memory layout in HBM (structure of arrays, index j = trajectory, coalesced across lanes):
//   y0, y_final, ybar, y0bar : [D][n]        element (d, j) at d*n + j
//   p, pbar                  : [P][n]
//   tape                     : [cap][n]      accepted step s of trajectory j at s*n + j
//   ckpt                     : [slots][DCK][n] state after m*C accepted steps at slot m
//                              (DCK = D, plus one row holding t for non-autonomous fields)
#pragma once

#include "gsx_platform.h"
#include "gsx_args.h"

// The computational code lives in GSX_NS. The library compiles it twice: once in namespace
// gsx with fused multiply-add contraction (the fast variant), and once in namespace
// gsx_strict with contraction off, whose forward rounds every operation exactly as the
// IEEE-sequential reference loop does.
#ifndef GSX_NS
#define GSX_NS gsx
#endif

namespace GSX_NS {

using ::gsx::AdjArgs;
using ::gsx::FwdArgs;
using ::gsx::ReplayArgs;
using ::gsx::SaveArgs;
using ::gsx::SaveVjpArgs;
using ::gsx::kDegenerate;
using ::gsx::kExhausted;
using ::gsx::kReached;
using ::gsx::kTapeFull;

// ---------------------------------------------------------------------------------------
// Tsitouras (2011) 5(4) pair (Comput. Math. Appl. 62(2), 770-775). The decimal strings and
// the binary64 expressions (e_i = b_i - bhat_i) are those of the published table as
// gradsolve holds it in gradsolve/solvers/tsit5_step.py:31-77.
// ---------------------------------------------------------------------------------------
namespace tab {
constexpr double c2 = 161.0 / 1000.0;
constexpr double c3 = 327.0 / 1000.0;
constexpr double c4 = 9.0 / 10.0;
constexpr double c5 = 0.9800255409045096857298102862870245954942137979563024768854764293221195950761080302604;
constexpr double c6 = 1.0;
constexpr double a21 = 161.0 / 1000.0;
constexpr double a31 = -0.8480655492356988544426874250230774675121177393430391537369234245294192976164141156943e-2;
constexpr double a32 = 0.3354806554923569885444268742502307746751211773934303915373692342452941929761641411569;
constexpr double a41 = 2.897153057105493432130432594192938764924887287701866490314866693455023795137503079289;
constexpr double a42 = -6.359448489975074843148159912383825625952700647415626703305928850207288721235210244366;
constexpr double a43 = 4.362295432869581411017727318190886861027813359713760212991062156752264926097707165077;
constexpr double a51 = 5.325864828439256604428877920840511317836476253097040101202360397727981648835607691791;
constexpr double a52 = -11.74888356406282787774717033978577296188744178259862899288666928009020615663593781589;
constexpr double a53 = 7.495539342889836208304604784564358155658679161518186721010132816213648793440552049753;
constexpr double a54 = -0.9249506636175524925650207933207191611349983406029535244034750452930469056411389539635e-1;
constexpr double a61 = 5.861455442946420028659251486982647890394337666164814434818157239052507339770711679748;
constexpr double a62 = -12.92096931784710929170611868178335939541780751955743459166312250439928519268343184452;
constexpr double a63 = 8.159367898576158643180400794539253485181918321135053305748355423955009222648673734986;
constexpr double a64 = -0.7158497328140099722453054252582973869127213147363544882721139659546372402303777878835e-1;
constexpr double a65 = -0.2826905039406838290900305721271224146717633626879770007617876201276764571291579142206e-1;
constexpr double b1 = 0.9646076681806522951816731316512876333711995238157997181903319145764851595234062815396e-1;
constexpr double b2 = 1.0 / 100.0;
constexpr double b3 = 0.4798896504144995747752495322905965199130404621990332488332634944254542060153074523509;
constexpr double b4 = 1.379008574103741893192274821856872770756462643091360525934940067397245698027561293331;
constexpr double b5 = -3.290069515436080679901047585711363850115683290894936158531296799594813811049925401677;
constexpr double b6 = 2.324710524099773982415355918398765796109060233222962411944060046314465391054716027841;
constexpr double e1 = 0.9646076681806522951816731316512876333711995238157997181903319145764851595234062815396e-1 - 0.9468075576583945807478876255758922856117527357724631226139574065785592789071067303271e-1;
constexpr double e2 = 1.0 / 100.0 - 0.9183565540343253096776363936645313759813746240984095238905939532922955247253608687270e-2;
constexpr double e3 = 0.4798896504144995747752495322905965199130404621990332488332634944254542060153074523509 - 0.4877705284247615707855642599631228241516691959761363774365216240304071651579571959813;
constexpr double e4 = 1.379008574103741893192274821856872770756462643091360525934940067397245698027561293331 - 1.234297566930478985655109673884237654035539930748192848315425833500484878378061439761;
constexpr double e5 = -3.290069515436080679901047585711363850115683290894936158531296799594813811049925401677 + 2.707712349983525454881109975059321670689605166938197378763992255714444407154902012702;
constexpr double e6 = 2.324710524099773982415355918398765796109060233222962411944060046314465391054716027841 - 1.866628418170587035753719399566211498666255505244122593996591602841258328965767580089;
constexpr double e7 = -1.0 / 66.0;
}  // namespace tab

// Workgroup sizes: forward and replay kernels run 256 lanes per workgroup; the segment
// adjoints run one wavefront per workgroup so that their LDS is allocated per wave.
constexpr int kFwdBlock = 256;
constexpr int kAdjBlock = 64;

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
// The largest v over the wavefront's 64 lanes; every lane of the wavefront must call it. The
// adjoint kernels pass the wavefront's largest segment count to their lanes, so that the lanes
// walk the segments in step and each checkpoint and tape read of the wavefront is one
// contiguous row segment.
__device__ __forceinline__ int32_t wave_max(int32_t v) {
  GSX_UNROLL for (int off = 32; off >= 1; off >>= 1) {
    const int32_t o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}
#endif

// Step-size controller constants (Hairer-Norsett-Wanner II.4; gradsolve/solvers/tsit5_step.py:26).
constexpr double kSafety = 0.9;
constexpr double kErrFloor = 1e-10;
constexpr double kFacMin = 0.2;
constexpr double kFacMax = 5.0;
constexpr double kTsit5Exp = -0.2;

// ---------------------------------------------------------------------------------------
// Controller power and error-norm reciprocal.
//
// The strict variant (GSX_STRICT_ARITH, contraction off) calls the correctly rounded IEEE
// division and the math library's pow, as the reference loop does. The default variant
// computes x^e for the controller exponents e = -0.2 (Tsit5) and e = -1/3 (Rosenbrock23)
// from a single-precision hardware seed (2^(e log2 x)) refined in double precision against
// g(y) = x y^m - 1 (m = 5 or 3; 10 for the squared norm), with the factor x^(e + 1/m)
// = exp((e + 1/m) ln x) that separates the binary64 exponent from -1/m; the result is
// within an ulp of the correctly rounded power. Inputs are first clamped to
// [1e-5, 1e4]: outside that range the controller factor 0.9 x^e saturates at the
// [0.2, 5] clamp for both exponents, so the step size is unchanged.
// ---------------------------------------------------------------------------------------
// min/max of operands that are never NaN (step sizes, error scales, controller factors):
// the hardware instructions in the default variant, Python's comparison form (b < a ? b : a)
// in the strict one; for non-NaN operands of equal zero sign both give the same value.
GSX_HD double nn_min(double a, double b) {
#if defined(GSX_STRICT_ARITH)
  return b < a ? b : a;
#else
  return std::fmin(a, b);
#endif
}
GSX_HD double nn_max(double a, double b) {
#if defined(GSX_STRICT_ARITH)
  return b > a ? b : a;
#else
  return std::fmax(a, b);
#endif
}

#if defined(__HIP_DEVICE_COMPILE__)
GSX_HD float fast_log2f(float x) { return __builtin_amdgcn_logf(x); }
GSX_HD float fast_exp2f(float x) { return __builtin_amdgcn_exp2f(x); }
#else
GSX_HD float fast_log2f(float x) { return std::log2(x); }
GSX_HD float fast_exp2f(float x) { return std::exp2(x); }
#endif

// x^(-1/m) * exp(-c ln x) for positive finite x, m in {3, 5, 10}: see above. For m = 10 the
// input is a squared norm and the clamp is squared with it.
//
// The refinement is one step of the binomial series rather than two Newton steps: with the
// seed y (relative error e below 1e-6) and u = x y^m - 1 (about m e), the root is
// y (1 + u)^(-1/m) = y (1 - u/m + (m+1)/(2 m^2) u^2 - O(u^3)), and the cubic term (below
// 1e-17 relative) is dropped. The exponent factor exp(-c ln x) ~ 1 - c ln x enters the same
// final fused multiply-add, so the step costs m's powers, three fused multiply-adds and the
// final one.
template <int M>
GSX_HD double ctrl_pow_newton(double x, double c) {
  constexpr double lo = M == 10 ? 1e-10 : 1e-5, hi = M == 10 ? 1e8 : 1e4;
  constexpr double c1 = -1.0 / M, c2 = (M + 1.0) / (2.0 * M * M);
  const double xc = nn_min(nn_max(x, lo), hi);
  const float lf = fast_log2f(static_cast<float>(xc));
  const double y = static_cast<double>(fast_exp2f(lf * (-1.0f / M)));
  const double y2 = y * y;
  const double y5 = y2 * y2 * y;
  const double ym = (M == 10) ? y5 * y5 : (M == 5) ? y5 : y2 * y;
  const double u = std::fma(xc, ym, -1.0);
  const double d = static_cast<double>(lf) * (-c * 0.6931471805599453);  // -c ln x
  const double q = std::fma(u, std::fma(u, c2, c1), d);
  return std::fma(y, q, y);
}

// x^(-0.2) with -0.2 in binary64: the double nearest 0.2 exceeds 1/5 by
// 1.1102230246251565e-17, so x^(-0.2) = x^(-1/5) * exp(-1.1102230246251565e-17 * ln x).
GSX_HD double ctrl_pow_tsit5(double x) {
#if defined(GSX_STRICT_ARITH)
  return std::pow(x, -0.2);
#else
  if (!(x < 1e300)) return std::pow(x, -0.2);  // inf (and NaN) keep the library's result
  return ctrl_pow_newton<5>(x, 1.1102230246251565e-17);
#endif
}

// m^(-0.1) for the squared norm m = err^2, equal to err^(-0.2) with -0.2 in binary64:
// m^(-1/10) * exp(-5.551115123125783e-18 * ln m) (half the correction above).
GSX_HD double ctrl_pow_tsit5_sq(double m) {
  if (!(m < 1e300)) return std::pow(m, -0.1);  // inf (and NaN) keep the library's result
  return ctrl_pow_newton<10>(m, 5.551115123125783e-18);
}

// x^(-1/3) with -1/3 in binary64: the double nearest 1/3 falls short of it by
// 1.850371707708594e-17, so x^(-1/3) = x^(-1/3 exact) * exp(+1.850371707708594e-17 * ln x).
GSX_HD double ctrl_pow_rb23(double x) {
#if defined(GSX_STRICT_ARITH)
  return std::pow(x, -1.0 / 3.0);
#else
  if (!(x < 1e300)) return std::pow(x, -1.0 / 3.0);
  return ctrl_pow_newton<3>(x, -1.850371707708594e-17);
#endif
}

// The seed of a refined reciprocal: the hardware's double-precision reciprocal on the device.
// Host builds seed with the single-precision reciprocal, of similar accuracy, so that the
// refinement can be tested on the CPU; where 1/b lies outside the single-precision normal
// range (which the device's seed covers) they take the double quotient instead.
GSX_HD double recip_seed(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcp(b);
#else
  const double q = 1.0 / b;
  const double m = std::fabs(q);
  return (m > 1.2e-38 && m < 3.4e38) ? static_cast<double>(static_cast<float>(q)) : q;
#endif
}

// a / b for the error norm (b >= atol > 0): the IEEE quotient in the strict variant; in the
// default variant the hardware reciprocal refined by one Newton step, the product a r, and
// one Markstein correction q + r (a - b q) of the quotient.
GSX_HD double norm_div(double a, double b) {
#if defined(GSX_STRICT_ARITH)
  return a / b;
#else
  double r = recip_seed(b);
  r = std::fma(r, std::fma(-b, r, 1.0), r);
  const double q = a * r;
  return std::fma(std::fma(-b, q, a), r, q);
#endif
}

// sin(x) and cos(x) together (the Kuramoto field's phases). For |x| < 2^20: the nearest
// integer k to x * 2/pi (added to 1.5 * 2^52 in one fused multiply-add, which rounds the sum to
// an integer; the low bits of that double are k mod 4) and r = x - k*pi/2 by a three-part split
// of pi/2 (hi + mid + lo, each product inside a fused multiply-add; the first subtraction is
// exact because k*hi and x are multiples of 2^-52 in that range), then polynomials on
// |r| <= pi/4: sin r = r + r^3 S(r^2) and cos r = 1 - r^2/2 + r^4 C(r^2) with S and C of degree
// 5 (Chebyshev fits to (sin r - r)/r^3 and (cos r - 1 + r^2/2)/r^4 over [0, (pi/4)^2], computed
// in 50-digit arithmetic and rounded to double; the fits' own errors are below 2^-55 and 2^-59
// of the results), and the quadrant k mod 4 swaps them and flips their signs. Within 2 ulp of
// the correctly rounded values. Larger |x| and non-finite x go to the math library. Every
// rounding step is an explicit operation, so host and device builds return the same bits.
// gsx_sincos_core is the evaluation without the range branch, for callers that keep a block of
// them branch-free (the Kuramoto field interleaves them with its MFMA products) and redo the
// evaluation with gsx_sincos when some |x| >= 2^20. For such x (and non-finite x) the results
// are meaningless but well defined.
GSX_HD void gsx_sincos_core(double x, double* s_out, double* c_out) {
  constexpr double kTwoOverPi = 0.6366197723675814;
  constexpr double kShift = 6755399441055744.0;  // 1.5 * 2^52
  constexpr double kPio2Hi = 1.5707963267948966;
  constexpr double kPio2Mid = 6.123233995736766e-17;
  constexpr double kPio2Lo = -1.4973849048591698e-33;
  const double kd = std::fma(x, kTwoOverPi, kShift);
  const double k = kd - kShift;
  const uint32_t q = static_cast<uint32_t>(__builtin_bit_cast(uint64_t, kd)) & 3u;
  double r = std::fma(-k, kPio2Hi, x);
  r = std::fma(-k, kPio2Mid, r);
  r = std::fma(-k, kPio2Lo, r);
  const double z = r * r;
  double ps = 1.5918129294866608e-10;
  ps = std::fma(z, ps, -2.5051131845003624e-08);
  ps = std::fma(z, ps, 2.755731610255244e-06);
  ps = std::fma(z, ps, -0.00019841269836758574);
  ps = std::fma(z, ps, 0.008333333333330948);
  ps = std::fma(z, ps, -0.16666666666666666);
  const double sr = std::fma(r * z, ps, r);
  double pc = -1.1382632425521717e-11;
  pc = std::fma(z, pc, 2.08761462684032e-09);
  pc = std::fma(z, pc, -2.7557317271729793e-07);
  pc = std::fma(z, pc, 2.480158729876569e-05);
  pc = std::fma(z, pc, -0.0013888888888887398);
  pc = std::fma(z, pc, 0.041666666666666664);
  pc = std::fma(z, pc, -0.5);
  const double cr = std::fma(z, pc, 1.0);
  const double sa = (q & 1u) ? cr : sr;
  const double ca = (q & 1u) ? sr : cr;
  // sin(r + q pi/2) = +-sa (negative for q = 2, 3), cos = +-ca (negative for q = 1, 2)
  const uint64_t sflip = static_cast<uint64_t>(q & 2u) << 62;
  const uint64_t cflip = static_cast<uint64_t>((q + 1u) & 2u) << 62;
  *s_out = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, sa) ^ sflip);
  *c_out = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, ca) ^ cflip);
}

GSX_HD void gsx_sincos(double x, double* s_out, double* c_out) {
  if (!(std::fabs(x) < 1048576.0)) {
    *s_out = std::sin(x);
    *c_out = std::cos(x);
    return;
  }
  gsx_sincos_core(x, s_out, c_out);
}

// Single-precision counterpart (the FP32 Kuramoto variant): for |x| < 2^14 the same reduction
// with 1.5 * 2^23 as the rounding constant and an FP32 split of pi/2 (hi = 0x1.921fb6p+0,
// mid = -0x1.777a5cp-25, lo = -0x1.ee59dap-50; the first subtraction is exact because k*hi and
// x are multiples of 2^-23 there), then sin r = r + r^3 S(r^2) with S of degree 3 and
// cos r = 1 - r^2/2 + r^4 C(r^2) with C of degree 2 (Chebyshev fits as above, rounded to
// single precision; fit errors below 2^-36 and 2^-30 of the results). Larger |x| and non-finite
// x go to the math library (gsx_sincosf); gsx_sincosf_core is the evaluation without that
// branch, as gsx_sincos_core.
GSX_HD void gsx_sincosf_core(float x, float* s_out, float* c_out) {
  constexpr float kTwoOverPi = 0.636619746685028f;
  constexpr float kShift = 12582912.0f;  // 1.5 * 2^23
  constexpr float kPio2Hi = 0x1.921fb6p+0f;
  constexpr float kPio2Mid = -0x1.777a5cp-25f;
  constexpr float kPio2Lo = -0x1.ee59dap-50f;
  const float kd = std::fma(x, kTwoOverPi, kShift);
  const float k = kd - kShift;
  const uint32_t q = __builtin_bit_cast(uint32_t, kd) & 3u;
  float r = std::fma(-k, kPio2Hi, x);
  r = std::fma(-k, kPio2Mid, r);
  r = std::fma(-k, kPio2Lo, r);
  const float z = r * r;
  float ps = 0x1.6dbe08p-19f;
  ps = std::fma(z, ps, -0x1.a013a8p-13f);
  ps = std::fma(z, ps, 0x1.11110ep-7f);
  ps = std::fma(z, ps, -0x1.555556p-3f);
  const float sr = std::fma(r * z, ps, r);
  float pc = 0x1.9bd89cp-16f;
  pc = std::fma(z, pc, -0x1.6c12d2p-10f);
  pc = std::fma(z, pc, 0x1.555554p-5f);
  pc = std::fma(z, pc, -0.5f);
  const float cr = std::fma(z, pc, 1.0f);
  const float sa = (q & 1u) ? cr : sr;
  const float ca = (q & 1u) ? sr : cr;
  const uint32_t sflip = (q & 2u) << 30;
  const uint32_t cflip = ((q + 1u) & 2u) << 30;
  *s_out = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, sa) ^ sflip);
  *c_out = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, ca) ^ cflip);
}

GSX_HD void gsx_sincosf(float x, float* s_out, float* c_out) {
  if (!(std::fabs(x) < 16384.0f)) {
    *s_out = std::sin(x);
    *c_out = std::cos(x);
    return;
  }
  gsx_sincosf_core(x, s_out, c_out);
}

// Python's builtin min/max argument semantics (min(a, b) is a unless b < a), which the
// reference controller uses (gradsolve/solvers/tsit5_replay.py:108,124). They propagate a NaN
// in the first operand, unlike fmin/fmax.
template <class T>
GSX_HD T py_min(T a, T b) { return b < a ? b : a; }
template <class T>
GSX_HD T py_max(T a, T b) { return b > a ? b : a; }


// ---------------------------------------------------------------------------------------
// Fields. f(t, y, p, dy) and the vector-Jacobian product vjp(t, y, p, v, ybar+=, pbar+=).
// Each field's textual association follows the registered field it restates.
// ---------------------------------------------------------------------------------------

// Lorenz (D=3, P=1: rho; sigma=10, beta=8/3): gradsolve/warp/_warp_kernel.py:111-130,
// gradsolve/cuda/_fields.py:12-18.
struct FLorenz {
  static constexpr int D = 3, P = 1;
  static constexpr bool autonomous = true;
  template <class T>
  GSX_HD static void f(T, const T* y, const T* p, T* dy, const double*) {
    const T sigma = T(10.0), beta = T(8.0 / 3.0), rho = p[0];
    dy[0] = sigma * (y[1] - y[0]);
    dy[1] = rho * y[0] - y[1] - y[0] * y[2];
    dy[2] = y[0] * y[1] - beta * y[2];
  }
  template <class T>
  GSX_HD static void vjp(T, const T* y, const T* p, const T* v, T* yb, T* pb, const double*) {
    const T sigma = T(10.0), beta = T(8.0 / 3.0), rho = p[0];
    yb[0] += -sigma * v[0] + (rho - y[2]) * v[1] + y[1] * v[2];
    yb[1] += sigma * v[0] - v[1] + y[0] * v[2];
    yb[2] += -y[0] * v[1] - beta * v[2];
    pb[0] += y[0] * v[1];
  }
};

// Lorenz with all three parameters free (D=3, P=3: sigma, beta, rho), the "lorenz" problem's
// [sigma, beta, rho] columns (gradsolve/warp/warp_ode.py:552-589). Same association as FLorenz,
// so at sigma=10, beta=8/3 both fields round identically.
struct FLorenz3 {
  static constexpr int D = 3, P = 3;
  static constexpr bool autonomous = true;
  template <class T>
  GSX_HD static void f(T, const T* y, const T* p, T* dy, const double*) {
    const T sigma = p[0], beta = p[1], rho = p[2];
    dy[0] = sigma * (y[1] - y[0]);
    dy[1] = rho * y[0] - y[1] - y[0] * y[2];
    dy[2] = y[0] * y[1] - beta * y[2];
  }
  template <class T>
  GSX_HD static void vjp(T, const T* y, const T* p, const T* v, T* yb, T* pb, const double*) {
    const T sigma = p[0], beta = p[1], rho = p[2];
    yb[0] += -sigma * v[0] + (rho - y[2]) * v[1] + y[1] * v[2];
    yb[1] += sigma * v[0] - v[1] + y[0] * v[2];
    yb[2] += -y[0] * v[1] - beta * v[2];
    pb[0] += (y[1] - y[0]) * v[0];
    pb[1] -= y[2] * v[2];
    pb[2] += y[0] * v[1];
  }
};

// Linear decay f = -p0*y (test field with a closed-form solution), fixed D.
template <int DD>
struct FLinearDecay {
  static constexpr int D = DD, P = 1;
  static constexpr bool autonomous = true;
  template <class T>
  GSX_HD static void f(T, const T* y, const T* p, T* dy, const double*) {
    GSX_UNROLL for (int i = 0; i < D; ++i) dy[i] = -p[0] * y[i];
  }
  template <class T>
  GSX_HD static void vjp(T, const T* y, const T* p, const T* v, T* yb, T* pb, const double*) {
    T acc = T(0);
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      yb[i] += -p[0] * v[i];
      acc += y[i] * v[i];
    }
    pb[0] -= acc;
  }
  // Analytic Jacobian (row-major) and d/d(y,p) of sum_ab Jbar_ab J_ab.
  template <class T>
  GSX_HD static void jac(T, const T*, const T* p, T* J, const double*) {
    GSX_UNROLL for (int i = 0; i < D * D; ++i) J[i] = T(0);
    GSX_UNROLL for (int i = 0; i < D; ++i) J[i * D + i] = -p[0];
  }
  template <class T>
  GSX_HD static void jac_vjp(T, const T*, const T*, const T* Jb, T*, T* pb, const double*) {
    T tr = T(0);
    GSX_UNROLL for (int i = 0; i < D; ++i) tr += Jb[i * D + i];
    pb[0] -= tr;
  }
};

// Robertson kinetics (D=3, P=3: k1, k2, k3), stiff: gradsolve/warp/_warp_rosenbrock.py:96-149
// (field and analytic Jacobian), gradsolve tests/_oracles/gradsolve_g2_spike_oracle.py:53-69.
struct FRobertson {
  static constexpr int D = 3, P = 3;
  static constexpr bool autonomous = true;
  template <class T>
  GSX_HD static void f(T, const T* y, const T* p, T* dy, const double*) {
    const T k1 = p[0], k2 = p[1], k3 = p[2];
    dy[0] = -k1 * y[0] + k3 * y[1] * y[2];
    dy[1] = k1 * y[0] - k3 * y[1] * y[2] - k2 * y[1] * y[1];
    dy[2] = k2 * y[1] * y[1];
  }
  template <class T>
  GSX_HD static void vjp(T, const T* y, const T* p, const T* v, T* yb, T* pb, const double*) {
    const T k1 = p[0], k2 = p[1], k3 = p[2];
    yb[0] += -k1 * v[0] + k1 * v[1];
    yb[1] += k3 * y[2] * v[0] - (k3 * y[2] + T(2.0) * k2 * y[1]) * v[1] +
             T(2.0) * k2 * y[1] * v[2];
    yb[2] += k3 * y[1] * v[0] - k3 * y[1] * v[1];
    pb[0] += y[0] * (v[1] - v[0]);
    pb[1] += y[1] * y[1] * (v[2] - v[1]);
    pb[2] += y[1] * y[2] * (v[0] - v[1]);
  }
  template <class T>
  GSX_HD static void jac(T, const T* y, const T* p, T* J, const double*) {
    const T k1 = p[0], k2 = p[1], k3 = p[2];
    const T y1 = y[1], y2 = y[2];
    J[0] = -k1;
    J[1] = k3 * y2;
    J[2] = k3 * y1;
    J[3] = k1;
    J[4] = -k3 * y2 - T(2.0) * k2 * y1;
    J[5] = -k3 * y1;
    J[6] = T(0.0);
    J[7] = T(2.0) * k2 * y1;
    J[8] = T(0.0);
  }
  template <class T>
  GSX_HD static void jac_vjp(T, const T* y, const T* p, const T* Jb, T* yb, T* pb,
                             const double*) {
    const T k2 = p[1], k3 = p[2];
    const T y1 = y[1], y2 = y[2];
    yb[1] += k3 * Jb[2] - T(2.0) * k2 * Jb[4] - k3 * Jb[5] + T(2.0) * k2 * Jb[7];
    yb[2] += k3 * Jb[1] - k3 * Jb[4];
    pb[0] += Jb[3] - Jb[0];
    pb[1] += T(2.0) * y1 * (Jb[7] - Jb[4]);
    pb[2] += y2 * Jb[1] + y1 * Jb[2] - y2 * Jb[4] - y1 * Jb[5];
  }
};

// A stiff field that reads t (D=2, P=3: a, b, c), for the Rodas5P lane's df/dt term:
//   f0 = -a (y0 - y1^2) + 0.4 sin 3t,  f1 = y0 - y1 - b y1^3 + 0.2 cos t + c sin(t) y0.
// At (a, b, c) = (30, 0.7, 0) it is the time-dependent stiff test field of gradsolve's
// tests/test_rodas5p_dense.py:94-103; c makes J depend on t and df/dt on y0 and c.
// dfdt gives df/dt at (t, y); dfdt_vjp adds d/d(y, p) of ftb . df/dt.
struct FForcedStiff {
  static constexpr int D = 2, P = 3;
  static constexpr bool autonomous = false;
  template <class T>
  GSX_HD static void f(T t, const T* y, const T* p, T* dy, const double*) {
    const T a = p[0], b = p[1], c = p[2];
    const T st = sin(t);
    dy[0] = -a * (y[0] - y[1] * y[1]) + T(0.4) * sin(T(3.0) * t);
    dy[1] = y[0] - y[1] - b * y[1] * y[1] * y[1] + T(0.2) * cos(t) + c * st * y[0];
  }
  template <class T>
  GSX_HD static void vjp(T t, const T* y, const T* p, const T* v, T* yb, T* pb, const double*) {
    const T a = p[0], b = p[1], c = p[2];
    const T st = sin(t);
    yb[0] += -a * v[0] + (T(1.0) + c * st) * v[1];
    yb[1] += T(2.0) * a * y[1] * v[0] - (T(1.0) + T(3.0) * b * y[1] * y[1]) * v[1];
    pb[0] -= (y[0] - y[1] * y[1]) * v[0];
    pb[1] -= y[1] * y[1] * y[1] * v[1];
    pb[2] += st * y[0] * v[1];
  }
  template <class T>
  GSX_HD static void jac(T t, const T* y, const T* p, T* J, const double*) {
    const T a = p[0], b = p[1], c = p[2];
    const T y1 = y[1];
    J[0] = -a;
    J[1] = T(2.0) * a * y1;
    J[2] = T(1.0) + c * sin(t);
    J[3] = -T(1.0) - T(3.0) * b * y1 * y1;
  }
  template <class T>
  GSX_HD static void jac_vjp(T t, const T* y, const T* p, const T* Jb, T* yb, T* pb,
                             const double*) {
    const T a = p[0], b = p[1];
    const T y1 = y[1];
    yb[1] += T(2.0) * a * Jb[1] - T(6.0) * b * y1 * Jb[3];
    pb[0] += T(2.0) * y1 * Jb[1] - Jb[0];
    pb[1] -= T(3.0) * y1 * y1 * Jb[3];
    pb[2] += sin(t) * Jb[2];
  }
  template <class T>
  GSX_HD static void dfdt(T t, const T* y, const T* p, T* ft, const double*) {
    const T c = p[2];
    ft[0] = T(1.2) * cos(T(3.0) * t);
    ft[1] = -T(0.2) * sin(t) + c * cos(t) * y[0];
  }
  template <class T>
  GSX_HD static void dfdt_vjp(T t, const T* y, const T* p, const T* ftb, T* yb, T* pb,
                              const double*) {
    const T ct = cos(t);
    yb[0] += p[2] * ct * ftb[1];
    pb[2] += ct * y[0] * ftb[1];
  }
};

// Van der Pol (D=2, P=1: mu): gradsolve/warp/_warp_kernel.py:158-174.
struct FVdp {
  static constexpr int D = 2, P = 1;
  static constexpr bool autonomous = true;
  template <class T>
  GSX_HD static void f(T, const T* y, const T* p, T* dy, const double*) {
    const T x = y[0], v = y[1];
    dy[0] = v;
    dy[1] = p[0] * ((T(1.0) - x * x) * v - x);
  }
  template <class T>
  GSX_HD static void vjp(T, const T* y, const T* p, const T* v, T* yb, T* pb, const double*) {
    const T x = y[0], w = y[1];
    yb[0] += p[0] * (T(-2.0) * x * w - T(1.0)) * v[1];
    yb[1] += v[0] + p[0] * (T(1.0) - x * x) * v[1];
    pb[0] += ((T(1.0) - x * x) * w - x) * v[1];
  }
};

// ---------------------------------------------------------------------------------------
// NumPy-order reductions: the reference's WRMS norm is np.mean((e/sc)**2)
// (gradsolve/solvers/tsit5_replay.py:118). Below 8 elements NumPy sums left to right; from 8
// to 128 it uses eight interleaved accumulators combined pairwise.
// ---------------------------------------------------------------------------------------
template <int D, class T>
GSX_HD T np_sum(const T* a) {
  if constexpr (D < 8) {
    T r = a[0];
    GSX_UNROLL for (int i = 1; i < D; ++i) r = r + a[i];
    return r;
  } else {
    static_assert(D <= 128, "pairwise block limit");
    T r[8];
    GSX_UNROLL for (int q = 0; q < 8; ++q) r[q] = a[q];
    constexpr int full = D - (D % 8);
    GSX_UNROLL for (int i = 8; i < full; i += 8) {
      GSX_UNROLL for (int q = 0; q < 8; ++q) r[q] = r[q] + a[i + q];
    }
    T res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    GSX_UNROLL for (int i = full; i < D; ++i) res = res + a[i];
    return res;
  }
}

template <int D>
GSX_HD double wrms(const double* e, const double* y, const double* ynew, double rtol,
                   double atol) {
  double q[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    const double sc = atol + rtol * py_max(std::fabs(y[i]), std::fabs(ynew[i]));
    const double r = norm_div(e[i], sc);
    q[i] = r * r;
  }
  return std::sqrt(np_sum<D>(q) / double(D));
}

// ---------------------------------------------------------------------------------------
// Tsit5 error measure and step factor (gradsolve/solvers/method.py:85-108). The strict
// variant follows the reference: err = sqrt(mean q^2) (the WRMS norm), accept iff err <= 1,
// factor 0.9 * max(err, 1e-10)^-0.2. The default variant carries m = mean q^2 = err^2
// instead and never forms the square root: m <= 1 accepts the same steps (the root is
// monotone and 1 at 1; the one exception is the double just above 1, whose rounded root is
// 1), and the factor m^(-0.1) (ctrl_pow_tsit5_sq) equals err^(-0.2) to within the
// controller power's own rounding. The mean divides by D with one Markstein correction
// (q = s/D to within an ulp) instead of the IEEE division sequence.
// ---------------------------------------------------------------------------------------
template <int D>
GSX_HD double wrms_sq(const double* e, const double* y, const double* ynew, double rtol,
                      double atol) {
  double q[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) {
    const double sc = atol + rtol * nn_max(std::fabs(y[i]), std::fabs(ynew[i]));
    const double r = norm_div(e[i], sc);
    q[i] = r * r;
  }
  const double s = np_sum<D>(q);
  constexpr double inv = 1.0 / D;
  if constexpr ((D & (D - 1)) == 0) return s * inv;
  const double m = s * inv;
  return std::fma(std::fma(-m, double(D), s), inv, m);
}

template <int D>
GSX_HD double tsit5_err(const double* e, const double* y, const double* ynew, double rtol,
                        double atol) {
#if defined(GSX_STRICT_ARITH)
  return wrms<D>(e, y, ynew, rtol, atol);
#else
  return wrms_sq<D>(e, y, ynew, rtol, atol);
#endif
}

// Step-size factor 0.9 * err^-0.2 from tsit5_err's measure.
GSX_HD double tsit5_fac(double em) {
#if defined(GSX_STRICT_ARITH)
  return kSafety * ctrl_pow_tsit5(py_max(em, kErrFloor));
#else
  return kSafety * ctrl_pow_tsit5_sq(nn_max(em, kErrFloor * kErrFloor));
#endif
}

// Next step size after a trial of size h with measure em: h * clamp(fac, 0.2, 5).
GSX_HD double tsit5_next_dt(double h, double em) {
  return h * nn_min(nn_max(tsit5_fac(em), kFacMin), kFacMax);
}

// ---------------------------------------------------------------------------------------
// Tsit5 stage block. Given k1 (FSAL) computes z2..z6 / k2..k6 and y5 = y + h*sum b_i k_i.
// Stage inputs follow tsit5_step.py:88-99: z_i = y + h*(a_i1 k1 + ... ).
// ---------------------------------------------------------------------------------------
// With `inc` non-null the increment h*sum(b_i k_i) is also returned (rounded once), and
// y5 = y + inc; the FP32 variant adds it to its compensated state.
template <class F, class T>
GSX_HD void tsit5_stages(T t, const T* y, T h, const T* p, T (&k)[7][F::D], T* y5,
                         const double* aux, T* inc = nullptr) {
  constexpr int D = F::D;
  using namespace tab;
  const T A21 = T(a21), A31 = T(a31), A32 = T(a32), A41 = T(a41), A42 = T(a42), A43 = T(a43);
  const T A51 = T(a51), A52 = T(a52), A53 = T(a53), A54 = T(a54);
  const T A61 = T(a61), A62 = T(a62), A63 = T(a63), A64 = T(a64), A65 = T(a65);
  const T B1 = T(b1), B2 = T(b2), B3 = T(b3), B4 = T(b4), B5 = T(b5), B6 = T(b6);
  T z[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) z[i] = y[i] + h * (A21 * k[0][i]);
  F::f(t + T(c2) * h, z, p, k[1], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) z[i] = y[i] + h * (A31 * k[0][i] + A32 * k[1][i]);
  F::f(t + T(c3) * h, z, p, k[2], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[i] = y[i] + h * (A41 * k[0][i] + A42 * k[1][i] + A43 * k[2][i]);
  F::f(t + T(c4) * h, z, p, k[3], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[i] = y[i] + h * (A51 * k[0][i] + A52 * k[1][i] + A53 * k[2][i] + A54 * k[3][i]);
  F::f(t + T(c5) * h, z, p, k[4], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[i] = y[i] + h * (A61 * k[0][i] + A62 * k[1][i] + A63 * k[2][i] + A64 * k[3][i] +
                       A65 * k[4][i]);
  F::f(t + T(c6) * h, z, p, k[5], aux);
  if (inc) {
    GSX_UNROLL for (int i = 0; i < D; ++i) {
      inc[i] = h * (B1 * k[0][i] + B2 * k[1][i] + B3 * k[2][i] + B4 * k[3][i] +
                    B5 * k[4][i] + B6 * k[5][i]);
      y5[i] = y[i] + inc[i];
    }
  } else {
    GSX_UNROLL for (int i = 0; i < D; ++i)
      y5[i] = y[i] + h * (B1 * k[0][i] + B2 * k[1][i] + B3 * k[2][i] + B4 * k[3][i] +
                          B5 * k[4][i] + B6 * k[5][i]);
  }
}

// ---------------------------------------------------------------------------------------
// Forward + accepted-h record for one trajectory (gradsolve/solvers/tsit5_replay.py:100-124,
// gradsolve/warp/_warp_kernel.py:338-399). The recorder writes h into the tape at row acc
// and the state into checkpoint slot acc/C every C accepted steps.
// ---------------------------------------------------------------------------------------

template <class F>
GSX_HD constexpr int ckpt_rows() { return F::D + (F::autonomous ? 0 : 1); }

// One checkpoint at `base` = &ck[slot][0][j] (rows n apart).
template <class F>
GSX_HD void ckpt_put(double* base, int64_t n, const double* y, double t) {
  GSX_UNROLL for (int d = 0; d < F::D; ++d) base[d * n] = y[d];
  if constexpr (!F::autonomous) base[F::D * n] = t;
  (void)t;
}

template <class F>
GSX_HD void ckpt_store(double* ck, int64_t slot, int64_t n, int64_t j, const double* y,
                       double t) {
  ckpt_put<F>(ck + slot * ckpt_rows<F>() * n + j, n, y, t);
}

// kRecCk = true is the recording launch (tape and checkpoints both given): the two flags
// become constants and the accept path carries no runtime tests of them.
template <class F, bool kRecCk = false>
GSX_HD void tsit5_fwd_lane(const FwdArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  using namespace tab;
  const int64_t n = a.n;
  double y[D], p[P], y5[D];
  double k[7][D];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  double t = a.t0;
  const double t1 = a.t1_per ? a.t1_per[j] : a.t1;
  double dt = (t1 - t) / 100.0;  // dt0 = (t1 - t0)/100 (gradsolve/warp/warp_replay.py:157)
  F::f(t, y, p, k[0], a.aux);
  int32_t acc = 0, rej = 0, status = kReached;
  const bool rec = kRecCk || a.tape != nullptr;
  const bool ck = kRecCk || a.ckpt != nullptr;
  // Tape rows available (the tape-full test is one 32-bit compare; no tape: unlimited).
  const int32_t cap = rec ? static_cast<int32_t>(py_min<int64_t>(a.tape_cap, INT32_MAX))
                          : INT32_MAX;
  int32_t ck_left = a.ckpt_interval;
  // Running write cursors: tape row acc and checkpoint slot acc / C of this lane.
  double* tp = rec ? a.tape + j : nullptr;
  double* cp = ck ? a.ckpt + j : nullptr;
  const int64_t cstride = static_cast<int64_t>(ckpt_rows<F>()) * n;
  if (ck) ckpt_put<F>(cp, n, y, t);
  for (int64_t it = 0; it < a.max_steps; ++it) {
    if (t >= t1) break;
    const double h = nn_min(dt, t1 - t);
    tsit5_stages<F, double>(t, y, h, p, k, y5, a.aux);
    F::f(t + h, y5, p, k[6], a.aux);
    double e[D];
    GSX_UNROLL for (int i = 0; i < D; ++i)
      e[i] = h * (e1 * k[0][i] + e2 * k[1][i] + e3 * k[2][i] + e4 * k[3][i] + e5 * k[4][i] +
                  e6 * k[5][i] + e7 * k[6][i]);
    const double err = tsit5_err<D>(e, y, y5, a.rtol, a.atol);
    if (err != err) {
      // A NaN error makes the next step size NaN; the reference loop then rejects every
      // remaining attempt (tsit5_replay.py:108-124) and ends with status 1 at the same state.
      rej += static_cast<int32_t>(a.max_steps - it);
      break;
    }
    if (err <= 1.0) {
      if (acc >= cap) {
        status = kTapeFull;
        break;
      }
      t = t + h;
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        y[d] = y5[d];
        k[0][d] = k[6][d];
      }
      if (rec) {
        *tp = h;
        tp += n;
      }
      ++acc;
      if (ck && --ck_left == 0) {
        cp += cstride;
        ckpt_put<F>(cp, n, y, t);
        ck_left = a.ckpt_interval;
      }
    } else {
      ++rej;
    }
    dt = tsit5_next_dt(h, err);
  }
  if (status == kReached && !(t >= t1)) status = kExhausted;
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * n + j] = y[d];
  // the counts and the status are optional outputs of a forward without a tape
  if (a.n_acc) a.n_acc[j] = acc;
  if (a.n_rej) a.n_rej[j] = rej;
  if (a.status) a.status[j] = status;
}

// ---------------------------------------------------------------------------------------
// Discrete adjoint of the frozen-h replay for one trajectory (the VJP of
// gradsolve/warp/warp_replay.py:30-55 / tsit5_step.py:80-102). Segments of C accepted steps
// are re-integrated forward from their checkpoint into `seg` (the states at the start of each
// step, stride `sstride` between consecutive doubles of one lane), then walked in reverse,
// recomputing each step's stages from its start state. k7 carries weight b7 = 0 and is not
// formed.
// ---------------------------------------------------------------------------------------

// One replay step without the error estimate and without k7 (y_next == y5).
template <class F>
GSX_HD void tsit5_advance(double t, double* y, double h, const double* p, const double* aux) {
  constexpr int D = F::D;
  double k[7][D];
  double y5[D];
  F::f(t, y, p, k[0], aux);
  tsit5_stages<F, double>(t, y, h, p, k, y5, aux);
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = y5[d];
}

// Stage inputs z1..z6 of the step (t, y, h) (z1 = y), and with ynext non-null the step's
// result y + h*(b1 k1 + ... + b6 k6), which equals tsit5_advance's.
template <class F>
GSX_HD void tsit5_stage_inputs(double t, const double* y, double h, const double* p,
                               double (&z)[6][F::D], double* ynext, const double* aux) {
  constexpr int D = F::D;
  using namespace tab;
  double k[6][D];
  GSX_UNROLL for (int i = 0; i < D; ++i) z[0][i] = y[i];
  F::f(t, z[0], p, k[0], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) z[1][i] = y[i] + h * (a21 * k[0][i]);
  F::f(t + c2 * h, z[1], p, k[1], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i) z[2][i] = y[i] + h * (a31 * k[0][i] + a32 * k[1][i]);
  F::f(t + c3 * h, z[2], p, k[2], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[3][i] = y[i] + h * (a41 * k[0][i] + a42 * k[1][i] + a43 * k[2][i]);
  F::f(t + c4 * h, z[3], p, k[3], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[4][i] = y[i] + h * (a51 * k[0][i] + a52 * k[1][i] + a53 * k[2][i] + a54 * k[3][i]);
  F::f(t + c5 * h, z[4], p, k[4], aux);
  GSX_UNROLL for (int i = 0; i < D; ++i)
    z[5][i] = y[i] + h * (a61 * k[0][i] + a62 * k[1][i] + a63 * k[2][i] + a64 * k[3][i] +
                          a65 * k[4][i]);
  if (ynext) {
    F::f(t + c6 * h, z[5], p, k[5], aux);
    GSX_UNROLL for (int i = 0; i < D; ++i)
      ynext[i] = y[i] + h * (b1 * k[0][i] + b2 * k[1][i] + b3 * k[2][i] + b4 * k[3][i] +
                             b5 * k[4][i] + b6 * k[5][i]);
  }
}

// The reverse half of the step VJP given the stage inputs z: lam <- (d y_next / d y)^T lam,
// pb += (d y_next / d p)^T lam.
template <class F>
GSX_HD void tsit5_step_vjp_z(double t, const double (&z)[6][F::D], double h, const double* p,
                             double* lam, double* pb, const double* aux) {
  constexpr int D = F::D;
  using namespace tab;
  // kbar_s = h*b_s*lam + sum_{r>s} a_rs * (h*zbar_r); zbar_s = J(z_s)^T kbar_s.
  constexpr double A[6][6] = {{0, 0, 0, 0, 0, 0},
                              {a21, 0, 0, 0, 0, 0},
                              {a31, a32, 0, 0, 0, 0},
                              {a41, a42, a43, 0, 0, 0},
                              {a51, a52, a53, a54, 0, 0},
                              {a61, a62, a63, a64, a65, 0}};
  constexpr double B[6] = {b1, b2, b3, b4, b5, b6};
  constexpr double Cn[6] = {0.0, c2, c3, c4, c5, c6};
  double hl[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) hl[i] = h * lam[i];
  double kb[6][D];
  GSX_UNROLL for (int s = 0; s < 6; ++s) {
    GSX_UNROLL for (int i = 0; i < D; ++i) kb[s][i] = B[s] * hl[i];
  }
  double lam_new[D];
  GSX_UNROLL for (int i = 0; i < D; ++i) lam_new[i] = lam[i];
  GSX_UNROLL for (int s = 5; s >= 0; --s) {
    double zb[D];
    GSX_UNROLL for (int i = 0; i < D; ++i) zb[i] = -0.0;  // x + (-0.0) == x: the first add folds
    F::vjp(t + Cn[s] * h, z[s], p, kb[s], zb, pb, aux);
    GSX_UNROLL for (int i = 0; i < D; ++i) lam_new[i] += zb[i];
    double hz[D];
    GSX_UNROLL for (int i = 0; i < D; ++i) hz[i] = h * zb[i];
    GSX_UNROLL for (int r = 0; r < s; ++r) {
      GSX_UNROLL for (int i = 0; i < D; ++i) kb[r][i] += A[s][r] * hz[i];
    }
  }
  GSX_UNROLL for (int i = 0; i < D; ++i) lam[i] = lam_new[i];
}

// lam <- (d y_next / d y)^T lam ; pb += (d y_next / d p)^T lam, at the step (t, y, h): the
// stage inputs are recomputed from y, then reversed.
template <class F>
GSX_HD void tsit5_step_vjp(double t, const double* y, double h, const double* p, double* lam,
                           double* pb, const double* aux) {
  double z[6][F::D];
  tsit5_stage_inputs<F>(t, y, h, p, z, nullptr, aux);
  tsit5_step_vjp_z<F>(t, z, h, p, lam, pb, aux);
}

// Frozen-mesh replay (the forward of gradsolve/warp/warp_replay.py:30-55) from possibly new
// (y0, p) over a recorded tape, writing the checkpoints the adjoint needs.

template <class F>
GSX_HD void tsit5_replay_lane(const ReplayArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  double y[D], p[P];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  double t = a.t0;
  const int32_t na = a.n_acc[j];
  const bool ck = a.ckpt != nullptr;
  if (ck) ckpt_store<F>(a.ckpt, 0, n, j, y, t);
  int32_t ck_left = a.ckpt_interval;
  for (int32_t s = 0; s < na; ++s) {
    const double h = a.tape[static_cast<int64_t>(s) * n + j];
    tsit5_advance<F>(t, y, h, p, a.aux);
    t = t + h;
    if (ck && --ck_left == 0) {
      ckpt_store<F>(a.ckpt, (s + 1) / a.ckpt_interval, n, j, y, t);
      ck_left = a.ckpt_interval;
    }
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y_final[d * n + j] = y[d];
}

// The one-step method a dense-output lane replays and re-steps with: M::advance<F> is the
// value-only step (y <- y + step), M::step_vjp<F> its VJP. Tsit5 here; Rosenbrock23's policy
// sits beside its step in gsx_rb23.h.
struct MethodTsit5 {
  template <class F>
  GSX_HD static void advance(double t, double* y, double h, const double* p, const double* aux) {
    tsit5_advance<F>(t, y, h, p, aux);
  }
  template <class F>
  GSX_HD static void step_vjp(double t, const double* y, double h, const double* p,
                              double* lam, double* pb, const double* aux) {
    tsit5_step_vjp<F>(t, y, h, p, lam, pb, aux);
  }
};

// Dense output for one trajectory (gradsolve/solvers/dense.py:106-163, restated as the C
// oracle's replay_saveat): replays the recorded mesh; a save time in (t_s, t_s + h_s] is the
// state one step of size ts[q] - t_s from (t_s, y_s), except that at or after the final time
// it is y_final itself, and at or before t0 it is y0. save_step / save_t / states record the
// bracket for the gradient (saveat_vjp_lane).
template <class F, class M>
GSX_HD void saveat_lane(const SaveArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n, k = a.nsave;
  double y[D], p[P];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y0[d * n + j];
  GSX_UNROLL for (int q = 0; q < P; ++q) p[q] = a.p[q * n + j];
  auto put = [&](int64_t q, const double* v, int32_t st, double tq, const double* ys0) {
    GSX_UNROLL for (int d = 0; d < D; ++d) a.ys[(q * D + d) * n + j] = v[d];
    a.save_step[q * n + j] = st;
    a.save_t[q * n + j] = tq;
    if (a.states) {
      GSX_UNROLL for (int d = 0; d < D; ++d) a.states[(q * D + d) * n + j] = ys0[d];
    }
  };
  double t = a.t0;
  int64_t q = 0;
  while (q < k && a.ts[q] <= a.t0) {
    put(q, y, -1, t, y);
    ++q;
  }
  const int32_t na = a.n_acc[j];
  for (int32_t s = 0; s < na; ++s) {
    const double h = a.tape[static_cast<int64_t>(s) * n + j];
    const double tn = t + h;
    while (q < k && a.ts[q] <= tn) {
      if (s == na - 1 && a.ts[q] >= tn) break;  // at the final time: y_final, below
      double yq[D];
      GSX_UNROLL for (int d = 0; d < D; ++d) yq[d] = y[d];
      M::template advance<F>(t, yq, a.ts[q] - t, p, a.aux);
      put(q, yq, s, t, y);
      ++q;
    }
    M::template advance<F>(t, y, h, p, a.aux);
    t = tn;
  }
  if (a.y_final) {
    GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.y_final[d * n + j];
  }
  for (; q < k; ++q) put(q, y, na, t, y);
}

// The save cotangents' part of the gradient for one trajectory: mu[q] = (d ys[q] / d y_s)^T
// ysbar[q] for a bracketed save (the VJP of its re-step, which also adds to the parameter
// gradient), ysbar[q] itself for a save at y0, and the final-time saves folded into the
// final-state cotangent. The adjoint sweep (AdjArgs with nsave) then adds mu at each bracket.
template <class F, class M>
GSX_HD void saveat_vjp_lane(const SaveVjpArgs& a, int64_t j) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n, k = a.nsave;
  double p[P], pb[P], lf[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = -0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lf[d] = a.ybar ? a.ybar[d * n + j] : 0.0;
  const int32_t na = a.n_acc[j];
  for (int64_t q = 0; q < k; ++q) {
    const int32_t st = a.save_step[q * n + j];
    double l[D];
    GSX_UNROLL for (int d = 0; d < D; ++d) l[d] = a.ysbar[(q * D + d) * n + j];
    if (st >= na) {
      GSX_UNROLL for (int d = 0; d < D; ++d) {
        lf[d] += l[d];
        l[d] = 0.0;
      }
    } else if (st >= 0) {
      double y[D];
      GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = a.states[(q * D + d) * n + j];
      const double ts0 = a.save_t[q * n + j];
      M::template step_vjp<F>(ts0, y, a.ts[q] - ts0, p, l, pb, a.aux);
    }
    GSX_UNROLL for (int d = 0; d < D; ++d) a.mu[(q * D + d) * n + j] = l[d];
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) a.ybar_eff[d * n + j] = lf[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q] + 0.0;
}

template <class F>
GSX_HD void tsit5_saveat_lane(const SaveArgs& a, int64_t j) {
  saveat_lane<F, MethodTsit5>(a, j);
}

template <class F>
GSX_HD void tsit5_saveat_vjp_lane(const SaveVjpArgs& a, int64_t j) {
  saveat_vjp_lane<F, MethodTsit5>(a, j);
}

// Adds the save cotangents mu[q] of the saves bracketed by step s (a cotangent on y_s) to lam,
// once the reverse sweep has turned lam into dL/dy_s; s = -1 takes the saves at or before t0.
// `q` walks the ascending saves downwards; final-time saves (save_step == n_acc) are skipped,
// their cotangent is already in ybar.
template <int D>
GSX_HD void inject_saves(const AdjArgs& a, int64_t j, int32_t s, int32_t& q, double* lam) {
  if (a.nsave == 0) return;
  const int64_t n = a.n;
  while (q >= 0) {
    const int32_t st = a.save_step[static_cast<int64_t>(q) * n + j];
    if (st < s) break;
    if (st == s) {
      GSX_UNROLL for (int d = 0; d < D; ++d)
        lam[d] += a.mu[(static_cast<int64_t>(q) * D + d) * n + j];
    }
    --q;
  }
}

// Segment count of lane j: checkpoint m starts accepted step m * C.
GSX_HD int32_t segment_count(const AdjArgs& a, int64_t j) {
  return (a.n_acc[j] + a.ckpt_interval - 1) / a.ckpt_interval;
}

// The segment adjoints walk from segment max(mtop, own count) - 1 down and skip the segments
// above their own count: with mtop the wavefront's largest count (wave_max) the lanes run in
// step; mtop = 0 walks the lane's own segments only (the CPU harness).
template <class F>
GSX_HD void tsit5_adj_lane(const AdjArgs& a, int64_t j, double* seg, int64_t sstride,
                           int32_t mtop = 0) {
  constexpr int D = F::D, P = F::P;
  constexpr int DCK = ckpt_rows<F>();
  const int64_t n = a.n;
  double p[P], pb[P], lam[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = a.pbar_in ? a.pbar_in[q * n + j] : 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = a.ybar[d * n + j];
  int32_t qs = static_cast<int32_t>(a.nsave) - 1;  // save cursor, walked downwards
  const int32_t na = a.n_acc[j];
  const int32_t C = a.ckpt_interval;
  const int32_t nseg = (na + C - 1) / C;
  for (int32_t m = (mtop > nseg ? mtop : nseg) - 1; m >= 0; --m) {
    if (m >= nseg) continue;
    const int32_t s0 = m * C;
    const int32_t s1 = (s0 + C < na) ? s0 + C : na;
    double y[D];
    double t;
    const double* base = a.ckpt + static_cast<int64_t>(m) * DCK * n + j;
    GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = base[d * n];
    if constexpr (!F::autonomous) {
      t = base[D * n];
    } else {
      t = 0.0;
    }
    // Re-integrate the segment, keeping the state at the start of each step.
    for (int32_t s = s0; s < s1; ++s) {
      double* slot = seg + static_cast<int64_t>(s - s0) * DCK * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d) slot[d * sstride] = y[d];
      if constexpr (!F::autonomous) slot[D * sstride] = t;
      if (s + 1 < s1) {
        const double h = a.tape[static_cast<int64_t>(s) * n + j];
        tsit5_advance<F>(t, y, h, p, a.aux);
        t = t + h;
      }
    }
    // Walk the segment in reverse.
    for (int32_t s = s1 - 1; s >= s0; --s) {
      const double* slot = seg + static_cast<int64_t>(s - s0) * DCK * sstride;
      GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = slot[d * sstride];
      if constexpr (!F::autonomous) t = slot[D * sstride];
      const double h = a.tape[static_cast<int64_t>(s) * n + j];
      tsit5_step_vjp<F>(t, y, h, p, lam, pb, a.aux);
      inject_saves<D>(a, j, s, qs, lam);
    }
  }
  inject_saves<D>(a, j, -1, qs, lam);
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y0bar[d * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

// The adjoint with checkpoints every 2 accepted steps and no re-integration overhead: for
// segment m (steps 2m, 2m+1) one advance from checkpoint m forms step 2m's stage inputs and the
// state y_{2m+1}; step 2m+1's stage inputs are formed for its own VJP, and step 2m's VJP reuses
// the ones kept from the advance, so each step's stages are computed once, as the minimal
// adjoint needs. A missing second step (odd n_acc) runs with h = 0, which leaves lam and pb
// unchanged exactly, so every lane follows the same control flow. `m` is wavefront-uniform in
// the kernel: all active lanes read checkpoint row m and tape rows 2m, 2m+1 together.
// kSaves = false drops the save-cotangent injection and pbar_in (the plain final-state
// gradient), which keeps the lane's register footprint at that of the bare adjoint.
template <class F, bool kSaves = true>
GSX_HD void tsit5_adj_pair_segment(const AdjArgs& a, int64_t j, int32_t m, int32_t na,
                                   const double* p, double* lam, double* pb, int32_t& qs) {
  constexpr int D = F::D;
  constexpr int DCK = ckpt_rows<F>();
  const int64_t n = a.n;
  const int32_t s0 = 2 * m;
  const double* base = a.ckpt + static_cast<int64_t>(m) * DCK * n + j;
  double y[D];
  GSX_UNROLL for (int d = 0; d < D; ++d) y[d] = base[d * n];
  double t = 0.0;
  if constexpr (!F::autonomous) t = base[D * n];
  const double h0 = a.tape[static_cast<int64_t>(s0) * n + j];
  double h1 = 0.0;
  if (s0 + 1 < na) h1 = a.tape[static_cast<int64_t>(s0 + 1) * n + j];
  double z0[6][D], y1[D];
  tsit5_stage_inputs<F>(t, y, h0, p, z0, y1, a.aux);
  {
    double z1[6][D];
    tsit5_stage_inputs<F>(t + h0, y1, h1, p, z1, nullptr, a.aux);
    tsit5_step_vjp_z<F>(t + h0, z1, h1, p, lam, pb, a.aux);
  }
  if constexpr (kSaves) {
    if (s0 + 1 < na) inject_saves<D>(a, j, s0 + 1, qs, lam);
  }
  tsit5_step_vjp_z<F>(t, z0, h0, p, lam, pb, a.aux);
  if constexpr (kSaves) inject_saves<D>(a, j, s0, qs, lam);
}

// The pair adjoint of one lane. `mmax` is at least the lane's segment count; the kernel passes
// the wavefront's largest, so every lane runs the loop in step and idles above its own count.
template <class F, bool kSaves = true>
GSX_HD void tsit5_adj_pair_lane(const AdjArgs& a, int64_t j, int32_t na, int32_t mmax) {
  constexpr int D = F::D, P = F::P;
  const int64_t n = a.n;
  double p[P], pb[P], lam[D];
  GSX_UNROLL for (int q = 0; q < P; ++q) {
    p[q] = a.p[q * n + j];
    pb[q] = (kSaves && a.pbar_in) ? a.pbar_in[q * n + j] : 0.0;
  }
  GSX_UNROLL for (int d = 0; d < D; ++d) lam[d] = a.ybar[d * n + j];
  int32_t qs = kSaves ? static_cast<int32_t>(a.nsave) - 1 : -1;
  const int32_t nseg = (na + 1) / 2;
  for (int32_t m = mmax - 1; m >= 0; --m) {
    if (m < nseg) tsit5_adj_pair_segment<F, kSaves>(a, j, m, na, p, lam, pb, qs);
  }
  if constexpr (kSaves) inject_saves<D>(a, j, -1, qs, lam);
  GSX_UNROLL for (int d = 0; d < D; ++d) a.y0bar[d * n + j] = lam[d];
  GSX_UNROLL for (int q = 0; q < P; ++q) a.pbar[q * n + j] = pb[q];
}

}  // namespace GSX_NS
