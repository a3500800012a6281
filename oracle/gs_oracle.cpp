// gs_oracle.cpp -- CPU oracle for the ensemble ODE record-and-replay path.
//
// TEST INFRASTRUCTURE ONLY: the checker that the HIP product path is compared with.
// Only tests/, the smoke check and bench.py's CPU-baseline leg may load it.
//
// Compiled with -ffp-contract=off and without -ffast-math, so every expression rounds
// exactly as the reference's NumPy loop does: each binary operation rounds once, in the
// textual order of the reference source (gradsolve/solvers/tsit5_replay.py:100-124,
// gradsolve/solvers/tsit5_step.py:80-102). pow and sqrt are the C library's, as NumPy's
// scalar ``**`` and ``np.sqrt`` are.
//
// The core is templated on the scalar type so that the same source, instantiated on an
// op-counting scalar, measures the algorithmic flop count of each step (gso_count_ops).
#include "gs_oracle.h"
#include "../gradsolve_amd/csrc/gsx_rodas5p_tab.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

constexpr int kMaxD = 64;

// ---------------------------------------------------------------------------------------
// Tsitouras (2011) 5(4) pair, transcribed as gradsolve/solvers/tsit5_step.py:31-77 holds it
// (the same decimal strings and the same double-precision expressions, e.g. e_i = b_i - bhat_i
// evaluated in binary64).
// ---------------------------------------------------------------------------------------
constexpr double C2 = 161.0 / 1000.0;
constexpr double C3 = 327.0 / 1000.0;
constexpr double C4 = 9.0 / 10.0;
constexpr double C5 = 0.9800255409045096857298102862870245954942137979563024768854764293221195950761080302604;
constexpr double C6 = 1.0;
constexpr double A21 = 161.0 / 1000.0;
constexpr double A31 = -0.8480655492356988544426874250230774675121177393430391537369234245294192976164141156943e-2;
constexpr double A32 = 0.3354806554923569885444268742502307746751211773934303915373692342452941929761641411569;
constexpr double A41 = 2.897153057105493432130432594192938764924887287701866490314866693455023795137503079289;
constexpr double A42 = -6.359448489975074843148159912383825625952700647415626703305928850207288721235210244366;
constexpr double A43 = 4.362295432869581411017727318190886861027813359713760212991062156752264926097707165077;
constexpr double A51 = 5.325864828439256604428877920840511317836476253097040101202360397727981648835607691791;
constexpr double A52 = -11.74888356406282787774717033978577296188744178259862899288666928009020615663593781589;
constexpr double A53 = 7.495539342889836208304604784564358155658679161518186721010132816213648793440552049753;
constexpr double A54 = -0.9249506636175524925650207933207191611349983406029535244034750452930469056411389539635e-1;
constexpr double A61 = 5.861455442946420028659251486982647890394337666164814434818157239052507339770711679748;
constexpr double A62 = -12.92096931784710929170611868178335939541780751955743459166312250439928519268343184452;
constexpr double A63 = 8.159367898576158643180400794539253485181918321135053305748355423955009222648673734986;
constexpr double A64 = -0.7158497328140099722453054252582973869127213147363544882721139659546372402303777878835e-1;
constexpr double A65 = -0.2826905039406838290900305721271224146717633626879770007617876201276764571291579142206e-1;
constexpr double B1 = 0.9646076681806522951816731316512876333711995238157997181903319145764851595234062815396e-1;
constexpr double B2 = 1.0 / 100.0;
constexpr double B3 = 0.4798896504144995747752495322905965199130404621990332488332634944254542060153074523509;
constexpr double B4 = 1.379008574103741893192274821856872770756462643091360525934940067397245698027561293331;
constexpr double B5 = -3.290069515436080679901047585711363850115683290894936158531296799594813811049925401677;
constexpr double B6 = 2.324710524099773982415355918398765796109060233222962411944060046314465391054716027841;
constexpr double B7 = 0.0;
constexpr double E1 = 0.9646076681806522951816731316512876333711995238157997181903319145764851595234062815396e-1 - 0.9468075576583945807478876255758922856117527357724631226139574065785592789071067303271e-1;
constexpr double E2 = 1.0 / 100.0 - 0.9183565540343253096776363936645313759813746240984095238905939532922955247253608687270e-2;
constexpr double E3 = 0.4798896504144995747752495322905965199130404621990332488332634944254542060153074523509 - 0.4877705284247615707855642599631228241516691959761363774365216240304071651579571959813;
constexpr double E4 = 1.379008574103741893192274821856872770756462643091360525934940067397245698027561293331 - 1.234297566930478985655109673884237654035539930748192848315425833500484878378061439761;
constexpr double E5 = -3.290069515436080679901047585711363850115683290894936158531296799594813811049925401677 + 2.707712349983525454881109975059321670689605166938197378763992255714444407154902012702;
constexpr double E6 = 2.324710524099773982415355918398765796109060233222962411944060046314465391054716027841 - 1.866628418170587035753719399566211498666255505244122593996591602841258328965767580089;
constexpr double E7 = -1.0 / 66.0;

// Controller constants (tsit5_step.py:26, method.py:35; Rosenbrock: _warp_rosenbrock.py:61-70).
constexpr double SAFETY = 0.9, ERR_FLOOR = 1e-10, FAC_MIN = 0.2, FAC_MAX = 5.0;
constexpr double TSIT5_EXP = -0.2;
constexpr double RB23_EXP = -1.0 / 3.0;

// ---------------------------------------------------------------------------------------
// Op-counting scalar. Every arithmetic operation and sqrt adds one flop; pow and the
// trigonometric functions are counted separately. abs / max / comparisons are free.
// ---------------------------------------------------------------------------------------
thread_local double g_flops = 0.0, g_pows = 0.0;
struct OpD {
  double v;
  OpD() : v(0.0) {}
  OpD(double x) : v(x) {}  // NOLINT: implicit from literals, like double
};
inline OpD operator+(OpD a, OpD b) { g_flops += 1; return OpD(a.v + b.v); }
inline OpD operator-(OpD a, OpD b) { g_flops += 1; return OpD(a.v - b.v); }
inline OpD operator*(OpD a, OpD b) { g_flops += 1; return OpD(a.v * b.v); }
inline OpD operator/(OpD a, OpD b) { g_flops += 1; return OpD(a.v / b.v); }
inline OpD operator-(OpD a) { return OpD(-a.v); }
inline bool operator<=(OpD a, OpD b) { return a.v <= b.v; }
inline bool operator<(OpD a, OpD b) { return a.v < b.v; }
inline bool operator>=(OpD a, OpD b) { return a.v >= b.v; }
inline bool operator>(OpD a, OpD b) { return a.v > b.v; }
inline OpD sqrt_(OpD a) { g_flops += 1; return OpD(std::sqrt(a.v)); }
inline OpD pow_(OpD a, OpD b) { g_pows += 1; return OpD(std::pow(a.v, b.v)); }
inline OpD abs_(OpD a) { return OpD(std::fabs(a.v)); }
inline OpD max_(OpD a, OpD b) { return a.v >= b.v ? a : b; }
inline OpD min_(OpD a, OpD b) { return a.v <= b.v ? a : b; }
inline OpD sin_(OpD a) { g_pows += 1; return OpD(std::sin(a.v)); }
inline OpD cos_(OpD a) { g_pows += 1; return OpD(std::cos(a.v)); }
inline double val(OpD a) { return a.v; }

inline double sqrt_(double a) { return std::sqrt(a); }
inline double pow_(double a, double b) { return std::pow(a, b); }
inline double abs_(double a) { return std::fabs(a); }
// Python's max(a, b) returns a unless b > a; min(a, b) returns a unless b < a. np.maximum
// returns the larger (NaN-propagating); all operands here are finite or the trajectory is
// already failing.
inline double max_(double a, double b) { return b > a ? b : a; }
inline double min_(double a, double b) { return b < a ? b : a; }
inline double sin_(double a) { return std::sin(a); }
inline double cos_(double a) { return std::cos(a); }
inline double val(double a) { return a; }

// ---------------------------------------------------------------------------------------
// Fields. Each textual expression follows the registered field's association.
// ---------------------------------------------------------------------------------------
struct FieldSpec {
  int id, D, P;
  const double* aux;
};

bool field_ok(const FieldSpec& fs) {
  switch (fs.id) {
    case GSO_F_LORENZ: return fs.D == 3 && fs.P >= 1;
    case GSO_F_LORENZ3: return fs.D == 3 && fs.P >= 3;
    case GSO_F_LINEAR_DECAY: return fs.D >= 1 && fs.D <= kMaxD && fs.P >= 1;
    case GSO_F_ROBERTSON: return fs.D == 3 && fs.P >= 3;
    case GSO_F_KURAMOTO: return fs.D >= 1 && fs.D <= kMaxD && fs.P >= fs.D && fs.aux != nullptr;
    case GSO_F_VDP: return fs.D == 2 && fs.P >= 1;
    case GSO_F_LINEAR_LADDER: return fs.D >= 1 && fs.D <= kMaxD;
    case GSO_F_FORCED_STIFF: return fs.D == 2 && fs.P >= 3;
    default: return false;
  }
}

// Fields whose right-hand side reads t (the Rodas5P stages then carry the df/dt term).
bool field_time_dependent(const FieldSpec& fs) { return fs.id == GSO_F_FORCED_STIFF; }

// Fields with an analytic Jacobian (field_jac), which the Rosenbrock methods need.
bool field_has_jac(const FieldSpec& fs) {
  return fs.id == GSO_F_ROBERTSON || fs.id == GSO_F_LINEAR_DECAY ||
         fs.id == GSO_F_LINEAR_LADDER || fs.id == GSO_F_FORCED_STIFF;
}

template <class T>
void field_eval(const FieldSpec& fs, T t, const T* y, const double* p, T* dy) {
  (void)t;
  switch (fs.id) {
    case GSO_F_LORENZ: {
      // _warp_kernel.py:120-128: [sigma*(y1-y0), rho*y0 - y1 - y0*y2, y0*y1 - beta*y2]
      const T sigma = 10.0, beta = T(8.0 / 3.0), rho = p[0];
      dy[0] = sigma * (y[1] - y[0]);
      dy[1] = rho * y[0] - y[1] - y[0] * y[2];
      dy[2] = y[0] * y[1] - beta * y[2];
      break;
    }
    case GSO_F_LORENZ3: {
      const T sigma = p[0], beta = p[1], rho = p[2];
      dy[0] = sigma * (y[1] - y[0]);
      dy[1] = rho * y[0] - y[1] - y[0] * y[2];
      dy[2] = y[0] * y[1] - beta * y[2];
      break;
    }
    case GSO_F_LINEAR_DECAY: {
      const T k = p[0];
      for (int i = 0; i < fs.D; ++i) dy[i] = -k * y[i];
      break;
    }
    case GSO_F_ROBERTSON: {
      // _warp_rosenbrock.py:121-130
      const T k1 = p[0], k2 = p[1], k3 = p[2];
      dy[0] = -k1 * y[0] + k3 * y[1] * y[2];
      dy[1] = k1 * y[0] - k3 * y[1] * y[2] - k2 * y[1] * y[1];
      dy[2] = k2 * y[1] * y[1];
      break;
    }
    case GSO_F_KURAMOTO: {
      // f_i = omega_i + (1/D) * ((K s)_i * c_i - (K c)_i * s_i), s = sin(theta), c = cos(theta)
      const int D = fs.D;
      T s[kMaxD], c[kMaxD];
      for (int j = 0; j < D; ++j) { s[j] = sin_(y[j]); c[j] = cos_(y[j]); }
      const T invD = T(1.0 / D);
      for (int i = 0; i < D; ++i) {
        T a = 0.0, b = 0.0;
        for (int j = 0; j < D; ++j) {
          const T kij = fs.aux[i * D + j];
          a = a + kij * s[j];
          b = b + kij * c[j];
        }
        dy[i] = T(p[i]) + invD * (a * c[i] - b * s[i]);
      }
      break;
    }
    case GSO_F_VDP: {
      // _warp_kernel.py:167-172: dx = v; dv = mu*((1 - x*x)*v - x)
      const T mu = p[0];
      const T x = y[0], v = y[1];
      dy[0] = v;
      dy[1] = mu * ((T(1.0) - x * x) * v - x);
      break;
    }
    case GSO_F_LINEAR_LADDER: {
      for (int i = 0; i < fs.D; ++i) dy[i] = T(1.01) * y[i];
      break;
    }
    case GSO_F_FORCED_STIFF: {
      const T a = p[0], b = p[1], c = p[2];
      const T st = sin_(t);
      dy[0] = -a * (y[0] - y[1] * y[1]) + T(0.4) * sin_(T(3.0) * t);
      dy[1] = y[0] - y[1] - b * y[1] * y[1] * y[1] + T(0.2) * cos_(t) + c * st * y[0];
      break;
    }
    default: break;
  }
}

// Vector-Jacobian product: ybar += (df/dy)^T v ; pbar += (df/dp)^T v. Derived by hand from
// the field expressions above.
template <class T>
void field_vjp(const FieldSpec& fs, T t, const T* y, const double* p, const T* v, T* ybar,
               T* pbar) {
  (void)t;
  switch (fs.id) {
    case GSO_F_LORENZ: {
      const T sigma = 10.0, beta = T(8.0 / 3.0), rho = p[0];
      // J = [[-s, s, 0], [rho - y2, -1, -y0], [y1, y0, -beta]]
      ybar[0] = ybar[0] + (-sigma * v[0] + (rho - y[2]) * v[1] + y[1] * v[2]);
      ybar[1] = ybar[1] + (sigma * v[0] - v[1] + y[0] * v[2]);
      ybar[2] = ybar[2] + (-y[0] * v[1] - beta * v[2]);
      pbar[0] = pbar[0] + y[0] * v[1];
      break;
    }
    case GSO_F_LORENZ3: {
      const T sigma = p[0], beta = p[1], rho = p[2];
      ybar[0] = ybar[0] + (-sigma * v[0] + (rho - y[2]) * v[1] + y[1] * v[2]);
      ybar[1] = ybar[1] + (sigma * v[0] - v[1] + y[0] * v[2]);
      ybar[2] = ybar[2] + (-y[0] * v[1] - beta * v[2]);
      pbar[0] = pbar[0] + (y[1] - y[0]) * v[0];
      pbar[1] = pbar[1] - y[2] * v[2];
      pbar[2] = pbar[2] + y[0] * v[1];
      break;
    }
    case GSO_F_LINEAR_DECAY: {
      const T k = p[0];
      T acc = 0.0;
      for (int i = 0; i < fs.D; ++i) {
        ybar[i] = ybar[i] - k * v[i];
        acc = acc + y[i] * v[i];
      }
      pbar[0] = pbar[0] - acc;
      break;
    }
    case GSO_F_ROBERTSON: {
      const T k1 = p[0], k2 = p[1], k3 = p[2];
      // J = [[-k1, k3 y2, k3 y1], [k1, -k3 y2 - 2 k2 y1, -k3 y1], [0, 2 k2 y1, 0]]
      ybar[0] = ybar[0] + (-k1 * v[0] + k1 * v[1]);
      ybar[1] = ybar[1] + (k3 * y[2] * v[0] - (k3 * y[2] + T(2.0) * k2 * y[1]) * v[1] +
                           T(2.0) * k2 * y[1] * v[2]);
      ybar[2] = ybar[2] + (k3 * y[1] * v[0] - k3 * y[1] * v[1]);
      // df/dk1 = [-y0, y0, 0]; df/dk2 = [0, -y1^2, y1^2]; df/dk3 = [y1 y2, -y1 y2, 0]
      pbar[0] = pbar[0] + y[0] * (v[1] - v[0]);
      pbar[1] = pbar[1] + y[1] * y[1] * (v[2] - v[1]);
      pbar[2] = pbar[2] + y[1] * y[2] * (v[0] - v[1]);
      break;
    }
    case GSO_F_KURAMOTO: {
      const int D = fs.D;
      T s[kMaxD], c[kMaxD], a[kMaxD], b[kMaxD];
      for (int j = 0; j < D; ++j) { s[j] = sin_(y[j]); c[j] = cos_(y[j]); }
      for (int i = 0; i < D; ++i) {
        T ai = 0.0, bi = 0.0;
        for (int j = 0; j < D; ++j) {
          ai = ai + T(fs.aux[i * D + j]) * s[j];
          bi = bi + T(fs.aux[i * D + j]) * c[j];
        }
        a[i] = ai;
        b[i] = bi;
      }
      const T invD = T(1.0 / D);
      // (J^T v)_m = (1/D) [ c_m sum_i K_im v_i c_i + s_m sum_i K_im v_i s_i - v_m (a_m s_m + b_m c_m) ]
      for (int m = 0; m < D; ++m) {
        T u = 0.0, w = 0.0;
        for (int i = 0; i < D; ++i) {
          const T kim = fs.aux[i * D + m];
          u = u + kim * (v[i] * c[i]);
          w = w + kim * (v[i] * s[i]);
        }
        ybar[m] = ybar[m] + invD * (c[m] * u + s[m] * w - v[m] * (a[m] * s[m] + b[m] * c[m]));
        pbar[m] = pbar[m] + v[m];
      }
      break;
    }
    case GSO_F_VDP: {
      const T mu = p[0];
      const T x = y[0], w = y[1];
      // J = [[0, 1], [mu*(-2 x w - 1), mu*(1 - x^2)]]
      ybar[0] = ybar[0] + mu * (T(-2.0) * x * w - T(1.0)) * v[1];
      ybar[1] = ybar[1] + v[0] + mu * (T(1.0) - x * x) * v[1];
      pbar[0] = pbar[0] + ((T(1.0) - x * x) * w - x) * v[1];
      break;
    }
    case GSO_F_LINEAR_LADDER: {
      for (int i = 0; i < fs.D; ++i) ybar[i] = ybar[i] + T(1.01) * v[i];
      break;
    }
    case GSO_F_FORCED_STIFF: {
      const T a = p[0], b = p[1];
      const T st = sin_(t);
      // J = [[-a, 2 a y1], [1 + c sin t, -1 - 3 b y1^2]]
      ybar[0] = ybar[0] + (-a * v[0] + (T(1.0) + T(p[2]) * st) * v[1]);
      ybar[1] = ybar[1] + (T(2.0) * a * y[1] * v[0] - (T(1.0) + T(3.0) * b * y[1] * y[1]) * v[1]);
      pbar[0] = pbar[0] - (y[0] - y[1] * y[1]) * v[0];
      pbar[1] = pbar[1] - y[1] * y[1] * y[1] * v[1];
      pbar[2] = pbar[2] + st * y[0] * v[1];
      break;
    }
    default: break;
  }
}

// Analytic Jacobian df/dy (row-major DxD) for the stiff fields (Rosenbrock23).
// Robertson: _warp_rosenbrock.py:132-149 / gradsolve_g2_spike_oracle.py:53-69.
template <class T>
bool field_jac(const FieldSpec& fs, T t, const T* y, const double* p, T* J) {
  const int D = fs.D;
  switch (fs.id) {
    case GSO_F_ROBERTSON: {
      const T k1 = p[0], k2 = p[1], k3 = p[2];
      const T y1 = y[1], y2 = y[2];
      J[0] = -k1;      J[1] = k3 * y2;                      J[2] = k3 * y1;
      J[3] = k1;       J[4] = -k3 * y2 - T(2.0) * k2 * y1;  J[5] = -k3 * y1;
      J[6] = 0.0;      J[7] = T(2.0) * k2 * y1;             J[8] = 0.0;
      return true;
    }
    case GSO_F_LINEAR_DECAY: {
      for (int i = 0; i < D * D; ++i) J[i] = 0.0;
      for (int i = 0; i < D; ++i) J[i * D + i] = -T(p[0]);
      return true;
    }
    case GSO_F_LINEAR_LADDER: {
      for (int i = 0; i < D * D; ++i) J[i] = 0.0;
      for (int i = 0; i < D; ++i) J[i * D + i] = 1.01;
      return true;
    }
    case GSO_F_FORCED_STIFF: {
      const T a = p[0], b = p[1], c = p[2];
      const T y1 = y[1];
      J[0] = -a;
      J[1] = T(2.0) * a * y1;
      J[2] = T(1.0) + c * sin_(t);
      J[3] = -T(1.0) - T(3.0) * b * y1 * y1;
      return true;
    }
    default: return false;
  }
}

// ybar += d/dy [ sum_ab Jbar_ab J_ab(y, p) ], pbar += d/dp [ same ] (second-derivative terms
// of the Rosenbrock adjoint, which differentiates through W = I - d h J(y, p)).
template <class T>
void field_jac_vjp(const FieldSpec& fs, T t, const T* y, const double* p, const T* Jb, T* ybar,
                   T* pbar) {
  switch (fs.id) {
    case GSO_F_FORCED_STIFF: {
      const T a = p[0], b = p[1];
      const T y1 = y[1];
      // dJ/dy1: J01 = 2 a, J11 = -6 b y1 ; dJ/da: J00 = -1, J01 = 2 y1 ; dJ/db: J11 = -3 y1^2 ;
      // dJ/dc: J10 = sin t
      ybar[1] = ybar[1] + (T(2.0) * a * Jb[1] - T(6.0) * b * y1 * Jb[3]);
      pbar[0] = pbar[0] + (T(2.0) * y1 * Jb[1] - Jb[0]);
      pbar[1] = pbar[1] - T(3.0) * y1 * y1 * Jb[3];
      pbar[2] = pbar[2] + sin_(t) * Jb[2];
      break;
    }
    case GSO_F_ROBERTSON: {
      const T k2 = p[1], k3 = p[2];
      const T y1 = y[1], y2 = y[2];
      // dJ/dy1: J02 = k3, J11 = -2 k2, J12 = -k3, J21 = 2 k2 ; dJ/dy2: J01 = k3, J11 = -k3
      ybar[1] = ybar[1] + (k3 * Jb[2] - T(2.0) * k2 * Jb[4] - k3 * Jb[5] + T(2.0) * k2 * Jb[7]);
      ybar[2] = ybar[2] + (k3 * Jb[1] - k3 * Jb[4]);
      // dJ/dk1: J00 = -1, J10 = 1 ; dJ/dk2: J11 = -2 y1, J21 = 2 y1 ;
      // dJ/dk3: J01 = y2, J02 = y1, J11 = -y2, J12 = -y1
      pbar[0] = pbar[0] + (Jb[3] - Jb[0]);
      pbar[1] = pbar[1] + T(2.0) * y1 * (Jb[7] - Jb[4]);
      pbar[2] = pbar[2] + (y2 * Jb[1] + y1 * Jb[2] - y2 * Jb[4] - y1 * Jb[5]);
      break;
    }
    case GSO_F_LINEAR_DECAY: {
      T tr = 0.0;
      for (int i = 0; i < fs.D; ++i) tr = tr + Jb[i * fs.D + i];
      pbar[0] = pbar[0] - tr;
      break;
    }
    default: break;
  }
}

// df/dt at (t, y) for the time-dependent fields (the Rodas5P stages' time-derivative term).
template <class T>
void field_dfdt(const FieldSpec& fs, T t, const T* y, const double* p, T* ft) {
  switch (fs.id) {
    case GSO_F_FORCED_STIFF: {
      const T c = p[2];
      ft[0] = T(1.2) * cos_(T(3.0) * t);
      ft[1] = -T(0.2) * sin_(t) + c * cos_(t) * y[0];
      break;
    }
    default:
      for (int i = 0; i < fs.D; ++i) ft[i] = T(0.0);
      break;
  }
}

// ybar += (d(df/dt)/dy)^T ftb, pbar += (d(df/dt)/dp)^T ftb.
template <class T>
void field_dfdt_vjp(const FieldSpec& fs, T t, const T* y, const double* p, const T* ftb,
                    T* ybar, T* pbar) {
  switch (fs.id) {
    case GSO_F_FORCED_STIFF: {
      const T ct = cos_(t);
      ybar[0] = ybar[0] + T(p[2]) * ct * ftb[1];
      pbar[2] = pbar[2] + ct * y[0] * ftb[1];
      break;
    }
    default: break;
  }
}

// NumPy's float64 sum of a contiguous 1-D array (pairwise summation; blocks of fewer than 8
// elements are summed left to right, blocks up to 128 use 8 interleaved accumulators).
template <class T>
T np_sum(const T* a, int n) {
  if (n < 8) {
    T r = 0.0;
    for (int i = 0; i < n; ++i) r = r + a[i];
    return r;
  }
  if (n <= 128) {
    T r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] = r[j] + a[i + j];
    T res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res = res + a[i];
    return res;
  }
  int n2 = n / 2;
  n2 -= n2 % 8;
  return np_sum(a, n2) + np_sum(a + n2, n - n2);
}

// WRMS norm sqrt(mean((e/sc)^2)), sc = atol + rtol*max(|y|, |ynew|) (tsit5_replay.py:116-117).
template <class T>
T wrms(int D, const T* e, const T* y, const T* ynew, T rtol, T atol) {
  T q[kMaxD];
  for (int i = 0; i < D; ++i) {
    const T sc = atol + rtol * max_(abs_(y[i]), abs_(ynew[i]));
    const T r = e[i] / sc;
    q[i] = r * r;
  }
  return sqrt_(np_sum(q, D) / T(double(D)));
}

// ---------------------------------------------------------------------------------------
// Tsit5 stages. `stage_inputs` forms z_i = y + h*(sum_j a_ij k_j) in the reference's
// textual order: ((a_i1*k1 + a_i2*k2) + ...), then times h, then y + (.).
// ---------------------------------------------------------------------------------------
template <class T>
struct Tsit5Work {
  T k[7][kMaxD];
  T z[7][kMaxD];
};

// Stage time t + c h. Only a field that reads t needs it: every other field ignores its time
// argument, so it gets the step's own t and the sum, which is no part of an autonomous
// field's algorithmic work, is neither formed nor counted (values are unchanged).
template <class T>
T stage_time(const FieldSpec& fs, T t, double c, T h) {
  return field_time_dependent(fs) ? t + T(c) * h : t;
}

template <class T>
T stage_time_end(const FieldSpec& fs, T t, T h) {
  return field_time_dependent(fs) ? t + h : t;
}

// Computes k2..k6 (k1 given in w.k[0]) and their stage inputs z2..z6, then y5 (no B7 term).
template <class T>
void tsit5_stages(const FieldSpec& fs, T t, const T* y, T h, const double* p, Tsit5Work<T>& w,
                  T* y5) {
  const int D = fs.D;
  T(&k)[7][kMaxD] = w.k;
  T(&z)[7][kMaxD] = w.z;
  for (int i = 0; i < D; ++i) z[0][i] = y[i];
  for (int i = 0; i < D; ++i) z[1][i] = y[i] + h * (T(A21) * k[0][i]);
  field_eval(fs, stage_time(fs, t, C2, h), z[1], p, k[1]);
  for (int i = 0; i < D; ++i) z[2][i] = y[i] + h * (T(A31) * k[0][i] + T(A32) * k[1][i]);
  field_eval(fs, stage_time(fs, t, C3, h), z[2], p, k[2]);
  for (int i = 0; i < D; ++i)
    z[3][i] = y[i] + h * (T(A41) * k[0][i] + T(A42) * k[1][i] + T(A43) * k[2][i]);
  field_eval(fs, stage_time(fs, t, C4, h), z[3], p, k[3]);
  for (int i = 0; i < D; ++i)
    z[4][i] = y[i] + h * (T(A51) * k[0][i] + T(A52) * k[1][i] + T(A53) * k[2][i] +
                          T(A54) * k[3][i]);
  field_eval(fs, stage_time(fs, t, C5, h), z[4], p, k[4]);
  for (int i = 0; i < D; ++i)
    z[5][i] = y[i] + h * (T(A61) * k[0][i] + T(A62) * k[1][i] + T(A63) * k[2][i] +
                          T(A64) * k[3][i] + T(A65) * k[4][i]);
  field_eval(fs, stage_time(fs, t, C6, h), z[5], p, k[5]);
  for (int i = 0; i < D; ++i)
    y5[i] = y[i] + h * (T(B1) * k[0][i] + T(B2) * k[1][i] + T(B3) * k[2][i] +
                        T(B4) * k[3][i] + T(B5) * k[4][i] + T(B6) * k[5][i]);
}

// One adaptive trial step (method.py:159-208): returns err; y5, k7 out.
template <class T>
T tsit5_trial(const FieldSpec& fs, T t, const T* y, T h, const double* p, Tsit5Work<T>& w,
              T* y5, T rtol, T atol) {
  const int D = fs.D;
  tsit5_stages(fs, t, y, h, p, w, y5);
  field_eval(fs, stage_time_end(fs, t, h), y5, p, w.k[6]);
  T e[kMaxD];
  const auto& k = w.k;
  for (int i = 0; i < D; ++i)
    e[i] = h * (T(E1) * k[0][i] + T(E2) * k[1][i] + T(E3) * k[2][i] + T(E4) * k[3][i] +
                T(E5) * k[4][i] + T(E6) * k[5][i] + T(E7) * k[6][i]);
  return wrms(D, e, y, y5, rtol, atol);
}

// The replay step tsit5_step (tsit5_step.py:80-102): all 7 stages, k7 at the b-weighted
// point, and the B7*k7 term kept (B7 == 0).
template <class T>
void tsit5_step(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* ynext) {
  const int D = fs.D;
  Tsit5Work<T> w;
  T y5[kMaxD];
  field_eval(fs, t, y, p, w.k[0]);
  tsit5_stages(fs, t, y, h, p, w, y5);
  // k7 input: y + dt*(B1 k1 + ... + B6 k6) -- the same expression as y5.
  field_eval(fs, stage_time(fs, t, C6, h), y5, p, w.k[6]);
  const auto& k = w.k;
  for (int i = 0; i < D; ++i)
    ynext[i] = y[i] + h * (T(B1) * k[0][i] + T(B2) * k[1][i] + T(B3) * k[2][i] +
                           T(B4) * k[3][i] + T(B5) * k[4][i] + T(B6) * k[5][i] +
                           T(B7) * k[6][i]);
}

// Adjoint of one replay step at (t, y, h): lam <- dL/dy given lam = dL/dy_next; pbar += dL/dp.
// The k7 branch carries weight B7 == 0 and contributes nothing for finite values.
template <class T>
void tsit5_step_vjp(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* lam,
                    T* pbar) {
  const int D = fs.D;
  Tsit5Work<T> w;
  T y5[kMaxD];
  field_eval(fs, t, y, p, w.k[0]);
  tsit5_stages(fs, t, y, h, p, w, y5);
  static const double A[6][6] = {
      {0, 0, 0, 0, 0, 0},
      {A21, 0, 0, 0, 0, 0},
      {A31, A32, 0, 0, 0, 0},
      {A41, A42, A43, 0, 0, 0},
      {A51, A52, A53, A54, 0, 0},
      {A61, A62, A63, A64, A65, 0}};
  static const double Bw[6] = {B1, B2, B3, B4, B5, B6};
  static const double Cn[6] = {0.0, C2, C3, C4, C5, C6};
  T kbar[6][kMaxD];
  for (int s = 0; s < 6; ++s)
    for (int i = 0; i < D; ++i) kbar[s][i] = h * (T(Bw[s]) * lam[i]);
  T ynew[kMaxD];
  for (int i = 0; i < D; ++i) ynew[i] = lam[i];
  for (int s = 5; s >= 0; --s) {
    T zbar[kMaxD];
    for (int i = 0; i < D; ++i) zbar[i] = 0.0;
    field_vjp(fs, stage_time(fs, t, Cn[s], h), w.z[s], p, kbar[s], zbar, pbar);
    for (int i = 0; i < D; ++i) ynew[i] = ynew[i] + zbar[i];
    for (int r = 0; r < s; ++r)
      for (int i = 0; i < D; ++i) kbar[r][i] = kbar[r][i] + h * (T(A[s][r]) * zbar[i]);
  }
  for (int i = 0; i < D; ++i) lam[i] = ynew[i];
}

// ---------------------------------------------------------------------------------------
// Rosenbrock23 (ode23s, autonomous): W = I - d*h*J, LU with partial pivoting once per step.
// ---------------------------------------------------------------------------------------
const double RB_D = 1.0 / (2.0 + std::sqrt(2.0));
const double RB_E32 = 6.0 + std::sqrt(2.0);

template <class T>
struct LU {
  T a[kMaxD * kMaxD];
  int piv[kMaxD];
  int D;
};

template <class T>
void lu_factor(LU<T>& f) {
  const int D = f.D;
  for (int i = 0; i < D; ++i) f.piv[i] = i;
  for (int c = 0; c < D; ++c) {
    int pr = c;
    T big = abs_(f.a[c * D + c]);
    for (int r = c + 1; r < D; ++r) {
      const T v = abs_(f.a[r * D + c]);
      if (v > big) { big = v; pr = r; }
    }
    if (pr != c) {
      for (int k = 0; k < D; ++k) std::swap(f.a[c * D + k], f.a[pr * D + k]);
      std::swap(f.piv[c], f.piv[pr]);
    }
    for (int r = c + 1; r < D; ++r) {
      const T l = f.a[r * D + c] / f.a[c * D + c];
      f.a[r * D + c] = l;
      for (int k = c + 1; k < D; ++k) f.a[r * D + k] = f.a[r * D + k] - l * f.a[c * D + k];
    }
  }
}

// Solve W x = b with the factorization P W = L U.
template <class T>
void lu_solve(const LU<T>& f, const T* b, T* x) {
  const int D = f.D;
  T w[kMaxD];
  for (int i = 0; i < D; ++i) w[i] = b[f.piv[i]];
  for (int i = 0; i < D; ++i)
    for (int k = 0; k < i; ++k) w[i] = w[i] - f.a[i * D + k] * w[k];
  for (int i = D - 1; i >= 0; --i) {
    T s = w[i];
    for (int k = i + 1; k < D; ++k) s = s - f.a[i * D + k] * x[k];
    x[i] = s / f.a[i * D + i];
  }
}

// Solve W^T x = b: U^T z = b, L^T v = z, x = P^T v.
template <class T>
void lu_solve_t(const LU<T>& f, const T* b, T* x) {
  const int D = f.D;
  T z[kMaxD], v[kMaxD];
  for (int i = 0; i < D; ++i) {
    T s = b[i];
    for (int k = 0; k < i; ++k) s = s - f.a[k * D + i] * z[k];
    z[i] = s / f.a[i * D + i];
  }
  for (int i = D - 1; i >= 0; --i) {
    T s = z[i];
    for (int k = i + 1; k < D; ++k) s = s - f.a[k * D + i] * v[k];
    v[i] = s;
  }
  for (int i = 0; i < D; ++i) x[f.piv[i]] = v[i];
}

template <class T>
void rb23_build_w(const FieldSpec& fs, T t, const T* y, T h, const double* p, LU<T>& lu, T* J) {
  const int D = fs.D;
  field_jac(fs, t, y, p, J);
  lu.D = D;
  const T dh = h * T(RB_D);
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) lu.a[i * D + j] = (i == j ? T(1.0) : T(0.0)) - dh * J[i * D + j];
  lu_factor(lu);
}

// One Rosenbrock23 trial step (gradsolve_g2_spike_oracle.py:172-183). Returns err.
template <class T>
T rb23_trial(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* ynew, T rtol,
             T atol) {
  const int D = fs.D;
  LU<T> lu;
  T J[kMaxD * kMaxD];
  rb23_build_w(fs, t, y, h, p, lu, J);
  T F0[kMaxD], F1[kMaxD], F2[kMaxD], k1[kMaxD], k2[kMaxD], k3[kMaxD], u[kMaxD], r[kMaxD];
  field_eval(fs, t, y, p, F0);
  lu_solve(lu, F0, k1);
  for (int i = 0; i < D; ++i) u[i] = y[i] + T(0.5) * h * k1[i];
  field_eval(fs, t, u, p, F1);
  for (int i = 0; i < D; ++i) r[i] = F1[i] - k1[i];
  lu_solve(lu, r, k2);
  for (int i = 0; i < D; ++i) k2[i] = k2[i] + k1[i];
  for (int i = 0; i < D; ++i) ynew[i] = y[i] + h * k2[i];
  field_eval(fs, t, ynew, p, F2);
  for (int i = 0; i < D; ++i)
    r[i] = F2[i] - T(RB_E32) * (k2[i] - F1[i]) - T(2.0) * (k1[i] - F0[i]);
  lu_solve(lu, r, k3);
  T e[kMaxD];
  for (int i = 0; i < D; ++i) e[i] = (h / T(6.0)) * (k1[i] - T(2.0) * k2[i] + k3[i]);
  return wrms(D, e, y, ynew, rtol, atol);
}

// Replay step rosenbrock23_step (rosenbrock23_step.py:47-66): no k3.
template <class T>
void rb23_step(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* ynew) {
  const int D = fs.D;
  LU<T> lu;
  T J[kMaxD * kMaxD];
  rb23_build_w(fs, t, y, h, p, lu, J);
  T F0[kMaxD], F1[kMaxD], k1[kMaxD], k2[kMaxD], u[kMaxD], r[kMaxD];
  field_eval(fs, t, y, p, F0);
  lu_solve(lu, F0, k1);
  for (int i = 0; i < D; ++i) u[i] = y[i] + T(0.5) * h * k1[i];
  field_eval(fs, t, u, p, F1);
  for (int i = 0; i < D; ++i) r[i] = F1[i] - k1[i];
  lu_solve(lu, r, k2);
  for (int i = 0; i < D; ++i) k2[i] = k2[i] + k1[i];
  for (int i = 0; i < D; ++i) ynew[i] = y[i] + h * k2[i];
}

// Adjoint of rb23_step: lam = dL/dynew in, dL/dy out; pbar += dL/dp.
template <class T>
void rb23_step_vjp(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* lam,
                   T* pbar) {
  const int D = fs.D;
  LU<T> lu;
  T J[kMaxD * kMaxD];
  rb23_build_w(fs, t, y, h, p, lu, J);
  T F0[kMaxD], F1[kMaxD], k1[kMaxD], x2[kMaxD], u[kMaxD], r[kMaxD];
  field_eval(fs, t, y, p, F0);
  lu_solve(lu, F0, k1);
  for (int i = 0; i < D; ++i) u[i] = y[i] + T(0.5) * h * k1[i];
  field_eval(fs, t, u, p, F1);
  for (int i = 0; i < D; ++i) r[i] = F1[i] - k1[i];
  lu_solve(lu, r, x2);  // k2 = x2 + k1
  // Reverse.
  T ybar[kMaxD], k1b[kMaxD], r2b[kMaxD], F0b[kMaxD], ub[kMaxD], Wb[kMaxD * kMaxD];
  for (int i = 0; i < D; ++i) { ybar[i] = lam[i]; k1b[i] = h * lam[i]; }  // k2bar = h*lam
  for (int i = 0; i < D * D; ++i) Wb[i] = 0.0;
  // k2 = x2 + k1: x2bar = k2bar, k1bar += k2bar (k1b already holds k2bar).
  T x2b[kMaxD];
  for (int i = 0; i < D; ++i) x2b[i] = h * lam[i];
  lu_solve_t(lu, x2b, r2b);  // x2 = W^{-1} r2
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) Wb[i * D + j] = Wb[i * D + j] - r2b[i] * x2[j];
  // r2 = F1 - k1
  for (int i = 0; i < D; ++i) k1b[i] = k1b[i] - r2b[i];
  // F1 = f(u)
  for (int i = 0; i < D; ++i) ub[i] = 0.0;
  field_vjp(fs, t, u, p, r2b, ub, pbar);
  // u = y + 0.5 h k1
  for (int i = 0; i < D; ++i) {
    ybar[i] = ybar[i] + ub[i];
    k1b[i] = k1b[i] + T(0.5) * h * ub[i];
  }
  // k1 = W^{-1} F0
  lu_solve_t(lu, k1b, F0b);
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) Wb[i * D + j] = Wb[i * D + j] - F0b[i] * k1[j];
  field_vjp(fs, t, y, p, F0b, ybar, pbar);
  // W = I - d h J(y, p): Jbar = -d h Wbar
  const T dh = h * T(RB_D);
  T Jb[kMaxD * kMaxD];
  for (int i = 0; i < D * D; ++i) Jb[i] = -dh * Wb[i];
  field_jac_vjp(fs, t, y, p, Jb, ybar, pbar);
  for (int i = 0; i < D; ++i) lam[i] = ybar[i];
}

// ---------------------------------------------------------------------------------------
// Rodas5P (gradsolve/solvers/rodas5p_step.py:119-168). The alpha/Gamma tableau is the
// library's generated header (gsx_rodas5p_tab.h: NumPy's conversion of the published a/C
// form, rodas5p_step.py:63-77); it is pinned separately by its generator test, the published
// row sums, the (a, C)-form stage vectors and the observed orders.
//   J = df/dy(t, y), f_t = df/dt(t, y) (zero for fields that do not read t);
//   W = I - (h gamma) J;  for i = 0..7:  u_i = y + sum_{j<i} alpha_ij k_j,
//   a_i = 0 + sum_{j<i} Gamma_ij k_j,
//   k_i = W \ (h f(t + c_i h, u_i) + h (J a_i) + (h h g_i) f_t),  g_i = sum_j Gamma_ij;
//   y_next = y + sum_i m_i k_i,  err vector = 0 + sum_i (m_i - mhat_i) k_i.
// ---------------------------------------------------------------------------------------
namespace r5 = gsx_rodas5p_tab;

template <class T>
struct R5Stages {
  T k[8][kMaxD];
  T J[kMaxD * kMaxD];
  LU<T> lu;
};

template <class T>
void r5_stage_input(int D, const T* y, const T (&k)[8][kMaxD], int i, T* u, T* acc) {
  for (int d = 0; d < D; ++d) { u[d] = y[d]; acc[d] = T(0.0); }
  for (int j = 0; j < i; ++j)
    for (int d = 0; d < D; ++d) {
      u[d] = u[d] + T(r5::kAlpha[i * 8 + j]) * k[j][d];
      acc[d] = acc[d] + T(r5::kGammaOff[i * 8 + j]) * k[j][d];
    }
}

template <class T>
void r5_stages(const FieldSpec& fs, T t, const T* y, T h, const double* p, R5Stages<T>& w) {
  const int D = fs.D;
  field_jac(fs, t, y, p, w.J);
  w.lu.D = D;
  const T hg = h * T(r5::kGamma);
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j)
      w.lu.a[i * D + j] = (i == j ? T(1.0) : T(0.0)) - hg * w.J[i * D + j];
  lu_factor(w.lu);
  const bool td = field_time_dependent(fs);
  T ft[kMaxD];
  if (td) field_dfdt(fs, t, y, p, ft);
  for (int i = 0; i < 8; ++i) {
    T u[kMaxD], acc[kMaxD], F[kMaxD], rhs[kMaxD];
    r5_stage_input(D, y, w.k, i, u, acc);
    field_eval(fs, stage_time(fs, t, r5::kC[i], h), u, p, F);
    for (int r = 0; r < D; ++r) {
      T ja = T(0.0);
      for (int c = 0; c < D; ++c) ja = ja + w.J[r * D + c] * acc[c];
      rhs[r] = h * F[r] + h * ja;
      if (td) rhs[r] = rhs[r] + (h * h * T(r5::kGammaSum[i])) * ft[r];
    }
    lu_solve(w.lu, rhs, w.k[i]);
  }
}

// One trial step: y_next and the WRMS error (rodas5p_step.py:158-168).
template <class T>
T r5_trial(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* ynew, T rtol,
           T atol, bool* finite) {
  const int D = fs.D;
  R5Stages<T> w;
  r5_stages(fs, t, y, h, p, w);
  T e[kMaxD];
  for (int d = 0; d < D; ++d) { ynew[d] = y[d]; e[d] = T(0.0); }
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < D; ++d) {
      ynew[d] = ynew[d] + T(r5::kM[i]) * w.k[i][d];
      e[d] = e[d] + T(r5::kE[i]) * w.k[i][d];
    }
  bool ok = true;
  for (int d = 0; d < D; ++d) ok = ok && std::isfinite(val(ynew[d])) && std::isfinite(val(e[d]));
  *finite = ok;
  return wrms(D, e, y, ynew, rtol, atol);
}

// Value-only step rodas5p_advance (rodas5p_step.py:149-155).
template <class T>
void r5_step(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* ynew) {
  const int D = fs.D;
  R5Stages<T> w;
  r5_stages(fs, t, y, h, p, w);
  for (int d = 0; d < D; ++d) ynew[d] = y[d];
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < D; ++d) ynew[d] = ynew[d] + T(r5::kM[i]) * w.k[i][d];
}

// The dense-output bundle of a step (rodas5p_step.py:198-226): D_r = 0 + sum_i hd[r][i] k_i.
template <class T>
void r5_dense(int D, const R5Stages<T>& w, T (&dv)[3][kMaxD]) {
  for (int r = 0; r < 3; ++r) {
    for (int d = 0; d < D; ++d) dv[r][d] = T(0.0);
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < D; ++d) dv[r][d] = dv[r][d] + T(r5::kHD[r * 8 + i]) * w.k[i][d];
  }
}

// y(theta) = (1 - th) y_l + th (y_r + (1 - th) (D1 + th (D2 + th D3))) (rodas5p_step.py:171-195).
template <class T>
void r5_dense_eval(int D, const T* yl, const T* yr, const T (&dv)[3][kMaxD], T th, T* out) {
  const T om = T(1.0) - th;
  for (int d = 0; d < D; ++d)
    out[d] = om * yl[d] + th * (yr[d] + om * (dv[0][d] + th * (dv[1][d] + th * dv[2][d])));
}

// Adjoint of r5_step with extra cotangents on the dense vectors: lam = dL/dynew in, dL/dy out;
// dbar (may be null) = dL/dD_r of this step's bundle; pbar += dL/dp.
template <class T>
void r5_step_vjp(const FieldSpec& fs, T t, const T* y, T h, const double* p, T* lam,
                 const T (*dbar)[kMaxD], T* pbar) {
  const int D = fs.D;
  R5Stages<T> w;
  r5_stages(fs, t, y, h, p, w);
  T kb[8][kMaxD], yb[kMaxD], Wb[kMaxD * kMaxD], Jb[kMaxD * kMaxD];
  for (int d = 0; d < D; ++d) yb[d] = lam[d];
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < D; ++d) {
      kb[i][d] = T(r5::kM[i]) * lam[d];
      if (dbar)
        for (int r = 0; r < 3; ++r) kb[i][d] = kb[i][d] + T(r5::kHD[r * 8 + i]) * dbar[r][d];
    }
  const bool td = field_time_dependent(fs);
  T ftb[kMaxD];
  for (int i = 0; i < D * D; ++i) { Wb[i] = T(0.0); Jb[i] = T(0.0); }
  for (int d = 0; d < D; ++d) ftb[d] = T(0.0);
  for (int i = 7; i >= 0; --i) {
    T u[kMaxD], acc[kMaxD], rb[kMaxD], Fb[kMaxD], ub[kMaxD], accb[kMaxD];
    r5_stage_input(D, y, w.k, i, u, acc);
    lu_solve_t(w.lu, kb[i], rb);  // k_i = W^-1 rhs_i
    if (td)
      for (int d = 0; d < D; ++d) ftb[d] = ftb[d] + (h * h * T(r5::kGammaSum[i])) * rb[d];
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) Wb[a * D + b] = Wb[a * D + b] - rb[a] * w.k[i][b];
    for (int d = 0; d < D; ++d) { Fb[d] = h * rb[d]; ub[d] = T(0.0); }
    field_vjp(fs, stage_time(fs, t, r5::kC[i], h), u, p, Fb, ub, pbar);
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) Jb[a * D + b] = Jb[a * D + b] + h * rb[a] * acc[b];
    for (int c = 0; c < D; ++c) {
      T s = T(0.0);
      for (int r = 0; r < D; ++r) s = s + w.J[r * D + c] * rb[r];
      accb[c] = h * s;
    }
    for (int d = 0; d < D; ++d) yb[d] = yb[d] + ub[d];
    for (int j = 0; j < i; ++j)
      for (int d = 0; d < D; ++d)
        kb[j][d] = kb[j][d] + T(r5::kAlpha[i * 8 + j]) * ub[d] +
                   T(r5::kGammaOff[i * 8 + j]) * accb[d];
  }
  const T hg = h * T(r5::kGamma);
  for (int i = 0; i < D * D; ++i) Jb[i] = Jb[i] - hg * Wb[i];
  field_jac_vjp(fs, t, y, p, Jb, yb, pbar);
  if (td) field_dfdt_vjp(fs, t, y, p, ftb, yb, pbar);
  for (int d = 0; d < D; ++d) lam[d] = yb[d];
}

// Record with the device recorder's Rodas5P semantics (gradsolve/solvers/record_jax.py:83-132
// with RODAS5P_METHOD): a step below the underflow floor (t + h == t or h <= floor) ends the
// trajectory with status 2; a non-finite trial is rejected and shrinks h by FAC_MIN without a
// controller update; the I-controller uses exponent -1/5; max_steps counts attempts.
void r5_record_one(const FieldSpec& fs, const double* y0, const double* p, double t0, double t1,
                   double dt0, double rtol, double atol, int64_t max_steps, double floor,
                   double* yf, int64_t* nacc, int64_t* nrej, int32_t* st, double* tape,
                   int64_t cap) {
  const int D = fs.D;
  double y[kMaxD], ynew[kMaxD];
  for (int i = 0; i < D; ++i) y[i] = y0[i];
  double t = t0, dt = dt0;
  int64_t acc = 0, rej = 0, attempts = 0;
  int32_t status = 0;
  if (tape)
    for (int64_t s = 0; s < cap; ++s) tape[s] = 0.0;
  while (t < t1 && status == 0) {
    const double h = min_(dt, t1 - t);
    if (t + h == t || h <= floor) {
      status = 2;
      break;
    }
    bool finite = true;
    const double err = r5_trial(fs, t, y, h, p, ynew, rtol, atol, &finite);
    const bool ok = finite && std::isfinite(err);
    ++attempts;
    if (ok && err <= 1.0) {
      if (tape && acc >= cap) { status = 3; break; }
      t = t + h;
      for (int i = 0; i < D; ++i) y[i] = ynew[i];
      if (tape) tape[acc] = h;
      ++acc;
      dt = h * min_(max_(SAFETY * pow_(max_(err, ERR_FLOOR), TSIT5_EXP), FAC_MIN), FAC_MAX);
    } else {
      ++rej;
      dt = ok ? h * min_(max_(SAFETY * pow_(max_(err, ERR_FLOOR), TSIT5_EXP), FAC_MIN), FAC_MAX)
              : h * FAC_MIN;
    }
    if (attempts >= max_steps && t < t1) status = 1;
  }
  for (int i = 0; i < D; ++i) yf[i] = y[i];
  *nacc = acc;
  *nrej = rej;
  *st = status;
}

// ---------------------------------------------------------------------------------------
// Ensemble drivers.
// ---------------------------------------------------------------------------------------
enum Method { kTsit5 = 0, kRb23 = 1, kRodas5p = 2 };

// The value-only step and its VJP of method M.
template <int M>
void m_step(const FieldSpec& fs, double t, const double* y, double h, const double* p,
            double* yn) {
  if (M == kTsit5)
    tsit5_step(fs, t, y, h, p, yn);
  else if (M == kRb23)
    rb23_step(fs, t, y, h, p, yn);
  else
    r5_step(fs, t, y, h, p, yn);
}

template <int M>
void m_step_vjp(const FieldSpec& fs, double t, const double* y, double h, const double* p,
                double* lam, double* pb) {
  if (M == kTsit5)
    tsit5_step_vjp(fs, t, y, h, p, lam, pb);
  else if (M == kRb23)
    rb23_step_vjp(fs, t, y, h, p, lam, pb);
  else
    r5_step_vjp<double>(fs, t, y, h, p, lam, nullptr, pb);
}

template <int M>
void record_one(const FieldSpec& fs, const double* y0, const double* p, double t0, double t1,
                double dt0, double rtol, double atol, int64_t max_steps, double* yf,
                int64_t* nacc, int64_t* nrej, int32_t* st, double* tape, int64_t cap) {
  const int D = fs.D;
  double y[kMaxD], ynew[kMaxD];
  for (int i = 0; i < D; ++i) y[i] = y0[i];
  double t = t0, dt = dt0;
  Tsit5Work<double> w;
  if (M == kTsit5) field_eval(fs, t, y, p, w.k[0]);
  int64_t acc = 0, rej = 0;
  int32_t status = 0;
  if (tape)
    for (int64_t s = 0; s < cap; ++s) tape[s] = 0.0;
  const double order_exp = (M == kTsit5) ? TSIT5_EXP : RB23_EXP;
  for (int64_t it = 0; it < max_steps; ++it) {
    if (t >= t1) break;
    const double h = min_(dt, t1 - t);
    double err;
    if (M == kTsit5)
      err = tsit5_trial(fs, t, y, h, p, w, ynew, rtol, atol);
    else
      err = rb23_trial(fs, t, y, h, p, ynew, rtol, atol);
    if (err <= 1.0) {
      if (tape && acc >= cap) { status = 3; break; }
      t = t + h;
      for (int i = 0; i < D; ++i) y[i] = ynew[i];
      if (M == kTsit5)
        for (int i = 0; i < D; ++i) w.k[0][i] = w.k[6][i];
      if (tape) tape[acc] = h;
      ++acc;
    } else {
      ++rej;
    }
    dt = h * min_(max_(SAFETY * pow_(max_(err, ERR_FLOOR), order_exp), FAC_MIN), FAC_MAX);
  }
  if (status == 0) status = (t >= t1) ? 0 : 1;
  for (int i = 0; i < D; ++i) yf[i] = y[i];
  *nacc = acc;
  *nrej = rej;
  *st = status;
}

template <int M>
int record_ensemble(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                    const double* p, double t0, const double* t1, const double* dt0,
                    double rtol, double atol, int64_t max_steps, double* y_final,
                    int64_t* n_acc, int64_t* n_rej, int32_t* status, double* tape,
                    int64_t cap, int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || (tape && cap < 0)) return -1;
  if (M == kRb23 && !field_has_jac(fs)) return -2;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j) {
    const double t1j = t1[j];
    const double dt0j = dt0 ? dt0[j] : (t1j - t0) / 100.0;
    record_one<M>(fs, y0 + j * D, p + j * P, t0, t1j, dt0j, rtol, atol, max_steps,
                  y_final + j * D, n_acc + j, n_rej + j, status + j,
                  tape ? tape + j * cap : nullptr, cap);
  }
  return 0;
}

template <int M>
int replay_ensemble(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                    const double* p, double t0, const double* tape, int64_t S, double* y_final,
                    int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || S < 0) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j) {
    double y[kMaxD], yn[kMaxD];
    for (int i = 0; i < D; ++i) y[i] = y0[j * D + i];
    double t = t0;
    for (int64_t s = 0; s < S; ++s) {
      const double h = tape[j * S + s];
      m_step<M>(fs, t, y, h, p + j * P, yn);
      for (int i = 0; i < D; ++i) y[i] = yn[i];
      t = t + h;
    }
    for (int i = 0; i < D; ++i) y_final[j * D + i] = y[i];
  }
  return 0;
}

// Reverse sweep with saved states. With saveat (ts != nullptr) the bracket-and-re-step
// outputs of dense.py:106-163 are differentiated too.
template <int M>
void adjoint_one(const FieldSpec& fs, const double* y0, const double* p, double t0,
                 const double* tape, int64_t S, const double* ybar, const double* ts,
                 int64_t k, const double* ysbar, double* y0bar, double* pbar) {
  const int D = fs.D, P = fs.P;
  std::vector<double> ys((S + 1) * D), tt(S + 1);
  for (int i = 0; i < D; ++i) ys[i] = y0[i];
  tt[0] = t0;
  for (int64_t s = 0; s < S; ++s) {
    m_step<M>(fs, tt[s], &ys[s * D], tape[s], p, &ys[(s + 1) * D]);
    tt[s + 1] = tt[s] + tape[s];
  }
  double lam[kMaxD], pb[kMaxD];
  for (int i = 0; i < D; ++i) lam[i] = ybar ? ybar[i] : 0.0;
  for (int i = 0; i < P; ++i) pb[i] = 0.0;
  // Dense-output brackets (the last bracketing step wins, as jnp.where overwrites).
  std::vector<int64_t> br;
  std::vector<double> y0_extra(D, 0.0);
  if (ts) {
    br.assign(k, -1);
    const double t_fin = tt[S];
    for (int64_t q = 0; q < k; ++q) {
      for (int64_t s = 0; s < S; ++s) {
        const double dt = tape[s];
        if (dt > 0.0 && ts[q] > tt[s] && ts[q] <= tt[s] + dt) br[q] = s;
      }
      if (ts[q] >= t_fin) {
        for (int i = 0; i < D; ++i) lam[i] += ysbar[q * D + i];
        br[q] = -2;  // resolved to y_final
      } else if (br[q] == -1) {
        for (int i = 0; i < D; ++i) y0_extra[i] += ysbar[q * D + i];
      }
    }
  }
  for (int64_t s = S - 1; s >= 0; --s) {
    m_step_vjp<M>(fs, tt[s], &ys[s * D], tape[s], p, lam, pb);
    if (ts) {
      for (int64_t q = 0; q < k; ++q) {
        if (br[q] != s) continue;
        double mu[kMaxD];
        for (int i = 0; i < D; ++i) mu[i] = ysbar[q * D + i];
        const double hq = ts[q] - tt[s];
        m_step_vjp<M>(fs, tt[s], &ys[s * D], hq, p, mu, pb);
        for (int i = 0; i < D; ++i) lam[i] += mu[i];
      }
    }
  }
  for (int i = 0; i < D; ++i) y0bar[i] = lam[i] + y0_extra[i];
  for (int i = 0; i < P; ++i) pbar[i] = pb[i];
}

template <int M>
int adjoint_ensemble(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                     const double* p, double t0, const double* tape, int64_t S,
                     const double* ybar, const double* ts, int64_t k, const double* ysbar,
                     double* y0bar, double* pbar, int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || S < 0 || P > kMaxD) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j)
    adjoint_one<M>(fs, y0 + j * D, p + j * P, t0, tape + j * S, S, ybar ? ybar + j * D : nullptr,
                   ts, k, ysbar ? ysbar + j * k * D : nullptr, y0bar + j * D, pbar + j * P);
  return 0;
}

// Bracket-and-re-step dense output (gradsolve/solvers/dense.py:106-163): while the mesh is
// replayed, each save time remembers the start state and time of the last step with
// dt > 0 and t < ts <= t + dt; afterwards a save at or past the final time is y_final, a
// bracketed save is one step of size ts - t_left from there, and any other save is y0.
template <int M>
int replay_saveat_ensemble(int field, int64_t n, int D, int P, const double* aux,
                           const double* y0, const double* p, double t0, const double* tape,
                           int64_t S, const double* ts, int64_t k, double* y_final, double* ys,
                           int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || S < 0 || k < 0) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j) {
    const double* tp = tape + j * S;
    const double* pj = p + j * P;
    std::vector<double> yl(k * D), tl(k, t0);
    std::vector<char> written(k, 0);
    double y[kMaxD], yn[kMaxD];
    for (int i = 0; i < D; ++i) y[i] = y0[j * D + i];
    for (int64_t q = 0; q < k; ++q)
      for (int i = 0; i < D; ++i) yl[q * D + i] = y[i];
    double t = t0;
    for (int64_t s = 0; s < S; ++s) {
      const double dt = tp[s];
      const double tn = t + dt;
      for (int64_t q = 0; q < k; ++q) {
        if (dt > 0.0 && ts[q] > t && ts[q] <= tn) {
          for (int i = 0; i < D; ++i) yl[q * D + i] = y[i];
          tl[q] = t;
          written[q] = 1;
        }
      }
      m_step<M>(fs, t, y, dt, pj, yn);
      for (int i = 0; i < D; ++i) y[i] = yn[i];
      t = tn;
    }
    for (int i = 0; i < D; ++i) y_final[j * D + i] = y[i];
    for (int64_t q = 0; q < k; ++q) {
      double* out = ys + (j * k + q) * D;
      if (ts[q] >= t) {
        for (int i = 0; i < D; ++i) out[i] = y[i];
      } else if (written[q]) {
        m_step<M>(fs, tl[q], &yl[q * D], ts[q] - tl[q], pj, out);
      } else {
        for (int i = 0; i < D; ++i) out[i] = y0[j * D + i];
      }
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int gso_tsit5_tableau(double* c, double* a, double* b, double* e) {
  const double cc[5] = {C2, C3, C4, C5, C6};
  const double aa[15] = {A21, A31, A32, A41, A42, A43, A51, A52,
                         A53, A54, A61, A62, A63, A64, A65};
  const double bb[6] = {B1, B2, B3, B4, B5, B6};
  const double ee[7] = {E1, E2, E3, E4, E5, E6, E7};
  std::memcpy(c, cc, sizeof cc);
  std::memcpy(a, aa, sizeof aa);
  std::memcpy(b, bb, sizeof bb);
  std::memcpy(e, ee, sizeof ee);
  return 0;
}

int gso_tsit5_record(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                     const double* p, double t0, const double* t1, const double* dt0,
                     double rtol, double atol, int64_t max_steps, double* y_final,
                     int64_t* n_acc, int64_t* n_rej, int32_t* status, double* tape,
                     int64_t cap, int nthreads) {
  return record_ensemble<kTsit5>(field, n, D, P, aux, y0, p, t0, t1, dt0, rtol, atol,
                                 max_steps, y_final, n_acc, n_rej, status, tape, cap, nthreads);
}

int gso_tsit5_teacher_forced(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0, const double* t1,
                             double rtol, double atol, const double* tape, int64_t S,
                             const int64_t* n_acc, double* h_oracle, int64_t* n_rej_oracle,
                             int64_t* rej_at, int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || S < 0) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j) {
    const double* tp = tape + j * S;
    const double* pj = p + j * P;
    double* ho = h_oracle + j * S;
    double y[kMaxD], y5[kMaxD];
    for (int i = 0; i < D; ++i) y[i] = y0[j * D + i];
    double t = t0;
    const double t1j = t1[j];
    double dt = (t1j - t0) / 100.0;
    Tsit5Work<double> w;
    field_eval(fs, t, y, pj, w.k[0]);
    int64_t rej = 0;
    for (int64_t s = 0; s < S; ++s) {
      ho[s] = 0.0;
      if (rej_at) rej_at[j * S + s] = 0;
      if (s >= n_acc[j]) continue;
      // The oracle's controller from this (teacher-forced) state: attempts until one is
      // accepted; that step size is the oracle's entry s.
      for (int64_t guard = 0; guard < 100000; ++guard) {
        if (t >= t1j) break;
        const double h = min_(dt, t1j - t);
        const double err = tsit5_trial(fs, t, y, h, pj, w, y5, rtol, atol);
        dt = h * min_(max_(SAFETY * pow_(max_(err, ERR_FLOOR), TSIT5_EXP), FAC_MIN), FAC_MAX);
        if (err <= 1.0) { ho[s] = h; break; }
        ++rej;
        if (rej_at) ++rej_at[j * S + s];
      }
      // Teacher forcing: advance with the device's step size and derive the next proposal
      // from the device's step.
      const double hg = tp[s];
      const double errg = tsit5_trial(fs, t, y, hg, pj, w, y5, rtol, atol);
      dt = hg * min_(max_(SAFETY * pow_(max_(errg, ERR_FLOOR), TSIT5_EXP), FAC_MIN), FAC_MAX);
      t = t + hg;
      for (int i = 0; i < D; ++i) { y[i] = y5[i]; w.k[0][i] = w.k[6][i]; }
    }
    n_rej_oracle[j] = rej;
  }
  return 0;
}

int gso_tsit5_replay(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                     const double* p, double t0, const double* tape, int64_t S, double* y_final,
                     int nthreads) {
  return replay_ensemble<kTsit5>(field, n, D, P, aux, y0, p, t0, tape, S, y_final, nthreads);
}

int gso_tsit5_adjoint(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                      const double* p, double t0, const double* tape, int64_t S,
                      const double* ybar, double* y0bar, double* pbar, int nthreads) {
  return adjoint_ensemble<kTsit5>(field, n, D, P, aux, y0, p, t0, tape, S, ybar, nullptr, 0,
                                  nullptr, y0bar, pbar, nthreads);
}

int gso_tsit5_replay_saveat(int field, int64_t n, int D, int P, const double* aux,
                            const double* y0, const double* p, double t0, const double* tape,
                            int64_t S, const double* ts, int64_t k, double* y_final,
                            double* ys, int nthreads) {
  return replay_saveat_ensemble<kTsit5>(field, n, D, P, aux, y0, p, t0, tape, S, ts, k,
                                        y_final, ys, nthreads);
}

int gso_tsit5_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                             const double* y0, const double* p, double t0, const double* tape,
                             int64_t S, const double* ts, int64_t k, const double* ysbar,
                             const double* yfbar, double* y0bar, double* pbar, int nthreads) {
  std::vector<double> zeros;
  if (!ysbar) {
    zeros.assign(static_cast<size_t>(n * k * D), 0.0);
    ysbar = zeros.data();
  }
  return adjoint_ensemble<kTsit5>(field, n, D, P, aux, y0, p, t0, tape, S, yfbar, ts, k, ysbar,
                                  y0bar, pbar, nthreads);
}

int gso_rb23_record(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                    const double* p, double t0, const double* t1, const double* dt0,
                    double rtol, double atol, int64_t max_steps, double* y_final,
                    int64_t* n_acc, int64_t* n_rej, int32_t* status, double* tape, int64_t cap,
                    int nthreads) {
  return record_ensemble<kRb23>(field, n, D, P, aux, y0, p, t0, t1, dt0, rtol, atol, max_steps,
                                y_final, n_acc, n_rej, status, tape, cap, nthreads);
}

int gso_rb23_replay(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                    const double* p, double t0, const double* tape, int64_t S, double* y_final,
                    int nthreads) {
  return replay_ensemble<kRb23>(field, n, D, P, aux, y0, p, t0, tape, S, y_final, nthreads);
}

int gso_rb23_adjoint(int field, int64_t n, int D, int P, const double* aux, const double* y0,
                     const double* p, double t0, const double* tape, int64_t S,
                     const double* ybar, double* y0bar, double* pbar, int nthreads) {
  return adjoint_ensemble<kRb23>(field, n, D, P, aux, y0, p, t0, tape, S, ybar, nullptr, 0,
                                 nullptr, y0bar, pbar, nthreads);
}

int gso_rb23_replay_saveat(int field, int64_t n, int D, int P, const double* aux,
                           const double* y0, const double* p, double t0, const double* tape,
                           int64_t S, const double* ts, int64_t k, double* y_final, double* ys,
                           int nthreads) {
  return replay_saveat_ensemble<kRb23>(field, n, D, P, aux, y0, p, t0, tape, S, ts, k, y_final,
                                       ys, nthreads);
}

int gso_rb23_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                            const double* y0, const double* p, double t0, const double* tape,
                            int64_t S, const double* ts, int64_t k, const double* ysbar,
                            const double* yfbar, double* y0bar, double* pbar, int nthreads) {
  std::vector<double> zeros;
  if (!ysbar) {
    zeros.assign(static_cast<size_t>(n * k * D), 0.0);
    ysbar = zeros.data();
  }
  return adjoint_ensemble<kRb23>(field, n, D, P, aux, y0, p, t0, tape, S, yfbar, ts, k, ysbar,
                                 y0bar, pbar, nthreads);
}

int gso_rodas5p_record(int field, int64_t n, int D, int P, const double* aux,
                       const double* y0, const double* p, double t0, const double* t1,
                       const double* dt0, double rtol, double atol, int64_t max_steps,
                       double* y_final, int64_t* n_acc, int64_t* n_rej, int32_t* status,
                       double* tape, int64_t cap, int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || (tape && cap < 0)) return -1;
  if (!field_has_jac(fs)) return -2;
  const double eps = std::numeric_limits<double>::epsilon();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j) {
    const double t1j = t1[j];
    const double dt0j = dt0 ? dt0[j] : (t1j - t0) / 100.0;
    const double floor = 4.0 * eps * max_(std::fabs(t0), std::fabs(t1j - t0));
    r5_record_one(fs, y0 + j * D, p + j * P, t0, t1j, dt0j, rtol, atol, max_steps, floor,
                  y_final + j * D, n_acc + j, n_rej + j, status + j,
                  tape ? tape + j * cap : nullptr, cap);
  }
  return 0;
}

int gso_rodas5p_replay(int field, int64_t n, int D, int P, const double* aux,
                       const double* y0, const double* p, double t0, const double* tape,
                       int64_t S, double* y_final, int nthreads) {
  return replay_ensemble<kRodas5p>(field, n, D, P, aux, y0, p, t0, tape, S, y_final, nthreads);
}

int gso_rodas5p_adjoint(int field, int64_t n, int D, int P, const double* aux,
                        const double* y0, const double* p, double t0, const double* tape,
                        int64_t S, const double* ybar, double* y0bar, double* pbar,
                        int nthreads) {
  return adjoint_ensemble<kRodas5p>(field, n, D, P, aux, y0, p, t0, tape, S, ybar, nullptr, 0,
                                    nullptr, y0bar, pbar, nthreads);
}

int gso_rodas5p_replay_saveat(int field, int64_t n, int D, int P, const double* aux,
                              const double* y0, const double* p, double t0, const double* tape,
                              int64_t S, const double* ts, int64_t k, double* y_final,
                              double* ys, int nthreads) {
  return replay_saveat_ensemble<kRodas5p>(field, n, D, P, aux, y0, p, t0, tape, S, ts, k,
                                          y_final, ys, nthreads);
}

int gso_rodas5p_adjoint_saveat(int field, int64_t n, int D, int P, const double* aux,
                               const double* y0, const double* p, double t0,
                               const double* tape, int64_t S, const double* ts, int64_t k,
                               const double* ysbar, const double* yfbar, double* y0bar,
                               double* pbar, int nthreads) {
  std::vector<double> zeros;
  if (!ysbar) {
    zeros.assign(static_cast<size_t>(n * k * D), 0.0);
    ysbar = zeros.data();
  }
  return adjoint_ensemble<kRodas5p>(field, n, D, P, aux, y0, p, t0, tape, S, yfbar, ts, k,
                                    ysbar, y0bar, pbar, nthreads);
}

// Dense output by the continuous extension (gradsolve/solvers/dense.py:166-227): a save time
// bracketed by step s (dt > 0, t_s < ts <= t_s + dt, the last such step) is the step's
// interpolant at theta = (ts - t_s) / dt; at or past the final time it is y_final, and
// unbracketed times (at or before t0) are y0.
int gso_rodas5p_replay_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                                    const double* y0, const double* p, double t0,
                                    const double* tape, int64_t S, const double* ts, int64_t k,
                                    double* y_final, double* ys, int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || S < 0 || k < 0) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j) {
    const double* tp = tape + j * S;
    const double* pj = p + j * P;
    double* out = ys + j * k * D;
    std::vector<char> written(k, 0);
    double y[kMaxD], yn[kMaxD];
    for (int i = 0; i < D; ++i) y[i] = y0[j * D + i];
    double t = t0;
    for (int64_t s = 0; s < S; ++s) {
      const double dt = tp[s];
      const double tn = t + dt;
      R5Stages<double> w;
      r5_stages(fs, t, y, dt, pj, w);
      for (int d = 0; d < D; ++d) yn[d] = y[d];
      for (int i = 0; i < 8; ++i)
        for (int d = 0; d < D; ++d) yn[d] = yn[d] + r5::kM[i] * w.k[i][d];
      double dv[3][kMaxD];
      bool any = false;
      for (int64_t q = 0; q < k; ++q) any = any || (dt > 0.0 && ts[q] > t && ts[q] <= tn);
      if (any) r5_dense(D, w, dv);
      for (int64_t q = 0; q < k; ++q) {
        if (dt > 0.0 && ts[q] > t && ts[q] <= tn) {
          r5_dense_eval(D, y, yn, dv, (ts[q] - t) / dt, out + q * D);
          written[q] = 1;
        }
      }
      for (int d = 0; d < D; ++d) y[d] = yn[d];
      t = tn;
    }
    for (int i = 0; i < D; ++i) y_final[j * D + i] = y[i];
    for (int64_t q = 0; q < k; ++q) {
      if (ts[q] >= t) {
        for (int i = 0; i < D; ++i) out[q * D + i] = y[i];
      } else if (!written[q]) {
        for (int i = 0; i < D; ++i) out[q * D + i] = y0[j * D + i];
      }
    }
  }
  return 0;
}

// Its adjoint: a bracketed save sends (1 - th) of its cotangent to y_s, th to y_{s+1} and
// th (1 - th) th^r to the step's dense vector D_{r+1}, which the step VJP carries to the stages.
int gso_rodas5p_adjoint_saveat_dense(int field, int64_t n, int D, int P, const double* aux,
                                     const double* y0, const double* p, double t0,
                                     const double* tape, int64_t S, const double* ts, int64_t k,
                                     const double* ysbar, const double* yfbar, double* y0bar,
                                     double* pbar, int nthreads) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || n < 0 || S < 0 || k < 0 || P > kMaxD) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j) {
    const double* tp = tape + j * S;
    const double* pj = p + j * P;
    const double* sb = ysbar + j * k * D;
    std::vector<double> st((S + 1) * D), tt(S + 1);
    for (int i = 0; i < D; ++i) st[i] = y0[j * D + i];
    tt[0] = t0;
    for (int64_t s = 0; s < S; ++s) {
      r5_step(fs, tt[s], &st[s * D], tp[s], pj, &st[(s + 1) * D]);
      tt[s + 1] = tt[s] + tp[s];
    }
    const double t_fin = tt[S];
    std::vector<int64_t> br(k, -1);
    double lam[kMaxD], pb[kMaxD], extra[kMaxD];
    for (int i = 0; i < D; ++i) {
      lam[i] = yfbar ? yfbar[j * D + i] : 0.0;
      extra[i] = 0.0;
    }
    for (int i = 0; i < P; ++i) pb[i] = 0.0;
    for (int64_t q = 0; q < k; ++q) {
      for (int64_t s = 0; s < S; ++s)
        if (tp[s] > 0.0 && ts[q] > tt[s] && ts[q] <= tt[s] + tp[s]) br[q] = s;
      if (ts[q] >= t_fin) {
        for (int i = 0; i < D; ++i) lam[i] += sb[q * D + i];
        br[q] = -2;
      } else if (br[q] == -1) {
        for (int i = 0; i < D; ++i) extra[i] += sb[q * D + i];
      }
    }
    for (int64_t s = S - 1; s >= 0; --s) {
      double dbar[3][kMaxD];
      bool any = false;
      for (int r = 0; r < 3; ++r)
        for (int i = 0; i < D; ++i) dbar[r][i] = 0.0;
      for (int64_t q = 0; q < k; ++q) {
        if (br[q] != s) continue;
        any = true;
        const double th = (ts[q] - tt[s]) / tp[s], om = 1.0 - th;
        for (int i = 0; i < D; ++i) {
          const double g = sb[q * D + i];
          lam[i] += th * g;
          dbar[0][i] += th * om * g;
          dbar[1][i] += th * om * th * g;
          dbar[2][i] += th * om * th * th * g;
        }
      }
      r5_step_vjp<double>(fs, tt[s], &st[s * D], tp[s], pj, lam, any ? dbar : nullptr, pb);
      for (int64_t q = 0; q < k; ++q) {
        if (br[q] != s) continue;
        const double om = 1.0 - (ts[q] - tt[s]) / tp[s];
        for (int i = 0; i < D; ++i) lam[i] += om * sb[q * D + i];
      }
    }
    for (int i = 0; i < D; ++i) y0bar[j * D + i] = lam[i] + extra[i];
    for (int i = 0; i < P; ++i) pbar[j * P + i] = pb[i];
  }
  return 0;
}

// The Rodas5P tableau the oracle and the library use (alpha, Gamma strictly lower, c, d, m,
// m - mhat, dense weights), for the tableau tests.
int gso_rodas5p_tableau(double* alpha, double* gamma_off, double* c, double* d, double* m,
                        double* e, double* hd, double* gamma) {
  std::memcpy(alpha, r5::kAlpha, sizeof r5::kAlpha);
  std::memcpy(gamma_off, r5::kGammaOff, sizeof r5::kGammaOff);
  std::memcpy(c, r5::kC, sizeof r5::kC);
  std::memcpy(d, r5::kD, sizeof r5::kD);
  std::memcpy(m, r5::kM, sizeof r5::kM);
  std::memcpy(e, r5::kE, sizeof r5::kE);
  std::memcpy(hd, r5::kHD, sizeof r5::kHD);
  *gamma = r5::kGamma;
  return 0;
}

int gso_rodas5p_stages(int field, int D, int P, const double* aux, double t, const double* y,
                       double h, const double* p, double* k) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs) || D > kMaxD) return -1;
  double J[kMaxD * kMaxD];
  if (!field_jac(fs, t, y, p, J)) return -1;
  R5Stages<double> w;
  r5_stages(fs, t, y, h, p, w);
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < D; ++d) k[i * D + d] = w.k[i][d];
  return 0;
}

int gso_field(int field, int D, int P, const double* aux, double t, const double* y,
              const double* p, double* dy) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs)) return -1;
  field_eval(fs, t, y, p, dy);
  return 0;
}

int gso_field_vjp(int field, int D, int P, const double* aux, double t, const double* y,
                  const double* p, const double* v, double* ybar, double* pbar) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs)) return -1;
  field_vjp(fs, t, y, p, v, ybar, pbar);
  return 0;
}

int gso_count_ops(int which, int field, int D, int P, const double* aux, double* flops,
                  double* pows) {
  FieldSpec fs{field, D, P, aux};
  if (!field_ok(fs)) return -1;
  OpD y[kMaxD], yn[kMaxD], lam[kMaxD], pb[kMaxD];
  std::vector<double> pv(P > 0 ? P : 1, 0.5);
  for (int i = 0; i < D; ++i) { y[i] = 0.1 * (i + 1); lam[i] = 1.0; }
  for (int i = 0; i < P; ++i) pb[i] = 0.0;
  const OpD t = 0.0, h = 1e-3, rtol = 1e-8, atol = 1e-8;
  g_flops = 0.0;
  g_pows = 0.0;
  switch (which) {
    case 0: {
      // One adaptive trial step with FSAL (k1 carried in), plus the controller update
      // (max/pow/min/clamp and h*fac) and the time update t + h.
      Tsit5Work<OpD> w;
      for (int i = 0; i < D; ++i) w.k[0][i] = 1.0;
      const OpD err = tsit5_trial(fs, t, y, h, pv.data(), w, yn, rtol, atol);
      const OpD fac = min_(max_(OpD(SAFETY) * pow_(max_(err, OpD(ERR_FLOOR)), OpD(TSIT5_EXP)),
                                OpD(FAC_MIN)), OpD(FAC_MAX));
      OpD dtn = h * fac;
      OpD tn = t + h;
      (void)dtn;
      (void)tn;
      break;
    }
    case 1: tsit5_step(fs, t, y, h, pv.data(), yn); break;
    case 2: tsit5_step_vjp(fs, t, y, h, pv.data(), lam, pb); break;
    case 3: {
      const OpD err = rb23_trial(fs, t, y, h, pv.data(), yn, rtol, atol);
      const OpD fac = min_(max_(OpD(SAFETY) * pow_(max_(err, OpD(ERR_FLOOR)), OpD(RB23_EXP)),
                                OpD(FAC_MIN)), OpD(FAC_MAX));
      OpD dtn = h * fac;
      (void)dtn;
      break;
    }
    case 4: rb23_step_vjp(fs, t, y, h, pv.data(), lam, pb); break;
    case 5: {
      // One Rodas5P trial step (J, one LU, 8 stages, solution and error combinations, the
      // WRMS norm) plus the controller update.
      bool finite = true;
      const OpD err = r5_trial(fs, t, y, h, pv.data(), yn, rtol, atol, &finite);
      const OpD fac = min_(max_(OpD(SAFETY) * pow_(max_(err, OpD(ERR_FLOOR)), OpD(TSIT5_EXP)),
                                OpD(FAC_MIN)), OpD(FAC_MAX));
      OpD dtn = h * fac;
      (void)dtn;
      break;
    }
    case 6: r5_step_vjp(fs, t, y, h, pv.data(), lam, static_cast<const OpD (*)[kMaxD]>(nullptr),
                        pb); break;
    default: return -1;
  }
  *flops = g_flops;
  *pows = g_pows;
  return 0;
}

}  // extern "C"
