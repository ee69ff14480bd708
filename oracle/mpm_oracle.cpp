// ORACLE — test infrastructure only. CPU restatement of DexDeform's differentiable
// MLS-MPM substep and its hand-written adjoint, used as the checker for the HIP
// kernels in dexdeform_amd/csrc. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product path never does.
//
// Every function cites the reference file:line it restates (paths relative to the
// DexDeform repository root). Templated on the storage/arithmetic type T:
//   T = float  : follows the reference's FP32 arithmetic (FP64 SVD internals, the
//                few double-precision constants and casts the reference has),
//   T = double : the same equations entirely in FP64 (ground truth for finite
//                differences and for judging FP32 rounding spread).
// Scatter order is the sequential particle order (deterministic) unless built with
// -fopenmp, in which case scatters use `omp atomic` (the timed CPU baseline).
//
// Adjoint modes: ADJ_REFERENCE reproduces the reference backward exactly, including
// two known departures from the true derivative (capsule SDF-gradient adjoint is
// identically zero, shape.h:136-147 uses the zero-initialised grad_in; the sticky
// ground branch ground_friction>=99 has no backward, integrator.cu:844-884).
// ADJ_EXACT fixes both, so finite differences agree with it.
//
// Memory safety: the stencil base index is clamped to [0, dim-3] per axis. For any
// particle inside the simulated domain this is the identity (positions are clamped to
// [3dx, (dim-3)dx] by g2p, integrator.cu:1100); it only changes out-of-domain inputs,
// which read and write out of bounds in the reference.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

enum { ADJ_REFERENCE = 0, ADJ_EXACT = 1 };

template <class T> struct V3 { T x, y, z; };
template <class T> struct M3 { T m[3][3]; };
template <class T> struct Q4 { T w, x, y, z; };

// ---------------------------------------------------------------- vec3 (vec3.h:26-120)
template <class T> inline V3<T> v3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <class T> inline V3<T> v3s(T a) { return V3<T>{a, a, a}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> operator+(V3<T> a, T s) { return {a.x + s, a.y + s, a.z + s}; }
template <class T> inline V3<T> operator-(V3<T> a, T s) { return {a.x - s, a.y - s, a.z - s}; }
template <class T> inline V3<T> operator-(T s, V3<T> a) { return {s - a.x, s - a.y, s - a.z}; }
template <class T> inline void operator+=(V3<T> &a, V3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline V3<T> pow2(V3<T> a) { return a * a; }
template <class T> inline T vsum(V3<T> a) { return a.x + a.y + a.z; }
template <class T> inline V3<T> vmax(V3<T> a, T b) { return {std::fmax(a.x, b), std::fmax(a.y, b), std::fmax(a.z, b)}; }
template <class T> inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class T> inline V3<T> vlog(V3<T> a) { return {std::log(a.x), std::log(a.y), std::log(a.z)}; }
template <class T> inline V3<T> vexp(V3<T> a) { return {std::exp(a.x), std::exp(a.y), std::exp(a.z)}; }

// ---------------------------------------------------------------- mat3 (mat3.h:8-186)
// Row-major m[row][col]. Constructor from three vectors builds COLUMNS (mat3.h:11-21).
template <class T> inline M3<T> mcols(V3<T> c0, V3<T> c1, V3<T> c2) {
  M3<T> r;
  r.m[0][0] = c0.x; r.m[1][0] = c0.y; r.m[2][0] = c0.z;
  r.m[0][1] = c1.x; r.m[1][1] = c1.y; r.m[2][1] = c1.z;
  r.m[0][2] = c2.x; r.m[1][2] = c2.y; r.m[2][2] = c2.z;
  return r;
}
template <class T> inline M3<T> mdiag(V3<T> d) {
  M3<T> r{};
  r.m[0][0] = d.x; r.m[1][1] = d.y; r.m[2][2] = d.z;
  return r;
}
template <class T> inline M3<T> meye(T d) { return mdiag(v3s(d)); }
template <class T> inline M3<T> mT(const M3<T> &a) {
  M3<T> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
// mat3::mul (mat3.h:83-93): accumulate from 0 in k order
template <class T> inline M3<T> mmul(const M3<T> &a, const M3<T> &b) {
  M3<T> r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) r.m[i][j] += a.m[i][k] * b.m[k][j];
  return r;
}
// mat3::mul(vec3) (mat3.h:95-98): col0*b.x + col1*b.y + col2*b.z
template <class T> inline V3<T> mvec(const M3<T> &a, V3<T> b) {
  V3<T> c0{a.m[0][0], a.m[1][0], a.m[2][0]}, c1{a.m[0][1], a.m[1][1], a.m[2][1]},
      c2{a.m[0][2], a.m[1][2], a.m[2][2]};
  return c0 * b.x + c1 * b.y + c2 * b.z;
}
// element-wise product: mat3 operator*(mat3, mat3) (mat3.h:117-129)
template <class T> inline M3<T> mhad(const M3<T> &a, const M3<T> &b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * b.m[i][j];
  return r;
}
template <class T> inline M3<T> mscale(const M3<T> &a, T s) {
  M3<T> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s;
  return r;
}
template <class T> inline M3<T> madd(const M3<T> &a, const M3<T> &b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
  return r;
}
template <class T> inline M3<T> msub(const M3<T> &a, const M3<T> &b) {
  M3<T> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j];
  return r;
}
template <class T> inline T msum(const M3<T> &a) {  // mat3::sum (mat3.h:69-77)
  T s = 0;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) s += a.m[i][j];
  return s;
}
template <class T> inline M3<T> outer(V3<T> a, V3<T> b) { return mcols(a * b.x, a * b.y, a * b.z); }  // mat3.h:175-177
template <class T> inline V3<T> mdiagv(const M3<T> &a) { return {a.m[0][0], a.m[1][1], a.m[2][2]}; }

template <class T> inline M3<T> load_m(const T *p) {
  M3<T> r;
  for (int i = 0; i < 9; ++i) r.m[i / 3][i % 3] = p[i];
  return r;
}
template <class T> inline void store_m(T *p, const M3<T> &a) {
  for (int i = 0; i < 9; ++i) p[i] = a.m[i / 3][i % 3];
}
template <class T> inline V3<T> load_v(const T *p) { return {p[0], p[1], p[2]}; }
template <class T> inline void store_v(T *p, V3<T> a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
template <class T> inline Q4<T> load_q(const T *p) { return {p[0], p[1], p[2], p[3]}; }

// ---------------------------------------------------------------- quat (quat.h:5-97)
template <class T> inline Q4<T> qinv(Q4<T> q) { return {q.w, -q.x, -q.y, -q.z}; }
// q * v = v + 2 (w (q x v) + q x (q x v))  (quat.h:14-19)
template <class T> inline V3<T> qrot(Q4<T> q, V3<T> v) {
  V3<T> qv{q.x, q.y, q.z};
  V3<T> uv = cross(qv, v);
  V3<T> uuv = cross(qv, uv);
  return v + T(2) * (q.w * uv + uuv);
}
// quat.h:25-47
template <class T>
inline void qmul_backward(Q4<T> q, V3<T> v, V3<T> g, Q4<T> &gq, V3<T> &gv) {
  V3<T> qv{q.x, q.y, q.z};
  V3<T> uv = cross(qv, v);
  gq.w += dot(uv, g) * T(2);
  V3<T> g_uv = T(2) * q.w * g;
  V3<T> g_uuv = T(2) * g;
  V3<T> g_qv = cross(uv, g_uuv);
  g_uv = g_uv + cross(g_uuv, qv);
  g_qv += cross(v, g_uv);
  gv += g + cross(g_uv, qv);
  gq.x += g_qv.x; gq.y += g_qv.y; gq.z += g_qv.z;
}
// quat.h:89-97
template <class T> inline V3<T> xform(V3<T> p, Q4<T> q, V3<T> pt) { return p + qrot(q, pt); }
template <class T> inline V3<T> xform_inv(V3<T> p, Q4<T> q, V3<T> pt) { return qrot(qinv(q), pt - p); }
// quat.h:50-62
template <class T>
inline void xform_backward(V3<T> p, Q4<T> q, V3<T> pt, V3<T> g, V3<T> &gp, Q4<T> &gq, V3<T> &gpt) {
  (void)p;
  gp += g;
  qmul_backward(q, pt, g, gq, gpt);
}
// quat.h:64-80
template <class T>
inline void xform_inv_backward(V3<T> p, Q4<T> q, V3<T> pt, V3<T> g, V3<T> &gp, Q4<T> &gq, V3<T> &gpt) {
  Q4<T> tq{0, 0, 0, 0};
  V3<T> tp{0, 0, 0};
  qmul_backward(qinv(q), pt - p, g, tq, tp);
  Q4<T> tqi = qinv(tq);
  gq = Q4<T>{gq.w + tqi.w, gq.x + tqi.x, gq.y + tqi.y, gq.z + tqi.z};
  gp = gp - tp;
  gpt = gpt + tp;
}

// ---------------------------------------------------------------- shape.h
// length with NORM_EPS2=1e-30; the sum is formed in double as in the reference
// (float dot + double literal), shape.h:4-9
template <class T> inline T length_e30(V3<T> v) { return (T)std::sqrt((double)dot(v, v) + 1e-30); }
template <class T> inline V3<T> normalized(V3<T> v) { return v / length_e30(v); }  // shape.h:13-15
// shape.h:17-25
template <class T> inline V3<T> normalized_backward(V3<T> v, V3<T> g) {
  T doted = (T)((double)dot(v, v) + 1e-30);
  V3<T> out = (g - v * dot(v / doted, g)) * (T)(1.0 / (double)std::sqrt(doted));
  return out;
}
// shape.h:27-31
template <class T> inline void abs_backward(V3<T> gx, V3<T> &g) {
  if (gx.x < 0) g.x *= -1;
  if (gx.y < 0) g.y *= -1;
  if (gx.z < 0) g.z *= -1;
}
template <class T> inline int get_type(const T *tfsr) { return (int)std::floor(tfsr[0] + T(0.1)); }  // shape.h:33-35

// shape.h:37-61. tfsr = (type, friction, softness, round); args = (a0,a1,a2,a3)
template <class T> inline T shape_sdf(const T *tfsr, const T *args, V3<T> gx) {
  int type = get_type(tfsr);
  T sdf = 0;
  if (type == 0) {
    V3<T> q = vabs(gx) - v3(args[0], args[1], args[2]);
    sdf = length_e30(vmax(q, T(0))) + std::fmin(std::fmax(std::fmax(q.x, q.y), q.z), T(0));
  } else if (type == 1) {
    V3<T> p2 = gx;
    T r = args[0], h = args[1];
    p2.y += h / 2;
    p2.y -= std::fmin(std::fmax(p2.y, T(0)), h);
    sdf = length_e30(p2) - r;
  }
  return sdf - tfsr[3];
}
// shape.h:65-102
template <class T> inline V3<T> shape_grad(const T *tfsr, const T *args, V3<T> gx) {
  int type = get_type(tfsr);
  V3<T> grad{0, 0, 0};
  if (type == 0) {
    V3<T> q = vabs(gx) - v3(args[0], args[1], args[2]);
    T inside = std::fmax(std::fmax(q.x, q.y), q.z);
    if (inside <= 0) {
      if (q.x == inside) grad.x += 1;
      if (q.y == inside) grad.y += 1;
      if (q.z == inside) grad.z += 1;
    } else {
      grad = normalized(vmax(q, T(0)));
    }
    abs_backward(gx, grad);
  } else if (type == 1) {
    V3<T> p2 = gx;
    T h = args[1];
    p2.y += h / 2;
    p2.y -= std::fmin(std::fmax(p2.y, T(0)), h);
    return normalized(p2);
  }
  return grad;
}
// shape.h:104-151; `mode` selects the reference (capsule term dropped) or exact adjoint
template <class T> inline V3<T> shape_grad_backward(const T *tfsr, const T *args, V3<T> gx, V3<T> g_out, int mode) {
  int type = get_type(tfsr);
  V3<T> g_in{0, 0, 0};
  if (type == 0) {
    V3<T> q = vabs(gx) - v3(args[0], args[1], args[2]);
    T inside = std::fmax(std::fmax(q.x, q.y), q.z);
    g_in = g_out;
    abs_backward(gx, g_in);
    if (inside <= 0) {
      g_in = v3s(T(0));
    } else {
      V3<T> q2 = vmax(q, T(0));
      g_in = normalized_backward(q2, g_in);
      if (q.x < 0) g_in.x = 0;
      if (q.y < 0) g_in.y = 0;
      if (q.z < 0) g_in.z = 0;
      abs_backward(gx, g_in);
    }
  } else if (type == 1) {
    V3<T> p2 = gx;
    T h = args[1];
    p2.y += h / 2;
    T on_cylinder = (p2.y >= T(0) && p2.y <= h) ? T(1) : T(0);
    p2.y -= std::fmin(std::fmax(p2.y, T(0)), h);
    if (mode == ADJ_EXACT) {
      // true derivative: on the cylinder section p2.y is clamped to 0, so d p2.y / d gx.y = 0
      g_in = normalized_backward(p2, g_out);
      g_in.y *= (T(1) - on_cylinder);
    } else {
      // reference: normalized_backward(p2, grad_in) with grad_in == 0, then the
      // y component scaled by the in-range flag (shape.h:141-146)
      g_in = normalized_backward(p2, g_in);
      g_in.y *= on_cylinder;
    }
    return g_in;
  }
  return g_in;
}

// ---------------------------------------------------------------- svd.h (FP64 McAdams et al. 2011)
// Float-literal constants exactly as the reference writes them (svd.h:17-20).
static const double kGamma = (double)5.82842712474619f;
static const double kCstar = (double)0.9238795325112867f;
static const double kSstar = (double)0.3826834323650897f;
static const double kEps = 1e-6;

inline double accurate_sqrt(double x) { return x / std::sqrt(x); }  // svd.h:22
inline void cond_swap(bool c, double &X, double &Y) { double Z = X; X = c ? Y : X; Y = c ? Z : Y; }          // svd.h:24-29
inline void cond_neg_swap(bool c, double &X, double &Y) { double Z = -X; X = c ? Y : X; Y = c ? Z : Y; }   // svd.h:31-36

// SVD mode: 0 = the reference's McAdams approximate Givens (default, parity);
// 1 = exact Jacobi rotations with a renormalised quaternion. Mode 1 is a smooth
// function of A (away from repeated singular values) and is used only by the
// finite-difference checks of the adjoints, which differentiate the exact SVD.
static int g_svd_mode = 0;

// svd.h:111-124
inline void approx_givens(double a11, double a12, double a22, double &ch, double &sh) {
  if (g_svd_mode == 1) {
    // inner rotation |phi| <= pi/4 (the classical convergent Jacobi choice)
    double d = a11 - a22;
    double phi = d != 0.0 ? 0.5 * std::atan(2 * a12 / d) : (a12 >= 0 ? 0.25 : -0.25) * 3.14159265358979323846;
    ch = std::cos(0.5 * phi);
    sh = std::sin(0.5 * phi);
    return;
  }
  ch = 2 * (a11 - a22);
  sh = a12;
  bool b = kGamma * sh * sh < ch * ch;
  double w = 1.0 / std::sqrt(ch * ch + sh * sh);
  ch = b ? w * ch : kCstar;
  sh = b ? w * sh : kSstar;
}
// svd.h:126-186 ; s = {s11,s21,s22,s31,s32,s33}, q = {x,y,z,w}
inline void jacobi_conj(int x, int y, int z, double s[6], double q[4]) {
  double ch, sh;
  approx_givens(s[0], s[1], s[2], ch, sh);
  double scale = ch * ch + sh * sh;
  double a = (ch * ch - sh * sh) / scale;
  double b = (2 * sh * ch) / scale;
  double s11 = s[0], s21 = s[1], s22 = s[2], s31 = s[3], s32 = s[4], s33 = s[5];
  double n11 = a * (a * s11 + b * s21) + b * (a * s21 + b * s22);
  double n21 = a * (-b * s11 + a * s21) + b * (-b * s21 + a * s22);
  double n22 = -b * (-b * s11 + a * s21) + a * (-b * s21 + a * s22);
  double n31 = a * s31 + b * s32;
  double n32 = -b * s31 + a * s32;
  double n33 = s33;
  double tmp[3] = {q[0] * sh, q[1] * sh, q[2] * sh};
  sh *= q[3];
  q[0] *= ch; q[1] *= ch; q[2] *= ch; q[3] *= ch;
  q[z] += sh;
  q[3] -= tmp[z];
  q[x] += tmp[y];
  q[y] -= tmp[x];
  // rotate the matrix indices for the next pair (svd.h:173-185)
  s[0] = n22; s[1] = n32; s[2] = n33; s[3] = n21; s[4] = n31; s[5] = n11;
}
// svd.h:81-109 ; q = {x,y,z,w}
inline void quat_to_mat(const double q[4], double M[3][3]) {
  double w = q[3], x = q[0], y = q[1], z = q[2];
  double qxx = x * x, qyy = y * y, qzz = z * z, qxz = x * z, qxy = x * y, qyz = y * z;
  double qwx = w * x, qwy = w * y, qwz = w * z;
  M[0][0] = 1 - 2 * (qyy + qzz); M[0][1] = 2 * (qxy - qwz); M[0][2] = 2 * (qxz + qwy);
  M[1][0] = 2 * (qxy + qwz); M[1][1] = 1 - 2 * (qxx + qzz); M[1][2] = 2 * (qyz - qwx);
  M[2][0] = 2 * (qxz - qwy); M[2][1] = 2 * (qyz + qwx); M[2][2] = 1 - 2 * (qxx + qyy);
}
// svd.h:248-262
inline void qr_givens(double a1, double a2, double &ch, double &sh) {
  double rho = accurate_sqrt(a1 * a1 + a2 * a2);
  sh = rho > kEps ? a2 : 0;
  ch = std::fabs(a1) + std::fmax(rho, kEps);
  bool b = a1 < 0;
  cond_swap(b, sh, ch);
  double w = 1.0 / std::sqrt(ch * ch + sh * sh);
  ch *= w;
  sh *= w;
}

// svd.h:346-413: A = U diag(sigma) V^T, all in double
inline void svd3_d(const double A[3][3], double U[3][3], double sig[3], double V[3][3]) {
  // A^T A (svd.h:61-79)
  double ATA[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ATA[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
  double s[6] = {ATA[0][0], ATA[1][0], ATA[1][1], ATA[2][0], ATA[2][1], ATA[2][2]};
  double q[4] = {0, 0, 0, 1};
  const int sweeps = g_svd_mode == 1 ? 12 : 8;
  for (int it = 0; it < sweeps; ++it) {  // svd.h:193-210 (8 sweeps)
    jacobi_conj(0, 1, 2, s, q);
    jacobi_conj(1, 2, 0, s, q);
    jacobi_conj(2, 0, 1, s, q);
  }
  if (g_svd_mode == 1) {
    double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) q[i] /= qn;
  }
  quat_to_mat(q, V);
  double B[3][3];  // B = A V (svd.h:39-58)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] = A[i][0] * V[0][j] + A[i][1] * V[1][j] + A[i][2] * V[2][j];
  // sortSingularValues (svd.h:212-246)
  double rho1 = B[0][0] * B[0][0] + B[1][0] * B[1][0] + B[2][0] * B[2][0];
  double rho2 = B[0][1] * B[0][1] + B[1][1] * B[1][1] + B[2][1] * B[2][1];
  double rho3 = B[0][2] * B[0][2] + B[1][2] * B[1][2] + B[2][2] * B[2][2];
  bool c = rho1 < rho2;
  for (int r = 0; r < 3; ++r) { cond_neg_swap(c, B[r][0], B[r][1]); cond_neg_swap(c, V[r][0], V[r][1]); }
  cond_swap(c, rho1, rho2);
  c = rho1 < rho3;
  for (int r = 0; r < 3; ++r) { cond_neg_swap(c, B[r][0], B[r][2]); cond_neg_swap(c, V[r][0], V[r][2]); }
  cond_swap(c, rho1, rho3);
  c = rho2 < rho3;
  for (int r = 0; r < 3; ++r) { cond_neg_swap(c, B[r][1], B[r][2]); cond_neg_swap(c, V[r][1], V[r][2]); }
  // QRDecomposition (svd.h:264-344)
  double b11 = B[0][0], b12 = B[0][1], b13 = B[0][2], b21 = B[1][0], b22 = B[1][1], b23 = B[1][2];
  double b31 = B[2][0], b32 = B[2][1], b33 = B[2][2];
  double ch1, sh1, ch2, sh2, ch3, sh3, a, b;
  double r11, r12, r13, r21, r22, r23, r31, r32, r33;
  qr_givens(b11, b21, ch1, sh1);
  a = 1 - 2 * sh1 * sh1;
  b = 2 * ch1 * sh1;
  r11 = a * b11 + b * b21; r12 = a * b12 + b * b22; r13 = a * b13 + b * b23;
  r21 = -b * b11 + a * b21; r22 = -b * b12 + a * b22; r23 = -b * b13 + a * b23;
  r31 = b31; r32 = b32; r33 = b33;
  qr_givens(r11, r31, ch2, sh2);
  a = 1 - 2 * sh2 * sh2;
  b = 2 * ch2 * sh2;
  b11 = a * r11 + b * r31; b12 = a * r12 + b * r32; b13 = a * r13 + b * r33;
  b21 = r21; b22 = r22; b23 = r23;
  b31 = -b * r11 + a * r31; b32 = -b * r12 + a * r32; b33 = -b * r13 + a * r33;
  qr_givens(b22, b32, ch3, sh3);
  a = 1 - 2 * sh3 * sh3;
  b = 2 * ch3 * sh3;
  r11 = b11;
  r22 = a * b22 + b * b32;
  r33 = -b * b23 + a * b33;
  double sh12 = sh1 * sh1, sh22 = sh2 * sh2, sh32 = sh3 * sh3;
  U[0][0] = (-1 + 2 * sh12) * (-1 + 2 * sh22);
  U[0][1] = 4 * ch2 * ch3 * (-1 + 2 * sh12) * sh2 * sh3 + 2 * ch1 * sh1 * (-1 + 2 * sh32);
  U[0][2] = 4 * ch1 * ch3 * sh1 * sh3 - 2 * ch2 * (-1 + 2 * sh12) * sh2 * (-1 + 2 * sh32);
  U[1][0] = 2 * ch1 * sh1 * (1 - 2 * sh22);
  U[1][1] = -8 * ch1 * ch2 * ch3 * sh1 * sh2 * sh3 + (-1 + 2 * sh12) * (-1 + 2 * sh32);
  U[1][2] = -2 * ch3 * sh3 + 4 * sh1 * (ch3 * sh1 * sh3 + ch1 * ch2 * sh2 * (-1 + 2 * sh32));
  U[2][0] = 2 * ch2 * sh2;
  U[2][1] = 2 * ch3 * (1 - 2 * sh22) * sh3;
  U[2][2] = (-1 + 2 * sh22) * (-1 + 2 * sh32);
  sig[0] = r11; sig[1] = r22; sig[2] = r33;
}

// svd3 on a T matrix (svd.h:390-413): promote to double, demote results to T
template <class T> inline void svd3(const M3<T> &A, M3<T> &U, V3<T> &s, M3<T> &V) {
  double a[3][3], u[3][3], v[3][3], sg[3];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[i][j] = (double)A.m[i][j];
  svd3_d(a, u, sg, v);
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) { U.m[i][j] = (T)u[i][j]; V.m[i][j] = (T)v[i][j]; }
  s = V3<T>{(T)sg[0], (T)sg[1], (T)sg[2]};
}

// ---------------------------------------------------------------- integrator.cu helpers
static const double NORM_EPS = 1e-8;     // integrator.cu:11
static const double SIG_CLIP_EPS = 0.05; // integrator.cu:12

// norm(a, eps) = sqrt(dot + eps) with eps passed as float (integrator.cu:28-31)
template <class T> inline T norm_eps(V3<T> a) { return std::sqrt(dot(a, a) + (T)NORM_EPS); }

// integrator.cu:33-40
template <class T> inline V3<T> dw(const M3<T> &w0, const M3<T> &w1, int i, int j, int k) {
  return {w1.m[0][i] * w0.m[1][j] * w0.m[2][k], w0.m[0][i] * w1.m[1][j] * w0.m[2][k],
          w0.m[0][i] * w0.m[1][j] * w1.m[2][k]};
}

// integrator.cu:42-67
template <class T>
inline T von_mises(const M3<T> &F, const M3<T> &U, V3<T> s, const M3<T> &V, T yield, T mu, M3<T> &out_F) {
  V3<T> s_new = vmax(s, (T)SIG_CLIP_EPS);
  V3<T> eps = vlog(s_new);
  V3<T> eps_hat = eps - (eps.x + eps.y + eps.z) / T(3);
  T eps_hat_norm = norm_eps(eps_hat);
  T delta_gamma = eps_hat_norm - yield / (2 * mu);
  if (delta_gamma > T(0)) {
    eps = eps - (delta_gamma / eps_hat_norm) * eps_hat;
    V3<T> e = vexp(eps);
    out_F = mmul(mmul(U, mdiag(e)), mT(V));
    return e.x * e.y * e.z;
  }
  out_F = F;
  return s.x * s.y * s.z;
}

// Quadratic B-spline weights (columns = offsets 0,1,2), integrator.cu:348-352
template <class T> inline M3<T> bspline_w(V3<T> fx) {
  return mcols(T(0.5) * pow2(T(1.5) - fx), T(0.75) - pow2(fx - T(1)), T(0.5) * pow2(fx - T(0.5)));
}
// derivative weights, integrator.cu:493
template <class T> inline M3<T> bspline_w1(V3<T> fx, T inv_dx) {
  return mcols(-inv_dx * (T(1.5) - fx), inv_dx * (T(-2) * fx + T(2)), -inv_dx * (fx * T(-1) + T(0.5)));
}

struct Grid {
  int gx, gy, gz;
};
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// cast_int(x*inv_dx - 0.5) (vec3.h:98-100) with the safety clamp described in the header
template <class T> inline void base_fx(V3<T> x, T inv_dx, const Grid &g, int base[3], V3<T> &fx) {
  V3<T> t = x * inv_dx - (T)0.5;
  base[0] = clampi((int)std::floor(t.x), 0, g.gx - 3);
  base[1] = clampi((int)std::floor(t.y), 0, g.gy - 3);
  base[2] = clampi((int)std::floor(t.z), 0, g.gz - 3);
  fx = x * inv_dx - v3((T)base[0], (T)base[1], (T)base[2]);
}
inline int gidx(const Grid &g, int x, int y, int z) { return (x * g.gy + y) * g.gz + z; }  // vec3.h:201-209

template <class T> inline void atomic_add(T *p, T v) {
#pragma omp atomic
  *p += v;
}

// ---------------------------------------------------------------- kernels
// compute_svd (integrator.cu:84-100)
template <class T>
void compute_svd(int n, const T *F, const T *C, T dt, T *Ftmp, T *U, T *V, T *sig) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    M3<T> f = mmul(madd(meye(T(1)), mscale(load_m(C + 9 * p), dt)), load_m(F + 9 * p));
    M3<T> u, v;
    V3<T> s;
    svd3(f, u, s, v);
    store_m(Ftmp + 9 * p, f);
    store_m(U + 9 * p, u);
    store_m(V + 9 * p, v);
    store_v(sig + 3 * p, s);
  }
}

// clamp(float a, float eps=1e-6) (integrator.cu:102-108): argument narrowed to float
inline float clamp_sep(double a) {
  float af = (float)a, e = 1e-6f;
  return af >= 0. ? std::max(af, e) : std::min(af, -e);
}

// compute_svd_grad (integrator.cu:110-186)
template <class T>
void compute_svd_grad(int n, const T *F, const T *C, const T *U, const T *V, const T *sig, T *Ftmp_grad,
                      const T *U_grad, const T *V_grad, const T *sig_grad, T *F_grad, T *C_grad, T dt) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    M3<T> u = load_m(U + 9 * p), v = load_m(V + 9 * p), gu = load_m(U_grad + 9 * p), gv = load_m(V_grad + 9 * p);
    V3<T> sigma = load_v(sig + 3 * p), gsig = load_v(sig_grad + 3 * p);
    M3<T> sg = mdiag(sigma), gsg = mdiag(gsig);
    M3<T> vt = mT(v), ut = mT(u);
    M3<T> sigma_term = mmul(mmul(u, gsg), vt);
    double s0 = sigma.x, s1 = sigma.y, s2 = sigma.z;
    s0 = s0 * s0; s1 = s1 * s1; s2 = s2 * s2;
    M3<T> FF;
    if (sizeof(T) == sizeof(float)) {
      FF.m[0][0] = 0; FF.m[0][1] = (T)(1.0 / clamp_sep(s1 - s0)); FF.m[0][2] = (T)(1.0 / clamp_sep(s2 - s0));
      FF.m[1][0] = (T)(1.0 / clamp_sep(s0 - s1)); FF.m[1][1] = 0; FF.m[1][2] = (T)(1.0 / clamp_sep(s2 - s1));
      FF.m[2][0] = (T)(1.0 / clamp_sep(s0 - s2)); FF.m[2][1] = (T)(1.0 / clamp_sep(s1 - s2)); FF.m[2][2] = 0;
    } else {
      auto cl = [](double a) { return a >= 0. ? std::max(a, 1e-6) : std::min(a, -1e-6); };
      FF.m[0][0] = 0; FF.m[0][1] = 1.0 / cl(s1 - s0); FF.m[0][2] = 1.0 / cl(s2 - s0);
      FF.m[1][0] = 1.0 / cl(s0 - s1); FF.m[1][1] = 0; FF.m[1][2] = 1.0 / cl(s2 - s1);
      FF.m[2][0] = 1.0 / cl(s0 - s2); FF.m[2][1] = 1.0 / cl(s1 - s2); FF.m[2][2] = 0;
    }
    M3<T> u_term = mmul(mmul(u, mmul(mhad(FF, msub(mmul(ut, gu), mmul(mT(gu), u))), sg)), vt);
    M3<T> v_term = mmul(u, mmul(sg, mmul(mhad(FF, msub(mmul(vt, gv), mmul(mT(gv), v))), vt)));
    M3<T> Ft_grad = madd(madd(madd(load_m(Ftmp_grad + 9 * p), u_term), sigma_term), v_term);
    store_m(Ftmp_grad + 9 * p, Ft_grad);
    M3<T> Fp = load_m(F + 9 * p), Cp = load_m(C + 9 * p);
    M3<T> cg = madd(load_m(C_grad + 9 * p), mscale(mmul(Ft_grad, mT(Fp)), dt));
    store_m(C_grad + 9 * p, cg);
    M3<T> fg = madd(load_m(F_grad + 9 * p), mmul(mT(madd(meye(T(1)), mscale(Cp, dt))), Ft_grad));
    store_m(F_grad + 9 * p, fg);
  }
}

// compute_dist (integrator.cu:188-237); bodies: pos[nb*3], rot[nb*4], tfsr[nb*4], args[nb*4]
template <class T>
void compute_dist(int n, const T *x, const T *bpos, const T *brot, const T *tfsr, const T *args, T *dist, int nb,
                  T *x_grad, T *bpos_grad, T *brot_grad, const T *dist_grad, int compute_grad) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    for (int b = 0; b < nb; ++b) {
      V3<T> bx = load_v(bpos + 3 * b);
      Q4<T> bq = load_q(brot + 4 * b);
      V3<T> gx_b = xform_inv(bx, bq, load_v(x + 3 * p));
      if (!compute_grad) {
        dist[p * nb + b] = shape_sdf(tfsr + 4 * b, args + 4 * b, gx_b);
      } else {
        T gd = dist_grad[p * nb + b];
        V3<T> g_gxb = shape_grad(tfsr + 4 * b, args + 4 * b, gx_b) * gd;
        V3<T> g_bx{0, 0, 0}, g_gx{0, 0, 0};
        Q4<T> g_bq{0, 0, 0, 0};
        xform_inv_backward(bx, bq, load_v(x + 3 * p), g_gxb, g_bx, g_bq, g_gx);
        for (int c = 0; c < 3; ++c) x_grad[3 * p + c] += (&g_gx.x)[c];
        for (int c = 0; c < 3; ++c) atomic_add(&bpos_grad[3 * b + c], (&g_bx.x)[c]);
        atomic_add(&brot_grad[4 * b + 0], g_bq.w);
        atomic_add(&brot_grad[4 * b + 1], g_bq.x);
        atomic_add(&brot_grad[4 * b + 2], g_bq.y);
        atomic_add(&brot_grad[4 * b + 3], g_bq.z);
      }
    }
  }
}

// particle2mass (integrator.cu:239-310)
template <class T>
void particle2mass(int n, const T *x, const T *pm, const Grid &g, T dx, T inv_dx, T *grid_m, const T *grid_m_grad,
                   T *x_grad, const int *ids, int id, int compute_grad) {
  (void)dx;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    if (id != -1 && ids[p] != id) continue;
    V3<T> xp = load_v(x + 3 * p);
    int base[3];
    V3<T> fx;
    base_fx(xp, inv_dx, g, base, fx);
    M3<T> w = bspline_w(fx);
    if (compute_grad) {
      M3<T> w1 = bspline_w1(fx, inv_dx);
      V3<T> gxp{0, 0, 0};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            int idx = gidx(g, base[0] + i, base[1] + j, base[2] + k);
            gxp += grid_m_grad[idx] * pm[p] * dw(w, w1, i, j, k);
          }
      for (int c = 0; c < 3; ++c) x_grad[3 * p + c] += (&gxp.x)[c];
    } else {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            T weight = w.m[0][i] * w.m[1][j] * w.m[2][k];
            int idx = gidx(g, base[0] + i, base[1] + j, base[2] + k);
            atomic_add(&grid_m[idx], pm[p] * weight);
          }
    }
  }
}

// p2g (integrator.cu:312-394)
template <class T>
void p2g(int n, const T *x, const T *v, const T *pm, const T *pvol, const T *Ftmp, const T *U, const T *sig,
         const T *V, const T *C, const T *mly, const Grid &g, T dx, T inv_dx, T dt, T *F_out, T *grid_mv, T *grid_m) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    V3<T> xp = load_v(x + 3 * p);
    M3<T> u = load_m(U + 9 * p), vv = load_m(V + 9 * p), ft = load_m(Ftmp + 9 * p);
    V3<T> sigma = load_v(sig + 3 * p);
    int base[3];
    V3<T> fx;
    base_fx(xp, inv_dx, g, base, fx);
    M3<T> w = bspline_w(fx);
    M3<T> new_F;
    T mu = mly[3 * p + 0], lam = mly[3 * p + 1], yield = mly[3 * p + 2];
    T J = von_mises(ft, u, sigma, vv, yield, mu, new_F);
    store_m(F_out + 9 * p, new_F);
    M3<T> r = mmul(u, mT(vv));
    M3<T> stress = madd(mscale(mmul(msub(new_F, r), mT(new_F)), T(2) * mu), meye(lam * J * (J - 1)));
    stress = mscale(stress, -dt * pvol[p] * T(4) * inv_dx * inv_dx);
    M3<T> affine = madd(stress, mscale(load_m(C + 9 * p), pm[p]));
    V3<T> mv = pm[p] * load_v(v + 3 * p);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          T weight = w.m[0][i] * w.m[1][j] * w.m[2][k];
          V3<T> dpos = (v3((T)i, (T)j, (T)k) - fx) * dx;
          int idx = gidx(g, base[0] + i, base[1] + j, base[2] + k);
          atomic_add(&grid_m[idx], pm[p] * weight);
          V3<T> c = (mv + mvec(affine, dpos)) * weight;
          atomic_add(&grid_mv[3 * idx + 0], c.x);
          atomic_add(&grid_mv[3 * idx + 1], c.y);
          atomic_add(&grid_mv[3 * idx + 2], c.z);
        }
  }
}

// p2g_grad (integrator.cu:396-627). Accumulates (+=) into x/v/C grads of the current
// state and into the temporary F_tmp/U/V/sigma grads.
template <class T>
void p2g_grad(int n, const T *x, const T *v, const T *pm, const T *pvol, const T *Ftmp, const T *U, const T *sig,
              const T *V, const T *C, const T *mly, const Grid &g, T dx, T inv_dx, T dt, T *x_grad, T *v_grad,
              T *Ftmp_grad, T *C_grad, T *U_grad, T *sig_grad, T *V_grad, const T *Fnext_grad, const T *grid_mv_grad,
              const T *grid_m_grad) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    M3<T> u = load_m(U + 9 * p), vv = load_m(V + 9 * p);
    V3<T> sigma = load_v(sig + 3 * p);
    T mu = mly[3 * p + 0], lam = mly[3 * p + 1], yield = mly[3 * p + 2];
    T J;
    M3<T> new_F;
    V3<T> s_new = vmax(sigma, (T)SIG_CLIP_EPS);
    V3<T> eps = vlog(s_new);
    V3<T> eps_hat = eps - (eps.x + eps.y + eps.z) / T(3);
    T ehn = norm_eps(eps_hat);
    T delta_gamma = ehn - yield / (2 * mu);
    if (delta_gamma > T(0)) {
      V3<T> e = vexp(eps - (delta_gamma / ehn) * eps_hat);
      new_F = mmul(mmul(u, mdiag(e)), mT(vv));
      J = e.x * e.y * e.z;
    } else {
      new_F = load_m(Ftmp + 9 * p);
      J = sigma.x * sigma.y * sigma.z;
    }
    M3<T> r = mmul(u, mT(vv));
    T gss = -dt * inv_dx * pvol[p] * T(4) * inv_dx;
    M3<T> stress = madd(mscale(mmul(msub(new_F, r), mT(new_F)), T(2) * mu), meye(lam * J * (J - 1)));
    M3<T> affine = madd(mscale(stress, gss), mscale(load_m(C + 9 * p), pm[p]));
    T m_p = pm[p];
    V3<T> v_p = load_v(v + 3 * p);
    V3<T> xp = load_v(x + 3 * p);
    int base[3];
    V3<T> fx;
    base_fx(xp, inv_dx, g, base, fx);
    M3<T> w = bspline_w(fx), w1 = bspline_w1(fx, inv_dx);
    M3<T> g_stress{}, g_C{};
    V3<T> g_x{0, 0, 0}, g_v{0, 0, 0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          int idx = gidx(g, base[0] + i, base[1] + j, base[2] + k);
          T N = w.m[0][i] * w.m[1][j] * w.m[2][k];
          V3<T> dpos = (v3((T)i, (T)j, (T)k) - fx) * dx;
          V3<T> ogv = load_v(grid_mv_grad + 3 * idx);
          V3<T> gN = dw(w, w1, i, j, k);
          M3<T> tmp = outer(ogv, dpos);
          g_stress = madd(g_stress, mscale(tmp, N * gss));
          g_C = madd(g_C, mscale(tmp, N * m_p));
          T gm = grid_m_grad[idx];
          g_v += N * m_p * ogv;
          g_x += gm * m_p * gN;
          g_x += dot(v_p, ogv) * m_p * gN;
          g_x += T(-1) * N * mvec(mT(affine), ogv) + dot(mvec(affine, dpos), ogv) * gN;
        }
    for (int c = 0; c < 3; ++c) {
      x_grad[3 * p + c] += (&g_x.x)[c];
      v_grad[3 * p + c] += (&g_v.x)[c];
    }
    store_m(C_grad + 9 * p, madd(load_m(C_grad + 9 * p), g_C));

    M3<T> g_r = mscale(mmul(g_stress, new_F), T(-2) * mu);
    M3<T> g_U = mmul(g_r, vv);
    M3<T> g_V = mmul(mT(g_r), u);
    M3<T> g_newF = madd(load_m(Fnext_grad + 9 * p),
                        mscale(madd(mmul(mT(g_stress), msub(new_F, r)), mmul(g_stress, new_F)), T(2) * mu));
    T g_J = ((2 * J - 1) * lam) * (g_stress.m[0][0] + g_stress.m[1][1] + g_stress.m[2][2]);
    V3<T> g_sig{0, 0, 0};
    M3<T> g_F{};
    if (delta_gamma > T(0)) {
      V3<T> e = vexp(eps - (delta_gamma / ehn) * eps_hat);
      M3<T> em = mdiag(e);
      g_U = madd(g_U, mmul(mmul(g_newF, vv), em));
      g_V = madd(g_V, mmul(mmul(mT(g_newF), u), em));
      V3<T> Fpart = mdiagv(mmul(mmul(mT(u), g_newF), vv));
      V3<T> Jpart{g_J * e.y * e.z, g_J * e.x * e.z, g_J * e.x * e.y};
      V3<T> g_eps = e * (Jpart + Fpart);
      V3<T> g_eh = -delta_gamma / ehn * g_eps;
      // formed in T, stored in a double (integrator.cu:585)
      double g_ehn = -dot(eps_hat / ehn, g_eps) * (yield / (2 * mu)) / ehn;
      g_eh += eps_hat / ehn * (T)g_ehn;
      g_eps += g_eh - (T)((double)vsum(g_eh) / 3.);
      if (sigma.x >= (T)SIG_CLIP_EPS) g_sig.x += g_eps.x / sigma.x;
      if (sigma.y >= (T)SIG_CLIP_EPS) g_sig.y += g_eps.y / sigma.y;
      if (sigma.z >= (T)SIG_CLIP_EPS) g_sig.z += g_eps.z / sigma.z;
    } else {
      g_sig += V3<T>{g_J * sigma.y * sigma.z, g_J * sigma.x * sigma.z, g_J * sigma.x * sigma.y};
      g_F = madd(g_F, g_newF);
    }
    store_m(U_grad + 9 * p, madd(load_m(U_grad + 9 * p), g_U));
    store_m(V_grad + 9 * p, madd(load_m(V_grad + 9 * p), g_V));
    for (int c = 0; c < 3; ++c) sig_grad[3 * p + c] += (&g_sig.x)[c];
    store_m(Ftmp_grad + 9 * p, madd(load_m(Ftmp_grad + 9 * p), g_F));
  }
}

// Contact of one grid node with one body (integrator.cu:694-730). Returns whether
// the contact branch was taken.
template <class T>
inline bool contact_fwd(V3<T> &v_out, V3<T> gx, V3<T> bx, Q4<T> bq, V3<T> nbx, Q4<T> nbq, const T *tfsr,
                        const T *args, T dt) {
  T friction = tfsr[1], softness = tfsr[2];
  V3<T> gx_b = xform_inv(bx, bq, gx);
  T dist = shape_sdf(tfsr, args, gx_b);
  T influence = std::min(std::exp(-dist * softness), T(1));
  if ((softness > 0 && (double)influence > 0.1) || dist <= 0) {
    V3<T> normal = normalized(shape_grad(tfsr, args, gx_b));
    normal = qrot(bq, normal);
    V3<T> bv = (xform(nbx, nbq, gx_b) - gx) / dt;
    V3<T> rel_v = v_out - bv;
    T nc = dot(rel_v, normal);
    V3<T> vt = rel_v - std::min(nc, T(0)) * normal;
    if (nc < 0. && (double)dot(vt, vt) > 1e-30) {
      T vtn = length_e30(vt);
      vt = vt * (T(1) / vtn) * std::max(T(0), vtn + nc * friction);
    }
    v_out = bv + rel_v * (T(1) - influence) + vt * influence;
    return true;
  }
  return false;
}

struct GridOpArgs {
  int gx, gy, gz, nb;
  double dx, inv_dx, dt, ground_friction;
  double gravity[3];
};

// grid_op_v2 (integrator.cu:647-777). `tape` may be null; when given it has
// G*(nb+1)*3 entries laid out as in the reference (cell*(nb+1)+slot).
template <class T>
void grid_op(const T *grid_m, const T *grid_mv, T *tape, const T *bpos, const T *brot, const T *nbpos,
             const T *nbrot, const T *tfsr, const T *args, T *grid_v_out, const Grid &g, int nb, T dx, T dt, T gf,
             V3<T> gravity) {
  const int G = g.gx * g.gy * g.gz;
#pragma omp parallel for schedule(static)
  for (int tid = 0; tid < G; ++tid) {
    if (!((double)grid_m[tid] > 1e-12)) continue;
    int gxi = tid / g.gz / g.gy, gyi = (tid / g.gz) % g.gy, gzi = tid % g.gz;
    T m = grid_m[tid];
    V3<T> mv = load_v(grid_mv + 3 * tid);
    V3<T> v_out = mv * (T(1) / m);
    v_out += dt * gravity;
    V3<T> gxw = v3((T)gxi, (T)gyi, (T)gzi) * dx;
    for (int b = 0; b < nb; ++b) {
      if (tape) store_v(tape + 3 * ((size_t)tid * (nb + 1) + b), v_out);
      contact_fwd(v_out, gxw, load_v(bpos + 3 * b), load_q(brot + 4 * b), load_v(nbpos + 3 * b),
                  load_q(nbrot + 4 * b), tfsr + 4 * b, args + 4 * b, dt);
    }
    if (tape) store_v(tape + 3 * ((size_t)tid * (nb + 1) + nb), v_out);
    const int bound = 3;
    if (gxi < bound && v_out.x < 0) v_out.x = 0;
    if (gxi > g.gx - bound && v_out.x > 0) v_out.x = 0;
    if (gyi < bound && v_out.y < 0) {
      if (gf > T(0)) {
        if (gf < T(99)) {
          T lin = v_out.y;
          V3<T> vit{v_out.x, 0, v_out.z};
          T lit = norm_eps(vit);
          v_out = vit * std::fmax((T)(1. + (double)(gf * lin / lit)), T(0));
        } else {
          v_out = v3s(T(0));
        }
      }
      v_out.y = 0;
    }
    if (gyi > g.gy - bound && v_out.y > 0) v_out.y = 0;
    if (gzi < bound && v_out.z < 0) v_out.z = 0;
    if (gzi > g.gz - bound && v_out.z > 0) v_out.z = 0;
    store_v(grid_v_out + 3 * tid, v_out);
  }
}

// grid_op_v2_grad (integrator.cu:779-1057). The body-velocity tape of the reference
// is recomputed here from the node input, which is the same sequence of operations.
template <class T>
void grid_op_grad(const T *grid_m, const T *grid_mv, const T *bpos, const T *brot, const T *nbpos, const T *nbrot,
                  const T *tfsr, const T *args, T *grid_m_grad, T *grid_mv_grad, T *bpos_grad, T *brot_grad,
                  T *nbpos_grad, T *nbrot_grad, const T *grid_v_out_grad, const Grid &g, int nb, T dx, T dt, T gf,
                  V3<T> gravity, int mode) {
  const int G = g.gx * g.gy * g.gz;
#pragma omp parallel for schedule(static)
  for (int tid = 0; tid < G; ++tid) {
    if (!((double)grid_m[tid] > 1e-12)) continue;
    int gxi = tid / g.gz / g.gy, gyi = (tid / g.gz) % g.gy, gzi = tid % g.gz;
    T m = grid_m[tid];
    V3<T> mv = load_v(grid_mv + 3 * tid);
    V3<T> gxw = v3((T)gxi, (T)gyi, (T)gzi) * dx;
    // forward recompute of the tape (integrator.cu:678-732)
    V3<T> tape[257];
    V3<T> v_cur = mv * (T(1) / m);
    v_cur += dt * gravity;
    for (int b = 0; b < nb; ++b) {
      tape[b] = v_cur;
      contact_fwd(v_cur, gxw, load_v(bpos + 3 * b), load_q(brot + 4 * b), load_v(nbpos + 3 * b),
                  load_q(nbrot + 4 * b), tfsr + 4 * b, args + 4 * b, dt);
    }
    tape[nb] = v_cur;

    V3<T> grad_v = load_v(grid_v_out_grad + 3 * tid);
    V3<T> vv = tape[nb];
    V3<T> v_in = vv;
    const int bound = 3;
    if (gxi > g.gx - bound && vv.x > 0) v_in.x = 0;
    if (gxi < bound && vv.x < 0) v_in.x = 0;
    T lin = 0, lit = 1;
    V3<T> vit{0, 0, 0};
    bool hit_ground = gyi < bound && v_in.y < 0;
    if (hit_ground) {
      lin = v_in.y;
      vit = V3<T>{v_in.x, 0, v_in.z};
      lit = norm_eps(vit);
      T flag = (T)(1. + (double)(gf * lin / lit));
      v_in = vit * std::fmax(flag, T(0));
    }
    if (gzi > g.gz - bound && v_in.z > 0) grad_v.z = 0;
    if (gzi < bound && v_in.z < 0) grad_v.z = 0;
    if (gyi > g.gy - bound && v_in.y > 0) grad_v.y = 0;
    if (hit_ground) {
      grad_v.y = 0;
      T flag = (T)(1. + (double)(gf * lin / lit));
      if (mode == ADJ_EXACT && gf >= T(99)) {
        grad_v = v3s(T(0));  // sticky ground: the forward zeroes the velocity
      } else if (flag >= 0.) {
        V3<T> g_vit = flag * grad_v;
        T g_lin = gf / lit * dot(vit, grad_v);
        T g_lit = -gf * lin / lit / lit * dot(vit, grad_v);
        g_vit += g_lit * (vit / lit);
        grad_v = V3<T>{g_vit.x, g_lin, g_vit.z};
      } else {
        grad_v = v3s(T(0));
      }
    }
    if (gxi > g.gx - bound && vv.x > 0) grad_v.x = 0;
    if (gxi < bound && vv.x < 0) grad_v.x = 0;

    for (int b = nb - 1; b >= 0; --b) {
      V3<T> v_o = tape[b];
      V3<T> bx = load_v(bpos + 3 * b);
      Q4<T> bq = load_q(brot + 4 * b);
      V3<T> nbx = load_v(nbpos + 3 * b);
      Q4<T> nbq = load_q(nbrot + 4 * b);
      const T *tf = tfsr + 4 * b;
      const T *ar = args + 4 * b;
      T friction = tf[1], softness = tf[2];
      V3<T> gx_b = xform_inv(bx, bq, gxw);
      T dist = shape_sdf(tf, ar, gx_b);
      T influence = std::min(std::exp(-dist * softness), T(1));
      if ((softness > 0 && (double)influence > 0.1) || dist <= 0) {
        V3<T> g_gxb{0, 0, 0}, g_bx{0, 0, 0};
        Q4<T> g_bq{0, 0, 0, 0};
        T g_infl = 0;
        V3<T> un = shape_grad(tf, ar, gx_b);
        V3<T> rn = normalized(un);
        V3<T> normal = qrot(bq, rn);
        V3<T> bv = (xform(nbx, nbq, gx_b) - gxw) / dt;
        V3<T> rel_v = v_o - bv;
        T nc = dot(rel_v, normal);
        V3<T> vt_in = rel_v - std::min(nc, T(0)) * normal;
        bool has_friction = nc < 0 && (double)dot(vt_in, vt_in) > 1e-30;
        V3<T> vt = vt_in;
        T vtn = length_e30(vt_in);
        if (has_friction) vt = vt_in * (T(1) / vtn) * std::max(vtn + nc * friction, T(0));
        T g_nc = 0;
        V3<T> g_bv = grad_v, g_rel = grad_v * (T(1) - influence), g_vt = grad_v * influence;
        g_infl += dot(vt - rel_v, grad_v);
        if (has_friction) {
          T bf = vtn + nc * friction;
          if (bf > 0.) {
            g_nc += dot(vt_in, g_vt) * friction / vtn;
            T g_vtn = -nc * g_nc / vtn;
            g_vt = g_vt * (T)(1. / (double)vtn) * bf + g_vtn * vt_in / vtn;
          } else {
            g_vt = v3s(T(0));
          }
        }
        V3<T> g_normal{0, 0, 0};
        g_rel += g_vt;
        if (nc < 0.) {
          g_nc += -dot(normal, g_vt);
          g_normal += -nc * g_vt;
        }
        g_rel += normal * g_nc;
        g_normal += rel_v * g_nc;
        grad_v = g_rel;
        g_bv = g_bv - g_rel;
        V3<T> g_npos{0, 0, 0};
        Q4<T> g_nrot{0, 0, 0, 0};
        xform_backward(nbx, nbq, gx_b, g_bv * (T(1) / dt), g_npos, g_nrot, g_gxb);
        for (int c = 0; c < 3; ++c) atomic_add(&nbpos_grad[3 * b + c], (&g_npos.x)[c]);
        atomic_add(&nbrot_grad[4 * b + 0], g_nrot.w);
        atomic_add(&nbrot_grad[4 * b + 1], g_nrot.x);
        atomic_add(&nbrot_grad[4 * b + 2], g_nrot.y);
        atomic_add(&nbrot_grad[4 * b + 3], g_nrot.z);
        V3<T> g_rn{0, 0, 0};
        qmul_backward(bq, rn, g_normal, g_bq, g_rn);
        g_gxb += shape_grad_backward(tf, ar, gx_b, normalized_backward(un, g_rn), mode);
        T expdist = std::exp(-dist * softness);
        if (expdist <= 1) {
          T g_dist = -softness * expdist * g_infl;
          g_gxb += un * g_dist;
        }
        V3<T> g_gx_tmp{0, 0, 0};
        xform_inv_backward(bx, bq, gxw, g_gxb, g_bx, g_bq, g_gx_tmp);
        for (int c = 0; c < 3; ++c) atomic_add(&bpos_grad[3 * b + c], (&g_bx.x)[c]);
        atomic_add(&brot_grad[4 * b + 0], g_bq.w);
        atomic_add(&brot_grad[4 * b + 1], g_bq.x);
        atomic_add(&brot_grad[4 * b + 2], g_bq.y);
        atomic_add(&brot_grad[4 * b + 3], g_bq.z);
      }
    }
    V3<T> gmv = grad_v * (T)(1. / (double)m);
    for (int c = 0; c < 3; ++c) grid_mv_grad[3 * tid + c] += (&gmv.x)[c];
    grid_m_grad[tid] += (T(-1) / m / m) * dot(mv, grad_v);
  }
}

// g2p (integrator.cu:1059-1109)
template <class T>
void g2p(int n, const T *x, const T *grid_v, const Grid &g, T dx, T inv_dx, T dt, T ground_height, T *v_out, T *C_out,
         T *x_out) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    V3<T> xp = load_v(x + 3 * p);
    int base[3];
    V3<T> fx;
    base_fx(xp, inv_dx, g, base, fx);
    M3<T> w = bspline_w(fx);
    V3<T> nv{0, 0, 0};
    M3<T> nC{};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          T weight = w.m[0][i] * w.m[1][j] * w.m[2][k];
          V3<T> dpos = v3((T)i, (T)j, (T)k) - fx;
          V3<T> gv = load_v(grid_v + 3 * gidx(g, base[0] + i, base[1] + j, base[2] + k));
          nv = nv + gv * weight;
          nC = madd(nC, mscale(outer(gv, dpos), weight * inv_dx * T(4)));
        }
    V3<T> upper = (v3((T)g.gx, (T)g.gy, (T)g.gz) - T(3)) * dx;
    V3<T> t = xp + nv * dt;
    V3<T> nx{std::fmax(std::fmin(t.x, upper.x), ground_height * dx), std::fmax(std::fmin(t.y, upper.y), ground_height * dx),
             std::fmax(std::fmin(t.z, upper.z), ground_height * dx)};
    store_v(x_out + 3 * p, nx);
    store_v(v_out + 3 * p, nv);
    store_m(C_out + 9 * p, nC);
  }
}

// g2p_grad (integrator.cu:1527-1614)
template <class T>
void g2p_grad(int n, const T *x, const T *grid_v, const Grid &g, T dx, T inv_dx, T dt, T ground_height,
              const T *v_next, T *x_grad, T *grid_v_grad, const T *v_next_grad, const T *C_next_grad,
              const T *x_next_grad) {
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    V3<T> xp = load_v(x + 3 * p);
    V3<T> g_x = load_v(x_next_grad + 3 * p);
    V3<T> g_nv = load_v(v_next_grad + 3 * p);
    M3<T> g_nC = load_m(C_next_grad + 9 * p);
    V3<T> nx = xp + load_v(v_next + 3 * p) * dt;
    V3<T> upper = (v3((T)g.gx, (T)g.gy, (T)g.gz) - T(3)) * dx;
    T lower_limit = ground_height * dx;
    if (nx.x > upper.x || nx.x < lower_limit) g_x.x = 0;
    if (nx.y > upper.y || nx.y < lower_limit) g_x.y = 0;
    if (nx.z > upper.z || nx.z < lower_limit) g_x.z = 0;
    g_nv += g_x * dt;
    int base[3];
    V3<T> fx;
    base_fx(xp, inv_dx, g, base, fx);
    M3<T> w = bspline_w(fx), w1 = bspline_w1(fx, inv_dx);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          T weight = w.m[0][i] * w.m[1][j] * w.m[2][k];
          V3<T> dpos = v3((T)i, (T)j, (T)k) - fx;
          int idx = gidx(g, base[0] + i, base[1] + j, base[2] + k);
          V3<T> gv = load_v(grid_v + 3 * idx);
          T xx = weight * inv_dx * T(4);
          V3<T> ggv = weight * g_nv + mvec(g_nC, dpos) * xx;
          atomic_add(&grid_v_grad[3 * idx + 0], ggv.x);
          atomic_add(&grid_v_grad[3 * idx + 1], ggv.y);
          atomic_add(&grid_v_grad[3 * idx + 2], ggv.z);
          g_x += -inv_dx * mvec(mT(g_nC), gv) * xx;
          T gw = dot(g_nv, gv) + (inv_dx * T(4)) * msum(mhad(outer(gv, dpos), g_nC));
          g_x += dw(w, w1, i, j, k) * gw;
        }
    for (int c = 0; c < 3; ++c) x_grad[3 * p + c] += (&g_x.x)[c];
  }
}


// ======================================================================== renderer
// CPU restatement of the renderer (FP32): particle_sdf (bake_volume, copy_volume and two
// smooth passes, integrator.cu:1438-1525,1908-1933) and render (ray_aabb_intersection,
// next_hit, render, integrator.cu:1111-1436; volume queries array.h:6-75). The reference
// draws its jitter from curand's XORWOW stream per pixel (integrator.cu:1306); this
// restatement, like the HIP renderer it checks, draws from Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) with one stream per
// (pixel, sample): draw k of sample s of pixel p is word k % 4 of
// Philox(counter = (k / 4, p, s, 0), key = (seed, 0)), mapped to (0, 1] as
// ((w >> 8) + 1) 2^-24. cos/sin of the bounce angle: Taylor series after an exact
// reduction (see sincos_taylor). Samples are traced in order and added as the reference's
// loop adds them.
namespace rnd {

inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t o0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t o1 = (uint32_t)p1;
    uint32_t o2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t o3 = (uint32_t)p0;
    c[0] = o0; c[1] = o1; c[2] = o2; c[3] = o3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct Stream {
  uint32_t seed, pixel, sample, block = 0, used = 4, w[4] = {0, 0, 0, 0};
  float uniform() {
    if (used == 4) {
      w[0] = block++; w[1] = pixel; w[2] = sample; w[3] = 0;
      philox4x32_10(w, seed, 0u);
      used = 0;
    }
    uint32_t x = w[used++];
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
  }
};

// sin and cos of phi in [0, 2 pi]: k = round(phi 2/pi), r = phi - k pi/2 with pi/2 =
// 1.5703125 + 4.83826794e-4 in two fused steps, then Taylor polynomials (sin to r^9, cos
// to r^10) in Horner form with fused multiply-adds; quadrant by k mod 4.
inline void sincos_taylor(float phi, float &s, float &c) {
  float k = std::rint(phi * 0.636619772f);
  float r = std::fmaf(-k, 1.5703125f, phi);
  r = std::fmaf(-k, 4.83826794e-4f, r);
  float r2 = r * r;
  float ps = std::fmaf(r2, 2.75573192e-6f, -1.98412698e-4f);
  ps = std::fmaf(r2, ps, 8.33333377e-3f);
  ps = std::fmaf(r2, ps, -1.66666672e-1f);
  float sn = std::fmaf(r * r2, ps, r);
  float pc = std::fmaf(r2, -2.75573188e-7f, 2.48015876e-5f);
  pc = std::fmaf(r2, pc, -1.38888892e-3f);
  pc = std::fmaf(r2, pc, 4.16666679e-2f);
  pc = std::fmaf(r2, pc, -0.5f);
  float cs = std::fmaf(r2, pc, 1.0f);
  switch (((int)k) & 3) {
    case 0: s = sn; c = cs; break;
    case 1: s = cs; c = -sn; break;
    case 2: s = -sn; c = -cs; break;
    default: s = -cs; c = sn; break;
  }
}

}  // namespace rnd

namespace rdr {
using F3 = V3<float>;

struct Vol {
  const float *sdf, *color;
  F3 lo, hi;
  int d[3];
};

// volume_query (array.h:6-49) for C channels
inline bool volume_query(const Vol &V, const float *values, int C, F3 point, float *out) {
  F3 ratio{(point.x - V.lo.x) / (V.hi.x - V.lo.x), (point.y - V.lo.y) / (V.hi.y - V.lo.y),
           (point.z - V.lo.z) / (V.hi.z - V.lo.z)};
  if (!(ratio.x >= 0 && ratio.y >= 0 && ratio.z >= 0 && ratio.x < 1. && ratio.y < 1. && ratio.z < 1.)) return false;
  F3 grid{ratio.x * (float)V.d[0], ratio.y * (float)V.d[1], ratio.z * (float)V.d[2]};
  int x0 = std::min(std::max((int)std::floor(grid.x), 0), V.d[0] - 1);
  int y0 = std::min(std::max((int)std::floor(grid.y), 0), V.d[1] - 1);
  int z0 = std::min(std::max((int)std::floor(grid.z), 0), V.d[2] - 1);
  float dx = grid.x - (float)x0, dy = grid.y - (float)y0, dz = grid.z - (float)z0;
  int base = (x0 * V.d[1] + y0) * V.d[2] + z0;
  int fz = z0 + 1 < V.d[2], fy = y0 + 1 < V.d[1], fx = x0 + 1 < V.d[0];
  int idx[8];
  idx[0] = base;
  idx[1] = base + fz;
  idx[2] = base + V.d[2] * fy;
  idx[3] = idx[2] + fz;
  idx[4] = base + V.d[1] * V.d[2] * fx;
  idx[5] = idx[4] + fz;
  idx[6] = idx[4] + V.d[2] * fy;
  idx[7] = idx[6] + fz;
  float wx[2] = {1 - dx, dx}, wy[2] = {1 - dy, dy}, wz[2] = {1 - dz, dz};
  for (int c = 0; c < C; ++c) {
    float acc = 0.f;
    for (int t = 0; t < 8; ++t) {
      float v = values[C * idx[t] + c] * wx[t >> 2] * wy[(t >> 1) & 1] * wz[t & 1];
      acc = t == 0 ? v : acc + v;
    }
    out[c] = acc;
  }
  return true;
}
inline float sample_sdf(const Vol &V, F3 p, float thr) {  // array.h:51-59
  float out = 0.f;
  if (volume_query(V, V.sdf, 1, p, &out)) out -= thr;
  return out;
}
inline F3 sample_normal(const Vol &V, F3 p) {  // array.h:61-70
  float e = 1e-3f;
  float x0 = sample_sdf(V, p + F3{-e, 0.f, 0.f}, 0.f), x1 = sample_sdf(V, p + F3{e, 0.f, 0.f}, 0.f);
  float y0 = sample_sdf(V, p + F3{0.f, -e, 0.f}, 0.f), y1 = sample_sdf(V, p + F3{0.f, e, 0.f}, 0.f);
  float z0 = sample_sdf(V, p + F3{0.f, 0.f, -e}, 0.f), z1 = sample_sdf(V, p + F3{0.f, 0.f, e}, 0.f);
  return F3{x1 - x0, y1 - y0, z1 - z0} * (float)(0.5 / (double)e);
}
inline F3 sample_color(const Vol &V, F3 p) {  // array.h:72-76
  float c[3] = {0.f, 0.f, 0.f};
  volume_query(V, V.color, 3, p, c);
  return F3{c[0], c[1], c[2]};
}
// ray_aabb_intersection (integrator.cu:1111-1143)
inline bool ray_aabb(F3 lo, F3 hi, F3 o, F3 d, float &tn, float &tf) {
  float dd[3] = {d.x, d.y, d.z}, oo[3] = {o.x, o.y, o.z}, bl[3] = {lo.x, lo.y, lo.z}, bh[3] = {hi.x, hi.y, hi.z};
  tn = -1e9f;
  tf = 1e9f;
  for (int i = 0; i < 3; ++i) {
    if (dd[i] == 0) {
      if (oo[i] < bl[i] || oo[i] > bh[i]) return false;
    } else {
      float i1 = (bl[i] - oo[i]) / dd[i], i2 = (bh[i] - oo[i]) / dd[i];
      tf = std::min(std::max(i1, i2), tf);
      tn = std::max(std::min(i1, i2), tn);
    }
  }
  return !(tn > tf);
}

struct Scene {
  Vol V;
  float thr, ground;
  int nb;
  bool shape;
  const float *pos, *rot, *tfsr, *args;
};

// next_hit (integrator.cu:1145-1267)
inline float next_hit(const Scene &S, F3 o, F3 d, F3 &color, F3 &normal, int &material) {
  const float limit = 100.f;
  float closest = 1e9f;
  if (d.y < 0) {
    float gd = (o.y - S.ground) / (-d.y);
    if (gd < limit && gd < closest) {
      closest = gd;
      normal = F3{0.f, 1.f, 0.f};
      color = F3{0.3f, 0.5f, 0.7f};
      F3 p = o + d * closest;
      if (p.x <= 1 && p.x >= 0 && p.z <= 1 && p.z >= 0)
        color = color * (float)((((int)(p.x / 0.25) + (int)(p.z / 0.25)) % 2) * 0.2 + 0.35);
      else
        color = color * 0.4f;
      material = 1;
    }
  }
  if (S.nb > 0) {
    int j = 0, id = 0;
    float dist = 0.f, val = 1e9f;
    F3 pp{0.f, 0.f, 0.f};
    while (j < 200 && dist < limit && val > 1e-8) {
      pp = o + d * dist;
      val = 1e9f;
      for (int i = 0; i < S.nb; ++i) {
        F3 gb = xform_inv(load_v(S.pos + 3 * i), load_q(S.rot + 4 * i), pp);
        float t = shape_sdf(S.tfsr + 4 * i, S.args + 4 * i, gb);
        if (t < val) {
          val = t;
          id = i;
        }
      }
      dist += val;
      j += 1;
    }
    if (dist < closest && dist < limit) {
      closest = dist;
      Q4<float> q = load_q(S.rot + 4 * id);
      F3 gb = xform_inv(load_v(S.pos + 3 * id), q, pp);
      normal = qrot(q, normalized(shape_grad(S.tfsr + 4 * id, S.args + 4 * id, gb)));
      color = F3{0.5f, 0.5f, 0.5f};
      material = 1;
    }
  }
  float tn = 0.f, tf = 0.f;
  if (S.shape && ray_aabb(S.V.lo, S.V.hi, o, d, tn, tf)) {
    tn = (float)std::max((double)tn, 0.);
    F3 pos = o + d * (float)((double)tn + 1e-4);
    F3 step{0.f, 0.f, 0.f};
    for (int j = 0; j < 500; ++j) {
      float s = sample_sdf(S.V, pos, S.thr);
      if (s < 0) {
        F3 back = step;
        for (int k = 0; k < 20; ++k) {
          back = back * 0.5f;
          if (sample_sdf(S.V, pos - back, S.thr) < 0) pos = pos - back;
        }
        float dist = std::sqrt(dot(o - pos, o - pos));
        if (dist < closest) {
          closest = dist;
          normal = normalized(sample_normal(S.V, pos));
          color = sample_color(S.V, pos);
          material = 1;
        }
        break;
      }
      step = d * (float)std::max((double)s * 0.05, 0.01);
      pos += step;
    }
  }
  return closest;
}

inline F3 matvec(const float *m, F3 b) {  // mat3::mul (mat3.h:95-98)
  F3 c0{m[0], m[3], m[6]}, c1{m[1], m[4], m[7]}, c2{m[2], m[5], m[8]};
  return c0 * b.x + c1 * b.y + c2 * b.z;
}

// one sample (integrator.cu:1315-1413 with use_directional_light = 1, noise 0.1)
inline void trace(const Scene &S, const float *K, const float *R, F3 cpos, F3 light, int depth_max, int u, int v,
                  rnd::Stream &rs, F3 &out, float &first) {
  float ju = rs.uniform();
  float jv = rs.uniform();
  F3 dir{(float)u + ju, (float)v + jv, 1.f};
  F3 contrib{0.f, 0.f, 0.f}, through{1.f, 1.f, 1.f};
  first = 0.f;
  dir = matvec(K, dir);
  F3 d = matvec(R, normalized(dir));
  F3 pos = cpos;
  int depth = 0;
  while (depth < depth_max) {
    F3 normal{0.f, 0.f, 0.f}, color{0.f, 0.f, 0.f};
    int material = 0;
    float closest = next_hit(S, pos, d, color, normal, material);
    if (depth == 0) first = closest;
    F3 hit = pos + d * closest;
    depth += 1;
    if (dot(normal, normal) > 1e-10) {
      F3 odir;
      if (material == 0) {
        odir = d - normal * (dot(d, normal) * 2.f);
      } else {
        F3 ub{1.f, 0.f, 0.f};
        if (std::fabs(normal.y) < 1 - 1e-3) ub = normalized(cross(normal, F3{0.f, 1.f, 0.f}));
        F3 vb = cross(normal, ub);
        float phi = (float)(2 * 3.14159265358979323846 * (double)rs.uniform());
        float r = rs.uniform();
        float ay = std::sqrt(r), ax = std::sqrt(1 - r);
        float sp, cp;
        rnd::sincos_taylor(phi, sp, cp);
        odir = (ub * cp + vb * sp) * ax + normal * ay;
      }
      d = normalized(odir + F3{0.f, 0.f, 0.f});
      pos = hit + d * 1e-4f;
      through = through * color;
      float n0 = rs.uniform() - 0.5f;
      float n1 = rs.uniform() - 0.5f;
      float n2 = rs.uniform() - 0.5f;
      F3 direct = normalized(light + F3{n0, n1, n2} * 0.1f);
      float dv = dot(direct, normal);
      if (dv > 0) {
        F3 c2 = color, nn = normal;
        int m2 = material;
        float dist = next_hit(S, pos, direct, c2, nn, m2);
        if (dist > 100.f) contrib += through * dv * F3{0.8f, 0.8f, 0.8f};
      }
    } else {
      break;
    }
  }
  contrib += through * F3{0.8f, 0.8f, 0.8f};
  out = contrib;
}

}  // namespace rdr

}  // namespace orc

// ======================================================================== C ABI
// Arrays are AoS as in the reference: vec3 = 3 values, mat3 = 9 values row-major,
// quat = (w,x,y,z). grid = {gx, gy, gz}. Suffix _f32 / _f64 = value type.
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

#define ORC_INSTANTIATE(T, SFX)                                                                                     \
  ORC_EXPORT void orc_svd3_##SFX(int n, const T *A, T *U, T *sig, T *V) {                                           \
    for (int p = 0; p < n; ++p) {                                                                                   \
      orc::M3<T> u, v;                                                                                              \
      orc::V3<T> s;                                                                                                 \
      orc::svd3(orc::load_m(A + 9 * p), u, s, v);                                                                   \
      orc::store_m(U + 9 * p, u);                                                                                   \
      orc::store_m(V + 9 * p, v);                                                                                   \
      orc::store_v(sig + 3 * p, s);                                                                                 \
    }                                                                                                               \
  }                                                                                                                 \
  ORC_EXPORT void orc_compute_svd_##SFX(int n, const T *F, const T *C, T dt, T *Ftmp, T *U, T *V, T *sig) {          \
    orc::compute_svd(n, F, C, dt, Ftmp, U, V, sig);                                                                 \
  }                                                                                                                 \
  ORC_EXPORT void orc_compute_svd_grad_##SFX(int n, const T *F, const T *C, const T *U, const T *V, const T *sig,   \
                                             T *Ftmp_grad, const T *U_grad, const T *V_grad, const T *sig_grad,     \
                                             T *F_grad, T *C_grad, T dt) {                                          \
    orc::compute_svd_grad(n, F, C, U, V, sig, Ftmp_grad, U_grad, V_grad, sig_grad, F_grad, C_grad, dt);             \
  }                                                                                                                 \
  ORC_EXPORT void orc_von_mises_##SFX(int n, const T *F, const T *U, const T *sig, const T *V, const T *mly,        \
                                      T *F_out, T *J_out) {                                                         \
    for (int p = 0; p < n; ++p) {                                                                                   \
      orc::M3<T> nf;                                                                                                \
      J_out[p] = orc::von_mises(orc::load_m(F + 9 * p), orc::load_m(U + 9 * p), orc::load_v(sig + 3 * p),           \
                                orc::load_m(V + 9 * p), mly[3 * p + 2], mly[3 * p + 0], nf);                        \
      orc::store_m(F_out + 9 * p, nf);                                                                              \
    }                                                                                                               \
  }                                                                                                                 \
  ORC_EXPORT void orc_p2g_##SFX(int n, const T *x, const T *v, const T *pm, const T *pvol, const T *Ftmp,            \
                                const T *U, const T *sig, const T *V, const T *C, const T *mly, const int *grid,    \
                                T dx, T inv_dx, T dt, T *F_out, T *grid_mv, T *grid_m) {                            \
    orc::Grid g{grid[0], grid[1], grid[2]};                                                                         \
    orc::p2g(n, x, v, pm, pvol, Ftmp, U, sig, V, C, mly, g, dx, inv_dx, dt, F_out, grid_mv, grid_m);                \
  }                                                                                                                 \
  ORC_EXPORT void orc_p2g_grad_##SFX(int n, const T *x, const T *v, const T *pm, const T *pvol, const T *Ftmp,       \
                                     const T *U, const T *sig, const T *V, const T *C, const T *mly,                \
                                     const int *grid, T dx, T inv_dx, T dt, T *x_grad, T *v_grad, T *Ftmp_grad,     \
                                     T *C_grad, T *U_grad, T *sig_grad, T *V_grad, const T *Fnext_grad,             \
                                     const T *grid_mv_grad, const T *grid_m_grad) {                                 \
    orc::Grid g{grid[0], grid[1], grid[2]};                                                                         \
    orc::p2g_grad(n, x, v, pm, pvol, Ftmp, U, sig, V, C, mly, g, dx, inv_dx, dt, x_grad, v_grad, Ftmp_grad, C_grad, \
                  U_grad, sig_grad, V_grad, Fnext_grad, grid_mv_grad, grid_m_grad);                                 \
  }                                                                                                                 \
  ORC_EXPORT void orc_grid_op_##SFX(const T *grid_m, const T *grid_mv, T *tape, const T *bpos, const T *brot,       \
                                    const T *nbpos, const T *nbrot, const T *tfsr, const T *args, T *grid_v_out,    \
                                    const int *grid, int nb, T dx, T dt, T gf, const T *gravity) {                  \
    orc::Grid g{grid[0], grid[1], grid[2]};                                                                         \
    orc::grid_op(grid_m, grid_mv, tape, bpos, brot, nbpos, nbrot, tfsr, args, grid_v_out, g, nb, dx, dt, gf,        \
                 orc::load_v(gravity));                                                                             \
  }                                                                                                                 \
  ORC_EXPORT void orc_grid_op_grad_##SFX(const T *grid_m, const T *grid_mv, const T *bpos, const T *brot,           \
                                         const T *nbpos, const T *nbrot, const T *tfsr, const T *args,              \
                                         T *grid_m_grad, T *grid_mv_grad, T *bpos_grad, T *brot_grad,               \
                                         T *nbpos_grad, T *nbrot_grad, const T *grid_v_out_grad, const int *grid,   \
                                         int nb, T dx, T dt, T gf, const T *gravity, int mode) {                    \
    orc::Grid g{grid[0], grid[1], grid[2]};                                                                         \
    orc::grid_op_grad(grid_m, grid_mv, bpos, brot, nbpos, nbrot, tfsr, args, grid_m_grad, grid_mv_grad, bpos_grad,  \
                      brot_grad, nbpos_grad, nbrot_grad, grid_v_out_grad, g, nb, dx, dt, gf,                        \
                      orc::load_v(gravity), mode);                                                                  \
  }                                                                                                                 \
  ORC_EXPORT void orc_g2p_##SFX(int n, const T *x, const T *grid_v, const int *grid, T dx, T inv_dx, T dt,          \
                                T ground_height, T *v_out, T *C_out, T *x_out) {                                    \
    orc::Grid g{grid[0], grid[1], grid[2]};                                                                         \
    orc::g2p(n, x, grid_v, g, dx, inv_dx, dt, ground_height, v_out, C_out, x_out);                                  \
  }                                                                                                                 \
  ORC_EXPORT void orc_g2p_grad_##SFX(int n, const T *x, const T *grid_v, const int *grid, T dx, T inv_dx, T dt,     \
                                     T ground_height, const T *v_next, T *x_grad, T *grid_v_grad,                   \
                                     const T *v_next_grad, const T *C_next_grad, const T *x_next_grad) {            \
    orc::Grid g{grid[0], grid[1], grid[2]};                                                                         \
    orc::g2p_grad(n, x, grid_v, g, dx, inv_dx, dt, ground_height, v_next, x_grad, grid_v_grad, v_next_grad,         \
                  C_next_grad, x_next_grad);                                                                        \
  }                                                                                                                 \
  ORC_EXPORT void orc_compute_dist_##SFX(int n, const T *x, const T *bpos, const T *brot, const T *tfsr,            \
                                         const T *args, T *dist, int nb, T *x_grad, T *bpos_grad, T *brot_grad,     \
                                         const T *dist_grad, int compute_grad) {                                    \
    orc::compute_dist(n, x, bpos, brot, tfsr, args, dist, nb, x_grad, bpos_grad, brot_grad, dist_grad,              \
                      compute_grad);                                                                                \
  }                                                                                                                 \
  ORC_EXPORT void orc_particle2mass_##SFX(int n, const T *x, const T *pm, const int *grid, T dx, T inv_dx,          \
                                          T *grid_m, const T *grid_m_grad, T *x_grad, const int *ids, int id,       \
                                          int compute_grad) {                                                       \
    orc::Grid g{grid[0], grid[1], grid[2]};                                                                         \
    orc::particle2mass(n, x, pm, g, dx, inv_dx, grid_m, grid_m_grad, x_grad, ids, id, compute_grad);                \
  }                                                                                                                 \
  ORC_EXPORT void orc_shape_sdf_##SFX(int n, const T *tfsr, const T *args, const T *pts, T *sdf, T *grad) {         \
    for (int p = 0; p < n; ++p) {                                                                                   \
      sdf[p] = orc::shape_sdf(tfsr, args, orc::load_v(pts + 3 * p));                                                \
      orc::store_v(grad + 3 * p, orc::shape_grad(tfsr, args, orc::load_v(pts + 3 * p)));                            \
    }                                                                                                               \
  }                                                                                                                 \
  ORC_EXPORT void orc_xform_##SFX(int n, const T *pos, const T *rot, const T *pts, T *out, int inverse) {           \
    for (int p = 0; p < n; ++p) {                                                                                   \
      orc::V3<T> r = inverse ? orc::xform_inv(orc::load_v(pos), orc::load_q(rot), orc::load_v(pts + 3 * p))         \
                             : orc::xform(orc::load_v(pos), orc::load_q(rot), orc::load_v(pts + 3 * p));            \
      orc::store_v(out + 3 * p, r);                                                                                 \
    }                                                                                                               \
  }

ORC_INSTANTIATE(float, f32)
ORC_INSTANTIATE(double, f64)


// ------------------------------------------------------------------ renderer C ABI (FP32)
ORC_EXPORT void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  orc::rnd::philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}
ORC_EXPORT void orc_uniforms(uint32_t seed, uint32_t pixel, uint32_t sample, int n, float *out) {
  orc::rnd::Stream st;
  st.seed = seed;
  st.pixel = pixel;
  st.sample = sample;
  for (int i = 0; i < n; ++i) out[i] = st.uniform();
}
ORC_EXPORT void orc_sincos(int n, const float *phi, float *s, float *c) {
  for (int i = 0; i < n; ++i) orc::rnd::sincos_taylor(phi[i], s[i], c[i]);
}
// particle_sdf (integrator.cu:1480-1525, 1463-1478, 1438-1461, 1908-1933): volume must be
// filled with (1 << 33) - 1 by the caller
ORC_EXPORT void orc_particle_sdf(int64_t *volume, const float *x, const int *color, const float *bbox_min,
                                 int bake, const int *grid, float inv_dx, float *sdf, float *rgb, float *sdf_tmp,
                                 int n) {
  using orc::V3;
  const int gx = grid[0], gy = grid[1], gz = grid[2], w = 2 * bake;
  V3<float> lo{bbox_min[0], bbox_min[1], bbox_min[2]};
  for (int p = 0; p < n; ++p)
    for (int i = -bake; i < w - bake; ++i)
      for (int j = -bake; j < w - bake; ++j)
        for (int k = -bake; k < w - bake; ++k) {
          V3<float> xs = (orc::load_v(x + 3 * p) - lo) * inv_dx;
          int ix = (int)std::floor(xs.x) + i, iy = (int)std::floor(xs.y) + j, iz = (int)std::floor(xs.z) + k;
          if (ix < 0 || iy < 0 || iz < 0 || ix >= gx || iy >= gy || iz >= gz) continue;
          V3<float> dv{(float)ix - xs.x, (float)iy - xs.y, (float)iz - xs.z};
          float dist = std::sqrt(orc::dot(dv, dv) + 0.f);
          dist = std::min(std::max(0.0f, 51.0f * dist), 255.0f);
          int64_t out = ((int64_t)dist << 24) + (int64_t)color[p];
          int64_t &cell = volume[((int64_t)ix * gy + iy) * gz + iz];
          if (out < cell) cell = out;
        }
  const int64_t dim = (int64_t)gx * gy * gz;
  for (int64_t t = 0; t < dim; ++t) {
    int64_t c = volume[t];
    float zc = (float)((double)(c & 255) / 255.);
    c >>= 8;
    float yc = (float)((double)(c & 255) / 255.);
    c >>= 8;
    float xc = (float)((double)(c & 255) / 255.);
    c >>= 8;
    rgb[3 * t] = xc;
    rgb[3 * t + 1] = yc;
    rgb[3 * t + 2] = zc;
    sdf[t] = (float)((double)(c & 255) / 255.);
  }
  for (int pass = 0; pass < 2; ++pass) {
    const float *in = pass == 0 ? sdf : sdf_tmp;
    float *out = pass == 0 ? sdf_tmp : sdf;
    for (int x0 = 0; x0 < gx; ++x0)
      for (int y0 = 0; y0 < gy; ++y0)
        for (int z0 = 0; z0 < gz; ++z0) {
          int64_t t = ((int64_t)x0 * gy + y0) * gz + z0;
          if (x0 >= 1 && y0 >= 1 && z0 >= 1 && x0 < gx - 1 && y0 < gy - 1 && z0 < gz - 1) {
            float s = 0.f;
            for (int i = -1; i < 2; ++i)
              for (int j = -1; j < 2; ++j)
                for (int k = -1; k < 2; ++k) s += in[((int64_t)(x0 + i) * gy + (y0 + j)) * gz + (z0 + k)];
            out[t] = (float)((double)s / 27.);
          } else {
            out[t] = 1.f;
          }
        }
  }
}
// render (integrator.cu:1269-1436, 1863-1906) for the pixels [p_begin, p_end) of an
// H x W image (p = row * W + column)
ORC_EXPORT void orc_render(const float *sdf, const float *box_min, const float *box_max, const float *rgb,
                           const float *bpos, const float *brot, const float *tfsr, const float *args,
                           const float *cam_rot, const float *cam_pos, const float *cam_K, float *color_out,
                           float *depth_out, float thr, const int *grid, int nb, int shape, int H, int W,
                           int max_depth, int spp, float ground, int seed, const float *light, int p_begin,
                           int p_end) {
  using orc::V3;
  orc::rdr::Scene S;
  S.V.sdf = sdf;
  S.V.color = rgb;
  S.V.lo = orc::load_v(box_min);
  S.V.hi = orc::load_v(box_max);
  S.V.d[0] = grid[0];
  S.V.d[1] = grid[1];
  S.V.d[2] = grid[2];
  S.thr = thr;
  S.ground = ground;
  S.nb = nb;
  S.shape = shape != 0;
  S.pos = bpos;
  S.rot = brot;
  S.tfsr = tfsr;
  S.args = args;
  V3<float> cp = orc::load_v(cam_pos), li = orc::load_v(light);
  for (int p = p_begin; p < p_end; ++p) {
    int u = p % W, v = p / W;
    V3<float> acc{0.f, 0.f, 0.f};
    float dacc = 0.f;
    for (int s = 0; s < spp; ++s) {
      orc::rnd::Stream rs;
      rs.seed = (uint32_t)seed;
      rs.pixel = (uint32_t)p;
      rs.sample = (uint32_t)s;
      V3<float> o;
      float first;
      orc::rdr::trace(S, cam_K, cam_rot, cp, li, max_depth, u, v, rs, o, first);
      acc += o / (float)spp;
      dacc += first / (float)spp;
    }
    float uu = (float)(1.0 * u / W), vv = (float)(1.0 * v / H);
    float rad = std::sqrt((uu - 0.5f) * (uu - 0.5f) + (vv - 0.5f) * (vv - 0.5f));
    float darken = (float)(1.0 - (double)(0.9f * std::max(rad - 0.f, 0.f)));
    V3<float> c = acc * darken * 1.5f;
    color_out[3 * p] = std::sqrt(c.x);
    color_out[3 * p + 1] = std::sqrt(c.y);
    color_out[3 * p + 2] = std::sqrt(c.z);
    depth_out[p] = dacc;
  }
}

ORC_EXPORT void orc_set_svd_mode(int mode) { orc::g_svd_mode = mode; }
ORC_EXPORT int orc_get_svd_mode(void) { return orc::g_svd_mode; }

ORC_EXPORT int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
