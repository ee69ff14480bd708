// Device-side math and per-particle / per-node physics of the differentiable MLS-MPM
// substep, for gfx950. Every function names the reference routine whose arithmetic it
// reproduces (paths relative to the DexDeform repository root). Values are FP32 as in
// the reference, with the reference's FP64 islands (SVD, SVD-adjoint denominators and a
// few double-precision constants) kept in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ddm {

#define DDM_HD __device__ __forceinline__

// ------------------------------------------------------------------ small types
struct f3 {
  float x, y, z;
};
struct q4 {
  float w, x, y, z;
};
// 3x3, row-major: a[3*r + c]
struct m33 {
  float a[9];
};

DDM_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
DDM_HD f3 splat3(float s) { return f3{s, s, s}; }
DDM_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DDM_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DDM_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
DDM_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
DDM_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
DDM_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
DDM_HD f3 operator-(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
DDM_HD f3 operator-(float s, f3 a) { return f3{s - a.x, s - a.y, s - a.z}; }
DDM_HD void operator+=(f3 &a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
DDM_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DDM_HD f3 cross3(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
DDM_HD f3 fmax3(f3 a, float b) { return f3{fmaxf(a.x, b), fmaxf(a.y, b), fmaxf(a.z, b)}; }
DDM_HD f3 fabs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }

DDM_HD m33 mzero() {
  m33 r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = 0.f;
  return r;
}
DDM_HD m33 mdiag(f3 d) {
  m33 r = mzero();
  r.a[0] = d.x; r.a[4] = d.y; r.a[8] = d.z;
  return r;
}
DDM_HD m33 mtrans(const m33 &m) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[3 * i + j] = m.a[3 * j + i];
  return r;
}
// matrix product accumulated in k order (as mat3::mul, mat3.h:83-93)
DDM_HD m33 mmul(const m33 &a, const m33 &b) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[3 * i + j] = a.a[3 * i + 0] * b.a[0 + j] + a.a[3 * i + 1] * b.a[3 + j] + a.a[3 * i + 2] * b.a[6 + j];
  return r;
}
// matrix-vector product (mat3::mul(vec3), mat3.h:95-98)
DDM_HD f3 mvec(const m33 &m, f3 v) {
  return f3{m.a[0] * v.x + m.a[1] * v.y + m.a[2] * v.z, m.a[3] * v.x + m.a[4] * v.y + m.a[5] * v.z,
            m.a[6] * v.x + m.a[7] * v.y + m.a[8] * v.z};
}
// transpose(m) * v without forming the transpose
DDM_HD f3 mtvec(const m33 &m, f3 v) {
  return f3{m.a[0] * v.x + m.a[3] * v.y + m.a[6] * v.z, m.a[1] * v.x + m.a[4] * v.y + m.a[7] * v.z,
            m.a[2] * v.x + m.a[5] * v.y + m.a[8] * v.z};
}
DDM_HD m33 madd(const m33 &a, const m33 &b) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = a.a[i] + b.a[i];
  return r;
}
DDM_HD m33 msub(const m33 &a, const m33 &b) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = a.a[i] - b.a[i];
  return r;
}
DDM_HD m33 mscale(const m33 &a, float s) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = a.a[i] * s;
  return r;
}
DDM_HD m33 mhad(const m33 &a, const m33 &b) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = a.a[i] * b.a[i];
  return r;
}
// outer(a, b)[i][j] = a_i b_j  (mat3.h:175-177)
DDM_HD m33 mouter(f3 a, f3 b) {
  return m33{{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}

DDM_HD f3 ld3(const float *p) { return f3{p[0], p[1], p[2]}; }
DDM_HD void st3(float *p, f3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
DDM_HD void add3(float *p, f3 v) { p[0] += v.x; p[1] += v.y; p[2] += v.z; }
DDM_HD m33 ld9(const float *p) {
  m33 r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.a[i] = p[i];
  return r;
}
DDM_HD void st9(float *p, const m33 &m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) p[i] = m.a[i];
}
DDM_HD void add9(float *p, const m33 &m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) p[i] += m.a[i];
}
DDM_HD q4 ldq(const float *p) { return q4{p[0], p[1], p[2], p[3]}; }
DDM_HD void stq(float *p, q4 q) { p[0] = q.w; p[1] = q.x; p[2] = q.y; p[3] = q.z; }

// ------------------------------------------------------------------ rotations (quat.h)
DDM_HD q4 qconj(q4 q) { return q4{q.w, -q.x, -q.y, -q.z}; }
// rotate v by unit quaternion q: v + 2 (w (q x v) + q x (q x v))  (quat.h:14-19)
DDM_HD f3 qrot(q4 q, f3 v) {
  f3 u{q.x, q.y, q.z};
  f3 uv = cross3(u, v);
  f3 uuv = cross3(u, uv);
  return v + 2.f * (q.w * uv + uuv);
}
// adjoint of qrot (quat.h:25-47); accumulates into gq, gv
DDM_HD void qrot_adj(q4 q, f3 v, f3 g, q4 &gq, f3 &gv) {
  f3 u{q.x, q.y, q.z};
  f3 uv = cross3(u, v);
  gq.w += dot3(uv, g) * 2.f;
  f3 g_uv = 2.f * q.w * g;
  f3 g_uuv = 2.f * g;
  f3 g_u = cross3(uv, g_uuv);
  g_uv = g_uv + cross3(g_uuv, u);
  g_u += cross3(v, g_uv);
  gv += g + cross3(g_uv, u);
  gq.x += g_u.x; gq.y += g_u.y; gq.z += g_u.z;
}
DDM_HD f3 to_world(f3 p, q4 q, f3 pt) { return p + qrot(q, pt); }        // quat.h:89-92
DDM_HD f3 to_body(f3 p, q4 q, f3 pt) { return qrot(qconj(q), pt - p); }  // quat.h:94-97
// adjoint of to_world w.r.t. (p, q, pt) (quat.h:50-62)
DDM_HD void to_world_adj(q4 q, f3 pt, f3 g, f3 &gp, q4 &gq, f3 &gpt) {
  gp += g;
  qrot_adj(q, pt, g, gq, gpt);
}
// adjoint of to_body w.r.t. (p, q, pt) (quat.h:64-80)
DDM_HD void to_body_adj(f3 p, q4 q, f3 pt, f3 g, f3 &gp, q4 &gq, f3 &gpt) {
  q4 tq{0.f, 0.f, 0.f, 0.f};
  f3 tp{0.f, 0.f, 0.f};
  qrot_adj(qconj(q), pt - p, g, tq, tp);
  gq.w += tq.w; gq.x -= tq.x; gq.y -= tq.y; gq.z -= tq.z;
  gp = gp - tp;
  gpt = gpt + tp;
}

// ------------------------------------------------------------------ primitive SDFs (shape.h)
enum { ADJ_REFERENCE = 0, ADJ_EXACT = 1 };

// |v| with 1e-30 added in double (shape.h:7-9)
DDM_HD float len30(f3 v) { return (float)sqrt((double)dot3(v, v) + 1e-30); }
DDM_HD f3 unit30(f3 v) { return v / len30(v); }  // shape.h:13-15
// adjoint of unit30 (shape.h:17-25)
DDM_HD f3 unit30_adj(f3 v, f3 g) {
  float d = (float)((double)dot3(v, v) + 1e-30);
  return (g - v * dot3(v / d, g)) * (float)(1.0 / (double)sqrtf(d));
}
DDM_HD void sign_flip(f3 x, f3 &g) {  // shape.h:27-31
  if (x.x < 0) g.x = -g.x;
  if (x.y < 0) g.y = -g.y;
  if (x.z < 0) g.z = -g.z;
}
// prim = (type, friction, softness, round), args = shape arguments (shape.h:33-35)
DDM_HD int prim_type(const float *prim) { return (int)floorf(prim[0] + 0.1f); }

// capsule: segment along body-frame y from -h/2 to h/2 (shape.h:50-56)
DDM_HD f3 capsule_offset(f3 p, float h, bool *on_cylinder) {
  p.y += h / 2;
  if (on_cylinder) *on_cylinder = (p.y >= 0.f && p.y <= h);
  p.y -= fminf(fmaxf(p.y, 0.f), h);
  return p;
}
// signed distance (shape.h:37-61)
DDM_HD float prim_sdf(const float *prim, const float *args, f3 p) {
  int t = prim_type(prim);
  float d = 0.f;
  if (t == 0) {
    f3 q = fabs3(p) - f3{args[0], args[1], args[2]};
    d = len30(fmax3(q, 0.f)) + fminf(fmaxf(fmaxf(q.x, q.y), q.z), 0.f);
  } else if (t == 1) {
    d = len30(capsule_offset(p, args[1], nullptr)) - args[0];
  }
  return d - prim[3];
}
// unnormalised SDF gradient (shape.h:65-102)
DDM_HD f3 prim_grad(const float *prim, const float *args, f3 p) {
  int t = prim_type(prim);
  f3 g{0.f, 0.f, 0.f};
  if (t == 0) {
    f3 q = fabs3(p) - f3{args[0], args[1], args[2]};
    float m = fmaxf(fmaxf(q.x, q.y), q.z);
    if (m <= 0.f) {
      if (q.x == m) g.x += 1.f;
      if (q.y == m) g.y += 1.f;
      if (q.z == m) g.z += 1.f;
    } else {
      g = unit30(fmax3(q, 0.f));
    }
    sign_flip(p, g);
  } else if (t == 1) {
    return unit30(capsule_offset(p, args[1], nullptr));
  }
  return g;
}
// adjoint of prim_grad (shape.h:104-151). ADJ_REFERENCE reproduces the reference,
// whose capsule branch back-propagates its zero-initialised buffer instead of the
// incoming gradient; ADJ_EXACT gives the true derivative.
DDM_HD f3 prim_grad_adj(const float *prim, const float *args, f3 p, f3 g_out, int mode) {
  int t = prim_type(prim);
  f3 g{0.f, 0.f, 0.f};
  if (t == 0) {
    f3 q = fabs3(p) - f3{args[0], args[1], args[2]};
    float m = fmaxf(fmaxf(q.x, q.y), q.z);
    if (m > 0.f) {
      g = g_out;
      sign_flip(p, g);
      g = unit30_adj(fmax3(q, 0.f), g);
      if (q.x < 0) g.x = 0.f;
      if (q.y < 0) g.y = 0.f;
      if (q.z < 0) g.z = 0.f;
      sign_flip(p, g);
    }
  } else if (t == 1) {
    bool on_cyl;
    f3 p2 = capsule_offset(p, args[1], &on_cyl);
    if (mode == ADJ_EXACT) {
      g = unit30_adj(p2, g_out);
      if (on_cyl) g.y = 0.f;
    } else {
      g = unit30_adj(p2, g);
      if (!on_cyl) g.y = 0.f;
    }
  }
  return g;
}

// ------------------------------------------------------------------ 3x3 SVD in FP64
// McAdams, Selle, Tamstorf, Teran, Sifakis, "Computing the Singular Value Decomposition
// of 3x3 matrices with minimal branching and elementary floating point operations"
// (2011): Jacobi eigen-analysis of A^T A with approximate Givens quaternions
// (8 sweeps), column sort by norm, Givens QR. The float-rounded constants of the
// reference (svd.h:17-20) are kept so that the same rotations are chosen.
namespace svd {
// The SVD's FP64 arithmetic contracts multiply-adds into FMAs (as nvcc does by default for
// the reference's svd.h); the factors are returned in FP32, where the rounding differences
// this makes, and those of rsq64 and of the Newton reciprocal in jacobi, do not show
// beyond FP32 rounding (the parity tests against the oracle are unchanged).
// 1/sqrt(x) in FP64: the hardware estimate (about 23 bits) refined by two Newton steps
// (about 46 bits; the reference divides by a correctly rounded sqrt)
DDM_HD double rsq64(double x) {
#pragma clang fp contract(fast)
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * (1.5 - h * (y * y));
  y = y * (1.5 - h * (y * y));
  return y;
}
constexpr double kGamma = (double)5.82842712474619f;
constexpr double kCstar = (double)0.9238795325112867f;
constexpr double kSstar = (double)0.3826834323650897f;
constexpr double kEps = 1e-6;

// one Jacobi conjugation on the symmetric matrix stored as (s11,s21,s22,s31,s32,s33),
// then rotate the storage so the next call targets the next index pair (svd.h:126-186)
template <int X, int Y, int Z>
DDM_HD void jacobi(double &s11, double &s21, double &s22, double &s31, double &s32, double &s33, double q[4]) {
#pragma clang fp contract(fast)
  double ch = 2.0 * (s11 - s22), sh = s21;
  bool b = kGamma * sh * sh < ch * ch;
  double w = rsq64(ch * ch + sh * sh);
  ch = b ? w * ch : kCstar;
  sh = b ? w * sh : kSstar;
  double scale = ch * ch + sh * sh;
  // 1 / scale (one reciprocal for both quotients, svd.h:133-135) by two Newton steps from 1:
  // scale is 1 up to rounding (normalised (ch, sh)) or 1 - 5.7e-8 (the float-rounded
  // cos/sin(pi/8)), so two steps reach FP64 rounding without a division
  double inv = 2.0 - scale;
  inv = inv * (2.0 - scale * inv);
  double a = (ch * ch - sh * sh) * inv;
  double bb = (2.0 * sh * ch) * inv;
  double n11 = a * (a * s11 + bb * s21) + bb * (a * s21 + bb * s22);
  double n21 = a * (-bb * s11 + a * s21) + bb * (-bb * s21 + a * s22);
  double n22 = -bb * (-bb * s11 + a * s21) + a * (-bb * s21 + a * s22);
  double n31 = a * s31 + bb * s32;
  double n32 = -bb * s31 + a * s32;
  double n33 = s33;
  double t0 = q[0] * sh, t1 = q[1] * sh, t2 = q[2] * sh;
  double tmp[3] = {t0, t1, t2};
  double shw = sh * q[3];
  q[0] *= ch; q[1] *= ch; q[2] *= ch; q[3] *= ch;
  q[Z] += shw;
  q[3] -= tmp[Z];
  q[X] += tmp[Y];
  q[Y] -= tmp[X];
  s11 = n22; s21 = n32; s22 = n33; s31 = n21; s32 = n31; s33 = n11;
}
DDM_HD void cswap(bool c, double &x, double &y) {
  double t = x;
  x = c ? y : x;
  y = c ? t : y;
}
DDM_HD void cnswap(bool c, double &x, double &y) {
  double t = -x;
  x = c ? y : x;
  y = c ? t : y;
}
DDM_HD void qr_rot(double a1, double a2, double &ch, double &sh) {  // svd.h:248-262
#pragma clang fp contract(fast)
  double r2 = a1 * a1 + a2 * a2;
  double rho = r2 * rsq64(r2);  // r2 / sqrt(r2); NaN for r2 = 0 as there
  sh = rho > kEps ? a2 : 0.0;
  ch = fabs(a1) + fmax(rho, kEps);
  cswap(a1 < 0.0, sh, ch);
  double w = rsq64(ch * ch + sh * sh);
  ch *= w;
  sh *= w;
}
}  // namespace svd

// q (x) r for r = (1 + s e_K) / sqrt(2), the rotation by s 90 degrees about axis K:
// right-multiplying the rotation matrix of q by that of r
template <int K>
DDM_HD q4 qmul_quarter(q4 q, float s) {
  constexpr float h = 0.70710678118654752f;
  if (K == 2) return q4{(q.w - s * q.z) * h, (q.x + s * q.y) * h, (q.y - s * q.x) * h, (q.z + s * q.w) * h};
  if (K == 1) return q4{(q.w - s * q.y) * h, (q.x - s * q.z) * h, (q.y + s * q.w) * h, (q.z + s * q.x) * h};
  return q4{(q.w - s * q.x) * h, (q.x + s * q.w) * h, (q.y + s * q.z) * h, (q.z - s * q.y) * h};
}
DDM_HD q4 qsel(bool c, q4 a, q4 b) { return c ? a : b; }
// Rotation matrix of a unit quaternion (w, x, y, z), in the form svd3 builds V with
DDM_HD m33 quat_mat(q4 q) {
  const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z, xz = q.x * q.z, xy = q.x * q.y, yz = q.y * q.z,
              wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
  return m33{{1.f - 2.f * (yy + zz), 2.f * (xy - wz), 2.f * (xz + wy), 2.f * (xy + wz), 1.f - 2.f * (xx + zz),
              2.f * (yz - wx), 2.f * (xz - wy), 2.f * (yz + wx), 1.f - 2.f * (xx + yy)}};
}

// F = U diag(sig) V^T ; F given in FP32, factors returned in FP32 (svd.h:390-413).
// Also returns U and V as quaternions (qu, qv; quat_mat(qu) = U and quat_mat(qv) = V to
// FP32 rounding), the compact form the adjoint's tape keeps: V is the Jacobi quaternion
// times a quarter turn per column swap of the sort (a swap with one sign flip is a
// rotation by +-90 degrees about the third axis), U = Rz Ry Rx of the three Givens
// rotations, whose half-angle cosines and sines the QR already has.
DDM_HD void svd3(const m33 &F, m33 &U, f3 &sig, m33 &V, q4 &qu, q4 &qv) {
#pragma clang fp contract(fast)
  using namespace svd;
  double a[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = (double)F.a[i];
  // A^T A
  double s11 = a[0] * a[0] + a[3] * a[3] + a[6] * a[6];
  double s21 = a[1] * a[0] + a[4] * a[3] + a[7] * a[6];
  double s22 = a[1] * a[1] + a[4] * a[4] + a[7] * a[7];
  double s31 = a[2] * a[0] + a[5] * a[3] + a[8] * a[6];
  double s32 = a[2] * a[1] + a[5] * a[4] + a[8] * a[7];
  double s33 = a[2] * a[2] + a[5] * a[5] + a[8] * a[8];
  double q[4] = {0.0, 0.0, 0.0, 1.0};  // (x, y, z, w)
  for (int it = 0; it < 8; ++it) {
    jacobi<0, 1, 2>(s11, s21, s22, s31, s32, s33, q);
    jacobi<1, 2, 0>(s11, s21, s22, s31, s32, s33, q);
    jacobi<2, 0, 1>(s11, s21, s22, s31, s32, s33, q);
  }
  double v[9];
  {
    double w = q[3], x = q[0], y = q[1], z = q[2];
    double xx = x * x, yy = y * y, zz = z * z, xz = x * z, xy = x * y, yz = y * z, wx = w * x, wy = w * y,
           wz = w * z;
    v[0] = 1 - 2 * (yy + zz); v[1] = 2 * (xy - wz); v[2] = 2 * (xz + wy);
    v[3] = 2 * (xy + wz); v[4] = 1 - 2 * (xx + zz); v[5] = 2 * (yz - wx);
    v[6] = 2 * (xz - wy); v[7] = 2 * (yz + wx); v[8] = 1 - 2 * (xx + yy);
  }
  double b[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) b[3 * i + j] = a[3 * i] * v[j] + a[3 * i + 1] * v[3 + j] + a[3 * i + 2] * v[6 + j];
  // order columns by decreasing norm, flipping one sign per swap
  double r1 = b[0] * b[0] + b[3] * b[3] + b[6] * b[6];
  double r2 = b[1] * b[1] + b[4] * b[4] + b[7] * b[7];
  double r3 = b[2] * b[2] + b[5] * b[5] + b[8] * b[8];
  q4 qs{(float)q[3], (float)q[0], (float)q[1], (float)q[2]};
  bool c = r1 < r2;
#pragma unroll
  for (int r = 0; r < 3; ++r) { cnswap(c, b[3 * r], b[3 * r + 1]); cnswap(c, v[3 * r], v[3 * r + 1]); }
  cswap(c, r1, r2);
  qs = qsel(c, qmul_quarter<2>(qs, 1.f), qs);  // columns (0, 1): +90 degrees about z
  c = r1 < r3;
#pragma unroll
  for (int r = 0; r < 3; ++r) { cnswap(c, b[3 * r], b[3 * r + 2]); cnswap(c, v[3 * r], v[3 * r + 2]); }
  cswap(c, r1, r3);
  qs = qsel(c, qmul_quarter<1>(qs, -1.f), qs);  // columns (0, 2): -90 degrees about y
  c = r2 < r3;
#pragma unroll
  for (int r = 0; r < 3; ++r) { cnswap(c, b[3 * r + 1], b[3 * r + 2]); cnswap(c, v[3 * r + 1], v[3 * r + 2]); }
  qs = qsel(c, qmul_quarter<0>(qs, 1.f), qs);  // columns (1, 2): +90 degrees about x
  qv = qs;
  // Givens QR of B: three rotations annihilating (2,1), (3,1), (3,2)
  double ch1, sh1, ch2, sh2, ch3, sh3, ca, sb;
  qr_rot(b[0], b[3], ch1, sh1);
  ca = 1 - 2 * sh1 * sh1;
  sb = 2 * ch1 * sh1;
  double r11 = ca * b[0] + sb * b[3], r12 = ca * b[1] + sb * b[4], r13 = ca * b[2] + sb * b[5];
  double r21 = -sb * b[0] + ca * b[3], r22 = -sb * b[1] + ca * b[4], r23 = -sb * b[2] + ca * b[5];
  double r31 = b[6], r32 = b[7], r33 = b[8];
  qr_rot(r11, r31, ch2, sh2);
  ca = 1 - 2 * sh2 * sh2;
  sb = 2 * ch2 * sh2;
  double c11 = ca * r11 + sb * r31;
  double c22 = r22, c23 = r23;
  double c32 = -sb * r12 + ca * r32, c33 = -sb * r13 + ca * r33;
  (void)r21;
  qr_rot(c22, c32, ch3, sh3);
  ca = 1 - 2 * sh3 * sh3;
  sb = 2 * ch3 * sh3;
  double d22 = ca * c22 + sb * c32;
  double d33 = -sb * c23 + ca * c33;
  double h1 = sh1 * sh1, h2 = sh2 * sh2, h3 = sh3 * sh3;
  double u[9];
  u[0] = (-1 + 2 * h1) * (-1 + 2 * h2);
  u[1] = 4 * ch2 * ch3 * (-1 + 2 * h1) * sh2 * sh3 + 2 * ch1 * sh1 * (-1 + 2 * h3);
  u[2] = 4 * ch1 * ch3 * sh1 * sh3 - 2 * ch2 * (-1 + 2 * h1) * sh2 * (-1 + 2 * h3);
  u[3] = 2 * ch1 * sh1 * (1 - 2 * h2);
  u[4] = -8 * ch1 * ch2 * ch3 * sh1 * sh2 * sh3 + (-1 + 2 * h1) * (-1 + 2 * h3);
  u[5] = -2 * ch3 * sh3 + 4 * sh1 * (ch3 * sh1 * sh3 + ch1 * ch2 * sh2 * (-1 + 2 * h3));
  u[6] = 2 * ch2 * sh2;
  u[7] = 2 * ch3 * (1 - 2 * h2) * sh3;
  u[8] = (-1 + 2 * h2) * (-1 + 2 * h3);
#pragma unroll
  for (int i = 0; i < 9; ++i) { U.a[i] = (float)u[i]; V.a[i] = (float)v[i]; }
  sig = f3{(float)c11, (float)d22, (float)d33};
  {  // U = G1^T G2^T G3^T: rotations by the Givens angles about z, -y and x
    const float c1 = (float)ch1, s1 = (float)sh1, c2 = (float)ch2, s2 = (float)sh2, c3 = (float)ch3,
                s3 = (float)sh3;
    const float pw = c1 * c2, px = s1 * s2, py = -c1 * s2, pz = s1 * c2;  // (c1, s1 z) (x) (c2, -s2 y)
    qu = q4{pw * c3 - px * s3, pw * s3 + px * c3, py * c3 + pz * s3, pz * c3 - py * s3};  // (x) (c3, s3 x)
  }
}
DDM_HD void svd3(const m33 &F, m33 &U, f3 &sig, m33 &V) {
  q4 qu, qv;
  svd3(F, U, sig, V, qu, qv);
}

// ------------------------------------------------------------------ particle physics
constexpr float kNormEps = 1e-8f;      // integrator.cu:11 (passed as float)
constexpr float kSigClip = 0.05f;      // integrator.cu:12

DDM_HD float norm8(f3 a) { return sqrtf(dot3(a, a) + kNormEps); }  // integrator.cu:28-31

// quadratic B-spline weights and their x-derivatives along one axis (integrator.cu:348-352, 493)
struct Stencil {
  int base[3];
  float fx[3];
  float w[3][3];  // w[axis][offset]
};
DDM_HD void stencil(f3 x, float inv_dx, const int dim[3], Stencil &s) {
  float t[3] = {x.x * inv_dx - 0.5f, x.y * inv_dx - 0.5f, x.z * inv_dx - 0.5f};
  float xs[3] = {x.x * inv_dx, x.y * inv_dx, x.z * inv_dx};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int b = (int)floorf(t[d]);
    // keep the 3^3 stencil inside the grid; identity for in-domain particles
    b = min(max(b, 0), dim[d] - 3);
    s.base[d] = b;
    float f = xs[d] - (float)b;
    s.fx[d] = f;
    float a0 = 1.5f - f, a1 = f - 1.f, a2 = f - 0.5f;
    s.w[d][0] = 0.5f * (a0 * a0);
    s.w[d][1] = 0.75f - a1 * a1;
    s.w[d][2] = 0.5f * (a2 * a2);
  }
}
DDM_HD void stencil_dw(const Stencil &s, float inv_dx, float w1[3][3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float f = s.fx[d];
    w1[d][0] = -inv_dx * (1.5f - f);
    w1[d][1] = inv_dx * (-2.f * f + 2.f);
    w1[d][2] = -inv_dx * (f * -1.f + 0.5f);
  }
}
// gradient of the tensor-product weight (integrator.cu:33-40)
DDM_HD f3 weight_grad(const Stencil &s, const float w1[3][3], int i, int j, int k) {
  return f3{w1[0][i] * s.w[1][j] * s.w[2][k], s.w[0][i] * w1[1][j] * s.w[2][k], s.w[0][i] * s.w[1][j] * w1[2][k]};
}

// von Mises return mapping on the singular values (integrator.cu:42-67)
struct Plastic {
  f3 eps, eps_hat;
  float eps_hat_norm, delta_gamma;
};
DDM_HD Plastic plastic_state(f3 sig, float yield, float mu) {
  Plastic p;
  f3 sc = fmax3(sig, kSigClip);
  p.eps = f3{logf(sc.x), logf(sc.y), logf(sc.z)};
  p.eps_hat = p.eps - (p.eps.x + p.eps.y + p.eps.z) / 3.f;
  p.eps_hat_norm = norm8(p.eps_hat);
  p.delta_gamma = p.eps_hat_norm - yield / (2.f * mu);
  return p;
}
DDM_HD f3 plastic_exp(const Plastic &p) {
  f3 e = p.eps - (p.delta_gamma / p.eps_hat_norm) * p.eps_hat;
  return f3{expf(e.x), expf(e.y), expf(e.z)};
}
// U diag(e) V^T
DDM_HD m33 usvt(const m33 &U, f3 e, const m33 &V) {
  m33 r;
  float ev[3] = {e.x, e.y, e.z};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[3 * i + j] = (U.a[3 * i] * ev[0]) * V.a[3 * j] + (U.a[3 * i + 1] * ev[1]) * V.a[3 * j + 1] +
                       (U.a[3 * i + 2] * ev[2]) * V.a[3 * j + 2];
  return r;
}
// returns J, writes the plastically projected F (integrator.cu:42-67)
DDM_HD float return_map(const m33 &Ftmp, const m33 &U, f3 sig, const m33 &V, float yield, float mu, m33 &Fnew,
                        Plastic &p) {
  p = plastic_state(sig, yield, mu);
  if (p.delta_gamma > 0.f) {
    f3 e = plastic_exp(p);
    Fnew = usvt(U, e, V);
    return e.x * e.y * e.z;
  }
  Fnew = Ftmp;
  return sig.x * sig.y * sig.z;
}
// Kirchhoff-type stress of the fixed-corotated model (integrator.cu:364-369)
DDM_HD m33 fc_stress(const m33 &Fnew, const m33 &R, float J, float mu, float lam) {
  m33 s = mscale(mmul(msub(Fnew, R), mtrans(Fnew)), 2.f * mu);
  float d = lam * J * (J - 1.f);
  s.a[0] += d; s.a[4] += d; s.a[8] += d;
  return s;
}
DDM_HD m33 mul_abt(const m33 &a, const m33 &b) {  // a * b^T
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[3 * i + j] = a.a[3 * i] * b.a[3 * j] + a.a[3 * i + 1] * b.a[3 * j + 1] + a.a[3 * i + 2] * b.a[3 * j + 2];
  return r;
}
DDM_HD m33 mul_atb(const m33 &a, const m33 &b) {  // a^T * b
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[3 * i + j] = a.a[i] * b.a[j] + a.a[3 + i] * b.a[3 + j] + a.a[6 + i] * b.a[6 + j];
  return r;
}

// F_tmp = (I + dt C) F  (integrator.cu:90). Written out here (the operations of mscale and
// mmul, in their order) and never contracted, also in sources built with FMA contraction,
// so every kernel that computes it gets the same bits.
DDM_HD m33 advect_F(const m33 &F, const m33 &C, float dt) {
#pragma clang fp contract(off)
  float A[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) A[i] = C.a[i] * dt;
  A[0] += 1.f; A[4] += 1.f; A[8] += 1.f;
  m33 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[3 * i + j] = A[3 * i + 0] * F.a[0 + j] + A[3 * i + 1] * F.a[3 + j] + A[3 * i + 2] * F.a[6 + j];
  return r;
}

// Per-particle P2G quantities (integrator.cu:340-375). Returns the affine matrix and
// momentum used by the scatter; writes the projected F.
struct P2GParticle {
  Stencil st;
  m33 affine;
  f3 mv;
  float m;
};
DDM_HD void p2g_particle(f3 x, f3 v, const m33 &C, const m33 &Ftmp, const m33 &U, f3 sig, const m33 &V, float m,
                         float vol, f3 mly, float dx, float inv_dx, float dt, const int dim[3], m33 &Fnew,
                         P2GParticle &out) {
  stencil(x, inv_dx, dim, out.st);
  Plastic pl;
  float J = return_map(Ftmp, U, sig, V, mly.z, mly.x, Fnew, pl);
  m33 R = mul_abt(U, V);
  m33 s = fc_stress(Fnew, R, J, mly.x, mly.y);
  s = mscale(s, -dt * vol * 4.f * inv_dx * inv_dx);
  out.affine = madd(s, mscale(C, m));
  out.mv = m * v;
  out.m = m;
  (void)dx;
}
// select among three values by a small runtime index (keeps stencil tables in registers)
DDM_HD float sel3(int i, float a, float b, float c) { return i == 0 ? a : (i == 1 ? b : c); }
// o in [0, 27) -> stencil offset (i, j, k), i-major as the reference loops
DDM_HD void offset_ijk(int o, int &i, int &j, int &k) {
  i = o / 9;
  j = (o / 3) % 3;
  k = o % 3;
}
// stencil weight for a runtime offset
DDM_HD float weight_dyn(const Stencil &s, int i, int j, int k) {
  return sel3(i, s.w[0][0], s.w[0][1], s.w[0][2]) * sel3(j, s.w[1][0], s.w[1][1], s.w[1][2]) *
         sel3(k, s.w[2][0], s.w[2][1], s.w[2][2]);
}

// contribution of one particle with stencil weight w to stencil node (i,j,k):
// (m w, w (m v + affine dpos))
DDM_HD void p2g_node_contrib(float w, f3 fx, float m, f3 mv, const m33 &affine, float dx, int i, int j, int k,
                             float &cm, f3 &cmv) {
  f3 dpos = f3{((float)i - fx.x) * dx, ((float)j - fx.y) * dx, ((float)k - fx.z) * dx};
  cm = m * w;
  cmv = (mv + mvec(affine, dpos)) * w;
}
DDM_HD void p2g_contrib(const P2GParticle &p, float dx, int i, int j, int k, float &cm, f3 &cmv) {
  p2g_node_contrib(weight_dyn(p.st, i, j, k), f3{p.st.fx[0], p.st.fx[1], p.st.fx[2]}, p.m, p.mv, p.affine, dx, i, j,
                   k, cm, cmv);
}

// ------------------------------------------------------------------ grid node update
struct Body {
  const float *prim;  // (type, friction, softness, round)
  const float *args;
  f3 pos, npos;       // pose at substep f and f+1
  q4 rot, nrot;
};
DDM_HD Body load_body(const float *prim, const float *args, const float *pos, const float *rot, const float *npos,
                      const float *nrot, int b) {
  Body B;
  B.prim = prim + 4 * b;
  B.args = args + 4 * b;
  B.pos = ld3(pos + 3 * b);
  B.rot = ldq(rot + 4 * b);
  B.npos = ld3(npos + 3 * b);
  B.nrot = ldq(nrot + 4 * b);
  return B;
}
// influence test of the contact model (integrator.cu:705-710); compares in double as
// the reference does (float vs. double literal)
DDM_HD bool in_contact(float dist, float softness, float &influence) {
  influence = fminf(expf(-dist * softness), 1.f);
  return (softness > 0.f && (double)influence > 0.1) || dist <= 0.f;
}
// contact of a node velocity with one moving primitive (integrator.cu:694-730)
DDM_HD bool contact(f3 &v, f3 gx, const Body &B, float dt) {
  f3 pb = to_body(B.pos, B.rot, gx);
  float dist = prim_sdf(B.prim, B.args, pb);
  float infl;
  if (!in_contact(dist, B.prim[2], infl)) return false;
  f3 n = qrot(B.rot, unit30(prim_grad(B.prim, B.args, pb)));
  f3 bv = (to_world(B.npos, B.nrot, pb) - gx) / dt;
  f3 rel = v - bv;
  float nc = dot3(rel, n);
  f3 vt = rel - fminf(nc, 0.f) * n;
  if (nc < 0.f && (double)dot3(vt, vt) > 1e-30) {
    float vtn = len30(vt);
    vt = vt * (1.f / vtn) * fmaxf(0.f, vtn + nc * B.prim[1]);
  }
  v = bv + rel * (1.f - infl) + vt * infl;
  return true;
}

#include "mpm_device_node.inc"

}  // namespace ddm
