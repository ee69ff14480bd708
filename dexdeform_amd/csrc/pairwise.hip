// Point-cloud scores on gfx950: all-pairs reductions between two point clouds without
// forming the N x M cost matrix (the shape-matching scores DexDeform evaluates between
// simulated particles and a target shape: the debiased Sinkhorn divergence that
// tutorials/4_computing_score.ipynb computes with geomloss, and the Chamfer distance).
//
//   softmin   out_i = -eps * log sum_j exp(h_j - C(x_i, y_j) / eps)
//   plan_grad out_i = sum_j exp(a_i + b_j - C(x_i, y_j) / eps) * dC/dx (x_i, y_j)
//   nearest   out_i = min_j |x_i - y_j|^2 and its argmin
//
// with C = |x - y| (p = 1) or |x - y|^2 / 2 (p = 2). Layout: points are float [n][3]
// (the particle layout), weights float [n]. Each workgroup stages a tile of y (with its
// weight) in LDS; each lane keeps R points of x in registers and streams the tile (all
// lanes read the same LDS entry: a broadcast), so each LDS read feeds R pairs. The j range
// is split into segments so that a launch has enough workgroups for 256 CUs even when n
// is small; per-segment partial results are combined in a second, deterministic pass
// (fixed segment order), so results do not depend on scheduling.
//
// The online log-sum-exp keeps a running maximum m and sum s per x point in base 2 (the
// weights are pre-scaled by log2 e when a tile is staged, so each pair costs one v_exp_f32
// and no multiply); a group of 8 pairs is reduced at a time (one rescale exp per group).
// The p = 1 distance is the hardware square root (v_sqrt_f32, 1 ulp) and |x - y|^2 and the
// exponent are fused multiply-adds: about 10 VALU instructions per pair, 2 of them
// transcendental. It is VALU and transcendental bound: no MFMA (the p = 1 cost needs a
// square root per pair).
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "../../include/dexdeform_mpm.h"

namespace ddm_pairwise {
namespace {

constexpr int kWG = 256;     // threads per workgroup
constexpr int kR = 2;        // x points per lane
constexpr int kTile = 256;   // y points per LDS tile
constexpr int kGroup = 8;    // pairs reduced together in the online log-sum-exp

constexpr float kLog2e = 1.44269504088896341f;
constexpr float kLn2 = 0.693147180559945309f;

template <int P>
__device__ __forceinline__ float cost(float r2) {
  if (P == 1) return __builtin_amdgcn_sqrtf(r2);
  return 0.5f * r2;
}

__device__ __forceinline__ float dist2(const float *p, float4 q) {
  const float dx = p[0] - q.x, dy = p[1] - q.y, dz = p[2] - q.z;
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Stages y[j0 .. j0+kTile) (and its weight times `scale`; missing points get weight
// `pad`) into LDS.
__device__ __forceinline__ void load_tile(float4 *t, const float *__restrict__ y, const float *__restrict__ w,
                                          int j0, int j_end, float pad, float scale = 1.f) {
  const int j = j0 + (int)threadIdx.x;
  float4 v = make_float4(0.f, 0.f, 0.f, pad);
  if (j < j_end) v = make_float4(y[3 * j], y[3 * j + 1], y[3 * j + 2], w ? w[j] * scale : 0.f);
  t[threadIdx.x] = v;
}

// Partial log-sum-exp of one j segment: part[seg][i] = (m, s) with
// sum_{j in seg} exp(h_j - C_ij / eps) = s * 2^m. inv_eps2 = log2(e) / eps.
template <int P>
__global__ __launch_bounds__(kWG) void k_softmin_part(const float *__restrict__ x, int n, const float *__restrict__ y,
                                                      const float *__restrict__ h, int m, int seg_len, float inv_eps2,
                                                      float2 *__restrict__ part) {
  __shared__ float4 tile[kTile];
  const int seg = blockIdx.y;
  const int j_begin = seg * seg_len, j_end = min(m, j_begin + seg_len);
  float px[kR][3];
  float mx[kR], sm[kR];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = (blockIdx.x * kR + r) * kWG + (int)threadIdx.x;
    const int ic = min(i, n - 1);
    px[r][0] = x[3 * ic];
    px[r][1] = x[3 * ic + 1];
    px[r][2] = x[3 * ic + 2];
    mx[r] = -FLT_MAX;
    sm[r] = 0.f;
  }
  for (int j0 = j_begin; j0 < j_end; j0 += kTile) {
    __syncthreads();
    load_tile(tile, y, h, j0, j_end, -INFINITY, kLog2e);
    __syncthreads();
    const int cnt = min(kTile, j_end - j0);
    for (int k0 = 0; k0 < cnt; k0 += kGroup) {
      float t[kR][kGroup];
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        const float4 q = tile[k0 + g];  // entries past cnt hold weight -inf
#pragma unroll
        for (int r = 0; r < kR; ++r) t[r][g] = fmaf(-cost<P>(dist2(px[r], q)), inv_eps2, q.w);
      }
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        float gm = t[r][0];
#pragma unroll
        for (int g = 1; g < kGroup; ++g) gm = fmaxf(gm, t[r][g]);
        if (gm > mx[r]) {
          sm[r] *= __builtin_amdgcn_exp2f(mx[r] - gm);
          mx[r] = gm;
        }
        float acc = 0.f;
#pragma unroll
        for (int g = 0; g < kGroup; ++g) acc += __builtin_amdgcn_exp2f(t[r][g] - mx[r]);
        sm[r] += acc;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = (blockIdx.x * kR + r) * kWG + (int)threadIdx.x;
    if (i < n) part[(size_t)seg * n + i] = make_float2(mx[r], sm[r]);
  }
}

// out_i = -eps ln2 (M + log2 S) over the segments in order
__global__ void k_softmin_combine(const float2 *__restrict__ part, int n, int segs, float eps,
                                  float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float M = -FLT_MAX;
  for (int s = 0; s < segs; ++s) M = fmaxf(M, part[(size_t)s * n + i].x);
  float S = 0.f;
  for (int s = 0; s < segs; ++s) {
    const float2 v = part[(size_t)s * n + i];
    S += v.y * exp2f(v.x - M);
  }
  out[i] = -eps * kLn2 * (M + log2f(S));
}

// Partial plan-weighted cost gradients of one j segment:
// part[seg][i] = sum_{j in seg} exp(a_i + b_j - C_ij / eps) dC/dx(x_i, y_j), evaluated as
// 2^((a_i + b_j) log2 e - C_ij log2 e / eps); inv_eps2 = log2(e) / eps.
template <int P>
__global__ __launch_bounds__(kWG) void k_plan_grad_part(const float *__restrict__ x, const float *__restrict__ a,
                                                        int n, const float *__restrict__ y,
                                                        const float *__restrict__ b, int m, int seg_len,
                                                        float inv_eps2, float4 *__restrict__ part) {
  __shared__ float4 tile[kTile];
  const int seg = blockIdx.y;
  const int j_begin = seg * seg_len, j_end = min(m, j_begin + seg_len);
  float px[kR][3], pa[kR], acc[kR][3];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = (blockIdx.x * kR + r) * kWG + (int)threadIdx.x;
    const int ic = min(i, n - 1);
    px[r][0] = x[3 * ic];
    px[r][1] = x[3 * ic + 1];
    px[r][2] = x[3 * ic + 2];
    pa[r] = a[ic] * kLog2e;
    acc[r][0] = acc[r][1] = acc[r][2] = 0.f;
  }
  for (int j0 = j_begin; j0 < j_end; j0 += kTile) {
    __syncthreads();
    load_tile(tile, y, b, j0, j_end, -INFINITY, kLog2e);
    __syncthreads();
    const int cnt = min(kTile, j_end - j0);
    for (int k = 0; k < cnt; ++k) {
      const float4 q = tile[k];
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        const float dx = px[r][0] - q.x, dy = px[r][1] - q.y, dz = px[r][2] - q.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        float wgt, gscale;
        if (P == 1) {
          // d|x - y|/dx = (x - y) / |x - y|; 0 at coincident points
          const bool pos = r2 > 0.f;
          const float inv_d = pos ? __builtin_amdgcn_rsqf(r2) : 0.f;
          wgt = __builtin_amdgcn_exp2f(fmaf(-r2 * inv_d, inv_eps2, pa[r] + q.w));
          gscale = wgt * inv_d;
        } else {
          wgt = __builtin_amdgcn_exp2f(fmaf(-0.5f * r2, inv_eps2, pa[r] + q.w));
          gscale = wgt;
        }
        acc[r][0] += gscale * dx;
        acc[r][1] += gscale * dy;
        acc[r][2] += gscale * dz;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = (blockIdx.x * kR + r) * kWG + (int)threadIdx.x;
    if (i < n) part[(size_t)seg * n + i] = make_float4(acc[r][0], acc[r][1], acc[r][2], 0.f);
  }
}

__global__ void k_grad_combine(const float4 *__restrict__ part, int n, int segs, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gx = 0.f, gy = 0.f, gz = 0.f;
  for (int s = 0; s < segs; ++s) {
    const float4 v = part[(size_t)s * n + i];
    gx += v.x;
    gy += v.y;
    gz += v.z;
  }
  out[3 * i] = gx;
  out[3 * i + 1] = gy;
  out[3 * i + 2] = gz;
}

// Partial nearest neighbour of one j segment: (squared distance, index); ties go to the
// smaller index.
__global__ __launch_bounds__(kWG) void k_nearest_part(const float *__restrict__ x, int n, const float *__restrict__ y,
                                                      int m, int seg_len, float2 *__restrict__ part) {
  __shared__ float4 tile[kTile];
  const int seg = blockIdx.y;
  const int j_begin = seg * seg_len, j_end = min(m, j_begin + seg_len);
  float px[kR][3], best[kR];
  int arg[kR];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = (blockIdx.x * kR + r) * kWG + (int)threadIdx.x;
    const int ic = min(i, n - 1);
    px[r][0] = x[3 * ic];
    px[r][1] = x[3 * ic + 1];
    px[r][2] = x[3 * ic + 2];
    best[r] = FLT_MAX;
    arg[r] = -1;
  }
  for (int j0 = j_begin; j0 < j_end; j0 += kTile) {
    __syncthreads();
    load_tile(tile, y, nullptr, j0, j_end, 0.f);
    __syncthreads();
    const int cnt = min(kTile, j_end - j0);
    for (int k = 0; k < cnt; ++k) {
      const float4 q = tile[k];
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        const float r2 = dist2(px[r], q);
        if (r2 < best[r]) {
          best[r] = r2;
          arg[r] = j0 + k;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int i = (blockIdx.x * kR + r) * kWG + (int)threadIdx.x;
    if (i < n) part[(size_t)seg * n + i] = make_float2(best[r], __int_as_float(arg[r]));
  }
}

__global__ void k_nearest_combine(const float2 *__restrict__ part, int n, int segs, float *__restrict__ d2,
                                  int32_t *__restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float best = FLT_MAX;
  int arg = -1;
  for (int s = 0; s < segs; ++s) {
    const float2 v = part[(size_t)s * n + i];
    const int a = __float_as_int(v.y);
    if (a >= 0 && (v.x < best || (v.x == best && a < arg))) {
      best = v.x;
      arg = a;
    }
  }
  d2[i] = best;
  idx[i] = arg;
}

inline int blocks_x(int n) { return (n + kR * kWG - 1) / (kR * kWG); }

}  // namespace
}  // namespace ddm_pairwise

using namespace ddm_pairwise;

#define DDM_API extern "C" __attribute__((visibility("default")))

DDM_API size_t ddm_pairwise_workspace_bytes(int n, int segments) {
  if (n <= 0 || segments <= 0) return 0;
  return (size_t)n * (size_t)segments * sizeof(float4);
}

static int check_pair_args(const float *x, int n, const float *y, int m, int segments, void *ws, size_t ws_bytes) {
  if (!x || !y || n <= 0 || m <= 0 || segments <= 0 || segments > m || !ws) return DDM_EINVAL;
  if (ws_bytes < ddm_pairwise_workspace_bytes(n, segments)) return DDM_EINVAL;
  if ((size_t)n * 3 >= (size_t)INT32_MAX || (size_t)m * 3 >= (size_t)INT32_MAX) return DDM_EINVAL;
  return 0;
}

DDM_API int ddm_softmin(const float *x, int n, const float *y, const float *h, int m, float eps, int p, int segments,
                        void *workspace, size_t workspace_bytes, float *out, hipStream_t stream) {
  int e = check_pair_args(x, n, y, m, segments, workspace, workspace_bytes);
  if (e) return e;
  if (!h || !out || !(eps > 0.f) || (p != 1 && p != 2)) return DDM_EINVAL;
  const int seg_len = (m + segments - 1) / segments;
  const int segs = (m + seg_len - 1) / seg_len;
  float2 *part = static_cast<float2 *>(workspace);
  dim3 grid(blocks_x(n), segs);
  const float inv_eps2 = kLog2e / eps;
  if (p == 1) k_softmin_part<1><<<grid, kWG, 0, stream>>>(x, n, y, h, m, seg_len, inv_eps2, part);
  else k_softmin_part<2><<<grid, kWG, 0, stream>>>(x, n, y, h, m, seg_len, inv_eps2, part);
  k_softmin_combine<<<(n + 255) / 256, 256, 0, stream>>>(part, n, segs, eps, out);
  return (int)hipGetLastError();
}

DDM_API int ddm_plan_grad(const float *x, const float *a, int n, const float *y, const float *b, int m, float eps,
                          int p, int segments, void *workspace, size_t workspace_bytes, float *grad,
                          hipStream_t stream) {
  int e = check_pair_args(x, n, y, m, segments, workspace, workspace_bytes);
  if (e) return e;
  if (!a || !b || !grad || !(eps > 0.f) || (p != 1 && p != 2)) return DDM_EINVAL;
  const int seg_len = (m + segments - 1) / segments;
  const int segs = (m + seg_len - 1) / seg_len;
  float4 *part = static_cast<float4 *>(workspace);
  dim3 grid(blocks_x(n), segs);
  const float inv_eps2 = kLog2e / eps;
  if (p == 1) k_plan_grad_part<1><<<grid, kWG, 0, stream>>>(x, a, n, y, b, m, seg_len, inv_eps2, part);
  else k_plan_grad_part<2><<<grid, kWG, 0, stream>>>(x, a, n, y, b, m, seg_len, inv_eps2, part);
  k_grad_combine<<<(n + 255) / 256, 256, 0, stream>>>(part, n, segs, grad);
  return (int)hipGetLastError();
}

DDM_API int ddm_nearest(const float *x, int n, const float *y, int m, int segments, void *workspace,
                        size_t workspace_bytes, float *dist2, int32_t *index, hipStream_t stream) {
  int e = check_pair_args(x, n, y, m, segments, workspace, workspace_bytes);
  if (e) return e;
  if (!dist2 || !index) return DDM_EINVAL;
  const int seg_len = (m + segments - 1) / segments;
  const int segs = (m + seg_len - 1) / seg_len;
  float2 *part = static_cast<float2 *>(workspace);
  dim3 grid(blocks_x(n), segs);
  k_nearest_part<<<grid, kWG, 0, stream>>>(x, n, y, m, seg_len, part);
  k_nearest_combine<<<(n + 255) / 256, 256, 0, stream>>>(part, n, segs, dist2, index);
  return (int)hipGetLastError();
}
