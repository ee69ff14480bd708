// Particle-volume renderer on gfx950: the SDF path tracer DexDeform uses to look at a
// simulation state (mpm/renderer.py; kernels mpm/csrc/integrator.cu:1111-1525, launchers
// particle_sdf / render at integrator.cu:1863-1933, trilinear volume queries array.h:6-75).
//
// particle_sdf bakes the particles into a voxel volume:
//   bake    every particle writes min(packed (quantised distance << 24 | colour)) into the
//           (2 bake)^3 voxels around it (64-bit atomic min, so the result does not depend
//           on the order of the writes); a voxel is read first and the atomic skipped when
//           it already holds a smaller value (the stored value only ever decreases);
//   unpack  distance byte -> sdf in [0, 1], colour bytes -> RGB;
//   smooth  two 3x3x3 box filters (fixed summation order).
// render path-traces the image: the ground plane, the bodies (sphere tracing over their
// SDFs) and the baked particle volume (ray marching with a bisection refinement), with
// a diffuse bounce and a jittered directional light, max_ray_depth bounces per sample.
//
// Layout and launch (render): a workgroup owns an 8 x 8 pixel tile and up to 16 of its
// samples at a time (one wave per sample, so a wave traces 64 neighbouring pixels of the
// same sample: coherent rays); the per-sample results go through LDS and one wave adds
// them in sample order, so the image is the same as a sequential loop over samples. The
// bodies' poses and shapes sit in LDS. Nothing is allocated: the random numbers come
// from a counter-based generator (Philox4x32-10, Salmon, Moraes, Dror and Shaw, SC'11)
// keyed by the seed, with one stream per (pixel, sample): draw k of sample s of pixel p
// is 32-bit word k % 4 of Philox(counter = (k / 4, p, s, 0), key = (seed, 0)), mapped to
// (0, 1] as ((w >> 8) + 1) 2^-24. The reference draws from curand's XORWOW stream per
// pixel instead (integrator.cu:1306), so its images differ from these by the random
// numbers; the estimator and every other operation follow the reference. cos/sin of the
// bounce angle come from Taylor polynomials evaluated with fused multiply-adds after an
// exact reduction, so the CPU restatement reproduces them bit for bit.
//
// The arithmetic is FP32 without contraction (this file is compiled with
// -ffp-contract=off), with the reference's double-precision literals and casts kept.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dexdeform_mpm.h"
#include "mpm_device.h"

namespace ddm_render {
namespace {

using ddm::f3;
using ddm::q4;

constexpr int kTile = 8;             // pixel tile edge (64 pixels = one wave)
constexpr int kMaxSampleWaves = 16;  // waves per workgroup (samples in flight)
constexpr int kMaxBodies = DDM_MAX_BODIES;

__device__ __forceinline__ f3 div3(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 i2f(int x, int y, int z) { return f3{(float)x, (float)y, (float)z}; }
// normalized (shape.h:13-15): v / (float)sqrt(double(v.v) + 1e-30)
__device__ __forceinline__ f3 normalized(f3 v) { return v / ddm::len30(v); }
// mat3::mul (mat3.h:95-98) of a row-major 3x3
__device__ __forceinline__ f3 matvec(const float *m, f3 b) {
  f3 c0{m[0], m[3], m[6]}, c1{m[1], m[4], m[7]}, c2{m[2], m[5], m[8]};
  return c0 * b.x + c1 * b.y + c2 * b.z;
}

// ------------------------------------------------------------------ random numbers
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct Rng {
  uint32_t key, pixel, sample, block, used;
  uint32_t w[4];
  __device__ float uniform() {
    if (used == 4) {
      w[0] = block++;
      w[1] = pixel;
      w[2] = sample;
      w[3] = 0u;
      philox4x32_10(w, key, 0u);
      used = 0;
    }
    const uint32_t x = w[used++];
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
  }
};

// cos and sin of phi in [0, 2 pi]: k = nearest multiple of pi/2, r = phi - k pi/2 in two
// fused steps (pi/2 split into an 11-significant-bit head and a tail), then the Taylor
// series of sin to r^9 and of cos to r^10 on |r| <= pi/4 (truncation below 2e-9).
__device__ __forceinline__ void sincos_taylor(float phi, float &s, float &c) {
  const float k = rintf(phi * 0.636619772f);
  float r = fmaf(-k, 1.5703125f, phi);
  r = fmaf(-k, 4.83826794e-4f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.75573192e-6f, -1.98412698e-4f);
  ps = fmaf(r2, ps, 8.33333377e-3f);
  ps = fmaf(r2, ps, -1.66666672e-1f);
  const float sn = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.75573188e-7f, 2.48015876e-5f);
  pc = fmaf(r2, pc, -1.38888892e-3f);
  pc = fmaf(r2, pc, 4.16666679e-2f);
  pc = fmaf(r2, pc, -0.5f);
  const float cs = fmaf(r2, pc, 1.0f);
  const int q = ((int)k) & 3;
  s = q == 0 ? sn : q == 1 ? cs : q == 2 ? -sn : -cs;
  c = q == 0 ? cs : q == 1 ? -sn : q == 2 ? -cs : sn;
}

// ------------------------------------------------------------------ volume queries (array.h)
struct Volume {
  const float *sdf;
  const float *color;  // vec3 per voxel
  f3 lo, hi;
  int dx, dy, dz;
};

// trilinear query (array.h:6-49); returns false outside [lo, hi)
template <int C>
__device__ __forceinline__ bool volume_query(const Volume &V, const float *values, f3 point, float *out) {
  const f3 ratio = div3(point - V.lo, V.hi - V.lo);
  if (!((double)ratio.x >= 0 && (double)ratio.y >= 0 && (double)ratio.z >= 0 && (double)ratio.x < 1. &&
        (double)ratio.y < 1. && (double)ratio.z < 1.))
    return false;
  const f3 grid = ratio * i2f(V.dx, V.dy, V.dz);
  const int x0 = min(max((int)floorf(grid.x), 0), V.dx - 1);
  const int y0 = min(max((int)floorf(grid.y), 0), V.dy - 1);
  const int z0 = min(max((int)floorf(grid.z), 0), V.dz - 1);
  const float dx = grid.x - (float)x0, dy = grid.y - (float)y0, dz = grid.z - (float)z0;
  const int i000 = (x0 * V.dy + y0) * V.dz + z0;
  const int fz = z0 + 1 < V.dz, fy = y0 + 1 < V.dy, fx = x0 + 1 < V.dx;
  const int i001 = i000 + fz;
  const int i010 = i000 + V.dz * fy;
  const int i011 = i010 + fz;
  const int i100 = i000 + V.dy * V.dz * fx;
  const int i101 = i100 + fz;
  const int i110 = i100 + V.dz * fy;
  const int i111 = i110 + fz;
  const float ex = 1 - dx, ey = 1 - dy, ez = 1 - dz;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const float v000 = values[C * i000 + c] * ex * ey * ez;
    const float v001 = values[C * i001 + c] * ex * ey * dz;
    const float v010 = values[C * i010 + c] * ex * dy * ez;
    const float v011 = values[C * i011 + c] * ex * dy * dz;
    const float v100 = values[C * i100 + c] * dx * ey * ez;
    const float v101 = values[C * i101 + c] * dx * ey * dz;
    const float v110 = values[C * i110 + c] * dx * dy * ez;
    const float v111 = values[C * i111 + c] * dx * dy * dz;
    out[c] = v000 + v001 + v010 + v011 + v100 + v101 + v110 + v111;
  }
  return true;
}

// sample_sdf (array.h:51-59): the trilinear value minus the threshold inside, 0 outside
__device__ __forceinline__ float sample_sdf(const Volume &V, f3 p, float threshold) {
  float out = 0.f;
  if (volume_query<1>(V, V.sdf, p, &out)) out -= threshold;
  return out;
}

// sample_normal (array.h:61-70): central differences with eps 1e-3
__device__ __forceinline__ f3 sample_normal(const Volume &V, f3 p) {
  const float e = 1e-3f;
  const float x0 = sample_sdf(V, p + f3{-e, 0.f, 0.f}, 0.f), x1 = sample_sdf(V, p + f3{e, 0.f, 0.f}, 0.f);
  const float y0 = sample_sdf(V, p + f3{0.f, -e, 0.f}, 0.f), y1 = sample_sdf(V, p + f3{0.f, e, 0.f}, 0.f);
  const float z0 = sample_sdf(V, p + f3{0.f, 0.f, -e}, 0.f), z1 = sample_sdf(V, p + f3{0.f, 0.f, e}, 0.f);
  return f3{x1 - x0, y1 - y0, z1 - z0} * (float)(0.5 / (double)e);
}

__device__ __forceinline__ f3 sample_color(const Volume &V, f3 p) {
  float c[3] = {0.f, 0.f, 0.f};
  volume_query<3>(V, V.color, p, c);
  return f3{c[0], c[1], c[2]};
}

// ray_aabb_intersection (integrator.cu:1111-1143)
__device__ __forceinline__ bool ray_aabb(f3 lo, f3 hi, f3 o, f3 d, float &tnear, float &tfar) {
  const float dd[3] = {d.x, d.y, d.z}, oo[3] = {o.x, o.y, o.z}, bl[3] = {lo.x, lo.y, lo.z}, bh[3] = {hi.x, hi.y, hi.z};
  tnear = -1e9f;
  tfar = 1e9f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (dd[i] == 0.f) {
      if (oo[i] < bl[i] || oo[i] > bh[i]) return false;
    } else {
      const float i1 = (bl[i] - oo[i]) / dd[i];
      const float i2 = (bh[i] - oo[i]) / dd[i];
      tfar = fminf(fmaxf(i1, i2), tfar);
      tnear = fmaxf(fminf(i1, i2), tnear);
    }
  }
  return !(tnear > tfar);
}

struct Scene {
  Volume V;
  float threshold;
  float ground;
  int n_bodies;
  bool shape;
  const float *pos;   // LDS: n_bodies x 3
  const float *rot;   // LDS: n_bodies x 4
  const float *prim;  // LDS: n_bodies x 4
  const float *args;  // LDS: n_bodies x 4
};

struct Hit {
  f3 color, normal;
  int material;
};

// next_hit (integrator.cu:1145-1267): the closest of ground, bodies and particle volume
// along o + t d; writes colour, normal and material of the hit it keeps.
__device__ float next_hit(const Scene &S, f3 o, f3 d, Hit &h) {
  const float dist_limit = 100.f;
  float closest = 1e9f;
  if (d.y < 0.f) {  // ground
    const float gd = (o.y - S.ground) / (-d.y);
    if (gd < dist_limit && gd < closest) {
      closest = gd;
      h.normal = f3{0.f, 1.f, 0.f};
      f3 col{0.3f, 0.5f, 0.7f};
      const f3 p = o + d * closest;
      if (p.x <= 1.f && p.x >= 0.f && p.z <= 1.f && p.z >= 0.f) {
        const int chk = ((int)((double)p.x / 0.25) + (int)((double)p.z / 0.25)) % 2;
        col = col * (float)(chk * 0.2 + 0.35);
      } else {
        col = col * 0.4f;
      }
      h.color = col;
      h.material = 1;
    }
  }
  if (S.n_bodies > 0) {  // sphere tracing over the bodies' SDFs
    int j = 0, id = 0;
    float dist = 0.f, val = 1e9f;
    f3 pp{0.f, 0.f, 0.f};
    while (j < 200 && dist < dist_limit && (double)val > 1e-8) {
      pp = o + d * dist;
      val = 1e9f;
      for (int i = 0; i < S.n_bodies; ++i) {
        const f3 gb = ddm::to_body(ddm::ld3(S.pos + 3 * i), ddm::ldq(S.rot + 4 * i), pp);
        const float t = ddm::prim_sdf(S.prim + 4 * i, S.args + 4 * i, gb);
        if (t < val) {
          val = t;
          id = i;
        }
      }
      dist += val;
      j += 1;
    }
    if (dist < closest && dist < dist_limit) {
      closest = dist;
      const q4 q = ddm::ldq(S.rot + 4 * id);
      const f3 gb = ddm::to_body(ddm::ld3(S.pos + 3 * id), q, pp);
      h.normal = ddm::qrot(q, normalized(ddm::prim_grad(S.prim + 4 * id, S.args + 4 * id, gb)));
      h.color = f3{0.5f, 0.5f, 0.5f};
      h.material = 1;
    }
  }
  float tnear = 0.f, tfar = 0.f;
  if (S.shape && ray_aabb(S.V.lo, S.V.hi, o, d, tnear, tfar)) {  // particle volume
    tnear = (float)fmax((double)tnear, 0.);
    f3 pos = o + d * (float)((double)tnear + 1e-4);
    f3 step{0.f, 0.f, 0.f};
    for (int j = 0; j < 500; ++j) {
      const float s = sample_sdf(S.V, pos, S.threshold);
      if (s < 0.f) {
        f3 back = step;
        for (int k = 0; k < 20; ++k) {
          back = back * 0.5f;
          if (sample_sdf(S.V, pos - back, S.threshold) < 0.f) pos = pos - back;
        }
        const f3 od = o - pos;
        const float dist = sqrtf(ddm::dot3(od, od));
        if (dist < closest) {
          closest = dist;
          h.normal = normalized(sample_normal(S.V, pos));
          h.color = sample_color(S.V, pos);
          h.material = 1;
        }
        break;
      }
      step = d * (float)fmax((double)s * 0.05, 0.01);
      pos += step;
    }
  }
  return closest;
}

struct Camera {
  float K[9], R[9];
  f3 pos, light;
  int H, W, depth, spp;
};

// One sample of one pixel (integrator.cu:1315-1413, use_directional_light = 1,
// light_direction_noise = 0.1): returns the sample's colour and first-hit distance.
__device__ void trace_sample(const Scene &S, const Camera &cam, int u, int v, Rng &rng, f3 &out, float &first) {
  const float ju = rng.uniform();
  const float jv = rng.uniform();
  f3 dir{(float)u + ju, (float)v + jv, 1.f};
  f3 contrib{0.f, 0.f, 0.f}, through{1.f, 1.f, 1.f};
  first = 0.f;
  dir = matvec(cam.K, dir);
  f3 d = matvec(cam.R, normalized(dir));
  f3 pos = cam.pos;
  for (int depth = 0; depth < cam.depth;) {
    Hit h{f3{0.f, 0.f, 0.f}, f3{0.f, 0.f, 0.f}, 0};
    const float closest = next_hit(S, pos, d, h);
    if (depth == 0) first = closest;
    const f3 hit = pos + d * closest;
    depth += 1;
    if (!((double)ddm::dot3(h.normal, h.normal) > 1e-10)) break;  // sky
    f3 odir;
    if (h.material == 0) {
      odir = d - h.normal * (ddm::dot3(d, h.normal) * 2.f);
    } else {
      f3 ub{1.f, 0.f, 0.f};
      if ((double)fabsf(h.normal.y) < 1 - 1e-3) ub = normalized(ddm::cross3(h.normal, f3{0.f, 1.f, 0.f}));
      const f3 vb = ddm::cross3(h.normal, ub);
      const float phi = (float)(2 * 3.14159265358979323846 * (double)rng.uniform());
      const float r = rng.uniform();
      const float ay = sqrtf(r);
      const float ax = sqrtf(1 - r);
      float sp, cp;
      sincos_taylor(phi, sp, cp);
      odir = (ub * cp + vb * sp) * ax + h.normal * ay;
    }
    d = normalized(odir + f3{0.f, 0.f, 0.f});
    pos = hit + d * 1e-4f;
    through = through * h.color;
    const float n0 = rng.uniform() - 0.5f;
    const float n1 = rng.uniform() - 0.5f;
    const float n2 = rng.uniform() - 0.5f;
    const f3 direct = normalized(cam.light + f3{n0, n1, n2} * 0.1f);
    const float dv = ddm::dot3(direct, h.normal);
    if (dv > 0.f) {
      Hit sh = h;
      const float dist = next_hit(S, pos, direct, sh);
      if (dist > 100.f) contrib += through * dv * f3{0.8f, 0.8f, 0.8f};
    }
  }
  contrib += through * f3{0.8f, 0.8f, 0.8f};  // the sky term is always added (hit_sky stays 1)
  out = contrib;
}

// ------------------------------------------------------------------ kernels
__global__ __launch_bounds__(256) void k_bake(unsigned long long *__restrict__ volume, const float *__restrict__ x,
                                              const int *__restrict__ color, const float *__restrict__ bbox_min,
                                              int bake, int gx, int gy, int gz, float inv_dx, int n) {
  const int w = 2 * bake;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * w * w) return;
  const int p = (int)(t / (w * w));
  const int ij = (int)(t % (w * w));
  const int i = ij / w - bake, j = ij % w - bake;
  const f3 lo = ddm::ld3(bbox_min);
  const f3 xs = (ddm::ld3(x + 3 * p) - lo) * inv_dx;
  const int cx = (int)floorf(xs.x) + i, cy = (int)floorf(xs.y) + j, cz0 = (int)floorf(xs.z);
  if (cx < 0 || cy < 0 || cx >= gx || cy >= gy) return;
  const unsigned long long col = (unsigned long long)(long long)color[p];
  for (int k = -bake; k < bake; ++k) {
    const int cz = cz0 + k;
    if (cz < 0 || cz >= gz) continue;
    const f3 dv = i2f(cx, cy, cz) - xs;
    float dist = sqrtf(ddm::dot3(dv, dv) + 0.f);  // norm(., 0) (integrator.cu:28-31)
    dist = fminf(fmaxf(0.0f, 51.0f * dist), 255.0f);
    const unsigned long long out = ((unsigned long long)(long long)dist << 24) + col;
    unsigned long long *cell = volume + ((size_t)cx * gy + cy) * gz + cz;
    if (__hip_atomic_load(cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > out) atomicMin(cell, out);
  }
}

// copy_volume (integrator.cu:1463-1478)
__global__ __launch_bounds__(256) void k_unpack(const unsigned long long *__restrict__ volume, float *__restrict__ sdf,
                                                float *__restrict__ color, int dim) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dim) return;
  long long c = (long long)volume[t];
  const float z = (float)((double)(c & 255) / 255.);
  c >>= 8;
  const float y = (float)((double)(c & 255) / 255.);
  c >>= 8;
  const float xx = (float)((double)(c & 255) / 255.);
  c >>= 8;
  color[3 * t] = xx;
  color[3 * t + 1] = y;
  color[3 * t + 2] = z;
  sdf[t] = (float)((double)(c & 255) / 255.);
}

// smooth (integrator.cu:1438-1461): 3x3x3 mean inside, 1 on the boundary layer
__global__ __launch_bounds__(256) void k_smooth(const float *__restrict__ in, float *__restrict__ out, int gx, int gy,
                                                int gz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= gx * gy * gz) return;
  const int x = t / gz / gy, y = (t / gz) % gy, z = t % gz;
  if (x >= 1 && y >= 1 && z >= 1 && x < gx - 1 && y < gy - 1 && z < gz - 1) {
    float s = 0.f;
    for (int i = -1; i < 2; ++i)
      for (int j = -1; j < 2; ++j)
        for (int k = -1; k < 2; ++k) s += in[((x + i) * gy + (y + j)) * gz + (z + k)];
    out[t] = (float)((double)s / 27.);
  } else {
    out[t] = 1.f;
  }
}

struct RenderArgs {
  const float *sdf, *box_min, *box_max, *color_volume, *body_pos, *body_rot, *prim, *args, *cam_rot, *cam_pos,
      *cam_K, *light;
  float *color_out, *depth_out;
  float threshold, ground;
  int gx, gy, gz, n_bodies, shape, H, W, max_depth, spp, seed, tiles_u;
};

__global__ __launch_bounds__(64 * kMaxSampleWaves) void k_render(RenderArgs a) {
  __shared__ float s_pos[3 * kMaxBodies], s_rot[4 * kMaxBodies], s_prim[4 * kMaxBodies], s_args[4 * kMaxBodies];
  __shared__ float4 s_out[kMaxSampleWaves][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  for (int i = threadIdx.x; i < a.n_bodies; i += blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; ++c) s_pos[3 * i + c] = a.body_pos[3 * i + c];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      s_rot[4 * i + c] = a.body_rot[4 * i + c];
      s_prim[4 * i + c] = a.prim[4 * i + c];
      s_args[4 * i + c] = a.args[4 * i + c];
    }
  }
  Scene S;
  S.V.sdf = a.sdf;
  S.V.color = a.color_volume;
  S.V.lo = ddm::ld3(a.box_min);
  S.V.hi = ddm::ld3(a.box_max);
  S.V.dx = a.gx;
  S.V.dy = a.gy;
  S.V.dz = a.gz;
  S.threshold = a.threshold;
  S.ground = a.ground;
  S.n_bodies = a.n_bodies;
  S.shape = a.shape != 0;
  S.pos = s_pos;
  S.rot = s_rot;
  S.prim = s_prim;
  S.args = s_args;
  Camera cam;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    cam.K[i] = a.cam_K[i];
    cam.R[i] = a.cam_rot[i];
  }
  cam.pos = ddm::ld3(a.cam_pos);
  cam.light = ddm::ld3(a.light);
  cam.H = a.H;
  cam.W = a.W;
  cam.depth = a.max_depth;
  cam.spp = a.spp;
  // pixel of this lane: column u, row v of an 8 x 8 tile (p = v W + u, integrator.cu:1310-1311)
  const int tile = blockIdx.x;
  const int u = (tile % a.tiles_u) * kTile + (lane & (kTile - 1));
  const int v = (tile / a.tiles_u) * kTile + (lane / kTile);
  const bool valid = u < a.W && v < a.H;
  const int p = v * a.W + u;
  __syncthreads();
  f3 acc{0.f, 0.f, 0.f};
  float dacc = 0.f;
  for (int s0 = 0; s0 < a.spp; s0 += waves) {
    const int s = s0 + wave;
    if (valid && s < a.spp) {
      Rng rng{(uint32_t)a.seed, (uint32_t)p, (uint32_t)s, 0u, 4u, {0u, 0u, 0u, 0u}};
      f3 o;
      float first;
      trace_sample(S, cam, u, v, rng, o, first);
      s_out[wave][lane] = make_float4(o.x, o.y, o.z, first);
    }
    __syncthreads();
    if (wave == 0 && valid) {
      const int cnt = min(waves, a.spp - s0);
      const float fs = (float)a.spp;
      for (int w = 0; w < cnt; ++w) {  // sample order, as a sequential loop adds them
        const float4 r = s_out[w][lane];
        acc += f3{r.x, r.y, r.z} / fs;
        dacc += r.w / fs;
      }
    }
    __syncthreads();
  }
  if (wave == 0 && valid) {
    // vignette, exposure, gamma 2 (integrator.cu:1415-1434): float uu, vv and radius
    const float uu = (float)(1.0 * u / a.W), vv = (float)(1.0 * v / a.H);
    const float rad = sqrtf((uu - 0.5f) * (uu - 0.5f) + (vv - 0.5f) * (vv - 0.5f));
    const float darken = (float)(1.0 - (double)(0.9f * fmaxf(rad - 0.f, 0.f)));
    const f3 c = acc * darken * 1.5f;
    a.color_out[3 * p] = sqrtf(c.x);
    a.color_out[3 * p + 1] = sqrtf(c.y);
    a.color_out[3 * p + 2] = sqrtf(c.z);
    a.depth_out[p] = dacc;
  }
}

inline unsigned blocks_for(long long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace
}  // namespace ddm_render

using namespace ddm_render;

#define DDM_API extern "C" __attribute__((visibility("default")))

DDM_API int particle_sdf(long long *volume, float *particle_x, int *particle_color, float *bbox_min, float *bbox_max,
                         int bake_size, const int *grid_dim, float inv_dx, float *sdf, float *color, float *sdf_tmp,
                         int n_particles, hipStream_t stream) {
  (void)bbox_max;
  if (!volume || !particle_x || !particle_color || !bbox_min || !grid_dim || !sdf || !color || !sdf_tmp)
    return DDM_EINVAL;
  const int gx = grid_dim[0], gy = grid_dim[1], gz = grid_dim[2];
  if (gx <= 0 || gy <= 0 || gz <= 0 || bake_size <= 0 || bake_size > 64 || n_particles < 0) return DDM_EINVAL;
  const long long dim = (long long)gx * gy * gz;
  if (dim >= (long long)INT32_MAX / 3) return DDM_EINVAL;
  const long long w2 = 4LL * bake_size * bake_size;
  if (n_particles > 0)
    k_bake<<<blocks_for(w2 * n_particles, 256), 256, 0, stream>>>(reinterpret_cast<unsigned long long *>(volume),
                                                                  particle_x, particle_color, bbox_min, bake_size, gx,
                                                                  gy, gz, inv_dx, n_particles);
  const unsigned nb = blocks_for(dim, 256);
  k_unpack<<<nb, 256, 0, stream>>>(reinterpret_cast<const unsigned long long *>(volume), sdf, color, (int)dim);
  k_smooth<<<nb, 256, 0, stream>>>(sdf, sdf_tmp, gx, gy, gz);
  k_smooth<<<nb, 256, 0, stream>>>(sdf_tmp, sdf, gx, gy, gz);
  return (int)hipGetLastError();
}

DDM_API int render(float *sdf_volume, float *box_min, float *box_max, float *color_volume, float *body_pos,
                   float *body_rot, float *body_type_friction_softness_round, float *body_args, float *camera_rot,
                   float *camera_pos, float *camera_intrinsic, float *color_buffer, float *depth_buffer,
                   float sdf_threshold, const int *grid_dim, int n_bodies, bool visualize_shape, const int *image_dim,
                   int max_ray_depth, int spp, float ground_height, int seed, float *light_direction,
                   hipStream_t stream) {
  if (!sdf_volume || !box_min || !box_max || !color_volume || !camera_rot || !camera_pos || !camera_intrinsic ||
      !color_buffer || !depth_buffer || !grid_dim || !image_dim || !light_direction)
    return DDM_EINVAL;
  if (n_bodies < 0 || (n_bodies > 0 && (!body_pos || !body_rot || !body_type_friction_softness_round || !body_args)))
    return DDM_EINVAL;
  if (n_bodies > kMaxBodies) return DDM_ETOOMANYBODIES;
  const int H = image_dim[0], W = image_dim[1];
  if (H <= 0 || W <= 0 || (long long)H * W >= (long long)INT32_MAX / 3 || spp <= 0 || max_ray_depth < 0)
    return DDM_EINVAL;
  if (grid_dim[0] <= 0 || grid_dim[1] <= 0 || grid_dim[2] <= 0) return DDM_EINVAL;
  if ((long long)grid_dim[0] * grid_dim[1] * grid_dim[2] >= (long long)INT32_MAX / 3) return DDM_EINVAL;
  RenderArgs a{sdf_volume, box_min, box_max, color_volume, body_pos, body_rot, body_type_friction_softness_round,
               body_args, camera_rot, camera_pos, camera_intrinsic, light_direction, color_buffer, depth_buffer,
               sdf_threshold, ground_height, grid_dim[0], grid_dim[1], grid_dim[2], n_bodies,
               visualize_shape ? 1 : 0, H, W, max_ray_depth, spp, seed, (W + kTile - 1) / kTile};
  const int tiles = a.tiles_u * ((H + kTile - 1) / kTile);
  const int waves = spp < kMaxSampleWaves ? spp : kMaxSampleWaves;
  k_render<<<tiles, 64 * waves, 0, stream>>>(a);
  return (int)hipGetLastError();
}

// create_volume / destroy_volume (integrator.cu:2074-2124, bound by mpm/types.py:121-124;
// nothing in the reference samples the volume). The reference uploads a float4 volume
// into a 3-D CUDA array with a linearly filtered texture object; gfx950 has no texture
// objects (hipMalloc3DArray and hipCreateTextureObject return hipErrorNotSupported), so
// the volume is a plain device copy (float4 per voxel, x fastest, then y, then z) and the
// texture handle is 0. Allocates, as the reference's does; on failure the handles are null.
DDM_API ddm_texture_resources create_volume(float *data, int x, int y, int z) {
  ddm_texture_resources r{nullptr, 0};
  if (!data || x <= 0 || y <= 0 || z <= 0) return r;
  const size_t bytes = (size_t)x * (size_t)y * (size_t)z * 4 * sizeof(float);
  void *d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return r;
  if (hipMemcpy(d, data, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return r;
  }
  r.volume = d;
  return r;
}

DDM_API int destroy_volume(ddm_texture_resources tex) { return tex.volume ? (int)hipFree(tex.volume) : 0; }
