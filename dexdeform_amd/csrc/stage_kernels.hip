// Per-stage kernels with the reference's launcher ABI (mpm/csrc/integrator.cu:1616-1985):
// one thread per particle or grid node, dense x-major grids, global atomics. These are
// the drop-in replacements of the reference's symbols; the fused multi-scene path lives
// in fused_kernels.hip and shares the device physics in mpm_device.h.
#include "mpm_device.h"
#include "../../include/dexdeform_mpm.h"

using namespace ddm;

namespace {

constexpr int kThreads = 256;
inline int nblk(int n) { return (n + kThreads - 1) / kThreads; }

__device__ __forceinline__ f3 grid_lower_offset(const int *gl, float dx) {
  return f3{(float)gl[0], (float)gl[1], (float)gl[2]} * dx;
}

struct DenseV3 {
  const float *v;
  DenseLayout L;
  __device__ f3 operator()(int x, int y, int z) const { return ld3(v + 3 * L.index(x, y, z)); }
};
struct DenseScatter3 {
  float *g;
  DenseLayout L;
  __device__ void operator()(int x, int y, int z, f3 val) const {
    float *p = g + 3 * L.index(x, y, z);
    atomicAdd(p + 0, val.x);
    atomicAdd(p + 1, val.y);
    atomicAdd(p + 2, val.z);
  }
};
struct DenseGatherMMV {
  const float *gmv, *gm;
  DenseLayout L;
  __device__ void operator()(int x, int y, int z, f3 &mv, float &m) const {
    int i = L.index(x, y, z);
    mv = ld3(gmv + 3 * i);
    m = gm[i];
  }
};

__global__ __launch_bounds__(kThreads) void k_compute_svd(const float *__restrict__ F, const float *__restrict__ C,
                                                          float *__restrict__ Ft, float *__restrict__ U,
                                                          float *__restrict__ V, float *__restrict__ sig, float dt,
                                                          int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  m33 f = advect_F(ld9(F + 9 * p), ld9(C + 9 * p), dt);
  m33 u, v;
  f3 s;
  svd3(f, u, s, v);
  st9(Ft + 9 * p, f);
  st9(U + 9 * p, u);
  st9(V + 9 * p, v);
  st3(sig + 3 * p, s);
}

__global__ __launch_bounds__(kThreads) void k_compute_svd_grad(
    const float *__restrict__ F, const float *__restrict__ C, const float *__restrict__ U,
    const float *__restrict__ V, const float *__restrict__ sig, float *__restrict__ Ft_grad,
    const float *__restrict__ U_grad, const float *__restrict__ V_grad, const float *__restrict__ sig_grad,
    float *__restrict__ F_grad, float *__restrict__ C_grad, float dt, int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  m33 gFt, gF, gC;
  svd_adj(ld9(F + 9 * p), ld9(C + 9 * p), ld9(U + 9 * p), ld9(V + 9 * p), ld3(sig + 3 * p), ld9(Ft_grad + 9 * p),
          ld9(U_grad + 9 * p), ld9(V_grad + 9 * p), ld3(sig_grad + 3 * p), dt, gFt, gF, gC);
  st9(Ft_grad + 9 * p, gFt);  // the reference keeps the total F_tmp gradient (integrator.cu:179)
  add9(C_grad + 9 * p, gC);
  add9(F_grad + 9 * p, gF);
}

__global__ __launch_bounds__(kThreads) void k_p2g(const float *__restrict__ x, const float *__restrict__ v,
                                                  const float *__restrict__ pm, const float *__restrict__ pvol,
                                                  const float *__restrict__ Ft, const float *__restrict__ U,
                                                  const float *__restrict__ sig, const float *__restrict__ V,
                                                  const float *__restrict__ C, const float *__restrict__ mly,
                                                  const int *__restrict__ gl, int3 dim, float dx, float inv_dx,
                                                  float dt, float *__restrict__ F_out, float *__restrict__ gmv,
                                                  float *__restrict__ gm, int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  int d[3] = {dim.x, dim.y, dim.z};
  DenseLayout L{{dim.x, dim.y, dim.z}};
  f3 xp = ld3(x + 3 * p) - grid_lower_offset(gl, dx);
  P2GParticle q;
  m33 Fn;
  p2g_particle(xp, ld3(v + 3 * p), ld9(C + 9 * p), ld9(Ft + 9 * p), ld9(U + 9 * p), ld3(sig + 3 * p),
               ld9(V + 9 * p), pm[p], pvol[p], ld3(mly + 3 * p), dx, inv_dx, dt, d, Fn, q);
  st9(F_out + 9 * p, Fn);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        float cm;
        f3 cmv;
        p2g_contrib(q, dx, i, j, k, cm, cmv);
        int gi = L.index(q.st.base[0] + i, q.st.base[1] + j, q.st.base[2] + k);
        atomicAdd(gm + gi, cm);
        atomicAdd(gmv + 3 * gi + 0, cmv.x);
        atomicAdd(gmv + 3 * gi + 1, cmv.y);
        atomicAdd(gmv + 3 * gi + 2, cmv.z);
      }
}

__global__ __launch_bounds__(kThreads) void k_p2g_grad(
    const float *__restrict__ x, const float *__restrict__ v, const float *__restrict__ pm,
    const float *__restrict__ pvol, const float *__restrict__ Ft, const float *__restrict__ U,
    const float *__restrict__ sig, const float *__restrict__ V, const float *__restrict__ C,
    const float *__restrict__ mly, const int *__restrict__ gl, int3 dim, float dx, float inv_dx, float dt,
    float *__restrict__ x_grad, float *__restrict__ v_grad, float *__restrict__ Ft_grad, float *__restrict__ C_grad,
    float *__restrict__ U_grad, float *__restrict__ sig_grad, float *__restrict__ V_grad,
    const float *__restrict__ Fn_grad, const float *__restrict__ gmv_grad, const float *__restrict__ gm_grad,
    int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  int d[3] = {dim.x, dim.y, dim.z};
  DenseGatherMMV g{gmv_grad, gm_grad, DenseLayout{{dim.x, dim.y, dim.z}}};
  P2GAdj a;
  f3 xp = ld3(x + 3 * p) - grid_lower_offset(gl, dx);
  p2g_particle_adj(xp, ld3(v + 3 * p), ld9(C + 9 * p), ld9(Ft + 9 * p), ld9(U + 9 * p), ld3(sig + 3 * p),
                   ld9(V + 9 * p), pm[p], pvol[p], ld3(mly + 3 * p), ld9(Fn_grad + 9 * p), g, dx, inv_dx, dt, d, a);
  add3(x_grad + 3 * p, a.x);
  add3(v_grad + 3 * p, a.v);
  add9(C_grad + 9 * p, a.C);
  add9(U_grad + 9 * p, a.U);
  add9(V_grad + 9 * p, a.V);
  add3(sig_grad + 3 * p, a.sig);
  add9(Ft_grad + 9 * p, a.Ft);
}

__global__ __launch_bounds__(kThreads) void k_grid_op(const float *__restrict__ gm, const float *__restrict__ gmv,
                                                      float *__restrict__ tape, const int *__restrict__ gl,
                                                      const float *__restrict__ gravity, NodeParams P,
                                                      float *__restrict__ gv_out, int G) {
  int tid = blockIdx.x * kThreads + threadIdx.x;
  if (tid >= G) return;
  float m = gm[tid];
  if (!node_active(m)) return;
  int ix = tid / P.dim[2] / P.dim[1], iy = (tid / P.dim[2]) % P.dim[1], iz = tid % P.dim[2];
  P.gravity = ld3(gravity);
  f3 v = ld3(gmv + 3 * tid) * (1.f / m);
  v += P.dt * P.gravity;
  f3 gx = grid_lower_offset(gl, P.dx) + node_position(ix, iy, iz, P.dx);
  for (int b = 0; b < P.nb; ++b) {
    if (tape) st3(tape + 3 * ((size_t)tid * (P.nb + 1) + b), v);
    contact(v, gx, load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b), P.dt);
  }
  if (tape) st3(tape + 3 * ((size_t)tid * (P.nb + 1) + P.nb), v);
  st3(gv_out + 3 * tid, boundary_fwd(v, ix, iy, iz, P));
}

__device__ __forceinline__ void atomic_pose(float *pos, float *rot, int b, f3 gp, q4 gq) {
  atomicAdd(pos + 3 * b + 0, gp.x);
  atomicAdd(pos + 3 * b + 1, gp.y);
  atomicAdd(pos + 3 * b + 2, gp.z);
  atomicAdd(rot + 4 * b + 0, gq.w);
  atomicAdd(rot + 4 * b + 1, gq.x);
  atomicAdd(rot + 4 * b + 2, gq.y);
  atomicAdd(rot + 4 * b + 3, gq.z);
}

__global__ __launch_bounds__(kThreads) void k_grid_op_grad(
    const float *__restrict__ gm, const float *__restrict__ gmv, const float *__restrict__ tape,
    const int *__restrict__ gl, const float *__restrict__ gravity, NodeParams P, float *__restrict__ gm_grad,
    float *__restrict__ gmv_grad, float *__restrict__ bpos_grad, float *__restrict__ brot_grad,
    float *__restrict__ nbpos_grad, float *__restrict__ nbrot_grad, const float *__restrict__ gv_out_grad, int G) {
  int tid = blockIdx.x * kThreads + threadIdx.x;
  if (tid >= G) return;
  float m = gm[tid];
  if (!node_active(m)) return;
  int ix = tid / P.dim[2] / P.dim[1], iy = (tid / P.dim[2]) % P.dim[1], iz = tid % P.dim[2];
  f3 mv = ld3(gmv + 3 * tid);
  f3 g = ld3(gv_out_grad + 3 * tid);
  const float *tp = tape + 3 * ((size_t)tid * (P.nb + 1));
  g = boundary_adj(ld3(tp + 3 * P.nb), g, ix, iy, iz, P, ADJ_REFERENCE);
  f3 gx = grid_lower_offset(gl, P.dx) + node_position(ix, iy, iz, P.dx);
  for (int b = P.nb - 1; b >= 0; --b) {
    Body B = load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b);
    f3 pb = to_body(B.pos, B.rot, gx);
    float infl;
    if (!in_contact(prim_sdf(B.prim, B.args, pb), B.prim[2], infl)) continue;
    PoseGrad pg = posegrad_zero();
    contact_adj(ld3(tp + 3 * b), gx, B, P.dt, ADJ_REFERENCE, g, pg);
    atomic_pose(nbpos_grad, nbrot_grad, b, pg.npos, pg.nrot);
    atomic_pose(bpos_grad, brot_grad, b, pg.pos, pg.rot);
  }
  add3(gmv_grad + 3 * tid, (float)(1.0 / (double)m) * g);
  gm_grad[tid] += (-1.f / m / m) * dot3(mv, g);
}

__global__ __launch_bounds__(kThreads) void k_g2p(const float *__restrict__ x, const float *__restrict__ gv,
                                                  const int *__restrict__ gl, float dx, float inv_dx, float dt,
                                                  int3 dim, float *__restrict__ v_out, float ground_height,
                                                  float *__restrict__ C_out, float *__restrict__ x_out, int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  int d[3] = {dim.x, dim.y, dim.z};
  DenseV3 g{gv, DenseLayout{{dim.x, dim.y, dim.z}}};
  f3 lower = grid_lower_offset(gl, dx);
  f3 xo, vo;
  m33 Co;
  g2p_particle(ld3(x + 3 * p) - lower, g, inv_dx, dx, dt, ground_height, d, xo, vo, Co);
  st3(x_out + 3 * p, xo + lower);
  st3(v_out + 3 * p, vo);
  st9(C_out + 9 * p, Co);
}

__global__ __launch_bounds__(kThreads) void k_g2p_grad(const float *__restrict__ x, const float *__restrict__ gv,
                                                       const int *__restrict__ gl, float dx, float inv_dx, float dt,
                                                       int3 dim, const float *__restrict__ v_next,
                                                       float ground_height, float *__restrict__ x_grad,
                                                       float *__restrict__ gv_grad,
                                                       const float *__restrict__ v_next_grad,
                                                       const float *__restrict__ C_next_grad,
                                                       const float *__restrict__ x_next_grad, int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  int d[3] = {dim.x, dim.y, dim.z};
  DenseLayout L{{dim.x, dim.y, dim.z}};
  DenseV3 g{gv, L};
  DenseScatter3 sc{gv_grad, L};
  f3 gx = g2p_particle_adj(ld3(x + 3 * p) - grid_lower_offset(gl, dx), ld3(v_next + 3 * p), ld3(x_next_grad + 3 * p),
                           ld3(v_next_grad + 3 * p), ld9(C_next_grad + 9 * p), g, sc, inv_dx, dx, dt, ground_height,
                           d);
  add3(x_grad + 3 * p, gx);
}

__global__ __launch_bounds__(kThreads) void k_compute_dist(const float *__restrict__ x,
                                                           const float *__restrict__ bpos,
                                                           const float *__restrict__ brot,
                                                           const float *__restrict__ prim,
                                                           const float *__restrict__ args, float *__restrict__ dist,
                                                           int nb, float *__restrict__ x_grad,
                                                           float *__restrict__ bpos_grad,
                                                           float *__restrict__ brot_grad,
                                                           const float *__restrict__ dist_grad, int compute_grad,
                                                           int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  f3 xp = ld3(x + 3 * p);
  for (int b = 0; b < nb; ++b) {
    f3 bx = ld3(bpos + 3 * b);
    q4 bq = ldq(brot + 4 * b);
    f3 pb = to_body(bx, bq, xp);
    if (!compute_grad) {
      dist[(size_t)p * nb + b] = prim_sdf(prim + 4 * b, args + 4 * b, pb);
    } else {
      float gd = dist_grad[(size_t)p * nb + b];
      f3 g_pb = prim_grad(prim + 4 * b, args + 4 * b, pb) * gd;
      f3 g_bx{0.f, 0.f, 0.f}, g_x{0.f, 0.f, 0.f};
      q4 g_bq{0.f, 0.f, 0.f, 0.f};
      to_body_adj(bx, bq, xp, g_pb, g_bx, g_bq, g_x);
      add3(x_grad + 3 * p, g_x);
      atomic_pose(bpos_grad, brot_grad, b, g_bx, g_bq);
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_particle2mass(const float *__restrict__ x, const float *__restrict__ pm,
                                                            const int *__restrict__ gl, int3 dim, float dx,
                                                            float inv_dx, float *__restrict__ gm,
                                                            const float *__restrict__ gm_grad,
                                                            float *__restrict__ x_grad, const int *__restrict__ ids,
                                                            int id, int compute_grad, int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  if (id != -1 && ids[p] != id) return;
  int d[3] = {dim.x, dim.y, dim.z};
  DenseLayout L{{dim.x, dim.y, dim.z}};
  Stencil s;
  stencil(ld3(x + 3 * p) - grid_lower_offset(gl, dx), inv_dx, d, s);
  if (compute_grad) {
    float w1[3][3];
    stencil_dw(s, inv_dx, w1);
    f3 g{0.f, 0.f, 0.f};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          g += gm_grad[L.index(s.base[0] + i, s.base[1] + j, s.base[2] + k)] * pm[p] * weight_grad(s, w1, i, j, k);
    add3(x_grad + 3 * p, g);
  } else {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          float w = s.w[0][i] * s.w[1][j] * s.w[2][k];
          atomicAdd(gm + L.index(s.base[0] + i, s.base[1] + j, s.base[2] + k), pm[p] * w);
        }
  }
}

inline int status() { return (int)hipGetLastError(); }
inline bool dims_ok(const int *d) { return d && d[0] >= 3 && d[1] >= 3 && d[2] >= 3; }

}  // namespace

#define DDM_API extern "C" __attribute__((visibility("default")))

DDM_API int compute_svd(float *F, float *C, float *newF, float *U, float *V, float *sig, float dt, int dim,
                        hipStream_t stream) {
  if (dim < 0) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_compute_svd<<<nblk(dim), kThreads, 0, stream>>>(F, C, newF, U, V, sig, dt, dim);
  return status();
}

DDM_API int compute_svd_grad(float *F, float *C, float *U, float *V, float *sig, float *newF_grad, float *U_grad,
                             float *V_grad, float *sig_grad, float *F_grad, float *C_grad, float dt, int dim,
                             hipStream_t stream) {
  if (dim < 0) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_compute_svd_grad<<<nblk(dim), kThreads, 0, stream>>>(F, C, U, V, sig, newF_grad, U_grad, V_grad, sig_grad,
                                                         F_grad, C_grad, dt, dim);
  return status();
}

DDM_API int p2g(float *particle_x, float *particle_v, float *particle_m, float *particle_vol, float *particle_F,
                float *particle_U, float *particle_sig, float *particle_V, float *particle_C,
                float *particle_mu_lam_yield, int *grid_lower, const int *grid_dim, float dx, float inv_dx, float dt,
                float *out_particle_F, float *out_grid_mv, float *out_grid_m, int dim, hipStream_t stream) {
  if (dim < 0 || !dims_ok(grid_dim)) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_p2g<<<nblk(dim), kThreads, 0, stream>>>(particle_x, particle_v, particle_m, particle_vol, particle_F, particle_U,
                                            particle_sig, particle_V, particle_C, particle_mu_lam_yield, grid_lower,
                                            make_int3(grid_dim[0], grid_dim[1], grid_dim[2]), dx, inv_dx, dt,
                                            out_particle_F, out_grid_mv, out_grid_m, dim);
  return status();
}

DDM_API int p2g_grad(float *particle_x, float *particle_v, float *particle_m, float *particle_vol, float *particle_F,
                     float *particle_U, float *particle_sig, float *particle_V, float *particle_C,
                     float *particle_mu_lam_yield, int *grid_lower, const int *grid_dim, float dx, float inv_dx,
                     float dt, float *out_particle_F, float *out_grid_mv, float *out_grid_m, float *particle_x_grad,
                     float *particle_v_grad, float *particle_F_grad, float *particle_C_grad, float *particle_U_grad,
                     float *particle_sig_grad, float *particle_V_grad, float *out_particle_F_grad,
                     float *out_grid_v_grad, float *out_grid_m_grad, int dim, hipStream_t stream) {
  (void)out_particle_F;
  (void)out_grid_mv;
  (void)out_grid_m;
  if (dim < 0 || !dims_ok(grid_dim)) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_p2g_grad<<<nblk(dim), kThreads, 0, stream>>>(
      particle_x, particle_v, particle_m, particle_vol, particle_F, particle_U, particle_sig, particle_V, particle_C,
      particle_mu_lam_yield, grid_lower, make_int3(grid_dim[0], grid_dim[1], grid_dim[2]), dx, inv_dx, dt,
      particle_x_grad, particle_v_grad, particle_F_grad, particle_C_grad, particle_U_grad, particle_sig_grad,
      particle_V_grad, out_particle_F_grad, out_grid_v_grad, out_grid_m_grad, dim);
  return status();
}

static NodeParams node_params(const int *grid_dim, int nb, float dx, float dt, float gf, float *body_pos,
                              float *body_rot, float *next_body_pos, float *next_body_rot, float *prim, float *args) {
  NodeParams P;
  P.dim[0] = grid_dim[0];
  P.dim[1] = grid_dim[1];
  P.dim[2] = grid_dim[2];
  P.nb = nb;
  P.dx = dx;
  P.dt = dt;
  P.gf = gf;
  P.gravity = f3{0.f, 0.f, 0.f};
  P.prim = prim;
  P.args = args;
  P.pos = body_pos;
  P.rot = body_rot;
  P.npos = next_body_pos;
  P.nrot = next_body_rot;
  return P;
}

DDM_API int grid_op_v2(float *grid_m, float *grid_v_in, float *grid_body_v_in, int *grid_lower, float *gravity,
                       float *body_pos, float *body_rot, float *next_body_pos, float *next_body_rot,
                       float *body_type_friction_softness_round, float *body_args, float dx, float inv_dx, float dt,
                       float ground_friction, float *out_grid_v, const int *grid_dim, int n_bodies,
                       hipStream_t stream) {
  (void)inv_dx;
  if (!dims_ok(grid_dim) || n_bodies < 0) return DDM_EINVAL;
  int G = grid_dim[0] * grid_dim[1] * grid_dim[2];
  NodeParams P = node_params(grid_dim, n_bodies, dx, dt, ground_friction, body_pos, body_rot, next_body_pos,
                             next_body_rot, body_type_friction_softness_round, body_args);
  k_grid_op<<<nblk(G), kThreads, 0, stream>>>(grid_m, grid_v_in, grid_body_v_in, grid_lower, gravity, P, out_grid_v,
                                              G);
  return status();
}

DDM_API int grid_op_v2_grad(float *grid_m, float *grid_v_in, float *grid_body_v_in, int *grid_lower, float *gravity,
                            float *body_pos, float *body_rot, float *next_body_pos, float *next_body_rot,
                            float *body_type_friction_softness_round, float *body_args, float *grid_m_grad,
                            float *grid_v_in_grad, float *body_pos_grad, float *body_rot_grad,
                            float *next_body_pos_grad, float *next_body_rot_grad, float dx, float inv_dx, float dt,
                            float ground_friction, float *out_grid_v, float *out_grid_v_grad, const int *grid_dim,
                            int n_bodies, hipStream_t stream) {
  (void)inv_dx;
  (void)out_grid_v;
  if (!dims_ok(grid_dim) || n_bodies < 0 || !grid_body_v_in) return DDM_EINVAL;
  int G = grid_dim[0] * grid_dim[1] * grid_dim[2];
  NodeParams P = node_params(grid_dim, n_bodies, dx, dt, ground_friction, body_pos, body_rot, next_body_pos,
                             next_body_rot, body_type_friction_softness_round, body_args);
  k_grid_op_grad<<<nblk(G), kThreads, 0, stream>>>(grid_m, grid_v_in, grid_body_v_in, grid_lower, gravity, P,
                                                   grid_m_grad, grid_v_in_grad, body_pos_grad, body_rot_grad,
                                                   next_body_pos_grad, next_body_rot_grad, out_grid_v_grad, G);
  return status();
}

DDM_API int g2p(float *particle_x, float *grid_v, int *grid_lower, float dx, float inv_dx, float dt,
                const int *grid_dim, float *out_particle_v, float ground_height, float *out_particle_C,
                float *out_particle_x, int dim, hipStream_t stream) {
  if (dim < 0 || !dims_ok(grid_dim)) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_g2p<<<nblk(dim), kThreads, 0, stream>>>(particle_x, grid_v, grid_lower, dx, inv_dx, dt,
                                            make_int3(grid_dim[0], grid_dim[1], grid_dim[2]), out_particle_v,
                                            ground_height, out_particle_C, out_particle_x, dim);
  return status();
}

DDM_API int g2p_grad(float *particle_x, float *grid_v, int *grid_lower, float dx, float inv_dx, float dt,
                     const int *grid_dim, float *out_particle_v, float ground_height, float *out_particle_C,
                     float *out_particle_x, int dim, float *particle_x_grad, float *grid_v_grad,
                     float *out_particle_v_grad, float *out_particle_C_grad, float *out_particle_x_grad,
                     hipStream_t stream) {
  (void)out_particle_C;
  (void)out_particle_x;
  if (dim < 0 || !dims_ok(grid_dim)) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_g2p_grad<<<nblk(dim), kThreads, 0, stream>>>(particle_x, grid_v, grid_lower, dx, inv_dx, dt,
                                                 make_int3(grid_dim[0], grid_dim[1], grid_dim[2]), out_particle_v,
                                                 ground_height, particle_x_grad, grid_v_grad, out_particle_v_grad,
                                                 out_particle_C_grad, out_particle_x_grad, dim);
  return status();
}

DDM_API int compute_dist(float *particle_x, float *body_pos, float *body_rot, float *body_type_friction_softness_round,
                         float *body_args, float *dist, int n_bodies, float *particle_x_grad, float *body_pos_grad,
                         float *body_rot_grad, float *dist_grad, int compute_grad, int dim, hipStream_t stream) {
  if (dim < 0 || n_bodies < 0) return DDM_EINVAL;
  if (dim == 0 || n_bodies == 0) return 0;
  k_compute_dist<<<nblk(dim), kThreads, 0, stream>>>(particle_x, body_pos, body_rot,
                                                     body_type_friction_softness_round, body_args, dist, n_bodies,
                                                     particle_x_grad, body_pos_grad, body_rot_grad, dist_grad,
                                                     compute_grad, dim);
  return status();
}

DDM_API int particle2mass(float *particle_x, float *particle_m, int *grid_lower, const int *grid_dim, float dx,
                          float inv_dx, float *out_grid_m, float *out_grid_m_grad, float *particle_x_grad,
                          int *particle_ids, int id, int compute_grad, int dim, hipStream_t stream) {
  if (dim < 0 || !dims_ok(grid_dim)) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_particle2mass<<<nblk(dim), kThreads, 0, stream>>>(particle_x, particle_m, grid_lower,
                                                      make_int3(grid_dim[0], grid_dim[1], grid_dim[2]), dx, inv_dx,
                                                      out_grid_m, out_grid_m_grad, particle_x_grad, particle_ids, id,
                                                      compute_grad, dim);
  return status();
}

DDM_API int ddm_version(void) { return 100; }
