// Fused multi-scene substep for gfx950: forward kernels (and the adjoint pass's
// recompute and restore kernels), launch sequences and the exported entry points. The
// layout and invariants are described in fused_common.h.
#include <mutex>

#include "fused_common.h"

namespace ddm_fused {
namespace {

// ---------------------------------------------------------------- P2G (+ F advection, SVD, return map)
enum P2GMode { kForward = 0, kRecompute = 1, kForwardTaped = 2, kRecomputeUntaped = 3 };
// kForward         : forward substep, writes next.F.
// kRecompute       : adjoint recompute, writes the F_tmp/U/V/sigma temporaries instead.
// kForwardTaped    : writes next.F, and the temporaries of scenes that keep a snapshot
//                    (the taped adjoint reads them).
// kRecomputeUntaped: kRecompute for the scenes whose snapshot was not restored (no
//                    snapshot, or tape_status[1] set by k_restore); other chunks exit.
// Taped forward: workgroup 0 of the P2G kernel empties every snapshot of the launch
// (scenes without particles included, so that no stale snapshot survives) before the
// grid update, the next kernel, fills them.
__device__ __forceinline__ void tape_reset(const ddm_batch &B) {
  for (int si = threadIdx.x; si < live_scenes(B); si += kWG)
    if (B.scenes[si].tape_mmv) B.scenes[si].tape_status[0] = 0;
}

// at most 96 VGPRs (a few bytes of spill): five waves per SIMD, as the LDS allows
// AHEAD (small forward launches): each particle first runs the G2P of the substep before
// (prev_tab -> cur_tab; the grid update of that substep has finished) and P2G then uses
// the x, v, C it computed instead of reading them back: one launch fewer per substep.
template <int MODE, bool AHEAD = false>
__global__ __launch_bounds__(kWG, AHEAD ? 1 : 5) void k_p2g(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                                          const ddm_state *__restrict__ next_tab, int scan,
                                                          const ddm_state *__restrict__ prev_tab = nullptr) {
  static_assert(!AHEAD || MODE == kForward || MODE == kForwardTaped, "G2P ahead of forward P2G only");
  __shared__ ScatterLDS<4, kTileP2G> sh;
  __shared__ StencilRecs rc;
  __shared__ Groups gr;
  __shared__ int s_box[6];
  float *const tm = sh.tile[0], *const tx = sh.tile[1], *const ty = sh.tile[2], *const tz = sh.tile[3];
  const int ci = chunk_index(B);
  if (ci < 0) {  // uniform: a fixed-size launch with fewer chunks
    if (MODE == kForwardTaped && blockIdx.x == 0) tape_reset(B);
    return;
  }
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[ci];
  const ddm_scene S = B.scenes[ch.x];
  if (MODE == kRecomputeUntaped && S.tape_mmv && S.tape_status[1] == 0) return;  // uniform: one scene per chunk
  const bool temps = MODE == kRecompute || MODE == kRecomputeUntaped || (MODE == kForwardTaped && S.tape_mmv);
  const ddm_state cur = cur_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  const bool valid = p < ch.z;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  P2GParticle q;
  int key = 0;
  if (valid) {
    f3 xa, va;
    m33 Ca;
    if (AHEAD) {  // G2P of the substep before, as k_g2p computes it
      BlockedV gv{reinterpret_cast<const float4 *>(B.grid_v), &S};
      g2p_particle_prefetch(ld3(prev_tab[ch.x].x + 3 * (size_t)p), gv, S.inv_dx, S.dx, S.dt, S.ground_height, dim,
                            xa, va, Ca);
      st3(cur.x + 3 * (size_t)p, xa);
      st3(cur.v + 3 * (size_t)p, va);
      st9(cur.C + 9 * (size_t)p, Ca);
    }
    const float *Cp = cur.C + 9 * (size_t)p;
    m33 F = ld9(cur.F + 9 * (size_t)p);
    m33 Ft = advect_F(F, AHEAD ? Ca : ld9(Cp), S.dt);
    m33 U, V;
    f3 sig;
    q4 qu, qv;
    svd3(Ft, U, sig, V, qu, qv);
    if (temps) {  // stored before the stress so that their registers can retire early
      stq(S.tmp_qu + 4 * (size_t)p, qu);
      stq(S.tmp_qv + 4 * (size_t)p, qv);
      st3(S.tmp_sig + 3 * (size_t)p, sig);
    }
    if (!AHEAD) asm volatile("" : "+v"(Cp));  // C is read again (from the cache) rather than held across the SVD
    m33 Fn;
    p2g_particle(AHEAD ? xa : ld3(cur.x + 3 * (size_t)p), AHEAD ? va : ld3(cur.v + 3 * (size_t)p),
                 AHEAD ? Ca : ld9(Cp), Ft, U, sig, V, S.mass[p], S.vol[p], ld3(S.mu_lam_yield + 3 * (size_t)p), S.dx,
                 S.inv_dx, S.dt, dim, Fn, q);
    if (MODE == kForward || MODE == kForwardTaped) st9(next_tab[ch.x].F + 9 * (size_t)p, Fn);
    key = pack_base(q.st.base);
  } else {
    q.st.base[0] = q.st.base[1] = q.st.base[2] = 0;
  }
  const Tile T = chunk_tile(valid, q.st.base, s_box);
  int pos;
  const int G = build_groups(valid, key, min(kWG, ch.z - ch.y), sh.grp, gr, pos);
  if (valid) stage_record(rc, pos, q.st, q.m, q.mv, q.affine);
  const float dx = S.dx;
  // contributions of record s to the nodes (i, j, 0..2): w (m, mv + A d) with the
  // affine matrix A = {b.w, c.x, c.y | c.z, c.w, d.x | d.y, d.z, d.w} and d = (i, j, k) dx
  // - fx dx (p2g_node_contrib's terms), the k-independent parts once per column and the
  // products accumulated with fused multiply-adds: w (mv + A_xy d_xy) + (w d_z) A_z
  auto contrib = [dx](const StencilRecs &r, int s, int i, int j, float (*acc)[4]) {
    const float4 a = r.f0[s], b = r.f1[s], c = r.f2[s], d = r.f3[s];
    const float wxy = bspline_wi(a.x, i) * bspline_wi(a.y, j);
    const float wz[3] = {bspline_w<0>(a.z), bspline_w<1>(a.z), bspline_w<2>(a.z)};
    const float d0 = ((float)i - a.x) * dx, d1 = ((float)j - a.y) * dx;
    const float s0 = b.x + fmaf(c.x, d1, b.w * d0), s1 = b.y + fmaf(c.w, d1, c.z * d0),
                s2 = b.z + fmaf(d.z, d1, d.y * d0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float w = wxy * wz[k];
      const float wd = w * (((float)k - a.z) * dx);
      acc[k][0] = fmaf(a.w, w, acc[k][0]);
      acc[k][1] = fmaf(s0, w, fmaf(c.y, wd, acc[k][1]));
      acc[k][2] = fmaf(s1, w, fmaf(d.x, wd, acc[k][2]));
      acc[k][3] = fmaf(s2, w, fmaf(d.w, wd, acc[k][3]));
    }
  };
  float *gmmv = B.grid_mmv;
  if (slabs_fit<kTileP2G>(T)) {
    const int sx = slab_width<kTileP2G>(T);
    for (int xs = T.lo[0]; xs < T.lo[0] + T.t[0]; xs += sx) {  // uniform over the workgroup
      const Tile L = slab(T, xs, sx);
      if (xs != T.lo[0]) __syncthreads();  // the previous pass has flushed the tile
      for (int i = threadIdx.x; i < L.n; i += kWG) tm[i] = tx[i] = ty[i] = tz[i] = 0.f;
      __syncthreads();  // also publishes the staged records
      group_column_tasks<4>(G, gr, rc, &L, contrib, [&](int X, int Y, int Z, const float *a) {
        const int li = L.local(X, Y, Z);
        atomicAdd(&tm[li], a[0]);
        atomicAdd(&tx[li], a[1]);
        atomicAdd(&ty[li], a[2]);
        atomicAdd(&tz[li], a[3]);
      });
      __syncthreads();
      for (int i = threadIdx.x; i < L.n; i += kWG) {
        float a = tm[i], b = tx[i], c = ty[i], d = tz[i];
        if (a != 0.f || b != 0.f || c != 0.f || d != 0.f) {
          int x = L.lo[0] + i / (L.t[1] * L.t[2]), y = L.lo[1] + (i / L.t[2]) % L.t[1], z = L.lo[2] + i % L.t[2];
          float *g = gmmv + 4 * (size_t)node_index(S, x, y, z);
          atomicAdd(g + 0, a);
          atomicAdd(g + 1, b);
          atomicAdd(g + 2, c);
          atomicAdd(g + 3, d);
        }
      }
    }
    if (scan) flag_tile_blocks_det(B, S, ch.x, T);  // uniform; the same plain stores of scene + 1
    else flag_tile_blocks(B, S, ch.x, T);
  } else {
    __syncthreads();  // publishes the staged records
    group_column_tasks<4>(G, gr, rc, nullptr, contrib, [&](int X, int Y, int Z, const float *a) {
      float *g = gmmv + 4 * (size_t)node_index(S, X, Y, Z);
      atomicAdd(g + 0, a[0]);
      atomicAdd(g + 1, a[1]);
      atomicAdd(g + 2, a[2]);
      atomicAdd(g + 3, a[3]);
    });
    for (int g = threadIdx.x; g < G; g += kWG) {
      int b[3];
      unpack_base(gr.key[g], b);
      if (scan) store_stencil_flags(B, S, ch.x, b);
      else flag_stencil_blocks(B, S, ch.x, b);
    }
  }
  // at the end, where few registers are live (at the start it costs the kernel spills)
  if (MODE == kForwardTaped && blockIdx.x == 0) tape_reset(B);
}

// ---------------------------------------------------------------- grid update
// ZERO=true (forward substep): consumes (m, mv) and clears them and the flags.
// SAVE=true (taped forward): also copies each listed block's (m, mv) and v into its
//   scene's snapshot, if it keeps one (blocks beyond its capacity are dropped; the count in
//   tape_status[0] keeps growing, so the taped adjoint sees the overflow).
// SKIP_RESTORED=true (taped adjoint's recompute): skips the blocks of scenes whose
//   snapshot was restored.
// clear_count (deterministic mode, forward): the bin counts of the listed blocks are cleared
// with the grid (the sorted gather was their last reader).
// SCAN=true (forward, small launches): no block list; every block of the launch is
//   visited and those whose flag P2G set (to scene + 1) are updated (ZERO only).
template <bool ZERO, bool SAVE, bool SKIP_RESTORED, bool SCAN = false>
__global__ __launch_bounds__(kWG) void k_grid(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                              const ddm_state *__restrict__ next_tab, int32_t *clear_count) {
  static_assert(ZERO || !SCAN, "the scan clears the flags it reads");
  __shared__ WaveBodies s_bodies[kWG / 64];
  const int wave = wave_id(), lane = threadIdx.x & 63;
  float4 *mmv = reinterpret_cast<float4 *>(B.grid_mmv);
  float4 *gv = reinterpret_cast<float4 *>(B.grid_v);
  const int n_items = SCAN ? B.n_blocks : *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  for (int e = blockIdx.x * (kWG / 64) + wave; e < n_items; e += gridDim.x * (kWG / 64)) {
    int blk, si;
    if (SCAN) {
      const uint32_t f = B.block_flags[e];
      if (f == 0u) continue;  // the same on every lane
      blk = e;
      si = (int)f - 1;
    } else {
      const int2 entry = list[e];
      blk = entry.x;
      si = entry.y;
    }
    const ddm_scene &S = B.scenes[si];
    if (SKIP_RESTORED && S.tape_mmv && S.tape_status[1] == 0) continue;
    BlockCoord c = block_node(S, blk, lane);
    size_t gi = ((size_t)blk << 6) | (size_t)lane;
    float4 a = mmv[gi];
    f3 v{0.f, 0.f, 0.f};
    NodeParams P = scene_node_params(S, cur_tab[si], next_tab[si]);
    stage_bodies(P, s_bodies[wave], lane);
    const unsigned long long near = near_bodies(P, c.ix, c.iy, c.iz, lane);
    if (node_active(a.x)) v = node_fwd(a.x, f3{a.y, a.z, a.w}, c.ix, c.iy, c.iz, P, near);
    gv[gi] = make_float4(v.x, v.y, v.z, 0.f);
    if (SAVE && S.tape_mmv) {
      int slot = 0;
      if (lane == 0) slot = atomicAdd(S.tape_status, 1);
      slot = __shfl(slot, 0, 64);
      if (slot < S.tape_capacity) {
        const size_t ti = ((size_t)slot << 6) | (size_t)lane;
        reinterpret_cast<float4 *>(S.tape_mmv)[ti] = a;
        reinterpret_cast<float4 *>(S.tape_v)[ti] = make_float4(v.x, v.y, v.z, 0.f);
        if (lane == 0) S.tape_list[slot] = blk;
      }
    }
    if (ZERO) {
      mmv[gi] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lane == 0) B.block_flags[blk] = 0u;
      if (clear_count) clear_count[gi] = 0;
    }
  }
}


__global__ __launch_bounds__(kWG) void k_restore(ddm_batch B) {
  restore_scene_part(B, (int)blockIdx.y, (int)blockIdx.x, (int)gridDim.x);
}

// ---------------------------------------------------------------- G2P
__global__ __launch_bounds__(kWG) void k_g2p(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                             const ddm_state *__restrict__ next_tab) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.block_count = 0;  // the grid kernels have finished
  const int ci = chunk_index(B);
  if (ci < 0) return;  // uniform: a fixed-size launch with fewer chunks
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[ci];
  const ddm_scene S = B.scenes[ch.x];
  const ddm_state cur = cur_tab[ch.x], next = next_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  if (p >= ch.z) return;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  BlockedV gv{reinterpret_cast<const float4 *>(B.grid_v), &S};
  f3 xo, vo;
  m33 Co;
  g2p_particle_prefetch(ld3(cur.x + 3 * (size_t)p), gv, S.inv_dx, S.dx, S.dt, S.ground_height, dim, xo, vo, Co);
  st3(next.x + 3 * (size_t)p, xo);
  st3(next.v + 3 * (size_t)p, vo);
  st9(next.C + 9 * (size_t)p, Co);
}

// ---------------------------------------------------------------- deterministic mode
// Per particle: k_p2g's particle stage (same modes and outputs), then its stencil record
// (fx, m | m v, affine) in its launch slot, its base node and its arrival rank among the
// node's particles; the chunk's blocks are flagged with scene + 1. A restored scene of
// the taped adjoint is only binned: its grid comes from the snapshot, but the G2P
// adjoint's sorted gather needs the bins.
template <int MODE>
__global__ __launch_bounds__(kWG) void k_det_prep(ddm_batch B, ddm_det D, const ddm_state *__restrict__ cur_tab,
                                                  const ddm_state *__restrict__ next_tab) {
  __shared__ int s_box[6];
  if (MODE == kForwardTaped && blockIdx.x == 0) tape_reset(B);
  const int ci = chunk_index(B);
  if (ci < 0) return;  // uniform: a fixed-size launch with fewer chunks
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[ci];
  const ddm_scene S = B.scenes[ch.x];
  const bool bin_only = MODE == kRecomputeUntaped && S.tape_mmv && S.tape_status[1] == 0;  // uniform
  const bool temps = MODE == kRecompute || MODE == kRecomputeUntaped || (MODE == kForwardTaped && S.tape_mmv);
  const ddm_state cur = cur_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  const bool valid = p < ch.z;
  const size_t slot = (size_t)ci * kWG + threadIdx.x;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  Stencil st;
  if (valid) {
    const f3 x = ld3(cur.x + 3 * (size_t)p);
    if (bin_only) {
      stencil(x, S.inv_dx, dim, st);
    } else {
      m33 F = ld9(cur.F + 9 * (size_t)p), C = ld9(cur.C + 9 * (size_t)p);
      m33 Ft = advect_F(F, C, S.dt);
      m33 U, V;
      f3 sig;
      q4 qu, qv;
      svd3(Ft, U, sig, V, qu, qv);
      m33 Fn;
      P2GParticle q;
      p2g_particle(x, ld3(cur.v + 3 * (size_t)p), C, Ft, U, sig, V, S.mass[p], S.vol[p],
                   ld3(S.mu_lam_yield + 3 * (size_t)p), S.dx, S.inv_dx, S.dt, dim, Fn, q);
      if (temps) {
        stq(S.tmp_qu + 4 * (size_t)p, qu);
        stq(S.tmp_qv + 4 * (size_t)p, qv);
        st3(S.tmp_sig + 3 * (size_t)p, sig);
      }
      if (MODE == kForward || MODE == kForwardTaped) st9(next_tab[ch.x].F + 9 * (size_t)p, Fn);
      st = q.st;
      const m33 &A = q.affine;
      float4 *r = reinterpret_cast<float4 *>(D.records) + 4 * slot;
      r[0] = make_float4(st.fx[0], st.fx[1], st.fx[2], q.m);
      r[1] = make_float4(q.mv.x, q.mv.y, q.mv.z, A.a[0]);
      r[2] = make_float4(A.a[1], A.a[2], A.a[3], A.a[4]);
      r[3] = make_float4(A.a[5], A.a[6], A.a[7], A.a[8]);
    }
    const int key = node_index(S, st.base[0], st.base[1], st.base[2]);
    const int rank = atomicAdd(&D.cell_count[key], 1);
    reinterpret_cast<int2 *>(D.key_slot)[slot] = make_int2(key, rank);
  } else {
    st.base[0] = st.base[1] = st.base[2] = 0;
  }
  const Tile T = chunk_tile(valid, st.base, s_box);
  flag_tile_blocks_det(B, S, ch.x, T);
}

// Block list rebuilt from the flags in block order: flagged blocks per range of kRange
// blocks, then each range writes its (block, scene) entries at the sum of the counts of
// the ranges before it. Grid: one workgroup per range.
constexpr int kRange = 4 * kWG;
__device__ __forceinline__ int wg_sum_i(int v, int *s_w) {
  v = wave_sum_i(v);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0;
#pragma unroll
  for (int w = 0; w < kWG / 64; ++w) t += s_w[w];
  __syncthreads();
  return t;
}
__global__ __launch_bounds__(kWG) void k_det_count(ddm_batch B, ddm_det D) {
  __shared__ int s_w[kWG / 64];
  const int b0 = (int)blockIdx.x * kRange + 4 * (int)threadIdx.x;
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) c += (b0 + k < B.n_blocks && B.block_flags[b0 + k] != 0u) ? 1 : 0;
  const int t = wg_sum_i(c, s_w);
  if (threadIdx.x == 0) D.range_count[blockIdx.x] = t;
}
__global__ __launch_bounds__(kWG) void k_det_relist(ddm_batch B, ddm_det D) {
  __shared__ int s_w[kWG / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int before = 0, total = 0;
  for (int r = threadIdx.x; r < (int)gridDim.x; r += kWG) {
    const int c = D.range_count[r];
    total += c;
    before += r < (int)blockIdx.x ? c : 0;
  }
  before = wg_sum_i(before, s_w);
  total = wg_sum_i(total, s_w);
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.block_count = total;
  const int b0 = (int)blockIdx.x * kRange + 4 * (int)threadIdx.x;
  uint32_t f[4];
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[k] = b0 + k < B.n_blocks ? B.block_flags[b0 + k] : 0u;
    c += f[k] != 0u ? 1 : 0;
  }
  int incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  int off = before + incl - c;
#pragma unroll
  for (int w = 0; w < kWG / 64; ++w) off += w < wave ? s_w[w] : 0;
  int2 *list = reinterpret_cast<int2 *>(B.block_list);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (f[k] != 0u) list[off++] = make_int2(b0 + k, (int)f[k] - 1);
}

// Sorted segments: each listed block takes a contiguous range of positions for its 64
// nodes' particles (ranges in any order; the order within a node is fixed afterwards).
__global__ __launch_bounds__(kWG) void k_det_scan(ddm_batch B, ddm_det D) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n_listed = *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  for (int e = blockIdx.x * (kWG / 64) + wave; e < n_listed; e += gridDim.x * (kWG / 64)) {
    const size_t gi = ((size_t)list[e].x << 6) | (size_t)lane;
    const int c = D.cell_count[gi];
    int incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    const int tot = __shfl(incl, 63, 64);
    int base = 0;
    if (lane == 0 && tot > 0) base = atomicAdd(D.counter, tot);
    base = __shfl(base, 0, 64);
    D.cell_start[gi] = base + incl - c;
  }
}
__global__ __launch_bounds__(kWG) void k_det_place(ddm_batch B, ddm_det D) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *D.counter = 0;  // the scan has finished
  const int ci = chunk_index(B);
  if (ci < 0) return;  // uniform: a fixed-size launch with fewer chunks
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[ci];
  const int p = ch.y + (int)threadIdx.x;
  if (p >= ch.z) return;
  const int slot = ci * kWG + (int)threadIdx.x;
  const int2 ks = reinterpret_cast<const int2 *>(D.key_slot)[slot];
  const int q = D.cell_start[ks.x] + ks.y;
  D.order[q] = slot;
  D.order_p[q] = p;
}
// Final order: per listed block (one workgroup), every entry's rank among its node's
// entries by particle index (particle indices are unique within a scene, and a node holds
// one scene's particles), counted over the node's entries held in LDS (read from global
// memory when the block's segment exceeds kSortCap). Each particle's record is copied to
// its sorted position, which replaces the arrival rank in key_slot.
constexpr int kSortCap = 3072;
__global__ __launch_bounds__(kWG) void k_det_order(ddm_batch B, ddm_det D) {
  __shared__ int s_slot[kSortCap], s_p[kSortCap];
  __shared__ int s_start[64], s_cnt[64];
  const int t = threadIdx.x;
  const int n_listed = *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  for (int e = blockIdx.x; e < n_listed; e += gridDim.x) {  // uniform over the workgroup
    const size_t g0 = (size_t)list[e].x << 6;
    if (t < 64) {
      s_start[t] = D.cell_start[g0 + t];
      s_cnt[t] = D.cell_count[g0 + t];
    }
    __syncthreads();
    const int base = s_start[0], tot = s_start[63] + s_cnt[63] - s_start[0];
    const bool fits = tot <= kSortCap;
    if (fits)
      for (int i = t; i < tot; i += kWG) {
        s_slot[i] = D.order[base + i];
        s_p[i] = D.order_p[base + i];
      }
    __syncthreads();
    const int *P = fits ? s_p : D.order_p + base;
    const int *L = fits ? s_slot : D.order + base;
    for (int i = t; i < tot; i += kWG) {
      int lo = 0, hi = 63;  // node of entry i: the last node whose segment starts at or before i
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_start[mid] - base <= i) lo = mid;
        else hi = mid - 1;
      }
      const int c0 = s_start[lo] - base, n = s_cnt[lo], pi = P[i];
      int r = 0;
      for (int j = c0; j < c0 + n; ++j) r += P[j] < pi ? 1 : 0;
      const int q = base + c0 + r, slot = L[i];
      const float4 *src = reinterpret_cast<const float4 *>(D.records) + 4 * (size_t)slot;
      float4 *dst = reinterpret_cast<float4 *>(D.sorted) + 4 * (size_t)q;
      dst[0] = src[0];
      dst[1] = src[1];
      dst[2] = src[2];
      dst[3] = src[3];
      D.key_slot[2 * (size_t)slot + 1] = q;
    }
    __syncthreads();  // LDS is reused by the next block
  }
}

// one particle's (m, m v + affine (node - x)) contribution to the node at stencil offset
// (I, J, K): k_p2g's per-node operations
struct P2GDetContrib {
  float dx;
  template <int I, int J, int K>
  __device__ __forceinline__ void add(float4 a, float4 b, float4 c, float4 d, float *acc) const {
    const float wxy = bspline_w<I>(a.x) * bspline_w<J>(a.y);
    const float d0 = ((float)I - a.x) * dx, d1 = ((float)J - a.y) * dx;
    const float p0 = b.w * d0 + c.x * d1, p1 = c.z * d0 + c.w * d1, p2 = d.y * d0 + d.z * d1;
    const float w = wxy * bspline_w<K>(a.z);
    const float d2 = ((float)K - a.z) * dx;
    acc[0] += a.w * w;
    acc[1] += (b.x + (p0 + c.y * d2)) * w;
    acc[2] += (b.y + (p1 + d.x * d2)) * w;
    acc[3] += (b.z + (p2 + d.w * d2)) * w;
  }
};
// (m, mv) of every node of the listed blocks, summed in the fixed order (plain stores);
// one workgroup per listed block
template <bool SKIP_RESTORED>
__global__ __launch_bounds__(kWG) void k_det_gather_p2g(ddm_batch B, ddm_det D) {
  __shared__ DetGatherLDS L;
  const int n_listed = *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  for (int e = blockIdx.x; e < n_listed; e += gridDim.x) {  // uniform over the workgroup
    const int2 entry = list[e];
    const ddm_scene &S = B.scenes[entry.y];
    if (SKIP_RESTORED && S.tape_mmv && S.tape_status[1] == 0) continue;
    P2GDetContrib k{S.dx};
    float acc[4];
    det_gather_block<4>(S, D, entry.x, L, k, acc);
    if (threadIdx.x < 64)
      reinterpret_cast<float4 *>(B.grid_mmv)[((size_t)entry.x << 6) | threadIdx.x] =
          make_float4(acc[0], acc[1], acc[2], acc[3]);
  }
}

__global__ void k_zero_i32(int32_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_zero(float4 *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_zero_u32(uint32_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

inline int check(const ddm_batch *b, const void *cur, const void *next) {
  if (!b || !cur || !next || b->n_chunks < 0 || b->n_blocks < 0 || b->n_scenes <= 0) return DDM_EINVAL;
  if (b->max_bodies > DDM_MAX_BODIES) return DDM_ETOOMANYBODIES;
  if (!b->scenes || !b->chunks || !b->grid_mmv || !b->grid_v || !b->block_flags || !b->block_list ||
      !b->block_count)
    return DDM_EINVAL;
  if (b->options & DDM_DETERMINISTIC) {
    const ddm_det *d = b->det;
    if (!d || !d->cell_count || !d->cell_start || !d->order || !d->order_p || !d->key_slot || !d->records ||
        !d->body_partial || !d->range_count || !d->counter)
      return DDM_EINVAL;
  }
  return 0;
}

// deterministic P2G: bin the particles (prep, block list in block order, segments,
// placement, per-node sort), then gather every listed node in the fixed order
template <int MODE>
void launch_det_p2g(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, hipStream_t stream) {
  const ddm_det &d = *b->det;
  const int ranges = b->n_blocks > 0 ? (b->n_blocks + kRange - 1) / kRange : 1;
  k_det_prep<MODE><<<b->n_chunks, kWG, 0, stream>>>(*b, d, cur, next);
  k_det_count<<<ranges, kWG, 0, stream>>>(*b, d);
  k_det_relist<<<ranges, kWG, 0, stream>>>(*b, d);
  k_det_scan<<<grid_wgs(b), kWG, 0, stream>>>(*b, d);
  k_det_place<<<b->n_chunks, kWG, 0, stream>>>(*b, d);
  k_det_order<<<grid_wgs(b) * 4, kWG, 0, stream>>>(*b, d);
  k_det_gather_p2g<MODE == kRecomputeUntaped><<<grid_wgs(b) * 4, kWG, 0, stream>>>(*b, d);
}

}  // namespace
}  // namespace ddm_fused

using namespace ddm_fused;

#define DDM_API extern "C" __attribute__((visibility("default")))

static inline void mark(hipEvent_t *ev, int i, hipStream_t stream) {
  if (ev) (void)hipEventRecord(ev[i], stream);
}

// prev (optional): the G2P of the substep before (prev -> cur) has not run yet; it runs
// first, inside this substep's P2G launch where the launch scans the block flags, else as
// its own launch. finish = 0: this substep's G2P is left to the next call (given prev).
static int launch_forward(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, hipStream_t stream,
                          hipEvent_t *ev, const ddm_state *prev = nullptr, int finish = 1) {
  int e = check(b, cur, next);
  if (e) return e;
  if (b->n_chunks == 0) return 0;
  const bool det = b->options & DDM_DETERMINISTIC, scan = scan_blocks(b);
  if (prev && !scan) k_g2p<<<b->n_chunks, kWG, 0, stream>>>(*b, prev, cur);
  mark(ev, 0, stream);
  if (det) launch_det_p2g<kForward>(b, cur, next, stream);
  else if (prev && scan) k_p2g<kForward, true><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, scan, prev);
  else k_p2g<kForward><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, scan);
  mark(ev, 1, stream);
  if (scan) k_grid<true, false, false, true><<<scan_wgs(b), kWG, 0, stream>>>(*b, cur, next, nullptr);
  else k_grid<true, false, false><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, det ? b->det->cell_count : nullptr);
  mark(ev, 2, stream);
  if (finish) k_g2p<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next);
  mark(ev, 3, stream);
  return (int)hipGetLastError();
}

// adjoint kernels after the grid update (both adjoint launch sequences)
static void launch_adjoint_tail(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream,
                                hipEvent_t *ev, int ev0, const ddm_scene *restore_prev = nullptr) {
  if (b->options & DDM_DETERMINISTIC) {
    launch_g2p_adj_det(*b, *b->det, cur, next, cur_grad, next_grad, stream);
    mark(ev, ev0, stream);
    launch_grid_adj_det(*b, *b->det, cur, next, cur_grad, next_grad, stream);
  } else {
    launch_g2p_adj(*b, cur, next, cur_grad, next_grad, stream);
    mark(ev, ev0, stream);
    launch_grid_adj(*b, cur, next, cur_grad, next_grad, stream);
  }
  mark(ev, ev0 + 1, stream);
  launch_p2g_adj(*b, cur, cur_grad, next_grad, stream, restore_prev);
  mark(ev, ev0 + 2, stream);
}

static int launch_backward(const ddm_batch *b_in, const ddm_state *cur, const ddm_state *next,
                           const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream, hipEvent_t *ev) {
  int e = check(b_in, cur, next);
  if (e) return e;
  ddm_batch bb = *b_in;  // DDM_TAPE_COMPLETE means nothing here (the grid adjoint must read the list)
  bb.options &= ~DDM_TAPE_COMPLETE;
  const ddm_batch *b = &bb;
  if (!cur_grad || !next_grad || !b->grid_dv || !b->grid_dmmv) return DDM_EINVAL;
  if (b->n_chunks == 0) return 0;
  mark(ev, 0, stream);
  if (b->options & DDM_DETERMINISTIC) launch_det_p2g<kRecompute>(b, cur, next, stream);
  else k_p2g<kRecompute><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, 0);
  mark(ev, 1, stream);
  k_grid<false, false, false><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, nullptr);
  mark(ev, 2, stream);
  launch_adjoint_tail(b, cur, next, cur_grad, next_grad, stream, ev, 3);
  return (int)hipGetLastError();
}

static int launch_forward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                hipStream_t stream, hipEvent_t *ev, const ddm_state *prev = nullptr,
                                int finish = 1) {
  int e = check(b, cur, next);
  if (e) return e;
  if (b->n_chunks == 0) return 0;
  const bool det = b->options & DDM_DETERMINISTIC, scan = scan_blocks(b);
  if (prev && !scan) k_g2p<<<b->n_chunks, kWG, 0, stream>>>(*b, prev, cur);
  mark(ev, 0, stream);
  if (det) launch_det_p2g<kForwardTaped>(b, cur, next, stream);
  else if (prev && scan) k_p2g<kForwardTaped, true><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, scan, prev);
  else k_p2g<kForwardTaped><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, scan);
  mark(ev, 1, stream);
  if (scan) k_grid<true, true, false, true><<<scan_wgs(b), kWG, 0, stream>>>(*b, cur, next, nullptr);
  else k_grid<true, true, false><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, det ? b->det->cell_count : nullptr);
  mark(ev, 2, stream);
  if (finish) k_g2p<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next);
  mark(ev, 3, stream);
  return (int)hipGetLastError();
}

static int launch_backward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                 const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream,
                                 hipEvent_t *ev, int restored = 0, const ddm_scene *restore_prev = nullptr) {
  int e = check(b, cur, next);
  if (e) return e;
  if (!cur_grad || !next_grad || !b->grid_dv || !b->grid_dmmv) return DDM_EINVAL;
  if ((restored || restore_prev) && (b->options & DDM_DETERMINISTIC)) return DDM_EINVAL;
  if (b->n_chunks == 0) return 0;
  mark(ev, 0, stream);
  if (!restored) k_restore<<<dim3(kRestoreWGs, b->n_scenes), kWG, 0, stream>>>(*b);
  // With a complete tape every scene restores its whole substep and both recompute kernels
  // would exit at once; the deterministic recompute P2G also rebuilds the per-cell particle
  // order that the sorted adjoint gathers read, so it always runs.
  if (!(b->options & DDM_TAPE_COMPLETE) || (b->options & DDM_DETERMINISTIC)) {
    if (b->options & DDM_DETERMINISTIC) launch_det_p2g<kRecomputeUntaped>(b, cur, next, stream);
    else k_p2g<kRecomputeUntaped><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, 0);
    k_grid<false, false, true><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, nullptr);
  }
  mark(ev, 1, stream);
  launch_adjoint_tail(b, cur, next, cur_grad, next_grad, stream, ev, 2, restore_prev);
  return (int)hipGetLastError();
}

DDM_API int ddm_substep_forward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                      hipStream_t stream) {
  return launch_forward_taped(b, cur, next, stream, nullptr);
}

DDM_API int ddm_substep_backward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                       const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream) {
  return launch_backward_taped(b, cur, next, cur_grad, next_grad, stream, nullptr);
}

DDM_API int ddm_substep_forward_ahead(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                      const ddm_state *prev, int finish, int taped, hipStream_t stream) {
  if (!finish && !next) return DDM_EINVAL;
  return taped ? launch_forward_taped(b, cur, next, stream, nullptr, prev, finish)
               : launch_forward(b, cur, next, stream, nullptr, prev, finish);
}

DDM_API int ddm_substep_backward_taped_ahead(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                             const ddm_state *cur_grad, const ddm_state *next_grad, int restored,
                                             const ddm_scene *prev_scenes, hipStream_t stream) {
  return launch_backward_taped(b, cur, next, cur_grad, next_grad, stream, nullptr, restored, prev_scenes);
}

DDM_API int ddm_substep_forward_taped_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                            hipStream_t stream, hipEvent_t *events) {
  if (!events) return DDM_EINVAL;
  return launch_forward_taped(b, cur, next, stream, events);
}

DDM_API int ddm_substep_backward_taped_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                             const ddm_state *cur_grad, const ddm_state *next_grad,
                                             hipStream_t stream, hipEvent_t *events) {
  if (!events) return DDM_EINVAL;
  return launch_backward_taped(b, cur, next, cur_grad, next_grad, stream, events);
}

DDM_API int ddm_substep_forward(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, hipStream_t stream) {
  return launch_forward(b, cur, next, stream, nullptr);
}

DDM_API int ddm_substep_backward(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                 const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream) {
  return launch_backward(b, cur, next, cur_grad, next_grad, stream, nullptr);
}

DDM_API int ddm_substep_forward_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                      hipStream_t stream, hipEvent_t *events) {
  if (!events) return DDM_EINVAL;
  return launch_forward(b, cur, next, stream, events);
}

DDM_API int ddm_substep_backward_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                       const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream,
                                       hipEvent_t *events) {
  if (!events) return DDM_EINVAL;
  return launch_backward(b, cur, next, cur_grad, next_grad, stream, events);
}

// ---------------------------------------------------------------- allocation-free forward
// Grid workspace of ddm_substep_forward_scratch, in the library's own device memory: one
// copy per GPU, allocated when the code object loads and held until the process ends.
// Layout: (m, mv) [kScratchBlocks*64] float4 | v [kScratchBlocks*64] float4 |
// flags [kScratchBlocks] u32 | list [kScratchBlocks][2] i32 | count (256 B).
constexpr size_t kScratchBlocks = 8192;  // e.g. two 64^3 grids, or one 128x128x32 grid
constexpr size_t kScratchNodes = kScratchBlocks * 64;
constexpr size_t kScratchOffV = kScratchNodes * 16;
constexpr size_t kScratchOffFlags = 2 * kScratchNodes * 16;
constexpr size_t kScratchOffList = kScratchOffFlags + kScratchBlocks * 4;
constexpr size_t kScratchOffCount = kScratchOffList + kScratchBlocks * 8;
constexpr size_t kScratchBytes = kScratchOffCount + 256;  // 16.1 MiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, looked up once per device (each GPU
// holds its own copy at its own address; the array's name in host code is not that address).
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

static int scratch_base(unsigned char **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return DDM_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return DDM_EINVAL;
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

DDM_API int ddm_substep_forward_scratch(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                        hipStream_t stream) {
  if (!b || b->n_blocks < 0) return DDM_EINVAL;
  if ((size_t)b->n_blocks > kScratchBlocks) return DDM_ESCRATCH;
  if (b->options & DDM_DETERMINISTIC) return DDM_EINVAL;  // its workspace is not part of the scratch
  unsigned char *base = nullptr;
  int e = scratch_base(&base);
  if (e) return e;
  ddm_batch sb = *b;
  sb.grid_mmv = reinterpret_cast<float *>(base);
  sb.grid_v = reinterpret_cast<float *>(base + kScratchOffV);
  sb.block_flags = reinterpret_cast<uint32_t *>(base + kScratchOffFlags);
  sb.block_list = reinterpret_cast<int32_t *>(base + kScratchOffList);
  sb.block_count = reinterpret_cast<int32_t *>(base + kScratchOffCount);
  sb.grid_dv = sb.grid_dmmv = nullptr;
  sb.det = nullptr;
  // The scratch keeps whatever the last call on this GPU left. A completed forward leaves
  // (m, mv), flags and count cleared, but a failed or interrupted one may not: clear what
  // must start at zero, for the blocks this launch addresses, with fill kernels.
  const size_t nodes = (size_t)sb.n_blocks * 64;
  if (nodes) {
    k_zero<<<1024, kWG, 0, stream>>>(reinterpret_cast<float4 *>(sb.grid_mmv), nodes);
    k_zero_u32<<<64, kWG, 0, stream>>>(sb.block_flags, (size_t)sb.n_blocks);
  }
  k_zero_u32<<<1, kWG, 0, stream>>>(reinterpret_cast<uint32_t *>(sb.block_count), 1);
  e = (int)hipGetLastError();
  if (e) return e;
  return launch_forward(&sb, cur, next, stream, nullptr);
}

DDM_API int ddm_batch_reset(const ddm_batch *b, hipStream_t stream) {
  if (!b || b->n_blocks < 0) return DDM_EINVAL;
  size_t nodes = (size_t)b->n_blocks * 64;
  if (nodes == 0) return 0;
  int wg = 1024;
  k_zero<<<wg, kWG, 0, stream>>>(reinterpret_cast<float4 *>(b->grid_mmv), nodes);
  k_zero<<<wg, kWG, 0, stream>>>(reinterpret_cast<float4 *>(b->grid_v), nodes);
  if (b->grid_dv) k_zero<<<wg, kWG, 0, stream>>>(reinterpret_cast<float4 *>(b->grid_dv), nodes);
  if (b->grid_dmmv) k_zero<<<wg, kWG, 0, stream>>>(reinterpret_cast<float4 *>(b->grid_dmmv), nodes);
  k_zero_u32<<<wg, kWG, 0, stream>>>(b->block_flags, (size_t)b->n_blocks);
  if (b->block_count) k_zero_u32<<<1, kWG, 0, stream>>>(reinterpret_cast<uint32_t *>(b->block_count), 1);
  if (b->det && b->det->cell_count) k_zero_i32<<<wg, kWG, 0, stream>>>(b->det->cell_count, nodes);
  if (b->det && b->det->counter) k_zero_i32<<<1, kWG, 0, stream>>>(b->det->counter, 1);
  return (int)hipGetLastError();
}
