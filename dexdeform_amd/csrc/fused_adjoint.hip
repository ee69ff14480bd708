// Adjoint kernels of the fused multi-scene substep. Compiled with FMA contraction
// (see fused_common.h); the recompute kernels of the adjoint pass stay in
// fused_kernels.hip with the forward kernels.
#include "fused_common.h"

namespace ddm_fused {
namespace {

// adjoint of the grid update; clears (m, mv), the velocity adjoint and the flags.
// DET (deterministic mode): each block writes its body-pose gradient sums to its list
// entry's partials (k_det_body_reduce adds them up in block order) instead of atomics,
// and clears its bin counts.
// SCAN (complete snapshots, small launches): every block of the launch is visited and those
// the restore marked (with scene + 1) are processed, instead of reading the block list.
template <bool DET, bool SCAN = false>
__global__ __launch_bounds__(kWG) void k_grid_adj(ddm_batch B, ddm_det D, const ddm_state *__restrict__ cur_tab,
                                                  const ddm_state *__restrict__ next_tab,
                                                  const ddm_state *__restrict__ cur_grad_tab,
                                                  const ddm_state *__restrict__ next_grad_tab) {
  __shared__ WaveBodies s_bodies[kWG / 64];
  const int wave = wave_id(), lane = threadIdx.x & 63;
  float4 *mmv = reinterpret_cast<float4 *>(B.grid_mmv);
  float4 *dv = reinterpret_cast<float4 *>(B.grid_dv);
  float4 *dmmv = reinterpret_cast<float4 *>(B.grid_dmmv);
  static_assert(!(DET && SCAN), "the deterministic mode reduces in list order");
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
    BlockCoord c = block_node(S, blk, lane);
    size_t gi = ((size_t)blk << 6) | (size_t)lane;
    float4 a = mmv[gi];
    float4 d = dv[gi];
    const float m = a.x;
    const f3 mv{a.y, a.z, a.w};
    const bool act = node_active(m);
    NodeParams P = scene_node_params(S, cur_tab[si], next_tab[si]);
    stage_bodies(P, s_bodies[wave], lane);
    const unsigned long long near = near_bodies(P, c.ix, c.iy, c.iz, lane);  // others: no contact here
    const f3 gx = node_position(c.ix, c.iy, c.iz, P.dx);
    const int mode = S.adjoint_mode;
    // forward replay keeping the velocities entering the node's first kTapeMax contacts
    // (integrator.cu:678-732) and the set of bodies in contact (bit b of hit: contact()
    // returned true, i.e. in_contact held for that body at this node)
    f3 tape_v[kTapeMax];
    unsigned long long hit = 0ull;
    int nt = 0;
    f3 v0{0.f, 0.f, 0.f}, g{0.f, 0.f, 0.f};
    if (act) {
      f3 v = mv * (1.f / m);
      v += P.dt * P.gravity;
      v0 = v;
      for (int b = 0; b < P.nb; ++b) {
        if (!((near >> b) & 1ull)) continue;
        f3 vin = v;
        if (contact(v, gx, load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b), P.dt)) {
          hit |= 1ull << b;
          if (nt < kTapeMax) tape_v[nt] = vin;
          ++nt;  // counts every contact; the tape keeps the first kTapeMax
        }
      }
      g = boundary_adj(v, f3{d.x, d.y, d.z}, c.ix, c.iy, c.iz, P, mode);
    }
    // bodies in reverse order, uniformly across the wave so pose gradients can be reduced;
    // a body no node of the block touched is skipped by the whole wave
    for (int b = P.nb - 1; b >= 0; --b) {
      const bool mine = (hit >> b) & 1ull;
      if (!__any(mine)) {  // uniform
        if (DET && lane < 4)  // its partial for this block is zero
          reinterpret_cast<float4 *>(D.body_partial + ((size_t)e * B.max_bodies + b) * 16)[lane] =
              make_float4(0.f, 0.f, 0.f, 0.f);
        continue;
      }
      PoseGrad pg = posegrad_zero();
      if (mine) {
        const Body Bd = load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b);
        f3 vin;
        --nt;  // contacts are popped in reverse order of the replay
        if (nt < kTapeMax) {
          vin = tape_v[nt];
        } else {  // beyond the tape: replay the contacts before b
          vin = v0;
          for (int b2 = 0; b2 < b; ++b2)
            if ((hit >> b2) & 1ull)
              contact(vin, gx, load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b2), P.dt);
        }
        contact_adj(vin, gx, Bd, P.dt, mode, g, pg);
      }
      float vals[14] = {pg.pos.x,  pg.pos.y,  pg.pos.z,  pg.rot.w,  pg.rot.x,  pg.rot.y,  pg.rot.z,
                        pg.npos.x, pg.npos.y, pg.npos.z, pg.nrot.w, pg.nrot.x, pg.nrot.y, pg.nrot.z};
#pragma unroll
      for (int q = 0; q < 14; ++q) vals[q] = wave_sum(vals[q]);
      if (DET) {  // this block's partial, reduced in block order by k_det_body_reduce
        float4 *part = reinterpret_cast<float4 *>(D.body_partial + ((size_t)e * B.max_bodies + b) * 16);
        if (lane == 0) {
          part[0] = make_float4(vals[0], vals[1], vals[2], vals[3]);
          part[1] = make_float4(vals[4], vals[5], vals[6], vals[7]);
          part[2] = make_float4(vals[8], vals[9], vals[10], vals[11]);
          part[3] = make_float4(vals[12], vals[13], 0.f, 0.f);
        }
      } else if (lane == 0) {
        const ddm_state &cg = cur_grad_tab[si];
        const ddm_state &ng = next_grad_tab[si];
        float *cp = cg.body_pos + 3 * b, *cr = cg.body_rot + 4 * b;
        float *np = ng.body_pos + 3 * b, *nr = ng.body_rot + 4 * b;
        atomicAdd(cp + 0, vals[0]);
        atomicAdd(cp + 1, vals[1]);
        atomicAdd(cp + 2, vals[2]);
        atomicAdd(cr + 0, vals[3]);
        atomicAdd(cr + 1, vals[4]);
        atomicAdd(cr + 2, vals[5]);
        atomicAdd(cr + 3, vals[6]);
        atomicAdd(np + 0, vals[7]);
        atomicAdd(np + 1, vals[8]);
        atomicAdd(np + 2, vals[9]);
        atomicAdd(nr + 0, vals[10]);
        atomicAdd(nr + 1, vals[11]);
        atomicAdd(nr + 2, vals[12]);
        atomicAdd(nr + 3, vals[13]);
      }
    }
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (act) {
      f3 gmv = (float)(1.0 / (double)m) * g;  // integrator.cu:1054-1055
      out = make_float4((-1.f / m / m) * dot3(mv, g), gmv.x, gmv.y, gmv.z);
    }
    dmmv[gi] = out;
    mmv[gi] = make_float4(0.f, 0.f, 0.f, 0.f);
    dv[gi] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (lane == 0) B.block_flags[blk] = 0u;
    if (DET) D.cell_count[gi] = 0;  // the sorted gathers of this substep are done
  }
}

// Deterministic mode: per scene (one wave), each body's pose gradients summed over the
// scene's listed blocks in block order (the rebuilt list is sorted by block; a scene's
// blocks are a contiguous range), then added once to the current and next substep's
// body-pose gradients.
__global__ __launch_bounds__(64) void k_det_body_reduce(ddm_batch B, ddm_det D,
                                                        const ddm_state *__restrict__ cur_grad_tab,
                                                        const ddm_state *__restrict__ next_grad_tab) {
  const int si = (int)blockIdx.x;
  if (si >= live_scenes(B)) return;
  const ddm_scene &S = B.scenes[si];
  const int lane = threadIdx.x;
  const int n_listed = *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  auto lower = [&](int blk) {
    int lo = 0, hi = n_listed;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (list[mid].x < blk) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int e0 = lower(S.blk_begin), e1 = lower(S.blk_begin + S.n_blocks);
  if (e0 == e1) return;
  const ddm_state cg = cur_grad_tab[si], ng = next_grad_tab[si];
  for (int b = 0; b < S.n_bodies; ++b) {
    float v[14];
#pragma unroll
    for (int q = 0; q < 14; ++q) v[q] = 0.f;
    for (int e = e0 + lane; e < e1; e += 64) {
      const float4 *part = reinterpret_cast<const float4 *>(D.body_partial + ((size_t)e * B.max_bodies + b) * 16);
      const float4 p0 = part[0], p1 = part[1], p2 = part[2], p3 = part[3];
      v[0] += p0.x, v[1] += p0.y, v[2] += p0.z, v[3] += p0.w;
      v[4] += p1.x, v[5] += p1.y, v[6] += p1.z, v[7] += p1.w;
      v[8] += p2.x, v[9] += p2.y, v[10] += p2.z, v[11] += p2.w;
      v[12] += p3.x, v[13] += p3.y;
    }
#pragma unroll
    for (int q = 0; q < 14; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
      float *cp = cg.body_pos + 3 * b, *cr = cg.body_rot + 4 * b;
      float *np = ng.body_pos + 3 * b, *nr = ng.body_rot + 4 * b;
      cp[0] += v[0], cp[1] += v[1], cp[2] += v[2];
      cr[0] += v[3], cr[1] += v[4], cr[2] += v[5], cr[3] += v[6];
      np[0] += v[7], np[1] += v[8], np[2] += v[9];
      nr[0] += v[10], nr[1] += v[11], nr[2] += v[12], nr[3] += v[13];
    }
  }
}

// adjoint of G2P: adds to the x adjoint and scatters the node-velocity adjoint. LAT: the
// variant for launches of at most kLatencyChunks chunks (about one wave per SIMD), where
// the kernel runs at one chunk's latency: the node velocities are gathered in one round
// trip (more registers, which such a launch has to spare).
template <bool LAT>
__global__ __launch_bounds__(kWG, LAT ? 1 : 5) void k_g2p_adj(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                                 const ddm_state *__restrict__ next_tab,
                                                 const ddm_state *__restrict__ cur_grad_tab,
                                                 const ddm_state *__restrict__ next_grad_tab) {
  __shared__ ScatterLDS<3, kTileG2P> sh;
  __shared__ StencilRecs rc;
  __shared__ Groups gr;
  __shared__ int s_box[6];
  float *const tx = sh.tile[0], *const ty = sh.tile[1], *const tz = sh.tile[2];
  const int ci = chunk_index(B);
  if (ci < 0) return;  // uniform: a fixed-size launch with fewer chunks
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[ci];
  const ddm_scene S = B.scenes[ch.x];
  const ddm_state cur = cur_tab[ch.x], next = next_tab[ch.x];
  const ddm_state cur_grad = cur_grad_tab[ch.x], next_grad = next_grad_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  const bool valid = p < ch.z;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  Stencil st;
  f3 g_nv{0.f, 0.f, 0.f};
  m33 gC;
  int key = 0;
  if (valid) {
    BlockedV gv{reinterpret_cast<const float4 *>(B.grid_v), &S};
    gC = ld9(next_grad.C + 9 * (size_t)p);
    f3 gxp = g2p_particle_adj_gather<LAT>(ld3(cur.x + 3 * (size_t)p), ld3(next.v + 3 * (size_t)p),
                                     ld3(next_grad.x + 3 * (size_t)p), ld3(next_grad.v + 3 * (size_t)p), gC, gv,
                                     S.inv_dx, S.dx, S.dt, S.ground_height, dim, g_nv, st);
    if (B.options & DDM_GRAD_OVERWRITE) st3(cur_grad.x + 3 * (size_t)p, gxp);
    else add3(cur_grad.x + 3 * (size_t)p, gxp);
    key = pack_base(st.base);
  } else {
    st.base[0] = st.base[1] = st.base[2] = 0;
  }
  const Tile T = chunk_tile(valid, st.base, s_box);
  int pos;
  const int G = build_groups(valid, key, min(kWG, ch.z - ch.y), sh.grp, gr, pos);
  if (valid) stage_record(rc, pos, st, 0.f, g_nv, gC);
  const float inv_dx = S.inv_dx;
  // contributions of record s to the nodes (i, j, 0..2): g2p_adj_contrib's operations, with
  // the k-independent ones done once (C adjoint = {b.w, c.x, c.y | c.z, c.w, d.x | d.y, d.z,
  // d.w}, node-velocity adjoint seed = b.xyz)
  auto contrib = [inv_dx](const StencilRecs &r, int s, int i, int j, float (*acc)[3]) {
    const float4 a = r.f0[s], b = r.f1[s], c = r.f2[s], d = r.f3[s];
    const float wxy = bspline_wi(a.x, i) * bspline_wi(a.y, j);
    const float wz[3] = {bspline_w<0>(a.z), bspline_w<1>(a.z), bspline_w<2>(a.z)};
    const float d0 = (float)i - a.x, d1 = (float)j - a.y;
    const float p0 = b.w * d0 + c.x * d1, p1 = c.z * d0 + c.w * d1, p2 = d.y * d0 + d.z * d1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float w = wxy * wz[k];
      const float d2 = (float)k - a.z;
      const float xx = w * inv_dx * 4.f;
      acc[k][0] += w * b.x + (p0 + c.y * d2) * xx;
      acc[k][1] += w * b.y + (p1 + d.x * d2) * xx;
      acc[k][2] += w * b.z + (p2 + d.w * d2) * xx;
    }
  };
  if (slabs_fit<kTileG2P>(T)) {
    const int sx = slab_width<kTileG2P>(T);
    for (int xs = T.lo[0]; xs < T.lo[0] + T.t[0]; xs += sx) {  // uniform over the workgroup
      const Tile L = slab(T, xs, sx);
      if (xs != T.lo[0]) __syncthreads();  // the previous pass has flushed the tile
      for (int i = threadIdx.x; i < L.n; i += kWG) tx[i] = ty[i] = tz[i] = 0.f;
      __syncthreads();  // also publishes the staged records
      group_column_tasks<3>(G, gr, rc, &L, contrib, [&](int X, int Y, int Z, const float *a) {
        const int li = L.local(X, Y, Z);
        atomicAdd(&tx[li], a[0]);
        atomicAdd(&ty[li], a[1]);
        atomicAdd(&tz[li], a[2]);
      });
      __syncthreads();
      for (int i = threadIdx.x; i < L.n; i += kWG) {
        float a = tx[i], b = ty[i], c = tz[i];
        if (a != 0.f || b != 0.f || c != 0.f) {
          int xx = L.lo[0] + i / (L.t[1] * L.t[2]), yy = L.lo[1] + (i / L.t[2]) % L.t[1], zz = L.lo[2] + i % L.t[2];
          float *g = B.grid_dv + 4 * (size_t)node_index(S, xx, yy, zz);
          atomicAdd(g + 0, a);
          atomicAdd(g + 1, b);
          atomicAdd(g + 2, c);
        }
      }
    }
  } else {
    __syncthreads();  // publishes the staged records
    group_column_tasks<3>(G, gr, rc, nullptr, contrib, [&](int X, int Y, int Z, const float *a) {
      float *g = B.grid_dv + 4 * (size_t)node_index(S, X, Y, Z);
      atomicAdd(g + 0, a[0]);
      atomicAdd(g + 1, a[1]);
      atomicAdd(g + 2, a[2]);
    });
  }
}

// Deterministic G2P adjoint, per particle: the x adjoint as k_g2p_adj, and the record
// (fx | node-velocity adjoint seed, C adjoint) for the sorted gather (the bins are the P2G
// pass's: same positions, same stencil bases).
__global__ __launch_bounds__(kWG) void k_g2p_adj_det(ddm_batch B, ddm_det D, const ddm_state *__restrict__ cur_tab,
                                                     const ddm_state *__restrict__ next_tab,
                                                     const ddm_state *__restrict__ cur_grad_tab,
                                                     const ddm_state *__restrict__ next_grad_tab) {
  const int ci = chunk_index(B);
  if (ci < 0) return;  // uniform: a fixed-size launch with fewer chunks
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[ci];
  const int p = ch.y + (int)threadIdx.x;
  if (p >= ch.z) return;
  const ddm_scene &S = B.scenes[ch.x];
  const ddm_state cur = cur_tab[ch.x], next = next_tab[ch.x];
  const ddm_state cur_grad = cur_grad_tab[ch.x], next_grad = next_grad_tab[ch.x];
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  BlockedV gv{reinterpret_cast<const float4 *>(B.grid_v), &S};
  NullScatter none;
  Stencil st;
  f3 g_nv{0.f, 0.f, 0.f};
  const m33 gC = ld9(next_grad.C + 9 * (size_t)p);
  f3 gxp = g2p_particle_adj(ld3(cur.x + 3 * (size_t)p), ld3(next.v + 3 * (size_t)p), ld3(next_grad.x + 3 * (size_t)p),
                            ld3(next_grad.v + 3 * (size_t)p), gC, gv, none, S.inv_dx, S.dx, S.dt, S.ground_height,
                            dim, &g_nv, &st);
  if (B.options & DDM_GRAD_OVERWRITE) st3(cur_grad.x + 3 * (size_t)p, gxp);
  else add3(cur_grad.x + 3 * (size_t)p, gxp);
  // at the particle's sorted position (the P2G pass of this adjoint sorted the same bins)
  const size_t slot = (size_t)ci * kWG + threadIdx.x;
  float4 *r = reinterpret_cast<float4 *>(D.sorted) + 4 * (size_t)D.key_slot[2 * slot + 1];
  r[0] = make_float4(st.fx[0], st.fx[1], st.fx[2], 0.f);
  r[1] = make_float4(g_nv.x, g_nv.y, g_nv.z, gC.a[0]);
  r[2] = make_float4(gC.a[1], gC.a[2], gC.a[3], gC.a[4]);
  r[3] = make_float4(gC.a[5], gC.a[6], gC.a[7], gC.a[8]);
}
// one particle's node-velocity adjoint at stencil offset (I, J, K): k_g2p_adj's per-node operations
struct G2PAdjDetContrib {
  float inv_dx;
  template <int I, int J, int K>
  __device__ __forceinline__ void add(float4 a, float4 b, float4 c, float4 d, float *acc) const {
    const float wxy = bspline_w<I>(a.x) * bspline_w<J>(a.y);
    const float d0 = (float)I - a.x, d1 = (float)J - a.y;
    const float p0 = b.w * d0 + c.x * d1, p1 = c.z * d0 + c.w * d1, p2 = d.y * d0 + d.z * d1;
    const float w = wxy * bspline_w<K>(a.z);
    const float d2 = (float)K - a.z;
    const float xx = w * inv_dx * 4.f;
    acc[0] += w * b.x + (p0 + c.y * d2) * xx;
    acc[1] += w * b.y + (p1 + d.x * d2) * xx;
    acc[2] += w * b.z + (p2 + d.w * d2) * xx;
  }
};
// the node-velocity adjoint of every node of the listed blocks; one workgroup per block
__global__ __launch_bounds__(kWG) void k_det_gather_g2p_adj(ddm_batch B, ddm_det D) {
  __shared__ DetGatherLDS L;
  const int n_listed = *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  for (int e = blockIdx.x; e < n_listed; e += gridDim.x) {  // uniform over the workgroup
    const int2 entry = list[e];
    const ddm_scene &S = B.scenes[entry.y];
    G2PAdjDetContrib k{S.inv_dx};
    float acc[3];
    det_gather_block<3>(S, D, entry.x, L, k, acc);
    if (threadIdx.x < 64)
      reinterpret_cast<float4 *>(B.grid_dv)[((size_t)entry.x << 6) | threadIdx.x] =
          make_float4(acc[0], acc[1], acc[2], 0.f);
  }
}

// adjoint of P2G + SVD adjoint + chain through F_tmp = (I + dt C) F
// at most 168 VGPRs (3 waves per SIMD instead of 2): 11 % faster. LAT: as k_g2p_adj's, the
// node gradients are gathered first, in one round trip that overlaps the particle's stress.
// restore_prev (taped adjoint, optional): kRestoreWGs workgroups per scene after the chunks'
// put back the snapshots of the substep before, from its scene table (the arena blocks,
// list and count they write are not read here; the grid adjoint before has finished).
template <bool LAT>
__global__ __launch_bounds__(kWG, LAT ? 1 : 3) void k_p2g_adj(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                                 const ddm_state *__restrict__ cur_grad_tab,
                                                 const ddm_state *__restrict__ next_grad_tab,
                                                 const ddm_scene *restore_prev) {
  if (restore_prev && (int)blockIdx.x >= B.n_chunks) {  // uniform
    ddm_batch Bp = B;
    Bp.scenes = restore_prev;
    const int r = (int)blockIdx.x - B.n_chunks;
    restore_scene_part(Bp, r / kRestoreWGs, r % kRestoreWGs, kRestoreWGs);
    return;
  }
  // the grid kernels have finished (a restore running beside sets the count itself)
  if (!restore_prev && blockIdx.x == 0 && threadIdx.x == 0) *B.block_count = 0;
  const int ci = chunk_index(B);
  if (ci < 0) return;  // uniform: a fixed-size launch with fewer chunks
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[ci];
  const ddm_scene S = B.scenes[ch.x];
  const ddm_state cur = cur_tab[ch.x], cur_grad = cur_grad_tab[ch.x], next_grad = next_grad_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  if (p >= ch.z) return;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  const size_t p3 = 3 * (size_t)p, p9 = 9 * (size_t)p;
  BlockedGatherDMMV gat{reinterpret_cast<const float4 *>(B.grid_dmmv), &S};
  const float *Fp = cur.F + p9;
  const m33 C = ld9(cur.C + p9);
  // F_tmp as the forward P2G computed it (advect_F is evaluated without contraction here
  // too, so it is the same bits); its SVD from the temporaries (U and V as quaternions)
  const m33 F0 = ld9(Fp);
  const m33 Ft = advect_F(F0, C, S.dt), U = quat_mat(ldq(S.tmp_qu + 4 * (size_t)p)),
            V = quat_mat(ldq(S.tmp_qv + 4 * (size_t)p));
  f3 sig = ld3(S.tmp_sig + p3);
  P2GAdj a;
  p2g_particle_adj<true, LAT>(ld3(cur.x + p3), ld3(cur.v + p3), C, Ft, U, sig, V, S.mass[p], S.vol[p],
                   ld3(S.mu_lam_yield + p3), ld9(next_grad.F + p9), gat, S.dx, S.inv_dx, S.dt, dim, a);
  m33 gFt, gF, gC;
  // F is read again (from the cache) rather than held across the P2G adjoint, except in
  // the latency variant, which has the registers and would wait for the read
  if (!LAT) asm volatile("" : "+v"(Fp));
  svd_adj<true>(LAT ? F0 : ld9(Fp), C, U, V, sig, a.Ft, a.U, a.V, a.sig, S.dt, gFt, gF, gC);
  const bool overwrite = B.options & DDM_GRAD_OVERWRITE;
  add3(cur_grad.x + p3, a.x);  // onto the G2P adjoint's contribution of this substep
  if (overwrite) st3(cur_grad.v + p3, a.v);
  else add3(cur_grad.v + p3, a.v);
  m33 cg = overwrite ? m33{} : ld9(cur_grad.C + p9);
  cg = madd(madd(cg, a.C), gC);  // p2g_grad adds first, then compute_svd_grad (simulator.py:584-585)
  st9(cur_grad.C + p9, cg);
  if (overwrite) st9(cur_grad.F + p9, gF);
  else add9(cur_grad.F + p9, gF);
}

}  // namespace

void launch_g2p_adj(const ddm_batch &b, const ddm_state *cur, const ddm_state *next, const ddm_state *cur_grad,
                    const ddm_state *next_grad, hipStream_t stream) {
  if (b.n_chunks <= kLatencyChunks) k_g2p_adj<true><<<b.n_chunks, kWG, 0, stream>>>(b, cur, next, cur_grad, next_grad);
  else k_g2p_adj<false><<<b.n_chunks, kWG, 0, stream>>>(b, cur, next, cur_grad, next_grad);
}
void launch_grid_adj(const ddm_batch &b, const ddm_state *cur, const ddm_state *next, const ddm_state *cur_grad,
                     const ddm_state *next_grad, hipStream_t stream) {
  if ((b.options & DDM_TAPE_COMPLETE) && scan_blocks(&b))  // every listed block was restored, and marked
    k_grid_adj<false, true><<<scan_wgs(&b), kWG, 0, stream>>>(b, ddm_det{}, cur, next, cur_grad, next_grad);
  else k_grid_adj<false><<<grid_wgs(&b), kWG, 0, stream>>>(b, ddm_det{}, cur, next, cur_grad, next_grad);
}
void launch_g2p_adj_det(const ddm_batch &b, const ddm_det &d, const ddm_state *cur, const ddm_state *next,
                        const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream) {
  k_g2p_adj_det<<<b.n_chunks, kWG, 0, stream>>>(b, d, cur, next, cur_grad, next_grad);
  k_det_gather_g2p_adj<<<grid_wgs(&b) * 4, kWG, 0, stream>>>(b, d);
}
void launch_grid_adj_det(const ddm_batch &b, const ddm_det &d, const ddm_state *cur, const ddm_state *next,
                         const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream) {
  k_grid_adj<true><<<grid_wgs(&b), kWG, 0, stream>>>(b, d, cur, next, cur_grad, next_grad);
  k_det_body_reduce<<<b.n_scenes, 64, 0, stream>>>(b, d, cur_grad, next_grad);
}
void launch_p2g_adj(const ddm_batch &b, const ddm_state *cur, const ddm_state *cur_grad, const ddm_state *next_grad,
                    hipStream_t stream, const ddm_scene *restore_prev) {
  const int wgs = b.n_chunks + (restore_prev ? kRestoreWGs * b.n_scenes : 0);
  if (b.n_chunks <= kLatencyChunks)
    k_p2g_adj<true><<<wgs, kWG, 0, stream>>>(b, cur, cur_grad, next_grad, restore_prev);
  else k_p2g_adj<false><<<wgs, kWG, 0, stream>>>(b, cur, cur_grad, next_grad, restore_prev);
}

}  // namespace ddm_fused
