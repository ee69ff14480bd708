// Fused multi-scene substep for gfx950.
//
// Layout and invariants
//  * Scenes are independent: per-scene records (ddm_scene, ddm_state) in device memory
//    give each scene's particle arrays, material, primitives and poses; a chunk table
//    maps each workgroup to <= 256 particles of one scene (local indices).
//  * Each scene's grid is stored in 4x4x4 node blocks (64 nodes = one wavefront),
//    blocks of all scenes concatenated; every grid array is float4 per node.
//  * Block-sparse: P2G marks the blocks it touches in block_flags and appends each newly
//    marked block, with its scene, to block_list; the grid kernels visit exactly the
//    listed blocks. Between substeps grid_mmv, grid_dv, block_flags and block_count are
//    all zero: the last grid kernel of a substep clears what it consumed and the kernel
//    after it resets the count, so no dense memset is ever needed.
//  * P2G and the G2P adjoint stage each particle's stencil data in LDS, group the chunk
//    into runs of consecutive particles with the same stencil base (cells, for
//    cell-sorted particles) and scatter one (run, stencil node) task per thread: the
//    task sums its run in registers and adds one value per component to an LDS tile
//    covering the chunk's stencil box. LDS float atomics are the throughput limit of
//    this scatter, so a run of n particles costs one atomic instead of n.
//  * The tile is flushed to the grid with global atomics for its non-zero nodes. A box
//    larger than the tile is processed in x-slabs, one tile pass per slab; only a box
//    whose single x-plane exceeds the tile sends the tasks' sums to global atomics.
//  * The contact adjoint reduces body-pose gradients across the wavefront before one
//    atomic per body per wave.
#include <limits.h>
#include "mpm_device.h"
#include "../../include/dexdeform_mpm.h"

using namespace ddm;

namespace {

constexpr int kWG = 256;
// LDS tile capacity in nodes. P2G keeps 4 floats per node and the G2P adjoint 3; with
// the staged particle records both stay under 40 KB of LDS per workgroup, so four
// workgroups (one wave per SIMD each) fit a CU.
constexpr int kTileP2G = 704;
constexpr int kTileG2P = 960;
constexpr int kHash = 512;       // open-addressing table of the chunk's stencil bases
constexpr int kTapeMax = 8;      // contact tape entries kept per node in the adjoint

__device__ __forceinline__ int node_index(const ddm_scene &S, int x, int y, int z) {
  int nby = S.dim[1] >> 2, nbz = S.dim[2] >> 2;
  int blk = ((x >> 2) * nby + (y >> 2)) * nbz + (z >> 2);
  return ((S.blk_begin + blk) << 6) | (((x & 3) << 4) | ((y & 3) << 2) | (z & 3));
}
__device__ __forceinline__ int block_of(const ddm_scene &S, int bx, int by, int bz) {
  return S.blk_begin + (bx * (S.dim[1] >> 2) + by) * (S.dim[2] >> 2) + bz;
}

struct BlockedV {
  const float4 *g;
  const ddm_scene *S;
  __device__ f3 operator()(int x, int y, int z) const {
    float4 a = g[node_index(*S, x, y, z)];
    return f3{a.x, a.y, a.z};
  }
};
struct BlockedGatherDMMV {
  const float4 *g;
  const ddm_scene *S;
  __device__ void operator()(int x, int y, int z, f3 &mv, float &m) const {
    float4 a = g[node_index(*S, x, y, z)];
    m = a.x;
    mv = f3{a.y, a.z, a.w};
  }
};

// ---------------------------------------------------------------- LDS tile helpers
struct Tile {
  int lo[3], t[3], n;
  __device__ int local(int x, int y, int z) const { return ((x - lo[0]) * t[1] + (y - lo[1])) * t[2] + (z - lo[2]); }
  __device__ bool has_x(int x) const { return x >= lo[0] && x < lo[0] + t[0]; }
};
// nodes per x-plane of the chunk box must fit the tile for the slab passes
template <int CAP>
__device__ __forceinline__ bool slabs_fit(const Tile &T) { return T.t[1] * T.t[2] <= CAP; }
// x-planes per slab pass
template <int CAP>
__device__ __forceinline__ int slab_width(const Tile &T) { return min(T.t[0], CAP / (T.t[1] * T.t[2])); }
// the slab of the chunk box starting at x-plane xs
__device__ __forceinline__ Tile slab(const Tile &T, int xs, int sx) {
  Tile L = T;
  L.lo[0] = xs;
  L.t[0] = min(sx, T.lo[0] + T.t[0] - xs);
  L.n = L.t[0] * T.t[1] * T.t[2];
  return L;
}

// ---------------------------------------------------------------- stencil groups
// Particles of a chunk that share a stencil base form a group. A scatter task is one
// (group, stencil node) pair: it sums the group's contributions to that node in
// registers and adds the sum to the tile with one atomic per component, instead of
// one atomic per particle.
//
// Per-particle stencil data staged in LDS, in group order. f1..f3 hold a vector and a
// row-major 3x3 matrix whose meaning is the kernel's (P2G: m v and the affine momentum
// matrix; G2P adjoint: node-velocity adjoint seed and C adjoint).
struct StencilRecs {
  float4 f0[kWG];    // fx (3), scalar
  float4 f1[kWG];    // vector (3), matrix[0]
  float4 f2[kWG];    // matrix[1..4]
  float4 f3[kWG];    // matrix[5..8]
  float w[kWG * 9];  // B-spline weights w[axis][offset], 9 consecutive per particle
};
__device__ __forceinline__ void stage_record(StencilRecs &rc, int s, const Stencil &st, float scalar, f3 vec,
                                             const m33 &M) {
  rc.f0[s] = make_float4(st.fx[0], st.fx[1], st.fx[2], scalar);
  rc.f1[s] = make_float4(vec.x, vec.y, vec.z, M.a[0]);
  rc.f2[s] = make_float4(M.a[1], M.a[2], M.a[3], M.a[4]);
  rc.f3[s] = make_float4(M.a[5], M.a[6], M.a[7], M.a[8]);
#pragma unroll
  for (int e = 0; e < 9; ++e) rc.w[9 * s + e] = st.w[e / 3][e % 3];
}
__device__ __forceinline__ f3 rec_vec(const StencilRecs &rc, int s) {
  float4 b = rc.f1[s];
  return f3{b.x, b.y, b.z};
}
__device__ __forceinline__ m33 rec_mat(const StencilRecs &rc, int s) {
  float4 b = rc.f1[s], c = rc.f2[s], d = rc.f3[s];
  return m33{{b.w, c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w}};
}
// stencil base packed into one non-negative int (grid dimensions are at most 1024)
__device__ __forceinline__ int pack_base(const int b[3]) { return (b[0] << 20) | (b[1] << 10) | b[2]; }
__device__ __forceinline__ void unpack_base(int key, int b[3]) {
  b[0] = key >> 20;
  b[1] = (key >> 10) & 1023;
  b[2] = key & 1023;
}

// LDS used only while grouping; it shares storage with the scatter tile.
struct GroupScratch {
  int hkey[kHash];  // base key per table entry, -1 when empty
  int hgrp[kHash];  // group index of the entry
  int cnt[kWG];     // group sizes
};
// Group layout shared with the scatter tasks.
struct Groups {
  int off[kWG + 1];  // group g holds record slots off[g] .. off[g+1]-1
  int key[kWG];      // packed stencil base of group g
  int wcnt[kWG / 64];
};

// Groups the chunk's valid slots (a prefix of the chunk) by stencil base. Returns the
// number of groups G (uniform) and, for a valid slot, the record slot pos where its
// data goes: groups are contiguous and in the order their first member was inserted.
// Ends with a barrier, after which the scratch is free.
__device__ int build_groups(bool valid, int key, int n_valid, GroupScratch &sc, Groups &gr, int &pos) {
  const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int i = t; i < kHash; i += kWG) sc.hkey[i] = -1;
  sc.cnt[t] = 0;
  __syncthreads();
  int h = 0;
  bool leader = false;
  if (valid) {
    // the low 3 bits of each coordinate: collision-free for bases spanning <= 8 cells per axis
    h = ((key >> 14) & 0x1c0) | ((key >> 7) & 0x38) | (key & 7);
    for (int probe = 0; probe < kHash; ++probe) {  // at most kWG keys: a free entry exists
      const int old = atomicCAS(&sc.hkey[h], -1, key);
      if (old == -1) {
        leader = true;
        break;
      }
      if (old == key) break;
      h = (h + 1) & (kHash - 1);
    }
  }
  unsigned long long m = __ballot(leader);
  if (lane == 0) gr.wcnt[wave] = __popcll(m);
  __syncthreads();
  int G = 0, off = 0;
#pragma unroll
  for (int w = 0; w < kWG / 64; ++w) {
    const int c = gr.wcnt[w];
    off += w < wave ? c : 0;
    G += c;
  }
  if (leader) {
    const int g = off + __popcll(m & ((1ull << lane) - 1ull));
    sc.hgrp[h] = g;
    gr.key[g] = key;
  }
  __syncthreads();
  int g = 0, rank = 0;
  if (valid) {
    g = sc.hgrp[h];
    rank = atomicAdd(&sc.cnt[g], 1);
  }
  __syncthreads();
  // exclusive scan of the group sizes
  const int c = t < G ? sc.cnt[t] : 0;
  int v = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) gr.wcnt[wave] = v;
  __syncthreads();
  off = 0;
#pragma unroll
  for (int w = 0; w < kWG / 64; ++w) off += w < wave ? gr.wcnt[w] : 0;
  if (t < G) gr.off[t] = off + v - c;
  if (t == 0) gr.off[G] = n_valid;
  __syncthreads();
  pos = valid ? gr.off[g] + rank : 0;
  return G;
}

// Every (group, stencil node) task whose node lies in slab L (all tasks when L is null):
// acc = sum over the group's records of contrib(rc, slot, i, j, k, w), then
// sink(X, Y, Z, acc). Tasks are spread over the workgroup's threads.
template <int NC, class Contrib, class Sink>
__device__ void group_tasks(int G, const Groups &gr, const StencilRecs &rc, const Tile *L, Contrib contrib,
                            Sink sink) {
  for (int task = (int)threadIdx.x; task < 27 * G; task += kWG) {
    const int g = task / 27, o = task - 27 * (task / 27);
    int i, j, k;
    offset_ijk(o, i, j, k);
    int b[3];
    unpack_base(gr.key[g], b);
    if (L && !L->has_x(b[0] + i)) continue;
    float acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.f;
    for (int q = gr.off[g], qe = gr.off[g + 1]; q < qe; ++q) {
      const float w = rc.w[9 * q + i] * rc.w[9 * q + 3 + j] * rc.w[9 * q + 6 + k];  // weight_dyn's order
      contrib(rc, q, i, j, k, w, acc);
    }
    sink(b[0] + i, b[1] + j, b[2] + k, acc);
  }
}

// the scatter tile shares LDS with the grouping scratch, which is dead by then
template <int NC, int CAP>
union ScatterLDS {
  float tile[NC][CAP];
  GroupScratch grp;
};

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// bounding box of the chunk's stencil bases -> tile covering all stencil nodes.
// Reduced across each wavefront first, then one LDS atomic per wave and bound.
__device__ Tile chunk_tile(bool valid, const int base[3], int *s_box) {
  if (threadIdx.x < 3) {
    s_box[threadIdx.x] = INT_MAX;
    s_box[3 + threadIdx.x] = INT_MIN;
  }
  __syncthreads();
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = wave_min(valid ? base[d] : INT_MAX);
    hi[d] = wave_max(valid ? base[d] : INT_MIN);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      atomicMin(&s_box[d], lo[d]);
      atomicMax(&s_box[3 + d], hi[d]);
    }
  }
  __syncthreads();
  Tile T;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T.lo[d] = s_box[d];
    T.t[d] = s_box[3 + d] - s_box[d] + 3;
  }
  T.n = T.t[0] * T.t[1] * T.t[2];
  __syncthreads();
  return T;
}

// mark a block of scene si; the first marker of a substep appends it to the block list
__device__ __forceinline__ void mark_block(const ddm_batch &B, int blk, int si) {
  if (atomicExch(&B.block_flags[blk], 1u) == 0u) {
    const int e = atomicAdd(B.block_count, 1);
    reinterpret_cast<int2 *>(B.block_list)[e] = make_int2(blk, si);
  }
}
// mark every node block that intersects the tile
__device__ void flag_tile_blocks(const ddm_batch &B, const ddm_scene &S, int si, const Tile &T) {
  int b0[3], nb[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b0[d] = T.lo[d] >> 2;
    nb[d] = ((T.lo[d] + T.t[d] - 1) >> 2) - b0[d] + 1;
  }
  int tot = nb[0] * nb[1] * nb[2];
  for (int i = threadIdx.x; i < tot; i += kWG) {
    int bx = i / (nb[1] * nb[2]), by = (i / nb[2]) % nb[1], bz = i % nb[2];
    mark_block(B, block_of(S, b0[0] + bx, b0[1] + by, b0[2] + bz), si);
  }
}
// mark the (at most 8) blocks of one stencil
__device__ void flag_stencil_blocks(const ddm_batch &B, const ddm_scene &S, int si, const int base[3]) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c)
        mark_block(B, block_of(S, (base[0] + 2 * a) >> 2, (base[1] + 2 * b) >> 2, (base[2] + 2 * c) >> 2), si);
}

// ---------------------------------------------------------------- P2G (+ F advection, SVD, return map)
enum P2GMode { kForward = 0, kRecompute = 1, kForwardTaped = 2 };
// kForward     : forward substep, writes next.F.
// kRecompute   : adjoint recompute, writes the F_tmp/U/V/sigma temporaries instead.
// kForwardTaped: writes both (the taped adjoint reads the temporaries).
// need (optional): the kernel runs only if *need != 0 (a taped adjoint that has to
// recompute decides this on the device).
template <int MODE>
__global__ __launch_bounds__(kWG) void k_p2g(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                             const ddm_state *__restrict__ next_tab, const int32_t *need) {
  if (need && *need == 0) return;  // uniform over the grid
  __shared__ ScatterLDS<4, kTileP2G> sh;
  __shared__ StencilRecs rc;
  __shared__ Groups gr;
  __shared__ int s_box[6];
  float *const tm = sh.tile[0], *const tx = sh.tile[1], *const ty = sh.tile[2], *const tz = sh.tile[3];
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[blockIdx.x];
  const ddm_scene S = B.scenes[ch.x];
  const ddm_state cur = cur_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  const bool valid = p < ch.z;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  P2GParticle q;
  int key = 0;
  if (valid) {
    m33 F = ld9(cur.F + 9 * (size_t)p), C = ld9(cur.C + 9 * (size_t)p);
    m33 Ft = advect_F(F, C, S.dt);
    m33 U, V;
    f3 sig;
    svd3(Ft, U, sig, V);
    m33 Fn;
    p2g_particle(ld3(cur.x + 3 * (size_t)p), ld3(cur.v + 3 * (size_t)p), C, Ft, U, sig, V, S.mass[p], S.vol[p],
                 ld3(S.mu_lam_yield + 3 * (size_t)p), S.dx, S.inv_dx, S.dt, dim, Fn, q);
    if (MODE != kForward) {
      st9(S.tmp_Ft + 9 * (size_t)p, Ft);
      st9(S.tmp_U + 9 * (size_t)p, U);
      st9(S.tmp_V + 9 * (size_t)p, V);
      st3(S.tmp_sig + 3 * (size_t)p, sig);
    }
    if (MODE != kRecompute) st9(next_tab[ch.x].F + 9 * (size_t)p, Fn);
    key = pack_base(q.st.base);
  } else {
    q.st.base[0] = q.st.base[1] = q.st.base[2] = 0;
  }
  const Tile T = chunk_tile(valid, q.st.base, s_box);
  int pos;
  const int G = build_groups(valid, key, min(kWG, ch.z - ch.y), sh.grp, gr, pos);
  if (valid) stage_record(rc, pos, q.st, q.m, q.mv, q.affine);
  const float dx = S.dx;
  auto contrib = [dx](const StencilRecs &r, int s, int i, int j, int k, float w, float *acc) {
    float cm;
    f3 cmv;
    float4 a = r.f0[s];
    p2g_node_contrib(w, f3{a.x, a.y, a.z}, a.w, rec_vec(r, s), rec_mat(r, s), dx, i, j, k, cm, cmv);
    acc[0] += cm;
    acc[1] += cmv.x;
    acc[2] += cmv.y;
    acc[3] += cmv.z;
  };
  float *gmmv = B.grid_mmv;
  if (slabs_fit<kTileP2G>(T)) {
    const int sx = slab_width<kTileP2G>(T);
    for (int xs = T.lo[0]; xs < T.lo[0] + T.t[0]; xs += sx) {  // uniform over the workgroup
      const Tile L = slab(T, xs, sx);
      if (xs != T.lo[0]) __syncthreads();  // the previous pass has flushed the tile
      for (int i = threadIdx.x; i < L.n; i += kWG) tm[i] = tx[i] = ty[i] = tz[i] = 0.f;
      __syncthreads();  // also publishes the staged records
      group_tasks<4>(G, gr, rc, &L, contrib, [&](int X, int Y, int Z, const float *a) {
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
    flag_tile_blocks(B, S, ch.x, T);
  } else {
    __syncthreads();  // publishes the staged records
    group_tasks<4>(G, gr, rc, nullptr, contrib, [&](int X, int Y, int Z, const float *a) {
      float *g = gmmv + 4 * (size_t)node_index(S, X, Y, Z);
      atomicAdd(g + 0, a[0]);
      atomicAdd(g + 1, a[1]);
      atomicAdd(g + 2, a[2]);
      atomicAdd(g + 3, a[3]);
    });
    for (int g = threadIdx.x; g < G; g += kWG) {
      int b[3];
      unpack_base(gr.key[g], b);
      flag_stencil_blocks(B, S, ch.x, b);
    }
  }
}

// ---------------------------------------------------------------- grid helpers
struct BlockCoord {
  int ix, iy, iz;
};
__device__ __forceinline__ BlockCoord block_node(const ddm_scene &S, int blk, int lane) {
  int lb = blk - S.blk_begin;
  int nby = S.dim[1] >> 2, nbz = S.dim[2] >> 2;
  int bx = lb / (nby * nbz), by = (lb / nbz) % nby, bz = lb % nbz;
  return BlockCoord{bx * 4 + (lane >> 4), by * 4 + ((lane >> 2) & 3), bz * 4 + (lane & 3)};
}
__device__ __forceinline__ NodeParams scene_node_params(const ddm_scene &S, const ddm_state &cur,
                                                        const ddm_state &next) {
  NodeParams P;
  P.dim[0] = S.dim[0];
  P.dim[1] = S.dim[1];
  P.dim[2] = S.dim[2];
  P.nb = S.n_bodies;
  P.dx = S.dx;
  P.dt = S.dt;
  P.gf = S.ground_friction;
  P.gravity = f3{S.gravity[0], S.gravity[1], S.gravity[2]};
  P.prim = S.body_prim;
  P.args = S.body_args;
  P.pos = cur.body_pos;
  P.rot = cur.body_rot;
  P.npos = next.body_pos;
  P.nrot = next.body_rot;
  return P;
}

// ---------------------------------------------------------------- grid update
// ZERO=true (forward substep): consumes (m, mv) and clears them and the flags.
// SAVE=true (taped forward): also copies each listed block's (m, mv) and v, and the list,
// into the tape (entries beyond its capacity are dropped; status[0] keeps the count).
// need (optional): run only if *need != 0.
template <bool ZERO, bool SAVE>
__global__ __launch_bounds__(kWG) void k_grid(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                              const ddm_state *__restrict__ next_tab, ddm_grid_tape tape,
                                              const int32_t *need) {
  if (need && *need == 0) return;  // uniform over the grid
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float4 *mmv = reinterpret_cast<float4 *>(B.grid_mmv);
  float4 *gv = reinterpret_cast<float4 *>(B.grid_v);
  const int n_listed = *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  if (SAVE && blockIdx.x == 0 && threadIdx.x == 0) tape.status[0] = n_listed;
  for (int e = blockIdx.x * (kWG / 64) + wave; e < n_listed; e += gridDim.x * (kWG / 64)) {
    const int2 entry = list[e];
    const int blk = entry.x, si = entry.y;
    const ddm_scene &S = B.scenes[si];
    BlockCoord c = block_node(S, blk, lane);
    size_t gi = ((size_t)blk << 6) | (size_t)lane;
    float4 a = mmv[gi];
    f3 v{0.f, 0.f, 0.f};
    if (node_active(a.x)) {
      NodeParams P = scene_node_params(S, cur_tab[si], next_tab[si]);
      v = node_fwd(a.x, f3{a.y, a.z, a.w}, c.ix, c.iy, c.iz, P);
    }
    gv[gi] = make_float4(v.x, v.y, v.z, 0.f);
    if (SAVE && e < tape.capacity) {
      const size_t ti = ((size_t)e << 6) | (size_t)lane;
      reinterpret_cast<float4 *>(tape.grid_mmv)[ti] = a;
      reinterpret_cast<float4 *>(tape.grid_v)[ti] = make_float4(v.x, v.y, v.z, 0.f);
      if (lane == 0) reinterpret_cast<int2 *>(tape.block_list)[e] = entry;
    }
    if (ZERO) {
      mmv[gi] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lane == 0) B.block_flags[blk] = 0u;
    }
  }
}

// Taped adjoint: puts a substep's grid back into the arena from its tape, or, when the
// forward listed more blocks than the tape holds, leaves the (clear) arena alone and sets
// status[1] so that the recompute kernels run.
__global__ __launch_bounds__(kWG) void k_restore(ddm_batch B, ddm_grid_tape tape) {
  const int n = tape.status[0];
  const bool use = n <= tape.capacity;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    tape.status[1] = use ? 0 : 1;
    if (use) *B.block_count = n;
  }
  if (!use) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int2 *saved = reinterpret_cast<const int2 *>(tape.block_list);
  for (int e = blockIdx.x * (kWG / 64) + wave; e < n; e += gridDim.x * (kWG / 64)) {
    const int2 entry = saved[e];
    const size_t gi = ((size_t)entry.x << 6) | (size_t)lane, ti = ((size_t)e << 6) | (size_t)lane;
    reinterpret_cast<float4 *>(B.grid_mmv)[gi] = reinterpret_cast<const float4 *>(tape.grid_mmv)[ti];
    reinterpret_cast<float4 *>(B.grid_v)[gi] = reinterpret_cast<const float4 *>(tape.grid_v)[ti];
    if (lane == 0) reinterpret_cast<int2 *>(B.block_list)[e] = entry;
  }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// adjoint of the grid update; clears (m, mv), the velocity adjoint and the flags
__global__ __launch_bounds__(kWG) void k_grid_adj(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                                  const ddm_state *__restrict__ next_tab,
                                                  const ddm_state *__restrict__ cur_grad_tab,
                                                  const ddm_state *__restrict__ next_grad_tab) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float4 *mmv = reinterpret_cast<float4 *>(B.grid_mmv);
  float4 *dv = reinterpret_cast<float4 *>(B.grid_dv);
  float4 *dmmv = reinterpret_cast<float4 *>(B.grid_dmmv);
  const int n_listed = *B.block_count;
  const int2 *list = reinterpret_cast<const int2 *>(B.block_list);
  for (int e = blockIdx.x * (kWG / 64) + wave; e < n_listed; e += gridDim.x * (kWG / 64)) {
    const int2 entry = list[e];
    const int blk = entry.x, si = entry.y;
    const ddm_scene &S = B.scenes[si];
    BlockCoord c = block_node(S, blk, lane);
    size_t gi = ((size_t)blk << 6) | (size_t)lane;
    float4 a = mmv[gi];
    float4 d = dv[gi];
    const float m = a.x;
    const f3 mv{a.y, a.z, a.w};
    const bool act = node_active(m);
    NodeParams P = scene_node_params(S, cur_tab[si], next_tab[si]);
    const f3 gx = node_position(c.ix, c.iy, c.iz, P.dx);
    const int mode = S.adjoint_mode;
    // forward replay keeping the velocities entering each contact (integrator.cu:678-732)
    f3 tape_v[kTapeMax];
    int tape_b[kTapeMax];
    int nt = 0;
    f3 v0{0.f, 0.f, 0.f}, g{0.f, 0.f, 0.f};
    if (act) {
      f3 v = mv * (1.f / m);
      v += P.dt * P.gravity;
      v0 = v;
      for (int b = 0; b < P.nb; ++b) {
        f3 vin = v;
        if (contact(v, gx, load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b), P.dt) && nt < kTapeMax) {
          tape_v[nt] = vin;
          tape_b[nt] = b;
          ++nt;
        }
      }
      g = boundary_adj(v, f3{d.x, d.y, d.z}, c.ix, c.iy, c.iz, P, mode);
    }
    // bodies in reverse order, uniformly across the wave so pose gradients can be reduced
    for (int b = P.nb - 1; b >= 0; --b) {
      PoseGrad pg = posegrad_zero();
      bool mine = false;
      if (act) {
        Body Bd = load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b);
        float infl;
        if (in_contact(prim_sdf(Bd.prim, Bd.args, to_body(Bd.pos, Bd.rot, gx)), Bd.prim[2], infl)) {
          mine = true;
          f3 vin;
          if (nt > 0 && tape_b[nt - 1] == b) {
            vin = tape_v[--nt];
          } else {  // tape overflow: replay the contacts before b
            vin = v0;
            for (int b2 = 0; b2 < b; ++b2)
              contact(vin, gx, load_body(P.prim, P.args, P.pos, P.rot, P.npos, P.nrot, b2), P.dt);
          }
          contact_adj(vin, gx, Bd, P.dt, mode, g, pg);
        }
      }
      if (__any(mine)) {
        float vals[14] = {pg.pos.x,  pg.pos.y,  pg.pos.z,  pg.rot.w,  pg.rot.x,  pg.rot.y,  pg.rot.z,
                          pg.npos.x, pg.npos.y, pg.npos.z, pg.nrot.w, pg.nrot.x, pg.nrot.y, pg.nrot.z};
#pragma unroll
        for (int e = 0; e < 14; ++e) vals[e] = wave_sum(vals[e]);
        if (lane == 0) {
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
  }
}

// ---------------------------------------------------------------- G2P
__global__ __launch_bounds__(kWG) void k_g2p(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                             const ddm_state *__restrict__ next_tab) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.block_count = 0;  // the grid kernels have finished
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[blockIdx.x];
  const ddm_scene S = B.scenes[ch.x];
  const ddm_state cur = cur_tab[ch.x], next = next_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  if (p >= ch.z) return;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  BlockedV gv{reinterpret_cast<const float4 *>(B.grid_v), &S};
  f3 xo, vo;
  m33 Co;
  g2p_particle(ld3(cur.x + 3 * (size_t)p), gv, S.inv_dx, S.dx, S.dt, S.ground_height, dim, xo, vo, Co);
  st3(next.x + 3 * (size_t)p, xo);
  st3(next.v + 3 * (size_t)p, vo);
  st9(next.C + 9 * (size_t)p, Co);
}

struct NullScatter {
  __device__ void operator()(int, int, int, f3) const {}
};

// adjoint of G2P: adds to the x adjoint and scatters the node-velocity adjoint
__global__ __launch_bounds__(kWG) void k_g2p_adj(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                                 const ddm_state *__restrict__ next_tab,
                                                 const ddm_state *__restrict__ cur_grad_tab,
                                                 const ddm_state *__restrict__ next_grad_tab) {
  __shared__ ScatterLDS<3, kTileG2P> sh;
  __shared__ StencilRecs rc;
  __shared__ Groups gr;
  __shared__ int s_box[6];
  float *const tx = sh.tile[0], *const ty = sh.tile[1], *const tz = sh.tile[2];
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[blockIdx.x];
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
    NullScatter none;
    gC = ld9(next_grad.C + 9 * (size_t)p);
    f3 gxp = g2p_particle_adj(ld3(cur.x + 3 * (size_t)p), ld3(next.v + 3 * (size_t)p),
                              ld3(next_grad.x + 3 * (size_t)p), ld3(next_grad.v + 3 * (size_t)p), gC, gv, none,
                              S.inv_dx, S.dx, S.dt, S.ground_height, dim, &g_nv, &st);
    add3(cur_grad.x + 3 * (size_t)p, gxp);
    key = pack_base(st.base);
  } else {
    st.base[0] = st.base[1] = st.base[2] = 0;
  }
  const Tile T = chunk_tile(valid, st.base, s_box);
  int pos;
  const int G = build_groups(valid, key, min(kWG, ch.z - ch.y), sh.grp, gr, pos);
  if (valid) stage_record(rc, pos, st, 0.f, g_nv, gC);
  const float inv_dx = S.inv_dx;
  auto contrib = [inv_dx](const StencilRecs &r, int s, int i, int j, int k, float w, float *acc) {
    float4 a = r.f0[s];
    f3 g = g2p_adj_contrib(w, f3{a.x, a.y, a.z}, rec_vec(r, s), rec_mat(r, s), inv_dx, i, j, k);
    acc[0] += g.x;
    acc[1] += g.y;
    acc[2] += g.z;
  };
  if (slabs_fit<kTileG2P>(T)) {
    const int sx = slab_width<kTileG2P>(T);
    for (int xs = T.lo[0]; xs < T.lo[0] + T.t[0]; xs += sx) {  // uniform over the workgroup
      const Tile L = slab(T, xs, sx);
      if (xs != T.lo[0]) __syncthreads();  // the previous pass has flushed the tile
      for (int i = threadIdx.x; i < L.n; i += kWG) tx[i] = ty[i] = tz[i] = 0.f;
      __syncthreads();  // also publishes the staged records
      group_tasks<3>(G, gr, rc, &L, contrib, [&](int X, int Y, int Z, const float *a) {
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
    group_tasks<3>(G, gr, rc, nullptr, contrib, [&](int X, int Y, int Z, const float *a) {
      float *g = B.grid_dv + 4 * (size_t)node_index(S, X, Y, Z);
      atomicAdd(g + 0, a[0]);
      atomicAdd(g + 1, a[1]);
      atomicAdd(g + 2, a[2]);
    });
  }
}

// adjoint of P2G + SVD adjoint + chain through F_tmp = (I + dt C) F
__global__ __launch_bounds__(kWG) void k_p2g_adj(ddm_batch B, const ddm_state *__restrict__ cur_tab,
                                                 const ddm_state *__restrict__ cur_grad_tab,
                                                 const ddm_state *__restrict__ next_grad_tab) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *B.block_count = 0;  // the grid kernels have finished
  const int4 ch = reinterpret_cast<const int4 *>(B.chunks)[blockIdx.x];
  const ddm_scene S = B.scenes[ch.x];
  const ddm_state cur = cur_tab[ch.x], cur_grad = cur_grad_tab[ch.x], next_grad = next_grad_tab[ch.x];
  const int p = ch.y + (int)threadIdx.x;
  if (p >= ch.z) return;
  const int dim[3] = {S.dim[0], S.dim[1], S.dim[2]};
  const size_t p3 = 3 * (size_t)p, p9 = 9 * (size_t)p;
  BlockedGatherDMMV gat{reinterpret_cast<const float4 *>(B.grid_dmmv), &S};
  m33 F = ld9(cur.F + p9), C = ld9(cur.C + p9);
  m33 Ft = ld9(S.tmp_Ft + p9), U = ld9(S.tmp_U + p9), V = ld9(S.tmp_V + p9);
  f3 sig = ld3(S.tmp_sig + p3);
  P2GAdj a;
  p2g_particle_adj(ld3(cur.x + p3), ld3(cur.v + p3), C, Ft, U, sig, V, S.mass[p], S.vol[p],
                   ld3(S.mu_lam_yield + p3), ld9(next_grad.F + p9), gat, S.dx, S.inv_dx, S.dt, dim, a);
  m33 gFt, gF, gC;
  svd_adj(F, C, U, V, sig, a.Ft, a.U, a.V, a.sig, S.dt, gFt, gF, gC);
  add3(cur_grad.x + p3, a.x);
  add3(cur_grad.v + p3, a.v);
  m33 cg = ld9(cur_grad.C + p9);
  cg = madd(madd(cg, a.C), gC);  // p2g_grad adds first, then compute_svd_grad (simulator.py:584-585)
  st9(cur_grad.C + p9, cg);
  add9(cur_grad.F + p9, gF);
}

__global__ void k_zero(float4 *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_zero_u32(uint32_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

inline int grid_wgs(const ddm_batch *b) {
  int w = (b->n_blocks + (kWG / 64) - 1) / (kWG / 64);
  return w < 4096 ? (w > 0 ? w : 1) : 4096;
}
inline int check(const ddm_batch *b, const void *cur, const void *next) {
  if (!b || !cur || !next || b->n_chunks < 0 || b->n_blocks < 0 || b->n_scenes <= 0) return DDM_EINVAL;
  if (b->max_bodies > DDM_MAX_BODIES) return DDM_ETOOMANYBODIES;
  if (!b->scenes || !b->chunks || !b->grid_mmv || !b->grid_v || !b->block_flags || !b->block_list ||
      !b->block_count)
    return DDM_EINVAL;
  return 0;
}

}  // namespace

#define DDM_API extern "C" __attribute__((visibility("default")))

static inline void mark(hipEvent_t *ev, int i, hipStream_t stream) {
  if (ev) (void)hipEventRecord(ev[i], stream);
}

static int launch_forward(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, hipStream_t stream,
                          hipEvent_t *ev) {
  int e = check(b, cur, next);
  if (e) return e;
  if (b->n_chunks == 0) return 0;
  mark(ev, 0, stream);
  k_p2g<kForward><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, nullptr);
  mark(ev, 1, stream);
  k_grid<true, false><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, ddm_grid_tape{}, nullptr);
  mark(ev, 2, stream);
  k_g2p<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next);
  mark(ev, 3, stream);
  return (int)hipGetLastError();
}

static int launch_backward(const ddm_batch *b, const ddm_state *cur, const ddm_state *next, const ddm_state *cur_grad,
                           const ddm_state *next_grad, hipStream_t stream, hipEvent_t *ev) {
  int e = check(b, cur, next);
  if (e) return e;
  if (!cur_grad || !next_grad || !b->grid_dv || !b->grid_dmmv) return DDM_EINVAL;
  if (b->n_chunks == 0) return 0;
  mark(ev, 0, stream);
  k_p2g<kRecompute><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, nullptr);
  mark(ev, 1, stream);
  k_grid<false, false><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, ddm_grid_tape{}, nullptr);
  mark(ev, 2, stream);
  k_g2p_adj<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, cur_grad, next_grad);
  mark(ev, 3, stream);
  k_grid_adj<<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, cur_grad, next_grad);
  mark(ev, 4, stream);
  k_p2g_adj<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, cur_grad, next_grad);
  mark(ev, 5, stream);
  return (int)hipGetLastError();
}

static inline int check_tape(const ddm_grid_tape *t) {
  if (!t || !t->grid_mmv || !t->grid_v || !t->block_list || !t->status || t->capacity < 0) return DDM_EINVAL;
  return 0;
}

static int launch_forward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                const ddm_grid_tape *tape, hipStream_t stream, hipEvent_t *ev) {
  int e = check(b, cur, next);
  if (!e) e = check_tape(tape);
  if (e) return e;
  if (b->n_chunks == 0) return 0;
  mark(ev, 0, stream);
  k_p2g<kForwardTaped><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, nullptr);
  mark(ev, 1, stream);
  k_grid<true, true><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, *tape, nullptr);
  mark(ev, 2, stream);
  k_g2p<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next);
  mark(ev, 3, stream);
  return (int)hipGetLastError();
}

static int launch_backward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                 const ddm_state *cur_grad, const ddm_state *next_grad, const ddm_grid_tape *tape,
                                 hipStream_t stream, hipEvent_t *ev) {
  int e = check(b, cur, next);
  if (!e) e = check_tape(tape);
  if (e) return e;
  if (!cur_grad || !next_grad || !b->grid_dv || !b->grid_dmmv) return DDM_EINVAL;
  if (b->n_chunks == 0) return 0;
  const int32_t *need = tape->status + 1;
  mark(ev, 0, stream);
  k_restore<<<grid_wgs(b), kWG, 0, stream>>>(*b, *tape);
  k_p2g<kRecompute><<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, need);
  k_grid<false, false><<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, ddm_grid_tape{}, need);
  mark(ev, 1, stream);
  k_g2p_adj<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, next, cur_grad, next_grad);
  mark(ev, 2, stream);
  k_grid_adj<<<grid_wgs(b), kWG, 0, stream>>>(*b, cur, next, cur_grad, next_grad);
  mark(ev, 3, stream);
  k_p2g_adj<<<b->n_chunks, kWG, 0, stream>>>(*b, cur, cur_grad, next_grad);
  mark(ev, 4, stream);
  return (int)hipGetLastError();
}

DDM_API int ddm_substep_forward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                      const ddm_grid_tape *tape, hipStream_t stream) {
  return launch_forward_taped(b, cur, next, tape, stream, nullptr);
}

DDM_API int ddm_substep_backward_taped(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                       const ddm_state *cur_grad, const ddm_state *next_grad,
                                       const ddm_grid_tape *tape, hipStream_t stream) {
  return launch_backward_taped(b, cur, next, cur_grad, next_grad, tape, stream, nullptr);
}

DDM_API int ddm_substep_forward_taped_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                            const ddm_grid_tape *tape, hipStream_t stream, hipEvent_t *events) {
  if (!events) return DDM_EINVAL;
  return launch_forward_taped(b, cur, next, tape, stream, events);
}

DDM_API int ddm_substep_backward_taped_timed(const ddm_batch *b, const ddm_state *cur, const ddm_state *next,
                                             const ddm_state *cur_grad, const ddm_state *next_grad,
                                             const ddm_grid_tape *tape, hipStream_t stream, hipEvent_t *events) {
  if (!events) return DDM_EINVAL;
  return launch_backward_taped(b, cur, next, cur_grad, next_grad, tape, stream, events);
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
  return (int)hipGetLastError();
}
