// Shared device code of the fused multi-scene substep (forward kernels in
// fused_kernels.hip, adjoint kernels in fused_adjoint.hip).
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
//  * P2G and the G2P adjoint group the chunk's particles by stencil base (a hash table
//    and a counting sort in LDS), stage each particle's stencil data in LDS in group
//    order and scatter one (group, stencil node) task per thread: the task sums its
//    group in registers and adds one value per component to an LDS tile covering the
//    chunk's stencil box. LDS float atomics are the throughput limit of this scatter, so
//    a group of n particles costs one atomic instead of n, whatever the particle order.
//  * The tile is flushed to the grid with global atomics for its non-zero nodes. A box
//    larger than the tile is processed in x-slabs, one tile pass per slab; only a box
//    whose single x-plane exceeds the tile sends the tasks' sums to global atomics.
//  * The contact adjoint reduces body-pose gradients across the wavefront before one
//    atomic per body per wave.
#pragma once
#include <limits.h>
#include "mpm_device.h"
#include "../../include/dexdeform_mpm.h"

namespace ddm_fused {

using namespace ddm;

namespace {

constexpr int kWG = 256;
// LDS tile capacity in nodes. P2G keeps 4 floats per node and the G2P adjoint 3; with
// the staged particle records both stay under 32 KB of LDS per workgroup, so five
// workgroups (one wave per SIMD each) fit a CU.
constexpr int kTileP2G = 704;
constexpr int kTileG2P = 960;
constexpr int kHash = 512;       // open-addressing table of the chunk's stencil bases
constexpr int kTapeMax = 8;      // contact tape entries kept per node in the adjoint
// Particle launches of at most this many chunks (256-thread workgroups) put at most one
// workgroup on each of the 256 CUs, one wave per SIMD: they run at one chunk's latency,
// not at a throughput bound, and use the latency variants of the gathering kernels (loads
// issued together; as many registers as one wave per SIMD allows).
constexpr int kLatencyChunks = 256;
// Forward launches of at most kLatencyChunks chunks over at most this many node blocks
// skip the block list: P2G marks its blocks with plain stores of scene + 1 instead of two
// dependent device-scope atomics with return (which mostly overlap the drain of its own
// tile flush), and the grid update scans every block's flag, one wave per block, instead
// of reading a list entry and then the scene: one dependent load fewer on its chain.
constexpr int kScanBlocks = 16384;

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
// The B-spline weights are recomputed from fx by the tasks (bspline_wi): fewer LDS reads
// and 9 KB less LDS per workgroup than staging them.
struct StencilRecs {
  float4 f0[kWG];  // fx (3), scalar
  float4 f1[kWG];  // vector (3), matrix[0]
  float4 f2[kWG];  // matrix[1..4]
  float4 f3[kWG];  // matrix[5..8]
};
// B-spline weight of offset i (0..2) for the fractional position f: stencil()'s expressions
__device__ __forceinline__ float bspline_wi(float f, int i) {
  const float a = i == 0 ? 1.5f - f : (i == 1 ? f - 1.f : f - 0.5f);
  return i == 1 ? 0.75f - a * a : 0.5f * (a * a);
}
__device__ __forceinline__ void stage_record(StencilRecs &rc, int s, const Stencil &st, float scalar, f3 vec,
                                             const m33 &M) {
  rc.f0[s] = make_float4(st.fx[0], st.fx[1], st.fx[2], scalar);
  rc.f1[s] = make_float4(vec.x, vec.y, vec.z, M.a[0]);
  rc.f2[s] = make_float4(M.a[1], M.a[2], M.a[3], M.a[4]);
  rc.f3[s] = make_float4(M.a[5], M.a[6], M.a[7], M.a[8]);
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

// Every (group, i, j) task -- the column of 3 stencil nodes (i, j, 0..2) of a group -- whose
// nodes lie in slab L (all tasks when L is null): acc[k] = sum over the group's records of
// the contribution to node (i, j, k), then sink(X, Y, Z, acc[k]) for the 3 nodes. The
// column shares each record read, the weight product w_x w_y and the x and y terms of the
// affine product between its nodes; contrib(rc, slot, i, j, acc) adds one record's three
// contributions with the same operations, in the same order, as per-node code would.
// When the 9 G tasks leave lanes of the workgroup idle (small chunks: a few dozen groups),
// each task's records are dealt to P parts (record q to part (q - first) mod P) on
// neighbouring lanes, whose sums meet in sink's atomics: the longest task, which sets how
// long a wave takes, is then about P times shorter. P = 1 when the tasks fill the lanes.
template <int NC, class Contrib, class Sink>
__device__ void group_column_tasks(int G, const Groups &gr, const StencilRecs &rc, const Tile *L, Contrib contrib,
                                   Sink sink) {
  const int P = G > 0 ? max(1, min(4, kWG / (9 * G))) : 1;  // uniform
  for (int task = (int)threadIdx.x; task < 9 * G * P; task += kWG) {
    const int gt = task / P, part = task - P * gt;
    const int g = gt / 9, ij = gt - 9 * g;
    const int i = ij / 3, j = ij - 3 * (ij / 3);
    const int q0 = gr.off[g] + part, qe = gr.off[g + 1];
    if (q0 >= qe) continue;  // a part with no records
    int b[3];
    unpack_base(gr.key[g], b);
    if (L && !L->has_x(b[0] + i)) continue;
    float acc[3][NC];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[k][c] = 0.f;
    for (int q = q0; q < qe; q += P) contrib(rc, q, i, j, acc);
#pragma unroll
    for (int k = 0; k < 3; ++k) sink(b[0] + i, b[1] + j, b[2] + k, acc[k]);
  }
}

// the scatter tile shares LDS with the grouping scratch, which is dead by then
template <int NC, int CAP>
union ScatterLDS {
  float tile[NC][CAP];
  GroupScratch grp;
};

// v from the lane that DPP control CTRL selects; identity on rows outside ROW_MASK
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_or(int v, int identity) {
  return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROW_MASK, 0xf, false);
}
// minimum / maximum over the wave, the same value on every lane (the DPP steps of wave_sum)
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, dpp_or<0xb1>(v, INT_MAX));
  v = min(v, dpp_or<0x4e>(v, INT_MAX));
  v = min(v, dpp_or<0x141>(v, INT_MAX));
  v = min(v, dpp_or<0x140>(v, INT_MAX));
  v = min(v, dpp_or<0x142, 0xa>(v, INT_MAX));
  v = min(v, dpp_or<0x143, 0xc>(v, INT_MAX));
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max(int v) {
  v = max(v, dpp_or<0xb1>(v, INT_MIN));
  v = max(v, dpp_or<0x4e>(v, INT_MIN));
  v = max(v, dpp_or<0x141>(v, INT_MIN));
  v = max(v, dpp_or<0x140>(v, INT_MIN));
  v = max(v, dpp_or<0x142, 0xa>(v, INT_MIN));
  v = max(v, dpp_or<0x143, 0xc>(v, INT_MIN));
  return __builtin_amdgcn_readlane(v, 63);
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
// mark the (at most 8) blocks of one stencil with scene + 1, plain stores, no list (for
// grid kernels that scan the flags)
__device__ void store_stencil_flags(const ddm_batch &B, const ddm_scene &S, int si, const int base[3]) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c)
        B.block_flags[block_of(S, (base[0] + 2 * a) >> 2, (base[1] + 2 * b) >> 2, (base[2] + 2 * c) >> 2)] =
            (uint32_t)(si + 1);
}
// mark the (at most 8) blocks of one stencil
__device__ void flag_stencil_blocks(const ddm_batch &B, const ddm_scene &S, int si, const int base[3]) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c)
        mark_block(B, block_of(S, (base[0] + 2 * a) >> 2, (base[1] + 2 * b) >> 2, (base[2] + 2 * c) >> 2), si);
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

// Body parameters of one scene, staged in LDS for one wave: (type, friction, softness,
// round), args, and the poses at the substep and the next, each as its own array as in
// global memory (so load_body and near_bodies read them unchanged)
constexpr int kBodyFloats = 22;
struct WaveBodies {
  float f[kBodyFloats * DDM_MAX_BODIES];
};
// Copies P's bodies into W (lane b copies body b) and points P at the copy. The grid
// kernels then read a body's parameters with LDS loads in their loops over bodies instead
// of a vector-memory round trip per body (they cannot be scalar loads: the compiler cannot
// prove that the kernel's own stores leave them alone). All 22 loads are issued before
// the first LDS store: written as load-then-store pairs, each load would wait for the
// store before it (the source pointers are generic and might alias the copy), 22
// dependent memory round trips per staging. Call with all 64 lanes of the wave; LDS
// operations of one wave execute in order, so its later reads see the copy.
__device__ __forceinline__ void stage_bodies(NodeParams &P, WaveBodies &W, int lane) {
  float *prim = W.f, *args = W.f + 4 * DDM_MAX_BODIES, *pos = W.f + 8 * DDM_MAX_BODIES,
        *rot = W.f + 11 * DDM_MAX_BODIES, *npos = W.f + 15 * DDM_MAX_BODIES, *nrot = W.f + 18 * DDM_MAX_BODIES;
  if (lane < P.nb) {
    float r[kBodyFloats];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      r[k] = P.prim[4 * lane + k];
      r[4 + k] = P.args[4 * lane + k];
      r[8 + k] = P.rot[4 * lane + k];
      r[12 + k] = P.nrot[4 * lane + k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      r[16 + k] = P.pos[3 * lane + k];
      r[19 + k] = P.npos[3 * lane + k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      prim[4 * lane + k] = r[k];
      args[4 * lane + k] = r[4 + k];
      rot[4 * lane + k] = r[8 + k];
      nrot[4 * lane + k] = r[12 + k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      pos[3 * lane + k] = r[16 + k];
      npos[3 * lane + k] = r[19 + k];
    }
  }
  P.prim = prim;
  P.args = args;
  P.pos = pos;
  P.rot = rot;
  P.npos = npos;
  P.nrot = nrot;
}

// Index of the calling wave in its workgroup, as a value the compiler knows to be the same
// on every lane (threadIdx.x >> 6 is, but divergence analysis cannot tell): list entries,
// scene records and body parameters addressed through it are then read with scalar loads
// into SGPRs instead of a vector load per lane.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// v combined with the lane that DPP control CTRL selects; rows (16 lanes) outside ROW_MASK
// keep 0, so the sum below adds nothing there
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_or_zero(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
// Sum over the 64 lanes of a wave, the same value on every lane. Data-parallel-primitive
// moves (VALU operations, no LDS round trip): lane ^ 1, lane ^ 2, the half-row and row
// mirrors give every lane its row's sum; row_bcast:15 and row_bcast:31 carry rows 0 and
// 0-1 up to row 3, whose lane 63 then holds the total. Fixed order: deterministic.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_or_zero<0xb1>(v);       // quad_perm [1, 0, 3, 2]
  v += dpp_or_zero<0x4e>(v);       // quad_perm [2, 3, 0, 1]
  v += dpp_or_zero<0x141>(v);      // row_half_mirror
  v += dpp_or_zero<0x140>(v);      // row_mirror
  v += dpp_or_zero<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_or_zero<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

struct NullScatter {
  __device__ void operator()(int, int, int, f3) const {}
};

inline int grid_wgs(const ddm_batch *b) {
  int w = (b->n_blocks + (kWG / 64) - 1) / (kWG / 64);
  return w < 4096 ? (w > 0 ? w : 1) : 4096;
}

// ---------------------------------------------------------------- deterministic mode
// Sorted reductions instead of atomics (options bit DDM_DETERMINISTIC): particles are
// binned by stencil-base node with a counting sort (ranks within a node sorted back into
// particle-index order), and each grid node sums its contributions in a fixed order --
// the 27 stencil offsets in (i, j, k) order, then the particles whose base is the node
// minus that offset, in ascending particle index. The block list is rebuilt from the
// flags in block order, which fixes the order of the body-pose reductions too.
__device__ __forceinline__ bool deterministic(const ddm_batch &B) { return (B.options & DDM_DETERMINISTIC) != 0; }

// B-spline weight of offset i for the fractional position f (stencil()'s expressions)
template <int I>
__device__ __forceinline__ float bspline_w(float f) {
  if (I == 0) {
    const float a = 1.5f - f;
    return 0.5f * (a * a);
  } else if (I == 1) {
    const float a = f - 1.f;
    return 0.75f - a * a;
  } else {
    const float a = f - 0.5f;
    return 0.5f * (a * a);
  }
}

// flag-only marking (the block list is rebuilt from the flags): the flag holds scene + 1
__device__ void flag_tile_blocks_det(const ddm_batch &B, const ddm_scene &S, int si, const Tile &T) {
  int b0[3], nb[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b0[d] = T.lo[d] >> 2;
    nb[d] = ((T.lo[d] + T.t[d] - 1) >> 2) - b0[d] + 1;
  }
  const int tot = nb[0] * nb[1] * nb[2];
  for (int i = threadIdx.x; i < tot; i += kWG) {
    const int bx = i / (nb[1] * nb[2]), by = (i / nb[2]) % nb[1], bz = i % nb[2];
    B.block_flags[block_of(S, b0[0] + bx, b0[1] + by, b0[2] + bz)] = (uint32_t)(si + 1);
  }
}

// Sorted gather of one listed block by one workgroup. The (count, start) of the 6x6x6
// stencil-base nodes that can reach the block are staged in LDS; wave w of the workgroup
// sums, for every node of the block (lane = node), the stencil offsets with index
// = w (mod 4) in (i, j, k) order, each over the offset's base node's particles in sorted
// order (their records are contiguous); the four partial sums are then added in wave
// order. contrib.add<I, J, K>(record, acc) adds one particle whose stencil base is the
// node minus (I, J, K).
struct DetGatherLDS {
  int cnt[216], start[216];
  float part[kWG / 64][64][4];
};
template <int I, int J, int K, class Contrib>
__device__ __forceinline__ void det_gather_offset(const DetGatherLDS &L, const ddm_det &D, int lx, int ly, int lz,
                                                  Contrib &contrib, float *acc) {
  const int c = ((lx - I + 2) * 6 + (ly - J + 2)) * 6 + (lz - K + 2);
  const int n = L.cnt[c];
  if (n == 0) return;
  const float4 *r = reinterpret_cast<const float4 *>(D.sorted) + 4 * (size_t)L.start[c];
  const float4 *const re = r + 4 * (size_t)n;
  for (; r + 4 < re; r += 8) {  // two records per iteration (loads overlap), summed in order
    const float4 a0 = r[0], b0 = r[1], c0 = r[2], d0 = r[3], a1 = r[4], b1 = r[5], c1 = r[6], d1 = r[7];
    contrib.template add<I, J, K>(a0, b0, c0, d0, acc);
    contrib.template add<I, J, K>(a1, b1, c1, d1, acc);
  }
  if (r < re) contrib.template add<I, J, K>(r[0], r[1], r[2], r[3], acc);
}
// returns, in the lanes of wave 0, the block's node sums (NC components); call with the
// whole workgroup (it synchronizes)
template <int NC, class Contrib>
__device__ void det_gather_block(const ddm_scene &S, const ddm_det &D, int blk, DetGatherLDS &L, Contrib &contrib,
                                 float *out) {
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int lb = blk - S.blk_begin, nby = S.dim[1] >> 2, nbz = S.dim[2] >> 2;
  const int bx = 4 * (lb / (nby * nbz)), by = 4 * ((lb / nbz) % nby), bz = 4 * (lb % nbz);
  if (t < 216) {
    const int x = bx - 2 + t / 36, y = by - 2 + (t / 6) % 6, z = bz - 2 + t % 6;
    int n = 0, q = 0;
    if (x >= 0 && y >= 0 && z >= 0 && x < S.dim[0] && y < S.dim[1] && z < S.dim[2]) {
      const int c = node_index(S, x, y, z);
      n = D.cell_count[c];
      if (n) q = D.cell_start[c];
    }
    L.cnt[t] = n;
    L.start[t] = q;
  }
  __syncthreads();
  const int lx = lane >> 4, ly = (lane >> 2) & 3, lz = lane & 3;
  float acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.f;
#define DDM_DET_OFF(n, i, j, k) \
  if (wave == (n) % 4) det_gather_offset<i, j, k>(L, D, lx, ly, lz, contrib, acc);
  DDM_DET_OFF(0, 0, 0, 0) DDM_DET_OFF(1, 0, 0, 1) DDM_DET_OFF(2, 0, 0, 2)
  DDM_DET_OFF(3, 0, 1, 0) DDM_DET_OFF(4, 0, 1, 1) DDM_DET_OFF(5, 0, 1, 2)
  DDM_DET_OFF(6, 0, 2, 0) DDM_DET_OFF(7, 0, 2, 1) DDM_DET_OFF(8, 0, 2, 2)
  DDM_DET_OFF(9, 1, 0, 0) DDM_DET_OFF(10, 1, 0, 1) DDM_DET_OFF(11, 1, 0, 2)
  DDM_DET_OFF(12, 1, 1, 0) DDM_DET_OFF(13, 1, 1, 1) DDM_DET_OFF(14, 1, 1, 2)
  DDM_DET_OFF(15, 1, 2, 0) DDM_DET_OFF(16, 1, 2, 1) DDM_DET_OFF(17, 1, 2, 2)
  DDM_DET_OFF(18, 2, 0, 0) DDM_DET_OFF(19, 2, 0, 1) DDM_DET_OFF(20, 2, 0, 2)
  DDM_DET_OFF(21, 2, 1, 0) DDM_DET_OFF(22, 2, 1, 1) DDM_DET_OFF(23, 2, 1, 2)
  DDM_DET_OFF(24, 2, 2, 0) DDM_DET_OFF(25, 2, 2, 1) DDM_DET_OFF(26, 2, 2, 2)
#undef DDM_DET_OFF
#pragma unroll
  for (int c = 0; c < NC; ++c) L.part[wave][lane][c] = acc[c];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = ((L.part[0][lane][c] + L.part[1][lane][c]) + L.part[2][lane][c]) + L.part[3][lane][c];
  }
  __syncthreads();  // LDS is reused by the next block
}

// live chunks / scenes of a launch (a captured launch of fixed size reads them on the device)
__device__ __forceinline__ int live_chunks(const ddm_batch &B) { return B.live ? B.live[0] : B.n_chunks; }
// Chunk of this workgroup, or -1 past the live chunks (uniform over the workgroup).
// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8); giving
// each XCD's workgroups one contiguous range of chunks keeps neighbouring chunks, which
// share grid nodes, on one XCD's L2 (P2G 0.42 -> 0.39 ms and G2P adjoint 0.33 -> 0.30 ms
// on 2 M particles in 64 scenes). A bijection of [0, live chunks).
__device__ __forceinline__ int chunk_index(const ddm_batch &B) {
  const int n = live_chunks(B), b = (int)blockIdx.x;
  if (b >= n) return -1;
  const int x = b & 7, k = b >> 3, q = n >> 3, r = n & 7;
  return x * q + min(x, r) + k;
}
__device__ __forceinline__ int live_scenes(const ddm_batch &B) { return B.live ? B.live[1] : B.n_scenes; }

// blocks a scene restores from its snapshot (0 without one, or when it overflowed)
__device__ __forceinline__ int restorable(const ddm_scene &S) {
  if (!S.tape_mmv) return 0;
  const int n = S.tape_status[0];
  return n <= S.tape_capacity ? n : 0;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Taped adjoint: for every scene with a snapshot that holds all its listed blocks, puts
// the blocks back into the arena and into the block list (scenes in order, each at the
// prefix sum of the restored counts before it), and clears tape_status[1]; a scene whose
// forward listed more blocks than its snapshot holds gets tape_status[1] = 1 and is
// recomputed by the kernels that follow (its arena blocks are untouched, i.e. clear; its
// recomputed blocks are appended after the restored ones), and marks each restored block
// with scene + 1 (outside the deterministic mode). Run by kRestoreWGs workgroups
// per scene: this is workgroup `part` of `parts` for scene si (k_restore's body; the taped
// adjoint's P2G adjoint can also run it, in extra workgroups, for the substep before).
// launches whose grid kernels visit every block's flag instead of the block list
// (non-deterministic, few chunks, few blocks)
inline bool scan_blocks(const ddm_batch *b) {
  return !(b->options & DDM_DETERMINISTIC) && b->n_chunks <= kLatencyChunks && b->n_blocks > 0 &&
         b->n_blocks <= kScanBlocks;
}
inline int scan_wgs(const ddm_batch *b) { return (b->n_blocks + kWG / 64 - 1) / (kWG / 64); }

constexpr int kRestoreWGs = 16;
__device__ void restore_scene_part(const ddm_batch &B, int si, int part, int parts) {
  const int n_live = live_scenes(B);
  if (si >= n_live) return;
  const ddm_scene &S = B.scenes[si];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // restored entries of the scenes before this one, and of all scenes
  int before = 0, total = 0;
  for (int s0 = 0; s0 < n_live; s0 += 64) {
    const int s2 = s0 + lane;
    const int c = s2 < n_live ? restorable(B.scenes[s2]) : 0;
    total += wave_sum_i(c);
    before += wave_sum_i(s2 < si ? c : 0);
  }
  if (si == 0 && part == 0 && threadIdx.x == 0) *B.block_count = total;
  if (!S.tape_mmv) return;
  const int n = S.tape_status[0];
  const bool use = n <= S.tape_capacity;
  if (part == 0 && threadIdx.x == 0) S.tape_status[1] = use ? 0 : 1;
  if (!use) return;
  int2 *list = reinterpret_cast<int2 *>(B.block_list) + before;
  for (int e = part * (kWG / 64) + wave; e < n; e += parts * (kWG / 64)) {
    const int blk = S.tape_list[e];
    const size_t gi = ((size_t)blk << 6) | (size_t)lane, ti = ((size_t)e << 6) | (size_t)lane;
    reinterpret_cast<float4 *>(B.grid_mmv)[gi] = reinterpret_cast<const float4 *>(S.tape_mmv)[ti];
    reinterpret_cast<float4 *>(B.grid_v)[gi] = reinterpret_cast<const float4 *>(S.tape_v)[ti];
    if (lane == 0) {
      list[e] = make_int2(blk, si);
      // the mark a scanning grid adjoint reads (it clears it, as the listing one does)
      if (!(B.options & DDM_DETERMINISTIC)) B.block_flags[blk] = (uint32_t)(si + 1);
    }
  }
}

}  // namespace

// Adjoint kernels (fused_adjoint.hip, compiled with FMA contraction: they feed only
// gradients, whose bar is 1e-3, while the forward kernels keep the operation order of the
// CPU restatement for 1e-5 state parity).
void launch_g2p_adj(const ddm_batch &b, const ddm_state *cur, const ddm_state *next, const ddm_state *cur_grad,
                    const ddm_state *next_grad, hipStream_t stream);
void launch_grid_adj(const ddm_batch &b, const ddm_state *cur, const ddm_state *next, const ddm_state *cur_grad,
                     const ddm_state *next_grad, hipStream_t stream);
void launch_p2g_adj(const ddm_batch &b, const ddm_state *cur, const ddm_state *cur_grad, const ddm_state *next_grad,
                    hipStream_t stream,
                    const ddm_scene *restore_prev = nullptr);
// deterministic mode: G2P adjoint (per-particle part, then the sorted node gather) and the
// grid adjoint with per-block body-pose partials reduced in block order
void launch_g2p_adj_det(const ddm_batch &b, const ddm_det &d, const ddm_state *cur, const ddm_state *next,
                        const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream);
void launch_grid_adj_det(const ddm_batch &b, const ddm_det &d, const ddm_state *cur, const ddm_state *next,
                         const ddm_state *cur_grad, const ddm_state *next_grad, hipStream_t stream);

}  // namespace ddm_fused
