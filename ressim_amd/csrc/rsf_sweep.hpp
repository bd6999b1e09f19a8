// Natural-order ILU(0) sweeps in one launch each: the forward and backward solves of the
// reference's own ordering (rs/fim/linear/gmres_block_jacobi.rs:190-237 for the 3x3 block
// ILU0 of the CPR post-smoother, :634-658 for the scalar ILU0 of the pressure stage).
//
// In the natural order a cell depends on its x-, y- and z- neighbours (forward solve) or on
// its x+, y+ and z+ neighbours (backward solve): a hyperplane schedule needs nx+ny+nz-2
// dependent levels, one launch each. Here one launch does the whole solve:
//  * a workgroup owns a tile of ty x tz grid lines along x, one line per thread, and marches
//    along x; thread (a, b) at step t handles x = t - a - b (forward; the backward marches
//    back from the far corner). Its y and z neighbours inside the tile were handled by
//    threads (a-1, b) and (a, b-1) one step earlier and are exchanged through LDS; its x
//    neighbour is its own previous result, in registers.
//  * the first line of a tile needs the last line of the tile before it in y (or z). That
//    tile handled the same cell ty-1 (tz-1) steps later in its own march and wrote it to a
//    halo buffer with device-coherent 64-bit stores, each holding 32 bits of the value and a
//    32-bit tag naming the launch and the step. The consumer loads its halo entries a few
//    steps ahead, with the coefficients, and checks the tags when it uses them: an entry
//    whose tag is not yet the expected one is reloaded until it is. No counters, no fences:
//    a value and its tag become visible together.
//  * tiles are numbered so that every dependency points to a lower workgroup index (the
//    backward solve reverses the numbering): a workgroup waits only on workgroups
//    dispatched before it. Every wait is bounded; a wait that gives up raises a flag that
//    ends every other wait, and the host turns it into an error.
//  * tags grow by 65536 per launch, and a halo position is rewritten by every launch of its
//    system that reaches it, so an entry left by an earlier launch never carries the tag
//    the current launch expects.
// The coefficients are packed once per preconditioner build in march order ("sweep
// layout"): row t of a tile holds, for every thread of the tile, the planes of the cell that
// thread handles at forward step t (the backward step t reads row steps-1-t), so every
// coefficient load of a step is one contiguous read by the workgroup. The sweeps do the
// arithmetic of the level-by-level kernels (k_bilu_fwd_level / k_pilu_fwd_level and their
// backward forms) in the same order, so the results are the same numbers.
#pragma once
#include <algorithm>
#include <climits>
#include <type_traits>

#include "rsf_kernels.hpp"

namespace rsf {

struct SweepTiles {
  int ty, tz;    // lines of a tile in y and z; threads per workgroup = ty * tz
  int nty, ntz;  // tiles in y and z
  int steps;     // nx + ty + tz - 2: the march of one tile (below 65536: it is part of the tags)
};

// Tiles of the one-launch natural-order sweeps: ty x tz lines per workgroup, at least
// `lanes` threads where the grid allows, grown in the direction with more tiles left.
inline SweepTiles sweep_tiles(int nx, int ny, int nz, int lanes) {
  int ty = 1, tz = 1;
  while (ty * tz < lanes) {
    const int cy = (ny + ty - 1) / ty, cz = (nz + tz - 1) / tz;
    if (ty < ny && (cy >= cz || tz >= nz))
      ty *= 2;
    else if (tz < nz)
      tz *= 2;
    else
      break;
  }
  SweepTiles t;
  t.ty = ty;
  t.tz = tz;
  t.nty = (ny + ty - 1) / ty;
  t.ntz = (nz + tz - 1) / tz;
  t.steps = nx + ty + tz - 2;
  return t;
}

constexpr int kSweepMaxLanes = 256;
constexpr int kSweepMaxSteps = 65535;
constexpr unsigned kSweepSpinLimit = 1u << 21;  // polls (each with a short sleep) before a wait gives up

// Sweep layout: per (tile, row, thread) a contiguous record of kSweepRec(nc) values, so a
// thread loads its coefficients with 16-byte loads. The forward solve reads the first half
// (D^-1 and the lower blocks), the backward solve the second (D^-1 and the upper blocks);
// both halves start on a 16-byte boundary.
//   nc = 1 (8 values):  [d, l_x-, l_y-, l_z-, d, u_x+, u_y+, u_z+]
//   nc = 3 (64 values): [L_x- L_y- L_z- (27), pad, D^-1 (9), U_x+ U_y+ U_z+ (27)]
__host__ __device__ constexpr int sweep_rec(int nc) { return nc == 1 ? 8 : 64; }
__host__ __device__ constexpr int sweep_fwd_off(int nc, int part) {  // part 0: D^-1, 1..3: x, y, z blocks
  return nc == 1 ? part : (part == 0 ? 28 : (part - 1) * 9);
}
__host__ __device__ constexpr int sweep_bwd_off(int nc, int part) {
  return nc == 1 ? 4 + part : (part == 0 ? 28 : 37 + (part - 1) * 9);
}
// values each solve loads per thread and step (whole 16-byte vectors)
__host__ __device__ constexpr int sweep_fwd_len(int nc) { return nc == 1 ? 4 : 40; }
__host__ __device__ constexpr int sweep_bwd_first(int nc) { return nc == 1 ? 4 : 28; }
__host__ __device__ constexpr int sweep_bwd_len(int nc) { return nc == 1 ? 4 : 36; }

// Packs D^-1 (natural order) and the off-diagonal stencil blocks of M (slot-major stencil
// store: M[(s * nc*nc + e) * n + i]) into the sweep layout; padding entries are zero.
template <int NC, class PT>
__global__ void k_sweep_pack(Grid g, SweepTiles st, const PT* __restrict__ M, const PT* __restrict__ dinv,
                             PT* __restrict__ sl) {
  constexpr int E = NC * NC;
  constexpr int R = sweep_rec(NC);
  const int NL = st.ty * st.tz;
  const long long total = (long long)st.nty * st.ntz * st.steps * NL;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const int L = (int)(k % NL);
  const long long tr = k / NL;  // tile * steps + row
  const int row = (int)(tr % st.steps);
  const int T = (int)(tr / st.steps);
  const int a = L % st.ty, b = L / st.ty;
  const int y = (T % st.nty) * st.ty + a, z = (T / st.nty) * st.tz + b;
  const int x = row - a - b;
  const bool valid = y < g.ny && z < g.nz && x >= 0 && x < g.nx;
  const int n = g.n;
  const int i = valid ? x + g.nx * (y + g.ny * z) : 0;
  PT rec[R];
#pragma unroll
  for (int e = 0; e < R; ++e) rec[e] = PT(0);
  if (valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      rec[sweep_fwd_off(NC, 0) + e] = dinv[(size_t)e * n + i];
      rec[sweep_bwd_off(NC, 0) + e] = dinv[(size_t)e * n + i];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        rec[sweep_fwd_off(NC, 1 + q) + e] = M[(size_t)((1 + 2 * q) * E + e) * n + i];  // slots 1, 3, 5
        rec[sweep_bwd_off(NC, 1 + q) + e] = M[(size_t)((2 + 2 * q) * E + e) * n + i];  // slots 2, 4, 6
      }
  }
  PT* dst = sl + (size_t)k * R;
#pragma unroll
  for (int e = 0; e < R; ++e) dst[e] = rec[e];
}

// Vectors of the sweeps use the same march order (plane c of a vector holds entry
// (tile * steps + row) * lanes + thread at that index), so every access of a step is
// contiguous over the workgroup. One parallel pass gathers the right-hand side into it
// and one scatters the result back to the natural order.
template <int NC, class PT>
__global__ void k_sweep_gather(Grid g, SweepTiles st, const PT* __restrict__ v, PT* __restrict__ vs,
                               const double* __restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const int NL = st.ty * st.tz;
  const long long cells = (long long)st.nty * st.ntz * st.steps * NL;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cells) return;
  const int L = (int)(k % NL);
  const long long tr = k / NL;
  const int row = (int)(tr % st.steps);
  const int T = (int)(tr / st.steps);
  const int a = L % st.ty, b = L / st.ty;
  const int y = (T % st.nty) * st.ty + a, z = (T / st.nty) * st.tz + b;
  const int x = row - a - b;
  const bool valid = y < g.ny && z < g.nz && x >= 0 && x < g.nx;
  const int i = valid ? x + g.nx * (y + g.ny * z) : 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) vs[c * cells + k] = valid ? v[(size_t)c * g.n + i] : PT(0);
}

template <int NC, class PT>
__global__ void k_sweep_scatter(Grid g, SweepTiles st, const PT* __restrict__ vs, PT* __restrict__ v,
                                const double* __restrict__ skip) {
  if (skip && *skip != 0.0) return;
  const int NL = st.ty * st.tz;
  const long long cells = (long long)st.nty * st.ntz * st.steps * NL;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cells) return;
  const int L = (int)(k % NL);
  const long long tr = k / NL;
  const int row = (int)(tr % st.steps);
  const int T = (int)(tr / st.steps);
  const int a = L % st.ty, b = L / st.ty;
  const int y = (T % st.nty) * st.ty + a, z = (T / st.nty) * st.tz + b;
  const int x = row - a - b;
  if (y >= g.ny || z >= g.nz || x < 0 || x >= g.nx) return;
  const int i = x + g.nx * (y + g.ny * z);
#pragma unroll
  for (int c = 0; c < NC; ++c) v[(size_t)c * g.n + i] = vs[c * cells + k];
}

template <class PT>
struct SweepArgs {
  Grid g;
  SweepTiles st;
  const PT* sl;    // sweep-layout coefficients
  const PT* r;     // forward: right-hand side (sweep layout, nc planes of `cells`)
  PT* out;         // forward: result w; backward: w in, x out (sweep layout)
  size_t cells;    // entries of one plane of a sweep-layout vector: tiles x steps x lanes
  unsigned long long* halo_y;  // tagged edge-line values for the next tile in y: [word][tile][step][b]
  unsigned long long* halo_z;  // for the next tile in z: [word][tile][step][a]
  unsigned tag0;               // tag of step s is tag0 + s + 1
  int* err;                    // set when a wait gave up
  const double* skip;          // optional: the whole launch returns when *skip != 0
};

__device__ __forceinline__ void store_coherent(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long load_coherent(const unsigned long long* p) {
  return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int load_coherent(const int* p) {
  return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A value in tagged 64-bit words: one word per float, two (low, high half) per double.
template <class PT> struct Tagged;
template <> struct Tagged<float> {
  static constexpr int kWords = 1;
  __device__ static void pack(float v, unsigned tag, unsigned long long* w) {
    w[0] = ((unsigned long long)tag << 32) | __float_as_uint(v);
  }
  __device__ static bool ok(const unsigned long long* w, unsigned tag) { return (unsigned)(w[0] >> 32) == tag; }
  __device__ static float value(const unsigned long long* w) { return __uint_as_float((unsigned)w[0]); }
  template <int NW>
  __device__ static bool ok_all(const unsigned long long (&w)[NW], unsigned tag) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NW; ++k) ok = ok && (unsigned)(w[k] >> 32) == tag;
    return ok;
  }
};
template <> struct Tagged<double> {
  static constexpr int kWords = 2;
  __device__ static void pack(double v, unsigned tag, unsigned long long* w) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    w[0] = ((unsigned long long)tag << 32) | (b & 0xffffffffull);
    w[1] = ((unsigned long long)tag << 32) | (b >> 32);
  }
  __device__ static bool ok(const unsigned long long* w, unsigned tag) {
    return (unsigned)(w[0] >> 32) == tag && (unsigned)(w[1] >> 32) == tag;
  }
  __device__ static double value(const unsigned long long* w) {
    return __longlong_as_double((long long)((w[1] << 32) | (w[0] & 0xffffffffull)));
  }
  template <int NW>
  __device__ static bool ok_all(const unsigned long long (&w)[NW], unsigned tag) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NW; ++k) ok = ok && (unsigned)(w[k] >> 32) == tag;
    return ok;
  }
};

// LDS-only barrier: the prefetched loads stay in flight across it
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Reloads the tagged words of one halo value until they carry `tag` (bounded; on giving up,
// or when another wait gave up, the flag is set and the words are used as they are).
// (Inlined: arrays passed to an outlined function would live in scratch memory, and a
// scratch access waits for every load issued before it, the prefetches included.)
template <int NW>
__device__ __forceinline__ void halo_wait(const unsigned long long* src, size_t plane, unsigned tag,
                                          unsigned long long (&w)[NW], int* err) {
  unsigned spins = 0;
#pragma nounroll
  while (true) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NW; ++k) ok = ok && (unsigned)(w[k] >> 32) == tag;
    if (ok) return;
    if (++spins > kSweepSpinLimit || load_coherent(err) != 0) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = load_coherent(src + k * plane);
  }
}

// 16-byte vector of PT
template <class PT> struct Vec16;
template <> struct Vec16<float> { using T = float4; static constexpr int kN = 4; };
template <> struct Vec16<double> { using T = double2; static constexpr int kN = 2; };

// Element i of a (uniform base, 32-bit byte offset per lane): the load or store takes the
// base in scalar registers and only the lane's offset in a vector register.
template <class T>
__device__ __forceinline__ const T& sw_at(const T* base, unsigned byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ T& sw_at(T* base, unsigned byte_off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off);
}

// One forward (BWD false) or backward (BWD true) solve of the natural-order DILU factor.
// NC = 1: pressure ILU0 (k_pilu_fwd_level / k_pilu_bwd_level); NC = 3: block ILU0
// (k_bilu_fwd_level / k_bilu_bwd_level). DEPTH steps of coefficients and right-hand sides
// and HD steps of halo entries are in flight ahead of the step that uses them. A halo entry
// can only be loaded once the tile before has published it, so every step of halo prefetch
// adds a step to the lag between consecutive tiles; HD is kept to about one load latency.
//
// The march runs a whole number of DEPTH-step groups: steps past the last one (t >= steps)
// handle no cell (x is outside the grid for every thread). With no exit inside a group,
// the loop head is reached only from the previous group's last step, where the slot of
// step t0 is the oldest load in flight, so the compiler's wait for it leaves the loads of
// the following DEPTH - 1 steps in flight (an exit inside the group joins the loop head on
// a path where that slot is the newest load, and every group then started with a wait for
// all loads). Addresses are a uniform base plus a 32-bit lane offset; threads that do not
// take a halo value, and cells outside the grid, read clamped in-range entries.
template <int NC, class PT, bool BWD, int DEPTH, int HD>
__global__ __launch_bounds__(kSweepMaxLanes) void k_ilu_sweep(SweepArgs<PT> A) {
  static_assert(DEPTH % HD == 0, "halo slots repeat within a group of DEPTH steps");
  if (A.skip && *A.skip != 0.0) return;
  constexpr int R = sweep_rec(NC);
  using V16 = typename Vec16<PT>::T;
  constexpr int VN = Vec16<PT>::kN;
  constexpr int FIRST = BWD ? sweep_bwd_first(NC) : 0;   // first value of this solve's half
  constexpr int LEN = BWD ? sweep_bwd_len(NC) : sweep_fwd_len(NC);
  constexpr int NV = LEN / VN;                            // 16-byte loads per step
  constexpr int OD = (BWD ? sweep_bwd_off(NC, 0) : sweep_fwd_off(NC, 0)) - FIRST;
  constexpr int TW = Tagged<PT>::kWords;
  constexpr int NW = NC * TW;  // tagged words of one halo value
  static_assert(LEN % VN == 0 && FIRST % VN == 0, "16-byte aligned halves");
  using U64 = unsigned long long;
  const Grid g = A.g;
  const SweepTiles st = A.st;
  const int nT = st.nty * st.ntz;
  const int T = BWD ? nT - 1 - (int)blockIdx.x : (int)blockIdx.x;
  const int tY = T % st.nty, tZ = T / st.nty;
  const int NL = st.ty * st.tz;
  const int L = threadIdx.x;
  const int a = L % st.ty, b = L / st.ty;
  const int y = tY * st.ty + a, z = tZ * st.tz + b;
  const bool line = y < g.ny && z < g.nz;
  const int S = st.steps;
  const int Tuy = BWD ? T + 1 : T - 1;  // the tile before in y (read only where one exists)
  const int Tuz = BWD ? T + st.nty : T - st.nty;
  const bool hy_line = BWD ? y + 1 < g.ny : y > 0;
  const bool hz_line = BWD ? z + 1 < g.nz : z > 0;
  // threads on the tile's reading edges take the neighbour from the halo, the others from LDS
  const bool in_edge_y = hy_line && (BWD ? a == st.ty - 1 : a == 0);
  const bool in_edge_z = hz_line && (BWD ? b == st.tz - 1 : b == 0);
  const bool out_edge_y = BWD ? (a == 0 && tY > 0) : (a == st.ty - 1 && tY + 1 < st.nty);
  const bool out_edge_z = BWD ? (b == 0 && tZ > 0) : (b == st.tz - 1 && tZ + 1 < st.ntz);
  // LDS lanes of the y and z neighbours (clamped: an edge thread's own lane is in range)
  const int lane_y = min(max(BWD ? L + 1 : L - 1, 0), NL - 1);
  const int lane_z = min(max(BWD ? L + st.ty : L - st.ty, 0), NL - 1);
  const size_t hy_plane = (size_t)nT * S * st.tz, hz_plane = (size_t)nT * S * st.ty;
  const int hrows = nT * S;  // rows of a halo plane
  const PT* vsrc = BWD ? A.out : A.r;
  // lane offsets (bytes)
  const unsigned off_c = (unsigned)(L * R + FIRST) * (unsigned)sizeof(PT);
  const unsigned off_v = (unsigned)L * (unsigned)sizeof(PT);
  const unsigned off_hy = (unsigned)b * 8u, off_hz = (unsigned)a * 8u;

  __shared__ PT xs[2][NC][kSweepMaxLanes];

  PT wprev[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) wprev[c] = PT(0);

  // rings: slot u holds step t with t % DEPTH == u. Every load is unconditional (a clamped
  // step or a clamped row where the value goes unused): a load under a branch has its
  // result copied at the branch's end, and the copy waits for every load in flight.
  V16 cf[DEPTH][NV];
  PT vv[DEPTH][NC];
  U64 hyr[HD][NW], hzr[HD][NW];
  auto issue = [&](int t, V16 (&c)[NV], PT (&v)[NC]) {
    t = min(t, S - 1);
    const int row = BWD ? S - 1 - t : t;
    const size_t e0 = (size_t)(T * S + row) * NL;  // uniform: the step's first entry
    const PT* cb = A.sl + e0 * R;
#pragma unroll
    for (int k = 0; k < NV; ++k) c[k] = sw_at((const V16*)cb, off_c + 16u * k);
#pragma unroll
    for (int q = 0; q < NC; ++q) v[q] = sw_at(vsrc + q * A.cells + e0, off_v);
  };
  auto issue_halo = [&](int t, U64 (&hy)[NW], U64 (&hz)[NW]) {
    t = min(t, S - 1);
    // halo entries of this step: written at step t + ty - 1 (t + tz - 1) of the tile before
    const int ry = min(max(Tuy * S + t + st.ty - 1, 0), hrows - 1);
    const int rz = min(max(Tuz * S + t + st.tz - 1, 0), hrows - 1);
    const U64* hyb = A.halo_y + (size_t)ry * st.tz;
    const U64* hzb = A.halo_z + (size_t)rz * st.ty;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      hy[k] = __hip_atomic_load(const_cast<U64*>(&sw_at(hyb + k * hy_plane, off_hy)), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
      hz[k] = __hip_atomic_load(const_cast<U64*>(&sw_at(hzb + k * hz_plane, off_hz)), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
    }
  };
#pragma unroll
  for (int u = 0; u < DEPTH; ++u) issue(u, cf[u], vv[u]);
#pragma unroll
  for (int u = 0; u < HD; ++u) issue_halo(u, hyr[u], hzr[u]);

  const int S_pad = (S + DEPTH - 1) / DEPTH * DEPTH;
  for (int t0 = 0; t0 < S_pad; t0 += DEPTH) {
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
      const int t = t0 + u;
      const int row = BWD ? S - 1 - t : t;
      const int x = row - a - b;
      const bool valid = line && x >= 0 && x < g.nx;
      const int cur = t & 1, prv = cur ^ 1;
      PT cc[LEN];
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const PT* p = (const PT*)&cf[u][k];
#pragma unroll
        for (int j = 0; j < VN; ++j) cc[k * VN + j] = p[j];
      }
      const PT* Dv = cc + OD;
      // neighbour values: own previous result (x), LDS or halo (y, z)
      const bool hx = BWD ? x + 1 < g.nx : x > 0;
      PT nb[3][NC];
      // the LDS neighbours first: their latency overlaps the halo checks
      PT ly[NC], lz[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        ly[q] = xs[prv][q][lane_y];
        lz[q] = xs[prv][q][lane_z];
      }
      U64 hw[NW], zw[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        hw[k] = hyr[u % HD][k];
        zw[k] = hzr[u % HD][k];
      }
      const unsigned tag_y = A.tag0 + (unsigned)(t + st.ty), tag_z = A.tag0 + (unsigned)(t + st.tz);
      // a halo value not yet published when its load was issued is reloaded (rare)
      if (valid && in_edge_y && !Tagged<PT>::ok_all(hw, tag_y)) {
        const U64* h = A.halo_y + (size_t)(Tuy * S + t + st.ty - 1) * st.tz + b;
        halo_wait<NW>(h, hy_plane, tag_y, hw, A.err);
      }
      if (valid && in_edge_z && !Tagged<PT>::ok_all(zw, tag_z)) {
        const U64* h = A.halo_z + (size_t)(Tuz * S + t + st.tz - 1) * st.ty + a;
        halo_wait<NW>(h, hz_plane, tag_z, zw, A.err);
      }
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        nb[0][q] = wprev[q];
        const PT hyv = Tagged<PT>::value(hw + q * TW), hzv = Tagged<PT>::value(zw + q * TW);
        nb[1][q] = hy_line ? (in_edge_y ? hyv : ly[q]) : PT(0);
        nb[2][q] = hz_line ? (in_edge_z ? hzv : lz[q]) : PT(0);
      }
      const bool has[3] = {hx, hy_line, hz_line};
      const PT* vr = vv[u];
      PT w[NC];
      if (!BWD) {
        // w = D^-1 (r - sum over x-, y-, z- of L_s w_s), in slot order 1, 3, 5
        double v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = double(vr[c]);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          if (!has[s]) continue;
          const PT* Ls = cc + sweep_fwd_off(NC, 1 + s) - FIRST;
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int q = 0; q < NC; ++q) v[c] -= double(Ls[c * NC + q]) * double(nb[s][q]);
        }
        if (NC == 1) {
          w[0] = PT(v[0] * double(Dv[0]));
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < NC; ++q) acc += double(Dv[c * NC + q]) * v[q];
            w[c] = PT(acc);
          }
        }
      } else {
        // x = w - D^-1 sum over x+, y+, z+ of U_s x_s, in slot order 2, 4, 6
        double v[NC];
        bool any = false;
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = 0.0;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          if (!has[s]) continue;
          any = true;
          const PT* Us = cc + sweep_bwd_off(NC, 1 + s) - FIRST;
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int q = 0; q < NC; ++q) v[c] += double(Us[c * NC + q]) * double(nb[s][q]);
        }
        if (NC == 1) {
          w[0] = any ? PT(double(vr[0]) - double(Dv[0]) * v[0]) : vr[0];
        } else {
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < NC; ++q) acc += double(Dv[c * NC + q]) * v[q];
            w[c] = any ? PT(double(vr[c]) - acc) : vr[c];
          }
        }
      }
      // a step outside the grid leaves 0 (never read by a cell inside it)
#pragma unroll
      for (int c = 0; c < NC; ++c) w[c] = valid ? w[c] : PT(0);
      // the next step's LDS neighbours first, then the global and halo stores
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        xs[cur][c][L] = w[c];
        wprev[c] = w[c];
      }
      {
        const size_t e0 = (size_t)(T * S + row) * NL;
        if (valid) {
#pragma unroll
          for (int c = 0; c < NC; ++c) sw_at(A.out + c * A.cells + e0, off_v) = w[c];
        }
      }
      const unsigned tag = A.tag0 + (unsigned)(t + 1);
      if (valid && out_edge_y) {
        U64* h = A.halo_y + (size_t)(T * S + t) * st.tz + b;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          U64 pw[TW];
          Tagged<PT>::pack(w[c], tag, pw);
#pragma unroll
          for (int k = 0; k < TW; ++k) store_coherent(h + (c * TW + k) * hy_plane, pw[k]);
        }
      }
      if (valid && out_edge_z) {
        U64* h = A.halo_z + (size_t)(T * S + t) * st.ty + a;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          U64 pw[TW];
          Tagged<PT>::pack(w[c], tag, pw);
#pragma unroll
          for (int k = 0; k < TW; ++k) store_coherent(h + (c * TW + k) * hz_plane, pw[k]);
        }
      }
      // this step's slots are free: load step t + DEPTH (t + HD for the halo) into them
      issue(t + DEPTH, cf[u], vv[u]);
      issue_halo(t + HD, hyr[u % HD], hzr[u % HD]);
      lds_barrier();
    }
  }
}

template <class FT>
struct FactorSweepArgs {
  Grid g;
  SweepTiles st;
  const double* A;  // NC = 3: the Jacobian's stencil blocks, A[(s * 9 + e) * n + i]
  const FT* P;      // NC = 1: the pressure matrix, P[s * n + i]
  FT* dinv;         // D^-1 per cell, nc * nc planes of n (natural order)
  unsigned long long* halo_y;  // tagged D^-1 of the edge lines for the next tile in y
  unsigned long long* halo_z;  // for the next tile in z
  unsigned tag0;
  int* err;
};

// The natural-order DILU factorization in one launch (factorize_block_ilu0,
// rs/fim/linear/gmres_block_jacobi.rs:239-337, for NC = 3; factorize_pressure_ilu0,
// :746-808, for NC = 1): the pivot of a cell needs the inverted pivots of its x-, y- and z-
// neighbours, the dependency pattern of the forward solve, so the march, the tiles and the
// tagged halo are those of k_ilu_sweep (forward); what travels between cells is D^-1 (nc x
// nc values). A thread loads its cell's operands in the natural order (its own diagonal
// and lower blocks, each lower neighbour's block towards it): over the march a thread's
// addresses advance by one cell per step. The arithmetic is the per-level kernels'
// (bilu_factor_sub / pilu_factor_sub), so the factors are the same numbers.
template <int NC, class FT, int DEPTH, int HD>
__global__ __launch_bounds__(kSweepMaxLanes) void k_ilu_factor_sweep(FactorSweepArgs<FT> A) {
  static_assert(DEPTH % HD == 0, "halo slots repeat within a group of DEPTH steps");
  constexpr int E = NC * NC;  // values of D^-1
  using OT = typename std::conditional<NC == 3, double, FT>::type;
  constexpr int NOP = NC == 3 ? 63 : 7;  // D, the x-, y-, z- blocks, the neighbours' blocks towards the cell
  constexpr int TW = Tagged<FT>::kWords;
  constexpr int NW = E * TW;
  using U64 = unsigned long long;
  const Grid g = A.g;
  const SweepTiles st = A.st;
  const int n = g.n, nxy = g.nx * g.ny;
  const int nT = st.nty * st.ntz;
  const int T = (int)blockIdx.x;
  const int tY = T % st.nty, tZ = T / st.nty;
  const int NL = st.ty * st.tz;
  const int L = threadIdx.x;
  const int a = L % st.ty, b = L / st.ty;
  const int y = tY * st.ty + a, z = tZ * st.tz + b;
  const bool line = y < g.ny && z < g.nz;
  const int S = st.steps;
  const int Tuy = T - 1, Tuz = T - st.nty;
  const bool hy_line = y > 0, hz_line = z > 0;
  const bool in_edge_y = hy_line && a == 0;
  const bool in_edge_z = hz_line && b == 0;
  const bool out_edge_y = a == st.ty - 1 && tY + 1 < st.nty;
  const bool out_edge_z = b == st.tz - 1 && tZ + 1 < st.ntz;
  const int lane_y = max(L - 1, 0), lane_z = max(L - st.ty, 0);
  const size_t hy_plane = (size_t)nT * S * st.tz, hz_plane = (size_t)nT * S * st.ty;
  const int hrows = nT * S;
  const unsigned off_hy = (unsigned)b * 8u, off_hz = (unsigned)a * 8u;
  const int nat0 = g.nx * (y + g.ny * z) - a - b;  // natural index at march row `row`: nat0 + row

  __shared__ FT xs[2][E][kSweepMaxLanes];
  FT wprev[E];
#pragma unroll
  for (int e = 0; e < E; ++e) wprev[e] = FT(0);

  OT op[DEPTH][NOP];
  U64 hyr[HD][NW], hzr[HD][NW];
  auto issue = [&](int t, OT (&o)[NOP]) {
    t = min(t, S - 1);
    const int i = min(max(nat0 + t, 0), n - 1);
    const int jn[3] = {max(i - 1, 0), max(i - g.nx, 0), max(i - nxy, 0)};
    const unsigned oi = (unsigned)i * (unsigned)sizeof(OT);
    if constexpr (NC == 3) {
#pragma unroll
      for (int e = 0; e < 9; ++e) o[e] = sw_at(A.A + (size_t)e * n, oi);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int sl = 1 + 2 * q;
        const unsigned oj = (unsigned)jn[q] * (unsigned)sizeof(OT);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
          o[9 + 9 * q + e] = sw_at(A.A + (size_t)(sl * 9 + e) * n, oi);
          o[36 + 9 * q + e] = sw_at(A.A + (size_t)(opposite(sl) * 9 + e) * n, oj);
        }
      }
    } else {
      o[0] = sw_at(A.P, oi);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int sl = 1 + 2 * q;
        o[1 + q] = sw_at(A.P + (size_t)sl * n, oi);
        o[4 + q] = sw_at(A.P + (size_t)opposite(sl) * n, (unsigned)jn[q] * (unsigned)sizeof(OT));
      }
    }
  };
  auto issue_halo = [&](int t, U64 (&hy)[NW], U64 (&hz)[NW]) {
    t = min(t, S - 1);
    const int ry = min(max(Tuy * S + t + st.ty - 1, 0), hrows - 1);
    const int rz = min(max(Tuz * S + t + st.tz - 1, 0), hrows - 1);
    const U64* hyb = A.halo_y + (size_t)ry * st.tz;
    const U64* hzb = A.halo_z + (size_t)rz * st.ty;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      hy[k] = __hip_atomic_load(const_cast<U64*>(&sw_at(hyb + k * hy_plane, off_hy)), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
      hz[k] = __hip_atomic_load(const_cast<U64*>(&sw_at(hzb + k * hz_plane, off_hz)), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
    }
  };
#pragma unroll
  for (int u = 0; u < DEPTH; ++u) issue(u, op[u]);
#pragma unroll
  for (int u = 0; u < HD; ++u) issue_halo(u, hyr[u], hzr[u]);

  const int S_pad = (S + DEPTH - 1) / DEPTH * DEPTH;
  for (int t0 = 0; t0 < S_pad; t0 += DEPTH) {
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
      const int t = t0 + u;
      const int x = t - a - b;
      const bool valid = line && x >= 0 && x < g.nx;
      const int cur = t & 1, prv = cur ^ 1;
      // the LDS neighbours first: their latency overlaps the halo checks
      FT ly[E], lz[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        ly[e] = xs[prv][e][lane_y];
        lz[e] = xs[prv][e][lane_z];
      }
      U64 hw[NW], zw[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        hw[k] = hyr[u % HD][k];
        zw[k] = hzr[u % HD][k];
      }
      const unsigned tag_y = A.tag0 + (unsigned)(t + st.ty), tag_z = A.tag0 + (unsigned)(t + st.tz);
      if (valid && in_edge_y && !Tagged<FT>::ok_all(hw, tag_y)) {
        const U64* h = A.halo_y + (size_t)(Tuy * S + t + st.ty - 1) * st.tz + b;
        halo_wait<NW>(h, hy_plane, tag_y, hw, A.err);
      }
      if (valid && in_edge_z && !Tagged<FT>::ok_all(zw, tag_z)) {
        const U64* h = A.halo_z + (size_t)(Tuz * S + t + st.tz - 1) * st.ty + a;
        halo_wait<NW>(h, hz_plane, tag_z, zw, A.err);
      }
      FT nb[3][E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        nb[0][e] = wprev[e];
        nb[1][e] = hy_line ? (in_edge_y ? Tagged<FT>::value(hw + e * TW) : ly[e]) : FT(0);
        nb[2][e] = hz_line ? (in_edge_z ? Tagged<FT>::value(zw + e * TW) : lz[e]) : FT(0);
      }
      const bool has[3] = {x > 0, hy_line, hz_line};
      const OT* o = op[u];
      FT w[E];
      if constexpr (NC == 3) {
        double D[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) D[e] = o[e];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (!has[q]) continue;
          double Lq[9], Dj[9], Uq[9];
#pragma unroll
          for (int e = 0; e < 9; ++e) {
            Lq[e] = o[9 + 9 * q + e];
            Dj[e] = double(nb[q][e]);
            Uq[e] = o[36 + 9 * q + e];
          }
          bilu_factor_sub(D, Lq, Dj, Uq);
        }
        double inv[9];
        inv3x3(D, inv);
#pragma unroll
        for (int e = 0; e < 9; ++e) w[e] = FT(inv[e]);
      } else {
        const double a0 = double(o[0]);
        double uu = a0;
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (has[q]) pilu_factor_sub(uu, double(o[1 + q]), double(nb[q][0]), double(o[4 + q]));
        w[0] = FT(pilu_factor_inv(a0, uu));
      }
      // a step outside the grid leaves 0 (never read by a cell inside it)
#pragma unroll
      for (int e = 0; e < E; ++e) w[e] = valid ? w[e] : FT(0);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        xs[cur][e][L] = w[e];
        wprev[e] = w[e];
      }
      if (valid) {
        const unsigned oi = (unsigned)(nat0 + t) * (unsigned)sizeof(FT);
#pragma unroll
        for (int e = 0; e < E; ++e) sw_at(A.dinv + (size_t)e * n, oi) = w[e];
      }
      const unsigned tag = A.tag0 + (unsigned)(t + 1);
      if (valid && out_edge_y) {
        U64* h = A.halo_y + (size_t)(T * S + t) * st.tz + b;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          U64 pw[TW];
          Tagged<FT>::pack(w[e], tag, pw);
#pragma unroll
          for (int k = 0; k < TW; ++k) store_coherent(h + (e * TW + k) * hy_plane, pw[k]);
        }
      }
      if (valid && out_edge_z) {
        U64* h = A.halo_z + (size_t)(T * S + t) * st.ty + a;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          U64 pw[TW];
          Tagged<FT>::pack(w[e], tag, pw);
#pragma unroll
          for (int k = 0; k < TW; ++k) store_coherent(h + (e * TW + k) * hz_plane, pw[k]);
        }
      }
      issue(t + DEPTH, op[u]);
      issue_halo(t + HD, hyr[u % HD], hzr[u % HD]);
      lds_barrier();
    }
  }
}

}  // namespace rsf
