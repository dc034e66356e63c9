// Device helpers shared by the tiled kernels (tiled_engine.hip) and the
// fused tile kernels (fused_engine.hip): tile-local union-find in LDS and the
// per-tile measurement partials.
#pragma once
#include "engine.h"

namespace {

constexpr int PARTS = 7;  // e, m, n_clusters, key1, key2, n_plus1, n_plus2

// Forest words in HBM: parent (low 32 bits) | wx:16 | wy:16 windings of the
// tree edge to the parent.
__device__ __forceinline__ uint64_t gpack(uint32_t p, int wx, int wy) {
  return (uint64_t)p | ((uint64_t)(uint16_t)(int16_t)wx << 32) |
         ((uint64_t)(uint16_t)(int16_t)wy << 48);
}
__device__ __forceinline__ uint32_t gpar(uint64_t v) { return (uint32_t)v; }
__device__ __forceinline__ int gwx(uint64_t v) { return (int)(int16_t)(uint16_t)(v >> 32); }
__device__ __forceinline__ int gwy(uint64_t v) { return (int)(int16_t)(uint16_t)(v >> 48); }

__device__ __forceinline__ uint64_t gload(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gstore(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Roots of a and b in the HBM forest while hooks are in flight, walked
// together (both chains' loads in flight at once): agent-scope loads so a
// retry sees other workgroups' hooks; path halving; the windings of the walked
// paths are added to (ax, ay) and (bx, by).
__device__ __forceinline__ void gfind2(uint64_t* uf, uint32_t& a, int& ax, int& ay, uint32_t& b,
                                       int& bx, int& by) {
  uint64_t va = gload(&uf[a]), vb = gload(&uf[b]);
  while (true) {
    const uint32_t pa = gpar(va), pb = gpar(vb);
    const bool da = pa == a, db = pb == b;
    if (da && db) return;
    const uint64_t wa = da ? va : gload(&uf[pa]);
    const uint64_t wb = db ? vb : gload(&uf[pb]);
    if (!da) {
      const uint32_t ga = gpar(wa);
      if (ga == pa) {
        ax += gwx(va);
        ay += gwy(va);
        a = pa;
        va = wa;
      } else {
        const int hx = gwx(va) + gwx(wa), hy = gwy(va) + gwy(wa);
        gstore(&uf[a], gpack(ga, hx, hy));
        ax += hx;
        ay += hy;
        a = ga;
        va = gload(&uf[ga]);
      }
    }
    if (!db) {
      const uint32_t gb = gpar(wb);
      if (gb == pb) {
        bx += gwx(vb);
        by += gwy(vb);
        b = pb;
        vb = wb;
      } else {
        const int hx = gwx(vb) + gwx(wb), hy = gwy(vb) + gwy(wb);
        gstore(&uf[b], gpack(gb, hx, hy));
        bx += hx;
        by += hy;
        b = gb;
        vb = gload(&uf[gb]);
      }
    }
  }
}

// Root in the settled forest (K3), then the start node is pointed straight at
// it.  Windings are no longer needed once K2 has decided wrapping, and a
// concurrent reader sees either the old ancestor chain or the root.
__device__ __forceinline__ uint32_t groot(uint64_t* gf, uint32_t x) {
  const uint32_t start = x;
  uint32_t p = gpar(gf[x]);
  while (p != x) {
    x = p;
    p = gpar(gf[x]);
  }
  if (x != start) gf[start] = gpack(x, 0, 0);
  return x;
}


// OR of the lanes' wrap bits (1: x, 2: y) into *flag, one atomic per
// wavefront; every lane that calls it works on the lattice of that flag.
template <typename P>
__device__ __forceinline__ void wave_or_wraps(P* flag, uint32_t w) {
  const uint64_t mx = __ballot((w & 1u) != 0u), my = __ballot((w & 2u) != 0u);
  if ((mx | my) != 0ull && __lane_id() == (unsigned)(__ffsll((long long)__ballot(true)) - 1))
    atomicOr(flag, (mx ? 1u : 0u) | (my ? 2u : 0u));
}

__device__ __forceinline__ bool lattice_active(const LatticeSet& s, int lat) {
  return s.sweep_ctr[lat] < lat_total(s, lat);
}

// Per-tile partial of one sweep's measurement: E, M, n_clusters and the two
// largest clusters (key = size << 32 | ~root, ties to the smaller root) with
// their n_plus carried along, so the partial is final even after the next
// sweep has reused the cluster counters.
struct TileMeas {
  int32_t e, m, ncl;  // one tile: |e| <= 2 T^2
  uint64_t k1, k2;
  uint32_t p1, p2;
};

__device__ __forceinline__ void top2_insert(TileMeas& a, uint64_t k, uint32_t p) {
  if (k > a.k1) {
    a.k2 = a.k1;
    a.p2 = a.p1;
    a.k1 = k;
    a.p1 = p;
  } else if (k > a.k2) {
    a.k2 = k;
    a.p2 = p;
  }
}

// Top-2 of two key pairs (0 = none), keys distinct or equal.
template <typename K>
__device__ __forceinline__ void top2_merge(K& k1, K& k2, K o1, K o2) {
  const K hi = max(k1, o1);
  const K lo = max(min(k1, o1), max(k2, o2));
  k1 = hi;
  k2 = lo;
}

// Data-parallel-primitive moves within a wavefront (gfx9 DPP): the value of
// another lane, 0 where that lane lies outside the pattern (and in the rows
// the row mask leaves out).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, true);
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint64_t dpp_key(uint64_t v) {
  return (uint64_t)dpp_u32<CTRL, ROWS>((uint32_t)v) |
         (uint64_t)dpp_u32<CTRL, ROWS>((uint32_t)(v >> 32)) << 32;
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp_key(uint32_t v) {
  return dpp_u32<CTRL, ROWS>(v);
}
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_SHR8 = 0x118, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// Top-2 keys of a wavefront (0 = none) into lane 63: shifts within each row of
// 16 lanes, then row 0 -> 1 and 2 -> 3, then row 1 -> 2, 3.
template <typename K>
__device__ __forceinline__ void wave_top2(K& k1, K& k2) {
  top2_merge(k1, k2, dpp_key<DPP_ROW_SHR1>(k1), dpp_key<DPP_ROW_SHR1>(k2));
  top2_merge(k1, k2, dpp_key<DPP_ROW_SHR2>(k1), dpp_key<DPP_ROW_SHR2>(k2));
  top2_merge(k1, k2, dpp_key<DPP_ROW_SHR4>(k1), dpp_key<DPP_ROW_SHR4>(k2));
  top2_merge(k1, k2, dpp_key<DPP_ROW_SHR8>(k1), dpp_key<DPP_ROW_SHR8>(k2));
  top2_merge(k1, k2, dpp_key<DPP_ROW_BCAST15, 0xA>(k1), dpp_key<DPP_ROW_BCAST15, 0xA>(k2));
  top2_merge(k1, k2, dpp_key<DPP_ROW_BCAST31, 0xC>(k1), dpp_key<DPP_ROW_BCAST31, 0xC>(k2));
}

// Sum over a wavefront into lane 63.
template <typename T>
__device__ __forceinline__ T wave_sum63(T v) {
  v += (T)dpp_u32<DPP_ROW_SHR1>((uint32_t)v);
  v += (T)dpp_u32<DPP_ROW_SHR2>((uint32_t)v);
  v += (T)dpp_u32<DPP_ROW_SHR4>((uint32_t)v);
  v += (T)dpp_u32<DPP_ROW_SHR8>((uint32_t)v);
  v += (T)dpp_u32<DPP_ROW_BCAST15, 0xA>((uint32_t)v);
  v += (T)dpp_u32<DPP_ROW_BCAST31, 0xC>((uint32_t)v);
  return v;
}

// Workgroup reduction of the lanes' partials in two halves, so the cluster
// part (read before the sweep overwrites the roots) and the spin part (summed
// while the sweep reads the spins) need not be live at the same time: each
// wave reduces into red[wave][.] and, after a barrier, thread 0 combines.
[[maybe_unused]] __device__ void wave_clusters(TileMeas a, int64_t (*red)[PARTS]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a.ncl += __shfl_xor(a.ncl, off);
    const uint64_t o1 = __shfl_xor(a.k1, off), o2 = __shfl_xor(a.k2, off);
    const uint32_t q1 = __shfl_xor(a.p1, off), q2 = __shfl_xor(a.p2, off);
    top2_insert(a, o1, q1);
    top2_insert(a, o2, q2);
  }
  if ((threadIdx.x & 63) == 0) {
    int64_t* r = red[threadIdx.x >> 6];
    r[2] = a.ncl;
    r[3] = (int64_t)a.k1;
    r[4] = (int64_t)a.k2;
    r[5] = a.p1;
    r[6] = a.p2;
  }
}

[[maybe_unused]] __device__ void wave_spins(int32_t e, int32_t m, int64_t (*red)[PARTS]) {
  e = wave_sum63(e);
  m = wave_sum63(m);
  if ((threadIdx.x & 63) == 63) {
    red[threadIdx.x >> 6][0] = e;
    red[threadIdx.x >> 6][1] = m;
  }
}

// Start of the in-row run holding a lane's even site.  Lane j's pair is
// (chunk + 2j, chunk + 2j + 1); thru: lanes whose pair is linked inside and to
// the next pair, ol: lanes whose odd site links to the next pair.  The run
// starts after the nearest lane below that does not pass the run through.
__device__ __forceinline__ uint32_t run_start(uint64_t thru, uint64_t ol, uint64_t below,
                                              uint32_t chunk) {
  const uint64_t stop = ~thru & below;
  if (stop == 0ull) return chunk;
  const uint32_t j = 63u - (uint32_t)__clzll(stop);
  return chunk + 2u * j + (((ol >> j) & 1ull) ? 1u : 2u);
}

// Tile-local union-find in LDS over tile indices (no windings: a tile never
// crosses the periodic boundary).  Two finds walked together with path
// halving; the larger root hooks under the smaller, so every root is the
// minimum index (= minimum site) of its tile-local cluster.
//
// A node is named by a handle: its index, or (BA) its byte offset 4 * index
// from the forest's base, which then also holds parents as byte offsets.  With
// the forest in static LDS a byte-offset walk needs no address arithmetic: the
// parent read from one word is the address of the next, and the array's base
// folds into the load's offset field.  Handles order like indices.
template <bool BA>
__device__ __forceinline__ uint32_t& lds_at(uint32_t* base, uint32_t h) {
  if constexpr (BA) return *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + h);
  else return base[h];
}
template <bool BA>
__device__ __forceinline__ uint32_t lds_get(const uint32_t* base, uint32_t h) {
  if constexpr (BA) return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + h);
  else return base[h];
}

// One step of a halving walk: x has the parent px and the grandparent g.  A
// non-root x is pointed at g and the walk moves there; at a root (px == x)
// the walk stays.  Only non-root words are written (hooks CAS roots only).
template <bool BA = false>
__device__ __forceinline__ void halve_step(uint32_t* par, uint32_t& x, uint32_t px, uint32_t g) {
  const bool halve = px != x && g != px;
  if (halve) lds_at<BA>(par, x) = g;
  x = halve ? g : px;
}

template <bool BA = false>
__device__ __forceinline__ void tfind2(uint32_t* par, uint32_t& a, uint32_t& b) {
  uint32_t pa = lds_get<BA>(par, a), pb = lds_get<BA>(par, b);
  while (pa != a || pb != b) {
    // a root re-reads its own word; a hook made meanwhile shows as a new parent
    const uint32_t ga = lds_get<BA>(par, pa), gb = lds_get<BA>(par, pb);
    halve_step<BA>(par, a, pa, ga);
    halve_step<BA>(par, b, pb, gb);
    pa = lds_get<BA>(par, a);
    pb = lds_get<BA>(par, b);
  }
}

template <bool BA = false>
__device__ __forceinline__ void tunion(uint32_t* par, uint32_t a, uint32_t b) {
  while (true) {
    tfind2<BA>(par, a, b);
    if (a == b) return;
    const uint32_t hi = max(a, b), lo = min(a, b);
    if (atomicCAS(&lds_at<BA>(par, hi), hi, lo) == hi) return;
    a = lo;
    b = hi;
  }
}

// Roots of x and y in the settled tile forest, walked together.
__device__ __forceinline__ void troot2(const uint32_t* par, uint32_t& x, uint32_t& y) {
  uint32_t px = par[x], py = par[y];
  while (px != x || py != y) {
    x = px;
    y = py;
    px = par[x];
    py = par[y];
  }
}

// Unions of one wavefront through a dense queue q (packed a | b << 16): every
// lane holds up to three requests (bit k of req: the pair (a_k, b_k) of
// handles), appended at the queue's length n (wavefront-uniform), which the
// call returns grown.  The queue is then worked 64 entries at a time, so a
// round has no idle lanes waiting on a neighbour's second request -- and,
// filled over several passes of a wavefront's pairs, fewer rounds than one
// queue per pass (the requests depend on the occupation bits only, never on
// the forest).  A pass adds at most 129 (two per lane, three for lane 63).
__device__ __forceinline__ uint32_t wave_queue_put(uint32_t* q, uint32_t n, uint32_t req,
                                                   const uint32_t (&a)[3], const uint32_t (&b)[3],
                                                   uint64_t below) {
  const uint32_t c = (uint32_t)__popc(req);
  const uint64_t c0 = __ballot(c & 1u), c1 = __ballot(c & 2u);
  uint32_t pos = n + (uint32_t)(__popcll(c0 & below) + 2 * __popcll(c1 & below));
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if ((req >> k) & 1u) q[pos++] = a[k] | (b[k] << 16);
  return n + (uint32_t)(__popcll(c0) + 2 * __popcll(c1));
}

// The unions of entries [0, n) of the wave's queue; afterwards the queue may
// be filled again.  PRIO: the wavefront works its queue at issue priority 1
// (its dependent LDS round trips go ahead of other wavefronts' arithmetic).
template <bool BA = false, bool PRIO = false>
__device__ __forceinline__ void wave_queue_run(uint32_t* par, const uint32_t* q, uint32_t n) {
  // the queue is written and read by the lanes of this wavefront only
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (PRIO && n != 0u) __builtin_amdgcn_s_setprio(1);
  for (uint32_t i = threadIdx.x & 63u; i < n; i += 64u) {
    const uint32_t e = q[i];
    tunion<BA>(par, e & 0xFFFFu, e >> 16);
  }
  if (PRIO && n != 0u) __builtin_amdgcn_s_setprio(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Most requests a pass of wave_queue_put adds.
constexpr uint32_t WAVE_QUEUE_PASS = 129u;

}  // namespace
