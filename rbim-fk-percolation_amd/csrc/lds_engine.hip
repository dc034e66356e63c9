// LDS-resident Swendsen-Wang engine for lattices up to 128 x 128 on gfx950.
//
// One workgroup owns one lattice for a whole block of sweeps: spins, couplings,
// the union-find forest and the per-cluster counters all live in LDS, so a
// sweep touches HBM only through the moment accumulators at bin boundaries.
// Each lane owns a site pair (2q, 2q+1) -- the unit of one Philox4x32-10 call
// -- in every phase.  Per sweep g (src/rbim/swendsen_wang.py:103-172):
//
//   A  activation  Philox draws of the pair; satisfied bonds are occupied when
//                  draw < thr31 (clusters.py:48-79).  Occupied in-row bonds
//                  (y -> y+1 inside the 128-site chunk of a wavefront) are
//                  merged without atomics: two ballots give every site the
//                  start of its run, which becomes its parent (winding 0).
//                  The same pass measures sweep g-1 (below).
//   B  union       the remaining occupied bonds (all x bonds, y bonds that
//                  wrap or leave the chunk) by hook-and-compress on LDS CAS.
//                  The forest word packs parent (14 bits) with the windings
//                  (wx, wy; 9 bits each) of the tree path to the parent; the
//                  larger root always hooks under the smaller, so every root
//                  is its cluster's minimum site.  A bond that finds both ends
//                  under one root is a non-tree bond: its loop winding
//                  c + w(b) - w(a) decides wrap_x / wrap_y exactly
//                  (_unionfind.py:101-124).
//   C  flip        s ^= flipbit(root); cnt[root] += (1 << 16) | [s = +1]
//                  (swendsen_wang.py:140-141, clusters.py:98-118).
//   measurement of sweep g (fused into A of sweep g+1, or a closing pass):
//                  satisfied bonds, up spins and roots are counted with
//                  ballot + popcount; the two largest clusters (size desc,
//                  min site asc) by a 64-bit key max-reduction.  Lane 0 then
//                  folds the sweep into its bin's moment sums
//                  (simulation.py:212-236).
#include "engine.h"
#include "tile_common.h"

#ifndef RBIM_UNROLL_A
#define RBIM_UNROLL_A 1
#endif
#ifndef RBIM_UNROLL_B
#define RBIM_UNROLL_B 1
#endif

namespace {

constexpr uint32_t PMASK = 0x3FFFu;  // parent field

__device__ __forceinline__ uint32_t uf_pack(uint32_t p, int wx, int wy) {
  return p | (((uint32_t)wx & 0x1FFu) << 14) | ((uint32_t)wy << 23);
}
__device__ __forceinline__ int uf_wx(uint32_t v) { return ((int)(v << 9)) >> 23; }
__device__ __forceinline__ int uf_wy(uint32_t v) { return ((int)v) >> 23; }

// Forest nodes are handles (lds_at in tile_common.h): site indices, or (BA)
// byte offsets 4 n, which the 14-bit parent field holds while N <= 4096.

// Roots of x and y in the settled forest (phase C), walked together.
template <bool BA>
__device__ __forceinline__ void lds_root2(const uint32_t* uf, uint32_t& x, uint32_t& y) {
  uint32_t px = lds_get<BA>(uf, x) & PMASK, py = lds_get<BA>(uf, y) & PMASK;
  while (px != x || py != y) {
    x = px;
    y = py;
    px = lds_get<BA>(uf, x) & PMASK;
    py = lds_get<BA>(uf, y) & PMASK;
  }
}

// Roots of four sites at once (the two pairs of a lane when it owns exactly two).
template <bool BA>
__device__ __forceinline__ void lds_root4(const uint32_t* uf, uint32_t (&x)[4]) {
  uint32_t p[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) p[c] = lds_get<BA>(uf, x[c]) & PMASK;
  while (p[0] != x[0] || p[1] != x[1] || p[2] != x[2] || p[3] != x[3]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = p[c];
#pragma unroll
    for (int c = 0; c < 4; ++c) p[c] = lds_get<BA>(uf, x[c]) & PMASK;
  }
}

// Walk a and b to their current roots together (the two chains' loads are in
// flight at the same time), accumulating tree-path windings and halving both
// paths.  a / b may be stale roots from a failed hook: the windings carried so
// far stay valid because relative positions inside a tree never change.
template <bool BA>
__device__ __forceinline__ void find2(uint32_t* uf, uint32_t& a, int& ax, int& ay, uint32_t& b,
                                      int& bx, int& by) {
  uint32_t va = lds_get<BA>(uf, a), vb = lds_get<BA>(uf, b);
  while (true) {
    const uint32_t pa = va & PMASK, pb = vb & PMASK;
    const bool da = pa == a, db = pb == b;
    if (da && db) return;
    const uint32_t wa = da ? va : lds_get<BA>(uf, pa);
    const uint32_t wb = db ? vb : lds_get<BA>(uf, pb);
    if (!da) {
      const uint32_t ga = wa & PMASK;
      if (ga == pa) {
        ax += uf_wx(va);
        ay += uf_wy(va);
        a = pa;
        va = wa;
      } else {
        const int hx = uf_wx(va) + uf_wx(wa), hy = uf_wy(va) + uf_wy(wa);
        lds_at<BA>(uf, a) = uf_pack(ga, hx, hy);
        ax += hx;
        ay += hy;
        a = ga;
        va = lds_get<BA>(uf, ga);
      }
    }
    if (!db) {
      const uint32_t gb = wb & PMASK;
      if (gb == pb) {
        bx += uf_wx(vb);
        by += uf_wy(vb);
        b = pb;
        vb = wb;
      } else {
        const int hx = uf_wx(vb) + uf_wx(wb), hy = uf_wy(vb) + uf_wy(wb);
        lds_at<BA>(uf, b) = uf_pack(gb, hx, hy);
        bx += hx;
        by += hy;
        b = gb;
        vb = lds_get<BA>(uf, gb);
      }
    }
  }
}

// Union of the occupied bond a -> b that crosses the periodic boundary
// (cx, cy) times.  Returns wrap bits (1: x, 2: y) when the bond closes a loop.
template <bool BA>
__device__ __forceinline__ uint32_t lds_union(uint32_t* uf, uint32_t a, uint32_t b, int cx,
                                              int cy) {
  int ax = 0, ay = 0, bx = 0, by = 0;
  while (true) {
    find2<BA>(uf, a, ax, ay, b, bx, by);
    if (a == b) return ((cx + bx - ax) != 0 ? 1u : 0u) | ((cy + by - ay) != 0 ? 2u : 0u);
    if (a < b) {
      if (atomicCAS(&lds_at<BA>(uf, b), b, uf_pack(a, ax - bx - cx, ay - by - cy)) == b) return 0u;
    } else {
      if (atomicCAS(&lds_at<BA>(uf, a), a, uf_pack(b, bx - ax + cx, by - ay + cy)) == a) return 0u;
    }
  }
}

// Winding-free forest (every word is a bare parent): roots of a and b,
// walked together with path halving.
template <bool BA>
__device__ __forceinline__ void find2_plain(uint32_t* uf, uint32_t& a, uint32_t& b) {
  uint32_t pa = lds_get<BA>(uf, a), pb = lds_get<BA>(uf, b);
  while (pa != a || pb != b) {
    const uint32_t ga = pa != a ? lds_get<BA>(uf, pa) : pa;
    const uint32_t gb = pb != b ? lds_get<BA>(uf, pb) : pb;
    if (pa != a) {
      if (ga == pa) {
        a = pa;
      } else {
        lds_at<BA>(uf, a) = ga;
        a = ga;
        pa = lds_get<BA>(uf, ga);
      }
    }
    if (pb != b) {
      if (gb == pb) {
        b = pb;
      } else {
        lds_at<BA>(uf, b) = gb;
        b = gb;
        pb = lds_get<BA>(uf, gb);
      }
    }
  }
}

// Union of a bond that does not cross the periodic boundary, before any bond
// that does: every tree path then lies inside the open L x L square, so all
// windings are 0 and a loop closed here cannot wind.
template <bool BA>
__device__ __forceinline__ void lds_union_plain(uint32_t* uf, uint32_t a, uint32_t b) {
  while (true) {
    find2_plain<BA>(uf, a, b);
    if (a == b) return;
    const uint32_t hi = max(a, b), lo = min(a, b);
    if (atomicCAS(&lds_at<BA>(uf, hi), hi, lo) == hi) return;
    a = lo;
    b = hi;
  }
}

// Cluster keys from the counter c = size << 16 | n_plus of a root (its handle:
// index, or byte offset when BA).  The max key is the largest cluster, ties to
// the smaller root; n_plus rides along.
//   wide   (any L):   size (bits 32..), ~root (16..29), n_plus (0..14)
//   narrow (N <= 1024): size (21..31), 1023 - root index (11..20), n_plus (0..10)
template <bool NARROW, bool BA>
struct ClusterKey {
  using T = uint64_t;
  static __device__ __forceinline__ T make(uint32_t c, uint32_t root) {
    return ((uint64_t)(c >> 16) << 32) | ((uint64_t)(PMASK - root) << 16) | (uint64_t)(c & 0xFFFFu);
  }
  static __device__ __forceinline__ int64_t size(T k) { return (int64_t)(k >> 32); }
  static __device__ __forceinline__ int64_t n_plus(T k) { return (int64_t)(k & 0xFFFFu); }
};
template <bool BA>
struct ClusterKey<true, BA> {
  using T = uint32_t;
  static __device__ __forceinline__ T make(uint32_t c, uint32_t root) {
    // (4092 - 4 r) << 9 == (1023 - r) << 11
    return ((c >> 16) << 21) | (BA ? (4092u - root) << 9 : (1023u - root) << 11) | (c & 0x7FFu);
  }
  static __device__ __forceinline__ int64_t size(T k) { return (int64_t)(k >> 21); }
  static __device__ __forceinline__ int64_t n_plus(T k) { return (int64_t)(k & 0x7FFu); }
};


__device__ __forceinline__ uint32_t popc_ballot(bool pred) {
  return (uint32_t)__popcll(__ballot(pred));
}

// Geometry of the site pair (2q, 2q+1).  LC > 0 is a compile-time L.
struct Pair {
  uint32_t n0, n1, x0, y0, x1, y1, nr0, nu0, nr1, nu1;
  bool v0, v1;
};

// ALL: every lane of every iteration owns a full pair (known at compile time).
template <int LC, bool ALL>
__device__ __forceinline__ Pair make_pair(uint32_t q, uint32_t N, uint32_t Lr, uint32_t magic) {
  const uint32_t L = LC ? (uint32_t)LC : Lr;
  Pair p;
  p.n0 = 2u * q;
  p.n1 = p.n0 + 1u;
  p.v0 = ALL || p.n0 < N;
  p.v1 = ALL || p.n1 < N;
  p.x0 = LC ? p.n0 / (uint32_t)LC : site_x(p.n0, magic);
  p.y0 = p.n0 - p.x0 * L;
  if (LC != 0 && (LC & 1) == 0) {  // even L: the pair never straddles a row
    p.x1 = p.x0;
    p.y1 = p.y0 + 1u;
  } else {
    const bool edge = p.y0 == L - 1u;
    p.x1 = edge ? p.x0 + 1u : p.x0;
    p.y1 = edge ? 0u : p.y0 + 1u;
  }
  if constexpr (LC > 0 && (LC & (LC - 1)) == 0) {
    // a power of two: the periodic neighbours by masks (x + 1 wraps with n + L
    // mod N, y + 1 within the row's bits)
    constexpr uint32_t NM = (uint32_t)(LC * LC) - 1u, YM = (uint32_t)LC - 1u;
    p.nr0 = (p.n0 + L) & NM;
    p.nu0 = (p.n0 & ~YM) | ((p.n0 + 1u) & YM);
    p.nr1 = (p.n1 + L) & NM;
    p.nu1 = (p.n1 & ~YM) | ((p.n1 + 1u) & YM);
  } else {
    p.nr0 = (p.x0 == L - 1u) ? p.y0 : p.n0 + L;
    p.nu0 = (p.y0 == L - 1u) ? p.n0 - p.y0 : p.n0 + 1u;
    p.nr1 = (p.x1 == L - 1u) ? p.y1 : p.n1 + L;
    p.nu1 = (p.y1 == L - 1u) ? p.n1 - p.y1 : p.n1 + 1u;
  }
  return p;
}

// LDS words of one lattice of N sites: forest, counters, state bytes.
__host__ __device__ constexpr uint32_t lds_words(uint32_t N) { return 2u * N + ((N + 15u) & ~15u) / 4u; }
// One lattice per workgroup at a compile-time side up to 64: static LDS, byte
// offsets as forest handles (4 N - 4 fits the 14-bit parent field).
__host__ __device__ constexpr bool lds_static(int P, int LC) { return P == 1 && LC > 0 && LC <= 64; }

// The sweeps of P lattices in one workgroup, NT threads (whole wavefronts)
// each: the workgroup's `page` holds lattices page * P .. page * P + P - 1 of
// the launch's list, each in its own LDS region, every one run exactly as a
// workgroup of NT threads of its own would run it.  The workgroup's barriers
// are shared, so the P lattices sweep in step: they must share their sweep
// counter and schedule (true of every lattice of a batch).  A page past the
// end of the list keeps its idle lattices in step on neutral LDS (all spins
// up, no bond drawn) and writes nothing.
template <int NT, int LC, int P>
__device__ __forceinline__ void lds_sweeps(const LatticeSet& s, int32_t sweeps, int page) {
  constexpr int NW = NT / 64;
  constexpr bool ALL = LC > 0 && (LC % 2) == 0 && ((LC * LC / 2) % NT) == 0;
  // exactly two pairs per lane: phase C walks the four roots together
  constexpr bool MERGED = ALL && LC * LC / 2 == 2 * NT;
  static_assert(P == 1 || LC > 0, "lattices share a workgroup only at a compile-time L");
  // One lattice of a compile-time side up to 64 sits in static LDS and names
  // its forest nodes by byte offset (handles, see lds_at): the walks then need
  // no address arithmetic.
  constexpr bool BA = lds_static(P, LC);
  using Key = ClusterKey<(LC > 0 && LC <= 32), BA>;
  using KT = typename Key::T;
  extern __shared__ uint32_t smem_all[];
  __shared__ uint32_t smem_fix[BA ? lds_words((uint32_t)(LC * LC)) : 1];
  __shared__ uint32_t red_c_all[P][NW][3];  // satisfied bonds, up spins, roots
  __shared__ KT red_k_all[P][NW][2];        // two largest cluster keys
  __shared__ uint32_t wrapflag_all[P][2];   // by sweep parity
  __shared__ uint64_t a_all[P][RBIM_ACC_SLOTS];  // running sums of the open bin (fold lane)

  // wavefront-uniform (NT is a multiple of 64): the lattice's LDS base is scalar
  const int sub = P > 1 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x / NT) : 0;
  const uint32_t N = LC ? (uint32_t)(LC * LC) : (uint32_t)s.N;
  const uint32_t L = LC ? (uint32_t)LC : (uint32_t)s.L;
  const uint32_t magic = s.magic;
  uint32_t* smem = BA ? smem_fix : smem_all + (size_t)sub * lds_words(N);
  uint32_t (&red_c)[NW][3] = red_c_all[sub];
  KT (&red_k)[NW][2] = red_k_all[sub];
  uint32_t (&wrapflag)[2] = wrapflag_all[sub];
  uint64_t (&a)[RBIM_ACC_SLOTS] = a_all[sub];
  uint32_t* uf = smem;
  uint32_t* cnt = smem + N;
  uint8_t* sb = reinterpret_cast<uint8_t*>(smem + 2 * N);

  const int ri = page * P + sub;
  const bool valid = P == 1 || ri < s.n_run;
  const int lat = run_lat(s, valid ? ri : page * P);  // an idle lattice follows the page's first
  // handle of site n, and site of handle h
  auto H = [](uint32_t n) -> uint32_t { return BA ? n << 2 : n; };
  auto IDX = [](uint32_t h) -> uint32_t { return BA ? h >> 2 : h; };
  const int64_t g0 = s.sweep_ctr[lat];
  const int64_t gend = min(g0 + (int64_t)sweeps, lat_total(s, lat));
  if (gend <= g0) return;  // finished (or empty) slot: its state stays as it is
  const int64_t n_therm = lat_therm(s, lat), spb = lat_spb(s, lat);
  const int tid = P > 1 ? (int)threadIdx.x % NT : (int)threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  const uint32_t npairs = (N + 1) >> 1;
  const int iters = (int)((npairs + NT - 1) / NT);
  // B1 unions through a per-wave queue in cnt when every wave's share of cnt
  // holds the up to 3 x 64 requests of one pass
  const uint32_t qw = N / NW;
  const bool compact = qw >= 192u;

  int8_t* gsp = s.spins + (size_t)lat * N;
  const uint8_t* gj = s.jbits + (size_t)lat * N;
  for (uint32_t n = tid; n < N; n += NT)
    sb[n] = valid ? (uint8_t)((gsp[n] > 0 ? 1u : 0u) | ((uint32_t)gj[n] & 3u) << 1) : (uint8_t)1u;
  if (tid == 0) {
    wrapflag[0] = 0u;
    wrapflag[1] = 0u;
  }

  const uint32_t thr = valid ? s.thr31[lat] : 0u;
  const uint64_t seed = s.seed[lat];
  const uint32_t lid = s.lat_id[lat];
  const bool given = s.given_occ != nullptr;
  uint64_t* gacc = s.acc + (size_t)lat * s.n_bins * RBIM_ACC_SLOTS;

  // The fold of each measured sweep runs while the first 2L threads unite
  // the boundary bonds (B2): on the first wavefront past them, else on the
  // last one.  Its bin state lives in registers of that lane.
  const int fold_tid = min((int)((2u * L + 63u) / 64u) * 64, NT - 64);
  // The B1 union queues (in cnt) are cleared during B2 as well, by the
  // wavefronts past the fold lane's (all of them when there are none).
  const int clr0 = fold_tid + 64 < NT ? fold_tid + 64 : 0;
  int64_t cur_bin = -1, bin_end = 0;  // lane fold_tid only
  __syncthreads();

  // Measurement bookkeeping for one pass (counts are wave-uniform).
  uint32_t m_sat = 0u, m_up = 0u, m_root = 0u;
  KT k1 = 0u, k2 = 0u;
  auto measure_site = [&](bool valid, uint32_t n, uint32_t b, uint32_t sat_h, uint32_t sat_v) {
    m_sat += popc_ballot(valid && sat_h) + popc_ballot(valid && sat_v);
    m_up += popc_ballot(valid && (b & 1u));
    const bool root = valid && (lds_get<BA>(uf, H(n)) & PMASK) == H(n);
    m_root += popc_ballot(root);
    const KT key = Key::make(cnt[n], H(n));
    top2_merge(k1, k2, root ? key : (KT)0u, (KT)0u);
  };
  auto publish = [&]() {
    wave_top2(k1, k2);
    if (lane == 63) {
      red_c[wave][0] = m_sat;
      red_c[wave][1] = m_up;
      red_c[wave][2] = m_root;
      red_k[wave][0] = k1;
      red_k[wave][1] = k2;
    }
    m_sat = m_up = m_root = 0u;
    k1 = k2 = 0u;
  };
  // One lane: fold the published partials of sweep gm into its bin.
  auto finalize = [&](int64_t gm) {
    if (!valid) return;
    uint32_t sat = 0u, up = 0u, roots = 0u;
    KT f1 = 0u, f2 = 0u;
    for (int w = 0; w < NW; ++w) {
      sat += red_c[w][0];
      up += red_c[w][1];
      roots += red_c[w][2];
      top2_merge(f1, f2, red_k[w][0], red_k[w][1]);
    }
    rbim_sweep_stats st;
    st.energy = 2 * (int64_t)N - 2 * (int64_t)sat;
    st.magnetization = 2 * (int64_t)up - (int64_t)N;
    st.n_clusters = roots;
    const int64_t s1 = Key::size(f1), mc1 = 2 * Key::n_plus(f1) - s1;
    st.largest_size = s1;
    st.largest_abs_mag = mc1 < 0 ? -mc1 : mc1;
    if (f2 != 0u) {
      const int64_t s2 = Key::size(f2), mc2 = 2 * Key::n_plus(f2) - s2;
      st.second_size = s2;
      st.second_abs_mag = mc2 < 0 ? -mc2 : mc2;
    } else {
      st.second_size = 0;
      st.second_abs_mag = 0;
    }
    const uint32_t wf = wrapflag[gm & 1];
    wrapflag[gm & 1] = 0u;
    st.wrap_x = (int32_t)(wf & 1u);
    st.wrap_y = (int32_t)((wf >> 1) & 1u);
    if (cur_bin < 0 || gm >= bin_end) {  // measured sweeps arrive in order
      const int64_t bin = (gm - n_therm) / spb;
      if (cur_bin >= 0)
        for (int k = 0; k < RBIM_ACC_SLOTS; ++k) gacc[cur_bin * RBIM_ACC_SLOTS + k] = a[k];
      for (int k = 0; k < RBIM_ACC_SLOTS; ++k) a[k] = gacc[bin * RBIM_ACC_SLOTS + k];
      cur_bin = bin;
      bin_end = n_therm + (bin + 1) * spb;
    }
    acc_sweep(a, st);
    if (s.last) s.last[lat] = st;
    series_put(s, lat, gm, st);
  };
  // satisfied flags of the pair's four bonds, from the current spins
  auto satisfied = [&](const Pair& p, uint32_t& b0, uint32_t& b1, uint32_t& sh0, uint32_t& sv0,
                       uint32_t& sh1, uint32_t& sv1) {
    b0 = b1 = sh0 = sv0 = sh1 = sv1 = 0u;
    if (p.v0) {
      b0 = sb[p.n0];
      sh0 = ~(b0 ^ sb[p.nr0] ^ (b0 >> 1)) & 1u;
      sv0 = ~(b0 ^ sb[p.nu0] ^ (b0 >> 2)) & 1u;
    }
    if (p.v1) {
      b1 = sb[p.n1];
      sh1 = ~(b1 ^ sb[p.nr1] ^ (b1 >> 1)) & 1u;
      sv1 = ~(b1 ^ sb[p.nu1] ^ (b1 >> 2)) & 1u;
    }
  };

  // MERGED: the run starts of the lane's four sites, kept from phase A (the
  // sites' parents then), where B1's unions and C's root walks start: a site
  // and its run start have one root, and halving from the run start compresses
  // the path every site of the run shares
  uint32_t rs[4] = {0u, 0u, 0u, 0u};
  bool pending = false;  // sweep g-1 still to be measured (block-uniform)
  for (int64_t g = g0; g < gend; ++g) {
    // ------------------------------------------------ A (+ measure g-1)
#pragma unroll RBIM_UNROLL_A
    for (int it = 0; it < iters; ++it) {
      const uint32_t q = (uint32_t)(it * NT + tid);
      const uint32_t chunk = 2u * (uint32_t)(it * NT + wave * 64);
      const Pair p = make_pair<LC, ALL>(q, N, L, magic);
      uint32_t b0, b1, sh0, sv0, sh1, sv1;
      satisfied(p, b0, b1, sh0, sv0, sh1, sv1);
      if (pending) {
        measure_site(p.v0, p.n0, b0, sh0, sv0);
        measure_site(p.v1, p.n1, b1, sh1, sv1);
      }
      uint32_t oh0, ov0, fl0, oh1, ov1, fl1;
      if (given) {
        const uint32_t o0 = p.v0 ? s.given_occ[p.n0] : 0u;
        const uint32_t o1 = p.v1 ? s.given_occ[p.n1] : 0u;
        oh0 = o0 & 1u;
        ov0 = (o0 >> 1) & 1u;
        fl0 = p.v0 ? s.given_flip[p.n0] & 1u : 0u;
        oh1 = o1 & 1u;
        ov1 = (o1 >> 1) & 1u;
        fl1 = p.v1 ? s.given_flip[p.n1] & 1u : 0u;
      } else {
        const rbim_u32x4 r = rbim_sweep_draws(seed, lid, (uint32_t)g, q);
        oh0 = sh0 & ((r.x >> 1) < thr ? 1u : 0u);
        ov0 = sv0 & ((r.y >> 1) < thr ? 1u : 0u);
        fl0 = r.x & 1u;
        oh1 = sh1 & ((r.z >> 1) < thr ? 1u : 0u);
        ov1 = sv1 & ((r.w >> 1) < thr ? 1u : 0u);
        fl1 = r.z & 1u;
      }
      // run pre-labelling of in-row bonds inside this wave's 128-site chunk
      const bool internal = p.v1 && ov0 && p.y0 != L - 1u;
      const bool outlink = p.v1 && ov1 && p.y1 != L - 1u;
      const uint32_t start0 = run_start(__ballot(internal && outlink), __ballot(outlink), below, chunk);
      if constexpr (MERGED) {
        const uint32_t s1 = internal ? start0 : p.n1;
        if (it == 0) {
          rs[0] = start0;
          rs[1] = s1;
        } else {
          rs[2] = start0;
          rs[3] = s1;
        }
      }
      if (p.v0) {
        sb[p.n0] = (uint8_t)((b0 & 7u) | (oh0 << 3) | (ov0 << 4) | (fl0 << 5));
        lds_at<BA>(uf, H(p.n0)) = H(start0);
        if (!compact) cnt[p.n0] = 0u;  // else B1's queue holds cnt until B2 clears it
      }
      if (p.v1) {
        sb[p.n1] = (uint8_t)((b1 & 7u) | (oh1 << 3) | (ov1 << 4) | (fl1 << 5));
        lds_at<BA>(uf, H(p.n1)) = internal ? H(start0) : H(p.n1);
        if (!compact) cnt[p.n1] = 0u;
      }
    }
    if (pending) publish();
    __syncthreads();
    // ----------------------------------------------------------- B1
    // Bonds that do not cross the periodic boundary: x bonds below row L-1 and
    // the in-row y bond that leaves the wave's chunk (lane 63).  The forest
    // has no windings yet, so plain parent chasing suffices.  With compact
    // queues the passes fill one queue per wave (in cnt, cleared in B2).
    uint32_t qn = 0u;
#pragma unroll RBIM_UNROLL_B
    for (int it = 0; it < iters; ++it) {
      const uint32_t q = (uint32_t)(it * NT + tid);
      const Pair p = make_pair<LC, ALL>(q, N, L, magic);
      // An occupied x bond (n -> n+L) is skipped when the other three bonds
      // of the plaquette to its left -- the x bond of n-1 and the in-row y
      // bonds n-1 -> n and n-1+L -> n+L -- are occupied: the four close a
      // contractible loop of winding 0, so the skipped bond changes neither
      // the clusters nor the winding of any cycle.  Skips chain along a row;
      // by induction the endpoints of every skipped bond stay connected
      // through kept bonds with the same displacement.
      uint32_t req = 0u;
      uint32_t b0 = 0u;
      if (p.v0) {
        b0 = sb[p.n0];
        uint32_t h0 = (b0 >> 3) & 1u & (p.x0 != L - 1u ? 1u : 0u);
        {
          const bool edge = p.y0 == 0u;
          const uint32_t left = sb[edge ? p.n0 : p.n0 - 1u], left_up = sb[edge ? p.nr0 : p.nr0 - 1u];
          h0 &= ~((left >> 3) & (left >> 4) & (left_up >> 4) & (edge ? 0u : 1u)) & 1u;
        }
        req |= h0;
      }
      if (p.v1) {
        const uint32_t b1 = sb[p.n1];
        uint32_t h1 = (b1 >> 3) & 1u & (p.x1 != L - 1u ? 1u : 0u);
        h1 &= ~((b0 >> 3) & (b0 >> 4) & (sb[p.nr0] >> 4) & (p.y1 != 0u ? 1u : 0u)) & 1u;
        req |= h1 << 1;
        req |= ((b1 >> 4) & 1u & ((lane == 63 && p.y1 != L - 1u) ? 1u : 0u)) << 2;
      }
      if (compact) {
        // the wave's requests go through a dense queue (in cnt, cleared in B2)
        uint32_t qa[3] = {H(p.n0), H(p.n1), H(p.n1)};
        uint32_t qb[3] = {H(p.nr0), H(p.nr1), H(p.nu1)};
        if constexpr (MERGED) {
          // from the run starts: the lane's own, and for n0 + L, n1 + L their
          // parent words, read here ahead of the dependent finds (any parent
          // word names an ancestor, and B1's forest carries no windings; the
          // reads of the lanes that ask for no union go unused)
          const uint32_t s0 = it == 0 ? rs[0] : rs[2], s1 = it == 0 ? rs[1] : rs[3];
          qa[0] = H(s0);
          qa[1] = qa[2] = H(s1);
          qb[0] = lds_get<BA>(uf, H(p.nr0));
          qb[1] = lds_get<BA>(uf, H(p.nr1));
        }
        // worked when the next pass could overflow it, and after the last
        if (qn + WAVE_QUEUE_PASS > qw) {
          wave_queue_run<BA, true>(uf, cnt + wave * qw, qn);
          qn = 0u;
        }
        qn = wave_queue_put(cnt + wave * qw, qn, req, qa, qb, below);
      } else {
        while (__ballot(req != 0u)) {
          if (req) {
            const uint32_t k = (uint32_t)__builtin_ctz(req);
            req &= req - 1u;
            lds_union_plain<BA>(uf, H(k == 0u ? p.n0 : p.n1),
                                H(k == 0u ? p.nr0 : k == 1u ? p.nr1 : p.nu1));
          }
        }
      }
    }
    if (compact) wave_queue_run<BA, true>(uf, cnt + wave * qw, qn);
    __syncthreads();
    // ----------------------------------------------------------- B2
    // The at most 2L bonds across the periodic boundary, with windings: x
    // bonds of row L-1 and y bonds of column L-1.  A bond that finds both ends
    // under one root closes a loop whose winding c + w(b) - w(a) decides
    // wrap_x / wrap_y.
    uint32_t wraps = 0u;
    // When one wavefront holds all of B2 (2L <= 64), the other wavefronts of
    // the lattice wait at the next barrier for it alone: it issues ahead of
    // the other workgroups' wavefronts on its SIMD meanwhile, above the B1
    // unions too (the wavefront's own issue priority; at L=32 1.310e11 ->
    // 1.341e11 site-updates/s, +1.6 % at L = 16 and 24; for two B2
    // wavefronts, L = 48 / 64, +1.2 % / -0.8 %).
    const bool b2_first = 2u * L <= 64u && wave == 0;
    if (b2_first) __builtin_amdgcn_s_setprio(2);
    for (uint32_t t = tid; t < 2u * L; t += NT) {
      // t < L: y bond of row t across y = L-1; t >= L: x bond of row L-1 at y = t-L
      const bool horiz = t >= L;
      const uint32_t a = horiz ? (L - 1u) * L + (t - L) : t * L + (L - 1u);
      uint32_t occ = (sb[a] >> (horiz ? 3 : 4)) & 1u;
      // B1's plaquette skip along the boundary: the y bond of row i (i = t)
      // is skipped when the y bond of row i - 1 and the x bonds of (i - 1,
      // L - 1) and (i - 1, 0) are occupied; the x bond at column i (i = t - L)
      // when the x bond at column i - 1 and the in-row bonds (L - 1, i - 1) ->
      // (L - 1, i) and (0, i - 1) -> (0, i) are.  The four bonds close a loop
      // that crosses the boundary once each way (winding 0); row 0 and column
      // 0 are never skipped, so by induction along the boundary a skipped
      // bond's ends stay joined through kept bonds with the same crossings.
      const uint32_t i = horiz ? t - L : t;
      if (occ && i != 0u) {
        // the partner bond's site, and the site of the plaquette's far bond
        const uint32_t p0 = horiz ? (L - 1u) * L + i - 1u : (i - 1u) * L + L - 1u;
        const uint32_t q0 = horiz ? i - 1u : (i - 1u) * L;
        const uint32_t bp = sb[p0];
        const uint32_t sh = horiz ? 3u : 4u, oh = horiz ? 4u : 3u;  // along / across
        occ &= ~((bp >> sh) & (bp >> oh) & (sb[q0] >> oh)) & 1u;
      }
      if (occ)
        wraps |= lds_union<BA>(uf, H(a), H(horiz ? t - L : t * L), horiz ? 1 : 0, horiz ? 0 : 1);
    }
    wave_or_wraps(&wrapflag[g & 1], wraps);
    if (b2_first) __builtin_amdgcn_s_setprio(0);
    // sweep g-1's partials were published before the barrier after A, and its
    // wrap slot is not the one B2 of sweep g sets
    if (pending && tid == fold_tid) finalize(g - 1);
    if (compact && tid >= clr0)
      for (uint32_t i = (uint32_t)(tid - clr0); i < N; i += (uint32_t)(NT - clr0)) cnt[i] = 0u;
    __syncthreads();
    // ------------------------------------------------------------ C
    // Counts go to cnt[root] by LDS atomics; the sites of the first lane's
    // cluster (at low temperature usually most of the wave, one giant cluster)
    // are summed by ballots and added by a single atomic.
    // The root walks are dependent LDS round trips: with one walk per site
    // pair in flight they go ahead of other wavefronts' arithmetic (+3.5 % at
    // L = 24 and 48); with four (MERGED) the walks overlap by themselves (-0.5 %
    // at L=32 with the priority).
    if constexpr (!MERGED) __builtin_amdgcn_s_setprio(1);
    uint32_t rr[4] = {0u, 0u, 0u, 0u};  // MERGED: roots of n0, n1 of both pairs
    uint32_t bb[2] = {0u, 0u};          // MERGED: state bytes of both n0
    if constexpr (MERGED) {  // even L: n1 is linked to n0 by the in-row bond bit
#pragma unroll
      for (int it = 0; it < 2; ++it) {
        const uint32_t n0 = 2u * (uint32_t)(it * NT + tid);
        bb[it] = sb[n0];
        rr[2 * it] = H(rs[2 * it]);  // n1's run start is n0's when they are linked
        rr[2 * it + 1] = H(rs[2 * it + 1]);
      }
      lds_root4<BA>(uf, rr);
    }
    // flip and count of the lane's pair of pass it
    auto c_pass = [&](int it) {
      const uint32_t q = (uint32_t)(it * NT + tid);
      const Pair p = make_pair<LC, ALL>(q, N, L, magic);
      uint32_t r0 = 0u, r1 = 0u, ns0 = 0u, ns1 = 0u;
      if (p.v0) {
        uint32_t b0;
        if constexpr (MERGED) {
          b0 = it == 0 ? bb[0] : bb[1];
          r0 = it == 0 ? rr[0] : rr[2];
          r1 = it == 0 ? rr[1] : rr[3];
        } else {
          b0 = sb[p.n0];
          // n1 shares n0's run (and root) when the in-row bond n0 -> n1 is occupied
          const bool linked = ((b0 >> 4) & 1u) && p.y0 != L - 1u;
          r0 = H(p.n0);
          r1 = H(p.v1 && !linked ? p.n1 : p.n0);
          lds_root2<BA>(uf, r0, r1);
        }
        ns0 = (b0 ^ (sb[IDX(r0)] >> 5)) & 1u;
        sb[p.n0] = (uint8_t)((b0 & ~1u) | ns0);
        if (s.canon_out) s.canon_out[p.n0] = (int32_t)IDX(r0);
        if (p.v1) {
          const uint32_t b1 = sb[p.n1];
          ns1 = (b1 ^ (sb[IDX(r1)] >> 5)) & 1u;
          sb[p.n1] = (uint8_t)((b1 & ~1u) | ns1);
          if (s.canon_out) s.canon_out[p.n1] = (int32_t)IDX(r1);
        }
      }
      const uint32_t R = __builtin_amdgcn_readfirstlane(r0);
      const bool in0 = p.v0 && r0 == R, in1 = p.v1 && r1 == R;
      const uint64_t m0 = __ballot(in0), m1 = __ballot(in1);
      const uint64_t u0 = __ballot(in0 && ns0), u1 = __ballot(in1 && ns1);
      if (__builtin_amdgcn_readfirstlane((uint32_t)lane) == (uint32_t)lane && (m0 | m1))
        atomicAdd(&lds_at<BA>(cnt, R), (uint32_t)(__popcll(m0) + __popcll(m1)) << 16 |
                                           (uint32_t)(__popcll(u0) + __popcll(u1)));
      const bool rest0 = p.v0 && !in0, rest1 = p.v1 && !in1;
      if (rest0 && rest1 && r0 == r1) {
        atomicAdd(&lds_at<BA>(cnt, r0), 0x20000u | (ns0 + ns1));
      } else {
        if (rest0) atomicAdd(&lds_at<BA>(cnt, r0), 0x10000u | ns0);
        if (rest1) atomicAdd(&lds_at<BA>(cnt, r1), 0x10000u | ns1);
      }
        };
    if constexpr (MERGED) {
      c_pass(0);
      c_pass(1);
    } else {
#pragma unroll 1
      for (int it = 0; it < iters; ++it) c_pass(it);
    }
    if constexpr (!MERGED) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    pending = g >= n_therm;
    // an unmeasured sweep's wrap bits are dropped; its slot is reused two
    // sweeps later, after two more barriers
    if (!pending && tid == 0) wrapflag[g & 1] = 0u;
  }

  if (pending) {  // measure the last sweep of this launch
#pragma unroll 1
    for (int it = 0; it < iters; ++it) {
      const uint32_t q = (uint32_t)(it * NT + tid);
      const Pair p = make_pair<LC, ALL>(q, N, L, magic);
      uint32_t b0, b1, sh0, sv0, sh1, sv1;
      satisfied(p, b0, b1, sh0, sv0, sh1, sv1);
      measure_site(p.v0, p.n0, b0, sh0, sv0);
      measure_site(p.v1, p.n1, b1, sh1, sv1);
    }
    publish();
    __syncthreads();
    if (tid == fold_tid) finalize(gend - 1);
  }

  if (!valid) return;
  if (tid == fold_tid && cur_bin >= 0)
    for (int k = 0; k < RBIM_ACC_SLOTS; ++k) gacc[cur_bin * RBIM_ACC_SLOTS + k] = a[k];
  if (tid == 0) s.sweep_ctr[lat] = gend;
  for (uint32_t n = tid; n < N; n += NT) gsp[n] = (sb[n] & 1u) ? (int8_t)1 : (int8_t)-1;
}

template <int NT, int LC>
__global__ __launch_bounds__(NT) void lds_sweep_kernel(LatticeSet s, int32_t sweeps) {
  lds_sweeps<NT, LC, 1>(s, sweeps, blockIdx.x);
}

// One launch over the LDS-resident lattices of several sides: workgroups of
// 1024 threads, each holding a page of lattices of one side -- one of L = 128,
// two of L = 64 (512 threads each), four of L = 32 (256 threads), sixteen of
// L = 16 (a wavefront each: the per-side kernels' own workgroup sizes, which
// keep every side's loop within the 128 registers of a 1024-thread
// workgroup) -- so one dynamic LDS size serves every side.  set[k] holds the
// lattices of side RAGGED_L[k]; page0[k] is its first page.  Workgroups take
// the pages from the last (page = grid size - 1 - workgroup id), so they go out
// largest first (L = 128, 64, 32, 16: 16384, 8192, 4096 and 4096 sites) and the
// small pages fill the last round of workgroups.
constexpr int RAGGED_L[4] = {16, 32, 64, 128};
constexpr int RAGGED_P[4] = {16, 4, 2, 1};
struct RaggedLaunch {
  LatticeSet set[4];
  int32_t page0[4];
};

__global__ __launch_bounds__(1024) void lds_ragged_kernel(RaggedLaunch r, int32_t sweeps) {
  const int b = (int)(gridDim.x - 1u - blockIdx.x);
  // (largest side first: the order the register allocator handles best)
  if (b >= r.page0[3]) {
    lds_sweeps<1024, 128, 1>(r.set[3], sweeps, b - r.page0[3]);
    return;
  }
  if (b >= r.page0[2]) {
    lds_sweeps<512, 64, 2>(r.set[2], sweeps, b - r.page0[2]);
    return;
  }
  if (b >= r.page0[1]) {
    lds_sweeps<256, 32, 4>(r.set[1], sweeps, b - r.page0[1]);
    return;
  }
  lds_sweeps<64, 16, 16>(r.set[0], sweeps, b - r.page0[0]);
}

__global__ void init_spins_kernel(LatticeSet s, int start) {
  const int lat = blockIdx.y;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= s.N) return;
  s.spins[(size_t)lat * s.N + n] =
      start == RBIM_START_COLD ? (int8_t)1 : (int8_t)rbim_init_spin(s.seed[lat], s.lat_id[lat], (uint32_t)n);
}

template <int NT, int LC>
hipError_t launch_lds_t(const LatticeSet& s, int sweeps, hipStream_t st) {
  static std::atomic<uint64_t> configured{0};
  if (!lds_static(1, LC)) {
    hipError_t e = ensure_dynamic_lds((const void*)lds_sweep_kernel<NT, LC>,
                                      (int)lds_bytes(RBIM_LDS_MAX_N), configured);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((lds_sweep_kernel<NT, LC>), dim3(s.n_run), dim3(NT),
                     lds_static(1, LC) ? 0 : lds_bytes(s.N), st, s, sweeps);
  return hipGetLastError();
}

template <int NT>
hipError_t launch_lds_l(const LatticeSet& s, int sweeps, hipStream_t st) {
  switch (s.L) {
    case 16: return launch_lds_t<NT, 16>(s, sweeps, st);
    case 24: return launch_lds_t<NT, 24>(s, sweeps, st);
    case 32: return launch_lds_t<NT, 32>(s, sweeps, st);
    case 48: return launch_lds_t<NT, 48>(s, sweeps, st);
    case 64: return launch_lds_t<NT, 64>(s, sweeps, st);
    case 128: return launch_lds_t<NT, 128>(s, sweeps, st);
    default: return launch_lds_t<NT, 0>(s, sweeps, st);
  }
}

}  // namespace

size_t lds_bytes(int N) { return (size_t)lds_words((uint32_t)N) * 4u; }

int ragged_slot(int L) {
  for (int k = 0; k < 4; ++k)
    if (RAGGED_L[k] == L) return k;
  return -1;
}

int ragged_threads(int L) {
  const int k = ragged_slot(L);
  return k < 0 ? 0 : 1024 / RAGGED_P[k];
}

hipError_t launch_lds_ragged(const LatticeSet* const sets[4], int sweeps, hipStream_t st) {
  static std::atomic<uint64_t> configured{0};
  RaggedLaunch r{};
  int pages = 0;
  size_t bytes = 0;
  for (int k = 0; k < 4; ++k) {
    r.page0[k] = pages;
    if (!sets[k]) continue;
    r.set[k] = *sets[k];
    pages += (r.set[k].n_run + RAGGED_P[k] - 1) / RAGGED_P[k];
    bytes = std::max(bytes, (size_t)RAGGED_P[k] * lds_bytes(RAGGED_L[k] * RAGGED_L[k]));
  }
  if (pages == 0) return hipSuccess;
  hipError_t e = ensure_dynamic_lds((const void*)lds_ragged_kernel,
                                    (int)lds_bytes(RBIM_LDS_MAX_N), configured);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(lds_ragged_kernel, dim3(pages), dim3(1024), bytes, st, r, sweeps);
  return hipGetLastError();
}

hipError_t launch_lds_sweeps(const LatticeSet& s, int threads, int sweeps, hipStream_t st) {
  switch (threads) {
    case 64: return launch_lds_l<64>(s, sweeps, st);
    case 128: return launch_lds_l<128>(s, sweeps, st);
    case 256: return launch_lds_l<256>(s, sweeps, st);
    case 512: return launch_lds_l<512>(s, sweeps, st);
    case 1024: return launch_lds_l<1024>(s, sweeps, st);
    default: return hipErrorInvalidValue;
  }
}

template <int NT>
int lds_blocks_per_cu_t(int L) {
  const int N = L * L;
  int n = 0;
  hipError_t e;
  switch (L) {
    case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lds_sweep_kernel<NT, 16>, NT, 0); break;
    case 24: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lds_sweep_kernel<NT, 24>, NT, 0); break;
    case 32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lds_sweep_kernel<NT, 32>, NT, 0); break;
    case 48: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lds_sweep_kernel<NT, 48>, NT, 0); break;
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lds_sweep_kernel<NT, 64>, NT, 0); break;
    case 128: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lds_sweep_kernel<NT, 128>, NT, lds_bytes(N)); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lds_sweep_kernel<NT, 0>, NT, lds_bytes(N)); break;
  }
  return e == hipSuccess ? n : 0;
}

int lds_blocks_per_cu(int L, int threads) {
  switch (threads) {
    case 64: return lds_blocks_per_cu_t<64>(L);
    case 128: return lds_blocks_per_cu_t<128>(L);
    case 256: return lds_blocks_per_cu_t<256>(L);
    case 512: return lds_blocks_per_cu_t<512>(L);
    case 1024: return lds_blocks_per_cu_t<1024>(L);
    default: return 0;
  }
}

hipError_t launch_init_spins(const LatticeSet& s, int start, hipStream_t st) {
  dim3 grid((s.N + 255) / 256, s.n_lat);
  hipLaunchKernelGGL(init_spins_kernel, grid, dim3(256), 0, st, s, start);
  return hipGetLastError();
}
