// Multi-kernel tiled Swendsen-Wang sweep for any L >= 3 on gfx950: lattices
// whose side is not a multiple of 32, the one-lattice parity entry points
// (given bits, canonical labels) and RBIM_ENGINE_TILED.  A sweep is four
// stream-ordered kernels (tiles of 64 x 64 when 64 divides L >= 128, else
// 32 x 32; partial tiles at the edge when the tile does not divide L):
//
//   K1 tile      one workgroup per tile: measures the previous sweep on the
//                tile (E, M from the spins; clusters from the root flag K3
//                left and the root's counters), Philox activation of every
//                site, occupation/flip bits to HBM, union-find of the bonds
//                inside the tile in LDS (no winding: a tile never crosses the
//                periodic boundary).  Every site gets its tile-local root (the
//                min site of its tile-local cluster, as a 16-bit tile index);
//                each local root opens a node of the HBM forest and records its
//                local size / n_plus.
//   K2 merge     one thread per tile-boundary bond (including the periodic
//                ones): union of the two local roots in the HBM forest with a
//                64-bit CAS on (parent | wx | wy), larger root hooked under
//                smaller; a bond that finds both ends under one root is a
//                non-tree bond and its loop winding decides wrap_x / wrap_y
//                (_unionfind.py:101-124).  Bonds inside a tile close only
//                windingless loops and need no test.
//   K3 flip      per tile: each local root resolves its global root and flip
//                bit into an LDS table, every site flips from it, each local
//                root adds its (size, n_plus after the flip) to its global root
//                -- one atomic per local cluster -- and global roots are
//                flagged for the next K1.
//   K5 finalize  per lattice: fold the previous sweep's tile partials into its
//                bin's moment sums (simulation.py:212-236), advance the counter.
//
// K4 + a flush kernel measure and fold the last sweep of a run.  Every kernel
// reads the lattice's own sweep counter from HBM, so a captured graph of
// sweeps replays without host arguments changing.
#include "engine.h"
#include "tile_common.h"

namespace {

// ------------------------------------------------------------ measurement
// Bits of the per-site occ byte: bonds and flip bit of the sweep, and -- set by
// K3 at the root of every cluster -- the root flag the next measurement reads.
constexpr uint32_t OCC_ROOT = 8u;

// Bond energy of site n's two forward bonds (couplings j: bit 0 horizontal
// negative, bit 1 vertical negative) with neighbour spins sr and su.
__device__ __forceinline__ int bond_energy(uint32_t j, int s0, int sr, int su) {
  return -(((j & 1u) ? -1 : 1) * s0 * sr) - (((j & 2u) ? -1 : 1) * s0 * su);
}

// Cluster rooted at site n with counters c = size << 32 | n_plus.
__device__ __forceinline__ void meas_root(TileMeas& a, uint64_t c, uint32_t n) {
  ++a.ncl;
  top2_insert(a, ((c >> 32) << 32) | (uint64_t)(0xFFFFFFFFu - n), (uint32_t)c);
}

template <int NTH>
__device__ void meas_combine(int64_t (*red)[PARTS], int64_t* out) {
  TileMeas a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
  int64_t e = 0, m = 0, ncl = 0;
#pragma unroll 1
  for (int w = 0; w < NTH / 64; ++w) {
    e += red[w][0];
    m += red[w][1];
    ncl += red[w][2];
    top2_insert(a, (uint64_t)red[w][3], (uint32_t)red[w][5]);
    top2_insert(a, (uint64_t)red[w][4], (uint32_t)red[w][6]);
  }
  out[0] = e;
  out[1] = m;
  out[2] = ncl;
  out[3] = (int64_t)a.k1;
  out[4] = (int64_t)a.k2;
  out[5] = a.p1;
  out[6] = a.p2;
}

// Both halves at once (one barrier).
template <int NTH>
__device__ void meas_publish(TileMeas a, int64_t (*red)[PARTS], int64_t* out) {
  wave_clusters(a, red);
  wave_spins(a.e, a.m, red);
  __syncthreads();
  if (threadIdx.x == 0) meas_combine<NTH>(red, out);
}

// ----------------------------------------------------------------- K1
__device__ __forceinline__ uint32_t lfind(uint32_t* par, uint32_t x) {
  uint32_t p = par[x];
  while (p != x) {
    const uint32_t gp = par[p];
    if (gp == p) return p;
    par[x] = gp;
    x = gp;
    p = par[x];
  }
  return x;
}

__device__ __forceinline__ void lunion(uint32_t* par, uint32_t a, uint32_t b) {
  while (true) {
    const uint32_t ra = lfind(par, a), rb = lfind(par, b);
    if (ra == rb) return;
    const uint32_t hi = max(ra, rb), lo = min(ra, rb);
    if (atomicCAS(&par[hi], hi, lo) == hi) return;
  }
}

template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k1_tile(LatticeSet s, TiledScratch t) {
  __shared__ int64_t red[NTH / 64][PARTS];
  __shared__ uint32_t par[T * T];
  __shared__ uint32_t lcnt[T * T];
  __shared__ uint8_t bits[T * T];  // bit0 spin, bit1 occ_h, bit2 occ_v
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const int L = s.L, N = s.N;
  const int tx = blockIdx.x % t.tiles_x, ty = blockIdx.x / t.tiles_x;
  const int x0 = tx * T, y0 = ty * T;
  const int w = min(T, L - x0), h = min(T, L - y0);
  const size_t base = (size_t)lat * N;
  const int8_t* sp = s.spins + base;
  const uint8_t* jb = s.jbits + base;
  const uint32_t sweep = (uint32_t)s.sweep_ctr[lat];
  const uint32_t thr = s.thr31[lat];
  const uint64_t seed = s.seed[lat];
  const uint32_t lid = s.lat_id[lat];
  const bool given = s.given_occ != nullptr;
  const bool even = (L & 1) == 0;
  const bool meas = t.pending[lat] != 0;  // the previous sweep is measured here
  TileMeas ms = {0, 0, 0, 0ull, 0ull, 0u, 0u};

  // activation: thread handles the local pair (lx, 2k), (lx, 2k+1)
  for (int i = threadIdx.x; i < T * T / 2; i += NTH) {
    const int lx = i / (T / 2), ly0 = 2 * (i % (T / 2));
    if (lx >= w || ly0 >= h) continue;
    const int x = x0 + lx;
    rbim_u32x4 r = {0u, 0u, 0u, 0u};
    const uint32_t n0 = (uint32_t)(x * L + y0 + ly0);
    if (!given && even) r = rbim_sweep_draws(seed, lid, sweep, n0 >> 1);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int ly = ly0 + j;
      if (ly >= h) break;
      const int y = y0 + ly;
      const uint32_t n = (uint32_t)(x * L + y);
      if (!given && !even) r = rbim_sweep_draws(seed, lid, sweep, n >> 1);
      const uint32_t nr = (x == L - 1) ? (uint32_t)y : n + L;
      const uint32_t nu = (y == L - 1) ? n - y : n + 1;
      const int s0 = sp[n];
      if (meas) {
        ms.e += bond_energy(jb[n], s0, sp[nr], sp[nu]);
        ms.m += s0;
        if (t.occ[base + n] & OCC_ROOT) meas_root(ms, t.st[base + n], n);
      }
      uint32_t oh, ov, fl;
      if (given) {
        const uint32_t o = s.given_occ[n];
        oh = o & 1u;
        ov = (o >> 1) & 1u;
        fl = s.given_flip[n] & 1u;
      } else {
        const uint32_t jbn = jb[n];
        const int jh = (jbn & 1u) ? -1 : 1, jv = (jbn & 2u) ? -1 : 1;
        const bool odd = (n & 1u) != 0;
        const uint32_t dh = odd ? r.z : r.x, dv = odd ? r.w : r.y;
        oh = (jh * s0 * sp[nr] > 0) && ((dh >> 1) < thr) ? 1u : 0u;
        ov = (jv * s0 * sp[nu] > 0) && ((dv >> 1) < thr) ? 1u : 0u;
        fl = dh & 1u;
      }
      t.occ[base + n] = (uint8_t)(oh | (ov << 1) | (fl << 2));
      const int li = lx * T + ly;
      bits[li] = (uint8_t)((s0 > 0 ? 1u : 0u) | (oh << 1) | (ov << 2));
      par[li] = (uint32_t)li;
      lcnt[li] = 0u;
    }
  }
  if (meas) meas_publish<NTH>(ms, red, t.part + ((size_t)lat * t.tiles_x * t.tiles_y + blockIdx.x) * PARTS);
  __syncthreads();
  // local union over bonds with both ends in the tile and no periodic crossing
  for (int li = threadIdx.x; li < T * T; li += NTH) {
    const int lx = li / T, ly = li % T;
    if (lx >= w || ly >= h) continue;
    const uint32_t b = bits[li];
    if ((b & 2u) && lx + 1 < w && x0 + lx + 1 < L) lunion(par, li, li + T);
    if ((b & 4u) && ly + 1 < h && y0 + ly + 1 < L) lunion(par, li, li + 1);
  }
  __syncthreads();
  for (int li = threadIdx.x; li < T * T; li += NTH) {
    const int lx = li / T, ly = li % T;
    if (lx >= w || ly >= h) continue;
    const uint32_t r = lfind(par, li);
    atomicAdd(&lcnt[r], 0x10000u | (bits[li] & 1u));
    const uint32_t n = (uint32_t)((x0 + lx) * L + y0 + ly);
    t.lroot[base + n] = (uint16_t)r;
  }
  __syncthreads();
  for (int li = threadIdx.x; li < T * T; li += NTH) {
    const int lx = li / T, ly = li % T;
    if (lx >= w || ly >= h || par[li] != (uint32_t)li) continue;
    const uint32_t n = (uint32_t)((x0 + lx) * L + y0 + ly);
    t.gf[base + n] = gpack(n, 0, 0);
    t.lst[base + n] = lcnt[li];
    t.st[base + n] = 0ull;
  }
}

// Full T x T tiles (L a multiple of T): the LDS kernel's scheme on an open
// tile.  Each lane owns a site pair (one Philox call); occupied in-row bonds
// inside a wavefront's 128-site chunk are merged by two ballots (every site
// points at the start of its run); the remaining in-tile bonds go through a
// winding-free union-find in LDS, skipping any x bond that closes an occupied
// plaquette with its left neighbour (a contractible loop: connectivity and
// windings are unchanged).  Bounded to 8 waves per SIMD: the LDS of four
// 512-thread 64 x 64 tiles fits a CU.
template <int T, int NTH>
__global__ __launch_bounds__(NTH, 8) void k1_tile_full(LatticeSet s, TiledScratch t) {
  __shared__ int64_t red[NTH / 64][PARTS];
  __shared__ uint32_t par[T * T];
  __shared__ uint32_t lcnt[T * T];
  __shared__ uint8_t bits[T * T];  // bit0 spin, bit1 occ_h, bit2 occ_v
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const uint32_t L = (uint32_t)s.L, N = (uint32_t)s.N;
  const uint32_t tx = blockIdx.x % t.tiles_x, ty = blockIdx.x / t.tiles_x;
  const uint32_t x0 = tx * T, y0 = ty * T;
  const size_t base = (size_t)lat * N;
  const int8_t* sp = s.spins + base;
  const uint8_t* jb = s.jbits + base;
  const uint32_t sweep = (uint32_t)s.sweep_ctr[lat];
  const uint32_t thr = s.thr31[lat];
  const uint64_t seed = s.seed[lat];
  const uint32_t lid = s.lat_id[lat];
  const bool given = s.given_occ != nullptr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  constexpr int ITERS = T * T / 2 / NTH;
  const bool meas = t.pending[lat] != 0;  // the previous sweep is measured here
  int64_t* const meas_out = t.part + ((size_t)lat * t.tiles_x * t.tiles_y + blockIdx.x) * PARTS;
  int32_t me = 0, mm = 0;
  if (meas) {
    TileMeas ms = {0, 0, 0, 0ull, 0ull, 0u, 0u};
    // clusters of the previous sweep: root flags of this thread's pairs first,
    // then the counters of the flagged roots, all loads in flight together
    uint32_t rf[ITERS];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
      const uint32_t ql = (uint32_t)(it * NTH + tid);
      const uint32_t n0 = (x0 + ql / (T / 2)) * L + y0 + (ql % (T / 2)) * 2u;
      rf[it] = *reinterpret_cast<const uint16_t*>(t.occ + base + n0);  // n0 even: aligned
    }
    // unflagged sites load the tile's first counter instead (one cached line),
    // so every load is issued before the first result is needed
    const uint32_t dummy = x0 * L + y0;
    uint64_t c[2 * ITERS];
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
      const uint32_t ql = (uint32_t)(it * NTH + tid);
      const uint32_t n0 = (x0 + ql / (T / 2)) * L + y0 + (ql % (T / 2)) * 2u;
      c[2 * it] = t.st[base + ((rf[it] & OCC_ROOT) ? n0 : dummy)];
      c[2 * it + 1] = t.st[base + ((rf[it] & (OCC_ROOT << 8)) ? n0 + 1u : dummy)];
    }
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
      const uint32_t ql = (uint32_t)(it * NTH + tid);
      const uint32_t n0 = (x0 + ql / (T / 2)) * L + y0 + (ql % (T / 2)) * 2u;
      if (rf[it] & OCC_ROOT) meas_root(ms, c[2 * it], n0);
      if (rf[it] & (OCC_ROOT << 8)) meas_root(ms, c[2 * it + 1], n0 + 1u);
    }
    wave_clusters(ms, red);
  }

#pragma unroll 1
  for (int it = 0; it < ITERS; ++it) {
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t lx = ql / (T / 2), ly0 = (ql % (T / 2)) * 2u;
    const uint32_t li0 = ql << 1;
    const uint32_t x = x0 + lx, y = y0 + ly0;
    const uint32_t n0 = x * L + y, n1 = n0 + 1u;
    const uint32_t nr0 = (x == L - 1u) ? y : n0 + L;
    const uint32_t nu1 = (y + 1u == L - 1u) ? n1 - (y + 1u) : n1 + 1u;
    const int s0 = sp[n0], s1 = sp[n1];
    const int sr0 = sp[nr0], sr1 = sp[nr0 + 1u], su1 = sp[nu1];
    const uint32_t j0 = jb[n0], j1 = jb[n1];
    if (meas) {
      me += bond_energy(j0, s0, sr0, s1) + bond_energy(j1, s1, sr1, su1);
      mm += s0 + s1;
    }
    uint32_t oh0, ov0, fl0, oh1, ov1, fl1;
    if (given) {
      const uint32_t o0 = s.given_occ[n0], o1 = s.given_occ[n1];
      oh0 = o0 & 1u;
      ov0 = (o0 >> 1) & 1u;
      fl0 = s.given_flip[n0] & 1u;
      oh1 = o1 & 1u;
      ov1 = (o1 >> 1) & 1u;
      fl1 = s.given_flip[n1] & 1u;
    } else {
      const rbim_u32x4 r = rbim_sweep_draws(seed, lid, sweep, n0 >> 1);
      oh0 = (((j0 & 1u) ? -1 : 1) * s0 * sr0 > 0) && ((r.x >> 1) < thr) ? 1u : 0u;
      ov0 = (((j0 & 2u) ? -1 : 1) * s0 * s1 > 0) && ((r.y >> 1) < thr) ? 1u : 0u;
      oh1 = (((j1 & 1u) ? -1 : 1) * s1 * sr1 > 0) && ((r.z >> 1) < thr) ? 1u : 0u;
      ov1 = (((j1 & 2u) ? -1 : 1) * s1 * su1 > 0) && ((r.w >> 1) < thr) ? 1u : 0u;
      fl0 = r.x & 1u;
      fl1 = r.z & 1u;
    }
    t.occ[base + n0] = (uint8_t)(oh0 | (ov0 << 1) | (fl0 << 2));
    t.occ[base + n1] = (uint8_t)(oh1 | (ov1 << 1) | (fl1 << 2));
    bits[li0] = (uint8_t)((s0 > 0 ? 1u : 0u) | (oh0 << 1) | (ov0 << 2));
    bits[li0 + 1u] = (uint8_t)((s1 > 0 ? 1u : 0u) | (oh1 << 1) | (ov1 << 2));
    const bool internal = ov0 != 0u;                       // n0 -> n1, same row and tile
    const bool outlink = ov1 != 0u && ly0 + 1u != T - 1u;  // n1 -> n1+1 inside the tile
    const uint64_t thru = __ballot(internal && outlink);
    const uint64_t ol = __ballot(outlink);
    const uint64_t stop = ~thru & below;
    const uint32_t chunk = 2u * (uint32_t)(it * NTH + wave * 64);
    uint32_t start0 = chunk;
    if (stop != 0ull) {
      const uint32_t j = 63u - (uint32_t)__clzll(stop);
      start0 = chunk + 2u * j + (((ol >> j) & 1ull) ? 1u : 2u);
    }
    par[li0] = start0;
    par[li0 + 1u] = internal ? start0 : li0 + 1u;
    lcnt[li0] = 0u;
    lcnt[li0 + 1u] = 0u;
  }
  if (meas) wave_spins(me, mm, red);
  __syncthreads();
  if (meas && tid == 0) meas_combine<NTH>(red, meas_out);
#pragma unroll 1
  for (int it = 0; it < ITERS; ++it) {
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t lx = ql / (T / 2), ly0 = (ql % (T / 2)) * 2u;
    const uint32_t li0 = ql << 1, li1 = li0 + 1u;
    const uint32_t b0 = bits[li0], b1 = bits[li1];
    uint32_t req = 0u;
    if (lx + 1u < T) {
      const bool first = ly0 == 0u;  // no plaquette to the left
      const uint32_t lf = first ? li0 : li0 - 1u;
      const uint32_t left = bits[lf], left_up = bits[lf + T];
      const uint32_t h0 = (b0 >> 1) & 1u &
                          ~((left >> 1) & (left >> 2) & (left_up >> 2) & (first ? 0u : 1u));
      const uint32_t h1 = (b1 >> 1) & 1u & ~((b0 >> 1) & (b0 >> 2) & (bits[li0 + T] >> 2));
      req = h0 | (h1 << 1);
    }
    req |= ((b1 >> 2) & 1u & ((lane == 63 && ly0 + 1u != T - 1u) ? 1u : 0u)) << 2;
    while (__ballot(req != 0u)) {
      if (req) {
        const uint32_t k = (uint32_t)__builtin_ctz(req);
        req &= req - 1u;
        tunion(par, k == 0u ? li0 : li1, k == 0u ? li0 + T : k == 1u ? li1 + T : li1 + 1u);
      }
    }
  }
  __syncthreads();
#pragma unroll 1
  for (int it = 0; it < ITERS; ++it) {
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t lx = ql / (T / 2), ly0 = (ql % (T / 2)) * 2u;
    const uint32_t li0 = ql << 1, li1 = li0 + 1u;
    const uint32_t b0 = bits[li0], b1 = bits[li1];
    uint32_t r0 = li0, r1 = (b0 & 4u) ? li0 : li1;  // li1 shares li0's run when linked
    troot2(par, r0, r1);
    if (r0 == r1) {
      atomicAdd(&lcnt[r0], 0x20000u | ((b0 & 1u) + (b1 & 1u)));
    } else {
      atomicAdd(&lcnt[r0], 0x10000u | (b0 & 1u));
      atomicAdd(&lcnt[r1], 0x10000u | (b1 & 1u));
    }
    const uint32_t n0 = (x0 + lx) * L + y0 + ly0;
    t.lroot[base + n0] = (uint16_t)r0;
    t.lroot[base + n0 + 1u] = (uint16_t)r1;
  }
  __syncthreads();
  for (uint32_t li = tid; li < T * T; li += NTH) {
    if (par[li] != li) continue;
    const uint32_t n = (x0 + li / T) * L + y0 + li % T;
    t.gf[base + n] = gpack(n, 0, 0);
    t.lst[base + n] = lcnt[li];
    t.st[base + n] = 0ull;
  }
}

// ----------------------------------------------------------------- K2
template <int T>
__global__ void k2_merge(LatticeSet s, TiledScratch t) {
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const int L = s.L;
  const int nh = t.tiles_x * L, nv = t.tiles_y * L;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nh + nv) return;
  int x, y, horiz;
  if (i < nh) {
    x = min((i / L) * T + T - 1, L - 1);
    y = i % L;
    horiz = 1;
  } else {
    const int k = i - nh;
    y = min((k / L) * T + T - 1, L - 1);
    x = k % L;
    horiz = 0;
  }
  const size_t base = (size_t)lat * s.N;
  const uint32_t a = (uint32_t)(x * L + y);
  const uint32_t o = t.occ[base + a];
  if (!(o & (horiz ? 1u : 2u))) return;
  const uint32_t b = horiz ? ((x == L - 1) ? (uint32_t)y : a + L) : ((y == L - 1) ? a - y : a + 1);
  const int cx = (horiz && x == L - 1) ? 1 : 0;
  const int cy = (!horiz && y == L - 1) ? 1 : 0;
  uint64_t* uf = t.gf + base;
  // a site -> its local root has winding 0 (inside one tile)
  auto local_root = [&](uint32_t site, uint32_t sx, uint32_t sy) {
    const uint32_t r = t.lroot[base + site];
    return ((sx & ~(uint32_t)(T - 1)) + (r / T)) * (uint32_t)L + (sy & ~(uint32_t)(T - 1)) + (r % T);
  };
  const uint32_t bxs = horiz ? (x == L - 1 ? 0u : (uint32_t)x + 1u) : (uint32_t)x;
  const uint32_t bys = horiz ? (uint32_t)y : (y == L - 1 ? 0u : (uint32_t)y + 1u);
  const uint32_t la = local_root(a, (uint32_t)x, (uint32_t)y), lb = local_root(b, bxs, bys);
  // after a failed hook the walks go on from the stale roots: the windings
  // carried so far stay valid (positions inside a tree never change)
  uint32_t ra = la, rb = lb;
  int ax = 0, ay = 0, bx = 0, by = 0;
  uint32_t wbits = 0u;
  while (true) {
    gfind2(uf, ra, ax, ay, rb, bx, by);
    if (ra == rb) {
      wbits = ((cx + bx - ax) != 0 ? 1u : 0u) | ((cy + by - ay) != 0 ? 2u : 0u);
      break;
    }
    if (ra < rb) {
      if (atomicCAS((unsigned long long*)&uf[rb], (unsigned long long)rb,
                    (unsigned long long)gpack(ra, ax - bx - cx, ay - by - cy)) == rb)
        break;
    } else {
      if (atomicCAS((unsigned long long*)&uf[ra], (unsigned long long)ra,
                    (unsigned long long)gpack(rb, bx - ax + cx, by - ay + cy)) == ra)
        break;
    }
  }
  wave_or_wraps(&t.wrapf[2 * lat + (s.sweep_ctr[lat] & 1)], wbits);
}

// ----------------------------------------------------------------- K3
// One workgroup per tile: each tile-local root resolves its global root and
// that root's flip bit once (into an LDS table); every site then flips from
// the table, and each local root adds its counts to the global root.
template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k3_flip(LatticeSet s, TiledScratch t) {
  __shared__ uint32_t rmap[T * T];  // global root | flip << 31, at local roots
  constexpr int PER = T * T / NTH;
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const uint32_t L = (uint32_t)s.L, N = (uint32_t)s.N;
  const uint32_t tx = blockIdx.x % t.tiles_x, ty = blockIdx.x / t.tiles_x;
  const uint32_t x0 = tx * T, y0 = ty * T;
  const uint32_t w = min((uint32_t)T, L - x0), h = min((uint32_t)T, L - y0);
  const size_t base = (size_t)lat * N;
  // every load this thread needs is issued up front: the site's local root
  // and spin, and at local roots the forest entry, occ byte and local counts
  // (other sites load the tile's first entries instead: one cached line)
  const uint32_t dummy = x0 * L + y0;
  uint32_t lr[PER], lc[PER], oc[PER];
  uint64_t gv[PER];
  int sv[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t li = (uint32_t)(k * NTH + threadIdx.x), lx = li / T, ly = li % T;
    const bool in = lx < w && ly < h;
    const uint32_t n = in ? (x0 + lx) * L + y0 + ly : dummy;
    lr[k] = in ? (uint32_t)t.lroot[base + n] : 0xFFFFFFFFu;
    sv[k] = s.spins[base + n];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t li = (uint32_t)(k * NTH + threadIdx.x), lx = li / T, ly = li % T;
    const uint32_t n = lr[k] == li ? (x0 + lx) * L + y0 + ly : dummy;
    gv[k] = t.gf[base + n];
    oc[k] = t.occ[base + n];
    lc[k] = t.lst[base + n];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t li = (uint32_t)(k * NTH + threadIdx.x), lx = li / T, ly = li % T;
    if (lr[k] != li) continue;
    const uint32_t n = (x0 + lx) * L + y0 + ly;
    uint32_t r = gpar(gv[k]), o = oc[k];
    if (r != n) {  // merged across tiles: finish the walk, point n at the root
      r = groot(t.gf + base, r);
      t.gf[base + n] = gpack(r, 0, 0);
      o = t.occ[base + r];
    }
    rmap[li] = r | ((o >> 2) & 1u) << 31;
    // other tiles read only the flip bit of this byte, which stays as it is
    if (r == n) t.occ[base + n] = (uint8_t)(o | OCC_ROOT);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    if (lr[k] == 0xFFFFFFFFu) continue;
    const uint32_t li = (uint32_t)(k * NTH + threadIdx.x), lx = li / T, ly = li % T;
    const uint32_t n = (x0 + lx) * L + y0 + ly;
    const uint32_t m = rmap[lr[k]];
    const uint32_t r = m & 0x7FFFFFFFu, fl = m >> 31;
    if (fl) s.spins[base + n] = (int8_t)-sv[k];
    if (lr[k] == li) {
      const uint64_t size = lc[k] >> 16, np = lc[k] & 0xFFFFu;
      const uint64_t np_new = fl ? size - np : np;
      atomicAdd((unsigned long long*)&t.st[base + r], (unsigned long long)((size << 32) | np_new));
    }
    if (s.canon_out) s.canon_out[n] = (int32_t)r;
  }
}

// ----------------------------------------------------------------- K4
template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k4_measure(LatticeSet s, TiledScratch t) {
  // measures the settled state after the last sweep of a run (earlier sweeps
  // are measured inside the next sweep's K1)
  __shared__ int64_t red[NTH / 64][PARTS];
  const int lat = run_lat(s, blockIdx.y);
  if (t.pending[lat] == 0) return;
  const int L = s.L, N = s.N;
  const int tx = blockIdx.x % t.tiles_x, ty = blockIdx.x / t.tiles_x;
  const int x0 = tx * T, y0 = ty * T;
  const int w = min(T, L - x0), h = min(T, L - y0);
  const size_t base = (size_t)lat * N;
  const int8_t* sp = s.spins + base;
  const uint8_t* jb = s.jbits + base;
  TileMeas ms = {0, 0, 0, 0ull, 0ull, 0u, 0u};
  for (int li = threadIdx.x; li < T * T; li += NTH) {
    const int lx = li / T, ly = li % T;
    if (lx >= w || ly >= h) continue;
    const int x = x0 + lx, y = y0 + ly;
    const uint32_t n = (uint32_t)(x * L + y);
    const uint32_t nr = (x == L - 1) ? (uint32_t)y : n + L;
    const uint32_t nu = (y == L - 1) ? n - y : n + 1;
    ms.e += bond_energy(jb[n], sp[n], sp[nr], sp[nu]);
    ms.m += sp[n];
    if (t.occ[base + n] & OCC_ROOT) meas_root(ms, t.st[base + n], n);
  }
  meas_publish<NTH>(ms, red, t.part + ((size_t)lat * t.tiles_x * t.tiles_y + blockIdx.x) * PARTS);
}

// ----------------------------------------------------------------- K5
// Fold the tile partials of sweep gm into its bin (single thread).
__device__ void finalize_sweep(const LatticeSet& s, const TiledScratch& t, int lat, int64_t gm) {
  const int nt = t.tiles_x * t.tiles_y;
  const int64_t* p = t.part + (size_t)lat * nt * PARTS;
  TileMeas a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
  int64_t e = 0, m = 0, ncl = 0;
  for (int i = 0; i < nt; ++i) {
    e += p[i * PARTS + 0];
    m += p[i * PARTS + 1];
    ncl += p[i * PARTS + 2];
    top2_insert(a, (uint64_t)p[i * PARTS + 3], (uint32_t)p[i * PARTS + 5]);
    top2_insert(a, (uint64_t)p[i * PARTS + 4], (uint32_t)p[i * PARTS + 6]);
  }
  rbim_sweep_stats st;
  st.energy = e;
  st.magnetization = m;
  st.n_clusters = ncl;
  const int64_t s1 = (int64_t)(a.k1 >> 32), mc1 = 2 * (int64_t)a.p1 - s1;
  st.largest_size = s1;
  st.largest_abs_mag = mc1 < 0 ? -mc1 : mc1;
  if (a.k2 != 0ull) {
    const int64_t s2 = (int64_t)(a.k2 >> 32), mc2 = 2 * (int64_t)a.p2 - s2;
    st.second_size = s2;
    st.second_abs_mag = mc2 < 0 ? -mc2 : mc2;
  } else {
    st.second_size = 0;
    st.second_abs_mag = 0;
  }
  uint32_t* wf = &t.wrapf[2 * lat + (gm & 1)];
  st.wrap_x = (int32_t)(*wf & 1u);
  st.wrap_y = (int32_t)((*wf >> 1) & 1u);
  *wf = 0u;
  const int64_t bin = (gm - lat_therm(s, lat)) / lat_spb(s, lat);
  acc_sweep(s.acc + ((size_t)lat * s.n_bins + bin) * RBIM_ACC_SLOTS, st);
  if (s.last) s.last[lat] = st;
  series_put(s, lat, gm, st);
}

// After K1..K3 of sweep g: finalize sweep g-1 (measured in K1 of g), mark
// sweep g pending (or drop its wrap bits) and advance the counter.
__global__ void k5_finalize(LatticeSet s, TiledScratch t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n_run) return;
  const int lat = run_lat(s, i);
  const int64_t g = s.sweep_ctr[lat];
  if (g >= lat_total(s, lat)) return;
  if (t.pending[lat]) finalize_sweep(s, t, lat, g - 1);
  const bool measured = g >= lat_therm(s, lat);
  t.pending[lat] = measured ? 1 : 0;
  if (!measured) t.wrapf[2 * lat + (g & 1)] = 0u;
  s.sweep_ctr[lat] = g + 1;
}

// End of a run: finalize the pending sweep measured by k4_measure.
__global__ void k5_flush(LatticeSet s, TiledScratch t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n_run) return;
  const int lat = run_lat(s, i);
  if (t.pending[lat] == 0) return;
  finalize_sweep(s, t, lat, s.sweep_ctr[lat] - 1);
  t.pending[lat] = 0;
}

// ------------------------------------------------------ one-lattice helpers
// labels[n] = rank of canon[n] among roots (roots are the sites with
// canon[n] == n), i.e. the reference's min-site-order numbering.
__global__ __launch_bounds__(1024) void label_rank_kernel(const int32_t* canon, int32_t N,
                                                          int32_t* labels, int32_t* rank,
                                                          int32_t* n_clusters) {
  __shared__ int32_t sums[1024];
  const int tid = threadIdx.x;
  const int chunk = (N + 1023) / 1024;
  const int lo = min(N, tid * chunk), hi = min(N, lo + chunk);
  int c = 0;
  for (int n = lo; n < hi; ++n) c += canon[n] == n;
  sums[tid] = c;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = tid >= off ? sums[tid - off] : 0;
    __syncthreads();
    sums[tid] += v;
    __syncthreads();
  }
  int r = sums[tid] - c;
  for (int n = lo; n < hi; ++n)
    if (canon[n] == n) rank[n] = r++;
  if (tid == 1023) *n_clusters = sums[1023];
  __syncthreads();
  __threadfence_block();
  for (int n = tid; n < N; n += 1024) labels[n] = rank[canon[n]];
}

__global__ void cluster_stats_kernel(const int32_t* labels, const int8_t* spins, int64_t n,
                                     unsigned long long* sizes, unsigned long long* mags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&sizes[labels[i]], 1ull);
  atomicAdd(&mags[labels[i]], (unsigned long long)(long long)spins[i]);  // two's complement
}

__global__ void energy_mag_kernel(const int8_t* spins, const uint8_t* jb, int32_t L,
                                  unsigned long long* em) {
  const int N = L * L;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  long long e = 0, m = 0;
  if (n < N) {
    const int x = n / L, y = n % L;
    const int nr = (x == L - 1) ? y : n + L;
    const int nu = (y == L - 1) ? n - y : n + 1;
    const int s0 = spins[n];
    e -= (long long)((jb[n] & 1u) ? -1 : 1) * s0 * spins[nr];
    e -= (long long)((jb[n] & 2u) ? -1 : 1) * s0 * spins[nu];
    m = s0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    m += __shfl_xor(m, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&em[0], (unsigned long long)e);
    atomicAdd(&em[1], (unsigned long long)m);
  }
}

}  // namespace

int tiled_kernels_per_sweep() { return 4; }

int tiled_tile_size(int L, int pref) {
  if (pref == 32) return 32;
  if (pref == 64 && L % 64 == 0) return 64;
  return (L % 64 == 0 && L >= 128) ? 64 : 32;
}

template <int T, int NTH>
hipError_t launch_tiled_t(const LatticeSet& s, const TiledScratch& t, hipStream_t st) {
  const int ntiles = t.tiles_x * t.tiles_y;
  if (s.L % T == 0)
    hipLaunchKernelGGL((k1_tile_full<T, NTH>), dim3(ntiles, s.n_run), dim3(NTH), 0, st, s, t);
  else
    hipLaunchKernelGGL((k1_tile<T, NTH>), dim3(ntiles, s.n_run), dim3(NTH), 0, st, s, t);
  const int nb = (t.tiles_x + t.tiles_y) * s.L;
  hipLaunchKernelGGL((k2_merge<T>), dim3((nb + 255) / 256, s.n_run), dim3(256), 0, st, s, t);
  hipLaunchKernelGGL((k3_flip<T, NTH>), dim3(ntiles, s.n_run), dim3(NTH), 0, st, s, t);
  hipLaunchKernelGGL(k5_finalize, dim3((s.n_run + 63) / 64), dim3(64), 0, st, s, t);
  return hipGetLastError();
}

template <int T, int NTH>
hipError_t launch_flush_t(const LatticeSet& s, const TiledScratch& t, hipStream_t st) {
  hipLaunchKernelGGL((k4_measure<T, NTH>), dim3(t.tiles_x * t.tiles_y, s.n_run), dim3(NTH), 0, st, s, t);
  hipLaunchKernelGGL(k5_flush, dim3((s.n_run + 63) / 64), dim3(64), 0, st, s, t);
  return hipGetLastError();
}

hipError_t launch_tiled_sweep(const LatticeSet& s, const TiledScratch& t, hipStream_t st) {
  if (t.tile == 64) return launch_tiled_t<64, 512>(s, t, st);
  return launch_tiled_t<32, 256>(s, t, st);
}

hipError_t launch_tiled_flush(const LatticeSet& s, const TiledScratch& t, hipStream_t st) {
  if (t.tile == 64) return launch_flush_t<64, 512>(s, t, st);
  return launch_flush_t<32, 256>(s, t, st);
}

hipError_t launch_label_rank(const int32_t* canon, int32_t N, int32_t* labels, int32_t* rank_tmp,
                             int32_t* n_clusters, hipStream_t st) {
  hipLaunchKernelGGL(label_rank_kernel, dim3(1), dim3(1024), 0, st, canon, N, labels, rank_tmp,
                     n_clusters);
  return hipGetLastError();
}

hipError_t launch_cluster_stats(const int32_t* labels, const int8_t* spins, int64_t n,
                                unsigned long long* sizes, unsigned long long* mags, hipStream_t st) {
  hipLaunchKernelGGL(cluster_stats_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     labels, spins, n, sizes, mags);
  return hipGetLastError();
}

hipError_t launch_energy_mag(const int8_t* spins, const uint8_t* jbits, int32_t L,
                             unsigned long long* em, hipStream_t st) {
  const int N = L * L;
  hipLaunchKernelGGL(energy_mag_kernel, dim3((N + 255) / 256), dim3(256), 0, st, spins, jbits, L,
                     em);
  return hipGetLastError();
}
