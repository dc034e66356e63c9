// LDS-resident Swendsen-Wang engine for lattices up to 128 x 128 on gfx950.
//
// One workgroup owns one lattice for a whole block of sweeps: spins, couplings,
// the union-find forest and the per-cluster counters all live in LDS, so a
// sweep touches HBM only through the moment accumulators at bin boundaries.
// Per sweep (src/rbim/swendsen_wang.py:103-172):
//   A  activation   Philox draws for each site pair; satisfied bonds are
//                   occupied when draw < thr31 (clusters.py:48-79)
//   B  union        hook-and-compress on LDS atomics; the forest word packs
//                   parent (14 bits) with the winding numbers (wx, wy, 9 bits
//                   each) of the tree path to the parent.  A bond whose ends
//                   already share a root is a non-tree bond; its loop winding
//                   c + w(b) - w(a) decides wrap_x / wrap_y exactly
//                   (_unionfind.py:101-124)
//   CD flip         root = min site of the cluster (larger roots always hook
//                   under smaller ones); s ^= flipbit(root); per-root counter
//                   += (1 << 16) | [s = +1]  (swendsen_wang.py:140-141,
//                   clusters.py:98-118)
//   E  measure      E, M, n_clusters and the two largest clusters (size desc,
//                   min site asc) by wavefront reductions; one lane folds the
//                   sweep into its bin's moment sums (simulation.py:212-236).
#include "engine.h"

namespace {

constexpr uint32_t PMASK = 0x3FFFu;  // parent field

__device__ __forceinline__ uint32_t uf_pack(uint32_t p, int wx, int wy) {
  return p | (((uint32_t)wx & 0x1FFu) << 14) | ((uint32_t)wy << 23);
}
__device__ __forceinline__ int uf_wx(uint32_t v) { return ((int)(v << 9)) >> 23; }
__device__ __forceinline__ int uf_wy(uint32_t v) { return ((int)v) >> 23; }

// Root of x with the winding of the tree path x -> root; halves the path.
// Stores only touch non-root words, which hooks (CAS on roots) never target.
__device__ __forceinline__ uint32_t lds_find(uint32_t* uf, uint32_t x, int& wx, int& wy) {
  wx = 0;
  wy = 0;
  uint32_t v = uf[x];
  while (true) {
    const uint32_t p = v & PMASK;
    if (p == x) return x;
    const uint32_t vp = uf[p];
    const uint32_t gp = vp & PMASK;
    if (gp == p) {
      wx += uf_wx(v);
      wy += uf_wy(v);
      return p;
    }
    const int hx = uf_wx(v) + uf_wx(vp);
    const int hy = uf_wy(v) + uf_wy(vp);
    uf[x] = uf_pack(gp, hx, hy);
    wx += hx;
    wy += hy;
    x = gp;
    v = uf[gp];
  }
}

__device__ __forceinline__ uint32_t lds_root(uint32_t* uf, uint32_t x) {
  uint32_t v = uf[x];
  while ((v & PMASK) != x) {
    const uint32_t p = v & PMASK;
    const uint32_t vp = uf[p];
    if ((vp & PMASK) == p) return p;
    x = vp & PMASK;
    v = uf[x];
  }
  return x;
}

// Union of the occupied bond a -> b that crosses the periodic boundary
// (cx, cy) times.  Returns wrap bits (1: x, 2: y) when the bond closes a loop.
__device__ __forceinline__ uint32_t lds_union(uint32_t* uf, uint32_t a, uint32_t b, int cx,
                                              int cy) {
  while (true) {
    int ax, ay, bx, by;
    const uint32_t ra = lds_find(uf, a, ax, ay);
    const uint32_t rb = lds_find(uf, b, bx, by);
    if (ra == rb) {
      return ((cx + bx - ax) != 0 ? 1u : 0u) | ((cy + by - ay) != 0 ? 2u : 0u);
    }
    if (ra < rb) {
      if (atomicCAS(&uf[rb], rb, uf_pack(ra, ax - bx - cx, ay - by - cy)) == rb) return 0u;
    } else {
      if (atomicCAS(&uf[ra], ra, uf_pack(rb, bx - ax + cx, by - ay + cy)) == ra) return 0u;
    }
  }
}

__device__ __forceinline__ void top2_merge(uint32_t& k1, uint32_t& k2, uint32_t o1, uint32_t o2) {
  const uint32_t hi = max(k1, o1);
  const uint32_t lo = max(min(k1, o1), max(k2, o2));
  k1 = hi;
  k2 = lo;
}

template <int NT>
__global__ __launch_bounds__(NT) void lds_sweep_kernel(LatticeSet s, int32_t sweeps) {
  constexpr int NW = NT / 64;
  extern __shared__ uint32_t smem[];
  __shared__ int32_t red_i[NW][3];
  __shared__ uint32_t red_k[NW][2];
  __shared__ uint32_t wrapflag;
  __shared__ uint64_t a[RBIM_ACC_SLOTS];  // running sums of the open bin (lane 0)

  const int N = s.N;
  const uint32_t L = (uint32_t)s.L;
  const uint32_t magic = s.magic;
  uint32_t* uf = smem;
  uint32_t* cnt = smem + N;
  uint8_t* sb = reinterpret_cast<uint8_t*>(smem + 2 * N);

  const int lat = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;

  int8_t* gsp = s.spins + (size_t)lat * N;
  const uint8_t* gj = s.jbits + (size_t)lat * N;
  for (int n = tid; n < N; n += NT) sb[n] = (uint8_t)((gsp[n] > 0 ? 1u : 0u) | ((uint32_t)gj[n] & 3u) << 1);
  if (tid == 0) wrapflag = 0u;

  const int64_t g0 = s.sweep_ctr[lat];
  const int64_t gend = min(g0 + (int64_t)sweeps, s.total);
  const uint32_t thr = s.thr31[lat];
  const uint64_t seed = s.seed[lat];
  const uint32_t lid = s.lat_id[lat];
  const bool given = s.given_occ != nullptr;
  uint64_t* gacc = s.acc + (size_t)lat * s.n_bins * RBIM_ACC_SLOTS;

  int64_t cur_bin = -1;
  __syncthreads();

  for (int64_t g = g0; g < gend; ++g) {
    // ---------------------------------------------------------------- A
    const int npairs = (N + 1) >> 1;
    for (int q = tid; q < npairs; q += NT) {
      rbim_u32x4 r = {0u, 0u, 0u, 0u};
      if (!given) r = rbim_sweep_draws(seed, lid, (uint32_t)g, (uint32_t)q);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t n = 2u * q + j;
        if (n >= (uint32_t)N) break;
        const uint32_t x = site_x(n, magic);
        const uint32_t y = n - x * L;
        const uint32_t nr = (x == L - 1) ? y : n + L;
        const uint32_t nu = (y == L - 1) ? n - y : n + 1;
        const uint32_t b = sb[n];
        const uint32_t sp = b & 1u;
        uint32_t oh, ov, fl;
        if (given) {
          const uint32_t o = s.given_occ[n];
          oh = o & 1u;
          ov = (o >> 1) & 1u;
          fl = s.given_flip[n] & 1u;
        } else {
          const uint32_t sat_h = ~(sp ^ sb[nr] ^ (b >> 1)) & 1u;
          const uint32_t sat_v = ~(sp ^ sb[nu] ^ (b >> 2)) & 1u;
          const uint32_t dh = j ? r.z : r.x;
          const uint32_t dv = j ? r.w : r.y;
          oh = sat_h & ((dh >> 1) < thr ? 1u : 0u);
          ov = sat_v & ((dv >> 1) < thr ? 1u : 0u);
          fl = dh & 1u;
        }
        sb[n] = (uint8_t)((b & 7u) | (oh << 3) | (ov << 4) | (fl << 5));
        uf[n] = n;
        cnt[n] = 0u;
      }
    }
    __syncthreads();
    // ---------------------------------------------------------------- B
    uint32_t wraps = 0u;
    for (int n = tid; n < N; n += NT) {
      const uint32_t b = sb[n];
      if ((b & 0x18u) == 0u) continue;
      const uint32_t x = site_x((uint32_t)n, magic);
      const uint32_t y = (uint32_t)n - x * L;
      if (b & 0x08u) wraps |= lds_union(uf, n, (x == L - 1) ? y : n + L, x == L - 1 ? 1 : 0, 0);
      if (b & 0x10u) wraps |= lds_union(uf, n, (y == L - 1) ? n - y : n + 1, 0, y == L - 1 ? 1 : 0);
    }
    if (wraps) atomicOr(&wrapflag, wraps);
    __syncthreads();
    // --------------------------------------------------------------- CD
    for (int n = tid; n < N; n += NT) {
      const uint32_t r = lds_root(uf, (uint32_t)n);
      const uint32_t b = sb[n];
      const uint32_t ns = (b ^ (sb[r] >> 5)) & 1u;
      sb[n] = (uint8_t)((b & ~1u) | ns);
      atomicAdd(&cnt[r], 0x10000u | ns);
      if (s.canon_out) s.canon_out[n] = (int32_t)r;
    }
    __syncthreads();
    const bool measure = g >= s.n_therm;
    if (!measure) {
      if (tid == 0) wrapflag = 0u;
      continue;  // the barrier at the top of A orders this reset
    }
    // ---------------------------------------------------------------- E
    int32_t e = 0, m = 0, ncl = 0;
    uint32_t k1 = 0u, k2 = 0u;
    for (int n = tid; n < N; n += NT) {
      const uint32_t x = site_x((uint32_t)n, magic);
      const uint32_t y = (uint32_t)n - x * L;
      const uint32_t b = sb[n];
      const uint32_t sp = b & 1u;
      const uint32_t sr = sb[(x == L - 1) ? y : n + L];
      const uint32_t su = sb[(y == L - 1) ? n - y : n + 1];
      const uint32_t sat_h = ~(sp ^ sr ^ (b >> 1)) & 1u;
      const uint32_t sat_v = ~(sp ^ su ^ (b >> 2)) & 1u;
      e += 2 - 2 * (int32_t)(sat_h + sat_v);
      m += 2 * (int32_t)sp - 1;
      if ((uf[n] & PMASK) == (uint32_t)n) {
        ++ncl;
        const uint32_t key = ((cnt[n] >> 16) << 14) | (PMASK - (uint32_t)n);
        top2_merge(k1, k2, key, 0u);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      e += __shfl_xor(e, off);
      m += __shfl_xor(m, off);
      ncl += __shfl_xor(ncl, off);
      const uint32_t o1 = __shfl_xor(k1, off);
      const uint32_t o2 = __shfl_xor(k2, off);
      top2_merge(k1, k2, o1, o2);
    }
    if (lane == 0) {
      red_i[wave][0] = e;
      red_i[wave][1] = m;
      red_i[wave][2] = ncl;
      red_k[wave][0] = k1;
      red_k[wave][1] = k2;
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < NW; ++w) {
        e += red_i[w][0];
        m += red_i[w][1];
        ncl += red_i[w][2];
        top2_merge(k1, k2, red_k[w][0], red_k[w][1]);
      }
      rbim_sweep_stats st;
      st.energy = e;
      st.magnetization = m;
      st.n_clusters = ncl;
      const uint32_t r1 = PMASK - (k1 & PMASK);
      const int32_t s1 = (int32_t)(k1 >> 14);
      const int32_t mc1 = 2 * (int32_t)(cnt[r1] & 0xFFFFu) - s1;
      st.largest_size = s1;
      st.largest_abs_mag = mc1 < 0 ? -mc1 : mc1;
      if (k2 != 0u) {
        const uint32_t r2 = PMASK - (k2 & PMASK);
        const int32_t s2 = (int32_t)(k2 >> 14);
        const int32_t mc2 = 2 * (int32_t)(cnt[r2] & 0xFFFFu) - s2;
        st.second_size = s2;
        st.second_abs_mag = mc2 < 0 ? -mc2 : mc2;
      } else {
        st.second_size = 0;
        st.second_abs_mag = 0;
      }
      st.wrap_x = (int32_t)(wrapflag & 1u);
      st.wrap_y = (int32_t)((wrapflag >> 1) & 1u);
      wrapflag = 0u;
      const int64_t bin = (g - s.n_therm) / s.spb;
      if (bin != cur_bin) {
        if (cur_bin >= 0) {
          for (int k = 0; k < RBIM_ACC_SLOTS; ++k) gacc[cur_bin * RBIM_ACC_SLOTS + k] = a[k];
        }
        for (int k = 0; k < RBIM_ACC_SLOTS; ++k) a[k] = gacc[bin * RBIM_ACC_SLOTS + k];
        cur_bin = bin;
      }
      acc_sweep(a, st);
      if (s.last) s.last[lat] = st;
    }
    __syncthreads();
  }

  if (tid == 0) {
    if (cur_bin >= 0) {
      for (int k = 0; k < RBIM_ACC_SLOTS; ++k) gacc[cur_bin * RBIM_ACC_SLOTS + k] = a[k];
    }
    s.sweep_ctr[lat] = gend > g0 ? gend : g0;
  }
  for (int n = tid; n < N; n += NT) gsp[n] = (sb[n] & 1u) ? (int8_t)1 : (int8_t)-1;
}

__global__ void init_spins_kernel(LatticeSet s, int start) {
  const int lat = blockIdx.y;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= s.N) return;
  s.spins[(size_t)lat * s.N + n] =
      start == RBIM_START_COLD ? (int8_t)1 : (int8_t)rbim_init_spin(s.seed[lat], s.lat_id[lat], (uint32_t)n);
}

template <int NT>
hipError_t launch_lds_t(const LatticeSet& s, int sweeps, hipStream_t st) {
  const size_t bytes = lds_bytes(s.N);
  hipError_t e = hipFuncSetAttribute((const void*)lds_sweep_kernel<NT>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(lds_sweep_kernel<NT>, dim3(s.n_lat), dim3(NT), bytes, st, s, sweeps);
  return hipGetLastError();
}

}  // namespace

size_t lds_bytes(int N) {
  return (size_t)N * 8u + (((size_t)N + 15u) & ~(size_t)15u);
}

hipError_t launch_lds_sweeps(const LatticeSet& s, int threads, int sweeps, hipStream_t st) {
  switch (threads) {
    case 64: return launch_lds_t<64>(s, sweeps, st);
    case 128: return launch_lds_t<128>(s, sweeps, st);
    case 256: return launch_lds_t<256>(s, sweeps, st);
    case 512: return launch_lds_t<512>(s, sweeps, st);
    case 1024: return launch_lds_t<1024>(s, sweeps, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_init_spins(const LatticeSet& s, int start, hipStream_t st) {
  dim3 grid((s.N + 255) / 256, s.n_lat);
  hipLaunchKernelGGL(init_spins_kernel, grid, dim3(256), 0, st, s, start);
  return hipGetLastError();
}
