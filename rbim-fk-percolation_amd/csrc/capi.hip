// C ABI of librbim_hip.so (declared in include/rbim_hip.h): contexts, the
// one-lattice entry points, and batches of lattices grouped by (L, engine).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"

struct rbim_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // guards err and the context stream (batches are created from several threads)
  std::string err;
};

// Device memory of rbim_sweep_noalloc: every buffer of its one-lattice sweep
// is carved out of this array, so that entry point never allocates.  Each GPU
// holds its own copy from module load until the process ends.  Host code
// never uses the array's name as an address: rbim_open resolves the device
// address once per device (hipGetSymbolAddress) into g_scratch.
static __device__ __align__(256) uint8_t scratch[RBIM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

// Per device: the resolved address of `scratch` and the lock that serialises
// the calls sharing it (they would race on the array otherwise).
struct ScratchSlot {
  std::mutex mu;
  uint8_t* ptr = nullptr;
};
ScratchSlot g_scratch[kMaxDevices];

// With `device` current: resolve the device address of `scratch` once.
hipError_t resolve_scratch(int device) {
  ScratchSlot& sl = g_scratch[device];
  std::lock_guard<std::mutex> lk(sl.mu);
  if (sl.ptr) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) sl.ptr = static_cast<uint8_t*>(p);
  return e;
}

std::mutex g_err_mu;
std::string g_err;

int set_err(rbim_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->err = buf;
  }
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = buf;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                   \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return set_err((ctx), e_ == hipErrorOutOfMemory ? RBIM_E_OOM : RBIM_E_HIP,             \
                     "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__,        \
                     __LINE__);                                                              \
  } while (0)

// One byte per site: bit 0 J_h < 0, bit 1 J_v < 0 (non-aliasing, so the loop
// vectorizes).
void pack_jbits(const int8_t* __restrict jh, const int8_t* __restrict jv, uint8_t* __restrict out,
                int n) {
  for (int i = 0; i < n; ++i) out[i] = (uint8_t)((jh[i] < 0 ? 1 : 0) | (jv[i] < 0 ? 2 : 0));
}

uint32_t magic_for(int L) {
  return (uint32_t)(((1ull << 32) + (uint64_t)L - 1) / (uint64_t)L);
}

// Failure injection (rbim_inject_alloc_failure): while >= 0, the number of
// device allocations that still succeed before one fails as out of memory.
std::atomic<int64_t> g_alloc_countdown{-1};

bool injected_alloc_failure() {
  int64_t c = g_alloc_countdown.load();
  while (c >= 0)
    if (g_alloc_countdown.compare_exchange_weak(c, c == 0 ? -1 : c - 1)) return c == 0;
  return false;
}

// Device allocations of one group, released together.
struct DevArena {
  std::vector<void*> ptrs;
  size_t bytes = 0;
  hipError_t alloc(void** p, size_t n) {
    if (n == 0) n = 16;
    if (injected_alloc_failure()) return hipErrorOutOfMemory;
    hipError_t e = hipMalloc(p, n);
    if (e == hipSuccess) {
      ptrs.push_back(*p);
      bytes += n;
    }
    return e;
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    bytes = 0;
  }
};

struct Group {
  int L = 0, N = 0, n_lat = 0;
  bool lds = false;
  bool fused = false;  // tiled group on the two-kernel edge-slot sweep
  int tile_pref = 0;   // requested tile side (0: by L)
  int threads = 256;
  int per_launch = 1;  // LDS sweeps per launch, or tiled sweeps per graph
  std::vector<int> idx;
  LatticeSet set{};
  TiledScratch tile{};
  FusedScratch fz{};
  DevArena arena;
  hipGraphExec_t graph = nullptr;
  hipStream_t stream = nullptr;  // batch groups run concurrently, one stream each
  hipEvent_t done = nullptr;     // recorded after the group's work of a run
};

size_t group_bytes(int L, int n_lat, bool lds, bool fused, int n_bins, int series, int tile_pref) {
  const size_t N = (size_t)L * L;
  size_t b = N * n_lat * 2 + (size_t)n_lat * (4 + 8 + 4 + 8 + sizeof(rbim_sweep_stats)) +
             (size_t)n_lat * n_bins * RBIM_ACC_SLOTS * 8 +
             (size_t)n_lat * (size_t)std::max(series, 0) * sizeof(rbim_sweep_stats);
  if (fused) {
    const int T = fused_tile_size(L, tile_pref);
    const size_t nt = (size_t)fused_tiles_per_side(L, T), tiles = nt * nt, ns = tiles * 4 * T;
    b += N * n_lat * 4 + (size_t)n_lat * (tiles * T * T / 4 + tiles * (2 * T / 64) * 8 + ns / 8 +
                                                 tiles * (2 * 5 + 2) * 8 + 2 * 6 * 8 + 8 + 4);
    if (fused_merge_in_hbm(L, T))
      b += (size_t)n_lat * (ns * (8 + 8 + 4) + (ns + 255) / 256 * 5 * 8 + 4);
    else
      b += (size_t)n_lat * ns * 8;
  } else if (!lds) {
    const size_t tx = (L + 31) / 32;
    b += N * n_lat * (1 + 2 + 8 + 4 + 8) + (size_t)n_lat * tx * tx * 7 * 8 + (size_t)n_lat * 12;
  }
  return b;
}

// LDS kernel workgroup size for a full batch of side L.  Four wavefronts up to
// 64 x 64: the kernel holds five per SIMD, so a CU keeps five such lattices at
// once (four at L = 64 by LDS), where eight-wavefront workgroups fit two and
// leave a 1280-lattice batch in three rounds (1280 lattices, 256 against 512
// threads: L = 40 9.4e10 / 7.2e10, 48 1.08e11 / 8.2e10, 56 1.10e11 / 8.9e10,
// 64 1.37e11 / 1.22e11 site-updates/s).
int default_threads(int L) {
  const int N = L * L;
  if (N <= 256) return 64;
  if (N <= 4096) return 256;
  return 1024;
}

// Workgroup sizes of a batch's LDS groups (threads not given): the per-L
// default, doubled -- fewer pairs per lane, more wavefronts in flight, up to
// one pair per lane and 1024 threads -- while the whole batch would give the
// SIMDs fewer than four wavefronts each, for lattices up to 64 x 64 (128
// lattices, 512 against 256 threads: L = 48 2.1e10 / 1.4e10, 64 3.2e10 /
// 2.1e10; beyond, the larger workgroups measured no faster).
void default_lds_threads(std::vector<Group>& groups, int n_cu) {
  const long target = 16l * (n_cu > 0 ? n_cu : 256);  // wavefronts
  long waves = 0;
  for (Group& g : groups)
    if (g.lds) {
      g.threads = default_threads(g.L);
      waves += (long)g.idx.size() * (g.threads / 64);
    }
  bool grew = true;
  while (waves < target && grew) {
    grew = false;
    for (Group& g : groups)
      if (g.lds && g.N <= 4096 && 2 * g.threads <= g.N / 2 && 2 * g.threads <= 1024) {
        waves += (long)g.idx.size() * (g.threads / 64);
        g.threads *= 2;
        grew = true;
      }
  }
}

int default_lds_sweeps(int L) {
  const int N = L * L;
  return N <= 1024 ? 2000 : N <= 4096 ? 500 : 200;
}

// Allocates and fills the device side of a group.  jh/jv/spins0 are the
// group's host slices (concatenated); thr/seed/lid per lattice.  With
// `empty` (slot pools) the host vectors are not read: every array is
// allocated and zeroed, and the slots are filled later on the device.
int group_alloc(rbim_ctx* ctx, Group& g, const std::vector<uint8_t>& jbits,
                const std::vector<uint32_t>& thr, const std::vector<uint64_t>& seed,
                const std::vector<uint32_t>& lid, const int8_t* spins0, int start,
                int64_t n_therm, int64_t spb, int64_t n_bins, int64_t g0, int series = 0,
                bool empty = false) {
  const size_t NN = (size_t)g.N * g.n_lat;
  LatticeSet& s = g.set;
  s.L = g.L;
  s.N = g.N;
  s.n_lat = g.n_lat;
  s.n_run = g.n_lat;
  s.sched = nullptr;
  s.lat_map = nullptr;
  s.magic = magic_for(g.L);
  s.n_therm = n_therm;
  s.spb = spb;
  s.n_bins = n_bins;
  s.total = n_therm + n_bins * spb;
  void* p;
  // host data into a fresh allocation, or zeros when the group starts empty
  auto upload = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    return empty ? hipMemsetAsync(dst, 0, bytes, ctx->stream)
                 : hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  };
  std::vector<int64_t> ctr0(empty ? 0 : g.n_lat, g0);
  HIP_TRY(ctx, g.arena.alloc(&p, NN)); s.spins = (int8_t*)p;
  HIP_TRY(ctx, g.arena.alloc(&p, NN)); s.jbits = (const uint8_t*)p;
  HIP_TRY(ctx, upload(p, jbits.data(), NN));
  HIP_TRY(ctx, g.arena.alloc(&p, 4 * g.n_lat)); s.thr31 = (const uint32_t*)p;
  HIP_TRY(ctx, upload(p, thr.data(), 4 * g.n_lat));
  HIP_TRY(ctx, g.arena.alloc(&p, 8 * g.n_lat)); s.seed = (const uint64_t*)p;
  HIP_TRY(ctx, upload(p, seed.data(), 8 * g.n_lat));
  HIP_TRY(ctx, g.arena.alloc(&p, 4 * g.n_lat)); s.lat_id = (const uint32_t*)p;
  HIP_TRY(ctx, upload(p, lid.data(), 4 * g.n_lat));
  HIP_TRY(ctx, g.arena.alloc(&p, 8 * g.n_lat)); s.sweep_ctr = (int64_t*)p;
  HIP_TRY(ctx, upload(p, ctr0.data(), 8 * g.n_lat));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const size_t accb = (size_t)g.n_lat * n_bins * RBIM_ACC_SLOTS * 8;
  HIP_TRY(ctx, g.arena.alloc(&p, accb)); s.acc = (uint64_t*)p;
  HIP_TRY(ctx, hipMemsetAsync(p, 0, accb, ctx->stream));
  HIP_TRY(ctx, g.arena.alloc(&p, sizeof(rbim_sweep_stats) * g.n_lat)); s.last = (rbim_sweep_stats*)p;
  HIP_TRY(ctx, hipMemsetAsync(p, 0, sizeof(rbim_sweep_stats) * g.n_lat, ctx->stream));
  s.series = nullptr;
  s.series_cap = series;
  if (series > 0) {
    const size_t sb = (size_t)g.n_lat * series * sizeof(rbim_sweep_stats);
    HIP_TRY(ctx, g.arena.alloc(&p, sb)); s.series = (rbim_sweep_stats*)p;
    HIP_TRY(ctx, hipMemsetAsync(p, 0, sb, ctx->stream));
  }
  if (g.fused) {
    FusedScratch& f = g.fz;
    f.tile = fused_tile_size(g.L, g.tile_pref);
    f.tiles_x = fused_tiles_per_side(g.L, f.tile);
    const size_t tiles = (size_t)f.tiles_x * f.tiles_x;
    f.nslots = (int32_t)(tiles * 4 * f.tile);
    const size_t nl = (size_t)g.n_lat;
    auto zalloc = [&](void** q, size_t bytes) -> hipError_t {
      hipError_t e = g.arena.alloc(q, bytes);
      return e == hipSuccess ? hipMemsetAsync(*q, 0, bytes, ctx->stream) : e;
    };
    HIP_TRY(ctx, zalloc(&p, 2 * NN * 2)); f.st = (uint16_t*)p;
    HIP_TRY(ctx, zalloc(&p, nl * (size_t)f.tiles_x * f.tiles_x * f.tile * f.tile / 4));
    f.jt = (uint32_t*)p;
    f.big = fused_merge_in_hbm(g.L, f.tile) ? 1 : 0;
    if (f.big) {
      HIP_TRY(ctx, zalloc(&p, nl * f.nslots * 8)); f.guf = (uint64_t*)p;
      HIP_TRY(ctx, zalloc(&p, nl * f.nslots * 8)); f.gcnt = (uint64_t*)p;
      HIP_TRY(ctx, zalloc(&p, nl * f.nslots * 4)); f.gmn = (uint32_t*)p;
      HIP_TRY(ctx, zalloc(&p, nl * ((f.nslots + 255) / 256) * 5 * 8)); f.gpart = (int64_t*)p;
      HIP_TRY(ctx, zalloc(&p, nl * 4)); f.wrapw = (uint32_t*)p;
    } else {
      HIP_TRY(ctx, zalloc(&p, nl * f.nslots * 8)); f.slots = (uint64_t*)p;
    }
    HIP_TRY(ctx, zalloc(&p, nl * tiles * (2 * f.tile / 64) * 8)); f.xbits = (uint64_t*)p;
    HIP_TRY(ctx, zalloc(&p, nl * f.nslots / 8)); f.flipbits = (uint32_t*)p;
    HIP_TRY(ctx, zalloc(&p, 2 * nl * tiles * 5 * 8)); f.part_int = (int64_t*)p;
    HIP_TRY(ctx, zalloc(&p, nl * tiles * 2 * 8)); f.part_em = (int64_t*)p;
    HIP_TRY(ctx, zalloc(&p, 2 * nl * 6 * 8)); f.bst = (int64_t*)p;
    HIP_TRY(ctx, g.arena.alloc(&p, nl * 8)); f.fin = (int64_t*)p;
    HIP_TRY(ctx, upload(p, ctr0.data(), nl * 8));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, zalloc(&p, nl * 4)); f.cur = (int32_t*)p;
  } else if (!g.lds) {
    TiledScratch& t = g.tile;
    t.tile = tiled_tile_size(g.L, g.tile_pref);
    t.tiles_x = t.tiles_y = (g.L + t.tile - 1) / t.tile;
    HIP_TRY(ctx, g.arena.alloc(&p, NN)); t.occ = (uint8_t*)p;
    HIP_TRY(ctx, g.arena.alloc(&p, NN * 2)); t.lroot = (uint16_t*)p;
    HIP_TRY(ctx, g.arena.alloc(&p, NN * 8)); t.gf = (uint64_t*)p;
    HIP_TRY(ctx, g.arena.alloc(&p, NN * 4)); t.lst = (uint32_t*)p;
    HIP_TRY(ctx, g.arena.alloc(&p, NN * 8)); t.st = (uint64_t*)p;
    const size_t pb = (size_t)g.n_lat * t.tiles_x * t.tiles_y * 7 * 8;
    HIP_TRY(ctx, g.arena.alloc(&p, pb)); t.part = (int64_t*)p;
    HIP_TRY(ctx, hipMemsetAsync(p, 0, pb, ctx->stream));
    HIP_TRY(ctx, g.arena.alloc(&p, 8 * g.n_lat)); t.wrapf = (uint32_t*)p;
    HIP_TRY(ctx, hipMemsetAsync(p, 0, 8 * g.n_lat, ctx->stream));
    HIP_TRY(ctx, g.arena.alloc(&p, 4 * g.n_lat)); t.pending = (int32_t*)p;
    HIP_TRY(ctx, hipMemsetAsync(p, 0, 4 * g.n_lat, ctx->stream));
  }
  if (empty) {
    HIP_TRY(ctx, hipMemsetAsync(s.spins, 0, NN, ctx->stream));
  } else if (spins0) {
    HIP_TRY(ctx, hipMemcpyAsync(s.spins, spins0, NN, hipMemcpyHostToDevice, ctx->stream));
  } else {
    HIP_TRY(ctx, launch_init_spins(s, start, ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return RBIM_OK;
}

// Queue the next launch of one group on `stream` -- one LDS dispatch or one
// graph of sweeps (single sweeps below a graph's length) -- of the `n` sweeps
// still to run; returns how many it queued and counts kernel launches.
int group_chunk(rbim_ctx* ctx, Group& g, int64_t n, hipStream_t stream, int64_t* launches,
                int64_t* queued) {
  *queued = 0;
  if (n <= 0) return RBIM_OK;
  if (g.lds) {
    const int k = (int)std::min<int64_t>(n, g.per_launch);
    HIP_TRY(ctx, launch_lds_sweeps(g.set, g.threads, k, stream));
    ++*launches;
    *queued = k;
    return RBIM_OK;
  }
  // one sweep of a tiled group (both layouts)
  auto sweep = [&g](hipStream_t st) {
    return g.fused ? launch_fused_sweep(g.set, g.fz, st) : launch_tiled_sweep(g.set, g.tile, st);
  };
  const int per_sweep = g.fused ? fused_kernels_per_sweep(g.set, g.fz) : tiled_kernels_per_sweep();
  const int S = std::max(1, g.per_launch);
  if (n >= S && S > 1) {
    if (!g.graph) {
      hipGraph_t graph;
      HIP_TRY(ctx, hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      for (int i = 0; i < S; ++i) {
        hipError_t e = sweep(stream);
        if (e != hipSuccess) {
          hipGraph_t junk;
          (void)hipStreamEndCapture(stream, &junk);
          return set_err(ctx, RBIM_E_HIP, "tiled sweep capture failed: %s", hipGetErrorString(e));
        }
      }
      HIP_TRY(ctx, hipStreamEndCapture(stream, &graph));
      HIP_TRY(ctx, hipGraphInstantiate(&g.graph, graph, nullptr, nullptr, 0));
      HIP_TRY(ctx, hipGraphDestroy(graph));
    }
    HIP_TRY(ctx, hipGraphLaunch(g.graph, stream));
    *launches += (int64_t)S * per_sweep;
    *queued = S;
    return RBIM_OK;
  }
  const int64_t k = S > 1 ? n : 1;
  for (int64_t i = 0; i < k; ++i) {
    HIP_TRY(ctx, sweep(stream));
    *launches += per_sweep;
  }
  *queued = k;
  return RBIM_OK;
}

// After the last sweep of a run of a tiled group: each sweep's measurement
// rides in the next sweep's kernels, so the last one of the run is measured
// here and the moment sums (and spins) are complete.
int group_close(rbim_ctx* ctx, Group& g, hipStream_t stream, int64_t* launches) {
  if (g.lds) return RBIM_OK;
  if (g.fused) {
    HIP_TRY(ctx, launch_fused_close(g.set, g.fz, stream));
    *launches += 3;
  } else {
    HIP_TRY(ctx, launch_tiled_flush(g.set, g.tile, stream));
    *launches += 2;
  }
  return RBIM_OK;
}

// Before the first sweep of a run of a group: the edge-slot sweep's state
// words from the group's spins.
int group_open(rbim_ctx* ctx, Group& g, hipStream_t stream, int64_t* launches) {
  if (!g.fused) return RBIM_OK;
  HIP_TRY(ctx, launch_fused_open(g.set, g.fz, stream));
  ++*launches;
  return RBIM_OK;
}

// Queue `n` sweeps of one group on `stream`.
int group_run(rbim_ctx* ctx, Group& g, int64_t n, hipStream_t stream, int64_t* launches) {
  const int rc0 = group_open(ctx, g, stream, launches);
  if (rc0 != RBIM_OK) return rc0;
  while (n > 0) {
    int64_t q = 0;
    const int rc = group_chunk(ctx, g, n, stream, launches, &q);
    if (rc != RBIM_OK) return rc;
    n -= q;
  }
  return group_close(ctx, g, stream, launches);
}

void group_free(Group& g) {
  if (g.stream) (void)hipStreamSynchronize(g.stream);
  if (g.graph) (void)hipGraphExecDestroy(g.graph);
  g.graph = nullptr;
  g.arena.release();
  if (g.done) (void)hipEventDestroy(g.done);
  if (g.stream) (void)hipStreamDestroy(g.stream);
  g.done = nullptr;
  g.stream = nullptr;
}

// Tiled sets of lattices are split over this many streams by default: their
// kernels are alternately ALU-bound (activation + tile labelling) and
// latency-bound (cross-tile merge, flip), so two halves interleave on the CUs
// -- except on whole tiles of 64, whose tile pass fills the GPU by itself
// (1280 lattices: one stream 1.384e11 site-updates/s at L = 256 against
// 1.353e11 on two, 1.133e11 against 1.097e11 at L = 384; two streams are
// ahead on tiles of 32 and cut tiles: 1.035e11 against 9.85e10 at L = 160,
// 7.92e10 against 7.63e10 at L = 333).
constexpr int kTiledStreams = 2;
int default_tiled_streams(const Group& g) {
  return g.fused && fused_tile_size(g.L, g.tile_pref) == 64 && g.L % 64 == 0 ? 1 : kTiledStreams;
}
// Streams of one batch at most: the hardware queues a process gets by default.
constexpr int kMaxStreams = 4;

bool valid_L(int L) { return L >= 3 && L <= 16384; }

// One lattice, one sweep (index `sweep`), optionally with given bits.
int run_single(rbim_ctx* ctx, int L, int engine, int8_t* spins, const int8_t* Jh, const int8_t* Jv,
               const uint8_t* occ_pairs, const uint8_t* flips, uint32_t thr, uint64_t seed,
               uint32_t lid, uint32_t sweep, int measure, int32_t* canon_host,
               rbim_sweep_stats* stats, uint8_t* occ_host) {
  if (!valid_L(L)) return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d)", L);
  const bool lds = engine == RBIM_ENGINE_BATCHED;
  if (engine != RBIM_ENGINE_PLAIN && engine != RBIM_ENGINE_BATCHED && engine != RBIM_ENGINE_AUTO)
    return set_err(ctx, RBIM_E_INVAL, "unknown engine %d", engine);
  if (lds && L > RBIM_LDS_MAX_L)
    return set_err(ctx, RBIM_E_INVAL, "the LDS engine holds L <= %d (got %d)", RBIM_LDS_MAX_L, L);
  const int N = L * L;
  Group g;
  g.L = L;
  g.N = N;
  g.n_lat = 1;
  g.lds = lds;
  g.threads = default_threads(L);
  g.per_launch = 1;
  std::vector<uint8_t> jb(N);
  for (int n = 0; n < N; ++n)
    jb[n] = (uint8_t)((Jh && Jh[n] < 0 ? 1 : 0) | (Jv && Jv[n] < 0 ? 2 : 0));
  const int64_t n_therm = measure ? (int64_t)sweep : (int64_t)sweep + 1;
  int rc = group_alloc(ctx, g, jb, {thr}, {seed}, {lid}, spins, RBIM_START_COLD, n_therm, 1, 1,
                       (int64_t)sweep);
  if (rc != RBIM_OK) { group_free(g); return rc; }
  DevArena extra;
  void* p;
  uint8_t* d_occ = nullptr;
  uint8_t* d_flip = nullptr;
  int32_t* d_canon = nullptr;
  auto fail = [&](int code) { extra.release(); group_free(g); return code; };
#define TRY2(expr)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(set_err(ctx, RBIM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)));  \
  } while (0)
  if (occ_pairs) {
    TRY2(extra.alloc(&p, N)); d_occ = (uint8_t*)p;
    TRY2(hipMemcpyAsync(d_occ, occ_pairs, N, hipMemcpyHostToDevice, ctx->stream));
    TRY2(extra.alloc(&p, N)); d_flip = (uint8_t*)p;
    TRY2(hipMemcpyAsync(d_flip, flips, N, hipMemcpyHostToDevice, ctx->stream));
    g.set.given_occ = d_occ;
    g.set.given_flip = d_flip;
  }
  if (canon_host) {
    TRY2(extra.alloc(&p, 4 * (size_t)N)); d_canon = (int32_t*)p;
    g.set.canon_out = d_canon;
  }
  int64_t launches = 0;
  rc = group_run(ctx, g, 1, ctx->stream, &launches);
  if (rc != RBIM_OK) return fail(rc);
  TRY2(hipMemcpyAsync(spins, g.set.spins, N, hipMemcpyDeviceToHost, ctx->stream));
  if (canon_host) TRY2(hipMemcpyAsync(canon_host, d_canon, 4 * (size_t)N, hipMemcpyDeviceToHost, ctx->stream));
  if (stats) TRY2(hipMemcpyAsync(stats, g.set.last, sizeof(rbim_sweep_stats), hipMemcpyDeviceToHost, ctx->stream));
  if (occ_host) TRY2(hipMemcpyAsync(occ_host, g.tile.occ, N, hipMemcpyDeviceToHost, ctx->stream));
  TRY2(hipStreamSynchronize(ctx->stream));
#undef TRY2
  extra.release();
  group_free(g);
  return RBIM_OK;
}

// ------------------------------------------ allocation-free one-lattice sweep
// Carves the device buffers of a one-lattice group -- the LDS kernel's
// (lds) or the multi-kernel tiled sweep's -- plus staging for the raw
// couplings out of [base, base + cap), 256-byte aligned, and points g's
// descriptors at them.  Returns the bytes spanned, or 0 when they exceed cap.
// base == nullptr only sizes the layout.
size_t carve_single(uint8_t* base, size_t cap, int L, bool lds, Group& g, int8_t** jh_raw,
                    int8_t** jv_raw) {
  const size_t N = (size_t)L * L;
  size_t used = 0;
  bool fits = true;
  auto take = [&](size_t n) -> uint8_t* {
    const size_t off = (used + 255) & ~(size_t)255;
    if (off + n > cap) fits = false;
    used = off + n;
    return base ? base + off : nullptr;
  };
  LatticeSet& s = g.set;
  s.spins = (int8_t*)take(N);
  s.jbits = take(N);
  s.thr31 = (const uint32_t*)take(4);
  s.seed = (const uint64_t*)take(8);
  s.lat_id = (const uint32_t*)take(4);
  s.sweep_ctr = (int64_t*)take(8);
  s.acc = (uint64_t*)take(RBIM_ACC_SLOTS * 8);
  s.last = (rbim_sweep_stats*)take(sizeof(rbim_sweep_stats));
  *jh_raw = (int8_t*)take(N);
  *jv_raw = (int8_t*)take(N);
  if (!lds) {
    TiledScratch& t = g.tile;
    t.tile = tiled_tile_size(L, 0);
    t.tiles_x = t.tiles_y = (L + t.tile - 1) / t.tile;
    t.occ = take(N);
    t.lroot = (uint16_t*)take(2 * N);
    t.gf = (uint64_t*)take(8 * N);
    t.lst = (uint32_t*)take(4 * N);
    t.st = (uint64_t*)take(8 * N);
    t.part = (int64_t*)take((size_t)t.tiles_x * t.tiles_y * 7 * 8);
    t.wrapf = (uint32_t*)take(8);
    t.pending = (int32_t*)take(4);
  }
  used = (used + 255) & ~(size_t)255;
  return fits ? used : 0;
}

__global__ void scratch_zero_kernel(uint4* p, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Packs the staged couplings into coupling bits and writes the lattice's
// scalars (the buffers are the caller's scratch, hence the casts).
__global__ void noalloc_setup_kernel(LatticeSet s, const int8_t* jh, const int8_t* jv,
                                     uint32_t thr, uint64_t seed, uint32_t lid, int64_t g0) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < s.N)
    const_cast<uint8_t*>(s.jbits)[n] = (uint8_t)((jh[n] < 0 ? 1 : 0) | (jv[n] < 0 ? 2 : 0));
  if (n == 0) {
    const_cast<uint32_t*>(s.thr31)[0] = thr;
    const_cast<uint64_t*>(s.seed)[0] = seed;
    const_cast<uint32_t*>(s.lat_id)[0] = lid;
    s.sweep_ctr[0] = g0;
  }
}

// One sweep of one lattice, as run_single without given bits, on buffers
// carved from the device's scratch array: no allocation of any kind.
int run_single_noalloc(rbim_ctx* ctx, uint8_t* base, int L, bool lds, int8_t* spins,
                       const int8_t* Jh, const int8_t* Jv, uint32_t thr, uint64_t seed,
                       uint32_t lid, uint32_t sweep, int measure, rbim_sweep_stats* stats) {
  const int N = L * L;
  Group g;  // descriptors only: its arena stays empty
  g.L = L;
  g.N = N;
  g.n_lat = 1;
  g.lds = lds;
  g.threads = default_threads(L);
  g.per_launch = 1;
  int8_t *jh_raw = nullptr, *jv_raw = nullptr;
  const size_t used = carve_single(base, RBIM_NOALLOC_SCRATCH_BYTES, L, lds, g, &jh_raw, &jv_raw);
  if (used == 0)
    return set_err(ctx, RBIM_E_INVAL, "L=%d needs more than the %u-byte scratch", L,
                   (unsigned)RBIM_NOALLOC_SCRATCH_BYTES);
  LatticeSet& s = g.set;
  s.L = L;
  s.N = N;
  s.n_lat = 1;
  s.n_run = 1;
  s.sched = nullptr;
  s.lat_map = nullptr;
  s.magic = magic_for(L);
  s.n_therm = measure ? (int64_t)sweep : (int64_t)sweep + 1;
  s.spb = 1;
  s.n_bins = 1;
  s.total = s.n_therm + 1;
  s.series = nullptr;
  s.series_cap = 0;
  s.given_occ = nullptr;
  s.given_flip = nullptr;
  s.canon_out = nullptr;
  hipStream_t st = ctx->stream;
  // the previous call's contents stay in the array: clear everything this one uses
  hipLaunchKernelGGL(scratch_zero_kernel, dim3(512), dim3(256), 0, st, (uint4*)base, used / 16);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(s.spins, spins, N, hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(jh_raw, Jh, N, hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(jv_raw, Jv, N, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(noalloc_setup_kernel, dim3((N + 255) / 256), dim3(256), 0, st, s, jh_raw,
                     jv_raw, thr, seed, lid, (int64_t)sweep);
  HIP_TRY(ctx, hipGetLastError());
  int64_t launches = 0;
  const int rc = group_run(ctx, g, 1, st, &launches);
  if (rc != RBIM_OK) return rc;
  HIP_TRY(ctx, hipMemcpyAsync(spins, s.spins, N, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipMemcpyAsync(stats, s.last, sizeof(rbim_sweep_stats), hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  return RBIM_OK;
}

// active[2N] in bond order -> per-site bits (bit0 horizontal, bit1 vertical)
std::vector<uint8_t> pack_active(const uint8_t* active, int N) {
  std::vector<uint8_t> o(N);
  for (int n = 0; n < N; ++n) o[n] = (uint8_t)((active[2 * n] ? 1 : 0) | (active[2 * n + 1] ? 2 : 0));
  return o;
}

// labels (min-site order) from canonical labels, on the device.
int ranks_from_canon(rbim_ctx* ctx, const int32_t* canon, int N, int32_t* labels, int32_t* ncl) {
  DevArena a;
  void* p;
  HIP_TRY(ctx, a.alloc(&p, 4 * (size_t)N * 3 + 16));
  int32_t* d_canon = (int32_t*)p;
  int32_t* d_lab = d_canon + N;
  int32_t* d_rank = d_lab + N;
  int32_t* d_n = d_rank + N;
  hipError_t e = hipMemcpyAsync(d_canon, canon, 4 * (size_t)N, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_label_rank(d_canon, N, d_lab, d_rank, d_n, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(labels, d_lab, 4 * (size_t)N, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && ncl) e = hipMemcpyAsync(ncl, d_n, 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  a.release();
  if (e != hipSuccess) return set_err(ctx, RBIM_E_HIP, "label ranking failed: %s", hipGetErrorString(e));
  return RBIM_OK;
}


// ------------------------------------------------------------- slot pools
// Per-slot device state a load clears: `planes` copies, plane_stride bytes
// apart, of per_lat bytes for each slot (multiples of 4 bytes).
struct Region {
  uint8_t* base;
  uint64_t per_lat, plane_stride;
  int32_t planes;
};
struct RegionTable {
  Region r[16];
  int32_t n;
};

__global__ void pool_reset_kernel(RegionTable t, const int32_t* slots) {
  const uint64_t slot = (uint64_t)slots[blockIdx.y];
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (int i = 0; i < t.n; ++i) {
    const Region rg = t.r[i];
    for (int pl = 0; pl < rg.planes; ++pl) {
      uint32_t* w = reinterpret_cast<uint32_t*>(rg.base + pl * rg.plane_stride + slot * rg.per_lat);
      for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < rg.per_lat / 4; k += step)
        w[k] = 0u;
    }
  }
}

// One staged chunk of lattices k = 0 .. n-1 on the device.
struct PoolStage {
  const uint8_t* jbits;  // [n][N]
  const int8_t* spins;   // [n][N], or NULL: the start's spins
  const uint32_t* thr;   // [n]
  const uint64_t* seed;  // [n]
  const uint32_t* lid;   // [n]
  const int32_t* slots;  // [n]
};

// Staged lattices into their slots: couplings, spins, scalars and schedule;
// the sweep counter and the edge-slot sweep's bookkeeping start at zero.
__global__ void pool_load_kernel(LatticeSet s, int64_t* sched, int64_t* fin, int32_t* cur,
                                 PoolStage st, int start, int64_t n_therm, int64_t spb,
                                 int64_t total) {
  const int k = blockIdx.y;
  const int slot = st.slots[k];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t base = (size_t)slot * s.N, src = (size_t)k * s.N;
  if (n < s.N) {
    const_cast<uint8_t*>(s.jbits)[base + n] = st.jbits[src + n];
    s.spins[base + n] = st.spins ? st.spins[src + n]
                        : start == RBIM_START_COLD
                            ? (int8_t)1
                            : (int8_t)rbim_init_spin(st.seed[k], st.lid[k], (uint32_t)n);
  }
  if (n == 0) {
    const_cast<uint32_t*>(s.thr31)[slot] = st.thr[k];
    const_cast<uint64_t*>(s.seed)[slot] = st.seed[k];
    const_cast<uint32_t*>(s.lat_id)[slot] = st.lid[k];
    s.sweep_ctr[slot] = 0;
    sched[3 * (size_t)slot] = n_therm;
    sched[3 * (size_t)slot + 1] = spb;
    sched[3 * (size_t)slot + 2] = total;
    if (fin) fin[slot] = 0;
    if (cur) cur[slot] = 0;
  }
}

// Lattices per staged load chunk: the whole pool while its staging stays
// under 32 MiB.
int pool_chunk(int N, int n_slots) {
  const size_t per = 2 * (size_t)N + 32;
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)n_slots, (32ull << 20) / per));
}

size_t pool_stage_bytes(int N, int chunk) {
  return (size_t)chunk * (2 * (size_t)N + 32) + 6 * 256;
}

}  // namespace

// =================================================================== batch
struct rbim_batch {
  rbim_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  std::vector<hipStream_t> streams;  // the groups' streams (groups may share one)
  bool ragged = false;               // LDS groups run in one launch (lds_ragged_kernel)
  int ragged_lead = -1;              // the LDS group whose turn queues that launch
  int ragged_per_launch = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, fork = nullptr;
  std::vector<Group> groups;
  int n_lat = 0;
  std::vector<int64_t> off;  // site offset of lattice i (batch order)
  std::vector<int> Ls;
  int64_t n_therm = 0, spb = 1, n_bins = 1, total = 0, issued = 0, launches = 0;
};

extern "C" {

int rbim_abi_version(void) { return RBIM_ABI_VERSION; }

const char* rbim_last_error(const rbim_ctx* ctx) {
  static thread_local std::string copy;
  if (ctx) {
    std::lock_guard<std::mutex> lk(const_cast<rbim_ctx*>(ctx)->mu);
    copy = ctx->err;
  } else {
    std::lock_guard<std::mutex> lk(g_err_mu);
    copy = g_err;
  }
  return copy.c_str();
}

int rbim_open(int device, rbim_ctx** out) {
  if (!out) return set_err(nullptr, RBIM_E_INVAL, "out is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    return set_err(nullptr, RBIM_E_HIP, "no HIP device available (%s)",
                   e != hipSuccess ? hipGetErrorString(e) : "device count is 0");
  if (device < 0 || device >= n)
    return set_err(nullptr, RBIM_E_INVAL, "device %d out of range (have %d)", device, n);
  if (device >= kMaxDevices)
    return set_err(nullptr, RBIM_E_INVAL, "device %d beyond the %d this library tracks", device,
                   kMaxDevices);
  rbim_ctx* c = new rbim_ctx();
  c->device = device;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = resolve_scratch(device);
  if (e != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return set_err(nullptr, RBIM_E_HIP, "context creation failed: %s", hipGetErrorString(e));
  }
  *out = c;
  return RBIM_OK;
}

int rbim_close(rbim_ctx* ctx) {
  if (!ctx) return RBIM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RBIM_OK;
}

int rbim_device_memory(rbim_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
  if (!ctx) return set_err(nullptr, RBIM_E_INVAL, "ctx is NULL");
  size_t f = 0, t = 0;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return RBIM_OK;
}

int rbim_label_and_wrap(rbim_ctx* ctx, int32_t L, const uint8_t* active, int32_t engine,
                        int32_t* labels, int32_t* canon, int32_t* n_clusters, int32_t* wrap_xy) {
  if (!ctx || !active) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  if (!valid_L(L)) return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d)", L);
  const int N = L * L;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::vector<uint8_t> occ = pack_active(active, N);
  std::vector<uint8_t> flips(N, 0);
  std::vector<int8_t> spins(N, 1);
  std::vector<int32_t> cn(N);
  rbim_sweep_stats st{};
  int rc = run_single(ctx, L, engine, spins.data(), nullptr, nullptr, occ.data(), flips.data(), 0,
                      0, 0, 0, 1, cn.data(), &st, nullptr);
  if (rc != RBIM_OK) return rc;
  if (canon) std::memcpy(canon, cn.data(), 4 * (size_t)N);
  if (n_clusters) *n_clusters = (int32_t)st.n_clusters;
  if (wrap_xy) {
    wrap_xy[0] = st.wrap_x;
    wrap_xy[1] = st.wrap_y;
  }
  if (labels) return ranks_from_canon(ctx, cn.data(), N, labels, nullptr);
  return RBIM_OK;
}

int rbim_activate(rbim_ctx* ctx, int32_t L, const int8_t* spins, const int8_t* Jh, const int8_t* Jv,
                  uint32_t thr31, uint64_t seed, uint32_t lattice_id, uint32_t sweep,
                  uint8_t* active, uint8_t* flipbits) {
  if (!ctx || !spins || !Jh || !Jv || !active) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  if (!valid_L(L)) return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d)", L);
  if (thr31 > (1u << 31)) return set_err(ctx, RBIM_E_INVAL, "thr31 exceeds 2^31");
  const int N = L * L;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::vector<int8_t> s(spins, spins + N);
  std::vector<uint8_t> occ(N);
  int rc = run_single(ctx, L, RBIM_ENGINE_PLAIN, s.data(), Jh, Jv, nullptr, nullptr, thr31, seed,
                      lattice_id, sweep, 0, nullptr, nullptr, occ.data());
  if (rc != RBIM_OK) return rc;
  for (int n = 0; n < N; ++n) {
    active[2 * n] = occ[n] & 1u;
    active[2 * n + 1] = (occ[n] >> 1) & 1u;
    if (flipbits) flipbits[n] = (occ[n] >> 2) & 1u;
  }
  return RBIM_OK;
}

int rbim_sweep_given_bits(rbim_ctx* ctx, int32_t L, int8_t* spins, const int8_t* Jh,
                          const int8_t* Jv, const uint8_t* active, const uint8_t* flip_by_minsite,
                          int32_t engine, int32_t* labels, rbim_sweep_stats* out) {
  if (!ctx || !spins || !Jh || !Jv || !active || !flip_by_minsite)
    return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  if (!valid_L(L)) return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d)", L);
  const int N = L * L;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::vector<uint8_t> occ = pack_active(active, N);
  std::vector<uint8_t> fl(N);
  for (int n = 0; n < N; ++n) fl[n] = flip_by_minsite[n] ? 1 : 0;
  std::vector<int32_t> cn(labels ? N : 0);
  rbim_sweep_stats st{};
  int rc = run_single(ctx, L, engine, spins, Jh, Jv, occ.data(), fl.data(), 0, 0, 0, 0, 1,
                      labels ? cn.data() : nullptr, &st, nullptr);
  if (rc != RBIM_OK) return rc;
  if (out) *out = st;
  if (labels) return ranks_from_canon(ctx, cn.data(), N, labels, nullptr);
  return RBIM_OK;
}

int rbim_sweep(rbim_ctx* ctx, int32_t L, int8_t* spins, const int8_t* Jh, const int8_t* Jv,
               uint32_t thr31, uint64_t seed, uint32_t lattice_id, uint32_t sweep, int32_t measure,
               int32_t engine, rbim_sweep_stats* out) {
  if (!ctx || !spins || !Jh || !Jv) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  if (thr31 > (1u << 31)) return set_err(ctx, RBIM_E_INVAL, "thr31 exceeds 2^31");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  rbim_sweep_stats st{};
  int rc = run_single(ctx, L, engine, spins, Jh, Jv, nullptr, nullptr, thr31, seed, lattice_id,
                      sweep, measure, nullptr, &st, nullptr);
  if (rc == RBIM_OK && out && measure) *out = st;
  return rc;
}

int rbim_sweep_noalloc(rbim_ctx* ctx, int32_t L, int8_t* spins, const int8_t* Jh,
                       const int8_t* Jv, uint32_t thr31, uint64_t seed, uint32_t lattice_id,
                       uint32_t sweep, int32_t measure, int32_t engine, rbim_sweep_stats* out) {
  if (!ctx || !spins || !Jh || !Jv) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  if (!valid_L(L)) return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d)", L);
  if (L > RBIM_NOALLOC_MAX_L)
    return set_err(ctx, RBIM_E_INVAL,
                   "rbim_sweep_noalloc holds L <= %d in its %u-byte scratch (got L=%d)",
                   RBIM_NOALLOC_MAX_L, (unsigned)RBIM_NOALLOC_SCRATCH_BYTES, L);
  if (engine != RBIM_ENGINE_PLAIN && engine != RBIM_ENGINE_BATCHED && engine != RBIM_ENGINE_AUTO)
    return set_err(ctx, RBIM_E_INVAL, "unknown engine %d", engine);
  const bool lds = engine == RBIM_ENGINE_BATCHED;
  if (lds && L > RBIM_LDS_MAX_L)
    return set_err(ctx, RBIM_E_INVAL, "the LDS engine holds L <= %d (got %d)", RBIM_LDS_MAX_L, L);
  if (thr31 > (1u << 31)) return set_err(ctx, RBIM_E_INVAL, "thr31 exceeds 2^31");
  ScratchSlot& sl = g_scratch[ctx->device];
  std::lock_guard<std::mutex> lk(sl.mu);  // calls on one device share its array
  if (!sl.ptr) return set_err(ctx, RBIM_E_STATE, "scratch of device %d not resolved", ctx->device);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  rbim_sweep_stats st{};
  const int rc = run_single_noalloc(ctx, sl.ptr, L, lds, spins, Jh, Jv, thr31, seed, lattice_id,
                                    sweep, measure, &st);
  if (rc == RBIM_OK && out && measure) *out = st;
  return rc;
}

int rbim_cluster_statistics(rbim_ctx* ctx, int64_t n_sites, const int32_t* labels,
                            const int8_t* spins, int32_t n_clusters, int64_t* sizes, int64_t* mags) {
  if (!ctx || !labels || !spins || !sizes || !mags) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  if (n_sites <= 0 || n_clusters <= 0) return set_err(ctx, RBIM_E_INVAL, "empty input");
  for (int64_t i = 0; i < n_sites; ++i)
    if (labels[i] < 0 || labels[i] >= n_clusters)
      return set_err(ctx, RBIM_E_INVAL, "label %d at site %lld outside [0, %d)", labels[i],
                     (long long)i, n_clusters);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  DevArena a;
  void* p;
  HIP_TRY(ctx, a.alloc(&p, (size_t)n_clusters * 16 + (size_t)n_sites * 5 + 16));
  unsigned long long* d_sz = (unsigned long long*)p;
  unsigned long long* d_mg = d_sz + n_clusters;
  int32_t* d_lab = (int32_t*)(d_mg + n_clusters);
  int8_t* d_sp = (int8_t*)(d_lab + n_sites);
  hipError_t e = hipMemsetAsync(d_sz, 0, (size_t)n_clusters * 16, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_lab, labels, 4 * (size_t)n_sites, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_sp, spins, (size_t)n_sites, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_cluster_stats(d_lab, d_sp, n_sites, d_sz, d_mg, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(sizes, d_sz, 8 * (size_t)n_clusters, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(mags, d_mg, 8 * (size_t)n_clusters, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  a.release();
  if (e != hipSuccess) return set_err(ctx, RBIM_E_HIP, "cluster statistics failed: %s", hipGetErrorString(e));
  return RBIM_OK;
}

int rbim_energy_magnetization(rbim_ctx* ctx, int32_t L, const int8_t* spins, const int8_t* Jh,
                              const int8_t* Jv, int64_t* energy, int64_t* magnetization) {
  if (!ctx || !spins || !Jh || !Jv) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  if (!valid_L(L)) return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d)", L);
  const int N = L * L;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::vector<uint8_t> jb(N);
  for (int n = 0; n < N; ++n) jb[n] = (uint8_t)((Jh[n] < 0 ? 1 : 0) | (Jv[n] < 0 ? 2 : 0));
  DevArena a;
  void* p;
  HIP_TRY(ctx, a.alloc(&p, 16 + 2 * (size_t)N));
  unsigned long long* d_em = (unsigned long long*)p;
  int8_t* d_sp = (int8_t*)(d_em + 2);
  uint8_t* d_jb = (uint8_t*)(d_sp + N);
  long long em[2] = {0, 0};
  hipError_t e = hipMemsetAsync(d_em, 0, 16, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_sp, spins, N, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_jb, jb.data(), N, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_energy_mag(d_sp, d_jb, L, d_em, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(em, d_em, 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  a.release();
  if (e != hipSuccess) return set_err(ctx, RBIM_E_HIP, "energy failed: %s", hipGetErrorString(e));
  if (energy) *energy = em[0];
  if (magnetization) *magnetization = em[1];
  return RBIM_OK;
}

// ------------------------------------------------------------------ batch
static int check_desc(rbim_ctx* ctx, const rbim_batch_desc* d) {
  if (!d) return set_err(ctx, RBIM_E_INVAL, "desc is NULL");
  if (d->n_lattices <= 0) return set_err(ctx, RBIM_E_INVAL, "n_lattices must be positive");
  if (!d->L || !d->Jh || !d->Jv || !d->thr31 || !d->seed || !d->lattice_id)
    return set_err(ctx, RBIM_E_INVAL, "NULL array in batch description");
  if (d->n_bins < 1) return set_err(ctx, RBIM_E_INVAL, "n_bins must be positive, got %d", d->n_bins);
  if (d->sweeps_per_bin < 1)
    return set_err(ctx, RBIM_E_INVAL, "sweeps_per_bin must be at least 1");
  if (d->n_therm < 0) return set_err(ctx, RBIM_E_INVAL, "n_therm must be non-negative");
  if (d->n_therm + (int64_t)d->n_bins * d->sweeps_per_bin > 0xFFFFFFFFll)
    return set_err(ctx, RBIM_E_INVAL, "more than 2^32 sweeps per lattice");
  if (d->engine < RBIM_ENGINE_AUTO || d->engine > RBIM_ENGINE_TILED)
    return set_err(ctx, RBIM_E_INVAL, "unknown engine %d", d->engine);
  if (d->series < 0) return set_err(ctx, RBIM_E_INVAL, "series must be >= 0");
  if (d->tile != 0 && d->tile != 32 && d->tile != 64)
    return set_err(ctx, RBIM_E_INVAL, "tile must be 0 (by L), 32 or 64");
  if (d->streams < 0 || d->streams > 16)
    return set_err(ctx, RBIM_E_INVAL, "streams must be 0 (default) .. 16");
  if (d->single_launch != 0 && d->single_launch != 1)
    return set_err(ctx, RBIM_E_INVAL, "single_launch must be 0 or 1");
  if (d->threads != 0 && d->threads != 64 && d->threads != 128 && d->threads != 256 &&
      d->threads != 512 && d->threads != 1024)
    return set_err(ctx, RBIM_E_INVAL, "threads must be 64, 128, 256, 512 or 1024");
  for (int i = 0; i < d->n_lattices; ++i) {
    if (!valid_L(d->L[i]))
      return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d for lattice %d)", d->L[i], i);
    if (d->engine == RBIM_ENGINE_BATCHED && d->L[i] > RBIM_LDS_MAX_L)
      return set_err(ctx, RBIM_E_INVAL, "the LDS engine holds L <= %d (lattice %d has L=%d)",
                     RBIM_LDS_MAX_L, i, d->L[i]);
    if (d->thr31[i] > (1u << 31)) return set_err(ctx, RBIM_E_INVAL, "thr31 exceeds 2^31");
  }
  return RBIM_OK;
}

// The auto engine's choice for a set of `count` lattices of side L: the
// LDS-resident kernel up to L = 64, and up to L = 128 while the set holds
// fewer than 2^23 sites; the edge-slot sweep above.  Measured on scans of 1280
// lattices (site-updates/s, LDS kernel against edge-slot sweep): L = 128
// 1.20e11 / 1.50e11, 112 9.6e10 / 9.7e10, 96 9.4e10 / 1.12e11, 80 8.2e10 /
// 9.5e10, 64 1.23e11 / 1.06e11, 48 8.2e10 / 5.7e10; and of 256: L = 128
// 1.10e11 / 8.4e10 (the edge-slot sweep's two launches per sweep against the
// LDS kernel's hundreds of sweeps per launch), L = 64 6.3e10 / 3.7e10.
static bool auto_lds(int L, int64_t count) {
  return L <= 64 || (L <= RBIM_LDS_MAX_L && count * (int64_t)L * L < (int64_t(1) << 23));
}

static bool use_lds(const rbim_batch_desc* d, int L, int64_t count) {
  return d->engine == RBIM_ENGINE_BATCHED || (d->engine == RBIM_ENGINE_AUTO && auto_lds(L, count));
}

// Lattices of each side in a batch.
static std::map<int, int64_t> side_counts(const rbim_batch_desc* d) {
  std::map<int, int64_t> n;
  for (int i = 0; i < d->n_lattices; ++i) n[d->L[i]]++;
  return n;
}

static bool use_fused(const rbim_batch_desc* d, int L, bool lds) {
  return !lds && d->engine != RBIM_ENGINE_TILED && fused_tile_size(L) > 0;
}

int rbim_batch_footprint(const rbim_batch_desc* d, uint64_t* bytes) {
  int rc = check_desc(nullptr, d);
  if (rc != RBIM_OK) return rc;
  std::map<std::pair<int, int>, int> counts;
  const std::map<int, int64_t> sides = side_counts(d);
  for (int i = 0; i < d->n_lattices; ++i)
    counts[{d->L[i], use_lds(d, d->L[i], sides.at(d->L[i])) ? 1 : 0}]++;
  uint64_t b = 0;
  for (auto& kv : counts) {
    const bool lds = kv.first.second != 0;
    b += group_bytes(kv.first.first, kv.second, lds, use_fused(d, kv.first.first, lds), d->n_bins,
                     d->series, d->tile);
  }
  if (bytes) *bytes = b;
  return RBIM_OK;
}

int rbim_batch_create(rbim_ctx* ctx, const rbim_batch_desc* d, rbim_batch** out) {
  if (!ctx || !out) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  int rc = check_desc(ctx, d);
  if (rc != RBIM_OK) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  uint64_t need = 0;
  rbim_batch_footprint(d, &need);
  size_t fr = 0, tot = 0;
  HIP_TRY(ctx, hipMemGetInfo(&fr, &tot));
  if (need + (64ull << 20) > fr)
    return set_err(ctx, RBIM_E_ADMIT, "batch needs %llu bytes, %llu free", (unsigned long long)need,
                   (unsigned long long)fr);
  rbim_batch* b = new rbim_batch();
  b->ctx = ctx;
  b->n_lat = d->n_lattices;
  b->n_therm = d->n_therm;
  b->spb = d->sweeps_per_bin;
  b->n_bins = d->n_bins;
  b->total = d->n_therm + (int64_t)d->n_bins * d->sweeps_per_bin;
  b->off.resize(d->n_lattices + 1);
  b->Ls.assign(d->L, d->L + d->n_lattices);
  b->off[0] = 0;
  for (int i = 0; i < d->n_lattices; ++i) b->off[i + 1] = b->off[i] + (int64_t)d->L[i] * d->L[i];
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e != hipSuccess) {
    rbim_batch_destroy(b);
    return set_err(ctx, RBIM_E_HIP, "stream/event creation failed: %s", hipGetErrorString(e));
  }
  std::map<std::pair<int, int>, int> gid;
  const std::map<int, int64_t> sides = side_counts(d);
  for (int i = 0; i < d->n_lattices; ++i) {
    const int L = d->L[i];
    const int lds = use_lds(d, L, sides.at(L)) ? 1 : 0;
    auto key = std::make_pair(L, lds);
    if (!gid.count(key)) {
      gid[key] = (int)b->groups.size();
      b->groups.emplace_back();
      Group& g = b->groups.back();
      g.L = L;
      g.N = L * L;
      g.lds = lds != 0;
      g.fused = use_fused(d, L, g.lds);
      g.tile_pref = d->tile;
      g.threads = d->threads ? d->threads : default_threads(L);
      g.per_launch = d->sweeps_per_launch > 0 ? d->sweeps_per_launch
                                              : (g.lds ? default_lds_sweeps(L) : 16);
    }
    b->groups[gid[key]].idx.push_back(i);
  }
  int n_cu = 0;
  (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device);
  if (d->threads == 0) default_lds_threads(b->groups, n_cu);
  // one launch for every LDS set when asked for and every LDS side is one of
  // the ragged kernel's (its page shapes fix the workgroup sizes)
  b->ragged = d->single_launch != 0;
  int lds_sets = 0;
  for (Group& g : b->groups)
    if (g.lds) {
      ++lds_sets;
      if (ragged_slot(g.L) < 0) b->ragged = false;
    }
  b->ragged = b->ragged && lds_sets > 0;
  if (b->ragged) {
    b->ragged_per_launch = d->sweeps_per_launch;
    for (Group& g : b->groups)
      if (g.lds) {
        g.threads = ragged_threads(g.L);
        if (d->sweeps_per_launch <= 0)
          b->ragged_per_launch = b->ragged_per_launch > 0
                                     ? std::min(b->ragged_per_launch, default_lds_sweeps(g.L))
                                     : default_lds_sweeps(g.L);
      }
  }
  // split groups into contiguous chunks, one stream each: tiled groups in two
  // (ALU- and latency-bound kernels of the halves interleave); LDS groups in
  // two when they fill the GPU more than once and the last round would run
  // part-empty, so the chunks' launches refill the freed CUs
  {
    std::vector<Group> split;
    for (Group& g : b->groups) {
      int want = d->streams > 0 ? d->streams : default_tiled_streams(g);
      if (g.lds && b->ragged) {
        want = 1;
      } else if (g.lds && d->streams <= 0) {
        const long cap = (long)lds_blocks_per_cu(g.L, g.threads) * n_cu;
        const long n = (long)g.idx.size();
        want = cap > 0 && n > cap && n % cap != 0 ? 2 : 1;
      }
      const int parts = std::max(1, std::min(want, (int)g.idx.size()));
      const size_t n = g.idx.size();
      for (int k = 0; k < parts; ++k) {
        Group h;
        h.L = g.L;
        h.N = g.N;
        h.lds = g.lds;
        h.fused = g.fused;
        h.tile_pref = g.tile_pref;
        h.threads = g.threads;
        h.per_launch = g.per_launch;
        h.idx.assign(g.idx.begin() + n * k / parts, g.idx.begin() + n * (k + 1) / parts);
        split.push_back(std::move(h));
      }
    }
    b->groups = std::move(split);
  }
  // Streams: at most kMaxStreams (the device's hardware queues; more streams
  // would share queues anyway), the largest groups first on streams of their
  // own; with one launch for the LDS sets, those share one stream.
  {
    std::vector<int> order;
    for (int i = 0; i < (int)b->groups.size(); ++i) {
      Group& g = b->groups[i];
      if (b->ragged && g.lds) {
        if (b->ragged_lead < 0) b->ragged_lead = i;
        if (b->ragged_lead != i) continue;
      }
      order.push_back(i);
    }
    auto work = [&](int i) {
      int64_t w = 0;  // a ragged launch carries the sites of every LDS set
      for (Group& g : b->groups)
        if (&g == &b->groups[i] || (b->ragged && g.lds && b->groups[i].lds))
          w += (int64_t)g.N * (int64_t)g.idx.size();
      return w;
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    const int n_streams = std::min<int>((int)order.size(), kMaxStreams);
    for (int k = 0; k < n_streams; ++k) {
      hipStream_t st = nullptr;
      hipError_t se = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
      if (se != hipSuccess) {
        rbim_batch_destroy(b);
        return set_err(ctx, RBIM_E_HIP, "stream creation failed: %s", hipGetErrorString(se));
      }
      b->streams.push_back(st);
    }
    for (int k = 0; k < (int)order.size(); ++k) b->groups[order[k]].stream = b->streams[k % n_streams];
    for (Group& g : b->groups)
      if (b->ragged && g.lds) g.stream = b->groups[b->ragged_lead].stream;
  }
  for (Group& g : b->groups) {
    hipError_t se = hipEventCreateWithFlags(&g.done, hipEventDisableTiming);
    if (se != hipSuccess) {
      rbim_batch_destroy(b);
      return set_err(ctx, RBIM_E_HIP, "event creation failed: %s", hipGetErrorString(se));
    }
  }
  for (Group& g : b->groups) {
    g.n_lat = (int)g.idx.size();
    std::vector<uint8_t> jb((size_t)g.N * g.n_lat);
    std::vector<uint32_t> thr(g.n_lat), lid(g.n_lat);
    std::vector<uint64_t> seed(g.n_lat);
    std::vector<int8_t> sp(d->spins0 ? (size_t)g.N * g.n_lat : 0);
    for (int k = 0; k < g.n_lat; ++k) {
      const int i = g.idx[k];
      const int64_t o = b->off[i];
      pack_jbits(d->Jh + o, d->Jv + o, jb.data() + (size_t)k * g.N, g.N);
      if (d->spins0) std::memcpy(&sp[(size_t)k * g.N], d->spins0 + o, g.N);
      thr[k] = d->thr31[i];
      seed[k] = d->seed[i];
      lid[k] = d->lattice_id[i];
    }
    rc = group_alloc(ctx, g, jb, thr, seed, lid, d->spins0 ? sp.data() : nullptr, d->start,
                     d->n_therm, d->sweeps_per_bin, d->n_bins, 0, d->series);
    if (rc != RBIM_OK) {
      rbim_batch_destroy(b);
      return rc;
    }
  }
  *out = b;
  return RBIM_OK;
}

int rbim_batch_run(rbim_batch* b, int64_t n_sweeps) {
  if (!b) return set_err(nullptr, RBIM_E_INVAL, "batch is NULL");
  rbim_ctx* ctx = b->ctx;
  if (n_sweeps < 0) return set_err(ctx, RBIM_E_INVAL, "n_sweeps must be non-negative");
  if (b->issued + n_sweeps > b->total)
    return set_err(ctx, RBIM_E_STATE, "run of %lld sweeps passes the end (%lld of %lld issued)",
                   (long long)n_sweeps, (long long)b->issued, (long long)b->total);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (b->groups.size() == 1 && !b->ragged) {
    int rc = group_run(ctx, b->groups[0], n_sweeps, b->stream, &b->launches);
    if (rc != RBIM_OK) return rc;
  } else {
    // fork onto the group streams and join back, so the batch stream (what
    // query, sync and the timing events see) completes after every group.
    // The groups' launches are queued in turn, one launch each per round: a
    // host that queued one group's whole run first would block on that
    // stream's full queue before the next group had any work on the device.
    HIP_TRY(ctx, hipEventRecord(b->fork, b->stream));
    for (Group& g : b->groups) {
      HIP_TRY(ctx, hipStreamWaitEvent(g.stream, b->fork, 0));
      const int rc = group_open(ctx, g, g.stream, &b->launches);
      if (rc != RBIM_OK) return rc;
    }
    // one launch for the LDS sets: the lead LDS group's turn queues it
    const LatticeSet* sets[4] = {nullptr, nullptr, nullptr, nullptr};
    if (b->ragged)
      for (Group& g : b->groups)
        if (g.lds) sets[ragged_slot(g.L)] = &g.set;
    std::vector<int64_t> left(b->groups.size(), n_sweeps);
    for (bool more = n_sweeps > 0; more;) {
      more = false;
      for (size_t i = 0; i < b->groups.size(); ++i) {
        if (left[i] <= 0) continue;
        Group& g = b->groups[i];
        int64_t q = 0;
        if (b->ragged && g.lds) {
          if ((int)i != b->ragged_lead) {
            left[i] = 0;
            continue;
          }
          q = std::min<int64_t>(left[i], b->ragged_per_launch);
          HIP_TRY(ctx, launch_lds_ragged(sets, (int)q, g.stream));
          ++b->launches;
        } else {
          const int rc = group_chunk(ctx, g, left[i], g.stream, &b->launches, &q);
          if (rc != RBIM_OK) return rc;
        }
        left[i] -= q;
        more = more || left[i] > 0;
      }
    }
    for (Group& g : b->groups) {
      const int rc = group_close(ctx, g, g.stream, &b->launches);
      if (rc != RBIM_OK) return rc;
      HIP_TRY(ctx, hipEventRecord(g.done, g.stream));
      HIP_TRY(ctx, hipStreamWaitEvent(b->stream, g.done, 0));
    }
  }
  b->issued += n_sweeps;
  return RBIM_OK;
}

int rbim_batch_run_timed(rbim_batch* b, int64_t n_sweeps, double* elapsed_ms) {
  if (!b) return set_err(nullptr, RBIM_E_INVAL, "batch is NULL");
  rbim_ctx* ctx = b->ctx;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipEventRecord(b->ev0, b->stream));
  int rc = rbim_batch_run(b, n_sweeps);
  if (rc != RBIM_OK) return rc;
  HIP_TRY(ctx, hipEventRecord(b->ev1, b->stream));
  HIP_TRY(ctx, hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  return RBIM_OK;
}

int rbim_batch_query(rbim_batch* b, int32_t* done) {
  if (!b || !done) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  hipError_t e = hipStreamQuery(b->stream);
  if (e == hipSuccess) { *done = 1; return RBIM_OK; }
  if (e == hipErrorNotReady) { *done = 0; return RBIM_OK; }
  return set_err(b->ctx, RBIM_E_HIP, "stream query failed: %s", hipGetErrorString(e));
}

int rbim_batch_sync(rbim_batch* b) {
  if (!b) return set_err(nullptr, RBIM_E_INVAL, "batch is NULL");
  HIP_TRY(b->ctx, hipStreamSynchronize(b->stream));
  return RBIM_OK;
}

int rbim_batch_moments(rbim_batch* b, uint64_t* out) {
  if (!b || !out) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  rbim_ctx* ctx = b->ctx;
  HIP_TRY(ctx, hipStreamSynchronize(b->stream));
  const size_t per = (size_t)b->n_bins * RBIM_ACC_SLOTS;
  for (Group& g : b->groups) {
    std::vector<uint64_t> h(per * g.n_lat);
    HIP_TRY(ctx, hipMemcpy(h.data(), g.set.acc, h.size() * 8, hipMemcpyDeviceToHost));
    for (int k = 0; k < g.n_lat; ++k) std::memcpy(out + per * g.idx[k], &h[per * k], per * 8);
  }
  return RBIM_OK;
}

int rbim_batch_spins(rbim_batch* b, int8_t* out) {
  if (!b || !out) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  rbim_ctx* ctx = b->ctx;
  HIP_TRY(ctx, hipStreamSynchronize(b->stream));
  for (Group& g : b->groups) {
    std::vector<int8_t> h((size_t)g.N * g.n_lat);
    HIP_TRY(ctx, hipMemcpy(h.data(), g.set.spins, h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < g.n_lat; ++k) std::memcpy(out + b->off[g.idx[k]], &h[(size_t)g.N * k], g.N);
  }
  return RBIM_OK;
}

int rbim_batch_info_get(rbim_batch* b, rbim_batch_info* o) {
  if (!b || !o) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  std::memset(o, 0, sizeof *o);
  o->n_groups = (int32_t)b->groups.size();
  for (Group& g : b->groups) {
    if (g.lds) ++o->lds_groups; else ++o->tiled_groups;
    o->device_bytes += (int64_t)g.arena.bytes;
  }
  if (!b->groups.empty()) {
    o->threads = b->groups[0].threads;
    o->sweeps_per_launch = b->groups[0].per_launch;
  }
  o->sweeps_done = b->issued;
  o->total_sweeps = b->total;
  o->kernel_launches = b->launches;
  return RBIM_OK;
}

int rbim_batch_series(rbim_batch* b, rbim_sweep_stats* out, int64_t* recorded) {
  if (!b || !out || !recorded) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  rbim_ctx* ctx = b->ctx;
  if (b->groups.empty() || b->groups[0].set.series == nullptr)
    return set_err(ctx, RBIM_E_STATE, "batch was created without a series ring");
  HIP_TRY(ctx, hipStreamSynchronize(b->stream));
  const int64_t cap = b->groups[0].set.series_cap;
  for (Group& g : b->groups) {
    std::vector<rbim_sweep_stats> h((size_t)g.n_lat * cap);
    HIP_TRY(ctx, hipMemcpy(h.data(), g.set.series, h.size() * sizeof(rbim_sweep_stats),
                           hipMemcpyDeviceToHost));
    for (int k = 0; k < g.n_lat; ++k)
      std::memcpy(out + (size_t)cap * g.idx[k], &h[(size_t)cap * k], cap * sizeof(rbim_sweep_stats));
  }
  *recorded = std::max<int64_t>(0, b->issued - b->n_therm);
  return RBIM_OK;
}

int rbim_batch_destroy(rbim_batch* b) {
  if (!b) return RBIM_OK;
  (void)hipSetDevice(b->ctx->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (hipStream_t st : b->streams) (void)hipStreamSynchronize(st);
  for (Group& g : b->groups) {
    g.stream = nullptr;  // the batch's, destroyed below
    group_free(g);
  }
  for (hipStream_t st : b->streams) (void)hipStreamDestroy(st);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return RBIM_OK;
}

}  // extern "C"

// ============================================================ slot pools
struct rbim_pool {
  rbim_ctx* ctx = nullptr;
  Group g;                      // every slot: n_lat = n_slots
  hipStream_t stream = nullptr;
  hipEvent_t staged = nullptr;  // the last load's host-to-device copy has run
  hipEvent_t mapped = nullptr;  // the last run's slot list copy has run
  int max_bins = 1, chunk = 1;
  int64_t* d_sched = nullptr;   // [n_slots][3]
  int32_t* d_map = nullptr;     // [n_slots] slots of the current run
  uint8_t* d_stage = nullptr;   // one load chunk
  uint8_t* h_stage = nullptr;   // pinned, one load chunk
  int32_t* h_map = nullptr;     // pinned [n_slots]
  size_t stage_bytes = 0;
  RegionTable regions{};
  std::vector<int64_t> done, total;  // host mirror of every slot's progress
  std::vector<uint8_t> loaded;
  int64_t launches = 0;
};

extern "C" {

static int check_pool_desc(rbim_ctx* ctx, const rbim_pool_desc* d) {
  if (!d) return set_err(ctx, RBIM_E_INVAL, "desc is NULL");
  if (!valid_L(d->L)) return set_err(ctx, RBIM_E_INVAL, "L must be at least 3 (got %d)", d->L);
  if (d->n_slots < 1 || d->n_slots > (1 << 20))
    return set_err(ctx, RBIM_E_INVAL, "n_slots must be in 1 .. 2^20 (got %d)", d->n_slots);
  if (d->max_bins < 1) return set_err(ctx, RBIM_E_INVAL, "max_bins must be positive");
  if (d->engine < RBIM_ENGINE_AUTO || d->engine > RBIM_ENGINE_TILED)
    return set_err(ctx, RBIM_E_INVAL, "unknown engine %d", d->engine);
  if (d->engine == RBIM_ENGINE_BATCHED && d->L > RBIM_LDS_MAX_L)
    return set_err(ctx, RBIM_E_INVAL, "the LDS engine holds L <= %d (got %d)", RBIM_LDS_MAX_L, d->L);
  if (d->tile != 0 && d->tile != 32 && d->tile != 64)
    return set_err(ctx, RBIM_E_INVAL, "tile must be 0 (by L), 32 or 64");
  if (d->threads != 0 && d->threads != 64 && d->threads != 128 && d->threads != 256 &&
      d->threads != 512 && d->threads != 1024)
    return set_err(ctx, RBIM_E_INVAL, "threads must be 64, 128, 256, 512 or 1024");
  if (d->sweeps_per_launch < 0) return set_err(ctx, RBIM_E_INVAL, "sweeps_per_launch < 0");
  return RBIM_OK;
}

static void pool_engine(const rbim_pool_desc* d, bool* lds, bool* fused) {
  *lds = d->engine == RBIM_ENGINE_BATCHED || (d->engine == RBIM_ENGINE_AUTO && d->L <= RBIM_LDS_MAX_L);
  *fused = !*lds && d->engine != RBIM_ENGINE_TILED && fused_tile_size(d->L, d->tile) > 0;
}

int rbim_pool_footprint(const rbim_pool_desc* d, uint64_t* bytes) {
  int rc = check_pool_desc(nullptr, d);
  if (rc != RBIM_OK) return rc;
  bool lds, fused;
  pool_engine(d, &lds, &fused);
  const int N = d->L * d->L;
  uint64_t b = group_bytes(d->L, d->n_slots, lds, fused, d->max_bins, 0, d->tile);
  b += (uint64_t)d->n_slots * (24 + 4) + pool_stage_bytes(N, pool_chunk(N, d->n_slots));
  if (bytes) *bytes = b;
  return RBIM_OK;
}

int rbim_pool_create(rbim_ctx* ctx, const rbim_pool_desc* d, rbim_pool** out) {
  if (!ctx || !out) return set_err(ctx, RBIM_E_INVAL, "NULL argument");
  int rc = check_pool_desc(ctx, d);
  if (rc != RBIM_OK) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  uint64_t need = 0;
  rbim_pool_footprint(d, &need);
  size_t fr = 0, tot = 0;
  HIP_TRY(ctx, hipMemGetInfo(&fr, &tot));
  if (need + (64ull << 20) > fr)
    return set_err(ctx, RBIM_E_ADMIT, "pool needs %llu bytes, %llu free", (unsigned long long)need,
                   (unsigned long long)fr);
  rbim_pool* p = new rbim_pool();
  p->ctx = ctx;
  p->max_bins = d->max_bins;
  Group& g = p->g;
  g.L = d->L;
  g.N = d->L * d->L;
  g.n_lat = d->n_slots;
  pool_engine(d, &g.lds, &g.fused);
  g.tile_pref = d->tile;
  g.threads = d->threads ? d->threads : default_threads(g.L);
  // tiled pools launch sweep by sweep: their grids follow the loaded slots
  g.per_launch = g.lds ? (d->sweeps_per_launch > 0 ? d->sweeps_per_launch : default_lds_sweeps(g.L)) : 1;
  p->done.assign(d->n_slots, 0);
  p->total.assign(d->n_slots, 0);
  p->loaded.assign(d->n_slots, 0);
  p->chunk = pool_chunk(g.N, d->n_slots);
  p->stage_bytes = pool_stage_bytes(g.N, p->chunk);
  auto fail = [&](int code) { rbim_pool_destroy(p); return code; };
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->staged, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->mapped, hipEventDisableTiming);
  if (e != hipSuccess)
    return fail(set_err(ctx, RBIM_E_HIP, "stream/event creation failed: %s", hipGetErrorString(e)));
  rc = group_alloc(ctx, g, {}, {}, {}, {}, nullptr, RBIM_START_HOT, 0, 1, d->max_bins, 0, 0, true);
  if (rc != RBIM_OK) return fail(rc);
  void* q;
  e = g.arena.alloc(&q, (size_t)d->n_slots * 24);
  if (e == hipSuccess) {
    p->d_sched = (int64_t*)q;
    e = hipMemsetAsync(q, 0, (size_t)d->n_slots * 24, ctx->stream);  // total 0: inactive
  }
  if (e == hipSuccess) e = g.arena.alloc(&q, (size_t)d->n_slots * 4);
  if (e == hipSuccess) {
    p->d_map = (int32_t*)q;
    e = g.arena.alloc(&q, p->stage_bytes);
  }
  if (e == hipSuccess) {
    p->d_stage = (uint8_t*)q;
    e = hipHostMalloc((void**)&p->h_stage, p->stage_bytes, hipHostMallocDefault);
  }
  if (e == hipSuccess) e = hipHostMalloc((void**)&p->h_map, (size_t)d->n_slots * 4, hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess)
    return fail(set_err(ctx, e == hipErrorOutOfMemory ? RBIM_E_OOM : RBIM_E_HIP,
                        "pool allocation failed: %s", hipGetErrorString(e)));
  g.set.sched = p->d_sched;
  // the per-slot state a load clears
  RegionTable& rt = p->regions;
  const uint64_t nl = (uint64_t)d->n_slots, N = (uint64_t)g.N;
  auto add = [&](void* base, uint64_t per_lat, int planes = 1, uint64_t plane_stride = 0) {
    if (base && per_lat) rt.r[rt.n++] = Region{(uint8_t*)base, per_lat, plane_stride, planes};
  };
  add(g.set.acc, (uint64_t)d->max_bins * RBIM_ACC_SLOTS * 8);
  add(g.set.last, sizeof(rbim_sweep_stats));
  if (g.fused) {
    FusedScratch& f = g.fz;
    const uint64_t tiles = (uint64_t)f.tiles_x * f.tiles_x, ns = (uint64_t)f.nslots;
    add(f.st, 2 * N, 2, nl * N * 2);
    add(f.jt, tiles * f.tile * f.tile / 4);
    add(f.slots, ns * 8);
    add(f.guf, ns * 8);
    add(f.gcnt, ns * 8);
    add(f.gmn, ns * 4);
    add(f.gpart, (ns + 255) / 256 * 5 * 8);
    add(f.wrapw, 4);
    add(f.xbits, tiles * (2 * f.tile / 64) * 8);
    add(f.flipbits, ns / 8);
    add(f.part_int, tiles * 5 * 8, 2, nl * tiles * 5 * 8);
    add(f.part_em, tiles * 2 * 8);
    add(f.bst, 6 * 8, 2, nl * 6 * 8);
  } else if (!g.lds) {
    TiledScratch& t = g.tile;
    add(t.part, (uint64_t)t.tiles_x * t.tiles_y * 7 * 8);
    add(t.wrapf, 8);
    add(t.pending, 4);
  }
  *out = p;
  return RBIM_OK;
}

int rbim_pool_load(rbim_pool* p, int32_t n, const int32_t* slots, const int8_t* Jh,
                   const int8_t* Jv, const uint32_t* thr31, const uint64_t* seed,
                   const uint32_t* lattice_id, const int8_t* spins0, int32_t start,
                   int64_t n_therm, int32_t n_bins, int64_t sweeps_per_bin) {
  if (!p) return set_err(nullptr, RBIM_E_INVAL, "pool is NULL");
  rbim_ctx* ctx = p->ctx;
  if (n < 1 || !slots || !Jh || !Jv || !thr31 || !seed || !lattice_id)
    return set_err(ctx, RBIM_E_INVAL, "NULL argument or n < 1");
  if (n_bins < 1 || n_bins > p->max_bins)
    return set_err(ctx, RBIM_E_INVAL, "n_bins must be in 1 .. %d (got %d)", p->max_bins, n_bins);
  if (sweeps_per_bin < 1) return set_err(ctx, RBIM_E_INVAL, "sweeps_per_bin must be at least 1");
  if (n_therm < 0) return set_err(ctx, RBIM_E_INVAL, "n_therm must be non-negative");
  const int64_t total = n_therm + (int64_t)n_bins * sweeps_per_bin;
  if (total > 0xFFFFFFFFll) return set_err(ctx, RBIM_E_INVAL, "more than 2^32 sweeps per lattice");
  const int n_slots = p->g.n_lat;
  for (int k = 0; k < n; ++k) {
    if (slots[k] < 0 || slots[k] >= n_slots)
      return set_err(ctx, RBIM_E_INVAL, "slot %d outside [0, %d)", slots[k], n_slots);
    if (p->loaded[slots[k]])
      return set_err(ctx, RBIM_E_STATE, "slot %d is in use (release it first)", slots[k]);
    if (thr31[k] > (1u << 31)) return set_err(ctx, RBIM_E_INVAL, "thr31 exceeds 2^31");
  }
  int dup = -1;  // every slot is free (0) here: mark them to find repeats
  for (int k = 0; k < n && dup < 0; ++k) {
    if (p->loaded[slots[k]] == 2) dup = slots[k];
    p->loaded[slots[k]] = 2;
  }
  for (int k = 0; k < n; ++k) p->loaded[slots[k]] = 0;
  if (dup >= 0) return set_err(ctx, RBIM_E_INVAL, "slot %d given twice", dup);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int N = p->g.N;
  const int C = p->chunk;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const int max_region = [&] {
    uint64_t m = 0;
    for (int i = 0; i < p->regions.n; ++i) m = std::max(m, p->regions.r[i].per_lat);
    return (int)std::min<uint64_t>(64, (m / 4 + 255) / 256 + 1);
  }();
  for (int k0 = 0; k0 < n; k0 += C) {
    const int m = std::min(C, n - k0);
    // staging layout of m lattices: jbits [m][N] | spins [m][N] | thr | seed | lid | slots
    const size_t o_sp = al((size_t)m * N), o_thr = o_sp + (spins0 ? al((size_t)m * N) : 0),
                 o_seed = o_thr + al(4 * (size_t)m), o_lid = o_seed + al(8 * (size_t)m),
                 o_slot = o_lid + al(4 * (size_t)m), bytes = o_slot + 4 * (size_t)m;
    HIP_TRY(ctx, hipEventSynchronize(p->staged));  // the previous chunk's copy has run
    uint8_t* h = p->h_stage;
    for (int k = 0; k < m; ++k) {
      const size_t o = (size_t)(k0 + k) * N;
      pack_jbits(Jh + o, Jv + o, h + (size_t)k * N, N);
      if (spins0) std::memcpy(h + o_sp + (size_t)k * N, spins0 + o, N);
    }
    std::memcpy(h + o_thr, thr31 + k0, 4 * (size_t)m);
    std::memcpy(h + o_seed, seed + k0, 8 * (size_t)m);
    std::memcpy(h + o_lid, lattice_id + k0, 4 * (size_t)m);
    std::memcpy(h + o_slot, slots + k0, 4 * (size_t)m);
    HIP_TRY(ctx, hipMemcpyAsync(p->d_stage, h, bytes, hipMemcpyHostToDevice, p->stream));
    HIP_TRY(ctx, hipEventRecord(p->staged, p->stream));
    const int32_t* d_slots = (const int32_t*)(p->d_stage + o_slot);
    hipLaunchKernelGGL(pool_reset_kernel, dim3(max_region, m), dim3(256), 0, p->stream, p->regions,
                       d_slots);
    HIP_TRY(ctx, hipGetLastError());
    PoolStage st;
    st.jbits = p->d_stage;
    st.spins = spins0 ? (const int8_t*)(p->d_stage + o_sp) : nullptr;
    st.thr = (const uint32_t*)(p->d_stage + o_thr);
    st.seed = (const uint64_t*)(p->d_stage + o_seed);
    st.lid = (const uint32_t*)(p->d_stage + o_lid);
    st.slots = d_slots;
    hipLaunchKernelGGL(pool_load_kernel, dim3((N + 255) / 256, m), dim3(256), 0, p->stream, p->g.set,
                       p->d_sched, p->g.fused ? p->g.fz.fin : nullptr,
                       p->g.fused ? p->g.fz.cur : nullptr, st, (int)start, n_therm,
                       sweeps_per_bin, total);
    HIP_TRY(ctx, hipGetLastError());
  }
  for (int k = 0; k < n; ++k) {
    p->loaded[slots[k]] = 1;
    p->done[slots[k]] = 0;
    p->total[slots[k]] = total;
  }
  return RBIM_OK;
}

int rbim_pool_run(rbim_pool* p, int64_t n_sweeps) {
  if (!p) return set_err(nullptr, RBIM_E_INVAL, "pool is NULL");
  rbim_ctx* ctx = p->ctx;
  if (n_sweeps < 0) return set_err(ctx, RBIM_E_INVAL, "n_sweeps must be non-negative");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipEventSynchronize(p->mapped));  // h_map is free again
  int cnt = 0;
  int64_t longest = 0;
  for (int i = 0; i < p->g.n_lat; ++i)
    if (p->loaded[i] && p->done[i] < p->total[i]) {
      p->h_map[cnt++] = i;
      longest = std::max(longest, p->total[i] - p->done[i]);
    }
  const int64_t n = std::min(n_sweeps, longest);
  if (cnt == 0 || n == 0) return RBIM_OK;
  HIP_TRY(ctx, hipMemcpyAsync(p->d_map, p->h_map, 4 * (size_t)cnt, hipMemcpyHostToDevice, p->stream));
  HIP_TRY(ctx, hipEventRecord(p->mapped, p->stream));
  p->g.set.lat_map = p->d_map;
  p->g.set.n_run = cnt;
  const int rc = group_run(ctx, p->g, n, p->stream, &p->launches);
  if (rc != RBIM_OK) return rc;
  for (int k = 0; k < cnt; ++k) {
    const int i = p->h_map[k];
    p->done[i] = std::min(p->done[i] + n, p->total[i]);
  }
  return RBIM_OK;
}

int rbim_pool_query(rbim_pool* p, int32_t* done) {
  if (!p || !done) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  hipError_t e = hipStreamQuery(p->stream);
  if (e == hipSuccess) { *done = 1; return RBIM_OK; }
  if (e == hipErrorNotReady) { *done = 0; return RBIM_OK; }
  return set_err(p->ctx, RBIM_E_HIP, "stream query failed: %s", hipGetErrorString(e));
}

int rbim_pool_sync(rbim_pool* p) {
  if (!p) return set_err(nullptr, RBIM_E_INVAL, "pool is NULL");
  HIP_TRY(p->ctx, hipStreamSynchronize(p->stream));
  return RBIM_OK;
}

int rbim_pool_progress(rbim_pool* p, int32_t n, const int32_t* slots, int64_t* done,
                       int64_t* total) {
  if (!p || !slots || n < 0) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  for (int k = 0; k < n; ++k) {
    if (slots[k] < 0 || slots[k] >= p->g.n_lat)
      return set_err(p->ctx, RBIM_E_INVAL, "slot %d outside [0, %d)", slots[k], p->g.n_lat);
    if (done) done[k] = p->loaded[slots[k]] ? p->done[slots[k]] : -1;
    if (total) total[k] = p->loaded[slots[k]] ? p->total[slots[k]] : -1;
  }
  return RBIM_OK;
}

int rbim_pool_read(rbim_pool* p, int32_t n, const int32_t* slots, uint64_t* moments, int8_t* spins) {
  if (!p || !slots || n < 0) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  rbim_ctx* ctx = p->ctx;
  int lo = p->g.n_lat, hi = -1;
  for (int k = 0; k < n; ++k) {
    if (slots[k] < 0 || slots[k] >= p->g.n_lat || !p->loaded[slots[k]])
      return set_err(ctx, RBIM_E_INVAL, "slot %d is not loaded", slots[k]);
    lo = std::min(lo, slots[k]);
    hi = std::max(hi, slots[k]);
  }
  if (n == 0) return RBIM_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(p->stream));
  const size_t per = (size_t)p->max_bins * RBIM_ACC_SLOTS;
  if (moments) {
    std::vector<uint64_t> h(per * (size_t)(hi - lo + 1));
    HIP_TRY(ctx, hipMemcpy(h.data(), p->g.set.acc + per * lo, h.size() * 8, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; ++k) std::memcpy(moments + per * k, &h[per * (slots[k] - lo)], per * 8);
  }
  if (spins)
    for (int k = 0; k < n; ++k)
      HIP_TRY(ctx, hipMemcpy(spins + (size_t)k * p->g.N, p->g.set.spins + (size_t)slots[k] * p->g.N,
                             p->g.N, hipMemcpyDeviceToHost));
  return RBIM_OK;
}

int rbim_pool_release(rbim_pool* p, int32_t n, const int32_t* slots) {
  if (!p || !slots || n < 0) return set_err(nullptr, RBIM_E_INVAL, "NULL argument");
  for (int k = 0; k < n; ++k)
    if (slots[k] < 0 || slots[k] >= p->g.n_lat)
      return set_err(p->ctx, RBIM_E_INVAL, "slot %d outside [0, %d)", slots[k], p->g.n_lat);
  for (int k = 0; k < n; ++k) p->loaded[slots[k]] = 0;
  return RBIM_OK;
}

int rbim_pool_destroy(rbim_pool* p) {
  if (!p) return RBIM_OK;
  (void)hipSetDevice(p->ctx->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  group_free(p->g);
  if (p->h_stage) (void)hipHostFree(p->h_stage);
  if (p->h_map) (void)hipHostFree(p->h_map);
  if (p->staged) (void)hipEventDestroy(p->staged);
  if (p->mapped) (void)hipEventDestroy(p->mapped);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return RBIM_OK;
}

int rbim_inject_alloc_failure(int64_t after) {
  g_alloc_countdown.store(after < 0 ? -1 : after);
  return RBIM_OK;
}

}  // extern "C"
