// Internal declarations shared by the gfx950 kernels and the C ABI.
#pragma once
#include <atomic>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rbim_hip.h"
#include "philox.h"

// Largest lattice the LDS-resident kernel holds: 14-bit parent field.
#define RBIM_LDS_MAX_L 128
#define RBIM_LDS_MAX_N (RBIM_LDS_MAX_L * RBIM_LDS_MAX_L)

// Per-lattice state common to both engines (device pointers).
struct LatticeSet {
  int32_t L, N, n_lat;
  uint32_t magic;           // ceil(2^32 / L): x = umulhi(n, magic) for n < 2^20
  int8_t* spins;            // [n_lat][N] +-1
  const uint8_t* jbits;     // [n_lat][N] bit0: J_h < 0, bit1: J_v < 0
  const uint32_t* thr31;    // [n_lat]
  const uint64_t* seed;     // [n_lat]
  const uint32_t* lat_id;   // [n_lat]
  int64_t* sweep_ctr;       // [n_lat] sweeps completed
  uint64_t* acc;            // [n_lat][n_bins][RBIM_ACC_SLOTS]
  rbim_sweep_stats* last;   // [n_lat] stats of the last measured sweep
  rbim_sweep_stats* series; // [n_lat][series_cap] ring of measured sweeps, or NULL
  int64_t series_cap;
  int64_t n_therm, spb, n_bins, total;
  // Slot pools (requests of different schedules sharing launches): per
  // lattice n_therm, sweeps per bin and total sweeps, [n_lat][3], or NULL when
  // the scalars above hold for every lattice (n_bins then only strides acc).
  const int64_t* sched;
  // Lattices a launch covers: lat_map[0 .. n_run) (slot indices), or NULL for
  // all n_lat (n_run == n_lat).  Grids span n_run, never n_lat.
  const int32_t* lat_map;
  int32_t n_run;
  // given-bits mode (one-lattice parity entry points); NULL in production
  const uint8_t* given_occ; // [N] bit0 horizontal, bit1 vertical
  const uint8_t* given_flip;// [N]
  int32_t* canon_out;       // [N] min-site label of each site
};

// Tiled-path scratch (device pointers), [n_lat][N] unless noted.  Only
// occ and lroot are written for every site; gf, lst and st are used at
// tile-local roots only.
struct TiledScratch {
  uint8_t* occ;        // bit0 occ_h, bit1 occ_v, bit2 flip
  uint16_t* lroot;     // tile-local root of every site, as a tile index lx * 32 + ly
  uint64_t* gf;        // forest over local roots: parent (low 32) | wx:16 | wy:16
  uint32_t* lst;       // local cluster: size << 16 | n_plus before the flip
  uint64_t* st;        // cluster: size << 32 | n_plus after the flip, at roots
  int64_t* part;       // [n_lat][n_tiles][7] partial measurements of one sweep
  uint32_t* wrapf;     // [n_lat][2] wrap flags by sweep parity
  int32_t* pending;    // [n_lat] the last completed sweep still awaits its measurement
  int32_t tile;        // tile side: 64 when L is a multiple of 64 (L >= 128), else 32
  int32_t tiles_x, tiles_y;
};

// Launchers (stream-ordered, no host synchronisation).
hipError_t launch_init_spins(const LatticeSet& s, int start, hipStream_t st);
hipError_t launch_lds_sweeps(const LatticeSet& s, int threads, int sweeps, hipStream_t st);
size_t lds_bytes(int N);
// One launch over LDS sets of L = 16, 32, 64, 128 (sets[k] of side 16 << k, or
// NULL): pages of 16384 sites per 1024-thread workgroup.  ragged_slot(L): k,
// or -1 for a side it does not hold.
int ragged_slot(int L);
int ragged_threads(int L);  // threads per lattice in its page
hipError_t launch_lds_ragged(const LatticeSet* const sets[4], int sweeps, hipStream_t st);
// Workgroups of the LDS kernel one CU holds at once (0 if unknown).
int lds_blocks_per_cu(int L, int threads);
// Two-kernel sweep for lattices of several full tiles whose tile-edge slots fit
// one workgroup's LDS (fused_engine.hip); device pointers, [n_lat][N] unless noted.
#define RBIM_FUSED_MAX_SLOTS 6400  // slots merged in one workgroup's LDS
struct FusedScratch {
  uint16_t* st;        // [2][n_lat][N] state words: spin << 15 | edge << 14 | the rep slot of
                       // the site's edge cluster (edge clusters' flips are pending)
  uint32_t* jt;        // [n_lat][tiles][T*T/16] couplings tile by tile, 2 bits a site
                       // (J_h < 0, J_v < 0) at local index lx * T + ly (0 where absent)
  uint64_t* slots;     // [n_lat][nslots] edge-slot entries of the last tile pass
  uint64_t* xbits;     // [n_lat][tiles][2T/64] occupied bonds leaving each tile (+x, +y)
  uint32_t* flipbits;  // [n_lat][nslots/32] flip bit of each slot's cluster
  int64_t* part_int;   // [2][n_lat][tiles][5] interior clusters: n, key1, key2, n_plus1, n_plus2
  int64_t* part_em;    // [n_lat][tiles][2] E, M of the state a tile pass started from
  int64_t* bst;        // [2][n_lat][6] edge clusters: n, key1, key2, n_plus1, n_plus2, wrap bits
  int64_t* fin;        // [n_lat] first sweep not yet folded into the moments
  int32_t* cur;        // [n_lat] buffer (0 / 1) holding the current state
  // slot merge in HBM (big != 0: the slots do not fit one workgroup's LDS)
  uint64_t* guf;       // [n_lat][nslots] slot forest: parent | wx:16 | wy:16
  uint64_t* gcnt;      // [n_lat][nslots] size << 32 | n_plus before the flip, at reps
  uint32_t* gmn;       // [n_lat][nslots] min site << 1 | its flip bit, at reps (else ~0)
  int64_t* gpart;      // [n_lat][nslots / 256][5] merged-cluster partials
  uint32_t* wrapw;     // [n_lat] wrap bits of the sweep
  int32_t tile, tiles_x, nslots, big;
};
// Tile side for L >= 3 (pref: 32, or 64 for L >= 64; else 64 for L >= 64,
// 32 below), the last tile row and column cut when the side does not divide
// L (below 32: the lattice is one cut tile); 0 for L < 3.
int fused_tile_size(int L, int pref = 0);
int fused_tiles_per_side(int L, int T);
bool fused_merge_in_hbm(int L, int T);
int fused_kernels_per_sweep(const LatticeSet& s, const FusedScratch& f);
// Before the first sweep of a run: the state words from the batch's spins.
hipError_t launch_fused_open(const LatticeSet& s, const FusedScratch& f, hipStream_t st);
hipError_t launch_fused_sweep(const LatticeSet& s, const FusedScratch& f, hipStream_t st);
// Folds the last sweep of a run and applies its pending flips into buffer 0.
hipError_t launch_fused_close(const LatticeSet& s, const FusedScratch& f, hipStream_t st);

hipError_t launch_tiled_sweep(const LatticeSet& s, const TiledScratch& t, hipStream_t st);
// Measures and folds the pending sweep at the end of a run of tiled sweeps.
hipError_t launch_tiled_flush(const LatticeSet& s, const TiledScratch& t, hipStream_t st);
hipError_t launch_label_rank(const int32_t* canon, int32_t N, int32_t* labels, int32_t* rank_tmp,
                             int32_t* n_clusters, hipStream_t st);
hipError_t launch_cluster_stats(const int32_t* labels, const int8_t* spins, int64_t n,
                                unsigned long long* sizes, unsigned long long* mags,
                                hipStream_t st);
hipError_t launch_energy_mag(const int8_t* spins, const uint8_t* jbits, int32_t L,
                             unsigned long long* em, hipStream_t st);
int tiled_kernels_per_sweep();
int tiled_tile_size(int L, int pref = 0);

// ---------------------------------------------------------------- helpers
// Lattice of grid index i (blockIdx of a per-lattice grid dimension).
__device__ __forceinline__ int run_lat(const LatticeSet& s, int i) {
  return s.lat_map ? s.lat_map[i] : i;
}
// Schedule of lattice lat.
__device__ __forceinline__ int64_t lat_therm(const LatticeSet& s, int lat) {
  return s.sched ? s.sched[3 * (size_t)lat] : s.n_therm;
}
__device__ __forceinline__ int64_t lat_spb(const LatticeSet& s, int lat) {
  return s.sched ? s.sched[3 * (size_t)lat + 1] : s.spb;
}
__device__ __forceinline__ int64_t lat_total(const LatticeSet& s, int lat) {
  return s.sched ? s.sched[3 * (size_t)lat + 2] : s.total;
}

__device__ __forceinline__ uint32_t site_x(uint32_t n, uint32_t magic) {
  return __umulhi(n, magic);
}

// Accumulate one measured sweep into a 20-slot bin (single thread).
__device__ __forceinline__ void acc_add_u128(uint64_t* lo, uint64_t* hi, uint64_t vlo,
                                             uint64_t vhi) {
  const uint64_t nlo = *lo + vlo;
  *hi += vhi + (nlo < vlo ? 1ull : 0ull);
  *lo = nlo;
}

// Raises a kernel's dynamic-LDS limit once per device (the attribute is per
// device; `done` is the kernel's own bitmask of devices already configured).
inline hipError_t ensure_dynamic_lds(const void* kernel, int bytes, std::atomic<uint64_t>& done) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const uint64_t bit = 1ull << (dev & 63);
  if (done.load(std::memory_order_acquire) & bit) return hipSuccess;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess) done.fetch_or(bit, std::memory_order_release);
  return e;
}

// Measured sweep gm of lattice lat into its series ring (when kept).
__device__ __forceinline__ void series_put(const LatticeSet& s, int lat, int64_t gm,
                                           const rbim_sweep_stats& st) {
  if (s.series) s.series[(size_t)lat * s.series_cap + (gm - lat_therm(s, lat)) % s.series_cap] = st;
}

__device__ __forceinline__ void acc_sweep(uint64_t* a, const rbim_sweep_stats& s) {
  const int64_t E = s.energy, M = s.magnetization;
  const uint64_t aE = (uint64_t)(E < 0 ? -E : E);
  const uint64_t aM = (uint64_t)(M < 0 ? -M : M);
  a[RBIM_ACC_E] = (uint64_t)((int64_t)a[RBIM_ACC_E] + E);
  acc_add_u128(&a[RBIM_ACC_E2_LO], &a[RBIM_ACC_E2_HI], aE * aE, __umul64hi(aE, aE));
  a[RBIM_ACC_M] = (uint64_t)((int64_t)a[RBIM_ACC_M] + M);
  a[RBIM_ACC_ABSM] += aM;
  const uint64_t m2 = aM * aM;  // |M| <= 2^20: M^2 < 2^40 exact
  acc_add_u128(&a[RBIM_ACC_M2_LO], &a[RBIM_ACC_M2_HI], m2, 0ull);
  acc_add_u128(&a[RBIM_ACC_M4_LO], &a[RBIM_ACC_M4_HI], m2 * m2, __umul64hi(m2, m2));
  a[RBIM_ACC_SMAX] += (uint64_t)s.largest_size;
  a[RBIM_ACC_ABSMC] += (uint64_t)s.largest_abs_mag;
  a[RBIM_ACC_S2] += (uint64_t)s.second_size;
  const double ratio =
      s.largest_size > 0 ? (double)s.largest_abs_mag / (double)s.largest_size : 0.0;
  a[RBIM_ACC_RATIO] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)a[RBIM_ACC_RATIO]) + ratio);
  a[RBIM_ACC_NCL] += (uint64_t)s.n_clusters;
  a[RBIM_ACC_WX] += (uint64_t)(s.wrap_x != 0);
  a[RBIM_ACC_WY] += (uint64_t)(s.wrap_y != 0);
  a[RBIM_ACC_WANY] += (uint64_t)((s.wrap_x | s.wrap_y) != 0);
  a[RBIM_ACC_COUNT] += 1ull;
}
