/*
 * rbim_hip.h -- C ABI of librbim_hip.so, the MI355X (gfx950) Swendsen-Wang
 * engine for the 2D +-J random-bond Ising model.
 *
 * The reference (Amirhossein-Manouchehri/rbim-fk-percolation) has no FFI: its
 * operator boundary is a set of Python functions, and its only compiled code is
 * the numba-jitted union-find.  Each entry point below replaces one of those
 * seams; the Python package `rbim` (rbim-fk-percolation_amd/rbim) binds them
 * with ctypes and keeps the reference's Python signatures on top.
 *
 * Conventions
 *   - Every function returns an int status: RBIM_OK (0) or a negative
 *     RBIM_E_* code; rbim_last_error() then holds a message.  No C++
 *     exception crosses the ABI.
 *   - Host buffers are owned by the caller and must be C-contiguous with the
 *     exact element types named.  Device memory is owned by the library and
 *     hidden behind rbim_ctx / rbim_batch handles.
 *   - A handle is used by one host thread at a time.
 *   - Geometry follows src/rbim/lattice.py:1-24 of the reference: site
 *     n = x*L + y (x slow); bond 2n joins n to (x+1, y), bond 2n+1 joins n to
 *     (x, y+1); J_h[x, y] sits on bond 2n and J_v[x, y] on bond 2n+1.
 *   - Random draws come from Philox4x32-10 keyed by the 64-bit lattice seed,
 *     counter (site>>1, sweep, lattice_id, 0); results never depend on launch
 *     shape, engine or batch composition.
 */
#ifndef RBIM_HIP_H
#define RBIM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBIM_ABI_VERSION 3

/* Moment slots per (lattice, bin); sums over the measured sweeps of a bin.
 * 128-bit sums are stored as (lo, hi) uint64 pairs.                       */
#define RBIM_ACC_SLOTS 20
enum rbim_acc_slot {
  RBIM_ACC_E = 0,          /* int64  sum E                     */
  RBIM_ACC_E2_LO = 1,      /* uint128 sum E^2 (lo, hi)          */
  RBIM_ACC_E2_HI = 2,
  RBIM_ACC_M = 3,          /* int64  sum M                     */
  RBIM_ACC_ABSM = 4,       /* uint64 sum |M|                   */
  RBIM_ACC_M2_LO = 5,      /* uint128 sum M^2                  */
  RBIM_ACC_M2_HI = 6,
  RBIM_ACC_M4_LO = 7,      /* uint128 sum M^4                  */
  RBIM_ACC_M4_HI = 8,
  RBIM_ACC_SMAX = 9,       /* uint64 sum S_max                 */
  RBIM_ACC_ABSMC = 10,     /* uint64 sum |M_C| of the largest  */
  RBIM_ACC_S2 = 11,        /* uint64 sum S_2                   */
  RBIM_ACC_RATIO = 12,     /* double sum |M_C|/S_max (bits)    */
  RBIM_ACC_NCL = 13,       /* uint64 sum n_clusters            */
  RBIM_ACC_WX = 14,        /* uint64 count wrap_x              */
  RBIM_ACC_WY = 15,        /* uint64 count wrap_y              */
  RBIM_ACC_WANY = 16,      /* uint64 count wrap_x || wrap_y    */
  RBIM_ACC_COUNT = 17      /* uint64 measured sweeps in the bin */
};

enum rbim_status {
  RBIM_OK = 0,
  RBIM_E_INVAL = -1, /* bad argument (the reference raises ValueError)   */
  RBIM_E_HIP = -2,   /* HIP runtime error                                */
  RBIM_E_OOM = -3,   /* device allocation failed                         */
  RBIM_E_ADMIT = -4, /* batch does not fit the device memory budget      */
  RBIM_E_STATE = -5  /* handle in the wrong state (e.g. run past the end) */
};

enum rbim_engine {
  RBIM_ENGINE_AUTO = 0,    /* a batch's lattices of side L: the LDS-resident
                              kernel when L <= 64, or L <= 128 and they hold
                              fewer than 2^23 sites; else as PLAIN.  A slot
                              pool: the LDS kernel when L <= 128.  The
                              one-lattice entry points: as PLAIN              */
  RBIM_ENGINE_PLAIN = 1,   /* tiled path, sweep by sweep: the edge-slot sweep
                              on tiles of 32 or 64, the last tile row and column
                              cut where the side does not divide L (slots merged
                              in LDS up to 6400 tile-edge slots, in HBM beyond) */
  RBIM_ENGINE_BATCHED = 2, /* LDS-resident persistent kernel (L <= 128)      */
  RBIM_ENGINE_TILED = 3    /* always the multi-kernel HBM-forest sweep        */
};

enum rbim_start { RBIM_START_HOT = 0, RBIM_START_COLD = 1 };

typedef struct rbim_ctx rbim_ctx;
typedef struct rbim_batch rbim_batch;

/* Observables of one sweep (src/rbim/swendsen_wang.py:24-53 SweepResult).
 * "largest" / "second" order clusters by size, ties by the smaller minimum
 * site index. */
typedef struct {
  int64_t energy, magnetization, n_clusters;
  int64_t largest_size, largest_abs_mag, second_size, second_abs_mag;
  int32_t wrap_x, wrap_y;
} rbim_sweep_stats;

/* A batch of independent lattices advanced together.  Lattice i has L[i]
 * sites per side; its couplings are the L[i]^2 entries of Jh / Jv starting at
 * sum_{j<i} L[j]^2 (x-major, +-1 int8).  Replaces the task list that
 * run_temperature_scan farms out to processes (src/rbim/simulation.py:384-399). */
typedef struct {
  int32_t n_lattices;
  const int32_t* L;
  const int8_t* Jh;
  const int8_t* Jv;
  const uint32_t* thr31;      /* round((1 - exp(-2/T)) * 2^31) per lattice       */
  const uint64_t* seed;       /* Philox key per lattice                           */
  const uint32_t* lattice_id; /* Philox counter word per lattice                  */
  const int8_t* spins0;       /* initial spins (concatenated) or NULL -> start    */
  int32_t start;              /* rbim_start, used when spins0 == NULL             */
  int32_t n_bins;             /* >= 1                                             */
  int64_t sweeps_per_bin;     /* >= 1                                             */
  int64_t n_therm;            /* >= 0                                             */
  int32_t engine;             /* rbim_engine                                      */
  int32_t sweeps_per_launch;  /* LDS kernel sweeps per launch / tiled sweeps per
                                 graph; 0 -> default                              */
  int32_t threads;            /* LDS kernel workgroup size (64..1024); 0 -> default:
                                 per L, doubled for L <= 32 while the batch would give
                                 the GPU's SIMDs fewer than four wavefronts each    */
  int32_t streams;            /* parts each set of lattices of one L is split into
                                 (parts run concurrently, on at most 4 streams per
                                 batch -- the device's hardware queues -- the largest
                                 parts first on streams of their own); 0 -> default:
                                 2 for tiled sets (L > 128) other than whole 64 x 64
                                 tiles of the edge-slot sweep, and for LDS sets
                                 that fill the GPU more than once with a partial
                                 last round; else 1                               */
  int32_t series;             /* > 0: keep the per-sweep stats of the last `series`
                                 measured sweeps of every lattice in a device ring
                                 (time series for autocorrelation / equilibration
                                 checks); 0 -> off                                */
  int32_t tile;               /* tile side of the tiled engines: 32, or 64 for L >= 64
                                 (the edge-slot sweep cuts the last tile row and
                                 column when the side does not divide L); 0 -> 64
                                 for L >= 64, 32 below                            */
  int32_t single_launch;      /* 1: the LDS-resident lattices (every side one of
                                 16, 32, 64, 128) run in one launch of 1024-thread
                                 workgroups, each a page of lattices of one side
                                 (1 of L = 128, 2 of 64, 4 of 32, 16 of 16); 0 -> one
                                 launch sequence per L, each on its own stream    */
} rbim_batch_desc;

typedef struct {
  int32_t n_groups;           /* lattices are grouped by (L, engine), tiled
                                 groups split further; one stream per group    */
  int32_t lds_groups, tiled_groups;
  int32_t threads;            /* LDS kernel workgroup size of the first group  */
  int32_t sweeps_per_launch;
  int64_t device_bytes;       /* device memory held by the batch               */
  int64_t sweeps_done;        /* sweeps issued so far (all lattices in step)   */
  int64_t total_sweeps;       /* n_therm + n_bins * sweeps_per_bin             */
  int64_t kernel_launches;    /* launches issued so far                        */
} rbim_batch_info;

/* ------------------------------------------------------------ context */
int rbim_abi_version(void);
/* Opens device `device`; fails loudly (RBIM_E_HIP) when no GPU is present. */
int rbim_open(int device, rbim_ctx** out);
int rbim_close(rbim_ctx* ctx);
/* Message of the last failure on ctx (ctx == NULL: last failure anywhere). */
const char* rbim_last_error(const rbim_ctx* ctx);
int rbim_device_memory(rbim_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* ------------------------------------------- one-lattice entry points */
/* Replaces label_and_wrap / _label_and_wrap (src/rbim/clusters.py:144-159,
 * src/rbim/_unionfind.py:91-136).  active: 2N bytes in bond order.
 * labels (N, or NULL): 0..n_clusters-1 numbered in increasing min-site order,
 * as the reference numbers them.  canon (N, or NULL): min site index of each
 * site's cluster.  wrap_xy[2]: exact winding flags.  engine: PLAIN (tiled
 * kernels, any L >= 3) or BATCHED (LDS kernel, L <= 128). */
int rbim_label_and_wrap(rbim_ctx* ctx, int32_t L, const uint8_t* active, int32_t engine,
                        int32_t* labels, int32_t* canon, int32_t* n_clusters,
                        int32_t* wrap_xy);

/* Replaces activate_bonds (src/rbim/clusters.py:48-79) with the engine's
 * Philox stream.  active: 2N bytes out; flipbits (N, or NULL): flip bit of the
 * cluster whose min site is n. */
int rbim_activate(rbim_ctx* ctx, int32_t L, const int8_t* spins, const int8_t* Jh,
                  const int8_t* Jv, uint32_t thr31, uint64_t seed, uint32_t lattice_id,
                  uint32_t sweep, uint8_t* active, uint8_t* flipbits);

/* One SW update with given occupation and flip bits (src/rbim/swendsen_wang.py:
 * 137-172 after activate_bonds): label, flip clusters whose min site m has
 * flip_by_minsite[m] != 0, then measure.  spins (N, int8 +-1) updated in place;
 * labels (N, or NULL) as in rbim_label_and_wrap. */
int rbim_sweep_given_bits(rbim_ctx* ctx, int32_t L, int8_t* spins, const int8_t* Jh,
                          const int8_t* Jv, const uint8_t* active,
                          const uint8_t* flip_by_minsite, int32_t engine, int32_t* labels,
                          rbim_sweep_stats* out);

/* One full SW update of one lattice (sw_sweep, src/rbim/swendsen_wang.py:103-172)
 * with Philox draws for sweep index `sweep`.  out is written when measure != 0. */
int rbim_sweep(rbim_ctx* ctx, int32_t L, int8_t* spins, const int8_t* Jh, const int8_t* Jv,
               uint32_t thr31, uint64_t seed, uint32_t lattice_id, uint32_t sweep,
               int32_t measure, int32_t engine, rbim_sweep_stats* out);

/* rbim_sweep that allocates nothing.  It has no counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, checks and results as
 * rbim_sweep (bit-identical spins and stats for the same engine); every device
 * buffer of the sweep is carved out of a RBIM_NOALLOC_SCRATCH_BYTES array
 * that each GPU holds from library load until the process ends, which limits
 * L to RBIM_NOALLOC_MAX_L (larger L -> RBIM_E_INVAL).  Calls on one GPU share
 * that array: they are serialised by a lock inside this process (calls that
 * overlapped on it would race), and each call clears what it uses first.   */
#define RBIM_NOALLOC_MAX_L 256
#define RBIM_NOALLOC_SCRATCH_BYTES (2u << 20)
int rbim_sweep_noalloc(rbim_ctx* ctx, int32_t L, int8_t* spins, const int8_t* Jh,
                       const int8_t* Jv, uint32_t thr31, uint64_t seed, uint32_t lattice_id,
                       uint32_t sweep, int32_t measure, int32_t engine, rbim_sweep_stats* out);

/* cluster_statistics (src/rbim/clusters.py:98-118): sizes and signed cluster
 * magnetizations from labels 0..n_clusters-1. */
int rbim_cluster_statistics(rbim_ctx* ctx, int64_t n_sites, const int32_t* labels,
                            const int8_t* spins, int32_t n_clusters, int64_t* sizes,
                            int64_t* mags);

/* energy / magnetization (src/rbim/swendsen_wang.py:81-89). */
int rbim_energy_magnetization(rbim_ctx* ctx, int32_t L, const int8_t* spins, const int8_t* Jh,
                              const int8_t* Jv, int64_t* energy, int64_t* magnetization);

/* -------------------------------------------------------------- batch */
/* Device bytes a batch would hold; the serving loop admits by this figure. */
int rbim_batch_footprint(const rbim_batch_desc* desc, uint64_t* bytes);
int rbim_batch_create(rbim_ctx* ctx, const rbim_batch_desc* desc, rbim_batch** out);
/* Queue n_sweeps sweeps of every lattice on the batch's stream (asynchronous).
 * Fails with RBIM_E_STATE past n_therm + n_bins*sweeps_per_bin. */
int rbim_batch_run(rbim_batch* b, int64_t n_sweeps);
/* Like rbim_batch_run, then waits; *elapsed_ms is the time between HIP events
 * recorded on the batch's stream around the queued work. */
int rbim_batch_run_timed(rbim_batch* b, int64_t n_sweeps, double* elapsed_ms);
int rbim_batch_query(rbim_batch* b, int32_t* done);
int rbim_batch_sync(rbim_batch* b);
/* Moment sums, [n_lattices][n_bins][RBIM_ACC_SLOTS] in batch order. */
int rbim_batch_moments(rbim_batch* b, uint64_t* out);
/* Current spins, concatenated in batch order. */
int rbim_batch_spins(rbim_batch* b, int8_t* out);
int rbim_batch_info_get(rbim_batch* b, rbim_batch_info* out);
/* Per-sweep stats ring of a batch created with series > 0:
 * out[n_lattices][series] in batch order; entry k of a lattice holds measured
 * sweep j with j % series == k.  *recorded = measured sweeps so far (the
 * oldest kept one is max(0, recorded - series)).  The per-sweep loop of
 * scripts/measure_autocorrelation.py:69-90 reads these series. */
int rbim_batch_series(rbim_batch* b, rbim_sweep_stats* out, int64_t* recorded);
int rbim_batch_destroy(rbim_batch* b);

/* --------------------------------------------------------- slot pools */
/* A pool keeps n_slots lattice slots of one side L resident on the device,
 * each with its own schedule, and advances every loaded, unfinished slot in
 * shared launches: the serving loop's continuous batching, which replaces the
 * reference's process pool (src/rbim/simulation.py:395-399) for a stream of
 * scans.  Requests are loaded into free slots between runs; a slot stops at
 * its own n_therm + n_bins * sweeps_per_bin.  A lattice's moments depend only
 * on its seed, lattice id, couplings, start and schedule -- never on the slot,
 * the other slots or the run boundaries -- so they equal a batch's, bit for
 * bit.  A pool handle is used by one host thread at a time.                 */
typedef struct rbim_pool rbim_pool;
typedef struct {
  int32_t L;                  /* side of every slot's lattice                      */
  int32_t n_slots;            /* 1 .. 2^20                                         */
  int32_t max_bins;           /* bins of each slot's moment sums (>= every load's) */
  int32_t engine;             /* rbim_engine, as for batches                       */
  int32_t threads;            /* LDS kernel workgroup size; 0 -> default           */
  int32_t sweeps_per_launch;  /* LDS kernel sweeps per launch; 0 -> default        */
  int32_t tile;               /* tiled engines' tile side; 0 -> by L               */
} rbim_pool_desc;

/* Device bytes a pool would hold (admission by HBM footprint). */
int rbim_pool_footprint(const rbim_pool_desc* desc, uint64_t* bytes);
int rbim_pool_create(rbim_ctx* ctx, const rbim_pool_desc* desc, rbim_pool** out);
/* Loads n lattices into the free slots `slots` (queued after the pool's
 * earlier work): couplings Jh / Jv (n x L^2 concatenated, +-1), thresholds,
 * seeds, lattice ids, initial spins (n x L^2) or NULL for `start`, and one
 * schedule for all n.  RBIM_E_STATE if a slot is in use.                    */
int rbim_pool_load(rbim_pool* p, int32_t n, const int32_t* slots, const int8_t* Jh,
                   const int8_t* Jv, const uint32_t* thr31, const uint64_t* seed,
                   const uint32_t* lattice_id, const int8_t* spins0, int32_t start,
                   int64_t n_therm, int32_t n_bins, int64_t sweeps_per_bin);
/* Queues up to n_sweeps sweeps of every loaded, unfinished slot (asynchronous). */
int rbim_pool_run(rbim_pool* p, int64_t n_sweeps);
int rbim_pool_query(rbim_pool* p, int32_t* done);
int rbim_pool_sync(rbim_pool* p);
/* Sweeps queued so far and total sweeps of each slot (-1: slot not loaded);
 * host bookkeeping, no device access.                                       */
int rbim_pool_progress(rbim_pool* p, int32_t n, const int32_t* slots, int64_t* done,
                       int64_t* total);
/* Waits for the pool, then copies each slot's moment sums, [n][max_bins][20]
 * (the first n_bins bins of its load are meaningful), and its spins
 * ([n][L^2], or NULL).                                                      */
int rbim_pool_read(rbim_pool* p, int32_t n, const int32_t* slots, uint64_t* moments,
                   int8_t* spins);
/* Frees slots for later loads (each load clears its slots first). */
int rbim_pool_release(rbim_pool* p, int32_t n, const int32_t* slots);
int rbim_pool_destroy(rbim_pool* p);

/* Test hook with no counterpart in the reference: after `after` more device
 * allocations succeed, the next one fails as out of memory (RBIM_E_OOM);
 * after < 0 disarms it.  Lets tests show that callers survive a failed
 * allocation.                                                               */
int rbim_inject_alloc_failure(int64_t after);

#ifdef __cplusplus
}
#endif
#endif /* RBIM_HIP_H */
