/*
 * CPU restatement of the reference's Swendsen-Wang sweep for the 2D +-J RBIM.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine and
 * the timed host-core baseline; nothing in the shipped package links, loads or
 * calls it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may use the library built from it (oracle/_build/libsw_oracle.so).
 *
 * What it restates, function by function:
 *   oracle_philox4x32_10   Philox4x32-10 (Salmon et al., SC'11), the published
 *                          round function; pinned by Random123 known-answer
 *                          vectors and rocRAND's host engine (tests/golden).
 *   oracle_label_and_wrap  src/rbim/_unionfind.py:59-136 (_find with two-pass
 *                          compression, _label_and_wrap: hook b's root under a's
 *                          root, loop test on int64 unwrapped displacements,
 *                          relabel in min-site order).  Bond order and geometry
 *                          from src/rbim/lattice.py:105-126.
 *   oracle_activate        src/rbim/clusters.py:48-79, with the reference's PCG64
 *                          uniforms replaced by the engine's Philox stream
 *                          (counter = (site>>1, sweep, lattice, 0)).
 *   oracle_sweep_given_bits src/rbim/swendsen_wang.py:103-172 given occupation
 *                          and flip bits: label, flip (:140-141), statistics
 *                          (clusters.py:98-118), two largest (:92-100, tie rule
 *                          fixed here: equal sizes -> smaller min site first),
 *                          energy / magnetization (:81-89).
 *   oracle_run             src/rbim/simulation.py:196-236: initial spins, n_therm
 *                          unmeasured sweeps, measured sweeps accumulated per bin
 *                          into the engine's 20-slot moment layout.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SLOTS 20

typedef struct {
  int64_t energy, magnetization, n_clusters;
  int64_t largest_size, largest_abs_mag, second_size, second_abs_mag;
  int32_t wrap_x, wrap_y;
} oracle_stats;

/* ---------------------------------------------------------------- Philox */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------- geometry */
static inline int64_t right_of(int64_t n, int64_t L) {
  int64_t x = n / L, y = n % L;
  return ((x + 1) % L) * L + y;
}
static inline int64_t up_of(int64_t n, int64_t L) {
  int64_t x = n / L, y = n % L;
  return x * L + (y + 1) % L;
}

/* ----------------------------------------------------------- union-find */
/* _unionfind.py:59-88 */
static int64_t uf_find(int64_t* parent, int64_t* rx, int64_t* ry, int64_t start,
                       int64_t* out_x, int64_t* out_y) {
  int64_t root = start, ax = 0, ay = 0;
  while (parent[root] != root) {
    ax += rx[root];
    ay += ry[root];
    root = parent[root];
  }
  int64_t node = start, bx = 0, by = 0;
  while (parent[node] != node) {
    int64_t next = parent[node];
    int64_t ox = rx[node], oy = ry[node];
    parent[node] = root;
    rx[node] = ax - bx;
    ry[node] = ay - by;
    bx += ox;
    by += oy;
    node = next;
  }
  *out_x = ax;
  *out_y = ay;
  return root;
}

/* _unionfind.py:91-136.  active[k] in the reference's bond order: bond 2n is
 * (n -> right(n), dx=1), bond 2n+1 is (n -> up(n), dy=1).  labels[] receives
 * labels numbered in increasing min-site order; returns n_clusters. */
typedef struct {
  int64_t *parent, *rx, *ry, *remap, *right, *up;
} uf_ws;

static void ws_alloc(uf_ws* w, int64_t L) {
  const int64_t N = L * L;
  w->parent = (int64_t*)malloc(sizeof(int64_t) * 6 * N);
  w->rx = w->parent + N;
  w->ry = w->parent + 2 * N;
  w->remap = w->parent + 3 * N;
  w->right = w->parent + 4 * N;
  w->up = w->parent + 5 * N;
  for (int64_t n = 0; n < N; ++n) { w->right[n] = right_of(n, L); w->up[n] = up_of(n, L); }
}
static void ws_free(uf_ws* w) { free(w->parent); }

static int64_t label_and_wrap_ws(int64_t L, const uint8_t* active, int64_t* labels,
                                 int32_t* wrap_x, int32_t* wrap_y, uf_ws* w) {
  const int64_t N = L * L;
  int64_t *parent = w->parent, *rx = w->rx, *ry = w->ry, *remap = w->remap;
  for (int64_t i = 0; i < N; ++i) { parent[i] = i; rx[i] = 0; ry[i] = 0; remap[i] = -1; }
  int wx = 0, wy = 0;
  for (int64_t k = 0; k < 2 * N; ++k) {
    if (!active[k]) continue;
    int64_t n = k >> 1;
    int64_t a = n, b = (k & 1) ? w->up[n] : w->right[n];
    int64_t ex = (k & 1) ? 0 : 1, ey = (k & 1) ? 1 : 0;
    int64_t ax, ay, bx, by;
    int64_t ra = uf_find(parent, rx, ry, a, &ax, &ay);
    int64_t rb = uf_find(parent, rx, ry, b, &bx, &by);
    if (ra != rb) {
      parent[rb] = ra;
      rx[rb] = ax - ex - bx;
      ry[rb] = ay - ey - by;
    } else {
      if (ex + bx - ax != 0) wx = 1;
      if (ey + by - ay != 0) wy = 1;
    }
  }
  int64_t n_clusters = 0;
  for (int64_t s = 0; s < N; ++s) {
    int64_t dx, dy;
    int64_t root = uf_find(parent, rx, ry, s, &dx, &dy);
    if (remap[root] == -1) remap[root] = n_clusters++;
    labels[s] = remap[root];
  }
  *wrap_x = wx;
  *wrap_y = wy;
  return n_clusters;
}

int64_t oracle_label_and_wrap(int64_t L, const uint8_t* active, int64_t* labels,
                              int32_t* wrap_x, int32_t* wrap_y) {
  uf_ws w;
  ws_alloc(&w, L);
  int64_t n = label_and_wrap_ws(L, active, labels, wrap_x, wrap_y, &w);
  ws_free(&w);
  return n;
}

/* ----------------------------------------------------------- activation */
/* clusters.py:48-79 with Philox draws: active = satisfied & (draw < p_add),
 * draw = word >> 1 against thr31 = round(p_add * 2^31).  flipbits[n] = low bit
 * of the horizontal word of site n. */
static void activate_ws(int64_t L, const int8_t* spins, const int8_t* Jh, const int8_t* Jv,
                        uint32_t thr31, uint64_t seed, uint32_t lattice_id, uint32_t sweep,
                        uint8_t* active, uint8_t* flipbits, const uf_ws* w) {
  const int64_t N = L * L;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4] = {0, 0, 0, 0};
  for (int64_t n = 0; n < N; ++n) {
    if ((n & 1) == 0) { /* one Philox call serves the pair (n, n+1) */
      uint32_t ctr[4] = {(uint32_t)(n >> 1), sweep, lattice_id, 0u};
      oracle_philox4x32_10(ctr, key, r);
    }
    uint32_t dh = (n & 1) ? r[2] : r[0];
    uint32_t dv = (n & 1) ? r[3] : r[1];
    int sh = Jh[n] * spins[n] * spins[w->right[n]] > 0;
    int sv = Jv[n] * spins[n] * spins[w->up[n]] > 0;
    active[2 * n] = (uint8_t)(sh && (dh >> 1) < thr31);
    active[2 * n + 1] = (uint8_t)(sv && (dv >> 1) < thr31);
    if (flipbits) flipbits[n] = (uint8_t)(dh & 1u);
  }
}

void oracle_activate(int64_t L, const int8_t* spins, const int8_t* Jh, const int8_t* Jv,
                     uint32_t thr31, uint64_t seed, uint32_t lattice_id, uint32_t sweep,
                     uint8_t* active, uint8_t* flipbits) {
  uf_ws w;
  ws_alloc(&w, L);
  activate_ws(L, spins, Jh, Jv, thr31, seed, lattice_id, sweep, active, flipbits, &w);
  ws_free(&w);
}

/* ------------------------------------------------- sweep given the bits */
/* swendsen_wang.py:103-172.  flip_by_minsite[m] is the flip decision of the
 * cluster whose minimum site index is m.  spins are updated in place. */
static void sweep_given_bits_impl(int64_t L, int8_t* spins, const int8_t* Jh, const int8_t* Jv,
                                  const uint8_t* active, const uint8_t* flip_by_minsite,
                                  int measure, oracle_stats* st, int64_t* labels_out,
                                  int64_t* labels, int64_t* minsite, int64_t* size,
                                  int64_t* mag, uf_ws* w) {
  const int64_t N = L * L;
  int32_t wx, wy;
  int64_t ncl = label_and_wrap_ws(L, active, labels, &wx, &wy, w);
  /* labels are in min-site order: the first site carrying label k is its min site */
  for (int64_t k = 0; k < ncl; ++k) minsite[k] = -1;
  for (int64_t s = 0; s < N; ++s)
    if (minsite[labels[s]] < 0) minsite[labels[s]] = s;
  for (int64_t s = 0; s < N; ++s)
    if (flip_by_minsite[minsite[labels[s]]]) spins[s] = (int8_t)-spins[s];
  if (labels_out) memcpy(labels_out, labels, sizeof(int64_t) * N);
  if (!measure) return;
  for (int64_t k = 0; k < ncl; ++k) { size[k] = 0; mag[k] = 0; }
  for (int64_t s = 0; s < N; ++s) { size[labels[s]] += 1; mag[labels[s]] += spins[s]; }
  /* two largest in the order (size desc, min site asc); labels are already in
   * min-site order, so a strict '>' scan keeps the smaller label on ties. */
  int64_t first = 0, second = -1;
  for (int64_t k = 1; k < ncl; ++k) {
    if (size[k] > size[first]) { second = first; first = k; }
    else if (second < 0 || size[k] > size[second]) second = k;
  }
  int64_t E = 0, M = 0;
  for (int64_t n = 0; n < N; ++n) {
    E -= (int64_t)Jh[n] * spins[n] * spins[w->right[n]];
    E -= (int64_t)Jv[n] * spins[n] * spins[w->up[n]];
    M += spins[n];
  }
  st->energy = E;
  st->magnetization = M;
  st->n_clusters = ncl;
  st->largest_size = size[first];
  st->largest_abs_mag = mag[first] < 0 ? -mag[first] : mag[first];
  st->second_size = second < 0 ? 0 : size[second];
  st->second_abs_mag = second < 0 ? 0 : (mag[second] < 0 ? -mag[second] : mag[second]);
  st->wrap_x = wx;
  st->wrap_y = wy;
}

int oracle_sweep_given_bits(int64_t L, int8_t* spins, const int8_t* Jh, const int8_t* Jv,
                            const uint8_t* active, const uint8_t* flip_by_minsite,
                            oracle_stats* st, int64_t* labels_out) {
  const int64_t N = L * L;
  int64_t* buf = (int64_t*)malloc(sizeof(int64_t) * 4 * N);
  uf_ws w;
  ws_alloc(&w, L);
  sweep_given_bits_impl(L, spins, Jh, Jv, active, flip_by_minsite, 1, st, labels_out, buf,
                        buf + N, buf + 2 * N, buf + 3 * N, &w);
  ws_free(&w);
  free(buf);
  return 0;
}

/* ------------------------------------------------------- initial spins */
/* swendsen_wang.py:66-78 with the Philox init stream (tag 1). */
void oracle_init_spins(int64_t L, uint64_t seed, uint32_t lattice_id, int start, int8_t* spins) {
  const int64_t N = L * L;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t n = 0; n < N; ++n) {
    if (start == 1) { spins[n] = 1; continue; }
    uint32_t ctr[4] = {(uint32_t)(n >> 7), 0u, lattice_id, 1u}, r[4];
    oracle_philox4x32_10(ctr, key, r);
    uint32_t w = r[(n >> 5) & 3];
    spins[n] = ((w >> (n & 31)) & 1u) ? 1 : -1;
  }
}

/* --------------------------------------------------- moment accumulation */
static void add_u128(uint64_t* lo, uint64_t* hi, unsigned __int128 v) {
  unsigned __int128 cur = ((unsigned __int128)(*hi) << 64) | *lo;
  cur += v;
  *lo = (uint64_t)cur;
  *hi = (uint64_t)(cur >> 64);
}

/* simulation.py:212-236: the 14 primitives of one measured sweep, summed into
 * the 20-slot layout (slot meanings are listed in include/rbim_hip.h). */
void oracle_accumulate(uint64_t* acc, const oracle_stats* st) {
  int64_t E = st->energy, M = st->magnetization;
  uint64_t aM = (uint64_t)(M < 0 ? -M : M);
  acc[0] = (uint64_t)((int64_t)acc[0] + E);
  add_u128(&acc[1], &acc[2], (unsigned __int128)(uint64_t)(E < 0 ? -E : E) * (uint64_t)(E < 0 ? -E : E));
  acc[3] = (uint64_t)((int64_t)acc[3] + M);
  acc[4] += aM;
  add_u128(&acc[5], &acc[6], (unsigned __int128)aM * aM);
  unsigned __int128 m2 = (unsigned __int128)aM * aM;
  add_u128(&acc[7], &acc[8], m2 * m2);
  acc[9] += (uint64_t)st->largest_size;
  acc[10] += (uint64_t)st->largest_abs_mag;
  acc[11] += (uint64_t)st->second_size;
  double ratio = st->largest_size > 0 ? (double)st->largest_abs_mag / (double)st->largest_size : 0.0;
  double cur;
  memcpy(&cur, &acc[12], sizeof(double));
  cur += ratio;
  memcpy(&acc[12], &cur, sizeof(double));
  acc[13] += (uint64_t)st->n_clusters;
  acc[14] += (uint64_t)(st->wrap_x != 0);
  acc[15] += (uint64_t)(st->wrap_y != 0);
  acc[16] += (uint64_t)(st->wrap_x != 0 || st->wrap_y != 0);
  acc[17] += 1;
}

/* ------------------------------------------------------------ full run */
/* simulation.py:196-236 for one lattice: sweeps g0 .. g0+n_sweeps-1 of a run
 * with n_therm unmeasured sweeps followed by n_bins*spb measured ones.
 * acc is [n_bins][SLOTS]; spins are updated in place. */
int oracle_run(int64_t L, int8_t* spins, const int8_t* Jh, const int8_t* Jv, uint32_t thr31,
               uint64_t seed, uint32_t lattice_id, int64_t g0, int64_t n_sweeps, int64_t n_therm,
               int64_t spb, int64_t n_bins, uint64_t* acc) {
  const int64_t N = L * L;
  uint8_t* active = (uint8_t*)malloc(2 * N);
  uint8_t* flips = (uint8_t*)malloc(N);
  int64_t* buf = (int64_t*)malloc(sizeof(int64_t) * 4 * N);
  uf_ws w;
  ws_alloc(&w, L);
  const int64_t total = n_therm + n_bins * spb;
  for (int64_t g = g0; g < g0 + n_sweeps && g < total; ++g) {
    activate_ws(L, spins, Jh, Jv, thr31, seed, lattice_id, (uint32_t)g, active, flips, &w);
    int measure = g >= n_therm;
    oracle_stats st;
    sweep_given_bits_impl(L, spins, Jh, Jv, active, flips, measure, &st, NULL, buf, buf + N,
                          buf + 2 * N, buf + 3 * N, &w);
    if (measure) oracle_accumulate(acc + ((g - n_therm) / spb) * SLOTS, &st);
  }
  ws_free(&w);
  free(active); free(flips); free(buf);
  return 0;
}
