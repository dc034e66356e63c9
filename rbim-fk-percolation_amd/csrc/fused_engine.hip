// Edge-slot Swendsen-Wang sweep for lattices of whole tiles (L a multiple of
// the tile side T = 64, else 32), on gfx950.  Two kernels while a lattice's
// tile-edge slots fit one workgroup's LDS (RBIM_FUSED_MAX_SLOTS: L <= 320 at
// T = 64, L <= 224 at T = 32); beyond that the slot merge runs in HBM
// (kg_* kernels, five kernels per sweep).
//
// The tiled path of tiled_engine.hip keeps a forest word, counters and a
// local-root index for every site in HBM (about 49 B of traffic per
// site-update).  Here only the spins (1 B) and a 16-bit label per site cross
// HBM; everything that couples tiles is reduced to the tile-edge sites, which
// one workgroup per lattice merges in LDS:
//
//   KA tile    one workgroup per T x T tile.  Applies the previous sweep's
//              pending flips (below) to its sites and to the row / column of
//              sites just beyond its upper edges, measures E and M of that
//              state, draws the sweep's bonds (Philox4x32-10, as everywhere),
//              labels the tile's interior bonds in LDS (ballot runs + plaquette
//              skip + hook-and-compress, no windings: a tile never crosses the
//              periodic boundary) and classifies every tile-local cluster:
//                interior  no site on the tile edge: the cluster is final.  Its
//                          flip (flip bit of its minimum site) is applied at
//                          once and it is measured here (size, n_plus after the
//                          flip, two largest keys).
//                edge      touches the tile edge: its sites keep their spin and
//                          get label 0x8000 | rep, rep = the cluster's smallest
//                          edge slot (4T slots per tile: x = x0, x = x0+T-1,
//                          y = y0, y = y0+T-1).  Every slot records its rep and,
//                          at reps, the cluster's size, n_plus before the flip,
//                          minimum site and that site's flip bit.
//              The occupation of the 2T bonds leaving the tile (+x, +y) goes out
//              as a bitmask.
//   KB lattice one workgroup per lattice, one 64-bit LDS word per slot.
//              Union-find over the lattice's edge
//              slots in LDS: slots start under their rep (winding 0: same
//              tile), every occupied tile-crossing bond unites its two slots
//              with its periodic winding; a bond that finds one root closes a
//              loop and c + w(b) - w(a) decides wrap_x / wrap_y exactly
//              (_unionfind.py:101-124).  Roots hook by minimum site, so a
//              merged cluster's root is the rep holding its minimum site and
//              the cluster flips by that site's flip bit, the rule of every
//              engine here; it sums its reps' sizes and n_plus.  Flip bits go
//              out per slot; the merged
//              clusters are measured; the previous sweep is folded into its
//              bin (E, M from this sweep's KA; interior clusters from the
//              previous KA; edge clusters and wraps from the previous KB).
//
// Spins and labels are double-buffered (KA reads one buffer and writes the
// other), so a tile reading its neighbours' edge sites never races their
// writes.  After a run of sweeps three small kernels measure the last state,
// apply the pending flips into buffer 0 and fold the last sweep, so the spins
// a caller reads are final and a later run starts from a clean state.
#include "engine.h"
#include "tile_common.h"

namespace {

// State word of a site between tile passes: its spin (1: +1), whether it
// belongs to an edge cluster, whose flip is pending, and that cluster's rep
// slot in the site's tile.
constexpr uint32_t ST_SPIN = 0x8000u;
constexpr uint32_t ST_EDGE = 0x4000u;
constexpr uint32_t ST_SLOT = 0x3FFFu;
constexpr uint32_t NONE16 = 0xFFFFu;

// Slot entry: rep (0..11) | min local index (12..23) | flip (24) |
// size (25..37) | n_plus before the flip (38..50).
__device__ __forceinline__ uint64_t slot_entry(uint32_t rep, uint32_t root, uint32_t flip,
                                               uint32_t size, uint32_t np) {
  return (uint64_t)rep | ((uint64_t)root << 12) | ((uint64_t)flip << 24) |
         ((uint64_t)size << 25) | ((uint64_t)np << 38);
}

// Local index of edge slot k of a T x T tile: side k / T (0: lx = 0,
// 1: lx = T-1, 2: ly = 0, 3: ly = T-1), position k % T along it.
template <int T>
__device__ __forceinline__ uint32_t slot_li(uint32_t k) {
  const uint32_t side = k / T, i = k % T;
  const uint32_t lx = side == 0 ? 0u : side == 1 ? (uint32_t)T - 1u : i;
  const uint32_t ly = side == 2 ? 0u : side == 3 ? (uint32_t)T - 1u : i;
  return lx * T + ly;
}

// Edge slot k of a tile whose valid part is wx x wy sites (wx = wy = T for a
// whole tile): its local index in rows of 2^rsh sites, and whether the
// position holds a site (the sides of a cut tile are shorter than T).
template <int T>
__device__ __forceinline__ uint32_t slot_li_cut(uint32_t k, uint32_t wx, uint32_t wy, uint32_t rsh,
                                                bool& ok) {
  const uint32_t side = k / T, i = k % T;
  ok = side < 2u ? i < wy : i < wx;
  const uint32_t lx = side == 0 ? 0u : side == 1 ? wx - 1u : i;
  const uint32_t ly = side == 2 ? 0u : side == 3 ? wy - 1u : i;
  return (lx << rsh) + ly;
}

// Slot entry of a position of a cut tile that holds no site.
constexpr uint64_t SLOT_ABSENT = 1ull << 63;

// Site pairs per row of KA's lane map, as a shift.  A whole row holds T / 2
// pairs; a tile whose rows are cut to wy < T sites packs each row into the
// smallest power of two of pairs that holds them, so that lanes, and whole
// wavefronts, follow the sites that are present.  A row never spans two
// wavefronts (T / 2 <= 64), and a wavefront never spans more than T rows (at
// least 64 / T pairs a row), so the rows of a wavefront that starts inside
// the tile end inside it.
template <int T>
__device__ __forceinline__ uint32_t pair_shift(uint32_t wy) {
  constexpr uint32_t lo = T >= 64 ? 0u : (uint32_t)__builtin_ctz(64 / T);
  const uint32_t pairs = (wy + 1u) >> 1;
  const uint32_t sh = pairs <= 1u ? 0u : 32u - (uint32_t)__builtin_clz(pairs - 1u);
  return sh > lo ? sh : lo;
}

__device__ __forceinline__ uint32_t flip_of(const uint32_t* fb, uint32_t slot) {
  return (fb[slot >> 5] >> (slot & 31u)) & 1u;
}

// State planes of lattice lat: the current one and the one a tile pass writes.
struct FusedView {
  const uint16_t* in;
  uint16_t* out;
};

__device__ __forceinline__ FusedView fused_view(const LatticeSet& s, const FusedScratch& f, int lat) {
  const int cur = f.cur[lat];
  FusedView v;
  v.in = f.st + ((size_t)cur * s.n_lat + lat) * s.N;
  v.out = f.st + ((size_t)(1 - cur) * s.n_lat + lat) * s.N;
  return v;
}

// Spin bit (1: +1) of a site of tile `tile` from its state word w, pending
// edge-cluster flip applied.
template <int T>
__device__ __forceinline__ uint32_t true_spin(uint32_t w, uint32_t tile, const uint32_t* fb) {
  uint32_t b = (w >> 15) & 1u;
  if (w & ST_EDGE) b ^= flip_of(fb, tile * 4u * T + (w & ST_SLOT));
  return b;
}

// The lattice's flip bits of the previous sweep: staged into LDS (lds) when
// the lattice's slots merge in LDS, read in place from HBM (L2) otherwise.
template <bool BIG>
__device__ __forceinline__ const uint32_t* flipbits_of(const FusedScratch& f, int lat, uint32_t* lds,
                                                       int nth) {
  const uint32_t words = (uint32_t)f.nslots / 32u;
  const uint32_t* src = f.flipbits + (size_t)lat * words;
  if (BIG) return src;
  for (uint32_t i = threadIdx.x; i < words; i += nth) lds[i] = src[i];
  return lds;
}

// LDS words for staged flip bits (none when they stay in HBM).
template <bool BIG>
constexpr int fb_words() { return BIG ? 1 : RBIM_FUSED_MAX_SLOTS / 32; }

// Spin bits of the column beyond x0+wx-1 (hsp[0..wy)) and the row beyond
// y0+wy-1 (hsp[T..T+wx)), pending flips applied (sites of the neighbour tiles);
// wx x wy is the tile's valid part.
template <int T>
__device__ __forceinline__ void load_halo(const LatticeSet& s, const FusedScratch& f,
                                          const FusedView& v, const uint32_t* fb, uint32_t tx,
                                          uint32_t ty, uint32_t wx, uint32_t wy, uint8_t* hsp) {
  const uint32_t tid = threadIdx.x;
  if (tid >= 2u * T) return;
  const uint32_t L = (uint32_t)s.L, nt = (uint32_t)f.tiles_x;
  const uint32_t x0 = tx * T, y0 = ty * T;
  const bool right = tid < (uint32_t)T;
  const uint32_t i = right ? tid : tid - T;
  if (i >= (right ? wy : wx)) return;
  const uint32_t hx = right ? (x0 + wx == L ? 0u : x0 + wx) : x0 + i;
  const uint32_t hy = right ? y0 + i : (y0 + wy == L ? 0u : y0 + wy);
  const uint32_t ntile = right ? ty * nt + (tx + 1u == nt ? 0u : tx + 1u)
                               : (ty + 1u == nt ? 0u : ty + 1u) * nt + tx;
  const uint32_t hn = hx * L + hy;
  hsp[tid] = (uint8_t)true_spin<T>(v.in[hn], ntile, fb);
}

// Roots of x and y in the settled tile forest (nodes named by byte-offset
// handles, lds_at), walked together, reading parents through a 16-bit mask (a
// root's word may already carry its rep slot above bit 16).
[[maybe_unused]] __device__ __forceinline__ void troot16_2(const uint32_t* par, uint32_t& x,
                                                          uint32_t& y) {
  uint32_t px = lds_get<true>(par, x) & 0xFFFFu, py = lds_get<true>(par, y) & 0xFFFFu;
  while (px != x || py != y) {
    x = px;
    y = py;
    px = lds_get<true>(par, x) & 0xFFFFu;
    py = lds_get<true>(par, y) & 0xFFFFu;
  }
}

// Roots of four sites at once (two site pairs of a lane), as troot16_2.
__device__ __forceinline__ void troot16_4(const uint32_t* par, uint32_t (&x)[4]) {
  uint32_t p[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) p[c] = lds_get<true>(par, x[c]) & 0xFFFFu;
  while (p[0] != x[0] || p[1] != x[1] || p[2] != x[2] || p[3] != x[3]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = p[c];
#pragma unroll
    for (int c = 0; c < 4; ++c) p[c] = lds_get<true>(par, x[c]) & 0xFFFFu;
  }
}

// The HBM slot forest's root words: the slot itself (low 32 bits, as in any
// word: gpar) and its cluster's minimum site in the upper half, where a
// non-root word keeps its windings; roots hook by that site (kg_union).
__device__ __forceinline__ uint64_t groot_word(uint32_t k, uint32_t site) {
  return (uint64_t)k | ((uint64_t)site << 32);
}
__device__ __forceinline__ uint32_t groot_site(uint64_t v) { return (uint32_t)(v >> 32); }

// ------------------------------------------------------------------- KA
// LDS of one tile pass over a whole tile.
template <int T, int NTH>
struct KaShared {
  uint32_t par[T * T];
  __align__(16) uint32_t lcnt[T * T];
  uint8_t bits[T * T];  // bit0 spin, bit1 occ_h, bit2 occ_v, bit3 flip draw, bit4/5 J<0
  uint8_t hsp[2 * T];
  int64_t red[NTH / 64][PARTS];
};

// The arrays of a tile pass: the whole-tile kernel's static ones, or a cut-tile
// kernel's dynamic ones sized for the cut (cap sites).  lcnt holds the waves'
// union queues (qw words each, a multiple of 4 and at least 192) while the
// tile is labelled, then the cluster counters.
template <int NTH>
struct KaLds {
  uint32_t* par;
  uint32_t* lcnt;  // 16-byte aligned
  uint8_t* bits;
  uint8_t* hsp;
  int64_t (*red)[PARTS];
  uint32_t qw;
};

// Sites of LDS a cut-tile pass needs: the cut rows of the last tile column
// (cut x T) and the packed rows of the last tile row (T x 2^(pair shift + 1)),
// and at least 192 queue words per wave; a multiple of 4 per wave.
template <int T>
__host__ __device__ constexpr uint32_t cut_lds_sites(uint32_t cut, uint32_t waves) {
  uint32_t pairs = (cut + 1u) / 2u, sh = 0u;
  while ((1u << sh) < pairs) ++sh;
  if (T < 64 && sh < 1u) sh = 1u;  // at least 64 / T pairs a row (T = 32)
  uint32_t cap = (uint32_t)T << (sh + 1u);
  if (cap < 192u * waves) cap = 192u * waves;
  return (cap + 4u * waves - 1u) / (4u * waves) * (4u * waves);
}

// PART: 0 -- a whole tile (wx = wy = T) of a lattice of even L; 1 -- L even,
// the last tile row and column cut (wx, wy < T; site pairs stay whole and
// aligned); 2 -- L odd, cut as well, so the last pair of a cut row holds one
// site and a pair of a row starting at an odd site straddles two Philox pairs
// (two draws); 3 -- a whole tile of a lattice of odd L: its odd rows start at
// odd sites, and the lane map gives every wavefront rows of one parity, so
// that only the odd-row wavefronts load site by site and draw twice.
template <int T, int NTH, bool BIG, int PART>
__device__ __forceinline__ void ka_body(const LatticeSet& s, const FusedScratch& f,
                                        const KaLds<NTH>& m, uint32_t tile) {
  constexpr int ITERS = T * T / 2 / NTH;
  constexpr uint32_t NS = 4u * T;  // edge slots per tile
  constexpr bool CUT = PART == 1 || PART == 2;  // a tile that may be cut: sites may be absent
  const uint32_t QW = m.qw;        // union queue entries per wave (a wave queues up to 192)
  uint32_t* par = m.par;
  uint32_t* lcnt = m.lcnt;
  uint8_t* bits = m.bits;
  uint8_t* hsp = m.hsp;
  int64_t (*red)[PARTS] = m.red;
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const uint32_t L = (uint32_t)s.L;
  const uint32_t nt = (uint32_t)f.tiles_x, tiles = nt * nt;
  const uint32_t tx = tile % nt, ty = tile / nt;
  const uint32_t x0 = tx * T, y0 = ty * T;
  // the tile's valid part (the last tile row / column is cut when T does not
  // divide L); sites beyond it are absent: no loads, stores, bonds or counts
  const uint32_t wx = CUT ? min((uint32_t)T, L - x0) : (uint32_t)T;
  const uint32_t wy = CUT ? min((uint32_t)T, L - y0) : (uint32_t)T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  // lane map: queue position ql = it * NTH + tid holds the site pair (lx, ly0),
  // (lx, ly0 + 1), local index li0 = lx * RS + ly0 in rows of RS = 2^rsh sites
  // (T in a whole tile, where li0 = 2 ql; a tile with cut rows packs them to
  // twice its pairs per row, so a wavefront's LDS accesses spread over the
  // banks as in a whole tile).  In a cut tile the wavefronts whose pairs all
  // lie past row wx - 1 skip every phase that works pair by pair (their sites
  // are absent, and no present site reads the local state of an absent row).
  // Slot entries and site numbers use the tile's own rows (stride T).
  const uint32_t psh = CUT ? pair_shift<T>(wy) : (uint32_t)__builtin_ctz(T / 2);
  const uint32_t rsh = psh + 1u, RS = 1u << rsh;
  const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane(wave);
  // PART 3: rows permuted within blocks of twice a wavefront's rows (RPW = 128
  // / T), a wavefront taking the block's even or its odd rows: the wavefront
  // index's parity is its rows' parity (x0 is even, and NTH / 64 is even)
  constexpr uint32_t RPW = 128u / T;
  auto row_of = [&](uint32_t ql) -> uint32_t {
    const uint32_t r = ql >> psh;
    if constexpr (PART == 3)
      return (r & ~(2u * RPW - 1u)) | ((r & (RPW - 1u)) << 1) | ((r / RPW) & 1u);
    else
      return r;
  };
  auto col_of = [&](uint32_t ql) -> uint32_t { return (ql & ((1u << psh) - 1u)) << 1; };
  auto li_of = [&](uint32_t ql) -> uint32_t {
    return PART ? (row_of(ql) << rsh) + col_of(ql) : ql << 1;
  };
  const bool odd_rows = PART == 3 && (wv & 1u);  // wavefront-uniform
  // The tile forest names its nodes by handles, byte offsets 4 li (lds_at): a
  // parent read is the address of the next step.  H: handle of local index li,
  // IDX: local index of handle h; lcnt is read and counted at handles as well.
  auto H = [](uint32_t li) -> uint32_t { return li << 2; };
  auto IDX = [](uint32_t h) -> uint32_t { return h >> 2; };
  // a local index as the tile-row index lx * T + ly
  auto tile_li = [&](uint32_t li) -> uint32_t {
    return CUT ? (li >> rsh) * T + (li & (RS - 1u)) : li;
  };
  auto skip = [&](int it) -> bool {
    return CUT && (((uint32_t)it * NTH + wv * 64u) >> psh) >= wx;
  };
  const int64_t g = s.sweep_ctr[lat];
  const uint32_t thr = s.thr31[lat];
  // a bond is drawn occupied when draw >> 1 < thr, i.e. draw <= 2 thr - 1
  const bool any_bond = thr != 0u;
  const uint32_t tlim = any_bond ? 2u * thr - 1u : 0u;
  const uint64_t seed = s.seed[lat];
  const uint32_t lid = s.lat_id[lat];
  const FusedView v = fused_view(s, f, lat);
  const uint32_t* jt = f.jt + ((size_t)lat * tiles + tile) * (T * T / 16);
  // E, M of the state this sweep starts from close sweep g-1 (unless a run's
  // closing pass has already folded it); the clusters of sweep g are measured
  // when g is past thermalisation
  const int64_t n_therm = lat_therm(s, lat);
  const bool meas_em = g >= 1 && f.fin[lat] == g - 1 && g - 1 >= n_therm;
  const bool meas_cl = g >= n_therm;

  // The state this sweep starts from: spins, labels and couplings of the
  // tile, the halo sites (the column beyond x0+T-1 and the row beyond
  // y0+T-1), and the flip words of the previous sweep's edge clusters that
  // they can need -- WPT words each of this tile, the right and the upper
  // neighbour tile, held by lanes 0 .. 3 WPT - 1 of every wavefront and read
  // by shuffles.  All loads are issued together, with no barrier before them.
  constexpr uint32_t WPT = 4u * T / 32u;
  const uint32_t nt_r = ty * nt + (tx + 1u == nt ? 0u : tx + 1u);
  const uint32_t nt_u = (ty + 1u == nt ? 0u : ty + 1u) * nt + tx;
  uint32_t fw = 0u;
  if ((uint32_t)lane < 3u * WPT) {
    const uint32_t w = (uint32_t)lane % WPT;
    const uint32_t src = lane < (int)WPT ? tile : lane < (int)(2u * WPT) ? nt_r : nt_u;
    fw = f.flipbits[(size_t)lat * ((uint32_t)f.nslots / 32u) + src * WPT + w];
  }
  // the pair's two state words, and its word of couplings (the pair's four
  // bits sit at bit 2 (lx T + ly0) mod 32: a pair never straddles two words)
  uint32_t st2[ITERS], j2[ITERS];
#pragma unroll
  for (int it = 0; it < ITERS; ++it) {
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t lx = row_of(ql), ly0 = col_of(ql);
    const uint32_t n0 = (x0 + lx) * L + y0 + ly0;  // even L: the pair loads are aligned
    const uint32_t lt = lx * T + ly0;
    if (PART == 0 || (PART == 3 && !odd_rows)) {
      st2[it] = *reinterpret_cast<const uint32_t*>(v.in + n0);
      j2[it] = jt[lt >> 4];
    } else if (PART == 3) {  // an odd row: the pair starts at an odd site
      st2[it] = (uint32_t)v.in[n0] | (uint32_t)v.in[n0 + 1u] << 16;
      j2[it] = jt[lt >> 4];
    } else {
      // absent sites read as spin +1, no label, J = +1
      const bool v0 = lx < wx && ly0 < wy, v1 = lx < wx && ly0 + 1u < wy;
      st2[it] = ST_SPIN | ST_SPIN << 16;
      j2[it] = 0u;
      if (skip(it)) continue;
      if (v0) j2[it] = jt[lt >> 4];
      if (PART == 1) {
        if (v0) st2[it] = *reinterpret_cast<const uint32_t*>(v.in + n0);
      } else {
        if (v0) st2[it] = (uint32_t)v.in[n0] | ST_SPIN << 16;
        if (v1) st2[it] = (st2[it] & 0xFFFFu) | (uint32_t)v.in[n0 + 1u] << 16;
      }
    }
  }
  const bool halo = tid < 2 * T;  // tid < T: right column, else upper row
  uint32_t hw = ST_SPIN;
  if (halo) {
    const bool right = tid < T;
    const uint32_t i = right ? (uint32_t)tid : (uint32_t)tid - T;
    const uint32_t hx = right ? (x0 + wx == L ? 0u : x0 + wx) : x0 + i;
    const uint32_t hy = right ? y0 + i : (y0 + wy == L ? 0u : y0 + wy);
    if (!CUT || i < (right ? wy : wx)) hw = v.in[hx * L + hy];
  }
  // spin bit of a site from its state word w, its edge cluster's pending flip
  // applied (wbase: the first flip word of the site's tile); the shuffle runs
  // on every lane
  auto spin_bit = [&](uint32_t w, uint32_t wbase) -> uint32_t {
    const uint32_t slot = w & ST_SLOT;
    const uint32_t word = (uint32_t)__shfl((int)fw, (int)(wbase + (slot >> 5)));
    const uint32_t flip = (w & ST_EDGE) ? (word >> (slot & 31u)) & 1u : 0u;
    return ((w >> 15) & 1u) ^ flip;
  };
#pragma unroll
  for (int it = 0; it < ITERS; ++it) {
    if (skip(it)) continue;
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t li0 = li_of(ql);
    const uint32_t j4 = j2[it] >> (((row_of(ql) * T + col_of(ql)) & 15u) * 2u);
    const uint32_t t0 = spin_bit(st2[it] & 0xFFFFu, 0u);
    const uint32_t t1 = spin_bit(st2[it] >> 16, 0u);
    bits[li0] = (uint8_t)(t0 | (j4 & 3u) << 4);
    bits[li0 + 1u] = (uint8_t)(t1 | ((j4 >> 2) & 3u) << 4);
  }
  {
    const uint32_t hb = spin_bit(hw, tid < T ? WPT : 2u * WPT);
    if (halo) hsp[tid] = (uint8_t)hb;
  }
  __syncthreads();

  // activation (+ E, M of the starting state) and run pre-labels.  In a whole
  // tile the run starts of each pair's two sites stay in registers (RSK): the
  // unions and root walks below start there, one LDS round trip shorter (a
  // site and its run start have one root, and halving from the run start
  // compresses the path its whole run shares); a cut tile's pass has no
  // registers to spare for them
  constexpr bool RSK = !CUT && ITERS % 2 == 0;  // (the four-root walk below keeps the roots)
  constexpr int UB = RSK ? ITERS : 1;  // unrolled where the registers are indexed
  int32_t me = 0, mm = 0;
  uint32_t rsa[ITERS], rsb[ITERS];
#pragma unroll
  for (int it = 0; it < ITERS; ++it) {
    if (skip(it)) continue;
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t lx = row_of(ql), ly0 = col_of(ql);
    const uint32_t li0 = li_of(ql), li1 = li0 + 1u;
    const uint32_t n0 = (x0 + lx) * L + y0 + ly0;
    const uint32_t b0 = bits[li0], b1 = bits[li1];
    // the x neighbour is in the tile below row wx-1, the y neighbour of n1 below
    // column wy-1 (of n0: n1 itself, unless n0 is a cut row's last site)
    const bool v0 = !CUT || (lx < wx && ly0 < wy), v1 = !CUT || (lx < wx && ly0 + 1u < wy);
    const uint32_t sr0 = lx + 1u < wx ? bits[li0 + RS] & 1u : hsp[ly0];
    const uint32_t sr1 = lx + 1u < wx ? bits[li1 + RS] & 1u : hsp[ly0 + 1u];
    const uint32_t su0 = PART == 2 && ly0 + 1u >= wy ? hsp[T + lx] : b1 & 1u;
    const uint32_t su1 = ly0 + 2u < wy ? bits[li1 + 1u] & 1u : hsp[T + lx];
    // satisfied: J s_a s_b = +1, i.e. equal spins xor J < 0 (absent: never)
    const uint32_t sh0 = v0 ? ~(b0 ^ sr0 ^ (b0 >> 4)) & 1u : 0u;
    const uint32_t sv0 = v0 ? ~(b0 ^ su0 ^ (b0 >> 5)) & 1u : 0u;
    const uint32_t sh1 = v1 ? ~(b1 ^ sr1 ^ (b1 >> 4)) & 1u : 0u;
    const uint32_t sv1 = v1 ? ~(b1 ^ su1 ^ (b1 >> 5)) & 1u : 0u;
    if (meas_em) {
      if (!CUT) {
        me += 4 - 2 * (int32_t)(sh0 + sv0 + sh1 + sv1);
        mm += 2 * (int32_t)((b0 & 1u) + (b1 & 1u)) - 2;
      } else {
        if (v0) me += 2 - 2 * (int32_t)(sh0 + sv0), mm += 2 * (int32_t)(b0 & 1u) - 1;
        if (v1) me += 2 - 2 * (int32_t)(sh1 + sv1), mm += 2 * (int32_t)(b1 & 1u) - 1;
      }
    }
    // horizontal / vertical draws of the two sites: words x, y of the even
    // and z, w of the odd site of a Philox pair; at odd L a pair starting at
    // an odd site takes z, w of its pair and x, y of the next one
    const rbim_u32x4 r = rbim_sweep_draws(seed, lid, (uint32_t)g, n0 >> 1);
    uint32_t dh0 = r.x, dv0 = r.y, dh1 = r.z, dv1 = r.w;
    if ((PART == 2 && (n0 & 1u)) || odd_rows) {
      const rbim_u32x4 r2 = rbim_sweep_draws(seed, lid, (uint32_t)g, (n0 >> 1) + 1u);
      dh0 = r.z;
      dv0 = r.w;
      dh1 = r2.x;
      dv1 = r2.y;
    }
    const uint32_t oh0 = any_bond && dh0 <= tlim ? sh0 : 0u;
    const uint32_t ov0 = any_bond && dv0 <= tlim ? sv0 : 0u;
    const uint32_t oh1 = any_bond && dh1 <= tlim ? sh1 : 0u;
    const uint32_t ov1 = any_bond && dv1 <= tlim ? sv1 : 0u;
    // bit 0 (the spin other lanes read as a neighbour) is rewritten unchanged
    bits[li0] = (uint8_t)((b0 & 1u) | (oh0 << 1) | (ov0 << 2) | ((dh0 & 1u) << 3));
    bits[li1] = (uint8_t)((b1 & 1u) | (oh1 << 1) | (ov1 << 2) | ((dh1 & 1u) << 3));
    // n0 -> n1, same row and tile (for a cut row's last site, the bond leaves)
    const bool internal = ov0 != 0u && (PART != 2 || ly0 + 1u < wy);
    const bool outlink = ov1 != 0u && ly0 + 2u < wy;  // n1 -> n1+1 inside the tile
    // local index of the wave's lane 0 as seen from this lane's row (a run
    // never leaves its row)
    const uint32_t chunk = PART ? li0 - 2u * (uint32_t)lane : 2u * (uint32_t)(it * NTH + wave * 64);
    const uint32_t start0 = run_start(__ballot(internal && outlink), __ballot(outlink), below, chunk);
    rsa[it] = start0;
    rsb[it] = internal ? start0 : li1;
    par[li0] = H(start0);
    par[li1] = rsb[it] << 2;
  }
  if (meas_em) wave_spins(me, mm, red);
  __syncthreads();
  if (meas_em && tid == 0) {
    int64_t e = 0, m = 0;
    for (int w = 0; w < NTH / 64; ++w) {
      e += red[w][0];
      m += red[w][1];
    }
    int64_t* o = f.part_em + ((size_t)lat * tiles + tile) * 2;
    o[0] = e;
    o[1] = m;
  }

  // remaining in-tile bonds: x bonds (plaquette skip) and the y bond leaving a
  // wave's chunk, through the wave's queue in lcnt (cleared before the counts
  // use it), worked when the next pass could overflow it and after the last
  uint32_t* wq = lcnt + wave * QW;
  uint32_t qn = 0u;
#pragma unroll UB
  for (int it = 0; it < ITERS; ++it) {
    if (skip(it)) continue;
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t lx = row_of(ql), ly0 = col_of(ql);
    const uint32_t li0 = li_of(ql), li1 = li0 + 1u;
    const uint32_t b0 = bits[li0], b1 = bits[li1];
    uint32_t req = 0u;
    if (lx + 1u < wx) {
      const bool first = ly0 == 0u;  // no plaquette to the left
      const uint32_t lf = first ? li0 : li0 - 1u;
      const uint32_t left = bits[lf], left_up = bits[lf + RS];
      const uint32_t h0 = (b0 >> 1) & 1u &
                          ~((left >> 1) & (left >> 2) & (left_up >> 2) & (first ? 0u : 1u));
      const uint32_t h1 = (b1 >> 1) & 1u & ~((b0 >> 1) & (b0 >> 2) & (bits[li0 + RS] >> 2));
      req = h0 | (h1 << 1);
    }
    req |= ((b1 >> 2) & 1u & ((lane == 63 && ly0 + 2u < wy) ? 1u : 0u)) << 2;
    // the unions run at issue priority 1 (L=256 +1.4 %, L=200 +2.8 %; the root
    // walks below gain nothing from it)
    uint32_t qa[3] = {H(li0), H(li1), H(li1)};
    uint32_t qb[3] = {H(li0 + RS), H(li1 + RS), H(li1 + 1u)};
    if constexpr (RSK) {
      qa[0] = H(rsa[it]);
      qa[1] = qa[2] = H(rsb[it]);
    }
    if constexpr (RSK) {
      // the row above from its parent words, read here ahead of the dependent
      // finds (any parent word names an ancestor; the forest has no rep bits
      // yet)
      if (lx + 1u < wx) {
        qb[0] = par[li0 + RS];
        qb[1] = par[li1 + RS];
      }
    }
    if (qn + WAVE_QUEUE_PASS > QW) {
      wave_queue_run<true, true>(par, wq, qn);
      qn = 0u;
    }
    qn = wave_queue_put(wq, qn, req, qa, qb, below);
  }
  wave_queue_run<true, true>(par, wq, qn);
  {  // this wave's queue, cleared with 16-byte stores (QW is a multiple of 4)
    uint4* q4 = reinterpret_cast<uint4*>(lcnt + wave * QW);
    for (uint32_t i = lane; i < QW / 4u; i += 64u) q4[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();
  // roots: counts (size << 16 | n_plus before the flip); every site is pointed
  // straight at its root (the forest is settled, so this is only compression)
  // (v0, v1: the sites are present; absent ones add to no count)
  auto count_and_point = [&](uint32_t li0, uint32_t b0, uint32_t b1, uint32_t r0, uint32_t r1,
                             bool v0, bool v1) {
    const uint32_t li1 = li0 + 1u;
    // the first lane's cluster (often most of the wave) is summed by ballots
    // and added by one atomic
    const uint32_t R = __builtin_amdgcn_readfirstlane(r0);
    const bool in0 = v0 && r0 == R, in1 = v1 && r1 == R;
    const uint64_t m0 = __ballot(in0), m1 = __ballot(in1);
    const uint64_t u0 = __ballot(in0 && (b0 & 1u)), u1 = __ballot(in1 && (b1 & 1u));
    if (lane == 0)
      atomicAdd(&lds_at<true>(lcnt, R), (uint32_t)(__popcll(m0) + __popcll(m1)) << 16 |
                              (uint32_t)(__popcll(u0) + __popcll(u1)));
    const bool o0 = v0 && !in0, o1 = v1 && !in1;
    if (o0 && o1 && r0 == r1) {
      atomicAdd(&lds_at<true>(lcnt, r0), 0x20000u | ((b0 & 1u) + (b1 & 1u)));
    } else {
      if (o0) atomicAdd(&lds_at<true>(lcnt, r0), 0x10000u | (b0 & 1u));
      if (o1) atomicAdd(&lds_at<true>(lcnt, r1), 0x10000u | (b1 & 1u));
    }
    // every site now points straight at its root; a root's word becomes
    // root | rep << 16 with rep = "none" until an edge slot claims it (the
    // walks read parents through a 16-bit mask)
    par[li0] = r0 != H(li0) ? r0 : H(li0) | (NONE16 << 16);
    par[li1] = r1 != H(li1) ? r1 : H(li1) | (NONE16 << 16);
  };
  if constexpr (ITERS % 2 == 0) {
    // two pairs per step: four root walks in flight together (from the run
    // starts in a whole tile)
#pragma unroll UB
    for (int it = 0; it < ITERS; it += 2) {
      if (skip(it)) continue;  // and so is it + 1
      const uint32_t qla = (uint32_t)(it * NTH + tid), qlb = qla + NTH;
      const uint32_t la = li_of(qla), lb = li_of(qlb);
      const uint32_t ba0 = bits[la], ba1 = bits[la + 1u];
      bool va0 = true, va1 = true, vb0 = true, vb1 = true;
      if (CUT) {
        const uint32_t xa = row_of(qla), ya = col_of(qla), xb = row_of(qlb), yb = col_of(qlb);
        va0 = xa < wx && ya < wy, va1 = xa < wx && ya + 1u < wy;
        vb0 = xb < wx && yb < wy, vb1 = xb < wx && yb + 1u < wy;
        if (skip(it + 1)) {  // the second pair's wavefront is past the cut
          uint32_t r0 = H(la), r1 = H((ba0 & 4u) && va1 ? la : la + 1u);
          troot16_2(par, r0, r1);
          count_and_point(la, ba0, ba1, r0, r1, va0, va1);
          continue;
        }
      }
      const uint32_t bb0 = bits[lb], bb1 = bits[lb + 1u];
      // li1 shares li0's run when the in-row bond is occupied
      uint32_t r[4] = {H(la), H((ba0 & 4u) && va1 ? la : la + 1u), H(lb),
                       H((bb0 & 4u) && vb1 ? lb : lb + 1u)};
      if constexpr (RSK) {
        r[0] = H(rsa[it]);
        r[1] = H(rsb[it]);
        r[2] = H(rsa[it + 1]);
        r[3] = H(rsb[it + 1]);
      }
      troot16_4(par, r);
      if constexpr (RSK) {  // the registers now keep the sites' roots (handles)
        rsa[it] = r[0];
        rsb[it] = r[1];
        rsa[it + 1] = r[2];
        rsb[it + 1] = r[3];
      }
      count_and_point(la, ba0, ba1, r[0], r[1], va0, va1);
      count_and_point(lb, bb0, bb1, r[2], r[3], vb0, vb1);
    }
  } else {
#pragma unroll 1
    for (int it = 0; it < ITERS; ++it) {
      if (skip(it)) continue;
      const uint32_t ql = (uint32_t)(it * NTH + tid);
      const uint32_t li0 = li_of(ql), li1 = li0 + 1u;
      const uint32_t b0 = bits[li0], b1 = bits[li1];
      const bool v0 = !CUT || (row_of(ql) < wx && col_of(ql) < wy);
      const bool v1 = !CUT || (row_of(ql) < wx && col_of(ql) + 1u < wy);
      uint32_t r0 = H(li0), r1 = H((b0 & 4u) && v1 ? li0 : li1);  // li1 shares li0's run when linked
      troot16_2(par, r0, r1);
      count_and_point(li0, b0, b1, r0, r1, v0, v1);
    }
  }
  __syncthreads();
  // rep slot of every edge cluster: the root's word = root | min slot << 16
  for (uint32_t k = tid; k < NS; k += NTH) {
    bool ok = true;
    const uint32_t li = CUT ? slot_li_cut<T>(k, wx, wy, rsh, ok) : slot_li<T>(k);
    if (!ok) continue;
    const uint32_t r = par[li] & 0xFFFFu;
    atomicMin(&lds_at<true>(par, r), r | (k << 16));
  }
  __syncthreads();

  // new state: interior clusters flipped now, edge clusters labelled.  Interior
  // clusters are ranked by a 32-bit key size << 12 | (4095 - root index): inside
  // one tile the local index orders like the site; the winners' n_plus is
  // looked up afterwards.
  uint32_t ik1 = 0u, ik2 = 0u, incl = 0u;
#pragma unroll
  for (int it = 0; it < ITERS; ++it) {
    if (skip(it)) continue;
    const uint32_t ql = (uint32_t)(it * NTH + tid);
    const uint32_t lx = row_of(ql), ly0 = col_of(ql);
    const uint32_t li0 = li_of(ql);
    const uint32_t n0 = (x0 + lx) * L + y0 + ly0;
    const bool vs[2] = {!CUT || (lx < wx && ly0 < wy), !CUT || (lx < wx && ly0 + 1u < wy)};
    uint32_t st2 = 0u;  // the pair's new state words
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t li = li0 + (uint32_t)j;
      const uint32_t r = RSK ? (j == 0 ? rsa[it] : rsb[it]) : par[li] & 0xFFFFu;
      const uint32_t rp = lds_get<true>(par, r);
      const uint32_t rep = rp >> 16;
      const uint32_t br = bits[IDX(r)];
      const uint32_t fl = (br >> 3) & 1u;
      const uint32_t sbit = bits[li] & 1u;
      const bool edge = rep != NONE16;
      const uint32_t ns = edge ? sbit : sbit ^ fl;
      st2 |= ((ns ? ST_SPIN : 0u) | (edge ? ST_EDGE | rep : 0u)) << (16 * j);
      if (meas_cl && !edge && r == H(li) && vs[j]) {  // an interior cluster, at its root
        const uint32_t key = ((lds_get<true>(lcnt, r) >> 16) << 12) | (0xFFFu - li);
        ++incl;
        ik2 = max(ik2, min(ik1, key));
        ik1 = max(ik1, key);
      }
    }
    if (PART == 0 || (PART == 1 && vs[0]) || (PART == 3 && !odd_rows)) {
      *reinterpret_cast<uint32_t*>(v.out + n0) = st2;
    } else if (PART == 3) {
      v.out[n0] = (uint16_t)(st2 & 0xFFFFu);
      v.out[n0 + 1u] = (uint16_t)(st2 >> 16);
    } else if (PART == 2) {
      if (vs[0]) v.out[n0] = (uint16_t)(st2 & 0xFFFFu);
      if (vs[1]) v.out[n0 + 1u] = (uint16_t)(st2 >> 16);
    }
  }
  // edge slots for KB, and the occupation of the bonds leaving the tile
  if (BIG) {
    // the slot forest in HBM starts here: every slot under its rep, reps
    // roots that carry their cluster's minimum site in the word's upper half,
    // their counts, and (min site << 1 | flip)
    const size_t sb = (size_t)lat * f.nslots + (size_t)tile * NS;
    for (uint32_t k = tid; k < NS; k += NTH) {
      bool ok = true;
      const uint32_t li = CUT ? slot_li_cut<T>(k, wx, wy, rsh, ok) : slot_li<T>(k);
      if (!ok) {  // no site here: a slot of its own, never folded or counted
        f.guf[sb + k] = gpack(tile * NS + k, 0, 0);
        f.gmn[sb + k] = 0xFFFFFFFFu;
        continue;
      }
      const uint32_t r = par[li] & 0xFFFFu;
      const uint32_t rep = lds_get<true>(par, r) >> 16;
      if (rep == k) {
        const uint32_t c = lds_get<true>(lcnt, r);
        f.gcnt[sb + k] = ((uint64_t)(c >> 16) << 32) | (c & 0xFFFFu);
        const uint32_t rt = tile_li(IDX(r));
        const uint32_t site = (x0 + rt / T) * L + y0 + rt % T;
        f.guf[sb + k] = groot_word(tile * NS + k, site);
        f.gmn[sb + k] = (site << 1) | ((bits[IDX(r)] >> 3) & 1u);
      } else {
        f.guf[sb + k] = gpack(tile * NS + rep, 0, 0);
        f.gmn[sb + k] = 0xFFFFFFFFu;
      }
    }
  } else {
    uint64_t* sl = f.slots + (size_t)lat * f.nslots + (size_t)tile * NS;
    for (uint32_t k = tid; k < NS; k += NTH) {
      bool ok = true;
      const uint32_t li = CUT ? slot_li_cut<T>(k, wx, wy, rsh, ok) : slot_li<T>(k);
      if (!ok) {
        sl[k] = SLOT_ABSENT;
        continue;
      }
      const uint32_t r = par[li] & 0xFFFFu;
      const uint32_t rep = lds_get<true>(par, r) >> 16;
      const uint32_t c = lds_get<true>(lcnt, r);
      sl[k] = slot_entry(rep, tile_li(IDX(r)), (bits[IDX(r)] >> 3) & 1u, c >> 16, c & 0xFFFFu);
    }
  }
  if (tid < 2 * T) {
    // tid < T: +x bond of (x0+wx-1, y0+tid); else +y bond of (x0+tid-T, y0+wy-1)
    // (a cut tile reads only the positions that hold sites: the rows and
    // columns past the cut lie beyond its LDS)
    const uint32_t i = tid < T ? (uint32_t)tid : (uint32_t)tid - T;
    uint32_t occ = 0u;
    if (!CUT || i < (tid < T ? wy : wx))
      occ = tid < T ? (bits[((wx - 1u) << rsh) + i] >> 1) & 1u
                    : (bits[(i << rsh) + wy - 1u] >> 2) & 1u;
    const uint64_t word = __ballot(occ != 0u);
    if (lane == 0) f.xbits[((size_t)lat * tiles + tile) * (2 * T / 64) + wave] = word;
  }
  if (meas_cl) {
    wave_top2(ik1, ik2);
    incl = wave_sum63(incl);
    if (lane == 63) {
      red[wave][2] = incl;
      red[wave][3] = ik1;
      red[wave][4] = ik2;
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t k1 = 0u, k2 = 0u;
      int64_t ncl = 0;
      for (int w = 0; w < NTH / 64; ++w) {
        ncl += red[w][2];
        const uint32_t o1 = (uint32_t)red[w][3], o2 = (uint32_t)red[w][4];
        k2 = max(min(k1, o1), max(k2, o2));
        k1 = max(k1, o1);
      }
      int64_t* o = f.part_int + (((size_t)(g & 1) * s.n_lat + lat) * tiles + tile) * 5;
      o[0] = ncl;
      const uint32_t ks[2] = {k1, k2};
      for (int q = 0; q < 2; ++q) {
        int64_t key = 0, np_after = 0;
        if (ks[q] != 0u) {
          const uint32_t li = 0xFFFu - (ks[q] & 0xFFFu);
          const uint32_t c = lcnt[li], size = c >> 16, np = c & 0xFFFFu;
          const uint32_t lt = tile_li(li);
          const uint32_t site = (x0 + lt / T) * L + y0 + lt % T;
          key = (int64_t)(((uint64_t)size << 32) | (uint64_t)(0xFFFFFFFFu - site));
          np_after = ((bits[li] >> 3) & 1u) ? size - np : np;
        }
        o[1 + q] = key;
        o[3 + q] = np_after;
      }
    }
  }
}

// The whole tiles of a lattice: all of them when T divides L (PART 0), else
// all but the last tile row and column (PART 0 at even L, 3 at odd L, whose
// odd rows start at odd sites), the grid then covering (nt - 1)^2 tiles.
// Eight wavefronts per SIMD: four 512-thread workgroups of 64 x 64 tiles (their
// LDS allows four per CU) or eight of 32 x 32 tiles per CU, which caps the
// scalar registers below what the code would otherwise take at odd L (100,
// enough for seven).
template <int T, int NTH, bool BIG, int PART>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(8, 8))) void ka_tile(
    LatticeSet s, FusedScratch f) {
  __shared__ KaShared<T, NTH> sm;
  const KaLds<NTH> m{sm.par, sm.lcnt, sm.bits, sm.hsp, sm.red, (uint32_t)(T * T / (NTH / 64))};
  const uint32_t nt = (uint32_t)f.tiles_x, b = blockIdx.x;
  const uint32_t tile = gridDim.x == nt * nt ? b : (b / (nt - 1u)) * nt + b % (nt - 1u);
  ka_body<T, NTH, BIG, PART>(s, f, m, tile);
}

// The cut tiles of a lattice T does not divide (2 nt - 1 per lattice: the last
// tile column, then the rest of the last tile row), in narrower workgroups
// whose LDS is sized for the cut (cap sites): most of a cut tile's wavefronts
// would have nothing to do, and a whole tile's LDS would hold the CU's
// workgroup slots as long as a whole tile does.
template <int T, int NTH, bool BIG, int PART>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(6, 8))) void ka_tile_cut(
    LatticeSet s, FusedScratch f, uint32_t cap) {
  extern __shared__ uint64_t ka_dyn[];
  KaLds<NTH> m;
  m.par = reinterpret_cast<uint32_t*>(ka_dyn);
  m.lcnt = m.par + cap;
  m.bits = reinterpret_cast<uint8_t*>(m.lcnt + cap);
  m.hsp = m.bits + cap;
  m.red = reinterpret_cast<int64_t(*)[PARTS]>(m.hsp + ((2u * T + 7u) & ~7u));
  m.qw = cap / (NTH / 64);
  const uint32_t nt = (uint32_t)f.tiles_x, c = blockIdx.x;
  const uint32_t tile = c < nt ? c * nt + (nt - 1u) : (nt - 1u) * nt + (c - nt);
  ka_body<T, NTH, BIG, PART>(s, f, m, tile);
}

// Dynamic LDS bytes of a cut-tile pass of cap sites.
template <int T, int NTH>
constexpr size_t cut_lds_bytes(uint32_t cap) {
  return (size_t)cap * 9u + ((2u * T + 7u) & ~7u) + (size_t)(NTH / 64) * PARTS * 8u;
}

// Fold measured sweep gm of lattice lat into its bin: E, M from part_em,
// interior clusters from part_int[gm & 1], edge clusters and wraps from
// bst[gm & 1].  Single thread.
// The sweep record of measured sweep gm from its totals (E, M, cluster count,
// the two largest clusters, the wrap word), folded into its bin.  Single
// thread.
__device__ void fused_fold(const LatticeSet& s, int lat, int64_t gm, int64_t e, int64_t m,
                           int64_t ncl, const TileMeas& a, int64_t wrapw) {
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
  st.wrap_x = (int32_t)(wrapw & 1);
  st.wrap_y = (int32_t)((wrapw >> 1) & 1);
  const int64_t bin = (gm - lat_therm(s, lat)) / lat_spb(s, lat);
  acc_sweep(s.acc + ((size_t)lat * s.n_bins + bin) * RBIM_ACC_SLOTS, st);
  if (s.last) s.last[lat] = st;
  series_put(s, lat, gm, st);
}

// Fold measured sweep gm of lattice lat into its bin: E, M from part_em,
// interior clusters from part_int[gm & 1], edge clusters and wraps from
// bst[gm & 1].  Single thread.
__device__ void fused_finalize(const LatticeSet& s, const FusedScratch& f, int lat, int64_t gm) {
  const int tiles = f.tiles_x * f.tiles_x;
  const int64_t* em = f.part_em + (size_t)lat * tiles * 2;
  const int64_t* pi = f.part_int + ((size_t)(gm & 1) * s.n_lat + lat) * tiles * 5;
  const int64_t* eb = f.bst + ((size_t)(gm & 1) * s.n_lat + lat) * 6;
  int64_t e = 0, m = 0, ncl = eb[0];
  TileMeas a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
  top2_insert(a, (uint64_t)eb[1], (uint32_t)eb[3]);
  top2_insert(a, (uint64_t)eb[2], (uint32_t)eb[4]);
  for (int i = 0; i < tiles; ++i) {
    e += em[2 * i];
    m += em[2 * i + 1];
    ncl += pi[5 * i];
    top2_insert(a, (uint64_t)pi[5 * i + 1], (uint32_t)pi[5 * i + 3]);
    top2_insert(a, (uint64_t)pi[5 * i + 2], (uint32_t)pi[5 * i + 4]);
  }
  fused_fold(s, lat, gm, e, m, ncl, a, eb[5]);
}

// Sum of a workgroup's (e, m, n, top-2) partials: wave shuffles, then thread
// 0 combines the waves' rows of red (NW <= 16).
template <int NTHR>
__device__ void block_meas(int64_t& e, int64_t& m, int64_t& n, TileMeas& a, int64_t (*red)[PARTS]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    m += __shfl_xor(m, off);
    n += __shfl_xor(n, off);
    const uint64_t o1 = __shfl_xor(a.k1, off), o2 = __shfl_xor(a.k2, off);
    const uint32_t q1 = __shfl_xor(a.p1, off), q2 = __shfl_xor(a.p2, off);
    top2_insert(a, o1, q1);
    top2_insert(a, o2, q2);
  }
  if ((threadIdx.x & 63) == 0) {
    int64_t* r = red[threadIdx.x >> 6];
    r[0] = e;
    r[1] = m;
    r[2] = n;
    r[3] = (int64_t)a.k1;
    r[4] = (int64_t)a.k2;
    r[5] = a.p1;
    r[6] = a.p2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    e = m = n = 0;
    a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
    for (int w = 0; w < NTHR / 64; ++w) {
      e += red[w][0];
      m += red[w][1];
      n += red[w][2];
      top2_insert(a, (uint64_t)red[w][3], (uint32_t)red[w][5]);
      top2_insert(a, (uint64_t)red[w][4], (uint32_t)red[w][6]);
    }
  }
}

// fused_finalize with its per-tile sums spread over a workgroup of NTHR
// threads, for lattices of many tiles (every thread calls it; red is
// overwritten).
template <int NTHR>
__device__ void fused_finalize_block(const LatticeSet& s, const FusedScratch& f, int lat,
                                     int64_t gm, int64_t (*red)[PARTS]) {
  const int tiles = f.tiles_x * f.tiles_x;
  const int64_t* em = f.part_em + (size_t)lat * tiles * 2;
  const int64_t* pi = f.part_int + ((size_t)(gm & 1) * s.n_lat + lat) * tiles * 5;
  int64_t e = 0, m = 0, ncl = 0;
  TileMeas a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
  for (int i = (int)threadIdx.x; i < tiles; i += NTHR) {
    e += em[2 * i];
    m += em[2 * i + 1];
    ncl += pi[5 * i];
    top2_insert(a, (uint64_t)pi[5 * i + 1], (uint32_t)pi[5 * i + 3]);
    top2_insert(a, (uint64_t)pi[5 * i + 2], (uint32_t)pi[5 * i + 4]);
  }
  block_meas<NTHR>(e, m, ncl, a, red);
  if (threadIdx.x == 0) {
    const int64_t* eb = f.bst + ((size_t)(gm & 1) * s.n_lat + lat) * 6;
    top2_insert(a, (uint64_t)eb[1], (uint32_t)eb[3]);
    top2_insert(a, (uint64_t)eb[2], (uint32_t)eb[4]);
    fused_fold(s, lat, gm, e, m, ncl + eb[0], a, eb[5]);
  }
}

// ------------------------------------------------------------------- KB
// One workgroup per lattice merges its edge slots in LDS, one 64-bit word per
// slot (at most 8192 slots and 131071 sites; 32 KB at L = 256's 4096 slots,
// so four workgroups share a CU):
//   non-root  parent (bits 0..12) | wx (13..28) | wy (29..44), the int16
//             windings of the tree path to the parent
//   root      itself (0..12) | its cluster's minimum site (13..29) | size
//             (30..46) | n_plus before the flip (47..63); size 0 marks a
//             position of a cut tile that holds no site
// Roots hook the one with the larger minimum site under the smaller, so every
// root is the rep holding its cluster's minimum site, whose flip bit (the
// cluster's) stays in the root's slot entry; the reps that end up below a root
// add their counts to its word.  A winding is that of a path in the spanning
// forest of the hooking bonds, a simple path, which crosses each of the L
// boundary bonds of a direction at most once: |w| <= L fits 16 bits.
constexpr uint32_t KC_P = 0x1FFFu;
// every lattice merged in LDS fits: at most 6400 slots, L <= 320 at T = 64
static_assert(RBIM_FUSED_MAX_SLOTS <= (int)KC_P + 1 && 320 * 320 < (1 << 17), "slot word fields");
__device__ __forceinline__ uint64_t kc_link(uint32_t p, int wx, int wy) {
  return (uint64_t)p | ((uint64_t)(uint16_t)(int16_t)wx << 13) |
         ((uint64_t)(uint16_t)(int16_t)wy << 29);
}
__device__ __forceinline__ uint64_t kc_root(uint32_t k, uint32_t site, uint32_t size, uint32_t np) {
  return (uint64_t)k | ((uint64_t)site << 13) | ((uint64_t)size << 30) | ((uint64_t)np << 47);
}
__device__ __forceinline__ uint32_t kc_par(uint64_t v) { return (uint32_t)v & KC_P; }
__device__ __forceinline__ int kc_wx(uint64_t v) { return (int)(int16_t)(uint16_t)(v >> 13); }
__device__ __forceinline__ int kc_wy(uint64_t v) { return (int)(int16_t)(uint16_t)(v >> 29); }
__device__ __forceinline__ uint32_t kc_site(uint64_t v) { return (uint32_t)(v >> 13) & 0x1FFFFu; }
__device__ __forceinline__ uint32_t kc_size(uint64_t v) { return (uint32_t)(v >> 30) & 0x1FFFFu; }
__device__ __forceinline__ uint32_t kc_np(uint64_t v) { return (uint32_t)(v >> 47) & 0x1FFFFu; }

// kfind2 on compact words; va / vb end as the roots' words.
__device__ __forceinline__ void kc_find2(uint64_t* uf, uint32_t& a, uint64_t& va, int& ax, int& ay,
                                         uint32_t& b, uint64_t& vb, int& bx, int& by) {
  va = uf[a];
  vb = uf[b];
  while (true) {
    const uint32_t pa = kc_par(va), pb = kc_par(vb);
    const bool da = pa == a, db = pb == b;
    if (da && db) return;
    const uint64_t wa = da ? va : uf[pa];
    const uint64_t wb = db ? vb : uf[pb];
    if (!da) {
      const uint32_t ga = kc_par(wa);
      if (ga == pa) {
        ax += kc_wx(va);
        ay += kc_wy(va);
        a = pa;
        va = wa;
      } else {
        const int hx = kc_wx(va) + kc_wx(wa), hy = kc_wy(va) + kc_wy(wa);
        uf[a] = kc_link(ga, hx, hy);
        ax += hx;
        ay += hy;
        a = ga;
        va = uf[ga];
      }
    }
    if (!db) {
      const uint32_t gb = kc_par(wb);
      if (gb == pb) {
        bx += kc_wx(vb);
        by += kc_wy(vb);
        b = pb;
        vb = wb;
      } else {
        const int hx = kc_wx(vb) + kc_wx(wb), hy = kc_wy(vb) + kc_wy(wb);
        uf[b] = kc_link(gb, hx, hy);
        bx += hx;
        by += hy;
        b = gb;
        vb = uf[gb];
      }
    }
  }
}

__device__ __forceinline__ uint32_t kc_rootof(const uint64_t* uf, uint32_t x) {
  uint32_t p = kc_par(uf[x]);
  while (p != x) {
    x = p;
    p = kc_par(uf[x]);
  }
  return x;
}

// Occupied bond a -> b crossing the periodic boundary (cx, cy) times; wrap
// bits when it closes a loop.  A root's word changes only when it is hooked,
// so a CAS against the word its find read fails exactly when it was.
__device__ uint32_t kc_union(uint64_t* uf, uint32_t a, uint32_t b, int cx, int cy) {
  int ax = 0, ay = 0, bx = 0, by = 0;
  while (true) {
    uint64_t va, vb;
    kc_find2(uf, a, va, ax, ay, b, vb, bx, by);
    if (a == b) return ((cx + bx - ax) != 0 ? 1u : 0u) | ((cy + by - ay) != 0 ? 2u : 0u);
    if (kc_site(va) < kc_site(vb)) {
      if (atomicCAS((unsigned long long*)&uf[b], (unsigned long long)vb,
                    (unsigned long long)kc_link(a, ax - bx - cx, ay - by - cy)) == vb)
        return 0u;
    } else {
      if (atomicCAS((unsigned long long*)&uf[a], (unsigned long long)va,
                    (unsigned long long)kc_link(b, bx - ax + cx, by - ay + cy)) == va)
        return 0u;
    }
  }
}

template <int T, int NTH>
__global__ __launch_bounds__(NTH) void kb_merge(LatticeSet s, FusedScratch f) {
  constexpr uint32_t NS = 4u * T;
  // the last wavefront folds the previous sweep into its bin (one thread, a
  // few tiles per lattice) while the others merge
  constexpr int WORK = NTH - 64;
  extern __shared__ uint64_t kbsm[];  // one word per slot; the reduction rows at the end
  const int lat = run_lat(s, blockIdx.x);
  if (!lattice_active(s, lat)) return;
  const uint32_t ns = (uint32_t)f.nslots, nt = (uint32_t)f.tiles_x, L = (uint32_t)s.L;
  uint64_t* uf = kbsm;
  const int tid = threadIdx.x;
  const int64_t g = s.sweep_ctr[lat];
  const uint64_t* ent = f.slots + (size_t)lat * ns;
  const bool worker = tid < WORK;
  if (tid == WORK && g >= 1 && f.fin[lat] == g - 1) {
    if (g - 1 >= lat_therm(s, lat)) fused_finalize(s, f, lat, g - 1);
    f.fin[lat] = g;
  }
  for (uint32_t k = tid; worker && k < ns; k += WORK) {
    const uint64_t e = ent[k];
    const uint32_t t = k / NS, loc = k % NS, rep = (uint32_t)(e & 0xFFFu);
    if (e & SLOT_ABSENT) {
      uf[k] = kc_root(k, 0u, 0u, 0u);
    } else if (rep == loc) {
      const uint32_t r = (uint32_t)(e >> 12) & 0xFFFu;
      const uint32_t size = (uint32_t)(e >> 25) & 0x1FFFu, np = (uint32_t)(e >> 38) & 0x1FFFu;
      const uint32_t site = ((t % nt) * T + r / T) * L + (t / nt) * T + r % T;
      uf[k] = kc_root(k, site, size, np);
    } else {
      uf[k] = kc_link(t * NS + rep, 0, 0);
    }
  }
  __syncthreads();
  // tile-crossing bonds: bond i of tile t (i < T: the +x bond of slot (side 1,
  // i) into the right neighbour's (side 0, i); else the +y bond of (side 3,
  // i - T) into the upper neighbour's (side 2, i - T)).  nb and WORK are
  // multiples of 64, so every lane of a working wavefront holds a bond, lane l
  // - 1 the bond b - 1.  An occupied bond whose two slots have the same reps
  // (from the tile pass's entries) as the occupied bond before it on the same
  // edge is skipped: with that bond (or the first kept one of the run) it
  // closes a loop through the two tiles' own clusters that crosses the edge
  // once each way -- winding 0, no new connection.
  const uint32_t nb = (ns / NS) * 2u * T;
  uint32_t wraps = 0u;
  const int lane = tid & 63;
  for (uint32_t b = tid; worker && b < nb; b += WORK) {
    const uint32_t t = b / (2u * T), i = b % (2u * T);
    const uint64_t word = f.xbits[((size_t)lat * (ns / NS) + t) * (2 * T / 64) + i / 64];
    const bool occ = (word >> (i % 64)) & 1ull;
    const uint32_t tx = t % nt, ty = t / nt;
    uint32_t a, c;
    int cx = 0, cy = 0;
    if (i < (uint32_t)T) {
      a = t * NS + T + i;
      c = (ty * nt + (tx + 1u == nt ? 0u : tx + 1u)) * NS + i;
      cx = tx + 1u == nt ? 1 : 0;
    } else {
      const uint32_t j = i - T;
      a = t * NS + 3u * T + j;
      c = ((ty + 1u == nt ? 0u : ty + 1u) * nt + tx) * NS + 2u * T + j;
      cy = ty + 1u == nt ? 1 : 0;
    }
    uint32_t ra = 0u, rc = 0u;
    if (occ) {
      ra = (uint32_t)(ent[a] & 0xFFFu);  // reps within the slots' own tiles
      rc = (uint32_t)(ent[c] & 0xFFFu);
    }
    const int pocc = __shfl_up(occ ? 1 : 0, 1);
    const uint32_t pra = (uint32_t)__shfl_up((int)ra, 1), prc = (uint32_t)__shfl_up((int)rc, 1);
    if (!occ || (lane != 0 && (i % T) != 0u && pocc && pra == ra && prc == rc)) continue;
    // the finds start at the reps (a slot sits under its rep with winding 0)
    wraps |= kc_union(uf, (a / NS) * NS + ra, (c / NS) * NS + rc, cx, cy);
  }
  const uint64_t wmx = __ballot((wraps & 1u) != 0u), wmy = __ballot((wraps & 2u) != 0u);
  const uint32_t wave_wraps = (wmx ? 1u : 0u) | (wmy ? 2u : 0u);
  __syncthreads();
  // reps below a root add their counts to its word and point straight at it
  // (the windings are no longer read); the root word's parent bits never
  // change, so walks meanwhile still end there
  for (uint32_t k = tid; worker && k < ns; k += WORK) {
    const uint64_t e = ent[k];
    if ((e & SLOT_ABSENT) || (uint32_t)(e & 0xFFFu) != k % NS) continue;
    const uint32_t r = kc_rootof(uf, k);
    if (r != k) {
      const uint64_t size = (e >> 25) & 0x1FFFu, np = (e >> 38) & 0x1FFFu;
      atomicAdd((unsigned long long*)&uf[r], (unsigned long long)((size << 30) | (np << 47)));
      uf[k] = kc_link(r, 0, 0);
    }
  }
  __syncthreads();
  // flip bits per slot; merged clusters measured at their roots
  const bool meas_cl = g >= lat_therm(s, lat);
  TileMeas ms = {0, 0, 0, 0ull, 0ull, 0u, 0u};
  uint32_t* fbo = f.flipbits + (size_t)lat * (ns / 32u);
  for (uint32_t k0 = 0; worker && k0 < ns; k0 += WORK) {
    const uint32_t k = k0 + tid;
    uint32_t fl = 0u;
    if (k < ns) {
      const uint32_t r = kc_rootof(uf, k);
      const uint64_t w = uf[r];
      const uint32_t size = kc_size(w);
      fl = (uint32_t)(ent[r] >> 24) & 1u;  // (0 at an absent position)
      if (meas_cl && r == k && size != 0u) {  // (absent slots are no cluster)
        const uint32_t np = kc_np(w);
        ++ms.ncl;
        top2_insert(ms, ((uint64_t)size << 32) | (uint64_t)(0xFFFFFFFFu - kc_site(w)),
                    fl ? size - np : np);
      }
    }
    const uint64_t word = __ballot(fl != 0u);
    if ((tid & 63) == 0 && k < ns) {
      fbo[k / 32] = (uint32_t)word;
      if (k + 32 < ns) fbo[k / 32 + 1] = (uint32_t)(word >> 32);
    }
  }
  __syncthreads();  // the slot words are dead: their first rows take the reduction
  int64_t (*red)[PARTS] = reinterpret_cast<int64_t(*)[PARTS]>(kbsm);
  if (meas_cl) wave_clusters(ms, red);
  if ((tid & 63) == 0) red[tid >> 6][0] = wave_wraps;
  __syncthreads();
  if (tid == 0) {
    if (meas_cl) {
      TileMeas a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
      int64_t ncl = 0;
      uint32_t wb = 0u;
      for (int w = 0; w < NTH / 64; ++w) {
        wb |= (uint32_t)red[w][0];
        ncl += red[w][2];
        top2_insert(a, (uint64_t)red[w][3], (uint32_t)red[w][5]);
        top2_insert(a, (uint64_t)red[w][4], (uint32_t)red[w][6]);
      }
      int64_t* o = f.bst + ((size_t)(g & 1) * s.n_lat + lat) * 6;
      o[0] = ncl;
      o[1] = (int64_t)a.k1;
      o[2] = (int64_t)a.k2;
      o[3] = a.p1;
      o[4] = a.p2;
      o[5] = wb;
    }
    f.cur[lat] ^= 1;
    s.sweep_ctr[lat] = g + 1;
  }
}

// The same merge as kb_merge for lattices whose slots do not fit one
// workgroup's LDS: the slot forest (parent | wx | wy, roots: the slot | its
// cluster's minimum site), the rep counts and the (min site << 1 | flip)
// words live in HBM, one thread per bond / slot.  Roots hook by minimum site,
// so a merged cluster's root is the rep holding its minimum site, whose flip
// bit is the cluster's: the fold adds counts only.

// gfind2 that also returns the roots' words (va, vb).
__device__ __forceinline__ void gfind2w(uint64_t* uf, uint32_t& a, uint64_t& va, int& ax, int& ay,
                                        uint32_t& b, uint64_t& vb, int& bx, int& by) {
  va = gload(&uf[a]);
  vb = gload(&uf[b]);
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
template <int T>
__global__ __launch_bounds__(256) void kg_union(LatticeSet s, FusedScratch f) {
  constexpr uint32_t NS = 4u * T;
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const uint32_t ns = (uint32_t)f.nslots, nt = (uint32_t)f.tiles_x;
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  // 2T bonds leave each tile: ns / 2 per lattice, a multiple of 64, so whole
  // wavefronts leave here and lane l - 1 holds bond i - 1 of the same tile edge
  // unless i is a multiple of T (a wavefront's first lane among them)
  if (b >= ns / 2u) return;
  const uint32_t t = b / (2u * T), i = b % (2u * T);
  const uint64_t word = f.xbits[((size_t)lat * (ns / NS) + t) * (2 * T / 64) + i / 64];
  const bool occ = (word >> (i % 64)) & 1ull;
  const uint32_t tx = t % nt, ty = t / nt;
  uint32_t a, c;
  int cx = 0, cy = 0;
  if (i < (uint32_t)T) {
    a = t * NS + T + i;
    c = (ty * nt + (tx + 1u == nt ? 0u : tx + 1u)) * NS + i;
    cx = tx + 1u == nt ? 1 : 0;
  } else {
    const uint32_t j = i - T;
    a = t * NS + 3u * T + j;
    c = ((ty + 1u == nt ? 0u : ty + 1u) * nt + tx) * NS + 2u * T + j;
    cy = ty + 1u == nt ? 1 : 0;
  }
  uint64_t* uf = f.guf + (size_t)lat * ns;
  // An occupied bond whose two slots have the same reps as the occupied bond
  // before it along the edge is skipped: with that bond (or the first kept one
  // of the run) it closes a loop through the two tiles' own clusters, which
  // crosses the edge once each way -- winding 0, no new connection.  A slot's
  // rep is itself (it carries a min-site word) or its parent as the tile pass
  // wrote it: a slot that is no rep is a leaf, whose word only its own bond's
  // find changes, after this read.
  const uint32_t* mn = f.gmn + (size_t)lat * ns;
  uint32_t ra = 0u, rc = 0u;
  if (occ) {
    ra = mn[a] != 0xFFFFFFFFu ? a : gpar(gload(&uf[a]));
    rc = mn[c] != 0xFFFFFFFFu ? c : gpar(gload(&uf[c]));
  }
  const int pocc = __shfl_up(occ ? 1 : 0, 1);
  const uint32_t pra = (uint32_t)__shfl_up((int)ra, 1), prc = (uint32_t)__shfl_up((int)rc, 1);
  if (!occ || ((i % T) != 0u && pocc && pra == ra && prc == rc)) return;
  // the finds start at the reps (a slot sits under its rep with winding 0)
  a = ra;
  c = rc;
  int ax = 0, ay = 0, bx = 0, by = 0;
  uint32_t w = 0u;
  while (true) {
    uint64_t va, vc;
    gfind2w(uf, a, va, ax, ay, c, vc, bx, by);
    if (a == c) {
      w = ((cx + bx - ax) != 0 ? 1u : 0u) | ((cy + by - ay) != 0 ? 2u : 0u);
      break;
    }
    // a root's word changes only when it is hooked: a CAS against the word
    // the find read fails exactly when it was
    if (groot_site(va) < groot_site(vc)) {
      if (atomicCAS((unsigned long long*)&uf[c], (unsigned long long)vc,
                    (unsigned long long)gpack(a, ax - bx - cx, ay - by - cy)) == vc)
        break;
    } else {
      if (atomicCAS((unsigned long long*)&uf[a], (unsigned long long)va,
                    (unsigned long long)gpack(c, bx - ax + cx, by - ay + cy)) == va)
        break;
    }
  }
  wave_or_wraps(&f.wrapw[lat], w);
}

// Reps fold their counts into their roots.
__global__ __launch_bounds__(256) void kg_fold(LatticeSet s, FusedScratch f) {
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const uint32_t ns = (uint32_t)f.nslots;
  // ns is a multiple of 128: a wavefront's 64 slots are all in or all out
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const size_t base = (size_t)lat * ns;
  const uint32_t m = f.gmn[base + k];
  bool fold = m != 0xFFFFFFFFu;
  const uint32_t r = fold ? groot(f.guf + base, k) : k;
  fold = fold && r != k;
  const uint64_t c = fold ? f.gcnt[base + k] : 0ull;
  // The folding reps of a wavefront under the first one's root (below the
  // transition nearly all: one giant cluster) add their counts by a wavefront
  // sum, and one lane makes the atomic -- per-lane atomics on one address
  // serialise in the L2.
  const uint64_t fm = __ballot(fold);
  if (fm == 0ull) return;
  const int first = __ffsll((long long)fm) - 1;
  const uint32_t R = (uint32_t)__shfl((int)r, first);
  const bool in = fold && r == R;
  unsigned long long cs = in ? c : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cs += __shfl_xor(cs, off);
  if ((int)__lane_id() == first) atomicAdd((unsigned long long*)&f.gcnt[base + R], cs);
  if (fold && !in) atomicAdd((unsigned long long*)&f.gcnt[base + r], (unsigned long long)c);
}

// Flip bit of every slot's cluster; merged clusters measured at their roots
// into per-block partials (n, key1, key2, n_plus1, n_plus2).
__global__ __launch_bounds__(256) void kg_flip(LatticeSet s, FusedScratch f) {
  __shared__ int64_t red[256 / 64][PARTS];
  const int lat = run_lat(s, blockIdx.y);
  if (!lattice_active(s, lat)) return;
  const uint32_t ns = (uint32_t)f.nslots;
  // ns is a multiple of 128: a wavefront's 64 slots are all in or all out
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = k < ns;
  const size_t base = (size_t)lat * ns;
  const bool meas_cl = s.sweep_ctr[lat] >= lat_therm(s, lat);
  // only reps walk: a site (and a halo read) names its edge cluster's rep slot,
  // so the other slots' flip bits are never read (they go out as 0)
  uint32_t r = k, m = 0xFFFFFFFFu;
  if (valid) {
    m = f.gmn[base + k];
    if (m != 0xFFFFFFFFu) {
      r = groot(f.guf + base, k);
      if (r != k) m = f.gmn[base + r];
    }
  }
  const uint32_t fl = m != 0xFFFFFFFFu ? m & 1u : 0u;
  const uint64_t word = __ballot(fl != 0u);
  if ((threadIdx.x & 63) == 0 && valid) {
    uint32_t* fbo = f.flipbits + (size_t)lat * (ns / 32u) + k / 32u;
    fbo[0] = (uint32_t)word;
    fbo[1] = (uint32_t)(word >> 32);
  }
  if (!meas_cl) return;  // block-uniform
  TileMeas ms = {0, 0, 0, 0ull, 0ull, 0u, 0u};
  if (valid && r == k && m != 0xFFFFFFFFu) {  // roots are reps (absent slots: no cluster)
    const uint64_t c = f.gcnt[base + k];
    const uint32_t size = (uint32_t)(c >> 32), np = (uint32_t)c;
    ++ms.ncl;
    top2_insert(ms, ((uint64_t)size << 32) | (uint64_t)(0xFFFFFFFFu - (m >> 1)), fl ? size - np : np);
  }
  wave_clusters(ms, red);
  __syncthreads();
  if (threadIdx.x == 0) {
    TileMeas a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
    int64_t ncl = 0;
    for (int w = 0; w < 4; ++w) {
      ncl += red[w][2];
      top2_insert(a, (uint64_t)red[w][3], (uint32_t)red[w][5]);
      top2_insert(a, (uint64_t)red[w][4], (uint32_t)red[w][6]);
    }
    int64_t* o = f.gpart + ((size_t)lat * gridDim.x + blockIdx.x) * 5;
    o[0] = ncl;
    o[1] = (int64_t)a.k1;
    o[2] = (int64_t)a.k2;
    o[3] = a.p1;
    o[4] = a.p2;
  }
}

// Per lattice (one workgroup): the merged clusters' partials and the wrap
// bits of sweep g, then the same bookkeeping as the end of kb_merge.  The
// per-block and per-tile partials are summed by all threads.
constexpr int KG_END_THREADS = 256;

__global__ __launch_bounds__(KG_END_THREADS) void kg_sweep_end(LatticeSet s, FusedScratch f,
                                                               int nblocks) {
  __shared__ int64_t red[KG_END_THREADS / 64][PARTS];
  const int lat = run_lat(s, blockIdx.x);
  if (!lattice_active(s, lat)) return;
  const int tid = threadIdx.x;
  const int64_t g = s.sweep_ctr[lat];
  if (g >= lat_therm(s, lat)) {
    int64_t e = 0, m = 0, ncl = 0;
    TileMeas a = {0, 0, 0, 0ull, 0ull, 0u, 0u};
    const int64_t* p = f.gpart + (size_t)lat * nblocks * 5;
    for (int i = tid; i < nblocks; i += KG_END_THREADS) {
      ncl += p[5 * i];
      top2_insert(a, (uint64_t)p[5 * i + 1], (uint32_t)p[5 * i + 3]);
      top2_insert(a, (uint64_t)p[5 * i + 2], (uint32_t)p[5 * i + 4]);
    }
    block_meas<KG_END_THREADS>(e, m, ncl, a, red);
    if (tid == 0) {
      int64_t* o = f.bst + ((size_t)(g & 1) * s.n_lat + lat) * 6;
      o[0] = ncl;
      o[1] = (int64_t)a.k1;
      o[2] = (int64_t)a.k2;
      o[3] = a.p1;
      o[4] = a.p2;
      o[5] = f.wrapw[lat];
    }
    __syncthreads();  // red is reused below
  }
  const bool fold = g >= 1 && f.fin[lat] == g - 1;
  if (fold && g - 1 >= lat_therm(s, lat)) fused_finalize_block<KG_END_THREADS>(s, f, lat, g - 1, red);
  __syncthreads();  // every thread has read the counters it updates
  if (tid == 0) {
    f.wrapw[lat] = 0u;
    if (fold) f.fin[lat] = g;
    f.cur[lat] ^= 1;
    s.sweep_ctr[lat] = g + 1;
  }
}

// ------------------------------------------------------------ run closing
// E, M of the current state (pending flips applied), when the last sweep
// still has to be folded.
template <int T, int NTH, bool BIG, int PART>
__global__ __launch_bounds__(NTH) void kf_measure(LatticeSet s, FusedScratch f) {
  __shared__ uint8_t bits[T * T];
  __shared__ uint8_t hsp[2 * T];
  __shared__ uint32_t fb_lds[fb_words<BIG>()];
  __shared__ int64_t red[NTH / 64][PARTS];
  const int lat = run_lat(s, blockIdx.y);
  const int64_t G = s.sweep_ctr[lat] - 1;
  if (!(G >= 0 && f.fin[lat] == G && G >= lat_therm(s, lat))) return;
  const uint32_t L = (uint32_t)s.L, N = (uint32_t)s.N;
  const uint32_t nt = (uint32_t)f.tiles_x, tiles = nt * nt;
  const uint32_t tile = blockIdx.x, tx = tile % nt, ty = tile / nt;
  const uint32_t x0 = tx * T, y0 = ty * T;
  const uint32_t wx = PART ? min((uint32_t)T, L - x0) : (uint32_t)T;
  const uint32_t wy = PART ? min((uint32_t)T, L - y0) : (uint32_t)T;
  const FusedView v = fused_view(s, f, lat);
  const uint8_t* jb = s.jbits + (size_t)lat * N;
  const uint32_t* fb = flipbits_of<BIG>(f, lat, fb_lds, NTH);
  __syncthreads();
  load_halo<T>(s, f, v, fb, tx, ty, wx, wy, hsp);
  for (uint32_t li = threadIdx.x; li < T * T; li += NTH) {
    if (PART && (li / T >= wx || li % T >= wy)) continue;  // absent site
    const uint32_t n = (x0 + li / T) * L + y0 + li % T;
    bits[li] = (uint8_t)(true_spin<T>(v.in[n], tile, fb) | (jb[n] & 3u) << 4);
  }
  __syncthreads();
  int32_t e = 0, m = 0;
  for (uint32_t li = threadIdx.x; li < T * T; li += NTH) {
    const uint32_t lx = li / T, ly = li % T;
    if (PART && (lx >= wx || ly >= wy)) continue;
    const uint32_t b = bits[li];
    const uint32_t sr = lx + 1u < wx ? bits[li + T] & 1u : hsp[ly];
    const uint32_t su = ly + 1u < wy ? bits[li + 1u] & 1u : hsp[T + lx];
    const uint32_t sh = ~(b ^ sr ^ (b >> 4)) & 1u, sv = ~(b ^ su ^ (b >> 5)) & 1u;
    e += 2 - 2 * (int32_t)(sh + sv);
    m += 2 * (int32_t)(b & 1u) - 1;
  }
  wave_spins(e, m, red);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t es = 0, msum = 0;
    for (int w = 0; w < NTH / 64; ++w) {
      es += red[w][0];
      msum += red[w][1];
    }
    int64_t* o = f.part_em + ((size_t)lat * tiles + tile) * 2;
    o[0] = es;
    o[1] = msum;
  }
}

// The batch's spins from the current state, pending flips applied.
template <int T, int NTH, bool BIG, int PART>
__global__ __launch_bounds__(NTH) void kf_settle(LatticeSet s, FusedScratch f) {
  __shared__ uint32_t fb_lds[fb_words<BIG>()];
  const int lat = run_lat(s, blockIdx.y);
  const uint32_t L = (uint32_t)s.L, N = (uint32_t)s.N;
  const uint32_t nt = (uint32_t)f.tiles_x;
  const uint32_t tile = blockIdx.x, tx = tile % nt, ty = tile / nt;
  const uint32_t x0 = tx * T, y0 = ty * T;
  const FusedView v = fused_view(s, f, lat);
  int8_t* s0 = s.spins + (size_t)lat * N;
  const uint32_t* fb = flipbits_of<BIG>(f, lat, fb_lds, NTH);
  const uint32_t wx = PART ? min((uint32_t)T, L - x0) : (uint32_t)T;
  const uint32_t wy = PART ? min((uint32_t)T, L - y0) : (uint32_t)T;
  __syncthreads();
  for (uint32_t li = threadIdx.x; li < T * T; li += NTH) {
    if (PART && (li / T >= wx || li % T >= wy)) continue;  // absent site
    const uint32_t n = (x0 + li / T) * L + y0 + li % T;
    s0[n] = true_spin<T>(v.in[n], tile, fb) ? (int8_t)1 : (int8_t)-1;
  }
}

// Before a run's first sweep: state plane 0 (the current one between runs)
// from the batch's spins, with no site in an edge cluster, and the couplings
// tile by tile (16 sites of one tile row per word).
__global__ void kf_open(LatticeSet s, FusedScratch f) {
  const int lat = run_lat(s, blockIdx.y);
  const size_t base = (size_t)lat * s.N;
  const int8_t* sp = s.spins + base;
  const uint8_t* jb = s.jbits + base;
  uint16_t* st = f.st + base;
  const uint32_t N = (uint32_t)s.N, L = (uint32_t)s.L, T = (uint32_t)f.tile;
  const uint32_t nt = (uint32_t)f.tiles_x, per_tile = T * T / 16u, words = nt * nt * per_tile;
  uint32_t* jt = f.jt + (size_t)lat * words;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += stride)
    st[n] = sp[n] > 0 ? (uint16_t)ST_SPIN : (uint16_t)0u;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
    const uint32_t tile = w / per_tile, lt = (w % per_tile) * 16u;
    const uint32_t x = (tile % nt) * T + lt / T, y0 = (tile / nt) * T + lt % T;
    uint32_t word = 0u;
    if (x < L)
      for (uint32_t k = 0; k < 16u && y0 + k < L; ++k)
        word |= (uint32_t)(jb[x * L + y0 + k] & 3u) << (2u * k);
    jt[w] = word;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) f.cur[lat] = 0;
}

__global__ void kf_close(LatticeSet s, FusedScratch f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n_run) return;
  const int lat = run_lat(s, i);
  const int64_t G = s.sweep_ctr[lat] - 1;
  if (G >= 0 && f.fin[lat] == G) {
    if (G >= lat_therm(s, lat)) fused_finalize(s, f, lat, G);
    f.fin[lat] = G + 1;
  }
  f.cur[lat] = 0;
}

// The slot merge: workgroup size and LDS (one word per slot, at least the
// reduction rows).  Measured on 1280-lattice scans: 512 threads at 4096 slots
// (L = 200, 256: four workgroups of 32 KB per CU; 256 threads 5 % and 1024
// threads 9 % slower at L = 256), 256 at 2304 (L = 129: 5 % ahead of 512).
int kb_threads(int nslots) { return nslots > 3072 ? 512 : 256; }
size_t kb_lds_bytes(int nslots, int threads) {
  return std::max((size_t)nslots * 8, (size_t)(threads / 64) * PARTS * 8);
}

// The tile passes of one sweep: every tile in one launch when T divides L,
// else the whole tiles and the cut ones (in workgroups of half the threads) in
// two.
template <int T, bool BIG, int PART>
void launch_ka(const LatticeSet& s, const FusedScratch& f, hipStream_t st) {
  constexpr int NTH = T == 64 ? 512 : 256;
  const int nt = f.tiles_x;
  if constexpr (PART == 0) {
    hipLaunchKernelGGL((ka_tile<T, NTH, BIG, 0>), dim3(nt * nt, s.n_run), dim3(NTH), 0, st, s, f);
  } else {
    constexpr int NTHC = NTH / 2;
    constexpr int PW = PART == 1 ? 0 : 3;  // whole tiles of an even / odd L
    if (nt > 1)
      hipLaunchKernelGGL((ka_tile<T, NTH, BIG, PW>), dim3((nt - 1) * (nt - 1), s.n_run), dim3(NTH), 0,
                         st, s, f);
    const uint32_t cap = cut_lds_sites<T>((uint32_t)(s.L - (nt - 1) * T), NTHC / 64);
    const size_t bytes = cut_lds_bytes<T, NTHC>(cap);
    hipLaunchKernelGGL((ka_tile_cut<T, NTHC, BIG, PART>), dim3(2 * nt - 1, s.n_run), dim3(NTHC),
                       bytes, st, s, f, cap);
  }
}

template <int T, int PART>
hipError_t launch_fused_t(const LatticeSet& s, const FusedScratch& f, hipStream_t st) {
  if (f.big) {
    const int nb = (f.nslots + 255) / 256;
    launch_ka<T, true, PART>(s, f, st);
    hipLaunchKernelGGL((kg_union<T>), dim3((f.nslots / 2 + 255) / 256, s.n_run), dim3(256), 0, st, s, f);
    hipLaunchKernelGGL(kg_fold, dim3(nb, s.n_run), dim3(256), 0, st, s, f);
    hipLaunchKernelGGL(kg_flip, dim3(nb, s.n_run), dim3(256), 0, st, s, f);
    hipLaunchKernelGGL(kg_sweep_end, dim3(s.n_run), dim3(KG_END_THREADS), 0, st, s, f, nb);
    return hipGetLastError();
  }
  launch_ka<T, false, PART>(s, f, st);
  if (kb_threads(f.nslots) == 512)
    hipLaunchKernelGGL((kb_merge<T, 512>), dim3(s.n_run), dim3(512), kb_lds_bytes(f.nslots, 512), st,
                       s, f);
  else
    hipLaunchKernelGGL((kb_merge<T, 256>), dim3(s.n_run), dim3(256), kb_lds_bytes(f.nslots, 256), st,
                       s, f);
  return hipGetLastError();
}

template <int T, bool BIG, int PART>
hipError_t launch_fused_close_tb(const LatticeSet& s, const FusedScratch& f, hipStream_t st) {
  constexpr int NTH = T == 64 ? 512 : 256;
  const int tiles = f.tiles_x * f.tiles_x;
  hipLaunchKernelGGL((kf_measure<T, NTH, BIG, PART>), dim3(tiles, s.n_run), dim3(NTH), 0, st, s, f);
  hipLaunchKernelGGL((kf_settle<T, NTH, BIG, PART>), dim3(tiles, s.n_run), dim3(NTH), 0, st, s, f);
  hipLaunchKernelGGL(kf_close, dim3((s.n_run + 63) / 64), dim3(64), 0, st, s, f);
  return hipGetLastError();
}

template <int T, int PART>
hipError_t launch_fused_close_t(const LatticeSet& s, const FusedScratch& f, hipStream_t st) {
  return f.big ? launch_fused_close_tb<T, true, PART>(s, f, st)
               : launch_fused_close_tb<T, false, PART>(s, f, st);
}

int fused_part(const LatticeSet& s, const FusedScratch& f) {
  return s.L % f.tile == 0 ? 0 : (s.L & 1) ? 2 : 1;
}

}  // namespace

// Tiles of 64 for every L >= 64, cut where 64 does not divide L; measured on
// 1280-lattice scans (site-updates/s, tiles of 64 against 32): L = 96 9.38e10
// against 9.37e10, L = 129 7.61e10 / 7.46e10, L = 160 1.048e11 / 1.038e11,
// L = 200 1.136e11 / 9.61e10, L = 224 1.209e11 / 1.056e11.  Below 32 the
// lattice is one cut tile of 32, its own neighbour across both edges.
int fused_tile_size(int L, int pref) {
  if (L < 3) return 0;
  if (pref == 32 || (pref == 64 && L >= 64)) return pref;  // cut tiles when T does not divide L
  return L >= 64 ? 64 : 32;
}

int fused_tiles_per_side(int L, int T) { return T > 0 ? (L + T - 1) / T : 0; }

bool fused_merge_in_hbm(int L, int T) {
  const size_t nt = (size_t)fused_tiles_per_side(L, T);
  return T > 0 && nt * nt * 4 * T > RBIM_FUSED_MAX_SLOTS;
}

int fused_kernels_per_sweep(const LatticeSet& s, const FusedScratch& f) {
  return (f.big ? 5 : 2) + (fused_part(s, f) != 0 ? 1 : 0);
}

hipError_t launch_fused_open(const LatticeSet& s, const FusedScratch& f, hipStream_t st) {
  if (s.n_run <= 0) return hipSuccess;
  const int blocks = s.N < 64 * 256 ? (s.N + 255) / 256 : 64;
  hipLaunchKernelGGL(kf_open, dim3(blocks, s.n_run), dim3(256), 0, st, s, f);
  return hipGetLastError();
}

hipError_t launch_fused_sweep(const LatticeSet& s, const FusedScratch& f, hipStream_t st) {
  const int part = fused_part(s, f);
  if (f.tile == 64)
    return part == 1 ? launch_fused_t<64, 1>(s, f, st)
         : part == 2 ? launch_fused_t<64, 2>(s, f, st)
                     : launch_fused_t<64, 0>(s, f, st);
  return part == 1 ? launch_fused_t<32, 1>(s, f, st)
       : part == 2 ? launch_fused_t<32, 2>(s, f, st)
                   : launch_fused_t<32, 0>(s, f, st);
}

hipError_t launch_fused_close(const LatticeSet& s, const FusedScratch& f, hipStream_t st) {
  const int part = fused_part(s, f);
  if (f.tile == 64)
    return part == 1 ? launch_fused_close_t<64, 1>(s, f, st)
         : part == 2 ? launch_fused_close_t<64, 2>(s, f, st)
                     : launch_fused_close_t<64, 0>(s, f, st);
  return part == 1 ? launch_fused_close_t<32, 1>(s, f, st)
       : part == 2 ? launch_fused_close_t<32, 2>(s, f, st)
                   : launch_fused_close_t<32, 0>(s, f, st);
}
