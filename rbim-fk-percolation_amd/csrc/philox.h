// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), written for gfx950
// device code and for host code from the published round function:
//
//   (hi0, lo0) = M0 * c0,  (hi1, lo1) = M1 * c2          (32x32 -> 64 products)
//   c' = (hi1 ^ c1 ^ k0,  lo1,  hi0 ^ c3 ^ k1,  lo0)
//   k' = (k0 + W0, k1 + W1)                               (key schedule, Weyl)
//
// applied ten times with the key bumped between rounds.
//
// Stream layout used by the sweep engine (fixed; every path that draws uses it,
// so results never depend on launch shape, tile size or batch composition):
//   key     = (seed_lo, seed_hi) of the 64-bit engine seed of a lattice
//   sweep   : counter = (site >> 1, sweep, lattice_id, RBIM_TAG_SWEEP)
//             out.x / out.y -> horizontal / vertical bond of the even site,
//             out.z / out.w -> horizontal / vertical bond of the odd site.
//             A bond draw is (word >> 1), compared with a 31-bit threshold;
//             the low bit of out.x / out.z is the cluster-flip bit of the
//             even / odd site (used when that site is a cluster's min site).
//   hot init: counter = (site >> 7, 0, lattice_id, RBIM_TAG_INIT)
//             spin(site) = +1 iff bit (site & 31) of word ((site >> 5) & 3).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RBIM_HD __host__ __device__ __forceinline__
#else
#define RBIM_HD static inline
#endif

#define RBIM_TAG_SWEEP 0u
#define RBIM_TAG_INIT 1u

struct rbim_u32x4 {
  uint32_t x, y, z, w;
};

RBIM_HD void rbim_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

// a ^ b ^ c; one v_bitop3_b32 (truth table 0x96) in device code
RBIM_HD uint32_t rbim_xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

RBIM_HD rbim_u32x4 rbim_philox4x32_10(rbim_u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    rbim_mulhilo32(M0, c.x, &hi0, &lo0);
    rbim_mulhilo32(M1, c.z, &hi1, &lo1);
    rbim_u32x4 n;
    // With a counter whose words 1-3 are uniform over a wavefront (sweep,
    // lattice id, tag) two of the three operands are uniform up to round 1
    // (n.x) and round 2 (n.z): there plain xors pair them on the scalar unit,
    // leaving one vector xor, and keep round 0's n.x (and the product it
    // feeds) scalar.
    n.x = r < 2 ? hi1 ^ c.y ^ k0 : rbim_xor3(hi1, c.y, k0);
    n.y = lo1;
    n.z = r < 3 ? hi0 ^ c.w ^ k1 : rbim_xor3(hi0, c.w, k1);
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Draws of one site pair for one sweep of one lattice.
RBIM_HD rbim_u32x4 rbim_sweep_draws(uint64_t seed, uint32_t lattice_id, uint32_t sweep,
                                    uint32_t pair) {
  rbim_u32x4 c;
  c.x = pair;
  c.y = sweep;
  c.z = lattice_id;
  c.w = RBIM_TAG_SWEEP;
  return rbim_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Hot-start spin of one site: +1 or -1.
RBIM_HD int rbim_init_spin(uint64_t seed, uint32_t lattice_id, uint32_t site) {
  rbim_u32x4 c;
  c.x = site >> 7;
  c.y = 0u;
  c.z = lattice_id;
  c.w = RBIM_TAG_INIT;
  const rbim_u32x4 r = rbim_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint32_t sel = (site >> 5) & 3u;
  const uint32_t word = sel == 0 ? r.x : sel == 1 ? r.y : sel == 2 ? r.z : r.w;
  return ((word >> (site & 31u)) & 1u) ? 1 : -1;
}
