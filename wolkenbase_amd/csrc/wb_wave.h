// Wave64 sorting network for 32-bit keys held in registers: element
// e = 64*s + lane of a wave lives in slot s of that lane, K slots per lane.
//
// The network is the bitonic sort in its "mirror" form: stage G (G = 2, 4,
// ..., 64K) first compares every element with its mirror image inside its
// aligned group of G elements (e against e ^ (G-1)), then runs the
// half-cleaner steps e against e ^ J for J = G/4, ..., 1; the lower element
// of every pair keeps the smaller key, so each stage leaves every group of G
// sorted ascending and no step needs a direction bit.  The mirror form lets
// most exchanges stay inside a row of 16 lanes, where they are DPP moves on
// the VALU instead of round trips through the LDS crossbar:
//   mirror in 2, 4, 8, 16 lanes: quad_perm 1,0,3,2 / quad_perm 3,2,1,0 /
//                                row_half_mirror / row_mirror
//   xor 1, 2, 8:                 quad_perm 1,0,3,2 / quad_perm 2,3,0,1 /
//                                row_ror 8
//   xor 4:                       row_half_mirror then quad_perm 3,2,1,0
// Mirror in 32 and 64 lanes and xor 16 and 32 go through ds_bpermute
// (three of the 21 exchanges of a 64-element sort).  Partners 64 or more
// elements apart are register exchanges between slots.
//
// Sorting stops after stage P (a wave-uniform power of two): the first P
// elements are then in ascending order.  Slots at or past P/64 are left as
// they are, so they must not hold any of the first P elements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wbw {

template <int CTRL>
__device__ __attribute__((always_inline)) inline uint32_t dpp_u(uint32_t v)
{
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}

// the value of lane (lane ^ (G - 1)): the mirror image inside aligned groups
// of G lanes
template <int G>
__device__ __attribute__((always_inline)) inline uint32_t lane_mirror(uint32_t v)
{
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group of 2..64 lanes");
  if constexpr (G == 2) return dpp_u<0xb1>(v);
  else if constexpr (G == 4) return dpp_u<0x1b>(v);
  else if constexpr (G == 8) return dpp_u<0x141>(v);
  else if constexpr (G == 16) return dpp_u<0x140>(v);
  else return (uint32_t)__shfl_xor((int)v, G - 1, 64);
}

// the value of lane (lane ^ J)
template <int J>
__device__ __attribute__((always_inline)) inline uint32_t lane_xor(uint32_t v)
{
  static_assert(J >= 1 && J <= 32 && (J & (J - 1)) == 0, "distance 1..32");
  if constexpr (J == 1) return dpp_u<0xb1>(v);
  else if constexpr (J == 2) return dpp_u<0x4e>(v);
  else if constexpr (J == 4) return dpp_u<0x1b>(dpp_u<0x141>(v));
  else if constexpr (J == 8) return dpp_u<0x128>(v);
  else return (uint32_t)__shfl_xor((int)v, J, 64);
}

// compare-exchange from one side: this element has key k (payload v), its
// partner ok (ov); the lower element of the pair keeps the smaller key
template <bool PAY>
__device__ __attribute__((always_inline)) inline void cx_side(uint32_t &k, uint32_t &v,
                                                              uint32_t ok, uint32_t ov, bool lower)
{
  const bool take = lower ? ok < k : ok > k;
  k = take ? ok : k;
  if constexpr (PAY) v = take ? ov : v;
}

// compare-exchange of two slots of the same lane: a keeps the smaller key
template <bool PAY>
__device__ __attribute__((always_inline)) inline void cx_slots(uint32_t &ka, uint32_t &va,
                                                               uint32_t &kb, uint32_t &vb)
{
  if (kb < ka) {
    const uint32_t tk = ka;
    ka = kb;
    kb = tk;
    if constexpr (PAY) {
      const uint32_t tv = va;
      va = vb;
      vb = tv;
    }
  }
}

// half-cleaner steps J, J/2, ..., 1: element e against e ^ J
template <int K, int J, bool PAY>
__device__ __attribute__((always_inline)) inline void half_clean(uint32_t (&c)[K], uint32_t (&v)[K],
                                                                 uint32_t live)
{
  if constexpr (J >= 64) {
#pragma unroll
    for (int s = 0; s < K; s++) {
      const int s2 = s ^ (J / 64);
      if (s2 > s && (uint32_t)s < live) cx_slots<PAY>(c[s], v[s], c[s2], v[s2]);
    }
  } else {
    const bool lower = (__lane_id() & J) == 0;
#pragma unroll
    for (int s = 0; s < K; s++) {
      if ((uint32_t)s >= live) continue;
      const uint32_t ok = lane_xor<J>(c[s]);
      uint32_t ov = 0;
      if constexpr (PAY) ov = lane_xor<J>(v[s]);
      cx_side<PAY>(c[s], v[s], ok, ov, lower);
    }
  }
  if constexpr (J > 1) half_clean<K, J / 2, PAY>(c, v, live);
}

// stages G, 2G, ..., up to P
template <int K, int G, bool PAY>
__device__ __attribute__((always_inline)) inline void sort_stages(uint32_t (&c)[K], uint32_t (&v)[K],
                                                                  uint32_t P, uint32_t live)
{
  if constexpr (G <= 64 * K) {
    if ((uint32_t)G > P) return;
    if constexpr (G <= 64) {
      const bool lower = (__lane_id() & (G / 2)) == 0;
#pragma unroll
      for (int s = 0; s < K; s++) {
        if ((uint32_t)s >= live) continue;
        const uint32_t ok = lane_mirror<G>(c[s]);
        uint32_t ov = 0;
        if constexpr (PAY) ov = lane_mirror<G>(v[s]);
        cx_side<PAY>(c[s], v[s], ok, ov, lower);
      }
    } else {
      // e = 64s + lane against 64(s ^ (G/64 - 1)) + (63 - lane): the upper
      // slot is mirrored across the wave, exchanged lane by lane, and
      // mirrored back
#pragma unroll
      for (int s = 0; s < K; s++) {
        const int s2 = s ^ (G / 64 - 1);
        if (s2 <= s || (uint32_t)s >= live) continue;
        uint32_t ok = lane_mirror<64>(c[s2]), ov = 0;
        if constexpr (PAY) ov = lane_mirror<64>(v[s2]);
        cx_slots<PAY>(c[s], v[s], ok, ov);
        c[s2] = lane_mirror<64>(ok);
        if constexpr (PAY) v[s2] = lane_mirror<64>(ov);
      }
    }
    if constexpr (G >= 4) half_clean<K, G / 4, PAY>(c, v, live);
    sort_stages<K, 2 * G, PAY>(c, v, P, live);
  }
}

// Sorts the first P elements (P a power of two, 1 <= P <= 64K, the same in
// every lane) by key, carrying the payload when PAY.
template <int K, bool PAY>
__device__ __attribute__((always_inline)) inline void wave_sort(uint32_t (&c)[K], uint32_t (&v)[K],
                                                                uint32_t P)
{
  sort_stages<K, 2, PAY>(c, v, P, (P + 63) / 64);
}

// the smallest power of two >= m (1 for m <= 1)
__device__ __attribute__((always_inline)) inline uint32_t pow2_ceil(uint32_t m)
{
  return m <= 1 ? 1u : 1u << (32 - __builtin_clz(m - 1));
}

}  // namespace wbw
