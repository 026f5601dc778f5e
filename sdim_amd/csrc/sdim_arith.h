// Device arithmetic of the engine.
//
// Odd d: four generators per 32-bit word, one uint8 lane each, every lane
// holding a value in [0, d).  Lane-wise mod-d arithmetic is done SWAR style:
// sums of two reduced lanes stay below 2d <= 22, so a single biased add
// exposes "lane >= d" in bit 7 of each byte without carries between lanes.
// Multiplication of a word by a wave-uniform scalar c is one v_perm_b32 byte
// lookup into an 8-byte table of c*v mod d (two lookups + a select for d = 11).
//
// d = 2: 64 generators per 64-bit word, one bit each; phases mod 4 as two
// bit-planes.  Everything is AND/XOR.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdim {

#define SDIM_DEV __device__ __forceinline__

constexpr uint32_t ONES = 0x01010101u;

// bytes in [0, 2d) -> bytes mod d
template <int D> SDIM_DEV uint32_t red1(uint32_t s) {
  uint32_t t = s + (0x80u - D) * ONES;
  uint32_t m = (t >> 7) & ONES;
  return s - m * (uint32_t)D;
}
template <int D> SDIM_DEV uint32_t addm(uint32_t a, uint32_t b) { return red1<D>(a + b); }
template <int D> SDIM_DEV uint32_t subm(uint32_t a, uint32_t b) { return red1<D>(a + (uint32_t)D * ONES - b); }
template <int D> SDIM_DEV uint32_t negm(uint32_t a) { return red1<D>((uint32_t)D * ONES - a); }

// one conditional subtraction of K*D per lane (lanes < 128)
template <int D, int K> SDIM_DEV uint32_t red_step(uint32_t v) {
  if constexpr (K < D) {
    uint32_t t = v + (0x80u - (uint32_t)(K * D)) * ONES;
    uint32_t m = (t >> 7) & ONES;
    v -= m * (uint32_t)(K * D);
  }
  return v;
}
// lanes in [0, d*d) -> lanes mod d
template <int D> SDIM_DEV uint32_t redsq(uint32_t v) {
  v = red_step<D, 8>(v);
  v = red_step<D, 4>(v);
  v = red_step<D, 2>(v);
  v = red_step<D, 1>(v);
  return v;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// lane-wise product of two reduced words, mod d
template <int D> SDIM_DEV uint32_t mulv(uint32_t a, uint32_t b) {
  u16x2 ae = __builtin_bit_cast(u16x2, a & 0x00FF00FFu);
  u16x2 be = __builtin_bit_cast(u16x2, b & 0x00FF00FFu);
  u16x2 ao = __builtin_bit_cast(u16x2, (a >> 8) & 0x00FF00FFu);
  u16x2 bo = __builtin_bit_cast(u16x2, (b >> 8) & 0x00FF00FFu);
  uint32_t pe = __builtin_bit_cast(uint32_t, (u16x2)(ae * be));
  uint32_t po = __builtin_bit_cast(uint32_t, (u16x2)(ao * bo));
  return redsq<D>(pe | (po << 8));
}

// Multiplication table of a scalar c: lo = lanes v*c for v = 0..7, hi = v = 8..15.
struct MulTab { uint32_t l0, l1, h0, h1; };

template <int D> SDIM_DEV MulTab make_tab(uint32_t c) {
  MulTab t{0, 0, 0, 0};
  #pragma unroll
  for (int v = 0; v < 16; v++) {
    uint32_t e = (v < D) ? (uint32_t)((v * c) % D) : 0u;
    if (v < 4) t.l0 |= e << (8 * v);
    else if (v < 8) t.l1 |= e << (8 * (v - 4));
    else if (v < 12) t.h0 |= e << (8 * (v - 8));
    else t.h1 |= e << (8 * (v - 12));
  }
  return t;
}

// lane-wise c * x mod d through the table of c (x lanes in [0, d))
template <int D> SDIM_DEV uint32_t mult(const MulTab& t, uint32_t x) {
  if constexpr (D <= 8) {
    return __builtin_amdgcn_perm(t.l1, t.l0, x);
  } else {
    uint32_t i = x & 0x07070707u;
    uint32_t lo = __builtin_amdgcn_perm(t.l1, t.l0, i);
    uint32_t hi = __builtin_amdgcn_perm(t.h1, t.h0, i);
    uint32_t sel = 0x03020100u | ((x & 0x08080808u) >> 1);
    return __builtin_amdgcn_perm(hi, lo, sel);
  }
}

// table of the P-gate phase increment v(v-1)/2 mod d (tableau_optimized.py:74)
template <int D> SDIM_DEV MulTab make_half_tab() {
  MulTab t{0, 0, 0, 0};
  #pragma unroll
  for (int v = 0; v < 16; v++) {
    uint32_t e = (v < D) ? (uint32_t)(((v * (v - 1)) / 2) % D) : 0u;
    if (v < 4) t.l0 |= e << (8 * v);
    else if (v < 8) t.l1 |= e << (8 * (v - 4));
    else if (v < 12) t.h0 |= e << (8 * (v - 8));
    else t.h1 |= e << (8 * (v - 12));
  }
  return t;
}

SDIM_DEV uint32_t byte_of(uint32_t w, int k) { return (w >> (8 * k)) & 0xFFu; }

// ---------------------------------------------------------------- Philox4x64-10
struct U64x4 { uint64_t v[4]; };

// Full 64 x 64 -> 128-bit product from four 32 x 32 -> 64 partial products, each a
// single multiply-add (v_mad_u64_u32).  No sum overflows: a 32-bit product is at
// most 2^64 - 2^33 + 1 and at most two 32-bit carries are added to it.
__host__ __device__ __forceinline__ void mul128(uint64_t a, uint64_t m, uint64_t& lo, uint64_t& hi) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), m0 = (uint32_t)m, m1 = (uint32_t)(m >> 32);
  const uint64_t p00 = (uint64_t)a0 * m0;
  const uint64_t p01 = (uint64_t)a0 * m1 + (p00 >> 32);
  const uint64_t p10 = (uint64_t)a1 * m0 + (uint32_t)p01;
  hi = (uint64_t)a1 * m1 + (p01 >> 32) + (p10 >> 32);
  lo = (p10 << 32) | (uint32_t)p00;
}

SDIM_DEV U64x4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1) {
  #pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B97F4A7C15ull; k1 += 0xBB67AE8584CAA73Bull; }
    uint64_t lo0, hi0, lo1, hi1;
    mul128(c0, 0xD2E7470EE14C6C93ull, lo0, hi0);
    mul128(c2, 0xCA5A826395121157ull, lo1, hi1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  U64x4 o; o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
  return o;
}

// Draw keyed by (seed, stream | op, shot); identical to oracle/philox.py's `block`.
SDIM_DEV U64x4 draw(uint64_t seed, uint64_t stream, uint64_t op, uint64_t shot) {
  return philox4x64_10(op, shot, 0, 0, seed, stream);
}

// Counter word 2 of the frame sampler's draws (initial and redrawn z components);
// the tableau path's draws use 0.
constexpr uint64_t FRAME_DOMAIN = 1;

SDIM_DEV uint32_t outcome_of(const U64x4& o, uint32_t d) { return (uint32_t)__umul64hi(o.v[0], (uint64_t)d); }

// (a, b) exponents of an N1 error, 0 when the coin says identity (program.py:484-503 semantics)
SDIM_DEV uint32_t noise_ab(const U64x4& o, uint32_t d, uint32_t channel, uint64_t keep53) {
  if ((o.v[0] >> 11) < keep53) return 0u;
  uint32_t a = 0, b = 0;
  if (channel == 0) {
    uint32_t r = 1u + (uint32_t)__umul64hi(o.v[1], (uint64_t)(d * d - 1));
    a = r % d; b = r / d;
  } else if (channel == 1) {
    a = 1u + (uint32_t)__umul64hi(o.v[1], (uint64_t)(d - 1));
  } else {
    b = 1u + (uint32_t)__umul64hi(o.v[1], (uint64_t)(d - 1));
  }
  return a | (b << 4) | 0x100u;  // bit 8 marks "applied"
}

}  // namespace sdim
