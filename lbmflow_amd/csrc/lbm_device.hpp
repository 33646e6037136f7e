// lbm_device.hpp — device helpers shared by the HIP translation units:
// compile-time loops, population storage codecs (fp32, and fp16/bf16 scaled
// by 1/w_q), wrapped neighbour offsets, the c_q . v dot product and the
// bit-exact moment sum of moments_row (kernels.rs:627-668).
#pragma once
#include <type_traits>

#include "lbm_internal.hpp"
#include "lbm_lattice.hpp"

namespace lbm {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
template <int B, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// Sums N doubles over a 256-thread block (four 64-lane waves).  Each wave
// reduces its lanes with a fixed xor butterfly through cross-lane shuffles
// (no LDS, no atomics), then thread 0 adds the four wave sums in wave order
// through 4N doubles of LDS.  The order of every addition is fixed by the
// launch shape, so results repeat bit for bit.  On return thread 0 holds the
// block sums in v; other threads hold their wave's sums.
template <int N>
__device__ __forceinline__ void block_sum256(double* v, double* lds) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[k] = x;
  }
  const int w = (int)(threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) lds[w * N + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = ((lds[k] + lds[N + k]) + lds[2 * N + k]) + lds[3 * N + k];
}

struct bf16s {
  uint16_t v;
};

template <typename S>
struct Store;
template <>
struct Store<float> {
  static constexpr bool kScaled = false;
  __device__ __forceinline__ static float ld(const float* p, uint32_t i) { return p[i]; }
  __device__ __forceinline__ static void st(float* p, uint32_t i, float v) { p[i] = v; }
};
template <>
struct Store<_Float16> {
  static constexpr bool kScaled = true;
  __device__ __forceinline__ static float ld(const _Float16* p, uint32_t i) { return (float)p[i]; }
  __device__ __forceinline__ static void st(_Float16* p, uint32_t i, float v) { p[i] = (_Float16)v; }
};
template <>
struct Store<bf16s> {
  static constexpr bool kScaled = true;
  __device__ __forceinline__ static float ld(const bf16s* p, uint32_t i) {
    return __uint_as_float((uint32_t)p[i].v << 16);
  }
  __device__ __forceinline__ static void st(bf16s* p, uint32_t i, float v) {
    uint32_t b = __float_as_uint(v);
    b += 0x7FFFu + ((b >> 16) & 1u);  // round to nearest even (finite values)
    p[i].v = (uint16_t)(b >> 16);
  }
};

// population value <-> stored value
template <typename S, int Q, int q>
__device__ __forceinline__ float pop_ld(const S* plane, uint32_t i) {
  float v = Store<S>::ld(plane, i);
  if constexpr (Store<S>::kScaled) v = v * lat_wf(Q, q);
  return v;
}
template <typename S, int Q, int q>
__device__ __forceinline__ void pop_st(S* plane, uint32_t i, float v) {
  if constexpr (Store<S>::kScaled) v = v * lat_inv_w(Q, q);
  Store<S>::st(plane, i, v);
}

// Block-id remap: hardware hands consecutive workgroups to the 8 XCDs
// round-robin; give each XCD a contiguous run of tiles so neighbouring rows
// share its L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
  constexpr uint32_t X = 8;
  uint32_t per = n / X, rem = n % X;
  uint32_t xcd = b % X, local = b / X;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Wrapped neighbour offsets of one cell (index shift for c = -1 and +1 on
// each axis).  Non-periodic axes never need the wrap for fluid cells
// (lbm_set_masks rejects fluid on a non-periodic face), so wrapping
// unconditionally keeps every address in bounds.
struct Nbr {
  int32_t xm, xp, ym, yp, zm, zp;
  __device__ __forceinline__ Nbr(const Dom& d, uint32_t x, uint32_t y, uint32_t z) {
    const int32_t nx = (int32_t)d.nx, nxy = (int32_t)(d.nx * d.ny);
    xm = x == 0 ? nx - 1 : -1;
    xp = x == d.nx - 1 ? -(nx - 1) : 1;
    ym = (y == 0 ? (int32_t)d.ny - 1 : -1) * nx;
    yp = (y == d.ny - 1 ? -((int32_t)d.ny - 1) : 1) * nx;
    zm = (z == 0 ? (int32_t)d.nz - 1 : -1) * nxy;
    zp = (z == d.nz - 1 ? -((int32_t)d.nz - 1) : 1) * nxy;
  }
  template <int cx, int cy, int cz>
  __device__ __forceinline__ int32_t off() const {
    int32_t o = 0;
    if constexpr (cx < 0) o += xm;
    if constexpr (cx > 0) o += xp;
    if constexpr (cy < 0) o += ym;
    if constexpr (cy > 0) o += yp;
    if constexpr (cz < 0) o += zm;
    if constexpr (cz > 0) o += zp;
    return o;
  }
};

// c_q . v with the terms of zero coefficient dropped (they only add signed
// zeros in the reference's seeded sum, collision.rs:53-65).
template <int q>
__device__ __forceinline__ float cdot(float v0, float v1, float v2) {
  constexpr int c0 = lat_c(q, 0), c1 = lat_c(q, 1), c2 = lat_c(q, 2);
  float acc = 0.0f;
  bool first = true;
  if constexpr (c0 != 0) {
    acc = c0 > 0 ? v0 : -v0;
    first = false;
  }
  if constexpr (c1 != 0) {
    if (first) {
      acc = c1 > 0 ? v1 : -v1;
      first = false;
    } else {
      acc = c1 > 0 ? acc + v1 : acc - v1;
    }
  }
  if constexpr (c2 != 0) {
    if (first) {
      acc = c2 > 0 ? v2 : -v2;
    } else {
      acc = c2 > 0 ? acc + v2 : acc - v2;
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// moments: moments_row (kernels.rs:627-668), bit-exact in fp32
// ---------------------------------------------------------------------------
template <int Q>
__device__ __forceinline__ void moments(const float* fv, const float Fm[3], bool fon, float& r, float& ux, float& uy,
                                        float& uz) {
  float dr = 0.0f;
  float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    dr = dr + fv[q];
    if constexpr (lat_c(q, 0) > 0) m0 = m0 + fv[q];
    if constexpr (lat_c(q, 0) < 0) m0 = m0 - fv[q];
    if constexpr (lat_c(q, 1) > 0) m1 = m1 + fv[q];
    if constexpr (lat_c(q, 1) < 0) m1 = m1 - fv[q];
    if constexpr (lat_c(q, 2) > 0) m2 = m2 + fv[q];
    if constexpr (lat_c(q, 2) < 0) m2 = m2 - fv[q];
  });
  r = 1.0f + dr;
  const float inv = 1.0f / r;
  if (fon) {
    const float half = 0.5f;
    ux = (m0 + half * Fm[0]) * inv;
    uy = (m1 + half * Fm[1]) * inv;
    uz = (m2 + half * Fm[2]) * inv;
  } else {
    ux = m0 * inv;
    uy = m1 * inv;
    uz = m2 * inv;
  }
}

// ---------------------------------------------------------------------------
// read a population of cell i in the reference sense (post-stream f_q)
// ---------------------------------------------------------------------------
// Even state: f_q(i) sits in the cell's own slot q.  Odd state (after an
// even step): in the upstream cell's slot opp(q), except across a wall link
// (cell i - c_q solid), where the bounced population stays in the cell's own
// slot q.  `m` is the cell's mask word.
template <int Q, typename S, int q>
__device__ __forceinline__ float read_pop(const S* base, size_t ps, const Nbr& nb, uint32_t i, uint32_t m,
                                          bool odd) {
  constexpr int o = lat_opp(q);
  // one load from a selected address: with 16-bit storage a load converted
  // inside either arm of a branch makes the caller's loads wait one by one
  const bool own = !odd || ((m >> q) & 1u);
  const uint32_t s = own ? i : i + (uint32_t)nb.template off<-lat_c(q, 0), -lat_c(q, 1), -lat_c(q, 2)>();
  return pop_ld<S, Q, q>(base + (size_t)(own ? q : o) * ps, s);
}

// write population q of cell i where read_pop finds it (the inverse map)
template <int Q, typename S, int q>
__device__ __forceinline__ void write_pop(S* base, size_t ps, const Nbr& nb, uint32_t i, uint32_t m, bool odd,
                                          float v) {
  constexpr int o = lat_opp(q);
  if (!odd || ((m >> q) & 1u)) {
    pop_st<S, Q, q>(base + (size_t)q * ps, i, v);
    return;
  }
  const uint32_t s = i + (uint32_t)nb.template off<-lat_c(q, 0), -lat_c(q, 1), -lat_c(q, 2)>();
  pop_st<S, Q, q>(base + (size_t)o * ps, s, v);
}

__device__ __forceinline__ void cell_xyz(const Dom& d, uint32_t i, uint32_t& x, uint32_t& y, uint32_t& z) {
  x = i % d.nx;
  const uint32_t t = i / d.nx;
  y = t % d.ny;
  z = t / d.ny;
}

}  // namespace lbm
