// lbm_les.hip — WALE large-eddy closure (Nicoud & Ducros 1999) as a
// producer of the per-cell omega_plus field the collide-stream kernels read.
//
// Two cell-parallel passes per update (WaleLes::update, les.rs:120-211):
//   k_wale_local  reads the cell's Q populations once and writes the cached
//     velocity of the current state (moments_row) and the off-diagonal
//     strain from the non-equilibrium stress (strain_rate_at,
//     solver.rs:5393-5447) into a 6 x N work buffer;
//   k_wale        takes the velocity gradient g = du_i/dx_j: diagonal from
//     finite differences of the cached velocity, off-diagonal S + the
//     antisymmetric part of the finite differences
//     (gather_velocity_gradient, solver.rs:5490-5581), then
//     nu_t = (Cw Delta)^2 (Sd:Sd)^(3/2) / ((S:S)^(5/2) + (Sd:Sd)^(5/4)),
//     clipped to nu_t_max when asked, omega = 1/(3 (nu0 + nu_t) + 1/2).
//     The second pass marches (x, y) columns along z with the z - 1, z,
//     z + 1 velocities in registers, so each velocity is read from memory
//     once; x and y neighbours come from the tile's own rows in L2.
// Bytes per fluid cell: 4Q + 4 (mask) + 24 in the first pass; 4 + 24 + the
// neighbours' velocities (mostly from L2) + 4 in the second, against two
// full population reads for a moment pass followed by a strain pass.
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

#pragma clang fp contract(off)

namespace lbm {

struct WaleArgs {
  Dom d;
  const void* f;
  const uint32_t* mask;
  const float* wall_u;  // N x 3 or null (velocity of each solid cell's wall)
  const float* F_cur;   // N x 3 or null (current force field, nonzero only on
                        // the force list, whose fluid cells carry kMaskForced)
  const float* F_mom;   // N x 3 or null (the field the cached moments were taken with)
  HydroParams p;        // the uniform force
  int odd;
  float tau_eff;      // 1/omega_p of the global rate (stress scale)
  float cw_delta_sq;  // (Cw Delta)^2
  float nu0;          // laminar viscosity
  float nut_max;      // < 0: no clip
  float* vel;         // 3 x N: ux, uy, uz of the current state (cached moments)
  float* strain;      // 3 x N: Sxy, Sxz, Syz from the non-equilibrium stress
  float* omega;       // N: omega_plus out
  float* nut;         // N or null: nu_t out (after the clip)
  uint32_t* stats;    // null or [clipped cells, bits of the largest raw nu_t]
  uint32_t zlen;      // cells each thread of the second pass marches along z
};

// Population loads marked non-temporal: each population is read once, so
// the L2 is left to the velocity planes the neighbour differences reuse.
template <typename S>
__device__ __forceinline__ float ld_nt(const S* p, uint32_t i);
template <>
__device__ __forceinline__ float ld_nt<float>(const float* p, uint32_t i) {
  return __builtin_nontemporal_load(p + i);
}
template <>
__device__ __forceinline__ float ld_nt<_Float16>(const _Float16* p, uint32_t i) {
  return (float)__builtin_nontemporal_load(p + i);
}
template <>
__device__ __forceinline__ float ld_nt<bf16s>(const bf16s* p, uint32_t i) {
  const uint16_t v = __builtin_nontemporal_load((const uint16_t*)p + i);
  return __uint_as_float((uint32_t)v << 16);
}
template <int Q, typename S, int q>
__device__ __forceinline__ float read_pop_nt(const S* base, size_t ps, const Nbr& nb, uint32_t i, uint32_t m,
                                             bool odd) {
  constexpr int o = lat_opp(q);
  const S* plane;
  uint32_t j;
  if (!odd || ((m >> q) & 1u)) {
    plane = base + (size_t)q * ps;
    j = i;
  } else {
    plane = base + (size_t)o * ps;
    j = i + (uint32_t)nb.template off<-lat_c(q, 0), -lat_c(q, 1), -lat_c(q, 2)>();
  }
  float v = ld_nt<S>(plane, j);
  if constexpr (Store<S>::kScaled) v = v * lat_wf(Q, q);
  return v;
}

// Pass 1, per cell: the cached moments (moments_row with the force the state
// was produced under, as the moment readback computes them) and the
// off-diagonal strain from the same populations, read once.
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_wale_local(WaleArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.d.N) return;
  const size_t N = a.d.N;
  const uint32_t m = a.mask[i];
  if (m & kMaskSolid) {
    a.vel[i] = a.vel[N + i] = a.vel[2 * N + i] = 0.0f;
    a.strain[i] = a.strain[N + i] = a.strain[2 * N + i] = 0.0f;
    return;
  }
  uint32_t x, y, z;
  cell_xyz(a.d, i, x, y, z);
  const Nbr nb(a.d, x, y, z);
  float fv[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    fv[q] = read_pop_nt<Q, S, q>((const S*)a.f, a.d.ps, nb, i, m, a.odd != 0);
  });
  float Fm[3] = {a.p.F[0], a.p.F[1], a.p.F[2]};
  bool fon = a.p.force_on != 0;
  if (a.F_mom && (m & kMaskForced)) {
    for (int c = 0; c < 3; ++c) Fm[c] = a.p.F[c] + a.F_mom[3 * (size_t)i + c];
    fon = true;
  }
  float r, u0, u1, u2;
  moments<Q>(fv, Fm, fon, r, u0, u1, u2);
  a.vel[i] = u0;
  a.vel[N + i] = u1;
  a.vel[2 * N + i] = u2;
  // Pi_neq = sum_q c c (f_q - feq_q), deviation-form equilibrium
  // (kernels.rs:103-121); terms with a zero factor add exact zeros in the
  // reference's sums and are dropped.
  const float usq = u0 * u0 + u1 * u1 + u2 * u2;
  const float drho = r - 1.0f;
  float pxy = 0.0f, pxz = 0.0f, pyz = 0.0f;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = lat_c(q, 0), cy = lat_c(q, 1), cz = lat_c(q, 2);
    const float cu = cdot<q>(u0, u1, u2);
    const float feq = lat_wf(Q, q) * (drho + r * (3.0f * cu + 4.5f * cu * cu - 1.5f * usq));
    const float fn = fv[q] - feq;
    if constexpr (cx * cy == 1) pxy = pxy + fn;
    if constexpr (cx * cy == -1) pxy = pxy - fn;
    if constexpr (cx * cz == 1) pxz = pxz + fn;
    if constexpr (cx * cz == -1) pxz = pxz - fn;
    if constexpr (cy * cz == 1) pyz = pyz + fn;
    if constexpr (cy * cz == -1) pyz = pyz - fn;
  });
  // Guo-force correction of the stress, + (u_a F_b + u_b F_a)/2 off the
  // diagonal, F = uniform force + current field
  float F0 = a.p.F[0], F1 = a.p.F[1], F2 = a.p.F[2];
  if (a.F_cur && (m & kMaskForced)) {  // the field is zero off the force list
    F0 = F0 + a.F_cur[3 * (size_t)i + 0];
    F1 = F1 + a.F_cur[3 * (size_t)i + 1];
    F2 = F2 + a.F_cur[3 * (size_t)i + 2];
  }
  pxy = pxy + 0.5f * (u0 * F1 + u1 * F0);
  pxz = pxz + 0.5f * (u0 * F2 + u2 * F0);
  pyz = pyz + 0.5f * (u1 * F2 + u2 * F1);
  // S = -Pi / (2 cs^2 rho tau), tau = 1/omega_p (the global rate)
  const float scale = -(1.0f / ((2.0f / 3.0f) * r * a.tau_eff));
  a.strain[i] = pxy * scale;
  a.strain[N + i] = pxz * scale;
  a.strain[2 * N + i] = pyz * scale;
}

// One neighbour of a finite difference: kind 0 none (non-periodic domain
// face), 1 fluid (v = its cached velocity), 2 solid (v = its wall velocity).
__device__ __forceinline__ int wale_nbr(const WaleArgs& a, bool none, uint32_t j, uint32_t mj, const float* uj,
                                        float v[3]) {
  if (none) {
    v[0] = v[1] = v[2] = 0.0f;
    return 0;
  }
  if (mj & kMaskSolid) {
    if (a.wall_u) {
      v[0] = a.wall_u[3 * (size_t)j + 0];
      v[1] = a.wall_u[3 * (size_t)j + 1];
      v[2] = a.wall_u[3 * (size_t)j + 2];
    } else {
      v[0] = v[1] = v[2] = 0.0f;
    }
    return 2;
  }
  v[0] = uj[0];
  v[1] = uj[1];
  v[2] = uj[2];
  return 1;
}

// nu_t of one fluid cell from its velocity, strain and the neighbours of
// each axis (kp/km kinds, pv/mv values)
__device__ __forceinline__ float wale_nut(const WaleArgs& a, const float own[3], float sxy, float sxz, float syz,
                                          const int kp[3], const int km[3], const float pv[3][3],
                                          const float mv[3][3]) {
  float fd[3][3];
  const bool central = kp[0] == 1 && km[0] == 1 && kp[1] == 1 && km[1] == 1 && kp[2] == 1 && km[2] == 1;
  if (central) {  // every neighbour fluid: the common case
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
      for (int k = 0; k < 3; ++k) fd[k][ax] = (pv[ax][k] - mv[ax][k]) * 0.5f;
  } else {
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float v = 0.0f;
      if (kp[ax] == 1 && km[ax] == 1)
        v = (pv[ax][k] - mv[ax][k]) * 0.5f;
      else if (kp[ax] == 1 && km[ax] == 2)
        v = -(4.0f / 3.0f) * mv[ax][k] + own[k] + (1.0f / 3.0f) * pv[ax][k];
      else if (kp[ax] == 2 && km[ax] == 1)
        v = (4.0f / 3.0f) * pv[ax][k] - own[k] - (1.0f / 3.0f) * mv[ax][k];
      else if (kp[ax] == 1 && km[ax] == 0)
        v = pv[ax][k] - own[k];
      else if (kp[ax] == 0 && km[ax] == 1)
        v = own[k] - mv[ax][k];
      fd[k][ax] = v;
    }
  }
  float g[3][3];
  g[0][0] = fd[0][0];
  g[1][1] = fd[1][1];
  g[2][2] = fd[2][2];
  g[0][1] = sxy + 0.5f * (fd[0][1] - fd[1][0]);
  g[1][0] = sxy + 0.5f * (fd[1][0] - fd[0][1]);
  g[0][2] = sxz + 0.5f * (fd[0][2] - fd[2][0]);
  g[2][0] = sxz + 0.5f * (fd[2][0] - fd[0][2]);
  g[1][2] = syz + 0.5f * (fd[1][2] - fd[2][1]);
  g[2][1] = syz + 0.5f * (fd[2][1] - fd[1][2]);
  // WALE invariants
  float s2 = 0.0f, g2[3][3];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float sv = 0.5f * (g[p][q] + g[q][p]);
      s2 = s2 + sv * sv;
      float acc = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) acc = acc + g[p][k] * g[k][q];
      g2[p][q] = acc;
    }
  const float tr = g2[0][0] + g2[1][1] + g2[2][2];
  float sd2 = 0.0f;
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float sd = 0.5f * (g2[p][q] + g2[q][p]);
      if (p == q) sd = sd - tr / 3.0f;
      sd2 = sd2 + sd * sd;
    }
  // the three powers from square roots (s2^2.5 = s2^2 sqrt(s2),
  // sd2^1.5 = sd2 sqrt(sd2), sd2^1.25 = sd2 sqrt(sqrt(sd2))): a few ulp from
  // the correctly rounded powers, at a fraction of a general powf's cost
  const float rs = sqrtf(s2), rd = sqrtf(sd2);
  const float denom = s2 * s2 * rs + sd2 * sqrtf(rd);
  return denom > 0.0f ? a.cw_delta_sq * (sd2 * rd) / denom : 0.0f;
}

// Pass 2: a block is a 64 x 4 tile of (x, y) columns, each thread marches
// its column over `zlen` cells keeping the mask and cached velocity of
// z - 1, z, z + 1 in registers, so every velocity is read from HBM about
// once; the x and y neighbours come from the rows the tile's own threads
// load (L2).  Each step of the march waits for its loads, so the march is
// kept short enough for the grid to hold many waves (wale_zlen).
constexpr uint32_t kWaleTileX = 64, kWaleTileY = 4;
static_assert(kWaleTileX * kWaleTileY == 256 && kWaleTileX == 64, "one wave per tile row, 256-thread blocks");

template <int Q, typename S>
__global__ __launch_bounds__(256) void k_wale(WaleArgs a) {
  const uint32_t nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  const uint32_t x = blockIdx.x * kWaleTileX + (threadIdx.x & (kWaleTileX - 1));
  const uint32_t y = blockIdx.y * kWaleTileY + threadIdx.x / kWaleTileX;
  const uint32_t z0 = blockIdx.z * a.zlen;
  const uint32_t z1 = min(z0 + a.zlen, nz);
  const size_t N = a.d.N;
  const uint32_t plane = nx * ny;
  uint32_t cnt = 0u, mx = 0u;
  if (x < nx && y < ny && z0 < nz) {
    uint32_t i = (z0 * ny + y) * nx + x;
    const uint32_t iwrap = (nz - 1) * plane;  // z = 0 <-> z = nz - 1
    uint32_t jm = z0 == 0 ? i + iwrap : i - plane;
    uint32_t mm = a.mask[jm], mc = a.mask[i];
    float um[3] = {a.vel[jm], a.vel[N + jm], a.vel[2 * N + jm]};
    float uc[3] = {a.vel[i], a.vel[N + i], a.vel[2 * N + i]};
    for (uint32_t z = z0; z < z1; ++z) {
      const uint32_t jp = z == nz - 1 ? i - iwrap : i + plane;
      const uint32_t mp = a.mask[jp];
      const float up[3] = {a.vel[jp], a.vel[N + jp], a.vel[2 * N + jp]};
      float nut_raw = 0.0f;
      if (!(mc & kMaskSolid)) {
        const Nbr nb(a.d, x, y, z);
        int kp[3], km[3];
        float pv[3][3], mv[3][3];
        {
          const uint32_t j1 = i + (uint32_t)nb.xp, j2 = i + (uint32_t)nb.xm;
          const float u1[3] = {a.vel[j1], a.vel[N + j1], a.vel[2 * N + j1]};
          const float u2[3] = {a.vel[j2], a.vel[N + j2], a.vel[2 * N + j2]};
          kp[0] = wale_nbr(a, x == nx - 1 && !a.d.per[0], j1, a.mask[j1], u1, pv[0]);
          km[0] = wale_nbr(a, x == 0 && !a.d.per[0], j2, a.mask[j2], u2, mv[0]);
        }
        {
          const uint32_t j1 = i + (uint32_t)nb.yp, j2 = i + (uint32_t)nb.ym;
          const float u1[3] = {a.vel[j1], a.vel[N + j1], a.vel[2 * N + j1]};
          const float u2[3] = {a.vel[j2], a.vel[N + j2], a.vel[2 * N + j2]};
          kp[1] = wale_nbr(a, y == ny - 1 && !a.d.per[1], j1, a.mask[j1], u1, pv[1]);
          km[1] = wale_nbr(a, y == 0 && !a.d.per[1], j2, a.mask[j2], u2, mv[1]);
        }
        kp[2] = wale_nbr(a, z == nz - 1 && !a.d.per[2], jp, mp, up, pv[2]);
        km[2] = wale_nbr(a, z == 0 && !a.d.per[2], jm, mm, um, mv[2]);
        nut_raw = wale_nut(a, uc, a.strain[i], a.strain[N + i], a.strain[2 * N + i], kp, km, pv, mv);
      }
      const bool clipped = a.nut_max >= 0.0f && nut_raw > a.nut_max;
      const float nut = clipped ? a.nut_max : nut_raw;
      a.omega[i] = 1.0f / (3.0f * (a.nu0 + nut) + 0.5f);
      if (a.nut) a.nut[i] = nut;
      cnt += clipped ? 1u : 0u;
      mx = max(mx, __float_as_uint(nut_raw));
      // slide the window one cell up
      jm = i;
      mm = mc;
      um[0] = uc[0];
      um[1] = uc[1];
      um[2] = uc[2];
      i = jp;
      mc = mp;
      uc[0] = up[0];
      uc[1] = up[1];
      uc[2] = up[2];
    }
  }
  if (a.stats) {
    // block totals: clipped count and the largest raw nu_t (non-negative, so
    // its bit pattern orders like the value)
    __shared__ uint32_t s_cnt[4], s_max[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      cnt += __shfl_xor(cnt, o);
      mx = max(mx, (uint32_t)__shfl_xor(mx, o));
    }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      s_cnt[w] = cnt;
      s_max[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t c = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
      const uint32_t b = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
      if (c) atomicAdd(&a.stats[0], c);
      if (b) atomicMax(&a.stats[1], b);
    }
  }
}

// March length of the second pass: the longest power of two up to 64 that
// still gives the grid 65,536 blocks, at least 4 (measured best: 4 cells at
// 256^3, 8 at 512^3; 64 cost 4.6 % and 1.5 % of the whole step there);
// LBM_WALE_ZLEN overrides it (tuning).  Any length gives the same values.
static uint32_t wale_zlen(const Dom& d) {
  if (const char* e = getenv("LBM_WALE_ZLEN")) {
    const long v = strtol(e, nullptr, 10);
    if (v >= 1 && v <= 1024) return (uint32_t)v;
  }
  const uint64_t cols = (uint64_t)((d.nx + kWaleTileX - 1) / kWaleTileX) * ((d.ny + kWaleTileY - 1) / kWaleTileY);
  uint32_t z = 64;
  while (z > 4 && cols * ((d.nz + z - 1) / z) < 65536) z >>= 1;
  return z;
}

template <typename S>
static hipError_t wale_s(int Q, WaleArgs a, hipStream_t s) {
  a.zlen = wale_zlen(a.d);
  const dim3 grid((a.d.N + 255) / 256), block(256);
  const dim3 tiles((a.d.nx + kWaleTileX - 1) / kWaleTileX, (a.d.ny + kWaleTileY - 1) / kWaleTileY,
                   (a.d.nz + a.zlen - 1) / a.zlen);
  if (Q == 19) {
    hipLaunchKernelGGL((k_wale_local<19, S>), grid, block, 0, s, a);
    hipLaunchKernelGGL((k_wale<19, S>), tiles, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_wale_local<27, S>), grid, block, 0, s, a);
    hipLaunchKernelGGL((k_wale<27, S>), tiles, block, 0, s, a);
  }
  return hipGetLastError();
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

hipError_t launch_wale(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                       const float* wall_u, const float* F_cur, const float* F_mom, float* work, const HydroParams& p,
                       float tau_eff, float cw_delta_sq, float nu0, float nut_max, float* omega, float* nut,
                       uint32_t* stats, hipStream_t s) {
  if (d.N == 0) return hipSuccess;
  if (stats) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, stats, 2u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  WaleArgs a{d,           f,       mask, wall_u, F_cur, F_mom, p,     odd, tau_eff, cw_delta_sq,
             nu0,         nut_max, work, work + 3 * (size_t)d.N, omega, nut, stats};
  if (storage == kF32) return wale_s<float>(Q, a, s);
  if (storage == kF16) return wale_s<_Float16>(Q, a, s);
  return wale_s<bf16s>(Q, a, s);
}

}  // namespace lbm
