// lbm_les.hip — WALE large-eddy closure (Nicoud & Ducros 1999) as a
// producer of the per-cell omega_plus field the collide-stream kernels read.
//
// One cell-parallel kernel per update, after the moment kernel has written
// the cached moments (rho, u) of the current state into a 4 x N buffer:
//   strain S from the non-equilibrium stress of the cell's own populations
//     (strain_rate_at, solver.rs:5393-5447),
//   velocity gradient g = du_i/dx_j: diagonal from finite differences of the
//     cached velocity, off-diagonal S + antisymmetric part of the finite
//     differences (gather_velocity_gradient, solver.rs:5490-5581),
//   nu_t = (Cw Delta)^2 (Sd:Sd)^(3/2) / ((S:S)^(5/2) + (Sd:Sd)^(5/4)),
//     clipped to nu_t_max when asked, omega = 1/(3 (nu0 + nu_t) + 1/2)
//     (WaleLes::update, les.rs:120-211).
// Each cell reads its Q populations once and the velocity of its six axis
// neighbours (served from L2: neighbouring cells of a block share them), so
// the kernel moves ~4Q + 16 + 4 + 4 bytes per cell; HBM-bound like the step.
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
  const float* mom;     // 4 x N: rho, ux, uy, uz (cached moments)
  HydroParams p;        // the uniform force
  int odd;
  float tau_eff;      // 1/omega_p of the global rate (stress scale)
  float cw_delta_sq;  // (Cw Delta)^2
  float nu0;          // laminar viscosity
  float nut_max;      // < 0: no clip
  float* omega;       // N: omega_plus out
  float* nut;         // N or null: nu_t out (after the clip)
  uint32_t* stats;    // null or [clipped cells, bits of the largest raw nu_t]
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

template <int Q, typename S>
__global__ __launch_bounds__(256) void k_wale(WaleArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < a.d.N;
  float nut_raw = 0.0f;
  if (live && !(a.mask[i] & kMaskSolid)) {
    const uint32_t m = a.mask[i];
    const size_t N = a.d.N;
    uint32_t x, y, z;
    cell_xyz(a.d, i, x, y, z);
    const Nbr nb(a.d, x, y, z);
    const float r = a.mom[i];
    const float u0 = a.mom[N + i], u1 = a.mom[2 * N + i], u2 = a.mom[3 * N + i];
    // Pi_neq = sum_q c c (f_q - feq_q), deviation-form equilibrium
    // (kernels.rs:103-121); terms with a zero factor add exact zeros in the
    // reference's sums and are dropped.
    const float usq = u0 * u0 + u1 * u1 + u2 * u2;
    const float drho = r - 1.0f;
    float pxx = 0.0f, pyy = 0.0f, pzz = 0.0f, pxy = 0.0f, pxz = 0.0f, pyz = 0.0f;
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = lat_c(q, 0), cy = lat_c(q, 1), cz = lat_c(q, 2);
      const float fq = read_pop_nt<Q, S, q>((const S*)a.f, a.d.ps, nb, i, m, a.odd != 0);
      const float cu = cdot<q>(u0, u1, u2);
      const float feq = lat_wf(Q, q) * (drho + r * (3.0f * cu + 4.5f * cu * cu - 1.5f * usq));
      const float fn = fq - feq;
      if constexpr (cx != 0) pxx = pxx + fn;
      if constexpr (cy != 0) pyy = pyy + fn;
      if constexpr (cz != 0) pzz = pzz + fn;
      if constexpr (cx * cy == 1) pxy = pxy + fn;
      if constexpr (cx * cy == -1) pxy = pxy - fn;
      if constexpr (cx * cz == 1) pxz = pxz + fn;
      if constexpr (cx * cz == -1) pxz = pxz - fn;
      if constexpr (cy * cz == 1) pyz = pyz + fn;
      if constexpr (cy * cz == -1) pyz = pyz - fn;
    });
    // Guo-force correction of the stress: + u F (diagonal), + (u_a F_b +
    // u_b F_a)/2 (off-diagonal), F = uniform force + current field
    float F0 = a.p.F[0], F1 = a.p.F[1], F2 = a.p.F[2];
    if (a.F_cur && (m & kMaskForced)) {  // the field is zero off the force list
      F0 = F0 + a.F_cur[3 * (size_t)i + 0];
      F1 = F1 + a.F_cur[3 * (size_t)i + 1];
      F2 = F2 + a.F_cur[3 * (size_t)i + 2];
    }
    pxx = pxx + u0 * F0;
    pyy = pyy + u1 * F1;
    pzz = pzz + u2 * F2;
    pxy = pxy + 0.5f * (u0 * F1 + u1 * F0);
    pxz = pxz + 0.5f * (u0 * F2 + u2 * F0);
    pyz = pyz + 0.5f * (u1 * F2 + u2 * F1);
    // S = -Pi / (2 cs^2 rho tau), tau = 1/omega_p (the global rate)
    const float scale = -(1.0f / ((2.0f / 3.0f) * r * a.tau_eff));
    const float sxy = pxy * scale, sxz = pxz * scale, syz = pyz * scale;
    (void)pxx;
    (void)pyy;
    (void)pzz;
    // finite differences of the cached velocity along each axis
    const float own[3] = {u0, u1, u2};
    const uint32_t coord[3] = {x, y, z};
    const uint32_t ext[3] = {a.d.nx, a.d.ny, a.d.nz};
    const int32_t offp[3] = {nb.xp, nb.yp, nb.zp}, offm[3] = {nb.xm, nb.ym, nb.zm};
    float fd[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      // kind 0: no neighbour (non-periodic domain face), 1: fluid, 2: solid
      auto probe = [&](bool at_face, int32_t off, float v[3]) -> int {
        if (at_face && !a.d.per[ax]) {
          v[0] = v[1] = v[2] = 0.0f;
          return 0;
        }
        const uint32_t j = i + (uint32_t)off;
        if (a.mask[j] & kMaskSolid) {
          if (a.wall_u) {
            v[0] = a.wall_u[3 * (size_t)j + 0];
            v[1] = a.wall_u[3 * (size_t)j + 1];
            v[2] = a.wall_u[3 * (size_t)j + 2];
          } else {
            v[0] = v[1] = v[2] = 0.0f;
          }
          return 2;
        }
        v[0] = a.mom[N + j];
        v[1] = a.mom[2 * N + j];
        v[2] = a.mom[3 * N + j];
        return 1;
      };
      float pv[3], mv[3];
      const int kp = probe(coord[ax] == ext[ax] - 1, offp[ax], pv);
      const int km = probe(coord[ax] == 0, offm[ax], mv);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float v = 0.0f;
        if (kp == 1 && km == 1)
          v = (pv[k] - mv[k]) * 0.5f;
        else if (kp == 1 && km == 2)
          v = -(4.0f / 3.0f) * mv[k] + own[k] + (1.0f / 3.0f) * pv[k];
        else if (kp == 2 && km == 1)
          v = (4.0f / 3.0f) * pv[k] - own[k] - (1.0f / 3.0f) * mv[k];
        else if (kp == 1 && km == 0)
          v = pv[k] - own[k];
        else if (kp == 0 && km == 1)
          v = own[k] - mv[k];
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
    const float denom = powf(s2, 2.5f) + powf(sd2, 1.25f);
    nut_raw = denom > 0.0f ? a.cw_delta_sq * powf(sd2, 1.5f) / denom : 0.0f;
  }
  const bool clipped = live && a.nut_max >= 0.0f && nut_raw > a.nut_max;
  if (live) {
    const float nut = clipped ? a.nut_max : nut_raw;
    a.omega[i] = 1.0f / (3.0f * (a.nu0 + nut) + 0.5f);
    if (a.nut) a.nut[i] = nut;
  }
  if (a.stats) {
    // block totals: clipped count and the largest raw nu_t (non-negative, so
    // its bit pattern orders like the value)
    __shared__ uint32_t s_cnt[4], s_max[4];
    uint32_t cnt = clipped ? 1u : 0u, mx = __float_as_uint(nut_raw);
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

template <typename S>
static hipError_t wale_s(int Q, const WaleArgs& a, hipStream_t s) {
  const dim3 grid((a.d.N + 255) / 256), block(256);
  if (Q == 19)
    hipLaunchKernelGGL((k_wale<19, S>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_wale<27, S>), grid, block, 0, s, a);
  return hipGetLastError();
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

hipError_t launch_wale(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                       const float* wall_u, const float* F_cur, const float* mom, const HydroParams& p,
                       float tau_eff, float cw_delta_sq, float nu0, float nut_max, float* omega, float* nut, uint32_t* stats,
                       hipStream_t s) {
  if (d.N == 0) return hipSuccess;
  if (stats) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, stats, 2u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  WaleArgs a{d, f, mask, wall_u, F_cur, mom, p, odd, tau_eff, cw_delta_sq, nu0, nut_max, omega, nut, stats};
  if (storage == kF32) return wale_s<float>(Q, a, s);
  if (storage == kF16) return wale_s<_Float16>(Q, a, s);
  return wale_s<bf16s>(Q, a, s);
}

}  // namespace lbm
