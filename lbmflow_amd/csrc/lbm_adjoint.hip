// lbm_adjoint.hip — reverse-mode (discrete adjoint) kernel of one D3Q19
// TRT/BGK collide-stream step with half-way bounce-back and Guo forcing.
//
// Forward map of one step at a fluid cell i (collision.rs:123-202 with the
// stale-moment force of solver.rs:2805-2822, stream_row kernels.rs:547-622):
//   rho = 1 + sum f,  m = sum c f,  u = (m + F_mom/2)/rho
//   feq_q = w_q (rho - 1 + rho (3 cu + 4.5 cu^2 - 1.5 u.u))
//   S_q   = w_q (3 (c_q.F - u.F) + 9 cu c_q.F)          (F = F_cur)
//   f*_q  = f_q - w+ (f+ - feq+) - w- (f- - feq-) + (1 - w+/2) S+ + (1 - w-/2) S-
//   f'_q(i) = f*_q(i - c_q), or f*_opp(q)(i) when i - c_q is solid.
// Forced cells (the force list) carry F_mom and F_cur; every other cell has
// no force.  w+ may vary per cell (omega field); w- is global.
//
// The kernel is the transposed Jacobian of that map at the forward state:
// given lam' (adjoint of f' at step n+1) it pulls lam* = adjoint of f*
// through the streaming (lam*_q(i) = lam'_q(i + c_q), or lam'_opp(q)(i)
// across a wall link) and returns the adjoint of f (lam), of w+ (added to
// omega_bar), of F_cur (Fbar) and of F_mom (Gout) at forced cells.  All
// adjoint arithmetic is f64; the forward state is the stored fp32 (or
// 16-bit) populations, read in the parity-dependent AA layout.
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

namespace lbm {

struct AdjArgs {
  Dom d;
  const void* f;
  const uint32_t* mask;
  const float* omega;  // per-cell omega_plus or null
  const float* F_cur;  // N x 3 or null
  const float* F_mom;  // N x 3 or null
  HydroParams p;
  const double* lam_in;  // Q x N: adjoint of the populations after the step
  double* lam_out;       // Q x N: adjoint of the populations before it
  double* omega_bar;     // N: accumulated adjoint of omega_plus
  double* Fbar;          // N x 3: adjoint of F_cur (forced cells; others untouched)
  double* Gout;          // N x 3: adjoint of F_mom (forced cells; others untouched)
};

template <typename S, bool ODD>
__global__ __launch_bounds__(256) void k_adj_collide_stream(AdjArgs a) {
  constexpr int Q = 19;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.d.N) return;
  const uint32_t m = a.mask[i];
  const size_t N = a.d.N;
  if (m & kMaskSolid) {
    sfor<0, Q>([&](auto qc) { a.lam_out[(size_t)decltype(qc)::value * N + i] = 0.0; });
    return;
  }
  uint32_t x, y, z;
  cell_xyz(a.d, i, x, y, z);
  const Nbr nb(a.d, x, y, z);
  // forward state
  double f[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[q] = (double)read_pop<Q, S, q>((const S*)a.f, a.d.ps, nb, i, m, ODD);
  });
  // adjoint of the post-collision populations, pulled through the streaming
  double ls[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = lat_opp(q);
    if ((m >> o) & 1u) {
      ls[q] = a.lam_in[(size_t)o * N + i];
    } else {
      const uint32_t j = i + (uint32_t)nb.template off<lat_c(q, 0), lat_c(q, 1), lat_c(q, 2)>();
      ls[q] = a.lam_in[(size_t)q * N + j];
    }
  });
  const bool forced = (m & kMaskForced) != 0 && a.F_cur && a.F_mom;
  double Fm[3] = {0.0, 0.0, 0.0}, Fc[3] = {0.0, 0.0, 0.0};
  if (forced)
    for (int c = 0; c < 3; ++c) {
      Fm[c] = (double)a.F_mom[3 * (size_t)i + c];
      Fc[c] = (double)a.F_cur[3 * (size_t)i + c];
    }
  const double op = a.omega ? (double)a.omega[i] : (double)a.p.omega_p;
  const double cp = 1.0 - 0.5 * op;
  const double om = (double)a.p.omega_m, cm = (double)a.p.cm;
  // moments
  double r = 1.0, mv[3] = {0.0, 0.0, 0.0};
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    r += f[q];
    if constexpr (lat_c(q, 0) != 0) mv[0] += (double)lat_c(q, 0) * f[q];
    if constexpr (lat_c(q, 1) != 0) mv[1] += (double)lat_c(q, 1) * f[q];
    if constexpr (lat_c(q, 2) != 0) mv[2] += (double)lat_c(q, 2) * f[q];
  });
  const double u[3] = {(mv[0] + 0.5 * Fm[0]) / r, (mv[1] + 0.5 * Fm[1]) / r, (mv[2] + 0.5 * Fm[2]) / r};
  const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double uf = u[0] * Fc[0] + u[1] * Fc[1] + u[2] * Fc[2];
  double lo[Q], eb[Q], sb[Q];  // adjoint of f (direct part), of feq, of S
  double obar = 0.0;
  auto feq = [&](double cu, double w) { return w * (r - 1.0 + r * (3.0 * cu + 4.5 * cu * cu - 1.5 * usq)); };
  auto src = [&](double cu, double cf, double w) { return w * (3.0 * (cf - uf) + 9.0 * cu * cf); };
  {  // rest population
    const double w0 = (double)lat_wf(Q, 0);
    const double e0 = feq(0.0, w0), s0 = forced ? src(0.0, 0.0, w0) : 0.0;
    lo[0] = ls[0] * (1.0 - op);
    eb[0] = ls[0] * op;
    sb[0] = ls[0] * cp;
    obar += ls[0] * (-(f[0] - e0) - 0.5 * s0);
  }
  sfor<0, (Q - 1) / 2>([&](auto kc) {
    constexpr int qa = 2 * decltype(kc)::value + 1;
    constexpr int qb = qa + 1;
    const double w = (double)lat_wf(Q, qa);
    const double cua = (double)lat_c(qa, 0) * u[0] + (double)lat_c(qa, 1) * u[1] + (double)lat_c(qa, 2) * u[2];
    const double cub = -cua;
    const double ea = feq(cua, w), ebb = feq(cub, w);
    double sa = 0.0, sbb = 0.0;
    if (forced) {
      const double cfa = (double)lat_c(qa, 0) * Fc[0] + (double)lat_c(qa, 1) * Fc[1] + (double)lat_c(qa, 2) * Fc[2];
      sa = src(cua, cfa, w);
      sbb = src(cub, -cfa, w);
    }
    const double Lp = ls[qa] + ls[qb], Lm = ls[qa] - ls[qb];
    lo[qa] = ls[qa] - 0.5 * op * Lp - 0.5 * om * Lm;
    lo[qb] = ls[qb] - 0.5 * op * Lp + 0.5 * om * Lm;
    eb[qa] = 0.5 * op * Lp + 0.5 * om * Lm;
    eb[qb] = 0.5 * op * Lp - 0.5 * om * Lm;
    sb[qa] = 0.5 * cp * Lp + 0.5 * cm * Lm;
    sb[qb] = 0.5 * cp * Lp - 0.5 * cm * Lm;
    const double fp = 0.5 * (f[qa] + f[qb]), ep = 0.5 * (ea + ebb), sp = 0.5 * (sa + sbb);
    obar += Lp * (-(fp - ep) - 0.5 * sp);
  });
  // feq and S back to rho, u, F_cur
  double rb = 0.0, ub[3] = {0.0, 0.0, 0.0}, fcb[3] = {0.0, 0.0, 0.0};
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const double w = (double)lat_wf(Q, q);
    const double c[3] = {(double)lat_c(q, 0), (double)lat_c(q, 1), (double)lat_c(q, 2)};
    const double cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
    rb += eb[q] * w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    for (int al = 0; al < 3; ++al) ub[al] += eb[q] * w * r * (3.0 * c[al] + 9.0 * cu * c[al] - 3.0 * u[al]);
    if (forced) {
      const double cf = c[0] * Fc[0] + c[1] * Fc[1] + c[2] * Fc[2];
      for (int al = 0; al < 3; ++al) {
        ub[al] += sb[q] * w * (-3.0 * Fc[al] + 9.0 * c[al] * cf);
        fcb[al] += sb[q] * w * (3.0 * (c[al] - u[al]) + 9.0 * cu * c[al]);
      }
    }
  });
  // u = (m + F_mom/2)/rho
  double mb[3];
  for (int al = 0; al < 3; ++al) {
    mb[al] = ub[al] / r;
    rb += -ub[al] * u[al] / r;
  }
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    double v = lo[q] + rb;
    if constexpr (lat_c(q, 0) != 0) v += (double)lat_c(q, 0) * mb[0];
    if constexpr (lat_c(q, 1) != 0) v += (double)lat_c(q, 1) * mb[1];
    if constexpr (lat_c(q, 2) != 0) v += (double)lat_c(q, 2) * mb[2];
    a.lam_out[(size_t)q * N + i] = v;
  });
  a.omega_bar[i] += obar;
  if (forced)
    for (int al = 0; al < 3; ++al) {
      a.Fbar[3 * (size_t)i + al] = fcb[al];
      a.Gout[3 * (size_t)i + al] = ub[al] / (2.0 * r);
    }
}

template <typename S>
static hipError_t adj_s(int odd, const AdjArgs& a, hipStream_t s) {
  const dim3 grid((a.d.N + 255) / 256), block(256);
  if (odd)
    hipLaunchKernelGGL((k_adj_collide_stream<S, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_adj_collide_stream<S, false>), grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adj_collide_stream(int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                                     const float* omega, const float* F_cur, const float* F_mom,
                                     const HydroParams& p, const double* lam_in, double* lam_out, double* omega_bar,
                                     double* Fbar, double* Gout, hipStream_t s) {
  AdjArgs a{d, f, mask, omega, F_cur, F_mom, p, lam_in, lam_out, omega_bar, Fbar, Gout};
  if (storage == kF32) return adj_s<float>(odd, a, s);
  if (storage == kF16) return adj_s<_Float16>(odd, a, s);
  return adj_s<bf16s>(odd, a, s);
}

__global__ void k_fill_d(double* p, size_t n, double v) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

hipError_t launch_fill_f64(double* p, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace lbm
