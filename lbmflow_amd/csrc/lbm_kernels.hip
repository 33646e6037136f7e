// lbm_kernels.hip — CDNA4 (gfx950) kernels for lbm-core's lattice-Boltzmann
// step: fused in-place (AA-pattern) collide-stream over SoA populations with
// half-way bounce-back, moving walls and Guo forcing; TRT/BGK and the
// factorised central-moment collision; masks, device geometry, staging,
// moments and f64 reductions.
//
// Storage: q-major planes f[q*ps + cell] of the deviation f - w_q in fp32,
// or scaled by 1/w_q in fp16 / bf16.  Arithmetic is fp32.  Reductions are
// fp64.  The whole file is compiled with -ffp-contract=off so the TRT path
// keeps the reference's separately rounded multiply/add sequence
// (collision.rs:123-202) and reproduces the CpuScalar f32 trajectory
// bit for bit; the central-moment path opts into contraction locally, per
// source expression (contract(on)), so every kernel that inlines it rounds
// identically (the batched launch reproduces the single-domain one).
//
// AA pattern (in-place streaming, one buffer):
//   even step  cell i reads f_q from slot (i, q), writes the post-collision
//              value moving along c_q to slot (i, opp q);
//   odd step   cell i reads f_q from slot (i - c_q, opp q), writes the value
//              moving along c_q to slot (i + c_q, q).
// A post-collision population heading into a solid neighbour s = i + c_q is
// reflected immediately: on even steps it is written to the solid's slot
// (s, q) -- where the odd step's uniform read of i looks for the incoming
// opp(q) population -- and on odd steps to (i, opp q), where the next even
// step reads it.  Each slot is read and written by exactly one cell per
// step, so cells are independent and no second buffer is needed.
#include <algorithm>
#include <type_traits>

#include "lbm_device.hpp"

namespace lbm {

// ---------------------------------------------------------------------------
// TRT/BGK with Guo forcing: TrtGuo::relax (collision.rs:123-202) with the
// equilibrium of SecondOrderEq::feq (81-102) and GuoForcing::source
// (104-121).  Same operand order, no contraction.
// ---------------------------------------------------------------------------
template <int Q>
__device__ __forceinline__ void collide_trt(float* fv, float r, float ux, float uy, float uz, const float Fc[3],
                                            bool fon, float op, float cp, const HydroParams& p) {
  const float half = 0.5f;
  float usq = ux * ux;
  usq = usq + uy * uy;
  usq = usq + uz * uz;
  float uf = 0.0f;
  if (fon) {
    uf = ux * Fc[0];
    uf = uf + uy * Fc[1];
    uf = uf + uz * Fc[2];
  }
  const float drho = r - 1.0f;
  const float f15_usq = 1.5f * usq;
  auto feq_of = [&](float cu, float w) {
    float three_cu = 3.0f * cu;
    float f45_cu = 4.5f * cu;
    float f45_cu2 = f45_cu * cu;
    float sum = three_cu + f45_cu2;
    float inner = sum - f15_usq;
    float rho_inner = r * inner;
    float body = drho + rho_inner;
    return w * body;
  };
  auto src_of = [&](float cu, float cf, float w) {
    float cf_minus_uf = cf - uf;
    float first = 3.0f * cf_minus_uf;
    float nine_cu = 9.0f * cu;
    float second = nine_cu * cf;
    float body = first + second;
    return w * body;
  };
  // rest population
  {
    const float w0 = lat_wf(Q, 0);
    float feq0 = feq_of(0.0f * ux, w0);  // cu of the rest velocity is a signed zero
    float f0 = fv[0];
    float relaxed = f0 - op * (f0 - feq0);
    if (fon) {
      float cf0 = 0.0f;
      float src0 = src_of(0.0f, cf0, w0);
      relaxed = relaxed + cp * src0;
    }
    fv[0] = relaxed;
  }
  const float om = p.omega_m, cm = p.cm;
  // pairs (q, q+1), q odd, ascending (lattice.rs PAIRS)
  sfor<0, (Q - 1) / 2>([&](auto kc) {
    constexpr int qa = 2 * decltype(kc)::value + 1;
    constexpr int qb = qa + 1;
    constexpr float w = lat_wf(Q, qa);
    const float cua = cdot<qa>(ux, uy, uz);
    const float cub = cdot<qb>(ux, uy, uz);
    const float ea = feq_of(cua, w);
    const float eb = feq_of(cub, w);
    const float fa = fv[qa], fb = fv[qb];
    const float fp = half * (fa + fb);
    const float fm = half * (fa - fb);
    const float ep = half * (ea + eb);
    const float em = half * (ea - eb);
    const float rp = op * (fp - ep);
    const float rm = om * (fm - em);
    float va = (fa - rp) - rm;
    float vb = (fb - rp) + rm;
    if (fon) {
      const float sa = src_of(cua, cdot<qa>(Fc[0], Fc[1], Fc[2]), w);
      const float sb = src_of(cub, cdot<qb>(Fc[0], Fc[1], Fc[2]), w);
      const float sp = half * (sa + sb);
      const float sm = half * (sa - sb);
      const float cp_sp = cp * sp;
      const float cm_sm = cm * sm;
      va = va + (cp_sp + cm_sm);
      vb = vb + (cp_sp - cm_sm);
    }
    fv[qa] = va;
    fv[qb] = vb;
  });
}

// ---------------------------------------------------------------------------
// Central-moment collision (collide_row_central_moment, kernels.rs:406-533)
// in non-equilibrium form.  With feq the deviation equilibrium and
// K = central moments of (f - feq), S = central moments of the Guo source,
// the reference's post-collision populations are feq + B(d), where B is the
// inverse central-moment transform and
//   order 0,1:       d = K + S
//   order 2, a!=b:   d = (1-os) K + (1-os/2) S
//   order 2, a==b:   d = (1-os)(K_aa - K_bulk) + S_bulk/2 + (1-os/2)(S_aa - S_bulk)
//   order >= 3:      d = S/2
// (the reference relaxes the trace and all orders >= 3 at rate 1).  The
// D3Q27 transforms factorise into three passes of 1-D three-point lines
// (tensor-product basis, exponents 0..2 per axis), so no 27x27 matrix
// is ever formed.  Working on f - feq keeps the fp32 arithmetic on small
// numbers; the rounding difference to the reference's f64 internals stays
// well below the 1e-6 population gate.
// ---------------------------------------------------------------------------
constexpr int gidx(int cx, int cy, int cz) { return (cx + 1) * 9 + (cy + 1) * 3 + (cz + 1); }

// forward 1-D central line: (g-, g0, g+) -> (k0, k1, k2)
__device__ __forceinline__ void fwd_line(float& a0, float& a1, float& a2, float u) {
#pragma clang fp contract(on)
  const float gm = a0, g0 = a1, gp = a2;
  const float s = gp + gm, t = gp - gm;
  const float k0 = s + g0;
  const float k1 = t - u * k0;
  const float k2 = (s - 2.0f * u * t) + u * u * k0;
  a0 = k0;
  a1 = k1;
  a2 = k2;
}
// inverse 1-D central line: (k0, k1, k2) -> (g-, g0, g+)
__device__ __forceinline__ void inv_line(float& a0, float& a1, float& a2, float u) {
#pragma clang fp contract(on)
  const float k0 = a0, k1 = a1, k2 = a2;
  const float m1 = k1 + u * k0;
  const float m2 = (k2 + 2.0f * u * k1) + u * u * k0;
  a0 = 0.5f * (m2 - m1);
  a1 = k0 - m2;
  a2 = 0.5f * (m1 + m2);
}

// Full 27-point transforms on G[(a*9 + b*3 + c)], a/b/c = x/y/z index 0..2.
__device__ __forceinline__ void fwd27(float* G, float ux, float uy, float uz) {
#pragma unroll
  for (int ab = 0; ab < 9; ++ab) fwd_line(G[ab * 3 + 0], G[ab * 3 + 1], G[ab * 3 + 2], uz);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) fwd_line(G[a * 9 + 0 + c], G[a * 9 + 3 + c], G[a * 9 + 6 + c], uy);
#pragma unroll
  for (int bc = 0; bc < 9; ++bc) fwd_line(G[bc], G[9 + bc], G[18 + bc], ux);
}
__device__ __forceinline__ void inv27(float* G, float ux, float uy, float uz) {
#pragma unroll
  for (int bc = 0; bc < 9; ++bc) inv_line(G[bc], G[9 + bc], G[18 + bc], ux);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) inv_line(G[a * 9 + 0 + c], G[a * 9 + 3 + c], G[a * 9 + 6 + c], uy);
#pragma unroll
  for (int ab = 0; ab < 9; ++ab) inv_line(G[ab * 3 + 0], G[ab * 3 + 1], G[ab * 3 + 2], uz);
}

// deviation equilibrium of all 27 directions (same formula as the TRT
// path, here with contraction allowed)
__device__ __forceinline__ void feq27(float* feq, float r, float ux, float uy, float uz) {
#pragma clang fp contract(on)
  const float usq = ux * ux + uy * uy + uz * uz;
  const float drho = r - 1.0f;
  const float b = drho - r * 1.5f * usq;
  sfor<0, 27>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float cu = cdot<q>(ux, uy, uz);
    feq[q] = lat_wf(27, q) * (b + r * cu * (3.0f + 4.5f * cu));
  });
}

__device__ __forceinline__ float cm_rate(float os_base, float ux, float uy, float uz) {
#pragma clang fp contract(on)
  const float usq = ux * ux + uy * uy + uz * uz;
  return fminf(os_base * (1.0f - 0.16f * usq), 2.0f);
}

// General form (any force): full transforms of f - feq and of the source.
__device__ __forceinline__ void collide_cm27_general(float* fv, float r, float ux, float uy, float uz,
                                                     const float F[3], bool fon, float os_base) {
#pragma clang fp contract(on)
  float feq[27];
  feq27(feq, r, ux, uy, uz);
  float K[27], S[27];
  sfor<0, 27>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int g = gidx(lat_c(q, 0), lat_c(q, 1), lat_c(q, 2));
    K[g] = fv[q] - feq[q];
  });
  fwd27(K, ux, uy, uz);
  if (fon) {
    const float uf = ux * F[0] + uy * F[1] + uz * F[2];
    sfor<0, 27>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int g = gidx(lat_c(q, 0), lat_c(q, 1), lat_c(q, 2));
      const float cu = cdot<q>(ux, uy, uz);
      const float cf = cdot<q>(F[0], F[1], F[2]);
      S[g] = lat_wf(27, q) * (3.0f * (cf - uf) + 9.0f * cu * cf);
    });
    fwd27(S, ux, uy, uz);
  } else {
#pragma unroll
    for (int g = 0; g < 27; ++g) S[g] = 0.0f;
  }
  const float os = cm_rate(os_base, ux, uy, uz);
  const float a1 = 1.0f - os, a2 = 1.0f - 0.5f * os;
  const int ixx = 2 * 9, iyy = 2 * 3, izz = 2;
  const float kb = (K[ixx] + K[iyy] + K[izz]) * (1.0f / 3.0f);
  const float sb = (S[ixx] + S[iyy] + S[izz]) * (1.0f / 3.0f);
  float D[27];
#pragma unroll
  for (int g = 0; g < 27; ++g) {
    const int ex = g / 9, ey = (g / 3) % 3, ez = g % 3;
    const int order = ex + ey + ez;
    if (order <= 1)
      D[g] = K[g] + S[g];
    else if (order == 2 && (ex == 2 || ey == 2 || ez == 2))
      D[g] = a1 * (K[g] - kb) + 0.5f * sb + a2 * (S[g] - sb);
    else if (order == 2)
      D[g] = a1 * K[g] + a2 * S[g];
    else
      D[g] = 0.5f * S[g];
  }
  inv27(D, ux, uy, uz);
  sfor<0, 27>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int g = gidx(lat_c(q, 0), lat_c(q, 1), lat_c(q, 2));
    fv[q] = feq[q] + D[g];
  });
}

// D3Q19 central moments.  The basis is the reference's (central_basis,
// kernels.rs:162-183): exponents (ex, ey, ez) in {0,1,2}^3, x outermost,
// without the eight with all three positive.  It is not a tensor product, so
// the transforms are the raw-moment matrix R (entries c^e in {-1, 0, 1}),
// the binomial shift by -u / +u (kernels.rs:332-370) and the inverse
// P = R^-1, whose entries (0, +-1/4, +-1/2, +-1) are exact in fp32.  All three
// are compile-time tables; zero terms vanish when the loops unroll.
struct Cm19Tables {
  int e[19][3];
  int idx[3][3][3];  // exponent -> moment, -1 when not in the basis
  double P[19][19];
};

constexpr int cm19_phi(int q, int m, const Cm19Tables& t) {
  int v = 1;
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < t.e[m][a]; ++k) v *= lat_c(q, a);
  return v;
}

constexpr Cm19Tables make_cm19() {
  Cm19Tables t{};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) t.idx[a][b][c] = -1;
  int n = 0;
  for (int ax = 0; ax <= 2; ++ax)
    for (int ay = 0; ay <= 2; ++ay)
      for (int az = 0; az <= 2; ++az) {
        if (ax > 0 && ay > 0 && az > 0) continue;
        t.e[n][0] = ax;
        t.e[n][1] = ay;
        t.e[n][2] = az;
        t.idx[ax][ay][az] = n;
        ++n;
      }
  // Gauss-Jordan inverse of R (partial pivoting), exact here: R is integer
  // with a power-of-two-denominator inverse
  double A[19][38] = {};
  for (int m = 0; m < 19; ++m) {
    for (int q = 0; q < 19; ++q) A[m][q] = cm19_phi(q, m, t);
    A[m][19 + m] = 1.0;
  }
  for (int col = 0; col < 19; ++col) {
    int piv = col;
    for (int r = col + 1; r < 19; ++r)
      if ((A[r][col] < 0 ? -A[r][col] : A[r][col]) > (A[piv][col] < 0 ? -A[piv][col] : A[piv][col])) piv = r;
    for (int j = 0; j < 38; ++j) {
      const double tmp = A[col][j];
      A[col][j] = A[piv][j];
      A[piv][j] = tmp;
    }
    const double inv = 1.0 / A[col][col];
    for (int j = 0; j < 38; ++j) A[col][j] *= inv;
    for (int r = 0; r < 19; ++r) {
      if (r == col || A[r][col] == 0.0) continue;
      const double f = A[r][col];
      for (int j = 0; j < 38; ++j) A[r][j] -= f * A[col][j];
    }
  }
  for (int q = 0; q < 19; ++q)
    for (int m = 0; m < 19; ++m) t.P[q][m] = A[q][19 + m];
  return t;
}
inline constexpr Cm19Tables kCm19 = make_cm19();

constexpr int binom2(int n, int k) { return (k == 0 || k == n) ? 1 : 2; }  // n <= 2

// u^k for k in 0..2
__device__ __forceinline__ float pow2u(float u, int k) { return k == 0 ? 1.0f : (k == 1 ? u : u * u); }

// central moments (shift by -u) of the raw moments R, or raw moments (shift
// by +u) of central moments: out[m] = sum_{k <= e_m} C(e,k) s^(e-k) in[idx(k)]
template <bool TO_CENTRAL>
__device__ __forceinline__ void cm19_shift(const float* in, float* out, float ux, float uy, float uz) {
#pragma clang fp contract(on)
  const float sx = TO_CENTRAL ? -ux : ux, sy = TO_CENTRAL ? -uy : uy, sz = TO_CENTRAL ? -uz : uz;
  sfor<0, 19>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    constexpr int ex = kCm19.e[m][0], ey = kCm19.e[m][1], ez = kCm19.e[m][2];
    float acc = in[m];  // k = e: coefficient 1
    sfor<0, (ex + 1) * (ey + 1) * (ez + 1)>([&](auto kc) {
      constexpr int kk = decltype(kc)::value;
      constexpr int kx = kk / ((ey + 1) * (ez + 1)), ky = (kk / (ez + 1)) % (ey + 1), kz = kk % (ez + 1);
      if constexpr (!(kx == ex && ky == ey && kz == ez)) {
        constexpr int j = kCm19.idx[kx][ky][kz];
        constexpr float coef = (float)(binom2(ex, kx) * binom2(ey, ky) * binom2(ez, kz));
        acc = acc + coef * (pow2u(sx, ex - kx) * pow2u(sy, ey - ky) * pow2u(sz, ez - kz)) * in[j];
      }
    });
    out[m] = acc;
  });
}

// raw moments of a population vector: R[m] = sum_q c_q^e_m v_q
__device__ __forceinline__ void cm19_raw(const float* v, float* R) {
  sfor<0, 19>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    float acc = 0.0f;
    sfor<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int c = cm19_phi(q, m, kCm19);
      if constexpr (c == 1) acc = acc + v[q];
      if constexpr (c == -1) acc = acc - v[q];
    });
    R[m] = acc;
  });
}

// D3Q19 central-moment collision, general form (any force): feq + B(d) as
// for D3Q27, with the matrix transforms above.
__device__ __forceinline__ void collide_cm19_general(float* fv, float r, float ux, float uy, float uz,
                                                     const float F[3], bool fon, float os_base) {
#pragma clang fp contract(on)
  const float usq = ux * ux + uy * uy + uz * uz;
  const float b = (r - 1.0f) - r * 1.5f * usq;
  float feq[19], G[19], K[19], Sm[19];
  sfor<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float cu = cdot<q>(ux, uy, uz);
    feq[q] = lat_wf(19, q) * (b + r * cu * (3.0f + 4.5f * cu));
    G[q] = fv[q] - feq[q];
  });
  cm19_raw(G, Sm);
  cm19_shift<true>(Sm, K, ux, uy, uz);
  if (fon) {
    const float uf = ux * F[0] + uy * F[1] + uz * F[2];
    sfor<0, 19>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const float cu = cdot<q>(ux, uy, uz);
      const float cf = cdot<q>(F[0], F[1], F[2]);
      G[q] = lat_wf(19, q) * (3.0f * (cf - uf) + 9.0f * cu * cf);
    });
    float Rs[19];
    cm19_raw(G, Rs);
    cm19_shift<true>(Rs, Sm, ux, uy, uz);
  } else {
#pragma unroll
    for (int m = 0; m < 19; ++m) Sm[m] = 0.0f;
  }
  const float os = cm_rate(os_base, ux, uy, uz);
  const float a1 = 1.0f - os, a2 = 1.0f - 0.5f * os;
  constexpr int ixx = kCm19.idx[2][0][0], iyy = kCm19.idx[0][2][0], izz = kCm19.idx[0][0][2];
  const float kb = (K[ixx] + K[iyy] + K[izz]) * (1.0f / 3.0f);
  const float sb = (Sm[ixx] + Sm[iyy] + Sm[izz]) * (1.0f / 3.0f);
  float D[19];
  sfor<0, 19>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    constexpr int ex = kCm19.e[m][0], ey = kCm19.e[m][1], ez = kCm19.e[m][2];
    constexpr int order = ex + ey + ez;
    if constexpr (order <= 1)
      D[m] = K[m] + Sm[m];
    else if constexpr (order == 2 && (ex == 2 || ey == 2 || ez == 2))
      D[m] = a1 * (K[m] - kb) + 0.5f * sb + a2 * (Sm[m] - sb);
    else if constexpr (order == 2)
      D[m] = a1 * K[m] + a2 * Sm[m];
    else
      D[m] = 0.5f * Sm[m];
  });
  cm19_shift<false>(D, K, ux, uy, uz);
  sfor<0, 19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    float acc = feq[q];
    sfor<0, 19>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      constexpr float pc = (float)kCm19.P[q][m];
      if constexpr (pc != 0.0f) acc = acc + pc * K[m];
    });
    fv[q] = acc;
  });
}

// Force-free D3Q27 central-moment collision straight from the populations.
// The ten raw moments of order <= 2 of f - w (partial sums z, y, x) give rho
// and u, and -- because the second-order equilibrium reproduces those moments
// exactly (sum w = 1, sum w c_a c_b = delta/3, sum w c_a^2 c_b^2 = 1/9 for
// a != b, 1/3 for a == b) -- also the non-equilibrium raw moments:
//   N_0 = R_0 - (rho - 1),  N_a = R_a - rho u_a,
//   N_aa = R_aa - (rho - 1)/3 - rho u_a^2,  N_ab = R_ab - rho u_a u_b.
// Writing N_0 and N_a out (instead of dropping them as zero) keeps the
// rounding of rho and u exactly as mass- and momentum-conserving as the
// reference's f64 path.  Returns rho for the moving-wall bounce-back term.
__device__ __forceinline__ float collide_cm27_free_direct(float* fv, float os_base) {
#pragma clang fp contract(on)
  float Z0[9], Z1[9], Z2[9];
  sfor<0, 9>([&](auto abc) {
    constexpr int ab = decltype(abc)::value;
    constexpr int cx = ab / 3 - 1, cy = ab % 3 - 1;
    // populations of this (cx, cy) line with cz = -1, 0, +1
    constexpr int qm = [] {
      for (int q = 0; q < 27; ++q)
        if (lat_c(q, 0) == cx && lat_c(q, 1) == cy && lat_c(q, 2) == -1) return q;
      return -1;
    }();
    constexpr int q0 = [] {
      for (int q = 0; q < 27; ++q)
        if (lat_c(q, 0) == cx && lat_c(q, 1) == cy && lat_c(q, 2) == 0) return q;
      return -1;
    }();
    constexpr int qp = [] {
      for (int q = 0; q < 27; ++q)
        if (lat_c(q, 0) == cx && lat_c(q, 1) == cy && lat_c(q, 2) == 1) return q;
      return -1;
    }();
    Z2[ab] = fv[qp] + fv[qm];
    Z1[ab] = fv[qp] - fv[qm];
    Z0[ab] = Z2[ab] + fv[q0];
  });
  float Y00[3], Y10[3], Y20[3], Y01[3], Y11[3], Y02[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float* z0 = Z0 + 3 * a;
    const float* z1 = Z1 + 3 * a;
    const float* z2 = Z2 + 3 * a;
    Y20[a] = z0[2] + z0[0];
    Y10[a] = z0[2] - z0[0];
    Y00[a] = Y20[a] + z0[1];
    Y11[a] = z1[2] - z1[0];
    Y01[a] = (z1[2] + z1[0]) + z1[1];
    Y02[a] = (z2[2] + z2[0]) + z2[1];
  }
  const float R200 = Y00[2] + Y00[0], R100 = Y00[2] - Y00[0], R000 = R200 + Y00[1];
  const float R110 = Y10[2] - Y10[0], R010 = (Y10[2] + Y10[0]) + Y10[1];
  const float R020 = (Y20[2] + Y20[0]) + Y20[1];
  const float R101 = Y01[2] - Y01[0], R001 = (Y01[2] + Y01[0]) + Y01[1];
  const float R011 = (Y11[2] + Y11[0]) + Y11[1];
  const float R002 = (Y02[2] + Y02[0]) + Y02[1];
  const float r = 1.0f + R000;
  const float inv = 1.0f / r;
  const float ux = R100 * inv, uy = R010 * inv, uz = R001 * inv;
  const float dr = r - 1.0f;
  const float third_dr = dr * (1.0f / 3.0f);
  const float rux = r * ux, ruy = r * uy, ruz = r * uz;
  // non-equilibrium raw moments
  const float N000 = R000 - dr;
  const float N100 = R100 - rux, N010 = R010 - ruy, N001 = R001 - ruz;
  const float N200 = R200 - (third_dr + rux * ux);
  const float N020 = R020 - (third_dr + ruy * uy);
  const float N002 = R002 - (third_dr + ruz * uz);
  const float N110 = R110 - rux * uy, N101 = R101 - rux * uz, N011 = R011 - ruy * uz;
  // central moments of the non-equilibrium part
  const float K000 = N000;
  const float K100 = N100 - ux * N000, K010 = N010 - uy * N000, K001 = N001 - uz * N000;
  const float K200 = (N200 - 2.0f * ux * N100) + ux * ux * N000;
  const float K020 = (N020 - 2.0f * uy * N010) + uy * uy * N000;
  const float K002 = (N002 - 2.0f * uz * N001) + uz * uz * N000;
  const float K110 = ((N110 - ux * N010) - uy * N100) + ux * uy * N000;
  const float K101 = ((N101 - ux * N001) - uz * N100) + ux * uz * N000;
  const float K011 = ((N011 - uy * N001) - uz * N010) + uy * uz * N000;
  const float os = cm_rate(os_base, ux, uy, uz);
  const float a1 = 1.0f - os;
  const float kb = (K200 + K020 + K002) * (1.0f / 3.0f);
  const float d200 = a1 * (K200 - kb), d020 = a1 * (K020 - kb), d002 = a1 * (K002 - kb);
  const float d110 = a1 * K110, d101 = a1 * K101, d011 = a1 * K011;
  float X[27];
#pragma unroll
  for (int g = 0; g < 27; ++g) X[g] = 0.0f;
  const float ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
  auto inv3 = [](float k0, float k1, float k2, float u, float u2, float& gm, float& g0, float& gp) {
    const float m1 = k1 + u * k0;
    const float m2 = (k2 + 2.0f * u * k1) + u2 * k0;
    gm = 0.5f * (m2 - m1);
    g0 = k0 - m2;
    gp = 0.5f * (m1 + m2);
  };
  auto inv2 = [](float k0, float k1, float u, float u2, float& gm, float& g0, float& gp) {
    const float m1 = k1 + u * k0;
    const float m2 = 2.0f * u * k1 + u2 * k0;
    gm = 0.5f * (m2 - m1);
    g0 = k0 - m2;
    gp = 0.5f * (m1 + m2);
  };
  auto inv1 = [](float k0, float u, float u2, float& gm, float& g0, float& gp) {
    const float m1 = u * k0;
    const float m2 = u2 * k0;
    gm = 0.5f * (m2 - m1);
    g0 = k0 - m2;
    gp = 0.5f * (m1 + m2);
  };
  inv3(K000, K100, d200, ux, ux2, X[0], X[9], X[18]);
  inv2(K010, d110, ux, ux2, X[3], X[12], X[21]);
  inv2(K001, d101, ux, ux2, X[1], X[10], X[19]);
  inv1(d020, ux, ux2, X[6], X[15], X[24]);
  inv1(d002, ux, ux2, X[2], X[11], X[20]);
  inv1(d011, ux, ux2, X[4], X[13], X[22]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float* L = X + 9 * a;
    float g_m, g_0, g_p;
    inv3(L[0], L[3], L[6], uy, uy2, g_m, g_0, g_p);
    L[0] = g_m;
    L[3] = g_0;
    L[6] = g_p;
    inv2(L[1], L[4], uy, uy2, g_m, g_0, g_p);
    L[1] = g_m;
    L[4] = g_0;
    L[7] = g_p;
    inv1(L[2], uy, uy2, g_m, g_0, g_p);
    L[2] = g_m;
    L[5] = g_0;
    L[8] = g_p;
  }
#pragma unroll
  for (int ab = 0; ab < 9; ++ab) {
    float g_m, g_0, g_p;
    inv3(X[ab * 3 + 0], X[ab * 3 + 1], X[ab * 3 + 2], uz, uz2, g_m, g_0, g_p);
    X[ab * 3 + 0] = g_m;
    X[ab * 3 + 1] = g_0;
    X[ab * 3 + 2] = g_p;
  }
  // f' = feq + B(d).  Opposite directions share feq but for the sign of the
  // linear term: feq_{q, opp q} = s -+ t, s = w (b + 4.5 r cu^2), t = 3 w r cu.
  const float usq = ux2 + uy2 + uz2;
  const float b = dr - r * 1.5f * usq;
  const float r45 = 4.5f * r, r3 = 3.0f * r;
  fv[0] = lat_wf(27, 0) * b + X[gidx(0, 0, 0)];
  sfor<0, 13>([&](auto kc) {
    constexpr int qa = 2 * decltype(kc)::value + 1;
    constexpr int qb = qa + 1;
    constexpr int ga = gidx(lat_c(qa, 0), lat_c(qa, 1), lat_c(qa, 2));
    constexpr int gb = gidx(lat_c(qb, 0), lat_c(qb, 1), lat_c(qb, 2));
    constexpr float w = lat_wf(27, qa);
    const float cu = cdot<qa>(ux, uy, uz);
    const float sym = w * (b + r45 * (cu * cu));
    const float lin = (w * r3) * cu;
    fv[qa] = (sym + lin) + X[ga];
    fv[qb] = (sym - lin) + X[gb];
  });
  return r;
}

// ---------------------------------------------------------------------------
// one cell: load (parity), moments, collide, bounce-back-aware store
// ---------------------------------------------------------------------------
// Population slot addressing.  Plane bases are wave-uniform (scalar
// registers); the per-lane part of every address is one 32-bit byte offset,
// so the loads and stores use the scalar-base + 32-bit-offset form.  When no
// axis wraps, the neighbour shift is uniform as well and folds into the
// scalar base: all Q accesses of a cell share one offset register.
template <typename S, bool WRAP>
struct Slots {
  char* base;
  size_t plane_bytes;
  int64_t sy, sz;  // element strides of y and z
  uint32_t voff;   // i * sizeof(S)
  uint32_t i;
  Nbr nb;
  __device__ __forceinline__ Slots(const HydroArgs& a, uint32_t i_, uint32_t x, uint32_t y, uint32_t z)
      : base((char*)a.f),
        plane_bytes((size_t)a.d.ps * sizeof(S)),
        sy(a.d.nx),
        sz((int64_t)a.d.nx * a.d.ny),
        voff(i_ * (uint32_t)sizeof(S)),
        i(i_),
        nb(a.d, x, y, z) {}
  // slot of plane p at cell i + (cx, cy, cz)
  template <int p, int cx, int cy, int cz>
  __device__ __forceinline__ S* at() const {
    if constexpr (!WRAP) {
      const int64_t shift = ((int64_t)cx + cy * sy + cz * sz) * (int64_t)sizeof(S);
      return (S*)(base + p * plane_bytes + shift + voff);
    } else {
      const uint32_t o = (i + (uint32_t)nb.template off<cx, cy, cz>()) * (uint32_t)sizeof(S);
      return (S*)(base + p * plane_bytes + o);
    }
  }
};

template <typename S, int Q, int q>
__device__ __forceinline__ float slot_ld(const S* p) {
  float v = Store<S>::ld(p, 0);
  if constexpr (Store<S>::kScaled) v = v * lat_wf(Q, q);
  return v;
}
template <typename S, int Q, int q>
__device__ __forceinline__ void slot_st(S* p, float v) {
  if constexpr (Store<S>::kScaled) v = v * lat_inv_w(Q, q);
  Store<S>::st(p, 0, v);
}

// Bounce-back layout.  A population reflected off the wall at i + c_q is
// stored where the plain AA store of direction q would put it in the cell's
// own slots: even steps write slot opp(q) of cell i (the plain target), odd
// steps write slot opp(q) of cell i instead of slot q of i + c_q.  The odd
// step then reads a wall-link direction q (i - c_q solid) from the cell's
// own slot q, not from the solid upstream cell.  So no fluid cell ever reads
// or writes a solid cell's slots, and a solid cell inside an active tile can
// rewrite its own slots freely: even steps do, which turns every store of an
// active tile into whole cache lines (partial lines cost the memory a
// read-modify-write).
//
// PLAIN: the whole tile is fluid with mask word 0 (no wall link, no force
// list, no moving wall), known from the tile list; the mask is not read.
template <int Q, typename S, int KIND, bool ODD, bool FON, bool WRAP, bool FILL, bool PLAIN = false>
__device__ __forceinline__ void cell_step(const HydroArgs& a, uint32_t i, uint32_t x, uint32_t y, uint32_t z,
                                          const uint32_t* mask_ptr, uint32_t skip, const float Fm[3],
                                          const float Fc[3], uint32_t xlo = 0u, uint32_t xhi = 0u,
                                          uint32_t xr = 0xFFFFFFFFu) {
  constexpr bool fon = FON;
  const Slots<S, WRAP> sl(a, i, x, y, z);
  // The mask word and the Q populations are loaded together: the loads do
  // not wait for the solid test (the buffer's guard bands keep a solid face
  // cell's neighbour addresses inside the allocation); only the stores are
  // predicated on it.
  const uint32_t m = PLAIN ? 0u : *mask_ptr;
  float fv[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = lat_opp(q);
    if constexpr (!ODD)
      fv[q] = slot_ld<S, Q, q>(sl.template at<q, 0, 0, 0>());
    else
      fv[q] = slot_ld<S, Q, q>(sl.template at<o, -lat_c(q, 0), -lat_c(q, 1), -lat_c(q, 2)>());
  });
  // Consume every loaded value before the solid test, so the compiler cannot
  // sink the population loads behind the mask branch: one memory latency per
  // cell instead of two.
#pragma unroll
  for (int q = 0; q < Q; ++q) asm volatile("" : "+v"(fv[q]));
  // A solid cell of an active tile runs along with its wave and stores zeros
  // into dead slots (even steps: its own; odd steps, D3Q27 with fp32
  // storage: those of its targets that no fluid cell writes), so the tile's
  // stores cover whole cache lines.  With 16-bit storage the odd-step fill
  // speeds up the next even step about as much as it slows the odd one, and
  // on D3Q19 (the batched tanks) it measured 1.3 % slower, so both keep the
  // plain odd step.
  constexpr bool kOddFill = sizeof(S) == 4 && Q == 27;
  const bool dead = FILL && (!ODD || kOddFill) && (m & kMaskSolid) != 0;
  if (m & skip & (dead ? ~kMaskSolid : ~0u)) return;
  if constexpr (ODD) {
    if (!dead && (m & kMaskLinks)) {  // wall links: the bounced population is in the own slot
      // Every reload is issued before any is used: a 16-bit value converted
      // inside its own conditional made each load wait for the one before
      // (one memory latency per link direction).  So each direction issues
      // one unconditional wave-wide load: the own slot q when some lane of
      // the wave needs it, else the own slot 0 the step has just read (an L1
      // hit); the choice is wave-uniform, no branch, no merge to wait for.
      S raw[Q];
      sfor<1, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const bool need = __builtin_amdgcn_ballot_w64(((m >> q) & 1u) != 0u) != 0u;
        raw[q] = *(need ? sl.template at<q, 0, 0, 0>() : sl.template at<0, 0, 0, 0>());
      });
      sfor<1, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if ((m >> q) & 1u) {
          float v = Store<S>::ld(&raw[q], 0);
          if constexpr (Store<S>::kScaled) v = v * lat_wf(Q, q);
          fv[q] = v;
        }
      });
    }
  }
  float op = a.p.omega_p, cp = a.p.cp, os = a.p.omega_shear;
  if (a.omega) {  // per-cell omega_plus (kernels.rs:151-156)
    op = a.omega[i];
    cp = 1.0f - op / 2.0f;
    os = op;
  }
  float r;
  if constexpr (KIND == 0) {
    float ux, uy, uz;
    moments<Q>(fv, Fm, fon, r, ux, uy, uz);
    collide_trt<Q>(fv, r, ux, uy, uz, Fc, fon, op, cp, a.p);
  } else if constexpr (Q == 19) {
    float ux, uy, uz;
    moments<Q>(fv, Fm, fon, r, ux, uy, uz);
    collide_cm19_general(fv, r, ux, uy, uz, Fc, fon, os);
  } else {
    static_assert(Q == 27, "central-moment kernels are D3Q19 or D3Q27");
    if constexpr (FON) {
      float ux, uy, uz;
      moments<Q>(fv, Fm, fon, r, ux, uy, uz);
      collide_cm27_general(fv, r, ux, uy, uz, Fc, true, os);
    } else {
      r = collide_cm27_free_direct(fv, os);
    }
  }
  // half-way bounce-back off a moving wall at i + c_q (stream_row,
  // kernels.rs:591-603): the reflected population enters i along c_o with
  // fin = fout + 6 w_o rho (c_o . u_wall)
  if (m & kMaskMoving) {
    sfor<1, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int o = lat_opp(q);
      if ((m >> o) & 1u) {
        // (no wrapping axis: a fluid cell's neighbours are inside the domain,
        // so the linear offset is the wrapped one)
        const uint32_t fwd =
            WRAP ? i + (uint32_t)sl.nb.template off<lat_c(q, 0), lat_c(q, 1), lat_c(q, 2)>()
                 : i + (uint32_t)(lat_c(q, 0) + lat_c(q, 1) * (int32_t)a.d.nx + lat_c(q, 2) * (int32_t)(a.d.nx * a.d.ny));
        const float* wu = a.wall_u + 3 * (size_t)fwd;
        const float cu = cdot<o>(wu[0], wu[1], wu[2]);
        constexpr float six_w = 6.0f * lat_wf(Q, o);
        fv[q] = fv[q] + six_w * r * cu;
      }
    });
  }
  // plain streaming store of direction q
  auto store_plain = [&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = lat_opp(q);
    if constexpr (!ODD)
      slot_st<S, Q, q>(sl.template at<o, 0, 0, 0>(), dead ? 0.0f : fv[q]);
    else
      slot_st<S, Q, q>(sl.template at<q, lat_c(q, 0), lat_c(q, 1), lat_c(q, 2)>(), fv[q]);
  };
  // even steps: the bounced population goes to the plain target as well
  if constexpr (!ODD || PLAIN) {
    sfor<0, Q>(store_plain);
  } else if constexpr (!FILL || !kOddFill) {
    // one exec-masked pair of stores per direction for the whole wave
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int o = lat_opp(q);
      if ((m >> o) & 1u)
        slot_st<S, Q, q>(sl.template at<o, 0, 0, 0>(), fv[q]);
      else
        store_plain(qc);
    });
  } else {
    // Odd steps of an active fp32 tile.  A wall-link lane stores the bounced
    // population in its own slot; a solid lane writes zeros to each of its
    // targets that no fluid cell writes (mask bit q clear), only inside the
    // tile's live x range [xlo, xhi], so no store reaches a line outside the
    // tile's own.  (Also writing a link lane's plain target, a dead slot,
    // measured slower.)
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int o = lat_opp(q);
      constexpr int cx = lat_c(q, 0);
      // xr: the coordinate the live range is in (x for row tiles, the cell
      // index for linear tiles)
      const uint32_t xc = xr == 0xFFFFFFFFu ? x : xr;
      const bool inr = cx == 0 || ((uint32_t)(xc + cx) >= xlo && (uint32_t)(xc + cx) <= xhi);
      const bool link = !dead && ((m >> o) & 1u);
      const bool plain = dead ? (!((m >> q) & 1u) && inr) : !link;
      if (link) slot_st<S, Q, q>(sl.template at<o, 0, 0, 0>(), fv[q]);
      if (plain)
        slot_st<S, Q, q>(sl.template at<q, lat_c(q, 0), lat_c(q, 1), lat_c(q, 2)>(), dead ? 0.0f : fv[q]);
    });
  }
}

template <int Q, typename S, int KIND, bool ODD, bool FON, bool WRAP>
__global__ __launch_bounds__(256) void k_hydro(HydroArgs a) {
  // each wave takes one active 64-cell tile; tiles of only solid (or
  // force-listed) cells are never launched
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t t = b * 4 + (threadIdx.x >> 6);
  if (t >= a.ntiles) return;
  const uint32_t te = __builtin_amdgcn_readfirstlane(a.tiles[t]);  // one tile per wave
  const uint32_t tile = te & kTileId;
  if (!WRAP && a.lin) {
    // linear tile: cells 64 tile .. 64 tile + 63 across row ends (no axis
    // wraps, so every neighbour shift is the linear one)
    const uint32_t i = tile * 64 + (threadIdx.x & 63);
    if (i >= a.d.N || (te & ((threadIdx.x & 32) ? kTileDeadHi : kTileDeadLo))) return;
    if (te & kTilePlain) {
      cell_step<Q, S, KIND, ODD, FON, WRAP, true, true>(a, i, 0u, 0u, 0u, nullptr, 0u, a.p.F, a.p.F);
    } else {
      const uint32_t c0 = tile * 64;
      const uint32_t lo = (te & kTileDeadLo) ? c0 + 32 : c0;
      const uint32_t hi = min((te & kTileDeadHi) ? c0 + 31 : c0 + 63, a.d.N - 1);
      cell_step<Q, S, KIND, ODD, FON, WRAP, true, false>(a, i, 0u, 0u, 0u, a.mask + i, kMaskSolid | kMaskForced,
                                                         a.p.F, a.p.F, lo, hi, i);
    }
    return;
  }
  const uint32_t row = tile / a.nseg, seg = tile - row * a.nseg;
  const uint32_t x = seg * 64 + (threadIdx.x & 63);
  if (x >= a.d.nx || (te & ((threadIdx.x & 32) ? kTileDeadHi : kTileDeadLo))) return;
  const uint32_t y = row % a.d.ny, z = row / a.d.ny;
  const uint32_t i = row * a.d.nx + x;
  if (te & kTilePlain) {
    cell_step<Q, S, KIND, ODD, FON, WRAP, true, true>(a, i, x, y, z, nullptr, 0u, a.p.F, a.p.F);
  } else {
    // the tile's live x range (an all-solid half is not visited)
    const uint32_t x0 = seg * 64;
    const uint32_t xlo = (te & kTileDeadLo) ? x0 + 32 : x0;
    const uint32_t xhi = min((te & kTileDeadHi) ? x0 + 31 : x0 + 63, a.d.nx - 1);
    cell_step<Q, S, KIND, ODD, FON, WRAP, true, false>(a, i, x, y, z, a.mask + i, kMaskSolid | kMaskForced,
                                                       a.p.F, a.p.F, xlo, xhi);
  }
}

// Cells carrying a per-cell force: the velocity uses the force of the cached
// moments (F_mom, set when moments were last refreshed), the Guo source the
// current field (F_cur), as the reference's stale-moment step does
// (solver.rs:2805-2822 keeps moments across a non-uniform clear).  After the
// step the cached moments carry F_cur.
template <int Q, typename S, int KIND, bool ODD>
__device__ __forceinline__ void list_cell(const HydroArgs& a, uint32_t i) {
  const uint32_t m = a.mask[i];
  if (m & kMaskSolid) return;
  const uint32_t x = i % a.d.nx, t = i / a.d.nx;
  const uint32_t y = t % a.d.ny, z = t / a.d.ny;
  float Fm[3], Fc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Fm[c] = a.p.F[c] + a.F_mom[3 * (size_t)i + c];
    Fc[c] = a.p.F[c] + a.F_cur[3 * (size_t)i + c];
  }
  cell_step<Q, S, KIND, ODD, true, true, false>(a, i, x, y, z, a.mask + i, kMaskSolid, Fm, Fc);
#pragma unroll
  for (int c = 0; c < 3; ++c) a.F_mom[3 * (size_t)i + c] = a.F_cur[3 * (size_t)i + c];
}

template <int Q, typename S, int KIND, bool ODD>
__global__ __launch_bounds__(256) void k_hydro_list(HydroArgs a) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nlist) return;
  list_cell<Q, S, KIND, ODD>(a, a.flist[k]);
}

// ---------------------------------------------------------------------------
// Batched domains: one persistent launch walks the concatenated tile lists
// of K independent (non-periodic) domains; entry e = (tile, domain).  Each
// wave strides over entries and reads its tile's domain arguments from the
// table (wave-uniform, so scalar loads).  Consecutive entries share an XCD.
// ---------------------------------------------------------------------------
template <int Q, typename S, int KIND, bool ODD>
__global__ __launch_bounds__(256) void k_hydro_batch(const HydroArgs* __restrict__ tab,
                                                     const uint32_t* __restrict__ ent, uint32_t nent) {
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t stride = gridDim.x * 4;
  for (uint32_t e = b * 4 + (threadIdx.x >> 6); e < nent; e += stride) {
    const uint32_t te = __builtin_amdgcn_readfirstlane(ent[2 * (size_t)e]);
    const uint32_t k = __builtin_amdgcn_readfirstlane(ent[2 * (size_t)e + 1]);
    const HydroArgs& a = tab[k];
    const uint32_t tile = te & kTileId;
    if (a.lin) {  // linear tile (members never wrap)
      const uint32_t i = tile * 64 + (threadIdx.x & 63);
      if (i < a.d.N && !(te & ((threadIdx.x & 32) ? kTileDeadHi : kTileDeadLo))) {
        if (te & kTilePlain) {
          cell_step<Q, S, KIND, ODD, false, false, true, true>(a, i, 0u, 0u, 0u, nullptr, 0u, a.p.F, a.p.F);
        } else {
          const uint32_t c0 = tile * 64;
          const uint32_t lo = (te & kTileDeadLo) ? c0 + 32 : c0;
          const uint32_t hi = min((te & kTileDeadHi) ? c0 + 31 : c0 + 63, a.d.N - 1);
          cell_step<Q, S, KIND, ODD, false, false, true, false>(a, i, 0u, 0u, 0u, a.mask + i,
                                                                kMaskSolid | kMaskForced, a.p.F, a.p.F, lo, hi, i);
        }
      }
      continue;
    }
    const uint32_t row = tile / a.nseg, seg = tile - row * a.nseg;
    const uint32_t x = seg * 64 + (threadIdx.x & 63);
    if (x < a.d.nx && !(te & ((threadIdx.x & 32) ? kTileDeadHi : kTileDeadLo))) {
      const uint32_t y = row % a.d.ny, z = row / a.d.ny;
      const uint32_t i = row * a.d.nx + x;
      if (te & kTilePlain) {
        cell_step<Q, S, KIND, ODD, false, false, true, true>(a, i, x, y, z, nullptr, 0u, a.p.F, a.p.F);
      } else {
        const uint32_t x0 = seg * 64;
        const uint32_t xlo = (te & kTileDeadLo) ? x0 + 32 : x0;
        const uint32_t xhi = min((te & kTileDeadHi) ? x0 + 31 : x0 + 63, a.d.nx - 1);
        cell_step<Q, S, KIND, ODD, false, false, true, false>(a, i, x, y, z, a.mask + i, kMaskSolid | kMaskForced,
                                                              a.p.F, a.p.F, xlo, xhi);
      }
    }
  }
}

// forced cells of all domains: entry = (cell, domain)
template <int Q, typename S, int KIND, bool ODD>
__global__ __launch_bounds__(256) void k_hydro_list_batch(const HydroArgs* __restrict__ tab,
                                                          const uint32_t* __restrict__ cells, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  list_cell<Q, S, KIND, ODD>(tab[cells[2 * (size_t)k + 1]], cells[2 * (size_t)k]);
}

// clear the force field at the listed cells of all domains
__global__ void k_clear_batch(const HydroArgs* __restrict__ tab, const uint32_t* __restrict__ cells, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float* F = const_cast<float*>(tab[cells[2 * (size_t)k + 1]].F_cur);
  const size_t i = cells[2 * (size_t)k];
  F[3 * i] = 0.0f;
  F[3 * i + 1] = 0.0f;
  F[3 * i + 2] = 0.0f;
}

// moments readback (moments_row with the cached-moment force F_u + F_mom)
template <int Q, typename S>
__global__ void k_moments(Dom d, const S* f, const uint32_t* mask, const float* F_mom, HydroParams p, int odd,
                          float* rho, float* ux, float* uy, float* uz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  const uint32_t m = mask[i];
  if (m & kMaskSolid) {
    rho[i] = 1.0f;
    ux[i] = uy[i] = uz[i] = 0.0f;
    return;
  }
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  const Nbr nb(d, x, y, z);
  float fv[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    fv[q] = read_pop<Q, S, q>(f, d.ps, nb, i, m, odd != 0);
  });
  // the field is zero off the force list, whose fluid cells carry kMaskForced
  float Fm[3] = {p.F[0], p.F[1], p.F[2]};
  bool fon = p.force_on != 0;
  if (F_mom && (m & kMaskForced)) {
    for (int c = 0; c < 3; ++c) Fm[c] = p.F[c] + F_mom[3 * (size_t)i + c];
    fon = true;
  }
  float r, a, b, c;
  moments<Q>(fv, Fm, fon, r, a, b, c);
  rho[i] = r;
  ux[i] = a;
  uy[i] = b;
  uz[i] = c;
}

// stage in (dir 0: host layout -> device, parity reset to even) /
// stage out (dir 1: device -> host layout, de-permuted by parity).
template <int Q, typename S>
__global__ void k_stage(Dom d, S* f, float* h, const uint32_t* mask, int dir, int odd) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  if (dir == 0) {
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      pop_st<S, Q, q>(f + (size_t)q * d.ps, i, h[(size_t)q * d.N + i]);
    });
    return;
  }
  const uint32_t m = mask[i];
  const bool solid = (m & kMaskSolid) != 0;
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  const Nbr nb(d, x, y, z);
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    h[(size_t)q * d.N + i] = solid ? 0.0f : read_pop<Q, S, q>(f, d.ps, nb, i, m, odd != 0);
  });
}

// Solver::init_with, solver.rs:2474-2623 (pass 2): feq of every cell plus
// the gradient non-equilibrium part on fluid cells.
template <int Q, typename S>
__global__ void k_init_with(Dom d, S* f, const uint32_t* mask, const float* rho, const float* u, float tau) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  const float r = rho[i];
  const float own[3] = {u[3 * (size_t)i], u[3 * (size_t)i + 1], u[3 * (size_t)i + 2]};
  // equilibrium(), kernels.rs:103-121
  const float three = 3.0f, f45 = 4.5f, f15 = 1.5f, half = 0.5f;
  float usq = own[0] * own[0];
  usq = usq + own[1] * own[1];
  usq = usq + own[2] * own[2];
  const float drho = r - 1.0f;
  float fv[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float cu = cdot<q>(own[0], own[1], own[2]);
    fv[q] = lat_wf(Q, q) * (drho + r * (three * cu + f45 * cu * cu - f15 * usq));
  });
  const bool solid = (mask[i] & kMaskSolid) != 0;
  if (!solid) {
    const uint32_t n3[3] = {d.nx, d.ny, d.nz};
    const uint32_t c3[3] = {x, y, z};
    float grad[3][3];
    for (int ax = 0; ax < 3; ++ax) {
      float pv[3], mv[3];
      bool hp = false, hm = false;
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        int64_t g[3] = {(int64_t)x, (int64_t)y, (int64_t)z};
        g[ax] += sgn;
        bool ok = true;
        if (g[ax] < 0 || g[ax] >= (int64_t)n3[ax]) {
          if (d.per[ax])
            g[ax] = (g[ax] + n3[ax]) % n3[ax];
          else
            ok = false;
        }
        if (ok) {
          const uint32_t j = (uint32_t)((g[2] * d.ny + g[1]) * d.nx + g[0]);
          if (mask[j] & kMaskSolid) ok = false;
          if (ok) {
            float* dst = sgn > 0 ? pv : mv;
            dst[0] = u[3 * (size_t)j];
            dst[1] = u[3 * (size_t)j + 1];
            dst[2] = u[3 * (size_t)j + 2];
          }
        }
        (sgn > 0 ? hp : hm) = ok;
      }
      (void)c3;
      for (int b = 0; b < 3; ++b) {
        if (hp && hm)
          grad[ax][b] = (pv[b] - mv[b]) * half;
        else if (hp)
          grad[ax][b] = pv[b] - own[b];
        else if (hm)
          grad[ax][b] = own[b] - mv[b];
        else
          grad[ax][b] = 0.0f;
      }
    }
    float div = grad[0][0];
    div = div + grad[1][1];
    div = div + grad[2][2];
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr float c0 = lat_c(q, 0), c1 = lat_c(q, 1), c2 = lat_c(q, 2);
      float ccgu = c0 * c0 * grad[0][0];
      ccgu = ccgu + c0 * c1 * (grad[0][1] + grad[1][0]);
      ccgu = ccgu + c0 * c2 * (grad[0][2] + grad[2][0]);
      ccgu = ccgu + c1 * c1 * grad[1][1];
      ccgu = ccgu + c1 * c2 * (grad[1][2] + grad[2][1]);
      ccgu = ccgu + c2 * c2 * grad[2][2];
      const float fneq = -lat_wf(Q, q) * r * tau * (three * ccgu - div);
      fv[q] = fv[q] + fneq;
    });
  }
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    pop_st<S, Q, q>(f + (size_t)q * d.ps, i, fv[q]);
  });
}

// f64 reductions (reduce_impl, backend.rs:920-970): one partial per block,
// fixed tree inside the block, partials summed in block order on the host.
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_reduce(Dom d, const S* f, const uint32_t* mask, const float* F_cur,
                                                HydroParams p, int odd, int kind, int axis, double* partials) {
  __shared__ double sm[4];
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < d.N; i += gridDim.x * 256) {
    const uint32_t m = mask[i];
    if (m & kMaskSolid) continue;
    if (kind == 0) {
      acc += 1.0;
      continue;
    }
    uint32_t x, y, z;
    cell_xyz(d, i, x, y, z);
    const Nbr nb(d, x, y, z);
    double s = 0.0, mom = 0.0;
    sfor<0, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const double v = (double)read_pop<Q, S, q>(f, d.ps, nb, i, m, odd != 0);
      s += v;
      if (kind == 2) {
        const int c = axis == 0 ? lat_c(q, 0) : (axis == 1 ? lat_c(q, 1) : lat_c(q, 2));
        mom += (double)c * v;
      }
    });
    if (kind == 1) {
      acc += s;
    } else {
      double fa = (double)p.F[axis];
      if (F_cur) fa = fa + ((double)F_cur[3 * (size_t)i + axis] + 0.0);
      acc += mom + 0.5 * fa;
    }
  }
  block_sum256<1>(&acc, sm);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// The previous step's post-collision populations f*_q(j) (what the
// reference's second buffer holds between steps: fields.rs ftmp, the
// checkpoint's STALE_STASH section), recovered from the current state by
// undoing the streaming: f*_q(j) = f_q(j + c_q) when j + c_q is fluid; the
// reflected population f_opp(q)(j), less the moving-wall term 6 w rho
// (c_opp . u_wall) (rho of the current state), when j + c_q is solid; 0 for
// solid cells and where j + c_q lies beyond a non-periodic face (a slot the
// next stream overwrites).  Exact copies on resting walls.
template <int Q, typename S>
__global__ void k_poststate(Dom d, const S* f, const uint32_t* mask, const float* wall_u, int odd, float* out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.N) return;
  const uint32_t m = mask[j];
  if (m & kMaskSolid) {
    for (int q = 0; q < Q; ++q) out[(size_t)q * d.N + j] = 0.0f;
    return;
  }
  uint32_t x, y, z;
  cell_xyz(d, j, x, y, z);
  const Nbr nb(d, x, y, z);
  float rho = 1.0f;
  if (wall_u && (m & kMaskMoving)) {
    float dr = 0.0f;
    sfor<0, Q>([&](auto qc) { dr = dr + read_pop<Q, S, decltype(qc)::value>(f, d.ps, nb, j, m, odd != 0); });
    rho = 1.0f + dr;
  }
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = lat_opp(q);
    constexpr int cx = lat_c(q, 0), cy = lat_c(q, 1), cz = lat_c(q, 2);
    float v = 0.0f;
    if (q != 0 && ((m >> o) & 1u)) {  // j + c_q solid: reflected into f_opp(j)
      v = read_pop<Q, S, o>(f, d.ps, nb, j, m, odd != 0);
      if (wall_u && (m & kMaskMoving)) {
        const uint32_t t = j + (uint32_t)nb.template off<cx, cy, cz>();
        const float* wu = wall_u + 3 * (size_t)t;
        v = v - 6.0f * lat_wf(Q, o) * rho * cdot<o>(wu[0], wu[1], wu[2]);
      }
    } else {
      const int64_t tx = (int64_t)x + cx, ty = (int64_t)y + cy, tz = (int64_t)z + cz;
      const bool out_x = (tx < 0 || tx >= (int64_t)d.nx) && !d.per[0];
      const bool out_y = (ty < 0 || ty >= (int64_t)d.ny) && !d.per[1];
      const bool out_z = (tz < 0 || tz >= (int64_t)d.nz) && !d.per[2];
      if (!(out_x || out_y || out_z)) {
        const uint32_t t = j + (uint32_t)nb.template off<cx, cy, cz>();
        uint32_t ux, uy, uz;
        cell_xyz(d, t, ux, uy, uz);
        const Nbr nt(d, ux, uy, uz);
        v = read_pop<Q, S, q>(f, d.ps, nt, t, mask[t], odd != 0);
      }
    }
    out[(size_t)q * d.N + j] = v;
  });
}

// masks: links in the pull convention plus cell flags
template <int Q>
__global__ void k_build_mask(Dom d, const uint8_t* solid, const float* wall_u, uint32_t* mask,
                             unsigned int* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  if (solid[i]) {
    // bit q: the odd-step store target of direction q (i + c_q, wrapped or
    // linear, as either addressing of the collide-stream kernel reaches it)
    // is a fluid cell, which writes that slot itself
    const Nbr nb(d, x, y, z);
    uint32_t m = kMaskSolid;
    sfor<1, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = lat_c(q, 0), cy = lat_c(q, 1), cz = lat_c(q, 2);
      const uint32_t jw = i + (uint32_t)nb.template off<cx, cy, cz>();
      const int64_t jl = (int64_t)i + cx + cy * (int64_t)d.nx + cz * (int64_t)d.nx * d.ny;
      const bool fluid_target = !solid[jw] || (jl >= 0 && jl < (int64_t)d.N && !solid[jl]);
      if (fluid_target) m |= 1u << q;
    });
    mask[i] = m;
    return;
  }
  const bool edge = (!d.per[0] && (x == 0 || x == d.nx - 1)) || (!d.per[1] && (y == 0 || y == d.ny - 1)) ||
                    (!d.per[2] && (z == 0 || z == d.nz - 1));
  if (edge) atomicAdd(err, 1u);
  const Nbr nb(d, x, y, z);
  uint32_t m = 0;
  sfor<1, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const uint32_t s = i + (uint32_t)nb.template off<-lat_c(q, 0), -lat_c(q, 1), -lat_c(q, 2)>();
    if (solid[s]) {
      m |= 1u << q;
      if (wall_u && (wall_u[3 * (size_t)s] != 0.0f || wall_u[3 * (size_t)s + 1] != 0.0f ||
                     wall_u[3 * (size_t)s + 2] != 0.0f))
        m |= kMaskMoving;
    }
  });
  mask[i] = m;
}

// flags per 64-cell x-segment: bit 0 some cell for the main kernel, bit 1
// all 64 cells plain fluid (mask word 0), bits 2/3 lanes 0-31 / 32-63 all
// solid (or past the row end)
// Linear tiles (`linear`): tile t is cells 64t .. 64t + 63 of the linear
// index, across row ends, with the same four flags.
__global__ void k_tile_flags(Dom d, const uint32_t* mask, uint8_t* flags, int linear) {
  const uint32_t nseg = (d.nx + 63) / 64;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c0, c1;
  if (linear) {
    if (t >= (d.N + 63) / 64) return;
    c0 = t * 64;
    c1 = min(c0 + 64, d.N);
  } else {
    if (t >= nseg * d.ny * d.nz) return;
    const uint32_t row = t / nseg, seg = t - row * nseg;
    c0 = row * d.nx + seg * 64;
    c1 = row * d.nx + min(seg * 64 + 64, d.nx);
  }
  bool any = false, plain = c1 - c0 == 64, dead[2] = {true, true};
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t m = mask[c];
    if (!(m & (kMaskSolid | kMaskForced))) any = true;
    if (m != 0u) plain = false;
    if (!(m & kMaskSolid)) dead[(c - c0) >> 5] = false;
  }
  flags[t] = (uint8_t)((any ? 1 : 0) | (plain ? 2 : 0) | (dead[0] ? 4 : 0) | (dead[1] ? 8 : 0));
}

__global__ void k_mark_forced(const uint32_t* flist, uint32_t n, uint32_t* mask) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = flist[k];
  if (!(mask[i] & kMaskSolid)) mask[i] |= kMaskForced;
}

// apply_cylindrical_tank_mask + apply_baffle_mask (geometry.rs:657-723),
// f64 with the same expressions (no contraction in this file).
__global__ void k_tank(Dom d, double dx, double diameter, double height, int baffles, double bw, double bt,
                       uint8_t* solid) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  const double radius = 0.5 * diameter;
  const double cx = 0.5 * (double)d.nx * dx, cy = 0.5 * (double)d.ny * dx;
  const double z_m = ((double)z + 0.5) * dx;
  const double pxm = ((double)x + 0.5) * dx, pym = ((double)y + 0.5) * dx;
  const double r = sqrt((pxm - cx) * (pxm - cx) + (pym - cy) * (pym - cy));
  const bool rim = x == 0 || y == 0 || z == 0 || x + 1 == d.nx || y + 1 == d.ny || z + 1 == d.nz;
  const bool outside = r >= radius || z_m >= height;
  bool s = rim || outside;
  if (baffles == 4 && !rim && !(z_m >= height)) {
    const double rx = pxm - cx, ry = pym - cy;
    const double rb = sqrt(rx * rx + ry * ry);
    const double radial_inner = radius - bw;
    if (!(rb >= radius || rb < radial_inner)) {
      if (fabs(ry) <= 0.5 * bt || fabs(rx) <= 0.5 * bt) s = true;
    }
  }
  if (s) solid[i] = 1;
}

// build_wall_rims, solver.rs:1357-1406: one launch per face in the
// reference's order; the faster wall keeps the corner.
__global__ void k_rim_face(Dom d, int face, double speed, float u0, float u1, float u2, uint8_t* solid,
                           float* wall_u, double* best) {
  const int a = face / 2;
  const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
  const uint32_t n3[3] = {d.nx, d.ny, d.nz};
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n3[t1] * n3[t2]) return;
  uint32_t pos[3];
  pos[a] = (face % 2 == 0) ? 0 : n3[a] - 1;
  pos[t1] = k % n3[t1];
  pos[t2] = k / n3[t1];
  const uint32_t i = (pos[2] * d.ny + pos[1]) * d.nx + pos[0];
  solid[i] = 1;
  // best[] holds the squared speed of the velocity stored so far (-1 = none)
  if (speed > best[i]) {
    best[i] = speed;
    wall_u[3 * (size_t)i + 0] = u0;
    wall_u[3 * (size_t)i + 1] = u1;
    wall_u[3 * (size_t)i + 2] = u2;
  }
}

__global__ void k_fill(float* p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_copy_list(const uint32_t* flist, uint32_t n, float* dst, const float* src) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t i = flist[k];
  dst[3 * i] = src ? src[3 * i] : 0.0f;
  dst[3 * i + 1] = src ? src[3 * i + 1] : 0.0f;
  dst[3 * i + 2] = src ? src[3 * i + 2] : 0.0f;
}

// STREAM copy (b[i] = a[i]) with 16-byte accesses: the measured HBM ceiling
// and the calibration run for the byte counters.
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) b[i] = a[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
Tiling make_tiling(const Dom& d, uint32_t tx_pref) {
  Tiling t;
  uint32_t tx = tx_pref ? tx_pref : 256;
  while (tx > 64 && tx / 2 >= d.nx) tx /= 2;
  t.tx = tx;
  t.ty = 256 / tx;
  t.bx = (d.nx + t.tx - 1) / t.tx;
  t.by = (d.ny + t.ty - 1) / t.ty;
  t.nblocks = t.bx * t.by * d.nz;
  return t;
}

static inline uint32_t blocks_for(size_t n, uint32_t bs = 256) { return (uint32_t)((n + bs - 1) / bs); }

template <int Q, typename S, int KIND>
static hipError_t hydro_q(int odd, const HydroArgs& a, hipStream_t s, bool list) {
  if (list) {
    if (a.nlist == 0) return hipSuccess;
    const dim3 grid(blocks_for(a.nlist)), block(256);
    if (odd)
      hipLaunchKernelGGL((k_hydro_list<Q, S, KIND, true>), grid, block, 0, s, a);
    else
      hipLaunchKernelGGL((k_hydro_list<Q, S, KIND, false>), grid, block, 0, s, a);
    return hipGetLastError();
  }
  if (a.ntiles == 0) return hipSuccess;
  const dim3 grid((a.ntiles + 3) / 4), block(256);
  const bool fon = a.p.force_on != 0;
  const bool wrap = a.d.per[0] || a.d.per[1] || a.d.per[2] || a.open;
  const int sel = (odd ? 4 : 0) | (fon ? 2 : 0) | (wrap ? 1 : 0);
  switch (sel) {
    case 0: hipLaunchKernelGGL((k_hydro<Q, S, KIND, false, false, false>), grid, block, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_hydro<Q, S, KIND, false, false, true>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_hydro<Q, S, KIND, false, true, false>), grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_hydro<Q, S, KIND, false, true, true>), grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_hydro<Q, S, KIND, true, false, false>), grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_hydro<Q, S, KIND, true, false, true>), grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_hydro<Q, S, KIND, true, true, false>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((k_hydro<Q, S, KIND, true, true, true>), grid, block, 0, s, a); break;
  }
  return hipGetLastError();
}

template <typename S>
static hipError_t hydro_s(int Q, int collision, int odd, const HydroArgs& a, hipStream_t s, bool list) {
  const int kind = collision == 2 ? 1 : 0;
  if (Q == 19 && kind == 0) return hydro_q<19, S, 0>(odd, a, s, list);
  if (Q == 27 && kind == 0) return hydro_q<27, S, 0>(odd, a, s, list);
  if (Q == 27 && kind == 1) return hydro_q<27, S, 1>(odd, a, s, list);
  if (Q == 19 && kind == 1) return hydro_q<19, S, 1>(odd, a, s, list);
  return hipErrorInvalidValue;
}

hipError_t launch_hydro(int Q, int storage, int collision, int odd, const HydroArgs& a, hipStream_t s) {
  if (storage == kF32) return hydro_s<float>(Q, collision, odd, a, s, false);
  if (storage == kF16) return hydro_s<_Float16>(Q, collision, odd, a, s, false);
  return hydro_s<bf16s>(Q, collision, odd, a, s, false);
}

hipError_t launch_hydro_list(int Q, int storage, int collision, int odd, const HydroArgs& a, hipStream_t s) {
  if (storage == kF32) return hydro_s<float>(Q, collision, odd, a, s, true);
  if (storage == kF16) return hydro_s<_Float16>(Q, collision, odd, a, s, true);
  return hydro_s<bf16s>(Q, collision, odd, a, s, true);
}

template <int Q, typename S, int KIND>
static hipError_t batch_q(int odd, const HydroArgs* tab, const uint32_t* ent, uint32_t nent, uint32_t grid,
                          const uint32_t* cells, uint32_t ncells, hipStream_t s) {
  if (ent) {
    if (nent == 0) return hipSuccess;
    if (odd)
      hipLaunchKernelGGL((k_hydro_batch<Q, S, KIND, true>), dim3(grid), dim3(256), 0, s, tab, ent, nent);
    else
      hipLaunchKernelGGL((k_hydro_batch<Q, S, KIND, false>), dim3(grid), dim3(256), 0, s, tab, ent, nent);
  } else {
    if (ncells == 0) return hipSuccess;
    if (odd)
      hipLaunchKernelGGL((k_hydro_list_batch<Q, S, KIND, true>), dim3(blocks_for(ncells)), dim3(256), 0, s, tab,
                         cells, ncells);
    else
      hipLaunchKernelGGL((k_hydro_list_batch<Q, S, KIND, false>), dim3(blocks_for(ncells)), dim3(256), 0, s, tab,
                         cells, ncells);
  }
  return hipGetLastError();
}

template <typename S>
static hipError_t batch_s(int Q, int collision, int odd, const HydroArgs* tab, const uint32_t* ent, uint32_t nent,
                          uint32_t grid, const uint32_t* cells, uint32_t ncells, hipStream_t s) {
  const int kind = collision == 2 ? 1 : 0;
  if (Q == 19 && kind == 0) return batch_q<19, S, 0>(odd, tab, ent, nent, grid, cells, ncells, s);
  if (Q == 27 && kind == 0) return batch_q<27, S, 0>(odd, tab, ent, nent, grid, cells, ncells, s);
  if (Q == 27 && kind == 1) return batch_q<27, S, 1>(odd, tab, ent, nent, grid, cells, ncells, s);
  if (Q == 19 && kind == 1) return batch_q<19, S, 1>(odd, tab, ent, nent, grid, cells, ncells, s);
  return hipErrorInvalidValue;
}

hipError_t launch_hydro_batch(int Q, int storage, int collision, int odd, const HydroArgs* tab, const uint32_t* ent,
                              uint32_t nent, uint32_t grid, hipStream_t s) {
  if (grid == 0) return hipErrorInvalidValue;
  if (storage == kF32) return batch_s<float>(Q, collision, odd, tab, ent, nent, grid, nullptr, 0, s);
  if (storage == kF16) return batch_s<_Float16>(Q, collision, odd, tab, ent, nent, grid, nullptr, 0, s);
  return batch_s<bf16s>(Q, collision, odd, tab, ent, nent, grid, nullptr, 0, s);
}

hipError_t launch_hydro_list_batch(int Q, int storage, int collision, int odd, const HydroArgs* tab,
                                   const uint32_t* cells, uint32_t n, hipStream_t s) {
  if (storage == kF32) return batch_s<float>(Q, collision, odd, tab, nullptr, 0, 0, cells, n, s);
  if (storage == kF16) return batch_s<_Float16>(Q, collision, odd, tab, nullptr, 0, 0, cells, n, s);
  return batch_s<bf16s>(Q, collision, odd, tab, nullptr, 0, 0, cells, n, s);
}

hipError_t launch_clear_batch(const HydroArgs* tab, const uint32_t* cells, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_clear_batch, dim3(blocks_for(n)), dim3(256), 0, s, tab, cells, n);
  return hipGetLastError();
}

hipError_t launch_build_mask(int Q, const Dom& d, const uint8_t* solid, const float* wall_u, uint32_t* mask,
                             unsigned int* err, hipStream_t s) {
  if (Q == 19)
    hipLaunchKernelGGL((k_build_mask<19>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, solid, wall_u, mask, err);
  else
    hipLaunchKernelGGL((k_build_mask<27>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, solid, wall_u, mask, err);
  return hipGetLastError();
}

hipError_t launch_tile_flags(const Dom& d, const uint32_t* mask, uint8_t* flags, int linear, hipStream_t s) {
  const size_t nt = linear ? (size_t)(d.N + 63) / 64 : (size_t)((d.nx + 63) / 64) * d.ny * d.nz;
  hipLaunchKernelGGL(k_tile_flags, dim3(blocks_for(nt)), dim3(256), 0, s, d, mask, flags, linear);
  return hipGetLastError();
}

hipError_t launch_mark_forced(const uint32_t* flist, uint32_t n, uint32_t* mask, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mark_forced, dim3(blocks_for(n)), dim3(256), 0, s, flist, n, mask);
  return hipGetLastError();
}

template <typename S>
static hipError_t stage_s(int Q, int dir, int odd, const Dom& d, void* f, float* h, const uint32_t* mask,
                          hipStream_t s) {
  if (Q == 19)
    hipLaunchKernelGGL((k_stage<19, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (S*)f, h, mask, dir, odd);
  else
    hipLaunchKernelGGL((k_stage<27, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (S*)f, h, mask, dir, odd);
  return hipGetLastError();
}

hipError_t launch_stage(int Q, int storage, int dir, int odd, const Dom& d, void* f, float* h, const uint32_t* mask,
                        hipStream_t s) {
  if (storage == kF32) return stage_s<float>(Q, dir, odd, d, f, h, mask, s);
  if (storage == kF16) return stage_s<_Float16>(Q, dir, odd, d, f, h, mask, s);
  return stage_s<bf16s>(Q, dir, odd, d, f, h, mask, s);
}

template <typename S>
static hipError_t init_s(int Q, const Dom& d, void* f, const uint32_t* mask, const float* rho, const float* u,
                         float tau, hipStream_t s) {
  if (Q == 19)
    hipLaunchKernelGGL((k_init_with<19, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (S*)f, mask, rho, u, tau);
  else
    hipLaunchKernelGGL((k_init_with<27, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (S*)f, mask, rho, u, tau);
  return hipGetLastError();
}

hipError_t launch_init_with(int Q, int storage, const Dom& d, void* f, const uint32_t* mask, const float* rho,
                            const float* u, float tau, hipStream_t s) {
  if (storage == kF32) return init_s<float>(Q, d, f, mask, rho, u, tau, s);
  if (storage == kF16) return init_s<_Float16>(Q, d, f, mask, rho, u, tau, s);
  return init_s<bf16s>(Q, d, f, mask, rho, u, tau, s);
}

template <typename S>
static hipError_t moments_s(int Q, int odd, const Dom& d, const void* f, const uint32_t* mask, const float* F_mom,
                            HydroParams p, float* rho, float* ux, float* uy, float* uz, hipStream_t s) {
  if (Q == 19)
    hipLaunchKernelGGL((k_moments<19, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (const S*)f, mask, F_mom, p,
                       odd, rho, ux, uy, uz);
  else
    hipLaunchKernelGGL((k_moments<27, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (const S*)f, mask, F_mom, p,
                       odd, rho, ux, uy, uz);
  return hipGetLastError();
}

template <typename S>
static hipError_t poststate_s(int Q, int odd, const Dom& d, const void* f, const uint32_t* mask, const float* wall_u,
                              float* out, hipStream_t s) {
  if (Q == 19)
    hipLaunchKernelGGL((k_poststate<19, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (const S*)f, mask, wall_u, odd,
                       out);
  else
    hipLaunchKernelGGL((k_poststate<27, S>), dim3(blocks_for(d.N)), dim3(256), 0, s, d, (const S*)f, mask, wall_u, odd,
                       out);
  return hipGetLastError();
}

hipError_t launch_poststate(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                            const float* wall_u, float* out, hipStream_t s) {
  if (storage == kF32) return poststate_s<float>(Q, odd, d, f, mask, wall_u, out, s);
  if (storage == kF16) return poststate_s<_Float16>(Q, odd, d, f, mask, wall_u, out, s);
  return poststate_s<bf16s>(Q, odd, d, f, mask, wall_u, out, s);
}

hipError_t launch_moments(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                          const float* F_mom, HydroParams p, float* rho, float* ux, float* uy, float* uz,
                          hipStream_t s) {
  if (storage == kF32) return moments_s<float>(Q, odd, d, f, mask, F_mom, p, rho, ux, uy, uz, s);
  if (storage == kF16) return moments_s<_Float16>(Q, odd, d, f, mask, F_mom, p, rho, ux, uy, uz, s);
  return moments_s<bf16s>(Q, odd, d, f, mask, F_mom, p, rho, ux, uy, uz, s);
}

template <typename S>
static hipError_t reduce_s(int Q, int odd, const Dom& d, const void* f, const uint32_t* mask, const float* F,
                           HydroParams p, int kind, int axis, double* partials, uint32_t np, hipStream_t s) {
  if (Q == 19)
    hipLaunchKernelGGL((k_reduce<19, S>), dim3(np), dim3(256), 0, s, d, (const S*)f, mask, F, p, odd, kind, axis,
                       partials);
  else
    hipLaunchKernelGGL((k_reduce<27, S>), dim3(np), dim3(256), 0, s, d, (const S*)f, mask, F, p, odd, kind, axis,
                       partials);
  return hipGetLastError();
}

hipError_t launch_reduce(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                         const float* F, HydroParams p, int kind, int axis, double* partials, uint32_t np,
                         hipStream_t s) {
  if (storage == kF32) return reduce_s<float>(Q, odd, d, f, mask, F, p, kind, axis, partials, np, s);
  if (storage == kF16) return reduce_s<_Float16>(Q, odd, d, f, mask, F, p, kind, axis, partials, np, s);
  return reduce_s<bf16s>(Q, odd, d, f, mask, F, p, kind, axis, partials, np, s);
}

hipError_t launch_tank(const Dom& d, double dx, double diameter, double height, int baffles, double bw, double bt,
                       uint8_t* solid, hipStream_t s) {
  hipLaunchKernelGGL(k_tank, dim3(blocks_for(d.N)), dim3(256), 0, s, d, dx, diameter, height, baffles, bw, bt,
                     solid);
  return hipGetLastError();
}

__global__ void k_filld(double* p, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_rims(const Dom& d, const int is_wall[6], const double u[18], uint8_t* solid, float* wall_u,
                       double* best, hipStream_t s) {
  hipLaunchKernelGGL(k_filld, dim3(blocks_for(d.N)), dim3(256), 0, s, best, (size_t)d.N, -1.0);
  const int order[6] = {2, 3, 0, 1, 4, 5};  // YNeg, YPos, XNeg, XPos, ZNeg, ZPos
  const uint32_t n3[3] = {d.nx, d.ny, d.nz};
  for (int k = 0; k < 6; ++k) {
    const int face = order[k];
    if (!is_wall[face]) continue;
    const int a = face / 2;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const float u0 = (float)u[3 * face], u1 = (float)u[3 * face + 1], u2 = (float)u[3 * face + 2];
    double speed = (double)u0 * (double)u0 + (double)u1 * (double)u1;
    speed += (double)u2 * (double)u2;
    const size_t cnt = (size_t)n3[t1] * n3[t2];
    hipLaunchKernelGGL(k_rim_face, dim3(blocks_for(cnt)), dim3(256), 0, s, d, face, speed, u0, u1, u2, solid,
                       wall_u, best);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Copy n elements of T, each lane one naturally aligned element per
// iteration, starting OFF elements past 256-byte-aligned bases: the access
// widths and line offsets of the collide-stream kernels (4-byte planes, and
// their x-shifted neighbour rows that put a wave across three lines; 2-byte
// planes of 16-bit storage) next to the 16-byte copy, for calibrating the
// memory counters per access form.
template <typename T, int OFF>
__global__ __launch_bounds__(256) void k_access_copy(const T* __restrict__ a, T* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    b[OFF + i] = a[OFF + i];
}

hipError_t launch_access_copy(int width, int offset_elems, const void* a, void* b, size_t n, hipStream_t s) {
  const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 256 * 8 * 64);
  if (width == 16 && offset_elems == 0)
    hipLaunchKernelGGL((k_access_copy<uint4, 0>), dim3(grid), dim3(256), 0, s, (const uint4*)a, (uint4*)b, n);
  else if (width == 4 && offset_elems == 0)
    hipLaunchKernelGGL((k_access_copy<uint32_t, 0>), dim3(grid), dim3(256), 0, s, (const uint32_t*)a, (uint32_t*)b, n);
  else if (width == 4 && offset_elems == 1)
    hipLaunchKernelGGL((k_access_copy<uint32_t, 1>), dim3(grid), dim3(256), 0, s, (const uint32_t*)a, (uint32_t*)b, n);
  else if (width == 2 && offset_elems == 0)
    hipLaunchKernelGGL((k_access_copy<uint16_t, 0>), dim3(grid), dim3(256), 0, s, (const uint16_t*)a, (uint16_t*)b, n);
  else if (width == 2 && offset_elems == 1)
    hipLaunchKernelGGL((k_access_copy<uint16_t, 1>), dim3(grid), dim3(256), 0, s, (const uint16_t*)a, (uint16_t*)b, n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_stream_copy(const void* a, void* b, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 256 * 8 * 64);
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, s, (const float4*)a, (float4*)b, n);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_copy_force_list(const uint32_t* flist, uint32_t n, float* dst, const float* src, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_list, dim3(blocks_for(n)), dim3(256), 0, s, flist, n, dst, src);
  return hipGetLastError();
}

}  // namespace lbm
