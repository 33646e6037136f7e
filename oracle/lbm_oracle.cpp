// lbm_oracle.cpp — CPU restatement of lbm-core's CpuScalar lattice-Boltzmann
// path (TakuTsuzuki/lbmflow).  TEST INFRASTRUCTURE ONLY: nothing in the
// product library links or calls this file; tests, the smoke check and the
// bench's CPU-baseline leg load it as the checker.
//
// Every routine cites the reference file:line it restates.  Floating-point
// expressions keep the reference's operand order and grouping; the build
// uses -ffp-contract=off so no multiply-add is fused behind our back.
//
// Pins (tests/test_oracle_*.py):
//   lattice invariants, omegas, equilibrium fixed point, CM rest fixed point,
//   uniform-u uniformity, closed-box mass conservation, TGV decay, geometry
//   counts/volumes, QOI closed forms, IBM single-marker exactness; WALE
//   closed forms; open faces (imposed velocity / density, fixed points,
//   channel flux and profile); volume-source mass and momentum; face
//   patches (the reference's impinging-jet and full-face acceptance tests,
//   exact per-rect rules); force-probe momentum balance.
#include <xmmintrin.h>
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Lattice tables — lattice.rs:295-442 (D3Q19/D3Q27 velocity order and
// weights), opp/pairs derived as in lattice.rs:158-191.  D3Q7 follows the
// scalar-lattice definition (rest + 6 axis links, w0 = 1/4, w1..6 = 1/8),
// docs/archive/2026-07-07-pivot/proposals/WSCAL_MULTICOMPONENT_SPEC.md:68-71.
// ---------------------------------------------------------------------------
static const int8_t C27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0},
    {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1}, {0, 1, -1},  {0, -1, 1}, {1, 1, 1},   {-1, -1, -1},
    {1, 1, -1},  {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};

struct Lat {
  int Q = 0;
  int D = 3;
  int8_t c[27][3];
  double w[27];
  int opp[27];
  int npairs = 0;
  int pairs[13][2];
};

static Lat make_lat(int Q) {
  Lat L;
  L.Q = Q;
  for (int q = 0; q < Q; ++q)
    for (int a = 0; a < 3; ++a) L.c[q][a] = C27[q][a];
  for (int q = 0; q < Q; ++q) {
    if (Q == 19) L.w[q] = q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
    if (Q == 27) L.w[q] = q == 0 ? 8.0 / 27.0 : (q < 7 ? 2.0 / 27.0 : (q < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    if (Q == 7) L.w[q] = q == 0 ? 1.0 / 4.0 : 1.0 / 8.0;
  }
  for (int q = 0; q < Q; ++q) {
    L.opp[q] = -1;
    for (int r = 0; r < Q; ++r)
      if (L.c[r][0] == -L.c[q][0] && L.c[r][1] == -L.c[q][1] && L.c[r][2] == -L.c[q][2]) L.opp[q] = r;
  }
  for (int q = 0; q < Q; ++q)
    if (q != 0 && q < L.opp[q]) {
      L.pairs[L.npairs][0] = q;
      L.pairs[L.npairs][1] = L.opp[q];
      L.npairs++;
    }
  return L;
}

// ---------------------------------------------------------------------------
// Central-moment transform — kernels.rs:162-392.
// ---------------------------------------------------------------------------
struct MomentTransform {
  int Q;
  uint8_t basis[27][3];
  int exp_index[3][3][3];
  double raw_from_pop[27][27];
  double pop_from_raw[27][27];
};

static double pow_upto2(double x, int e) { return e == 0 ? 1.0 : (e == 1 ? x : x * x); }
static double binom_upto2(int n, int k) {
  if (k == 0) return 1.0;
  if (n == 1 && k == 1) return 1.0;
  if (n == 2 && k == 1) return 2.0;
  return 1.0;  // (2,2)
}

static MomentTransform build_transform(const Lat& L) {
  MomentTransform t;
  t.Q = L.Q;
  int n = 0;
  // central_basis, kernels.rs:162-183
  for (int ax = 0; ax <= 2; ++ax)
    for (int ay = 0; ay <= 2; ++ay)
      for (int az = 0; az <= 2; ++az) {
        if (L.Q == 19 && ax > 0 && ay > 0 && az > 0) continue;
        if (n < L.Q) {
          t.basis[n][0] = ax;
          t.basis[n][1] = ay;
          t.basis[n][2] = az;
          n++;
        }
      }
  for (int i = 0; i < 27; ++i) t.exp_index[i / 9][(i / 3) % 3][i % 3] = -1;
  // build_moment_transform, kernels.rs:279-297; raw_phi kernels.rs:195-201
  for (int m = 0; m < L.Q; ++m) {
    const uint8_t* e = t.basis[m];
    t.exp_index[e[0]][e[1]][e[2]] = m;
    for (int q = 0; q < L.Q; ++q) {
      double v = pow_upto2((double)L.c[q][0], e[0]);
      v *= pow_upto2((double)L.c[q][1], e[1]);
      v *= pow_upto2((double)L.c[q][2], e[2]);
      t.raw_from_pop[m][q] = v;
    }
  }
  // invert_raw_moment_matrix, kernels.rs:227-277 (Gauss-Jordan, partial pivot)
  const int N = L.Q;
  static double a[27][54];
  for (int r = 0; r < N; ++r) {
    for (int c = 0; c < 2 * N; ++c) a[r][c] = 0.0;
    for (int c = 0; c < N; ++c) a[r][c] = t.raw_from_pop[r][c];
    a[r][N + r] = 1.0;
  }
  for (int col = 0; col < N; ++col) {
    int pivot = col;
    for (int row = col + 1; row < N; ++row)
      if (std::fabs(a[row][col]) > std::fabs(a[pivot][col])) pivot = row;
    if (pivot != col)
      for (int j = 0; j < 2 * N; ++j) std::swap(a[pivot][j], a[col][j]);
    double inv = 1.0 / a[col][col];
    for (int j = col; j < 2 * N; ++j) a[col][j] *= inv;
    for (int row = 0; row < N; ++row) {
      if (row == col) continue;
      double factor = a[row][col];
      if (factor == 0.0) continue;
      for (int j = col; j < 2 * N; ++j) a[row][j] -= factor * a[col][j];
    }
  }
  for (int r = 0; r < N; ++r)
    for (int c = 0; c < N; ++c) t.pop_from_raw[r][c] = a[r][N + c];
  return t;
}

static const MomentTransform& transform_for(int Q) {
  static MomentTransform t19 = build_transform(make_lat(19));
  static MomentTransform t27 = build_transform(make_lat(27));
  return Q == 19 ? t19 : t27;
}

// shifted_moments, kernels.rs:332-363
static void shifted_moments(const MomentTransform& t, const double u[3], const double* in, double* out,
                            bool raw_to_central) {
  for (int m = 0; m < t.Q; ++m) {
    const uint8_t* e = t.basis[m];
    double acc = 0.0;
    for (int kx = 0; kx <= e[0]; ++kx)
      for (int ky = 0; ky <= e[1]; ++ky)
        for (int kz = 0; kz <= e[2]; ++kz) {
          int idx = t.exp_index[kx][ky][kz];
          double coef = binom_upto2(e[0], kx) * binom_upto2(e[1], ky) * binom_upto2(e[2], kz);
          double sx = raw_to_central ? -u[0] : u[0];
          double sy = raw_to_central ? -u[1] : u[1];
          double sz = raw_to_central ? -u[2] : u[2];
          coef *= pow_upto2(sx, e[0] - kx);
          coef *= pow_upto2(sy, e[1] - ky);
          coef *= pow_upto2(sz, e[2] - kz);
          acc += coef * in[idx];
        }
    out[m] = acc;
  }
}

// Factorised D3Q27 transforms (test infrastructure speed-up, not a
// reference function).  The D3Q27 basis is the tensor product of the 1-D
// exponents 0..2 (central_basis, kernels.rs:162-183: basis index m =
// 9 ex + 3 ey + ez), so the central moments sum_q (c - u)^e f_q factor into
// three passes of 1-D three-point lines, and the inverse (raw moments by
// the binomial shift +u, then populations) into three inverse passes.  Same
// mathematics as the generic matrix path in exact arithmetic; the two agree
// to f64 round-off (pinned in tests).  On by default for D3Q27.
static bool g_cm_factorised = true;

// G[(cx+1)*9 + (cy+1)*3 + (cz+1)] = population of velocity c  ->
// G[9 ex + 3 ey + ez] = central moment of exponent e
static void chimera_fwd(double* G, const double u[3]) {
  auto line = [](double& a0, double& a1, double& a2, double uu) {
    const double gm = a0, g0 = a1, gp = a2;
    const double k0 = (gm + g0) + gp;
    const double k1 = (gp - gm) - uu * k0;
    const double k2 = ((gp + gm) - 2.0 * uu * (gp - gm)) + uu * uu * k0;
    a0 = k0;
    a1 = k1;
    a2 = k2;
  };
  for (int ab = 0; ab < 9; ++ab) line(G[ab * 3 + 0], G[ab * 3 + 1], G[ab * 3 + 2], u[2]);
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) line(G[a * 9 + c], G[a * 9 + 3 + c], G[a * 9 + 6 + c], u[1]);
  for (int bc = 0; bc < 9; ++bc) line(G[bc], G[9 + bc], G[18 + bc], u[0]);
}
// central moments -> populations (inverse of chimera_fwd)
static void chimera_inv(double* G, const double u[3]) {
  auto line = [](double& a0, double& a1, double& a2, double uu) {
    const double k0 = a0, k1 = a1, k2 = a2;
    const double m1 = k1 + uu * k0;
    const double m2 = (k2 + 2.0 * uu * k1) + uu * uu * k0;
    a0 = 0.5 * (m2 - m1);
    a1 = k0 - m2;
    a2 = 0.5 * (m1 + m2);
  };
  for (int bc = 0; bc < 9; ++bc) line(G[bc], G[9 + bc], G[18 + bc], u[0]);
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) line(G[a * 9 + c], G[a * 9 + 3 + c], G[a * 9 + 6 + c], u[1]);
  for (int ab = 0; ab < 9; ++ab) line(G[ab * 3 + 0], G[ab * 3 + 1], G[ab * 3 + 2], u[2]);
}

// populations_to_central_moments, kernels.rs:318-330
static void pops_to_central(const MomentTransform& t, const double u[3], const double* pop, double* out) {
  double raw[27];
  for (int m = 0; m < t.Q; ++m) {
    raw[m] = 0.0;
    for (int q = 0; q < t.Q; ++q) raw[m] += t.raw_from_pop[m][q] * pop[q];
  }
  shifted_moments(t, u, raw, out, true);
}

// ---------------------------------------------------------------------------
// Solver state — fields.rs:214-303 (q-major padded SoA deviation storage,
// compact moments), subdomain.rs:31-44 (monolithic, halo 1).
// ---------------------------------------------------------------------------
template <typename T>
struct ScalarDist {
  double diffusivity;
  Lat L;  // hydro lattice (reference) or D3Q7 (new)
  bool d3q7;
  std::vector<T> h, htmp, conc;
};

template <typename T>
struct Solver {
  Lat L;
  int nx, ny, nz;
  bool periodic[3];
  int px, py, pz;
  size_t np, nc;
  std::vector<T> f, ftmp;
  std::vector<T> rho, ux, uy, uz;
  std::vector<uint8_t> solid;  // padded
  std::vector<T> wall_u;       // padded x3
  std::vector<uint8_t> probe;  // padded, empty = no force probe
  T probed[3] = {T(0), T(0), T(0)};
  bool has_ff = false;
  std::vector<T> ff;  // compact x3
  bool has_omega = false;
  std::vector<T> omega_field;  // compact
  int collision;               // 0 BGK, 1 TRT, 2 CM
  double nu, omega_p, omega_m, omega_shear_in;
  T force[3];
  // KParams, params.rs:350-374
  T kp_op, kp_om, kp_os, kp_cp, kp_cm;
  T cr[27][3], wr[27];
  std::vector<ScalarDist<T>> scalars;
  long long time = 0;
  // open faces (FaceBC, params.rs:232-252), Face::ALL order XNeg..ZPos:
  // kind 0 closed, 1 Zou-He velocity, 2 Zou-He pressure, 3 outflow,
  // 4 convective
  // volume sources (VolumeSource, params.rs:262-283): box lo..hi inclusive,
  // kind 0 mass flow, 1 jet with velocity u
  struct Src {
    long lo[3], hi[3];
    int kind;
    T q_lu;
    T u[3];
  };
  std::vector<Src> sources;
  int face_kind[6] = {0, 0, 0, 0, 0, 0};
  T face_u[6][3] = {};
  T face_rho[6] = {};
  T face_uconv[6] = {};
  // one face rule: kind as above with its parameters
  struct Rule {
    int kind;
    T u[3];
    T rho;
    T uconv;
  };
  // FacePatch (params.rs:286-296): inclusive in-face rect lo..hi over the
  // face's tangents (ascending axis order) with its own rule
  struct Patch {
    int face;
    long lo[2], hi[2];
    Rule bc;
  };
  std::vector<Patch> patches;
  // inlet profiles (FieldSet::inlet_profiles, fields.rs:267): per velocity
  // face, u[3] per face cell in face order (c2 * n1 + c1); empty = none
  std::vector<T> profile[6];

  size_t pidx(long x, long y, long z) const {
    return ((size_t)(z + 1) * py + (size_t)(y + 1)) * px + (size_t)(x + 1);
  }
  size_t cidx(long x, long y, long z) const { return ((size_t)z * ny + y) * nx + x; }

  bool force_on() const {  // params.rs:376-383 (no gravity on this path)
    return force[0] != T(0) || force[1] != T(0) || force[2] != T(0) || has_ff;
  }
  void force_at(size_t c, T out[3]) const {  // params.rs:385-405
    if (has_ff) {
      out[0] = force[0] + ff[3 * c + 0];
      out[1] = force[1] + ff[3 * c + 1];
      out[2] = force[2] + ff[3 * c + 2];
    } else {
      out[0] = force[0];
      out[1] = force[1];
      out[2] = force[2];
    }
  }

  // Fill halo cells of periodic axes with the wrapped core values (the
  // LocalPeriodic exchange, halo.rs:77-100; every plane copied).
  template <typename V>
  void wrap_halo(std::vector<V>& arr, int planes, int comps) {
    for (long z = -1; z <= nz; ++z)
      for (long y = -1; y <= ny; ++y)
        for (long x = -1; x <= nx; ++x) {
          bool in_halo = x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz;
          if (!in_halo) continue;
          long s[3] = {x, y, z};
          long n[3] = {nx, ny, nz};
          bool ok = true;
          for (int a = 0; a < 3; ++a) {
            if (s[a] < 0 || s[a] >= n[a]) {
              if (!periodic[a]) {
                ok = false;
                break;
              }
              s[a] = (s[a] + n[a]) % n[a];
            }
          }
          if (!ok) continue;
          size_t di = pidx(x, y, z), si = pidx(s[0], s[1], s[2]);
          for (int p = 0; p < planes; ++p)
            for (int k = 0; k < comps; ++k) arr[(p * np + di) * comps + k] = arr[(p * np + si) * comps + k];
        }
  }

  void init(int Q, const int dims[3], const int per[3], int coll, double nu_, double magic, double os,
            const double frc[3]) {
    L = make_lat(Q);
    nx = dims[0];
    ny = dims[1];
    nz = dims[2];
    for (int a = 0; a < 3; ++a) periodic[a] = per[a] != 0;
    px = nx + 2;
    py = ny + 2;
    pz = nz + 2;
    np = (size_t)px * py * pz;
    nc = (size_t)nx * ny * nz;
    f.assign((size_t)Q * np, T(0));
    ftmp.assign((size_t)Q * np, T(0));
    rho.assign(nc, T(1));
    ux.assign(nc, T(0));
    uy.assign(nc, T(0));
    uz.assign(nc, T(0));
    solid.assign(np, 0);
    wall_u.assign(np * 3, T(0));
    collision = coll;
    nu = nu_;
    // CollisionKind::omegas, params.rs:205-217
    double tau = 3.0 * nu + 0.5;
    omega_p = 1.0 / tau;
    if (coll == 1) {
      double lam_p = tau - 0.5;
      omega_m = 1.0 / (magic / lam_p + 0.5);
    } else {
      omega_m = omega_p;
    }
    omega_shear_in = coll == 2 ? os : omega_p;
    for (int a = 0; a < 3; ++a) force[a] = (T)frc[a];
    kp_op = (T)omega_p;
    kp_om = (T)omega_m;
    kp_os = (T)omega_shear_in;
    kp_cp = (T)(1.0 - omega_p / 2.0);
    kp_cm = (T)(1.0 - omega_m / 2.0);
    for (int q = 0; q < 27; ++q) {
      for (int a = 0; a < 3; ++a) cr[q][a] = q < Q ? (T)(double)L.c[q][a] : T(0);
      wr[q] = q < Q ? (T)L.w[q] : T(0);
    }
  }

  // Solver::build tail, solver.rs:1829-1891: masks into padded core,
  // exchange_masks, update_moments.
  void set_masks(const uint8_t* solid_c, const double* wall_c) {
    for (long z = 0; z < nz; ++z)
      for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) {
          size_t gi = cidx(x, y, z), pi = pidx(x, y, z);
          solid[pi] = solid_c ? solid_c[gi] : 0;
          for (int a = 0; a < 3; ++a) wall_u[3 * pi + a] = wall_c ? (T)wall_c[3 * gi + a] : T(0);
        }
    wrap_halo(solid, 1, 1);
    wrap_halo(wall_u, 1, 3);
    update_moments();
  }

  // ---- kernels ------------------------------------------------------------
  // TrtGuo::relax via collide_row, kernels.rs:129-160, collision.rs:123-202.
  void collide_trt_cell(long x, long y, long z) {
    size_t i = pidx(x, y, z), c = cidx(x, y, z);
    const int Q = L.Q;
    T r = rho[c];
    T u[3] = {ux[c], uy[c], uz[c]};
    T fv[3];
    force_at(c, fv);
    T op = has_omega ? omega_field[c] : kp_op;
    T cp = has_omega ? (T(1) - op / T(2.0)) : kp_cp;
    bool fon = force_on();
    const T half = T(0.5);
    T usq = u[0] * u[0];
    for (int d = 1; d < 3; ++d) {
      T ud2 = u[d] * u[d];
      usq = usq + ud2;
    }
    T uf = u[0] * fv[0];
    for (int d = 1; d < 3; ++d) {
      T udfd = u[d] * fv[d];
      uf = uf + udfd;
    }
    T feq[27], src[27];
    for (int q = 0; q < Q; ++q) {
      // dot_u, collision.rs:53-65
      T cu = cr[q][0] * u[0];
      for (int d = 1; d < 3; ++d) {
        T term = cr[q][d] * u[d];
        cu = cu + term;
      }
      // SecondOrderEq::feq, collision.rs:81-102
      T drho = r - T(1.0);
      T three_cu = T(3.0) * cu;
      T f45_cu = T(4.5) * cu;
      T f45_cu2 = f45_cu * cu;
      T f15_usq = T(1.5) * usq;
      T sum = three_cu + f45_cu2;
      T inner = sum - f15_usq;
      T rho_inner = r * inner;
      T body = drho + rho_inner;
      feq[q] = wr[q] * body;
      // GuoForcing::source, collision.rs:104-121
      if (!fon) {
        src[q] = T(0.0);
      } else {
        T cf = cr[q][0] * fv[0];
        for (int d = 1; d < 3; ++d) {
          T term = cr[q][d] * fv[d];
          cf = cf + term;
        }
        T cf_minus_uf = cf - uf;
        T first = T(3.0) * cf_minus_uf;
        T nine_cu = T(9.0) * cu;
        T second = nine_cu * cf;
        T b = first + second;
        src[q] = wr[q] * b;
      }
    }
    T f0 = f[0 * np + i];
    T f0_minus_e = f0 - feq[0];
    T op_delta = op * f0_minus_e;
    T relaxed = f0 - op_delta;
    T cp_src = cp * src[0];
    f[0 * np + i] = relaxed + cp_src;
    for (int k = 0; k < L.npairs; ++k) {
      int qa = L.pairs[k][0], qb = L.pairs[k][1];
      T fa = f[qa * np + i], fb = f[qb * np + i];
      T fp = half * (fa + fb);
      T fm = half * (fa - fb);
      T ep = half * (feq[qa] + feq[qb]);
      T em = half * (feq[qa] - feq[qb]);
      T sp = half * (src[qa] + src[qb]);
      T sm = half * (src[qa] - src[qb]);
      T rp = op * (fp - ep);
      T rm = kp_om * (fm - em);
      T cp_sp = cp * sp;
      T cm_sm = kp_cm * sm;
      T fa_rp = fa - rp;
      T fa_rp_rm = fa_rp - rm;
      T src_a = cp_sp + cm_sm;
      T va = fa_rp_rm + src_a;
      T fb_rp = fb - rp;
      T fb_rp_rm = fb_rp + rm;
      T src_b = cp_sp - cm_sm;
      T vb = fb_rp_rm + src_b;
      f[qa * np + i] = va;
      f[qb * np + i] = vb;
    }
  }

  // collide_row_central_moment, kernels.rs:406-533 (f64 internally).
  void collide_cm_cell(long x, long y, long z) {
    const MomentTransform& tr = transform_for(L.Q);
    const int Q = L.Q;
    size_t i = pidx(x, y, z), c = cidx(x, y, z);
    T r_t = rho[c];
    double r = (double)r_t;
    double u[3] = {(double)ux[c], (double)uy[c], (double)uz[c]};
    T fv_t[3];
    force_at(c, fv_t);
    double fv[3] = {(double)fv_t[0], (double)fv_t[1], (double)fv_t[2]};
    bool fon = fv[0] != 0.0 || fv[1] != 0.0 || fv[2] != 0.0;
    double phys[27], src[27];
    for (int q = 0; q < 27; ++q) src[q] = 0.0;
    bool feq_rest = u[0] == 0.0 && u[1] == 0.0 && u[2] == 0.0 && !fon;
    bool exact_rest = feq_rest;
    for (int q = 0; q < Q; ++q) {
      T fq = f[q * np + i];
      phys[q] = (double)fq + L.w[q];
      if (exact_rest) {
        T expected = (T)(L.w[q] * (r - 1.0));
        double a = (double)fq, b = (double)expected;
        uint64_t ba, bb;
        std::memcpy(&ba, &a, 8);
        std::memcpy(&bb, &b, 8);
        exact_rest = exact_rest && ba == bb;
      }
    }
    if (exact_rest) return;
    if (fon) {
      double uf = u[0] * fv[0];
      for (int d = 1; d < 3; ++d) uf += u[d] * fv[d];
      for (int q = 0; q < Q; ++q) {
        double cu = (double)L.c[q][0] * u[0];
        double cf = (double)L.c[q][0] * fv[0];
        for (int d = 1; d < 3; ++d) {
          cu += (double)L.c[q][d] * u[d];
          cf += (double)L.c[q][d] * fv[d];
        }
        src[q] = L.w[q] * (3.0 * (cf - uf) + 9.0 * cu * cf);
      }
    }
    double usq = u[0] * u[0];
    for (int d = 1; d < 3; ++d) usq += u[d] * u[d];
    double feq_phys[27];
    for (int q = 0; q < Q; ++q) {
      double cu = (double)L.c[q][0] * u[0];
      for (int d = 1; d < 3; ++d) cu += (double)L.c[q][d] * u[d];
      feq_phys[q] = L.w[q] * r * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    }
    double mom[27], eq[27], src_mom[27];
    const bool fact = Q == 27 && g_cm_factorised;
    if (fact) {
      for (int q = 0; q < 27; ++q) {
        const int g = (L.c[q][0] + 1) * 9 + (L.c[q][1] + 1) * 3 + (L.c[q][2] + 1);
        mom[g] = phys[q];
        eq[g] = feq_phys[q];
        src_mom[g] = src[q];
      }
      chimera_fwd(mom, u);
      chimera_fwd(eq, u);
      chimera_fwd(src_mom, u);
    } else {
      pops_to_central(tr, u, phys, mom);
      pops_to_central(tr, u, feq_phys, eq);
      pops_to_central(tr, u, src, src_mom);
    }
    double os_base = has_omega ? (double)omega_field[c] : (double)kp_os;
    double usq2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double vc = -0.16 * usq2;
    double os = std::min(os_base * (1.0 + vc), 2.0);
    double post[27];
    int diag[3] = {-1, -1, -1};
    for (int m = 0; m < Q; ++m) {
      const uint8_t* e = tr.basis[m];
      int order = e[0] + e[1] + e[2];
      if (order == 2)
        for (int a = 0; a < 3; ++a) {
          uint8_t de[3] = {0, 0, 0};
          de[a] = 2;
          if (e[0] == de[0] && e[1] == de[1] && e[2] == de[2]) diag[a] = m;
        }
      double rate = order <= 1 ? 0.0 : (order == 2 ? os : 1.0);
      post[m] = mom[m] - rate * (mom[m] - eq[m]) + (1.0 - 0.5 * rate) * src_mom[m];
    }
    double inv_d = 1.0 / 3.0;
    double trace_neq = 0.0, trace_src = 0.0;
    for (int a = 0; a < 3; ++a) {
      trace_neq += mom[diag[a]] - eq[diag[a]];
      trace_src += src_mom[diag[a]];
    }
    double bulk_neq = trace_neq * inv_d;
    double bulk_src = trace_src * inv_d;
    for (int a = 0; a < 3; ++a) {
      int idx = diag[a];
      double dev_neq = mom[idx] - eq[idx] - bulk_neq;
      double dev_src = src_mom[idx] - bulk_src;
      post[idx] = eq[idx] + (1.0 - os) * dev_neq + 0.5 * bulk_src + (1.0 - 0.5 * os) * dev_src;
    }
    if (fact) {
      chimera_inv(post, u);
      for (int q = 0; q < 27; ++q) {
        const int g = (L.c[q][0] + 1) * 9 + (L.c[q][1] + 1) * 3 + (L.c[q][2] + 1);
        f[q * np + i] = (T)(post[g] - L.w[q]);
      }
      return;
    }
    double raw_post[27];
    shifted_moments(tr, u, post, raw_post, false);
    for (int q = 0; q < Q; ++q) {
      double o = 0.0;
      for (int m = 0; m < Q; ++m) o += tr.pop_from_raw[q][m] * raw_post[m];
      f[q * np + i] = (T)(o - L.w[q]);
    }
  }

  void collide() {
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)ny * nz; ++r) {
      long y = r % ny, z = r / ny;
      for (long x = 0; x < nx; ++x) {
        if (solid[pidx(x, y, z)]) continue;
        if (collision == 2)
          collide_cm_cell(x, y, z);
        else
          collide_trt_cell(x, y, z);
      }
    }
  }

  // stream_row, kernels.rs:547-622 (no probe: the momentum-exchange probe
  // is not on this path).
  void stream() {
    const T six = T(6.0), two = T(2.0);
    const int Q = L.Q;
    const bool prb = !probe.empty();
    std::vector<T> part(prb ? 3 * (size_t)ny * nz : 0, T(0));
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)ny * nz; ++r) {
      long y = r % ny, z = r / ny;
      T pf[3] = {T(0), T(0), T(0)};
      for (long x = 0; x < nx; ++x) {
        size_t i = pidx(x, y, z);
        if (solid[i]) continue;
        T rho_x = rho[cidx(x, y, z)];
        for (int q = 0; q < Q; ++q) {
          long s[3] = {x - L.c[q][0], y - L.c[q][1], z - L.c[q][2]};
          long n[3] = {nx, ny, nz};
          bool skip = false;
          for (int a = 0; a < 3; ++a) {
            if (s[a] < 0) {
              if (!periodic[a]) skip = true;
            } else if (s[a] >= n[a] && !periodic[a]) {
              skip = true;
            }
          }
          if (skip) continue;
          size_t si = pidx(s[0], s[1], s[2]);
          if (solid[si]) {
            T fout = f[L.opp[q] * np + i];
            const T* wu = &wall_u[3 * si];
            T cu = cr[q][0] * wu[0];
            for (int d = 1; d < 3; ++d) cu = cu + cr[q][d] * wu[d];
            T fin = fout + six * wr[q] * rho_x * cu;
            ftmp[q * np + i] = fin;
            if (prb && probe[si]) {  // momentum given to the wall (stream_row, kernels.rs:605-615)
              const T ftot = fout + fin + two * wr[q];
              pf[0] = pf[0] - cr[q][0] * ftot;
              pf[1] = pf[1] - cr[q][1] * ftot;
              pf[2] = pf[2] - cr[q][2] * ftot;
            }
          } else {
            ftmp[q * np + i] = f[q * np + si];
          }
        }
      }
      if (prb)
        for (int a = 0; a < 3; ++a) part[3 * r + a] = pf[a];
    }
    // deterministic total: row partials folded in row order (backend.rs:521-544)
    probed[0] = probed[1] = probed[2] = T(0);
    if (prb)
      for (long r = 0; r < (long)ny * nz; ++r)
        for (int a = 0; a < 3; ++a) probed[a] = probed[a] + part[3 * r + a];
  }
  // Solver::set_force_probe, solver.rs:2741-2760 (mask over global cells,
  // exchanged into the halo like the solid mask)
  void set_probe(const uint8_t* m) {
    if (!m) {
      probe.clear();
      return;
    }
    probe.assign(np, 0);
    for (long z = 0; z < nz; ++z)
      for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) probe[pidx(x, y, z)] = m[cidx(x, y, z)];
    wrap_halo(probe, 1, 1);
  }

  // moments_row via update_moments_impl, kernels.rs:627-668, backend.rs:858-916
  void update_moments() {
    const T half = T(0.5);
    const int Q = L.Q;
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)ny * nz; ++r) {
      long y = r % ny, z = r / ny;
      for (long x = 0; x < nx; ++x) {
        size_t i = pidx(x, y, z), c = cidx(x, y, z);
        if (solid[i]) continue;
        T dr = T(0);
        T m[3] = {T(0), T(0), T(0)};
        for (int q = 0; q < Q; ++q) {
          T fq = f[q * np + i];
          dr = dr + fq;
          m[0] = m[0] + cr[q][0] * fq;
          m[1] = m[1] + cr[q][1] * fq;
          m[2] = m[2] + cr[q][2] * fq;
        }
        T rr = T(1) + dr;
        T fv[3];
        force_at(c, fv);
        rho[c] = rr;
        T inv = T(1) / rr;
        ux[c] = (m[0] + half * fv[0]) * inv;
        uy[c] = (m[1] + half * fv[1]) * inv;
        uz[c] = (m[2] + half * fv[2]) * inv;
      }
    }
  }

  // ---- open faces ----------------------------------------------------------
  // Face geometry, lattice.rs:38-90: axis = face/2, inward normal +1 on the
  // low face, tangents ascending; unknowns = {q : c_q . n_in > 0} ascending
  // (face_unknowns, lattice.rs:193-214).
  int dir_of(int cx, int cy, int cz) const {
    for (int q = 0; q < L.Q; ++q)
      if (L.c[q][0] == cx && L.c[q][1] == cy && L.c[q][2] == cz) return q;
    return -1;
  }
  struct FaceGeo {
    int a, t1, t2, nin[3];
    long fixed;
    int unk[9];
    int nunk;
  };
  FaceGeo face_geo(int face) const {
    FaceGeo g{};
    g.a = face / 2;
    const bool neg = face % 2 == 0;
    g.nin[g.a] = neg ? 1 : -1;
    g.t1 = g.a == 0 ? 1 : 0;
    g.t2 = g.a == 2 ? 1 : 2;
    const long n[3] = {nx, ny, nz};
    g.fixed = neg ? 0 : n[g.a] - 1;
    g.nunk = 0;
    for (int q = 0; q < L.Q; ++q) {
      int dot = L.c[q][0] * g.nin[0] + L.c[q][1] * g.nin[1] + L.c[q][2] * g.nin[2];
      if (dot > 0) g.unk[g.nunk++] = q;
    }
    return g;
  }
  // FaceCellSelection (kernels.rs:718-740): kind 0 All, 1 Rect lo..hi,
  // 2 Excluding every rect of `ex`
  struct Sel {
    int kind = 0;
    long lo[2] = {0, 0}, hi[2] = {0, 0};
    const std::vector<std::array<long, 4>>* ex = nullptr;
    bool operator()(long c1, long c2) const {
      if (kind == 1) return lo[0] <= c1 && c1 <= hi[0] && lo[1] <= c2 && c2 <= hi[1];
      if (kind == 2)
        for (const auto& r : *ex)
          if (r[0] <= c1 && c1 <= r[2] && r[1] <= c2 && c2 <= r[3]) return false;
      return true;
    }
  };
  // for_face_cells_selected (kernels.rs:742-766): t2 outer, t1 inner
  long coord_ = 0;  // face coordinate of the cell for_face_cells visits
  template <class F>
  void for_face_cells(const FaceGeo& g, const Sel& sel, F&& body) {
    const long n[3] = {nx, ny, nz};
    for (long c2 = 0; c2 < n[g.t2]; ++c2)
      for (long c1 = 0; c1 < n[g.t1]; ++c1) {
        if (!sel(c1, c2)) continue;
        coord_ = c2 * n[g.t1] + c1;
        long pos[3];
        pos[g.a] = g.fixed;
        pos[g.t1] = c1;
        pos[g.t2] = c2;
        body(pos);
      }
  }

  // zou_he_face_3d (D3Q19, kernels.rs:1085-1219) and zou_he_face_d3q27
  // (kernels.rs:958-1083); kind 1 velocity, 2 pressure.  Gravity is not on
  // this path (g = 0); F0 = params force + field (force_independent_at).
  void zou_he_face(int face, const Rule& bc, const Sel& sel, const T* prof = nullptr) {
    const FaceGeo g = face_geo(face);
    const int a = g.a, t1 = g.t1, t2 = g.t2;
    const T nsign = T(g.nin[a]);
    const T half = T(0.5), two = T(2.0), one = T(1);
    const bool vel = bc.kind == 1;
    const T* ub = bc.u;
    const T rho_bc = bc.rho;
    if (L.Q == 27) {
      double s2a = 0.0, s2b = 0.0;
      for (int k = 0; k < g.nunk; ++k) {
        const int q = g.unk[k];
        s2a += L.w[q] * (double)L.c[q][t1] * (double)L.c[q][t1];
        s2b += L.w[q] * (double)L.c[q][t2] * (double)L.c[q][t2];
      }
      const T s2_t1 = T(s2a), s2_t2 = T(s2b), six = T(6.0);
      for_face_cells(g, sel, [&](const long* pos) {
        const size_t i = pidx(pos[0], pos[1], pos[2]);
        if (solid[i]) return;
        T f0[3];
        force_at(cidx(pos[0], pos[1], pos[2]), f0);
        T s0 = T(0), sneg = T(0), q0a = T(0), q0b = T(0);
        for (int q = 0; q < L.Q; ++q) {
          const int dot = L.c[q][a] * g.nin[a];
          const T fq = f[q * np + i];
          if (dot == 0) {
            s0 = s0 + fq;
            q0a = q0a + T(L.c[q][t1]) * fq;
            q0b = q0b + T(L.c[q][t2]) * fq;
          } else if (dot < 0) {
            sneg = sneg + fq;
          }
        }
        const T closure = s0 + two * sneg + one;
        T r, vn, vt1, vt2;
        // profile.map_or(u, |p| p[coord]) (kernels.rs:1021-1034)
        const T* uc = (vel && prof) ? prof + 3 * coord_ : ub;
        zh_solve(vel, uc, rho_bc, closure, f0, a, t1, t2, nsign, r, vn, vt1, vt2);
        const T c_t1 = r * vt1 * (one - six * s2_t1) - q0a;
        const T c_t2 = r * vt2 * (one - six * s2_t2) - q0b;
        for (int k = 0; k < g.nunk; ++k) {
          const int q = g.unk[k];
          const T cdotu = T(L.c[q][a]) * nsign * vn + T(L.c[q][t1]) * vt1 + T(L.c[q][t2]) * vt2;
          const T w = T(L.w[q]);
          const T delta = c_t1 * w * T(L.c[q][t1]) / s2_t1 + c_t2 * w * T(L.c[q][t2]) / s2_t2;
          f[q * np + i] = f[L.opp[q] * np + i] + six * w * r * cdotu + delta;
        }
      });
      return;
    }
    int e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
    e1[t1] = 1;
    e2[t2] = 1;
    const int* n = g.nin;
    const int q_n = dir_of(n[0], n[1], n[2]);
    const int q_p1 = dir_of(n[0] + e1[0], n[1] + e1[1], n[2] + e1[2]);
    const int q_m1 = dir_of(n[0] - e1[0], n[1] - e1[1], n[2] - e1[2]);
    const int q_p2 = dir_of(n[0] + e2[0], n[1] + e2[1], n[2] + e2[2]);
    const int q_m2 = dir_of(n[0] - e2[0], n[1] - e2[1], n[2] - e2[2]);
    const int q_t1 = dir_of(e1[0], e1[1], e1[2]), q_mt1 = dir_of(-e1[0], -e1[1], -e1[2]);
    const int q_t2 = dir_of(e2[0], e2[1], e2[2]), q_mt2 = dir_of(-e2[0], -e2[1], -e2[2]);
    const int q_pp = dir_of(e1[0] + e2[0], e1[1] + e2[1], e1[2] + e2[2]);
    const int q_pm = dir_of(e1[0] - e2[0], e1[1] - e2[1], e1[2] - e2[2]);
    const int q_mp = dir_of(-e1[0] + e2[0], -e1[1] + e2[1], -e1[2] + e2[2]);
    const int q_mm = dir_of(-e1[0] - e2[0], -e1[1] - e2[1], -e1[2] - e2[2]);
    const T c13 = T(1.0 / 3.0), c16 = T(1.0 / 6.0);
    for_face_cells(g, sel, [&](const long* pos) {
      const size_t i = pidx(pos[0], pos[1], pos[2]);
      if (solid[i]) return;
      T f0[3];
      force_at(cidx(pos[0], pos[1], pos[2]), f0);
      auto F = [&](int q) { return f[q * np + i]; };
      const T s0 = F(0) + F(q_t1) + F(q_mt1) + F(q_t2) + F(q_mt2) + F(q_pp) + F(q_pm) + F(q_mp) + F(q_mm);
      const T sneg = F(L.opp[q_n]) + F(L.opp[q_p1]) + F(L.opp[q_m1]) + F(L.opp[q_p2]) + F(L.opp[q_m2]);
      const T closure = s0 + two * sneg + one;
      T r, vn, vt1, vt2;
      const T* uc = (vel && prof) ? prof + 3 * coord_ : ub;  // kernels.rs:1164-1180
      zh_solve(vel, uc, rho_bc, closure, f0, a, t1, t2, nsign, r, vn, vt1, vt2);
      const T qt1 = F(q_t1) - F(q_mt1) + F(q_pp) + F(q_pm) - F(q_mp) - F(q_mm);
      const T qt2 = F(q_t2) - F(q_mt2) + F(q_pp) - F(q_pm) + F(q_mp) - F(q_mm);
      const T n1 = c13 * r * vt1 - half * qt1;
      const T n2 = c13 * r * vt2 - half * qt2;
      f[q_n * np + i] = F(L.opp[q_n]) + c13 * r * vn;
      f[q_p1 * np + i] = F(L.opp[q_p1]) + c16 * r * (vn + vt1) + n1;
      f[q_m1 * np + i] = F(L.opp[q_m1]) + c16 * r * (vn - vt1) - n1;
      f[q_p2 * np + i] = F(L.opp[q_p2]) + c16 * r * (vn + vt2) + n2;
      f[q_m2 * np + i] = F(L.opp[q_m2]) + c16 * r * (vn - vt2) - n2;
    });
  }
  // density and raw (pre-force) velocity of a Zou-He face cell
  // (kernels.rs:1154-1197 / 1017-1060, g = 0)
  void zh_solve(bool vel, const T* u, T rho_bc, T closure, const T f0[3], int a, int t1, int t2, T nsign, T& r,
                T& vn, T& vt1, T& vt2) const {
    const T one = T(1), half = T(0.5);
    if (f0[0] == T(0) && f0[1] == T(0) && f0[2] == T(0)) {
      if (vel) {
        const T un = u[a] * nsign;
        r = closure / (one - un);
        vn = un;
        vt1 = u[t1];
        vt2 = u[t2];
      } else {
        r = rho_bc;
        vn = one - closure / rho_bc;
        vt1 = T(0);
        vt2 = T(0);
      }
      return;
    }
    const T f0n = f0[a] * nsign, gn = T(0);
    if (vel) {
      const T un_phys = u[a] * nsign;
      r = (closure - half * f0n) / (one - un_phys + half * gn);
      const T inv_r = one / r;
      vn = un_phys - half * (f0n * inv_r + gn);
      vt1 = u[t1] - half * (f0[t1] * inv_r + T(0));
      vt2 = u[t2] - half * (f0[t2] * inv_r + T(0));
    } else {
      const T inv_r = one / rho_bc;
      r = rho_bc;
      vn = one - closure * inv_r;
      vt1 = -half * (f0[t1] * inv_r + T(0));
      vt2 = -half * (f0[t2] * inv_r + T(0));
    }
  }
  // outflow_face_selected, kernels.rs:1221-1248
  void outflow_face(int face, const Sel& sel) {
    const FaceGeo g = face_geo(face);
    for_face_cells(g, sel, [&](const long* pos) {
      const size_t i = pidx(pos[0], pos[1], pos[2]);
      const size_t j = pidx(pos[0] + g.nin[0], pos[1] + g.nin[1], pos[2] + g.nin[2]);
      if (solid[i] || solid[j]) return;
      for (int k = 0; k < g.nunk; ++k) f[g.unk[k] * np + i] = f[g.unk[k] * np + j];
    });
  }
  // convective_face_selected, kernels.rs:1254-1299
  void convective_face(int face, T lam, const Sel& sel) {
    const FaceGeo g = face_geo(face);
    const T inv = T(1) / (T(1) + lam);
    double ws = 0.0;
    for (int k = 0; k < g.nunk; ++k) ws += L.w[g.unk[k]];
    const T wsum = T(ws);
    for_face_cells(g, sel, [&](const long* pos) {
      const size_t i = pidx(pos[0], pos[1], pos[2]);
      const size_t j = pidx(pos[0] + g.nin[0], pos[1] + g.nin[1], pos[2] + g.nin[2]);
      if (solid[i] || solid[j]) return;
      for (int k = 0; k < g.nunk; ++k) {
        const int q = g.unk[k];
        const T prev = f[q * np + i];
        f[q * np + i] = (prev + lam * f[q * np + j]) * inv;
      }
      T di = T(0), dj = T(0);
      for (int q = 0; q < L.Q; ++q) {
        di = di + f[q * np + i];
        dj = dj + f[q * np + j];
      }
      const T corr = dj - di;
      for (int k = 0; k < g.nunk; ++k) {
        const int q = g.unk[k];
        f[q * np + i] = f[q * np + i] + corr * T(L.w[q]) / wsum;
      }
    });
  }
  void apply_rule(int face, const Rule& bc, const Sel& sel, const T* prof = nullptr) {
    if (bc.kind == 0) {  // a Closed patch: impermeable lid, zero-velocity Zou-He
      Rule z{1, {T(0), T(0), T(0)}, T(0), T(0)};
      zou_he_face(face, z, sel);
    } else if (bc.kind == 1 || bc.kind == 2) {
      zou_he_face(face, bc, sel, prof);
    } else if (bc.kind == 3) {
      outflow_face(face, sel);
    } else if (bc.kind == 4) {
      convective_face(face, bc.uconv, sel);
    }
  }
  // apply_open_faces_impl, backend.rs:602-766 (single part: the patch rects
  // are already local; no inlet profiles on this path): the faces in
  // Face::ALL order; on each, the base rule over the cells outside every
  // patch (zero-velocity Zou-He when the base is closed and the face has
  // patches), then each patch's rule over its rect, in patch order
  void apply_open_faces() {
    for (int face = 0; face < 6; ++face) {
      std::vector<std::array<long, 4>> ex;
      std::vector<const Patch*> mine;
      for (const Patch& pt : patches)
        if (pt.face == face) {
          mine.push_back(&pt);
          ex.push_back({pt.lo[0], pt.lo[1], pt.hi[0], pt.hi[1]});
        }
      const int k = face_kind[face];
      if (!k && mine.empty()) continue;
      Sel excl;
      excl.kind = 2;
      excl.ex = &ex;
      Rule base{k, {face_u[face][0], face_u[face][1], face_u[face][2]}, face_rho[face], face_uconv[face]};
      // kind 0 here: the closed base with patches; the inlet profile only
      // on the face's own velocity rule (backend.rs:664-675)
      apply_rule(face, base, excl, profile[face].empty() ? nullptr : profile[face].data());
      for (const Patch* pt : mine) {
        Sel r;
        r.kind = 1;
        r.lo[0] = pt->lo[0];
        r.lo[1] = pt->lo[1];
        r.hi[0] = pt->hi[0];
        r.hi[1] = pt->hi[1];
        apply_rule(face, pt->bc, r);
      }
    }
  }
  // apply_volume_sources_impl, backend.rs:799-856 (single part: global ==
  // local coordinates), sources in order
  void apply_volume_sources() {
    const T one = T(1), three = T(3.0), f45 = T(4.5), f15 = T(1.5);
    for (const Src& src : sources) {
      const long count = (src.hi[0] - src.lo[0] + 1) * (src.hi[1] - src.lo[1] + 1) * (src.hi[2] - src.lo[2] + 1);
      const T q_cell = src.q_lu / T((double)count);
      T u[3] = {T(0), T(0), T(0)};
      if (src.kind == 1)
        for (int a = 0; a < 3; ++a) u[a] = src.u[a];
      T usq = u[0] * u[0];
      for (int a = 1; a < 3; ++a) usq = usq + u[a] * u[a];
      for (long z = src.lo[2]; z <= src.hi[2]; ++z)
        for (long y = src.lo[1]; y <= src.hi[1]; ++y)
          for (long x = src.lo[0]; x <= src.hi[0]; ++x) {
            const size_t i = pidx(x, y, z);
            if (solid[i]) continue;
            for (int q = 0; q < L.Q; ++q) {
              T cu = T((double)L.c[q][0]) * u[0];
              for (int a = 1; a < 3; ++a) cu = cu + T((double)L.c[q][a]) * u[a];
              const T delta = T(L.w[q]) * q_cell * (one + three * cu + f45 * cu * cu - f15 * usq);
              f[q * np + i] = f[q * np + i] + delta;
            }
          }
    }
  }
  bool any_open() const {
    for (int k : face_kind)
      if (k) return true;
    return !patches.empty();
  }

  // Backend::run_span default body, backend.rs:266-328 (single part,
  // single pass): collide -> exchange_f -> stream -> swap -> open faces ->
  // moments.  Sources beyond a non-periodic face are skipped by the stream,
  // so those slots keep the other buffer's previous post-collide values.
  void hydro_step() {
    collide();
    wrap_halo(f, L.Q, 1);
    stream();
    std::swap(f, ftmp);
    if (any_open()) apply_open_faces();
    if (!sources.empty()) apply_volume_sources();
    update_moments();
  }

  // Solver::step, solver.rs:2237-2280 (hydro, then scalars).
  void step() {
    hydro_step();
    time += 1;
    step_scalars();
  }

  // reduce_impl, backend.rs:920-970
  double reduce(int kind, int axis) const {
    double acc = 0.0;
    for (long z = 0; z < nz; ++z)
      for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) {
          size_t i = pidx(x, y, z), c = cidx(x, y, z);
          if (solid[i]) continue;
          if (kind == 0) {
            acc += 1.0;
          } else if (kind == 1) {
            for (int q = 0; q < L.Q; ++q) acc += (double)f[q * np + i];
          } else {
            double m = 0.0;
            for (int q = 0; q < L.Q; ++q) m += (double)L.c[q][axis] * (double)f[q * np + i];
            double fa = has_ff ? (double)force[axis] + ((double)ff[3 * c + axis] + 0.0) : (double)force[axis] + 0.0;
            acc += m + 0.5 * fa;
          }
        }
    return acc;
  }

  bool ff_uniform() const {  // solver.rs:2058-2073
    if (!has_ff || nc == 0) return false;
    for (size_t c = 1; c < nc; ++c)
      for (int a = 0; a < 3; ++a)
        if (ff[3 * c + a] != ff[a]) return false;
    return true;
  }
  // set_body_force_field, solver.rs:2776-2801
  void set_force_field(const double* v) {
    ff.assign(nc * 3, T(0));
    for (size_t k = 0; k < nc * 3; ++k) ff[k] = (T)v[k];
    has_ff = true;
    if (ff_uniform()) update_moments();
  }
  // clear_body_force_field, solver.rs:2805-2822
  void clear_force_field() {
    if (!has_ff) return;
    bool refresh = ff_uniform();
    has_ff = false;
    ff.clear();
    if (refresh) update_moments();
  }

  // equilibrium(), kernels.rs:103-121
  void equilibrium(T r, const T u[3], T* feq) const {
    const T three = T(3.0), f45 = T(4.5), f15 = T(1.5);
    T usq = u[0] * u[0];
    for (int d = 1; d < 3; ++d) usq = usq + u[d] * u[d];
    T drho = r - T(1);
    for (int q = 0; q < L.Q; ++q) {
      T cu = cr[q][0] * u[0];
      for (int d = 1; d < 3; ++d) cu = cu + cr[q][d] * u[d];
      feq[q] = wr[q] * (drho + r * (three * cu + f45 * cu * cu - f15 * usq));
    }
  }

  // strain_rate_at, solver.rs:5393-5447: S = -(Pi_neq + Pi_force) /
  // (2 cs^2 rho tau) from the stored populations and cached moments, with
  // Pi_neq = sum_q c c (f_q - feq_q(rho, u)), Pi_force = u F (diagonal) and
  // (u_a F_b + u_b F_a)/2 (off-diagonal), F = force_at (the current field)
  // and tau = 1/omega_p (the global rate, also under an omega field).
  // Order [Sxx, Syy, Szz, Sxy, Sxz, Syz]; solid cells give zeros.
  void strain_at(long x, long y, long z, T s[6]) const {
    for (int k = 0; k < 6; ++k) s[k] = T(0);
    const size_t pi = pidx(x, y, z);
    if (solid[pi]) return;
    const size_t c = cidx(x, y, z);
    const T r = rho[c];
    const T u[3] = {ux[c], uy[c], uz[c]};
    T feq[27];
    equilibrium(r, u, feq);
    for (int q = 0; q < L.Q; ++q) {
      const T fn = f[q * np + pi] - feq[q];
      const T cx = cr[q][0], cy = cr[q][1], cz = cr[q][2];
      s[0] = s[0] + cx * cx * fn;
      s[1] = s[1] + cy * cy * fn;
      s[2] = s[2] + cz * cz * fn;
      s[3] = s[3] + cx * cy * fn;
      s[4] = s[4] + cx * cz * fn;
      s[5] = s[5] + cy * cz * fn;
    }
    T fv[3];
    force_at(c, fv);
    const T half = T(0.5);
    s[0] = s[0] + u[0] * fv[0];
    s[1] = s[1] + u[1] * fv[1];
    s[2] = s[2] + u[2] * fv[2];
    s[3] = s[3] + half * (u[0] * fv[1] + u[1] * fv[0]);
    s[4] = s[4] + half * (u[0] * fv[2] + u[2] * fv[0]);
    s[5] = s[5] + half * (u[1] * fv[2] + u[2] * fv[1]);
    const T tau_eff = T(1.0 / omega_p);
    const T scale = -(T(1) / (T(2.0 / 3.0) * r * tau_eff));
    for (int k = 0; k < 6; ++k) s[k] = s[k] * scale;
  }

  // gather_velocity_gradient, solver.rs:5490-5581: g[i][j] = du_i/dx_j.
  // Diagonal: finite differences of the cached velocity; off-diagonal: the
  // stress-based strain plus the antisymmetric part of the finite
  // differences.  Along each axis: central difference between two fluid
  // neighbours; next to a solid neighbour the one-sided second-order form
  // through the wall velocity stored at that solid (a half-way wall); a
  // first-order difference at a non-periodic domain face; zero otherwise.
  void vgrad_at(long x, long y, long z, T g[3][3]) const {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) g[a][b] = T(0);
    if (solid[pidx(x, y, z)]) return;
    T sv[6];
    strain_at(x, y, z, sv);
    const T sm[3][3] = {{sv[0], sv[3], sv[4]}, {sv[3], sv[1], sv[5]}, {sv[4], sv[5], sv[2]}};
    const size_t c = cidx(x, y, z);
    const T own[3] = {ux[c], uy[c], uz[c]};
    const long n[3] = {nx, ny, nz};
    // 0: no neighbour, 1: fluid, 2: solid (value = its wall velocity)
    auto nbr = [&](int a, int da, T v[3]) -> int {
      long p[3] = {x, y, z};
      p[a] += da;
      if (p[a] < 0 || p[a] >= n[a]) {
        if (!periodic[a]) return 0;
        p[a] = (p[a] + n[a]) % n[a];
      }
      const size_t pi = pidx(p[0], p[1], p[2]);
      if (solid[pi]) {
        for (int k = 0; k < 3; ++k) v[k] = wall_u[3 * pi + k];
        return 2;
      }
      const size_t cc = cidx(p[0], p[1], p[2]);
      v[0] = ux[cc];
      v[1] = uy[cc];
      v[2] = uz[cc];
      return 1;
    };
    T fd[3][3];
    const T c43 = T(4.0 / 3.0), c13 = T(1.0 / 3.0), half = T(0.5);
    for (int a = 0; a < 3; ++a) {
      T pv[3], mv[3];
      const int kp = nbr(a, 1, pv), km = nbr(a, -1, mv);
      for (int k = 0; k < 3; ++k) {
        T v = T(0);
        if (kp == 1 && km == 1)
          v = (pv[k] - mv[k]) * half;
        else if (kp == 1 && km == 2)
          v = -c43 * mv[k] + own[k] + c13 * pv[k];
        else if (kp == 2 && km == 1)
          v = c43 * pv[k] - own[k] - c13 * mv[k];
        else if (kp == 1 && km == 0)
          v = pv[k] - own[k];
        else if (kp == 0 && km == 1)
          v = own[k] - mv[k];
        fd[k][a] = v;
      }
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) g[a][b] = a == b ? fd[a][b] : sm[a][b] + half * (fd[a][b] - fd[b][a]);
  }

  // WaleLes::update, les.rs:120-211: nu_t = (Cw Delta)^2 (Sd:Sd)^(3/2) /
  // ((S:S)^(5/2) + (Sd:Sd)^(5/4)) with Delta = 1 and 0/0 -> 0, optionally
  // clipped to (tau_eff_max - tau0)/3, then omega = 1/(3 (nu0 + nu_t) + 1/2)
  // for every cell (solid cells get the laminar rate).  Returns the clipped
  // cell count; *max_nut = the largest raw nu_t.
  long wale(T cw, double tau_eff_max, T* omega_out, T* nut_out, T* max_nut) const {
    const T cw_delta_sq = (cw * T(1)) * (cw * T(1));
    const T nu0 = T(nu);
    const T tau0 = T(0.5) + T(3.0) * nu0;
    const bool clip = tau_eff_max > 0.0;
    const T nut_max = clip ? (T(tau_eff_max) - tau0) / T(3.0) : T(0);
    long clipped = 0;
    T mx = T(0);
    for (long z = 0; z < nz; ++z)
      for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) {
          T g[3][3];
          vgrad_at(x, y, z, g);
          T s[3][3], g2[3][3], sd[3][3];
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              s[a][b] = T(0.5) * (g[a][b] + g[b][a]);
              g2[a][b] = T(0);
              for (int k = 0; k < 3; ++k) g2[a][b] = g2[a][b] + g[a][k] * g[k][b];
            }
          const T tr = g2[0][0] + g2[1][1] + g2[2][2];
          T ss = T(0), sdsd = T(0);
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              sd[a][b] = T(0.5) * (g2[a][b] + g2[b][a]);
              if (a == b) sd[a][b] = sd[a][b] - tr / T(3.0);
            }
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              ss = ss + s[a][b] * s[a][b];
              sdsd = sdsd + sd[a][b] * sd[a][b];
            }
          const T denom = std::pow(ss, T(2.5)) + std::pow(sdsd, T(1.25));
          T nut = denom > T(0) ? cw_delta_sq * std::pow(sdsd, T(1.5)) / denom : T(0);
          mx = std::max(mx, nut);
          if (clip && nut > nut_max) {
            ++clipped;
            nut = nut_max;
          }
          const size_t c = cidx(x, y, z);
          if (nut_out) nut_out[c] = nut;
          omega_out[c] = T(1) / (T(3.0) * (nu0 + nut) + T(0.5));
        }
    *max_nut = mx;
    return clipped;
  }

  // Solver::init_with, solver.rs:2474-2623: f = feq + f_neq(grad u), then
  // update_moments.  rho_in/u_in are compact global arrays.
  void init_with(const double* rho_in, const double* u_in) {
    const T tau = (T)(3.0 * nu + 0.5);
    const T three = T(3.0), half = T(0.5);
    for (size_t c = 0; c < nc; ++c) {
      rho[c] = (T)rho_in[c];
      ux[c] = (T)u_in[3 * c + 0];
      uy[c] = (T)u_in[3 * c + 1];
      uz[c] = (T)u_in[3 * c + 2];
    }
    long dims[3] = {nx, ny, nz};
    for (long z = 0; z < nz; ++z)
      for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) {
          size_t c = cidx(x, y, z), pi = pidx(x, y, z);
          T uu[3] = {ux[c], uy[c], uz[c]};
          T feq[27];
          equilibrium(rho[c], uu, feq);
          for (int q = 0; q < L.Q; ++q) f[q * np + pi] = feq[q];
          if (solid[pi]) continue;
          auto sample = [&](const long da[3], T out[3]) -> bool {
            long g[3] = {x + da[0], y + da[1], z + da[2]};
            for (int a = 0; a < 3; ++a)
              if (g[a] < 0 || g[a] >= dims[a]) {
                if (periodic[a])
                  g[a] = (g[a] + dims[a]) % dims[a];
                else
                  return false;
              }
            if (solid[pidx(x + da[0], y + da[1], z + da[2])]) return false;
            size_t gc = cidx(g[0], g[1], g[2]);
            out[0] = (T)u_in[3 * gc + 0];
            out[1] = (T)u_in[3 * gc + 1];
            out[2] = (T)u_in[3 * gc + 2];
            return true;
          };
          T grad[3][3];
          for (int a = 0; a < 3; ++a) {
            long dp[3] = {0, 0, 0}, dm[3] = {0, 0, 0};
            dp[a] = 1;
            dm[a] = -1;
            T pv[3], mv[3];
            bool hp = sample(dp, pv), hm = sample(dm, mv);
            for (int b = 0; b < 3; ++b) {
              if (hp && hm)
                grad[a][b] = (pv[b] - mv[b]) * half;
              else if (hp)
                grad[a][b] = pv[b] - uu[b];
              else if (hm)
                grad[a][b] = uu[b] - mv[b];
              else
                grad[a][b] = T(0);
            }
          }
          T div = grad[0][0];
          for (int a = 1; a < 3; ++a) div = div + grad[a][a];
          for (int q = 0; q < L.Q; ++q) {
            const T* cq = cr[q];
            T ccgu = cq[0] * cq[0] * grad[0][0];
            for (int a = 0; a < 3; ++a)
              for (int b = a; b < 3; ++b) {
                if (a == 0 && b == 0) continue;
                if (a == b)
                  ccgu = ccgu + cq[a] * cq[a] * grad[a][a];
                else
                  ccgu = ccgu + cq[a] * cq[b] * (grad[a][b] + grad[b][a]);
              }
            T fneq = -wr[q] * rho[c] * tau * (three * ccgu - div);
            f[q * np + pi] = f[q * np + pi] + fneq;
          }
        }
    update_moments();
  }

  // ---- rotating IBM, solver.rs:3016-3282 -----------------------------------
  struct SP {
    long x, y, z;
    double w;
  };
  std::vector<long> ibm_slot;  // IBM: compact index of each stencil cell, -1 elsewhere
  // marker_stencil, rotating_ibm.rs:163-224.  `d` is the lattice's
  // dimension: axes a >= d take the single point (0, 1.0) (rotating_ibm.rs:
  // 176-179), the reference's 2D case, used here by one-cell-deep boxes
  // that restate the reference's D2Q9 IBM tests.
  static void marker_stencil(const double p[3], const int dims[3], int radius, std::vector<SP>& out, int d = 3) {
    std::vector<std::pair<long, double>> axes[3];
    for (int a = 0; a < 3; ++a) {
      if (a >= d) {
        axes[a].push_back({0, 1.0});
        continue;
      }
      long lo = (long)std::floor(p[a]) - radius + 1;
      long hi = (long)std::floor(p[a]) + radius;
      for (long i = lo; i <= hi; ++i) {
        if (i < 0 || i >= dims[a]) continue;
        double rr = std::fabs(p[a] - (double)i);
        double w;
        if (radius == 1) {
          w = std::max(1.0 - rr, 0.0);
        } else {
          if (rr < 0.5)
            w = 0.75 - rr * rr;
          else if (rr < 1.5)
            w = 0.5 * (1.5 - rr) * (1.5 - rr);
          else
            w = 0.0;
        }
        if (w != 0.0) axes[a].push_back({i, w});
      }
    }
    out.clear();
    for (auto& ax : axes[0])
      for (auto& ay : axes[1])
        for (auto& az : axes[2]) out.push_back({ax.first, ay.first, az.first, ax.second * ay.second * az.second});
  }

  static double norm3(const double v[3]) { return std::sqrt(0.0 + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

  // diag layout: iterations, slip_max, slip_rms, slip_max_rel, slip_rms_rel,
  // torque[3], fluid_force[3], marker_force[3], momentum_error_rel
  int apply_ibm(int nm, const double* pos, const double* weight, const double center[3], const double omega[3],
                int max_iter, double tol, int radius, double relax, double* diag_out, int stencil_dim = 3) {
    int dims[3] = {nx, ny, nz};
    // Only the stencil cells enter the update: u_now, du, spread and the
    // kernel sums live in compact arrays over the distinct stencil cells
    // (slot[gi] = compact index), with the reference's per-cell arithmetic
    // and accumulation order (markers in order, stencil points in order).
    std::vector<std::vector<SP>> st(nm);
    for (int m = 0; m < nm; ++m) {
      marker_stencil(&pos[3 * m], dims, radius, st[m], stencil_dim);
      if (st[m].empty()) return -1;
    }
    if (ibm_slot.size() != nc) ibm_slot.assign(nc, -1);
    std::vector<size_t> cells;
    for (int m = 0; m < nm; ++m)
      for (auto& sp : st[m]) {
        const size_t gi = cidx(sp.x, sp.y, sp.z);
        if (ibm_slot[gi] < 0) {
          ibm_slot[gi] = (long)cells.size();
          cells.push_back(gi);
        }
      }
    const size_t nk = cells.size();
    struct Clear {  // leaves slot all -1 on every return path
      std::vector<long>& slot;
      const std::vector<size_t>& cells;
      ~Clear() {
        for (size_t gi : cells) slot[gi] = -1;
      }
    } clear_slots{ibm_slot, cells};
    std::vector<double> u_now(3 * nk, 0.0);
    std::vector<uint8_t> sol(nk, 0);
    std::vector<double> rho_k(nk, 1.0);
    for (size_t k = 0; k < nk; ++k) {
      const size_t gi = cells[k];
      const long x = (long)(gi % (size_t)nx), y = (long)((gi / (size_t)nx) % (size_t)ny),
                 z = (long)(gi / ((size_t)nx * ny));
      const size_t pi = pidx(x, y, z);
      sol[k] = solid[pi];
      rho_k[k] = (double)rho[gi];
      if (sol[k]) continue;
      double mom[3] = {0.0, 0.0, 0.0};
      for (int q = 0; q < L.Q; ++q) {
        double fq = (double)f[q * np + pi];
        for (int a = 0; a < 3; ++a) mom[a] += (double)L.c[q][a] * fq;
      }
      double frc[3] = {(double)force[0], (double)force[1], (double)force[2]};
      if (has_ff)
        for (int a = 0; a < 3; ++a) frc[a] += (double)ff[3 * gi + a];
      double inv_rho = 1.0 / std::max(rho_k[k], 1.0e-30);
      for (int a = 0; a < 3; ++a) u_now[3 * k + a] = (mom[a] + 0.5 * frc[a]) * inv_rho;
    }
    std::vector<double> du(3 * nk, 0.0), spread(3 * nk, 0.0);
    double iterations = 0, slip_max = 0, slip_rms = 0, slip_max_rel = 0, slip_rms_rel = 0;
    double torque[3] = {0, 0, 0}, fluid_force[3] = {0, 0, 0}, marker_force[3] = {0, 0, 0};
    double mom_err = 0;
    auto target = [&](const double* p, double out[3]) {  // rotating_ibm.rs:89-101
      double r[3] = {p[0] - center[0], p[1] - center[1], p[2] - center[2]};
      out[0] = omega[1] * r[2] - omega[2] * r[1];
      out[1] = omega[2] * r[0] - omega[0] * r[2];
      out[2] = omega[0] * r[1] - omega[1] * r[0];
    };
    double max_target = 0.0;
    for (int m = 0; m < nm; ++m) {
      double t[3];
      target(&pos[3 * m], t);
      max_target = std::max(max_target, norm3(t));
    }
    double scale = std::max(max_target, 1.0);
    auto K = [&](const SP& sp) { return (size_t)ibm_slot[cidx(sp.x, sp.y, sp.z)]; };
    std::vector<double> ksum(nk, 0.0);
    for (int m = 0; m < nm; ++m)
      for (auto& sp : st[m]) {
        size_t k = K(sp);
        if (sol[k]) continue;
        ksum[k] += sp.w;
      }
    std::vector<double> imp(3 * (size_t)nm);
    for (int iter = 0; iter < max_iter; ++iter) {
      std::fill(imp.begin(), imp.end(), 0.0);
      for (int m = 0; m < nm; ++m) {
        double t[3];
        target(&pos[3 * m], t);
        double um[3] = {0, 0, 0}, mob = 0.0;
        for (auto& sp : st[m]) {
          size_t k = K(sp);
          if (sol[k]) continue;
          double rc = rho_k[k];
          for (int a = 0; a < 3; ++a) um[a] += sp.w * (u_now[3 * k + a] + du[3 * k + a]);
          mob += sp.w * ksum[k] / std::max(rc, 1.0e-30);
        }
        if (mob == 0.0) continue;
        for (int a = 0; a < 3; ++a) imp[3 * m + a] = relax * (t[a] - um[a]) / mob;
      }
      for (int m = 0; m < nm; ++m) {
        double rep[3] = {0, 0, 0};
        for (auto& sp : st[m]) {
          size_t k = K(sp);
          if (sol[k]) continue;
          double cf[3] = {imp[3 * m + 0] * sp.w, imp[3 * m + 1] * sp.w, imp[3 * m + 2] * sp.w};
          for (int a = 0; a < 3; ++a) {
            spread[3 * k + a] += cf[a];
            fluid_force[a] += cf[a];
            rep[a] += cf[a];
          }
          double inv_rho = 1.0 / std::max(rho_k[k], 1.0e-30);
          for (int a = 0; a < 3; ++a) du[3 * k + a] += cf[a] * inv_rho;
        }
        for (int a = 0; a < 3; ++a) marker_force[a] += rep[a];
        const double* p = &pos[3 * m];
        double r[3] = {p[0] - center[0], p[1] - center[1], p[2] - center[2]};
        double tq[3] = {r[1] * rep[2] - r[2] * rep[1], r[2] * rep[0] - r[0] * rep[2], r[0] * rep[1] - r[1] * rep[0]};
        for (int a = 0; a < 3; ++a) torque[a] -= tq[a];
      }
      iterations = iter + 1;
      slip_max = 0.0;
      double ssq = 0.0, wsum = 0.0;
      for (int m = 0; m < nm; ++m) {
        double t[3];
        target(&pos[3 * m], t);
        double um[3] = {0, 0, 0};
        bool active = false;
        for (auto& sp : st[m]) {
          size_t k = K(sp);
          if (sol[k]) continue;
          active = true;
          for (int a = 0; a < 3; ++a) um[a] += sp.w * (u_now[3 * k + a] + du[3 * k + a]);
        }
        if (!active) continue;
        double s[3] = {t[0] - um[0], t[1] - um[1], t[2] - um[2]};
        double sm = norm3(s);
        slip_max = std::max(slip_max, sm);
        ssq += weight[m] * sm * sm;
        wsum += weight[m];
      }
      slip_rms = wsum > 0.0 ? std::sqrt(ssq / wsum) : 0.0;
      slip_max_rel = slip_max / scale;
      slip_rms_rel = slip_rms / scale;
      double err[3] = {fluid_force[0] - marker_force[0], fluid_force[1] - marker_force[1],
                       fluid_force[2] - marker_force[2]};
      mom_err = norm3(err) / std::max(norm3(marker_force), 1.0e-12);
      if (slip_max_rel <= tol) break;
    }
    // force_field += T(spread): every cell, as the reference adds the whole
    // spread field (zero off the stencils)
    if (!has_ff) {
      ff.assign(nc * 3, T(0));
      has_ff = true;
    }
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)nc; ++c) {
      const long k = ibm_slot[c];
      for (int a = 0; a < 3; ++a) ff[3 * c + a] = ff[3 * c + a] + (T)(k >= 0 ? spread[3 * k + a] : 0.0);
    }
    double d[15] = {iterations,    slip_max,      slip_rms,        slip_max_rel,    slip_rms_rel,
                    torque[0],     torque[1],     torque[2],       fluid_force[0],  fluid_force[1],
                    fluid_force[2], marker_force[0], marker_force[1], marker_force[2], mom_err};
    std::memcpy(diag_out, d, sizeof(d));
    return 0;
  }

  // ---- checkpoint rank file (write_part_checkpoint, solver.rs:4536-4630) ----
  // The reference's v3 layout from this solver's own buffers: header, section
  // table, F_PRIMARY (f), STALE_STASH (ftmp), MOMENTS, SOLID, FORCE_FIELD (if
  // a field is set), SCALAR_FIELDS (if any; names "scalar<k>"), FNV-1a 64
  // trailer of the payload.  Returns 0, or -1 if the file cannot be written.
  int write_checkpoint(const char* path) const {
    auto u32 = [](std::vector<uint8_t>& b, uint32_t v) { b.insert(b.end(), (uint8_t*)&v, (uint8_t*)&v + 4); };
    auto u64 = [](std::vector<uint8_t>& b, uint64_t v) { b.insert(b.end(), (uint8_t*)&v, (uint8_t*)&v + 8); };
    auto reals = [](std::vector<uint8_t>& b, const T* v, size_t n) {
      b.insert(b.end(), (const uint8_t*)v, (const uint8_t*)(v + n));
    };
    std::vector<std::pair<uint32_t, std::vector<uint8_t>>> sec;
    std::vector<uint8_t> b;
    reals(b, f.data(), f.size());
    sec.push_back({1, b});
    b.clear();
    reals(b, ftmp.data(), ftmp.size());
    sec.push_back({2, b});
    b.clear();
    for (const std::vector<T>* m : {&rho, &ux, &uy, &uz}) reals(b, m->data(), m->size());
    sec.push_back({3, b});
    b.clear();
    for (uint8_t v : solid) b.push_back(v ? 1 : 0);
    sec.push_back({4, b});
    if (has_ff) {
      b.clear();
      reals(b, ff.data(), ff.size());
      sec.push_back({5, b});
    }
    if (!scalars.empty()) {
      b.clear();
      u32(b, (uint32_t)scalars.size());
      for (size_t k = 0; k < scalars.size(); ++k) {
        const ScalarDist<T>& sc = scalars[k];
        const std::string name = "scalar" + std::to_string(k);
        u32(b, (uint32_t)name.size());
        b.insert(b.end(), name.begin(), name.end());
        const double D = sc.diffusivity;
        b.insert(b.end(), (const uint8_t*)&D, (const uint8_t*)&D + 8);
        u64(b, sc.h.size());
        reals(b, sc.h.data(), sc.h.size());
        u64(b, sc.htmp.size());
        reals(b, sc.htmp.data(), sc.htmp.size());
        u64(b, sc.conc.size());
        reals(b, sc.conc.data(), sc.conc.size());
      }
      sec.push_back({6, b});
    }
    std::vector<uint8_t> payload, head;
    head.insert(head.end(), (const uint8_t*)"LBMKPT\0\0", (const uint8_t*)"LBMKPT\0\0" + 8);
    u32(head, 3);
    head.push_back(0);
    head.push_back(sizeof(T) == 4 ? 0 : 1);
    const uint16_t lat = L.Q == 19 ? 319 : 327, q16 = (uint16_t)L.Q, d16 = 3;
    head.insert(head.end(), (const uint8_t*)&lat, (const uint8_t*)&lat + 2);
    head.insert(head.end(), (const uint8_t*)&q16, (const uint8_t*)&q16 + 2);
    head.insert(head.end(), (const uint8_t*)&d16, (const uint8_t*)&d16 + 2);
    u64(head, np);
    u64(head, nc);
    u32(head, (uint32_t)sec.size());
    for (auto& e : sec) {
      u32(head, e.first);
      u64(head, payload.size());
      u64(head, e.second.size());
      payload.insert(payload.end(), e.second.begin(), e.second.end());
    }
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint8_t v : payload) {
      h ^= v;
      h *= 0x100000001b3ull;
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) return -1;
    bool ok = fwrite(head.data(), 1, head.size(), fp) == head.size();
    ok = ok && fwrite(payload.data(), 1, payload.size(), fp) == payload.size();
    ok = ok && fwrite(&h, 1, 8, fp) == 8;
    ok = fclose(fp) == 0 && ok;
    return ok ? 0 : -1;
  }

  // ---- passive scalar ADE, solver.rs:2113-2205 / scalar.rs:6-15 --------------
  void scalar_eq(const ScalarDist<T>& s, T conc, const T u[3], T* out) const {
    if (!s.d3q7) {
      const T three = T(3.0);
      for (int q = 0; q < s.L.Q; ++q) {
        T cu = (T)(double)s.L.c[q][0] * u[0] + (T)(double)s.L.c[q][1] * u[1] + (T)(double)s.L.c[q][2] * u[2];
        out[q] = (T)s.L.w[q] * conc * (T(1) + three * cu);
      }
    } else {
      // D3Q7: h_eq = w C (1 + c.u / cs^2), cs^2 = 1/4
      const T four = T(4.0);
      for (int q = 0; q < 7; ++q) {
        T cu = (T)(double)s.L.c[q][0] * u[0] + (T)(double)s.L.c[q][1] * u[1] + (T)(double)s.L.c[q][2] * u[2];
        out[q] = (T)s.L.w[q] * conc * (T(1) + four * cu);
      }
    }
  }
  int scalar_enable(double D, bool d3q7) {
    ScalarDist<T> s;
    s.diffusivity = D;
    s.d3q7 = d3q7;
    s.L = make_lat(d3q7 ? 7 : L.Q);
    s.h.assign((size_t)s.L.Q * np, T(0));
    s.htmp.assign((size_t)s.L.Q * np, T(0));
    s.conc.assign(nc, T(0));
    scalars.push_back(s);
    return (int)scalars.size() - 1;
  }
  // set_scalar_concentration, solver.rs:2910-2965
  void scalar_set(int k, const double* vals) {
    ScalarDist<T>& s = scalars[k];
    for (long z = 0; z < nz; ++z)
      for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) {
          size_t c = cidx(x, y, z), pi = pidx(x, y, z);
          s.conc[c] = (T)vals[c];
          T u[3] = {ux[c], uy[c], uz[c]};
          T heq[27];
          scalar_eq(s, (T)vals[c], u, heq);
          for (int q = 0; q < s.L.Q; ++q) {
            s.h[q * np + pi] = heq[q];
            s.htmp[q * np + pi] = T(0);
          }
        }
  }
  void step_scalars() {
    for (auto& s : scalars) {
      const int Q = s.L.Q;
      T omega = s.d3q7 ? (T)(1.0 / (4.0 * s.diffusivity + 0.5)) : (T)(1.0 / (3.0 * s.diffusivity + 0.5));
#pragma omp parallel for schedule(static)
      for (long r = 0; r < (long)ny * nz; ++r) {
        long y = r % ny, z = r / ny;
        for (long x = 0; x < nx; ++x) {
          size_t pi = pidx(x, y, z), c = cidx(x, y, z);
          if (solid[pi]) continue;
          T u[3] = {ux[c], uy[c], uz[c]};
          T heq[27];
          scalar_eq(s, s.conc[c], u, heq);
          for (int q = 0; q < Q; ++q) {
            size_t k = q * np + pi;
            s.h[k] = s.h[k] - omega * (s.h[k] - heq[q]);
          }
        }
      }
      wrap_halo(s.h, Q, 1);
      std::fill(s.htmp.begin(), s.htmp.end(), T(0));
#pragma omp parallel for schedule(static)
      for (long r = 0; r < (long)ny * nz; ++r) {
        long y = r % ny, z = r / ny;
        for (long x = 0; x < nx; ++x) {
          size_t pi = pidx(x, y, z);
          if (solid[pi]) {
            s.conc[cidx(x, y, z)] = T(0);
            continue;
          }
          for (int q = 0; q < Q; ++q) {
            size_t src = pidx(x - s.L.c[q][0], y - s.L.c[q][1], z - s.L.c[q][2]);
            T v = solid[src] ? s.h[s.L.opp[q] * np + pi] : s.h[q * np + src];
            s.htmp[q * np + pi] = v;
          }
        }
      }
      std::swap(s.h, s.htmp);
#pragma omp parallel for schedule(static)
      for (long r = 0; r < (long)ny * nz; ++r) {
        long y = r % ny, z = r / ny;
        for (long x = 0; x < nx; ++x) {
          size_t pi = pidx(x, y, z), c = cidx(x, y, z);
          if (solid[pi]) {
            s.conc[c] = T(0);
            continue;
          }
          T sum = T(0);
          for (int q = 0; q < Q; ++q) sum = sum + s.h[q * np + pi];
          s.conc[c] = sum;
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Geometry — solver.rs:1357-1406 (wall rims); geometry.rs:657-723 (tank and
// baffles), 815-920 (radial impeller, wall-velocity stamp), 922-947 helpers.
// ---------------------------------------------------------------------------
static void build_wall_rims(const int dims[3], const int is_wall[6], const double u[6][3], uint8_t* solid,
                            double* wall_u) {
  size_t n = (size_t)dims[0] * dims[1] * dims[2];
  std::vector<double> best(n, -1.0);
  for (size_t i = 0; i < n; ++i) {
    solid[i] = 0;
    wall_u[3 * i] = wall_u[3 * i + 1] = wall_u[3 * i + 2] = 0.0;
  }
  // Face index order XNeg=0, XPos=1, YNeg=2, YPos=3, ZNeg=4, ZPos=5;
  // application order YNeg, YPos, XNeg, XPos, ZNeg, ZPos.
  const int order[6] = {2, 3, 0, 1, 4, 5};
  for (int k = 0; k < 6; ++k) {
    int face = order[k];
    int a = face / 2;
    if (!is_wall[face]) continue;
    const double* uu = u[face];
    double speed = uu[0] * uu[0] + uu[1] * uu[1];
    speed += uu[2] * uu[2];
    long fixed = (face % 2 == 0) ? 0 : dims[a] - 1;
    int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    for (long c2 = 0; c2 < dims[t2]; ++c2)
      for (long c1 = 0; c1 < dims[t1]; ++c1) {
        long pos[3];
        pos[a] = fixed;
        pos[t1] = c1;
        pos[t2] = c2;
        size_t i = ((size_t)pos[2] * dims[1] + pos[1]) * dims[0] + pos[0];
        solid[i] = 1;
        if (speed > best[i]) {
          best[i] = speed;
          for (int d = 0; d < 3; ++d) wall_u[3 * i + d] = uu[d];
        }
      }
  }
}

static void tank_mask(const int dims[3], double dx, double diameter, double height, uint8_t* solid) {
  double radius = 0.5 * diameter;
  double cx = 0.5 * (double)dims[0] * dx, cy = 0.5 * (double)dims[1] * dx;
  for (long z = 0; z < dims[2]; ++z) {
    double z_m = ((double)z + 0.5) * dx;
    for (long y = 0; y < dims[1]; ++y)
      for (long x = 0; x < dims[0]; ++x) {
        double pxm = ((double)x + 0.5) * dx, pym = ((double)y + 0.5) * dx;
        double r = std::sqrt((pxm - cx) * (pxm - cx) + (pym - cy) * (pym - cy));
        bool rim = x == 0 || y == 0 || z == 0 || x + 1 == dims[0] || y + 1 == dims[1] || z + 1 == dims[2];
        bool outside = r >= radius || z_m >= height;
        if (rim || outside) solid[((size_t)z * dims[1] + y) * dims[0] + x] = 1;
      }
  }
}

static void baffle_mask(const int dims[3], double dx, double diameter, double height, double width,
                        double thickness, uint8_t* solid, uint8_t* baffle) {
  double radius = 0.5 * diameter;
  double cx = 0.5 * (double)dims[0] * dx, cy = 0.5 * (double)dims[1] * dx;
  double radial_inner = radius - width;
  for (long z = 1; z < dims[2] - 1; ++z) {
    double z_m = ((double)z + 0.5) * dx;
    if (z_m >= height) continue;
    for (long y = 1; y < dims[1] - 1; ++y)
      for (long x = 1; x < dims[0] - 1; ++x) {
        double pxm = ((double)x + 0.5) * dx, pym = ((double)y + 0.5) * dx;
        double rx = pxm - cx, ry = pym - cy;
        double r = std::sqrt(rx * rx + ry * ry);
        if (r >= radius || r < radial_inner) continue;
        bool nxw = std::fabs(ry) <= 0.5 * thickness;
        bool nyw = std::fabs(rx) <= 0.5 * thickness;
        if (nxw || nyw) {
          size_t i = ((size_t)z * dims[1] + y) * dims[0] + x;
          solid[i] = 1;
          if (baffle) baffle[i] = 1;
        }
      }
  }
}

// build_radial_impeller (blade angle 0 => Rushton), geometry.rs:815-906.
// Returns the marker count; fills pos (lattice units), weight, and the
// target velocity used by the wall-velocity stamp.
static int radial_impeller(const int dims[3], double dx, double vessel_diameter, double liquid_height,
                           double diameter, double clearance, int blades, double blade_angle_deg, double omega_z,
                           double thickness_fraction, int inner_kind, std::vector<double>& pos,
                           std::vector<double>& weight, std::vector<double>& vel, double center_out[3]) {
  const double TAU = 6.283185307179586;
  double cxy[2] = {0.5 * (double)dims[0] * dx, 0.5 * (double)dims[1] * dx};
  double radius = 0.5 * diameter;
  double center[3] = {cxy[0], cxy[1], clearance};
  double radial_inner = inner_kind == 0 ? 0.65 * radius : (inner_kind == 1 ? 0.35 * radius : 0.0);
  double dr = std::max(dx, radius / 20.0);
  double dz = dx;
  double tstep = dx;
  double thickness = std::max(thickness_fraction * diameter, dx);
  pos.clear();
  weight.clear();
  vel.clear();
  for (int b = 0; b < blades; ++b) {
    double th = TAU * (double)b / (double)blades;
    double er[3] = {std::cos(th), std::sin(th), 0.0};
    double et[3] = {-std::sin(th), std::cos(th), 0.0};
    double pitch = blade_angle_deg * (3.141592653589793 / 180.0);
    double z_span = blade_angle_deg == 0.0 ? thickness : thickness + (radius - radial_inner) * std::sin(pitch) * 0.18;
    long nr = (long)std::max(std::ceil((radius - radial_inner) / dr), 1.0);
    long nt = (long)std::max(std::ceil(thickness / tstep), 1.0);
    long nzs = (long)std::max(std::ceil(z_span / dz), 1.0);
    for (long ir = 0; ir <= nr; ++ir) {
      double rr = radial_inner + (radius - radial_inner) * (double)ir / (double)nr;
      for (long it = 0; it <= nt; ++it) {
        double offset_t = -0.5 * thickness + thickness * (double)it / (double)nt;
        for (long iz = 0; iz <= nzs; ++iz) {
          double offset_z = -0.5 * z_span + z_span * (double)iz / (double)nzs;
          double pitched_z = offset_z + (rr - radial_inner) * std::sin(pitch) * 0.18;
          double p[3] = {center[0] + rr * er[0] + offset_t * et[0], center[1] + rr * er[1] + offset_t * et[1],
                         center[2] + pitched_z};
          double rel_r = std::sqrt((p[0] - cxy[0]) * (p[0] - cxy[0]) + (p[1] - cxy[1]) * (p[1] - cxy[1]));
          if (rel_r >= 0.5 * vessel_diameter || p[2] <= 0.0 || p[2] >= liquid_height) continue;
          double r[3] = {p[0] - center[0], p[1] - center[1], p[2] - center[2]};
          double om[3] = {0.0, 0.0, omega_z};
          vel.push_back(om[1] * r[2] - om[2] * r[1]);
          vel.push_back(om[2] * r[0] - om[0] * r[2]);
          vel.push_back(om[0] * r[1] - om[1] * r[0]);
          pos.push_back(p[0] / dx);
          pos.push_back(p[1] / dx);
          pos.push_back(p[2] / dx);
          weight.push_back(dx * dx);
        }
      }
    }
  }
  center_out[0] = center[0] / dx;
  center_out[1] = center[1] / dx;
  center_out[2] = center[2] / dx;
  return (int)weight.size();
}

// Seeded uniform surface sampling of the blade boxes of
// build_radial_impeller (the benchmark's marker set; it has no reference
// counterpart, so this restates its specification: per blade
// N_b = max(1, round(A_b / (area dx^2))) markers, each drawing three
// SplitMix64 uniforms u1, u2, u3 from one stream: u1 A_b picks the face by
// cumulative area in the order r_in, r_out, t-, t+, z-, z+, and (u2, u3)
// place the point on it; weight A_b / N_b; markers outside the liquid are
// dropped; the axial shear (r - r_in) sin(pitch) 0.18 is applied after).
struct SplitMix64 {
  uint64_t s;
  double next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;  // 2^53
  }
};

static int impeller_surface(const int dims[3], double dx, double vessel_diameter, double liquid_height,
                            double diameter, double clearance, int blades, double blade_angle_deg, double omega_z,
                            double thickness_fraction, double area_per_marker, uint64_t seed,
                            std::vector<double>& pos, std::vector<double>& weight, std::vector<double>& vel,
                            double center_out[3]) {
  const double R = 0.5 * diameter;
  const double rin = blade_angle_deg == 0.0 ? 0.65 * R : 0.35 * R;
  const double t = std::max(thickness_fraction * diameter, dx);
  const double ang = blade_angle_deg * (3.141592653589793 / 180.0);
  const double h = blade_angle_deg == 0.0 ? t : t + (R - rin) * std::sin(ang) * 0.18;
  const double len = R - rin;
  const double face_area[6] = {t * h, t * h, len * h, len * h, len * t, len * t};
  double bound[6], total = 0.0;
  for (int f = 0; f < 6; ++f) bound[f] = (total += face_area[f]);
  const long n_per_blade = std::max(1L, (long)std::llround(total / (area_per_marker * dx * dx)));
  const double w_marker = total / (double)n_per_blade;
  const double cx = 0.5 * (double)dims[0] * dx, cy = 0.5 * (double)dims[1] * dx;
  SplitMix64 rng{seed};
  pos.clear();
  weight.clear();
  vel.clear();
  for (int b = 0; b < blades; ++b) {
    const double phi = 6.283185307179586 * (double)b / (double)blades;
    const double c = std::cos(phi), sn = std::sin(phi);
    for (long j = 0; j < n_per_blade; ++j) {
      const double a = rng.next() * total, u = rng.next(), v = rng.next();
      int f = 0;
      while (f < 5 && !(a < bound[f])) ++f;
      // (radial, tangential, axial) in the blade frame
      double r, tt, zz;
      switch (f) {
        case 0: r = rin; tt = -0.5 * t + t * u; zz = -0.5 * h + h * v; break;
        case 1: r = R; tt = -0.5 * t + t * u; zz = -0.5 * h + h * v; break;
        case 2: tt = -0.5 * t; r = rin + len * u; zz = -0.5 * h + h * v; break;
        case 3: tt = 0.5 * t; r = rin + len * u; zz = -0.5 * h + h * v; break;
        case 4: zz = -0.5 * h; r = rin + len * u; tt = -0.5 * t + t * v; break;
        default: zz = 0.5 * h; r = rin + len * u; tt = -0.5 * t + t * v; break;
      }
      const double x = cx + r * c + tt * -sn;
      const double y = cy + r * sn + tt * c;
      const double z = clearance + (zz + (r - rin) * std::sin(ang) * 0.18);
      const double rho = std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
      if (rho >= 0.5 * vessel_diameter || z <= 0.0 || z >= liquid_height) continue;
      const double rx = x - cx, ry = y - cy, rz = z - clearance;
      vel.push_back(0.0 * rz - omega_z * ry);
      vel.push_back(omega_z * rx - 0.0 * rz);
      vel.push_back(0.0 * ry - 0.0 * rx);
      pos.push_back(x / dx);
      pos.push_back(y / dx);
      pos.push_back(z / dx);
      weight.push_back(w_marker);
    }
  }
  center_out[0] = cx / dx;
  center_out[1] = cy / dx;
  center_out[2] = clearance / dx;
  return (int)weight.size();
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------
// C ABI for the Python checker (ctypes).  All arrays are compact global
// order (z, y, x) with q-major planes for populations.
// ---------------------------------------------------------------------------
struct Handle {
  int prec;
  Solver<float>* s32;
  Solver<double>* s64;
};

#define DISPATCH(h, ...)         \
  do {                           \
    if ((h)->prec == 4) {        \
      auto* S = (h)->s32;        \
      __VA_ARGS__;               \
    } else {                     \
      auto* S = (h)->s64;        \
      __VA_ARGS__;               \
    }                            \
  } while (0)

extern "C" {

int orc_lattice(int Q, int8_t* c_out, double* w_out, int* opp_out) {
  if (Q != 7 && Q != 19 && Q != 27) return -1;
  Lat L = make_lat(Q);
  for (int q = 0; q < Q; ++q) {
    for (int a = 0; a < 3; ++a) c_out[3 * q + a] = L.c[q][a];
    w_out[q] = L.w[q];
    opp_out[q] = L.opp[q];
  }
  return L.npairs;
}

void orc_omegas(int collision, double nu, double magic, double* op, double* om) {
  double tau = 3.0 * nu + 0.5;
  *op = 1.0 / tau;
  if (collision == 1) {
    double lam_p = tau - 0.5;
    *om = 1.0 / (magic / lam_p + 0.5);
  } else {
    *om = *op;
  }
}

void orc_cm_transform(int Q, double* raw_from_pop, double* pop_from_raw, uint8_t* basis) {
  const MomentTransform& t = transform_for(Q);
  for (int m = 0; m < Q; ++m) {
    for (int q = 0; q < Q; ++q) {
      raw_from_pop[m * Q + q] = t.raw_from_pop[m][q];
      pop_from_raw[m * Q + q] = t.pop_from_raw[m][q];
    }
    for (int a = 0; a < 3; ++a) basis[3 * m + a] = t.basis[m][a];
  }
}

// Flush subnormal results and operands to zero on every OpenMP worker and
// the calling thread (x86 MXCSR FTZ|DAZ).  Long f64 runs otherwise crawl
// once decaying deviations and tracer tails underflow below 2.2e-308, where
// x86 arithmetic takes a microcode path; values that small cannot change a
// quantity of interest.  Off unless a caller asks (fixture generation).
void orc_set_cm_factorised(int on) { g_cm_factorised = on != 0; }

int orc_write_checkpoint(void* p, const char* path) {
  Handle* h = (Handle*)p;
  int r = 0;
  DISPATCH(h, r = S->write_checkpoint(path));
  return r;
}

// threads an OpenMP region of the oracle runs on (OMP_NUM_THREADS, else
// the cores the process may use)
int orc_max_threads() { return omp_get_max_threads(); }

void orc_flush_subnormals(int on) {
  const unsigned bits = 0x8040u;  // FTZ (bit 15) | DAZ (bit 6)
#pragma omp parallel
  {
    const unsigned c = _mm_getcsr();
    _mm_setcsr(on ? (c | bits) : (c & ~bits));
  }
  const unsigned c = _mm_getcsr();
  _mm_setcsr(on ? (c | bits) : (c & ~bits));
}

void* orc_new(int prec, int Q, const int dims[3], const int periodic[3], int collision, double nu, double magic,
              double omega_shear, const double force[3], const uint8_t* solid, const double* wall_u) {
  Handle* h = new Handle{prec, nullptr, nullptr};
  if (prec == 4) {
    h->s32 = new Solver<float>();
    h->s32->init(Q, dims, periodic, collision, nu, magic, omega_shear, force);
    h->s32->set_masks(solid, wall_u);
  } else {
    h->s64 = new Solver<double>();
    h->s64->init(Q, dims, periodic, collision, nu, magic, omega_shear, force);
    h->s64->set_masks(solid, wall_u);
  }
  return h;
}

void orc_free(void* p) {
  Handle* h = (Handle*)p;
  delete h->s32;
  delete h->s64;
  delete h;
}

void orc_step(void* p, int n) {
  Handle* h = (Handle*)p;
  DISPATCH(h, for (int k = 0; k < n; ++k) S->step());
}

void orc_collide_only(void* p) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S->collide());
}

void orc_update_moments(void* p) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S->update_moments());
}

// Populations, q-major compact core planes (deviation form).
// Population exchange in compact order f[q][cell] (f64 or f32 on the host
// side), one OpenMP thread per (q, z) plane row block.
extern "C++" {
template <typename H>
static void get_f_any(void* p, H* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    const long Q = S->L.Q, nz = S->nz;
    _Pragma("omp parallel for collapse(2) schedule(static)")
    for (long q = 0; q < Q; ++q)
      for (long z = 0; z < nz; ++z)
        for (long y = 0; y < S->ny; ++y)
          for (long x = 0; x < S->nx; ++x)
            out[q * S->nc + S->cidx(x, y, z)] = (H)S->f[q * S->np + S->pidx(x, y, z)];
  });
}

template <typename H>
static void set_f_any(void* p, const H* in) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    const long Q = S->L.Q, nz = S->nz;
    _Pragma("omp parallel for collapse(2) schedule(static)")
    for (long q = 0; q < Q; ++q)
      for (long z = 0; z < nz; ++z)
        for (long y = 0; y < S->ny; ++y)
          for (long x = 0; x < S->nx; ++x)
            S->f[q * S->np + S->pidx(x, y, z)] = (TT)in[q * S->nc + S->cidx(x, y, z)];
  });
}
}  // extern "C++"

void orc_get_f(void* p, double* out) { get_f_any(p, out); }
void orc_set_f(void* p, const double* in) { set_f_any(p, in); }
// f32 host arrays: an f32 oracle exchanges its populations without the f64
// copy (512^3 D3Q27: 14.5 GB instead of 29 GB)
void orc_get_f32(void* p, float* out) { get_f_any(p, out); }
void orc_set_f32(void* p, const float* in) { set_f_any(p, in); }

void orc_get_moments(void* p, double* rho, double* ux, double* uy, double* uz) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    for (size_t c = 0; c < S->nc; ++c) {
      rho[c] = (double)S->rho[c];
      ux[c] = (double)S->ux[c];
      uy[c] = (double)S->uy[c];
      uz[c] = (double)S->uz[c];
    }
  });
}

void orc_set_moments(void* p, const double* rho, const double* ux, const double* uy, const double* uz) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    for (size_t c = 0; c < S->nc; ++c) {
      S->rho[c] = (TT)rho[c];
      S->ux[c] = (TT)ux[c];
      S->uy[c] = (TT)uy[c];
      S->uz[c] = (TT)uz[c];
    }
  });
}

double orc_reduce(void* p, int kind, int axis) {
  Handle* h = (Handle*)p;
  double r = 0;
  DISPATCH(h, r = S->reduce(kind, axis));
  return r;
}

void orc_init_with(void* p, const double* rho, const double* u) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S->init_with(rho, u));
}

// equilibrium() for one (rho, u); out has Q entries.
void orc_equilibrium(void* p, double r, const double u[3], double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    TT uu[3] = {(TT)u[0], (TT)u[1], (TT)u[2]};
    TT feq[27];
    S->equilibrium((TT)r, uu, feq);
    for (int q = 0; q < S->L.Q; ++q) out[q] = (double)feq[q];
  });
}

void orc_set_force_field(void* p, const double* ff) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S->set_force_field(ff));
}

void orc_clear_force_field(void* p) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S->clear_force_field());
}

int orc_get_force_field(void* p, double* out) {
  Handle* h = (Handle*)p;
  int has = 0;
  DISPATCH(h, {
    has = S->has_ff ? 1 : 0;
    for (size_t k = 0; k < S->nc * 3; ++k) out[k] = S->has_ff ? (double)S->ff[k] : 0.0;
  });
  return has;
}

void orc_set_omega_field(void* p, const double* om) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    if (!om) {
      S->has_omega = false;
      S->omega_field.clear();
    } else {
      S->has_omega = true;
      S->omega_field.resize(S->nc);
      for (size_t c = 0; c < S->nc; ++c) S->omega_field[c] = (TT)om[c];
    }
  });
}

// Volume source list: n entries of 10 doubles (lo[3], hi[3], kind, q_lu,
// u[3] packed as lo0 lo1 lo2 hi0 hi1 hi2 kind q_lu u0 u1 u2 -> 11 values).
void orc_set_sources(void* p, int n, const double* v) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    S->sources.clear();
    for (int k = 0; k < n; ++k) {
      const double* e = v + 11 * k;
      typename std::remove_reference_t<decltype(*S)>::Src src{};
      for (int a = 0; a < 3; ++a) {
        src.lo[a] = (long)e[a];
        src.hi[a] = (long)e[3 + a];
        src.u[a] = (TT)e[8 + a];
      }
      src.kind = (int)e[6];
      src.q_lu = (TT)e[7];
      S->sources.push_back(src);
    }
  });
}

// Force probe over the solids marked in `mask` (compact, NULL clears) and
// the momentum-exchange force of the most recent step.
void orc_set_probe(void* p, const uint8_t* mask) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S->set_probe(mask));
}
void orc_probed_force(void* p, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    for (int a = 0; a < 3; ++a) out[a] = (double)S->probed[a];
  });
}

// Open face `face` (0..5 = XNeg..ZPos): kind 0 closed, 1 velocity (v = u),
// 2 pressure (v[0] = rho), 3 outflow, 4 convective (v[0] = u_conv).  Returns
// -1 for an open face on a periodic axis (GlobalSpec::validate).
int orc_set_face(void* p, int face, int kind, const double* v) {
  Handle* h = (Handle*)p;
  int rc = 0;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    if (face < 0 || face > 5 || kind < 0 || kind > 4 || (kind && S->periodic[face / 2])) {
      rc = -1;
    } else {
      S->face_kind[face] = kind;
      for (int a = 0; a < 3; ++a) S->face_u[face][a] = kind == 1 ? (TT)v[a] : TT(0);
      S->face_rho[face] = kind == 2 ? (TT)v[0] : TT(0);
      S->face_uconv[face] = kind == 4 ? (TT)v[0] : TT(0);
    }
  });
  return rc;
}

// Face patches (FacePatch, params.rs:286-296; validate_face_patches,
// solver.rs:690-740): n rows of 9 doubles {face, lo1, lo2, hi1, hi2, kind,
// v[3]} with v as for orc_set_face; replaces the list.  Returns -1 for a
// face out of range, a rect outside the face or inverted, overlapping rects
// on one face, or a patch on a periodic axis.
int orc_set_patches(void* p, int n, const double* rows) {
  Handle* h = (Handle*)p;
  int rc = 0;
  DISPATCH(h, {
    using SS = std::remove_reference_t<decltype(*S)>;
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    std::vector<typename SS::Patch> list;
    const long ext[3] = {S->nx, S->ny, S->nz};
    for (int k = 0; k < n && !rc; ++k) {
      const double* r = rows + 9 * k;
      typename SS::Patch pt{};
      pt.face = (int)r[0];
      const int kind = (int)r[5];
      if (pt.face < 0 || pt.face > 5 || kind < 0 || kind > 4 || S->periodic[pt.face / 2]) {
        rc = -1;
        break;
      }
      const int a = pt.face / 2, t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
      for (int c = 0; c < 2; ++c) {
        pt.lo[c] = (long)r[1 + c];
        pt.hi[c] = (long)r[3 + c];
      }
      if (pt.lo[0] < 0 || pt.lo[1] < 0 || pt.lo[0] > pt.hi[0] || pt.lo[1] > pt.hi[1] || pt.hi[0] >= ext[t1] ||
          pt.hi[1] >= ext[t2]) {
        rc = -1;
        break;
      }
      pt.bc.kind = kind;
      for (int c = 0; c < 3; ++c) pt.bc.u[c] = kind == 1 ? (TT)r[6 + c] : TT(0);
      pt.bc.rho = kind == 2 ? (TT)r[6] : TT(0);
      pt.bc.uconv = kind == 4 ? (TT)r[6] : TT(0);
      for (const auto& o : list)
        if (o.face == pt.face && pt.lo[0] <= o.hi[0] && o.lo[0] <= pt.hi[0] && pt.lo[1] <= o.hi[1] &&
            o.lo[1] <= pt.hi[1])
          rc = -1;
      list.push_back(pt);
    }
    if (!rc) S->patches = list;
  });
  return rc;
}

// Inlet profile of a velocity face (Solver::set_inlet_profile,
// solver.rs:3322-3350): values = 3 x (face cells) in face order
// (c2 * n1 + c1), or NULL to clear.  Returns -1 unless the face is a
// velocity face.
int orc_set_inlet_profile(void* p, int face, const double* values) {
  Handle* h = (Handle*)p;
  int rc = 0;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    if (face < 0 || face > 5 || S->face_kind[face] != 1) {
      rc = -1;
    } else if (!values) {
      S->profile[face].clear();
    } else {
      const long ext[3] = {S->nx, S->ny, S->nz};
      const int a = face / 2, t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
      const long nf = ext[t1] * ext[t2];
      S->profile[face].assign(3 * nf, TT(0));
      for (long k = 0; k < 3 * nf; ++k) S->profile[face][k] = (TT)values[k];
    }
  });
  return rc;
}

// Strain rate (N x 6) and velocity gradient (N x 9, row-major du_i/dx_j) of
// the current state, compact order.
void orc_strain(void* p, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    for (long z = 0; z < S->nz; ++z)
      for (long y = 0; y < S->ny; ++y)
        for (long x = 0; x < S->nx; ++x) {
          TT s[6];
          S->strain_at(x, y, z, s);
          for (int k = 0; k < 6; ++k) out[6 * S->cidx(x, y, z) + k] = (double)s[k];
        }
  });
}

void orc_vgrad(void* p, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    for (long z = 0; z < S->nz; ++z)
      for (long y = 0; y < S->ny; ++y)
        for (long x = 0; x < S->nx; ++x) {
          TT g[3][3];
          S->vgrad_at(x, y, z, g);
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) out[9 * S->cidx(x, y, z) + 3 * a + b] = (double)g[a][b];
        }
  });
}

// WALE omega field of the current state (tau_eff_max <= 0: no clip);
// nut may be null.  Returns the clipped cell count.
long orc_wale(void* p, double cw, double tau_eff_max, double* omega, double* nut, double* max_nut) {
  Handle* h = (Handle*)p;
  long clipped = 0;
  DISPATCH(h, {
    using TT = std::remove_reference_t<decltype(S->f[0])>;
    std::vector<TT> om(S->nc), nt(S->nc);
    TT mx;
    clipped = S->wale((TT)cw, tau_eff_max, om.data(), nt.data(), &mx);
    for (size_t c = 0; c < S->nc; ++c) {
      omega[c] = (double)om[c];
      if (nut) nut[c] = (double)nt[c];
    }
    *max_nut = (double)mx;
  });
  return clipped;
}

int orc_apply_ibm(void* p, int nm, const double* pos, const double* weight, const double center[3],
                  const double omega[3], int max_iter, double tol, int radius, double relax, double* diag,
                  int stencil_dim) {
  Handle* h = (Handle*)p;
  int r = 0;
  DISPATCH(h, r = S->apply_ibm(nm, pos, weight, center, omega, max_iter, tol, radius, relax, diag, stencil_dim));
  return r;
}

int orc_scalar_enable(void* p, double D, int d3q7) {
  Handle* h = (Handle*)p;
  int r = 0;
  DISPATCH(h, r = S->scalar_enable(D, d3q7 != 0));
  return r;
}

void orc_scalar_set(void* p, int k, const double* vals) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S->scalar_set(k, vals));
}

void orc_scalar_get(void* p, int k, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    for (size_t c = 0; c < S->nc; ++c) out[c] = (double)S->scalars[k].conc[c];
  });
}

void orc_scalar_get_h(void* p, int k, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    auto& s = S->scalars[k];
    for (int q = 0; q < s.L.Q; ++q)
      for (long z = 0; z < S->nz; ++z)
        for (long y = 0; y < S->ny; ++y)
          for (long x = 0; x < S->nx; ++x)
            out[q * S->nc + S->cidx(x, y, z)] = (double)s.h[q * S->np + S->pidx(x, y, z)];
  });
}

void orc_build_wall_rims(const int dims[3], const int is_wall[6], const double* u6x3, uint8_t* solid,
                         double* wall_u) {
  double u[6][3];
  for (int f = 0; f < 6; ++f)
    for (int a = 0; a < 3; ++a) u[f][a] = u6x3[3 * f + a];
  build_wall_rims(dims, is_wall, u, solid, wall_u);
}

void orc_tank_mask(const int dims[3], double dx, double diameter, double height, uint8_t* solid) {
  tank_mask(dims, dx, diameter, height, solid);
}

void orc_baffle_mask(const int dims[3], double dx, double diameter, double height, double width, double thickness,
                     uint8_t* solid, uint8_t* baffle) {
  baffle_mask(dims, dx, diameter, height, width, thickness, solid, baffle);
}

// Returns marker count; if pos == NULL only counts.
int orc_impeller_surface(const int dims[3], double dx, double vessel_diameter, double liquid_height, double diameter,
                         double clearance, int blades, double blade_angle_deg, double omega_z,
                         double thickness_fraction, double area_per_marker, uint64_t seed, double* pos,
                         double* weight, double* vel, double* center) {
  std::vector<double> P, W, V;
  double c[3];
  int n = impeller_surface(dims, dx, vessel_diameter, liquid_height, diameter, clearance, blades, blade_angle_deg,
                           omega_z, thickness_fraction, area_per_marker, seed, P, W, V, c);
  if (pos) {
    std::memcpy(pos, P.data(), P.size() * 8);
    std::memcpy(weight, W.data(), W.size() * 8);
    std::memcpy(vel, V.data(), V.size() * 8);
    std::memcpy(center, c, 24);
  }
  return n;
}

// Returns marker count; if pos == NULL only counts.
int orc_radial_impeller(const int dims[3], double dx, double vessel_diameter, double liquid_height, double diameter,
                        double clearance, int blades, double blade_angle_deg, double omega_z,
                        double thickness_fraction, int inner_kind, double* pos, double* weight, double* vel,
                        double* center) {
  std::vector<double> P, W, V;
  double c[3];
  int n = radial_impeller(dims, dx, vessel_diameter, liquid_height, diameter, clearance, blades, blade_angle_deg,
                          omega_z, thickness_fraction, inner_kind, P, W, V, c);
  if (pos) {
    std::memcpy(pos, P.data(), P.size() * 8);
    std::memcpy(weight, W.data(), W.size() * 8);
    std::memcpy(vel, V.data(), V.size() * 8);
    std::memcpy(center, c, 24);
  }
  return n;
}

// stamp_impeller_wall_velocity, geometry.rs:908-920
void orc_stamp_wall_velocity(const int dims[3], int nm, const double* pos, const double* vel, double* wall_u) {
  for (int m = 0; m < nm; ++m) {
    long ix[3];
    for (int a = 0; a < 3; ++a) {
      double v = std::round(pos[3 * m + a]);
      v = std::min(std::max(v, 0.0), (double)(dims[a] - 1));
      ix[a] = (long)v;
    }
    size_t i = ((size_t)ix[2] * dims[1] + ix[1]) * dims[0] + ix[0];
    for (int a = 0; a < 3; ++a) wall_u[3 * i + a] = vel[3 * m + a];
  }
}

// marker_stencil for one marker: writes up to 64 (x,y,z) triples and weights.
int orc_marker_stencil(const double p[3], const int dims[3], int radius, long* cells, double* w) {
  std::vector<Solver<double>::SP> st;
  Solver<double>::marker_stencil(p, dims, radius, st);
  for (size_t k = 0; k < st.size(); ++k) {
    cells[3 * k + 0] = st[k].x;
    cells[3 * k + 1] = st[k].y;
    cells[3 * k + 2] = st[k].z;
    w[k] = st[k].w;
  }
  return (int)st.size();
}

// ---- QOI reducers, qoi.rs:477-556, 699-739 ----------------------------------
// power_qois: out = [power, n_hz, np, p_over_v]
int orc_power_qois(double torque, double omega, double rho, double d, double volume, double* out) {
  if (!(std::isfinite(omega) && std::isfinite(rho) && std::isfinite(d) && std::isfinite(volume) && rho > 0.0 &&
        d > 0.0 && volume > 0.0))
    return -1;
  double n_hz = omega / 6.283185307179586;
  if (!(n_hz > 0.0)) return -2;
  double power = omega * torque;
  double denom = rho * (n_hz * n_hz * n_hz) * (d * d * d * d * d);
  out[0] = power;
  out[1] = n_hz;
  out[2] = power / denom;
  out[3] = power / volume;
  return 0;
}

// scalar_cv: returns 1 and writes *cv if defined, else 0.
int orc_scalar_cv(long n, const double* values, const uint8_t* include, double* cv) {
  long cnt = 0;
  double sum = 0.0;
  for (long i = 0; i < n; ++i)
    if (include[i]) {
      cnt++;
      sum += values[i];
    }
  if (cnt == 0) return 0;
  double mean = sum / (double)cnt;
  if (mean == 0.0) return 0;
  double var = 0.0;
  for (long i = 0; i < n; ++i)
    if (include[i]) {
      double d = values[i] - mean;
      var += d * d;
    }
  *cv = std::sqrt(var / (double)cnt) / std::fabs(mean);
  return 1;
}

// mixing_time_from_cv: out = [cv0, t95 (NaN if none), t99 (NaN if none)]
int orc_mixing_time(long n, const double* t, const double* cv, double* out) {
  long first = -1;
  for (long i = 0; i < n; ++i)
    if (std::isfinite(cv[i]) && cv[i] > 0.0) {
      first = i;
      break;
    }
  if (first < 0) return 0;
  double cv0 = cv[first];
  double t95 = NAN, t99 = NAN;
  bool h95 = false, h99 = false;
  for (long i = 0; i < n; ++i) {
    if (!std::isfinite(cv[i])) continue;
    if (!h95 && cv[i] <= 0.05 * cv0) {
      t95 = t[i];
      h95 = true;
    }
    if (!h99 && cv[i] <= 0.01 * cv0) {
      t99 = t[i];
      h99 = true;
    }
  }
  out[0] = cv0;
  out[1] = t95;
  out[2] = t99;
  return 1;
}

// compartment_cv: out = [cv_top, cv_bulk, cv_near] (NaN if undefined),
// counts = [n_top, n_bulk, n_near]
void orc_compartment_cv(const int dims[3], const double* values, const uint8_t* solid, int has_imp, double imp_z,
                        double* out, long* counts) {
  size_t n = (size_t)dims[0] * dims[1] * dims[2];
  std::vector<uint8_t> mask[3];
  for (auto& m : mask) m.assign(n, 0);
  for (long z = 0; z < dims[2]; ++z)
    for (long y = 0; y < dims[1]; ++y)
      for (long x = 0; x < dims[0]; ++x) {
        size_t i = ((size_t)z * dims[1] + y) * dims[0] + x;
        if (solid[i]) continue;
        double zn = ((double)z + 0.5) / (double)dims[2];
        if (zn >= 2.0 / 3.0)
          mask[0][i] = 1;
        else if (has_imp && std::fabs(((double)z + 0.5) - imp_z) <= (double)dims[2] / 6.0)
          mask[2][i] = 1;
        else
          mask[1][i] = 1;
      }
  for (int k = 0; k < 3; ++k) {
    double cv;
    out[k] = orc_scalar_cv((long)n, values, mask[k].data(), &cv) ? cv : NAN;
    long c = 0;
    for (size_t i = 0; i < n; ++i) c += mask[k][i];
    counts[k] = c;
  }
}

}  // extern "C"
