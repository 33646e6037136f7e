// lbm_open.hip — open-face boundary conditions (Zou-He velocity / pressure,
// zero-gradient outflow, convective outflow) on the AA-pattern layout, and
// the stale-slot ("stash") contract they need.
//
// Reference step order (backend.rs:266-328, kernels.rs:547-622, 849-1299):
// collide -> stream (a population whose source lies beyond a non-periodic
// face is skipped, so its slot keeps the other buffer's value: the cell's
// own post-collide population of the previous step) -> swap -> open faces in
// Face::ALL order -> moments.
//
// Here the collide-stream kernel wraps every neighbour offset, so the
// populations that would enter a face cell from outside arrive from the
// opposite face (or across a wrapped edge) and the ones that leave land in
// slots nobody else reads: in the AA layout each logical population f_q(i)
// has exactly one slot, and the slot of an unknown f_q(i) of an open face is
// that of a population whose source is outside the domain.  After every
// step, in this order:
//   capture  (all faces): stash_next[k][cell] = f*_q(i), the post-collide
//            population the face cell just sent inward along unknown q
//            (found where the next state keeps f_q(i + c_q), or in the
//            cell's own bounce-back slot when i + c_q is solid);
//   restore  (all faces): f_q(i) = stash_cur[k][cell] for every unknown q,
//            which is the reference's post-swap state of the skipped slots;
//   BC       (one launch per face, Face::ALL order): the face's rule, on
//            the logical populations read and written in place.
// Face patches give the cells of their rect their own rule; a face with
// patches is open even when its own rule is closed, and its other cells then
// take a zero-velocity Zou-He lid (backend.rs:602-766).  Within one face no
// rule reads another face cell, so the reference's base-then-patches passes
// are one pass with a per-cell rule.
// When every open face lies on one axis the faces share no cell and no
// thread reads what another writes, so one launch does all three per cell.
// The stash is double-buffered and starts at zero (the reference's ftmp).
// Face kernels are templated on the face so every lattice direction is a
// compile-time constant (no indexed register arrays).
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

#pragma clang fp contract(off)

namespace lbm {

template <int FACE>
struct FaceC {
  static constexpr int a = FACE / 2;
  static constexpr int nin = (FACE % 2 == 0) ? 1 : -1;
  static constexpr int t1 = a == 0 ? 1 : 0;
  static constexpr int t2 = a == 2 ? 1 : 2;
  template <int q>
  static constexpr int dot() {
    return lat_c(q, a) * nin;
  }
};

// index of unknown q in the face's ascending unknown list
template <int Q, int FACE, int q>
constexpr int unk_index() {
  int k = 0;
  for (int p = 0; p < q; ++p)
    if (lat_c(p, FACE / 2) * ((FACE % 2 == 0) ? 1 : -1) > 0) ++k;
  return k;
}

constexpr int find_dir(int Q, int cx, int cy, int cz) {
  for (int q = 0; q < Q; ++q)
    if (lat_c(q, 0) == cx && lat_c(q, 1) == cy && lat_c(q, 2) == cz) return q;
  return -1;
}

// sum over the unknowns of w_q c_{q,t}^2 in f64, then fp32 (kernels.rs:980-988)
template <int Q, int FACE>
constexpr float unk_s2(int t) {
  double s = 0.0;
  for (int q = 0; q < Q; ++q)
    if (lat_c(q, FACE / 2) * ((FACE % 2 == 0) ? 1 : -1) > 0) s += lat_w(Q, q) * lat_c(q, t) * lat_c(q, t);
  return (float)s;
}
template <int Q, int FACE>
constexpr float unk_wsum() {  // kernels.rs:1272-1277
  double s = 0.0;
  for (int q = 0; q < Q; ++q)
    if (lat_c(q, FACE / 2) * ((FACE % 2 == 0) ? 1 : -1) > 0) s += lat_w(Q, q);
  return (float)s;
}

// cell of face coordinate `coord` (t2 outer, t1 inner: for_face_cells,
// kernels.rs:742-766)
template <int FACE>
__device__ __forceinline__ uint32_t face_cell(const Dom& d, uint32_t coord, uint32_t& x, uint32_t& y, uint32_t& z) {
  using F = FaceC<FACE>;
  const uint32_t ext[3] = {d.nx, d.ny, d.nz};
  const uint32_t n1 = ext[F::t1];
  const uint32_t c1 = coord % n1, c2 = coord / n1;
  uint32_t p[3];
  p[F::a] = F::nin > 0 ? 0u : ext[F::a] - 1;
  p[F::t1] = c1;
  p[F::t2] = c2;
  x = p[0];
  y = p[1];
  z = p[2];
  return (z * d.ny + y) * d.nx + x;
}

template <int FACE>
__host__ __device__ inline uint32_t face_cells(const Dom& d) {
  using F = FaceC<FACE>;
  const uint32_t ext[3] = {d.nx, d.ny, d.nz};
  return ext[F::t1] * ext[F::t2];
}

// ---- capture / restore --------------------------------------------------------
template <int Q, typename S, int FACE>
__device__ __forceinline__ void capture_cell(const FaceArgs& a, uint32_t coord) {
  uint32_t x, y, z;
  const uint32_t i = face_cell<FACE>(a.d, coord, x, y, z);
  const uint32_t m = a.mask[i];
  const uint32_t nf = face_cells<FACE>(a.d);
  float* out = a.stash_next[FACE];
  const Nbr nb(a.d, x, y, z);
  const S* f = (const S*)a.f;
  sfor<1, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (FaceC<FACE>::template dot<q>() > 0) {
      constexpr int k = unk_index<Q, FACE, q>();
      constexpr int o = lat_opp(q);
      float v = 0.0f;
      if (!(m & kMaskSolid)) {
        if ((m >> o) & 1u) {  // i + c_q solid: the bounce-back slot
          v = read_pop<Q, S, o>(f, a.d.ps, nb, i, m, a.odd != 0);
        } else {
          const uint32_t j = i + (uint32_t)nb.template off<lat_c(q, 0), lat_c(q, 1), lat_c(q, 2)>();
          uint32_t jx, jy, jz;
          cell_xyz(a.d, j, jx, jy, jz);
          const Nbr nj(a.d, jx, jy, jz);
          v = read_pop<Q, S, q>(f, a.d.ps, nj, j, a.mask[j], a.odd != 0);
        }
      }
      out[(size_t)k * nf + coord] = v;
    }
  });
}

template <int Q, typename S, int FACE>
__device__ __forceinline__ void restore_cell(const FaceArgs& a, uint32_t coord) {
  uint32_t x, y, z;
  const uint32_t i = face_cell<FACE>(a.d, coord, x, y, z);
  const uint32_t m = a.mask[i];
  if (m & kMaskSolid) return;
  const uint32_t nf = face_cells<FACE>(a.d);
  const float* in = a.stash_cur[FACE];
  const Nbr nb(a.d, x, y, z);
  S* f = (S*)a.f;
  sfor<1, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (FaceC<FACE>::template dot<q>() > 0) {
      constexpr int k = unk_index<Q, FACE, q>();
      write_pop<Q, S, q>(f, a.d.ps, nb, i, m, a.odd != 0, in[(size_t)k * nf + coord]);
    }
  });
}

// one launch for every open face: blockIdx.y = slot in a.faces
template <int Q, typename S, bool RESTORE>
__global__ __launch_bounds__(256) void k_face_stash(FaceArgs a) {
  const int face = a.faces[blockIdx.y];
  const uint32_t coord = blockIdx.x * blockDim.x + threadIdx.x;
  switch (face) {
#define LBM_FACE_CASE(F)                                   \
  case F:                                                  \
    if (coord < face_cells<F>(a.d)) {                      \
      if constexpr (RESTORE)                               \
        restore_cell<Q, S, F>(a, coord);                   \
      else                                                 \
        capture_cell<Q, S, F>(a, coord);                   \
    }                                                      \
    break;
    LBM_FACE_CASE(0)
    LBM_FACE_CASE(1)
    LBM_FACE_CASE(2)
    LBM_FACE_CASE(3)
    LBM_FACE_CASE(4)
    LBM_FACE_CASE(5)
#undef LBM_FACE_CASE
    default:
      break;
  }
}

// ---- boundary rules ----------------------------------------------------------
// density and raw (pre-force) velocity of a Zou-He face cell
// (kernels.rs:1154-1197 / 1017-1060; no gravity on this path)
__device__ __forceinline__ void zh_solve(int kind, const float* u, float rho_bc, float closure, const float f0[3],
                                         int ax, int t1, int t2, float nsign, float& r, float& vn, float& vt1,
                                         float& vt2) {
  const float one = 1.0f, half = 0.5f;
  if (f0[0] == 0.0f && f0[1] == 0.0f && f0[2] == 0.0f) {
    if (kind == kFaceVelocity) {
      const float un = u[ax] * nsign;
      r = closure / (one - un);
      vn = un;
      vt1 = u[t1];
      vt2 = u[t2];
    } else {
      r = rho_bc;
      vn = one - closure / rho_bc;
      vt1 = 0.0f;
      vt2 = 0.0f;
    }
    return;
  }
  const float f0n = f0[ax] * nsign, gn = 0.0f;
  if (kind == kFaceVelocity) {
    const float un_phys = u[ax] * nsign;
    r = (closure - half * f0n) / (one - un_phys + half * gn);
    const float inv_r = one / r;
    vn = un_phys - half * (f0n * inv_r + gn);
    vt1 = u[t1] - half * (f0[t1] * inv_r + 0.0f);
    vt2 = u[t2] - half * (f0[t2] * inv_r + 0.0f);
  } else {
    const float inv_r = one / rho_bc;
    r = rho_bc;
    vn = one - closure * inv_r;
    vt1 = -half * (f0[t1] * inv_r + 0.0f);
    vt2 = -half * (f0[t2] * inv_r + 0.0f);
  }
}

// FUSED: capture and restore of this cell in the same thread, then the rule
// (legal when every open face lies on one axis: the faces then share no
// cell, and no thread reads a population another thread writes)
template <int Q, typename S, int FACE, bool FUSED>
__device__ __forceinline__ void bc_cell(const FaceArgs& a, uint32_t coord) {
  using F = FaceC<FACE>;
  constexpr int ax = F::a, t1 = F::t1, t2 = F::t2;
  if constexpr (FUSED) capture_cell<Q, S, FACE>(a, coord);
  uint32_t x, y, z;
  const uint32_t i = face_cell<FACE>(a.d, coord, x, y, z);
  const uint32_t m = a.mask[i];
  if (m & kMaskSolid) return;
  const bool odd = a.odd != 0;
  const Nbr nb(a.d, x, y, z);
  S* f = (S*)a.f;
  float fv[Q];
  const uint32_t nf = face_cells<FACE>(a.d);
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (FUSED && F::template dot<q>() > 0) {
      // restore: the stale slot takes the previous post-collide value
      const float v = a.stash_cur[FACE][(size_t)unk_index<Q, FACE, q>() * nf + coord];
      write_pop<Q, S, q>(f, a.d.ps, nb, i, m, odd, v);
      fv[q] = v;
    } else {
      fv[q] = read_pop<Q, S, q>(f, a.d.ps, nb, i, m, odd);
    }
  });
  // the cell's rule: the patch whose rect holds it, else the face's own
  // (patches on one face never overlap)
  int kind = a.kind[FACE];
  float ub[3] = {a.u[FACE][0], a.u[FACE][1], a.u[FACE][2]};
  float rho_bc = a.rho[FACE], lam = a.uconv[FACE];
  bool own_rule = true;
  if (a.npatch) {
    const uint32_t ext[3] = {a.d.nx, a.d.ny, a.d.nz};
    const uint32_t c1 = coord % ext[t1], c2 = coord / ext[t1];
    for (int p = 0; p < a.npatch; ++p) {
      const FacePatchArgs& pt = a.patch[p];
      if (pt.face == FACE && c1 >= pt.lo[0] && c1 <= pt.hi[0] && c2 >= pt.lo[1] && c2 <= pt.hi[1]) {
        kind = pt.kind;
        ub[0] = pt.u[0];
        ub[1] = pt.u[1];
        ub[2] = pt.u[2];
        rho_bc = pt.rho;
        lam = pt.uconv;
        own_rule = false;
        break;
      }
    }
  }
  // the inlet profile replaces u on the face's own velocity rule
  // (profile.map_or(u, |p| p[coord]), kernels.rs:1021-1034, 1164-1180)
  if (own_rule && kind == kFaceVelocity && a.prof[FACE]) {
    const float* pu = a.prof[FACE] + 3 * (size_t)coord;
    ub[0] = pu[0];
    ub[1] = pu[1];
    ub[2] = pu[2];
  }
  if (kind == kFaceVelocity || kind == kFacePressure) {
    float f0[3] = {a.p.F[0], a.p.F[1], a.p.F[2]};
    if (a.F_cur && (m & kMaskForced)) {  // force_independent_at: field nonzero only on the force list
      f0[0] = f0[0] + a.F_cur[3 * (size_t)i + 0];
      f0[1] = f0[1] + a.F_cur[3 * (size_t)i + 1];
      f0[2] = f0[2] + a.F_cur[3 * (size_t)i + 2];
    }
    const float nsign = (float)F::nin, one = 1.0f, two = 2.0f, half = 0.5f;
    float r, vn, vt1, vt2;
    if constexpr (Q == 27) {  // zou_he_face_d3q27, kernels.rs:958-1083
      float s0 = 0.0f, sneg = 0.0f, q0a = 0.0f, q0b = 0.0f;
      sfor<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int dt = F::template dot<q>();
        if constexpr (dt == 0) {
          s0 = s0 + fv[q];
          q0a = q0a + (float)lat_c(q, t1) * fv[q];
          q0b = q0b + (float)lat_c(q, t2) * fv[q];
        } else if constexpr (dt < 0) {
          sneg = sneg + fv[q];
        }
      });
      const float closure = s0 + two * sneg + one;
      zh_solve(kind, ub, rho_bc, closure, f0, ax, t1, t2, nsign, r, vn, vt1, vt2);
      constexpr float s2a = unk_s2<Q, FACE>(t1), s2b = unk_s2<Q, FACE>(t2);
      const float six = 6.0f;
      const float c_t1 = r * vt1 * (one - six * s2a) - q0a;
      const float c_t2 = r * vt2 * (one - six * s2b) - q0b;
      sfor<1, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (F::template dot<q>() > 0) {
          const float cdotu =
              (float)lat_c(q, ax) * nsign * vn + (float)lat_c(q, t1) * vt1 + (float)lat_c(q, t2) * vt2;
          constexpr float w = lat_wf(Q, q);
          const float delta = c_t1 * w * (float)lat_c(q, t1) / s2a + c_t2 * w * (float)lat_c(q, t2) / s2b;
          const float v = fv[lat_opp(q)] + six * w * r * cdotu + delta;
          write_pop<Q, S, q>(f, a.d.ps, nb, i, m, odd, v);
        }
      });
    } else {  // zou_he_face_3d, kernels.rs:1085-1219
      constexpr int n0 = ax == 0 ? F::nin : 0, n1 = ax == 1 ? F::nin : 0, n2 = ax == 2 ? F::nin : 0;
      constexpr int e10 = t1 == 0, e11 = t1 == 1, e12 = t1 == 2;
      constexpr int e20 = t2 == 0, e21 = t2 == 1, e22 = t2 == 2;
      constexpr int q_n = find_dir(Q, n0, n1, n2);
      constexpr int q_p1 = find_dir(Q, n0 + e10, n1 + e11, n2 + e12);
      constexpr int q_m1 = find_dir(Q, n0 - e10, n1 - e11, n2 - e12);
      constexpr int q_p2 = find_dir(Q, n0 + e20, n1 + e21, n2 + e22);
      constexpr int q_m2 = find_dir(Q, n0 - e20, n1 - e21, n2 - e22);
      constexpr int q_t1 = find_dir(Q, e10, e11, e12), q_mt1 = find_dir(Q, -e10, -e11, -e12);
      constexpr int q_t2 = find_dir(Q, e20, e21, e22), q_mt2 = find_dir(Q, -e20, -e21, -e22);
      constexpr int q_pp = find_dir(Q, e10 + e20, e11 + e21, e12 + e22);
      constexpr int q_pm = find_dir(Q, e10 - e20, e11 - e21, e12 - e22);
      constexpr int q_mp = find_dir(Q, -e10 + e20, -e11 + e21, -e12 + e22);
      constexpr int q_mm = find_dir(Q, -e10 - e20, -e11 - e21, -e12 - e22);
      static_assert(q_n > 0 && q_p1 > 0 && q_m1 > 0 && q_p2 > 0 && q_m2 > 0 && q_pp > 0 && q_pm > 0 && q_mp > 0 &&
                        q_mm > 0,
                    "face directions");
      const float s0 = fv[0] + fv[q_t1] + fv[q_mt1] + fv[q_t2] + fv[q_mt2] + fv[q_pp] + fv[q_pm] + fv[q_mp] +
                       fv[q_mm];
      const float sneg =
          fv[lat_opp(q_n)] + fv[lat_opp(q_p1)] + fv[lat_opp(q_m1)] + fv[lat_opp(q_p2)] + fv[lat_opp(q_m2)];
      const float closure = s0 + two * sneg + one;
      zh_solve(kind, ub, rho_bc, closure, f0, ax, t1, t2, nsign, r, vn, vt1, vt2);
      const float qt1 = fv[q_t1] - fv[q_mt1] + fv[q_pp] + fv[q_pm] - fv[q_mp] - fv[q_mm];
      const float qt2 = fv[q_t2] - fv[q_mt2] + fv[q_pp] - fv[q_pm] + fv[q_mp] - fv[q_mm];
      const float c13 = (float)(1.0 / 3.0), c16 = (float)(1.0 / 6.0);
      const float nn1 = c13 * r * vt1 - half * qt1;
      const float nn2 = c13 * r * vt2 - half * qt2;
      write_pop<Q, S, q_n>(f, a.d.ps, nb, i, m, odd, fv[lat_opp(q_n)] + c13 * r * vn);
      write_pop<Q, S, q_p1>(f, a.d.ps, nb, i, m, odd, fv[lat_opp(q_p1)] + c16 * r * (vn + vt1) + nn1);
      write_pop<Q, S, q_m1>(f, a.d.ps, nb, i, m, odd, fv[lat_opp(q_m1)] + c16 * r * (vn - vt1) - nn1);
      write_pop<Q, S, q_p2>(f, a.d.ps, nb, i, m, odd, fv[lat_opp(q_p2)] + c16 * r * (vn + vt2) + nn2);
      write_pop<Q, S, q_m2>(f, a.d.ps, nb, i, m, odd, fv[lat_opp(q_m2)] + c16 * r * (vn - vt2) - nn2);
    }
    return;
  }
  // outflow / convective: the interior cell one step inward
  // (outflow_face_selected kernels.rs:1221-1248, convective_face_selected
  // kernels.rs:1254-1299)
  const uint32_t j = i + (uint32_t)nb.template off<ax == 0 ? F::nin : 0, ax == 1 ? F::nin : 0, ax == 2 ? F::nin : 0>();
  const uint32_t mj = a.mask[j];
  if (mj & kMaskSolid) return;
  uint32_t jx, jy, jz;
  cell_xyz(a.d, j, jx, jy, jz);
  const Nbr nj(a.d, jx, jy, jz);
  if (kind == kFaceOutflow) {
    sfor<1, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (F::template dot<q>() > 0)
        write_pop<Q, S, q>(f, a.d.ps, nb, i, m, odd, read_pop<Q, S, q>(f, a.d.ps, nj, j, mj, odd));
    });
    return;
  }
  // convective
  const float inv = 1.0f / (1.0f + lam);
  float fj[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    fj[q] = read_pop<Q, S, q>(f, a.d.ps, nj, j, mj, odd);
  });
  sfor<1, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (F::template dot<q>() > 0) fv[q] = (fv[q] + lam * fj[q]) * inv;
  });
  float di = 0.0f, dj = 0.0f;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    di = di + fv[q];
    dj = dj + fj[q];
  });
  const float corr = dj - di;
  constexpr float wsum = unk_wsum<Q, FACE>();
  sfor<1, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (F::template dot<q>() > 0)
      write_pop<Q, S, q>(f, a.d.ps, nb, i, m, odd, fv[q] + corr * lat_wf(Q, q) / wsum);
  });
}

// FUSED: every face of a.faces at once (blockIdx.y = slot); otherwise the
// one face `face`
template <int Q, typename S, bool FUSED>
__global__ __launch_bounds__(256) void k_face_bc(FaceArgs a, int face) {
  const uint32_t coord = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (FUSED) face = a.faces[blockIdx.y];
  switch (face) {
#define LBM_FACE_CASE(F)                                                  \
  case F:                                                                 \
    if (coord < face_cells<F>(a.d)) bc_cell<Q, S, F, FUSED>(a, coord);    \
    break;
    LBM_FACE_CASE(0)
    LBM_FACE_CASE(1)
    LBM_FACE_CASE(2)
    LBM_FACE_CASE(3)
    LBM_FACE_CASE(4)
    LBM_FACE_CASE(5)
#undef LBM_FACE_CASE
    default:
      break;
  }
}

// ---- volume sources (apply_volume_sources_impl, backend.rs:799-856) ----------
// after the open faces, per source in order: every fluid cell of the box
// gains w_q q_cell (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), q_cell = q_lu/count
template <int Q, typename S>
__global__ __launch_bounds__(256) void k_volume_source(SourceArgs a) {
  const uint32_t ex = a.hi[0] - a.lo[0] + 1, ey = a.hi[1] - a.lo[1] + 1, ez = a.hi[2] - a.lo[2] + 1;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ex * ey * ez) return;
  const uint32_t x = a.lo[0] + k % ex, y = a.lo[1] + (k / ex) % ey, z = a.lo[2] + k / (ex * ey);
  const uint32_t i = (z * a.d.ny + y) * a.d.nx + x;
  const uint32_t m = a.mask[i];
  if (m & kMaskSolid) return;
  const Nbr nb(a.d, x, y, z);
  S* f = (S*)a.f;
  const bool odd = a.odd != 0;
  const float q_cell = a.q_lu / (float)((double)ex * ey * ez);
  const float u0 = a.u[0], u1 = a.u[1], u2 = a.u[2];
  float usq = u0 * u0;
  usq = usq + u1 * u1;
  usq = usq + u2 * u2;
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    float cu = (float)lat_c(q, 0) * u0;  // every term kept, as the reference's loop
    cu = cu + (float)lat_c(q, 1) * u1;
    cu = cu + (float)lat_c(q, 2) * u2;
    const float delta = lat_wf(Q, q) * q_cell * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * usq);
    write_pop<Q, S, q>(f, a.d.ps, nb, i, m, odd, read_pop<Q, S, q>(f, a.d.ps, nb, i, m, odd) + delta);
  });
}

hipError_t launch_volume_source(int Q, int storage, const SourceArgs& a, hipStream_t s) {
  const uint32_t n = (a.hi[0] - a.lo[0] + 1) * (a.hi[1] - a.lo[1] + 1) * (a.hi[2] - a.lo[2] + 1);
  const dim3 g((n + 255) / 256), b(256);
  if (storage == kF32) {
    if (Q == 19)
      hipLaunchKernelGGL((k_volume_source<19, float>), g, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_volume_source<27, float>), g, b, 0, s, a);
  } else if (storage == kF16) {
    if (Q == 19)
      hipLaunchKernelGGL((k_volume_source<19, _Float16>), g, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_volume_source<27, _Float16>), g, b, 0, s, a);
  } else {
    if (Q == 19)
      hipLaunchKernelGGL((k_volume_source<19, bf16s>), g, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_volume_source<27, bf16s>), g, b, 0, s, a);
  }
  return hipGetLastError();
}

// fluid cells of non-periodic faces that no open face covers
__global__ void k_uncovered(Dom d, const uint8_t* solid, uint32_t open_bits, unsigned int* count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N || solid[i]) return;
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  const uint32_t c[3] = {x, y, z}, n[3] = {d.nx, d.ny, d.nz};
  bool bad = false;
  for (int ax = 0; ax < 3; ++ax) {
    if (d.per[ax]) continue;
    if (c[ax] == 0 && !((open_bits >> (2 * ax)) & 1u)) bad = true;
    if (c[ax] == n[ax] - 1 && !((open_bits >> (2 * ax + 1)) & 1u)) bad = true;
  }
  if (bad) atomicAdd(count, 1u);
}

static uint32_t max_face_cells(const Dom& d) {
  const uint32_t a = d.ny * d.nz, b = d.nx * d.nz, c = d.nx * d.ny;
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

template <typename S>
static hipError_t open_s(int Q, const FaceArgs& a, hipStream_t s) {
  if (a.nfaces == 0) return hipSuccess;
  const dim3 block(256), grid((max_face_cells(a.d) + 255) / 256, (unsigned)a.nfaces);
  bool one_axis = true;
  for (int k = 1; k < a.nfaces; ++k) one_axis = one_axis && a.faces[k] / 2 == a.faces[0] / 2;
  if (one_axis) {
    if (Q == 19)
      hipLaunchKernelGGL((k_face_bc<19, S, true>), grid, block, 0, s, a, 0);
    else
      hipLaunchKernelGGL((k_face_bc<27, S, true>), grid, block, 0, s, a, 0);
    return hipGetLastError();
  }
  if (Q == 19) {
    hipLaunchKernelGGL((k_face_stash<19, S, false>), grid, block, 0, s, a);
    hipLaunchKernelGGL((k_face_stash<19, S, true>), grid, block, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_face_stash<27, S, false>), grid, block, 0, s, a);
    hipLaunchKernelGGL((k_face_stash<27, S, true>), grid, block, 0, s, a);
  }
  for (int k = 0; k < a.nfaces; ++k) {
    const int face = a.faces[k];
    const dim3 g1((max_face_cells(a.d) + 255) / 256);
    if (Q == 19)
      hipLaunchKernelGGL((k_face_bc<19, S, false>), g1, block, 0, s, a, face);
    else
      hipLaunchKernelGGL((k_face_bc<27, S, false>), g1, block, 0, s, a, face);
  }
  return hipGetLastError();
}

hipError_t launch_open_faces(int Q, int storage, const FaceArgs& a, hipStream_t s) {
  if (storage == kF32) return open_s<float>(Q, a, s);
  if (storage == kF16) return open_s<_Float16>(Q, a, s);
  return open_s<bf16s>(Q, a, s);
}

hipError_t launch_uncovered(const Dom& d, const uint8_t* solid, uint32_t open_bits, unsigned int* count,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_uncovered, dim3((d.N + 255) / 256), dim3(256), 0, s, d, solid, open_bits, count);
  return hipGetLastError();
}

}  // namespace lbm
