// lbm_aux.hip — rotating immersed-boundary direct forcing and the passive
// scalar advection-diffusion distributions on the device.
//
// IBM (Solver::apply_rotating_ibm, solver.rs:3016-3282, rotating_ibm.rs):
// marker positions are frozen, as in the reference, so the marker stencils
// (rotating_ibm.rs:163-224) are built once on the host.  Every per-cell
// quantity the reference accumulates over markers in marker order
// (cell_kernel_sum, spread, du) is gathered here cell-centrically over a
// marker-ordered CSR list, so the f64 sums follow the reference's order and
// the force field is reproduced bit for bit.  A step is two launches: the
// gather of u_now and rho at the stencil cells, then one block per body
// that runs every sweep (per-marker interpolation, mobility and impulse;
// per-cell spread; fixed-tree marker reductions), commits the force field
// and records the diagnostics.  The slip-tolerance early exit is taken
// inside that block, so the host never waits inside a sweep.
//
// Scalar ADE: BGK collide toward h_eq = w C (1 + c.u/cs^2) with the cached
// concentration and the post-step hydro velocity, pull streaming with
// half-way bounce-back at solids, C = sum h (solver.rs:2113-2205,
// scalar.rs:6-15).  The hydro lattice (reference) and D3Q7 (cs^2 = 1/4) are
// both supported.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lbm_aux.hpp"
#include "lbm_device.hpp"

namespace lbm {

namespace {

inline uint32_t nblk(size_t n, uint32_t bs = 256) { return (uint32_t)((n + bs - 1) / bs); }

#define AUXCHK(v, expr)                                                      \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      if ((v).ctx_err) *(v).ctx_err = std::string(#expr ": ") + hipGetErrorString(_e); \
      return LBM_ERR_DEVICE;                                                 \
    }                                                                        \
  } while (0)

}  // namespace

// ===========================================================================
// rotating IBM
// ===========================================================================
// A device IBM state holds K bodies, one per domain ("instance"), with every
// array concatenated instance after instance; a single domain is K = 1.
// Markers and stencil cells carry their instance, the per-instance
// reductions run one block per instance, and the early exit is per instance.
struct IbmDevDiag {
  int iterations;
  int converged;
  double slip_max, slip_rms, slip_max_rel, slip_rms_rel;
  double torque[3], fluid_force[3], marker_force[3];
  double mom_err;
};

// host description of one body (stencils built once: markers are frozen)
struct IbmHost {
  int64_t M = 0;
  double center[3], omega[3], scale = 1.0;
  std::vector<double> pos, weight, target;
  std::vector<uint32_t> mk_off, pt_u, pt_m, uc_cell, uc_off, uc_ent;
  std::vector<double> pt_w, ksum;
};

struct IbmState {
  int K = 0;
  int64_t M = 0;
  uint32_t P = 0, U = 0;
  lbm_ibm_config cfg{};
  IbmHost host;  // K == 1: the body this state was built from
  // per instance
  uint32_t* moff = nullptr;  // K + 1 marker offsets
  uint32_t* uoff = nullptr;  // K + 1 stencil-cell offsets
  // fused path: dynamic LDS of k_ibm_sweeps (0: the wide path), sized for
  // the largest body (mmax markers, umax stencil cells)
  size_t lds_bytes = 0;
  int level = 0;
  uint32_t mmax = 0, umax = 0, pmax = 0;
  double* cen = nullptr;     // K x 3
  double* scale = nullptr;   // K
  IbmDomain* dom = nullptr;  // K
  std::vector<IbmDomain> dom_h;
  bool dom_valid = false;
  // marker-major
  double* pos = nullptr;       // M x 3
  double* weight = nullptr;    // M
  double* target = nullptr;    // M x 3
  uint32_t* mk_off = nullptr;  // M + 1
  uint32_t* mk_inst = nullptr; // M
  uint32_t* pt_u = nullptr;    // P: unique-cell index
  uint32_t* pt_m = nullptr;    // P: marker index
  double* pt_w = nullptr;      // P
  // cell-major
  uint32_t* uc_cell = nullptr;  // U: cell index within its instance
  uint32_t* uc_inst = nullptr;  // U
  uint32_t* uc_off = nullptr;   // U + 1
  uint32_t* uc_ent = nullptr;   // P, marker order within a cell
  uint32_t* uc_m = nullptr;     // P: marker of entry e (pt_m[uc_ent[e]])
  double* uc_w = nullptr;       // P: weight of entry e (pt_w[uc_ent[e]])
  double* ksum = nullptr;       // U
  // per-apply work
  double* unow = nullptr;    // U x 3
  double* rho = nullptr;     // U
  double* du = nullptr;      // U x 3
  double* spread = nullptr;  // U x 3
  double* imp = nullptr;     // M x 3
  double* imp_next = nullptr;  // M x 3: the next sweep's impulses, computed with this sweep's slip
  double* mrep = nullptr;    // M x 3
  double* mtq = nullptr;     // M x 3
  double* mslip = nullptr;   // M
  uint8_t* mact = nullptr;   // M
  IbmDevDiag* diag = nullptr;  // K
  double* integ = nullptr;     // K x 4: running sums of the torque (x, y, z) and the update count
  // adjoint (single body): per-sweep impulses and du recorded by the
  // forward update, and the reverse-sweep work arrays
  double* imp_hist = nullptr;  // max_iterations x M x 3
  double* du_hist = nullptr;   // max_iterations x U x 3
  double* a_T = nullptr;       // M x 3  adjoint of the target velocity
  double* a_um = nullptr;      // M x 3  adjoint of the interpolated velocity
  double* a_mob = nullptr;     // M      adjoint of the mobility
  double* a_du = nullptr;      // U x 3
  double* a_unow = nullptr;    // U x 3
  double* a_rho = nullptr;     // U
};

// u_now = (m + F/2) / rho at the stencil cells, rho the cached density
// (solver.rs:3040-3084: f64 momentum of the populations, force without the
// just-cleared field).
// With reset_K > 0 the first reset_K threads also clear the instances'
// diagnostics (no other kernel of the gather reads them).  With nclr > 0
// the first nclr threads also clear the force field at the listed cells
// (the runner's per-step clear, runner.rs:1079-1129): the gather does not
// read the field then (field_present is 0), and the commit that adds the new
// spread runs in a later launch.
template <int Q, typename S>
__global__ void k_ibm_gather(const IbmDomain* dom, int odd, const uint32_t* uc_cell, const uint32_t* uc_inst,
                             uint32_t U, int field_present, double* unow, double* rho, double* du, double* spread,
                             IbmDevDiag* diag, uint32_t reset_K, const uint32_t* clr, uint32_t nclr, float* clrF) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < reset_K) memset(diag + u, 0, sizeof(IbmDevDiag));
  if (u < nclr) {
    const size_t i = clr[u];
    clrF[3 * i + 0] = 0.0f;
    clrF[3 * i + 1] = 0.0f;
    clrF[3 * i + 2] = 0.0f;
  }
  if (u >= U) return;
  const IbmDomain& D = dom[uc_inst[u]];
  const Dom d = D.d;
  const uint32_t i = uc_cell[u];
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  const Nbr nb(d, x, y, z);
  const uint32_t m = D.mask[i];
  const S* f = (const S*)D.f;
  float dr = 0.0f;
  double mom[3] = {0.0, 0.0, 0.0};
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float v = read_pop<Q, S, q>(f, d.ps, nb, i, m, odd != 0);
    dr = dr + v;
    if constexpr (lat_c(q, 0) != 0) mom[0] += (double)lat_c(q, 0) * (double)v;
    if constexpr (lat_c(q, 1) != 0) mom[1] += (double)lat_c(q, 1) * (double)v;
    if constexpr (lat_c(q, 2) != 0) mom[2] += (double)lat_c(q, 2) * (double)v;
  });
  const float r = 1.0f + dr;
  double force[3] = {(double)D.F[0], (double)D.F[1], (double)D.F[2]};
  if (field_present && D.F_cur)
    for (int a = 0; a < 3; ++a) force[a] += (double)D.F_cur[3 * (size_t)i + a];
  const double inv_rho = 1.0 / fmax((double)r, 1.0e-30);
  for (int a = 0; a < 3; ++a) {
    unow[3 * (size_t)u + a] = (mom[a] + 0.5 * force[a]) * inv_rho;
    du[3 * (size_t)u + a] = 0.0;
    spread[3 * (size_t)u + a] = 0.0;
  }
  rho[u] = (double)r;
}

// ---- wide path (bodies whose stencil data do not fit in LDS) -----------------
__global__ void k_ibm_reset(IbmDevDiag* diag, int K) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K) memset(diag + k, 0, sizeof(IbmDevDiag));
}

// The wide path's per-marker and per-cell sums walk short lists (a marker's
// stencil points, a cell's marker entries) whose entries point at data
// written by another kernel, usually in another XCD's L2: one thread per
// marker made each entry two or three dependent far-cache round trips.  Here
// one wave takes a marker (or cell): each lane loads one entry, all loads in
// flight at once, and the sum then runs over the entries in list order in
// every lane from values read out of the loading lanes (v_readlane with a
// wave-uniform index), each term in the statement form of the one-thread
// loop, so the f64 results are the same bits.
__device__ __forceinline__ double lane_f64(double v, int j) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), j);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Per marker over its stencil points in point order: the interpolated
// velocity sum_p w (u_now + du) and, with `mob`, the mobility row sum
// sum_p w ksum / rho (solver.rs:3107-3208).  One wave per marker.
__device__ __forceinline__ void marker_point_sums(uint32_t p0, uint32_t p1, const uint32_t* pt_u, const double* pt_w,
                                                  const double* unow, const double* du, const double* rho,
                                                  const double* ksum, bool want_mob, double um[3], double* mob,
                                                  double* w_of_lane) {
  const int lane = threadIdx.x & 63;
  for (uint32_t base = p0; base < p1; base += 64) {
    const uint32_t p = base + lane;
    double w = 0.0, t[3] = {0.0, 0.0, 0.0}, q = 0.0;
    if (p < p1) {
      const uint32_t u = pt_u[p];
      w = pt_w[p];
      for (int a = 0; a < 3; ++a) t[a] = unow[3 * (size_t)u + a] + du[3 * (size_t)u + a];
      if (want_mob) q = w * ksum[u] / fmax(rho[u], 1.0e-30);
    }
    if (w_of_lane && base == p0) *w_of_lane = w;
    const int n = (int)min(64u, p1 - base);
    for (int j = 0; j < n; ++j) {
      const double wj = lane_f64(w, j);
      for (int a = 0; a < 3; ++a) um[a] += wj * lane_f64(t[a], j);
      if (want_mob) *mob += lane_f64(q, j);
    }
  }
}

// per marker: interpolated velocity, mobility row sum, impulse (one wave
// per marker)
__global__ __launch_bounds__(256) void k_ibm_impulse(int64_t M, const uint32_t* mk_off, const uint32_t* mk_inst,
                                                     const uint32_t* pt_u, const double* pt_w, const double* target,
                                                     const double* unow, const double* du, const double* rho,
                                                     const double* ksum, double relax, double* imp,
                                                     const IbmDevDiag* diag) {
  const int64_t m = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= M || diag[mk_inst[m]].converged) return;
  double um[3] = {0.0, 0.0, 0.0}, mob = 0.0;
  marker_point_sums(mk_off[m], mk_off[m + 1], pt_u, pt_w, unow, du, rho, ksum, true, um, &mob, nullptr);
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      const double slip = target[3 * m + a] - um[a];
      imp[3 * m + a] = mob == 0.0 ? 0.0 : relax * slip / mob;
    }
}

// per cell: spread and du accumulated over the cell's markers in marker
// order (one wave per cell)
__global__ __launch_bounds__(256) void k_ibm_spread(uint32_t U, const uint32_t* uc_off, const uint32_t* uc_inst,
                                                    const uint32_t* uc_ent, const uint32_t* pt_m, const double* pt_w,
                                                    const double* imp, const double* rho, double* spread, double* du,
                                                    const IbmDevDiag* diag) {
  const uint64_t u = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (u >= U || diag[uc_inst[u]].converged) return;
  const int lane = threadIdx.x & 63;
  const double inv_rho = 1.0 / fmax(rho[u], 1.0e-30);
  double s[3], dd[3];
  for (int a = 0; a < 3; ++a) {
    s[a] = spread[3 * (size_t)u + a];
    dd[a] = du[3 * (size_t)u + a];
  }
  const uint32_t e0 = uc_off[u], e1 = uc_off[u + 1];
  for (uint32_t base = e0; base < e1; base += 64) {
    const uint32_t e = base + lane;
    double w = 0.0, im[3] = {0.0, 0.0, 0.0};
    if (e < e1) {
      const uint32_t p = uc_ent[e];
      const uint32_t m = pt_m[p];
      w = pt_w[p];
      for (int a = 0; a < 3; ++a) im[a] = imp[3 * (size_t)m + a];
    }
    const int n = (int)min(64u, e1 - base);
    for (int j = 0; j < n; ++j) {
      const double wj = lane_f64(w, j);
      for (int a = 0; a < 3; ++a) {
        const double cf = lane_f64(im[a], j) * wj;
        s[a] += cf;
        dd[a] += cf * inv_rho;
      }
    }
  }
  if (lane == 0)
    for (int a = 0; a < 3; ++a) {
      spread[3 * (size_t)u + a] = s[a];
      du[3 * (size_t)u + a] = dd[a];
    }
}

// per marker: represented force, torque contribution, post-sweep slip; with
// imp_next, also the next sweep's impulse from the same interpolated
// velocity (what k_ibm_impulse of the next sweep computes, operation for
// operation), which k_ibm_reduce installs if the instance goes on (one wave
// per marker)
__global__ __launch_bounds__(256) void k_ibm_marker_diag(int64_t M, const uint32_t* mk_off, const uint32_t* mk_inst,
                                                         const uint32_t* pt_u, const double* pt_w, const double* pos,
                                                         const double* target, const double* unow, const double* du,
                                                         const double* imp, const double* cen, double* mrep,
                                                         double* mtq, double* mslip, uint8_t* mact,
                                                         const IbmDevDiag* diag, const double* rho, const double* ksum,
                                                         double relax, double* imp_next) {
  const int64_t m = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= M) return;
  const uint32_t k = mk_inst[m];
  if (diag[k].converged) return;
  const uint32_t p0 = mk_off[m], p1 = mk_off[m + 1];
  double rep[3] = {0.0, 0.0, 0.0}, um[3] = {0.0, 0.0, 0.0};
  double mob = 0.0;
  double w0 = 0.0;  // this lane's point weight of the first chunk
  marker_point_sums(p0, p1, pt_u, pt_w, unow, du, rho, ksum, imp_next != nullptr, um, &mob, &w0);
  double im[3];
  for (int a = 0; a < 3; ++a) im[a] = imp[3 * m + a];
  for (uint32_t base = p0; base < p1; base += 64) {
    const uint32_t p = base + (threadIdx.x & 63);
    const double w = base == p0 ? w0 : (p < p1 ? pt_w[p] : 0.0);
    const int n = (int)min(64u, p1 - base);
    for (int j = 0; j < n; ++j) {
      const double wj = lane_f64(w, j);
      for (int a = 0; a < 3; ++a) rep[a] += im[a] * wj;
    }
  }
  if ((threadIdx.x & 63) != 0) return;
  if (imp_next)
    for (int a = 0; a < 3; ++a) {
      const double slip = target[3 * m + a] - um[a];
      imp_next[3 * m + a] = mob == 0.0 ? 0.0 : relax * slip / mob;
    }
  const double r[3] = {pos[3 * m] - cen[3 * k], pos[3 * m + 1] - cen[3 * k + 1], pos[3 * m + 2] - cen[3 * k + 2]};
  mtq[3 * m + 0] = r[1] * rep[2] - r[2] * rep[1];
  mtq[3 * m + 1] = r[2] * rep[0] - r[0] * rep[2];
  mtq[3 * m + 2] = r[0] * rep[1] - r[1] * rep[0];
  for (int a = 0; a < 3; ++a) mrep[3 * m + a] = rep[a];
  const double s[3] = {target[3 * m] - um[0], target[3 * m + 1] - um[1], target[3 * m + 2] - um[2]};
  mslip[m] = sqrt(0.0 + s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  mact[m] = p1 > p0 ? 1 : 0;
}

// one block per instance: fixed-tree reduction over its markers and the
// convergence test
__global__ __launch_bounds__(256) void k_ibm_reduce(const uint32_t* moff, const double* mrep, const double* mtq,
                                                    const double* mslip, const uint8_t* mact, const double* weight,
                                                    int iter, const double* scale, double tol, IbmDevDiag* diags,
                                                    const double* imp_next, double* imp) {
  __shared__ double sh[9][256];
  __shared__ int conv;
  const int k = blockIdx.x;
  IbmDevDiag* diag = diags + k;
  if (threadIdx.x == 0) conv = diag->converged;
  __syncthreads();
  if (conv) return;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // tq[3], rep[3], slipmax, ssq, wsum
  // each thread's markers in their order, four at a time with every load
  // issued before the first add (the loads are independent; one far-cache
  // latency per four markers instead of one per marker)
  const int64_t m1 = moff[k + 1];
  for (int64_t mb = (int64_t)moff[k] + threadIdx.x; mb < m1; mb += 4 * 256) {
    double tq[4][3], rp[4][3], sl[4], wt[4];
    bool ac[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t m = mb + 256 * j;
      ac[j] = false;
      sl[j] = wt[j] = 0.0;
      for (int a = 0; a < 3; ++a) tq[j][a] = rp[j][a] = 0.0;
      if (m < m1) {
        for (int a = 0; a < 3; ++a) {
          tq[j][a] = mtq[3 * m + a];
          rp[j][a] = mrep[3 * m + a];
        }
        ac[j] = mact[m] != 0;
        sl[j] = mslip[m];
        wt[j] = weight[m];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (mb + 256 * j >= m1) break;
      for (int a = 0; a < 3; ++a) {
        acc[a] += tq[j][a];
        acc[3 + a] += rp[j][a];
      }
      if (ac[j]) {
        acc[6] = fmax(acc[6], sl[j]);
        acc[7] += wt[j] * sl[j] * sl[j];
        acc[8] += wt[j];
      }
    }
  }
  for (int j = 0; j < 9; ++j) sh[j][threadIdx.x] = acc[j];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int j = 0; j < 9; ++j)
        sh[j][threadIdx.x] = j == 6 ? fmax(sh[j][threadIdx.x], sh[j][threadIdx.x + w])
                                    : sh[j][threadIdx.x] + sh[j][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) {
      diag->torque[a] -= sh[a][0];
      diag->marker_force[a] += sh[3 + a][0];
      diag->fluid_force[a] += sh[3 + a][0];
    }
    diag->iterations = iter + 1;
    diag->slip_max = sh[6][0];
    diag->slip_rms = sh[8][0] > 0.0 ? sqrt(sh[7][0] / sh[8][0]) : 0.0;
    diag->slip_max_rel = diag->slip_max / scale[k];
    diag->slip_rms_rel = diag->slip_rms / scale[k];
    const double* ff = diag->fluid_force;
    const double* mf = diag->marker_force;
    const double e[3] = {ff[0] - mf[0], ff[1] - mf[1], ff[2] - mf[2]};
    const double ne = sqrt(0.0 + e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    const double nm = sqrt(0.0 + mf[0] * mf[0] + mf[1] * mf[1] + mf[2] * mf[2]);
    diag->mom_err = ne / fmax(nm, 1.0e-12);
    if (diag->slip_max_rel <= tol) diag->converged = 1;
    conv = diag->converged;
  }
  // an instance that goes on takes the next sweep's impulses: 32 loads
  // issued before their 32 stores (the two arrays could alias as far as the
  // compiler knows, so a plain loop waits for each load in turn)
  __syncthreads();
  if (imp_next && !conv) {
    constexpr int kBatch = 32;
    const int64_t e1 = 3 * (int64_t)moff[k + 1];
    for (int64_t e = 3 * (int64_t)moff[k] + threadIdx.x; e < e1; e += kBatch * 256) {
      double v[kBatch];
#pragma unroll
      for (int j = 0; j < kBatch; ++j) v[j] = e + 256 * j < e1 ? imp_next[e + 256 * j] : 0.0;
#pragma unroll
      for (int j = 0; j < kBatch; ++j)
        if (e + 256 * j < e1) imp[e + 256 * j] = v[j];
    }
  }
}

// force_field += T(spread)

__global__ void k_zero_integ(double* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// one block: append the K diagnostics of this update to a device ring (read
// back after a span of steps, so the step loop needs no host
// synchronisation), then advance the ring position (ring may be null); the
// running f64 torque integral of each instance advances by this update's
// torque
__device__ __forceinline__ void ibm_record_block(const IbmDevDiag* diag, int K, IbmDevDiag* ring, uint32_t cap,
                                                 uint32_t* pos, double* integ) {
  __shared__ uint32_t slot;
  if (threadIdx.x == 0 && ring) slot = *pos;
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    if (ring && slot < cap) ring[(size_t)slot * K + k] = diag[k];
    for (int c = 0; c < 3; ++c) integ[4 * k + c] += diag[k].torque[c];
    integ[4 * k + 3] += 1.0;
  }
  __syncthreads();
  if (threadIdx.x == 0 && ring) *pos = slot + 1;
}

// The wide path's last launch of an update, two independent jobs: blocks
// 0 .. nb-1 add the spread to the force field at the stencil cells
// (force_field += T(spread)), the last block records the diagnostics.
__global__ __launch_bounds__(256) void k_ibm_commit_record(uint32_t U, uint32_t nb, const uint32_t* uc_cell,
                                                           const uint32_t* uc_inst, const IbmDomain* dom,
                                                           const double* spread, const IbmDevDiag* diag, int K,
                                                           IbmDevDiag* ring, uint32_t cap, uint32_t* pos,
                                                           double* integ) {
  if (blockIdx.x == nb) {
    ibm_record_block(diag, K, ring, cap, pos, integ);
    return;
  }
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  float* F = dom[uc_inst[u]].F_cur;
  const size_t i = uc_cell[u];
  for (int a = 0; a < 3; ++a) F[3 * i + a] = F[3 * i + a] + (float)spread[3 * (size_t)u + a];
}

// ---------------------------------------------------------------------------
// Fused path: all direct-forcing sweeps of one body in one block (1024
// threads), after the gather: the impulses, the spread and du update, the
// per-marker diagnostics and their reductions for every sweep (solver.rs:
// 3107-3208), then the force-field commit and the diagnostics record.  One
// launch per step, one block per body, every body of a batch at once; the
// block leaves the sweep loop as soon as its body meets the slip tolerance
// (device-side test, no host wait).
//
// LDS staging, copied in once per step: always the body's per-cell state
// (u_now, du, spread, rho, kernel sum: 11 doubles per stencil cell) and its
// marker impulses (3 per marker); when they fit as well (level 2), the
// markers' stencils (point offsets, cell and weight per point) with their
// targets, positions and weights, so the marker phases never leave the CU.
// Bodies whose state does not fit take the wide path.
// Order of every f64 operation is the reference's and the wide path's: per
// cell over its markers in marker order, per marker over its stencil points
// in point order, and the marker reductions on a fixed 256-lane tree (lane t
// takes markers t, t + 256, ... in order, then a halving tree), so the force
// field is the reference's bit for bit and both paths agree to the bit.
// ---------------------------------------------------------------------------
struct IbmSweepArgs {
  const uint32_t *moff, *uoff, *mk_off, *pt_u, *uc_off, *uc_m, *uc_cell;
  const double *pt_w, *uc_w, *target, *pos, *cen, *weight, *scale, *ksum, *unow, *rho;
  IbmDevDiag* diag;
  double* integ;  // K x 4 running torque sums and update count
  const IbmDomain* dom;
  double relax, tol;
  int max_iter;
  int level;                  // 1: cell state + impulses in LDS; 2: also the marker stencils
  uint32_t mmax, umax, pmax;  // LDS capacity in markers, cells and stencil points
  // adjoint recording (single body): du before and imp of every sweep
  double *du_hist, *imp_hist;
  int64_t M;
  uint32_t U;
  // diagnostics ring (null: none); ring_pos[0] = slot, ring_pos[1] = blocks done
  IbmDevDiag* ring;
  uint32_t cap;
  uint32_t* ring_pos;
  int K;
};

// Dynamic LDS bytes of k_ibm_sweeps at a staging level.
static inline size_t ibm_lds_bytes(int level, uint32_t mmax, uint32_t umax, uint32_t pmax) {
  size_t b = ((size_t)mmax * 3 + (size_t)umax * 11) * sizeof(double);
  if (level >= 2) b += ((size_t)mmax * 7 + pmax) * sizeof(double) + ((size_t)mmax + 1 + pmax) * sizeof(uint32_t);
  return b;
}

// One body's view: LDS arrays (cell u at u - u0, marker m at m - m0, point p
// at p - p0) or, for what is not staged, the global arrays.
struct IbmView {
  double *imp, *unow, *du, *spread, *rho, *ksum;
  const uint32_t *mk_off, *pt_u;
  const double *pt_w, *target, *pos, *weight;
  int64_t m0;
  uint32_t u0, p0;  // p0: subtracted from point indices (0 when unstaged)
  int64_t mb;       // subtracted from marker indices of the marker arrays
};

// Per-marker sums over the stencil points in point order (solver.rs:
// 3118-3140): the interpolated velocity sum w (u_now + du), with MOB the
// mobility row sum w ksum / rho, and the represented force sum imp w.
template <bool MOB>
__device__ __forceinline__ void marker_sums(const IbmView& V, int64_t m, const double im[3], double um[3],
                                            double& mob, double rep[3]) {
  const uint32_t p0 = V.mk_off[m - V.mb] - V.p0, p1 = V.mk_off[m - V.mb + 1] - V.p0;
  for (uint32_t p = p0; p < p1; ++p) {
    const size_t u = V.pt_u[p] - V.u0;
    const double w = V.pt_w[p];
    for (int c = 0; c < 3; ++c) {
      rep[c] += im[c] * w;
      um[c] += w * (V.unow[3 * u + c] + V.du[3 * u + c]);
    }
    if (MOB) mob += w * V.ksum[u] / fmax(V.rho[u], 1.0e-30);
  }
}

__global__ __launch_bounds__(1024) void k_ibm_sweeps(IbmSweepArgs a) {
  extern __shared__ double dyn[];
  __shared__ double sh[9][256];
  __shared__ int hand[256];  // lanes 256-511 hand their marker's presence to lanes 0-255
  __shared__ int conv;
  const int k = blockIdx.x;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int64_t m0 = a.moff[k], m1 = a.moff[k + 1];
  const uint32_t u0 = a.uoff[k], u1 = a.uoff[k + 1];
  const uint32_t nm = (uint32_t)(m1 - m0), nu = u1 - u0;
  IbmDevDiag* diag = a.diag + k;
  IbmView V;
  V.m0 = m0;
  V.u0 = u0;
  V.imp = dyn;
  V.unow = V.imp + 3 * (size_t)a.mmax;
  V.du = V.unow + 3 * (size_t)a.umax;
  V.spread = V.du + 3 * (size_t)a.umax;
  V.rho = V.spread + 3 * (size_t)a.umax;
  V.ksum = V.rho + a.umax;
  if (tid == 0) {
    memset(diag, 0, sizeof(IbmDevDiag));
    conv = 0;
  }
  // stage: cell state (du, spread start at zero); at level 2 the markers too
  for (uint32_t l = tid; l < nu; l += nth) {
    const size_t u = u0 + l;
    for (int c = 0; c < 3; ++c) {
      V.unow[3 * l + c] = a.unow[3 * u + c];
      V.du[3 * l + c] = 0.0;
      V.spread[3 * l + c] = 0.0;
    }
    V.rho[l] = a.rho[u];
    V.ksum[l] = a.ksum[u];
  }
  if (a.level >= 2) {
    double* tg = V.ksum + a.umax;
    double* ps = tg + 3 * (size_t)a.mmax;
    double* wt = ps + 3 * (size_t)a.mmax;
    double* pw = wt + a.mmax;
    uint32_t* mo = (uint32_t*)(pw + a.pmax);
    uint32_t* pu = mo + a.mmax + 1;
    const uint32_t pbeg = a.mk_off[m0], np = a.mk_off[m1] - pbeg;
    for (uint32_t l = tid; l < nm; l += nth) {
      const int64_t m = m0 + l;
      for (int c = 0; c < 3; ++c) {
        tg[3 * l + c] = a.target[3 * m + c];
        ps[3 * l + c] = a.pos[3 * m + c];
      }
      wt[l] = a.weight[m];
    }
    for (uint32_t l = tid; l <= nm; l += nth) mo[l] = a.mk_off[m0 + l];
    for (uint32_t l = tid; l < np; l += nth) {
      pu[l] = a.pt_u[pbeg + l];
      pw[l] = a.pt_w[pbeg + l];
    }
    V.mk_off = mo;
    V.pt_u = pu;
    V.pt_w = pw;
    V.target = tg;
    V.pos = ps;
    V.weight = wt;
    V.mb = m0;
    V.p0 = pbeg;
  } else {
    V.mk_off = a.mk_off;
    V.pt_u = a.pt_u;
    V.pt_w = a.pt_w;
    V.target = a.target;
    V.pos = a.pos;
    V.weight = a.weight;
    V.mb = 0;
    V.p0 = 0;
  }
  __syncthreads();
  for (int it = 0; it < a.max_iter; ++it) {
    if (a.du_hist)
      for (uint32_t l = tid; l < nu; l += nth)
        for (int c = 0; c < 3; ++c) a.du_hist[3 * ((size_t)a.U * it + u0 + l) + c] = V.du[3 * (size_t)l + c];
    if (it == 0) {  // later sweeps take the impulses the previous sweep's diagnostics computed
      for (int64_t m = m0 + tid; m < m1; m += nth) {
        double um[3] = {0.0, 0.0, 0.0}, mob = 0.0, rep[3] = {0.0, 0.0, 0.0};
        const double none[3] = {0.0, 0.0, 0.0};
        marker_sums<true>(V, m, none, um, mob, rep);
        const double* t = V.target + 3 * (m - V.mb);
        for (int c = 0; c < 3; ++c) {
          const double slip = t[c] - um[c];
          V.imp[3 * (m - m0) + c] = mob == 0.0 ? 0.0 : a.relax * slip / mob;
        }
      }
      __syncthreads();
    }
    if (a.imp_hist)
      for (uint32_t l = tid; l < nm; l += nth)
        for (int c = 0; c < 3; ++c) a.imp_hist[3 * ((size_t)a.M * it + m0 + l) + c] = V.imp[3 * (size_t)l + c];
    // spread and du per cell, over the cell's markers in marker order
    for (uint32_t u = u0 + tid; u < u1; u += nth) {
      const uint32_t l = u - u0;
      const double inv_rho = 1.0 / fmax(V.rho[l], 1.0e-30);
      double sp[3], dd[3];
      for (int c = 0; c < 3; ++c) {
        sp[c] = V.spread[3 * l + c];
        dd[c] = V.du[3 * l + c];
      }
      const uint32_t e0 = a.uc_off[u], e1 = a.uc_off[u + 1];
      for (uint32_t eb = e0; eb < e1; eb += 4) {
        int64_t mm[4];
        double ww[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t e = min(eb + (uint32_t)j, e1 - 1);
          mm[j] = (int64_t)a.uc_m[e] - m0;
          ww[j] = a.uc_w[e];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (eb + (uint32_t)j >= e1) break;
          for (int c = 0; c < 3; ++c) {
            const double cf = V.imp[3 * mm[j] + c] * ww[j];
            sp[c] += cf;
            dd[c] += cf * inv_rho;
          }
        }
      }
      for (int c = 0; c < 3; ++c) {
        V.spread[3 * l + c] = sp[c];
        V.du[3 * l + c] = dd[c];
      }
    }
    __syncthreads();
    // per marker, on the reduction's 256 lanes (lane t: markers t, t + 256,
    // ... in order): represented force, torque, post-sweep slip, and (when a
    // sweep follows) the next sweep's impulse from the same interpolated
    // velocity, written over this sweep's impulse of the same marker; the
    // lane accumulates its reduction terms as it goes
    const bool next = it + 1 < a.max_iter;
    // One marker's terms: torque and represented force, post-sweep slip
    // (and, when a sweep follows, its next impulse from the same
    // interpolated velocity, written over this sweep's impulse).
    auto marker_terms = [&](int64_t m, double* tm, bool& haspts) {
      double rep[3] = {0.0, 0.0, 0.0}, um[3] = {0.0, 0.0, 0.0};
      double im[3];
      for (int c = 0; c < 3; ++c) im[c] = V.imp[3 * (m - m0) + c];
      double mob = 0.0;
      if (next)
        marker_sums<true>(V, m, im, um, mob, rep);
      else
        marker_sums<false>(V, m, im, um, mob, rep);
      const double* t = V.target + 3 * (m - V.mb);
      if (next)
        for (int c = 0; c < 3; ++c) {
          const double slip = t[c] - um[c];
          V.imp[3 * (m - m0) + c] = mob == 0.0 ? 0.0 : a.relax * slip / mob;
        }
      const double* pm = V.pos + 3 * (m - V.mb);
      const double r[3] = {pm[0] - a.cen[3 * k], pm[1] - a.cen[3 * k + 1], pm[2] - a.cen[3 * k + 2]};
      tm[0] = r[1] * rep[2] - r[2] * rep[1];
      tm[1] = r[2] * rep[0] - r[0] * rep[2];
      tm[2] = r[0] * rep[1] - r[1] * rep[0];
      for (int c = 0; c < 3; ++c) tm[3 + c] = rep[c];
      const double sl[3] = {t[0] - um[0], t[1] - um[1], t[2] - um[2]};
      const double slip = sqrt(0.0 + sl[0] * sl[0] + sl[1] * sl[1] + sl[2] * sl[2]);
      haspts = V.mk_off[m - V.mb + 1] > V.mk_off[m - V.mb];
      const double w = haspts ? V.weight[m - V.mb] : 0.0;
      tm[6] = slip;
      tm[7] = w * slip * slip;
      tm[8] = w;
    };
    auto accumulate = [&](double* acc, const double* tm, bool haspts) {
      for (int c = 0; c < 6; ++c) acc[c] += tm[c];
      if (haspts) {
        acc[6] = fmax(acc[6], tm[6]);
        acc[7] += tm[7];
        acc[8] += tm[8];
      }
    };
    // The 256 accumulators of the reduction (lane t: markers t, t + 256,
    // t + 512, ... in that order), with the markers computed on 512 lanes:
    // per round, lanes 0-255 take chunk 2r (markers 512r + t) and lanes
    // 256-511 chunk 2r + 1, whose terms lane t adds after its own through
    // LDS; the sums are those of the one-group loop.
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // tq[3], rep[3], slipmax, ssq, wsum
    for (uint32_t r0 = 0; r0 < nm; r0 += 512) {
      double tm[9];
      bool haspts = false, have = false;
      if (tid < 512) {
        const uint32_t l = r0 + (uint32_t)tid;
        if (l < nm) {
          have = true;
          marker_terms(m0 + (int64_t)l, tm, haspts);
        }
      }
      if (tid >= 256 && tid < 512) {
        for (int j = 0; j < 9; ++j) sh[j][tid - 256] = have ? tm[j] : 0.0;
        hand[tid - 256] = (have ? 1 : 0) | (haspts ? 2 : 0);
      }
      __syncthreads();
      if (tid < 256) {
        if (have) accumulate(acc, tm, haspts);
        if (hand[tid] & 1) {
          double o[9];
          for (int j = 0; j < 9; ++j) o[j] = sh[j][tid];
          accumulate(acc, o, (hand[tid] & 2) != 0);
        }
      }
      __syncthreads();
    }
    if (tid < 256)
      for (int j = 0; j < 9; ++j) sh[j][tid] = acc[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w)
        for (int j = 0; j < 9; ++j)
          sh[j][tid] = j == 6 ? fmax(sh[j][tid], sh[j][tid + w]) : sh[j][tid] + sh[j][tid + w];
      __syncthreads();
    }
    if (tid == 0) {
      for (int c = 0; c < 3; ++c) {
        diag->torque[c] -= sh[c][0];
        diag->marker_force[c] += sh[3 + c][0];
        diag->fluid_force[c] += sh[3 + c][0];
      }
      diag->iterations = it + 1;
      diag->slip_max = sh[6][0];
      diag->slip_rms = sh[8][0] > 0.0 ? sqrt(sh[7][0] / sh[8][0]) : 0.0;
      diag->slip_max_rel = diag->slip_max / a.scale[k];
      diag->slip_rms_rel = diag->slip_rms / a.scale[k];
      const double* ff = diag->fluid_force;
      const double* mf = diag->marker_force;
      const double e[3] = {ff[0] - mf[0], ff[1] - mf[1], ff[2] - mf[2]};
      const double ne = sqrt(0.0 + e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      const double nmf = sqrt(0.0 + mf[0] * mf[0] + mf[1] * mf[1] + mf[2] * mf[2]);
      diag->mom_err = ne / fmax(nmf, 1.0e-12);
      if (diag->slip_max_rel <= a.tol) diag->converged = 1;
      conv = diag->converged;
    }
    __syncthreads();
    if (conv) break;
  }
  // force_field += T(spread)
  if (u1 > u0) {
    float* F = a.dom[k].F_cur;
    for (uint32_t u = u0 + tid; u < u1; u += nth) {
      const size_t i = a.uc_cell[u];
      const uint32_t l = u - u0;
      for (int c = 0; c < 3; ++c) F[3 * i + c] = F[3 * i + c] + (float)V.spread[3 * l + c];
    }
  }
  // the running f64 torque integral of this body
  if (tid == 0) {
    for (int c = 0; c < 3; ++c) a.integ[4 * k + c] += diag->torque[c];
    a.integ[4 * k + 3] += 1.0;
  }
  // append this body's diagnostics to the ring at the current slot; the last
  // block to finish advances the slot (every block has read it by then)
  if (a.ring && tid == 0) {
    // (the slot's value is in hand before the count is bumped; the ring and
    // both words are read only by later launches)
    const uint32_t slot = a.ring_pos[0];
    if (slot < a.cap) a.ring[(size_t)slot * a.K + k] = *diag;
    if (atomicAdd(&a.ring_pos[1], 1u) == (uint32_t)a.K - 1) {
      a.ring_pos[0] = slot + 1;
      a.ring_pos[1] = 0u;
    }
  }
}

// marker_stencil, rotating_ibm.rs:163-224 (host, f64)
static void marker_stencil(const double p[3], const uint32_t dims[3], int radius, std::vector<int64_t>& cells,
                           std::vector<double>& w) {
  std::vector<std::pair<int64_t, double>> axes[3];
  for (int a = 0; a < 3; ++a) {
    const int64_t lo = (int64_t)std::floor(p[a]) - radius + 1;
    const int64_t hi = (int64_t)std::floor(p[a]) + radius;
    for (int64_t i = lo; i <= hi; ++i) {
      if (i < 0 || i >= (int64_t)dims[a]) continue;
      const double r = std::fabs(p[a] - (double)i);
      double k;
      if (radius == 1) {
        k = std::max(1.0 - r, 0.0);
      } else if (r < 0.5) {
        k = 0.75 - r * r;
      } else if (r < 1.5) {
        k = 0.5 * (1.5 - r) * (1.5 - r);
      } else {
        k = 0.0;
      }
      if (k != 0.0) axes[a].push_back({i, k});
    }
  }
  cells.clear();
  w.clear();
  for (auto& ax : axes[0])
    for (auto& ay : axes[1])
      for (auto& az : axes[2]) {
        cells.push_back((az.first * (int64_t)dims[1] + ay.first) * (int64_t)dims[0] + ax.first);
        w.push_back(ax.second * ay.second * az.second);
      }
}

template <typename T>
static hipError_t upload(T** dst, const std::vector<T>& src) {
  const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
  hipError_t e = hipMalloc((void**)dst, bytes);
  if (e != hipSuccess) return e;
  if (!src.empty()) e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

template <typename T>
static hipError_t scratch(T** dst, size_t n) {
  return hipMalloc((void**)dst, std::max<size_t>(1, n) * sizeof(T));
}

void ibm_destroy(IbmState* s) {
  if (!s) return;
  void* ptrs[] = {s->integ,   s->uc_m,    s->uc_w,   s->moff,    s->uoff,   s->cen,    s->scale,   s->dom,    s->pos,    s->weight, s->target, s->mk_off,
                  s->mk_inst, s->pt_u,   s->pt_m,    s->pt_w,   s->uc_cell, s->uc_inst, s->uc_off, s->uc_ent,
                  s->ksum,    s->unow,   s->rho,     s->du,     s->spread, s->imp,    s->imp_next, s->mrep, s->mtq,
                  s->mslip,   s->mact,   s->diag,    s->imp_hist, s->du_hist, s->a_T, s->a_um, s->a_mob,
                  s->a_du,    s->a_unow, s->a_rho};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
}

// Concatenate bodies into one device state (instance k = parts[k]).
static IbmState* ibm_build(const std::vector<const IbmHost*>& parts, const lbm_ibm_config& cfg, std::string* err) {
  IbmState* s = new IbmState();
  s->K = (int)parts.size();
  s->cfg = cfg;
  std::vector<double> P, W, T, cen, scale;
  std::vector<uint32_t> moff{0}, uoff{0}, mk_off{0}, mk_inst, pt_u, pt_m, uc_cell, uc_inst, uc_off, uc_ent;
  std::vector<double> pt_w, ksum;
  for (int k = 0; k < s->K; ++k) {
    const IbmHost& h = *parts[k];
    const uint32_t m0 = (uint32_t)mk_inst.size(), p0 = (uint32_t)pt_u.size(), u0 = (uint32_t)uc_cell.size();
    P.insert(P.end(), h.pos.begin(), h.pos.end());
    W.insert(W.end(), h.weight.begin(), h.weight.end());
    T.insert(T.end(), h.target.begin(), h.target.end());
    for (int a = 0; a < 3; ++a) cen.push_back(h.center[a]);
    scale.push_back(h.scale);
    for (int64_t m = 0; m < h.M; ++m) {
      mk_off.push_back(p0 + h.mk_off[m + 1]);
      mk_inst.push_back((uint32_t)k);
    }
    moff.push_back((uint32_t)mk_inst.size());
    for (size_t p = 0; p < h.pt_u.size(); ++p) {
      pt_u.push_back(u0 + h.pt_u[p]);
      pt_m.push_back(m0 + h.pt_m[p]);
      pt_w.push_back(h.pt_w[p]);
      uc_ent.push_back(p0 + h.uc_ent[p]);
    }
    for (size_t u = 0; u < h.uc_cell.size(); ++u) {
      uc_cell.push_back(h.uc_cell[u]);
      uc_inst.push_back((uint32_t)k);
      uc_off.push_back(p0 + h.uc_off[u]);
      ksum.push_back(h.ksum[u]);
    }
    uoff.push_back((uint32_t)uc_cell.size());
  }
  uc_off.push_back((uint32_t)pt_u.size());
  std::vector<uint32_t> uc_m(uc_ent.size());
  std::vector<double> uc_w(uc_ent.size());
  for (size_t e = 0; e < uc_ent.size(); ++e) {
    uc_m[e] = pt_m[uc_ent[e]];
    uc_w[e] = pt_w[uc_ent[e]];
  }
  s->M = (int64_t)mk_inst.size();
  s->P = (uint32_t)pt_u.size();
  s->U = (uint32_t)uc_cell.size();
  if (s->K == 1) s->host = *parts[0];
  s->dom_h.assign(s->K, IbmDomain{});
  hipError_t e = hipSuccess;
  auto chain = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  chain(upload(&s->moff, moff));
  chain(upload(&s->uoff, uoff));
  // the fused path when the largest body's state fits in LDS beside the
  // kernel's 18 KB of reduction arrays: level 2 (state and marker stencils)
  // if possible, else level 1 (state only); LBM_IBM_PATH=wide forces the
  // wide path (tuning and A/B tests)
  {
    for (size_t k = 0; k + 1 < moff.size(); ++k) {
      s->mmax = std::max(s->mmax, moff[k + 1] - moff[k]);
      s->pmax = std::max(s->pmax, mk_off[moff[k + 1]] - mk_off[moff[k]]);
    }
    for (size_t k = 0; k + 1 < uoff.size(); ++k) s->umax = std::max(s->umax, uoff[k + 1] - uoff[k]);
    int dev = 0, lim = 0;
    const char* path = getenv("LBM_IBM_PATH");
    const bool wide = path && strcmp(path, "wide") == 0;
    const int force_level = path && strcmp(path, "fused1") == 0 ? 1 : 0;
    if (!wide && s->mmax > 0 && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess) {
      const size_t fixed = 9 * 256 * sizeof(double) + 256 * sizeof(int) + 64;  // the kernel's static LDS
      for (int level = force_level ? force_level : 2; level >= 1 && !s->lds_bytes; --level) {
        const size_t need = ibm_lds_bytes(level, s->mmax, s->umax, s->pmax);
        if (need + fixed <= (size_t)lim) {
          s->lds_bytes = need;
          s->level = level;
        }
      }
    }
  }
  chain(upload(&s->cen, cen));
  chain(upload(&s->scale, scale));
  chain(scratch(&s->dom, s->K));
  chain(upload(&s->pos, P));
  chain(upload(&s->weight, W));
  chain(upload(&s->target, T));
  chain(upload(&s->mk_off, mk_off));
  chain(upload(&s->mk_inst, mk_inst));
  chain(upload(&s->pt_u, pt_u));
  chain(upload(&s->pt_m, pt_m));
  chain(upload(&s->pt_w, pt_w));
  chain(upload(&s->uc_cell, uc_cell));
  chain(upload(&s->uc_inst, uc_inst));
  chain(upload(&s->uc_off, uc_off));
  chain(upload(&s->uc_ent, uc_ent));
  chain(upload(&s->uc_m, uc_m));
  chain(upload(&s->uc_w, uc_w));
  chain(upload(&s->ksum, ksum));
  chain(scratch(&s->unow, 3 * (size_t)s->U));
  chain(scratch(&s->rho, s->U));
  chain(scratch(&s->du, 3 * (size_t)s->U));
  chain(scratch(&s->spread, 3 * (size_t)s->U));
  chain(scratch(&s->imp, 3 * (size_t)s->M));
  chain(scratch(&s->imp_next, 3 * (size_t)s->M));
  chain(scratch(&s->mrep, 3 * (size_t)s->M));
  chain(scratch(&s->mtq, 3 * (size_t)s->M));
  chain(scratch(&s->mslip, s->M));
  chain(scratch(&s->mact, s->M));
  chain(scratch(&s->diag, s->K));
  chain(scratch(&s->integ, 4 * (size_t)s->K));
  if (e == hipSuccess) e = hipMemset(s->integ, 0, 4 * (size_t)s->K * sizeof(double));
  if (e != hipSuccess) {
    *err = std::string("IBM device setup failed: ") + hipGetErrorString(e);
    ibm_destroy(s);
    return nullptr;
  }
  return s;
}

IbmState* ibm_create(const FieldsView& v, int64_t n, const double* pos, const double* weight, const double center[3],
                     const double omega[3], const lbm_ibm_config& cfg, std::vector<uint32_t>* cells,
                     std::string* err) {
  const uint32_t dims[3] = {v.d.nx, v.d.ny, v.d.nz};
  for (int64_t m = 0; m < n; ++m) {
    if (!(std::isfinite(pos[3 * m]) && std::isfinite(pos[3 * m + 1]) && std::isfinite(pos[3 * m + 2]) &&
          std::isfinite(weight[m]) && weight[m] > 0.0)) {
      *err = "IBM marker must have finite position and positive finite weight";
      return nullptr;
    }
  }
  // solid flags of the domain, for the stencil's solid skip
  std::vector<uint32_t> mask(v.d.N);
  if (hipMemcpyAsync(mask.data(), v.mask, (size_t)v.d.N * 4, hipMemcpyDeviceToHost, v.stream) != hipSuccess ||
      hipStreamSynchronize(v.stream) != hipSuccess) {
    *err = "mask readback failed";
    return nullptr;
  }
  IbmHost h;
  h.M = n;
  for (int a = 0; a < 3; ++a) {
    h.center[a] = center[a];
    h.omega[a] = omega[a];
  }
  // targets (RotatingBody::target_velocity, rotating_ibm.rs:89-101) and scale
  h.target.assign(3 * n, 0.0);
  h.pos.assign(pos, pos + 3 * n);
  h.weight.assign(weight, weight + n);
  double max_target = 0.0;
  for (int64_t m = 0; m < n; ++m) {
    const double r[3] = {pos[3 * m] - center[0], pos[3 * m + 1] - center[1], pos[3 * m + 2] - center[2]};
    h.target[3 * m + 0] = omega[1] * r[2] - omega[2] * r[1];
    h.target[3 * m + 1] = omega[2] * r[0] - omega[0] * r[2];
    h.target[3 * m + 2] = omega[0] * r[1] - omega[1] * r[0];
    const double* t = &h.target[3 * m];
    max_target = std::max(max_target, std::sqrt(0.0 + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
  }
  h.scale = std::max(max_target, 1.0);
  // stencils, solid cells dropped
  h.mk_off.assign(n + 1, 0);
  std::vector<uint32_t> pt_cell;
  std::vector<int64_t> sc;
  std::vector<double> sw;
  for (int64_t m = 0; m < n; ++m) {
    marker_stencil(&pos[3 * m], dims, cfg.kernel_radius, sc, sw);
    if (sc.empty()) {
      *err = "IBM marker stencil is outside the domain";
      return nullptr;
    }
    for (size_t k = 0; k < sc.size(); ++k) {
      if (mask[sc[k]] & kMaskSolid) continue;
      pt_cell.push_back((uint32_t)sc[k]);
      h.pt_m.push_back((uint32_t)m);
      h.pt_w.push_back(sw[k]);
    }
    h.mk_off[m + 1] = (uint32_t)pt_cell.size();
  }
  const uint32_t Pn = (uint32_t)pt_cell.size();
  // unique cells, and per cell the points in marker order (stable sort)
  h.uc_ent.resize(Pn);
  for (uint32_t k = 0; k < Pn; ++k) h.uc_ent[k] = k;
  std::stable_sort(h.uc_ent.begin(), h.uc_ent.end(),
                   [&](uint32_t a, uint32_t b) { return pt_cell[a] < pt_cell[b]; });
  h.pt_u.assign(Pn, 0);
  for (uint32_t k = 0; k < Pn; ++k) {
    const uint32_t p = h.uc_ent[k];
    if (h.uc_cell.empty() || h.uc_cell.back() != pt_cell[p]) {
      h.uc_cell.push_back(pt_cell[p]);
      h.uc_off.push_back(k);
      h.ksum.push_back(0.0);
    }
    h.pt_u[p] = (uint32_t)h.uc_cell.size() - 1;
  }
  h.uc_off.push_back(Pn);
  // cell_kernel_sum in marker order (solver.rs:3107-3116)
  for (uint32_t p = 0; p < Pn; ++p) h.ksum[h.pt_u[p]] += h.pt_w[p];
  IbmState* s = ibm_build({&h}, cfg, err);
  if (!s) return nullptr;
  *cells = h.uc_cell;
  return s;
}

IbmState* ibm_concat(const std::vector<const IbmState*>& parts, std::string* err) {
  std::vector<const IbmHost*> hosts;
  for (const IbmState* p : parts) {
    if (!p || p->K != 1) {
      *err = "batched IBM needs one single-body state per instance";
      return nullptr;
    }
    const lbm_ibm_config& a = p->cfg;
    const lbm_ibm_config& b = parts[0]->cfg;
    if (a.max_iterations != b.max_iterations || a.kernel_radius != b.kernel_radius ||
        a.relaxation != b.relaxation || a.slip_tolerance != b.slip_tolerance) {
      *err = "batched IBM bodies must share one IBM configuration";
      return nullptr;
    }
    hosts.push_back(&p->host);
  }
  if (hosts.empty()) {
    *err = "empty IBM batch";
    return nullptr;
  }
  return ibm_build(hosts, parts[0]->cfg, err);
}

int ibm_instances(const IbmState* s) { return s ? s->K : 0; }

int ibm_set_domains(const FieldsView& v, IbmState* s, const IbmDomain* doms) {
  bool same = s->dom_valid;
  for (int k = 0; k < s->K && same; ++k) same = memcmp(&doms[k], &s->dom_h[k], sizeof(IbmDomain)) == 0;
  if (same) return LBM_OK;
  for (int k = 0; k < s->K; ++k) s->dom_h[k] = doms[k];
  AUXCHK(v, hipMemcpyAsync(s->dom, s->dom_h.data(), s->K * sizeof(IbmDomain), hipMemcpyHostToDevice, v.stream));
  AUXCHK(v, hipStreamSynchronize(v.stream));
  s->dom_valid = true;
  return LBM_OK;
}

IbmDomain ibm_domain_of(const FieldsView& v) {
  IbmDomain D{};
  D.d = v.d;
  D.f = v.f;
  D.mask = v.mask;
  D.F_cur = v.F_cur;
  for (int a = 0; a < 3; ++a) D.F[a] = v.hp.F[a];
  return D;
}

template <int Q, typename S>
static void gather_q(const FieldsView& v, IbmState* s, uint32_t reset_K, const uint32_t* clr, uint32_t nclr) {
  hipLaunchKernelGGL((k_ibm_gather<Q, S>), dim3(nblk(std::max(s->U, nclr))), dim3(256), 0, v.stream, s->dom, v.parity,
                     s->uc_cell, s->uc_inst, s->U, v.field_present ? 1 : 0, s->unow, s->rho, s->du, s->spread, s->diag,
                     reset_K, clr, nclr, v.F_cur);
}

static void diag_to_abi(const IbmDevDiag& h, lbm_ibm_diag* out) {
  out->iterations = h.iterations;
  out->slip_max = h.slip_max;
  out->slip_rms = h.slip_rms;
  out->slip_max_rel = h.slip_max_rel;
  out->slip_rms_rel = h.slip_rms_rel;
  for (int a = 0; a < 3; ++a) {
    out->torque[a] = h.torque[a];
    out->fluid_force[a] = h.fluid_force[a];
    out->marker_force[a] = h.marker_force[a];
  }
  out->momentum_error_rel = h.mom_err;
}

static int ibm_enqueue_impl(const FieldsView& v, IbmState* s, void* ring, uint32_t cap, uint32_t* pos,
                            bool record, const uint32_t* clr, uint32_t nclr);

int ibm_enqueue(const FieldsView& v, IbmState* s, void* ring, uint32_t cap, uint32_t* pos, const uint32_t* clear_list,
                uint32_t clear_n) {
  return ibm_enqueue_impl(v, s, ring, cap, pos, false, clear_list, clear_n);
}

__global__ void k_copy_d(double* dst, const double* src, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

// wide path: one kernel per phase across the whole GPU (12 launches per
// step at three sweeps: gather, impulse, three times spread, diagnostics and
// reduction, then commit with record), for bodies whose state does not fit
// in LDS
static int ibm_enqueue_wide(const FieldsView& v, IbmState* s, void* ring, uint32_t cap, uint32_t* pos, bool record,
                            const uint32_t* clr, uint32_t nclr) {
  if (!s->dom_valid) {
    if (v.ctx_err) *v.ctx_err = "IBM domain table not set";
    return LBM_ERR_STATE;
  }
  for (int k = 0; k < s->K; ++k)
    if (!s->dom_h[k].F_cur) {
      if (v.ctx_err) *v.ctx_err = "force arrays missing";
      return LBM_ERR_STATE;
    }
  hipStream_t st = v.stream;
  // the gather clears the diagnostics when it has a thread per instance
  const uint32_t reset_K = s->U >= (uint32_t)s->K ? (uint32_t)s->K : 0u;
  if (s->U || nclr) {
    if (v.Q == 19) {
      if (v.storage == kF32) gather_q<19, float>(v, s, reset_K, clr, nclr);
      else if (v.storage == kF16) gather_q<19, _Float16>(v, s, reset_K, clr, nclr);
      else gather_q<19, bf16s>(v, s, reset_K, clr, nclr);
    } else {
      if (v.storage == kF32) gather_q<27, float>(v, s, reset_K, clr, nclr);
      else if (v.storage == kF16) gather_q<27, _Float16>(v, s, reset_K, clr, nclr);
      else gather_q<27, bf16s>(v, s, reset_K, clr, nclr);
    }
    AUXCHK(v, hipGetLastError());
  }
  if (!s->U || !reset_K) {
    hipLaunchKernelGGL(k_ibm_reset, dim3(nblk(s->K, 64)), dim3(64), 0, st, s->diag, s->K);
    AUXCHK(v, hipGetLastError());
  }
  const uint32_t gm = nblk((size_t)s->M), gu = nblk(s->U);
  const uint32_t gm4 = nblk((size_t)s->M, 4), gu4 = nblk(s->U, 4);  // one wave per marker / cell
  for (int it = 0; it < s->cfg.max_iterations; ++it) {
    if (record && s->U)
      hipLaunchKernelGGL(k_copy_d, dim3(nblk(3 * (size_t)s->U)), dim3(256), 0, st,
                         s->du_hist + 3 * (size_t)s->U * it, s->du, 3 * (size_t)s->U);
    if (it == 0)  // later sweeps take the impulses the previous sweep's diagnostics computed
      hipLaunchKernelGGL(k_ibm_impulse, dim3(gm4), dim3(256), 0, st, s->M, s->mk_off, s->mk_inst, s->pt_u, s->pt_w,
                         s->target, s->unow, s->du, s->rho, s->ksum, s->cfg.relaxation, s->imp, s->diag);
    double* nxt = it + 1 < s->cfg.max_iterations ? s->imp_next : nullptr;
    if (record)
      hipLaunchKernelGGL(k_copy_d, dim3(nblk(3 * (size_t)s->M)), dim3(256), 0, st,
                         s->imp_hist + 3 * (size_t)s->M * it, s->imp, 3 * (size_t)s->M);
    if (s->U)
      hipLaunchKernelGGL(k_ibm_spread, dim3(gu4), dim3(256), 0, st, s->U, s->uc_off, s->uc_inst, s->uc_ent, s->pt_m,
                         s->pt_w, s->imp, s->rho, s->spread, s->du, s->diag);
    hipLaunchKernelGGL(k_ibm_marker_diag, dim3(gm4), dim3(256), 0, st, s->M, s->mk_off, s->mk_inst, s->pt_u, s->pt_w,
                       s->pos, s->target, s->unow, s->du, s->imp, s->cen, s->mrep, s->mtq, s->mslip, s->mact,
                       s->diag, s->rho, s->ksum, s->cfg.relaxation, nxt);
    hipLaunchKernelGGL(k_ibm_reduce, dim3(s->K), dim3(256), 0, st, s->moff, s->mrep, s->mtq, s->mslip, s->mact,
                       s->weight, it, s->scale, s->cfg.slip_tolerance, s->diag, nxt, s->imp);
    AUXCHK(v, hipGetLastError());
  }
  const uint32_t nb = s->U ? gu : 0u;
  hipLaunchKernelGGL(k_ibm_commit_record, dim3(nb + 1), dim3(256), 0, st, s->U, nb, s->uc_cell, s->uc_inst, s->dom,
                     s->spread, s->diag, s->K, (IbmDevDiag*)ring, cap, pos, s->integ);
  AUXCHK(v, hipGetLastError());
  return LBM_OK;
}

static int ibm_enqueue_impl(const FieldsView& v, IbmState* s, void* ring, uint32_t cap, uint32_t* pos,
                            bool record, const uint32_t* clr, uint32_t nclr) {
  if (!s->lds_bytes) return ibm_enqueue_wide(v, s, ring, cap, pos, record, clr, nclr);
  if (!s->dom_valid) {
    if (v.ctx_err) *v.ctx_err = "IBM domain table not set";
    return LBM_ERR_STATE;
  }
  for (int k = 0; k < s->K; ++k)
    if (!s->dom_h[k].F_cur) {
      if (v.ctx_err) *v.ctx_err = "force arrays missing";
      return LBM_ERR_STATE;
    }
  hipStream_t st = v.stream;
  const uint32_t reset_K = 0;  // the sweep kernel clears the diagnostics
  if (s->U || nclr) {
    if (v.Q == 19) {
      if (v.storage == kF32) gather_q<19, float>(v, s, reset_K, clr, nclr);
      else if (v.storage == kF16) gather_q<19, _Float16>(v, s, reset_K, clr, nclr);
      else gather_q<19, bf16s>(v, s, reset_K, clr, nclr);
    } else {
      if (v.storage == kF32) gather_q<27, float>(v, s, reset_K, clr, nclr);
      else if (v.storage == kF16) gather_q<27, _Float16>(v, s, reset_K, clr, nclr);
      else gather_q<27, bf16s>(v, s, reset_K, clr, nclr);
    }
    AUXCHK(v, hipGetLastError());
  }
  IbmSweepArgs a{};
  a.moff = s->moff;
  a.uoff = s->uoff;
  a.mk_off = s->mk_off;
  a.pt_u = s->pt_u;
  a.uc_off = s->uc_off;
  a.uc_m = s->uc_m;
  a.uc_cell = s->uc_cell;
  a.pt_w = s->pt_w;
  a.uc_w = s->uc_w;
  a.target = s->target;
  a.pos = s->pos;
  a.cen = s->cen;
  a.weight = s->weight;
  a.scale = s->scale;
  a.ksum = s->ksum;
  a.unow = s->unow;
  a.rho = s->rho;
  a.diag = s->diag;
  a.integ = s->integ;
  a.dom = s->dom;
  a.relax = s->cfg.relaxation;
  a.tol = s->cfg.slip_tolerance;
  a.max_iter = s->cfg.max_iterations;
  a.level = s->level;
  a.mmax = s->mmax;
  a.umax = s->umax;
  a.pmax = s->pmax;
  a.du_hist = record ? s->du_hist : nullptr;
  a.imp_hist = record ? s->imp_hist : nullptr;
  a.M = s->M;
  a.U = s->U;
  a.ring = (IbmDevDiag*)ring;
  a.cap = cap;
  a.ring_pos = pos;
  a.K = s->K;
  hipLaunchKernelGGL(k_ibm_sweeps, dim3(s->K), dim3(1024), s->lds_bytes, st, a);
  AUXCHK(v, hipGetLastError());
  return LBM_OK;
}

// ---------------------------------------------------------------------------
// Adjoint of one direct-forcing update (single body).  Forward, per sweep s
// (solver.rs:3107-3208): um = sum_p w (unow + du^s); mob = sum_p w ksum/rho;
// imp^s = relax (T - um)/mob; spread += imp^s w; du^{s+1} = du^s + imp^s w/rho;
// shaft torque += sum_m W_m (r x imp^s)_z, W_m = sum_p w.  Output F = spread.
// Reverse, for the recorded sweeps s = S-1..0, with Fbar = adjoint of F and
// tq_bar = adjoint of the shaft torque; then through unow = m/rho and
// rho = 1 + sum f, m = sum c f into the population adjoint.
// ---------------------------------------------------------------------------
__global__ void k_ibm_adj_zero(uint32_t U, int64_t M, double* a_du, double* a_unow, double* a_rho, double* a_T,
                               double* a_mob) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < (int64_t)U) {
    for (int a = 0; a < 3; ++a) a_du[3 * k + a] = a_unow[3 * k + a] = 0.0;
    a_rho[k] = 0.0;
  }
  if (k < M) {
    for (int a = 0; a < 3; ++a) a_T[3 * k + a] = 0.0;
    a_mob[k] = 0.0;
  }
}

// per marker, sweep s: impulse adjoint -> target, interpolated velocity and
// mobility adjoints
__global__ void k_ibm_adj_marker(int s, int64_t M, const uint32_t* mk_off, const uint32_t* pt_u, const double* pt_w,
                                 const uint32_t* uc_cell, const double* Fbar, const double* Gbar, const double* rho,
                                 const double* ksum, const double* a_du, const double* imp_hist, const double* pos,
                                 const double* cen, double relax, double tq_bar, const IbmDevDiag* diag, double* a_T,
                                 double* a_um, double* a_mob) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M || s >= diag->iterations) return;
  double ib[3] = {0.0, 0.0, 0.0}, W = 0.0, mob = 0.0;
  for (uint32_t p = mk_off[m]; p < mk_off[m + 1]; ++p) {
    const uint32_t u = pt_u[p];
    const double w = pt_w[p];
    const size_t i = uc_cell[u];
    const double ir = 1.0 / fmax(rho[u], 1.0e-30);
    for (int a = 0; a < 3; ++a) {
      double fb = Fbar[3 * i + a];
      if (Gbar) fb += Gbar[3 * i + a];
      ib[a] += w * (fb + a_du[3 * (size_t)u + a] * ir);
    }
    W += w;
    mob += w * ksum[u] / fmax(rho[u], 1.0e-30);
  }
  const double rx = pos[3 * m] - cen[0], ry = pos[3 * m + 1] - cen[1];
  ib[0] += tq_bar * W * -ry;
  ib[1] += tq_bar * W * rx;
  const double* imp = imp_hist + 3 * (size_t)M * s + 3 * m;
  if (mob == 0.0) {
    for (int a = 0; a < 3; ++a) a_um[3 * m + a] = 0.0;
    return;
  }
  double dot = 0.0;
  for (int a = 0; a < 3; ++a) {
    a_T[3 * m + a] += relax / mob * ib[a];
    a_um[3 * m + a] = -relax / mob * ib[a];
    dot += imp[a] * ib[a];
  }
  a_mob[m] += -dot / mob;
}

// per stencil cell, sweep s: rho adjoint of the du update, then the
// interpolated-velocity adjoint gathered in marker order
__global__ void k_ibm_adj_cell(int s, uint32_t U, int64_t M, const uint32_t* uc_off, const uint32_t* uc_ent,
                               const uint32_t* pt_m, const double* pt_w, const double* rho, const double* imp_hist,
                               const double* a_um, const IbmDevDiag* diag, double* a_du, double* a_unow,
                               double* a_rho) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U || s >= diag->iterations) return;
  const double r = fmax(rho[u], 1.0e-30);
  double sw[3] = {0.0, 0.0, 0.0}, g[3] = {0.0, 0.0, 0.0};
  for (uint32_t e = uc_off[u]; e < uc_off[u + 1]; ++e) {
    const uint32_t p = uc_ent[e];
    const uint32_t m = pt_m[p];
    const double w = pt_w[p];
    for (int a = 0; a < 3; ++a) {
      sw[a] += imp_hist[3 * (size_t)M * s + 3 * (size_t)m + a] * w;
      g[a] += w * a_um[3 * (size_t)m + a];
    }
  }
  double rb = 0.0;
  for (int a = 0; a < 3; ++a) rb += a_du[3 * (size_t)u + a] * sw[a];
  a_rho[u] += -rb / (r * r);
  for (int a = 0; a < 3; ++a) {
    a_unow[3 * (size_t)u + a] += g[a];
    a_du[3 * (size_t)u + a] += g[a];
  }
}

// per stencil cell, after the sweeps: mobility adjoint into rho, then
// unow = m/rho and the moments into the population adjoint (q-major f64)
template <int Q>
__global__ void k_ibm_adj_final(uint32_t U, uint32_t N, const uint32_t* uc_cell, const uint32_t* uc_off,
                                const uint32_t* uc_ent, const uint32_t* pt_m, const double* pt_w, const double* rho,
                                const double* unow, const double* ksum, const double* a_mob, const double* a_unow,
                                const double* a_rho, double* lam) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  const double r = fmax(rho[u], 1.0e-30);
  double rb = a_rho[u];
  for (uint32_t e = uc_off[u]; e < uc_off[u + 1]; ++e) {
    const uint32_t p = uc_ent[e];
    rb += -pt_w[p] * ksum[u] * a_mob[pt_m[p]] / (r * r);
  }
  double mb[3];
  double ud = 0.0;
  for (int a = 0; a < 3; ++a) {
    ud += unow[3 * (size_t)u + a] * a_unow[3 * (size_t)u + a];
    mb[a] = a_unow[3 * (size_t)u + a] / r;
  }
  rb += -ud / r;
  const size_t i = uc_cell[u];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    double v = rb;
    if constexpr (lat_c(q, 0) != 0) v += (double)lat_c(q, 0) * mb[0];
    if constexpr (lat_c(q, 1) != 0) v += (double)lat_c(q, 1) * mb[1];
    if constexpr (lat_c(q, 2) != 0) v += (double)lat_c(q, 2) * mb[2];
    lam[(size_t)q * N + i] += v;
  });
}

// one block: rotation-rate adjoint, T = Omega x r with Omega = (0, 0, W):
// dT/dW = (-r_y, r_x, 0); fixed-tree sum over markers, added to *out
__global__ __launch_bounds__(256) void k_ibm_adj_omega(int64_t M, const double* pos, const double* cen,
                                                       const double* a_T, double* out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t m = threadIdx.x; m < M; m += 256)
    acc += -(pos[3 * m + 1] - cen[1]) * a_T[3 * m] + (pos[3 * m] - cen[0]) * a_T[3 * m + 1];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out += sh[0];
}

int ibm_adjoint_prepare(const FieldsView& v, IbmState* s) {
  if (s->K != 1) {
    if (v.ctx_err) *v.ctx_err = "the IBM adjoint takes a single body";
    return LBM_ERR_UNSUPPORTED;
  }
  if (s->imp_hist) return LBM_OK;
  const size_t it = (size_t)s->cfg.max_iterations;
  AUXCHK(v, scratch(&s->imp_hist, it * 3 * (size_t)s->M));
  AUXCHK(v, scratch(&s->du_hist, it * 3 * (size_t)s->U));
  AUXCHK(v, scratch(&s->a_T, 3 * (size_t)s->M));
  AUXCHK(v, scratch(&s->a_um, 3 * (size_t)s->M));
  AUXCHK(v, scratch(&s->a_mob, s->M));
  AUXCHK(v, scratch(&s->a_du, 3 * (size_t)s->U));
  AUXCHK(v, scratch(&s->a_unow, 3 * (size_t)s->U));
  AUXCHK(v, scratch(&s->a_rho, s->U));
  return LBM_OK;
}

int ibm_enqueue_recorded(const FieldsView& v, IbmState* s) {
  int rc = ibm_adjoint_prepare(v, s);
  if (rc) return rc;
  return ibm_enqueue_impl(v, s, nullptr, 0, nullptr, true, nullptr, 0);
}

int ibm_adjoint(const FieldsView& v, IbmState* s, const double* Fbar, const double* Gbar, double tq_bar, double* lam,
                double* omega_bar) {
  int rc = ibm_adjoint_prepare(v, s);
  if (rc) return rc;
  hipStream_t st = v.stream;
  const uint32_t gm = nblk((size_t)s->M), gu = nblk(s->U), gz = nblk(std::max<size_t>(s->U, (size_t)s->M));
  hipLaunchKernelGGL(k_ibm_adj_zero, dim3(gz), dim3(256), 0, st, s->U, s->M, s->a_du, s->a_unow, s->a_rho, s->a_T,
                     s->a_mob);
  for (int it = s->cfg.max_iterations - 1; it >= 0; --it) {
    hipLaunchKernelGGL(k_ibm_adj_marker, dim3(gm), dim3(256), 0, st, it, s->M, s->mk_off, s->pt_u, s->pt_w,
                       s->uc_cell, Fbar, Gbar, s->rho, s->ksum, s->a_du, s->imp_hist, s->pos, s->cen,
                       s->cfg.relaxation, tq_bar, s->diag, s->a_T, s->a_um, s->a_mob);
    if (s->U)
      hipLaunchKernelGGL(k_ibm_adj_cell, dim3(gu), dim3(256), 0, st, it, s->U, s->M, s->uc_off, s->uc_ent, s->pt_m,
                         s->pt_w, s->rho, s->imp_hist, s->a_um, s->diag, s->a_du, s->a_unow, s->a_rho);
  }
  if (s->U) {
    if (v.Q == 19)
      hipLaunchKernelGGL((k_ibm_adj_final<19>), dim3(gu), dim3(256), 0, st, s->U, v.d.N, s->uc_cell, s->uc_off,
                         s->uc_ent, s->pt_m, s->pt_w, s->rho, s->unow, s->ksum, s->a_mob, s->a_unow, s->a_rho, lam);
    else
      hipLaunchKernelGGL((k_ibm_adj_final<27>), dim3(gu), dim3(256), 0, st, s->U, v.d.N, s->uc_cell, s->uc_off,
                         s->uc_ent, s->pt_m, s->pt_w, s->rho, s->unow, s->ksum, s->a_mob, s->a_unow, s->a_rho, lam);
  }
  hipLaunchKernelGGL(k_ibm_adj_omega, dim3(1), dim3(256), 0, st, s->M, s->pos, s->cen, s->a_T, omega_bar);
  AUXCHK(v, hipGetLastError());
  return LBM_OK;
}

int ibm_apply(const FieldsView& v, IbmState* s, lbm_ibm_diag* out) {
  int rc = ibm_enqueue(v, s, nullptr, 0, nullptr);
  if (rc) return rc;
  std::vector<IbmDevDiag> h(s->K);
  AUXCHK(v, hipMemcpyAsync(h.data(), s->diag, s->K * sizeof(IbmDevDiag), hipMemcpyDeviceToHost, v.stream));
  AUXCHK(v, hipStreamSynchronize(v.stream));
  for (int k = 0; k < s->K; ++k) diag_to_abi(h[k], out + k);
  return LBM_OK;
}

int ibm_torque_integral(const FieldsView& v, IbmState* s, int reset, double* out) {
  AUXCHK(v, hipMemcpyAsync(out, s->integ, 4 * (size_t)s->K * sizeof(double), hipMemcpyDeviceToHost, v.stream));
  AUXCHK(v, hipStreamSynchronize(v.stream));
  if (reset) {
    hipLaunchKernelGGL(k_zero_integ, dim3(1), dim3(64), 0, v.stream, s->integ, 4 * s->K);
    AUXCHK(v, hipGetLastError());
  }
  return LBM_OK;
}

size_t ibm_ring_entry_bytes() { return sizeof(IbmDevDiag); }

int ibm_ring_read(const FieldsView& v, const void* ring, uint32_t n, lbm_ibm_diag* out) {
  std::vector<IbmDevDiag> h(n);
  if (n) {
    AUXCHK(v, hipMemcpyAsync(h.data(), ring, n * sizeof(IbmDevDiag), hipMemcpyDeviceToHost, v.stream));
    AUXCHK(v, hipStreamSynchronize(v.stream));
  }
  for (uint32_t k = 0; k < n; ++k) diag_to_abi(h[k], out + k);
  return LBM_OK;
}

// ===========================================================================
// passive scalar ADE
// ===========================================================================
struct ScalarState {
  int Qs = 7;
  double D = 0.0;
  float omega = 1.0f;
  float* h = nullptr;
  float* h2 = nullptr;
  float* C = nullptr;
  float* mom = nullptr;  // rho, ux, uy, uz (4 x N)
  double* partials = nullptr;
  uint32_t npartials = 1024;
};

// h_eq of scalar_equilibrium (scalar.rs:6-15) for the hydro lattice; the
// D3Q7 form uses 1/cs^2 = 4.
template <int Qs, int q>
__device__ __forceinline__ float heq_q(float conc, float ux, float uy, float uz) {
  const float cu = cdot<q>(ux, uy, uz);
  constexpr float k = Qs == 7 ? 4.0f : 3.0f;
  return lat_wf(Qs, q) * conc * (1.0f + k * cu);
}

template <int Qs>
__global__ void k_scalar_set(Dom d, const float* conc, const float* mom, float* h, float* C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  const float c = conc[i];
  const float ux = mom[d.N + i], uy = mom[2 * (size_t)d.N + i], uz = mom[3 * (size_t)d.N + i];
  C[i] = c;
  sfor<0, Qs>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    h[(size_t)q * d.N + i] = heq_q<Qs, q>(c, ux, uy, uz);
  });
}

template <int Qs>
__global__ void k_scalar_collide(Dom d, const uint32_t* mask, float* h, const float* C, const float* mom,
                                 float omega) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N || (mask[i] & kMaskSolid)) return;
  const float c = C[i];
  const float ux = mom[d.N + i], uy = mom[2 * (size_t)d.N + i], uz = mom[3 * (size_t)d.N + i];
  sfor<0, Qs>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    float* p = h + (size_t)q * d.N + i;
    const float hq = *p;
    *p = hq - omega * (hq - heq_q<Qs, q>(c, ux, uy, uz));
  });
}

template <int Qs>
__global__ void k_scalar_stream(Dom d, const uint32_t* mask, const float* h, float* h2, float* C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  const uint32_t m = mask[i];
  if (m & kMaskSolid) {
    C[i] = 0.0f;
    sfor<0, Qs>([&](auto qc) { h2[(size_t)decltype(qc)::value * d.N + i] = 0.0f; });
    return;
  }
  uint32_t x, y, z;
  cell_xyz(d, i, x, y, z);
  const Nbr nb(d, x, y, z);
  float sum = 0.0f;
  sfor<0, Qs>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = lat_opp(q);
    float v;
    if ((m >> q) & 1u) {  // source i - c_q is solid: bounce back own h_opp
      v = h[(size_t)o * d.N + i];
    } else {
      const uint32_t s = i + (uint32_t)nb.template off<-lat_c(q, 0), -lat_c(q, 1), -lat_c(q, 2)>();
      v = h[(size_t)q * d.N + s];
    }
    h2[(size_t)q * d.N + i] = v;
    sum = sum + v;
  });
  C[i] = sum;
}

// Two-pass CV over fluid cells (qoi::scalar_cv, qoi.rs:507-532), for the
// whole fluid (parts = 1) or the three compartments of compartment_cv
// (qoi.rs:699-739, parts = 3): a fluid cell of plane z is "top" when
// (z + 0.5)/nz >= 2/3, else "near_impeller" when an impeller height is given
// and |(z + 0.5) - z_imp| <= nz/6, else "bulk" (f64, the reference's
// expressions).  Pass 0 sums counts and values per compartment, pass 1 the
// squared deviations from the compartment means.  Per block: 6 doubles
// (count, sum per compartment; pass 1 fills the first 3 with the variances).
__global__ __launch_bounds__(256) void k_cv_pass(Dom d, const uint32_t* mask, const float* C, int pass, int parts,
                                                 int has_imp, double imp_z, double m0, double m1, double m2,
                                                 double* partials) {
  __shared__ double lds[4 * 6];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double nz = (double)d.nz;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < d.N; i += gridDim.x * 256) {
    if (mask[i] & kMaskSolid) continue;
    int k = 0;
    if (parts == 3) {
      const double zc = (double)(i / (d.nx * d.ny)) + 0.5;
      if (zc / nz >= 2.0 / 3.0)
        k = 0;
      else if (has_imp && fabs(zc - imp_z) <= nz / 6.0)
        k = 2;
      else
        k = 1;
    }
    const double v = (double)C[i];
    if (pass == 0) {
      acc[2 * k] += 1.0;
      acc[2 * k + 1] += v;
    } else {
      const double dv = v - (k == 0 ? m0 : (k == 1 ? m1 : m2));
      acc[k] += dv * dv;
    }
  }
  block_sum256<6>(acc, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 6; ++k) partials[6 * blockIdx.x + k] = acc[k];
}

ScalarState* scalar_create(const FieldsView& v, double diffusivity, int Qs, std::string* err) {
  ScalarState* s = new ScalarState();
  s->Qs = Qs;
  s->D = diffusivity;
  // omega = T(1/(3D + 1/2)) on the hydro lattice (solver.rs:2131);
  // D3Q7 maps tau = D/cs^2 + 1/2 = 4D + 1/2.
  s->omega = Qs == 7 ? (float)(1.0 / (4.0 * diffusivity + 0.5)) : (float)(1.0 / (3.0 * diffusivity + 0.5));
  const size_t N = v.d.N;
  hipError_t e = hipMalloc((void**)&s->h, (size_t)Qs * N * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&s->h2, (size_t)Qs * N * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&s->C, N * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&s->mom, 4 * N * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&s->partials, 6 * s->npartials * sizeof(double));
  if (e == hipSuccess) e = hipMemsetAsync(s->h, 0, (size_t)Qs * N * 4, v.stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->h2, 0, (size_t)Qs * N * 4, v.stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->C, 0, N * 4, v.stream);
  if (e != hipSuccess) {
    *err = std::string("scalar allocation failed: ") + hipGetErrorString(e);
    scalar_destroy(s);
    return nullptr;
  }
  return s;
}

void scalar_destroy(ScalarState* s) {
  if (!s) return;
  for (void* p : {(void*)s->h, (void*)s->h2, (void*)s->C, (void*)s->mom, (void*)s->partials})
    if (p) (void)hipFree(p);
  delete s;
}

static int refresh_mom(const FieldsView& v, ScalarState* s) {
  const size_t N = v.d.N;
  AUXCHK(v, launch_moments(v.Q, v.storage, v.parity, v.d, v.f, v.mask, v.F_mom, v.hp, s->mom, s->mom + N,
                           s->mom + 2 * N, s->mom + 3 * N, v.stream));
  return LBM_OK;
}

int scalar_set(const FieldsView& v, ScalarState* s, const float* conc) {
  const size_t N = v.d.N;
  float* dc = nullptr;
  AUXCHK(v, hipMalloc((void**)&dc, N * 4));
  AUXCHK(v, hipMemcpyAsync(dc, conc, N * 4, hipMemcpyHostToDevice, v.stream));
  int rc = refresh_mom(v, s);
  if (rc) return rc;
  if (s->Qs == 7)
    hipLaunchKernelGGL((k_scalar_set<7>), dim3(nblk(N)), dim3(256), 0, v.stream, v.d, dc, s->mom, s->h, s->C);
  else if (s->Qs == 19)
    hipLaunchKernelGGL((k_scalar_set<19>), dim3(nblk(N)), dim3(256), 0, v.stream, v.d, dc, s->mom, s->h, s->C);
  else
    hipLaunchKernelGGL((k_scalar_set<27>), dim3(nblk(N)), dim3(256), 0, v.stream, v.d, dc, s->mom, s->h, s->C);
  AUXCHK(v, hipGetLastError());
  AUXCHK(v, hipStreamSynchronize(v.stream));
  (void)hipFree(dc);
  return LBM_OK;
}

// checkpoint access: the distribution h (current), its ping-pong partner h2
// (after a step: the post-collision populations), the concentration, and
// the lattice size and diffusivity
int scalar_info(const ScalarState* s, int* Qs, double* D) {
  *Qs = s->Qs;
  *D = s->D;
  return LBM_OK;
}

int scalar_get_state(const FieldsView& v, ScalarState* s, float* h, float* h2, float* C) {
  const size_t N = v.d.N, bytes = (size_t)s->Qs * N * 4;
  AUXCHK(v, hipMemcpyAsync(h, s->h, bytes, hipMemcpyDeviceToHost, v.stream));
  AUXCHK(v, hipMemcpyAsync(h2, s->h2, bytes, hipMemcpyDeviceToHost, v.stream));
  AUXCHK(v, hipMemcpyAsync(C, s->C, N * 4, hipMemcpyDeviceToHost, v.stream));
  AUXCHK(v, hipStreamSynchronize(v.stream));
  return LBM_OK;
}

int scalar_put_state(const FieldsView& v, ScalarState* s, const float* h, const float* h2, const float* C) {
  const size_t N = v.d.N, bytes = (size_t)s->Qs * N * 4;
  AUXCHK(v, hipMemcpyAsync(s->h, h, bytes, hipMemcpyHostToDevice, v.stream));
  AUXCHK(v, hipMemcpyAsync(s->h2, h2, bytes, hipMemcpyHostToDevice, v.stream));
  AUXCHK(v, hipMemcpyAsync(s->C, C, N * 4, hipMemcpyHostToDevice, v.stream));
  AUXCHK(v, hipStreamSynchronize(v.stream));
  return LBM_OK;
}

int scalar_gather(const FieldsView& v, ScalarState* s, float* conc) {
  AUXCHK(v, hipMemcpyAsync(conc, s->C, (size_t)v.d.N * 4, hipMemcpyDeviceToHost, v.stream));
  AUXCHK(v, hipStreamSynchronize(v.stream));
  return LBM_OK;
}

template <int Qs>
static void scalar_step_q(const FieldsView& v, ScalarState* s) {
  const uint32_t g = nblk(v.d.N);
  hipLaunchKernelGGL((k_scalar_collide<Qs>), dim3(g), dim3(256), 0, v.stream, v.d, v.mask, s->h, s->C, s->mom,
                     s->omega);
  hipLaunchKernelGGL((k_scalar_stream<Qs>), dim3(g), dim3(256), 0, v.stream, v.d, v.mask, s->h, s->h2, s->C);
}

int scalar_step(const FieldsView& v, ScalarState* s) {
  int rc = refresh_mom(v, s);
  if (rc) return rc;
  if (s->Qs == 7)
    scalar_step_q<7>(v, s);
  else if (s->Qs == 19)
    scalar_step_q<19>(v, s);
  else
    scalar_step_q<27>(v, s);
  AUXCHK(v, hipGetLastError());
  std::swap(s->h, s->h2);
  return LBM_OK;
}

// Both passes of k_cv_pass; results per compartment (parts = 1 or 3).
static int cv_parts(const FieldsView& v, ScalarState* s, int parts, int has_imp, double imp_z, double* cv,
                    int* defined, int64_t* counts) {
  const uint32_t np = std::min<uint32_t>(s->npartials, nblk(v.d.N));
  std::vector<double> h(6 * (size_t)np);
  auto run = [&](int pass, const double* mean) -> int {
    hipLaunchKernelGGL(k_cv_pass, dim3(np), dim3(256), 0, v.stream, v.d, v.mask, s->C, pass, parts, has_imp, imp_z,
                       mean[0], mean[1], mean[2], s->partials);
    AUXCHK(v, hipGetLastError());
    AUXCHK(v, hipMemcpyAsync(h.data(), s->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, v.stream));
    AUXCHK(v, hipStreamSynchronize(v.stream));
    return LBM_OK;
  };
  const double zero[3] = {0.0, 0.0, 0.0};
  int rc = run(0, zero);
  if (rc) return rc;
  double n[3] = {0, 0, 0}, sum[3] = {0, 0, 0}, mean[3] = {0, 0, 0};
  for (uint32_t b = 0; b < np; ++b)
    for (int k = 0; k < parts; ++k) {
      n[k] += h[6 * (size_t)b + 2 * k];
      sum[k] += h[6 * (size_t)b + 2 * k + 1];
    }
  bool any = false;
  for (int k = 0; k < parts; ++k) {
    counts[k] = (int64_t)n[k];
    defined[k] = 0;
    cv[k] = 0.0;
    if (n[k] == 0.0) continue;
    mean[k] = sum[k] / n[k];
    if (mean[k] == 0.0) continue;
    defined[k] = 1;
    any = true;
  }
  if (!any) return LBM_OK;
  rc = run(1, mean);
  if (rc) return rc;
  double var[3] = {0, 0, 0};
  for (uint32_t b = 0; b < np; ++b)
    for (int k = 0; k < parts; ++k) var[k] += h[6 * (size_t)b + k];
  for (int k = 0; k < parts; ++k)
    if (defined[k]) cv[k] = std::sqrt(var[k] / n[k]) / std::fabs(mean[k]);
  return LBM_OK;
}

int scalar_cv(const FieldsView& v, ScalarState* s, double* cv, int* defined) {
  int64_t count = 0;
  return cv_parts(v, s, 1, 0, 0.0, cv, defined, &count);
}

int scalar_compartment_cv(const FieldsView& v, ScalarState* s, int has_imp, double imp_z, double cv[3],
                          int defined[3], int64_t counts[3]) {
  return cv_parts(v, s, 3, has_imp, imp_z, cv, defined, counts);
}

}  // namespace lbm
