// lbm_probe.hip — momentum-exchange force on probed solids (the force probe
// of stream_row, kernels.rs:572-615, with the row-ordered fold of
// backend.rs:521-544).
//
// The probe is a side computation of the step: for every half-way
// bounce-back link (i, q) whose wall cell i - c_q is probed, the wall takes
// -c_q (f_out + f_in + 2 w_q) in physical populations, f_out the fluid
// cell's post-collision population towards the wall and f_in the bounced one.
// Three small launches per step, over the link list only:
//   k_probe_rho  (before the step) the density of each probed fluid cell,
//                the rho_row the reference's stream uses;
//   k_probe_rows (after the step, before the open faces) one thread per row
//                holding links: f_in is the bounced population where the new
//                state keeps it, f_out = f_in - 6 w_q rho (c_q . u_wall);
//                the row partial sums links in (x, q) order;
//   k_probe_fold one thread adds the row partials in row order.
// With walls at rest f_out = f_in and the result is the reference's to the
// bit; for a moving probed wall f_out is recovered by one subtraction.
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

#pragma clang fp contract(off)

namespace lbm {

template <int Q, typename S>
__global__ __launch_bounds__(256) void k_probe_rho(ProbeArgs a) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.ncells) return;
  const uint32_t i = a.cells[k];
  const uint32_t m = a.mask[i];
  uint32_t x, y, z;
  cell_xyz(a.d, i, x, y, z);
  const Nbr nb(a.d, x, y, z);
  float fv[Q];
  sfor<0, Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    fv[q] = read_pop<Q, S, q>((const S*)a.f, a.d.ps, nb, i, m, a.odd != 0);
  });
  const float F0[3] = {0.0f, 0.0f, 0.0f};
  float r, u0, u1, u2;
  moments<Q>(fv, F0, false, r, u0, u1, u2);  // rho = 1 + sum f in q order
  a.rho[k] = r;
}

template <int Q, typename S>
__global__ __launch_bounds__(256) void k_probe_rows(ProbeArgs a) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nrows) return;
  float pf[3] = {0.0f, 0.0f, 0.0f};
  const float six = 6.0f, two = 2.0f;
  for (uint32_t l = a.row_start[r]; l < a.row_start[r + 1]; ++l) {
    const uint32_t k = a.lcell[l];
    const int ql = a.lq[l];
    const uint32_t i = a.cells[k];
    const uint32_t m = a.mask[i];
    uint32_t x, y, z;
    cell_xyz(a.d, i, x, y, z);
    const Nbr nb(a.d, x, y, z);
    const float rho = a.rho[k];
    sfor<1, Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if (q != ql) return;
      const float fin = read_pop<Q, S, q>((const S*)a.f, a.d.ps, nb, i, m, a.odd != 0);
      float cu = 0.0f;
      if (a.wall_u) {
        const uint32_t s = i + (uint32_t)nb.template off<-lat_c(q, 0), -lat_c(q, 1), -lat_c(q, 2)>();
        const float* wu = a.wall_u + 3 * (size_t)s;
        // the reference's seeded dot product, every term kept
        cu = (float)lat_c(q, 0) * wu[0];
        cu = cu + (float)lat_c(q, 1) * wu[1];
        cu = cu + (float)lat_c(q, 2) * wu[2];
      }
      constexpr float w = lat_wf(Q, q);
      const float fout = cu == 0.0f ? fin : fin - six * w * rho * cu;
      const float ftot = fout + fin + two * w;
      pf[0] = pf[0] - (float)lat_c(q, 0) * ftot;
      pf[1] = pf[1] - (float)lat_c(q, 1) * ftot;
      pf[2] = pf[2] - (float)lat_c(q, 2) * ftot;
    });
  }
  a.partial[3 * (size_t)r + 0] = pf[0];
  a.partial[3 * (size_t)r + 1] = pf[1];
  a.partial[3 * (size_t)r + 2] = pf[2];
}

__global__ void k_probe_fold(ProbeArgs a) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float t[3] = {0.0f, 0.0f, 0.0f};
  for (uint32_t r = 0; r < a.nrows; ++r)
    for (int c = 0; c < 3; ++c) t[c] = t[c] + a.partial[3 * (size_t)r + c];
  for (int c = 0; c < 3; ++c) a.out[c] = t[c];
}

template <typename S>
static hipError_t probe_s(int Q, int post, const ProbeArgs& a, hipStream_t s) {
  if (!post) {
    if (a.ncells == 0) return hipSuccess;
    const dim3 g((a.ncells + 255) / 256);
    if (Q == 19)
      hipLaunchKernelGGL((k_probe_rho<19, S>), g, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_probe_rho<27, S>), g, dim3(256), 0, s, a);
    return hipGetLastError();
  }
  if (a.nrows) {
    const dim3 g((a.nrows + 255) / 256);
    if (Q == 19)
      hipLaunchKernelGGL((k_probe_rows<19, S>), g, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_probe_rows<27, S>), g, dim3(256), 0, s, a);
  }
  hipLaunchKernelGGL(k_probe_fold, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_probe(int Q, int storage, int post, const ProbeArgs& a, hipStream_t s) {
  if (storage == kF32) return probe_s<float>(Q, post, a, s);
  if (storage == kF16) return probe_s<_Float16>(Q, post, a, s);
  return probe_s<bf16s>(Q, post, a, s);
}

}  // namespace lbm
