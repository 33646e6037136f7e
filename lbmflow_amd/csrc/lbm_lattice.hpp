// lbm_lattice.hpp — compile-time lattice tables for device and host code.
//
// Velocity order and weights follow lbm-core's lattice.rs (D3Q19 at
// lattice.rs:295-337, D3Q27 at lattice.rs:384-442: the first 19 directions
// are D3Q19, the 8 corners are appended as opposite pairs, so OPP[q] = q±1).
// D3Q7 (rest + 6 axis links, w0 = 1/4, w1..6 = 1/8, cs^2 = 1/4) is the
// scalar lattice of the WSCAL spec
// (docs/archive/2026-07-07-pivot/proposals/WSCAL_MULTICOMPONENT_SPEC.md:68-71).
// The tables are the reference's by necessity (the direction order is the
// storage contract); everything is constexpr so unrolled kernels see
// literal constants.
#pragma once
#include <cstdint>

#ifndef __HIPCC__
#define LBM_HD
#else
#define LBM_HD __host__ __device__
#endif

namespace lbm {

// c_q for the 27-velocity set; D3Q19 is the prefix 0..18, D3Q7 the prefix 0..6.
LBM_HD constexpr int lat_c(int q, int a) {
  constexpr int8_t C[27][3] = {
      {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
      {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},   {0, -1, -1}, {1, -1, 0},
      {-1, 1, 0}, {1, 0, -1},  {-1, 0, 1}, {0, 1, -1},  {0, -1, 1},  {1, 1, 1},  {-1, -1, -1},
      {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
  return C[q][a];
}

// OPP[q] = q±1 for every moving direction (lattice.rs:291, 377).
LBM_HD constexpr int lat_opp(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }

// Weights as f64 (lattice.rs:317-337, 414-442; D3Q7 per the WSCAL spec).
LBM_HD constexpr double lat_w(int Q, int q) {
  if (Q == 19) return q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  if (Q == 27) return q == 0 ? 8.0 / 27.0 : (q < 7 ? 2.0 / 27.0 : (q < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
  return q == 0 ? 1.0 / 4.0 : 1.0 / 8.0;  // Q == 7
}

// Weight rounded to fp32 exactly as T::r(W[q]) (params.rs:357).
LBM_HD constexpr float lat_wf(int Q, int q) { return (float)lat_w(Q, q); }

// Storage scale for 16-bit populations: s_q = 1/w_q (exact in fp32 for all
// three lattices).  FP32 storage uses s_q = 1.
LBM_HD constexpr float lat_inv_w(int Q, int q) {
  if (Q == 19) return q == 0 ? 3.0f : (q < 7 ? 18.0f : 36.0f);
  if (Q == 27) return q == 0 ? 3.375f : (q < 7 ? 13.5f : (q < 19 ? 54.0f : 216.0f));
  return q == 0 ? 4.0f : 8.0f;
}

}  // namespace lbm
