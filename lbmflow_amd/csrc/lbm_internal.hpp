// lbm_internal.hpp — structures shared by the C-ABI layer (lbm_api.cpp) and
// the HIP kernels (lbm_kernels.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace lbm {

// Per-cell mask word.  Bits 0..26: half-way bounce-back links in the pull
// convention of stream_row (kernels.rs:572-603): bit q is set on a fluid
// cell i when i - c_q is solid.  High bits are cell flags.
constexpr uint32_t kMaskSolid = 1u << 31;
constexpr uint32_t kMaskForced = 1u << 30;  // handled by the force-list kernel
constexpr uint32_t kMaskMoving = 1u << 29;  // some linked solid has a nonzero wall velocity
constexpr uint32_t kMaskLinks = (1u << 27) - 1u;
// tile-list entry flags: every cell of the 64-cell tile has mask word 0;
// lanes 0-31 / 32-63 are all solid (their slots are dead: skip them)
constexpr uint32_t kTilePlain = 1u << 31;
constexpr uint32_t kTileDeadLo = 1u << 30;
constexpr uint32_t kTileDeadHi = 1u << 29;
constexpr uint32_t kTileId = kTileDeadHi - 1u;

enum Storage : int { kF32 = 0, kF16 = 1, kBF16 = 2 };

struct Dom {
  uint32_t nx, ny, nz;
  uint32_t N;        // nx*ny*nz cells
  uint32_t ps;       // population plane stride in elements (>= N)
  uint32_t per[3];   // periodic axes
};

// Launch geometry of the cell-parallel kernels: blocks of TX x TY threads
// tile the (x, y) plane; block ids are remapped so each XCD walks a
// contiguous slab of the domain.
struct Tiling {
  uint32_t tx, ty;
  uint32_t bx, by;   // blocks along x and y
  uint32_t nblocks;  // bx * by * nz
};

struct HydroParams {
  float omega_p, omega_m, cp, cm, omega_shear;
  float F[3];        // uniform Guo force (params.force)
  int force_on;      // F != 0
};

struct HydroArgs {
  Dom d;
  Tiling t;
  // active wave tiles: 64-cell x-segments holding at least one cell of the
  // main kernel; tile id = (z*ny + y)*nseg + segment
  const uint32_t* tiles;
  uint32_t ntiles;
  uint32_t nseg;
  void* f;                  // Q planes of N storage elements
  const uint32_t* mask;     // N
  const float* wall_u;      // N x 3 or null
  const float* omega;       // N or null (per-cell omega_plus)
  HydroParams p;
  // force-list kernel only
  const uint32_t* flist;    // listed cells
  uint32_t nlist;
  float* F_mom;             // N x 3: force of the cached moments
  const float* F_cur;       // N x 3: current field
  int open;                 // fluid on a non-periodic (open) face: wrap every offset
  int lin;                  // tiles are 64-cell blocks of the linear index (no axis wraps)
};

// Open faces (lbm_open.hip): FaceBC kinds of params.rs:232-252, faces in
// Face::ALL order (XNeg, XPos, YNeg, YPos, ZNeg, ZPos).
constexpr int kFaceClosed = 0, kFaceVelocity = 1, kFacePressure = 2, kFaceOutflow = 3, kFaceConvective = 4;
// MAX_SPEED (params.rs:11): the low-Mach limit on prescribed speeds, lattice units
constexpr double kMaxSpeed = 0.3;
// One face patch: inclusive rect lo..hi over the face's tangents (ascending
// axis order) with its own rule (a closed patch arrives as a zero velocity).
constexpr int kMaxFacePatches = 16;
struct FacePatchArgs {
  int face;
  uint32_t lo[2], hi[2];
  int kind;
  float u[3];
  float rho;
  float uconv;
};

struct FaceArgs {
  Dom d;
  void* f;
  const uint32_t* mask;
  int odd;              // parity of the state the step just produced
  HydroParams p;        // uniform force
  const float* F_cur;   // N x 3 or null (current field, nonzero only on the force list)
  int nfaces;           // open faces, in Face::ALL order
  int faces[6];
  int kind[6];
  float u[6][3];
  float rho[6];
  float uconv[6];
  float* stash_cur[6];   // per face: unknowns x face cells, previous step's post-collide values
  float* stash_next[6];  // this step's, captured before the restore
  int npatch;            // rect overrides, at most one per face cell
  FacePatchArgs patch[kMaxFacePatches];
  const float* prof[6];  // per face: u[3] per face cell (inlet profile) or null
};

// One volume source (lbm_open.hip): inclusive box, q_lu per step, velocity u
// (zero for a mass flow).
struct SourceArgs {
  Dom d;
  void* f;
  const uint32_t* mask;
  int odd;
  uint32_t lo[3], hi[3];
  float q_lu;
  float u[3];
};
hipError_t launch_volume_source(int Q, int storage, const SourceArgs& a, hipStream_t s);

// Force probe (lbm_probe.hip): links (i, q) with i - c_q a probed solid,
// grouped by row in (row, x, q) order; cells = the distinct fluid cells.
struct ProbeArgs {
  Dom d;
  const void* f;
  const uint32_t* mask;
  int odd;
  const float* wall_u;  // N x 3 or null
  const uint32_t* cells;
  uint32_t ncells;
  float* rho;  // ncells: pre-step density
  const uint32_t* lcell;
  const uint8_t* lq;
  const uint32_t* row_start;  // nrows + 1
  uint32_t nrows;
  float* partial;  // 3 x nrows
  float* out;      // 3: the step's probed force
};

// Host-side launchers (lbm_kernels.hip).  All enqueue on `s`.
hipError_t launch_hydro(int Q, int storage, int collision, int odd, const HydroArgs& a, hipStream_t s);
hipError_t launch_hydro_list(int Q, int storage, int collision, int odd, const HydroArgs& a, hipStream_t s);
// Batched domains: `tab` holds K HydroArgs on the device, `ent` 2 x nent
// (tile entry, domain) pairs, `cells` 2 x n (cell, domain) pairs.
hipError_t launch_hydro_batch(int Q, int storage, int collision, int odd, const HydroArgs* tab, const uint32_t* ent,
                              uint32_t nent, uint32_t grid, hipStream_t s);
hipError_t launch_hydro_list_batch(int Q, int storage, int collision, int odd, const HydroArgs* tab,
                                   const uint32_t* cells, uint32_t n, hipStream_t s);
hipError_t launch_clear_batch(const HydroArgs* tab, const uint32_t* cells, uint32_t n, hipStream_t s);
// Adjoint of one D3Q19 TRT/BGK collide-stream step (lbm_adjoint.hip).
hipError_t launch_adj_collide_stream(int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                                     const float* omega, const float* F_cur, const float* F_mom,
                                     const HydroParams& p, const double* lam_in, double* lam_out, double* omega_bar,
                                     double* Fbar, double* Gout, hipStream_t s);
hipError_t launch_fill_f64(double* p, size_t n, double v, hipStream_t s);
// probe passes (lbm_probe.hip): post = 0 before the step (densities), 1 after
hipError_t launch_probe(int Q, int storage, int post, const ProbeArgs& a, hipStream_t s);
// capture, restore and the per-face rules after a step (lbm_open.hip)
hipError_t launch_open_faces(int Q, int storage, const FaceArgs& a, hipStream_t s);
hipError_t launch_uncovered(const Dom& d, const uint8_t* solid, uint32_t open_bits, unsigned int* count,
                            hipStream_t s);
// WALE omega_plus field from the state and its cached moments `mom` (4 x N:
// rho, ux, uy, uz) (lbm_les.hip).  `stats` (2 words, cleared here by a
// kernel): clipped cell count, bits of the largest raw nu_t.
// work: 6 x N floats (cached velocity, off-diagonal strain)
hipError_t launch_wale(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                       const float* wall_u, const float* F_cur, const float* F_mom, float* work, const HydroParams& p,
                       float tau_eff, float cw_delta_sq, float nu0, float nut_max, float* omega, float* nut,
                       uint32_t* stats, hipStream_t s);
hipError_t launch_build_mask(int Q, const Dom& d, const uint8_t* solid, const float* wall_u, uint32_t* mask,
                             unsigned int* err_count, hipStream_t s);
hipError_t launch_mark_forced(const uint32_t* flist, uint32_t n, uint32_t* mask, hipStream_t s);
hipError_t launch_tile_flags(const Dom& d, const uint32_t* mask, uint8_t* flags, int linear, hipStream_t s);
hipError_t launch_stage(int Q, int storage, int dir, int odd, const Dom& d, void* f, float* host_layout,
                        const uint32_t* mask, hipStream_t s);
hipError_t launch_init_with(int Q, int storage, const Dom& d, void* f, const uint32_t* mask, const float* rho,
                            const float* u, float tau, hipStream_t s);
// previous step's post-collision populations (compact Q x N), see k_poststate
hipError_t launch_poststate(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                            const float* wall_u, float* out, hipStream_t s);
hipError_t launch_moments(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                          const float* F_mom, HydroParams p, float* rho, float* ux, float* uy, float* uz,
                          hipStream_t s);
hipError_t launch_reduce(int Q, int storage, int odd, const Dom& d, const void* f, const uint32_t* mask,
                         const float* F_mom, HydroParams p, int kind, int axis, double* partials,
                         uint32_t npartials, hipStream_t s);
hipError_t launch_tank(const Dom& d, double dx, double diameter, double height, int baffles, double bw,
                       double bt, uint8_t* solid, hipStream_t s);
hipError_t launch_rims(const Dom& d, const int is_wall[6], const double u[18], uint8_t* solid, float* wall_u,
                       double* best, hipStream_t s);
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
hipError_t launch_stream_copy(const void* a, void* b, size_t bytes, hipStream_t s);
// n elements of `width` bytes (16, 4 or 2) from a + offset_elems to
// b + offset_elems (offset 0, or 1 for widths 4 and 2)
hipError_t launch_access_copy(int width, int offset_elems, const void* a, void* b, size_t n, hipStream_t s);
hipError_t launch_copy_force_list(const uint32_t* flist, uint32_t n, float* F_dst, const float* F_src,
                                  hipStream_t s);

Tiling make_tiling(const Dom& d, uint32_t tx_pref);

}  // namespace lbm
