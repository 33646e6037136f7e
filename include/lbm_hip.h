/*
 * lbm_hip.h — C ABI of the MI355X (gfx950) lattice-Boltzmann hot path.
 *
 * Drop-in boundary for lbm-core's `Backend` trait (TakuTsuzuki/lbmflow,
 * crates/lbm-core/src/backend.rs:126-363) plus the Solver-side hooks that
 * the reference keeps in host code (rotating IBM, scalar ADE, reducers).
 * Plain C: pointers, sizes and status codes only.  Every entry point
 * returns 0 on success or a negative LBM_ERR_* code; the message of the
 * most recent failure on a context is available from lbm_last_error().
 * No call throws or aborts across the ABI.
 *
 * Array conventions (host side):
 *   cell     = (z * ny + y) * nx + x            compact global order
 *   f[q][cell]                                  q-major planes, deviation form f - w_q
 *   vec3[cell][3]                               interleaved xyz
 * This is the reference's compact-core order (fields.rs:143-148) with the
 * padded q-major planes (fields.rs:1-16) stripped of their halo.
 *
 * Threading: one HIP stream per context; calls on one context are
 * serialized by the caller (the reference Solver is single-threaded too,
 * gpu/backend.rs:38).  Readback calls block until the stream drains.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (UnsupportedReason-like, solver.rs:172-180) ---------- */
#define LBM_OK 0
#define LBM_ERR_INVALID (-1)     /* bad argument / shape */
#define LBM_ERR_UNSUPPORTED (-2) /* feature not implemented on this backend */
#define LBM_ERR_DEVICE (-3)      /* HIP runtime failure */
#define LBM_ERR_NOMEM (-4)       /* device allocation failed */
#define LBM_ERR_STATE (-5)       /* call out of order (e.g. no markers set) */

/* ---- enums -------------------------------------------------------------- */
/* Population storage precision (gpu/backend.rs:143-159 has f32|f16; bf16 is
 * new).  Storage always holds the deviation f - w_q; arithmetic is fp32. */
#define LBM_STORAGE_F32 0
#define LBM_STORAGE_F16 1
#define LBM_STORAGE_BF16 2

/* CollisionKind, params.rs:22-37 */
#define LBM_COLLISION_BGK 0
#define LBM_COLLISION_TRT 1
#define LBM_COLLISION_CM 2

/* Reduction, params.rs:411-419 */
#define LBM_REDUCE_FLUID_CELLS 0
#define LBM_REDUCE_MASS_DEVIATION 1
#define LBM_REDUCE_MOMENTUM 2

/* scalar distribution lattice: 0 = the hydro lattice (reference,
 * solver.rs:2113-2205), 7 = D3Q7 (WSCAL spec; cs^2 = 1/4) */
#define LBM_SCALAR_HYDRO_LATTICE 0
#define LBM_SCALAR_D3Q7 7

/* capability bits (backend.rs:157-199 probes) */
#define LBM_CAP_TRT 0x1u
#define LBM_CAP_CENTRAL_MOMENT 0x2u
#define LBM_CAP_F16_STORAGE 0x4u
#define LBM_CAP_BF16_STORAGE 0x8u
#define LBM_CAP_FORCE_FIELD 0x10u
#define LBM_CAP_OMEGA_FIELD 0x20u
#define LBM_CAP_ROTATING_IBM 0x40u
#define LBM_CAP_SCALAR 0x80u
#define LBM_CAP_PERIODIC 0x100u
#define LBM_CAP_OPEN_FACES 0x200u    /* Zou-He velocity/pressure, outflow, convective faces */
#define LBM_CAP_TWO_PASS 0x400u      /* not set: one fused pass per step */
#define LBM_CAP_GRAVITY 0x800u       /* not set */
#define LBM_CAP_PROBE 0x1000u        /* momentum-exchange force probe */
#define LBM_CAP_WALE 0x2000u         /* WALE omega_plus producer (les.rs) */
#define LBM_CAP_LOCALIZED 0x4000u    /* volume sources and face patches (supports_localized_features) */

typedef struct lbm_ctx lbm_ctx;
typedef struct lbm_fields lbm_fields;

/* StepParams, params.rs:302-320 (the subset on this path).  Relaxation rates
 * stay f64 here and are rounded to fp32 when kernel constants are built,
 * exactly as KParams::new does (params.rs:350-374). */
typedef struct lbm_step_params {
  int collision;      /* LBM_COLLISION_* */
  double omega_p;     /* 1/tau */
  double omega_m;     /* TRT antisymmetric rate (== omega_p for BGK/CM) */
  double omega_shear; /* CM shear rate */
  float force[3];     /* uniform Guo body force */
} lbm_step_params;

/* DirectForcingConfig, rotating_ibm.rs:104-130 */
typedef struct lbm_ibm_config {
  int max_iterations;    /* default 3 */
  double slip_tolerance; /* default 1e-2 */
  int kernel_radius;     /* 1 (linear) or 2 (quadratic B-spline) */
  double relaxation;     /* default 1.0 */
} lbm_ibm_config;

/* IbmDiagnostics, rotating_ibm.rs:132-153 */
typedef struct lbm_ibm_diag {
  int iterations;
  double slip_max, slip_rms, slip_max_rel, slip_rms_rel;
  double torque[3];
  double fluid_force[3];
  double marker_force[3];
  double momentum_error_rel;
} lbm_ibm_diag;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart (wgpu device creation, gpu/backend.rs:1895+). */
int lbm_ctx_create(int device, lbm_ctx** out);
void lbm_ctx_destroy(lbm_ctx* ctx);
const char* lbm_last_error(const lbm_ctx* ctx);
/* Backend capability probes, backend.rs:157-199. */
int lbm_capabilities(const lbm_ctx* ctx, uint32_t* flags);
/* Host-visible description of the compiled kernels (for diagnostics). */
const char* lbm_build_info(void);
/* Diagnostic with no reference counterpart: STREAM copy of `bytes` between
 * two device buffers, `iterations` times; *gbps counts read + write bytes.
 * The measured HBM ceiling that the collide-stream roofline is read against. */
int lbm_stream_copy(lbm_ctx* ctx, size_t bytes, int iterations, double* gbps);
/* Diagnostic with no reference counterpart: copy of `bytes` with one
 * element of `width` bytes per lane access (16; 4; 2), the element run
 * starting `offset_elems` elements (0, or 1 for widths 4 and 2) past
 * 256-byte-aligned bases: the access forms of the collide-stream kernels
 * (4-byte planes, their x-shifted neighbour rows, 2-byte planes), against
 * which the HBM traffic counters are calibrated.  *gbps counts read + write
 * bytes. */
int lbm_access_copy(lbm_ctx* ctx, int width, int offset_elems, size_t bytes, int iterations, double* gbps);

/* ---- fields ------------------------------------------------------------- */
/* Backend::alloc, backend.rs:138-139: quiescent fields (f = 0, rho = 1,
 * u = 0, no solids).  Q is 19 or 27.  periodic[a] != 0 wraps axis a
 * (Subdomain::monolithic, subdomain.rs:31-44). */
int lbm_fields_alloc(lbm_ctx* ctx, int Q, int storage, const int dims[3], const int periodic[3],
                     lbm_fields** out);
void lbm_fields_free(lbm_fields* fl);
/* Device bytes held by this domain. */
size_t lbm_fields_device_bytes(const lbm_fields* fl);

/* Solver::build mask distribution, solver.rs:1829-1848: solid (1 byte per
 * cell, nonzero = solid) and the per-cell wall velocity (meaningful on
 * solids).  Either pointer may be NULL (all fluid / resting walls).
 * A domain with a fluid cell on a non-periodic face cannot be stepped
 * unless an open-face rule (lbm_set_face_bc) or face patches
 * (lbm_set_face_patches) cover that face: lbm_run_span fails with
 * LBM_ERR_UNSUPPORTED until the masks, a face rule or patches close it. */
int lbm_set_masks(lbm_fields* fl, const uint8_t* solid, const float* wall_u);
int lbm_read_solid(lbm_fields* fl, uint8_t* solid);
/* Half-way bounce-back link list (implicit in stream_row, kernels.rs:572-603):
 * all (cell, q) with cell fluid and cell - c_q solid, ordered by cell then q.
 * With cells == NULL only *count is written. */
int lbm_boundary_links(lbm_fields* fl, int64_t* count, int64_t* cells, uint8_t* qs, int64_t capacity);

/* StepParams / KParams, params.rs:302-374. */
int lbm_set_params(lbm_fields* fl, const lbm_step_params* p);
/* Per-cell omega_plus (docs/KERNEL_EXTENSION_POINTS.md:22-73): NULL clears. */
int lbm_set_omega_field(lbm_fields* fl, const float* omega);

/* Open faces: FaceBC, params.rs:232-252, on the faces in Face::ALL order
 * (lattice.rs:38-45).  Applied after every collide-stream step in face
 * order, after the populations skipped by the stream have been restored to
 * their previous post-collide values (apply_open_faces_impl,
 * backend.rs:602-705; kernels.rs:849-1299).  No inlet profiles on this
 * path. */
#define LBM_FACE_XNEG 0
#define LBM_FACE_XPOS 1
#define LBM_FACE_YNEG 2
#define LBM_FACE_YPOS 3
#define LBM_FACE_ZNEG 4
#define LBM_FACE_ZPOS 5
#define LBM_FACE_CLOSED 0     /* no open-face pass (wall rim or periodic) */
#define LBM_FACE_VELOCITY 1   /* Zou-He, value = u[3] */
#define LBM_FACE_PRESSURE 2   /* Zou-He, value[0] = rho */
#define LBM_FACE_OUTFLOW 3    /* zero-gradient copy from one cell inward */
#define LBM_FACE_CONVECTIVE 4 /* value[0] = u_conv, mass pinned to the inner neighbour */
/* GlobalSpec faces[face] = kind (params.rs:302-320): an open face on a
 * periodic axis, a velocity above MAX_SPEED = 0.3 (params.rs:11) or a
 * non-positive density is LBM_ERR_INVALID.  Stepping refuses fluid cells on a
 * non-periodic face that no open face covers, and scalar distributions
 * together with open faces (LBM_ERR_UNSUPPORTED). */
int lbm_set_face_bc(lbm_fields* fl, int face, int kind, const double value[3]);

/* Solver::set_inlet_profile (solver.rs:3322-3350): a per-cell velocity on
 * a LBM_FACE_VELOCITY face, u[3] per face cell in face order c2 * n1 + c1
 * (t1, t2 = the face's other axes ascending; n1 = extent of t1), replacing
 * the face's uniform u in its own Zou-He rule (not on its patches).  NULL
 * clears.  LBM_ERR_INVALID on a face that is not a velocity face or a
 * non-finite value.  The profile is used while the face stays a velocity
 * face. */
int lbm_set_inlet_profile(lbm_fields* fl, int face, const double* values);

/* FacePatch (params.rs:286-296) and the patch arms of apply_open_faces_impl
 * (backend.rs:602-766): each patch gives an inclusive rect lo..hi of face
 * `face` (in-face coordinates over the two other axes in ascending order:
 * x faces (y, z), y faces (x, z), z faces (x, y)) its own rule, kind and
 * value as for lbm_set_face_bc; a LBM_FACE_CLOSED patch is a resting
 * (zero-velocity Zou-He) lid.  A face holding patches takes the open-face
 * pass even when its own rule is closed: its cells outside every patch then
 * get the resting lid.  lbm_set_face_patches replaces the list (n = 0
 * clears).  LBM_ERR_INVALID (validate_face_patches, solver.rs:690-740): bad
 * face or kind, a rect outside its face or inverted, two rects of one face
 * overlapping, a rule parameter out of range, a patch on a periodic axis,
 * more than 16 patches. */
typedef struct lbm_face_patch {
  int face;
  int lo[2], hi[2];
  int kind;
  double value[3];
} lbm_face_patch;
int lbm_set_face_patches(lbm_fields* fl, int n, const lbm_face_patch* patches);

/* VolumeSource / SourceKind (params.rs:262-283): after the open faces of
 * every step, each fluid cell of the inclusive box lo..hi gains
 * w_q q_cell (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u) with q_cell = q_lu / box
 * cells (apply_volume_sources_impl, backend.rs:799-856): q_lu of mass per
 * step, carrying u for a jet (negative q_lu: a sink).  Sources apply in
 * order; lbm_set_volume_sources replaces the list (n = 0 clears). */
#define LBM_SOURCE_MASS_FLOW 0
#define LBM_SOURCE_JET 1
typedef struct lbm_volume_source {
  int lo[3], hi[3];
  int kind;
  double q_lu;
  double u[3];
} lbm_volume_source;
int lbm_set_volume_sources(lbm_fields* fl, int n, const lbm_volume_source* src);

/* Solver::set_force_probe (solver.rs:2741-2760): `mask` marks the solid
 * cells whose momentum-exchange force is accumulated every step (N bytes,
 * nonzero = probed; NULL removes the probe). */
int lbm_set_force_probe(lbm_fields* fl, const uint8_t* mask);
/* Solver::probed_force / read_probed_force (solver.rs:5065-5079): force on
 * the probed solids during the most recent step, -sum c_q (f_out + f_in +
 * 2 w_q) over their half-way links with the reference's row-ordered fp32
 * sum (stream_row kernels.rs:605-615, backend.rs:521-544).  Zero without a
 * probe.  Bit-exact for probed walls at rest. */
int lbm_probed_force(lbm_fields* fl, double out[3]);

/* WALE large-eddy closure as an omega_plus producer.
 * WaleLesDiagnostics, les.rs:18-31 (tau_eff_max / nu_t_max are 0 when the
 * update does not clip). */
typedef struct lbm_wale_diag {
  int64_t clipped_cells;
  double clipped_fraction;
  double max_nu_t_before_clipping;
  double tau_eff_max;
  double nu_t_max;
} lbm_wale_diag;
/* WaleLes::update, les.rs:120-211 (replaces WaleLes::new().update(&mut
 * solver), or with_tau_eff_max(t).update): computes the velocity gradient of
 * the current state (Solver::gather_velocity_gradient, solver.rs:5490-5581),
 * nu_t with Delta = 1 and installs omega = 1/(3 (nu0 + nu_t) + 1/2) as the
 * per-cell omega_plus field of the next steps.  cw: model coefficient (the
 * reference fixes 0.325); tau_eff_max <= 0: no clip, otherwise it must exceed
 * 0.5 and the laminar tau (LBM_ERR_INVALID).  nu_t (host, N floats, after
 * the clip) and diag may be NULL. */
int lbm_wale_update(lbm_fields* fl, double cw, double tau_eff_max, float* nu_t, lbm_wale_diag* diag);
/* On-device WALE every step (GpuSolver::set_wale, gpu/solver.rs:60-73): with
 * enabled != 0 each lbm_run_span / lbm_run_ibm_span step first recomputes the
 * omega_plus field from the pre-collision state (after the IBM spread), as
 * lbm_wale_update does, then collides with it.  Disabling leaves the last
 * field installed; lbm_set_omega_field(fl, NULL) removes it
 * (WaleLes::clear, les.rs:213-220). */
int lbm_set_wale(lbm_fields* fl, int enabled, double cw, double tau_eff_max);
/* Installed omega_plus field to the host (GpuSolver::gather_wale_omega,
 * gpu/solver.rs:228-235): N floats; *present = 0 (and omega untouched)
 * when no field is installed. */
int lbm_gather_omega_field(lbm_fields* fl, float* omega, int* present);

/* Backend::stage_in / stage_out, backend.rs:141-155 (compact q-major).
 * stage_in makes the cached moments those of the staged populations with
 * the current force field (as Solver::build's update_moments does);
 * stage_out writes zero for solid cells. */
int lbm_stage_in(lbm_fields* fl, const float* f);
int lbm_stage_out(lbm_fields* fl, float* f);
/* Solver::init_with, solver.rs:2474-2623: f = feq(rho, u) + f_neq(grad u),
 * computed on the device from compact rho[N] and u[N][3]. */
int lbm_init_with(lbm_fields* fl, const float* rho, const float* u);
/* Backend::update_moments, backend.rs:243-249 (refresh the cached moments
 * from the populations with the current force field). */
int lbm_update_moments(lbm_fields* fl);

/* Solver::set_body_force_field / clear_body_force_field,
 * solver.rs:2776-2822 (compact vec3; refreshes moments only when uniform). */
int lbm_set_force_field(lbm_fields* fl, const float* ff);
int lbm_clear_force_field(lbm_fields* fl);
int lbm_read_force_field(lbm_fields* fl, float* ff, int* present);

/* Backend::run_span, backend.rs:262-328: `steps` whole steps (collide,
 * stream with half-way bounce-back and moving walls, moments), plus the
 * enabled scalar distributions after each hydro step (Solver::step,
 * solver.rs:2237-2280).  Asynchronous: returns after enqueueing. */
int lbm_run_span(lbm_fields* fl, int steps);
int lbm_synchronize(lbm_fields* fl);
/* Kernel timing (no reference counterpart; the reference times whole runs
 * with Instant, runner.rs:1076): with timing on, HIP events bracket every
 * main collide-stream launch on the context stream; _read returns the
 * summed elapsed time and the launch count since the last read. */
int lbm_kernel_timing(lbm_fields* fl, int enable);
int lbm_kernel_timing_read(lbm_fields* fl, double* total_ms, int64_t* launches);
/* Cells handled by the per-cell-force kernel (IBM stencils, force fields). */
int64_t lbm_force_list_size(const lbm_fields* fl);
/* Completed steps since alloc. */
int64_t lbm_time(const lbm_fields* fl);

/* Backend::read_moments, backend.rs:361-362 (compact; solids keep the
 * values of the last explicit write, as in the reference). */
int lbm_read_moments(lbm_fields* fl, float* rho, float* ux, float* uy, float* uz);
/* No counterpart in the reference.  lbm_read_moments for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernel and the same results, but
 * the device staging buffer is a fixed array inside the library (one per
 * GPU, held until the process ends) instead of a per-call hipMalloc.  At most
 * LBM_NOALLOC_MAX_CELLS cells (64 MiB of staging); larger domains return
 * LBM_ERR_INVALID.  The array is shared by every call on one GPU: calls that
 * overlap on one GPU (two contexts on two host threads) race. */
#define LBM_NOALLOC_MAX_CELLS 4194304
int lbm_read_moments_noalloc(lbm_fields* fl, float* rho, float* ux, float* uy, float* uz);
/* Backend::reduce, backend.rs:351-359: f64 accumulation over fluid cells. */
int lbm_reduce(lbm_fields* fl, int kind, int axis, double* out);

/* ---- device geometry (bit-exact with geometry.rs) ----------------------- */
/* build_wall_rims, solver.rs:1357-1406; is_wall[6] and u[6][3] in Face
 * order XNeg, XPos, YNeg, YPos, ZNeg, ZPos. Writes solid + wall_u. */
int lbm_geom_wall_rims(lbm_fields* fl, const int is_wall[6], const double u[18]);
/* apply_cylindrical_tank_mask + apply_baffle_mask (count 4),
 * geometry.rs:657-723.  Lengths in metres with cell size dx; baffle_count
 * 0 or 4.  ORs into the current solid mask, then rebuilds links. */
int lbm_geom_stirred_tank(lbm_fields* fl, double dx, double vessel_diameter, double liquid_height,
                          int baffle_count, double baffle_width, double baffle_thickness);
/* build_radial_impeller, geometry.rs:815-906 (host, libm trig).  blade
 * angle 0 = Rushton (inner radius 0.65 R), else pitched (0.35 R).  Writes
 * up to `capacity` markers (pos[n][3] in lattice units, weight[n],
 * velocity[n][3]) and center[3]; *count receives the full count. */
int lbm_geom_radial_impeller(const int dims[3], double dx, double vessel_diameter, double liquid_height,
                             double diameter, double clearance, int blades, double blade_angle_deg,
                             double omega_z, double thickness_fraction, double* pos, double* weight,
                             double* velocity, double* center, int64_t capacity, int64_t* count);
/* No counterpart in the reference: a seeded surface sampler for the same
 * blades (the marker set of the stirred-tank benchmark configuration).  The
 * blade boxes of build_radial_impeller (geometry.rs:815-906) are sampled at
 * uniform areal density: per blade N_b = max(1, round(A_b / (area_per_marker
 * dx^2))) markers, each on a face chosen with probability proportional to its
 * area at a uniform point of it, weight A_b / N_b (m^2), three SplitMix64
 * draws per marker from `seed`.  Output layout as lbm_geom_radial_impeller. */
int lbm_geom_impeller_surface(const int dims[3], double dx, double vessel_diameter, double liquid_height,
                              double diameter, double clearance, int blades, double blade_angle_deg, double omega_z,
                              double thickness_fraction, double area_per_marker, uint64_t seed, double* pos,
                              double* weight, double* velocity, double* center, int64_t capacity, int64_t* count);

/* ---- rotating IBM (Solver::apply_rotating_ibm, solver.rs:3016-3282) ----- */
/* Register a rigid marker body (RotatingBody::from_markers,
 * rotating_ibm.rs:31-52).  Marker positions are frozen, as in the
 * reference; the stencil cell sets are built here once. */
int lbm_ibm_set_body(lbm_fields* fl, int64_t n, const double* pos, const double* weight, const double center[3],
                     const double omega[3], const lbm_ibm_config* cfg);
/* One direct-forcing update: adds the spread impulses to the force field
 * (the caller clears it first, as the runner does, runner.rs:1080-1084). */
int lbm_ibm_apply(lbm_fields* fl, lbm_ibm_diag* diag);
/* `steps` runner steps (runner.rs:1079-1129: clear the force field, one
 * rotating-IBM update, one collide-stream step, the scalar distributions)
 * with no host synchronisation inside the span: the per-step diagnostics
 * are kept on the device and read back at the end.  torque (nullable)
 * receives steps x 3 values (IbmDiagnostics::torque of each step); last
 * (nullable) the final step's diagnostics.  use_graph != 0 replays pairs of
 * steps from one captured HIP graph (recaptured whenever masks, parameters,
 * force list, IBM body or scalars change); the results are identical to the
 * eager launches.  The reference's equivalent seam is Backend::run_chunk_size
 * / finish_run_chunk (backend.rs:340-349). */
int lbm_run_ibm_span(lbm_fields* fl, int steps, int use_graph, double* torque, lbm_ibm_diag* last);
/* Autotuner of the collide-stream launch on one domain: the visit order of
 * the active tiles (slabs of zb planes, y-rows within a slab, then the
 * planes) for each parity.  Times `trial_steps` (2..64, whole even/odd
 * pairs) main-kernel launches for each of six candidate depths per parity on
 * a scratch copy of the populations (HIP events on the library stream),
 * keeps the fastest for each parity and rebuilds the tile lists; the state
 * is untouched (the scratch copy needs device memory for a second
 * population buffer; LBM_ERR_NOMEM if it cannot be had, with the tile
 * lists unchanged).  zb_out = {even, odd}; ms_out (nullable) their mean
 * launch times.  No reference counterpart: the reference's GPU prototype fixes its
 * workgroup shapes (backend.rs:126-363 exposes no tuning). */
int lbm_autotune_tiles(lbm_fields* fl, int trial_steps, uint32_t zb_out[2], double ms_out[2]);
/* Running FP64 torque integral of the registered body, kept on the device
 * and advanced by every rotating-IBM update (lbm_ibm_apply, spans, batches):
 * out[0..2] = sum of IbmDiagnostics::torque over the updates since the body
 * was registered or the integral last reset, out[3] = their count.  reset
 * != 0 clears it after the read.  Replaces the host-side accumulation the
 * runner does for the power number (runner.rs:1514-1530, mean shaft torque
 * over the QOI window: qoi.rs:477-505), so a span needs no per-step torque
 * readback. */
int lbm_ibm_torque_integral(lbm_fields* fl, int reset, double out[4]);

/* ---- batched domains ------------------------------------------------------
 * No single reference counterpart: lbm-cli's sweep runs its cases one after
 * another through the runner (crates/lbm-cli/src/sweep.rs:61-106, grid
 * expansion crates/lbm-scenario/src/sweep.rs:52-71).  A batch steps several
 * independent closed domains (same lattice, storage and collision kind; each
 * with its own grid, masks, rates and rotating body) together: one persistent
 * collide-stream launch over their concatenated tile lists, one forced-cell
 * launch, one set of IBM launches for all bodies.  Members must not be
 * changed while a batch over them exists; the batch checks this. */
typedef struct lbm_batch lbm_batch;
int lbm_batch_create(lbm_ctx* ctx, lbm_fields* const* members, int n, lbm_batch** out);
void lbm_batch_destroy(lbm_batch* b);
int lbm_batch_info(const lbm_batch* b, int* members, uint32_t* tiles, uint32_t* forced_cells, uint32_t* grid);
/* Persistent grid (workgroups of 256) of the collide-stream launch. */
int lbm_batch_set_grid(lbm_batch* b, uint32_t grid);
/* `steps` runner steps of every member (as lbm_run_ibm_span, or plain
 * collide-stream steps when the members have no rotating body).  torque
 * (nullable) receives members x steps x 3 values, member-major. */
int lbm_batch_run_span(lbm_batch* b, int steps, int use_graph, double* torque);
/* Times the collide-stream launch over a set of persistent grid sizes on
 * scratch copies of the populations (the members' state is untouched) and
 * keeps the fastest. */
int lbm_batch_autotune(lbm_batch* b, int trial_steps, uint32_t* best_grid, double* best_ms);
/* The running FP64 torque integrals of the members' bodies (as
 * lbm_ibm_torque_integral), out = members x 4 in member order (zeros for a
 * member without a body); they advance only while the members step through
 * this batch. */
int lbm_batch_torque_integral(lbm_batch* b, int reset, double* out);

/* ---- adjoint of the time-averaged shaft torque --------------------------
 * No reference counterpart (the reference differentiates only by finite
 * differences, crates/lbm-core/src/uq.rs:47-70).  Runs `steps` runner steps
 * from the current state (as lbm_run_ibm_span) and sets *J to the mean shaft
 * torque (-torque_z) of the last `window` steps; then, if dJ_domega or
 * dJ_dnu is non-null, the discrete adjoint of those steps (D3Q19 TRT/BGK
 * collide-stream with half-way bounce-back, stale-moment Guo forcing and the
 * direct-forcing sweeps), recomputing the forward trajectory from
 * `snapshots` binomial checkpoints, gives dJ/dOmega_z of the rotating body
 * and dJ/dnu per cell (nu = (1/omega_plus - 1/2)/3; omega_minus fixed, as
 * the omega field does).  Adjoint arithmetic is f64.  The fields end in the
 * state after the `steps` forward steps.  *forward_steps (nullable) counts
 * every forward step taken, the first pass included. */
int lbm_adjoint_torque(lbm_fields* fl, int steps, int window, int snapshots, double* J, double* dJ_domega,
                       double* dJ_dnu, int64_t* forward_steps);
/* The checkpoint schedule used above, for inspection: `count` (action,
 * argument) pairs; action 0 advance arg steps, 1 save the state to slot arg,
 * 2 restore slot arg, 3 adjoint of step arg.  actions may be null to get the
 * count only. */
int lbm_revolve_schedule(int steps, int snapshots, int* actions, int capacity, int* count);

/* ---- checkpoint rank file (format v3) -------------------------------------
 * The binary part of Solver::save / Solver::restore for one monolithic part
 * (crates/lbm-core/src/solver.rs:4536-4630 writer, 1113-1215 reader; layout
 * constants 29-45): header, section table, F_PRIMARY / STALE_STASH /
 * MOMENTS / SOLID / FORCE_FIELD sections in the reference's padded host
 * layout, FNV-1a 64 payload hash.  The manifest and the scenario /
 * decomposition hashes are the caller's (lbmflow_amd/checkpoint.py).
 * Errors carry the reference's CKPT_* codes in lbm_last_error. */
typedef struct lbm_ckpt_info {
  uint64_t bytes;         /* rank file size */
  uint64_t payload_hash;  /* FNV-1a 64 of the payload */
  uint64_t mask_hash;     /* part_mask_hash of the current masks (solver.rs:967-979) */
  int64_t step;
} lbm_ckpt_info;
int lbm_checkpoint_write_part(lbm_fields* fl, const char* path, lbm_ckpt_info* info);
/* Installs the state of `path` (masks, parameters, IBM body must already be
 * those of the saved run) and sets the step counter to `step`.  On entry
 * info->mask_hash and info->payload_hash hold the manifest's values for the
 * part: the current masks are checked against the first before the file is
 * read (CKPT_MASK_MISMATCH, solver.rs:4954-4962) and the file's payload hash
 * against the second (CKPT_PAYLOAD_CORRUPT) before any state is installed.
 * On return info holds the file's values. */
int lbm_checkpoint_read_part(lbm_fields* fl, const char* path, int64_t step, lbm_ckpt_info* info);
/* FNV-1a 64 (offset basis 0xcbf29ce484222325, prime 0x100000001b3). */
uint64_t lbm_fnv1a64(const void* data, size_t n);

/* ---- passive scalar ADE (solver.rs:2113-2205, scalar.rs:6-15) ------------ */
int lbm_scalar_enable(lbm_fields* fl, double diffusivity, int lattice, int* index);
/* set_scalar_concentration, solver.rs:2910-2965 (h = h_eq(C, u)). */
int lbm_scalar_set(lbm_fields* fl, int index, const float* conc);
/* gather_scalar_concentration, solver.rs:2887-2907. */
int lbm_scalar_gather(lbm_fields* fl, int index, float* conc);
/* qoi::scalar_cv over fluid cells (qoi.rs:507-532), f64 on the device.
 * *defined = 0 when the mean is zero or there are no fluid cells. */
int lbm_scalar_cv(lbm_fields* fl, int index, double* cv, int* defined);
/* qoi::compartment_cv (qoi.rs:699-739; the runner's final mixing QOIs,
 * runner.rs:1818): the CV of scalar `index` over the fluid cells of three
 * horizontal compartments, in the order top, bulk, near_impeller.  A cell of
 * plane z is top when (z + 0.5)/nz >= 2/3, else near_impeller when
 * has_impeller_z and |(z + 0.5) - impeller_z| <= nz/6, else bulk.
 * defined[k] = 0 where the compartment is empty or its mean is zero (the
 * reference's None); cell_count[k] counts its fluid cells. */
int lbm_scalar_compartment_cv(lbm_fields* fl, int index, int has_impeller_z, double impeller_z, double cv[3],
                              int defined[3], int64_t cell_count[3]);

/* ---- QOI reducers (qoi.rs:477-556) --------------------------------------- */
/* power_qois: out = {power, n_hz, np, p_over_v}. */
int lbm_qoi_power(double torque, double omega, double rho, double diameter, double volume, double out[4]);
/* mixing_time_from_cv: out = {cv0, t95, t99}; NaN where not reached.
 * Returns LBM_ERR_STATE when no finite positive CV exists. */
int lbm_qoi_mixing_time(int64_t n, const double* t, const double* cv, double out[3]);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
