// lbm_api.cpp — C ABI over the HIP kernels: the Backend-phase host side
// (context, fields, masks, staging, force fields, stepping, readback,
// reductions, geometry, IBM, scalar ADE, QOI helpers).
//
// Reference interface: lbm-core's Backend trait (backend.rs:126-363) and the
// Solver methods that drive it (solver.rs).  Errors never cross the ABI as
// exceptions: each entry point returns an LBM_ERR_* code and records a
// message on the context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_aux.hpp"
#include "lbm_internal.hpp"
#include "lbm_lattice.hpp"

using namespace lbm;

struct lbm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

struct lbm_fields {
  lbm_ctx* ctx = nullptr;
  int Q = 19;
  int storage = kF32;
  Dom d{};
  Tiling t{};
  size_t elem = 4;
  void* f = nullptr;        // first element of plane 0
  void* f_alloc = nullptr;  // allocation start (guard band before plane 0)
  size_t f_bytes = 0;
  uint32_t* mask = nullptr;
  uint8_t* solid = nullptr;
  float* wall_u = nullptr;
  float* omega = nullptr;
  int parity = 0;
  int64_t time = 0;
  int collision = LBM_COLLISION_TRT;
  HydroParams hp{};
  lbm_step_params params{};
  bool params_set = false;
  // force fields
  float* F_cur = nullptr;
  float* F_mom = nullptr;
  bool field_present = false;
  unsigned int open_face_cells = 0;
  bool field_uniform = false;
  std::vector<uint32_t> flist_host;
  // active wave tiles of the main kernel
  uint32_t* tiles = nullptr;
  size_t tiles_cap = 0;
  uint32_t ntiles = 0;
  std::vector<uint32_t> tiles_host;  // the even-step visit order
  uint32_t* tiles_odd = nullptr;     // the same tiles in the odd-step visit order
  uint32_t zb_set[2] = {0, 0};       // visit-order slab depth per parity (0: default), set by the autotuner
  bool tiles_linear = false;         // tiles are 64-cell blocks of the linear index
  uint32_t* flist = nullptr;
  size_t flist_cap = 0;
  // reduction scratch
  double* partials = nullptr;
  uint32_t npartials = 0;
  // kernel timing (HIP events around each main hydro launch)
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  // rotating IBM and scalar distributions
  IbmState* ibm = nullptr;
  std::vector<ScalarState*> scalars;
  // device-resident step loop: per-step IBM diagnostics ring, and a HIP
  // graph of two consecutive steps (even + odd) with the key it was
  // captured under (every argument the launches bake in)
  void* ring = nullptr;
  uint32_t ring_cap = 0;
  uint32_t* ring_pos = nullptr;
  uint64_t gen = 0;  // bumped by every change a captured step would not see
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  HydroArgs g_args{};
  int g_parity = -1;
  uint64_t g_gen = 0;
  // open faces (Face::ALL order): rule and values, the per-face stash
  // (two generations of unknowns x face cells), which generation is current,
  // and the fluid face cells no open face covers
  int face_kind[6] = {0, 0, 0, 0, 0, 0};
  std::vector<FacePatchArgs> patches;  // rect overrides (closed ones as zero velocity)
  float* prof[6] = {};                 // inlet profiles: u[3] per face cell, or null
  int patched[6] = {0, 0, 0, 0, 0, 0};  // faces holding at least one patch
  float face_u[6][3] = {};
  float face_rho[6] = {};
  float face_uconv[6] = {};
  float* stash[6][2] = {};
  size_t stash_bytes[6] = {};
  int stash_sel = 0;
  unsigned int uncovered = 0;
  int g_stash = -1;
  // volume sources, applied after the open faces in order
  std::vector<SourceArgs> sources;
  // force probe: the probed-solid mask (compact, empty = none) and its link
  // list on the device, rebuilt after a mask change
  std::vector<uint8_t> probe_host;
  bool probe_built = false;
  ProbeArgs pr{};
  std::vector<std::pair<void*, size_t>> pr_allocs;
  // WALE LES: velocity and strain work array (6 x N), diagnostics words, and the
  // per-step mode's coefficients
  float* wale_mom = nullptr;
  uint32_t* wale_stats = nullptr;
  bool wale_on = false;
  float wale_cw2 = 0.0f, wale_nut_max = -1.0f;
  size_t bytes = 0;
};

// Staging array of lbm_read_moments_noalloc: rho, ux, uy, uz of at most
// LBM_NOALLOC_MAX_CELLS cells.  Each GPU holds its own copy at its own
// address; host code never passes this name to a kernel or a copy (in host
// code it is a stand-in, not a device address) but the per-device address
// that lbm_ctx_create resolves with hipGetSymbolAddress.
static __device__ __align__(256) float scratch[4 * (size_t)LBM_NOALLOC_MAX_CELLS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Once per device, with that device current: the device address of
// `scratch`.  A failure is recorded and reported by the entry point.
void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? (float*)p : nullptr;
  g_scratch_done[device] = true;
}

float* scratch_of(int device, hipError_t* err) {
  if (device < 0 || device >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  *err = g_scratch_done[device] ? g_scratch_err[device] : hipErrorNotInitialized;
  return g_scratch_ptr[device];
}

int fail(lbm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                              \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess) return fail((ctx), LBM_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

int dev_alloc(lbm_fields* fl, void** p, size_t bytes) {
  if (bytes == 0) bytes = 1;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(fl->ctx, LBM_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  fl->bytes += bytes;
  return LBM_OK;
}

void dev_free(lbm_fields* fl, void* p, size_t bytes) {
  if (p) {
    (void)hipFree(p);
    fl->bytes -= std::min(fl->bytes, bytes);
  }
}

size_t N3(const lbm_fields* fl) { return (size_t)fl->d.N * 3; }

// Destroys a captured step graph once the launches already queued on `s`
// (replays of that graph among them) have finished: an executable graph is
// not destroyed under a replay still in flight.
hipError_t drop_graph(hipStream_t s, hipGraphExec_t* gexec, hipGraph_t* graph) {
  hipError_t e = hipSuccess;
  if (*gexec || *graph) e = hipStreamSynchronize(s);
  if (*gexec) (void)hipGraphExecDestroy(*gexec);
  if (*graph) (void)hipGraphDestroy(*graph);
  *gexec = nullptr;
  *graph = nullptr;
  return e;
}

// Plane padding in elements; LBM_PLANE_PAD overrides the default (tuning).
uint64_t plane_pad_elems() {
  const char* e = getenv("LBM_PLANE_PAD");
  if (!e || !*e) return UINT64_MAX;
  return strtoull(e, nullptr, 10);
}

HydroParams build_hp(const lbm_step_params& p) {
  // KParams::new, params.rs:350-374: rates rounded to fp32, Guo prefactors
  // computed in f64 then rounded.
  HydroParams h{};
  h.omega_p = (float)p.omega_p;
  h.omega_m = (float)p.omega_m;
  h.omega_shear = (float)(p.collision == LBM_COLLISION_CM ? p.omega_shear : p.omega_p);
  h.cp = (float)(1.0 - p.omega_p / 2.0);
  h.cm = (float)(1.0 - p.omega_m / 2.0);
  for (int a = 0; a < 3; ++a) h.F[a] = p.force[a];
  h.force_on = (p.force[0] != 0.0f || p.force[1] != 0.0f || p.force[2] != 0.0f) ? 1 : 0;
  return h;
}

int ensure_force_arrays(lbm_fields* fl) {
  if (fl->F_cur) return LBM_OK;
  int rc = dev_alloc(fl, (void**)&fl->F_cur, N3(fl) * 4);
  if (rc) return rc;
  rc = dev_alloc(fl, (void**)&fl->F_mom, N3(fl) * 4);
  if (rc) return rc;
  HIPCHK(fl->ctx, hipMemsetAsync(fl->F_cur, 0, N3(fl) * 4, fl->ctx->stream));
  HIPCHK(fl->ctx, hipMemsetAsync(fl->F_mom, 0, N3(fl) * 4, fl->ctx->stream));
  return LBM_OK;
}

int rebuild_tiles(lbm_fields* fl);

// Upload the sorted union of the current list and `extra` cells, and flag
// them in the mask so the main kernel leaves them to the list kernel.
int grow_force_list(lbm_fields* fl, const std::vector<uint32_t>& extra) {
  std::vector<uint32_t> merged;
  merged.reserve(fl->flist_host.size() + extra.size());
  std::merge(fl->flist_host.begin(), fl->flist_host.end(), extra.begin(), extra.end(), std::back_inserter(merged));
  merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
  if (merged.size() == fl->flist_host.size()) return LBM_OK;
  fl->flist_host.swap(merged);
  if (fl->flist_host.size() > fl->flist_cap) {
    dev_free(fl, fl->flist, fl->flist_cap * 4);
    fl->flist = nullptr;
    fl->flist_cap = std::max<size_t>(fl->flist_host.size(), 2 * fl->flist_cap);
    int rc = dev_alloc(fl, (void**)&fl->flist, fl->flist_cap * 4);
    if (rc) return rc;
  }
  HIPCHK(fl->ctx, hipMemcpyAsync(fl->flist, fl->flist_host.data(), fl->flist_host.size() * 4,
                                 hipMemcpyHostToDevice, fl->ctx->stream));
  HIPCHK(fl->ctx, launch_mark_forced(fl->flist, (uint32_t)fl->flist_host.size(), fl->mask, fl->ctx->stream));
  // the host vector must outlive the async copy
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  return rebuild_tiles(fl);
}

// Rebuild the list of 64-cell tiles the main kernel must visit.  Row tiles
// are x-segments of one row; when no axis wraps and the row length is not
// a multiple of 64, tiles are 64-cell blocks of the linear index instead
// (rows packed end to end: on the 64-tank sweep, edges 48 to 160, 14 %
// fewer tiles for the same cells), which is exact because without wrapping
// every neighbour shift is the linear one.
int rebuild_tiles(lbm_fields* fl) {
  const uint32_t nseg = (fl->d.nx + 63) / 64;
  const bool linear = fl->d.nx % 64 != 0 && !fl->d.per[0] && !fl->d.per[1] && !fl->d.per[2] && !fl->open_face_cells &&
                      !(getenv("LBM_ROW_TILES") && getenv("LBM_ROW_TILES")[0] == '1');
  const size_t nt = linear ? ((size_t)fl->d.N + 63) / 64 : (size_t)nseg * fl->d.ny * fl->d.nz;
  uint8_t* dflags = nullptr;
  int rc = dev_alloc(fl, (void**)&dflags, nt);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  std::vector<uint8_t> flags(nt);
  HIPCHK(fl->ctx, launch_tile_flags(fl->d, fl->mask, dflags, linear ? 1 : 0, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(flags.data(), dflags, nt, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, dflags, nt);
  // Visit order: slabs of zb planes, y-rows within a slab, then the planes of
  // the slab, then the x-segments, one order per parity.  Even steps read
  // and write each cell's own slots and stream best in plain row order
  // (zb = 1).  Odd steps write a row's slots from the row itself
  // (bounce-back) and from its upstream rows; keeping z-neighbour rows a
  // few tiles apart lets one XCD's L2 merge those writes into whole lines
  // before they reach memory (512^3 D3Q27 tank, odd step: fp32 4.68 ms at
  // zb = 1, 4.48 at zb = 32; fp16 3.71 at zb = 1, 3.11 at zb = 16).
  uint32_t zb_even = 1, zb_odd = fl->storage == kF32 ? 32 : 16;
  if (const char* e = getenv("LBM_TILE_ZBLOCK")) {  // "n" (both parities) or "even,odd"
    zb_even = zb_odd = (uint32_t)std::max(1, atoi(e));
    if (const char* c = strchr(e, ',')) zb_odd = (uint32_t)std::max(1, atoi(c + 1));
  }
  if (fl->zb_set[0]) zb_even = fl->zb_set[0];
  if (fl->zb_set[1]) zb_odd = fl->zb_set[1];
  const uint32_t ny = fl->d.ny, nz = fl->d.nz, nx = fl->d.nx;
  auto build = [&](uint32_t zb) {
    std::vector<uint32_t> list;
    list.reserve(nt);
    for (uint32_t z0 = 0; z0 < nz; z0 += zb)
      for (uint32_t y = 0; y < ny; ++y)
        for (uint32_t z = z0; z < std::min(nz, z0 + zb); ++z) {
          // the row's tiles: its x-segments, or the linear tiles whose
          // first cell lies in the row
          const size_t r0 = ((size_t)z * ny + y) * nx;
          const size_t ta = linear ? (r0 + 63) / 64 : ((size_t)z * ny + y) * nseg;
          const size_t tb = linear ? (r0 + nx + 63) / 64 : ta + nseg;
          for (size_t t = ta; t < tb; ++t) {
            const uint8_t fg = flags[t];
            if (!(fg & 1)) continue;
            uint32_t e = (uint32_t)t | ((fg & 2) ? kTilePlain : 0u);
            // an all-solid half of an fp32 tile is one whole line per plane
            // nobody needs; with 16-bit storage it is half a line, and
            // leaving it out (a 64-byte partial write) measured slower than
            // keeping its zero stores (2.83 vs 2.75 ms on the 512^3 tank)
            if (fl->storage == kF32) e |= ((fg & 4) ? kTileDeadLo : 0u) | ((fg & 8) ? kTileDeadHi : 0u);
            list.push_back(e);
          }
        }
    return list;
  };
  std::vector<uint32_t> list = build(zb_even);
  const std::vector<uint32_t> list_odd = build(zb_odd);
  if (list.size() > fl->tiles_cap) {
    dev_free(fl, fl->tiles, fl->tiles_cap * 4);
    dev_free(fl, fl->tiles_odd, fl->tiles_cap * 4);
    fl->tiles = fl->tiles_odd = nullptr;
    fl->tiles_cap = std::max<size_t>(list.size(), 1);
    rc = dev_alloc(fl, (void**)&fl->tiles, fl->tiles_cap * 4);
    if (!rc) rc = dev_alloc(fl, (void**)&fl->tiles_odd, fl->tiles_cap * 4);
    if (rc) return rc;
  }
  if (!list.empty()) {
    HIPCHK(fl->ctx, hipMemcpyAsync(fl->tiles, list.data(), list.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(fl->ctx, hipMemcpyAsync(fl->tiles_odd, list_odd.data(), list_odd.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(fl->ctx, hipStreamSynchronize(s));
  }
  fl->ntiles = (uint32_t)list.size();
  fl->tiles_host.swap(list);
  fl->tiles_linear = linear;
  fl->gen++;
  return LBM_OK;
}

// The odd-state layout depends on the wall links (a bounced population sits
// in the fluid cell's own slot), so before the links change the populations
// are brought to the even state (every cell's own slots) under the old mask.
static int normalize_parity(lbm_fields* fl) {
  if (!fl->parity) return LBM_OK;
  const size_t bytes = (size_t)fl->Q * fl->d.N * 4;
  float* tmp = nullptr;
  int rc = dev_alloc(fl, (void**)&tmp, bytes);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, launch_stage(fl->Q, fl->storage, 1, 1, fl->d, fl->f, tmp, fl->mask, s));
  HIPCHK(fl->ctx, launch_stage(fl->Q, fl->storage, 0, 0, fl->d, fl->f, tmp, fl->mask, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, tmp, bytes);
  fl->parity = 0;
  return LBM_OK;
}

// faces with an open-face pass: an open rule, or patches (a closed face
// with patches is a zero-velocity lid outside them)
uint32_t open_bits(const lbm_fields* fl) {
  uint32_t b = 0;
  for (int k = 0; k < 6; ++k)
    if (fl->face_kind[k] || fl->patched[k]) b |= 1u << k;
  return b;
}

// the face's two stash generations, zeroed (the reference's second buffer)
static int ensure_stash(lbm_fields* fl, int face) {
  if (fl->stash[face][0]) return LBM_OK;
  const int ax = face / 2;
  const uint32_t ext[3] = {fl->d.nx, fl->d.ny, fl->d.nz};
  const uint32_t t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const size_t bytes = (size_t)ext[t1] * ext[t2] * (fl->Q == 27 ? 9 : 5) * sizeof(float);
  hipStream_t s = fl->ctx->stream;
  for (int g = 0; g < 2; ++g) {
    int rc = dev_alloc(fl, (void**)&fl->stash[face][g], bytes);
    if (rc) return rc;
    HIPCHK(fl->ctx, hipMemsetAsync(fl->stash[face][g], 0, bytes, s));
  }
  fl->stash_bytes[face] = bytes;
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  return LBM_OK;
}

// Fluid cells on a non-periodic face that no open face rule covers (their
// skipped populations would stay stale forever): with no open face these are
// all the face fluid cells.
int count_uncovered(lbm_fields* fl) {
  const uint32_t bits = open_bits(fl);
  if (!bits || !fl->open_face_cells) {
    fl->uncovered = fl->open_face_cells;
    return LBM_OK;
  }
  unsigned int* cnt = nullptr;
  int rc = dev_alloc(fl, (void**)&cnt, sizeof(unsigned int));
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, hipMemsetAsync(cnt, 0, sizeof(unsigned int), s));
  HIPCHK(fl->ctx, launch_uncovered(fl->d, fl->solid, bits, cnt, s));
  unsigned int n = 0;
  HIPCHK(fl->ctx, hipMemcpyAsync(&n, cnt, sizeof n, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, cnt, sizeof(unsigned int));
  fl->uncovered = n;
  return LBM_OK;
}

int rebuild_mask(lbm_fields* fl) {
  int rc = normalize_parity(fl);
  if (rc) return rc;
  unsigned int* err = nullptr;
  rc = dev_alloc(fl, (void**)&err, sizeof(unsigned int));
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, hipMemsetAsync(err, 0, sizeof(unsigned int), s));
  HIPCHK(fl->ctx, launch_build_mask(fl->Q, fl->d, fl->solid, fl->wall_u, fl->mask, err, s));
  if (!fl->flist_host.empty())
    HIPCHK(fl->ctx, launch_mark_forced(fl->flist, (uint32_t)fl->flist_host.size(), fl->mask, s));
  unsigned int nerr = 0;
  HIPCHK(fl->ctx, hipMemcpyAsync(&nerr, err, sizeof nerr, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, err, sizeof(unsigned int));
  // Fluid on a non-periodic face needs an open-face BC, which is not on this
  // path; stepping such a domain is refused (lbm_run_span).
  fl->open_face_cells = nerr;
  fl->probe_built = false;
  int rc2 = count_uncovered(fl);
  if (rc2) return rc2;
  return rebuild_tiles(fl);
}

HydroArgs hydro_args(lbm_fields* fl) {
  HydroArgs a{};
  a.d = fl->d;
  a.t = fl->t;
  a.tiles = fl->parity ? fl->tiles_odd : fl->tiles;
  a.ntiles = fl->ntiles;
  a.nseg = (fl->d.nx + 63) / 64;
  a.f = fl->f;
  a.mask = fl->mask;
  a.wall_u = fl->wall_u;
  a.omega = fl->omega;
  a.p = fl->hp;
  a.flist = fl->flist;
  a.nlist = (uint32_t)fl->flist_host.size();
  a.F_mom = fl->F_mom;
  a.F_cur = fl->F_cur;
  a.open = fl->open_face_cells ? 1 : 0;
  a.lin = fl->tiles_linear ? 1 : 0;
  return a;
}

// Cached moments now carry the current field (Backend::update_moments).
int refresh_moments(lbm_fields* fl) {
  if (fl->F_mom && !fl->flist_host.empty())
    HIPCHK(fl->ctx, launch_copy_force_list(fl->flist, (uint32_t)fl->flist_host.size(), fl->F_mom,
                                           fl->field_present ? fl->F_cur : nullptr, fl->ctx->stream));
  return LBM_OK;
}

}  // namespace

// Accessors used by the auxiliary translation unit (IBM / scalar ADE).
namespace lbm {
FieldsView view_of(lbm_fields* fl) {
  FieldsView v{};
  v.ctx_err = &fl->ctx->err;
  v.stream = fl->ctx->stream;
  v.Q = fl->Q;
  v.storage = fl->storage;
  v.d = fl->d;
  v.f = fl->f;
  v.mask = fl->mask;
  v.parity = fl->parity;
  v.hp = fl->hp;
  v.F_cur = fl->F_cur;
  v.F_mom = fl->F_mom;
  v.field_present = fl->field_present;
  return v;
}
}  // namespace lbm

extern "C" {

const char* lbm_build_info(void) {
  return "lbmflow_amd liblbm_hip: gfx950, AA-pattern fused collide-stream, TRT/BGK (D3Q19/D3Q27), "
         "central moment (D3Q19/D3Q27), storage f32|f16|bf16 (shifted, 16-bit scaled by 1/w)";
}

int lbm_ctx_create(int device, lbm_ctx** out) {
  if (!out) return LBM_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return LBM_ERR_DEVICE;
  if (device < 0 || device >= n) return LBM_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return LBM_ERR_DEVICE;
  resolve_scratch(device);
  lbm_ctx* c = new (std::nothrow) lbm_ctx();
  if (!c) return LBM_ERR_NOMEM;
  c->device = device;
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return LBM_ERR_DEVICE;
  }
  *out = c;
  return LBM_OK;
}

void lbm_ctx_destroy(lbm_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* lbm_last_error(const lbm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lbm_stream_copy(lbm_ctx* ctx, size_t bytes, int iterations, double* gbps) {
  if (!ctx || !gbps || iterations <= 0 || bytes < 16) return LBM_ERR_INVALID;
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess) return fail(ctx, LBM_ERR_NOMEM, "stream copy: hipMalloc failed");
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(ctx, LBM_ERR_NOMEM, "stream copy: hipMalloc failed");
  }
  hipStream_t s = ctx->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LBM_OK;
  float ms = 0.0f;
  if (hipMemsetAsync(a, 0, bytes, s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess || launch_stream_copy(a, b, bytes, s) != hipSuccess ||
      hipEventRecord(e0, s) != hipSuccess)
    rc = LBM_ERR_DEVICE;
  for (int k = 0; rc == LBM_OK && k < iterations; ++k)
    if (launch_stream_copy(a, b, bytes, s) != hipSuccess) rc = LBM_ERR_DEVICE;
  if (rc == LBM_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = LBM_ERR_DEVICE;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return fail(ctx, rc, "stream copy failed");
  *gbps = 2.0 * (double)bytes * iterations / (ms * 1e-3) / 1e9;
  return LBM_OK;
}

int lbm_access_copy(lbm_ctx* ctx, int width, int offset_elems, size_t bytes, int iterations, double* gbps) {
  if (!ctx || !gbps || iterations <= 0 || bytes < 256) return LBM_ERR_INVALID;
  if (!((width == 16 && offset_elems == 0) || ((width == 4 || width == 2) && (offset_elems == 0 || offset_elems == 1))))
    return fail(ctx, LBM_ERR_INVALID, "access copy: width %d offset %d", width, offset_elems);
  const size_t n = bytes / (size_t)width;
  const size_t alloc = n * (size_t)width + 256;  // room for the offset element
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, alloc) != hipSuccess) return fail(ctx, LBM_ERR_NOMEM, "access copy: hipMalloc failed");
  if (hipMalloc(&b, alloc) != hipSuccess) {
    (void)hipFree(a);
    return fail(ctx, LBM_ERR_NOMEM, "access copy: hipMalloc failed");
  }
  hipStream_t s = ctx->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LBM_OK;
  float ms = 0.0f;
  if (hipMemsetAsync(a, 0, alloc, s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess || launch_access_copy(width, offset_elems, a, b, n, s) != hipSuccess ||
      hipEventRecord(e0, s) != hipSuccess)
    rc = LBM_ERR_DEVICE;
  for (int k = 0; rc == LBM_OK && k < iterations; ++k)
    if (launch_access_copy(width, offset_elems, a, b, n, s) != hipSuccess) rc = LBM_ERR_DEVICE;
  if (rc == LBM_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = LBM_ERR_DEVICE;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return fail(ctx, rc, "access copy failed");
  *gbps = 2.0 * (double)(n * (size_t)width) * iterations / (ms * 1e-3) / 1e9;
  return LBM_OK;
}

int lbm_capabilities(const lbm_ctx* ctx, uint32_t* flags) {
  if (!ctx || !flags) return LBM_ERR_INVALID;
  *flags = LBM_CAP_TRT | LBM_CAP_CENTRAL_MOMENT | LBM_CAP_F16_STORAGE | LBM_CAP_BF16_STORAGE | LBM_CAP_FORCE_FIELD |
           LBM_CAP_OMEGA_FIELD | LBM_CAP_ROTATING_IBM | LBM_CAP_SCALAR | LBM_CAP_PERIODIC | LBM_CAP_WALE | LBM_CAP_OPEN_FACES | LBM_CAP_PROBE |
           LBM_CAP_LOCALIZED;
  return LBM_OK;
}

int lbm_fields_alloc(lbm_ctx* ctx, int Q, int storage, const int dims[3], const int periodic[3],
                     lbm_fields** out) {
  if (!ctx || !out || !dims) return LBM_ERR_INVALID;
  *out = nullptr;
  if (Q != 19 && Q != 27) return fail(ctx, LBM_ERR_UNSUPPORTED, "lattice D3Q%d is not supported (19 or 27)", Q);
  if (storage < 0 || storage > 2) return fail(ctx, LBM_ERR_INVALID, "storage %d", storage);
  for (int a = 0; a < 3; ++a)
    if (dims[a] < 1) return fail(ctx, LBM_ERR_INVALID, "dims must be positive");
  const uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
  // byte offsets inside one population plane are 32-bit in the kernels
  if (n >= (1ull << 30)) return fail(ctx, LBM_ERR_INVALID, "domain of %llu cells exceeds 2^30", (unsigned long long)n);
  lbm_fields* fl = new (std::nothrow) lbm_fields();
  if (!fl) return LBM_ERR_NOMEM;
  fl->ctx = ctx;
  fl->Q = Q;
  fl->storage = storage;
  fl->d.nx = dims[0];
  fl->d.ny = dims[1];
  fl->d.nz = dims[2];
  fl->d.N = (uint32_t)n;
  for (int a = 0; a < 3; ++a) fl->d.per[a] = periodic ? (periodic[a] != 0) : 0;
  // Plane stride: the cell count rounded up to 256 elements, plus an odd
  // multiple of 256 elements so the Q concurrent plane streams do not share
  // the same low address bits (a 512^3 fp32 plane is exactly 2^29 bytes).
  // Measured on the 512^3 D3Q27 tank (mean collide-stream launch, three
  // runs each): fp32 4.271 ms at 1280 elements vs 4.290 at 2304; fp16
  // 2.526 ms at 256 vs 2.550 at 2304.
  {
    uint64_t ps = (n + 255) / 256 * 256;
    uint64_t pad = plane_pad_elems();
    if (pad == UINT64_MAX) pad = storage == kF32 ? 256 * 5 : 256;
    ps += pad;
    if (ps * 4 >= (1ull << 32)) ps = n;
    fl->d.ps = (uint32_t)ps;
  }
  fl->t = make_tiling(fl->d, 0);
  fl->elem = storage == kF32 ? 4 : 2;
  // Guard bands of one z-slab (+ row + cell) before plane 0 and after the
  // last plane: the collide-stream kernel issues its neighbour loads before
  // it knows whether a cell is solid, and a solid cell on the domain face
  // would otherwise address just outside the buffer.
  const size_t guard = ((size_t)fl->d.nx * fl->d.ny + fl->d.nx + 1 + 255) / 256 * 256;
  fl->f_bytes = ((size_t)Q * fl->d.ps + 2 * guard) * fl->elem;
  int rc = dev_alloc(fl, &fl->f_alloc, fl->f_bytes);
  if (!rc) fl->f = (char*)fl->f_alloc + guard * fl->elem;
  if (!rc) rc = dev_alloc(fl, (void**)&fl->mask, n * 4);
  if (!rc) rc = dev_alloc(fl, (void**)&fl->solid, n);
  fl->npartials = 1024;
  if (!rc) rc = dev_alloc(fl, (void**)&fl->partials, fl->npartials * sizeof(double));
  if (rc) {
    lbm_fields_free(fl);
    return rc;
  }
  hipStream_t s = ctx->stream;
  if (hipMemsetAsync(fl->f_alloc, 0, fl->f_bytes, s) != hipSuccess ||
      hipMemsetAsync(fl->solid, 0, n, s) != hipSuccess) {
    lbm_fields_free(fl);
    return fail(ctx, LBM_ERR_DEVICE, "memset failed");
  }
  // default parameters: TRT with nu = 1/6 (tau = 1)
  lbm_step_params p{};
  p.collision = LBM_COLLISION_TRT;
  p.omega_p = 1.0;
  p.omega_m = 1.0 / ((3.0 / 16.0) / 0.5 + 0.5);
  p.omega_shear = 1.0;
  fl->params = p;
  fl->hp = build_hp(p);
  fl->collision = p.collision;
  // all-fluid default (steppable only when every axis is periodic)
  rc = rebuild_mask(fl);
  if (rc) {
    lbm_fields_free(fl);
    return rc;
  }
  ctx->err.clear();
  *out = fl;
  return LBM_OK;
}

void lbm_fields_free(lbm_fields* fl) {
  if (!fl) return;
  (void)hipStreamSynchronize(fl->ctx->stream);
  if (fl->ibm) ibm_destroy(fl->ibm);
  for (ScalarState* s : fl->scalars) scalar_destroy(s);
  for (hipEvent_t e : fl->ev_pool) (void)hipEventDestroy(e);
  if (fl->gexec) (void)hipGraphExecDestroy(fl->gexec);
  if (fl->graph) (void)hipGraphDestroy(fl->graph);
  for (void* p : {(void*)fl->f_alloc, (void*)fl->mask, (void*)fl->solid, (void*)fl->wall_u, (void*)fl->omega,
                  (void*)fl->F_cur, (void*)fl->F_mom, (void*)fl->flist, (void*)fl->partials, (void*)fl->tiles, (void*)fl->tiles_odd,
                  (void*)fl->ring, (void*)fl->ring_pos, (void*)fl->wale_mom, (void*)fl->wale_stats})
    if (p) (void)hipFree(p);
  for (int k = 0; k < 6; ++k) {
    for (int g = 0; g < 2; ++g)
      if (fl->stash[k][g]) (void)hipFree(fl->stash[k][g]);
    if (fl->prof[k]) (void)hipFree(fl->prof[k]);
  }
  for (auto& pa : fl->pr_allocs) (void)hipFree(pa.first);
  delete fl;
}

size_t lbm_fields_device_bytes(const lbm_fields* fl) { return fl ? fl->bytes : 0; }

int lbm_set_masks(lbm_fields* fl, const uint8_t* solid, const float* wall_u) {
  if (!fl) return LBM_ERR_INVALID;
  hipStream_t s = fl->ctx->stream;
  const size_t n = fl->d.N;
  if (solid) {
    HIPCHK(fl->ctx, hipMemcpyAsync(fl->solid, solid, n, hipMemcpyHostToDevice, s));
  } else {
    HIPCHK(fl->ctx, hipMemsetAsync(fl->solid, 0, n, s));
  }
  bool any = false;
  if (wall_u)
    for (size_t k = 0; k < 3 * n && !any; ++k) any = wall_u[k] != 0.0f;
  if (any) {
    if (!fl->wall_u) {
      int rc = dev_alloc(fl, (void**)&fl->wall_u, 3 * n * 4);
      if (rc) return rc;
    }
    HIPCHK(fl->ctx, hipMemcpyAsync(fl->wall_u, wall_u, 3 * n * 4, hipMemcpyHostToDevice, s));
  } else if (fl->wall_u) {
    dev_free(fl, fl->wall_u, 3 * n * 4);
    fl->wall_u = nullptr;
  }
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  int rc = rebuild_mask(fl);
  if (rc) return rc;
  return refresh_moments(fl);
}

int lbm_read_solid(lbm_fields* fl, uint8_t* solid) {
  if (!fl || !solid) return LBM_ERR_INVALID;
  HIPCHK(fl->ctx, hipMemcpyAsync(solid, fl->solid, fl->d.N, hipMemcpyDeviceToHost, fl->ctx->stream));
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  return LBM_OK;
}

int lbm_boundary_links(lbm_fields* fl, int64_t* count, int64_t* cells, uint8_t* qs, int64_t capacity) {
  if (!fl || !count) return LBM_ERR_INVALID;
  std::vector<uint32_t> m(fl->d.N);
  HIPCHK(fl->ctx, hipMemcpyAsync(m.data(), fl->mask, (size_t)fl->d.N * 4, hipMemcpyDeviceToHost, fl->ctx->stream));
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  int64_t k = 0;
  for (uint32_t i = 0; i < fl->d.N; ++i) {
    if (m[i] & kMaskSolid) continue;
    for (int q = 1; q < fl->Q; ++q)
      if ((m[i] >> q) & 1u) {
        if (cells && k < capacity) {
          cells[k] = i;
          qs[k] = (uint8_t)q;
        }
        ++k;
      }
  }
  *count = k;
  if (cells && k > capacity) return fail(fl->ctx, LBM_ERR_INVALID, "capacity %lld < %lld links", (long long)capacity,
                                         (long long)k);
  return LBM_OK;
}

int lbm_set_params(lbm_fields* fl, const lbm_step_params* p) {
  if (!fl || !p) return LBM_ERR_INVALID;
  if (p->collision < 0 || p->collision > 2) return fail(fl->ctx, LBM_ERR_INVALID, "collision %d", p->collision);
  if (!(p->omega_p > 0.0 && p->omega_p < 2.0) || !(p->omega_m > 0.0 && p->omega_m < 2.0))
    return fail(fl->ctx, LBM_ERR_INVALID, "relaxation rates must lie in (0, 2)");
  fl->params = *p;
  fl->hp = build_hp(*p);
  fl->collision = p->collision;
  fl->params_set = true;
  fl->gen++;
  return LBM_OK;
}

int lbm_set_omega_field(lbm_fields* fl, const float* omega) {
  if (!fl) return LBM_ERR_INVALID;
  const size_t n = fl->d.N;
  if (!omega) {
    HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
    dev_free(fl, fl->omega, n * 4);
    fl->omega = nullptr;
    fl->wale_on = false;  // the per-step WALE field goes with it
    fl->gen++;
    return LBM_OK;
  }
  if (!fl->omega) {
    int rc = dev_alloc(fl, (void**)&fl->omega, n * 4);
    if (rc) return rc;
    fl->gen++;
  }
  HIPCHK(fl->ctx, hipMemcpyAsync(fl->omega, omega, n * 4, hipMemcpyHostToDevice, fl->ctx->stream));
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  return LBM_OK;
}

int lbm_stage_in(lbm_fields* fl, const float* f) {
  if (!fl || !f) return LBM_ERR_INVALID;
  const size_t bytes = (size_t)fl->Q * fl->d.N * 4;
  float* tmp = nullptr;
  int rc = dev_alloc(fl, (void**)&tmp, bytes);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, hipMemcpyAsync(tmp, f, bytes, hipMemcpyHostToDevice, s));
  HIPCHK(fl->ctx, launch_stage(fl->Q, fl->storage, 0, 0, fl->d, fl->f, tmp, fl->mask, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, tmp, bytes);
  fl->parity = 0;
  return refresh_moments(fl);
}

int lbm_stage_out(lbm_fields* fl, float* f) {
  if (!fl || !f) return LBM_ERR_INVALID;
  const size_t bytes = (size_t)fl->Q * fl->d.N * 4;
  float* tmp = nullptr;
  int rc = dev_alloc(fl, (void**)&tmp, bytes);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, launch_stage(fl->Q, fl->storage, 1, fl->parity, fl->d, fl->f, tmp, fl->mask, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(f, tmp, bytes, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, tmp, bytes);
  return LBM_OK;
}

int lbm_init_with(lbm_fields* fl, const float* rho, const float* u) {
  if (!fl || !rho || !u) return LBM_ERR_INVALID;
  const size_t n = fl->d.N;
  float *dr = nullptr, *du = nullptr;
  int rc = dev_alloc(fl, (void**)&dr, n * 4);
  if (!rc) rc = dev_alloc(fl, (void**)&du, 3 * n * 4);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, hipMemcpyAsync(dr, rho, n * 4, hipMemcpyHostToDevice, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(du, u, 3 * n * 4, hipMemcpyHostToDevice, s));
  // tau = T::r(3 nu + 1/2) with nu recovered from omega_p (solver.rs:2478)
  const double nu = (1.0 / fl->params.omega_p - 0.5) / 3.0;
  const float tau = (float)(3.0 * nu + 0.5);
  HIPCHK(fl->ctx, launch_init_with(fl->Q, fl->storage, fl->d, fl->f, fl->mask, dr, du, tau, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, dr, n * 4);
  dev_free(fl, du, 3 * n * 4);
  fl->parity = 0;
  return refresh_moments(fl);
}

int lbm_update_moments(lbm_fields* fl) {
  if (!fl) return LBM_ERR_INVALID;
  return refresh_moments(fl);
}

int lbm_set_force_field(lbm_fields* fl, const float* ff) {
  if (!fl) return LBM_ERR_INVALID;
  if (!ff) return lbm_clear_force_field(fl);
  int rc = ensure_force_arrays(fl);
  if (rc) return rc;
  const size_t n = fl->d.N;
  std::vector<uint32_t> extra;
  bool uniform = true;
  for (size_t i = 0; i < n; ++i) {
    const float* v = ff + 3 * i;
    if (v[0] != 0.0f || v[1] != 0.0f || v[2] != 0.0f) extra.push_back((uint32_t)i);
    if (uniform && (v[0] != ff[0] || v[1] != ff[1] || v[2] != ff[2])) uniform = false;
  }
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, hipMemcpyAsync(fl->F_cur, ff, 3 * n * 4, hipMemcpyHostToDevice, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  rc = grow_force_list(fl, extra);
  if (rc) return rc;
  fl->field_present = true;
  fl->field_uniform = uniform;
  if (uniform) return refresh_moments(fl);
  return LBM_OK;
}

int lbm_clear_force_field(lbm_fields* fl) {
  if (!fl) return LBM_ERR_INVALID;
  if (!fl->field_present) return LBM_OK;
  const bool refresh = fl->field_uniform;
  if (!fl->flist_host.empty())
    HIPCHK(fl->ctx, launch_copy_force_list(fl->flist, (uint32_t)fl->flist_host.size(), fl->F_cur, nullptr,
                                           fl->ctx->stream));
  fl->field_present = false;
  fl->field_uniform = false;
  if (refresh) return refresh_moments(fl);
  return LBM_OK;
}

int lbm_read_force_field(lbm_fields* fl, float* ff, int* present) {
  if (!fl || !ff) return LBM_ERR_INVALID;
  if (present) *present = fl->field_present ? 1 : 0;
  if (!fl->field_present || !fl->F_cur) {
    std::fill(ff, ff + N3(fl), 0.0f);
    return LBM_OK;
  }
  HIPCHK(fl->ctx, hipMemcpyAsync(ff, fl->F_cur, N3(fl) * 4, hipMemcpyDeviceToHost, fl->ctx->stream));
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  return LBM_OK;
}

// ---- WALE LES ------------------------------------------------------------
// Laminar viscosity and the f32 quantities WaleLes::update derives from it
// (les.rs:136-147): nu0, tau0 = 1/2 + 3 nu0, nu_t_max = (tau_eff_max -
// tau0)/3.
static double wale_nu(const lbm_fields* fl) { return (1.0 / fl->params.omega_p - 0.5) / 3.0; }

static int wale_coeffs(lbm_fields* fl, double cw, double tau_eff_max, float* cw2, float* nut_max) {
  if (!fl->params_set) return fail(fl->ctx, LBM_ERR_STATE, "set the step parameters before WALE");
  if (!(cw > 0.0) || !std::isfinite(cw)) return fail(fl->ctx, LBM_ERR_INVALID, "WALE coefficient %g", cw);
  const float c = (float)cw;
  *cw2 = c * c;
  *nut_max = -1.0f;
  if (tau_eff_max > 0.0) {
    // set_tau_eff_max, les.rs:82-95, and the laminar bound of update, les.rs:141-147
    if (!std::isfinite(tau_eff_max) || !(tau_eff_max > 0.5))
      return fail(fl->ctx, LBM_ERR_INVALID, "tau_eff_max must be finite and greater than 0.5");
    const float nu0 = (float)wale_nu(fl);
    const float tau0 = 0.5f + 3.0f * nu0;
    const float t = (float)tau_eff_max;
    if (!(t >= tau0))
      return fail(fl->ctx, LBM_ERR_INVALID, "tau_eff_max %g is below the laminar tau %g", tau_eff_max, (double)tau0);
    *nut_max = (t - tau0) / 3.0f;
  }
  return LBM_OK;
}

static int wale_buffers(lbm_fields* fl) {
  const size_t n = fl->d.N;
  if (!fl->omega) {
    int rc = dev_alloc(fl, (void**)&fl->omega, n * 4);
    if (rc) return rc;
    fl->gen++;
  }
  if (!fl->wale_mom) {
    int rc = dev_alloc(fl, (void**)&fl->wale_mom, 6 * n * 4);
    if (rc) return rc;
  }
  if (!fl->wale_stats) {
    int rc = dev_alloc(fl, (void**)&fl->wale_stats, 2 * sizeof(uint32_t));
    if (rc) return rc;
  }
  return LBM_OK;
}

// Cached velocity and strain of the current state, then the omega_plus
// field; launches only (graph-capturable: the statistics words are cleared
// by a kernel).
static int wale_enqueue(lbm_fields* fl, float cw2, float nut_max, float* nut, uint32_t* stats) {
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, launch_wale(fl->Q, fl->storage, fl->parity, fl->d, fl->f, fl->mask, fl->wall_u,
                              fl->field_present ? fl->F_cur : nullptr, fl->F_mom, fl->wale_mom, fl->hp,
                              (float)(1.0 / fl->params.omega_p),
                              cw2, (float)wale_nu(fl), nut_max, fl->omega, nut, stats, s));
  return LBM_OK;
}

int lbm_wale_update(lbm_fields* fl, double cw, double tau_eff_max, float* nu_t, lbm_wale_diag* diag) {
  if (!fl) return LBM_ERR_INVALID;
  float cw2, nut_max;
  int rc = wale_coeffs(fl, cw, tau_eff_max, &cw2, &nut_max);
  if (!rc) rc = wale_buffers(fl);
  if (rc) return rc;
  const size_t n = fl->d.N;
  float* dnut = nullptr;
  if (nu_t) {
    rc = dev_alloc(fl, (void**)&dnut, n * 4);
    if (rc) return rc;
  }
  hipStream_t s = fl->ctx->stream;
  rc = wale_enqueue(fl, cw2, nut_max, dnut, fl->wale_stats);
  uint32_t st[2] = {0, 0};
  if (!rc) {
    HIPCHK(fl->ctx, hipMemcpyAsync(st, fl->wale_stats, sizeof st, hipMemcpyDeviceToHost, s));
    if (nu_t) HIPCHK(fl->ctx, hipMemcpyAsync(nu_t, dnut, n * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(fl->ctx, hipStreamSynchronize(s));
  }
  dev_free(fl, dnut, n * 4);
  if (rc) return rc;
  if (diag) {
    diag->clipped_cells = st[0];
    diag->clipped_fraction = n ? (double)((float)st[0] / (float)n) : 0.0;
    float mx;
    memcpy(&mx, &st[1], sizeof mx);
    diag->max_nu_t_before_clipping = mx;
    diag->tau_eff_max = nut_max >= 0.0f ? (double)(float)tau_eff_max : 0.0;
    diag->nu_t_max = nut_max >= 0.0f ? (double)nut_max : 0.0;
  }
  return LBM_OK;
}

int lbm_gather_omega_field(lbm_fields* fl, float* omega, int* present) {
  if (!fl || !omega || !present) return LBM_ERR_INVALID;
  *present = fl->omega ? 1 : 0;
  if (!fl->omega) return LBM_OK;
  HIPCHK(fl->ctx, hipMemcpyAsync(omega, fl->omega, (size_t)fl->d.N * 4, hipMemcpyDeviceToHost, fl->ctx->stream));
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  return LBM_OK;
}

int lbm_set_wale(lbm_fields* fl, int enabled, double cw, double tau_eff_max) {
  if (!fl) return LBM_ERR_INVALID;
  if (!enabled) {
    if (fl->wale_on) fl->gen++;
    fl->wale_on = false;
    return LBM_OK;
  }
  float cw2, nut_max;
  int rc = wale_coeffs(fl, cw, tau_eff_max, &cw2, &nut_max);
  if (!rc) rc = wale_buffers(fl);
  if (rc) return rc;
  fl->wale_on = true;
  fl->wale_cw2 = cw2;
  fl->wale_nut_max = nut_max;
  fl->gen++;
  return LBM_OK;
}

// ---- force probe -------------------------------------------------------------
static void probe_release(lbm_fields* fl) {
  for (auto& pa : fl->pr_allocs) dev_free(fl, pa.first, pa.second);
  fl->pr_allocs.clear();
  fl->pr = ProbeArgs{};
  fl->probe_built = false;
}

extern "C++" template <typename T>
static int probe_upload(lbm_fields* fl, T** dst, const std::vector<T>& v) {
  const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
  int rc = dev_alloc(fl, (void**)dst, bytes);
  if (rc) return rc;
  fl->pr_allocs.push_back({(void*)*dst, bytes});
  if (!v.empty())
    HIPCHK(fl->ctx, hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, fl->ctx->stream));
  return LBM_OK;
}

// Links (i, q): fluid i, probed solid i - c_q, sources beyond a non-periodic
// face skipped as the stream skips them (kernels.rs:572-590); ordered by
// (row, x, q), the reference's per-row summation order.
static int probe_build(lbm_fields* fl) {
  probe_release(fl);
  const Dom& d = fl->d;
  const size_t n = d.N;
  std::vector<uint8_t> solid(n);
  HIPCHK(fl->ctx, hipMemcpyAsync(solid.data(), fl->solid, n, hipMemcpyDeviceToHost, fl->ctx->stream));
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  std::vector<uint64_t> keys;  // (cell << 8) | q, cell order = (row, x)
  const long ext[3] = {(long)d.nx, (long)d.ny, (long)d.nz};
  for (size_t sidx = 0; sidx < n; ++sidx) {
    if (!fl->probe_host[sidx] || !solid[sidx]) continue;
    const long sx = (long)(sidx % d.nx), t = (long)(sidx / d.nx);
    const long sy = t % (long)d.ny, sz = t / (long)d.ny;
    for (int q = 1; q < fl->Q; ++q) {
      long p[3] = {sx + lat_c(q, 0), sy + lat_c(q, 1), sz + lat_c(q, 2)};
      bool ok = true;
      for (int a = 0; a < 3 && ok; ++a)
        if (p[a] < 0 || p[a] >= ext[a]) {
          if (d.per[a])
            p[a] = (p[a] + ext[a]) % ext[a];
          else
            ok = false;
        }
      if (!ok) continue;
      const size_t i = ((size_t)p[2] * d.ny + p[1]) * d.nx + p[0];
      if (solid[i]) continue;
      keys.push_back(((uint64_t)i << 8) | (uint64_t)q);
    }
  }
  std::sort(keys.begin(), keys.end());
  std::vector<uint32_t> cells, lcell, row_start;
  std::vector<uint8_t> lq;
  uint32_t last_row = UINT32_MAX;
  for (uint64_t k : keys) {
    const uint32_t i = (uint32_t)(k >> 8);
    if (cells.empty() || cells.back() != i) cells.push_back(i);
    const uint32_t row = i / d.nx;
    if (row != last_row) {
      row_start.push_back((uint32_t)lcell.size());
      last_row = row;
    }
    lcell.push_back((uint32_t)cells.size() - 1);
    lq.push_back((uint8_t)(k & 0xff));
  }
  const uint32_t nrows = (uint32_t)row_start.size();
  row_start.push_back((uint32_t)lcell.size());
  ProbeArgs a{};
  uint32_t *dc = nullptr, *dl = nullptr, *dr = nullptr;
  uint8_t* dq = nullptr;
  float *drho = nullptr, *dpart = nullptr, *dout = nullptr;
  int rc = probe_upload(fl, &dc, cells);
  if (!rc) rc = probe_upload(fl, &dl, lcell);
  if (!rc) rc = probe_upload(fl, &dq, lq);
  if (!rc) rc = probe_upload(fl, &dr, row_start);
  if (!rc) rc = probe_upload(fl, &drho, std::vector<float>(cells.size(), 0.0f));
  if (!rc) rc = probe_upload(fl, &dpart, std::vector<float>(3 * (size_t)nrows, 0.0f));
  if (!rc) rc = probe_upload(fl, &dout, std::vector<float>(3, 0.0f));
  if (rc) return rc;
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  a.cells = dc;
  a.ncells = (uint32_t)cells.size();
  a.rho = drho;
  a.lcell = dl;
  a.lq = dq;
  a.row_start = dr;
  a.nrows = nrows;
  a.partial = dpart;
  a.out = dout;
  fl->pr = a;
  fl->probe_built = true;
  fl->gen++;
  return LBM_OK;
}

int lbm_set_volume_sources(lbm_fields* fl, int n, const lbm_volume_source* src) {
  if (!fl || n < 0 || (n > 0 && !src)) return LBM_ERR_INVALID;
  std::vector<SourceArgs> out;
  const uint32_t ext[3] = {fl->d.nx, fl->d.ny, fl->d.nz};
  for (int k = 0; k < n; ++k) {
    const lbm_volume_source& v = src[k];
    if (v.kind != LBM_SOURCE_MASS_FLOW && v.kind != LBM_SOURCE_JET)
      return fail(fl->ctx, LBM_ERR_INVALID, "volume source %d: unknown kind %d", k, v.kind);
    SourceArgs a{};
    for (int c = 0; c < 3; ++c) {
      if (v.lo[c] < 0 || v.hi[c] < v.lo[c] || (uint32_t)v.hi[c] >= ext[c])
        return fail(fl->ctx, LBM_ERR_INVALID, "volume source %d: region outside the grid", k);
      a.lo[c] = (uint32_t)v.lo[c];
      a.hi[c] = (uint32_t)v.hi[c];
      a.u[c] = v.kind == LBM_SOURCE_JET ? (float)v.u[c] : 0.0f;
    }
    if (!std::isfinite(v.q_lu)) return fail(fl->ctx, LBM_ERR_INVALID, "volume source %d: q_lu", k);
    a.q_lu = (float)v.q_lu;
    out.push_back(a);
  }
  fl->sources = out;
  fl->gen++;
  return LBM_OK;
}

int lbm_set_force_probe(lbm_fields* fl, const uint8_t* mask) {
  if (!fl) return LBM_ERR_INVALID;
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  probe_release(fl);
  fl->gen++;
  if (!mask) {
    fl->probe_host.clear();
    return LBM_OK;
  }
  fl->probe_host.assign(mask, mask + fl->d.N);
  return probe_build(fl);
}

int lbm_probed_force(lbm_fields* fl, double out[3]) {
  if (!fl || !out) return LBM_ERR_INVALID;
  out[0] = out[1] = out[2] = 0.0;
  if (fl->probe_host.empty() || !fl->pr.out) return LBM_OK;
  float v[3];
  HIPCHK(fl->ctx, hipMemcpyAsync(v, fl->pr.out, sizeof v, hipMemcpyDeviceToHost, fl->ctx->stream));
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  for (int c = 0; c < 3; ++c) out[c] = v[c];
  return LBM_OK;
}

static int probe_launch(lbm_fields* fl, int post) {
  ProbeArgs a = fl->pr;
  a.d = fl->d;
  a.f = fl->f;
  a.mask = fl->mask;
  a.odd = post ? fl->parity ^ 1 : fl->parity;
  a.wall_u = fl->wall_u;
  HIPCHK(fl->ctx, launch_probe(fl->Q, fl->storage, post, a, fl->ctx->stream));
  return LBM_OK;
}

static int check_steppable(lbm_fields* fl) {
  if (fl->uncovered)
    return fail(fl->ctx, LBM_ERR_UNSUPPORTED,
                "%u fluid cells lie on a non-periodic domain face with no open-face rule "
                "(set one with lbm_set_face_bc, close the face with solid rims or make the axis periodic)",
                fl->uncovered);
  if (fl->open_face_cells && !fl->scalars.empty())
    return fail(fl->ctx, LBM_ERR_UNSUPPORTED, "scalar distributions do not take open faces on this path");
  if (!fl->probe_host.empty() && !fl->probe_built) return probe_build(fl);
  return LBM_OK;
}

// ---- open faces ------------------------------------------------------------
// validate_face_bc / validate_velocity (solver.rs:745-780) and the open-axis
// extent rule (solver.rs:566-575) for one rule on `face`
static int face_rule_check(lbm_fields* fl, int face, int kind, const double* value) {
  if (kind == LBM_FACE_VELOCITY) {
    double sq = 0.0;
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(value[a])) return fail(fl->ctx, LBM_ERR_INVALID, "face velocity must be finite");
      sq += value[a] * value[a];
    }
    if (!(std::sqrt(sq) <= kMaxSpeed))
      return fail(fl->ctx, LBM_ERR_INVALID, "prescribed speed %g exceeds the low-Mach limit %g", std::sqrt(sq),
                  kMaxSpeed);
  }
  if (kind == LBM_FACE_PRESSURE && !(value[0] > 0.0 && std::isfinite(value[0])))
    return fail(fl->ctx, LBM_ERR_INVALID, "face density must be finite and positive");
  if (kind == LBM_FACE_CONVECTIVE && !(value[0] >= 0.0 && std::isfinite(value[0])))
    return fail(fl->ctx, LBM_ERR_INVALID, "convective speed must be finite and non-negative");
  if (kind != LBM_FACE_CLOSED && fl->Q != 19 && fl->Q != 27) return LBM_ERR_UNSUPPORTED;
  const uint32_t ext[3] = {fl->d.nx, fl->d.ny, fl->d.nz};
  if (kind != LBM_FACE_CLOSED && ext[face / 2] < 3)
    return fail(fl->ctx, LBM_ERR_INVALID, "an open face needs at least 3 cells along its axis");
  return LBM_OK;
}

int lbm_set_face_bc(lbm_fields* fl, int face, int kind, const double value[3]) {
  if (!fl || face < 0 || face > 5 || kind < LBM_FACE_CLOSED || kind > LBM_FACE_CONVECTIVE)
    return LBM_ERR_INVALID;
  const int ax = face / 2;
  if (kind != LBM_FACE_CLOSED && fl->d.per[ax])
    return fail(fl->ctx, LBM_ERR_INVALID, "face %d is on a periodic axis and cannot be open", face);
  if (kind != LBM_FACE_CLOSED && kind != LBM_FACE_OUTFLOW && !value) return LBM_ERR_INVALID;
  int rc = face_rule_check(fl, face, kind, value);
  if (rc) return rc;
  if (kind != LBM_FACE_CLOSED) {
    rc = ensure_stash(fl, face);
    if (rc) return rc;
  }
  fl->face_kind[face] = kind;
  for (int a = 0; a < 3; ++a) fl->face_u[face][a] = kind == LBM_FACE_VELOCITY ? (float)value[a] : 0.0f;
  fl->face_rho[face] = kind == LBM_FACE_PRESSURE ? (float)value[0] : 0.0f;
  fl->face_uconv[face] = kind == LBM_FACE_CONVECTIVE ? (float)value[0] : 0.0f;
  fl->gen++;
  return count_uncovered(fl);
}

// FacePatch list (params.rs:286-296) with validate_face_patches
// (solver.rs:690-740): face in range, rect inside the face and not inverted,
// rule parameters in range, no two rects of one face overlapping; a patch on
// a periodic axis is refused.
int lbm_set_face_patches(lbm_fields* fl, int n, const lbm_face_patch* patches) {
  if (!fl || n < 0 || (n && !patches)) return LBM_ERR_INVALID;
  if (n > kMaxFacePatches) return fail(fl->ctx, LBM_ERR_INVALID, "at most %d face patches", kMaxFacePatches);
  const uint32_t ext[3] = {fl->d.nx, fl->d.ny, fl->d.nz};
  std::vector<FacePatchArgs> list;
  for (int k = 0; k < n; ++k) {
    const lbm_face_patch& p = patches[k];
    if (p.face < 0 || p.face > 5 || p.kind < LBM_FACE_CLOSED || p.kind > LBM_FACE_CONVECTIVE)
      return fail(fl->ctx, LBM_ERR_INVALID, "patch %d: bad face or kind", k);
    const int ax = p.face / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    if (fl->d.per[ax]) return fail(fl->ctx, LBM_ERR_INVALID, "patch %d lies on a periodic axis", k);
    if (p.lo[0] < 0 || p.lo[1] < 0 || p.lo[0] > p.hi[0] || p.lo[1] > p.hi[1] || (uint32_t)p.hi[0] >= ext[t1] ||
        (uint32_t)p.hi[1] >= ext[t2])
      return fail(fl->ctx, LBM_ERR_INVALID, "patch %d: rect outside face %d", k, p.face);
    int rc = face_rule_check(fl, p.face, p.kind, p.value);
    if (rc) return rc;
    if (fl->Q != 19 && fl->Q != 27) return LBM_ERR_UNSUPPORTED;
    for (int j = 0; j < k; ++j) {
      const lbm_face_patch& o = patches[j];
      if (o.face == p.face && p.lo[0] <= o.hi[0] && o.lo[0] <= p.hi[0] && p.lo[1] <= o.hi[1] && o.lo[1] <= p.hi[1])
        return fail(fl->ctx, LBM_ERR_INVALID, "patches %d and %d overlap on face %d", j, k, p.face);
    }
    FacePatchArgs a{};
    a.face = p.face;
    for (int c = 0; c < 2; ++c) {
      a.lo[c] = (uint32_t)p.lo[c];
      a.hi[c] = (uint32_t)p.hi[c];
    }
    a.kind = p.kind == LBM_FACE_CLOSED ? kFaceVelocity : p.kind;  // a closed patch: a resting lid
    for (int c = 0; c < 3; ++c) a.u[c] = p.kind == LBM_FACE_VELOCITY ? (float)p.value[c] : 0.0f;
    a.rho = p.kind == LBM_FACE_PRESSURE ? (float)p.value[0] : 0.0f;
    a.uconv = p.kind == LBM_FACE_CONVECTIVE ? (float)p.value[0] : 0.0f;
    list.push_back(a);
  }
  for (const FacePatchArgs& a : list) {
    int rc = ensure_stash(fl, a.face);
    if (rc) return rc;
  }
  fl->patches = list;
  for (int k = 0; k < 6; ++k) fl->patched[k] = 0;
  for (const FacePatchArgs& a : list) fl->patched[a.face] = 1;
  fl->gen++;
  return count_uncovered(fl);
}

// Solver::set_inlet_profile (solver.rs:3322-3350): u[3] per cell of a
// velocity face in face order (c2 * n1 + c1); NULL clears.
int lbm_set_inlet_profile(lbm_fields* fl, int face, const double* values) {
  if (!fl || face < 0 || face > 5) return LBM_ERR_INVALID;
  if (fl->face_kind[face] != LBM_FACE_VELOCITY)
    return fail(fl->ctx, LBM_ERR_INVALID, "set_inlet_profile: face %d is not a velocity face", face);
  const uint32_t ext[3] = {fl->d.nx, fl->d.ny, fl->d.nz};
  const int ax = face / 2;
  const size_t nf = (size_t)ext[ax == 0 ? 1 : 0] * ext[ax == 2 ? 1 : 2];
  if (!values) {
    dev_free(fl, fl->prof[face], 3 * nf * sizeof(float));
    fl->prof[face] = nullptr;
    fl->gen++;
    return LBM_OK;
  }
  std::vector<float> h(3 * nf);
  for (size_t k = 0; k < 3 * nf; ++k) {
    if (!std::isfinite(values[k])) return fail(fl->ctx, LBM_ERR_INVALID, "inlet profile value %zu is not finite", k);
    h[k] = (float)values[k];
  }
  if (!fl->prof[face]) {
    int rc = dev_alloc(fl, (void**)&fl->prof[face], 3 * nf * sizeof(float));
    if (rc) return rc;
  }
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, hipMemcpyAsync(fl->prof[face], h.data(), 3 * nf * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  fl->gen++;
  return LBM_OK;
}

static int open_faces_step(lbm_fields* fl) {
  FaceArgs a{};
  a.d = fl->d;
  a.f = fl->f;
  a.mask = fl->mask;
  a.odd = fl->parity ^ 1;  // the state the step just produced
  a.p = fl->hp;
  a.F_cur = fl->field_present ? fl->F_cur : nullptr;
  for (int k = 0; k < 6; ++k) {
    if (!fl->face_kind[k] && !fl->patched[k]) continue;
    a.faces[a.nfaces++] = k;
    // a closed face with patches: zero-velocity Zou-He outside them
    a.kind[k] = fl->face_kind[k] ? fl->face_kind[k] : kFaceVelocity;
    for (int c = 0; c < 3; ++c) a.u[k][c] = fl->face_u[k][c];
    a.rho[k] = fl->face_rho[k];
    a.uconv[k] = fl->face_uconv[k];
    a.prof[k] = fl->face_kind[k] == LBM_FACE_VELOCITY ? fl->prof[k] : nullptr;
    a.stash_cur[k] = fl->stash[k][fl->stash_sel];
    a.stash_next[k] = fl->stash[k][fl->stash_sel ^ 1];
  }
  if (!a.nfaces) return LBM_OK;
  a.npatch = (int)fl->patches.size();
  for (int p = 0; p < a.npatch; ++p) a.patch[p] = fl->patches[p];
  HIPCHK(fl->ctx, launch_open_faces(fl->Q, fl->storage, a, fl->ctx->stream));
  fl->stash_sel ^= 1;
  return LBM_OK;
}

// One collide-stream step (main kernel, forced-cell list kernel) plus the
// scalar distributions; launches only.
static int hydro_step(lbm_fields* fl) {
  hipStream_t s = fl->ctx->stream;
  if (fl->wale_on) {
    int rc = wale_enqueue(fl, fl->wale_cw2, fl->wale_nut_max, nullptr, nullptr);
    if (rc) return rc;
  }
  if (fl->probe_built) {
    int rc = probe_launch(fl, 0);
    if (rc) return rc;
  }
  HydroArgs a = hydro_args(fl);
  if (fl->timing) {
    if (fl->ev_used + 2 > fl->ev_pool.size()) {
      for (int k2 = 0; k2 < 2; ++k2) {
        hipEvent_t e;
        HIPCHK(fl->ctx, hipEventCreate(&e));
        fl->ev_pool.push_back(e);
      }
    }
    HIPCHK(fl->ctx, hipEventRecord(fl->ev_pool[fl->ev_used], s));
  }
  HIPCHK(fl->ctx, launch_hydro(fl->Q, fl->storage, fl->collision, fl->parity, a, s));
  if (fl->timing) {
    HIPCHK(fl->ctx, hipEventRecord(fl->ev_pool[fl->ev_used + 1], s));
    fl->ev_used += 2;
  }
  if (a.nlist) HIPCHK(fl->ctx, launch_hydro_list(fl->Q, fl->storage, fl->collision, fl->parity, a, s));
  if (fl->probe_built) {  // before the open faces, as the stream accumulates it
    int rc = probe_launch(fl, 1);
    if (rc) return rc;
  }
  if (fl->open_face_cells) {
    int rc = open_faces_step(fl);
    if (rc) return rc;
  }
  for (const SourceArgs& src : fl->sources) {
    SourceArgs sa = src;
    sa.d = fl->d;
    sa.f = fl->f;
    sa.mask = fl->mask;
    sa.odd = fl->parity ^ 1;
    HIPCHK(fl->ctx, launch_volume_source(fl->Q, fl->storage, sa, s));
  }
  fl->parity ^= 1;
  fl->time += 1;
  for (ScalarState* sc : fl->scalars) {
    int rc = scalar_step(view_of(fl), sc);
    if (rc) return rc;
  }
  return LBM_OK;
}

int lbm_run_span(lbm_fields* fl, int steps) {
  if (!fl || steps < 0) return LBM_ERR_INVALID;
  int rc = check_steppable(fl);
  if (rc) return rc;
  for (int k = 0; k < steps; ++k) {
    rc = hydro_step(fl);
    if (rc) return rc;
  }
  return LBM_OK;
}

// The runner's step (runner.rs:1079-1129): clear the force field, rotating
// IBM, collide-stream, scalars.  Launches only; the IBM diagnostics go to the
// device ring.
static int ibm_step(lbm_fields* fl) {
  // clear the field at its listed cells (in the gather's launch), then the IBM
  fl->field_present = false;
  fl->field_uniform = false;
  int rc = ibm_enqueue(view_of(fl), fl->ibm, fl->ring, fl->ring_cap, fl->ring_pos,
                       fl->flist_host.empty() ? nullptr : fl->flist, (uint32_t)fl->flist_host.size());
  if (rc) return rc;
  fl->field_present = true;
  return hydro_step(fl);
}

int lbm_run_ibm_span(lbm_fields* fl, int steps, int use_graph, double* torque, lbm_ibm_diag* last) {
  if (!fl || steps < 0) return LBM_ERR_INVALID;
  if (!fl->ibm) return fail(fl->ctx, LBM_ERR_STATE, "no IBM body registered");
  int rc = check_steppable(fl);
  if (rc) return rc;
  if (steps == 0) return LBM_OK;
  if (fl->field_present && fl->field_uniform) {  // the clear must refresh the cached moments
    rc = lbm_clear_force_field(fl);
    if (rc) return rc;
  }
  const IbmDomain D = ibm_domain_of(view_of(fl));
  rc = ibm_set_domains(view_of(fl), fl->ibm, &D);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  if (fl->ring_cap < (uint32_t)steps) {
    // the captured steps write the old ring: retire them before it goes
    HIPCHK(fl->ctx, drop_graph(s, &fl->gexec, &fl->graph));
    if (fl->ring) dev_free(fl, fl->ring, (size_t)fl->ring_cap * ibm_ring_entry_bytes());
    fl->ring = nullptr;
    fl->ring_cap = 0;
    rc = dev_alloc(fl, &fl->ring, (size_t)steps * ibm_ring_entry_bytes());
    if (rc) return rc;
    fl->ring_cap = (uint32_t)steps;
    fl->gen++;
  }
  if (!fl->ring_pos) {
    rc = dev_alloc(fl, (void**)&fl->ring_pos, 256);
    if (rc) return rc;
  }
  // a fill kernel, not a memset, on the stream the graph runs on
  HIPCHK(fl->ctx, launch_fill_f32((float*)fl->ring_pos, 2, 0.0f, s));  // slot, blocks done
  int k = 0;
  if (use_graph && !fl->timing && steps >= 2) {
    const HydroArgs a = hydro_args(fl);
    const bool stale = !fl->gexec || fl->g_parity != fl->parity || fl->g_gen != fl->gen ||
                       fl->g_stash != fl->stash_sel || memcmp(&a, &fl->g_args, sizeof a) != 0;
    if (stale) {
      HIPCHK(fl->ctx, drop_graph(s, &fl->gexec, &fl->graph));
      const int parity0 = fl->parity;
      const int64_t time0 = fl->time;
      HIPCHK(fl->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc1 = ibm_step(fl);
      if (!rc1) rc1 = ibm_step(fl);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      // capture records the launches without running them: two steps flip
      // the parity (and the scalar buffers) back, the clock does not move
      fl->time = time0;
      if (rc1) {
        if (g) (void)hipGraphDestroy(g);
        return rc1;
      }
      HIPCHK(fl->ctx, e);
      if (fl->parity != parity0) return fail(fl->ctx, LBM_ERR_STATE, "captured step pair changed parity");
      fl->graph = g;
      HIPCHK(fl->ctx, hipGraphInstantiateWithFlags(&fl->gexec, g, 0));
      fl->g_args = a;
      fl->g_parity = fl->parity;
      fl->g_gen = fl->gen;
      fl->g_stash = fl->stash_sel;
    }
    for (; k + 2 <= steps; k += 2) {
      HIPCHK(fl->ctx, hipGraphLaunch(fl->gexec, s));
      fl->time += 2;
    }
    fl->field_present = true;
    fl->field_uniform = false;
  }
  for (; k < steps; ++k) {
    rc = ibm_step(fl);
    if (rc) return rc;
  }
  std::vector<lbm_ibm_diag> d(steps);
  rc = ibm_ring_read(view_of(fl), fl->ring, (uint32_t)steps, d.data());
  if (rc) return rc;
  if (torque)
    for (int j = 0; j < steps; ++j)
      for (int c = 0; c < 3; ++c) torque[3 * (size_t)j + c] = d[j].torque[c];
  if (last) *last = d[steps - 1];
  return LBM_OK;
}

int lbm_kernel_timing(lbm_fields* fl, int enable) {
  if (!fl) return LBM_ERR_INVALID;
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  fl->timing = enable != 0;
  fl->ev_used = 0;
  return LBM_OK;
}

int lbm_kernel_timing_read(lbm_fields* fl, double* total_ms, int64_t* launches) {
  if (!fl || !total_ms || !launches) return LBM_ERR_INVALID;
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  double acc = 0.0;
  for (size_t k = 0; k + 1 < fl->ev_used; k += 2) {
    float ms = 0.0f;
    HIPCHK(fl->ctx, hipEventElapsedTime(&ms, fl->ev_pool[k], fl->ev_pool[k + 1]));
    acc += ms;
  }
  *total_ms = acc;
  *launches = (int64_t)(fl->ev_used / 2);
  fl->ev_used = 0;
  return LBM_OK;
}

int64_t lbm_force_list_size(const lbm_fields* fl) { return fl ? (int64_t)fl->flist_host.size() : -1; }

int lbm_synchronize(lbm_fields* fl) {
  if (!fl) return LBM_ERR_INVALID;
  HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  return LBM_OK;
}

int64_t lbm_time(const lbm_fields* fl) { return fl ? fl->time : -1; }

int lbm_read_moments(lbm_fields* fl, float* rho, float* ux, float* uy, float* uz) {
  if (!fl || !rho || !ux || !uy || !uz) return LBM_ERR_INVALID;
  const size_t n = fl->d.N;
  float* buf = nullptr;
  int rc = dev_alloc(fl, (void**)&buf, 4 * n * 4);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, launch_moments(fl->Q, fl->storage, fl->parity, fl->d, fl->f, fl->mask, fl->F_mom, fl->hp, buf,
                                 buf + n, buf + 2 * n, buf + 3 * n, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(rho, buf, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(ux, buf + n, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(uy, buf + 2 * n, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(uz, buf + 3 * n, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, buf, 4 * n * 4);
  return LBM_OK;
}

int lbm_read_moments_noalloc(lbm_fields* fl, float* rho, float* ux, float* uy, float* uz) {
  if (!fl || !rho || !ux || !uy || !uz) return LBM_ERR_INVALID;
  const size_t n = fl->d.N;
  if (n > (size_t)LBM_NOALLOC_MAX_CELLS)
    return fail(fl->ctx, LBM_ERR_INVALID, "lbm_read_moments_noalloc: %zu cells exceed the staging maximum of %d cells",
                n, LBM_NOALLOC_MAX_CELLS);
  hipError_t se = hipSuccess;
  float* buf = scratch_of(fl->ctx->device, &se);
  if (!buf)
    return fail(fl->ctx, LBM_ERR_DEVICE, "lbm_read_moments_noalloc: no staging address for device %d: %s",
                fl->ctx->device, hipGetErrorString(se));
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, launch_moments(fl->Q, fl->storage, fl->parity, fl->d, fl->f, fl->mask, fl->F_mom, fl->hp, buf,
                                 buf + n, buf + 2 * n, buf + 3 * n, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(rho, buf, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(ux, buf + n, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(uy, buf + 2 * n, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(uz, buf + 3 * n, n * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  return LBM_OK;
}

int lbm_reduce(lbm_fields* fl, int kind, int axis, double* out) {
  if (!fl || !out || kind < 0 || kind > 2 || axis < 0 || axis > 2) return LBM_ERR_INVALID;
  hipStream_t s = fl->ctx->stream;
  const uint32_t np = std::min<uint32_t>(fl->npartials, (fl->d.N + 255) / 256);
  HIPCHK(fl->ctx, launch_reduce(fl->Q, fl->storage, fl->parity, fl->d, fl->f, fl->mask,
                                fl->field_present ? fl->F_cur : nullptr, fl->hp, kind, axis, fl->partials, np, s));
  std::vector<double> h(np);
  HIPCHK(fl->ctx, hipMemcpyAsync(h.data(), fl->partials, np * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  double acc = 0.0;
  for (double v : h) acc += v;
  *out = acc;
  return LBM_OK;
}

int lbm_geom_wall_rims(lbm_fields* fl, const int is_wall[6], const double u[18]) {
  if (!fl || !is_wall || !u) return LBM_ERR_INVALID;
  const size_t n = fl->d.N;
  bool moving = false;
  for (int fc = 0; fc < 6; ++fc)
    if (is_wall[fc] && (u[3 * fc] != 0.0 || u[3 * fc + 1] != 0.0 || u[3 * fc + 2] != 0.0)) moving = true;
  hipStream_t s = fl->ctx->stream;
  if (moving && !fl->wall_u) {
    int rc = dev_alloc(fl, (void**)&fl->wall_u, 3 * n * 4);
    if (rc) return rc;
    HIPCHK(fl->ctx, hipMemsetAsync(fl->wall_u, 0, 3 * n * 4, s));
  }
  float* wu = fl->wall_u;
  float* scratch_wu = nullptr;
  if (!wu) {  // resting walls: velocities are all zero, write into a throwaway buffer
    int rc = dev_alloc(fl, (void**)&scratch_wu, 3 * n * 4);
    if (rc) return rc;
    wu = scratch_wu;
  }
  double* best = nullptr;
  int rc = dev_alloc(fl, (void**)&best, n * 8);
  if (rc) return rc;
  HIPCHK(fl->ctx, launch_rims(fl->d, is_wall, u, fl->solid, wu, best, s));
  HIPCHK(fl->ctx, hipStreamSynchronize(s));
  dev_free(fl, best, n * 8);
  if (scratch_wu) dev_free(fl, scratch_wu, 3 * n * 4);
  rc = rebuild_mask(fl);
  if (rc) return rc;
  return refresh_moments(fl);
}

int lbm_geom_stirred_tank(lbm_fields* fl, double dx, double vessel_diameter, double liquid_height, int baffle_count,
                          double baffle_width, double baffle_thickness) {
  if (!fl) return LBM_ERR_INVALID;
  if (!(dx > 0.0) || !(vessel_diameter > 0.0) || !(liquid_height > 0.0))
    return fail(fl->ctx, LBM_ERR_INVALID, "tank dimensions must be finite and positive");
  if (fl->d.nx < 3 || fl->d.ny < 3 || fl->d.nz < 3)
    return fail(fl->ctx, LBM_ERR_INVALID, "stirred-tank grid dimensions must each be >= 3");
  if (baffle_count != 0 && baffle_count != 4)
    return fail(fl->ctx, LBM_ERR_UNSUPPORTED, "custom baffle placement is not implemented; use count = 4");
  HIPCHK(fl->ctx, launch_tank(fl->d, dx, vessel_diameter, liquid_height, baffle_count, baffle_width,
                              baffle_thickness, fl->solid, fl->ctx->stream));
  int rc = rebuild_mask(fl);
  if (rc) return rc;
  return refresh_moments(fl);
}

int lbm_geom_radial_impeller(const int dims[3], double dx, double vessel_diameter, double liquid_height,
                             double diameter, double clearance, int blades, double blade_angle_deg, double omega_z,
                             double thickness_fraction, double* pos, double* weight, double* velocity,
                             double* center, int64_t capacity, int64_t* count) {
  if (!dims || !count) return LBM_ERR_INVALID;
  if (!(blade_angle_deg >= 0.0 && blade_angle_deg <= 90.0) || blades <= 0 || !(diameter > 0.0) || !(dx > 0.0))
    return LBM_ERR_INVALID;
  // build_radial_impeller, geometry.rs:815-906
  const double TAU = 6.283185307179586;
  const double cxy[2] = {0.5 * (double)dims[0] * dx, 0.5 * (double)dims[1] * dx};
  const double radius = 0.5 * diameter;
  const double ctr[3] = {cxy[0], cxy[1], clearance};
  const double radial_inner = blade_angle_deg == 0.0 ? 0.65 * radius : 0.35 * radius;
  const double dr = std::max(dx, radius / 20.0);
  const double dz = dx, tstep = dx;
  const double thickness = std::max(thickness_fraction * diameter, dx);
  int64_t k = 0;
  for (int b = 0; b < blades; ++b) {
    const double th = TAU * (double)b / (double)blades;
    const double er[2] = {std::cos(th), std::sin(th)};
    const double et[2] = {-std::sin(th), std::cos(th)};
    const double pitch = blade_angle_deg * (3.141592653589793 / 180.0);
    const double z_span =
        blade_angle_deg == 0.0 ? thickness : thickness + (radius - radial_inner) * std::sin(pitch) * 0.18;
    const long nr = (long)std::max(std::ceil((radius - radial_inner) / dr), 1.0);
    const long nt = (long)std::max(std::ceil(thickness / tstep), 1.0);
    const long nzs = (long)std::max(std::ceil(z_span / dz), 1.0);
    for (long ir = 0; ir <= nr; ++ir) {
      const double rr = radial_inner + (radius - radial_inner) * (double)ir / (double)nr;
      for (long it = 0; it <= nt; ++it) {
        const double ot = -0.5 * thickness + thickness * (double)it / (double)nt;
        for (long iz = 0; iz <= nzs; ++iz) {
          const double oz = -0.5 * z_span + z_span * (double)iz / (double)nzs;
          const double pz = oz + (rr - radial_inner) * std::sin(pitch) * 0.18;
          const double p[3] = {ctr[0] + rr * er[0] + ot * et[0], ctr[1] + rr * er[1] + ot * et[1], ctr[2] + pz};
          const double rel = std::sqrt((p[0] - cxy[0]) * (p[0] - cxy[0]) + (p[1] - cxy[1]) * (p[1] - cxy[1]));
          if (rel >= 0.5 * vessel_diameter || p[2] <= 0.0 || p[2] >= liquid_height) continue;
          if (pos && k < capacity) {
            const double r3[3] = {p[0] - ctr[0], p[1] - ctr[1], p[2] - ctr[2]};
            const double om[3] = {0.0, 0.0, omega_z};
            velocity[3 * k + 0] = om[1] * r3[2] - om[2] * r3[1];
            velocity[3 * k + 1] = om[2] * r3[0] - om[0] * r3[2];
            velocity[3 * k + 2] = om[0] * r3[1] - om[1] * r3[0];
            pos[3 * k + 0] = p[0] / dx;
            pos[3 * k + 1] = p[1] / dx;
            pos[3 * k + 2] = p[2] / dx;
            weight[k] = dx * dx;
          }
          ++k;
        }
      }
    }
  }
  if (center) {
    center[0] = ctr[0] / dx;
    center[1] = ctr[1] / dx;
    center[2] = ctr[2] / dx;
  }
  *count = k;
  return k == 0 ? LBM_ERR_INVALID : LBM_OK;
}

int lbm_geom_impeller_surface(const int dims[3], double dx, double vessel_diameter, double liquid_height,
                              double diameter, double clearance, int blades, double blade_angle_deg, double omega_z,
                              double thickness_fraction, double area_per_marker, uint64_t seed, double* pos,
                              double* weight, double* velocity, double* center, int64_t capacity, int64_t* count) {
  if (!dims || !count) return LBM_ERR_INVALID;
  if (!(blade_angle_deg >= 0.0 && blade_angle_deg <= 90.0) || blades <= 0 || !(diameter > 0.0) || !(dx > 0.0) ||
      !(area_per_marker > 0.0) || !std::isfinite(area_per_marker))
    return LBM_ERR_INVALID;
  // The blade solids of build_radial_impeller (geometry.rs:815-906): per
  // blade a box in the blade frame, radial [r_in, R], tangential [-t/2, t/2],
  // axial [-z_span/2, z_span/2], sheared axially by (r - r_in) sin(pitch) 0.18.
  // Markers are drawn at uniform areal density over the six faces of each
  // box (in the unsheared frame): a face with probability proportional to its
  // area, then a uniform point on it; N_b = max(1, round(A_b / area)) per
  // blade, each marker weighted A_b / N_b (m^2).  Three SplitMix64 draws per
  // marker, one stream from `seed`, blades in order.  Markers outside the
  // liquid are dropped as in the reference generator.
  const double TAU = 6.283185307179586;
  const double cxy[2] = {0.5 * (double)dims[0] * dx, 0.5 * (double)dims[1] * dx};
  const double radius = 0.5 * diameter;
  const double ctr[3] = {cxy[0], cxy[1], clearance};
  const double radial_inner = blade_angle_deg == 0.0 ? 0.65 * radius : 0.35 * radius;
  const double thickness = std::max(thickness_fraction * diameter, dx);
  const double pitch = blade_angle_deg * (3.141592653589793 / 180.0);
  const double z_span =
      blade_angle_deg == 0.0 ? thickness : thickness + (radius - radial_inner) * std::sin(pitch) * 0.18;
  const double L = radius - radial_inner;
  // faces: r_in, r_out (t x z), t-, t+ (L x z), z-, z+ (L x t)
  const double fa[6] = {thickness * z_span, thickness * z_span, L * z_span, L * z_span, L * thickness,
                        L * thickness};
  double cum[6];
  double ab = 0.0;
  for (int f = 0; f < 6; ++f) {
    ab += fa[f];
    cum[f] = ab;
  }
  const double area_lu = area_per_marker * dx * dx;
  const int64_t nb = std::max<int64_t>(1, (int64_t)std::llround(ab / area_lu));
  const double w = ab / (double)nb;
  uint64_t state = seed;
  auto uniform = [&state]() {  // SplitMix64, top 53 bits
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;
  };
  int64_t k = 0;
  for (int b = 0; b < blades; ++b) {
    const double th = TAU * (double)b / (double)blades;
    const double er[2] = {std::cos(th), std::sin(th)};
    const double et[2] = {-std::sin(th), std::cos(th)};
    for (int64_t j = 0; j < nb; ++j) {
      const double u1 = uniform() * ab, u2 = uniform(), u3 = uniform();
      int f = 0;
      while (f < 5 && u1 >= cum[f]) ++f;
      double rr, ot, oz;
      if (f < 2) {
        rr = f == 0 ? radial_inner : radius;
        ot = -0.5 * thickness + thickness * u2;
        oz = -0.5 * z_span + z_span * u3;
      } else if (f < 4) {
        ot = f == 2 ? -0.5 * thickness : 0.5 * thickness;
        rr = radial_inner + L * u2;
        oz = -0.5 * z_span + z_span * u3;
      } else {
        oz = f == 4 ? -0.5 * z_span : 0.5 * z_span;
        rr = radial_inner + L * u2;
        ot = -0.5 * thickness + thickness * u3;
      }
      const double pz = oz + (rr - radial_inner) * std::sin(pitch) * 0.18;
      const double p[3] = {ctr[0] + rr * er[0] + ot * et[0], ctr[1] + rr * er[1] + ot * et[1], ctr[2] + pz};
      const double rel = std::sqrt((p[0] - cxy[0]) * (p[0] - cxy[0]) + (p[1] - cxy[1]) * (p[1] - cxy[1]));
      if (rel >= 0.5 * vessel_diameter || p[2] <= 0.0 || p[2] >= liquid_height) continue;
      if (pos && k < capacity) {
        const double r3[3] = {p[0] - ctr[0], p[1] - ctr[1], p[2] - ctr[2]};
        const double om[3] = {0.0, 0.0, omega_z};
        velocity[3 * k + 0] = om[1] * r3[2] - om[2] * r3[1];
        velocity[3 * k + 1] = om[2] * r3[0] - om[0] * r3[2];
        velocity[3 * k + 2] = om[0] * r3[1] - om[1] * r3[0];
        pos[3 * k + 0] = p[0] / dx;
        pos[3 * k + 1] = p[1] / dx;
        pos[3 * k + 2] = p[2] / dx;
        weight[k] = w;
      }
      ++k;
    }
  }
  if (center) {
    center[0] = ctr[0] / dx;
    center[1] = ctr[1] / dx;
    center[2] = ctr[2] / dx;
  }
  *count = k;
  return k == 0 ? LBM_ERR_INVALID : LBM_OK;
}

// ---- rotating IBM -----------------------------------------------------------
int lbm_ibm_set_body(lbm_fields* fl, int64_t n, const double* pos, const double* weight, const double center[3],
                     const double omega[3], const lbm_ibm_config* cfg) {
  if (!fl || n <= 0 || !pos || !weight || !center || !omega) return LBM_ERR_INVALID;
  lbm_ibm_config c{3, 1e-2, 1, 1.0};
  if (cfg) c = *cfg;
  if (c.max_iterations <= 0 || !(c.slip_tolerance >= 0.0) || !(c.relaxation > 0.0 && c.relaxation <= 1.0) ||
      (c.kernel_radius != 1 && c.kernel_radius != 2))
    return fail(fl->ctx, LBM_ERR_INVALID, "invalid direct-forcing configuration");
  if (fl->ibm) {
    ibm_destroy(fl->ibm);
    fl->ibm = nullptr;
  }
  std::vector<uint32_t> cells;
  std::string err;
  IbmState* st = ibm_create(view_of(fl), n, pos, weight, center, omega, c, &cells, &err);
  if (!st) return fail(fl->ctx, LBM_ERR_INVALID, "%s", err.c_str());
  fl->ibm = st;
  fl->gen++;
  int rc = ensure_force_arrays(fl);
  if (rc) return rc;
  return grow_force_list(fl, cells);
}

int lbm_ibm_torque_integral(lbm_fields* fl, int reset, double out[4]) {
  if (!fl || !out) return LBM_ERR_INVALID;
  if (!fl->ibm) return fail(fl->ctx, LBM_ERR_STATE, "no IBM body registered");
  return ibm_torque_integral(view_of(fl), fl->ibm, reset, out);
}

int lbm_ibm_apply(lbm_fields* fl, lbm_ibm_diag* diag) {
  if (!fl || !diag) return LBM_ERR_INVALID;
  if (!fl->ibm) return fail(fl->ctx, LBM_ERR_STATE, "no IBM body registered");
  const IbmDomain D = ibm_domain_of(view_of(fl));
  int rc = ibm_set_domains(view_of(fl), fl->ibm, &D);
  if (rc) return rc;
  rc = ibm_apply(view_of(fl), fl->ibm, diag);
  if (rc) return rc;
  fl->field_present = true;
  fl->field_uniform = false;
  return LBM_OK;
}

// ---- scalar ADE -------------------------------------------------------------
int lbm_scalar_enable(lbm_fields* fl, double diffusivity, int lattice, int* index) {
  if (!fl || !index) return LBM_ERR_INVALID;
  if (!std::isfinite(diffusivity) || !(diffusivity > 0.0))
    return fail(fl->ctx, LBM_ERR_INVALID, "diffusivity must be finite > 0");
  if (lattice != LBM_SCALAR_HYDRO_LATTICE && lattice != LBM_SCALAR_D3Q7)
    return fail(fl->ctx, LBM_ERR_INVALID, "scalar lattice %d", lattice);
  std::string err;
  ScalarState* s = scalar_create(view_of(fl), diffusivity, lattice == LBM_SCALAR_D3Q7 ? 7 : fl->Q, &err);
  if (!s) return fail(fl->ctx, LBM_ERR_NOMEM, "%s", err.c_str());
  fl->scalars.push_back(s);
  fl->gen++;
  *index = (int)fl->scalars.size() - 1;
  return LBM_OK;
}

int lbm_scalar_set(lbm_fields* fl, int index, const float* conc) {
  if (!fl || !conc || index < 0 || index >= (int)fl->scalars.size()) return LBM_ERR_INVALID;
  return scalar_set(view_of(fl), fl->scalars[index], conc);
}

int lbm_scalar_gather(lbm_fields* fl, int index, float* conc) {
  if (!fl || !conc || index < 0 || index >= (int)fl->scalars.size()) return LBM_ERR_INVALID;
  return scalar_gather(view_of(fl), fl->scalars[index], conc);
}

int lbm_scalar_cv(lbm_fields* fl, int index, double* cv, int* defined) {
  if (!fl || !cv || !defined || index < 0 || index >= (int)fl->scalars.size()) return LBM_ERR_INVALID;
  return scalar_cv(view_of(fl), fl->scalars[index], cv, defined);
}

int lbm_scalar_compartment_cv(lbm_fields* fl, int index, int has_impeller_z, double impeller_z, double cv[3],
                              int defined[3], int64_t cell_count[3]) {
  if (!fl || !cv || !defined || !cell_count || index < 0 || index >= (int)fl->scalars.size())
    return LBM_ERR_INVALID;
  if (has_impeller_z && !std::isfinite(impeller_z))
    return fail(fl->ctx, LBM_ERR_INVALID, "impeller height must be finite");
  return scalar_compartment_cv(view_of(fl), fl->scalars[index], has_impeller_z ? 1 : 0, impeller_z, cv, defined,
                               cell_count);
}

// ---- QOIs -------------------------------------------------------------------
int lbm_qoi_power(double torque, double omega, double rho, double diameter, double volume, double out[4]) {
  // power_qois, qoi.rs:477-505
  if (!out) return LBM_ERR_INVALID;
  if (!(std::isfinite(omega) && std::isfinite(rho) && std::isfinite(diameter) && std::isfinite(volume) && rho > 0.0 &&
        diameter > 0.0 && volume > 0.0))
    return LBM_ERR_INVALID;
  const double n_hz = omega / 6.283185307179586;
  if (!(n_hz > 0.0)) return LBM_ERR_INVALID;
  const double power = omega * torque;
  const double denom = rho * (n_hz * n_hz * n_hz) * (diameter * diameter * diameter * diameter * diameter);
  out[0] = power;
  out[1] = n_hz;
  out[2] = power / denom;
  out[3] = power / volume;
  return LBM_OK;
}

int lbm_qoi_mixing_time(int64_t n, const double* t, const double* cv, double out[3]) {
  // mixing_time_from_cv, qoi.rs:534-556
  if (!out || n < 0 || (n > 0 && (!t || !cv))) return LBM_ERR_INVALID;
  int64_t first = -1;
  for (int64_t i = 0; i < n; ++i)
    if (std::isfinite(cv[i]) && cv[i] > 0.0) {
      first = i;
      break;
    }
  if (first < 0) return LBM_ERR_STATE;
  const double cv0 = cv[first];
  double t95 = NAN, t99 = NAN;
  bool h95 = false, h99 = false;
  for (int64_t i = 0; i < n; ++i) {
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
  return LBM_OK;
}

}  // extern "C"

// ---- batched domains ---------------------------------------------------------
// K independent stirred tanks stepped together (lbm-cli's sweep runs its
// cases one after another through the runner, crates/lbm-cli/src/sweep.rs:
// 61-106; here they share every launch).  The members keep their own
// populations, masks, parameters and IBM bodies; the batch owns only the
// concatenated tile and force lists, the device argument table, the joined
// IBM state and the captured step graph.
struct lbm_batch {
  lbm_ctx* ctx = nullptr;
  std::vector<lbm_fields*> mem;
  std::vector<uint64_t> mem_gen;
  int Q = 19, storage = kF32, collision = LBM_COLLISION_TRT;
  HydroArgs* tab = nullptr;
  uint32_t* ent = nullptr;  // 2 x nent: (tile entry, member)
  uint32_t nent = 0;
  uint32_t* cells = nullptr;  // 2 x ncells: (forced cell, member)
  uint32_t ncells = 0;
  IbmState* ibm = nullptr;
  void* ring = nullptr;
  uint32_t ring_cap = 0;
  uint32_t* ring_pos = nullptr;
  uint32_t grid = 0;  // persistent blocks of the collide-stream launch
  int ncu = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_parity = -1;
  uint32_t g_grid = 0;
};

namespace {

FieldsView batch_view(lbm_batch* b) {
  FieldsView v = view_of(b->mem[0]);
  v.ctx_err = &b->ctx->err;
  v.field_present = false;
  return v;
}

template <typename T>
int batch_upload(lbm_batch* b, T** dst, const std::vector<T>& src) {
  const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
  HIPCHK(b->ctx, hipMalloc((void**)dst, bytes));
  if (!src.empty())
    HIPCHK(b->ctx, hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, b->ctx->stream));
  HIPCHK(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return LBM_OK;
}

int batch_check(lbm_batch* b) {
  for (size_t k = 0; k < b->mem.size(); ++k)
    if (b->mem[k]->gen != b->mem_gen[k] || b->mem[k]->parity != b->mem[0]->parity)
      return fail(b->ctx, LBM_ERR_STATE, "batch member %zu changed after the batch was made; make a new batch", k);
  return LBM_OK;
}

// one collide-stream step of every member (launches only)
int batch_hydro(lbm_batch* b, const HydroArgs* tab, uint32_t grid) {
  hipStream_t s = b->ctx->stream;
  const int odd = b->mem[0]->parity;
  HIPCHK(b->ctx, launch_hydro_batch(b->Q, b->storage, b->collision, odd, tab, b->ent, b->nent, grid, s));
  HIPCHK(b->ctx, launch_hydro_list_batch(b->Q, b->storage, b->collision, odd, tab, b->cells, b->ncells, s));
  for (lbm_fields* fl : b->mem) {
    fl->parity ^= 1;
    fl->time += 1;
  }
  return LBM_OK;
}

// the runner's step for every member: clear, rotating IBM, collide-stream
int batch_step(lbm_batch* b) {
  hipStream_t s = b->ctx->stream;
  if (b->ibm) {
    HIPCHK(b->ctx, launch_clear_batch(b->tab, b->cells, b->ncells, s));
    int rc = ibm_enqueue(batch_view(b), b->ibm, b->ring, b->ring_cap, b->ring_pos);
    if (rc) return rc;
  }
  return batch_hydro(b, b->tab, b->grid);
}

}  // namespace

extern "C" {

int lbm_batch_create(lbm_ctx* ctx, lbm_fields* const* members, int n, lbm_batch** out) {
  if (!ctx || !members || n <= 0 || !out) return LBM_ERR_INVALID;
  *out = nullptr;
  lbm_fields* f0 = members[0];
  if (!f0) return LBM_ERR_INVALID;
  int with_ibm = -1;
  for (int k = 0; k < n; ++k) {
    lbm_fields* fl = members[k];
    if (!fl || fl->ctx != ctx) return fail(ctx, LBM_ERR_INVALID, "batch member %d is null or on another context", k);
    for (int j = 0; j < k; ++j)
      if (members[j] == fl) return fail(ctx, LBM_ERR_INVALID, "batch member %d appears twice", k);
    if (fl->Q != f0->Q || fl->storage != f0->storage || fl->collision != f0->collision)
      return fail(ctx, LBM_ERR_INVALID, "batch members must share lattice, storage and collision kind");
    if (fl->d.per[0] || fl->d.per[1] || fl->d.per[2])
      return fail(ctx, LBM_ERR_UNSUPPORTED, "batch member %d has a periodic axis; batches hold closed domains", k);
    if (fl->open_face_cells) return fail(ctx, LBM_ERR_UNSUPPORTED, "batch member %d has fluid on a face", k);
    if (!fl->scalars.empty())
      return fail(ctx, LBM_ERR_UNSUPPORTED, "batch member %d carries scalar distributions", k);
    if (fl->parity != f0->parity) return fail(ctx, LBM_ERR_STATE, "batch members must be at the same step parity");
    if (!fl->probe_host.empty())
      return fail(ctx, LBM_ERR_UNSUPPORTED, "batch member %d has a force probe", k);
    if (fl->hp.force_on || fl->omega)
      return fail(ctx, LBM_ERR_UNSUPPORTED, "batch member %d has a body force or an omega field", k);
    if (fl->field_present && fl->field_uniform && !fl->ibm)
      return fail(ctx, LBM_ERR_UNSUPPORTED, "batch member %d has a uniform force field", k);
    const int has = fl->ibm ? 1 : 0;
    if (with_ibm >= 0 && has != with_ibm)
      return fail(ctx, LBM_ERR_INVALID, "either every batch member has a rotating body or none");
    with_ibm = has;
  }
  lbm_batch* b = new (std::nothrow) lbm_batch();
  if (!b) return LBM_ERR_NOMEM;
  b->ctx = ctx;
  b->Q = f0->Q;
  b->storage = f0->storage;
  b->collision = f0->collision;
  int rc = LBM_OK;
  for (int k = 0; k < n && !rc; ++k) {
    lbm_fields* fl = members[k];
    b->mem.push_back(fl);
    // a uniform clear must refresh moments; do it before the members freeze
    if (fl->field_present && fl->field_uniform) rc = lbm_clear_force_field(fl);
  }
  std::vector<HydroArgs> tab;
  std::vector<uint32_t> ent, cells;
  for (int k = 0; k < n && !rc; ++k) {
    lbm_fields* fl = members[k];
    tab.push_back(hydro_args(fl));
    for (uint32_t te : fl->tiles_host) {
      ent.push_back(te);
      ent.push_back((uint32_t)k);
    }
    for (uint32_t c : fl->flist_host) {
      cells.push_back(c);
      cells.push_back((uint32_t)k);
    }
  }
  if (!rc) rc = batch_upload(b, &b->tab, tab);
  if (!rc) rc = batch_upload(b, &b->ent, ent);
  if (!rc) rc = batch_upload(b, &b->cells, cells);
  b->nent = (uint32_t)(ent.size() / 2);
  b->ncells = (uint32_t)(cells.size() / 2);
  if (!rc && with_ibm == 1) {
    std::vector<const IbmState*> parts;
    std::vector<IbmDomain> doms;
    for (lbm_fields* fl : b->mem) {
      parts.push_back(fl->ibm);
      doms.push_back(ibm_domain_of(view_of(fl)));
    }
    std::string err;
    b->ibm = ibm_concat(parts, &err);
    if (!b->ibm) rc = fail(ctx, LBM_ERR_INVALID, "%s", err.c_str());
    if (!rc) rc = ibm_set_domains(batch_view(b), b->ibm, doms.data());
  }
  if (!rc) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) prop.multiProcessorCount = 256;
    b->ncu = prop.multiProcessorCount;
    // default: every entry's wave resident at once, capped at 8 blocks per CU
    b->grid = std::max<uint32_t>(1, std::min<uint32_t>((b->nent + 3) / 4, (uint32_t)b->ncu * 8));
  }
  for (lbm_fields* fl : b->mem) b->mem_gen.push_back(fl->gen);
  if (rc) {
    lbm_batch_destroy(b);
    return rc;
  }
  *out = b;
  return LBM_OK;
}

void lbm_batch_destroy(lbm_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->ibm) ibm_destroy(b->ibm);
  for (void* p : {(void*)b->tab, (void*)b->ent, (void*)b->cells, b->ring, (void*)b->ring_pos})
    if (p) (void)hipFree(p);
  delete b;
}

int lbm_batch_info(const lbm_batch* b, int* members, uint32_t* tiles, uint32_t* forced_cells, uint32_t* grid) {
  if (!b) return LBM_ERR_INVALID;
  if (members) *members = (int)b->mem.size();
  if (tiles) *tiles = b->nent;
  if (forced_cells) *forced_cells = b->ncells;
  if (grid) *grid = b->grid;
  return LBM_OK;
}

int lbm_batch_set_grid(lbm_batch* b, uint32_t grid) {
  if (!b || grid == 0) return LBM_ERR_INVALID;
  b->grid = grid;
  return LBM_OK;
}

int lbm_batch_run_span(lbm_batch* b, int steps, int use_graph, double* torque) {
  if (!b || steps < 0) return LBM_ERR_INVALID;
  int rc = batch_check(b);
  if (rc) return rc;
  if (steps == 0) return LBM_OK;
  const size_t K = b->mem.size();
  hipStream_t s = b->ctx->stream;
  if (b->ibm && b->ring_cap < (uint32_t)steps) {
    // the captured steps write the old ring: retire them before it goes
    HIPCHK(b->ctx, drop_graph(s, &b->gexec, &b->graph));
    if (b->ring) (void)hipFree(b->ring);
    b->ring = nullptr;
    b->ring_cap = 0;
    HIPCHK(b->ctx, hipMalloc(&b->ring, (size_t)steps * K * ibm_ring_entry_bytes()));
    b->ring_cap = (uint32_t)steps;
  }
  if (b->ibm && !b->ring_pos) HIPCHK(b->ctx, hipMalloc((void**)&b->ring_pos, 256));
  if (b->ibm) HIPCHK(b->ctx, launch_fill_f32((float*)b->ring_pos, 2, 0.0f, s));  // slot, blocks done
  int k = 0;
  if (use_graph && steps >= 2) {
    const int parity = b->mem[0]->parity;
    if (!b->gexec || b->g_parity != parity || b->g_grid != b->grid) {
      HIPCHK(b->ctx, drop_graph(s, &b->gexec, &b->graph));
      std::vector<int64_t> t0;
      for (lbm_fields* fl : b->mem) t0.push_back(fl->time);
      HIPCHK(b->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc1 = batch_step(b);
      if (!rc1) rc1 = batch_step(b);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      for (size_t j = 0; j < K; ++j) b->mem[j]->time = t0[j];  // capture ran nothing
      if (rc1) {
        if (g) (void)hipGraphDestroy(g);
        return rc1;
      }
      HIPCHK(b->ctx, e);
      b->graph = g;
      HIPCHK(b->ctx, hipGraphInstantiateWithFlags(&b->gexec, g, 0));
      b->g_parity = parity;
      b->g_grid = b->grid;
    }
    for (; k + 2 <= steps; k += 2) {
      HIPCHK(b->ctx, hipGraphLaunch(b->gexec, s));
      for (lbm_fields* fl : b->mem) fl->time += 2;
    }
  }
  for (; k < steps; ++k) {
    rc = batch_step(b);
    if (rc) return rc;
  }
  for (lbm_fields* fl : b->mem) {
    if (b->ibm) {
      fl->field_present = true;
      fl->field_uniform = false;
    }
  }
  if (b->ibm) {
    std::vector<lbm_ibm_diag> d((size_t)steps * K);
    rc = ibm_ring_read(batch_view(b), b->ring, (uint32_t)(steps * K), d.data());
    if (rc) return rc;
    if (torque)
      for (size_t j = 0; j < K; ++j)
        for (int t = 0; t < steps; ++t)
          for (int c = 0; c < 3; ++c) torque[(j * steps + t) * 3 + c] = d[(size_t)t * K + j].torque[c];
  } else {
    HIPCHK(b->ctx, hipStreamSynchronize(s));
  }
  return LBM_OK;
}

int lbm_autotune_tiles(lbm_fields* fl, int trial_steps, uint32_t zb_out[2], double ms_out[2]) {
  if (!fl || trial_steps < 2 || !zb_out) return LBM_ERR_INVALID;
  int rc = check_steppable(fl);
  if (rc) return rc;
  if (fl->ntiles == 0) return fail(fl->ctx, LBM_ERR_STATE, "no active tiles to time");
  hipStream_t s = fl->ctx->stream;
  // trial steps run on a copy of the populations: the state is untouched
  void* copy = nullptr;
  if (hipMalloc(&copy, fl->f_bytes) != hipSuccess) return fail(fl->ctx, LBM_ERR_NOMEM, "autotune copy");
  hipEvent_t ev[2 * 64 + 1] = {};
  const int T = std::min(trial_steps, 64) & ~1;  // whole even/odd pairs
  hipError_t err = hipMemcpyAsync(copy, fl->f_alloc, fl->f_bytes, hipMemcpyDeviceToDevice, s);
  for (int k = 0; k <= T && err == hipSuccess; ++k) err = hipEventCreate(&ev[k]);
  static const uint32_t cand_even[] = {1, 2, 4, 8, 16, 32};
  static const uint32_t cand_odd[] = {1, 4, 8, 16, 32, 64};
  constexpr int NC = 6;
  double best_ms[2] = {1e300, 1e300};
  uint32_t best_zb[2] = {1, 1};
  const uint32_t keep[2] = {fl->zb_set[0], fl->zb_set[1]};
  const size_t off = (char*)fl->f - (char*)fl->f_alloc;
  for (int c = 0; c < NC && err == hipSuccess && !rc; ++c) {
    fl->zb_set[0] = cand_even[c];
    fl->zb_set[1] = cand_odd[c];
    rc = rebuild_tiles(fl);
    if (rc) break;
    HydroArgs a = hydro_args(fl);
    a.f = (char*)copy + off;
    const int p0 = fl->parity;
    double sum[2] = {0.0, 0.0};
    // one untimed pair, then T launches with an event between each
    for (int t = 0; t < 2 && err == hipSuccess; ++t) {
      const int p = (p0 + t) & 1;
      a.tiles = p ? fl->tiles_odd : fl->tiles;
      err = launch_hydro(fl->Q, fl->storage, fl->collision, p, a, s);
    }
    if (err == hipSuccess) err = hipEventRecord(ev[0], s);
    for (int t = 0; t < T && err == hipSuccess; ++t) {
      const int p = (p0 + t) & 1;
      a.tiles = p ? fl->tiles_odd : fl->tiles;
      err = launch_hydro(fl->Q, fl->storage, fl->collision, p, a, s);
      if (err == hipSuccess) err = hipEventRecord(ev[t + 1], s);
    }
    if (err == hipSuccess) err = hipEventSynchronize(ev[T]);
    for (int t = 0; t < T && err == hipSuccess; ++t) {
      float ms = 0.0f;
      err = hipEventElapsedTime(&ms, ev[t], ev[t + 1]);
      sum[(p0 + t) & 1] += ms;
    }
    if (err != hipSuccess) break;
    for (int p = 0; p < 2; ++p) {
      const double per = sum[p] / (T / 2);
      if (per < best_ms[p]) {
        best_ms[p] = per;
        best_zb[p] = p ? cand_odd[c] : cand_even[c];
      }
    }
  }
  (void)hipStreamSynchronize(s);
  for (int k = 0; k <= T; ++k)
    if (ev[k]) (void)hipEventDestroy(ev[k]);
  (void)hipFree(copy);
  if (err != hipSuccess && !rc) rc = fail(fl->ctx, LBM_ERR_DEVICE, "autotune trial: %s", hipGetErrorString(err));
  if (rc) {
    fl->zb_set[0] = keep[0];
    fl->zb_set[1] = keep[1];
    (void)rebuild_tiles(fl);
    fl->gen++;
    return rc;
  }
  fl->zb_set[0] = best_zb[0];
  fl->zb_set[1] = best_zb[1];
  rc = rebuild_tiles(fl);
  fl->gen++;
  if (rc) return rc;
  zb_out[0] = best_zb[0];
  zb_out[1] = best_zb[1];
  if (ms_out) {
    ms_out[0] = best_ms[0];
    ms_out[1] = best_ms[1];
  }
  return LBM_OK;
}

int lbm_batch_torque_integral(lbm_batch* b, int reset, double* out) {
  if (!b || !out) return LBM_ERR_INVALID;
  const size_t K = b->mem.size();
  if (!b->ibm) {
    std::fill(out, out + 4 * K, 0.0);
    return LBM_OK;
  }
  return ibm_torque_integral(batch_view(b), b->ibm, reset, out);
}

int lbm_batch_autotune(lbm_batch* b, int trial_steps, uint32_t* best_grid, double* best_ms) {
  if (!b || trial_steps < 2) return LBM_ERR_INVALID;
  int rc = batch_check(b);
  if (rc) return rc;
  hipStream_t s = b->ctx->stream;
  const size_t K = b->mem.size();
  // trial copies of the populations: the members' state is left untouched
  std::vector<void*> copies(K, nullptr);
  std::vector<HydroArgs> tab;
  auto release = [&]() {
    for (void* p : copies)
      if (p) (void)hipFree(p);
  };
  for (size_t j = 0; j < K; ++j) {
    lbm_fields* fl = b->mem[j];
    if (hipMalloc(&copies[j], fl->f_bytes) != hipSuccess) {
      release();
      return fail(b->ctx, LBM_ERR_NOMEM, "autotune copy of member %zu", j);
    }
    if (hipMemcpyAsync(copies[j], fl->f_alloc, fl->f_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
      release();
      return fail(b->ctx, LBM_ERR_DEVICE, "autotune copy of member %zu", j);
    }
    HydroArgs a = hydro_args(fl);
    a.f = (char*)copies[j] + ((char*)fl->f - (char*)fl->f_alloc);
    tab.push_back(a);
  }
  HydroArgs* dtab = nullptr;
  if (hipMalloc((void**)&dtab, K * sizeof(HydroArgs)) != hipSuccess ||
      hipMemcpyAsync(dtab, tab.data(), K * sizeof(HydroArgs), hipMemcpyHostToDevice, s) != hipSuccess) {
    release();
    if (dtab) (void)hipFree(dtab);
    return fail(b->ctx, LBM_ERR_NOMEM, "autotune table");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const uint32_t full = std::max<uint32_t>(1, (b->nent + 3) / 4);
  std::vector<uint32_t> cand;
  for (uint32_t per_cu : {1u, 2u, 4u, 8u, 16u}) cand.push_back(std::min(full, (uint32_t)b->ncu * per_cu));
  cand.push_back(full);
  std::sort(cand.begin(), cand.end());
  cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  const int parity0 = b->mem[0]->parity;
  double best = 1e300;
  uint32_t bg = b->grid;
  for (uint32_t g : cand) {
    hipError_t err = hipSuccess;
    for (int w = 0; w < 2 && err == hipSuccess; ++w)
      err = launch_hydro_batch(b->Q, b->storage, b->collision, w & 1, dtab, b->ent, b->nent, g, s);
    if (err == hipSuccess) err = hipEventRecord(e0, s);
    for (int t = 0; t < trial_steps && err == hipSuccess; ++t)
      err = launch_hydro_batch(b->Q, b->storage, b->collision, (parity0 + t) & 1, dtab, b->ent, b->nent, g, s);
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err != hipSuccess) {
      rc = fail(b->ctx, LBM_ERR_DEVICE, "autotune trial: %s", hipGetErrorString(err));
      break;
    }
    const double per = ms / trial_steps;
    if (per < best) {
      best = per;
      bg = g;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(dtab);
  release();
  if (rc) return rc;
  b->grid = bg;
  if (best_grid) *best_grid = bg;
  if (best_ms) *best_ms = best;
  return LBM_OK;
}

}  // extern "C"

// ---- adjoint of the time-averaged shaft torque --------------------------------
// No reference counterpart (the reference has only finite differences,
// crates/lbm-core/src/uq.rs:47-70).  Forward: `steps` runner steps from the
// current state; J = mean shaft torque (-torque_z of the IBM diagnostics,
// runner.rs:1396-1425) over the last `window` steps.  Reverse: discrete
// adjoint of every step (collide-stream with bounce-back, the stale-moment
// force, the direct-forcing sweeps) with a binomial checkpoint schedule.

namespace {

enum { ACT_ADVANCE = 0, ACT_SAVE = 1, ACT_RESTORE = 2, ACT_ADJOINT = 3 };

// C(s + r, s): the most steps s free snapshots can reverse with at most r
// forward recomputations of any step (Griewank's binomial bound)
double binom_reach(int s, int r) {
  double b = 1.0;
  for (int k = 1; k <= s; ++k) b = b * (double)(r + k) / (double)k;
  return b;
}

// Reverse steps [lo, hi); the state at lo is in `slot`, `free_slots` more
// slots (slot + 1, slot + 2, ...) may be used.  The first checkpoint goes
// m = n - C(s - 1 + r, s - 1) steps ahead (r the smallest repetition count
// with C(s + r, s) >= n), so every step is recomputed at most r times.
void revolve(int lo, int hi, int slot, int free_slots, std::vector<int>& out) {
  while (hi > lo) {
    const int n = hi - lo;
    if (n == 1) {
      out.insert(out.end(), {ACT_RESTORE, slot, ACT_ADJOINT, lo});
      return;
    }
    if (free_slots == 0) {
      for (int k = hi - 1; k >= lo; --k) {
        out.insert(out.end(), {ACT_RESTORE, slot});
        if (k > lo) out.insert(out.end(), {ACT_ADVANCE, k - lo});
        out.insert(out.end(), {ACT_ADJOINT, k});
      }
      return;
    }
    int r = 1;
    while (binom_reach(free_slots, r) < (double)n) ++r;
    const double right = binom_reach(free_slots - 1, r);
    const int m = right >= (double)(n - 1) ? 1 : n - (int)right;
    out.insert(out.end(), {ACT_RESTORE, slot, ACT_ADVANCE, m, ACT_SAVE, slot + 1});
    revolve(lo + m, hi, slot + 1, free_slots - 1, out);
    hi = lo + m;
  }
}

struct Snapshot {
  void* f = nullptr;
  float* F_mom = nullptr;
  int parity = 0;
  int64_t time = 0;
};

int snap_save(lbm_fields* fl, Snapshot& sn) {
  hipStream_t s = fl->ctx->stream;
  if (!sn.f) HIPCHK(fl->ctx, hipMalloc(&sn.f, fl->f_bytes));
  if (!sn.F_mom) HIPCHK(fl->ctx, hipMalloc((void**)&sn.F_mom, N3(fl) * 4));
  HIPCHK(fl->ctx, hipMemcpyAsync(sn.f, fl->f_alloc, fl->f_bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(sn.F_mom, fl->F_mom, N3(fl) * 4, hipMemcpyDeviceToDevice, s));
  sn.parity = fl->parity;
  sn.time = fl->time;
  return LBM_OK;
}

int snap_restore(lbm_fields* fl, const Snapshot& sn) {
  hipStream_t s = fl->ctx->stream;
  HIPCHK(fl->ctx, hipMemcpyAsync(fl->f_alloc, sn.f, fl->f_bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(fl->ctx, hipMemcpyAsync(fl->F_mom, sn.F_mom, N3(fl) * 4, hipMemcpyDeviceToDevice, s));
  fl->parity = sn.parity;
  fl->time = sn.time;
  return LBM_OK;
}

void snap_free(Snapshot& sn) {
  if (sn.f) (void)hipFree(sn.f);
  if (sn.F_mom) (void)hipFree(sn.F_mom);
  sn = Snapshot{};
}

}  // namespace

extern "C" {

int lbm_revolve_schedule(int steps, int snapshots, int* actions, int capacity, int* count) {
  if (steps < 1 || snapshots < 0 || !count) return LBM_ERR_INVALID;
  std::vector<int> out;
  revolve(0, steps, 0, snapshots, out);
  *count = (int)(out.size() / 2);
  if (actions) {
    if (capacity < *count) return LBM_ERR_INVALID;
    std::copy(out.begin(), out.end(), actions);
  }
  return LBM_OK;
}

int lbm_adjoint_torque(lbm_fields* fl, int steps, int window, int snapshots, double* J, double* dJ_domega,
                       double* dJ_dnu, int64_t* forward_steps) {
  if (!fl || steps < 1 || window < 1 || window > steps || snapshots < 0 || !J) return LBM_ERR_INVALID;
  if (!fl->ibm) return fail(fl->ctx, LBM_ERR_STATE, "the torque adjoint needs a rotating body");
  if (fl->Q != 19 || fl->collision == LBM_COLLISION_CM)
    return fail(fl->ctx, LBM_ERR_UNSUPPORTED, "the adjoint is implemented for D3Q19 TRT/BGK");
  if (fl->wall_u) return fail(fl->ctx, LBM_ERR_UNSUPPORTED, "the adjoint does not cover moving walls");
  if (fl->wale_on)
    return fail(fl->ctx, LBM_ERR_UNSUPPORTED, "the adjoint does not differentiate the per-step WALE closure");
  if (fl->open_face_cells)
    return fail(fl->ctx, LBM_ERR_UNSUPPORTED, "the adjoint covers closed domains (no open faces)");
  if (fl->hp.force_on) return fail(fl->ctx, LBM_ERR_UNSUPPORTED, "the adjoint does not cover a body force");
  int rc = check_steppable(fl);
  if (rc) return rc;
  hipStream_t s = fl->ctx->stream;
  const size_t N = fl->d.N, Q = (size_t)fl->Q;
  if (fl->field_present && fl->field_uniform) {
    rc = lbm_clear_force_field(fl);
    if (rc) return rc;
  }
  {
    const IbmDomain D = ibm_domain_of(view_of(fl));
    rc = ibm_set_domains(view_of(fl), fl->ibm, &D);
    if (rc) return rc;
  }
  std::vector<Snapshot> snap((size_t)snapshots + 1);
  Snapshot fin;
  double *lamA = nullptr, *lamB = nullptr, *obar = nullptr, *Fbar = nullptr, *G0 = nullptr, *G1 = nullptr,
         *wbar = nullptr;
  // a failure after the forward pass puts the final state back (best
  // effort) before every snapshot and adjoint buffer is freed
  bool reversing = false;
  auto cleanup = [&]() {
    const std::string err = fl->ctx->err;  // the failure being reported
    if (reversing && fin.f) (void)snap_restore(fl, fin);
    (void)hipStreamSynchronize(s);
    fl->ctx->err = err;
    for (Snapshot& sn : snap) snap_free(sn);
    snap_free(fin);
    for (double* p : {lamA, lamB, obar, Fbar, G0, G1, wbar})
      if (p) (void)hipFree(p);
  };
  auto hip_rc = [&](hipError_t e, const char* what) -> int {
    return e == hipSuccess ? LBM_OK : fail(fl->ctx, LBM_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  };
#define ADJ_TRY(expr)   \
  do {                  \
    int _r = (expr);    \
    if (_r) {           \
      cleanup();        \
      return _r;        \
    }                   \
  } while (0)
#define ADJ_HIP(expr) ADJ_TRY(hip_rc((expr), #expr))
  // forward pass: J
  ADJ_TRY(snap_save(fl, snap[0]));
  std::vector<double> tq(3 * (size_t)steps);
  ADJ_TRY(lbm_run_ibm_span(fl, steps, 0, tq.data(), nullptr));
  double acc = 0.0;
  for (int k = steps - window; k < steps; ++k) acc += -tq[3 * (size_t)k + 2];
  *J = acc / window;
  ADJ_TRY(snap_save(fl, fin));
  if (!dJ_domega && !dJ_dnu) {
    cleanup();
    return LBM_OK;
  }
  // reverse pass
  reversing = true;
  auto dalloc = [&](double** p, size_t n) -> int {
    if (hipMalloc((void**)p, n * sizeof(double)) != hipSuccess)
      return fail(fl->ctx, LBM_ERR_NOMEM, "adjoint buffer of %zu doubles", n);
    HIPCHK(fl->ctx, launch_fill_f64(*p, n, 0.0, s));
    return LBM_OK;
  };
  ADJ_TRY(dalloc(&lamA, Q * N));
  ADJ_TRY(dalloc(&lamB, Q * N));
  ADJ_TRY(dalloc(&obar, 1));
  ADJ_TRY(dalloc(&Fbar, 3 * N));
  ADJ_TRY(dalloc(&G0, 3 * N));
  ADJ_TRY(dalloc(&G1, 3 * N));
  ADJ_TRY(dalloc(&wbar, N));
  std::vector<int> plan;
  revolve(0, steps, 0, snapshots, plan);
  int64_t advanced = steps;
  const double tq_bar = 1.0 / window;
  // the snapshot slot the device state still equals (the schedule restores
  // the slot it has just saved; that copy back is skipped), -1 for none
  int held = -1;
  for (size_t a = 0; a < plan.size(); a += 2) {
    const int op = plan[a], arg = plan[a + 1];
    if (op == ACT_RESTORE) {
      if (arg != held) ADJ_TRY(snap_restore(fl, snap[arg]));
      held = arg;
    } else if (op == ACT_SAVE) {
      ADJ_TRY(snap_save(fl, snap[arg]));
      held = arg;
    } else if (op == ACT_ADVANCE) {
      held = -1;
      for (int k = 0; k < arg; ++k) ADJ_TRY(ibm_step(fl));
      advanced += arg;
    } else {  // adjoint of step `arg`; the solver holds the state before it
      held = -1;
      if (!fl->flist_host.empty())
        ADJ_HIP(launch_copy_force_list(fl->flist, (uint32_t)fl->flist_host.size(), fl->F_cur, nullptr, s));
      fl->field_present = false;
      ADJ_TRY(ibm_enqueue_recorded(view_of(fl), fl->ibm));
      fl->field_present = true;
      ADJ_HIP(launch_adj_collide_stream(fl->storage, fl->parity, fl->d, fl->f, fl->mask, fl->omega,
                                                fl->F_cur, fl->F_mom, fl->hp, lamA, lamB, wbar, Fbar, G1, s));
      ADJ_TRY(ibm_adjoint(view_of(fl), fl->ibm, Fbar, G0, arg >= steps - window ? tq_bar : 0.0, lamB, obar));
      std::swap(lamA, lamB);
      std::swap(G0, G1);
    }
  }
  double ob = 0.0;
  std::vector<double> wb(dJ_dnu ? N : 0);
  ADJ_HIP(hipMemcpyAsync(&ob, obar, sizeof ob, hipMemcpyDeviceToHost, s));
  if (dJ_dnu) ADJ_HIP(hipMemcpyAsync(wb.data(), wbar, N * sizeof(double), hipMemcpyDeviceToHost, s));
  std::vector<float> om(fl->omega && dJ_dnu ? N : 0);
  if (!om.empty()) ADJ_HIP(hipMemcpyAsync(om.data(), fl->omega, N * 4, hipMemcpyDeviceToHost, s));
  ADJ_HIP(hipStreamSynchronize(s));
  if (dJ_domega) *dJ_domega = ob;
  if (dJ_dnu) {
    // omega_plus = 1 / (3 nu + 1/2): d omega_plus / d nu = -3 omega_plus^2
    for (size_t i = 0; i < N; ++i) {
      const double op = om.empty() ? (double)fl->hp.omega_p : (double)om[i];
      dJ_dnu[i] = wb[i] * (-3.0 * op * op);
    }
  }
  if (forward_steps) *forward_steps = advanced;
  ADJ_TRY(snap_restore(fl, fin));
  ADJ_HIP(hipStreamSynchronize(s));
  reversing = false;
  cleanup();
#undef ADJ_HIP
#undef ADJ_TRY
  return LBM_OK;
}

}  // extern "C"

// ---- checkpoint rank file, format v3 ------------------------------------------
// The binary part of Solver::save / Solver::restore for one monolithic part
// (solver.rs:4536-4630 writer, 1113-1215 reader, format constants 29-45):
// "LBMKPT\0\0", u32 version 3, u8 endianness 0, u8 dtype (0 = f32), u16
// lattice id (319 / 327), u16 Q, u16 D, u64 padded cells, u64 core cells,
// u32 section count, (u32 id, u64 offset, u64 length) per section, the
// payload, and the FNV-1a 64 hash of the payload.  Arrays use the
// reference's host layout: q-major planes over the box padded by one halo
// cell per side (fields.rs:1-16, subdomain.rs:41); moments and the force
// field are compact.  Halo cells of periodic axes carry the copies the mask
// exchange would make (halo.rs:156-196, 456-492), other halo cells zero.
// The stream stash section holds the populations again: on the closed-wall
// path no slot it keeps is ever read.  The manifest (JSON) and the scenario
// and decomposition hashes belong to the caller.

namespace {

constexpr uint32_t kCkptVersion = 3;
constexpr uint32_t kSecF = 1, kSecStash = 2, kSecMoments = 3, kSecSolid = 4, kSecForce = 5, kSecScalars = 6;

// Open-face stash geometry on the host (the device's face_cell / unk_index
// of lbm_open.hip): face cells in (t2 outer, t1 inner) order, unknowns are
// the directions entering the domain through the face, ascending q.
struct FaceMap {
  int a, nin, t1, t2;
  uint32_t ext[3], nf;
  FaceMap(const Dom& d, int face)
      : a(face / 2), nin(face % 2 == 0 ? 1 : -1), t1(face / 2 == 0 ? 1 : 0), t2(face / 2 == 2 ? 1 : 2) {
    ext[0] = d.nx;
    ext[1] = d.ny;
    ext[2] = d.nz;
    nf = ext[t1] * ext[t2];
  }
  void cell(uint32_t coord, int64_t xyz[3]) const {
    xyz[a] = nin > 0 ? 0 : (int64_t)ext[a] - 1;
    xyz[t1] = coord % ext[t1];
    xyz[t2] = coord / ext[t1];
  }
  std::vector<int> unknowns(int Q) const {
    std::vector<int> u;
    for (int q = 0; q < Q; ++q)
      if (lat_c(q, a) * nin > 0) u.push_back(q);
    return u;
  }
};

void put_u32(std::vector<uint8_t>& b, uint32_t v) { b.insert(b.end(), (uint8_t*)&v, (uint8_t*)&v + 4); }
void put_u64(std::vector<uint8_t>& b, uint64_t v) { b.insert(b.end(), (uint8_t*)&v, (uint8_t*)&v + 8); }
void put_f64(std::vector<uint8_t>& b, double v) { b.insert(b.end(), (uint8_t*)&v, (uint8_t*)&v + 8); }
void put_f32s(std::vector<uint8_t>& b, const float* v, size_t n) {
  b.insert(b.end(), (const uint8_t*)v, (const uint8_t*)(v + n));
}

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
  const uint8_t* p = (const uint8_t*)data;
  for (size_t k = 0; k < n; ++k) {
    h ^= p[k];
    h *= 0x100000001b3ull;
  }
  return h;
}

struct PadGeom {
  int64_t n[3], p[3];
  size_t np, nc;
  explicit PadGeom(const Dom& d) {
    n[0] = d.nx;
    n[1] = d.ny;
    n[2] = d.nz;
    for (int a = 0; a < 3; ++a) p[a] = n[a] + 2;
    np = (size_t)(p[0] * p[1] * p[2]);
    nc = (size_t)(n[0] * n[1] * n[2]);
  }
  size_t pidx(int64_t x, int64_t y, int64_t z) const { return (size_t)(((z + 1) * p[1] + (y + 1)) * p[0] + (x + 1)); }
  size_t cidx(int64_t x, int64_t y, int64_t z) const { return (size_t)((z * n[1] + y) * n[0] + x); }
};

// one padded plane from a compact plane, halo filled as a single periodic
// part's mask exchange does: axes in order x, y, z, low face then high face;
// tangents of earlier axes span the halo (forwarding corners)
template <typename T>
void pad_plane(const PadGeom& g, const int per[3], const T* core, T* out) {
  std::fill(out, out + g.np, T{});
  for (int64_t z = 0; z < g.n[2]; ++z)
    for (int64_t y = 0; y < g.n[1]; ++y)
      for (int64_t x = 0; x < g.n[0]; ++x) out[g.pidx(x, y, z)] = core[g.cidx(x, y, z)];
  for (int a = 0; a < 3; ++a) {
    if (!per[a]) continue;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    auto range = [&](int t, int64_t& lo, int64_t& hi) {
      lo = t < a ? -1 : 0;
      hi = t < a ? g.n[t] + 1 : g.n[t];
    };
    int64_t l1, h1, l2, h2;
    range(t1, l1, h1);
    range(t2, l2, h2);
    for (int side = 0; side < 2; ++side) {
      const int64_t dst = side == 0 ? -1 : g.n[a];
      const int64_t src = side == 0 ? g.n[a] - 1 : 0;
      for (int64_t c2 = l2; c2 < h2; ++c2)
        for (int64_t c1 = l1; c1 < h1; ++c1) {
          int64_t ps[3], pd[3];
          ps[a] = src;
          pd[a] = dst;
          ps[t1] = pd[t1] = c1;
          ps[t2] = pd[t2] = c2;
          out[g.pidx(pd[0], pd[1], pd[2])] = out[g.pidx(ps[0], ps[1], ps[2])];
        }
    }
  }
}

// part_mask_hash (solver.rs:967-979): padded solid flags, then padded wall
// velocities as f32 bit patterns
uint64_t mask_hash(const PadGeom& g, const int per[3], const std::vector<uint8_t>& solid,
                   const std::vector<float>& wall) {
  std::vector<uint8_t> ps(g.np);
  pad_plane(g, per, solid.data(), ps.data());
  uint64_t h = 0xcbf29ce484222325ull;
  const uint64_t np = g.np;
  h = fnv1a(&np, 8, h);
  for (size_t k = 0; k < g.np; ++k) {
    const uint64_t v = ps[k] ? 1 : 0;
    h = fnv1a(&v, 8, h);
  }
  std::vector<float> comp(g.nc), pw(g.np);
  std::vector<float> padded(3 * g.np, 0.0f);
  if (!wall.empty())
    for (int c = 0; c < 3; ++c) {
      for (size_t k = 0; k < g.nc; ++k) comp[k] = wall[3 * k + c];
      pad_plane(g, per, comp.data(), pw.data());
      for (size_t k = 0; k < g.np; ++k) padded[3 * k + c] = pw[k];
    }
  return fnv1a(padded.data(), padded.size() * 4, h);
}

struct CkptWriter {
  FILE* fp = nullptr;
  uint64_t h = 0xcbf29ce484222325ull;
  bool ok = true;
  void raw(const void* p, size_t n) {
    if (ok && fwrite(p, 1, n, fp) != n) ok = false;
  }
  void payload(const void* p, size_t n) {
    h = fnv1a(p, n, h);
    raw(p, n);
  }
};

int ckpt_host_state(lbm_fields* fl, std::vector<float>& f, std::vector<uint8_t>& solid, std::vector<float>& wall,
                    std::vector<float>& mom, std::vector<float>& force, int* present) {
  const size_t N = fl->d.N;
  f.resize((size_t)fl->Q * N);
  solid.resize(N);
  mom.resize(4 * N);
  force.resize(3 * N);
  int rc = lbm_stage_out(fl, f.data());
  if (!rc) rc = lbm_read_solid(fl, solid.data());
  if (!rc) rc = lbm_read_moments(fl, mom.data(), mom.data() + N, mom.data() + 2 * N, mom.data() + 3 * N);
  if (!rc) rc = lbm_read_force_field(fl, force.data(), present);
  if (rc) return rc;
  wall.clear();
  if (fl->wall_u) {
    wall.resize(3 * N);
    HIPCHK(fl->ctx, hipMemcpyAsync(wall.data(), fl->wall_u, 3 * N * 4, hipMemcpyDeviceToHost, fl->ctx->stream));
    HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  }
  return LBM_OK;
}

}  // namespace

extern "C" {

uint64_t lbm_fnv1a64(const void* data, size_t n) { return fnv1a(data, n); }

int lbm_checkpoint_write_part(lbm_fields* fl, const char* path, lbm_ckpt_info* info) {
  if (!fl || !path || !info) return LBM_ERR_INVALID;
  const PadGeom g(fl->d);
  const int per[3] = {(int)fl->d.per[0], (int)fl->d.per[1], (int)fl->d.per[2]};
  std::vector<float> f, wall, mom, force;
  std::vector<uint8_t> solid;
  int present = 0;
  int rc = ckpt_host_state(fl, f, solid, wall, mom, force, &present);
  if (rc) return rc;
  const size_t Q = (size_t)fl->Q;
  // STALE_STASH: the previous step's post-collision populations (the
  // reference's second buffer); at open faces the entering unknowns are the
  // stash the next step restores
  std::vector<float> post(Q * g.nc);
  {
    float* dpost = nullptr;
    rc = dev_alloc(fl, (void**)&dpost, Q * g.nc * 4);
    if (rc) return rc;
    hipError_t e = launch_poststate(fl->Q, fl->storage, fl->parity, fl->d, fl->f, fl->mask, fl->wall_u, dpost,
                                    fl->ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(post.data(), dpost, Q * g.nc * 4, hipMemcpyDeviceToHost, fl->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(fl->ctx->stream);
    dev_free(fl, dpost, Q * g.nc * 4);
    if (e != hipSuccess) return fail(fl->ctx, LBM_ERR_DEVICE, "post-collision state: %s", hipGetErrorString(e));
  }
  for (int face = 0; face < 6; ++face) {
    if (!fl->stash[face][0]) continue;
    const FaceMap fm(fl->d, face);
    const std::vector<int> unk = fm.unknowns(fl->Q);
    std::vector<float> st(unk.size() * fm.nf);
    HIPCHK(fl->ctx, hipMemcpyAsync(st.data(), fl->stash[face][fl->stash_sel], st.size() * 4, hipMemcpyDeviceToHost,
                                   fl->ctx->stream));
    HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
    for (uint32_t coord = 0; coord < fm.nf; ++coord) {
      int64_t c[3];
      fm.cell(coord, c);
      const size_t i = g.cidx(c[0], c[1], c[2]);
      if (solid[i]) continue;
      for (size_t k = 0; k < unk.size(); ++k) post[(size_t)unk[k] * g.nc + i] = st[k * fm.nf + coord];
    }
  }
  // SCALAR_FIELDS (scalar_checkpoint_bytes, solver.rs:1202-1215): count, then
  // per scalar name, diffusivity, h (padded planes), its ping-pong partner,
  // and the compact concentration
  std::vector<uint8_t> scal;
  if (!fl->scalars.empty()) {
    put_u32(scal, (uint32_t)fl->scalars.size());
    for (size_t k = 0; k < fl->scalars.size(); ++k) {
      int Qs = 0;
      double D = 0.0;
      scalar_info(fl->scalars[k], &Qs, &D);
      std::vector<float> h((size_t)Qs * g.nc), h2((size_t)Qs * g.nc), C(g.nc), plane(g.np);
      rc = scalar_get_state(view_of(fl), fl->scalars[k], h.data(), h2.data(), C.data());
      if (rc) return rc;
      const std::string name = "scalar" + std::to_string(k);
      put_u32(scal, (uint32_t)name.size());
      scal.insert(scal.end(), name.begin(), name.end());
      put_f64(scal, D);
      for (const std::vector<float>* buf : {&h, &h2}) {
        put_u64(scal, (uint64_t)Qs * g.np);
        for (int q = 0; q < Qs; ++q) {
          pad_plane(g, per, buf->data() + (size_t)q * g.nc, plane.data());
          put_f32s(scal, plane.data(), g.np);
        }
      }
      put_u64(scal, g.nc);
      put_f32s(scal, C.data(), g.nc);
    }
  }
  std::vector<std::pair<uint32_t, uint64_t>> sec = {{kSecF, Q * g.np * 4}, {kSecStash, Q * g.np * 4},
                                                    {kSecMoments, 4 * g.nc * 4}, {kSecSolid, g.np}};
  if (present) sec.push_back({kSecForce, 3 * g.nc * 4});
  if (!scal.empty()) sec.push_back({kSecScalars, scal.size()});
  CkptWriter w;
  w.fp = fopen(path, "wb");
  if (!w.fp) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_IO: cannot create %s", path);
  const uint32_t ver = kCkptVersion, nsec = (uint32_t)sec.size();
  const uint8_t endian = 0, dtype = 0;
  const uint16_t lat = fl->Q == 19 ? 319 : 327, q16 = (uint16_t)fl->Q, d16 = 3;
  const uint64_t np = g.np, nc = g.nc;
  w.raw("LBMKPT\0\0", 8);
  w.raw(&ver, 4);
  w.raw(&endian, 1);
  w.raw(&dtype, 1);
  w.raw(&lat, 2);
  w.raw(&q16, 2);
  w.raw(&d16, 2);
  w.raw(&np, 8);
  w.raw(&nc, 8);
  w.raw(&nsec, 4);
  uint64_t off = 0;
  for (auto& e : sec) {
    w.raw(&e.first, 4);
    w.raw(&off, 8);
    w.raw(&e.second, 8);
    off += e.second;
  }
  std::vector<float> plane(g.np);
  for (int pass = 0; pass < 2; ++pass)  // populations, then the post-collision stash
    for (size_t q = 0; q < Q; ++q) {
      pad_plane(g, per, (pass == 0 ? f.data() : post.data()) + q * g.nc, plane.data());
      w.payload(plane.data(), g.np * 4);
    }
  w.payload(mom.data(), 4 * g.nc * 4);
  std::vector<uint8_t> ps(g.np);
  pad_plane(g, per, solid.data(), ps.data());
  w.payload(ps.data(), g.np);
  if (present) w.payload(force.data(), 3 * g.nc * 4);
  if (!scal.empty()) w.payload(scal.data(), scal.size());
  const uint64_t ph = w.h;
  w.raw(&ph, 8);
  const long bytes = ftell(w.fp);
  if (fclose(w.fp) != 0) w.ok = false;
  if (!w.ok) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_IO: short write to %s", path);
  info->bytes = (uint64_t)bytes;
  info->payload_hash = ph;
  info->mask_hash = mask_hash(g, per, solid, wall);
  info->step = fl->time;
  return LBM_OK;
}

int lbm_checkpoint_read_part(lbm_fields* fl, const char* path, int64_t step, lbm_ckpt_info* info) {
  if (!fl || !path || !info || step < 0) return LBM_ERR_INVALID;
  const PadGeom g(fl->d);
  const int per[3] = {(int)fl->d.per[0], (int)fl->d.per[1], (int)fl->d.per[2]};
  // the manifest's mask hash against the current masks, before the rank file
  // is read (solver.rs:4954-4962)
  std::vector<uint8_t> solid(g.nc);
  int rc = lbm_read_solid(fl, solid.data());
  if (rc) return rc;
  {
    std::vector<float> wall;
    if (fl->wall_u) {
      wall.resize(3 * g.nc);
      HIPCHK(fl->ctx, hipMemcpyAsync(wall.data(), fl->wall_u, 3 * g.nc * 4, hipMemcpyDeviceToHost, fl->ctx->stream));
      HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
    }
    const uint64_t cur = mask_hash(g, per, solid, wall);
    if (cur != info->mask_hash)
      return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_MASK_MISMATCH: mask_hash differed: checkpoint=fnv1a64:%016llx "
                  "current=fnv1a64:%016llx", (unsigned long long)info->mask_hash, (unsigned long long)cur);
  }
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_IO: cannot open %s", path);
  std::vector<uint8_t> b;
  {
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, fp)) > 0) b.insert(b.end(), buf, buf + k);
    fclose(fp);
  }
  size_t pos = 0;
  auto take = [&](void* dst, size_t n) -> bool {
    if (b.size() - pos < n) return false;
    memcpy(dst, b.data() + pos, n);
    pos += n;
    return true;
  };
  char magic[8];
  uint32_t ver = 0, nsec = 0;
  uint8_t endian = 0, dtype = 0;
  uint16_t lat = 0, q16 = 0, d16 = 0;
  uint64_t np = 0, nc = 0;
  if (!take(magic, 8) || memcmp(magic, "LBMKPT\0\0", 8) != 0)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_BAD_MAGIC: rank file magic is not LBMKPT");
  if (!take(&ver, 4) || ver != kCkptVersion)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_VERSION_MISMATCH: format_version %u, supported %u", ver,
                kCkptVersion);
  if (!take(&endian, 1) || endian != 0)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_BAD_MAGIC: cross-endian checkpoints are not supported");
  if (!take(&dtype, 1) || !take(&lat, 2) || !take(&q16, 2) || !take(&d16, 2) || !take(&np, 8) || !take(&nc, 8) ||
      !take(&nsec, 4))
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: rank file ended inside the header");
  if (dtype != 0) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_DTYPE_MISMATCH: checkpoint dtype is not f32");
  if (q16 != fl->Q || d16 != 3 || lat != (fl->Q == 19 ? 319 : 327))
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_LATTICE_MISMATCH: checkpoint lattice %u (Q %u)", lat, q16);
  if (np != g.np || nc != g.nc)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_GEOM_MISMATCH: %llu/%llu cells, run has %zu/%zu",
                (unsigned long long)np, (unsigned long long)nc, g.np, g.nc);
  struct Ent {
    uint32_t id;
    uint64_t off, len;
  };
  std::vector<Ent> tab(nsec);
  for (auto& e : tab)
    if (!take(&e.id, 4) || !take(&e.off, 8) || !take(&e.len, 8))
      return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: rank file ended inside the section table");
  if (b.size() - pos < 8) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: missing payload hash trailer");
  const size_t pend = b.size() - 8;
  uint64_t ph;
  memcpy(&ph, b.data() + pend, 8);
  const uint8_t* payload = b.data() + pos;
  const size_t plen = pend - pos;
  if (ph != info->payload_hash)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_PAYLOAD_CORRUPT: rank payload hash fnv1a64:%016llx differs from "
                "manifest fnv1a64:%016llx", (unsigned long long)ph, (unsigned long long)info->payload_hash);
  if (fnv1a(payload, plen) != ph)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_PAYLOAD_CORRUPT: payload hash does not match its contents");
  auto section = [&](uint32_t id, uint64_t want, const uint8_t** out) -> int {
    for (auto& e : tab)
      if (e.id == id) {
        if (e.off > plen || e.len > plen - e.off)
          return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: section %u exceeds the payload", id);
        if (e.len != want)
          return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_GEOM_MISMATCH: section %u has %llu bytes, expected %llu",
                      id, (unsigned long long)e.len, (unsigned long long)want);
        *out = payload + e.off;
        return LBM_OK;
      }
    *out = nullptr;
    return LBM_OK;
  };
  const size_t Q = (size_t)fl->Q;
  const uint8_t *sf = nullptr, *sm = nullptr, *ss = nullptr, *sff = nullptr, *sst = nullptr, *ssc = nullptr;
  uint64_t sclen = 0;
  rc = section(kSecF, Q * g.np * 4, &sf);
  if (!rc) rc = section(kSecMoments, 4 * g.nc * 4, &sm);
  if (!rc) rc = section(kSecSolid, g.np, &ss);
  if (!rc) rc = section(kSecForce, 3 * g.nc * 4, &sff);
  if (!rc) rc = section(kSecStash, Q * g.np * 4, &sst);
  if (rc) return rc;
  for (auto& e : tab)
    if (e.id == kSecScalars) {
      if (e.off > plen || e.len > plen - e.off)
        return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: section %u exceeds the payload", e.id);
      ssc = payload + e.off;
      sclen = e.len;
    }
  if (!sf) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: checkpoint is missing required section F_PRIMARY");
  if (!sm) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: checkpoint is missing required section MOMENTS");
  bool any_stash = false;
  for (int face = 0; face < 6; ++face) any_stash = any_stash || fl->stash[face][0] != nullptr;
  if (any_stash && !sst)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: open faces need the STALE_STASH section");
  // scalars (restore_scalar_checkpoint, solver.rs:1217-1270): parsed and
  // checked before anything is installed; this run must hold the same number
  // of scalars on the same lattices with the same diffusivities
  struct ScalarIn {
    std::vector<float> h, h2, C;
  };
  std::vector<ScalarIn> sin_(fl->scalars.size());
  if (!fl->scalars.empty() && !ssc)
    return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: checkpoint holds no scalar fields for this run's scalars");
  if (ssc) {
    size_t pos2 = 0;
    auto need = [&](size_t n) { return sclen - pos2 >= n; };
    auto rd = [&](void* dst, size_t n) {
      memcpy(dst, ssc + pos2, n);
      pos2 += n;
    };
    uint32_t count = 0;
    if (!need(4)) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: scalar section");
    rd(&count, 4);
    if (count != fl->scalars.size())
      return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_SCENARIO_MISMATCH: checkpoint holds %u scalars, this run %zu", count,
                  fl->scalars.size());
    for (uint32_t k = 0; k < count; ++k) {
      int Qs = 0;
      double D = 0.0;
      scalar_info(fl->scalars[k], &Qs, &D);
      uint32_t nl = 0;
      if (!need(4)) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: scalar section");
      rd(&nl, 4);
      if (!need((size_t)nl + 8)) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: scalar section");
      pos2 += nl;
      double Df = 0.0;
      rd(&Df, 8);
      if (Df != D)
        return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_SCENARIO_MISMATCH: scalar %u diffusivity %.17g, this run %.17g", k,
                    Df, D);
      ScalarIn& si = sin_[k];
      for (std::vector<float>* buf : {&si.h, &si.h2}) {
        uint64_t len = 0;
        if (!need(8)) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: scalar section");
        rd(&len, 8);
        if (len != (uint64_t)Qs * g.np)
          return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_GEOM_MISMATCH: scalar checkpoint section length differs");
        if (!need(len * 4)) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: scalar section");
        buf->resize((size_t)Qs * g.nc);
        for (int q = 0; q < Qs; ++q)
          for (int64_t z = 0; z < g.n[2]; ++z)
            for (int64_t y = 0; y < g.n[1]; ++y)
              for (int64_t x = 0; x < g.n[0]; ++x)
                memcpy(&(*buf)[(size_t)q * g.nc + g.cidx(x, y, z)], ssc + pos2 + 4 * ((size_t)q * g.np + g.pidx(x, y, z)),
                       4);
        pos2 += len * 4;
      }
      uint64_t clen = 0;
      if (!need(8)) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: scalar section");
      rd(&clen, 8);
      if (clen != g.nc)
        return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_GEOM_MISMATCH: scalar checkpoint section length differs");
      if (!need(clen * 4)) return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_TRUNCATED: scalar section");
      si.C.resize(g.nc);
      rd(si.C.data(), g.nc * 4);
    }
    if (pos2 != sclen)
      return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_PAYLOAD_CORRUPT: scalar checkpoint section has trailing bytes");
  }
  if (ss) {
    std::vector<uint8_t> mine(g.np);
    pad_plane(g, per, solid.data(), mine.data());
    for (size_t k = 0; k < g.np; ++k)
      if ((ss[k] != 0) != (mine[k] != 0))
        return fail(fl->ctx, LBM_ERR_INVALID, "CKPT_MASK_MISMATCH: serialized solid mask differs from rebuilt mask");
  }
  // install: populations (core cells), force field, cached-moment force
  std::vector<float> f(Q * g.nc);
  for (size_t q = 0; q < Q; ++q)
    for (int64_t z = 0; z < g.n[2]; ++z)
      for (int64_t y = 0; y < g.n[1]; ++y)
        for (int64_t x = 0; x < g.n[0]; ++x)
          memcpy(&f[q * g.nc + g.cidx(x, y, z)], sf + 4 * (q * g.np + g.pidx(x, y, z)), 4);
  rc = lbm_stage_in(fl, f.data());
  if (rc) return rc;
  if (sff) {
    std::vector<float> ff(3 * g.nc);
    memcpy(ff.data(), sff, ff.size() * 4);
    rc = lbm_set_force_field(fl, ff.data());
  } else {
    rc = lbm_clear_force_field(fl);
  }
  if (!rc) rc = refresh_moments(fl);
  if (rc) return rc;
  for (size_t k = 0; k < sin_.size(); ++k) {
    rc = scalar_put_state(view_of(fl), fl->scalars[k], sin_[k].h.data(), sin_[k].h2.data(), sin_[k].C.data());
    if (rc) return rc;
  }
  // open faces: the entering unknowns of the face cells are the stash the
  // next step restores
  for (int face = 0; face < 6 && sst; ++face) {
    if (!fl->stash[face][0]) continue;
    const FaceMap fm(fl->d, face);
    const std::vector<int> unk = fm.unknowns(fl->Q);
    std::vector<float> st(unk.size() * fm.nf, 0.0f);
    for (uint32_t coord = 0; coord < fm.nf; ++coord) {
      int64_t c[3];
      fm.cell(coord, c);
      for (size_t k = 0; k < unk.size(); ++k)
        memcpy(&st[k * fm.nf + coord], sst + 4 * ((size_t)unk[k] * g.np + g.pidx(c[0], c[1], c[2])), 4);
    }
    HIPCHK(fl->ctx, hipMemcpyAsync(fl->stash[face][fl->stash_sel], st.data(), st.size() * 4, hipMemcpyHostToDevice,
                                   fl->ctx->stream));
    HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  }
  fl->gen++;
  fl->time = step;
  // the cached moments recomputed from the installed state must be the saved ones
  std::vector<float> mom(4 * g.nc);
  rc = lbm_read_moments(fl, mom.data(), mom.data() + g.nc, mom.data() + 2 * g.nc, mom.data() + 3 * g.nc);
  if (rc) return rc;
  for (size_t k = 0; k < g.nc; ++k) {
    if (solid[k]) continue;
    for (int c = 0; c < 4; ++c)
      if (memcmp(&mom[c * g.nc + k], sm + 4 * (c * g.nc + k), 4) != 0)
        return fail(fl->ctx, LBM_ERR_STATE, "CKPT_MOMENTS_MISMATCH: cell %zu moment %d differs after restore", k, c);
  }
  std::vector<float> wall;
  if (fl->wall_u) {
    wall.resize(3 * g.nc);
    HIPCHK(fl->ctx, hipMemcpyAsync(wall.data(), fl->wall_u, 3 * g.nc * 4, hipMemcpyDeviceToHost, fl->ctx->stream));
    HIPCHK(fl->ctx, hipStreamSynchronize(fl->ctx->stream));
  }
  info->bytes = b.size();
  info->payload_hash = ph;
  info->mask_hash = mask_hash(g, per, solid, wall);
  info->step = step;
  return LBM_OK;
}

}  // extern "C"
