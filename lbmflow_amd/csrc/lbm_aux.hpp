// lbm_aux.hpp — host interface of the rotating-IBM and scalar-ADE device
// paths (lbm_aux.hip), used by the C-ABI layer.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_internal.hpp"

namespace lbm {

// The slice of a domain's state the auxiliary paths need.
struct FieldsView {
  std::string* ctx_err;
  hipStream_t stream;
  int Q, storage;
  Dom d;
  void* f;
  const uint32_t* mask;
  int parity;
  HydroParams hp;
  float* F_cur;
  float* F_mom;
  bool field_present;
};

struct IbmState;
struct ScalarState;

// What the IBM kernels need of one domain (one entry per batched instance).
struct IbmDomain {
  Dom d;
  const void* f;
  const uint32_t* mask;
  float* F_cur;
  float F[3];  // uniform body force
};

// Rotating IBM (Solver::apply_rotating_ibm, solver.rs:3016-3282).  `cells`
// receives the sorted fluid cells the body can force.
IbmState* ibm_create(const FieldsView& v, int64_t n, const double* pos, const double* weight, const double center[3],
                     const double omega[3], const lbm_ibm_config& cfg, std::vector<uint32_t>* cells,
                     std::string* err);
void ibm_destroy(IbmState* s);
// One state may hold several bodies, one per domain ("instance"); a state
// made by ibm_create holds one.  ibm_concat joins single-body states (same
// IBM configuration) into one state whose instance k is parts[k].
IbmState* ibm_concat(const std::vector<const IbmState*>& parts, std::string* err);
int ibm_instances(const IbmState* s);
// The per-instance domain table the kernels read (populations, mask, force
// field, body force).  Uploads only when it changed; not inside a capture.
int ibm_set_domains(const FieldsView& v, IbmState* s, const IbmDomain* doms);
IbmDomain ibm_domain_of(const FieldsView& v);
// One update of every instance; `diag` receives one entry per instance.
// v supplies the stream, lattice, storage, parity and field flag.
int ibm_apply(const FieldsView& v, IbmState* s, lbm_ibm_diag* diag);
// The same update without the host read-back: the K diagnostics go to entry
// *pos of a device ring of `cap` entries (K x ibm_ring_entry_bytes each) and
// *pos advances.  Launches kernels only, so a stream capture can record it.
// With clear_list, the force field is first cleared at those clear_n cells
// (single body), in the same launch as the gather.
int ibm_enqueue(const FieldsView& v, IbmState* s, void* ring, uint32_t cap, uint32_t* pos,
                const uint32_t* clear_list = nullptr, uint32_t clear_n = 0);
size_t ibm_ring_entry_bytes();
// Running f64 torque integral per instance: out[4k + c] = sum of torque c
// over the updates since the last reset, out[4k + 3] = their count.
int ibm_torque_integral(const FieldsView& v, IbmState* s, int reset, double* out);
// Adjoint support (single body).  ibm_enqueue_recorded is ibm_enqueue that
// also keeps every sweep's impulses and du for ibm_adjoint, which runs the
// sweeps backwards: Fbar (+ Gbar, nullable) is the adjoint of the force
// field (dense N x 3 f64), tq_bar the adjoint of the update's shaft torque
// (-torque_z); it adds the population adjoint at the stencil cells to lam
// (Q x N f64, q-major) and the rotation-rate adjoint to *omega_bar (device).
int ibm_enqueue_recorded(const FieldsView& v, IbmState* s);
int ibm_adjoint(const FieldsView& v, IbmState* s, const double* Fbar, const double* Gbar, double tq_bar, double* lam,
                double* omega_bar);
// n ring records (entry-major, instance-minor) to ABI diagnostics
int ibm_ring_read(const FieldsView& v, const void* ring, uint32_t n, lbm_ibm_diag* out);

// Passive scalar ADE (Solver::step_scalar_distributions, solver.rs:2113-2205).
ScalarState* scalar_create(const FieldsView& v, double diffusivity, int Qs, std::string* err);
void scalar_destroy(ScalarState* s);
int scalar_set(const FieldsView& v, ScalarState* s, const float* conc);
int scalar_gather(const FieldsView& v, ScalarState* s, float* conc);
int scalar_step(const FieldsView& v, ScalarState* s);
int scalar_cv(const FieldsView& v, ScalarState* s, double* cv, int* defined);
int scalar_info(const ScalarState* s, int* Qs, double* D);
int scalar_get_state(const FieldsView& v, ScalarState* s, float* h, float* h2, float* C);
int scalar_put_state(const FieldsView& v, ScalarState* s, const float* h, const float* h2, const float* C);
int scalar_compartment_cv(const FieldsView& v, ScalarState* s, int has_imp, double imp_z, double cv[3],
                          int defined[3], int64_t counts[3]);

}  // namespace lbm
