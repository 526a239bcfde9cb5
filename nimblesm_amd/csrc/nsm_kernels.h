// nsm_kernels.h -- host-side launchers of the HIP kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsm {

// Per-(node, coordinate) resolved kinematic condition.  kind: 0 none,
// 1 prescribed velocity, 2 prescribed displacement.
struct BcEntry {
  int32_t kind[3];
  int32_t pad;
  double  value[3];
};

// Times of one explicit step, computed on the host exactly as
// src/integrators/explicit_time_integrator.cc:192-195 does.
struct StepTimes {
  double t_prev, t_cur, dt, half_dt;
};

struct NodeArrays {
  int64_t         num_nodes;
  const uint32_t* node_ptr;  // [N+1] CSR offsets into node_ent
  const uint32_t* node_ent;  // 8*element + local node, ascending element
  const double*   mass;      // [N]
  double*         u;         // [N][3]
  double*         v;         // [N][3]
  double*         a;         // [N][3]
  double*         f;         // [N][3] internal force
  const uint8_t*  bc_code;   // [N], 0 = free, else index+1 into bc_table
  const BcEntry*  bc_table;
};

// ---- element kernels (element range [e_begin, e_end) of one block) ----
void launch_element_fp64(int64_t e_begin, int64_t e_end, const int* conn, const double* X, const double* u, double K,
                         double G, double* Fip, double* sip, double* slots, int* status, hipStream_t s);
void launch_element_mixed(int64_t e_begin, int64_t e_end, const int* conn, const double* X, const double* u, float K,
                          float G, float* Fip, float* sip, float* slots, int* status, hipStream_t s);

// ---- element / material seams on gathered arrays ----
void launch_seam_defgrad(int64_t E, const double* Xg, const double* ug, double* F, int* status, hipStream_t s);
void launch_seam_forces(int64_t E, const double* Xg, const double* ug, const double* sig, double* f, int* status,
                        hipStream_t s);
void launch_seam_stress(int64_t P, double K, double G, const double* F, double* sig, hipStream_t s);
void launch_seam_lumped_mass(int64_t E, double rho, const double* Xg, double* m, int* status, hipStream_t s);
void launch_seam_char_length(int64_t E, const double* xg, double* out, hipStream_t s);

// ---- model setup ----
void launch_element_mass(int64_t e_begin, int64_t e_end, const int* conn, const double* X, double rho,
                         double* elem_mass, int* status, hipStream_t s);
// per-workgroup minimum of charlen/c over [e_begin, e_end) into wg_min[blockIdx]
int  launch_element_dt(int64_t e_begin, int64_t e_end, const int* conn, const double* X, const double* u,
                       double sound_speed, double* wg_min, hipStream_t s);
void launch_min_reduce(const double* in, int64_t n, double* out, hipStream_t s);
void launch_mass_gather(int64_t N, const uint32_t* node_ptr, const uint32_t* node_ent, const double* elem_mass,
                        double* m, hipStream_t s);

// ---- node kernels ----
// Kinematic conditions on v (and nothing else): the explicit scheme's
// ApplyKinematicBC.
void launch_apply_bc(const NodeArrays& na, StepTimes t, hipStream_t s);
// first half kick + BC + drift + BC of one step
void launch_prekick(const NodeArrays& na, StepTimes t, hipStream_t s);
// gather f from element slots, a = f/m, second half kick (+BC on output
// steps); if fuse_next, also the next step's first half kick, BC, drift, BC.
template <typename SlotT>
void launch_node_update(const NodeArrays& na, const SlotT* slots, StepTimes cur, bool output_bc, bool fuse_next,
                        StepTimes next, bool write_a, bool update_v, hipStream_t s);
// gather only (ComputeInternalForce without the integrator)
template <typename SlotT>
void launch_force_gather(const NodeArrays& na, const SlotT* slots, hipStream_t s);
void launch_reaction(const int* nodes, int64_t n, int coord, const double* f, double* out, hipStream_t s);

}  // namespace nsm
