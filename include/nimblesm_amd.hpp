// nimblesm_amd.hpp -- header-only C++ adapters over the C ABI with the
// reference's class and method names, for C++ hosts that drive NimbleSM's
// explicit integrator:
//   nimble_amd::HexElement          <- nimble::HexElement (src/nimble_element.h:225)
//   nimble_amd::NeohookeanMaterial  <- nimble::NeohookeanMaterial (src/nimble_material.h:425)
//   nimble_amd::ModelData           <- nimble::ModelDataBase / nimble_kokkos::ModelData
//                                      (src/nimble_model_data_base.h:87-345)
//   nimble_amd::ExplicitTimeIntegrator <- the loop of
//                                      src/integrators/explicit_time_integrator.cc:177-278
// Errors are thrown as nimble_amd::Error (the reference throws
// std::invalid_argument or aborts, src/nimble.cc:131-137).
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "nimblesm_amd.h"

namespace nimble_amd {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status)
{
  if (status != NSM_OK) throw Error(status, std::string(nsm_status_string(status)) + ": " + nsm_last_error());
}

// Device-pointer seams on gathered [E][8][3] arrays.
struct HexElement {
  static constexpr int dim = 3, num_nodes = 8, num_int_pts = 8;
  static void ComputeDeformationGradients(int64_t num_elem, const double* X, const double* u, double* F,
                                          void* stream = nullptr)
  {
    check(nsm_hex_compute_deformation_gradients(num_elem, X, u, F, stream));
  }
  static void ComputeNodalForces(int64_t num_elem, const double* X, const double* u, const double* sigma, double* f,
                                 void* stream = nullptr)
  {
    check(nsm_hex_compute_nodal_forces(num_elem, X, u, sigma, f, stream));
  }
  static void ComputeLumpedMass(int64_t num_elem, double density, const double* X, double* m,
                                void* stream = nullptr)
  {
    check(nsm_hex_compute_lumped_mass(num_elem, density, X, m, stream));
  }
  static void ComputeCharacteristicLength(int64_t num_elem, const double* x, double* out, void* stream = nullptr)
  {
    check(nsm_hex_compute_characteristic_length(num_elem, x, out, stream));
  }
};

class NeohookeanMaterial {
 public:
  NeohookeanMaterial(double density, double bulk_modulus, double shear_modulus)
      : density_(density), bulk_modulus_(bulk_modulus), shear_modulus_(shear_modulus)
  {
  }
  double GetDensity() const { return density_; }
  double GetBulkModulus() const { return bulk_modulus_; }
  double GetShearModulus() const { return shear_modulus_; }
  // Device pointers: F_np1 [num_points][9], sigma_np1 [num_points][6].
  void GetStress(int64_t num_points, const double* F_np1, double* sigma_np1, void* stream = nullptr) const
  {
    check(nsm_neohookean_get_stress(num_points, bulk_modulus_, shear_modulus_, F_np1, sigma_np1, stream));
  }
  // Device pointers at (element, point, component) strides: the STEP_NP1
  // views of one block where they lie.
  void GetStressStrided(int64_t num_elem, int num_ipt, const double* F_np1, const int64_t F_strides[3],
                        double* sigma_np1, const int64_t sigma_strides[3], void* stream = nullptr) const
  {
    check(nsm_neohookean_get_stress_strided(num_elem, num_ipt, bulk_modulus_, shear_modulus_, F_np1, F_strides[0],
                                            F_strides[1], F_strides[2], sigma_np1, sigma_strides[0],
                                            sigma_strides[1], sigma_strides[2], stream));
  }
  // Arrays with elements in groups of 8 at (group, element, point, component)
  // strides, elements [first_elem, first_elem + num_elem): a block of an FP64
  // model's STEP_NP1 arrays.
  void GetStressGrouped(int64_t first_elem, int64_t num_elem, int num_ipt, const double* F_np1,
                        const int64_t F_strides[4], double* sigma_np1, const int64_t sigma_strides[4],
                        void* stream = nullptr) const
  {
    check(nsm_neohookean_get_stress_grouped(first_elem, num_elem, num_ipt, bulk_modulus_, shear_modulus_, F_np1,
                                            F_strides, sigma_np1, sigma_strides, stream));
  }

 private:
  double density_, bulk_modulus_, shear_modulus_;
};

struct BlockSpec {
  int                block_id;
  std::vector<int>   conn;  // [E_b][8]
  NeohookeanMaterial material;
};

class ModelData {
 public:
  ModelData(const std::vector<double>& reference_coordinates, std::vector<BlockSpec> blocks,
            nsm_precision precision = NSM_FP64)
  {
    std::map<int, const BlockSpec*> ordered;  // ascending block id, as the reference's std::map
    for (const auto& b : blocks) ordered[b.block_id] = &b;
    std::vector<int>     ids, conn;
    std::vector<int64_t> ne;
    std::vector<double>  rho, K, G;
    for (const auto& kv : ordered) {
      const BlockSpec& b = *kv.second;
      ids.push_back(b.block_id);
      ne.push_back((int64_t)b.conn.size() / 8);
      conn.insert(conn.end(), b.conn.begin(), b.conn.end());
      rho.push_back(b.material.GetDensity());
      K.push_back(b.material.GetBulkModulus());
      G.push_back(b.material.GetShearModulus());
    }
    num_nodes_ = (int64_t)reference_coordinates.size() / 3;
    num_elem_  = (int64_t)conn.size() / 8;
    check(nsm_model_create(num_nodes_, reference_coordinates.data(), (int)ids.size(), ids.data(), ne.data(),
                           conn.data(), rho.data(), K.data(), G.data(), precision, &h_));
  }
  ~ModelData() { nsm_model_destroy(h_); }
  ModelData(const ModelData&)            = delete;
  ModelData& operator=(const ModelData&) = delete;

  void AddBoundaryCondition(nsm_bc_kind kind, const std::vector<int>& node_set, int coordinate, double magnitude)
  {
    check(nsm_model_add_boundary_condition(h_, kind, coordinate, magnitude, (int64_t)node_set.size(),
                                           node_set.data()));
  }
  void ApplyInitialConditions() { check(nsm_model_apply_initial_conditions(h_)); }
  void ApplyKinematicConditions(double t_cur, double t_prev)
  {
    check(nsm_model_apply_kinematic_conditions(h_, t_cur, t_prev));
  }
  void   ComputeLumpedMass() { check(nsm_model_compute_lumped_mass(h_)); }
  double GetCriticalTimeStep()
  {
    double dt = 0.0;
    check(nsm_model_get_critical_time_step(h_, &dt));
    return dt;
  }
  void ComputeInternalForce(double t_prev, double t_cur, bool is_output_step)
  {
    check(nsm_model_compute_internal_force(h_, t_prev, t_cur, is_output_step ? 1 : 0));
  }
  void UpdateStates() { check(nsm_model_update_states(h_)); }
  // ComputeInternalForce split around a stress plug-in (FP64 models)
  void ComputeDeformationGradients() { check(nsm_model_compute_deformation_gradients(h_)); }
  void ComputeInternalForceFromStress() { check(nsm_model_compute_internal_force_from_stress(h_)); }
  void* Stream()
  {
    void* s = nullptr;
    check(nsm_model_get_stream(h_, &s));
    return s;
  }
  // device address of a field (see nsm_model_field_device_pointer)
  void* DevicePointer(nsm_field f, int* elem_bytes = nullptr)
  {
    void* p = nullptr;
    check(nsm_model_field_device_pointer(h_, f, &p, elem_bytes));
    return p;
  }
  void SetScalarNodeData(nsm_field f, const std::vector<double>& v) { check(nsm_model_set_field(h_, f, v.data())); }
  std::vector<double> GetVectorNodeData(nsm_field f)
  {
    std::vector<double> out((size_t)(3 * num_nodes_));
    check(nsm_model_get_field(h_, f, out.data()));
    return out;
  }
  std::vector<double> GetScalarNodeData(nsm_field f)
  {
    std::vector<double> out((size_t)num_nodes_);
    check(nsm_model_get_field(h_, f, out.data()));
    return out;
  }
  void SetVectorNodeData(nsm_field f, const std::vector<double>& v) { check(nsm_model_set_field(h_, f, v.data())); }
  // F ([E][8][9]) or sigma ([E][8][6]) at the 8 Gauss points, mesh element order
  std::vector<double> GetIntegrationPointData(nsm_field f)
  {
    const int64_t width = f == NSM_FIELD_DEFORMATION_GRADIENT ? 9 : 6;
    std::vector<double> out((size_t)(num_elem_ * 8 * width));
    check(nsm_model_get_field(h_, f, out.data()));
    return out;
  }
  // mesh element stored at each device element position (layout of the
  // device ipt arrays returned by nsm_model_field_device_pointer)
  std::vector<int64_t> ElementOrder()
  {
    std::vector<int64_t> out((size_t)num_elem_);
    check(nsm_model_get_element_order(h_, out.data()));
    return out;
  }
  void SetReactionNodes(const std::vector<int>& nodes, int coordinate)
  {
    check(nsm_model_set_reaction_nodes(h_, (int64_t)nodes.size(), nodes.data(), coordinate));
  }
  std::vector<double> GetReactionHistory(int64_t first, int64_t count)
  {
    std::vector<double> out((size_t)count);
    check(nsm_model_get_reaction_history(h_, first, count, out.data()));
    return out;
  }
  // off-nominal ensemble: K, G [num_exact][num_blocks] (ascending block
  // id), closure coefficients [num_approx][num_exact + 1]
  void SetUqSamples(int num_exact, const std::vector<double>& K, const std::vector<double>& G, int num_approx,
                    const std::vector<double>& coefficients)
  {
    check(nsm_model_set_uq_samples(h_, num_exact, K.data(), G.data(), num_approx, coefficients.data()));
  }
  // node vector of UQ sample `sample` (exact samples first)
  std::vector<double> GetUqNodeData(int sample, nsm_field f)
  {
    std::vector<double> out((size_t)(3 * num_nodes_));
    check(nsm_model_get_uq_field(h_, sample, f, out.data()));
    return out;
  }
  nsm_model* handle() { return h_; }
  int64_t    num_nodes() const { return num_nodes_; }
  int64_t    num_elem() const { return num_elem_; }

 private:
  nsm_model* h_ = nullptr;
  int64_t    num_nodes_ = 0, num_elem_ = 0;
};

// Central-difference driver; all steps run on the device.
class ExplicitTimeIntegrator {
 public:
  ExplicitTimeIntegrator(ModelData& model, double initial_time, double final_time, int64_t num_load_steps,
                         int64_t output_frequency)
      : model_(model),
        num_load_steps_(num_load_steps),
        output_frequency_(output_frequency),
        dt_user_((final_time - initial_time) / (double)num_load_steps),
        time_current_(initial_time)
  {
    if (final_time < initial_time) throw std::invalid_argument("final time is less than initial time");
  }
  void Initialize()
  {
    model_.ComputeLumpedMass();
    critical_time_step_ = model_.GetCriticalTimeStep();
    model_.ApplyInitialConditions();
    model_.ApplyKinematicConditions(time_current_, time_current_);
  }
  void Advance(int64_t nsteps)
  {
    const int64_t end = step_ + nsteps < num_load_steps_ ? step_ + nsteps : num_load_steps_;
    check(nsm_model_explicit_steps(model_.handle(), step_, end, num_load_steps_, output_frequency_, dt_user_,
                                   &time_current_));
    step_ = end;
  }
  void    Integrate() { Advance(num_load_steps_ - step_); }
  double  TimeCurrent() const { return time_current_; }
  double  CriticalTimeStep() const { return critical_time_step_; }
  int64_t Step() const { return step_; }

 private:
  ModelData& model_;
  int64_t    num_load_steps_, output_frequency_, step_ = 0;
  double     dt_user_, time_current_, critical_time_step_ = 0.0;
};

}  // namespace nimble_amd
