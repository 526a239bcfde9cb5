// nimblesm_amd_nimble.hpp -- the NimbleSM side of the drop-in: classes a
// NimbleSM build compiles with NimbleSM's src/ on its include path, so that
// its explicit integrator and its stress plug-in seam run on the MI355X
// library (include/nimblesm_amd.h) unchanged.
//
//   nimble_amd::AmdModelData
//       a nimble::ModelDataBase (src/nimble_model_data_base.h:87-345) whose
//       blocks, lumped mass, critical time step and internal force live on
//       the device.  The integrator's node vectors stay host views, as the
//       integrator reads and writes them
//       (src/integrators/explicit_time_integrator.cc:177-278); each
//       ComputeInternalForce uploads the displacement and returns f_int.
//       DataManager::Initialize (src/nimble_data_manager.cc:106-114)
//       constructs it in place of nimble_kokkos::ModelData.
//   nimble_amd::AmdBlockMaterialInterface / AmdBlockMaterialInterfaceFactory
//       the stress plug-in seam (src/nimble_block_material_interface_base.h:76-84,
//       src/nimble_block_material_interface_factory_base.h:59-72): per block,
//       the Neo-Hookean stress of the model's STEP_NP1 deformation gradients
//       on the device (nsm_neohookean_get_stress_grouped).  Returned by
//       NimbleApplication::CreateBlockMaterialInterfaceFactory
//       (src/nimble.h:85, src/nimble.cc:280-283) in a NimbleSM application
//       that installs it.
//
// When a block material interface factory is installed, AmdModelData runs
// the force evaluation split around it, as nimble_kokkos::ModelData does
// (src/nimble_kokkos_model_data.cc:1196-1257): F_np1 on the device, the
// plug-in's ComputeStress(), then the forces from the stress it left.
// Without one (or for a mixed-precision model) the fused element kernel
// evaluates the Neo-Hookean stress itself.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nimble_block_material_interface_base.h"
#include "nimble_block_material_interface_factory_base.h"
#include "nimble_data_manager.h"
#include "nimble_data_utils.h"
#include "nimble_genesis_mesh.h"
#include "nimble_material.h"
#include "nimble_material_factory_base.h"
#include "nimble_model_data_base.h"
#include "nimble_parser.h"
#include "nimble_view.h"
#include "nimblesm_amd.hpp"

namespace nimble_amd {

class AmdModelData : public nimble::ModelDataBase {
 public:
  explicit AmdModelData(nsm_precision precision = NSM_FP64) : precision_(precision) {}

  // -- node data: host vectors, as nimble::ModelData keeps them
  //    (src/nimble_model_data.cc:79-91)
  int AllocateNodeData(nimble::Length length, std::string label, int num_objects) override
  {
    const int id           = (int)labels_.size();
    labels_[id]            = label;
    node_data_[id]         = std::vector<double>((size_t)num_objects * (size_t)nimble::LengthToInt(length, dim_));
    return id;
  }

  int GetFieldId(const std::string& label) const override
  {
    for (const auto& kv : labels_)
      if (kv.second == label) return kv.first;
    return -1;
  }

  nimble::Viewify<1> GetScalarNodeData(int field_id) override
  {
    auto& v = node_data_.at(field_id);
    return {v.data(), {(int)v.size()}, {1}};
  }

  nimble::Viewify<2> GetVectorNodeData(int field_id) override
  {
    auto& v = node_data_.at(field_id);
    return {v.data(), {(int)v.size() / 3, 3}, {3, 1}};
  }

  // -- blocks: one device model over the mesh, each block's Neo-Hookean
  //    parameters from the material factory as nimble::ModelData::EmplaceBlocks
  //    creates them (src/nimble_model_data.cc:416-437)
  void InitializeBlocks(nimble::DataManager&                                data_manager,
                        const std::shared_ptr<nimble::MaterialFactoryBase>& material_factory) override
  {
    const nimble::GenesisMesh& mesh   = data_manager.GetMesh();
    const nimble::Parser&      parser = data_manager.GetParser();
    const int                  N      = mesh.GetNumNodes();
    std::vector<double>        X((size_t)(3 * N));
    for (int n = 0; n < N; ++n) {
      X[(size_t)(3 * n)]     = mesh.GetCoordinatesX()[n];
      X[(size_t)(3 * n + 1)] = mesh.GetCoordinatesY()[n];
      X[(size_t)(3 * n + 2)] = mesh.GetCoordinatesZ()[n];
    }
    std::vector<BlockSpec> blocks;
    block_data_.clear();
    materials_.clear();
    for (int id : mesh.GetBlockIds()) {  // ascending, the reference's block map order
      material_factory->parse_and_create(parser.GetModelMaterialParameters(id), 8);
      std::shared_ptr<nimble::Material> mat = material_factory->get_material();
      if (mat->NumStateVariables() != 0 || mat->GetShearModulus() < 0.0)
        throw std::invalid_argument("AmdModelData: block " + std::to_string(id) + " is not a Neo-Hookean material");
      const int  ne = mesh.GetNumElementsInBlock(id);
      const int* c  = mesh.GetConnectivity(id);
      blocks.push_back({id, std::vector<int>(c, c + 8 * ne),
                        NeohookeanMaterial(mat->GetDensity(), mat->GetBulkModulus(), mat->GetShearModulus())});
      materials_.push_back(mat);
      block_data_.emplace_back(nullptr, mat.get(), id, ne, 8);
      block_first_elem_[id] = num_elem_;
      num_elem_ += ne;
    }
    dev_ = std::unique_ptr<ModelData>(new ModelData(X, std::move(blocks), precision_));
  }

  // -- lumped mass and the critical time step; the integrator reads the
  //    latter through the non-virtual ModelDataBase::GetCriticalTimeStep
  //    (src/nimble_model_data_base.h:296-300), so it is stored with
  //    SetCriticalTimeStep (:276-279)
  void ComputeLumpedMass(nimble::DataManager&) override
  {
    device().ComputeLumpedMass();
    SetCriticalTimeStep(device().GetCriticalTimeStep());
    auto m  = device().GetScalarNodeData(NSM_FIELD_LUMPED_MASS);
    auto& h = node_data_.at(GetFieldIdChecked("lumped_mass"));
    h.assign(m.begin(), m.end());
  }

  // -- the hot path: f_int of the given displacement
  void ComputeInternalForce(nimble::DataManager& data_manager, double time_previous, double time_current,
                            bool is_output_step, const nimble::Viewify<2>& displacement,
                            nimble::Viewify<2>& force) override
  {
    const int           N = (int)node_data_.at(GetFieldIdChecked("displacement")).size() / 3;
    std::vector<double> u((size_t)(3 * N));
    for (int n = 0; n < N; ++n)
      for (int c = 0; c < 3; ++c) u[(size_t)(3 * n + c)] = displacement(n, c);
    device().SetVectorNodeData(NSM_FIELD_DISPLACEMENT, u);
    const auto& factory = data_manager.GetBlockMaterialInterfaceFactory();
    if (factory && precision_ == NSM_FP64) {
      device().ComputeDeformationGradients();
      factory->create(time_previous, time_current, data_manager.GetFieldIDs(), block_data_, this)->ComputeStress();
      device().ComputeInternalForceFromStress();
    } else {
      device().ComputeInternalForce(time_previous, time_current, is_output_step);
    }
    const std::vector<double> f = device().GetVectorNodeData(NSM_FIELD_INTERNAL_FORCE);
    for (int n = 0; n < N; ++n)
      for (int c = 0; c < 3; ++c) force(n, c) = f[(size_t)(3 * n + c)];
  }

  // stateless Neo-Hookean: nothing is carried from step to step
  // (src/nimble_kokkos_model_data.cc:963-1021 copies STEP_NP1 to STEP_N)
  void UpdateStates(const nimble::DataManager&) override { device().UpdateStates(); }

  // -- for the stress plug-in: the device STEP_NP1 arrays of block `id`
  //    (FP64, elements of all blocks in groups of 8) and the stream to run on
  struct BlockIpt {
    const double* F;
    double*       sigma;
    int64_t       F_strides[4];      // group of 8 elements, element, point, component
    int64_t       sigma_strides[4];
    int64_t       first_elem, num_elem;
  };
  BlockIpt BlockIntegrationPointData(int id)
  {
    if (precision_ != NSM_FP64) throw std::invalid_argument("AmdModelData: plug-in stress needs an FP64 model");
    const int64_t e0 = block_first_elem_.at(id);
    auto*         F  = static_cast<double*>(device().DevicePointer(NSM_FIELD_DEFORMATION_GRADIENT));
    auto*         S  = static_cast<double*>(device().DevicePointer(NSM_FIELD_STRESS));
    int64_t       ne = 0;
    for (const auto& b : block_data_)
      if (b.id == id) ne = b.num_elems;
    return {F, S, {576, 1, 8, 64}, {384, 1, 8, 64}, e0, ne};
  }
  void* Stream() { return device().Stream(); }
  ModelData& device()
  {
    if (!dev_) throw std::logic_error("AmdModelData: InitializeBlocks has not run");
    return *dev_;
  }

 private:
  nsm_precision                                 precision_;
  std::map<int, std::string>                    labels_;
  std::map<int, std::vector<double>>            node_data_;
  std::vector<nimble::BlockData>                block_data_;
  std::vector<std::shared_ptr<nimble::Material>> materials_;
  std::map<int, int64_t>                        block_first_elem_;
  int64_t                                       num_elem_ = 0;
  std::unique_ptr<ModelData>                    dev_;
};

// The plug-in: Neo-Hookean stress of every block on the device, with the
// block's moduli from its material (BlockData::material_device).
class AmdBlockMaterialInterface : public nimble::BlockMaterialInterfaceBase {
 public:
  AmdBlockMaterialInterface(AmdModelData* model, const std::vector<nimble::BlockData>& blocks)
      : model_(model), blocks_(blocks)
  {
  }

  void ComputeStress() const override
  {
    for (const auto& b : blocks_) {
      const AmdModelData::BlockIpt ip = model_->BlockIntegrationPointData(b.id);
      const NeohookeanMaterial     mat(b.material_device->GetDensity(), b.material_device->GetBulkModulus(),
                                       b.material_device->GetShearModulus());
      mat.GetStressGrouped(ip.first_elem, ip.num_elem, b.num_points_per_elem, ip.F, ip.F_strides, ip.sigma,
                           ip.sigma_strides, model_->Stream());
    }
  }

 private:
  AmdModelData*                  model_;
  std::vector<nimble::BlockData> blocks_;
};

class AmdBlockMaterialInterfaceFactory : public nimble::BlockMaterialInterfaceFactoryBase {
 public:
  std::shared_ptr<nimble::BlockMaterialInterfaceBase> create(double, double, const nimble::FieldIds&,
                                                             const std::vector<nimble::BlockData>& blocks,
                                                             nimble::ModelDataBase* model_data_ptr) const override
  {
    auto* model = dynamic_cast<AmdModelData*>(model_data_ptr);
    if (!model) throw std::invalid_argument("AmdBlockMaterialInterfaceFactory needs an AmdModelData");
    return std::make_shared<AmdBlockMaterialInterface>(model, blocks);
  }
};

}  // namespace nimble_amd
