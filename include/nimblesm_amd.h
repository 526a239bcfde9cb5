/*
 * nimblesm_amd.h -- C ABI of the MI355X (gfx950) explicit internal-force
 * library.  Plain C types only: sizes are int64_t, arrays are raw pointers,
 * a stream is an opaque void* (a hipStream_t, NULL = the null stream).
 *
 * Every entry point returns an nsm_status; on failure nsm_last_error()
 * returns a message for the calling thread.  No entry point throws.
 *
 * Conventions (identical to NimbleSM):
 *   node vectors       [N][3] double
 *   connectivity       [E][8] int32, 0-based, HEX8 node order of
 *                      src/nimble_element.cc:100-121
 *   gathered element   [E][8][3] double (the Kokkos "gathered" views,
 *   arrays             src/nimble_kokkos_model_data.cc:1636-1658)
 *   deformation grad.  [.][8][9] order XX,YY,ZZ,XY,YZ,ZX,YX,ZY,XZ
 *   Cauchy stress      [.][8][6] order XX,YY,ZZ,XY,YZ,ZX
 *                      (src/nimble_utils.h:91-109)
 *   internal force     carries the reference's minus sign
 *                      (src/nimble_element.h:614-616)
 */
#ifndef NIMBLESM_AMD_H
#define NIMBLESM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nsm_status {
  NSM_OK               = 0,
  NSM_ERR_ARG          = 1, /* invalid argument / size                        */
  NSM_ERR_HIP          = 2, /* HIP runtime error                              */
  NSM_ERR_NEG_JACOBIAN = 3, /* det <= 0 in Invert3x3 (src/nimble_utils.h:1243) */
  NSM_ERR_STATE        = 4, /* call out of order (e.g. no lumped mass yet)    */
  NSM_ERR_NOMEM        = 5  /* device allocation failed                       */
} nsm_status;

typedef enum nsm_precision {
  NSM_FP64  = 0, /* plain FP64, reference operation order                     */
  NSM_MIXED = 1  /* FP32 F / stress / B^T.sigma, FP64 node state and sums     */
} nsm_precision;

typedef enum nsm_field {
  NSM_FIELD_REFERENCE_COORDINATE = 0, /* [N][3] */
  NSM_FIELD_DISPLACEMENT         = 1, /* [N][3] */
  NSM_FIELD_VELOCITY             = 2, /* [N][3] */
  NSM_FIELD_ACCELERATION         = 3, /* [N][3] */
  NSM_FIELD_INTERNAL_FORCE       = 4, /* [N][3] */
  NSM_FIELD_LUMPED_MASS          = 5, /* [N]    */
  NSM_FIELD_DEFORMATION_GRADIENT = 6, /* [E][8][9], STEP_NP1 */
  NSM_FIELD_STRESS               = 7  /* [E][8][6], STEP_NP1 */
} nsm_field;

typedef enum nsm_bc_kind {
  NSM_BC_INITIAL_VELOCITY        = 0,
  NSM_BC_PRESCRIBED_VELOCITY     = 1,
  NSM_BC_PRESCRIBED_DISPLACEMENT = 2
} nsm_bc_kind;

/* kernels whose launch durations nsm_model_kernel_time() reports */
typedef enum nsm_kernel {
  NSM_KERNEL_ELEMENT = 0, /* fused F -> stress -> B^T.sigma element kernel
                             (tiled: also the update of tile-interior nodes)   */
  NSM_KERNEL_NODE    = 1, /* CSR gather + central-difference node kernel
                             (tiled: update of nodes shared between tiles)     */
  NSM_KERNEL_REACTION = 2, /* reaction-force sum                                 */
  NSM_KERNEL_CLOSURE  = 3  /* UQ closure + update of the approximate samples      */
} nsm_kernel;

const char* nsm_status_string(int status);
const char* nsm_last_error(void);
/* Library version and the device it was built for ("gfx950"). */
const char* nsm_version(void);

/* ---------------------------------------------------------------------------
 * Element seam: HexElement on gathered per-element arrays (device pointers).
 * ------------------------------------------------------------------------- */

/* Replaces HexElement::ComputeDeformationGradients (Kokkos overload,
 * src/nimble_element.h:430-502 via :353-361): F at the 8 Gauss points with
 * current coordinates x = X + u. */
int nsm_hex_compute_deformation_gradients(int64_t num_elem, const double* X_gathered, const double* u_gathered,
                                          double* F, void* stream);

/* Replaces HexElement::ComputeNodalForces (src/nimble_element.h:539-625 via
 * :644-653): f = -sum_q (dN/dx . sigma_q) det(J_q) w_q, per element. */
int nsm_hex_compute_nodal_forces(int64_t num_elem, const double* X_gathered, const double* u_gathered,
                                 const double* sigma, double* f, void* stream);

/* Replaces HexElement::ComputeLumpedMass (src/nimble_element.cc:173-193,
 * src/nimble_element.h:266-309): row sums of the consistent mass, [E][8]. */
int nsm_hex_compute_lumped_mass(int64_t num_elem, double density, const double* X_gathered, double* m,
                                void* stream);

/* Replaces HexElement::ComputeCharacteristicLength
 * (src/nimble_element.cc:221-260) on current coordinates, [E]. */
int nsm_hex_compute_characteristic_length(int64_t num_elem, const double* x_gathered, double* out, void* stream);

/* ---------------------------------------------------------------------------
 * Material seam.
 * ------------------------------------------------------------------------- */

/* Replaces NeohookeanMaterial::GetStress (array form
 * src/nimble_material.h:240-253, per point src/nimble_material.cc:252-310;
 * device per-point form src/nimble_material.h:269-286) for num_points
 * points: sigma_np1 from F_np1 through Polar_Decomp. */
int nsm_neohookean_get_stress(int64_t num_points, double bulk_modulus, double shear_modulus, const double* F_np1,
                              double* sigma_np1, void* stream);

/* The per-point NeohookeanMaterial::GetStress over the STEP_NP1 views of one
 * block where they lie, as a block material interface applies it
 * (compute_block_stress, src/nimble_kokkos_block_material_interface.cc:65-119;
 * per point src/nimble_material.h:269-286): num_elem x num_ipt points, F
 * value (e, q, c) at F_np1[e*F_elem_stride + q*F_ipt_stride + c*F_comp_stride],
 * sigma value at the same form with the sigma strides (device pointers).
 * Kokkos LayoutLeft [E][8][9] views have strides (1, E, 8E), LayoutRight
 * ones (72, 9, 1).  A model's FP64 arrays store elements in groups of 8: use
 * nsm_neohookean_get_stress_grouped on them. */
int nsm_neohookean_get_stress_strided(int64_t num_elem, int num_ipt, double bulk_modulus, double shear_modulus,
                                      const double* F_np1, int64_t F_elem_stride, int64_t F_ipt_stride,
                                      int64_t F_comp_stride, double* sigma_np1, int64_t sigma_elem_stride,
                                      int64_t sigma_ipt_stride, int64_t sigma_comp_stride, void* stream);
/* The same GetStress over views whose elements are stored in groups of 8
 * (the same reference interface as nsm_neohookean_get_stress_strided):
 * elements e in [first_elem, first_elem + num_elem), F value (e, q, c) at
 * F_np1[(e/8)*g + (e%8)*s_e + q*s_q + c*s_c] with F_strides = {g, s_e, s_q,
 * s_c}, sigma likewise.  A model's FP64 arrays (nsm_model_field_device_pointer,
 * base pointers) have F strides {576, 1, 8, 64} and sigma strides
 * {384, 1, 8, 64}; a block's elements [e0, e0 + n) are first_elem = e0,
 * num_elem = n.  Plain strides are the case g = 8 s_e, first_elem = 0. */
int nsm_neohookean_get_stress_grouped(int64_t first_elem, int64_t num_elem, int num_ipt, double bulk_modulus,
                                      double shear_modulus, const double* F_np1, const int64_t* F_strides,
                                      double* sigma_np1, const int64_t* sigma_strides, void* stream);

/* ---------------------------------------------------------------------------
 * Model-data seam: replaces nimble::ModelDataBase / nimble_kokkos::ModelData
 * (src/nimble_model_data_base.h:87-345, src/nimble_kokkos_model_data.cc) and
 * the central-difference loop of ExplicitTimeIntegrator::Integrate
 * (src/integrators/explicit_time_integrator.cc:177-278).  The model owns its
 * device arrays and its HIP stream; host pointers below are host memory.
 * ------------------------------------------------------------------------- */
typedef struct nsm_model nsm_model;

/* Blocks are given in the order they are stored (the reference iterates
 * blocks in ascending block id, src/nimble_model_data.cc:636-659); conn holds
 * the connectivity of all blocks concatenated in that order. */
int nsm_model_create(int64_t num_nodes, const double* reference_coordinates, int num_blocks, const int* block_ids,
                     const int64_t* block_num_elem, const int* conn, const double* density,
                     const double* bulk_modulus, const double* shear_modulus, int precision, nsm_model** out);
int nsm_model_destroy(nsm_model* model);

/* Replaces BoundaryConditionManager's constant-magnitude initial / prescribed
 * velocity and prescribed displacement conditions
 * (src/nimble_boundary_condition_manager.h:93-217).  Conditions apply in the
 * order they are added, as the reference's list does. */
int nsm_model_add_boundary_condition(nsm_model* model, int kind, int coordinate, double magnitude, int64_t num_nodes,
                                     const int* node_ids);
/* ModelDataBase::ApplyInitialConditions / ApplyKinematicConditions
 * (src/nimble_model_data_base.cc:84-100). */
int nsm_model_apply_initial_conditions(nsm_model* model);
int nsm_model_apply_kinematic_conditions(nsm_model* model, double time_current, double time_previous);

/* ModelData::ComputeLumpedMass (src/nimble_kokkos_model_data.cc:1067-1152):
 * lumped mass by a deterministic node gather in (block, element) order, and
 * the critical time step min_e charlen_e / sqrt(K/rho)
 * (src/nimble_block_base.cc:52-84). */
int nsm_model_compute_lumped_mass(nsm_model* model);
int nsm_model_get_critical_time_step(nsm_model* model, double* dt);

/* ModelData::ComputeInternalForce (src/nimble_kokkos_model_data.cc:1154-1286):
 * f_int, F_np1 and sigma_np1 from the model's current displacement. */
int nsm_model_compute_internal_force(nsm_model* model, double time_previous, double time_current,
                                     int is_output_step);

/* ComputeInternalForce split around a block material interface, as the
 * reference's Kokkos ModelData runs it when a BlockMaterialInterfaceFactory
 * is installed (src/nimble_kokkos_model_data.cc:1196-1257, the interface's
 * ComputeStress() between them, src/nimble_block_material_interface_base.h:76-84):
 * nsm_model_compute_deformation_gradients writes F_np1 from the current
 * displacement; the caller's stress computation then fills sigma_np1
 * (nsm_model_field_device_pointer; e.g. nsm_neohookean_get_stress_strided on
 * the model's stream, nsm_model_get_stream); nsm_model_compute_internal_force_from_stress
 * forms f_int from it.  FP64 models without UQ samples (NSM_ERR_STATE
 * otherwise).  With the reference's own Neo-Hookean stress the result is
 * bit-identical to nsm_model_compute_internal_force. */
int nsm_model_compute_deformation_gradients(nsm_model* model);
int nsm_model_compute_internal_force_from_stress(nsm_model* model);
/* The HIP stream (hipStream_t) the model's work is ordered on. */
int nsm_model_get_stream(nsm_model* model, void** stream);

/* ModelData::UpdateStates (src/nimble_kokkos_model_data.cc:963-1021).  The
 * stateless Neo-Hookean model never reads STEP_N, so this is a no-op kept for
 * interface parity. */
int nsm_model_update_states(nsm_model* model);

/* Reaction-force history: after every step, -sum over node_ids of
 * f_int[node][coordinate] (FP64).  The reference has no such output; this is
 * the quantity of interest of the synthetic benchmarks. */
int nsm_model_set_reaction_nodes(nsm_model* model, int64_t num_nodes, const int* node_ids, int coordinate);
int nsm_model_get_reaction_history(nsm_model* model, int64_t first_step, int64_t count, double* out);
/* The quantity of interest of the synthetic workloads: the reaction history
 * of steps [0, num_steps) sampled on the device at num_samples equally
 * spaced times.  Sample k is taken at p = (k+1)*num_steps/num_samples steps,
 * linearly interpolated between integer steps j (value after step j-1;
 * step 0 is the model at rest, value 0).  out: num_samples doubles (host). */
int nsm_model_sample_reaction_history(nsm_model* model, int64_t num_steps, int num_samples, double* out);
/* The same sampling, allocating nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc), where
 * nsm_model_sample_reaction_history allocates a device staging buffer per
 * call.  The values are staged in a fixed device array of
 * NSM_NOALLOC_MAX_SAMPLES doubles that each GPU holds for the life of the
 * process; num_samples above that maximum fails with NSM_ERR_ARG.  All calls
 * on one GPU share that array: calls that overlap on one GPU (other threads,
 * other models' streams) race. */
#define NSM_NOALLOC_MAX_SAMPLES 4096
int nsm_model_sample_reaction_history_noalloc(nsm_model* model, int64_t num_steps, int num_samples, double* out);

/* Steps [step_begin, step_end) of ExplicitTimeIntegrator::Integrate with
 * t += dt_user, delta = t - t_prev; *time_current is read and updated.  All
 * work stays on the device.  Requires nsm_model_compute_lumped_mass. */
int nsm_model_explicit_steps(nsm_model* model, int64_t step_begin, int64_t step_end, int64_t num_load_steps,
                             int64_t output_frequency, double dt_user, double* time_current);

/* Host copies of fields (ModelDataBase::GetVectorNodeData /
 * GetScalarNodeData, src/nimble_model_data_base.h:145-173).  Integration
 * point fields are returned as double whatever the precision. */
int nsm_model_get_field(nsm_model* model, int field, double* out);
int nsm_model_set_field(nsm_model* model, int field, const double* in);
/* Device address of a field (node fields double; ipt fields double in FP64
 * mode and float in mixed mode) and its element size in bytes.  Node fields
 * are [N][3] / [N] in the node order of nsm_model_get_node_order.  Ipt
 * fields (nc = 9 or 6 components, device elements in the order of
 * nsm_model_get_element_order) are stored in groups of 8 elements in FP64
 * models: value (c, q) of element e at 64 nc (e/8) + 64 c + 8 q + e%8 (E
 * rounded up to 8 elements allocated); untiled mixed models store them
 * element-fastest [comp][ipt][E]; tiled mixed models per element tile: the
 * tile holding device elements [e0, e0 + n) owns [8 nc e0, 8 nc (e0 + n)),
 * value (c, q) of its element e0 + l at 8 nc e0 + (4 c + q/2) 2n + 2l + q%2
 * (rows of 2n values: the element's points q and q+1, q even, side by side). */
int nsm_model_field_device_pointer(nsm_model* model, int field, void** ptr, int* elem_bytes);
/* Node order of the device node arrays: out[i] (i < N) is the mesh node
 * stored at device position i.  FP64 models keep the mesh numbering; tiled
 * mixed models store each element tile's interior nodes contiguously, then
 * the nodes shared between tiles.  Host copies (get/set_field) and node lists
 * (boundary conditions, reaction nodes) always use mesh numbering.  No
 * reference counterpart: the reference's node views keep the mesh order. */
int nsm_model_get_node_order(nsm_model* model, int64_t* out);
/* Element order of the device ipt arrays: out[i] (i < E) is the element, as
 * numbered by the conn given to nsm_model_create, stored at device position
 * i.  FP64 models keep that order; tiled mixed models store each block's
 * elements tile by tile.  No reference counterpart: the reference's views
 * keep the mesh order (src/nimble_kokkos_model_data.cc:894-961). */
int nsm_model_get_element_order(nsm_model* model, int64_t* out);
/* How the model assembles nodal forces: *tiled = 1 for tile-local assembly
 * (mixed precision), with the number of element tiles, of nodes updated
 * inside their tile and of nodes updated from per-tile partial sums, and
 * the number of tiles that sum each node's element forces per corner index
 * (the others sum them per element colour);
 * *tiled = 0 for element force slots + CSR node gather (all nodes "boundary").
 * No reference counterpart (the reference scatters element forces,
 * src/nimble_kokkos_model_data.cc:1259-1280). */
int nsm_model_assembly_info(nsm_model* model, int* tiled, int64_t* num_tiles, int64_t* interior_nodes,
                            int64_t* boundary_nodes, int64_t* gather_tiles);
/* Elements per element tile (= threads per workgroup of the tiled kernel) of
 * a tiled mixed model, 0 otherwise.  Chosen at creation: NSM_TILE_ELEMS (64,
 * 128, 256, 512) when set, 512 when that gives every CU eight tiles, else by
 * the autotuner from a few timed steps with each size (cached per mesh shape
 * for the process, and in the file NSM_AUTOTUNE_CACHE names).  No reference
 * counterpart (the reference's Kokkos policies are fixed,
 * src/nimble_kokkos_block_material_interface.cc:76-80). */
int nsm_model_tile_elems(nsm_model* model, int* tile_elems);

int nsm_model_synchronize(nsm_model* model);
/* Launch-duration statistics from HIP events of launches of `kernel` while
 * timing is enabled: average milliseconds and launch count.  enable = n > 0
 * times every n-th launch of each kernel (1: every launch), 0 disables.  The
 * element and node launches record their events themselves at the kernel's
 * start and end; the reaction and closure launches are bracketed by events
 * on the stream, which add a few microseconds each. */
int nsm_model_enable_kernel_timing(nsm_model* model, int enable);
int nsm_model_kernel_time(nsm_model* model, int kernel, double* avg_ms, int64_t* launches);
/* Sizes of the model: nodes, elements, CSR entries, bytes of device memory. */
int nsm_model_info(nsm_model* model, int64_t* num_nodes, int64_t* num_elem, int64_t* csr_entries,
                   int64_t* device_bytes);

/* ---------------------------------------------------------------------------
 * Off-nominal ensemble: the reference's UQ model (src/uq/nimble_uq.cc:145-345,
 * src/uq/nimble_uq_model_data.cc:145-248, src/uq/nimble_uq_block.cc:57-164).
 * Beside the nominal trajectory the model advances num_exact "exact" samples,
 * each with its own bulk and shear modulus per block (the off-nominal stress
 * of NeohookeanMaterial::GetOffNominalStress, src/nimble_material.cc:312-342),
 * and num_approx "approximate" samples whose internal force is the closure
 * (UqModel::ApplyClosure, src/uq/nimble_uq.cc:327-346)
 *     f_s = c[s][0] f_nominal + sum_k c[s][k+1] f_exact_k.
 * All samples share the lumped mass; initial and prescribed velocities apply
 * to every sample, prescribed displacements only to the nominal trajectory
 * (src/nimble_boundary_condition_manager.h:93-217).  Samples start at rest.
 * Samples are numbered exact first, then approximate (UqModel::Forces()).
 * ------------------------------------------------------------------------- */

/* Replaces UqModel::Initialize / Allocate / Setup (src/uq/nimble_uq.cc:270-325)
 * with the parameters already scaled (UqModel::ScaleParameters, :220-244):
 * bulk_modulus, shear_modulus [num_exact][num_blocks] in the block order of
 * nsm_model_create; coefficients [num_approx][num_exact + 1].  Replaces any
 * previous ensemble and zeroes every sample's state; 0, 0 removes it. */
int nsm_model_set_uq_samples(nsm_model* model, int num_exact, const double* bulk_modulus,
                             const double* shear_modulus, int num_approx, const double* coefficients);
/* Host copies of a sample's node fields (NSM_FIELD_DISPLACEMENT, _VELOCITY,
 * _ACCELERATION, _INTERNAL_FORCE) in mesh node order: the reference's
 * exact_* / sample_* node data (src/uq/nimble_uq.cc:281-300).  The reference
 * keeps no sample acceleration; here it is (1/m) f_s of the last step. */
int nsm_model_get_uq_field(nsm_model* model, int sample, int field, double* out);
int nsm_model_set_uq_field(nsm_model* model, int sample, int field, const double* in);

/* ---------------------------------------------------------------------------
 * Batched ensemble: independent samples advanced together on one stream.
 * Each sample is a copy of one of the ensemble's meshes with its own block
 * materials (density, bulk and shear modulus), its own magnitudes of the
 * mesh's boundary conditions and its own time step.  One step of every
 * sample is one element launch over all samples (mixed precision: a
 * persistent launch over the element tiles of every sample, which also
 * updates tile-interior nodes; FP64: element chunks of every sample) and one
 * node launch; a whole run can be captured in a HIP graph and replayed.
 *
 * The reference's counterpart is the serial loop over the exact samples of
 * its UQ model (src/uq/nimble_uq_model_data.cc:145-248,
 * src/uq/nimble_uq_block.cc:57-164), where all samples share one mesh, one
 * lumped mass and one time step and differ only in K and G.  Here each sample
 * has its own lumped mass (ModelData::ComputeLumpedMass with the sample's
 * densities, src/nimble_kokkos_model_data.cc:1067-1152) and critical time
 * step (src/nimble_block_base.cc:52-84), and its trajectory is the one the
 * nsm_model_* entry points give for that sample alone (bit-identical in
 * FP64; in mixed precision the same arithmetic, and every sample's result
 * is independent of the other samples).
 * ------------------------------------------------------------------------- */
typedef struct nsm_ensemble nsm_ensemble;

/* Kernels whose launch durations nsm_ensemble_kernel_time reports (eager
 * runs with timing enabled). */
typedef enum nsm_ensemble_kernel {
  NSM_ENSEMBLE_KERNEL_ELEMENT = 0, /* element launch of a step (all samples)   */
  NSM_ENSEMBLE_KERNEL_NODE    = 1  /* node launch of a step (all samples)      */
} nsm_ensemble_kernel;

/* Limits: conditions combine into at most NSM_ENSEMBLE_MAX_BC_CODES distinct
 * per-node combinations per mesh; a mesh has at most 8 blocks. */
#define NSM_ENSEMBLE_MAX_BC_CODES 8

/* tile_elems: elements per element tile of the mixed-precision kernel (64,
 * 128, 256 or 512), or 0 to let the autotuner choose at setup from timed
 * launches (the choice is cached per ensemble shape for the process). */
int nsm_ensemble_create(int precision, int tile_elems, nsm_ensemble** out);
int nsm_ensemble_destroy(nsm_ensemble* ens);
/* A mesh, with arguments as nsm_model_create's (blocks in storage order,
 * conn concatenated); *mesh_id receives its index. */
int nsm_ensemble_add_mesh(nsm_ensemble* ens, int64_t num_nodes, const double* reference_coordinates, int num_blocks,
                          const int64_t* block_num_elem, const int* conn, int* mesh_id);
/* A condition of a mesh, kinds as nsm_model_add_boundary_condition; the
 * magnitude is given per sample (nsm_ensemble_add_samples), in the order the
 * mesh's conditions are added, and conditions apply in that order. */
int nsm_ensemble_add_boundary_condition(nsm_ensemble* ens, int mesh_id, int kind, int coordinate, int64_t num_nodes,
                                        const int* node_ids);
/* Reaction-force history of every sample of the mesh: -sum over node_ids of
 * f_int[node][coordinate] after every step (as nsm_model_set_reaction_nodes). */
int nsm_ensemble_set_reaction_nodes(nsm_ensemble* ens, int mesh_id, int64_t num_nodes, const int* node_ids,
                                    int coordinate);
/* num_samples samples of mesh mesh_id: density, bulk_modulus, shear_modulus
 * [num_samples][num_blocks of the mesh]; bc_magnitude [num_samples][conditions
 * of the mesh] (may be NULL when the mesh has none).  *first_sample receives
 * the index of the first; samples are numbered in the order they are added.
 * Before nsm_ensemble_setup only. */
int nsm_ensemble_add_samples(nsm_ensemble* ens, int mesh_id, int num_samples, const double* density,
                             const double* bulk_modulus, const double* shear_modulus, const double* bc_magnitude,
                             int* first_sample);
/* New parameters for samples [first_sample, first_sample + num_samples), all
 * of one mesh, arrays as in nsm_ensemble_add_samples.  After setup: their
 * lumped mass and critical time step are recomputed at once; the next run
 * (or a replay of a captured run) uses them. */
int nsm_ensemble_set_samples(nsm_ensemble* ens, int first_sample, int num_samples, const double* density,
                             const double* bulk_modulus, const double* shear_modulus, const double* bc_magnitude);
/* Device data of every sample, element tiles (mixed), and each sample's
 * lumped mass and critical time step.  Once, after the last add_*. */
int nsm_ensemble_setup(nsm_ensemble* ens);
/* Each sample's critical time step (GetCriticalTimeStep), [num_samples]. */
int nsm_ensemble_get_critical_time_steps(nsm_ensemble* ens, double* dt);
/* Every sample from rest: u = v = a = 0, initial velocities, the kinematic
 * conditions at t = 0, then num_steps steps of
 * ExplicitTimeIntegrator::Integrate (src/integrators/explicit_time_integrator.cc:123-278)
 * with dt_user = final_time[s] / num_steps, t += dt_user, on output steps as
 * in nsm_model_explicit_steps (output_frequency 0: none).  use_graph: the run
 * is captured once per (num_steps, output_frequency) into a HIP graph and
 * replayed; per-sample parameters and times live in device memory, so a
 * replay after nsm_ensemble_set_samples uses the new ones. */
int nsm_ensemble_run(nsm_ensemble* ens, int64_t num_steps, int64_t output_frequency, const double* final_time,
                     int use_graph);
/* After a run: each sample's reaction history sampled at num_points equally
 * spaced times (the rule of nsm_model_sample_reaction_history), out
 * [num_samples][num_points]; and a sample's per-step history. */
int nsm_ensemble_sample_reaction_history(nsm_ensemble* ens, int num_points, double* out);
int nsm_ensemble_get_reaction_history(nsm_ensemble* ens, int sample, int64_t first_step, int64_t count, double* out);
/* Host copy of a sample's node field (NSM_FIELD_DISPLACEMENT, _VELOCITY,
 * _ACCELERATION, _INTERNAL_FORCE, _LUMPED_MASS) in its mesh's node order;
 * acceleration and internal force are those of the latest run's last step
 * (the only step that stores them). */
int nsm_ensemble_get_field(nsm_ensemble* ens, int sample, int field, double* out);
/* Samples, nodes and elements over all samples, element tiles of all samples
 * (mixed; element chunks in FP64), the tile size in use, device bytes. */
int nsm_ensemble_info(nsm_ensemble* ens, int64_t* num_samples, int64_t* num_nodes, int64_t* num_elem,
                      int64_t* work_items, int* tile_elems, int64_t* device_bytes);
/* What the autotuner measured at setup (tile_elems 0 at create): count
 * candidates, their tile sizes and seconds per step (negative: unusable);
 * count 0 when the choice came from the cache or was given. */
int nsm_ensemble_autotune_timings(nsm_ensemble* ens, int capacity, int* count, int* tile_elems,
                                  double* seconds_per_step);
int nsm_ensemble_enable_kernel_timing(nsm_ensemble* ens, int enable);
int nsm_ensemble_kernel_time(nsm_ensemble* ens, int kernel, double* avg_ms, int64_t* launches);

/* ---------------------------------------------------------------------------
 * Surrogate network of the ensemble estimator (no counterpart in the
 * reference): fully connected widths[0] -> widths[1] -> ... ->
 * widths[num_layers], GELU (error-function form) after every layer but the
 * last.  Weights and activations are BF16, products accumulate in FP32 on the
 * matrix cores; FP32 master parameters are trained by Adam on the mean
 * squared error.  Parameters are ordered W_1 [d_1][d_0] (row-major), b_1
 * [d_1], W_2, b_2, ...  All arrays below are host arrays of FP32 rows.
 * ------------------------------------------------------------------------- */
typedef struct nsm_mlp nsm_mlp;

/* Weights uniform in +-sqrt(6 / (fan_in + fan_out)) from SplitMix64(seed),
 * zero biases. */
int nsm_mlp_create(int num_layers, const int* widths, uint64_t seed, nsm_mlp** out);
int nsm_mlp_destroy(nsm_mlp* net);
int nsm_mlp_num_parameters(nsm_mlp* net, int64_t* count);
int nsm_mlp_get_parameters(nsm_mlp* net, float* out);
/* Replaces the parameters and resets the optimizer state. */
int nsm_mlp_set_parameters(nsm_mlp* net, const float* in);
/* y [batch][d_L] from x [batch][d_0]. */
int nsm_mlp_forward(nsm_mlp* net, int64_t batch, const float* x, float* y);
/* Mean over batch and outputs of (y - target)^2 and its gradient with
 * respect to every parameter (in parameter order). */
int nsm_mlp_loss_gradients(nsm_mlp* net, int64_t batch, const float* x, const float* target, double* loss,
                           float* grad);
/* `iterations` full-batch Adam steps (beta 0.9, 0.999, eps 1e-8) with the
 * learning rate decaying from lr to lr_final on a cosine; the loss before
 * each step goes to loss_history [iterations] (may be NULL). */
int nsm_mlp_train(nsm_mlp* net, int64_t batch, const float* x, const float* target, int iterations, double lr,
                  double lr_final, double* loss_history);

/* ---------------------------------------------------------------------------
 * Synthetic input: n^3 HEX8 unit cube, node (i,j,k) -> i+(n+1)(j+(n+1)k),
 * interior nodes jittered by U(-f,f)/n per axis from SplitMix64 keyed by
 * (seed, 3*node+axis).  Host arrays X [(n+1)^3][3], conn [n^3][8].
 * ------------------------------------------------------------------------- */
int nsm_jittered_cube(int n, uint64_t seed, double jitter_fraction, double* X, int* conn);

#ifdef __cplusplus
}
#endif
#endif
