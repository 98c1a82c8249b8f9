/*
 * aperi_hip.h — C ABI of the MI355X (gfx950) explicit internal-force step.
 *
 * This library is a drop-in for the hot path of aperi-mech's explicit solver:
 * the per-step internal-force evaluation of an element block and the
 * central-difference update around it.  Every entry point names the
 * reference interface it replaces (paths relative to the aperi-mech tree).
 *
 * Conventions
 *   - Every call returns APHIP_OK (0) or a negative error code; the message of
 *     the last failure on the calling thread is returned by aphip_last_error().
 *   - Host arrays are in the caller's node / element numbering and in the
 *     reference's field layout (include/Field.h:129-138): node vectors are
 *     n_nodes x 3 row-major, element tensors n_elem x 9 row-major (3*i+j),
 *     connectivity n_elem x nodes_per_element.  The library renumbers nodes
 *     and elements internally (space-filling orders grouped by its tile plan)
 *     and converts on the way in/out.
 *   - A context owns one HIP stream and all device memory it allocates.  Calls
 *     on one context are stream-ordered and must come from one thread at a time.
 *   - All arithmetic is FP64.
 */
#ifndef APERI_HIP_H
#define APERI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APHIP_OK 0
#define APHIP_ERR_INVALID (-1)   /* bad argument or call order            */
#define APHIP_ERR_HIP (-2)       /* HIP runtime error                      */
#define APHIP_ERR_ALLOC (-3)     /* device or host allocation failed       */
#define APHIP_ERR_NUMERIC (-4)   /* inverted element / non-finite result   */
#define APHIP_ERR_UNSUPPORTED (-5)

typedef struct aphip_ctx aphip_ctx;

/* Element topology: include/Constants.h:23-25 (ElementTopology). */
enum { APHIP_TET4 = 0, APHIP_HEX8 = 1 };

/* Material kinds: include/Materials/{LinearElastic,NeoHooke,Plastic,Elastic,NeoHookeWithDamage,
 * PowerLawCreep,DruckerPrager}.h */
enum {
    APHIP_LINEAR_ELASTIC = 0,
    APHIP_NEO_HOOKEAN = 1,
    APHIP_J2 = 2,
    APHIP_SVK = 3,
    APHIP_NEO_HOOKEAN_DAMAGE = 4,
    APHIP_POWER_LAW_CREEP = 5,
    APHIP_DRUCKER_PRAGER = 6
};

/* Fields for aphip_get_field / aphip_set_field.
 * Node fields (n_nodes x 3): U, V, A, F; MASS is n_nodes x 1.
 * Element fields (n_elem x 9): H (displacement_gradient), P (pk1_stress);
 * STATE is n_elem x 1 (J2 equivalent plastic strain; n_elem x 8 with state_per_qp) or
 * n_elem x 2 (neo-Hookean with damage: damage, I1; power-law creep: effective creep strain,
 * effective stress rate; Drucker-Prager: plastic strain, volumetric plastic strain;
 * n_elem x 16 with state_per_qp);
 * VOLUME is n_elem x 1. */
enum {
    APHIP_FIELD_U = 0,
    APHIP_FIELD_V = 1,
    APHIP_FIELD_A = 2,
    APHIP_FIELD_F = 3,
    APHIP_FIELD_MASS = 4,
    APHIP_FIELD_H = 5,
    APHIP_FIELD_P = 6,
    APHIP_FIELD_STATE = 7,
    APHIP_FIELD_VOLUME = 8,
    /* two-pass RK method only, read-only: the physical displacement Phi d of the last
     * force evaluation ("displacement", src/Solvers/Base.cpp:58-63) and the cell force
     * before its scatter ("force", include/Solvers/Explicit.h:63) */
    APHIP_FIELD_RK_DISPLACEMENT = 9,
    APHIP_FIELD_RK_FORCE_LOCAL = 10
};

/* Kinematic formulation: the values of LagrangianFormulationType (include/Constants.h:33-36;
 * Incremental = 1 is not supported). */
enum { APHIP_TOTAL_LAGRANGIAN = 0, APHIP_UPDATED_LAGRANGIAN = 2, APHIP_SEMI_LAGRANGIAN = 3 };

/* aphip_config.flags */
#define APHIP_FLAG_TWO_PASS 1 /* element partials through HBM + node CSR gather (no LDS tiles) */
#define APHIP_FLAG_PIPELINED 2 /* Tet4 step as a persistent, software-pipelined kernel instead of
                                  one tile per workgroup (bitwise identical results) */

typedef struct {
    int32_t device; /* HIP device ordinal */
    int32_t flags;  /* APHIP_FLAG_* */
} aphip_config;

/* Material parameters.  Replaces the YAML-built MaterialProperties
 * (include/Materials/Base.h:163-168) consumed by CreateMaterial
 * (include/Materials/Factory.h:43-160). */
typedef struct {
    double density;
    double youngs_modulus;
    double poissons_ratio;
    double yield_stress;      /* J2 only */
    double hardening_modulus; /* J2 only */
    int32_t state_per_qp;     /* J2 or damage + Hex8: one state per Gauss point (extension; 0 = reference) */
    int32_t reserved;
    /* neo-Hookean with damage (include/Materials/NeoHookeWithDamage.h; YAML keys of
     * include/Materials/Factory.h:70-101, whose defaults for the optional ones are
     * maximum_damage 0.99, maximum_damage_increment 1, the two times -1).  The two times
     * only label separations for the reference's optional material separator and do
     * not change any field here. */
    double I1_critical;
    double I1_failure;
    double alpha;
    double maximum_damage;
    double maximum_damage_increment;
    double stop_failing_at_time;
    double fail_all_at_time;
    /* power-law creep (include/Materials/PowerLawCreep.h, YAML keys of Factory.h:134-153):
     * creep constant A, exponent n, thermal constant m; the reference requires a constant
     * temperature (use_constant_temperature = 1 with constant_temperature). */
    double creep_A;
    double creep_n;
    double creep_m;
    double constant_temperature;
    int32_t use_constant_temperature;
    int32_t reserved2;
    /* Drucker-Prager (include/Materials/DruckerPrager.h, Factory.h:118-132): yield function
     * s / sqrt(2) = A1 - A2 z of the Lode components, flow direction A2G */
    double dp_A1;
    double dp_A2;
    double dp_A2G;
} aphip_material_params;

/* ---- lifetime ---------------------------------------------------------- */

/* Replaces construction of ElementBase + ComputeInternalForceBase
 * (include/ElementBase.h:27-117, include/ComputeInternalForceBase.h:34-89). */
int aphip_create(aphip_ctx **ctx, const aphip_config *cfg);
int aphip_destroy(aphip_ctx *ctx);
const char *aphip_last_error(void);
/* Number of visible HIP devices (0 when none); no reference counterpart. */
int aphip_device_count(int32_t *count);

/* ---- mesh, material, state ----------------------------------------------- */

/* Element block: replaces the STK mesh/field binding of
 * NgpMeshData::ForEachElementAndConnectedNodes (include/NgpMeshData.h:131-153)
 * and the element volume functor (include/ComputeElementVolumeFunctor.h:47-69). */
int aphip_set_mesh(aphip_ctx *ctx, int32_t elem_type, int64_t n_nodes, const double *X, int64_t n_elem,
                   const int32_t *conn);
/* perm[i] = internal index of caller node i. */
int aphip_get_node_permutation(aphip_ctx *ctx, int32_t *perm);

/* Material of the block: replaces CreateMaterial (include/Materials/Factory.h:43-160)
 * and LinearElasticProperties::CompleteProperties (include/Materials/Base.h:47-126). */
int aphip_set_material(aphip_ctx *ctx, int32_t kind, const aphip_material_params *params);
/* Per-element Young's modulus / Poisson ratio (extension; no reference counterpart:
 * the reference holds one material per part).  NULL, NULL clears them. */
int aphip_set_material_fields(aphip_ctx *ctx, const double *E, const double *nu);

/* Kinematic formulation of the Gaussian elements (the reference's part formulation,
 * include/InternalForceContributionParameters.h lagrangian_formulation_type).
 * APHIP_UPDATED_LAGRANGIAN: each step integrates over the start-of-step configuration
 * x_n with the step's displacement increment (ExplicitTimeIntegratorUpdated,
 * include/ExplicitTimeIntegrator.h:193-226, :466-486), H = dH + H_n + dH H_n
 * (include/ComputeInternalForceBase.h:158-166) and the force with B F_r and w / det F_r
 * (include/ComputeInternalForceGaussian.h:109-123).  Call after aphip_set_mesh and
 * aphip_set_material; the current coordinates start at the mesh coordinates.  Runs on
 * the element-partials kernels; not with RKPM cells, ensembles, the power method or
 * the adjoint (errors). */
int aphip_set_formulation(aphip_ctx *ctx, int32_t formulation);
/* APHIP_SEMI_LAGRANGIAN with its reference-configuration update interval (the reference's
 * semi-Lagrangian configuration update interval, src/Application.cpp:301-307, 394-398):
 * the force integrates over the reference configuration with the displacement increment
 * accumulated since its last update and the reference displacement gradient
 * (ExplicitTimeIntegratorSemi, include/ExplicitTimeIntegrator.h:488-546); after the
 * steps n = interval, 2 interval, ... (counted from aphip_initialize) the reference
 * configuration becomes the current one, the increment restarts at zero and the reference
 * gradient takes H.  Graph replay is not used; rate-form materials are refused (their
 * velocity gradient needs the previous step's H beside the reference one).
 * aphip_set_formulation(APHIP_SEMI_LAGRANGIAN) uses interval 1. */
int aphip_set_formulation_semi(aphip_ctx *ctx, int32_t update_interval);

/* Lumped mass rho V_e / n_nodes per node: replaces ComputeMassMatrixForPart +
 * FinishComputingMassMatrix (src/MassUtils.cpp:57-112).  After aphip_rk_setup the
 * mass is scattered through the RK functions (MassUtils.cpp:87-97). */
int aphip_compute_mass(aphip_ctx *ctx, double *total_mass);

/* Velocity/acceleration held at zero on masked DOFs (n_nodes x 3 bytes, 1 = fixed):
 * replaces BoundaryCondition::ApplyVelocity/ApplyAcceleration for a fixed set
 * (src/BoundaryCondition.cpp:33-72).  NULL clears the mask. */
int aphip_set_fixed(aphip_ctx *ctx, const uint8_t *mask);

/* Upload / download a field in caller numbering and reference layout. */
int aphip_set_field(aphip_ctx *ctx, int32_t field, const double *data);
int aphip_get_field(aphip_ctx *ctx, int32_t field, double *out);

/* Device-resident field transfer, with no host memory touched: the same fields, caller
 * numbering and layout as aphip_set_field / aphip_get_field, but `data` / `out` are
 * device arrays on the context's device (hipMalloc'd; checked to be device memory of
 * that device and long enough for the field, else APHIP_ERR_INVALID).  The renumbering
 * runs on the GPU.  Both calls are queued on the context's stream and return without
 * waiting: the caller's writes to `data` must be complete before the call (for example
 * by synchronizing the stream that wrote them), `data` must stay unchanged until the
 * context's stream has passed the call, and `out` holds the field once it has
 * (aphip_synchronize, or any later call that waits for the stream, such as
 * aphip_internal_force).  Setting MASS waits for the stream to validate the values.
 * Replaces the field access of ComputeInternalForceBase::UpdateField and the NGP field
 * views the reference's force functors read and write on the device
 * (include/ComputeInternalForceBase.h:110-135, include/ElementTetrahedron4.h:184-192,
 * which marks the fields modified on the device). */
int aphip_set_field_device(aphip_ctx *ctx, int32_t field, const double *data);
int aphip_get_field_device(aphip_ctx *ctx, int32_t field, double *out);

/* ---- the hot path -------------------------------------------------------- */

/* f = -sum_e int B^T P dV from the current displacement; writes H, P and the NP1
 * state.  Replaces ElementTetrahedron4/ElementHexahedron8::ComputeInternalForceAllElements
 * (include/ElementTetrahedron4.h:184-192, include/ElementHexahedron8.h:174-182) and,
 * after aphip_rk_setup, ElementReproducingKernel::ComputeInternalForceAllElements
 * (include/ElementReproducingKernel.h:214-224). */
int aphip_internal_force(aphip_ctx *ctx, double time, double time_increment);

/* Initial force and acceleration a = f/m, as before the time loop of
 * ExplicitSolver::Solve (src/Solvers/Explicit.cpp:362-376). */
int aphip_initialize(aphip_ctx *ctx);

/* n_steps of the explicit loop body (src/Solvers/Explicit.cpp:383-452) with a fixed
 * increment (DirectTimeStepper, include/TimeStepper.h:97-128): state swap, v += dt/2 a,
 * velocity BC, u += dt v, internal force, a = f/m, acceleration BC, v += dt/2 a. */
int aphip_step(aphip_ctx *ctx, double time_increment, int32_t n_steps);

/* Stable time increment 2 / sqrt(lambda_max(M^-1 K)) by the power method about the
 * current displacement: replaces PowerMethodProcessor::ComputeStableTimeIncrement
 * (include/PowerMethodProcessor.h:331-470) of the power-method time stepper
 * (include/TimeStepper.h:138-229); the reference's defaults are 50 iterations,
 * epsilon 1e-5, tolerance 1e-3, variation tolerance 1e-7.  previous_time_increment is
 * the stepper's last full (unscaled) increment, 0 on the first call: the first
 * iteration's convergence test compares against it.  The first call after the mesh is
 * set starts from a vector uniform in [-1, 1] per DOF (a counter-based generator keyed
 * by `seed` and the caller's DOF index, times epsilon and the node's max edge length);
 * every later call starts from the eigenvector the previous call ended with, as the
 * reference's processor does.  u, f, H, P and the state are as before on return.  The
 * max edge lengths are computed on the first call and kept with the mesh; the call's
 * work buffers (8 (12 n_nodes + 18 n_elem + state values) bytes) are kept in the context
 * between calls and released with the mesh. */
int aphip_power_method_time_increment(aphip_ctx *ctx, double previous_time_increment, int32_t num_iterations,
                                      double epsilon, double tolerance,
                                      double variation_tolerance, uint64_t seed, double *time_increment,
                                      int32_t *iterations, int32_t *converged);

/* aphip_power_method_time_increment without any device allocation: same arguments, same
 * algorithm and the same results (it shares the kernels and the driver).  No counterpart
 * in the reference: it is for callers that cannot allocate device memory, such as a time
 * loop under HIP graph capture or a serving loop that forbids hipMalloc.  Its work buffers
 * (the displacement and force on entry, F(u), the iterate, the targets of H, P and the
 * state, and the norm's partial sums: 8 (12 n_nodes + 18 n_elem + state values) bytes
 * plus a few KiB) live in a fixed 64 MiB array per GPU, held from library load to process end;
 * a mesh that needs more is refused with APHIP_ERR_INVALID.  Every call on a GPU shares
 * that GPU's array: calls on contexts of the same GPU must not overlap in time (they
 * would race on it).  The eigenvector kept between calls stays in the context, as for
 * the allocating entry point. */
int aphip_power_method_time_increment_noalloc(aphip_ctx *ctx, double previous_time_increment,
                                              int32_t num_iterations, double epsilon, double tolerance,
                                              double variation_tolerance, uint64_t seed, double *time_increment,
                                              int32_t *iterations, int32_t *converged);

/* Exodus II results file (netCDF 64-bit-offset format, written by this library with
 * no netCDF dependency): replaces the reference's results output, IoMesh::
 * CreateFieldResultsFile / AddFieldResultsOutput / WriteFieldResults / CloseFieldResults
 * File (src/IoMesh.cpp:190-240) as driven by Solver::WriteOutput (src/Solvers/Base.cpp:
 * 114-127).  One element block with the caller's node and element numbering
 * (connectivity 1-based).  Node variables, with the reference's output names
 * (src/Solvers/Explicit.cpp:47-53, src/Solvers/Base.cpp:79-80): displacement, velocity,
 * acceleration, force, mass and essential_boundary, each as _x, _y, _z.  Element
 * variables: displacement_gradient and pk1_stress as _xx, _xy, _xz, _yx, ..., _zz (the
 * row-major components), volume, and, for J2 with one state per element,
 * equivalent_plastic_strain.  aphip_results_write appends one time step (the current
 * fields); a later aphip_results_open or aphip_destroy closes the file. */
int aphip_results_open(aphip_ctx *ctx, const char *path, const char *title);
int aphip_results_write(aphip_ctx *ctx, double time);
int aphip_results_close(aphip_ctx *ctx);

/* Element-tile shape of the fused step: elements per tile and workgroup size.
 * Tet4: one element per thread (64, 128, 256, 512 or 1024).  Hex8: one element per
 * thread (64, 128 or 256), or eight lanes per element, one per Gauss point
 * (threads = 8 x elements, 8..64 elements).  Rebuilds the tile plan; results do
 * not depend on it beyond the rounding of the summation grouping.  No reference
 * counterpart. */
int aphip_set_tiling(aphip_ctx *ctx, int32_t elems_per_tile, int32_t threads);

/* Replay aphip_step through a captured HIP graph (two steps per graph, so the
 * alternating state buffers return to their captured roles).  No reference
 * counterpart (the reference launches every kernel eagerly). */
int aphip_set_graph(aphip_ctx *ctx, int32_t enable);

/* Times the candidate tile shapes (and, for Tet4, the pipelined kernel) on the
 * context's own workload (trial_steps steps each, the simulation state is restored
 * afterwards) and keeps the fastest.  Writes elements_per_tile * 10000 + threads
 * (+ 100000000 for the pipelined kernel) of the choice to *chosen.  No reference
 * counterpart. */
int aphip_autotune(aphip_ctx *ctx, double time_increment, int32_t trial_steps, int32_t *chosen);

/* Wait for all work queued on the context's stream. */
int aphip_synchronize(aphip_ctx *ctx);

/* ---- RKPM (one-pass, element-as-cell strain smoothing) ------------------- */

/* Max edge length, kernel radius r = scale * l, variable-radius ball search with the
 * 40-neighbor cap, RK shape functions and the smoothed-cell CSR.  Replaces
 * MaxEdgeLengthProcessor::ComputeMaxEdgeLength (src/MaxEdgeLengthProcessor.cpp:87-104),
 * NeighborSearchProcessor::AddNeighborsWithinVariableSizedBall
 * (include/NeighborSearchProcessor.h:196-317, src/NeighborSearchProcessor.cpp:154-240),
 * FunctionCreationProcessor::compute_and_store_function_values
 * (include/FunctionCreationProcessor.h:73-134) and
 * SmoothedCellDataProcessor::SetSubcellNodeIndices / ComputeFunctionDerivatives
 * (include/SmoothedCellDataProcessor.h:407-675).  One smoothed cell per element of a Tet4
 * mesh (ElementReproducingKernelTet4) or of a Hex8 mesh (ElementReproducingKernelHex8,
 * include/ElementReproducingKernel.h:317-378; face-smoothed Hex8 B of QuadratureSmoothed.h:96-150). */
int aphip_rk_setup(aphip_ctx *ctx, double kernel_radius_scale);
/* aphip_rk_setup with the reference's force-method switch
 * (IntegrationSchemeStrainSmoothingParameters, include/InternalForceContributionParameters.h:
 * 115-117, "force_two_pass_method"): use_one_pass_method = 1 is aphip_rk_setup;
 * 0 selects the two-pass method of ExplicitSolver::ComputeForce (src/Solvers/Explicit.cpp:
 * 148-186): each cell lists its element's own nodes with derivative (B_k V)/V
 * (include/SmoothedCellDataProcessor.h:474-486, :637-648), every force evaluation
 * gathers the physical displacement Phi d (FunctionEvaluationProcessor::ComputeValues,
 * include/FunctionEvaluationProcessor.h:70-117), computes the cell force on it and
 * scatters it back, f = Phi^T f_local (ScatterValues, :130-181).  Both methods give
 * the same force up to rounding. */
int aphip_rk_setup_method(aphip_ctx *ctx, double kernel_radius_scale, int32_t use_one_pass_method);
/* Finite-element strain smoothing with one cell per Tet4 (ElementSmoothedTetrahedron4,
 * include/ElementSmoothedTetrahedron4.h, created by include/Element.h:44-47): each cell's
 * derivative is the face-smoothed B of its own four nodes (QuadratureSmoothed.h:32-62,
 * SmoothedCellDataProcessor.h:623-668 with the element's shape functions), the force
 * f_k -= V P B_k, the lumped mass rho V / 4.  Replaces the element's construction and
 * its ComputeInternalForceAllElements; a Hex8 mesh gets the reference's error
 * (include/Element.h:66-69).  The functions read back through aphip_rk_get_* are then the
 * identity (one neighbor, the node itself, value 1). */
int aphip_strain_smoothing_setup(aphip_ctx *ctx);
/* Nodal strain smoothing on a "thex" mesh (a Hex8 mesh made by splitting every tetrahedron
 * of a tet mesh into four hexahedra): the reference's nodal_smoothing_cell integration
 * (include/MeshLabelerProcessor.h:227-330 active nodes, :690-827 cells and subcells;
 * src/MaxEdgeLengthProcessor.cpp:54-55 edges to inactive nodes doubled; the search of
 * include/NeighborSearchProcessor.h:196-260 with the active nodes as the only candidates;
 * include/SmoothedCellDataProcessor.h:407-675 subcell lists and derivatives).  Every hex's
 * node of lowest caller id is active (the tet vertices); a mesh in which some hex then holds
 * another number of active nodes is refused with the reference's message.  One cell per
 * active node, made of the hexes around it; subdomains 0: every hex its own subcell,
 * subdomains 1: the whole cell one subcell.  use_f_bar (the reference's strain_smoothing
 * use_f_bar) applies F-bar over the subcells of each cell
 * (include/ComputeInternalForceSmoothedCell.h:120-413), which the reference does only when
 * cells hold more than one subcell (:458-459); F-bar takes the linear elastic, neo-Hookean,
 * J2 and SVK materials.  Only active nodes carry degrees of freedom: the others keep their
 * values through aphip_step, and their mass is zero.  Element fields H, P and STATE are per
 * subcell: every element reads its subcell's values (the reference's output population);
 * VOLUME is each element's smoothed volume. */
int aphip_rk_nodal_setup(aphip_ctx *ctx, double kernel_radius_scale, int32_t subdomains, int32_t use_f_bar);
/* The nodal cells of aphip_rk_nodal_setup: counts, cell_subcells[n_cells+1] (the subcells of
 * cell c are cell_subcells[c] .. cell_subcells[c+1]-1; cells ordered by their active node's
 * caller id, subcells by caller element id within a cell), element_subcell[n_elem] (caller
 * element -> subcell), active[n_nodes] (1 = the node carries degrees of freedom) and whether
 * F-bar is applied.  Any output may be NULL.  aphip_rk_get_cells then lists the subcells in
 * this order (start[n_subcells+1], volume[n_subcells]). */
int aphip_rk_get_nodal_cells(aphip_ctx *ctx, int64_t *n_cells, int64_t *n_subcells, int64_t *cell_subcells,
                             int32_t *element_subcell, uint8_t *active, int32_t *f_bar);
/* Neighbor lists in caller numbering: num[n_nodes], nbr[n_nodes x 40] (-1 padded),
 * the 40-cap truncation count (NeighborSearchProcessor.h:307-316). */
int aphip_rk_get_neighbors(aphip_ctx *ctx, int32_t *num, int32_t *nbr, int64_t *n_truncated);
/* RK function values phi[n_nodes x 40] aligned with the neighbor lists. */
int aphip_rk_get_function_values(aphip_ctx *ctx, double *phi);
/* Cell CSR in caller numbering: start[n_elem+1], nodes[nnz], dphi[nnz x 3], volume[n_elem].
 * Pass NULL arrays to query nnz only. */
int aphip_rk_get_cells(aphip_ctx *ctx, int64_t *nnz, int64_t *start, int32_t *nodes, double *dphi,
                       double *volume);

/* ---- ensembles (extension; no reference counterpart) ---------------------- */

/* Uneven-size specimens with their own material parameters and time increment,
 * advanced together by the same kernels as one element block.  Tiles never
 * span two specimens, so every specimen's results are bitwise identical to
 * running it alone in its own context.  Between begin and end, add specimens
 * in caller numbering local to the specimen; afterwards the context behaves as
 * one block whose caller numbering is the concatenation of the specimens (see
 * aphip_batch_ranges).  All specimens share the element type and material kind;
 * J2 yield stress and hardening are taken from the first specimen.  With a
 * batch, aphip_step ignores its time_increment argument. */
int aphip_batch_begin(aphip_ctx *ctx, int32_t elem_type, int32_t material_kind);
int aphip_batch_add(aphip_ctx *ctx, int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn,
                    const aphip_material_params *params, double time_increment, int32_t *specimen);
int aphip_batch_end(aphip_ctx *ctx);
int aphip_batch_ranges(aphip_ctx *ctx, int32_t specimen, int64_t *first_node, int64_t *n_nodes,
                       int64_t *first_elem, int64_t *n_elem);

/* ---- adjoints (extension; no reference counterpart) ----------------------- */

/* Vector-Jacobian product of the internal force at the current displacement:
 * given f_bar (n_nodes x 3) returns u_bar = (df/du)^T f_bar (n_nodes x 3) and the
 * per-element E_bar, nu_bar = (df/dE_e)^T f_bar, (df/dnu_e)^T f_bar (n_elem each).
 * Linear elastic and neo-Hookean only.  Any output may be NULL. */
int aphip_internal_force_vjp(aphip_ctx *ctx, const double *f_bar, double *u_bar, double *E_bar, double *nu_bar);

/* Runs n_steps of aphip_step from the context's current state (u, v, a after
 * aphip_initialize) keeping (u, v, a) checkpoints every checkpoint_interval steps. */
int aphip_adjoint_forward(aphip_ctx *ctx, double time_increment, int32_t n_steps, int32_t checkpoint_interval);

/* Reverse sweep of the last aphip_adjoint_forward: given the adjoint of the final
 * displacement (and optionally velocity), returns dL/dE_e, dL/dnu_e (n_elem, caller
 * order) and dL/dv0 (n_nodes x 3), recomputing each segment from its checkpoint.
 * The initial acceleration is taken as f(u0)/m.  Leaves the final forward state.  A new
 * tile plan (aphip_set_tiling, aphip_autotune) renumbers the nodes and drops the recording:
 * the call then fails until aphip_adjoint_forward runs again. */
int aphip_adjoint_backward(aphip_ctx *ctx, const double *u_bar_T, const double *v_bar_T, double *E_bar,
                           double *nu_bar, double *v0_bar);

/* ---- profiling ----------------------------------------------------------- */

/* Record HIP events around every launch of the step kernels on the context's stream. */
int aphip_set_profiling(aphip_ctx *ctx, int32_t enable);
/* Accumulated launches and device milliseconds of kernel class `which`
 * (0 = whole step, 1 = element/force kernel, 2 = node update kernel) since the last reset. */
int aphip_get_profile(aphip_ctx *ctx, int32_t which, int64_t *launches, double *milliseconds);
int aphip_reset_profile(aphip_ctx *ctx);
/* Measurement yardstick with no counterpart in the reference: times `reps`
 * device-to-device copies of `nbytes` (a multiple of 16) on `device` and returns
 * the device milliseconds per copy (read + write of nbytes each).  Used to state
 * achievable HBM bandwidth and to scale the FETCH_SIZE / WRITE_SIZE counters. */
int aphip_bandwidth_probe(int32_t device, int64_t nbytes, int32_t reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
