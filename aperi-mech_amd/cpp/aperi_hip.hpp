// C++ host adapter over the C ABI of libaperi_hip (include/aperi_hip.h).
//
// Mirrors the host-side interfaces of aperi-mech's explicit force path so a caller
// (and its tests) can drive the MI355X library the way it drives aperi-mech:
//   MaterialType, LinearElasticProperties, MaterialProperties  include/Materials/Base.h:20-169
//   Material, CreateMaterial                                   include/Materials/Base.h:174-323,
//                                                              include/Materials/Factory.h:43-160
//   ElementBase                                                include/ElementBase.h:27-117
//   CreateElement (supported combinations, error messages)     include/Element.h:32-86
//   InternalForceContribution::ComputeForce                    include/InternalForceContribution.h:26-146
//   TimeIncrementScheduler, StepScheduler                      include/Scheduler.h:10-121
//   TimeStepper, DirectTimeStepper, PowerMethodTimeStepper     include/TimeStepper.h:24-229
//   ExplicitSolver::Solve (time loop, output schedule)         src/Solvers/Explicit.cpp:308-498
//   PowerMethodProcessor::ComputeStableTimeIncrement           include/PowerMethodProcessor.h:331-470
//   IoMesh results output (Exodus II)                          src/IoMesh.cpp:190-240
// Errors are reported the way the reference reports them: std::runtime_error.
// Header-only; link against libaperi_hip.so.
#pragma once

#include <aperi_hip.h>

#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace aperi_hip {

inline void Check(int rc) {
    if (rc != APHIP_OK) throw std::runtime_error(aphip_last_error());
}

// ---- materials -------------------------------------------------------------

// Same enumerators and order as include/Materials/Base.h:20-29.
enum MaterialType {
    ELASTIC,
    LINEAR_ELASTIC,
    NEO_HOOKEAN,
    NEO_HOOKEAN_WITH_DAMAGE,
    PLASTIC,
    DRUCKER_PRAGER,
    POWER_LAW_CREEP,
    NONE
};

// Exactly two of the five isotropic moduli define the other three; the physical
// bounds are checked afterwards (include/Materials/Base.h:40-160).  The conversion
// formulas are the textbook isotropic-elasticity relations.
struct LinearElasticProperties {
    static constexpr double kUnset = std::numeric_limits<double>::quiet_NaN();
    double shear_modulus = kUnset;
    double lambda = kUnset;
    double youngs_modulus = kUnset;
    double poisson_ratio = kUnset;
    double bulk_modulus = kUnset;

    void CompleteProperties() {
        const bool K = !std::isnan(bulk_modulus), E = !std::isnan(youngs_modulus), L = !std::isnan(lambda),
                   G = !std::isnan(shear_modulus), nu = !std::isnan(poisson_ratio);
        const int defined = K + E + L + G + nu;
        if (defined != 2)
            throw std::runtime_error("Exactly two properties must be defined. Defined: " + std::to_string(defined));
        double &k = bulk_modulus, &e = youngs_modulus, &l = lambda, &g = shear_modulus, &v = poisson_ratio;
        if (K && E) {
            v = (3 * k - e) / (6 * k);
            g = e / (2 * (1 + v));
            l = (e * v) / ((1 + v) * (1 - 2 * v));
        } else if (K && L) {
            g = (3 * (k - l)) / 2;
            v = l / (2 * (l + g));
            e = 2 * g * (1 + v);
        } else if (K && G) {
            l = k - (2.0 / 3.0) * g;
            v = l / (2 * (l + g));
            e = 2 * g * (1 + v);
        } else if (K && nu) {
            e = 3 * k * (1 - 2 * v);
            g = e / (2 * (1 + v));
            l = (e * v) / ((1 + v) * (1 - 2 * v));
        } else if (E && L) {
            // nu from 2 l nu^2 + (l + E) nu - l = 0, the root in (-1, 0.5)
            const double a = 2 * l, b = l + e, c = -l;
            v = (-b + std::sqrt(b * b - 4 * a * c)) / (2 * a);
            g = e / (2 * (1 + v));
            k = l + (2.0 / 3.0) * g;
        } else if (E && G) {
            v = e / (2 * g) - 1;
            l = (e * v) / ((1 + v) * (1 - 2 * v));
            k = l + (2.0 / 3.0) * g;
        } else if (E && nu) {
            g = e / (2 * (1 + v));
            l = (e * v) / ((1 + v) * (1 - 2 * v));
            k = e / (3 * (1 - 2 * v));
        } else if (L && G) {
            k = l + (2.0 / 3.0) * g;
            v = l / (2 * (l + g));
            e = 2 * g * (1 + v);
        } else if (L && nu) {
            e = (l * (1 + v) * (1 - 2 * v)) / v;
            g = e / (2 * (1 + v));
            k = l + (2.0 / 3.0) * g;
        } else {  // G and nu
            e = 2 * g * (1 + v);
            l = (2 * g * v) / (1 - 2 * v);
            k = (2 * g * (1 + v)) / (3 * (1 - 2 * v));
        }
        CheckPhysicalBounds();
    }

   private:
    void CheckPhysicalBounds() const {
        if (poisson_ratio < -1.0 || poisson_ratio >= 0.5)
            throw std::runtime_error("Poisson's ratio violation: " + std::to_string(poisson_ratio) +
                                     " must be in [-1, 0.5)");
        auto positive = [](double x, const char *name) {
            if (x <= 0.0) throw std::runtime_error(std::string(name) + " must be positive: " + std::to_string(x));
        };
        positive(youngs_modulus, "Young's modulus");
        positive(shear_modulus, "Shear modulus");
        positive(bulk_modulus, "Bulk modulus");
        if (lambda <= -(2.0 / 3.0) * shear_modulus)
            throw std::runtime_error("Lamé's parameter violation: λ=" + std::to_string(lambda) +
                                     " must be > -2G/3=" + std::to_string(-(2.0 / 3.0) * shear_modulus));
    }
};

struct MaterialProperties {
    MaterialType material_type = NONE;
    double density = std::numeric_limits<double>::quiet_NaN();
    LinearElasticProperties linear_elastic_properties;
    // "yield_stress", "hardening_modulus" for PLASTIC; "I1_critical", "I1_failure", "alpha" and
    // the optional "maximum_damage", "maximum_damage_increment", "stop_failing_at_time",
    // "fail_all_at_time" for NEO_HOOKEAN_WITH_DAMAGE (include/Materials/Factory.h:70-101)
    std::map<std::string, double> properties;
};

// Host side of a material: its kind on the device and its completed constants.
class Material {
   public:
    explicit Material(std::shared_ptr<MaterialProperties> p) : m_props(std::move(p)) {
        if (!m_props) throw std::runtime_error("Material properties are null");
        m_props->linear_elastic_properties.CompleteProperties();
        switch (m_props->material_type) {
            case LINEAR_ELASTIC: m_kind = APHIP_LINEAR_ELASTIC; break;
            case NEO_HOOKEAN: m_kind = APHIP_NEO_HOOKEAN; break;
            case PLASTIC: m_kind = APHIP_J2; break;
            case ELASTIC: m_kind = APHIP_SVK; break;
            case NEO_HOOKEAN_WITH_DAMAGE: m_kind = APHIP_NEO_HOOKEAN_DAMAGE; break;
            case POWER_LAW_CREEP: m_kind = APHIP_POWER_LAW_CREEP; break;
            case DRUCKER_PRAGER: m_kind = APHIP_DRUCKER_PRAGER; break;
            default: throw std::runtime_error("Material type not supported by the MI355X force step");
        }
        if (m_kind == APHIP_J2) {
            m_yield = Required("yield_stress");
            m_hardening = Required("hardening_modulus");
        }
        if (m_kind == APHIP_NEO_HOOKEAN_DAMAGE) {
            m_damage[0] = Required("I1_critical");
            m_damage[1] = Required("I1_failure");
            m_damage[2] = Required("alpha");
            m_damage[3] = Optional("maximum_damage", 0.99);
            m_damage[4] = Optional("maximum_damage_increment", 1.0);
            m_damage[5] = Optional("stop_failing_at_time", -1.0);
            m_damage[6] = Optional("fail_all_at_time", -1.0);
        }
        if (m_kind == APHIP_POWER_LAW_CREEP) {  // Materials/Factory.h:134-153
            m_creep[0] = Required("A");
            m_creep[1] = Required("n");
            m_creep[2] = Required("m");
            m_creep[3] = Optional("constant_temperature", 0.0);
            m_creep_const_t = Optional("use_constant_temperature", 0.0) != 0.0 ? 1 : 0;
        }
        if (m_kind == APHIP_DRUCKER_PRAGER) {  // Materials/Factory.h:118-132
            m_dp[0] = Required("A1");
            m_dp[1] = Required("A2");
            m_dp[2] = Required("A2G");
        }
    }
    MaterialType GetMaterialType() const { return m_props->material_type; }
    std::shared_ptr<MaterialProperties> GetMaterialProperties() const { return m_props; }
    double GetDensity() const { return m_props->density; }
    bool HasState() const {
        return m_kind == APHIP_J2 || m_kind == APHIP_NEO_HOOKEAN_DAMAGE || m_kind == APHIP_POWER_LAW_CREEP ||
               m_kind == APHIP_DRUCKER_PRAGER;
    }
    // the reference's counts (NeoHookeWithDamage.h:279-282 reports 1, the damage;
    // PowerLawCreep.h:186-189 and DruckerPrager.h:224-227 report 2)
    size_t NumberOfStateVariables() const {
        return (m_kind == APHIP_POWER_LAW_CREEP || m_kind == APHIP_DRUCKER_PRAGER) ? 2 : HasState() ? 1 : 0;
    }
    bool NeedsVelocityGradient() const { return m_kind == APHIP_POWER_LAW_CREEP || m_kind == APHIP_DRUCKER_PRAGER; }
    bool NeedsOldStress() const { return NeedsVelocityGradient(); }
    // values per element of the state field (the damage model also stores I1)
    size_t StateFieldWidth() const { return m_kind == APHIP_NEO_HOOKEAN_DAMAGE ? 2 : NumberOfStateVariables(); }
    int Kind() const { return m_kind; }
    aphip_material_params Params(int state_per_qp = 0) const {
        const LinearElasticProperties &le = m_props->linear_elastic_properties;
        return aphip_material_params{m_props->density, le.youngs_modulus, le.poisson_ratio, m_yield, m_hardening,
                                     state_per_qp, 0, m_damage[0], m_damage[1], m_damage[2], m_damage[3],
                                     m_damage[4], m_damage[5], m_damage[6], m_creep[0], m_creep[1], m_creep[2],
                                     m_creep[3], m_creep_const_t, 0, m_dp[0], m_dp[1], m_dp[2]};
    }

   private:
    double Required(const char *name) const {
        const auto it = m_props->properties.find(name);
        if (it == m_props->properties.end()) throw std::runtime_error(std::string("Missing material property: ") + name);
        return it->second;
    }
    double Optional(const char *name, double fallback) const {
        const auto it = m_props->properties.find(name);
        return it == m_props->properties.end() ? fallback : it->second;
    }
    std::shared_ptr<MaterialProperties> m_props;
    int m_kind = -1;
    double m_yield = 0.0, m_hardening = 0.0;
    double m_damage[7] = {0, 0, 0, 0, 0, 0, 0};
    double m_creep[4] = {0, 0, 0, 0};
    double m_dp[3] = {0, 0, 0};
    int32_t m_creep_const_t = 0;
};

inline std::shared_ptr<Material> CreateMaterial(std::shared_ptr<MaterialProperties> p) {
    return std::make_shared<Material>(std::move(p));
}

// ---- device context --------------------------------------------------------

// One library context: one HIP stream on one device.
class Device {
   public:
    explicit Device(int device = 0, int flags = 0) {
        const aphip_config cfg{device, flags};
        Check(aphip_create(&m_ctx, &cfg));
    }
    ~Device() { aphip_destroy(m_ctx); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    aphip_ctx *get() const { return m_ctx; }
    static int Count() {
        int32_t n = 0;
        Check(aphip_device_count(&n));
        return n;
    }

   private:
    aphip_ctx *m_ctx = nullptr;
};

// ---- elements --------------------------------------------------------------

enum class ElementTopology { Tetrahedron4, Hexahedron8 };       // include/Constants.h:23-25
enum class ApproximationSpaceType { FiniteElement, ReproducingKernel };
enum class IntegrationSchemeType { GaussQuadrature, StrainSmoothing };
// include/Constants.h:33-36 (Incremental is not supported here)
enum class LagrangianFormulationType { Total, Incremental, Updated, Semi };
// include/MeshLabelerParameters.h:15-29: the smoothing cells of a strain-smoothed part and
// their subdomains (nodal cells: 0 = every element a subcell, 1 = the cell one subcell)
enum class SmoothingCellType { None, Nodal, Element };
struct MeshLabelerParameters {
    SmoothingCellType smoothing_cell_type = SmoothingCellType::None;
    size_t num_subcells = 1;
};

// An element block in the caller's numbering: coordinates n x 3, connectivity
// n_elem x nodes_per_element (0-based).
struct MeshBlock {
    std::vector<double> coordinates;
    std::vector<int32_t> connectivity;
    size_t NumNodes() const { return coordinates.size() / 3; }
};

// include/ElementBase.h:27-117.  Fields are exchanged in the caller's numbering and
// the reference's layout (node vectors n x 3, element tensors n_elem x 9 row-major).
class ElementBase {
   public:
    ElementBase(size_t num_nodes, std::shared_ptr<Device> device, std::shared_ptr<Material> material)
        : m_num_nodes(num_nodes), m_device(std::move(device)), m_material(std::move(material)) {}
    virtual ~ElementBase() = default;
    virtual bool CheckIfUpdateIsNeeded() const { return false; }
    virtual void UpdateShapeFunctions() {}
    size_t GetNumNodes() const { return m_num_nodes; }
    virtual void ComputeInternalForceAllElements(double time_increment, double time) = 0;
    virtual void PopulateElementOutputs() {}
    virtual void FinishPreprocessing() {}
    std::shared_ptr<Material> GetMaterial() const { return m_material; }
    aphip_ctx *Context() const { return m_device->get(); }
    size_t NumMeshNodes() const { return m_mesh_nodes; }
    size_t NumElements() const { return m_elements; }

    void SetField(int field, const std::vector<double> &values) {
        Check(FieldSize(field) == values.size() ? APHIP_OK : APHIP_ERR_INVALID, "field size mismatch");
        Check(aphip_set_field(Context(), field, values.data()));
    }
    std::vector<double> GetField(int field) const {
        std::vector<double> out(FieldSize(field));
        Check(aphip_get_field(Context(), field, out.data()));
        return out;
    }
    // Velocity and acceleration held at zero on the DOFs marked 1 (n x 3 bytes).
    void SetFixedDofs(const std::vector<uint8_t> &mask) {
        Check(mask.size() == 3 * m_mesh_nodes ? APHIP_OK : APHIP_ERR_INVALID, "fixed-DOF mask size mismatch");
        Check(aphip_set_fixed(Context(), mask.data()));
    }
    double ComputeMass() {
        double total = 0.0;
        Check(aphip_compute_mass(Context(), &total));
        return total;
    }

   protected:
    static void Check(int rc, const char *msg = nullptr) {
        if (rc == APHIP_OK) return;
        throw std::runtime_error(msg ? std::string(msg) : std::string(aphip_last_error()));
    }
    size_t FieldSize(int field) const {
        switch (field) {
            case APHIP_FIELD_U: case APHIP_FIELD_V: case APHIP_FIELD_A: case APHIP_FIELD_F:
            case APHIP_FIELD_RK_DISPLACEMENT: case APHIP_FIELD_RK_FORCE_LOCAL: return 3 * m_mesh_nodes;
            case APHIP_FIELD_MASS: return m_mesh_nodes;
            case APHIP_FIELD_H: case APHIP_FIELD_P: return 9 * m_elements;
            case APHIP_FIELD_VOLUME: return m_elements;
            case APHIP_FIELD_STATE: return m_elements * m_material->StateFieldWidth();
            default: throw std::runtime_error("unknown field");
        }
    }
    void BindMesh(int etype, const MeshBlock &mesh) {
        const size_t npe = etype == APHIP_TET4 ? 4 : 8;
        if (mesh.coordinates.size() % 3 || mesh.connectivity.size() % npe)
            throw std::runtime_error("mesh arrays do not match the element topology");
        m_mesh_nodes = mesh.NumNodes();
        m_elements = mesh.connectivity.size() / npe;
        Check(aphip_set_mesh(Context(), etype, (int64_t)m_mesh_nodes, mesh.coordinates.data(), (int64_t)m_elements,
                             mesh.connectivity.data()));
        const aphip_material_params p = m_material->Params();
        Check(aphip_set_material(Context(), m_material->Kind(), &p));
    }

    size_t m_num_nodes;
    std::shared_ptr<Device> m_device;
    std::shared_ptr<Material> m_material;
    size_t m_mesh_nodes = 0, m_elements = 0;
};

// ElementTetrahedron4 / ElementHexahedron8 with Gauss quadrature
// (include/ElementTetrahedron4.h:33-204, include/ElementHexahedron8.h:31-192).
class ElementGaussianHip : public ElementBase {
   public:
    ElementGaussianHip(ElementTopology topology, const MeshBlock &mesh, std::shared_ptr<Device> device,
                       std::shared_ptr<Material> material,
                       LagrangianFormulationType formulation = LagrangianFormulationType::Total,
                       int semi_update_interval = 1)
        : ElementBase(topology == ElementTopology::Tetrahedron4 ? 4 : 8, std::move(device), std::move(material)) {
        BindMesh(topology == ElementTopology::Tetrahedron4 ? APHIP_TET4 : APHIP_HEX8, mesh);
        if (formulation == LagrangianFormulationType::Updated)
            Check(aphip_set_formulation(Context(), APHIP_UPDATED_LAGRANGIAN));
        else if (formulation == LagrangianFormulationType::Semi)
            Check(aphip_set_formulation_semi(Context(), semi_update_interval));
        else if (formulation != LagrangianFormulationType::Total)
            throw std::runtime_error("Lagrangian formulation not supported by the MI355X force step");
    }
    void ComputeInternalForceAllElements(double time_increment, double time) override {
        Check(aphip_internal_force(Context(), time, time_increment));
    }
};

// ElementReproducingKernelTet4 / ElementReproducingKernelHex8, element-as-cell strain
// smoothing (include/ElementReproducingKernel.h:67-80, :214-224, :317-378), one-pass by
// default (IntegrationSchemeParameters::uses_one_pass_method); false selects the
// two-pass force (gather, cell force on the physical displacement, scatter).
class ElementReproducingKernelHip : public ElementBase {
   public:
    ElementReproducingKernelHip(const MeshBlock &mesh, double kernel_radius_scale_factor,
                                std::shared_ptr<Device> device, std::shared_ptr<Material> material,
                                bool use_one_pass_method = true,
                                ElementTopology topology = ElementTopology::Tetrahedron4)
        : ElementBase(topology == ElementTopology::Tetrahedron4 ? 4 : 8, std::move(device), std::move(material)) {
        BindMesh(topology == ElementTopology::Tetrahedron4 ? APHIP_TET4 : APHIP_HEX8, mesh);
        Check(aphip_rk_setup_method(Context(), kernel_radius_scale_factor, use_one_pass_method ? 1 : 0));
    }
    void ComputeInternalForceAllElements(double time_increment, double time) override {
        Check(aphip_internal_force(Context(), time, time_increment));
    }
    // Neighbor lists in caller numbering: counts per node, ids n x 40 (-1 padded);
    // returns how many nodes had more than 40 candidates.
    int64_t GetNeighbors(std::vector<int32_t> &counts, std::vector<int32_t> &neighbors) const {
        counts.assign(m_mesh_nodes, 0);
        neighbors.assign(m_mesh_nodes * 40, -1);
        int64_t truncated = 0;
        Check(aphip_rk_get_neighbors(Context(), counts.data(), neighbors.data(), &truncated));
        return truncated;
    }
};

// ElementReproducingKernelHex8 with nodal smoothing cells on a thex mesh (the reference's
// nodal_smoothing_cell, include/MeshLabelerProcessor.h:227-330, :690-827) and, with several
// subcells per cell, optionally F-bar (include/ComputeInternalForceSmoothedCell.h:120-413).
class ElementNodalReproducingKernelHip : public ElementBase {
   public:
    ElementNodalReproducingKernelHip(const MeshBlock &mesh, double kernel_radius_scale_factor,
                                     std::shared_ptr<Device> device, std::shared_ptr<Material> material,
                                     size_t num_subcells, bool use_f_bar)
        : ElementBase(8, std::move(device), std::move(material)) {
        BindMesh(APHIP_HEX8, mesh);
        Check(aphip_rk_nodal_setup(Context(), kernel_radius_scale_factor, (int32_t)num_subcells, use_f_bar ? 1 : 0));
    }
    void ComputeInternalForceAllElements(double time_increment, double time) override {
        Check(aphip_internal_force(Context(), time, time_increment));
    }
    // Nodes carrying degrees of freedom (1) and whether F-bar is applied (only when the
    // cells hold more than one subcell).
    bool UsesFBar(std::vector<uint8_t> *active = nullptr) const {
        int32_t fbar = 0;
        if (active) active->assign(m_mesh_nodes, 0);
        Check(aphip_rk_get_nodal_cells(Context(), nullptr, nullptr, nullptr, nullptr,
                                       active ? active->data() : nullptr, &fbar));
        return fbar != 0;
    }
};

// ElementSmoothedTetrahedron4: finite-element strain smoothing, one cell per Tet4
// (include/ElementSmoothedTetrahedron4.h; created by include/Element.h:44-47).
class ElementSmoothedTetrahedron4Hip : public ElementBase {
   public:
    ElementSmoothedTetrahedron4Hip(const MeshBlock &mesh, std::shared_ptr<Device> device,
                                   std::shared_ptr<Material> material)
        : ElementBase(4, std::move(device), std::move(material)) {
        BindMesh(APHIP_TET4, mesh);
        Check(aphip_strain_smoothing_setup(Context()));
    }
    void ComputeInternalForceAllElements(double time_increment, double time) override {
        Check(aphip_internal_force(Context(), time, time_increment));
    }
};

// include/Element.h:32-86: the supported (topology, approximation, integration)
// combinations and the reference's messages for the unsupported ones.
inline std::shared_ptr<ElementBase> CreateElement(ElementTopology topology, ApproximationSpaceType space,
                                                  IntegrationSchemeType scheme, const MeshBlock &mesh,
                                                  std::shared_ptr<Device> device, std::shared_ptr<Material> material,
                                                  double kernel_radius_scale_factor = 0.0,
                                                  bool use_one_pass_method = true,
                                                  LagrangianFormulationType formulation = LagrangianFormulationType::Total,
                                                  const MeshLabelerParameters &mesh_labeler = MeshLabelerParameters{},
                                                  bool use_f_bar = false) {
    if (space == ApproximationSpaceType::ReproducingKernel) {
        if (scheme != IntegrationSchemeType::StrainSmoothing)
            throw std::runtime_error("Gauss Quadrature is not supported for Reproducing Kernel");
        if (formulation != LagrangianFormulationType::Total)
            throw std::runtime_error("the updated formulation applies to the Gaussian elements, not to RKPM cells");
        if (mesh_labeler.smoothing_cell_type == SmoothingCellType::Nodal) {
            if (topology != ElementTopology::Hexahedron8)
                throw std::runtime_error(
                    "Nodal integration will only work correctly when being built from hexahedral elements. Use "
                    "'thex' to divide a tetrahedral mesh into hexahedral elements.");
            return std::make_shared<ElementNodalReproducingKernelHip>(mesh, kernel_radius_scale_factor,
                                                                      std::move(device), std::move(material),
                                                                      mesh_labeler.num_subcells, use_f_bar);
        }
        return std::make_shared<ElementReproducingKernelHip>(mesh, kernel_radius_scale_factor, std::move(device),
                                                             std::move(material), use_one_pass_method, topology);
    }
    if (scheme == IntegrationSchemeType::StrainSmoothing) {
        if (topology == ElementTopology::Hexahedron8)
            throw std::runtime_error("Strain Smoothing is not supported for finite elements and Hexahedron8");
        if (formulation != LagrangianFormulationType::Total)
            throw std::runtime_error("the updated formulation applies to the Gaussian elements, not to smoothed cells");
        return std::make_shared<ElementSmoothedTetrahedron4Hip>(mesh, std::move(device), std::move(material));
    }
    return std::make_shared<ElementGaussianHip>(topology, mesh, std::move(device), std::move(material), formulation);
}

// include/InternalForceContribution.h:26-146: the force contribution of one block.
class InternalForceContribution {
   public:
    explicit InternalForceContribution(std::shared_ptr<ElementBase> element) : m_element(std::move(element)) {
        if (!m_element) throw std::runtime_error("InternalForceContribution needs an element");
    }
    void Preprocess() {}
    void FinishPreprocessing() { m_element->FinishPreprocessing(); }
    bool CheckIfUpdateIsNeeded() const { return m_element->CheckIfUpdateIsNeeded(); }
    void UpdateShapeFunctions() { m_element->UpdateShapeFunctions(); }
    void ComputeForce(double time, double time_increment) {
        m_element->ComputeInternalForceAllElements(time_increment, time);
    }
    std::shared_ptr<ElementBase> GetElement() const { return m_element; }

   private:
    std::shared_ptr<ElementBase> m_element;
};

// The explicit loop of src/Solvers/Explicit.cpp:308-498 with a fixed increment
// (DirectTimeStepper, include/TimeStepper.h:97-128), run on the device: the
// initial force and acceleration, then central-difference steps.
// ---- results output ------------------------------------------------------------

// The results half of the reference's IoMesh (src/IoMesh.cpp:190-240): an Exodus II
// file of the element's mesh with the displacement, velocity, acceleration, force, mass
// and essential_boundary node variables and the displacement_gradient, pk1_stress and
// volume element variables (written by the library, netCDF 64-bit-offset format).
class IoMesh {
   public:
    explicit IoMesh(std::shared_ptr<ElementBase> element) : m_element(std::move(element)) {}
    ~IoMesh() {
        if (m_open) (void)aphip_results_close(m_element->Context());
    }
    void CreateFieldResultsFile(const std::string &filename, const std::string &title = "") {
        Check(aphip_results_open(m_element->Context(), filename.c_str(), title.c_str()));
        m_open = true;
    }
    void WriteFieldResults(double time) const { Check(aphip_results_write(m_element->Context(), time)); }
    void CloseFieldResultsFile() {
        Check(aphip_results_close(m_element->Context()));
        m_open = false;
    }

   private:
    std::shared_ptr<ElementBase> m_element;
    bool m_open = false;
};

// ---- schedulers and time steppers --------------------------------------------

// An event every time_increment from time_start; several events that one call has
// passed count once (include/Scheduler.h:37-63).
class TimeIncrementScheduler {
   public:
    TimeIncrementScheduler(double time_start, double time_end, double time_increment)
        : m_time_end(time_end), m_next(time_start), m_increment(time_increment) {}
    bool AtNextEvent(double time) {
        const double tolerance = 1.0e-15;
        const bool at_event = time + tolerance >= m_next;
        if (at_event) {
            m_next += m_increment;
            AtNextEvent(time);
            if (m_next > m_time_end) m_next = std::numeric_limits<double>::max();
        }
        return at_event;
    }
    double GetTimeEnd() const { return m_time_end; }
    double GetNextEventTime() const { return m_next; }

   private:
    double m_time_end, m_next, m_increment;
};

// An event at steps start, start + increment, ... (include/Scheduler.h:100-121).
class StepScheduler {
   public:
    StepScheduler(size_t start, size_t increment) : m_next(start), m_increment(increment) {}
    bool AtNextEvent(size_t step) {
        const bool at_event = step == m_next;
        if (at_event) m_next += m_increment;
        return at_event;
    }
    size_t GetNextEventTime() const { return m_next; }

   private:
    size_t m_next, m_increment;
};

class ExplicitSolver;

// include/TimeStepper.h:24-90
class TimeStepper {
   public:
    explicit TimeStepper(double time_end) : m_time_end(time_end) {}
    virtual ~TimeStepper() = default;
    virtual void Initialize(ExplicitSolver & /*solver*/) {}
    double GetTimeEnd() const { return m_time_end; }
    bool AtEnd(double time) const { return time >= GetTimeEnd(); }
    virtual double GetTimeIncrement(double time, size_t increment) = 0;

   protected:
    // the step that ends at time_end exactly (TimeStepper.h:113-121, :207-214)
    double Clamp(double time, double time_increment) const {
        if (time + time_increment <= GetTimeEnd()) return time_increment;
        if (time > GetTimeEnd()) return 0.0;
        return GetTimeEnd() - time;
    }

   private:
    double m_time_end;
};

// A fixed increment (include/TimeStepper.h:97-128).
class DirectTimeStepper : public TimeStepper {
   public:
    DirectTimeStepper(double time_increment, double time_end) : TimeStepper(time_end), m_time_increment(time_increment) {}
    double GetTimeIncrement(double time, size_t /*increment*/) override { return Clamp(time, m_time_increment); }

   private:
    double m_time_increment;
};

// scale_factor x the power-method increment, recomputed every time_update_interval
// steps from step 0 (include/TimeStepper.h:138-216).
class PowerMethodTimeStepper : public TimeStepper {
   public:
    PowerMethodTimeStepper(double scale_factor, size_t time_update_interval, double time_end)
        : TimeStepper(time_end), m_scale_factor(scale_factor), m_scheduler(0, time_update_interval) {}
    void Initialize(ExplicitSolver &solver) override { m_solver = &solver; }
    double GetTimeIncrement(double time, size_t increment) override;
    // (increment, full increment) of every update so far
    const std::vector<std::pair<size_t, double>> &Updates() const { return m_updates; }

   private:
    double m_scale_factor;
    double m_time_increment = 0.0;
    StepScheduler m_scheduler;
    ExplicitSolver *m_solver = nullptr;
    std::vector<std::pair<size_t, double>> m_updates;
};

class ExplicitSolver {
   public:
    explicit ExplicitSolver(std::shared_ptr<InternalForceContribution> contribution)
        : m_contribution(std::move(contribution)) {}
    // Mass, the initial force and a = f/m (Explicit.cpp:362-376).
    double Initialize() {
        const double mass = m_contribution->GetElement()->ComputeMass();
        Check(aphip_initialize(Context()));
        m_initialized = true;
        return mass;
    }
    // n steps of the loop body (Explicit.cpp:383-452) with a fixed increment; returns the
    // time reached.
    double Solve(double time_increment, int n_steps) {
        if (!m_initialized) Initialize();
        Check(aphip_step(Context(), time_increment, n_steps));
        m_time += time_increment * n_steps;
        return m_time;
    }
    // The time loop of ExplicitSolver::Solve (src/Solvers/Explicit.cpp:308-452): the first
    // increment is requested before the initial force; then, until the stepper's end time,
    // the increment for step n, one step, time += increment; write_output(time, n) runs at
    // t = 0 and after every step at which the output scheduler fires.  Returns the final time.
    double Solve(TimeStepper &time_stepper, TimeIncrementScheduler &output_scheduler,
                 const std::function<void(double, size_t)> &write_output) {
        double time = 0.0;
        size_t n = 0;
        if (!m_initialized) m_contribution->GetElement()->ComputeMass();  // BuildMassMatrix (:313)
        time_stepper.Initialize(*this);
        double time_increment = time_stepper.GetTimeIncrement(time, n);
        if (!m_initialized) {
            Check(aphip_initialize(Context()));
            m_initialized = true;
        }
        if (output_scheduler.AtNextEvent(time)) write_output(time, n);
        while (!time_stepper.AtEnd(time)) {
            time_increment = time_stepper.GetTimeIncrement(time, n);
            Check(aphip_step(Context(), time_increment, 1));
            time += time_increment;
            ++n;
            if (output_scheduler.AtNextEvent(time)) write_output(time, n);
        }
        m_time = time;
        return time;
    }
    double GetTime() const { return m_time; }
    // PowerMethodProcessor::ComputeStableTimeIncrement (include/PowerMethodProcessor.h:331-470)
    // with the reference's defaults about the current displacement; time_increment is the
    // last full increment (0 on the first call); the simulation state is left as it was.
    double ComputeStableTimeIncrement(double time_increment = 0.0, int num_iterations = 50, double epsilon = 1e-5,
                                      double tolerance = 1e-3, double variation_tolerance = 1e-7,
                                      uint64_t seed = 20261202) {
        double dt = 0.0;
        Check(aphip_power_method_time_increment(Context(), time_increment, num_iterations, epsilon, tolerance,
                                                variation_tolerance, seed, &dt, nullptr, nullptr));
        return dt;
    }
    aphip_ctx *Context() const { return m_contribution->GetElement()->Context(); }

   private:
    std::shared_ptr<InternalForceContribution> m_contribution;
    bool m_initialized = false;
    double m_time = 0.0;
};

inline double PowerMethodTimeStepper::GetTimeIncrement(double time, size_t increment) {
    if (m_solver == nullptr) throw std::runtime_error("PowerMethodTimeStepper used before Initialize");
    if (m_scheduler.AtNextEvent(increment)) {
        const double full = m_solver->ComputeStableTimeIncrement(m_time_increment / m_scale_factor);
        m_time_increment = full * m_scale_factor;
        m_updates.emplace_back(increment, full);
    }
    return Clamp(time, m_time_increment);
}

}  // namespace aperi_hip
