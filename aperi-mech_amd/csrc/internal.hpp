// Internal context and helpers of the aperi-hip library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "aperi_hip.h"
#include "device_math.hpp"
#include "exodus_format.hpp"

namespace aphip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define APHIP_HIP(x)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess)                                                                         \
            throw ::aphip::Error(APHIP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define APHIP_REQUIRE(cond, msg)                                                                      \
    do {                                                                                              \
        if (!(cond)) throw ::aphip::Error(APHIP_ERR_INVALID, msg);                                    \
    } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        if (hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * count) != hipSuccess) {
            p = nullptr;
            throw Error(APHIP_ERR_ALLOC, "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed");
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void zero(hipStream_t s) {
        if (p) APHIP_HIP(hipMemsetAsync(p, 0, sizeof(T) * n, s));
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) APHIP_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t count, hipStream_t s) const {
        if (count) APHIP_HIP(hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    }
};

// Material constants resolved on the host (Lame form).
struct MaterialConst {
    int kind = -1;
    double density = 0.0;
    double lambda = 0.0, mu = 0.0;
    double yield_stress = 0.0, hardening = 0.0;
    double E = 0.0, nu = 0.0;
    int state_per_qp = 0;
    // neo-Hookean with damage (increment and maximum already clamped to [0, 1])
    double I1_critical = 0.0, I1_failure = 0.0, alpha = 0.0, max_damage = 0.0, max_damage_increment = 0.0;
    // power-law creep and Drucker-Prager: bulk modulus; creep rate factor A exp(-m / T), exponent
    double bulk = 0.0, creep_rate = 0.0, creep_n = 0.0;
    double dp_A1 = 0.0, dp_A2 = 0.0, dp_A2G = 0.0;
    MatParams params() const {
        return MatParams{lambda, mu,     yield_stress, hardening, I1_critical, I1_failure, alpha, max_damage,
                         max_damage_increment, bulk, creep_rate, creep_n, dp_A1, dp_A2, dp_A2G};
    }
    // reads the previous displacement gradient and stress (velocity gradient, old stress)
    bool rate_form() const { return kind == APHIP_POWER_LAW_CREEP || kind == APHIP_DRUCKER_PRAGER; }
    // state values per evaluation point
    int state_width() const {
        return kind == APHIP_J2 ? 1 : (kind == APHIP_NEO_HOOKEAN_DAMAGE || rate_form()) ? 2 : 0;
    }
};

// Checkpointed forward run for the adjoint (adjoint.hip).
struct AdjointData {
    bool ready = false;
    double dt = 0.0;
    int n_steps = 0, interval = 0;
    std::vector<DevBuf<double> *> ckpt;  // (u, v, a) at steps 0, K, 2K, ...; 9 doubles per node
    DevBuf<double> final_state;          // (u, v, a) at the last step
    DevBuf<double> seg_u;                // recomputed u of one segment, K x 3 x nn
    DevBuf<double> ub, vb, ab;           // adjoint node state
    DevBuf<double> lam_bar, mu_bar;      // per-element parameter adjoints
    void reset() {
        for (auto *b : ckpt) delete b;
        ckpt.clear();
        final_state.release();
        seg_u.release();
        ub.release(); vb.release(); ab.release(); lam_bar.release(); mu_bar.release();
        ready = false;
    }
    ~AdjointData() { reset(); }
};

// Smoothed-cell (RKPM) data in internal numbering.
struct CellData {
    bool ready = false;
    int64_t nnz = 0;
    int64_t n_truncated = 0;
    DevBuf<int32_t> num_nbr;    // [nn]
    DevBuf<int32_t> nbr;        // [nn * 40] internal ids, -1 padded
    DevBuf<double> phi;         // [nn * 40]
    DevBuf<double> radius;      // [nn]
    DevBuf<int64_t> start;      // [ne + 1] cell entry ranges
    DevBuf<int32_t> nodes;      // [nnz] internal node ids
    DevBuf<double> dphi;        // [nnz * 3]
    DevBuf<double> volume;      // [ne]
    DevBuf<int32_t> inc_cell;   // [nnz] cell owning each entry
    DevBuf<int64_t> node_start; // [nn + 1] node -> entries naming it
    DevBuf<int32_t> node_inc;   // unused (kept for layout stability)
    DevBuf<int32_t> entry_pos;  // [nnz] position of each entry in node order (by node, then ascending entry)
    DevBuf<double> g;           // [nnz * 3] two-pass: per-entry force contributions of the last evaluation, node order
    DevBuf<int64_t> rstart;     // [nn + 1] reverse neighbor lists: points i that list node J
    DevBuf<int32_t> rsrc;       // [nrev] i, ascending per node
    DevBuf<double> rphi;        // [nrev] phi_iJ
    int64_t ncells = 0;         // cells of the CSR above: the elements, or the subcells of nodal cells
    // nodal smoothing cells on a thex mesh: the CSR cells above are subcells, grouped into
    // one cell per active node; nodes with degrees of freedom come first (Ctx::n_dof)
    bool nodal = false;
    bool fbar = false;              // F-bar over the subcells of each cell
    int64_t n_fcells = 0;           // nodal cells
    DevBuf<int64_t> fcell_sub;      // [n_fcells + 1] subcell range of each nodal cell
    DevBuf<double> fcell_volume;    // [n_fcells]
    DevBuf<int32_t> sub_fcell;      // [ncells] F-bar: nodal cell of each subcell
    DevBuf<double> evol;            // [ne] smoothed volume of every element (nodal cells)
    std::vector<int32_t> sub_of_elem;     // [ne] subcell of each internal element (nodal cells)
    std::vector<int32_t> sub_first_elem;  // [ncells] first internal element of each subcell
    std::vector<int64_t> fcell_sub_h;     // host copy of fcell_sub
    // two-pass method: cells list the element's own nodes; the cell force acts on
    // u_phys = Phi d and f = Phi^T f_local
    bool two_pass = false;
    DevBuf<double> u_phys;      // [nn * 3] physical displacement of the last evaluation
    DevBuf<double> f_local;     // [nn * 3] cell force before the scatter
    // one-pass method on cell tiles (runs of consecutive cells, one wave each)
    bool tiled = false;
    int tile_cells = 0;             // cells per tile
    int tile_maxtn = 0;             // largest tile node set
    int tile_lpc = 8;               // lanes per cell of the tile kernel
    int64_t ntiles = 0, tile_total = 0;
    DevBuf<uint16_t> lidx;          // [nnz] entry -> index in its tile's node set
    int tile_maxsub = 0;            // most cells in one tile
    DevBuf<int64_t> t_cstart;       // [ntiles + 1] first cell of each tile
    DevBuf<int32_t> t_fstart;       // F-bar: [ntiles + 1] first nodal cell of each tile
    DevBuf<int32_t> t_nstart;       // [ntiles + 1]
    DevBuf<int32_t> t_node;         // [tile_total] internal node ids, ascending per tile
    DevBuf<int64_t> nslot_start;    // [nn + 1] slot range of each node
    DevBuf<int32_t> nslot;          // [tile_total] slots of each node, ascending tile
    DevBuf<double> slots;           // [tile_total * 3] tile force rows
    void reset() {
        ready = false;
        two_pass = false;
        tiled = false;
        tile_cells = tile_maxtn = 0;
        ntiles = tile_total = 0;
        lidx.release(); t_nstart.release(); t_node.release(); nslot_start.release(); nslot.release(); slots.release();
        t_cstart.release(); t_fstart.release();
        tile_maxsub = 0;
        rstart.release(); rsrc.release(); rphi.release(); u_phys.release(); f_local.release();
        nnz = n_truncated = 0;
        ncells = n_fcells = 0;
        nodal = fbar = false;
        fcell_sub.release(); fcell_volume.release(); evol.release(); sub_fcell.release();
        sub_of_elem.clear(); sub_first_elem.clear(); fcell_sub_h.clear();
        num_nbr.release(); nbr.release(); phi.release(); radius.release(); start.release();
        nodes.release(); dphi.release(); volume.release(); inc_cell.release(); node_start.release();
        node_inc.release(); entry_pos.release(); g.release();
    }
};

// Element-tile plan of the fused step (tile.hip), internal numbering.  The plan
// renumbers nodes: interior nodes of each tile are contiguous, shared nodes follow.
struct TilePlan {
    bool ready = false;
    int te = 0;                       // elements per tile (= threads)
    int threads = 0;                  // workgroup size
    int64_t ntiles = 0;
    int maxtn = 0;                    // most nodes in one tile
    int max_nint = 0;                 // most interior nodes in one tile
    int f_rounds = 1;                 // LDS rounds of the element forces (tile kernel)
    int64_t nb = 0;                   // shared nodes (finished by the second launch)
    int64_t n_interior = 0;           // first shared node id
    int64_t total_tile_nodes = 0;
    int64_t nslot = 0;                // shared-node slots (tile sums parked for the second launch)
    size_t lds_bytes = 0;
    int stream_grid[4] = {0, 0, 0, 0};  // persistent Tet4 step grid per material kind (0 = not sized yet)
    bool use_stream = false;           // persistent pipelined Tet4 step (APHIP_FLAG_PIPELINED)
    DevBuf<int32_t> tn_start;         // [ntiles + 1] offsets into the tile-node arrays
    DevBuf<int32_t> tile_e0;          // [ntiles + 1] first element of each tile
    DevBuf<int32_t> tn_ninterior;     // [ntiles] interior nodes come first in each tile
    DevBuf<int32_t> int_base;         // [ntiles] id of the tile's first interior node
    DevBuf<int32_t> tn_node;          // [total] global node id
    DevBuf<uint32_t> tn_csr;          // [total] (local CSR start << 16) | count
    DevBuf<int32_t> tn_bdst;          // [total] slot of a shared node's tile sum (-1 interior)
    DevBuf<uint16_t> lcsr;            // [ntiles * te * npe] local element * npe + slot
    DevBuf<uint16_t> lconn;           // [ne * npe] local node of each element slot
    DevBuf<double> bpart;             // [slots * 3] tile sums of shared nodes
    DevBuf<double> upred;             // [nb * 3] shared nodes' predicted displacement for the next step
    DevBuf<int32_t> bstart;           // [nb + 1] slot range of each shared node
    DevBuf<double> zero;              // [8] zeros: stands in for absent optional arrays
    DevBuf<int32_t> meta;             // [ntiles * 8] packed per-tile scalars of the pipelined step
    // the tile kernel's copies at a fixed stride per tile (rows past a tile's end repeat
    // its last row), so their addresses need no per-tile scalar
    int tpad = 0;                     // node rows per tile (= maxtn)
    DevBuf<int32_t> tnp_node;         // [ntiles * tpad]
    DevBuf<uint32_t> tnp_csr;         // [ntiles * tpad]
    DevBuf<int32_t> tnp_bdst;         // [ntiles * tpad]
    DevBuf<uint16_t> lconn_p;         // [ntiles * te * npe]
    void reset() {
        ready = false;
        te = threads = maxtn = max_nint = 0;
        f_rounds = 1;
        ntiles = nb = n_interior = total_tile_nodes = nslot = 0;
        lds_bytes = 0;
        for (int &g : stream_grid) g = 0;
        tn_start.release(); tile_e0.release(); tn_ninterior.release(); int_base.release(); tn_node.release(); tn_csr.release();
        tn_bdst.release(); lcsr.release(); lconn.release(); bpart.release(); bstart.release(); upred.release();
        zero.release(); meta.release();
        tpad = 0;
        tnp_node.release(); tnp_csr.release(); tnp_bdst.release(); lconn_p.release();
    }
};

struct Profile {
    bool enabled = false;
    std::vector<hipEvent_t> ev;   // pairs (start, stop) tagged by which
    std::vector<int> tag;
    int64_t launches[3] = {0, 0, 0};
    double ms[3] = {0, 0, 0};
};

// Specimens collected between aphip_batch_begin / aphip_batch_end (caller numbering).
struct BatchBuild {
    bool open = false;
    int etype = -1;
    int kind = -1;
    std::vector<double> X;
    std::vector<int32_t> conn;
    std::vector<int64_t> node_seg{0}, elem_seg{0};
    std::vector<aphip_material_params> params;
    std::vector<double> dt;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;

    // mesh
    int etype = -1;
    int npe = 0;
    int64_t nn = 0, ne = 0;
    int64_t n_dof = 0;                // nodes with degrees of freedom: internal ids 0 .. n_dof - 1
    std::vector<int32_t> node_perm;   // caller -> internal
    std::vector<int32_t> node_iperm;  // internal -> caller
    std::vector<int32_t> elem_perm;   // caller -> internal
    std::vector<int32_t> elem_iperm;  // internal -> caller
    std::vector<int32_t> conn_h;      // internal connectivity (host copy, for tile plans)
    std::vector<int64_t> elem_seg{0}; // element blocks (specimens); tiles never cross them
    int flags = 0;
    BatchBuild batch;
    int64_t n_specimens = 0;
    std::vector<int64_t> spec_node_seg, spec_elem_seg;  // caller-numbering ranges per specimen

    DevBuf<int32_t> d_node_iperm, d_elem_iperm;  // device copies of node_iperm / elem_iperm
    DevBuf<int32_t> d_count;             // one counter (validation of device-side field input)

    DevBuf<double> X, u, v, a, f, mass;  // node vectors AoS n x 3 (mass n)
    DevBuf<uint8_t> fixed;               // bit d set = DOF d held at zero
    DevBuf<int32_t> conn;                // AoS ne x npe, internal node ids
    DevBuf<double> H, P;                 // SoA 9 x ne
    DevBuf<double> H1, P1;               // rate-form materials: the second of the N / NP1 pairs
    DevBuf<double> state[2];             // material state (N / NP1)
    int state_np1 = 1;                   // index of the NP1 state buffer
    double rate_dt = 0.0;                // increment for force-only evaluations of rate-form materials
    int formulation = APHIP_TOTAL_LAGRANGIAN;
    DevBuf<double> xc[2];                // updated Lagrangian: current coordinates (N / NP1 with the state);
                                         // semi-Lagrangian: xc[0] is the reference configuration
    DevBuf<double> dinc;                 // the step's displacement increments (semi: accumulated)
    DevBuf<double> Href;                 // semi-Lagrangian: reference displacement gradient (SoA 9 x ne)
    int64_t semi_interval = 1, steps_done = 0;
    DevBuf<double> &xcn() { return formulation == APHIP_SEMI_LAGRANGIAN ? xc[0] : xc[1 - state_np1]; }
    DevBuf<double> &xcnp1() { return xc[state_np1]; }
    // H of the previous step is read by the rate-form materials and the updated formulation
    bool h_pair() const { return mat.rate_form() || formulation == APHIP_UPDATED_LAGRANGIAN; }
    // H and P of the current evaluation (NP1) and of the previous step (N); the pairs
    // swap with the state buffers
    DevBuf<double> &Hnp1() { return h_pair() && state_np1 == 0 ? H1 : H; }
    DevBuf<double> &Pnp1() { return mat.rate_form() && state_np1 == 0 ? P1 : P; }
    DevBuf<double> &Hn() { return state_np1 == 0 ? H : H1; }
    DevBuf<double> &Pn() { return state_np1 == 0 ? P : P1; }
    DevBuf<double> volume;               // ne
    DevBuf<int32_t> node_start;          // nn + 1
    DevBuf<int32_t> node_inc;            // ne * npe: element * npe + slot, ascending element
    DevBuf<double> partial;              // ne * npe * 3 element force partials

    MaterialConst mat;
    bool has_fields = false;             // per-element lambda / mu
    DevBuf<double> elem_lambda, elem_mu;
    DevBuf<double> elem_rho;             // per-element density (ensembles), else mat.density
    std::vector<double> elem_E_h, elem_nu_h;  // per-element E / nu (internal order) when fields are set
    DevBuf<double> node_dt;              // per-node time increment (ensembles), else the step argument

    // captured step sequences (aphip_set_graph)
    bool graph_enabled = false;
    hipGraphExec_t graph_exec = nullptr;
    double graph_dt = 0.0;
    int graph_parity = -1;
    int graph_steps = 0;
    void graph_reset() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
        graph_steps = 0;
    }
    bool has_mass = false;
    bool has_fixed = false;
    bool all_fixed = false;              // every DOF of every node is fixed

    std::unique_ptr<exo::Writer> results;  // open Exodus II results file, if any

    // power method: the eigenvector of the last call, the next call's start vector; the
    // per-node max edge length, computed on the first call after the mesh is set (it
    // depends on X and the connectivity only; a node renumbering permutes it); the work
    // buffers of the allocating entry point, kept between its calls
    DevBuf<double> pm_eig;               // 3 nn, internal numbering, allocated with the mesh
    DevBuf<double> pm_l;                 // nn, internal numbering, allocated with the mesh
    DevBuf<double> pm_work[8];
    bool pm_eig_valid = false;
    bool pm_l_valid = false;

    CellData cells;
    TilePlan tiles;
    AdjointData adj;
    Profile prof;

    int64_t state_count() const {
        return mat.state_width() * ne * ((etype == APHIP_HEX8 && mat.state_per_qp) ? 8 : 1);
    }
};

void set_last_error(const std::string &msg);

// ---- kernels (fem.hip) ----
void fem_element_force(Ctx &c);
void fem_predict(Ctx &c, double dt);
// semi-Lagrangian reference update: x_ref += increment, increment = 0, H_ref = H
void fem_semi_update(Ctx &c);
// mode 0: f only; 1: f and a = f/m (initial); 2: f, a, BC, v += dt/2 a (step corrector)
void fem_node_update(Ctx &c, double dt, int mode);
void fem_volumes(Ctx &c);
void fem_lumped_mass(Ctx &c, double rho);

// ---- fused tile step (tile.hip) ----
void tile_build(Ctx &c, int te, int threads);
// Reorders the internal nodes (old_of_new[i] = current id of new node i) and every
// node-indexed array of the context (capi.hip).
void renumber_nodes(Ctx &c, const std::vector<int32_t> &old_of_new);
// One tiled pass (MODE 0 force only, 1 force and a = f/m, 2 the full step).  In MODE 0 the
// displacement read and the force written may be other arrays than the context's u and f
// (the two-pass RKPM force: physical displacement in, local force out).
// pred_in (mode 2): shared nodes take the displacement predicted by the previous step's
// shared-node pass (it must have run with pred_out and the same dt in this same call);
// pred_out (mode 2): the shared-node pass stores that prediction for the next step.
void tile_run(Ctx &c, double dt, int mode, const double *u_in = nullptr, double *f_out = nullptr,
              bool pred_in = false, bool pred_out = false);

// ---- adjoint (adjoint.hip) ----
void adjoint_forward(Ctx &c, double dt, int n_steps, int interval);
void adjoint_backward(Ctx &c, const double *ub_T, const double *vb_T, double *E_bar, double *nu_bar,
                      double *v0_bar);
void force_vjp(Ctx &c, const double *f_bar, double *u_bar, double *E_bar, double *nu_bar);
void step_once_tiles(Ctx &c, double dt);

// ---- RK (rk.hip) ----
// nodal_subdomains >= 0: nodal smoothing cells of a thex mesh (0: every hex a subcell,
// 1: every cell one subcell), with F-bar over the subcells when fbar
void rk_setup(Ctx &c, double scale, bool two_pass, bool fe_cells = false, int nodal_subdomains = -1,
              bool fbar = false);
void rk_cell_force(Ctx &c);
void rk_node_update(Ctx &c, double dt, int mode);
void rk_mass(Ctx &c, double *total);

// ---- profiling helpers (capi.hip) ----
void prof_begin(Ctx &c, int which);
void prof_end(Ctx &c, int which);

// ---- power-method stable increment (power.hip) ----
void pm_max_edge(Ctx &c, double *l);
void pm_start(Ctx &c, const int32_t *caller_id, uint64_t seed, double eps, const double *l, double *v);
void pm_perturb(Ctx &c, const double *u_in, const double *v, const double *l, double eps);
// doubles of the partial-sum buffer pm_next needs (one per block of nodes, plus the total)
int64_t pm_partials(const Ctx &c);
// v = (f - f0) / (l eps m), returns |v|_2 (partial: pm_partials(c) doubles)
double pm_next(Ctx &c, const double *f0, const double *l, double eps, double *v, double *partial);
void pm_scale(Ctx &c, double norm, double *v);  // v /= norm

// ---- field transfer with caller device arrays (field.hip) ----
// dst[i, :] = src[iperm[i], :] / dst[iperm[i], :] = src[i, :], w doubles per entity
void field_gather_rows(Ctx &c, int64_t n, int w, const int32_t *iperm, const double *src, double *dst);
void field_scatter_rows(Ctx &c, int64_t n, int w, const int32_t *iperm, const double *src, double *dst);
// SoA 9 x ne tensor planes -> caller row-major n_elem x 9
void field_scatter_tensor(Ctx &c, const double *planes, double *dst);
// caller-numbered nodal mass into c.mass; returns how many values are not positive
int32_t field_gather_mass(Ctx &c, const double *src);

// ---- copy yardstick (probe.hip): device milliseconds per copy of nbytes ----
double bandwidth_probe(int device, int64_t nbytes, int reps);

}  // namespace aphip
