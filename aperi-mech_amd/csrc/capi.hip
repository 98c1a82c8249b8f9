// C ABI of the aperi-hip library: context lifetime, mesh set-up (Morton
// renumbering, node-to-element CSR), material resolution, field transfer and
// the step driver.  Host C++ around the kernels in fem.hip and rk.hip.
#include <algorithm>
#include <functional>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <utility>

#include "internal.hpp"

namespace aphip {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

// ---------------- profiling ----------------

struct ProfRecord {
    hipEvent_t start, stop;
    int which;
};
static thread_local std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    APHIP_HIP(hipEventCreate(&e));
    return e;
}

void prof_begin(Ctx &c, int which) {
    if (!c.prof.enabled) return;
    hipEvent_t e = take_event();
    APHIP_HIP(hipEventRecord(e, c.stream));
    c.prof.ev.push_back(e);
    c.prof.tag.push_back(which);
}

void prof_end(Ctx &c, int which) {
    if (!c.prof.enabled) return;
    hipEvent_t e = take_event();
    APHIP_HIP(hipEventRecord(e, c.stream));
    c.prof.ev.push_back(e);
    c.prof.tag.push_back(-1 - which);
}

static void prof_collect(Ctx &c) {
    if (c.prof.ev.empty()) return;
    APHIP_HIP(hipStreamSynchronize(c.stream));
    // match each begin with the next end of the same class
    std::vector<int> open(3, -1);
    for (size_t k = 0; k < c.prof.ev.size(); ++k) {
        const int t = c.prof.tag[k];
        if (t >= 0) {
            open[t] = (int)k;
        } else {
            const int w = -1 - t;
            if (open[w] >= 0) {
                float ms = 0.0f;
                APHIP_HIP(hipEventElapsedTime(&ms, c.prof.ev[open[w]], c.prof.ev[k]));
                c.prof.ms[w] += ms;
                c.prof.launches[w] += 1;
                open[w] = -1;
            }
        }
    }
    for (hipEvent_t e : c.prof.ev) g_event_pool.push_back(e);
    c.prof.ev.clear();
    c.prof.tag.clear();
}

// ---------------- mesh set-up helpers ----------------

static inline uint64_t spread3(uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// 63-bit Morton code of each point over the points' bounding box.
static std::vector<uint64_t> morton_codes(const std::vector<double> &pts, int64_t n) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], pts[3 * i + d]);
            hi[d] = std::max(hi[d], pts[3 * i + d]);
        }
    double ext = 0.0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d]);
    const double scale = ext > 0.0 ? (double)((1u << 21) - 1) / ext : 0.0;
    std::vector<uint64_t> code(n);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t q[3];
        for (int d = 0; d < 3; ++d) {
            double t = (pts[3 * i + d] - lo[d]) * scale;
            t = std::min(std::max(t, 0.0), (double)((1u << 21) - 1));
            q[d] = (uint64_t)t;
        }
        code[i] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    }
    return code;
}

// Stable ordering of points by Morton code.
static std::vector<int32_t> morton_order(const std::vector<double> &pts, int64_t n) {
    const std::vector<uint64_t> code = morton_codes(pts, n);
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
    return order;
}

// Recursive coordinate bisection: a set of m points is split across its longest extent
// at a multiple of a, the largest power of two below m and at most 1024, nearest m / 2;
// sets of at most 64 points are kept in Morton order.  Every aligned run of 64..1024
// consecutive points of the result is therefore one subtree, a compact box, where a run
// of the Morton order may straddle octree cells.  Ties are broken by index, so the
// order is deterministic.
static std::vector<int32_t> rcb_order(const std::vector<double> &pts, int64_t n) {
    const std::vector<uint64_t> code = morton_codes(pts, n);
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    // the halves of the top splits run on threads of their own (disjoint ranges of idx)
    const int par_depth = std::thread::hardware_concurrency() >= 16 ? 4 : (std::thread::hardware_concurrency() >= 4 ? 2 : 0);
    std::function<void(int64_t, int64_t, int)> split = [&](int64_t b, int64_t e, int depth) {
        const int64_t m = e - b;
        if (m <= 64) {
            std::sort(idx.begin() + b, idx.begin() + e, [&](int32_t x, int32_t y) {
                return code[x] != code[y] ? code[x] < code[y] : x < y;
            });
            return;
        }
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = b; i < e; ++i)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], pts[3 * (int64_t)idx[i] + d]);
                hi[d] = std::max(hi[d], pts[3 * (int64_t)idx[i] + d]);
            }
        int ax = 0;
        for (int d = 1; d < 3; ++d)
            if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
        int64_t a = 1024;
        while (a >= m) a /= 2;
        int64_t k = ((m / 2 + a / 2) / a) * a;
        k = std::min(std::max(k, a), m - 1);
        std::nth_element(idx.begin() + b, idx.begin() + b + k, idx.begin() + e, [&](int32_t x, int32_t y) {
            const double px = pts[3 * (int64_t)x + ax], py = pts[3 * (int64_t)y + ax];
            return px != py ? px < py : x < y;
        });
        if (depth < par_depth && m > 65536) {
            std::thread left([&, b, k, depth] { split(b, b + k, depth + 1); });
            split(b + k, e, depth + 1);
            left.join();
        } else {
            split(b, b + k, depth + 1);
            split(b + k, e, depth + 1);
        }
    };
    split(0, n, 0);
    return idx;
}

static void resolve_material(int kind, const aphip_material_params *p, MaterialConst &m) {
    APHIP_REQUIRE(kind >= APHIP_LINEAR_ELASTIC && kind <= APHIP_DRUCKER_PRAGER, "unknown material kind");
    APHIP_REQUIRE(p != nullptr, "material params are NULL");
    const double E = p->youngs_modulus, nu = p->poissons_ratio;
    // LinearElasticProperties::CompleteProperties, E/nu branch and bound checks
    // (include/Materials/Base.h:102-107, :130-158)
    APHIP_REQUIRE(nu >= -1.0 && nu < 0.5, "Poisson's ratio must be in [-1, 0.5)");
    APHIP_REQUIRE(E > 0.0, "Young's modulus must be positive");
    m.kind = kind;
    m.density = p->density;
    m.mu = E / (2 * (1 + nu));
    m.lambda = (E * nu) / ((1 + nu) * (1 - 2 * nu));
    APHIP_REQUIRE(m.lambda > -(2.0 / 3.0) * m.mu, "lambda must exceed -2G/3");
    m.yield_stress = p->yield_stress;
    m.hardening = p->hardening_modulus;
    m.E = E;
    m.nu = nu;
    m.state_per_qp = p->state_per_qp ? 1 : 0;
    if (kind == APHIP_NEO_HOOKEAN_DAMAGE) {
        // NeoHookeanWithDamageGetStressFunctor's checks and clamps (NeoHookeWithDamage.h:126-168)
        APHIP_REQUIRE(m.lambda >= 0.0, "NeoHookeanWithDamage: lambda must be >= 0");
        APHIP_REQUIRE(m.mu >= 0.0, "NeoHookeanWithDamage: mu must be >= 0");
        APHIP_REQUIRE(p->I1_critical >= 0.0, "NeoHookeanWithDamage: I1_critical must be >= 0");
        APHIP_REQUIRE(p->I1_failure >= 0.0, "NeoHookeanWithDamage: I1_failure must be >= 0");
        APHIP_REQUIRE(p->I1_critical < p->I1_failure, "NeoHookeanWithDamage: I1_critical must be < I1_failure");
        APHIP_REQUIRE(p->alpha >= 0.0, "NeoHookeanWithDamage: alpha must be >= 0");
        APHIP_REQUIRE(p->maximum_damage >= 0.0 && p->maximum_damage <= 1.0,
                      "NeoHookeanWithDamage: max_effective_damage must be in [0, 1]");
        m.I1_critical = p->I1_critical;
        m.I1_failure = p->I1_failure;
        m.alpha = p->alpha;
        m.max_damage = p->maximum_damage;
        m.max_damage_increment = std::min(std::max(p->maximum_damage_increment, 0.0), 1.0);
    }
    if (kind == APHIP_POWER_LAW_CREEP) {
        // PowerLawCreepGetStressFunctor's checks (PowerLawCreep.h:80-90); bulk modulus of the
        // E / nu pair (LinearElasticProperties::CompleteProperties)
        APHIP_REQUIRE(p->creep_A >= 0.0, "PowerLawCreep: A must be > 0");
        APHIP_REQUIRE(p->creep_n >= 0.0, "PowerLawCreep: n must be > 0");
        APHIP_REQUIRE(p->creep_m >= 0.0, "PowerLawCreep: m must be > 0");
        APHIP_REQUIRE(p->use_constant_temperature != 0,
                      "PowerLawCreep: currently have to use constant temperature, but use_constant_temperature is "
                      "false");
        const double T = p->constant_temperature;
        m.bulk = E / (3 * (1 - 2 * nu));
        m.creep_rate = T == 0.0 ? p->creep_A * std::exp(-p->creep_m) : p->creep_A * std::exp(-p->creep_m / T);
        m.creep_n = p->creep_n;
    }
    if (kind == APHIP_DRUCKER_PRAGER) {  // DruckerPragerMaterial::CreateStressFunctor (DruckerPrager.h:52-68)
        m.bulk = E / (3 * (1 - 2 * nu));
        m.dp_A1 = p->dp_A1;
        m.dp_A2 = p->dp_A2;
        m.dp_A2G = p->dp_A2G;
    }
}

static void alloc_state(Ctx &c) {
    const int64_t ns = c.state_count();
    for (int s = 0; s < 2; ++s) {
        if (ns) {
            c.state[s].alloc(ns);
            c.state[s].zero(c.stream);
        } else {
            c.state[s].release();
        }
    }
    if (c.h_pair()) {  // the N halves of the H / P pairs start at zero
        c.H1.alloc(9 * std::max<int64_t>(c.ne, 1));
        c.H1.zero(c.stream);
    } else {
        c.H1.release();
    }
    if (c.mat.rate_form()) {
        c.P1.alloc(9 * std::max<int64_t>(c.ne, 1));
        c.P1.zero(c.stream);
    } else {
        c.P1.release();
    }
    c.state_np1 = 1;
}

static void require_ready(Ctx &c, bool need_mass) {
    APHIP_REQUIRE(c.etype >= 0, "mesh not set");
    APHIP_REQUIRE(c.mat.kind >= 0, "material not set");
    if (need_mass) APHIP_REQUIRE(c.has_mass, "mass not computed or set");
}

// node -> (element, slot) incidences in ascending internal element order (two-pass path, mass)
static void upload_node_csr(Ctx &c, const std::vector<int32_t> &ci) {
    const int64_t n_nodes = c.nn, n_inc = c.ne * c.npe;
    std::vector<int32_t> start(n_nodes + 1, 0), inc((size_t)std::max<int64_t>(n_inc, 1));
    for (int64_t t = 0; t < n_inc; ++t) start[ci[t] + 1]++;
    for (int64_t i = 0; i < n_nodes; ++i) start[i + 1] += start[i];
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int64_t t = 0; t < n_inc; ++t) inc[fill[ci[t]]++] = (int32_t)t;
    c.node_start.upload(start.data(), start.size(), c.stream);
    c.node_inc.upload(inc.data(), inc.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

template <typename T>
static void permute_device(Ctx &c, DevBuf<T> &b, int comps, const std::vector<int32_t> &old_of_new) {
    if (!b.p) return;
    const int64_t n = c.nn;
    std::vector<T> h((size_t)n * comps), o((size_t)n * comps);
    b.download(h.data(), h.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < comps; ++d) o[i * comps + d] = h[(int64_t)old_of_new[i] * comps + d];
    b.upload(o.data(), o.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

// device copies of the internal -> caller permutations (device-side field transfer)
static void upload_perms(Ctx &c) {
    c.d_node_iperm.upload(c.node_iperm.data(), c.node_iperm.size(), c.stream);
    c.d_elem_iperm.upload(c.elem_iperm.data(), c.elem_iperm.size(), c.stream);
    c.d_count.alloc(1);
}

void renumber_nodes(Ctx &c, const std::vector<int32_t> &old_of_new) {
    const int64_t n = c.nn;
    std::vector<int32_t> new_of_old(n);
    for (int64_t i = 0; i < n; ++i) new_of_old[old_of_new[i]] = (int32_t)i;
    std::vector<int32_t> iperm(n);
    for (int64_t i = 0; i < n; ++i) iperm[i] = c.node_iperm[old_of_new[i]];
    c.node_iperm.swap(iperm);
    for (int64_t i = 0; i < n; ++i) c.node_perm[c.node_iperm[i]] = (int32_t)i;
    upload_perms(c);
    for (auto &x : c.conn_h) x = new_of_old[x];
    c.conn.upload(c.conn_h.data(), c.conn_h.size(), c.stream);
    upload_node_csr(c, c.conn_h);
    permute_device(c, c.X, 3, old_of_new);
    permute_device(c, c.u, 3, old_of_new);
    permute_device(c, c.v, 3, old_of_new);
    permute_device(c, c.a, 3, old_of_new);
    permute_device(c, c.f, 3, old_of_new);
    if (c.has_mass) permute_device(c, c.mass, 1, old_of_new);
    if (c.has_fixed) permute_device(c, c.fixed, 1, old_of_new);
    if (c.pm_eig_valid) permute_device(c, c.pm_eig, 3, old_of_new);
    if (c.pm_l_valid) permute_device(c, c.pm_l, 1, old_of_new);
    permute_device(c, c.node_dt, 1, old_of_new);  // ensembles: each specimen's increment
    // recorded adjoint checkpoints hold node state in the old numbering
    if (c.adj.ready) c.adj.reset();
    if (c.formulation != APHIP_TOTAL_LAGRANGIAN) {
        permute_device(c, c.xc[0], 3, old_of_new);
        permute_device(c, c.xc[1], 3, old_of_new);
        permute_device(c, c.dinc, 3, old_of_new);
    }
}

static bool use_tiles(const Ctx &c) {
    return !c.cells.ready && c.tiles.ready && !(c.flags & APHIP_FLAG_TWO_PASS) &&
           c.formulation == APHIP_TOTAL_LAGRANGIAN;
}

static void default_tiling(Ctx &c) {
    if (c.ne == 0 || (c.flags & APHIP_FLAG_TWO_PASS)) return;
    if (c.etype == APHIP_TET4) {
        // 1024-element tiles (fewer shared nodes) when the mesh's tiles fit a CU's LDS
        try {
            tile_build(c, 1024, 1024);
        } catch (const Error &) {
            tile_build(c, 512, 512);
        }
    } else {
        tile_build(c, 128, 128);
    }
}


// Mesh set-up shared by aphip_set_mesh and the ensemble builder.  node_seg / elem_seg
// delimit blocks of caller nodes / elements (one block per specimen) that are
// ordered independently and never share a tile.
static void setup_mesh(Ctx &c, int32_t elem_type, int64_t n_nodes, const double *X, int64_t n_elem,
                       const int32_t *conn, const std::vector<int64_t> &node_seg, const std::vector<int64_t> &elem_seg) {
    APHIP_REQUIRE(elem_type == APHIP_TET4 || elem_type == APHIP_HEX8, "element type must be TET4 or HEX8");
    APHIP_REQUIRE(n_nodes > 0 && n_nodes < (int64_t)INT32_MAX, "n_nodes out of range");
    APHIP_REQUIRE(n_elem >= 0, "n_elem out of range");
    APHIP_REQUIRE(X != nullptr && (n_elem == 0 || conn != nullptr), "mesh arrays are NULL");
    const int npe = elem_type == APHIP_TET4 ? 4 : 8;
    APHIP_REQUIRE(n_elem * npe < (int64_t)INT32_MAX, "n_elem too large for int32 incidences");
    for (int64_t i = 0; i < n_elem * npe; ++i)
        APHIP_REQUIRE(conn[i] >= 0 && conn[i] < n_nodes, "connectivity references a node out of range");
    c.cells.reset();
    c.formulation = APHIP_TOTAL_LAGRANGIAN;
    c.xc[0].release();
    c.xc[1].release();
    c.dinc.release();
    c.Href.release();
    c.steps_done = 0;
    c.pm_eig_valid = false;
    c.pm_l_valid = false;
    for (DevBuf<double> &b : c.pm_work) b.release();
    c.tiles.reset();
    c.graph_reset();
    c.etype = elem_type;
    c.npe = npe;
    c.nn = n_nodes;
    c.n_dof = n_nodes;
    c.ne = n_elem;
    c.elem_seg = elem_seg;

    // Renumbering per block: nodes in Morton order of position; elements by centroid, in
    // Morton order (Hex8: on lattices its tile-sized runs are octree blocks already) or,
    // for Tet4, by recursive bisection (runs of 1024 tets touch 11 % fewer tile nodes and
    // leave 15 % fewer of them shared than Morton runs on a jittered lattice)
    c.node_iperm.assign(n_nodes, 0);
    c.elem_iperm.assign(n_elem, 0);
    for (size_t sgm = 0; sgm + 1 < node_seg.size(); ++sgm) {
        const int64_t n0 = node_seg[sgm], n1 = node_seg[sgm + 1];
        std::vector<double> xs(X + 3 * n0, X + 3 * n1);
        const std::vector<int32_t> ord = morton_order(xs, n1 - n0);
        for (int64_t i = 0; i < n1 - n0; ++i) c.node_iperm[n0 + i] = (int32_t)(n0 + ord[i]);
        const int64_t e0 = elem_seg[sgm], e1 = elem_seg[sgm + 1];
        std::vector<double> cent(3 * (e1 - e0), 0.0);
        for (int64_t e = e0; e < e1; ++e) {
            for (int k = 0; k < npe; ++k) {
                APHIP_REQUIRE(conn[e * npe + k] >= n0 && conn[e * npe + k] < n1, "an element references another specimen's node");
                for (int d = 0; d < 3; ++d) cent[3 * (e - e0) + d] += X[3 * (int64_t)conn[e * npe + k] + d] / npe;
            }
        }
        const std::vector<int32_t> eord = npe == 4 ? rcb_order(cent, e1 - e0) : morton_order(cent, e1 - e0);
        for (int64_t e = 0; e < e1 - e0; ++e) c.elem_iperm[e0 + e] = (int32_t)(e0 + eord[e]);
    }
    c.node_perm.assign(n_nodes, 0);
    for (int64_t i = 0; i < n_nodes; ++i) c.node_perm[c.node_iperm[i]] = (int32_t)i;
    c.elem_perm.assign(n_elem, 0);
    for (int64_t e = 0; e < n_elem; ++e) c.elem_perm[c.elem_iperm[e]] = (int32_t)e;

    std::vector<double> Xi(3 * n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i)
        for (int d = 0; d < 3; ++d) Xi[3 * i + d] = X[3 * (int64_t)c.node_iperm[i] + d];
    std::vector<int32_t> ci((size_t)n_elem * npe);
    for (int64_t e = 0; e < n_elem; ++e)
        for (int k = 0; k < npe; ++k) ci[e * npe + k] = c.node_perm[conn[(int64_t)c.elem_iperm[e] * npe + k]];

    c.X.upload(Xi.data(), Xi.size(), c.stream);
    c.conn.upload(ci.data(), ci.size(), c.stream);
    upload_perms(c);
    upload_node_csr(c, ci);
    for (DevBuf<double> *b : {&c.u, &c.v, &c.a, &c.f}) {
        b->alloc(3 * n_nodes);
        b->zero(c.stream);
    }
    c.pm_eig.alloc(3 * n_nodes);  // the power method's eigenvector, kept between its calls
    c.pm_l.alloc(n_nodes);        // and its max edge lengths
    c.H.alloc(9 * std::max<int64_t>(n_elem, 1));
    c.P.alloc(9 * std::max<int64_t>(n_elem, 1));
    c.H.zero(c.stream);
    c.P.zero(c.stream);
    c.partial.alloc((size_t)std::max<int64_t>(n_elem, 1) * npe * 3);
    c.mass.release();
    c.has_mass = false;
    c.fixed.release();
    c.has_fixed = false;
    c.all_fixed = false;
    c.has_fields = false;
    c.elem_lambda.release();
    c.elem_mu.release();
    c.elem_rho.release();
    c.node_dt.release();
    fem_volumes(c);
    if (c.mat.kind >= 0) alloc_state(c);
    c.conn_h.swap(ci);
    default_tiling(c);
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

static void force_pass(Ctx &c) {
    prof_begin(c, 1);
    if (c.cells.ready) rk_cell_force(c);
    else fem_element_force(c);
    prof_end(c, 1);
}

static void step_once(Ctx &c, double dt, bool pred_in = false, bool pred_out = false);

static void node_pass(Ctx &c, double dt, int mode) {
    prof_begin(c, 2);
    if (c.cells.ready) rk_node_update(c, dt, mode);
    else fem_node_update(c, dt, mode);
    prof_end(c, 2);
}

void step_once_tiles(Ctx &c, double dt);

// one step of the explicit loop body (src/Solvers/Explicit.cpp:391-452).  On the tile
// path, pred_out has the shared-node pass also store the shared nodes' predicted
// displacement for the next step, and pred_in has the tile kernel take it instead of
// predicting from u, v and a (the same arithmetic on the same values: bitwise the same
// step, 57 B less per shared tile-node row).  Only steps of one aphip_step call pass a
// prediction on, so nothing between calls can leave it stale.
static void step_once(Ctx &c, double dt, bool pred_in, bool pred_out) {
    c.state_np1 = 1 - c.state_np1;  // UpdateFieldStates (Explicit.cpp:393)
    c.rate_dt = dt;                 // SetTimeIncrement (ElementTetrahedron4.h:187)
    if (use_tiles(c)) {
        tile_run(c, dt, 2, nullptr, nullptr, pred_in, pred_out);
    } else {
        fem_predict(c, dt);
        force_pass(c);
        node_pass(c, dt, 2);
    }
}

void step_once_tiles(Ctx &c, double dt) { step_once(c, dt); }

// f, H, P and the NP1 state from the current u (ComputeInternalForceAllElements).
static void internal_force_eval(Ctx &c) {
    if (use_tiles(c)) {
        tile_run(c, 0.0, 0);
    } else {
        force_pass(c);
        node_pass(c, 0.0, 0);
    }
}

struct PowerMethodResult {
    double dt;
    int iterations;
    bool converged;
};

// Stable time increment 2 / sqrt(lambda_max(M^-1 K)) by the reference's power
// method (include/PowerMethodProcessor.h:331-470, ComputeStableTimeIncrement): the
// force is linearised about the current displacement by a perturbation of eps
// times the node's max edge length; convergence when successive increments agree
// to `tolerance`, or when the variance of the last five is below
// `variation_tolerance` x their mean; otherwise the smallest of the last five.
// The reference's oscillation test (:424-437) takes its else branch at i = 0 and so
// never reports oscillation; that outcome is kept.  The first comparison is against
// prev_dt (1e-16 when 0, :363-364).  Like the reference's processor, the context keeps
// the final eigenvector and starts the next call from it (:350-361, :468-469); only the
// first call after the mesh is set draws the random start.  The force evaluations write H, P
// and the NP1 state into the call's work buffers (the context's pointers are exchanged for
// the call and exchanged back on every exit), and u and f are restored afterwards, so the
// call leaves the simulation state untouched.  The max edge lengths are the context's.
// Device buffers of one power-method call (internal numbering): the displacement and
// force on entry, F(u), the iterate, the targets of H / P / state and the norm's partial
// sums.
struct PowerMethodWork {
    double *u_in, *f_in, *f0, *v, *H, *P, *st, *partial;
};

// doubles each buffer of PowerMethodWork needs for the context's mesh and material
static std::array<int64_t, 8> power_method_sizes(const Ctx &c) {
    const int64_t n3 = 3 * c.nn, nt = c.ne ? 9 * c.ne : 0;
    return {n3, n3, n3, n3, nt, nt, c.state_count(), pm_partials(c)};
}

// points context buffers at other memory until it goes out of scope
class Redirect {
  public:
    Redirect() = default;
    Redirect(const Redirect &) = delete;
    Redirect &operator=(const Redirect &) = delete;
    void to(double *&ptr, double *target) {
        saved_[n_++] = {&ptr, ptr};
        ptr = target;
    }
    ~Redirect() {
        while (n_ > 0) {
            --n_;
            *saved_[n_].first = saved_[n_].second;
        }
    }

  private:
    std::array<std::pair<double **, double *>, 3> saved_{};
    int n_ = 0;
};

static PowerMethodResult power_method(Ctx &c, const PowerMethodWork &w, double prev_dt, int n_iter, double eps,
                                      double tol, double vtol, uint64_t seed) {
    const int64_t n3 = 3 * c.nn;
    APHIP_REQUIRE(c.formulation == APHIP_TOTAL_LAGRANGIAN, "the power method runs with the total formulation");
    APHIP_REQUIRE(!(c.cells.ready && c.cells.nodal), "the power method does not run on nodal cells");
    // PowerMethodProcessor::CheckEssentialBoundaries (:159-173)
    APHIP_REQUIRE(!(c.has_fixed && c.all_fixed),
                  "All nodes are in the essential boundary set. Cannot run the power method. Use the "
                  "direct_time_stepper instead.");
    const int64_t ns = c.state_count();
    APHIP_HIP(hipMemcpyAsync(w.u_in, c.u.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    APHIP_HIP(hipMemcpyAsync(w.f_in, c.f.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    Redirect redirect;
    if (c.ne) {
        redirect.to(c.Hnp1().p, w.H);
        redirect.to(c.Pnp1().p, w.P);
    }
    if (ns) redirect.to(c.state[c.state_np1].p, w.st);
    c.graph_reset();
    if (!c.pm_l_valid) {
        pm_max_edge(c, c.pm_l.p);
        c.pm_l_valid = true;
    }
    const double *l = c.pm_l.p;
    if (c.pm_eig_valid)
        APHIP_HIP(hipMemcpyAsync(w.v, c.pm_eig.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    else
        pm_start(c, c.d_node_iperm.p, seed, eps, l, w.v);
    internal_force_eval(c);  // F(u)
    APHIP_HIP(hipMemcpyAsync(w.f0, c.f.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));

    constexpr int window = 5;
    std::array<double, window> recent{};
    double dt_np1 = prev_dt == 0.0 ? 1.0e-16 : prev_dt, dt_n = 0.0;
    bool converged = false, oscillating = false;
    int used = n_iter;
    for (int k = 0; k < n_iter; ++k) {
        dt_n = dt_np1;
        pm_perturb(c, w.u_in, w.v, l, eps);
        internal_force_eval(c);
        const double lambda = pm_next(c, w.f0, l, eps, w.v, w.partial);
        pm_scale(c, lambda, w.v);
        dt_np1 = 2.0 / std::sqrt(lambda);
        recent[k % window] = dt_np1;
        if (std::fabs(dt_np1 - dt_n) < tol * dt_np1) {
            converged = true;
            used = k + 1;
            break;
        }
        if (k >= window) {
            double mean = 0.0;
            for (double x : recent) mean += x;
            mean /= window;
            double var = 0.0;
            for (double x : recent) var += (x - mean) * (x - mean);
            var /= (window - 1);
            if (k > window + 1) {
                double alt = 0.0;
                for (int i = 0; i < window - 1; ++i) {
                    const double diff = recent[i + 1] - recent[i];
                    if (i > 0 && diff * alt < 0) {
                        oscillating = true;
                    } else {
                        oscillating = false;
                        break;
                    }
                    alt = diff;
                }
            }
            if (var < vtol * mean || (oscillating && var < vtol * mean * 2.0)) {
                converged = true;
                used = k + 1;
                break;
            }
        }
    }
    const double wmin = *std::min_element(recent.begin(), recent.end());
    const double dt = converged ? (oscillating ? wmin : dt_np1) : wmin;
    APHIP_HIP(hipMemcpyAsync(c.pm_eig.p, w.v, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    c.pm_eig_valid = true;

    APHIP_HIP(hipMemcpyAsync(c.u.p, w.u_in, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    APHIP_HIP(hipMemcpyAsync(c.f.p, w.f_in, sizeof(double) * n3, hipMemcpyDeviceToDevice, c.stream));
    APHIP_HIP(hipStreamSynchronize(c.stream));
    return PowerMethodResult{dt, used, converged};
}

static void get_node_field(Ctx &c, const DevBuf<double> &b, double *out) {
    std::vector<double> h(3 * c.nn);
    b.download(h.data(), h.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    for (int64_t i = 0; i < c.nn; ++i)
        for (int d = 0; d < 3; ++d) out[3 * (int64_t)c.node_iperm[i] + d] = h[3 * i + d];
}

static void set_node_field(Ctx &c, DevBuf<double> &b, const double *in) {
    std::vector<double> h(3 * c.nn);
    for (int64_t i = 0; i < c.nn; ++i)
        for (int d = 0; d < 3; ++d) h[3 * i + d] = in[3 * (int64_t)c.node_iperm[i] + d];
    b.upload(h.data(), h.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace aphip

using namespace aphip;

struct aphip_ctx : Ctx {};

#define APHIP_API_BEGIN try {
#define APHIP_API_END                                                   \
    return APHIP_OK;                                                    \
    }                                                                   \
    catch (const Error &e) {                                            \
        set_last_error(e.what());                                       \
        return e.code;                                                  \
    }                                                                   \
    catch (const std::bad_alloc &) {                                    \
        set_last_error("host allocation failed");                       \
        return APHIP_ERR_ALLOC;                                         \
    }                                                                   \
    catch (const std::exception &e) {                                   \
        set_last_error(e.what());                                       \
        return APHIP_ERR_INVALID;                                       \
    }

// Scratch memory of the allocation-free power method
// (aphip_power_method_time_increment_noalloc): a fixed array in the code object, one copy
// per GPU, present from module load to process end.  Host code never passes `scratch`
// itself to a kernel or copy (on the host the name is only a stand-in for the symbol):
// aphip_create looks up each device's copy once, with that device current, and every
// later use goes through the address kept in g_scratch_ptr.
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Looks up this device's copy of `scratch` (the device must be current).
static void scratch_resolve(int device) {
    APHIP_REQUIRE(device >= 0 && device < kMaxDevices, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        throw Error(APHIP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
}

static void *scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

extern "C" {

const char *aphip_last_error(void) { return g_last_error.c_str(); }

int aphip_device_count(int32_t *count) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(count != nullptr, "NULL argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    APHIP_API_END
}

int aphip_create(aphip_ctx **out, const aphip_config *cfg) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(out != nullptr, "ctx out-pointer is NULL");
    *out = nullptr;
    int ndev = 0;
    APHIP_HIP(hipGetDeviceCount(&ndev));
    const int dev = cfg ? cfg->device : 0;
    APHIP_REQUIRE(dev >= 0 && dev < ndev, "invalid HIP device ordinal");
    APHIP_HIP(hipSetDevice(dev));
    scratch_resolve(dev);
    aphip_ctx *c = new aphip_ctx();
    c->device = dev;
    c->flags = cfg ? cfg->flags : 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        throw Error(APHIP_ERR_HIP, "hipStreamCreate failed");
    }
    *out = c;
    APHIP_API_END
}

int aphip_destroy(aphip_ctx *c) {
    APHIP_API_BEGIN
    if (!c) return APHIP_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->graph_reset();
    for (hipEvent_t e : c->prof.ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    APHIP_API_END
}

int aphip_set_mesh(aphip_ctx *c, int32_t elem_type, int64_t n_nodes, const double *X, int64_t n_elem,
                   const int32_t *conn) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_HIP(hipSetDevice(c->device));
    c->batch = BatchBuild();
    setup_mesh(*c, elem_type, n_nodes, X, n_elem, conn, {0, n_nodes}, {0, n_elem});
    APHIP_API_END
}

int aphip_set_tiling(aphip_ctx *c, int32_t te, int32_t threads) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    tile_build(*c, te, threads);
    APHIP_API_END
}

int aphip_get_node_permutation(aphip_ctx *c, int32_t *perm) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c && perm, "NULL argument");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    std::memcpy(perm, c->node_perm.data(), sizeof(int32_t) * c->nn);
    APHIP_API_END
}

int aphip_set_material(aphip_ctx *c, int32_t kind, const aphip_material_params *p) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    MaterialConst m;
    resolve_material(kind, p, m);
    APHIP_REQUIRE(c->formulation != APHIP_SEMI_LAGRANGIAN || !m.rate_form(),
                  "the semi-Lagrangian formulation does not take the rate-form materials");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    c->mat = m;
    if (c->etype >= 0) alloc_state(*c);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

static void set_formulation(Ctx &c, int formulation, int64_t interval) {
    APHIP_REQUIRE(c.etype >= 0 && c.mat.kind >= 0, "mesh and material must be set");
    APHIP_REQUIRE(formulation == APHIP_TOTAL_LAGRANGIAN || formulation == APHIP_UPDATED_LAGRANGIAN ||
                      formulation == APHIP_SEMI_LAGRANGIAN,
                  "formulation must be total, updated or semi-Lagrangian");
    APHIP_REQUIRE(formulation == APHIP_TOTAL_LAGRANGIAN || !c.cells.ready,
                  "the updated formulation applies to the Gaussian elements, not to RKPM cells");
    APHIP_REQUIRE(formulation == APHIP_TOTAL_LAGRANGIAN || c.n_specimens == 0,
                  "ensembles use the total Lagrangian formulation");
    APHIP_REQUIRE(formulation != APHIP_SEMI_LAGRANGIAN || !c.mat.rate_form(),
                  "the semi-Lagrangian formulation does not take the rate-form materials");
    APHIP_REQUIRE(interval >= 1, "the reference update interval must be at least 1");
    APHIP_HIP(hipSetDevice(c.device));
    APHIP_HIP(hipStreamSynchronize(c.stream));
    c.graph_reset();
    c.formulation = formulation;
    c.semi_interval = interval;
    c.steps_done = 0;
    c.Href.release();
    if (formulation != APHIP_TOTAL_LAGRANGIAN) {
        for (int s = 0; s < 2; ++s) {
            c.xc[s].alloc(3 * c.nn);
            APHIP_HIP(hipMemcpyAsync(c.xc[s].p, c.X.p, sizeof(double) * 3 * c.nn, hipMemcpyDeviceToDevice, c.stream));
        }
        c.dinc.alloc(3 * c.nn);
        c.dinc.zero(c.stream);
        if (formulation == APHIP_SEMI_LAGRANGIAN) {
            c.Href.alloc(9 * std::max<int64_t>(c.ne, 1));
            c.Href.zero(c.stream);
        }
    } else {
        c.xc[0].release();
        c.xc[1].release();
        c.dinc.release();
    }
    alloc_state(c);
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

int aphip_set_formulation(aphip_ctx *c, int32_t formulation) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    set_formulation(*c, formulation, 1);
    APHIP_API_END
}

int aphip_set_formulation_semi(aphip_ctx *c, int32_t update_interval) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    set_formulation(*c, APHIP_SEMI_LAGRANGIAN, update_interval);
    APHIP_API_END
}

int aphip_set_material_fields(aphip_ctx *c, const double *E, const double *nu) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_REQUIRE(!(c->cells.ready && c->cells.nodal), "nodal cells take one material per block");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    if (!E && !nu) {
        c->has_fields = false;
        c->elem_E_h.clear();
        c->elem_nu_h.clear();
        c->elem_lambda.release();
        c->elem_mu.release();
        return APHIP_OK;
    }
    APHIP_REQUIRE(E && nu, "E and nu must both be given");
    std::vector<double> lam(c->ne), mu(c->ne);
    c->elem_E_h.assign(c->ne, 0.0);
    c->elem_nu_h.assign(c->ne, 0.0);
    for (int64_t e = 0; e < c->ne; ++e) {
        const int64_t src = c->elem_iperm[e];
        const double Ee = E[src], ne = nu[src];
        APHIP_REQUIRE(ne >= -1.0 && ne < 0.5 && Ee > 0.0, "per-element E/nu out of bounds");
        mu[e] = Ee / (2 * (1 + ne));
        lam[e] = (Ee * ne) / ((1 + ne) * (1 - 2 * ne));
        c->elem_E_h[e] = Ee;
        c->elem_nu_h[e] = ne;
    }
    c->elem_lambda.upload(lam.data(), lam.size(), c->stream);
    c->elem_mu.upload(mu.data(), mu.size(), c->stream);
    c->has_fields = true;
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_compute_mass(aphip_ctx *c, double *total) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, false);
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    double tot = 0.0;
    if (c->cells.ready) {
        rk_mass(*c, &tot);
    } else {
        fem_lumped_mass(*c, c->mat.density);
        std::vector<double> m(c->nn);
        c->mass.download(m.data(), m.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (double x : m) tot += x;
    }
    c->has_mass = true;
    if (total) *total = tot;
    APHIP_API_END
}

int aphip_set_fixed(aphip_ctx *c, const uint8_t *mask) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    if (!mask) {
        c->fixed.release();
        c->has_fixed = false;
        c->all_fixed = false;
        return APHIP_OK;
    }
    std::vector<uint8_t> bits(c->nn, 0);
    bool all = true;
    for (int64_t i = 0; i < c->nn; ++i) {
        const int64_t src = c->node_iperm[i];
        for (int d = 0; d < 3; ++d)
            if (mask[3 * src + d]) bits[i] |= (uint8_t)(1u << d);
        all = all && bits[i] == 7u;
    }
    c->fixed.upload(bits.data(), bits.size(), c->stream);
    c->has_fixed = true;
    c->all_fixed = all;
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_set_field(aphip_ctx *c, int32_t field, const double *data) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && data != nullptr, "NULL argument");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    if (field <= APHIP_FIELD_F) {
        DevBuf<double> *b = field == APHIP_FIELD_U ? &c->u : field == APHIP_FIELD_V ? &c->v
                            : field == APHIP_FIELD_A ? &c->a : &c->f;
        std::vector<double> h(3 * c->nn);
        for (int64_t i = 0; i < c->nn; ++i)
            for (int d = 0; d < 3; ++d) h[3 * i + d] = data[3 * (int64_t)c->node_iperm[i] + d];
        b->upload(h.data(), h.size(), c->stream);
    } else if (field == APHIP_FIELD_MASS) {
        std::vector<double> h(c->nn);
        for (int64_t i = 0; i < c->nn; ++i) {
            h[i] = data[c->node_iperm[i]];
            APHIP_REQUIRE(h[i] > 0.0, "nodal mass must be positive");
        }
        c->mass.upload(h.data(), h.size(), c->stream);
        c->has_mass = true;
    } else if (field == APHIP_FIELD_STATE) {
        const int64_t ns = c->state_count();
        APHIP_REQUIRE(ns > 0, "material has no state");
        const int64_t per = ns / c->ne;
        std::vector<double> h(ns, 0.0);
        if (c->cells.ready && c->cells.nodal) {  // per subcell: the value of its first element
            for (int64_t sc = 0; sc < c->cells.ncells; ++sc)
                for (int q = 0; q < per; ++q)
                    h[sc * per + q] = data[(int64_t)c->elem_iperm[c->cells.sub_first_elem[sc]] * per + q];
        } else {
            for (int64_t e = 0; e < c->ne; ++e)
                for (int q = 0; q < per; ++q) h[e * per + q] = data[(int64_t)c->elem_iperm[e] * per + q];
        }
        c->state[c->state_np1].upload(h.data(), h.size(), c->stream);
    } else {
        throw Error(APHIP_ERR_INVALID, "field cannot be set");
    }
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_get_field(aphip_ctx *c, int32_t field, double *out) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && out != nullptr, "NULL argument");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
    if (field <= APHIP_FIELD_F) {
        const DevBuf<double> *b = field == APHIP_FIELD_U ? &c->u : field == APHIP_FIELD_V ? &c->v
                                  : field == APHIP_FIELD_A ? &c->a : &c->f;
        std::vector<double> h(3 * c->nn);
        b->download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < c->nn; ++i)
            for (int d = 0; d < 3; ++d) out[3 * (int64_t)c->node_iperm[i] + d] = h[3 * i + d];
    } else if (field == APHIP_FIELD_MASS) {
        APHIP_REQUIRE(c->has_mass, "mass not computed or set");
        std::vector<double> h(c->nn);
        c->mass.download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < c->nn; ++i) out[c->node_iperm[i]] = h[i];
    } else if (field == APHIP_FIELD_H || field == APHIP_FIELD_P) {
        // with nodal cells the values live per subcell: every element reads its subcell's
        // (the reference's PopulateElementOutputs, SmoothedCellDataProcessor.h:676-720)
        const bool nodal = c->cells.ready && c->cells.nodal;
        const int64_t stride = nodal ? c->cells.ncells : c->ne;
        const DevBuf<double> &b = field == APHIP_FIELD_H ? c->Hnp1() : c->Pnp1();
        std::vector<double> h(9 * c->ne);
        b.download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < c->ne; ++e) {
            const int64_t src = nodal ? c->cells.sub_of_elem[e] : e;
            for (int k = 0; k < 9; ++k) out[9 * (int64_t)c->elem_iperm[e] + k] = h[k * stride + src];
        }
    } else if (field == APHIP_FIELD_STATE) {
        const int64_t ns = c->state_count();
        APHIP_REQUIRE(ns > 0, "material has no state");
        const bool nodal = c->cells.ready && c->cells.nodal;
        const int64_t per = ns / c->ne;
        std::vector<double> h(ns);
        c->state[c->state_np1].download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < c->ne; ++e) {
            const int64_t src = nodal ? c->cells.sub_of_elem[e] : e;
            for (int q = 0; q < per; ++q) out[(int64_t)c->elem_iperm[e] * per + q] = h[src * per + q];
        }
    } else if (field == APHIP_FIELD_RK_DISPLACEMENT || field == APHIP_FIELD_RK_FORCE_LOCAL) {
        APHIP_REQUIRE(c->cells.ready && c->cells.two_pass, "field exists with the two-pass RK method only");
        const DevBuf<double> &b = field == APHIP_FIELD_RK_DISPLACEMENT ? c->cells.u_phys : c->cells.f_local;
        std::vector<double> h(3 * c->nn);
        b.download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < c->nn; ++i)
            for (int d = 0; d < 3; ++d) out[3 * (int64_t)c->node_iperm[i] + d] = h[3 * i + d];
    } else if (field == APHIP_FIELD_VOLUME) {
        const DevBuf<double> &b = c->cells.ready ? (c->cells.nodal ? c->cells.evol : c->cells.volume) : c->volume;
        std::vector<double> h(c->ne);
        b.download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < c->ne; ++e) out[c->elem_iperm[e]] = h[e];
    } else {
        throw Error(APHIP_ERR_INVALID, "unknown field");
    }
    APHIP_API_END
}

}  // extern "C"

// A caller array must be device memory of the context's device holding at least
// `bytes` from p on (checked against the allocation that contains p), so that no
// kernel of the transfer reads or writes outside the caller's allocation.
static void require_device_array(const Ctx &c, const void *p, size_t bytes) {
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(APHIP_ERR_INVALID, "array is not device memory known to HIP");
    }
    APHIP_REQUIRE(at.type == hipMemoryTypeDevice && at.device == c.device,
                  "array is not device memory of the context's device");
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(APHIP_ERR_INVALID, "array is not inside a device allocation");
    }
    APHIP_REQUIRE((const char *)p + bytes <= (const char *)base + size,
                  "device array is shorter than the field");
}

static DevBuf<double> &node_vector(Ctx &c, int32_t field) {
    return field == APHIP_FIELD_U ? c.u : field == APHIP_FIELD_V ? c.v : field == APHIP_FIELD_A ? c.a : c.f;
}

extern "C" {

int aphip_set_field_device(aphip_ctx *c, int32_t field, const double *data) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && data != nullptr, "NULL argument");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
    if (c->graph_exec) {  // a captured step sequence is re-captured after a field changes
        APHIP_HIP(hipStreamSynchronize(c->stream));
        c->graph_reset();
    }
    if (field >= APHIP_FIELD_U && field <= APHIP_FIELD_F) {
        require_device_array(*c, data, sizeof(double) * 3 * c->nn);
        field_gather_rows(*c, c->nn, 3, c->d_node_iperm.p, data, node_vector(*c, field).p);
    } else if (field == APHIP_FIELD_MASS) {
        require_device_array(*c, data, sizeof(double) * c->nn);
        c->mass.alloc(c->nn);
        c->has_mass = false;
        APHIP_REQUIRE(field_gather_mass(*c, data) == 0, "nodal mass must be positive");
        c->has_mass = true;
    } else if (field == APHIP_FIELD_STATE) {
        const int64_t ns = c->state_count();
        APHIP_REQUIRE(ns > 0, "material has no state");
        APHIP_REQUIRE(!(c->cells.ready && c->cells.nodal),
                      "element fields of nodal cells are transferred through host arrays");
        require_device_array(*c, data, sizeof(double) * ns);
        field_gather_rows(*c, c->ne, (int)(ns / c->ne), c->d_elem_iperm.p, data, c->state[c->state_np1].p);
    } else {
        throw Error(APHIP_ERR_INVALID, "field cannot be set");
    }
    APHIP_API_END
}

int aphip_get_field_device(aphip_ctx *c, int32_t field, double *out) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && out != nullptr, "NULL argument");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_REQUIRE(!(c->cells.ready && c->cells.nodal) || field <= APHIP_FIELD_MASS ||
                      field == APHIP_FIELD_RK_DISPLACEMENT || field == APHIP_FIELD_RK_FORCE_LOCAL,
                  "element fields of nodal cells are transferred through host arrays");
    APHIP_HIP(hipSetDevice(c->device));
    if (field >= APHIP_FIELD_U && field <= APHIP_FIELD_F) {
        require_device_array(*c, out, sizeof(double) * 3 * c->nn);
        field_scatter_rows(*c, c->nn, 3, c->d_node_iperm.p, node_vector(*c, field).p, out);
    } else if (field == APHIP_FIELD_MASS) {
        APHIP_REQUIRE(c->has_mass, "mass not computed or set");
        require_device_array(*c, out, sizeof(double) * c->nn);
        field_scatter_rows(*c, c->nn, 1, c->d_node_iperm.p, c->mass.p, out);
    } else if (field == APHIP_FIELD_H || field == APHIP_FIELD_P) {
        require_device_array(*c, out, sizeof(double) * 9 * c->ne);
        field_scatter_tensor(*c, (field == APHIP_FIELD_H ? c->Hnp1() : c->Pnp1()).p, out);
    } else if (field == APHIP_FIELD_STATE) {
        const int64_t ns = c->state_count();
        APHIP_REQUIRE(ns > 0, "material has no state");
        require_device_array(*c, out, sizeof(double) * ns);
        field_scatter_rows(*c, c->ne, (int)(ns / c->ne), c->d_elem_iperm.p, c->state[c->state_np1].p, out);
    } else if (field == APHIP_FIELD_RK_DISPLACEMENT || field == APHIP_FIELD_RK_FORCE_LOCAL) {
        APHIP_REQUIRE(c->cells.ready && c->cells.two_pass, "field exists with the two-pass RK method only");
        require_device_array(*c, out, sizeof(double) * 3 * c->nn);
        field_scatter_rows(*c, c->nn, 3, c->d_node_iperm.p,
                           (field == APHIP_FIELD_RK_DISPLACEMENT ? c->cells.u_phys : c->cells.f_local).p, out);
    } else if (field == APHIP_FIELD_VOLUME) {
        require_device_array(*c, out, sizeof(double) * c->ne);
        field_scatter_rows(*c, c->ne, 1, c->d_elem_iperm.p, (c->cells.ready ? c->cells.volume : c->volume).p, out);
    } else {
        throw Error(APHIP_ERR_INVALID, "unknown field");
    }
    APHIP_API_END
}

int aphip_internal_force(aphip_ctx *c, double time, double dt) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, false);
    (void)time;
    APHIP_REQUIRE(std::isfinite(dt) && dt >= 0.0, "time increment must be finite and non-negative");
    c->rate_dt = dt;
    APHIP_HIP(hipSetDevice(c->device));
    internal_force_eval(*c);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_power_method_time_increment(aphip_ctx *c, double previous_time_increment, int32_t num_iterations,
                                      double epsilon, double tolerance, double variation_tolerance, uint64_t seed,
                                      double *time_increment, int32_t *iterations, int32_t *converged) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && time_increment != nullptr, "NULL argument");
    require_ready(*c, true);
    APHIP_REQUIRE(num_iterations > 0 && epsilon > 0.0 && tolerance > 0.0 && variation_tolerance >= 0.0,
                  "power method needs positive iterations, epsilon and tolerance");
    APHIP_REQUIRE(previous_time_increment >= 0.0, "previous time increment must be >= 0");
    APHIP_HIP(hipSetDevice(c->device));
    const auto n = power_method_sizes(*c);
    DevBuf<double> *b = c->pm_work;  // kept between calls; alloc keeps a buffer of the same size
    for (int k = 0; k < 8; ++k) b[k].alloc((size_t)n[k]);
    const PowerMethodWork w{b[0].p, b[1].p, b[2].p, b[3].p, b[4].p, b[5].p, b[6].p, b[7].p};
    const PowerMethodResult r = power_method(*c, w, previous_time_increment, num_iterations, epsilon, tolerance,
                                             variation_tolerance, seed);
    *time_increment = r.dt;
    if (iterations) *iterations = r.iterations;
    if (converged) *converged = r.converged ? 1 : 0;
    APHIP_API_END
}

int aphip_power_method_time_increment_noalloc(aphip_ctx *c, double previous_time_increment, int32_t num_iterations,
                                              double epsilon, double tolerance, double variation_tolerance,
                                              uint64_t seed, double *time_increment, int32_t *iterations,
                                              int32_t *converged) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && time_increment != nullptr, "NULL argument");
    require_ready(*c, true);
    APHIP_REQUIRE(num_iterations > 0 && epsilon > 0.0 && tolerance > 0.0 && variation_tolerance >= 0.0,
                  "power method needs positive iterations, epsilon and tolerance");
    APHIP_REQUIRE(previous_time_increment >= 0.0, "previous time increment must be >= 0");
    // the work buffers, each on a 256-byte boundary inside this device's scratch
    const auto n = power_method_sizes(*c);
    size_t off[9] = {0};
    for (int k = 0; k < 8; ++k) off[k + 1] = off[k] + (((size_t)n[k] * sizeof(double) + 255) & ~size_t(255));
    if (off[8] > kScratchBytes)
        throw Error(APHIP_ERR_INVALID, "the allocation-free power method holds its work buffers in " +
                                           std::to_string(kScratchBytes >> 20) + " MiB of scratch per GPU; this mesh needs " +
                                           std::to_string(off[8]) + " bytes (use aphip_power_method_time_increment)");
    APHIP_HIP(hipSetDevice(c->device));
    unsigned char *base = static_cast<unsigned char *>(scratch_address(c->device));
    APHIP_REQUIRE(base != nullptr, "scratch of this device not resolved");
    double *p[8];
    for (int k = 0; k < 8; ++k) p[k] = reinterpret_cast<double *>(base + off[k]);
    const PowerMethodWork w{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
    const PowerMethodResult r = power_method(*c, w, previous_time_increment, num_iterations, epsilon, tolerance,
                                             variation_tolerance, seed);
    *time_increment = r.dt;
    if (iterations) *iterations = r.iterations;
    if (converged) *converged = r.converged ? 1 : 0;
    APHIP_API_END
}

int aphip_initialize(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, true);
    c->steps_done = 0;
    APHIP_HIP(hipSetDevice(c->device));
    if (use_tiles(*c)) {
        tile_run(*c, 0.0, 1);
    } else {
        force_pass(*c);
        node_pass(*c, 0.0, 1);
    }
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_step(aphip_ctx *c, double dt, int32_t n_steps) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, true);
    APHIP_REQUIRE(n_steps >= 0, "n_steps must be non-negative");
    APHIP_REQUIRE(std::isfinite(dt) && dt >= 0.0, "time increment must be finite and non-negative");
    APHIP_HIP(hipSetDevice(c->device));
    int32_t done = 0;
    if (c->graph_enabled && !c->prof.enabled && n_steps >= 2 && c->formulation != APHIP_SEMI_LAGRANGIAN) {
        // two steps per graph: the N/NP1 state buffers swap twice and return to their roles
        if (!c->graph_exec || c->graph_dt != dt || c->graph_parity != c->state_np1) {
            c->graph_reset();
            const int parity = c->state_np1;
            hipGraph_t g = nullptr;
            APHIP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            try {  // the second step takes the first one's prediction
                step_once(*c, dt, false, true);
                step_once(*c, dt, true, false);
            } catch (...) {
                (void)hipStreamEndCapture(c->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                c->state_np1 = parity;
                throw;
            }
            APHIP_HIP(hipStreamEndCapture(c->stream, &g));
            const hipError_t ie = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            APHIP_HIP(ie);
            c->graph_dt = dt;
            c->graph_parity = parity;
        }
        for (; done + 2 <= n_steps; done += 2) APHIP_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    }
    for (bool pred = false; done < n_steps; ++done) {
        const bool more = done + 1 < n_steps && c->formulation != APHIP_SEMI_LAGRANGIAN;
        prof_begin(*c, 0);
        step_once(*c, dt, pred, more);
        prof_end(*c, 0);
        pred = more;
        // the reference's StepScheduler(interval, interval) after step n (Explicit.cpp:195-212)
        if (c->formulation == APHIP_SEMI_LAGRANGIAN && ++c->steps_done % c->semi_interval == 0) fem_semi_update(*c);
    }
    APHIP_API_END
}

int aphip_set_graph(aphip_ctx *c, int32_t enable) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    c->graph_enabled = enable != 0;
    APHIP_API_END
}

int aphip_autotune(aphip_ctx *c, double dt, int32_t trial_steps, int32_t *chosen) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, true);
    APHIP_REQUIRE(!c->cells.ready && !(c->flags & APHIP_FLAG_TWO_PASS), "autotuning applies to the tiled element path");
    APHIP_REQUIRE(trial_steps > 0, "trial_steps must be positive");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_REQUIRE(c->formulation == APHIP_TOTAL_LAGRANGIAN, "autotuning applies to the total Lagrangian formulation");
    // save the simulation state: node fields in caller numbering (tilings renumber the
    // nodes), element fields (never reordered by a tiling) as they lie, both halves of
    // every N / NP1 pair
    std::vector<double> su(3 * c->nn), sv(3 * c->nn), sa(3 * c->nn), sf(3 * c->nn);
    const int64_t ns = c->state_count();
    const int parity = c->state_np1;
    const double rate_dt = c->rate_dt;
    get_node_field(*c, c->u, su.data());
    get_node_field(*c, c->v, sv.data());
    get_node_field(*c, c->a, sa.data());
    get_node_field(*c, c->f, sf.data());
    const int prev_te = c->tiles.te;
    struct Saved {
        DevBuf<double> *b;
        std::vector<double> h;
    };
    std::vector<Saved> esave;
    for (DevBuf<double> *b : {&c->H, &c->P, &c->H1, &c->P1})
        if (b->p) esave.push_back({b, std::vector<double>(9 * c->ne)});
    if (ns)
        for (DevBuf<double> *b : {&c->state[0], &c->state[1]}) esave.push_back({b, std::vector<double>(ns)});
    for (auto &s : esave) s.b->download(s.h.data(), s.h.size(), c->stream);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    struct Cand {
        int te, threads;
        bool pipelined;
    };
    std::vector<Cand> cand;
    if (c->etype == APHIP_TET4)
        cand = {{64, 64, false},  {128, 128, false}, {256, 256, false}, {512, 512, false},
                {1024, 1024, false}, {128, 128, true}, {256, 256, true}};
    else
        cand = {{64, 64, false}, {128, 128, false}, {256, 256, false}, {16, 128, false}, {32, 256, false},
                {64, 512, false}};
    const bool was_prof = c->prof.enabled;
    Cand best{prev_te, c->tiles.threads, c->tiles.use_stream};
    double best_ms = 1e300;
    hipEvent_t e0, e1;
    APHIP_HIP(hipEventCreate(&e0));
    APHIP_HIP(hipEventCreate(&e1));
    c->prof.enabled = false;
    for (const auto &tt : cand) {
        try {
            tile_build(*c, tt.te, tt.threads);
        } catch (const Error &) {
            continue;  // this shape does not fit the mesh (e.g. a tile needs more LDS than a CU has)
        }
        c->tiles.use_stream = tt.pipelined;
        step_once(*c, dt);  // warm-up
        APHIP_HIP(hipEventRecord(e0, c->stream));
        for (int k = 0; k < trial_steps; ++k) step_once(*c, dt);
        APHIP_HIP(hipEventRecord(e1, c->stream));
        APHIP_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        APHIP_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = tt;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->prof.enabled = was_prof;
    tile_build(*c, best.te, best.threads);
    c->tiles.use_stream = best.pipelined;
    set_node_field(*c, c->u, su.data());
    set_node_field(*c, c->v, sv.data());
    set_node_field(*c, c->a, sa.data());
    set_node_field(*c, c->f, sf.data());
    for (auto &s : esave) s.b->upload(s.h.data(), s.h.size(), c->stream);
    c->state_np1 = parity;
    c->rate_dt = rate_dt;
    APHIP_HIP(hipStreamSynchronize(c->stream));
    if (chosen) *chosen = (best.pipelined ? 100000000 : 0) + best.te * 10000 + best.threads;
    APHIP_API_END
}

int aphip_batch_begin(aphip_ctx *c, int32_t elem_type, int32_t kind) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(elem_type == APHIP_TET4 || elem_type == APHIP_HEX8, "element type must be TET4 or HEX8");
    APHIP_REQUIRE(kind >= APHIP_LINEAR_ELASTIC && kind <= APHIP_SVK,
                  "ensembles take the linear-elastic, neo-Hookean, J2 and SVK materials");
    c->batch = BatchBuild();
    c->batch.open = true;
    c->batch.etype = elem_type;
    c->batch.kind = kind;
    APHIP_API_END
}

int aphip_batch_add(aphip_ctx *c, int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn,
                    const aphip_material_params *p, double dt, int32_t *specimen) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && X != nullptr && conn != nullptr && p != nullptr, "NULL argument");
    BatchBuild &B = c->batch;
    APHIP_REQUIRE(B.open, "aphip_batch_begin has not been called");
    APHIP_REQUIRE(n_nodes > 0 && n_elem > 0, "empty specimen");
    APHIP_REQUIRE(std::isfinite(dt) && dt >= 0.0, "time increment must be finite and non-negative");
    MaterialConst m;
    resolve_material(B.kind, p, m);  // validates the parameters
    const int npe = B.etype == APHIP_TET4 ? 4 : 8;
    const int64_t base = B.node_seg.back();
    APHIP_REQUIRE(base + n_nodes < (int64_t)INT32_MAX, "ensemble too large");
    for (int64_t i = 0; i < n_elem * npe; ++i)
        APHIP_REQUIRE(conn[i] >= 0 && conn[i] < n_nodes, "connectivity references a node out of range");
    B.X.insert(B.X.end(), X, X + 3 * n_nodes);
    for (int64_t i = 0; i < n_elem * npe; ++i) B.conn.push_back((int32_t)(base + conn[i]));
    B.node_seg.push_back(base + n_nodes);
    B.elem_seg.push_back(B.elem_seg.back() + n_elem);
    B.params.push_back(*p);
    B.dt.push_back(dt);
    if (specimen) *specimen = (int32_t)(B.params.size() - 1);
    APHIP_API_END
}

int aphip_batch_end(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    BatchBuild B = std::move(c->batch);
    c->batch = BatchBuild();
    APHIP_REQUIRE(B.open, "aphip_batch_begin has not been called");
    APHIP_REQUIRE(!B.params.empty(), "the ensemble has no specimens");
    APHIP_HIP(hipSetDevice(c->device));
    // material: kind and J2 constants of the first specimen, E / nu / rho per element
    MaterialConst m;
    resolve_material(B.kind, &B.params[0], m);
    c->mat = m;
    const int64_t nn = B.node_seg.back(), ne = B.elem_seg.back();
    setup_mesh(*c, B.etype, nn, B.X.data(), ne, B.conn.data(), B.node_seg, B.elem_seg);
    const int64_t nspec = (int64_t)B.params.size();
    std::vector<double> lam(ne), mu(ne), rho(ne), ndt(nn);
    for (int64_t sp = 0; sp < nspec; ++sp) {
        MaterialConst ms;
        resolve_material(B.kind, &B.params[sp], ms);
        for (int64_t e = B.elem_seg[sp]; e < B.elem_seg[sp + 1]; ++e) {
            const int64_t ie = c->elem_perm[e];
            lam[ie] = ms.lambda;
            mu[ie] = ms.mu;
            rho[ie] = ms.density;
        }
        for (int64_t i = B.node_seg[sp]; i < B.node_seg[sp + 1]; ++i) ndt[c->node_perm[i]] = B.dt[sp];
    }
    c->elem_lambda.upload(lam.data(), ne, c->stream);
    c->elem_mu.upload(mu.data(), ne, c->stream);
    c->elem_rho.upload(rho.data(), ne, c->stream);
    c->node_dt.upload(ndt.data(), nn, c->stream);
    c->has_fields = true;
    c->n_specimens = nspec;
    c->spec_node_seg = B.node_seg;
    c->spec_elem_seg = B.elem_seg;
    alloc_state(*c);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_batch_ranges(aphip_ctx *c, int32_t sp, int64_t *first_node, int64_t *n_nodes, int64_t *first_elem,
                       int64_t *n_elem) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(sp >= 0 && sp < c->n_specimens, "specimen index out of range");
    if (first_node) *first_node = c->spec_node_seg[sp];
    if (n_nodes) *n_nodes = c->spec_node_seg[sp + 1] - c->spec_node_seg[sp];
    if (first_elem) *first_elem = c->spec_elem_seg[sp];
    if (n_elem) *n_elem = c->spec_elem_seg[sp + 1] - c->spec_elem_seg[sp];
    APHIP_API_END
}

int aphip_synchronize(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}


int aphip_rk_setup(aphip_ctx *c, double scale) { return aphip_rk_setup_method(c, scale, 1); }

int aphip_rk_setup_method(aphip_ctx *c, double scale, int32_t use_one_pass_method) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, false);
    APHIP_REQUIRE(c->etype == APHIP_TET4 || c->etype == APHIP_HEX8,
                  "RKPM element-as-cell smoothing needs a Tet4 or Hex8 background mesh");
    APHIP_REQUIRE(c->mat.state_per_qp == 0, "RKPM cells keep one material state per cell, not per Gauss point");
    APHIP_REQUIRE(c->formulation == APHIP_TOTAL_LAGRANGIAN,
                  "the updated formulation applies to the Gaussian elements, not to RKPM cells");
    APHIP_REQUIRE(scale >= 0.0, "Kernel radius scale factor must be non-negative.");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    rk_setup(*c, scale, use_one_pass_method == 0);
    c->has_mass = false;
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_strain_smoothing_setup(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, false);
    APHIP_REQUIRE(c->etype != APHIP_HEX8, "Strain Smoothing is not supported for finite elements and Hexahedron8");
    APHIP_REQUIRE(c->mat.state_per_qp == 0, "smoothed cells keep one material state per cell, not per Gauss point");
    APHIP_REQUIRE(c->formulation == APHIP_TOTAL_LAGRANGIAN,
                  "the updated formulation applies to the Gaussian elements, not to smoothed cells");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    rk_setup(*c, 0.0, false, true);
    c->has_mass = false;
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_rk_nodal_setup(aphip_ctx *c, double scale, int32_t subdomains, int32_t use_f_bar) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, false);
    APHIP_REQUIRE(c->mat.state_per_qp == 0, "RKPM cells keep one material state per cell, not per Gauss point");
    APHIP_REQUIRE(c->formulation == APHIP_TOTAL_LAGRANGIAN,
                  "the updated formulation applies to the Gaussian elements, not to RKPM cells");
    APHIP_REQUIRE(!c->has_fields, "nodal cells take one material per block");
    APHIP_REQUIRE(scale >= 0.0, "Kernel radius scale factor must be non-negative.");
    APHIP_REQUIRE(!use_f_bar || c->mat.kind == APHIP_LINEAR_ELASTIC || c->mat.kind == APHIP_NEO_HOOKEAN ||
                      c->mat.kind == APHIP_J2 || c->mat.kind == APHIP_SVK,
                  "F-bar takes the linear elastic, neo-Hookean, J2 and SVK materials");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    c->graph_reset();
    rk_setup(*c, scale, false, false, subdomains, use_f_bar != 0);
    c->has_mass = false;
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_rk_get_nodal_cells(aphip_ctx *c, int64_t *n_cells, int64_t *n_subcells, int64_t *cell_subcells,
                             int32_t *element_subcell, uint8_t *active, int32_t *f_bar) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(c->cells.ready && c->cells.nodal, "aphip_rk_nodal_setup has not run");
    const CellData &D = c->cells;
    if (n_cells) *n_cells = D.n_fcells;
    if (n_subcells) *n_subcells = D.ncells;
    if (f_bar) *f_bar = D.fbar ? 1 : 0;
    if (cell_subcells) std::memcpy(cell_subcells, D.fcell_sub_h.data(), sizeof(int64_t) * (D.n_fcells + 1));
    if (element_subcell)
        for (int64_t e = 0; e < c->ne; ++e) element_subcell[c->elem_iperm[e]] = D.sub_of_elem[e];
    if (active)
        for (int64_t i = 0; i < c->nn; ++i) active[c->node_iperm[i]] = i < c->n_dof ? 1 : 0;
    APHIP_API_END
}

int aphip_rk_get_neighbors(aphip_ctx *c, int32_t *num, int32_t *nbr, int64_t *n_truncated) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(c->cells.ready, "aphip_rk_setup has not run");
    APHIP_HIP(hipSetDevice(c->device));
    const int64_t M = 40;
    std::vector<int32_t> hn(c->nn), hb(c->nn * M);
    c->cells.num_nbr.download(hn.data(), hn.size(), c->stream);
    c->cells.nbr.download(hb.data(), hb.size(), c->stream);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->nn; ++i) {
        const int64_t dst = c->node_iperm[i];
        if (num) num[dst] = hn[i];
        if (nbr)
            for (int64_t k = 0; k < M; ++k) {
                const int32_t j = hb[i * M + k];
                nbr[dst * M + k] = (k < hn[i] && j >= 0) ? c->node_iperm[j] : -1;
            }
    }
    if (n_truncated) *n_truncated = c->cells.n_truncated;
    APHIP_API_END
}

int aphip_rk_get_function_values(aphip_ctx *c, double *phi) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && phi != nullptr, "NULL argument");
    APHIP_REQUIRE(c->cells.ready, "aphip_rk_setup has not run");
    APHIP_HIP(hipSetDevice(c->device));
    const int64_t M = 40;
    std::vector<double> h(c->nn * M);
    c->cells.phi.download(h.data(), h.size(), c->stream);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->nn; ++i)
        std::memcpy(&phi[(int64_t)c->node_iperm[i] * M], &h[i * M], sizeof(double) * M);
    APHIP_API_END
}

int aphip_rk_get_cells(aphip_ctx *c, int64_t *nnz, int64_t *start, int32_t *nodes, double *dphi, double *volume) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(c->cells.ready, "aphip_rk_setup has not run");
    APHIP_HIP(hipSetDevice(c->device));
    if (nnz) *nnz = c->cells.nnz;
    if (!start && !nodes && !dphi && !volume) return APHIP_OK;
    if (c->cells.nodal) {  // subcells, in their own order (aphip_rk_get_nodal_cells)
        const int64_t ns = c->cells.ncells;
        std::vector<int64_t> hs(ns + 1);
        std::vector<int32_t> hn(c->cells.nnz);
        c->cells.start.download(hs.data(), hs.size(), c->stream);
        c->cells.nodes.download(hn.data(), hn.size(), c->stream);
        if (dphi) c->cells.dphi.download(dphi, 3 * c->cells.nnz, c->stream);
        if (volume) c->cells.volume.download(volume, ns, c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        if (start) std::memcpy(start, hs.data(), sizeof(int64_t) * (ns + 1));
        if (nodes)
            for (int64_t t = 0; t < c->cells.nnz; ++t) nodes[t] = c->node_iperm[hn[t]];
        return APHIP_OK;
    }
    std::vector<int64_t> hs(c->ne + 1);
    std::vector<int32_t> hn(c->cells.nnz);
    std::vector<double> hd(3 * c->cells.nnz), hv(c->ne);
    c->cells.start.download(hs.data(), hs.size(), c->stream);
    c->cells.nodes.download(hn.data(), hn.size(), c->stream);
    c->cells.dphi.download(hd.data(), hd.size(), c->stream);
    c->cells.volume.download(hv.data(), hv.size(), c->stream);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    // caller element order
    std::vector<int64_t> cs(c->ne + 1, 0);
    for (int64_t e = 0; e < c->ne; ++e) cs[c->elem_iperm[e] + 1] = hs[e + 1] - hs[e];
    for (int64_t e = 0; e < c->ne; ++e) cs[e + 1] += cs[e];
    if (start) std::memcpy(start, cs.data(), sizeof(int64_t) * (c->ne + 1));
    for (int64_t e = 0; e < c->ne; ++e) {
        const int64_t ce = c->elem_iperm[e];
        const int64_t len = hs[e + 1] - hs[e];
        for (int64_t t = 0; t < len; ++t) {
            if (nodes) nodes[cs[ce] + t] = c->node_iperm[hn[hs[e] + t]];
            if (dphi)
                for (int d = 0; d < 3; ++d) dphi[3 * (cs[ce] + t) + d] = hd[3 * (hs[e] + t) + d];
        }
        if (volume) volume[ce] = hv[e];
    }
    APHIP_API_END
}

int aphip_internal_force_vjp(aphip_ctx *c, const double *f_bar, double *u_bar, double *E_bar, double *nu_bar) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && f_bar != nullptr, "NULL argument");
    APHIP_HIP(hipSetDevice(c->device));
    force_vjp(*c, f_bar, u_bar, E_bar, nu_bar);
    APHIP_API_END
}

int aphip_adjoint_forward(aphip_ctx *c, double dt, int32_t n_steps, int32_t interval) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(std::isfinite(dt) && dt >= 0.0, "time increment must be finite and non-negative");
    APHIP_HIP(hipSetDevice(c->device));
    adjoint_forward(*c, dt, n_steps, interval);
    APHIP_API_END
}

int aphip_adjoint_backward(aphip_ctx *c, const double *ub_T, const double *vb_T, double *E_bar, double *nu_bar,
                           double *v0_bar) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_HIP(hipSetDevice(c->device));
    adjoint_backward(*c, ub_T, vb_T, E_bar, nu_bar, v0_bar);
    APHIP_API_END
}

int aphip_set_profiling(aphip_ctx *c, int32_t enable) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    c->prof.enabled = enable != 0;
    APHIP_API_END
}

int aphip_get_profile(aphip_ctx *c, int32_t which, int64_t *launches, double *ms) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && which >= 0 && which < 3, "bad argument");
    APHIP_HIP(hipSetDevice(c->device));
    prof_collect(*c);
    if (launches) *launches = c->prof.launches[which];
    if (ms) *ms = c->prof.ms[which];
    APHIP_API_END
}

int aphip_reset_profile(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_HIP(hipSetDevice(c->device));
    prof_collect(*c);
    for (int w = 0; w < 3; ++w) {
        c->prof.launches[w] = 0;
        c->prof.ms[w] = 0.0;
    }
    APHIP_API_END
}

int aphip_bandwidth_probe(int32_t device, int64_t nbytes, int32_t reps, double *ms_per_copy) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(ms_per_copy != nullptr, "ms_per_copy is NULL");
    *ms_per_copy = bandwidth_probe(device, nbytes, reps);
    APHIP_API_END
}

// ---- Exodus II results (the reference's field output) --------------------------

static const char *const kTensorSuffix[9] = {"xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz"};  // row-major

int aphip_results_open(aphip_ctx *c, const char *path, const char *title) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && path != nullptr, "NULL argument");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
    c->results.reset();
    // coordinates and connectivity in the caller's numbering
    std::vector<double> Xi(3 * c->nn), x(c->nn), y(c->nn), z(c->nn);
    c->X.download(Xi.data(), Xi.size(), c->stream);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->nn; ++i) {
        const int64_t o = c->node_iperm[i];
        x[o] = Xi[3 * i];
        y[o] = Xi[3 * i + 1];
        z[o] = Xi[3 * i + 2];
    }
    std::vector<int32_t> conn((size_t)c->ne * c->npe);
    for (int64_t e = 0; e < c->ne; ++e)
        for (int k = 0; k < c->npe; ++k)
            conn[(size_t)c->elem_iperm[e] * c->npe + k] = c->node_iperm[c->conn_h[(size_t)e * c->npe + k]];
    exo::Mesh m;
    m.title = title ? title : "";
    m.n_nodes = c->nn;
    m.x = x.data();
    m.y = y.data();
    m.z = z.data();
    m.n_elem = c->ne;
    m.npe = c->npe;
    m.conn = conn.data();
    m.elem_type = c->npe == 4 ? "TETRA4" : "HEX8";
    std::vector<std::string> nv, ev;
    for (const char *f : {"displacement", "velocity", "acceleration", "force", "mass", "essential_boundary"})
        for (const char *d : {"_x", "_y", "_z"}) nv.push_back(std::string(f) + d);
    for (const char *f : {"displacement_gradient", "pk1_stress"})
        for (const char *t : kTensorSuffix) ev.push_back(std::string(f) + "_" + t);
    ev.push_back("volume");
    if (c->state_count() == c->ne) ev.push_back("equivalent_plastic_strain");
    c->results.reset(new exo::Writer(path, m, nv, ev));
    APHIP_API_END
}

int aphip_results_write(aphip_ctx *c, double time) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "NULL argument");
    APHIP_REQUIRE(c->results != nullptr, "no results file open");
    const int64_t nn = c->nn, ne = c->ne;
    std::vector<double> nodal(18 * nn, 0.0), elem(19 * ne + ne, 0.0), tmp(9 * std::max(nn, ne));
    const int32_t vec_fields[4] = {APHIP_FIELD_U, APHIP_FIELD_V, APHIP_FIELD_A, APHIP_FIELD_F};
    for (int q = 0; q < 4; ++q) {
        const int rc = aphip_get_field(c, vec_fields[q], tmp.data());
        if (rc != APHIP_OK) return rc;
        for (int64_t i = 0; i < nn; ++i)
            for (int d = 0; d < 3; ++d) nodal[(3 * q + d) * nn + i] = tmp[3 * i + d];
    }
    if (c->has_mass) {  // one scalar per node, written as the reference's three equal components
        const int rc = aphip_get_field(c, APHIP_FIELD_MASS, tmp.data());
        if (rc != APHIP_OK) return rc;
        for (int64_t i = 0; i < nn; ++i)
            for (int d = 0; d < 3; ++d) nodal[(12 + d) * nn + i] = tmp[i];
    }
    if (c->has_fixed) {
        std::vector<uint8_t> fx(nn);
        c->fixed.download(fx.data(), nn, c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < nn; ++i)
            for (int d = 0; d < 3; ++d) nodal[(15 + d) * nn + c->node_iperm[i]] = (fx[i] >> d) & 1u ? 1.0 : 0.0;
    }
    const int32_t ten_fields[2] = {APHIP_FIELD_H, APHIP_FIELD_P};
    for (int q = 0; q < 2; ++q) {
        const int rc = aphip_get_field(c, ten_fields[q], tmp.data());
        if (rc != APHIP_OK) return rc;
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < 9; ++k) elem[(9 * q + k) * ne + e] = tmp[9 * e + k];
    }
    {
        const int rc = aphip_get_field(c, APHIP_FIELD_VOLUME, tmp.data());
        if (rc != APHIP_OK) return rc;
        std::copy(tmp.begin(), tmp.begin() + ne, elem.begin() + 18 * ne);
    }
    const bool state = c->state_count() == ne;
    if (state) {
        const int rc = aphip_get_field(c, APHIP_FIELD_STATE, tmp.data());
        if (rc != APHIP_OK) return rc;
        std::copy(tmp.begin(), tmp.begin() + ne, elem.begin() + 19 * ne);
    }
    std::vector<const double *> nvals, evals;
    for (int q = 0; q < 18; ++q) nvals.push_back(nodal.data() + q * nn);
    for (int q = 0; q < (state ? 20 : 19); ++q) evals.push_back(elem.data() + q * ne);
    c->results->write_step(time, nvals, evals);
    APHIP_API_END
}

int aphip_results_close(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "NULL argument");
    c->results.reset();
    APHIP_API_END
}

}  // extern "C"
