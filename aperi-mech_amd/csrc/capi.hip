// C ABI of the aperi-hip library: context lifetime, mesh set-up (Morton
// renumbering, node-to-element CSR), material resolution, field transfer and
// the step driver.  Host C++ around the kernels in fem.hip and rk.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

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

// Stable ordering of points by 63-bit Morton code over their bounding box.
static std::vector<int32_t> morton_order(const std::vector<double> &pts, int64_t n) {
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
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return code[a] < code[b]; });
    return order;
}

static void resolve_material(int kind, const aphip_material_params *p, MaterialConst &m) {
    APHIP_REQUIRE(kind >= APHIP_LINEAR_ELASTIC && kind <= APHIP_SVK, "unknown material kind");
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
    m.state_per_qp = p->state_per_qp ? 1 : 0;
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
    c.state_np1 = 1;
}

static void require_ready(Ctx &c, bool need_mass) {
    APHIP_REQUIRE(c.etype >= 0, "mesh not set");
    APHIP_REQUIRE(c.mat.kind >= 0, "material not set");
    if (need_mass) APHIP_REQUIRE(c.has_mass, "mass not computed or set");
}

static void force_pass(Ctx &c) {
    prof_begin(c, 1);
    if (c.cells.ready) rk_cell_force(c);
    else fem_element_force(c);
    prof_end(c, 1);
}

static void node_pass(Ctx &c, double dt, int mode) {
    prof_begin(c, 2);
    if (c.cells.ready) rk_node_update(c, dt, mode);
    else fem_node_update(c, dt, mode);
    prof_end(c, 2);
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
    aphip_ctx *c = new aphip_ctx();
    c->device = dev;
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
    for (hipEvent_t e : c->prof.ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    APHIP_API_END
}

int aphip_set_mesh(aphip_ctx *c, int32_t elem_type, int64_t n_nodes, const double *X, int64_t n_elem,
                   const int32_t *conn) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(elem_type == APHIP_TET4 || elem_type == APHIP_HEX8, "element type must be TET4 or HEX8");
    APHIP_REQUIRE(n_nodes > 0 && n_nodes < (int64_t)INT32_MAX, "n_nodes out of range");
    APHIP_REQUIRE(n_elem >= 0, "n_elem out of range");
    APHIP_REQUIRE(X != nullptr && (n_elem == 0 || conn != nullptr), "mesh arrays are NULL");
    const int npe = elem_type == APHIP_TET4 ? 4 : 8;
    APHIP_REQUIRE(n_elem * npe < (int64_t)INT32_MAX, "n_elem too large for int32 incidences");
    for (int64_t i = 0; i < n_elem * npe; ++i)
        APHIP_REQUIRE(conn[i] >= 0 && conn[i] < n_nodes, "connectivity references a node out of range");
    APHIP_HIP(hipSetDevice(c->device));
    c->cells.reset();
    c->etype = elem_type;
    c->npe = npe;
    c->nn = n_nodes;
    c->ne = n_elem;

    // Morton renumbering: nodes by position, elements by centroid
    std::vector<double> xs(X, X + 3 * n_nodes);
    c->node_iperm = morton_order(xs, n_nodes);
    c->node_perm.assign(n_nodes, 0);
    for (int64_t i = 0; i < n_nodes; ++i) c->node_perm[c->node_iperm[i]] = (int32_t)i;
    std::vector<double> cent(3 * n_elem, 0.0);
    for (int64_t e = 0; e < n_elem; ++e)
        for (int k = 0; k < npe; ++k)
            for (int d = 0; d < 3; ++d) cent[3 * e + d] += X[3 * (int64_t)conn[e * npe + k] + d] / npe;
    c->elem_iperm = morton_order(cent, n_elem);
    c->elem_perm.assign(n_elem, 0);
    for (int64_t e = 0; e < n_elem; ++e) c->elem_perm[c->elem_iperm[e]] = (int32_t)e;

    std::vector<double> Xi(3 * n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i)
        for (int d = 0; d < 3; ++d) Xi[3 * i + d] = X[3 * (int64_t)c->node_iperm[i] + d];
    std::vector<int32_t> ci((size_t)n_elem * npe);
    for (int64_t e = 0; e < n_elem; ++e)
        for (int k = 0; k < npe; ++k) ci[e * npe + k] = c->node_perm[conn[(int64_t)c->elem_iperm[e] * npe + k]];

    // node -> (element, slot) incidences in ascending internal element order
    std::vector<int32_t> start(n_nodes + 1, 0), inc((size_t)n_elem * npe);
    for (int64_t t = 0; t < n_elem * npe; ++t) start[ci[t] + 1]++;
    for (int64_t i = 0; i < n_nodes; ++i) start[i + 1] += start[i];
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int64_t t = 0; t < n_elem * npe; ++t) inc[fill[ci[t]]++] = (int32_t)t;

    c->X.upload(Xi.data(), Xi.size(), c->stream);
    c->conn.upload(ci.data(), ci.size(), c->stream);
    c->node_start.upload(start.data(), start.size(), c->stream);
    c->node_inc.upload(inc.data(), inc.size(), c->stream);
    for (DevBuf<double> *b : {&c->u, &c->v, &c->a, &c->f}) {
        b->alloc(3 * n_nodes);
        b->zero(c->stream);
    }
    c->H.alloc(9 * std::max<int64_t>(n_elem, 1));
    c->P.alloc(9 * std::max<int64_t>(n_elem, 1));
    c->H.zero(c->stream);
    c->P.zero(c->stream);
    c->partial.alloc((size_t)std::max<int64_t>(n_elem, 1) * npe * 3);
    c->mass.release();
    c->has_mass = false;
    c->fixed.release();
    c->has_fixed = false;
    c->has_fields = false;
    c->elem_lambda.release();
    c->elem_mu.release();
    fem_volumes(*c);
    if (c->mat.kind >= 0) alloc_state(*c);
    APHIP_HIP(hipStreamSynchronize(c->stream));
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
    APHIP_HIP(hipSetDevice(c->device));
    c->mat = m;
    if (c->etype >= 0) alloc_state(*c);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_set_material_fields(aphip_ctx *c, const double *E, const double *nu) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
    if (!E && !nu) {
        c->has_fields = false;
        c->elem_lambda.release();
        c->elem_mu.release();
        return APHIP_OK;
    }
    APHIP_REQUIRE(E && nu, "E and nu must both be given");
    std::vector<double> lam(c->ne), mu(c->ne);
    for (int64_t e = 0; e < c->ne; ++e) {
        const int64_t src = c->elem_iperm[e];
        const double Ee = E[src], ne = nu[src];
        APHIP_REQUIRE(ne >= -1.0 && ne < 0.5 && Ee > 0.0, "per-element E/nu out of bounds");
        mu[e] = Ee / (2 * (1 + ne));
        lam[e] = (Ee * ne) / ((1 + ne) * (1 - 2 * ne));
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
    if (!mask) {
        c->fixed.release();
        c->has_fixed = false;
        return APHIP_OK;
    }
    std::vector<uint8_t> bits(c->nn, 0);
    for (int64_t i = 0; i < c->nn; ++i) {
        const int64_t src = c->node_iperm[i];
        for (int d = 0; d < 3; ++d)
            if (mask[3 * src + d]) bits[i] |= (uint8_t)(1u << d);
    }
    c->fixed.upload(bits.data(), bits.size(), c->stream);
    c->has_fixed = true;
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_set_field(aphip_ctx *c, int32_t field, const double *data) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr && data != nullptr, "NULL argument");
    APHIP_REQUIRE(c->etype >= 0, "mesh not set");
    APHIP_HIP(hipSetDevice(c->device));
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
        std::vector<double> h(ns);
        for (int64_t e = 0; e < c->ne; ++e)
            for (int q = 0; q < per; ++q) h[e * per + q] = data[(int64_t)c->elem_iperm[e] * per + q];
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
        const DevBuf<double> &b = field == APHIP_FIELD_H ? c->H : c->P;
        std::vector<double> h(9 * c->ne);
        b.download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < c->ne; ++e)
            for (int k = 0; k < 9; ++k) out[9 * (int64_t)c->elem_iperm[e] + k] = h[k * c->ne + e];
    } else if (field == APHIP_FIELD_STATE) {
        const int64_t ns = c->state_count();
        APHIP_REQUIRE(ns > 0, "material has no state");
        const int64_t per = ns / c->ne;
        std::vector<double> h(ns);
        c->state[c->state_np1].download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < c->ne; ++e)
            for (int q = 0; q < per; ++q) out[(int64_t)c->elem_iperm[e] * per + q] = h[e * per + q];
    } else if (field == APHIP_FIELD_VOLUME) {
        const DevBuf<double> &b = c->cells.ready ? c->cells.volume : c->volume;
        std::vector<double> h(c->ne);
        b.download(h.data(), h.size(), c->stream);
        APHIP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < c->ne; ++e) out[c->elem_iperm[e]] = h[e];
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
    (void)dt;
    APHIP_HIP(hipSetDevice(c->device));
    force_pass(*c);
    node_pass(*c, 0.0, 0);
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}

int aphip_initialize(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, true);
    APHIP_HIP(hipSetDevice(c->device));
    force_pass(*c);
    node_pass(*c, 0.0, 1);
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
    for (int32_t s = 0; s < n_steps; ++s) {
        prof_begin(*c, 0);
        c->state_np1 = 1 - c->state_np1;  // UpdateFieldStates (Explicit.cpp:393)
        fem_predict(*c, dt);
        force_pass(*c);
        node_pass(*c, dt, 2);
        prof_end(*c, 0);
    }
    APHIP_API_END
}

int aphip_synchronize(aphip_ctx *c) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    APHIP_HIP(hipSetDevice(c->device));
    APHIP_HIP(hipStreamSynchronize(c->stream));
    APHIP_API_END
}


int aphip_rk_setup(aphip_ctx *c, double scale) {
    APHIP_API_BEGIN
    APHIP_REQUIRE(c != nullptr, "ctx is NULL");
    require_ready(*c, false);
    APHIP_REQUIRE(c->etype == APHIP_TET4, "RKPM element-as-cell smoothing needs a Tet4 background mesh");
    APHIP_REQUIRE(scale >= 0.0, "Kernel radius scale factor must be non-negative.");
    APHIP_HIP(hipSetDevice(c->device));
    rk_setup(*c, scale);
    c->has_mass = false;
    APHIP_HIP(hipStreamSynchronize(c->stream));
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

}  // extern "C"
