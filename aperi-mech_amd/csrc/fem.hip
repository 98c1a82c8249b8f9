// Gaussian-quadrature element force (Tet4 1-point, Hex8 2x2x2), deterministic
// node assembly and the fused central-difference node updates, for gfx950.
//
// Reference path replaced (aperi-mech):
//   ComputeInternalForceGaussian::operator()      include/ComputeInternalForceGaussian.h:57-131
//   Quadrature::ComputeBMatrixAndWeight            include/QuadratureGaussian.h:17-38
//   ShapeFunctionsFunctorTet4/Hex8::Derivatives    include/ShapeFunctionsFunctorTet4.h:27-35,
//                                                  include/ShapeFunctionsFunctorHex8.h:51-62
//   Field::AtomicAdd force scatter                 include/Field.h:235-246  (replaced by a
//                                                  node-to-element CSR gather in fixed order)
//   ExplicitTimeIntegrator node functors           include/ExplicitTimeIntegrator.h:116-126,
//                                                  :267-276, :308-318, :348-357
//
// Data layout (internal Morton numbering):
//   node vectors  AoS  x,y,z per node (24 B), gathered per element
//   connectivity  AoS  npe int32 per element (one 16 B / 32 B load)
//   H, P          SoA  9 planes of n_elem doubles (coalesced stores)
//   partials      AoS  npe*3 doubles per element, read back by the node gather
#include "element.hpp"
#include "internal.hpp"

namespace aphip {

namespace {

constexpr int kElemBlock = 256;
constexpr int kNodeBlock = 256;
constexpr int kNumXcd = 8;

// Consecutive logical blocks go to the same XCD (the dispatcher places block b on
// XCD b % 8), so neighbouring elements share that XCD's L2 for their node gathers.
// The grid is padded to a multiple of 8 so the map is a bijection.
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks_padded) {
    const int64_t b = blockIdx.x;
    const int64_t per = nblocks_padded / kNumXcd;
    return (b % kNumXcd) * per + b / kNumXcd;
}

inline int64_t padded_blocks(int64_t n, int block) {
    int64_t nb = (n + block - 1) / block;
    return ((nb + kNumXcd - 1) / kNumXcd) * kNumXcd;
}

struct ElemArgs {
    const double *X;
    const double *u;
    const int32_t *conn;
    int64_t ne;
    int64_t nblocks;
    double *H;
    double *P;
    const double *Hn, *Pn;  // rate-form materials: H and P of the previous step
    double rate_dt;
    const double *state_n;
    double *state_np1;
    double *partial;
    MatParams mat;
    const double *elem_lambda;
    const double *elem_mu;
    double gauss_a;  // 1/sqrt(3), computed on the host as the reference does
};

__device__ __forceinline__ void load3(const double *p, int64_t i, double *o) {
    o[0] = p[3 * i];
    o[1] = p[3 * i + 1];
    o[2] = p[3 * i + 2];
}

// UL: updated Lagrangian (A.X the start-of-step coordinates, A.u the increments)
template <int KIND, bool UL>
__global__ __launch_bounds__(kElemBlock) void k_tet4_force(ElemArgs A) {
    const int64_t e = xcd_block(A.nblocks) * kElemBlock + threadIdx.x;
    if (e >= A.ne) return;
    const int4 c = reinterpret_cast<const int4 *>(A.conn)[e];
    const int cn[4] = {c.x, c.y, c.z, c.w};
    double X[4][3], U[4][3], fe[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        load3(A.X, cn[k], X[k]);
        load3(A.u, cn[k], U[k]);
    }
    MatParams m = A.mat;
    if (A.elem_lambda) {
        m.lambda = A.elem_lambda[e];
        m.mu = A.elem_mu[e];
    }
    double snp1 = 0.0, aux = 0.0, H[9], P[9];
    const double sn = load_state<KIND>(A.state_n, e), sn2 = load_state2<KIND>(A.state_n, e);
    RateIn rin;
    load_rate<KIND, UL>(rin, A.Hn, A.Pn, A.ne, e, A.rate_dt);
    tet4_element<KIND, UL>(X, U, m, sn, snp1, aux, H, P, fe, &rin, sn2);
    store_state<KIND>(A.state_np1, e, snp1, aux);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        store_once(A.H + i * A.ne + e, H[i]);
        store_once(A.P + i * A.ne + e, P[i]);
    }
    double *fp = A.partial + e * 12;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) fp[3 * k + j] = fe[k][j];
}

template <int KIND, bool PER_QP, bool UL>
__global__ __launch_bounds__(128) void k_hex8_force(ElemArgs A) {
    const int64_t e = xcd_block(A.nblocks) * 128 + threadIdx.x;
    if (e >= A.ne) return;
    const int4 c0 = reinterpret_cast<const int4 *>(A.conn)[2 * e];
    const int4 c1 = reinterpret_cast<const int4 *>(A.conn)[2 * e + 1];
    const int cn[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    double X[8][3], U[8][3], fe[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        load3(A.X, cn[k], X[k]);
        load3(A.u, cn[k], U[k]);
    }
    MatParams m = A.mat;
    if (A.elem_lambda) {
        m.lambda = A.elem_lambda[e];
        m.mu = A.elem_mu[e];
    }
    const int64_t so = state_width<KIND>() * (PER_QP ? e * 8 : e);
    double H[9], P[9];
    RateIn rin;
    load_rate<KIND, UL>(rin, A.Hn, A.Pn, A.ne, e, A.rate_dt);
    hex8_element<KIND, PER_QP, UL>(X, U, m, A.gauss_a, state_width<KIND>() ? A.state_n + so : nullptr,
                                   state_width<KIND>() ? A.state_np1 + so : nullptr, H, P, fe, &rin);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        store_once(A.H + i * A.ne + e, H[i]);
        store_once(A.P + i * A.ne + e, P[i]);
    }
    double *fp = A.partial + e * 24;
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j) fp[3 * k + j] = fe[k][j];
}

struct NodeArgs {
    int64_t nn;
    const int32_t *node_start;
    const int32_t *node_inc;
    const double *partial;
    const double *mass;
    const uint8_t *fixed;
    double *f;
    double *a;
    double *v;
    double *u;
    const double *node_dt;
    double dt;
    // updated Lagrangian: the step's increments and the current coordinates at n, n+1;
    // semi-Lagrangian (semi = 1): the increments accumulate, no coordinates are moved
    double *dinc;
    const double *xn;
    double *xnp1;
    int semi;
};

// v1/2 = v + dt/2 a; velocity BC; u += dt v1/2  (ExplicitTimeIntegrator.h:116-126, :348-357)
__global__ __launch_bounds__(kNodeBlock) void k_predict(NodeArgs A) {
    const int64_t i = (int64_t)blockIdx.x * kNodeBlock + threadIdx.x;
    if (i >= A.nn) return;
    double u[3], v[3], a[3], vh[3], u1[3];
    load3(A.u, i, u);
    load3(A.v, i, v);
    load3(A.a, i, a);
    cd_predict(u, v, a, A.fixed ? A.fixed[i] : 0u, A.node_dt ? A.node_dt[i] : A.dt, vh, u1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        A.v[3 * i + d] = vh[d];
        A.u[3 * i + d] = u1[d];
    }
}

// Updated Lagrangian predictor (ExplicitTimeIntegrator.h:193-226): v1/2 as above, the
// increment dt v1/2, u += increment, x_{n+1} = x_n + increment; semi-Lagrangian: the
// increment accumulates since the last reference update and no coordinates move.
__global__ __launch_bounds__(kNodeBlock) void k_predict_ul(NodeArgs A) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kNodeBlock + threadIdx.x;
    if (i >= A.nn) return;
    double u[3], v[3], a[3], vh[3], u1[3], x[3], di[3];
    load3(A.u, i, u);
    load3(A.v, i, v);
    load3(A.a, i, a);
    load3(A.xn, i, x);
    load3(A.dinc, i, di);
    cd_predict(u, v, a, A.fixed ? A.fixed[i] : 0u, A.dt, vh, u1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double inc = A.dt * vh[d];
        A.v[3 * i + d] = vh[d];
        A.u[3 * i + d] = u[d] + inc;
        if (A.semi) {
            A.dinc[3 * i + d] = di[d] + inc;
        } else {
            A.dinc[3 * i + d] = inc;
            A.xnp1[3 * i + d] = x[d] + inc;
        }
    }
}

// Semi-Lagrangian reference update (ExplicitTimeIntegratorSemi::UpdateReferenceConfiguration,
// ExplicitTimeIntegrator.h:508-526): x_ref += increment, increment = 0 (nodes), and
// H_ref = H (elements; i indexes 9 x ne doubles).
__global__ __launch_bounds__(kNodeBlock) void k_semi_update(int64_t nn3, double *xref, double *dinc, int64_t nh,
                                                            double *href, const double *H) {
    const int64_t i = (int64_t)blockIdx.x * kNodeBlock + threadIdx.x;
    if (i < nn3) {
        xref[i] = xref[i] + dinc[i];
        dinc[i] = 0.0;
    }
    if (i < nh) href[i] = H[i];
}

// f = sum of the node's element partials in ascending element order, then per mode:
// 1: a = f/m; 2: a = f/m, acceleration BC, v += dt/2 a (ExplicitTimeIntegrator.h:267-276, :308-318)
template <int MODE>
__global__ __launch_bounds__(kNodeBlock) void k_node(NodeArgs A) {
    const int64_t i = (int64_t)blockIdx.x * kNodeBlock + threadIdx.x;
    if (i >= A.nn) return;
    const int32_t b = A.node_start[i], eidx = A.node_start[i + 1];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int32_t t = b; t < eidx; ++t) {
        const int64_t s = A.node_inc[t];
        const double *p = A.partial + 3 * s;
        f0 += p[0];
        f1 += p[1];
        f2 += p[2];
    }
    A.f[3 * i] = f0;
    A.f[3 * i + 1] = f1;
    A.f[3 * i + 2] = f2;
    if constexpr (MODE == 1) {
        const double m = A.mass[i];
        A.a[3 * i] = f0 / m;
        A.a[3 * i + 1] = f1 / m;
        A.a[3 * i + 2] = f2 / m;
    } else if constexpr (MODE == 2) {
        const double f[3] = {f0, f1, f2};
        double vh[3], a1[3], v1[3];
        load3(A.v, i, vh);
        cd_correct(f, A.mass[i], A.fixed ? A.fixed[i] : 0u, A.node_dt ? A.node_dt[i] : A.dt, vh, a1, v1);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            A.a[3 * i + d] = a1[d];
            A.v[3 * i + d] = v1[d];
        }
    }
}

struct VolArgs {
    const double *X;
    const int32_t *conn;
    int64_t ne;
    int npe;
    double gauss_a;
    double *vol;
};

__global__ __launch_bounds__(128) void k_volume(VolArgs A) {
    const int64_t e = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (e >= A.ne) return;
    double X[8][3];
    for (int k = 0; k < A.npe; ++k) load3(A.X, A.conn[e * A.npe + k], X[k]);
    double v = 0.0;
    if (A.npe == 4) {
        double J[9], Ji[9];
        for (int i = 0; i < 3; ++i) {
            J[3 * i + 0] = X[1][i] - X[0][i];
            J[3 * i + 1] = X[2][i] - X[0][i];
            J[3 * i + 2] = X[3][i] - X[0][i];
        }
        v = (1.0 / 6.0) * inverse3(J, Ji);
    } else {
        for (int q = 0; q < 8; ++q) {
            const int sq0 = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : -1;
            const int sq1 = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : -1;
            const int sq2 = (q >= 4) ? 1 : -1;
            const double p0 = sq0 * A.gauss_a, p1 = sq1 * A.gauss_a, p2 = sq2 * A.gauss_a;
            double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Ji[9];
            for (int k = 0; k < 8; ++k) {
                const int s0 = (k == 1 || k == 2 || k == 5 || k == 6) ? 1 : -1;
                const int s1 = (k == 2 || k == 3 || k == 6 || k == 7) ? 1 : -1;
                const int s2 = (k >= 4) ? 1 : -1;
                const double f0 = s0 < 0 ? (1.0 - p0) : (1.0 + p0);
                const double f1 = s1 < 0 ? (1.0 - p1) : (1.0 + p1);
                const double f2 = s2 < 0 ? (1.0 - p2) : (1.0 + p2);
                const double d0 = (s0 * 0.125) * f1 * f2, d1 = (s1 * 0.125) * f0 * f2, d2 = (s2 * 0.125) * f0 * f1;
                for (int i = 0; i < 3; ++i) {
                    J[3 * i] += X[k][i] * d0;
                    J[3 * i + 1] += X[k][i] * d1;
                    J[3 * i + 2] += X[k][i] * d2;
                }
            }
            v += inverse3(J, Ji);
        }
    }
    A.vol[e] = v;
}

// m_I = sum over incident elements (ascending) of rho V_e / npe  (src/MassUtils.cpp:26-39)
__global__ __launch_bounds__(kNodeBlock) void k_lumped_mass(int64_t nn, const int32_t *node_start,
                                                            const int32_t *node_inc, int npe, const double *vol,
                                                            double rho, const double *elem_rho, double *mass) {
    const int64_t i = (int64_t)blockIdx.x * kNodeBlock + threadIdx.x;
    if (i >= nn) return;
    double m = 0.0;
    for (int32_t t = node_start[i]; t < node_start[i + 1]; ++t) {
        const int64_t e = node_inc[t] / npe;
        m += (vol[e] * (elem_rho ? elem_rho[e] : rho)) / npe;
    }
    mass[i] = m;
}

template <int KIND, bool UL>
void launch_elem_kind_ul(Ctx &c, const ElemArgs &A) {
    if (c.etype == APHIP_TET4) {
        hipLaunchKernelGGL((k_tet4_force<KIND, UL>), dim3((unsigned)A.nblocks), dim3(kElemBlock), 0, c.stream, A);
    } else if (state_width<KIND>() && c.mat.state_per_qp) {
        hipLaunchKernelGGL((k_hex8_force<KIND, true, UL>), dim3((unsigned)A.nblocks), dim3(128), 0, c.stream, A);
    } else {
        hipLaunchKernelGGL((k_hex8_force<KIND, false, UL>), dim3((unsigned)A.nblocks), dim3(128), 0, c.stream, A);
    }
}

template <int KIND>
void launch_elem_kind(Ctx &c, const ElemArgs &A) {
    if (c.formulation != APHIP_TOTAL_LAGRANGIAN) launch_elem_kind_ul<KIND, true>(c, A);
    else launch_elem_kind_ul<KIND, false>(c, A);
}

}  // namespace

void fem_element_force(Ctx &c) {
    const bool ul = c.formulation != APHIP_TOTAL_LAGRANGIAN;
    ElemArgs A;
    A.X = ul ? c.xcn().p : c.X.p;
    A.u = ul ? c.dinc.p : c.u.p;
    A.conn = c.conn.p;
    A.ne = c.ne;
    A.nblocks = padded_blocks(c.ne, c.etype == APHIP_TET4 ? kElemBlock : 128);
    A.H = c.Hnp1().p;
    A.P = c.Pnp1().p;
    A.Hn = c.formulation == APHIP_SEMI_LAGRANGIAN ? c.Href.p : c.h_pair() ? c.Hn().p : nullptr;
    A.Pn = c.mat.rate_form() ? c.Pn().p : nullptr;
    A.rate_dt = c.rate_dt;
    A.state_n = c.state[1 - c.state_np1].p;
    A.state_np1 = c.state[c.state_np1].p;
    A.partial = c.partial.p;
    A.mat = c.mat.params();
    A.elem_lambda = c.has_fields ? c.elem_lambda.p : nullptr;
    A.elem_mu = c.has_fields ? c.elem_mu.p : nullptr;
    A.gauss_a = 1.0 / std::sqrt(3.0);
    if (c.ne == 0) return;
    switch (c.mat.kind) {
        case APHIP_LINEAR_ELASTIC: launch_elem_kind<APHIP_LINEAR_ELASTIC>(c, A); break;
        case APHIP_NEO_HOOKEAN: launch_elem_kind<APHIP_NEO_HOOKEAN>(c, A); break;
        case APHIP_J2: launch_elem_kind<APHIP_J2>(c, A); break;
        case APHIP_SVK: launch_elem_kind<APHIP_SVK>(c, A); break;
        case APHIP_NEO_HOOKEAN_DAMAGE: launch_elem_kind<APHIP_NEO_HOOKEAN_DAMAGE>(c, A); break;
        case APHIP_POWER_LAW_CREEP: launch_elem_kind<APHIP_POWER_LAW_CREEP>(c, A); break;
        case APHIP_DRUCKER_PRAGER: launch_elem_kind<APHIP_DRUCKER_PRAGER>(c, A); break;
        default: throw Error(APHIP_ERR_INVALID, "material not set");
    }
    APHIP_HIP(hipGetLastError());
}

static NodeArgs node_args(Ctx &c, double dt) {
    NodeArgs A;
    A.nn = c.nn;
    A.node_start = c.node_start.p;
    A.node_inc = c.node_inc.p;
    A.partial = c.partial.p;
    A.mass = c.mass.p;
    A.fixed = c.has_fixed ? c.fixed.p : nullptr;
    A.f = c.f.p;
    A.a = c.a.p;
    A.v = c.v.p;
    A.u = c.u.p;
    A.node_dt = c.node_dt.p;
    A.dt = dt;
    A.dinc = c.dinc.p;
    A.xn = c.formulation != APHIP_TOTAL_LAGRANGIAN ? c.xcn().p : nullptr;
    A.xnp1 = c.formulation == APHIP_UPDATED_LAGRANGIAN ? c.xcnp1().p : nullptr;
    A.semi = c.formulation == APHIP_SEMI_LAGRANGIAN ? 1 : 0;
    return A;
}

void fem_predict(Ctx &c, double dt) {
    if (c.n_dof == 0) return;
    NodeArgs A = node_args(c, dt);
    A.nn = c.n_dof;  // nodes without degrees of freedom (last in the numbering) keep their values
    const dim3 grid((unsigned)((c.n_dof + kNodeBlock - 1) / kNodeBlock));
    if (c.formulation != APHIP_TOTAL_LAGRANGIAN) hipLaunchKernelGGL(k_predict_ul, grid, dim3(kNodeBlock), 0, c.stream, A);
    else hipLaunchKernelGGL(k_predict, grid, dim3(kNodeBlock), 0, c.stream, A);
    APHIP_HIP(hipGetLastError());
}

void fem_semi_update(Ctx &c) {
    const int64_t n3 = 3 * c.nn, nh = 9 * c.ne, n = std::max(n3, nh);
    if (n == 0) return;
    hipLaunchKernelGGL(k_semi_update, dim3((unsigned)((n + kNodeBlock - 1) / kNodeBlock)), dim3(kNodeBlock), 0,
                       c.stream, n3, c.xc[0].p, c.dinc.p, nh, c.Href.p, c.Hnp1().p);
    APHIP_HIP(hipGetLastError());
}

void fem_node_update(Ctx &c, double dt, int mode) {
    if (c.nn == 0) return;
    NodeArgs A = node_args(c, dt);
    const dim3 grid((unsigned)((c.nn + kNodeBlock - 1) / kNodeBlock));
    if (mode == 0) hipLaunchKernelGGL(k_node<0>, grid, dim3(kNodeBlock), 0, c.stream, A);
    else if (mode == 1) hipLaunchKernelGGL(k_node<1>, grid, dim3(kNodeBlock), 0, c.stream, A);
    else hipLaunchKernelGGL(k_node<2>, grid, dim3(kNodeBlock), 0, c.stream, A);
    APHIP_HIP(hipGetLastError());
}

void fem_volumes(Ctx &c) {
    c.volume.alloc(c.ne);
    if (c.ne == 0) return;
    VolArgs A{c.X.p, c.conn.p, c.ne, c.npe, 1.0 / std::sqrt(3.0), c.volume.p};
    hipLaunchKernelGGL(k_volume, dim3((unsigned)((c.ne + 127) / 128)), dim3(128), 0, c.stream, A);
    APHIP_HIP(hipGetLastError());
}

void fem_lumped_mass(Ctx &c, double rho) {
    c.mass.alloc(c.nn);
    if (c.nn == 0) return;
    hipLaunchKernelGGL(k_lumped_mass, dim3((unsigned)((c.nn + kNodeBlock - 1) / kNodeBlock)), dim3(kNodeBlock), 0,
                       c.stream, c.nn, c.node_start.p, c.node_inc.p, c.npe, c.volume.p, rho, c.elem_rho.p, c.mass.p);
    APHIP_HIP(hipGetLastError());
}

}  // namespace aphip
