// Discrete adjoint of the explicit step for gfx950 (an extension: the reference
// has no differentiation).  Hand-written vector-Jacobian products:
//   stress      H_bar = (dP/dH)^T : P_bar and (lambda_bar, mu_bar) for linear elastic
//               (include/Materials/LinearElastic.h:70-88) and neo-Hookean
//               (include/Materials/NeoHooke.h:67-92):
//                 P = tau F^-T, tau = lambda ln J I + mu (F F^T - I):
//                 T = P_bar F^-1,
//                 H_bar = lambda tr(T) F^-T + mu (T + T^T) F - F^-T P_bar^T tau F^-T,
//                 lambda_bar = ln J tr(T),  mu_bar = T : (F F^T - I)
//   force       f_k = -w B_k P^T  ->  P_bar = -w sum_k f_bar_k (x) B_k,  u_bar_l = H_bar B_l^T
//   time step   reverse of  v1/2 = mask(v + h a), u1 = u + dt v1/2, a1 = mask(f(u1)/m),
//               v1 = v1/2 + h a1 (h = dt/2):
//                 f_bar = mask(a_bar1 + h v_bar1) / m,  u_bar = u_bar1 + K^T f_bar,
//                 v_bar1/2 = mask(v_bar1 + dt u_bar),  v_bar = v_bar1/2,  a_bar = h v_bar1/2
// The node assembly of u_bar reuses the forward tile plan (same deterministic
// gather); the forward trajectory is recomputed segment by segment from
// checkpoints (checkpoint-and-recompute).
#include <cmath>
#include <vector>

#include "element.hpp"
#include "internal.hpp"

namespace aphip {

namespace {

constexpr int kNumXcd = 8;

__device__ __forceinline__ int64_t xcd_tile(int64_t nblocks_padded) {
    const int64_t b = blockIdx.x;
    const int64_t per = nblocks_padded / kNumXcd;
    return (b % kNumXcd) * per + b / kNumXcd;
}

__device__ __forceinline__ void ld3(const double *p, int64_t i, double *o) {
    o[0] = p[3 * i];
    o[1] = p[3 * i + 1];
    o[2] = p[3 * i + 2];
}

__device__ __forceinline__ void st3(double *p, int64_t i, const double *o) {
    p[3 * i] = o[0];
    p[3 * i + 1] = o[1];
    p[3 * i + 2] = o[2];
}

// H_bar and parameter adjoints of P(H; lambda, mu) for a given P_bar.
template <int KIND>
__device__ __forceinline__ void stress_vjp(const double *H, const double *Pb, const MatParams &m, double *Hb,
                                           double &lb, double &mb) {
    if constexpr (KIND == APHIP_LINEAR_ELASTIC) {
        const double trP = Pb[0] + Pb[4] + Pb[8];
        const double trE = H[0] + H[4] + H[8];
        double se = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double sp = 0.5 * (Pb[3 * i + j] + Pb[3 * j + i]);
                Hb[3 * i + j] = 2.0 * m.mu * sp;
                se += Pb[3 * i + j] * 0.5 * (H[3 * i + j] + H[3 * j + i]);
            }
#pragma unroll
        for (int i = 0; i < 3; ++i) Hb[4 * i] += m.lambda * trP;
        lb = trP * trE;
        mb = 2.0 * se;
    } else {  // neo-Hookean
        double F[9], Fi[9], tau[9], BmI[9], T[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) F[i] = H[i];
        F[0] += 1.0;
        F[4] += 1.0;
        F[8] += 1.0;
        inverse3(F, Fi);
        const double lnJ = log1p(det_ap_im1(H));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                BmI[3 * i + j] = H[3 * i] * H[3 * j] + H[3 * i + 1] * H[3 * j + 1] + H[3 * i + 2] * H[3 * j + 2] +
                                 H[3 * j + i] + H[3 * i + j];
                tau[3 * i + j] = m.mu * BmI[3 * i + j];
            }
#pragma unroll
        for (int i = 0; i < 3; ++i) tau[4 * i] += m.lambda * lnJ;
        // T = P_bar F^-1
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                T[3 * i + j] = Pb[3 * i] * Fi[j] + Pb[3 * i + 1] * Fi[3 + j] + Pb[3 * i + 2] * Fi[6 + j];
        const double trT = T[0] + T[4] + T[8];
        // W = P_bar^T tau ; Z = F^-T W F^-T
        double W[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                W[3 * i + j] = Pb[i] * tau[j] + Pb[3 + i] * tau[3 + j] + Pb[6 + i] * tau[6 + j];
        double FW[9];  // F^-T W: (F^-T)_ik = Fi_ki
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) FW[3 * i + j] = Fi[i] * W[j] + Fi[3 + i] * W[3 + j] + Fi[6 + i] * W[6 + j];
        double sm = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                // (T + T^T) F
                double tf = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) tf += (T[3 * i + k] + T[3 * k + i]) * F[3 * k + j];
                // (F^-T W) F^-T: sum_k FW_ik Fi_jk
                const double z = FW[3 * i] * Fi[3 * j] + FW[3 * i + 1] * Fi[3 * j + 1] + FW[3 * i + 2] * Fi[3 * j + 2];
                Hb[3 * i + j] = m.lambda * trT * Fi[3 * j + i] + m.mu * tf - z;
                sm += T[3 * i + j] * BmI[3 * i + j];
            }
        lb = lnJ * trT;
        mb = sm;
    }
}

template <int KIND>
__device__ void tet4_vjp(const double (&X)[4][3], const double (&U)[4][3], const double (&Fb)[4][3],
                         const MatParams &m, double (&ub)[4][3], double &lb, double &mb) {
    double J[9], Ji[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        J[3 * i + 0] = X[1][i] - X[0][i];
        J[3 * i + 1] = X[2][i] - X[0][i];
        J[3 * i + 2] = X[3][i] - X[0][i];
    }
    const double w = (1.0 / 6.0) * inverse3(J, Ji);
    double B[4][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        B[1][j] = Ji[j];
        B[2][j] = Ji[3 + j];
        B[3][j] = Ji[6 + j];
        B[0][j] = ((-Ji[j]) - Ji[3 + j]) - Ji[6 + j];
    }
    double H[9], Pb[9], Hb[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            H[3 * i + j] = U[0][i] * B[0][j] + U[1][i] * B[1][j] + U[2][i] * B[2][j] + U[3][i] * B[3][j];
            Pb[3 * i + j] = -w * (Fb[0][i] * B[0][j] + Fb[1][i] * B[1][j] + Fb[2][i] * B[2][j] + Fb[3][i] * B[3][j]);
        }
    stress_vjp<KIND>(H, Pb, m, Hb, lb, mb);
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int a = 0; a < 3; ++a) ub[l][a] = Hb[3 * a] * B[l][0] + Hb[3 * a + 1] * B[l][1] + Hb[3 * a + 2] * B[l][2];
}

template <int KIND>
__device__ void hex8_vjp(const double (&X)[8][3], const double (&U)[8][3], const double (&Fb)[8][3],
                         const MatParams &m, double ga, double (&ub)[8][3], double &lb, double &mb) {
#pragma unroll
    for (int k = 0; k < 8; ++k) ub[k][0] = ub[k][1] = ub[k][2] = 0.0;
    lb = mb = 0.0;
    // Gauss points one at a time: unrolled, the eight need over 500 registers per lane
    // (one wave per SIMD, with spills)
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
        int sq0, sq1, sq2;
        hex8_sign(q, sq0, sq1, sq2);
        const double p0 = sq0 * ga, p1 = sq1 * ga, p2 = sq2 * ga;
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, FG[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double dN[8][3];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int s0, s1, s2;
            hex8_sign(k, s0, s1, s2);
            const double f0 = s0 < 0 ? (1.0 - p0) : (1.0 + p0);
            const double f1 = s1 < 0 ? (1.0 - p1) : (1.0 + p1);
            const double f2 = s2 < 0 ? (1.0 - p2) : (1.0 + p2);
            dN[k][0] = (s0 * 0.125) * f1 * f2;
            dN[k][1] = (s1 * 0.125) * f0 * f2;
            dN[k][2] = (s2 * 0.125) * f0 * f1;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    J[3 * i + d] += X[k][i] * dN[k][d];
                    G[3 * i + d] += U[k][i] * dN[k][d];
                    FG[3 * i + d] += Fb[k][i] * dN[k][d];
                }
        }
        double Ji[9];
        const double w = inverse3(J, Ji);
        double H[9], Pb[9], Hb[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                H[3 * i + j] = G[3 * i] * Ji[j] + G[3 * i + 1] * Ji[3 + j] + G[3 * i + 2] * Ji[6 + j];
                Pb[3 * i + j] = -w * (FG[3 * i] * Ji[j] + FG[3 * i + 1] * Ji[3 + j] + FG[3 * i + 2] * Ji[6 + j]);
            }
        double lq, mq;
        stress_vjp<KIND>(H, Pb, m, Hb, lq, mq);
        lb += lq;
        mb += mq;
        // R = H_bar J^-T ; u_bar_l += R dN_l
        double R[9];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int n = 0; n < 3; ++n)
                R[3 * a + n] = Hb[3 * a] * Ji[3 * n] + Hb[3 * a + 1] * Ji[3 * n + 1] + Hb[3 * a + 2] * Ji[3 * n + 2];
#pragma unroll
        for (int l = 0; l < 8; ++l)
#pragma unroll
            for (int a = 0; a < 3; ++a) ub[l][a] += R[3 * a] * dN[l][0] + R[3 * a + 1] * dN[l][1] + R[3 * a + 2] * dN[l][2];
    }
}

struct AdjArgs {
    const double *X;
    const double *u1;      // displacement at which the force was evaluated
    const double *fbar;    // MODE 0: given force adjoint per node
    double *ub, *vb, *ab;  // adjoint node state
    const double *mass;
    const uint8_t *fixed;
    const double *node_dt;
    const int32_t *tn_start, *tile_e0, *tn_ninterior, *int_base, *tn_node;
    const uint32_t *tn_csr;
    const int32_t *tn_bdst;
    const uint16_t *lcsr, *lconn;
    const double *zero;  // 8 zero doubles: stands in for an absent optional array
    double *bpart;
    double *lam_bar, *mu_bar;
    MatParams mat;
    const double *elem_lambda, *elem_mu;
    double gauss_a;
    int64_t ne, ntiles, nblocks;
    int te, maxtn;
    double dt;
};

// MODE 0: u_bar = K^T f_bar (f_bar given);  MODE 1: one reverse central-difference step;
// MODE 2: f_bar = a_bar / m (initial acceleration), parameter adjoints only.
template <int NPE, int KIND, int MODE>
__global__ __launch_bounds__(NPE == 4 ? 512 : 128) void k_tile_adj(AdjArgs A) {
    extern __shared__ __align__(16) char lds_raw[];
    const int64_t t = xcd_tile(A.nblocks);
    if (t >= A.ntiles) return;
    double *sX = reinterpret_cast<double *>(lds_raw);
    double *sU = sX + 3 * A.maxtn;
    double *sB = sU + 3 * A.maxtn;  // f_bar per tile node
    double *sF = sB + 3 * A.maxtn;  // u_bar partials per element slot
    uint32_t *sC = reinterpret_cast<uint32_t *>(sF + 3 * A.te * NPE);
    int32_t *sD = reinterpret_cast<int32_t *>(sC + A.maxtn);
    uint16_t *sE = reinterpret_cast<uint16_t *>(sD + A.maxtn);
    const int tid = threadIdx.x;
    const int tb = A.tn_start[t];
    const int tn = A.tn_start[t + 1] - tb;
    const int nint = A.tn_ninterior[t];
    const int64_t ibase = A.int_base[t];
    const int64_t te0 = A.tile_e0[t];
    const int nel = A.tile_e0[t + 1] - (int)te0;
    const int64_t e = te0 + tid;
    const bool has_elem = tid < nel;
    // Staging: the CSR words, then every node row's index and data, all issued before
    // the first LDS write (a load and its write per iteration would each wait a full
    // memory latency).  The kernel runs one element per thread (te == blockDim), so a
    // thread holds NPE / 2 CSR words; the first NPT rows of a thread load together.
    constexpr int EW = NPE / 2;
    constexpr int NPT = NPE == 4 ? 1 : 2;
    const int nw = A.te * NPE / 2;
    const uint32_t *ent_src = reinterpret_cast<const uint32_t *>(A.lcsr) + t * (int64_t)nw;
    uint32_t ew[EW];
#pragma unroll
    for (int r = 0; r < EW; ++r) ew[r] = ent_src[min(tid + r * (int)blockDim.x, nw - 1)];
    auto load_row = [&](int ln, double (&x)[3], double (&u)[3], double (&fb)[3]) {
        const int64_t g = A.tn_node[tb + ln];  // interior nodes are listed as ibase + ln
        ld3(A.X, g, x);
        ld3(A.u1, g, u);
        if constexpr (MODE == 0) {
            ld3(A.fbar, g, fb);
        } else if constexpr (MODE == 1) {
            double a[3], v[3];
            ld3(A.ab, g, a);
            ld3(A.vb, g, v);
            const double *pd = A.node_dt ? A.node_dt : A.zero;
            const uint8_t *pf = A.fixed ? A.fixed : reinterpret_cast<const uint8_t *>(A.zero);
            const double d8 = pd[A.node_dt ? g : 0];
            const unsigned f8 = pf[A.fixed ? g : 0];
            const double h = 0.5 * (A.node_dt ? d8 : A.dt);
            const unsigned fx = A.fixed ? f8 : 0u;
            const double m = A.mass[g];
#pragma unroll
            for (int d = 0; d < 3; ++d) fb[d] = (fx & (1u << d)) ? 0.0 : __fma_rn(h, v[d], a[d]) / m;
        } else {
            double a[3];
            ld3(A.ab, g, a);
            const double m = A.mass[g];
#pragma unroll
            for (int d = 0; d < 3; ++d) fb[d] = a[d] / m;
        }
    };
    double xq[NPT][3], uq[NPT][3], bq[NPT][3];
    uint32_t cq[NPT];
    int32_t dq[NPT];
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int ln = min(tid + q * (int)blockDim.x, tn - 1);  // clamped: unconditional loads
        cq[q] = A.tn_csr[tb + ln];
        dq[q] = A.tn_bdst[tb + ln];
        load_row(ln, xq[q], uq[q], bq[q]);
    }
    {
        uint32_t *dst = reinterpret_cast<uint32_t *>(sE);
#pragma unroll
        for (int r = 0; r < EW; ++r) {
            const int i = tid + r * (int)blockDim.x;
            if (i < nw) dst[i] = ew[r];
        }
    }
    auto stage = [&](int ln, uint32_t csr, int32_t bdst, const double (&x)[3], const double (&u)[3],
                     const double (&fb)[3]) {
        sC[ln] = csr;
        sD[ln] = bdst;  // -1 for interior nodes, never read for them
        st3(sX, ln, x);
        st3(sU, ln, u);
        st3(sB, ln, fb);
    };
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int ln = tid + q * (int)blockDim.x;
        if (ln < tn) stage(ln, cq[q], dq[q], xq[q], uq[q], bq[q]);
    }
    for (int ln = tid + NPT * (int)blockDim.x; ln < tn; ln += blockDim.x) {
        double x[3], u[3], fb[3];
        load_row(ln, x, u, fb);
        stage(ln, A.tn_csr[tb + ln], A.tn_bdst[tb + ln], x, u, fb);
    }
    __syncthreads();
    if (has_elem) {
        int lc[NPE];
        if constexpr (NPE == 4) {
            const ushort4 c = reinterpret_cast<const ushort4 *>(A.lconn)[e];
            lc[0] = c.x; lc[1] = c.y; lc[2] = c.z; lc[3] = c.w;
        } else {
            const ushort4 c0 = reinterpret_cast<const ushort4 *>(A.lconn)[2 * e];
            const ushort4 c1 = reinterpret_cast<const ushort4 *>(A.lconn)[2 * e + 1];
            lc[0] = c0.x; lc[1] = c0.y; lc[2] = c0.z; lc[3] = c0.w;
            lc[4] = c1.x; lc[5] = c1.y; lc[6] = c1.z; lc[7] = c1.w;
        }
        MatParams m = A.mat;
        if (A.elem_lambda) {
            m.lambda = A.elem_lambda[e];
            m.mu = A.elem_mu[e];
        }
        double X[NPE][3], U[NPE][3], Fb[NPE][3], ub[NPE][3];
#pragma unroll
        for (int k = 0; k < NPE; ++k) {
            ld3(sX, lc[k], X[k]);
            ld3(sU, lc[k], U[k]);
            ld3(sB, lc[k], Fb[k]);
        }
        double lb, mb;
        if constexpr (NPE == 4) tet4_vjp<KIND>(X, U, Fb, m, ub, lb, mb);
        else hex8_vjp<KIND>(X, U, Fb, m, A.gauss_a, ub, lb, mb);
        A.lam_bar[e] += lb;
        A.mu_bar[e] += mb;
#pragma unroll
        for (int k = 0; k < NPE; ++k) st3(sF, tid * NPE + k, ub[k]);
    }
    __syncthreads();
    if constexpr (MODE == 2) return;
    for (int ln = tid; ln < tn; ln += blockDim.x) {
        const uint32_t pk = sC[ln];
        const int st = (int)(pk >> 16), cnt = (int)(pk & 0xffffu);
        double s[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < cnt; ++j) {
            const int q = sE[st + j];
            s[0] += sF[3 * q];
            s[1] += sF[3 * q + 1];
            s[2] += sF[3 * q + 2];
        }
        if (ln < nint) {
            const int64_t g = ibase + ln;
            if constexpr (MODE == 0) {
                st3(A.ub, g, s);
            } else {
                double u[3], v[3], vh[3], a[3];
                ld3(A.ub, g, u);
                ld3(A.vb, g, v);
                const double dt = A.node_dt ? A.node_dt[g] : A.dt;
                const unsigned fx = A.fixed ? A.fixed[g] : 0u;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    u[d] += s[d];
                    vh[d] = (fx & (1u << d)) ? 0.0 : __fma_rn(dt, u[d], v[d]);
                    a[d] = (0.5 * dt) * vh[d];
                }
                st3(A.ub, g, u);
                st3(A.vb, g, vh);
                st3(A.ab, g, a);
            }
        } else {
            st3(A.bpart, sD[ln], s);
        }
    }
}

struct AdjBoundArgs {
    int64_t nb, first;
    const int32_t *bstart;
    const double *bpart;
    double *ub, *vb, *ab;
    const uint8_t *fixed;
    const double *node_dt;
    double dt;
};

template <int MODE>
__global__ __launch_bounds__(256) void k_boundary_adj(AdjBoundArgs A) {
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= A.nb) return;
    const int64_t g = A.first + b;
    double s[3] = {0.0, 0.0, 0.0};
    for (int32_t k = A.bstart[b]; k < A.bstart[b + 1]; ++k) {
        s[0] += A.bpart[3 * (int64_t)k];
        s[1] += A.bpart[3 * (int64_t)k + 1];
        s[2] += A.bpart[3 * (int64_t)k + 2];
    }
    if constexpr (MODE == 0) {
        st3(A.ub, g, s);
    } else {
        double u[3], v[3], vh[3], a[3];
        ld3(A.ub, g, u);
        ld3(A.vb, g, v);
        const double dt = A.node_dt ? A.node_dt[g] : A.dt;
        const unsigned fx = A.fixed ? A.fixed[g] : 0u;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            u[d] += s[d];
            vh[d] = (fx & (1u << d)) ? 0.0 : __fma_rn(dt, u[d], v[d]);
            a[d] = (0.5 * dt) * vh[d];
        }
        st3(A.ub, g, u);
        st3(A.vb, g, vh);
        st3(A.ab, g, a);
    }
}

inline int64_t padded(int64_t n) { return ((n + kNumXcd - 1) / kNumXcd) * kNumXcd; }

template <int NPE, int MODE>
void launch_adj(Ctx &c, const AdjArgs &A, size_t lds) {
    auto run = [&](auto kern) {
        if (lds > 65536)
            APHIP_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3((unsigned)A.nblocks), dim3(c.tiles.threads), lds, c.stream, A);
    };
    if (c.mat.kind == APHIP_LINEAR_ELASTIC) run(k_tile_adj<NPE, APHIP_LINEAR_ELASTIC, MODE>);
    else run(k_tile_adj<NPE, APHIP_NEO_HOOKEAN, MODE>);
}

// mode: 0 force VJP, 1 reverse step, 2 initial-acceleration parameter adjoint
void adj_run(Ctx &c, const double *u1, const double *fbar, int mode, double dt) {
    TilePlan &P = c.tiles;
    APHIP_REQUIRE(P.threads == P.te && (c.npe == 4 || P.te <= 128), "adjoint needs a one-element-per-thread tiling");
    AdjointData &D = c.adj;
    AdjArgs A;
    A.X = c.X.p;
    A.u1 = u1;
    A.fbar = fbar;
    A.ub = D.ub.p;
    A.vb = D.vb.p;
    A.ab = D.ab.p;
    A.mass = c.mass.p;
    A.fixed = c.has_fixed ? c.fixed.p : nullptr;
    A.node_dt = c.node_dt.p;
    A.tn_start = P.tn_start.p;
    A.tile_e0 = P.tile_e0.p;
    A.tn_ninterior = P.tn_ninterior.p;
    A.int_base = P.int_base.p;
    A.tn_node = P.tn_node.p;
    A.tn_csr = P.tn_csr.p;
    A.tn_bdst = P.tn_bdst.p;
    A.lcsr = P.lcsr.p;
    A.lconn = P.lconn.p;
    A.zero = P.zero.p;
    A.bpart = P.bpart.p;
    A.lam_bar = D.lam_bar.p;
    A.mu_bar = D.mu_bar.p;
    A.mat = c.mat.params();
    A.elem_lambda = c.has_fields ? c.elem_lambda.p : nullptr;
    A.elem_mu = c.has_fields ? c.elem_mu.p : nullptr;
    A.gauss_a = 1.0 / std::sqrt(3.0);
    A.ne = c.ne;
    A.ntiles = P.ntiles;
    A.nblocks = padded(P.ntiles);
    A.te = P.te;
    A.maxtn = P.maxtn;
    A.dt = dt;
    const size_t lds = sizeof(double) * (9 * (size_t)P.maxtn + 3 * (size_t)P.te * c.npe) + 8 * (size_t)P.maxtn +
                       2 * (size_t)P.te * c.npe + 16;
    if (c.npe == 4) {
        if (mode == 0) launch_adj<4, 0>(c, A, lds);
        else if (mode == 1) launch_adj<4, 1>(c, A, lds);
        else launch_adj<4, 2>(c, A, lds);
    } else {
        if (mode == 0) launch_adj<8, 0>(c, A, lds);
        else if (mode == 1) launch_adj<8, 1>(c, A, lds);
        else launch_adj<8, 2>(c, A, lds);
    }
    APHIP_HIP(hipGetLastError());
    if (mode != 2 && P.nb > 0) {
        AdjBoundArgs B{P.nb, P.n_interior, P.bstart.p, P.bpart.p, D.ub.p, D.vb.p, D.ab.p,
                       c.has_fixed ? c.fixed.p : nullptr, c.node_dt.p, dt};
        const dim3 grid((unsigned)((P.nb + 255) / 256));
        if (mode == 0) hipLaunchKernelGGL(k_boundary_adj<0>, grid, dim3(256), 0, c.stream, B);
        else hipLaunchKernelGGL(k_boundary_adj<1>, grid, dim3(256), 0, c.stream, B);
        APHIP_HIP(hipGetLastError());
    }
}

void require_adjoint_ready(Ctx &c) {
    APHIP_REQUIRE(c.etype >= 0 && c.mat.kind >= 0, "mesh and material must be set");
    APHIP_REQUIRE(c.mat.kind == APHIP_LINEAR_ELASTIC || c.mat.kind == APHIP_NEO_HOOKEAN,
                  "adjoints are provided for linear elastic and neo-Hookean materials");
    APHIP_REQUIRE(c.tiles.ready && !c.cells.ready, "adjoints run on the tiled Gaussian element path");
    // the adjoint kernels evaluate one element per thread, at most 512 Tet4 or 128 Hex8
    // per workgroup
    if (c.tiles.threads != c.tiles.te || c.tiles.te > (c.npe == 8 ? 128 : 512)) tile_build(c, 64, 64);
}

void alloc_adjoint(Ctx &c) {
    AdjointData &D = c.adj;
    for (DevBuf<double> *b : {&D.ub, &D.vb, &D.ab}) {
        b->alloc(3 * c.nn);
        b->zero(c.stream);
    }
    D.lam_bar.alloc(std::max<int64_t>(c.ne, 1));
    D.mu_bar.alloc(std::max<int64_t>(c.ne, 1));
    D.lam_bar.zero(c.stream);
    D.mu_bar.zero(c.stream);
}

// d(lambda, mu)/d(E, nu) chain rule per element, in caller element order
void parameter_gradients(Ctx &c, double *E_bar, double *nu_bar) {
    std::vector<double> lb(c.ne), mb(c.ne);
    c.adj.lam_bar.download(lb.data(), c.ne, c.stream);
    c.adj.mu_bar.download(mb.data(), c.ne, c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    for (int64_t e = 0; e < c.ne; ++e) {
        const double E = c.has_fields && !c.elem_E_h.empty() ? c.elem_E_h[e] : c.mat.E;
        const double nu = c.has_fields && !c.elem_nu_h.empty() ? c.elem_nu_h[e] : c.mat.nu;
        const double q = (1 + nu) * (1 - 2 * nu);
        const double dl_dE = nu / q, dm_dE = 1.0 / (2 * (1 + nu));
        const double dl_dnu = E * (1 + 2 * nu * nu) / (q * q), dm_dnu = -E / (2 * (1 + nu) * (1 + nu));
        const int64_t ce = c.elem_iperm[e];
        if (E_bar) E_bar[ce] = lb[e] * dl_dE + mb[e] * dm_dE;
        if (nu_bar) nu_bar[ce] = lb[e] * dl_dnu + mb[e] * dm_dnu;
    }
}

void copy_state(Ctx &c, double *dst) {
    const size_t n3 = 3 * (size_t)c.nn * sizeof(double);
    APHIP_HIP(hipMemcpyAsync(dst, c.u.p, n3, hipMemcpyDeviceToDevice, c.stream));
    APHIP_HIP(hipMemcpyAsync(dst + 3 * c.nn, c.v.p, n3, hipMemcpyDeviceToDevice, c.stream));
    APHIP_HIP(hipMemcpyAsync(dst + 6 * c.nn, c.a.p, n3, hipMemcpyDeviceToDevice, c.stream));
}

void restore_state(Ctx &c, const double *src) {
    const size_t n3 = 3 * (size_t)c.nn * sizeof(double);
    APHIP_HIP(hipMemcpyAsync(c.u.p, src, n3, hipMemcpyDeviceToDevice, c.stream));
    APHIP_HIP(hipMemcpyAsync(c.v.p, src + 3 * c.nn, n3, hipMemcpyDeviceToDevice, c.stream));
    APHIP_HIP(hipMemcpyAsync(c.a.p, src + 6 * c.nn, n3, hipMemcpyDeviceToDevice, c.stream));
}

}  // namespace

void force_vjp(Ctx &c, const double *f_bar, double *u_bar, double *E_bar, double *nu_bar) {
    require_adjoint_ready(c);
    alloc_adjoint(c);
    std::vector<double> h(3 * c.nn);
    for (int64_t i = 0; i < c.nn; ++i)
        for (int d = 0; d < 3; ++d) h[3 * i + d] = f_bar[3 * (int64_t)c.node_iperm[i] + d];
    DevBuf<double> fb;
    fb.upload(h.data(), h.size(), c.stream);
    adj_run(c, c.u.p, fb.p, 0, 0.0);
    c.adj.ub.download(h.data(), h.size(), c.stream);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    if (u_bar)
        for (int64_t i = 0; i < c.nn; ++i)
            for (int d = 0; d < 3; ++d) u_bar[3 * (int64_t)c.node_iperm[i] + d] = h[3 * i + d];
    parameter_gradients(c, E_bar, nu_bar);
}

void adjoint_forward(Ctx &c, double dt, int n_steps, int interval) {
    require_adjoint_ready(c);
    APHIP_REQUIRE(c.has_mass, "mass not computed or set");
    APHIP_REQUIRE(n_steps >= 1 && interval >= 1, "n_steps and checkpoint interval must be positive");
    AdjointData &D = c.adj;
    D.reset();
    const int nck = (n_steps + interval - 1) / interval;
    for (int k = 0; k < nck; ++k) {
        D.ckpt.push_back(new DevBuf<double>());
        D.ckpt.back()->alloc(9 * c.nn);
    }
    D.final_state.alloc(9 * c.nn);
    for (int s = 0; s < n_steps; ++s) {
        if (s % interval == 0) copy_state(c, D.ckpt[s / interval]->p);
        step_once_tiles(c, dt);
    }
    copy_state(c, D.final_state.p);
    APHIP_HIP(hipStreamSynchronize(c.stream));
    D.dt = dt;
    D.n_steps = n_steps;
    D.interval = interval;
    D.ready = true;
}

void adjoint_backward(Ctx &c, const double *ub_T, const double *vb_T, double *E_bar, double *nu_bar,
                      double *v0_bar) {
    require_adjoint_ready(c);
    AdjointData &D = c.adj;
    APHIP_REQUIRE(D.ready, "aphip_adjoint_forward has not run since the mesh was set or its nodes renumbered");
    APHIP_REQUIRE(ub_T != nullptr, "u_bar_T is NULL");
    alloc_adjoint(c);
    const int64_t n3 = 3 * c.nn;
    std::vector<double> h(n3);
    for (int64_t i = 0; i < c.nn; ++i)
        for (int d = 0; d < 3; ++d) h[3 * i + d] = ub_T[3 * (int64_t)c.node_iperm[i] + d];
    D.ub.upload(h.data(), n3, c.stream);
    if (vb_T) {
        for (int64_t i = 0; i < c.nn; ++i)
            for (int d = 0; d < 3; ++d) h[3 * i + d] = vb_T[3 * (int64_t)c.node_iperm[i] + d];
        D.vb.upload(h.data(), n3, c.stream);
    }
    const int K = D.interval;
    D.seg_u.alloc((size_t)K * n3);
    const int nseg = (int)D.ckpt.size();
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * K, len = std::min(K, D.n_steps - s0);
        restore_state(c, D.ckpt[sg]->p);
        for (int k = 0; k < len; ++k) {
            step_once_tiles(c, D.dt);
            APHIP_HIP(hipMemcpyAsync(D.seg_u.p + (size_t)k * n3, c.u.p, sizeof(double) * n3, hipMemcpyDeviceToDevice,
                                     c.stream));
        }
        for (int k = len - 1; k >= 0; --k) adj_run(c, D.seg_u.p + (size_t)k * n3, nullptr, 1, D.dt);
    }
    // the initial acceleration a0 = f(u0)/m also depends on the material parameters
    adj_run(c, D.ckpt[0]->p, nullptr, 2, D.dt);
    if (v0_bar) {
        D.vb.download(h.data(), n3, c.stream);
        APHIP_HIP(hipStreamSynchronize(c.stream));
        for (int64_t i = 0; i < c.nn; ++i)
            for (int d = 0; d < 3; ++d) v0_bar[3 * (int64_t)c.node_iperm[i] + d] = h[3 * i + d];
    }
    parameter_gradients(c, E_bar, nu_bar);
    restore_state(c, D.final_state.p);
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace aphip
