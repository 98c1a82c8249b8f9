// Fused element-tile step for gfx950: one workgroup owns a Morton-contiguous tile
// of elements (one element per thread) and, in one launch,
//   A. stages the tile's nodes in LDS (reference coordinates and the predicted
//      displacement u + dt (v + dt/2 a) of the central-difference step) together
//      with the tile's local CSR, masses and boundary flags,
//   B. evaluates every element of the tile (B, H, P, stress update, -B P^T w) and
//      writes its force partials to LDS,
//   C. sums each tile node's partials through the local CSR in ascending element
//      order (deterministic; no atomics) and finishes the nodes that only this
//      tile touches (a = f/m, BCs, v = v1/2 + dt/2 a, writes u, v, a, f).
// Nodes shared with other tiles leave their tile sum in a slot of a per-node
// contiguous slot list; a second, much smaller launch adds each shared node's
// slots in ascending tile order and finishes it.  This replaces the reference's
// per-element atomic scatter (include/ComputeInternalForceGaussian.h:127-130,
// include/Field.h:235-246) and its separate node loops (src/Solvers/Explicit.cpp:405-452).
//
// Node numbering (set by the plan): the interior nodes of tile t are the
// contiguous range [int_base[t], int_base[t] + ninterior[t]); all shared nodes
// follow as one contiguous range [n_interior, n_nodes).  Interior nodes are
// therefore read and written without indirection, and the shared-node pass is
// coalesced.
#include <algorithm>
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

struct TileArgs {
    const double *X;
    double *u, *v, *a, *f;
    const double *mass;
    const uint8_t *fixed;
    const int32_t *tn_start;
    const int32_t *tile_e0;
    const int32_t *tn_ninterior;
    const int32_t *int_base;
    const int32_t *tn_node;
    const uint32_t *tn_csr;
    const int32_t *tn_bdst;
    const uint16_t *lcsr;
    const uint16_t *lconn;
    double *bpart;
    double *H, *P;
    const double *state_n;
    double *state_np1;
    MatParams mat;
    const double *elem_lambda, *elem_mu;
    const double *node_dt;  // per-node increment (ensembles) or null
    double gauss_a;
    int64_t ne, ntiles, nblocks;
    int te, maxtn;
    double dt;
};

// LDS carve-up (bytes), sized on the host from maxtn and te.
struct TileLds {
    double *X, *U, *V, *F, *M;
    uint32_t *csr;
    int32_t *bdst;
    uint16_t *ent;
    uint8_t *fx;
};

__device__ __forceinline__ TileLds carve(char *base, int maxtn, int te, int npe) {
    TileLds L;
    L.X = reinterpret_cast<double *>(base);
    L.U = L.X + 3 * maxtn;
    L.V = L.U + 3 * maxtn;
    L.M = L.V + 3 * maxtn;
    L.F = L.M + maxtn;
    L.csr = reinterpret_cast<uint32_t *>(L.F + 3 * te * npe);
    L.bdst = reinterpret_cast<int32_t *>(L.csr + maxtn);
    L.ent = reinterpret_cast<uint16_t *>(L.bdst + maxtn);
    L.fx = reinterpret_cast<uint8_t *>(L.ent + te * npe);
    return L;
}

// MODE 0: force only; 1: force + a = f/m (initial); 2: full central-difference step.
// QPAR (Hex8 only): eight lanes per element, one per Gauss point.
template <int NPE, int KIND, bool PER_QP, int MODE, bool QPAR>
__global__ __launch_bounds__(NPE == 4 ? 512 : (QPAR ? 512 : 128)) void k_tile(TileArgs A) {
    extern __shared__ __align__(16) char lds_raw[];
    const int64_t t = xcd_tile(A.nblocks);
    if (t >= A.ntiles) return;
    const TileLds L = carve(lds_raw, A.maxtn, A.te, NPE);
    const int tid = threadIdx.x;
    const int tb = A.tn_start[t];
    const int tn = A.tn_start[t + 1] - tb;
    const int nint = A.tn_ninterior[t];
    const int64_t ibase = A.int_base[t];

    // element-side prefetch: independent of the node staging below
    const int64_t te0 = A.tile_e0[t];
    const int nel = A.tile_e0[t + 1] - (int)te0;
    const int le = QPAR ? (tid >> 3) : tid;
    const int64_t e = te0 + le;
    const bool has_elem = le < nel;
    int lc[NPE];
    MatParams m = A.mat;
    double sn = 0.0;
    if (has_elem) {
        if constexpr (NPE == 4) {
            const ushort4 c = reinterpret_cast<const ushort4 *>(A.lconn)[e];
            lc[0] = c.x; lc[1] = c.y; lc[2] = c.z; lc[3] = c.w;
        } else {
            const ushort4 c0 = reinterpret_cast<const ushort4 *>(A.lconn)[2 * e];
            const ushort4 c1 = reinterpret_cast<const ushort4 *>(A.lconn)[2 * e + 1];
            lc[0] = c0.x; lc[1] = c0.y; lc[2] = c0.z; lc[3] = c0.w;
            lc[4] = c1.x; lc[5] = c1.y; lc[6] = c1.z; lc[7] = c1.w;
        }
        if (A.elem_lambda) {
            m.lambda = A.elem_lambda[e];
            m.mu = A.elem_mu[e];
        }
        if constexpr (KIND == APHIP_J2 && !PER_QP && !QPAR) sn = A.state_n[e];
    }

    // A: tile CSR entries and node data into LDS
    {
        const int nent = A.te * NPE;  // entries of a full tile; a partial tile reads zeros past its end
        const uint32_t *src = reinterpret_cast<const uint32_t *>(A.lcsr + t * (int64_t)nent);
        uint32_t *dst = reinterpret_cast<uint32_t *>(L.ent);
        for (int i = tid; i < nent / 2; i += blockDim.x) dst[i] = src[i];
    }
    for (int ln = tid; ln < tn; ln += blockDim.x) {
        const bool interior = ln < nint;
        const int64_t g = interior ? ibase + ln : (int64_t)A.tn_node[tb + ln];
        L.csr[ln] = A.tn_csr[tb + ln];
        if (!interior) L.bdst[ln] = A.tn_bdst[tb + ln];
        double x[3], u[3];
        ld3(A.X, g, x);
        ld3(A.u, g, u);
        st3(L.X, ln, x);
        const unsigned fx = A.fixed ? A.fixed[g] : 0u;
        L.fx[ln] = (uint8_t)fx;
        if constexpr (MODE >= 1) {
            if (interior) L.M[ln] = A.mass[g];
        }
        if constexpr (MODE == 2) {
            double v[3], a[3], vh[3], u1[3];
            ld3(A.v, g, v);
            ld3(A.a, g, a);
            cd_predict(u, v, a, fx, A.node_dt ? A.node_dt[g] : A.dt, vh, u1);
            st3(L.U, ln, u1);
            st3(L.V, ln, vh);
        } else {
            st3(L.U, ln, u);
        }
    }
    __syncthreads();

    // B: one element per thread, or one Gauss point per lane
    if constexpr (QPAR && NPE == 8) {
        if (has_elem) {
            const int q = tid & 7;
            const int64_t so = PER_QP ? e * 8 : e;
            double H[9], P[9], fk[3];
            hex8_qp_lane<KIND, PER_QP>(q, lc, L.X, L.U, m, A.gauss_a, KIND == APHIP_J2 ? A.state_n + so : nullptr,
                                       KIND == APHIP_J2 ? A.state_np1 + so : nullptr, H, P, fk);
            if (q == 7) {
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    A.H[i * A.ne + e] = H[i];
                    A.P[i * A.ne + e] = P[i];
                }
            }
            st3(L.F, le * NPE + q, fk);
        }
    } else if (has_elem) {
        double X[NPE][3], U[NPE][3], fe[NPE][3];
#pragma unroll
        for (int k = 0; k < NPE; ++k) {
            ld3(L.X, lc[k], X[k]);
            ld3(L.U, lc[k], U[k]);
        }
        double H[9], P[9];
        if constexpr (NPE == 4) {
            double snp1 = 0.0;
            tet4_element<KIND>(X, U, m, sn, snp1, H, P, fe);
            if constexpr (KIND == APHIP_J2) A.state_np1[e] = snp1;
        } else {
            const int64_t so = PER_QP ? e * 8 : e;
            hex8_element<KIND, PER_QP>(X, U, m, A.gauss_a, KIND == APHIP_J2 ? A.state_n + so : nullptr,
                                       KIND == APHIP_J2 ? A.state_np1 + so : nullptr, H, P, fe);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            A.H[i * A.ne + e] = H[i];
            A.P[i * A.ne + e] = P[i];
        }
#pragma unroll
        for (int k = 0; k < NPE; ++k) st3(L.F, tid * NPE + k, fe[k]);
    }
    __syncthreads();

    // C: deterministic per-node gather from LDS; finish interior nodes, park shared-node sums
    for (int ln = tid; ln < tn; ln += blockDim.x) {
        const uint32_t pk = L.csr[ln];
        const int st = (int)(pk >> 16), cnt = (int)(pk & 0xffffu);
        double f[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < cnt; ++j) {
            const int s = L.ent[st + j];
            f[0] += L.F[3 * s];
            f[1] += L.F[3 * s + 1];
            f[2] += L.F[3 * s + 2];
        }
        if (ln < nint) {
            const int64_t g = ibase + ln;
            st3(A.f, g, f);
            if constexpr (MODE == 1) {
                const double mm = L.M[ln];
                const double acc[3] = {f[0] / mm, f[1] / mm, f[2] / mm};
                st3(A.a, g, acc);
            } else if constexpr (MODE == 2) {
                double vh[3], u1[3], a1[3], v1[3];
                ld3(L.V, ln, vh);
                ld3(L.U, ln, u1);
                cd_correct(f, L.M[ln], L.fx[ln], A.node_dt ? A.node_dt[g] : A.dt, vh, a1, v1);
                st3(A.u, g, u1);
                st3(A.v, g, v1);
                st3(A.a, g, a1);
            }
        } else {
            st3(A.bpart, L.bdst[ln], f);
        }
    }
}

struct BoundArgs {
    int64_t nb;
    int64_t first;  // global id of the first shared node
    const int32_t *bstart;
    const double *bpart;
    double *u, *v, *a, *f;
    const double *mass;
    const uint8_t *fixed;
    const double *node_dt;
    double dt;
};

template <int MODE>
__global__ __launch_bounds__(256) void k_boundary(BoundArgs A) {
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (b >= A.nb) return;
    const int64_t g = A.first + b;
    double f[3] = {0.0, 0.0, 0.0};
    const int32_t s1 = A.bstart[b + 1];
    for (int32_t s = A.bstart[b]; s < s1; ++s) {
        f[0] += A.bpart[3 * (int64_t)s];
        f[1] += A.bpart[3 * (int64_t)s + 1];
        f[2] += A.bpart[3 * (int64_t)s + 2];
    }
    st3(A.f, g, f);
    if constexpr (MODE == 1) {
        const double m = A.mass[g];
        const double acc[3] = {f[0] / m, f[1] / m, f[2] / m};
        st3(A.a, g, acc);
    } else if constexpr (MODE == 2) {
        double u[3], v[3], a[3], vh[3], u1[3], a1[3], v1[3];
        ld3(A.u, g, u);
        ld3(A.v, g, v);
        ld3(A.a, g, a);
        const unsigned fx = A.fixed ? A.fixed[g] : 0u;
        const double dt = A.node_dt ? A.node_dt[g] : A.dt;
        cd_predict(u, v, a, fx, dt, vh, u1);
        cd_correct(f, A.mass[g], fx, dt, vh, a1, v1);
        st3(A.u, g, u1);
        st3(A.v, g, v1);
        st3(A.a, g, a1);
    }
}

inline int64_t padded(int64_t n) { return ((n + kNumXcd - 1) / kNumXcd) * kNumXcd; }

template <int NPE, int KIND, bool PER_QP, int MODE>
void launch_tile(Ctx &c, const TileArgs &A, size_t lds) {
    if constexpr (NPE == 8) {
        if (c.tiles.threads == 8 * c.tiles.te) {
            auto kq = k_tile<NPE, KIND, PER_QP, MODE, true>;
            if (lds > 65536)
                APHIP_HIP(hipFuncSetAttribute((const void *)kq, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(kq, dim3((unsigned)A.nblocks), dim3(c.tiles.threads), lds, c.stream, A);
            return;
        }
    }
    auto kern = k_tile<NPE, KIND, PER_QP, MODE, false>;
    if (lds > 65536)
        APHIP_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)A.nblocks), dim3(c.tiles.threads), lds, c.stream, A);
}

template <int NPE, int MODE>
void dispatch_kind(Ctx &c, const TileArgs &A, size_t lds) {
    const bool pq = c.mat.state_per_qp && NPE == 8;
    switch (c.mat.kind) {
        case APHIP_LINEAR_ELASTIC: launch_tile<NPE, APHIP_LINEAR_ELASTIC, false, MODE>(c, A, lds); break;
        case APHIP_NEO_HOOKEAN: launch_tile<NPE, APHIP_NEO_HOOKEAN, false, MODE>(c, A, lds); break;
        case APHIP_J2:
            if (pq) launch_tile<NPE, APHIP_J2, true, MODE>(c, A, lds);
            else launch_tile<NPE, APHIP_J2, false, MODE>(c, A, lds);
            break;
        case APHIP_SVK: launch_tile<NPE, APHIP_SVK, false, MODE>(c, A, lds); break;
        default: throw Error(APHIP_ERR_INVALID, "material not set");
    }
}

}  // namespace

// Builds the tile plan for the current internal connectivity and renumbers the
// nodes so that interior nodes are contiguous per tile and shared nodes follow.
void tile_build(Ctx &c, int te, int threads) {
    const int npe = c.npe;
    const int64_t ne = c.ne, nn = c.nn;
    APHIP_REQUIRE(!c.cells.ready, "tiling applies to the Gaussian element path");
    c.graph_reset();
    if (npe == 4) {
        APHIP_REQUIRE(te == threads && (threads == 64 || threads == 128 || threads == 256 || threads == 512),
                      "Tet4 tiles: one element per thread, 64, 128, 256 or 512 threads");
    } else {
        const bool per_elem = te == threads && (threads == 64 || threads == 128);
        const bool per_qp = threads == 8 * te && (te == 8 || te == 16 || te == 32 || te == 64);
        APHIP_REQUIRE(per_elem || per_qp,
                      "Hex8 tiles: one element per thread (64 or 128) or eight lanes per element (8..64 elements)");
    }
    const std::vector<int32_t> &ci = c.conn_h;
    // tiles: consecutive chunks of te elements inside each element block
    std::vector<int32_t> te0v;
    for (size_t sg = 0; sg + 1 < c.elem_seg.size(); ++sg)
        for (int64_t e = c.elem_seg[sg]; e < c.elem_seg[sg + 1]; e += te) te0v.push_back((int32_t)e);
    const int64_t nt = (int64_t)te0v.size();
    te0v.push_back((int32_t)ne);

    // 1. node set of every tile and the number of tiles touching each node
    std::vector<int32_t> ntouch(nn, 0), stamp(nn, -1);
    std::vector<int32_t> tstart(nt + 1, 0), tnodes;
    tnodes.reserve((size_t)(ne * npe / 3 + 16));
    std::vector<int32_t> tmp;
    for (int64_t t = 0; t < nt; ++t) {
        tmp.clear();
        const int64_t e0 = te0v[t], e1 = te0v[t + 1];
        for (int64_t e = e0; e < e1; ++e)
            for (int k = 0; k < npe; ++k) {
                const int32_t n = ci[e * npe + k];
                if (stamp[n] != (int32_t)t) {
                    stamp[n] = (int32_t)t;
                    tmp.push_back(n);
                    ntouch[n]++;
                }
            }
        std::sort(tmp.begin(), tmp.end());
        tnodes.insert(tnodes.end(), tmp.begin(), tmp.end());
        APHIP_REQUIRE(tnodes.size() < (size_t)INT32_MAX, "tile node count overflow");
        tstart[t + 1] = (int32_t)tnodes.size();
    }

    // 2. new numbering: interior nodes tile by tile, then every other node in old order
    std::vector<int32_t> new_of_old(nn, -1), int_base(nt), ninterior(nt);
    int32_t next = 0;
    for (int64_t t = 0; t < nt; ++t) {
        int_base[t] = next;
        for (int32_t i = tstart[t]; i < tstart[t + 1]; ++i)
            if (ntouch[tnodes[i]] == 1) new_of_old[tnodes[i]] = next++;
        ninterior[t] = next - int_base[t];
    }
    const int64_t n_interior = next;
    for (int64_t n = 0; n < nn; ++n)
        if (new_of_old[n] < 0) new_of_old[n] = next++;
    const int64_t nb = nn - n_interior;

    // 3. tile node lists in new ids (interior first, each part ascending), local conn and CSR
    for (auto &x : tnodes) x = new_of_old[x];
    int maxtn = 0;
    std::vector<int32_t> loc(nn, -1);
    std::vector<uint16_t> lconn((size_t)ne * npe), lcsr((size_t)nt * te * npe, 0);
    std::vector<uint32_t> tcsr(tnodes.size());
    std::vector<int32_t> cnt, fillp;
    for (int64_t t = 0; t < nt; ++t) {
        auto b = tnodes.begin() + tstart[t], en = tnodes.begin() + tstart[t + 1];
        const int tn = (int)(en - b);
        APHIP_REQUIRE(tn <= 65535, "tile has more than 65535 nodes");
        maxtn = std::max(maxtn, tn);
        std::sort(b, en);  // interior ids are < n_interior, so they come first
        for (int i = 0; i < tn; ++i) loc[*(b + i)] = i;
        cnt.assign(tn, 0);
        const int64_t e0 = te0v[t], e1 = te0v[t + 1];
        for (int64_t e = e0; e < e1; ++e)
            for (int k = 0; k < npe; ++k) {
                const int ln = loc[new_of_old[ci[e * npe + k]]];
                lconn[e * npe + k] = (uint16_t)ln;
                cnt[ln]++;
            }
        fillp.assign(tn, 0);
        int run = 0;
        for (int ln = 0; ln < tn; ++ln) {
            fillp[ln] = run;
            tcsr[tstart[t] + ln] = ((uint32_t)run << 16) | (uint32_t)cnt[ln];
            run += cnt[ln];
        }
        uint16_t *dst = &lcsr[(size_t)t * te * npe];
        for (int64_t e = e0; e < e1; ++e)
            for (int k = 0; k < npe; ++k) {
                const int ln = loc[new_of_old[ci[e * npe + k]]];
                dst[fillp[ln]++] = (uint16_t)((e - e0) * npe + k);
            }
    }

    // 4. shared-node slots: contiguous per node, ascending tile order
    std::vector<int32_t> bstart(nb + 1, 0), tbdst(tnodes.size(), -1);
    for (int64_t t = 0; t < nt; ++t)
        for (int32_t i = tstart[t] + ninterior[t]; i < tstart[t + 1]; ++i) bstart[tnodes[i] - n_interior + 1]++;
    for (int64_t b = 0; b < nb; ++b) bstart[b + 1] += bstart[b];
    std::vector<int32_t> bfill(bstart.begin(), bstart.end() - 1);
    for (int64_t t = 0; t < nt; ++t)
        for (int32_t i = tstart[t] + ninterior[t]; i < tstart[t + 1]; ++i)
            tbdst[i] = bfill[tnodes[i] - n_interior]++;

    // 5. renumber the context's nodes, then upload the plan
    std::vector<int32_t> old_of_new(nn);
    for (int64_t n = 0; n < nn; ++n) old_of_new[new_of_old[n]] = (int32_t)n;
    renumber_nodes(c, old_of_new);

    TilePlan &P = c.tiles;
    P.te = te;
    P.threads = threads;
    P.ntiles = nt;
    P.maxtn = maxtn;
    P.nb = nb;
    P.n_interior = n_interior;
    P.total_tile_nodes = (int64_t)tnodes.size();
    P.tn_start.upload(tstart.data(), tstart.size(), c.stream);
    P.tile_e0.upload(te0v.data(), te0v.size(), c.stream);
    P.tn_ninterior.upload(ninterior.data(), ninterior.size(), c.stream);
    P.int_base.upload(int_base.data(), int_base.size(), c.stream);
    P.tn_node.upload(tnodes.data(), tnodes.size(), c.stream);
    P.tn_csr.upload(tcsr.data(), tcsr.size(), c.stream);
    P.tn_bdst.upload(tbdst.data(), tbdst.size(), c.stream);
    P.lcsr.upload(lcsr.data(), lcsr.size(), c.stream);
    P.lconn.upload(lconn.data(), lconn.size(), c.stream);
    P.bpart.alloc(3 * std::max<int64_t>(1, (int64_t)bstart[nb]));
    P.bstart.upload(bstart.data(), bstart.size(), c.stream);
    // LDS: X, U, V (3 doubles each) + mass per node, partials, csr/bdst per node, entries, flags
    P.lds_bytes = sizeof(double) * (10 * (size_t)maxtn + 3 * (size_t)te * npe) + 8 * (size_t)maxtn +
                  2 * (size_t)te * npe + (size_t)maxtn + 16;
    APHIP_REQUIRE(P.lds_bytes <= 160 * 1024, "tile needs more LDS than a CU has; use a smaller tile");
    P.ready = true;
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

void tile_run(Ctx &c, double dt, int mode) {
    TilePlan &P = c.tiles;
    APHIP_REQUIRE(P.ready, "tile plan not built");
    if (c.ne == 0) return;
    TileArgs A;
    A.X = c.X.p;
    A.u = c.u.p;
    A.v = c.v.p;
    A.a = c.a.p;
    A.f = c.f.p;
    A.mass = c.mass.p;
    A.fixed = c.has_fixed ? c.fixed.p : nullptr;
    A.tn_start = P.tn_start.p;
    A.tile_e0 = P.tile_e0.p;
    A.tn_ninterior = P.tn_ninterior.p;
    A.int_base = P.int_base.p;
    A.tn_node = P.tn_node.p;
    A.tn_csr = P.tn_csr.p;
    A.tn_bdst = P.tn_bdst.p;
    A.lcsr = P.lcsr.p;
    A.lconn = P.lconn.p;
    A.bpart = P.bpart.p;
    A.H = c.H.p;
    A.P = c.P.p;
    A.state_n = c.state[1 - c.state_np1].p;
    A.state_np1 = c.state[c.state_np1].p;
    A.mat = MatParams{c.mat.lambda, c.mat.mu, c.mat.yield_stress, c.mat.hardening};
    A.elem_lambda = c.has_fields ? c.elem_lambda.p : nullptr;
    A.elem_mu = c.has_fields ? c.elem_mu.p : nullptr;
    A.node_dt = c.node_dt.p;
    A.gauss_a = 1.0 / std::sqrt(3.0);
    A.ne = c.ne;
    A.ntiles = P.ntiles;
    A.nblocks = padded(P.ntiles);
    A.te = P.te;
    A.maxtn = P.maxtn;
    A.dt = dt;
    prof_begin(c, 1);
    if (c.npe == 4) {
        if (mode == 0) dispatch_kind<4, 0>(c, A, P.lds_bytes);
        else if (mode == 1) dispatch_kind<4, 1>(c, A, P.lds_bytes);
        else dispatch_kind<4, 2>(c, A, P.lds_bytes);
    } else {
        if (mode == 0) dispatch_kind<8, 0>(c, A, P.lds_bytes);
        else if (mode == 1) dispatch_kind<8, 1>(c, A, P.lds_bytes);
        else dispatch_kind<8, 2>(c, A, P.lds_bytes);
    }
    APHIP_HIP(hipGetLastError());
    prof_end(c, 1);
    if (P.nb > 0) {
        BoundArgs B{P.nb, P.n_interior, P.bstart.p, P.bpart.p, c.u.p, c.v.p, c.a.p, c.f.p, c.mass.p,
                    c.has_fixed ? c.fixed.p : nullptr, c.node_dt.p, dt};
        const dim3 grid((unsigned)((P.nb + 255) / 256));
        prof_begin(c, 2);
        if (mode == 0) hipLaunchKernelGGL(k_boundary<0>, grid, dim3(256), 0, c.stream, B);
        else if (mode == 1) hipLaunchKernelGGL(k_boundary<1>, grid, dim3(256), 0, c.stream, B);
        else hipLaunchKernelGGL(k_boundary<2>, grid, dim3(256), 0, c.stream, B);
        APHIP_HIP(hipGetLastError());
        prof_end(c, 2);
    }
}

}  // namespace aphip
