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
    const int32_t *tnp_node;  // fixed-stride copies of tn_node / tn_csr / tn_bdst / lconn
    const uint32_t *tnp_csr;
    const int32_t *tnp_bdst;
    const uint16_t *lconn_p;
    int tpad;
    double *bpart;
    double *H, *P;
    const double *Hn, *Pn;  // rate-form materials: H and P of the previous step
    double rate_dt;         // and the increment of this evaluation
    const double *state_n;
    double *state_np1;
    MatParams mat;
    const double *elem_lambda, *elem_mu;
    const double *node_dt;  // per-node increment (ensembles) or null
    const double *zero;     // 8 zero doubles: stands in for an absent optional array
    const int32_t *meta;    // [ntiles * 8] packed per-tile scalars (TileMeta)
    double gauss_a;
    int64_t ne, ntiles, nblocks;
    int64_t nn, nslot;      // nodes; shared-node slots (bpart holds 3 doubles each)
    int te, maxtn;
    int max_nint;  // most interior nodes in one tile (Hex8: LDS rows of the corrector data)
    double dt;
    // mode 2: the shared nodes' displacement predicted by the previous step's shared-node
    // pass (row first_shared + b at upred[3 b]), or null: predicted here from u, v, a
    const double *upred;
    int64_t first_shared;
};

// LDS carve-up (bytes), sized on the host (tile_lds_bytes) from maxtn, max_nint, te
// and the number of force rounds.  X and U (the predicted displacement) are kept for
// every tile node; v_half, m, dt and the fixed mask for max_nint rows (Hex8: the
// interior nodes, the only ones the kernel finishes; Tet4: every tile node); F holds
// the element forces of one round.
struct TileLds {
    double *X, *U, *V, *F, *M, *DT;
    uint32_t *csr;
    int32_t *bdst;
    uint16_t *ent;
    uint8_t *fx;
};

// Rows between the slot blocks of the element-force rows: te_round + pad with
// stride = 4 (mod 32), so that the 24-byte rows of one element's slots start 24 banks
// apart (mod 64): the 4 or 8 slots of an element, which the nodes of that element
// read at the same time in the gather, never share a bank.
__host__ __device__ __forceinline__ int f_stride(int te_round) { return te_round + ((4 - te_round) & 31); }

__host__ __device__ __forceinline__ TileLds carve(char *base, int maxtn, int max_nint, int te, int f_rounds,
                                                  int npe) {
    TileLds L;
    L.X = reinterpret_cast<double *>(base);
    L.U = L.X + 3 * maxtn;
    L.V = L.U + 3 * maxtn;
    L.M = L.V + 3 * max_nint;
    L.DT = L.M + max_nint;
    L.F = L.DT + max_nint;
    L.csr = reinterpret_cast<uint32_t *>(L.F + 3 * (size_t)f_stride(te / f_rounds) * npe);
    L.bdst = reinterpret_cast<int32_t *>(L.csr + maxtn);
    L.ent = reinterpret_cast<uint16_t *>(L.bdst + maxtn);
    L.fx = reinterpret_cast<uint8_t *>(L.ent + te * npe);
    return L;
}

__host__ inline size_t tile_lds_bytes(int maxtn, int max_nint, int te, int f_rounds, int npe) {
    return sizeof(double) * (6 * (size_t)maxtn + 5 * (size_t)max_nint + 3 * (size_t)f_stride(te / f_rounds) * npe) +
           8 * (size_t)maxtn + 2 * (size_t)te * npe + (size_t)max_nint + 16;
}

// Optional arrays are read without branches: an absent array is replaced by the
// zero buffer at index 0 (a uniform select), and the value is then selected.  A
// load under a branch ends a basic block, and at the join the compiler waits for
// every outstanding memory operation of the wave, stores included.
__device__ __forceinline__ unsigned load_fixed(const TileArgs &A, int64_t g) {
    const uint8_t *p = A.fixed ? A.fixed : reinterpret_cast<const uint8_t *>(A.zero);
    return p[A.fixed ? g : 0];
}

__device__ __forceinline__ double load_node_dt(const TileArgs &A, int64_t g) {
    const double *p = A.node_dt ? A.node_dt : A.zero;
    const double d = p[A.node_dt ? g : 0];
    return A.node_dt ? d : A.dt;
}

__device__ __forceinline__ void load_elem_lame(const TileArgs &A, int64_t e, MatParams &m) {
    const double *pl = A.elem_lambda ? A.elem_lambda : A.zero;
    const double *pm = A.elem_lambda ? A.elem_mu : A.zero;
    const int64_t i = A.elem_lambda ? e : 0;
    const double l = pl[i], mu = pm[i];
    m.lambda = A.elem_lambda ? l : m.lambda;
    m.mu = A.elem_lambda ? mu : m.mu;
}

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt at their maxima: not waited on).  At a
// point where no store is outstanding this costs nothing beyond the loads' latency.
__device__ __forceinline__ void wait_vmem_loads() { __builtin_amdgcn_s_waitcnt(0x0f70); }

// Node ln's force f is complete: an interior node is finished and written (MODE 0: f;
// 1: f and a = f/m; 2: the corrector); a shared node parks f in its slot for the
// shared-node pass.
template <int MODE>
__device__ __forceinline__ void finish_node(const TileArgs &A, const TileLds &L, int ln, int nint, int64_t ibase,
                                            const double (&f)[3]) {
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
            cd_correct(f, L.M[ln], L.fx[ln], L.DT[ln], vh, a1, v1);
            st3(A.u, g, u1);
            st3(A.v, g, v1);
            st3(A.a, g, a1);
        }
    } else {
        st3(A.bpart, L.bdst[ln], f);
    }
}

// MODE 0: force only; 1: force + a = f/m (initial); 2: full central-difference step.
// QPAR (Hex8 only): eight lanes per element, one per Gauss point.
template <int NPE, int KIND, bool PER_QP, int MODE, bool QPAR>
__global__ __launch_bounds__(NPE == 4 ? 1024 : (QPAR ? 512 : 256), NPE == 8 && !QPAR ? 2 : 1) void k_tile(TileArgs A) {
    extern __shared__ __align__(16) char lds_raw[];
    const int64_t t = xcd_tile(A.nblocks);
    if (t >= A.ntiles) return;
    // One-element-per-thread Hex8 tiles pass their forces through LDS in two rounds
    // (24 doubles per element otherwise keep a 128-element tile at one wave per SIMD);
    // Tet4 and QPAR tiles in one.  The host sizes the LDS with the same counts.
    constexpr int ROUNDS = (NPE == 8 && !QPAR) ? 2 : 1;
    const TileLds L = carve(lds_raw, A.maxtn, NPE == 4 ? A.maxtn : A.max_nint, A.te, ROUNDS, NPE);
    const int tid = threadIdx.x;
    const int le = QPAR ? (tid >> 3) : tid;
    int lc[NPE];
    // (le < te always; the fixed-stride copy needs no per-tile scalar)
    const int64_t ep = t * (int64_t)A.te + le;
    if constexpr (NPE == 4) {
        const ushort4 c = reinterpret_cast<const ushort4 *>(A.lconn_p)[ep];
        lc[0] = c.x; lc[1] = c.y; lc[2] = c.z; lc[3] = c.w;
    } else {
        const ushort4 c0 = reinterpret_cast<const ushort4 *>(A.lconn_p)[2 * ep];
        const ushort4 c1 = reinterpret_cast<const ushort4 *>(A.lconn_p)[2 * ep + 1];
        lc[0] = c0.x; lc[1] = c0.y; lc[2] = c0.z; lc[3] = c0.w;
        lc[4] = c1.x; lc[5] = c1.y; lc[6] = c1.z; lc[7] = c1.w;
    }

    // A: tile CSR entries and node data into LDS.  Every global load of the stage is
    // issued before the first LDS write: the CSR words and the tile-node indices
    // together, then the node data of the first NPT node rows of each thread.  (A load
    // followed by its LDS write in a loop waits for that load, and vmcnt counts in
    // order, so each iteration would add a full memory latency.)
    constexpr int EW = NPE == 4 ? 2 : (QPAR ? 1 : 4);  // CSR words per thread: nent / 2 <= EW x blockDim
    constexpr int NPT = NPE == 4 ? 1 : (QPAR ? 1 : 2); // node rows per thread loaded together
    const int nw = A.te * NPE / 2;  // words of a full tile; a partial tile reads zeros past its end
    const uint32_t *ent_src = reinterpret_cast<const uint32_t *>(A.lcsr) + t * (int64_t)nw;
    uint32_t ew[EW];
#pragma unroll
    for (int r = 0; r < EW; ++r) ew[r] = ent_src[min(tid + r * (int)blockDim.x, nw - 1)];
    int64_t gq[NPT];
    uint32_t csrq[NPT];
    int32_t bdq[NPT];
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        // all loads unconditional (tn_node also lists the interior nodes, as ibase + ln)
        const int64_t row = t * (int64_t)A.tpad + min(tid + q * (int)blockDim.x, A.tpad - 1);
        gq[q] = A.tnp_node[row];
        csrq[q] = A.tnp_csr[row];
        bdq[q] = A.tnp_bdst[row];
    }
    double xq[NPT][3], uq[NPT][3], vq[NPT][3], aq[NPT][3], mq[NPT], dq[NPT];
    unsigned fq[NPT];
    bool pq[NPT];  // shared row with a stored prediction: X and the predicted u only
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int64_t g = gq[q];
        pq[q] = MODE == 2 && A.upred != nullptr && bdq[q] >= 0;
        ld3(A.X, g, xq[q]);
        mq[q] = 1.0;
        dq[q] = A.dt;
        if (pq[q]) {
            ld3(A.upred, g - A.first_shared, uq[q]);
            fq[q] = 0u;
#pragma unroll
            for (int d = 0; d < 3; ++d) vq[q][d] = aq[q][d] = 0.0;
        } else {
            ld3(A.u, g, uq[q]);
            fq[q] = load_fixed(A, g);
            if constexpr (MODE >= 1) mq[q] = A.mass[g];
            if constexpr (MODE == 2) {
                ld3(A.v, g, vq[q]);
                ld3(A.a, g, aq[q]);
                dq[q] = load_node_dt(A, g);
            }
        }
    }
    // per-tile scalars, read after the staging loads are out (their addresses need none)
    const int tb = A.tn_start[t];
    const int tn = A.tn_start[t + 1] - tb;
    const int nint = A.tn_ninterior[t];
    const int64_t ibase = A.int_base[t];
    const int64_t te0 = A.tile_e0[t];
    const int nel = A.tile_e0[t + 1] - (int)te0;
    const int64_t e = te0 + le;
    const bool has_elem = le < nel;
    const int64_t ec = te0 + min(le, nel - 1);  // clamped: the loads below need no branch
    MatParams m = A.mat;
    double sn = 0.0;
    load_elem_lame(A, ec, m);
    double sn2 = 0.0;
    if constexpr (NPE == 4) {
        sn = load_state<KIND>(A.state_n, ec);
        sn2 = load_state2<KIND>(A.state_n, ec);
    }
    RateIn rin;
    load_rate<KIND>(rin, A.Hn, A.Pn, A.ne, ec, A.rate_dt);

    {
        uint32_t *dst = reinterpret_cast<uint32_t *>(L.ent);
#pragma unroll
        for (int r = 0; r < EW; ++r) {
            const int i = tid + r * (int)blockDim.x;
            if (i < nw) dst[i] = ew[r];
        }
    }
    // Hex8 keeps the corrector data of interior nodes only (the only ones it finishes);
    // Tet4 writes it for every node (measured faster with its 512-element tiles).
    // pre: u is already the predicted displacement (a shared row; the corrector data
    // written for it are never read)
    auto stage = [&](int ln, uint32_t csr, int32_t bdst, const double *x, const double *u, const double *v,
                     const double *a, unsigned fx, double mm, double nd, bool pre) {
        L.csr[ln] = csr;
        L.bdst[ln] = bdst;
        st3(L.X, ln, x);
        const bool interior = NPE == 4 || ln < nint;
        if (interior) {
            L.fx[ln] = (uint8_t)fx;
            L.M[ln] = mm;
            L.DT[ln] = nd;
        }
        if constexpr (MODE == 2) {
            double vh[3] = {0.0, 0.0, 0.0}, u1[3] = {u[0], u[1], u[2]};
            if (!pre) cd_predict(u, v, a, fx, nd, vh, u1);
            st3(L.U, ln, u1);
            if (interior) st3(L.V, ln, vh);
        } else {
            st3(L.U, ln, u);
        }
    };
#pragma unroll
    for (int q = 0; q < NPT; ++q) {
        const int ln = tid + q * (int)blockDim.x;
        if (ln < tn) stage(ln, csrq[q], bdq[q], xq[q], uq[q], vq[q], aq[q], fq[q], mq[q], dq[q], pq[q]);
    }
    // rows past NPT x blockDim (tiles with unusually many nodes)
    for (int ln = tid + NPT * (int)blockDim.x; ln < tn; ln += blockDim.x) {
        const int64_t g = A.tn_node[tb + ln];
        double x[3], u[3], v[3], a[3];
        ld3(A.X, g, x);
        ld3(A.u, g, u);
        const unsigned fx = load_fixed(A, g);
        double mm = 1.0, nd = A.dt;
        if constexpr (MODE >= 1) mm = A.mass[g];
        if constexpr (MODE == 2) {
            ld3(A.v, g, v);
            ld3(A.a, g, a);
            nd = load_node_dt(A, g);
        }
        stage(ln, A.tn_csr[tb + ln], A.tn_bdst[tb + ln], x, u, v, a, fx, mm, nd, false);
    }
    // Every load of the wave completes here, on every path (a wave whose lanes all
    // skip the staging loop still has its element prefetch in flight).  Otherwise the
    // join after the element phase merges a path with pending loads into the path
    // with the H / P stores, and the first register reuse in the gather waits for
    // vmcnt(0): for every store of the wave.
    wait_vmem_loads();
    __syncthreads();

    // B: one element per thread, or one Gauss point per lane.  The element's nodal
    // forces (QPAR: this lane's node) stay in registers until their round's LDS pass.
    constexpr int KF = (QPAR && NPE == 8) ? 1 : NPE;
    double fe[KF][3];
    // Tet4: the element phase stores H and keeps P for the start of the gather phase.
    // The 144 B per element of H and P, all issued during the element phase, leave a CU
    // faster than the memory takes them while many CUs are in that phase; split between
    // the two phases they overlap the gather (tile kernel 274 -> 250 us on the 6M-tet
    // block; H and P both after the barrier: 277 us; P after the gather: 256 us).  Hex8
    // tiles, whose element phase is arithmetic-bound, measured the same either way.
    double pdef[NPE == 4 ? 9 : 1];
    const int te_round = ROUNDS == 1 ? A.te : A.te / ROUNDS;
    const int my_round = ROUNDS == 1 ? 0 : le / te_round;
    // LDS row of entry (element j of the round, slot k).  One element per thread:
    // slot-major with padded slot blocks, so a wave's writes of one slot are
    // consecutive 24-byte rows (no bank conflicts; element-major rows 96 bytes apart
    // hit 16 of the 64 banks) and the slots of one element sit in distinct banks for
    // the gather.  One Gauss point per lane: element-major, consecutive for those lanes.
    const int fs = f_stride(te_round);
    auto frow = [fs](int j, int k) { return QPAR ? j * NPE + k : k * fs + j; };
    if constexpr (QPAR && NPE == 8) {
        if (has_elem) {
            const int q = tid & 7;
            const int64_t so = state_width<KIND>() * (PER_QP ? e * 8 : e);
            double H[9], P[9];
            hex8_qp_lane<KIND, PER_QP>(q, lc, L.X, L.U, m, A.gauss_a, state_width<KIND>() ? A.state_n + so : nullptr,
                                       state_width<KIND>() ? A.state_np1 + so : nullptr, H, P, fe[0], &rin);
            if (q == 7) {
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    store_once(A.H + i * A.ne + e, H[i]);
                    store_once(A.P + i * A.ne + e, P[i]);
                }
            }
            if (my_round == 0) st3(L.F, frow(le, q), fe[0]);
        }
    } else if (has_elem) {
        double X[NPE][3], U[NPE][3];
#pragma unroll
        for (int k = 0; k < NPE; ++k) {
            ld3(L.X, lc[k], X[k]);
            ld3(L.U, lc[k], U[k]);
        }
        double H[9], P[9];
        if constexpr (NPE == 4) {
            double snp1 = 0.0, aux = 0.0;
            tet4_element<KIND>(X, U, m, sn, snp1, aux, H, P, fe, &rin, sn2);
            store_state<KIND>(A.state_np1, e, snp1, aux);
        } else {
            const int64_t so = state_width<KIND>() * (PER_QP ? e * 8 : e);
            hex8_element<KIND, PER_QP>(X, U, m, A.gauss_a, state_width<KIND>() ? A.state_n + so : nullptr,
                                       state_width<KIND>() ? A.state_np1 + so : nullptr, H, P, fe, &rin);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            store_once(A.H + i * A.ne + e, H[i]);
            if constexpr (NPE == 4) pdef[i] = P[i];
            else store_once(A.P + i * A.ne + e, P[i]);
        }
        if (my_round == 0) {
#pragma unroll
            for (int k = 0; k < NPE; ++k) st3(L.F, frow(le, k), fe[k]);
        }
    }
    __syncthreads();
    if constexpr (NPE == 4) {
        if (has_elem) {
#pragma unroll
            for (int i = 0; i < 9; ++i) store_once(A.P + i * A.ne + e, pdef[i]);
        }
    }

    // C: deterministic per-node gather from LDS (each node's entries in ascending order:
    // round r adds the entries of elements [r, r + 1) x te / ROUNDS, a contiguous run of
    // the node's sorted list, so the sum is the same sequence for every round count);
    // then finish interior nodes and park shared-node sums.
    if constexpr (ROUNDS == 1) {
        for (int ln = tid; ln < tn; ln += blockDim.x) {
            const uint32_t pk = L.csr[ln];
            const int st = (int)(pk >> 16), cnt = (int)(pk & 0xffffu);
            double f[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < cnt; ++j) {
                const int s = L.ent[st + j];
                const int rw = 3 * frow(s / NPE, s % NPE);
                f[0] += L.F[rw];
                f[1] += L.F[rw + 1];
                f[2] += L.F[rw + 2];
            }
            finish_node<MODE>(A, L, ln, nint, ibase, f);
        }
        return;
    }
    // Several rounds: a thread owns the tile nodes tid + q * blockDim (maxtn <= NPE x te
    // <= NPE x blockDim) and keeps their partial sums in registers between rounds.
    double acc[NPE][3];
    int jpos[NPE];
#pragma unroll
    for (int q = 0; q < NPE; ++q) {
        acc[q][0] = acc[q][1] = acc[q][2] = 0.0;
        jpos[q] = 0;
    }
    for (int r = 0; r < ROUNDS; ++r) {
        if (r > 0) {
            __syncthreads();  // the previous round's gather is done with F
            if (has_elem && my_round == r) {
                if constexpr (QPAR && NPE == 8) {
                    st3(L.F, frow(le - r * te_round, tid & 7), fe[0]);
                } else {
#pragma unroll
                    for (int k = 0; k < NPE; ++k) st3(L.F, frow(le - r * te_round, k), fe[k]);
                }
            }
            __syncthreads();
        }
        const int off = r * te_round * NPE, lim = off + te_round * NPE;
#pragma unroll
        for (int q = 0; q < NPE; ++q) {
            const int ln = tid + q * (int)blockDim.x;
            if (ln < tn) {
                const uint32_t pk = L.csr[ln];
                const int st = (int)(pk >> 16), cnt = (int)(pk & 0xffffu);
                int j = jpos[q];
                for (; j < cnt; ++j) {
                    const int s = L.ent[st + j];
                    if (s >= lim) break;
                    const int rw = 3 * frow((s - off) / NPE, (s - off) % NPE);
                    acc[q][0] += L.F[rw];
                    acc[q][1] += L.F[rw + 1];
                    acc[q][2] += L.F[rw + 2];
                }
                jpos[q] = j;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NPE; ++q) {
        const int ln = tid + q * (int)blockDim.x;
        if (ln < tn) finish_node<MODE>(A, L, ln, nint, ibase, acc[q]);
    }
}

// ---- persistent, software-pipelined Tet4 step -------------------------------
// Same arithmetic and summation order as k_tile<4, KIND, false, 2, false>, so the
// results are bitwise identical.  A block owns the tile sequence
//   tbeg(xcd) + kx, + bx, + 2 bx, ...   (xcd = block % 8, kx = block / 8)
// inside the Morton range of its XCD and overlaps the global loads of the next
// tiles with the work of the current one: while tile j is evaluated, the node
// data of tile j+1, the node slots of tile j+2 and the scalars of tile j+3 are in
// flight.  Node staging is double-buffered in LDS and tile j+1 is committed to the
// idle buffer after tile j's element arithmetic but before tile j's H / P stores.
// gfx9 counts loads and stores on one in-order counter, so this order lets the
// wait for the prefetched data leave the fresh stores in flight.  Every global
// access in the loop is unconditional (idle lanes use clamped loads and dropped
// buffer stores): a memory instruction under a branch would make the compiler's
// wait counts conservative at the join, and a conservative count drains the stores.
// Requires one element per thread and at most one tile node per thread.

// Per-tile scalars, packed so that one 32-byte load brings all of them.
struct TileMeta {
    int tb, tn, nint, nel;  // first tile-node slot, tile nodes, interior nodes, elements
    int ibase, e0, pad0, pad1;  // first interior node id, first element
};

__device__ __forceinline__ TileMeta load_meta(const TileArgs &A, int64_t t) {
    const int4 *p = reinterpret_cast<const int4 *>(A.meta + 8 * t);
    const int4 a = p[0], b = p[1];
    return TileMeta{a.x, a.y, a.z, a.w, b.x, b.y, 0, 0};
}

// Raw buffer stores (stride 0): a store whose offset lies past the resource's
// num_records is discarded by the hardware's range check, which turns a
// per-lane condition into an unconditional instruction with no memory traffic.
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
constexpr uint32_t kDropOffset = 0x80000000u;  // beyond every resource (all are < 2 GiB)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

// aux: cache policy (2 = nt, the non-temporal hint used for the H / P stream)
template <int AUX = 0>
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t off, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, off, 0, AUX);
}

__device__ __forceinline__ void bstore3(__amdgpu_buffer_rsrc_t r, uint32_t off, const double *v) {
    bstore(r, off, v[0]);
    bstore(r, off + 8, v[1]);
    bstore(r, off + 16, v[2]);
}

struct StageIdx {  // per-thread node slot of a tile
    int g;
    uint32_t csr;
    int bdst;
};

struct StageData {  // one thread's share of a tile, from global memory
    TileMeta mt;
    uint32_t csr;
    int bdst;
    double x[3], u[3], v[3], a[3];
    double m, ndt, lam, mu, sn;
    unsigned fx;
    ushort4 lc;
    uint32_t w0, w1;
};

// One staging buffer: the tile's nodes and elements as the element and node phases read them.
struct StageLds {
    double *X, *U, *V, *M, *DT, *LAM, *MU, *SN;
    uint32_t *csr;
    int32_t *bdst;
    ushort4 *lc;
    uint32_t *ent;  // the tile's CSR entries (u16 pairs)
    uint8_t *fx;
};

__host__ __device__ inline size_t stage_bytes(int te, int maxtn) {
    const size_t d = 8 * (11 * (size_t)maxtn + 3 * (size_t)te);  // X U V M DT LAM MU SN
    const size_t w = 4 * (2 * (size_t)maxtn) + 8 * (size_t)te + 8 * (size_t)te;  // csr bdst lc ent
    return (d + w + (size_t)maxtn + 15) & ~(size_t)15;
}

__host__ __device__ inline size_t stream_lds_bytes(int te, int maxtn) {
    return 96 * (size_t)te + 2 * stage_bytes(te, maxtn);
}

__device__ __forceinline__ StageLds stage_carve(char *base, int te, int maxtn) {
    StageLds S;
    S.X = reinterpret_cast<double *>(base);
    S.U = S.X + 3 * maxtn;
    S.V = S.U + 3 * maxtn;
    S.M = S.V + 3 * maxtn;
    S.DT = S.M + maxtn;
    S.LAM = S.DT + maxtn;
    S.MU = S.LAM + te;
    S.SN = S.MU + te;
    S.lc = reinterpret_cast<ushort4 *>(S.SN + te);
    S.ent = reinterpret_cast<uint32_t *>(S.lc + te);
    S.csr = S.ent + 2 * te;
    S.bdst = reinterpret_cast<int32_t *>(S.csr + maxtn);
    S.fx = reinterpret_cast<uint8_t *>(S.bdst + maxtn);
    return S;
}

__device__ __forceinline__ void stage_idx(const TileArgs &A, const TileMeta &mt, int tid, StageIdx &s) {
    const int slot = mt.tb + min(tid, mt.tn - 1);  // clamped: no branch around the loads
    s.g = A.tn_node[slot];
    s.csr = A.tn_csr[slot];
    s.bdst = A.tn_bdst[slot];
}

// Optional per-node increment and per-element moduli are read from the zero buffer
// when absent and added to the uniform value, so the loaded value is used on both
// paths and the compiler keeps the loads unconditional.
template <int KIND>
__device__ __forceinline__ void stage_data(const TileArgs &A, int64_t t, const TileMeta &mt, int tid,
                                           const StageIdx &si, StageData &d) {
    d.mt = mt;
    d.csr = si.csr;
    d.bdst = si.bdst;
    const int64_t g = si.g;
    ld3(A.X, g, d.x);
    ld3(A.u, g, d.u);
    ld3(A.v, g, d.v);
    ld3(A.a, g, d.a);
    d.fx = load_fixed(A, g);
    d.m = A.mass[g];
    const double *pd = A.node_dt ? A.node_dt : A.zero;
    d.ndt = pd[A.node_dt ? g : 0] + (A.node_dt ? 0.0 : A.dt);
    const int64_t e = (int64_t)mt.e0 + min(tid, mt.nel - 1);
    d.lc = reinterpret_cast<const ushort4 *>(A.lconn)[e];
    const int64_t ie = A.elem_lambda ? e : 0;
    const double *pl = A.elem_lambda ? A.elem_lambda : A.zero;
    const double *pm = A.elem_lambda ? A.elem_mu : A.zero;
    d.lam = pl[ie] + (A.elem_lambda ? 0.0 : A.mat.lambda);
    d.mu = pm[ie] + (A.elem_lambda ? 0.0 : A.mat.mu);
    d.sn = 0.0;
    if constexpr (KIND == APHIP_J2) d.sn = A.state_n[e];
    // the tile's CSR entries: te * 4 u16 = 2 words per thread
    const uint32_t *src = reinterpret_cast<const uint32_t *>(A.lcsr + t * (int64_t)(A.te * 4));
    d.w0 = src[tid];
    d.w1 = src[tid + A.te];
}

// Writes one thread's share of a staged tile into an LDS buffer (predictor applied).
__device__ __forceinline__ void stage_commit(const StageData &d, const StageLds &S, int tid, int te) {
    if (tid < d.mt.tn) {
        S.csr[tid] = d.csr;
        S.bdst[tid] = d.bdst;
        st3(S.X, tid, d.x);
        S.fx[tid] = (uint8_t)d.fx;
        S.M[tid] = d.m;
        S.DT[tid] = d.ndt;
        double vh[3], u1[3];
        cd_predict(d.u, d.v, d.a, d.fx, d.ndt, vh, u1);
        st3(S.U, tid, u1);
        st3(S.V, tid, vh);
    }
    S.lc[tid] = d.lc;
    S.LAM[tid] = d.lam;
    S.MU[tid] = d.mu;
    S.SN[tid] = d.sn;
    S.ent[tid] = d.w0;
    S.ent[tid + te] = d.w1;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_tile_stream(TileArgs A, int nblk) {
    extern __shared__ __align__(16) char lds_raw[];
    const int te = A.te;
    double *F = reinterpret_cast<double *>(lds_raw);
    const size_t sb = stage_bytes(te, A.maxtn);
    const StageLds S0 = stage_carve(lds_raw + 96 * (size_t)te, te, A.maxtn);
    const StageLds S1 = stage_carve(lds_raw + 96 * (size_t)te + sb, te, A.maxtn);
    const int tid = threadIdx.x;
    const int xcd = blockIdx.x % kNumXcd, kx = blockIdx.x / kNumXcd;
    const int bx = nblk / kNumXcd;
    const int64_t per = (A.ntiles + kNumXcd - 1) / kNumXcd;
    const int64_t tend = min(A.ntiles, (int64_t)(xcd + 1) * per);
    int64_t t = (int64_t)xcd * per + kx;
    if (t >= tend) return;
    // every prefetch is unconditional, on tile ids clamped to the block's range: a
    // branch around a load would make the compiler wait for it at the join
    const int64_t tlast = tend - 1;
    const __amdgpu_buffer_rsrc_t rH = buffer_rsrc(A.H, 72 * A.ne), rP = buffer_rsrc(A.P, 72 * A.ne);
    const __amdgpu_buffer_rsrc_t rS = buffer_rsrc(A.state_np1, KIND == APHIP_J2 ? 8 * A.ne : 0);
    const __amdgpu_buffer_rsrc_t rF = buffer_rsrc(A.f, 24 * A.nn), rU = buffer_rsrc(A.u, 24 * A.nn);
    const __amdgpu_buffer_rsrc_t rV = buffer_rsrc(A.v, 24 * A.nn), rA = buffer_rsrc(A.a, 24 * A.nn);
    const __amdgpu_buffer_rsrc_t rB = buffer_rsrc(A.bpart, 24 * A.nslot);

    StageIdx si;
    StageData d;
    {
        const TileMeta m0 = load_meta(A, t);
        stage_idx(A, m0, tid, si);
        stage_data<KIND>(A, t, m0, tid, si, d);
        stage_commit(d, S0, tid, te);
    }
    TileMeta m1 = load_meta(A, min(t + bx, tlast));
    stage_idx(A, m1, tid, si);
    TileMeta m2 = load_meta(A, min(t + 2 * bx, tlast));
    TileMeta c = d.mt;  // the tile in the current buffer
    bool odd = false;
    __syncthreads();

    // One tile; returns false after the block's last tile.  The first call is peeled
    // off the loop so that the loop header merges two identical states and the
    // compiler's wait counts for the in-flight prefetch stay exact.
    auto tile = [&]() __attribute__((always_inline)) -> bool {
        const StageLds &S = odd ? S1 : S0;
        const StageLds &Sn = odd ? S0 : S1;
        const int64_t t1 = t + bx;
        // prefetch (data of t1, slots of t2, scalars of t3)
        stage_data<KIND>(A, min(t1, tlast), m1, tid, si, d);
        stage_idx(A, m2, tid, si);
        m1 = m2;
        m2 = load_meta(A, min(t + 3 * bx, tlast));
        // keep the scheduler from sinking the prefetch towards its use
        __builtin_amdgcn_sched_barrier(0);
        // B: element (every lane computes; the stores of lanes past the last element are dropped)
        const bool has_elem = tid < c.nel;
        const int64_t e = (int64_t)c.e0 + tid;
        {
            const ushort4 q = S.lc[tid];
            const int lc[4] = {q.x, q.y, q.z, q.w};
            MatParams m = A.mat;
            m.lambda = S.LAM[tid];
            m.mu = S.MU[tid];
            double X[4][3], U[4][3], fe[4][3];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ld3(S.X, lc[k], X[k]);
                ld3(S.U, lc[k], U[k]);
            }
            static_assert(state_width<KIND>() <= 1, "the pipelined kernel keeps one state value per element");
            double H[9], P[9], snp1 = 0.0, aux = 0.0;
            tet4_element<KIND>(X, U, m, S.SN[tid], snp1, aux, H, P, fe);
#pragma unroll
            for (int k = 0; k < 4; ++k) st3(F, tid * 4 + k, fe[k]);
            // commit the prefetched tile before this tile's stores are issued
            stage_commit(d, Sn, tid, te);
            const uint32_t oe = has_elem ? (uint32_t)(8 * e) : kDropOffset;
            const uint32_t oi = (uint32_t)(8 * A.ne);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                bstore<2>(rH, oe + i * oi, H[i]);
                bstore<2>(rP, oe + i * oi, P[i]);
            }
            if constexpr (KIND == APHIP_J2) bstore<2>(rS, oe, snp1);
        }
        __syncthreads();

        // C: gather and finish (same order as k_tile).  Stores are unconditional; those
        // of lanes past the tile's nodes, and the u, v, a of a shared node (finished by
        // the shared-node pass), are dropped.
        {
            const bool live = tid < c.tn, interior = tid < c.nint;
            const int ln = min(tid, c.tn - 1);
            const uint32_t pk = S.csr[ln];
            const int st = (int)(pk >> 16), cnt = (int)(pk & 0xffffu);
            const uint16_t *ent = reinterpret_cast<const uint16_t *>(S.ent);
            double f[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < cnt; ++j) {
                const int s = ent[st + j];
                f[0] += F[3 * s];
                f[1] += F[3 * s + 1];
                f[2] += F[3 * s + 2];
            }
            double vh[3], u1[3], a1[3], v1[3];
            ld3(S.V, ln, vh);
            ld3(S.U, ln, u1);
            cd_correct(f, S.M[ln], S.fx[ln], S.DT[ln], vh, a1, v1);
            const int64_t g = (int64_t)c.ibase + tid;
            const uint32_t og = (live && interior) ? (uint32_t)(24 * g) : kDropOffset;
            const uint32_t ob = (live && !interior) ? (uint32_t)(24 * (int64_t)S.bdst[ln]) : kDropOffset;
            bstore3(rF, og, f);
            bstore3(rB, ob, f);
            bstore3(rU, og, u1);
            bstore3(rV, og, v1);
            bstore3(rA, og, a1);
        }
        if (t1 >= tend) return false;
        c = d.mt;
        odd = !odd;
        t = t1;
        __syncthreads();
        return true;
    };
    if (tile())
        while (tile()) {
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
    const double *zero;  // 8 zero doubles: stands in for an absent optional array
    double dt;
    double *upred;  // mode 2: store the next step's predicted displacement (same dt), or null
};

template <int MODE>
__global__ __launch_bounds__(256) void k_boundary(BoundArgs A) {
    const int64_t b_raw = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool live = b_raw < A.nb;
    const int64_t b = live ? b_raw : A.nb - 1;  // idle lanes redo the last node and store nothing
    const int64_t g = A.first + b;
    const int32_t s0 = A.bstart[b], s1 = A.bstart[b + 1];
    // node data first: independent of the slot sums, so its loads overlap them
    double u[3], v[3], a[3], m = 1.0, dt = A.dt;
    unsigned fx = 0u;
    if constexpr (MODE >= 1) m = A.mass[g];
    if constexpr (MODE == 2) {
        ld3(A.u, g, u);
        ld3(A.v, g, v);
        ld3(A.a, g, a);
        const uint8_t *pf = A.fixed ? A.fixed : reinterpret_cast<const uint8_t *>(A.zero);
        const double *pd = A.node_dt ? A.node_dt : A.zero;
        const unsigned f8 = pf[A.fixed ? g : 0];
        const double d8 = pd[A.node_dt ? g : 0];
        fx = A.fixed ? f8 : 0u;
        dt = A.node_dt ? d8 : A.dt;
    }
    // slot sums in ascending tile order, four slots' loads at a time (clamped; an
    // absent slot adds zero)
    double f[3] = {0.0, 0.0, 0.0};
    for (int32_t sb = s0; sb < s1; sb += 4) {
        double p[4][3];
#pragma unroll
        for (int r = 0; r < 4; ++r) ld3(A.bpart, min(sb + r, s1 - 1), p[r]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const bool in = sb + r < s1;
#pragma unroll
            for (int d = 0; d < 3; ++d) f[d] += in ? p[r][d] : 0.0;
        }
    }
    if (!live) return;
    st3(A.f, g, f);
    if constexpr (MODE == 1) {
        const double acc[3] = {f[0] / m, f[1] / m, f[2] / m};
        st3(A.a, g, acc);
    } else if constexpr (MODE == 2) {
        double vh[3], u1[3], a1[3], v1[3];
        cd_predict(u, v, a, fx, dt, vh, u1);
        cd_correct(f, m, fx, dt, vh, a1, v1);
        st3(A.u, g, u1);
        st3(A.v, g, v1);
        st3(A.a, g, a1);
        if (A.upred) {  // the predictor of the next step from the values just stored
            double vh2[3], u2[3];
            cd_predict(u1, v1, a1, fx, dt, vh2, u2);
            st3(A.upred, b, u2);
        }
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

template <int KIND>
void launch_stream(Ctx &c, const TileArgs &A, size_t) {
    auto kern = k_tile_stream<KIND>;
    TilePlan &P = c.tiles;
    const size_t lds = stream_lds_bytes(P.te, P.maxtn);
    int &grid = P.stream_grid[KIND];
    if (grid == 0) {
        if (lds > 65536)
            APHIP_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int per_cu = 0, cus = 0;
        APHIP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, P.threads, lds));
        APHIP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
        const int64_t want = std::max<int64_t>(1, std::min<int64_t>((P.ntiles + kNumXcd - 1) / kNumXcd,
                                                                   (int64_t)std::max(per_cu, 1) * cus / kNumXcd));
        grid = (int)(want * kNumXcd);
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(P.threads), lds, c.stream, A, grid);
}

bool stream_eligible(const Ctx &c) {
    const TilePlan &P = c.tiles;
    // buffer-store offsets are 32-bit and must stay below the drop offset (2 GiB)
    const int64_t lim = (int64_t)1 << 31;
    return P.use_stream && c.npe == 4 && P.te == P.threads && P.threads <= 256 && P.maxtn <= P.threads &&
           stream_lds_bytes(P.te, P.maxtn) <= 160 * 1024 && 72 * c.ne < lim && 24 * c.nn < lim &&
           24 * P.nslot < lim;
}

template <int NPE, int MODE>
void dispatch_kind(Ctx &c, const TileArgs &A, size_t lds) {
    if constexpr (NPE == 4 && MODE == 2) {
        if (stream_eligible(c)) {
            switch (c.mat.kind) {
                case APHIP_LINEAR_ELASTIC: launch_stream<APHIP_LINEAR_ELASTIC>(c, A, lds); return;
                case APHIP_NEO_HOOKEAN: launch_stream<APHIP_NEO_HOOKEAN>(c, A, lds); return;
                case APHIP_J2: launch_stream<APHIP_J2>(c, A, lds); return;
                case APHIP_SVK: launch_stream<APHIP_SVK>(c, A, lds); return;
                default: break;
            }
        }
    }
    const bool pq = c.mat.state_per_qp && NPE == 8;
    switch (c.mat.kind) {
        case APHIP_LINEAR_ELASTIC: launch_tile<NPE, APHIP_LINEAR_ELASTIC, false, MODE>(c, A, lds); break;
        case APHIP_NEO_HOOKEAN: launch_tile<NPE, APHIP_NEO_HOOKEAN, false, MODE>(c, A, lds); break;
        case APHIP_J2:
            if (pq) launch_tile<NPE, APHIP_J2, true, MODE>(c, A, lds);
            else launch_tile<NPE, APHIP_J2, false, MODE>(c, A, lds);
            break;
        case APHIP_SVK: launch_tile<NPE, APHIP_SVK, false, MODE>(c, A, lds); break;
        case APHIP_NEO_HOOKEAN_DAMAGE:
            if (pq) launch_tile<NPE, APHIP_NEO_HOOKEAN_DAMAGE, true, MODE>(c, A, lds);
            else launch_tile<NPE, APHIP_NEO_HOOKEAN_DAMAGE, false, MODE>(c, A, lds);
            break;
        case APHIP_POWER_LAW_CREEP:
            if (pq) launch_tile<NPE, APHIP_POWER_LAW_CREEP, true, MODE>(c, A, lds);
            else launch_tile<NPE, APHIP_POWER_LAW_CREEP, false, MODE>(c, A, lds);
            break;
        case APHIP_DRUCKER_PRAGER:
            if (pq) launch_tile<NPE, APHIP_DRUCKER_PRAGER, true, MODE>(c, A, lds);
            else launch_tile<NPE, APHIP_DRUCKER_PRAGER, false, MODE>(c, A, lds);
            break;
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
        APHIP_REQUIRE(te == threads && (threads == 64 || threads == 128 || threads == 256 || threads == 512 ||
                                        threads == 1024),
                      "Tet4 tiles: one element per thread, 64, 128, 256, 512 or 1024 threads");
    } else {
        const bool per_elem = te == threads && (threads == 64 || threads == 128 || threads == 256);
        const bool per_qp = threads == 8 * te && (te == 8 || te == 16 || te == 32 || te == 64);
        APHIP_REQUIRE(per_elem || per_qp,
                      "Hex8 tiles: one element per thread (64, 128 or 256) or eight lanes per element (8..64 elements)");
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
    {  // fixed-stride copies for the tile kernel
        P.tpad = std::max(maxtn, 1);
        std::vector<int32_t> pn((size_t)nt * P.tpad), pb((size_t)nt * P.tpad);
        std::vector<uint32_t> pc((size_t)nt * P.tpad);
        std::vector<uint16_t> pl((size_t)nt * te * npe);
        for (int64_t t = 0; t < nt; ++t) {
            const int tn = tstart[t + 1] - tstart[t];
            for (int r = 0; r < P.tpad; ++r) {
                const int32_t i = tstart[t] + std::min(r, tn - 1);
                pn[(size_t)t * P.tpad + r] = tnodes[i];
                pc[(size_t)t * P.tpad + r] = tcsr[i];
                pb[(size_t)t * P.tpad + r] = tbdst[i];
            }
            const int64_t e0 = te0v[t], nel = te0v[t + 1] - e0;
            for (int le = 0; le < te; ++le)
                for (int k = 0; k < npe; ++k)
                    pl[((size_t)t * te + le) * npe + k] = lconn[(e0 + std::min<int64_t>(le, nel - 1)) * npe + k];
        }
        P.tnp_node.upload(pn.data(), pn.size(), c.stream);
        P.tnp_csr.upload(pc.data(), pc.size(), c.stream);
        P.tnp_bdst.upload(pb.data(), pb.size(), c.stream);
        P.lconn_p.upload(pl.data(), pl.size(), c.stream);
    }
    P.nslot = bstart[nb];
    P.bpart.alloc(3 * std::max<int64_t>(1, (int64_t)bstart[nb]));
    P.upred.alloc(3 * std::max<int64_t>(1, nb));
    P.bstart.upload(bstart.data(), bstart.size(), c.stream);
    P.zero.alloc(8);
    P.zero.zero(c.stream);
    {
        std::vector<int32_t> meta((size_t)nt * 8, 0);
        for (int64_t t = 0; t < nt; ++t) {
            int32_t *m = &meta[(size_t)t * 8];
            m[0] = tstart[t];
            m[1] = tstart[t + 1] - tstart[t];
            m[2] = ninterior[t];
            m[3] = te0v[t + 1] - te0v[t];
            m[4] = int_base[t];
            m[5] = te0v[t];
        }
        P.meta.upload(meta.data(), meta.size(), c.stream);
    }
    // LDS: X, U, V (3 doubles each) + mass + increment per node, partials, csr/bdst per node, entries, flags
    int max_nint = 0;
    for (int64_t t = 0; t < nt; ++t) max_nint = std::max(max_nint, ninterior[t]);
    P.max_nint = npe == 4 ? maxtn : max_nint;  // LDS rows of the corrector data (the kernel's choice)
    // One-element-per-thread Hex8 tiles pass their forces through LDS in two rounds of
    // te / 2 elements: half the F rows, so twice the workgroups fit per CU.
    P.f_rounds = (npe == 8 && te == threads) ? 2 : 1;  // the kernel's ROUNDS
    P.lds_bytes = tile_lds_bytes(P.maxtn, P.max_nint, te, P.f_rounds, npe);  // the kernel carves with these
    APHIP_REQUIRE(P.lds_bytes <= 160 * 1024, "tile needs more LDS than a CU has; use a smaller tile");
    {  // the kernel's carve-up must end inside the allocation
        const TileLds L = carve(nullptr, P.maxtn, P.max_nint, te, P.f_rounds, npe);
        APHIP_REQUIRE((size_t)(reinterpret_cast<const char *>(L.fx + P.max_nint) - static_cast<const char *>(nullptr)) <=
                          P.lds_bytes,
                      "tile LDS layout exceeds its allocation");
    }
    APHIP_REQUIRE(maxtn <= npe * threads, "tile has more nodes than its gather covers");
    // the staging loads te * npe / 2 CSR words with at most 2 (Tet4), 4 (Hex8, one element
    // per thread) or 1 (Hex8, eight lanes per element) per thread
    APHIP_REQUIRE(te * npe / 2 <= (npe == 4 ? 2 : (threads == te ? 4 : 1)) * threads, "tile CSR exceeds the staging");
    P.use_stream = (c.flags & APHIP_FLAG_PIPELINED) != 0;
    for (int &g : P.stream_grid) g = 0;  // sized (and LDS attribute set) at the first launch
    P.ready = true;
    APHIP_HIP(hipStreamSynchronize(c.stream));
}

void tile_run(Ctx &c, double dt, int mode, const double *u_in, double *f_out, bool pred_in, bool pred_out) {
    TilePlan &P = c.tiles;
    APHIP_REQUIRE(P.ready, "tile plan not built");
    APHIP_REQUIRE(mode == 0 || (u_in == nullptr && f_out == nullptr), "displacement / force arrays only in force mode");
    APHIP_REQUIRE(mode == 2 || !(pred_in || pred_out), "predicted displacements only in step mode");
    pred_in = pred_in && P.nb > 0;
    pred_out = pred_out && P.nb > 0;
    if (c.ne == 0) return;
    TileArgs A;
    A.X = c.X.p;
    A.u = u_in ? const_cast<double *>(u_in) : c.u.p;
    A.v = c.v.p;
    A.a = c.a.p;
    A.f = f_out ? f_out : c.f.p;
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
    A.tnp_node = P.tnp_node.p;
    A.tnp_csr = P.tnp_csr.p;
    A.tnp_bdst = P.tnp_bdst.p;
    A.lconn_p = P.lconn_p.p;
    A.tpad = P.tpad;
    A.bpart = P.bpart.p;
    A.H = c.Hnp1().p;
    A.P = c.Pnp1().p;
    A.Hn = c.mat.rate_form() ? c.Hn().p : nullptr;
    A.Pn = c.mat.rate_form() ? c.Pn().p : nullptr;
    A.rate_dt = mode == 2 ? dt : c.rate_dt;
    A.state_n = c.state[1 - c.state_np1].p;
    A.state_np1 = c.state[c.state_np1].p;
    A.mat = c.mat.params();
    A.elem_lambda = c.has_fields ? c.elem_lambda.p : nullptr;
    A.elem_mu = c.has_fields ? c.elem_mu.p : nullptr;
    A.node_dt = c.node_dt.p;
    A.zero = P.zero.p;
    A.meta = P.meta.p;
    A.gauss_a = 1.0 / std::sqrt(3.0);
    A.ne = c.ne;
    A.nn = c.nn;
    A.nslot = P.nslot;
    A.ntiles = P.ntiles;
    A.nblocks = padded(P.ntiles);
    A.te = P.te;
    A.maxtn = P.maxtn;
    A.max_nint = P.max_nint;
    A.dt = dt;
    A.upred = pred_in ? P.upred.p : nullptr;
    A.first_shared = P.n_interior;
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
        BoundArgs B{P.nb, P.n_interior, P.bstart.p, P.bpart.p, c.u.p, c.v.p, c.a.p, A.f, c.mass.p,
                    c.has_fixed ? c.fixed.p : nullptr, c.node_dt.p, P.zero.p, dt,
                    pred_out ? P.upred.p : nullptr};
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
