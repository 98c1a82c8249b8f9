// Internal context and helpers of the aperi-hip library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "aperi_hip.h"

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
    int state_per_qp = 0;
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
    DevBuf<int64_t> start;      // [ne + 1]
    DevBuf<int32_t> nodes;      // [nnz] internal node ids
    DevBuf<double> dphi;        // [nnz * 3]
    DevBuf<double> volume;      // [ne]
    DevBuf<int64_t> node_start; // [nn + 1] node -> (cell entry) incidences
    DevBuf<int64_t> node_inc;   // [nnz] entry index t (cell found by binary search in start)
    DevBuf<int32_t> inc_cell;   // [nnz] owning cell of each incidence
    void reset() {
        ready = false;
        nnz = n_truncated = 0;
        num_nbr.release(); nbr.release(); phi.release(); radius.release(); start.release();
        nodes.release(); dphi.release(); volume.release(); node_start.release(); node_inc.release();
        inc_cell.release();
    }
};

struct Profile {
    bool enabled = false;
    std::vector<hipEvent_t> ev;   // pairs (start, stop) tagged by which
    std::vector<int> tag;
    int64_t launches[3] = {0, 0, 0};
    double ms[3] = {0, 0, 0};
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;

    // mesh
    int etype = -1;
    int npe = 0;
    int64_t nn = 0, ne = 0;
    std::vector<int32_t> node_perm;   // caller -> internal
    std::vector<int32_t> node_iperm;  // internal -> caller
    std::vector<int32_t> elem_perm;   // caller -> internal
    std::vector<int32_t> elem_iperm;  // internal -> caller

    DevBuf<double> X, u, v, a, f, mass;  // node vectors AoS n x 3 (mass n)
    DevBuf<uint8_t> fixed;               // bit d set = DOF d held at zero
    DevBuf<int32_t> conn;                // AoS ne x npe, internal node ids
    DevBuf<double> H, P;                 // SoA 9 x ne
    DevBuf<double> state[2];             // J2 equivalent plastic strain (N / NP1)
    int state_np1 = 1;                   // index of the NP1 state buffer
    DevBuf<double> volume;               // ne
    DevBuf<int32_t> node_start;          // nn + 1
    DevBuf<int32_t> node_inc;            // ne * npe: element * npe + slot, ascending element
    DevBuf<double> partial;              // ne * npe * 3 element force partials

    MaterialConst mat;
    bool has_fields = false;             // per-element lambda / mu
    DevBuf<double> elem_lambda, elem_mu;
    bool has_mass = false;
    bool has_fixed = false;

    CellData cells;
    Profile prof;

    int64_t state_count() const {
        if (mat.kind != APHIP_J2) return 0;
        return ne * ((etype == APHIP_HEX8 && mat.state_per_qp) ? 8 : 1);
    }
};

void set_last_error(const std::string &msg);

// ---- kernels (fem.hip) ----
void fem_element_force(Ctx &c);
void fem_predict(Ctx &c, double dt);
// mode 0: f only; 1: f and a = f/m (initial); 2: f, a, BC, v += dt/2 a (step corrector)
void fem_node_update(Ctx &c, double dt, int mode);
void fem_volumes(Ctx &c);
void fem_lumped_mass(Ctx &c, double rho);

// ---- RK (rk.hip) ----
void rk_setup(Ctx &c, double scale);
void rk_cell_force(Ctx &c);
void rk_node_update(Ctx &c, double dt, int mode);
void rk_mass(Ctx &c, double *total);

// ---- profiling helpers (capi.cpp) ----
void prof_begin(Ctx &c, int which);
void prof_end(Ctx &c, int which);

}  // namespace aphip
