// mgx_common.h -- shared host/device definitions for the mgx HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "mgx.h"

namespace mgx {

// Errors never cross the C ABI as exceptions: every entry point catches
// mgx::Error and turns it into an MGX_ERR_* status plus a message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MGX_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::mgx::Error(MGX_ERR_HIP, std::string(#call " failed: ") +           \
                                                hipGetErrorString(e_) + " at " +       \
                                                __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

#define MGX_REQUIRE(cond, code, msg)                       \
    do {                                                   \
        if (!(cond)) throw ::mgx::Error((code), (msg));    \
    } while (0)

// Device buffer owned by the library (caller-owned fields are raw pointers).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) return;
        MGX_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Geometry of one multigrid level of the nodal problem.
//   cells n[d]; nodes nn[d] = periodic ? n : n+1; cell coefficient stored
//   with one ghost layer: sg[d] = n[d] + 2 (ghosts: periodic image or 0).
struct NodalGeom {
    int n[3];
    int nn[3];
    int per[3];
    int dlo[3], dhi[3];  // Dirichlet faces (nodes fixed at 0)
    double dx[3];
    int sg[3];
    long nnode() const { return (long)nn[0] * nn[1] * nn[2]; }
    long nsig() const { return (long)sg[0] * sg[1] * sg[2]; }
    bool singular() const {
        for (int d = 0; d < 3; ++d)
            if (!per[d] && (dlo[d] || dhi[d])) return false;
        return true;
    }
    bool iso() const { return dx[0] == dx[1] && dx[1] == dx[2]; }
};

inline NodalGeom make_nodal_geom(const int n[3], const int per[3], const int bclo[3],
                                 const int bchi[3], const double dx[3]) {
    NodalGeom g{};
    for (int d = 0; d < 3; ++d) {
        g.n[d] = n[d];
        g.per[d] = per[d] ? 1 : 0;
        g.nn[d] = per[d] ? n[d] : n[d] + 1;
        g.dlo[d] = (!per[d] && bclo[d] == MGX_BC_DIRICHLET) ? 1 : 0;
        g.dhi[d] = (!per[d] && bchi[d] == MGX_BC_DIRICHLET) ? 1 : 0;
        g.dx[d] = dx[d];
        g.sg[d] = n[d] + 2;
    }
    return g;
}

inline NodalGeom coarsen(const NodalGeom& f) {
    NodalGeom c = f;
    for (int d = 0; d < 3; ++d) {
        c.n[d] = f.n[d] / 2;
        c.nn[d] = f.per[d] ? c.n[d] : c.n[d] + 1;
        c.dx[d] = 2.0 * f.dx[d];
        c.sg[d] = c.n[d] + 2;
    }
    return c;
}

// Weighted-Jacobi factor for a level: omega is meant for isotropic cells, where
// the Jacobi-preconditioned constant-sigma Q1 symbol peaks at 3/2; on
// anisotropic cells the peak (the symbol at the corners of [0,pi]^3) is
//   lmax = max(4 max ih2, 4/3 max pair sum, 4/9 S) / (8/9 S)
// and omega is scaled by min(1, 1.5/lmax).
inline double nodal_omega(const NodalGeom& g, double omega) {
    double ih2[3], S = 0.0;
    for (int d = 0; d < 3; ++d) { ih2[d] = 1.0 / (g.dx[d] * g.dx[d]); S += ih2[d]; }
    const double m1 = std::max(ih2[0], std::max(ih2[1], ih2[2]));
    const double m2 = std::max(ih2[0] + ih2[1], std::max(ih2[0] + ih2[2], ih2[1] + ih2[2]));
    const double lmax =
        std::max(4.0 * m1, std::max(4.0 / 3.0 * m2, 4.0 / 9.0 * S)) / (8.0 / 9.0 * S);
    return omega * std::min(1.0, 1.5 / lmax);
}

// Coarsen while every cell count is even and the coarse count stays >= 4.
inline int nodal_nlevels(const NodalGeom& g, int max_coarsening_level) {
    int n[3] = {g.n[0], g.n[1], g.n[2]};
    int nl = 1;
    while (nl - 1 < max_coarsening_level) {
        bool ok = true;
        for (int d = 0; d < 3; ++d)
            if (n[d] % 2 != 0 || n[d] / 2 < 4) ok = false;
        if (!ok) break;
        for (int d = 0; d < 3; ++d) n[d] /= 2;
        ++nl;
    }
    return nl;
}

}  // namespace mgx
