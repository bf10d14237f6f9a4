// mac_mg.h -- cell-centred multigrid for the MAC projection.
#pragma once

#include <vector>

#include "mac_kernels.h"
#include "mgx_common.h"
#include "nodal_mg.h"

namespace mgx {

struct MacGeom {
    int n[3];
    int per[3];
    int dlo[3], dhi[3];
    double dx[3];
    long ncell() const { return (long)n[0] * n[1] * n[2]; }
    long nface(int d) const {
        return (long)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
    }
    bool singular() const {
        for (int d = 0; d < 3; ++d)
            if (!per[d] && (dlo[d] || dhi[d])) return false;
        return true;
    }
    MGeomK k() const {
        MGeomK g{};
        for (int d = 0; d < 3; ++d) {
            g.n[d] = n[d];
            g.per[d] = per[d];
            g.dlo[d] = dlo[d];
            g.dhi[d] = dhi[d];
        }
        return g;
    }
    MacGeom coarse() const {
        MacGeom c = *this;
        for (int d = 0; d < 3; ++d) { c.n[d] = n[d] / 2; c.dx[d] = 2 * dx[d]; }
        return c;
    }
};

template <class T>
inline MCoef<T> make_mcoef(const double dx[3]) {
    MCoef<T> c;
    for (int d = 0; d < 3; ++d) c.ih2[d] = (T)(1.0 / (dx[d] * dx[d]));
    return c;
}

template <class T>
struct MacLevel {
    MacGeom g;
    MGeomK kg;
    MCoef<T> c;
    DevBuf beta_buf[3], b_buf, x_buf, r_buf;
    T* beta[3] = {nullptr, nullptr, nullptr};
    T* b = nullptr;
    T* x = nullptr;
    T* r = nullptr;
};

template <class T>
class MacHierarchy {
   public:
    std::vector<MacLevel<T>> L;
    DevBuf bottom_work;
    bool built = false;
    void build(const MacGeom& g0, int nlev);
    // level-0 beta from FP64 face arrays, then coarsen
    void set_beta(const double* const beta0[3], hipStream_t s);
    void vcycle(const MGOpts& o, hipStream_t s, int* bottom_iters);
};

class MacSolver {
   public:
    MacGeom g0;
    MGOpts opts;
    int nlev = 1;
    MacHierarchy<double> H64;
    MacHierarchy<float> H32;
    DevBuf beta0[3];      // level-0 face coefficients (FP64)
    DevBuf phi_a, phi_b;  // outer iterate
    DevBuf res64, res32;
    DevBuf scal, part, ibuf;
    unsigned long long* h_pinned = nullptr;
    double* phi = nullptr;
    bool beta_set = false;
    bool h32_beta = false, h64_beta = false;

    ~MacSolver();
    void setup(const MacGeom& g, const MGOpts& o);
    double* beta(int d) { return beta0[d].as<double>(); }
    void beta_changed() { beta_set = true; h32_beta = false; h64_beta = false; }
    // Solve L phi = b (b modified in place: mean removed when singular);
    // initial guess this->phi.
    void solve(double* b, double rtol, double atol, int mode, mgx_solve_info* info, hipStream_t s);
    double read_norm(int slot, hipStream_t s);
};

}  // namespace mgx
