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
    int dlo[3], dhi[3];  // Dirichlet-form sides (Dirichlet, or Robin with scaled beta)
    int rlo[3], rhi[3];  // of those, the Robin sides
    bool robin_fixes;    // some Robin face has a > 0
    bool has_a;          // Helmholtz form: a cell coefficient A > 0 (never singular)
    double dx[3];
    long ncell() const { return (long)n[0] * n[1] * n[2]; }
    long nface(int d) const {
        return (long)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
    }
    bool singular() const {
        if (has_a) return false;
        for (int d = 0; d < 3; ++d)
            if (!per[d] && ((dlo[d] && !rlo[d]) || (dhi[d] && !rhi[d]))) return false;
        return !robin_fixes;
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
    DevBuf beta_buf[3], a_buf, b_buf, x_buf, r_buf;
    T* beta[3] = {nullptr, nullptr, nullptr};
    T* a = nullptr;  // Helmholtz cell coefficient (null: none)
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
    // level-0 x holds an initial guess (the correction so far): the next cycle
    // smooths from it instead of from zero
    bool x0_guess = false;
    // on (default): each V-cycle is captured once per (level-0 x and b,
    // initial guess, options, bottom counter) as a HIP graph and replayed,
    // which removes the host launch gaps between the many small kernels of
    // the coarse levels; off: the same kernels launched one by one
    bool graphs = true;
    MacHierarchy() = default;
    MacHierarchy(const MacHierarchy&) = delete;
    MacHierarchy& operator=(const MacHierarchy&) = delete;
    ~MacHierarchy();
    void drop_graphs();
    void build(const MacGeom& g0, int nlev);
    // level-0 beta from FP64 face arrays, then coarsen; with Robin sides the
    // boundary planes of each level are then scaled by rscale(l, d, side)
    // (null: not a Robin side)
    void set_beta(const double* const beta0[3], hipStream_t s,
                  const std::vector<const double*>* rscale = nullptr);
    // level-0 Helmholtz coefficient A from an FP64 cell array; coarse levels
    // take the mean of their 8 fine cells (needs g0.has_a)
    void set_acoef(const double* a0, hipStream_t s);
    void vcycle(const MGOpts& o, hipStream_t s, int* bottom_iters);

   private:
    struct CycleGraph {
        T* x0;
        const T* b0;
        bool guess0;
        int* bottom_iters;
        long coarse;  // mac_coarse_cells() at capture
        MGOpts o;
        hipGraphExec_t exec;
    };
    std::vector<CycleGraph> cache_;
    hipStream_t cap_stream_ = nullptr;  // capture stream (the caller's may be the null stream)
    void vcycle_body(const MGOpts& o, hipStream_t s, int* bottom_iters);
};

class MacSolver {
   public:
    MacGeom g0;
    MGOpts opts;
    int nlev = 1;
    MacHierarchy<double> H64;
    MacHierarchy<float> H32;
    DevBuf beta0[3];      // level-0 face coefficients (FP64)
    DevBuf acoef0;        // level-0 Helmholtz cell coefficient A (FP64; g0.has_a)
    bool acoef_set = false, h32_a = false, h64_a = false;
    double* acoef() { return g0.has_a ? acoef0.as<double>() : nullptr; }
    void acoef_changed() { acoef_set = true; h32_a = false; h64_a = false; }
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
    void beta_changed() { beta_set = true; h32_beta = false; h64_beta = false; beta0s_ok = false; }
    // Robin data changed whether the problem is singular: rebuild the levels
    void set_robin_fixes(bool v) {
        g0.robin_fixes = v;
        H64.built = H32.built = false;
    }
    // Robin sides: level-0 a and b per boundary face (x fastest over the two
    // remaining directions).  Each level's boundary coefficient is its
    // face-averaged beta times a_l h_l / (a_l h_l + 2 b_l), with a_l, b_l the
    // face averages of a, b (the Robin relation on that level's mesh).
    void set_robin(int d, int side, std::vector<double> a, std::vector<double> b);
    void clear_robin();
    // level-0 beta with the Robin planes scaled (the operator's face coefficient)
    const double* fine_beta(int d, hipStream_t s);

   private:
    bool robin_on = false;
    std::vector<double> rob_a[3][2], rob_b[3][2];
    std::vector<DevBuf> rob_scale[3][2];  // per level
    std::vector<const double*> rob_ptr;   // [l * 6 + 2 d + side]
    int rob_levels = 0;                   // levels rob_scale was built for
    DevBuf beta0s[3];
    bool beta0s_ok = false;
    void robin_prepare(hipStream_t s);

   public:
    // Solve L phi = b (b modified in place: mean removed when singular);
    // initial guess this->phi.
    void solve(double* b, double rtol, double atol, int mode, mgx_solve_info* info, hipStream_t s);
    double read_norm(int slot, hipStream_t s);
};

}  // namespace mgx
