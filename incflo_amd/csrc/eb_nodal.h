// eb_nodal.h -- embedded-boundary nodal multigrid (FP64): geometry tables per
// level, Galerkin coarse cut elements, V-cycle and MLMG outer loop.
#pragma once

#include <vector>

#include "eb_kernels.h"
#include "mgx_common.h"
#include "nodal_mg.h"

namespace mgx {

// Fluid moments of a z-aligned cylinder (fluid outside) for every cell of a
// level: type 0 covered / 1 regular / 2 cut and 27 moments
// m[p + 3q + 9r] = int over the fluid part of the unit cell of xi^p eta^q zeta^r.
// Computed from the boundary integral (Green's theorem) of the covered part:
// Gauss-Legendre along the circular arcs, exact along the cell edges.
void eb_cylinder_moments(const NodalGeom& g, const double prob_lo[3], double cx, double cy,
                         double r, std::vector<int>& type, std::vector<double>& mom);

// Body-surface weights of the same cylinder, 24 per cell ([corner][d], zero
// outside cut cells): (1/V) int over the surface in the cell of N_p n_d dA with
// n out of the fluid, for the EB inflow velocity term of the divergence.
void eb_cylinder_surface_weights(const NodalGeom& g, const double prob_lo[3], double cx,
                                 double cy, double r, const std::vector<int>& type,
                                 std::vector<double>& B);

// Face apertures (fluid fraction of each face, canonical face arrays with
// n_d + 1 faces along d) and covered-cell flags of the same cylinder.
void eb_cylinder_apertures(const int n[3], const double prob_lo[3], const double dx[3], double cx,
                           double cy, double r, std::vector<double> ap[3],
                           std::vector<unsigned char>& covered);

// FD-normalised element matrix of a cell from its fluid moments (sigma excluded)
void eb_element_from_moments(const double* mom, const double dx[3], double E[64]);
// regular element of spacing dx (same normalisation)
void eb_regular_element(const double dx[3], double E[64]);

struct EBLevelData {
    NodalGeom g;
    std::vector<int> cidx_h;        // -1 regular, -2 covered, >= 0 cut index
    std::vector<int> cut_cells;     // linear cell index of every cut cell
    std::vector<int> children;      // coarse levels: 8 fine cells per cut cell
    std::vector<unsigned char> mask_h;
    DevBuf cidx, cutc, child, mask, K, Egeom, W, V;
    DevBuf sig_buf, b_buf, x_buf, x2_buf, r_buf;
    double E[64];
    int ncut = 0;
};

class EBNodal {
   public:
    std::vector<EBLevelData> L;
    DevBuf bottom_work;
    bool ready = false;
    // level 0 geometry from per-cell types and moments (host arrays, n0*n1*n2 cells)
    void setup(const NodalGeom& g0, int nlev, const std::vector<int>& type,
               const std::vector<double>& mom, hipStream_t s);
    // sigma of level 0 (ghosted, device) -> coarse sigma and all cut elements
    void set_sigma(const double* sig0, hipStream_t s);
    EBLev lev(int l) const;
    // MLMG solve of L phi = b (b modified: inactive rows zeroed, mean removed when singular)
    void solve(double* b, double* phi_a, double* phi_b, double*& phi, double rtol, double atol,
               const MGOpts& o, DevBuf& scal, DevBuf& part, unsigned long long* h_pinned,
               mgx_solve_info* info, hipStream_t s);
    const unsigned char* mask0() const { return L[0].mask.as<unsigned char>(); }
    const int* cidx0() const { return L[0].cidx.as<int>(); }
    const double* W0() const { return L[0].W.as<double>(); }
    const double* V0() const { return L[0].V.as<double>(); }
    const double* sig0 = nullptr;

   private:
    void vcycle(const MGOpts& o, hipStream_t s);
    void smooth(int l, int sweeps, bool from_zero, const MGOpts& o, hipStream_t s);
};

}  // namespace mgx
