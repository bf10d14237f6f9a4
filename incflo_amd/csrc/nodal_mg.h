// nodal_mg.h -- nodal multigrid hierarchy and the outer MLMG iteration.
#pragma once

#include <array>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "box_kernels.h"
#include "unet.h"
#include "mgx_common.h"
#include "nodal_kernels.h"

namespace mgx {

struct MGOpts {
    int maxiter = 100;
    int bottom_maxiter = 100;
    double bottom_rtol = 1e-4;
    int bottom_solver = MGX_BOTTOM_BICGCG;
    int max_coarsening_level = 100;
    int nu1 = 2, nu2 = 2;
    double omega = 8.0 / 9.0;
    int verbose = 0;
    int stall_iters = 3;
    int cycles_per_update = 3;  // most FP32 V-cycles per FP64 refinement step (mixed modes)
    int smoother = MGX_SMOOTHER_DEFAULT;
    int learned_steps = 1;      // refinement steps started from the network (-1: all)
};

// Chebyshev sweeps for a hierarchy of FP32 (fp32) or FP64 storage
inline bool use_chebyshev(const MGOpts& o, bool fp32) {
    return o.smoother == MGX_SMOOTHER_CHEBYSHEV || (o.smoother == MGX_SMOOTHER_DEFAULT && fp32);
}

// Relaxation factor of sweep q (0-based) of n smoothing sweeps: weighted
// Jacobi's omega, or the Chebyshev roots of degree n for the interval [0.5,
// 1.5] of the Jacobi-scaled 27-point operator (its high-frequency range and
// top at constant sigma on a cubic grid), smallest factor first; both scaled
// down on anisotropic grids as nodal_omega does.
inline double sweep_omega(const NodalGeom& g, const MGOpts& o, bool cheb, int q, int n) {
    if (!cheb || n < 1) return nodal_omega(g, o.omega);
    const double t = std::cos(3.14159265358979323846 * (2.0 * q + 1.0) / (2.0 * n));
    return nodal_omega(g, 1.0 / (1.0 + 0.5 * t));
}

// The C options struct -> MGOpts, validated the same way for every solver
// handle.  cycles_per_update 0 keeps the built-in default.
inline void to_opts(const mgx_mlmg_opts* o, MGOpts& m) {
    MGX_REQUIRE(o != nullptr, MGX_ERR_ARG, "null MLMG options");
    MGX_REQUIRE(o->maxiter >= 1 && o->bottom_maxiter >= 1, MGX_ERR_ARG, "maxiter must be >= 1");
    MGX_REQUIRE(o->nu1 >= 0 && o->nu2 >= 0 && o->nu1 + o->nu2 >= 1, MGX_ERR_ARG,
                "smoothing sweeps must be non-negative with nu1+nu2 >= 1");
    MGX_REQUIRE(o->omega > 0.0 && o->omega < 2.0, MGX_ERR_ARG, "omega must lie in (0,2)");
    MGX_REQUIRE(o->cycles_per_update >= 0 && o->cycles_per_update <= 8, MGX_ERR_ARG,
                "cycles_per_update must lie in 0..8");
    MGX_REQUIRE(o->bottom_solver != MGX_BOTTOM_HYPRE, MGX_ERR_UNSUPPORTED,
                "bottom_solver hypre is not available (bicgcg, bicgstab, cg, cgbicg, smoother)");
    MGX_REQUIRE(o->bottom_solver >= MGX_BOTTOM_DEFAULT && o->bottom_solver <= MGX_BOTTOM_SMOOTHER,
                MGX_ERR_ARG, "unknown bottom_solver");
    m.verbose = o->verbose;
    m.bottom_solver = o->bottom_solver == MGX_BOTTOM_DEFAULT ? MGOpts{}.bottom_solver
                                                             : o->bottom_solver;
    m.maxiter = o->maxiter;
    m.bottom_maxiter = o->bottom_maxiter;
    m.bottom_rtol = o->bottom_rtol;
    m.max_coarsening_level = o->max_coarsening_level < 0 ? 0 : o->max_coarsening_level;
    m.nu1 = o->nu1;
    m.nu2 = o->nu2;
    m.omega = o->omega;
    m.stall_iters = o->stall_iters;
    m.cycles_per_update = o->cycles_per_update == 0 ? MGOpts{}.cycles_per_update
                                                     : o->cycles_per_update;
    MGX_REQUIRE(o->smoother >= MGX_SMOOTHER_DEFAULT && o->smoother <= MGX_SMOOTHER_CHEBYSHEV,
                MGX_ERR_ARG, "unknown smoother");
    m.smoother = o->smoother;
    MGX_REQUIRE(o->learned_steps >= -1 && o->learned_steps <= 100, MGX_ERR_ARG,
                "learned_steps must lie in -1..100");
    m.learned_steps = o->learned_steps == 0 ? MGOpts{}.learned_steps : o->learned_steps;
}

inline KGeom kgeom(const NodalGeom& g) {
    KGeom k{};
    for (int d = 0; d < 3; ++d) {
        k.nn[d] = g.nn[d];
        k.per[d] = g.per[d];
        k.dlo[d] = g.dlo[d];
        k.dhi[d] = g.dhi[d];
        k.sg[d] = g.sg[d];
        k.n[d] = g.n[d];
    }
    return k;
}

template <class T>
struct MGLevel {
    NodalGeom g;
    KGeom kg;
    Coef<T> coef;
    bool iso = true;
    DevBuf sig_buf, b_buf, x_buf, x2_buf, r_buf;
    T* sig = nullptr;
    T* b = nullptr;
    T* x = nullptr;
    T* x2 = nullptr;
    T* r = nullptr;
};

// The nodes of a level that take part in a computation (an AMR fine level:
// the nodes in or on the boundary of the fine region) and, built on first use
// per tile shape, the lists of the launch tiles that hold any of them.
// Launches over such a level visit those tiles only; the rest of its arrays
// must hold finite values that nothing reads (zeroed once).
class TileSet {
   public:
    // the block of the point kernels (one plane each), and the stencil tiles:
    // 16 x 16 nodes (narrow: skip more) or 64 x 8 (wide: the dense kernel's
    // shape) per the tuning table's amr_tile_width(32), kKZ planes by default
    static constexpr int kTX = 16, kTY = 16, kKZ = 8;
    void build(const NodalGeom& g, const std::vector<unsigned char>& active);
    // the same over any dims[0] x dims[1] x dims[2] array of points (e.g. cells)
    void build(const int dims[3], const std::vector<unsigned char>& active);
    bool built() const { return nn_[0] > 0; }
    // stencil tiles of tx x ty nodes x kz planes on a gx x gy x gz tile grid
    // (linear indices; tx a multiple of 16); count 0: none
    const int* tiles(int tx, int ty, int kz, int gx, int gy, int gz, int& count);
    // kTX x kTY blocks of one plane for the point kernels: (bx, by, k, 0)
    const int4* blocks(int& count);

   private:
    int nn_[3] = {0, 0, 0};
    int g16_ = 0;
    std::vector<unsigned char> col_;  // any active node in (16-column block, row, plane)
    std::map<std::array<int, 3>, std::pair<DevBuf, int>> lists_;  // (tx, ty, kz) -> tiles
    DevBuf blk_;
    int nblk_ = -1;
};

// cell-centred V-cycles: the most cells of a level below the first that the
// one-workgroup coarse cycle takes (tuning table mac_coarse_cells; 0: none)
long mac_coarse_cells();

// Full weighting of a fine level's nodal array onto every node of the level
// below it (dense, z-streamed as the V-cycle's restriction; Dirichlet nodes 0)
template <class T>
void restrict_full(const NodalGeom& gf, const NodalGeom& gc, const T* rf, T* rc, hipStream_t s);

// Launch one 27-point stencil pass (Jacobi / residual / update+residual);
// ts (optional) restricts it to the tiles that hold active nodes.
template <class T, class TC, int MODE>
void launch_stencil(const NodalGeom& g, SArgs<T, TC> a, hipStream_t s, bool fine,
                    TileSet* ts = nullptr);

// Optional stream-ordered timing of selected launches (event pairs).
struct LaunchTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    double total_ms = 0.0;
    long long count = 0;
    void begin(hipStream_t s);
    void end(hipStream_t s);
    void reserve(size_t pairs);  // events for `pairs` more launches (before a capture)
    void collect();  // after the stream has been synchronised
    void reset() { total_ms = 0.0; count = 0; used = 0; }
    ~LaunchTimer();
};

template <class T>
class NodalHierarchy {
   public:
    std::vector<MGLevel<T>> L;
    DevBuf bottom_work;
    LaunchTimer* timer = nullptr;  // times the finest-level smoother sweeps
    void build(const NodalGeom& g0, int nlev);
    // level-0 sigma (ghosted, same layout) from a double array, then coarsen
    void set_sigma(const double* sig0, hipStream_t s);
    // V-cycle on the correction equation: input L[0].b, output L[0].x.
    void vcycle(const MGOpts& o, hipStream_t s, int* bottom_iters);
    // x_ready: L[lev].x already holds the first sweep from zero (written by the
    // pass that produced L[lev].b), so that sweep is skipped
    void smooth(int lev, int nsweeps, bool from_zero, const MGOpts& o, hipStream_t s,
                bool x_ready = false);
    // set by the outer iteration when its residual pass also wrote L[0].x
    bool x0_ready = false;
    // level-0 x holds an initial guess (the correction so far): the next cycle
    // pre-smooths from it instead of from zero
    bool x0_guess = false;
    // Chebyshev factors for the per-level sweeps (set by vcycle from the options)
    bool cheb = false;
    // optional overset mask per level (1 solve, 0 fixed at 0; empty: none)
    std::vector<const unsigned char*> masks;
    bool bottom_fixed = false;  // the bottom level's overset mask holds a node at zero
    const unsigned char* mask(int l) const { return masks.empty() ? nullptr : masks[l]; }
    bool bottom_singular() const { return L.back().g.singular() && !(mask(0) && bottom_fixed); }
    bool built = false;
    // Replay the V-cycle as a HIP graph: the launch sequence is the same every
    // cycle once each level starts from its own pair of buffers, so it is
    // captured once per (level-0 buffers, options) and replayed with one
    // launch, which removes the host launch gaps between the small coarse-level
    // kernels.  Off: plain stream launches (identical kernels and results).
    bool graphs = true;
    ~NodalHierarchy() {
        drop_graphs();
        if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
    }
    void drop_graphs();

   private:
    struct CycleGraph {
        T* x0 = nullptr;
        const T* b0 = nullptr;
        bool ready0 = false;
        bool guess0 = false;
        bool timed = false;
        int* bottom_iters = nullptr;
        MGOpts o;
        hipGraphExec_t exec = nullptr;
        std::vector<T*> xs, x2s;  // buffer roles after the cycle
    };
    std::vector<CycleGraph> cache_;
    // capture happens on a private stream (the caller's may be the null
    // stream, which cannot be captured); the graph is then launched on the
    // caller's stream
    hipStream_t cap_stream_ = nullptr;
    void vcycle_body(const MGOpts& o, hipStream_t s, int* bottom_iters);
};

class EBNodal;

// Outer MLMG iteration: FP64 residual, V-cycle correction (FP64 or FP32).
class NodalSolver {
   public:
    NodalGeom g0;
    MGOpts opts;
    int nlev = 1;
    NodalHierarchy<double> H64;
    NodalHierarchy<float> H32;
    DevBuf sig0_buf;         // level-0 sigma (ghosted, FP64)
    DevBuf phi_a, phi_b;     // outer iterate (ping-pong)
    DevBuf res64;            // FP64 residual (plain mode's level-0 V-cycle rhs)
    DevBuf res32;            // FP32 residual (mixed mode's level-0 V-cycle rhs)
    DevBuf scal;             // device scalars: [0..7] u64 norms, doubles after
    DevBuf part;             // partial sums
    DevBuf ibuf;             // device ints (bottom iteration counter)
    unsigned long long* h_pinned = nullptr;
    double* phi = nullptr;   // current outer iterate (points into phi_a / phi_b)
    // the caller zeroed phi before this solve: the first residual is b itself
    bool phi_zero = false;
    LaunchTimer fine_timer;  // finest-level smoother launches
    bool sigma_set = false;
    bool h32_sigma = false;
    // embedded boundary (owned by the projector handle; nullptr: none)
    EBNodal* eb = nullptr;
    // learned coarse-level predictor (MGX_MODE_LEARNED)
    UNet* net = nullptr;
    int net_level = 2;
    DevBuf net_in, net_out;
    int logsig_lvl = -1;  // level whose mean log(sigma) is in scal[40]

    ~NodalSolver();
    void setup(const NodalGeom& g, const MGOpts& o);
    double* sigma0() { return sig0_buf.as<double>(); }
    // Mixed-mode V-cycle with an FP16-stored finest level (FP32 arithmetic):
    // the operator is scaled to L' = L dx0^2 / sigma_max and the residual to
    // r / s (s: the last residual norm), which keeps every stored value in
    // FP16's normal range; the levels below are an FP32 hierarchy on the
    // once-coarsened grid with the same scaling.
    struct HalfTop {
        NodalGeom g;            // level 0, dx scaled by 1 / dx0
        KGeom kg;
        Coef<float> coef;
        DevBuf sig, xa, xb, b, r, sub_b, sub_sig;
        __half* x = nullptr;
        __half* x2 = nullptr;
        NodalHierarchy<float> sub;
        double sigmax = 1.0, dx0 = 1.0;
        bool ready = false;     // x holds the first sweep from zero
        bool built = false, sigma_ok = false;
    } H16;
    bool use_half = false;  // FP32 storage by default: the sweep is issue-bound, FP16 gains < 1%
    // overset mask (MLNodeLaplacian::setOversetMask): per level, 1 = solve,
    // 0 = node fixed at phi = 0; coarse levels by injection
    std::vector<DevBuf> mask_lv;
    bool mask_on = false, mask_fixed = false, mask_bottom_fixed = false;
    // level-0 mask (canonical nodes, device); nullptr removes it
    void set_mask(const unsigned char* m0, hipStream_t s);
    void prepare_h16(hipStream_t s);
    void vcycle16(hipStream_t s);
    void sigma_changed() {
        sigma_set = true;
        h32_sigma = false;
        H16.sigma_ok = false;
        logsig_lvl = -1;
    }
    // FP32 correction in H32.L[0].x from the FP32 residual in H32.L[0].b via
    // the learned predictor at level net_level.
    void learned_correction(hipStream_t s);
    void check_learned() const;
    // features of the current H32.L[0].b at level `lvl` -> in[8*nodes]; scalars in scal
    void learned_features(int lvl, float* in, hipStream_t s);
    void prepare_h32(hipStream_t s);
    void training_pair(const double* b, int lvl, float* in, float* target, hipStream_t s);
    // Solve L phi = b for the canonical rhs b (device, FP64; modified in place:
    // Dirichlet rows zeroed, mean removed for singular problems).  The initial
    // guess is this->phi (zeroed by the caller if wanted).
    void solve(double* b, double rtol, double atol, int mode, mgx_solve_info* info,
               hipStream_t s);
    double read_norm(int slot, hipStream_t s);
    void zero_norm(int slot, hipStream_t s);
    double sum(const double* a, long n, hipStream_t s);
    // time candidate launch shapes on this problem's levels (mode's
    // hierarchy) and keep the fastest in the process-wide table; returns a
    // text report (one line per tuned pass)
    std::string autotune(int mode, hipStream_t s);
};

// the process-wide launch-shape table as text, and loading one back
std::string tuning_table();
void set_tuning_table(const std::string& text);

// Box list of one level, its device box-offset table and tile tables.
class BoxLayout {
   public:
    std::vector<int> boxes;  // 6 ints per box
    int nboxes = 0;
    DevBuf dboxes;
    DevBuf dtiles[REG_NREG];
    int ntiles[REG_NREG] = {0, 0, 0, 0, 0, 0};
    // partial: the boxes need not tile the domain (an AMR fine level)
    void build(const int* boxes6, int nb, const NodalGeom& g, bool partial = false);
    const BoxInfo* dev_boxes() const { return dboxes.as<BoxInfo>(); }
    const Tile* tiles(int region) const { return dtiles[region].as<Tile>(); }
};

}  // namespace mgx
