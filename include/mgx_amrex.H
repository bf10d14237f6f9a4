// mgx_amrex.H -- header-only C++ facade over the C ABI (mgx.h) with the call
// surface incflo uses on AMReX-Hydro's projectors and AMReX's MLMG, so the
// GPU solver drops in at the same call sites:
//   NodalProjector  src/projection/incflo_apply_nodal_projection.cpp:181-219
//   MacProjector    src/convection/incflo_compute_MAC_projected_velocities.cpp:96-300,
//                   src/incflo.cpp:286-291,
//                   src/projection/incflo_apply_cc_projection.cpp:321,387-389
//   MLMG            src/diffusion/DiffusionTensorOp.cpp:188-215
// Field containers are views (device pointer + layout) of caller-owned
// buffers in the boxed layout: boxes back to back, per box component-major,
// x fastest, ngrow ghost layers (AMReX Array4 order).  Errors end the process
// with the library's message, as amrex::Abort does at the reference's call
// sites (src/projection/incflo_projection_bc.cpp:36,74).
#pragma once

#include <array>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mgx.h"

namespace mgx {

[[noreturn]] inline void Abort(const std::string& msg) {
    std::fprintf(stderr, "mgx::Abort: %s\n", msg.c_str());
    std::fflush(stderr);
    std::abort();
}

// amrex::LinOpBCType values used by incflo's BC maps
enum class LinOpBCType : int {
    Periodic = MGX_BC_PERIODIC,
    Dirichlet = MGX_BC_DIRICHLET,
    Neumann = MGX_BC_NEUMANN,
    inflow = MGX_BC_INFLOW,
    Robin = MGX_BC_ROBIN,
};

struct Box {
    std::array<int, 3> lo, hi;  // cell indices, inclusive
};

struct BoxArray {
    std::vector<Box> boxes;
    std::vector<int> flat() const {
        std::vector<int> f;
        for (const Box& b : boxes)
            f.insert(f.end(), {b.lo[0], b.lo[1], b.lo[2], b.hi[0], b.hi[1], b.hi[2]});
        return f;
    }
};

struct Geometry {
    std::array<int, 3> n_cell;
    std::array<double, 3> prob_lo{0.0, 0.0, 0.0}, prob_hi{1.0, 1.0, 1.0};
    std::array<int, 3> is_periodic{1, 1, 1};
};

// A MultiFab view: device buffer of one level in the boxed layout.
struct MultiFab {
    double* p = nullptr;
    const BoxArray* ba = nullptr;
    int ncomp = 1;
    int ngrow = 0;
};

struct LPInfo {
    int max_coarsening_level = 100;
    LPInfo& setMaxCoarseningLevel(int n) {
        max_coarsening_level = n;
        return *this;
    }
};

// Owns an mgx context (device + stream).  hipStream_t is passed as void*.
class Context {
   public:
    explicit Context(int device = 0, void* stream = nullptr) {
        if (mgx_ctx_create(device, stream, &h_) != MGX_OK)
            Abort("mgx_ctx_create failed for device " + std::to_string(device));
    }
    ~Context() { mgx_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    mgx_ctx get() const { return h_; }
    void check(int st, const char* what) const {
        if (st != MGX_OK) Abort(std::string(what) + ": " + mgx_ctx_last_error(h_));
    }
    void synchronize() const { check(mgx_ctx_synchronize(h_), "synchronize"); }

    // The context used by constructors called without one (incflo's call
    // shapes): the one made current with make_default(), else device 0 on its
    // null stream, created on first use.
    static Context& current() {
        if (!slot()) {
            static Context dev0(0);
            slot() = &dev0;
        }
        return *slot();
    }
    static void make_default(Context& c) { slot() = &c; }

   private:
    static Context*& slot() {
        static Context* c = nullptr;
        return c;
    }
    mgx_ctx h_ = nullptr;
};

namespace detail {
inline mgx_level_desc level_desc(const Geometry& g, const std::vector<int>& flat) {
    mgx_level_desc d{};
    for (int k = 0; k < 3; ++k) {
        d.n_cell[k] = g.n_cell[k];
        d.prob_lo[k] = g.prob_lo[k];
        d.prob_hi[k] = g.prob_hi[k];
        d.is_periodic[k] = g.is_periodic[k];
    }
    d.nboxes = (int)(flat.size() / 6);
    d.boxes = flat.data();
    return d;
}
inline void one_level(size_t n, const char* who) {
    if (n != 1) Abort(std::string(who) + ": only single-level hierarchies are supported");
}
// device buffer owned by a projector (getPhi / getGradPhi results)
struct DevField {
    mgx_ctx ctx = nullptr;
    void* p = nullptr;
    MultiFab view;
    void alloc(mgx_ctx c, const BoxArray* ba, int ncomp, bool nodal) {
        if (p) return;
        ctx = c;
        long long n = 0;
        for (const Box& b : ba->boxes) {
            long long v = 1;
            for (int k = 0; k < 3; ++k) v *= (b.hi[k] - b.lo[k] + 1) + (nodal ? 1 : 0);
            n += v * ncomp;
        }
        if (mgx_dev_malloc(c, (unsigned long long)n * sizeof(double), &p) != MGX_OK)
            Abort("device allocation failed");
        view.p = (double*)p;
        view.ba = ba;
        view.ncomp = ncomp;
        view.ngrow = 0;
    }
    ~DevField() {
        if (p) mgx_dev_free(ctx, p);
    }
};
}  // namespace detail

// Hydro::NodalProjector over one level, or up to MGX_AMR_MAX_LEVELS
// (refinement ratio 2: the composite projection of mgx_amr_nodal_*,
// src/projection/incflo_apply_nodal_projection.cpp:124-215).  project()
// overwrites vel in place.
class NodalProjector {
   public:
    NodalProjector(Context& ctx, std::vector<MultiFab*> vel, std::vector<MultiFab const*> sigma,
                   std::vector<Geometry> geom, LPInfo info = LPInfo())
        : ctx_(ctx), vel_(vel), geom_(geom), sigma_(sigma) {
        init(info);
        set_sigma();
    }
    NodalProjector(Context& ctx, std::vector<MultiFab*> vel, double const_sigma,
                   std::vector<Geometry> geom, LPInfo info = LPInfo())
        : ctx_(ctx), vel_(vel), geom_(geom) {
        init(info);
        if (ha_) Abort("NodalProjector: AMR levels need a sigma field per level");
        ctx_.check(mgx_nodal_set_const_sigma(h_, const_sigma), "set sigma");
    }
    NodalProjector(std::vector<MultiFab*> vel, std::vector<MultiFab const*> sigma,
                   std::vector<Geometry> geom, LPInfo info = LPInfo())
        : NodalProjector(Context::current(), vel, sigma, geom, info) {}
    NodalProjector(std::vector<MultiFab*> vel, double const_sigma, std::vector<Geometry> geom,
                   LPInfo info = LPInfo())
        : NodalProjector(Context::current(), vel, const_sigma, geom, info) {}
    ~NodalProjector() {
        if (h_) mgx_nodal_destroy(h_);
        if (ha_) mgx_amr_nodal_destroy(ha_);
    }
    NodalProjector(const NodalProjector&) = delete;
    NodalProjector& operator=(const NodalProjector&) = delete;

    void setDomainBC(std::array<LinOpBCType, 3> lo, std::array<LinOpBCType, 3> hi) {
        int l[3], r[3];
        for (int k = 0; k < 3; ++k) {
            l[k] = (int)lo[k];
            r[k] = (int)hi[k];
        }
        if (ha_) {
            // the AMR handle re-sets up its levels: hand sigma over again
            ctx_.check(mgx_amr_nodal_set_domain_bc(ha_, l, r), "setDomainBC");
            set_sigma();
        } else {
            ctx_.check(mgx_nodal_set_domain_bc(h_, l, r), "setDomainBC");
        }
    }
    // solve mode: MGX_MODE_FP64 (default), MGX_MODE_MIXED, MGX_MODE_LEARNED (one level)
    void setSolveMode(int mode) { mode_ = mode; }
    void setVerbose(int v) {
        opts_.verbose = v;
        if (ha_) {
            ctx_.check(mgx_amr_nodal_set_options(ha_, &opts_), "set options");
            set_sigma();
        } else {
            ctx_.check(mgx_nodal_set_options(h_, &opts_), "set options");
        }
    }
    void project(double rtol, double atol) {
        if (ha_) {
            const size_t n = vel_.size();
            std::vector<double*> v(n), ph(n), gp(n);
            for (size_t l = 0; l < n; ++l) {
                phis_[l].alloc(ctx_.get(), vel_[l]->ba, 1, true);
                gphis_[l].alloc(ctx_.get(), vel_[l]->ba, 3, false);
                if (vel_[l]->ngrow != vel_[0]->ngrow) Abort("NodalProjector: velocity ghost widths differ");
                v[l] = vel_[l]->p;
                ph[l] = phis_[l].view.p;
                gp[l] = gphis_[l].view.p;
            }
            ctx_.check(mgx_amr_nodal_project(ha_, v.data(), vel_[0]->ngrow, ph.data(), 0, gp.data(), 0,
                                             rtol, atol, mode_, &info_),
                       "NodalProjector::project");
            projected_ = true;
            return;
        }
        ctx_.check(mgx_nodal_project(h_, vel_[0]->p, vel_[0]->ngrow, rtol, atol, mode_, &info_),
                   "NodalProjector::project");
    }
    std::vector<MultiFab*> getPhi() {
        if (ha_) {
            if (!projected_) Abort("getPhi before project");
            std::vector<MultiFab*> out;
            for (size_t l = 0; l < vel_.size(); ++l) out.push_back(&phis_[l].view);
            return out;
        }
        phi_.alloc(ctx_.get(), vel_[0]->ba, 1, true);
        ctx_.check(mgx_nodal_get_phi(h_, phi_.view.p, 0), "getPhi");
        return {&phi_.view};
    }
    std::vector<MultiFab*> getGradPhi() {
        if (ha_) {
            if (!projected_) Abort("getGradPhi before project");
            std::vector<MultiFab*> out;
            for (size_t l = 0; l < vel_.size(); ++l) out.push_back(&gphis_[l].view);
            return out;
        }
        gphi_.alloc(ctx_.get(), vel_[0]->ba, 3, false);
        ctx_.check(mgx_nodal_get_gradphi(h_, gphi_.view.p, 0), "getGradPhi");
        return {&gphi_.view};
    }
    const mgx_solve_info& info() const { return info_; }
    mgx_nodal handle() const { return h_; }
    Context& context() const { return ctx_; }

   private:
    void init(const LPInfo& info) {
        if (vel_.size() >= 2 && vel_.size() == geom_.size() && vel_.size() <= MGX_AMR_MAX_LEVELS) {
            const size_t n = vel_.size();
            flats_.resize(n);
            std::vector<mgx_level_desc> d(n);
            for (size_t l = 0; l < n; ++l) {
                flats_[l] = vel_[l]->ba->flat();
                d[l] = detail::level_desc(geom_[l], flats_[l]);
            }
            ctx_.check(mgx_amr_nodal_create(ctx_.get(), d.data(), (int)n, &ha_), "NodalProjector");
            mgx_default_mlmg_opts(&opts_);
            opts_.max_coarsening_level = info.max_coarsening_level;
            ctx_.check(mgx_amr_nodal_set_options(ha_, &opts_), "set options");
            return;
        }
        detail::one_level(vel_.size(), "NodalProjector");
        detail::one_level(geom_.size(), "NodalProjector");
        flat_ = vel_[0]->ba->flat();
        mgx_level_desc d = detail::level_desc(geom_[0], flat_);
        ctx_.check(mgx_nodal_create(ctx_.get(), &d, &h_), "NodalProjector");
        mgx_default_mlmg_opts(&opts_);
        opts_.max_coarsening_level = info.max_coarsening_level;
        ctx_.check(mgx_nodal_set_options(h_, &opts_), "set options");
    }
    void set_sigma() {
        if (sigma_.empty()) return;
        if (ha_) {
            std::vector<const double*> sg;
            for (const MultiFab* m : sigma_) {
                if (m->ngrow != sigma_[0]->ngrow) Abort("NodalProjector: sigma ghost widths differ");
                sg.push_back(m->p);
            }
            if (sg.size() != vel_.size()) Abort("NodalProjector: AMR levels need a sigma field per level");
            ctx_.check(mgx_amr_nodal_set_sigma(ha_, sg.data(), sigma_[0]->ngrow), "set sigma");
        } else {
            detail::one_level(sigma_.size(), "NodalProjector");
            ctx_.check(mgx_nodal_set_sigma(h_, sigma_[0]->p, sigma_[0]->ngrow), "set sigma");
        }
    }
    Context& ctx_;
    std::vector<MultiFab*> vel_;
    std::vector<Geometry> geom_;
    std::vector<MultiFab const*> sigma_;
    std::vector<int> flat_;
    std::vector<std::vector<int>> flats_;
    mgx_nodal h_ = nullptr;
    mgx_amr_nodal ha_ = nullptr;  // AMR levels
    mgx_mlmg_opts opts_{};
    mgx_solve_info info_{};
    int mode_ = MGX_MODE_FP64;
    bool projected_ = false;
    detail::DevField phi_, gphi_;
    detail::DevField phis_[MGX_AMR_MAX_LEVELS], gphis_[MGX_AMR_MAX_LEVELS];
};

// Hydro::MacProjector (no EB) over one level, or up to MGX_AMR_MAX_LEVELS
// (refinement ratio 2: the composite projection of mgx_amr_mac_*,
// src/convection/incflo_compute_MAC_projected_velocities.cpp:96-133, 272-302).
class MacProjector {
   public:
    MacProjector(Context& ctx, std::vector<Geometry> geom) : ctx_(ctx), geom_(geom) {
        if (geom_.empty() || geom_.size() > MGX_AMR_MAX_LEVELS)
            Abort("MacProjector: one to MGX_AMR_MAX_LEVELS levels (refinement ratio 2)");
    }
    explicit MacProjector(std::vector<Geometry> geom) : MacProjector(Context::current(), geom) {}
    ~MacProjector() {
        if (h_) mgx_mac_destroy(h_);
        if (ha_) mgx_amr_mac_destroy(ha_);
    }
    MacProjector(const MacProjector&) = delete;
    MacProjector& operator=(const MacProjector&) = delete;

    bool needInitialization() const { return h_ == nullptr && ha_ == nullptr; }
    // initProjector(lp_info, beta) (:112): the BoxArrays are the coefficients'
    void initProjector(LPInfo info, std::vector<std::array<MultiFab const*, 3>> beta) {
        if (beta.size() >= 2 && beta.size() == geom_.size()) {
            std::vector<const BoxArray*> bas;
            for (auto& b : beta) bas.push_back(b[0]->ba);
            create_amr(bas, info);
            updateCoeffs(beta);
            return;
        }
        detail::one_level(beta.size(), "MacProjector");
        initProjector(*beta[0][0]->ba, info, beta);
    }
    // initProjector(lp_info, beta) with face coefficients
    void initProjector(const BoxArray& ba, LPInfo info,
                       std::vector<std::array<MultiFab const*, 3>> beta) {
        create(ba, info);
        updateCoeffs(beta);
    }
    // initProjector(ba, dm, lp_info, beta) with a constant coefficient
    void initProjector(const BoxArray& ba, LPInfo info, double beta) {
        create(ba, info);
        updateBeta(beta);
    }
    void setDomainBC(std::array<LinOpBCType, 3> lo, std::array<LinOpBCType, 3> hi) {
        int l[3], r[3];
        for (int k = 0; k < 3; ++k) {
            l[k] = (int)lo[k];
            r[k] = (int)hi[k];
        }
        if (ha_) {
            // the AMR handle re-sets up its levels: hand beta over again
            ctx_.check(mgx_amr_mac_set_domain_bc(ha_, l, r), "setDomainBC");
            if (!beta_.empty()) updateCoeffs(beta_);
            return;
        }
        ctx_.check(mgx_mac_set_domain_bc(need(), l, r), "setDomainBC");
    }
    // setLevelBC(lev, nullptr, &robin_a, &robin_b, &robin_f): Robin (mixed) face data
    void setLevelBC(int lev, const MultiFab* levelbcdata, const MultiFab* robin_a,
                    const MultiFab* robin_b, const MultiFab* robin_f) {
        if (lev != 0 || levelbcdata || ha_)
            Abort("MacProjector::setLevelBC: one level, level 0, no Dirichlet data");
        ctx_.check(mgx_mac_set_level_bc(need(), robin_a->p, robin_b->p, robin_f->p,
                                        robin_a->ngrow),
                   "setLevelBC");
    }
    void updateBeta(double beta) {
        if (ha_) Abort("MacProjector::updateBeta: AMR levels need beta fields");
        ctx_.check(mgx_mac_set_const_beta(need(), beta), "updateBeta");
    }
    void updateCoeffs(std::vector<std::array<MultiFab const*, 3>> beta) {
        if (ha_) {
            if (beta.size() != geom_.size()) Abort("MacProjector: a beta field per level");
            beta_ = beta;
            std::vector<const double*> b;
            for (auto& lv : beta)
                for (int d = 0; d < 3; ++d) {
                    if (lv[d]->ngrow != beta[0][0]->ngrow)
                        Abort("MacProjector: beta fields of equal ghost width");
                    b.push_back(lv[d]->p);
                }
            ctx_.check(mgx_amr_mac_set_beta(ha_, b.data(), beta[0][0]->ngrow), "updateCoeffs");
            return;
        }
        detail::one_level(beta.size(), "MacProjector");
        ctx_.check(mgx_mac_set_beta(need(), beta[0][0]->p, beta[0][1]->p, beta[0][2]->p,
                                    beta[0][0]->ngrow),
                   "updateCoeffs");
    }
    void setUMAC(std::vector<std::array<MultiFab*, 3>> umac) {
        if (ha_) {
            if (umac.size() != geom_.size()) Abort("MacProjector::setUMAC: a face velocity per level");
            umacs_ = umac;
            return;
        }
        detail::one_level(umac.size(), "MacProjector");
        umac_ = umac[0];
    }
    void setSolveMode(int mode) { mode_ = mode; }
    void project(double rtol, double atol) { run(nullptr, 0, rtol, atol); }
    void project(std::vector<MultiFab*> phi, double rtol, double atol) {
        if (ha_) {
            if (phi.size() != geom_.size()) Abort("MacProjector::project: phi per level");
            std::vector<double*> ph;
            for (MultiFab* m : phi) {
                if (m->ngrow != phi[0]->ngrow) Abort("MacProjector::project: phi ghost widths differ");
                ph.push_back(m->p);
            }
            run_amr(ph.data(), phi[0]->ngrow, rtol, atol);
            return;
        }
        detail::one_level(phi.size(), "MacProjector");
        run(phi[0]->p, phi[0]->ngrow, rtol, atol);
    }
    // fluxes -beta grad(phi) of the last solve (MLMG::Location::FaceCenter)
    void getFluxes(std::vector<std::array<MultiFab*, 3>> flux) {
        if (ha_) Abort("MacProjector::getFluxes: one level only");
        detail::one_level(flux.size(), "MacProjector");
        ctx_.check(mgx_mac_get_fluxes(need(), flux[0][0]->p, flux[0][1]->p, flux[0][2]->p,
                                      flux[0][0]->ngrow),
                   "getFluxes");
    }
    const mgx_solve_info& info() const { return info_; }
    mgx_mac handle() const { return h_; }
    Context& context() const { return ctx_; }

   private:
    void create(const BoxArray& ba, const LPInfo& info) {
        if (h_) mgx_mac_destroy(h_);
        flat_ = ba.flat();
        mgx_level_desc d = detail::level_desc(geom_[0], flat_);
        ctx_.check(mgx_mac_create(ctx_.get(), &d, &h_), "MacProjector");
        mgx_mlmg_opts o;
        mgx_default_mlmg_opts(&o);
        o.maxiter = 200;
        o.bottom_maxiter = 200;
        o.max_coarsening_level = info.max_coarsening_level;
        ctx_.check(mgx_mac_set_options(h_, &o), "set options");
    }
    void create_amr(const std::vector<const BoxArray*>& bas, const LPInfo& info) {
        if (ha_) mgx_amr_mac_destroy(ha_);
        flats_.resize(bas.size());
        std::vector<mgx_level_desc> d(bas.size());
        for (size_t l = 0; l < bas.size(); ++l) {
            flats_[l] = bas[l]->flat();
            d[l] = detail::level_desc(geom_[l], flats_[l]);
        }
        ctx_.check(mgx_amr_mac_create(ctx_.get(), d.data(), (int)d.size(), &ha_), "MacProjector");
        mgx_mlmg_opts o;
        mgx_default_mlmg_opts(&o);
        o.maxiter = 200;
        o.bottom_maxiter = 200;
        o.max_coarsening_level = info.max_coarsening_level;
        ctx_.check(mgx_amr_mac_set_options(ha_, &o), "set options");
    }
    void run_amr(double* const* phi, int phi_ng, double rtol, double atol) {
        if (umacs_.size() != geom_.size()) Abort("MacProjector::project: setUMAC has not been called");
        std::vector<double*> u;
        for (auto& lv : umacs_)
            for (int d = 0; d < 3; ++d) {
                if (!lv[d] || lv[d]->ngrow != umacs_[0][0]->ngrow)
                    Abort("MacProjector: face velocities missing or of unequal ghost widths");
                u.push_back(lv[d]->p);
            }
        ctx_.check(mgx_amr_mac_project(ha_, u.data(), umacs_[0][0]->ngrow, phi, phi_ng, rtol, atol, mode_,
                                       &info_),
                   "MacProjector::project");
    }
    mgx_mac need() const {
        if (!h_) Abort("MacProjector: initProjector has not been called");
        return h_;
    }
    void run(double* phi, int phi_ng, double rtol, double atol) {
        if (ha_) {
            run_amr(nullptr, 0, rtol, atol);
            return;
        }
        if (!umac_[0]) Abort("MacProjector::project: setUMAC has not been called");
        ctx_.check(mgx_mac_project(need(), umac_[0]->p, umac_[1]->p, umac_[2]->p, umac_[0]->ngrow,
                                   phi, phi_ng, rtol, atol, mode_, &info_),
                   "MacProjector::project");
    }
    Context& ctx_;
    std::vector<Geometry> geom_;
    std::vector<int> flat_;
    std::vector<std::vector<int>> flats_;
    mgx_mac h_ = nullptr;
    mgx_amr_mac ha_ = nullptr;  // AMR levels
    std::vector<std::array<MultiFab const*, 3>> beta_;
    std::array<MultiFab*, 3> umac_{nullptr, nullptr, nullptr};
    std::vector<std::array<MultiFab*, 3>> umacs_;  // AMR levels
    mgx_solve_info info_{};
    int mode_ = MGX_MODE_FP64;
};

// One level of incflo's LevelData (src/incflo.H:618-656) as the projections use it.
struct ProjectionFields {
    MultiFab* velocity = nullptr;          // 3 comps, >= 1 ghost
    const MultiFab* velocity_o = nullptr;  // incremental projections
    const MultiFab* density = nullptr;     // or const_rho > 0
    double const_rho = 0.0;
    MultiFab* gp = nullptr;                // 3 comps
    MultiFab* p_nd = nullptr;              // nodal (nodal projection)
    MultiFab* p_cc = nullptr;              // cells (cell-centred projection)
};

// incflo::ApplyProjection (src/incflo.cpp:213-231): with use_cc_proj the
// cell-centred projection on incflo's persistent MAC projector
// (ApplyCCProjection, src/projection/incflo_apply_cc_projection.cpp:199-487;
// the projector is initialised with a constant beta and the MAC boundary
// types when it needs it, :282-310), otherwise the nodal projection with a
// per-call projector (ApplyNodalProjection, src/projection/incflo_apply_nodal_projection.cpp:29-267).
// inflow: the 6x3 inflow velocities [2*dir + side][comp] or nullptr.
inline mgx_solve_info ApplyProjection(Context& ctx, bool use_cc_proj, MacProjector& macproj,
                                      const Geometry& geom, const ProjectionFields& f,
                                      std::array<LinOpBCType, 3> nodal_lo,
                                      std::array<LinOpBCType, 3> nodal_hi,
                                      std::array<LinOpBCType, 3> mac_lo,
                                      std::array<LinOpBCType, 3> mac_hi, double scaling_factor,
                                      bool incremental, const double* inflow, double rtol,
                                      double atol, int mode = MGX_MODE_FP64) {
    mgx_solve_info info{};
    if (!f.velocity || !f.gp || (!f.density && !(f.const_rho > 0.0)))
        Abort("ApplyProjection: velocity, gp and a density are required");
    const double* rho = f.density ? f.density->p : nullptr;
    const int rho_ng = f.density ? f.density->ngrow : 0;
    const double* vo = incremental && f.velocity_o ? f.velocity_o->p : nullptr;
    const int vo_ng = f.velocity_o ? f.velocity_o->ngrow : 0;
    if (use_cc_proj) {
        if (!f.p_cc) Abort("ApplyProjection: the cell-centred projection writes p_cc");
        if (macproj.needInitialization()) {
            macproj.initProjector(*f.velocity->ba, LPInfo(), scaling_factor);
            macproj.setDomainBC(mac_lo, mac_hi);
        }
        ctx.check(mgx_apply_cc_projection(macproj.handle(), rho, rho_ng, f.const_rho,
                                          f.velocity->p, f.velocity->ngrow, vo, vo_ng, f.gp->p,
                                          f.gp->ngrow, f.p_cc->p, f.p_cc->ngrow, scaling_factor,
                                          incremental ? 1 : 0, inflow, nullptr, nullptr, nullptr,
                                          0, rtol, atol, mode, &info),
                  "ApplyCCProjection");
        return info;
    }
    if (!f.p_nd) Abort("ApplyProjection: the nodal projection writes p_nd");
    const std::vector<int> flat = f.velocity->ba->flat();
    mgx_level_desc d = detail::level_desc(geom, flat);
    mgx_nodal h = nullptr;
    ctx.check(mgx_nodal_create(ctx.get(), &d, &h), "NodalProjector");
    int l[3], r[3];
    for (int k = 0; k < 3; ++k) {
        l[k] = (int)nodal_lo[k];
        r[k] = (int)nodal_hi[k];
    }
    int st = mgx_nodal_set_domain_bc(h, l, r);
    if (st == MGX_OK)
        st = mgx_apply_nodal_projection(h, rho, rho_ng, f.const_rho, f.velocity->p,
                                        f.velocity->ngrow, vo, vo_ng, f.gp->p, f.gp->ngrow,
                                        f.p_nd->p, f.p_nd->ngrow, scaling_factor,
                                        incremental ? 1 : 0, inflow, rtol, atol, mode, &info);
    if (st != MGX_OK) {
        const std::string msg = mgx_ctx_last_error(ctx.get());
        mgx_nodal_destroy(h);
        Abort("ApplyNodalProjection: " + msg);
    }
    mgx_nodal_destroy(h);
    return info;
}

// amrex::MLMG driving one of the two operators directly (setMaxIter, ...,
// solve(sol, rhs, rtol, atol)); the operator's coefficients and BCs come from
// the projector it wraps.
class MLMG {
   public:
    explicit MLMG(NodalProjector& op) : ctx_(op.context()), nodal_(op.handle()) { defaults(); }
    explicit MLMG(MacProjector& op) : ctx_(op.context()), mac_(op.handle()) {
        defaults();
        o_.maxiter = 200;
        o_.bottom_maxiter = 200;
    }
    void setMaxIter(int n) { o_.maxiter = n; }
    void setBottomMaxIter(int n) { o_.bottom_maxiter = n; }
    void setVerbose(int v) { o_.verbose = v; }
    void setBottomVerbose(int v) { o_.bottom_verbose = v; }
    void setBottomTolerance(double t) { o_.bottom_rtol = t; }
    void setBottomToleranceAbs(double t) { o_.bottom_atol = t; }
    // MLMG::BottomSolver; hypre is refused at solve time (MGX_ERR_UNSUPPORTED)
    enum class BottomSolver { Default = MGX_BOTTOM_DEFAULT, bicgstab = MGX_BOTTOM_BICGSTAB,
                              cg = MGX_BOTTOM_CG, bicgcg = MGX_BOTTOM_BICGCG,
                              cgbicg = MGX_BOTTOM_CGBICG, smoother = MGX_BOTTOM_SMOOTHER,
                              hypre = MGX_BOTTOM_HYPRE };
    void setBottomSolver(BottomSolver b) { o_.bottom_solver = static_cast<int>(b); }
    void setSolveMode(int mode) { mode_ = mode; }
    // returns the final residual norm, as amrex::MLMG::solve does
    double solve(std::vector<MultiFab*> sol, std::vector<MultiFab const*> rhs, double rtol,
                 double atol) {
        detail::one_level(sol.size(), "MLMG::solve");
        detail::one_level(rhs.size(), "MLMG::solve");
        if (nodal_) {
            ctx_.check(mgx_nodal_set_options(nodal_, &o_), "MLMG options");
            ctx_.check(mgx_nodal_solve(nodal_, sol[0]->p, sol[0]->ngrow, rhs[0]->p, rhs[0]->ngrow,
                                       rtol, atol, mode_, &info_),
                       "MLMG::solve");
        } else {
            if (!mac_) Abort("MLMG: the operator has not been initialised");
            ctx_.check(mgx_mac_set_options(mac_, &o_), "MLMG options");
            ctx_.check(mgx_mac_solve(mac_, sol[0]->p, sol[0]->ngrow, rhs[0]->p, rhs[0]->ngrow, rtol,
                                     atol, mode_, &info_),
                       "MLMG::solve");
        }
        return info_.resnorm;
    }
    int getNumIters() const { return info_.iters; }
    const mgx_solve_info& info() const { return info_; }

   private:
    void defaults() { mgx_default_mlmg_opts(&o_); }
    Context& ctx_;
    mgx_nodal nodal_ = nullptr;
    mgx_mac mac_ = nullptr;
    mgx_mlmg_opts o_{};
    mgx_solve_info info_{};
    int mode_ = MGX_MODE_FP64;
};

}  // namespace mgx
