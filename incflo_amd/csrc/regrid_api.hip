// regrid_api.hip -- dynamic refinement of a two-level hierarchy (ratio 2):
// ErrorEst tagging on the device, Berger-Rigoutsos clustering of the tags
// into fine boxes on the host, and the fill of a new or remade fine level
// from the old fine data and the coarse level on the device.
//
// incflo: ErrorEst (src/incflo_tagging.cpp:11-198); MakeNewLevelFromCoarse /
// RemakeLevel (src/incflo_regrid.cpp:9-113) through fillcoarsepatch_* /
// fillpatch_* with cell_cons_interp (src/boundary_conditions/incflo_fillpatch.cpp).
// AMReX's grid generation and interpolater are not vendored; both are
// restated from their published algorithms:
//   * clustering: M. Berger and I. Rigoutsos, "An algorithm for point
//     clustering and grid generation", IEEE Trans. SMC 21(5), 1991, with
//     AMReX's cut rules (a hole nearest the middle, else the largest
//     sign change of the signature's second difference, else bisection);
//   * cell_cons_interp (CellConservativeLinear with linear limiting): central
//     coarse slopes (a one-sided 4-point slope next to ext_dir / hoextrap
//     faces), each direction scaled by one factor shared by the components,
//     min over components of min(1, 2 min(|forward|, |backward|) / |central|)
//     (0 where forward and backward differ in sign); fine = coarse + offset
//     . slope with offsets +-1/4 of a coarse cell.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "box_kernels.h"
#include "mgx_common.h"
#include "mgx_ctx.h"
#include "nodal_mg.h"

using namespace mgx;

namespace {

NodalGeom level_geom(const mgx_level_desc* lv, double dx[3]) {
    MGX_REQUIRE(lv && lv->boxes && lv->nboxes > 0, MGX_ERR_ARG, "bad level description");
    int n[3], per[3], z[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
        n[d] = lv->n_cell[d];
        per[d] = lv->is_periodic[d];
        dx[d] = (lv->prob_hi[d] - lv->prob_lo[d]) / n[d];
        MGX_REQUIRE(n[d] > 0 && dx[d] > 0.0, MGX_ERR_ARG, "bad level geometry");
    }
    return make_nodal_geom(n, per, z, z, dx);
}

KGeom cell_kgeom(const NodalGeom& g) {
    KGeom k{};
    for (int d = 0; d < 3; ++d) {
        k.n[d] = g.n[d];
        k.per[d] = g.per[d];
        k.nn[d] = g.n[d];
    }
    return k;
}

// ---- ErrorEst ---------------------------------------------------------------
struct TagArgs {
    KGeom g;
    const BoxInfo* boxes;
    const Tile* tiles;
    const double* rho;
    int rho_ng;
    int tag_rho, tag_grad, tag_region;
    double rhoerr, gradrhoerr;
    double rlo[3], rhi[3];
    double prob_lo[3], dx[3];
    unsigned char* tags;  // canonical cells
};

// src/incflo_tagging.cpp:60-125: rho > rhoerr; the largest one-sided density
// difference along any direction >= gradrhoerr (the ghost cells are read);
// the cell centre inside the closed tag region.  Untagged cells are left
// as they are (tags accumulate, as TagBox::SET does).
__global__ void __launch_bounds__(256) tag_kernel(TagArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL, bi, i, j, k)) return;
    bool tag = false;
    if (a.tag_rho || a.tag_grad) {
        const BoxInfo& b = a.boxes[bi];
        const long o = field_offset(b, a.rho_ng, 1, false);
        auto r = [&](int di, int dj, int dk) {
            return a.rho[o + box_index(b, a.rho_ng, false, i + di, j + dj, k + dk, 0)];
        };
        const double c = r(0, 0, 0);
        if (a.tag_rho && c > a.rhoerr) tag = true;
        if (a.tag_grad) {
            const double ax = fmax(fabs(r(1, 0, 0) - c), fabs(c - r(-1, 0, 0)));
            const double ay = fmax(fabs(r(0, 1, 0) - c), fabs(c - r(0, -1, 0)));
            const double az = fmax(fabs(r(0, 0, 1) - c), fabs(c - r(0, 0, -1)));
            if (fmax(ax, fmax(ay, az)) >= a.gradrhoerr) tag = true;
        }
    }
    if (a.tag_region) {
        const double x = a.prob_lo[0] + (i + 0.5) * a.dx[0];
        const double y = a.prob_lo[1] + (j + 0.5) * a.dx[1];
        const double z = a.prob_lo[2] + (k + 0.5) * a.dx[2];
        if (x >= a.rlo[0] && x <= a.rhi[0] && y >= a.rlo[1] && y <= a.rhi[1] && z >= a.rlo[2] &&
            z <= a.rhi[2])
            tag = true;
    }
    if (tag) a.tags[i + (long)a.g.n[0] * (j + (long)a.g.n[1] * k)] = 1;
}

// ---- FillPatch of a fine level ------------------------------------------------
// canonical ghosted coarse array: (n0+2)(n1+2)(n2+2) cells per component,
// cell (i,j,k) at (i+1) + (n0+2)((j+1) + (n1+2)(k+1))
__device__ __forceinline__ long gidx(const KGeom& g, int i, int j, int k) {
    return (i + 1) + (long)(g.n[0] + 2) * ((j + 1) + (long)(g.n[1] + 2) * (k + 1));
}

// valid cells and the ghost cells outside non-periodic faces (the caller's)
__global__ void __launch_bounds__(256) gather_ghosted_kernel(KGeom g, const BoxInfo* boxes,
                                                             const Tile* tiles,
                                                             const double* __restrict__ in, int ng,
                                                             int ncomp, double* __restrict__ out) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL_GROWN, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long base = field_offset(b, ng, ncomp, false);
    const long ng3 = (long)(g.n[0] + 2) * (g.n[1] + 2) * (g.n[2] + 2);
    for (int q = 0; q < ncomp; ++q)
        out[q * ng3 + gidx(g, i, j, k)] = in[base + box_index(b, ng, false, i, j, k, q)];
}

// ghost cells across periodic faces: copies of their periodic images
__global__ void __launch_bounds__(256) wrap_ghosted_kernel(KGeom g, int ncomp,
                                                           double* __restrict__ a) {
    const long ng3 = (long)(g.n[0] + 2) * (g.n[1] + 2) * (g.n[2] + 2);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ng3) return;
    int p[3];
    p[0] = (int)(t % (g.n[0] + 2)) - 1;
    p[1] = (int)((t / (g.n[0] + 2)) % (g.n[1] + 2)) - 1;
    p[2] = (int)(t / ((long)(g.n[0] + 2) * (g.n[1] + 2))) - 1;
    int s[3] = {p[0], p[1], p[2]};
    bool wrapped = false;
    for (int d = 0; d < 3; ++d)
        if (g.per[d] && (p[d] < 0 || p[d] >= g.n[d])) {
            s[d] = p[d] < 0 ? p[d] + g.n[d] : p[d] - g.n[d];
            wrapped = true;
        }
    if (!wrapped) return;
    // the source lies inside the domain along every periodic direction, so it
    // is never itself written here
    for (int q = 0; q < ncomp; ++q) a[q * ng3 + t] = a[q * ng3 + gidx(g, s[0], s[1], s[2])];
}

// old fine level -> dense canonical fine arrays and an ownership mask
__global__ void __launch_bounds__(256) old_fine_kernel(KGeom gf, const BoxInfo* boxes,
                                                       const Tile* tiles,
                                                       const double* __restrict__ in, int ng,
                                                       int ncomp, double* __restrict__ dense,
                                                       unsigned char* __restrict__ own) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long base = field_offset(b, ng, ncomp, false);
    const long nf = (long)gf.n[0] * gf.n[1] * gf.n[2];
    const long c = i + (long)gf.n[0] * (j + (long)gf.n[1] * k);
    for (int q = 0; q < ncomp; ++q) dense[q * nf + c] = in[base + box_index(b, ng, false, i, j, k, q)];
    own[c] = 1;
}

constexpr int kMaxInterpComp = 8;

struct FillArgs {
    KGeom gf, gc;
    const BoxInfo* boxes;
    const Tile* tiles;
    double* out;
    int ng, ncomp;
    const double* cg;             // ghosted canonical coarse
    const double* dense;          // old fine (dense canonical), or null
    const unsigned char* own;
    int bc[kMaxInterpComp][6];    // amrex::BCType per component: lo x,y,z, hi x,y,z
};

__device__ __forceinline__ bool one_sided(int bc) {
    return bc == 3 /* ext_dir */ || bc == 4 /* hoextrap */;
}

__global__ void __launch_bounds__(256) fill_fine_kernel(FillArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = a.boxes[bi];
    const long base = field_offset(b, a.ng, a.ncomp, false);
    const long nf = (long)a.gf.n[0] * a.gf.n[1] * a.gf.n[2];
    const long c = i + (long)a.gf.n[0] * (j + (long)a.gf.n[1] * k);
    if (a.own && a.own[c]) {
        for (int q = 0; q < a.ncomp; ++q)
            a.out[base + box_index(b, a.ng, false, i, j, k, q)] = a.dense[q * nf + c];
        return;
    }
    const KGeom& g = a.gc;
    const int P[3] = {i >> 1, j >> 1, k >> 1};
    const double off[3] = {(i & 1) ? 0.25 : -0.25, (j & 1) ? 0.25 : -0.25, (k & 1) ? 0.25 : -0.25};
    const long ng3 = (long)(g.n[0] + 2) * (g.n[1] + 2) * (g.n[2] + 2);
    auto u = [&](int q, int d, int s) {
        int p[3] = {P[0], P[1], P[2]};
        p[d] += s;
        return a.cg[q * ng3 + gidx(g, p[0], p[1], p[2])];
    };
    double cen[3][kMaxInterpComp], fac[3] = {1.0, 1.0, 1.0};
    for (int d = 0; d < 3; ++d)
        for (int q = 0; q < a.ncomp; ++q) {
            const double um = u(q, d, -1), u0 = u(q, d, 0), up = u(q, d, 1);
            double cs = 0.5 * (up - um);
            if (!g.per[d] && P[d] == 0 && one_sided(a.bc[q][d]))
                cs = -16.0 / 15.0 * um + 0.5 * u0 + 2.0 / 3.0 * up - 0.1 * u(q, d, 2);
            else if (!g.per[d] && P[d] == g.n[d] - 1 && one_sided(a.bc[q][3 + d]))
                cs = 16.0 / 15.0 * up - 0.5 * u0 - 2.0 / 3.0 * um + 0.1 * u(q, d, -2);
            const double fw = 2.0 * (up - u0), bw = 2.0 * (u0 - um);
            const double slp = fw * bw >= 0.0 ? fmin(fabs(fw), fabs(bw)) : 0.0;
            if (cs != 0.0) fac[d] = fmin(fac[d], fmin(1.0, slp / fabs(cs)));
            cen[d][q] = cs;
        }
    for (int q = 0; q < a.ncomp; ++q) {
        const double v = u(q, 0, 0) + off[0] * fac[0] * cen[0][q] + off[1] * fac[1] * cen[1][q] +
                         off[2] * fac[2] * cen[2][q];
        a.out[base + box_index(b, a.ng, false, i, j, k, q)] = v;
    }
}

// ---- Berger-Rigoutsos ------------------------------------------------------
struct IBox {
    int lo[3], hi[3];
    long vol() const {
        long v = 1;
        for (int d = 0; d < 3; ++d) v *= hi[d] - lo[d] + 1;
        return v;
    }
    int len(int d) const { return hi[d] - lo[d] + 1; }
};

struct TagField {
    int n[3];
    const std::vector<unsigned char>* t;
    bool at(int i, int j, int k) const { return (*t)[i + (long)n[0] * (j + (long)n[1] * k)] != 0; }
};

// the bounding box of the tags inside b and their count (0: none)
long min_box(const TagField& f, IBox& b) {
    IBox s{{1 << 30, 1 << 30, 1 << 30}, {-1, -1, -1}};
    long cnt = 0;
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                if (f.at(i, j, k)) {
                    ++cnt;
                    const int p[3] = {i, j, k};
                    for (int d = 0; d < 3; ++d) {
                        s.lo[d] = std::min(s.lo[d], p[d]);
                        s.hi[d] = std::max(s.hi[d], p[d]);
                    }
                }
    if (cnt) b = s;
    return cnt;
}

// tag count of every plane of b normal to d
std::vector<long> signature(const TagField& f, const IBox& b, int d) {
    std::vector<long> s(b.len(d), 0);
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                if (f.at(i, j, k)) {
                    const int p[3] = {i, j, k};
                    ++s[p[d] - b.lo[d]];
                }
    return s;
}

enum CutKind { CUT_NONE = 0, CUT_BISECT = 1, CUT_STEEP = 2, CUT_HOLE = 3 };

// where to cut a signature (offset of the first plane of the upper part)
int find_cut(const std::vector<long>& h, CutKind& kind) {
    const int len = (int)h.size();
    kind = CUT_NONE;
    if (len <= 1) return 0;
    const int mid = len / 2;
    // the hole nearest the middle
    int cut = -1;
    for (int q = 0; q < len; ++q)
        if (h[q] == 0 && (cut < 0 || std::abs(q - mid) < std::abs(cut - mid))) cut = q;
    if (cut >= 0) {
        kind = CUT_HOLE;
        return cut;
    }
    // the largest sign change of the second difference (away from the ends),
    // ties to the one nearest the middle
    std::vector<long> lap(len, 0);
    for (int q = 1; q < len - 1; ++q) lap[q] = h[q + 1] - 2 * h[q] + h[q - 1];
    const int kMinOff = 2;
    const long kCutThresh = 2;
    long best = -1;
    for (int q = kMinOff; q < len - kMinOff; ++q) {
        const long a = lap[q - 1], b = lap[q];
        const long dif = std::labs(a - b);
        if (a * b < 0 && (dif > best || (dif == best && std::abs(q - mid) < std::abs(cut - mid)))) {
            best = dif;
            cut = q;
        }
    }
    if (best > kCutThresh) {
        kind = CUT_STEEP;
        return cut;
    }
    kind = CUT_BISECT;
    return mid;
}

void cluster(const TagField& f, IBox b, double eff, std::vector<IBox>& out) {
    const long nt = min_box(f, b);
    if (nt == 0) return;
    if ((double)nt >= eff * (double)b.vol()) {
        out.push_back(b);
        return;
    }
    // the best cut over the three directions: a hole before a steep cut
    // before a bisection, the longest direction among equals
    int dir = -1, cut = 0;
    CutKind best = CUT_NONE;
    for (int d = 0; d < 3; ++d) {
        if (b.len(d) < 2) continue;
        CutKind kind;
        const int c = find_cut(signature(f, b, d), kind);
        if (c <= 0 || c >= b.len(d)) continue;
        if (kind > best || (kind == best && dir >= 0 && b.len(d) > b.len(dir))) {
            best = kind;
            dir = d;
            cut = c;
        }
    }
    if (dir < 0) {  // a single cell in every direction that could be cut
        out.push_back(b);
        return;
    }
    IBox l = b, r = b;
    l.hi[dir] = b.lo[dir] + cut - 1;
    r.lo[dir] = b.lo[dir] + cut;
    cluster(f, l, eff, out);
    cluster(f, r, eff, out);
}

// BoxList::maxSize: split each direction longer than m into ceil(len/m)
// nearly equal pieces (the first len % pieces one longer)
void max_size(const IBox& b, int m, std::vector<IBox>& out) {
    std::vector<IBox> cur{b};
    for (int d = 0; d < 3; ++d) {
        std::vector<IBox> nxt;
        for (const IBox& c : cur) {
            const int len = c.len(d);
            if (len <= m) {
                nxt.push_back(c);
                continue;
            }
            const int np = (len + m - 1) / m, sz = len / np, extra = len % np;
            int lo = c.lo[d];
            for (int q = 0; q < np; ++q) {
                IBox p = c;
                p.lo[d] = lo;
                p.hi[d] = lo + sz + (q < extra ? 1 : 0) - 1;
                lo = p.hi[d] + 1;
                nxt.push_back(p);
            }
        }
        cur.swap(nxt);
    }
    out.insert(out.end(), cur.begin(), cur.end());
}

}  // namespace

extern "C" {

MGX_API int mgx_tag_cells(mgx_ctx ctx, const mgx_level_desc* level, const double* rho, int rho_ng,
                          double rhoerr, double gradrhoerr, const double* region_lo,
                          const double* region_hi, unsigned char* tags) {
    if (!ctx) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        MGX_REQUIRE(tags, MGX_ERR_ARG, "tag_cells: no tag array");
        const bool tr = std::isfinite(rhoerr);
        const bool tg = std::isfinite(gradrhoerr);
        MGX_REQUIRE(!(tr || tg) || (rho && rho_ng >= (tg ? 1 : 0)), MGX_ERR_ARG,
                    "tag_cells: the density criteria need the density (gradrhoerr: with one "
                    "filled ghost cell)");
        MGX_REQUIRE((region_lo == nullptr) == (region_hi == nullptr), MGX_ERR_ARG,
                    "tag_cells: give both region corners or neither");
        double dx[3];
        const NodalGeom g = level_geom(level, dx);
        BoxLayout layout;
        layout.build(level->boxes, level->nboxes, g, true);
        TagArgs a{};
        a.g = cell_kgeom(g);
        for (int d = 0; d < 3; ++d) {
            a.prob_lo[d] = level->prob_lo[d];
            a.dx[d] = dx[d];
            if (region_lo) {
                a.rlo[d] = region_lo[d];
                a.rhi[d] = region_hi[d];
            }
        }
        a.boxes = layout.dev_boxes();
        a.tiles = layout.tiles(REG_CELL);
        a.rho = rho;
        a.rho_ng = rho_ng;
        a.tag_rho = tr;
        a.tag_grad = tg;
        a.tag_region = region_lo ? 1 : 0;
        a.rhoerr = rhoerr;
        a.gradrhoerr = gradrhoerr;
        a.tags = tags;
        tag_kernel<<<layout.ntiles[REG_CELL], 256, 0, ctx->stream>>>(a);
        MGX_HIP(hipGetLastError());
        MGX_HIP(hipStreamSynchronize(ctx->stream));  // the layout's tables are freed on return
    });
}

MGX_API int mgx_cluster_tags(const unsigned char* tags, const int n_cell[3],
                             const int is_periodic[3], int n_error_buf, int blocking_factor,
                             int max_grid_size, double grid_eff, int* boxes, int max_boxes,
                             int* nboxes) {
    try {
        MGX_REQUIRE(tags && n_cell && boxes && nboxes && n_error_buf >= 0, MGX_ERR_ARG,
                    "cluster_tags: bad arguments");
        MGX_REQUIRE(blocking_factor >= 2 && blocking_factor % 2 == 0 &&
                        max_grid_size >= blocking_factor && max_grid_size % blocking_factor == 0,
                    MGX_ERR_ARG,
                    "cluster_tags: blocking_factor even, max_grid_size a multiple of it");
        MGX_REQUIRE(grid_eff > 0.0 && grid_eff <= 1.0, MGX_ERR_ARG,
                    "cluster_tags: grid_eff must lie in (0, 1]");
        const int bfc = blocking_factor / 2;  // coarse cells per block (ratio 2)
        int n[3], nc[3];
        for (int d = 0; d < 3; ++d) {
            n[d] = n_cell[d];
            MGX_REQUIRE(n[d] > 0 && n[d] % bfc == 0, MGX_ERR_ARG,
                        "cluster_tags: the coarse extents must be multiples of blocking_factor/2");
            nc[d] = n[d] / bfc;
        }
        // TagBoxArray::buffer (n_error_buf cells, periodic directions wrap)
        // and coarsening by the blocking factor: a block is tagged when any
        // of its buffered cells is
        std::vector<unsigned char> blk((size_t)nc[0] * nc[1] * nc[2], 0);
        const int nb = n_error_buf;
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) {
                    if (!tags[i + (long)n[0] * (j + (long)n[1] * k)]) continue;
                    for (int dk = -nb; dk <= nb; ++dk)
                        for (int dj = -nb; dj <= nb; ++dj)
                            for (int di = -nb; di <= nb; ++di) {
                                int p[3] = {i + di, j + dj, k + dk};
                                bool ok = true;
                                for (int d = 0; d < 3; ++d) {
                                    if (p[d] >= 0 && p[d] < n[d]) continue;
                                    if (is_periodic && is_periodic[d]) p[d] = (p[d] + n[d]) % n[d];
                                    else ok = false;
                                }
                                if (ok)
                                    blk[p[0] / bfc +
                                        (long)nc[0] * (p[1] / bfc + (long)nc[1] * (p[2] / bfc))] = 1;
                            }
                }
        TagField f{{nc[0], nc[1], nc[2]}, &blk};
        std::vector<IBox> clusters;
        cluster(f, IBox{{0, 0, 0}, {nc[0] - 1, nc[1] - 1, nc[2] - 1}}, grid_eff, clusters);
        // max_grid_size in blocks, then refine blocks to fine cells
        std::vector<IBox> chopped;
        for (const IBox& c : clusters) max_size(c, max_grid_size / blocking_factor, chopped);
        MGX_REQUIRE((int)chopped.size() <= max_boxes, MGX_ERR_ARG,
                    "cluster_tags: more boxes (" + std::to_string(chopped.size()) +
                        ") than max_boxes");
        for (size_t q = 0; q < chopped.size(); ++q)
            for (int d = 0; d < 3; ++d) {
                boxes[6 * q + d] = chopped[q].lo[d] * blocking_factor;
                boxes[6 * q + 3 + d] = (chopped[q].hi[d] + 1) * blocking_factor - 1;
            }
        *nboxes = (int)chopped.size();
        return MGX_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (const std::exception&) {
        return MGX_ERR_ARG;
    }
}

MGX_API int mgx_fill_fine_level(mgx_ctx ctx, const mgx_level_desc* crse, const double* crse_field,
                                int c_ng, int ncomp, const int* bcrec,
                                const mgx_level_desc* fine, double* fine_field, int f_ng,
                                const mgx_level_desc* old_fine, const double* old_field,
                                int of_ng) {
    if (!ctx) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        MGX_REQUIRE(crse_field && fine_field && ncomp >= 1 && ncomp <= kMaxInterpComp &&
                        c_ng >= 0 && f_ng >= 0 && of_ng >= 0,
                    MGX_ERR_ARG, "fill_fine_level: bad fields (1..8 components)");
        MGX_REQUIRE((old_fine == nullptr) == (old_field == nullptr), MGX_ERR_ARG,
                    "fill_fine_level: give the old fine level with its field, or neither");
        double dxc[3], dxf[3];
        const NodalGeom gc = level_geom(crse, dxc);
        const NodalGeom gf = level_geom(fine, dxf);
        bool bounded = false;
        for (int d = 0; d < 3; ++d) {
            MGX_REQUIRE(gf.n[d] == 2 * gc.n[d] && gf.per[d] == gc.per[d] && gc.n[d] >= 2,
                        MGX_ERR_ARG, "fill_fine_level: the fine level must refine the coarse by 2");
            bounded |= !gc.per[d];
        }
        MGX_REQUIRE(!bounded || c_ng >= 1, MGX_ERR_ARG,
                    "fill_fine_level: a bounded domain needs one filled coarse ghost cell");
        const KGeom kc = cell_kgeom(gc), kf = cell_kgeom(gf);
        hipStream_t s = ctx->stream;
        BoxLayout lc, lf, lo;
        lc.build(crse->boxes, crse->nboxes, gc, true);
        lf.build(fine->boxes, fine->nboxes, gf, true);
        {
            // the coarse level may itself be a fine level: it must cover every
            // coarse cell the interpolation reads (the parents of the fine
            // cells and their neighbours)
            std::vector<unsigned char> cov((size_t)gc.n[0] * gc.n[1] * gc.n[2], 0);
            for (int b = 0; b < crse->nboxes; ++b) {
                const int* q = crse->boxes + 6 * b;
                for (int k = q[2]; k <= q[5]; ++k)
                    for (int j = q[1]; j <= q[4]; ++j)
                        for (int i = q[0]; i <= q[3]; ++i)
                            cov[i + (size_t)gc.n[0] * (j + (size_t)gc.n[1] * k)] = 1;
            }
            for (int b = 0; b < fine->nboxes; ++b) {
                const int* q = fine->boxes + 6 * b;
                int lo3[3], hi3[3];
                for (int d = 0; d < 3; ++d) {
                    lo3[d] = q[d] / 2 - 1;
                    hi3[d] = q[3 + d] / 2 + 1;
                    if (!gc.per[d]) {
                        lo3[d] = std::max(lo3[d], 0);
                        hi3[d] = std::min(hi3[d], gc.n[d] - 1);
                    }
                }
                for (int k = lo3[2]; k <= hi3[2]; ++k)
                    for (int j = lo3[1]; j <= hi3[1]; ++j)
                        for (int i = lo3[0]; i <= hi3[0]; ++i) {
                            const int w[3] = {(i + gc.n[0]) % gc.n[0], (j + gc.n[1]) % gc.n[1],
                                              (k + gc.n[2]) % gc.n[2]};
                            MGX_REQUIRE(cov[w[0] + (size_t)gc.n[0] * (w[1] + (size_t)gc.n[1] * w[2])],
                                        MGX_ERR_ARG,
                                        "fill_fine_level: the fine level is not properly nested in "
                                        "the coarse level (a parent cell or its neighbour is missing)");
                        }
            }
        }
        DevBuf cg, dense, own;
        const long ng3 = (long)(gc.n[0] + 2) * (gc.n[1] + 2) * (gc.n[2] + 2);
        cg.alloc(ng3 * ncomp * sizeof(double));
        gather_ghosted_kernel<<<lc.ntiles[REG_CELL_GROWN], 256, 0, s>>>(
            kc, lc.dev_boxes(), lc.tiles(REG_CELL_GROWN), crse_field, c_ng, ncomp, cg.as<double>());
        wrap_ghosted_kernel<<<grid1d(ng3), 256, 0, s>>>(kc, ncomp, cg.as<double>());
        FillArgs a{};
        if (old_fine) {
            double dxo[3];
            const NodalGeom go = level_geom(old_fine, dxo);
            for (int d = 0; d < 3; ++d)
                MGX_REQUIRE(go.n[d] == gf.n[d], MGX_ERR_ARG,
                            "fill_fine_level: the old fine level has another domain");
            lo.build(old_fine->boxes, old_fine->nboxes, go, true);
            const long nf = (long)gf.n[0] * gf.n[1] * gf.n[2];
            dense.alloc(nf * ncomp * sizeof(double));
            own.alloc(nf);
            MGX_HIP(hipMemsetAsync(own.p, 0, nf, s));
            old_fine_kernel<<<lo.ntiles[REG_CELL], 256, 0, s>>>(
                kf, lo.dev_boxes(), lo.tiles(REG_CELL), old_field, of_ng, ncomp, dense.as<double>(),
                own.as<unsigned char>());
            a.dense = dense.as<double>();
            a.own = own.as<unsigned char>();
        }
        a.gf = kf;
        a.gc = kc;
        a.boxes = lf.dev_boxes();
        a.tiles = lf.tiles(REG_CELL);
        a.out = fine_field;
        a.ng = f_ng;
        a.ncomp = ncomp;
        a.cg = cg.as<double>();
        for (int q = 0; q < ncomp; ++q)
            for (int e = 0; e < 6; ++e) a.bc[q][e] = bcrec ? bcrec[6 * q + e] : 0;
        fill_fine_kernel<<<lf.ntiles[REG_CELL], 256, 0, s>>>(a);
        MGX_HIP(hipGetLastError());
        MGX_HIP(hipStreamSynchronize(s));  // the staging buffers are freed on return
    });
}

}  // extern "C"
