// io_api.hip -- AMReX-format plotfiles and checkpoints of device fields
// (host code: fields are copied from the device and written by the host).
//
// Formats (as incflo writes them through AMReX, src/utilities/io.cpp):
//  * VisMF multifab (VisMF::Write, "NFiles" form, version 1): <prefix>_H holds
//      1 \n 1 \n ncomp \n ngrow \n BoxArray \n nfabs \n "FabOnDisk: <file> <offset>" per
//      fab \n then the per-fab minima and maxima ("nfab,ncomp" and one line of
//      "v," per fab), and <prefix>_D_00000 holds per fab the FAB header
//      "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))<box> <ncomp>\n"
//      (little-endian IEEE doubles) followed by the fab's values, component
//      major, x fastest, over the box grown by ngrow.
//  * BoxArray::writeOn: "(<n> 0\n" then "((lo) (hi) (type))" per box and ")".
//  * Plotfile (WriteMultiLevelPlotfile, "HyperCLaw-V1.1") with Level_<l>/Cell.
//  * incflo's checkpoint Header (incflo::WriteHeader, src/utilities/io.cpp:16-68).
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "mgx_common.h"
#include "mgx_ctx.h"

using namespace mgx;

namespace {

const char* kRealDesc = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))";

struct Fab {
    int lo[3], hi[3];  // cell (or node) index box of the valid region
};

std::vector<Fab> fabs_of(const mgx_level_desc* lv, bool nodal) {
    std::vector<Fab> out(lv->nboxes);
    for (int b = 0; b < lv->nboxes; ++b)
        for (int d = 0; d < 3; ++d) {
            out[b].lo[d] = lv->boxes[6 * b + d];
            out[b].hi[d] = lv->boxes[6 * b + 3 + d] + (nodal ? 1 : 0);
        }
    return out;
}

long fab_size(const Fab& f, int ng) {
    long v = 1;
    for (int d = 0; d < 3; ++d) v *= (long)(f.hi[d] - f.lo[d] + 1 + 2 * ng);
    return v;
}

void mkdir_p(const std::string& path) {
    std::string cur;
    std::istringstream in(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(in, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            throw Error(MGX_ERR_ARG, "cannot create directory " + cur);
    }
}

void write_box(std::ostream& os, const int lo[3], const int hi[3], int type) {
    os << "((" << lo[0] << "," << lo[1] << "," << lo[2] << ") (" << hi[0] << "," << hi[1] << ","
       << hi[2] << ") (" << type << "," << type << "," << type << "))";
}

void write_boxarray(std::ostream& os, const std::vector<Fab>& fabs, int type) {
    os << "(" << fabs.size() << " 0\n";
    for (const Fab& f : fabs) {
        write_box(os, f.lo, f.hi, type);
        os << "\n";
    }
    os << ")";
}

std::vector<double> download(mgx_ctx_s* ctx, const double* d, long n) {
    std::vector<double> h(n);
    MGX_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MGX_HIP(hipStreamSynchronize(ctx->stream));
    return h;
}

// VisMF::Write of one boxed field (ngrow written = the field's ghost width)
void write_vismf(mgx_ctx_s* ctx, const std::string& prefix, const mgx_level_desc* lv,
                 const double* field, int ng, int ncomp, bool nodal) {
    const std::vector<Fab> fabs = fabs_of(lv, nodal);
    long total = 0;
    for (const Fab& f : fabs) total += ncomp * fab_size(f, ng);
    const std::vector<double> h = download(ctx, field, total);
    const std::string dname = prefix + "_D_00000";
    std::ofstream D(dname, std::ios::binary | std::ios::trunc);
    MGX_REQUIRE(D.good(), MGX_ERR_ARG, "cannot open " + dname);
    std::vector<long> offsets;
    std::vector<std::vector<double>> mins, maxs;
    long pos = 0;
    const int type = nodal ? 1 : 0;
    for (const Fab& f : fabs) {
        offsets.push_back((long)D.tellp());
        int glo[3], ghi[3];
        for (int d = 0; d < 3; ++d) { glo[d] = f.lo[d] - ng; ghi[d] = f.hi[d] + ng; }
        std::ostringstream hdr;
        hdr << "FAB " << kRealDesc;
        write_box(hdr, glo, ghi, type);
        hdr << " " << ncomp << "\n";
        D << hdr.str();
        const long n = fab_size(f, ng);
        D.write(reinterpret_cast<const char*>(h.data() + pos), ncomp * n * sizeof(double));
        // min / max of each component over the valid region
        std::vector<double> mn(ncomp, std::numeric_limits<double>::max()),
            mx(ncomp, -std::numeric_limits<double>::max());
        const int e0 = f.hi[0] - f.lo[0] + 1 + 2 * ng, e1 = f.hi[1] - f.lo[1] + 1 + 2 * ng;
        for (int c = 0; c < ncomp; ++c)
            for (int k = ng; k <= f.hi[2] - f.lo[2] + ng; ++k)
                for (int j = ng; j <= f.hi[1] - f.lo[1] + ng; ++j)
                    for (int i = ng; i <= f.hi[0] - f.lo[0] + ng; ++i) {
                        const double v = h[pos + c * n + i + (long)e0 * (j + (long)e1 * k)];
                        mn[c] = std::min(mn[c], v);
                        mx[c] = std::max(mx[c], v);
                    }
        mins.push_back(mn);
        maxs.push_back(mx);
        pos += ncomp * n;
    }
    D.close();
    MGX_REQUIRE(!D.fail(), MGX_ERR_ARG, "write failed: " + dname);
    std::ofstream H(prefix + "_H", std::ios::trunc);
    MGX_REQUIRE(H.good(), MGX_ERR_ARG, "cannot open " + prefix + "_H");
    H << std::setprecision(17);
    H << 1 << "\n" << 1 << "\n" << ncomp << "\n" << ng << "\n";
    std::vector<Fab> vb = fabs;
    write_boxarray(H, vb, type);
    H << "\n" << fabs.size() << "\n";
    const std::string base = dname.substr(dname.find_last_of('/') + 1);
    for (size_t b = 0; b < fabs.size(); ++b) H << "FabOnDisk: " << base << " " << offsets[b] << "\n";
    H << "\n";
    for (const auto* mm : {&mins, &maxs}) {
        H << fabs.size() << "," << ncomp << "\n";
        for (const auto& v : *mm) {
            for (double x : v) H << x << ",";
            H << "\n";
        }
        H << "\n";
    }
    H.close();
    MGX_REQUIRE(!H.fail(), MGX_ERR_ARG, "write failed: " + prefix + "_H");
}

// Box "((lo) (hi) (type))" from a stream
bool read_box(std::istream& is, int lo[3], int hi[3], int& type) {
    char c;
    int t[3];
    is >> c;  // (
    if (c != '(') return false;
    is >> c >> lo[0] >> c >> lo[1] >> c >> lo[2] >> c;
    is >> c >> hi[0] >> c >> hi[1] >> c >> hi[2] >> c;
    is >> c >> t[0] >> c >> t[1] >> c >> t[2] >> c;
    is >> c;  // )
    type = t[0];
    return !is.fail();
}

// VisMF::Read of one field into a boxed device field of the same box list
void read_vismf(mgx_ctx_s* ctx, const std::string& prefix, const mgx_level_desc* lv, double* field,
                int ng, int ncomp, bool nodal) {
    std::ifstream H(prefix + "_H");
    MGX_REQUIRE(H.good(), MGX_ERR_ARG, "cannot open " + prefix + "_H");
    int vers, how, nc, ngf;
    H >> vers >> how >> nc >> ngf;
    MGX_REQUIRE(!H.fail() && vers == 1 && nc == ncomp, MGX_ERR_ARG,
                prefix + "_H: expected a version-1 multifab with " + std::to_string(ncomp) +
                    " components");
    char c;
    long nb;
    int hash;
    H >> c >> nb >> hash;
    const std::vector<Fab> fabs = fabs_of(lv, nodal);
    MGX_REQUIRE(c == '(' && nb == (long)fabs.size(), MGX_ERR_ARG,
                prefix + "_H: box count differs from the level's");
    for (long b = 0; b < nb; ++b) {
        int lo[3], hi[3], type;
        MGX_REQUIRE(read_box(H, lo, hi, type), MGX_ERR_ARG, prefix + "_H: bad box");
        for (int d = 0; d < 3; ++d)
            MGX_REQUIRE(lo[d] == fabs[b].lo[d] && hi[d] == fabs[b].hi[d], MGX_ERR_ARG,
                        prefix + "_H: box " + std::to_string(b) + " differs from the level's");
    }
    H >> c;  // )
    long nfod;
    H >> nfod;
    MGX_REQUIRE(nfod == nb, MGX_ERR_ARG, prefix + "_H: FabOnDisk count");
    std::vector<std::string> files(nb);
    std::vector<long> offs(nb);
    for (long b = 0; b < nb; ++b) {
        std::string tag;
        H >> tag >> files[b] >> offs[b];
        MGX_REQUIRE(tag == "FabOnDisk:", MGX_ERR_ARG, prefix + "_H: bad FabOnDisk line");
    }
    const std::string dir = prefix.substr(0, prefix.find_last_of('/') + 1);
    long total = 0;
    for (const Fab& f : fabs) total += ncomp * fab_size(f, ng);
    std::vector<double> h(total, 0.0);
    long pos = 0;
    for (long b = 0; b < nb; ++b) {
        std::ifstream D(dir + files[b], std::ios::binary);
        MGX_REQUIRE(D.good(), MGX_ERR_ARG, "cannot open " + dir + files[b]);
        D.seekg(offs[b]);
        std::string line;
        std::getline(D, line);
        MGX_REQUIRE(line.rfind("FAB ", 0) == 0 && line.find(kRealDesc) != std::string::npos,
                    MGX_ERR_UNSUPPORTED, "FAB header is not little-endian IEEE double: " + line);
        const Fab& f = fabs[b];
        const long nf = fab_size(f, ngf), nt = fab_size(f, ng);
        std::vector<double> buf(ncomp * nf);
        D.read(reinterpret_cast<char*>(buf.data()), buf.size() * sizeof(double));
        MGX_REQUIRE(!D.fail(), MGX_ERR_ARG, "short read of " + files[b]);
        // copy the common region (valid cells plus min(ng, ngf) ghosts)
        const int gm = std::min(ng, ngf);
        const int ef0 = f.hi[0] - f.lo[0] + 1 + 2 * ngf, ef1 = f.hi[1] - f.lo[1] + 1 + 2 * ngf;
        const int et0 = f.hi[0] - f.lo[0] + 1 + 2 * ng, et1 = f.hi[1] - f.lo[1] + 1 + 2 * ng;
        for (int q = 0; q < ncomp; ++q)
            for (int k = -gm; k <= f.hi[2] - f.lo[2] + gm; ++k)
                for (int j = -gm; j <= f.hi[1] - f.lo[1] + gm; ++j)
                    for (int i = -gm; i <= f.hi[0] - f.lo[0] + gm; ++i)
                        h[pos + q * nt + (i + ng) + (long)et0 * ((j + ng) + (long)et1 * (k + ng))] =
                            buf[q * nf + (i + ngf) + (long)ef0 * ((j + ngf) + (long)ef1 * (k + ngf))];
        pos += ncomp * nt;
    }
    MGX_HIP(hipMemcpyAsync(field, h.data(), total * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream));
    MGX_HIP(hipStreamSynchronize(ctx->stream));
}

std::string level_dir(const std::string& dir, int lev) {
    return dir + "/Level_" + std::to_string(lev);
}

}  // namespace

extern "C" {

MGX_API int mgx_write_multifab(mgx_ctx ctx, const char* prefix, const mgx_level_desc* level,
                               const double* field, int ng, int ncomp, int nodal) {
    if (!ctx) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        MGX_REQUIRE(prefix && level && field && ng >= 0 && ncomp >= 1, MGX_ERR_ARG,
                    "write_multifab: bad arguments");
        const std::string p(prefix);
        const size_t slash = p.find_last_of('/');
        if (slash != std::string::npos) mkdir_p(p.substr(0, slash));
        write_vismf(ctx, p, level, field, ng, ncomp, nodal != 0);
    });
}

MGX_API int mgx_read_multifab(mgx_ctx ctx, const char* prefix, const mgx_level_desc* level,
                              double* field, int ng, int ncomp, int nodal) {
    if (!ctx) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        MGX_REQUIRE(prefix && level && field && ng >= 0 && ncomp >= 1, MGX_ERR_ARG,
                    "read_multifab: bad arguments");
        read_vismf(ctx, prefix, level, field, ng, ncomp, nodal != 0);
    });
}

MGX_API int mgx_write_checkpoint_header(const char* dir, int nlevels, const mgx_level_desc* levels,
                                        int nstep, double cur_time, double dt, double prev_dt,
                                        double prev_prev_dt) {
    try {
        MGX_REQUIRE(dir && levels && nlevels >= 1, MGX_ERR_ARG, "checkpoint header: bad arguments");
        mkdir_p(dir);
        std::ofstream H(std::string(dir) + "/Header", std::ios::trunc);
        MGX_REQUIRE(H.good(), MGX_ERR_ARG, "cannot open the checkpoint Header");
        H << std::setprecision(17);
        H << "Checkpoint version: 1\n" << nlevels - 1 << "\n" << nstep << "\n" << cur_time << "\n"
          << dt << "\n" << prev_dt << "\n" << prev_prev_dt << "\n";
        for (int d = 0; d < 3; ++d) H << levels[0].prob_lo[d] << ' ';
        H << '\n';
        for (int d = 0; d < 3; ++d) H << levels[0].prob_hi[d] << ' ';
        H << '\n';
        for (int l = 0; l < nlevels; ++l) {
            write_boxarray(H, fabs_of(&levels[l], false), 0);
            H << '\n';
        }
        H.close();
        MGX_REQUIRE(!H.fail(), MGX_ERR_ARG, "checkpoint Header write failed");
        return MGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

MGX_API int mgx_read_checkpoint_header(const char* dir, int* finest_level, int* nstep,
                                       double* times /* cur_time, dt, prev_dt, prev_prev_dt */,
                                       double* prob_lo, double* prob_hi, int* nboxes_per_level,
                                       int max_levels, int* boxes, int max_boxes) {
    try {
        MGX_REQUIRE(dir && finest_level && nstep && times && prob_lo && prob_hi, MGX_ERR_ARG,
                    "checkpoint header: bad arguments");
        std::ifstream H(std::string(dir) + "/Header");
        MGX_REQUIRE(H.good(), MGX_ERR_ARG, "cannot open the checkpoint Header");
        std::string line;
        std::getline(H, line);
        MGX_REQUIRE(line == "Checkpoint version: 1", MGX_ERR_ARG, "not an incflo checkpoint Header");
        H >> *finest_level >> *nstep >> times[0] >> times[1] >> times[2] >> times[3];
        for (int d = 0; d < 3; ++d) H >> prob_lo[d];
        for (int d = 0; d < 3; ++d) H >> prob_hi[d];
        MGX_REQUIRE(!H.fail() && *finest_level >= 0 && *finest_level < max_levels, MGX_ERR_ARG,
                    "checkpoint Header: bad level count");
        int used = 0;
        for (int l = 0; l <= *finest_level; ++l) {
            char c;
            int nb, hash;
            H >> c >> nb >> hash;
            MGX_REQUIRE(c == '(' && nb >= 0 && used + nb <= max_boxes, MGX_ERR_ARG,
                        "checkpoint Header: box list does not fit");
            nboxes_per_level[l] = nb;
            for (int b = 0; b < nb; ++b) {
                int lo[3], hi[3], type;
                MGX_REQUIRE(read_box(H, lo, hi, type), MGX_ERR_ARG, "checkpoint Header: bad box");
                for (int d = 0; d < 3; ++d) {
                    boxes[6 * (used + b) + d] = lo[d];
                    boxes[6 * (used + b) + 3 + d] = hi[d];
                }
            }
            H >> c;
            used += nb;
        }
        return MGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

MGX_API int mgx_write_plotfile(mgx_ctx ctx, const char* dir, int nlevels,
                               const mgx_level_desc* levels, const double* const* data, int ng,
                               int nvars, const char* const* varnames, double time,
                               const int* level_steps, const int* ref_ratio) {
    if (!ctx) return MGX_ERR_ARG;
    return guarded(ctx, [&] {
        MGX_REQUIRE(dir && levels && data && varnames && nlevels >= 1 && nvars >= 1 && ng >= 0,
                    MGX_ERR_ARG, "write_plotfile: bad arguments");
        const std::string d(dir);
        mkdir_p(d);
        for (int l = 0; l < nlevels; ++l) {
            mkdir_p(level_dir(d, l));
            write_vismf(ctx, level_dir(d, l) + "/Cell", &levels[l], data[l], ng, nvars, false);
        }
        std::ofstream H(d + "/Header", std::ios::trunc);
        MGX_REQUIRE(H.good(), MGX_ERR_ARG, "cannot open the plotfile Header");
        H << std::setprecision(17);
        H << "HyperCLaw-V1.1\n" << nvars << "\n";
        for (int v = 0; v < nvars; ++v) H << varnames[v] << "\n";
        H << 3 << "\n" << time << "\n" << nlevels - 1 << "\n";
        for (int q = 0; q < 3; ++q) H << levels[0].prob_lo[q] << ' ';
        H << "\n";
        for (int q = 0; q < 3; ++q) H << levels[0].prob_hi[q] << ' ';
        H << "\n";
        for (int l = 0; l + 1 < nlevels; ++l) H << (ref_ratio ? ref_ratio[l] : 2) << ' ';
        H << "\n";
        for (int l = 0; l < nlevels; ++l) {
            const int lo[3] = {0, 0, 0};
            const int hi[3] = {levels[l].n_cell[0] - 1, levels[l].n_cell[1] - 1,
                               levels[l].n_cell[2] - 1};
            write_box(H, lo, hi, 0);
            H << ' ';
        }
        H << "\n";
        for (int l = 0; l < nlevels; ++l) H << (level_steps ? level_steps[l] : 0) << ' ';
        H << "\n";
        for (int l = 0; l < nlevels; ++l) {
            for (int q = 0; q < 3; ++q)
                H << (levels[l].prob_hi[q] - levels[l].prob_lo[q]) / levels[l].n_cell[q] << ' ';
            H << "\n";
        }
        H << 0 << "\n" << 0 << "\n";  // coordinate system (Cartesian), boundary width
        for (int l = 0; l < nlevels; ++l) {
            const mgx_level_desc& lv = levels[l];
            H << l << ' ' << lv.nboxes << ' ' << time << "\n";
            H << (level_steps ? level_steps[l] : 0) << "\n";
            for (int b = 0; b < lv.nboxes; ++b)
                for (int q = 0; q < 3; ++q) {
                    const double h = (lv.prob_hi[q] - lv.prob_lo[q]) / lv.n_cell[q];
                    H << lv.prob_lo[q] + h * lv.boxes[6 * b + q] << ' '
                      << lv.prob_lo[q] + h * (lv.boxes[6 * b + 3 + q] + 1) << "\n";
                }
            H << "Level_" << l << "/Cell\n";
        }
        H.close();
        MGX_REQUIRE(!H.fail(), MGX_ERR_ARG, "plotfile Header write failed");
    });
}

}  // extern "C"
