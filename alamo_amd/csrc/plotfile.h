// AMReX plotfile writer for nodal fields (Integrator::WritePlotFile, Integrator.cpp:808-1054, which
// hands the nodal MultiFabs of every level to amrex::WriteMultiLevelPlotfile as "<prefix>node"
// and then writes the level box arrays to "<prefix>node/Checkpoint").
//
// Layout written (the HyperCLaw-V1.1 plotfile of AMReX, one I/O rank):
//   <dir>/Header              version, component names, dimension, time, finest level, problem
//                             extent, refinement ratios, level domains, level steps, cell sizes,
//                             coordinate system, boundary width, then per level its boxes'
//                             physical extents and the path of its multifab header;
//   <dir>/Level_<l>/Cell_H    the VisMF header: version, layout, components, ghost width, the
//                             nodal box array, the (file, offset) of every box's data, and the
//                             per-box, per-component minima and maxima;
//   <dir>/Level_<l>/Cell_D_00000  per box a FAB header line naming the real format (IEEE double,
//                             little-endian) and the box, then the box's values component by
//                             component, x fastest;
//   <dir>/Checkpoint          the cell box array of every level.
// Host code only; the caller gives host copies of the fields.
#pragma once
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace alamo {
namespace plot {

struct CellBox {
    int lo[3], hi[3];  // inclusive cell indices in the level's index space
};

inline std::string ivec(const int* v) {
    return "(" + std::to_string(v[0]) + "," + std::to_string(v[1]) + "," + std::to_string(v[2]) + ")";
}
// ((lo) (hi) (type)) of a box; type 0 cell-centred, 1 nodal in every direction
inline std::string box_str(const int* lo, const int* hi, int type) {
    const int t[3] = {type, type, type};
    return "(" + ivec(lo) + " " + ivec(hi) + " " + ivec(t) + ")";
}
inline std::string box_array_str(const std::vector<CellBox>& bs, int type) {
    std::string s = "(" + std::to_string(bs.size()) + " 0\n";
    for (const CellBox& b : bs) {
        int hi[3];
        for (int d = 0; d < 3; d++) hi[d] = b.hi[d] + type;  // nodal box: one more node than cells
        s += box_str(b.lo, hi, type) + "\n";
    }
    return s + ")";
}
inline std::string real_str(double v) {
    char buf[40];
    snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}
inline bool make_dir(const std::string& p) { return mkdir(p.c_str(), 0755) == 0 || errno == EEXIST; }

// One level: n_cell cells per direction, its boxes, and ncomp SoA planes of its (n + 1)^3 nodes.
struct Level {
    int n_cell[3];
    double dx[3];
    std::vector<CellBox> boxes;
    const double* data;  // host
};

// Returns an empty string on success, else the error.
inline std::string write_node_plotfile(const std::string& dir, const std::vector<Level>& L, int ncomp,
                                       const std::vector<std::string>& names, const double* prob_lo,
                                       const double* prob_hi, double time, const std::vector<int>& steps) {
    const int nl = (int)L.size();
    if (!make_dir(dir)) return "cannot create " + dir;
    std::string H = "HyperCLaw-V1.1\n" + std::to_string(ncomp) + "\n";
    for (const std::string& s : names) H += s + "\n";
    H += "3\n" + real_str(time) + "\n" + std::to_string(nl - 1) + "\n";
    for (int d = 0; d < 3; d++) H += real_str(prob_lo[d]) + " ";
    H += "\n";
    for (int d = 0; d < 3; d++) H += real_str(prob_hi[d]) + " ";
    H += "\n";
    for (int l = 0; l + 1 < nl; l++) H += std::to_string(L[l + 1].n_cell[0] / L[l].n_cell[0]) + " ";
    H += "\n";
    for (int l = 0; l < nl; l++) {
        const int lo[3] = {0, 0, 0};
        const int hi[3] = {L[l].n_cell[0] - 1, L[l].n_cell[1] - 1, L[l].n_cell[2] - 1};
        H += box_str(lo, hi, 0) + " ";
    }
    H += "\n";
    for (int l = 0; l < nl; l++) H += std::to_string(steps[l]) + " ";
    H += "\n";
    for (int l = 0; l < nl; l++) {
        for (int d = 0; d < 3; d++) H += real_str(L[l].dx[d]) + " ";
        H += "\n";
    }
    H += "0\n0\n";  // Cartesian coordinates, no boundary layer
    for (int l = 0; l < nl; l++) {
        const Level& lv = L[l];
        H += std::to_string(l) + " " + std::to_string(lv.boxes.size()) + " " + real_str(time) + "\n";
        H += std::to_string(steps[l]) + "\n";
        for (const CellBox& b : lv.boxes)
            for (int d = 0; d < 3; d++)  // a nodal box spans lo dx .. hi_node dx = (hi + 1) dx
                H += real_str(prob_lo[d] + lv.dx[d] * b.lo[d]) + " " + real_str(prob_lo[d] + lv.dx[d] * (b.hi[d] + 1)) +
                     "\n";
        H += "Level_" + std::to_string(l) + "/Cell\n";

        const std::string ldir = dir + "/Level_" + std::to_string(l);
        if (!make_dir(ldir)) return "cannot create " + ldir;
        const std::string dname = "Cell_D_00000";
        FILE* f = fopen((ldir + "/" + dname).c_str(), "wb");
        if (!f) return "cannot write " + ldir + "/" + dname;
        std::vector<long> offs;
        std::vector<std::vector<double>> mins, maxs;
        const int64_t N0 = lv.n_cell[0] + 1, N1 = lv.n_cell[1] + 1, tot = N0 * N1 * (lv.n_cell[2] + 1);
        std::vector<double> buf;
        for (const CellBox& b : lv.boxes) {
            offs.push_back(ftell(f));
            const int hi[3] = {b.hi[0] + 1, b.hi[1] + 1, b.hi[2] + 1};
            const std::string fh = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))" +
                                   box_str(b.lo, hi, 1) + " " + std::to_string(ncomp) + "\n";
            fwrite(fh.data(), 1, fh.size(), f);
            std::vector<double> mn(ncomp, 0.0), mx(ncomp, 0.0);
            for (int c = 0; c < ncomp; c++) {
                buf.clear();
                for (int k = b.lo[2]; k <= hi[2]; k++)
                    for (int j = b.lo[1]; j <= hi[1]; j++) {
                        const double* row = lv.data + c * tot + N0 * (j + N1 * (int64_t)k);
                        buf.insert(buf.end(), row + b.lo[0], row + hi[0] + 1);
                    }
                mn[c] = mx[c] = buf[0];
                for (double v : buf) {
                    mn[c] = v < mn[c] ? v : mn[c];
                    mx[c] = v > mx[c] ? v : mx[c];
                }
                fwrite(buf.data(), sizeof(double), buf.size(), f);
            }
            mins.push_back(mn);
            maxs.push_back(mx);
        }
        if (fclose(f) != 0) return "write error in " + ldir + "/" + dname;
        std::string V = "1\n0\n" + std::to_string(ncomp) + "\n0\n" + box_array_str(lv.boxes, 1) + "\n" +
                        std::to_string(lv.boxes.size()) + "\n";
        for (long o : offs) V += "FabOnDisk: " + dname + " " + std::to_string(o) + "\n";
        V += "\n";
        for (int pass = 0; pass < 2; pass++) {
            const auto& mm = pass == 0 ? mins : maxs;
            V += std::to_string(mm.size()) + "," + std::to_string(ncomp) + "\n";
            for (const auto& r : mm) {
                for (double v : r) V += real_str(v) + ",";
                V += "\n";
            }
            if (pass == 0) V += "\n";
        }
        FILE* hf = fopen((ldir + "/Cell_H").c_str(), "w");
        if (!hf || fwrite(V.data(), 1, V.size(), hf) != V.size() || fclose(hf) != 0)
            return "cannot write " + ldir + "/Cell_H";
    }
    FILE* hf = fopen((dir + "/Header").c_str(), "w");
    if (!hf || fwrite(H.data(), 1, H.size(), hf) != H.size() || fclose(hf) != 0) return "cannot write " + dir + "/Header";
    std::string C;
    for (const Level& lv : L) C += box_array_str(lv.boxes, 0);
    FILE* cf = fopen((dir + "/Checkpoint").c_str(), "w");
    if (!cf || fwrite(C.data(), 1, C.size(), cf) != C.size() || fclose(cf) != 0)
        return "cannot write " + dir + "/Checkpoint";
    return "";
}

}  // namespace plot
}  // namespace alamo
