// islands.cpp -- host-side island setup of the streamfunction solver.
//
// Replaces the island part of streamfunction_init
// (for_src/external/streamfct_init.f90:35-267) and the MOM "isleperim"
// flood fill (for_src/external/island_module.f90:7-273).  It runs once per
// basin on the CPU, as in the reference; the perimeter lists it produces
// define the solver's Dirichlet set and the island line integrals.  Unlike
// the reference there are no fixed limits (maxq, maxipp, mnisle): the queue
// and the lists grow as needed.
#include <stdint.h>

#include <deque>
#include <utility>
#include <vector>

#include "../../include/psisolve.h"

namespace {

// Map indices: (I, J) 1-based over [1, imt] x [1, jmt] (I = i + 2 for the
// Fortran index i in [-1, nx+2]).
struct LandMap {
    int imt, jmt;
    std::vector<int> v;
    int& at(int I, int J) { return v[(size_t)(J - 1) * imt + (I - 1)]; }
};

// isleperim/expand: label 8-connected land masses 1..nisle and their ocean
// perimeter cells -1..-nisle; other ocean 0.  Land masses are discovered
// scanning J from north to south, I from west to east (island_module.f90:78-95),
// so a perimeter cell shared by two masses belongs to the first discovered.
int label_land(const int32_t* kmt, int imt, int jmt, bool cyclic, LandMap& m) {
    constexpr int OFF = -1;
    m.imt = imt;
    m.jmt = jmt;
    m.v.assign((size_t)imt * jmt, 0);
    for (int J = 1; J <= jmt; ++J)
        for (int I = 1; I <= imt; ++I) m.at(I, J) = (kmt[(size_t)(J - 1) * imt + (I - 1)] > 0) ? 0 : 1;
    const int iwest = cyclic ? 2 : 1, ieast = cyclic ? imt - 1 : imt;
    auto east = [&](int i) {
        if (cyclic) return (i < imt - 1) ? i + 1 : (i + 1) - imt + 2;
        return (i < imt) ? i + 1 : OFF;
    };
    auto west = [&](int i) {
        if (cyclic) return (i > 2) ? i - 1 : (i - 1) + imt - 2;
        return (i > 1) ? i - 1 : OFF;
    };
    int label = 2;
    std::deque<std::pair<int, int>> q;
    for (int J = jmt; J >= 1; --J) {
        for (int I = iwest; I <= ieast; ++I) {
            if (m.at(I, J) != 1) continue;
            q.clear();
            q.emplace_back(I, J);
            while (!q.empty()) {
                const auto [i, j] = q.front();
                q.pop_front();
                if (i == OFF || j == OFF) continue;
                int& cell = m.at(i, j);
                if (cell == label || cell == -label) continue;
                if (cell == 1) {
                    cell = label;
                    const int jn = (j < jmt) ? j + 1 : OFF, js = (j > 1) ? j - 1 : OFF;
                    const int ie = east(i), iw = west(i);
                    // neighbour order of expand (island_module.f90:207-214)
                    q.emplace_back(i, jn);
                    q.emplace_back(ie, jn);
                    q.emplace_back(ie, j);
                    q.emplace_back(ie, js);
                    q.emplace_back(i, js);
                    q.emplace_back(iw, js);
                    q.emplace_back(iw, j);
                    q.emplace_back(iw, jn);
                } else if (cell == 0) {
                    cell = -label;
                }
                // cell < 0: perimeter of an earlier mass; cell > 1: another mass
            }
            ++label;
        }
    }
    for (int I = iwest; I <= ieast; ++I)
        for (int J = 1; J <= jmt; ++J) {
            int& c = m.at(I, J);
            if (c > 0) c -= 1;
            else if (c < 0) c += 1;
        }
    if (cyclic)
        for (int J = 1; J <= jmt; ++J) {
            m.at(1, J) = m.at(imt - 1, J);
            m.at(imt, J) = m.at(2, J);
        }
    return label - 2;
}

struct Walk {
    std::vector<int32_t> bi, bj, di, dj;
};

// Perimeter walk of one island on its own map (1 land, -1 perimeter,
// 0 ocean): start-point rule streamfct_init.f90:116-165, wall-following
// rules :184-258.  Returns 0, or -1 (no start point) / -2 (lost track).
int walk_island(LandMap& m, int nx, int ny, bool cyclic, Walk& w) {
    auto M = [&](int i, int j) -> int {
        if (i < -1 || i > nx + 2 || j < -1 || j > ny + 2) return 0x7fffffff;
        return m.at(i + 2, j + 2);
    };
    int ij0 = 0, ij1 = 0, d0 = 0, d1 = 0;
    bool found = false;
    auto scan = [&](int i) {
        for (int j = 0; j <= ny; ++j) {
            if (M(i, j) == 1 && M(i, j + 1) == -1) {  // eastward, from the west
                ij0 = i; ij1 = j; d0 = 1; d1 = 0;
                w.bi.push_back(ij0 - 1); w.bj.push_back(ij1);
                return true;
            }
            if (M(i, j) == -1 && M(i, j + 1) == 1) {  // westward, from the east
                ij0 = i - 1; ij1 = j; d0 = -1; d1 = 0;
                w.bi.push_back(ij0 + 1); w.bj.push_back(ij1);
                return true;
            }
        }
        return false;
    };
    for (int i = nx / 2 + 1; i <= nx && !found; ++i) found = scan(i);
    for (int i = nx / 2; i >= 1 && !found; --i) found = scan(i);
    if (!found) return -1;
    w.di.push_back(d0);
    w.dj.push_back(d1);
    w.bi.push_back(ij0);
    w.bj.push_back(ij1);
    const int si = w.bi[0], sj = w.bj[0];
    for (;;) {
        int p0, p1, q0, q1;   // cell ahead, cell ahead-right
        if (d0 == 0 && d1 == 1)       { p0 = ij0;     p1 = ij1 + 1; q0 = ij0 + 1; q1 = ij1 + 1; }
        else if (d0 == -1 && d1 == 0) { p0 = ij0;     p1 = ij1;     q0 = ij0;     q1 = ij1 + 1; }
        else if (d0 == 0 && d1 == -1) { p0 = ij0 + 1; p1 = ij1;     q0 = ij0;     q1 = ij1; }
        else                          { p0 = ij0 + 1; p1 = ij1 + 1; q0 = ij0 + 1; q1 = ij1; }
        const int a = M(p0, p1), b = M(q0, q1);
        if (a == -1 && b == 1) {
            // straight on
        } else if (a == -1 && b == -1) {
            const int t = d0; d0 = d1; d1 = -t;        // turn right
        } else if (a == 1 && (b == 1 || b == -1)) {
            const int t = d0; d0 = -d1; d1 = t;        // turn left
        } else {
            return -2;
        }
        w.di.push_back(d0);
        w.dj.push_back(d1);
        ij0 += d0;
        ij1 += d1;
        bool cont = !(si == ij0 && sj == ij1);
        if (cyclic && d0 == 1 && d1 == 0 && ij0 > nx) ij0 -= nx;
        if (cyclic && d0 == -1 && d1 == 0 && ij0 < 1) ij0 += nx;
        if (si == ij0 && sj == ij1) cont = false;
        if (!cont) break;
        w.bi.push_back(ij0);
        w.bj.push_back(ij1);
        if (w.bi.size() > (size_t)8 * (nx + 4) * (ny + 4)) return -2;
    }
    // the direction list has one entry per point
    w.di.resize(w.bi.size());
    w.dj.resize(w.bi.size());
    return 0;
}

}  // namespace

extern "C" int psi_islands_trace(int32_t nx, int32_t ny, int32_t cyclic_x, const int32_t* kbot,
                                 int32_t* nisle, int32_t* offsets, int32_t cap_isles, int32_t* bi,
                                 int32_t* bj, int32_t* di, int32_t* dj, int32_t cap_points,
                                 int32_t* npoints) {
    if (nx < 1 || ny < 1 || !kbot || !nisle || !npoints) return PSI_BAD_ARG;
    const int imt = nx + 4, jmt = ny + 4;
    const bool cyclic = cyclic_x != 0;
    // kmt = 5 on interior ocean, 0 elsewhere; cyclic halo copy
    // (streamfct_init.f90:35-48)
    std::vector<int32_t> kmt((size_t)imt * jmt, 0);
    for (int J = 3; J <= ny + 2; ++J)
        for (int I = 3; I <= nx + 2; ++I)
            if (kbot[(size_t)(J - 1) * imt + (I - 1)] > 0) kmt[(size_t)(J - 1) * imt + (I - 1)] = 5;
    auto copy_cyclic = [&](std::vector<int32_t>& a) {
        for (int J = 1; J <= jmt; ++J)
            for (int k = 1; k <= 2; ++k) {
                a[(size_t)(J - 1) * imt + (nx + k + 2 - 1)] = a[(size_t)(J - 1) * imt + (k + 2 - 1)];
                a[(size_t)(J - 1) * imt + (1 - k + 2 - 1)] = a[(size_t)(J - 1) * imt + (nx - k + 1 + 2 - 1)];
            }
    };
    if (cyclic) copy_cyclic(kmt);
    LandMap all;
    const int n = label_land(kmt.data(), imt, jmt, cyclic, all);
    if (cyclic) copy_cyclic(all.v);
    std::vector<Walk> walks((size_t)n);
    size_t total = 0;
    for (int isle = 1; isle <= n; ++isle) {
        std::vector<int32_t> k1((size_t)imt * jmt);
        for (size_t q = 0; q < k1.size(); ++q) k1[q] = (all.v[q] == isle) ? 0 : 1;
        LandMap one;
        label_land(k1.data(), imt, jmt, cyclic, one);
        if (walk_island(one, nx, ny, cyclic, walks[(size_t)isle - 1]) != 0) return PSI_TRACE_FAILED;
        total += walks[(size_t)isle - 1].bi.size();
    }
    *nisle = n;
    *npoints = (int32_t)total;
    if (cap_isles < n + 1 || (size_t)cap_points < total || !offsets || !bi || !bj || !di || !dj)
        return PSI_BAD_ARG;
    size_t o = 0;
    offsets[0] = 0;
    for (int k = 0; k < n; ++k) {
        const Walk& w = walks[(size_t)k];
        for (size_t q = 0; q < w.bi.size(); ++q, ++o) {
            bi[o] = w.bi[q];
            bj[o] = w.bj[q];
            di[o] = w.di[q];
            dj[o] = w.dj[q];
        }
        offsets[k + 1] = (int32_t)o;
    }
    return PSI_OK;
}
