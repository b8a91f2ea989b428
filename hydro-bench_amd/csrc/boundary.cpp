// boundary.cpp -- boundary-series and rating-curve preparation on the host:
// MeshData::take_boundary_for_two_d and BOUNDRYinterp
// (cuda_native_impl/hydro-cal-src/src/mesh.cpp:374-533), with the expansion
// from boundary entries to per-cell arrays of CellView's constructor
// (cuda_native_impl/hydro-cal-src/src/cell.cpp:92-126), in Real = double or
// float.  The series files are read the way the reference reads them (a
// line, then `>>` into Real), so an fp32 build parses its numbers straight to
// float.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "prep.hpp"

namespace hydro {
namespace {

template <class R> struct Table {
    std::vector<R> a, b;   // first and second column
};

std::string series_path(const std::string &dir, const char *kind, int k) {
    char name[64];
    std::snprintf(name, sizeof(name), "%s/%s%04d.DAT", kind, kind, k);
    return dir + "/" + name;
}

// A count on the first line, then `n` lines of two numbers; a short line
// leaves 0 in what it lacks, as `>>` does (mesh.cpp:389-400, 421-437, 478-492).
template <class R> bool read_table(const std::string &path, int cap, Table<R> &t, std::string &err) {
    std::ifstream f(path);
    if (!f.is_open()) {
        err = "cannot open " + path;
        return false;
    }
    std::string line;
    std::getline(f, line);
    std::istringstream head(line);
    int n = 0;
    if (!(head >> n) || n < 0) {
        err = "parse error on the first line of " + path;
        return false;
    }
    if (cap >= 0 && n > cap) {
        err = path + " holds " + std::to_string(n) + " points, more than NHQ = " + std::to_string(cap);
        return false;
    }
    t.a.assign(n, R(0));
    t.b.assign(n, R(0));
    for (int i = 0; i < n; i++) {
        std::getline(f, line);
        std::istringstream iss(line);
        iss >> t.a[i];
        iss >> t.b[i];
    }
    return true;
}

// BOUNDRYinterp (mesh.cpp:523-533): the last bracket [t_{i-1}, t_i] holding
// `t` gives the line through its end points, in Real; 0 when none does.  The
// reference's loop also tests the bracket ending one entry past the table,
// reading outside it; that bracket is not tested here.
template <class R> R interp(R t, const Table<R> &s) {
    R r = 0;
    const int n = (int)s.a.size();
    for (int i = 1; i < n; i++)
        if (t >= s.a[i - 1] && t <= s.a[i])
            r = s.b[i - 1] + (s.b[i] - s.b[i - 1]) / (s.a[i] - s.a[i - 1]) * (t - s.a[i - 1]);
    return r;
}

// Rows 0..nrow-1 of one series file at STIME1 = i / (24 * 3600 / MDT) days,
// through the float temporary of mesh.cpp:405 and :441; a discharge is
// divided by the number of entries of its group in float (mesh.cpp:445).
template <class R> std::vector<R> rows_of(const Table<R> &s, int nrow, int MDT, int divide_by) {
    std::vector<R> v(nrow);
    for (int i = 0; i < nrow; i++) {
        const R t = (R)(0.0 + i / (24.0 * 3600.0 / MDT));
        const float tmp = (float)interp(t, s);
        v[i] = divide_by > 0 ? (R)(tmp / divide_by) : (R)tmp;
    }
    return v;
}

template <class R>
bool take_series(const BoundaryIn &d, const std::string &dir, const char *kind, int n, int nfiles,
                 const int32_t *MB, const int32_t *NN, bool discharge, R *V, R *D, std::string &err) {
    const int C = d.CELL, rows = d.rows, nrow = rows + 1 < d.NDAYS ? rows + 1 : d.NDAYS;
    const int K0 = (int)((R)d.MDT / (R)d.DT);
    std::vector<std::vector<R>> val(n);   // per entry, rows 0..nrow-1 (mesh.cpp:380, 413)
    for (int j = 0; j < n; j++) val[j].assign(nrow, R(0));
    for (int k = 1; k <= nfiles; k++) {
        Table<R> s;
        if (!read_table(series_path(dir, kind, k), -1, s, err)) return false;
        int kl = 0;
        for (int j = 0; j < n; j++) kl += NN[j] == k;
        const std::vector<R> v = rows_of(s, nrow, d.MDT, discharge ? kl : 0);
        for (int j = 0; j < n; j++)
            if (NN[j] == k) val[j] = v;
    }
    // increments over K0 steps, 0 on the last row of the series (mesh.cpp:449-464);
    // column-major expansion to cells, a later entry of a cell winning (cell.cpp:110-126)
    for (int j = 0; j < n; j++) {
        const int pos = MB[j] - 1;
        for (int jt = 0; jt < rows; jt++) {
            V[(size_t)jt * C + pos] = val[j][jt];
            D[(size_t)jt * C + pos] = jt != d.NDAYS - 1 ? (val[j][jt + 1] - val[j][jt]) / K0 : R(0);
        }
    }
    return true;
}

// Rating curves (mesh.cpp:468-520; cell.cpp:103-109): per curve k and point
// i, the wetted width of every entry j on the curve is the depth above its bed
// (at least HM2) times the side of its last KLAS 3 slot, 0 for a cell with no
// such slot; each entry takes the share of the point's discharge its width
// holds of the curve's total.
template <class R>
bool take_rating(const BoundaryIn &d, const std::string &dir, const R *ZBC, const R *SIDE, const int32_t *KLAS,
                 int32_t *NHQ1, R *ZW, R *QW, std::string &err) {
    const int n = d.NZQ, NHQ = d.NHQ;
    const R HM2 = (R)d.HM2;
    std::vector<int32_t> nhq1(n, 0);
    std::vector<R> zw((size_t)NHQ * n, R(0)), qw((size_t)NHQ * n, R(0));   // [point][entry]
    for (int k = 1; k <= d.NNZQ0; k++) {
        Table<R> s;
        if (!read_table(series_path(dir, "NZQ", k), NHQ, s, err)) return false;
        const int npt = (int)s.a.size();
        std::vector<R> aqzh0(n, R(0)), aqzh1(NHQ, R(0));
        for (int i = 0; i < npt; i++) {
            for (int j = 0; j < n; j++) {
                if (d.NNZQ[j] != k) continue;
                const int cell = d.MBZQ[j] - 1;
                nhq1[j] = npt;
                for (int jj = 0; jj < 4; jj++)
                    if (KLAS[4 * cell + jj] == 3)
                        aqzh0[j] = std::fmax(s.a[i] - ZBC[cell], HM2) * SIDE[4 * cell + jj];
                aqzh1[i] += aqzh0[j];
            }
            for (int j = 0; j < n; j++) {
                if (d.NNZQ[j] != k) continue;
                zw[(size_t)i * n + j] = s.a[i];
                qw[(size_t)i * n + j] = s.b[i] * aqzh0[j] / aqzh1[i];
            }
        }
    }
    for (int j = 0; j < n; j++) {
        const int pos = d.MBZQ[j] - 1;
        for (int i = 0; i < NHQ; i++) {
            ZW[(size_t)pos * NHQ + i] = zw[(size_t)i * n + j];
            QW[(size_t)pos * NHQ + i] = qw[(size_t)i * n + j];
        }
        NHQ1[pos] = nhq1[j];
    }
    return true;
}

}  // namespace

template <class R>
bool take_boundary(const BoundaryIn &d, const char *bounde_dir, R *QT, R *DQT, R *ZT, R *DZT, int32_t *NHQ1, R *ZW,
                   R *QW, std::string &err) {
    const size_t nser = (size_t)d.rows * d.CELL;
    std::fill(QT, QT + nser, R(0));
    std::fill(DQT, DQT + nser, R(0));
    std::fill(ZT, ZT + nser, R(0));
    std::fill(DZT, DZT + nser, R(0));
    if (NHQ1) std::fill(NHQ1, NHQ1 + d.CELL, 0);
    if (ZW) std::fill(ZW, ZW + (size_t)d.CELL * d.NHQ, R(0));
    if (QW) std::fill(QW, QW + (size_t)d.CELL * d.NHQ, R(0));
    const std::string dir(bounde_dir);
    if (!take_series<R>(d, dir, "NZ", d.NZ, d.NNZ0, d.MBZ, d.NNZ, false, ZT, DZT, err)) return false;
    if (!take_series<R>(d, dir, "NQ", d.NQ, d.NNQ0, d.MBQ, d.NNQ, true, QT, DQT, err)) return false;
    if (d.NZQ > 0 && !take_rating<R>(d, dir, (const R *)d.ZBC, (const R *)d.SIDE, d.KLAS, NHQ1, ZW, QW, err))
        return false;
    return true;
}

template <class R> void boundary_increments(int CELL, int rows_v, int rows_d, int K0, const R *V, R *D) {
    for (int jt = 0; jt < rows_d; jt++)
        for (int c = 0; c < CELL; c++)
            D[(size_t)jt * CELL + c] =
                jt + 1 < rows_v ? (V[(size_t)(jt + 1) * CELL + c] - V[(size_t)jt * CELL + c]) / K0 : R(0);
}

template bool take_boundary<double>(const BoundaryIn &, const char *, double *, double *, double *, double *,
                                    int32_t *, double *, double *, std::string &);
template bool take_boundary<float>(const BoundaryIn &, const char *, float *, float *, float *, float *, int32_t *,
                                   float *, float *, std::string &);
template void boundary_increments<double>(int, int, int, int, const double *, double *);
template void boundary_increments<float>(int, int, int, int, const float *, float *);

}  // namespace hydro
