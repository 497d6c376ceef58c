// analysis.cc -- host setup of the preconditioners (see analysis.h).
#include "analysis.h"

#include <algorithm>
#include <numeric>

namespace hflmp {

int multicoloring(int n, const int *row, const int *col, int *ncolors, int *color_of,
                  std::vector<int> &color_sizes, int *perm) {
    // First fit: a row takes colour 1 and moves up while an off-diagonal
    // neighbour (coloured earlier) holds its colour.  The reference rescans
    // the row (row length) times, which reaches the same fixed point: the
    // smallest colour not held by a lower-numbered neighbour.
    int hicol = 0;
    std::vector<int> &nc = color_sizes;  // reused as scratch below
    std::fill(color_of, color_of + n, 0);
    std::vector<int> mark;  // mark[c] == j: colour c is taken around row j
    for (int j = 0; j < n; ++j) {
        int c = 1;
        bool has_off = false;
        for (int q = row[j]; q < row[j + 1]; ++q) {
            const int d = col[q];
            if (d == j) continue;
            has_off = true;
            const int cd = color_of[d];
            if (cd > 0) {
                if (static_cast<int>(mark.size()) <= cd) mark.resize(cd + 1, -1);
                mark[cd] = j;
            }
        }
        while (c < static_cast<int>(mark.size()) && mark[c] == j) ++c;
        color_of[j] = c;
        if (has_off) hicol = std::max(hicol, c);
    }
    // the reference only counts colours met while scanning off-diagonals; a
    // pattern without any off-diagonal entry keeps 0 colours and is rejected
    if (hicol == 0) return -1;
    *ncolors = hicol;
    nc.assign(hicol, 0);
    for (int i = 0; i < n; ++i) {
        if (color_of[i] > hicol) return -1;
        ++nc[color_of[i] - 1];
    }
    std::vector<int> start(hicol + 1, 0);
    for (int c = 0; c < hicol; ++c) start[c + 1] = start[c] + nc[c];
    for (int j = 0; j < n; ++j) perm[start[color_of[j] - 1]++] = j;
    return 0;
}

static void bucket_by_level(int n, const std::vector<int> &lev, std::vector<int> &ptr, int *rows) {
    const int nl = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
    ptr.assign(nl + 1, 0);
    for (int i = 0; i < n; ++i) ++ptr[lev[i] + 1];
    for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
}

int levels_lower(int n, const int *row, const int *col, std::vector<int> &ptr, int *rows) {
    std::vector<int> lev(n, 0);
    for (int i = 0; i < n; ++i) {
        int l = 0;
        for (int q = row[i]; q < row[i + 1]; ++q)
            if (col[q] < i) l = std::max(l, lev[col[q]] + 1);
        lev[i] = l;
    }
    bucket_by_level(n, lev, ptr, rows);
    return 0;
}

int levels_upper(int n, const int *row, const int *col, std::vector<int> &ptr, int *rows) {
    std::vector<int> lev(n, 0);
    for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        for (int q = row[i]; q < row[i + 1]; ++q)
            if (col[q] > i) l = std::max(l, lev[col[q]] + 1);
        lev[i] = l;
    }
    bucket_by_level(n, lev, ptr, rows);
    return 0;
}

int ilu0(int n, const int *row, const int *col, double *val) {
    // position of each column of row k, for the a_kj lookups
    std::vector<int> pos(n, -1);
    std::vector<int> dpos(n, -1);
    int missing = 0;
    for (int i = 0; i < n; ++i)
        for (int q = row[i]; q < row[i + 1]; ++q)
            if (col[q] == i) dpos[i] = q;
    for (int i = 1; i < n; ++i) {
        for (int ik = row[i]; ik < row[i + 1] && col[ik] < i; ++ik) {
            const int k = col[ik];
            if (dpos[k] < 0) {
                ++missing;
                continue;
            }
            val[ik] = val[ik] / (1.0 * val[dpos[k]]);
            for (int q = row[k]; q < row[k + 1]; ++q) pos[col[q]] = q;
            for (int ij = ik + 1; ij < row[i + 1]; ++ij) {
                const int p = pos[col[ij]];
                // entries of row i without a partner in row k are unchanged
                if (p >= 0) val[ij] = val[ij] - val[ik] * val[p];
            }
            for (int q = row[k]; q < row[k + 1]; ++q) pos[col[q]] = -1;
        }
    }
    return missing;
}

int csr_permute(int n, const int *row, const int *col, const double *val, const int *perm, int *prow, int *pcol,
                double *pval) {
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) {
        if (perm[i] < 0 || perm[i] >= n) return -1;
        inv[perm[i]] = i;
    }
    prow[0] = 0;
    for (int i = 0; i < n; ++i) prow[i + 1] = prow[i] + (row[perm[i] + 1] - row[perm[i]]);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < n; ++i) {
        const int o = perm[i];
        const int len = row[o + 1] - row[o];
        std::vector<std::pair<int, double>> e(len);
        for (int q = 0; q < len; ++q) e[q] = {inv[col[row[o] + q]], val[row[o] + q]};
        std::sort(e.begin(), e.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (int q = 0; q < len; ++q) {
            pcol[prow[i] + q] = e[q].first;
            pval[prow[i] + q] = e[q].second;
        }
    }
    return 0;
}

}  // namespace hflmp

namespace hflmp {

void levelscheduling(int n, const int *row, const int *col, int *nlevels, int *level_sizes, int *perm) {
    std::vector<int> depth(n, 0), ends;
    std::vector<char> placed(n, 0);
    int done = 0, nlev = 0;
    for (int L = 0; L < n; ++L) {
        if (done == n) {
            nlev = static_cast<int>(ends.size());
            break;
        }
        for (int k = 0; k < n; ++k) {
            if (placed[k] || depth[k] != L) continue;
            bool ready = true;
            for (int j = row[k]; j < row[k + 1] && ready; ++j)
                if (col[j] < k && !placed[col[j]]) ready = false;
            if (!ready) continue;
            perm[done++] = k;
            placed[k] = 1;
            for (int j = row[k]; j < row[k + 1]; ++j)
                if (col[j] > k && depth[col[j]] <= depth[k]) depth[col[j]] = depth[k] + 1;
        }
        ends.push_back(done);
    }
    *nlevels = nlev;
    for (int i = 0, prev = 0; i < nlev; ++i) {
        level_sizes[i] = ends[i] - prev;
        prev = ends[i];
    }
}

}  // namespace hflmp
