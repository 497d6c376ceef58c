// mmio.cc -- MatrixMarket matrices and plain-text vectors on the host
// (CPU_COO_lMatrix::ReadFile / WriteFile, lmatrix_coo_cpu.cc:133-309, with the
// banner rules of lmp/mmio.cc; CPU_lVector::ReadFile / WriteFile,
// lvector_cpu.cc:582-660).
//
// Reading accepts coordinate real or integer matrices, general or symmetric,
// as the reference does; a symmetric file's off-diagonal entries are mirrored
// (each mirror follows its entry, as in the reference).  The COO entries
// then become CSR with ascending columns per row (CPU_CSR_lMatrix::
// TransformFromCOO, lmatrix_csr_cpu.cc:951-1044).  Writing emits the
// standard "row col value" triples; the reference's WriteFile prints the
// column index first (lmatrix_coo_cpu.cc:302-303), i.e. the transpose.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/hflmp.h"

namespace {

struct Coo {
    int nrow = 0, ncol = 0;
    std::vector<int> r, c;
    std::vector<double> v;
};

std::string lower(std::string s) {
    for (auto &ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    return s;
}

int read_coo(const char *path, Coo *out) {
    FILE *f = std::fopen(path, "r");
    if (!f) {
        hflmp_set_error("hfl_mm_read: cannot open the file");
        return -1;
    }
    char line[1024];
    if (!std::fgets(line, sizeof line, f)) {
        std::fclose(f);
        hflmp_set_error("hfl_mm_read: empty file");
        return -2;
    }
    char tag[64] = {0}, obj[64] = {0}, fmt[64] = {0}, field[64] = {0}, sym[64] = {0};
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", tag, obj, fmt, field, sym) != 5 ||
        std::string(tag) != "%%MatrixMarket" || lower(obj) != "matrix") {
        std::fclose(f);
        hflmp_set_error("hfl_mm_read: not a MatrixMarket matrix banner");
        return -2;
    }
    const std::string F = lower(fmt), T = lower(field), S = lower(sym);
    if (F != "coordinate" || (T != "real" && T != "integer") || (S != "general" && S != "symmetric")) {
        std::fclose(f);
        hflmp_set_error("hfl_mm_read: only coordinate real/integer matrices, general or symmetric, are supported");
        return -3;
    }
    // skip comments
    long long nnz = -1;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '%') continue;
        if (std::sscanf(line, "%d %d %lld", &out->nrow, &out->ncol, &nnz) == 3) break;
    }
    if (nnz < 0 || out->nrow < 0 || out->ncol < 0) {
        std::fclose(f);
        hflmp_set_error("hfl_mm_read: missing size line");
        return -2;
    }
    out->r.reserve(static_cast<size_t>(nnz) * (S == "symmetric" ? 2 : 1));
    for (long long e = 0; e < nnz; ++e) {
        int i, j;
        double v;
        if (std::fscanf(f, " %d %d %lf", &i, &j, &v) != 3) {
            std::fclose(f);
            hflmp_set_error("hfl_mm_read: fewer entries than the size line states");
            return -2;
        }
        if (i < 1 || i > out->nrow || j < 1 || j > out->ncol) {
            std::fclose(f);
            hflmp_set_error("hfl_mm_read: entry index out of range");
            return -2;
        }
        out->r.push_back(i - 1);
        out->c.push_back(j - 1);
        out->v.push_back(v);
        if (S == "symmetric" && i != j) {
            out->r.push_back(j - 1);
            out->c.push_back(i - 1);
            out->v.push_back(v);
        }
    }
    std::fclose(f);
    return 0;
}

}  // namespace

extern "C" int hfl_mm_read(const char *path, int *nrow, int *ncol, long long *nnz, int *row, int *col,
                           double *val) {
    if (!path || !nrow || !ncol || !nnz) {
        hflmp_set_error("hfl_mm_read: null argument");
        return -1;
    }
    Coo m;
    int rc = read_coo(path, &m);
    if (rc) return rc;
    *nrow = m.nrow;
    *ncol = m.ncol;
    *nnz = static_cast<long long>(m.v.size());
    if (!row) return 0;
    if (!col || !val) {
        hflmp_set_error("hfl_mm_read: null output array");
        return -1;
    }
    // counting sort by row, entries of a row in file order, then ascending columns
    std::vector<int> cnt(m.nrow + 1, 0);
    for (int r : m.r) ++cnt[r + 1];
    for (int i = 0; i < m.nrow; ++i) cnt[i + 1] += cnt[i];
    for (int i = 0; i <= m.nrow; ++i) row[i] = cnt[i];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    std::vector<std::pair<int, double>> e(m.v.size());
    for (size_t k = 0; k < m.v.size(); ++k) e[fill[m.r[k]]++] = {m.c[k], m.v[k]};
    for (int i = 0; i < m.nrow; ++i)
        std::stable_sort(e.begin() + row[i], e.begin() + row[i + 1],
                         [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
                             return a.first < b.first;
                         });
    for (size_t k = 0; k < e.size(); ++k) {
        col[k] = e[k].first;
        val[k] = e[k].second;
    }
    return 0;
}

extern "C" int hfl_mm_write(const char *path, int nrow, int ncol, const int *row, const int *col, const double *val) {
    if (!path || nrow < 0 || ncol < 0 || !row || (row[nrow] > 0 && (!col || !val))) {
        hflmp_set_error("hfl_mm_write: bad argument");
        return -1;
    }
    FILE *f = std::fopen(path, "w");
    if (!f) {
        hflmp_set_error("hfl_mm_write: cannot open the file");
        return -1;
    }
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%d %d %d\n", nrow, ncol, row[nrow]);
    for (int i = 0; i < nrow; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j) std::fprintf(f, "%d %d %24.14e\n", i + 1, col[j] + 1, val[j]);
    const int bad = std::ferror(f);
    std::fclose(f);
    if (bad) {
        hflmp_set_error("hfl_mm_write: write failed");
        return -2;
    }
    return 0;
}

// CPU_lVector::ReadFile: whitespace-separated values, as many as can be read.
extern "C" int hfl_vec_read(const char *path, int *n, double *values) {
    if (!path || !n) {
        hflmp_set_error("hfl_vec_read: null argument");
        return -1;
    }
    std::ifstream in(path);
    if (!in.is_open()) {
        hflmp_set_error("hfl_vec_read: cannot open the file");
        return -1;
    }
    int k = 0;
    double v;
    while (in >> v) {
        if (values) {
            if (k >= *n) break;
            values[k] = v;
        }
        ++k;
    }
    *n = k;
    return 0;
}

// CPU_lVector::WriteFile: one value per line, std::scientific at the stream's
// default precision (6 digits after the point), as the reference writes.
extern "C" int hfl_vec_write(const char *path, int n, const double *values) {
    if (!path || n < 0 || (n > 0 && !values)) {
        hflmp_set_error("hfl_vec_write: bad argument");
        return -1;
    }
    std::ofstream out(path);
    if (!out.is_open()) {
        hflmp_set_error("hfl_vec_write: cannot open the file");
        return -1;
    }
    for (int i = 0; i < n; ++i) out << std::scientific << values[i] << '\n';
    return out.good() ? 0 : -2;
}
