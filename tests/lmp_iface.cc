// lmp_iface.cc -- the lmp interface of the product (lmp.h) exercised the way
// the reference's own C++ tests exercise theirs
// (test/LinearAlgebra/CPUsimple_CSR_lMatrix.cc), through base-class
// pointers from the platform factories, and checked against the CPU oracle
// (oracle/libhfo.so) or against known answers.
//
//   lmp_iface cpu   host containers and the interface itself (no GPU needed)
//   lmp_iface hip   every lVector / lMatrix virtual on the HIP platform
//
// Prints one line per failed check and exits with the number of failures.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "coupled.h"
#include "hfo.h"
#include "lmp.h"
#include "solvers.h"

using namespace hflmp;

static int g_fail = 0, g_checks = 0;
static std::string g_test;

#define EXPECT(cond)                                                                            \
    do {                                                                                        \
        ++g_checks;                                                                             \
        if (!(cond)) {                                                                          \
            ++g_fail;                                                                           \
            std::printf("FAIL %s:%d [%s] %s\n", __FILE__, __LINE__, g_test.c_str(), #cond);    \
        }                                                                                       \
    } while (0)

template <typename F>
static void run(const char *name, F f) {
    g_test = name;
    try {
        f();
    } catch (const std::exception &e) {
        ++g_fail;
        std::printf("FAIL [%s] exception: %s\n", name, e.what());
    }
}

template <typename F>
static bool throws(F f) {
    try {
        f();
    } catch (const Error &) {
        return true;
    }
    return false;
}

// the HIP classes implement every pure virtual of the interface
static_assert(!std::is_abstract<HIP_lVector<double>>::value, "HIP_lVector<double> is abstract");
static_assert(!std::is_abstract<HIP_lVector<float>>::value, "HIP_lVector<float> is abstract");
static_assert(!std::is_abstract<HIP_CSR_lMatrix<double>>::value, "HIP_CSR_lMatrix<double> is abstract");
static_assert(!std::is_abstract<HIP_CSR_lMatrix<float>>::value, "HIP_CSR_lMatrix<float> is abstract");
static_assert(!std::is_abstract<CPU_lVector<double>>::value, "CPU_lVector<double> is abstract");
static_assert(!std::is_abstract<CPU_CSR_lMatrix<double>>::value, "CPU_CSR_lMatrix<double> is abstract");

template <typename T>
static bool same(const std::vector<T> &a, const std::vector<T> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (!(a[i] == b[i]) && !(std::isnan(a[i]) && std::isnan(b[i]))) return false;
    return true;
}

template <typename T>
static std::vector<T> host_of(const lVector<T> &v) {
    std::vector<T> h(v.get_size());
    v.GetBlockValues(0, v.get_size(), h.data());
    return h;
}

// FV 7-point matrix on an nx*ny*nz grid (k = 1 + smooth variation), CSR
static void fv_grid(int nx, int ny, int nz, CPU_CSR_lMatrix<double> &A) {
    const int n = nx * ny * nz;
    std::vector<int> r, c;
    std::vector<double> v;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int row = i + nx * (j + ny * k);
                const int nb[6] = {row - nx * ny, row - nx, row - 1, row + 1, row + nx, row + nx * ny};
                const bool has[6] = {k > 0, j > 0, i > 0, i < nx - 1, j < ny - 1, k < nz - 1};
                double d = 0;
                for (int q = 0; q < 6; ++q) {
                    const double w = 1.0 + 0.25 * std::sin(0.3 * row + q);
                    d += w;
                    if (!has[q]) continue;
                    r.push_back(row);
                    c.push_back(nb[q]);
                    v.push_back(-w);
                }
                r.push_back(row);
                c.push_back(row);
                v.push_back(d + 0.5);
            }
    A.Init(static_cast<int>(r.size()), n, n, "A");
    A.TransformFromCOO(r.data(), c.data(), v.data(), n, n, static_cast<int>(r.size()));
}

// ---------------------------------------------------------------- CPU (no GPU)

static void cpu_tests() {
    // test/LinearAlgebra/CPUsimple_CSR_lMatrix.cc: transpose_2x2, transpose_1x3
    run("transpose_2x2", [] {
        CPU_CSR_lMatrix<double> A;
        A.Init(4, 2, 2, "A");
        int row[4] = {0, 0, 1, 1}, col[4] = {0, 1, 0, 1};
        A.init_structure(row, col);
        A.Zeros();
        A.add_value(0, 0, 1.0);
        A.add_value(0, 1, 2.0);
        A.add_value(1, 0, 3.0);
        A.add_value(1, 1, 4.0);
        A.transpose_me();
        double v = 0;
        A.get_value(0, 0, &v);
        EXPECT(v == 1.0);
        A.get_value(0, 1, &v);
        EXPECT(v == 3.0);
        A.get_value(1, 0, &v);
        EXPECT(v == 2.0);
        A.get_value(1, 1, &v);
        EXPECT(v == 4.0);
    });
    run("transpose_1x3", [] {
        CPU_CSR_lMatrix<double> A;
        A.Init(3, 1, 3, "A");
        int row[3] = {0, 0, 0}, col[3] = {0, 1, 2};
        A.init_structure(row, col);
        A.Zeros();
        A.add_value(0, 0, 1.0);
        A.add_value(0, 1, 2.0);
        A.add_value(0, 2, 3.0);
        A.transpose_me();
        EXPECT(A.get_num_row() == 3 && A.get_num_col() == 1);
        double v = 0;
        A.get_value(0, 0, &v);
        EXPECT(v == 1.0);
        A.get_value(1, 0, &v);
        EXPECT(v == 2.0);
        A.get_value(2, 0, &v);
        EXPECT(v == 3.0);
    });
    // CPUsimple_CSR_lMatrix.cc:79-101 (add_values)
    run("add_values", [] {
        int rows[] = {0, 0, 1, 1, 1, 2, 2}, cols[] = {0, 1, 0, 1, 2, 1, 2};
        double values[] = {1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0};
        CPU_CSR_lMatrix<double> A;
        A.Init(7, 3, 3, "A");
        A.init_structure(rows, cols);
        A.Zeros();
        int r2[] = {0, 1, 2}, c2[] = {0, 1, 2};
        A.add_values(r2, 3, c2, 3, values);
        const double ref[] = {1.0, 2.0, 4.0, 5.0, 6.0, 8.0, 9.0};
        for (int j = 0; j < 7; ++j) EXPECT(A.val[j] == ref[j]);
    });
    // CPUsimple_CSR_lMatrix.cc:159-200 (UnorderedInitStructure_square / _rect)
    run("UnorderedInitStructure", [] {
        int ar[] = {1, 2, 0, 0, 1}, ac[] = {2, 1, 0, 1, 1};
        CPU_CSR_lMatrix<double> A;
        A.Init(5, 3, 3, "A");
        A.init_structure(ar, ac);
        EXPECT((A.row == std::vector<int>{0, 2, 4, 5}));
        EXPECT((A.col == std::vector<int>{0, 1, 1, 2, 1}));
        int br[] = {2, 0, 0, 1}, bc[] = {1, 0, 1, 1};
        CPU_CSR_lMatrix<double> B;
        B.Init(4, 3, 2, "B");
        B.init_structure(br, bc);
        EXPECT((B.row == std::vector<int>{0, 2, 3, 4}));
        EXPECT((B.col == std::vector<int>{0, 1, 1, 1}));
    });
    run("TransformFromCOO_vs_oracle", [] {
        std::mt19937 g(3);
        const int n = 40, nnz = 300;
        std::vector<int> r(nnz), c(nnz);
        std::vector<double> v(nnz);
        for (int k = 0; k < nnz; ++k) {
            r[k] = g() % n;
            c[k] = g() % n;
            v[k] = k;
        }
        CPU_CSR_lMatrix<double> A;
        A.Init(nnz, n, n, "A");
        A.TransformFromCOO(r.data(), c.data(), v.data(), n, n, nnz);
        std::vector<int> orow(n + 1), ocol(nnz);
        std::vector<double> oval(nnz);
        hfo_csr_from_coo_d(n, nnz, r.data(), c.data(), v.data(), orow.data(), ocol.data(), oval.data());
        EXPECT(A.row == orow);
        EXPECT(A.col == ocol);
        EXPECT(same(A.val, oval));   // repeated columns in the reference's order too
    });
    run("ZeroRows_ZeroCols_get_value", [] {
        CPU_CSR_lMatrix<double> A;
        fv_grid(4, 3, 2, A);
        std::vector<double> ref = A.val;
        int rows[] = {0, 5, 5, 23};
        hfo_zero_rows_d(A.row.data(), A.col.data(), ref.data(), rows, 4, 3.5);
        CPU_CSR_lMatrix<double> B = A;
        B.ZeroRows(rows, 4, 3.5);
        EXPECT(same(B.val, ref));
        ref = A.val;
        int cols[] = {2, 7};
        hfo_zero_cols_d(A.get_num_row(), A.row.data(), A.col.data(), ref.data(), cols, 2, -1.0);
        B = A;
        B.ZeroCols(cols, 2, -1.0);
        EXPECT(same(B.val, ref));
        double v = 42.0;
        A.get_value(0, 23, &v);   // not in the pattern: left as it was
        EXPECT(v == 42.0);
        A.get_value(0, 0, &v);
        EXPECT(v == A.val[A.row[0]]);
    });
    run("Levelscheduling_vs_oracle", [] {
        for (int shape = 0; shape < 3; ++shape) {
            CPU_CSR_lMatrix<double> A;
            if (shape == 0) fv_grid(5, 4, 3, A);
            else if (shape == 1) fv_grid(17, 1, 1, A);   // a chain: the reference reports 0 levels
            else {
                // non-symmetric pattern: lower entries only
                fv_grid(6, 5, 1, A);
                std::vector<int> r, c;
                std::vector<double> v;
                for (int i = 0; i < A.get_num_row(); ++i)
                    for (int j = A.row[i]; j < A.row[i + 1]; ++j)
                        if (A.col[j] <= i) {
                            r.push_back(i);
                            c.push_back(A.col[j]);
                            v.push_back(A.val[j]);
                        }
                const int n = A.get_num_row();
                A.Init(static_cast<int>(r.size()), n, n, "L");
                A.TransformFromCOO(r.data(), c.data(), v.data(), n, n, static_cast<int>(r.size()));
            }
            const int n = A.get_num_row();
            int nlev = -1, *sizes = nullptr, *perm = nullptr;
            A.Levelscheduling(nlev, &sizes, &perm);
            int onl = -1;
            std::vector<int> osz(n), operm(n);
            hfo_levelscheduling(n, A.row.data(), A.col.data(), &onl, osz.data(), operm.data());
            EXPECT(nlev == onl);
            EXPECT(std::vector<int>(perm, perm + n) == operm);
            for (int l = 0; l < nlev; ++l) EXPECT(sizes[l] == osz[l]);
            if (shape == 1) EXPECT(nlev == 0);
            delete[] sizes;
            delete[] perm;
        }
    });
    run("Multicoloring_vs_oracle", [] {
        CPU_CSR_lMatrix<double> A;
        fv_grid(6, 5, 4, A);
        const int n = A.get_num_row();
        int nc = 0, *sizes = nullptr, *perm = nullptr;
        A.Multicoloring(nc, &sizes, &perm);
        int onc = 0;
        std::vector<int> color(n), osz(n), operm(n);
        hfo_multicoloring(n, A.row.data(), A.col.data(), &onc, color.data(), osz.data(), operm.data());
        EXPECT(nc == onc && nc == 2);
        EXPECT(std::vector<int>(perm, perm + n) == operm);
        for (int c = 0; c < nc; ++c) EXPECT(sizes[c] == osz[c]);
        delete[] sizes;
        delete[] perm;
    });
    run("extract_submatrix_and_diag", [] {
        CPU_CSR_lMatrix<double> A;
        fv_grid(5, 4, 2, A);
        std::unique_ptr<lMatrix<double>> S(A.extract_submatrix(3, 2, 17, 30));
        CPU_CSR_lMatrix<double> &B = dynamic_cast<CPU_CSR_lMatrix<double> &>(*S);
        EXPECT(B.get_num_row() == 14 && B.get_num_col() == 28);
        int cnt = 0;
        for (int i = 3; i < 17; ++i)
            for (int j = A.row[i]; j < A.row[i + 1]; ++j)
                if (A.col[j] >= 2 && A.col[j] < 30) {
                    double v = -1;
                    B.get_value(i - 3, A.col[j] - 2, &v);
                    EXPECT(v == A.val[j]);
                    ++cnt;
                }
        EXPECT(cnt == B.get_nnz());
        CPU_lVector<double> d(10, "d");
        A.extract_invdiagelements(5, 15, &d);
        std::vector<double> inv(A.get_num_row());
        hfo_invdiag_d(A.get_num_row(), A.row.data(), A.col.data(), A.val.data(), inv.data());
        for (int i = 0; i < 10; ++i) EXPECT(d.buffer[i] == inv[5 + i]);
    });
    run("Rotg_Rotmg", [] {
        CPU_lVector<double> x(1, "x");
        double a = 3, b = 4, c = 0, s = 0;
        x.Rotg(&a, &b, &c, &s);
        double oa = 3, ob = 4, oc = 0, os = 0;
        hfo_rotg_d(&oa, &ob, &oc, &os);
        EXPECT(a == oa && b == ob && c == oc && s == os);
        // Rotmg: H (x1, y1) has a zero second component, and the weighted
        // squares are kept: d1' x1'^2 = d1 x1^2 + d2 y1^2
        const double cases[][4] = {{2.0, 3.0, 1.5, -0.7}, {0.5, 4.0, 0.2, 1.9}, {1e-9, 1e9, 3.0, 2.0},
                                   {3.0, 1e-12, -1.0, 1e-3}, {1.0, 1.0, 0.0, 2.0}};
        for (const auto &cs : cases) {
            double d1 = cs[0], d2 = cs[1], x1 = cs[2], p[5] = {0, 0, 0, 0, 0};
            const double y1 = cs[3];
            x.Rotmg(&d1, &d2, &x1, y1, p);
            double h11, h12, h21, h22;
            if (p[0] == -1) h11 = p[1], h21 = p[2], h12 = p[3], h22 = p[4];
            else if (p[0] == 0) h11 = 1, h21 = p[2], h12 = p[3], h22 = 1;
            else if (p[0] == 1) h11 = p[1], h21 = -1, h12 = 1, h22 = p[4];
            else h11 = 1, h21 = 0, h12 = 0, h22 = 1;
            const double nx = h11 * cs[2] + h12 * y1, ny = h21 * cs[2] + h22 * y1;
            EXPECT(std::fabs(ny) <= 1e-12 * (std::fabs(cs[2]) + std::fabs(y1)));
            EXPECT(std::fabs(nx - x1) <= 1e-12 * std::fabs(x1));
            const double w0 = cs[0] * cs[2] * cs[2] + cs[1] * y1 * y1;
            EXPECT(std::fabs(d1 * x1 * x1 - w0) <= 1e-12 * w0);
        }
        double d1 = 1, d2 = 0, x1 = 2, p[5] = {0, 0, 0, 0, 0};   // d2 y1 = 0: H = I
        x.Rotmg(&d1, &d2, &x1, 3.0, p);
        EXPECT(p[0] == -2);
    });
    run("CPU_lVector_storage", [] {
        CPU_lVector<double> v(10, "v");
        int idx[] = {1, 4, 1};
        double val[] = {1e16, 2.0, -1e16};
        v.add_values(idx, 3, val);
        EXPECT(v.buffer[1] == 0.0 && v.buffer[4] == 2.0);
        int is[] = {4, 1, 9};
        v.set_indexset(is, 3);
        EXPECT(v.get_indexset_size() == 3);
        double got[3];
        v.GetIndexedValues(got);
        EXPECT(got[0] == 2.0 && got[1] == 0.0 && got[2] == 0.0);
        double put[3] = {5.0, 6.0, 7.0};
        v.SetIndexedValues(put);
        EXPECT(v.buffer[4] == 5.0 && v.buffer[1] == 6.0 && v.buffer[9] == 7.0);
        CPU_lVector<double> w(10, "w");
        w.CopyFromIndexset(v);
        EXPECT(w.get_indexset_size() == 3);
        std::unique_ptr<lVector<double>> sub(v.extract_subvector(4, 10));
        EXPECT(sub->get_size() == 6 && dynamic_cast<CPU_lVector<double> &>(*sub).buffer[0] == 5.0);
        EXPECT(throws([&] { v.Dot(w); }));
    });
}

// ---------------------------------------------------------------- HIP

template <typename T>
static std::vector<T> rnd(int n, unsigned seed, double scale = 1.0) {
    std::mt19937 g(seed);
    std::normal_distribution<double> d(0.0, scale);
    std::vector<T> v(n);
    for (auto &x : v) x = static_cast<T>(d(g));
    return v;
}

template <typename T>
static std::unique_ptr<lVector<T>> hip_vec(const std::vector<T> &h, const std::string &name = "v") {
    std::unique_ptr<lVector<T>> v(init_vector<T>(static_cast<int>(h.size()), name, HIP));
    v->SetBlockValues(0, static_cast<int>(h.size()), h.data());
    return v;
}

template <typename T> struct O;
template <> struct O<double> {
    static double dot(int n, const double *x, const double *y) { return hfo_dot_d(n, x, y); }
    static double norm1(int n, const double *x) { return hfo_norm1_d(n, x); }
    static int argmax(int n, const double *x) { return hfo_argmax_d(n, x); }
    static int argmin(int n, const double *x) { return hfo_argmin_d(n, x); }
    static void axpy(int n, double *y, const double *x, double a) { hfo_axpy_d(n, y, x, a); }
    static void scaleadd(int n, double *y, double a, const double *x) { hfo_scaleadd_d(n, y, a, x); }
    static void scale(int n, double *y, double a) { hfo_scale_d(n, y, a); }
    static void ewmult(int n, double *y, const double *x) { hfo_ewmult_d(n, y, x); }
    static void rot(int n, double *a, double *b, double c, double s) { hfo_rot_d(n, a, b, c, s); }
    static void rotm(int n, double *u, double *v, const double *p) { hfo_rotm_d(n, u, v, p); }
    static void add_values(double *y, const int *i, int m, const double *v) { hfo_add_values_d(y, i, m, v); }
    static void add_range(double *y, int s, const double *x, int ss, int m, double w) {
        hfo_add_range_d(y, s, x, ss, m, w);
    }
    static void spmv(int n, const int *r, const int *c, const double *v, const double *x, double *y) {
        hfo_spmv_d(n, r, c, v, x, y);
    }
    static void spmv_add(int n, const int *r, const int *c, const double *v, const double *x, double *y) {
        hfo_spmv_add_d(n, r, c, v, x, y);
    }
    static void nodiag(int n, const int *r, const int *c, const double *v, const double *x, double *y) {
        hfo_spmv_nodiag_d(n, r, c, v, x, y);
    }
    static int ilu0(int n, const int *r, const int *c, double *v) { return hfo_ilu0_d(n, r, c, v); }
    static void ilu_solve(int n, const int *r, const int *c, const double *lu, const double *in, double *out) {
        hfo_ilu_solve_d(n, r, c, lu, in, out);
    }
    static void sgs(int n, const int *r, const int *c, const double *v, const double *in, double *out) {
        hfo_sgs_d(n, r, c, v, in, out);
    }
    static void pjacobi(int n, const int *r, const int *c, const double *v, const double *in, double *out) {
        hfo_pjacobi_d(n, r, c, v, in, out);
    }
};
template <> struct O<float> {
    static float dot(int n, const float *x, const float *y) { return hfo_dot_f(n, x, y); }
    static float norm1(int n, const float *x) { return hfo_norm1_f(n, x); }
    static int argmax(int n, const float *x) { return hfo_argmax_f(n, x); }
    static int argmin(int n, const float *x) { return hfo_argmin_f(n, x); }
    static void axpy(int n, float *y, const float *x, float a) { hfo_axpy_f(n, y, x, a); }
    static void scaleadd(int n, float *y, float a, const float *x) { hfo_scaleadd_f(n, y, a, x); }
    static void scale(int n, float *y, float a) { hfo_scale_f(n, y, a); }
    static void ewmult(int n, float *y, const float *x) { hfo_ewmult_f(n, y, x); }
    static void rot(int n, float *a, float *b, float c, float s) { hfo_rot_f(n, a, b, c, s); }
    static void rotm(int n, float *u, float *v, const float *p) { hfo_rotm_f(n, u, v, p); }
    static void add_values(float *y, const int *i, int m, const float *v) { hfo_add_values_f(y, i, m, v); }
    static void add_range(float *y, int s, const float *x, int ss, int m, float w) {
        hfo_add_range_f(y, s, x, ss, m, w);
    }
    static void spmv(int n, const int *r, const int *c, const float *v, const float *x, float *y) {
        hfo_spmv_f(n, r, c, v, x, y);
    }
    static void spmv_add(int n, const int *r, const int *c, const float *v, const float *x, float *y) {
        hfo_spmv_add_f(n, r, c, v, x, y);
    }
    static void nodiag(int n, const int *r, const int *c, const float *v, const float *x, float *y) {
        hfo_spmv_nodiag_f(n, r, c, v, x, y);
    }
    static int ilu0(int n, const int *r, const int *c, float *v) { return hfo_ilu0_f(n, r, c, v); }
    static void ilu_solve(int n, const int *r, const int *c, const float *lu, const float *in, float *out) {
        hfo_ilu_solve_f(n, r, c, lu, in, out);
    }
    static void sgs(int n, const int *r, const int *c, const float *v, const float *in, float *out) {
        hfo_sgs_f(n, r, c, v, in, out);
    }
    static void pjacobi(int n, const int *r, const int *c, const float *v, const float *in, float *out) {
        hfo_pjacobi_f(n, r, c, v, in, out);
    }
};

template <typename T>
static std::unique_ptr<lVector<T>> Y_of(int n) {
    return std::unique_ptr<lVector<T>>(init_vector<T>(n, "y", HIP));
}

template <typename T>
static bool close_rel(double a, double b, double tol) {
    return std::fabs(a - b) <= tol * std::max(std::fabs(b), 1e-300);
}

template <typename T>
static void hip_vector_tests(const char *tname) {
    const double tol = sizeof(T) == 8 ? 1e-14 : 2e-6;
    const int n = 5003;
    const std::vector<T> a = rnd<T>(n, 1), b = rnd<T>(n, 2);
    std::string pre = std::string(tname) + " ";
    run((pre + "blas1").c_str(), [&] {
        auto X = hip_vec(a), Y = hip_vec(b);
        EXPECT(X->get_platform() == HIP);
        EXPECT(close_rel<T>(X->Dot(*Y), O<T>::dot(n, a.data(), b.data()), tol));
        EXPECT(close_rel<T>(X->Norm1(), O<T>::norm1(n, a.data()), tol));
        EXPECT(close_rel<T>(X->Norm2(), std::sqrt(O<T>::dot(n, a.data(), a.data())), tol));
        EXPECT(X->ArgMax() == O<T>::argmax(n, a.data()));
        EXPECT(X->ArgMin() == O<T>::argmin(n, a.data()));
        EXPECT(X->NormMax() == std::fabs(a[O<T>::argmax(n, a.data())]));
        std::vector<T> ref = b;
        Y->Axpy(*X, T(0.3));
        O<T>::axpy(n, ref.data(), a.data(), T(0.3));
        EXPECT(same(host_of(*Y), ref));
        Y->ScaleAdd(T(-1.7), *X);
        O<T>::scaleadd(n, ref.data(), T(-1.7), a.data());
        EXPECT(same(host_of(*Y), ref));
        Y->Scale(T(0.25));
        O<T>::scale(n, ref.data(), T(0.25));
        EXPECT(same(host_of(*Y), ref));
        const lVector<T> &cy = *Y;
        cy.ElementWiseMult(*X);   // const, as lvector.h:248 declares it
        O<T>::ewmult(n, ref.data(), a.data());
        EXPECT(same(host_of(*Y), ref));
    });
    run((pre + "rotations").c_str(), [&] {
        auto X = hip_vec(a), Y = hip_vec(b);
        std::vector<T> ra = a, rb = b;
        const T c = T(0.6), s = T(0.8);
        X->Rot(Y.get(), c, s);
        O<T>::rot(n, ra.data(), rb.data(), c, s);
        EXPECT(same(host_of(*X), ra) && same(host_of(*Y), rb));
        const T params[4][5] = {{T(-1), T(0.5), T(-2), T(1.5), T(0.25)},
                                {T(0), T(9), T(-0.5), T(2), T(9)},
                                {T(1), T(0.75), T(9), T(9), T(-1.25)},
                                {T(-2), T(3), T(3), T(3), T(3)}};
        for (const auto &p : params) {
            X->Rotm(Y.get(), p[0]);
            O<T>::rotm(n, ra.data(), rb.data(), p);
            EXPECT(same(host_of(*X), ra) && same(host_of(*Y), rb));
        }
        T sa = 3, sb = -4, sc = 0, ss = 0;
        X->Rotg(&sa, &sb, &sc, &ss);
        EXPECT(std::fabs(sc - T(-0.6)) < T(1e-6) && std::fabs(ss - T(0.8)) < T(1e-6));
    });
    run((pre + "values_and_indexset").c_str(), [&] {
        auto X = hip_vec(a);
        std::vector<T> ref = a;
        const int idx[] = {7, 0, 7, n - 1, 7};
        const T val[] = {T(1e6), T(2), T(-1e6), T(3), T(0.5)};
        X->add_values(idx, 5, val);
        O<T>::add_values(ref.data(), idx, 5, val);
        X->add_value(9, T(4));
        ref[9] += T(4);
        EXPECT(same(host_of(*X), ref));
        X->SetValues(idx, 5, val);   // a repeated index keeps its last value
        for (int k = 0; k < 5; ++k) ref[idx[k]] = val[k];
        EXPECT(same(host_of(*X), ref));
        T got[5];
        X->GetValues(idx, 5, got);
        for (int k = 0; k < 5; ++k) EXPECT(got[k] == ref[idx[k]]);
        const int iset[] = {3, 100, 3, n - 2};
        X->set_indexset(iset, 4);
        EXPECT(X->get_indexset_size() == 4);
        int back[4] = {0, 0, 0, 0};
        X->get_indexset(back);
        for (int k = 0; k < 4; ++k) EXPECT(back[k] == iset[k]);
        X->GetIndexedValues(got);
        for (int k = 0; k < 4; ++k) EXPECT(got[k] == ref[iset[k]]);
        const T put[] = {T(11), T(12), T(13), T(14)};
        X->SetIndexedValues(put);
        for (int k = 0; k < 4; ++k) ref[iset[k]] = put[k];
        EXPECT(same(host_of(*X), ref));
        auto Y = hip_vec(b);
        Y->CopyFromIndexset(*X);
        EXPECT(Y->get_indexset_size() == 4);
        std::unique_ptr<lVector<T>> sub(X->extract_subvector(10, 4000));
        EXPECT(sub->get_size() == 3990 && sub->get_platform() == HIP);
        EXPECT(same(host_of(*sub), std::vector<T>(ref.begin() + 10, ref.begin() + 4000)));
        X->partial_add_subvector(100, 5, 2000, T(0.37), *Y);
        O<T>::add_range(ref.data(), 100, b.data(), 5, 2000, T(0.37));
        EXPECT(same(host_of(*X), ref));
        X->partial_replace_subvector(0, 50, 70, *Y);
        std::copy(b.begin() + 50, b.begin() + 120, ref.begin());
        EXPECT(same(host_of(*X), ref));
    });
    run((pre + "copies_and_casts").c_str(), [&] {
        auto X = hip_vec(a);
        CPU_lVector<T> h(n, "h");
        X->CopyTo(h);
        EXPECT(same(h.buffer, a));
        std::unique_ptr<lVector<T>> Z(X->CloneWithoutContent());
        EXPECT(Z->get_size() == n && Z->get_platform() == HIP && Z->Norm1() == T(0));
        *Z = *X;
        EXPECT(same(host_of(*Z), a));
        Z->Zeros();
        Z->CopyFrom(h);
        EXPECT(same(host_of(*Z), a));
        std::unique_ptr<lVector<T>> W(init_vector<T>(1, "w", HIP));
        W->CloneFrom(*X);
        EXPECT(W->get_size() == n && same(host_of(*W), a));
        W->CopyStructureFrom(*Y_of<T>(n));
        EXPECT(W->get_size() == n);
        std::unique_ptr<lVector<double>> D(init_vector<double>(n, "d", HIP));
        std::unique_ptr<lVector<float>> F(init_vector<float>(n, "f", HIP));
        X->CastTo(*D);
        X->CastTo(*F);
        std::vector<double> dref(a.begin(), a.end());
        std::vector<float> fref(n);
        for (int i = 0; i < n; ++i) fref[i] = static_cast<float>(a[i]);
        EXPECT(same(host_of(*D), dref) && same(host_of(*F), fref));
        CPU_lVector<double> hd(n, "hd");
        X->CastTo(hd);
        EXPECT(same(hd.buffer, dref));
        Z->CastFrom(*F);
        std::vector<T> zref(n);
        for (int i = 0; i < n; ++i) zref[i] = static_cast<T>(fref[i]);
        EXPECT(same(host_of(*Z), zref));
        Z->CastFrom(hd);
        EXPECT(same(host_of(*Z), a));
        std::vector<int> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = (i * 7 + 3) % n;
        Z->Reorder(perm.data());
        std::vector<T> pr(n);
        for (int i = 0; i < n; ++i) pr[i] = a[perm[i]];
        EXPECT(same(host_of(*Z), pr));
        Z->Sync();
        const std::string path = std::string("/tmp/lmp_iface_vec_") + tname + ".txt";
        X->WriteFile(path.c_str());
        std::unique_ptr<lVector<T>> R(init_vector<T>(0, "r", HIP));
        R->ReadFile(path.c_str());
        EXPECT(R->get_size() == n);
        std::remove(path.c_str());
    });
}

template <typename T>
static void hip_matrix_tests(const char *tname) {
    std::string pre = std::string(tname) + " ";
    CPU_CSR_lMatrix<double> A64;
    fv_grid(9, 8, 7, A64);
    CPU_CSR_lMatrix<T> A;
    A.CastFrom(A64);
    const int n = A.get_num_row(), nnz = A.get_nnz();
    const std::vector<T> x = rnd<T>(n, 5);
    run((pre + "structure_copies").c_str(), [&] {
        std::unique_ptr<lMatrix<T>> M(init_matrix<T>(nnz, n, n, "M", HIP));
        EXPECT(M->get_platform() == HIP && M->get_matrix_format() == CSR);
        M->CopyStructureFrom(A);
        M->CopyFrom(A);
        CPU_CSR_lMatrix<T> back;
        M->CopyStructureTo(back);
        EXPECT(back.row == A.row && back.col == A.col);
        M->CopyTo(back);
        EXPECT(same(back.val, A.val));
        std::vector<T> v;
        std::vector<int> c, r;
        M->get_as_coo(v, c, r);
        EXPECT(same(v, A.val) && c == A.col && static_cast<int>(r.size()) == nnz && r[A.row[5]] == 5);
        std::unique_ptr<lMatrix<T>> M2(M->CloneWithoutContent());
        EXPECT(M2->get_platform() == HIP && M2->get_nnz() == nnz);
        *M2 = *M;
        CPU_CSR_lMatrix<T> b2;
        b2.CloneFrom(*M2);
        EXPECT(same(b2.val, A.val));
        std::unique_ptr<lMatrix<T>> M3(init_matrix<T>(1, 1, 1, "M3", HIP));
        M3->ConvertFrom(A);
        CPU_CSR_lMatrix<T> b3;
        b3.CloneFrom(*M3);
        EXPECT(b3.row == A.row && same(b3.val, A.val));
        std::unique_ptr<lMatrix<double>> Md(init_matrix<double>(nnz, n, n, "Md", HIP));
        std::unique_ptr<lMatrix<float>> Mf(init_matrix<float>(nnz, n, n, "Mf", HIP));
        M->CastTo(*Md);
        M->CastTo(*Mf);
        CPU_CSR_lMatrix<float> bf;
        bf.CloneFrom(*Mf);
        for (int j = 0; j < nnz; ++j) EXPECT(bf.val[j] == static_cast<float>(A.val[j]));
        M2->CastFrom(*Md);
        b2.CloneFrom(*M2);
        EXPECT(same(b2.val, A.val));
        CPU_CSR_lMatrix<double> bd;
        M->CastTo(bd);
        for (int j = 0; j < nnz; ++j) EXPECT(bd.val[j] == static_cast<double>(A.val[j]));
        M->Sync();
        M->Zeros();
        M->CopyTo(back);
        for (int j = 0; j < nnz; ++j) EXPECT(back.val[j] == T(0));
    });
    run((pre + "products").c_str(), [&] {
        std::unique_ptr<lMatrix<T>> M(init_matrix<T>(nnz, n, n, "M", HIP));
        M->CloneFrom(A);
        auto X = hip_vec(x);
        auto Yv = hip_vec(rnd<T>(n, 9));
        std::vector<T> y = host_of(*Yv), ref(n);
        M->VectorMult(*X, Yv.get());
        O<T>::spmv(n, A.row.data(), A.col.data(), A.val.data(), x.data(), ref.data());
        EXPECT(same(host_of(*Yv), ref));
        M->VectorMultAdd(*X, Yv.get());
        O<T>::spmv_add(n, A.row.data(), A.col.data(), A.val.data(), x.data(), ref.data());
        EXPECT(same(host_of(*Yv), ref));
        M->VectorMultNoDiag(*X, Yv.get());
        O<T>::nodiag(n, A.row.data(), A.col.data(), A.val.data(), x.data(), ref.data());
        EXPECT(same(host_of(*Yv), ref));
        auto B = hip_vec(rnd<T>(n, 11));
        std::vector<T> bh = host_of(*B), r2(n);
        M->Residual(*B, *X, Yv.get());
        O<T>::spmv(n, A.row.data(), A.col.data(), A.val.data(), x.data(), r2.data());
        O<T>::scaleadd(n, r2.data(), T(-1), bh.data());
        EXPECT(same(host_of(*Yv), r2));
        M->Pjacobi(*X, Yv.get());
        O<T>::pjacobi(n, A.row.data(), A.col.data(), A.val.data(), x.data(), ref.data());
        EXPECT(same(host_of(*Yv), ref));
        M->Psgauss_seidel(*X, Yv.get());
        O<T>::sgs(n, A.row.data(), A.col.data(), A.val.data(), x.data(), ref.data());
        EXPECT(same(host_of(*Yv), ref));
    });
    run((pre + "value_ops").c_str(), [&] {
        std::unique_ptr<lMatrix<T>> M(init_matrix<T>(nnz, n, n, "M", HIP));
        M->CloneFrom(A);
        CPU_CSR_lMatrix<T> ref = A, back;
        const int rows[] = {0, 33, 33, n - 1};
        M->ZeroRows(rows, 4, T(2));
        ref.ZeroRows(rows, 4, T(2));
        M->CopyTo(back);
        EXPECT(same(back.val, ref.val));
        const int cols[] = {5, 77};
        M->ZeroCols(cols, 2, T(-3));
        ref.ZeroCols(cols, 2, T(-3));
        M->CopyTo(back);
        EXPECT(same(back.val, ref.val));
        M->Scale(T(0.5));
        ref.Scale(T(0.5));
        M->ScaleOffdiag(T(-1.25));
        ref.ScaleOffdiag(T(-1.25));
        M->CopyTo(back);
        EXPECT(same(back.val, ref.val));
        auto D = hip_vec(std::vector<T>(n - 20, T(0)));
        M->extract_diagelements(10, n - 10, D.get());
        CPU_lVector<T> dh(n - 20, "dh");
        ref.extract_diagelements(10, n - 10, &dh);
        EXPECT(same(host_of(*D), dh.buffer));
        M->extract_invdiagelements(10, n - 10, D.get());
        ref.extract_invdiagelements(10, n - 10, &dh);
        EXPECT(same(host_of(*D), dh.buffer));
        std::unique_ptr<lMatrix<T>> S(M->extract_submatrix(4, 0, 200, 180));
        std::unique_ptr<lMatrix<T>> Sr(ref.extract_submatrix(4, 0, 200, 180));
        CPU_CSR_lMatrix<T> sb;
        sb.CloneFrom(*S);
        const CPU_CSR_lMatrix<T> &sr = dynamic_cast<const CPU_CSR_lMatrix<T> &>(*Sr);
        EXPECT(S->get_platform() == HIP && sb.row == sr.row && sb.col == sr.col && same(sb.val, sr.val));
    });
    run((pre + "analysis_and_ilu").c_str(), [&] {
        std::unique_ptr<lMatrix<T>> M(init_matrix<T>(nnz, n, n, "M", HIP));
        M->CloneFrom(A);
        int nc = 0, *cs = nullptr, *cp = nullptr, onc = 0, *ocs = nullptr, *ocp = nullptr;
        M->Multicoloring(nc, &cs, &cp);
        A.Multicoloring(onc, &ocs, &ocp);
        EXPECT(nc == onc && std::vector<int>(cp, cp + n) == std::vector<int>(ocp, ocp + n));
        int nl = 0, *ls = nullptr, *lp = nullptr, onl = 0, *ols = nullptr, *olp = nullptr;
        M->Levelscheduling(nl, &ls, &lp);
        A.Levelscheduling(onl, &ols, &olp);
        EXPECT(nl == onl && std::vector<int>(lp, lp + n) == std::vector<int>(olp, olp + n));
        delete[] cs; delete[] cp; delete[] ocs; delete[] ocp;
        delete[] ls; delete[] lp; delete[] ols; delete[] olp;
        std::vector<int> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = n - 1 - i;
        M->Reorder(perm.data());
        CPU_CSR_lMatrix<T> pr = A, back;
        pr.Reorder(perm.data());
        M->CopyTo(back);
        EXPECT(back.row == pr.row && back.col == pr.col && same(back.val, pr.val));
        // ilu0 in place on the device, then ilu_solve with the factors
        std::unique_ptr<lMatrix<T>> L(init_matrix<T>(nnz, n, n, "LU", HIP));
        L->CloneFrom(A);
        L->ilu0();
        std::vector<T> lu = A.val;
        O<T>::ilu0(n, A.row.data(), A.col.data(), lu.data());
        L->CopyTo(back);
        EXPECT(same(back.val, lu));
        auto X = hip_vec(x), Z = hip_vec(std::vector<T>(n, T(0)));
        L->ilu_solve(*X, Z.get());
        std::vector<T> ref(n);
        O<T>::ilu_solve(n, A.row.data(), A.col.data(), lu.data(), x.data(), ref.data());
        EXPECT(same(host_of(*Z), ref));
        L->transpose_me();
        L->CopyTo(back);
        CPU_CSR_lMatrix<T> tr;
        tr.CloneFrom(A);
        tr.val = lu;
        tr.transpose_me();
        EXPECT(back.row == tr.row && back.col == tr.col && same(back.val, tr.val));
    });
}

template <typename T>
static void coupled_tests(const char *tname) {
    std::string pre = std::string(tname) + " ";
    run((pre + "diagonalize_rows").c_str(), [&] {
        // CoupledMatrix::diagonalize_rows (coupled_matrix.cc:550-575) on one
        // rank: ZeroRows of the diagonal block with the given value
        CPU_CSR_lMatrix<double> A64;
        fv_grid(6, 5, 4, A64);
        CPU_CSR_lMatrix<T> A;
        A.CastFrom(A64);
        CoupledMatrix<T> M;
        M.InitFrom(A);
        const int rows[] = {0, 1, 2, 119};
        M.diagonalize_rows(rows, 4, T(1));
        A.ZeroRows(rows, 4, T(1));
        CPU_CSR_lMatrix<T> back;
        M.diagonal().CopyTo(back);
        EXPECT(same(back.val, A.val));
    });
}

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "cpu") {
        cpu_tests();
    } else if (mode == "hip") {
        hip_vector_tests<double>("f64");
        hip_vector_tests<float>("f32");
        hip_matrix_tests<double>("f64");
        hip_matrix_tests<float>("f32");
        coupled_tests<double>("f64");
        coupled_tests<float>("f32");
        stop_platform();
    } else {
        std::printf("usage: lmp_iface cpu|hip\n");
        return 2;
    }
    std::printf("%s: %d checks, %d failed\n", mode.c_str(), g_checks, g_fail);
    return g_fail > 255 ? 255 : g_fail;
}
