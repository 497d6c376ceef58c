// lmp.h -- the local multi-platform layer (lVector / lMatrix /
// lPreconditioner) with a HIP platform, mirroring HiFlow3's lmp interface
// (src/linear_algebra/lmp/{la_global,lvector,lmatrix,lpreconditioner}.h).
// lVector and lMatrix declare every virtual of the reference's classes with
// the reference's signatures (lvector.h:39-336, lmatrix.h:40-448), so a
// platform class written against one compiles against the other.
//
// Platforms:
//   CPU : host containers for assembly and transfer (CPU_lVector,
//         CPU_CSR_lMatrix) and the host-side preconditioner setup that the
//         reference also runs on the CPU (colouring, level schedules, ILU(0)).
//         Arithmetic on CPU objects is not provided: solves run on HIP.
//   HIP : device objects reached only through the C ABI of libhfhip.so
//         (include/hfh.h).  Every compute call goes to a gfx950 kernel.
// Errors from the C ABI are thrown as hflmp::Error (the reference exit()s);
// an operation a class does not provide throws Error(HFH_ERR_UNSUPPORTED),
// where the reference logs "not supported" and exits.
#pragma once

#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hfh.h"

namespace hflmp {

// la_global.h:40-60 with the HIP platform added
enum PLATFORM { CPU = 0, GPU = 1, OPENCL = 2, HIP = 3 };
enum IMPLEMENTATION { NAIVE, BLAS, OPENMP, MKL, SCALAR, SCALAR_TEX, OPEN_CL, CUBLAS2, GFX950 };
enum MATRIX_FORMAT { DENSE, CSR, COO, ELL };

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void check(int rc, const char *what);
[[noreturn]] void unsupported(const std::string &cls, const char *op);

// platform_management.cc:135-187: one HIP context (device + stream) per process
void init_platform(int device = 0);
void stop_platform();
hfh_ctx *hip_context();

template <typename T> struct dtype_of;
template <> struct dtype_of<double> { static constexpr int value = HFH_F64; };
template <> struct dtype_of<float> { static constexpr int value = HFH_F32; };

// Rotg / Rotmg: the scalar plane rotations of BLAS level 1 (no vector data):
// Rotg follows CPUsimple_lvector.cc:492-536 (the LINPACK drotg construction),
// Rotmg the modified Givens construction of Hanson & Krogh as in the
// reference BLAS drotmg (the reference's CPUsimple platform leaves it
// unimplemented, CPUsimple_lvector.cc:623-628).
template <typename T> void rotg(T *sa, T *sb, T *sc, T *ss);
template <typename T> void rotmg(T *sd1, T *sd2, T *sx1, T sy1, T *param);

// ---------------------------------------------------------------- lVector

template <typename T>
class lVector {
  public:
    lVector() {}
    virtual ~lVector() {}
    int get_size() const { return size_; }
    std::string get_name() const { return name_; }
    void print(std::ostream &out = std::cout) const;
    IMPLEMENTATION get_implementation() const { return impl_; }
    std::string get_implementation_name() const { return implementation_name_; }
    PLATFORM get_platform() const { return platform_; }
    std::string get_platform_name() const { return platform_name_; }

    virtual void add_value(const int i, T val) = 0;
    virtual void add_values(const int *indices, int length, const T *values) = 0;
    virtual void SetValues(const int *index, const int size, const T *values) = 0;
    virtual void GetValues(const int *index, const int size, T *values) const = 0;
    virtual void SetBlockValues(const int start_i, const int end_i, const T *values) = 0;
    virtual void GetBlockValues(const int start_i, const int end_i, T *values) const = 0;
    virtual void set_indexset(const int *indexset, const int size) = 0;
    virtual void get_indexset(int *indexset) const = 0;
    virtual int get_indexset_size() const = 0;
    virtual void CopyFromIndexset(const lVector<T> &vec);
    virtual void GetIndexedValues(T *values) const = 0;
    virtual void SetIndexedValues(const T *values) = 0;
    virtual lVector<T> *extract_subvector(const int start_i, const int end_i) const = 0;
    virtual void partial_replace_subvector(const int start_i, const int start_sub_vec, const int size,
                                           const lVector<T> &sub_vec) = 0;
    virtual void partial_add_subvector(const int start_i, const int start_sub_vec, const int size, const T weight,
                                       const lVector<T> &sub_vec) = 0;
    virtual void Init(const int size, const std::string name) = 0;
    virtual void Clear() = 0;
    virtual void Zeros() = 0;
    virtual void ReadFile(const char *filename);
    virtual void WriteFile(const char *filename);
    virtual void Reorder(const int *index) = 0;
    virtual lVector<T> &operator=(const lVector<T> &vec2) = 0;
    virtual void CopyFrom(const lVector<T> &vec) = 0;
    virtual void CopyTo(lVector<T> &vec) const = 0;
    virtual void CopyStructureFrom(const lVector<T> &vec2);
    virtual void CastFrom(const lVector<double> &vec) = 0;
    virtual void CastFrom(const lVector<float> &vec) = 0;
    virtual void CastTo(lVector<double> &vec) const = 0;
    virtual void CastTo(lVector<float> &vec) const = 0;
    virtual void CloneFrom(const lVector<T> &vec);
    virtual lVector<T> *CloneWithoutContent() const;
    virtual void Sync() const {}
    virtual void ElementWiseMult(const lVector<T> &vec) const = 0;
    virtual int ArgMin() const = 0;
    virtual int ArgMax() const = 0;
    virtual T Norm1() const = 0;
    virtual T Norm2() const = 0;
    virtual T NormMax() const = 0;
    virtual T Dot(const lVector<T> &vec) const = 0;
    virtual void Scale(const T scalar) = 0;
    virtual void Axpy(const lVector<T> &vec, const T scalar) = 0;
    virtual void ScaleAdd(const T scalar, const lVector<T> &vec) = 0;  // this = scalar*this + vec
    virtual void Rot(lVector<T> *vec, const T &sc, const T &ss) = 0;
    virtual void Rotg(T *sa, T *sb, T *sc, T *ss) const = 0;
    virtual void Rotm(lVector<T> *vec, const T &sparam) = 0;
    virtual void Rotmg(T *sd1, T *sd2, T *x1, const T &x2, T *sparam) const = 0;
    virtual T *GetBuffer() const { return 0; }

  protected:
    int size_ = 0;
    std::string name_;
    PLATFORM platform_ = CPU;
    IMPLEMENTATION impl_ = NAIVE;
    std::string platform_name_, implementation_name_;
};

// Host vector: storage, assembly and transfer.
template <typename T>
class CPU_lVector : public lVector<T> {
  public:
    CPU_lVector();
    CPU_lVector(int size, const std::string &name) : CPU_lVector() { Init(size, name); }
    CPU_lVector(const CPU_lVector &v) : CPU_lVector() { *this = v; }
    CPU_lVector &operator=(const CPU_lVector &v);
    void add_value(const int i, T val) override;
    void add_values(const int *indices, int length, const T *values) override;
    void SetValues(const int *index, const int size, const T *values) override;
    void GetValues(const int *index, const int size, T *values) const override;
    void SetBlockValues(const int start_i, const int end_i, const T *values) override;
    void GetBlockValues(const int start_i, const int end_i, T *values) const override;
    void set_indexset(const int *indexset, const int size) override;
    void get_indexset(int *indexset) const override;
    int get_indexset_size() const override { return static_cast<int>(indexset_.size()); }
    void GetIndexedValues(T *values) const override;
    void SetIndexedValues(const T *values) override;
    lVector<T> *extract_subvector(const int start_i, const int end_i) const override;
    void partial_replace_subvector(const int start_i, const int start_sub_vec, const int size,
                                   const lVector<T> &sub_vec) override;
    void partial_add_subvector(const int start_i, const int start_sub_vec, const int size, const T weight,
                               const lVector<T> &sub_vec) override;
    void Init(const int size, const std::string name) override;
    void Clear() override;
    void Zeros() override;
    void ReadFile(const char *filename) override;
    void WriteFile(const char *filename) override;
    void Reorder(const int *index) override;
    lVector<T> &operator=(const lVector<T> &vec2) override;
    void CopyFrom(const lVector<T> &vec) override;
    void CopyTo(lVector<T> &vec) const override { vec.CopyFrom(*this); }
    void CastFrom(const lVector<double> &vec) override;
    void CastFrom(const lVector<float> &vec) override;
    void CastTo(lVector<double> &vec) const override { vec.CastFrom(*this); }
    void CastTo(lVector<float> &vec) const override { vec.CastFrom(*this); }
    lVector<T> *CloneWithoutContent() const override;
    void ElementWiseMult(const lVector<T> &) const override { no_compute("ElementWiseMult"); }
    int ArgMin() const override { no_compute("ArgMin"); }
    int ArgMax() const override { no_compute("ArgMax"); }
    T Norm1() const override { no_compute("Norm1"); }
    T Norm2() const override { no_compute("Norm2"); }
    T NormMax() const override { no_compute("NormMax"); }
    T Dot(const lVector<T> &) const override { no_compute("Dot"); }
    void Scale(const T) override { no_compute("Scale"); }
    void Axpy(const lVector<T> &, const T) override { no_compute("Axpy"); }
    void ScaleAdd(const T, const lVector<T> &) override { no_compute("ScaleAdd"); }
    void Rot(lVector<T> *, const T &, const T &) override { no_compute("Rot"); }
    void Rotg(T *sa, T *sb, T *sc, T *ss) const override { rotg(sa, sb, sc, ss); }
    void Rotm(lVector<T> *, const T &) override { no_compute("Rotm"); }
    void Rotmg(T *sd1, T *sd2, T *x1, const T &x2, T *sparam) const override { rotmg(sd1, sd2, x1, x2, sparam); }
    T *GetBuffer() const override { return const_cast<T *>(buffer.data()); }
    std::vector<T> buffer;

  private:
    std::vector<int> indexset_;
    [[noreturn]] static void no_compute(const char *op);
};

// Device vector on the gfx950 platform.
template <typename T>
class HIP_lVector : public lVector<T> {
  public:
    HIP_lVector();
    HIP_lVector(int size, const std::string &name) : HIP_lVector() { Init(size, name); }
    ~HIP_lVector() override { Clear(); }
    HIP_lVector(const HIP_lVector &) = delete;
    void add_value(const int i, T val) override;
    void add_values(const int *indices, int length, const T *values) override;
    void SetValues(const int *index, const int size, const T *values) override;
    void GetValues(const int *index, const int size, T *values) const override;
    void SetBlockValues(const int start_i, const int end_i, const T *values) override;
    void GetBlockValues(const int start_i, const int end_i, T *values) const override;
    void set_indexset(const int *indexset, const int size) override;
    void get_indexset(int *indexset) const override;
    int get_indexset_size() const override;
    void GetIndexedValues(T *values) const override;
    void SetIndexedValues(const T *values) override;
    lVector<T> *extract_subvector(const int start_i, const int end_i) const override;
    void partial_replace_subvector(const int start_i, const int start_sub_vec, const int size,
                                   const lVector<T> &sub_vec) override;
    void partial_add_subvector(const int start_i, const int start_sub_vec, const int size, const T weight,
                               const lVector<T> &sub_vec) override;
    void Init(const int size, const std::string name) override;
    void Clear() override;
    void Zeros() override;
    void ReadFile(const char *filename) override;
    void WriteFile(const char *filename) override;
    void Reorder(const int *index) override;
    lVector<T> &operator=(const lVector<T> &vec2) override;
    HIP_lVector &operator=(const HIP_lVector &vec2) {
        CopyFrom(vec2);
        return *this;
    }
    void CopyFrom(const lVector<T> &vec) override;
    void CopyTo(lVector<T> &vec) const override;
    void CastFrom(const lVector<double> &vec) override;
    void CastFrom(const lVector<float> &vec) override;
    void CastTo(lVector<double> &vec) const override;
    void CastTo(lVector<float> &vec) const override;
    lVector<T> *CloneWithoutContent() const override;
    void Sync() const override;
    void ElementWiseMult(const lVector<T> &vec) const override;
    int ArgMin() const override;
    int ArgMax() const override;
    T Norm1() const override;
    T Norm2() const override;
    T NormMax() const override;
    T Dot(const lVector<T> &vec) const override;
    void Scale(const T scalar) override;
    void Axpy(const lVector<T> &vec, const T scalar) override;
    void ScaleAdd(const T scalar, const lVector<T> &vec) override;
    void Rot(lVector<T> *vec, const T &sc, const T &ss) override;
    void Rotg(T *sa, T *sb, T *sc, T *ss) const override { rotg(sa, sb, sc, ss); }
    void Rotm(lVector<T> *vec, const T &sparam) override;
    void Rotmg(T *sd1, T *sd2, T *x1, const T &x2, T *sparam) const override { rotmg(sd1, sd2, x1, x2, sparam); }
    hfh_vec *handle() const { return h_; }

  private:
    hfh_vec *h_ = nullptr;
};

// ---------------------------------------------------------------- lMatrix

template <typename T>
class lMatrix {
  public:
    lMatrix() {}
    virtual ~lMatrix() {}
    virtual void Compress() {}
    int get_num_col() const { return ncol_; }
    int get_num_row() const { return nrow_; }
    void swap_dimensions() { std::swap(nrow_, ncol_); }
    int get_nnz() const { return nnz_; }
    MATRIX_FORMAT get_matrix_format() const { return format_; }
    std::string get_matrix_format_name() const;
    // get_as_coo: values, then the column indices, then the row indices of
    // the entries in stored order (the argument order CPU_CSR_lMatrix fills,
    // lmatrix_csr_cpu.cc:79-98, and coupled_matrix.cc:1147 reads)
    virtual void get_as_coo(std::vector<T> &vals, std::vector<int> &cols, std::vector<int> &rows) const;
    std::string get_name() const { return name_; }
    virtual void print(std::ostream &out = std::cout) const;
    IMPLEMENTATION get_implementation() const { return impl_; }
    std::string get_implementation_name() const { return implementation_name_; }
    PLATFORM get_platform() const { return platform_; }
    std::string get_platform_name() const { return platform_name_; }

    virtual void Init(const int init_nnz, const int init_num_row, const int init_num_col,
                      const std::string init_name) = 0;
    virtual void Clear() = 0;
    virtual void ZeroRows(const int *index_set, const int size, const T alpha);
    virtual void ZeroCols(const int *index_set, const int size, const T alpha);
    virtual void Zeros() = 0;
    virtual void Reorder(const int *index);
    virtual void Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const;
    virtual void Levelscheduling(int &nlevels, int **level_sizes, int **permut_index) const;
    virtual void ilu0();
    virtual void ilup(const int p);
    virtual void ilup(const lMatrix<T> &mat, const int p);
    virtual void ilu_solve(const lVector<T> &invec, lVector<T> *outvec) const;
    virtual void Scale(const T alpha);
    virtual void ScaleOffdiag(const T alpha);
    virtual void init_structure(const int *rows, const int *cols);
    virtual void add_value(const int row, const int col, const T val);
    virtual void add_values(const int *rows, int num_rows, const int *cols, int num_cols, const T *values);
    virtual void get_value(const int row, const int col, T *val) const;
    virtual void get_add_values(const int *rows, int num_rows, const int *cols, int num_cols, const int *cols_target,
                                int num_cols_target, T *values) const;
    virtual void VectorMultAdd_submatrix(const int *rows, int num_rows, const int *cols, int num_cols,
                                         const int *cols_input, const T *in_values, T *out_values) const;
    virtual void VectorMultAdd_submatrix_vanka(const int *rows, int num_rows, const lVector<T> &invec,
                                               T *out_values) const;
    virtual void transpose_me();
    virtual bool issymmetric();
    virtual void compress_me();
    virtual void delete_diagonal();
    virtual void delete_offdiagonal();
    virtual void delete_lower_triangular();
    virtual void delete_strictly_lower_triangular();
    virtual void delete_upper_triangular();
    virtual void delete_strictly_upper_triangular();
    virtual lMatrix<T> *extract_submatrix(const int start_row, const int start_col, const int end_row,
                                          const int end_col) const;
    virtual void extract_diagelements(const int start_i, const int end_i, lVector<T> *vec) const;
    virtual void extract_invdiagelements(const int start_i, const int end_i, lVector<T> *vec) const;
    virtual lMatrix<T> &operator=(const lMatrix<T> &mat2) = 0;
    virtual void CopyFrom(const lMatrix<T> &mat2) = 0;
    virtual void CopyTo(lMatrix<T> &mat2) const = 0;
    virtual void CastFrom(const lMatrix<double> &other) = 0;
    virtual void CastFrom(const lMatrix<float> &other) = 0;
    virtual void CastTo(lMatrix<double> &other) const = 0;
    virtual void CastTo(lMatrix<float> &other) const = 0;
    virtual void ConvertFrom(const lMatrix<T> &mat2) = 0;
    virtual void CopyStructureFrom(const lMatrix<T> &mat2) = 0;
    virtual void CopyStructureTo(lMatrix<T> &mat2) const = 0;
    virtual void CloneFrom(const lMatrix<T> &mat2);
    virtual lMatrix<T> *CloneWithoutContent() const;
    virtual void Sync() const {}
    virtual void ReadFile(const char *filename);
    virtual void WriteFile(const char *filename) const;
    virtual void VectorMult(const lVector<T> &invec, lVector<T> *outvec) const = 0;
    virtual void VectorMultAdd(const lVector<T> &invec, lVector<T> *outvec) const = 0;
    virtual void VectorMultNoDiag(const lVector<T> &in, lVector<T> *out) const = 0;
    // Residual (lmatrix.cc:453-462): res = A x, then res = -res + b
    virtual void Residual(const lVector<T> &b, const lVector<T> &x, lVector<T> *res) const;
    virtual lMatrix<T> *MatrixMult(const lMatrix<T> &inmat) const;
    virtual lMatrix<T> *MatrixMultSupStructure(const lMatrix<T> &inmat) const;
    virtual lMatrix<T> *MatrixSupSPower(const int p) const;
    virtual void MatrixAdd(const lMatrix<T> &inmat);
    virtual void GershgorinSpectrum(T *lambda_min, T *lambda_max) const;
    virtual void Pjacobi(const lVector<T> &invec, lVector<T> *outvec) const;
    virtual void Pgauss_seidel(const lVector<T> &invec, lVector<T> *outvec) const;
    virtual void Psgauss_seidel(const lVector<T> &invec, lVector<T> *outvec) const;
    virtual void BlockPsgauss_seidel(const lVector<T> &invec, lVector<T> *outvec, const int start_i,
                                     const int end_i) const;
    virtual void BlocksPsgauss_seidel(const lVector<T> &invec, lVector<T> *outvec, const int num_blocks) const;
    virtual void Psor(const T omega, const lVector<T> &invec, lVector<T> *outvec) const;
    virtual void Pssor(const T omega, const lVector<T> &invec, lVector<T> *outvec) const;

  protected:
    [[noreturn]] void no(const char *op) const { unsupported("lMatrix '" + name_ + "'", op); }
    int nrow_ = 0, ncol_ = 0, nnz_ = 0;
    std::string name_;
    MATRIX_FORMAT format_ = CSR;
    PLATFORM platform_ = CPU;
    IMPLEMENTATION impl_ = NAIVE;
    std::string platform_name_, implementation_name_;
    bool symmetric_ = false;
};

// Host CSR: assembly (init_structure, add_value(s), ZeroRows) and the
// CPU-side analysis used by the preconditioner setup.
template <typename T>
class CPU_CSR_lMatrix : public lMatrix<T> {
  public:
    CPU_CSR_lMatrix();
    CPU_CSR_lMatrix(const CPU_CSR_lMatrix &m) : CPU_CSR_lMatrix() { *this = m; }
    CPU_CSR_lMatrix &operator=(const CPU_CSR_lMatrix &m);
    void get_as_coo(std::vector<T> &vals, std::vector<int> &cols, std::vector<int> &rows) const override;
    void Init(const int nnz, const int nrow, const int ncol, const std::string name) override;
    void Clear() override;
    void Zeros() override;
    lMatrix<T> &operator=(const lMatrix<T> &mat2) override;
    void CopyFrom(const lMatrix<T> &mat) override;
    void CopyTo(lMatrix<T> &mat) const override { mat.CopyFrom(*this); }
    void CastFrom(const lMatrix<double> &other) override;
    void CastFrom(const lMatrix<float> &other) override;
    void CastTo(lMatrix<double> &other) const override { other.CastFrom(*this); }
    void CastTo(lMatrix<float> &other) const override { other.CastFrom(*this); }
    void ConvertFrom(const lMatrix<T> &mat2) override;
    void CopyStructureFrom(const lMatrix<T> &mat) override;
    void CopyStructureTo(lMatrix<T> &mat) const override { mat.CopyStructureFrom(*this); }
    lMatrix<T> *CloneWithoutContent() const override;
    // TransformFromCOO / init_structure (lmatrix_csr_cpu.cc:951-1044, 1101-1118): COO -> CSR with each
    // row's columns ordered by the reference's selection sort
    void init_structure(const int *rows, const int *cols) override;
    void TransformFromCOO(const int *rows, const int *cols, const T *data, int nrow, int ncol, int nnz);
    void add_value(const int row, const int col, const T val) override;
    void add_values(const int *rows, int nr, const int *cols, int nc, const T *values) override;
    void get_value(const int row, const int col, T *val) const override;
    // ZeroRows / ZeroCols (lmatrix_csr_cpu.cc:1047-1098): Dirichlet rows / columns -> alpha on the diagonal
    void ZeroRows(const int *index_set, const int size, const T alpha) override;
    void ZeroCols(const int *index_set, const int size, const T alpha) override;
    void Scale(const T alpha) override;
    void ScaleOffdiag(const T alpha) override;
    void Multicoloring(int &ncolors, std::vector<int> &color_sizes, std::vector<int> &perm) const;
    void Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const override;
    void Levelscheduling(int &nlevels, int **level_sizes, int **permut_index) const override;
    void Reorder(const int *perm) override;
    // transpose_me (lmatrix_csr_cpu.cc:1866-1878): through the COO triplets
    void transpose_me() override;
    void ilu0() override;
    lMatrix<T> *extract_submatrix(const int start_row, const int start_col, const int end_row,
                                  const int end_col) const override;
    void extract_diagelements(const int start_i, const int end_i, lVector<T> *vec) const override;
    void extract_invdiagelements(const int start_i, const int end_i, lVector<T> *vec) const override;
    void ReadFile(const char *filename) override;
    void WriteFile(const char *filename) const override;
    void VectorMult(const lVector<T> &, lVector<T> *) const override { no_compute("VectorMult"); }
    void VectorMultAdd(const lVector<T> &, lVector<T> *) const override { no_compute("VectorMultAdd"); }
    void VectorMultNoDiag(const lVector<T> &, lVector<T> *) const override { no_compute("VectorMultNoDiag"); }
    std::vector<int> row, col;
    std::vector<T> val;

  private:
    [[noreturn]] static void no_compute(const char *op);
};

template <typename T>
class HIP_CSR_lMatrix : public lMatrix<T> {
  public:
    HIP_CSR_lMatrix();
    ~HIP_CSR_lMatrix() override { Clear(); }
    HIP_CSR_lMatrix(const HIP_CSR_lMatrix &) = delete;
    void get_as_coo(std::vector<T> &vals, std::vector<int> &cols, std::vector<int> &rows) const override;
    void Init(const int nnz, const int nrow, const int ncol, const std::string name) override;
    void Clear() override;
    void Zeros() override;
    lMatrix<T> &operator=(const lMatrix<T> &mat2) override;
    HIP_CSR_lMatrix &operator=(const HIP_CSR_lMatrix &mat2) {
        CopyFrom(mat2);
        return *this;
    }
    void CopyFrom(const lMatrix<T> &mat) override;
    void CopyTo(lMatrix<T> &mat) const override;
    // CastFrom / CastTo (lmatrix.h:297-301): f64 <-> f32, the structure with it
    void CastFrom(const lMatrix<double> &other) override;
    void CastFrom(const lMatrix<float> &other) override;
    void CastTo(lMatrix<double> &other) const override;
    void CastTo(lMatrix<float> &other) const override;
    void ConvertFrom(const lMatrix<T> &mat2) override;
    void CopyStructureFrom(const lMatrix<T> &mat) override;
    void CopyStructureTo(lMatrix<T> &mat) const override;
    lMatrix<T> *CloneWithoutContent() const override;
    void Sync() const override;
    void ZeroRows(const int *index_set, const int size, const T alpha) override;
    void ZeroCols(const int *index_set, const int size, const T alpha) override;
    void Scale(const T alpha) override;
    void ScaleOffdiag(const T alpha) override;
    void Reorder(const int *index) override;
    void Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const override;
    void Levelscheduling(int &nlevels, int **level_sizes, int **permut_index) const override;
    void transpose_me() override;
    void ilu0() override;
    void ilu_solve(const lVector<T> &invec, lVector<T> *outvec) const override;
    lMatrix<T> *extract_submatrix(const int start_row, const int start_col, const int end_row,
                                  const int end_col) const override;
    void extract_diagelements(const int start_i, const int end_i, lVector<T> *vec) const override;
    void extract_invdiagelements(const int start_i, const int end_i, lVector<T> *vec) const override;
    void VectorMult(const lVector<T> &in, lVector<T> *out) const override;
    void VectorMultAdd(const lVector<T> &in, lVector<T> *out) const override;
    void VectorMultNoDiag(const lVector<T> &in, lVector<T> *out) const override;
    void Pjacobi(const lVector<T> &invec, lVector<T> *outvec) const override;
    void Psgauss_seidel(const lVector<T> &invec, lVector<T> *outvec) const override;
    hfh_csr *handle() const { return h_; }
    const std::vector<int> &host_row() const { return hrow_; }
    const std::vector<int> &host_col() const { return hcol_; }

  private:
    void adopt(hfh_csr *h, const std::vector<int> &row, const std::vector<int> &col);
    void drop_applies() const;
    hfh_csr *h_ = nullptr;
    std::vector<int> hrow_, hcol_;  // pattern kept for host-side analysis
    bool structure_ = false;
    // device applies built on first use: ilu_solve on these factors, and
    // Psgauss_seidel on this operator (both follow later value changes)
    mutable hfh_prec *ilu_ = nullptr, *sgs_ = nullptr;
    template <typename U> friend class HIP_CSR_lMatrix;
};

// init_vec_mat.cc:47-415: platform factories
template <typename T>
lVector<T> *init_vector(int size, const std::string &name, PLATFORM platform, IMPLEMENTATION impl = NAIVE);
template <typename T>
lMatrix<T> *init_matrix(int nnz, int nrow, int ncol, const std::string &name, PLATFORM platform,
                        IMPLEMENTATION impl = NAIVE, MATRIX_FORMAT format = CSR);

// ---------------------------------------------------------------- lPreconditioner

template <typename T>
class lPreconditioner {
  public:
    virtual ~lPreconditioner() {}
    virtual void SetupOperator(const lMatrix<T> &op) { op_ = &op; state_ = false; }
    virtual void Build() = 0;
    virtual void ApplylPreconditioner(const lVector<T> &input, lVector<T> *output) = 0;
    bool GetState() const { return state_; }
    std::string name() const { return name_; }
    // device handle for the fused drivers (nullptr for identity)
    virtual hfh_prec *handle() const { return nullptr; }

  protected:
    const lMatrix<T> *op_ = nullptr;
    bool state_ = false;
    std::string name_;
};

// Shared implementation: a device preconditioner built from the operator.
template <typename T>
class lPreconditioner_Device : public lPreconditioner<T> {
  public:
    ~lPreconditioner_Device() override { Clear(); }
    void Clear();
    void ApplylPreconditioner(const lVector<T> &input, lVector<T> *output) override;
    hfh_prec *handle() const override { return p_; }

  protected:
    const HIP_CSR_lMatrix<T> &hip_op() const;
    hfh_prec *p_ = nullptr;
};

// lpreconditioner.cc:104-166
template <typename T>
class lPreconditioner_Jacobi : public lPreconditioner_Device<T> {
  public:
    lPreconditioner_Jacobi() { this->name_ = "Jacobi"; }
    void Build() override;
};

// lpreconditioner.cc:200-230 (natural order, level-scheduled on the device)
template <typename T>
class lPreconditioner_SymmetricGaussSeidel : public lPreconditioner_Device<T> {
  public:
    lPreconditioner_SymmetricGaussSeidel() { this->name_ = "Symmetric Gauss-Seidel"; }
    void Build() override;
};

// lpreconditioner.cc:354-425 with p = 0: host ilu0, device ilu_solve
template <typename T>
class lPreconditioner_ILUp : public lPreconditioner_Device<T> {
  public:
    lPreconditioner_ILUp() { this->name_ = "ILU(p)"; }
    void Init(int p) { if (p != 0) throw Error(HFH_ERR_UNSUPPORTED, "ILU(p): only p = 0 is provided"); }
    void Build() override;

  private:
    std::unique_ptr<HIP_CSR_lMatrix<T>> LU_;
};

// lpreconditioner_mc.cc + lpreconditioner_mc_sgs.cc: multicolour SGS on the
// colour-permuted operator (the permutation comes from Analyse on the
// original operator, as PreconditionerMultiColoring::Preprocess does).
template <typename T>
class lPreconditioner_MultiColoring_SymmetricGaussSeidel : public lPreconditioner_Device<T> {
  public:
    lPreconditioner_MultiColoring_SymmetricGaussSeidel() {
        this->name_ = "Multi-Coloring / Symmetric Gauss-Seidel";
    }
    void SetRelax(T omega) { omega_ = omega; }
    // colour sizes of the permuted operator (from Multicoloring)
    void SetColors(const std::vector<int> &color_sizes) { sizes_ = color_sizes; }
    void Build() override;

  private:
    std::vector<int> sizes_;
    T omega_ = 1;
};

}  // namespace hflmp
