// lmp.h -- the local multi-platform layer (lVector / lMatrix /
// lPreconditioner) with a HIP platform, mirroring HiFlow3's lmp interface
// (src/linear_algebra/lmp/{la_global,lvector,lmatrix,lpreconditioner}.h).
//
// Platforms:
//   CPU : host containers for assembly and transfer (CPU_lVector,
//         CPU_CSR_lMatrix) and the host-side preconditioner setup that the
//         reference also runs on the CPU (colouring, level schedules, ILU(0)).
//         Arithmetic on CPU objects is not provided: solves run on HIP.
//   HIP : device objects reached only through the C ABI of libhfhip.so
//         (include/hfh.h).  Every compute call goes to a gfx950 kernel.
// Errors from the C ABI are thrown as hflmp::Error (the reference exit()s).
#pragma once

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

// platform_management.cc:135-187: one HIP context (device + stream) per process
void init_platform(int device = 0);
void stop_platform();
hfh_ctx *hip_context();

template <typename T> struct dtype_of;
template <> struct dtype_of<double> { static constexpr int value = HFH_F64; };
template <> struct dtype_of<float> { static constexpr int value = HFH_F32; };

// ---------------------------------------------------------------- lVector

template <typename T>
class lVector {
  public:
    virtual ~lVector() {}
    int get_size() const { return size_; }
    std::string get_name() const { return name_; }
    PLATFORM get_platform() const { return platform_; }
    IMPLEMENTATION get_implementation() const { return impl_; }

    virtual void Init(int size, const std::string &name) = 0;
    virtual void Clear() = 0;
    virtual void Zeros() = 0;
    virtual void SetBlockValues(int start_i, int end_i, const T *values) = 0;
    virtual void GetBlockValues(int start_i, int end_i, T *values) const = 0;
    virtual void SetValues(const int *index, int size, const T *values) = 0;
    virtual void GetValues(const int *index, int size, T *values) const = 0;
    virtual void CopyFrom(const lVector<T> &vec) = 0;
    virtual void CopyTo(lVector<T> &vec) const { vec.CopyFrom(*this); }
    virtual void CastFrom(const lVector<double> &vec) = 0;
    virtual void CastFrom(const lVector<float> &vec) = 0;
    virtual lVector<T> *CloneWithoutContent() const = 0;
    void CloneFrom(const lVector<T> &vec) {
        if (get_size() != vec.get_size()) Init(vec.get_size(), vec.get_name());
        CopyFrom(vec);
    }
    virtual void partial_replace_subvector(int start_i, int start_sub, int size, const lVector<T> &sub) = 0;
    virtual void Reorder(const int *index) = 0;
    virtual void Sync() const {}
    // BLAS-1 (lvector.h:262-291)
    virtual T Dot(const lVector<T> &vec) const = 0;
    virtual T Norm2() const = 0;
    virtual void Axpy(const lVector<T> &vec, T scalar) = 0;
    virtual void ScaleAdd(T scalar, const lVector<T> &vec) = 0;  // this = scalar*this + vec
    virtual void Scale(T scalar) = 0;
    virtual void ElementWiseMult(const lVector<T> &vec) = 0;

  protected:
    int size_ = 0;
    std::string name_;
    PLATFORM platform_ = CPU;
    IMPLEMENTATION impl_ = NAIVE;
};

// Host vector: storage and transfer only.
template <typename T>
class CPU_lVector : public lVector<T> {
  public:
    CPU_lVector() { this->platform_ = CPU; this->impl_ = NAIVE; }
    CPU_lVector(int size, const std::string &name) : CPU_lVector() { Init(size, name); }
    void Init(int size, const std::string &name) override;
    void Clear() override { buffer.clear(); this->size_ = 0; }
    void Zeros() override;
    void SetBlockValues(int s, int e, const T *v) override;
    void GetBlockValues(int s, int e, T *v) const override;
    void SetValues(const int *idx, int n, const T *v) override;
    void GetValues(const int *idx, int n, T *v) const override;
    void add_value(int i, T v);
    void CopyFrom(const lVector<T> &vec) override;
    void CastFrom(const lVector<double> &vec) override;
    void CastFrom(const lVector<float> &vec) override;
    lVector<T> *CloneWithoutContent() const override;
    void partial_replace_subvector(int s, int ss, int n, const lVector<T> &sub) override;
    void Reorder(const int *index) override;
    T Dot(const lVector<T> &) const override { no_compute("Dot"); return 0; }
    T Norm2() const override { no_compute("Norm2"); return 0; }
    void Axpy(const lVector<T> &, T) override { no_compute("Axpy"); }
    void ScaleAdd(T, const lVector<T> &) override { no_compute("ScaleAdd"); }
    void Scale(T) override { no_compute("Scale"); }
    void ElementWiseMult(const lVector<T> &) override { no_compute("ElementWiseMult"); }
    std::vector<T> buffer;

  private:
    [[noreturn]] static void no_compute(const char *op);
};

// Device vector on the gfx950 platform.
template <typename T>
class HIP_lVector : public lVector<T> {
  public:
    HIP_lVector() { this->platform_ = HIP; this->impl_ = GFX950; }
    HIP_lVector(int size, const std::string &name) : HIP_lVector() { Init(size, name); }
    ~HIP_lVector() override { Clear(); }
    HIP_lVector(const HIP_lVector &) = delete;
    HIP_lVector &operator=(const HIP_lVector &) = delete;
    void Init(int size, const std::string &name) override;
    void Clear() override;
    void Zeros() override;
    void SetBlockValues(int s, int e, const T *v) override;
    void GetBlockValues(int s, int e, T *v) const override;
    void SetValues(const int *idx, int n, const T *v) override;
    void GetValues(const int *idx, int n, T *v) const override;
    void CopyFrom(const lVector<T> &vec) override;
    void CastFrom(const lVector<double> &vec) override;
    void CastFrom(const lVector<float> &vec) override;
    lVector<T> *CloneWithoutContent() const override;
    void partial_replace_subvector(int s, int ss, int n, const lVector<T> &sub) override;
    void Reorder(const int *index) override;
    void Sync() const override;
    T Dot(const lVector<T> &vec) const override;
    T Norm2() const override;
    void Axpy(const lVector<T> &vec, T scalar) override;
    void ScaleAdd(T scalar, const lVector<T> &vec) override;
    void Scale(T scalar) override;
    void ElementWiseMult(const lVector<T> &vec) override;
    hfh_vec *handle() const { return h_; }

  private:
    hfh_vec *h_ = nullptr;
};

// ---------------------------------------------------------------- lMatrix

template <typename T>
class lMatrix {
  public:
    virtual ~lMatrix() {}
    int get_num_row() const { return nrow_; }
    int get_num_col() const { return ncol_; }
    int get_nnz() const { return nnz_; }
    PLATFORM get_platform() const { return platform_; }
    virtual void Init(int nnz, int nrow, int ncol, const std::string &name) = 0;
    virtual void Clear() = 0;
    virtual void CopyStructureFrom(const lMatrix<T> &mat) = 0;
    virtual void CopyFrom(const lMatrix<T> &mat) = 0;
    void CloneFrom(const lMatrix<T> &mat) {
        Init(mat.get_nnz(), mat.get_num_row(), mat.get_num_col(), mat.name_);
        CopyStructureFrom(mat);
        CopyFrom(mat);
    }
    virtual lMatrix<T> *CloneWithoutContent() const = 0;
    virtual void VectorMult(const lVector<T> &in, lVector<T> *out) const = 0;
    virtual void VectorMultAdd(const lVector<T> &in, lVector<T> *out) const = 0;
    virtual void extract_invdiagelements(int start_i, int end_i, lVector<T> *vec) const = 0;

  protected:
    int nrow_ = 0, ncol_ = 0, nnz_ = 0;
    std::string name_;
    PLATFORM platform_ = CPU;
};

// Host CSR: assembly (init_structure, add_value(s), ZeroRows) and the
// CPU-side analysis used by the preconditioner setup.
template <typename T>
class CPU_CSR_lMatrix : public lMatrix<T> {
  public:
    CPU_CSR_lMatrix() { this->platform_ = CPU; }
    void Init(int nnz, int nrow, int ncol, const std::string &name) override;
    void Clear() override;
    void CopyStructureFrom(const lMatrix<T> &mat) override;
    void CopyFrom(const lMatrix<T> &mat) override;
    lMatrix<T> *CloneWithoutContent() const override;
    // TransformFromCOO / init_structure (lmatrix_csr_cpu.cc:951-1044, 1101-1118): COO -> sorted CSR
    void init_structure(const int *rows, const int *cols);
    void TransformFromCOO(const int *rows, const int *cols, const T *data, int nrow, int ncol, int nnz);
    void Zeros();
    void add_value(int row, int col, T val);
    void add_values(const int *rows, int nr, const int *cols, int nc, const T *values);
    void get_value(int row, int col, T *val) const;
    // ZeroRows (lmatrix_csr_cpu.cc:1047-1068): Dirichlet rows -> alpha on the diagonal
    void ZeroRows(const int *index_set, int size, T alpha);
    void Multicoloring(int &ncolors, std::vector<int> &color_sizes, std::vector<int> &perm) const;
    void Reorder(const int *perm);
    void ilu0();
    void VectorMult(const lVector<T> &, lVector<T> *) const override;
    void VectorMultAdd(const lVector<T> &, lVector<T> *) const override;
    void extract_invdiagelements(int, int, lVector<T> *) const override;
    std::vector<int> row, col;
    std::vector<T> val;
};

template <typename T>
class HIP_CSR_lMatrix : public lMatrix<T> {
  public:
    HIP_CSR_lMatrix() { this->platform_ = HIP; }
    ~HIP_CSR_lMatrix() override { Clear(); }
    HIP_CSR_lMatrix(const HIP_CSR_lMatrix &) = delete;
    HIP_CSR_lMatrix &operator=(const HIP_CSR_lMatrix &) = delete;
    void Init(int nnz, int nrow, int ncol, const std::string &name) override;
    void Clear() override;
    void CopyStructureFrom(const lMatrix<T> &mat) override;
    void CopyFrom(const lMatrix<T> &mat) override;
    lMatrix<T> *CloneWithoutContent() const override;
    // CastFrom (lmatrix.h:297-298)
    void CastFrom(const HIP_CSR_lMatrix<double> &other);
    void CastFrom(const HIP_CSR_lMatrix<float> &other);
    void VectorMult(const lVector<T> &in, lVector<T> *out) const override;
    void VectorMultAdd(const lVector<T> &in, lVector<T> *out) const override;
    void extract_invdiagelements(int start_i, int end_i, lVector<T> *vec) const override;
    hfh_csr *handle() const { return h_; }
    const std::vector<int> &host_row() const { return hrow_; }
    const std::vector<int> &host_col() const { return hcol_; }

  private:
    hfh_csr *h_ = nullptr;
    std::vector<int> hrow_, hcol_;  // pattern kept for host-side analysis
    bool structure_ = false;
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
