// coupled.h -- single-rank CoupledVector / CoupledMatrix and the LA
// descriptors the solver templates are instantiated with
// (src/linear_algebra/{coupled_vector,coupled_matrix,la_descriptor}.h).
// With one rank the ghost part is empty, Update() exchanges nothing and the
// scalar reductions need no allreduce.
#pragma once

#include <cmath>
#include <memory>

#include "lmp.h"

namespace hflmp {

template <typename T>
class CoupledVector {
  public:
    typedef T DataType;
    CoupledVector() {}
    // CoupledVector::Init (coupled_vector.cc:70-120): interior on the platform
    void Init(int size, PLATFORM platform = HIP, IMPLEMENTATION impl = GFX950) {
        interior_.reset(init_vector<T>(size, "interior", platform, impl));
    }
    // CloneFromWithoutContent (coupled_vector.cc:856-896)
    void CloneFromWithoutContent(const CoupledVector<T> &v) { interior_.reset(v.interior().CloneWithoutContent()); }
    void CloneFrom(const CoupledVector<T> &v) {
        CloneFromWithoutContent(v);
        CopyFrom(v);
    }
    void CopyFrom(const CoupledVector<T> &v) { interior_->CopyFrom(v.interior()); }
    void Zeros() { interior_->Zeros(); }
    void Update() {}
    // Dot / Norm2 (coupled_vector.cc:330-344, 432-436): local dot, allreduce over one rank
    T Dot(const CoupledVector<T> &v) const { return interior_->Dot(v.interior()); }
    T Norm2() const { return std::sqrt(Dot(*this)); }
    // Axpy (coupled_vector.cc:364-371): this += s v
    void Axpy(const CoupledVector<T> &v, T s) { interior_->Axpy(v.interior(), s); }
    // ScaleAdd (coupled_vector.cc:391-398): this = s this + v
    void ScaleAdd(const CoupledVector<T> &v, T s) { interior_->ScaleAdd(s, v.interior()); }
    void Scale(T s) { interior_->Scale(s); }
    // CastInteriorFrom (coupled_vector.cc:755-778)
    template <typename U>
    void CastInteriorFrom(const CoupledVector<U> &v) { interior_->CastFrom(v.interior()); }
    lVector<T> &interior() { return *interior_; }
    const lVector<T> &interior() const { return *interior_; }
    int size_local() const { return interior_ ? interior_->get_size() : 0; }

  private:
    std::unique_ptr<lVector<T>> interior_;
};

template <typename T>
class CoupledMatrix {
  public:
    typedef T DataType;
    // Init + InitStructure + CopyFrom of an assembled host matrix
    // (coupled_matrix.cc:254-339; CopyStructureFrom/CopyFrom to the device)
    void InitFrom(const CPU_CSR_lMatrix<T> &host, PLATFORM platform = HIP) {
        diagonal_.reset(init_matrix<T>(host.get_nnz(), host.get_num_row(), host.get_num_col(), "diagonal", platform));
        diagonal_->CopyStructureFrom(host);
        diagonal_->CopyFrom(host);
    }
    // CastFrom of a matrix in the other precision (MPIR's low-precision operator)
    template <typename U>
    void CastFrom(const CoupledMatrix<U> &other) {
        HIP_CSR_lMatrix<T> *m = new HIP_CSR_lMatrix<T>();
        m->CastFrom(dynamic_cast<const HIP_CSR_lMatrix<U> &>(other.diagonal()));
        diagonal_.reset(m);
    }
    // CoupledMatrix::Zeros (coupled_matrix.cc:540-547)
    void Zeros() { diagonal_->Zeros(); }
    // CoupledMatrix::diagonalize_rows (coupled_matrix.cc:550-575): the listed
    // (global = local on one rank) rows zeroed, diagonal_value on the diagonal;
    // the off-diagonal block is empty on one rank
    void diagonalize_rows(const int *row_indices, int num_rows, T diagonal_value) {
        diagonal_->ZeroRows(row_indices, num_rows, diagonal_value);
    }
    // CoupledMatrix::VectorMult (coupled_matrix.cc:433-468): diagonal block only on one rank
    void VectorMult(const CoupledVector<T> &in, CoupledVector<T> *out) const {
        diagonal_->VectorMult(in.interior(), &out->interior());
    }
    lMatrix<T> &diagonal() { return *diagonal_; }
    const lMatrix<T> &diagonal() const { return *diagonal_; }
    int num_rows_local() const { return diagonal_->get_num_row(); }

  private:
    std::unique_ptr<lMatrix<T>> diagonal_;
};

// la_descriptor.h:51-67 with HIP platform objects
struct LADescriptorHipD {
    typedef CoupledMatrix<double> MatrixType;
    typedef CoupledVector<double> VectorType;
    typedef double DataType;
};

struct LADescriptorHipS {
    typedef CoupledMatrix<float> MatrixType;
    typedef CoupledVector<float> VectorType;
    typedef float DataType;
};

}  // namespace hflmp
