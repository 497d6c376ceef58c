// lmp_host.cc -- CPU containers and the HIP platform classes of lmp.h.
// HIP objects forward every operation to the C ABI of libhfhip.so.
#include "lmp.h"

#include <algorithm>
#include <cstdio>
#include <numeric>

#include "analysis.h"

namespace hflmp {

void check(int rc, const char *what) {
    if (rc != HFH_OK) throw Error(rc, std::string(what) + ": " + hfh_last_error());
}

static hfh_ctx *g_ctx = nullptr;
static int g_device = -1;

void init_platform(int device) {
    if (g_ctx && g_device == device) return;
    stop_platform();
    check(hfh_ctx_create(device, &g_ctx), "init_platform");
    g_device = device;
}

void stop_platform() {
    if (g_ctx) hfh_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    g_device = -1;
}

hfh_ctx *hip_context() {
    if (!g_ctx) init_platform(0);
    return g_ctx;
}

template <typename T>
static const HIP_lVector<T> &as_hip(const lVector<T> &v, const char *what) {
    const HIP_lVector<T> *h = dynamic_cast<const HIP_lVector<T> *>(&v);
    if (!h) throw Error(HFH_ERR_INVALID, std::string(what) + ": vector is not on the HIP platform");
    return *h;
}

template <typename T>
static const CPU_lVector<T> *as_cpu(const lVector<T> &v) {
    return dynamic_cast<const CPU_lVector<T> *>(&v);
}

// ---------------------------------------------------------------- CPU_lVector

template <typename T>
void CPU_lVector<T>::Init(int size, const std::string &name) {
    if (size < 0) throw Error(HFH_ERR_INVALID, "CPU_lVector::Init: negative size");
    buffer.assign(size, T(0));
    this->size_ = size;
    this->name_ = name;
}

template <typename T>
void CPU_lVector<T>::Zeros() { std::fill(buffer.begin(), buffer.end(), T(0)); }

template <typename T>
void CPU_lVector<T>::SetBlockValues(int s, int e, const T *v) {
    if (s < 0 || e > this->size_ || s > e) throw Error(HFH_ERR_INVALID, "CPU_lVector::SetBlockValues: bad range");
    std::copy(v, v + (e - s), buffer.begin() + s);
}

template <typename T>
void CPU_lVector<T>::GetBlockValues(int s, int e, T *v) const {
    if (s < 0 || e > this->size_ || s > e) throw Error(HFH_ERR_INVALID, "CPU_lVector::GetBlockValues: bad range");
    std::copy(buffer.begin() + s, buffer.begin() + e, v);
}

template <typename T>
void CPU_lVector<T>::SetValues(const int *idx, int n, const T *v) {
    for (int i = 0; i < n; ++i) {
        if (idx[i] < 0 || idx[i] >= this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::SetValues: index out of range");
        buffer[idx[i]] = v[i];
    }
}

template <typename T>
void CPU_lVector<T>::GetValues(const int *idx, int n, T *v) const {
    for (int i = 0; i < n; ++i) {
        if (idx[i] < 0 || idx[i] >= this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::GetValues: index out of range");
        v[i] = buffer[idx[i]];
    }
}

template <typename T>
void CPU_lVector<T>::add_value(int i, T v) {
    if (i < 0 || i >= this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::add_value: index out of range");
    buffer[i] += v;
}

template <typename T>
void CPU_lVector<T>::CopyFrom(const lVector<T> &vec) {
    if (vec.get_size() != this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::CopyFrom: size mismatch");
    if (const CPU_lVector<T> *c = as_cpu(vec)) {
        buffer = c->buffer;
        return;
    }
    vec.GetBlockValues(0, this->size_, buffer.data());  // device -> host
}

template <typename T>
void CPU_lVector<T>::CastFrom(const lVector<double> &vec) {
    if (vec.get_size() != this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::CastFrom: size mismatch");
    std::vector<double> tmp(this->size_);
    vec.GetBlockValues(0, this->size_, tmp.data());
    for (int i = 0; i < this->size_; ++i) buffer[i] = static_cast<T>(tmp[i]);
}

template <typename T>
void CPU_lVector<T>::CastFrom(const lVector<float> &vec) {
    if (vec.get_size() != this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::CastFrom: size mismatch");
    std::vector<float> tmp(this->size_);
    vec.GetBlockValues(0, this->size_, tmp.data());
    for (int i = 0; i < this->size_; ++i) buffer[i] = static_cast<T>(tmp[i]);
}

template <typename T>
lVector<T> *CPU_lVector<T>::CloneWithoutContent() const {
    return new CPU_lVector<T>(this->size_, this->name_);
}

template <typename T>
void CPU_lVector<T>::partial_replace_subvector(int s, int ss, int n, const lVector<T> &sub) {
    std::vector<T> tmp(n);
    sub.GetBlockValues(ss, ss + n, tmp.data());
    SetBlockValues(s, s + n, tmp.data());
}

template <typename T>
void CPU_lVector<T>::Reorder(const int *index) {
    std::vector<T> old(buffer);
    for (int i = 0; i < this->size_; ++i) buffer[i] = old[index[i]];
}

template <typename T>
void CPU_lVector<T>::no_compute(const char *op) {
    throw Error(HFH_ERR_UNSUPPORTED, std::string("CPU_lVector::") + op +
                                         ": arithmetic runs on the HIP platform (init_vector(..., HIP))");
}

// ---------------------------------------------------------------- HIP_lVector

template <typename T>
void HIP_lVector<T>::Init(int size, const std::string &name) {
    Clear();
    check(hfh_vec_create(hip_context(), size, dtype_of<T>::value, &h_), "HIP_lVector::Init");
    this->size_ = size;
    this->name_ = name;
}

template <typename T>
void HIP_lVector<T>::Clear() {
    if (h_) hfh_vec_destroy(h_);
    h_ = nullptr;
    this->size_ = 0;
}

template <typename T>
void HIP_lVector<T>::Zeros() { check(hfh_vec_zeros(h_), "HIP_lVector::Zeros"); }

template <typename T>
void HIP_lVector<T>::SetBlockValues(int s, int e, const T *v) {
    check(hfh_vec_upload(h_, v, s, e - s), "HIP_lVector::SetBlockValues");
}

template <typename T>
void HIP_lVector<T>::GetBlockValues(int s, int e, T *v) const {
    check(hfh_vec_download(h_, v, s, e - s), "HIP_lVector::GetBlockValues");
}

template <typename T>
void HIP_lVector<T>::SetValues(const int *idx, int n, const T *v) {
    check(hfh_vec_set_values(h_, idx, n, v), "HIP_lVector::SetValues");
}

template <typename T>
void HIP_lVector<T>::GetValues(const int *idx, int n, T *v) const {
    check(hfh_vec_get_values(h_, idx, n, v), "HIP_lVector::GetValues");
}

template <typename T>
void HIP_lVector<T>::CopyFrom(const lVector<T> &vec) {
    if (const CPU_lVector<T> *c = as_cpu(vec)) {  // host -> device
        if (c->get_size() != this->size_) throw Error(HFH_ERR_INVALID, "HIP_lVector::CopyFrom: size mismatch");
        SetBlockValues(0, this->size_, c->buffer.data());
        return;
    }
    check(hfh_vec_copy(h_, as_hip(vec, "HIP_lVector::CopyFrom").h_), "HIP_lVector::CopyFrom");
}

template <typename T>
void HIP_lVector<T>::CastFrom(const lVector<double> &vec) {
    check(hfh_vec_cast(h_, as_hip(vec, "HIP_lVector::CastFrom").handle()), "HIP_lVector::CastFrom");
}

template <typename T>
void HIP_lVector<T>::CastFrom(const lVector<float> &vec) {
    check(hfh_vec_cast(h_, as_hip(vec, "HIP_lVector::CastFrom").handle()), "HIP_lVector::CastFrom");
}

template <typename T>
lVector<T> *HIP_lVector<T>::CloneWithoutContent() const {
    return new HIP_lVector<T>(this->size_, this->name_);
}

template <typename T>
void HIP_lVector<T>::partial_replace_subvector(int s, int ss, int n, const lVector<T> &sub) {
    check(hfh_vec_copy_range(h_, s, as_hip(sub, "partial_replace_subvector").h_, ss, n),
          "HIP_lVector::partial_replace_subvector");
}

template <typename T>
void HIP_lVector<T>::Reorder(const int *index) {
    HIP_lVector<T> tmp(this->size_, this->name_);
    tmp.CopyFrom(*this);
    check(hfh_vec_reorder(h_, tmp.h_, index), "HIP_lVector::Reorder");
}

template <typename T>
void HIP_lVector<T>::Sync() const { check(hfh_ctx_sync(hip_context()), "HIP_lVector::Sync"); }

template <typename T>
T HIP_lVector<T>::Dot(const lVector<T> &vec) const {
    double d = 0;
    check(hfh_dot(h_, as_hip(vec, "HIP_lVector::Dot").h_, &d), "HIP_lVector::Dot");
    return static_cast<T>(d);
}

template <typename T>
T HIP_lVector<T>::Norm2() const {
    double d = 0;
    check(hfh_nrm2(h_, &d), "HIP_lVector::Norm2");
    return static_cast<T>(d);
}

template <typename T>
void HIP_lVector<T>::Axpy(const lVector<T> &vec, T scalar) {
    check(hfh_axpy(h_, as_hip(vec, "HIP_lVector::Axpy").h_, scalar), "HIP_lVector::Axpy");
}

template <typename T>
void HIP_lVector<T>::ScaleAdd(T scalar, const lVector<T> &vec) {
    check(hfh_scaleadd(h_, scalar, as_hip(vec, "HIP_lVector::ScaleAdd").h_), "HIP_lVector::ScaleAdd");
}

template <typename T>
void HIP_lVector<T>::Scale(T scalar) { check(hfh_scale(h_, scalar), "HIP_lVector::Scale"); }

template <typename T>
void HIP_lVector<T>::ElementWiseMult(const lVector<T> &vec) {
    check(hfh_ewmult(h_, as_hip(vec, "HIP_lVector::ElementWiseMult").h_), "HIP_lVector::ElementWiseMult");
}

// ---------------------------------------------------------------- CPU_CSR_lMatrix

template <typename T>
void CPU_CSR_lMatrix<T>::Init(int nnz, int nrow, int ncol, const std::string &name) {
    if (nnz < 0 || nrow < 0 || ncol < 0) throw Error(HFH_ERR_INVALID, "CPU_CSR_lMatrix::Init: negative size");
    this->nnz_ = nnz;
    this->nrow_ = nrow;
    this->ncol_ = ncol;
    this->name_ = name;
    row.assign(nrow + 1, 0);
    col.assign(nnz, 0);
    val.assign(nnz, T(0));
}

template <typename T>
void CPU_CSR_lMatrix<T>::Clear() {
    row.clear();
    col.clear();
    val.clear();
    this->nnz_ = this->nrow_ = this->ncol_ = 0;
}

template <typename T>
void CPU_CSR_lMatrix<T>::CopyStructureFrom(const lMatrix<T> &mat) {
    const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat);
    if (!m) throw Error(HFH_ERR_UNSUPPORTED, "CPU_CSR_lMatrix::CopyStructureFrom: source must be a CPU CSR matrix");
    row = m->row;
    col = m->col;
    val.assign(col.size(), T(0));
}

template <typename T>
void CPU_CSR_lMatrix<T>::CopyFrom(const lMatrix<T> &mat) {
    const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat);
    if (!m) throw Error(HFH_ERR_UNSUPPORTED, "CPU_CSR_lMatrix::CopyFrom: source must be a CPU CSR matrix");
    if (m->col.size() != col.size()) throw Error(HFH_ERR_INVALID, "CPU_CSR_lMatrix::CopyFrom: structure mismatch");
    val = m->val;
}

template <typename T>
lMatrix<T> *CPU_CSR_lMatrix<T>::CloneWithoutContent() const {
    CPU_CSR_lMatrix<T> *m = new CPU_CSR_lMatrix<T>();
    m->Init(this->nnz_, this->nrow_, this->ncol_, this->name_);
    return m;
}

template <typename T>
void CPU_CSR_lMatrix<T>::TransformFromCOO(const int *rows, const int *cols, const T *data, int nrow, int ncol, int nnz) {
    Init(nnz, nrow, ncol, this->name_);
    // counting sort by row (stable), then columns ascending inside each row
    for (int i = 0; i < nnz; ++i) {
        if (rows[i] < 0 || rows[i] >= nrow || cols[i] < 0 || cols[i] >= ncol)
            throw Error(HFH_ERR_INVALID, "TransformFromCOO: entry out of range");
        ++row[rows[i] + 1];
    }
    for (int r = 0; r < nrow; ++r) row[r + 1] += row[r];
    std::vector<int> fill(row.begin(), row.end() - 1);
    for (int i = 0; i < nnz; ++i) {
        const int p = fill[rows[i]]++;
        col[p] = cols[i];
        val[p] = data ? data[i] : T(0);
    }
    std::vector<int> idx;
    for (int r = 0; r < nrow; ++r) {
        const int a = row[r], b = row[r + 1];
        idx.resize(b - a);
        std::iota(idx.begin(), idx.end(), a);
        std::sort(idx.begin(), idx.end(), [&](int p, int q) { return col[p] < col[q]; });
        std::vector<int> c2(b - a);
        std::vector<T> v2(b - a);
        for (int k = 0; k < b - a; ++k) {
            c2[k] = col[idx[k]];
            v2[k] = val[idx[k]];
        }
        std::copy(c2.begin(), c2.end(), col.begin() + a);
        std::copy(v2.begin(), v2.end(), val.begin() + a);
    }
}

template <typename T>
void CPU_CSR_lMatrix<T>::init_structure(const int *rows, const int *cols) {
    TransformFromCOO(rows, cols, nullptr, this->nrow_, this->ncol_, this->nnz_);
}

template <typename T>
void CPU_CSR_lMatrix<T>::Zeros() { std::fill(val.begin(), val.end(), T(0)); }

template <typename T>
void CPU_CSR_lMatrix<T>::add_value(int r, int c, T v) {
    // entries outside the pattern are ignored, as in the reference
    for (int j = row[r]; j < row[r + 1]; ++j)
        if (col[j] == c) {
            val[j] += v;
            return;
        }
}

template <typename T>
void CPU_CSR_lMatrix<T>::add_values(const int *rows, int nr, const int *cols, int nc, const T *values) {
    for (int i = 0; i < nr; ++i) {
        const int r = rows[i];
        // both column lists ascending: one merge pass per row
        int k = 0;
        for (int j = row[r]; j < row[r + 1] && k < nc; ++j) {
            while (k < nc && cols[k] < col[j]) ++k;
            if (k < nc && cols[k] == col[j]) {
                val[j] += values[i * nc + k];
                ++k;
            }
        }
    }
}

template <typename T>
void CPU_CSR_lMatrix<T>::get_value(int r, int c, T *v) const {
    *v = T(0);
    for (int j = row[r]; j < row[r + 1]; ++j)
        if (col[j] == c) *v = val[j];
}

template <typename T>
void CPU_CSR_lMatrix<T>::ZeroRows(const int *index_set, int size, T alpha) {
    for (int i = 0; i < size; ++i) {
        const int r = index_set[i];
        for (int j = row[r]; j < row[r + 1]; ++j) val[j] = col[j] == r ? alpha : T(0);
    }
}

template <typename T>
void CPU_CSR_lMatrix<T>::Multicoloring(int &ncolors, std::vector<int> &sizes, std::vector<int> &perm) const {
    std::vector<int> color(this->nrow_);
    perm.assign(this->nrow_, 0);
    if (multicoloring(this->nrow_, row.data(), col.data(), &ncolors, color.data(), sizes, perm.data()))
        throw Error(HFH_ERR_INVALID, "Multicoloring: pattern without off-diagonal entries");
}

template <typename T>
void CPU_CSR_lMatrix<T>::Reorder(const int *perm) {
    std::vector<int> prow(row.size()), pcol(col.size());
    std::vector<double> v(val.begin(), val.end()), pv(val.size());
    if (csr_permute(this->nrow_, row.data(), col.data(), v.data(), perm, prow.data(), pcol.data(), pv.data()))
        throw Error(HFH_ERR_INVALID, "Reorder: invalid permutation");
    row.swap(prow);
    col.swap(pcol);
    for (size_t j = 0; j < val.size(); ++j) val[j] = static_cast<T>(pv[j]);
}

template <typename T>
void CPU_CSR_lMatrix<T>::ilu0() {
    std::vector<double> v(val.begin(), val.end());
    const int miss = hflmp::ilu0(this->nrow_, row.data(), col.data(), v.data());
    if (miss) throw Error(HFH_ERR_ZERO_PIVOT, "ilu0: vanishing diagonal elements");
    for (size_t j = 0; j < val.size(); ++j) val[j] = static_cast<T>(v[j]);
}

template <typename T>
void CPU_CSR_lMatrix<T>::VectorMult(const lVector<T> &, lVector<T> *) const {
    throw Error(HFH_ERR_UNSUPPORTED, "CPU_CSR_lMatrix::VectorMult: arithmetic runs on the HIP platform");
}

template <typename T>
void CPU_CSR_lMatrix<T>::VectorMultAdd(const lVector<T> &, lVector<T> *) const {
    throw Error(HFH_ERR_UNSUPPORTED, "CPU_CSR_lMatrix::VectorMultAdd: arithmetic runs on the HIP platform");
}

template <typename T>
void CPU_CSR_lMatrix<T>::extract_invdiagelements(int s, int e, lVector<T> *vec) const {
    std::vector<T> d(e - s, T(0));
    for (int i = s; i < e; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j)
            if (col[j] == i) d[i - s] = T(1) / val[j];
    vec->SetBlockValues(0, e - s, d.data());
}

// ---------------------------------------------------------------- HIP_CSR_lMatrix

template <typename T>
void HIP_CSR_lMatrix<T>::Init(int nnz, int nrow, int ncol, const std::string &name) {
    Clear();
    check(hfh_csr_create(hip_context(), nrow, ncol, nnz, dtype_of<T>::value, &h_), "HIP_CSR_lMatrix::Init");
    this->nnz_ = nnz;
    this->nrow_ = nrow;
    this->ncol_ = ncol;
    this->name_ = name;
}

template <typename T>
void HIP_CSR_lMatrix<T>::Clear() {
    if (h_) hfh_csr_destroy(h_);
    h_ = nullptr;
    structure_ = false;
    hrow_.clear();
    hcol_.clear();
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyStructureFrom(const lMatrix<T> &mat) {
    if (const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat)) {
        if (!h_ || m->get_nnz() != this->nnz_ || m->get_num_row() != this->nrow_)
            Init(m->get_nnz(), m->get_num_row(), m->get_num_col(), this->name_);
        check(hfh_csr_upload(h_, m->row.data(), m->col.data(), m->val.data()), "HIP_CSR_lMatrix::CopyStructureFrom");
        hrow_ = m->row;
        hcol_ = m->col;
        structure_ = true;
        return;
    }
    const HIP_CSR_lMatrix<T> *d = dynamic_cast<const HIP_CSR_lMatrix<T> *>(&mat);
    if (!d) throw Error(HFH_ERR_UNSUPPORTED, "HIP_CSR_lMatrix::CopyStructureFrom: unsupported source");
    std::vector<T> v(d->nnz_);
    check(hfh_csr_download(d->h_, nullptr, nullptr, v.data()), "HIP_CSR_lMatrix::CopyStructureFrom");
    if (!h_ || d->nnz_ != this->nnz_) Init(d->nnz_, d->nrow_, d->ncol_, this->name_);
    check(hfh_csr_upload(h_, d->hrow_.data(), d->hcol_.data(), v.data()), "HIP_CSR_lMatrix::CopyStructureFrom");
    hrow_ = d->hrow_;
    hcol_ = d->hcol_;
    structure_ = true;
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyFrom(const lMatrix<T> &mat) {
    if (!structure_) CopyStructureFrom(mat);
    if (const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat)) {
        check(hfh_csr_upload_values(h_, m->val.data()), "HIP_CSR_lMatrix::CopyFrom");
        return;
    }
    const HIP_CSR_lMatrix<T> *d = dynamic_cast<const HIP_CSR_lMatrix<T> *>(&mat);
    if (!d) throw Error(HFH_ERR_UNSUPPORTED, "HIP_CSR_lMatrix::CopyFrom: unsupported source");
    check(hfh_csr_cast(h_, d->h_), "HIP_CSR_lMatrix::CopyFrom");
}

template <typename T>
lMatrix<T> *HIP_CSR_lMatrix<T>::CloneWithoutContent() const {
    HIP_CSR_lMatrix<T> *m = new HIP_CSR_lMatrix<T>();
    m->Init(this->nnz_, this->nrow_, this->ncol_, this->name_);
    return m;
}

template <typename T>
void HIP_CSR_lMatrix<T>::CastFrom(const HIP_CSR_lMatrix<double> &o) {
    if (!h_) Init(o.get_nnz(), o.get_num_row(), o.get_num_col(), this->name_);
    check(hfh_csr_cast(h_, o.h_), "HIP_CSR_lMatrix::CastFrom");
    hrow_ = o.hrow_;
    hcol_ = o.hcol_;
    structure_ = true;
}

template <typename T>
void HIP_CSR_lMatrix<T>::CastFrom(const HIP_CSR_lMatrix<float> &o) {
    if (!h_) Init(o.get_nnz(), o.get_num_row(), o.get_num_col(), this->name_);
    check(hfh_csr_cast(h_, o.h_), "HIP_CSR_lMatrix::CastFrom");
    hrow_ = o.hrow_;
    hcol_ = o.hcol_;
    structure_ = true;
}

template <typename T>
void HIP_CSR_lMatrix<T>::VectorMult(const lVector<T> &in, lVector<T> *out) const {
    check(hfh_spmv(h_, as_hip(in, "VectorMult").handle(), as_hip(*out, "VectorMult").handle()),
          "HIP_CSR_lMatrix::VectorMult");
}

template <typename T>
void HIP_CSR_lMatrix<T>::VectorMultAdd(const lVector<T> &in, lVector<T> *out) const {
    check(hfh_spmv_add(h_, as_hip(in, "VectorMultAdd").handle(), as_hip(*out, "VectorMultAdd").handle()),
          "HIP_CSR_lMatrix::VectorMultAdd");
}

template <typename T>
void HIP_CSR_lMatrix<T>::extract_invdiagelements(int s, int e, lVector<T> *vec) const {
    if (s != 0 || e != this->nrow_)
        throw Error(HFH_ERR_UNSUPPORTED, "extract_invdiagelements: only the full row range is provided on HIP");
    check(hfh_csr_invdiag(h_, as_hip(*vec, "extract_invdiagelements").handle()), "extract_invdiagelements");
}

// ---------------------------------------------------------------- factories

template <typename T>
lVector<T> *init_vector(int size, const std::string &name, PLATFORM platform, IMPLEMENTATION) {
    if (platform == CPU) return new CPU_lVector<T>(size, name);
    if (platform == HIP) return new HIP_lVector<T>(size, name);
    throw Error(HFH_ERR_UNSUPPORTED, "init_vector: platform not provided (CPU or HIP)");
}

template <typename T>
lMatrix<T> *init_matrix(int nnz, int nrow, int ncol, const std::string &name, PLATFORM platform, IMPLEMENTATION,
                        MATRIX_FORMAT format) {
    if (format != CSR) throw Error(HFH_ERR_UNSUPPORTED, "init_matrix: only CSR is provided");
    lMatrix<T> *m = nullptr;
    if (platform == CPU) m = new CPU_CSR_lMatrix<T>();
    else if (platform == HIP) m = new HIP_CSR_lMatrix<T>();
    else throw Error(HFH_ERR_UNSUPPORTED, "init_matrix: platform not provided (CPU or HIP)");
    m->Init(nnz, nrow, ncol, name);
    return m;
}

// ---------------------------------------------------------------- preconditioners

template <typename T>
void lPreconditioner_Device<T>::Clear() {
    if (p_) hfh_prec_destroy(p_);
    p_ = nullptr;
}

template <typename T>
const HIP_CSR_lMatrix<T> &lPreconditioner_Device<T>::hip_op() const {
    const HIP_CSR_lMatrix<T> *m = dynamic_cast<const HIP_CSR_lMatrix<T> *>(this->op_);
    if (!m) throw Error(HFH_ERR_INVALID, this->name_ + ": operator must be a HIP CSR matrix");
    return *m;
}

template <typename T>
void lPreconditioner_Device<T>::ApplylPreconditioner(const lVector<T> &input, lVector<T> *output) {
    if (!this->state_) this->Build();
    check(hfh_prec_apply(p_, as_hip(input, "ApplylPreconditioner").handle(),
                         as_hip(*output, "ApplylPreconditioner").handle()),
          "ApplylPreconditioner");
}

template <typename T>
void lPreconditioner_Jacobi<T>::Build() {
    this->Clear();
    check(hfh_jacobi_create(this->hip_op().handle(), &this->p_), "lPreconditioner_Jacobi::Build");
    this->state_ = true;
}

template <typename T>
static void levels_of(const HIP_CSR_lMatrix<T> &A, std::vector<int> &lp, std::vector<int> &lr, std::vector<int> &up,
                      std::vector<int> &ur) {
    const int n = A.get_num_row();
    lr.assign(n, 0);
    ur.assign(n, 0);
    levels_lower(n, A.host_row().data(), A.host_col().data(), lp, lr.data());
    levels_upper(n, A.host_row().data(), A.host_col().data(), up, ur.data());
}

template <typename T>
void lPreconditioner_SymmetricGaussSeidel<T>::Build() {
    this->Clear();
    const HIP_CSR_lMatrix<T> &A = this->hip_op();
    std::vector<int> lp, lr, up, ur;
    levels_of(A, lp, lr, up, ur);
    check(hfh_sgs_create(A.handle(), static_cast<int>(lp.size()) - 1, lp.data(), lr.data(),
                         static_cast<int>(up.size()) - 1, up.data(), ur.data(), &this->p_),
          "lPreconditioner_SymmetricGaussSeidel::Build");
    this->state_ = true;
}

template <typename T>
void lPreconditioner_ILUp<T>::Build() {
    this->Clear();
    const HIP_CSR_lMatrix<T> &A = this->hip_op();
    // factorise on the host (lpreconditioner.cc:388-411), solve on the device
    CPU_CSR_lMatrix<T> cpu;
    cpu.Init(A.get_nnz(), A.get_num_row(), A.get_num_col(), "LU");
    cpu.row = A.host_row();
    cpu.col = A.host_col();
    check(hfh_csr_download(A.handle(), nullptr, nullptr, cpu.val.data()), "lPreconditioner_ILUp::Build");
    cpu.ilu0();
    LU_.reset(new HIP_CSR_lMatrix<T>());
    LU_->Init(cpu.get_nnz(), cpu.get_num_row(), cpu.get_num_col(), "LU");
    LU_->CopyStructureFrom(cpu);
    std::vector<int> lp, lr, up, ur;
    levels_of(*LU_, lp, lr, up, ur);
    check(hfh_ilu0_create(LU_->handle(), static_cast<int>(lp.size()) - 1, lp.data(), lr.data(),
                          static_cast<int>(up.size()) - 1, up.data(), ur.data(), &this->p_),
          "lPreconditioner_ILUp::Build");
    this->state_ = true;
}

template <typename T>
void lPreconditioner_MultiColoring_SymmetricGaussSeidel<T>::Build() {
    this->Clear();
    const HIP_CSR_lMatrix<T> &A = this->hip_op();
    if (sizes_.empty()) throw Error(HFH_ERR_INVALID, "MC-SGS: colour sizes not set (Analyse/Preprocess first)");
    check(hfh_mcsgs_create(A.handle(), static_cast<int>(sizes_.size()), sizes_.data(), omega_, &this->p_),
          "lPreconditioner_MultiColoring_SymmetricGaussSeidel::Build");
    this->state_ = true;
}

#define HFLMP_INSTANTIATE(T)                                                                                \
    template class CPU_lVector<T>;                                                                          \
    template class HIP_lVector<T>;                                                                          \
    template class CPU_CSR_lMatrix<T>;                                                                      \
    template class HIP_CSR_lMatrix<T>;                                                                      \
    template class lPreconditioner_Device<T>;                                                               \
    template class lPreconditioner_Jacobi<T>;                                                               \
    template class lPreconditioner_SymmetricGaussSeidel<T>;                                                 \
    template class lPreconditioner_ILUp<T>;                                                                 \
    template class lPreconditioner_MultiColoring_SymmetricGaussSeidel<T>;                                   \
    template lVector<T> *init_vector<T>(int, const std::string &, PLATFORM, IMPLEMENTATION);                \
    template lMatrix<T> *init_matrix<T>(int, int, int, const std::string &, PLATFORM, IMPLEMENTATION,       \
                                        MATRIX_FORMAT);
HFLMP_INSTANTIATE(double)
HFLMP_INSTANTIATE(float)

}  // namespace hflmp
