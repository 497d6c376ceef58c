// lmp_host.cc -- the lVector / lMatrix base behaviour, the CPU containers and
// the HIP platform classes of lmp.h.  HIP objects forward every operation to
// the C ABI of libhfhip.so.
#include "lmp.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

#include "../../include/hflmp.h"
#include "analysis.h"

namespace hflmp {

void check(int rc, const char *what) {
    if (rc != HFH_OK) throw Error(rc, std::string(what) + ": " + hfh_last_error());
}

void unsupported(const std::string &cls, const char *op) {
    throw Error(HFH_ERR_UNSUPPORTED, cls + ": " + op + " is not supported by this platform / format");
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
static HIP_lVector<T> &as_hip(lVector<T> &v, const char *what) {
    HIP_lVector<T> *h = dynamic_cast<HIP_lVector<T> *>(&v);
    if (!h) throw Error(HFH_ERR_INVALID, std::string(what) + ": vector is not on the HIP platform");
    return *h;
}

static void check_range(int start, int end, int size, const char *what) {
    if (start < 0 || end > size || start > end) throw Error(HFH_ERR_INVALID, std::string(what) + ": bad range");
}

// ---------------------------------------------------------------- Rotg / Rotmg

template <typename T>
void rotg(T *sa, T *sb, T *sc, T *ss) {
    const T a = *sa, b = *sb;
    const T aa = std::fabs(a), ab = std::fabs(b);
    const T roe = aa > ab ? a : b;
    const T scale = aa + ab;
    if (scale == T(0)) {
        *sc = T(1);
        *ss = *sa = *sb = T(0);
        return;
    }
    const T t0 = aa / scale, t1 = ab / scale;
    T r = scale * std::sqrt(t0 * t0 + t1 * t1);
    if (roe < T(0)) r = -r;
    *sc = a / r;
    *ss = b / r;
    T z;
    if (aa > ab) z = *ss;
    else if (*sc != T(0)) z = T(1) / *sc;
    else z = T(1);
    *sa = r;
    *sb = z;
}

// Modified Givens rotation (Hanson & Krogh; the reference BLAS xROTMG):
// finds H with flag -1 (full), 0 (unit diagonal), 1 (unit off-diagonal
// pattern h12 = 1, h21 = -1) or -2 (identity) such that the second
// component of H (x1, y1) is zero, with the weights d1, d2 rescaled into
// [1/gam^2, gam^2], gam = 4096.
template <typename T>
void rotmg(T *sd1, T *sd2, T *sx1, T sy1, T *param) {
    const T gam = T(4096), gamsq = gam * gam, rgamsq = T(1) / gamsq;
    T flag, h11 = 0, h12 = 0, h21 = 0, h22 = 0;
    T d1 = *sd1, d2 = *sd2, x1 = *sx1;
    auto zero_all = [&]() {
        flag = T(-1);
        h11 = h12 = h21 = h22 = T(0);
        d1 = d2 = x1 = T(0);
    };
    if (d1 < T(0)) {
        zero_all();
    } else {
        const T p2 = d2 * sy1;
        if (p2 == T(0)) {
            param[0] = T(-2);
            return;
        }
        const T p1 = d1 * x1, q2 = p2 * sy1, q1 = p1 * x1;
        if (std::fabs(q1) > std::fabs(q2)) {
            h21 = -sy1 / x1;
            h12 = p2 / p1;
            const T u = T(1) - h12 * h21;
            if (u > T(0)) {
                flag = T(0);
                d1 /= u;
                d2 /= u;
                x1 *= u;
            } else {
                zero_all();
            }
        } else if (q2 < T(0)) {
            zero_all();
        } else {
            flag = T(1);
            h11 = p1 / p2;
            h22 = x1 / sy1;
            const T u = T(1) + h11 * h22;
            const T t = d2 / u;
            d2 = d1 / u;
            d1 = t;
            x1 = sy1 * u;
        }
        // rescaling works on the full form of H
        auto full = [&]() {
            if (flag == T(0)) {
                h11 = h22 = T(1);
            } else if (flag == T(1)) {
                h21 = T(-1);
                h12 = T(1);
            }
            flag = T(-1);
        };
        if (d1 != T(0)) {
            while (d1 <= rgamsq || d1 >= gamsq) {
                full();
                if (d1 <= rgamsq) {
                    d1 *= gamsq;
                    x1 /= gam;
                    h11 /= gam;
                    h12 /= gam;
                } else {
                    d1 /= gamsq;
                    x1 *= gam;
                    h11 *= gam;
                    h12 *= gam;
                }
            }
        }
        if (d2 != T(0)) {
            while (std::fabs(d2) <= rgamsq || std::fabs(d2) >= gamsq) {
                full();
                if (std::fabs(d2) <= rgamsq) {
                    d2 *= gamsq;
                    h21 /= gam;
                    h22 /= gam;
                } else {
                    d2 /= gamsq;
                    h21 *= gam;
                    h22 *= gam;
                }
            }
        }
    }
    if (flag < T(0)) {
        param[1] = h11;
        param[2] = h21;
        param[3] = h12;
        param[4] = h22;
    } else if (flag == T(0)) {
        param[2] = h21;
        param[3] = h12;
    } else {
        param[1] = h11;
        param[4] = h22;
    }
    param[0] = flag;
    *sd1 = d1;
    *sd2 = d2;
    *sx1 = x1;
}

// ---------------------------------------------------------------- lVector (lvector.cc)

template <typename T>
void lVector<T>::print(std::ostream &out) const {
    out << "lVector name='" << name_ << "', Elements=" << size_ << ", Precision=" << sizeof(T) * 8
        << "bit, Platform:" << platform_name_ << ", Implementation:" << implementation_name_ << std::endl;
}

template <typename T>
void lVector<T>::CopyFromIndexset(const lVector<T> &vec) {
    std::vector<int> idx(vec.get_indexset_size());
    vec.get_indexset(idx.data());
    set_indexset(idx.data(), static_cast<int>(idx.size()));
}

template <typename T>
void lVector<T>::ReadFile(const char *) { unsupported("lVector '" + name_ + "'", "ReadFile"); }

template <typename T>
void lVector<T>::WriteFile(const char *) { unsupported("lVector '" + name_ + "'", "WriteFile"); }

template <typename T>
void lVector<T>::CopyStructureFrom(const lVector<T> &vec2) {
    if (this != &vec2) Init(vec2.get_size(), vec2.get_name());
}

template <typename T>
void lVector<T>::CloneFrom(const lVector<T> &other) {
    if (this == &other) return;
    Clear();
    CopyStructureFrom(other);
    CopyFrom(other);
}

template <typename T>
lVector<T> *lVector<T>::CloneWithoutContent() const {
    lVector<T> *v = init_vector<T>(size_, "clone from " + name_, platform_, impl_);
    v->Zeros();
    return v;
}

// ---------------------------------------------------------------- CPU_lVector

template <typename T>
CPU_lVector<T>::CPU_lVector() {
    this->platform_ = CPU;
    this->impl_ = NAIVE;
    this->platform_name_ = "CPU (x86)";
    this->implementation_name_ = "host container";
}

template <typename T>
CPU_lVector<T> &CPU_lVector<T>::operator=(const CPU_lVector &v) {
    if (this == &v) return *this;
    buffer = v.buffer;
    indexset_ = v.indexset_;
    this->size_ = v.size_;
    this->name_ = v.name_;
    return *this;
}

template <typename T>
void CPU_lVector<T>::Init(const int size, const std::string name) {
    if (size < 0) throw Error(HFH_ERR_INVALID, "CPU_lVector::Init: negative size");
    Clear();
    buffer.assign(size, T(0));
    this->size_ = size;
    this->name_ = name;
}

template <typename T>
void CPU_lVector<T>::Clear() {
    buffer.clear();
    indexset_.clear();
    this->size_ = 0;
    this->name_ = "";
}

template <typename T>
void CPU_lVector<T>::Zeros() { std::fill(buffer.begin(), buffer.end(), T(0)); }

template <typename T>
void CPU_lVector<T>::SetBlockValues(const int s, const int e, const T *v) {
    check_range(s, e, this->size_, "CPU_lVector::SetBlockValues");
    std::copy(v, v + (e - s), buffer.begin() + s);
}

template <typename T>
void CPU_lVector<T>::GetBlockValues(const int s, const int e, T *v) const {
    check_range(s, e, this->size_, "CPU_lVector::GetBlockValues");
    std::copy(buffer.begin() + s, buffer.begin() + e, v);
}

template <typename T>
void CPU_lVector<T>::SetValues(const int *idx, const int n, const T *v) {
    for (int i = 0; i < n; ++i) {
        if (idx[i] < 0 || idx[i] >= this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::SetValues: index out of range");
        buffer[idx[i]] = v[i];
    }
}

template <typename T>
void CPU_lVector<T>::GetValues(const int *idx, const int n, T *v) const {
    for (int i = 0; i < n; ++i) {
        if (idx[i] < 0 || idx[i] >= this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::GetValues: index out of range");
        v[i] = buffer[idx[i]];
    }
}

template <typename T>
void CPU_lVector<T>::add_value(const int i, T v) {
    if (i < 0 || i >= this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::add_value: index out of range");
    buffer[i] += v;
}

template <typename T>
void CPU_lVector<T>::add_values(const int *idx, int n, const T *v) {
    for (int i = 0; i < n; ++i) add_value(idx[i], v[i]);
}

template <typename T>
void CPU_lVector<T>::set_indexset(const int *indexset, const int size) {
    if (size < 0) throw Error(HFH_ERR_INVALID, "CPU_lVector::set_indexset: negative size");
    for (int i = 0; i < size; ++i)
        if (indexset[i] < 0 || indexset[i] >= this->size_)
            throw Error(HFH_ERR_INVALID, "CPU_lVector::set_indexset: index out of range");
    indexset_.assign(indexset, indexset + size);
}

template <typename T>
void CPU_lVector<T>::get_indexset(int *indexset) const {
    if (!indexset_.empty()) std::copy(indexset_.begin(), indexset_.end(), indexset);
}

template <typename T>
void CPU_lVector<T>::GetIndexedValues(T *values) const {
    if (!indexset_.empty() && this->size_ > 0) GetValues(indexset_.data(), get_indexset_size(), values);
}

template <typename T>
void CPU_lVector<T>::SetIndexedValues(const T *values) {
    if (!indexset_.empty() && this->size_ > 0 && values) SetValues(indexset_.data(), get_indexset_size(), values);
}

template <typename T>
lVector<T> *CPU_lVector<T>::extract_subvector(const int start_i, const int end_i) const {
    if (start_i < 0 || start_i >= end_i || end_i > this->size_)
        throw Error(HFH_ERR_INVALID, "CPU_lVector::extract_subvector: bad range");
    CPU_lVector<T> *sub = new CPU_lVector<T>(end_i - start_i, "sub vector from " + this->name_);
    GetBlockValues(start_i, end_i, sub->buffer.data());
    return sub;
}

template <typename T>
void CPU_lVector<T>::partial_replace_subvector(const int s, const int ss, const int n, const lVector<T> &sub) {
    check_range(s, s + n, this->size_, "CPU_lVector::partial_replace_subvector");
    sub.GetBlockValues(ss, ss + n, buffer.data() + s);
}

template <typename T>
void CPU_lVector<T>::partial_add_subvector(const int s, const int ss, const int n, const T w, const lVector<T> &sub) {
    check_range(s, s + n, this->size_, "CPU_lVector::partial_add_subvector");
    std::vector<T> tmp(n);
    sub.GetBlockValues(ss, ss + n, tmp.data());
    for (int i = 0; i < n; ++i) buffer[s + i] += w * tmp[i];
}

template <typename T>
void CPU_lVector<T>::ReadFile(const char *filename) {
    int n = 0;
    if (hfl_vec_read(filename, &n, nullptr)) throw Error(HFH_ERR_INVALID, std::string("ReadFile: ") + hfl_last_error());
    std::vector<double> v(n);
    if (hfl_vec_read(filename, &n, v.data())) throw Error(HFH_ERR_INVALID, std::string("ReadFile: ") + hfl_last_error());
    Init(n, filename);
    for (int i = 0; i < n; ++i) buffer[i] = static_cast<T>(v[i]);
}

template <typename T>
void CPU_lVector<T>::WriteFile(const char *filename) {
    std::vector<double> v(buffer.begin(), buffer.end());
    if (hfl_vec_write(filename, this->size_, v.data()))
        throw Error(HFH_ERR_INVALID, std::string("WriteFile: ") + hfl_last_error());
}

template <typename T>
lVector<T> &CPU_lVector<T>::operator=(const lVector<T> &vec2) {
    if (this != &vec2) CopyFrom(vec2);
    return *this;
}

template <typename T>
void CPU_lVector<T>::CopyFrom(const lVector<T> &vec) {
    if (this == &vec) return;
    if (const CPU_lVector<T> *c = dynamic_cast<const CPU_lVector<T> *>(&vec)) {
        if (c->get_size() != this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::CopyFrom: size mismatch");
        buffer = c->buffer;
        return;
    }
    vec.CopyTo(*this);  // the other platform writes into the host buffer
}

template <typename T>
void CPU_lVector<T>::CastFrom(const lVector<double> &vec) {
    if (vec.get_size() != this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::CastFrom: size mismatch");
    if (const CPU_lVector<double> *c = dynamic_cast<const CPU_lVector<double> *>(&vec)) {
        for (int i = 0; i < this->size_; ++i) buffer[i] = static_cast<T>(c->buffer[i]);
        return;
    }
    vec.CastTo(*this);
}

template <typename T>
void CPU_lVector<T>::CastFrom(const lVector<float> &vec) {
    if (vec.get_size() != this->size_) throw Error(HFH_ERR_INVALID, "CPU_lVector::CastFrom: size mismatch");
    if (const CPU_lVector<float> *c = dynamic_cast<const CPU_lVector<float> *>(&vec)) {
        for (int i = 0; i < this->size_; ++i) buffer[i] = static_cast<T>(c->buffer[i]);
        return;
    }
    vec.CastTo(*this);
}

template <typename T>
lVector<T> *CPU_lVector<T>::CloneWithoutContent() const {
    return new CPU_lVector<T>(this->size_, "clone from " + this->name_);
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
HIP_lVector<T>::HIP_lVector() {
    this->platform_ = HIP;
    this->impl_ = GFX950;
    this->platform_name_ = "HIP (MI355X)";
    this->implementation_name_ = "gfx950 kernels";
}

template <typename T>
void HIP_lVector<T>::Init(const int size, const std::string name) {
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
    this->name_ = "";
}

template <typename T>
void HIP_lVector<T>::Zeros() { check(hfh_vec_zeros(h_), "HIP_lVector::Zeros"); }

template <typename T>
void HIP_lVector<T>::SetBlockValues(const int s, const int e, const T *v) {
    check(hfh_vec_upload(h_, v, s, e - s), "HIP_lVector::SetBlockValues");
}

template <typename T>
void HIP_lVector<T>::GetBlockValues(const int s, const int e, T *v) const {
    check(hfh_vec_download(h_, v, s, e - s), "HIP_lVector::GetBlockValues");
}

template <typename T>
void HIP_lVector<T>::SetValues(const int *idx, const int n, const T *v) {
    check(hfh_vec_set_values(h_, idx, n, v), "HIP_lVector::SetValues");
}

template <typename T>
void HIP_lVector<T>::GetValues(const int *idx, const int n, T *v) const {
    check(hfh_vec_get_values(h_, idx, n, v), "HIP_lVector::GetValues");
}

template <typename T>
void HIP_lVector<T>::add_value(const int i, T v) {
    check(hfh_vec_add_values(h_, &i, 1, &v), "HIP_lVector::add_value");
}

template <typename T>
void HIP_lVector<T>::add_values(const int *idx, int n, const T *v) {
    check(hfh_vec_add_values(h_, idx, n, v), "HIP_lVector::add_values");
}

template <typename T>
void HIP_lVector<T>::set_indexset(const int *indexset, const int size) {
    check(hfh_vec_set_indexset(h_, indexset, size), "HIP_lVector::set_indexset");
}

template <typename T>
void HIP_lVector<T>::get_indexset(int *indexset) const {
    int n = 0;
    check(hfh_vec_get_indexset(h_, nullptr, &n), "HIP_lVector::get_indexset");
    if (n > 0) check(hfh_vec_get_indexset(h_, indexset, &n), "HIP_lVector::get_indexset");
}

template <typename T>
int HIP_lVector<T>::get_indexset_size() const {
    int n = 0;
    check(hfh_vec_get_indexset(h_, nullptr, &n), "HIP_lVector::get_indexset_size");
    return n;
}

template <typename T>
void HIP_lVector<T>::GetIndexedValues(T *values) const {
    check(hfh_vec_get_indexed_values(h_, values), "HIP_lVector::GetIndexedValues");
}

template <typename T>
void HIP_lVector<T>::SetIndexedValues(const T *values) {
    check(hfh_vec_set_indexed_values(h_, values), "HIP_lVector::SetIndexedValues");
}

template <typename T>
lVector<T> *HIP_lVector<T>::extract_subvector(const int start_i, const int end_i) const {
    if (start_i < 0 || start_i >= end_i || end_i > this->size_)
        throw Error(HFH_ERR_INVALID, "HIP_lVector::extract_subvector: bad range");
    HIP_lVector<T> *sub = new HIP_lVector<T>(end_i - start_i, "sub vector from " + this->name_);
    check(hfh_vec_copy_range(sub->h_, 0, h_, start_i, end_i - start_i), "HIP_lVector::extract_subvector");
    return sub;
}

template <typename T>
void HIP_lVector<T>::partial_replace_subvector(const int s, const int ss, const int n, const lVector<T> &sub) {
    if (const CPU_lVector<T> *c = dynamic_cast<const CPU_lVector<T> *>(&sub)) {
        check_range(ss, ss + n, c->get_size(), "HIP_lVector::partial_replace_subvector");
        SetBlockValues(s, s + n, c->buffer.data() + ss);
        return;
    }
    check(hfh_vec_copy_range(h_, s, as_hip(sub, "partial_replace_subvector").h_, ss, n),
          "HIP_lVector::partial_replace_subvector");
}

template <typename T>
void HIP_lVector<T>::partial_add_subvector(const int s, const int ss, const int n, const T w, const lVector<T> &sub) {
    check(hfh_vec_add_range(h_, s, as_hip(sub, "partial_add_subvector").h_, ss, n, w),
          "HIP_lVector::partial_add_subvector");
}

template <typename T>
void HIP_lVector<T>::ReadFile(const char *filename) {
    CPU_lVector<T> host;
    host.ReadFile(filename);
    Init(host.get_size(), filename);
    CopyFrom(host);
}

template <typename T>
void HIP_lVector<T>::WriteFile(const char *filename) {
    CPU_lVector<T> host(this->size_, this->name_);
    CopyTo(host);
    host.WriteFile(filename);
}

template <typename T>
void HIP_lVector<T>::Reorder(const int *index) {
    HIP_lVector<T> tmp(this->size_, this->name_);
    tmp.CopyFrom(*this);
    check(hfh_vec_reorder(h_, tmp.h_, index), "HIP_lVector::Reorder");
}

template <typename T>
lVector<T> &HIP_lVector<T>::operator=(const lVector<T> &vec2) {
    if (this != &vec2) CopyFrom(vec2);
    return *this;
}

template <typename T>
void HIP_lVector<T>::CopyFrom(const lVector<T> &vec) {
    if (this == &vec) return;
    if (const CPU_lVector<T> *c = dynamic_cast<const CPU_lVector<T> *>(&vec)) {  // host -> device
        if (c->get_size() != this->size_) throw Error(HFH_ERR_INVALID, "HIP_lVector::CopyFrom: size mismatch");
        SetBlockValues(0, this->size_, c->buffer.data());
        return;
    }
    check(hfh_vec_copy(h_, as_hip(vec, "HIP_lVector::CopyFrom").h_), "HIP_lVector::CopyFrom");
}

template <typename T>
void HIP_lVector<T>::CopyTo(lVector<T> &vec) const {
    if (CPU_lVector<T> *c = dynamic_cast<CPU_lVector<T> *>(&vec)) {  // device -> host
        if (c->get_size() != this->size_) throw Error(HFH_ERR_INVALID, "HIP_lVector::CopyTo: size mismatch");
        GetBlockValues(0, this->size_, c->buffer.data());
        return;
    }
    if (dynamic_cast<HIP_lVector<T> *>(&vec)) {
        vec.CopyFrom(*this);
        return;
    }
    unsupported("HIP_lVector", "CopyTo (destination platform)");
}

template <typename T, typename U>
static void cast_from(HIP_lVector<T> &self, const lVector<U> &vec) {
    if (const CPU_lVector<U> *c = dynamic_cast<const CPU_lVector<U> *>(&vec)) {
        if (c->get_size() != self.get_size()) throw Error(HFH_ERR_INVALID, "HIP_lVector::CastFrom: size mismatch");
        std::vector<T> tmp(c->buffer.begin(), c->buffer.end());
        self.SetBlockValues(0, self.get_size(), tmp.data());
        return;
    }
    if (const HIP_lVector<U> *h = dynamic_cast<const HIP_lVector<U> *>(&vec)) {
        check(hfh_vec_cast(self.handle(), h->handle()), "HIP_lVector::CastFrom");
        return;
    }
    vec.CastTo(self);
}

template <typename T, typename U>
static void cast_to(const HIP_lVector<T> &self, lVector<U> &vec) {
    if (CPU_lVector<U> *c = dynamic_cast<CPU_lVector<U> *>(&vec)) {
        if (c->get_size() != self.get_size()) throw Error(HFH_ERR_INVALID, "HIP_lVector::CastTo: size mismatch");
        std::vector<T> tmp(self.get_size());
        self.GetBlockValues(0, self.get_size(), tmp.data());
        for (int i = 0; i < self.get_size(); ++i) c->buffer[i] = static_cast<U>(tmp[i]);
        return;
    }
    if (HIP_lVector<U> *h = dynamic_cast<HIP_lVector<U> *>(&vec)) {
        check(hfh_vec_cast(h->handle(), self.handle()), "HIP_lVector::CastTo");
        return;
    }
    unsupported("HIP_lVector", "CastTo (destination platform)");
}

template <typename T>
void HIP_lVector<T>::CastFrom(const lVector<double> &vec) { cast_from<T, double>(*this, vec); }

template <typename T>
void HIP_lVector<T>::CastFrom(const lVector<float> &vec) { cast_from<T, float>(*this, vec); }

template <typename T>
void HIP_lVector<T>::CastTo(lVector<double> &vec) const { cast_to<T, double>(*this, vec); }

template <typename T>
void HIP_lVector<T>::CastTo(lVector<float> &vec) const { cast_to<T, float>(*this, vec); }

template <typename T>
lVector<T> *HIP_lVector<T>::CloneWithoutContent() const {
    return new HIP_lVector<T>(this->size_, "clone from " + this->name_);
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
T HIP_lVector<T>::Norm1() const {
    double d = 0;
    check(hfh_norm1(h_, &d), "HIP_lVector::Norm1");
    return static_cast<T>(d);
}

template <typename T>
T HIP_lVector<T>::Norm2() const {
    double d = 0;
    check(hfh_nrm2(h_, &d), "HIP_lVector::Norm2");
    return static_cast<T>(d);
}

template <typename T>
T HIP_lVector<T>::NormMax() const {
    double d = 0;
    check(hfh_normmax(h_, &d), "HIP_lVector::NormMax");
    return static_cast<T>(d);
}

template <typename T>
int HIP_lVector<T>::ArgMin() const {
    int i = 0;
    check(hfh_iamin(h_, &i), "HIP_lVector::ArgMin");
    return i;
}

template <typename T>
int HIP_lVector<T>::ArgMax() const {
    int i = 0;
    check(hfh_iamax(h_, &i), "HIP_lVector::ArgMax");
    return i;
}

template <typename T>
void HIP_lVector<T>::Axpy(const lVector<T> &vec, const T scalar) {
    check(hfh_axpy(h_, as_hip(vec, "HIP_lVector::Axpy").h_, scalar), "HIP_lVector::Axpy");
}

template <typename T>
void HIP_lVector<T>::ScaleAdd(const T scalar, const lVector<T> &vec) {
    check(hfh_scaleadd(h_, scalar, as_hip(vec, "HIP_lVector::ScaleAdd").h_), "HIP_lVector::ScaleAdd");
}

template <typename T>
void HIP_lVector<T>::Scale(const T scalar) { check(hfh_scale(h_, scalar), "HIP_lVector::Scale"); }

template <typename T>
void HIP_lVector<T>::ElementWiseMult(const lVector<T> &vec) const {
    check(hfh_ewmult(h_, as_hip(vec, "HIP_lVector::ElementWiseMult").h_), "HIP_lVector::ElementWiseMult");
}

template <typename T>
void HIP_lVector<T>::Rot(lVector<T> *vec, const T &sc, const T &ss) {
    check(hfh_rot(h_, as_hip(*vec, "HIP_lVector::Rot").h_, sc, ss), "HIP_lVector::Rot");
}

template <typename T>
void HIP_lVector<T>::Rotm(lVector<T> *vec, const T &sparam) {
    // sparam is the first of five: flag, h11, h21, h12, h22 (CPUsimple_lvector.cc:540-600)
    const T *p = &sparam;
    const double q[5] = {p[0], p[1], p[2], p[3], p[4]};
    check(hfh_rotm(h_, as_hip(*vec, "HIP_lVector::Rotm").h_, q), "HIP_lVector::Rotm");
}

// ---------------------------------------------------------------- lMatrix (lmatrix.cc)

template <typename T>
std::string lMatrix<T>::get_matrix_format_name() const {
    switch (format_) {
    case DENSE: return "DENSE";
    case CSR: return "CSR";
    case COO: return "COO";
    default: return "ELL";
    }
}

template <typename T>
void lMatrix<T>::print(std::ostream &out) const {
    out << "lMatrix name='" << name_ << "' Matrix format=" << get_matrix_format_name() << " nnz=" << nnz_
        << ", num_row=" << nrow_ << ", num_col=" << ncol_ << ", Precision=" << sizeof(T) * 8
        << "bit, Platform:" << platform_name_ << ", Implementation:" << implementation_name_ << std::endl;
}

template <typename T>
void lMatrix<T>::get_as_coo(std::vector<T> &, std::vector<int> &, std::vector<int> &) const { no("get_as_coo"); }
template <typename T>
void lMatrix<T>::ZeroRows(const int *, const int, const T) { no("ZeroRows"); }
template <typename T>
void lMatrix<T>::ZeroCols(const int *, const int, const T) { no("ZeroCols"); }
template <typename T>
void lMatrix<T>::Reorder(const int *) { no("Reorder"); }
template <typename T>
void lMatrix<T>::Multicoloring(int &, int **, int **) const { no("Multicoloring"); }
template <typename T>
void lMatrix<T>::Levelscheduling(int &, int **, int **) const { no("Levelscheduling"); }
template <typename T>
void lMatrix<T>::ilu0() { no("ilu0"); }
template <typename T>
void lMatrix<T>::ilup(const int) { no("ilup"); }
template <typename T>
void lMatrix<T>::ilup(const lMatrix<T> &, const int) { no("ilup"); }
template <typename T>
void lMatrix<T>::ilu_solve(const lVector<T> &, lVector<T> *) const { no("ilu_solve"); }
template <typename T>
void lMatrix<T>::Scale(const T) { no("Scale"); }
template <typename T>
void lMatrix<T>::ScaleOffdiag(const T) { no("ScaleOffdiag"); }
template <typename T>
void lMatrix<T>::init_structure(const int *, const int *) { no("init_structure"); }
template <typename T>
void lMatrix<T>::add_value(const int, const int, const T) { no("add_value"); }
template <typename T>
void lMatrix<T>::add_values(const int *, int, const int *, int, const T *) { no("add_values"); }
template <typename T>
void lMatrix<T>::get_value(const int, const int, T *) const { no("get_value"); }
template <typename T>
void lMatrix<T>::get_add_values(const int *, int, const int *, int, const int *, int, T *) const {
    no("get_add_values");
}
template <typename T>
void lMatrix<T>::VectorMultAdd_submatrix(const int *, int, const int *, int, const int *, const T *, T *) const {
    no("VectorMultAdd_submatrix");
}
template <typename T>
void lMatrix<T>::VectorMultAdd_submatrix_vanka(const int *, int, const lVector<T> &, T *) const {
    no("VectorMultAdd_submatrix_vanka");
}
template <typename T>
void lMatrix<T>::transpose_me() { no("transpose_me"); }
template <typename T>
bool lMatrix<T>::issymmetric() { no("issymmetric"); }
template <typename T>
void lMatrix<T>::compress_me() { no("compress_me"); }
template <typename T>
void lMatrix<T>::delete_diagonal() { no("delete_diagonal"); }
template <typename T>
void lMatrix<T>::delete_offdiagonal() { no("delete_offdiagonal"); }
template <typename T>
void lMatrix<T>::delete_lower_triangular() { no("delete_lower_triangular"); }
template <typename T>
void lMatrix<T>::delete_strictly_lower_triangular() { no("delete_strictly_lower_triangular"); }
template <typename T>
void lMatrix<T>::delete_upper_triangular() { no("delete_upper_triangular"); }
template <typename T>
void lMatrix<T>::delete_strictly_upper_triangular() { no("delete_strictly_upper_triangular"); }
template <typename T>
void lMatrix<T>::extract_diagelements(const int, const int, lVector<T> *) const { no("extract_diagelements"); }
template <typename T>
void lMatrix<T>::extract_invdiagelements(const int, const int, lVector<T> *) const { no("extract_invdiagelements"); }
template <typename T>
void lMatrix<T>::ReadFile(const char *) { no("ReadFile"); }
template <typename T>
void lMatrix<T>::WriteFile(const char *) const { no("WriteFile"); }
template <typename T>
lMatrix<T> *lMatrix<T>::MatrixMult(const lMatrix<T> &) const { no("MatrixMult"); }
template <typename T>
lMatrix<T> *lMatrix<T>::MatrixMultSupStructure(const lMatrix<T> &) const { no("MatrixMultSupStructure"); }
template <typename T>
lMatrix<T> *lMatrix<T>::MatrixSupSPower(const int) const { no("MatrixSupSPower"); }
template <typename T>
void lMatrix<T>::MatrixAdd(const lMatrix<T> &) { no("MatrixAdd"); }
template <typename T>
void lMatrix<T>::GershgorinSpectrum(T *, T *) const { no("GershgorinSpectrum"); }
template <typename T>
void lMatrix<T>::Pjacobi(const lVector<T> &, lVector<T> *) const { no("Pjacobi"); }
template <typename T>
void lMatrix<T>::Pgauss_seidel(const lVector<T> &, lVector<T> *) const { no("Pgauss_seidel"); }
template <typename T>
void lMatrix<T>::Psgauss_seidel(const lVector<T> &, lVector<T> *) const { no("Psgauss_seidel"); }
template <typename T>
void lMatrix<T>::BlockPsgauss_seidel(const lVector<T> &, lVector<T> *, const int, const int) const {
    no("BlockPsgauss_seidel");
}
template <typename T>
void lMatrix<T>::BlocksPsgauss_seidel(const lVector<T> &, lVector<T> *, const int) const {
    no("BlocksPsgauss_seidel");
}
template <typename T>
void lMatrix<T>::Psor(const T, const lVector<T> &, lVector<T> *) const { no("Psor"); }
template <typename T>
void lMatrix<T>::Pssor(const T, const lVector<T> &, lVector<T> *) const { no("Pssor"); }

template <typename T>
void lMatrix<T>::Residual(const lVector<T> &b, const lVector<T> &x, lVector<T> *res) const {
    VectorMult(x, res);
    res->ScaleAdd(T(-1), b);
}

template <typename T>
void lMatrix<T>::CloneFrom(const lMatrix<T> &mat2) {
    if (this == &mat2) return;
    Init(mat2.get_nnz(), mat2.get_num_row(), mat2.get_num_col(), mat2.get_name());
    CopyStructureFrom(mat2);
    CopyFrom(mat2);
}

template <typename T>
lMatrix<T> *lMatrix<T>::CloneWithoutContent() const {
    lMatrix<T> *m = init_matrix<T>(nnz_, nrow_, ncol_, "clone from " + name_, platform_, impl_, format_);
    m->Zeros();
    return m;
}

// lMatrix::extract_submatrix (lmatrix.cc:315-352): through a host CSR copy
template <typename T>
lMatrix<T> *lMatrix<T>::extract_submatrix(const int start_row, const int start_col, const int end_row,
                                          const int end_col) const {
    CPU_CSR_lMatrix<T> host;
    host.CloneFrom(*this);
    std::unique_ptr<lMatrix<T>> sub(host.extract_submatrix(start_row, start_col, end_row, end_col));
    lMatrix<T> *out = init_matrix<T>(sub->get_nnz(), sub->get_num_row(), sub->get_num_col(),
                                     "sub matrix from " + name_, platform_, impl_, format_);
    out->CloneFrom(*sub);
    return out;
}

// ---------------------------------------------------------------- CPU_CSR_lMatrix

template <typename T>
CPU_CSR_lMatrix<T>::CPU_CSR_lMatrix() {
    this->platform_ = CPU;
    this->impl_ = NAIVE;
    this->format_ = CSR;
    this->platform_name_ = "CPU (x86)";
    this->implementation_name_ = "host container";
}

template <typename T>
CPU_CSR_lMatrix<T> &CPU_CSR_lMatrix<T>::operator=(const CPU_CSR_lMatrix &m) {
    if (this == &m) return *this;
    row = m.row;
    col = m.col;
    val = m.val;
    this->nnz_ = m.nnz_;
    this->nrow_ = m.nrow_;
    this->ncol_ = m.ncol_;
    this->name_ = m.name_;
    return *this;
}

template <typename T>
void CPU_CSR_lMatrix<T>::Init(const int nnz, const int nrow, const int ncol, const std::string name) {
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
    this->name_ = "";
}

template <typename T>
void CPU_CSR_lMatrix<T>::get_as_coo(std::vector<T> &vals, std::vector<int> &cols, std::vector<int> &rows) const {
    vals = val;
    cols = col;
    rows.assign(this->nnz_, 0);
    for (int i = 0; i < this->nrow_; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j) rows[j] = i;
}

template <typename T>
lMatrix<T> &CPU_CSR_lMatrix<T>::operator=(const lMatrix<T> &mat2) {
    if (this != &mat2) CopyFrom(mat2);
    return *this;
}

template <typename T>
void CPU_CSR_lMatrix<T>::CopyStructureFrom(const lMatrix<T> &mat) {
    if (this == &mat) return;
    if (const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat)) {
        Init(m->get_nnz(), m->get_num_row(), m->get_num_col(), this->name_);
        row = m->row;
        col = m->col;
        return;
    }
    mat.CopyStructureTo(*this);
}

template <typename T>
void CPU_CSR_lMatrix<T>::CopyFrom(const lMatrix<T> &mat) {
    if (this == &mat) return;
    if (const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat)) {
        if (m->col.size() != col.size()) throw Error(HFH_ERR_INVALID, "CPU_CSR_lMatrix::CopyFrom: structure mismatch");
        val = m->val;
        return;
    }
    mat.CopyTo(*this);
}

template <typename T, typename U>
static void cpu_cast_from(CPU_CSR_lMatrix<T> &self, const lMatrix<U> &other) {
    if (const CPU_CSR_lMatrix<U> *m = dynamic_cast<const CPU_CSR_lMatrix<U> *>(&other)) {
        self.Init(m->get_nnz(), m->get_num_row(), m->get_num_col(), self.get_name());
        self.row = m->row;
        self.col = m->col;
        for (size_t j = 0; j < m->val.size(); ++j) self.val[j] = static_cast<T>(m->val[j]);
        return;
    }
    other.CastTo(self);
}

template <typename T>
void CPU_CSR_lMatrix<T>::CastFrom(const lMatrix<double> &other) { cpu_cast_from<T, double>(*this, other); }

template <typename T>
void CPU_CSR_lMatrix<T>::CastFrom(const lMatrix<float> &other) { cpu_cast_from<T, float>(*this, other); }

template <typename T>
void CPU_CSR_lMatrix<T>::ConvertFrom(const lMatrix<T> &mat2) {
    if (this == &mat2) return;
    std::vector<T> v;
    std::vector<int> c, r;
    mat2.get_as_coo(v, c, r);
    Init(mat2.get_nnz(), mat2.get_num_row(), mat2.get_num_col(), mat2.get_name());
    TransformFromCOO(r.data(), c.data(), v.data(), this->nrow_, this->ncol_, this->nnz_);
}

template <typename T>
lMatrix<T> *CPU_CSR_lMatrix<T>::CloneWithoutContent() const {
    CPU_CSR_lMatrix<T> *m = new CPU_CSR_lMatrix<T>();
    m->Init(this->nnz_, this->nrow_, this->ncol_, "clone from " + this->name_);
    return m;
}

template <typename T>
void CPU_CSR_lMatrix<T>::TransformFromCOO(const int *rows, const int *cols, const T *data, int nrow, int ncol,
                                          int nnz) {
    if (nrow != this->nrow_ || ncol != this->ncol_ || nnz != this->nnz_) Init(nnz, nrow, ncol, this->name_);
    std::fill(row.begin(), row.end(), 0);
    for (int i = 0; i < nnz; ++i) {
        if (rows[i] < 0 || rows[i] >= nrow || cols[i] < 0 || cols[i] >= ncol)
            throw Error(HFH_ERR_INVALID, "TransformFromCOO: entry out of range");
        ++row[rows[i] + 1];
    }
    for (int r = 0; r < nrow; ++r) row[r + 1] += row[r];
    // scatter in input order, then each row ordered by a selection sort whose
    // swaps are those of the reference (so repeated columns land as there)
    std::vector<int> fill(row.begin(), row.end() - 1);
    for (int i = 0; i < nnz; ++i) {
        const int p = fill[rows[i]]++;
        col[p] = cols[i];
        val[p] = data ? data[i] : T(0);
    }
    for (int r = 0; r < nrow; ++r)
        for (int i = row[r]; i < row[r + 1]; ++i) {
            int m = i;
            for (int j = i + 1; j < row[r + 1]; ++j)
                if (col[j] < col[m]) m = j;
            std::swap(col[i], col[m]);
            std::swap(val[i], val[m]);
        }
}

template <typename T>
void CPU_CSR_lMatrix<T>::init_structure(const int *rows, const int *cols) {
    TransformFromCOO(rows, cols, nullptr, this->nrow_, this->ncol_, this->nnz_);
}

template <typename T>
void CPU_CSR_lMatrix<T>::Zeros() { std::fill(val.begin(), val.end(), T(0)); }

template <typename T>
void CPU_CSR_lMatrix<T>::add_value(const int r, const int c, const T v) {
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
        // one merge pass of the row's columns against the ascending cols
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
void CPU_CSR_lMatrix<T>::get_value(const int r, const int c, T *v) const {
    // *v is left as it is when (r, c) is not in the pattern, as in the reference
    for (int j = row[r]; j < row[r + 1]; ++j)
        if (col[j] == c) {
            *v = val[j];
            return;
        }
}

template <typename T>
void CPU_CSR_lMatrix<T>::ZeroRows(const int *index_set, const int size, const T alpha) {
    for (int i = 0; i < size; ++i) {
        const int r = index_set[i];
        if (r < 0 || r >= this->nrow_) throw Error(HFH_ERR_INVALID, "ZeroRows: row out of range");
        for (int j = row[r]; j < row[r + 1]; ++j) val[j] = col[j] == r ? alpha : T(0);
    }
}

template <typename T>
void CPU_CSR_lMatrix<T>::ZeroCols(const int *index_set, const int size, const T alpha) {
    std::vector<char> listed(this->ncol_, 0);
    for (int i = 0; i < size; ++i) {
        if (index_set[i] < 0 || index_set[i] >= this->ncol_) throw Error(HFH_ERR_INVALID, "ZeroCols: column out of range");
        listed[index_set[i]] = 1;
    }
    for (int k = 0; k < this->nrow_; ++k)
        for (int j = row[k]; j < row[k + 1]; ++j)
            if (listed[col[j]]) val[j] = col[j] == k ? alpha : T(0);
}

template <typename T>
void CPU_CSR_lMatrix<T>::Scale(const T alpha) {
    for (auto &v : val) v *= alpha;
}

template <typename T>
void CPU_CSR_lMatrix<T>::ScaleOffdiag(const T alpha) {
    for (int i = 0; i < this->nrow_; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j)
            if (col[j] != i) val[j] *= alpha;
}

template <typename T>
void CPU_CSR_lMatrix<T>::Multicoloring(int &ncolors, std::vector<int> &sizes, std::vector<int> &perm) const {
    std::vector<int> color(this->nrow_);
    perm.assign(this->nrow_, 0);
    if (multicoloring(this->nrow_, row.data(), col.data(), &ncolors, color.data(), sizes, perm.data()))
        throw Error(HFH_ERR_INVALID, "Multicoloring: pattern without off-diagonal entries");
}

template <typename T>
void CPU_CSR_lMatrix<T>::Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const {
    std::vector<int> sizes, perm;
    Multicoloring(ncolors, sizes, perm);
    *color_sizes = new int[ncolors];
    *permut_index = new int[this->nrow_];
    std::copy(sizes.begin(), sizes.begin() + ncolors, *color_sizes);
    std::copy(perm.begin(), perm.end(), *permut_index);
}

template <typename T>
void CPU_CSR_lMatrix<T>::Levelscheduling(int &nlevels, int **level_sizes, int **permut_index) const {
    std::vector<int> sizes(std::max(this->nrow_, 1)), perm(std::max(this->nrow_, 1));
    levelscheduling(this->nrow_, row.data(), col.data(), &nlevels, sizes.data(), perm.data());
    *level_sizes = new int[nlevels];
    *permut_index = new int[this->nrow_];
    std::copy(sizes.begin(), sizes.begin() + nlevels, *level_sizes);
    std::copy(perm.begin(), perm.begin() + this->nrow_, *permut_index);
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
void CPU_CSR_lMatrix<T>::transpose_me() {
    std::vector<T> v;
    std::vector<int> c, r;
    get_as_coo(v, c, r);
    const int nr = this->nrow_, nc = this->ncol_, nnz = this->nnz_;
    Init(nnz, nc, nr, this->name_);
    TransformFromCOO(c.data(), r.data(), v.data(), nc, nr, nnz);
}

template <typename T>
void CPU_CSR_lMatrix<T>::ilu0() {
    std::vector<double> v(val.begin(), val.end());
    const int miss = hflmp::ilu0(this->nrow_, row.data(), col.data(), v.data());
    if (miss) throw Error(HFH_ERR_ZERO_PIVOT, "ilu0: vanishing diagonal elements");
    for (size_t j = 0; j < val.size(); ++j) val[j] = static_cast<T>(v[j]);
}

// CPU_CSR_lMatrix::extract_submatrix (lmatrix_csr_cpu.cc:1768-1844): the
// entries of rows [start_row, end_row) with columns in [start_col, end_col),
// renumbered from 0, in a matrix of the same platform
template <typename T>
lMatrix<T> *CPU_CSR_lMatrix<T>::extract_submatrix(const int start_row, const int start_col, const int end_row,
                                                  const int end_col) const {
    if (start_row < 0 || start_col < 0 || end_row > this->nrow_ || end_col > this->ncol_ || end_row <= start_row ||
        end_col <= start_col)
        throw Error(HFH_ERR_INVALID, "extract_submatrix: bad range");
    std::vector<int> r, c;
    std::vector<T> v;
    for (int i = start_row; i < end_row; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j)
            if (col[j] >= start_col && col[j] < end_col) {
                r.push_back(i - start_row);
                c.push_back(col[j] - start_col);
                v.push_back(val[j]);
            }
    CPU_CSR_lMatrix<T> *sub = new CPU_CSR_lMatrix<T>();
    sub->Init(static_cast<int>(r.size()), end_row - start_row, end_col - start_col, "sub matrix from " + this->name_);
    if (!r.empty()) {
        sub->init_structure(r.data(), c.data());
        for (size_t k = 0; k < r.size(); ++k) sub->add_value(r[k], c[k], v[k]);
    }
    return sub;
}

template <typename T>
void CPU_CSR_lMatrix<T>::extract_diagelements(const int s, const int e, lVector<T> *vec) const {
    check_range(s, e, this->nrow_, "extract_diagelements");
    std::vector<T> d(e - s, T(0));
    for (int i = s; i < e; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j)
            if (col[j] == i) d[i - s] = val[j];
    vec->SetBlockValues(0, e - s, d.data());
}

template <typename T>
void CPU_CSR_lMatrix<T>::extract_invdiagelements(const int s, const int e, lVector<T> *vec) const {
    check_range(s, e, this->nrow_, "extract_invdiagelements");
    std::vector<T> d(e - s, T(0));
    for (int i = s; i < e; ++i)
        for (int j = row[i]; j < row[i + 1]; ++j)
            if (col[j] == i) d[i - s] = T(1) / val[j];
    vec->SetBlockValues(0, e - s, d.data());
}

template <typename T>
void CPU_CSR_lMatrix<T>::ReadFile(const char *filename) {
    int nr = 0, nc = 0;
    long long nnz = 0;
    if (hfl_mm_read(filename, &nr, &nc, &nnz, nullptr, nullptr, nullptr))
        throw Error(HFH_ERR_INVALID, std::string("ReadFile: ") + hfl_last_error());
    std::vector<int> r(nr + 1), c(nnz);
    std::vector<double> v(nnz);
    if (hfl_mm_read(filename, &nr, &nc, &nnz, r.data(), c.data(), v.data()))
        throw Error(HFH_ERR_INVALID, std::string("ReadFile: ") + hfl_last_error());
    Init(static_cast<int>(nnz), nr, nc, filename);
    row = r;
    col = c;
    for (long long j = 0; j < nnz; ++j) val[j] = static_cast<T>(v[j]);
}

template <typename T>
void CPU_CSR_lMatrix<T>::WriteFile(const char *filename) const {
    std::vector<double> v(val.begin(), val.end());
    if (hfl_mm_write(filename, this->nrow_, this->ncol_, row.data(), col.data(), v.data()))
        throw Error(HFH_ERR_INVALID, std::string("WriteFile: ") + hfl_last_error());
}

template <typename T>
void CPU_CSR_lMatrix<T>::no_compute(const char *op) {
    throw Error(HFH_ERR_UNSUPPORTED, std::string("CPU_CSR_lMatrix::") + op + ": arithmetic runs on the HIP platform");
}

// ---------------------------------------------------------------- HIP_CSR_lMatrix

template <typename T>
HIP_CSR_lMatrix<T>::HIP_CSR_lMatrix() {
    this->platform_ = HIP;
    this->impl_ = GFX950;
    this->format_ = CSR;
    this->platform_name_ = "HIP (MI355X)";
    this->implementation_name_ = "gfx950 kernels";
}

template <typename T>
void HIP_CSR_lMatrix<T>::drop_applies() const {
    if (ilu_) hfh_prec_destroy(ilu_);
    if (sgs_) hfh_prec_destroy(sgs_);
    ilu_ = sgs_ = nullptr;
}

template <typename T>
void HIP_CSR_lMatrix<T>::Init(const int nnz, const int nrow, const int ncol, const std::string name) {
    Clear();
    check(hfh_csr_create(hip_context(), nrow, ncol, nnz, dtype_of<T>::value, &h_), "HIP_CSR_lMatrix::Init");
    this->nnz_ = nnz;
    this->nrow_ = nrow;
    this->ncol_ = ncol;
    this->name_ = name;
}

template <typename T>
void HIP_CSR_lMatrix<T>::Clear() {
    drop_applies();
    if (h_) hfh_csr_destroy(h_);
    h_ = nullptr;
    structure_ = false;
    hrow_.clear();
    hcol_.clear();
    this->nnz_ = this->nrow_ = this->ncol_ = 0;
}

// takes over a device matrix holding this pattern (the previous one is destroyed)
template <typename T>
void HIP_CSR_lMatrix<T>::adopt(hfh_csr *h, const std::vector<int> &row, const std::vector<int> &col) {
    drop_applies();
    if (h_ && h_ != h) hfh_csr_destroy(h_);
    h_ = h;
    hrow_ = row;
    hcol_ = col;
    structure_ = true;
}

template <typename T>
void HIP_CSR_lMatrix<T>::Zeros() {
    if (structure_) check(hfh_csr_zeros(h_), "HIP_CSR_lMatrix::Zeros");
}

template <typename T>
void HIP_CSR_lMatrix<T>::get_as_coo(std::vector<T> &vals, std::vector<int> &cols, std::vector<int> &rows) const {
    if (!structure_) throw Error(HFH_ERR_INVALID, "get_as_coo: matrix has no structure");
    vals.resize(this->nnz_);
    check(hfh_csr_download(h_, nullptr, nullptr, vals.data()), "HIP_CSR_lMatrix::get_as_coo");
    cols = hcol_;
    rows.assign(this->nnz_, 0);
    for (int i = 0; i < this->nrow_; ++i)
        for (int j = hrow_[i]; j < hrow_[i + 1]; ++j) rows[j] = i;
}

template <typename T>
lMatrix<T> &HIP_CSR_lMatrix<T>::operator=(const lMatrix<T> &mat2) {
    if (this != &mat2) CopyFrom(mat2);
    return *this;
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyStructureFrom(const lMatrix<T> &mat) {
    if (this == &mat) return;
    if (const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat)) {
        if (!h_ || m->get_nnz() != this->nnz_ || m->get_num_row() != this->nrow_ || m->get_num_col() != this->ncol_)
            Init(m->get_nnz(), m->get_num_row(), m->get_num_col(), this->name_);
        drop_applies();
        // the pattern with zero values; CopyFrom brings the values
        std::vector<T> zero(m->col.size(), T(0));
        check(hfh_csr_upload(h_, m->row.data(), m->col.data(), zero.data()), "HIP_CSR_lMatrix::CopyStructureFrom");
        hrow_ = m->row;
        hcol_ = m->col;
        structure_ = true;
        return;
    }
    if (const HIP_CSR_lMatrix<T> *d = dynamic_cast<const HIP_CSR_lMatrix<T> *>(&mat)) {
        if (!d->structure_) throw Error(HFH_ERR_INVALID, "CopyStructureFrom: source has no structure");
        if (!h_ || d->nnz_ != this->nnz_ || d->nrow_ != this->nrow_ || d->ncol_ != this->ncol_)
            Init(d->nnz_, d->nrow_, d->ncol_, this->name_);
        drop_applies();
        check(hfh_csr_cast(h_, d->h_), "HIP_CSR_lMatrix::CopyStructureFrom");
        check(hfh_csr_zeros(h_), "HIP_CSR_lMatrix::CopyStructureFrom");
        hrow_ = d->hrow_;
        hcol_ = d->hcol_;
        structure_ = true;
        return;
    }
    mat.CopyStructureTo(*this);
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyStructureTo(lMatrix<T> &mat) const {
    if (CPU_CSR_lMatrix<T> *m = dynamic_cast<CPU_CSR_lMatrix<T> *>(&mat)) {
        if (!structure_) throw Error(HFH_ERR_INVALID, "CopyStructureTo: matrix has no structure");
        m->Init(this->nnz_, this->nrow_, this->ncol_, m->get_name());
        m->row = hrow_;
        m->col = hcol_;
        return;
    }
    if (dynamic_cast<HIP_CSR_lMatrix<T> *>(&mat)) {
        mat.CopyStructureFrom(*this);
        return;
    }
    unsupported("HIP_CSR_lMatrix", "CopyStructureTo (destination format)");
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyFrom(const lMatrix<T> &mat) {
    if (this == &mat) return;
    if (const CPU_CSR_lMatrix<T> *m = dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat)) {
        if (!structure_) CopyStructureFrom(mat);
        if (m->col.size() != static_cast<size_t>(this->nnz_))
            throw Error(HFH_ERR_INVALID, "HIP_CSR_lMatrix::CopyFrom: structure mismatch");
        check(hfh_csr_upload_values(h_, m->val.data()), "HIP_CSR_lMatrix::CopyFrom");
        return;
    }
    if (const HIP_CSR_lMatrix<T> *d = dynamic_cast<const HIP_CSR_lMatrix<T> *>(&mat)) {
        if (!d->structure_) throw Error(HFH_ERR_INVALID, "HIP_CSR_lMatrix::CopyFrom: source has no structure");
        if (!h_ || d->nnz_ != this->nnz_ || d->nrow_ != this->nrow_) Init(d->nnz_, d->nrow_, d->ncol_, this->name_);
        check(hfh_csr_cast(h_, d->h_), "HIP_CSR_lMatrix::CopyFrom");
        if (!structure_ || hcol_ != d->hcol_ || hrow_ != d->hrow_) {
            drop_applies();
            hrow_ = d->hrow_;
            hcol_ = d->hcol_;
        }
        structure_ = true;
        return;
    }
    mat.CopyTo(*this);
}

template <typename T>
void HIP_CSR_lMatrix<T>::CopyTo(lMatrix<T> &mat) const {
    if (CPU_CSR_lMatrix<T> *m = dynamic_cast<CPU_CSR_lMatrix<T> *>(&mat)) {
        if (!structure_) throw Error(HFH_ERR_INVALID, "HIP_CSR_lMatrix::CopyTo: matrix has no structure");
        if (m->col.size() != static_cast<size_t>(this->nnz_)) CopyStructureTo(*m);
        check(hfh_csr_download(h_, nullptr, nullptr, m->val.data()), "HIP_CSR_lMatrix::CopyTo");
        return;
    }
    if (dynamic_cast<HIP_CSR_lMatrix<T> *>(&mat)) {
        mat.CopyFrom(*this);
        return;
    }
    unsupported("HIP_CSR_lMatrix", "CopyTo (destination format)");
}

template <typename T, typename U>
static void hip_cast_from(HIP_CSR_lMatrix<T> &self, const lMatrix<U> &other) {
    if (const HIP_CSR_lMatrix<U> *d = dynamic_cast<const HIP_CSR_lMatrix<U> *>(&other)) {
        if (!d->handle()) throw Error(HFH_ERR_INVALID, "CastFrom: source has no structure");
        self.Init(d->get_nnz(), d->get_num_row(), d->get_num_col(), self.get_name());
        check(hfh_csr_cast(self.handle(), d->handle()), "HIP_CSR_lMatrix::CastFrom");
        return;
    }
    if (const CPU_CSR_lMatrix<U> *m = dynamic_cast<const CPU_CSR_lMatrix<U> *>(&other)) {
        CPU_CSR_lMatrix<T> host;
        host.CastFrom(*m);
        self.CloneFrom(host);
        return;
    }
    other.CastTo(self);
}

template <typename T>
void HIP_CSR_lMatrix<T>::CastFrom(const lMatrix<double> &other) {
    hip_cast_from<T, double>(*this, other);
    if (const HIP_CSR_lMatrix<double> *d = dynamic_cast<const HIP_CSR_lMatrix<double> *>(&other)) {
        drop_applies();
        hrow_ = d->hrow_;
        hcol_ = d->hcol_;
        structure_ = true;
    }
}

template <typename T>
void HIP_CSR_lMatrix<T>::CastFrom(const lMatrix<float> &other) {
    hip_cast_from<T, float>(*this, other);
    if (const HIP_CSR_lMatrix<float> *d = dynamic_cast<const HIP_CSR_lMatrix<float> *>(&other)) {
        drop_applies();
        hrow_ = d->hrow_;
        hcol_ = d->hcol_;
        structure_ = true;
    }
}

template <typename T, typename U>
static void hip_cast_to(const HIP_CSR_lMatrix<T> &self, lMatrix<U> &other) {
    if (dynamic_cast<HIP_CSR_lMatrix<U> *>(&other)) {
        other.CastFrom(self);
        return;
    }
    if (CPU_CSR_lMatrix<U> *m = dynamic_cast<CPU_CSR_lMatrix<U> *>(&other)) {
        CPU_CSR_lMatrix<T> host;
        host.CloneFrom(self);
        m->CastFrom(host);
        return;
    }
    unsupported("HIP_CSR_lMatrix", "CastTo (destination format)");
}

template <typename T>
void HIP_CSR_lMatrix<T>::CastTo(lMatrix<double> &other) const { hip_cast_to<T, double>(*this, other); }

template <typename T>
void HIP_CSR_lMatrix<T>::CastTo(lMatrix<float> &other) const { hip_cast_to<T, float>(*this, other); }

template <typename T>
void HIP_CSR_lMatrix<T>::ConvertFrom(const lMatrix<T> &mat2) {
    if (this == &mat2) return;
    if (dynamic_cast<const CPU_CSR_lMatrix<T> *>(&mat2) || dynamic_cast<const HIP_CSR_lMatrix<T> *>(&mat2)) {
        this->CloneFrom(mat2);
        return;
    }
    CPU_CSR_lMatrix<T> host;
    host.ConvertFrom(mat2);
    this->CloneFrom(host);
}

template <typename T>
lMatrix<T> *HIP_CSR_lMatrix<T>::CloneWithoutContent() const {
    HIP_CSR_lMatrix<T> *m = new HIP_CSR_lMatrix<T>();
    m->Init(this->nnz_, this->nrow_, this->ncol_, "clone from " + this->name_);
    return m;
}

template <typename T>
void HIP_CSR_lMatrix<T>::Sync() const { check(hfh_ctx_sync(hip_context()), "HIP_CSR_lMatrix::Sync"); }

template <typename T>
void HIP_CSR_lMatrix<T>::ZeroRows(const int *index_set, const int size, const T alpha) {
    check(hfh_csr_zero_rows(h_, index_set, size, alpha), "HIP_CSR_lMatrix::ZeroRows");
}

template <typename T>
void HIP_CSR_lMatrix<T>::ZeroCols(const int *index_set, const int size, const T alpha) {
    check(hfh_csr_zero_cols(h_, index_set, size, alpha), "HIP_CSR_lMatrix::ZeroCols");
}

template <typename T>
void HIP_CSR_lMatrix<T>::Scale(const T alpha) { check(hfh_csr_scale(h_, alpha, 0), "HIP_CSR_lMatrix::Scale"); }

template <typename T>
void HIP_CSR_lMatrix<T>::ScaleOffdiag(const T alpha) {
    check(hfh_csr_scale(h_, alpha, 1), "HIP_CSR_lMatrix::ScaleOffdiag");
}

// CPU_CSR_lMatrix::Reorder (lmatrix_csr_cpu.cc:1847-1862) on the device: the
// symmetric permutation, perm[new] = old
template <typename T>
void HIP_CSR_lMatrix<T>::Reorder(const int *index) {
    if (!structure_) throw Error(HFH_ERR_INVALID, "Reorder: matrix has no structure");
    hfh_csr *P = nullptr;
    check(hfh_csr_create(hip_context(), this->nrow_, this->ncol_, this->nnz_, dtype_of<T>::value, &P),
          "HIP_CSR_lMatrix::Reorder");
    int rc = hfh_csr_permute(h_, index, P);
    if (rc) {
        hfh_csr_destroy(P);
        check(rc, "HIP_CSR_lMatrix::Reorder");
    }
    std::vector<int> r(this->nrow_ + 1), c(this->nnz_);
    check(hfh_csr_download(P, r.data(), c.data(), nullptr), "HIP_CSR_lMatrix::Reorder");
    adopt(P, r, c);
}

template <typename T>
void HIP_CSR_lMatrix<T>::Multicoloring(int &ncolors, int **color_sizes, int **permut_index) const {
    std::vector<int> sizes(std::max(this->nrow_, 1)), perm(std::max(this->nrow_, 1));
    check(hfh_csr_multicoloring(h_, &ncolors, nullptr, sizes.data(), perm.data()), "HIP_CSR_lMatrix::Multicoloring");
    *color_sizes = new int[ncolors];
    *permut_index = new int[this->nrow_];
    std::copy(sizes.begin(), sizes.begin() + ncolors, *color_sizes);
    std::copy(perm.begin(), perm.begin() + this->nrow_, *permut_index);
}

template <typename T>
void HIP_CSR_lMatrix<T>::Levelscheduling(int &nlevels, int **level_sizes, int **permut_index) const {
    if (!structure_) throw Error(HFH_ERR_INVALID, "Levelscheduling: matrix has no structure");
    std::vector<int> sizes(std::max(this->nrow_, 1)), perm(std::max(this->nrow_, 1));
    levelscheduling(this->nrow_, hrow_.data(), hcol_.data(), &nlevels, sizes.data(), perm.data());
    *level_sizes = new int[nlevels];
    *permut_index = new int[this->nrow_];
    std::copy(sizes.begin(), sizes.begin() + nlevels, *level_sizes);
    std::copy(perm.begin(), perm.begin() + this->nrow_, *permut_index);
}

// level schedules of the forward and backward sweeps of this pattern
template <typename T>
static void sweep_levels(const HIP_CSR_lMatrix<T> &A, std::vector<int> &lp, std::vector<int> &lr,
                         std::vector<int> &up, std::vector<int> &ur) {
    const int n = A.get_num_row();
    lr.assign(std::max(n, 1), 0);
    ur.assign(std::max(n, 1), 0);
    levels_lower(n, A.host_row().data(), A.host_col().data(), lp, lr.data());
    levels_upper(n, A.host_row().data(), A.host_col().data(), up, ur.data());
}

template <typename T>
void HIP_CSR_lMatrix<T>::transpose_me() {
    CPU_CSR_lMatrix<T> host;
    host.CloneFrom(*this);
    host.transpose_me();
    this->CloneFrom(host);
}

// CPU_CSR_lMatrix::ilu0 (lmatrix_csr_cpu.cc:1880-1930) in place on the device
template <typename T>
void HIP_CSR_lMatrix<T>::ilu0() {
    if (!structure_) throw Error(HFH_ERR_INVALID, "ilu0: matrix has no structure");
    std::vector<int> lp, lr, up, ur;
    sweep_levels(*this, lp, lr, up, ur);
    int missing = 0;
    check(hfh_csr_ilu0(h_, static_cast<int>(lp.size()) - 1, lp.data(), lr.data(), &missing), "HIP_CSR_lMatrix::ilu0");
    if (missing) throw Error(HFH_ERR_ZERO_PIVOT, "ilu0: vanishing diagonal elements");
}

// ilu_solve (lmatrix_csr_cpu.cc:2108-2171): this matrix holds the factors
template <typename T>
void HIP_CSR_lMatrix<T>::ilu_solve(const lVector<T> &invec, lVector<T> *outvec) const {
    if (!ilu_) {
        std::vector<int> lp, lr, up, ur;
        sweep_levels(*this, lp, lr, up, ur);
        check(hfh_ilu0_create(h_, static_cast<int>(lp.size()) - 1, lp.data(), lr.data(), static_cast<int>(up.size()) - 1,
                              up.data(), ur.data(), &ilu_),
              "HIP_CSR_lMatrix::ilu_solve");
    }
    check(hfh_prec_apply(ilu_, as_hip(invec, "ilu_solve").handle(), as_hip(*outvec, "ilu_solve").handle()),
          "HIP_CSR_lMatrix::ilu_solve");
}

// Psgauss_seidel (lmatrix_csr_cpu.cc:542-616)
template <typename T>
void HIP_CSR_lMatrix<T>::Psgauss_seidel(const lVector<T> &invec, lVector<T> *outvec) const {
    if (!sgs_) {
        std::vector<int> lp, lr, up, ur;
        sweep_levels(*this, lp, lr, up, ur);
        check(hfh_sgs_create(h_, static_cast<int>(lp.size()) - 1, lp.data(), lr.data(), static_cast<int>(up.size()) - 1,
                             up.data(), ur.data(), &sgs_),
              "HIP_CSR_lMatrix::Psgauss_seidel");
    }
    check(hfh_prec_apply(sgs_, as_hip(invec, "Psgauss_seidel").handle(), as_hip(*outvec, "Psgauss_seidel").handle()),
          "HIP_CSR_lMatrix::Psgauss_seidel");
}

template <typename T>
void HIP_CSR_lMatrix<T>::Pjacobi(const lVector<T> &invec, lVector<T> *outvec) const {
    check(hfh_csr_pjacobi(h_, as_hip(invec, "Pjacobi").handle(), as_hip(*outvec, "Pjacobi").handle()),
          "HIP_CSR_lMatrix::Pjacobi");
}

template <typename T>
lMatrix<T> *HIP_CSR_lMatrix<T>::extract_submatrix(const int start_row, const int start_col, const int end_row,
                                                  const int end_col) const {
    CPU_CSR_lMatrix<T> host;
    host.CloneFrom(*this);
    std::unique_ptr<lMatrix<T>> sub(host.extract_submatrix(start_row, start_col, end_row, end_col));
    HIP_CSR_lMatrix<T> *out = new HIP_CSR_lMatrix<T>();
    out->Init(sub->get_nnz(), sub->get_num_row(), sub->get_num_col(), "sub matrix from " + this->name_);
    if (sub->get_nnz() > 0) out->CloneFrom(*sub);
    return out;
}

template <typename T>
static void diag_into(const HIP_CSR_lMatrix<T> &A, int s, int e, int invert, lVector<T> *vec) {
    if (HIP_lVector<T> *h = dynamic_cast<HIP_lVector<T> *>(vec)) {
        check(hfh_csr_diag(A.handle(), s, e, invert, h->handle()), "extract_diagelements");
        return;
    }
    HIP_lVector<T> tmp(e - s, "diag");
    check(hfh_csr_diag(A.handle(), s, e, invert, tmp.handle()), "extract_diagelements");
    std::vector<T> d(e - s);
    tmp.GetBlockValues(0, e - s, d.data());
    vec->SetBlockValues(0, e - s, d.data());
}

template <typename T>
void HIP_CSR_lMatrix<T>::extract_diagelements(const int s, const int e, lVector<T> *vec) const {
    diag_into(*this, s, e, 0, vec);
}

template <typename T>
void HIP_CSR_lMatrix<T>::extract_invdiagelements(const int s, const int e, lVector<T> *vec) const {
    diag_into(*this, s, e, 1, vec);
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

// VectorMultNoDiag (lmatrix_csr_cpu.cc:498-540): the row sums without the
// entries where column == row
template <typename T>
void HIP_CSR_lMatrix<T>::VectorMultNoDiag(const lVector<T> &in, lVector<T> *out) const {
    check(hfh_spmv_nodiag(h_, as_hip(in, "VectorMultNoDiag").handle(), as_hip(*out, "VectorMultNoDiag").handle()),
          "HIP_CSR_lMatrix::VectorMultNoDiag");
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
void lPreconditioner_SymmetricGaussSeidel<T>::Build() {
    this->Clear();
    const HIP_CSR_lMatrix<T> &A = this->hip_op();
    std::vector<int> lp, lr, up, ur;
    sweep_levels(A, lp, lr, up, ur);
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
    cpu.CloneFrom(A);
    cpu.ilu0();
    LU_.reset(new HIP_CSR_lMatrix<T>());
    LU_->CloneFrom(cpu);
    std::vector<int> lp, lr, up, ur;
    sweep_levels(*LU_, lp, lr, up, ur);
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
    template void rotg<T>(T *, T *, T *, T *);                                                              \
    template void rotmg<T>(T *, T *, T *, T, T *);                                                          \
    template class lVector<T>;                                                                              \
    template class lMatrix<T>;                                                                              \
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
