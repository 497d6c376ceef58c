// solvers.h -- LinearSolver<LAD> templates (src/linear_solver/*) for the
// single-rank LA descriptors of coupled.h.  The iteration of each method is
// the reference's step for step (same vector operations in the same order),
// so the HIP platform reproduces the CPU path's iterates; each Dot returns a
// host scalar as the lVector interface requires.  The device-resident
// drivers of include/hfh.h (hfh_pcg_solve, hfh_gmres_solve, hfh_dc_solve)
// run the same iterations without the per-Dot synchronisation.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "coupled.h"

namespace hflmp {

// preconditioner.h:32-37
enum LinearSolverState { kSolverSuccess = 0, kSolverExceeded, kSolverError };

// common/iterate_control.{h,cc}
class IterateControl {
  public:
    enum State { kIterate = 0, kSuccessAbsoluteTol, kSuccessRelativeTol, kFailureDivergenceTol, kFailureMaxitsExceeded };
    IterateControl() { Init(1000, 1.e-12, 1.e-8, 1.e8); }
    void Init(int maxits, double atol, double rtol, double dtol) {
        if (maxits <= 0 || atol < 0 || rtol < 0 || !(dtol > 1))
            throw Error(HFH_ERR_INVALID, "IterateControl::Init: need maxits > 0, atol >= 0, rtol >= 0, dtol > 1");
        maxits_ = maxits;
        atol_ = atol;
        rtol_ = rtol;
        dtol_ = dtol;
    }
    // iterate_control.cc:84-110: step 0 records the first value; abs, rel,
    // divergence and maxits tests in that order, strict comparisons
    State Check(int step, double value) {
        if (step == 0) first_ = value;
        if (value < atol_) return status_ = kSuccessAbsoluteTol;
        if (value < rtol_ * first_) return status_ = kSuccessRelativeTol;
        if (value > dtol_ * first_) return status_ = kFailureDivergenceTol;
        if (step >= maxits_) return status_ = kFailureMaxitsExceeded;
        return status_ = kIterate;
    }
    int maxits() const { return maxits_; }
    State status() const { return status_; }
    double first_value() const { return first_; }

  private:
    int maxits_;
    double atol_, rtol_, dtol_, first_ = 0;
    State status_ = kIterate;
};

template <class LAD>
class Preconditioner {
  public:
    typedef typename LAD::MatrixType OperatorType;
    typedef typename LAD::VectorType VectorType;
    typedef typename LAD::DataType DataType;
    virtual ~Preconditioner() {}
    virtual void SetupOperator(OperatorType &op) { op_ = &op; state_ = false; }
    virtual void Build() { state_ = true; }
    virtual LinearSolverState ApplyPreconditioner(const VectorType &b, VectorType *x) = 0;
    const OperatorType *GetOperator() const { return op_; }
    bool GetState() const { return state_; }
    void SetState(bool s) { state_ = s; }
    bool GetReuse() const { return reuse_; }
    void SetReuse(bool r) { reuse_ = r; }
    // the device handle behind a local preconditioner, for the fused drivers
    virtual hfh_prec *device_handle() const { return nullptr; }

  protected:
    OperatorType *op_ = nullptr;
    bool state_ = false, reuse_ = true;
};

template <class LAD>
class LinearSolver : public Preconditioner<LAD> {
  public:
    typedef typename LAD::MatrixType OperatorType;
    typedef typename LAD::VectorType VectorType;
    typedef typename LAD::DataType DataType;
    LinearSolver() { control_.Init(1000, 1e-15, 1e-8, 1e6); }  // linear_solver.h:48-52
    // LinearSolver::InitControl (linear_solver.h:290-305)
    virtual void InitControl(int maxits, double abstol, double reltol, double divtol) {
        control_.Init(maxits, abstol, reltol, divtol);
    }
    virtual void SetupPreconditioner(Preconditioner<LAD> &precond) {
        precond_ = &precond;
        method_ = "RightPreconditioning";
    }
    void SetupOperator(OperatorType &op) override {
        this->op_ = &op;
        this->state_ = false;
    }
    // linear_solver.h:169-182
    virtual void BuildPreconditioner() {
        if (!precond_->GetOperator()) precond_->SetupOperator(*this->op_);
        precond_->Build();
        precond_->SetState(true);
    }
    virtual LinearSolverState Solve(const VectorType &b, VectorType *x) = 0;
    LinearSolverState ApplyPreconditioner(const VectorType &b, VectorType *x) override { return Solve(b, x); }
    IterateControl &control() { return control_; }
    int iter() const { return iter_; }
    double res() const { return res_; }
    const std::string &Method() const { return method_; }
    const std::vector<double> &history() const { return hist_; }

  protected:
    static LinearSolverState state_of(IterateControl::State s) {
        return (s == IterateControl::kFailureDivergenceTol || s == IterateControl::kFailureMaxitsExceeded)
                   ? kSolverExceeded
                   : kSolverSuccess;
    }
    IterateControl control_;
    Preconditioner<LAD> *precond_ = nullptr;
    std::string method_ = "NoPreconditioning";
    int iter_ = 0;
    double res_ = 0;
    std::vector<double> hist_;
};

// linear_solver/cg.cc:84-248
template <class LAD>
class CG : public LinearSolver<LAD> {
  public:
    typedef typename LAD::VectorType VectorType;
    typedef typename LAD::DataType DataType;
    void InitParameter(const std::string &method) { this->method_ = method; }

    LinearSolverState Solve(const VectorType &b, VectorType *x) override {
        if (this->method_ == "NoPreconditioning") return SolveNoPrecond(b, x);
        if (this->method_ == "Preconditioning" || this->method_ == "RightPreconditioning" ||
            this->method_ == "LeftPreconditioning")
            return SolvePrecond(b, x);
        return kSolverError;
    }

  private:
    LinearSolverState SolveNoPrecond(const VectorType &b, VectorType *x) {
        IterateControl::State conv;
        VectorType r, p, Ap;
        r.CloneFromWithoutContent(b);
        p.CloneFromWithoutContent(b);
        Ap.CloneFromWithoutContent(b);
        DataType alpha, beta, rr;
        this->iter_ = 0;
        this->hist_.clear();
        this->op_->VectorMult(*x, &r);
        r.ScaleAdd(b, static_cast<DataType>(-1.));
        rr = r.Dot(r);
        this->res_ = std::sqrt(rr);
        this->hist_.push_back(this->res_);
        conv = this->control_.Check(this->iter_, this->res_);
        p.CopyFrom(r);
        beta = rr;
        while (conv == IterateControl::kIterate) {
            ++this->iter_;
            this->op_->VectorMult(p, &Ap);
            alpha = beta / Ap.Dot(p);
            x->Axpy(p, alpha);
            r.Axpy(Ap, -alpha);
            rr = r.Dot(r);
            this->res_ = std::sqrt(rr);
            this->hist_.push_back(this->res_);
            conv = this->control_.Check(this->iter_, this->res_);
            if (conv != IterateControl::kIterate) break;
            beta = rr / beta;
            p.ScaleAdd(r, beta);
            beta = rr;
        }
        return this->state_of(conv);
    }

    LinearSolverState SolvePrecond(const VectorType &b, VectorType *x) {
        if (!this->precond_) return kSolverError;
        if (!this->precond_->GetState() || !this->precond_->GetReuse()) this->BuildPreconditioner();
        IterateControl::State conv;
        VectorType r, p, z, Ap;
        r.CloneFromWithoutContent(b);
        p.CloneFromWithoutContent(b);
        z.CloneFromWithoutContent(b);
        Ap.CloneFromWithoutContent(b);
        DataType alpha, beta, gamma, rr;
        this->iter_ = 0;
        this->hist_.clear();
        this->op_->VectorMult(*x, &r);
        r.ScaleAdd(b, static_cast<DataType>(-1.));
        rr = r.Dot(r);
        this->res_ = std::sqrt(rr);
        this->hist_.push_back(this->res_);
        conv = this->control_.Check(this->iter_, this->res_);
        z.Zeros();
        this->precond_->ApplyPreconditioner(r, &z);
        p.CopyFrom(z);
        beta = r.Dot(z);
        while (conv == IterateControl::kIterate) {
            ++this->iter_;
            this->op_->VectorMult(p, &Ap);
            alpha = beta / Ap.Dot(p);
            x->Axpy(p, alpha);
            r.Axpy(Ap, -alpha);
            rr = r.Dot(r);
            this->res_ = std::sqrt(rr);
            this->hist_.push_back(this->res_);
            conv = this->control_.Check(this->iter_, this->res_);
            if (conv != IterateControl::kIterate) break;
            z.Zeros();
            this->precond_->ApplyPreconditioner(r, &z);
            gamma = r.Dot(z);
            beta = gamma / beta;
            p.ScaleAdd(z, beta);
            beta = gamma;
        }
        return this->state_of(conv);
    }
};

// Givens rotations (gmres.h:80-97)
template <typename T>
inline void apply_plane_rotation(T cs, T sn, T *dx, T *dy) {
    const T t = cs * (*dx) + sn * (*dy);
    *dy = -sn * (*dx) + cs * (*dy);
    *dx = t;
}
template <typename T>
inline void generate_plane_rotation(T dx, T dy, T *cs, T *sn) {
    const T beta = std::sqrt(dx * dx + dy * dy);
    *cs = dx / beta;
    *sn = dy / beta;
}

// GMRES<LAD>::SolveRight (gmres.cc:285-510) and FGMRES<LAD>::SolveRight
// (fgmres.cc:75-262): right preconditioning, modified Gram-Schmidt, Givens
// rotations; the Hessenberg column j is H[j][0..j+1].
template <class LAD, bool FLEXIBLE>
class GMRESBase : public LinearSolver<LAD> {
  public:
    typedef typename LAD::VectorType VectorType;
    typedef typename LAD::DataType DataType;
    void InitParameter(int size_basis, const std::string &method) {
        size_basis_ = size_basis;
        this->method_ = method;
    }
    // GMRES::Solve (gmres.cc:80-98): dispatch on the method; left
    // preconditioning is not implemented in the reference either
    // (gmres.cc:267-279, fgmres.cc:59-70) and returns kSolverError
    LinearSolverState Solve(const VectorType &b, VectorType *x) override {
        if (size_basis_ <= 0) return kSolverError;
        if (this->method_ == "NoPreconditioning") return SolveNoPrecond(b, x);
        if (this->method_ != "RightPreconditioning" || !this->precond_) return kSolverError;
        return SolveRight(b, x);
    }

  private:
    void UpdateSolution(const std::vector<std::unique_ptr<VectorType>> &V, const std::vector<std::vector<DataType>> &H,
                        const std::vector<DataType> &g, int k, VectorType *x) const {
        std::vector<DataType> y(g);
        for (int i = k + 1; i--;) {
            y[i] /= H[i][i];
            const DataType t = y[i];
            for (int j = 0; j < i; ++j) y[j] -= H[i][j] * t;
        }
        for (int j = 0; j <= k; ++j) x->Axpy(*V[j], y[j]);
    }

    // GMRES::SolveNoPrecond (gmres.cc:103-264): the Arnoldi process on A
    // itself, x += V y at convergence and at every restart (FGMRES uses the
    // same path without a preconditioner)
    LinearSolverState SolveNoPrecond(const VectorType &b, VectorType *x) {
        IterateControl::State conv;
        const int m = std::min(size_basis_, this->control_.maxits());
        std::vector<std::vector<DataType>> H(m, std::vector<DataType>(m + 1, DataType(0)));
        std::vector<std::unique_ptr<VectorType>> V(m + 1);
        for (int i = 0; i <= m; ++i) V[i].reset(new VectorType());
        VectorType w;
        w.CloneFromWithoutContent(b);
        std::vector<DataType> g(m + 1), cs(m + 1), sn(m + 1);
        int iter = 0;
        this->hist_.clear();
        V[0]->CloneFromWithoutContent(b);
        this->op_->VectorMult(*x, V[0].get());
        V[0]->ScaleAdd(b, static_cast<DataType>(-1.));
        this->res_ = V[0]->Norm2();
        this->hist_.push_back(this->res_);
        conv = this->control_.Check(iter, this->res_);
        while (conv == IterateControl::kIterate) {
            std::fill(g.begin(), g.end(), DataType(0));
            for (auto &h : H) std::fill(h.begin(), h.end(), DataType(0));
            V[0]->Scale(static_cast<DataType>(1.) / static_cast<DataType>(this->res_));
            g[0] = static_cast<DataType>(this->res_);
            for (int j = 0; j != m; ++j) {
                ++iter;
                this->op_->VectorMult(*V[j], &w);
                for (int i = 0; i <= j; ++i) {
                    H[j][i] = w.Dot(*V[i]);
                    w.Axpy(*V[i], static_cast<DataType>(-1.) * H[j][i]);
                }
                H[j][j + 1] = w.Norm2();
                if (H[j][j + 1] == DataType(0)) return kSolverError;
                w.Scale(static_cast<DataType>(1.) / H[j][j + 1]);
                V[j + 1]->CloneFrom(w);
                for (int k = 0; k < j; ++k) apply_plane_rotation(cs[k], sn[k], &H[j][k], &H[j][k + 1]);
                generate_plane_rotation(H[j][j], H[j][j + 1], &cs[j], &sn[j]);
                apply_plane_rotation(cs[j], sn[j], &H[j][j], &H[j][j + 1]);
                apply_plane_rotation(cs[j], sn[j], &g[j], &g[j + 1]);
                this->res_ = std::abs(g[j + 1]);
                this->hist_.push_back(this->res_);
                conv = this->control_.Check(iter, this->res_);
                if (conv != IterateControl::kIterate) {
                    UpdateSolution(V, H, g, j, x);
                    break;
                }
            }
            if (conv == IterateControl::kIterate) {
                UpdateSolution(V, H, g, m - 1, x);
                this->op_->VectorMult(*x, V[0].get());
                V[0]->ScaleAdd(b, static_cast<DataType>(-1.));
                // the restart normalises with res = |g[m]| (gmres.cc:163)
            }
        }
        this->iter_ = iter;
        return this->state_of(conv);
    }

    LinearSolverState SolveRight(const VectorType &b, VectorType *x) {
        if (!this->precond_->GetState() || !this->precond_->GetReuse()) this->BuildPreconditioner();
        IterateControl::State conv;
        const int m = std::min(size_basis_, this->control_.maxits());
        std::vector<std::vector<DataType>> H(m, std::vector<DataType>(m + 1, DataType(0)));
        std::vector<std::unique_ptr<VectorType>> V(m + 1), Z(FLEXIBLE ? m + 1 : 0);
        for (int i = 0; i <= m; ++i) {
            V[i].reset(new VectorType());
            if (FLEXIBLE) Z[i].reset(new VectorType());
        }
        VectorType w, z;
        w.CloneFromWithoutContent(b);
        z.CloneFromWithoutContent(b);
        std::vector<DataType> g(m + 1), cs(m + 1), sn(m + 1);
        int iter = 0;
        this->hist_.clear();
        V[0]->CloneFromWithoutContent(b);
        this->op_->VectorMult(*x, V[0].get());
        V[0]->ScaleAdd(b, static_cast<DataType>(-1.));
        this->res_ = V[0]->Norm2();
        this->hist_.push_back(this->res_);
        conv = this->control_.Check(iter, this->res_);
        while (conv == IterateControl::kIterate) {
            std::fill(g.begin(), g.end(), DataType(0));
            for (auto &h : H) std::fill(h.begin(), h.end(), DataType(0));
            V[0]->Scale(static_cast<DataType>(1.) / static_cast<DataType>(this->res_));
            g[0] = static_cast<DataType>(this->res_);
            for (int j = 0; j != m; ++j) {
                ++iter;
                VectorType *zj = &z;
                if (FLEXIBLE) {
                    Z[j]->CloneFromWithoutContent(b);
                    zj = Z[j].get();
                } else {
                    z.Zeros();
                }
                this->precond_->ApplyPreconditioner(*V[j], zj);
                this->op_->VectorMult(*zj, &w);
                for (int i = 0; i <= j; ++i) {
                    H[j][i] = w.Dot(*V[i]);
                    w.Axpy(*V[i], static_cast<DataType>(-1.) * H[j][i]);
                }
                H[j][j + 1] = w.Norm2();
                if (H[j][j + 1] == DataType(0)) return kSolverError;
                w.Scale(static_cast<DataType>(1.) / H[j][j + 1]);
                V[j + 1]->CloneFrom(w);
                for (int k = 0; k < j; ++k) apply_plane_rotation(cs[k], sn[k], &H[j][k], &H[j][k + 1]);
                generate_plane_rotation(H[j][j], H[j][j + 1], &cs[j], &sn[j]);
                apply_plane_rotation(cs[j], sn[j], &H[j][j], &H[j][j + 1]);
                apply_plane_rotation(cs[j], sn[j], &g[j], &g[j + 1]);
                this->res_ = std::abs(g[j + 1]);
                this->hist_.push_back(this->res_);
                conv = this->control_.Check(iter, this->res_);
                if (conv != IterateControl::kIterate) {
                    if (FLEXIBLE) {
                        UpdateSolution(Z, H, g, j, x);
                    } else {
                        z.Zeros();
                        UpdateSolution(V, H, g, j, &z);
                        w.Zeros();
                        this->precond_->ApplyPreconditioner(z, &w);
                        x->Axpy(w, static_cast<DataType>(1.));
                    }
                    break;
                }
            }
            if (conv == IterateControl::kIterate) {
                if (FLEXIBLE) {
                    UpdateSolution(Z, H, g, m - 1, x);
                    this->op_->VectorMult(*x, V[0].get());
                    V[0]->Scale(static_cast<DataType>(-1.));
                    V[0]->Axpy(b, static_cast<DataType>(1.));
                } else {
                    z.Zeros();
                    UpdateSolution(V, H, g, m - 1, &z);
                    w.Zeros();
                    this->precond_->ApplyPreconditioner(z, &w);
                    x->Axpy(w, static_cast<DataType>(1.));
                    this->op_->VectorMult(*x, V[0].get());
                    V[0]->ScaleAdd(b, static_cast<DataType>(-1.));
                }
                // the restart normalises with res = |g[m]| (gmres.cc:356-358)
            }
        }
        this->iter_ = iter;
        return this->state_of(conv);
    }

    int size_basis_ = 0;
};

template <class LAD> using GMRES = GMRESBase<LAD, false>;
template <class LAD> using FGMRES = GMRESBase<LAD, true>;

// Identity preconditioner (GMRES needs one for right preconditioning)
template <class LAD>
class PreconditionerIdentity : public Preconditioner<LAD> {
  public:
    LinearSolverState ApplyPreconditioner(const typename LAD::VectorType &b, typename LAD::VectorType *x) override {
        x->CopyFrom(b);
        return kSolverSuccess;
    }
};

// preconditioner_bjacobi_standard.cc:45-150
template <class LAD>
class PreconditionerBlockJacobiStand : public Preconditioner<LAD> {
  public:
    typedef typename LAD::MatrixType OperatorType;
    typedef typename LAD::VectorType VectorType;
    typedef typename LAD::DataType DataType;
    void Init_Jacobi(const VectorType &) { local_.reset(new lPreconditioner_Jacobi<DataType>()); }
    void Init_SymmetricGaussSeidel() { local_.reset(new lPreconditioner_SymmetricGaussSeidel<DataType>()); }
    void Init_ILUp(int p) {
        lPreconditioner_ILUp<DataType> *lp = new lPreconditioner_ILUp<DataType>();
        lp->Init(p);
        local_.reset(lp);
    }
    void SetupOperator(OperatorType &op) override {
        this->op_ = &op;
        local_->SetupOperator(op.diagonal());
        this->state_ = false;
    }
    void Build() override {
        local_->Build();
        this->state_ = true;
    }
    LinearSolverState ApplyPreconditioner(const VectorType &b, VectorType *x) override {
        if (!this->state_) Build();
        local_->ApplylPreconditioner(b.interior(), &x->interior());
        return kSolverSuccess;
    }
    hfh_prec *device_handle() const override { return local_ ? local_->handle() : nullptr; }

  private:
    std::unique_ptr<lPreconditioner<DataType>> local_;
};

// preconditioner_mc.cc:49-194.  Preprocess colours the operator and returns
// the permutation (perm[new] = old) with which the caller renumbers the
// system, as PermuteDofNumbering does in the reference; SetupOperator then
// takes the permuted operator.
template <class LAD>
class PreconditionerMultiColoring : public Preconditioner<LAD> {
  public:
    typedef typename LAD::MatrixType OperatorType;
    typedef typename LAD::VectorType VectorType;
    typedef typename LAD::DataType DataType;
    void Init_SymmetricGaussSeidel() { local_.reset(new lPreconditioner_MultiColoring_SymmetricGaussSeidel<DataType>()); }
    void SetRelax(DataType omega) { local_->SetRelax(omega); }
    std::vector<int> Preprocess(const CPU_CSR_lMatrix<DataType> &op) {
        int nc = 0;
        std::vector<int> sizes, perm;
        op.Multicoloring(nc, sizes, perm);
        local_->SetColors(sizes);
        return perm;
    }
    void SetupOperator(OperatorType &op) override {
        this->op_ = &op;
        local_->SetupOperator(op.diagonal());
        this->state_ = false;
    }
    void Build() override {
        local_->Build();
        this->state_ = true;
    }
    LinearSolverState ApplyPreconditioner(const VectorType &b, VectorType *x) override {
        if (!this->state_) Build();
        local_->ApplylPreconditioner(b.interior(), &x->interior());
        return kSolverSuccess;
    }
    hfh_prec *device_handle() const override { return local_ ? local_->handle() : nullptr; }

  private:
    std::unique_ptr<lPreconditioner_MultiColoring_SymmetricGaussSeidel<DataType>> local_;
};

// MPIR<LAD_high, LAD_low>::Solve (mpir.cc:69-327) without smoothers: FP64
// residual, cast to FP32, FP32 error-correction solve, cast back, x += c.
// res0 is reset to the last outer residual (mpir.cc:275).
template <class LADh, class LADl>
class MPIR : public LinearSolver<LADh> {
  public:
    typedef typename LADh::VectorType VectorH;
    typedef typename LADl::VectorType VectorL;
    typedef typename LADh::MatrixType MatrixH;
    typedef typename LADl::MatrixType MatrixL;
    typedef typename LADh::DataType DataH;
    typedef typename LADl::DataType DataL;
    void SetHighPrecMatrix(MatrixH &A) { Ah_ = &A; }
    void SetLowPrecMatrix(MatrixL &A) { Al_ = &A; }
    void SetLowPrecErrCorSolver(LinearSolver<LADl> &s) { ecl_ = &s; }
    void SetHighPrecErrCorSolver(LinearSolver<LADh> &s) { ech_ = &s; }
    void SetMpirParameters(int max_ech, int max_ecl, double abs_tol, double rel_tol, double div_tol) {
        max_ech_ = max_ech;
        max_ecl_ = max_ecl;
        abs_ = abs_tol;
        rel_ = rel_tol;
        div_ = div_tol;
    }
    // the low precision tolerances are stored in DataType_low, as the
    // reference's members are (mpir.h:225-240): the gates and the inner
    // solver's InitControl see them rounded to that type
    void SetLowPrecErrCorSolverParameters(int maxit, DataL atol, DataL rtol, DataL dtol) {
        ecl_max_ = maxit;
        ecl_atol_ = atol;
        ecl_rtol_ = rtol;
        ecl_dtol_ = dtol;
    }
    void SetHighPrecErrCorSolverParameters(int maxit, DataH atol, DataH rtol, DataH dtol) {
        ech_max_ = maxit;
        ech_atol_ = atol;
        ech_rtol_ = rtol;
        ech_dtol_ = dtol;
    }
    int n_ecl() const { return n_ecl_; }
    int n_ech() const { return n_ech_; }

    LinearSolverState Solve(const VectorH &b, VectorH *x) override {
        if (!Ah_ || !Al_ || !ecl_) return kSolverError;
        this->iter_ = 0;
        this->hist_.clear();
        ecl_->SetupOperator(*Al_);
        ecl_->InitControl(ecl_max_, ecl_atol_, ecl_rtol_, ecl_dtol_);
        if (ech_) {
            ech_->SetupOperator(*Ah_);
            ech_->InitControl(ech_max_, ech_atol_, ech_rtol_, ech_dtol_);
        }
        VectorH resh, corh;
        VectorL resl, corl;
        resh.CloneFromWithoutContent(b);
        corh.CloneFromWithoutContent(b);
        resl.Init(b.size_local(), HIP);
        corl.Init(b.size_local(), HIP);
        Ah_->VectorMult(*x, &resh);
        resh.ScaleAdd(b, -1.0);
        this->res_ = resh.Norm2();
        double res0 = this->res_;
        this->hist_.push_back(this->res_);
        // res0 is reset to the residual just computed after every correction
        // (mpir.cc:275): with relative tolerances below 1 the relative tests
        // compare the residual with itself and never end the loop
        auto ecl_ok = [&]() {
            return n_ecl_ < max_ecl_ && this->res_ > ecl_atol_ && this->res_ > res0 * ecl_rtol_ &&
                   this->res_ < ecl_dtol_;
        };
        auto ech_ok = [&]() {
            return ech_ && n_ech_ < max_ech_ && this->res_ > ech_atol_ && this->res_ > res0 * ech_rtol_ &&
                   this->res_ < ech_dtol_;
        };
        while (this->res_ > abs_ && this->res_ > res0 * rel_ && this->res_ < div_ && (ech_ok() || ecl_ok())) {
            if (ecl_ok()) {
                ++n_ecl_;
                resl.CastInteriorFrom(resh);
                corl.Zeros();
                ecl_->Solve(resl, &corl);
                this->iter_ += ecl_->iter();
                corh.CastInteriorFrom(corl);
            } else {
                ++n_ech_;
                corh.Zeros();
                ech_->Solve(resh, &corh);
                this->iter_ += ech_->iter();
            }
            x->Axpy(corh, 1.0);
            Ah_->VectorMult(*x, &resh);
            resh.ScaleAdd(b, -1.0);
            this->res_ = resh.Norm2();
            res0 = this->res_;
            this->hist_.push_back(this->res_);
        }
        return (this->res_ < abs_ || this->res_ < res0 * rel_) ? kSolverSuccess : kSolverExceeded;
    }

  private:
    MatrixH *Ah_ = nullptr;
    MatrixL *Al_ = nullptr;
    LinearSolver<LADl> *ecl_ = nullptr;
    LinearSolver<LADh> *ech_ = nullptr;
    int max_ech_ = 0, max_ecl_ = 0, n_ech_ = 0, n_ecl_ = 0;
    double abs_ = 0, rel_ = 0, div_ = 0;
    int ecl_max_ = 1000, ech_max_ = 1000;
    DataL ecl_atol_ = DataL(1e-6), ecl_rtol_ = DataL(1e-1), ecl_dtol_ = DataL(1e6);
    DataH ech_atol_ = DataH(1e-12), ech_rtol_ = DataH(1e-12), ech_dtol_ = DataH(1e6);
};

}  // namespace hflmp
