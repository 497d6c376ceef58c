// solve_capi.cc -- C entry points that run the LinearSolver templates of
// solvers.h on the HIP platform, assembling the system the way the
// reference applications do: host CSR -> device CoupledMatrix, right-hand
// side and start vector copied to the device, Solve, solution copied back
// (examples/convection_diffusion/convdiff_benchmark.cc:589-634).
#include <cstring>
#include <vector>

#include "../../include/hflmp.h"
#include "solvers.h"

using namespace hflmp;

namespace {

int fail(const Error &e) {
    hflmp_set_error(e.what());
    return e.code;
}

template <class LAD>
int run(int solver, int prec, int size_basis, double omega, int n, const int *row, const int *col, const double *val,
        const double *b, double *x, const hfl_solve_params *prm, hfl_solve_result *out) {
    typedef typename LAD::DataType T;
    CPU_CSR_lMatrix<T> host;
    host.Init(row[n], n, n, "A");
    host.row.assign(row, row + n + 1);
    host.col.assign(col, col + row[n]);
    for (int j = 0; j < row[n]; ++j) host.val[j] = static_cast<T>(val[j]);
    std::vector<T> bh(b, b + n), xh(x, x + n);
    std::vector<int> perm;
    PreconditionerBlockJacobiStand<LAD> bj;
    PreconditionerMultiColoring<LAD> mc;
    PreconditionerIdentity<LAD> id;
    Preconditioner<LAD> *P = nullptr;
    if (prec == HFL_PREC_MCSGS) {
        // Preprocess: colour, then renumber the system (PermuteDofNumbering)
        mc.Init_SymmetricGaussSeidel();
        mc.SetRelax(static_cast<T>(omega));
        perm = mc.Preprocess(host);
        host.Reorder(perm.data());
        std::vector<T> bp(n), xp(n);
        for (int i = 0; i < n; ++i) {
            bp[i] = bh[perm[i]];
            xp[i] = xh[perm[i]];
        }
        bh.swap(bp);
        xh.swap(xp);
        P = &mc;
    } else if (prec == HFL_PREC_JACOBI || prec == HFL_PREC_SGS || prec == HFL_PREC_ILU0) {
        if (prec == HFL_PREC_JACOBI) bj.Init_Jacobi(typename LAD::VectorType());
        else if (prec == HFL_PREC_SGS) bj.Init_SymmetricGaussSeidel();
        else bj.Init_ILUp(0);
        P = &bj;
    } else if (prec != HFL_PREC_NONE) {
        throw Error(HFH_ERR_INVALID, "hfl_lmp_solve: unknown preconditioner");
    }
    typename LAD::MatrixType A;
    A.InitFrom(host);
    typename LAD::VectorType vb, vx;
    vb.Init(n);
    vx.Init(n);
    vb.interior().SetBlockValues(0, n, bh.data());
    vx.interior().SetBlockValues(0, n, xh.data());
    std::unique_ptr<LinearSolver<LAD>> S;
    if (solver == HFL_SOLVER_CG) {
        CG<LAD> *cg = new CG<LAD>();
        cg->InitParameter(P ? "Preconditioning" : "NoPreconditioning");
        S.reset(cg);
    } else if (solver == HFL_SOLVER_GMRES || solver == HFL_SOLVER_FGMRES) {
        // without a preconditioner the reference's NoPreconditioning path
        // (gmres.cc:103-264) runs
        const char *method = P ? "RightPreconditioning" : "NoPreconditioning";
        if (solver == HFL_SOLVER_GMRES) {
            GMRES<LAD> *g = new GMRES<LAD>();
            g->InitParameter(size_basis, method);
            S.reset(g);
        } else {
            FGMRES<LAD> *g = new FGMRES<LAD>();
            g->InitParameter(size_basis, method);
            S.reset(g);
        }
    } else {
        throw Error(HFH_ERR_INVALID, "hfl_lmp_solve: unknown solver");
    }
    S->InitControl(prm->maxits, prm->atol, prm->rtol, prm->dtol);
    S->SetupOperator(A);
    if (P) {
        P->SetupOperator(A);
        S->SetupPreconditioner(*P);
        if (solver == HFL_SOLVER_CG) static_cast<CG<LAD> *>(S.get())->InitParameter("Preconditioning");
    }
    const LinearSolverState st = S->Solve(vb, &vx);
    vx.interior().GetBlockValues(0, n, xh.data());
    if (!perm.empty())
        for (int i = 0; i < n; ++i) x[perm[i]] = xh[i];
    else
        for (int i = 0; i < n; ++i) x[i] = xh[i];
    out->iters = S->iter();
    out->ret = st;
    out->state = S->control().status();
    out->res = S->res();
    const std::vector<double> &h = S->history();
    out->nhist = static_cast<int>(h.size());
    if (out->hist)
        for (int i = 0; i < out->nhist && i < out->hist_cap; ++i) out->hist[i] = h[i];
    return HFH_OK;
}

}  // namespace

extern "C" int hfl_lmp_solve(int solver, int precond, int size_basis, double omega, int single, int n,
                             const int *row, const int *col, const double *val, const double *b, double *x,
                             const hfl_solve_params *prm, hfl_solve_result *out) {
    if (n <= 0 || !row || !col || !val || !b || !x || !prm || !out) {
        hflmp_set_error("hfl_lmp_solve: bad argument");
        return HFH_ERR_INVALID;
    }
    try {
        init_platform(0);
        if (single)
            return run<LADescriptorHipS>(solver, precond, size_basis, omega, n, row, col, val, b, x, prm, out);
        return run<LADescriptorHipD>(solver, precond, size_basis, omega, n, row, col, val, b, x, prm, out);
    } catch (const Error &e) {
        return fail(e);
    } catch (const std::exception &e) {
        hflmp_set_error(e.what());
        return HFH_ERR_INVALID;
    }
}

extern "C" int hfl_lmp_mpir(int precond_low, int n, const int *row, const int *col, const double *val,
                            const double *b, double *x, int max_ecl, double abs_tol, double rel_tol, double div_tol,
                            const hfl_solve_params *ecl, hfl_solve_result *out) {
    if (n <= 0 || !row || !col || !val || !b || !x || !ecl || !out) {
        hflmp_set_error("hfl_lmp_mpir: bad argument");
        return HFH_ERR_INVALID;
    }
    try {
        init_platform(0);
        CPU_CSR_lMatrix<double> host;
        host.Init(row[n], n, n, "A");
        host.row.assign(row, row + n + 1);
        host.col.assign(col, col + row[n]);
        host.val.assign(val, val + row[n]);
        CoupledMatrix<double> Ah;
        Ah.InitFrom(host);
        CoupledMatrix<float> Al;
        Al.CastFrom(Ah);
        CoupledVector<double> vb, vx;
        vb.Init(n);
        vx.Init(n);
        vb.interior().SetBlockValues(0, n, b);
        vx.interior().SetBlockValues(0, n, x);
        CG<LADescriptorHipS> cgl;
        PreconditionerBlockJacobiStand<LADescriptorHipS> pl;
        if (precond_low == HFL_PREC_JACOBI) {
            pl.Init_Jacobi(CoupledVector<float>());
            pl.SetupOperator(Al);
            cgl.SetupPreconditioner(pl);
            cgl.InitParameter("Preconditioning");
        } else {
            cgl.InitParameter("NoPreconditioning");
        }
        MPIR<LADescriptorHipD, LADescriptorHipS> mpir;
        mpir.SetHighPrecMatrix(Ah);
        mpir.SetLowPrecMatrix(Al);
        mpir.SetLowPrecErrCorSolver(cgl);
        mpir.SetMpirParameters(0, max_ecl, abs_tol, rel_tol, div_tol);
        mpir.SetLowPrecErrCorSolverParameters(ecl->maxits, ecl->atol, ecl->rtol, ecl->dtol);
        const LinearSolverState st = mpir.Solve(vb, &vx);
        vx.interior().GetBlockValues(0, n, x);
        out->iters = mpir.iter();
        out->ret = st;
        out->state = mpir.n_ecl();
        out->res = mpir.res();
        const std::vector<double> &h = mpir.history();
        out->nhist = static_cast<int>(h.size());
        if (out->hist)
            for (int i = 0; i < out->nhist && i < out->hist_cap; ++i) out->hist[i] = h[i];
        return HFH_OK;
    } catch (const Error &e) {
        return fail(e);
    } catch (const std::exception &e) {
        hflmp_set_error(e.what());
        return HFH_ERR_INVALID;
    }
}
