// Implementation of the C++ drop-in API (ml_prec.h) on top of the C ABI.
#include "ml_prec.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <mutex>

namespace mlp_dropin {

mlp_ctx context() {
    static mlp_ctx ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        int rc = mlp_ctx_create(0, &ctx);
        if (rc) throw Exception(std::string("mlp_ctx_create: ") + mlp_last_error());
    });
    return ctx;
}

void check(int rc) {
    if (rc) throw Exception(mlp_last_error());
}

// a device buffer of doubles for copying host Vectors in and out
struct DevVec {
    double* d = nullptr;
    uint n = 0;
    explicit DevVec(uint n_) : n(n_) { check(mlp_dev_alloc(context(), std::max<size_t>(n, 1) * 8, (void**)&d)); }
    DevVec(const Vector& v) : DevVec(v.size()) { put(v); }
    ~DevVec() { if (d) mlp_dev_free(context(), d); }
    void put(const Vector& v) { check(mlp_copy_h2d(context(), d, v.data(), size_t(n) * 8)); }
    void get(Vector& v) const {
        v.resize(n);
        check(mlp_copy_d2h(context(), v.data(), d, size_t(n) * 8));
    }
    DevVec(const DevVec&) = delete;
    DevVec& operator=(const DevVec&) = delete;
};

}  // namespace mlp_dropin

using namespace mlp_dropin;

// ---------------------------------------------------------------- matrix ---
CSRMatrix::~CSRMatrix() { invalidate(); }

const CSRMatrix& CSRMatrix::operator=(const CSRMatrix& o) {
    invalidate();
    nrow = o.nrow; ncol = o.ncol; ia = o.ia; ja = o.ja; a = o.a;
    return *this;
}

void CSRMatrix::invalidate() const {
    if (dev_) {
        mlp_mat_destroy(dev_);
        dev_ = nullptr;
    }
}

mlp_mat CSRMatrix::device() const {
    if (!dev_) check(mlp_csr_upload(context(), nrow, ia.data(), ja.data(), a.data(), &dev_));
    return dev_;
}

double CSRMatrix::operator()(uint i, uint j) const {
    for (uint p = ia[i]; p < ia[i + 1]; p++)
        if (ja[p] == j) return a[p];
    return 0;
}

// CSR BINARY / ASCII / MatrixMarket readers (csrmatrix.cc:172-246)
void CSRMatrix::load(const std::string& filename, DumpType type) {
    invalidate();
    if (type == BINARY) {
        std::ifstream is(filename.c_str(), std::ifstream::binary);
        if (!is.good()) THROW_EXCEPTION("Problem reading file \"" << filename << "\"");
        is.read(reinterpret_cast<char*>(&nrow), sizeof(uint));
        is.read(reinterpret_cast<char*>(&ncol), sizeof(uint));
        ia.resize(nrow + 1);
        is.read(reinterpret_cast<char*>(ia.data()), (nrow + 1) * sizeof(uint));
        uint nz = ia[nrow];
        ja.resize(nz);
        a.resize(nz);
        is.read(reinterpret_cast<char*>(ja.data()), nz * sizeof(uint));
        is.read(reinterpret_cast<char*>(a.data()), nz * sizeof(double));
        if (!is) THROW_EXCEPTION("Truncated matrix file \"" << filename << "\"");
    } else if (type == ASCII) {
        std::ifstream is(filename.c_str());
        if (!is.good()) THROW_EXCEPTION("Problem reading file \"" << filename << "\"");
        std::string line;
        uint nz = 0;
        std::getline(is, line);
        is >> nrow >> ncol >> nz;
        std::getline(is, line);
        std::getline(is, line);
        ia.resize(nrow + 1);
        for (uint i = 0; i <= nrow; i++) is >> ia[i];
        std::getline(is, line);
        std::getline(is, line);
        ja.resize(nz);
        for (uint i = 0; i < nz; i++) is >> ja[i];
        std::getline(is, line);
        std::getline(is, line);
        a.resize(nz);
        for (uint i = 0; i < nz; i++) is >> a[i];
    } else if (type == MATRIX_MARKET) {
        std::ifstream is(filename.c_str());
        if (!is.good()) THROW_EXCEPTION("Problem reading file \"" << filename << "\"");
        std::string line;
        std::getline(is, line);
        std::getline(is, line);
        uint nz = 0;
        is >> nrow >> ncol >> nz;
        std::vector<std::vector<std::pair<uint, double>>> rowsv(nrow);
        for (uint k = 0; k < nz; k++) {
            uint i, j;
            double v;
            is >> i >> j >> v;
            auto& r = rowsv[i - 1];
            auto it = std::lower_bound(r.begin(), r.end(), std::make_pair(j - 1, -HUGE_VAL));
            if (it != r.end() && it->first == j - 1) it->second = v;     // later entries win, as in a map
            else r.insert(it, std::make_pair(j - 1, v));
        }
        ia.resize(nrow + 1);
        ja.clear();
        a.clear();
        ia[0] = 0;
        for (uint i = 0; i < nrow; i++) {
            for (auto& e : rowsv[i]) { ja.push_back(e.first); a.push_back(e.second); }
            ia[i + 1] = ja.size();
        }
    } else {
        THROW_EXCEPTION("Loading format HYPRE is not applicable to CSRMatrix");
    }
}

void SkylineMatrix::load(const std::string& filename, DumpType type, bool transform) {
    CSRMatrix::load(filename, type);
    if (!transform) return;
    for (uint i = 0; i < nrow; i++) {
        uint start = ia[i], end = ia[i + 1], d = end;
        for (uint j = start; j < end; j++)
            if (ja[j] == i) { d = j; break; }
        if (d == end) THROW_EXCEPTION("No diagonal element in row " << i);
        double v = a[d];
        for (uint j = d; j > start; j--) { ja[j] = ja[j - 1]; a[j] = a[j - 1]; }
        ja[start] = i;
        a[start] = v;
    }
}

void dump(const std::string& filename, const CSRMatrix& A, DumpType type) {
    const uvector<uint>& ia = A.get_ia();
    const uvector<uint>& ja = A.get_ja();
    const uvector<double>& a = A.get_a();
    uint m = A.rows(), n = A.cols(), nz = A.nnz();
    if (type == BINARY) {
        std::ofstream os(filename.c_str(), std::ofstream::binary);
        if (!os.good()) THROW_EXCEPTION("Could not open \"" << filename << "\" for write");
        os.write(reinterpret_cast<const char*>(&m), sizeof(uint));
        os.write(reinterpret_cast<const char*>(&n), sizeof(uint));
        os.write(reinterpret_cast<const char*>(ia.data()), (m + 1) * sizeof(uint));
        os.write(reinterpret_cast<const char*>(ja.data()), nz * sizeof(uint));
        os.write(reinterpret_cast<const char*>(a.data()), nz * sizeof(double));
    } else if (type == ASCII) {
        std::ofstream os(filename.c_str());
        os << "# rows cols nonzeros" << std::endl << m << " " << n << " " << nz << std::endl;
        os << "# row_ptr" << std::endl;
        for (uint i = 0; i <= m; i++) os << ia[i] << std::endl;
        os << "# col_ind" << std::endl;
        for (uint j = 0; j < nz; j++) os << ja[j] << std::endl;
        os << "# value" << std::endl << std::scientific << std::setprecision(15);
        for (uint j = 0; j < nz; j++) os << a[j] << std::endl;
    } else if (type == MATRIX_MARKET) {
        std::ofstream os(filename.c_str());
        os << std::fixed << std::setprecision(15);
        os << "%%MatrixMarket matrix coordinate real general" << std::endl << m << " " << n << " " << nz << std::endl;
        for (uint i = 0; i < m; i++)
            for (uint j = ia[i]; j < ia[i + 1]; j++) os << i + 1 << " " << ja[j] + 1 << " " << a[j] << std::endl;
    } else if (type == HYPRE) {
        // IJ format of a Skyline matrix: "ilower iupper jlower jupper", then "i j a" per entry in storage order
        if (!dynamic_cast<const SkylineMatrix*>(&A))
            THROW_EXCEPTION("Dumping matrix in HYPRE format is not supported for CSRMatrix");
        std::ofstream os(filename.c_str());
        if (!os.good()) THROW_EXCEPTION("Could not open \"" << filename << "\" for write");
        os << "0 " << m - 1 << " 0 " << n - 1 << std::endl << std::scientific << std::setprecision(15);
        for (uint i = 0; i < m; i++)
            for (uint j = ia[i]; j < ia[i + 1]; j++) os << i << " " << ja[j] << " " << a[j] << std::endl;
    } else {
        THROW_EXCEPTION("Unknown dump type");
    }
}

void multiply(const CSRMatrix& A, const Vector& v, Vector& res, char type) {
    if (type != 'o' && type != 's') THROW_EXCEPTION("Only A*v ('o') and symmetric ('s') products run on the device");
    if (v.size() != A.cols() || A.rows() == 0) THROW_EXCEPTION("Different sizes");
    DevVec dv(v), dr(A.rows());
    check(mlp_spmv(context(), A.device(), dv.d, dr.d));
    dr.get(res);
}

void residual(const CSRMatrix& A, const Vector& b, const Vector& x, Vector& r) {
    if (b.size() != A.size() || x.size() != A.size()) THROW_EXCEPTION("Wrong sizes");
    DevVec db(b), dx(x), dr(A.size());
    check(mlp_residual(context(), A.device(), db.d, dx.d, dr.d));
    dr.get(r);
}

void daxpy(double alpha, const Vector& x, Vector& y) {
    if (x.size() != y.size()) THROW_EXCEPTION("Different sizes: x (" << x.size() << "), y (" << y.size() << ")");
    if (!x.size()) return;
    DevVec dx(x), dy(y);
    check(mlp_daxpy(context(), x.size(), alpha, dx.d, dy.d));
    dy.get(y);
}

void dscal(double alpha, Vector& x) {
    if (!x.size()) return;
    DevVec dx(x);
    check(mlp_dscal(context(), x.size(), alpha, dx.d));
    dx.get(x);
}

double ddot(const Vector& x, const Vector& y) {
    if (x.size() != y.size()) THROW_EXCEPTION("Different sizes: x (" << x.size() << "), y (" << y.size() << ")");
    DevVec dx(x), dy(y);
    double r = 0;
    check(mlp_ddot(context(), x.size(), dx.d, dy.d, &r));
    return r;
}

double dnrm2(const Vector& x) {
    DevVec dx(x);
    double r = 0;
    check(mlp_dnrm2(context(), x.size(), dx.d, &r));
    return r;
}

// ----------------------------------------------------------- preconditioners
void PrecBase::solve(Vector& f, Vector& x) const {
    DevVec df(f), dx(f.size());
    check(mlp_prec_apply(context(), device(), df.d, dx.d));
    dx.get(x);
}

Prec::Prec(const SkylineMatrix& A, const Config& cfg) : level0_A(A) {
    for (double s : cfg.sigmas)
        if (s <= 1) THROW_EXCEPTION("All sigmas must be > 1");
    check(mlp_hier_build_uh_cheb(A.size(), A.get_ia().data(), A.get_ja().data(), A.get_a().data(),
                                 cfg.sigmas.data(), uint32_t(cfg.sigmas.size()), cfg.niters.data(),
                                 uint32_t(cfg.niters.size()), cfg.use_tails, &hier_));
    check(mlp_hier_nlevels(hier_, &nlevels_));
    mlp_level_info li;
    check(mlp_hier_level_info(hier_, 0, &li));
    lmin_ = li.lmin;
    lmax_ = li.lmax;
    check(mlp_prec_from_hier(context(), hier_, A.device(), cfg.gpu_precision, &prec_));
    precision_ = cfg.gpu_precision;
}

Prec::~Prec() {
    if (prec_) mlp_prec_destroy(prec_);
    if (hier_) mlp_hier_destroy(hier_);
}

// MultiSplitPrec's hierarchy from a Config (with its prec config file, if any)
static mlp_hier build_msplit(const SkylineMatrix& A, const Config& cfg) {
    // without a config file the command-line values apply (multi_split_constr.cc:234-244)
    mlp_msplit_params p;
    std::memset(&p, 0, sizeof(p));
    p.q = cfg.sigmas.data();
    p.nq = uint32_t(cfg.sigmas.size());
    p.niters = cfg.niters.data();
    p.nniters = uint32_t(cfg.niters.size());
    static const double eps_default = 1e-2;
    p.eps = &eps_default;
    p.neps = 1;
    p.inner_dynamic = 0;
    p.order = cfg.use_tails ? MLP_ORDER_ORIGINAL : MLP_ORDER_NONE;
    p.use_tails = cfg.use_tails;
    p.degree_max = 3;
    p.max_levels = cfg.max_levels;
    p.coarse_n = cfg.coarse_n;
    const char* path = cfg.prec_conf_file.empty() ? nullptr : cfg.prec_conf_file.c_str();
    mlp_hier h = nullptr;
    check(mlp_hier_build_multi_split(A.size(), A.get_ia().data(), A.get_ja().data(), A.get_a().data(), path, &p,
                                     &h));
    return h;
}

MultiSplitPrec::MultiSplitPrec(const SkylineMatrix& A, const Config& cfg) : level0_A(A) {
    hier_ = build_msplit(A, cfg);
    check(mlp_hier_nlevels(hier_, &nlevels_));
    check(mlp_prec_from_hier(context(), hier_, A.device(), cfg.gpu_precision, &prec_));
    precision_ = cfg.gpu_precision;
}

MultiSplitPrec::~MultiSplitPrec() {
    if (prec_) mlp_prec_destroy(prec_);
    if (hier_) mlp_hier_destroy(hier_);
}

DiagPrec::DiagPrec(const SkylineMatrix& A) { check(mlp_prec_simple(context(), MLP_PREC_DIAG, A.device(), &prec_)); }
DiagPrec::~DiagPrec() { if (prec_) mlp_prec_destroy(prec_); }
IdPrec::IdPrec(const SkylineMatrix& A) { check(mlp_prec_simple(context(), MLP_PREC_ID, A.device(), &prec_)); }
IdPrec::~IdPrec() { if (prec_) mlp_prec_destroy(prec_); }

GSPrec::GSPrec(const SkylineMatrix& A) { check(mlp_prec_gs(context(), A.device(), &prec_)); }
GSPrec::~GSPrec() { if (prec_) mlp_prec_destroy(prec_); }

CompositePrec::CompositePrec(const SkylineMatrix& A, const Config& cfg) {
    hier_ = build_msplit(A, cfg);
    check(mlp_prec_composite(context(), A.device(), hier_, &prec_));
}
CompositePrec::~CompositePrec() {
    if (prec_) mlp_prec_destroy(prec_);
    if (hier_) mlp_hier_destroy(hier_);
}

SymPrec::SymPrec(const SkylineMatrix& A, const Config& cfg) {
    // the inner preconditioner is Prec(Asym, cfg) with nonsym_matrix = false (sym_prec.cc:33-36)
    check(mlp_prec_sym(context(), A.device(), cfg.sigmas.data(), uint32_t(cfg.sigmas.size()), cfg.niters.data(),
                       uint32_t(cfg.niters.size()), cfg.use_tails, &prec_));
}
SymPrec::~SymPrec() { if (prec_) mlp_prec_destroy(prec_); }
void SymPrec::set_eps(double eps) { check(mlp_prec_sym_set_eps(prec_, eps)); }

// ----------------------------------------------------------------- solvers --
std::ostream& operator<<(std::ostream& os, const SolverStats& stats) {
    os << std::fixed << std::setprecision(3) << std::endl;
    os << "Residual time       :  avg = " << stats.t_resid << "\t total = " << stats.t_resid_total << std::endl;
    os << "Prec inversion time :  avg = " << stats.t_prec << "\t total = " << stats.t_prec_total << std::endl;
    if (stats.t_const) os << "Time of (possible) construction: " << stats.t_const << std::endl;
    os << "Number of iter      : " << stats.niter << std::endl;
    os << "Time of solution    : " << stats.t_sol << std::endl;
    return os;
}

enum class SolverKind { PCG, GMRES, CHEB, SIMPLE };

static void run_solver(SolverKind kind, const CSRMatrix& A, const Vector& b, const PrecBase& B, Vector& x,
                       SolverStats& stats, double eps, NormType norm_type, double lmin = 0, double lmax = 0) {
    if (norm_type != NORM_L2) THROW_EXCEPTION("Only the L2 stopping norm runs on the device");
    if (b.size() != A.size() || x.size() != A.size()) THROW_EXCEPTION("Wrong sizes");
    DevVec db(b), dx(A.size());
    mlp_solve_opts o;
    std::memset(&o, 0, sizeof(o));
    o.eps = eps;
    o.x0_mode = MLP_X0_REFERENCE;        // the reference always starts from generate_x0 (pcg.cc:24)
    o.use_graph = 1;
    o.check_every = 8;
    if (kind == SolverKind::PCG && B.precision() == MLP_FP32) {
        o.flexible = 1;
        o.true_res_every = 20;
    }
    mlp_solve_stats st;
    std::memset(&st, 0, sizeof(st));
    int rc = 0;
    switch (kind) {
        case SolverKind::PCG: rc = mlp_pcg(context(), A.device(), B.device(), db.d, dx.d, &o, &st); break;
        case SolverKind::GMRES: rc = mlp_gmres(context(), A.device(), B.device(), db.d, dx.d, &o, &st); break;
        case SolverKind::CHEB:
            rc = mlp_cheb_solver(context(), A.device(), B.device(), lmin, lmax, db.d, dx.d, &o, &st);
            break;
        case SolverKind::SIMPLE: rc = mlp_simple_solver(context(), A.device(), B.device(), db.d, dx.d, &o, &st); break;
    }
    // SimpleSolver computes x before it reports "No need to iterate" (simple_solver.cc:52-82)
    if (rc == MLP_ERR_NO_ITER && kind == SolverKind::SIMPLE) dx.get(x);
    if (rc) THROW_EXCEPTION(mlp_last_error());
    dx.get(x);
    stats.niter = st.niter;
    stats.t_sol = st.t_solve_ms * 1e-3;
    stats.true_rel_res = st.true_rel_res;
}

void PCGSolver(const CSRMatrix& A, const Vector& b, const PrecBase& B, Vector& x, SolverStats& stats, double eps,
               NormType norm_type, bool silent) {
    run_solver(SolverKind::PCG, A, b, B, x, stats, eps, norm_type);
    if (!silent) std::cout << "#" << stats.niter << ": relative -> " << std::scientific << stats.true_rel_res << std::endl;
}

void GMRESSolver(const CSRMatrix& A, const Vector& b, const PrecBase& B, Vector& x, SolverStats& stats, double eps,
                 NormType norm_type, bool silent) {
    run_solver(SolverKind::GMRES, A, b, B, x, stats, eps, norm_type);
    if (!silent) std::cout << "#" << stats.niter << ": relative -> " << std::scientific << stats.true_rel_res << std::endl;
}

void ChebSolver(const CSRMatrix& A, double lmin, double lmax, const Vector& b, const PrecBase& B, Vector& x,
                double eps, NormType norm_type, bool silent) {
    SolverStats stats;
    run_solver(SolverKind::CHEB, A, b, B, x, stats, eps, norm_type, lmin, lmax);
    if (!silent) std::cout << "#" << stats.niter << ": relative -> " << std::scientific << stats.true_rel_res << std::endl;
}

void SimpleSolver(const CSRMatrix& A, const Vector& b, const PrecBase& B, Vector& x, SolverStats& stats, double eps,
                  NormType norm_type, bool silent) {
    run_solver(SolverKind::SIMPLE, A, b, B, x, stats, eps, norm_type);
    if (!silent) std::cout << "#" << stats.niter << ": relative -> " << std::scientific << stats.true_rel_res << std::endl;
}
