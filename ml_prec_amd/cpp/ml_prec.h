// C++ drop-in API: the reference's matrix, vector, preconditioner, solver and
// config classes with the same names and signatures, running the solve path
// on the GPU through the C ABI (include/mlp.h).
//
//   reference                                  here
//   Vector = uvector<double>                   modules/vector/vector.h:9
//   CSRMatrix / SkylineMatrix                  modules/matrix/matrix.h:187-291
//   multiply / residual / dump / load          modules/matrix/csrmatrix.cc:172-355
//   daxpy / dscal / ddot / dnrm2               modules/vector/vector.h:13-22
//   PrecBase::solve(Vector& f, Vector& x)      modules/prec/prec_base.h:8-12
//   Prec(A, cfg)  ("uh_cheb")                  modules/prec/cheb/cheb_prec.h:129
//   MultiSplitPrec(A, cfg)                     modules/prec/multi_split/multi_split_prec.h:105
//   DiagPrec / IdPrec                          modules/prec/diag/diag_prec.h, id/id_prec.h
//   GSPrec / CompositePrec / SymPrec           modules/prec/gs, composite, sym
//   Config                                     project/config.h:54-86
//   SolverStats, PCG/GMRES/Cheb/SimpleSolver   modules/solvers/solvers.h:15-38
//   ConfigFile                                 modules/config_file/config_file.h:54-116
//
// Matrices and preconditioners upload themselves to device 0 on first use and
// keep the device copy; the solvers run entirely on the device and
// copy x back.  Errors surface as Exception, like the reference's
// THROW_EXCEPTION.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "config_file.h"

extern "C" {
#include "../../include/mlp.h"
}

typedef unsigned int uint;

class Exception : public std::exception {
public:
    explicit Exception(const std::string& msg) : msg_(msg) {}
    const char* what() const noexcept override { return msg_.c_str(); }
private:
    std::string msg_;
};

#define THROW_EXCEPTION(msg)                                                          \
    do {                                                                              \
        std::ostringstream os_;                                                       \
        os_ << __FILE__ << ":" << __LINE__ << ": " << msg;                            \
        throw Exception(os_.str());                                                   \
    } while (0)

// malloc-backed vector with a uint size (include/uvector.h:19-194)
template <typename T> class uvector {
public:
    typedef T value_type;
    typedef uint size_type;
    uvector(size_type n = 0) : v_(n) {}
    uvector(size_type n, const T& val) : v_(n, val) {}
    size_type size() const { return size_type(v_.size()); }
    bool empty() const { return v_.empty(); }
    T& operator[](size_type i) { return v_[i]; }
    const T& operator[](size_type i) const { return v_[i]; }
    void resize(size_type n) { v_.resize(n); }
    void reserve(size_type n) { v_.reserve(n); }
    void push_back(const T& x) { v_.push_back(x); }
    void clear() { v_.clear(); }
    void swap(uvector& o) { v_.swap(o.v_); }
    T& back() { return v_.back(); }
    const T& back() const { return v_.back(); }
    T* data() { return v_.data(); }
    const T* data() const { return v_.data(); }
    typename std::vector<T>::iterator begin() { return v_.begin(); }
    typename std::vector<T>::iterator end() { return v_.end(); }
    typename std::vector<T>::const_iterator begin() const { return v_.begin(); }
    typename std::vector<T>::const_iterator end() const { return v_.end(); }
private:
    std::vector<T> v_;
};

typedef uvector<double> Vector;

enum DumpType { ASCII, BINARY, HYPRE, MATRIX_MARKET };
enum NormType { NORM_L2, NORM_A, NORM_B_1 };

// ---------------------------------------------------------------- config ---
enum SolverType { PCG_SOLVER, CHEB_SOLVER, SIMPLE_SOLVER, GMRES_SOLVER };
enum PrecType { AMG_PREC, DIAG_PREC, COMP_PREC, GS_PREC, ID_PREC, SYM_SPLIT_PREC, MULTI_SPLIT_PREC, UH_CHEB_PREC };

struct Config {
    std::vector<uint> niters{2};           // defaults of project/main_helper.cc:62-79
    std::vector<double> sigmas{3};
    double c = 1.;
    uint nx = 60, ny = 220, nz = 85;
    uint ntests = 1;
    bool use_tails = true;
    bool optimize_storage = true;
    bool nonsym_matrix = false;
    double nonsym_shift = 0.1;
    SolverType solver = PCG_SOLVER;
    PrecType prec = UH_CHEB_PREC;
    uint coarse_n = 0;
    uint max_levels = 30;
    bool dump_data = false;
    std::string prec_conf_file;
    std::string dir = "results/";
    std::string matrix, vector;
    // GPU-only: precision of the preconditioner hierarchy (MLP_FP64 or MLP_FP32)
    int gpu_precision = MLP_FP64;
};

// ---------------------------------------------------------------- device ---
namespace mlp_dropin {
mlp_ctx context();                         // process-wide context on device 0
void check(int rc);                        // throws Exception with mlp_last_error()
}

// ---------------------------------------------------------------- matrix ---
class CSRMatrix {
public:
    CSRMatrix() : nrow(0), ncol(0) {}
    virtual ~CSRMatrix();
    CSRMatrix(const CSRMatrix& o) : nrow(o.nrow), ncol(o.ncol), ia(o.ia), ja(o.ja), a(o.a) {}
    const CSRMatrix& operator=(const CSRMatrix& o);

    uint rows() const { return nrow; }
    uint cols() const { return ncol; }
    uint size() const { return nrow; }
    void set_size(uint r, uint c) { nrow = r; ncol = c; invalidate(); }
    uvector<uint>& get_ia() { invalidate(); return ia; }
    uvector<uint>& get_ja() { invalidate(); return ja; }
    uvector<double>& get_a() { invalidate(); return a; }
    const uvector<uint>& get_ia() const { return ia; }
    const uvector<uint>& get_ja() const { return ja; }
    const uvector<double>& get_a() const { return a; }
    virtual uint nnz() const { return (nrow && ncol) ? ia.back() : 0; }
    double operator()(uint i, uint j) const;
    virtual void load(const std::string& filename, DumpType type = BINARY);

    // device copy (uploaded on first use, dropped when the host data may change)
    mlp_mat device() const;

    friend void multiply(const CSRMatrix& A, const Vector& v, Vector& res, char type);

protected:
    uint nrow, ncol;
    uvector<uint> ia, ja;
    uvector<double> a;
    mutable mlp_mat dev_ = nullptr;
    void invalidate() const;
};

class SkylineMatrix : public CSRMatrix {
public:
    SkylineMatrix() {}
    // CSR -> Skyline: move each row's diagonal to the front (skyline.cc:30-58)
    void load(const std::string& filename, DumpType type, bool transform);
    void load(const std::string& filename, DumpType type = BINARY) override { load(filename, type, true); }
    void optimize_storage(char type = 's') { (void)type; }   // the device layout keeps full rows
};

void multiply(const CSRMatrix& A, const Vector& v, Vector& res, char type = 'o');
void residual(const CSRMatrix& A, const Vector& b, const Vector& x, Vector& r);
void dump(const std::string& filename, const CSRMatrix& A, DumpType type);

void daxpy(double alpha, const Vector& x, Vector& y);
void dscal(double alpha, Vector& x);
double ddot(const Vector& x, const Vector& y);
double dnrm2(const Vector& x);

// ----------------------------------------------------------- preconditioners
class PrecBase {
public:
    virtual ~PrecBase() {}
    virtual void solve(Vector& f, Vector& x) const;
    virtual mlp_prec device() const = 0;
    // an FP32 hierarchy is an inexact preconditioner: the solvers then use
    // flexible PCG with periodic true-residual replacement
    int precision() const { return precision_; }
protected:
    int precision_ = MLP_FP64;
};

class Prec : public PrecBase {                       // "uh_cheb"
public:
    Prec(const SkylineMatrix& A, const Config& cfg);
    ~Prec() override;
    double lmin() const { return lmin_; }
    double lmax() const { return lmax_; }
    uint nlevels() const { return nlevels_; }
    mlp_prec device() const override { return prec_; }
private:
    [[maybe_unused]] const SkylineMatrix& level0_A;   // kept by reference, as cheb_prec.h:45
    mlp_hier hier_ = nullptr;
    mlp_prec prec_ = nullptr;
    double lmin_ = 0, lmax_ = 0;
    uint nlevels_ = 0;
};

class MultiSplitPrec : public PrecBase {
public:
    MultiSplitPrec(const SkylineMatrix& A, const Config& cfg);
    ~MultiSplitPrec() override;
    uint nlevels() const { return nlevels_; }
    mlp_prec device() const override { return prec_; }
private:
    [[maybe_unused]] const SkylineMatrix& level0_A;   // kept by reference, as multi_split_prec.h:70
    mlp_hier hier_ = nullptr;
    mlp_prec prec_ = nullptr;
    uint nlevels_ = 0;
};

class DiagPrec : public PrecBase {
public:
    explicit DiagPrec(const SkylineMatrix& A);
    ~DiagPrec() override;
    mlp_prec device() const override { return prec_; }
private:
    mlp_prec prec_ = nullptr;
};

class IdPrec : public PrecBase {
public:
    explicit IdPrec(const SkylineMatrix& A);
    ~IdPrec() override;
    mlp_prec device() const override { return prec_; }
private:
    mlp_prec prec_ = nullptr;
};

class GSPrec : public PrecBase {                     // modules/prec/gs/gs_prec.h
public:
    explicit GSPrec(const SkylineMatrix& A);
    ~GSPrec() override;
    mlp_prec device() const override { return prec_; }
private:
    mlp_prec prec_ = nullptr;
};

class CompositePrec : public PrecBase {              // GSPrec then MultiSplitPrec, composite_prec.h
public:
    CompositePrec(const SkylineMatrix& A, const Config& cfg);
    ~CompositePrec() override;
    mlp_prec device() const override { return prec_; }
private:
    mlp_hier hier_ = nullptr;
    mlp_prec prec_ = nullptr;
};

class SymPrec : public PrecBase {                    // "sym_split", sym_prec.h
public:
    SymPrec(const SkylineMatrix& A, const Config& cfg);
    ~SymPrec() override;
    void set_eps(double eps);
    mlp_prec device() const override { return prec_; }
private:
    mlp_prec prec_ = nullptr;
};

// ----------------------------------------------------------------- solvers --
struct SolverStats {
    double t_resid = 0, t_resid_total = 0, t_prec = 0, t_prec_total = 0, t_const = 0, t_sol = 0;
    uint niter = 0;
    double true_rel_res = 0;     // GPU addition: ||b - A x|| / ||b - A x0|| in FP64
};
std::ostream& operator<<(std::ostream& os, const SolverStats& stats);

void PCGSolver(const CSRMatrix& A, const Vector& b, const PrecBase& B, Vector& x, SolverStats& stats,
               double eps = 1e-10, NormType norm_type = NORM_L2, bool silent = false);
void GMRESSolver(const CSRMatrix& A, const Vector& b, const PrecBase& B, Vector& x, SolverStats& stats,
                 double eps = 1e-10, NormType norm_type = NORM_L2, bool silent = false);
void ChebSolver(const CSRMatrix& A, double lmin, double lmax, const Vector& b, const PrecBase& B, Vector& x,
                double eps = 1e-10, NormType norm_type = NORM_L2, bool silent = false);
void SimpleSolver(const CSRMatrix& A, const Vector& b, const PrecBase& B, Vector& x, SolverStats& stats,
                  double eps = 1e-10, NormType norm_type = NORM_L2, bool silent = false);
