// Hierarchy construction for the uh_cheb and multi_split preconditioners, in
// FP64, reproducing the reference's integer decisions (dropped links, tails,
// permutations, sparsity) and floating-point values exactly.  The host builder
// is here; a uh_cheb hierarchy can also be built level by level on a device
// (ChebLevelBuilder) into the same arrays.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "par.h"

namespace mlp {

// Skyline CSR (diagonal first, then ascending columns): modules/matrix/matrix.h:269-291
struct HostCSR {
    uint32_t nrow = 0, ncol = 0;
    HostVec<uint32_t> ia{0}, ja;
    HostVec<double> a;
    uint32_t nnz() const { return ia.empty() ? 0 : ia.back(); }
};

// One level of the hierarchy; fields follow Prec::Level (modules/prec/cheb/cheb_prec.h:17-40)
// and MultiSplitPrec::Level (modules/prec/multi_split/multi_split_prec.h:19-44).
struct HostLevel {
    uint32_t N = 0, nnz = 0, M = 0, Md = 0;
    uint32_t ncheb = 0;          // uh_cheb: Chebyshev steps on the level
    uint32_t niter = 0;          // multi_split: fixed inner iterations
    double alpha = 1.0, beta = 0.0, lmin = 0.0, lmax = 0.0;   // uh_cheb
    double q = 0.0, eps = 0.0;                                 // multi_split
    std::vector<uint32_t> map, rmap;   // permuted -> original, original -> permuted
    HostCSR L;                   // (N-Md) x (N-Md), strictly lower, unit diagonal implied
    HostCSR U;                   // M x (N-Md), Skyline (pivot first)
    std::vector<double> dval;    // Md reciprocals of the isolated-node diagonals
    std::vector<double> aux;     // row budget c (uh_cheb) / scaled c (multi_split)
    HostCSR A;                   // the level matrix (level 0: empty, the caller's matrix is used)
};

enum class PrecKind : int { UH_CHEB = 1, MULTI_SPLIT = 2 };
enum class ElimOrder : int { NONE = 0, ORIGINAL = 1, BLOCK = 2, SIMPLE_1 = 3 };

struct ChebParams {              // Config fields used by Prec (project/config.h:54-86)
    std::vector<double> sigmas{3.0};
    std::vector<uint32_t> niters{2};
    bool use_tails = true;
    bool optimize_storage = true;
};

struct MultiSplitParams {        // ConfigFile keys (modules/prec/multi_split/multi_split_constr.cc:185-228)
    std::vector<double> q{0.8};
    std::vector<uint32_t> niters{5};
    std::vector<double> eps{1e-2};
    bool inner_dynamic = false;
    ElimOrder order = ElimOrder::ORIGINAL;
    bool use_tails = true;
    uint32_t degree_max = 3;
    double level_ratio = 0.0;
    uint32_t coarse_n = 0;
    uint32_t max_levels = 15;
};

struct Hierarchy {
    PrecKind kind = PrecKind::UH_CHEB;
    uint32_t nlevels = 0;
    std::vector<HostLevel> levels;
    const HostCSR* A0 = nullptr;     // level-0 matrix, kept by reference like level0_A
};

// Every level of a uh_cheb hierarchy built somewhere other than the host
// threads (a device): given the level-0 matrix and H with each level's alpha
// and beta set, fill H.levels[l] (N, nnz, M, Md, map, rmap, aux, L, U, dval)
// and H.levels[l + 1].A for every level, and H.nlevels, throwing the host
// builder's errors ("Too many levels: l" past max_levels).  Every array must
// equal the host builder's bit for bit.
struct ChebLevelBuilder {
    virtual ~ChebLevelBuilder() = default;
    virtual void run(const HostCSR& A, bool use_tails, uint32_t max_levels, Hierarchy& H) = 0;
};

// Throws mlp::Error (derived from std::runtime_error) with the reference's messages.
void build_uh_cheb(const HostCSR& A, const ChebParams& p, Hierarchy& H, ChebLevelBuilder* dev = nullptr);
void build_multi_split(const HostCSR& A, const MultiSplitParams& p, Hierarchy& H);

// cheb(x,k): modules/prec/misc/misc.cc:232-242 (host libm for k >= 4)
double cheb_poly(double x, uint32_t k);

// Parse a prec config file (modules/config_file/config_file.cc:68-142) into MultiSplitParams,
// with the keys and defaults of multi_split_constr.cc:185-228.  Returns false when the file is
// absent (the reference then keeps its command-line parameters).
bool read_multi_split_config(const std::string& path, MultiSplitParams& p);

}  // namespace mlp
