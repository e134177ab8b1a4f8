// Host-side planning of the device layouts for one hierarchy level.
//
// The L forward solve and U back solve of a level (cheb_solve.cc:33-38 and
// :108-115; multi_split_solve.cc:25-53) only couple the eliminated rows
// 0..M-1 within small connected components.  The planner packs whole
// components, together with the kept rows they feed (their L children), into
// workgroup-sized blocks, and orders each block's rows by dependency depth.
// One kernel launch then stages a block's inputs in LDS and runs every depth
// level with a workgroup barrier between levels.  The arithmetic per row
// keeps the reference's operand order, so results are unchanged by the
// schedule.
#pragma once

#include <cstdint>
#include <vector>

namespace mlp {

// Raw arrays of one level (pointers into host memory).
struct LevelArrays {
    uint32_t N = 0, M = 0, Md = 0;
    const uint32_t* map = nullptr;                                      // N
    const uint32_t *L_ia = nullptr, *L_ja = nullptr; const double* L_a = nullptr;   // N-Md rows
    const uint32_t *U_ia = nullptr, *U_ja = nullptr; const double* U_a = nullptr;   // M rows
    const double* dval = nullptr;                                       // Md
};

constexpr uint32_t kGlobalRef = 0x80000000u;   // refers to a kept row (index into the next level)
constexpr uint32_t kBlkFields = 8;

// Component-blocked program of one level.
//
// Nodes are the eliminated rows (i < M) and the kept rows with L entries
// ("roots", M <= i < N-Md).  Rows that depend on each other form small
// connected components; whole components are packed into blocks of at most
// `cap` nodes.  Inside a block the nodes are ordered by L depth (a node's L
// columns lie in earlier depth levels), and the eliminated rows are listed a
// second time ordered by U depth.  A block runs its depth levels with a
// barrier between levels.
// Block b: blk[8b] = first node, blk[8b+1] = nodes, blk[8b+2..3] = its L
// depth levels in lvlL (local end offsets), blk[8b+4..5] = its U levels in
// lvlU (local end offsets into its U list), blk[8b+6] / blk[8b+7] = first L /
// U entry; ublk[2b] / ublk[2b+1] = first U-list position and length.
struct LevelProgram {
    uint32_t M = 0, n = 0, Md = 0, N = 0;
    uint32_t nblk = 0, nodes = 0;
    uint32_t max_nodes = 0, max_Lent = 0, max_Uent = 0, max_Unodes = 0, max_levels = 0;
    uint32_t max_depth = 0, max_comp_size = 0;
    std::vector<uint32_t> blk;           // kBlkFields per block
    std::vector<uint32_t> ublk;          // 2 per block
    std::vector<uint32_t> lvlL, lvlU;
    // L order (size nodes)
    std::vector<uint32_t> fidx;          // map[i]: where the row reads the level input
    std::vector<uint32_t> out;           // eliminated: local position; kept row r = i-M: kGlobalRef | r
    std::vector<uint32_t> Lptr;          // nodes+1 (global entry offsets)
    std::vector<uint32_t> Lloc;          // local position of the L column within the block
    std::vector<double> Lval;
    // U order (eliminated rows only, size M)
    std::vector<uint32_t> uloc;          // local position (L order) of the row
    std::vector<uint32_t> uxidx;         // map[i]: where the row writes the level output
    std::vector<uint32_t> Uptr;          // M+1
    std::vector<uint32_t> Ucol;          // local position, or kGlobalRef | (j - M) for kept columns
    std::vector<double> Uval, Udiag;
    // kept rows without L entries (their right-hand side is a plain gather)
    std::vector<uint32_t> flat;          // r = i - M
};

// 64-row slices, column-major inside a slice: entry k of row 64s+t sits at
// (off64[s] + k) * 64 + t.  Padding has col = 0xffffffff and val = 0.
// With a row order `rows`, slot q holds matrix row rows[q] (entries and their
// order unchanged, columns in the original numbering).
struct SellLayout {
    uint32_t n = 0, nslices = 0;
    uint64_t nent = 0;
    std::vector<uint32_t> off64;          // nslices+1
    std::vector<uint32_t> col;
    std::vector<double> val;
};

void plan_level(const LevelArrays& lv, uint32_t cap, LevelProgram& out);
void plan_sell(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, SellLayout& out,
               const uint32_t* rows = nullptr);
// The order in which the level kernels read the rows of a level's input:
// block nodes (fidx), kept rows without L entries, then the isolated rows.
std::vector<uint32_t> consumption_order(const LevelProgram& P, const uint32_t* map);

}  // namespace mlp
