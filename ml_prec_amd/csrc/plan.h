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
#include <functional>
#include <memory>
#include <vector>

#include "par.h"

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
constexpr uint32_t kGlobalBlock = 0x80000000u; // header word 1 flag: the block's values live in global memory
constexpr uint32_t kBlkFields = 12;

// Component-blocked program of one level.
//
// Nodes are the eliminated rows (i < M) and the kept rows with L entries
// ("roots", M <= i < N-Md).  Rows that depend on each other form small
// connected components; whole components are packed into blocks of at most
// `cap` nodes, and a component larger than `cap` gets a block of its own.  Inside a block the nodes are ordered by L depth (a node's L
// columns lie in earlier depth levels), and the eliminated rows are listed a
// second time ordered by U depth.  A block runs its depth levels with a
// barrier between levels.
//
// The kernels read the level input in "consumption order": position q runs
// over the block nodes, then the kept rows without L entries (flat), then
// the isolated rows; qrow[q] is the level row read at q.
//
// A block that the `fits` predicate of plan_level rejects (a component too
// large for the workgroup's LDS) keeps its values in global memory instead:
// kGlobalBlock is set in its header word 1, and the LDS maxima (max_nodes,
// ...) cover the other blocks only.
//
// Block b header blk[12b ..]: [0] first node, [1] nodes, [2..3] its L depth
// levels in lvlL (local end offsets), [4..5] its U levels in lvlU, [6..7] its
// L entry range, [8..9] its U entry range, [10] first U-list position,
// [11] U-list length.
struct LevelProgram {
    uint32_t M = 0, n = 0, Md = 0, N = 0;
    uint32_t nblk = 0, nodes = 0;
    uint32_t max_nodes = 0, max_Lent = 0, max_Uent = 0, max_Unodes = 0, max_levels = 0;
    uint32_t max_depth = 0, max_comp_size = 0;
    bool has_global = false;             // some block runs from global memory (the maxima exclude it)
    std::vector<uint32_t> blk;           // kBlkFields per block
    std::vector<uint32_t> lvlL, lvlU;
    std::vector<uint32_t> qrow;          // N: level row read at consumption position q
    // L order (size nodes)
    std::vector<uint32_t> out;           // eliminated: local position; root: kGlobalRef | (F slot of kept row i-M)
    std::vector<uint32_t> Lptr;          // nodes+1 (global entry offsets)
    std::vector<uint32_t> Lloc;          // local position of the L column within the block
    std::vector<double> Lval;
    std::vector<uint32_t> Ldep;          // depth level of the node within its block (0-based)
    // U order (eliminated rows only, size M)
    std::vector<uint32_t> uloc;          // local position (L order) of the row
    std::vector<uint32_t> uxidx;         // map[i]: where the row writes the level output
    std::vector<uint32_t> Udep;          // U depth level of the row within its block (0-based)
    std::vector<uint32_t> Uptr;          // M+1
    std::vector<uint32_t> Ucol;          // local position, or kGlobalRef | (j - M) for kept columns
    std::vector<double> Uval, Udiag;
    // kept rows without L entries (their right-hand side is a plain read)
    std::vector<uint32_t> flat;          // r = i - M
    std::vector<uint32_t> flatq;         // F slot the flat row writes (r, or its position in the next level)
};


// 64-row slices, column-major inside a slice: entry k of row 64s+t sits at
// (off64[s] + k) * 64 + t.  Padding has col = 0xffffffff and val = 0.
// With a row order `rows`, slot q holds matrix row rows[q] (entries and their
// order unchanged, columns in the original numbering).
struct SellLayout {
    uint32_t n = 0, nslices = 0;
    uint64_t nent = 0;
    std::vector<uint32_t> off64;          // nslices+1
    HostVec<uint32_t> col;                // absent when dcol is used
    HostVec<int16_t> dcol;                // column - row, when every entry fits (rows unpermuted); -32768 = padding
    HostVec<double> val;
};

// Whether a block with these sizes (nodes, L entries, U entries, U-list rows,
// depth levels) runs from LDS.
using BlockFits = std::function<bool(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t)>;
void plan_level(const LevelArrays& lv, uint32_t cap, LevelProgram& out, const BlockFits& fits = nullptr);

// Seven-point layout of a symmetric matrix on a full nx*ny*nz grid (every
// dimension >= 2; row r = i + nx*(j + ny*k), neighbours at r -+ 1, r -+ sy,
// r -+ sz with sy = nx, sz = nx*ny): per row the diagonal and the values of
// its three lower neighbours; a row's upper values are its upper neighbours'
// lower values (the matrix is bitwise symmetric).  pres bits: 1 z-, 2 y-,
// 4 x-, 8 x+, 16 y+, 32 z+.
struct StencilLayout {
    uint32_t n = 0, sy = 0, sz = 0;
    HostVec<uint8_t> pres;
    HostVec<double> d, lx, ly, lz;       // lx[r] = a(r, r-1), ly[r] = a(r, r-sy), lz[r] = a(r, r-sz); 0 if absent
};
// True (and S filled) when the Skyline CSR matrix has exactly that structure,
// diagonal first then columns ascending, and a(r, c) == a(c, r) bitwise.
bool plan_stencil(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, StencilLayout& S);
void plan_sell(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, SellLayout& out,
               const uint32_t* rows = nullptr);
// Makes level l write its kept-row right-hand side F in the consumption order
// of level l+1 (F slot of kept row r = position of row r in next.qrow), so
// that level l+1 reads it with plain coalesced loads.
void link_levels(LevelProgram& P, const LevelProgram& next);
// The same for a next level whose consumption order is next_qrow[0..n).
void link_levels(LevelProgram& P, const uint32_t* next_qrow, size_t n);

// Fused program of a uh_cheb level l and the coarsest level l+1 (which keeps
// no rows).  A visit of level l with two Chebyshev steps then takes two
// launches instead of four:
//   K1: level l's L sweep, then, for the kept rows the block produced, step 1
//       on level l+1 (its L sweep, U sweep and diagonal): u1 = B_{l+1} F;
//   K2: step 2 on level l+1 from the residual A_{l+1} (a1 u1) - F, the
//       Chebyshev combine of the kept rows, and level l's U sweep.
// Every level-(l+1) component must lie in one block together with the
// level-l components whose roots (or flat kept rows) it consists of: blocks
// are formed from the connected components of level l's L and U couplings
// joined with level (l+1)'s.  Kept rows of level l that are isolated on
// level l+1 and feed no level-l component ("flat") run outside the blocks.
//
// Block b header pblk[12b ..]: [0] first level-l node, [1] level-l nodes,
// [2] L depth levels, [3] first level-l U-list entry, [4] U-list length,
// [5] U depth levels, [6] first level-(l+1) node, [7] level-(l+1) nodes (the
// block's kept rows), [8] their L depth levels, [9] first level-(l+1) U-list
// entry, [10] its length, [11] its U depth levels.
//
// Consumption orders: level l reads its input at positions [block nodes]
// [flat kept rows][isolated rows] (qrow); level l+1 at [block level-(l+1)
// nodes][flat kept rows] (qrow2), which is also where F is stored.
constexpr uint32_t kPairFields = 12;
constexpr uint32_t kIso2 = 0x80000000u;   // row2 flag: the row is isolated on level l+1
struct PairProgram {
    uint32_t N = 0, M = 0, Md = 0, n = 0;
    uint32_t nblk = 0, nodes = 0, nodes2 = 0, nflat2 = 0, nU = 0, nU2 = 0;
    uint32_t max_nodes = 0, max_nodes2 = 0, max_U = 0, max_U2 = 0;
    std::vector<uint32_t> blk;
    std::vector<uint32_t> qrow, qrow2;
    // level-l nodes: eliminated rows and kept rows of the block, by L depth
    std::vector<uint32_t> Ldep, Lptr, Lloc, kept2;   // kept2: level-(l+1) local index of a kept node, else kNoRef
    std::vector<double> Lval;
    // level-l U list: eliminated rows by U depth; kept columns as kGlobalRef | level-(l+1) local index
    std::vector<uint32_t> Udep, uloc, uxidx, Uptr, Ucol;
    std::vector<double> Uval, Udiag;
    // level-(l+1) nodes: src (level-l local index), row2 (level-(l+1) row, kIso2 when isolated there),
    // xl (level-l output row of the kept row), L' by depth
    std::vector<uint32_t> src2, row2, xl2, Ldep2, Lptr2, Lloc2;
    std::vector<double> Lval2, d2;                   // d2: 1 / diagonal of a row isolated on level l+1
    // level-(l+1) U list (rows eliminated there), by U depth, columns as local indices
    std::vector<uint32_t> Udep2, uloc2, Uptr2, Ucol2;
    std::vector<double> Uval2, Udiag2;
    // flat kept rows (outside the blocks): level-(l+1) row, level-l output row, d2
    std::vector<uint32_t> frow2, fxl;
    std::vector<double> fd2;
};
constexpr uint32_t kNoRef = 0xffffffffu;
// Plans the pair (la = level l, lb = level l+1, which must keep no rows).
// Returns false (P untouched) when a merged component exceeds max_nodes
// level-l nodes, max_nodes level-(l+1) nodes or max_nodes U-list entries.
bool plan_pair(const LevelArrays& la, const LevelArrays& lb, uint32_t max_nodes, PairProgram& P);

// Chain program of the uh_cheb levels a..c (c the coarsest, which keeps no
// rows; ncheb = 2 on a..c-1).  The recursion of Prec::solve
// (cheb_solve.cc:13-123) then runs as alternating launches:
//   descent from level j: the L sweep of level j from its input (the
//     parent's F, or the Chebyshev residual of a second visit), then the L
//     sweeps of the first visits of levels j+1..c-1, each from the F of the
//     level above kept in LDS, then step 1 on level c (its L and U sweeps);
//   ascent to level k: step 2 on level c from the residual A_c (a u1) - F,
//     then for c-1, c-2, ..., k the Chebyshev combine of the kept rows and
//     the U sweep, as long as the visit just completed was a second visit.
// A level-(k+1) row is a kept row of level k, so every row has one identity
// across the levels (its row on level a).  Blocks are formed from the
// connected components of the L and U couplings of all levels a..c joined
// over these identities; a block holds all levels' rows of its components.
// Per level, a block's rows ("slots") are ordered: nodes (eliminated rows and
// kept rows with L entries) by L depth, kept rows without L entries, isolated
// rows.  Rows of no component with a node ("passthrough" rows: kept rows
// without L entries down to the level where they are isolated) run outside
// the blocks.
//
// Chain order of level k (the order its input vector F_{k-1} and the rows of
// A_k are stored in): the slots of block 0, block 1, ..., then the
// passthrough rows that exist on level k.  Blocks are sorted by their deepest
// level (descending), so the blocks with rows on level >= k are a prefix
// (blk_from); passthrough rows likewise (npt of each level).
//
// Block header (kChainHdr words per block and level): [0] chain position of
// slot 0, [1] slots, [2] nodes, [3] kept rows without L entries, [4] L depth
// levels, [5] first U-list entry, [6] U-list length, [7] U depth levels.
constexpr uint32_t kChainHdr = 8;
// (HostVec: the planner's threads write, and so first touch, every element)
struct ChainLevelProgram {
    uint32_t N = 0, M = 0, Md = 0;
    uint32_t nslots = 0, npt = 0;        // block slots; passthrough rows at positions nslots + t
    HostVec<uint32_t> hdr;               // kChainHdr per block
    HostVec<uint32_t> qrow;              // N: natural row at chain position q
    HostVec<uint32_t> src;               // nslots: slot of the same row on the level above (not on level a)
    HostVec<uint32_t> xl;                // nslots: natural row on the level above (where its combine goes)
    HostVec<uint32_t> Ldep, Lptr;        // nslots (+1): node slots only have entries
    HostVec<uint32_t> Lloc;              // slot of the L column
    HostVec<double> Lval;
    HostVec<uint32_t> Udep, uloc, uxidx, Uptr, Ucol;   // U list; Ucol: slot (eliminated or kept row)
    HostVec<double> Uval, Udiag;
    HostVec<double> dval;                // N: isolated-row factor at its chain position (0 elsewhere)
    // The same node data packed for the kernels (8-bit fields: depths, counts and slots are below 256 in
    // blocks of at most 256 slots): Lmeta per slot = L depth | entries << 8 | slot of entry 0 << 16 | slot of
    // entry 1 << 24, Lv2 its first two values; Umeta per U-list entry = U depth | off-diagonal entries << 8 |
    // loc << 16 | slot of entry 0 << 24, Uv1 that entry's value.  Entries past those run through Lptr / Uptr.
    HostVec<uint32_t> Lmeta, Umeta;
    HostVec<double> Lv2, Uv1;
};
struct ChainProgram {
    uint32_t a = 0, nlev = 0, nblk = 0;
    uint32_t max_slots = 0, max_nodes = 0;
    std::vector<ChainLevelProgram> lv;   // levels a..c
    std::vector<uint32_t> blk_from;      // nlev: blocks with rows on level a+i
};
// Plans the chain over levels L[0..] = a..c.  Components are packed into
// blocks of at most max_nodes nodes and max_nodes slots per level; returns
// false when one component alone has more than max_nodes nodes or max_slots
// slots on some level.
// The most nodes (eliminated rows and kept rows with L entries) in one connected
// component of a single level's L and U couplings.  A chain merges components
// across levels, so a chain through a level whose largest component exceeds
// the chain's block size cannot be planned.
uint32_t max_level_component(const LevelArrays& lv);
bool plan_chain(const std::vector<LevelArrays>& L, uint32_t a, uint32_t max_nodes, uint32_t max_slots,
                ChainProgram& P);

}  // namespace mlp
