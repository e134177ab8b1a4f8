// Host-side planning of the device layouts for one hierarchy level.
//
// The L forward solve and U back solve of a level (cheb_solve.cc:33-38 and
// :108-115; multi_split_solve.cc:25-53) only couple the eliminated rows
// 0..M-1 within small connected components.  The planner packs whole
// components into workgroup-sized blocks and orders each block's rows by
// dependency depth, so one kernel launch runs every depth level of a block
// with a workgroup barrier between levels.  The arithmetic per row keeps the
// reference's operand order, so results are unchanged by the schedule.
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

constexpr uint32_t kGlobalRef = 0x80000000u;   // U column refers to the level output x (global)

// Blocked program for the eliminated rows.  Block b owns positions
// [blk[6b], blk[6b+1]) of the L-ordered node list; its L depth levels are
// lvlL[blk[6b+2] .. blk[6b+3]] (absolute node positions, one past each level
// end) and its U levels lvlU[blk[6b+4] .. blk[6b+5]] (positions in U order).
struct TailLayout {
    uint32_t M = 0, nblk = 0, max_nodes = 0, max_levels = 0;
    std::vector<uint32_t> blk;
    std::vector<uint32_t> lvlL, lvlU;
    // L order (size M)
    std::vector<uint32_t> fidx;          // map[i]: where the row reads the level input
    std::vector<uint32_t> Lptr;          // M+1
    std::vector<uint32_t> Lloc;          // local position of the child within the block
    std::vector<double> Lval;
    // U order (size M)
    std::vector<uint32_t> uloc;          // local position (L order) of the row
    std::vector<uint32_t> uxidx;         // map[i]: where the row writes the level output
    std::vector<uint32_t> Uptr;          // M+1
    std::vector<uint32_t> Ucol;          // local position, or kGlobalRef | map[j] for rows >= M
    std::vector<double> Uval, Udiag;
    // row i (< M) -> global L-order position
    std::vector<uint32_t> pos_of_row;
};

// Rows M..N-Md-1 (the next level's right-hand side) as a gather + short dot.
struct SchurLayout {
    std::vector<uint32_t> ptr, col;      // col = L-order position of the eliminated row
    std::vector<double> val;
};

// 64-row slices, column-major inside a slice: entry k of row 64s+t sits at
// (off64[s] + k) * 64 + t.  Padding has col = 0xffffffff and val = 0.
struct SellLayout {
    uint32_t n = 0, nslices = 0;
    uint64_t nent = 0;
    std::vector<uint32_t> off64;          // nslices+1
    std::vector<uint32_t> col;
    std::vector<double> val;
};

void plan_tail(const LevelArrays& lv, uint32_t cap, TailLayout& out);
void plan_schur(const LevelArrays& lv, const TailLayout& tl, SchurLayout& out);
void plan_sell(uint32_t n, const uint32_t* ia, const uint32_t* ja, const double* a, SellLayout& out);

}  // namespace mlp
