// Host check of the uh_cheb tail walk (cheb_mark.cc:24-64) on link counters that
// disagree with the degrees: a node reported with one link whose links are all
// removed must stop construction with an error instead of walking on from a
// stale neighbour.  Built and run by tests/test_host_checks.py.
#include "../../ml_prec_amd/csrc/hierarchy.cc"

#include <cstdio>

int main() {
    using namespace mlp;
    // path 0 - 1 - 2, Skyline rows (diagonal first, then ascending columns)
    HostCSR A;
    A.nrow = A.ncol = 3;
    A.ia = {0, 2, 5, 7};
    A.ja = {0, 1, 1, 0, 2, 2, 1};
    A.a = {2.0, -1.0, 3.0, -1.0, -1.0, 2.0, -1.0};
    // consistent counters: upper entries 2 (present), lower entries point at their twin
    HostVec<uint32_t> cnt(A.nnz());
    cnt[0] = 2; cnt[1] = 2; cnt[2] = 2; cnt[3] = 1; cnt[4] = 2; cnt[5] = 2; cnt[6] = 4;
    {
        Hierarchy H;
        H.levels.assign(2, HostLevel());
        H.levels[0].N = 3;
        ChebBuilder b{H, true, 30};
        TwoSidedLinks links(A, HostVec<uint32_t>(cnt));
        std::vector<int> deg = {1, 2, 1};
        b.permutation(A, links, deg, H.levels[0]);
        const auto& m = H.levels[0].map;
        if (!(m == std::vector<uint32_t>{0, 1, 2} && H.levels[0].M == 3 && H.levels[0].Md == 0)) {
            std::printf("FAIL consistent: map %u %u %u M %u\n", m[0], m[1], m[2], H.levels[0].M);
            return 1;
        }
    }
    // inconsistent: node 0 claims one link, but both links are removed
    cnt[1] = 0;
    cnt[4] = 0;
    {
        Hierarchy H;
        H.levels.assign(2, HostLevel());
        H.levels[0].N = 3;
        ChebBuilder b{H, true, 30};
        TwoSidedLinks links(A, HostVec<uint32_t>(cnt));
        std::vector<int> deg = {1, 2, 1};
        try {
            b.permutation(A, links, deg, H.levels[0]);
        } catch (const Error& e) {
            std::printf("OK %d %s\n", e.code(), e.what());
            return 0;
        }
        std::printf("FAIL inconsistent counters were accepted\n");
        return 1;
    }
}
