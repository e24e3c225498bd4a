// Host-side checks of the grid barrier's group arithmetic (grid_sync.h) and a
// thread model of its arrival/release protocol.  Built and run by
// tests/test_grid_sync_model.py; exits non-zero on the first violation.
//
// Arithmetic: for every grid size up to 8192 workgroups, each group's member
// count from gs_members equals the number of workgroups w with w % ng == g,
// and the counts sum to the grid.
//
// Protocol: G threads stand for G workgroups and run R barriers with the
// device code's steps on std::atomic words: add to the group counter; the
// group's last arrival resets it and adds to the root; the root's last
// arrival resets the root and advances every group's release word; the
// others wait for their group's word to change.  Each thread records the
// round it has finished; no thread may leave barrier r before every thread
// has arrived at it, and every counter is zero after the last barrier.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "grid_sync.h"

using perks::gs_groups;
using perks::gs_members;
using perks::kSyncGroups;

static int check_arithmetic()
{
    for (unsigned int G = 1; G <= 8192; ++G) {
        const unsigned int ng = gs_groups(G);
        if (ng < 1 || ng > (unsigned)kSyncGroups || ng > G) return std::printf("ng(%u) = %u\n", G, ng), 1;
        std::vector<unsigned int> count(ng, 0);
        for (unsigned int w = 0; w < G; ++w) ++count[w % ng];
        unsigned int sum = 0;
        for (unsigned int g = 0; g < ng; ++g) {
            if (gs_members(G, g) != count[g])
                return std::printf("G=%u group %u: members %u, workgroups %u\n", G, g, gs_members(G, g), count[g]), 1;
            sum += gs_members(G, g);
        }
        if (sum != G) return std::printf("G=%u: members sum to %u\n", G, sum), 1;
    }
    return 0;
}

struct Model {
    std::vector<std::atomic<unsigned int>> count, gen;
    std::atomic<unsigned int> root{0};
    std::vector<std::atomic<int>> arrived;  // per round
    std::atomic<int> bad{0};
    Model(unsigned int ng, int rounds) : count(ng), gen(ng), arrived(rounds)
    {
        for (auto& c : count) c = 0;
        for (auto& c : gen) c = 0;
        for (auto& a : arrived) a = 0;
    }
};

static void workgroup(Model& m, unsigned int G, unsigned int w, int rounds)
{
    const unsigned int ng = gs_groups(G);
    const unsigned int g = w % ng;
    const unsigned int members = gs_members(G, g);
    for (int r = 0; r < rounds; ++r) {
        m.arrived[r].fetch_add(1);
        const unsigned int g0 = m.gen[g].load();
        const unsigned int before = m.count[g].fetch_add(1);
        bool released = false;
        if (before == members - 1) {
            m.count[g].exchange(0);
            const unsigned int rb = m.root.fetch_add(1);
            if (rb == ng - 1) {
                m.root.exchange(0);
                for (unsigned int j = 0; j < ng; ++j) m.gen[j].fetch_add(1);
                released = true;
            }
        }
        if (!released)
            while (m.gen[g].load() == g0) std::this_thread::yield();
        if (m.arrived[r].load() != (int)G) m.bad.fetch_add(1);  // left before everyone arrived
    }
}

static int check_protocol(unsigned int G, int rounds)
{
    Model m(gs_groups(G), rounds);
    std::vector<std::thread> t;
    t.reserve(G);
    for (unsigned int w = 0; w < G; ++w) t.emplace_back(workgroup, std::ref(m), G, w, rounds);
    for (auto& x : t) x.join();
    if (m.bad.load()) return std::printf("G=%u: %d early releases\n", G, m.bad.load()), 1;
    if (m.root.load()) return std::printf("G=%u: root left at %u\n", G, m.root.load()), 1;
    for (auto& c : m.count)
        if (c.load()) return std::printf("G=%u: a group counter left at %u\n", G, c.load()), 1;
    return 0;
}

int main()
{
    if (check_arithmetic()) return 1;
    const unsigned int sizes[] = {1, 2, 3, 31, 32, 33, 63, 255, 256, 1024, 1792, 2048};
    for (unsigned int G : sizes)
        if (check_protocol(G, G >= 1024 ? 6 : 40)) return 1;
    std::printf("ok\n");
    return 0;
}
