// Grouped launches: one kernel launch covers many independent work items
// (members of an ensemble, components of a field, or both).
//
// Every kernel of the library takes
//     const Args *__restrict__ tab     per-item arguments (device memory)
//     const int2 *__restrict__ bmap    per block: (item, block index within the item)
// so a launch over items of different sizes is a single dispatch whose blocks
// index their own item's arguments.  A Plan is an ordered list of such
// launches with all tables and block maps in one device arena; it is built
// once per (member set, operation) and replayed, eagerly or as a captured
// hipGraph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace epic {

// device side: the work item and local block of this work-group
struct ItemBlock {
    int item, blk;
};
__device__ inline ItemBlock item_block(const int2 *__restrict__ bmap)
{
    const int2 e = bmap[blockIdx.x];
    return ItemBlock{e.x, e.y};
}

// phases of the per-step work, for the event timers (the public timer ids
// are derived from these)
enum Phase : int {
    PH_AXES = 0,
    PH_SORT,
    PH_P2G,
    PH_VOR2VEL,
    PH_VTEND,
    PH_PUSH,
    PH_GRID2PAR,
    PH_P2G_TILES,   // the tile scatter of par2grid alone (also counted in par2grid)
    PH_OTHER,
    PH_COUNT
};

class Plan {
public:
    struct Rec {
        const void *fn;
        unsigned nblocks, threads;
        size_t lds;
        size_t tab_off, map_off;
        int phase;
    };

    // one launch: items[i] gets blocks[i] work-groups (0: skipped)
    template <class Args>
    void add(const void *fn, unsigned threads, size_t lds, int phase, const std::vector<Args> &items,
             const std::vector<int> &blocks)
    {
        std::vector<int> bm;
        for (size_t i = 0; i < items.size(); ++i)
            for (int b = 0; b < blocks[i]; ++b) {
                bm.push_back((int)i);
                bm.push_back(b);
            }
        if (bm.empty()) return;
        Rec r;
        r.fn = fn;
        r.nblocks = (unsigned)(bm.size() / 2);
        r.threads = threads;
        r.lds = lds;
        r.phase = phase;
        r.tab_off = put(items.data(), sizeof(Args) * items.size());
        r.map_off = put_map(bm);
        recs_.push_back(r);
        dev_ready_ = false;
    }

    bool empty() const { return recs_.empty(); }
    size_t size() const { return recs_.size(); }
    const std::vector<Rec> &recs() const { return recs_; }

    // copies the arena to the device on s (once; again after new launches)
    hipError_t upload(hipStream_t s);
    // enqueues every launch on s; with ev != nullptr records an event pair
    // around every run of launches of one phase (ev_phase gets the phases)
    hipError_t run(hipStream_t s, std::vector<hipEvent_t> *ev = nullptr, std::vector<int> *ev_phase = nullptr);
    // captures the launches into a graph (once) and launches it on s; a
    // timed capture puts a one-thread kernel that stores the device wall
    // clock before and after every run of launches of one phase, so a replay
    // times its phases (a change of `timed` captures again)
    hipError_t run_graph(hipStream_t s, bool timed = false);
    // uploads and captures (once) without launching
    hipError_t prepare_graph(hipStream_t s, bool timed = false);
    // after a timed replay has completed: (phase, ms) of every phase run
    hipError_t graph_phase_times(std::vector<std::pair<int, float>> &out);
    // when o is the same launch sequence as this uploaded plan (kernels,
    // grids, work-group sizes, LDS and arena layout), takes o's argument
    // tables into the device arena on s and keeps the captured graph, whose
    // launches read the tables in place (*same = true); otherwise changes
    // nothing (*same = false)
    hipError_t refresh_from(const Plan &o, hipStream_t s, bool *same);
    void swap(Plan &o);
    void release();
    ~Plan() { release(); }
    Plan() = default;
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;

private:
    size_t put(const void *p, size_t n)
    {
        const size_t off = (host_.size() + 255) & ~size_t(255);
        host_.resize(off + n);
        std::memcpy(host_.data() + off, p, n);
        return off;
    }
    size_t put_map(const std::vector<int> &bm)
    {
        auto it = maps_.find(bm);
        if (it != maps_.end()) return it->second;
        const size_t off = put(bm.data(), bm.size() * sizeof(int));
        maps_.emplace(bm, off);
        return off;
    }

    std::vector<char> host_;
    std::map<std::vector<int>, size_t> maps_;
    std::vector<Rec> recs_;
    char *dev_ = nullptr;
    size_t dev_size_ = 0;
    hipStream_t last_stream_ = nullptr;   // the arena returns to the pool behind this stream's work
    bool dev_ready_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    bool timed_ = false;
    unsigned long long *stamps_ = nullptr;   // device: two wall-clock stamps per phase run
    std::vector<int> stamp_phase_;
    std::vector<unsigned long long> stamps_h_;
    double clock_khz_ = 0.0;
    void drop_graph();
    hipError_t run_stamped(hipStream_t s);
};

}  // namespace epic
