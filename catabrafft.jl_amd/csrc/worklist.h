// worklist.h -- host-side construction of work-list launches (uneven batches, serving rounds).
#pragma once
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "library.h"
#include "multi.h"

namespace cfft {

// One distinct transform plan inside a work list, with its descriptor indices.
struct PlanEntry {
    const Plan* plan = nullptr;
    int mplan = -1;  // descriptor of the transform (rows / four-step / fused conv / conv ends)
    int msub = -1;   // large convolution: descriptor of its length-L FFT (carries K)
};

// Where one transform of a plan entry reads and writes.
struct XformRef {
    uint64_t in = 0, out = 0;  // addresses or offsets, see in_sel / out_sel
    int in_sel = kAbs, out_sel = kAbs;
    uint64_t scratch = 0;      // four-step middle buffer (absolute); 0 = use `out`
    uint64_t w1 = 0, w2 = 0;   // large convolution work lines (absolute)
    uint64_t dc = 0;           // large convolution slots {A^[0], x0} (absolute)
    bool launch_dir = true;    // direction / normalisation from the launch arguments
    int sgn = 1;
    double scale = 1.0;
    double cost() const { return 0; }
};

class WorkList {
   public:
    // Registers a plan (Direct, FourStep, ConvSmall or ConvLarge); returns its entry.
    PlanEntry add_plan(const Plan& p);
    // Adds every item of the given transforms of one plan entry.
    void add_transforms(const PlanEntry& e, const std::vector<XformRef>& xs);
    // Uploads descriptors, line tables and items; returns the launches in execution order.
    // Uploads the tables (growing the device buffers when too small) and fills `out`.  With
    // async, the uploads are enqueued on `st` from `stage` (pinned host memory that must stay
    // untouched until they complete) or, without one, from this work list's own host buffer.
    int finalize(DevBuf& dplans, DevBuf& ditems, DevBuf& dtable, std::vector<MultiLaunch>& out,
                 hipStream_t st = nullptr, bool async = false, HostStage* stage = nullptr);
    size_t item_count() const;

   private:
    struct Group {
        int phase, kernel, prec, cprec;
        int block;
        size_t lds;
        std::vector<MItem> items;
        // runs of consecutive items of equal cost: (cost, first item); a launch orders its
        // items by decreasing cost, equal costs in insertion order
        std::vector<std::pair<double, size_t>> runs;
    };
    using Key = std::tuple<int, int, int, int, int, size_t>;
    std::vector<MPlan> mplans_;
    std::map<Key, Group> groups_;
    // the groups found by the latest lookups (std::map nodes do not move)
    std::pair<Key, Group*> recent_[4];
    int recent_next_ = 0;
    std::vector<uint64_t> table_;
    std::vector<unsigned char> host_;  // staging for async uploads (kept alive by the caller)

    int add_mplan(const MPlan& m) {
        mplans_.push_back(m);
        return (int)mplans_.size() - 1;
    }
    void add_item(int phase, MultiKernel k, Prec prec, Prec cprec, int block, size_t lds,
                  const MItem& it, double cost);
    void add_lines(int phase, MultiKernel k, const Plan& lp, int mplan, Prec prec, int C,
                   int block, size_t lds, const std::vector<XformRef>& xs, double line_cost,
                   bool epilogue, int variant);

   public:
    const std::vector<unsigned char>& staging() const { return host_; }
};

MLine make_mline(const LinePlan& lp);

}  // namespace cfft
