// uneven.cpp -- many transforms of different lengths in one plan (cfft_plan_many_uneven), and
// the work-list builder it shares with the serving loop.
//
// Transforms are grouped by length; each distinct length gets an ordinary plan (its tables are
// shared by every transform of that length).  Execution is a short, fixed sequence of work-list
// launches, each covering every transform that needs that kernel in that phase:
//   0 gather (large convolutions)
//   1 rows | fused convolution | four-step pass A (incl. first FFT_L of large convolutions)
//   2 four-step pass B (incl. first FFT_L, with the convolution epilogue)
//   3 rows | four-step pass A of the second FFT_L
//   4 four-step pass B of the second FFT_L
//   5 scatter (large convolutions)
// split further by precision and by workgroup shape (threads, LDS).  The sequence is captured
// once into a hipGraph per (input, output, direction, normalisation) and replayed.
#include "worklist.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <system_error>
#include <thread>

#include "catabrafft.h"

namespace cfft {

// Threads per workgroup that the four-step tiles of an uneven batch's four-step plans are
// widened to: the default tiles give the 128- to 512-point lines of these lengths workgroups of
// 32-128 threads (measured on MI355X, 10,000 lengths: 9.09 ms with the default tiles,
// 8.96-8.99 ms at 256 threads, 9.30 ms at 512; widening the large convolutions' tiles as well
// made them slower).
constexpr int kFsMinThreads = 256;

#define CFFT_HIP(expr)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(CFFT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static size_t esz(Prec p) { return p == Prec::F32 ? 8 : 16; }

MLine make_mline(const LinePlan& lp) {
    MLine m{};
    m.N = lp.N;
    m.Tt = lp.Tt;
    m.Ls = lp.Ls;
    m.pt = lp.pt;
    m.tw = lp.dtw.p;
    return m;
}

static int pow2_ceil(int v) {
    int r = 64;
    while (r < v) r *= 2;
    return r;
}

static size_t lds_class(size_t b) {
    const size_t q = 16 * 1024;
    return (b + q - 1) / q * q;
}

void WorkList::add_item(int phase, MultiKernel k, Prec prec, Prec cprec, int block, size_t lds,
                        const MItem& it, double cost) {
    const int bc = (k == MultiKernel::Gather || k == MultiKernel::Scatter) ? 256 : pow2_ceil(block);
    const size_t lc = lds_class(lds);
    const Key key{phase, (int)k, (int)prec, (int)cprec, bc, lc};
    Group* gp = nullptr;
    for (auto& r : recent_)
        if (r.second && r.first == key) gp = r.second;
    if (!gp) {
        gp = &groups_[key];
        recent_[recent_next_] = {key, gp};
        recent_next_ = (recent_next_ + 1) & 3;
    }
    Group& g = *gp;
    if (g.items.empty()) {
        g.phase = phase;
        g.kernel = (int)k;
        g.prec = (int)prec;
        g.cprec = (int)cprec;
        g.block = bc;
        g.lds = 0;
    }
    g.lds = std::max(g.lds, lds);
    if (g.runs.empty() || g.runs.back().first != cost) g.runs.emplace_back(cost, g.items.size());
    g.items.push_back(it);
}

size_t WorkList::item_count() const {
    size_t n = 0;
    for (const auto& kv : groups_) n += kv.second.items.size();
    return n;
}

PlanEntry WorkList::add_plan(const Plan& p) {
    PlanEntry e;
    e.plan = &p;
    MPlan m{};
    m.n = (int)p.n;
    m.C = p.C;
    switch (p.kind) {
        case Kind::Direct:
            m.line = make_mline(p.line);
            break;
        case Kind::FourStep:
            m.N1 = p.fs.N1;
            m.N2 = p.fs.N2;
            m.CA = p.fs.CA;
            m.CB = p.fs.CB;
            m.tbits = p.fs.tbits;
            m.tnc = p.fs.tnc;
            m.la = make_mline(p.fs.a);
            m.lb = make_mline(p.fs.b);
            m.tw3 = p.fs.tw3.p;
            break;
        case Kind::ConvSmall:
        case Kind::ConvLarge:
            m.line = make_mline(p.line);
            m.mode = (int)p.mode;
            m.L = (int)p.L;
            m.Ls = p.LsC;
            m.perm = (const int*)p.perm.p;
            m.iperm = (const int*)p.iperm.p;
            m.chirp = p.chirp.p;
            m.K = p.K.p;
            if (p.kind == Kind::ConvLarge && p.convl) {  // its FFT_L four-step, for the fused launches
                const FourStepPlan& f = p.sub->fs;
                m.N1 = f.N1;
                m.N2 = f.N2;
                m.CA = p.convl_ca;
                m.CB = p.convl_cb;
                m.CC = p.convl_cc;
                m.LsC = p.convl_lsc;
                m.tbits = f.tbits;
                m.tnc = f.tnc;
                m.la = make_mline(f.a);
                m.la.Ls = p.convl_lsa;
                m.lb = make_mline(f.b);
                m.lb.Ls = p.convl_lsb;
                m.tw3 = f.tw3.p;
            }
            break;
        default:
            break;
    }
    e.mplan = add_mplan(m);
    if (p.kind == Kind::ConvLarge) {
        const Plan& s = *p.sub;
        MPlan ms{};
        ms.n = (int)s.n;
        ms.C = s.C;
        if (s.kind == Kind::Direct) {
            ms.line = make_mline(s.line);
        } else {
            ms.N1 = s.fs.N1;
            ms.N2 = s.fs.N2;
            ms.CA = s.fs.CA;
            ms.CB = s.fs.CB;
            ms.tbits = s.fs.tbits;
            ms.tnc = s.fs.tnc;
            ms.la = make_mline(s.fs.a);
            ms.lb = make_mline(s.fs.b);
            ms.tw3 = s.fs.tw3.p;
        }
        ms.K = p.K.p;
        e.msub = add_mplan(ms);
    }
    return e;
}

// Lines of a rows / fused-conv plan, C per item, through a line table.
// variant: 0 = transform in -> out, 1 = conv FFT_L #1 (w1 -> w2, epilogue), 2 = FFT_L #2 (w2 -> w1)
void WorkList::add_lines(int phase, MultiKernel k, const Plan& lp, int mplan, Prec prec, int C,
                         int block, size_t lds, const std::vector<XformRef>& xs,
                         double line_cost, bool epilogue, int variant) {
    // direction groups: lines of one item share sgn / scale / launch_dir
    size_t i = 0;
    while (i < xs.size()) {
        MItem it{};
        it.plan = mplan;
        const XformRef& x0 = xs[i];
        int count = 0;
        const size_t tbase = table_.size();
        int in_sel = kAbs, out_sel = kAbs;
        while (i < xs.size() && count < C) {
            const XformRef& x = xs[i];
            if (variant == 0) {
                if (count > 0 && (x.launch_dir != x0.launch_dir || x.sgn != x0.sgn ||
                                  x.scale != x0.scale || x.in_sel != x0.in_sel ||
                                  x.out_sel != x0.out_sel))
                    break;
                table_.push_back(x.in);
                table_.push_back(x.out);
                table_.push_back(0);
                in_sel = x.in_sel;
                out_sel = x.out_sel;
            } else if (variant == 1) {
                table_.push_back(x.w1);
                table_.push_back(x.w2);
                table_.push_back(x.dc);
            } else {
                table_.push_back(x.w2);
                table_.push_back(x.w1);
                table_.push_back(0);
            }
            ++count;
            ++i;
        }
        it.count = count;
        it.flags = in_sel | (out_sel << kSelOutShift) | kLineTable;
        if (variant == 0) {
            if (x0.launch_dir) it.flags |= kLaunchDir;
            it.sgn = x0.sgn;
            it.scale = x0.scale;
        } else {
            it.sgn = 1;
            it.scale = 1.0;
            if (epilogue) it.flags |= kEpilogue;
        }
        it.aux = (uint64_t)(tbase * sizeof(uint64_t));  // relocated by finalize()
        add_item(phase, k, prec, prec, block, lds, it, line_cost * count);
    }
    (void)lp;
}

void WorkList::add_transforms(const PlanEntry& e, const std::vector<XformRef>& xs) {
    if (xs.empty()) return;
    const Plan& p = *e.plan;
    const Prec pr = p.prec;
    const double lg = std::log2((double)std::max<int64_t>(2, p.n));
    switch (p.kind) {
        case Kind::Direct: {
            const int block = p.C * p.line.Tt;
            const size_t lds = 2 * (esz(pr) / 2) * (size_t)p.C * p.line.Ls;
            add_lines(1, MultiKernel::Rows, p, e.mplan, pr, p.C, block, lds, xs,
                      (double)p.n * lg, false, 0);
            break;
        }
        case Kind::ConvSmall: {
            const int block = p.C * p.line.Tt;
            const size_t lds = 2 * (esz(pr) / 2) * (size_t)p.C * p.LsC + esz(pr) * p.C;
            add_lines(1, MultiKernel::ConvSmall, p, e.mplan, pr, p.C, block, lds, xs,
                      4.0 * p.L * std::log2((double)p.L), false, 0);
            break;
        }
        case Kind::FourStep: {
            const FourStepPlan& f = p.fs;
            const int ta = (f.N2 + f.CA - 1) / f.CA, tb = (f.N1 + f.CB - 1) / f.CB;
            const size_t lda = 2 * (esz(pr) / 2) * (size_t)f.CA * f.a.Ls;
            const size_t ldb = fsb_tw_offset(esz(pr) / 2, f.CB, f.b.Ls) +
                               (size_t)tw3_entries(f.tbits, f.tnc) * esz(pr);
            for (const XformRef& x : xs) {
                const uint64_t mid = x.scratch ? x.scratch : x.out;
                const int mid_sel = x.scratch ? kAbs : x.out_sel;
                for (int t = 0; t < ta; ++t) {
                    MItem it{};
                    it.plan = e.mplan;
                    it.count = t;
                    it.in = x.in;
                    it.out = mid;
                    it.flags = x.in_sel | (mid_sel << kSelOutShift) | (x.launch_dir ? kLaunchDir : 0);
                    it.sgn = x.sgn;
                    it.scale = x.scale;
                    add_item(1, MultiKernel::FourStepA, pr, pr, f.CA * f.a.Tt, lda, it,
                             (double)f.CA * f.N1 * lg);
                }
                for (int t = 0; t < tb; ++t) {
                    MItem it{};
                    it.plan = e.mplan;
                    it.count = t;
                    it.in = mid;
                    it.out = x.out;
                    it.flags = mid_sel | (x.out_sel << kSelOutShift) | (x.launch_dir ? kLaunchDir : 0);
                    it.sgn = x.sgn;
                    it.scale = x.scale;
                    add_item(2, MultiKernel::FourStepB, pr, pr, f.CB * f.b.Tt, ldb, it,
                             (double)f.CB * f.N2 * lg);
                }
            }
            break;
        }
        case Kind::ConvLarge: {
            const Plan& s = *p.sub;
            const Prec cp = p.cprec;
            if (p.convl) {
                // three fused launches (phases 1-3), one item per tile of each transform
                const FourStepPlan& f = s.fs;
                const size_t ce = esz(cp) / 2;
                const size_t table = (size_t)tw3_entries(f.tbits, f.tnc) * esz(cp);
                const int CA = p.convl_ca, CB = p.convl_cb;
                const size_t lda = 2 * ce * (size_t)CA * p.convl_lsa;
                const size_t ldb = fsb_tw_offset(ce, CB, p.convl_lsb) + table;
                const size_t ldc = fsb_tw_offset(ce, p.convl_cc, p.convl_lsc) + table;
                const int ta = (f.N2 + CA - 1) / CA, tb = (f.N1 + CB - 1) / CB;
                const int tc = (f.N2 + p.convl_cc - 1) / p.convl_cc;
                const double la = std::log2((double)f.N1), lb = std::log2((double)f.N2);
                for (const XformRef& x : xs) {
                    MItem it{};
                    it.plan = e.mplan;
                    it.dc = x.dc;
                    it.sgn = x.sgn;
                    it.scale = x.scale;
                    const int dirf = x.launch_dir ? kLaunchDir : 0;
                    for (int t = 0; t < ta; ++t) {
                        it.count = t;
                        it.in = x.in;
                        it.out = x.w1;
                        it.flags = x.in_sel | (kAbs << kSelOutShift) | dirf;
                        add_item(1, MultiKernel::ConvLA, pr, cp, CA * f.a.Tt, lda, it,
                                 (double)CA * f.N1 * la);
                    }
                    for (int t = 0; t < tb; ++t) {
                        it.count = t;
                        it.in = x.w1;
                        it.out = x.w2;
                        it.flags = kAbs | (kAbs << kSelOutShift) | dirf;
                        add_item(2, MultiKernel::ConvLBA, pr, cp, CB * f.b.Tt, ldb, it,
                                 2.0 * CB * f.N2 * lb);
                    }
                    for (int t = 0; t < tc; ++t) {
                        it.count = t;
                        it.in = x.w2;
                        it.out = x.out;
                        it.flags = kAbs | (x.out_sel << kSelOutShift) | dirf;
                        add_item(3, MultiKernel::ConvLC, pr, cp, p.convl_cc * f.a.Tt, ldc, it,
                                 (double)p.convl_cc * f.N1 * la);
                    }
                }
                break;
            }
            const int64_t gchunks = (p.L + kChunkHost - 1) / kChunkHost;
            const int64_t schunks = (p.n + kChunkHost - 1) / kChunkHost;
            for (const XformRef& x : xs) {
                for (int64_t c = 0; c < gchunks; ++c) {
                    MItem it{};
                    it.plan = e.mplan;
                    it.count = (int)c;
                    it.in = x.in;
                    it.out = x.w1;
                    it.dc = x.dc;
                    it.flags = x.in_sel | (kAbs << kSelOutShift) | (x.launch_dir ? kLaunchDir : 0);
                    it.sgn = x.sgn;
                    it.scale = x.scale;
                    add_item(0, MultiKernel::Gather, pr, cp, 256, 0, it, (double)kChunkHost);
                }
                for (int64_t c = 0; c < schunks; ++c) {
                    MItem it{};
                    it.plan = e.mplan;
                    it.count = (int)c;
                    it.in = x.w1;
                    it.out = x.out;
                    it.dc = x.dc;
                    it.flags = kAbs | (x.out_sel << kSelOutShift) | (x.launch_dir ? kLaunchDir : 0);
                    it.sgn = x.sgn;
                    it.scale = x.scale;
                    add_item(5, MultiKernel::Scatter, pr, cp, 256, 0, it, (double)kChunkHost);
                }
            }
            const double lgL = std::log2((double)p.L);
            if (s.kind == Kind::Direct) {
                const int block = s.C * s.line.Tt;
                const size_t lds = 2 * (esz(cp) / 2) * (size_t)s.C * s.line.Ls;
                add_lines(1, MultiKernel::Rows, s, e.msub, cp, s.C, block, lds, xs,
                          (double)p.L * lgL, true, 1);
                add_lines(3, MultiKernel::Rows, s, e.msub, cp, s.C, block, lds, xs,
                          (double)p.L * lgL, false, 2);
            } else {
                const FourStepPlan& f = s.fs;
                const int ta = (f.N2 + f.CA - 1) / f.CA, tb = (f.N1 + f.CB - 1) / f.CB;
                const size_t lda = 2 * (esz(cp) / 2) * (size_t)f.CA * f.a.Ls;
                const size_t ldb = fsb_tw_offset(esz(cp) / 2, f.CB, f.b.Ls) +
                                   (size_t)tw3_entries(f.tbits, f.tnc) * esz(cp);
                for (const XformRef& x : xs) {
                    for (int pass = 0; pass < 2; ++pass) {
                        const uint64_t src = pass == 0 ? x.w1 : x.w2;
                        const uint64_t dst = pass == 0 ? x.w2 : x.w1;
                        for (int t = 0; t < ta; ++t) {
                            MItem it{};
                            it.plan = e.msub;
                            it.count = t;
                            it.in = src;
                            it.out = dst;
                            it.flags = kAbs | (kAbs << kSelOutShift);
                            it.sgn = 1;
                            it.scale = 1.0;
                            add_item(pass == 0 ? 1 : 3, MultiKernel::FourStepA, cp, cp,
                                     f.CA * f.a.Tt, lda, it, (double)f.CA * f.N1 * lgL);
                        }
                        for (int t = 0; t < tb; ++t) {
                            MItem it{};
                            it.plan = e.msub;
                            it.count = t;
                            it.in = dst;
                            it.out = dst;
                            it.flags = kAbs | (kAbs << kSelOutShift) | (pass == 0 ? kEpilogue : 0);
                            it.dc = pass == 0 ? x.dc : 0;
                            it.sgn = 1;
                            it.scale = 1.0;
                            add_item(pass == 0 ? 2 : 4, MultiKernel::FourStepB, cp, cp,
                                     f.CB * f.b.Tt, ldb, it, (double)f.CB * f.N2 * lgL);
                        }
                    }
                }
            }
            break;
        }
        default:
            break;
    }
}

int WorkList::finalize(DevBuf& dplans, DevBuf& ditems, DevBuf& dtable,
                       std::vector<MultiLaunch>& out, hipStream_t st, bool async,
                       HostStage* stage) {
    int rc;
    out.clear();
    // Groups of one phase, kernel and precision pair differ only in workgroup size and LDS.  A
    // group too small to fill the GPU is tail-bound as a launch of its own: it is folded into the
    // next group of the family (ascending block, then LDS), whose launch then takes the larger
    // block and LDS of the two (threads beyond an item's lines idle, or repeat a butterfly of
    // their line with the same values).
    std::vector<Group> merged;
    {
        std::vector<const Group*> fam;
        for (auto& kv : groups_) fam.push_back(&kv.second);
        std::stable_sort(fam.begin(), fam.end(), [](const Group* a, const Group* b) {
            return std::tie(a->phase, a->kernel, a->prec, a->cprec, a->block, a->lds) <
                   std::tie(b->phase, b->kernel, b->prec, b->cprec, b->block, b->lds);
        });
        constexpr size_t kMinItems = 2048;
        // groups join only if neither block nor LDS grows by more than 2x: the merged launch runs
        // every item at the larger of both, which costs occupancy
        auto close = [](const Group& a, const Group& b) {
            return std::max(a.block, b.block) <= 2 * std::min(a.block, b.block) &&
                   std::max(a.lds, b.lds) <= 2 * std::min(a.lds, b.lds);
        };
        for (size_t i = 0; i < fam.size(); ++i) {
            const Group& g = *fam[i];
            const bool same_family =
                !merged.empty() && merged.back().phase == g.phase &&
                merged.back().kernel == g.kernel && merged.back().prec == g.prec &&
                merged.back().cprec == g.cprec && close(merged.back(), g);
            if (same_family && merged.back().items.size() < kMinItems) {
                Group& m = merged.back();
                const size_t off = m.items.size();
                m.items.insert(m.items.end(), g.items.begin(), g.items.end());
                for (const auto& r : g.runs) m.runs.emplace_back(r.first, r.second + off);
                m.block = std::max(m.block, g.block);
                m.lds = std::max(m.lds, g.lds);
                continue;
            }
            merged.push_back(g);
        }
        // a small last group of a family joins the group before it
        for (size_t i = merged.size(); i-- > 1;) {
            Group& g = merged[i];
            Group& prev = merged[i - 1];
            const bool last_of_family =
                i + 1 == merged.size() || merged[i + 1].phase != g.phase ||
                merged[i + 1].kernel != g.kernel || merged[i + 1].prec != g.prec ||
                merged[i + 1].cprec != g.cprec;
            if (!last_of_family || g.items.size() >= kMinItems || prev.phase != g.phase ||
                prev.kernel != g.kernel || prev.prec != g.prec || prev.cprec != g.cprec ||
                !close(prev, g))
                continue;
            const size_t off = prev.items.size();
            prev.items.insert(prev.items.end(), g.items.begin(), g.items.end());
            for (const auto& r : g.runs) prev.runs.emplace_back(r.first, r.second + off);
            prev.block = std::max(prev.block, g.block);
            prev.lds = std::max(prev.lds, g.lds);
            merged.erase(merged.begin() + (std::ptrdiff_t)i);
        }
    }
    // order: phase, then kernel; inside a group, most expensive items first
    std::vector<const Group*> gs;
    for (const Group& g : merged) gs.push_back(&g);
    std::stable_sort(gs.begin(), gs.end(), [](const Group* a, const Group* b) {
        return std::tie(a->phase, a->kernel) < std::tie(b->phase, b->kernel);
    });
    size_t nitems = 0;
    for (const Group* g : gs) nitems += g->items.size();
    const size_t bp = mplans_.size() * sizeof(MPlan);
    const size_t bt = std::max<size_t>(1, table_.size()) * sizeof(uint64_t);
    const size_t bi = std::max<size_t>(1, nitems) * sizeof(MItem);
    // grow with headroom, so that buffers reused round after round are rarely reallocated
    if (dplans.bytes < bp && (rc = dplans.alloc(bp + bp / 2))) return rc;
    if (dtable.bytes < bt && (rc = dtable.alloc(bt + bt / 2))) return rc;
    if (ditems.bytes < bi && (rc = ditems.alloc(bi + bi / 2))) return rc;
    std::vector<MItem> all;
    all.reserve(nitems);
    for (const Group* g : gs) {
        // most expensive items first, equal costs in insertion order: a stable sort of the runs
        // (run k ends where run k + 1 starts)
        std::vector<size_t> ri(g->runs.size());
        std::iota(ri.begin(), ri.end(), 0);
        std::stable_sort(ri.begin(), ri.end(), [g](size_t a, size_t b) {
            return g->runs[a].first > g->runs[b].first;
        });
        std::vector<size_t> order;
        order.reserve(g->items.size());
        for (size_t k : ri) {
            const size_t end = k + 1 < g->runs.size() ? g->runs[k + 1].second : g->items.size();
            for (size_t i = g->runs[k].second; i < end; ++i) order.push_back(i);
        }
        MultiLaunch L{};
        L.kernel = g->kernel;
        L.prec = g->prec;
        L.cprec = g->cprec;
        L.block = g->block;
        L.lds = g->lds;
        L.items = reinterpret_cast<const MItem*>(ditems.p) + all.size();
        L.nitems = (int64_t)order.size();
        L.plans = reinterpret_cast<const MPlan*>(dplans.p);
        for (size_t k : order) {
            MItem it = g->items[k];
            if (it.flags & kLineTable) it.aux += (uint64_t)dtable.p;
            all.push_back(it);
        }
        out.push_back(L);
    }
    if (async) {
        unsigned char* h;
        if (stage) {
            if (stage->reserve(bp + bt + bi)) return fail(CFFT_ERR_NOMEM, "pinned staging failed");
            h = static_cast<unsigned char*>(stage->p);
        } else {
            host_.resize(bp + bt + bi);
            h = host_.data();
        }
        std::memcpy(h, mplans_.data(), bp);
        if (!table_.empty()) std::memcpy(h + bp, table_.data(), table_.size() * 8);
        if (!all.empty()) std::memcpy(h + bp + bt, all.data(), all.size() * sizeof(MItem));
        CFFT_HIP(hipMemcpyAsync(dplans.p, h, bp, hipMemcpyHostToDevice, st));
        CFFT_HIP(hipMemcpyAsync(dtable.p, h + bp, bt, hipMemcpyHostToDevice, st));
        CFFT_HIP(hipMemcpyAsync(ditems.p, h + bp + bt, bi, hipMemcpyHostToDevice, st));
        return CFFT_OK;
    }
    if (bp) CFFT_HIP(hipMemcpy(dplans.p, mplans_.data(), bp, hipMemcpyHostToDevice));
    if (!table_.empty())
        CFFT_HIP(hipMemcpy(dtable.p, table_.data(), table_.size() * 8, hipMemcpyHostToDevice));
    if (!all.empty())
        CFFT_HIP(hipMemcpy(ditems.p, all.data(), all.size() * sizeof(MItem), hipMemcpyHostToDevice));
    return CFFT_OK;
}

// ----------------------------------------------------------------------------- uneven plans
struct UnevenState {
    std::vector<std::unique_ptr<Plan>> plans;
    DevBuf dplans, ditems, dtable, work;
    std::vector<MultiLaunch> launches;
    hipStream_t cap = nullptr;
    std::mutex mu;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t g_in = 0, g_out = 0;
    int g_sgn = 0, g_norm = -1;
    bool use_graph = true;
    int device = 0;
    ~UnevenState() {
        (void)hipSetDevice(device);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

static int run_launches(const std::vector<MultiLaunch>& ls, const MultiArgs& ma, hipStream_t st) {
    for (const MultiLaunch& L : ls) CFFT_HIP(launch_multi(L, ma, st));
    return CFFT_OK;
}

int uneven_create(Ctx* ctx, int64_t count, const int64_t* lengths, const int64_t* offsets,
                  int dtype, int flags, std::shared_ptr<Plan>* out) {
    if (count < 0) return fail(CFFT_ERR_INVALID, "count must be >= 0");
    if (dtype != CFFT_C32 && dtype != CFFT_C64) return fail(CFFT_ERR_INVALID, "bad dtype");
    if (flags < CFFT_NO_FLAG || flags > CFFT_ENCHANT) return fail(CFFT_ERR_INVALID, "bad flags");
    const Prec pr = dtype == CFFT_C32 ? Prec::F32 : Prec::F64;
    // validate: positive lengths, non-negative offsets, no two transforms overlap
    std::vector<int64_t> idx(count);
    std::iota(idx.begin(), idx.end(), 0);
    int64_t span = 0;
    for (int64_t i = 0; i < count; ++i) {
        if (lengths[i] < 1) return fail(CFFT_ERR_INVALID, "lengths must be >= 1");
        if (offsets[i] < 0) return fail(CFFT_ERR_INVALID, "offsets must be >= 0");
        if (lengths[i] > ((int64_t)1 << 28)) return fail(CFFT_ERR_UNSUPPORTED, "length above 2^28");
        span = std::max(span, offsets[i] + lengths[i]);
    }
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return offsets[a] < offsets[b]; });
    for (int64_t k = 1; k < count; ++k)
        if (offsets[idx[k - 1]] + lengths[idx[k - 1]] > offsets[idx[k]])
            return fail(CFFT_ERR_INVALID, "transforms overlap");
    CFFT_HIP(hipSetDevice(ctx->device));
    auto P = std::make_shared<Plan>();
    P->ctx = ctx;
    P->kind = Kind::Uneven;
    P->dtype = dtype;
    P->flags = flags;
    P->prec = pr;
    P->batch = count;
    P->uneven_span = span;
    auto st = std::make_shared<UnevenState>();
    st->device = ctx->device;
    const char* ev = std::getenv("CFFT_UNEVEN_GRAPH");
    st->use_graph = !(ev && ev[0] == '0');
    // one plan per distinct length
    std::map<int64_t, std::vector<int64_t>> bylen;
    for (int64_t k = 0; k < count; ++k) bylen[lengths[idx[k]]].push_back(idx[k]);
    const size_t es = esz(pr);
    WorkList W;
    size_t work_bytes = 0;
    struct Pending {
        PlanEntry e;
        std::vector<int64_t> ids;
        size_t scratch_off, conv_off;
    };
    std::vector<Pending> pend;
    int ndirect = 0, nfour = 0, nconvs = 0, nconvl = 0;
    // The plans are built on several host threads: their twiddle, chirp and permutation tables
    // are computed on the host in long double, which dominates the planning of thousands of
    // lengths.  Plan building shares only mutex-guarded state (plan and line-stride caches);
    // each thread makes the context's device current first.
    std::vector<std::pair<int64_t, std::vector<int64_t>>> lens(bylen.begin(), bylen.end());
    std::vector<std::unique_ptr<Plan>> built(lens.size());
    std::vector<int> codes(lens.size(), CFFT_OK);
    std::vector<std::string> msgs(lens.size());
    std::atomic<size_t> next{0};
    std::atomic<int> setup_rc{CFFT_OK};
    auto build_one = [&](size_t k) {
        const int64_t n = lens[k].first;
        std::unique_ptr<Plan> p(new Plan());
        p->ctx = ctx;
        p->n = n;
        p->batch = (int64_t)lens[k].second.size();
        p->idist = p->odist = n;
        p->prec = pr;
        p->dtype = dtype;
        p->flags = flags;
        int rc;
        if (n == 1) {
            if ((rc = make_line(p->line, 1, pr, 16, 1024))) return rc;
            p->kind = Kind::Direct;
            p->C = 256;
        } else {
            // lines planned for the work-list kernels' per-thread capacity, so that a workgroup
            // has the threads those kernels use (a line planned for E = 8 would get twice as many)
            PlanOptions o;
            o.E = kMultiEHost;
            o.min_threads = kFsMinThreads;
            if ((rc = plan_build(ctx, *p, o))) return rc;
            if (p->kind == Kind::Large || (p->sub && p->sub->kind == Kind::Large))
                return fail(CFFT_ERR_UNSUPPORTED, "uneven batches take lengths the four-step "
                                                  "covers (n=" + std::to_string(n) + " needs six steps)");
        }
        built[k] = std::move(p);
        return (int)CFFT_OK;
    };
    auto worker = [&]() {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            setup_rc = CFFT_ERR_HIP;
            return;
        }
        for (size_t k; (k = next++) < lens.size();) {
            codes[k] = build_one(k);
            if (codes[k]) msgs[k] = cfft_last_error();
        }
    };
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const size_t nthr = std::min<size_t>(lens.size(), std::max(1u, std::min(8u, hw)));
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nthr; ++t) {
            try {
                pool.emplace_back(worker);
            } catch (const std::system_error&) {
                break;  // no more threads: the ones running (and this one) take the rest
            }
        }
        worker();
        for (auto& th : pool) th.join();
    }
    if (setup_rc != CFFT_OK) return fail(setup_rc, "hipSetDevice failed on a planning thread");
    for (size_t k = 0; k < lens.size(); ++k)
        if (codes[k]) return fail(codes[k], msgs[k]);
    for (size_t k = 0; k < lens.size(); ++k) {
        auto& kv = lens[k];
        const int64_t n = kv.first;
        const int64_t m = (int64_t)kv.second.size();
        std::unique_ptr<Plan> p = std::move(built[k]);
        Pending q;
        q.ids = kv.second;
        q.scratch_off = q.conv_off = 0;
        if (p->kind == Kind::FourStep) {
            q.scratch_off = work_bytes;
            work_bytes += (size_t)m * n * es;
            ++nfour;
        } else if (p->kind == Kind::ConvLarge) {
            ++nconvl;
        } else if (p->kind == Kind::ConvSmall) {
            ++nconvs;
        } else {
            ++ndirect;
        }
        st->plans.push_back(std::move(p));
        pend.push_back(std::move(q));
    }
    int rc;
    if ((rc = st->work.alloc(std::max<size_t>(work_bytes, 1)))) return rc;
    for (size_t k = 0; k < pend.size(); ++k) {
        const Plan& p = *st->plans[k];
        Pending& q = pend[k];
        q.e = W.add_plan(p);
        std::vector<XformRef> xs;
        const size_t ces = esz(p.cprec);
        for (size_t j = 0; j < q.ids.size(); ++j) {
            const int64_t id = q.ids[j];
            XformRef x;
            x.in = (uint64_t)offsets[id] * es;
            x.out = (uint64_t)offsets[id] * es;
            x.in_sel = kInBase;
            x.out_sel = kOutBase;
            x.launch_dir = true;
            if (p.kind == Kind::FourStep)
                x.scratch = (uint64_t)st->work.p + q.scratch_off + (uint64_t)j * p.n * es;
            if (p.kind == Kind::ConvLarge) {  // the length's own work buffers, line j
                x.w1 = (uint64_t)p.w1.p + (uint64_t)j * p.L * ces;
                x.w2 = (uint64_t)p.w2.p + (uint64_t)j * p.L * ces;
                x.dc = (uint64_t)p.dc.p + (uint64_t)j * 2 * ces;
            }
            xs.push_back(x);
        }
        W.add_transforms(q.e, xs);
    }
    if ((rc = W.finalize(st->dplans, st->ditems, st->dtable, st->launches))) return rc;
    CFFT_HIP(hipStreamCreateWithFlags(&st->cap, hipStreamNonBlocking));
    size_t bytes = st->dplans.bytes + st->ditems.bytes + st->dtable.bytes + st->work.bytes;
    for (auto& p : st->plans) bytes += plan_device_bytes(*p);
    P->uneven_bytes = bytes;
    std::ostringstream os;
    os << "uneven count=" << count << " lengths=" << bylen.size() << " (direct " << ndirect
       << ", four-step " << nfour << ", fused conv " << nconvs << ", large conv " << nconvl
       << ") launches=" << st->launches.size() << " items=" << W.item_count()
       << (st->use_graph ? " graph" : "");
    P->desc = os.str();
    P->uneven = st;
    *out = P;
    return CFFT_OK;
}

int uneven_exec(const Plan& p, const void* in, void* out, int direction, int normalize,
                hipStream_t stream) {
    auto* st = static_cast<UnevenState*>(p.uneven.get());
    if (!st) return fail(CFFT_ERR_INVALID, "not an uneven plan");
    if (p.batch == 0) return CFFT_OK;
    if (!in || !out) return fail(CFFT_ERR_INVALID, "null buffer");
    MultiArgs ma{(uint64_t)in, (uint64_t)out, direction > 0 ? -1 : 1, normalize ? 1 : 0};
    if (!st->use_graph) return run_launches(st->launches, ma, stream);
    std::lock_guard<std::mutex> g(st->mu);
    if (!st->gexec || st->g_in != ma.in_base || st->g_out != ma.out_base || st->g_sgn != ma.sgn ||
        st->g_norm != ma.normalize) {
        if (st->gexec) (void)hipGraphExecDestroy(st->gexec);
        if (st->graph) (void)hipGraphDestroy(st->graph);
        st->gexec = nullptr;
        st->graph = nullptr;
        CFFT_HIP(hipStreamBeginCapture(st->cap, hipStreamCaptureModeThreadLocal));
        int rc = run_launches(st->launches, ma, st->cap);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st->cap, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(CFFT_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        st->graph = graph;
        CFFT_HIP(hipGraphInstantiate(&st->gexec, st->graph, nullptr, nullptr, 0));
        st->g_in = ma.in_base;
        st->g_out = ma.out_base;
        st->g_sgn = ma.sgn;
        st->g_norm = ma.normalize;
    }
    CFFT_HIP(hipGraphLaunch(st->gexec, stream));
    return CFFT_OK;
}

}  // namespace cfft
