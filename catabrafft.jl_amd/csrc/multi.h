// multi.h -- work-list descriptors for launches that mix transforms of different lengths.
//
// A launch gets a table of MPlan (one per distinct transform plan) and a table of MItem (one
// per workgroup).  Workgroup b reads items[b], the plan it names, and runs that plan's lines on
// the item's buffers: one launch covers transforms of uneven lengths, each with its own pass
// table, twiddles and tile shape.  Items are sorted by estimated cost (longest first) so that
// the hardware dispatcher, which hands out workgroups in index order as slots free up, balances
// the launch.
#pragma once
#include <cstdint>

#include "engine.h"

namespace cfft {

// A line FFT as device-visible data (the PassTable is read with uniform scalar loads).
struct MLine {
    int N, Tt, Ls, pad;
    PassTable pt;
    const void* tw;
};

struct MPlan {
    int n, C;                   // transform length, lines per workgroup (rows / fused conv)
    int CC, LsC;                // large convolution in fused launches: the last launch's tile
                                // and LDS line stride (its FFT_L four-step fields below)
    MLine line;                 // rows: the transform; fused conv and conv sub-FFTs: length L
    // four-step pass A (la, N1, column tile CA) and pass B (lb, N2, row tile CB)
    int N1, N2, CA, CB, tbits, tnc;
    MLine la, lb;
    const void* tw3;            // pass B's three-level twiddle table (see FourStepBArgs)
    // convolution (Rader / Bluestein)
    int mode, L, Ls;            // Ls: LDS line stride (input and convolution work share it)
    const int* perm;
    const int* iperm;
    const void* chirp;
    const void* K;              // epilogue spectrum of a convolution's forward FFT_L
};

// Address selectors: an item address is absolute (0) or an offset from the launch's input (1)
// or output (2) base address.
enum : int { kAbs = 0, kInBase = 1, kOutBase = 2 };

struct MItem {
    int plan;       // index into the MPlan table
    int count;      // rows / fused conv: lines in this item; four-step: tile index;
                    // gather / scatter: element chunk index
    int sgn;        // -1 conjugates on load and store (backward)
    int flags;      // MItemFlags
    double scale;   // output scale
    uint64_t in, out;      // first line (or the transform) in / out
    uint64_t aux;          // kLineTable: address of the line table; scatter: unused
    uint64_t dc;           // convolution slots of the (first) line: {A^[0], x0}
    int64_t idist, odist;  // element distance between the item's lines (no line table)
};

enum MItemFlags : int {
    kSelInMask = 0xF,          // bits 0-3: selector of `in` (and of the table's in entries)
    kSelOutShift = 4,          // bits 4-7: selector of `out` (and of the table's out entries)
    kEpilogue = 1 << 8,        // forward half of a convolution: v -> conj(v K), dc = A^[0]
    kLaunchDir = 1 << 9,       // sign and 1/n normalisation come from the launch arguments
    kLineTable = 1 << 10,      // line c uses table entry {in, out, dc} at aux + 24 c
};

__host__ __device__ inline uint64_t mresolve(uint64_t a, int sel, uint64_t in_base,
                                             uint64_t out_base) {
    return a + (sel == kInBase ? in_base : (sel == kOutBase ? out_base : 0));
}

// Elements per gather / scatter work item (host copy of the device constant kChunk).
constexpr int64_t kChunkHost = 4096;
// Per-thread capacity E of the work-list kernels (host copy of the device constant kMultiE):
// plans whose lines run in work lists are made with it.
constexpr int kMultiEHost = 16;

// One launch of a work-list kernel.
struct MultiLaunch {
    int kernel;          // MultiKernel
    int prec, cprec;     // Prec of the data (and of the work buffer for gather / scatter)
    int block;           // threads per workgroup
    size_t lds;          // dynamic LDS bytes
    const MItem* items;  // device table
    int64_t nitems;
    const MPlan* plans;  // device table
};

enum class MultiKernel : int {
    Rows = 0, FourStepA, FourStepB, ConvSmall, Gather, Scatter,
    ConvLA, ConvLBA, ConvLC  // the three fused launches of a large convolution
};

// Launch-wide arguments of a work-list launch.
struct MultiArgs {
    uint64_t in_base, out_base;  // bases of the kInBase / kOutBase selectors
    int sgn, normalize;          // for items flagged kLaunchDir
};

hipError_t launch_multi(const MultiLaunch& L, const MultiArgs& a, hipStream_t st);

// Per-precision entry points (kernels_f32.hip / kernels_f64.hip).
hipError_t multi_f32(const MultiLaunch& L, const MultiArgs& a, hipStream_t st);
hipError_t multi_f64(const MultiLaunch& L, const MultiArgs& a, hipStream_t st);
hipError_t gs_multi_f32_f32(const MultiLaunch& L, const MultiArgs& a, hipStream_t st);
hipError_t gs_multi_f32_f64(const MultiLaunch& L, const MultiArgs& a, hipStream_t st);
hipError_t gs_multi_f64_f64(const MultiLaunch& L, const MultiArgs& a, hipStream_t st);

}  // namespace cfft
