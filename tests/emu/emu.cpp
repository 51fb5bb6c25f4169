// CPU emulation runtime for tests/emu/include/hip/hip_runtime.h (TEST INFRASTRUCTURE ONLY).
#include <hip/hip_runtime.h>
#include <ucontext.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace emu {

struct Fiber {
    ucontext_t ctx;
    bool done = false;
    Idx tid{0, 0, 0};
};

struct State {
    Idx thread{0, 0, 0}, block{0, 0, 0}, bdim{1, 1, 1}, gdim{1, 1, 1};
    std::vector<unsigned char> lds;
    ucontext_t main;
    std::vector<Fiber> fibers;
    std::vector<std::vector<char>> stacks;
    int cur = -1;
    const std::function<void()>* body = nullptr;
};
static thread_local State st;

Idx& thread_idx() { return st.thread; }
Idx& block_idx() { return st.block; }
Idx& block_dim() { return st.bdim; }
Idx& grid_dim() { return st.gdim; }
unsigned char* smem() { return st.lds.data(); }

void barrier() {
    if (st.cur < 0) return;  // not inside a launch
    swapcontext(&st.fibers[st.cur].ctx, &st.main);
}

static void entry(int idx) {
    (*st.body)();
    st.fibers[idx].done = true;
}

struct Stream {
    bool capturing = false;
    std::vector<std::function<void()>>* graph = nullptr;
};

static void run_block(const std::function<void()>& body, unsigned nthreads) {
    const size_t stack = 256 * 1024;
    if (st.stacks.size() < nthreads) st.stacks.resize(nthreads);
    st.fibers.assign(nthreads, Fiber());
    st.body = &body;
    for (unsigned i = 0; i < nthreads; ++i) {
        if (st.stacks[i].size() < stack) st.stacks[i].resize(stack);
        Fiber& f = st.fibers[i];
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = st.stacks[i].data();
        f.ctx.uc_stack.ss_size = stack;
        f.ctx.uc_link = &st.main;
        makecontext(&f.ctx, (void (*)())entry, 1, (int)i);
        const unsigned bx = st.bdim.x, by = st.bdim.y;
        f.tid = Idx{i % bx, (i / bx) % by, i / (bx * by)};
    }
    for (;;) {
        bool any = false;
        for (unsigned i = 0; i < nthreads; ++i) {
            Fiber& f = st.fibers[i];
            if (f.done) continue;
            any = true;
            st.cur = (int)i;
            st.thread = f.tid;
            swapcontext(&st.main, &f.ctx);
        }
        if (!any) break;
    }
    st.cur = -1;
}

void launch(std::function<void()> body, dim3 grid, dim3 block, size_t lds, Stream* s) {
    if (s && s->capturing) {
        s->graph->push_back([=]() { launch(body, grid, block, lds, nullptr); });
        return;
    }
    if (block.x * block.y * block.z > 1024) {
        std::fprintf(stderr, "emu: block of %u threads\n", block.x * block.y * block.z);
        std::abort();
    }
    st.gdim = Idx{grid.x, grid.y, grid.z};
    st.bdim = Idx{block.x, block.y, block.z};
    st.lds.assign(lds + 64, 0xCD);  // garbage-filled like real LDS
    for (unsigned z = 0; z < grid.z; ++z)
        for (unsigned y = 0; y < grid.y; ++y)
            for (unsigned x = 0; x < grid.x; ++x) {
                st.block = Idx{x, y, z};
                run_block(body, block.x * block.y * block.z);
            }
}

}  // namespace emu

int cfft_emu_lane_xor(int v, int d) {
    static thread_local std::vector<int> buf;
    const unsigned n = emu::block_dim().x * emu::block_dim().y * emu::block_dim().z;
    if (buf.size() < n) buf.resize(n);
    const unsigned i = emu::thread_idx().x;
    const unsigned j = i ^ (unsigned)d;
    if (j >= n || (i >> 6) != (j >> 6)) {
        std::fprintf(stderr, "emu: lane exchange %u ^ %d leaves the wavefront\n", i, d);
        std::abort();
    }
    buf[i] = v;
    emu::barrier();
    const int r = buf[j];
    emu::barrier();
    return r;
}

struct emu_event {
    double t = 0;
};
struct emu_graph {
    std::vector<std::function<void()>> ops;
};
struct emu_graph_exec {
    std::vector<std::function<void()>> ops;
};

static double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    // 2 CUs: persistent kernels then cover the tests' small batches with a few workgroups that
    // each loop over several line groups
    *v = a == hipDeviceAttributeMaxSharedMemoryPerBlock ? 160 * 1024 : 2;
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
// Resident workgroups per CU as the LDS allows (160 KiB per CU), at most 2: small enough that
// the emulated grids of the tests loop several times in the persistent kernels.
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t lds) {
    const size_t per_cu = 160 * 1024;
    int v = lds ? (int)(per_cu / lds) : 2;
    *n = v < 2 ? v : 2;
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emulated error"; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
    *p = std::aligned_alloc(256, (n + 255) / 256 * 256);
    if (*p) std::memset(*p, 0xAB, n);  // uninitialised device memory is not zero
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h,
                            hipMemcpyKind, hipStream_t) {
    for (size_t r = 0; r < h; ++r) std::memmove((char*)d + r * dp, (const char*)s + r * sp, w);
    return hipSuccess;
}
hipError_t hipMemset(void* p, int v, size_t n) { std::memset(p, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t) { return hipMemset(p, v, n); }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = new emu::Stream(); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new emu_event(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = now_ms(); return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(b->t - a->t); return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!s) return hipErrorInvalidValue;
    s->capturing = true;
    s->graph = new std::vector<std::function<void()>>();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    *g = new emu_graph();
    (*g)->ops = std::move(*s->graph);
    delete s->graph;
    s->graph = nullptr;
    s->capturing = false;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new emu_graph_exec();
    (*e)->ops = g->ops;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto& op : e->ops) op();
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
