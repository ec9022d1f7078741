// Standalone grid build on caller-owned device buffers: the cell hash, the
// stable (cell, index) sort and the cell bounds of DEMGrid::NeighborSearch
// (src/dem_grid.cu:649-678), with the same kernels as a context's rebuild.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/nbc.h"
#include "nbc_kernels.h"

namespace nbc {
void set_global_error(const std::string& m);
}

using namespace nbc;

namespace {

constexpr int RBITS = 8;  // radix digit of the cell sort (4 passes over 32-bit keys at most)

int bits_for(long long maxkey) {
    int b = 1;
    while (b < 31 && (1ll << b) <= maxkey) ++b;
    return b;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Temporary buffers of the sort for n keys, carved out of one block.
struct GridScratch {
    uint32_t* tk;   // second key buffer of the ping-pong
    int* tv;        // second value buffer
    int* hist;      // per-tile digit counts
    int* hscan;     // their exclusive scan
    int* stmp;      // scan block sums
    int* stotal;    // scan total
};

size_t grid_scratch_layout(long long n, char* base, GridScratch* g) {
    const size_t nh = (size_t)radix_tmp_ints((int)n, RBITS);
    const size_t sizes[6] = {sizeof(uint32_t) * (size_t)n, sizeof(int) * (size_t)n, sizeof(int) * nh,
                             sizeof(int) * nh, sizeof(int) * (size_t)scan_tmp_ints((int)nh), sizeof(int) * 4};
    size_t off = 0;
    void* p[6];
    for (int k = 0; k < 6; ++k) {
        p[k] = base ? base + off : nullptr;
        off += align256(sizes[k]);
    }
    if (g) {
        g->tk = (uint32_t*)p[0]; g->tv = (int*)p[1]; g->hist = (int*)p[2]; g->hscan = (int*)p[3];
        g->stmp = (int*)p[4]; g->stotal = (int*)p[5];
    }
    return off;
}

int fail(int code, const std::string& m) {
    set_global_error(m);
    return code;
}

bool is_device_ptr(const void* p) {
    hipPointerAttribute_t a;
    std::memset(&a, 0, sizeof(a));
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Arguments shared by both entry points; fills the hash parameters.
int check_args(const char* who, const nbc_params* p, const double* pos4, long long n, const uint32_t* cell,
               const int32_t* perm, const int32_t* beg, const int32_t* end, DevParams* dp) {
    const std::string w(who);
    if (!p || !pos4 || !cell || !perm || !beg || !end) return fail(NBC_E_BADINPUT, w + ": null argument");
    if (!p->finalized) return fail(NBC_E_BADINPUT, w + ": call nbc_finalize_params first");
    if (n < 1 || n > 0x7fffffffll) return fail(NBC_E_BADINPUT, w + ": sphere count out of range");
    if (p->CellNum < 1) return fail(NBC_E_BADINPUT, w + ": grid has no cells");
    const void* ptrs[5] = {pos4, cell, perm, beg, end};
    for (const void* q : ptrs)
        if (!is_device_ptr(q)) return fail(NBC_E_BADINPUT, w + ": every buffer must be device memory");
    std::memset(dp, 0, sizeof(*dp));
    dp->N = (int)n;
    dp->CellNum = p->CellNum;
    for (int a = 0; a < 3; ++a) {
        dp->G[a] = p->GridDim[a];
        dp->origin[a] = p->WorldOrigin[a];
        dp->cell[a] = p->CellSize[a];
        dp->icell[a] = p->i_CellSize[a];
        dp->boxlo[a] = p->BoxLowLmt[a];
        dp->boxup[a] = p->BoxUpLmt[a];
    }
    dp->L = p->Levels > 1 ? p->Levels : 1;
    if (dp->L > MAX_LEVELS) return fail(NBC_E_BADINPUT, w + ": too many grid levels");
    dp->nested = dp->L > 1 && p->LevelNested ? 1 : 0;
    dp->lblock = dp->nested ? p->LevelBlock : 0;
    if (dp->nested && (dp->L > 4 || p->LevelBlock < 1)) return fail(NBC_E_BADINPUT, w + ": bad nested grid table");
    for (int t = 0; t < dp->L && dp->L > 1; ++t) {
        dp->loff[t] = p->LevelOff[t];
        for (int a = 0; a < 3; ++a) dp->ldim[t][a] = p->LevelDim[t][a];
        dp->lcell[t] = p->LevelCell[t];
        dp->lrmin[t] = p->LevelRmin[t];
    }
    return 0;
}

// hash -> stable LSD radix sort -> cell bounds, all on `st`
int grid_build(const DevParams& dp, const double* pos4, uint32_t* cell, int32_t* perm, int32_t* beg, int32_t* end,
               const GridScratch& g, hipStream_t st) {
    const int n = dp.N;
    int e = launch_hash(st, dp, (const D4*)pos4, nullptr, cell, perm);
    if (e) return e;
    int second = 0;
    e = launch_radix_sort(st, cell, perm, g.tk, g.tv, n, bits_for(dp.CellNum - 1), RBITS, g.hist, g.hscan, g.stmp,
                          g.stotal, &second);
    if (e) return e;
    if (second) {
        hipError_t h = hipMemcpyAsync(cell, g.tk, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, st);
        if (h == hipSuccess) h = hipMemcpyAsync(perm, g.tv, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st);
        if (h != hipSuccess) return (int)h;
    }
    return launch_seg_bounds(st, cell, n, dp.CellNum, beg, end);
}

int hip_fail(const char* who, int h) {
    return fail(NBC_E_HIP, std::string(who) + ": " + hipGetErrorString((hipError_t)h));
}

}  // namespace

extern "C" int nbc_grid_build(const nbc_params* p, const double* d_pos4, int32_t n, uint32_t* d_cell,
                              int32_t* d_perm, int32_t* d_beg, int32_t* d_end, void* stream) {
    DevParams dp;
    int e = check_args("nbc_grid_build", p, d_pos4, n, d_cell, d_perm, d_beg, d_end, &dp);
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    const size_t bytes = grid_scratch_layout(n, nullptr, nullptr);
    char* buf = nullptr;
    hipError_t h = hipMalloc((void**)&buf, bytes);
    if (h != hipSuccess) return hip_fail("nbc_grid_build: hipMalloc", h);
    GridScratch g;
    grid_scratch_layout(n, buf, &g);
    e = grid_build(dp, d_pos4, d_cell, d_perm, d_beg, d_end, g, st);
    // the temporaries live until the stream has used them
    h = hipStreamSynchronize(st);
    (void)hipFree(buf);
    if (e) return hip_fail("nbc_grid_build: launch", e);
    if (h != hipSuccess) return hip_fail("nbc_grid_build: hipStreamSynchronize", h);
    return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free variant for callers that cannot call hipMalloc (code under
// HIP graph capture, tight serving loops): the sort's temporaries live in a
// file-scope device array of fixed size, one copy per GPU for the life of
// the process. Every scratch word a kernel reads was written earlier in the
// same call (digit counts, their scan, block sums, the second key/value
// buffer), so nothing carries over from a previous call and nothing needs
// clearing. All calls on one GPU share the array: calls that overlap in time
// on one GPU (concurrent streams, threads) race and must be serialized by
// the caller.
// 4 buffers of NBC_GRID_NOALLOC_MAX_SPHERES ints (second keys, second
// values, digit counts and their scan at 8-bit digits) plus the scan block
// sums and totals: 32 MiB + 64 KiB
static __device__ __align__(256) unsigned char scratch[4 * sizeof(int) * (size_t)NBC_GRID_NOALLOC_MAX_SPHERES +
                                                        64 * 1024];

namespace {

constexpr int32_t NOALLOC_MAX = NBC_GRID_NOALLOC_MAX_SPHERES;
constexpr size_t SCRATCH_BYTES = sizeof(scratch);

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU holds its own copy).
int scratch_on_current_device(void** out) {
    int dev = -1;
    hipError_t h = hipGetDevice(&dev);
    if (h != hipSuccess) return hip_fail("nbc_grid_build_noalloc: hipGetDevice", h);
    if (dev < 0 || dev >= MAX_DEVICES) return fail(NBC_E_NODEVICE, "nbc_grid_build_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        h = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (h != hipSuccess || !p) return hip_fail("nbc_grid_build_noalloc: hipGetSymbolAddress", h);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" int nbc_grid_build_noalloc(const nbc_params* p, const double* d_pos4, int32_t n, uint32_t* d_cell,
                                      int32_t* d_perm, int32_t* d_beg, int32_t* d_end, void* stream) {
    DevParams dp;
    int e = check_args("nbc_grid_build_noalloc", p, d_pos4, n, d_cell, d_perm, d_beg, d_end, &dp);
    if (e) return e;
    if (n > NOALLOC_MAX) {
        char b[160];
        std::snprintf(b, sizeof b, "nbc_grid_build_noalloc: %d spheres exceed the scratch limit of %d", n, NOALLOC_MAX);
        return fail(NBC_E_BADINPUT, b);
    }
    if (grid_scratch_layout(n, nullptr, nullptr) > SCRATCH_BYTES)
        return fail(NBC_E_BADINPUT, "nbc_grid_build_noalloc: scratch layout exceeds the static array");
    void* base = nullptr;
    e = scratch_on_current_device(&base);
    if (e) return e;
    GridScratch g;
    grid_scratch_layout(n, (char*)base, &g);
    e = grid_build(dp, d_pos4, d_cell, d_perm, d_beg, d_end, g, (hipStream_t)stream);
    if (e) return hip_fail("nbc_grid_build_noalloc: launch", e);
    return 0;
}
