// Autotuner: times candidate options for one bed through the public C ABI.
//
// Coordinate descent: each parameter in turn is set to every candidate value
// with the others at their best so far; a candidate is a fresh context
// (finalize, create, upload), warm-up steps including the first rebuild, then
// `steps` timed steps closed by a stream synchronize.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nbc.h"

namespace nbc {
void set_global_error(const std::string& m);
}

namespace {

// One bed, or nbeds beds advanced by one batch context.
struct Bed {
    int nbeds;
    const nbc_params* params;     // nbeds blocks, not finalized
    const nbc_spheres_in* s;      // nbeds
    const nbc_clumps_in* c;       // nbeds
    const nbc_wall* wall;         // nbeds (or one; may be NULL)
    const nbc_penetrator* pene;   // single bed only
    int device;
    int steps;
};

// ms per step of one candidate, < 0 when it fails
double run_candidate(const Bed& b, const nbc_options& o) {
    std::vector<nbc_params> p(b.params, b.params + b.nbeds);
    for (int k = 0; k < b.nbeds; ++k)
        if (nbc_finalize_params(&p[k], b.s[k].radius, b.s[k].n, &o) != 0) return -1.0;
    nbc_ctx* ctx = nullptr;
    const int ec = b.nbeds > 1 ? nbc_create_batch(&ctx, b.nbeds, p.data(), &o, b.device)
                               : nbc_create(&ctx, p.data(), &o, b.device);
    if (ec != 0) return -1.0;
    double ms = -1.0;
    const int warm = b.steps / 10 > 20 ? b.steps / 10 : 20;
    const int eu = b.nbeds > 1 ? nbc_upload_batch(ctx, b.s, b.c, b.wall) : nbc_upload(ctx, b.s, b.c, b.wall, b.pene);
    if (eu == 0 && nbc_step(ctx, warm) == 0 && nbc_synchronize(ctx) == 0) {
        const auto t0 = std::chrono::steady_clock::now();
        if (nbc_step(ctx, b.steps) == 0 && nbc_synchronize(ctx) == 0) {
            const auto t1 = std::chrono::steady_clock::now();
            ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / b.steps;
        }
    }
    nbc_destroy(ctx);
    return ms;
}

int tune(const Bed& b, const nbc_options* base, int32_t flags, nbc_options* best, nbc_tune_entry* entries,
         int32_t cap, int32_t* n_entries) {
    int32_t ndev = 0;
    if (nbc_device_count(&ndev) != 0 || ndev <= 0) return NBC_E_NODEVICE;
    nbc_options cur;
    if (base) cur = *base; else std::memset(&cur, 0, sizeof(cur));
    double rmax = 0.0;
    for (int k = 0; k < b.nbeds; ++k)
        for (int i = 0; i < b.s[k].n; ++i) rmax = std::max(rmax, b.s[k].radius[i]);
    int count = 0;
    double best_ms = -1.0;
    auto record = [&](const nbc_options& o, double ms) {
        if (entries && count < cap) { entries[count].opt = o; entries[count].ms_per_step = ms; }
        ++count;
    };
    auto sweep = [&](auto set, const std::vector<double>& values) {
        nbc_options pick = cur;
        double pick_ms = best_ms;
        for (double v : values) {
            nbc_options o = cur;
            set(o, v);
            const double ms = run_candidate(b, o);
            record(o, ms);
            if (ms > 0.0 && (pick_ms < 0.0 || ms < pick_ms)) { pick = o; pick_ms = ms; }
        }
        cur = pick;
        best_ms = pick_ms;
    };
    // near_skin 0: the per-sphere skin (near_horizon), else a fixed one
    sweep([](nbc_options& o, double v) { o.near_skin = v; }, {0.0, 0.125, 0.25, 0.5, 1.0});
    sweep([](nbc_options& o, double v) { o.batch_steps = (int32_t)v; }, {2, 4, 8, 16});
    sweep([](nbc_options& o, double v) { o.radix_bits = (int32_t)v; }, {4, 6, 8});
    sweep([](nbc_options& o, double v) { o.block_size = (int32_t)v; }, {128, 256});
    sweep([](nbc_options& o, double v) { o.graphs = (int32_t)v; }, {-1, 1});
    if (flags & NBC_TUNE_CELL)
        sweep([rmax](nbc_options& o, double v) { o.cell_size = v * 3.0 * rmax; o.hgrid_levels = 0; },
              {1.0, 1.25, 1.5});
    if ((flags & NBC_TUNE_HGRID) && b.nbeds == 1)
        // negative: nested levels
        sweep([](nbc_options& o, double v) {
                  o.hgrid_levels = (int32_t)std::fabs(v);
                  o.hgrid_nested = v < 0.0 ? 1 : 0;
                  if (o.hgrid_levels > 1) o.cell_size = 0.0;
              },
              {0, 2, 3, -2, -3});
    if (n_entries) *n_entries = count;
    if (best_ms < 0.0) {
        nbc::set_global_error("nbc_autotune: no candidate ran");
        return NBC_E_BADINPUT;
    }
    *best = cur;
    return 0;
}

int check_raw(const nbc_params* p, int n, const char* who) {
    for (int k = 0; k < n; ++k)
        if (p[k].finalized) {
            nbc::set_global_error(std::string(who) + ": pass world.par values before nbc_finalize_params");
            return NBC_E_BADINPUT;
        }
    return 0;
}

}  // namespace

extern "C" int nbc_autotune(const nbc_params* params, const nbc_options* base, const nbc_spheres_in* s,
                            const nbc_clumps_in* c, const nbc_wall* wall, const nbc_penetrator* pene, int32_t device,
                            int32_t steps, int32_t flags, nbc_options* best, nbc_tune_entry* entries, int32_t cap,
                            int32_t* n_entries) {
    if (!params || !s || !c || !best) {
        nbc::set_global_error("nbc_autotune: null argument");
        return NBC_E_BADINPUT;
    }
    if (int e = check_raw(params, 1, "nbc_autotune")) return e;
    const Bed b{1, params, s, c, wall, pene, device, steps > 0 ? steps : 300};
    return tune(b, base, flags, best, entries, cap, n_entries);
}

extern "C" int nbc_autotune_batch(int32_t nbeds, const nbc_params* params, const nbc_options* base,
                                  const nbc_spheres_in* s, const nbc_clumps_in* c, const nbc_wall* walls,
                                  int32_t device, int32_t steps, int32_t flags, nbc_options* best,
                                  nbc_tune_entry* entries, int32_t cap, int32_t* n_entries) {
    if (nbeds < 1 || !params || !s || !c || !best) {
        nbc::set_global_error("nbc_autotune_batch: null argument or no beds");
        return NBC_E_BADINPUT;
    }
    if (int e = check_raw(params, nbeds, "nbc_autotune_batch")) return e;
    const Bed b{nbeds, params, s, c, walls, nullptr, device, steps > 0 ? steps : 300};
    return tune(b, base, flags & ~NBC_TUNE_HGRID, best, entries, cap, n_entries);
}
