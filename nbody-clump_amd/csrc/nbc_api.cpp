// extern "C" entry points of libnbc that touch the device (see include/nbc.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

#include "../../include/nbc.h"
#include "nbc_engine.h"

struct nbc_ctx {
    nbc::EngineBase* eng = nullptr;
};

namespace {
thread_local std::string g_err;

int ctx_err(nbc_ctx* c, int code) {
    if (code && c && c->eng) g_err = c->eng->err;
    return code;
}
}  // namespace

namespace nbc {
void set_global_error(const std::string& m) { g_err = m; }
}

extern "C" {

const char* nbc_version(void) { return "nbc 0.1 (gfx950, HIP)"; }

const char* nbc_last_error(nbc_ctx* ctx) {
    if (ctx && ctx->eng && !ctx->eng->err.empty()) return ctx->eng->err.c_str();
    return g_err.c_str();
}

int nbc_device_count(int32_t* out) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        if (out) *out = 0;
        g_err = std::string("hipGetDeviceCount: ") + hipGetErrorString(e);
        return NBC_E_NODEVICE;
    }
    if (out) *out = n;
    return 0;
}

int nbc_create(nbc_ctx** out, const nbc_params* params, const nbc_options* opt, int32_t device) {
    if (!out || !params) { g_err = "nbc_create: null argument"; return NBC_E_BADINPUT; }
    *out = nullptr;
    if (!params->finalized) { g_err = "nbc_create: call nbc_finalize_params first"; return NBC_E_BADINPUT; }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_err = "nbc_create: no HIP device available";
        return NBC_E_NODEVICE;
    }
    if (device < 0 || device >= n) { g_err = "nbc_create: device index out of range"; return NBC_E_NODEVICE; }
    nbc_options o;
    if (opt) o = *opt; else std::memset(&o, 0, sizeof(o));
    nbc_ctx* c = new (std::nothrow) nbc_ctx();
    if (!c) { g_err = "nbc_create: out of host memory"; return NBC_E_BADINPUT; }
    c->eng = nbc::make_engine(o.precision);
    int e = c->eng->init(*params, o, device);
    if (e) {
        g_err = c->eng->err;
        delete c->eng;
        delete c;
        return e;
    }
    *out = c;
    return 0;
}

int nbc_create_batch(nbc_ctx** out, int32_t nbeds, const nbc_params* params, const nbc_options* opt,
                     int32_t device) {
    if (!out || !params || nbeds < 1) { g_err = "nbc_create_batch: null argument"; return NBC_E_BADINPUT; }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        g_err = "nbc_create_batch: no HIP device available";
        return NBC_E_NODEVICE;
    }
    if (device < 0 || device >= n) { g_err = "nbc_create_batch: device index out of range"; return NBC_E_NODEVICE; }
    nbc_options o;
    if (opt) o = *opt; else std::memset(&o, 0, sizeof(o));
    nbc_ctx* c = new (std::nothrow) nbc_ctx();
    if (!c) { g_err = "nbc_create_batch: out of host memory"; return NBC_E_BADINPUT; }
    c->eng = nbc::make_engine(o.precision);
    int e = c->eng->init_batch(params, nbeds, o, device);
    if (e) {
        g_err = c->eng->err;
        delete c->eng;
        delete c;
        return e;
    }
    *out = c;
    return 0;
}

int nbc_destroy(nbc_ctx* ctx) {
    if (!ctx) return 0;
    delete ctx->eng;
    delete ctx;
    return 0;
}

#define ENG(ctx)                                                        \
    if (!(ctx) || !(ctx)->eng) { g_err = "null context"; return NBC_E_BADINPUT; }

int nbc_upload(nbc_ctx* ctx, const nbc_spheres_in* s, const nbc_clumps_in* c, const nbc_wall* wall,
               const nbc_penetrator* pene) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->upload(s, c, wall, pene));
}
int nbc_upload_batch(nbc_ctx* ctx, const nbc_spheres_in* spheres, const nbc_clumps_in* clumps,
                     const nbc_wall* walls) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->upload_batch(spheres, clumps, walls));
}
int nbc_upload_particles(nbc_ctx* ctx, const nbc_particles_in* p, const nbc_wall* wall,
                         const nbc_penetrator* pene) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->upload_particles(p, wall, pene));
}
int nbc_get_particles(nbc_ctx* ctx, int32_t* sorted_id, int32_t* file_id, int32_t* type, double* pos, double* vel,
                      double* quat, double* angv, double* prop, double* radius) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_particles(sorted_id, file_id, type, pos, vel, quat, angv, prop, radius));
}
int nbc_get_batch_ranges(nbc_ctx* ctx, int32_t* sphere_off, int32_t* clump_off) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_batch_ranges(sphere_off, clump_off));
}
int nbc_step(nbc_ctx* ctx, int32_t nsteps) {
    ENG(ctx);
    if (nsteps < 0) return NBC_E_BADINPUT;
    return ctx_err(ctx, ctx->eng->step(nsteps));
}
int nbc_rebuild(nbc_ctx* ctx) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->force_rebuild());
}
int nbc_set_record(nbc_ctx* ctx, int32_t on) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->set_record(on));
}
int nbc_synchronize(nbc_ctx* ctx) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->sync());
}
int nbc_get_stats(nbc_ctx* ctx, nbc_stats* out) {
    ENG(ctx);
    if (!out) return NBC_E_BADINPUT;
    return ctx_err(ctx, ctx->eng->get_stats(out));
}
int nbc_get_spheres(nbc_ctx* ctx, int32_t* file_id, int32_t* clump_id, double* pos, double* cum, double* local,
                    double* mass, double* radius) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_spheres(file_id, clump_id, pos, cum, local, mass, radius));
}
int nbc_get_grid(nbc_ctx* ctx, int32_t* cell, int32_t* perm, int32_t* cell_beg, int32_t* cell_end) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_grid(cell, perm, cell_beg, cell_end));
}
int nbc_get_pair_counts(nbc_ctx* ctx, int64_t* nfront, int64_t* nback) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_pair_counts(nfront, nback));
}
int nbc_get_pairs(nbc_ctx* ctx, int32_t* front_off, int32_t* front_j, uint8_t* front_flags, int32_t* back_off,
                  int32_t* back_i, int32_t* back_slot) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_pairs(front_off, front_j, front_flags, back_off, back_i, back_slot));
}
int nbc_get_pair_history(nbc_ctx* ctx, double* hist9) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_pair_history(hist9));
}
int nbc_get_sphere_forces(nbc_ctx* ctx, int32_t kind, double* force, double* torque) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_sphere_forces(kind, force, torque));
}
int nbc_get_clumps(nbc_ctx* ctx, double* pos, double* vel, double* quat, double* angv, double* force,
                   double* torque) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_clumps(pos, vel, quat, angv, force, torque));
}
int nbc_get_penetrator(nbc_ctx* ctx, double* pos, double* vel, double* quat, double* angv, double* cache_f,
                       double* cache_t, int32_t* contact_state) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_penetrator(pos, vel, quat, angv, cache_f, cache_t, contact_state));
}
int nbc_get_pene_exports(nbc_ctx* ctx, double* fn, double* ft, double* tn, double* tt) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_pene_exports(fn, ft, tn, tt));
}
int nbc_get_pair_exports(nbc_ctx* ctx, double* fcn, double* cpt) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->get_pair_exports(fcn, cpt));
}
int nbc_profile_enable(nbc_ctx* ctx, int32_t on) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->profile_enable(on));
}
int nbc_profile_get(nbc_ctx* ctx, int32_t kernel_class, double* total_ms, int64_t* launches) {
    ENG(ctx);
    return ctx_err(ctx, ctx->eng->profile_get(kernel_class, total_ms, launches));
}

}  // extern "C"
