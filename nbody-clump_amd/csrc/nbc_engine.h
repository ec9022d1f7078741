// Precision-independent interface of the device engine (FP64 or FP32-mixed).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/nbc.h"

namespace nbc {

class EngineBase {
  public:
    virtual ~EngineBase() {}
    virtual int init(const nbc_params& p, const nbc_options& o, int device) = 0;
    virtual int upload(const nbc_spheres_in* s, const nbc_clumps_in* c, const nbc_wall* w,
                       const nbc_penetrator* pe) = 0;
    // batch of independent beds (uniform grids, no penetrator)
    virtual int init_batch(const nbc_params* beds, int nbeds, const nbc_options& o, int device) = 0;
    virtual int upload_batch(const nbc_spheres_in* s, const nbc_clumps_in* c, const nbc_wall* w) = 0;
    virtual int get_batch_ranges(int32_t* sphere_off, int32_t* clump_off) = 0;
    // DEMmain single-sphere loops
    virtual int upload_particles(const nbc_particles_in* p, const nbc_wall* w, const nbc_penetrator* pe) = 0;
    virtual int get_particles(int32_t* sid, int32_t* fid, int32_t* type, double* pos, double* vel, double* quat,
                              double* angv, double* prop, double* radius) = 0;
    virtual int step(int nsteps) = 0;
    virtual int force_rebuild() = 0;
    virtual int set_record(int on) = 0;
    virtual int sync() = 0;
    virtual int get_stats(nbc_stats* s) = 0;
    virtual int get_spheres(int32_t* fid, int32_t* cid, double* pos, double* cum, double* local, double* mass,
                            double* radius) = 0;
    virtual int get_grid(int32_t* cell, int32_t* perm, int32_t* cb, int32_t* ce) = 0;
    virtual int get_pair_counts(int64_t* nf, int64_t* nb) = 0;
    virtual int get_pairs(int32_t* fo, int32_t* fj, uint8_t* ff, int32_t* bo, int32_t* bi, int32_t* bs) = 0;
    virtual int get_pair_history(double* h9) = 0;
    virtual int get_sphere_forces(int kind, double* f, double* t) = 0;
    virtual int get_clumps(double* pos, double* vel, double* quat, double* angv, double* F, double* T) = 0;
    virtual int get_penetrator(double* pos, double* vel, double* quat, double* angv, double* cf, double* ct,
                               int32_t* state) = 0;
    virtual int get_pene_exports(double* fn, double* ft, double* tn, double* tt) = 0;
    virtual int get_pair_exports(double* fcn, double* cpt) = 0;
    virtual int profile_enable(int on) = 0;
    virtual int profile_get(int cls, double* ms, int64_t* n) = 0;

    int set_hip_error(hipError_t e, const char* what, int line);
    int fail(int code, const std::string& m);

    std::string err;
    int device = -1;
};

EngineBase* make_engine(int precision);

}  // namespace nbc
