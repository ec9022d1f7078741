// FP64 host oracle for the clumped-sphere DEM timestep.
//
// TEST INFRASTRUCTURE ONLY. This file restates the reference's step logic
// (qubajin/nbody-clump, DEMClumpMain::RunSimulationClump and the kernels it
// drives) as plain FP64 C++ with OpenMP loops where the reference launches a
// kernel. It keeps the reference's slot-padded dataflow ("faithful" mode:
// Anei/Aref rows of BranchLmtLen entries, per-slot caches and ping-pong
// histories sized N*BranchLmtLen/2, per-step memsets), so it is also the CPU
// baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it, and only as the checker; libnbc never links it.
//
// Build: oracle/Makefile, g++ -O3 -fopenmp -ffp-contract=off (no fast-math):
// every expression below is evaluated in the reference's order so that the
// HIP FP64 path (also built with contraction off) can be compared bit for bit.
//
// Deviations from the reference, all in code paths the reference leaves
// undefined (each is also listed in DESIGN.md):
//  * InheritContHistory leaves unmatched new slots with two-step-old values
//    (src/dem_grid.cu:603-647 writes only matches into a buffer that is not
//    cleared); the oracle clears them first. Under the Verlet-skin invariant
//    such pairs are not touching and the value is never used.
//  * A back slot whose owner did not list the pair keeps a stale Aref entry
//    in the reference (src/dem_grid.cu:591-599, :183-188); the oracle resets
//    Aref to -1 at every rebuild and skips such entries in the gather.
//  * Sums whose association order the reference leaves to Thrust/CUB
//    (src/dem_clumpsphere.cu:139-150, src/dem_force_penetrator.cu:189-248)
//    are sequential in ascending index order.
//  * The clump x cylinder wall (src/dem_force_cylind_wall.cu:152-237 geometry
//    with the clump kinematics of src/dem_force_cube_wall_clump.cu:16-171)
//    has no reference counterpart and is the oracle's own composition.
//
// DEMmain's single-sphere loops (src/dem_simulation_{cube,cylind,split,hgrid}.cu)
// run as one-sphere clumps (orc_upload_particles) with their own laws, chosen
// by nbc_options.driver: the sphere-spin Vrot (src/dem_force_particle.cu:87-88),
// the per-sphere displacement as a sum of drifts (src/dem_integrate.cu:36-38),
// the DEMmain cube wall (src/dem_force_cube_wall.cu: no history, no rolling or
// twisting torque), the glue types (src/dem_integrate_withglue.cu) and the
// cylinder loops' history exchange after a rebuild.
//
// Parity pins: tests/test_oracle_*.py check this file against closed forms
// (Hertz normal force, Coulomb/rolling/twisting caps, free fall, torque-free
// rotation, restitution band), brute-force O(N^2) neighbour sets,
// std::stable_sort order, Newton's third law and the shipped input fixtures.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/nbc.h"
#include "orc_math.h"

using namespace orc;

namespace {

struct Hist3 {  // one (Dv, Dt, Dr) history set: tangential, twisting, rolling
    std::vector<d3> v, t, r;
    void resize(size_t n) {
        v.assign(n, d3{0, 0, 0});
        t.assign(n, d3{0, 0, 0});
        r.assign(n, d3{0, 0, 0});
    }
    void zero() {
        std::fill(v.begin(), v.end(), d3{0, 0, 0});
        std::fill(t.begin(), t.end(), d3{0, 0, 0});
        std::fill(r.begin(), r.end(), d3{0, 0, 0});
    }
    void swap(Hist3& o) {
        v.swap(o.v);
        t.swap(o.t);
        r.swap(o.r);
    }
};

void zero_range(Hist3& h, long long n) {
    std::fill(h.v.begin(), h.v.begin() + n, d3{0, 0, 0});
    std::fill(h.t.begin(), h.t.begin() + n, d3{0, 0, 0});
    std::fill(h.r.begin(), h.r.begin() + n, d3{0, 0, 0});
}

void zero_range(std::vector<d3>& v, long long n) { std::fill(v.begin(), v.begin() + n, d3{0, 0, 0}); }

template <class T>
void gather(std::vector<T>& a, const std::vector<int>& perm) {
    std::vector<T> c(a);
    const int n = (int)a.size();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) a[i] = c[perm[i]];
}

}  // namespace

struct orc_ctx {
    nbc_params P;
    nbc_options O;
    int N = 0, Nc = 0, B = 0;
    long long S = 0;
    int wall_type = NBC_WALL_NONE;
    double pene_frcoe = 0.0;
    std::string err;

    // spheres, current sorted order (src/dem_clump.h:519-529)
    std::vector<int> ClumpId;
    std::vector<d3> Pos, LocalPos, CumPos;
    std::vector<double> Radius, Mass;
    std::vector<d3> totF, totT;
    // clumps (src/dem_clumpsphere.h:345-353)
    std::vector<d3> CPos, CVel, CAngSpd, CFc, CTc;
    std::vector<d4> CQuat, CProp;
    // grid (src/dem_grid.h:19-50)
    std::vector<int> part_cel, part_idx_pre, part_fix, cell_beg, cell_end;
    std::vector<int> Ng, Ni, Nsum, Nsum_pre, Lpa, Lpb, Lpb_pre;
    std::vector<int> Anei, Aref;  // N x B
    // sphere-sphere force (src/dem_force.h:14-40)
    std::vector<d3> Fc, Tc;
    Hist3 D, Dpre;
    std::vector<d3> CacheFc, CacheTc, CacheTc2, CacheFcn, CacheCtp;
    // walls: cube 6 faces (UpX LowX UpY LowY UpZ LowZ) or cylinder 3 surfaces (cap up side)
    nbc_wall W;
    std::vector<d3> WFc, WTc;
    Hist3 WH[6], WHpre[6];
    // penetrator (src/dem_penetrator.h)
    nbc_penetrator Pn;
    std::vector<double> pene_tri;
    double pene_bound = 0.0;
    std::vector<d3> PFc, PTc, PpeneTc, PFcn, PFct, PTcn, PTct, PRVo, PRVn;
    std::vector<int> PLcFlg;
    // DEMmain loops: glue types and the step's drift, per particle (= clump)
    int driver = NBC_DRIVER_CLUMP;
    std::vector<int> Type;
    std::vector<d3> CDist;
    Hist3 PH, PHpre;
    d3 pene_pos{}, pene_vel{}, pene_angv{};
    d4 pene_quat{};
    d3 totFc{}, totTc{};
    std::vector<d3> CacheFi, CacheTi;
    // run state
    unsigned Update = 0;
    long long step = 0, rebuilds = 0;
    int count = 0;
    long long touching_ss = 0, touching_wall = 0;
    bool first = true;
    bool rebuild_next = false;
    bool rebuilt_between = false;  // orc_rebuild ran since the last step
    // compact mode: per-step work and clears cover the used slot range
    // [0, S_used) and the used row entries only (a CSR pair list), instead of
    // all N*BranchLmtLen/2 slots and N*BranchLmtLen row entries
    bool compact = false;
    long long S_used = 0, S_high = 0;
};

static int fail(orc_ctx* c, int code, const std::string& m) {
    c->err = m;
    return code;
}

// ---------------------------------------------------------------------------
// Mesh set-up: src/dem_subsphere_io.cu:143-145 (Damping) and :197-230.
// Hierarchical grid table (no reference counterpart: the reference reads it
// from grid.bt, src/dem_io.cu:204-248): L levels split [rmin, rmax]
// geometrically, level 0 the largest spheres, cell = 3 x the level's largest
// admissible radius, one origin for all levels.
extern "C" int orc_finalize_params(nbc_params* p, const double* radius, int n, double cell_override, int levels) {
    double lnCOR = std::log(p->NormRes);
    p->Damping = 2.0 * std::sqrt(5.0 / 6.0) * lnCOR / std::sqrt(lnCOR * lnCOR + 3.14159265359 * 3.14159265359);
    double MeshSize = 0.0;
    for (int i = 0; i < n; ++i)
        if (radius[i] > MeshSize) MeshSize = radius[i];
    MeshSize = 3.0 * MeshSize;
    if (cell_override > 0.0) MeshSize = cell_override;
    for (int a = 0; a < 3; ++a) {
        p->CellSize[a] = MeshSize;
        p->i_CellSize[a] = 1.0 / MeshSize;
        p->WorldOrigin[a] = p->WorldOrigin[a] - p->CellSize[a];
        p->WorldSize[a] = p->WorldSize[a] + 2 * p->CellSize[a];
        p->GridDim[a] = (uint32_t)std::ceil(p->WorldSize[a] / p->CellSize[a]);
    }
    p->CellNum = (int32_t)(p->GridDim[0] * p->GridDim[1] * p->GridDim[2]);
    p->Levels = 1;
    p->LevelNested = 0;
    p->LevelBlock = 0;
    if (levels > 1) {
        double rmin = radius[0], rmax = radius[0];
        for (int i = 1; i < n; ++i) {
            if (radius[i] < rmin) rmin = radius[i];
            if (radius[i] > rmax) rmax = radius[i];
        }
        const double ratio = std::pow(rmin / rmax, 1.0 / levels);
        int off = 0;
        for (int t = 0; t < levels; ++t) {
            const double top = t == 0 ? rmax : p->LevelRmin[t - 1];
            p->LevelRmin[t] = (t == levels - 1) ? -1.0 : rmax * std::pow(ratio, t + 1);
            p->LevelCell[t] = 3.0 * top;
            p->LevelOff[t] = off;
            int cells = 1;
            for (int a = 0; a < 3; ++a) {
                p->LevelDim[t][a] = (int32_t)std::ceil(p->WorldSize[a] / p->LevelCell[t]);
                cells *= p->LevelDim[t][a];
            }
            off += cells;
        }
        p->Levels = levels;
        p->CellNum = off + 1;
    }
    p->finalized = 1;
    return 0;
}

// Nested hierarchical grid (no counterpart in the reference): level t has
// cells of 3 r_max / 2^t on the uniform grid's origin and margin, a sphere
// sits on the finest level whose cell is at least 3 r, and every coarse cell
// owns a block of sum_t 8^t keys (level t at offset sum_{u<t} 8^u, then its
// 8^t sub-cells z/y/x-major), so the sort keeps a coarse cell's spheres
// together.
extern "C" int orc_finalize_params_nested(nbc_params* p, const double* radius, int n, int levels) {
    if (levels < 2 || levels > 4) return NBC_E_BADINPUT;
    const int e = orc_finalize_params(p, radius, n, 0.0, 0);
    if (e) return e;
    const double s0 = p->CellSize[0];
    int block = 0;
    for (int t = 0; t < levels; ++t) {
        p->LevelCell[t] = s0 / (double)(1 << t);
        p->LevelOff[t] = block;
        block += 1 << (3 * t);
        for (int a = 0; a < 3; ++a) p->LevelDim[t][a] = (int32_t)(p->GridDim[a] << t);
    }
    for (int t = 0; t < levels; ++t) p->LevelRmin[t] = t + 1 < levels ? p->LevelCell[t + 1] / 3.0 : -1.0;
    p->Levels = levels;
    p->LevelNested = 1;
    p->LevelBlock = block;
    p->CellNum = (int32_t)((long long)p->GridDim[0] * p->GridDim[1] * p->GridDim[2] * block) + 1;
    return 0;
}

extern "C" orc_ctx* orc_create(const nbc_params* p, const nbc_options* o) {
    orc_ctx* c = new orc_ctx();
    c->P = *p;
    if (o) c->O = *o; else std::memset(&c->O, 0, sizeof(c->O));
    c->N = p->ParticleNum;
    c->Nc = p->ClumpNum;
    c->B = p->BranchLmtLen;
    c->S = (long long)c->N * (c->B / 2);
    c->wall_type = p->isWallForce ? (c->O.wall_type ? c->O.wall_type : NBC_WALL_CUBE) : NBC_WALL_NONE;
    c->pene_frcoe = (c->O.pene_frcoe > 0.0) ? c->O.pene_frcoe : p->Frcoe;
    c->driver = c->O.driver;
    // the DEMmain loops apply their wall unconditionally
    if (c->driver == NBC_DRIVER_CUBE) c->wall_type = NBC_WALL_CUBE;
    else if (c->driver != NBC_DRIVER_CLUMP) c->wall_type = NBC_WALL_CYLINDER;
    return c;
}

extern "C" void orc_destroy(orc_ctx* c) { delete c; }
extern "C" const char* orc_last_error(orc_ctx* c) { return c->err.c_str(); }
extern "C" void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
extern "C" int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

extern "C" int orc_upload(orc_ctx* c, const nbc_spheres_in* s, const nbc_clumps_in* cl,
                          const nbc_wall* w, const nbc_penetrator* pe) {
    const int N = c->N, Nc = c->Nc, B = c->B;
    if (!s || !cl || s->n != N || cl->n != Nc || N <= 0 || Nc <= 0 || B < 2)
        return fail(c, NBC_E_BADINPUT, "orc_upload: size mismatch");
    // An out-of-range clump id would be read out of bounds by UpdatePos
    // (src/dem_clump.cu:120) before the wall kernel's check
    // (src/dem_force_cube_wall_clump.cu:191-194) sees it; refuse it here.
    for (int i = 0; i < N; ++i)
        if (s->clump_id[i] < 0 || s->clump_id[i] >= Nc) {
            char m[160];
            std::snprintf(m, sizeof m, "FATAL: Invalid input (clump id) at particle index %d in file order. "
                                       "Simulation stopped.", i);
            return fail(c, NBC_E_NONFINITE, m);
        }
    c->ClumpId.assign(s->clump_id, s->clump_id + N);
    c->Pos.resize(N);
    c->LocalPos.resize(N);
    c->Radius.assign(s->radius, s->radius + N);
    c->Mass.assign(s->mass, s->mass + N);
    for (int i = 0; i < N; ++i) {
        c->Pos[i] = mk3(s->pos[3 * i], s->pos[3 * i + 1], s->pos[3 * i + 2]);
        c->LocalPos[i] = mk3(s->local[3 * i], s->local[3 * i + 1], s->local[3 * i + 2]);
    }
    c->CumPos.assign(N, d3{0, 0, 0});
    c->totF.assign(N, d3{0, 0, 0});
    c->totT.assign(N, d3{0, 0, 0});
    c->CPos.resize(Nc); c->CVel.resize(Nc); c->CAngSpd.resize(Nc);
    c->CQuat.resize(Nc); c->CProp.resize(Nc);
    c->CFc.assign(Nc, d3{0, 0, 0});
    c->CTc.assign(Nc, d3{0, 0, 0});
    for (int i = 0; i < Nc; ++i) {
        c->CPos[i] = mk3(cl->pos[3 * i], cl->pos[3 * i + 1], cl->pos[3 * i + 2]);
        c->CVel[i] = mk3(cl->vel[3 * i], cl->vel[3 * i + 1], cl->vel[3 * i + 2]);
        c->CAngSpd[i] = mk3(cl->angv[3 * i], cl->angv[3 * i + 1], cl->angv[3 * i + 2]);
        c->CQuat[i] = d4{cl->quat[4 * i], cl->quat[4 * i + 1], cl->quat[4 * i + 2], cl->quat[4 * i + 3]};
        c->CProp[i] = d4{cl->prop[4 * i], cl->prop[4 * i + 1], cl->prop[4 * i + 2], cl->prop[4 * i + 3]};
    }
    // grid (src/dem_grid.cu:212-271): part_fix = identity, Nsum/Nsum_pre = 0, Lpb_pre = -1
    c->part_cel.assign(N, 0);
    c->part_idx_pre.resize(N);
    c->part_fix.resize(N);
    for (int i = 0; i < N; ++i) { c->part_idx_pre[i] = i; c->part_fix[i] = i; }
    c->cell_beg.assign(c->P.CellNum, -1);
    c->cell_end.assign(c->P.CellNum, -1);
    c->Ng.assign(N, 0); c->Ni.assign(N, 0);
    c->Nsum.assign(N, 0); c->Nsum_pre.assign(N, 0);
    c->Lpa.assign(c->S, -1); c->Lpb.assign(c->S, -1); c->Lpb_pre.assign(c->S, -1);
    c->Anei.assign((size_t)N * B, -1);
    c->Aref.assign((size_t)N * B, -1);
    // force (src/dem_force.cu:32-66)
    c->Fc.assign(N, d3{0, 0, 0}); c->Tc.assign(N, d3{0, 0, 0});
    c->D.resize(c->S); c->Dpre.resize(c->S);
    c->CacheFc.assign(c->S, d3{0, 0, 0}); c->CacheTc.assign(c->S, d3{0, 0, 0});
    c->CacheTc2.assign(c->S, d3{0, 0, 0}); c->CacheFcn.assign(c->S, d3{0, 0, 0});
    c->CacheCtp.assign(c->S, d3{0, 0, 0});
    // walls
    if (w) c->W = *w; else std::memset(&c->W, 0, sizeof(c->W));
    c->WFc.assign(N, d3{0, 0, 0}); c->WTc.assign(N, d3{0, 0, 0});
    for (int f = 0; f < 6; ++f) { c->WH[f].resize(N); c->WHpre[f].resize(N); }
    // penetrator
    if (pe) {
        c->Pn = *pe;
        if (pe->shape == NBC_PENE_CONE && pe->tri && pe->ntri > 0) {
            c->pene_tri.assign(pe->tri, pe->tri + 9 * (size_t)pe->ntri);
            double b2 = 0.0;
            for (size_t k = 0; k < c->pene_tri.size(); k += 3)
                b2 = std::max(b2, c->pene_tri[k] * c->pene_tri[k] + c->pene_tri[k + 1] * c->pene_tri[k + 1] +
                                      c->pene_tri[k + 2] * c->pene_tri[k + 2]);
            c->pene_bound = std::sqrt(b2) * (1.0 + 1e-12);
        }
        c->Pn.tri = nullptr;
        c->pene_pos = mk3(pe->pos[0], pe->pos[1], pe->pos[2]);
        c->pene_vel = mk3(pe->vel[0], pe->vel[1], pe->vel[2]);
        c->pene_angv = mk3(pe->angv[0], pe->angv[1], pe->angv[2]);
        c->pene_quat = d4{pe->quat[0], pe->quat[1], pe->quat[2], pe->quat[3]};
    } else {
        std::memset(&c->Pn, 0, sizeof(c->Pn));
    }
    for (auto* v : {&c->PFc, &c->PTc, &c->PpeneTc, &c->PFcn, &c->PFct, &c->PTcn, &c->PTct, &c->PRVo, &c->PRVn})
        v->assign(N, d3{0, 0, 0});
    c->PLcFlg.assign(N, 0);
    c->PH.resize(N); c->PHpre.resize(N);
    c->totFc = d3{0, 0, 0}; c->totTc = d3{0, 0, 0};
    const int oi = c->P.OutputInterval > 0 ? c->P.OutputInterval : 1;
    c->CacheFi.assign(oi, d3{0, 0, 0}); c->CacheTi.assign(oi, d3{0, 0, 0});
    c->Update = (unsigned)c->P.Update;
    c->step = 0; c->rebuilds = 0; c->count = 0; c->first = true;
    c->S_used = 0; c->S_high = 0;
    c->CDist.assign(Nc, d3{0, 0, 0});
    if (c->Type.size() != (size_t)Nc) c->Type.assign(Nc, 0);
    return 0;
}

// DEMmain::CopyDataFromHostToDevice (src/dem.cu:62-72): particle i is clump i
// with one sphere at its centre.
extern "C" int orc_upload_particles(orc_ctx* c, const nbc_particles_in* q, const nbc_wall* w,
                                    const nbc_penetrator* pe) {
    if (c->driver == NBC_DRIVER_CLUMP) return fail(c, NBC_E_BADINPUT, "orc_upload_particles: clump driver");
    if (!q || q->n != c->N || c->Nc != c->N) return fail(c, NBC_E_BADINPUT, "orc_upload_particles: size mismatch");
    const int N = c->N;
    std::vector<int> cid(N);
    std::vector<double> local(3 * (size_t)N, 0.0), mass(N);
    for (int i = 0; i < N; ++i) { cid[i] = i; mass[i] = q->prop[4 * (size_t)i + 3]; }
    c->Type.assign(N, 0);
    if (q->type) c->Type.assign(q->type, q->type + N);
    nbc_spheres_in si{N, 0, cid.data(), q->pos, local.data(), mass.data(), q->radius};
    nbc_clumps_in ci{N, 0, q->pos, q->vel, q->quat, q->angv, q->prop};
    return orc_upload(c, &si, &ci, w, pe);
}

// ---------------------------------------------------------------------------
// Clump integration: src/dem_integrate_clumpsphere.cu:3-43 and :45-90.
static void integrate_before(orc_ctx* c) {
    const double Dt = c->P.StepSize, g = c->P.Gravity;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->Nc; ++i) {
        d3 pos = c->CPos[i], vel = c->CVel[i], w = c->CAngSpd[i];
        const int ty = c->driver == NBC_DRIVER_SPLIT ? c->Type[i] : 0;
        if (ty == 2) {  // src/dem_integrate_withglue.cu:7
            c->CDist[i] = d3{0, 0, 0};
            continue;
        }
        if (ty == 1) {  // src/dem_integrate_withglue.cu:46-57
            double Omega = 0.2;
            vel.x = -Omega * pos.y;
            vel.y = -Omega * pos.x;
            vel.z = 0.0;
            d3 dist = Dt * vel;
            pos = pos + dist;
            c->CPos[i] = pos; c->CVel[i] = vel;
            c->CDist[i] = dist;
            continue;
        }
        d3 F = c->CFc[i], T = c->CTc[i];
        const d4 pr = c->CProp[i];
        F = (1.0 / pr.w) * F;
        F.z = F.z + g;
        T.x = (1.0 / pr.x) * T.x;
        T.y = (1.0 / pr.y) * T.y;
        T.z = (1.0 / pr.z) * T.z;
        d3 dist = Dt * vel + 0.5 * Dt * Dt * F;
        pos = pos + dist;
        vel = vel + 0.5 * Dt * F;
        w = w + 0.5 * Dt * T;
        c->CPos[i] = pos; c->CVel[i] = vel; c->CAngSpd[i] = w;
        c->CDist[i] = dist;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->Nc; ++i) {
        if (c->driver == NBC_DRIVER_SPLIT && c->Type[i] != 0) continue;  // src/dem_integrate_withglue.cu:68
        d3 W = c->CAngSpd[i], T = c->CTc[i];
        const d4 pr = c->CProp[i];
        T.x = (1.0 / pr.x) * T.x;
        T.y = (1.0 / pr.y) * T.y;
        T.z = (1.0 / pr.z) * T.z;
        d4 q = c->CQuat[i];
        double DCM[3][3];
        Quat2DCM(q, DCM);
        d3 Wb = Map(DCM, W), Wd = Map(DCM, T);
        d4 qd, qdd;
        qd.x = 0.50 * (Wb.x * q.w - Wb.y * q.z + Wb.z * q.y);
        qd.y = 0.50 * (Wb.x * q.z + Wb.y * q.w - Wb.z * q.x);
        qd.z = 0.50 * (-Wb.x * q.y + Wb.y * q.x + Wb.z * q.w);
        qd.w = 0.50 * (-Wb.x * q.x - Wb.y * q.y - Wb.z * q.z);
        qdd.x = (Wd.x * q.w - Wd.y * q.z + Wd.z * q.y) * 0.50 + (Wb.x * qd.w - Wb.y * qd.z + Wb.z * qd.y) * 0.50;
        qdd.y = (Wd.x * q.z + Wd.y * q.w - Wd.z * q.x) * 0.50 + (Wb.x * qd.z + Wb.y * qd.w - Wb.z * qd.x) * 0.50;
        qdd.z = (-Wd.x * q.y + Wd.y * q.x + Wd.z * q.w) * 0.50 + (-Wb.x * qd.y + Wb.y * qd.x + Wb.z * qd.w) * 0.50;
        qdd.w = (-Wd.x * q.x - Wd.y * q.y - Wd.z * q.z) * 0.50 + (-Wb.x * qd.x - Wb.y * qd.y - Wb.z * qd.z) * 0.50;
        q = q + c->P.StepSize * qd + 0.5 * c->P.StepSize * c->P.StepSize * qdd;
        double norm = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
        q = (1.0 / norm) * q;
        c->CQuat[i] = q;
    }
}

// src/dem_clump.cu:113-138
static void update_pos(orc_ctx* c) {
    unsigned flag = 0;
    const double th = c->P.Thresh;
#pragma omp parallel for schedule(static) reduction(| : flag)
    for (int i = 0; i < c->N; ++i) {
        const int cl = c->ClumpId[i];
        double DCM[3][3];
        Quat2IvDCM(c->CQuat[cl], DCM);
        d3 arm = Map(DCM, c->LocalPos[i]);
        d3 up = c->CPos[cl] + arm;
        d3 cum;
        if (c->driver) {
            // DEMmain: cumpos = cumpos + dist (src/dem_integrate.cu:37); the glue
            // loop rebuilds past 0.1 r for type 0 only (src/dem_integrate_withglue.cu:40)
            cum = c->CumPos[i] + c->CDist[cl];
            if (c->driver == NBC_DRIVER_SPLIT) {
                if (c->Type[cl] == 0 && Length(cum) > 0.10 * c->Radius[i]) flag = 1u;
            } else if (Length(cum) > th) {
                flag = 1u;
            }
        } else {
            cum = c->CumPos[i] + up - c->Pos[i];
            if (Length(cum) > th) flag = 1u;
        }
        c->CumPos[i] = cum;
        c->Pos[i] = up;
    }
    if (flag && c->Update == 0) c->Update = 1;
}

// src/dem_integrate_penetrator.cu:3-41
static void pene_integrate_before(orc_ctx* c) {
    const double Dt = c->P.StepSize;
    nbc_penetrator& p = c->Pn;
    if (p.motion == NBC_PENE_GRAVITY && (!p.wait_for_contact || p.contact_state != 0)) {
        double mass = p.prop[3];
        d3 peneF = (-1.0 / mass) * c->totFc;
        peneF.z = peneF.z + c->P.Gravity;
        c->pene_pos = c->pene_pos + Dt * c->pene_vel + 0.5 * Dt * Dt * peneF;
        c->pene_vel = c->pene_vel + 0.5 * Dt * peneF;
    } else {
        c->pene_pos = c->pene_pos + Dt * c->pene_vel;
    }
}

// ---------------------------------------------------------------------------
// Grid: src/dem_grid.cu:155-202 (hash), :278-305 (stable sort), :105-153 (cell bounds).
static int cell_hash(const nbc_params& P, const d3& p) {
    if (p.z >= P.BoxUpLmt[2] || p.y >= P.BoxUpLmt[1] || p.x >= P.BoxUpLmt[0] || p.z < P.BoxLowLmt[2] ||
        p.y < P.BoxLowLmt[1] || p.x < P.BoxLowLmt[0])
        return P.CellNum - 1;
    int cx = (int)std::floor((p.x - P.WorldOrigin[0]) / P.CellSize[0]);
    int cy = (int)std::floor((p.y - P.WorldOrigin[1]) / P.CellSize[1]);
    int cz = (int)std::floor((p.z - P.WorldOrigin[2]) / P.CellSize[2]);
    // max(0, min(GridDim - 1, c)) with GridDim unsigned: the clamp happens in
    // unsigned arithmetic, so a negative index maps to GridDim - 1.
    auto clampu = [](uint32_t G, int v) -> int {
        uint32_t u = (uint32_t)v, hi = G - 1u;
        return (int)(u < hi ? u : hi);
    };
    cx = clampu(P.GridDim[0], cx);
    cy = clampu(P.GridDim[1], cy);
    cz = clampu(P.GridDim[2], cz);
    int idx = (int)((uint32_t)cz * P.GridDim[1] * P.GridDim[0] + (uint32_t)cy * P.GridDim[0] + (uint32_t)cx);
    if (idx < 0) idx = 0;
    if (idx >= P.CellNum) idx = P.CellNum - 1;
    return idx;
}

// Hierarchical grid (src/dem_hgrid.cu:205-242): level 0 holds the largest
// spheres; a sphere's level is the first t with r > LevelRmin[t]. Corrected
// semantics: out-of-box spheres go to the last cell (CellNum - 1, never
// searched) instead of CellNum, and the per-level index is clamped into the
// level's box so no index leaves its level.
static int sphere_level(const nbc_params& P, double r) {
    int t = 0;
    while (t < P.Levels - 1 && !(r > P.LevelRmin[t])) ++t;
    return t;
}

// Level of sorted sphere i: by radius, or, in the DEMmain hgrid loop with
// grid.bt levels, the particle's type (src/dem_hgrid.cu:213).
static int level_of(const orc_ctx* c, int i) {
    return c->O.level_from_type ? c->Type[c->ClumpId[i]] : sphere_level(c->P, c->Radius[i]);
}

// Key of level-t cell (x, y, z): level-major (src/dem_hgrid.cu:228) or nested.
static int level_key(const nbc_params& P, int t, int x, int y, int z) {
    if (!P.LevelNested) return P.LevelOff[t] + (z * P.LevelDim[t][1] + y) * P.LevelDim[t][0] + x;
    const int n = 1 << t;
    const int coarse = ((z / n) * P.LevelDim[0][1] + (y / n)) * P.LevelDim[0][0] + (x / n);
    return coarse * P.LevelBlock + P.LevelOff[t] + ((z % n) * n + (y % n)) * n + (x % n);
}

static int cell_hash_h(const nbc_params& P, const d3& p, int t) {
    if (p.z > P.BoxUpLmt[2] || p.y > P.BoxUpLmt[1] || p.x > P.BoxUpLmt[0] || p.z < P.BoxLowLmt[2] ||
        p.y < P.BoxLowLmt[1] || p.x < P.BoxLowLmt[0])
        return P.CellNum - 1;
    const double sz = P.LevelCell[t];
    const double x[3] = {p.x, p.y, p.z};
    int ci[3];
    for (int a = 0; a < 3; ++a) {
        int v = (int)std::floor((x[a] - P.WorldOrigin[a]) / sz);
        ci[a] = v < 0 ? 0 : (v > P.LevelDim[t][a] - 1 ? P.LevelDim[t][a] - 1 : v);
    }
    return level_key(P, t, ci[0], ci[1], ci[2]);
}

static void neighbor_search(orc_ctx* c) {
    const int N = c->N;
    std::vector<int> cel(N);
    const bool hg = c->P.Levels > 1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) cel[i] = hg ? cell_hash_h(c->P, c->Pos[i], level_of(c, i)) : cell_hash(c->P, c->Pos[i]);
    // Stable sort of (cell, index): CUB SortPairs is a stable LSD radix sort
    // over the identity values written by GetParticleCellIndex_cuk.
    std::vector<int> order(N);
    for (int i = 0; i < N; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cel[a] < cel[b]; });
    std::vector<int> fix_old = c->part_fix;
    for (int i = 0; i < N; ++i) {
        c->part_cel[i] = cel[order[i]];
        c->part_idx_pre[i] = order[i];
        c->part_fix[i] = fix_old[order[i]];
    }
    std::fill(c->cell_beg.begin(), c->cell_beg.end(), -1);
    std::fill(c->cell_end.begin(), c->cell_end.end(), -1);
    const int CN = c->P.CellNum;
    for (int i = 0; i < N; ++i) {
        int ci = c->part_cel[i];
        if (ci < 0 || ci >= CN) continue;
        int prev = i > 0 ? c->part_cel[i - 1] : -1;
        if (i == 0 || ci != prev) {
            c->cell_beg[ci] = i;
            if (i > 0 && prev >= 0 && prev < CN) c->cell_end[prev] = i;
        }
        if (i == N - 1) c->cell_end[ci] = i + 1;
    }
}

// src/dem_grid.cu:466-552 (27-cell search, z-outer / x-inner, split by index)
static int create_anei(orc_ctx* c) {
    const int N = c->N, B = c->B, half = B / 2;
    const nbc_params& P = c->P;
    int overflow = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
    for (int index = 0; index < N; ++index) {
        if (c->part_cel[index] == P.CellNum - 1) continue;
        const d3 pi = c->Pos[index];
        const double ri = c->Radius[index];
        int cix = (int)std::floor((pi.x - P.WorldOrigin[0]) * P.i_CellSize[0]);
        int ciy = (int)std::floor((pi.y - P.WorldOrigin[1]) * P.i_CellSize[1]);
        int ciz = (int)std::floor((pi.z - P.WorldOrigin[2]) * P.i_CellSize[2]);
        int* row = &c->Anei[(size_t)index * B];
        int ng = 0, ni = 0;
        for (int cz = -1; cz <= 1; cz++)
            for (int cy = -1; cy <= 1; cy++)
                for (int cx = -1; cx <= 1; cx++) {
                    int jx = cix + cx, jy = ciy + cy, jz = ciz + cz;
                    if (jx < 0 || (uint32_t)jx >= P.GridDim[0] || jy < 0 || (uint32_t)jy >= P.GridDim[1] || jz < 0 ||
                        (uint32_t)jz >= P.GridDim[2])
                        continue;
                    int cidx = (int)(((uint32_t)jz * P.GridDim[1] + (uint32_t)jy) * P.GridDim[0] + (uint32_t)jx);
                    int start = c->cell_beg[cidx];
                    if (start == -1) continue;
                    int end = c->cell_end[cidx];
                    for (int j = start; j < end; j++) {
                        if (j == index) continue;
                        const d3 pj = c->Pos[j];
                        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                        double ds = dx * dx + dy * dy + dz * dz;
                        double lim = 1.50 * (ri + c->Radius[j]);
                        if (ds < lim * lim) {
                            if (index < j) {
                                if (ng < half) row[ng++] = j;
                                else overflow = 1;
                            } else {
                                if (ni < half) { row[B - 1 - ni] = j; ni++; }
                                else overflow = 1;
                            }
                        }
                    }
                }
        c->Ng[index] = ng;
        c->Ni[index] = ni;
    }
    return overflow;
}

// Hierarchical search, src/dem_hgrid_contlist.cu:57-161: a sphere of level t
// searches levels t..0 (27 cells each, z outer / x inner, cut-off
// 1.5 (ri + rj)). Same level: split by index as in the uniform grid. Coarser
// level T < t: the coarser partner j owns the pair (with level-major cells it
// has the lower sorted index; with nested cells it owns it either way), so i
// lists j in its back row and j lists i in its front row.
// Corrected semantics (the reference appends with atomicAdd into a row its
// owner is writing without one): j's own same-level fronts come first, then its
// finer partners in ascending index; both halves are checked for overflow.
static int create_anei_h(orc_ctx* c) {
    const int N = c->N, B = c->B, half = B / 2;
    const nbc_params& P = c->P;
    int overflow = 0;
    std::vector<std::vector<std::pair<int, int>>> cross(orc_get_threads());
#pragma omp parallel reduction(| : overflow)
    {
#ifdef _OPENMP
        std::vector<std::pair<int, int>>& mine = cross[omp_get_thread_num()];
#else
        std::vector<std::pair<int, int>>& mine = cross[0];
#endif
#pragma omp for schedule(dynamic, 256)
        for (int index = 0; index < N; ++index) {
            if (c->part_cel[index] == P.CellNum - 1) continue;
            const d3 pi = c->Pos[index];
            const double ri = c->Radius[index];
            const int my = level_of(c, index);
            int* row = &c->Anei[(size_t)index * B];
            int ng = 0, ni = 0;
            for (int T = my; T >= 0; --T) {
                const double sz = P.LevelCell[T];
                const int* nd = P.LevelDim[T];
                const int cix = (int)std::floor((pi.x - P.WorldOrigin[0]) / sz);
                const int ciy = (int)std::floor((pi.y - P.WorldOrigin[1]) / sz);
                const int ciz = (int)std::floor((pi.z - P.WorldOrigin[2]) / sz);
                for (int cz = -1; cz <= 1; cz++)
                    for (int cy = -1; cy <= 1; cy++)
                        for (int cx = -1; cx <= 1; cx++) {
                            const int jx = cix + cx, jy = ciy + cy, jz = ciz + cz;
                            if (jx < 0 || jx >= nd[0] || jy < 0 || jy >= nd[1] || jz < 0 || jz >= nd[2]) continue;
                            const int cidx = level_key(P, T, jx, jy, jz);
                            const int start = c->cell_beg[cidx];
                            if (start == -1) continue;
                            const int end = c->cell_end[cidx];
                            for (int j = start; j < end; j++) {
                                if (j == index) continue;
                                const d3 pj = c->Pos[j];
                                double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                                double ds = dx * dx + dy * dy + dz * dz;
                                double lim = 1.5 * (ri + c->Radius[j]);
                                if (!(ds < lim * lim)) continue;
                                if (T == my && index < j) {
                                    if (ng < half) row[ng++] = j;
                                    else overflow = 1;
                                } else {
                                    if (ni < half) { row[B - 1 - ni] = j; ni++; }
                                    else overflow = 1;
                                    if (T != my) mine.emplace_back(j, index);
                                }
                            }
                        }
            }
            c->Ng[index] = ng;
            c->Ni[index] = ni;
        }
    }
    std::vector<std::pair<int, int>> all;
    for (auto& v : cross) all.insert(all.end(), v.begin(), v.end());
    std::sort(all.begin(), all.end());
    for (const auto& e : all) {
        const int j = e.first;
        if (c->Ng[j] < half) c->Anei[(size_t)j * B + c->Ng[j]++] = e.second;
        else overflow = 1;
    }
    return overflow;
}

// src/dem_grid.cu:680-750 plus the sort/gather of src/dem_simulation_clump.cu:126-175.
static int rebuild(orc_ctx* c) {
    const int N = c->N, B = c->B;
    neighbor_search(c);
    const std::vector<int>& perm = c->part_idx_pre;
    gather(c->Pos, perm);
    gather(c->LocalPos, perm);
    gather(c->Radius, perm);
    gather(c->Mass, perm);
    gather(c->ClumpId, perm);
    const int nf = c->wall_type == NBC_WALL_CUBE ? 6 : (c->wall_type == NBC_WALL_CYLINDER ? 3 : 0);
    for (int f = 0; f < nf; ++f) {
        gather(c->WH[f].v, perm); gather(c->WH[f].t, perm); gather(c->WH[f].r, perm);
    }
    if (c->P.isIntruder) {
        gather(c->PH.v, perm); gather(c->PH.t, perm); gather(c->PH.r, perm);
    }
    // CreateContList: member swaps Lpb<->Lpb_pre and Nsum<->Nsum_pre; the
    // caller's Dv/Dv_pre pair is swapped locally so that inheritance reads the
    // last written history (caller's Dv) and writes the caller's Dv_pre.
    c->Lpb.swap(c->Lpb_pre);
    c->Nsum.swap(c->Nsum_pre);
    if (c->compact) {
        // only the slots and row entries any list has used can hold values
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            int* rowA = &c->Anei[(size_t)i * B];
            int* rowR = &c->Aref[(size_t)i * B];
            for (int g = 0; g < c->Ng[i]; ++g) rowA[g] = rowR[g] = -1;
            for (int g = B - c->Ni[i]; g < B; ++g) rowA[g] = rowR[g] = -1;
        }
        std::fill(c->Lpa.begin(), c->Lpa.begin() + c->S_high, -1);
        std::fill(c->Lpb.begin(), c->Lpb.begin() + c->S_high, -1);
    } else {
        std::fill(c->Lpa.begin(), c->Lpa.end(), -1);
        std::fill(c->Lpb.begin(), c->Lpb.end(), -1);
        std::fill(c->Anei.begin(), c->Anei.end(), -1);
        std::fill(c->Aref.begin(), c->Aref.end(), -1);
    }
    std::fill(c->Ng.begin(), c->Ng.end(), 0);
    std::fill(c->Ni.begin(), c->Ni.end(), 0);
    if (c->P.Levels > 1 ? create_anei_h(c) : create_anei(c)) {
        char m[256];
        std::snprintf(m, sizeof m,
                      "FATAL: Neighbor list overflow (BranchLmtLen=%d). Simulation stopped before using the "
                      "incomplete list. Review neighbor capacity and GPU memory before rerunning.",
                      B);
        return fail(c, NBC_E_NEIGHBOR_OVERFLOW, m);
    }
    // inclusive scan
    long long acc = 0;
    for (int i = 0; i < N; ++i) { acc += c->Ng[i]; c->Nsum[i] = (int)acc; }
    // CreateContList_cuk (src/dem_grid.cu:561-601)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        int* rowA = &c->Anei[(size_t)i * B];
        int* rowR = &c->Aref[(size_t)i * B];
        for (int g = 0; g < c->Ng[i]; g++) {
            int Nlist = (i == 0) ? g : c->Nsum[i - 1] + g;
            rowR[g] = Nlist;
            int j = rowA[g];
            c->Lpa[Nlist] = i;
            c->Lpb[Nlist] = j;
            const int* rowA2 = &c->Anei[(size_t)j * B];
            for (int k = 0; k < c->Ni[j]; k++)
                if (i == rowA2[B - 1 - k]) c->Aref[(size_t)j * B + (B - 1 - k)] = Nlist;
        }
    }
    // InheritContHistory_cuk (src/dem_grid.cu:603-647): read the caller's Dv
    // (last written), write the caller's Dv_pre.
    Hist3& src = c->D;
    Hist3& dst = c->Dpre;
    const long long np = acc;
    c->S_used = np;
    c->S_high = std::max(c->S_high, np);
    if (c->compact) zero_range(dst, c->S_high);
    else dst.zero();
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < np; ++k) {
        const int i = c->Lpa[k], j = c->Lpb[k];
        const int ip = perm[i], jp = perm[j];
        // nested levels: a pair on two levels stays in the coarser sphere's row
        const bool cross = c->P.Levels > 1 && c->P.LevelNested && level_of(c, i) != level_of(c, j);
        if (cross || jp > ip) {
            const int b = ip == 0 ? 0 : c->Nsum_pre[ip - 1];
            for (int q = b; q < c->Nsum_pre[ip]; q++)
                if (jp == c->Lpb_pre[q]) { dst.v[k] = src.v[q]; dst.t[k] = src.t[q]; dst.r[k] = src.r[q]; }
        } else if (jp < ip) {
            const int b = jp == 0 ? 0 : c->Nsum_pre[jp - 1];
            for (int q = b; q < c->Nsum_pre[jp]; q++)
                if (ip == c->Lpb_pre[q]) {
                    dst.v[k] = -1.0 * src.v[q]; dst.t[k] = -1.0 * src.t[q]; dst.r[k] = -1.0 * src.r[q];
                }
        }
    }
    // the cylinder loops exchange right after this and read D (the last
    // written histories, old slots) as their Dpre; the others start D clean
    if (c->driver != NBC_DRIVER_CYLINDER && c->driver != NBC_DRIVER_SPLIT) {
        if (c->compact) zero_range(c->D, c->S_high);
        else c->D.zero();
    }
    c->Update = 0;
    std::fill(c->CumPos.begin(), c->CumPos.end(), d3{0, 0, 0});
    c->rebuilds++;
    return 0;
}

// ---------------------------------------------------------------------------
// Sphere-sphere contact, per slot: src/dem_force_clump.cu:3-163.
static bool ss_force(orc_ctx* c, long long k) {
    const nbc_params& P = c->P;
    const int I = c->Lpa[k], J = c->Lpb[k];
    if (I < 0 || J < 0) return false;
    const int cI = c->ClumpId[I], cJ = c->ClumpId[J];
    if (cI == cJ) return false;
    d3 Dist = c->Pos[J] - c->Pos[I];
    d3 H = c->Dpre.v[k], Mr = c->Dpre.r[k], Mt = c->Dpre.t[k];
    double DistS = Dist.x * Dist.x + Dist.y * Dist.y + Dist.z * Dist.z;
    double DistL = std::sqrt(DistS);
    const double rI = c->Radius[I], rJ = c->Radius[J];
    double Dn = rI + rJ - DistL;
    if (!(Dn > 0.0)) return false;
    double DistR = 1.0 / DistL;
    d3 DistU = DistR * Dist;
    double Ap = 0.50 * (rI * rI - rJ * rJ + DistS) * DistR;
    double An = DistL - Ap;
    double Rij = rI * rJ / (rI + rJ);
    double Mij = c->Mass[I] * c->Mass[J] / (c->Mass[I] + c->Mass[J]);
    double Kn = 2.0 * P.Ymod * std::sqrt(Rij * Dn) / (3.0 * (1.0 - P.Prat * P.Prat));
    double Cn = P.Damping * std::sqrt(Kn * Mij);
    double Ks = 2.0 * P.Ymod / (1.0 + P.Prat) / (2.0 - P.Prat) * std::sqrt(Rij * Dn);
    double Cs = P.Damping * std::sqrt(Ks * Mij);
    double Kr = 0.25 * Kn * (P.Beta * Rij) * (P.Beta * Rij);
    double Cr = 0.25 * Cn * (P.Beta * Rij) * (P.Beta * Rij);
    double Kt = 0.50 * Ks * (P.Beta * Rij) * (P.Beta * Rij);
    double Ct = 0.50 * Cs * (P.Beta * Rij) * (P.Beta * Rij);
    double Mi[3][3], Mj[3][3];
    Quat2IvDCM(c->CQuat[cI], Mi);
    Quat2IvDCM(c->CQuat[cJ], Mj);
    d3 ARMi = Map(Mi, c->LocalPos[I]);
    d3 ARMj = Map(Mj, c->LocalPos[J]);
    ARMi = ARMi + Ap * DistU;
    ARMj = ARMj - An * DistU;
    d3 Vrel = c->CVel[cJ] - c->CVel[cI];
    d3 Vrot = c->driver ? An * Cross(DistU, c->CAngSpd[cJ]) + Ap * Cross(DistU, c->CAngSpd[cI])  // particle.cu:88
                        : Cross(c->CAngSpd[cJ], ARMj) - Cross(c->CAngSpd[cI], ARMi);
    d3 Vtot = Vrel + Vrot;
    double ERR = Dot(DistU, Vtot);
    d3 Vnor = ERR * DistU;
    d3 Vtan = Vtot - Vnor;
    d3 nf = Kn * Dn * DistU + Cn * Vnor;
    double nfL = Length(nf);
    d3 Ds = P.StepSize * Vtan;
    d3 tf = -Ks * Ds + Cs * Vtan + H;
    double tfL = Length(tf);
    d3 th;
    if (tfL > nfL * P.Frcoe) {
        tf = P.Frcoe * nfL * (1.0 / tfL) * tf;
        th = tf;
    } else {
        th = tf - Cs * Vtan;
    }
    d3 cf = P.Cohesion * (P.Beta * Rij) * (P.Beta * Rij) * DistU;
    d3 cfc = -1.0 * nf - tf + cf;
    d3 ctc = -Ap * Cross(DistU, tf);
    d3 ctc2 = -An * Cross(DistU, tf);
    d3 Dth = P.StepSize * (c->CAngSpd[cI] - c->CAngSpd[cJ]);
    double DthL = Dot(Dth, DistU);
    d3 DthT = DthL * DistU;
    d3 DthR = Dth - DthT;
    d3 rm = Kr * DthR + Mr;
    double rmL = Length(rm);
    d3 rh;
    if (rmL > 0.525 * P.Beta * Rij * nfL) {
        rm = 0.525 * P.Beta * Rij * nfL * (1.0 / rmL) * rm;
        rh = rm;
    } else {
        rh = rm;
        rm = rm - Cr * (1.0 / P.StepSize) * DthR;
    }
    d3 tm = Kt * DthT + Mt;
    double tmL = Length(tm);
    d3 twh;
    if (tmL > 0.65 * P.Frcoe * P.Beta * Rij * nfL) {
        tm = 0.65 * P.Frcoe * P.Beta * Rij * nfL * (1.0 / tmL) * tm;
        twh = tm;
    } else {
        twh = tm;
        tm = tm - Ct * (1.0 / P.StepSize) * DthT;
    }
    ctc = ctc - rm - tm;
    ctc2 = ctc2 + rm + tm;
    c->CacheFc[k] = cfc;
    c->CacheTc[k] = ctc;
    c->CacheTc2[k] = ctc2;
    c->D.v[k] = th;
    c->D.t[k] = twh;
    c->D.r[k] = rh;
    c->CacheFcn[k] = -1.0 * nf;
    c->CacheCtp[k] = c->Pos[I] + Ap * DistU;
    return true;
}

static void sphere_forces(orc_ctx* c) {
    const int N = c->N, B = c->B;
    long long touch = 0;
    const long long S = c->compact ? c->S_used : c->S;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : touch)
    for (long long k = 0; k < S; ++k)
        if (ss_force(c, k)) touch++;
    c->touching_ss = touch;
    // SumContForce_clump (src/dem_force_clump.cu:165-191)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        const int* rowR = &c->Aref[(size_t)i * B];
        d3 F = c->Fc[i], T = c->Tc[i];
        for (int g = 0; g < c->Ng[i]; g++) {
            F = F + c->CacheFc[rowR[g]];
            T = T + c->CacheTc[rowR[g]];
        }
        for (int g = B - c->Ni[i]; g < B; g++) {
            const int s = rowR[g];
            if (s < 0) continue;
            F = F - c->CacheFc[s];
            T = T + c->CacheTc2[s];
        }
        c->Fc[i] = F;
        c->Tc[i] = T;
    }
}

// ---------------------------------------------------------------------------
// Wall contact for one face / surface: src/dem_force_cube_wall_clump.cu:16-171
// (cube) with Dn and the normal either fixed (cube, n_in) or from RVn
// (cylinder, src/dem_force_cylind_wall.cu:28-39).
struct WallKin { d3 pos, vel, angv; };

static bool wall_face(orc_ctx* c, int i, const WallKin& wk, const d3& RVo, const d3& RVn, const d3* n_in,
                      double Dn_cyl, Hist3& Hp, Hist3& Hc, d3& Fc, d3& Tc) {
    const nbc_params& P = c->P;
    const double Frcoe = P.Frcoe;
    const int clumpI = c->ClumpId[i];
    d3 DistU;
    double Ap, Dn;
    if (n_in) {
        DistU = *n_in;
        Ap = Dot(RVn, DistU);
        Dn = c->Radius[i] - Ap;
        if (Dn <= 0.0) return false;
    } else {
        d3 Dist = RVn;
        double DistS = Dist.x * Dist.x + Dist.y * Dist.y + Dist.z * Dist.z;
        double DistL = std::sqrt(DistS);
        double DistR = 1.0 / DistL;
        DistU = DistR * Dist;
        Ap = DistL;
        Dn = Dn_cyl;
    }
    d3 H = Hp.v[i], Mr = Hp.r[i], Mt = Hp.t[i];
    double Rij = c->Radius[i], Mij = c->Mass[i];
    double Kn = 2.0 * P.Ymod * std::sqrt(Rij * Dn) / (3.0 * (1.0 - P.Prat * P.Prat));
    double Cn = P.Damping * std::sqrt(Kn * Mij);
    double Ks = 2.0 * P.Ymod / (1.0 + P.Prat) / (2.0 - P.Prat) * std::sqrt(Rij * Dn);
    double Cs = 2.0 * P.Damping * std::sqrt(Ks * Mij);
    double Kr = 0.25 * Kn * (P.Beta * Rij) * (P.Beta * Rij);
    double Cr = 0.25 * Cn * (P.Beta * Rij) * (P.Beta * Rij);
    double Kt = 0.50 * Ks * (P.Beta * Rij) * (P.Beta * Rij);
    double Ct = 0.50 * Cs * (P.Beta * Rij) * (P.Beta * Rij);
    double M[3][3];
    Quat2IvDCM(c->CQuat[clumpI], M);
    d3 Arm = Map(M, c->LocalPos[i]);
    Arm = -1.0 * Arm + Ap * DistU;
    d3 Vwall = Cross(wk.angv, RVo);
    d3 Vrel = c->CVel[clumpI] - wk.vel;
    d3 Vrot = c->driver ? Ap * Cross(DistU, c->CAngSpd[clumpI]) - Vwall  // src/dem_force_cylind_wall.cu:55
                        : Cross(Arm, c->CAngSpd[clumpI]) - Vwall;
    d3 Vtot = Vrel + Vrot;
    double ERR = Dot(DistU, Vtot);
    d3 Vnor = ERR * DistU;
    d3 Vtan = Vtot - Vnor;
    d3 nf = Kn * Dn * DistU + Cn * Vnor;
    double nfL = Length(nf);
    d3 Ds = P.StepSize * Vtan;
    d3 tf = -Ks * Ds + Cs * Vtan + H;
    double tfL = Length(tf);
    d3 th;
    if (tfL > nfL * Frcoe) {
        tf = Frcoe * nfL * (1.0 / tfL) * tf;
        th = tf;
    } else {
        th = tf - Cs * Vtan;
    }
    Fc = Fc + nf + tf;
    Tc = Tc - Ap * Cross(DistU, tf);
    d3 Dth = P.StepSize * (c->CAngSpd[clumpI] - wk.angv);
    double DthL = Dot(Dth, DistU);
    d3 DthT = DthL * DistU;
    d3 DthR = Dth - DthT;
    d3 rm = Kr * DthR + Mr;
    double rmL = Length(rm);
    d3 rh;
    if (rmL > 0.525 * P.Beta * Rij * nfL) {
        rm = 0.525 * P.Beta * Rij * nfL * (1.0 / rmL) * rm;
        rh = rm;
    } else {
        rh = rm;
        rm = rm - Cr * (1.0 / P.StepSize) * DthR;
    }
    d3 tm = Kt * DthT + Mt;
    double tmL = Length(tm);
    d3 twh;
    if (tmL > 0.65 * Frcoe * P.Beta * Rij * nfL) {
        tm = 0.65 * Frcoe * P.Beta * Rij * nfL * (1.0 / tmL) * tm;
        twh = tm;
    } else {
        twh = tm;
        tm = tm - Ct * (1.0 / P.StepSize) * DthT;
    }
    Tc = Tc - rm - tm;
    Hc.v[i] = th;
    Hc.t[i] = twh;
    Hc.r[i] = rh;
    return true;
}

// DEMmain cube wall for one face: src/dem_force_cube_wall.cu:3-148 as it
// runs: normal and Ap from RVn, Dn = r - |RVn| from the caller (:170-183),
// histories read from arrays that are never written (zero), rolling and
// twisting moments computed but not applied.
static void wall_face_dm_cube(orc_ctx* c, int i, const WallKin& wk, const d3& RVo, const d3& RVn, double Dn,
                              d3& Fc, d3& Tc) {
    const nbc_params& P = c->P;
    const int clumpI = c->ClumpId[i];
    d3 Dist = RVn;
    double DistS = Dist.x * Dist.x + Dist.y * Dist.y + Dist.z * Dist.z;
    double DistL = std::sqrt(DistS);
    d3 H = d3{0, 0, 0};
    double DistR = 1.0 / DistL;
    d3 DistU = DistR * Dist;
    double Ap = DistL;
    double Rij = c->Radius[i], Mij = c->Mass[i];
    double Kn = 2.0 * P.Ymod * std::sqrt(Rij * Dn) / (3.0 * (1.0 - P.Prat * P.Prat));
    double Cn = P.Damping * std::sqrt(Kn * Mij);
    double Ks = 2.0 * P.Ymod / (1.0 + P.Prat) / (2.0 - P.Prat) * std::sqrt(Rij * Dn);
    double Cs = 2.0 * P.Damping * std::sqrt(Ks * Mij);
    d3 Vwall = Cross(wk.angv, RVo);
    d3 Vrel = c->CVel[clumpI] - wk.vel;
    d3 Vrot = Ap * Cross(DistU, c->CAngSpd[clumpI]) - Vwall;
    d3 Vtot = Vrel + Vrot;
    double ERR = Dot(DistU, Vtot);
    d3 Vnor = ERR * DistU;
    d3 Vtan = Vtot - Vnor;
    d3 nf = Kn * Dn * DistU + Cn * Vnor;
    double nfL = Length(nf);
    d3 Ds = P.StepSize * Vtan;
    d3 tf = -Ks * Ds + Cs * Vtan + H;
    double tfL = Length(tf);
    if (tfL > nfL * P.Frcoe) tf = P.Frcoe * nfL * (1.0 / tfL) * tf;
    Fc = Fc + nf + tf;
    Tc = Tc - Ap * Cross(DistU, tf);
}

static bool sphere_valid(orc_ctx* c, int i) {
    const int cid = c->ClumpId[i];
    if (cid < 0 || cid >= c->Nc) return false;
    const nbc_wall& w = c->W;
    d3 wp = mk3(w.pos[0], w.pos[1], w.pos[2]), wv = mk3(w.vel[0], w.vel[1], w.vel[2]),
       wa = mk3(w.angv[0], w.angv[1], w.angv[2]);
    bool geo;
    if (c->wall_type == NBC_WALL_CUBE)
        geo = std::isfinite(w.side[0]) && std::isfinite(w.side[1]) && std::isfinite(w.side[2]) && w.side[0] > 0.0 &&
              w.side[1] > 0.0 && w.side[2] > 0.0;
    else
        geo = std::isfinite(w.radius) && std::isfinite(w.height) && w.radius > 0.0 && w.height > 0.0;
    return finite3(c->Pos[i]) && finite3(c->LocalPos[i]) && finite3(c->CVel[cid]) && finite3(c->CAngSpd[cid]) &&
           finite4(c->CQuat[cid]) && std::isfinite(c->Radius[i]) && c->Radius[i] > 0.0 && std::isfinite(c->Mass[i]) &&
           c->Mass[i] > 0.0 && finite3(wp) && finite3(wv) && finite3(wa) && geo;
}

// src/dem_force_cube_wall_clump.cu:173-298, cylinder geometry from
// src/dem_force_cylind_wall.cu:152-237. Returns the smallest bad sorted index or -1.
static int wall_forces(orc_ctx* c) {
    const int N = c->N;
    const nbc_wall& w = c->W;
    WallKin wk{mk3(w.pos[0], w.pos[1], w.pos[2]), mk3(w.vel[0], w.vel[1], w.vel[2]),
               mk3(w.angv[0], w.angv[1], w.angv[2])};
    int bad = std::numeric_limits<int>::max();
    long long touch = 0;
#pragma omp parallel for schedule(static) reduction(min : bad) reduction(+ : touch)
    for (int i = 0; i < N; ++i) {
        if (!sphere_valid(c, i)) { bad = std::min(bad, i); continue; }
        d3 F = c->WFc[i], T = c->WTc[i];
        const double r = c->Radius[i];
        d3 RV = c->Pos[i] - wk.pos;
        if (c->wall_type == NBC_WALL_CUBE) {
            const d3 ex = mk3(1.0, 0.0, 0.0), ey = mk3(0.0, 1.0, 0.0), ez = mk3(0.0, 0.0, 1.0);
            const double ERRx = RV.x, ERRy = RV.y, ERRz = RV.z;
            d3 RVbx = RV - ERRx * ex, RVtx = ERRx * ex;
            d3 RVby = RV - ERRy * ey, RVty = ERRy * ey;
            d3 RVbz = RV - ERRz * ez, RVtz = ERRz * ez;
            const double* S = w.side;
            d3 RVo, RVn;
            const d3 nUx = mk3(-1.0, 0.0, 0.0), nLx = mk3(1.0, 0.0, 0.0);
            const d3 nUy = mk3(0.0, -1.0, 0.0), nLy = mk3(0.0, 1.0, 0.0);
            const d3 nUz = mk3(0.0, 0.0, -1.0), nLz = mk3(0.0, 0.0, 1.0);
            if (c->driver == NBC_DRIVER_CUBE) {
                // src/dem_force_cube_wall.cu:170-268, faces UpX DownX UpY DownY UpZ DownZ
                const double E[3] = {ERRx, ERRy, ERRz};
                const d3 RVt[3] = {RVtx, RVty, RVtz}, RVb[3] = {RVbx, RVby, RVbz};
                for (int a = 0; a < 3; ++a) {
                    for (int up = 1; up >= 0; --up) {
                        const bool cand = up ? (E[a] > 0 && E[a] + r - S[a] > 0) : (E[a] < 0 && E[a] - r + S[a] < 0);
                        if (!cand) continue;
                        RVo = (S[a] / std::fabs(E[a])) * RVt[a] + RVb[a];
                        RVn = RV - RVo;
                        const double DistL = Length(RVn);
                        wall_face_dm_cube(c, i, wk, RVo, RVn, r - DistL, F, T);
                        touch++;
                    }
                }
                c->WFc[i] = F;
                c->WTc[i] = T;
                if (!finite3(F) || !finite3(T)) bad = std::min(bad, i);
                continue;
            }
            // face order and history sets: 0 UpX, 1 LowX, 2 UpY, 3 LowY, 4 UpZ, 5 LowZ
            if (ERRx > 0 && ERRx + r - S[0] > 0) {
                RVo = (S[0] / std::fabs(ERRx)) * RVtx + RVbx; RVn = RV - RVo;
                touch += wall_face(c, i, wk, RVo, RVn, &nUx, 0, c->WHpre[0], c->WH[0], F, T);
            }
            if (ERRx < 0 && ERRx - r + S[0] < 0) {
                RVo = (S[0] / std::fabs(ERRx)) * RVtx + RVbx; RVn = RV - RVo;
                touch += wall_face(c, i, wk, RVo, RVn, &nLx, 0, c->WHpre[1], c->WH[1], F, T);
            }
            if (ERRy > 0 && ERRy + r - S[1] > 0) {
                RVo = (S[1] / std::fabs(ERRy)) * RVty + RVby; RVn = RV - RVo;
                touch += wall_face(c, i, wk, RVo, RVn, &nUy, 0, c->WHpre[2], c->WH[2], F, T);
            }
            if (ERRy < 0 && ERRy - r + S[1] < 0) {
                RVo = (S[1] / std::fabs(ERRy)) * RVty + RVby; RVn = RV - RVo;
                touch += wall_face(c, i, wk, RVo, RVn, &nLy, 0, c->WHpre[3], c->WH[3], F, T);
            }
            if (ERRz > 0 && ERRz + r - S[2] > 0) {
                RVo = (S[2] / std::fabs(ERRz)) * RVtz + RVbz; RVn = RV - RVo;
                touch += wall_face(c, i, wk, RVo, RVn, &nUz, 0, c->WHpre[4], c->WH[4], F, T);
            }
            if (ERRz < 0 && ERRz - r + S[2] < 0) {
                RVo = (S[2] / std::fabs(ERRz)) * RVtz + RVbz; RVn = RV - RVo;
                touch += wall_face(c, i, wk, RVo, RVn, &nLz, 0, c->WHpre[5], c->WH[5], F, T);
            }
        } else {
            const d3 Wv = mk3(w.dir[0], w.dir[1], w.dir[2]);
            double ERR = Dot(RV, Wv);
            d3 RVb = RV - ERR * Wv, RVt = ERR * Wv;
            double norm = Length(RVb);
            double d_cap = std::fabs(ERR) - w.height + r;
            double d_side = norm - w.radius + r;
            // 0 bottom cap (ERR < 0), 1 top (ERR > 0), 2 side
            if (d_cap >= 1e-14 && ERR < 0) {
                d3 RVo = (w.height / std::fabs(ERR)) * RVt + RVb, RVn = RV - RVo;
                double Dn = r - Length(RVn);
                if (Dn > 0.0) touch += wall_face(c, i, wk, RVo, RVn, nullptr, Dn, c->WHpre[0], c->WH[0], F, T);
            }
            if (d_cap >= 1e-14 && ERR > 0) {
                d3 RVo = (w.height / std::fabs(ERR)) * RVt + RVb, RVn = RV - RVo;
                double Dn = r - Length(RVn);
                if (Dn > 0.0) touch += wall_face(c, i, wk, RVo, RVn, nullptr, Dn, c->WHpre[1], c->WH[1], F, T);
            }
            if (d_side >= 1e-14) {
                d3 RVo = (w.radius / norm) * RVb + RVt, RVn = RV - RVo;
                double Dn = r - Length(RVn);
                if (Dn > 0.0) touch += wall_face(c, i, wk, RVo, RVn, nullptr, Dn, c->WHpre[2], c->WH[2], F, T);
            }
        }
        c->WFc[i] = F;
        c->WTc[i] = T;
        if (!finite3(F) || !finite3(T)) bad = std::min(bad, i);
    }
    c->touching_wall += touch;
    return bad == std::numeric_limits<int>::max() ? -1 : bad;
}

// ---------------------------------------------------------------------------
// Penetrator: src/dem_contact_penetrator.cu:8-84, src/dem_sphere_intruder.cuh:20-51,
// src/dem_force_penetrator_clump.cu:4-182, src/dem_force_penetrator.cu:189-248,
// src/dem_integrate_penetrator.cu:79-115. The cone is a triangle mesh (new).
struct SphereContact { d3 normal, surfaceArm, normalGapVector; double gap, overlap, particleArm; };

static SphereContact sphere_intruder(const d3& pp, double pr, const d3& ip, double ir, const d3& fb) {
    SphereContact k;
    d3 cv = pp - ip;
    double cd = Length(cv);
    if (cd > 1.0e-15) {
        k.normal = (1.0 / cd) * cv;
    } else {
        k.normal = Norm(fb);
        if (Length(k.normal) < 1.0e-15) k.normal = mk3(0.0, 0.0, 1.0);
    }
    k.surfaceArm = ir * k.normal;
    k.gap = cd - ir;
    k.overlap = pr - k.gap;
    k.particleArm = k.gap;
    k.normalGapVector = k.particleArm * k.normal;
    return k;
}

// Closest point on triangle abc to p (Ericson, Real-Time Collision
// Detection, 5.1.5), all in FP64.
static d3 closest_on_tri(const d3& p, const d3& a, const d3& b, const d3& cc) {
    d3 ab = b - a, ac = cc - a, ap = p - a;
    double d1 = Dot(ab, ap), d2 = Dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    d3 bp = p - b;
    double d3v = Dot(ab, bp), d4v = Dot(ac, bp);
    if (d3v >= 0.0 && d4v <= d3v) return b;
    double vc = d1 * d4v - d3v * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3v <= 0.0) { double v = d1 / (d1 - d3v); return a + v * ab; }
    d3 cp = p - cc;
    double d5 = Dot(ab, cp), d6 = Dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return cc;
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) { double w = d2 / (d2 - d6); return a + w * ac; }
    double va = d3v * d6 - d5 * d4v;
    if (va <= 0.0 && (d4v - d3v) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4v - d3v) / ((d4v - d3v) + (d5 - d6));
        return b + w * (cc - b);
    }
    double den = 1.0 / (va + vb + vc);
    double v = vb * den, w = vc * den;
    return a + v * ab + w * ac;
}

static void pene_precontact(orc_ctx* c) {
    const int N = c->N;
    nbc_penetrator& p = c->Pn;
    const d3 pp = c->pene_pos;
    const d3 dir = mk3(p.dir[0], p.dir[1], p.dir[2]);
    int newly = 0;
#pragma omp parallel for schedule(static) reduction(| : newly)
    for (int i = 0; i < N; ++i) {
        c->PLcFlg[i] = 0;
        if (p.shape == NBC_PENE_SPHERE) {
            SphereContact k = sphere_intruder(c->Pos[i], c->Radius[i], pp, p.radius, dir);
            if (k.overlap > 0.0) {
                c->PRVo[i] = k.surfaceArm;
                c->PRVn[i] = k.normalGapVector;
                c->PLcFlg[i] = 1;
                if (p.wait_for_contact) newly = 1;
            }
            continue;
        }
        d3 RV = c->Pos[i] - pp;
        if (p.shape == NBC_PENE_CONE) {
            // closest point on the mesh (penetrator frame = world frame offset by pos)
            const double reach = c->pene_bound + c->Radius[i];
            if (Dot(RV, RV) >= reach * reach) continue;
            const int nt = p.ntri;
            double best = std::numeric_limits<double>::infinity();
            d3 bq{0, 0, 0};
            for (int t = 0; t < nt; ++t) {
                const double* v = &c->pene_tri[9 * (size_t)t];
                d3 q = closest_on_tri(RV, mk3(v[0], v[1], v[2]), mk3(v[3], v[4], v[5]), mk3(v[6], v[7], v[8]));
                d3 dq = RV - q;
                double ds = Dot(dq, dq);
                if (ds < best) { best = ds; bq = q; }
            }
            d3 RVn = RV - bq;
            double Dn = c->Radius[i] - std::sqrt(Dot(RVn, RVn));
            if (nt > 0 && Dn > 0.0) {
                c->PRVo[i] = bq;
                c->PRVn[i] = RVn;
                c->PLcFlg[i] = 1;
            }
            continue;
        }
        const double LatDx = c->P.CellSize[0];
        double ERR = Dot(RV, dir);
        d3 RVn = mk3(0.0, 0.0, 0.0), RVo = mk3(0.0, 0.0, 0.0);
        bool enter = false;
        if (std::fabs(ERR) < (p.height + 0.80 * LatDx)) {
            d3 RVb = RV - ERR * dir;
            double norm = Length(RVb);
            if (norm < (p.radius + 0.80 * LatDx)) {
                d3 RVt = ERR * dir;
                if (std::fabs(ERR) < p.height) {
                    RVo = (1.0 / norm) * p.radius * RVb + RVt;
                } else if (norm < p.radius) {
                    RVo = (1.0 / std::fabs(ERR)) * p.height * RVt + RVb;
                } else {
                    RVo = (1.0 / norm) * p.radius * RVb + (1.0 / std::fabs(ERR)) * p.height * RVt;
                }
                RVn = RV - RVo;
                enter = true;
            }
        }
        if (enter) {
            double DistS = RVn.x * RVn.x + RVn.y * RVn.y + RVn.z * RVn.z;
            double Dn = c->Radius[i] - std::sqrt(DistS);
            if (Dn > 0.0) {
                c->PRVo[i] = RVo;
                c->PRVn[i] = RVn;
                c->PLcFlg[i] = 1;
            }
        }
    }
    if (newly && p.contact_state == 0) p.contact_state = 1;
}

static void pene_forces(orc_ctx* c) {
    const nbc_params& P = c->P;
    const int N = c->N;
    const nbc_penetrator& p = c->Pn;
    const double Frcoe = c->pene_frcoe;
    const d3 pp = c->pene_pos, pv = c->pene_vel, pw = c->pene_angv;
    const d3 dir = mk3(p.dir[0], p.dir[1], p.dir[2]);
    long long touch = 0;
#pragma omp parallel for schedule(static) reduction(+ : touch)
    for (int i = 0; i < N; ++i) {
        if (c->PLcFlg[i] == 0) continue;
        const int clumpI = c->ClumpId[i];
        d3 Dist = c->PRVn[i];
        d3 H = c->PHpre.v[i], Mr = c->PHpre.r[i], Mt = c->PHpre.t[i];
        double DistS = Dist.x * Dist.x + Dist.y * Dist.y + Dist.z * Dist.z;
        double DistL = std::sqrt(DistS);
        double Dn = c->Radius[i] - DistL;
        SphereContact sc{};
        if (p.shape == NBC_PENE_SPHERE) {
            sc = sphere_intruder(c->Pos[i], c->Radius[i], pp, p.radius, dir);
            Dist = sc.normalGapVector;
            DistL = sc.particleArm;
            Dn = sc.overlap;
        }
        if (!(Dn > 0.0)) continue;
        touch++;
        d3 DistU = (p.shape == NBC_PENE_SPHERE) ? sc.normal : (1.0 / DistL) * Dist;
        double Ap = DistL;
        double Rij = c->Radius[i];
        if (p.shape == NBC_PENE_SPHERE) Rij = c->Radius[i] * p.radius / (c->Radius[i] + p.radius);
        double Mij = c->Mass[i];
        if (p.motion == NBC_PENE_GRAVITY) Mij = c->Mass[i] * p.prop[3] / (c->Mass[i] + p.prop[3]);
        double Kn = 2.0 * P.Ymod * std::sqrt(Rij * Dn) / (3.0 * (1.0 - P.Prat * P.Prat));
        double Cn = P.Damping * std::sqrt(Kn * Mij);
        double Ks = 2.0 * P.Ymod / (1.0 + P.Prat) / (2.0 - P.Prat) * std::sqrt(Rij * Dn);
        double Cs = P.Damping * std::sqrt(Ks * Mij);
        double Kr = 0.25 * Kn * (P.Beta * Rij) * (P.Beta * Rij);
        double Cr = 0.25 * Cn * (P.Beta * Rij) * (P.Beta * Rij);
        double Kt = 0.50 * Ks * (P.Beta * Rij) * (P.Beta * Rij);
        double Ct = 0.50 * Cs * (P.Beta * Rij) * (P.Beta * Rij);
        double M[3][3];
        Quat2IvDCM(c->CQuat[clumpI], M);
        d3 Arm = Map(M, c->LocalPos[i]);
        Arm = -1.0 * Arm + Ap * DistU;
        d3 Vwall = Cross(pw, c->PRVo[i]);
        d3 Vrel = c->CVel[clumpI] - pv;
        // DEMmain loops: the sphere's own spin and the cohesion term
        // (src/dem_force_penetrator.cu:91, 119-121)
        d3 Vrot = c->driver ? Ap * Cross(DistU, c->CAngSpd[clumpI]) - Vwall : Cross(Arm, c->CAngSpd[clumpI]) - Vwall;
        d3 Vtot = Vrel + Vrot;
        double ERR = Dot(DistU, Vtot);
        d3 Vnor = ERR * DistU;
        d3 Vtan = Vtot - Vnor;
        d3 nf = Kn * Dn * DistU + Cn * Vnor;
        double nfL = Length(nf);
        d3 Ds = P.StepSize * Vtan;
        d3 tf = -Ks * Ds + Cs * Vtan + H;
        double tfL = Length(tf);
        d3 th;
        if (tfL > nfL * Frcoe) {
            tf = Frcoe * nfL * (1.0 / tfL) * tf;
            th = tf;
        } else {
            th = tf - Cs * Vtan;
        }
        d3 cf = c->driver ? -P.Cohesion * (P.Beta * Rij) * (P.Beta * Rij) * DistU : mk3(0.0, 0.0, 0.0);
        d3 Fc = c->driver ? nf + tf + cf : nf + tf;
        d3 Tc = -Ap * Cross(DistU, tf);
        d3 pTc = -1.0 * Cross(c->PRVo[i], (nf + tf + cf));
        c->PFcn[i] = nf;
        c->PFct[i] = tf;
        d3 Dth = P.StepSize * (c->CAngSpd[clumpI] - pw);
        double DthL = Dot(Dth, DistU);
        d3 DthT = DthL * DistU;
        d3 DthR = Dth - DthT;
        d3 rm = Kr * DthR + Mr;
        double rmL = Length(rm);
        d3 rh;
        if (rmL > 0.525 * P.Beta * Rij * nfL) {
            rm = 0.525 * P.Beta * Rij * nfL * (1.0 / rmL) * rm;
            rh = rm;
        } else {
            rh = rm;
            rm = rm - Cr * (1.0 / P.StepSize) * DthR;
        }
        d3 tm = Kt * DthT + Mt;
        double tmL = Length(tm);
        d3 twh;
        if (tmL > 0.65 * Frcoe * P.Beta * Rij * nfL) {
            tm = 0.65 * Frcoe * P.Beta * Rij * nfL * (1.0 / tmL) * tm;
            twh = tm;
        } else {
            twh = tm;
            tm = tm - Ct * (1.0 / P.StepSize) * DthT;
        }
        Tc = Tc - rm - tm;
        pTc = pTc + rm + tm;
        c->PFc[i] = Fc;
        c->PTc[i] = Tc;
        c->PpeneTc[i] = pTc;
        c->PH.v[i] = th;
        c->PH.t[i] = twh;
        c->PH.r[i] = rh;
        c->PTcn[i] = Cross(nf, Arm);
        c->PTct[i] = Cross(tf, Arm) - rm - tm;
    }
    c->touching_wall += touch;
    // SumForce_cuk: FP64 global sums (sequential ascending order here).
    d3 sF{0, 0, 0}, sT{0, 0, 0};
    for (int i = 0; i < N; ++i) { sF = sF + c->PFc[i]; sT = sT + c->PpeneTc[i]; }
    c->totFc = sF;
    c->totTc = sT;
}

// src/dem_integrate_penetrator.cu:79-115
static void pene_integrate_after(orc_ctx* c) {
    nbc_penetrator& p = c->Pn;
    if (p.wait_for_contact && p.contact_state == 1) p.contact_state = 2;
    d3 reaction = -1.0 * c->totFc;
    if (p.motion == NBC_PENE_GRAVITY && (!p.wait_for_contact || p.contact_state != 0)) {
        const double Dt = c->P.StepSize;
        d3 peneF = (1.0 / p.prop[3]) * reaction;
        peneF.z = peneF.z + c->P.Gravity;
        c->pene_vel = c->pene_vel + 0.5 * Dt * peneF;
    }
    const int idx = c->count % (int)c->CacheFi.size();
    c->CacheFi[idx] = reaction;
    c->CacheTi[idx] = c->totTc;
}

// ---------------------------------------------------------------------------
static void clump_kick(orc_ctx* c);

// src/dem_clump.cu:89-101 and src/dem_clumpsphere.cu:86-151, then
// src/dem_integrate_clumpsphere.cu:110-141.
static void clump_reduce_and_kick(orc_ctx* c) {
    const int N = c->N, Nc = c->Nc;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        c->totF[i] = c->Fc[i] + c->WFc[i] + c->PFc[i];
        c->totT[i] = c->Tc[i] + c->WTc[i] + c->PTc[i];
    }
    std::vector<d3> tcon(N);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        const int cid = c->ClumpId[i];
        const d3 ps = c->Pos[i], pc = c->CPos[cid];
        d3 r = mk3(ps.x - pc.x, ps.y - pc.y, ps.z - pc.z);
        d3 fs = c->totF[i], ts = c->totT[i];
        tcon[i] = mk3(ts.x + (r.y * fs.z - r.z * fs.y), ts.y + (r.z * fs.x - r.x * fs.z),
                      ts.z + (r.x * fs.y - r.y * fs.x));
    }
    // per clump, ascending sorted sphere index; clumps without spheres keep F,T
    std::vector<char> seen(Nc, 0);
    std::vector<d3> F(Nc), T(Nc);
    for (int i = 0; i < N; ++i) {
        const int cid = c->ClumpId[i];
        if (cid < 0) continue;
        if (!seen[cid]) { seen[cid] = 1; F[cid] = c->totF[i]; T[cid] = tcon[i]; }
        else { F[cid] = F[cid] + c->totF[i]; T[cid] = T[cid] + tcon[i]; }
    }
    for (int k = 0; k < Nc; ++k)
        if (seen[k]) { c->CFc[k] = F[k]; c->CTc[k] = T[k]; }
    clump_kick(c);
}

// src/dem_integrate_clumpsphere.cu:110-141: second half-kick from CFc/CTc.
static void clump_kick(orc_ctx* c) {
    const int N = c->N, Nc = c->Nc;
    const double Dt = c->P.StepSize, g = c->P.Gravity;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nc; ++i) {
        if (c->driver == NBC_DRIVER_SPLIT && c->Type[i] != 0) continue;  // src/dem_integrate_withglue.cu:139
        d3 Fi = c->CFc[i], Ti = c->CTc[i];
        const d4 pr = c->CProp[i];
        Fi = (1.0 / pr.w) * Fi;
        Fi.z = Fi.z + g;
        Ti.x = (1.0 / pr.x) * Ti.x;
        Ti.y = (1.0 / pr.y) * Ti.y;
        Ti.z = (1.0 / pr.z) * Ti.z;
        c->CVel[i] = c->CVel[i] + 0.5 * Dt * Fi;
        c->CAngSpd[i] = c->CAngSpd[i] + 0.5 * Dt * Ti;
    }
    if (c->driver && c->P.isIntruder) {
        // DEMmain loops: the next drift's half-kick and the attitude take
        // Fc + Fw and Tc + Tw only (src/dem_integrate.cu:23-25, 74-75); the
        // clump is its one sphere at its centre
        for (int i = 0; i < N; ++i) {
            const int cid = c->ClumpId[i];
            c->CFc[cid] = c->Fc[i] + c->WFc[i];
            c->CTc[cid] = c->Tc[i] + c->WTc[i];
        }
    }
}

static void zero3(std::vector<d3>& v) { std::fill(v.begin(), v.end(), d3{0, 0, 0}); }

// The rebuild branch of the loop body (src/dem_simulation_clump.cu:121-175).
static int rebuild_in_step(orc_ctx* c) {
    int e = rebuild(c);
    if (e) return e;
    if (c->driver == NBC_DRIVER_CYLINDER || c->driver == NBC_DRIVER_SPLIT) {
        // src/dem_simulation_cylind.cu:141: ForceExchange follows the rebuild
        // too, swapping the inherited histories out (and clearing them)
        c->D.swap(c->Dpre);
        if (c->compact) zero_range(c->D, c->S_high);
        else c->D.zero();
    }
    return 0;
}

// One pass of the loop body of src/dem_simulation_clump.cu:110-226.
static int step_once(orc_ctx* c) {
    integrate_before(c);
    update_pos(c);
    if (c->P.isIntruder) pene_integrate_before(c);
    const int nf = c->wall_type == NBC_WALL_CUBE ? 6 : (c->wall_type == NBC_WALL_CYLINDER ? 3 : 0);
    if (c->Update != 0 || c->first || c->rebuild_next) {
        c->rebuild_next = false;
        int e = rebuild_in_step(c);
        if (e) return e;
    } else if (c->rebuilt_between) {
        // orc_rebuild stood in for this step's rebuild: the inherited
        // histories are already the ones this step reads
    } else {
        c->D.swap(c->Dpre);  // DEMForce::ForceExchange
        if (c->compact) zero_range(c->D, c->S_used);
        else c->D.zero();
    }
    c->first = false;
    c->rebuilt_between = false;
    for (int f = 0; f < nf; ++f) { c->WH[f].swap(c->WHpre[f]); c->WH[f].zero(); }
    if (c->P.isIntruder) { c->PH.swap(c->PHpre); c->PH.zero(); }
    // ForceZero (src/dem_force.cu:75-86, src/dem_cube_wall.cu:208-211)
    zero3(c->Fc); zero3(c->Tc);
    if (c->compact) {
        for (auto* v : {&c->CacheFc, &c->CacheTc, &c->CacheTc2, &c->CacheFcn, &c->CacheCtp}) zero_range(*v, c->S_used);
    } else {
        zero3(c->CacheFc); zero3(c->CacheTc); zero3(c->CacheTc2); zero3(c->CacheFcn); zero3(c->CacheCtp);
    }
    zero3(c->WFc); zero3(c->WTc);
    c->touching_wall = 0;
    sphere_forces(c);
    if (nf) {
        int bad = wall_forces(c);
        if (bad >= 0) {
            char m[200];
            std::snprintf(m, sizeof m,
                          "FATAL: Invalid input or nonfinite wall force/torque at sorted particle index %d. "
                          "Simulation stopped.",
                          bad);
            return fail(c, NBC_E_NONFINITE, m);
        }
    }
    if (c->P.isIntruder) {
        for (auto* v : {&c->PFc, &c->PTc, &c->PpeneTc, &c->PFcn, &c->PFct, &c->PTcn, &c->PTct}) zero3(*v);
        pene_precontact(c);
        pene_forces(c);
        pene_integrate_after(c);
    }
    clump_reduce_and_kick(c);
    c->count++;
    if (c->P.OutputInterval > 0 && c->count == c->P.OutputInterval) c->count = 0;
    c->step++;
    return 0;
}

extern "C" int orc_step(orc_ctx* c, int n) {
    for (int k = 0; k < n; ++k) {
        int e = step_once(c);
        if (e) return e;
    }
    return 0;
}

// Compact mode (CSR pair range instead of the slot-padded layout); the
// results are identical, only the per-step clears and slot loops shrink.
extern "C" void orc_set_compact(orc_ctx* c, int on) { c->compact = on != 0; }

// Test hook for the integrator alone (no contacts): nsteps of the clump
// half-kick + drift + attitude (src/dem_integrate_clumpsphere.cu:3-90) and the
// second half-kick (:110-141) with each clump's force and torque held at the
// given values (3*Nc doubles each, world frame). Sphere poses are not updated.
extern "C" int orc_integrate_only(orc_ctx* c, int nsteps, const double* F, const double* T) {
    for (int k = 0; k < nsteps; ++k) {
        for (int i = 0; i < c->Nc; ++i) {
            c->CFc[i] = mk3(F[3 * i], F[3 * i + 1], F[3 * i + 2]);
            c->CTc[i] = mk3(T[3 * i], T[3 * i + 1], T[3 * i + 2]);
        }
        integrate_before(c);
        clump_kick(c);
        c->step++;
    }
    return 0;
}

// The next step rebuilds whatever the displacements say: a batch of beds that
// share one device context rebuilds all of them when any one needs it, and
// this is how a single-bed oracle follows it.
extern "C" void orc_request_rebuild(orc_ctx* c) { c->rebuild_next = true; }

// A rebuild between steps, on the positions the last step left: the next step
// then skips its history exchange, as a step that rebuilds does, so the
// inherited histories carry on (what a rebuild inside that step would give,
// on these positions).
extern "C" int orc_rebuild(orc_ctx* c) {
    int e = rebuild_in_step(c);
    c->first = false;
    c->rebuilt_between = e == 0;
    return e;
}

// ---------------------------------------------------------------------------
// getters (sorted order; double3 packed)
static void put3(double* o, const std::vector<d3>& v) {
    if (!o) return;
    for (size_t i = 0; i < v.size(); ++i) { o[3 * i] = v[i].x; o[3 * i + 1] = v[i].y; o[3 * i + 2] = v[i].z; }
}
static void put4(double* o, const std::vector<d4>& v) {
    if (!o) return;
    for (size_t i = 0; i < v.size(); ++i) {
        o[4 * i] = v[i].x; o[4 * i + 1] = v[i].y; o[4 * i + 2] = v[i].z; o[4 * i + 3] = v[i].w;
    }
}

extern "C" int orc_get_spheres(orc_ctx* c, int32_t* fid, int32_t* cid, double* pos, double* cum, double* local,
                               double* mass, double* radius) {
    if (fid) std::copy(c->part_fix.begin(), c->part_fix.end(), fid);
    if (cid) std::copy(c->ClumpId.begin(), c->ClumpId.end(), cid);
    put3(pos, c->Pos);
    put3(cum, c->CumPos);
    put3(local, c->LocalPos);
    if (mass) std::copy(c->Mass.begin(), c->Mass.end(), mass);
    if (radius) std::copy(c->Radius.begin(), c->Radius.end(), radius);
    return 0;
}

extern "C" int orc_get_grid(orc_ctx* c, int32_t* cell, int32_t* perm, int32_t* beg, int32_t* end) {
    if (cell) std::copy(c->part_cel.begin(), c->part_cel.end(), cell);
    if (perm) std::copy(c->part_idx_pre.begin(), c->part_idx_pre.end(), perm);
    if (beg) std::copy(c->cell_beg.begin(), c->cell_beg.end(), beg);
    if (end) std::copy(c->cell_end.begin(), c->cell_end.end(), end);
    return 0;
}

extern "C" int orc_get_pair_counts(orc_ctx* c, int64_t* nf, int64_t* nb) {
    long long a = 0, b = 0;
    for (int i = 0; i < c->N; ++i) { a += c->Ng[i]; b += c->Ni[i]; }
    if (nf) *nf = a;
    if (nb) *nb = b;
    return 0;
}

// Front rows in discovery order; back rows in the gather order of
// SumContForce_clump (slot index B-Ni .. B-1).
extern "C" int orc_get_pairs(orc_ctx* c, int32_t* foff, int32_t* fj, uint8_t* fflag, int32_t* boff, int32_t* bi,
                             int32_t* bslot) {
    const int N = c->N, B = c->B;
    int32_t a = 0, b = 0;
    for (int i = 0; i < N; ++i) {
        if (foff) foff[i] = a;
        if (boff) boff[i] = b;
        const int* row = &c->Anei[(size_t)i * B];
        const int* ref = &c->Aref[(size_t)i * B];
        for (int g = 0; g < c->Ng[i]; ++g, ++a) {
            if (fj) fj[a] = row[g];
            if (fflag) fflag[a] = (uint8_t)(c->ClumpId[i] == c->ClumpId[row[g]] ? 1 : 0);
        }
        for (int g = B - c->Ni[i]; g < B; ++g, ++b) {
            if (bi) bi[b] = row[g];
            if (bslot) bslot[b] = ref[g];
        }
    }
    if (foff) foff[N] = a;
    if (boff) boff[N] = b;
    return 0;
}

extern "C" int orc_get_pair_history(orc_ctx* c, double* h9) {
    long long np = 0;
    for (int i = 0; i < c->N; ++i) np += c->Ng[i];
    for (long long k = 0; k < np; ++k) {
        const d3 v = c->D.v[k], t = c->D.t[k], r = c->D.r[k];
        double* o = h9 + 9 * k;
        o[0] = v.x; o[1] = v.y; o[2] = v.z;
        o[3] = t.x; o[4] = t.y; o[5] = t.z;
        o[6] = r.x; o[7] = r.y; o[8] = r.z;
    }
    return 0;
}

// kind: 0 total (contact + wall + penetrator), 1 contact only, 2 wall only, 3 penetrator only
extern "C" int orc_get_sphere_forces(orc_ctx* c, int kind, double* f, double* t) {
    switch (kind) {
        case 0: put3(f, c->totF); put3(t, c->totT); break;
        case 1: put3(f, c->Fc); put3(t, c->Tc); break;
        case 2: put3(f, c->WFc); put3(t, c->WTc); break;
        case 3: put3(f, c->PFc); put3(t, c->PTc); break;
        default: return NBC_E_BADINPUT;
    }
    return 0;
}

extern "C" int orc_get_clumps(orc_ctx* c, double* pos, double* vel, double* quat, double* angv, double* F,
                              double* T) {
    put3(pos, c->CPos); put3(vel, c->CVel); put4(quat, c->CQuat); put3(angv, c->CAngSpd);
    put3(F, c->CFc); put3(T, c->CTc);
    return 0;
}

extern "C" int orc_get_penetrator(orc_ctx* c, double* pos, double* vel, double* cache_f, double* cache_t,
                                  int32_t* state, double* totf, double* tott) {
    if (pos) { pos[0] = c->pene_pos.x; pos[1] = c->pene_pos.y; pos[2] = c->pene_pos.z; }
    if (vel) { vel[0] = c->pene_vel.x; vel[1] = c->pene_vel.y; vel[2] = c->pene_vel.z; }
    put3(cache_f, c->CacheFi);
    put3(cache_t, c->CacheTi);
    if (state) *state = c->Pn.contact_state;
    if (totf) { totf[0] = c->totFc.x; totf[1] = c->totFc.y; totf[2] = c->totFc.z; }
    if (tott) { tott[0] = c->totTc.x; tott[1] = c->totTc.y; tott[2] = c->totTc.z; }
    return 0;
}

extern "C" int orc_get_pene_exports(orc_ctx* c, double* fn, double* ft, double* tn, double* tt) {
    put3(fn, c->PFcn); put3(ft, c->PFct); put3(tn, c->PTcn); put3(tt, c->PTct);
    return 0;
}

// Per front slot normal force and contact point (CacheFcn/CacheCtp), CSR order.
extern "C" int orc_get_pair_exports(orc_ctx* c, double* fcn, double* cpt) {
    long long np = 0;
    for (int i = 0; i < c->N; ++i) np += c->Ng[i];
    for (long long k = 0; k < np; ++k) {
        if (fcn) { fcn[3 * k] = c->CacheFcn[k].x; fcn[3 * k + 1] = c->CacheFcn[k].y; fcn[3 * k + 2] = c->CacheFcn[k].z; }
        if (cpt) { cpt[3 * k] = c->CacheCtp[k].x; cpt[3 * k + 1] = c->CacheCtp[k].y; cpt[3 * k + 2] = c->CacheCtp[k].z; }
    }
    return 0;
}

extern "C" int orc_get_stats(orc_ctx* c, nbc_stats* s) {
    std::memset(s, 0, sizeof(*s));
    s->step = c->step;
    s->rebuilds = c->rebuilds;
    int64_t a, b;
    orc_get_pair_counts(c, &a, &b);
    s->front_pairs = a;
    s->back_entries = b;
    s->touching_ss = c->touching_ss;
    s->touching_wall = c->touching_wall;
    double ke = 0.0;
    for (int i = 0; i < c->Nc; ++i) {
        const d3 v = c->CVel[i], w = c->CAngSpd[i];
        const d4 p = c->CProp[i];
        ke += 0.5 * p.w * (v.x * v.x + v.y * v.y + v.z * v.z) + 0.5 * (p.x * w.x * w.x + p.y * w.y * w.y + p.z * w.z * w.z);
    }
    s->kinetic_energy = ke;
    s->pene_force[0] = -c->totFc.x; s->pene_force[1] = -c->totFc.y; s->pene_force[2] = -c->totFc.z;
    s->pene_torque[0] = c->totTc.x; s->pene_torque[1] = c->totTc.y; s->pene_torque[2] = c->totTc.z;
    return 0;
}

// Brute-force O(N^2) list for pinning the grid search on small inputs: pairs
// (i<j) with d^2 < (1.5 (ri+rj))^2, both in sorted order, row-major by i then j.
extern "C" long long orc_bruteforce_pairs(orc_ctx* c, int32_t* out_i, int32_t* out_j, long long cap) {
    long long n = 0;
    for (int i = 0; i < c->N; ++i) {
        if (c->part_cel[i] == c->P.CellNum - 1) continue;
        for (int j = i + 1; j < c->N; ++j) {
            const d3 pi = c->Pos[i], pj = c->Pos[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            double ds = dx * dx + dy * dy + dz * dz;
            double lim = 1.50 * (c->Radius[i] + c->Radius[j]);
            if (ds < lim * lim) {
                if (n < cap) { out_i[n] = i; out_j[n] = j; }
                n++;
            }
        }
    }
    return n;
}
