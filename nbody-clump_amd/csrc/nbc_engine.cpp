// Device context and step orchestration behind the nbc C ABI.
//
// The loop body of src/dem_simulation_clump.cu:110-226 becomes, per step:
//   K2 k_sphere_pose (sets the rebuild / near-list flags),
//   k_contacts (sphere-sphere contacts and per-sphere totals; the wall-list
//   spheres' wall and penetrator contacts in extra blocks of the same launch),
//   k_pene_finalize (penetrator runs only),
//   K4 k_clump_after, which also runs K1 (k_clump_before) of the next step
//   inside one nbc_step call.
// Steps are enqueued in chunks; status words copied back after each chunk
// report the first step that needs a rebuild (or new near or wall lists),
// which is then done and the step replayed from its contact kernel, and the
// invalid-input index of the reference's fatal wall checks
// (src/dem_force_cube_wall_clump.cu:322-331), which surfaces as NBC_E_NONFINITE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "nbc_engine.h"
#include "nbc_kernels.h"

namespace nbc {

#define HIPC(x)                                                                 \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) return set_hip_error(e_, #x, __LINE__);           \
    } while (0)
#define KC(x)                                                                   \
    do {                                                                        \
        int e_ = (x);                                                           \
        if (e_ != 0) return set_hip_error((hipError_t)e_, #x, __LINE__);        \
    } while (0)

int EngineBase::set_hip_error(hipError_t e, const char* what, int line) {
    char b[512];
    std::snprintf(b, sizeof b, "HIP error %d (%s) in %s [nbc_engine.cpp:%d]", (int)e, hipGetErrorString(e), what, line);
    err = b;
    return NBC_E_HIP;
}
int EngineBase::fail(int code, const std::string& m) {
    err = m;
    return code;
}

static void dfree(void* p) {
    if (p) (void)hipFree(p);
}

template <class T>
static hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

static int bits_for(long long maxkey) {
    int b = 1;
    while (b < 31 && (1ll << b) <= maxkey) ++b;
    return b;
}

template <class R>
class Engine : public EngineBase {
  public:
    typedef typename Vec4T<R>::type R4;

    Engine() {}
    ~Engine() override { release(); }

    int init(const nbc_params& p, const nbc_options& o, int dev) override {
        P = p;
        O = o;
        device = dev;
        HIPC(hipSetDevice(device));
        HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        N = P.ParticleNum;
        Nc = P.ClumpNum;
        if (N <= 0 || Nc <= 0 || P.BranchLmtLen < 2 || P.CellNum <= 0)
            return fail(NBC_E_BADINPUT, "nbc_create: ParticleNum, ClumpNum, BranchLmtLen and the mesh must be set "
                                        "(call nbc_finalize_params first)");
        if ((long long)N > (long long)JMASK)
            return fail(NBC_E_BADINPUT, "nbc_create: too many spheres for the 29-bit pair index");
        if (P.Levels > 1 && P.LevelNested) {
            // nested levels: each coarse cell owns a block of sum 8^t keys
            bool ok = P.Levels <= 4;
            long long blk = 0;
            for (int t = 0; ok && t < P.Levels; ++t) {
                ok = P.LevelOff[t] == blk && P.LevelCell[t] > 0.0;
                for (int a = 0; a < 3; ++a) ok = ok && P.LevelDim[t][a] == (int)(P.GridDim[a] << t);
                blk += 1ll << (3 * t);
            }
            const long long coarse = (long long)P.GridDim[0] * P.GridDim[1] * P.GridDim[2];
            if (!ok || blk != P.LevelBlock || coarse * blk + 1 != P.CellNum)
                return fail(NBC_E_BADINPUT, "nbc_create: inconsistent nested hierarchical grid table (call "
                                            "nbc_finalize_params with hgrid_levels and hgrid_nested)");
            if (O.level_from_type)
                return fail(NBC_E_BADINPUT, "nbc_create: grid.bt level tables are not nested");
        } else if (P.Levels > 1) {
            // the kernels index cells through this table: it must tile [0, CellNum - 1)
            bool ok = P.Levels <= MAX_LEVELS;
            long long off = 0;
            for (int t = 0; ok && t < P.Levels; ++t) {
                ok = P.LevelOff[t] == off && P.LevelCell[t] > 0.0;
                long long c = 1;
                for (int a = 0; a < 3; ++a) { ok = ok && P.LevelDim[t][a] > 0; c *= P.LevelDim[t][a]; }
                off += c;
            }
            if (!ok || off + 1 != P.CellNum)
                return fail(NBC_E_BADINPUT, "nbc_create: inconsistent hierarchical grid table (call "
                                            "nbc_finalize_params with hgrid_levels)");
        }
        block = O.block_size > 0 ? O.block_size : 256;
        if (block < 64 || block > 1024 || (block % 64) != 0)
            return fail(NBC_E_BADINPUT, "nbc_create: block_size must be a multiple of 64 in [64, 1024]");
        if (block > 256) block = 256;  // the per-sphere kernels are tuned and tested up to 256 threads
        rbits = O.radix_bits > 0 ? O.radix_bits : 8;
        if (rbits > 8) rbits = 8;
        wall_type = P.isWallForce ? (O.wall_type ? O.wall_type : NBC_WALL_CUBE) : NBC_WALL_NONE;
        driver = O.driver;
        if (driver < NBC_DRIVER_CLUMP || driver > NBC_DRIVER_HGRID)
            return fail(NBC_E_BADINPUT, "nbc_create: unknown driver");
        if (driver != NBC_DRIVER_CLUMP) {
            // DEMmain loops: one sphere per particle, their wall force is applied
            // unconditionally (src/dem_simulation_cube.cu:175, src/dem_simulation_cylind.cu:157)
            if (Nc != N) return fail(NBC_E_BADINPUT, "nbc_create: DEMmain drivers need ClumpNum == ParticleNum");
            if ((driver == NBC_DRIVER_HGRID) != (P.Levels > 1))
                return fail(NBC_E_BADINPUT, "nbc_create: the HGRID driver, and only it, uses the hierarchical grid "
                                            "(nbc_finalize_params with hgrid_levels)");
            wall_type = driver == NBC_DRIVER_CUBE ? NBC_WALL_CUBE : NBC_WALL_CYLINDER;
        }
        if (O.level_from_type && driver != NBC_DRIVER_HGRID)
            return fail(NBC_E_BADINPUT, "nbc_create: level_from_type belongs to the HGRID driver");
        nfaces = wall_type == NBC_WALL_CUBE ? 6 : (wall_type == NBC_WALL_CYLINDER ? 3 : 0);
        pair_cap_per = O.pair_capacity > 0 ? O.pair_capacity : 16;
        // small beds are launch-bound: their steady chunks replay a captured
        // graph of 16 steps; large beds keep 4-step eager chunks (a halt then
        // wastes fewer no-op steps)
        use_graphs = O.graphs > 0 || (O.graphs == 0 && P.ParticleNum <= 262144);
        batch = O.batch_steps > 0 ? O.batch_steps : (use_graphs ? 16 : 4);
        // device parameters
        std::memset(&dp, 0, sizeof(dp));
        dp.N = N; dp.Nc = Nc; dp.B = P.BranchLmtLen; dp.CellNum = P.CellNum;
        for (int a = 0; a < 3; ++a) {
            dp.G[a] = P.GridDim[a];
            dp.origin[a] = P.WorldOrigin[a];
            dp.cell[a] = P.CellSize[a];
            dp.icell[a] = P.i_CellSize[a];
            dp.boxlo[a] = P.BoxLowLmt[a];
            dp.boxup[a] = P.BoxUpLmt[a];
        }
        dp.L = P.Levels > 1 ? P.Levels : 1;
        for (int t = 0; t < MAX_LEVELS; ++t) {
            const bool on = t < dp.L && dp.L > 1;
            dp.loff[t] = on ? P.LevelOff[t] : 0;
            for (int a = 0; a < 3; ++a) dp.ldim[t][a] = on ? P.LevelDim[t][a] : 0;
            dp.lcell[t] = on ? P.LevelCell[t] : 0.0;
            dp.lrmin[t] = on ? P.LevelRmin[t] : 0.0;
        }
        dp.nested = dp.L > 1 && P.LevelNested ? 1 : 0;
        dp.lblock = dp.nested ? P.LevelBlock : 0;
        dp.dt = P.StepSize; dp.Ymod = P.Ymod; dp.Prat = P.Prat; dp.Frcoe = P.Frcoe; dp.Beta = P.Beta;
        dp.Cohesion = P.Cohesion; dp.Gravity = P.Gravity; dp.Damping = P.Damping; dp.Thresh = P.Thresh;
        dp.pene_frcoe = O.pene_frcoe > 0.0 ? O.pene_frcoe : P.Frcoe;
        dp.cKn = 2.0 * P.Ymod / (3.0 * (1.0 - P.Prat * P.Prat));
        dp.cKs = 2.0 * P.Ymod / (1.0 + P.Prat) / (2.0 - P.Prat);
        dp.idt = 1.0 / P.StepSize;
        dp.half = 0.0;
        for (int a = 0; a < 3; ++a) {
            dp.pc[a] = 0.5 * (P.BoxLowLmt[a] + P.BoxUpLmt[a]);
            dp.half = std::max(dp.half, 0.5 * (P.BoxUpLmt[a] - P.BoxLowLmt[a]));
        }
        dp.f32_marg = (float)(4.0 * dp.half * std::ldexp(1.0, -23));
        // pairs farther apart than 2 * Thresh at a rebuild cannot touch before the next one
        const bool thresh_ok = std::isfinite(P.Thresh) && P.Thresh >= 0.0;
        const double inf = std::numeric_limits<double>::infinity();
        // near-list skin delta (a fraction of Thresh): smaller lists for the
        // contact kernel, refreshed when a sphere leaves its allowance. By
        // default each sphere's delta is the distance its clump's speed bound
        // covers in near_horizon steps: slow spheres keep short lists, fast
        // ones ask for fewer refreshes.
        const double skin = O.near_skin > 0.0 ? std::min(O.near_skin, 1.0) : 0.25;
        dp.near_delta = thresh_ok ? skin * P.Thresh : inf;
        dp.near_horizon = O.near_horizon > 0.0 ? O.near_horizon
                          : (O.near_horizon == 0.0 && !(O.near_skin > 0.0)) ? 500.0 : 0.0;
        if (!thresh_ok || !(std::isfinite(dp.near_horizon) && std::isfinite(dp.dt))) dp.near_horizon = 0.0;
        dp.near_lo = 0.03 * P.Thresh;
        dp.near_hi = P.Thresh;
        // wall list: spheres move at most 2 * Thresh after a list build; the
        // penetrator may move pene_move before the list is rebuilt
        dp.wl_slack = thresh_ok ? 2.0 * P.Thresh * (1.0 + 1e-9) + 1e-12 : inf;
        dp.pene_move = thresh_ok ? 8.0 * P.Thresh : inf;
        if (driver == NBC_DRIVER_SPLIT) {
            // glue loop: rebuilds at 0.1 r, and driven spheres move without
            // asking for one: every list pair is near, every sphere wall-listed
            dp.near_delta = inf;
            dp.near_horizon = 0.0;
            dp.wl_slack = inf;
        }
        dp.wall_type = wall_type; dp.nfaces = nfaces;
        dp.driver = driver;
        dp.pene_on = P.isIntruder ? 1 : 0;
        dp.pene_face = nfaces;
        nwh = nfaces + (dp.pene_on ? 1 : 0);
        oi = P.OutputInterval > 0 ? P.OutputInterval : 1;
        // spheres (double-buffered for the rebuild gather)
        for (int b = 0; b < 2; ++b) {
            SphereBuf<R>& s = S[b];
            HIPC(dalloc(&s.pos, N)); HIPC(dalloc(&s.cum, N)); HIPC(dalloc(&s.loc, N)); HIPC(dalloc(&s.kin, N));
            HIPC(dalloc(&s.cid, N)); HIPC(dalloc(&s.fid, N)); HIPC(dalloc(&s.wflag, N));
            HIPC(dalloc(&s.wh, (size_t)std::max(nwh, 1) * N * HSTRIDE));
            if (b == 0) { HIPC(dalloc(&s.sft, 2 * (size_t)N)); HIPC(dalloc(&s.slot, N)); HIPC(dalloc(&s.nthr, N)); }
            else { s.sft = S[0].sft; s.slot = S[0].slot; s.nthr = S[0].nthr; }
        }
        HIPC(dalloc(&C.pq, 2 * (size_t)Nc)); HIPC(dalloc(&C.vw, 2 * (size_t)Nc)); HIPC(dalloc(&C.prop, Nc));
        HIPC(dalloc(&C.ft, 2 * (size_t)Nc)); HIPC(dalloc(&C.snap, 2 * (size_t)Nc));
        C.dist = nullptr;
        C.type = nullptr;
        if (driver) { HIPC(dalloc(&C.dist, (size_t)Nc)); HIPC(dalloc(&C.type, (size_t)Nc)); }
        // second set for the clump renumbering at rebuilds
        HIPC(dalloc(&C2.pq, 2 * (size_t)Nc)); HIPC(dalloc(&C2.vw, 2 * (size_t)Nc)); HIPC(dalloc(&C2.prop, Nc));
        HIPC(dalloc(&C2.ft, 2 * (size_t)Nc)); HIPC(dalloc(&C2.snap, 2 * (size_t)Nc));
        C2.dist = nullptr;
        C2.type = nullptr;
        if (driver) { HIPC(dalloc(&C2.dist, (size_t)Nc)); HIPC(dalloc(&C2.type, (size_t)Nc)); }
        HIPC(dalloc(&cfile, Nc)); HIPC(dalloc(&cfile2, Nc)); HIPC(dalloc(&cfirst, Nc)); HIPC(dalloc(&cnew, Nc));
        // lists
        for (int b = 0; b < 2; ++b) {
            HIPC(dalloc(&Lf[b].foff, (size_t)N + 1));
            Lf[b].cap = 0; Lf[b].fword = nullptr; Lf[b].hist[0] = Lf[b].hist[1] = nullptr;
        }
        HIPC(dalloc(&boff, (size_t)N + 1));
        bcap = 0; bent = nullptr;
        HIPC(dalloc(&nfoff, (size_t)N + 1)); HIPC(dalloc(&nboff, (size_t)N + 1));
        HIPC(dalloc(&nslack, (size_t)N));
        HIPC(grow_front(0, (long long)N * pair_cap_per));
        HIPC(grow_front(1, (long long)N * pair_cap_per));
        HIPC(grow_back((long long)N * pair_cap_per));
        // sort / grid scratch
        const int nk = std::max(N, Nc);
        HIPC(dalloc(&key0, nk)); HIPC(dalloc(&key1, nk)); HIPC(dalloc(&val0, nk)); HIPC(dalloc(&val1, nk));
        HIPC(dalloc(&hist, radix_tmp_ints(nk, 8))); HIPC(dalloc(&hscan, radix_tmp_ints(nk, 8)));
        HIPC(dalloc(&stmp, scan_tmp_ints(std::max(radix_tmp_ints(nk, 8), N + 1))));
        HIPC(dalloc(&stotal, 4));
        HIPC(dalloc(&cellkey, N)); HIPC(dalloc(&perm, N)); HIPC(dalloc(&iperm, N));
        HIPC(dalloc(&cbeg, P.CellNum)); HIPC(dalloc(&cend, P.CellNum));
        HIPC(dalloc(&fcnt, (size_t)N + 1)); HIPC(dalloc(&bcnt, (size_t)N + 1));
        if (dp.L > 1) {
            HIPC(dalloc(&xcnt, (size_t)N + 1)); HIPC(dalloc(&xoff, (size_t)N + 1));
            HIPC(dalloc(&xbeg, (size_t)N)); HIPC(dalloc(&xend, (size_t)N));
        }
        HIPC(dalloc(&mbeg, Nc)); HIPC(dalloc(&mend, Nc)); HIPC(dalloc(&mem, N));
        HIPC(dalloc(&dstatus, 4));
        HIPC(dalloc(&ellmax, 1));
        ell_k = O.single_pass < 0 ? 0 : std::min(P.BranchLmtLen / 2, 32);
        HIPC(dalloc(&counters, 2));  // [1]: touching wall contacts of the last step ([0] unused)
        nwss = 4 * (((size_t)N + 255) / 256);  // one per wave of the contact launch's sphere blocks
        HIPC(dalloc(&wss, nwss));
        dp.wss = wss;
        HIPC(dalloc(&pst, 1));
        nbk = contacts_blocks(N, block);
        HIPC(dalloc(&pene_partial, 6 * (((size_t)N + 63) / 64)));  // one per wave of k_walls
        if (nfaces > 0 || dp.pene_on) {
            HIPC(dalloc(&wb.list, N)); HIPC(dalloc(&wb.pos_of, N)); HIPC(dalloc(&wb.res, 5 * (size_t)N));
            HIPC(dalloc(&wl_off, (size_t)N + 1));
        }
        HIPC(dalloc(&cache_f, 3 * (size_t)oi)); HIPC(dalloc(&cache_t, 3 * (size_t)oi));
        HIPC(dalloc(&ke_partial, (Nc + 255) / 256));
        HIPC(hipHostMalloc((void**)&hstatus, 4 * sizeof(uint32_t), hipHostMallocDefault));
        HIPC(hipHostMalloc((void**)&hring, 4 * sizeof(uint32_t) * (kRing + 1), hipHostMallocDefault));
        HIPC(dalloc(&dsbase, 1));
        for (int k = 0; k < kRing; ++k) HIPC(hipEventCreateWithFlags(&ev_stat[k], hipEventDisableTiming));
        std::memset(&rec, 0, sizeof(rec));
        return 0;
    }

    hipError_t grow_front(int b, long long cap) {
        FrontList& L = Lf[b];
        if (cap <= L.cap) return hipSuccess;
        dfree(L.fword);
        dfree(L.hist[0]);
        dfree(L.hist[1]);
        L.fword = nullptr; L.hist[0] = L.hist[1] = nullptr; L.cap = 0;
        hipError_t e = dalloc(&L.fword, (size_t)cap);
        if (e != hipSuccess) return e;
        e = dalloc(&L.hist[0], (size_t)cap * HSTRIDE);
        if (e != hipSuccess) return e;
        e = dalloc(&L.hist[1], (size_t)cap * HSTRIDE);
        if (e != hipSuccess) return e;
        L.cap = cap;
        return hipSuccess;
    }
    hipError_t grow_back(long long cap) {
        if (cap <= bcap) return hipSuccess;
        dfree(bent);
        bent = nullptr; bcap = 0;
        hipError_t e = dalloc(&bent, (size_t)cap);
        if (e != hipSuccess) return e;
        bcap = cap;
        return hipSuccess;
    }

    // cross-level entries of the hierarchical grid (key = owner j, value = i)
    hipError_t grow_cross(long long cap) {
        if (cap <= xcap) return hipSuccess;
        dfree(xk0); dfree(xk1); dfree(xv0); dfree(xv1); dfree(xhist); dfree(xhscan); dfree(xstmp);
        xk0 = xk1 = nullptr; xv0 = xv1 = xhist = xhscan = xstmp = nullptr;
        xcap = 0;
        const int n = (int)cap;
        hipError_t e;
        if ((e = dalloc(&xk0, (size_t)n)) != hipSuccess) return e;
        if ((e = dalloc(&xk1, (size_t)n)) != hipSuccess) return e;
        if ((e = dalloc(&xv0, (size_t)n)) != hipSuccess) return e;
        if ((e = dalloc(&xv1, (size_t)n)) != hipSuccess) return e;
        if ((e = dalloc(&xhist, (size_t)radix_tmp_ints(n, 8))) != hipSuccess) return e;
        if ((e = dalloc(&xhscan, (size_t)radix_tmp_ints(n, 8))) != hipSuccess) return e;
        if ((e = dalloc(&xstmp, (size_t)scan_tmp_ints(radix_tmp_ints(n, 8)))) != hipSuccess) return e;
        xcap = cap;
        return hipSuccess;
    }

    // entry-major rows of the single-pass search for capacity K; K shrinks to
    // what fits in a fixed budget of device memory (0: multi-pass search)
    int size_ell(int K) {
        const int tables = dp.L > 1 ? 3 : 2;
        const long long budget = 16ll << 30;
        const long long kmax = budget / ((long long)N * 4 * tables);
        if (K > kmax) K = (int)kmax;
        if (K < 8) return 0;
        const long long need = (long long)K * N;
        if (need > ell_cap) {
            dfree(ell_f); dfree(ell_b); dfree(ell_x);
            ell_f = nullptr; ell_b = ell_x = nullptr; ell_cap = 0;
            const long long cap = need + need / 4;
            if (dalloc(&ell_f, (size_t)cap) != hipSuccess || dalloc(&ell_b, (size_t)cap) != hipSuccess ||
                (tables == 3 && dalloc(&ell_x, (size_t)cap) != hipSuccess)) {
                (void)hipGetLastError();
                dfree(ell_f); dfree(ell_b); dfree(ell_x);
                ell_f = nullptr; ell_b = ell_x = nullptr;
                return 0;
            }
            ell_cap = cap;
        }
        return K;
    }

    // near pairs are a subset of the front list: every near array fits in the
    // front-list capacity (one entry per near pair on each side)
    hipError_t grow_near(long long fcap) {
        if (fcap <= ncap) return hipSuccess;
        dfree(nfw); dfree(nfs); dfree(nbo); dfree(nbe); dfree(slot2near);
        nfw = nullptr; nfs = nbo = nbe = slot2near = nullptr;
        ncap = 0;
        hipError_t e;
        if ((e = dalloc(&nfw, (size_t)fcap)) != hipSuccess) return e;
        if ((e = dalloc(&nfs, (size_t)fcap)) != hipSuccess) return e;
        if ((e = dalloc(&nbo, (size_t)fcap)) != hipSuccess) return e;
        if ((e = dalloc(&nbe, (size_t)fcap)) != hipSuccess) return e;
        if ((e = dalloc(&slot2near, (size_t)fcap)) != hipSuccess) return e;
        ncap = fcap;
        return hipSuccess;
    }

    void release() {
        if (device < 0) return;
        (void)hipSetDevice(device);
        if (st) (void)hipStreamSynchronize(st);
        auto F = [](void* p) { dfree(p); };
        for (int b = 0; b < 2; ++b) {
            SphereBuf<R>& s = S[b];
            F(s.pos); F(s.cum); F(s.loc); F(s.kin); F(s.cid); F(s.fid); F(s.wflag); F(s.wh);
            if (b == 0) { F(s.sft); F(s.slot); F(s.nthr); }
            F(Lf[b].foff); F(Lf[b].fword); F(Lf[b].hist[0]); F(Lf[b].hist[1]);
        }
        F(C.pq); F(C.vw); F(C.prop); F(C.ft); F(C.snap); F(C.dist); F(C.type);
        F(C2.pq); F(C2.vw); F(C2.prop); F(C2.ft); F(C2.snap); F(C2.dist); F(C2.type);
        F(cfile); F(cfile2); F(cfirst); F(cnew);
        F(boff); F(bent); F(nfoff); F(nboff); F(nfw); F(nfs); F(nbo); F(nbe); F(slot2near); F(hn[0]); F(hn[1]);
        F(key0); F(key1); F(val0); F(val1); F(hist); F(hscan); F(stmp); F(stotal);
        F(cellkey); F(perm); F(iperm); F(cbeg); F(cend); F(fcnt); F(bcnt); F(mbeg); F(mend); F(mem);
        F(dstatus); F(counters); F(wss); F(pst); F(pene_partial); F(wb.list); F(wb.pos_of); F(wb.res); F(wl_off); F(cache_f); F(cache_t); F(ke_partial); F(dtri);
        F(rec.tot); F(rec.parts); F(rec.pene); F(rec.pair_fcn); F(rec.pair_cpt);
        F(xcnt); F(xoff); F(xbeg); F(xend); F(xk0); F(xk1); F(xv0); F(xv1); F(xhist); F(xhscan); F(xstmp);
        F(ell_f); F(ell_b); F(ell_x); F(ellmax);
        F(dbeds); F(dbid); F(nslack);
        if (hstatus) (void)hipHostFree(hstatus);
        if (hring) (void)hipHostFree(hring);
        for (int k = 0; k < 2; ++k)
            if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
        F(dsbase);
        for (int k = 0; k < kRing; ++k)
            if (ev_stat[k]) (void)hipEventDestroy(ev_stat[k]);
        for (auto& ev : ev_pool) (void)hipEventDestroy(ev);
        ev_pool.clear();
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
        device = -1;
    }

    static WallGeo wall_geo(const nbc_wall& w) {
        WallGeo g;
        for (int a = 0; a < 3; ++a) {
            g.side[a] = w.side[a]; g.dir[a] = w.dir[a]; g.pos[a] = w.pos[a];
            g.vel[a] = w.vel[a]; g.angv[a] = w.angv[a];
        }
        g.rad = w.radius; g.height = w.height;
        return g;
    }

    // ------------------------------------------------------------------
    // Batch of beds: one context whose sphere and clump arrays are the beds
    // concatenated in bed order. Every bed keeps its own uniform grid (a range
    // of the shared cell arrays), box and wall; sorting by cell keeps each
    // bed's spheres in its own index range. A rebuild flagged by any bed
    // rebuilds every bed.
    int init_batch(const nbc_params* beds, int nb, const nbc_options& o, int dev) override {
        if (!beds || nb < 1) return fail(NBC_E_BADINPUT, "nbc_create_batch: no beds");
        if (o.driver != NBC_DRIVER_CLUMP) return fail(NBC_E_BADINPUT, "nbc_create_batch: beds run the clump loop");
        nbc_params all = beds[0];
        long long ns = 0, nc = 0, cells = 0;
        soff.assign(1, 0);
        coff.assign(1, 0);
        for (int b = 0; b < nb; ++b) {
            const nbc_params& q = beds[b];
            if (!q.finalized) return fail(NBC_E_BADINPUT, "nbc_create_batch: call nbc_finalize_params on every bed");
            if (q.Levels > 1 || q.isIntruder)
                return fail(NBC_E_BADINPUT, "nbc_create_batch: beds use the uniform grid and no penetrator");
            if (q.ParticleNum <= 0 || q.ClumpNum <= 0 || q.CellNum <= 0)
                return fail(NBC_E_BADINPUT, "nbc_create_batch: empty bed");
            const bool same = q.StepSize == all.StepSize && q.Ymod == all.Ymod && q.Prat == all.Prat &&
                              q.Frcoe == all.Frcoe && q.NormRes == all.NormRes && q.Beta == all.Beta &&
                              q.Cohesion == all.Cohesion && q.Gravity == all.Gravity && q.Thresh == all.Thresh &&
                              q.BranchLmtLen == all.BranchLmtLen && q.isWallForce == all.isWallForce &&
                              q.Damping == all.Damping && q.Update == all.Update &&
                              q.OutputInterval == all.OutputInterval;
            if (!same)
                return fail(NBC_E_BADINPUT, "nbc_create_batch: beds must share material, StepSize, Thresh, "
                                            "BranchLmtLen, Update, OutputInterval and isWallForce");
            ns += q.ParticleNum;
            nc += q.ClumpNum;
            cells += q.CellNum;
            if (ns > (1ll << 30) || nc > (1ll << 30) || cells > 2000000000ll)
                return fail(NBC_E_BADINPUT, "nbc_create_batch: batch too large");
            soff.push_back((int)ns);
            coff.push_back((int)nc);
            for (int a = 0; a < 3; ++a) {
                all.BoxLowLmt[a] = std::min(all.BoxLowLmt[a], q.BoxLowLmt[a]);
                all.BoxUpLmt[a] = std::max(all.BoxUpLmt[a], q.BoxUpLmt[a]);
            }
        }
        all.ParticleNum = (int32_t)ns;
        all.ClumpNum = (int32_t)nc;
        all.CellNum = (int32_t)cells;
        int e = init(all, o, dev);
        if (e) return e;
        nbeds = nb;
        hbeds.assign(nb, BedGeo{});
        int off = 0;
        for (int b = 0; b < nb; ++b) {
            const nbc_params& q = beds[b];
            BedGeo& g = hbeds[b];
            for (int a = 0; a < 3; ++a) {
                g.origin[a] = q.WorldOrigin[a]; g.cell[a] = q.CellSize[a]; g.icell[a] = q.i_CellSize[a];
                g.boxlo[a] = q.BoxLowLmt[a]; g.boxup[a] = q.BoxUpLmt[a]; g.G[a] = q.GridDim[a];
            }
            g.cell_off = off;
            g.cell_num = q.CellNum;
            off += q.CellNum;
        }
        HIPC(dalloc(&dbeds, (size_t)nb));
        HIPC(dalloc(&dbid, (size_t)N));
        std::vector<int> bid(N);
        for (int b = 0; b < nb; ++b)
            for (int i = soff[b]; i < soff[b + 1]; ++i) bid[i] = b;
        HIPC(hipMemcpy(dbid, bid.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(dbeds, hbeds.data(), sizeof(BedGeo) * nb, hipMemcpyHostToDevice));
        dp.nbeds = nb;
        dp.beds = dbeds;
        dp.bid = dbid;
        return 0;
    }

    int upload_batch(const nbc_spheres_in* s, const nbc_clumps_in* c, const nbc_wall* w) override {
        if (nbeds < 1 || soff.empty()) return fail(NBC_E_BADINPUT, "nbc_upload_batch: context is not a batch");
        if (!s || !c || (nfaces > 0 && !w)) return fail(NBC_E_BADINPUT, "nbc_upload_batch: null argument");
        std::vector<int32_t> cid(N);
        std::vector<double> pos(3 * (size_t)N), loc(3 * (size_t)N), mass(N), rad(N);
        std::vector<double> cp(3 * (size_t)Nc), cv(3 * (size_t)Nc), cq(4 * (size_t)Nc), cw(3 * (size_t)Nc),
            pr(4 * (size_t)Nc);
        for (int b = 0; b < nbeds; ++b) {
            const int n = soff[b + 1] - soff[b], m = coff[b + 1] - coff[b];
            if (s[b].n != n || c[b].n != m)
                return fail(NBC_E_BADINPUT, "nbc_upload_batch: sphere/clump counts differ from the bed parameters");
            for (int k = 0; k < n; ++k) {
                const int i = soff[b] + k;
                const int q = s[b].clump_id[k];
                cid[i] = (q >= 0 && q < m) ? q + coff[b] : -1;  // -1 is rejected by upload
                for (int a = 0; a < 3; ++a) {
                    pos[3 * (size_t)i + a] = s[b].pos[3 * (size_t)k + a];
                    loc[3 * (size_t)i + a] = s[b].local[3 * (size_t)k + a];
                }
                mass[i] = s[b].mass[k];
                rad[i] = s[b].radius[k];
            }
            for (int k = 0; k < m; ++k) {
                const size_t j = (size_t)coff[b] + k;
                for (int a = 0; a < 3; ++a) {
                    cp[3 * j + a] = c[b].pos[3 * (size_t)k + a];
                    cv[3 * j + a] = c[b].vel[3 * (size_t)k + a];
                    cw[3 * j + a] = c[b].angv[3 * (size_t)k + a];
                }
                for (int a = 0; a < 4; ++a) {
                    cq[4 * j + a] = c[b].quat[4 * (size_t)k + a];
                    pr[4 * j + a] = c[b].prop[4 * (size_t)k + a];
                }
            }
            if (nfaces > 0) {
                bool okw = true;
                for (int a = 0; a < 3; ++a)
                    okw = okw && std::isfinite(w[b].pos[a]) && std::isfinite(w[b].vel[a]) && std::isfinite(w[b].angv[a]);
                if (wall_type == NBC_WALL_CUBE)
                    for (int a = 0; a < 3; ++a) okw = okw && std::isfinite(w[b].side[a]) && w[b].side[a] > 0.0;
                else
                    okw = okw && std::isfinite(w[b].radius) && std::isfinite(w[b].height) && w[b].radius > 0.0 &&
                          w[b].height > 0.0;
                if (!okw)
                    return fail(NBC_E_NONFINITE, "FATAL: Invalid input or nonfinite wall force/torque at sorted "
                                                 "particle index 0. Simulation stopped.");
                hbeds[b].W = wall_geo(w[b]);
            }
        }
        HIPC(hipMemcpy(dbeds, hbeds.data(), sizeof(BedGeo) * nbeds, hipMemcpyHostToDevice));
        nbc_spheres_in sa;
        std::memset(&sa, 0, sizeof(sa));
        sa.n = N; sa.clump_id = cid.data(); sa.pos = pos.data(); sa.local = loc.data(); sa.mass = mass.data();
        sa.radius = rad.data();
        nbc_clumps_in ca;
        std::memset(&ca, 0, sizeof(ca));
        ca.n = Nc; ca.pos = cp.data(); ca.vel = cv.data(); ca.quat = cq.data(); ca.angv = cw.data(); ca.prop = pr.data();
        batch_upload = true;
        const int e = upload(&sa, &ca, nfaces > 0 ? &w[0] : nullptr, nullptr);
        batch_upload = false;
        return e;
    }

    int get_batch_ranges(int32_t* so, int32_t* co) override {
        for (size_t b = 0; b < soff.size(); ++b) {
            if (so) so[b] = soff[b];
            if (co) co[b] = coff[b];
        }
        return 0;
    }

    // ------------------------------------------------------------------
    int upload(const nbc_spheres_in* s, const nbc_clumps_in* c, const nbc_wall* w, const nbc_penetrator* pe) override {
        HIPC(hipSetDevice(device));
        if (nbeds > 1 && !batch_upload) return fail(NBC_E_BADINPUT, "nbc_upload: a batch context takes nbc_upload_batch");
        if (driver && !particle_upload)
            return fail(NBC_E_BADINPUT, "nbc_upload: a DEMmain driver context takes nbc_upload_particles");
        if (!s || !c || s->n != N || c->n != Nc)
            return fail(NBC_E_BADINPUT, "nbc_upload: sphere/clump counts differ from the context parameters");
        for (int i = 0; i < N; ++i)
            if (s->clump_id[i] < 0 || s->clump_id[i] >= Nc) {
                char b[160];
                std::snprintf(b, sizeof b,
                              "FATAL: Invalid input or nonfinite wall force/torque at sorted particle index %d. "
                              "Simulation stopped.", i);
                return fail(NBC_E_NONFINITE, b);
            }
        {
            // Invalid sphere or clump state is refused here, before anything is
            // copied to the device. The reference finds it one step later in the
            // cube-wall kernel (src/dem_force_cube_wall_clump.cu:191-203) and
            // reports the sorted index there; the first bad index in file order
            // is what the host can name.
            auto fin = [](const double* v, int k) {
                for (int a = 0; a < k; ++a)
                    if (!std::isfinite(v[a])) return false;
                return true;
            };
            for (int i = 0; i < N; ++i) {
                const int q = s->clump_id[i];
                const char* what = nullptr;
                if (!fin(s->pos + 3 * (size_t)i, 3)) what = "sphere position";
                else if (!fin(s->local + 3 * (size_t)i, 3)) what = "sphere body-frame offset";
                else if (!(std::isfinite(s->radius[i]) && s->radius[i] > 0.0)) what = "sphere radius";
                else if (!(std::isfinite(s->mass[i]) && s->mass[i] > 0.0)) what = "sphere mass";
                else if (sizeof(R) == sizeof(float) &&
                         !(std::isfinite((float)s->local[3 * (size_t)i]) && std::isfinite((float)s->local[3 * (size_t)i + 1]) &&
                           std::isfinite((float)s->local[3 * (size_t)i + 2]) && (float)s->mass[i] > 0.0f))
                    what = "sphere offset or mass outside the FP32 range";
                else if (!fin(c->pos + 3 * (size_t)q, 3)) what = "clump position";
                else if (!fin(c->vel + 3 * (size_t)q, 3)) what = "clump velocity";
                else if (!fin(c->angv + 3 * (size_t)q, 3)) what = "clump angular velocity";
                else if (!fin(c->quat + 4 * (size_t)q, 4)) what = "clump quaternion";
                else {
                    const double* pr = c->prop + 4 * (size_t)q;
                    if (!(fin(pr, 4) && pr[0] > 0.0 && pr[1] > 0.0 && pr[2] > 0.0 && pr[3] > 0.0))
                        what = "clump inertia or mass";
                }
                if (what) {
                    char b[200];
                    std::snprintf(b, sizeof b,
                                  "FATAL: Invalid input (%s) at particle index %d in file order. Simulation "
                                  "stopped.", what, i);
                    return fail(NBC_E_NONFINITE, b);
                }
            }
        }
        {
            // every clump holds a sphere (clumps are renumbered by their first sphere)
            std::vector<char> has(Nc, 0);
            for (int i = 0; i < N; ++i) has[s->clump_id[i]] = 1;
            for (int k = 0; k < Nc; ++k)
                if (!has[k]) {
                    char b[120];
                    std::snprintf(b, sizeof b, "nbc_upload: clump %d has no sphere", k);
                    return fail(NBC_E_BADINPUT, b);
                }
        }
        if (nfaces > 0) {
            if (!w) return fail(NBC_E_BADINPUT, "nbc_upload: isWallForce is set but no wall was given");
            bool okw = true;
            for (int a = 0; a < 3; ++a)
                okw = okw && std::isfinite(w->pos[a]) && std::isfinite(w->vel[a]) && std::isfinite(w->angv[a]);
            if (wall_type == NBC_WALL_CUBE)
                for (int a = 0; a < 3; ++a) okw = okw && std::isfinite(w->side[a]) && w->side[a] > 0.0;
            else
                okw = okw && std::isfinite(w->radius) && std::isfinite(w->height) && w->radius > 0.0 && w->height > 0.0;
            if (!okw)
                return fail(NBC_E_NONFINITE, "FATAL: Invalid input or nonfinite wall force/torque at sorted particle "
                                             "index 0. Simulation stopped.");
            dp.W = wall_geo(*w);
        }
        std::vector<D4> pos(N), cum(N);
        std::vector<R4> loc(N), kin(N);
        std::vector<int> fid(N);
        for (int i = 0; i < N; ++i) {
            pos[i].x = s->pos[3 * i]; pos[i].y = s->pos[3 * i + 1]; pos[i].z = s->pos[3 * i + 2]; pos[i].w = s->radius[i];
            cum[i].x = cum[i].y = cum[i].z = cum[i].w = 0.0;
            if (ref_disp<R>(driver)) cum[i] = pos[i];
            loc[i].x = (R)s->local[3 * i]; loc[i].y = (R)s->local[3 * i + 1]; loc[i].z = (R)s->local[3 * i + 2];
            loc[i].w = (R)s->mass[i];
            kin[i].x = kin[i].y = kin[i].z = R(0); kin[i].w = (R)s->mass[i];
            fid[i] = i;
        }
        SphereBuf<R>& s0 = S[scur = 0];
        HIPC(hipMemcpy(s0.pos, pos.data(), sizeof(D4) * N, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s0.cum, cum.data(), sizeof(D4) * N, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s0.loc, loc.data(), sizeof(R4) * N, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s0.kin, kin.data(), sizeof(R4) * N, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s0.cid, s->clump_id, sizeof(int) * N, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(s0.fid, fid.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        HIPC(hipMemset(s0.wflag, 0, sizeof(uint32_t) * N));
        HIPC(hipMemset(s0.wh, 0, sizeof(R) * (size_t)std::max(nwh, 1) * N * HSTRIDE));
        HIPC(hipMemset(S[1].wflag, 0, sizeof(uint32_t) * N));
        HIPC(hipMemset(s0.sft, 0, sizeof(R4) * 2 * (size_t)N));
        // near-list request thresholds: +inf until the first list build, so the
        // first step's k_sphere_pose (before any rebuild) reads defined values
        KC(launch_fill_u32(st, reinterpret_cast<uint32_t*>(s0.nthr), N, 0x7f800000u));
        std::vector<D4> pq(2 * (size_t)Nc), vw(2 * (size_t)Nc), prop(Nc), ft(2 * (size_t)Nc);
        for (int i = 0; i < Nc; ++i) {
            pq[2 * i] = D4{c->pos[3 * i], c->pos[3 * i + 1], c->pos[3 * i + 2], 0.0};
            pq[2 * i + 1] = D4{c->quat[4 * i], c->quat[4 * i + 1], c->quat[4 * i + 2], c->quat[4 * i + 3]};
            vw[2 * i] = D4{c->vel[3 * i], c->vel[3 * i + 1], c->vel[3 * i + 2], 0.0};
            vw[2 * i + 1] = D4{c->angv[3 * i], c->angv[3 * i + 1], c->angv[3 * i + 2], 0.0};
            prop[i] = D4{c->prop[4 * i], c->prop[4 * i + 1], c->prop[4 * i + 2], c->prop[4 * i + 3]};
            ft[2 * i] = D4{0, 0, 0, 0};
            ft[2 * i + 1] = D4{0, 0, 0, 0};
        }
        HIPC(hipMemcpy(C.pq, pq.data(), sizeof(D4) * pq.size(), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(C.vw, vw.data(), sizeof(D4) * vw.size(), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(C.prop, prop.data(), sizeof(D4) * prop.size(), hipMemcpyHostToDevice));
        HIPC(hipMemcpy(C.ft, ft.data(), sizeof(D4) * ft.size(), hipMemcpyHostToDevice));
        HIPC(hipMemset(C.snap, 0, sizeof(R4) * 2 * (size_t)Nc));
        {
            std::vector<int> iota(Nc);
            for (int k = 0; k < Nc; ++k) iota[k] = k;
            HIPC(hipMemcpy(cfile, iota.data(), sizeof(int) * (size_t)Nc, hipMemcpyHostToDevice));
        }
        if (driver) {
            if (O.level_from_type) {
                // the particle's type is its grid level (src/dem_hgrid.cu:213); the
                // 27-cell search of its level is complete when 3 r fits the cell
                for (int k = 0; k < Nc; ++k) {
                    const int t = htype[k];
                    if (t < 0 || t >= P.Levels) return fail(NBC_E_BADINPUT, "nbc_upload_particles: type is not a grid level");
                    if (!(3.0 * s->radius[k] <= P.LevelCell[t] * (1.0 + 1e-12))) {
                        char b[160];
                        std::snprintf(b, sizeof b, "nbc_upload_particles: particle %d is too large for grid level %d", k, t);
                        return fail(NBC_E_BADINPUT, b);
                    }
                }
                dp.clevel = C.type;
            } else {
                dp.clevel = nullptr;
            }
            HIPC(hipMemset(C.dist, 0, sizeof(D4) * (size_t)Nc));
            HIPC(hipMemcpy(C.type, htype.data(), sizeof(int) * (size_t)Nc, hipMemcpyHostToDevice));
        }
        // empty lists (src/dem_grid.cu:228-247: Nsum = 0, Lpb_pre = -1)
        lcur = 0;
        HIPC(hipMemset(Lf[0].foff, 0, sizeof(int) * ((size_t)N + 1)));
        HIPC(hipMemset(Lf[1].foff, 0, sizeof(int) * ((size_t)N + 1)));
        HIPC(hipMemset(boff, 0, sizeof(int) * ((size_t)N + 1)));
        HIPC(hipMemset(nfoff, 0, sizeof(int) * ((size_t)N + 1)));
        near_total = 0;
        wl_len = 0;
        HIPC(hipMemset(nboff, 0, sizeof(int) * ((size_t)N + 1)));
        nfront = nback = 0;
        HIPC(hipMemset(perm, 0, sizeof(int) * N));
        HIPC(hipMemset(cellkey, 0, sizeof(uint32_t) * N));
        KC(launch_fill_i32(st, cbeg, P.CellNum, -1));
        KC(launch_fill_i32(st, cend, P.CellNum, -1));
        KC(launch_fill_i32(st, mbeg, Nc, -1));
        KC(launch_fill_i32(st, mend, Nc, -1));
        {
            const int init[4] = {HALT_NONE, HALT_NONE, 0, 0};
            HIPC(hipMemcpy(dstatus, init, sizeof(init), hipMemcpyHostToDevice));
        }
        HIPC(hipMemset(counters, 0, 2 * sizeof(unsigned long long)));
        KC(launch_fill_u32(st, wss, (long long)nwss, 0u));
        // penetrator
        PeneState ps;
        std::memset(&ps, 0, sizeof(ps));
        if (dp.pene_on) {
            if (!pe) return fail(NBC_E_BADINPUT, "nbc_upload: isIntruder is set but no penetrator was given");
            for (int a = 0; a < 3; ++a) {
                ps.pos[a] = pe->pos[a]; ps.vel[a] = pe->vel[a]; ps.angv[a] = pe->angv[a]; dp.pdir[a] = pe->dir[a];
                ps.lpos[a] = pe->pos[a];
            }
            for (int a = 0; a < 4; ++a) ps.quat[a] = pe->quat[a];
            ps.contact_state = pe->contact_state;
            dp.pshape = pe->shape; dp.pmotion = pe->motion; dp.pwait = pe->wait_for_contact;
            dp.prad = pe->radius; dp.pheight = pe->height; dp.pmass = pe->prop[3];
            dp.ntri = 0; dp.ptri = nullptr;
            if (pe->shape == NBC_PENE_CONE) {
                if (!pe->tri || pe->ntri <= 0) return fail(NBC_E_BADINPUT, "nbc_upload: cone penetrator without a mesh");
                dfree(dtri);
                HIPC(dalloc(&dtri, 9 * (size_t)pe->ntri));
                HIPC(hipMemcpy(dtri, pe->tri, sizeof(double) * 9 * pe->ntri, hipMemcpyHostToDevice));
                dp.ntri = pe->ntri;
                dp.ptri = dtri;
                double b2 = 0.0;
                for (int k = 0; k < 3 * pe->ntri; ++k) {
                    const double* v = pe->tri + 3 * k;
                    b2 = std::max(b2, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                }
                dp.pbound = std::sqrt(b2) * (1.0 + 1e-12);
            }
        }
        HIPC(hipMemcpy(pst, &ps, sizeof(ps), hipMemcpyHostToDevice));
        HIPC(hipMemset(cache_f, 0, sizeof(double) * 3 * oi));
        HIPC(hipMemset(cache_t, 0, sizeof(double) * 3 * oi));
        init_update = P.Update != 0.0;
        first = true;
        k1_next = false;
        gen++;
        nstep = 0;
        nrebuild = 0;
        nnear = 0;
        nreplay = 0;
        ell_fills = 0;
        count = 0;
        uploaded = true;
        HIPC(hipDeviceSynchronize());
        return 0;
    }

    // ------------------------------------------------------------------
    // Neighbour-list rebuild for step parity p.
    int rebuild(int p) {
        gen++;
        SphereBuf<R>& so = S[scur];
        SphereBuf<R>& sn = S[scur ^ 1];
        const double t0 = prof ? 1.0 : 0.0;
        (void)t0;
        hipEvent_t e0 = prof_begin();
        // 1. hash + stable sort by cell (src/dem_grid.cu:155-305)
        KC(launch_hash(st, dp, so.pos, so.cid, key0, val0));
        int second = 0;
        KC(launch_radix_sort(st, key0, val0, key1, val1, N, bits_for(P.CellNum - 1), rbits, hist, hscan, stmp,
                             stotal, &second));
        const uint32_t* skey = second ? key1 : key0;
        const int* sval = second ? val1 : val0;
        HIPC(hipMemcpyAsync(cellkey, skey, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, st));
        HIPC(hipMemcpyAsync(perm, sval, sizeof(int) * N, hipMemcpyDeviceToDevice, st));
        // 2. cell bounds (src/dem_grid.cu:105-153)
        KC(launch_seg_bounds(st, cellkey, N, P.CellNum, cbeg, cend));
        // 3. gather every per-sphere array (src/dem_simulation_clump.cu:126-172)
        KC(launch_gather<R>(st, dp, nwh, perm, so, sn));
        scur ^= 1;
        SphereBuf<R>& s = S[scur];
        // 3b. clumps renumbered by their first sphere (no counterpart in the
        //     reference): neighbouring spheres' clump data in neighbouring lines
        {
            KC(launch_fill_i32(st, cfirst, Nc, N));
            KC(launch_clump_renumber_first(st, N, s.cid, cfirst, block));
            KC(launch_keys_from_int(st, cfirst, key0, val0, Nc));
            int sec = 0;
            KC(launch_radix_sort(st, key0, val0, key1, val1, Nc, bits_for(N), rbits, hist, hscan, stmp, stotal, &sec));
            KC(launch_clump_gather<R>(st, Nc, sec ? val1 : val0, C, C2, cfile, cfile2, cnew, s.cid, N, block));
            std::swap(C, C2);
            std::swap(cfile, cfile2);
            if (dp.clevel) dp.clevel = C.type;
        }
        // 4. neighbour counts, overflow check (src/dem_grid.cu:466-552, 709-719)
        FrontList& Lo = Lf[lcur];
        FrontList& Ln = Lf[lcur ^ 1];
        KC(launch_fill_i32(st, fcnt + N, 1, 0));
        KC(launch_fill_i32(st, bcnt + N, 1, 0));
        KC(launch_fill_u32(st, dstatus + 2, 1, 0u));
        KC(launch_fill_u32(st, ellmax, 1, 0u));
        const uint32_t* xkey = nullptr;
        const int* xsrt = nullptr;
        // single pass into entry-major rows when they fit (bit 2 of the
        // overflow word: a row was longer than K, fill with the search passes)
        const int K = size_ell(ell_k);
        const int* sidx = nullptr;
        uint32_t ovf_ell = 0;
        if (dp.L > 1) {
            // hierarchical grid: counts, cross entries (j, i) sorted by j, then
            // the front count of each owner
            KC(launch_fill_i32(st, xcnt + N, 1, 0));
            if (K > 0) {
                if (dp.nested) {
                    // threads in level order: a wave's lanes scan the same levels
                    KC(launch_level_keys(st, dp, s.pos, s.cid, key0, val0, block));
                    int sec = 0;
                    KC(launch_radix_sort(st, key0, val0, key1, val1, N, bits_for(dp.L - 1), rbits, hist, hscan, stmp,
                                         stotal, &sec));
                    sidx = sec ? val1 : val0;
                }
                KC(launch_hneighbors_ell(st, dp, s.pos, cellkey, cbeg, cend, s.cid, sidx, K, fcnt, bcnt, xcnt, ell_f,
                                         ell_b, ell_x, dstatus + 2, ellmax, block));
            } else {
                KC(launch_hneighbors(st, dp, 0, s.pos, cellkey, cbeg, cend, s.cid, fcnt, bcnt, xcnt, nullptr, nullptr,
                                     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                     dstatus + 2, block));
            }
            KC(launch_exclusive_scan(st, xcnt, xoff, N + 1, stmp, stotal + 3));
            int X = 0;
            HIPC(hipMemcpyAsync(&X, stotal + 3, sizeof(int), hipMemcpyDeviceToHost, st));
            HIPC(hipMemcpyAsync(&ovf_ell, dstatus + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            if (X < 0) return fail(NBC_E_NEIGHBOR_OVERFLOW, "cross-level pair list exceeds 2^31 entries");
            HIPC(grow_cross(std::max<long long>(X, 1)));
            if (K > 0 && !(ovf_ell & 2u))
                KC(launch_cross_from_ell(st, N, sidx, xcnt, xoff, ell_x, xk0, xv0, block));
            else
                KC(launch_hneighbors(st, dp, 1, s.pos, cellkey, cbeg, cend, s.cid, fcnt, bcnt, xcnt, xoff, xk0, xv0,
                                     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, dstatus + 2,
                                     block));
            int sec = 0;
            if (X > 0)
                KC(launch_radix_sort(st, xk0, xv0, xk1, xv1, X, bits_for(N - 1), rbits, xhist, xhscan, xstmp,
                                     stotal + 3, &sec));
            xkey = sec ? xk1 : xk0;
            xsrt = sec ? xv1 : xv0;
            KC(launch_seg_bounds(st, xkey, X, N, xbeg, xend));
            KC(launch_hfront_total(st, N, P.BranchLmtLen / 2, fcnt, bcnt, xbeg, xend, dstatus + 2, block));
        } else if (K > 0) {
            KC(launch_neighbors_ell(st, dp, s.pos, cellkey, cbeg, cend, s.cid, K, fcnt, bcnt, ell_f, ell_b,
                                    dstatus + 2, ellmax, block));
        } else {
            KC(launch_neighbors(st, dp, false, s.pos, cellkey, cbeg, cend, s.cid, fcnt, bcnt, nullptr, nullptr,
                                nullptr, nullptr, dstatus + 2, block));
        }
        KC(launch_exclusive_scan(st, fcnt, Ln.foff, N + 1, stmp, stotal));
        KC(launch_exclusive_scan(st, bcnt, boff, N + 1, stmp, stotal + 1));
        int tot[2];
        uint32_t ovf = 0;
        HIPC(hipMemcpyAsync(tot, stotal, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&ovf, dstatus + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        unsigned longest = 0;
        HIPC(hipMemcpyAsync(&longest, ellmax, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        // next rebuild's capacity: the longest row with a quarter to spare
        if (O.single_pass >= 0)
            ell_k = std::min(P.BranchLmtLen / 2, std::max(8, (int)((longest + longest / 4 + 8) & ~7u)));
        if (ovf & 1u) {
            char b[320];
            std::snprintf(b, sizeof b,
                          "FATAL: Neighbor list overflow (BranchLmtLen=%d). Simulation stopped before using the "
                          "incomplete list. Review neighbor capacity and GPU memory before rerunning.",
                          P.BranchLmtLen);
            return fail(NBC_E_NEIGHBOR_OVERFLOW, b);
        }
        if (tot[0] < 0 || tot[1] < 0) return fail(NBC_E_NEIGHBOR_OVERFLOW, "neighbour list exceeds 2^31 entries");
        if ((long long)tot[0] > Ln.cap) {
            const long long nc = (long long)(tot[0] * 1.25) + 1024;
            HIPC(grow_front(lcur ^ 1, nc));
            if (rec.pair_fcn) { dfree(rec.pair_fcn); dfree(rec.pair_cpt); rec.pair_fcn = rec.pair_cpt = nullptr; rec_cap = 0; }
        }
        if ((long long)tot[1] > bcap) HIPC(grow_back((long long)(tot[1] * 1.25) + 1024));
        // 5. fill rows, resolve back slots, remap histories (src/dem_grid.cu:561-647)
        const bool from_ell = K > 0 && !(ovf & 2u);
        if (dp.L > 1 && from_ell)
            KC(launch_hrows_from_ell(st, N, sidx, s.cid, Ln.foff, bcnt, boff, xbeg, xend, xsrt, ell_f, ell_b, Ln.fword,
                                     bent, block));
        else if (dp.L > 1)
            KC(launch_hneighbors(st, dp, 2, s.pos, cellkey, cbeg, cend, s.cid, fcnt, bcnt, xcnt, xoff, nullptr,
                                 nullptr, Ln.foff, boff, xbeg, xend, xsrt, Ln.fword, bent, dstatus + 2, block));
        else if (from_ell)
            KC(launch_rows_from_ell(st, N, fcnt, bcnt, Ln.foff, boff, ell_f, ell_b, Ln.fword, bent, block));
        else
            KC(launch_neighbors(st, dp, true, s.pos, cellkey, cbeg, cend, s.cid, fcnt, bcnt, Ln.foff, boff, Ln.fword,
                                bent, dstatus + 2, block));
        ell_fills += from_ell ? 1 : 0;
        KC(launch_back_resolve(st, N, Ln.foff, Ln.fword, boff, bent, block));
        if (driver == NBC_DRIVER_CYLINDER || driver == NBC_DRIVER_SPLIT) {
            // history-valid bits of the old near list back into the old pair words
            KC(launch_near_sync<R>(st, N, nfoff, nfw, nfs, Lo.fword, hn[p], Lo.hist[p], p, block));
            KC(launch_hist_carry<R>(st, tot[0], nfront, p, Ln.fword, Ln.hist[p], Lo.fword, Lo.hist[p], block));
        } else {
            // valid pairs of the old near list hand their histories to the new list
            KC(launch_hist_push<R>(st, dp, s.pos, s.cid, p, perm, iperm, nfoff, nfw, hn[p], Ln.foff, Ln.fword,
                                   Ln.hist[p], block));
        }
        lcur ^= 1;
        // near lists of the new list (the contact kernel's candidates)
        HIPC(grow_near(Ln.cap));
        KC(launch_near_build(st, dp, s.pos, s.cum, ref_disp<R>(driver) ? s.nthr : nullptr, nslack, s.cid, C.pq, C.vw, Ln.foff, Ln.fword,
                             boff, bent, fcnt, nfoff, nfw, nfs, slot2near, bcnt, nboff, nbo, nbe, stmp, stotal + 2,
                             block));
        HIPC(hipMemcpyAsync(&near_total, nfoff + N, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        { int e = near_hist_in(Ln.hist[p], p); if (e) return e; }
        nfront = tot[0];
        nback = tot[1];
        // 6. clump member lists in ascending sorted order (replaces the per-step
        //    thrust::sort_by_key of src/dem_clumpsphere.cu:123-150)
        KC(launch_keys_from_int(st, s.cid, key0, val0, N));
        KC(launch_radix_sort(st, key0, val0, key1, val1, N, bits_for(Nc - 1), rbits, hist, hscan, stmp, stotal,
                             &second));
        const uint32_t* ck = second ? key1 : key0;
        const int* cv = second ? val1 : val0;
        HIPC(hipMemcpyAsync(mem, cv, sizeof(int) * N, hipMemcpyDeviceToDevice, st));
        KC(launch_seg_bounds(st, ck, N, Nc, mbeg, mend));
        KC(launch_member_slot(st, N, mem, s.slot));
        // 7. wall list of the new order
        { int e = build_wall_list(); if (e) return e; }
        nrebuild++;
        prof_end(3, e0, -1);
        return 0;
    }

    int ensure_record() {
        gen++;
        if (!rec.tot) {
            HIPC(dalloc(&rec.tot, 2 * (size_t)N));
            HIPC(dalloc(&rec.parts, 6 * (size_t)N));
            HIPC(dalloc(&rec.pene, 4 * (size_t)N));
        }
        if (!rec.pair_fcn || rec_cap < Lf[lcur].cap) {
            if (rec.pair_fcn) { dfree(rec.pair_fcn); dfree(rec.pair_cpt); }
            rec_cap = std::max(Lf[0].cap, Lf[1].cap);
            HIPC(dalloc(&rec.pair_fcn, (size_t)rec_cap));
            HIPC(dalloc(&rec.pair_cpt, (size_t)rec_cap));
        }
        return 0;
    }

    static constexpr int HALT_NONE = 0x7fffffff;

    int check_status() {
        const int bad = (int)hstatus[1];
        if (bad != HALT_NONE) {
            char b[200];
            std::snprintf(b, sizeof b,
                          "FATAL: Invalid input or nonfinite wall force/torque at sorted particle index %d. "
                          "Simulation stopped.", bad);
            return fail(NBC_E_NONFINITE, b);
        }
        return 0;
    }

    int read_status() {
        HIPC(hipMemcpyAsync(hstatus, dstatus, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        return 0;
    }

    // K1 + K2 of step n (src/dem_simulation_clump.cu:113-119)
    int enqueue_head(long long n) {
        hipEvent_t e0 = prof_begin();
        if (!k1_next) KC(launch_clump_before<R>(st, dp, C, pst, counters, block, (int*)dstatus, karg(n)));
        k1_next = false;
        KC(launch_sphere_pose<R>(st, dp, S[scur], C, (int*)dstatus, karg(n), block));
        prof_end(0, e0, n);
        return 0;
    }

    // K3, penetrator sum, K4 of step n (src/dem_simulation_clump.cu:178-225)
    // With `next`, k_clump_after also runs K1 of step n + 1 and the next
    // enqueue_head skips it.
    int enqueue_tail(long long n, bool record, bool next) {
        const int p = (int)(n & 1);
        if (record) {
            int e = ensure_record();
            if (e) return e;
            KC(launch_fill_f64(st, (double*)rec.pene, 16 * (long long)N, 0.0));
            KC(launch_fill_f64(st, (double*)rec.pair_fcn, 4 * (long long)rec_cap, 0.0));
            KC(launch_fill_f64(st, (double*)rec.pair_cpt, 4 * (long long)rec_cap, 0.0));
        }
        ListBuf<R> L;
        L.nfoff = nfoff; L.nfw = nfw; L.nfs = nfs; L.nboff = nboff; L.nbo = nbo; L.nbe = nbe;
        L.hist[0] = hn[0]; L.hist[1] = hn[1];
        const int* halt = (const int*)dstatus;
        int* bad = (int*)(dstatus + 1);
        hipEvent_t e0 = prof_begin();
        KC(launch_contacts<R>(st, dp, S[scur], C, L, p, wb, wl_len, pst, pene_partial, counters, bad, rec, record, 1,
                              halt, karg(n)));
        prof_end(1, e0, n);
        if (dp.pene_on) {
            e0 = prof_begin();
            KC(launch_pene_finalize(st, dp, pst, pene_partial, (wl_len + 63) / 64, (int)(n % oi), cache_f, cache_t,
                                    halt, karg(n)));
            prof_end(4, e0, n);
        }
        e0 = prof_begin();
        KC(launch_clump_after<R>(st, dp, S[scur], C, wb, rec, record, mbeg, mend, mem, block, (int*)dstatus, karg(n),
                                 next, pst, counters));
        prof_end(2, e0, n);
        k1_next = next;
        return 0;
    }

    int clear_halt() {
        KC(launch_clear_halt(st, (int*)dstatus, HALT_NONE));
        return 0;
    }

    // Near lists of the current Verlet list at the current positions (a
    // sphere left its near-list allowance between rebuilds).
    int near_refresh(int q) {
        gen++;
        hipEvent_t e0 = prof_begin();
        FrontList& L = Lf[lcur];
        KC(launch_near_sync<R>(st, N, nfoff, nfw, nfs, L.fword, hn[q], L.hist[q], q, block));
        SphereBuf<R>& s = S[scur];
        KC(launch_near_build(st, dp, s.pos, s.cum, ref_disp<R>(driver) ? s.nthr : nullptr, nslack, s.cid, C.pq, C.vw, L.foff, L.fword,
                             boff, bent, fcnt, nfoff, nfw, nfs, slot2near, bcnt, nboff, nbo, nbe, stmp, stotal + 2,
                             block));
        HIPC(hipMemcpyAsync(&near_total, nfoff + N, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        { int e = near_hist_in(L.hist[q], q); if (e) return e; }
        nnear++;
        prof_end(6, e0, -1);
        return 0;
    }

    // The contact kernel keeps the pair histories in near order (dense: the
    // touching pairs of a wave sit in its contiguous near segment); after a
    // near build the latest buffer q moves in from front-slot order, and
    // k_near_sync moves it back before the lists change or are read.
    int near_hist_in(const R* hs_q, int q) {
        if ((long long)near_total > hn_cap) {
            dfree(hn[0]); dfree(hn[1]);
            hn[0] = hn[1] = nullptr;
            hn_cap = 0;
            const long long cap = (long long)(near_total * 1.25) + 1024;
            HIPC(dalloc(&hn[0], (size_t)cap * HSTRIDE));
            HIPC(dalloc(&hn[1], (size_t)cap * HSTRIDE));
            hn_cap = cap;
        }
        KC(launch_near_hist_gather<R>(st, near_total, nfw, nfs, hs_q, hn[q], q));
        return 0;
    }

    // Wall list at the current positions and penetrator position.
    int build_wall_list() {
        gen++;
        if (!(nfaces > 0 || dp.pene_on)) return 0;
        const SphereBuf<R>& s = S[scur];
        KC(launch_wlist_build(st, dp, s.pos, s.wflag, pst, fcnt, wl_off, wb.list, wb.pos_of, stmp, stotal + 2, block));
        int len = 0;
        HIPC(hipMemcpyAsync(&len, wl_off + N, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        wl_len = len;
        return 0;
    }

    // Steps are enqueued in chunks of `batch` steps without host round trips.
    // When a step needs a rebuild (or new near or wall lists), the device halt
    // step stops every later kernel at entry; the host then rebuilds and
    // replays from that step's contact kernel, so rebuilds happen at exactly
    // the reference's steps. The status words are copied back after every
    // chunk and read while up to kDepth chunks are in flight, so a halt wastes
    // at most a few no-op steps and the GPU never waits for the host.
    struct Chunk {
        int slot;  // event
        int sw;    // status copy
    };

    // Steady chunk of m steps from step n as a graph replay (captured on first
    // use for this parity and generation): every step's K1 is fused into the
    // previous step's k_clump_after, the last one included.
    int launch_chunk_graph(long long n, long long m) {
        const int par = (int)(n & 1);
        if (!gexec[par] || ggen[par] != gen) {
            if (gexec[par]) { (void)hipGraphExecDestroy(gexec[par]); gexec[par] = nullptr; }
            DevParams saved = dp;
            dp.sbase = dsbase;
            graph_base = n;
            HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int e = 0;
            for (long long k = 0; k < m && !e; ++k) {
                if (!e) e = enqueue_head(n + k);
                if (!e) e = enqueue_tail(n + k, false, true);
            }
            hipError_t ce = e ? hipSuccess
                              : hipMemcpyAsync(hring + 4 * kRing, dstatus, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(st, &g);
            graph_base = -1;
            dp = saved;
            if (e) { if (g) (void)hipGraphDestroy(g); return e; }
            HIPC(ce);
            HIPC(ee);
            hipError_t ie = hipGraphInstantiate(&gexec[par], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPC(ie);
            ggen[par] = gen;
        }
        KC(launch_set_int(st, dsbase, (int)n));
        HIPC(hipGraphLaunch(gexec[par], st));
        k1_next = true;
        nreplay++;
        return 0;
    }

    // Rebuild, near-list refresh or wall list for the step h that stopped,
    // then the rest of step h. Requires an idle stream and fresh hstatus.
    int handle_halt(long long h, long long end, bool rec) {
        int e;
        prof_drop(h);
        k1_next = false;  // the K1 fused into the stopped K4 did not run
        const uint32_t why = hstatus[3];
        if ((why & ST_REBUILD) || !(why & (ST_WLIST | ST_NEAR))) {
            if ((e = rebuild((int)(h & 1)))) return e;
        } else {
            if ((why & ST_NEAR) && (e = near_refresh((int)(h & 1)))) return e;
            if ((why & ST_WLIST) && (e = build_wall_list())) return e;
        }
        if ((e = clear_halt())) return e;
        return enqueue_tail(h, rec && h == end - 1, h + 1 < end);
    }

    int step(int nsteps) override {
        if (!uploaded) return fail(NBC_E_BADINPUT, "nbc_step: no state uploaded");
        HIPC(hipSetDevice(device));
        long long n = nstep;
        const long long end = nstep + nsteps;
        const bool rec = record_next;
        int e;
        std::vector<Chunk> q;  // chunks in flight, oldest first
        int next_slot = 0;
        while (n < end || !q.empty()) {
            if (n < end && (first || init_update)) {
                // the first step always rebuilds (src/dem_simulation_clump.cu:124)
                if ((e = enqueue_head(n))) return e;
                if ((e = rebuild((int)(n & 1)))) return e;
                if ((e = clear_halt())) return e;
                first = init_update = false;
                if ((e = enqueue_tail(n, rec && n == end - 1, n + 1 < end))) return e;
                ++n;
                continue;
            }
            if (n < end && (int)q.size() < kDepth) {
                const long long m = std::min<long long>(batch, end - n);
                const int slot = next_slot;
                next_slot = (next_slot + 1) % kRing;
                int sw = slot;  // ring slot holding this chunk's status copy
                if (use_graphs && !prof && !dp.pene_on && k1_next && m == batch && n + m < end) {
                    if ((e = launch_chunk_graph(n, m))) return e;
                    sw = kRing;
                } else {
                    for (long long k = 0; k < m; ++k) {
                        if ((e = enqueue_head(n + k))) return e;
                        if ((e = enqueue_tail(n + k, rec && n + k == end - 1, n + k + 1 < end))) return e;
                    }
                    HIPC(hipMemcpyAsync(hring + 4 * slot, dstatus, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
                }
                HIPC(hipEventRecord(ev_stat[slot], st));
                q.push_back(Chunk{slot, sw});
                n += m;
                continue;
            }
            // the oldest chunk in flight
            const Chunk c = q.front();
            q.erase(q.begin());
            HIPC(hipEventSynchronize(ev_stat[c.slot]));
            const uint32_t* sw = hring + 4 * c.sw;
            if (sw[0] == (uint32_t)HALT_NONE && sw[1] == (uint32_t)HALT_NONE) continue;
            // a halt or a bad sphere: let the queued (no-op) chunks finish
            q.clear();
            if ((e = read_status())) return e;
            if ((e = check_status())) return e;
            const long long h = (long long)(int)hstatus[0];
            if ((e = handle_halt(h, end, rec))) return e;
            n = h + 1;
        }
        nstep = end;
        if ((e = read_status())) return e;
        prof_collect();
        if (rec && nsteps > 0) {
            record_last = true;
            record_next = false;
        } else if (nsteps > 0) {
            record_last = false;
        }
        return check_status();
    }

    int force_rebuild() override {
        if (!uploaded) return fail(NBC_E_BADINPUT, "nbc_rebuild: no state uploaded");
        HIPC(hipSetDevice(device));
        int e = rebuild((int)(nstep & 1));
        if (e) return e;
        first = false;
        init_update = false;
        if ((e = clear_halt())) return e;
        HIPC(hipStreamSynchronize(st));
        prof_collect();
        return 0;
    }

    // ------------------------------------------------------------------
    // profiling with events on the library stream
    hipEvent_t prof_begin(hipStream_t s = nullptr) {
        if (!prof) return nullptr;
        hipEvent_t e = get_event();
        (void)hipEventRecord(e, s ? s : st);
        return e;
    }
    void prof_end(int cls, hipEvent_t b, long long step = -1, hipStream_t s = nullptr) {
        if (!prof || !b) return;
        hipEvent_t e = get_event();
        (void)hipEventRecord(e, s ? s : st);
        pending.push_back({cls, step, b, e});
    }
    // forget the timings of launches that stood down after a halt at step h
    void prof_drop(long long h) {
        std::vector<Pending> keep;
        for (auto& q : pending) {
            const bool skipped = q.step >= 0 && ((q.cls == 0 && q.step > h) || (q.cls != 0 && q.cls != 3 && q.step >= h));
            if (skipped) { ev_free.push_back(q.b); ev_free.push_back(q.e); }
            else keep.push_back(q);
        }
        pending.swap(keep);
    }
    hipEvent_t get_event() {
        if (ev_free.empty()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            ev_pool.push_back(e);
            return e;
        }
        hipEvent_t e = ev_free.back();
        ev_free.pop_back();
        return e;
    }
    void prof_collect() {
        for (auto& q : pending) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, q.b, q.e);
            prof_ms[q.cls] += ms;
            prof_n[q.cls] += 1;
            ev_free.push_back(q.b);
            ev_free.push_back(q.e);
        }
        pending.clear();
    }
    int profile_enable(int on) override {
        prof = on != 0;
        for (int k = 0; k < 8; ++k) { prof_ms[k] = 0.0; prof_n[k] = 0; }
        return 0;
    }
    int profile_get(int cls, double* ms, int64_t* n) override {
        if (cls < 0 || cls >= 8) return fail(NBC_E_BADINPUT, "nbc_profile_get: bad kernel class");
        if (ms) *ms = prof_ms[cls];
        if (n) *n = prof_n[cls];
        return 0;
    }

    // ------------------------------------------------------------------
    // getters
    int set_record(int on) override {
        record_next = on != 0;
        return 0;
    }
    int sync() override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        return 0;
    }

    int get_spheres(int32_t* fid, int32_t* cid, double* pos, double* cum, double* local, double* mass,
                    double* radius) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        SphereBuf<R>& s = S[scur];
        if (fid) HIPC(hipMemcpy(fid, s.fid, sizeof(int) * N, hipMemcpyDeviceToHost));
        if (cid) {
            // device clump indices are renumbered at rebuilds: report file indices
            HIPC(hipMemcpy(cid, s.cid, sizeof(int) * N, hipMemcpyDeviceToHost));
            std::vector<int> cf(Nc);
            HIPC(hipMemcpy(cf.data(), cfile, sizeof(int) * (size_t)Nc, hipMemcpyDeviceToHost));
            for (int i = 0; i < N; ++i) cid[i] = cf[cid[i]];
        }
        if (pos || radius) {
            std::vector<D4> v(N);
            HIPC(hipMemcpy(v.data(), s.pos, sizeof(D4) * N, hipMemcpyDeviceToHost));
            for (int i = 0; i < N; ++i) {
                if (pos) { pos[3 * i] = v[i].x; pos[3 * i + 1] = v[i].y; pos[3 * i + 2] = v[i].z; }
                if (radius) radius[i] = v[i].w;
            }
        }
        if (cum) {
            std::vector<D4> v(N);
            HIPC(hipMemcpy(v.data(), s.cum, sizeof(D4) * N, hipMemcpyDeviceToHost));
            if (ref_disp<R>(driver)) {
                // v holds the positions at the last rebuild
                std::vector<D4> p(N);
                HIPC(hipMemcpy(p.data(), s.pos, sizeof(D4) * N, hipMemcpyDeviceToHost));
                for (int i = 0; i < N; ++i) {
                    v[i].x = p[i].x - v[i].x; v[i].y = p[i].y - v[i].y; v[i].z = p[i].z - v[i].z;
                }
            }
            for (int i = 0; i < N; ++i) { cum[3 * i] = v[i].x; cum[3 * i + 1] = v[i].y; cum[3 * i + 2] = v[i].z; }
        }
        if (local || mass) {
            std::vector<R4> v(N);
            HIPC(hipMemcpy(v.data(), s.loc, sizeof(R4) * N, hipMemcpyDeviceToHost));
            for (int i = 0; i < N; ++i) {
                if (local) { local[3 * i] = v[i].x; local[3 * i + 1] = v[i].y; local[3 * i + 2] = v[i].z; }
                if (mass) mass[i] = v[i].w;
            }
        }
        return 0;
    }

    int get_grid(int32_t* cell, int32_t* pm, int32_t* cb, int32_t* ce) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        if (cell) HIPC(hipMemcpy(cell, cellkey, sizeof(int) * N, hipMemcpyDeviceToHost));
        if (pm) HIPC(hipMemcpy(pm, perm, sizeof(int) * N, hipMemcpyDeviceToHost));
        if (cb) HIPC(hipMemcpy(cb, cbeg, sizeof(int) * P.CellNum, hipMemcpyDeviceToHost));
        if (ce) HIPC(hipMemcpy(ce, cend, sizeof(int) * P.CellNum, hipMemcpyDeviceToHost));
        return 0;
    }

    int get_pair_counts(int64_t* nf, int64_t* nb) override {
        if (nf) *nf = nfront;
        if (nb) *nb = nback;
        return 0;
    }

    int get_pairs(int32_t* fo, int32_t* fj, uint8_t* ff, int32_t* bo, int32_t* bi, int32_t* bs) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        const FrontList& L = Lf[lcur];
        if (fo) HIPC(hipMemcpy(fo, L.foff, sizeof(int) * ((size_t)N + 1), hipMemcpyDeviceToHost));
        if (fj || ff) {
            std::vector<uint32_t> w(std::max<long long>(nfront, 1));
            if (nfront) HIPC(hipMemcpy(w.data(), L.fword, sizeof(uint32_t) * nfront, hipMemcpyDeviceToHost));
            for (long long k = 0; k < nfront; ++k) {
                if (fj) fj[k] = (int32_t)(w[k] & JMASK);
                if (ff) ff[k] = (w[k] & INTRA) ? 1 : 0;
            }
        }
        if (bo) HIPC(hipMemcpy(bo, boff, sizeof(int) * ((size_t)N + 1), hipMemcpyDeviceToHost));
        if (bi || bs) {
            std::vector<int2> v(std::max<long long>(nback, 1));
            if (nback) HIPC(hipMemcpy(v.data(), bent, sizeof(int2) * nback, hipMemcpyDeviceToHost));
            for (long long k = 0; k < nback; ++k) {
                if (bi) bi[k] = v[k].x & ~BINTRA;
                if (bs) bs[k] = v[k].y;
            }
        }
        return 0;
    }

    int get_pair_history(double* h9) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        if (!nfront) return 0;
        const FrontList& L = Lf[lcur];
        const int q = (int)(nstep & 1);  // step nstep-1 (parity q^1) wrote buffer q
        KC(launch_near_sync<R>(st, N, nfoff, nfw, nfs, L.fword, hn[q], L.hist[q], q, block));
        HIPC(hipStreamSynchronize(st));
        std::vector<uint32_t> w(nfront);
        std::vector<R> h((size_t)nfront * HSTRIDE);
        HIPC(hipMemcpy(w.data(), L.fword, sizeof(uint32_t) * nfront, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(h.data(), L.hist[q], sizeof(R) * h.size(), hipMemcpyDeviceToHost));
        for (long long k = 0; k < nfront; ++k)
            for (int t = 0; t < HSTRIDE; ++t)
                h9[k * HSTRIDE + t] = (w[k] & VALID(q)) ? (double)h[k * HSTRIDE + t] : 0.0;
        return 0;
    }

    int get_sphere_forces(int kind, double* f, double* t) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        if (!record_last || !rec.tot)
            return fail(NBC_E_BADINPUT, "nbc_get_sphere_forces: enable nbc_set_record before the step");
        std::vector<D4> v;
        if (kind == 0) {
            v.resize(2 * (size_t)N);
            HIPC(hipMemcpy(v.data(), rec.tot, sizeof(D4) * v.size(), hipMemcpyDeviceToHost));
            for (int i = 0; i < N; ++i) {
                if (f) { f[3 * i] = v[2 * i].x; f[3 * i + 1] = v[2 * i].y; f[3 * i + 2] = v[2 * i].z; }
                if (t) { t[3 * i] = v[2 * i + 1].x; t[3 * i + 1] = v[2 * i + 1].y; t[3 * i + 2] = v[2 * i + 1].z; }
            }
        } else if (kind >= 1 && kind <= 3) {
            v.resize(6 * (size_t)N);
            HIPC(hipMemcpy(v.data(), rec.parts, sizeof(D4) * v.size(), hipMemcpyDeviceToHost));
            const int o = 2 * (kind - 1);
            for (int i = 0; i < N; ++i) {
                const D4 a = v[6 * i + o], b = v[6 * i + o + 1];
                if (f) { f[3 * i] = a.x; f[3 * i + 1] = a.y; f[3 * i + 2] = a.z; }
                if (t) { t[3 * i] = b.x; t[3 * i + 1] = b.y; t[3 * i + 2] = b.z; }
            }
        } else {
            return fail(NBC_E_BADINPUT, "nbc_get_sphere_forces: bad kind");
        }
        return 0;
    }

    // DEMmain loops: particle i is clump i with one sphere at its centre
    // (DEMmain::CopyDataFromHostToDevice, src/dem.cu:62-72)
    int upload_particles(const nbc_particles_in* q, const nbc_wall* w, const nbc_penetrator* pe) override {
        if (!driver) return fail(NBC_E_BADINPUT, "nbc_upload_particles: the context runs the clump loop "
                                                 "(set nbc_options.driver)");
        if (!q || q->n != N || !q->pos || !q->vel || !q->quat || !q->angv || !q->prop || !q->radius)
            return fail(NBC_E_BADINPUT, "nbc_upload_particles: particle count or arrays missing");
        if (!w) return fail(NBC_E_BADINPUT, "nbc_upload_particles: the DEMmain loops need their wall");
        std::vector<int> cid(N), ty(N, 0);
        std::vector<double> local(3 * (size_t)N, 0.0), mass(N);
        for (int i = 0; i < N; ++i) {
            cid[i] = i;
            mass[i] = q->prop[4 * (size_t)i + 3];
            if (q->type) {
                ty[i] = q->type[i];
                if (driver == NBC_DRIVER_SPLIT && (ty[i] < 0 || ty[i] > 2))
                    return fail(NBC_E_BADINPUT, "nbc_upload_particles: glue type not 0, 1 or 2");
            }
        }
        std::vector<int> ty_old;
        ty_old.swap(htype);
        htype = ty;  // upload() reads the types
        const nbc_spheres_in si{N, 0, cid.data(), q->pos, local.data(), mass.data(), q->radius};
        const nbc_clumps_in ci{N, 0, q->pos, q->vel, q->quat, q->angv, q->prop};
        particle_upload = true;
        const int e = upload(&si, &ci, w, pe);
        particle_upload = false;
        if (e) htype.swap(ty_old);  // a refused upload leaves the context as it was
        else hprop.assign(q->prop, q->prop + 4 * (size_t)N);
        return e;
    }

    // DEMmain::CopyDataToIo (src/dem.cu:113-127), current sorted order
    int get_particles(int32_t* sid, int32_t* fid, int32_t* type, double* pos, double* vel, double* quat,
                      double* angv, double* prop, double* radius) override {
        if (!driver) return fail(NBC_E_BADINPUT, "nbc_get_particles: the context runs the clump loop");
        if (!uploaded) return fail(NBC_E_BADINPUT, "nbc_get_particles: no state uploaded");
        std::vector<int32_t> f(N), c(N);
        std::vector<double> p3(3 * (size_t)N), r(N);
        int e = get_spheres(f.data(), c.data(), p3.data(), nullptr, nullptr, nullptr, r.data());
        if (e) return e;
        std::vector<double> cv(3 * (size_t)N), cq(4 * (size_t)N), cw(3 * (size_t)N);
        e = get_clumps(nullptr, cv.data(), cq.data(), cw.data(), nullptr, nullptr);
        if (e) return e;
        for (int i = 0; i < N; ++i) {
            const int k = c[i];
            if (sid) sid[i] = i;
            if (fid) fid[i] = f[i];
            // the cube loop does not sort Type (src/dem_simulation_cube.cu:102-108)
            if (type) type[i] = driver == NBC_DRIVER_CUBE ? htype[i] : htype[k];
            if (pos) for (int a = 0; a < 3; ++a) pos[3 * (size_t)i + a] = p3[3 * (size_t)i + a];
            if (vel) for (int a = 0; a < 3; ++a) vel[3 * (size_t)i + a] = cv[3 * (size_t)k + a];
            if (quat) for (int a = 0; a < 4; ++a) quat[4 * (size_t)i + a] = cq[4 * (size_t)k + a];
            if (angv) for (int a = 0; a < 3; ++a) angv[3 * (size_t)i + a] = cw[3 * (size_t)k + a];
            if (prop) for (int a = 0; a < 4; ++a) prop[4 * (size_t)i + a] = hprop[4 * (size_t)k + a];
            if (radius) radius[i] = r[i];
        }
        return 0;
    }

    int get_clumps(double* pos, double* vel, double* quat, double* angv, double* F, double* T) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        // file order: device clump c is file clump cf[c]
        std::vector<int> cf(Nc);
        HIPC(hipMemcpy(cf.data(), cfile, sizeof(int) * (size_t)Nc, hipMemcpyDeviceToHost));
        std::vector<D4> a(2 * (size_t)Nc), d(2 * (size_t)Nc);
        auto fetch = [&](const D4* src) -> hipError_t {
            hipError_t e = hipMemcpy(d.data(), src, sizeof(D4) * d.size(), hipMemcpyDeviceToHost);
            for (int c = 0; c < Nc; ++c) { a[2 * (size_t)cf[c]] = d[2 * (size_t)c]; a[2 * (size_t)cf[c] + 1] = d[2 * (size_t)c + 1]; }
            return e;
        };
        if (pos || quat) {
            HIPC(fetch(C.pq));
            for (int i = 0; i < Nc; ++i) {
                if (pos) { pos[3 * i] = a[2 * i].x; pos[3 * i + 1] = a[2 * i].y; pos[3 * i + 2] = a[2 * i].z; }
                if (quat) {
                    quat[4 * i] = a[2 * i + 1].x; quat[4 * i + 1] = a[2 * i + 1].y;
                    quat[4 * i + 2] = a[2 * i + 1].z; quat[4 * i + 3] = a[2 * i + 1].w;
                }
            }
        }
        if (vel || angv) {
            HIPC(fetch(C.vw));
            for (int i = 0; i < Nc; ++i) {
                if (vel) { vel[3 * i] = a[2 * i].x; vel[3 * i + 1] = a[2 * i].y; vel[3 * i + 2] = a[2 * i].z; }
                if (angv) { angv[3 * i] = a[2 * i + 1].x; angv[3 * i + 1] = a[2 * i + 1].y; angv[3 * i + 2] = a[2 * i + 1].z; }
            }
        }
        if (F || T) {
            HIPC(fetch(C.ft));
            for (int i = 0; i < Nc; ++i) {
                if (F) { F[3 * i] = a[2 * i].x; F[3 * i + 1] = a[2 * i].y; F[3 * i + 2] = a[2 * i].z; }
                if (T) { T[3 * i] = a[2 * i + 1].x; T[3 * i + 1] = a[2 * i + 1].y; T[3 * i + 2] = a[2 * i + 1].z; }
            }
        }
        return 0;
    }

    int get_penetrator(double* pos, double* vel, double* quat, double* angv, double* cf, double* ct,
                       int32_t* state) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        PeneState ps;
        HIPC(hipMemcpy(&ps, pst, sizeof(ps), hipMemcpyDeviceToHost));
        if (pos) for (int a = 0; a < 3; ++a) pos[a] = ps.pos[a];
        if (vel) for (int a = 0; a < 3; ++a) vel[a] = ps.vel[a];
        if (quat) for (int a = 0; a < 4; ++a) quat[a] = ps.quat[a];
        if (angv) for (int a = 0; a < 3; ++a) angv[a] = ps.angv[a];
        if (cf) HIPC(hipMemcpy(cf, cache_f, sizeof(double) * 3 * oi, hipMemcpyDeviceToHost));
        if (ct) HIPC(hipMemcpy(ct, cache_t, sizeof(double) * 3 * oi, hipMemcpyDeviceToHost));
        if (state) *state = ps.contact_state;
        return 0;
    }

    int get_pene_exports(double* fn, double* ft, double* tn, double* tt) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        if (!record_last || !rec.pene) return fail(NBC_E_BADINPUT, "nbc_get_pene_exports: no recorded step");
        std::vector<D4> v(4 * (size_t)N);
        HIPC(hipMemcpy(v.data(), rec.pene, sizeof(D4) * v.size(), hipMemcpyDeviceToHost));
        double* o[4] = {fn, ft, tn, tt};
        for (int k = 0; k < 4; ++k)
            if (o[k])
                for (int i = 0; i < N; ++i) {
                    o[k][3 * i] = v[4 * i + k].x; o[k][3 * i + 1] = v[4 * i + k].y; o[k][3 * i + 2] = v[4 * i + k].z;
                }
        return 0;
    }

    int get_pair_exports(double* fcn, double* cpt) override {
        HIPC(hipSetDevice(device));
        HIPC(hipStreamSynchronize(st));
        if (!record_last || !rec.pair_fcn) return fail(NBC_E_BADINPUT, "nbc_get_pair_exports: no recorded step");
        if (!nfront) return 0;
        std::vector<D4> v(nfront);
        if (fcn) {
            HIPC(hipMemcpy(v.data(), rec.pair_fcn, sizeof(D4) * nfront, hipMemcpyDeviceToHost));
            for (long long k = 0; k < nfront; ++k) { fcn[3 * k] = v[k].x; fcn[3 * k + 1] = v[k].y; fcn[3 * k + 2] = v[k].z; }
        }
        if (cpt) {
            HIPC(hipMemcpy(v.data(), rec.pair_cpt, sizeof(D4) * nfront, hipMemcpyDeviceToHost));
            for (long long k = 0; k < nfront; ++k) { cpt[3 * k] = v[k].x; cpt[3 * k + 1] = v[k].y; cpt[3 * k + 2] = v[k].z; }
        }
        return 0;
    }

    int get_stats(nbc_stats* s) override {
        HIPC(hipSetDevice(device));
        std::memset(s, 0, sizeof(*s));
        KC(launch_kinetic(st, Nc, C.vw, C.prop, ke_partial));
        std::vector<double> kp((Nc + 255) / 256);
        unsigned long long ctr[2];
        std::vector<unsigned> wc(nwss);
        PeneState ps;
        HIPC(hipMemcpyAsync(kp.data(), ke_partial, sizeof(double) * kp.size(), hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(ctr, counters, sizeof(ctr), hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(wc.data(), wss, sizeof(unsigned) * nwss, hipMemcpyDeviceToHost, st));
        HIPC(hipMemcpyAsync(&ps, pst, sizeof(ps), hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        double ke = 0.0;
        for (double v : kp) ke += v;
        s->step = nstep;
        s->rebuilds = nrebuild;
        s->front_pairs = nfront;
        s->back_entries = nback;
        unsigned long long tss = 0;
        for (unsigned v : wc) tss += v;
        s->touching_ss = (int64_t)tss;
        s->touching_wall = (int64_t)ctr[1];
        s->near_refreshes = nnear;
        s->near_pairs = near_total;
        s->graph_replays = nreplay;
        s->single_pass_rebuilds = ell_fills;
        s->kinetic_energy = ke;
        for (int a = 0; a < 3; ++a) { s->pene_force[a] = -ps.totF[a]; s->pene_torque[a] = ps.totT[a]; }
        return 0;
    }

  private:
    struct FrontList {
        int* foff = nullptr;
        uint32_t* fword = nullptr;
        R* hist[2] = {nullptr, nullptr};
        long long cap = 0;
    };
    struct Pending {
        int cls;
        long long step;
        hipEvent_t b, e;
    };
    nbc_params P;
    nbc_options O;
    DevParams dp;
    hipStream_t st = nullptr;
    int N = 0, Nc = 0, block = 256, rbits = 8, wall_type = 0, nfaces = 0, nwh = 0, oi = 1, nbk = 1, batch = 64;
    int pair_cap_per = 16;
    SphereBuf<R> S[2] = {};
    int scur = 0;
    ClumpBuf<R> C = {};
    FrontList Lf[2];
    int lcur = 0;
    int* boff = nullptr;
    int2* bent = nullptr;
    long long bcap = 0, nfront = 0, nback = 0, rec_cap = 0;
    uint32_t *key0 = nullptr, *key1 = nullptr, *cellkey = nullptr;
    // hierarchical grid: per-sphere cross counts/offsets, per-owner segment
    // bounds, cross entries and their sort scratch
    int *xcnt = nullptr, *xoff = nullptr, *xbeg = nullptr, *xend = nullptr;
    uint32_t *xk0 = nullptr, *xk1 = nullptr;
    int *xv0 = nullptr, *xv1 = nullptr, *xhist = nullptr, *xhscan = nullptr, *xstmp = nullptr;
    long long xcap = 0;
    // single-pass neighbour search: entry-major rows of capacity ell_k
    // (fronts, backs, cross entries), sized from the longest row of the
    // previous rebuild; ell_k = 0 falls back to the multi-pass search
    uint32_t* ell_f = nullptr;
    int *ell_b = nullptr, *ell_x = nullptr;
    long long ell_cap = 0;  // entries per table
    int ell_k = 0;
    unsigned* ellmax = nullptr;
    long long ell_fills = 0;
    // batch of beds: sphere / clump offsets per bed (nbeds + 1), per-bed grid
    // and wall table (host copy and device copy), bed of each sphere
    int nbeds = 1;
    bool batch_upload = false;
    // DEMmain loops
    int driver = NBC_DRIVER_CLUMP;
    bool particle_upload = false;
    std::vector<int> htype;     // glue types, file order
    // clump renumbering: second clump buffer set, file index of each device
    // clump (cfile), first sphere per clump, old -> new index
    ClumpBuf<R> C2{};
    int *cfile = nullptr, *cfile2 = nullptr, *cfirst = nullptr, *cnew = nullptr;
    std::vector<double> hprop;  // particle Property (Ix Iy Iz m), file order
    std::vector<int> soff, coff;
    std::vector<BedGeo> hbeds;
    BedGeo* dbeds = nullptr;
    int* dbid = nullptr;
    int *val0 = nullptr, *val1 = nullptr, *hist = nullptr, *hscan = nullptr, *stmp = nullptr, *stotal = nullptr;
    int *perm = nullptr, *iperm = nullptr, *cbeg = nullptr, *cend = nullptr, *fcnt = nullptr, *bcnt = nullptr;
    int *mbeg = nullptr, *mend = nullptr, *mem = nullptr;
    uint32_t* dstatus = nullptr;  // [0] halt step, [1] smallest bad sphere, [2] overflow, [3] clump state not finite
    uint32_t* hstatus = nullptr;
    // status words copied back after every chunk of steps (pinned ring) and
    // the events that mark each copy complete
    // graph replays of steady chunks: one executable graph per parity of the
    // chunk's first step, captured at generation ggen[par] of everything its
    // kernel arguments point to (gen moves at every upload, rebuild, near or
    // wall list build); the step base of a replay is written to dsbase by a
    // one-thread kernel before it, and its status copy goes to the extra
    // ring slot kRing (status words only accumulate until the host clears
    // them, so a later chunk's copy is as good as this chunk's)
    bool use_graphs = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    long long ggen[2] = {-1, -1};
    long long gen = 0;
    long long graph_base = -1;  // >= 0 while capturing: kernel step arguments are offsets from it
    int* dsbase = nullptr;
    int karg(long long n) const { return (int)(graph_base >= 0 ? n - graph_base : n); }
    static constexpr int kDepth = 3;            // chunks in flight
    static constexpr int kRing = kDepth + 1;
    uint32_t* hring = nullptr;
    hipEvent_t ev_stat[kRing] = {};
    unsigned long long* counters = nullptr;
    unsigned* wss = nullptr;
    size_t nwss = 0;
    PeneState* pst = nullptr;
    double *pene_partial = nullptr, *cache_f = nullptr, *cache_t = nullptr, *ke_partial = nullptr, *dtri = nullptr;
    WallBuf<R> wb = {nullptr, nullptr, nullptr};
    int* wl_off = nullptr;
    int wl_len = 0;  // spheres on the wall list
    int *nfoff = nullptr, *nboff = nullptr, *slot2near = nullptr;
    uint32_t* nfw = nullptr;
    int *nfs = nullptr, *nbo = nullptr, *nbe = nullptr;
    long long ncap = 0;
    R* hn[2] = {nullptr, nullptr};  // [hn_cap][HSTRIDE] pair histories in near order
    long long hn_cap = 0;
    int near_total = 0;  // near pairs (front and back entries alike)
    RecordBuf rec;
    bool uploaded = false, first = true, init_update = false, record_next = false, record_last = false, prof = false;
    bool k1_next = false;  // the last enqueued k_clump_after also ran K1 of the next step
    long long nstep = 0, nrebuild = 0, nnear = 0, nreplay = 0;
    double* nslack = nullptr;  // per-sphere near-list slack of the last near build
    int count = 0;
    double prof_ms[8] = {0};
    int64_t prof_n[8] = {0};
    std::vector<hipEvent_t> ev_pool, ev_free;
    std::vector<Pending> pending;
};

EngineBase* make_engine(int precision) {
    if (precision == NBC_MIXED) return new Engine<float>();
    return new Engine<double>();
}

}  // namespace nbc
