// nbc_run_simulation_clump: the run loop of DEMClumpMain::RunSimulationClump
// (src/dem_simulation_clump.cu:3-261) over the device API. Same inputs from
// the working directory (world.par, IniConFile, ch.00000000.bt, Rocks/), same
// console lines and the same OutPart/OutClump/OutForce/OutCnt files; the
// optional nbc.par selects precision, wall type and tuning.
#include <sys/stat.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/nbc.h"

namespace nbc {
void set_global_error(const std::string& m);
}

namespace {

struct Run {
    std::string dir;
    nbc_params P{};
    nbc_options O{};
    nbc_ctx* ctx = nullptr;
    nbc_wall wall{};
    nbc_penetrator pene{};
    bool have_pene = false;
    std::vector<int32_t> cid;
    std::vector<double> pos, local, mass, radius;
    std::vector<double> cpos, cvel, cquat, cangv, cprop;

    std::string path(const std::string& p) const { return dir.empty() ? p : dir + "/" + p; }
    ~Run() {
        if (ctx) nbc_destroy(ctx);
        if (have_pene) nbc_free_mesh(&pene);
    }
};

bool exists(const std::string& p) {
    struct stat s;
    return ::stat(p.c_str(), &s) == 0;
}

int save_outputs(Run& r, int step, bool intruder) {
    const int N = r.P.ParticleNum, Nc = r.P.ClumpNum;
    std::vector<int32_t> fid(N), cid(N), sid(N);
    std::vector<double> pos(3 * (size_t)N), local(3 * (size_t)N), mass(N), radius(N);
    int e = nbc_get_spheres(r.ctx, fid.data(), cid.data(), pos.data(), nullptr, local.data(), mass.data(),
                            radius.data());
    if (e) return e;
    for (int i = 0; i < N; ++i) sid[i] = i;
    e = nbc_write_particles_bin(r.path("OutPart").c_str(), step, N, sid.data(), fid.data(), cid.data(), pos.data(),
                                local.data(), mass.data(), radius.data());
    if (e) return e;
    std::vector<double> cp(3 * (size_t)Nc), cv(3 * (size_t)Nc), cq(4 * (size_t)Nc), cw(3 * (size_t)Nc);
    e = nbc_get_clumps(r.ctx, cp.data(), cv.data(), cq.data(), cw.data(), nullptr, nullptr);
    if (e) return e;
    if (intruder) {
        const int oi = r.P.OutputInterval;
        std::vector<double> cf(3 * (size_t)oi), ct(3 * (size_t)oi);
        double pp[3], pv[3], pq[4], pw[3];
        e = nbc_get_penetrator(r.ctx, pp, pv, pq, pw, cf.data(), ct.data(), nullptr);
        if (e) return e;
        e = nbc_write_force_history_bin(r.path("OutForce").c_str(), step, oi, cf.data(), ct.data());
        if (e) return e;
        std::vector<double> fn(3 * (size_t)N), ft(3 * (size_t)N), tn(3 * (size_t)N), tt(3 * (size_t)N);
        e = nbc_get_pene_exports(r.ctx, fn.data(), ft.data(), tn.data(), tt.data());
        if (e) return e;
        e = nbc_write_yi_bin(r.path("OutForce").c_str(), step, N, cid.data(), fn.data(), ft.data(), tn.data(),
                             tt.data());
        if (e) return e;
        e = nbc_write_clumps_bin(r.path("OutClump").c_str(), step, Nc, cp.data(), cv.data(), cq.data(), cw.data(),
                                 r.cprop.data(), pp, pv, pq, pw);
    } else {
        e = nbc_write_clumps_bin(r.path("OutClump").c_str(), step, Nc, cp.data(), cv.data(), cq.data(), cw.data(),
                                 r.cprop.data(), nullptr, nullptr, nullptr, nullptr);
    }
    return e;
}

// Nei/Fcn/Cpt rows of BranchLmtLen entries: front pairs first, back pairs at
// the end of the row, -1 / zero elsewhere (src/dem_force_particle.cu:238-268).
int save_contacts(Run& r, int step) {
    const int N = r.P.ParticleNum, B = r.P.BranchLmtLen;
    int64_t nf = 0, nb = 0;
    int e = nbc_get_pair_counts(r.ctx, &nf, &nb);
    if (e) return e;
    std::vector<int32_t> fo(N + 1), fj(nf > 0 ? nf : 1), bo(N + 1), bi(nb > 0 ? nb : 1), bs(nb > 0 ? nb : 1);
    e = nbc_get_pairs(r.ctx, fo.data(), fj.data(), nullptr, bo.data(), bi.data(), bs.data());
    if (e) return e;
    std::vector<double> fcn(3 * (size_t)(nf > 0 ? nf : 1)), cpt(3 * (size_t)(nf > 0 ? nf : 1));
    e = nbc_get_pair_exports(r.ctx, fcn.data(), cpt.data());
    if (e) return e;
    std::vector<int32_t> nei((size_t)N * B, -1);
    std::vector<double> cf((size_t)N * B * 3, 0.0), cc((size_t)N * B * 3, 0.0);
    for (int i = 0; i < N; ++i) {
        for (int s = fo[i], g = 0; s < fo[i + 1]; ++s, ++g) {
            const size_t o = (size_t)i * B + g;
            nei[o] = fj[s];
            for (int a = 0; a < 3; ++a) { cf[3 * o + a] = fcn[3 * (size_t)s + a]; cc[3 * o + a] = cpt[3 * (size_t)s + a]; }
        }
        const int ni = bo[i + 1] - bo[i];
        for (int m = 0; m < ni; ++m) {
            const int k = bo[i] + m;
            const size_t o = (size_t)i * B + (B - ni + m);
            nei[o] = bi[k];
            const int s = bs[k];
            if (s < 0) continue;
            for (int a = 0; a < 3; ++a) { cf[3 * o + a] = -1.0 * fcn[3 * (size_t)s + a]; cc[3 * o + a] = cpt[3 * (size_t)s + a]; }
        }
    }
    return nbc_write_contacts_bin(r.path("OutCnt").c_str(), step, N, B, nei.data(), cf.data(), cc.data());
}

}  // namespace

extern "C" int nbc_run_simulation_clump(const char* workdir, int32_t device) {
    Run r;
    r.dir = workdir ? workdir : "";
    char ini[4096];
    std::cout << "Load Parameters ..." << std::endl;
    int e = nbc_load_world_par(r.path("world.par").c_str(), &r.P, ini, sizeof ini);
    if (e) return e;
    std::cout << "========================= Finish loading parameters =========================" << std::endl;
    const int N = r.P.ParticleNum, Nc = r.P.ClumpNum;
    if (N <= 0 || Nc <= 0 || r.P.OutputInterval <= 0) {
        nbc::set_global_error("world.par: NumberBody, ClumpNum and OutputInterval must be positive");
        return NBC_E_PARAM;
    }
    std::cout << "========================= Load Phase =========================" << std::endl;
    r.cid.resize(N); r.pos.resize(3 * (size_t)N); r.local.resize(3 * (size_t)N); r.mass.resize(N); r.radius.resize(N);
    e = nbc_load_phase(r.path(ini).c_str(), N, r.cid.data(), r.pos.data(), r.local.data(), r.mass.data(),
                       r.radius.data());
    if (e) return e;
    std::cout << "========================= Finish loading particle data =========================" << std::endl;
    e = nbc_load_options(r.path("nbc.par").c_str(), &r.O);
    if (e) return e;
    e = nbc_finalize_params(&r.P, r.radius.data(), N, &r.O);
    if (e) return e;
    std::cout << "Configuration of Grid ..." << std::endl;
    std::cout << "CellSize: " << r.P.CellSize[0] << " " << r.P.CellSize[1] << " " << r.P.CellSize[2] << std::endl;
    std::cout << "WorldOrigin: " << r.P.WorldOrigin[0] << " " << r.P.WorldOrigin[1] << " " << r.P.WorldOrigin[2] << std::endl;
    std::cout << "WorldSize: " << r.P.WorldSize[0] << " " << r.P.WorldSize[1] << " " << r.P.WorldSize[2] << std::endl;
    std::cout << "CellSize: " << r.P.CellSize[0] << " " << r.P.CellSize[1] << " " << r.P.CellSize[2] << std::endl;
    std::cout << "GridDim: " << r.P.GridDim[0] << " " << r.P.GridDim[1] << " " << r.P.GridDim[2] << std::endl;
    std::cout << "BoxUpLmt: " << r.P.BoxUpLmt[0] << " " << r.P.BoxUpLmt[1] << " " << r.P.BoxUpLmt[2] << std::endl;
    std::cout << "BoxLowLmt: " << r.P.BoxLowLmt[0] << " " << r.P.BoxLowLmt[1] << " " << r.P.BoxLowLmt[2] << std::endl;
    std::cout << "========================= Load Phase =========================" << std::endl;
    r.cpos.resize(3 * (size_t)Nc); r.cvel.resize(3 * (size_t)Nc); r.cquat.resize(4 * (size_t)Nc);
    r.cangv.resize(3 * (size_t)Nc); r.cprop.resize(4 * (size_t)Nc);
    e = nbc_load_clumps(r.path("ch.00000000.bt").c_str(), Nc, r.cpos.data(), r.cvel.data(), r.cquat.data(),
                        r.cangv.data(), r.cprop.data());
    if (e) return e;
    std::cout << "========================= Finish loading particle data =========================" << std::endl;
    const int wall_type = r.O.wall_type ? r.O.wall_type : NBC_WALL_CUBE;
    if (wall_type == NBC_WALL_CYLINDER) e = nbc_load_cylinder_wall(r.path("Rocks/CylindWall.bt").c_str(), &r.wall);
    else e = nbc_load_cube_wall(r.path("Rocks/CubeWall.bt").c_str(), &r.wall);
    if (e) return e;
    e = nbc_load_penetrator(r.path("Rocks/Penetrator.bt").c_str(), &r.pene);
    if (e) return e;
    r.have_pene = true;
    const bool intruder = r.P.isIntruder != 0;
    const char* folders_i[] = {"OutPart", "OutForce", "OutClump", "OutCnt"};
    const char* folders_n[] = {"OutPart", "OutClump", "OutCnt"};
    const int nfold = intruder ? 4 : 3;
    for (int k = 0; k < nfold; ++k) {
        const std::string f = r.path(intruder ? folders_i[k] : folders_n[k]);
        if (!exists(f)) {
            std::cout << "Create results folder: ./" << (intruder ? folders_i[k] : folders_n[k]) << "/" << std::endl;
            ::mkdir(f.c_str(), 0755);
        }
    }
    if (intruder) {
        std::ofstream shape(r.path("OutClump/PenetratorShape.bt"));
        if (r.pene.shape == NBC_PENE_SPHERE) shape << "sphere " << r.pene.radius << std::endl;
        else if (r.pene.shape == NBC_PENE_CONE) shape << "cone " << r.pene.radius << " " << r.pene.height << std::endl;
        else shape << "cylinder " << r.pene.radius << " " << r.pene.height << std::endl;
    }
    e = nbc_create(&r.ctx, &r.P, &r.O, device);
    if (e) return e;
    nbc_spheres_in si{N, 0, r.cid.data(), r.pos.data(), r.local.data(), r.mass.data(), r.radius.data()};
    nbc_clumps_in ci{Nc, 0, r.cpos.data(), r.cvel.data(), r.cquat.data(), r.cangv.data(), r.cprop.data()};
    e = nbc_upload(r.ctx, &si, &ci, r.P.isWallForce ? &r.wall : nullptr, intruder ? &r.pene : nullptr);
    if (e) return e;
    const int oi = r.P.OutputInterval, nth = 2 * oi;
    int count = 0, count2 = 0;
    int ii = r.P.StartStep;
    while (ii < r.P.EndStep) {
        // advance in one call up to the next step that writes output
        const int to_out = oi - count, to_cnt = nth - count2;
        int chunk = std::min(std::min(to_out, to_cnt), r.P.EndStep - ii);
        if (chunk > 1) {
            e = nbc_step(r.ctx, chunk - 1);
            if (e) return e;
            ii += chunk - 1; count += chunk - 1; count2 += chunk - 1;
        }
        const bool out_step = count + 1 == oi, cnt_step = count2 + 1 == nth;
        if ((out_step && intruder) || cnt_step) nbc_set_record(r.ctx, 1);
        e = nbc_step(r.ctx, 1);
        if (e) return e;
        count++; count2++;
        if (count == oi) {
            std::printf("StepSize = %d\n", ii + 1);
            std::fflush(stdout);
            e = save_outputs(r, ii + 1, intruder);
            if (e) return e;
            count = 0;
        }
        if (count2 == nth) {
            e = save_contacts(r, ii + 1);
            if (e) return e;
            count2 = 0;
        }
        ++ii;
    }
    return 0;
}

// nbc_run_simulation: DEMmain's single-sphere loops (src/dem_simulation_cube.cu,
// src/dem_simulation_cylind.cu, src/dem_simulation_split.cu,
// src/dem_simulation_hgrid.cu). Same inputs (world.par, IniConFile with the
// 21-column rows of src/dem_io.cu:156-201, Rocks/CubeWall.bt or
// Rocks/CylindWall.bt) and the OutPart/ph.<step> files of
// DEMSphereIo::SaveParticles_Bin every OutputInterval steps.
extern "C" int nbc_run_simulation(const char* workdir, int32_t device, int32_t driver) {
    Run r;
    r.dir = workdir ? workdir : "";
    int e = nbc_load_options(r.path("nbc.par").c_str(), &r.O);
    if (e) return e;
    if (driver > 0) r.O.driver = driver;
    if (r.O.driver == NBC_DRIVER_CLUMP) return nbc_run_simulation_clump(workdir, device);
    if (r.O.driver < 0 || r.O.driver > NBC_DRIVER_HGRID) {
        nbc::set_global_error("nbc_run_simulation: unknown driver");
        return NBC_E_BADINPUT;
    }
    char ini[4096];
    std::cout << "Load Parameters ..." << std::endl;
    e = nbc_load_world_par(r.path("world.par").c_str(), &r.P, ini, sizeof ini);
    if (e) return e;
    std::cout << "========================= Finish loading parameters =========================" << std::endl;
    const int N = r.P.ParticleNum;
    if (N <= 0 || r.P.OutputInterval <= 0) {
        nbc::set_global_error("world.par: NumberBody and OutputInterval must be positive");
        return NBC_E_PARAM;
    }
    // one sphere per particle
    r.P.ClumpNum = N;
    std::cout << "========================= Load Phase =========================" << std::endl;
    std::vector<int32_t> type(N);
    std::vector<double> pos(3 * (size_t)N), vel(3 * (size_t)N), quat(4 * (size_t)N), angv(3 * (size_t)N),
        prop(4 * (size_t)N), radius(N);
    e = nbc_load_particles(r.path(ini).c_str(), N, type.data(), pos.data(), vel.data(), quat.data(), angv.data(),
                           prop.data(), radius.data());
    if (e) return e;
    std::cout << "========================= Finish loading particle data =========================" << std::endl;
    if (r.O.driver == NBC_DRIVER_HGRID && exists(r.path("grid.bt"))) {
        // the reference's hgrid loop: levels from grid.bt, a particle's level
        // is its type (src/dem_simulation_hgrid.cu:15, src/dem_hgrid.cu:213)
        std::cout << "========================= Load HGrid =========================" << std::endl;
        e = nbc_load_grid_levels(r.path("grid.bt").c_str(), &r.P);
        if (e) return e;
        r.O.level_from_type = 1;
        std::cout << "========================= Finish loading hgrid data =========================" << std::endl;
    } else {
        if (r.O.driver == NBC_DRIVER_HGRID && r.O.hgrid_levels <= 1)
            r.O.hgrid_levels = r.P.GridLevels > 1 ? (int)r.P.GridLevels : 3;
        if (r.O.driver != NBC_DRIVER_HGRID) r.O.hgrid_levels = 0;
        e = nbc_finalize_params(&r.P, radius.data(), N, &r.O);
        if (e) return e;
    }
    std::cout << "Configuration of Grid ..." << std::endl;
    std::cout << "CellSize: " << r.P.CellSize[0] << " " << r.P.CellSize[1] << " " << r.P.CellSize[2] << std::endl;
    std::cout << "WorldOrigin: " << r.P.WorldOrigin[0] << " " << r.P.WorldOrigin[1] << " " << r.P.WorldOrigin[2] << std::endl;
    std::cout << "WorldSize: " << r.P.WorldSize[0] << " " << r.P.WorldSize[1] << " " << r.P.WorldSize[2] << std::endl;
    std::cout << "GridDim: " << r.P.GridDim[0] << " " << r.P.GridDim[1] << " " << r.P.GridDim[2] << std::endl;
    const bool cube = r.O.driver == NBC_DRIVER_CUBE;
    if (cube) e = nbc_load_cube_wall(r.path("Rocks/CubeWall.bt").c_str(), &r.wall);
    else e = nbc_load_cylinder_wall(r.path("Rocks/CylindWall.bt").c_str(), &r.wall);
    if (e) return e;
    // the loops load the penetrator file whether or not it is used (src/dem_simulation_cube.cu:56-62)
    e = nbc_load_penetrator(r.path("Rocks/Penetrator.bt").c_str(), &r.pene);
    if (e) return e;
    r.have_pene = true;
    const bool intruder = r.P.isIntruder != 0;
    // OutPart, and OutCnt for the cube loop's contact files; OutForce (written
    // with the intruder, src/dem_binary_io.h:138) is created here as well
    std::vector<std::string> folders = {"OutPart"};
    if (intruder) {
        folders.push_back("OutForce");
        if (cube) folders.push_back("OutCnt");
    }
    for (const std::string& fo : folders)
        if (!exists(r.path(fo))) {
            std::cout << "Create results folder: ./" << fo << "/" << std::endl;
            ::mkdir(r.path(fo).c_str(), 0755);
        }
    e = nbc_create(&r.ctx, &r.P, &r.O, device);
    if (e) return e;
    const nbc_particles_in pin{N, 0, type.data(), pos.data(), vel.data(), quat.data(), angv.data(), prop.data(),
                               radius.data()};
    e = nbc_upload_particles(r.ctx, &pin, &r.wall, intruder ? &r.pene : nullptr);
    if (e) return e;
    const int oi = r.P.OutputInterval, nth = 2 * oi;
    const bool contacts = intruder && cube;  // src/dem_simulation_cube.cu:210-221
    std::vector<int32_t> sid(N), fid(N);
    int count = 0, count2 = 0;
    int ii = r.P.StartStep;
    while (ii < r.P.EndStep) {
        const int to_out = oi - count, to_cnt = contacts ? nth - count2 : oi;
        const int chunk = std::min(std::min(to_out, to_cnt), r.P.EndStep - ii);
        if (chunk > 1) {
            e = nbc_step(r.ctx, chunk - 1);
            if (e) return e;
            ii += chunk - 1; count += chunk - 1; count2 += chunk - 1;
        }
        if (contacts && count2 + 1 == nth) nbc_set_record(r.ctx, 1);
        e = nbc_step(r.ctx, 1);
        if (e) return e;
        count++; count2++;
        if (count == oi) {
            std::printf("StepSize = %d\n", ii + 1);
            std::fflush(stdout);
            const std::string out = r.path("OutPart");
            e = nbc_get_particles(r.ctx, sid.data(), fid.data(), type.data(), pos.data(), vel.data(), quat.data(),
                                  angv.data(), prop.data(), radius.data());
            if (e) return e;
            e = nbc_write_spheres_bin(out.c_str(), ii + 1, N, sid.data(), fid.data(), type.data(), pos.data(),
                                      vel.data(), quat.data(), angv.data(), prop.data(), radius.data());
            if (e) return e;
            if (intruder) {
                std::vector<double> cf(3 * (size_t)oi), ct(3 * (size_t)oi);
                double pp[3], pv[3], pq[4], pw[3];
                e = nbc_get_penetrator(r.ctx, pp, pv, pq, pw, cf.data(), ct.data(), nullptr);
                if (e) return e;
                e = nbc_write_intruder_bin(out.c_str(), ii + 1, pp, pv, pq, pw);
                if (e) return e;
                e = nbc_write_force_history_bin(r.path("OutForce").c_str(), ii + 1, oi, cf.data(), ct.data());
                if (e) return e;
            }
            count = 0;
        }
        if (contacts && count2 == nth) {
            e = save_contacts(r, ii + 1);
            if (e) return e;
            count2 = 0;
        }
        if (!contacts) count2 = 0;
        ++ii;
    }
    return 0;
}
