// Host side of the drop-in boundary: the reference's file formats and the
// run loop of DEMClumpMain::RunSimulationClump on top of the device API.
// Nothing in this file computes the timestep; it parses, writes and drives.
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nbc.h"

namespace nbc {
void set_global_error(const std::string& m);
}
using nbc::set_global_error;

static int err(int code, const std::string& m) {
    set_global_error(m);
    return code;
}

// ---------------------------------------------------------------------------
// world.par, src/dem_subsphere_io.cu:40-149: '#' starts a comment, tabs and
// spaces are removed, "name=value", 38 recognised keys counted (duplicates
// count twice), unknown key -> error, fewer/more than 38 -> error. Numbers go
// through atof and are then converted to the field type.
extern "C" int nbc_load_world_par(const char* path, nbc_params* out, char* iniconfile, size_t iniconfile_len) {
    if (!path || !out) return err(NBC_E_BADINPUT, "nbc_load_world_par: null argument");
    std::ifstream in(path);
    if (!in) return err(NBC_E_IO, std::string("Cannot open ") + path);
    nbc_params P;
    std::memset(&P, 0, sizeof(P));
    std::string line, ini = "ph.00000000.bt";
    int flag = 0;
    while (std::getline(in, line)) {
        size_t h = line.find('#');
        std::string core = line.substr(0, h);
        core.erase(std::remove(core.begin(), core.end(), '\t'), core.end());
        core.erase(std::remove(core.begin(), core.end(), ' '), core.end());
        core.erase(std::remove(core.begin(), core.end(), '\r'), core.end());
        if (core.empty()) continue;
        size_t id = core.find('=');
        if (id == std::string::npos || id == 0 || id + 1 >= core.size())
            return err(NBC_E_PARAM, "Malformed line in world.par: " + line);
        std::string name = core.substr(0, id), value = core.substr(id + 1);
        const double v = std::atof(value.c_str());
        if (name == "NumberBody") P.ParticleNum = (int32_t)v;
        else if (name == "BodyFileDir") { /* parsed, unused (src/dem_subsphere_io.cu:84) */ }
        else if (name == "IniConFile") ini = value;
        else if (name == "BranchLmtLen") P.BranchLmtLen = (int32_t)v;
        else if (name == "StepSize") P.StepSize = v;
        else if (name == "EndStep") P.EndStep = (int32_t)v;
        else if (name == "StartStep") P.StartStep = (int32_t)v;
        else if (name == "OutputInterval") P.OutputInterval = (int32_t)v;
        else if (name == "Ymod") P.Ymod = v;
        else if (name == "Prat") P.Prat = v;
        else if (name == "Frcoe") P.Frcoe = v;
        else if (name == "NormRes") P.NormRes = v;
        else if (name == "Frvel") P.Frvel = v;
        else if (name == "Beta") P.Beta = v;
        else if (name == "MU_T") P.MU_T = v;
        else if (name == "MU_R") P.MU_R = v;
        else if (name == "Cohesion") P.Cohesion = v;
        else if (name == "ClumpNum") P.ClumpNum = (int32_t)v;
        else if (name == "Thresh") P.Thresh = v;
        else if (name == "GridLevels") P.GridLevels = v;
        else if (name == "Gravity") P.Gravity = v;
        else if (name == "Update") P.Update = (double)(unsigned)v;
        else if (name == "WorldSize_x") P.WorldSize[0] = v;
        else if (name == "WorldSize_y") P.WorldSize[1] = v;
        else if (name == "WorldSize_z") P.WorldSize[2] = v;
        else if (name == "WorldOrigin_x") P.WorldOrigin[0] = v;
        else if (name == "WorldOrigin_y") P.WorldOrigin[1] = v;
        else if (name == "WorldOrigin_z") P.WorldOrigin[2] = v;
        else if (name == "BoxLowLmt_x") P.BoxLowLmt[0] = v;
        else if (name == "BoxLowLmt_y") P.BoxLowLmt[1] = v;
        else if (name == "BoxLowLmt_z") P.BoxLowLmt[2] = v;
        else if (name == "BoxUpLmt_x") P.BoxUpLmt[0] = v;
        else if (name == "BoxUpLmt_y") P.BoxUpLmt[1] = v;
        else if (name == "BoxUpLmt_z") P.BoxUpLmt[2] = v;
        else if (name == "isIntruder") P.isIntruder = v != 0.0;
        else if (name == "isSelfGravity") P.isSelfGravity = v != 0.0;
        else if (name == "isWallForce") P.isWallForce = v != 0.0;
        else if (name == "isQuaternion") P.isQuaternion = v != 0.0;
        else return err(NBC_E_PARAM, "Unrecoganized item " + name + "!");
        flag++;
    }
    if (flag != 38) {
        std::ostringstream m;
        m << flag << "\nImportant parameter setting missing!";
        return err(NBC_E_PARAM, m.str());
    }
    const double lnCOR = std::log(P.NormRes);
    P.Damping = 2.0 * std::sqrt(5.0 / 6.0) * lnCOR / std::sqrt(lnCOR * lnCOR + 3.14159265359 * 3.14159265359);
    P.finalized = 0;
    *out = P;
    if (iniconfile && iniconfile_len) {
        std::strncpy(iniconfile, ini.c_str(), iniconfile_len - 1);
        iniconfile[iniconfile_len - 1] = 0;
    }
    return 0;
}

// SetMeshSize, src/dem_subsphere_io.cu:195-241 (Damping is recomputed here
// too so that parameters built in memory are complete).
extern "C" int nbc_finalize_params(nbc_params* p, const double* radius, int32_t n, const nbc_options* opt) {
    if (!p || (!radius && n > 0)) return err(NBC_E_BADINPUT, "nbc_finalize_params: null argument");
    if (p->finalized) return err(NBC_E_BADINPUT, "nbc_finalize_params: parameters already finalized");
    const double lnCOR = std::log(p->NormRes);
    p->Damping = 2.0 * std::sqrt(5.0 / 6.0) * lnCOR / std::sqrt(lnCOR * lnCOR + 3.14159265359 * 3.14159265359);
    double mesh = 0.0;
    for (int i = 0; i < n; ++i)
        if (radius[i] > mesh) mesh = radius[i];
    mesh = 3.0 * mesh;
    if (opt && opt->cell_size > 0.0) {
        if (opt->cell_size < mesh)
            return err(NBC_E_BADINPUT, "nbc_finalize_params: cell_size below 3 x max radius misses list pairs");
        mesh = opt->cell_size;
    }
    if (!(mesh > 0.0)) return err(NBC_E_BADINPUT, "nbc_finalize_params: radii must be positive");
    for (int a = 0; a < 3; ++a) {
        p->CellSize[a] = mesh;
        p->i_CellSize[a] = 1.0 / mesh;
        p->WorldOrigin[a] = p->WorldOrigin[a] - p->CellSize[a];
        p->WorldSize[a] = p->WorldSize[a] + 2 * p->CellSize[a];
        p->GridDim[a] = (uint32_t)std::ceil(p->WorldSize[a] / p->CellSize[a]);
    }
    const double cells = (double)p->GridDim[0] * p->GridDim[1] * p->GridDim[2];
    if (cells > 2.0e9) return err(NBC_E_BADINPUT, "nbc_finalize_params: grid too large");
    p->CellNum = (int32_t)(p->GridDim[0] * p->GridDim[1] * p->GridDim[2]);
    p->Levels = 1;
    p->LevelNested = 0;
    p->LevelBlock = 0;
    const int L = opt ? opt->hgrid_levels : 0;
    if (opt && opt->hgrid_nested && L <= 1)
        return err(NBC_E_BADINPUT, "nbc_finalize_params: hgrid_nested needs hgrid_levels 2..4");
    if (L > 1 && opt->hgrid_nested) {
        // nested power-of-two levels in coarse-cell-major order: level 0 is
        // the uniform grid of 3 x max radius, level t halves it t times
        if (L > 4) return err(NBC_E_BADINPUT, "nbc_finalize_params: a nested hierarchical grid has at most 4 levels");
        if (opt->cell_size > 0.0) return err(NBC_E_BADINPUT, "nbc_finalize_params: CellSize and HGridLevels exclude each other");
        int block = 0;
        for (int t = 0; t < L; ++t) {
            p->LevelCell[t] = mesh / (double)(1 << t);
            p->LevelOff[t] = block;
            block += 1 << (3 * t);
            for (int a = 0; a < 3; ++a) p->LevelDim[t][a] = (int32_t)(p->GridDim[a] << t);
        }
        for (int t = 0; t < L; ++t) p->LevelRmin[t] = t + 1 < L ? p->LevelCell[t + 1] / 3.0 : -1.0;
        if (cells * block + 1.0 > 2.0e9) return err(NBC_E_BADINPUT, "nbc_finalize_params: hierarchical grid too large");
        for (int a = 0; a < 3; ++a)
            if ((double)p->GridDim[a] * (1 << (L - 1)) > 1.0e9)
                return err(NBC_E_BADINPUT, "nbc_finalize_params: hierarchical grid too large");
        p->Levels = L;
        p->LevelNested = 1;
        p->LevelBlock = block;
        p->CellNum = (int32_t)(cells * block) + 1;  // + the out-of-box bucket
    } else if (L > 1) {
        // hierarchical grid: geometric radius boundaries between the smallest
        // and the largest radius, level 0 the largest spheres; every level
        // shares the origin set above (margin = the coarsest cell)
        if (L > NBC_MAX_LEVELS) return err(NBC_E_BADINPUT, "nbc_finalize_params: at most 8 grid levels");
        if (opt->cell_size > 0.0) return err(NBC_E_BADINPUT, "nbc_finalize_params: CellSize and HGridLevels exclude each other");
        double rmax = 0.0, rmin = 1e300;
        for (int i = 0; i < n; ++i) {
            rmax = std::max(rmax, radius[i]);
            rmin = std::min(rmin, radius[i]);
        }
        const double q = std::pow(rmin / rmax, 1.0 / L);
        double total = 0.0;
        double upper = rmax;
        for (int t = 0; t < L; ++t) {
            p->LevelRmin[t] = t + 1 < L ? rmax * std::pow(q, t + 1) : -1.0;
            p->LevelCell[t] = 3.0 * upper;
            p->LevelOff[t] = (int32_t)total;
            double c = 1.0;
            for (int a = 0; a < 3; ++a) {
                p->LevelDim[t][a] = (int32_t)std::ceil(p->WorldSize[a] / p->LevelCell[t]);
                c *= p->LevelDim[t][a];
            }
            total += c;
            upper = p->LevelRmin[t];
        }
        if (total + 1.0 > 2.0e9) return err(NBC_E_BADINPUT, "nbc_finalize_params: hierarchical grid too large");
        p->Levels = L;
        p->CellNum = (int32_t)total + 1;  // + the out-of-box bucket
    }
    p->finalized = 1;
    return 0;
}

// nbc.par: options that are not world.par keys.
extern "C" int nbc_load_options(const char* path, nbc_options* out) {
    if (!out) return err(NBC_E_BADINPUT, "nbc_load_options: null argument");
    std::memset(out, 0, sizeof(*out));
    if (!path) return 0;
    std::ifstream in(path);
    if (!in) return 0;
    std::string line;
    while (std::getline(in, line)) {
        std::string core = line.substr(0, line.find('#'));
        core.erase(std::remove_if(core.begin(), core.end(), [](unsigned char c) { return std::isspace(c); }),
                   core.end());
        if (core.empty()) continue;
        size_t id = core.find('=');
        if (id == std::string::npos || id == 0) return err(NBC_E_PARAM, "Malformed line in nbc.par: " + line);
        std::string k = core.substr(0, id), v = core.substr(id + 1);
        std::string lv = v;
        std::transform(lv.begin(), lv.end(), lv.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (k == "Precision") {
            if (lv == "fp64" || lv == "double") out->precision = NBC_FP64;
            else if (lv == "mixed" || lv == "fp32") out->precision = NBC_MIXED;
            else return err(NBC_E_PARAM, "nbc.par: Precision must be fp64 or mixed");
        } else if (k == "WallType") {
            if (lv == "cube") out->wall_type = NBC_WALL_CUBE;
            else if (lv == "cylinder") out->wall_type = NBC_WALL_CYLINDER;
            else return err(NBC_E_PARAM, "nbc.par: WallType must be cube or cylinder");
        } else if (k == "BlockSize") out->block_size = std::atoi(v.c_str());
        else if (k == "RadixBits") out->radix_bits = std::atoi(v.c_str());
        else if (k == "PairCapacity") out->pair_capacity = std::atoi(v.c_str());
        else if (k == "PeneFrcoe") out->pene_frcoe = std::atof(v.c_str());
        else if (k == "CellSize") out->cell_size = std::atof(v.c_str());
        else if (k == "HGridLevels") out->hgrid_levels = std::atoi(v.c_str());
        else if (k == "HGridNested") out->hgrid_nested = std::atoi(v.c_str());
        else if (k == "SinglePass") out->single_pass = std::atoi(v.c_str());
        else if (k == "NearSkin") out->near_skin = std::atof(v.c_str());
        else if (k == "NearHorizon") out->near_horizon = std::atof(v.c_str());
        else if (k == "Graphs") out->graphs = std::atoi(v.c_str());
        else if (k == "BatchSteps") out->batch_steps = std::atoi(v.c_str());
        else if (k == "Driver") {
            if (lv == "clump") out->driver = NBC_DRIVER_CLUMP;
            else if (lv == "cube") out->driver = NBC_DRIVER_CUBE;
            else if (lv == "cylinder") out->driver = NBC_DRIVER_CYLINDER;
            else if (lv == "split") out->driver = NBC_DRIVER_SPLIT;
            else if (lv == "hgrid") out->driver = NBC_DRIVER_HGRID;
            else return err(NBC_E_PARAM, "nbc.par: Driver must be clump, cube, cylinder, split or hgrid");
        } else return err(NBC_E_PARAM, "nbc.par: unknown key " + k);
    }
    return 0;
}

// grid.bt, DEMSphereIo::LoadGridLevels (src/dem_io.cu:204-248): GridLevels rows
// "level cellSize nx ny nz offset n". The level table replaces the margin and
// mesh of nbc_finalize_params (the reference's hgrid loop calls LoadGridLevels
// instead of SetMeshSize); the last cell is the out-of-box bucket.
extern "C" int nbc_load_grid_levels(const char* path, nbc_params* p) {
    if (!p) return err(NBC_E_BADINPUT, "nbc_load_grid_levels: null parameters");
    std::ifstream in(path ? path : "");
    if (!in) return err(NBC_E_IO, std::string("Error! Cannot open ") + (path ? path : "(null)") + "!");
    const int L = (int)p->GridLevels;
    if (L < 2 || L > NBC_MAX_LEVELS) return err(NBC_E_BADINPUT, "nbc_load_grid_levels: GridLevels must be 2..8");
    long long total = 0;
    for (int t = 0; t < L; ++t) {
        int id = 0, nx = 0, ny = 0, nz = 0, off = 0, tmp = 0;
        double cs = 0.0;
        if (!(in >> id >> cs >> nx >> ny >> nz >> off >> tmp)) {
            std::ostringstream m;
            m << "Error! " << path << " ends or is malformed at row " << t;
            return err(NBC_E_IO, m.str());
        }
        if (!(cs > 0.0) || nx <= 0 || ny <= 0 || nz <= 0 || off != total || (t > 0 && !(cs < p->LevelCell[t - 1])))
            return err(NBC_E_BADINPUT, "nbc_load_grid_levels: levels must have positive sizes, contiguous offsets "
                                       "and shrinking cells (level 0 the coarsest)");
        p->LevelCell[t] = cs;
        p->LevelDim[t][0] = nx; p->LevelDim[t][1] = ny; p->LevelDim[t][2] = nz;
        p->LevelOff[t] = off;
        p->LevelRmin[t] = t + 1 < L ? cs / 3.0 : -1.0;
        total += (long long)nx * ny * nz;
        if (total > (1ll << 30)) return err(NBC_E_BADINPUT, "nbc_load_grid_levels: too many cells");
    }
    const double lnCOR = std::log(p->NormRes);
    p->Damping = 2.0 * std::sqrt(5.0 / 6.0) * lnCOR / std::sqrt(lnCOR * lnCOR + 3.14159265359 * 3.14159265359);
    for (int a = 0; a < 3; ++a) {
        p->CellSize[a] = p->LevelCell[0];
        p->i_CellSize[a] = 1.0 / p->LevelCell[0];
        p->GridDim[a] = (uint32_t)p->LevelDim[0][a];
    }
    p->Levels = L;
    p->LevelNested = 0;
    p->LevelBlock = 0;
    p->CellNum = (int32_t)total + 1;
    p->finalized = 1;
    return 0;
}

// ph text, src/dem_subsphere_io.cu:166-189: "id sub clumpId x y z lx ly lz mass radius".
extern "C" int nbc_load_phase(const char* path, int32_t n, int32_t* clump_id, double* pos, double* local,
                              double* mass, double* radius) {
    std::ifstream in(path ? path : "");
    if (!in) return err(NBC_E_IO, std::string("Error! Cannot open ") + (path ? path : "(null)") + "!");
    double tmp;
    for (int i = 0; i < n; ++i) {
        int c;
        in >> tmp >> tmp >> c >> pos[3 * i] >> pos[3 * i + 1] >> pos[3 * i + 2] >> local[3 * i] >> local[3 * i + 1] >>
            local[3 * i + 2] >> mass[i] >> radius[i];
        clump_id[i] = c;
        if (!in) {
            std::ostringstream m;
            m << "Error! " << path << " ends or is malformed at row " << i;
            return err(NBC_E_IO, m.str());
        }
    }
    return 0;
}

// ch text, src/dem_clumpsphere_io.cu:40-58.
extern "C" int nbc_load_clumps(const char* path, int32_t n, double* pos, double* vel, double* quat, double* angv,
                               double* prop) {
    std::ifstream in(path ? path : "");
    if (!in) return err(NBC_E_IO, std::string("Error! Cannot open ") + (path ? path : "(null)") + "!");
    double tmp;
    for (int i = 0; i < n; ++i) {
        in >> tmp >> pos[3 * i] >> pos[3 * i + 1] >> pos[3 * i + 2] >> vel[3 * i] >> vel[3 * i + 1] >> vel[3 * i + 2] >>
            quat[4 * i] >> quat[4 * i + 1] >> quat[4 * i + 2] >> quat[4 * i + 3] >> angv[3 * i] >> angv[3 * i + 1] >>
            angv[3 * i + 2] >> prop[4 * i] >> prop[4 * i + 1] >> prop[4 * i + 2] >> prop[4 * i + 3];
        if (!in) {
            std::ostringstream m;
            m << "Error! " << path << " ends or is malformed at row " << i;
            return err(NBC_E_IO, m.str());
        }
    }
    return 0;
}

// DEMmain ph text, DEMSphereIo::LoadPhase (src/dem_io.cu:156-201):
// "id sub type x y z vx vy vz qx qy qz qw wx wy wz Ix Iy Iz mass r".
extern "C" int nbc_load_particles(const char* path, int32_t n, int32_t* type, double* pos, double* vel,
                                  double* quat, double* angv, double* prop, double* radius) {
    std::ifstream in(path ? path : "");
    if (!in) return err(NBC_E_IO, std::string("Error! Cannot open ") + (path ? path : "(null)") + "!");
    if (n < 0 || !type || !pos || !vel || !quat || !angv || !prop || !radius)
        return err(NBC_E_BADINPUT, "nbc_load_particles: bad arguments");
    double tmp;
    for (int i = 0; i < n; ++i) {
        in >> tmp >> tmp >> type[i] >> pos[3 * i] >> pos[3 * i + 1] >> pos[3 * i + 2] >> vel[3 * i] >>
            vel[3 * i + 1] >> vel[3 * i + 2] >> quat[4 * i] >> quat[4 * i + 1] >> quat[4 * i + 2] >> quat[4 * i + 3] >>
            angv[3 * i] >> angv[3 * i + 1] >> angv[3 * i + 2] >> prop[4 * i] >> prop[4 * i + 1] >> prop[4 * i + 2] >>
            prop[4 * i + 3] >> radius[i];
        if (!in) {
            std::ostringstream m;
            m << "Error! " << path << " ends or is malformed at row " << i;
            return err(NBC_E_IO, m.str());
        }
    }
    return 0;
}

static bool read_n(std::istream& in, double* v, int n) {
    for (int k = 0; k < n; ++k)
        if (!(in >> v[k])) return false;
    return true;
}

// Rocks/CubeWall.bt, src/dem_cube_wall_io.cu:43-48.
extern "C" int nbc_load_cube_wall(const char* path, nbc_wall* w) {
    std::ifstream in(path ? path : "");
    if (!in || !w) return err(NBC_E_IO, std::string("Error! Cannot open ") + (path ? path : "(null)") + "!");
    std::memset(w, 0, sizeof(*w));
    w->type = NBC_WALL_CUBE;
    if (!read_n(in, w->side, 3) || !read_n(in, w->pos, 3) || !read_n(in, w->vel, 3) || !read_n(in, w->quat, 4) ||
        !read_n(in, w->angv, 3))
        return err(NBC_E_IO, std::string("Error! Malformed ") + path);
    return 0;
}

// Rocks/CylindWall.bt, src/dem_cylind_wall_io.cu:41-49.
extern "C" int nbc_load_cylinder_wall(const char* path, nbc_wall* w) {
    std::ifstream in(path ? path : "");
    if (!in || !w) return err(NBC_E_IO, std::string("Error! Cannot open ") + (path ? path : "(null)") + "!");
    std::memset(w, 0, sizeof(*w));
    w->type = NBC_WALL_CYLINDER;
    if (!(in >> w->radius >> w->height) || !read_n(in, w->dir, 3) || !read_n(in, w->prop, 4) ||
        !read_n(in, w->pos, 3) || !read_n(in, w->vel, 3) || !read_n(in, w->quat, 4) || !read_n(in, w->angv, 3))
        return err(NBC_E_IO, std::string("Error! Malformed ") + path);
    return 0;
}

// Cone of base radius R and height H in the penetrator frame: apex at
// -H * dir-axis (here -z, the driving direction), base disc at z = 0.
extern "C" int nbc_make_cone_mesh(double radius, double height, int32_t nseg, double* tri) {
    if (!(radius > 0.0) || !(height > 0.0) || nseg < 3) return -1;
    const double pi = 3.14159265358979323846;
    int t = 0;
    for (int k = 0; k < nseg; ++k) {
        const double a0 = 2.0 * pi * k / nseg, a1 = 2.0 * pi * (k + 1) / nseg;
        const double x0 = radius * std::cos(a0), y0 = radius * std::sin(a0);
        const double x1 = radius * std::cos(a1), y1 = radius * std::sin(a1);
        if (tri) {
            double* s = tri + 9 * t;  // lateral face: apex, rim k+1, rim k
            s[0] = 0; s[1] = 0; s[2] = -height; s[3] = x1; s[4] = y1; s[5] = 0; s[6] = x0; s[7] = y0; s[8] = 0;
            double* b = tri + 9 * (t + 1);  // base face
            b[0] = 0; b[1] = 0; b[2] = 0; b[3] = x0; b[4] = y0; b[5] = 0; b[6] = x1; b[7] = y1; b[8] = 0;
        }
        t += 2;
    }
    return t;
}

// Rocks/Penetrator.bt, src/dem_penetrator_io.cu:21-116; "cone R H [nseg]" is
// an extension token for the triangulated cone.
extern "C" int nbc_load_penetrator(const char* path, nbc_penetrator* p) {
    std::ifstream in(path ? path : "");
    if (!in || !p) return err(NBC_E_IO, std::string("Error! Cannot open ") + (path ? path : "(null)") + "!");
    std::memset(p, 0, sizeof(*p));
    p->motion = NBC_PENE_CONSTANT_VELOCITY;
    std::string tok;
    in >> tok;
    std::string low = tok;
    std::transform(low.begin(), low.end(), low.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    int nseg = 0;
    if (low == "sphere") {
        p->shape = NBC_PENE_SPHERE;
        in >> p->radius;
        p->height = p->radius;
        p->dir[0] = 0.0; p->dir[1] = 0.0; p->dir[2] = 1.0;
    } else if (low == "cylinder") {
        p->shape = NBC_PENE_CYLINDER;
        in >> p->radius >> p->height;
        read_n(in, p->dir, 3);
    } else if (low == "cone") {
        p->shape = NBC_PENE_CONE;
        in >> p->radius >> p->height >> nseg;
        p->dir[0] = 0.0; p->dir[1] = 0.0; p->dir[2] = 1.0;
    } else {
        char* end = nullptr;
        p->radius = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0')
            return err(NBC_E_BADINPUT, "Error! Unknown penetrator shape '" + tok + "' in " + path + ".");
        p->shape = NBC_PENE_CYLINDER;
        in >> p->height;
        read_n(in, p->dir, 3);
    }
    read_n(in, p->prop, 4);
    read_n(in, p->pos, 3);
    read_n(in, p->vel, 3);
    read_n(in, p->quat, 4);
    read_n(in, p->angv, 3);
    if (!in || !(p->radius > 0.0) || !(p->prop[3] > 0.0)) {
        std::ostringstream m;
        m << "Error! Invalid penetrator data in " << path << " (radius=" << p->radius << ", mass=" << p->prop[3]
          << ", stream_ok=" << (bool)in << ").";
        return err(NBC_E_BADINPUT, m.str());
    }
    std::string opt;
    if (in >> opt) {
        std::transform(opt.begin(), opt.end(), opt.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (opt != "motion") return err(NBC_E_BADINPUT, "Error! Unknown penetrator option '" + opt + "' in " + path + ".");
        std::string m;
        in >> m;
        std::transform(m.begin(), m.end(), m.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (m == "constant_velocity" || m == "constant" || m == "kinematic") p->motion = NBC_PENE_CONSTANT_VELOCITY;
        else if (m == "gravity" || m == "dynamic") p->motion = NBC_PENE_GRAVITY;
        else if (m == "gravity_after_contact") { p->motion = NBC_PENE_GRAVITY; p->wait_for_contact = 1; }
        else return err(NBC_E_BADINPUT, "Error! Unknown penetrator motion mode '" + m + "' in " + path + ".");
    }
    if (p->shape == NBC_PENE_CONE) {
        if (!(p->height > 0.0)) return err(NBC_E_BADINPUT, "Error! cone penetrator needs R H");
        if (nseg < 3) nseg = 64;
        const int nt = nbc_make_cone_mesh(p->radius, p->height, nseg, nullptr);
        double* t = (double*)std::malloc(sizeof(double) * 9 * nt);
        if (!t) return err(NBC_E_BADINPUT, "out of memory");
        nbc_make_cone_mesh(p->radius, p->height, nseg, t);
        p->tri = t;
        p->ntri = nt;
    }
    return 0;
}

extern "C" void nbc_free_mesh(nbc_penetrator* p) {
    if (p && p->tri) {
        std::free((void*)p->tri);
        p->tri = nullptr;
        p->ntri = 0;
    }
}

// ---------------------------------------------------------------------------
// Binary writers, src/dem_binary_io.h:19-181: native little-endian, int32 and
// FP64, packed double3/double4, no header. Files are <dir>/<prefix>.%08d.bin.
static int mkdirs(const std::string& d) {
    std::string cur;
    for (size_t i = 0; i <= d.size(); ++i) {
        if (i == d.size() || d[i] == '/') {
            if (!cur.empty() && cur != ".") {
                if (::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return -1;
            }
        }
        if (i < d.size()) cur.push_back(d[i]);
    }
    return 0;
}

static std::string bin_name(const char* dir, const char* prefix, int step) {
    std::ostringstream n;
    n << dir << '/' << prefix << '.' << std::setw(8) << std::setfill('0') << step << ".bin";
    return n.str();
}

struct BinFile {
    FILE* f = nullptr;
    bool ok = true;
    explicit BinFile(const std::string& p) { f = std::fopen(p.c_str(), "wb"); ok = f != nullptr; }
    template <class T> void w(const T* d, size_t n) {
        if (ok && n) ok = std::fwrite(d, sizeof(T), n, f) == n;
    }
    template <class T> void s(const T& v) { w(&v, 1); }
    int close() {
        if (f && std::fclose(f) != 0) ok = false;
        f = nullptr;
        return ok ? 0 : -1;
    }
    ~BinFile() { if (f) std::fclose(f); }
};

static int open_out(const char* dir, int step, BinFile*& bf, const char* prefix) {
    if (step < 0) return err(NBC_E_BADINPUT, "Negative binary output count.");
    if (mkdirs(dir) != 0) return err(NBC_E_IO, std::string("Cannot create ") + dir);
    bf = new BinFile(bin_name(dir, prefix, step));
    if (!bf->ok) {
        delete bf;
        return err(NBC_E_IO, "Cannot open binary output: " + bin_name(dir, prefix, step));
    }
    return 0;
}

// OutPart/ph.%08d.bin, src/dem_subsphere_io.cu:262-275
extern "C" int nbc_write_particles_bin(const char* dir, int32_t step, int32_t n, const int32_t* sorted_id,
                                       const int32_t* file_id, const int32_t* clump_id, const double* pos,
                                       const double* local, const double* mass, const double* radius) {
    if (n < 0) return err(NBC_E_BADINPUT, "Negative binary output count.");
    BinFile* f;
    int e = open_out(dir, step, f, "ph");
    if (e) return e;
    f->s(n);
    f->w(sorted_id, n); f->w(file_id, n); f->w(clump_id, n);
    f->w(pos, 3 * (size_t)n); f->w(local, 3 * (size_t)n); f->w(mass, n); f->w(radius, n);
    e = f->close();
    delete f;
    return e ? err(NBC_E_IO, "Cannot write binary output") : 0;
}

// OutPart/ph.%08d.bin of the DEMmain loops, DEMSphereIo::SaveParticles_Bin
// (src/dem_io.cu:290-305): count, sortedParticleID, ParticleID, Type, Pos,
// Vel, Quat, AngSpd, Property, Radius.
extern "C" int nbc_write_spheres_bin(const char* dir, int32_t step, int32_t n, const int32_t* sorted_id,
                                     const int32_t* file_id, const int32_t* type, const double* pos,
                                     const double* vel, const double* quat, const double* angv, const double* prop,
                                     const double* radius) {
    if (n < 0) return err(NBC_E_BADINPUT, "Negative binary output count.");
    BinFile* f;
    int e = open_out(dir, step, f, "ph");
    if (e) return e;
    f->s(n);
    f->w(sorted_id, n); f->w(file_id, n); f->w(type, n);
    f->w(pos, 3 * (size_t)n); f->w(vel, 3 * (size_t)n); f->w(quat, 4 * (size_t)n); f->w(angv, 3 * (size_t)n);
    f->w(prop, 4 * (size_t)n); f->w(radius, n);
    e = f->close();
    delete f;
    return e ? err(NBC_E_IO, "Cannot write binary output") : 0;
}

// OutClump/ch.%08d.bin, src/dem_binary_io.h:148-180 (penetrator first as id -1)
static int write_clumps(const char* dir, const char* prefix, int32_t step, int32_t nc, const double* pos,
                        const double* vel, const double* quat, const double* angv, const double* prop,
                        const double* ppos, const double* pvel, const double* pquat, const double* pangv) {
    const bool hp = ppos != nullptr;
    if (hp && (!pvel || !pquat || !pangv)) return err(NBC_E_BADINPUT, "Incomplete intruder output data.");
    BinFile* f;
    int e = open_out(dir, step, f, prefix);
    if (e) return e;
    const int32_t count = nc + (hp ? 1 : 0);
    std::vector<int32_t> ids(count);
    if (hp) ids[0] = -1;
    for (int i = 0; i < nc; ++i) ids[i + (hp ? 1 : 0)] = i;
    const double zero4[4] = {0, 0, 0, 0};
    f->s(count);
    f->w(ids.data(), ids.size());
    if (hp) f->w(ppos, 3);
    f->w(pos, 3 * (size_t)nc);
    if (hp) f->w(pvel, 3);
    f->w(vel, 3 * (size_t)nc);
    if (hp) f->w(pquat, 4);
    f->w(quat, 4 * (size_t)nc);
    if (hp) f->w(pangv, 3);
    f->w(angv, 3 * (size_t)nc);
    if (hp) f->w(zero4, 4);
    f->w(prop, 4 * (size_t)nc);
    e = f->close();
    delete f;
    return e ? err(NBC_E_IO, "Cannot write binary output") : 0;
}

extern "C" int nbc_write_clumps_bin(const char* dir, int32_t step, int32_t nc, const double* pos, const double* vel,
                                    const double* quat, const double* angv, const double* prop,
                                    const double* ppos, const double* pvel, const double* pquat,
                                    const double* pangv) {
    return write_clumps(dir, "ch", step, nc, pos, vel, quat, angv, prop, ppos, pvel, pquat, pangv);
}

// OutPart/intruder.%08d.bin of the DEMmain loops (src/dem_io.cu:307-317): the
// clump layout with no clumps and the intruder as id -1
extern "C" int nbc_write_intruder_bin(const char* dir, int32_t step, const double* pos, const double* vel,
                                      const double* quat, const double* angv) {
    if (!pos) return err(NBC_E_BADINPUT, "Incomplete intruder output data.");
    return write_clumps(dir, "intruder", step, 0, nullptr, nullptr, nullptr, nullptr, nullptr, pos, vel, quat, angv);
}

// OutForce/fc.%08d.bin, src/dem_binary_io.h:123-144 (every 100th entry)
extern "C" int nbc_write_force_history_bin(const char* dir, int32_t step, int32_t interval, const double* force,
                                           const double* torque) {
    if (interval < 0) return err(NBC_E_BADINPUT, "Negative binary output count.");
    const int skip = 100;
    const int32_t count = interval / skip + (interval % skip != 0 ? 1 : 0);
    if (count != 0 && (!force || !torque)) return err(NBC_E_BADINPUT, "Null force history data.");
    BinFile* f;
    int e = open_out(dir, step, f, "fc");
    if (e) return e;
    std::vector<int32_t> samples(count);
    std::vector<double> fs(3 * (size_t)count), ts(3 * (size_t)count);
    for (int i = 0; i < count; ++i) {
        samples[i] = i * skip;
        for (int a = 0; a < 3; ++a) {
            fs[3 * i + a] = force[3 * (size_t)samples[i] + a];
            ts[3 * i + a] = torque[3 * (size_t)samples[i] + a];
        }
    }
    f->s(count);
    f->w(samples.data(), samples.size());
    f->w(fs.data(), fs.size());
    f->w(ts.data(), ts.size());
    e = f->close();
    delete f;
    return e ? err(NBC_E_IO, "Cannot write binary output") : 0;
}

// OutForce/yi.%08d.bin, src/dem_subsphere_io.cu:277-304 (|Fn|^2 > 1e-28)
extern "C" int nbc_write_yi_bin(const char* dir, int32_t step, int32_t n, const int32_t* clump_id, const double* fn,
                                const double* ft, const double* tn, const double* tt) {
    std::vector<int32_t> ids;
    std::vector<double> a, b, c, d;
    for (int j = 0; j < n; ++j) {
        const double* v = fn + 3 * (size_t)j;
        if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 1e-28) {
            ids.push_back(clump_id[j]);
            for (int k = 0; k < 3; ++k) {
                a.push_back(fn[3 * (size_t)j + k]); b.push_back(ft[3 * (size_t)j + k]);
                c.push_back(tn[3 * (size_t)j + k]); d.push_back(tt[3 * (size_t)j + k]);
            }
        }
    }
    BinFile* f;
    int e = open_out(dir, step, f, "yi");
    if (e) return e;
    const int32_t m = (int32_t)ids.size();
    f->s(m);
    f->w(ids.data(), ids.size());
    f->w(a.data(), a.size()); f->w(b.data(), b.size()); f->w(c.data(), c.size()); f->w(d.data(), d.size());
    e = f->close();
    delete f;
    return e ? err(NBC_E_IO, "Cannot write binary output") : 0;
}

// OutCnt/{Nei,Fcn,Cpt}.%08d.bin, src/dem_binary_io.h:100-121
extern "C" int nbc_write_contacts_bin(const char* dir, int32_t step, int32_t n, int32_t branch, const int32_t* nei,
                                      const double* fcn, const double* cpt) {
    if (n < 0 || branch < 0) return err(NBC_E_BADINPUT, "Negative binary output count.");
    const size_t count = (size_t)n * (size_t)branch;
    const char* pre[3] = {"Nei", "Fcn", "Cpt"};
    for (int k = 0; k < 3; ++k) {
        BinFile* f;
        int e = open_out(dir, step, f, pre[k]);
        if (e) return e;
        f->s(n);
        f->s(branch);
        if (k == 0) f->w(nei, count);
        else f->w(k == 1 ? fcn : cpt, 3 * count);
        e = f->close();
        delete f;
        if (e) return err(NBC_E_IO, "Cannot write binary output");
    }
    return 0;
}
