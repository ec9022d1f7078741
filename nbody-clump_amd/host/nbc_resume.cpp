// Binary -> text state converter for resuming a clump run. The reference's
// loaders read the text .bt formats (src/dem_subsphere_io.cu:174-175,
// src/dem_clumpsphere_io.cu:46-48) while its outputs are binary
// (src/dem_binary_io.h:19-181), so a run cannot be continued from its own
// outputs without this step; the reference ships no such converter.
//
// OutPart/ph.<step>.bin (src/dem_subsphere_io.cu:262-275, sorted order with
// each sphere's file index) and OutClump/ch.<step>.bin (src/dem_binary_io.h:
// 148-180, the penetrator first as id -1) become the IniConFile ph text (rows
// in file order) and ch.00000000.bt of a new run directory; world.par is
// copied with StartStep = step, nbc.par and Rocks/ are copied, and the
// penetrator's record replaces position, velocity, quaternion and angular
// velocity in Rocks/Penetrator.bt. Pair, wall and penetrator histories are not
// in the outputs: the resumed run starts its springs from zero.
#include <sys/stat.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nbc.h"

namespace nbc {
void set_global_error(const std::string& m);
int write_penetrator_text(const std::string& path, const nbc_penetrator* pene);
}  // namespace nbc

namespace {

int fail(int code, const std::string& m) {
    nbc::set_global_error("nbc_resume_from_bin: " + m);
    return code;
}

bool exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

std::string bin_path(const std::string& dir, const char* prefix, int step) {
    std::ostringstream n;
    n << dir << '/' << prefix << '.' << std::setw(8) << std::setfill('0') << step << ".bin";
    return n.str();
}

struct Reader {
    FILE* f = nullptr;
    bool ok = true;
    explicit Reader(const std::string& p) { f = std::fopen(p.c_str(), "rb"); ok = f != nullptr; }
    ~Reader() { if (f) std::fclose(f); }
    template <class T> void r(T* d, size_t n) {
        if (ok && n) ok = std::fread(d, sizeof(T), n, f) == n;
    }
    bool at_end() {
        if (!ok) return false;
        const int c = std::fgetc(f);
        return c == EOF;
    }
};

bool copy_file(const std::string& a, const std::string& b) {
    std::ifstream in(a, std::ios::binary);
    if (!in) return false;
    std::ofstream out(b, std::ios::binary);
    out << in.rdbuf();
    return (bool)out;
}

}  // namespace

extern "C" int nbc_resume_from_bin(const char* rundir, int32_t step, const char* outdir) {
    if (!rundir || !outdir || step < 0) return fail(NBC_E_BADINPUT, "bad arguments");
    const std::string rd(rundir), od(outdir);
    nbc_params P;
    char ini[4096];
    int e = nbc_load_world_par((rd + "/world.par").c_str(), &P, ini, sizeof ini);
    if (e) return e;
    const int N = P.ParticleNum, Nc = P.ClumpNum;
    ::mkdir(od.c_str(), 0755);
    if (exists(od + "/" + ini) || exists(od + "/ch.00000000.bt"))
        return fail(NBC_E_IO, od + " already holds ph/ch files; refusing to overwrite");

    // spheres, sorted order -> file order
    Reader ph(bin_path(rd + "/OutPart", "ph", step));
    if (!ph.ok) return fail(NBC_E_IO, "cannot open " + bin_path(rd + "/OutPart", "ph", step));
    int32_t n = -1;
    ph.r(&n, 1);
    if (!ph.ok || n != N) return fail(NBC_E_IO, "sphere count in the ph output differs from NumberBody");
    std::vector<int32_t> sid(n), fid(n), cid(n);
    std::vector<double> pos(3 * (size_t)n), loc(3 * (size_t)n), mass(n), rad(n);
    ph.r(sid.data(), n); ph.r(fid.data(), n); ph.r(cid.data(), n);
    ph.r(pos.data(), 3 * (size_t)n); ph.r(loc.data(), 3 * (size_t)n); ph.r(mass.data(), n); ph.r(rad.data(), n);
    if (!ph.ok || !ph.at_end()) return fail(NBC_E_IO, "ph output truncated or longer than its header says");
    std::vector<int> row(n, -1);
    for (int k = 0; k < n; ++k) {
        if (fid[k] < 0 || fid[k] >= n || row[fid[k]] >= 0) return fail(NBC_E_IO, "ph output file indices are not a permutation");
        if (cid[k] < 0 || cid[k] >= Nc) return fail(NBC_E_IO, "ph output clump id out of range");
        row[fid[k]] = k;
    }

    // clumps (and the penetrator record)
    Reader ch(bin_path(rd + "/OutClump", "ch", step));
    if (!ch.ok) return fail(NBC_E_IO, "cannot open " + bin_path(rd + "/OutClump", "ch", step));
    int32_t cnt = -1;
    ch.r(&cnt, 1);
    if (!ch.ok || cnt < Nc || cnt > Nc + 1) return fail(NBC_E_IO, "clump count in the ch output differs from ClumpNum");
    std::vector<int32_t> ids(cnt);
    std::vector<double> cpos(3 * (size_t)cnt), cvel(3 * (size_t)cnt), cq(4 * (size_t)cnt), cw(3 * (size_t)cnt),
        cprop(4 * (size_t)cnt);
    ch.r(ids.data(), cnt);
    ch.r(cpos.data(), cpos.size()); ch.r(cvel.data(), cvel.size()); ch.r(cq.data(), cq.size());
    ch.r(cw.data(), cw.size()); ch.r(cprop.data(), cprop.size());
    if (!ch.ok || !ch.at_end()) return fail(NBC_E_IO, "ch output truncated or longer than its header says");
    const int pk = cnt == Nc + 1 ? 0 : -1;  // the penetrator record
    if (pk == 0 && ids[0] != -1) return fail(NBC_E_IO, "ch output holds one record too many and no penetrator");
    std::vector<int> crow(Nc, -1);
    for (int k = 0; k < cnt; ++k) {
        if (k == pk) continue;
        if (ids[k] < 0 || ids[k] >= Nc || crow[ids[k]] >= 0) return fail(NBC_E_IO, "ch output clump ids are not a permutation");
        crow[ids[k]] = k;
    }

    // ph text in file order (id, index in its clump, clump, x, local, mass, radius)
    {
        FILE* f = std::fopen((od + "/" + ini).c_str(), "w");
        if (!f) return fail(NBC_E_IO, "cannot write " + od + "/" + ini);
        std::vector<int> sub(Nc, 0);
        for (int i = 0; i < n; ++i) {
            const size_t k = (size_t)row[i];
            std::fprintf(f, "%d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", i, sub[cid[k]]++, cid[k],
                         pos[3 * k], pos[3 * k + 1], pos[3 * k + 2], loc[3 * k], loc[3 * k + 1], loc[3 * k + 2],
                         mass[k], rad[k]);
        }
        if (std::fclose(f) != 0) return fail(NBC_E_IO, "cannot write " + od + "/" + ini);
    }
    {
        FILE* f = std::fopen((od + "/ch.00000000.bt").c_str(), "w");
        if (!f) return fail(NBC_E_IO, "cannot write " + od + "/ch.00000000.bt");
        for (int c = 0; c < Nc; ++c) {
            const size_t k = (size_t)crow[c];
            std::fprintf(f,
                         "%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g "
                         "%.17g %.17g %.17g\n",
                         c, cpos[3 * k], cpos[3 * k + 1], cpos[3 * k + 2], cvel[3 * k], cvel[3 * k + 1], cvel[3 * k + 2],
                         cq[4 * k], cq[4 * k + 1], cq[4 * k + 2], cq[4 * k + 3], cw[3 * k], cw[3 * k + 1],
                         cw[3 * k + 2], cprop[4 * k], cprop[4 * k + 1], cprop[4 * k + 2], cprop[4 * k + 3]);
        }
        if (std::fclose(f) != 0) return fail(NBC_E_IO, "cannot write " + od + "/ch.00000000.bt");
    }

    // world.par with StartStep = step (the rest as it was), nbc.par
    {
        std::ifstream in(rd + "/world.par");
        std::ofstream out(od + "/world.par");
        std::string line;
        bool done = false;
        while (std::getline(in, line)) {
            std::string core = line.substr(0, line.find('#'));
            const size_t eq = core.find('=');
            std::string key = eq == std::string::npos ? "" : core.substr(0, eq);
            key.erase(0, key.find_first_not_of(" \t"));
            key.erase(key.find_last_not_of(" \t") + 1);
            if (key == "StartStep") {
                out << "StartStep = " << step << "\n";
                done = true;
            } else {
                out << line << "\n";
            }
        }
        if (!done || !out) return fail(NBC_E_IO, "cannot write " + od + "/world.par");
    }
    if (exists(rd + "/nbc.par") && !copy_file(rd + "/nbc.par", od + "/nbc.par"))
        return fail(NBC_E_IO, "cannot copy nbc.par");

    // Rocks/: walls as they were, the penetrator at its recorded state
    ::mkdir((od + "/Rocks").c_str(), 0755);
    for (const char* w : {"CubeWall.bt", "CylindWall.bt"})
        if (exists(rd + "/Rocks/" + w) && !copy_file(rd + "/Rocks/" + w, od + "/Rocks/" + w))
            return fail(NBC_E_IO, std::string("cannot copy Rocks/") + w);
    if (exists(rd + "/Rocks/Penetrator.bt")) {
        if (pk == 0) {
            nbc_penetrator pe;
            e = nbc_load_penetrator((rd + "/Rocks/Penetrator.bt").c_str(), &pe);
            if (e) return e;
            for (int a = 0; a < 3; ++a) {
                pe.pos[a] = cpos[a];
                pe.vel[a] = cvel[a];
                pe.angv[a] = cw[a];
            }
            for (int a = 0; a < 4; ++a) pe.quat[a] = cq[a];
            const int w = nbc::write_penetrator_text(od + "/Rocks/Penetrator.bt", &pe);
            nbc_free_mesh(&pe);
            if (w) return fail(NBC_E_IO, "cannot write Rocks/Penetrator.bt");
        } else if (!copy_file(rd + "/Rocks/Penetrator.bt", od + "/Rocks/Penetrator.bt")) {
            return fail(NBC_E_IO, "cannot copy Rocks/Penetrator.bt");
        }
    }
    return 0;
}
