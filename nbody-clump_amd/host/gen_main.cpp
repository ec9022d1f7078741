// nbc_generate: writes a synthetic bed as a run directory for nbody_clump
// (world.par, ph.00000000.bt, ch.00000000.bt, Rocks/, nbc.par), the role
// the reference's MATLAB generator plays for its own runs
// (球簇生成器/generate_dem_bed.m:1-219, which writes ph/ch only and refuses to
// overwrite them; so does this tool).
//
//   nbc_generate --kind 1 --clumps 200000 --seed 1 --out run_c1
//   nbc_generate --kind 1 --clumps 200000 --seed 3 --radius 0.08 --cone 0.01 0.01732 64 0.1 --out run_c3
//
// kinds: 0 C0 (3-sphere rods in a 40x40x80 mm cube), 1 C1 (8 templates of
// 2-8 spheres, cylinder), 2 C2 (1-6 spheres, radii 0.2-2 mm, cube, 3-level
// hierarchical grid), 3 single spheres (DEMmain loops).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/nbc.h"

namespace {

void usage() {
    std::fprintf(stderr,
                 "usage: nbc_generate --out DIR [--kind 0|1|2|3] [--clumps N] [--seed S] [--radius R]\n"
                 "                    [--size X Y Z] [--steps N] [--output-interval N]\n"
                 "                    [--cone R H NSEG SPEED | --sphere-penetrator R SPEED]\n"
                 "  --radius R   container radius (kinds 1, 3) or cube section edge (kind 2), metres\n"
                 "  --size X Y Z cube edges (kind 0), metres\n"
                 "  --cone       triangulated cone above the bed centre, driven down at SPEED m/s\n");
}

bool exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

}  // namespace

int main(int argc, char** argv) {
    nbc_bed_spec spec;
    std::memset(&spec, 0, sizeof(spec));
    spec.kind = 1;
    spec.seed = 1;
    std::string out;
    long steps = -1, interval = -1;
    int pene = 0;  // 1 cone, 2 sphere
    double pr = 0, ph = 0, pspeed = 0.1;
    int nseg = 64;
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        auto need = [&](int n) {
            if (k + n >= argc) {
                usage();
                std::exit(2);
            }
        };
        if (a == "--out") { need(1); out = argv[++k]; }
        else if (a == "--kind") { need(1); spec.kind = std::atoi(argv[++k]); }
        else if (a == "--clumps") { need(1); spec.nclumps = std::atoi(argv[++k]); }
        else if (a == "--seed") { need(1); spec.seed = std::strtoull(argv[++k], nullptr, 10); }
        else if (a == "--radius") { need(1); spec.container[0] = std::atof(argv[++k]); }
        else if (a == "--size") {
            need(3);
            for (int d = 0; d < 3; ++d) spec.container[d] = std::atof(argv[++k]);
        } else if (a == "--steps") { need(1); steps = std::atol(argv[++k]); }
        else if (a == "--output-interval") { need(1); interval = std::atol(argv[++k]); }
        else if (a == "--cone") {
            need(4);
            pene = 1;
            pr = std::atof(argv[++k]);
            ph = std::atof(argv[++k]);
            nseg = std::atoi(argv[++k]);
            pspeed = std::atof(argv[++k]);
        } else if (a == "--sphere-penetrator") {
            need(2);
            pene = 2;
            pr = std::atof(argv[++k]);
            pspeed = std::atof(argv[++k]);
        } else {
            usage();
            return 2;
        }
    }
    if (out.empty() || spec.kind < 0 || spec.kind > 3) {
        usage();
        return 2;
    }
    ::mkdir(out.c_str(), 0755);
    if (exists(out + "/ph.00000000.bt") || exists(out + "/ch.00000000.bt")) {
        std::fprintf(stderr, "nbc_generate: %s already holds ph/ch files; refusing to overwrite\n", out.c_str());
        return 1;
    }
    nbc_bed_out bed;
    std::memset(&bed, 0, sizeof(bed));
    int e = nbc_generate_bed(&spec, &bed);
    if (e) {
        std::fprintf(stderr, "nbc_generate: %s\n", nbc_last_error(nullptr));
        return 1;
    }
    if (steps > 0) bed.params.EndStep = (int32_t)steps;
    if (interval > 0) bed.params.OutputInterval = (int32_t)interval;
    nbc_penetrator p;
    std::memset(&p, 0, sizeof(p));
    double* tri = nullptr;
    if (pene) {
        // on the container axis, its lowest point 20 um above the highest sphere
        double top = -1e300;
        for (int i = 0; i < bed.n; ++i) {
            const double x = bed.pos[3 * i], y = bed.pos[3 * i + 1], z = bed.pos[3 * i + 2];
            if (std::hypot(x - bed.wall.pos[0], y - bed.wall.pos[1]) < pr + bed.radius[i])
                top = std::fmax(top, z + bed.radius[i]);
        }
        if (top < -1e299) top = bed.wall.pos[2];
        p.motion = NBC_PENE_CONSTANT_VELOCITY;
        p.dir[2] = 1.0;
        p.prop[0] = p.prop[1] = p.prop[2] = p.prop[3] = 1.0;
        p.quat[3] = 1.0;
        p.vel[2] = -pspeed;
        p.pos[0] = bed.wall.pos[0];
        p.pos[1] = bed.wall.pos[1];
        p.radius = pr;
        if (pene == 1) {
            p.shape = NBC_PENE_CONE;
            p.height = ph;
            const int nt = nbc_make_cone_mesh(pr, ph, nseg, nullptr);
            if (nt <= 0) {
                std::fprintf(stderr, "nbc_generate: bad cone\n");
                return 1;
            }
            tri = new double[9 * (size_t)nt];
            nbc_make_cone_mesh(pr, ph, nseg, tri);
            p.ntri = nt;
            p.tri = tri;
            p.pos[2] = top + 20e-6 + ph;  // apex at z = -H in the penetrator frame
        } else {
            p.shape = NBC_PENE_SPHERE;
            p.height = pr;
            p.pos[2] = top + 20e-6 + pr;
        }
        bed.params.isIntruder = 1;
    }
    e = nbc_save_bed_text(out.c_str(), &bed, pene ? &p : nullptr);
    if (e) std::fprintf(stderr, "nbc_generate: %s\n", nbc_last_error(nullptr));
    else
        std::printf("nbc_generate: kind %d, %d clumps, %d spheres, seed %llu -> %s\n", spec.kind, bed.nc, bed.n,
                    (unsigned long long)spec.seed, out.c_str());
    delete[] tri;
    nbc_free_bed(&bed);
    return e ? 1 : 0;
}
