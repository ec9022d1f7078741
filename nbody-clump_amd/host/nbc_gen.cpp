// Deterministic synthetic clumped beds (host only).
//
// Semantics follow the reference's offline generator (球簇生成器/
// generate_dem_bed.m): lattice or deposition placement, uniform random
// orientation quaternions (x y z w, body->world = Quat2IvDCM), sub-sphere mass
// an equal share of the clump's union mass (make_clump_template_lambda.m:245),
// principal moments of the union. Union mass properties are integrated on a
// voxel grid instead of the MATLAB Monte-Carlo estimate. Sphere constants
// (local offset, mass, radius) are rounded to FP32-representable doubles so
// that the FP64 and FP32-mixed forms start from identical inputs.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nbc.h"

namespace nbc {
void set_global_error(const std::string& m);
}

namespace {

struct V { double x, y, z; };
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(double s, V a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double f32(double v) { return (double)(float)v; }

struct Template {
    std::vector<V> off;       // principal-frame offsets from the centre of mass
    std::vector<double> r;
    double mass = 0, I[3] = {0, 0, 0};
    double rbound = 0;        // max |off| + r
};

// Jacobi eigen-decomposition of a symmetric 3x3 matrix (columns of v = eigenvectors).
void jacobi3(double a[3][3], double d[3], double v[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (std::fabs(a[p][q]) < 1e-300) continue;
                const double th = 0.5 * (a[q][q] - a[p][p]) / a[p][q];
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 3; ++i) d[i] = a[i][i];
}

// Union mass properties by voxel integration, then the principal frame.
Template make_template(const std::vector<V>& pos, const std::vector<double>& rad, double density) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rmin = 1e300;
    for (size_t k = 0; k < pos.size(); ++k) {
        const double c[3] = {pos[k].x, pos[k].y, pos[k].z};
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], c[a] - rad[k]); hi[a] = std::max(hi[a], c[a] + rad[k]); }
        rmin = std::min(rmin, rad[k]);
    }
    const double h = rmin / 24.0;
    const int n[3] = {(int)std::ceil((hi[0] - lo[0]) / h), (int)std::ceil((hi[1] - lo[1]) / h),
                      (int)std::ceil((hi[2] - lo[2]) / h)};
    double m = 0, mx = 0, my = 0, mz = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
    const double dv = h * h * h * density;
    for (int i = 0; i < n[0]; ++i) {
        const double x = lo[0] + (i + 0.5) * h;
        for (int j = 0; j < n[1]; ++j) {
            const double y = lo[1] + (j + 0.5) * h;
            for (int k = 0; k < n[2]; ++k) {
                const double z = lo[2] + (k + 0.5) * h;
                bool in = false;
                for (size_t s = 0; s < pos.size() && !in; ++s) {
                    const double dx = x - pos[s].x, dy = y - pos[s].y, dz = z - pos[s].z;
                    in = dx * dx + dy * dy + dz * dz <= rad[s] * rad[s];
                }
                if (!in) continue;
                m += dv; mx += dv * x; my += dv * y; mz += dv * z;
                sxx += dv * x * x; syy += dv * y * y; szz += dv * z * z;
                sxy += dv * x * y; sxz += dv * x * z; syz += dv * y * z;
            }
        }
    }
    const V com = {mx / m, my / m, mz / m};
    // central second moments
    const double cxx = sxx - m * com.x * com.x, cyy = syy - m * com.y * com.y, czz = szz - m * com.z * com.z;
    const double cxy = sxy - m * com.x * com.y, cxz = sxz - m * com.x * com.z, cyz = syz - m * com.y * com.z;
    double I[3][3] = {{cyy + czz, -cxy, -cxz}, {-cxy, cxx + czz, -cyz}, {-cxz, -cyz, cxx + cyy}};
    double d[3], v[3][3];
    jacobi3(I, d, v);
    Template t;
    t.mass = m;
    for (int a = 0; a < 3; ++a) t.I[a] = d[a];
    for (size_t k = 0; k < pos.size(); ++k) {
        const V p = pos[k] - com;
        // principal-frame coordinates: columns of v are the axes
        V b = {p.x * v[0][0] + p.y * v[1][0] + p.z * v[2][0], p.x * v[0][1] + p.y * v[1][1] + p.z * v[2][1],
               p.x * v[0][2] + p.y * v[1][2] + p.z * v[2][2]};
        b = {f32(b.x), f32(b.y), f32(b.z)};
        t.off.push_back(b);
        t.r.push_back(f32(rad[k]));
        t.rbound = std::max(t.rbound, std::sqrt(dot(b, b)) + t.r.back());
    }
    return t;
}

// Body -> world rotation (the reference's Quat2IvDCM).
void ivdcm(const double q[4], double m[3][3]) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    m[0][0] = x * x - y * y - z * z + w * w;
    m[1][0] = 2.0 * (x * y + z * w);
    m[2][0] = 2.0 * (x * z - y * w);
    m[0][1] = 2.0 * (x * y - z * w);
    m[1][1] = -x * x + y * y - z * z + w * w;
    m[2][1] = 2.0 * (y * z + x * w);
    m[0][2] = 2.0 * (x * z + y * w);
    m[1][2] = 2.0 * (y * z - x * w);
    m[2][2] = -x * x - y * y + z * z + w * w;
}
V mapv(const double m[3][3], V u) {
    const double a = m[0][0] * u.x + m[0][1] * u.y + m[0][2] * u.z;
    const double b = m[1][0] * u.x + m[1][1] * u.y + m[1][2] * u.z;
    const double c = m[2][0] * u.x + m[2][1] * u.y + m[2][2] * u.z;
    return {a, b, c};
}

// Uniform random unit quaternion (Shoemake), stored x y z w.
void rand_quat(std::mt19937_64& g, double q[4]) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double u1 = U(g), u2 = U(g), u3 = U(g);
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1), tp = 2.0 * 3.14159265358979323846;
    q[0] = a * std::sin(tp * u2);
    q[1] = a * std::cos(tp * u2);
    q[2] = b * std::sin(tp * u3);
    q[3] = b * std::cos(tp * u3);
}

struct Placed { double z, x, y, r; };

// drop points tried per clump in the deposition (the lowest landing wins)
constexpr int kDropTries = 8;
constexpr double kPi = 3.14159265358979323846;

}  // namespace

namespace nbc {
// Rocks/Penetrator.bt in the format nbc_load_penetrator reads
// (src/dem_penetrator_io.cu:21-116, "cone R H NSEG" for the triangulated
// cone); without a penetrator, a far-away dummy sphere.
int write_penetrator_text(const std::string& path, const nbc_penetrator* pene) {
    std::ofstream f(path);
    if (!f) return -1;
    f << std::setprecision(17);
    if (pene) {
        if (pene->shape == NBC_PENE_SPHERE) f << "sphere " << pene->radius << "\n";
        else if (pene->shape == NBC_PENE_CONE)  // the mesh holds two triangles per rim segment
            f << "cone " << pene->radius << " " << pene->height << " " << (pene->ntri > 0 ? pene->ntri / 2 : 64)
              << "\n";
        else f << "cylinder " << pene->radius << " " << pene->height << "\n" << pene->dir[0] << " " << pene->dir[1]
               << " " << pene->dir[2] << "\n";
        f << pene->prop[0] << " " << pene->prop[1] << " " << pene->prop[2] << " " << pene->prop[3] << "\n"
          << pene->pos[0] << " " << pene->pos[1] << " " << pene->pos[2] << "\n" << pene->vel[0] << " "
          << pene->vel[1] << " " << pene->vel[2] << "\n" << pene->quat[0] << " " << pene->quat[1] << " "
          << pene->quat[2] << " " << pene->quat[3] << "\n" << pene->angv[0] << " " << pene->angv[1] << " "
          << pene->angv[2] << "\n" << "motion "
          << (pene->motion == NBC_PENE_GRAVITY ? (pene->wait_for_contact ? "gravity_after_contact" : "gravity")
                                               : "constant_velocity")
          << "\n";
    } else {
        f << "sphere 0.001\n1 1 1 1\n0 0 100\n0 0 0\n0 0 0 1\n0 0 0\nmotion constant_velocity\n";
    }
    return f ? 0 : -1;
}
}  // namespace nbc

extern "C" void nbc_free_bed(nbc_bed_out* o) {
    if (!o) return;
    void* p[] = {o->clump_id, o->pos, o->local, o->mass, o->radius, o->cpos, o->cvel, o->cquat, o->cangv, o->cprop};
    for (void* q : p) std::free(q);
    o->clump_id = nullptr; o->pos = o->local = o->mass = o->radius = nullptr;
    o->cpos = o->cvel = o->cquat = o->cangv = o->cprop = nullptr;
    o->n = o->nc = 0;
}

extern "C" int nbc_generate_bed(const nbc_bed_spec* spec, nbc_bed_out* out) {
    if (!spec || !out) { nbc::set_global_error("nbc_generate_bed: null argument"); return NBC_E_BADINPUT; }
    std::memset(out, 0, sizeof(*out));
    std::mt19937_64 g(spec->seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double rho = 2650.0;
    std::vector<Template> tpl;
    std::vector<int> ctpl;
    int nc = spec->nclumps;
    nbc_params& P = out->params;
    std::memset(&P, 0, sizeof(P));
    P.Ymod = 1e8; P.Prat = 0.3; P.Frcoe = 0.5; P.NormRes = 0.5; P.Frvel = 5e-4; P.Beta = 0.5;
    P.MU_T = 0.5; P.MU_R = 0.5; P.Cohesion = 0.0; P.Gravity = -9.81; P.GridLevels = 1; P.Update = 0;
    P.isIntruder = 0; P.isSelfGravity = 0; P.isWallForce = 1; P.isQuaternion = 1;
    P.StartStep = 0;
    if (spec->kind == 0) {
        // C0: 3 collinear spheres at 0.8 r spacing, r = 1 mm * U[0.9, 1.1] per clump
        if (nc <= 0) nc = 2000;
        const std::vector<V> base = {{-0.8, 0, 0}, {0, 0, 0}, {0.8, 0, 0}};
        Template unit = make_template(base, {1.0, 1.0, 1.0}, 1.0);
        ctpl.resize(nc);
        for (int k = 0; k < nc; ++k) {
            const double r = 1.0e-3 * (0.9 + 0.2 * U(g));
            Template t;
            t.mass = f32(unit.mass * rho * r * r * r);
            for (int a = 0; a < 3; ++a) t.I[a] = unit.I[a] * rho * r * r * r * r * r;
            t.off = {{f32(-0.8 * r), 0, 0}, {0, 0, 0}, {f32(0.8 * r), 0, 0}};
            t.r = {f32(r), f32(r), f32(r)};
            t.rbound = 0.8 * r + r;
            tpl.push_back(t);
            ctpl[k] = k;
        }
        P.StepSize = 1e-6; P.EndStep = 20000; P.OutputInterval = 10000; P.BranchLmtLen = 128; P.Thresh = 0.4e-3;
    } else if (spec->kind == 1) {
        // C1: 8 templates of 2..8 spheres, lognormal radii (median 1 mm, sigma 0.15, clipped to [0.6, 1.6] mm)
        if (nc <= 0) nc = 200000;
        std::lognormal_distribution<double> LN(std::log(1.0e-3), 0.15);
        for (int t = 0; t < 8; ++t) {
            const int ns = 2 + (int)(g() % 7);
            std::vector<V> ps;
            std::vector<double> rs;
            for (int s = 0; s < ns; ++s) {
                double r = LN(g);
                r = std::min(1.6e-3, std::max(0.6e-3, r));
                if (s == 0) { ps.push_back({0, 0, 0}); rs.push_back(r); continue; }
                // attach to a random earlier sphere at 0.8 x the mean radius
                const int par = (int)(g() % s);
                double q[4];
                rand_quat(g, q);
                double m[3][3];
                ivdcm(q, m);
                const V dir = mapv(m, V{1, 0, 0});
                ps.push_back(ps[par] + (0.8 * 0.5 * (r + rs[par])) * dir);
                rs.push_back(r);
            }
            Template T = make_template(ps, rs, rho);
            T.mass = f32(T.mass);
            tpl.push_back(T);
        }
        ctpl.resize(nc);
        for (int k = 0; k < nc; ++k) ctpl[k] = (int)(g() % 8);
        P.StepSize = 5e-7; P.EndStep = 10000; P.OutputInterval = 1000; P.BranchLmtLen = 128; P.Thresh = 0.25e-3;
    } else if (spec->kind == 2) {
        // C2: 32 unit templates of 1..6 spheres (relative radii U[0.8, 1]),
        // each clump scaled by s log-uniform over [0.25, 2] mm, so radii fall
        // in [0.2, 2.0] mm with a log-uniform spread
        if (nc <= 0) nc = 571429;
        std::vector<Template> unit;
        for (int t = 0; t < 32; ++t) {
            const int ns = 1 + t % 6;
            std::vector<V> ps;
            std::vector<double> rs;
            for (int q = 0; q < ns; ++q) {
                const double r = 0.8 + 0.2 * U(g);
                if (q == 0) { ps.push_back({0, 0, 0}); rs.push_back(r); continue; }
                const int par = (int)(g() % q);
                double qq[4];
                rand_quat(g, qq);
                double m[3][3];
                ivdcm(qq, m);
                ps.push_back(ps[par] + (0.8 * 0.5 * (r + rs[par])) * mapv(m, V{1, 0, 0}));
                rs.push_back(r);
            }
            unit.push_back(make_template(ps, rs, 1.0));
        }
        const double lo = std::log(0.25e-3), hi = std::log(2.0e-3);
        ctpl.resize(nc);
        for (int k = 0; k < nc; ++k) {
            const Template& u = unit[g() % 32];
            const double sc = std::exp(lo + (hi - lo) * U(g));
            Template t;
            t.mass = f32(u.mass * rho * sc * sc * sc);
            for (int a = 0; a < 3; ++a) t.I[a] = u.I[a] * rho * sc * sc * sc * sc * sc;
            for (size_t q = 0; q < u.off.size(); ++q) {
                t.off.push_back({f32(u.off[q].x * sc), f32(u.off[q].y * sc), f32(u.off[q].z * sc)});
                t.r.push_back(f32(u.r[q] * sc));
                t.rbound = std::max(t.rbound, std::sqrt(dot(t.off.back(), t.off.back())) + t.r.back());
            }
            tpl.push_back(t);
            ctpl[k] = k;
        }
        P.Frcoe = 0.6;
        P.GridLevels = 3;
        P.StepSize = 2e-7; P.EndStep = 5000; P.OutputInterval = 1000; P.BranchLmtLen = 256; P.Thresh = 0.08e-3;
    } else if (spec->kind == 3) {
        // DEMmain sphere bed: single spheres, lognormal radii as C1, solid
        // sphere mass and inertia; deposited like C1 into a cylinder
        if (nc <= 0) nc = 100000;
        std::lognormal_distribution<double> LN(std::log(1.0e-3), 0.15);
        ctpl.resize(nc);
        for (int k = 0; k < nc; ++k) {
            const double r = std::min(1.6e-3, std::max(0.6e-3, LN(g)));
            Template t;
            t.mass = f32(rho * 4.0 / 3.0 * kPi * r * r * r);
            for (int a = 0; a < 3; ++a) t.I[a] = 0.4 * t.mass * r * r;
            t.off = {{0, 0, 0}};
            t.r = {f32(r)};
            t.rbound = r;
            tpl.push_back(t);
            ctpl[k] = k;
        }
        P.StepSize = 5e-7; P.EndStep = 10000; P.OutputInterval = 1000; P.BranchLmtLen = 64; P.Thresh = 0.25e-3;
    } else {
        nbc::set_global_error("nbc_generate_bed: unknown kind");
        return NBC_E_BADINPUT;
    }
    int n = 0;
    for (int k = 0; k < nc; ++k) n += (int)tpl[ctpl[k]].off.size();
    out->n = n;
    out->nc = nc;
    out->clump_id = (int32_t*)std::malloc(sizeof(int32_t) * n);
    out->pos = (double*)std::malloc(sizeof(double) * 3 * n);
    out->local = (double*)std::malloc(sizeof(double) * 3 * n);
    out->mass = (double*)std::malloc(sizeof(double) * n);
    out->radius = (double*)std::malloc(sizeof(double) * n);
    out->cpos = (double*)std::malloc(sizeof(double) * 3 * nc);
    out->cvel = (double*)std::calloc(3 * (size_t)nc, sizeof(double));
    out->cquat = (double*)std::malloc(sizeof(double) * 4 * nc);
    out->cangv = (double*)std::calloc(3 * (size_t)nc, sizeof(double));
    out->cprop = (double*)std::malloc(sizeof(double) * 4 * nc);
    if (!out->clump_id || !out->pos || !out->local || !out->mass || !out->radius || !out->cpos || !out->cvel ||
        !out->cquat || !out->cangv || !out->cprop) {
        nbc_free_bed(out);
        nbc::set_global_error("nbc_generate_bed: out of memory");
        return NBC_E_BADINPUT;
    }
    std::vector<double> quats(4 * (size_t)nc);
    for (int k = 0; k < nc; ++k) rand_quat(g, &quats[4 * (size_t)k]);
    std::vector<V> cpos(nc);
    nbc_wall& W = out->wall;
    std::memset(&W, 0, sizeof(W));
    W.quat[3] = 1.0;
    if (spec->kind == 0) {
        // cube 40 x 40 x 80 mm, lattice of 4 mm cells, shuffled sites
        const double Lx = spec->container[0] > 0 ? spec->container[0] : 0.04;
        const double Ly = spec->container[1] > 0 ? spec->container[1] : 0.04;
        double Lz = spec->container[2] > 0 ? spec->container[2] : 0.08;
        double rb = 0;
        for (const Template& t : tpl) rb = std::max(rb, t.rbound);
        const double sp = 2.0 * rb + 2e-5;
        const int nx = (int)std::floor(Lx / sp), ny = (int)std::floor(Ly / sp);
        int nz = (int)std::floor(Lz / sp);
        if ((long long)nx * ny * nz < nc) {
            nz = (int)((nc + (long long)nx * ny - 1) / ((long long)nx * ny));
            Lz = nz * sp;
        }
        std::vector<int> sites((size_t)nx * ny * nz);
        for (size_t s = 0; s < sites.size(); ++s) sites[s] = (int)s;
        std::shuffle(sites.begin(), sites.end(), g);
        std::sort(sites.begin(), sites.begin() + nc);
        for (int k = 0; k < nc; ++k) {
            const int s = sites[k], ix = s % nx, iy = (s / nx) % ny, iz = s / (nx * ny);
            cpos[k] = {(ix + 0.5) * (Lx / nx), (iy + 0.5) * (Ly / ny), (iz + 0.5) * sp};
        }
        W.type = NBC_WALL_CUBE;
        W.side[0] = Lx / 2; W.side[1] = Ly / 2; W.side[2] = Lz / 2;
        W.pos[0] = Lx / 2; W.pos[1] = Ly / 2; W.pos[2] = Lz / 2;
        P.WorldOrigin[0] = 0; P.WorldOrigin[1] = 0; P.WorldOrigin[2] = 0;
        P.WorldSize[0] = Lx; P.WorldSize[1] = Ly; P.WorldSize[2] = Lz;
    } else {
        // C1 / C2: sequential deposition into a cylinder (C1, axis z, centred
        // on x = y = 0) or a square section [0, L]^2 (C2), floor at z = 0.
        // Every clump drops at a random (x, y) and stops at its first contact
        // with the floor or an earlier sphere, so the bed is contact-connected
        // without initial overlaps.
        const bool cyl = spec->kind == 1 || spec->kind == 3;
        const double Rc = spec->container[0] > 0 ? spec->container[0] : (cyl ? 0.06 : 0.15);
        const double ext = cyl ? 2.0 * Rc : Rc;  // extent of the section along x and y
        const double org = cyl ? -Rc : 0.0;      // its lower x / y
        double rmax = 0;
        for (const Template& t : tpl)
            for (double r : t.r) rmax = std::max(rmax, r);
        const double cs = 2.0 * rmax;
        const int nco = (int)std::ceil(ext / cs) + 1;
        std::vector<std::vector<Placed>> col((size_t)nco * nco);
        auto colidx = [&](double x, double y, int& cx, int& cy) {
            cx = std::min(nco - 1, std::max(0, (int)std::floor((x - org) / cs)));
            cy = std::min(nco - 1, std::max(0, (int)std::floor((y - org) / cs)));
        };
        std::vector<V> off;
        double zmax = 0;
        for (int k = 0; k < nc; ++k) {
            const Template& t = tpl[ctpl[k]];
            double m[3][3];
            ivdcm(&quats[4 * (size_t)k], m);
            off.resize(t.off.size());
            double rxy = 0;
            for (size_t s = 0; s < t.off.size(); ++s) {
                off[s] = mapv(m, t.off[s]);
                rxy = std::max(rxy, std::sqrt(off[s].x * off[s].x + off[s].y * off[s].y) + t.r[s]);
            }
            const double rr = (cyl ? Rc : 0.5 * Rc) - rxy - 1e-6;
            auto landing = [&](double x, double y) {
                double best = -1e300;
                for (size_t s = 0; s < t.off.size(); ++s) {
                    const double rs = t.r[s], xs = x + off[s].x, ys = y + off[s].y;
                    best = std::max(best, rs - off[s].z);
                    int cx, cy;
                    colidx(xs, ys, cx, cy);
                    for (int a = std::max(0, cx - 1); a <= std::min(nco - 1, cx + 1); ++a)
                        for (int b = std::max(0, cy - 1); b <= std::min(nco - 1, cy + 1); ++b) {
                            const std::vector<Placed>& c = col[(size_t)a * nco + b];
                            for (const Placed& p : c) {  // descending z
                                if (p.z + rs + rmax - off[s].z <= best) break;
                                const double dx = xs - p.x, dy = ys - p.y, d2 = dx * dx + dy * dy, rsum = rs + p.r;
                                if (d2 >= rsum * rsum) continue;
                                best = std::max(best, p.z + std::sqrt(rsum * rsum - d2) - off[s].z);
                            }
                        }
                }
                return best;
            };
            // the lowest of a few random drop points: fills pockets, denser bed
            double x = 0, y = 0, best = 1e300;
            for (int tries = 0; tries < kDropTries; ++tries) {
                double cx_, cy_;
                if (cyl) {
                    do { cx_ = (2 * U(g) - 1) * rr; cy_ = (2 * U(g) - 1) * rr; } while (cx_ * cx_ + cy_ * cy_ > rr * rr);
                } else {
                    cx_ = 0.5 * Rc + (2 * U(g) - 1) * rr;
                    cy_ = 0.5 * Rc + (2 * U(g) - 1) * rr;
                }
                const double h = landing(cx_, cy_);
                if (h < best) { best = h; x = cx_; y = cy_; }
            }
            cpos[k] = {x, y, best};
            for (size_t s = 0; s < t.off.size(); ++s) {
                const Placed p{best + off[s].z, x + off[s].x, y + off[s].y, t.r[s]};
                int cx, cy;
                colidx(p.x, p.y, cx, cy);
                std::vector<Placed>& c = col[(size_t)cx * nco + cy];
                auto it = std::lower_bound(c.begin(), c.end(), p, [](const Placed& u, const Placed& v) { return u.z > v.z; });
                c.insert(it, p);
                zmax = std::max(zmax, p.z + p.r);
            }
        }
        const double H = 0.5 * (zmax + 4.0 * rmax);  // half height: top cap above the bed
        if (!cyl) {
            W.type = NBC_WALL_CUBE;
            W.side[0] = 0.5 * Rc; W.side[1] = 0.5 * Rc; W.side[2] = H;
            W.pos[0] = 0.5 * Rc; W.pos[1] = 0.5 * Rc; W.pos[2] = H;
            P.WorldOrigin[0] = 0; P.WorldOrigin[1] = 0; P.WorldOrigin[2] = 0;
            P.WorldSize[0] = Rc; P.WorldSize[1] = Rc; P.WorldSize[2] = 2 * H;
        } else {
            W.type = NBC_WALL_CYLINDER;
            W.radius = Rc;
            W.height = H;
            W.dir[2] = 1.0;
            W.prop[0] = W.prop[1] = W.prop[2] = W.prop[3] = 1.0;
            W.pos[0] = 0; W.pos[1] = 0; W.pos[2] = H;
            P.WorldOrigin[0] = -Rc; P.WorldOrigin[1] = -Rc; P.WorldOrigin[2] = 0;
            P.WorldSize[0] = 2 * Rc; P.WorldSize[1] = 2 * Rc; P.WorldSize[2] = 2 * H;
        }
    }
    for (int a = 0; a < 3; ++a) {
        P.BoxLowLmt[a] = P.WorldOrigin[a];
        P.BoxUpLmt[a] = P.WorldOrigin[a] + P.WorldSize[a];
    }
    int si = 0;
    for (int k = 0; k < nc; ++k) {
        const Template& t = tpl[ctpl[k]];
        const double* q = &quats[4 * (size_t)k];
        double m[3][3];
        ivdcm(q, m);
        for (int a = 0; a < 4; ++a) out->cquat[4 * (size_t)k + a] = q[a];
        out->cpos[3 * (size_t)k] = cpos[k].x;
        out->cpos[3 * (size_t)k + 1] = cpos[k].y;
        out->cpos[3 * (size_t)k + 2] = cpos[k].z;
        out->cprop[4 * (size_t)k] = t.I[0];
        out->cprop[4 * (size_t)k + 1] = t.I[1];
        out->cprop[4 * (size_t)k + 2] = t.I[2];
        out->cprop[4 * (size_t)k + 3] = t.mass;
        const double ms = f32(t.mass / (double)t.off.size());
        for (size_t s = 0; s < t.off.size(); ++s, ++si) {
            const V w = mapv(m, t.off[s]);
            out->clump_id[si] = k;
            out->pos[3 * (size_t)si] = cpos[k].x + w.x;
            out->pos[3 * (size_t)si + 1] = cpos[k].y + w.y;
            out->pos[3 * (size_t)si + 2] = cpos[k].z + w.z;
            out->local[3 * (size_t)si] = t.off[s].x;
            out->local[3 * (size_t)si + 1] = t.off[s].y;
            out->local[3 * (size_t)si + 2] = t.off[s].z;
            out->mass[si] = ms;
            out->radius[si] = t.r[s];
        }
    }
    P.ParticleNum = n;
    P.ClumpNum = nc;
    return 0;
}

// Writes world.par (38 keys), ph/ch text files, Rocks/ wall and penetrator
// files and nbc.par into dir, in the reference's input formats.
extern "C" int nbc_save_bed_text(const char* dir, const nbc_bed_out* b, const nbc_penetrator* pene) {
    if (!dir || !b) return NBC_E_BADINPUT;
    const std::string d(dir);
    std::string rocks = d + "/Rocks";
    std::string cmd;
    {
        std::ofstream f(d + "/world.par");
        if (!f) { nbc::set_global_error("cannot write " + d + "/world.par"); return NBC_E_IO; }
        const nbc_params& P = b->params;
        f << std::setprecision(17);
        f << "# synthetic bed\nNumberBody = " << P.ParticleNum << "\nClumpNum = " << P.ClumpNum
          << "\nBranchLmtLen = " << P.BranchLmtLen << "\nBodyFileDir = ./Rocks/\nIniConFile = ph.00000000.bt\n"
          << "StepSize = " << P.StepSize << "\nEndStep = " << P.EndStep << "\nStartStep = " << P.StartStep
          << "\nOutputInterval = " << P.OutputInterval << "\nYmod = " << P.Ymod << "\nPrat = " << P.Prat
          << "\nFrcoe = " << P.Frcoe << "\nNormRes = " << P.NormRes << "\nFrvel = " << P.Frvel << "\nBeta = " << P.Beta
          << "\nMU_T = " << P.MU_T << "\nMU_R = " << P.MU_R << "\nCohesion = " << P.Cohesion
          << "\nGravity = " << P.Gravity << "\nThresh = " << P.Thresh << "\nGridLevels = " << P.GridLevels
          << "\nUpdate = " << P.Update << "\n";
        const char* ax = "xyz";
        for (int a = 0; a < 3; ++a) f << "WorldSize_" << ax[a] << " = " << P.WorldSize[a] << "\n";
        for (int a = 0; a < 3; ++a) f << "WorldOrigin_" << ax[a] << " = " << P.WorldOrigin[a] << "\n";
        for (int a = 0; a < 3; ++a) f << "BoxUpLmt_" << ax[a] << " = " << P.BoxUpLmt[a] << "\n";
        for (int a = 0; a < 3; ++a) f << "BoxLowLmt_" << ax[a] << " = " << P.BoxLowLmt[a] << "\n";
        f << "isIntruder = " << P.isIntruder << "\nisSelfGravity = " << P.isSelfGravity
          << "\nisWallForce = " << P.isWallForce << "\nisQuaternion = " << P.isQuaternion << "\n";
    }
    {
        FILE* f = std::fopen((d + "/ph.00000000.bt").c_str(), "w");
        if (!f) return NBC_E_IO;
        for (int i = 0, prev = -1, sub = 0; i < b->n; ++i) {
            sub = b->clump_id[i] == prev ? sub + 1 : 0;
            prev = b->clump_id[i];
            std::fprintf(f, "%d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", i, sub, b->clump_id[i],
                         b->pos[3 * i], b->pos[3 * i + 1], b->pos[3 * i + 2], b->local[3 * i], b->local[3 * i + 1],
                         b->local[3 * i + 2], b->mass[i], b->radius[i]);
        }
        std::fclose(f);
    }
    {
        FILE* f = std::fopen((d + "/ch.00000000.bt").c_str(), "w");
        if (!f) return NBC_E_IO;
        for (int k = 0; k < b->nc; ++k)
            std::fprintf(f,
                         "%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g "
                         "%.17g %.17g %.17g\n",
                         k, b->cpos[3 * k], b->cpos[3 * k + 1], b->cpos[3 * k + 2], b->cvel[3 * k], b->cvel[3 * k + 1],
                         b->cvel[3 * k + 2], b->cquat[4 * k], b->cquat[4 * k + 1], b->cquat[4 * k + 2],
                         b->cquat[4 * k + 3], b->cangv[3 * k], b->cangv[3 * k + 1], b->cangv[3 * k + 2],
                         b->cprop[4 * k], b->cprop[4 * k + 1], b->cprop[4 * k + 2], b->cprop[4 * k + 3]);
        std::fclose(f);
    }
    ::mkdir(rocks.c_str(), 0755);
    const nbc_wall& W = b->wall;
    {
        std::ofstream f(rocks + "/CubeWall.bt");
        f << std::setprecision(17);
        if (W.type == NBC_WALL_CUBE)
            f << W.side[0] << " " << W.side[1] << " " << W.side[2] << "\n" << W.pos[0] << " " << W.pos[1] << " "
              << W.pos[2] << "\n" << W.vel[0] << " " << W.vel[1] << " " << W.vel[2] << "\n" << W.quat[0] << " "
              << W.quat[1] << " " << W.quat[2] << " " << W.quat[3] << "\n" << W.angv[0] << " " << W.angv[1] << " "
              << W.angv[2] << "\n";
        else
            f << "1 1 1\n0 0 0\n0 0 0\n0 0 0 1\n0 0 0\n";
    }
    {
        std::ofstream f(rocks + "/CylindWall.bt");
        f << std::setprecision(17);
        if (W.type == NBC_WALL_CYLINDER)
            f << W.radius << " " << W.height << "\n" << W.dir[0] << " " << W.dir[1] << " " << W.dir[2] << "\n"
              << W.prop[0] << " " << W.prop[1] << " " << W.prop[2] << " " << W.prop[3] << "\n" << W.pos[0] << " "
              << W.pos[1] << " " << W.pos[2] << "\n" << W.vel[0] << " " << W.vel[1] << " " << W.vel[2] << "\n"
              << W.quat[0] << " " << W.quat[1] << " " << W.quat[2] << " " << W.quat[3] << "\n" << W.angv[0] << " "
              << W.angv[1] << " " << W.angv[2] << "\n";
        else
            f << "1 1\n0 0 1\n1 1 1 1\n0 0 0\n0 0 0\n0 0 0 1\n0 0 0\n";
    }
    if (nbc::write_penetrator_text(rocks + "/Penetrator.bt", pene)) {
        nbc::set_global_error("cannot write " + rocks + "/Penetrator.bt");
        return NBC_E_IO;
    }
    {
        std::ofstream f(d + "/nbc.par");
        f << "WallType = " << (W.type == NBC_WALL_CYLINDER ? "cylinder" : "cube") << "\n";
        if (b->params.GridLevels > 1) {
            f << "HGridLevels = " << (int)b->params.GridLevels << "\n";
            // nested levels keep a coarse cell's spheres together (faster steps)
            if (b->params.GridLevels <= 4) f << "HGridNested = 1\n";
        }
    }
    (void)cmd;
    return 0;
}
