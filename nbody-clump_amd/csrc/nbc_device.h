// Device-side types and math for the MI355X clumped-sphere DEM kernels.
//
// Arithmetic helpers evaluate in the same order as the reference's
// src/type.h operators so that the FP64 instantiation, compiled with FP
// contraction off, reproduces the reference formulas bit for bit. The FP32
// ("mixed") instantiation uses the same code with T = float.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace nbc {

struct alignas(32) D4 { double x, y, z, w; };
struct alignas(16) F4 { float x, y, z, w; };

template <class T> struct Vec4T;
template <> struct Vec4T<double> { typedef D4 type; };
template <> struct Vec4T<float> { typedef F4 type; };

template <class T>
struct V3 {
    T x, y, z;
};

template <class T> __host__ __device__ __forceinline__ V3<T> mk(T x, T y, T z) { V3<T> v; v.x = x; v.y = y; v.z = z; return v; }
template <class T> __host__ __device__ __forceinline__ V3<T> operator+(const V3<T>& a, const V3<T>& b) { return mk<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> __host__ __device__ __forceinline__ V3<T> operator-(const V3<T>& a, const V3<T>& b) { return mk<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> __host__ __device__ __forceinline__ V3<T> operator*(T s, const V3<T>& u) { return mk<T>(s * u.x, s * u.y, s * u.z); }
template <class T> __host__ __device__ __forceinline__ T Dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __host__ __device__ __forceinline__ V3<T> Cross(const V3<T>& a, const V3<T>& b) {
    T x = a.y * b.z - a.z * b.y;
    T y = a.z * b.x - a.x * b.z;
    T z = a.x * b.y - a.y * b.x;
    return mk<T>(x, y, z);
}
// |v| with the reference's cut-off: zero at or below |v|^2 = 1e-32.
template <class T> __host__ __device__ __forceinline__ T Length(const V3<T>& v) {
    T t = v.x * v.x + v.y * v.y + v.z * v.z;
    return t > T(1e-32) ? sqrt(t) : T(0);
}
template <class T> __host__ __device__ __forceinline__ V3<T> cvt3(const V3<double>& v) { return mk<T>((T)v.x, (T)v.y, (T)v.z); }

// Body -> world rotation of quaternion (x, y, z, w).
template <class T> __host__ __device__ __forceinline__ void QuatIvDCM(T qx, T qy, T qz, T qw, T m[3][3]) {
    m[0][0] = qx * qx - qy * qy - qz * qz + qw * qw;
    m[1][0] = T(2.0) * (qx * qy + qz * qw);
    m[2][0] = T(2.0) * (qx * qz - qy * qw);
    m[0][1] = T(2.0) * (qx * qy - qz * qw);
    m[1][1] = -qx * qx + qy * qy - qz * qz + qw * qw;
    m[2][1] = T(2.0) * (qy * qz + qx * qw);
    m[0][2] = T(2.0) * (qx * qz + qy * qw);
    m[1][2] = T(2.0) * (qy * qz - qx * qw);
    m[2][2] = -qx * qx - qy * qy + qz * qz + qw * qw;
}
// World -> body rotation.
template <class T> __host__ __device__ __forceinline__ void QuatDCM(T qx, T qy, T qz, T qw, T m[3][3]) {
    m[0][0] = qx * qx - qy * qy - qz * qz + qw * qw;
    m[1][0] = T(2.0) * (qx * qy - qz * qw);
    m[2][0] = T(2.0) * (qx * qz + qy * qw);
    m[0][1] = T(2.0) * (qx * qy + qz * qw);
    m[1][1] = -qx * qx + qy * qy - qz * qz + qw * qw;
    m[2][1] = T(2.0) * (qy * qz - qx * qw);
    m[0][2] = T(2.0) * (qx * qz - qy * qw);
    m[1][2] = T(2.0) * (qy * qz + qx * qw);
    m[2][2] = -qx * qx - qy * qy + qz * qz + qw * qw;
}
template <class T> __host__ __device__ __forceinline__ V3<T> Map(const T m[3][3], const V3<T>& u) {
    T a = m[0][0] * u.x + m[0][1] * u.y + m[0][2] * u.z;
    T b = m[1][0] * u.x + m[1][1] * u.y + m[1][2] * u.z;
    T c = m[2][0] * u.x + m[2][1] * u.y + m[2][2] * u.z;
    return mk<T>(a, b, c);
}

// ---- pair-word layout of the front list ------------------------------------
// bits 0..28  partner sphere index j (sorted order)
// bit  29     intra-clump pair (listed, never a contact)
// bit  30/31  history in buffer 0/1 is valid (written by the step that wrote it)
constexpr uint32_t JMASK = 0x1FFFFFFFu;
constexpr uint32_t INTRA = 1u << 29;
__host__ __device__ __forceinline__ uint32_t VALID(int p) { return 1u << (30 + p); }

// back entries (owner, owner slot): bit 30 of the owner marks an intra-clump pair
constexpr int BINTRA = 1 << 30;

constexpr int HSTRIDE = 9;   // tangential(3) twisting(3) rolling(3) per history record
constexpr int MAX_FACES = 7; // cube 6 (or cylinder 3) + penetrator 1

// ---- run parameters passed by value ----------------------------------------
constexpr int MAX_LEVELS = 8;

// Container wall: cube (half-sides `side`) or cylinder (radius, half height,
// axis `dir`), centre `pos`, rigid-body velocity and angular velocity.
struct WallGeo {
    double side[3], rad, height, dir[3], pos[3], vel[3], angv[3];
};

// One bed of a batch: its uniform grid (cells cell_off .. cell_off + cell_num - 1
// of the shared cell arrays, the last one taking out-of-box spheres), box and wall.
struct BedGeo {
    double origin[3], cell[3], icell[3], boxlo[3], boxup[3];
    uint32_t G[3];
    int cell_off, cell_num, pad;
    WallGeo W;
};

struct DevParams {
    int N, Nc, B, CellNum;
    uint32_t G[3];
    double origin[3], cell[3], icell[3], boxlo[3], boxup[3];
    // hierarchical grid (L > 1): level 0 holds the largest spheres; cells are
    // numbered level-major from loff[t] (or nested, below); the last cell is
    // the out-of-box bucket
    int L;
    int loff[MAX_LEVELS];
    int ldim[MAX_LEVELS][3];
    double lcell[MAX_LEVELS];
    double lrmin[MAX_LEVELS];  // a sphere's level is the first t with r > lrmin[t]
    const int* clevel;         // or, when set, clevel[clump index] (DEMmain hgrid loop: the particle's type)
    int nested;                // 1: nested power-of-two levels numbered coarse-cell-major,
    int lblock;                //    lblock keys per coarse cell (loff[t] is the offset inside it)
    double dt, Ymod, Prat, Frcoe, Beta, Cohesion, Gravity, Damping, Thresh;
    double pene_frcoe;
    double cKn, cKs, idt; // 2E/(3(1-nu^2)), 2E/(1+nu)/(2-nu), 1/dt (FP32 force path)
    double pc[3];         // box centre: origin of the FP32 position copy
    double half;          // largest box half-extent
    float f32_marg;       // absolute slack of the FP32 distance pre-filter
    double near_delta;    // near-list skin: a sphere's slack is min(2a + delta, Thresh + a), a = |cum| at the build
    double near_horizon;  // > 0: per-sphere delta = speed bound * dt * near_horizon, clamped to [near_lo, near_hi]
    double near_lo, near_hi;
    double wl_slack;      // wall list: distance a sphere may move after the list is built (2 * Thresh)
    double pene_move;     // wall list: penetrator travel allowed before the list is rebuilt
    int wall_type;        // NBC_WALL_*
    int driver;           // NBC_DRIVER_*: 0 the clump loop, > 0 a DEMmain single-sphere loop
    int nfaces;           // wall faces in use (6 cube, 3 cylinder, 0 none)
    int pene_on;          // isIntruder
    int pene_face;        // history slot of the penetrator (= nfaces)
    // wall (single bed)
    WallGeo W;
    // batch of beds (nbeds > 1): per-bed grid, box and wall; bid[i] = bed of sphere i
    int nbeds;
    const BedGeo* beds;
    const int* bid;
    // graph replays: the step kernels' step argument is an offset from *sbase
    // (read when the kernel runs); nullptr: the argument is the step itself
    const int* sbase;
    // touching sphere-sphere pairs of the last contact launch, one count per
    // wave of its sphere blocks (the host adds them up; no same-address atomics)
    unsigned* wss;
    // penetrator (static part)
    int pshape, pmotion, pwait, ntri;
    double prad, pheight, pdir[3], pmass;
    double pbound;        // radius of a sphere about the penetrator origin enclosing the mesh
    const double* ptri;   // 9*ntri, penetrator frame
};

// Penetrator dynamic state, device resident.
struct PeneState {
    double pos[3], vel[3], angv[3], quat[4];
    double totF[3], totT[3];
    double lpos[3];       // position when the wall list was last built
    int contact_state;
    int pad;
};

// bits of the status word next to the halt step
constexpr int ST_CLUMP_BAD = 1;  // a clump state is not finite (sticky)
constexpr int ST_REBUILD = 2;    // a sphere moved more than Thresh: rebuild the lists
constexpr int ST_WLIST = 4;      // the penetrator left its allowance: rebuild the wall list
constexpr int ST_NEAR = 8;       // a sphere left its near-list allowance: rebuild the near lists

}  // namespace nbc
