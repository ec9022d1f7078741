// HIP kernels of the clumped-sphere DEM timestep for gfx950 (MI355X).
//
// One steady-state step is three launches:
//   k_sphere_pose    per sphere  FP64 sphere pose, displacement, rebuild and
//                                near-list flags (src/dem_clump.cu:113-138)
//   k_contacts_f32 / k_contacts_f64
//                    per sphere  sphere-sphere contacts, totF/totT, clump torque
//                                (src/dem_force_clump.cu:3-191,
//                                src/dem_clump.cu:89-101,
//                                src/dem_clumpsphere.cu:86-107); extra blocks
//                                of the same launch take the wall-list spheres'
//                                wall and penetrator contacts
//                                (src/dem_force_cube_wall_clump.cu:16-298,
//                                src/dem_force_cylind_wall.cu:152-237,
//                                src/dem_contact_penetrator.cu:8-84,
//                                src/dem_force_penetrator_clump.cu:4-182)
//   k_clump_after    per clump   deterministic member sum + second half-kick
//                                (src/dem_clumpsphere.cu:110-151,
//                                src/dem_integrate_clumpsphere.cu:110-141),
//                                then the next step's half-kick, drift and
//                                quaternion update (k_clump_before,
//                                src/dem_integrate_clumpsphere.cu:3-90)
// plus k_pene_finalize (one block) when a penetrator is present.
//
// A rebuild (hash, stable LSD radix sort, cell bounds, gather, two-pass
// neighbour search into CSR lists, back-slot resolution, history remap,
// clump member lists) replaces src/dem_grid.cu:105-750; it also builds the
// near lists (the pairs that can touch before the next rebuild) and the wall
// list (the spheres that can touch a wall or the penetrator).
//
// Every sphere-sphere pair is stored once (owner = lower sorted index) and
// evaluated by both spheres' waves from the same inputs, so each wave only
// writes its own spheres: no float atomics, results independent of
// scheduling. The owner writes the new
// history into the other buffer of a ping-pong pair; bits 30/31 of the pair
// word say whether each buffer holds a valid history, which replaces the
// reference's per-step memsets of every slot.
#include <hip/hip_runtime.h>

#include "nbc_device.h"
#include "nbc_kernels.h"

#pragma clang fp contract(off)

namespace nbc {

// ---------------------------------------------------------------------------
// small helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
// Step index of a step kernel: the argument itself, or an offset from the base
// a graph replay reads from device memory.
__device__ __forceinline__ int step_of(const DevParams& P, int n) { return P.sbase ? *P.sbase + n : n; }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Waves per block of the contact kernels. A block keeps its LDS until its
// slowest wave ends, and the waves of one block differ in how many flush
// rounds they run: one-wave blocks let each wave's slot refill as soon as it
// is done (FP32 K3 72.4 -> 69.9 us on the settled 1.12 M-sphere bed).
#ifndef K3_NW_F32
#define K3_NW_F32 1
#endif
#ifndef K3_NW_F64
#define K3_NW_F64 1
#endif
// Waves per SIMD the contact kernels are compiled for: at five the FP32 form
// needs 96 VGPRs without spilling on its contact path; four (72.5 us) leaves
// latency unhidden, six (78.4 us) spills in the flush. The FP64 form spills
// above three.
#ifndef K3_WAVES_F32
#define K3_WAVES_F32 5
#endif
#ifndef K3_WAVES_F64
#define K3_WAVES_F64 3
#endif

// Logical block index giving each XCD runs of G consecutive blocks, the runs
// of the 8 XCDs interleaved: neighbouring blocks (which share spheres) meet in
// one L2, and every XCD still gets work from the whole bed.
// The mapping permutes the first nb - nb % (8 G) blocks; the rest keep the
// hardware order, so every logical block is run exactly once.
__device__ __forceinline__ int xcd_group_block(int b, int nb, int G) {
    const int nfull = nb - nb % (8 * G);
    if (b >= nfull) return b;
    const int x = b & 7, k = b >> 3;
    return (k / G) * (8 * G) + x * G + (k % G);
}

template <class R> struct Ld4;
template <> struct Ld4<double> {
    static __device__ __forceinline__ V3<double> xyz(const D4& v) { return mk<double>(v.x, v.y, v.z); }
};
template <> struct Ld4<float> {
    static __device__ __forceinline__ V3<float> xyz(const F4& v) { return mk<float>(v.x, v.y, v.z); }
};
template <class R> __device__ __forceinline__ typename Vec4T<R>::type mk4(R x, R y, R z, R w) {
    typename Vec4T<R>::type v;
    v.x = x; v.y = y; v.z = z; v.w = w;
    return v;
}
__device__ __forceinline__ V3<double> xyz(const D4& v) { return mk<double>(v.x, v.y, v.z); }
__device__ __forceinline__ bool finite3(const V3<double>& v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// FP32 copy of a sphere position, relative to the box centre. Coordinates
// farther than two half-extents from the centre, or not finite, get w = -1:
// the contact kernel then skips the FP32 pre-filter for that sphere.
__device__ __forceinline__ F4 pos32(const DevParams& P, const V3<double>& x, double r) {
    const double dx = x.x - P.pc[0], dy = x.y - P.pc[1], dz = x.z - P.pc[2], lim = 2.0 * P.half;
    F4 o;
    o.x = (float)dx; o.y = (float)dy; o.z = (float)dz;
    o.w = (fabs(dx) <= lim && fabs(dy) <= lim && fabs(dz) <= lim) ? (float)r : -1.0f;
    return o;
}

// FP32 pre-filter of the sphere-sphere test Dn = ri + rj - |xj - xi| > 0.
// Each coordinate of pos32 is within half * 2^-23 of the exact value, so the
// FP32 distance differs from the FP64 one by less than
// 2*sqrt(3)*half*2^-23 + a few ulp of the distance; f32_marg = 4*half*2^-23
// and the relative slack 1e-5 cover both. Returns false only when the FP64
// test would fail too.
__device__ __forceinline__ bool maybe_touch(const DevParams& P, const F4& a, const F4& b) {
    if (a.w < 0.0f || b.w < 0.0f) return true;
    const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float rs = (a.w + b.w) * 1.00001f + P.f32_marg;
    return !(d2 > rs * rs);
}

// ---------------------------------------------------------------------------
// K1: clump half-kick, drift and attitude (FP64). Thread 0 of block 0 also
// advances the penetrator (src/dem_integrate_penetrator.cu:3-41) and clears
// the step counters.
// Thread 0 of K1: clears the step counters and advances the penetrator
// (src/dem_integrate_penetrator.cu:3-41) for step n.
__device__ __forceinline__ void k1_thread0(const DevParams& P, PeneState* pst, unsigned long long* counters,
                                           int* halt, int n) {
    counters[1] = 0ull;  // touching wall contacts (the sphere-sphere count is per wave, DevParams::wss)
    if (!P.pene_on) return;
    const double Dt = P.dt;
    if (P.pmotion == NBC_PENE_GRAVITY && (!P.pwait || pst->contact_state != 0)) {
        V3<double> F = mk<double>(pst->totF[0], pst->totF[1], pst->totF[2]);
        V3<double> pf = (-1.0 / P.pmass) * F;
        pf.z = pf.z + P.Gravity;
        V3<double> pos = mk<double>(pst->pos[0], pst->pos[1], pst->pos[2]);
        V3<double> vel = mk<double>(pst->vel[0], pst->vel[1], pst->vel[2]);
        pos = pos + Dt * vel + 0.5 * Dt * Dt * pf;
        vel = vel + 0.5 * Dt * pf;
        pst->pos[0] = pos.x; pst->pos[1] = pos.y; pst->pos[2] = pos.z;
        pst->vel[0] = vel.x; pst->vel[1] = vel.y; pst->vel[2] = vel.z;
    } else {
        pst->pos[0] = pst->pos[0] + Dt * pst->vel[0];
        pst->pos[1] = pst->pos[1] + Dt * pst->vel[1];
        pst->pos[2] = pst->pos[2] + Dt * pst->vel[2];
    }
    // the wall list covers the penetrator's reach plus pene_move: past that,
    // stop this step after k_sphere_pose for a new list
    const double dx = pst->pos[0] - pst->lpos[0], dy = pst->pos[1] - pst->lpos[1], dz = pst->pos[2] - pst->lpos[2];
    if (!(dx * dx + dy * dy + dz * dz <= P.pene_move * P.pene_move)) {
        atomicOr(&halt[3], ST_WLIST);
        atomicMin(&halt[0], n);
    }
}

// Clump force and torque scaled to accelerations (src/dem_integrate_clumpsphere.cu:12-20).
__device__ __forceinline__ void scale_ft(const DevParams& P, const D4& pr, V3<double>& F, V3<double>& T) {
    F = (1.0 / pr.w) * F;
    F.z = F.z + P.Gravity;
    T.x = (1.0 / pr.x) * T.x;
    T.y = (1.0 / pr.y) * T.y;
    T.z = (1.0 / pr.z) * T.z;
}

// K1 of one clump: half-kick, drift and attitude (FP64) from the scaled force
// and torque of the previous step.
template <class R>
__device__ __forceinline__ void clump_drift(const DevParams& P, const ClumpBuf<R>& C, int i, const D4& pc,
                                            const D4& qc, V3<double> vel, V3<double> w, const V3<double>& F,
                                            const V3<double>& T, int* halt) {
    const double Dt = P.dt;
    const int ty = P.driver == NBC_DRIVER_SPLIT ? C.type[i] : 0;
    if (ty != 0) {
        // glue kinematic types (src/dem_integrate_withglue.cu:3-63): type 1 is
        // driven at v = (-0.2 y, -0.2 x, 0) from its position before the
        // drift, type 2 stays put; neither is kicked or turned
        V3<double> dist = mk<double>(0.0, 0.0, 0.0);
        D4 po = pc;
        if (ty == 1) {
            const double Omega = 0.2;
            vel.x = -Omega * pc.y;
            vel.y = -Omega * pc.x;
            vel.z = 0.0;
            dist = Dt * vel;
            po.x = pc.x + dist.x; po.y = pc.y + dist.y; po.z = pc.z + dist.z;
        }
        D4 vo; vo.x = vel.x; vo.y = vel.y; vo.z = vel.z; vo.w = 0.0;
        D4 wo; wo.x = w.x; wo.y = w.y; wo.z = w.z; wo.w = 0.0;
        D4 dd; dd.x = dist.x; dd.y = dist.y; dd.z = dist.z; dd.w = 0.0;
        C.pq[2 * i] = po;
        C.vw[2 * i] = vo;
        C.vw[2 * i + 1] = wo;
        C.dist[i] = dd;
        const bool ok = finite3(vel) && finite3(w);
        if (!ok) atomicOr(&halt[3], ST_CLUMP_BAD);
        C.snap[2 * i] = mk4<R>((R)vel.x, (R)vel.y, (R)vel.z, ok ? R(1) : R(0));
        C.snap[2 * i + 1] = mk4<R>((R)w.x, (R)w.y, (R)w.z, R(0));
        return;
    }
    V3<double> pos = xyz(pc);
    V3<double> dist = Dt * vel + 0.5 * Dt * Dt * F;
    pos = pos + dist;
    vel = vel + 0.5 * Dt * F;
    w = w + 0.5 * Dt * T;
    if (P.driver) {
        // DEMmain keeps each sphere's displacement as the sum of its drifts
        // (src/dem_integrate.cu:36-38); k_sphere_pose adds this one
        D4 dd; dd.x = dist.x; dd.y = dist.y; dd.z = dist.z; dd.w = 0.0;
        C.dist[i] = dd;
    }
    // attitude, src/dem_integrate_clumpsphere.cu:45-90 (T is the same scaled torque)
    double M[3][3];
    QuatDCM<double>(qc.x, qc.y, qc.z, qc.w, M);
    V3<double> Wb = Map<double>(M, w), Wd = Map<double>(M, T);
    const double qx = qc.x, qy = qc.y, qz = qc.z, qw = qc.w;
    double dx = 0.50 * (Wb.x * qw - Wb.y * qz + Wb.z * qy);
    double dy = 0.50 * (Wb.x * qz + Wb.y * qw - Wb.z * qx);
    double dz = 0.50 * (-Wb.x * qy + Wb.y * qx + Wb.z * qw);
    double dw = 0.50 * (-Wb.x * qx - Wb.y * qy - Wb.z * qz);
    double ddx = (Wd.x * qw - Wd.y * qz + Wd.z * qy) * 0.50 + (Wb.x * dw - Wb.y * dz + Wb.z * dy) * 0.50;
    double ddy = (Wd.x * qz + Wd.y * qw - Wd.z * qx) * 0.50 + (Wb.x * dz + Wb.y * dw - Wb.z * dx) * 0.50;
    double ddz = (-Wd.x * qy + Wd.y * qx + Wd.z * qw) * 0.50 + (-Wb.x * dy + Wb.y * dx + Wb.z * dw) * 0.50;
    double ddw = (-Wd.x * qx - Wd.y * qy - Wd.z * qz) * 0.50 + (-Wb.x * dx - Wb.y * dy - Wb.z * dz) * 0.50;
    const double h2 = 0.5 * Dt * Dt;
    double nx = qx + Dt * dx + h2 * ddx;
    double ny = qy + Dt * dy + h2 * ddy;
    double nz = qz + Dt * dz + h2 * ddz;
    double nw = qw + Dt * dw + h2 * ddw;
    double norm = sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
    const double s = 1.0 / norm;
    nx = s * nx; ny = s * ny; nz = s * nz; nw = s * nw;
    D4 po; po.x = pos.x; po.y = pos.y; po.z = pos.z; po.w = 0.0;
    D4 qo; qo.x = nx; qo.y = ny; qo.z = nz; qo.w = nw;
    D4 vo; vo.x = vel.x; vo.y = vel.y; vo.z = vel.z; vo.w = 0.0;
    D4 wo; wo.x = w.x; wo.y = w.y; wo.z = w.z; wo.w = 0.0;
    C.pq[2 * i] = po;
    C.pq[2 * i + 1] = qo;
    C.vw[2 * i] = vo;
    C.vw[2 * i + 1] = wo;
    // reduced-precision kinematics snapshot for the contact kernel; .w of the
    // first record flags a finite clump state (checked as in
    // src/dem_force_cube_wall_clump.cu:196-197)
    const bool ok = finite3(vel) && finite3(w) && isfinite(nx) && isfinite(ny) && isfinite(nz) && isfinite(nw);
    if (!ok) atomicOr(&halt[3], ST_CLUMP_BAD);  // k_contacts then checks every sphere exactly
    C.snap[2 * i] = mk4<R>((R)vel.x, (R)vel.y, (R)vel.z, ok ? R(1) : R(0));
    C.snap[2 * i + 1] = mk4<R>((R)w.x, (R)w.y, (R)w.z, R(0));
}

// ---------------------------------------------------------------------------
// K1: clump half-kick, drift and attitude (FP64). Thread 0 of block 0 also
// advances the penetrator and clears the step counters.
template <class R>
__global__ void k_clump_before(DevParams P, ClumpBuf<R> C, PeneState* pst, unsigned long long* counters,
                               int* halt, int n) {
    if (*halt < n) return;  // a previous step asked for a rebuild: replayed later
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) k1_thread0(P, pst, counters, halt, n);
    if (i >= P.Nc) return;
    const D4 pr = C.prop[i];
    const D4 Fr = C.ft[2 * i], Tr = C.ft[2 * i + 1];
    const D4 pc = C.pq[2 * i], qc = C.pq[2 * i + 1];
    const D4 vr = C.vw[2 * i], wr = C.vw[2 * i + 1];
    V3<double> F = xyz(Fr), T = xyz(Tr);
    scale_ft(P, pr, F, T);
    clump_drift<R>(P, C, i, pc, qc, xyz(vr), xyz(wr), F, T, halt);
}

// ---------------------------------------------------------------------------
// K2: sphere pose from the clump pose, cumulative displacement, rebuild flag.
template <class R>
__global__ void k_sphere_pose(DevParams P, SphereBuf<R> S, ClumpBuf<R> C, int* halt, int n_arg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool over = false, near = false;
    int n;
    if (ref_disp<R>(P.driver)) {
        // FP32-mixed clump loop: the loads that do not need the clump index
        // are issued with it, and ahead of the halt check (from a clamped
        // index; the arrays are valid whatever the halt), so a wave waits for
        // one gather level only
        const int ii = i < P.N ? i : P.N - 1;
        const int c = S.cid[ii];
        const typename Vec4T<R>::type lo = S.loc[ii];
        const D4 cu = S.cum[ii];  // x_rebuild, radius
        const float th = S.nthr[ii];
        n = step_of(P, n_arg);
        const int hv = *halt;
        __builtin_amdgcn_sched_barrier(0);  // keep these loads ahead of the waits
        if (hv < n) return;
        const D4 pc = C.pq[2 * c], q = C.pq[2 * c + 1];
        double M[3][3];
        QuatIvDCM<double>(q.x, q.y, q.z, q.w, M);
        const V3<double> arm = Map<double>(M, mk<double>((double)lo.x, (double)lo.y, (double)lo.z));
        const V3<double> up = xyz(pc) + arm;
        // displacement = x - x_rebuild
        const V3<double> cum = up - xyz(cu);
        over = Length<double>(cum) > P.Thresh;
        near = cum.x * cum.x + cum.y * cum.y + cum.z * cum.z > (double)th;
        D4 pn; pn.x = up.x; pn.y = up.y; pn.z = up.z; pn.w = cu.w;
        if (i < P.N) {
            S.pos[i] = pn;
            S.kin[i] = mk4<R>((R)arm.x, (R)arm.y, (R)arm.z, lo.w);
        } else {
            over = near = false;
        }
    } else {
        n = step_of(P, n_arg);
        if (*halt < n) return;
    }
    if (!ref_disp<R>(P.driver) && i < P.N) {
        const int c = S.cid[i];
        const typename Vec4T<R>::type lo = S.loc[i];
        const D4 cu = S.cum[i];
        const D4 po = S.pos[i];
        __builtin_amdgcn_sched_barrier(0);  // these loads ahead of the wait for c
        const D4 pc = C.pq[2 * c], q = C.pq[2 * c + 1];
        double M[3][3];
        QuatIvDCM<double>(q.x, q.y, q.z, q.w, M);
        V3<double> arm = Map<double>(M, mk<double>((double)lo.x, (double)lo.y, (double)lo.z));
        V3<double> up = xyz(pc) + arm;
        V3<double> cum;
        if (P.driver) {
            const D4 dd = C.dist[c];
            cum = xyz(cu) + xyz(dd);
        } else {
            cum = xyz(cu) + up - xyz(po);
        }
        // glue loop: type 0 rebuilds past 0.1 r, the kinematic types never ask
        // (src/dem_integrate_withglue.cu:40-43)
        if (P.driver == NBC_DRIVER_SPLIT) over = C.type[c] == 0 && Length<double>(cum) > 0.10 * po.w;
        else over = Length<double>(cum) > P.Thresh;
        // the sphere left its near-list allowance (threshold kept in cum.w)
        near = cum.x * cum.x + cum.y * cum.y + cum.z * cum.z > cu.w;
        D4 pn; pn.x = up.x; pn.y = up.y; pn.z = up.z; pn.w = po.w;
        D4 cn; cn.x = cum.x; cn.y = cum.y; cn.z = cum.z; cn.w = cu.w;
        S.pos[i] = pn;
        S.cum[i] = cn;
        S.kin[i] = mk4<R>((R)arm.x, (R)arm.y, (R)arm.z, lo.w);
    }
    // the rebuild request (src/dem_clump.cu:132-134) records the step index:
    // later kernels of this step and all kernels of later steps stand down
    const uint64_t b = __ballot(over);
    if (b != 0ull && lane_id() == (int)__ffsll((unsigned long long)b) - 1) {
        atomicOr(&halt[3], ST_REBUILD);
        atomicMin(halt, n);
    }
    const uint64_t bn = __ballot(near);
    if (b == 0ull && bn != 0ull && lane_id() == (int)__ffsll((unsigned long long)bn) - 1) {
        atomicOr(&halt[3], ST_NEAR);
        atomicMin(halt, n);
    }
}

// ---------------------------------------------------------------------------
// Grid hash, src/dem_grid.cu:155-202. The clamp of the reference runs in
// unsigned arithmetic (GridDim is unsigned), so a negative index maps to GridDim-1.
__device__ __forceinline__ int clamp_u(uint32_t G, int v) {
    const uint32_t u = (uint32_t)v, hi = G - 1u;
    return (int)(u < hi ? u : hi);
}

// Uniform-grid cell of p, src/dem_grid.cu:155-202 (out-of-box spheres go to
// the last cell).
__device__ __forceinline__ int uniform_cell(const double* O, const double* cs, const uint32_t* G, const double* lo,
                                            const double* up, int cell_num, const D4& p) {
    if (p.z >= up[2] || p.y >= up[1] || p.x >= up[0] || p.z < lo[2] || p.y < lo[1] || p.x < lo[0])
        return cell_num - 1;
    int cx = (int)floor((p.x - O[0]) / cs[0]);
    int cy = (int)floor((p.y - O[1]) / cs[1]);
    int cz = (int)floor((p.z - O[2]) / cs[2]);
    cx = clamp_u(G[0], cx);
    cy = clamp_u(G[1], cy);
    cz = clamp_u(G[2], cz);
    int idx = (int)((uint32_t)cz * G[1] * G[0] + (uint32_t)cy * G[0] + (uint32_t)cx);
    if (idx < 0) idx = 0;
    if (idx >= cell_num) idx = cell_num - 1;
    return idx;
}

// Level of a sphere in the hierarchical grid: the first t with r > lrmin[t]
// (the last level takes the rest).
__device__ __forceinline__ int level_of(const DevParams& P, double r) {
    int t = 0;
    while (t < P.L - 1 && !(r > P.lrmin[t])) ++t;
    return t;
}

// Cell coordinate of x on a level of cell size s (src/dem_hgrid.cu:225-227).
__device__ __forceinline__ int hcell(double x, double o, double s) { return (int)floor((x - o) / s); }

// Key of level-t cell (cx, cy, cz), inside the level's box. Level-major
// (src/dem_hgrid.cu:228): loff[t] + (cz ny + cy) nx + cx. Nested: the coarse
// cell (the coordinates shifted right by t) picks a block of lblock keys, the
// level picks loff[t] inside it and the low t bits of each coordinate the
// sub-cell, so the spheres of one coarse cell are contiguous after the sort.
__device__ __forceinline__ int hkey(const DevParams& P, int t, int cx, int cy, int cz) {
    if (!P.nested) return P.loff[t] + (cz * P.ldim[t][1] + cy) * P.ldim[t][0] + cx;
    const int m = (1 << t) - 1;
    const int coarse = ((cz >> t) * P.ldim[0][1] + (cy >> t)) * P.ldim[0][0] + (cx >> t);
    return coarse * P.lblock + P.loff[t] + ((((cz & m) << t) | (cy & m)) << t | (cx & m));
}

__global__ void k_hash(DevParams P, const D4* pos, const int* cid, uint32_t* key, int* val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    const D4 p = pos[i];
    val[i] = i;
    int idx;
    if (P.L > 1) {
        // hierarchical grid, src/dem_hgrid.cu:205-242; the out-of-box bucket
        // is the last cell (CellNum - 1) and each coordinate is clamped into
        // its level's box
        if (p.z > P.boxup[2] || p.y > P.boxup[1] || p.x > P.boxup[0] || p.z < P.boxlo[2] || p.y < P.boxlo[1] ||
            p.x < P.boxlo[0]) {
            idx = P.CellNum - 1;
        } else {
            const int t = P.clevel ? P.clevel[cid[i]] : level_of(P, p.w);
            const double s = P.lcell[t];
            const int* nd = P.ldim[t];
            int cx = hcell(p.x, P.origin[0], s), cy = hcell(p.y, P.origin[1], s), cz = hcell(p.z, P.origin[2], s);
            cx = cx < 0 ? 0 : (cx > nd[0] - 1 ? nd[0] - 1 : cx);
            cy = cy < 0 ? 0 : (cy > nd[1] - 1 ? nd[1] - 1 : cy);
            cz = cz < 0 ? 0 : (cz > nd[2] - 1 ? nd[2] - 1 : cz);
            idx = hkey(P, t, cx, cy, cz);
        }
        key[i] = (uint32_t)idx;
        return;
    }
    if (P.nbeds > 1) {
        // batch: the bed's own grid, numbered from its first cell
        const BedGeo& g = P.beds[P.bid[i]];
        idx = g.cell_off + uniform_cell(g.origin, g.cell, g.G, g.boxlo, g.boxup, g.cell_num, p);
    } else {
        idx = uniform_cell(P.origin, P.cell, P.G, P.boxlo, P.boxup, P.CellNum, p);
    }
    key[i] = (uint32_t)idx;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (uint32 key, int value), 256-thread tiles, one key
// per thread. Within a wave the lanes holding the same digit are found with
// one ballot per digit bit; ranks combine wave counts in thread order, so
// equal digits keep their input order (stable, like cub::DeviceRadixSort).
__device__ __forceinline__ uint64_t digit_match(int dig, bool valid, int rbits) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < rbits; ++b) {
        const bool bit = (dig >> b) & 1;
        const uint64_t bb = __ballot(valid && bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t* keys, int n, int shift, int rbits, int* hist, int nb) {
    __shared__ int cnt[4][256];
    const int t = threadIdx.x, wv = t >> 6;
    for (int k = t; k < 1024; k += 256) (&cnt[0][0])[k] = 0;
    __syncthreads();
    const int i = blockIdx.x * 256 + t;
    const bool valid = i < n;
    const int mask = (1 << rbits) - 1;
    const int dig = valid ? (int)((keys[i] >> shift) & (uint32_t)mask) : 0;
    const uint64_t m = digit_match(dig, valid, rbits);
    if (valid && (m & lanemask_lt()) == 0ull) cnt[wv][dig] = __popcll(m);
    __syncthreads();
    for (int d = t; d <= mask; d += 256) hist[d * nb + blockIdx.x] = cnt[0][d] + cnt[1][d] + cnt[2][d] + cnt[3][d];
}

__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t* kin, const int* vin, uint32_t* kout, int* vout,
                                                    int n, int shift, int rbits, const int* hscan, int nb) {
    __shared__ int cnt[4][256];
    const int t = threadIdx.x, wv = t >> 6;
    for (int k = t; k < 1024; k += 256) (&cnt[0][0])[k] = 0;
    __syncthreads();
    const int i = blockIdx.x * 256 + t;
    const bool valid = i < n;
    const int mask = (1 << rbits) - 1;
    uint32_t key = 0;
    int val = 0;
    if (valid) { key = kin[i]; val = vin[i]; }
    const int dig = (int)((key >> shift) & (uint32_t)mask);
    const uint64_t m = digit_match(dig, valid, rbits);
    if (valid && (m & lanemask_lt()) == 0ull) cnt[wv][dig] = __popcll(m);
    __syncthreads();
    for (int d = t; d <= mask; d += 256) {
        const int a = cnt[0][d], b = cnt[1][d], c = cnt[2][d];
        cnt[0][d] = 0; cnt[1][d] = a; cnt[2][d] = a + b; cnt[3][d] = a + b + c;
    }
    __syncthreads();
    if (valid) {
        const int pos = hscan[dig * nb + blockIdx.x] + cnt[wv][dig] + __popcll(m & lanemask_lt());
        kout[pos] = key;
        vout[pos] = val;
    }
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of int32 (tiles of 2048 = 256 threads x 8).
__device__ __forceinline__ int block_excl_scan256(int v, int* lds /*>= 4*/, int* total) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wv] = x;
    __syncthreads();
    int add = 0, tot = 0;
    for (int k = 0; k < 4; ++k) {
        const int s = lds[k];
        if (k < wv) add += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return x - v + add;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int* in, int n, int* bsum) {
    __shared__ int lds[4];
    const int base = blockIdx.x * 2048 + threadIdx.x * 8;
    int s = 0;
    for (int k = 0; k < 8; ++k) {
        const int i = base + k;
        if (i < n) s += in[i];
    }
    int tot;
    block_excl_scan256(s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_top(int* bsum, int nb, int* total_out) {
    __shared__ int lds[4];
    int carry = 0;
    for (int base = 0; base < nb; base += 256) {
        const int i = base + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        int tot;
        const int ex = block_excl_scan256(v, lds, &tot);
        if (i < nb) bsum[i] = ex + carry;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(256) void k_scan_down(const int* in, int* out, int n, const int* bsum) {
    __shared__ int lds[4];
    const int base = blockIdx.x * 2048 + threadIdx.x * 8;
    int v[8];
    int s = 0;
    for (int k = 0; k < 8; ++k) {
        const int i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_excl_scan256(s, lds, &tot) + bsum[blockIdx.x];
    for (int k = 0; k < 8; ++k) {
        const int i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
}

__global__ void k_keys_from_int(const int* a, uint32_t* k, int* v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    k[i] = (uint32_t)a[i];
    v[i] = i;
}

__global__ void k_fill_i32(int* a, long long n, int v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_u32(uint32_t* a, long long n, uint32_t v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_f64(double* a, long long n, double v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Segment bounds of a sorted key array (src/dem_grid.cu:105-153): beg/end of
// each key value, -1 where absent (caller fills -1 first).
__global__ void k_seg_bounds(const uint32_t* key, int n, int nkeys, int* beg, int* end) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)key[i];
    if (c < 0 || c >= nkeys) return;
    if (i == 0 || (int)key[i - 1] != c) beg[c] = i;
    if (i == n - 1 || (int)key[i + 1] != c) end[c] = i + 1;
}

// Gather of every per-sphere array into the new sorted order
// (src/dem_simulation_clump.cu:126-164); the displacement restarts at zero
// (src/dem_simulation_clump.cu:172).
template <class R>
__global__ void k_gather(DevParams P, int nface, const int* perm, SphereBuf<R> src, SphereBuf<R> dst) {
    const int N = P.N;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int o = perm[i];
    const D4 po = src.pos[o];
    dst.pos[i] = po;
    D4 z; z.x = 0.0; z.y = 0.0; z.z = 0.0; z.w = 0.0;
    dst.cum[i] = ref_disp<R>(P.driver) ? po : z;
    dst.loc[i] = src.loc[o];
    dst.kin[i] = src.kin[o];
    dst.cid[i] = src.cid[o];
    dst.fid[i] = src.fid[o];
    const uint32_t wf = src.wflag[o];
    dst.wflag[i] = wf;
    for (int f = 0; f < nface; ++f) {
        if (!((wf >> f) & 1u)) continue;
        const R* a = src.wh + ((size_t)f * N + o) * HSTRIDE;
        R* b = dst.wh + ((size_t)f * N + i) * HSTRIDE;
        for (int k = 0; k < HSTRIDE; ++k) b[k] = a[k];
    }
}

// Clump renumbering at a rebuild (no counterpart in the reference, whose
// clump arrays stay in file order): clumps take the order of their first
// sphere in the new sphere order, so the clump data of neighbouring spheres
// sits in neighbouring lines. first[c] = smallest sorted index of c's spheres.
__global__ void k_clump_first(int N, const int* cid, int* first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) atomicMin(&first[cid[i]], i);
}

// New clump c takes old clump order[c]: every clump array moves, newid maps
// old to new, cfile follows (file index of each clump).
template <class R>
__global__ void k_clump_gather(int Nc, const int* order, ClumpBuf<R> src, ClumpBuf<R> dst, const int* cfile,
                               int* cfile_out, int* newid) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= Nc) return;
    const int o = order[c];
    dst.pq[2 * c] = src.pq[2 * o];
    dst.pq[2 * c + 1] = src.pq[2 * o + 1];
    dst.vw[2 * c] = src.vw[2 * o];
    dst.vw[2 * c + 1] = src.vw[2 * o + 1];
    dst.prop[c] = src.prop[o];
    dst.ft[2 * c] = src.ft[2 * o];
    dst.ft[2 * c + 1] = src.ft[2 * o + 1];
    dst.snap[2 * c] = src.snap[2 * o];
    dst.snap[2 * c + 1] = src.snap[2 * o + 1];
    if (src.dist) dst.dist[c] = src.dist[o];
    if (src.type) dst.type[c] = src.type[o];
    cfile_out[c] = cfile[o];
    newid[o] = c;
}

__global__ void k_remap_cid(int N, int* cid, const int* newid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) cid[i] = newid[cid[i]];
}

// ---------------------------------------------------------------------------
// Neighbour search, src/dem_grid.cu:466-552: 27 cells (z outer, x inner),
// ascending j inside a cell, cut-off 1.5 (ri + rj). The search cell uses the
// reciprocal cell size (not the hash's division), unclamped. FILL = false
// counts, FILL = true writes the CSR rows: front j > i in discovery order,
// back j < i in reverse discovery order (the gather order of
// src/dem_force_clump.cu:183-188).
// Neighbour search over the 27 cells of (cix, ciy, ciz), z outer, x inner
// (src/dem_grid.cu:466-552). The three cells of one (z, y) row have
// consecutive keys, so their spheres are one contiguous run of the sorted
// order: each row is walked as the run [s0, s1) in ascending index, which is
// the reference's cell-by-cell order, with four candidate positions loaded
// ahead. visit(j, pj) is called for every j != i within 1.5 (ri + rj).
// The list passes work like the contact kernel: a wave takes 64 consecutive
// spheres, whose rows form one contiguous segment of the CSR list, and reads
// it 64 entries at a time (coalesced); an entry finds its row by binary search
// over the 65 row offsets kept in LDS.
__device__ __forceinline__ int row_of(const int* fo, int e) {
    int r = 0;
#pragma unroll
    for (int st = 32; st > 0; st >>= 1)
        if (fo[r + st] <= e) r += st;
    return r;
}

// Loads the row offsets of the wave's spheres i0 .. i0+63 into fo[0..64].
__device__ __forceinline__ void load_rows(int* fo, const int* off, int i0, int N) {
    const int lane = lane_id();
    fo[lane] = off[min(i0 + lane, N)];
    if (lane == 0) fo[64] = off[min(i0 + 64, N)];
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Index of partner j in the front row [t0, t1), or -1: eight words per round
// in flight instead of one dependent load per entry (pairs are listed once,
// so any match is the only one).
__device__ __forceinline__ int find_in_row(const uint32_t* fword, int t0, int t1, int j) {
    for (int t = t0; t < t1; t += 8) {
        uint32_t w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = fword[min(t + u, t1 - 1)];
        int hit = -1;
#pragma unroll
        for (int u = 7; u >= 0; --u)
            if ((int)(w[u] & JMASK) == j) hit = t + u;
        if (hit >= 0) return hit < t1 ? hit : -1;
    }
    return -1;
}

template <class Visit>
__device__ __forceinline__ void search_27(const D4* pos, const int* cbeg, const int* cend, const uint32_t* G,
                                          int coff, int i, const D4& pi, int cix, int ciy, int ciz, Visit visit) {
    const double ri = pi.w;
    const int jx0 = max(cix - 1, 0), jx1 = min(cix + 1, (int)G[0] - 1);
    if (jx0 > jx1) return;
    for (int cz = -1; cz <= 1; cz++)
        for (int cy = -1; cy <= 1; cy++) {
            const int jy = ciy + cy, jz = ciz + cz;
            if (jy < 0 || (uint32_t)jy >= G[1] || jz < 0 || (uint32_t)jz >= G[2]) continue;
            const int rowc = coff + (int)(((uint32_t)jz * G[1] + (uint32_t)jy) * G[0]);
            int s0 = -1, s1 = -1;
            for (int jx = jx0; jx <= jx1; ++jx) {
                const int b = cbeg[rowc + jx];
                if (b == -1) continue;
                if (s0 < 0) s0 = b;
                s1 = cend[rowc + jx];
            }
            if (s0 < 0) continue;
            for (int j0 = s0; j0 < s1; j0 += 4) {
                D4 q[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) q[u] = pos[min(j0 + u, s1 - 1)];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = j0 + u;
                    if (j >= s1 || j == i) continue;
                    const D4 pj = q[u];
                    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    const double ds = dx * dx + dy * dy + dz * dz;
                    const double lim = 1.50 * (ri + pj.w);
                    if (ds < lim * lim) visit(j);
                }
            }
        }
}

template <bool FILL>
__global__ void k_neighbors(DevParams P, const D4* pos, const uint32_t* cell, const int* cbeg, const int* cend,
                            const int* cid, int* fcnt, int* bcnt, const int* foff, const int* boff, uint32_t* fword,
                            int2* bent, unsigned* overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    const int half = P.B / 2;
    // grid of the sphere's bed (batch) or the single grid
    const double* O = P.origin;
    const double* ic = P.icell;
    const uint32_t* G = P.G;
    int coff = 0, last = P.CellNum - 1;
    if (P.nbeds > 1) {
        const BedGeo& g = P.beds[P.bid[i]];
        O = g.origin; ic = g.icell; G = g.G;
        coff = g.cell_off;
        last = g.cell_off + g.cell_num - 1;
    }
    if ((int)cell[i] == last) {
        if (!FILL) { fcnt[i] = 0; bcnt[i] = 0; }
        return;
    }
    const D4 pi = pos[i];
    const int cix = (int)floor((pi.x - O[0]) * ic[0]);
    const int ciy = (int)floor((pi.y - O[1]) * ic[1]);
    const int ciz = (int)floor((pi.z - O[2]) * ic[2]);
    int ng = 0, ni = 0;
    int fb = 0, bb = 0, nbi = 0, ci = 0;
    if (FILL) { fb = foff[i]; bb = boff[i]; nbi = bcnt[i]; ci = cid[i]; }
    search_27(pos, cbeg, cend, G, coff, i, pi, cix, ciy, ciz, [&](int j) {
        if (i < j) {
            if (FILL && ng < half) fword[fb + ng] = (uint32_t)j | (cid[j] == ci ? INTRA : 0u);
            ng++;
        } else {
            if (FILL && ni < half) bent[bb + (nbi - 1 - ni)] = make_int2(j, -1);
            ni++;
        }
    });
    if (!FILL) {
        fcnt[i] = ng;
        bcnt[i] = ni;
        if (ng > half || ni > half) atomicOr(overflow, 1u);
    }
}

// Hierarchical search, src/dem_hgrid_contlist.cu:57-161 with deterministic
// rows. A sphere of level t scans levels t..0 (27 cells each, z outer / x
// inner, ascending j, cut-off 1.5 (ri + rj)). Same level: j > i front, j < i
// back. Coarser level: the coarser sphere j owns the pair (with level-major
// numbering it sorts first; with nested numbering it may sort after i and
// owns the pair all the same); i lists j in its back row and emits the cross
// entry (j, i), which becomes a front entry of j. PH 0 counts (same-level fronts,
// backs, cross entries), PH 1 writes the cross entries at xoff[i], PH 2 writes
// the rows: same-level fronts in discovery order, then the cross partners
// xval[xbeg[i] .. xend[i]) (ascending i after the stable sort by j); backs in
// reverse discovery order.
template <int PH>
__global__ void k_hneighbors(DevParams P, const D4* pos, const uint32_t* cell, const int* cbeg, const int* cend,
                             const int* cid, int* fcnt, int* bcnt, int* xcnt, const int* xoff, uint32_t* xkey,
                             int* xval, const int* foff, const int* boff, const int* xbeg, const int* xend,
                             const int* xsrt, uint32_t* fword, int2* bent, unsigned* overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    if ((int)cell[i] == P.CellNum - 1) {
        if (PH == 0) { fcnt[i] = 0; bcnt[i] = 0; xcnt[i] = 0; }
        return;
    }
    const D4 pi = pos[i];
    const double ri = pi.w;
    const int my = P.clevel ? P.clevel[cid[i]] : level_of(P, ri);
    int ng = 0, ni = 0, nx = 0;
    int fb = 0, nf = 0, bb = 0, nbi = 0, xb = 0, xn = 0, ci = 0;
    if (PH == 1) { xb = xoff[i]; xn = xcnt[i]; }
    if (PH == 2) { fb = foff[i]; nf = foff[i + 1] - fb; bb = boff[i]; nbi = bcnt[i]; ci = cid[i]; }
    for (int T = my; T >= 0; --T) {
        const double s = P.lcell[T];
        const int nxd = P.ldim[T][0], nyd = P.ldim[T][1], nzd = P.ldim[T][2];
        const int cix = hcell(pi.x, P.origin[0], s), ciy = hcell(pi.y, P.origin[1], s);
        const int ciz = hcell(pi.z, P.origin[2], s);
        for (int cz = -1; cz <= 1; cz++)
            for (int cy = -1; cy <= 1; cy++)
                for (int cx = -1; cx <= 1; cx++) {
                    const int jx = cix + cx, jy = ciy + cy, jz = ciz + cz;
                    if (jx < 0 || jx >= nxd || jy < 0 || jy >= nyd || jz < 0 || jz >= nzd) continue;
                    const int cidx = hkey(P, T, jx, jy, jz);
                    const int start = cbeg[cidx];
                    if (start == -1) continue;
                    const int end = cend[cidx];
                    for (int j = start; j < end; j++) {
                        if (j == i) continue;
                        const D4 pj = pos[j];
                        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                        const double ds = dx * dx + dy * dy + dz * dz;
                        const double lim = 1.5 * (ri + pj.w);
                        if (!(ds < lim * lim)) continue;
                        if (T == my && i < j) {
                            if (PH == 2 && ng < nf) fword[fb + ng] = (uint32_t)j | (cid[j] == ci ? INTRA : 0u);
                            ng++;
                        } else {
                            if (PH == 2 && ni < nbi) bent[bb + (nbi - 1 - ni)] = make_int2(j, -1);
                            ni++;
                            if (T != my) {
                                if (PH == 1 && nx < xn) { xkey[xb + nx] = (uint32_t)j; xval[xb + nx] = i; }
                                nx++;
                            }
                        }
                    }
                }
    }
    if (PH == 0) {
        fcnt[i] = ng;
        bcnt[i] = ni;
        xcnt[i] = nx;
    }
    if (PH == 2) {
        const int b = xbeg[i];
        if (b >= 0) {
            const int e = xend[i];
            for (int k = b; k < e && ng < nf; ++k, ++ng) {
                const int j = xsrt[k];
                fword[fb + ng] = (uint32_t)j | (cid[j] == ci ? INTRA : 0u);
            }
        }
    }
}

// Front count = same-level fronts + cross entries owned by i; both halves of
// a row hold at most BranchLmtLen / 2 entries.
__global__ void k_hfront_total(int N, int half, int* fcnt, const int* bcnt, const int* xbeg, const int* xend,
                               unsigned* overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int b = xbeg[i];
    const int f = fcnt[i] + (b >= 0 ? xend[i] - b : 0);
    fcnt[i] = f;
    if (f > half || bcnt[i] > half) atomicOr(overflow, 1u);
}

// ---------------------------------------------------------------------------
// Single-pass neighbour search. The count pass of k_neighbors / k_hneighbors
// also stores each row into a fixed-capacity table kept entry-major (entry g
// of thread k at g * N + k, so a wave's lanes write neighbouring words); once
// the counts are scanned, a compaction kernel moves the rows into the CSR
// arrays in the order the fill pass would write them, so the second (and, on
// the hierarchical grid, third) search pass is not needed. A row longer than
// K sets bit 2 of *overflow: the counts stay exact, only the stores stop, and
// the host then fills the rows with the multi-pass kernels. ellmax receives
// the longest row (front or back), from which the host sizes K for the next
// rebuild.
// Same-address device atomics are serialised at the memory side (~13 ns each
// on this part), so a wave only issues its atomicMax when its value exceeds
// the maximum recorded so far (a device-coherent load; a stale read only
// costs an extra atomic).
__device__ __forceinline__ void wave_max_to(unsigned* dst, int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0 && v > 0 &&
        (unsigned)v > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(dst, (unsigned)v);
}

// Single-pass uniform-grid search (src/dem_grid.cu:466-552, the loop of
// k_neighbors) into entry-major tables of capacity K: fronts as pair words,
// backs in discovery order (reversed by the compaction). One wave per cell:
// the spheres of a cell share its 27-cell stencil, whose candidates (nine
// x-runs of three cells, each run contiguous in sorted order, the runs
// ascending) the wave loads 64 at a time, one per lane. Each home sphere then
// tests the 64 loaded candidates at once, its position broadcast from its own
// lane, and keeps the hits as a 64-bit mask; its lane writes them in
// candidate order, the order in which the per-sphere loop visits them. A
// per-sphere loop reloads every candidate once per home sphere (about 450
// loads of 32 bytes per sphere on a settled bed); here a candidate is loaded
// once per cell. A sphere whose cell coordinates, computed from its position
// as the per-sphere loop does, differ from those of its hashed cell runs the
// per-sphere loop (search_27) instead.
__device__ __forceinline__ double readlane_f64(double v, int k) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__global__ __launch_bounds__(256) void k_neighbors_cell(DevParams P, const D4* pos, const int* cbeg, const int* cend,
                                                        const int* cid, int K, int* fcnt, int* bcnt, uint32_t* ef,
                                                        int* eb, unsigned* overflow, unsigned* ellmax) {
    __shared__ int s_j[4][64];
    __shared__ int s_c[4][64];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int h = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + wv));
    if (h >= P.CellNum) return;
    const int hb = cbeg[h];
    if (hb < 0) return;
    const int he = cend[h];
    const int half = P.B / 2;
    const double* O = P.origin;
    const double* ic = P.icell;
    const uint32_t* G = P.G;
    int coff = 0, last = P.CellNum - 1;
    if (P.nbeds > 1) {
        const BedGeo& g = P.beds[P.bid[hb]];
        O = g.origin; ic = g.icell; G = g.G;
        coff = g.cell_off;
        last = g.cell_off + g.cell_num - 1;
    }
    if (h == last) {
        // out-of-box bucket: never searched
        for (int i = hb + lane; i < he; i += 64) { fcnt[i] = 0; bcnt[i] = 0; }
        return;
    }
    const int G0 = (int)G[0], G1 = (int)G[1], G2 = (int)G[2];
    const int l = h - coff;
    const int hx = l % G0, hy = (l / G0) % G1, hz = l / (G0 * G1);
    // stencil row r = 3 (dz + 1) + (dy + 1): its run [s0, s1) over the cells
    // hx-1 .. hx+1 that exist
    int rs = 0, rl = 0;
    if (lane < 9) {
        const int jy = hy + lane % 3 - 1, jz = hz + lane / 3 - 1;
        const int jx0 = max(hx - 1, 0), jx1 = min(hx + 1, G0 - 1);
        if (jy >= 0 && jy < G1 && jz >= 0 && jz < G2) {
            const int rowc = coff + (jz * G1 + jy) * G0;
            int s0 = -1, s1 = -1;
            for (int jx = jx0; jx <= jx1; ++jx) {
                const int b = cbeg[rowc + jx];
                if (b == -1) continue;
                if (s0 < 0) s0 = b;
                s1 = cend[rowc + jx];
            }
            if (s0 >= 0) { rs = s0; rl = s1 - s0; }
        }
    }
    int rst[9], roff[10];
    roff[0] = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        rst[r] = __builtin_amdgcn_readlane(rs, r);
        roff[r + 1] = roff[r] + __builtin_amdgcn_readlane(rl, r);
    }
    const int M = roff[9];
    int gmax = 0;
    for (int g0 = hb; g0 < he; g0 += 64) {
        const int i = g0 + lane;
        const bool act = i < he;
        const D4 pi = pos[act ? i : hb];
        const int ci = cid[act ? i : hb];
        const int cix = (int)floor((pi.x - O[0]) * ic[0]);
        const int ciy = (int)floor((pi.y - O[1]) * ic[1]);
        const int ciz = (int)floor((pi.z - O[2]) * ic[2]);
        const bool reg = act && cix == hx && ciy == hy && ciz == hz;
        const uint64_t regm = __ballot(reg);
        int ng = 0, ni = 0;
        for (int c0 = 0; c0 < M && regm; c0 += 64) {
            const int t = c0 + lane;
            const bool cv = t < M;
            int j = hb;
#pragma unroll
            for (int r = 0; r < 9; ++r)
                if (t >= roff[r] && t < roff[r + 1]) j = rst[r] + (t - roff[r]);
            const D4 pj = pos[j];
            const int cj = cid[j];
            uint64_t mym = 0;
            // home spheres in turn, each broadcast from its own lane
            for (uint64_t rm = regm; rm; rm &= rm - 1ull) {
                const int k = (int)__builtin_ctzll(rm);
                const double kx = readlane_f64(pi.x, k), ky = readlane_f64(pi.y, k);
                const double kz = readlane_f64(pi.z, k), kr = readlane_f64(pi.w, k);
                const double dx = kx - pj.x, dy = ky - pj.y, dz = kz - pj.z;
                const double ds = dx * dx + dy * dy + dz * dz;
                const double lim = 1.50 * (kr + pj.w);
                const uint64_t m = __ballot(cv && j != g0 + k && ds < lim * lim);
                if (lane == k) mym = m;
            }
            s_j[wv][lane] = j;
            s_c[wv][lane] = cj;
            wave_sync();
            while (mym) {
                const int b = __builtin_ctzll(mym);
                mym &= mym - 1ull;
                const int jj = s_j[wv][b];
                if (jj > i) {
                    if (ng < K) ef[(size_t)ng * P.N + i] = (uint32_t)jj | (s_c[wv][b] == ci ? INTRA : 0u);
                    ng++;
                } else {
                    if (ni < K) eb[(size_t)ni * P.N + i] = jj;
                    ni++;
                }
            }
            wave_sync();
        }
        if (act && !reg) {
            int ng2 = 0, ni2 = 0;
            search_27(pos, cbeg, cend, G, coff, i, pi, cix, ciy, ciz, [&](int j) {
                if (i < j) {
                    if (ng2 < K) ef[(size_t)ng2 * P.N + i] = (uint32_t)j | (cid[j] == ci ? INTRA : 0u);
                    ng2++;
                } else {
                    if (ni2 < K) eb[(size_t)ni2 * P.N + i] = j;
                    ni2++;
                }
            });
            ng = ng2;
            ni = ni2;
        }
        if (act) {
            fcnt[i] = ng;
            bcnt[i] = ni;
            if (ng > half || ni > half) atomicOr(overflow, 1u);
            if (ng > K || ni > K) atomicOr(overflow, 2u);
            gmax = max(gmax, max(ng, ni));
        }
    }
    wave_max_to(ellmax, gmax);
}

// A wave takes 64 consecutive spheres, whose front rows (and back rows) form
// one contiguous CSR segment, and writes it 64 entries at a time (coalesced);
// an entry finds its row by binary search over the 65 row offsets in LDS.
// Back rows hold the owners in descending order (the search found them
// ascending).
__global__ __launch_bounds__(256) void k_rows_from_ell(int N, const int* foff, const int* boff, const uint32_t* ef,
                                                       const int* eb, uint32_t* fword, int2* bent) {
    __shared__ int s_fo[4][65];
    __shared__ int s_bo[4][65];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i0 = (blockIdx.x * 4 + wv) * 64;
    if (i0 >= N) return;
    load_rows(s_fo[wv], foff, i0, N);
    load_rows(s_bo[wv], boff, i0, N);
    const int fend = s_fo[wv][64];
    for (int k0 = s_fo[wv][0]; k0 < fend; k0 += 64) {
        const int e = k0 + lane;
        if (e >= fend) continue;
        const int r = row_of(s_fo[wv], e);
        fword[e] = ef[(size_t)(e - s_fo[wv][r]) * N + i0 + r];
    }
    const int bend = s_bo[wv][64];
    for (int k0 = s_bo[wv][0]; k0 < bend; k0 += 64) {
        const int e = k0 + lane;
        if (e >= bend) continue;
        const int r = row_of(s_bo[wv], e);
        const int g = s_bo[wv][r + 1] - 1 - e;
        bent[e] = make_int2(eb[(size_t)g * N + i0 + r], -1);
    }
}

// Hierarchical grid: the loop of k_hneighbors (phase 0) storing own fronts,
// backs and cross entries; thread k takes sphere sidx[k] (the spheres of one
// level together, so that a wave's lanes run the same number of levels), or
// sphere k without sidx.
__global__ void k_hneighbors_ell(DevParams P, const D4* pos, const uint32_t* cell, const int* cbeg, const int* cend,
                                 const int* cid, const int* sidx, int K, int* fcnt, int* bcnt, int* xcnt, uint32_t* ef,
                                 int* eb, int* ex, unsigned* overflow, unsigned* ellmax) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int ng = 0, ni = 0;
    if (k < P.N) {
        const int i = sidx ? sidx[k] : k;
        int nx = 0;
        if ((int)cell[i] != P.CellNum - 1) {
            const D4 pi = pos[i];
            const double ri = pi.w;
            const int ci = cid[i];
            const int my = P.clevel ? P.clevel[ci] : level_of(P, ri);
            for (int T = my; T >= 0; --T) {
                const double s = P.lcell[T];
                const int nxd = P.ldim[T][0], nyd = P.ldim[T][1], nzd = P.ldim[T][2];
                const int cix = hcell(pi.x, P.origin[0], s), ciy = hcell(pi.y, P.origin[1], s);
                const int ciz = hcell(pi.z, P.origin[2], s);
                for (int cz = -1; cz <= 1; cz++)
                    for (int cy = -1; cy <= 1; cy++)
                        for (int cx = -1; cx <= 1; cx++) {
                            const int jx = cix + cx, jy = ciy + cy, jz = ciz + cz;
                            if (jx < 0 || jx >= nxd || jy < 0 || jy >= nyd || jz < 0 || jz >= nzd) continue;
                            const int cidx = hkey(P, T, jx, jy, jz);
                            const int start = cbeg[cidx];
                            if (start == -1) continue;
                            const int end = cend[cidx];
                            for (int j = start; j < end; j++) {
                                if (j == i) continue;
                                const D4 pj = pos[j];
                                const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                                const double ds = dx * dx + dy * dy + dz * dz;
                                const double lim = 1.5 * (ri + pj.w);
                                if (!(ds < lim * lim)) continue;
                                if (T == my && i < j) {
                                    if (ng < K) ef[(size_t)ng * P.N + k] = (uint32_t)j | (cid[j] == ci ? INTRA : 0u);
                                    ng++;
                                } else {
                                    if (ni < K) eb[(size_t)ni * P.N + k] = j;
                                    ni++;
                                    if (T != my) {
                                        if (nx < K) ex[(size_t)nx * P.N + k] = j;
                                        nx++;
                                    }
                                }
                            }
                        }
            }
        }
        fcnt[i] = ng;
        bcnt[i] = ni;
        xcnt[i] = nx;
        if (ng > K || ni > K) atomicOr(overflow, 2u);
    }
    wave_max_to(ellmax, max(ng, ni));
}

// cross entries (j, i) at xoff[i], in discovery order (as phase 1 writes them)
__global__ void k_cross_from_ell(int N, const int* sidx, const int* xcnt, const int* xoff, const int* ex,
                                 uint32_t* xkey, int* xval) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = sidx ? sidx[k] : k;
    const int n = xcnt[i], o = xoff[i];
    for (int g = 0; g < n; ++g) {
        xkey[o + g] = (uint32_t)ex[(size_t)g * N + k];
        xval[o + g] = i;
    }
}

// rows as phase 2 writes them: own fronts, then the cross partners
// xsrt[xbeg[i] .. xend[i]); backs in reverse discovery order
__global__ void k_hrows_from_ell(int N, const int* sidx, const int* cid, const int* foff, const int* bcnt,
                                 const int* boff, const int* xbeg, const int* xend, const int* xsrt, const uint32_t* ef,
                                 const int* eb, uint32_t* fword, int2* bent) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int i = sidx ? sidx[k] : k;
    const int fo = foff[i], nf = foff[i + 1] - fo;
    const int b = xbeg[i];
    const int nxc = b >= 0 ? xend[i] - b : 0;
    const int own = nf - nxc;
    for (int g = 0; g < own; ++g) fword[fo + g] = ef[(size_t)g * N + k];
    const int ci = cid[i];
    for (int q = 0; q < nxc; ++q) {
        const int j = xsrt[b + q];
        fword[fo + own + q] = (uint32_t)j | (cid[j] == ci ? INTRA : 0u);
    }
    const int nb = bcnt[i], bo = boff[i];
    for (int g = 0; g < nb; ++g) bent[bo + nb - 1 - g] = make_int2(eb[(size_t)g * N + k], -1);
}

// level of every sphere as a sort key (value: the sphere), for the
// level-grouped thread order of the nested hierarchical search
__global__ void k_level_keys(DevParams P, const D4* pos, const int* cid, uint32_t* key, int* val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    key[i] = (uint32_t)(P.clevel ? P.clevel[cid[i]] : level_of(P, pos[i].w));
    val[i] = i;
}

// Owner slot of every back entry: src/dem_grid.cu:591-599 (Aref of the back
// rows), -1 when the owner's own search did not list the pair. A wave takes
// the back rows of 64 consecutive spheres (one contiguous segment), one entry
// per lane, and each lane scans its owner's front row (pairs are listed once,
// so the first match is the only one).
__global__ __launch_bounds__(256) void k_back_resolve(int N, const int* foff, const uint32_t* fword, const int* boff,
                                                      int2* bent);

// History after a rebuild in DEMmain's cylinder and glue loops
// (src/dem_simulation_cylind.cu:130-141): CreateContList inherits the
// histories into the array that the unconditional ForceExchange then swaps
// out and clears, so new slot s reads the history written at old slot s in
// the last step (zero where that pair was not touching or s is past the old list).
template <class R>
__global__ void k_hist_carry(long long nnew, long long nold, int p, uint32_t* nfword, R* nhist,
                             const uint32_t* ofword, const R* ohist) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnew) return;
    uint32_t w = nfword[s] & ~(VALID(0) | VALID(1));
    if (s < nold && (ofword[s] & VALID(p)) && !(w & INTRA)) {
        const R* h = ohist + (size_t)s * HSTRIDE;
        R* d = nhist + (size_t)s * HSTRIDE;
        for (int k = 0; k < HSTRIDE; ++k) d[k] = h[k];
        w |= VALID(p);
    }
    nfword[s] = w;
}

// ---------------------------------------------------------------------------
// Near lists: the contact kernel's view of the Verlet list. Each sphere gets a
// slack at the near-list build: with a = |cum| (its displacement since the
// last rebuild) and delta its near skin (fixed, or the distance its clump's
// speed bound covers in a set number of steps), slack = 2a + delta, and k_sphere_pose
// asks for a new near list (ST_NEAR) once |cum| exceeds a + delta, so the
// sphere has moved at most |cum| + a <= slack since the build. When
// Thresh + a is smaller, the slack is Thresh + a and no request is needed
// (|cum| <= Thresh until the next rebuild). A pair whose gap at the build is at
// least the sum of the two slacks therefore cannot touch before the next near
// build, and an intra-clump pair never does. The near list keeps the other
// front entries, in slot order, as (slot, pair word); between builds it holds
// the history-valid bits of its pairs (k_near_sync writes them back into the
// pair words). The back near list keeps, in back-row order, (owner, index of
// the pair in the owner's near row). With Thresh not a finite non-negative
// number every slack is infinite: every pair is near.
__device__ __forceinline__ bool near_pair(const D4& a, const D4& b, uint32_t w, double sa, double sb) {
    if (w & INTRA) return false;
    if (w & (VALID(0) | VALID(1))) return true;
    const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const double reach = (a.w + b.w + sa + sb) * (1.0 + 1e-9);
    return !(dx * dx + dy * dy + dz * dz >= reach * reach);
}

// Slack of every sphere and its near-list request threshold (|cum|^2 above
// it asks for a new near list), written into cum.w, or with nthr given
// (cum = position at the last rebuild) into nthr rounded down, so that the
// request comes no later than with the exact threshold.
__global__ void k_near_slack(DevParams P, const D4* pos, D4* cum, float* nthr, double* slack, const int* cid,
                             const D4* pq, const D4* vw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    D4 c = cum[i];
    const D4 p = pos[i];
    if (nthr) {
        c.x = p.x - c.x; c.y = p.y - c.y; c.z = p.z - c.z;
    }
    const double a = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
    double d = P.near_delta;
    if (P.near_horizon > 0.0) {
        // the distance the clump's speed bound |v| + |w| |x - x_clump| covers
        // in near_horizon steps; any d >= 0 keeps the slack and the request
        // threshold consistent, so this only trades list length for refreshes
        const int k = cid[i];
        const D4 v = vw[2 * k], w = vw[2 * k + 1], o = pq[2 * k];
        const double rx = p.x - o.x, ry = p.y - o.y, rz = p.z - o.z;
        const double sp = sqrt(v.x * v.x + v.y * v.y + v.z * v.z) +
                          sqrt(w.x * w.x + w.y * w.y + w.z * w.z) * sqrt(rx * rx + ry * ry + rz * rz);
        d = sp * P.dt * P.near_horizon;
        d = d > P.near_lo ? (d < P.near_hi ? d : P.near_hi) : P.near_lo;
        if (!(sp * 0.0 == 0.0)) d = P.near_hi;  // speed bound not finite
    }
    double sl, trig;
    if (!(a == a) || !(d < 1e300)) {
        sl = __longlong_as_double(0x7ff0000000000000ll);  // no bound: every pair is near
        trig = sl;
    } else if (2.0 * a + d <= P.Thresh + a) {
        sl = (2.0 * a + d) * (1.0 + 1e-9) + 1e-15;
        const double t = (a + d) * (1.0 - 1e-9);
        trig = t * t;
    } else {
        sl = (P.Thresh + a) * (1.0 + 1e-9) + 1e-15;
        trig = __longlong_as_double(0x7ff0000000000000ll);
    }
    slack[i] = sl;
    if (nthr) {
        nthr[i] = __double2float_rd(trig);
    } else {
        c.w = trig;
        cum[i] = c;
    }
}

__global__ __launch_bounds__(256) void k_back_resolve(int N, const int* foff, const uint32_t* fword, const int* boff,
                                                      int2* bent) {
    __shared__ int s_fo[4][65];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i0 = (blockIdx.x * 4 + wv) * 64;
    if (i0 >= N) return;
    load_rows(s_fo[wv], boff, i0, N);
    const int end = s_fo[wv][64];
    for (int k0 = s_fo[wv][0]; k0 < end; k0 += 64) {
        const int e = k0 + lane;
        if (e >= end) continue;
        const int i = i0 + row_of(s_fo[wv], e);
        const int o = bent[e].x;
        const int slot = find_in_row(fword, foff[o], foff[o + 1], i);
        int2 be;
        be.x = (slot >= 0 && (fword[slot] & INTRA)) ? (o | BINTRA) : o;
        be.y = slot;
        bent[e] = be;
    }
}

// History inheritance, src/dem_grid.cu:603-647: a new pair takes the history
// (buffer p) of the same pair in the old list, negated when its orientation
// flipped; unmatched pairs start from zero. The reference pulls: every new
// pair searches the old row of its lower old index. Here the old pairs push:
// the old near list holds every pair with a valid history (a valid pair is
// always near) and, between builds, its latest history in near order, so each
// old near pair (a, b) valid in buffer p finds its pair in the new list (new
// indices inv[a], inv[b]; owner the lower one, or with nested grid levels the
// coarser sphere, which owned it before too and keeps it whatever the
// indices), copies the history there and sets the valid bit. The same pairs
// inherit the same values, but only the ~0.3 touching pairs per sphere search
// a row instead of the ~12 new pairs per sphere. The new pair words must come
// without valid bits (as the neighbour search writes them); inv is the
// inverse of perm (new index -> old index).
template <class R>
__global__ void k_hist_push(DevParams P, const D4* pos, const int* cid, int p, const int* inv, const int* onfoff,
                            const uint32_t* onfw, const R* ohn, const int* nfoff, uint32_t* nfword, R* nhist) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.N) return;
    const int oe = onfoff[a + 1];
    for (int o = onfoff[a]; o < oe; ++o) {
        const uint32_t w = onfw[o];
        if (!(w & VALID(p))) continue;
        const int na = inv[a], nb = inv[(int)(w & JMASK)];
        bool keep = na < nb;
        if (!keep && P.nested) {
            const int la = P.clevel ? P.clevel[cid[na]] : level_of(P, pos[na].w);
            keep = la != (P.clevel ? P.clevel[cid[nb]] : level_of(P, pos[nb].w));
        }
        const int i = keep ? na : nb, j = keep ? nb : na;
        const int s = find_in_row(nfword, nfoff[i], nfoff[i + 1], j);
        if (s < 0) continue;
        const uint32_t nw = nfword[s];
        if (nw & INTRA) continue;
        const R* h = ohn + (size_t)o * HSTRIDE;
        R* d = nhist + (size_t)s * HSTRIDE;
        for (int k = 0; k < HSTRIDE; ++k) d[k] = keep ? h[k] : R(-1.0) * h[k];
        nfword[s] = nw | VALID(p);
    }
}

__global__ void k_inverse_perm(int N, const int* perm, int* inv) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) inv[perm[n]] = n;
}

// Pass 1: near flag of every front entry (into flag[]) and near count per sphere.
__global__ __launch_bounds__(256) void k_near_scan(DevParams P, const D4* pos, const double* slack, const int* foff,
                                                   const uint32_t* fword, int* ncnt, int* flag) {
    __shared__ int s_fo[4][65];
    __shared__ int s_cnt[4][64];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i0 = (blockIdx.x * 4 + wv) * 64;
    if (i0 >= P.N) return;
    const int i = i0 + lane;
    const bool act = i < P.N;
    const D4 a = act ? pos[i] : D4{0.0, 0.0, 0.0, 0.0};
    const double sa = act ? slack[i] : 0.0;
    s_cnt[wv][lane] = 0;
    load_rows(s_fo[wv], foff, i0, P.N);
    const int end = s_fo[wv][64];
    for (int k0 = s_fo[wv][0]; k0 < end; k0 += 64) {
        const int e = k0 + lane;
        int r = 0;
        bool nr = false;
        uint32_t w = 0;
        int j = 0;
        if (e < end) {
            r = row_of(s_fo[wv], e);
            w = fword[e];
            j = (int)(w & JMASK);
        }
        // the row's own centre and slack come from the lane that owns the row
        const double ax = __shfl(a.x, r), ay = __shfl(a.y, r), az = __shfl(a.z, r), aw = __shfl(a.w, r);
        const double asl = __shfl(sa, r);
        if (e < end) {
            D4 ar; ar.x = ax; ar.y = ay; ar.z = az; ar.w = aw;
            nr = near_pair(ar, pos[j], w, asl, slack[j]);
            flag[e] = nr ? 1 : 0;
            if (nr) atomicAdd(&s_cnt[wv][r], 1);
        }
    }
    wave_sync();
    if (act) ncnt[i] = s_cnt[wv][lane];
}

// Pass 2: near entries in slot order at nfoff[i] + (near entries before them
// in the row); slot2near (which held the flags) becomes the near index or -1.
__global__ __launch_bounds__(256) void k_near_put(int N, const int* foff, const uint32_t* fword, const int* nfoff,
                                                  uint32_t* nfw, int* nfs, int* slot2near) {
    __shared__ int s_fo[4][65];
    __shared__ int s_carry[4][64];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i0 = (blockIdx.x * 4 + wv) * 64;
    if (i0 >= N) return;
    const int i = i0 + lane;
    const int no = nfoff[min(i, N)];
    s_carry[wv][lane] = 0;
    load_rows(s_fo[wv], foff, i0, N);
    const int end = s_fo[wv][64];
    for (int k0 = s_fo[wv][0]; k0 < end; k0 += 64) {
        const int e = k0 + lane;
        int r = 0, f = 0;
        if (e < end) {
            r = row_of(s_fo[wv], e);
            f = slot2near[e];
        }
        const uint64_t m = __ballot(f != 0);
        const int lo = max(s_fo[wv][r] - k0, 0);  // first lane of this row in the chunk
        const uint64_t below = lanemask_lt() & ~((1ull << lo) - 1ull);
        const int rank = __popcll(m & below);
        const int base = __shfl(no, r) + s_carry[wv][r];
        if (e < end) {
            if (f) {
                const int o = base + rank;
                nfw[o] = fword[e];
                nfs[o] = e;
                slot2near[e] = o;
            } else {
                slot2near[e] = -1;
            }
        }
        wave_sync();
        if (e < end && f) atomicAdd(&s_carry[wv][r], 1);
        wave_sync();
    }
}

// Back near lists: a back entry is near when its pair is on the owner's near
// row. Pass 1 counts, pass 2 writes (owner, near index) in back-row order.
__global__ __launch_bounds__(256) void k_nback_scan(int N, const int* boff, const int2* bent, const int* slot2near,
                                                    int* nbcnt) {
    __shared__ int s_fo[4][65];
    __shared__ int s_cnt[4][64];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i0 = (blockIdx.x * 4 + wv) * 64;
    if (i0 >= N) return;
    s_cnt[wv][lane] = 0;
    load_rows(s_fo[wv], boff, i0, N);
    const int end = s_fo[wv][64];
    for (int k0 = s_fo[wv][0]; k0 < end; k0 += 64) {
        const int e = k0 + lane;
        if (e < end) {
            const int r = row_of(s_fo[wv], e);
            const int2 be = bent[e];
            if (be.y >= 0 && slot2near[be.y] >= 0) atomicAdd(&s_cnt[wv][r], 1);
        }
    }
    wave_sync();
    if (i0 + lane < N) nbcnt[i0 + lane] = s_cnt[wv][lane];
}

__global__ __launch_bounds__(256) void k_nback_put(int N, const int* boff, const int2* bent, const int* slot2near,
                                                   const int* nboff, int* nbo, int* nbe) {
    __shared__ int s_fo[4][65];
    __shared__ int s_carry[4][64];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i0 = (blockIdx.x * 4 + wv) * 64;
    if (i0 >= N) return;
    const int no = nboff[min(i0 + lane, N)];
    s_carry[wv][lane] = 0;
    load_rows(s_fo[wv], boff, i0, N);
    const int end = s_fo[wv][64];
    for (int k0 = s_fo[wv][0]; k0 < end; k0 += 64) {
        const int e = k0 + lane;
        int r = 0, q = -1, own = 0;
        if (e < end) {
            r = row_of(s_fo[wv], e);
            const int2 be = bent[e];
            if (be.y >= 0) q = slot2near[be.y];
            own = be.x & ~BINTRA;
        }
        const uint64_t m = __ballot(q >= 0);
        const int lo = max(s_fo[wv][r] - k0, 0);
        const uint64_t below = lanemask_lt() & ~((1ull << lo) - 1ull);
        const int o = __shfl(no, r) + s_carry[wv][r] + __popcll(m & below);
        if (q >= 0) {
            nbo[o] = own;
            nbe[o] = q;
        }
        wave_sync();
        if (q >= 0) atomicAdd(&s_carry[wv][r], 1);
        wave_sync();
    }
}

// history-valid bits of the near pairs back into the pair words, and the
// latest history (buffer q, kept in near order between near builds) back to
// the pairs' front slots
template <class R>
__global__ void k_near_sync(int N, const int* nfoff, const uint32_t* nfw, const int* nfs, uint32_t* fword,
                            const R* hn, R* hs, int q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t V = VALID(0) | VALID(1);
    for (int o = nfoff[i]; o < nfoff[i + 1]; ++o) {
        const int s = nfs[o];
        const uint32_t w = nfw[o];
        fword[s] = (fword[s] & ~V) | (w & V);
        if (w & VALID(q))
            for (int k = 0; k < HSTRIDE; ++k) hs[(size_t)s * HSTRIDE + k] = hn[(size_t)o * HSTRIDE + k];
    }
}

// The latest history (buffer q) of the new near pairs, from their front slots
// into near order.
template <class R>
__global__ void k_near_hist_gather(int nnear, const uint32_t* nfw, const int* nfs, const R* hs, R* hn, int q) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nnear) return;
    if (!(nfw[o] & VALID(q))) return;
    const int s = nfs[o];
    for (int k = 0; k < HSTRIDE; ++k) hn[(size_t)o * HSTRIDE + k] = hs[(size_t)s * HSTRIDE + k];
}

// ---------------------------------------------------------------------------
// Hertz-Mindlin sphere-sphere contact of the pair (I, J), I < J, as in
// src/dem_force_clump.cu:54-161. Geometry (Dist = x_J - x_I, DistS, DistL,
// Dn > 0) comes in FP64; R is the arithmetic type of the force law.
template <class R>
struct PairIn {
    V3<R> Dist;
    R DistS, DistL, Dn, rI, rJ, mI, mJ;
    V3<R> armI, armJ;  // R(q) * local of each sphere
    V3<R> vI, wI, vJ, wJ;
    V3<R> H, Mt, Mr;   // tangential, twisting, rolling history
};
template <class R>
struct PairOut {
    V3<R> cfc, ctc, ctc2, ht, htw, hr, fcn;
    R Ap;
    V3<R> DistU;
};

// Hertz-Mindlin stiffness and damping. FP64: the reference's expressions in
// its evaluation order (src/dem_force_clump.cu:68-77). FP32: the same
// quantities with the material constants hoisted into DevParams.
template <class R>
struct Coef {
    R Kn, Cn, Ks, Cs, Kr, Cr, Kt, Ct, idt;
};
template <class R>
__device__ __forceinline__ void hm_coef(const DevParams& P, R Rij, R Mij, R Dn, R CsFactor, Coef<R>& c) {
    if constexpr (sizeof(R) == sizeof(double)) {
        const R Ymod = P.Ymod, Prat = P.Prat, Damping = P.Damping, Beta = P.Beta;
        c.Kn = R(2.0) * Ymod * sqrt(Rij * Dn) / (R(3.0) * (R(1.0) - Prat * Prat));
        c.Cn = Damping * sqrt(c.Kn * Mij);
        c.Ks = R(2.0) * Ymod / (R(1.0) + Prat) / (R(2.0) - Prat) * sqrt(Rij * Dn);
        c.Cs = CsFactor == R(1.0) ? Damping * sqrt(c.Ks * Mij) : CsFactor * Damping * sqrt(c.Ks * Mij);
        c.Kr = R(0.25) * c.Kn * (Beta * Rij) * (Beta * Rij);
        c.Cr = R(0.25) * c.Cn * (Beta * Rij) * (Beta * Rij);
        c.Kt = R(0.50) * c.Ks * (Beta * Rij) * (Beta * Rij);
        c.Ct = R(0.50) * c.Cs * (Beta * Rij) * (Beta * Rij);
        c.idt = R(1.0) / (R)P.dt;
    } else {
        const R sq = sqrt(Rij * Dn), D = (R)P.Damping, bR = (R)P.Beta * Rij, b2 = bR * bR;
        c.Kn = (R)P.cKn * sq;
        c.Cn = D * sqrt(c.Kn * Mij);
        c.Ks = (R)P.cKs * sq;
        c.Cs = CsFactor * D * sqrt(c.Ks * Mij);
        c.Kr = R(0.25) * c.Kn * b2;
        c.Cr = R(0.25) * c.Cn * b2;
        c.Kt = R(0.50) * c.Ks * b2;
        c.Ct = R(0.50) * c.Cs * b2;
        c.idt = (R)P.idt;
    }
}

// DM: a DEMmain single-sphere loop (spin about the sphere centre), fixed at
// compile time so that the clump loop's kernels carry no extra registers
template <class R, bool DM>
__device__ __forceinline__ void pair_force(const DevParams& P, const PairIn<R>& in, PairOut<R>& o) {
    const R Beta = (R)P.Beta, Frcoe = (R)P.Frcoe;
    const R dt = (R)P.dt, Coh = (R)P.Cohesion;
    const R DistR = R(1.0) / in.DistL;
    const V3<R> DistU = DistR * in.Dist;
    const R Ap = R(0.50) * (in.rI * in.rI - in.rJ * in.rJ + in.DistS) * DistR;
    const R An = in.DistL - Ap;
    const R Rij = in.rI * in.rJ / (in.rI + in.rJ);
    const R Mij = in.mI * in.mJ / (in.mI + in.mJ);
    Coef<R> k;
    hm_coef<R>(P, Rij, Mij, in.Dn, R(1.0), k);
    const R Kn = k.Kn, Cn = k.Cn, Ks = k.Ks, Cs = k.Cs, Kr = k.Kr, Cr = k.Cr, Kt = k.Kt, Ct = k.Ct;
    const V3<R> ARMi = in.armI + Ap * DistU;
    const V3<R> ARMj = in.armJ - An * DistU;
    const V3<R> Vrel = in.vJ - in.vI;
    // DEMmain loops: the sphere's own spin about its centre
    // (src/dem_force_particle.cu:87-88); the clump form otherwise
    const V3<R> Vrot = DM ? An * Cross<R>(DistU, in.wJ) + Ap * Cross<R>(DistU, in.wI)
                          : Cross<R>(in.wJ, ARMj) - Cross<R>(in.wI, ARMi);
    const V3<R> Vtot = Vrel + Vrot;
    const R ERR = Dot<R>(DistU, Vtot);
    const V3<R> Vnor = ERR * DistU;
    const V3<R> Vtan = Vtot - Vnor;
    const V3<R> nf = Kn * in.Dn * DistU + Cn * Vnor;
    const R nfL = Length<R>(nf);
    const V3<R> Ds = dt * Vtan;
    V3<R> tf = -Ks * Ds + Cs * Vtan + in.H;
    const R tfL = Length<R>(tf);
    V3<R> th;
    if (tfL > nfL * Frcoe) {
        tf = Frcoe * nfL * (R(1.0) / tfL) * tf;
        th = tf;
    } else {
        th = tf - Cs * Vtan;
    }
    const V3<R> cf = Coh * (Beta * Rij) * (Beta * Rij) * DistU;
    V3<R> cfc = R(-1.0) * nf - tf + cf;
    V3<R> ctc = -Ap * Cross<R>(DistU, tf);
    V3<R> ctc2 = -An * Cross<R>(DistU, tf);
    const V3<R> Dth = dt * (in.wI - in.wJ);
    const R DthL = Dot<R>(Dth, DistU);
    const V3<R> DthT = DthL * DistU;
    const V3<R> DthR = Dth - DthT;
    V3<R> rm = Kr * DthR + in.Mr;
    const R rmL = Length<R>(rm);
    V3<R> rh;
    if (rmL > R(0.525) * Beta * Rij * nfL) {
        rm = R(0.525) * Beta * Rij * nfL * (R(1.0) / rmL) * rm;
        rh = rm;
    } else {
        rh = rm;
        rm = rm - Cr * k.idt * DthR;
    }
    V3<R> tm = Kt * DthT + in.Mt;
    const R tmL = Length<R>(tm);
    V3<R> twh;
    if (tmL > R(0.65) * Frcoe * Beta * Rij * nfL) {
        tm = R(0.65) * Frcoe * Beta * Rij * nfL * (R(1.0) / tmL) * tm;
        twh = tm;
    } else {
        twh = tm;
        tm = tm - Ct * k.idt * DthT;
    }
    ctc = ctc - rm - tm;
    ctc2 = ctc2 + rm + tm;
    o.cfc = cfc;
    o.ctc = ctc;
    o.ctc2 = ctc2;
    o.ht = th;
    o.htw = twh;
    o.hr = rh;
    o.fcn = R(-1.0) * nf;
    o.Ap = Ap;
    o.DistU = DistU;
}

// Wall / penetrator contact of one sphere (src/dem_force_cube_wall_clump.cu:49-165
// for walls with CsFactor = 2, src/dem_force_penetrator_clump.cu:35-177 for the
// penetrator with CsFactor = 1). Writes the force/torque terms and the new history.
template <class R>
struct WallOut {
    V3<R> nf, tf, rm, tm, ht, htw, hr, arm;
};
template <class R, bool DM>
__device__ __forceinline__ void wall_force(const DevParams& P, R Frcoe, R CsFactor, const V3<R>& DistU, R Ap, R Dn,
                                           R Rij, R Mij, const V3<R>& arm0, const V3<R>& RVo, const V3<R>& cv,
                                           const V3<R>& cw, const V3<R>& wv, const V3<R>& ww, const V3<R>& H,
                                           const V3<R>& Mt, const V3<R>& Mr, WallOut<R>& o) {
    const R Beta = (R)P.Beta, dt = (R)P.dt;
    Coef<R> k;
    hm_coef<R>(P, Rij, Mij, Dn, CsFactor, k);
    const R Kn = k.Kn, Cn = k.Cn, Ks = k.Ks, Cs = k.Cs, Kr = k.Kr, Cr = k.Cr, Kt = k.Kt, Ct = k.Ct;
    const V3<R> Arm = R(-1.0) * arm0 + Ap * DistU;
    const V3<R> Vwall = Cross<R>(ww, RVo);
    const V3<R> Vrel = cv - wv;
    // DEMmain walls spin the sphere about its centre (src/dem_force_cube_wall.cu:54,
    // src/dem_force_cylind_wall.cu:55)
    const V3<R> Vrot = DM ? Ap * Cross<R>(DistU, cw) - Vwall : Cross<R>(Arm, cw) - Vwall;
    const V3<R> Vtot = Vrel + Vrot;
    const R ERR = Dot<R>(DistU, Vtot);
    const V3<R> Vnor = ERR * DistU;
    const V3<R> Vtan = Vtot - Vnor;
    const V3<R> nf = Kn * Dn * DistU + Cn * Vnor;
    const R nfL = Length<R>(nf);
    const V3<R> Ds = dt * Vtan;
    V3<R> tf = -Ks * Ds + Cs * Vtan + H;
    const R tfL = Length<R>(tf);
    V3<R> th;
    if (tfL > nfL * Frcoe) {
        tf = Frcoe * nfL * (R(1.0) / tfL) * tf;
        th = tf;
    } else {
        th = tf - Cs * Vtan;
    }
    const V3<R> Dth = dt * (cw - ww);
    const R DthL = Dot<R>(Dth, DistU);
    const V3<R> DthT = DthL * DistU;
    const V3<R> DthR = Dth - DthT;
    V3<R> rm = Kr * DthR + Mr;
    const R rmL = Length<R>(rm);
    V3<R> rh;
    if (rmL > R(0.525) * Beta * Rij * nfL) {
        rm = R(0.525) * Beta * Rij * nfL * (R(1.0) / rmL) * rm;
        rh = rm;
    } else {
        rh = rm;
        rm = rm - Cr * k.idt * DthR;
    }
    V3<R> tm = Kt * DthT + Mt;
    const R tmL = Length<R>(tm);
    V3<R> twh;
    if (tmL > R(0.65) * Frcoe * Beta * Rij * nfL) {
        tm = R(0.65) * Frcoe * Beta * Rij * nfL * (R(1.0) / tmL) * tm;
        twh = tm;
    } else {
        twh = tm;
        tm = tm - Ct * k.idt * DthT;
    }
    o.nf = nf; o.tf = tf; o.rm = rm; o.tm = tm; o.ht = th; o.htw = twh; o.hr = rh; o.arm = Arm;
}

template <class R>
__device__ __forceinline__ void load_hist(const R* h, V3<R>& ht, V3<R>& htw, V3<R>& hr) {
    ht = mk<R>(h[0], h[1], h[2]);
    htw = mk<R>(h[3], h[4], h[5]);
    hr = mk<R>(h[6], h[7], h[8]);
}
template <class R>
__device__ __forceinline__ void store_hist(R* h, const V3<R>& ht, const V3<R>& htw, const V3<R>& hr) {
    h[0] = ht.x; h[1] = ht.y; h[2] = ht.z;
    h[3] = htw.x; h[4] = htw.y; h[5] = htw.z;
    h[6] = hr.x; h[7] = hr.y; h[8] = hr.z;
}

// Closest point on triangle abc to p (Ericson, Real-Time Collision Detection 5.1.5).
__device__ __forceinline__ V3<double> closest_on_tri(const V3<double>& p, const V3<double>& a, const V3<double>& b,
                                                     const V3<double>& c) {
    const V3<double> ab = b - a, ac = c - a, ap = p - a;
    const double d1 = Dot<double>(ab, ap), d2 = Dot<double>(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const V3<double> bp = p - b;
    const double d3 = Dot<double>(ab, bp), d4 = Dot<double>(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) { const double v = d1 / (d1 - d3); return a + v * ab; }
    const V3<double> cp = p - c;
    const double d5 = Dot<double>(ab, cp), d6 = Dot<double>(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) { const double w = d2 / (d2 - d6); return a + w * ac; }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return b + w * (c - b);
    }
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    return a + v * ab + w * ac;
}

// ---------------------------------------------------------------------------
// K3: all contacts of sphere i, then its force and clump-torque contribution.
// Candidate tests for k_walls. With sl = 0 they are the exact FP64 face
// predicates of the wall code (a face is evaluated only when its predicate
// holds) and the penetrator culls. Every predicate is a threshold on a
// distance-like function of the position with Lipschitz constant at most
// (1 + |dir|^2), so lowering the thresholds by sl * (1 + |dir|^2) gives a test
// that holds for every position within sl of xi (the wall list uses the
// distance a sphere can move before the list is rebuilt). A sphere failing it
// has no wall or penetrator contact.
__device__ __forceinline__ bool wall_candidate(const DevParams& P, const WallGeo& W, const V3<double>& xi, double ri,
                                               double sl) {
    const V3<double> RV = xi - mk<double>(W.pos[0], W.pos[1], W.pos[2]);
    if (P.wall_type == NBC_WALL_CUBE) {
        bool c = false;
        for (int a = 0; a < 3; ++a) {
            const double E = a == 0 ? RV.x : (a == 1 ? RV.y : RV.z);
            const double Sd = W.side[a];
            c = c || (E > -sl && E + ri - Sd > -sl) || (E < sl && E - ri + Sd < sl);
        }
        return c;
    }
    const V3<double> Wv = mk<double>(W.dir[0], W.dir[1], W.dir[2]);
    const double s2 = sl * (1.0 + Dot<double>(Wv, Wv));
    const double ERR = Dot<double>(RV, Wv);
    const V3<double> RVb = RV - ERR * Wv;
    const double norm = Length<double>(RVb);
    const double d_cap = fabs(ERR) - W.height + ri;
    const double d_side = norm - W.rad + ri;
    return (d_cap >= 1e-14 - s2 && (ERR != 0.0 || sl > 0.0)) || d_side >= 1e-14 - s2 || !(d_cap == d_cap) ||
           !(d_side == d_side);
}

__device__ __forceinline__ bool pene_candidate(const DevParams& P, const PeneState* pst, const V3<double>& xi,
                                               double ri, double sl) {
    const V3<double> RV = xi - mk<double>(pst->pos[0], pst->pos[1], pst->pos[2]);
    if (P.pshape == NBC_PENE_SPHERE) return ri - (Length<double>(RV) - P.prad) > -sl;
    if (P.pshape == NBC_PENE_CONE) {
        const double reach = P.pbound + ri + sl;
        return P.ntri > 0 && Dot<double>(RV, RV) < reach * reach;
    }
    const V3<double> pdir = mk<double>(P.pdir[0], P.pdir[1], P.pdir[2]);
    const double s2 = sl * (1.0 + Dot<double>(pdir, pdir));
    const double LatDx = P.cell[0];
    const double ERR = Dot<double>(RV, pdir);
    if (!(fabs(ERR) < (P.pheight + 0.80 * LatDx) + s2)) return false;
    const double norm = Length<double>(RV - ERR * pdir);
    return norm < (P.prad + 0.80 * LatDx) + s2;
}

// Queue entry of the contact kernel: the other sphere's FP64 position as the
// scan gathered it, overwritten by the lane that evaluates the pair with its
// six result values (force, torque).
template <class R> struct QEnt;
template <> struct alignas(16) QEnt<float> { float v[8]; };
template <> struct alignas(16) QEnt<double> { double v[6]; };
template <class R>
__device__ __forceinline__ void qent_put_pos(QEnt<R>& q, const D4& p) { __builtin_memcpy(q.v, &p, sizeof(D4)); }
template <class R>
__device__ __forceinline__ D4 qent_pos(const QEnt<R>& q) {
    D4 p;
    __builtin_memcpy(&p, q.v, sizeof(D4));
    return p;
}

// K3: sphere-sphere contacts of 64 spheres per wave (src/dem_force_clump.cu:3-191),
// then each sphere's totals with its wall and penetrator terms
// (src/dem_clump.cu:89-101) and the clump torque term T + (x_s - x_c) x F
// (src/dem_clumpsphere.cu:86-107). The wave's 64 spheres own the contiguous
// near-list segments [nfoff[wbase], nfoff[wbase + 64]) and [nboff[wbase],
// nboff[wbase + 64]); the wave walks each segment 64 entries at a time
// (coalesced list reads, list words loaded one chunk ahead, equal work per
// lane), finds an entry's row by binary search over the row offsets in LDS,
// applies the FP32 pre-filter and queues the candidates in LDS. A flush takes
// the queue one pair per lane: the FP64 touching test, then the force law.
// Every pair is evaluated by both spheres' waves from the same inputs, so each
// wave writes only its own spheres (no float atomics, no second kernel); the
// owner writes the new history into the other ping-pong buffer and the pair's
// history-valid bits. Each lane adds its sphere's results in queue order,
// which is its slot order: front entries ascending, then back entries in
// back-row order.
template <class R, bool RECORD, bool DM, int NW>
__device__ __forceinline__ void contacts_body(const DevParams& P, const SphereBuf<R>& S, const ClumpBuf<R>& C,
                                              const ListBuf<R>& L, int p, const WallBuf<R>& WL,
                                              unsigned long long* counters, int* bad, const RecordBuf& rec, int remap,
                                              const int* halt, int vblock, int nvblocks) {
    typedef typename Vec4T<R>::type R4;
    const int* status = halt + 3;  // ST_* bits
    constexpr int Q = 128;
    __shared__ int q_e[NW][Q];    // front near index of the pair; -1 once found not touching
    __shared__ int q_oth[NW][Q];  // the other sphere
    __shared__ int q_key[NW][Q];  // owner lane, +64 for back entries (non-decreasing)
    // FP64 form: the scan also gathers the other sphere's clump and the pair's
    // history-valid bits, so no load of the flush waits on a clump index or a
    // pair word (121 -> 117.6 us on the settled 1.12 M-sphere bed); the FP32
    // form, slower with more loads in flight during the scan (+0.5 us), loads
    // them in the flush
    constexpr bool CARRY = sizeof(R) == sizeof(double);
    __shared__ uint32_t q_cx[NW][CARRY ? Q : 1];  // other sphere's clump | the pair word's history-valid bits
    __shared__ QEnt<R> q_ent[NW][Q];  // other sphere's position, then the pair's results
    // row offsets stay in each lane's registers; the binary search reads them
    // with lane shuffles, which need no LDS allocation (6.7 instead of 7.2 KB
    // per block; 69.8 -> 68.5 us on the settled 1.12 M-sphere bed)
    int fo_l = 0, bo_l = 0, fo_e = 0, bo_e = 0;
    int t_slot = 0, t_wk = -1;  // the tail's member slot and wall-list position, loaded with the prologue
    __shared__ F4 s_p32[NW][64];
    __shared__ int s_cid[NW][CARRY ? 64 : 1];
    // XCD runs of 16,384 spheres (runs of 1,024 / 4,096 / 65,536 / 131,072
    // spheres: 69.3 / 68.5 / 67.9 / 69.6 us against 67.8 on the settled
    // 1.12 M-sphere bed)
    const int lb = remap ? xcd_group_block(vblock, nvblocks, 256 / NW) : vblock;
    const int i = lb * (64 * NW) + threadIdx.x;
    const bool active = i < P.N;
    const int lane = threadIdx.x & 63, wv = NW > 1 ? threadIdx.x >> 6 : 0;
    const int wbase = i - lane;
    R* hcur = L.hist[p];
    R* hnew = L.hist[p ^ 1];
    int n_ss = 0;
    V3<R> Fss = mk<R>(R(0), R(0), R(0)), Tss = Fss;
    {
        const int ir = i < P.N ? i : P.N;
        fo_l = L.nfoff[ir];
        bo_l = L.nboff[ir];
        if (lane == 63) {
            const int ie = i + 1 < P.N ? i + 1 : P.N;
            fo_e = L.nfoff[ie];
            bo_e = L.nboff[ie];
        }
        const D4 pa = S.pos[i < P.N ? i : P.N - 1];  // clamped load, then a value select
        t_slot = S.slot[i < P.N ? i : P.N - 1];
        if (P.nfaces > 0 || P.pene_on) t_wk = WL.pos_of[i < P.N ? i : P.N - 1];
        if (CARRY) s_cid[wv][lane] = S.cid[i < P.N ? i : P.N - 1];
        F4 a = pos32(P, xyz(pa), pa.w);
        if (!active) a.w = -1.0f;
        s_p32[wv][lane] = a;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    int qn = 0;  // wave-uniform queue length

    auto flush = [&]() __attribute__((always_inline)) {
        for (int r0 = 0; r0 < qn; r0 += 64) {
            const int k = r0 + lane;
            if (k < qn) {
                const int key = q_key[wv][k];
                const int side = key >> 6;
                const int me = wbase + (key & 63);
                const int oth = q_oth[wv][k];
                const int e = q_e[wv][k];
                const int I = side ? oth : me, J = side ? me : oth;
                // the other sphere's position, clump and the pair's valid bits
                // come from the scan's gathers; an owner's near word is its
                // partner and those bits (intra-clump pairs are never near)
                const D4 pM = S.pos[me], pO = qent_pos<R>(q_ent[wv][k]);
                const D4 pI = side ? pO : pM, pJ = side ? pM : pO;
                const uint32_t cx = CARRY ? q_cx[wv][k] : 0u;
                const uint32_t w = CARRY ? ((uint32_t)oth | (cx & (VALID(0) | VALID(1)))) : L.nfw[e];
                const V3<double> Dd = xyz(pJ) - xyz(pI);
                const double DistS = Dd.x * Dd.x + Dd.y * Dd.y + Dd.z * Dd.z;
                const double rs = pI.w + pJ.w;
                bool touch = false;
                if (!(DistS > rs * rs * (1.0 + 1e-9))) touch = rs - sqrt(DistS) > 0.0;
                if (!touch) {
                    if (side == 0 && (w & VALID(p ^ 1))) L.nfw[e] = w & ~VALID(p ^ 1);
                    q_e[wv][k] = -1;
                } else {
                    if (side == 0) n_ss++;
                    const R4 kI = S.kin[I], kJ = S.kin[J];
                    int cI, cJ;
                    if (CARRY) {
                        const int cO = (int)(cx & JMASK), cM = s_cid[wv][key & 63];
                        cI = side ? cO : cM;
                        cJ = side ? cM : cO;
                    } else {
                        cI = S.cid[I];
                        cJ = S.cid[J];
                    }
                    const R4 svI = C.snap[2 * cI], swI = C.snap[2 * cI + 1];
                    const R4 svJ = C.snap[2 * cJ], swJ = C.snap[2 * cJ + 1];
                    const int s = RECORD ? L.nfs[e] : 0;  // front slot: record index
                    const double DistL = sqrt(DistS);
                    const double Dn = pI.w + pJ.w - DistL;
                    PairIn<R> in;
                    in.Dist = cvt3<R>(Dd); in.DistS = (R)DistS; in.DistL = (R)DistL; in.Dn = (R)Dn;
                    in.rI = (R)pI.w; in.rJ = (R)pJ.w; in.mI = kI.w; in.mJ = kJ.w;
                    in.armI = Ld4<R>::xyz(kI); in.armJ = Ld4<R>::xyz(kJ);
                    in.vI = Ld4<R>::xyz(svI); in.wI = Ld4<R>::xyz(swI); in.vJ = Ld4<R>::xyz(svJ);
                    in.wJ = Ld4<R>::xyz(swJ);
                    if (w & VALID(p)) load_hist<R>(hcur + (size_t)e * HSTRIDE, in.H, in.Mt, in.Mr);
                    else { in.H = mk<R>(R(0), R(0), R(0)); in.Mt = in.H; in.Mr = in.H; }
                    PairOut<R> o;
                    pair_force<R, DM>(P, in, o);
                    V3<R> f, t;
                    if (side == 0) {
                        store_hist<R>(hnew + (size_t)e * HSTRIDE, o.ht, o.htw, o.hr);
                        if (!(w & VALID(p ^ 1))) L.nfw[e] = w | VALID(p ^ 1);
                        f = o.cfc;
                        t = o.ctc;
                        if (RECORD) {
                            D4 a; a.x = (double)o.fcn.x; a.y = (double)o.fcn.y; a.z = (double)o.fcn.z; a.w = 0.0;
                            const V3<double> cpt = xyz(pI) + (double)o.Ap * mk<double>((double)o.DistU.x,
                                                                                      (double)o.DistU.y,
                                                                                      (double)o.DistU.z);
                            D4 bq; bq.x = cpt.x; bq.y = cpt.y; bq.z = cpt.z; bq.w = 0.0;
                            rec.pair_fcn[s] = a;
                            rec.pair_cpt[s] = bq;
                        }
                    } else {
                        f = R(-1.0) * o.cfc;  // x - y == x + (-y) exactly: same sums as subtracting
                        t = o.ctc2;
                    }
                    R* rr = q_ent[wv][k].v;
                    rr[0] = f.x; rr[1] = f.y; rr[2] = f.z; rr[3] = t.x; rr[4] = t.y; rr[5] = t.z;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // each lane's front and back entries are two runs of the sorted keys
        auto lbound = [&](int key) __attribute__((always_inline)) {
            int lo = 0;
            for (int st = Q; st > 0; st >>= 1)
                if (lo + st <= qn && q_key[wv][lo + st - 1] < key) lo += st;
            return lo;
        };
        const int f0 = lbound(lane), f1 = lbound(lane + 1), b0 = lbound(lane + 64), b1 = lbound(lane + 65);
        for (int k = f0; k < f1; ++k) {
            if (q_e[wv][k] < 0) continue;
            const R* rr = q_ent[wv][k].v;
            Fss = Fss + mk<R>(rr[0], rr[1], rr[2]);
            Tss = Tss + mk<R>(rr[3], rr[4], rr[5]);
        }
        for (int k = b0; k < b1; ++k) {
            if (q_e[wv][k] < 0) continue;
            const R* rr = q_ent[wv][k].v;
            Fss = Fss + mk<R>(rr[0], rr[1], rr[2]);
            Tss = Tss + mk<R>(rr[3], rr[4], rr[5]);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        qn = 0;
    };
    auto push = [&](bool cand, int key, int e, int oth, const D4& po, uint32_t cx) __attribute__((always_inline)) {
        const uint64_t m = __ballot(cand);
        if (cand) {
            const int pos = qn + __popcll(m & lanemask_lt());
            q_e[wv][pos] = e;
            q_oth[wv][pos] = oth;
            q_key[wv][pos] = key;
            if (CARRY) q_cx[wv][pos] = cx;
            qent_put_pos<R>(q_ent[wv][pos], po);
        }
        qn += __popcll(m);
        if (qn > Q - 64) flush();
    };

    // the back list's first chunk is fetched before the front scan
    const int B0 = __builtin_amdgcn_readfirstlane(bo_l), B1 = __builtin_amdgcn_readlane(bo_e, 63);
    int onext = 0, enext = 0;
    if (B0 + lane < B1) {
        onext = __builtin_nontemporal_load(L.nbo + B0 + lane);
        enext = __builtin_nontemporal_load(L.nbe + B0 + lane);
    }
    // front entries: pairs the row's sphere owns (it writes their history)
    {
        const int F0 = __builtin_amdgcn_readfirstlane(fo_l), F1 = __builtin_amdgcn_readlane(fo_e, 63);
        uint32_t wnext = F0 + lane < F1 ? __builtin_nontemporal_load(L.nfw + F0 + lane) : 0u;
        for (int c = F0; c < F1; c += 64) {
            const int e = c + lane;
            const uint32_t w = wnext;
            if (e + 64 < F1) wnext = __builtin_nontemporal_load(L.nfw + e + 64);
            bool cand = false;
            int r = 0, j = 0, cj = 0;
            D4 pj;
            pj.x = pj.y = pj.z = pj.w = 0.0;
            // the entry's row; every lane of the wave takes part in the shuffles
#pragma unroll
            for (int st = 32; st > 0; st >>= 1)
                if (__shfl(fo_l, r + st, 64) <= e) r += st;
            if (e < F1) {
                j = (int)(w & JMASK);
                // the partner's FP32 copy is made here from its FP64 position: a
                // gather fetches the same DRAM sector for 16 or 32 bytes
                pj = S.pos[j];
                if (CARRY) cj = S.cid[j];
                cand = maybe_touch(P, s_p32[wv][r], pos32(P, xyz(pj), pj.w));
                if (!cand && (w & VALID(p ^ 1))) L.nfw[e] = w & ~VALID(p ^ 1);
            }
            push(cand, r, e, j, pj, (uint32_t)cj | (w & (VALID(0) | VALID(1))));
        }
    }
    // back entries: the owner holds the history; the same pair is evaluated
    // again from the same inputs and the row's sphere takes the reaction
    {
        for (int c = B0; c < B1; c += 64) {
            const int e = c + lane;
            const int o = onext, ne = enext;
            if (e + 64 < B1) {
                onext = __builtin_nontemporal_load(L.nbo + e + 64);
                enext = __builtin_nontemporal_load(L.nbe + e + 64);
            }
            bool cand = false;
            int r = 0, co = 0;
            uint32_t wo = 0;
            D4 po;
            po.x = po.y = po.z = po.w = 0.0;
#pragma unroll
            for (int st = 32; st > 0; st >>= 1)
                if (__shfl(bo_l, r + st, 64) <= e) r += st;
            if (e < B1) {
                po = S.pos[o];
                if (CARRY) {
                    co = S.cid[o];
                    wo = L.nfw[ne];
                }
                cand = maybe_touch(P, s_p32[wv][r], pos32(P, xyz(po), po.w));
            }
            push(cand, r + 64, ne, o, po, (uint32_t)co | (wo & (VALID(0) | VALID(1))));
        }
    }
    if (qn > 0) flush();
    if (active) {
        const R4 ki = S.kin[i];
        const size_t so = (size_t)t_slot;  // k_clump_after reads the totals in member order
        const int wk = t_wk;
        if (P.nfaces > 0 && wk < 0) {
            // input check of the wall kernel (src/dem_force_cube_wall_clump.cu:196-197)
            // for spheres off the wall list: exact only when a cheap sign is off
            const F4 q = s_p32[wv][lane];
            if (!(q.w > 0.0f) || !isfinite((double)ki.x) || !isfinite((double)ki.y) || !isfinite((double)ki.z) ||
                !isfinite((double)ki.w) || !(ki.w > R(0)) || (*status & ST_CLUMP_BAD)) {
                const D4 pi4 = S.pos[i];
                const R4 li = S.loc[i];
                const int c = S.cid[i];
                const bool cid_ok = c >= 0 && c < P.Nc;
                const bool ok = cid_ok && C.snap[2 * (cid_ok ? c : 0)].w != R(0) && finite3(xyz(pi4)) &&
                                isfinite((double)li.x) && isfinite((double)li.y) && isfinite((double)li.z) &&
                                isfinite(pi4.w) && pi4.w > 0.0 && isfinite((double)ki.w) && ki.w > R(0);
                if (!ok) atomicMin(bad, i);
            }
        }
        if (wk >= 0) {
            // wall-list sphere: k_walls runs beside this kernel, so the raw sums
            // go out (.w = 1 marks them) and k_clump_after adds the wall and
            // penetrator terms and the clump torque term
            S.sft[2 * so] = mk4<R>(Fss.x, Fss.y, Fss.z, R(0));
            S.sft[2 * so + 1] = mk4<R>(Tss.x, Tss.y, Tss.z, R(1));
        } else {
            // totF/totT with zero wall and penetrator terms (src/dem_clump.cu:97-98)
            // and the clump torque term T + (x_s - x_c) x F (src/dem_clumpsphere.cu:86-107)
            const V3<R> Z = mk<R>(R(0), R(0), R(0));
            const V3<R> tF = Fss + Z + Z;
            const V3<R> tT = Tss + Z + Z;
            V3<R> r;
            if (sizeof(R) == sizeof(double)) {
                const D4 pi4 = S.pos[i];
                const D4 pc = C.pq[2 * S.cid[i]];
                r = cvt3<R>(mk<double>(pi4.x - pc.x, pi4.y - pc.y, pi4.z - pc.z));
            } else {
                r = Ld4<R>::xyz(ki);  // x_s - x_c = R(q) * local up to FP64 rounding
            }
            const V3<R> Tq = mk<R>(tT.x + (r.y * tF.z - r.z * tF.y), tT.y + (r.z * tF.x - r.x * tF.z),
                                   tT.z + (r.x * tF.y - r.y * tF.x));
            S.sft[2 * so] = mk4<R>(tF.x, tF.y, tF.z, R(0));
            S.sft[2 * so + 1] = mk4<R>(Tq.x, Tq.y, Tq.z, R(0));
            if (RECORD) {
                D4 a; a.x = tF.x; a.y = tF.y; a.z = tF.z; a.w = 0.0;
                D4 b; b.x = tT.x; b.y = tT.y; b.z = tT.z; b.w = 0.0;
                rec.tot[2 * (size_t)i] = a;
                rec.tot[2 * (size_t)i + 1] = b;
                const V3<R> parts[6] = {Fss, Tss, Z, Z, Z, Z};
                for (int k = 0; k < 6; ++k) {
                    D4 q; q.x = parts[k].x; q.y = parts[k].y; q.z = parts[k].z; q.w = 0.0;
                    rec.parts[6 * (size_t)i + k] = q;
                }
            }
        }
    }
    // touching sphere-sphere count of this wave's spheres, stored in the
    // wave's own slot: one device-scope atomicAdd per block on a single
    // counter had cost the FP32 kernel 17 us per launch (89.5 -> 72.4 us on
    // the settled 1.12 M-sphere bed), same-address atomics being serialised
    // at the memory side
    unsigned a0 = (unsigned)n_ss;
    for (int o = 32; o > 0; o >>= 1) a0 += __shfl_down(a0, o, 64);
    if (lane == 0) P.wss[wbase >> 6] = a0;
}

// ---------------------------------------------------------------------------
// Wall and penetrator contacts of the spheres on the wall list
// (src/dem_force_cube_wall_clump.cu:16-298, src/dem_force_cylind_wall.cu:152-237,
// src/dem_contact_penetrator.cu:8-84, src/dem_force_penetrator_clump.cu:4-182).
// They depend only on each sphere's own state, so this runs before the
// sphere-sphere kernel and leaves (Fw, Tw, Fp, Tp) per list entry for it. One
// lane per list entry; the list is in sphere order, so the per-wave FP64
// penetrator partial sums keep the order of the sphere index.
template <class R, bool RECORD, bool DM>
__device__ __forceinline__ void walls_body(const DevParams& P, const SphereBuf<R>& S, const ClumpBuf<R>& C,
                                           const WallBuf<R>& WL, int len, PeneState* pst, double* pene_partial,
                                           unsigned long long* counters, int* bad, const RecordBuf& rec, int vblock) {
    typedef typename Vec4T<R>::type R4;
    const int lane = threadIdx.x & 63;
    const int k = vblock * (int)blockDim.x + threadIdx.x;
    const int gw = k >> 6;  // wave index: penetrator partial slot
    const bool active = k < len;
    int n_w = 0;
    int i = 0;
    V3<R> Fw = mk<R>(R(0), R(0), R(0)), Tw = Fw, Fp = Fw, Tp = Fw;
    V3<double> pf_sum = mk<double>(0.0, 0.0, 0.0), pt_sum = pf_sum;
    D4 pi4;
    pi4.x = pi4.y = pi4.z = pi4.w = 0.0;
    if (active) {
        i = WL.list[k];
        pi4 = S.pos[i];
    }
    const V3<double> xi = xyz(pi4);
    const double ri = pi4.w;
    if (active) {
        const R4 ki = S.kin[i];
        const int ci_raw = S.cid[i];
        const bool cid_ok = ci_raw >= 0 && ci_raw < P.Nc;
        const int ci = cid_ok ? ci_raw : 0;
        const R4 sv = C.snap[2 * ci], sw = C.snap[2 * ci + 1];
        const V3<R> armi = Ld4<R>::xyz(ki), vi = Ld4<R>::xyz(sv), wi = Ld4<R>::xyz(sw);
        const R mi = ki.w;
        // walls and penetrator
        uint32_t wf = S.wflag[i];
        const uint32_t wf0 = wf;
        const bool clump_ok = sv.w != R(0);
        if (P.nfaces > 0 || P.pene_on) {
            const R4 li = S.loc[i];
            const bool ok = cid_ok && clump_ok && finite3(xi) && isfinite((double)li.x) &&
                            isfinite((double)li.y) && isfinite((double)li.z) && isfinite(ri) && ri > 0.0 &&
                            isfinite((double)mi) && mi > R(0);
            if (P.nfaces > 0 && !ok) atomicMin(bad, i);
            const R Rr = (R)ri;
            if (P.nfaces > 0 && ok) {
                const WallGeo& W = P.nbeds > 1 ? P.beds[P.bid[i]].W : P.W;
                const V3<double> wpos = mk<double>(W.pos[0], W.pos[1], W.pos[2]);
                const V3<R> wv = mk<R>((R)W.vel[0], (R)W.vel[1], (R)W.vel[2]);
                const V3<R> ww = mk<R>((R)W.angv[0], (R)W.angv[1], (R)W.angv[2]);
                const V3<double> RV = xi - wpos;
                if (P.wall_type == NBC_WALL_CUBE) {
                    // faces 0 UpX, 1 LowX, 2 UpY, 3 LowY, 4 UpZ, 5 LowZ
                    for (int f = 0; f < 6; ++f) {
                        const int a = f >> 1;
                        const bool up = (f & 1) == 0;
                        const double E = a == 0 ? RV.x : (a == 1 ? RV.y : RV.z);
                        const double Sd = W.side[a];
                        const bool cand = up ? (E > 0 && E + ri - Sd > 0) : (E < 0 && E - ri + Sd < 0);
                        bool touch = false;
                        if (cand) {
                            const V3<double> ea = mk<double>(a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0);
                            const V3<double> RVb = RV - E * ea, RVt = E * ea;
                            const V3<double> RVo = (Sd / fabs(E)) * RVt + RVb;
                            const V3<double> RVn = RV - RVo;
                            if (DM && P.driver == NBC_DRIVER_CUBE) {
                                // DEMmain cube wall (src/dem_force_cube_wall.cu:3-148, :170-183):
                                // the normal and Ap come from RVn, Dn = r - |RVn| unchecked;
                                // its histories are never written (zero), and its rolling
                                // and twisting moments are not applied
                                const double DistL = Length<double>(RVn);
                                const double Dn = ri - DistL;
                                const double DistS = RVn.x * RVn.x + RVn.y * RVn.y + RVn.z * RVn.z;
                                const R DistLr = (R)sqrt(DistS);
                                const V3<R> DistU = (R(1.0) / DistLr) * cvt3<R>(RVn);
                                const V3<R> Z = mk<R>(R(0), R(0), R(0));
                                WallOut<R> o;
                                wall_force<R, DM>(P, (R)P.Frcoe, R(2.0), DistU, DistLr, (R)Dn, Rr, mi, armi,
                                              cvt3<R>(RVo), vi, wi, wv, ww, Z, Z, Z, o);
                                Fw = Fw + o.nf + o.tf;
                                Tw = Tw - DistLr * Cross<R>(DistU, o.tf);
                                n_w++;
                                continue;
                            }
                            const double sgn = up ? -1.0 : 1.0;
                            const V3<double> nin = mk<double>(a == 0 ? sgn : 0.0, a == 1 ? sgn : 0.0, a == 2 ? sgn : 0.0);
                            const double Ap = Dot<double>(RVn, nin);
                            const double Dn = ri - Ap;
                            if (Dn > 0.0) {
                                touch = true;
                                V3<R> H, Mt, Mr;
                                R* hp = S.wh + ((size_t)f * P.N + i) * HSTRIDE;
                                if ((wf >> f) & 1u) load_hist<R>(hp, H, Mt, Mr);
                                else { H = mk<R>(R(0), R(0), R(0)); Mt = H; Mr = H; }
                                WallOut<R> o;
                                wall_force<R, DM>(P, (R)P.Frcoe, R(2.0), cvt3<R>(nin), (R)Ap, (R)Dn, Rr, mi, armi,
                                              cvt3<R>(RVo), vi, wi, wv, ww, H, Mt, Mr, o);
                                Fw = Fw + o.nf + o.tf;
                                Tw = Tw - (R)Ap * Cross<R>(cvt3<R>(nin), o.tf);
                                Tw = Tw - o.rm - o.tm;
                                store_hist<R>(hp, o.ht, o.htw, o.hr);
                                n_w++;
                            }
                        }
                        if (touch) wf |= 1u << f; else wf &= ~(1u << f);
                    }
                } else {
                    // cylinder surfaces 0 bottom cap, 1 top cap, 2 side
                    const V3<double> Wv = mk<double>(W.dir[0], W.dir[1], W.dir[2]);
                    const double ERR = Dot<double>(RV, Wv);
                    const V3<double> RVb = RV - ERR * Wv, RVt = ERR * Wv;
                    const double norm = Length<double>(RVb);
                    const double d_cap = fabs(ERR) - W.height + ri;
                    const double d_side = norm - W.rad + ri;
                    for (int f = 0; f < 3; ++f) {
                        bool cand;
                        V3<double> RVo;
                        if (f == 0) { cand = d_cap >= 1e-14 && ERR < 0; RVo = (W.height / fabs(ERR)) * RVt + RVb; }
                        else if (f == 1) { cand = d_cap >= 1e-14 && ERR > 0; RVo = (W.height / fabs(ERR)) * RVt + RVb; }
                        else { cand = d_side >= 1e-14; RVo = (W.rad / norm) * RVb + RVt; }
                        bool touch = false;
                        if (cand) {
                            const V3<double> RVn = RV - RVo;
                            const double Dn = ri - Length<double>(RVn);
                            if (Dn > 0.0) {
                                touch = true;
                                const double DistS = RVn.x * RVn.x + RVn.y * RVn.y + RVn.z * RVn.z;
                                const double DistL = sqrt(DistS);
                                const R DistLr = (R)DistL;
                                const V3<R> DistU = (R(1.0) / DistLr) * cvt3<R>(RVn);
                                V3<R> H, Mt, Mr;
                                R* hp = S.wh + ((size_t)f * P.N + i) * HSTRIDE;
                                if ((wf >> f) & 1u) load_hist<R>(hp, H, Mt, Mr);
                                else { H = mk<R>(R(0), R(0), R(0)); Mt = H; Mr = H; }
                                WallOut<R> o;
                                wall_force<R, DM>(P, (R)P.Frcoe, R(2.0), DistU, DistLr, (R)Dn, Rr, mi, armi, cvt3<R>(RVo),
                                              vi, wi, wv, ww, H, Mt, Mr, o);
                                Fw = Fw + o.nf + o.tf;
                                Tw = Tw - DistLr * Cross<R>(DistU, o.tf);
                                Tw = Tw - o.rm - o.tm;
                                store_hist<R>(hp, o.ht, o.htw, o.hr);
                                n_w++;
                            }
                        }
                        if (touch) wf |= 1u << f; else wf &= ~(1u << f);
                    }
                }
                if (!isfinite((double)Fw.x) || !isfinite((double)Fw.y) || !isfinite((double)Fw.z) ||
                    !isfinite((double)Tw.x) || !isfinite((double)Tw.y) || !isfinite((double)Tw.z))
                    atomicMin(bad, i);
            }
            if (P.pene_on) {
                const int pf = P.pene_face;
                const V3<double> pp = mk<double>(pst->pos[0], pst->pos[1], pst->pos[2]);
                const V3<double> pdir = mk<double>(P.pdir[0], P.pdir[1], P.pdir[2]);
                const V3<double> RV = xi - pp;
                bool touch = false;
                V3<double> RVo = mk<double>(0.0, 0.0, 0.0), Dist = RVo, DistU64 = RVo;
                double DistL = 0.0, Dn = 0.0;
                bool have_u = false;
                if (P.pshape == NBC_PENE_SPHERE) {
                    // src/dem_sphere_intruder.cuh:27-50
                    const double cd = Length<double>(RV);
                    V3<double> nrm;
                    if (cd > 1.0e-15) {
                        nrm = (1.0 / cd) * RV;
                    } else {
                        const double t = sqrt(pdir.x * pdir.x + pdir.y * pdir.y + pdir.z * pdir.z);
                        nrm = t > 1e-20 ? mk<double>(pdir.x / t, pdir.y / t, pdir.z / t) : mk<double>(0.0, 0.0, 0.0);
                        if (Length<double>(nrm) < 1.0e-15) nrm = mk<double>(0.0, 0.0, 1.0);
                    }
                    const double gap = cd - P.prad;
                    const double ov = ri - gap;
                    if (ov > 0.0) {
                        touch = true;
                        RVo = P.prad * nrm;
                        Dist = gap * nrm;
                        DistL = gap;
                        Dn = ov;
                        DistU64 = nrm;
                        have_u = true;
                        if (P.pwait) atomicCAS(&pst->contact_state, 0, 1);
                    }
                } else if (P.pshape == NBC_PENE_CONE) {
                    // spheres outside the mesh's bounding sphere cannot touch it
                    const double reach = P.pbound + ri;
                    if (Dot<double>(RV, RV) < reach * reach) {
                    double best = 1e300;
                    V3<double> bq = mk<double>(0.0, 0.0, 0.0);
                    for (int t = 0; t < P.ntri; ++t) {
                        const double* v = P.ptri + 9 * t;
                        const V3<double> q = closest_on_tri(RV, mk<double>(v[0], v[1], v[2]), mk<double>(v[3], v[4], v[5]),
                                                            mk<double>(v[6], v[7], v[8]));
                        const V3<double> dq = RV - q;
                        const double ds = Dot<double>(dq, dq);
                        if (ds < best) { best = ds; bq = q; }
                    }
                    const V3<double> RVn = RV - bq;
                    const double DistS = RVn.x * RVn.x + RVn.y * RVn.y + RVn.z * RVn.z;
                    const double dl = sqrt(DistS);
                    if (P.ntri > 0 && ri - dl > 0.0) {
                        touch = true; RVo = bq; Dist = RVn; DistL = dl; Dn = ri - dl;
                    }
                    }
                } else {
                    // src/dem_contact_penetrator.cu:42-83
                    const double LatDx = P.cell[0];
                    const double ERR = Dot<double>(RV, pdir);
                    if (fabs(ERR) < (P.pheight + 0.80 * LatDx)) {
                        const V3<double> RVb = RV - ERR * pdir;
                        const double norm = Length<double>(RVb);
                        if (norm < (P.prad + 0.80 * LatDx)) {
                            const V3<double> RVt = ERR * pdir;
                            V3<double> o3;
                            if (fabs(ERR) < P.pheight) o3 = (1.0 / norm) * P.prad * RVb + RVt;
                            else if (norm < P.prad) o3 = (1.0 / fabs(ERR)) * P.pheight * RVt + RVb;
                            else o3 = (1.0 / norm) * P.prad * RVb + (1.0 / fabs(ERR)) * P.pheight * RVt;
                            const V3<double> RVn = RV - o3;
                            const double DistS = RVn.x * RVn.x + RVn.y * RVn.y + RVn.z * RVn.z;
                            const double dl = sqrt(DistS);
                            if (ri - dl > 0.0) { touch = true; RVo = o3; Dist = RVn; DistL = dl; Dn = ri - dl; }
                        }
                    }
                }
                if (touch) {
                    const R DistLr = (R)DistL;
                    const V3<R> DistU = have_u ? cvt3<R>(DistU64) : (R(1.0) / DistLr) * cvt3<R>(Dist);
                    R Rij = Rr;
                    if (P.pshape == NBC_PENE_SPHERE) Rij = Rr * (R)P.prad / (Rr + (R)P.prad);
                    R Mij = mi;
                    if (P.pmotion == NBC_PENE_GRAVITY) Mij = mi * (R)P.pmass / (mi + (R)P.pmass);
                    V3<R> H, Mt, Mr;
                    R* hp = S.wh + ((size_t)pf * P.N + i) * HSTRIDE;
                    if ((wf >> pf) & 1u) load_hist<R>(hp, H, Mt, Mr);
                    else { H = mk<R>(R(0), R(0), R(0)); Mt = H; Mr = H; }
                    const V3<R> pv = mk<R>((R)pst->vel[0], (R)pst->vel[1], (R)pst->vel[2]);
                    const V3<R> pw = mk<R>((R)pst->angv[0], (R)pst->angv[1], (R)pst->angv[2]);
                    const V3<R> RVor = cvt3<R>(RVo);
                    WallOut<R> o;
                    wall_force<R, DM>(P, (R)P.pene_frcoe, R(1.0), DistU, DistLr, (R)Dn, Rij, Mij, armi, RVor, vi, wi, pv,
                                      pw, H, Mt, Mr, o);
                    // DEMmain's penetrator keeps the cohesion term (src/dem_force_penetrator.cu:119-123)
                    const V3<R> cf = DM ? (-(R)P.Cohesion * ((R)P.Beta * Rij) * ((R)P.Beta * Rij)) * DistU
                                        : mk<R>(R(0), R(0), R(0));
                    Fp = DM ? o.nf + o.tf + cf : o.nf + o.tf;
                    Tp = -DistLr * Cross<R>(DistU, o.tf);
                    V3<R> pTc = R(-1.0) * Cross<R>(RVor, (o.nf + o.tf + cf));
                    Tp = Tp - o.rm - o.tm;
                    pTc = pTc + o.rm + o.tm;
                    store_hist<R>(hp, o.ht, o.htw, o.hr);
                    pf_sum = mk<double>((double)Fp.x, (double)Fp.y, (double)Fp.z);
                    pt_sum = mk<double>((double)pTc.x, (double)pTc.y, (double)pTc.z);
                    n_w++;
                    if (RECORD) {
                        const V3<R> tn = Cross<R>(o.nf, o.arm);
                        const V3<R> tt = Cross<R>(o.tf, o.arm) - o.rm - o.tm;
                        D4 a; a.x = o.nf.x; a.y = o.nf.y; a.z = o.nf.z; a.w = 0.0;
                        D4 b; b.x = o.tf.x; b.y = o.tf.y; b.z = o.tf.z; b.w = 0.0;
                        D4 c; c.x = tn.x; c.y = tn.y; c.z = tn.z; c.w = 0.0;
                        D4 d; d.x = tt.x; d.y = tt.y; d.z = tt.z; d.w = 0.0;
                        rec.pene[4 * (size_t)i] = a; rec.pene[4 * (size_t)i + 1] = b;
                        rec.pene[4 * (size_t)i + 2] = c; rec.pene[4 * (size_t)i + 3] = d;
                    }
                }
                if (touch) wf |= 1u << pf; else wf &= ~(1u << pf);
            }
        }
        if (wf != wf0) S.wflag[i] = wf;
        // results at the sphere's member slot, with the clump torque arm, so
        // that k_clump_after finds them next to the sphere's contact totals
        R4* o = WL.res + 5 * (size_t)S.slot[i];
        o[0] = mk4<R>(Fw.x, Fw.y, Fw.z, R(0));
        o[1] = mk4<R>(Tw.x, Tw.y, Tw.z, R(0));
        o[2] = mk4<R>(Fp.x, Fp.y, Fp.z, R(0));
        o[3] = mk4<R>(Tp.x, Tp.y, Tp.z, R(0));
        V3<R> arm;
        if (sizeof(R) == sizeof(double)) {
            const D4 pcl = C.pq[2 * ci];
            arm = cvt3<R>(mk<double>(pi4.x - pcl.x, pi4.y - pcl.y, pi4.z - pcl.z));
        } else {
            arm = Ld4<R>::xyz(ki);  // x_s - x_c = R(q) * local up to FP64 rounding
        }
        o[4] = mk4<R>(arm.x, arm.y, arm.z, R(0));
    }
    // wave reductions: touching count and the FP64 penetrator partial sums of
    // this wave's 64 entries (a fixed tree over the lanes)
    double v[6] = {pf_sum.x, pf_sum.y, pf_sum.z, pt_sum.x, pt_sum.y, pt_sum.z};
    unsigned a1 = (unsigned)n_w;
    for (int o = 32; o > 0; o >>= 1) {
        for (int q = 0; q < 6; ++q) v[q] += __shfl_down(v[q], o, 64);
        a1 += __shfl_down(a1, o, 64);
    }
    if (lane == 0 && (k - lane) < len) {
        if (a1) atomicAdd(&counters[1], (unsigned long long)a1);
        if (P.pene_on)
            for (int q = 0; q < 6; ++q) pene_partial[6 * (size_t)gw + q] = v[q];
    }
}

// K3 of a step: the sphere-sphere contacts, and, when the wall list is not
// empty, the wall and penetrator contacts of the listed spheres in the same
// launch (the first nwb blocks take the wall list; both parts only read what
// the other writes after k_clump_after). The FP32 form asks for five waves per
// SIMD, which the contact blocks need and the rarer wall blocks pay for with
// a few spills.
template <class R, bool RECORD, bool DM, int NW>
__device__ __forceinline__ void step3_body(const DevParams& P, const SphereBuf<R>& S, const ClumpBuf<R>& C,
                                           const ListBuf<R>& L, int p, const WallBuf<R>& WL, int len, int nwb,
                                           PeneState* pst, double* pene_partial, unsigned long long* counters,
                                           int* bad, const RecordBuf& rec, int remap, const int* halt, int n_arg) {
    if (*halt <= step_of(P, n_arg)) return;  // a rebuild of step n comes first
    if ((int)blockIdx.x < nwb)
        walls_body<R, RECORD, DM>(P, S, C, WL, len, pst, pene_partial, counters, bad, rec, blockIdx.x);
    else
        contacts_body<R, RECORD, DM, NW>(P, S, C, L, p, WL, counters, bad, rec, remap, halt, (int)blockIdx.x - nwb,
                                 (int)gridDim.x - nwb);
}

template <bool RECORD, bool DM>
__global__ __launch_bounds__(64 * K3_NW_F32, K3_WAVES_F32) void k_contacts_f32(DevParams P, SphereBuf<float> S, ClumpBuf<float> C,
                                                         ListBuf<float> L, int p, WallBuf<float> WL, int len, int nwb,
                                                         PeneState* pst, double* pene_partial,
                                                         unsigned long long* counters, int* bad, RecordBuf rec,
                                                         int remap, const int* halt, int n) {
    step3_body<float, RECORD, DM, K3_NW_F32>(P, S, C, L, p, WL, len, nwb, pst, pene_partial, counters, bad, rec, remap,
                                            halt, n);
}

template <bool RECORD, bool DM>
__global__ __launch_bounds__(64 * K3_NW_F64, K3_WAVES_F64) void k_contacts_f64(DevParams P, SphereBuf<double> S, ClumpBuf<double> C,
                                                      ListBuf<double> L, int p, WallBuf<double> WL, int len, int nwb,
                                                      PeneState* pst, double* pene_partial,
                                                      unsigned long long* counters, int* bad, RecordBuf rec,
                                                      int remap, const int* halt, int n) {
    step3_body<double, RECORD, DM, K3_NW_F64>(P, S, C, L, p, WL, len, nwb, pst, pene_partial, counters, bad, rec,
                                              remap, halt, n);
}

// ---------------------------------------------------------------------------
// Wall list: spheres that may touch a wall or the penetrator before the list
// is rebuilt. A sphere moves at most 2 * Thresh from its position at the
// build (at most Thresh from the last rebuild, whichever came later), the
// walls do not move and the penetrator moves at most pene_move, so the
// candidate tests with that slack cover every contact until the next build.
// Spheres with a live wall history or a position that is not finite are
// listed too.
__device__ __forceinline__ bool wall_listed(const DevParams& P, const PeneState* pst, const D4& x, uint32_t wf,
                                            int i) {
    const V3<double> xi = xyz(x);
    if ((wf & 0xFFu) != 0u || !finite3(xi) || !isfinite(x.w)) return true;
    if (P.nfaces > 0 && wall_candidate(P, P.nbeds > 1 ? P.beds[P.bid[i]].W : P.W, xi, x.w, P.wl_slack)) return true;
    return P.pene_on && pene_candidate(P, pst, xi, x.w, P.wl_slack + P.pene_move);
}

__global__ void k_wlist_count(DevParams P, const D4* pos, const uint32_t* wflag, const PeneState* pst, int* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.N) return;
    cnt[i] = wall_listed(P, pst, pos[i], wflag[i], i) ? 1 : 0;
}

__global__ void k_wlist_fill(DevParams P, const D4* pos, const uint32_t* wflag, PeneState* pst, const int* off,
                             int* list, int* pos_of) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        for (int a = 0; a < 3; ++a) pst->lpos[a] = pst->pos[a];
    if (i >= P.N) return;
    if (wall_listed(P, pst, pos[i], wflag[i], i)) {
        list[off[i]] = i;
        pos_of[i] = off[i];
    } else {
        pos_of[i] = -1;
    }
}

// Penetrator reaction: FP64 sum of the per-block partials in block order, then
// src/dem_integrate_penetrator.cu:79-115.
__global__ __launch_bounds__(256) void k_pene_finalize(DevParams P, PeneState* pst, const double* partial, int nb,
                                                       int count, double* cache_f, double* cache_t, const int* halt,
                                                       int n) {
    if (*halt <= n) return;
    __shared__ double red[6][256];
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int k = 0; k < 6; ++k) v[k] += partial[6 * (size_t)b + k];
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) { pst->totF[k] = red[k][0]; pst->totT[k] = red[3 + k][0]; }
        if (P.pwait && pst->contact_state == 1) pst->contact_state = 2;
        const V3<double> reaction = -1.0 * mk<double>(pst->totF[0], pst->totF[1], pst->totF[2]);
        if (P.pmotion == NBC_PENE_GRAVITY && (!P.pwait || pst->contact_state != 0)) {
            V3<double> pf = (1.0 / P.pmass) * reaction;
            pf.z = pf.z + P.Gravity;
            const double h = 0.5 * P.dt;
            pst->vel[0] = pst->vel[0] + h * pf.x;
            pst->vel[1] = pst->vel[1] + h * pf.y;
            pst->vel[2] = pst->vel[2] + h * pf.z;
        }
        cache_f[3 * count] = reaction.x; cache_f[3 * count + 1] = reaction.y; cache_f[3 * count + 2] = reaction.z;
        cache_t[3 * count] = pst->totT[0]; cache_t[3 * count + 1] = pst->totT[1]; cache_t[3 * count + 2] = pst->totT[2];
    }
}

// ---------------------------------------------------------------------------
// K4: clump force/torque as the sum of its spheres in ascending sorted order
// (FP64 accumulation), then the second half-kick.
// A wall-list sphere's record holds its raw sphere-sphere sums (.w = 1 on the
// torque): add the wall and penetrator terms the wall blocks stored at the
// same member slot and the clump torque term, in the order of the other
// spheres' totals.
template <class R>
__device__ __forceinline__ void finish_listed(const DevParams& P, const SphereBuf<R>& S, const ClumpBuf<R>& C,
                                              const WallBuf<R>& WL, const RecordBuf& rec, int record, int k,
                                              const int* mem, int c, typename Vec4T<R>::type& f,
                                              typename Vec4T<R>::type& t) {
    const typename Vec4T<R>::type* w4 = WL.res + 5 * (size_t)k;  // member slot k
    const V3<R> Fss = Ld4<R>::xyz(f), Tss = Ld4<R>::xyz(t);
    const V3<R> Fw = Ld4<R>::xyz(w4[0]), Tw = Ld4<R>::xyz(w4[1]), Fp = Ld4<R>::xyz(w4[2]), Tp = Ld4<R>::xyz(w4[3]);
    const V3<R> tF = Fss + Fw + Fp;
    const V3<R> tT = Tss + Tw + Tp;
    const V3<R> r = Ld4<R>::xyz(w4[4]);
    const int m = record ? mem[k] : 0;
    const V3<R> Tq = mk<R>(tT.x + (r.y * tF.z - r.z * tF.y), tT.y + (r.z * tF.x - r.x * tF.z),
                           tT.z + (r.x * tF.y - r.y * tF.x));
    f = mk4<R>(tF.x, tF.y, tF.z, R(0));
    t = mk4<R>(Tq.x, Tq.y, Tq.z, R(0));
    if (record) {
        D4 a; a.x = tF.x; a.y = tF.y; a.z = tF.z; a.w = 0.0;
        D4 b; b.x = tT.x; b.y = tT.y; b.z = tT.z; b.w = 0.0;
        rec.tot[2 * (size_t)m] = a;
        rec.tot[2 * (size_t)m + 1] = b;
        const V3<R> parts[6] = {Fss, Tss, Fw, Tw, Fp, Tp};
        for (int u = 0; u < 6; ++u) {
            D4 q; q.x = parts[u].x; q.y = parts[u].y; q.z = parts[u].z; q.w = 0.0;
            rec.parts[6 * (size_t)m + u] = q;
        }
    }
}

// The member records of a wave's 64 clumps are one contiguous run of member
// slots (members are sorted by clump). The wave reads the run in chunks of
// K4Chunk records with coalesced loads (lane k takes record base + k, then
// base + 64 + k, ...), finishes the wall-list records among them and leaves
// the chunk in LDS; each lane then adds its clump's records of the chunk in
// slot order, carrying the sums from chunk to chunk. The addition order is
// the per-lane one (first member starts the sum, the rest in slot order).
// Launched with one wave per block.
template <class R>
struct K4Chunk {
    static constexpr int value = sizeof(R) == sizeof(double) ? 192 : 384;  // 12 KB of LDS either way
};

template <class R, bool NEXT>
__global__ __launch_bounds__(64) void k_clump_after(DevParams P, SphereBuf<R> S, ClumpBuf<R> C, WallBuf<R> WL, RecordBuf rec, int record,
                              const int* mbeg, const int* mend, const int* mem, int* halt, int n_arg, PeneState* pst,
                              unsigned long long* counters) {
    typedef typename Vec4T<R>::type R4;
    constexpr int CH = K4Chunk<R>::value;
    __shared__ R4 s_rec[CH][2];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = lane_id();
    // the wave's member run [r0, r1) from its first and last clump, requested
    // first, so that the member records can be requested while the clump
    // state below is still on its way (over the lanes when one of the two
    // clumps has no members)
    const int c0 = blockIdx.x * blockDim.x, cl = min(c0 + (int)blockDim.x, P.Nc) - 1;
    int r0 = mbeg[c0], r1 = mend[cl];
    const bool valid = c < P.Nc;
    // clump state: independent of the member sums (loaded without a branch,
    // from a clamped index, so that no wait for it precedes the member loads)
    const int cc = valid ? c : P.Nc - 1;
    const int bb = mbeg[cc], ee = mend[cc];
    const D4 pr = C.prop[cc];
    const D4 vr = C.vw[2 * cc], wr = C.vw[2 * cc + 1];
    D4 pc, qc;
    if (NEXT) { pc = C.pq[2 * cc]; qc = C.pq[2 * cc + 1]; }
    // the halt check overlaps these loads (the arrays are valid whatever it says)
    const int n = step_of(P, n_arg);
    const int hv = *halt;
    __builtin_amdgcn_sched_barrier(0);
    if (hv <= n) return;  // the same value for the whole wave
    // NEXT: K1 of step n + 1 follows in the same thread (it returns when K4 of
    // step n does, at a halt at or before n)
    if (NEXT && c == 0) k1_thread0(P, pst, counters, halt, n + 1);
    const int b = valid ? bb : -1, e = valid ? ee : -1;
    const bool has = b >= 0 && e > b;
    if (r0 < 0 || r1 < 0) {
        r0 = has ? b : 0x7fffffff;
        r1 = has ? e : -1;
        for (int off = 32; off > 0; off >>= 1) {
            r0 = min(r0, __shfl_xor(r0, off));
            r1 = max(r1, __shfl_xor(r1, off));
        }
    }
    V3<double> Fs = mk<double>(0.0, 0.0, 0.0), Ts = Fs;
    auto chunk = [&](int base) __attribute__((always_inline)) {
        const int cnt = min(CH, r1 - base);
        R4 fr[CH / 64], tr[CH / 64];
        // clamped indices: the loads need no branch, so nothing waits for the
        // clump state before they issue
#pragma unroll
        for (int r = 0; r < CH / 64; ++r) {
            const size_t k = (size_t)min(base + r * 64 + lane, r1 - 1);
            fr[r] = S.sft[2 * k];
            tr[r] = S.sft[2 * k + 1];
        }
#pragma unroll
        for (int r = 0; r < CH / 64; ++r) {
            const int k = r * 64 + lane;
            if (k < cnt) {
                if (tr[r].w != R(0)) finish_listed<R>(P, S, C, WL, rec, record, base + k, mem, 0, fr[r], tr[r]);
                s_rec[k][0] = fr[r];
                s_rec[k][1] = tr[r];
            }
        }
        __syncthreads();
        if (has) {
            const int lo = max(b, base), hi = min(e, base + cnt);
            for (int k = lo; k < hi; ++k) {
                const R4 f = s_rec[k - base][0], t = s_rec[k - base][1];
                const V3<double> fu = mk<double>(f.x, f.y, f.z), tu = mk<double>(t.x, t.y, t.z);
                if (k == b) { Fs = fu; Ts = tu; }
                else { Fs = Fs + fu; Ts = Ts + tu; }
            }
        }
        __syncthreads();
    };
    if (r0 < r1) {
        chunk(r0);  // the first chunk outside the loop: straight-line after the loads above
        for (int base = r0 + CH; base < r1; base += CH) chunk(base);
    }
    if (!valid) return;
    D4 Fr, Tr;
    D4 F1r, T1r;  // what the next K1 reads (C.ft)
    if (has) {
        Fr.x = Fs.x; Fr.y = Fs.y; Fr.z = Fs.z; Fr.w = 0.0;
        Tr.x = Ts.x; Tr.y = Ts.y; Tr.z = Ts.z; Tr.w = 0.0;
        if (P.driver && P.pene_on) {
            // DEMmain loops: the next drift's half-kick and the attitude leave
            // the penetrator term out (src/dem_integrate.cu:23-25, 74-75); the
            // clump is its one sphere, whose contact (+ wall) totals these are
            const R4 f0 = S.sft[2 * (size_t)b], t0 = S.sft[2 * (size_t)b + 1];
            V3<R> a = Ld4<R>::xyz(f0), q = Ld4<R>::xyz(t0);
            if (t0.w != R(0)) {
                const R4* w4 = WL.res + 5 * (size_t)b;
                a = a + Ld4<R>::xyz(w4[0]);
                q = q + Ld4<R>::xyz(w4[1]);
            }
            F1r.x = a.x; F1r.y = a.y; F1r.z = a.z; F1r.w = 0.0;
            T1r.x = q.x; T1r.y = q.y; T1r.z = q.z; T1r.w = 0.0;
        } else {
            F1r = Fr;
            T1r = Tr;
        }
        // a K1 launched on its own reads C.ft; the fused one below takes the
        // values from registers, and the last step of a call runs without it
        if (!NEXT) {
            C.ft[2 * c] = F1r;
            C.ft[2 * c + 1] = T1r;
        }
    } else {
        Fr = C.ft[2 * c];
        Tr = C.ft[2 * c + 1];
        F1r = Fr;
        T1r = Tr;
    }
    V3<double> F = xyz(Fr), T = xyz(Tr);
    scale_ft(P, pr, F, T);
    // glue kinematic types keep their velocities (src/dem_integrate_withglue.cu:139)
    const bool kick = P.driver != NBC_DRIVER_SPLIT || C.type[c] == 0;
    const V3<double> v = kick ? xyz(vr) + 0.5 * P.dt * F : xyz(vr);
    const V3<double> w = kick ? xyz(wr) + 0.5 * P.dt * T : xyz(wr);
    if (NEXT) {
        // K1 of step n + 1 reads this force, torque and velocity (what it
        // stored in C.ft, for the DEMmain loops with a penetrator)
        if (P.driver && P.pene_on) {
            V3<double> F1 = xyz(F1r), T1 = xyz(T1r);
            scale_ft(P, pr, F1, T1);
            clump_drift<R>(P, C, c, pc, qc, v, w, F1, T1, halt);
            return;
        }
        clump_drift<R>(P, C, c, pc, qc, v, w, F, T, halt);
        return;
    }
    D4 vo; vo.x = v.x; vo.y = v.y; vo.z = v.z; vo.w = 0.0;
    D4 wo; wo.x = w.x; wo.y = w.y; wo.z = w.z; wo.w = 0.0;
    C.vw[2 * c] = vo;
    C.vw[2 * c + 1] = wo;
}

// Member slot of every sphere: the sphere's position in the clump member
// order, where k_contacts writes its totals for k_clump_after.
__global__ void k_member_slot(int N, const int* mem, int* slot) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) slot[mem[k]] = k;
}

// Kinetic energy partial sums (FP64), one value per block.
__global__ __launch_bounds__(256) void k_kinetic(int Nc, const D4* vw, const D4* prop, double* partial) {
    __shared__ double red[256];
    const int c = blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (c < Nc) {
        const D4 v = vw[2 * c], w = vw[2 * c + 1], p = prop[c];
        e = 0.5 * p.w * (v.x * v.x + v.y * v.y + v.z * v.z) + 0.5 * (p.x * w.x * w.x + p.y * w.y * w.y + p.z * w.z * w.z);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ===========================================================================
// host-side launchers
static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

#define LAUNCH_CHECK()                                    \
    do {                                                  \
        hipError_t e_ = hipGetLastError();                \
        if (e_ != hipSuccess) return (int)e_;             \
    } while (0)

__global__ void k_set_int(int* a, int v) { *a = v; }

__global__ void k_clear_halt(int* status, int none) {
    status[0] = none;
    status[3] &= ST_CLUMP_BAD;
}
int launch_set_int(hipStream_t st, int* a, int v) {
    k_set_int<<<1, 1, 0, st>>>(a, v);
    LAUNCH_CHECK();
    return 0;
}

int launch_clear_halt(hipStream_t st, int* status, int none) {
    k_clear_halt<<<1, 1, 0, st>>>(status, none);
    LAUNCH_CHECK();
    return 0;
}
int launch_fill_i32(hipStream_t st, int* a, long long n, int v) {
    if (n <= 0) return 0;
    k_fill_i32<<<nblk(n, 256), 256, 0, st>>>(a, n, v);
    LAUNCH_CHECK();
    return 0;
}
int launch_fill_u32(hipStream_t st, uint32_t* a, long long n, uint32_t v) {
    if (n <= 0) return 0;
    k_fill_u32<<<nblk(n, 256), 256, 0, st>>>(a, n, v);
    LAUNCH_CHECK();
    return 0;
}
int launch_fill_f64(hipStream_t st, double* a, long long n, double v) {
    if (n <= 0) return 0;
    k_fill_f64<<<nblk(n, 256), 256, 0, st>>>(a, n, v);
    LAUNCH_CHECK();
    return 0;
}

int launch_keys_from_int(hipStream_t st, const int* a, uint32_t* k, int* v, int n) {
    if (n <= 0) return 0;
    k_keys_from_int<<<nblk(n, 256), 256, 0, st>>>(a, k, v, n);
    LAUNCH_CHECK();
    return 0;
}

int scan_tmp_ints(int n) { return (int)nblk(n, 2048) + 1; }

int launch_exclusive_scan(hipStream_t st, const int* in, int* out, int n, int* tmp, int* total) {
    const int nb = (int)nblk(n, 2048);
    if (n <= 0) return launch_fill_i32(st, total, 1, 0);
    k_scan_reduce<<<nb, 256, 0, st>>>(in, n, tmp);
    LAUNCH_CHECK();
    k_scan_top<<<1, 256, 0, st>>>(tmp, nb, total);
    LAUNCH_CHECK();
    k_scan_down<<<nb, 256, 0, st>>>(in, out, n, tmp);
    LAUNCH_CHECK();
    return 0;
}

int radix_tmp_ints(int n, int rbits) { return (1 << rbits) * (int)nblk(n, 256); }

// Sorts (keys, vals) by the low `bits` bits of the keys. Result in (k0, v0)
// when the pass count is even, else in (k1, v1); *in_second reports which.
int launch_radix_sort(hipStream_t st, uint32_t* k0, int* v0, uint32_t* k1, int* v1, int n, int bits, int rbits,
                      int* hist, int* hscan, int* stmp, int* stotal, int* in_second) {
    const int nb = (int)nblk(n, 256);
    if (rbits < 1) rbits = 1;
    if (rbits > 8) rbits = 8;
    int passes = (bits + rbits - 1) / rbits;
    if (passes < 1) passes = 1;
    uint32_t* ki = k0; int* vi = v0; uint32_t* ko = k1; int* vo = v1;
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = ps * rbits;
        const int rb = (shift + rbits <= 32) ? rbits : 32 - shift;
        const int nh = (1 << rb) * nb;
        k_rs_hist<<<nb, 256, 0, st>>>(ki, n, shift, rb, hist, nb);
        LAUNCH_CHECK();
        int e = launch_exclusive_scan(st, hist, hscan, nh, stmp, stotal);
        if (e) return e;
        k_rs_scatter<<<nb, 256, 0, st>>>(ki, vi, ko, vo, n, shift, rb, hscan, nb);
        LAUNCH_CHECK();
        uint32_t* tk = ki; ki = ko; ko = tk;
        int* tv = vi; vi = vo; vo = tv;
    }
    *in_second = (passes & 1) ? 1 : 0;
    return 0;
}

int launch_hash(hipStream_t st, const DevParams& P, const D4* pos, const int* cid, uint32_t* key, int* val) {
    k_hash<<<nblk(P.N, 256), 256, 0, st>>>(P, pos, cid, key, val);
    LAUNCH_CHECK();
    return 0;
}

int launch_seg_bounds(hipStream_t st, const uint32_t* key, int n, int nkeys, int* beg, int* end) {
    int e = launch_fill_i32(st, beg, nkeys, -1);
    if (e) return e;
    e = launch_fill_i32(st, end, nkeys, -1);
    if (e || n <= 0) return e;
    k_seg_bounds<<<nblk(n, 256), 256, 0, st>>>(key, n, nkeys, beg, end);
    LAUNCH_CHECK();
    return 0;
}

int launch_neighbors(hipStream_t st, const DevParams& P, bool fill, const D4* pos, const uint32_t* cell,
                     const int* cbeg, const int* cend, const int* cid, int* fcnt, int* bcnt, const int* foff,
                     const int* boff, uint32_t* fword, int2* bent, unsigned* overflow, int block) {
    if (fill)
        k_neighbors<true><<<nblk(P.N, block), block, 0, st>>>(P, pos, cell, cbeg, cend, cid, fcnt, bcnt, foff, boff,
                                                              fword, bent, overflow);
    else
        k_neighbors<false><<<nblk(P.N, block), block, 0, st>>>(P, pos, cell, cbeg, cend, cid, fcnt, bcnt, foff, boff,
                                                               fword, bent, overflow);
    LAUNCH_CHECK();
    return 0;
}

int launch_hneighbors(hipStream_t st, const DevParams& P, int phase, const D4* pos, const uint32_t* cell,
                      const int* cbeg, const int* cend, const int* cid, int* fcnt, int* bcnt, int* xcnt,
                      const int* xoff, uint32_t* xkey, int* xval, const int* foff, const int* boff, const int* xbeg,
                      const int* xend, const int* xsrt, uint32_t* fword, int2* bent, unsigned* overflow, int block) {
    const unsigned g = nblk(P.N, block);
    if (phase == 0)
        k_hneighbors<0><<<g, block, 0, st>>>(P, pos, cell, cbeg, cend, cid, fcnt, bcnt, xcnt, xoff, xkey, xval, foff,
                                             boff, xbeg, xend, xsrt, fword, bent, overflow);
    else if (phase == 1)
        k_hneighbors<1><<<g, block, 0, st>>>(P, pos, cell, cbeg, cend, cid, fcnt, bcnt, xcnt, xoff, xkey, xval, foff,
                                             boff, xbeg, xend, xsrt, fword, bent, overflow);
    else
        k_hneighbors<2><<<g, block, 0, st>>>(P, pos, cell, cbeg, cend, cid, fcnt, bcnt, xcnt, xoff, xkey, xval, foff,
                                             boff, xbeg, xend, xsrt, fword, bent, overflow);
    LAUNCH_CHECK();
    return 0;
}

int launch_neighbors_ell(hipStream_t st, const DevParams& P, const D4* pos, const uint32_t* cell, const int* cbeg,
                         const int* cend, const int* cid, int K, int* fcnt, int* bcnt, uint32_t* ef, int* eb,
                         unsigned* overflow, unsigned* ellmax, int block) {
    (void)cell; (void)block;
    k_neighbors_cell<<<nblk(P.CellNum, 4), 256, 0, st>>>(P, pos, cbeg, cend, cid, K, fcnt, bcnt, ef, eb, overflow,
                                                         ellmax);
    LAUNCH_CHECK();
    return 0;
}

int launch_rows_from_ell(hipStream_t st, int N, const int* fcnt, const int* bcnt, const int* foff, const int* boff,
                         const uint32_t* ef, const int* eb, uint32_t* fword, int2* bent, int block) {
    (void)fcnt; (void)bcnt; (void)block;
    k_rows_from_ell<<<nblk(N, 256), 256, 0, st>>>(N, foff, boff, ef, eb, fword, bent);
    LAUNCH_CHECK();
    return 0;
}

int launch_hneighbors_ell(hipStream_t st, const DevParams& P, const D4* pos, const uint32_t* cell, const int* cbeg,
                          const int* cend, const int* cid, const int* sidx, int K, int* fcnt, int* bcnt, int* xcnt,
                          uint32_t* ef, int* eb, int* ex, unsigned* overflow, unsigned* ellmax, int block) {
    k_hneighbors_ell<<<nblk(P.N, block), block, 0, st>>>(P, pos, cell, cbeg, cend, cid, sidx, K, fcnt, bcnt, xcnt, ef,
                                                          eb, ex, overflow, ellmax);
    LAUNCH_CHECK();
    return 0;
}

int launch_cross_from_ell(hipStream_t st, int N, const int* sidx, const int* xcnt, const int* xoff, const int* ex,
                          uint32_t* xkey, int* xval, int block) {
    k_cross_from_ell<<<nblk(N, block), block, 0, st>>>(N, sidx, xcnt, xoff, ex, xkey, xval);
    LAUNCH_CHECK();
    return 0;
}

int launch_hrows_from_ell(hipStream_t st, int N, const int* sidx, const int* cid, const int* foff, const int* bcnt,
                          const int* boff, const int* xbeg, const int* xend, const int* xsrt, const uint32_t* ef,
                          const int* eb, uint32_t* fword, int2* bent, int block) {
    k_hrows_from_ell<<<nblk(N, block), block, 0, st>>>(N, sidx, cid, foff, bcnt, boff, xbeg, xend, xsrt, ef, eb, fword,
                                                        bent);
    LAUNCH_CHECK();
    return 0;
}

int launch_level_keys(hipStream_t st, const DevParams& P, const D4* pos, const int* cid, uint32_t* key, int* val,
                      int block) {
    k_level_keys<<<nblk(P.N, block), block, 0, st>>>(P, pos, cid, key, val);
    LAUNCH_CHECK();
    return 0;
}

int launch_hfront_total(hipStream_t st, int N, int half, int* fcnt, const int* bcnt, const int* xbeg,
                        const int* xend, unsigned* overflow, int block) {
    k_hfront_total<<<nblk(N, block), block, 0, st>>>(N, half, fcnt, bcnt, xbeg, xend, overflow);
    LAUNCH_CHECK();
    return 0;
}

int launch_near_build(hipStream_t st, const DevParams& P, const D4* pos, D4* cum, float* nthr, double* slack,
                      const int* cid, const D4* pq, const D4* vw, const int* foff, const uint32_t* fword, const int* boff, const int2* bent, int* ncnt, int* nfoff,
                      uint32_t* nfw, int* nfs, int* slot2near, int* nbcnt, int* nboff, int* nbo, int* nbe, int* stmp,
                      int* stotal, int block) {
    const int N = P.N;
    const unsigned g = nblk(N, block);
    int e;
    if ((e = launch_fill_i32(st, ncnt + N, 1, 0))) return e;
    if ((e = launch_fill_i32(st, nbcnt + N, 1, 0))) return e;
    k_near_slack<<<g, block, 0, st>>>(P, pos, cum, nthr, slack, cid, pq, vw);
    LAUNCH_CHECK();
    const unsigned gw = nblk(N, 256);  // 4 waves of 64 spheres per block
    k_near_scan<<<gw, 256, 0, st>>>(P, pos, slack, foff, fword, ncnt, slot2near);
    LAUNCH_CHECK();
    if ((e = launch_exclusive_scan(st, ncnt, nfoff, N + 1, stmp, stotal))) return e;
    k_near_put<<<gw, 256, 0, st>>>(N, foff, fword, nfoff, nfw, nfs, slot2near);
    LAUNCH_CHECK();
    k_nback_scan<<<gw, 256, 0, st>>>(N, boff, bent, slot2near, nbcnt);
    LAUNCH_CHECK();
    if ((e = launch_exclusive_scan(st, nbcnt, nboff, N + 1, stmp, stotal + 1))) return e;
    k_nback_put<<<gw, 256, 0, st>>>(N, boff, bent, slot2near, nboff, nbo, nbe);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_near_sync(hipStream_t st, int N, const int* nfoff, const uint32_t* nfw, const int* nfs, uint32_t* fword,
                     const R* hn, R* hs, int q, int block) {
    k_near_sync<R><<<nblk(N, block), block, 0, st>>>(N, nfoff, nfw, nfs, fword, hn, hs, q);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_near_hist_gather(hipStream_t st, int nnear, const uint32_t* nfw, const int* nfs, const R* hs, R* hn, int q) {
    if (nnear <= 0) return 0;
    k_near_hist_gather<R><<<nblk(nnear, 256), 256, 0, st>>>(nnear, nfw, nfs, hs, hn, q);
    LAUNCH_CHECK();
    return 0;
}
int launch_back_resolve(hipStream_t st, int N, const int* foff, const uint32_t* fword, const int* boff, int2* bent,
                        int block) {
    (void)block;
    k_back_resolve<<<nblk(N, 256), 256, 0, st>>>(N, foff, fword, boff, bent);
    LAUNCH_CHECK();
    return 0;
}

int launch_kinetic(hipStream_t st, int Nc, const D4* vw, const D4* prop, double* partial) {
    k_kinetic<<<nblk(Nc, 256), 256, 0, st>>>(Nc, vw, prop, partial);
    LAUNCH_CHECK();
    return 0;
}

int launch_pene_finalize(hipStream_t st, const DevParams& P, PeneState* pst, const double* partial, int nb, int count,
                         double* cf, double* ct, const int* halt, int n) {
    k_pene_finalize<<<1, 256, 0, st>>>(P, pst, partial, nb, count, cf, ct, halt, n);
    LAUNCH_CHECK();
    return 0;
}

template <class R>
int launch_clump_before(hipStream_t st, const DevParams& P, ClumpBuf<R> C, PeneState* pst, unsigned long long* ctr,
                        int block, int* halt, int step) {
    const int n = P.Nc > 1 ? P.Nc : 1;
    k_clump_before<R><<<nblk(n, block), block, 0, st>>>(P, C, pst, ctr, halt, step);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_sphere_pose(hipStream_t st, const DevParams& P, SphereBuf<R> S, ClumpBuf<R> C, int* halt, int step,
                       int block) {
    k_sphere_pose<R><<<nblk(P.N, block), block, 0, st>>>(P, S, C, halt, step);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_gather(hipStream_t st, const DevParams& P, int nface, const int* perm, SphereBuf<R> src, SphereBuf<R> dst) {
    k_gather<R><<<nblk(P.N, 256), 256, 0, st>>>(P, nface, perm, src, dst);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_hist_push(hipStream_t st, const DevParams& P, const D4* pos, const int* cid, int p, const int* perm,
                     int* inv, const int* onfoff, const uint32_t* onfw, const R* ohn, const int* nfoff,
                     uint32_t* nfword, R* nhist, int block) {
    if (P.N <= 0) return 0;
    k_inverse_perm<<<nblk(P.N, block), block, 0, st>>>(P.N, perm, inv);
    LAUNCH_CHECK();
    k_hist_push<R><<<nblk(P.N, block), block, 0, st>>>(P, pos, cid, p, inv, onfoff, onfw, ohn, nfoff, nfword, nhist);
    LAUNCH_CHECK();
    return 0;
}
int launch_clump_renumber_first(hipStream_t st, int N, const int* cid, int* first, int block) {
    k_clump_first<<<nblk(N, block), block, 0, st>>>(N, cid, first);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_clump_gather(hipStream_t st, int Nc, const int* order, ClumpBuf<R> src, ClumpBuf<R> dst,
                        const int* cfile, int* cfile_out, int* newid, int* cid, int N, int block) {
    k_clump_gather<R><<<nblk(Nc, block), block, 0, st>>>(Nc, order, src, dst, cfile, cfile_out, newid);
    LAUNCH_CHECK();
    k_remap_cid<<<nblk(N, block), block, 0, st>>>(N, cid, newid);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_hist_carry(hipStream_t st, long long nnew, long long nold, int p, uint32_t* nfword, R* nhist,
                      const uint32_t* ofword, const R* ohist, int block) {
    if (nnew <= 0) return 0;
    k_hist_carry<R><<<nblk(nnew, block), block, 0, st>>>(nnew, nold, p, nfword, nhist, ofword, ohist);
    LAUNCH_CHECK();
    return 0;
}
int contacts_blocks(int N, int block) { return (int)nblk(N, block); }
template <class R>
int launch_contacts(hipStream_t st, const DevParams& P, SphereBuf<R> S, ClumpBuf<R> C, ListBuf<R> L, int p,
                    WallBuf<R> WL, int wlen, PeneState* pst, double* partial, unsigned long long* ctr, int* bad,
                    RecordBuf rec, bool record, int remap, const int* halt, int step) {
    const int bs = 64 * (sizeof(R) == sizeof(float) ? K3_NW_F32 : K3_NW_F64);
    const int nwb = wlen > 0 ? (int)nblk(wlen, bs) : 0;
    const unsigned nb = nblk(P.N, bs) + (unsigned)nwb;
#define NBC_K3(K, REC, DM) K<REC, DM><<<nb, bs, 0, st>>>(P, S, C, L, p, WL, wlen, nwb, pst, partial, ctr, bad, rec, \
                                                          remap, halt, step)
    const bool dm = P.driver != NBC_DRIVER_CLUMP;
    if constexpr (sizeof(R) == sizeof(float)) {
        if (dm) { if (record) NBC_K3(k_contacts_f32, true, true); else NBC_K3(k_contacts_f32, false, true); }
        else { if (record) NBC_K3(k_contacts_f32, true, false); else NBC_K3(k_contacts_f32, false, false); }
    } else {
        if (dm) { if (record) NBC_K3(k_contacts_f64, true, true); else NBC_K3(k_contacts_f64, false, true); }
        else { if (record) NBC_K3(k_contacts_f64, true, false); else NBC_K3(k_contacts_f64, false, false); }
    }
#undef NBC_K3
    LAUNCH_CHECK();
    return 0;
}
int launch_wlist_build(hipStream_t st, const DevParams& P, const D4* pos, const uint32_t* wflag, PeneState* pst,
                       int* cnt, int* off, int* list, int* pos_of, int* stmp, int* stotal, int block) {
    const unsigned g = nblk(P.N, block);
    int e;
    if ((e = launch_fill_i32(st, cnt + P.N, 1, 0))) return e;
    k_wlist_count<<<g, block, 0, st>>>(P, pos, wflag, pst, cnt);
    LAUNCH_CHECK();
    if ((e = launch_exclusive_scan(st, cnt, off, P.N + 1, stmp, stotal))) return e;
    k_wlist_fill<<<g, block, 0, st>>>(P, pos, wflag, pst, off, list, pos_of);
    LAUNCH_CHECK();
    return 0;
}
template <class R>
int launch_clump_after(hipStream_t st, const DevParams& P, SphereBuf<R> S, ClumpBuf<R> C, WallBuf<R> WL,
                       RecordBuf rec, bool record, const int* mbeg, const int* mend, const int* mem, int block,
                       int* halt, int step, bool next, PeneState* pst, unsigned long long* counters) {
    // one wave per block: there are few clumps (C1: 200,000 = 3,125 waves), and
    // 64-thread blocks spread them over the CUs with a tail of one wave instead
    // of one 256-thread block (K4 34 vs 39 us on C1)
    (void)block;
    const int kb = 64;
    const unsigned g = nblk(P.Nc, kb);
    if (next)
        k_clump_after<R, true><<<g, kb, 0, st>>>(P, S, C, WL, rec, record ? 1 : 0, mbeg, mend, mem, halt, step, pst,
                                                 counters);
    else
        k_clump_after<R, false><<<g, kb, 0, st>>>(P, S, C, WL, rec, record ? 1 : 0, mbeg, mend, mem, halt, step,
                                                  pst, counters);
    LAUNCH_CHECK();
    return 0;
}

int launch_member_slot(hipStream_t st, int N, const int* mem, int* slot) {
    k_member_slot<<<nblk(N, 256), 256, 0, st>>>(N, mem, slot);
    LAUNCH_CHECK();
    return 0;
}

#define INST(R)                                                                                                      \
    template int launch_clump_before<R>(hipStream_t, const DevParams&, ClumpBuf<R>, PeneState*, unsigned long long*, \
                                        int, int*, int);                                                             \
    template int launch_sphere_pose<R>(hipStream_t, const DevParams&, SphereBuf<R>, ClumpBuf<R>, int*, int, int);    \
    template int launch_gather<R>(hipStream_t, const DevParams&, int, const int*, SphereBuf<R>, SphereBuf<R>);                    \
    template int launch_near_sync<R>(hipStream_t, int, const int*, const uint32_t*, const int*, uint32_t*,          \
                                     const R*, R*, int, int);                                                       \
    template int launch_near_hist_gather<R>(hipStream_t, int, const uint32_t*, const int*, const R*, R*, int);      \
    template int launch_hist_push<R>(hipStream_t, const DevParams&, const D4*, const int*, int, const int*, int*,     \
                                     const int*, const uint32_t*, const R*, const int*, uint32_t*, R*, int);         \
    template int launch_clump_gather<R>(hipStream_t, int, const int*, ClumpBuf<R>, ClumpBuf<R>, const int*, int*,   \
                                        int*, int*, int, int);                                                         \
    template int launch_hist_carry<R>(hipStream_t, long long, long long, int, uint32_t*, R*, const uint32_t*,       \
                                      const R*, int);                                                                \
    template int launch_contacts<R>(hipStream_t, const DevParams&, SphereBuf<R>, ClumpBuf<R>, ListBuf<R>, int,       \
                                    WallBuf<R>, int, PeneState*, double*, unsigned long long*, int*, RecordBuf,      \
                                    bool, int, const int*, int);                                                     \
    template int launch_clump_after<R>(hipStream_t, const DevParams&, SphereBuf<R>, ClumpBuf<R>, WallBuf<R>,         \
                                       RecordBuf, bool, const int*, const int*, const int*, int, int*, int, bool,     \
                                       PeneState*, unsigned long long*);
INST(double)
INST(float)

}  // namespace nbc
