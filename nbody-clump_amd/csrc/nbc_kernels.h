// Device buffer views and launcher declarations shared by the kernels and the
// engine. All launchers return 0 or a hipError_t value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nbc.h"
#include "nbc_device.h"

namespace nbc {

// Per-sphere arrays, sorted order. pos.w holds the FP64 radius so that one
// 32-byte load gives a neighbour's centre and size; loc/kin .w hold the mass.
template <class R>
struct SphereBuf {
    D4* pos;                          // x y z radius (FP64)
    D4* cum;                          // displacement since the last rebuild (FP64); with ref_disp the
                                      // position at the last rebuild, w = radius
    typename Vec4T<R>::type* loc;     // local offset (body frame), mass
    typename Vec4T<R>::type* kin;     // R(q) * local (world frame), mass
    int* cid;                         // clump index
    int* fid;                         // file index (part_fix)
    uint32_t* wflag;                  // bit f: wall/penetrator history f valid
    R* wh;                            // [MAX_FACES][N][HSTRIDE] wall/penetrator histories
    typename Vec4T<R>::type* sft;     // [2N] force, torque about the clump centre, in member-slot order
    int* slot;                        // member slot of the sphere (its index in the clump member order)
    float* nthr;                      // with ref_disp: near-list request threshold (|disp|^2), rounded down
};

// FP32-mixed clump loops keep the position at the last rebuild instead of the
// accumulated displacement: the displacement is x - x_rebuild, so the pose
// kernel neither reads the previous position nor writes the displacement
// back (68 instead of 128 bytes of pose traffic per sphere). The FP64 form and
// the DEMmain loops accumulate it as the reference does
// (src/dem_clump.cu:139-141), so their rebuild steps stay bit-exact.
template <class R>
__host__ __device__ inline bool ref_disp(int driver) { return sizeof(R) == 4 && driver == 0; }

// Per-clump arrays (FP64 state; snap is the contact kernel's copy in R).
template <class R>
struct ClumpBuf {
    D4* pq;                           // [2Nc] position, quaternion (x y z w)
    D4* vw;                           // [2Nc] velocity, angular velocity
    D4* prop;                         // Ix Iy Iz mass
    D4* ft;                           // [2Nc] force, torque (FP64)
    typename Vec4T<R>::type* snap;    // [2Nc] velocity (w = state finite), angular velocity
    D4* dist;                         // [Nc] DEMmain loops: the step's drift (cum += dist), else null
    int* type;                        // [Nc] DEMmain loops: glue kinematic type, else null
};

// Verlet list: CSR front rows (owner i, partner j > i) with ping-pong history,
// CSR back rows (i, owner j < i, owner slot).
template <class R>
struct ListBuf {
    int* nfoff;                       // [N + 1] near front rows
    uint32_t* nfw;                    // [near] pair word with the history-valid bits
    int* nfs;                         // [near] front slot (history, records)
    int* nboff;                       // [N + 1] near back rows
    int* nbo;                         // [near] owner of the back entry's pair
    int* nbe;                         // [near] the pair's index in the owner's near row
    R* hist[2];                       // [near front entries][HSTRIDE] ping-pong histories in near order
                                      // (in front-slot order outside the step kernels)
};

// Wall list: spheres that may touch a wall or the penetrator until the list
// is rebuilt, in sphere order, with their wall/penetrator results per step.
template <class R>
struct WallBuf {
    int* list;                        // [len] sphere index
    int* pos_of;                      // [N] index into list, or -1
    typename Vec4T<R>::type* res;     // [5 * N] Fw, Tw, Fp, Tp, clump torque arm, at the member slot
};

struct RecordBuf {
    D4* tot;       // [2N] total force, total torque (before the clump term)
    D4* parts;     // [6N] contact F,T; wall F,T; penetrator F,T
    D4* pene;      // [4N] penetrator Fn, Ft, Tn, Tt
    D4* pair_fcn;  // [front slots] -Fn of the pair (CacheFcn)
    D4* pair_cpt;  // [front slots] contact point (CacheCtp)
};

int launch_set_int(hipStream_t st, int* a, int v);  // one-thread store (no memset: safe in graphs)
int launch_clear_halt(hipStream_t st, int* status, int none);
int launch_fill_i32(hipStream_t st, int* a, long long n, int v);
int launch_fill_u32(hipStream_t st, uint32_t* a, long long n, uint32_t v);
int launch_fill_f64(hipStream_t st, double* a, long long n, double v);
int scan_tmp_ints(int n);
int launch_exclusive_scan(hipStream_t st, const int* in, int* out, int n, int* tmp, int* total);
int radix_tmp_ints(int n, int rbits);
int launch_radix_sort(hipStream_t st, uint32_t* k0, int* v0, uint32_t* k1, int* v1, int n, int bits, int rbits,
                      int* hist, int* hscan, int* stmp, int* stotal, int* in_second);
int launch_keys_from_int(hipStream_t st, const int* a, uint32_t* k, int* v, int n);
int launch_hash(hipStream_t st, const DevParams& P, const D4* pos, const int* cid, uint32_t* key, int* val);
int launch_seg_bounds(hipStream_t st, const uint32_t* key, int n, int nkeys, int* beg, int* end);
int launch_neighbors(hipStream_t st, const DevParams& P, bool fill, const D4* pos, const uint32_t* cell,
                     const int* cbeg, const int* cend, const int* cid, int* fcnt, int* bcnt, const int* foff,
                     const int* boff, uint32_t* fword, int2* bent, unsigned* overflow, int block);
int launch_back_resolve(hipStream_t st, int N, const int* foff, const uint32_t* fword, const int* boff, int2* bent,
                        int block);
int launch_kinetic(hipStream_t st, int Nc, const D4* vw, const D4* prop, double* partial);
int launch_pene_finalize(hipStream_t st, const DevParams& P, PeneState* pst, const double* partial, int nb, int count,
                         double* cf, double* ct, const int* halt, int step);
int contacts_blocks(int N, int block);

// Step kernels take their absolute step index and the device "halt step":
// k_sphere_pose lowers it to the step that needs a rebuild; every later
// kernel of that step, and every kernel of later steps, returns at entry.
template <class R>
int launch_clump_before(hipStream_t st, const DevParams& P, ClumpBuf<R> C, PeneState* pst, unsigned long long* ctr,
                        int block, int* halt, int step);
template <class R>
int launch_sphere_pose(hipStream_t st, const DevParams& P, SphereBuf<R> S, ClumpBuf<R> C, int* halt, int step,
                       int block);
template <class R>
int launch_gather(hipStream_t st, const DevParams& P, int nface, const int* perm, SphereBuf<R> src, SphereBuf<R> dst);
// History inheritance at a rebuild from the old near list (onfoff/onfw,
// histories ohn in near order) into the new front list, whose pair words
// carry no valid bits yet; inv receives the inverse of perm.
template <class R>
int launch_hist_push(hipStream_t st, const DevParams& P, const D4* pos, const int* cid, int p, const int* perm,
                     int* inv, const int* onfoff, const uint32_t* onfw, const R* ohn, const int* nfoff,
                     uint32_t* nfword, R* nhist, int block);
// Clump renumbering at a rebuild: first[c] = first sphere of clump c (first
// filled with N before), then the clump arrays move to the order of first
// (order[new] = old) and every sphere's clump index follows.
int launch_clump_renumber_first(hipStream_t st, int N, const int* cid, int* first, int block);
template <class R>
int launch_clump_gather(hipStream_t st, int Nc, const int* order, ClumpBuf<R> src, ClumpBuf<R> dst,
                        const int* cfile, int* cfile_out, int* newid, int* cid, int N, int block);
// DEMmain cylinder/glue loops: new slot s takes old slot s's last history
template <class R>
int launch_hist_carry(hipStream_t st, long long nnew, long long nold, int p, uint32_t* nfword, R* nhist,
                      const uint32_t* ofword, const R* ohist, int block);
// Hierarchical grid search, phase 0 = counts, 1 = cross entries, 2 = rows.
int launch_hneighbors(hipStream_t st, const DevParams& P, int phase, const D4* pos, const uint32_t* cell,
                      const int* cbeg, const int* cend, const int* cid, int* fcnt, int* bcnt, int* xcnt,
                      const int* xoff, uint32_t* xkey, int* xval, const int* foff, const int* boff, const int* xbeg,
                      const int* xend, const int* xsrt, uint32_t* fword, int2* bent, unsigned* overflow, int block);
// single-pass neighbour search into entry-major rows of capacity K, and the
// compactions into the CSR rows (see nbc_kernels.hip)
int launch_neighbors_ell(hipStream_t st, const DevParams& P, const D4* pos, const uint32_t* cell, const int* cbeg,
                         const int* cend, const int* cid, int K, int* fcnt, int* bcnt, uint32_t* ef, int* eb,
                         unsigned* overflow, unsigned* ellmax, int block);
int launch_rows_from_ell(hipStream_t st, int N, const int* fcnt, const int* bcnt, const int* foff, const int* boff,
                         const uint32_t* ef, const int* eb, uint32_t* fword, int2* bent, int block);
int launch_hneighbors_ell(hipStream_t st, const DevParams& P, const D4* pos, const uint32_t* cell, const int* cbeg,
                          const int* cend, const int* cid, const int* sidx, int K, int* fcnt, int* bcnt, int* xcnt,
                          uint32_t* ef, int* eb, int* ex, unsigned* overflow, unsigned* ellmax, int block);
int launch_cross_from_ell(hipStream_t st, int N, const int* sidx, const int* xcnt, const int* xoff, const int* ex,
                          uint32_t* xkey, int* xval, int block);
int launch_hrows_from_ell(hipStream_t st, int N, const int* sidx, const int* cid, const int* foff, const int* bcnt,
                          const int* boff, const int* xbeg, const int* xend, const int* xsrt, const uint32_t* ef,
                          const int* eb, uint32_t* fword, int2* bent, int block);
int launch_level_keys(hipStream_t st, const DevParams& P, const D4* pos, const int* cid, uint32_t* key, int* val,
                      int block);
int launch_hfront_total(hipStream_t st, int N, int half, int* fcnt, const int* bcnt, const int* xbeg,
                        const int* xend, unsigned* overflow, int block);
// Near lists of the current list at the current positions; also sets each
// sphere's slack and its near-list request threshold (cum.w). cid, pq and vw
// (clump index, clump position, clump velocities) give the per-sphere skin
// when P.near_horizon > 0.
int launch_near_build(hipStream_t st, const DevParams& P, const D4* pos, D4* cum, float* nthr, double* slack,
                      const int* cid, const D4* pq, const D4* vw, const int* foff, const uint32_t* fword, const int* boff, const int2* bent, int* ncnt, int* nfoff,
                      uint32_t* nfw, int* nfs, int* slot2near, int* nbcnt, int* nboff, int* nbo, int* nbe, int* stmp,
                      int* stotal, int block);
template <class R>
int launch_near_sync(hipStream_t st, int N, const int* nfoff, const uint32_t* nfw, const int* nfs, uint32_t* fword,
                     const R* hn, R* hs, int q, int block);
// Latest history (buffer q) of the near pairs from front-slot order into near order.
template <class R>
int launch_near_hist_gather(hipStream_t st, int nnear, const uint32_t* nfw, const int* nfs, const R* hs, R* hn, int q);
// Sphere-sphere contacts of every sphere and, in the same launch, the wall and
// penetrator contacts of the wlen spheres on the wall list.
template <class R>
int launch_contacts(hipStream_t st, const DevParams& P, SphereBuf<R> S, ClumpBuf<R> C, ListBuf<R> L, int p,
                    WallBuf<R> WL, int wlen, PeneState* pst, double* partial, unsigned long long* ctr, int* bad,
                    RecordBuf rec, bool record, int remap, const int* halt, int step);

int launch_wlist_build(hipStream_t st, const DevParams& P, const D4* pos, const uint32_t* wflag, PeneState* pst,
                       int* cnt, int* off, int* list, int* pos_of, int* stmp, int* stotal, int block);
template <class R>
// K4 of step `step`; with `next`, K1 of step + 1 runs in the same kernel.
int launch_clump_after(hipStream_t st, const DevParams& P, SphereBuf<R> S, ClumpBuf<R> C, WallBuf<R> WL,
                       RecordBuf rec, bool record, const int* mbeg, const int* mend, const int* mem, int block,
                       int* halt, int step, bool next, PeneState* pst, unsigned long long* counters);
int launch_member_slot(hipStream_t st, int N, const int* mem, int* slot);

}  // namespace nbc
