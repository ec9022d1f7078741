/*
 * nbc.h — C ABI of the MI355X clumped-sphere DEM timestep (libnbc.so).
 *
 * Plain C: fixed-width integers, doubles, pointers and sizes. No HIP or torch
 * types cross this boundary; the library owns its device memory and its HIP
 * stream. Every entry point returns an nbc error code (0 = success) and never
 * calls exit(); nbc_last_error() gives the message. The reference aborts the
 * process in the same three situations (neighbour-list overflow, invalid or
 * non-finite wall input, bad input files); a caller that wants the reference's
 * behaviour prints the message and exits with EXIT_FAILURE, which is what the
 * bundled nbody_clump driver does.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   nbc_run_simulation_clump   DEMClumpMain::RunSimulationClump   src/dem_clump.h:95,
 *                                                                 src/dem_simulation_clump.cu:3-261
 *   nbc_load_world_par         DEMSubSphereIo::LoadParameters     src/dem_subsphere_io.cu:40-149
 *   nbc_finalize_params        DEMSubSphereIo::SetMeshSize        src/dem_subsphere_io.cu:195-241
 *   nbc_load_phase             DEMSubSphereIo::LoadPhase          src/dem_subsphere_io.cu:151-193
 *   nbc_load_clumps            DEMClumpSphereIo::LoadPhase        src/dem_clumpsphere_io.cu:24-62
 *   nbc_load_cube_wall         DEMWallCubeIo::LoadInfo            src/dem_cube_wall_io.cu:21-53
 *   nbc_load_cylinder_wall     DEMWallCylinderIo::LoadInfo        src/dem_cylind_wall_io.cu:20-56
 *   nbc_load_penetrator        DEMPeneCylinderIo::LoadInfo        src/dem_penetrator_io.cu:21-116
 *   nbc_create / nbc_upload    DEMClumpMain::AllocateGpuMemory + CopyDataFromHostToDevice,
 *                              DEMClumpSphere::AllocateGpuMemory, DEMGrid::GridConfig,
 *                              DEMForce::ForceConfig, DEMWallCube::AllocateMemory,
 *                              DEMPeneCylinder::AllocateMemory    src/dem_clump.cu:49-82,
 *                              src/dem_clumpsphere.cu:28-61, src/dem_grid.cu:212-271,
 *                              src/dem_force.cu:32-66, src/dem_cube_wall.cu:213-319,
 *                              src/dem_penetrator.cu:89-163
 *   nbc_step                   one pass of the loop body          src/dem_simulation_clump.cu:110-226
 *   nbc_rebuild                DEMGrid::NeighborSearch + SortArray + CreateContList
 *                                                                 src/dem_grid.cu:649-750
 *   nbc_get_*                  CopyDataToIo / CopyInfoToIo        src/dem_clump.cu:151-191,
 *                                                                 src/dem_clumpsphere.cu:63-70,
 *                                                                 src/dem_penetrator.cu:207-218
 *   nbc_write_*                dem_binary::* writers              src/dem_binary_io.h:100-180,
 *                                                                 src/dem_subsphere_io.cu:262-311
 */
#ifndef NBC_H
#define NBC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
enum {
    NBC_OK = 0,
    NBC_E_BADINPUT = 1,          /* invalid argument / malformed input file          */
    NBC_E_NEIGHBOR_OVERFLOW = 2, /* a sphere has more than BranchLmtLen/2 neighbours
                                    on one side (src/dem_grid.cu:709-719)            */
    NBC_E_NONFINITE = 3,         /* invalid or non-finite sphere/wall state
                                    (src/dem_force_cube_wall_clump.cu:191-203,295-297) */
    NBC_E_HIP = 4,               /* a HIP runtime call failed                        */
    NBC_E_NODEVICE = 5,          /* no usable GPU                                    */
    NBC_E_IO = 6,                /* file could not be opened or written              */
    NBC_E_PARAM = 7              /* world.par: unknown key or missing keys
                                    (src/dem_subsphere_io.cu:130,136-141)            */
};

/* ---- enumerations ------------------------------------------------------ */
enum { NBC_WALL_NONE = 0, NBC_WALL_CUBE = 1, NBC_WALL_CYLINDER = 2 };
enum { NBC_FP64 = 0, NBC_MIXED = 1 };
/* PenetratorShape (src/dem_parameters.h:100) plus the triangulated cone. */
enum { NBC_PENE_CYLINDER = 0, NBC_PENE_SPHERE = 1, NBC_PENE_CONE = 2 };
/* PenetratorMotionMode (src/dem_parameters.h:101). */
enum { NBC_PENE_CONSTANT_VELOCITY = 0, NBC_PENE_GRAVITY = 1 };
/* Which of the reference's run loops a context reproduces (nbc_options.driver).
 * CLUMP: DEMClumpMain::RunSimulationClump (src/dem_simulation_clump.cu:3-261).
 * The others are DEMmain's single-sphere loops, run as one-sphere clumps with
 * the DEMmain force and integration laws (nbc_upload_particles):
 * CUBE     RunSimulation2 (src/dem_simulation_cube.cu:3-226), cube wall of
 *          src/dem_force_cube_wall.cu (no wall history, no wall rolling or
 *          twisting torque);
 * CYLINDER RunSimulation (src/dem_simulation_cylind.cu:3-211), cylinder wall
 *          of src/dem_force_cylind_wall.cu; the history exchange that follows
 *          every rebuild (:141) leaves new list slot k with the history last
 *          written at old slot k;
 * SPLIT    RunSimulation_Split (src/dem_simulation_split.cu:3-211): CYLINDER
 *          with the glue kinematic types of src/dem_integrate_withglue.cu
 *          (type 0 free, rebuild when |cum| > 0.1 r; type 1 driven at
 *          v = (-0.2 y, -0.2 x, 0); type 2 fixed);
 * HGRID    RunSimulation_HGrid (src/dem_simulation_hgrid.cu:3-201): cylinder
 *          wall, inherited histories, the hierarchical grid: the level table
 *          of grid.bt (nbc_load_grid_levels) with each particle's type as its
 *          level (level_from_type), or one generated by radius (hgrid_levels).
 * With a penetrator (isIntruder) the DEMmain loops use its sphere law
 * (src/dem_force_penetrator.cu:8-168: the sphere's own spin, cohesion) and
 * leave its force and torque out of the drift's half-kick and the attitude
 * (src/dem_integrate.cu:23-25, 74-75); nbc_get_clumps' F and T then hold the
 * contact + wall totals that the next drift uses. */
enum { NBC_DRIVER_CLUMP = 0, NBC_DRIVER_CUBE = 1, NBC_DRIVER_CYLINDER = 2, NBC_DRIVER_SPLIT = 3,
       NBC_DRIVER_HGRID = 4 };

/* ---- run parameters ----------------------------------------------------
 * The 38 world.par keys (src/world.par:18-64, parsed by
 * src/dem_subsphere_io.cu:83-128 into Parameters/GlobalFlg,
 * src/dem_parameters.h:7-78) followed by the derived fields that
 * LoadParameters/SetMeshSize fill in. */
#define NBC_MAX_LEVELS 8

typedef struct nbc_params {
    int32_t ParticleNum;     /* NumberBody   */
    int32_t ClumpNum;
    int32_t BranchLmtLen;
    int32_t StartStep;
    int32_t EndStep;
    int32_t OutputInterval;
    double StepSize;
    double Ymod, Prat, Frcoe, NormRes, Frvel, Beta, MU_T, MU_R, Cohesion, Gravity;
    double Thresh, GridLevels;
    double Update;           /* initial rebuild flag (key "Update")            */
    double WorldSize[3], WorldOrigin[3], BoxUpLmt[3], BoxLowLmt[3];
    int32_t isIntruder, isSelfGravity, isWallForce, isQuaternion;
    /* derived: Damping (src/dem_subsphere_io.cu:143-145), uniform mesh
       (src/dem_subsphere_io.cu:197-230). WorldOrigin/WorldSize above are the
       values after SetMeshSize once nbc_finalize_params has run. */
    double Damping;
    double CellSize[3];
    double i_CellSize[3];
    uint32_t GridDim[3];
    int32_t CellNum;
    int32_t finalized;
    /* hierarchical grid (nbc_options.hgrid_levels > 1; src/dem_hgrid.cu:205-242
       with the level table generated, not read from grid.bt): level t holds
       the spheres with LevelRmin[t] < r <= LevelRmin[t-1] (level 0 the
       largest), cells of LevelCell[t] = 3 x the level's upper radius, cell
       index LevelOff[t] + (cz*ny + cy)*nx + cx; CellNum = all level cells + 1,
       the last cell holding out-of-box spheres. Levels = 1: uniform grid. */
    int32_t Levels;
    int32_t LevelOff[NBC_MAX_LEVELS];
    int32_t LevelDim[NBC_MAX_LEVELS][3];
    double LevelCell[NBC_MAX_LEVELS];
    double LevelRmin[NBC_MAX_LEVELS];
    /* Nested levels (nbc_options.hgrid_nested; no counterpart in the
       reference): LevelCell[t] = LevelCell[0] / 2^t with LevelCell[0] =
       3 x max radius, a sphere on the finest level whose cell is at least
       3 r (LevelRmin[t] = LevelCell[t+1] / 3), LevelDim[t] = GridDim * 2^t.
       Each coarse (level-0) cell owns a block of LevelBlock = sum 8^t keys:
       key = coarse * LevelBlock + LevelOff[t] + (sz * 2^t + sy) * 2^t + sx,
       with coarse = (Z * ny0 + Y) * nx0 + X, (X, Y, Z) = the level-t cell
       coordinates shifted right by t and (sx, sy, sz) their low t bits. The
       spheres of a coarse cell are contiguous in sorted order; a pair on two
       levels belongs to the front row of the coarser sphere. */
    int32_t LevelNested;
    int32_t LevelBlock;
} nbc_params;

/* Options that are not world.par keys (read from the optional nbc.par file by
 * the driver so that world.par stays 38-key compatible). Zero = default. */
typedef struct nbc_options {
    int32_t precision;       /* NBC_FP64 (default) or NBC_MIXED                */
    int32_t wall_type;       /* 0 = NBC_WALL_CUBE default when isWallForce     */
    int32_t block_size;      /* threads per block for per-sphere kernels (64..1024, used up to 256;
                                the contact kernels run fixed one-wave blocks) */
    int32_t radix_bits;      /* LSD radix digit width for the cell sort (1..8) */
    int32_t pair_capacity;   /* front-pair slots reserved per sphere (average) */
    int32_t batch_steps;     /* steps per chunk between status checks (default 16 with graphs, else 4) */
    int32_t hgrid_levels;    /* > 1: hierarchical grid with this many levels   */
    int32_t graphs;          /* steady chunks as HIP graph replays: 0 auto (beds up to 262144
                                spheres), 1 on, -1 off */
    double pene_frcoe;       /* penetrator friction; <= 0 means Frcoe          */
    double cell_size;        /* > 0 overrides 3*max radius (requires >= 3*rmax) */
    double near_skin;        /* fixed near-list skin as a fraction of Thresh (0: the
                                per-sphere skin of near_horizon; >= 1: gap 2*Thresh,
                                refreshed at rebuilds only) */
    int32_t driver;          /* NBC_DRIVER_* (0: the clump loop)                */
    int32_t level_from_type; /* HGRID driver: a particle's grid level is its type
                                (src/dem_hgrid.cu:213), with the level table of
                                nbc_load_grid_levels; 0: levels by radius      */
    int32_t hgrid_nested;    /* with hgrid_levels 2..4: nested power-of-two levels
                                numbered coarse-cell-major (see nbc_params)    */
    int32_t single_pass;     /* neighbour search: 0 a single pass into fixed-capacity
                                rows when they fit (else the count and fill
                                passes), -1 always the count and fill passes */
    double near_horizon;     /* per-sphere near-list skin: the sphere's speed bound
                                (|v| + |w| |x - x_clump| of its clump) times dt times
                                this many steps, clamped to [0.03, 1] Thresh. 0: 500
                                steps when near_skin is 0, else the fixed near_skin;
                                < 0: the fixed near_skin (0.25 when 0) */
} nbc_options;

/* ---- boundaries -------------------------------------------------------- */
/* Cube (Rocks/CubeWall.bt, src/dem_cube_wall_io.cu:43-48): half side lengths,
 * pose and motion. Cylinder (Rocks/CylindWall.bt,
 * src/dem_cylind_wall_io.cu:41-49): radius, half height, axis, property. */
typedef struct nbc_wall {
    int32_t type;            /* NBC_WALL_* */
    int32_t reserved;
    double side[3];          /* cube: SideLength (half-extents)               */
    double radius, height;   /* cylinder: Radius, Height (half height)        */
    double dir[3];           /* cylinder: Dirct                               */
    double prop[4];          /* cylinder: Property                            */
    double pos[3], vel[3], quat[4], angv[3];
} nbc_wall;

/* Penetrator (Rocks/Penetrator.bt, src/dem_penetrator_io.cu:21-116,
 * struct Cylinder src/dem_parameters.h:85-98). A cone is given as a closed
 * triangle mesh in the penetrator frame (positions relative to pos). */
typedef struct nbc_penetrator {
    int32_t shape;           /* NBC_PENE_* */
    int32_t motion;          /* NBC_PENE_CONSTANT_VELOCITY / NBC_PENE_GRAVITY */
    int32_t wait_for_contact;
    int32_t contact_state;
    double radius, height;
    double dir[3];
    double prop[4];          /* Ix Iy Iz mass */
    double pos[3], vel[3], quat[4], angv[3];
    int32_t ntri;            /* cone only: number of triangles                */
    int32_t reserved;
    const double* tri;       /* cone only: 9*ntri doubles (v0 v1 v2), may be NULL otherwise */
} nbc_penetrator;

/* ---- host-side state (file order, double3 packed as 3 doubles) -------- */
typedef struct nbc_spheres_in {      /* ph.00000000.bt columns */
    int32_t n;
    int32_t reserved;
    const int32_t* clump_id;
    const double* pos;       /* 3n */
    const double* local;     /* 3n */
    const double* mass;      /* n  */
    const double* radius;    /* n  */
} nbc_spheres_in;

typedef struct nbc_clumps_in {       /* ch.00000000.bt columns */
    int32_t n;
    int32_t reserved;
    const double* pos;       /* 3n */
    const double* vel;       /* 3n */
    const double* quat;      /* 4n (x y z w) */
    const double* angv;      /* 3n */
    const double* prop;      /* 4n (Ix Iy Iz mass) */
} nbc_clumps_in;

/* DEMmain single-sphere state (ph.00000000.bt of src/dem_io.cu:156-201:
 * rows "id sub type x y z vx vy vz qx qy qz qw wx wy wz Ix Iy Iz mass r").
 * Property = (Ix, Iy, Iz, mass) as in the reference's double4. */
typedef struct nbc_particles_in {
    int32_t n;
    int32_t reserved;
    const int32_t* type;     /* n, glue kinematic type (SPLIT driver); may be NULL (all 0) */
    const double* pos;       /* 3n */
    const double* vel;       /* 3n */
    const double* quat;      /* 4n (x y z w) */
    const double* angv;      /* 3n */
    const double* prop;      /* 4n (Ix Iy Iz mass) */
    const double* radius;    /* n  */
} nbc_particles_in;

/* Per-step statistics of the last step (FP64 global sums). */
typedef struct nbc_stats {
    int64_t step;            /* completed steps since upload                  */
    int64_t rebuilds;        /* neighbour-list rebuilds since upload          */
    int64_t front_pairs;     /* list pairs (i<j) incl. intra-clump            */
    int64_t back_entries;
    int64_t touching_ss;     /* touching sphere-sphere contacts, different clumps */
    int64_t touching_wall;   /* touching wall + penetrator contacts          */
    double kinetic_energy;   /* sum over clumps of 1/2 m v^2 + 1/2 sum I w^2 */
    double pene_force[3];    /* reaction force on the penetrator (-sum F)    */
    double pene_torque[3];
    double max_cum_disp;     /* reserved */
    int64_t near_refreshes;  /* near-list refreshes between rebuilds since upload */
    int64_t near_pairs;      /* pairs on the contact kernel's near list         */
    int64_t graph_replays;   /* chunks of steps run as HIP graph replays since upload */
    int64_t single_pass_rebuilds; /* rebuilds whose rows came from the single-pass search */
} nbc_stats;

typedef struct nbc_ctx nbc_ctx;

/* ---- library / errors -------------------------------------------------- */
const char* nbc_version(void);
const char* nbc_last_error(nbc_ctx* ctx);    /* ctx may be NULL: thread-global message */
int nbc_device_count(int32_t* out);

/* ---- drop-in file formats (host only, no device use) ------------------ */
/* world.par: 38 keys exactly, unknown key -> NBC_E_PARAM, fewer -> NBC_E_PARAM.
 * iniconfile receives the IniConFile value. Damping is derived. */
int nbc_load_world_par(const char* path, nbc_params* out, char* iniconfile, size_t iniconfile_len);
/* SetMeshSize: cell = 3*max radius (or opt->cell_size), origin -= cell,
 * size += 2*cell, GridDim = ceil(size/cell). opt may be NULL. With
 * opt->hgrid_levels > 1 it also fills the level table (see nbc_params); the
 * margin and the uniform fields use the coarsest level's cell. */
int nbc_finalize_params(nbc_params* p, const double* radius, int32_t n, const nbc_options* opt);
/* grid.bt of the DEMmain hgrid loop (DEMSphereIo::LoadGridLevels,
 * src/dem_io.cu:204-248): p->GridLevels rows "level cellSize nx ny nz offset
 * n"; fills the level table of p (level 0 the coarsest, offsets contiguous)
 * in place of nbc_finalize_params, without its origin margin. Use with
 * nbc_options.level_from_type = 1; every particle's radius must fit its level
 * (3 r <= cellSize). */
int nbc_load_grid_levels(const char* path, nbc_params* p);
/* Optional extension file nbc.par ("key = value"); missing file leaves defaults. */
int nbc_load_options(const char* path, nbc_options* out);
/* ph text: n rows "id sub clumpId x y z lx ly lz mass radius". Buffers caller-owned. */
int nbc_load_phase(const char* path, int32_t n, int32_t* clump_id, double* pos, double* local,
                   double* mass, double* radius);
/* DEMmain ph text (src/dem_io.cu:156-201): n rows of 21 columns, see
 * nbc_particles_in. Buffers caller-owned. */
int nbc_load_particles(const char* path, int32_t n, int32_t* type, double* pos, double* vel, double* quat,
                       double* angv, double* prop, double* radius);
/* ch text: n rows "id x y z vx vy vz qx qy qz qw wx wy wz Ix Iy Iz mass". */
int nbc_load_clumps(const char* path, int32_t n, double* pos, double* vel, double* quat,
                    double* angv, double* prop);
int nbc_load_cube_wall(const char* path, nbc_wall* out);
int nbc_load_cylinder_wall(const char* path, nbc_wall* out);
/* Penetrator.bt; a "cone R H" shape token is an extension. The cone's mesh
 * buffer is library-owned and lives until nbc_free_mesh. */
int nbc_load_penetrator(const char* path, nbc_penetrator* out);
void nbc_free_mesh(nbc_penetrator* p);
/* Triangulate a cone of base radius R and height H in the penetrator frame:
 * base disc at z = 0, apex at z = -H. nseg rim segments give 2*nseg
 * triangles (9 doubles each) written to tri_out when it is not NULL.
 * Returns the triangle count or < 0 on error. */
int nbc_make_cone_mesh(double radius, double height, int32_t nseg, double* tri_out);

/* ---- synthetic beds --------------------------------------------------- */
/* kind 0 = C0: nclumps (default 2 000) collinear 3-sphere clumps, r = 1 mm
 *          x U[0.9, 1.1], 0.8 r spacing, shuffled lattice in a cube
 *          40 x 40 x 80 mm (container[0..2] override the edge lengths);
 * kind 1 = C1: nclumps (default 200 000) clumps from 8 random templates of
 *          2-8 spheres, lognormal radii (median 1 mm, sigma 0.15, clipped to
 *          [0.6, 1.6] mm), sequential deposition into a cylinder of radius
 *          60 mm (container[0] overrides) whose height fits the bed;
 * kind 2 = C2: nclumps (default 571 429) clumps of 1-6 spheres from 32 unit
 *          templates, each clump scaled so that radii are log-uniform over
 *          [0.2, 2.0] mm, sequential deposition into a cube wall of section
 *          150 x 150 mm (container[0] overrides) whose height fits the bed;
 *          friction 0.6, meant for the hierarchical grid (3 levels);
 * kind 3 = single spheres for the DEMmain drivers: nclumps (default 100 000)
 *          one-sphere clumps, radii as C1, solid-sphere mass and inertia,
 *          deposited like C1 into a cylinder of radius 60 mm (container[0]).
 * The generator allocates the output arrays; release them with nbc_free_bed. */
typedef struct nbc_bed_spec {
    int32_t kind;
    int32_t nclumps;
    uint64_t seed;
    double container[4];
} nbc_bed_spec;

typedef struct nbc_bed_out {
    int32_t n, nc;           /* sphere and clump counts                       */
    int32_t* clump_id;       /* n   */
    double* pos;             /* 3n  */
    double* local;           /* 3n  */
    double* mass;            /* n   */
    double* radius;          /* n   */
    double* cpos;            /* 3nc */
    double* cvel;            /* 3nc */
    double* cquat;           /* 4nc */
    double* cangv;           /* 3nc */
    double* cprop;           /* 4nc */
    nbc_wall wall;           /* container                                      */
    nbc_params params;       /* world.par values for the bed (not finalized)   */
} nbc_bed_out;

int nbc_generate_bed(const nbc_bed_spec* spec, nbc_bed_out* out);
void nbc_free_bed(nbc_bed_out* out);
/* Writes world.par, ph.00000000.bt, ch.00000000.bt, Rocks/{CubeWall,CylindWall,
 * Penetrator}.bt and nbc.par for a generated bed (pene may be NULL). */
int nbc_save_bed_text(const char* dir, const nbc_bed_out* bed, const nbc_penetrator* pene);

/* ---- device context ---------------------------------------------------- */
int nbc_create(nbc_ctx** out, const nbc_params* params, const nbc_options* opt, int32_t device);
int nbc_destroy(nbc_ctx* ctx);
/* Batch of independent beds advanced by one context: every step is one set of
 * kernel launches over all beds. params points to nbeds finalized parameter
 * blocks; beds share the material, StepSize, Thresh, BranchLmtLen, Update,
 * OutputInterval and isWallForce, and each keeps its own uniform grid, box and
 * wall (all walls of one type, opt->wall_type); no hierarchical grid and no
 * penetrator. The spheres and clumps of bed b occupy the index ranges
 * [sphere_off[b], sphere_off[b+1]) and [clump_off[b], clump_off[b+1]) of the
 * getters (nbc_get_batch_ranges; sorting never moves a sphere out of its
 * bed's range); file indices and clump ids are global (offset by bed). A
 * rebuild flagged by one bed rebuilds every bed.
 * No counterpart in the reference, which runs one bed per process
 * (src/main.cpp:7-8). */
int nbc_create_batch(nbc_ctx** out, int32_t nbeds, const nbc_params* params, const nbc_options* opt,
                     int32_t device);
/* spheres, clumps and walls are arrays of nbeds (walls may be NULL when
 * isWallForce == 0); each bed's clump ids are local to the bed. */
int nbc_upload_batch(nbc_ctx* ctx, const nbc_spheres_in* spheres, const nbc_clumps_in* clumps,
                     const nbc_wall* walls);
/* nbeds + 1 offsets each (either pointer may be NULL). */
int nbc_get_batch_ranges(nbc_ctx* ctx, int32_t* sphere_off, int32_t* clump_off);
/* Copy host state in (file order) and reset histories; wall may be NULL when
 * isWallForce == 0; pene may be NULL when isIntruder == 0. */
int nbc_upload(nbc_ctx* ctx, const nbc_spheres_in* s, const nbc_clumps_in* c,
               const nbc_wall* wall, const nbc_penetrator* pene);
/* DEMmain drivers (opt->driver > 0): copy single-sphere state in, file order
 * (DEMmain::CopyDataFromHostToDevice, src/dem.cu:62-72); pene may be NULL when
 * isIntruder == 0. Particle i becomes
 * clump i with one sphere at its centre; params->ClumpNum must equal
 * ParticleNum. The clump getters then return the particles in file order,
 * nbc_get_particles in the current sorted order. */
int nbc_upload_particles(nbc_ctx* ctx, const nbc_particles_in* p, const nbc_wall* wall,
                         const nbc_penetrator* pene);
/* DEMmain::CopyDataToIo (src/dem.cu:113-127): the device arrays in the current
 * sorted order; sorted_id is the reference's part_idx (identity), file_id its
 * part_fix. type follows the sort in the CYLINDER, SPLIT and HGRID loops and
 * stays in file order in the CUBE loop, which does not sort it
 * (src/dem_simulation_cube.cu:102-108). Any pointer may be NULL. */
int nbc_get_particles(nbc_ctx* ctx, int32_t* sorted_id, int32_t* file_id, int32_t* type, double* pos,
                      double* vel, double* quat, double* angv, double* prop, double* radius);
/* Advance nsteps full timesteps (src/dem_simulation_clump.cu:110-226). */
int nbc_step(nbc_ctx* ctx, int32_t nsteps);
/* Force a neighbour-list rebuild on the current positions. */
int nbc_rebuild(nbc_ctx* ctx);
/* Record per-contact exports (penetrator per-sphere Fn/Ft/Tn/Tt, pair normal
 * force and contact point) during the next step. */
int nbc_set_record(nbc_ctx* ctx, int32_t on);
int nbc_synchronize(nbc_ctx* ctx);
int nbc_get_stats(nbc_ctx* ctx, nbc_stats* out);

/* ---- getters (sorted order; any pointer may be NULL) ------------------ */
int nbc_get_spheres(nbc_ctx* ctx, int32_t* file_id, int32_t* clump_id, double* pos, double* cum,
                    double* local, double* mass, double* radius);
int nbc_get_grid(nbc_ctx* ctx, int32_t* cell, int32_t* perm, int32_t* cell_beg, int32_t* cell_end);
int nbc_get_pair_counts(nbc_ctx* ctx, int64_t* nfront, int64_t* nback);
/* Front list (j > i, reference search order), flags bit0 = intra-clump;
 * back list of each sphere (j < i, reverse search order) with the owning
 * front slot (-1 if the owner's search did not find the pair). */
int nbc_get_pairs(nbc_ctx* ctx, int32_t* front_off, int32_t* front_j, uint8_t* front_flags,
                  int32_t* back_off, int32_t* back_i, int32_t* back_slot);
/* Tangential/twisting/rolling history written by the last step per front
 * slot (9 doubles; zero where the pair was not touching). */
int nbc_get_pair_history(nbc_ctx* ctx, double* hist9);
/* Per-sphere force / torque of the last step recorded with nbc_set_record:
 * kind 0 = total (contact + wall + penetrator, src/dem_clump.cu:97-98),
 * 1 = sphere-sphere, 2 = wall, 3 = penetrator. */
int nbc_get_sphere_forces(nbc_ctx* ctx, int32_t kind, double* force, double* torque);
int nbc_get_clumps(nbc_ctx* ctx, double* pos, double* vel, double* quat, double* angv,
                   double* force, double* torque);
/* Penetrator pose and the reaction history cache (OutputInterval entries). */
int nbc_get_penetrator(nbc_ctx* ctx, double* pos, double* vel, double* quat, double* angv,
                       double* cache_f, double* cache_t, int32_t* contact_state);
/* Recorded exports of the last step with recording on. */
int nbc_get_pene_exports(nbc_ctx* ctx, double* fn, double* ft, double* tn, double* tt);
int nbc_get_pair_exports(nbc_ctx* ctx, double* fcn, double* cpt);

/* ---- standalone grid build on device buffers ------------------------- */
/* DEMGrid::NeighborSearch (src/dem_grid.cu:649-678) without a context: the
 * cell hash of GetParticleCellIndex_cuk (:155-202; the hierarchical grid of
 * src/dem_hgrid.cu:205-242 when p->Levels > 1, levels by radius), the stable
 * sort of SortParticleByCellIndex (:278-305) and FindBegEndInCell_cuk
 * (:105-153). p is finalized; every buffer is device memory: d_pos4 holds n
 * rows (x, y, z, radius); d_cell[n] receives the sorted cell keys (part_cel),
 * d_perm[n] the new -> old index (part_idx_pre), d_beg / d_end[CellNum] the
 * first and one-past-last sorted index of each cell (-1 when empty).
 * stream is a hipStream_t (NULL: the default stream). The call allocates the
 * sort's temporaries (about 16 n + 8 n + small bytes), runs on the stream and
 * synchronizes it before freeing them. */
int nbc_grid_build(const nbc_params* p, const double* d_pos4, int32_t n, uint32_t* d_cell, int32_t* d_perm,
                   int32_t* d_beg, int32_t* d_end, void* stream);
/* nbc_grid_build without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results; the sort's temporaries are a fixed device array of
 * the library (about 32 MiB per GPU, held until the process ends), so n is at
 * most NBC_GRID_NOALLOC_MAX_SPHERES (larger n: NBC_E_BADINPUT). The call is
 * asynchronous on `stream` and does not synchronize. Every call on one GPU
 * shares that array: calls that overlap in time on one GPU race. */
#define NBC_GRID_NOALLOC_MAX_SPHERES 2097152
int nbc_grid_build_noalloc(const nbc_params* p, const double* d_pos4, int32_t n, uint32_t* d_cell,
                           int32_t* d_perm, int32_t* d_beg, int32_t* d_end, void* stream);

/* ---- timing of library kernels on the library stream ----------------- */
/* Kernel classes: 0 sphere kinematics (K2; K1 of the first step of a call),
 * 1 sphere-sphere contacts, 2 clump reduction and second half-kick (with K1
 * of the next step inside the call), 3 rebuild, 4 penetrator sum, 5 wall and
 * penetrator contacts (side stream), 6 near-list refresh between rebuilds. */
int nbc_profile_enable(nbc_ctx* ctx, int32_t on);
int nbc_profile_get(nbc_ctx* ctx, int32_t kernel_class, double* total_ms, int64_t* launches);

/* ---- autotuner --------------------------------------------------------- */
/* Picks options for one bed by timing candidates on the device: coordinate
 * descent over near_skin, batch_steps, radix_bits, block_size and graphs (none
 * of them changes any result), then, when asked, the cell size (1, 1.25 or 1.5 times
 * 3*max radius) and the hierarchical grid (off, 2 or 3 levels, level-major
 * or nested), which change
 * the sphere order and so the FP64 sums by rounding. Each candidate gets its
 * own context: upload, warm-up steps (the first rebuild included), then
 * `steps` timed steps. params are world.par values NOT yet finalized (every
 * candidate finalizes its own copy); base gives precision, wall type and
 * penetrator friction. best receives the fastest options; entries (may be
 * NULL) receives up to cap measured candidates and *n_entries their count.
 * No counterpart in the reference, whose launch shapes are fixed. */
#define NBC_TUNE_CELL 1
#define NBC_TUNE_HGRID 2
typedef struct nbc_tune_entry {
    nbc_options opt;
    double ms_per_step;      /* < 0: the candidate failed (e.g. neighbour overflow) */
} nbc_tune_entry;
int nbc_autotune(const nbc_params* params, const nbc_options* base, const nbc_spheres_in* s,
                 const nbc_clumps_in* c, const nbc_wall* wall, const nbc_penetrator* pene, int32_t device,
                 int32_t steps, int32_t flags, nbc_options* best, nbc_tune_entry* entries, int32_t cap,
                 int32_t* n_entries);
/* The same coordinate descent for a batch context of nbeds beds (arrays of
 * nbeds parameter blocks, not finalized, spheres, clumps and walls, as for
 * nbc_create_batch / nbc_upload_batch): near_skin, batch_steps, radix_bits,
 * block_size, graphs and, with NBC_TUNE_CELL, the cell size; every candidate
 * is one batch context stepped as a whole. No counterpart in the reference,
 * which runs one bed per process. */
int nbc_autotune_batch(int32_t nbeds, const nbc_params* params, const nbc_options* base,
                       const nbc_spheres_in* s, const nbc_clumps_in* c, const nbc_wall* walls, int32_t device,
                       int32_t steps, int32_t flags, nbc_options* best, nbc_tune_entry* entries, int32_t cap,
                       int32_t* n_entries);

/* ---- binary writers (src/dem_binary_io.h layout, native little-endian) */
int nbc_write_particles_bin(const char* dir, int32_t step, int32_t n, const int32_t* sorted_id,
                            const int32_t* file_id, const int32_t* clump_id, const double* pos,
                            const double* local, const double* mass, const double* radius);
/* DEMSphereIo::SaveParticles_Bin (src/dem_io.cu:290-305): ph.<step> of the
 * DEMmain loops. */
int nbc_write_spheres_bin(const char* dir, int32_t step, int32_t n, const int32_t* sorted_id,
                          const int32_t* file_id, const int32_t* type, const double* pos, const double* vel,
                          const double* quat, const double* angv, const double* prop, const double* radius);
int nbc_write_clumps_bin(const char* dir, int32_t step, int32_t nc, const double* pos,
                         const double* vel, const double* quat, const double* angv,
                         const double* prop, const double* pene_pos, const double* pene_vel,
                         const double* pene_quat, const double* pene_angv);
/* DEMSphereIo::SaveParticles_Bin's intruder file (src/dem_io.cu:307-317):
 * intruder.<step>.bin in the clump layout, no clumps, the intruder as id -1. */
int nbc_write_intruder_bin(const char* dir, int32_t step, const double* pos, const double* vel,
                           const double* quat, const double* angv);
int nbc_write_force_history_bin(const char* dir, int32_t step, int32_t interval,
                                const double* force, const double* torque);
int nbc_write_yi_bin(const char* dir, int32_t step, int32_t n, const int32_t* clump_id,
                     const double* fn, const double* ft, const double* tn, const double* tt);
int nbc_write_contacts_bin(const char* dir, int32_t step, int32_t n, int32_t branch,
                           const int32_t* nei, const double* fcn, const double* cpt);

/* ---- the drop-in entry point ------------------------------------------ */
/* Binary -> text state for resuming a clump run: OutPart/ph.<step>.bin and
 * OutClump/ch.<step>.bin under rundir (layouts of src/dem_binary_io.h:19-181)
 * become outdir/<IniConFile> and outdir/ch.00000000.bt in the text formats the
 * loaders read (src/dem_subsphere_io.cu:174-175, src/dem_clumpsphere_io.cu:
 * 46-48), rows in file order; world.par is copied with StartStep = step,
 * nbc.par and Rocks/ are copied, and the penetrator's output record replaces
 * its state in Rocks/Penetrator.bt. Pair, wall and penetrator histories are
 * not in the outputs, so the resumed run starts them from zero. Refuses to
 * overwrite ph/ch files in outdir. No counterpart in the reference, whose
 * outputs cannot be fed back without such a conversion. */
int nbc_resume_from_bin(const char* rundir, int32_t step, const char* outdir);
/* Runs the whole simulation like DEMClumpMain::RunSimulationClump from the
 * files in workdir (world.par, IniConFile, ch.00000000.bt, Rocks/, optional
 * nbc.par), writing OutPart/OutClump/OutForce/OutCnt there. */
int nbc_run_simulation_clump(const char* workdir, int32_t device);
/* Runs one of DEMmain's single-sphere loops (driver = NBC_DRIVER_CUBE:
 * RunSimulation2, CYLINDER: RunSimulation, SPLIT: RunSimulation_Split, HGRID:
 * RunSimulation_HGrid, which reads grid.bt when present and then takes each
 * particle's type as its level) from world.par, IniConFile (21-column rows),
 * Rocks/CubeWall.bt or Rocks/CylindWall.bt, Rocks/Penetrator.bt and the
 * optional nbc.par, writing OutPart/ph.<step> every OutputInterval steps and,
 * with isIntruder in the cube loop, OutPart/intruder.<step>, OutForce/fc.<step>
 * and, every 2 OutputInterval steps, OutCnt/{Nei,Fcn,Cpt}.<step>
 * (src/dem_simulation_cube.cu:182-222). driver 0 takes nbc.par's Driver. */
int nbc_run_simulation(const char* workdir, int32_t device, int32_t driver);

#ifdef __cplusplus
}
#endif
#endif /* NBC_H */
