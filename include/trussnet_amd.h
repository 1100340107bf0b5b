/*
 * trussnet_amd.h -- C ABI of the MI355X (gfx950) truss relaxation.
 *
 * The relaxation moves mesh nodes inside a multi-label volume until the truss
 * of their K nearest neighbours is in equilibrium, trapping nodes on the smooth
 * label interfaces. It is the drop-in for v2mesh's relaxation entry points:
 *
 *   void tn::relax_cl (const Grid&, const RelaxParams&, Nodes&, RelaxStats&, int device)
 *        (src/v2m_gpu.h:18, implemented at src/v2m_gpu.cpp:116-553)
 *   void tn::relax_cpu(const Grid&, const RelaxParams&, Nodes&, RelaxStats&)
 *        (src/v2m_particles.h:99, implemented at src/v2m_particles.cpp:728-976)
 *
 * The grid is read-only. The nodes are mutated in place: P, typ and part; lab
 * and the node count are unchanged (src/v2m_gpu.cpp:514-517). Every entry point
 * returns TN_OK or an error status; tn_last_error() gives the message of the
 * calling thread's last failure. The C++ wrapper (trussnet_amd.hpp) turns a
 * status into std::runtime_error, which is what relax_cl throws
 * (src/v2m_gpu.h:16-17) and what the pipeline catches to fall back to the CPU
 * (src/v2m_pipeline.cpp:497-526).
 *
 * Plain pointers and sizes only: no HIP or torch types cross this boundary.
 */
#ifndef TRUSSNET_AMD_H
#define TRUSSNET_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TN_ABI_VERSION 3

/* Sizes fixed by the node and field formats. */
#define TN_K 24          /* truss neighbours per node: V2M_K, src/opencl/v2m_particle_body.cl:18 */
#define TN_MAXLEV 8      /* hash levels: V2M_MAXLEV, src/opencl/v2m_particle_body.cl:19 */
#define TN_BS 8          /* brick side: V2M_BS, src/opencl/v2m_grid_body.cl:20 */
#define TN_SLOT 512      /* phi values per (brick, label) slot: V2M_SLOT, :21 */
#define TN_BL 16         /* labels per brick: V2M_BL, :22 */
#define TN_NOLAB 0xFFFF  /* no label: V2M_NOLAB, src/opencl/v2m_seed_body.cl:19 */

/* Node types (src/opencl/v2m_seed_body.cl:15-18). */
enum { TN_INTERIOR = 0, TN_INTERFACE = 1, TN_JUNCTION = 2, TN_CORNER = 3 };

/* Status codes. */
enum {
    TN_OK = 0,
    TN_ERR_ARG = 1,         /* a null pointer, a size mismatch, an out-of-range value */
    TN_ERR_NODEVICE = 2,    /* no HIP device, or the device index is out of range */
    TN_ERR_HIP = 3,         /* a HIP runtime call failed */
    TN_ERR_UNSUPPORTED = 4, /* a mode this build does not run on the device */
    TN_ERR_INTERNAL = 5
};

/*
 * The field the nodes move in: v2mesh's Grid (src/v2m_grid.h:45-81) as raw
 * arrays, x fastest. Voxel (i,j,k) is centred at (i*vs[0], j*vs[1], k*vs[2])
 * "grid mm". Bricks are 8^3 voxels; nb = nbx*nby*nbz.
 */
typedef struct tn_grid_view {
    int32_t nx, ny, nz;       /* voxels */
    int32_t nbx, nby, nbz;    /* bricks: ceil(n / 8) per axis */
    float vs[3];              /* voxel size (mm) */
    const uint16_t* L;        /* [nx*ny*nz] labels (Grid::L) */
    const int32_t* bl_cnt;    /* [nb] labels within reach of the brick (TN_BL + 1 = overflow) */
    const uint16_t* bl_lab;   /* [nb*TN_BL] those labels, ascending, TN_NOLAB padded */
    const int32_t* bl_slot;   /* [nb] first slot of the brick, -1 = uniform */
    const float* phi;         /* [nslot*TN_SLOT] smoothed indicators per (brick, label) slot */
    int64_t nslot;
    const float* h;           /* [nx*ny*nz] sizing field (mm) */
    float hmin, hmax;         /* sizing range (Grid::hmin/hmax): the hash geometry */
    const float* gI;          /* gray-scale intensity per voxel (gm > 0), else null */
    const float* gTW;         /* gm > 0: [t_0..t_{m-1}, W_0..W_{m-1}], the ascending thresholds and
                                 their membership widths (Grid::gTW); gm < 0: the labels' compiled
                                 analytic-field program (src/opencl/v2m_sdf_body.cl layout), checked
                                 on the host before use */
    int64_t ngTW;             /* values in gTW (>= 2 gm for a gray-scale grid) */
    int32_t gm;               /* 0: label volume; m > 0: gray-scale volume with m thresholds;
                                 < 0: shape fields */
    const float* feat;        /* [nfeat] the shape input's feature curves (Grid::feat: 1 x y z
                                 corners; 2 x0 y0 z0 x1 y1 z1 segments; 3 cx cy cz nx ny nz r
                                 circles; 4 n then n points, polylines), or null: seeding adds
                                 their pinned CORNER nodes (add_feature_nodes,
                                 src/v2m_particles.cpp:304-620) */
    int64_t nfeat;
} tn_grid_view;

/* RelaxParams (src/v2m_particles.h:21-41) plus the rebuild divisor that the
 * reference reads from the environment (V2M_REBUILD_DIV, src/v2m_particles.cpp:731). */
typedef struct tn_relax_params {
    int32_t nseed;
    float t, skin, fscale, fsurf, dt, maxstep, snap;
    int32_t max_iters;
    float dptol, jseed;
    int32_t corners, voxel_trap;
    float thin;
    int32_t fire;
    float fire_dtmax;
    int32_t verbose;
    int32_t rebuild_div;      /* <= 0: 300, the reference's default */
} tn_relax_params;

/* Nodes (src/v2m_particles.h:78-87). P is 3 floats per node (AoS). */
typedef struct tn_nodes_view {
    int64_t n;
    float* P;                 /* [3n] in/out */
    const uint16_t* lab;      /* [n] in */
    uint8_t* typ;             /* [n] in/out */
    uint16_t* part;           /* [2n] in/out */
} tn_nodes_view;

/* RelaxStats (src/v2m_particles.h:89-96). The ms_* fields are device times while
 * kernel timing is on (tn_plan_set_timing), else 0. power_fallbacks (no counterpart
 * in RelaxStats) counts the FIRE iterations whose power sign the device could not
 * certify from its tree sum and took from the serial sum in node order, the CPU's
 * own summation (src/v2m_particles.cpp:849-855). Setting TN_SERIAL_POWER=1 in the
 * environment before a plan is created makes every FIRE iteration take that path. */
typedef struct tn_relax_stats {
    int32_t iters, rebuilds;
    float last_move, last_p99;
    int32_t fire_resets;
    float fire_dt;
    int64_t n_interior, n_interface, n_junction, n_corner;
    double ms_seed, ms_hash, ms_force, ms_move;
    int32_t power_fallbacks;
} tn_relax_stats;

/* The RelaxParams defaults (src/v2m_particles.h:22-39), rebuild_div 300. */
void tn_relax_params_default(tn_relax_params* prm);

/* Message of the calling thread's last failed call ("" if none). */
const char* tn_last_error(void);

/* Number of HIP devices (0 and TN_ERR_NODEVICE when there is none). */
int tn_device_count(int32_t* count);

/*
 * Relax the nodes on HIP device `device` (-1 = the first). Same contract as
 * relax_cl (src/v2m_gpu.h:16-18) and relax_cpu (src/v2m_particles.h:99):
 * uploads the grid and nodes, runs the loop on the device, writes back P, typ
 * and part, fills *st. On error the nodes are left unchanged.
 */
int tn_relax(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, tn_relax_stats* st,
             int32_t device);

/*
 * tn_relax without allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, results (the same bits
 * as tn_relax) and errors, but every device buffer (the grid, the nodes, lists,
 * hash, queues, control block) is carved from one static device array of
 * TN_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until the process
 * ends; no hipMalloc, hipFree, pinned host memory, stream or event is created. The
 * work runs on the null stream with the move passes one after another. A
 * relaxation that needs more than the array returns TN_ERR_ARG with the size in
 * tn_last_error(). All calls on one GPU share its array, which keeps the last
 * call's contents (each call zeroes it first); calls that overlap on one GPU would
 * race on it, so the library serialises them with a per-device lock.
 */
#define TN_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int tn_relax_noalloc(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, tn_relax_stats* st,
                     int32_t device);

/*
 * Relax `count` independent volumes on one device: volume k is relaxed exactly as
 * tn_relax(&grids[k], &prms[k], &nodes[k], &stats[k], device) would relax it, to
 * the same bits. No counterpart in the reference, which relaxes one volume per
 * relax_cl call (src/v2m_gpu.h:18); for callers with many volumes. Each volume's
 * plan runs on its own HIP stream and one host thread advances them as their
 * iterations complete, so small volumes overlap. stats may be null. On error no
 * volume's nodes are written.
 */
int tn_relax_batch(const tn_grid_view* grids, const tn_relax_params* prms, tn_nodes_view* nodes,
                   tn_relax_stats* stats, int32_t count, int32_t device);

/*
 * Seeding on the device: seed_cpu (src/v2m_particles.h:98; its body at
 * src/v2m_particles.cpp:637-726), node for node. Each voxel claims the points of
 * its seed level's HCP lattice inside its box (v2m_seed_voxel, count -> scan ->
 * fill, node order = voxel order) and every seed is classified and projected onto
 * a nearby interface or junction curve (v2m_seed_classify). With prm->jseed > 0
 * the junction-line seeds follow (seed_junctions, src/v2m_particles.cpp:80-244):
 * one seed per (seed level, cell of jseed x level spacing, label triple) at the
 * 3-label grid vertices, the lattice nodes within 0.4 h of one dropped, the seeds
 * appended in key order. With prm->corners the fixed CORNER nodes at the 4+-label
 * vertices come last (v2m_corner_vertex). With g->nfeat > 0 (a shape grid's
 * feature curves) the seeds within 0.4 h of a feature node are dropped and the
 * feature nodes, pinned CORNER nodes with the labels round them, appended
 * (add_feature_nodes, src/v2m_particles.cpp:304-620), as seed_cpu does
 * (:625-635). grade[nx*ny*nz] is Grid::grade (src/v2m_grid.h:77). tn_seed_count
 * gives the seed count; tn_seed writes exactly
 * that many nodes into the arrays (P[3n], lab[n], typ[n], part[2n], and part3[n],
 * the third partner of a corner node, unless part3 is null).
 */
int tn_seed_count(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int32_t device,
                  int64_t* n);
int tn_seed(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int32_t device, int64_t n,
            float* P, uint16_t* lab, uint8_t* typ, uint16_t* part, uint16_t* part3);

/*
 * Node thinning on the device: thin_nodes (src/v2m_particles.h:101; body
 * src/v2m_particles.cpp:1223-1357), which run_pipeline applies to the seeds
 * before the relaxation when prm->thin > 0 (src/v2m_pipeline.cpp:467-476).
 * Junction and corner nodes are kept; interface nodes, then interior nodes, in
 * the order of h and then index, are each dropped when a node kept before them
 * lies within thin * h on the same interface (or, for an interior node, carries
 * its label). The n nodes (P[3n], lab, typ, part[2n], and part3[n], the corners'
 * third partners; null: none) are compacted in place, in order, exactly as the
 * reference compacts them; *n_kept is their new count. prm->thin <= 0 leaves them.
 * Only the grid's sizes (h, dims, voxel sizes, hmin, hmax) are read.
 */
int tn_thin_nodes(const tn_grid_view* g, const tn_relax_params* prm, int64_t n, float* P, uint16_t* lab, uint8_t* typ,
                  uint16_t* part, uint16_t* part3, int64_t* n_kept, int32_t device);

/*
 * Feature protection on the device: protect_features (src/v2m_particles.h:105;
 * body src/v2m_particles.cpp:1002-1221), run on the relaxed nodes of a shape
 * input (src/v2m_pipeline.cpp:549-555). feat[nfeat] is the grid's Grid::feat (the
 * shapes' feature curves: points, segments, circles, polylines). Between
 * neighbouring pinned (corner) nodes on each curve no node may sit inside the
 * pair's diametral ball; those nodes are dropped and the rest compacted in place
 * as above (*n_kept). The balls are built on the host as the reference builds
 * them; the test of every node and the compaction run on the device.
 */
int tn_protect_features(const tn_grid_view* g, const float* feat, int64_t nfeat, int64_t n, float* P, uint16_t* lab,
                        uint8_t* typ, uint16_t* part, uint16_t* part3, int64_t* n_kept, int32_t device);

/*
 * The grid build on the device: build_grid_cpu (src/v2m_grid.cpp:47-722, the
 * Grid of src/v2m_grid.h:45-81 from a LabelVolume and GridParams) for a label
 * volume, a tissue-probability input (prob) or a gray-scale input (gray and its
 * ascending thresholds; L then holds the number of thresholds <= each intensity,
 * apply_thresholds). It makes the labels near each 8^3 brick, the slots of the
 * mixed bricks, the interface fields (smoothed indicators or probabilities and
 * their thickness-adaptive blend, or the gray-scale memberships), the sizing field
 * (curvature, per-label and interface sizes, thin layers, the user's field), its
 * gradient limit and the grades, each array identical to the reference's. Shape
 * inputs (analytic fields) are not built here. The grid stays on the device until
 * tn_grid_download; tn_grid_get_info gives the sizes (nslot, the phi slots) and
 * the scalars of the Grid. Arrays of tn_grid_download may be null to skip them:
 * bl_cnt[nb], bl_lab[16 nb], bl_slot[nb], phi[512 nslot], h[nx*ny*nz],
 * grade[nx*ny*nz], gTW[2 gm] (gray-scale: thresholds, then widths).
 */
typedef struct tn_grid_params {   /* GridParams, src/v2m_grid.h:21-43 */
    float sigma, sigma_thin, thin_lo, thin_hi, thick, thin_floor, preserve, sigma_curv;
    float hbase, hmin, hmax, K, g;
    const float* hlab;      /* per-label size (mm), index = label, 0 = hbase */
    int32_t nhlab;
    float iglob;            /* interface sizes (InterfaceSizes, src/v2m_isize.h:25-28): every interface */
    const float* ilab;      /* per label */
    int32_t nilab;
    const float* ipair;     /* (a, b, h) triples, a < b */
    int32_t npair;
} tn_grid_params;

typedef struct tn_volume_view {   /* LabelVolume, src/v2m_volume.h */
    int32_t nx, ny, nz;
    float vs[3];                  /* voxel size (mm) */
    const uint16_t* L;            /* [nx*ny*nz] labels, x fastest */
    const float* prob;            /* optional [nprob][nx*ny*nz]: label l's probability (label 0 the exterior) */
    int32_t nprob;
    const float* gray;            /* optional [nx*ny*nz] gray-scale intensity */
    const float* thresholds;      /* [nthr] ascending */
    int32_t nthr;
    const float* hvox;            /* optional [nx*ny*nz] user sizing field (mm; > 0 replaces the automatic size) */
} tn_volume_view;

typedef struct tn_grid_info {
    int32_t nx, ny, nz, nbx, nby, nbz;
    int64_t nslot;
    float hbase, hmin, hmax, hcurv;
    int32_t R;                    /* the brick label radius */
    int32_t limit_sweeps;         /* gradient-limiting sweeps run */
    int32_t overflow_bricks;      /* bricks with more than 16 labels nearby */
    int32_t gm;                   /* gray-scale thresholds (0: none) */
    int32_t prob_fields;          /* Grid::prob_fields: the fields are probabilities (a tissue-probability
                                     input, src/v2m_grid.cpp:343-345); tessellation reads it */
} tn_grid_info;

typedef struct tn_grid tn_grid;
void tn_grid_params_default(tn_grid_params* p);
int tn_grid_build(const tn_volume_view* vol, const tn_grid_params* prm, int32_t device, tn_grid** out);
int tn_grid_get_info(const tn_grid* g, tn_grid_info* info);
int tn_grid_download(const tn_grid* g, int32_t* bl_cnt, uint16_t* bl_lab, int32_t* bl_slot, float* phi, float* h,
                     uint8_t* grade, float* gTW);
int tn_grid_destroy(tn_grid* g);


/*
 * A relaxation plan: the grid and the seed nodes resident on one device, with
 * every buffer the loop needs allocated once. tn_plan_run restores the seeds on
 * the device and relaxes them, so repeated runs do the same work with no host
 * traffic; the relaxed nodes stay on the device until tn_plan_download. This
 * is relax_cl split at its host/device crossings (upload at
 * src/v2m_gpu.cpp:187-204, loop at :303-405, read-back at :514-517).
 */
typedef struct tn_plan tn_plan;

int tn_plan_create(const tn_grid_view* g, const tn_relax_params* prm, const tn_nodes_view* seeds, int32_t device,
                   tn_plan** out);
int tn_plan_run(tn_plan* plan, tn_relax_stats* st);
int tn_plan_download(tn_plan* plan, tn_nodes_view* nd);
int tn_plan_destroy(tn_plan* plan);

/*
 * The steps after the grid build without leaving the device: seed_cpu on a built
 * grid (as tn_seed_count / tn_seed, its grades included), and a relaxation plan
 * over it (as tn_plan_create; the plan copies the grid device to device, so the
 * grid may be destroyed afterwards). No counterpart in the reference, whose
 * pipeline hands the Grid from build_grid_cpu to seed_cpu and relax_cl on the host
 * (src/v2m_pipeline.cpp:425-500). The count call keeps its nodes on the device, in
 * the grid's handle, until the next count call or the grid's destruction; a
 * tn_seed_from_grid with the same parameters (compared byte for byte) downloads
 * those instead of seeding again. Calls on one grid from several threads race.
 */
int tn_seed_count_from_grid(const tn_grid* g, const tn_relax_params* prm, int64_t* n);
int tn_seed_from_grid(const tn_grid* g, const tn_relax_params* prm, int64_t n, float* P, uint16_t* lab, uint8_t* typ,
                      uint16_t* part, uint16_t* part3);
int tn_plan_create_from_grid(const tn_grid* g, const tn_relax_params* prm, const tn_nodes_view* seeds,
                             tn_plan** out);

/* The hash geometry of a plan (V2mHash, src/opencl/v2m_particle_body.cl:21-27;
 * built at src/v2m_particles.cpp:738-761): b0, origin, levels, dims, offsets. */
typedef struct tn_hash_info {
    float b0, ox, oy, oz;
    int32_t nlev;
    int32_t dim[TN_MAXLEV][3];
    int32_t off[TN_MAXLEV];
    int32_t nkeys;
} tn_hash_info;

int tn_plan_hash_info(const tn_plan* plan, tn_hash_info* hi);

/*
 * Stages of one relaxation, for checking each against the CPU path. They act
 * on the plan's current device state (the seeds after tn_plan_create or
 * tn_plan_reset) and copy results to host arrays (any pointer may be null).
 *   tn_plan_rebuild: keys, the stable sort, cell starts, the K-nearest lists
 *     (the rebuild lambda, src/v2m_particles.cpp:771-820). key[n], sorted[n],
 *     cstart[nkeys+1], nbr[n*TN_K], nnb[n].
 *   tn_plan_force: the spring forces for the current lists (v2m_force,
 *     src/opencl/v2m_particle_body.cl:161-184). F[4n].
 *   tn_plan_step: `iters` whole iterations of the loop (src/v2m_particles.cpp:828-947),
 *     rebuilds included; iteration numbering continues across calls.
 */
int tn_plan_reset(tn_plan* plan);

/*
 * Per-kernel device times of a plan, from HIP event pairs recorded on the
 * stream of each launch (the OpenCL host's V2M_CL_PROFILE timings,
 * src/v2m_gpu.cpp:250-258, without its extra synchronisations). Off by default
 * (on with prm->verbose or TN_KERNEL_TIMING in the environment); switching
 * clears the totals. Kernel classes: keys, scan, radix_sort, permute, neighbors
 * (the rebuild), force, move_easy, move, move_int, move_rest, move_def, stats,
 * refresh, control (an iteration). move, move_int and move_rest run at the same
 * time on three streams, so their times overlap. items: the nodes the timed
 * launches processed (cells for scan; for move and move_int, the interface and
 * interior nodes queued for the trapped step; for move_rest, the junction and
 * other nodes queued for the general step; for move_def, the nodes the move and
 * move_int passes handed to it; for refresh, the listed nodes). The
 * tn_relax_stats ms_* fields are sums of these when timing is on.
 */
typedef struct tn_kernel_time {
    char name[16];
    int64_t launches;
    double ms;
    int64_t items;
} tn_kernel_time;

int tn_plan_set_timing(tn_plan* plan, int32_t on);
/*
 * The same for a subset of the classes: bit k of mask times class k in the order
 * above (keys = bit 0 ... control = bit 13); the others record no events. A caller
 * that needs one class's times in a timed region pays for that class's events only.
 * No counterpart in the reference (its profiling is all or nothing).
 */
int tn_plan_set_timing_mask(tn_plan* plan, uint32_t mask);
int tn_plan_timings(const tn_plan* plan, tn_kernel_time* out, int32_t cap, int32_t* count);
int tn_plan_rebuild(tn_plan* plan, int32_t* key, int32_t* sorted, int32_t* cstart, int32_t* nbr, int32_t* nnb);
int tn_plan_force(tn_plan* plan, float* F);
int tn_plan_step(tn_plan* plan, int32_t iters, tn_relax_stats* st);

/*
 * The field queries of the plan's grid at m points X[3m] (host arrays), as the
 * move evaluates them on the device: phi[k] = phi_a (v2m_phi_f,
 * src/opencl/v2m_seed_body.cl:33-39), psi[k] and grad[3k..3k+2] = psi_ab and its
 * gradient (v2m_psi_grad, :211-247), label[k] and second[k] = the owner and the
 * runner-up (v2m_label_of, :116-160, over v2m_labels_near, :42-59; TN_NOLAB when
 * there is none). Any output may be null.
 */
int tn_plan_field_points(tn_plan* plan, int32_t a, int32_t b, int64_t m, const float* X, float* phi, float* psi,
                         float* grad, int32_t* label, int32_t* second);

/*
 * A differentiable relaxation and its adjoint. No counterpart in the reference
 * (its relaxation is not differentiable: hinge bars, nearest-voxel sizes,
 * projections and FIRE resets); for callers that fit a sizing field or seed
 * positions by gradient descent. The surrogate keeps the reference's bar force and
 * Jacobi step (src/opencl/v2m_particle_body.cl:161-184, 844-870) with a softplus
 * bar law, a smooth active-bar count and a sizing grid (trilinear in smoothstep cell
 * fractions); interior nodes
 * move, the others stay; lists and tets are fixed. After `iters` steps the loss is
 * the mean truss energy plus wq x the mean (1 - smoothed Joe-Liu quality) of the
 * tets (src/v2m_tetra.cpp:100-120). The adjoint recomputes the steps from
 * checkpoints every `ckpt_every` steps. x0, theta, xT and the gradients are DEVICE
 * pointers (x as 3n floats, AoS; theta[(z gy + y) gx + x]); stream may be null (the
 * null stream). tn_adj_forward returns the loss on the host; tn_adj_backward needs
 * the forward of the same inputs before it and gives dL/dx0 and dL/dtheta.
 */
typedef struct tn_adj_problem {
    int64_t n;
    const int32_t* nbr;     /* [n * TN_K] host: node-major lists (entries past nnb ignored) */
    const int32_t* nnb;     /* [n] host */
    const uint8_t* typ;     /* [n] host: TN_INTERIOR nodes move */
    int64_t ntet;
    const int32_t* tet;     /* [4 ntet] host: node indices */
    int32_t gx, gy, gz;     /* the sizing grid */
    float origin[3];        /* position of theta[0] (mm) */
    float spacing[3];       /* grid spacing (mm) */
    float dt, fscale, fsurf;
    float beta;             /* softplus sharpness (1/mm) */
    float wq;               /* weight of the quality term */
    float eps_v;            /* volume smoothing (mm^3) */
    int32_t iters, ckpt_every;
} tn_adj_problem;

typedef struct tn_adj tn_adj;
int tn_adj_create(const tn_adj_problem* p, int32_t device, tn_adj** out);
int tn_adj_forward(tn_adj* a, const float* x0, const float* theta, float* xT, double* loss, void* stream);
int tn_adj_backward(tn_adj* a, const float* theta, float* grad_x0, float* grad_theta, void* stream);
int tn_adj_loss_terms(const tn_adj* a, double* terms);   /* loss, energy, quality term of the last forward */
int tn_adj_destroy(tn_adj* a);

#ifdef __cplusplus
}
#endif

#endif /* TRUSSNET_AMD_H */
