/*
 * nuvort.h -- C-ABI of the MI355X filament velocity and timestep library.
 *
 * The reference (NuVort, Fortran) has no plugin or FFI layer: its hot path
 * is a set of argument-less module procedures that read and write module
 * globals (the point array f(:) of src/cdata.f90:31, run.in parameters of
 * src/cdata.f90:76-127, the octree vtree of src/tree.f90:34) and are called
 * in a fixed order from the main loop (src/run.f90:24-65).  Each entry
 * point below replaces one of those procedures; the citation on each says
 * which.  A caller (the Fortran shim nuvort_amd/fortran/nuvort_hip.f90, or
 * ctypes) keeps the reference's call sequence and only swaps the callee.
 *
 * Conventions
 *  - Points are 1-based as in the reference: point p lives in row p-1 of
 *    every array; infront/behind/closest hold 1-based indices; infront == 0
 *    marks an empty slot (src/cdata.f90:19-29).  Vector fields are
 *    [rows][3] doubles, row-major.
 *  - The library copies host arrays in and out inside nv_set_state /
 *    nv_get_state and never keeps a host pointer (the reference reallocates
 *    f with move_alloc while inserting points, src/line.f90:51-58).
 *  - State lives on the device between calls.  Every call is synchronous
 *    with respect to the host: it returns after its kernels finished.
 *  - Every call returns NV_OK or an error code; nv_last_error() describes
 *    the last error.  The Fortran shim maps non-zero codes to the
 *    reference's fatal_error (src/cdata.f90:245-255).
 *  - One context per host thread; contexts are independent.
 */
#ifndef NUVORT_H
#define NUVORT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NV_OK = 0,
    NV_ERR_ARG = 1,          /* bad argument or parameter */
    NV_ERR_HIP = 2,          /* HIP runtime failure */
    NV_ERR_CAPACITY = 3,     /* caller's array too small (nv_get_state) */
    NV_ERR_ZERO_LENGTH = 4,  /* no active point: src/line.f90:25-27 */
    NV_ERR_SINGULAR = 5,     /* tree singularity: src/tree.f90:289-295 */
    NV_ERR_NAN = 6,          /* NaN in a position: src/general.f90:172-191 */
    NV_ERR_UNSUPPORTED = 7,  /* solid walls / pinned points / other modes */
    NV_ERR_TREE_DEPTH = 8,   /* octree deeper than NV_TREE_MAX_LEVELS (points closer than box / 2^48) */
    NV_ERR_TIMESTEP = 9,     /* dt above the stability bound: src/initial.f90:169-180 */
    NV_ERR_STATE = 10        /* call out of order (e.g. walk without a tree) */
};

/* Octree levels the build supports: box / 2^48 is 7e-16 cm in a 0.2 cm box,
 * a few units in the last place of a position; points closer than that are
 * coincident in practice (the reference's recursion does not end on them). */
#define NV_TREE_MAX_LEVELS 48

enum { NV_VEL_LIA = 0, NV_VEL_BS = 1, NV_VEL_TREE = 2 };
enum { NV_BC_OPEN = 0, NV_BC_PERIODIC = 1, NV_BC_SOLID = 2 };
/* normal_velocity models of src/normal_fluid.f90:83-118 that have a closed
 * form or use the noise table (the DNS fields laizetDNS / turbDNS need data
 * files the reference does not ship) */
enum { NV_NF_ZERO = 0, NV_NF_XFLOW = 1, NV_NF_NOISEY_XFLOW = 2, NV_NF_SHEAR_XFLOW = 3, NV_NF_YCHANNEL = 4,
       NV_NF_CHANNEL = 5 };
enum { NV_PREC_FP64 = 0, NV_PREC_FP32C = 1 };
enum { NV_CLOSE_AUTO = 0, NV_CLOSE_BRUTE = 1, NV_CLOSE_TREE = 2 };

/* run.in parameters of src/cdata.f90:76-127 that the path reads, followed by
 * this library's own extensions. */
typedef struct {
    double delta;                /* run.in delta */
    double dt;                   /* run.in dt */
    double quant_circ;           /* run.in quant_circ (default 9.97e-4) */
    double corea;                /* run.in corea (default 8.244023e-9) */
    double box_size[3];          /* run.in box_size */
    int32_t boundary[3];         /* run.in boundary_x/y/z: NV_BC_* */
    int32_t velocity;            /* run.in velocity: NV_VEL_* */
    double tree_theta;           /* run.in tree_theta (0 = no tree) */
    int32_t tree_extra_correction; /* run.in tree_extra_correction (default 1) */
    int32_t normal_velocity;     /* run.in normal_velocity: NV_NF_* */
    double norm_vel_xflow;       /* run.in norm_vel_xflow (default 0.5) */
    double alpha[2];             /* run.in alpha */
    double super_velocity[3];    /* run.in super_velocity */
    double normal_fluid_cutoff;  /* run.in normal_fluid_cutoff (default 1e8) */
    double t_zero_normal_fluid;  /* run.in t_zero_normal_fluid (default 1e8) */
    double norm_shear_omega;     /* run.in norm_shear_omega (default 0), shear_xflow */
    int32_t recon_shots;         /* run.in recon_shots (default 1) */
    int32_t overide_timestep_check; /* run.in overide_timestep_check */
    /* extensions (no reference counterpart) */
    int32_t precision;           /* NV_PREC_FP64 (reference arithmetic) or NV_PREC_FP32C */
    double near_radius;          /* FP32C: FP64 near field radius, in units of delta */
    int32_t close_search;        /* NV_CLOSE_*: AUTO = tree when tree_theta > 0 */
    int32_t exact_order;         /* 1: sum each point's terms in the reference's order (one
                                    chunk per point, bit-exact FP64, slower); 0: chunked */
    int32_t capacity;            /* initial device rows (0 = 3 x pcount + 1024; at least 2 x pcount + 256) */
    int32_t targets_per_lane;    /* FP32C direct sum: points per lane, 1 or 2 (0 = tuned) */
    int32_t kahan_block;         /* FP32C: terms summed in plain FP32 between two compensated
                                    adds, 1 (every term), 2, 4, 8, 16 or 32 (0 = default 8) */
    int32_t step_launch;         /* nv_step: 0 = record one step's launches in a HIP graph and
                                    replay it every step, 1 = eager launches */
    int32_t source_tile;         /* FP32C direct sum: sources staged per LDS tile, 128 or 256
                                    (0 = tuned; default 256) */
} nv_params;

/* Host view of the filament point array, structure of arrays (the fields
 * of type qvort, src/cdata.f90:19-29, that the path reads or writes). */
typedef struct {
    int32_t pcount;      /* rows in use (the reference's pcount) */
    int32_t capacity;    /* rows the arrays below can hold */
    int32_t itime;       /* the reference's itime */
    int32_t recon_count; /* the reference's recon_count */
    int32_t remove_count;/* the reference's remove_count */
    int32_t wall_recon_count; /* the reference's wall_recon_count (src/cdata.f90:45) */
    double t;            /* the reference's t */
    double *x, *u, *u1, *u2, *u_sup, *u_mf, *ghosti, *ghostb; /* [capacity][3] */
    int32_t *infront, *behind, *closest;                      /* [capacity] */
    double *closestd;                                         /* [capacity] */
    int32_t *pinnedi, *pinnedb;                               /* [capacity] 0/1 */
    int32_t *wpinned;                                         /* [capacity][3] */
    double *curv, *stretch;  /* [capacity] diagnostics (nv_diagnostics); may be NULL */
} nv_state;

/* The quantities of calculate_diagnostics (src/diagnostic.f90:9-74) and of
 * the last pinsert (src/line.f90:38-39,95) that print_info writes to ts.log
 * (src/output.f90:20-44), for a single tangle. */
typedef struct {
    int32_t active;          /* count(f%infront > 0) */
    int32_t recon_count, remove_count, itime;
    double t;
    double maxu, maxdu;      /* max |u|, max |u1 - u2| over f(1:pcount) */
    double mean_u_sup, mean_u_mf;   /* sum |u_sup|, |u_mf| over f(1:pcount) / active */
    double kappa_bar, kappa_min, kappa_max;   /* curvature: mean over active, min over > 0, max */
    double total_length;     /* pinsert: sum of |x - ghosti| over the points it started with */
    double avg_sep;          /* total_length / (pcount at its start - empty slots at its end) */
    int64_t tree_eval;       /* interactions of the last tree walk */
    int32_t wall_recon_count;/* reconnections with solid walls (src/reconnection.f90:66-150) */
    int32_t pad;
} nv_diag;

/* One reconnection, as src/reconnection.f90:46-49 logs it plus the pair
 * (i, j local to the tangle, 1-based; tangle = 0 for a single system). */
typedef struct {
    int32_t itime, i, j, tangle;
    double t, mid[3], angle, dl;
} nv_recon_event;

/* One octree node in depth-first pre-order (child order of
 * src/tree.f90:311-314), the fields of type node (src/tree.f90:20-32). */
typedef struct {
    double pos[3], width, cent[3], circ[3];
    int32_t pcount, leaf_infront;
} nv_tree_node;

typedef struct {
    int64_t pair_evals;        /* direct sum (target, segment, image) evaluations done */
    int64_t near_evals;        /* of which evaluated in FP64 by the FP32C kernel */
    int64_t tree_evals;        /* tree interactions (the reference's tree_eval) */
    int64_t tree_nodes_visited;/* node records read by tree walks */
    int32_t tree_nodes;        /* nodes of the last tree, empty ones included */
    int32_t tree_levels;       /* levels of the last tree */
    double ms_velocity;        /* last velocity kernel time (HIP events) */
    double ms_tree_build;      /* last tree build time */
    double ms_close;           /* last close-point search time */
    double ms_topology;        /* last pinsert+precon+premove time */
    double ms_step;            /* last nv_step call, per step */
    int64_t point_updates;     /* sum over steps of the points moved by pmotion */
    int64_t pair_launches;     /* direct-sum kernel launches */
    double pair_ms_total;      /* summed duration of those launches (HIP events) */
    int64_t pair_evals_total;  /* summed pair evaluations of those launches */
    int64_t walk_launches;     /* tree velocity walk launches timed (eager steps and calls; recorded
                                  octree steps are not timed by phase) */
    double walk_ms_total;      /* summed duration of those launches */
    int64_t walk_nodes_total;  /* summed node records read by those walks */
    int32_t targets_per_lane;  /* direct-sum variant in use: points per lane */
    int32_t kahan_block;       /* FP32C terms between compensated adds */
    int32_t items_per_cu;      /* direct-sum work items per CU (sets the chunk size) */
    int32_t tuned;             /* 0 defaults, 1 timed by nv_autotune, 2 read from its cache */
    double tune_ms;            /* direct-sum launch + chunk reduction time of the chosen variant */
    int32_t tree_builds;       /* octree builds */
    int32_t tree_sync_builds;  /* of which ran the level loop with a host round trip per level (the
                                  first build, and a tree that outgrew the previous build's levels) */
    int32_t tree_replays;      /* octree steps nv_step took as a recorded step (one synchronisation) */
    int32_t tree_redone;       /* recorded octree steps whose build outgrew the recorded plan and
                                  that were taken again eagerly */
    int32_t source_tile;       /* direct-sum variant in use: sources per LDS tile */
    int32_t pad;
} nv_stats;

/* lifecycle -------------------------------------------------------------- */
/* Replaces the setup part of init_setup (src/initial.f90:14-146): checks the
 * parameters (timestep_check, open boundaries with a tree, ...) and creates
 * the periodic shift table (src/boundary.f90:143-175). */
int nv_create(const nv_params *params, void **ctx);
int nv_destroy(void *ctx);
const char *nv_last_error(void *ctx);

/* state ------------------------------------------------------------------ */
/* Upload f(1:pcount) (replaces data_restore / the ICs, src/initial.f90:151-166). */
int nv_set_state(void *ctx, const nv_state *host);
/* Download into caller arrays; on NV_ERR_CAPACITY host->pcount holds the rows needed. */
int nv_get_state(void *ctx, nv_state *host);
int nv_set_time(void *ctx, double t, int32_t itime);
/* The noise table of normal_velocity noisey_xflow (replaces load_noise,
 * src/normal_fluid.f90:156-162, which reads 100000 values from noise_file):
 * the library copies the n values and restarts the table's counter at its
 * first entry.  pmotion advances the counter by one per call and wraps it
 * after the last entry (randomise_normal_fluid, src/normal_fluid.f90:340-347). */
int nv_set_noise(void *ctx, int32_t n, const double *noise);
/* Batches of independent tangles (no reference counterpart: the ensemble
 * form of the same path).  The arrays hold the tangles one after another,
 * tangle k with pcounts[k] rows and links local to it (1-based); every
 * phase and nv_step then advance all tangles together, the direct sums of
 * all of them in one launch.  With tree_theta > 0 each tangle gets its own
 * octree and device buffers on a stream of its own; nv_step launches every
 * tangle's recorded step before waiting for any.  The calls that return
 * one tangle's data (nv_get_state, nv_get_tree, the on-disk formats and the
 * diagnostics) take a single-tangle state. */
int nv_set_state_batch(void *ctx, int32_t ntangles, const int32_t *pcounts, const nv_state *host);
int nv_get_state_batch(void *ctx, int32_t ntangles, int32_t *pcounts, nv_state *host);

/* phases (in the order src/run.f90:24-65 calls them) -------------------- */
int nv_ghostp(void *ctx);           /* src/boundary.f90:10-19 ghostp */
int nv_construct_tree(void *ctx);   /* src/tree.f90:38-59 construct_tree */
int nv_pmotion(void *ctx);          /* src/timestep.f90:16-45 pmotion */
int nv_velocity(void *ctx);         /* src/timestep.f90:21-27: velocities only, no AB update */
int nv_pinsert(void *ctx);          /* src/line.f90:18-96 pinsert */
int nv_pclose(void *ctx);           /* src/line.f90:135-160 pclose */
int nv_pclose_tree(void *ctx);      /* src/tree.f90:190-203 pclose_tree */
int nv_precon(void *ctx);           /* src/reconnection.f90:11-61 precon */
int nv_premove(void *ctx);          /* src/line.f90:102-130 premove */
int nv_wall_recon(void *ctx);       /* src/reconnection.f90:66-150 wall_recon (src/run.f90:43) */
int nv_enforce_boundary(void *ctx); /* src/boundary.f90:91-138 enforce_boundary */
int nv_empty_tree(void *ctx);       /* src/tree.f90:160-186 empty_tree, src/run.f90:58-62 */
int nv_nan_check(void *ctx);        /* src/general.f90:172-191 NAN_finder */

/* nsteps iterations of the src/run.f90:24-65 loop body (diagnostics and
 * file output excluded), state kept on the device throughout.  With
 * step_launch 0 each step is one recorded HIP graph replayed with one host
 * synchronisation (the octree step too, when its velocity is the tree walk
 * and its close search the tree's); a fatal condition raised on the device
 * is returned at the end of the step in which it arose. */
int nv_step(void *ctx, int32_t nsteps);

/* diagnostics, output and restart (src/diagnostic.f90, src/output.f90,
 * src/initial.f90:151-166) ------------------------------------------------ */
/* calculate_diagnostics (src/diagnostic.f90:9-18): ghostp, then each point's
 * curvature and stretching (kept with the state, written by nv_data_dump
 * and nv_printf) and the ts.log quantities into *d. */
int nv_diagnostics(void *ctx, nv_diag *d);
/* data_dump (src/output.f90:85-94): the point array as var.dat, Fortran
 * unformatted sequential records (pcount, recon_count, itime, t, then f as
 * 248-byte qvort records) that the reference's data_restore reads. */
int nv_data_dump(void *ctx, const char *path);
/* data_restore (src/initial.f90:151-166): a var.dat (the reference's or
 * nv_data_dump's) uploaded as the state; the next step is itime + 1. */
int nv_data_restore(void *ctx, const char *path);
/* printf (src/output.f90:47-70): the var%04d.log stream file the MATLAB
 * readers load (t, pcount, x, y, z, infront, u, u_mf, curv, stretch). */
int nv_printf(void *ctx, const char *path);
/* print_info's ts.log line (src/output.f90:23-37), appended to path, with
 * the header line first when itime == shots; d from nv_diagnostics. */
int nv_print_info(void *ctx, const nv_diag *d, int32_t shots, const char *path);
/* tree_print (src/tree.f90:96-102): one line per leaf of the current tree in
 * depth-first order, posx, posx+w, posy, posy+w, posz, posz+w, pcount,
 * appended to path (the reference's tree_mesh%03d.log). */
int nv_tree_print(void *ctx, const char *path);
/* The file formats alone, on host arrays (no context, no GPU): write and
 * read var.dat (on NV_ERR_CAPACITY host->pcount holds the rows needed), and
 * write a var%04d.log. */
int nv_var_write(const char *path, const nv_state *host);
int nv_var_read(const char *path, nv_state *host);
int nv_var_log_write(const char *path, const nv_state *host);

/* outputs ---------------------------------------------------------------- */
/* Reconnections since the last call (the pairs of src/reconnection.f90:11-61). */
int nv_get_recon_events(void *ctx, nv_recon_event *buf, int32_t cap, int32_t *n);
/* The current tree in depth-first order (what tree_print writes,
 * src/tree.f90:96-102, plus internal nodes and moments). */
int nv_get_tree(void *ctx, nv_tree_node *buf, int32_t cap, int32_t *n);
int nv_get_stats(void *ctx, nv_stats *stats);

/* No counterpart in the reference.  The brute close-point search of
 * nv_pclose (src/line.f90:135-160 pclose, one tangle) for callers that
 * cannot allocate device memory: code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc.  It takes no context and calls no
 * allocator.  Inputs and outputs are device memory the caller holds:
 * positions x component-major with leading dimension ld >= n (component k of
 * point p at x[k*ld + p-1]), infront/behind 1-based (0 = empty slot), and
 * closest/closestd, which receive what pclose writes (closestd = 100 and
 * closest left as it was where no candidate exists).  A point whose links
 * leave 1..n is treated as empty.  The search's working memory is a static
 * device array of NV_NOALLOC_SCRATCH_MIB MiB per GPU, held until the process
 * ends, which bounds n at NV_NOALLOC_MAX_POINTS (NV_ERR_CAPACITY above).
 * Kernels are launched on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns without waiting for them.  Calls that overlap on one
 * GPU share its array and race. */
#define NV_NOALLOC_SCRATCH_MIB 64
#define NV_NOALLOC_MAX_POINTS 5341952
int nv_pclose_noalloc(int32_t n, int32_t ld, const double *x, const int32_t *infront, const int32_t *behind,
                      int32_t *closest, double *closestd, void *stream);

/* Autotuner (no reference counterpart).  Times the direct-sum variants on
 * the uploaded state -- points per lane (targets_per_lane), Kahan block,
 * sources per LDS tile (source_tile) and work items per CU, each left at 0
 * in nv_params -- one coordinate at a time, and keeps the fastest.  The state is not modified (only scratch
 * buffers are written).  With a cache_path, a choice is looked up first
 * under (GPU architecture, precision, tangle count, log2 of the total and
 * of the largest tangle's points) and a new one is appended to the file;
 * NULL tunes without a cache.  Call it between steps. */
int nv_autotune(void *ctx, const char *cache_path);

#ifdef __cplusplus
}
#endif
#endif
