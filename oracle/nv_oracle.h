/*
 * Test infrastructure only: the oracle.
 *
 * A plain-C, FP64, loop-by-loop restatement of NuVort's filament velocity and
 * timestep path (src/run.f90:24-65 and the routines it calls).  The HIP
 * library's parity tests compare against it, and bench.py times it as the CPU
 * baseline; nothing in the product links or calls it.  Every function cites the
 * reference lines it restates; evaluation order follows the Fortran source
 * term by term (built with -ffp-contract=off) so that results are
 * bit-identical to the reference compiled without FMA.
 *
 * Conventions (same as the reference's type qvort, src/cdata.f90:19-29):
 * point p (1-based) is stored at row p-1; links infront/behind/closest hold
 * 1-based indices, 0 marks an empty slot.  Vectors are [cap][3] doubles.
 */
#ifndef NV_ORACLE_H
#define NV_ORACLE_H
#include <stdint.h>

#define NVO_OK 0
#define NVO_ERR_CAPACITY 1
#define NVO_ERR_ZERO_LENGTH 2
#define NVO_ERR_SINGULAR 3
#define NVO_ERR_NAN 4
#define NVO_ERR_UNSUPPORTED 5
#define NVO_ERR_TREE_DEPTH 6

typedef struct {
    int32_t n;            /* pcount */
    int32_t cap;          /* rows allocated by the caller */
    int32_t itime;
    int32_t recon_count;
    int32_t remove_count;
    int32_t wall_recon_count; /* wall_recon's counter (src/reconnection.f90:66-150) */
    double t;
    double *x, *u, *u1, *u2, *u_sup, *u_mf, *gi, *gb; /* [cap][3] */
    int32_t *infront, *behind, *closest;              /* [cap] */
    double *closestd;                                 /* [cap] */
    int32_t *pinnedi, *pinnedb;                       /* [cap] 0/1 */
    int32_t *wpinned;                                 /* [cap][3] */
} nvo_state;

typedef struct {
    double delta, dt, quant_circ, corea;
    double box[3];
    int32_t periodic[3];          /* 1 = periodic, 0 = open or solid */
    int32_t velocity;             /* 0 LIA, 1 BS, 2 Tree */
    double tree_theta;
    int32_t tree_extra_correction;
    int32_t normal_velocity;      /* 0 zero, 1 xflow, 2 noisey_xflow, 3 shear_xflow, 4 ychannel, 5 channel */
    double norm_vel_xflow;
    double alpha[2];
    double super_velocity[3];
    double normal_fluid_cutoff, t_zero_normal_fluid;
    int32_t recon_shots;
    int32_t threads;              /* OpenMP threads, 0 = runtime default */
    double norm_shear_omega;      /* shear_xflow frequency */
    const double *noise;          /* noisey_xflow table (noise_n values) */
    int32_t noise_n;
    int32_t *noise_counter;       /* 1-based current entry; advanced by every pmotion */
    int32_t solid[3];             /* 1 = solid walls at +-box/2 in that direction */
    int32_t pad_solid;
} nvo_params;

/* One octree node as the reference holds it (src/tree.f90:20-32), stored in
 * depth-first pre-order with the reference's child order. */
typedef struct {
    double pos[3], width, cent[3], circ[3];
    int32_t pcount;
    int32_t leaf_infront;   /* parray(1)%infront at build time, 0 if empty */
    int32_t child[8];       /* node indices, -1 if none */
} nvo_node;

typedef struct {
    nvo_node *nodes;
    int32_t nnodes, cap;
} nvo_tree;

typedef struct {
    int32_t itime, i, j, pad;
    double t, mid[3], angle, dl;
} nvo_recon_event;

typedef struct {
    nvo_recon_event *ev;
    int32_t n, cap;
} nvo_recon_log;

int nvo_shift_table(const nvo_params *p, double shifts[26][3]);
int nvo_mirror_table(const nvo_params *p, int32_t reflect[26][3]);
void nvo_ghostp(nvo_state *s, const nvo_params *p);
int nvo_tree_build(const nvo_state *s, const nvo_params *p, nvo_tree *tr);
void nvo_tree_free(nvo_tree *tr);
int nvo_velocity(nvo_state *s, const nvo_params *p, const nvo_tree *tr,
                 int32_t first, int32_t count, double *u_out);
int nvo_pmotion(nvo_state *s, const nvo_params *p, const nvo_tree *tr);
void nvo_ab_update(nvo_state *s, const nvo_params *p);
int nvo_pinsert(nvo_state *s, const nvo_params *p);
double nvo_total_length(void); /* total_length of the last nvo_pinsert (src/line.f90:29,38) */
double nvo_avg_sep(void);      /* avg_sep of the last nvo_pinsert (src/line.f90:95) */
void nvo_pclose(nvo_state *s, const nvo_params *p);
void nvo_pclose_tree(nvo_state *s, const nvo_params *p, const nvo_tree *tr);
int nvo_precon(nvo_state *s, const nvo_params *p, nvo_recon_log *log);
void nvo_premove(nvo_state *s, const nvo_params *p);
void nvo_loop_killer(nvo_state *s, int32_t particle);
void nvo_enforce_boundary(nvo_state *s, const nvo_params *p);
void nvo_wall_recon(nvo_state *s, const nvo_params *p);
int nvo_nan_check(const nvo_state *s);
int nvo_step(nvo_state *s, const nvo_params *p, int32_t nsteps, nvo_recon_log *log);
int64_t nvo_tree_eval_count(void);

#endif
