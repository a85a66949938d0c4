// Launch wrappers of every kernel, and the device control block.
#pragma once
#include "nv_device.h"
#include "../../include/nuvort.h"

// status bits set by kernels (mapped to NV_ERR_* by the host)
#define NVK_NAN 1
#define NVK_ZERO_LENGTH 2
#define NVK_CAPACITY 4
#define NVK_SINGULAR 8
#define NVK_TREE_DEPTH 16

// device-resident counters shared by the topology kernels
struct DCtl {
    int n;             // pcount
    int status;        // NVK_* bits
    int recon_count;
    int remove_count;
    int nevents;       // reconnection events logged since the host last read them
    int nactive;       // active points counted by the pinsert pre-filter
    int ncand;         // rows in the current candidate list
    int nempty;        // rows in the current empty-slot list
    int nvel;          // points whose velocity the last velocity launch computed
    int itime;         // the step being taken (src/run.f90 loop index)
    double t;          // simulation time
    unsigned long long nvel_total;  // nvel summed over the steps k_clock_tick closed
    int noise_ctr;     // noisey_xflow: index of the noise table's current value (0-based)
    int n_old;         // pinsert: pcount when it started (summed over tangles)
    double total_length;   // pinsert: sum of |x - ghosti| over the points it started with
    int empty_after;   // pinsert: empty slots below pcount when it ended (summed over tangles)
    int abort;         // a recorded octree step found its build incomplete: its later phases do nothing
};

// src/diagnostic.f90:9-18 calculate_diagnostics, reduced on the device
struct DDiag {
    double maxu, maxdu, sum_usup, sum_umf, sum_curv, curv_min, curv_max;
    int active, pad;
};

// velocity (k_velocity.hip)
hipError_t nvk_src64(hipStream_t st, DState s, int rows, double *rec);
hipError_t nvk_bs64(hipStream_t st, DState s, Tangles tg, const Item *items, int nitems, int *queue, int grid,
                    const double *rec, const Shifts *sh, double kappa, double *part);
// spatial order of the FP32 source records: perm (in v0) and its inverse
hipError_t nvk_order(hipStream_t st, DState s, Tangles tg, int rows, const double *box, unsigned long long *k0,
                     int *v0, unsigned long long *k1, int *v1, int *iperm, int *hist, int *scan_tmp);
// ts: sources per tile of the FP32C direct sum (128 or 256)
hipError_t nvk_src32(hipStream_t st, DState s, int rows, const int *perm, void *rec, void *tiles, int ts);
hipError_t nvk_bs32c(hipStream_t st, DState s, Tangles tg, const Item *items, int nitems, int *queue, int grid,
                     const void *rec, const void *tiles, const int *iperm, const double *rec64, const Shifts *sh,
                     double kappa, float rn2, double *part, int tpt, int kb, int ts);
hipError_t nvk_finish(hipStream_t st, DState s, Phys p, Tangles tg, int rows, const double *part, int mode,
                      int do_ab, DCtl *ctl, const double *mir);
// solid walls (k_walls.hip, k_topology.hip): the image vortices' sum
// mir[3][rows] of the direct sum, the brute close search with pinned points,
// and the topology phases as the reference's serial loops
hipError_t nvk_mirror_bs64(hipStream_t st, DState s, Phys p, Tangles tg, int rows, double *mir);
hipError_t nvk_pclose_walls(hipStream_t st, DState s, Tangles tg, int rows);
hipError_t nvk_pinsert_walls(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl);
hipError_t nvk_precon_walls(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl, nv_recon_event *ev, int ev_cap);
hipError_t nvk_wall_recon(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl);
hipError_t nvk_premove_walls(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl);
hipError_t nvk_reduce_probe(hipStream_t st, DState s, Tangles tg, int rows, const double *part, double *sink);
size_t nvk_rec32_bytes();
size_t nvk_tileinfo_bytes(int rows);
int nvk_occupancy_bs64();
int nvk_occupancy_bs32c(int tpt, int kb, int ts);

// move the state to a wider layout (k_topology.hip)
hipError_t nvk_relayout(hipStream_t st, DState o, DState nu, const int *row2t, const int *base_old,
                        const int *base_new, int rows_old);

// device clock (k_topology.hip): the step's itime and t live in DCtl so that a
// captured step graph reads them at replay time
hipError_t nvk_clock_set(hipStream_t st, DCtl *ctl, double t, int itime);
hipError_t nvk_clock_begin_step(hipStream_t st, DCtl *ctl);
hipError_t nvk_clock_end_step(hipStream_t st, DCtl *ctl, double dt);
// recorded-step head (itime, counters, queue) and tail (wrap, t, NaN test)
hipError_t nvk_step_begin(hipStream_t st, DCtl *ctl, int *tnact, int K, int *queue);
hipError_t nvk_step_end(hipStream_t st, DState s, Phys p, int n, DCtl *ctl, double dt);

// boundary / topology / close search (k_topology.hip)
hipError_t nvk_ghosts(hipStream_t st, DState s, Phys p, int n);
// curvature and stretching of every row and the reductions of
// velocity_info / curv_info (src/diagnostic.f90:21-74) over rows [0, n) of a
// single tangle; part holds nvk_diag_parts(n) doubles
int nvk_diag_parts(int n);
hipError_t nvk_diagnostics(hipStream_t st, DState s, int n, double *part, DDiag *out);
hipError_t nvk_wrap(hipStream_t st, DState s, Phys p, int n);
hipError_t nvk_nan(hipStream_t st, DState s, int n, DCtl *ctl);
hipError_t nvk_pinsert_pre(hipStream_t st, DState s, Phys p, Tangles tg, int n, int *cand, int *empty, DCtl *ctl);
hipError_t nvk_pinsert_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, const int *empty,
                              DCtl *ctl);
hipError_t nvk_precon_pre(hipStream_t st, DState s, Phys p, int n, int *cand);
// precon: every candidate staged at once into rs (rs_cap records of
// nvk_recstage_bytes()), then the serial pass
size_t nvk_recstage_bytes();
hipError_t nvk_precon_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl,
                             nv_recon_event *ev, int ev_cap, void *rs, int rs_cap);
hipError_t nvk_premove_pre(hipStream_t st, DState s, Phys p, int n, int *cand);
hipError_t nvk_premove_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl);
hipError_t nvk_pclose(hipStream_t st, DState s, Tangles tg, int rows, const Item *items, int nitems, void *spheres,
                      double *pd, int *pj);
size_t nvk_sphere_bytes();
// the brute close search of one tangle in caller memory: positions [3][ld],
// working memory `work` of nvk_pclose_flat_bytes(n, S) bytes (S chunks per target)
size_t nvk_pclose_flat_bytes(int n, int S);
hipError_t nvk_pclose_flat(hipStream_t st, const double *x, int ld, const int *infront, const int *behind,
                           int *closest, double *closestd, int n, int S, void *work);
