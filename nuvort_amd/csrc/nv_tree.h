// Barnes-Hut octree on the device (src/tree.f90): build, walks, dump.
#pragma once
#include "nv_kernels.h"

struct TreeBuf;

TreeBuf *nv_tree_new();
void nv_tree_free(TreeBuf *t);
// construct_tree (src/tree.f90:38-157) from the current positions/ghosts
int nv_tree_build(TreeBuf *t, hipStream_t st, DState d, Phys p, int n, DCtl *ctl, nv_stats *stats);
// tree_walk (src/tree.f90:250-316) for every point and image: writes the
// unshifted and image sums to part[0..6)[npad] (the layout k_finish reads)
// (precision FP32C: FP32 node terms, FP64 leaves within sqrt(rn2))
int nv_tree_walk_velocity(TreeBuf *t, hipStream_t st, DState d, Phys p, const Shifts &sh, int n, int npad,
                          int precision, double rn2, double *part, DCtl *ctl, nv_stats *stats);
// bytes of one walk record read by the velocity walk of that precision
int nv_tree_walk_record_bytes(int precision, double theta);
// the layout of its per-image partials, as nvk_finish's mode: 4 for the FP32
// walk (images other than the unshifted one as floats), 3 for the FP64 walk
int nv_tree_walk_partial_mode(int precision, double theta);
int nv_tree_walk_counters(TreeBuf *t, hipStream_t st, nv_stats *stats);
int nv_tree_walk_counters_async(TreeBuf *t, hipStream_t st);
void nv_tree_walk_counters_take(const TreeBuf *t, nv_stats *stats);
// pclose_tree / closest_tree (src/tree.f90:190-244) over the (stale) tree
int nv_tree_walk_close(TreeBuf *t, hipStream_t st, DState d, double bsize, int n, nv_stats *stats);
// every node, depth-first, empty ones included
int nv_tree_dump(TreeBuf *t, hipStream_t st, nv_tree_node *buf, int cap, int *n);

// Recorded octree steps (launches only between prepare and take): the build,
// velocity walk and close search on device-held counts (*dn rows of the row
// range prepared; the node ranges in the build's level table), within bounds
// planned from the last build's level sizes.  A build that outgrows them sets
// ctl->abort, which switches off every later phase of the step.
// prepare: plan and grow every buffer (outside capture); NV_ERR_STATE before
// a first build
int nv_tree_graph_prepare(TreeBuf *t, hipStream_t st, int rows);
int nv_tree_graph_build(TreeBuf *t, hipStream_t st, DState d, Phys p, const int *dn, DCtl *ctl);
int nv_tree_graph_walk(TreeBuf *t, hipStream_t st, DState d, Phys p, const Shifts &sh, int precision, double rn2,
                       double *part, DCtl *ctl, const int *dn);
int nv_tree_graph_close(TreeBuf *t, hipStream_t st, DState d, double bsize, const int *dn, DCtl *ctl);
// changes whenever a buffer the recorded launches hold moves
int nv_tree_graph_gen(const TreeBuf *t);
// after a replay: queue the level table and walk counters to pinned memory,
// then (the stream past them) take them; true when the plan should be redone
int nv_tree_graph_take_async(TreeBuf *t, hipStream_t st);
bool nv_tree_graph_take(TreeBuf *t, nv_stats *stats);

// Solid walls: tree_walk_mirror (src/tree.f90:321-392) over every reflection
// into mir[3][npad] (FP64, the reference's arithmetic), and pclose_tree with
// pinned points (src/tree.f90:190-244); eagerly (host counts) or in a
// recorded step (device counts, as the graph functions above)
int nv_tree_walk_mirror(TreeBuf *t, hipStream_t st, DState d, Phys p, int n, int npad, double *mir, DCtl *ctl);
int nv_tree_close_walls(TreeBuf *t, hipStream_t st, DState d, int n);
int nv_tree_graph_mirror(TreeBuf *t, hipStream_t st, DState d, Phys p, double *mir, DCtl *ctl, const int *dn);
int nv_tree_graph_close_walls(TreeBuf *t, hipStream_t st, DState d, const int *dn, DCtl *ctl);
