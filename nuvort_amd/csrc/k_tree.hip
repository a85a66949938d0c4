// Barnes-Hut octree of src/tree.f90 built and walked on the device.
//
// Build (construct_tree / build_tree / new_tree, src/tree.f90:38-157):
//  1. every active point computes its path through the octree from the
//     exact FP64 boxes the reference builds (corner accumulated as
//     pos + width/2, child width width/2, half-open membership
//     [pos, pos+width), child order fbl fbr ftl ftr bbl bbr btl btr i.e.
//     c = 4*y + 2*z + x).  Four bits per level: 0..7 the child, 8 = in the
//     parent but in no child (the sliver a rounded upper bound leaves).
//  2. a stable LSD radix sort orders the points by path; every node is then
//     a contiguous range and its children are sub-ranges in child order.
//  3. nodes are created level by level: a node with more than one point
//     gets 8 children (empty ones included), ranges found by binary search.
//  4. moments (centroid, circulation) bottom-up; sizes and depth-first
//     positions give a flat array of the non-empty nodes in the reference's
//     recursion order, each with the index just past its subtree ("skip").
//
// Walks (tree_walk :250-316, closest_tree :206-244): one wavefront takes 64
// consecutive points (neighbours along a filament, so their traversals
// overlap) and follows ONE cursor through the depth-first array.  A lane is
// active at the cursor unless it already accepted an ancestor; the cursor
// descends when any active lane opens the node and skips the subtree
// otherwise.  Every lane therefore accepts exactly the nodes its own
// recursive walk would, in the same order, while all node reads are
// wave-uniform (scalar loads).
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <vector>

#include "nv_tree.h"
#include "nv_util.h"

namespace {

// A point's path is kept as up to KEY_WORDS 64-bit words of KEY_LEVELS
// 4-bit digits.  A build uses one word (16 levels) until a node at its
// deepest level still holds more than one point; the tree is then built
// again with one more word (the reference recurses without a bound,
// src/tree.f90:64-104).
constexpr int KEY_LEVELS = 16;
constexpr int KEY_WORDS = 3;
constexpr int LEVELS = KEY_LEVELS * KEY_WORDS;   // NV_TREE_MAX_LEVELS
static_assert(LEVELS == NV_TREE_MAX_LEVELS, "the header states the depth bound");

struct __align__(16) WNode {
    double cent[3];
    double circ[3];
    double width, geo;
    int pcount, leaf_infront, skip;
    int leaf_j;   // a leaf's behind(leaf_infront) when the walk runs (k_leaf_links)
};

// Walk record of the FP32 walk: FP64 centroid, the circulation in FP32 with
// its squared length (the FP32 term), the opening radius squared (FP32,
// rounded up), and one word: the skip index of an internal node, or minus
// one minus a leaf's j = behind(leaf_infront), resolved when the record is
// packed, right before the walk.  The FP64 circulation, which only the FP64
// cases read, is in a side array (three doubles per record), so that the
// records a wave keeps in flight fit in scalar registers.
struct __align__(8) Node48 {
    double cent[3];
    float r2;
    int word;
    float e[3];
    float ee;
};
// The part of a record the walk fetches ahead for both possible successors,
// and the part the FP32 term reads for the current node.
struct __align__(8) NodeHot {
    double cent[3];
    float r2;
    int word;
};
struct __align__(16) NodeCold {
    float e[3];
    float ee;
};
static_assert(sizeof(Node48) == 48 && sizeof(NodeHot) == 32 && offsetof(Node48, e) == sizeof(NodeHot),
              "the hot part leads the record");

// c ? a : b for a wave-uniform c, selected as 64-bit words: every field
// stays in scalar registers (selected field by field, a record's FP32 field
// is moved to a vector register and selected there)
template <typename R>
__device__ __forceinline__ R pick64(bool c, const R &a, const R &b) {
    static_assert(sizeof(R) % 8 == 0, "whole 64-bit words");
    constexpr int NW = sizeof(R) / 8;
    unsigned long long ua[NW], ub[NW], uo[NW];
    __builtin_memcpy(ua, &a, sizeof(R));
    __builtin_memcpy(ub, &b, sizeof(R));
#pragma unroll
    for (int k = 0; k < NW; ++k) uo[k] = c ? ua[k] : ub[k];
    R o;
    __builtin_memcpy(&o, uo, sizeof(R));
    return o;
}

// A row or node count the kernels take either from the host (dn null) or
// from the device, bounded by the launch (a recorded step's counts change
// between replays; its launches cover the bounds).
__device__ __forceinline__ int dev_count(int nmax, const int *dn) { return dn ? min(*dn, nmax) : nmax; }

// The sorted paths: word 0 (levels 1-16) and, in deep builds, words 1 and 2
// (levels 17-32, 33-48), each in the sorted order.
struct Keys {
    const unsigned long long *w0, *w1, *w2;
};

// the digit of level `level` (1-based) of the path at sorted position pos
__device__ __forceinline__ int key_digit(const Keys &k, int pos, int level) {
    const int lv = level - 1, word = lv / KEY_LEVELS, dg = lv - word * KEY_LEVELS;
    const unsigned long long v = word == 0 ? k.w0[pos] : (word == 1 ? k.w1[pos] : k.w2[pos]);
    return (int)((v >> (4 * (KEY_LEVELS - 1 - dg))) & 0xFull);
}

// The paths of every row through nwords * 16 levels: word 0 into keys,
// words 1.. into deep[0..] (by row); inactive rows get all ones, which
// sorts them last.
__global__ void k_keys(DState s, int nmax, const int *dn, double bsize, int nwords, unsigned long long *keys,
                       unsigned long long *deep0, unsigned long long *deep1, int *vals) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dev_count(nmax, dn)) return;
    vals[r] = r;
    // empty points and points pinned in front are not part of the tree
    // (src/tree.f90:126-128)
    if (s.infront[r] == 0 || (s.pin[r] & NV_PINNED_I)) {
        keys[r] = ~0ull;
        if (nwords > 1) deep0[r] = ~0ull;
        if (nwords > 2) deep1[r] = ~0ull;
        return;
    }
    const double x[3] = {s.x[r], s.x[s.cap + r], s.x[2 * s.cap + r]};
    double pos[3] = {-bsize / 2.0, -bsize / 2.0, -bsize / 2.0};
    double w = bsize;
    unsigned long long key = 0;
    int word = 0;
    bool out = false;
    for (int lev = 1; lev <= KEY_LEVELS * nwords; ++lev) {
        const int dg = lev - 1 - word * KEY_LEVELS;   // digit within the word
        int dig = 0;
        bool in = true;
        double lo[3];
        const double cw = w / 2.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double split = pos[k] + (w / 2);
            const int b = x[k] >= split ? 1 : 0;
            lo[k] = b ? split : pos[k];
            in = in && (x[k] >= lo[k]) && (x[k] < (lo[k] + cw));
            dig |= b << (k == 0 ? 0 : (k == 1 ? 2 : 1));
        }
        if (!in) {
            key |= 8ull << (4 * (KEY_LEVELS - 1 - dg));
            out = true;
        } else {
            key |= (unsigned long long)dig << (4 * (KEY_LEVELS - 1 - dg));
            pos[0] = lo[0]; pos[1] = lo[1]; pos[2] = lo[2];
            w = cw;
        }
        if (out || dg == KEY_LEVELS - 1) {
            // the word is complete (the rest of the path is zeros)
            if (word == 0) keys[r] = key;
            else if (word == 1) deep0[r] = key;
            else deep1[r] = key;
            key = 0;
            ++word;
            if (out) break;
        }
    }
    for (; word < nwords; ++word) {
        if (word == 0) keys[r] = 0ull;
        else if (word == 1) deep0[r] = 0ull;
        else deep1[r] = 0ull;
    }
}

// out[i] = in[perm[i]]
__global__ void k_gather_keys(const unsigned long long *__restrict__ in, const int *__restrict__ perm, int nmax,
                              const int *dn, unsigned long long *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < dev_count(nmax, dn)) out[i] = in[perm[i]];
}

// The number of active points, from the sorted keys: an active key's top
// nibble is at most 8, so active keys sort before the inactive ~0 ones. One
// thread finds the boundary and writes it; *nact was cleared beforehand.
__global__ void k_count_active(const unsigned long long *__restrict__ keys, int nmax, const int *dn, int *nact) {
    const int n = dev_count(nmax, dn);
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    if (keys[r] != ~0ull && (r + 1 == n || keys[r + 1] == ~0ull)) *nact = r + 1;
}

// ------------------------------------------------------------ nodes
struct Nodes {
    int *beg, *cnt, *pcount, *dropb, *first, *parent, *minrow, *size, *dfs, *leafinf;
    double *pos, *w, *sum, *circ;   // pos/sum/circ: [3][ncap]
    double *rad;                     // bound on |centroid of a descendant or point - centroid|
    double *cent;                    // [3][ncap] centroid: sum / count (the root: / active points)
    int ncap;
};

// Level bookkeeping of a build whose level loop runs without host round
// trips: each level's kernels read their node range here, and the launch
// bounds come from the previous build's level sizes.
struct LevelCtl {
    int lvl[LEVELS + 3];    // level starts: level l holds nodes [lvl[l], lvl[l + 1])
    int wide[LEVELS + 1];   // level's children made by k_children_wide
    int overflow;           // a level outgrew its launch bound or the node arrays
    int ovf_level;          // the first such level: the levels above it are complete
    int depth;              // a node of the deepest level holds more than one point
    int nact;
    int nwalk;              // non-empty nodes, the depth-first walk's length
};

__global__ void k_root(Nodes nd, DState s, int nmax, const int *dn, double bsize, const int *nact, LevelCtl *lc) {
    const int n = dev_count(nmax, dn);
    if (lc) {
        lc->lvl[0] = 0;
        lc->lvl[1] = 1;
        lc->overflow = 0;
        lc->ovf_level = LEVELS + 1;
        lc->depth = 0;
        lc->nact = *nact;
    }
    nd.beg[0] = 0;
    nd.cnt[0] = *nact;
    nd.pcount[0] = n;   // the root holds every slot (src/tree.f90:48)
    nd.parent[0] = -1;
    nd.first[0] = -1;
    for (int k = 0; k < 3; ++k) nd.pos[k * nd.ncap] = -bsize / 2.0;
    nd.w[0] = bsize;
    nd.dropb[0] = *nact;
}

__global__ void k_split_flags(Nodes nd, int L0, int L1, int *flags, int at_max, DCtl *ctl) {
    int q = L0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= L1) return;
    const int sp = nd.pcount[q] > 1 ? 1 : 0;
    flags[q - L0] = sp;
    if (sp && at_max) atomicOr(&ctl->status, NVK_TREE_DEPTH);
}

// k_split_flags for a level whose range is on the device: flags[0, bound),
// zero past the level's end so that a scan over the bound is the level's
__global__ void k_split_flags_d(Nodes nd, const LevelCtl *lc, int level, int bound, int *flags, DCtl *ctl) {
    // (only the deepest level any build has, LEVELS, reports the depth as an
    // error on the device; shallower builds are redone deeper by the host)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bound) return;
    const int L0 = lc->lvl[level], m = lc->lvl[level + 1] - L0;
    int sp = 0;
    if (i < m) {
        sp = nd.pcount[L0 + i] > 1 ? 1 : 0;
        if (sp && level == LEVELS) atomicOr(&ctl->status, NVK_TREE_DEPTH);
    }
    flags[i] = sp;
}

// after a level's scan: the next level's start, the children kernel, and
// whether the level fit its bounds (one thread)
__global__ void k_level_end(Nodes nd, LevelCtl *lc, int level, int maxlev, int bound, int wgrid, const int *offs,
                            const int *tsum) {
    const int L0 = lc->lvl[level], L1 = lc->lvl[level + 1], m = L1 - L0;
    int next = L1, wide = 0;
    if (m > bound && !lc->overflow) {
        lc->overflow = 1;
        lc->ovf_level = level;
    } else if (m > 0 && !lc->overflow) {
        const int nsplit = offs[m - 1] + tsum[(m - 1) / NV_SCAN_TILE] + (nd.pcount[L1 - 1] > 1 ? 1 : 0);
        if (nsplit && level == maxlev) {
            lc->depth = 1;
        } else if (nsplit) {
            if ((long long)L1 + 8ll * nsplit > (long long)nd.ncap) {
                lc->overflow = 1;
                lc->ovf_level = level;
            } else {
                next = L1 + 8 * nsplit;
                wide = (m <= wgrid && m <= 4096 && nsplit * 64 <= lc->nact) ? 1 : 0;
            }
        }
    }
    lc->wide[level] = wide;
    lc->lvl[level + 2] = next;
}

__device__ __forceinline__ int lower_digit(const Keys &keys, int lo, int hi, int level, int d) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (key_digit(keys, mid, level) < d) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// child c of split node q, its keys [lo, hi)
__device__ __forceinline__ void write_child(const Nodes &nd, int q, int base, int c, int lo, int hi) {
    const double w = nd.w[q];
    const double p0 = nd.pos[q], p1 = nd.pos[nd.ncap + q], p2 = nd.pos[2 * nd.ncap + q];
    const int ch = base + c;
    const int xb = c & 1, zb = (c >> 1) & 1, yb = (c >> 2) & 1;
    nd.pos[ch] = xb ? p0 + (w / 2) : p0;
    nd.pos[nd.ncap + ch] = yb ? p1 + (w / 2) : p1;
    nd.pos[2 * nd.ncap + ch] = zb ? p2 + (w / 2) : p2;
    nd.w[ch] = w / 2.0;
    nd.beg[ch] = lo;
    nd.cnt[ch] = hi - lo;
    nd.pcount[ch] = hi - lo;
    nd.parent[ch] = q;
    nd.first[ch] = -1;
    nd.dropb[ch] = hi;
}

// Work on the split nodes among a wave's nodes (one per lane) with eight
// lanes per node, so that reads and writes of the eight children (adjacent
// nodes) are one child per lane, consecutive lanes on consecutive
// children, instead of eight strided accesses per lane.  The split lanes'
// nodes are listed in lane order and taken eight per round: group
// g = lane / 8 gets the round's g-th node, lane c = lane % 8 its child c.
// fn(q, c) runs on all eight lanes of a group that has a node, so it may
// exchange values within the group.  Every lane of the wave must call this.
constexpr int SPLIT8_WAVES = 4;   // waves per workgroup of the kernels that use it

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename F>
__device__ __forceinline__ void for_split8(int q, bool split, F fn) {
    __shared__ int lst[SPLIT8_WAVES][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long m = wave_ballot(split);
    if (m == 0ull) return;
    const int k = __popcll(m);
    if (split) lst[wv][__popcll(m & ((1ull << lane) - 1ull))] = q;
    wave_lds_sync();
    for (int r = 0; r < k; r += 8) {
        const int i = r + (lane >> 3);
        if (i < k) fn(lst[wv][i], lane & 7);
    }
    wave_lds_sync();
}

// children of the split nodes (pcount > 1) of level `level` (nodes [L0,
// L1)) into L1 + 8 * off, one thread per node and eight lanes per split node
// off[]: the split flags' exclusive scan, or its tile scan with the tile
// prefixes in tsum[] (the two-level scan)
__device__ __forceinline__ int scanned(const int *off, const int *tsum, int i) {
    return off[i] + (tsum ? tsum[i / NV_SCAN_TILE] : 0);
}

__device__ __forceinline__ void children_of(const Nodes &nd, int q, int L0, int L1, const int *off,
                                            const int *tsum, const Keys &keys, int level) {
    for_split8(q, q < L1 && nd.pcount[q] > 1, [&](int p, int c) {
        // lane c: the first key of p's range with digit >= c + 1 at the next
        // level (the digits ascend over the range), the end of child c
        const int base = L1 + 8 * scanned(off, tsum, p - L0);
        const int b0 = nd.beg[p], b1 = nd.beg[p] + nd.cnt[p];
        const int hi = lower_digit(keys, b0, b1, level + 1, c + 1);
        const int lo = __shfl_up(hi, 1, 8);
        if (c == 7) {
            nd.first[p] = base;
            nd.dropb[p] = hi;
        }
        write_child(nd, p, base, c, c == 0 ? b0 : lo, hi);
    });
}

__global__ __launch_bounds__(256) void k_children(Nodes nd, int L0, int L1, const int *off, Keys keys, int level) {
    children_of(nd, L0 + blockIdx.x * blockDim.x + threadIdx.x, L0, L1, off, nullptr, keys, level);
}

__global__ __launch_bounds__(256) void k_children_d(Nodes nd, const LevelCtl *lc, const int *off, const int *tsum,
                                                    Keys keys, int level) {
    if (lc->overflow || lc->wide[level]) return;
    const int L0 = lc->lvl[level], L1 = lc->lvl[level + 1];
    children_of(nd, L0 + blockIdx.x * blockDim.x + threadIdx.x, L0, L1, off, tsum, keys, level);
}

// The same for a level of few nodes with long key ranges, one workgroup of
// 256 per node: 32 lanes find each of the 8 child boundaries by a 32-way
// search (a node's keys are ascending in the next level's digit).
__device__ __forceinline__ void children_wide_of(const Nodes &nd, int q, int L0, int L1, const int *off,
                                                 const int *tsum, const Keys &keys, int level) {
    __shared__ int bnd[9];
    if (q >= L1 || nd.pcount[q] <= 1) return;
    const int b0 = nd.beg[q], b1 = nd.beg[q] + nd.cnt[q];
    const int d = 1 + (threadIdx.x >> 5), li = threadIdx.x & 31, sh = threadIdx.x & 32;
    // first position in [b0, b1) with digit >= d
    int lo = b0, hi = b1;
    while (hi > lo) {
        const int step = (hi - lo + 31) / 32;
        const int pos = lo + li * step;
        const bool below = pos < hi && key_digit(keys, pos, level + 1) < d;
        const int c = __popcll((wave_ballot(below) >> sh) & 0xffffffffull);
        if (c == 0) { hi = lo; break; }
        const int nlo = lo + (c - 1) * step + 1, nhi = min(hi, lo + c * step);
        lo = nlo;
        hi = nhi;
    }
    if (li == 0) bnd[d] = lo;
    if (threadIdx.x == 0) bnd[0] = b0;
    __syncthreads();
    const int base = L1 + 8 * scanned(off, tsum, q - L0);
    if (threadIdx.x < 8) write_child(nd, q, base, threadIdx.x, bnd[threadIdx.x], bnd[threadIdx.x + 1]);
    if (threadIdx.x == 0) {
        nd.first[q] = base;
        nd.dropb[q] = bnd[8];
    }
}

__global__ __launch_bounds__(256) void k_children_wide(Nodes nd, int L0, int L1, const int *off, Keys keys,
                                                       int level) {
    children_wide_of(nd, L0 + blockIdx.x, L0, L1, off, nullptr, keys, level);
}

__global__ __launch_bounds__(256) void k_children_wide_d(Nodes nd, const LevelCtl *lc, const int *off,
                                                         const int *tsum, Keys keys, int level) {
    if (lc->overflow || !lc->wide[level]) return;
    const int L0 = lc->lvl[level], L1 = lc->lvl[level + 1];
    children_wide_of(nd, L0 + blockIdx.x, L0, L1, off, tsum, keys, level);
}

// moments, minimum row and non-empty subtree size, children before parents
__device__ __forceinline__ double dist3(double a0, double a1, double a2, double b0, double b1, double b2) {
    const double d0 = a0 - b0, d1 = a1 - b1, d2 = a2 - b2;
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// psum (solid walls, else null): the positions and count of the points
// pinned in front, which the root's centroid counts (its masked sum over
// infront > 0, src/tree.f90:50-52) and no other node does
__device__ __forceinline__ void up_of(const Nodes &nd, const DState &s, int q, const int *vals, int nact,
                                      const double *psum) {
    const int cap = s.cap, nc = nd.ncap;
    double sm[3] = {0.0, 0.0, 0.0}, ci[3] = {0.0, 0.0, 0.0};
    int mr = 0x7fffffff, size;
    if (nd.first[q] < 0) {
        // a leaf: the reference's new_tree sums over its points in index
        // order; with at most one point that is exact: cent = x, circ = (0 + gi) - x
        if (nd.cnt[q] == 1 && nd.pcount[q] == 1) {
            const int p = vals[nd.beg[q]];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double xk = s.x[k * cap + p];
                sm[k] = 0.0 + xk;
                ci[k] = (0.0 + s.gi[k * cap + p]) - xk;
            }
            mr = p;
        } else {
            for (int j = nd.beg[q]; j < nd.beg[q] + nd.cnt[q]; ++j) {
                const int p = vals[j];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double xk = s.x[k * cap + p];
                    sm[k] = sm[k] + xk;
                    ci[k] = ci[k] + s.gi[k * cap + p] - xk;
                }
                mr = min(mr, p);
            }
        }
        size = nd.pcount[q] > 0 ? 1 : 0;
    } else {
        size = 1;
        for (int c = 0; c < 8; ++c) {
            const int ch = nd.first[q] + c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                sm[k] = sm[k] + nd.sum[k * nc + ch];
                ci[k] = ci[k] + nd.circ[k * nc + ch];
            }
            mr = min(mr, nd.minrow[ch]);
            size += nd.size[ch];
        }
        for (int j = nd.dropb[q]; j < nd.beg[q] + nd.cnt[q]; ++j) {
            const int p = vals[j];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double xk = s.x[k * cap + p];
                sm[k] = sm[k] + xk;
                ci[k] = ci[k] + s.gi[k * cap + p] - xk;
            }
            mr = min(mr, p);
        }
    }
    // The centroid k_down gives this node, and a bound on the distance from
    // it to every centroid and point below it (the triangle inequality over
    // the children, plus the points kept at this node), inflated by 2^-40
    // to cover the rounding of the norms: the close search's exact pruning.
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        nd.sum[k * nc + q] = sm[k];
        nd.circ[k * nc + q] = ci[k];
    }
    nd.minrow[q] = mr;
    nd.size[q] = size;
    // empty nodes: no centroid (nothing reads it)
    if (nd.pcount[q] == 0) return;
    double cntd = q == 0 ? (double)nact : (double)nd.pcount[q];
    double cs[3] = {sm[0], sm[1], sm[2]};
    if (q == 0 && psum) {
#pragma unroll
        for (int k = 0; k < 3; ++k) cs[k] = cs[k] + psum[k];
        cntd = cntd + psum[3];
    }
    const double c0 = cs[0] / cntd, c1 = cs[1] / cntd, c2 = cs[2] / cntd;
    double rad = 0.0;
    int p0 = nd.beg[q];
    if (nd.first[q] >= 0) {
        for (int c = 0; c < 8; ++c) {
            const int ch = nd.first[q] + c;
            if (nd.pcount[ch] > 0)
                rad = fmax(rad, dist3(nd.cent[ch], nd.cent[nc + ch], nd.cent[2 * nc + ch], c0, c1, c2) + nd.rad[ch]);
        }
        p0 = nd.dropb[q];
    }
    // a one-point leaf is its point (c = (0 + x) / 1 = x): radius 0
    if (!(nd.first[q] < 0 && nd.cnt[q] == 1 && nd.pcount[q] == 1))
        for (int j = p0; j < nd.beg[q] + nd.cnt[q]; ++j) {
            const int pt = vals[j];
            rad = fmax(rad, dist3(s.x[pt], s.x[cap + pt], s.x[2 * cap + pt], c0, c1, c2));
        }
    nd.rad[q] = rad * (1.0 + 0x1p-40);
    nd.cent[q] = c0;
    nd.cent[nc + q] = c1;
    nd.cent[2 * nc + q] = c2;
}

__global__ void k_up(Nodes nd, DState s, int L0, int L1, const int *vals, int nact, const double *psum) {
    const int q = L0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (q < L1) up_of(nd, s, q, vals, nact, psum);
}

__global__ void k_up_d(Nodes nd, DState s, const LevelCtl *lc, int level, const int *vals, const double *psum) {
    if (lc->overflow) return;
    const int q = lc->lvl[level] + blockIdx.x * blockDim.x + threadIdx.x;
    if (q < lc->lvl[level + 1]) up_of(nd, s, q, vals, lc->nact, psum);
}

// the sum of the positions of the points pinned in front and their count
// (rows [0, n), or [0, *dn)), one workgroup, a fixed reduction order
__global__ __launch_bounds__(256) void k_pinned_sum(DState s, int nmax, const int *dn, double *psum) {
    const int n = dev_count(nmax, dn);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int r = threadIdx.x; r < n; r += 256)
        if (s.infront[r] != 0 && (s.pin[r] & NV_PINNED_I)) {
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] += s.x[k * s.cap + r];
            v[3] += 1.0;
        }
    __shared__ double red[4][256];
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) psum[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_build_tail(const Nodes nd, LevelCtl *lc) { lc->nwalk = lc->overflow ? 0 : nd.size[0]; }

// A recorded step's build is complete when no level outgrew its bound or the
// node arrays, no node of the deepest level holds several points, the tree
// ends within the levels launched (levels 0 .. top) and its walk fits the
// records (nwalk_max).  Otherwise the step's later phases are switched off
// (ctl->abort, no walk records) and the host takes the step again eagerly.
__global__ void k_build_check(LevelCtl *lc, DCtl *ctl, int top, int nwalk_max, const int *dn) {
    // (no points at all: the eager build's error)
    if (lc->overflow || lc->depth || lc->lvl[top + 2] > lc->lvl[top + 1] || lc->nwalk > nwalk_max || *dn <= 0) {
        ctl->abort = 1;
        lc->nwalk = 0;
    }
}

// depth-first positions of the children of node q and its walk record
__device__ __forceinline__ void down_of(const Nodes &nd, const DState &s, int q, WNode *walk, float *wrad) {
    const int nc = nd.ncap;
    const int pc = nd.pcount[q];
    if (q == 0) nd.dfs[0] = 0;
    // leaf_infront: parray(1)%infront at build time; the root's parray(1)
    // is f(1) whatever it holds (src/tree.f90:49)
    int li = 0;
    if (q == 0) li = s.infront[0];
    else if (pc >= 1 && nd.minrow[q] != 0x7fffffff) li = s.infront[nd.minrow[q]];
    nd.leafinf[q] = li;
    if (pc == 0) return;
    const int my = nd.dfs[q];
    if (nd.first[q] >= 0) {
        int acc = my + 1;
        for (int c = 0; c < 8; ++c) {
            const int ch = nd.first[q] + c;
            nd.dfs[ch] = acc;
            acc += nd.size[ch];
        }
    }
    WNode w;
    double cent[3];   // sum / count, from k_up
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cent[k] = nd.cent[k * nc + q];
        w.cent[k] = cent[k];
        w.circ[k] = (q == 0) ? 0.0 : nd.circ[k * nc + q];
    }
    w.width = nd.w[q];
    const double g0 = cent[0] - (nd.pos[q] + nd.w[q] / 2.0);
    const double g1 = cent[1] - (nd.pos[nc + q] + nd.w[q] / 2.0);
    const double g2 = cent[2] - (nd.pos[2 * nc + q] + nd.w[q] / 2.0);
    w.geo = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    w.pcount = pc;
    w.leaf_infront = li;
    w.skip = my + nd.size[q];
    w.leaf_j = 0;
    walk[my] = w;
    // the node radius as a float rounded up
    const double rd = nd.rad[q];
    float rf = (float)rd;
    if ((double)rf < rd) rf = nextafterf(rf, INFINITY);
    wrad[my] = rf;
}

// down_of over the nodes of level-`L`
__global__ void k_down(Nodes nd, DState s, int L0, int L1, WNode *walk, float *wrad) {
    const int q = L0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (q < L1) down_of(nd, s, q, walk, wrad);
}

// k_down on a level whose range is on the device; nothing once a recorded
// step's build has been found incomplete (k_build_check)
__global__ void k_down_d(Nodes nd, DState s, const LevelCtl *lc, int level, WNode *walk, float *wrad,
                         const DCtl *ctl) {
    if (ctl->abort) return;
    const int q = lc->lvl[level] + blockIdx.x * blockDim.x + threadIdx.x;
    if (q < lc->lvl[level + 1]) down_of(nd, s, q, walk, wrad);
}

// j = behind(leaf_infront) of every leaf record, read once before a walk
// (links do not change while it runs), so that the walk loop has no load
// that depends on the node it evaluates
__global__ void k_leaf_links(DState s, WNode *__restrict__ walk, int nwmax, const int *dnw) {
    const int nwalk = dev_count(nwmax, dnw);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nwalk) return;
    const int pc = walk[q].pcount, li = walk[q].leaf_infront;
    if (pc == 1) walk[q].leaf_j = li > 0 ? s.behind[li - 1] : 0;
}

__global__ void k_zero_counters(unsigned long long *c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------ walks
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Work queues of the persistent walks, one per XCD.  The nb blocks of 64
// targets (in Morton order) are split into XCDS contiguous ranges; queue k
// hands out (image, block) items of range k, the images in order (the
// unshifted one, the longest, first).  A workgroup takes items from queue
// blockIdx.x % XCDS (workgroups are dispatched round-robin over the XCDs, so
// that is its own XCD, whose L2 then holds one region's nodes), then from the
// others in turn, so every queue drains whatever the mapping.
constexpr int XCDS = 8;

struct WalkQueues {
    int *cnt;   // XCDS counters, zero at launch
    int nb;     // blocks of targets (of `per` targets each)
    int nimg;   // images per block
    const int *dn = nullptr;   // the target count on the device (then nb is derived from it) ...
    int nmax = 0;              // ... bounded by this
    int per = 64;              // targets per block
};

// the queues with their block count taken from the device when it is there
__device__ __forceinline__ WalkQueues resolve(WalkQueues wq) {
    if (wq.dn) wq.nb = (dev_count(wq.nmax, wq.dn) + wq.per - 1) / wq.per;
    return wq;
}

// The next item of the calling wave (every lane active): false once all
// queues are empty.  h: the queues this wave has emptied.
__device__ __forceinline__ bool next_walk_item(const WalkQueues &wq, int &h, int &m, int &tb) {
    const int home = blockIdx.x % XCDS;
    while (h < XCDS) {
        const int k = (home + h) % XCDS;
        const int b0 = (int)((long)wq.nb * k / XCDS), len = (int)((long)wq.nb * (k + 1) / XCDS) - b0;
        int it = 0;
        if ((threadIdx.x & 63) == 0) it = atomicAdd(&wq.cnt[k], 1);
        it = wave_uniform(it);   // lane 0 is the first active lane
        if (it < len * wq.nimg) {
            m = it / len;
            tb = b0 + (it - m * len);
            return true;
        }
        ++h;
    }
    return false;
}

constexpr double OPEN_BAND = 4.8e-7;
constexpr int OPEN_BAND_ULPS = 16;

// WNode -> Node48.  The reference accepts a node when pcount == 1 or
// width / (dist - geo) < theta (width / dist when dist <= geo or without the
// correction, src/tree.f90:266-278); for theta <= 1.1 (so that w / theta >
// geo) that is dist > geo + width / theta (width / theta without the
// correction), compared here on squares.
__global__ void k_pack48(DState s, const WNode *__restrict__ walk, int nwmax, const int *dnw, double theta, int corr,
                         Node48 *__restrict__ out, double *__restrict__ circ64) {
    const int nwalk = dev_count(nwmax, dnw);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > nwalk) return;
    if (q == nwalk) {   // the spare record
        out[q] = Node48{};
        return;
    }
    const WNode w = walk[q];
    Node48 o;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o.cent[k] = w.cent[k];
        circ64[3 * q + k] = w.circ[k];
    }
    const double r = (corr ? w.geo : 0.0) + w.width / theta;
    o.r2 = __double2float_ru(r * r * (1.0 + OPEN_BAND));   // the upper bound of the exact band
    o.word = w.pcount == 1 ? -(w.leaf_infront > 0 ? s.behind[w.leaf_infront - 1] : 0) - 1 : w.skip;
#pragma unroll
    for (int k = 0; k < 3; ++k) o.e[k] = (float)w.circ[k];
    o.ee = __fmaf_rn(o.e[0], o.e[0], __fmaf_rn(o.e[1], o.e[1], o.e[2] * o.e[2]));
    out[q] = o;
}

// The reference's node term (src/tree.f90:297-307) in FP64, a = dist**2.
__device__ __forceinline__ void tree_term64(double v0, double v1, double v2, const double c[3], double dd,
                                            double kappa, double nr[3], unsigned long long &nev) {
    const double dist = sqrt(dd);
    const double a = dist * dist;
    const double b = 2.0 * (v0 * c[0] + v1 * c[1] + v2 * c[2]);
    const double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const double D = 4.0 * a * cc - b * b;
    if (!(D < NV_EPS)) {
        const double den = (2.0 * NV_PI) * D;
        const double fac = (2.0 * cc + b) / sqrt(a + b + cc) - (b / sqrt(a));
        nr[0] = nr[0] + ((v1 * c[2] - v2 * c[1]) * kappa / den) * fac;
        nr[1] = nr[1] + ((v2 * c[0] - v0 * c[2]) * kappa / den) * fac;
        nr[2] = nr[2] + ((v0 * c[1] - v1 * c[0]) * kappa / den) * fac;
        ++nev;
    }
}

// The node term for the FP32 walk's FP64 cases (near leaves, ill-conditioned
// far nodes): the reference's expression (src/tree.f90:297-307) with
// a = dist**2 taken as dd, both inverse distances from rsqrt and one
// division, accurate to a few FP64 ulps (the FP32 walk is not bit-exact).
__device__ __forceinline__ void near_term64(double v0, double v1, double v2, const double c[3], double dd,
                                            double kappa, double nr[3], unsigned long long &nev) {
    const double b = 2.0 * (v0 * c[0] + v1 * c[1] + v2 * c[2]);
    const double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const double D = 4.0 * dd * cc - b * b;
    if (!(D < NV_EPS)) {
        const double fac = (2.0 * cc + b) * rsqrt(dd + b + cc) - b * rsqrt(dd);
        const double sc = kappa * fac / ((2.0 * NV_PI) * D);
        nr[0] = nr[0] + (v1 * c[2] - v2 * c[1]) * sc;
        nr[1] = nr[1] + (v2 * c[0] - v0 * c[2]) * sc;
        nr[2] = nr[2] + (v0 * c[1] - v1 * c[0]) * sc;
        ++nev;
    }
}

// The walk of k_walk_vel with FP32 interactions: displacements in FP64, the
// node term in FP32 in the well-conditioned form (as the direct sum's far
// field), the terms accumulated in FP64 (on this hardware an FP64 add issues
// at the unpacked FP32 rate: two instructions per term and component, half
// of a Kahan update).  The reference's FP64 term is kept for
// leaves closer than r_near and for nodes whose pseudo-segment (centroid to
// centroid + circulation) passes close by the target, where the FP32 form's
// r1 r2 + R1.R2 cancels.  Same traversal, exclusions and singularity test.
//
// Persistent: a grid the size of the resident waves, each wave taking items
// (image, block of 64 targets) from a queue, the unshifted image's (longest)
// walks first, so that no SIMD idles while long walks finish.
// targets per lane of the FP32 walk
constexpr int WALK_TPT = 2;
// The FP32 walk addresses its records (48-byte walk records, 80-byte full
// records, 24-byte FP64 circulations) by 32-bit byte offsets: at most this
// many records (about 25M points).
constexpr long long WALK32_MAX_RECORDS = 0xFFFFFFFFll / (long long)sizeof(WNode);

// The band around a node's opening radius squared r^2 inside which the walk
// decides the reference's opening test in FP64 exactly.  The record holds
// its upper bound, r^2 (1 + 4.8e-7) rounded up to FP32; its lower bound is
// 16 float ulps below that, at least 9.5e-7 r^2 lower.  The walk compares its
// squared distance rounded to FP32 (2^-24 = 6e-8 relative; it differs from
// the reference's FP64 distance squared by under 1e-11 relative) with both.
// (OPEN_BAND, OPEN_BAND_ULPS: defined before k_pack48)

// The reference's opening test for target row `row`, node centroid `cent`,
// shift `sm` (src/tree.f90:263-278), with the FP64 walk's arithmetic: true
// when the node is accepted as a whole.
__device__ __forceinline__ bool ref_accepts(const DState &s, int row, const double cent[3], const double sm[3],
                                            double width, double geo, double theta_c, int corr) {
    const double e0 = s.x[row] - (cent[0] + sm[0]);
    const double e1 = s.x[s.cap + row] - (cent[1] + sm[1]);
    const double e2 = s.x[2 * s.cap + row] - (cent[2] + sm[2]);
    const double dist = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    double theta;
    if (corr) {
        if (dist - geo > 0.0) theta = width / (dist - geo);
        else theta = width / dist;
    } else {
        theta = width / dist;
    }
    return theta < theta_c;
}

// Five waves per SIMD (96 VGPRs): the exact opening test of the band is
// rare (about a hundred lanes per walk of 1M points), and the few registers
// the compiler spills to fit are spilled there.  That test, although rarely
// taken, costs the loop about 10% (more scalar registers spilled to vector
// lanes); decided after the walk instead, by walking the items it touched
// again, it cost more, since one item's walk is a long chain of dependent
// record fetches.
template <int BT, int TPT>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(5))) void k_walk_vel32(DState s, Phys p, Shifts sh, int n, int npad,
                                                   const int *__restrict__ order, int nord, WalkQueues wq,
                                                   const Node48 *__restrict__ walk,
                                                   const double *__restrict__ circ64, const WNode *__restrict__ wn,
                                                   int nwalk, double rn2, double sing2, double *__restrict__ part,
                                                   DCtl *ctl, unsigned long long *counters, const int *dn,
                                                   const int *dnw) {
    if (ctl->abort) return;
    if (dn) nord = n = dev_count(n, dn);   // a recorded step: this step's count (its build's rows)
    if (dnw) nwalk = *dnw;
    wq = resolve(wq);
    const int lane = threadIdx.x & 63;
    int h = 0;   // queues emptied
    const double kappa = p.kappa;
    unsigned long long nev = 0, nvis = 0;   // nev: FP64 terms, per lane
    unsigned long long nevw = 0;            // FP32 terms, per wave
    // visits and FP32 terms of this lane: counted per lane on the vector
    // unit (a popcount of the wave's mask per block and iteration kept the
    // scalar unit, which the walk's mask algebra already loads, 4% longer)
    unsigned visl = 0, evl = 0;
    bool sing = false;
    for (;;) {
    int m, tb;
    if (!next_walk_item(wq, h, m, tb)) break;
    // TPT targets per lane: those of TPT consecutive blocks of 64 (in the
    // Morton order), which share the wave's traversal and record reads
    int r[TPT], beh[TPT], me[TPT], lskip[TPT];
    double xs[TPT][3];
    double fa[TPT][3];   // FP32 terms, accumulated in FP64
    double nr[TPT][3];   // FP64 terms
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int q = (tb * TPT + t) * 64 + lane;
        r[t] = q < nord ? order[q] : q;
        const bool act = r[t] < n && s.infront[r[t]] != 0;
        double xi[3] = {0, 0, 0};
        beh[t] = -5;
        me[t] = -5;
        if (act) {
            xi[0] = s.x[r[t]]; xi[1] = s.x[s.cap + r[t]]; xi[2] = s.x[2 * s.cap + r[t]];
            beh[t] = s.behind[r[t]];
            me[t] = r[t] + 1;
        }
        // the shifted target: displacements are centroid - xs (in the
        // unshifted image exactly the reference's cent - x)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xs[t][k] = xi[k] - sh.s[m][k];
            fa[t][k] = 0.0;
            nr[t][k] = 0.0;
        }
        lskip[t] = act ? 0 : 0x7fffffff;
    }
    const bool img0 = (m == 0);
    const unsigned long long img0m = img0 ? ~0ull : 0ull;
    int cur = 0;
    const char *wb = (const char *)walk;
    NodeHot nd = *(const NodeHot *)wb;
    // Control flow on wave-uniform conditions only (the node is uniform):
    // lane conditions become selects, so that the scalar unit, which the
    // four SIMDs of a CU share, does not spend the walk on exec masks.
    while (cur < nwalk) {
        cur = wave_uniform(cur);
        const int w = nd.word;
        const bool leaf = w < 0;
        // both possible successors are fetched while this node is evaluated
        // (the records end with one spare, so neither index needs a clamp)
        const int n1 = cur + 1, n2 = leaf ? cur + 1 : w;
        // (byte offsets in 32 bits: the scalar loads take them as their
        // offset register, without a 64-bit multiply and add on the scalar
        // unit, which the walk keeps busy; records stay below 2^31 bytes)
        const NodeHot a = *(const NodeHot *)(wb + (unsigned)n1 * (unsigned)sizeof(Node48));
        const NodeHot b = *(const NodeHot *)(wb + (unsigned)n2 * (unsigned)sizeof(Node48));
        // (this file is compiled without machine sinking, which would move
        // these loads down to their use at the end of the iteration)
        // Every lane condition is a 64-bit mask built from ballots of raw
        // compares (each beside its compare), combined on the scalar unit;
        // a per-lane select reads its mask back with inverse_ballot, which
        // costs nothing. Bools combined per lane would be kept twice, as lane
        // masks and as ballots, and a ballot of such a bool is rebuilt in
        // vector registers.
        unsigned long long mfar[TPT], mneed[TPT];
        double v[TPT][3], dd[TPT];
        bool down = false;
        unsigned long long anyfar = 0ull, anyneed = 0ull;
        // accepted surely above r2hi, opened surely below r2lo; between them
        // the reference's own test decides
        // the record holds the band's upper bound; the lower one lies
        // OPEN_BAND_ULPS float ulps below it (an integer step on the scalar unit)
        const float r2hi = nd.r2;
        const float r2lo = __int_as_float(__float_as_int(nd.r2) - OPEN_BAND_ULPS);
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            const unsigned long long mvis = wave_ballot(cur >= lskip[t]);
            mfar[t] = 0ull;
            mneed[t] = 0ull;
            // a block none of whose lanes visits this node (all accepted an
            // ancestor) changes nothing here: skip its work (a wave-uniform
            // branch; about a sixth of the iterations for each block)
            if (TPT > 1 && mvis == 0ull) continue;
            visl += lane_in(mvis) ? 1u : 0u;
#pragma unroll
            for (int k = 0; k < 3; ++k) v[t][k] = nd.cent[k] - xs[t][k];
            dd[t] = __fma_rn(v[t][0], v[t][0], __fma_rn(v[t][1], v[t][1], v[t][2] * v[t][2]));
            const float ddf = (float)dd[t];
            unsigned long long mgt = wave_ballot(ddf > r2hi);
            unsigned long long macc, muse, mnear;
            if (leaf) {
                const int j = -w - 1;
                const unsigned long long mexcl = (wave_ballot(j == me[t]) | wave_ballot(j == beh[t])) & img0m;
                // dd < sing2 <=> sqrt(dd) < epsilon (sqrt is correctly rounded)
                const unsigned long long msg = mvis & ~mexcl & wave_ballot(dd[t] < sing2);
                const unsigned long long mrn = wave_ballot(dd[t] < rn2);
                sing = sing || msg != 0ull;
                macc = mvis;
                muse = mvis & ~mexcl & ~msg;
                mnear = muse & mrn & (j > 0 ? ~0ull : 0ull);
            } else {
                const unsigned long long mband = mvis & ~mgt & ~wave_ballot(ddf < r2lo);
                if (mband != 0ull) {
                    const WNode *wc = (const WNode *)((const char *)wn + (unsigned)cur * (unsigned)sizeof(WNode));
                    const double wdt = wc->width, geo = wc->geo;
                    bool ok = false;
                    if (lane_in(mband)) ok = ref_accepts(s, r[t], nd.cent, sh.s[m], wdt, geo, p.tree_theta, p.extra_corr);
                    mgt |= wave_ballot(ok) & mband;
                }
                macc = mvis & mgt;
                muse = macc;
                mnear = 0ull;
                down = down || (mvis & ~mgt) != 0ull;   // a lane opens the node
            }
            lskip[t] = lane_in(macc) ? n2 : lskip[t];   // a leaf's successor, or the node's skip
            mfar[t] = muse & ~mnear;
            mneed[t] = mnear;
            anyfar |= mfar[t];
        }
        if (anyfar != 0ull) {
            const NodeCold cd = *(const NodeCold *)(wb + (unsigned)cur * (unsigned)sizeof(Node48) + sizeof(NodeHot));
            const float e0 = cd.e[0], e1 = cd.e[1], e2 = cd.e[2];
            const float kee = 64.0f * 5.96e-8f * cd.ee;
#pragma unroll
            for (int t = 0; t < TPT; ++t) {
                if (TPT > 1 && mfar[t] == 0ull) continue;
                const float d0 = (float)v[t][0], d1 = (float)v[t][1], d2 = (float)v[t][2];
                const float x0 = __fmaf_rn(d1, e2, -(d2 * e1));
                const float x1 = __fmaf_rn(d2, e0, -(d0 * e2));
                const float x2 = __fmaf_rn(d0, e1, -(d1 * e0));
                const float r1sq = (float)dd[t];
                const float de = __fmaf_rn(d0, e0, __fmaf_rn(d1, e1, d2 * e2));
                const float r2sq = __fmaf_rn(2.0f, de, r1sq) + cd.ee;
                const float inv1 = __builtin_amdgcn_rsqf(r1sq);
                const float inv2 = __builtin_amdgcn_rsqf(r2sq);
                const float r12 = (r1sq * inv1) * (r2sq * inv2);
                const float den = r12 + (r1sq + de);
                // the reference skips 4ac - b^2 = 4 |d x e|^2 below epsilon
                // (x 4 is exact: |d x e|^2 < epsilon / 4)
                const float xx = __fmaf_rn(x0, x0, __fmaf_rn(x1, x1, x2 * x2));
                // FP64 where the FP32 form loses digits: R1.R2 cancelling r1 r2, or
                // the pseudo-segment ending close to the target (r2 < r1 / 4, where
                // the FP32 r2^2 can even round to <= 0), or |d x e|^2 within the FP32
                // form's rounding (64 ulps of |d|^2 |e|^2) of the skip threshold or
                // under it, where the FP64 term decides the skip; NaN goes there too
                const unsigned long long mill = wave_ballot(!(den >= 0.0625f * r12)) |
                                                wave_ballot(!(r2sq >= 0.0625f * r1sq)) |
                                                wave_ballot(!(xx >= __fmaf_rn(kee, r1sq, 0.25f * (float)NV_EPS)));
                const unsigned long long mok = mfar[t] & ~mill;
                // x is finite for every lane, so x * 0 = 0 where the term is not taken
                const float fac = lane_in(mok) ? (inv1 + inv2) * __builtin_amdgcn_rcpf(den) : 0.0f;
                fa[t][0] += (double)(x0 * fac);
                fa[t][1] += (double)(x1 * fac);
                fa[t][2] += (double)(x2 * fac);
                evl += lane_in(mok) ? 1u : 0u;
                mneed[t] |= mfar[t] & mill;
            }
        }
#pragma unroll
        for (int t = 0; t < TPT; ++t) anyneed |= mneed[t];
        if (anyneed != 0ull) {
            const double *cc = (const double *)((const char *)circ64 + (unsigned)cur * 24u);
            const double c64[3] = {cc[0], cc[1], cc[2]};
#pragma unroll
            for (int t = 0; t < TPT; ++t)
                if (lane_in(mneed[t])) near_term64(v[t][0], v[t][1], v[t][2], c64, dd[t], kappa, nr[t], nev);
        }
        cur = down ? n1 : n2;
        nd = pick64(down, a, b);
    }
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        if (r[t] < n) {
            // image 0 (the unshifted walk, near field included) in FP64, the
            // other images' partials in FP32 after its block
            const double k4 = kappa / (4.0 * NV_PI);
            if (m == 0) {
#pragma unroll
                for (int k = 0; k < 3; ++k) part[k * npad + r[t]] = fa[t][k] * k4 + nr[t][k];
            } else {
                float *pp = (float *)(part + (size_t)3 * npad) + (size_t)(m - 1) * 3 * npad;
#pragma unroll
                for (int k = 0; k < 3; ++k) pp[k * npad + r[t]] = (float)(fa[t][k] * k4 + nr[t][k]);
            }
        }
    }
    }
    if (sing && lane == 0) atomicOr(&ctl->status, NVK_SINGULAR);
    nev += evl;
    nvis = visl;
    for (int off = 32; off > 0; off >>= 1) {
        nev += __shfl_down(nev, off);
        nvis += __shfl_down(nvis, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counters[0], nev + nevw);
        atomicAdd(&counters[1], nvis);
    }
}

// Persistent like k_walk_vel32: items (image, block of 64 targets); partial
// block [images][3][npad], from which k_finish (mode 3) rebuilds the
// reference's two sums, image 0 and the images in order
// (src/timestep.f90:107-115)
template <int BT>
__global__ __launch_bounds__(BT) void k_walk_vel(DState s, Phys p, Shifts sh, int n, int npad,
                                                 const int *__restrict__ order, int nord, WalkQueues wq,
                                                 const WNode *__restrict__ walk, int nwalk, double *__restrict__ part,
                                                 DCtl *ctl, unsigned long long *counters, const int *dn,
                                                 const int *dnw) {
    if (ctl->abort) return;
    if (dn) nord = n = dev_count(n, dn);
    if (dnw) nwalk = *dnw;
    wq = resolve(wq);
    const int lane = threadIdx.x & 63;
    int h = 0;   // queues emptied
    unsigned long long nev = 0, nvis = 0;
    bool sing = false;
    for (;;) {
    int m, tb;
    if (!next_walk_item(wq, h, m, tb)) break;
    const int q = tb * 64 + lane;
    const int r = q < nord ? order[q] : q;
    const bool act = r < n && s.infront[r] != 0;
    double xi[3] = {0, 0, 0};
    int beh = -5, me = -5;
    if (act) {
        xi[0] = s.x[r]; xi[1] = s.x[s.cap + r]; xi[2] = s.x[2 * s.cap + r];
        beh = s.behind[r];
        me = r + 1;
    }
    const double kappa = p.kappa, theta_c = p.tree_theta;
    const int corr = p.extra_corr;
    double acc[3] = {0, 0, 0};
    const double s0 = sh.s[m][0], s1 = sh.s[m][1], s2 = sh.s[m][2];
    const bool img0 = (m == 0);
    int lskip = act ? 0 : 0x7fffffff;
    int cur = 0;
    WNode nd = nwalk > 0 ? walk[0] : WNode{};
    while (cur < nwalk) {
        cur = wave_uniform(cur);
        // both possible successors are fetched while this node is evaluated
        const int n1 = cur + 1, n2 = nd.skip;
        const WNode a = walk[n1 < nwalk ? n1 : cur];
        const WNode b = walk[n2 < nwalk ? n2 : cur];
        bool open = false;
        if (cur >= lskip) {
            ++nvis;
            const double e0 = xi[0] - (nd.cent[0] + s0);
            const double e1 = xi[1] - (nd.cent[1] + s1);
            const double e2 = xi[2] - (nd.cent[2] + s2);
            const double dist = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            double theta;
            if (corr) {
                if (dist - nd.geo > 0.0) theta = nd.width / (dist - nd.geo);
                else theta = nd.width / dist;
            } else {
                theta = nd.width / dist;
            }
            if (nd.pcount == 1 || theta < theta_c) {
                lskip = nd.skip;
                bool use = true;
                if (nd.pcount == 1) {
                    const int j = nd.leaf_j;
                    if (img0 && (j == me || j == beh)) use = false;
                    if (use && dist < NV_EPS) { sing = true; use = false; }
                }
                if (use) {
                    const double v0 = ((nd.cent[0] + s0) - xi[0]);
                    const double v1 = ((nd.cent[1] + s1) - xi[1]);
                    const double v2 = ((nd.cent[2] + s2) - xi[2]);
                    const double a = dist * dist;
                    const double b = 2.0 * (v0 * nd.circ[0] + v1 * nd.circ[1] + v2 * nd.circ[2]);
                    const double c = nd.circ[0] * nd.circ[0] + nd.circ[1] * nd.circ[1] + nd.circ[2] * nd.circ[2];
                    const double D = 4.0 * a * c - b * b;
                    if (!(D < NV_EPS)) {
                        const double den = (2.0 * NV_PI) * D;
                        const double fac = (2.0 * c + b) / sqrt(a + b + c) - (b / sqrt(a));
                        acc[0] = acc[0] + ((v1 * nd.circ[2] - v2 * nd.circ[1]) * kappa / den) * fac;
                        acc[1] = acc[1] + ((v2 * nd.circ[0] - v0 * nd.circ[2]) * kappa / den) * fac;
                        acc[2] = acc[2] + ((v0 * nd.circ[1] - v1 * nd.circ[0]) * kappa / den) * fac;
                        ++nev;
                    }
                }
            } else {
                open = true;
            }
        }
        const bool down = wave_any(open);
        cur = down ? n1 : n2;
        nd = down ? a : b;
    }
    if (r < n) {
        double *pp = part + (size_t)m * 3 * npad;
#pragma unroll
        for (int k = 0; k < 3; ++k) pp[k * npad + r] = acc[k];
    }
    }
    if (sing) atomicOr(&ctl->status, NVK_SINGULAR);
    // wave-level counter partials
    for (int off = 32; off > 0; off >>= 1) {
        nev += __shfl_down(nev, off);
        nvis += __shfl_down(nvis, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counters[0], nev);
        atomicAdd(&counters[1], nvis);
    }
}

// 30-bit Morton keys of the current positions (1024 cells per box edge),
// inactive rows last: the target order of the close search, which runs
// after insertions the tree build did not see.
__global__ void k_target_keys(DState s, int nmax, const int *dn, double bsize, unsigned long long *keys, int *vals) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dev_count(nmax, dn)) return;
    vals[r] = r;
    unsigned code = 0xffffffffu;
    if (s.infront[r] != 0) {
        unsigned c[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int q = (int)floor((s.x[a * s.cap + r] / bsize + 0.5) * 1024.0);
            c[a] = (unsigned)(q < 0 ? 0 : q > 1023 ? 1023 : q);
        }
        code = nv_spread10(c[0]) | (nv_spread10(c[1]) << 1) | (nv_spread10(c[2]) << 2);
    }
    keys[r] = code;
}

// Walk record of the close search (32 bytes): FP64 centroid, one word and
// the node radius. The word is minus one minus a leaf's j =
// behind(leaf_infront), resolved when packed, or for an internal node the
// skip index with the node's level in bits 27-30 (internal nodes are at
// levels 0-15; skip < 2^27). The level gives the width, bsize 2^-level
// exactly, and so the bounds on dist^2 that decide the reference's opening
// test width / dist < 0.5 without its division (below lo the node is
// opened, above hi accepted; between them the test is evaluated as the
// reference does).
struct CNode {
    double cent[3];
    int word;
    float rad;   // every centroid and point below lies within rad of cent
};
constexpr int CLOSE_SKIP_BITS = 27;
constexpr int CLOSE_DEEP = 15;   // the largest level code the 4 bits above the skip index hold


// With d = RN(sqrt(dd)) and q = RN(w / d): q < 0.5 <=> w / d < 0.5 - 2^-55
// (0.5 is a power of two); dd > 4 w^2 (1 + 2^-46) makes d > 2 w (1 + 2^-48),
// so q < 0.5, and dd < 4 w^2 (1 - 2^-46) makes d < 2 w, so q >= 0.5.  The
// factors are far above the rounding of dd, of 4 w^2 and of the bounds.
constexpr double CLOSE_BAND = 0x1p-46;

__global__ void k_pack_close(DState s, const WNode *__restrict__ walk, const float *__restrict__ wrad, int nwmax,
                             const int *dnw, double bsize, CNode *__restrict__ out) {
    const int nwalk = dev_count(nwmax, dnw);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > nwalk) return;
    if (q == nwalk) {   // the spare record
        out[q] = CNode{};
        return;
    }
    const WNode w = walk[q];
    CNode o;
#pragma unroll
    for (int k = 0; k < 3; ++k) o.cent[k] = w.cent[k];
    // width = bsize 2^-level exactly (the build halves it level by level);
    // levels from CLOSE_DEEP on (deep trees only) are marked CLOSE_DEEP and
    // the walk reads their width from the full record
    const int level = min(ilogb(bsize / w.width), CLOSE_DEEP);
    o.word = w.pcount == 1 ? -(w.leaf_infront > 0 ? s.behind[w.leaf_infront - 1] : 0) - 1
                           : w.skip | (level << CLOSE_SKIP_BITS);
    o.rad = wrad[q];
    out[q] = o;
}

// src/tree.f90:190-244 pclose_tree on the stale tree: each lane accepts the
// nodes the reference's recursion accepts (leaves, or width / dist < 0.5);
// an accepted node closer than the best leaf so far sets closest = 0 (an
// internal node) or, not being i, infront(i) or behind(i), becomes the best
// (a leaf).  dist < best is decided from dd against best^2 as the opening
// test is, and sqrt runs only where it cannot be: the stored closestd is the
// reference's sqrt.  Control flow on wave-uniform conditions only, as in
// k_walk_vel32.  Persistent over 64-target items.
template <int BT>
__global__ __launch_bounds__(BT) void k_walk_close(DState s, int n, const int *__restrict__ order, int nord,
                                                   WalkQueues wq, const CNode *__restrict__ walk,
                                                   const WNode *__restrict__ full, int nwalk, double pm,
                                                   double b4, const DCtl *ctl, const int *dn, const int *dnw) {
    if (ctl && ctl->abort) return;
    if (dn) nord = n = dev_count(n, dn);
    if (dnw) nwalk = *dnw;
    wq = resolve(wq);
    const int lane = threadIdx.x & 63;
    int h = 0;   // queues emptied
    for (;;) {
    int m, tb;
    if (!next_walk_item(wq, h, m, tb)) break;
    const int q = tb * 64 + lane;
    const int r = q < nord ? order[q] : q;
    if (r < n) s.closestd[r] = 100.0;   // set for every slot (src/tree.f90:196)
    const bool act = r < n && s.infront[r] != 0;
    double xi[3] = {0, 0, 0};
    int inf = -5, beh = -5, me = -5;
    if (act) {
        xi[0] = s.x[r]; xi[1] = s.x[s.cap + r]; xi[2] = s.x[2 * s.cap + r];
        inf = s.infront[r];
        beh = s.behind[r];
        me = r + 1;
    }
    double best = 100.0;
    double blo = best * best * (1.0 - CLOSE_BAND), bhi = best * best * (1.0 + CLOSE_BAND);
    double bpm = best + pm;
    int closest = act ? s.closest[r] : 0;
    int lskip = act ? 0 : 0x7fffffff;
    int cur = 0;
    CNode nd = nwalk > 0 ? walk[0] : CNode{};
    while (cur < nwalk) {
        cur = wave_uniform(cur);
        const int w = nd.word;
        const bool leaf = w < 0;
        // both possible successors are fetched while this node is evaluated
        // (the records end with one spare, so neither index needs a clamp)
        const int n1 = cur + 1, n2 = leaf ? cur + 1 : (w & ((1 << CLOSE_SKIP_BITS) - 1));
        const CNode a = walk[n1];
        const CNode b = walk[n2];
        // Lane conditions as 64-bit masks from ballots of raw compares, as
        // in the velocity walk; per-lane updates read them with lane_in.
        const unsigned long long mvis = wave_ballot(cur >= lskip);
        const double e0 = xi[0] - nd.cent[0], e1 = xi[1] - nd.cent[1], e2 = xi[2] - nd.cent[2];
        const double dd = e0 * e0 + e1 * e1 + e2 * e2;
        // dist < best: sure below blo, excluded above bhi
        const unsigned long long mlo = wave_ballot(dd < blo), mhi = wave_ballot(dd < bhi);
        bool down;
        if (!leaf) {
            // accepted (width / dist < 0.5): sure above hi, excluded below
            // lo, decided by the reference's division in between
            // 4 width^2 = 4 RN(bsize^2) 2^-2 level exactly, as k_pack_close and
            // the reference's width give it
            const int lcode = w >> CLOSE_SKIP_BITS;
            double w4;
            if (lcode < CLOSE_DEEP) {
                w4 = ldexp(b4, -2 * lcode);
            } else {
                const double fw = full[cur].width;
                w4 = 4.0 * (fw * fw);   // = 4 RN(bsize^2) 2^-2 level: scaling by 2^-level is exact
            }
            const double lo = w4 * (1.0 - CLOSE_BAND), hi = w4 * (1.0 + CLOSE_BAND);
            const unsigned long long mfar = wave_ballot(dd > hi), mnear = wave_ballot(dd < lo);
            const unsigned long long mband = mvis & ~mfar & ~mnear;
            unsigned long long myes = mfar;
            if (mband != 0ull) {
                const double width = full[cur].width;
                myes |= mband & wave_ballot(width / sqrt(dd) < .5);
            }
            const unsigned long long macc = mvis & myes;
            // Exact pruning: every centroid and point below lies within rad
            // of cent, so a lane farther than best + rad (and a margin far
            // above the rounding of the distances) finds nothing there with
            // dist < best, and dist < best is the only test that changes
            // its closest or best: the subtree is skipped with no effect.
            const double thr = bpm + (double)nd.rad;
            const unsigned long long mprune = mvis & ~macc & wave_ballot(dd > thr * thr * (1.0 + 0x1p-30));
            // an accepted node closer than the best leaf so far
            unsigned long long mlt = macc & mlo;
            const unsigned long long mmaybe = macc & ~mlo & mhi;
            if (mmaybe != 0ull) mlt |= mmaybe & wave_ballot(sqrt(dd) < best);
            closest = lane_in(mlt) ? 0 : closest;
            lskip = lane_in(macc | mprune) ? n2 : lskip;
            down = (mvis & ~macc & ~mprune) != 0ull;
        } else {
            const int j = -w - 1;
            const unsigned long long mcand =
                mvis & mhi & ~(wave_ballot(me == j) | wave_ballot(inf == j) | wave_ballot(beh == j));
            if (mcand != 0ull) {
                const double dist = sqrt(dd);
                if (lane_in(mcand & wave_ballot(dist < best))) {
                    closest = j;
                    best = dist;
                    blo = best * best * (1.0 - CLOSE_BAND);
                    bhi = best * best * (1.0 + CLOSE_BAND);
                    bpm = best + pm;
                }
            }
            lskip = lane_in(mvis) ? n2 : lskip;
            down = false;
        }
        cur = down ? n1 : n2;
        nd = pick64(down, a, b);
    }
    if (act) {
        s.closest[r] = closest;
        s.closestd[r] = best;
    }
    }
}

// ------------------------------------------------------------ solid walls
// src/tree.f90:321-392 tree_walk_mirror for target row r, summed over the
// reflections in order (src/timestep.f90:116-120) into mir[3][npad]: each
// lane walks the depth-first records on its own (wall runs are small and
// the mirrored walks short).  A node is accepted when it is a leaf or
// width / dist < theta, dist being to the reflected end of the node's
// pseudo-segment, centroid + circ (the extra correction's theta is
// overwritten there, :356); a leaf whose segment goes into a wall, and a
// pinned target's own image, are skipped (:360-372).
__global__ __launch_bounds__(64) void k_walk_mirror(DState s, Phys p, int n, const WNode *__restrict__ walk, int nwalk,
                                                    double *__restrict__ mir, int npad, DCtl *ctl,
                                                    unsigned long long *counters, const int *dn, const int *dnw) {
    if (ctl->abort) return;
    if (dn) n = dev_count(npad, dn);
    if (dnw) nwalk = *dnw;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npad) return;
    double u[3] = {0.0, 0.0, 0.0};
    unsigned long long nev = 0;
    bool sing = false;
    if (r < n && s.infront[r] != 0) {
        const int cap = s.cap;
        const double x0 = s.x[r], x1 = s.x[cap + r], x2 = s.x[2 * cap + r];
        const int pin = s.pin[r], beh = s.behind[r];
        const int w0 = s.wp[r], w1 = s.wp[cap + r], w2 = s.wp[2 * cap + r];
        for (int m = 0; m < p.nmir; ++m) {
            const int rf[3] = {p.mir[m][0], p.mir[m][1], p.mir[m][2]};
            const bool own = ((w0 - rf[0]) + (w1 - rf[1]) + (w2 - rf[2])) == 0 && (w0 + w1 + w2) < 2;
            int cur = 0;
            while (cur < nwalk) {
                const WNode nd = walk[cur];
                double sim[3], sgi[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    sgi[k] = reflect1(nd.cent[k], rf[k], p.box[k]);
                    sim[k] = reflect1(nd.cent[k] + nd.circ[k], rf[k], p.box[k]);
                }
                const double v0 = sim[0] - x0, v1 = sim[1] - x1, v2 = sim[2] - x2;
                const double e0 = x0 - sim[0], e1 = x1 - sim[1], e2 = x2 - sim[2];
                const double dd = e0 * e0 + e1 * e1 + e2 * e2;
                const double dist = sqrt(dd);
                const double theta = nd.width / dist;
                if (!(nd.pcount == 1 || theta < p.tree_theta)) {
                    ++cur;   // open: the first child follows
                    continue;
                }
                cur = nd.skip;
                if (nd.pcount == 1) {
                    const int j = s.behind[nd.leaf_infront - 1];
                    if (s.pin[j - 1] & NV_PINNED_I) continue;
                    if (own) {
                        if ((pin & NV_PINNED_B) && j == r + 1) continue;
                        if ((pin & NV_PINNED_I) && j == beh) continue;
                    }
                    if (dist < NV_EPS) {
                        sing = true;
                        continue;
                    }
                }
                const double c[3] = {sgi[0] - sim[0], sgi[1] - sim[1], sgi[2] - sim[2]};
                tree_term64(v0, v1, v2, c, dd, p.kappa, u, nev);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) mir[(size_t)k * npad + r] = u[k];
    if (sing) atomicOr(&ctl->status, NVK_SINGULAR);
    for (int off = 32; off > 0; off >>= 1) nev += __shfl_down(nev, off);
    if ((threadIdx.x & 63) == 0 && nev) atomicAdd(&counters[0], nev);
}

// src/tree.f90:190-244 pclose_tree / closest_tree with pinned points, each
// lane walking the stale tree's depth-first records on its own: every slot
// gets closestd = 100, a pinned target no search, a pinned leaf is no
// candidate (:226-234).
__global__ __launch_bounds__(64) void k_close_walls(DState s, int n, const WNode *__restrict__ walk, int nwalk,
                                                    const DCtl *ctl, const int *dn, const int *dnw, int npad) {
    if (ctl && ctl->abort) return;
    if (dn) n = dev_count(npad, dn);
    if (dnw) nwalk = *dnw;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    s.closestd[r] = 100.0;
    if (s.infront[r] == 0 || s.pin[r]) return;
    const int cap = s.cap;
    const double x0 = s.x[r], x1 = s.x[cap + r], x2 = s.x[2 * cap + r];
    const int inf = s.infront[r], beh = s.behind[r];
    double best = 100.0;
    int cl = s.closest[r];
    int cur = 0;
    while (cur < nwalk) {
        const WNode nd = walk[cur];
        const double e0 = x0 - nd.cent[0], e1 = x1 - nd.cent[1], e2 = x2 - nd.cent[2];
        const double dist = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
        const double theta = nd.width / dist;
        if (!(nd.pcount == 1 || theta < 0.5)) {
            ++cur;
            continue;
        }
        cur = nd.skip;
        if (!(dist < best)) continue;
        if (nd.pcount > 1) {
            cl = 0;
            continue;
        }
        const int j = s.behind[nd.leaf_infront - 1];
        if (j == r + 1 || inf == j || beh == j) continue;
        if (s.pin[j - 1]) continue;
        cl = j;
        best = dist;
    }
    s.closest[r] = cl;
    s.closestd[r] = best;
}

}  // namespace

// ------------------------------------------------------------------ host
struct TreeBuf {
    int cap = 0;                         // rows the key buffers hold
    unsigned long long *k0 = nullptr, *k1 = nullptr;
    int *v0 = nullptr, *v1 = nullptr;
    // deep builds: words 0.. of the paths by row (raw) and words 1.. sorted
    int kwords = 1;                      // path words of the builds (grows when a tree is deeper)
    int maxlev = KEY_LEVELS;             // = KEY_LEVELS * kwords
    unsigned long long *raw[KEY_WORDS] = {}, *srt[KEY_WORDS - 1] = {};
    int deep_cap = 0;
    int *hist = nullptr, *tmp = nullptr;
    int hist_cap = 0;
    int *nact = nullptr;                 // device
    Nodes nd{};
    int *flags = nullptr, *offs = nullptr, *scan_tmp = nullptr;
    int flag_cap = 0;
    WNode *walk = nullptr;
    float *wrad = nullptr;               // node radii in walk order (close-search pruning)
    int walk_cap = 0;
    Node48 *walk48 = nullptr;            // FP32 walk records, packed per walk
    double *circ64 = nullptr;            // their FP64 circulations
    int walk48_cap = 0;
    CNode *walkc = nullptr;              // close-search records, packed per search
    int walkc_cap = 0;
    int nwalk = 0, nnodes = 0, nact_h = 0;
    int built_rows = 0;                  // rows the last build sorted into vals
    int grid32 = 0, grid64 = 0, gridc = 0;   // resident workgroups of the persistent walks
    std::vector<int> lvl;                // level start offsets
    unsigned long long *counters = nullptr;
    int *wq = nullptr;                   // the walks' per-XCD work queues
    int *h_int = nullptr;                // pinned scratch
    const int *vals = nullptr;           // sorted rows of the last build
    LevelCtl *lc = nullptr, *lc_h = nullptr;   // level loop control, device and pinned copy
    int sync_builds = 0;                 // builds that ran the synchronous level loop
    long gen = 0;                        // bumped whenever a buffer above moves (recorded steps hold them)
    double *psum = nullptr;              // solid walls: the pinned-in-front points' position sum and count
    bool walls = false;                  // the builds count them in the root's centroid
    // the plan of the recorded steps' builds: levels 0 .. g_top launched on
    // these bounds, over g_rows rows
    int g_top = -1, g_rows = 0;
    int g_bound[LEVELS + 1] = {}, g_wgrid[LEVELS + 1] = {}, g_hint[LEVELS + 1] = {};
};

#define TCHK(x)                                     \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return NV_ERR_HIP;    \
    } while (0)

TreeBuf *nv_tree_new() {
    TreeBuf *t = new TreeBuf();
    if (hipMalloc(&t->nact, sizeof(int)) != hipSuccess || hipMalloc(&t->counters, 4 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&t->wq, XCDS * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&t->h_int, 64 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&t->lc, sizeof(LevelCtl)) != hipSuccess || hipMalloc(&t->psum, 4 * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void **)&t->lc_h, sizeof(LevelCtl), hipHostMallocDefault) != hipSuccess) {
        delete t;
        return nullptr;
    }
    return t;
}

static void free_nodes(Nodes &nd) {
    int **ip[] = {&nd.beg, &nd.cnt, &nd.pcount, &nd.dropb, &nd.first, &nd.parent, &nd.minrow, &nd.size, &nd.dfs,
                  &nd.leafinf};
    for (int **p : ip)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    double **dp[] = {&nd.pos, &nd.w, &nd.sum, &nd.circ, &nd.rad, &nd.cent};
    for (double **p : dp)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    nd.ncap = 0;
}

void nv_tree_free(TreeBuf *t) {
    if (!t) return;
    (void)hipFree(t->k0); (void)hipFree(t->k1); (void)hipFree(t->v0); (void)hipFree(t->v1); (void)hipFree(t->hist); (void)hipFree(t->tmp);
    (void)hipFree(t->nact); (void)hipFree(t->flags); (void)hipFree(t->offs); (void)hipFree(t->scan_tmp); (void)hipFree(t->walk);
    (void)hipFree(t->wrad);
    (void)hipFree(t->walk48);
    (void)hipFree(t->circ64);
    (void)hipFree(t->walkc);
    (void)hipFree(t->counters);
    (void)hipFree(t->wq);
    (void)hipHostFree(t->h_int);
    (void)hipFree(t->lc);
    (void)hipFree(t->psum);
    for (auto *q : t->raw) (void)hipFree(q);
    for (auto *q : t->srt) (void)hipFree(q);
    (void)hipHostFree(t->lc_h);
    free_nodes(t->nd);
    delete t;
}

// grow node arrays to hold `want` nodes, keeping the first `keep`
static int grow_nodes(TreeBuf *t, hipStream_t st, int want, int keep) {
    if (want <= t->nd.ncap) return NV_OK;
    int ncap = std::max(want, t->nd.ncap * 2);
    Nodes o = t->nd, n{};
    n.ncap = ncap;
    int **ip[] = {&n.beg, &n.cnt, &n.pcount, &n.dropb, &n.first, &n.parent, &n.minrow, &n.size, &n.dfs, &n.leafinf};
    int *oi[] = {o.beg, o.cnt, o.pcount, o.dropb, o.first, o.parent, o.minrow, o.size, o.dfs, o.leafinf};
    for (int q = 0; q < 10; ++q) {
        TCHK(hipMalloc(ip[q], sizeof(int) * (size_t)ncap));
        if (oi[q] && keep) TCHK(hipMemcpyAsync(*ip[q], oi[q], sizeof(int) * keep, hipMemcpyDeviceToDevice, st));
    }
    TCHK(hipMalloc(&n.w, sizeof(double) * (size_t)ncap));
    if (o.w && keep) TCHK(hipMemcpyAsync(n.w, o.w, sizeof(double) * keep, hipMemcpyDeviceToDevice, st));
    TCHK(hipMalloc(&n.rad, sizeof(double) * (size_t)ncap));
    TCHK(hipMalloc(&n.cent, sizeof(double) * 3 * (size_t)ncap));
    double **dp[] = {&n.pos, &n.sum, &n.circ};
    double *od[] = {o.pos, o.sum, o.circ};
    for (int q = 0; q < 3; ++q) {
        TCHK(hipMalloc(dp[q], sizeof(double) * 3 * (size_t)ncap));
        if (od[q] && keep)
            for (int k = 0; k < 3; ++k)
                TCHK(hipMemcpyAsync(*dp[q] + (size_t)k * ncap, od[q] + (size_t)k * o.ncap, sizeof(double) * keep,
                                    hipMemcpyDeviceToDevice, st));
    }
    TCHK(hipStreamSynchronize(st));
    free_nodes(o);
    t->nd = n;
    ++t->gen;
    return NV_OK;
}

static int ensure_rows(TreeBuf *t, int n) {
    if (n <= t->cap) return NV_OK;
    int cap = std::max(n, t->cap + t->cap / 2);
    (void)hipFree(t->k0); (void)hipFree(t->k1); (void)hipFree(t->v0); (void)hipFree(t->v1); (void)hipFree(t->hist); (void)hipFree(t->tmp);
    TCHK(hipMalloc(&t->k0, sizeof(unsigned long long) * cap));
    TCHK(hipMalloc(&t->k1, sizeof(unsigned long long) * cap));
    TCHK(hipMalloc(&t->v0, sizeof(int) * cap));
    TCHK(hipMalloc(&t->v1, sizeof(int) * cap));
    int nblk = nv_radix_blocks(cap);
    TCHK(hipMalloc(&t->hist, sizeof(int) * 256 * (size_t)nblk));
    TCHK(hipMalloc(&t->tmp, sizeof(int) * (size_t)nv_scan_tmp_ints(256 * nblk)));
    t->cap = cap;
    ++t->gen;
    return NV_OK;
}

static int ensure_flags(TreeBuf *t, int m) {
    if (m <= t->flag_cap) return NV_OK;
    int cap = std::max(m, 2 * t->flag_cap);
    (void)hipFree(t->flags); (void)hipFree(t->offs); (void)hipFree(t->scan_tmp);
    TCHK(hipMalloc(&t->flags, sizeof(int) * cap));
    TCHK(hipMalloc(&t->offs, sizeof(int) * cap));
    TCHK(hipMalloc(&t->scan_tmp, sizeof(int) * (size_t)nv_scan_tmp_ints(cap)));
    t->flag_cap = cap;
    ++t->gen;
    return NV_OK;
}

Keys sorted_keys(const TreeBuf *t) {
    return Keys{t->k0, t->kwords > 1 ? t->srt[0] : nullptr, t->kwords > 2 ? t->srt[1] : nullptr};
}

// buffers of the path words of a deep build
static int ensure_deep(TreeBuf *t, int n) {
    if (n <= t->deep_cap) return NV_OK;
    const int cap = std::max(n, t->cap);
    for (auto *&q : t->raw) { (void)hipFree(q); q = nullptr; }
    for (auto *&q : t->srt) { (void)hipFree(q); q = nullptr; }
    for (auto *&q : t->raw) TCHK(hipMalloc(&q, sizeof(unsigned long long) * (size_t)cap));
    for (auto *&q : t->srt) TCHK(hipMalloc(&q, sizeof(unsigned long long) * (size_t)cap));
    t->deep_cap = cap;
    ++t->gen;
    return NV_OK;
}

// The paths of rows [0, n) sorted (stable LSD: the last word first), word 0
// in k0 and the rows in v0 (t->vals); deeper words gathered into srt[].
// With dn the row count is *dn (<= n), read on the device.
static int sort_paths(TreeBuf *t, hipStream_t st, DState d, int n, const int *dn, double bsize) {
    const int K = t->kwords;
    const dim3 g((n + 255) / 256), b(256);
    if (K == 1) {
        hipLaunchKernelGGL(k_keys, g, b, 0, st, d, n, dn, bsize, 1, t->k0, (unsigned long long *)nullptr,
                           (unsigned long long *)nullptr, t->v0);
        TCHK(hipGetLastError());
        // stable LSD radix sort of the 64-bit paths; the result is back in k0/v0
        TCHK(nv_radix_sort_pairs_dn(st, t->k0, t->v0, t->k1, t->v1, n, dn, 64, t->hist, t->tmp));
        return NV_OK;
    }
    if (int rc = ensure_deep(t, n)) return rc;
    hipLaunchKernelGGL(k_keys, g, b, 0, st, d, n, dn, bsize, K, t->raw[0], t->raw[1], t->raw[2], t->v0);
    TCHK(hipGetLastError());
    // v0 is the identity here: each pass sorts the rows by one more word
    for (int w = K - 1; w >= 0; --w) {
        hipLaunchKernelGGL(k_gather_keys, g, b, 0, st, (const unsigned long long *)t->raw[w], (const int *)t->v0, n,
                           dn, t->k0);
        TCHK(hipGetLastError());
        TCHK(nv_radix_sort_pairs_dn(st, t->k0, t->v0, t->k1, t->v1, n, dn, 64, t->hist, t->tmp));
    }
    for (int w = 1; w < K; ++w)
        hipLaunchKernelGGL(k_gather_keys, g, b, 0, st, (const unsigned long long *)t->raw[w], (const int *)t->v0, n,
                           dn, t->srt[w - 1]);
    TCHK(hipGetLastError());
    return NV_OK;
}

// Level loop with a host round trip per level: the level sizes decide the
// launches. Leaves t->lvl and t->nwalk; nodes are moments-complete.
// From level `level` on, t->lvl holding the starts of levels 0 .. level + 1
// and the nodes of the levels above complete (children included).
static int level_loop_sync(TreeBuf *t, hipStream_t st, DState d, DCtl *ctl, int level) {
    const int nact = t->nact_h;
    while (true) {
        const int L0 = t->lvl[level], L1 = t->lvl[level + 1], m = L1 - L0;
        if (m == 0) break;
        if (int rc = ensure_flags(t, m)) return rc;
        // split flags, scanned in place into the children offsets
        hipLaunchKernelGGL(k_split_flags, dim3((m + 255) / 256), dim3(256), 0, st, t->nd, L0, L1, t->offs,
                           level == LEVELS ? 1 : 0, ctl);
        const Keys keys = sorted_keys(t);
        TCHK(hipGetLastError());
        TCHK(nv_exclusive_scan(st, t->offs, m, t->scan_tmp));
        TCHK(hipMemcpyAsync(t->h_int, t->offs + (m - 1), sizeof(int), hipMemcpyDeviceToHost, st));
        TCHK(hipMemcpyAsync(t->h_int + 1, t->nd.pcount + (L1 - 1), sizeof(int), hipMemcpyDeviceToHost, st));
        TCHK(hipStreamSynchronize(st));
        const int nsplit = t->h_int[0] + (t->h_int[1] > 1 ? 1 : 0);
        if (nsplit == 0 || level == t->maxlev) {
            if (nsplit && level == t->maxlev) return NV_ERR_TREE_DEPTH;
            break;
        }
        if (int rc = grow_nodes(t, st, L1 + 8 * nsplit, L1)) return rc;
        if (m <= 4096 && nsplit * 64 <= t->nact_h)   // few nodes holding many keys each
            hipLaunchKernelGGL(k_children_wide, dim3(m), dim3(256), 0, st, t->nd, L0, L1, t->offs, keys, level);
        else
            hipLaunchKernelGGL(k_children, dim3((m + 255) / 256), dim3(256), 0, st, t->nd, L0, L1, t->offs, keys,
                               level);
        TCHK(hipGetLastError());
        t->lvl.push_back(L1 + 8 * nsplit);
        ++level;
    }
    const int nlev = (int)t->lvl.size() - 1;
    for (int lv = nlev - 1; lv >= 0; --lv) {
        const int L0 = t->lvl[lv], L1 = t->lvl[lv + 1];
        if (L1 > L0)
            hipLaunchKernelGGL(k_up, dim3((L1 - L0 + 255) / 256), dim3(256), 0, st, t->nd, d, L0, L1, t->vals, nact,
                               (const double *)(t->walls ? t->psum : nullptr));
        TCHK(hipGetLastError());
    }
    TCHK(hipMemcpyAsync(t->h_int, t->nd.size, sizeof(int), hipMemcpyDeviceToHost, st));
    TCHK(hipStreamSynchronize(st));
    t->nwalk = t->h_int[0];
    return NV_OK;
}

static int build_levels_sync(TreeBuf *t, hipStream_t st, DState d, int n, double bsize, DCtl *ctl) {
    TCHK(hipMemcpyAsync(t->h_int, t->nact, sizeof(int), hipMemcpyDeviceToHost, st));
    TCHK(hipStreamSynchronize(st));
    t->nact_h = t->h_int[0];
    if (int rc = grow_nodes(t, st, std::max(1024, 4 * t->nact_h + 64), 0)) return rc;
    hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, st, t->nd, d, n, (const int *)nullptr, bsize, t->nact,
                       (LevelCtl *)nullptr);
    TCHK(hipGetLastError());
    t->lvl.assign(1, 0);
    t->lvl.push_back(1);
    return level_loop_sync(t, st, d, ctl, 0);
}

constexpr int BUILD_RETRY = -1000;

// Launch bounds of levels 0 .. top from the previous build's level sizes
// (levels 0 .. prev - 1; 128 for the level past them): the size + 1/8 + 256,
// rounded up to 256.  Returns the node count holding those levels and the
// children of level top.
static long long plan_levels(const TreeBuf *t, int prev, int top, int *bound, int *wgrid, int *fmax) {
    long long total = 1;
    *fmax = 1;
    for (int l = 0; l <= top; ++l) {
        const int hint = l < prev ? t->lvl[l + 1] - t->lvl[l] : 128;
        bound[l] = (hint + hint / 8 + 256 + 255) / 256 * 256;
        wgrid[l] = hint <= 4096 ? std::min(4096, bound[l]) : 0;
        total += bound[l];
        *fmax = std::max(*fmax, bound[l]);
    }
    // children of the deepest launched level land past its bound
    return total + 8ll * bound[top];
}

// The level loop on node ranges held in LevelCtl (launches only): levels
// 0 .. top on their bounds, the moments upwards, the walk length.  Rows
// [0, n), or [0, *dn) when dn is given.
static int enqueue_levels(TreeBuf *t, hipStream_t st, DState d, int n, const int *dn, double bsize, DCtl *ctl,
                          int top, const int *bound, const int *wgrid) {
    const Keys keys = sorted_keys(t);
    hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, st, t->nd, d, n, dn, bsize, t->nact, t->lc);
    TCHK(hipGetLastError());
    for (int l = 0; l <= top; ++l) {
        hipLaunchKernelGGL(k_split_flags_d, dim3(bound[l] / 256), dim3(256), 0, st, t->nd, (const LevelCtl *)t->lc, l,
                           bound[l], t->offs, ctl);
        TCHK(hipGetLastError());
        TCHK(nv_scan_tiles_sums(st, t->offs, bound[l], t->scan_tmp));
        hipLaunchKernelGGL(k_level_end, dim3(1), dim3(1), 0, st, t->nd, t->lc, l, t->maxlev, bound[l], wgrid[l],
                           (const int *)t->offs, (const int *)t->scan_tmp);
        TCHK(hipGetLastError());
        if (wgrid[l] > 0)
            hipLaunchKernelGGL(k_children_wide_d, dim3(wgrid[l]), dim3(256), 0, st, t->nd, (const LevelCtl *)t->lc,
                               (const int *)t->offs, (const int *)t->scan_tmp, keys, l);
        hipLaunchKernelGGL(k_children_d, dim3(bound[l] / 256), dim3(256), 0, st, t->nd, (const LevelCtl *)t->lc,
                           (const int *)t->offs, (const int *)t->scan_tmp, keys, l);
        TCHK(hipGetLastError());
    }
    for (int l = top; l >= 0; --l) {
        hipLaunchKernelGGL(k_up_d, dim3(bound[l] / 256), dim3(256), 0, st, t->nd, d, (const LevelCtl *)t->lc, l,
                           t->vals, (const double *)(t->walls ? t->psum : nullptr));
        TCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_build_tail, dim3(1), dim3(1), 0, st, t->nd, t->lc);
    TCHK(hipGetLastError());
    return NV_OK;
}

// The same level loop with no host round trip until its end: every level is
// launched on a bound taken from the previous build's level sizes (one more
// level than it had), the kernels take their node range from LevelCtl, and a
// level that outgrows its bound, the node arrays or the levels launched makes
// the build return BUILD_RETRY, after which the caller runs the synchronous
// loop from the root again. Both loops run the same kernels on the same
// ranges, so the trees are identical.
static int build_levels_async(TreeBuf *t, hipStream_t st, DState d, int n, double bsize, DCtl *ctl) {
    const int prev = (int)t->lvl.size() - 1;   // levels of the previous build
    if (prev < 1 || !t->lc) return BUILD_RETRY;
    const int top = std::min(prev, t->maxlev);   // deepest level launched
    int bound[LEVELS + 1], wgrid[LEVELS + 1], fmax = 1;
    const long long total = plan_levels(t, prev, top, bound, wgrid, &fmax);
    if (total > (1ll << 30)) return BUILD_RETRY;
    if (int rc = grow_nodes(t, st, (int)total, 0)) return rc;
    if (int rc = ensure_flags(t, fmax)) return rc;
    if (int rc = enqueue_levels(t, st, d, n, nullptr, bsize, ctl, top, bound, wgrid)) return rc;
    TCHK(hipMemcpyAsync(t->lc_h, t->lc, sizeof(LevelCtl), hipMemcpyDeviceToHost, st));
    TCHK(hipStreamSynchronize(st));
    const LevelCtl &h = *t->lc_h;
    if (h.depth) return NV_ERR_TREE_DEPTH;
    t->nact_h = h.nact;
    // A level that outgrew its bound, or a tree deeper than the levels
    // launched (children past the last one): the levels above are complete,
    // and the synchronous loop goes on from there.
    const int resume = h.overflow ? h.ovf_level : h.lvl[top + 2] > h.lvl[top + 1] ? top + 1 : -1;
    if (resume >= 0) {
        t->lvl.assign(h.lvl, h.lvl + resume + 2);
        ++t->sync_builds;
        return level_loop_sync(t, st, d, ctl, resume);
    }
    int nlev = 0;
    while (nlev <= top && h.lvl[nlev + 1] > h.lvl[nlev]) ++nlev;
    t->lvl.assign(h.lvl, h.lvl + nlev + 1);
    t->nwalk = h.nwalk;
    return NV_OK;
}

int nv_tree_build(TreeBuf *t, hipStream_t st, DState d, Phys p, int n, DCtl *ctl, nv_stats *stats) {
    if (n <= 0) return NV_ERR_ZERO_LENGTH;
    if (int rc = ensure_rows(t, n)) return rc;
    const double bsize = std::max(p.box[0], std::max(p.box[1], p.box[2]));
    int rc = NV_OK;
    for (;;) {
        t->maxlev = KEY_LEVELS * t->kwords;
        TCHK(nv_fill_int(st, t->nact, 0, 1));
        if ((rc = sort_paths(t, st, d, n, nullptr, bsize))) return rc;
        t->walls = p.walls != 0;
        if (t->walls) hipLaunchKernelGGL(k_pinned_sum, dim3(1), dim3(256), 0, st, d, n, (const int *)nullptr, t->psum);
        hipLaunchKernelGGL(k_count_active, dim3((n + 255) / 256), dim3(256), 0, st, t->k0, n, (const int *)nullptr,
                           t->nact);
        TCHK(hipGetLastError());
        t->vals = t->v0;
        t->built_rows = n;
        rc = build_levels_async(t, st, d, n, bsize, ctl);
        if (rc == BUILD_RETRY) {
            ++t->sync_builds;
            rc = build_levels_sync(t, st, d, n, bsize, ctl);
        }
        if (rc == NV_ERR_TREE_DEPTH && t->kwords < KEY_WORDS) {
            // a node of the deepest level still holds several points: build
            // again with 16 more levels (later builds keep the depth)
            ++t->kwords;
            t->lvl.clear();
            continue;
        }
        break;
    }
    if (rc) {
        t->lvl.clear();   // no hints from a failed build
        return rc;
    }
    const int nlev = (int)t->lvl.size() - 1;
    t->nnodes = t->lvl.back();
    if (t->nwalk > t->walk_cap) {
        (void)hipFree(t->walk);
        (void)hipFree(t->wrad);
        t->walk = nullptr;
        t->wrad = nullptr;
        t->walk_cap = std::max(t->nwalk, 2 * t->walk_cap);
        TCHK(hipMalloc(&t->walk, sizeof(WNode) * (size_t)t->walk_cap));
        TCHK(hipMalloc(&t->wrad, sizeof(float) * (size_t)t->walk_cap));
        ++t->gen;
    }
    for (int lv = 0; lv < nlev; ++lv) {
        const int L0 = t->lvl[lv], L1 = t->lvl[lv + 1];
        if (L1 > L0)
            hipLaunchKernelGGL(k_down, dim3((L1 - L0 + 255) / 256), dim3(256), 0, st, t->nd, d, L0, L1, t->walk,
                               t->wrad);
        TCHK(hipGetLastError());
    }
    stats->tree_nodes = t->nnodes;
    stats->tree_levels = nlev;
    stats->tree_builds += 1;
    stats->tree_sync_builds = t->sync_builds;
    return NV_OK;
}

// The least double d with sqrt(d) >= e: for d >= 0, sqrt(d) < e <=> d < it
// (a correctly rounded square root is monotone).
static double sqrt_threshold(double e) {
    double lo = 0.0, hi = e * e * 4.0;   // sqrt(lo) < e <= sqrt(hi)
    while (true) {
        const double mid = lo + (hi - lo) / 2.0;
        if (mid <= lo || mid >= hi) break;
        (std::sqrt(mid) < e ? lo : hi) = mid;
    }
    while (std::sqrt(lo) >= e) lo = std::nextafter(lo, 0.0);
    double d = lo;
    while (std::sqrt(d) < e) d = std::nextafter(d, 1.0);
    return d;
}

// workgroups of 256 lanes resident on the device at once for kernel k
template <typename K>
static int resident_blocks(K k) {
    int per_cu = 0, ncu = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess)
        return 0;
    return std::max(1, per_cu) * std::max(1, ncu);
}

int nv_tree_walk_velocity(TreeBuf *t, hipStream_t st, DState d, Phys p, const Shifts &sh, int n, int npad,
                          int precision, double rn2, double *part, DCtl *ctl, nv_stats *stats) {
    (void)stats;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(4), 0, st, t->counters);
    if (n <= 0) return NV_OK;
    // Targets are taken in the tree's Morton order (the build's sorted rows,
    // active first), so that the 64 lanes of a wave are close together and
    // their traversals overlap; rows the build did not see follow in row
    // order.  Each lane's sequence of accepted nodes does not depend on its
    // wave-mates, so the order changes no result.  Rows >= n hold no target
    // and k_finish reads no partial there.
    const int *order = t->vals;
    const int nord = t->vals && t->built_rows <= n ? t->built_rows : 0;
    const int nb64 = (n + 63) / 64, nitems = nb64 * sh.n;
    // the FP32 walk's single opening radius needs width / theta > geo
    const bool fp32 = precision == NV_PREC_FP32C && p.tree_theta <= 1.1;
    if (fp32 && (long long)t->nwalk + 1 > WALK32_MAX_RECORDS) return NV_ERR_CAPACITY;
    if (fp32) {
        // one spare record past the last: the walk fetches a successor index
        // up to nwalk without a clamp
        if (t->nwalk + 1 > t->walk48_cap) {
            (void)hipFree(t->walk48);
            t->walk48 = nullptr;
            t->walk48_cap = std::max(t->nwalk + 1, 2 * t->walk48_cap);
            TCHK(hipMalloc(&t->walk48, sizeof(Node48) * (size_t)t->walk48_cap));
            (void)hipFree(t->circ64);
            t->circ64 = nullptr;
            TCHK(hipMalloc(&t->circ64, 3 * sizeof(double) * (size_t)t->walk48_cap));
            ++t->gen;
        }
        hipLaunchKernelGGL(k_pack48, dim3((t->nwalk + 1 + 255) / 256), dim3(256), 0, st, d, t->walk, t->nwalk,
                           (const int *)nullptr, p.tree_theta, p.extra_corr, t->walk48, t->circ64);
        // items: (image, WALK_TPT blocks of 64 targets)
        const int nbt = (n + 64 * WALK_TPT - 1) / (64 * WALK_TPT), nit32 = nbt * sh.n;
        if (t->grid32 == 0 && (t->grid32 = resident_blocks(k_walk_vel32<256, WALK_TPT>)) == 0) return NV_ERR_HIP;
        const int grid = std::min(t->grid32, (nit32 + 3) / 4);
        TCHK(nv_fill_int(st, t->wq, 0, XCDS));
        hipLaunchKernelGGL((k_walk_vel32<256, WALK_TPT>), dim3(grid), dim3(256), 0, st, d, p, sh, n, npad, order, nord,
                           WalkQueues{t->wq, nbt, sh.n}, t->walk48, t->circ64, t->walk, t->nwalk, rn2,
                           sqrt_threshold(NV_EPS), part, ctl, t->counters, (const int *)nullptr, (const int *)nullptr);
    } else {
        if (t->nwalk > 0)
            hipLaunchKernelGGL(k_leaf_links, dim3((t->nwalk + 255) / 256), dim3(256), 0, st, d, t->walk, t->nwalk,
                               (const int *)nullptr);
        if (t->grid64 == 0 && (t->grid64 = resident_blocks(k_walk_vel<256>)) == 0) return NV_ERR_HIP;
        const int grid = std::min(t->grid64, (nitems + 3) / 4);
        TCHK(nv_fill_int(st, t->wq, 0, XCDS));
        hipLaunchKernelGGL(k_walk_vel<256>, dim3(grid), dim3(256), 0, st, d, p, sh, n, npad, order, nord,
                           WalkQueues{t->wq, nb64, sh.n}, t->walk, t->nwalk, part, ctl, t->counters, (const int *)nullptr,
                           (const int *)nullptr);
    }
    TCHK(hipGetLastError());
    return NV_OK;
}

int nv_tree_walk_partial_mode(int precision, double theta) {
    return precision == NV_PREC_FP32C && theta <= 1.1 ? 4 : 3;
}

int nv_tree_walk_record_bytes(int precision, double theta) {
    return precision == NV_PREC_FP32C && theta <= 1.1 ? (int)sizeof(Node48) : (int)sizeof(WNode);
}

// queue a copy of the last velocity walk's interaction and node-visit
// counters into pinned memory (ints 32..35 of h_int)
int nv_tree_walk_counters_async(TreeBuf *t, hipStream_t st) {
    TCHK(hipMemcpyAsync(t->h_int + 32, t->counters, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    return NV_OK;
}

// the counters that copy brought, once the stream has passed it
void nv_tree_walk_counters_take(const TreeBuf *t, nv_stats *stats) {
    const unsigned long long *h = (const unsigned long long *)(t->h_int + 32);
    stats->tree_evals = (int64_t)h[0];
    stats->tree_nodes_visited = (int64_t)h[1];
}

// both of the above (synchronises)
int nv_tree_walk_counters(TreeBuf *t, hipStream_t st, nv_stats *stats) {
    if (int rc = nv_tree_walk_counters_async(t, st)) return rc;
    TCHK(hipStreamSynchronize(st));
    nv_tree_walk_counters_take(t, stats);
    return NV_OK;
}

int nv_tree_walk_close(TreeBuf *t, hipStream_t st, DState d, double bsize, int n, nv_stats *stats) {
    (void)stats;
    if (n <= 0) return NV_OK;
    // Targets in the Morton order of their current positions.  The build's
    // order holds points inserted since the build (at its end, or among the
    // inactive rows), whose waves would be scattered over the whole tangle;
    // the sort replaces it (the build's keys and rows are not needed after
    // the build; the velocity walk accepts any order of the rows).
    if (int rc = ensure_rows(t, n)) return rc;
    hipLaunchKernelGGL(k_target_keys, dim3((n + 255) / 256), dim3(256), 0, st, d, n, (const int *)nullptr, bsize,
                       t->k0, t->v0);
    TCHK(hipGetLastError());
    TCHK(nv_radix_sort_pairs(st, t->k0, t->v0, t->k1, t->v1, n, 32, t->hist, t->tmp));
    t->vals = t->v0;
    t->built_rows = n;
    // close-search records, one spare past the last
    if (t->nwalk + 1 > t->walkc_cap) {
        (void)hipFree(t->walkc);
        t->walkc = nullptr;
        t->walkc_cap = std::max(t->nwalk + 1, 2 * t->walkc_cap);
        TCHK(hipMalloc(&t->walkc, sizeof(CNode) * (size_t)t->walkc_cap));
        ++t->gen;
    }
    if (t->nwalk >= (1 << CLOSE_SKIP_BITS)) return NV_ERR_UNSUPPORTED;   // the skip index field
    hipLaunchKernelGGL(k_pack_close, dim3((t->nwalk + 1 + 255) / 256), dim3(256), 0, st, d, t->walk,
                       (const float *)t->wrad, t->nwalk, (const int *)nullptr, bsize, t->walkc);
    if (t->gridc == 0 && (t->gridc = resident_blocks(k_walk_close<256>)) == 0) return NV_ERR_HIP;
    const int nitems = (n + 63) / 64;
    TCHK(nv_fill_int(st, t->wq, 0, XCDS));
    hipLaunchKernelGGL(k_walk_close<256>, dim3(std::min(t->gridc, (nitems + 3) / 4)), dim3(256), 0, st, d, n, t->vals,
                       n, WalkQueues{t->wq, nitems, 1}, t->walkc, t->walk, t->nwalk, bsize * 0x1p-40,
                       4.0 * (bsize * bsize), (const DCtl *)nullptr, (const int *)nullptr, (const int *)nullptr);
    TCHK(hipGetLastError());
    return NV_OK;
}

// ------------------------------------------------------ recorded steps
// A recorded octree step reads no size on the host: its build sorts rows
// [0, *dn) of a fixed row range, launches levels 0 .. top (one more than the
// last build had) on bounds with slack over that build's level sizes (as the
// eager build's first attempt does), and holds walk records for every node
// those bounds allow.  k_build_check switches the step's later phases off
// when the build outgrows them; the host then takes the step again eagerly
// and records anew from that build's sizes.

// resident workgroups of each persistent walk
static int walk_grids(TreeBuf *t) {
    if (t->grid32 == 0 && (t->grid32 = resident_blocks(k_walk_vel32<256, WALK_TPT>)) == 0) return NV_ERR_HIP;
    if (t->grid64 == 0 && (t->grid64 = resident_blocks(k_walk_vel<256>)) == 0) return NV_ERR_HIP;
    if (t->gridc == 0 && (t->gridc = resident_blocks(k_walk_close<256>)) == 0) return NV_ERR_HIP;
    return NV_OK;
}

// grow a record array to hold `want` records
template <typename T>
static int ensure_records(TreeBuf *t, T **p, int *cap, int want) {
    if (want <= *cap) return NV_OK;
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    TCHK(hipMalloc(p, sizeof(T) * (size_t)want));
    *cap = want;
    ++t->gen;
    return NV_OK;
}

int nv_tree_graph_prepare(TreeBuf *t, hipStream_t st, int rows) {
    const int prev = (int)t->lvl.size() - 1;   // levels of the last build
    if (prev < 1 || rows <= 0) return NV_ERR_STATE;
    if (int rc = ensure_rows(t, rows)) return rc;
    if (t->kwords > 1)
        if (int rc = ensure_deep(t, rows)) return rc;
    t->maxlev = KEY_LEVELS * t->kwords;
    const int top = std::min(prev, t->maxlev);
    int fmax = 1;
    const long long total = plan_levels(t, prev, top, t->g_bound, t->g_wgrid, &fmax);
    // the close search's records index their successors in CLOSE_SKIP_BITS
    // bits, the FP32 walk its records by 32-bit byte offsets
    if (total + 1 >= (1ll << CLOSE_SKIP_BITS) || total + 1 > WALK32_MAX_RECORDS) return NV_ERR_UNSUPPORTED;
    for (int l = 0; l <= top; ++l) t->g_hint[l] = l < prev ? t->lvl[l + 1] - t->lvl[l] : 0;
    if (int rc = grow_nodes(t, st, (int)total, 0)) return rc;
    if (int rc = ensure_flags(t, fmax)) return rc;
    const int nrec = (int)total + 1;   // every node a build within the bounds makes, and the spare
    if (nrec > t->walk_cap) {
        (void)hipFree(t->wrad);
        t->wrad = nullptr;
        int cap = t->walk_cap;
        if (int rc = ensure_records(t, &t->walk, &cap, nrec)) return rc;
        TCHK(hipMalloc(&t->wrad, sizeof(float) * (size_t)cap));
        t->walk_cap = cap;
    }
    if (nrec > t->walk48_cap) {
        (void)hipFree(t->circ64);
        t->circ64 = nullptr;
        int cap = t->walk48_cap;
        if (int rc = ensure_records(t, &t->walk48, &cap, nrec)) return rc;
        TCHK(hipMalloc(&t->circ64, 3 * sizeof(double) * (size_t)cap));
        t->walk48_cap = cap;
    }
    if (int rc = ensure_records(t, &t->walkc, &t->walkc_cap, nrec)) return rc;
    if (int rc = walk_grids(t)) return rc;
    t->g_top = top;
    t->g_rows = rows;
    return NV_OK;
}

int nv_tree_graph_build(TreeBuf *t, hipStream_t st, DState d, Phys p, const int *dn, DCtl *ctl) {
    const int n = t->g_rows, top = t->g_top;
    if (top < 0) return NV_ERR_STATE;
    const double bsize = std::max(p.box[0], std::max(p.box[1], p.box[2]));
    TCHK(nv_fill_int(st, t->nact, 0, 1));
    if (int rc = sort_paths(t, st, d, n, dn, bsize)) return rc;
    t->walls = p.walls != 0;
    if (t->walls) hipLaunchKernelGGL(k_pinned_sum, dim3(1), dim3(256), 0, st, d, n, dn, t->psum);
    hipLaunchKernelGGL(k_count_active, dim3((n + 255) / 256), dim3(256), 0, st, t->k0, n, dn, t->nact);
    TCHK(hipGetLastError());
    t->vals = t->v0;
    if (int rc = enqueue_levels(t, st, d, n, dn, bsize, ctl, top, t->g_bound, t->g_wgrid)) return rc;
    const int nwalk_max = std::min(t->walk_cap, std::min(t->walk48_cap, t->walkc_cap)) - 1;
    hipLaunchKernelGGL(k_build_check, dim3(1), dim3(1), 0, st, t->lc, ctl, top, nwalk_max, dn);
    TCHK(hipGetLastError());
    for (int l = 0; l <= top; ++l)
        hipLaunchKernelGGL(k_down_d, dim3(t->g_bound[l] / 256), dim3(256), 0, st, t->nd, d, (const LevelCtl *)t->lc,
                           l, t->walk, t->wrad, (const DCtl *)ctl);
    TCHK(hipGetLastError());
    return NV_OK;
}

int nv_tree_graph_walk(TreeBuf *t, hipStream_t st, DState d, Phys p, const Shifts &sh, int precision, double rn2,
                       double *part, DCtl *ctl, const int *dn) {
    const int npad = t->g_rows;
    const int *dnw = &t->lc->nwalk;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(4), 0, st, t->counters);
    TCHK(nv_fill_int(st, t->wq, 0, XCDS));
    if (precision == NV_PREC_FP32C && p.tree_theta <= 1.1) {
        hipLaunchKernelGGL(k_pack48, dim3((t->walk48_cap + 255) / 256), dim3(256), 0, st, d, t->walk,
                           t->walk48_cap - 1, dnw, p.tree_theta, p.extra_corr, t->walk48, t->circ64);
        WalkQueues wq{t->wq, 0, sh.n, dn, npad, 64 * WALK_TPT};
        hipLaunchKernelGGL((k_walk_vel32<256, WALK_TPT>), dim3(t->grid32), dim3(256), 0, st, d, p, sh, npad, npad,
                           (const int *)t->v0, 0, wq, t->walk48, t->circ64, t->walk, 0, rn2, sqrt_threshold(NV_EPS),
                           part, ctl, t->counters, dn, dnw);
    } else {
        hipLaunchKernelGGL(k_leaf_links, dim3((t->walk_cap + 255) / 256), dim3(256), 0, st, d, t->walk, t->walk_cap,
                           dnw);
        WalkQueues wq{t->wq, 0, sh.n, dn, npad, 64};
        hipLaunchKernelGGL(k_walk_vel<256>, dim3(t->grid64), dim3(256), 0, st, d, p, sh, npad, npad,
                           (const int *)t->v0, 0, wq, t->walk, 0, part, ctl, t->counters, dn, dnw);
    }
    TCHK(hipGetLastError());
    return NV_OK;
}

int nv_tree_graph_close(TreeBuf *t, hipStream_t st, DState d, double bsize, const int *dn, DCtl *ctl) {
    const int rows = t->g_rows;
    const int *dnw = &t->lc->nwalk;
    hipLaunchKernelGGL(k_target_keys, dim3((rows + 255) / 256), dim3(256), 0, st, d, rows, dn, bsize, t->k0, t->v0);
    TCHK(hipGetLastError());
    TCHK(nv_radix_sort_pairs_dn(st, t->k0, t->v0, t->k1, t->v1, rows, dn, 32, t->hist, t->tmp));
    hipLaunchKernelGGL(k_pack_close, dim3((t->walkc_cap + 255) / 256), dim3(256), 0, st, d, t->walk,
                       (const float *)t->wrad, t->walkc_cap - 1, dnw, bsize, t->walkc);
    TCHK(nv_fill_int(st, t->wq, 0, XCDS));
    hipLaunchKernelGGL(k_walk_close<256>, dim3(t->gridc), dim3(256), 0, st, d, rows, (const int *)t->v0, 0,
                       WalkQueues{t->wq, 0, 1, dn, rows, 64}, t->walkc, t->walk, 0, bsize * 0x1p-40,
                       4.0 * (bsize * bsize), (const DCtl *)ctl, dn, dnw);
    TCHK(hipGetLastError());
    return NV_OK;
}

int nv_tree_graph_gen(const TreeBuf *t) { return (int)(t->gen & 0x7fffffff); }

int nv_tree_graph_take_async(TreeBuf *t, hipStream_t st) {
    TCHK(hipMemcpyAsync(t->lc_h, t->lc, sizeof(LevelCtl), hipMemcpyDeviceToHost, st));
    return nv_tree_walk_counters_async(t, st);
}

bool nv_tree_graph_take(TreeBuf *t, nv_stats *stats) {
    const LevelCtl &h = *t->lc_h;
    const int top = t->g_top;
    int nlev = 0;
    while (nlev <= top && h.lvl[nlev + 1] > h.lvl[nlev]) ++nlev;
    t->lvl.assign(h.lvl, h.lvl + nlev + 1);
    t->nwalk = h.nwalk;
    t->nact_h = h.nact;
    t->nnodes = t->lvl.back();
    t->built_rows = 0;   // the close search reordered the rows
    nv_tree_walk_counters_take(t, stats);
    stats->tree_nodes = t->nnodes;
    stats->tree_levels = nlev;
    stats->tree_builds += 1;
    // record anew before the next build may outgrow the plan: a level past
    // half its slack, or a tree as deep as the levels launched
    bool stale = nlev > top;
    for (int l = 0; l < nlev && l <= top && !stale; ++l) {
        const int m = h.lvl[l + 1] - h.lvl[l];
        if (2 * (m - t->g_hint[l]) > t->g_bound[l] - t->g_hint[l]) stale = true;
    }
    return stale;
}

int nv_tree_dump(TreeBuf *t, hipStream_t st, nv_tree_node *buf, int cap, int *nout) {
    const int nn = t->nnodes, nc = t->nd.ncap;
    *nout = nn;
    if (!buf || cap < nn) return NV_OK;
    std::vector<int> first(nn), pc(nn), li(nn);
    std::vector<double> pos(3 * (size_t)nc), w(nn), sum(3 * (size_t)nc), circ(3 * (size_t)nc);
    TCHK(hipMemcpyAsync(first.data(), t->nd.first, sizeof(int) * nn, hipMemcpyDeviceToHost, st));
    TCHK(hipMemcpyAsync(pc.data(), t->nd.pcount, sizeof(int) * nn, hipMemcpyDeviceToHost, st));
    TCHK(hipMemcpyAsync(li.data(), t->nd.leafinf, sizeof(int) * nn, hipMemcpyDeviceToHost, st));
    TCHK(hipMemcpyAsync(w.data(), t->nd.w, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
    TCHK(hipMemcpyAsync(pos.data(), t->nd.pos, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, st));
    TCHK(hipMemcpyAsync(sum.data(), t->nd.sum, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, st));
    TCHK(hipMemcpyAsync(circ.data(), t->nd.circ, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, st));
    double ps[4] = {0.0, 0.0, 0.0, 0.0};   // the root also counts the points pinned in front
    if (t->walls) TCHK(hipMemcpyAsync(ps, t->psum, sizeof(ps), hipMemcpyDeviceToHost, st));
    TCHK(hipStreamSynchronize(st));
    int o = 0;
    std::vector<int> stack;
    stack.push_back(0);
    while (!stack.empty()) {
        int q = stack.back();
        stack.pop_back();
        nv_tree_node &r = buf[o++];
        const double cntd = q == 0 ? (double)t->nact_h + ps[3] : (double)pc[q];
        for (int k = 0; k < 3; ++k) {
            r.pos[k] = pos[(size_t)k * nc + q];
            r.cent[k] = pc[q] > 0 ? (sum[(size_t)k * nc + q] + (q == 0 ? ps[k] : 0.0)) / cntd : 0.0;
            r.circ[k] = q == 0 ? 0.0 : circ[(size_t)k * nc + q];
        }
        r.width = w[q];
        r.pcount = pc[q];
        r.leaf_infront = li[q];
        if (first[q] >= 0)
            for (int c = 7; c >= 0; --c) stack.push_back(first[q] + c);
    }
    return NV_OK;
}

// ------------------------------------------------------ solid walls (host)
int nv_tree_walk_mirror(TreeBuf *t, hipStream_t st, DState d, Phys p, int n, int npad, double *mir, DCtl *ctl) {
    if (npad <= 0) return NV_OK;
    hipLaunchKernelGGL(k_walk_mirror, dim3((npad + 63) / 64), dim3(64), 0, st, d, p, n, (const WNode *)t->walk,
                       t->nwalk, mir, npad, ctl, t->counters, (const int *)nullptr, (const int *)nullptr);
    TCHK(hipGetLastError());
    return NV_OK;
}

int nv_tree_close_walls(TreeBuf *t, hipStream_t st, DState d, int n) {
    if (n <= 0) return NV_OK;
    hipLaunchKernelGGL(k_close_walls, dim3((n + 63) / 64), dim3(64), 0, st, d, n, (const WNode *)t->walk, t->nwalk,
                       (const DCtl *)nullptr, (const int *)nullptr, (const int *)nullptr, n);
    TCHK(hipGetLastError());
    return NV_OK;
}

int nv_tree_graph_mirror(TreeBuf *t, hipStream_t st, DState d, Phys p, double *mir, DCtl *ctl, const int *dn) {
    const int npad = t->g_rows;
    hipLaunchKernelGGL(k_walk_mirror, dim3((npad + 63) / 64), dim3(64), 0, st, d, p, npad, (const WNode *)t->walk, 0,
                       mir, npad, ctl, t->counters, dn, (const int *)&t->lc->nwalk);
    TCHK(hipGetLastError());
    return NV_OK;
}

int nv_tree_graph_close_walls(TreeBuf *t, hipStream_t st, DState d, const int *dn, DCtl *ctl) {
    const int npad = t->g_rows;
    hipLaunchKernelGGL(k_close_walls, dim3((npad + 63) / 64), dim3(64), 0, st, d, npad, (const WNode *)t->walk, 0,
                       (const DCtl *)ctl, dn, (const int *)&t->lc->nwalk, npad);
    TCHK(hipGetLastError());
    return NV_OK;
}
