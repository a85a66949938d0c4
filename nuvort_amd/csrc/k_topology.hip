// Per-point boundary kernels, the brute-force close-point search and the
// serial line-topology operations.
//
// pinsert / precon / premove (src/line.f90:18-130, src/reconnection.f90:11-61)
// mutate links in index order and later decisions read earlier mutations,
// so their semantics are serial.  Each runs as a parallel pre-filter (every
// row tests the reference's trigger predicate on the state at entry), a
// stable compaction of the rows that fire, and one device thread that walks
// those rows in ascending order, re-evaluating each against the current
// state.  Rows whose inputs a mutation changes after the pre-filter saw them
// (and that the reference loop has not reached yet) are queued as "dirty"
// and visited in order too, so the visited set always contains every row
// the reference loop would act on.
#include "nv_device.h"
#include "nv_kernels.h"

#include <algorithm>

namespace {

constexpr int DIRTY_MAX = 8192;

__device__ __forceinline__ void clear_row(const DState &s, int r) {
    // src/general.f90:10-20 clear_particle (u_sup, u_mf untouched), the
    // pinned flags included
    const int cap = s.cap;
    if (s.pin) {
        s.pin[r] = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) s.wp[k * cap + r] = 0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s.x[k * cap + r] = 0.0;
        s.u[k * cap + r] = 0.0;
        s.u1[k * cap + r] = 0.0;
        s.u2[k * cap + r] = 0.0;
        s.gi[k * cap + r] = 0.0;
        s.gb[k * cap + r] = 0.0;
    }
    s.infront[r] = 0;
    s.behind[r] = 0;
    s.closest[r] = 0;
    s.closestd[r] = 0.0;
}

// the periodic image of neighbour coordinate g seen from coordinate xi, as
// get_ghost_p does per component (src/boundary.f90:44-83)
__device__ __forceinline__ void wrap_ghost(double xi, double L, double &g) {
    if ((xi - g) > (L / 2.0)) g = g + L;
    else if ((xi - g) < (-L / 2.0)) g = g - L;
}

__device__ __forceinline__ void store_ghosts(const DState &s, const Phys &p, int r) {
    double gi[3], gb[3];
    ghost_of(s, p, r, gi, gb);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s.gi[k * s.cap + r] = gi[k];
        s.gb[k * s.cap + r] = gb[k];
    }
}

// src/reconnection.f90:155-182 loop_killer.  The reference walks the whole
// loop, but its outcome only depends on whether the loop closes within three
// hops (counter < 4); the walk stops as soon as counter reaches 4, which
// gives the same outcome (no pinned points exist on this path).
__device__ void loop_killer(const DState &s, int particle, int n) {
    int counter = 1, next = particle;
    bool closed = false;
    for (int it = 1; it <= n; ++it) {
        next = s.infront[next] - 1;
        if (next == particle) { closed = true; break; }
        if (next < 0) return;
        counter = counter + 1;
        if (counter >= 4) return;
    }
    (void)closed;
    if (counter < 4) {
        next = particle;
        for (int it = 1; it <= n; ++it) {
            int store_next = s.infront[next] - 1;
            clear_row(s, next);
            next = store_next;
            if (next == particle || next < 0) break;
        }
    }
}

struct Dirty {
    int v[DIRTY_MAX];
    int hd, nd;
};

__device__ bool dirty_push(Dirty &d, int val) {
    for (int q = d.hd; q < d.nd; ++q)
        if (d.v[q] == val) return true;
    if (d.nd >= DIRTY_MAX) {
        if (d.hd == 0) return false;
        for (int q = d.hd; q < d.nd; ++q) d.v[q - d.hd] = d.v[q];
        d.nd -= d.hd;
        d.hd = 0;
    }
    int q = d.nd++;
    while (q > d.hd && d.v[q - 1] > val) {
        d.v[q] = d.v[q - 1];
        --q;
    }
    d.v[q] = val;
    return true;
}

// merge cursor over the ascending candidate list and the dirty queue
__device__ __forceinline__ int next_row(const int *cand, int ncand, int &ci, Dirty &d) {
    int rc = ci < ncand ? cand[ci] : 0x7fffffff;
    int rd = d.hd < d.nd ? d.v[d.hd] : 0x7fffffff;
    int r = rc < rd ? rc : rd;
    if (rc == r && r != 0x7fffffff) ++ci;
    if (rd == r && r != 0x7fffffff) ++d.hd;
    return r;
}

// [lo, hi) of the ascending row list `list` (length n) with rows in [a, b)
__device__ __forceinline__ int lower_row(const int *list, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (list[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// --------------------------------------------------------------- per row
__global__ void k_ghosts(DState s, Phys p, int n) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || s.infront[r] == 0) return;
    store_ghosts(s, p, r);
}

// src/boundary.f90:124-135: a pinned point back onto its wall, the first
// component of largest |wpinned| (maxloc)
__device__ __forceinline__ void pin_to_wall(const DState &s, const Phys &p, int r) {
    if (!p.walls || s.pin[r] == 0) return;
    int w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = s.wp[k * s.cap + r];
    int pc = 0;
    if (abs(w[1]) > abs(w[pc])) pc = 1;
    if (abs(w[2]) > abs(w[pc])) pc = 2;
    if (w[pc] > 0) s.x[pc * s.cap + r] = 0.499999999 * p.box[pc];
    else if (w[pc] < 0) s.x[pc * s.cap + r] = -0.499999999 * p.box[pc];
}

__global__ void k_wrap(DState s, Phys p, int n) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || s.infront[r] == 0) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (!p.periodic[k]) continue;
        double L = p.box[k];
        double x = s.x[k * s.cap + r];
        if (x > (L / 2.0)) x = x - L;
        else if (x < (-L / 2.0)) x = x + L;
        s.x[k * s.cap + r] = x;
    }
    pin_to_wall(s, p, r);
}

__global__ void k_nan(DState s, int n, DCtl *ctl) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    if (isnan(s.x[r]) || isnan(s.x[s.cap + r]) || isnan(s.x[2 * s.cap + r])) atomicOr(&ctl->status, NVK_NAN);
}

// ------------------------------------------------------------ diagnostics
// src/diagnostic.f90:21-74 for rows [0, n): each row's curvature |s''|
// (curv_info, general.f90:54-61) and stretching u . s'' (stretching_rate),
// zero for empty rows, and per workgroup the maxima of |u| and |u1 - u2|,
// the sums of |u_sup|, |u_mf| and curv, the smallest positive and the largest
// curvature and the active count (velocity_info, curv_info).
constexpr int DIAG_VALS = 8;

__global__ __launch_bounds__(256) void k_diag(DState s, int n, double *part) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int cap = s.cap;
    double v[DIAG_VALS] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0e300, 0.0, 0.0};
    if (r < n) {
        double u[3], u1[3], u2[3], us[3], um[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u[k] = ld(s.u, cap, k, r);
            u1[k] = ld(s.u1, cap, k, r);
            u2[k] = ld(s.u2, cap, k, r);
            us[k] = ld(s.usup, cap, k, r);
            um[k] = ld(s.umf, cap, k, r);
        }
        v[0] = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double d0 = u1[0] - u2[0], d1 = u1[1] - u2[1], d2 = u1[2] - u2[2];
        v[1] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        v[2] = sqrt(us[0] * us[0] + us[1] * us[1] + us[2] * us[2]);
        v[3] = sqrt(um[0] * um[0] + um[1] * um[1] + um[2] * um[2]);
        double curv = 0.0, stretch = 0.0;
        if (s.infront[r] != 0) {
            double x[3], gi[3], gb[3], fd[3], fdd[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = ld(s.x, cap, k, r);
                gi[k] = ld(s.gi, cap, k, r);
                gb[k] = ld(s.gb, cap, k, r);
            }
            derivs(x, gi, gb, fd, fdd);
            curv = sqrt(fdd[0] * fdd[0] + fdd[1] * fdd[1] + fdd[2] * fdd[2]);
            stretch = u[0] * fdd[0] + u[1] * fdd[1] + u[2] * fdd[2];
            v[7] = 1.0;
        }
        s.curv[r] = curv;
        s.stretch[r] = stretch;
        v[4] = curv;
        if (curv > 0.0) v[5] = curv;
        v[6] = curv;
    }
    // workgroup reduction: maxima, sums, minimum, maximum, count
    __shared__ double red[DIAG_VALS][256];
#pragma unroll
    for (int k = 0; k < DIAG_VALS; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const int o = threadIdx.x + w;
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][o]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][o]);
            red[2][threadIdx.x] += red[2][o];
            red[3][threadIdx.x] += red[3][o];
            red[4][threadIdx.x] += red[4][o];
            red[5][threadIdx.x] = fmin(red[5][threadIdx.x], red[5][o]);
            red[6][threadIdx.x] = fmax(red[6][threadIdx.x], red[6][o]);
            red[7][threadIdx.x] += red[7][o];
        }
        __syncthreads();
    }
    if (threadIdx.x < DIAG_VALS) part[(size_t)blockIdx.x * DIAG_VALS + threadIdx.x] = red[threadIdx.x][0];
}

// the workgroups' partials reduced in a fixed order by one workgroup
__global__ __launch_bounds__(256) void k_diag_fin(const double *part, int nblk, DDiag *out) {
    double v[DIAG_VALS] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0e300, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += 256) {
        const double *q = part + (size_t)b * DIAG_VALS;
        v[0] = fmax(v[0], q[0]);
        v[1] = fmax(v[1], q[1]);
        v[2] += q[2];
        v[3] += q[3];
        v[4] += q[4];
        v[5] = fmin(v[5], q[5]);
        v[6] = fmax(v[6], q[6]);
        v[7] += q[7];
    }
    __shared__ double red[DIAG_VALS][256];
#pragma unroll
    for (int k = 0; k < DIAG_VALS; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const int o = threadIdx.x + w;
            red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][o]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][o]);
            red[2][threadIdx.x] += red[2][o];
            red[3][threadIdx.x] += red[3][o];
            red[4][threadIdx.x] += red[4][o];
            red[5][threadIdx.x] = fmin(red[5][threadIdx.x], red[5][o]);
            red[6][threadIdx.x] = fmax(red[6][threadIdx.x], red[6][o]);
            red[7][threadIdx.x] += red[7][o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out->maxu = red[0][0];
        out->maxdu = red[1][0];
        out->sum_usup = red[2][0];
        out->sum_umf = red[3][0];
        out->sum_curv = red[4][0];
        out->curv_min = red[5][0];
        out->curv_max = red[6][0];
        out->active = (int)red[7][0];
    }
}

// ---------------------------------------------------------------- pinsert
__global__ void k_pinsert_pre(DState s, Phys p, Tangles tg, int n, int *cand, int *empty, DCtl *ctl) {
    if (ctl->abort) return;   // (a recorded octree step whose build was incomplete: as k_finish)
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = r < n && s.infront[r] != 0;
    if (r < n) {
        // a free slot of the reference's f(1:pcount): empty and below pcount
        const int k = tg.row2t[r];
        empty[r] = (!act && r - tg.base[k] < tg.n[k]) ? 1 : 0;
    }
    int c = 0;
    double disti = 0.0;
    if (act) {
        disti = dist_gen(s.x[r], s.x[s.cap + r], s.x[2 * s.cap + r], s.gi[r], s.gi[s.cap + r],
                         s.gi[2 * s.cap + r]);
        c = disti > p.delta ? 1 : 0;
    }
    if (r < n) cand[r] = c;
    // src/line.f90:38-39 total_length (a diagnostic): the sum of disti over
    // the points present when pinsert starts, one atomic per workgroup
    __shared__ double wsum[4];
    double v = disti;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
        if (b != 0.0) atomicAdd(&ctl->total_length, b);
    }
}

// Inputs of one insertion at row r (src/line.f90:30-91): r's position and
// ghosts, its links, its infront's position, infront and velocity, its
// behind's position, and its own velocity.
struct InsIn {
    double x[3], gi[3], gb[3], xf[3], xff[3], xb[3], ur[3], uf[3];
    int fi, bb, act;
};

__device__ __forceinline__ void load_insertion(const DState &s, int r, InsIn &in) {
    const int cap = s.cap;
    in.act = s.infront[r] != 0;
    in.fi = in.bb = r;
    if (!in.act) return;
    in.fi = s.infront[r] - 1;
    in.bb = s.behind[r] - 1;
    const int ffi = s.infront[in.fi] - 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        in.x[k] = s.x[k * cap + r];
        in.gi[k] = s.gi[k * cap + r];
        in.gb[k] = s.gb[k * cap + r];
        in.xf[k] = s.x[k * cap + in.fi];
        in.xff[k] = s.x[k * cap + ffi];
        in.xb[k] = s.x[k * cap + in.bb];
        in.ur[k] = s.u[k * cap + r];
        in.uf[k] = s.u[k * cap + in.fi];
    }
}

// What an insertion leaves behind for later visits: the new point and the
// refreshed infront ghosts of the new point and of r's old infront.
struct InsOut {
    double xn[3], gs_i[3], gf_i[3], gf_b[3];
};

// The arithmetic of one visit (src/line.f90:30-91), which depends on the
// visit's inputs only: whether r inserts, the new point, its velocity, and
// the ghosts of the new point, of r and of r's old infront fi from the
// positions the insertion leaves in place, as store_ghosts
// (src/boundary.f90:26-84) computes them (fi's infront and r's behind are
// unchanged by the insertion).
struct InsPlan {
    double xn[3], un[3], gs_i[3], gs_b[3], gr_i[3], gr_b[3], gf_i[3], gf_b[3];
    int go;
};

__device__ void plan_insertion(const Phys &p, const InsIn &in, InsPlan &pl) {
    pl.go = 0;
    if (!in.act) return;
    const double disti = dist_gen(in.x[0], in.x[1], in.x[2], in.gi[0], in.gi[1], in.gi[2]);
    if (!(disti > p.delta)) return;
    pl.go = 1;
    double fd[3], fdd[3];
    derivs(in.x, in.gi, in.gb, fd, fdd);
    double curv = sqrt(fdd[0] * fdd[0] + fdd[1] * fdd[1] + fdd[2] * fdd[2]);
    double xn[3];
    if (curv > 1e-5) {
        curv = 1.0 / curv;
        if (curv * curv - 0.25 * (disti * disti) > 0.0) {
            const double rr = (sqrt(curv * curv - 0.25 * (disti * disti)) - curv) * curv;
#pragma unroll
            for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (in.x[k] + in.gi[k]) + rr * fdd[k];
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (in.x[k] + in.gi[k]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (in.x[k] + in.gi[k]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pl.xn[k] = xn[k];
        pl.un[k] = 0.5 * (in.ur[k] + in.uf[k]);
        pl.gs_i[k] = in.xf[k]; pl.gs_b[k] = in.x[k];
        pl.gr_i[k] = xn[k]; pl.gr_b[k] = in.xb[k];
        pl.gf_i[k] = in.xff[k]; pl.gf_b[k] = xn[k];
        if (!p.periodic[k]) continue;
        const double L = p.box[k];
        wrap_ghost(xn[k], L, pl.gs_i[k]);
        wrap_ghost(xn[k], L, pl.gs_b[k]);
        wrap_ghost(in.x[k], L, pl.gr_i[k]);
        wrap_ghost(in.x[k], L, pl.gr_b[k]);
        wrap_ghost(in.xf[k], L, pl.gf_i[k]);
        wrap_ghost(in.xf[k], L, pl.gf_b[k]);
    }
}

// One visit of the reference loop at row r from its plan: returns the slot
// used (or -1 when r does not insert); -2 when the tangle's rows are
// exhausted.
// The stores of a planned insertion at r (old infront fi) into slot; r's
// and fi's ghosts unless a later visit stores them again.
__device__ __forceinline__ void write_insertion(const DState &s, int r, int fi, const InsPlan &pl, int slot,
                                                bool ghosts_r = true, bool ghosts_fi = true) {
    const int cap = s.cap;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s.x[k * cap + slot] = pl.xn[k];
        s.u[k * cap + slot] = pl.un[k];
        s.u1[k * cap + slot] = 0.0;
        s.u2[k * cap + slot] = 0.0;
        s.gi[k * cap + slot] = pl.gs_i[k];
        s.gb[k * cap + slot] = pl.gs_b[k];
        // fi before r, the reference's refresh order (they coincide only for
        // a one-point loop)
        if (ghosts_fi) {
            s.gi[k * cap + fi] = pl.gf_i[k];
            s.gb[k * cap + fi] = pl.gf_b[k];
        }
        if (ghosts_r) {
            s.gi[k * cap + r] = pl.gr_i[k];
            s.gb[k * cap + r] = pl.gr_b[k];
        }
    }
    s.behind[slot] = r + 1;
    s.infront[slot] = fi + 1;
    s.behind[fi] = slot + 1;
    s.infront[r] = slot + 1;
}

// the slot of a grown row, as the reference's new array element: zeroed
__device__ __forceinline__ void clear_grown(const DState &s, int slot) {
    const int cap = s.cap;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s.usup[k * cap + slot] = 0.0;
        s.umf[k * cap + slot] = 0.0;
    }
    clear_row(s, slot);
}

__device__ int insert_at(const DState &s, int r, const InsIn &in, const InsPlan &pl, const int *empty, int &ei,
                         int e1, int &nl, int base, int tcap, InsOut &out, const int *est, int es0, int es1) {
    const int cap = s.cap;
    if (!pl.go) return -1;
    int slot;
    if (ei < e1) {
        // the empty list's entries [es0, es1) are staged in est
        slot = (ei >= es0 && ei < es1) ? est[ei - es0] : empty[ei];
        ++ei;
    } else {
        if (nl >= tcap) return -2;
        slot = base + nl;
        ++nl;
        clear_grown(s, slot);
    }
    write_insertion(s, r, in.fi, pl, slot);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        out.xn[k] = pl.xn[k];
        out.gs_i[k] = pl.gs_i[k];
        out.gf_i[k] = pl.gf_i[k];
        out.gf_b[k] = pl.gf_b[k];
    }
    return slot;
}

__device__ __forceinline__ double dist2_gen(const double a[3], const double b[3]) {
    const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

// First index q in [lo, hi) with v[q] >= key (v ascending), hi if none.
__device__ __forceinline__ int lower_index(const int *v, int lo, int hi, int key) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Index of row v among the staged rows not yet visited, strow[k, m)
// (ascending), or m if it is not one of them: most rows looked up lie
// before the next staged row or after the last, which two reads settle.
__device__ __forceinline__ int staged_index(const int *strow, int k, int m, int v) {
    if (k >= m) return m;
    const int first = strow[k];
    if (v <= first) return v == first ? k : m;
    if (v > strow[m - 1]) return m;
    const int q = lower_index(strow, k + 1, m, v);
    return (q < m && strow[q] == v) ? q : m;
}

// After an insertion at r (old infront fi, behind bb, new point in slot):
// the staged inputs of later candidates fi and bb take the new state in
// place (fi: ghosts and behind; bb: its infront's infront is the new
// point), and a row the reference loop will revisit is queued only when
// that visit inserts -- a revisit tests |x - ghost_infront| > delta, and the
// ghost it will see is the one just computed (a row's infront ghost changes
// only when the row itself inserts).
// The new point in a slot above r and below the loop's end is visited by
// the reference loop too: its |x - ghost_infront| (fixed from now on) joins
// total_length (src/line.f90:38-39), which the pre-filter summed only over
// the points present at the start; `len` collects it.
__device__ bool after_insert(const DState &s, const Phys &p, int r, const InsIn &in, int slot, const InsOut &o,
                             InsIn *st, unsigned char *stale, const int *strow, int k, int m, Dirty &d, int n_old,
                             double &len) {
    const int fi = in.fi, bb = in.bb;
    if (slot > r && slot < n_old) len += sqrt(dist2_gen(o.xn, o.gs_i));
    // the staged rows are ascending and those not yet visited are above r:
    // find fi and bb among them (usually fi is the next one; either can be
    // far from r, e.g. a new slot); their plans are then stale
    const int qf = fi <= r ? m : staged_index(strow, k, m, fi);
    if (qf < m && strow[qf] == fi) {
        stale[qf] = 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            st[qf].gi[c] = o.gf_i[c];
            st[qf].gb[c] = o.gf_b[c];
            st[qf].xb[c] = o.xn[c];
        }
        st[qf].bb = slot;
    }
    const int qb = bb <= r ? m : staged_index(strow, k, m, bb);
    if (qb < m && strow[qb] == bb) {
        stale[qb] = 1;
#pragma unroll
        for (int c = 0; c < 3; ++c) st[qb].xff[c] = o.xn[c];
    }
    // A row the loop has still to reach is queued when its visit may insert;
    // a staged candidate is visited anyway.  The test is on squared
    // distances with a margin, so it queues a superset of the rows that
    // insert (a visit that does not insert changes nothing).
    const double lim2 = p.delta * p.delta * (1.0 - 1.0e-9);
    if (fi > r && fi < n_old && !(qf < m && strow[qf] == fi) &&
        dist2_gen(in.xf, o.gf_i) > lim2 && !dirty_push(d, fi))
        return false;
    if (slot > r && slot < n_old && dist2_gen(o.xn, o.gs_i) > lim2 && !dirty_push(d, slot))
        return false;
    return true;
}

// the reference's loops, which the staged kernels hand over to when their
// queue of rows to revisit is full (defined with the solid-wall kernels)
__device__ int pinsert_loop(const DState &s, const Phys &p, const Tangles &tg, DCtl *ctl, int tk, int from, int end,
                            bool len_at_visit);
__device__ void premove_loop(const DState &s, const Phys &p, const Tangles &tg, int tk, int from);

constexpr int INS_CHUNK = 128;

// One workgroup per tangle.  Candidates are taken INS_CHUNK at a time: every
// thread loads one candidate's inputs into LDS, then thread 0 makes the
// reference's visits in row order (candidates merged with the dirty queue);
// after_insert keeps the staged inputs current.
__global__ __launch_bounds__(INS_CHUNK) void k_pinsert_serial(DState s, Phys p, Tangles tg, const int *cand,
                                                              const int *empty, DCtl *ctl) {
    if (ctl->abort) return;
    __shared__ Dirty d;
    __shared__ InsIn st[INS_CHUNK];
    __shared__ InsPlan pl[INS_CHUNK];
    __shared__ unsigned char stale[INS_CHUNK];   // the staged inputs changed after the plan
    __shared__ int strow[INS_CHUNK];
    __shared__ int nl_sh, stop_sh, ei_sh, res_sh;
    __shared__ int est[INS_CHUNK];   // the next entries of the empty list
    __shared__ unsigned long long gomask[INS_CHUNK / 64];
    __shared__ double lenv[INS_CHUNK];
    __shared__ unsigned char lenq[INS_CHUNK];
    __shared__ unsigned char fin[INS_CHUNK];     // the plan is final
    __shared__ int skipg[INS_CHUNK];             // ghosts a later visit stores again (1: r's, 2: fi's)
    __shared__ int slotv[INS_CHUNK];
    __shared__ int pushv[2 * INS_CHUNK];          // revisits queued after a segment's stores
    __shared__ int a_sh, vmin_sh, cut_sh;         // the segment's first candidate, first queued row, end
    __shared__ unsigned char rel[INS_CHUNK];      // inputs to load again after a segment
    __shared__ unsigned long long lenmask[INS_CHUNK / 64], pushmask[INS_CHUNK / 64];
    const int tid = threadIdx.x;
    const int tk = blockIdx.x;
    const int base = tg.base[tk], tcap = tg.cap[tk];
    const int n_old = base + tg.n[tk];     // rows the reference loop visits end here
    const int c0 = lower_row(cand, ctl->ncand, base), c1 = lower_row(cand, ctl->ncand, base + tcap);
    const int e0 = lower_row(empty, ctl->nempty, base), e1 = lower_row(empty, ctl->nempty, base + tcap);
    // src/line.f90:25-27: no active point in this tangle (active points =
    // pcount less the empty slots below it)
    if (tg.n[tk] - (e1 - e0) == 0) {
        if (tid == 0) atomicOr(&ctl->status, NVK_ZERO_LENGTH);
        return;
    }
    const int *cl = cand + c0;
    const int ncand = c1 - c0;
    if (tid == 0) {
        d.hd = d.nd = 0;
        nl_sh = tg.n[tk];
        stop_sh = 0;
        ei_sh = e0;
        res_sh = -1;
    }
    int ei = e0;   // used by thread 0 only
    double len = 0.0;   // thread 0: total_length of the points inserted above the cursor
    for (int cs = 0; cs <= ncand; cs += INS_CHUNK) {
        const int m = min(INS_CHUNK, ncand - cs);
        __syncthreads();
        const int es0 = ei_sh, es1 = min(e1, es0 + INS_CHUNK);
        if (tid < m) {
            const int r = cl[cs + tid];
            strow[tid] = r;
            load_insertion(s, r, st[tid]);
            plan_insertion(p, st[tid], pl[tid]);
            stale[tid] = 0;
        }
        if (tid < INS_CHUNK) rel[tid] = 0;
        if (es0 + tid < es1) est[tid] = empty[es0 + tid];
        __syncthreads();
        const int limit = cs + m < ncand ? cl[cs + m] : 0x7fffffff;
        if (tid == 0) a_sh = 0;
        // The chunk in segments of candidates whose visits are made at once,
        // between visits made one at a time.
        //
        // Whether a candidate inserts depends on its own position and its
        // infront's ghost, which no other insertion moves; so, while no row
        // queued for a revisit comes first, the visits of candidates a, a+1,
        // ... insert where their plans say, and the k-th insertion takes the
        // next empty slot but k, then new rows.  An insertion at r changes
        // the inputs of its old infront fi and behind bb (fi: ghosts and
        // behind; bb: its infront's infront): a later candidate that is fi or
        // bb of an earlier insertion of the segment is planned again, in
        // rounds, once that insertion is final.  Where two insertions store
        // the ghosts of the same row (r's own, and those of another
        // insertion's fi), the later visit's stores are kept.  The segment
        // ends before the first queued row that is not a candidate, before a
        // candidate that finds no room or whose plan is not final after the
        // rounds, and after a candidate that queues a revisit of a row of the
        // chunk; revisits of later rows are queued after its stores.  The
        // serial thread then makes the visits one at a time until the next
        // visit is a candidate again.
        while (true) {
            __syncthreads();
            const int a = a_sh;
            if (stop_sh) break;
            if (tid == 0) {
                vmin_sh = 0x7fffffff;
                cut_sh = m;
            }
            if (tid < INS_CHUNK) {
                lenq[tid] = 0;
                skipg[tid] = 0;
                pushv[2 * tid] = pushv[2 * tid + 1] = -1;
            }
            __syncthreads();
            // the first queued row of the chunk's rows that is not one of its candidates
            for (int q = d.hd + tid; q < d.nd && d.v[q] < limit; q += INS_CHUNK)
                if (staged_index(strow, a, m, d.v[q]) == m) atomicMin(&vmin_sh, d.v[q]);
            const bool in_seg0 = tid >= a && tid < m;
            if (in_seg0 && stale[tid]) {
                plan_insertion(p, st[tid], pl[tid]);
                stale[tid] = 0;
            }
            __syncthreads();
            const int vmin = vmin_sh;
            const int hi = vmin == 0x7fffffff ? m : lower_index(strow, a, m, vmin);
            const bool in_seg = tid >= a && tid < hi;
            const bool go = in_seg && pl[tid].go;
            const unsigned long long bm = __ballot(go);
            const int lane = tid & 63;
            if (lane == 0) gomask[tid >> 6] = bm;
            __syncthreads();
            int kq = __popcll(bm & ((1ull << lane) - 1ull));
#pragma unroll
            for (int w = 0; w < INS_CHUNK / 64; ++w)
                if (w < (tid >> 6)) kq += __popcll(gomask[w]);
            const int e_next = ei_sh, n_next = nl_sh;
            bool stop_here = false;   // this candidate cannot be visited in the segment
            int slot = -1, r = -1, predb = -1, predf = -1;
            if (go) {
                r = strow[tid];
                const int e = e_next + kq;
                if (e < e1) {
                    slot = (e >= es0 && e < es1) ? est[e - es0] : empty[e];
                } else {
                    const int a2 = e - e1;
                    if (n_next + a2 >= tcap) stop_here = true;
                    else slot = base + n_next + a2;
                }
                // earlier insertions of the segment whose fi (predb) or bb (predf) this row is
                const int bb = st[tid].bb, fi = st[tid].fi;
                const int qb = bb < r ? staged_index(strow, a, tid, bb) : tid;
                if (qb < tid && pl[qb].go) predb = qb;
                const int qf = fi < r ? staged_index(strow, a, tid, fi) : tid;
                if (qf < tid && pl[qf].go) predf = qf;
            }
            slotv[tid] = slot;
            fin[tid] = !(predb >= 0 || predf >= 0);
            // plan again, in rounds, the insertions whose inputs earlier ones change
            bool more = __syncthreads_or(!fin[tid]);
            for (int it = 0; more && it < 16; ++it) {
                const bool ready = !fin[tid] && (predb < 0 || fin[predb]) && (predf < 0 || fin[predf]);
                __syncthreads();
                if (ready) {
                    if (predb >= 0) {   // this row is predb's old infront
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            st[tid].gi[c] = pl[predb].gf_i[c];
                            st[tid].gb[c] = pl[predb].gf_b[c];
                            st[tid].xb[c] = pl[predb].xn[c];
                        }
                        st[tid].bb = slotv[predb];
                    }
                    if (predf >= 0) {   // this row is predf's old behind
#pragma unroll
                        for (int c = 0; c < 3; ++c) st[tid].xff[c] = pl[predf].xn[c];
                    }
                    plan_insertion(p, st[tid], pl[tid]);
                }
                __syncthreads();
                if (ready) fin[tid] = 1;
                more = __syncthreads_or(!fin[tid]);
            }
            stop_here = stop_here || !fin[tid];
            // revisits (after_insert's)
            int pushf = -1, pushs = -1;
            if (go && !stop_here) {
                const int fi = st[tid].fi;
                const double lim2 = p.delta * p.delta * (1.0 - 1.0e-9);
                if (fi > r && fi < n_old && staged_index(strow, tid + 1, m, fi) == m &&
                    dist2_gen(st[tid].xf, pl[tid].gf_i) > lim2)
                    pushf = fi;
                if (slot > r && slot < n_old && dist2_gen(pl[tid].xn, pl[tid].gs_i) > lim2) pushs = slot;
            }
            if (in_seg && stop_here) atomicMin(&cut_sh, tid);
            if ((pushf >= 0 && pushf < limit) || (pushs >= 0 && pushs < limit)) atomicMin(&cut_sh, tid + 1);
            __syncthreads();
            const int seg_end = min(cut_sh, hi);
            const bool mine = go && tid < seg_end;
            if (mine) {
                if (predf >= 0) atomicOr(&skipg[predf], 1);   // predf's own ghosts, stored again by this visit
                if (predb >= 0) atomicOr(&skipg[predb], 2);   // predb's fi ghosts: this row's own
                pushv[2 * tid] = pushf;
                pushv[2 * tid + 1] = pushs;
            }
            __syncthreads();
            if (mine) {
                if (e_next + kq >= e1) clear_grown(s, slot);
                const int fi = st[tid].fi, bb = st[tid].bb;
                write_insertion(s, r, fi, pl[tid], slot, !(skipg[tid] & 1), !(skipg[tid] & 2));
                if (slot > r && slot < n_old) {
                    lenq[tid] = 1;
                    lenv[tid] = sqrt(dist2_gen(pl[tid].xn, pl[tid].gs_i));
                }
                // candidates after the segment whose inputs this insertion changes
                const int qf = fi > r ? staged_index(strow, seg_end, m, fi) : m;
                if (qf < m) rel[qf] = 1;
                const int qb = bb > r ? staged_index(strow, seg_end, m, bb) : m;
                if (qb < m) rel[qb] = 1;
            }
            // and those the rounds planned again on insertions not made
            if (tid >= seg_end && tid < hi && (predb >= 0 || predf >= 0)) rel[tid] = 1;
            // per wave: the segment's insertions, those whose length counts,
            // those that queue revisits
            // (ballots of the whole wave, outside any branch)
            const unsigned long long mins = __ballot(mine);
            const unsigned long long mlen = __ballot(lenq[tid] != 0);
            const unsigned long long mpush = __ballot(mine && (pushf >= 0 || pushs >= 0));
            if (lane == 0) {
                gomask[tid >> 6] = (unsigned long long)__popcll(mins);
                lenmask[tid >> 6] = mlen;
                pushmask[tid >> 6] = mpush;
            }
            __syncthreads();
            // their inputs from memory, which now holds the segment's insertions
            if (tid < m && rel[tid]) {
                load_insertion(s, strow[tid], st[tid]);
                stale[tid] = 1;
                rel[tid] = 0;
            }
            if (tid == 0) {
                int total = 0;
#pragma unroll
                for (int w = 0; w < INS_CHUNK / 64; ++w) total += (int)gomask[w];
                // queued rows visited as candidates of the segment
                if (seg_end > a)
                    while (d.hd < d.nd && d.v[d.hd] <= strow[seg_end - 1]) ++d.hd;
                // total_length in visit order
                for (int w = 0; w < INS_CHUNK / 64; ++w)
                    for (unsigned long long b = lenmask[w]; b; b &= b - 1) len += lenv[64 * w + __ffsll(b) - 1];
                const int from_list = min(total, max(0, e1 - e_next));
                ei = e_next + from_list;
                ei_sh = ei;
                nl_sh = n_next + (total - from_list);
                bool room = true;
                for (int w = 0; w < INS_CHUNK / 64 && room; ++w)
                    for (unsigned long long b = pushmask[w]; b && room; b &= b - 1) {
                        const int q = 64 * w + __ffsll(b) - 1;
                        room = (pushv[2 * q] < 0 || dirty_push(d, pushv[2 * q])) &&
                               (pushv[2 * q + 1] < 0 || dirty_push(d, pushv[2 * q + 1]));
                    }
                if (!room) {
                    // no room: the plain loop takes over after the segment
                    res_sh = strow[seg_end - 1] + 1;
                    stop_sh = 1;
                }
                a_sh = seg_end;
            }
            __syncthreads();
            if (stop_sh) break;
            const int a2 = a_sh;
            const int rd0 = d.hd < d.nd ? d.v[d.hd] : 0x7fffffff;
            if (a2 >= m && rd0 >= limit) break;   // the chunk is done
            const bool serial = a2 == a || (rd0 < limit && (a2 >= m || rd0 < strow[a2]));
            if (!serial) continue;
            if (tid == 0) {
                // one visit at a time, until the next visit is a candidate
                int nl = nl_sh;
                int k = a2;
                bool visited = false;
                while (true) {
                    const int rc = k < m ? strow[k] : 0x7fffffff;
                    const int rd = d.hd < d.nd ? d.v[d.hd] : 0x7fffffff;
                    const int r = rc < rd ? rc : rd;
                    if (r == 0x7fffffff || r >= limit) break;
                    if (visited && rc <= rd) break;
                    visited = true;
                    InsIn in;
                    InsPlan plan;
                    const bool staged = rc == r;
                    if (rc == r) ++k;
                    if (rd == r) ++d.hd;
                    if (staged && stale[k - 1]) {
                        in = st[k - 1];
                        plan_insertion(p, in, plan);
                    } else if (staged) {
                        // a planned visit: insert_at and after_insert read only
                        // these of its inputs
                        in.fi = st[k - 1].fi;
                        in.bb = st[k - 1].bb;
#pragma unroll
                        for (int c = 0; c < 3; ++c) in.xf[c] = st[k - 1].xf[c];
                        plan = pl[k - 1];
                    } else {
                        load_insertion(s, r, in);
                        plan_insertion(p, in, plan);
                    }
                    InsOut o;
                    const int slot2 = insert_at(s, r, in, plan, empty, ei, e1, nl, base, tcap, o, est, es0, es1);
                    if (slot2 == -2) { atomicOr(&ctl->status, NVK_CAPACITY); stop_sh = 1; break; }
                    if (slot2 < 0) continue;
                    if (!after_insert(s, p, r, in, slot2, o, st, stale, strow, k, m, d, n_old, len)) {
                        // no room for a row to revisit: the plain loop takes
                        // over after r
                        res_sh = r + 1;
                        stop_sh = 1;
                        break;
                    }
                }
                nl_sh = nl;
                ei_sh = ei;
                a_sh = k;
            }
        }
        if (cs + m >= ncand) {
            // the last chunk: dirty rows beyond it are visited from memory
            __syncthreads();
            if (tid == 0 && !stop_sh) {
                int nl = nl_sh;
                while (d.hd < d.nd) {
                    const int r = d.v[d.hd++];
                    InsIn in;
                    InsPlan plan;
                    load_insertion(s, r, in);
                    plan_insertion(p, in, plan);
                    InsOut o;
                    const int slot = insert_at(s, r, in, plan, empty, ei, e1, nl, base, tcap, o, est, es0, es1);
                    if (slot == -2) { atomicOr(&ctl->status, NVK_CAPACITY); break; }
                    if (slot < 0) continue;
                    if (!after_insert(s, p, r, in, slot, o, st, stale, strow, 0, 0, d, n_old, len)) {
                        res_sh = r + 1;
                        break;
                    }
                }
                nl_sh = nl;
            }
            break;
        }
    }
    __syncthreads();
    if (tid == 0) {
        // avg_sep's denominator (src/line.f90:95): pcount at the start less
        // the empty slots left at the end
        atomicAdd(&ctl->n_old, tg.n[tk]);
        if (len != 0.0) atomicAdd(&ctl->total_length, len);
        tg.n[tk] = nl_sh;
        if (res_sh < 0)
            atomicAdd(&ctl->empty_after, e1 - ei);
        else   // the rest of the loop as the reference writes it (it counts the empty slots at its end)
            pinsert_loop(s, p, tg, ctl, tk, res_sh, n_old, false);
    }
}

// ----------------------------------------------------------------- precon
__global__ void k_precon_pre(DState s, Phys p, int n, int *cand) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double cd = s.closestd[r];
    cand[r] = (s.infront[r] != 0 && (cd < p.delta / 2.0) && (cd > NV_EPS)) ? 1 : 0;
}

// src/general.f90:106-117 norm_tanf
__device__ __forceinline__ void norm_tanf(const DState &s, int r, double t[3]) {
    double x[3], gi[3], gb[3], fdd[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = s.x[k * s.cap + r];
        gi[k] = s.gi[k * s.cap + r];
        gb[k] = s.gb[k * s.cap + r];
    }
    derivs(x, gi, gb, t, fdd);
    double length = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = t[k] / length;
}

__device__ __forceinline__ double dist_rows(const DState &s, const double *a, int ra, const double *b, int rb) {
    return dist_gen(s.x[ra], s.x[s.cap + ra], s.x[2 * s.cap + ra], b[rb], b[s.cap + rb], b[2 * s.cap + rb]);
}

// Inputs of one precon visit at row r (src/reconnection.f90:18-54).
struct ReconIn {
    double x[3], gi[3], gb[3];       // r
    double xj[3], gij[3], gbj[3];    // its closest point rj
    double xpi[3], xpjb[3];          // infront(r) and behind(rj)
    double cd;
    int act, pari, behr, j, actj, parjb;
    // the visit's tests, evaluated on these inputs
    double dot, lb, la;
    int go;
};

__device__ __forceinline__ void load_recon(const DState &s, int r, ReconIn &in) {
    const int cap = s.cap;
    in.act = s.infront[r] != 0;
    in.pari = s.infront[r];
    in.behr = s.behind[r];
    in.cd = s.closestd[r];
    in.j = s.closest[r];
    in.actj = 0;
    in.parjb = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        in.x[k] = s.x[k * cap + r];
        in.gi[k] = s.gi[k * cap + r];
        in.gb[k] = s.gb[k * cap + r];
    }
    if (in.j <= 0 || !in.act) return;
    const int rj = in.j - 1;
    in.actj = s.infront[rj] != 0;
    in.parjb = s.behind[rj];
    const int pi = in.pari > 0 ? in.pari - 1 : r, pjb = in.parjb > 0 ? in.parjb - 1 : rj;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        in.xj[k] = s.x[k * cap + rj];
        in.gij[k] = s.gi[k * cap + rj];
        in.gbj[k] = s.gb[k * cap + rj];
        in.xpi[k] = s.x[k * cap + pi];
        in.xpjb[k] = s.x[k * cap + pjb];
    }
}

// src/general.f90:106-117 norm_tanf from a point's position and ghosts
__device__ __forceinline__ void norm_tanf_of(const double x[3], const double gi[3], const double gb[3], double t[3]) {
    double fdd[3];
    derivs(x, gi, gb, t, fdd);
    const double length = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = t[k] / length;
}

// LDS hash size of premove's changed-row table (below)
constexpr int MOD_SLOTS = 4096;

// src/reconnection.f90:20-43: the tests of one visit; go = reconnect
__device__ void recon_tests(const Phys &p, ReconIn &in) {
    in.go = 0;
    if (!in.act) return;
    const double cd = in.cd;
    if (!((cd < p.delta / 2.0) && (cd > NV_EPS))) return;
    const int j = in.j;
    if (j == in.pari) return;
    if (j == in.behr) return;
    if (j == 0) return;
    if (!in.actj) return;
    double t1[3], t2[3];
    norm_tanf_of(in.x, in.gi, in.gb, t1);
    norm_tanf_of(in.xj, in.gij, in.gbj, t2);
    in.dot = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
    if (in.dot > 0.9) return;
    in.lb = dist_gen(in.x[0], in.x[1], in.x[2], in.gi[0], in.gi[1], in.gi[2]) +
            dist_gen(in.x[0], in.x[1], in.x[2], in.gb[0], in.gb[1], in.gb[2]) +
            dist_gen(in.xj[0], in.xj[1], in.xj[2], in.gij[0], in.gij[1], in.gij[2]) +
            dist_gen(in.xj[0], in.xj[1], in.xj[2], in.gbj[0], in.gbj[1], in.gbj[2]);
    in.la = dist_gen(in.x[0], in.x[1], in.x[2], in.gb[0], in.gb[1], in.gb[2]) +
            dist_gen(in.x[0], in.x[1], in.x[2], in.xj[0], in.xj[1], in.xj[2]) +
            dist_gen(in.xj[0], in.xj[1], in.xj[2], in.gij[0], in.gij[1], in.gij[2]) +
            dist_gen(in.xpi[0], in.xpi[1], in.xpi[2], in.xpjb[0], in.xpjb[1], in.xpjb[2]);
    in.go = in.la <= in.lb;
}

constexpr int REC_CHUNK = 256;
constexpr int REC_PARALLEL_MAX = 16;

// What one visit of the staged precon reads, loaded for a whole chunk of
// candidates at once.  The parts of the reference's tests that no
// reconnection can change are evaluated at staging: the tangents' dot
// product, l_before, and the first three terms of l_after (positions and
// stale ghosts of r and its closest point j are not written by precon,
// except by clearing, which empties r or j and ends the visit early).
// What a reconnection earlier in the chunk can change -- the links of r and
// j, and so infront(r), behind(j) and their positions -- is read through
// the chunk's overlay.
struct RecStage {
    double x[3], xj[3], xpi[3], xpjb[3];   // r, j = closest(r), infront(r), behind(j)
    double cd, dot, lb, s3, la;             // closestd, t_r . t_j, l_before, l_after's first three terms,
                                            // l_after as staged
    int act, pari, behr, j, actj, parjb;    // 1-based links as staged
    int ii, iii, iv, ji, jii;               // infront chains from pari and from j, for loop_killer's walks
    int go;
};

// stage one candidate from the current state in memory
__device__ void stage_recon(const DState &s, const Phys &p, int r, RecStage &in) {
    const int cap = s.cap;
    in.act = s.infront[r] != 0;
    in.pari = s.infront[r];
    in.behr = s.behind[r];
    in.cd = s.closestd[r];
    in.j = s.closest[r];
    in.actj = 0;
    in.parjb = 0;
    in.ii = in.iii = in.iv = in.ji = in.jii = 0;
    in.dot = 0.0;
    in.lb = in.s3 = in.la = 0.0;
    in.go = 0;
    double gi[3], gb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        in.x[k] = s.x[k * cap + r];
        gi[k] = s.gi[k * cap + r];
        gb[k] = s.gb[k * cap + r];
        in.xj[k] = in.xpi[k] = in.xpjb[k] = 0.0;
    }
    if (in.j <= 0 || !in.act) return;
    const int rj = in.j - 1;
    in.actj = s.infront[rj] != 0;
    in.parjb = s.behind[rj];
    const int pi = in.pari - 1, pjb = in.parjb > 0 ? in.parjb - 1 : rj;
    double gij[3], gbj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        in.xj[k] = s.x[k * cap + rj];
        gij[k] = s.gi[k * cap + rj];
        gbj[k] = s.gb[k * cap + rj];
        in.xpi[k] = s.x[k * cap + pi];
        in.xpjb[k] = s.x[k * cap + pjb];
    }
    in.ii = s.infront[pi];
    if (in.ii > 0) in.iii = s.infront[in.ii - 1];
    if (in.iii > 0) in.iv = s.infront[in.iii - 1];
    in.ji = s.infront[rj];
    if (in.ji > 0) in.jii = s.infront[in.ji - 1];
    if (!in.actj) return;
    double t1[3], t2[3];
    norm_tanf_of(in.x, gi, gb, t1);
    norm_tanf_of(in.xj, gij, gbj, t2);
    in.dot = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
    in.lb = dist_gen(in.x[0], in.x[1], in.x[2], gi[0], gi[1], gi[2]) +
            dist_gen(in.x[0], in.x[1], in.x[2], gb[0], gb[1], gb[2]) +
            dist_gen(in.xj[0], in.xj[1], in.xj[2], gij[0], gij[1], gij[2]) +
            dist_gen(in.xj[0], in.xj[1], in.xj[2], gbj[0], gbj[1], gbj[2]);
    in.s3 = dist_gen(in.x[0], in.x[1], in.x[2], gb[0], gb[1], gb[2]) +
            dist_gen(in.x[0], in.x[1], in.x[2], in.xj[0], in.xj[1], in.xj[2]) +
            dist_gen(in.xj[0], in.xj[1], in.xj[2], gij[0], gij[1], gij[2]);
}

// src/reconnection.f90:20-43 on the staged inputs with the links as given:
// returns l_after in la when the visit reconnects
__device__ __forceinline__ bool recon_decide(const Phys &p, const RecStage &in, int pari, int behr, bool actj,
                                             const double xpi[3], const double xpjb[3], double &la) {
    if (!in.act) return false;
    if (!((in.cd < p.delta / 2.0) && (in.cd > NV_EPS))) return false;
    const int j = in.j;
    if (j == pari) return false;
    if (j == behr) return false;
    if (j == 0) return false;
    if (!actj) return false;
    if (in.dot > 0.9) return false;
    la = in.s3 + dist_gen(xpi[0], xpi[1], xpi[2], xpjb[0], xpjb[1], xpjb[2]);
    return la <= in.lb;
}

// The rows the pass's reconnections have changed so far, with their current
// links: one 16-byte slot per row (key, inf, beh, cleared), read in one LDS
// access.  inf / beh hold the 1-based link (0 once cleared), or -1 where
// that link is still the one the visits were staged with; xs is the row's
// own position (a changed row is r, j, infront(r) or behind(j) of some
// reconnection, whose positions the visit had).  Open addressing over
// OV_SLOTS slots, at most OV_MAX of them used.
constexpr int OV_SLOTS = 2048;
constexpr int OV_MAX = 1024;

struct __align__(16) OvSlot {
    int key, inf, beh, cl;
};

struct Overlay {
    OvSlot hd[OV_SLOTS];
    double xs[OV_SLOTS][3];
    int n;
};

__device__ __forceinline__ unsigned ov_hash(int r) { return ((unsigned)r * 2654435761u) >> (32 - 11); }
static_assert(OV_SLOTS == 1 << 11, "ov_hash keeps 11 bits");

// the slot of row r (its contents in h), or -1
__device__ __forceinline__ int ov_find(const Overlay &o, int r, OvSlot &h) {
    unsigned i = ov_hash(r);
    for (;;) {
        h = o.hd[i];
        if (h.key == r) return (int)i;
        if (h.key == -1) return -1;
        i = (i + 1) & (OV_SLOTS - 1);
    }
}

__device__ __forceinline__ bool ov_has(const Overlay &o, int r) {
    OvSlot h;
    return ov_find(o, r, h) >= 0;
}

// the slot of row r, created (links unchanged, position xs) when absent;
// the caller has checked that a slot is free
__device__ int ov_get(Overlay &o, int r, const double xs[3]) {
    unsigned i = ov_hash(r);
    for (;;) {
        const int k = o.hd[i].key;
        if (k == r) return (int)i;
        if (k == -1) break;
        i = (i + 1) & (OV_SLOTS - 1);
    }
    ++o.n;
    o.hd[i] = OvSlot{r, -1, -1, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) o.xs[i][k] = xs[k];
    return (int)i;
}

// The serial thread's view of one visit: infront of any row (1-based), from
// the overlay, then from the chains staged for this visit, then from memory.
struct VisitView {
    const DState &s;
    const Overlay &o;
    const RecStage &in;
    bool tracked;
    __device__ int inf(int row) const {
        if (!tracked) return s.infront[row];
        OvSlot h;
        if (ov_find(o, row, h) >= 0 && h.inf >= 0) return h.inf;
        if (in.pari > 0 && row == in.pari - 1) return in.ii;
        if (in.ii > 0 && row == in.ii - 1) return in.iii;
        if (in.iii > 0 && row == in.iii - 1) return in.iv;
        if (in.j > 0 && row == in.j - 1) return in.ji;
        if (in.ji > 0 && row == in.ji - 1) return in.jii;
        return s.infront[row];
    }
};

__device__ void clear_tracked(const DState &s, Overlay &o, bool tracked, int row) {
    clear_row(s, row);
    if (tracked) {
        const double z[3] = {0.0, 0.0, 0.0};
        const int i = ov_get(o, row, z);
        o.hd[i] = OvSlot{row, 0, 0, 1};
    }
}

// src/reconnection.f90:155-182 loop_killer(particle): as loop_killer above,
// reading links through the visit's view and recording what it clears
__device__ void loop_killer_view(const DState &s, Overlay &o, const VisitView &v, int particle) {
    int counter = 1, next = particle;
    for (int it = 1; it <= 3; ++it) {
        next = v.inf(next) - 1;
        if (next == particle) break;
        if (next < 0) return;
        counter = counter + 1;
        if (counter >= 4) return;
    }
    next = particle;
    for (int it = 0; it < 4; ++it) {
        const int store_next = v.inf(next) - 1;
        clear_tracked(s, o, v.tracked, next);
        next = store_next;
        if (next == particle || next < 0) break;
    }
}

// position of the row behind the 1-based link `link`, where the staged visit
// had link `staged` with position xst
__device__ __forceinline__ void pos_of(const DState &s, const Overlay &o, int link, int staged, const double xst[3],
                                       double out[3]) {
    if (link == staged) {
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = xst[k];
        return;
    }
    OvSlot h;
    const int i = ov_find(o, link - 1, h);
    if (i >= 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = o.xs[i][k];
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = s.x[k * s.cap + link - 1];
}

// One visit of the reference's precon loop (src/reconnection.f90:11-61) at
// staged index kk, by the serial thread.  With the overlay complete
// (tracked), the visit decides from its staged inputs and the overlay's
// links; otherwise it stages itself again from memory.  A reconnection is
// written to memory (links, cleared loops) and to the overlay, and queued
// for the chunk's event log.
__device__ void precon_visit(const DState &s, const Phys &p, int r, RecStage &sin, Overlay &o, bool &tracked,
                             int &nrec, int kk, int *evk, double *evla, int &nev) {
    int pari, behr, parjb;
    bool actj;
    double xpi[3], xpjb[3], la = 0.0;
    RecStage in;
    if (tracked) {
        // The decision depends on the links of r and j alone (the positions
        // it reads are those of rows linked to r and j, which only a clear of
        // their loop changes): where neither row has changed since staging,
        // the staged decision stands.
        OvSlot hr, hj;
        const int sj = sin.j;
        const int er = ov_find(o, r, hr);
        const int ej = sj > 0 ? ov_find(o, sj - 1, hj) : -1;
        if (er < 0 && ej < 0 && !sin.go) return;
        if (er >= 0 && hr.cl) return;    // emptied by a loop_killer of this pass
        // the staged record in registers: one LDS read of all of it, not
        // one dependent read per field
        in = sin;
        pari = (er >= 0 && hr.inf >= 0) ? hr.inf : in.pari;
        behr = (er >= 0 && hr.beh >= 0) ? hr.beh : in.behr;
        actj = in.actj;
        parjb = in.parjb;
        if (ej >= 0) {
            actj = actj && !hj.cl;
            if (hj.beh >= 0) parjb = hj.beh;
        }
        if (er < 0 && ej < 0) {
            la = in.la;   // unchanged, and its staged tests say reconnect
#pragma unroll
            for (int k = 0; k < 3; ++k) { xpi[k] = in.xpi[k]; xpjb[k] = in.xpjb[k]; }
        } else {
            if (!in.act || !actj || in.j == pari || in.j == behr) return;
            pos_of(s, o, pari, in.pari, in.xpi, xpi);
            pos_of(s, o, parjb, in.parjb, in.xpjb, xpjb);
            if (!recon_decide(p, in, pari, behr, actj, xpi, xpjb, la)) return;
        }
    } else {
        stage_recon(s, p, r, in);
        pari = in.pari; behr = in.behr; parjb = in.parjb; actj = in.actj;
#pragma unroll
        for (int k = 0; k < 3; ++k) { xpi[k] = in.xpi[k]; xpjb[k] = in.xpjb[k]; }
        if (!recon_decide(p, in, pari, behr, actj, xpi, xpjb, la)) return;
        sin = in;   // staged again from memory: the event log reads it
    }
    const int j = in.j, rj = j - 1;
    ++nrec;
    evk[nev] = kk;
    evla[nev] = la;
    ++nev;
    // four link slots and up to 3 + 3 cleared rows
    if (tracked && o.n + 10 > OV_MAX) tracked = false;
    s.infront[parjb - 1] = pari;
    s.behind[pari - 1] = parjb;
    s.infront[r] = j;
    s.behind[rj] = r + 1;
    if (tracked) {
        o.hd[ov_get(o, parjb - 1, xpjb)].inf = pari;
        o.hd[ov_get(o, pari - 1, xpi)].beh = parjb;
        o.hd[ov_get(o, r, in.x)].inf = j;
        o.hd[ov_get(o, rj, in.xj)].beh = r + 1;
    }
    const VisitView v{s, o, in, tracked};
    loop_killer_view(s, o, v, pari - 1);
    loop_killer_view(s, o, v, r);
}

// Every candidate of the pass staged at once, before any reconnection:
// candidate q into rs[q] for q < rs_cap.
__global__ void k_precon_stage(DState s, Phys p, const int *cand, const DCtl *ctl, RecStage *rs, int rs_cap) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl->abort || q >= ctl->ncand || q >= rs_cap) return;
    RecStage in;
    stage_recon(s, p, cand[q], in);
    double la = 0.0;
    in.go = recon_decide(p, in, in.pari, in.behr, in.actj != 0, in.xpi, in.xpjb, la) ? 1 : 0;
    in.la = la;
    rs[q] = in;
}

// One workgroup per tangle.  Candidates are taken REC_CHUNK at a time into
// LDS, as staged before the pass (k_precon_stage) while the overlay holds
// every change of the pass, else staged again from memory at the chunk's
// start with the overlay emptied.  The serial thread makes the visits in
// row order from the staged inputs and the overlay.  With few staged
// reconnections in a chunk, the workgroup finds the next visit that can
// change anything -- its staged tests say reconnect, or r or j (or the rows
// it read through them) is in the overlay -- and the serial thread makes
// only those.  The chunk's reconnections are logged at its end, in visit
// order, with one atomic for the whole chunk.
__global__ __launch_bounds__(REC_CHUNK) void k_precon_serial(DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl,
                                                             nv_recon_event *ev, int ev_cap, const RecStage *rs,
                                                             int rs_cap) {
    if (ctl->abort) return;
    __shared__ RecStage st[REC_CHUNK];
    __shared__ int strow[REC_CHUNK];
    __shared__ Overlay ov;
    __shared__ int evk[REC_CHUNK];
    __shared__ double evla[REC_CHUNK];
    __shared__ int nk[2];          // the next visit that needs the serial thread
    __shared__ int tracked_sh;     // the overlay holds every change since staging
    __shared__ int prestaged;      // ... since the pass's staging
    __shared__ int ngo;            // staged visits whose tests say "reconnect"
    __shared__ int nev_sh, ev0_sh;
    const int tid = threadIdx.x;
    const double t = ctl->t;
    const int itime = ctl->itime;
    const int tk = blockIdx.x;
    const int base = tg.base[tk];
    const int c0 = lower_row(cand, ctl->ncand, base), c1 = lower_row(cand, ctl->ncand, base + tg.cap[tk]);
    int nrec = 0;   // reconnections (the serial thread's)
    for (int q = tid; q < OV_SLOTS; q += REC_CHUNK) ov.hd[q].key = -1;
    if (tid == 0) {
        ov.n = 0;
        tracked_sh = 1;
        prestaged = 1;
    }
    for (int cs = c0; cs < c1; cs += REC_CHUNK) {
        const int m = min(REC_CHUNK, c1 - cs);
        __syncthreads();
        // the staging of the pass is usable while the overlay holds every
        // change since, and it has room for this chunk's
        const bool pre = prestaged && tracked_sh && ov.n <= OV_MAX / 2 && cs + m <= rs_cap;
        __syncthreads();
        if (!pre) {
            for (int q = tid; q < OV_SLOTS; q += REC_CHUNK) ov.hd[q].key = -1;
        }
        if (tid == 0) {
            if (!pre) {
                ov.n = 0;
                prestaged = 0;
                tracked_sh = 1;
            }
            nk[0] = nk[1] = m;
            ngo = 0;
            nev_sh = 0;
        }
        __syncthreads();
        if (tid < m) {
            RecStage &in = st[tid];
            strow[tid] = cand[cs + tid];
            if (pre) {
                in = rs[cs + tid];
            } else {
                stage_recon(s, p, strow[tid], in);
                double la = 0.0;
                in.go = recon_decide(p, in, in.pari, in.behr, in.actj != 0, in.xpi, in.xpjb, la) ? 1 : 0;
                in.la = la;
            }
            if (in.go) atomicAdd(&ngo, 1);
        }
        __syncthreads();
        if (ngo > REC_PARALLEL_MAX) {
            // many reconnections: the serial thread visits every candidate
            if (tid == 0) {
                bool tracked = tracked_sh != 0;
                int nev = 0;
                for (int kk = 0; kk < m; ++kk)
                    precon_visit(s, p, strow[kk], st[kk], ov, tracked, nrec, kk, evk, evla, nev);
                nev_sh = nev;
                tracked_sh = tracked ? 1 : 0;
            }
        } else {
            // A visit whose staged tests say "no reconnection" and none of
            // whose rows a reconnection of this pass has changed changes
            // nothing: the block finds the next visit that is not such a
            // one, and the serial thread runs it.
            int k = 0, par = 0;
            while (true) {
                if (tid >= k && tid < m) {
                    const RecStage &in = st[tid];
                    bool need = in.go;
                    if (!need && ov.n > 0) {
                        need = !tracked_sh || ov_has(ov, strow[tid]) ||
                               (in.act && in.j > 0 &&
                                (ov_has(ov, in.j - 1) || (in.pari > 0 && ov_has(ov, in.pari - 1)) ||
                                 (in.parjb > 0 && ov_has(ov, in.parjb - 1))));
                    }
                    if (need) atomicMin(&nk[par], tid);
                }
                __syncthreads();
                const int kk = nk[par];
                if (kk >= m) break;
                if (tid == 0) {
                    nk[par ^ 1] = m;
                    bool tracked = tracked_sh != 0;
                    int nev = nev_sh;
                    precon_visit(s, p, strow[kk], st[kk], ov, tracked, nrec, kk, evk, evla, nev);
                    tracked_sh = tracked ? 1 : 0;
                    nev_sh = nev;
                }
                k = kk + 1;
                par ^= 1;
                __syncthreads();
            }
        }
        // the chunk's reconnection log, in visit order
        __syncthreads();
        const int nev = nev_sh;
        if (nev == 0) continue;
        if (tid == 0) ev0_sh = atomicAdd(&ctl->nevents, nev);
        __syncthreads();
        for (int q = tid; q < nev; q += REC_CHUNK) {
            const int slot = ev0_sh + q;
            if (slot >= ev_cap) break;
            const RecStage &in = st[evk[q]];
            nv_recon_event e;
            e.itime = itime; e.i = strow[evk[q]] + 1 - base; e.j = in.j - base; e.tangle = tk; e.t = t;
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) e.mid[k2] = 0.5 * (in.x[k2] + in.xj[k2]);
            e.angle = acos(in.dot);
            e.dl = in.lb - evla[q];
            ev[slot] = e;
        }
    }
    if (tid == 0 && nrec) tg.recon[tk] += nrec;
}

// ---------------------------------------------------------------- premove
__global__ void k_premove_pre(DState s, Phys p, int n, int *cand) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int c = 0;
    if (s.infront[r] != 0) {
        const int ii = s.infront[s.infront[r] - 1] - 1;
        c = (ii >= 0 && dist_rows(s, s.x, r, s.x, ii) < .99 * p.delta) ? 1 : 0;
    }
    cand[r] = c;
}

// The inputs of one premove visit (src/line.f90:108-113) of row r: its
// infront inf, that one's infront ii, and the removal test; for a removal
// also r's behind and the next two rows of the loop after ii (iii, iv),
// which are what loop_killer(r) and the revisit of behind(r) read.
struct RmIn {
    int inf, ii, go, beh, iii, iv;
};

__device__ __forceinline__ void load_rm(const DState &s, const Phys &p, int r, RmIn &in) {
    in.inf = in.ii = in.beh = in.iii = in.iv = -1;
    in.go = 0;
    if (s.infront[r] == 0) return;
    in.inf = s.infront[r] - 1;
    in.ii = s.infront[in.inf] - 1;
    if (in.ii < 0) return;
    in.go = dist_rows(s, s.x, r, s.x, in.ii) < .99 * p.delta;
    if (!in.go) return;
    in.beh = s.behind[r] - 1;
    in.iii = s.infront[in.ii] - 1;
    if (in.iii >= 0) in.iv = s.infront[in.iii] - 1;
}

// loop_killer's count (src/reconnection.f90:155-170) of r's loop right after
// the removal of r's infront, from the staged rows: the loop runs r -> ii
// -> iii -> iv, and it has fewer than 4 points when one of those is r (an
// empty row on the way ends the count with no clearing)
__device__ __forceinline__ bool rm_loop_small(int r, const RmIn &in) {
    if (in.ii == r) return true;
    if (in.iii < 0) return false;
    if (in.iii == r) return true;
    return in.iv == r;
}

constexpr int RM_CHUNK = 256;

// The rows whose infront a chunk's removals change (r) or clear (r's
// infront), each with the index q in the chunk of the first removal that
// does: an LDS hash of RM_HASH keys and values (RM_CHUNK - q, so that the
// first removal is the largest value), laid over a ModSet's slots, all -1
// when empty.
constexpr int RM_HASH = MOD_SLOTS / 2;

__device__ __forceinline__ void rm_hash_put(int *keys, int *vals, int row, int q) {
    unsigned h = ((unsigned)row * 2654435761u) & (RM_HASH - 1);
    for (int t = 0; t < RM_HASH; ++t) {
        const int old = atomicCAS(&keys[h], -1, row);
        if (old == -1 || old == row) {
            atomicMax(&vals[h], RM_CHUNK - q);
            return;
        }
        h = (h + 1) & (RM_HASH - 1);
    }
}

// the first removal of the chunk that changes row's infront, or RM_CHUNK
__device__ __forceinline__ int rm_hash_get(const int *keys, const int *vals, int row) {
    unsigned h = ((unsigned)row * 2654435761u) & (RM_HASH - 1);
    for (int t = 0; t < RM_HASH; ++t) {
        const int k = keys[h];
        if (k == row) return RM_CHUNK - vals[h];
        if (k == -1) break;
        h = (h + 1) & (RM_HASH - 1);
    }
    return RM_CHUNK;
}

// loop_killer (src/reconnection.f90:155-182): clears particle's loop when it
// has fewer than 4 points (the count stops at 4, which decides the same)
__device__ void loop_killer_plain(const DState &s, int particle, int n) {
    int counter = 1, next = particle;
    for (int it = 1; it <= n; ++it) {
        next = s.infront[next] - 1;
        if (next == particle) break;
        if (next < 0) return;
        counter = counter + 1;
        if (counter >= 4) return;
    }
    next = particle;
    for (int it = 1; it <= n; ++it) {
        const int store_next = s.infront[next] - 1;
        clear_row(s, next);
        next = store_next;
        if (next == particle || next < 0) break;
    }
}

// The reference's loop over the candidates in row order (merged with the
// rows queued for a revisit), RM_CHUNK candidates at a time, in segments of
// candidates whose visits are made at once, between visits made one at a
// time.
//
// A removal at r changes infront(r), clears inf = infront(r), and sets
// behind(ii) to r.  The visits of a segment are made at once when none reads
// what an earlier removal of the segment changes: the infront of its r,
// inf, ii (and iii, whose infront is iv, when it removes); behind(ii) is read
// only by a later visit of ii, to decide whether to queue its behind, and r,
// visited first, lies below ii: no revisit.  The segment ends before the
// first queued row that is not a candidate, before a visit that reads what
// an earlier removal of the segment changed or that clears a loop, and after
// a removal that queues a revisit of a row of the chunk; revisits of later
// rows are queued after its stores.  The serial thread then makes the visits
// one at a time, from memory, until the next visit is a candidate again; the
// next segment loads its inputs again.
__global__ __launch_bounds__(RM_CHUNK) void k_premove_serial(DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl) {
    if (ctl->abort) return;
    __shared__ Dirty d;
    __shared__ RmIn st[RM_CHUNK];
    __shared__ int strow[RM_CHUNK];
    __shared__ int hkeys[RM_HASH], hvals[RM_HASH];
    __shared__ int pushv[RM_CHUNK];
    __shared__ unsigned long long nmask[RM_CHUNK / 64], pmask[RM_CHUNK / 64];
    __shared__ int stop_sh, res_sh, a_sh, vmin_sh, cut_sh;
    const int tid = threadIdx.x, lane = tid & 63;
    const int tk = blockIdx.x;
    const int base = tg.base[tk];
    const int c0 = lower_row(cand, ctl->ncand, base), c1 = lower_row(cand, ctl->ncand, base + tg.cap[tk]);
    const int *cl = cand + c0;
    const int ncand = c1 - c0;
    const int n = tg.n[tk];
    if (tid == 0) {
        d.hd = d.nd = 0;
        stop_sh = 0;
        res_sh = -1;
    }
    int nrm = 0;   // removals (thread 0's count)
    for (int cs = 0; cs < ncand; cs += RM_CHUNK) {
        const int m = min(RM_CHUNK, ncand - cs);
        __syncthreads();
        if (tid < m) {
            strow[tid] = cl[cs + tid];
            load_rm(s, p, strow[tid], st[tid]);
        }
        // rows below the next chunk's first candidate are this chunk's; the
        // last chunk also takes every queued row left
        const int limit = cs + m < ncand ? cl[cs + m] : 0x7fffffff;
        if (tid == 0) a_sh = 0;
        bool fresh = true;   // the staged inputs are current
        while (true) {
            __syncthreads();
            const int a = a_sh;
            if (stop_sh) break;
            if (tid == 0) {
                vmin_sh = 0x7fffffff;
                cut_sh = m;
            }
            for (int q = tid; q < RM_HASH; q += RM_CHUNK) hkeys[q] = hvals[q] = -1;
            pushv[tid] = -1;
            if (!fresh && tid >= a && tid < m) load_rm(s, p, strow[tid], st[tid]);
            __syncthreads();
            // the first queued row of the chunk's rows that is not one of its candidates
            for (int q = d.hd + tid; q < d.nd && d.v[q] < limit; q += RM_CHUNK)
                if (staged_index(strow, a, m, d.v[q]) == m) atomicMin(&vmin_sh, d.v[q]);
            const bool in0 = tid >= a && tid < m;
            RmIn in;
            in.go = 0;
            in.inf = in.ii = in.beh = in.iii = in.iv = -1;
            if (in0) in = st[tid];
            const int r = in0 ? strow[tid] : -1;
            if (in0 && in.go) {
                rm_hash_put(hkeys, hvals, r, tid);
                rm_hash_put(hkeys, hvals, in.inf, tid);
            }
            __syncthreads();
            const int vmin = vmin_sh;
            const int hi = vmin == 0x7fffffff ? m : lower_index(strow, a, m, vmin);
            const bool in_seg = tid >= a && tid < hi;
            int cut = 0x7fffffff, push = -1;
            if (in_seg) {
                const int rows[4] = {r, in.inf, in.ii, in.go ? in.iii : -1};
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (rows[q] >= 0 && rm_hash_get(hkeys, hvals, rows[q]) < tid) cut = tid;
                if (in.go) {
                    if (rm_loop_small(r, in)) {
                        cut = tid;
                    } else if (in.beh > r) {
                        push = in.beh;
                        if (in.beh < limit) cut = min(cut, tid + 1);
                    }
                }
            }
            if (cut != 0x7fffffff) atomicMin(&cut_sh, cut);
            __syncthreads();
            const int seg_end = min(cut_sh, hi);
            const bool mine = in_seg && tid < seg_end && in.go;
            if (mine) {
                s.behind[in.ii] = r + 1;
                s.infront[r] = in.ii + 1;
                clear_row(s, in.inf);
                pushv[tid] = push;
            }
            // (ballots of the whole wave, outside any branch)
            const unsigned long long mins = __ballot(mine);
            const unsigned long long mpush = __ballot(mine && push >= 0);
            if (lane == 0) {
                nmask[tid >> 6] = mins;
                pmask[tid >> 6] = mpush;
            }
            __syncthreads();
            if (tid == 0) {
#pragma unroll
                for (int w = 0; w < RM_CHUNK / 64; ++w) nrm += __popcll(nmask[w]);
                // queued rows visited as candidates of the segment
                if (seg_end > a)
                    while (d.hd < d.nd && d.v[d.hd] <= strow[seg_end - 1]) ++d.hd;
                bool room = true;
                for (int w = 0; w < RM_CHUNK / 64 && room; ++w)
                    for (unsigned long long b = pmask[w]; b && room; b &= b - 1)
                        room = dirty_push(d, pushv[64 * w + __ffsll(b) - 1]);
                if (!room) {
                    // no room: the plain loop takes over after the segment
                    res_sh = strow[seg_end - 1] + 1;
                    stop_sh = 1;
                }
                a_sh = seg_end;
            }
            __syncthreads();
            fresh = false;
            if (stop_sh) break;
            const int a2 = a_sh;
            const int rd0 = d.hd < d.nd ? d.v[d.hd] : 0x7fffffff;
            if (a2 >= m && rd0 >= limit) break;   // the chunk is done
            const bool serial = a2 == a || (rd0 < limit && (a2 >= m || rd0 < strow[a2]));
            if (!serial || tid != 0) continue;
            // one visit at a time, from memory, until the next visit is a candidate
            int k = a2;
            bool visited = false;
            while (true) {
                const int rc = k < m ? strow[k] : 0x7fffffff;
                const int rd = d.hd < d.nd ? d.v[d.hd] : 0x7fffffff;
                const int rv = rc < rd ? rc : rd;
                if (rv == 0x7fffffff || rv >= limit) break;
                if (visited && rc <= rd) break;
                visited = true;
                if (rc == rv) ++k;
                if (rd == rv) ++d.hd;
                RmIn iv;
                load_rm(s, p, rv, iv);
                if (!iv.go) continue;
                s.behind[iv.ii] = rv + 1;
                s.infront[rv] = iv.ii + 1;
                clear_row(s, iv.inf);
                ++nrm;
                if (rm_loop_small(rv, iv)) {
                    loop_killer_plain(s, rv, n);
                } else if (iv.beh > rv && !dirty_push(d, iv.beh)) {
                    // no room for a row to revisit: the plain loop takes over after rv
                    res_sh = rv + 1;
                    stop_sh = 1;
                    break;
                }
            }
            a_sh = k;
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (nrm) tg.remove[tk] += nrm;
        // the rest of the loop as the reference writes it
        if (res_sh >= 0) premove_loop(s, p, tg, tk, res_sh);
    }
}

// ------------------------------------------------------ brute close search
// Bounding sphere of the active points of each 64-row segment (raw
// coordinates, FP64); rad < 0 marks a segment without active points.  The
// radius is rounded up so that it bounds the exact distances.
struct Sphere64 {
    double c[3], rad;
};

__device__ __forceinline__ double wave_min64(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ double wave_max64(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// one wave per segment (rows is a multiple of 256)
__global__ __launch_bounds__(256) void k_seg_spheres(DState s, int rows, Sphere64 *__restrict__ sp) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const bool act = r < rows && s.infront[r] != 0;
    double lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double v = act ? s.x[k * s.cap + r] : 0.0;
        lo[k] = wave_min64(act ? v : 1.0e300);
        hi[k] = wave_max64(act ? v : -1.0e300);
    }
    if ((threadIdx.x & 63) != 0) return;
    Sphere64 o;
    if (lo[0] > hi[0]) {
        o.c[0] = o.c[1] = o.c[2] = 0.0;
        o.rad = -1.0;
    } else {
        double d2 = 0.0, m = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o.c[k] = 0.5 * (lo[k] + hi[k]);
            d2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
            m += fabs(o.c[k]);
        }
        const double rad = 0.5 * sqrt(d2);
        o.rad = rad * (1.0 + 1.0e-12) + 1.0e-15 * (m + rad);
    }
    sp[r / 64] = o;
}

// src/line.f90:135-160 pclose: the first (lowest) j with the smallest
// distf(i,j) among non-empty j of the same tangle not in {i, infront(i),
// behind(i)}.  Work items (tangle, target block, chunk of source tiles): each
// chunk keeps its first minimum, k_pclose_fin keeps the first chunk's on
// ties.  Distances are the reference's correctly rounded sqrt, compared
// strictly.  Partial block of tangle k: [S][cap_k].
//
// Pruning, exact: every target starts from the distance to a valid
// candidate, infront(infront(i)), as an upper bound just above it (so that
// candidate, and any point at least as close, is still accepted by the
// strict comparison; a chunk holding none of them reports no candidate).  A
// wave scans a 64-row source segment only if the segment's bounding sphere
// comes within the current best of one of its targets; a 256-row tile is
// staged in LDS only if some wave needs one of its segments.
__global__ __launch_bounds__(256) void k_pclose(DState s, Tangles tg, const Item *__restrict__ items,
                                                const Sphere64 *__restrict__ seg, double *__restrict__ pd,
                                                int *__restrict__ pj) {
    constexpr int T = 256;
    __shared__ double tile[T * 4];
    const Item I = items[blockIdx.x];
    const int tid = threadIdx.x;
    const int base = tg.base[I.k], n = tg.n[I.k], capk = tg.cap[I.k], S = tg.S[I.k];
    if (I.tb * 256 >= n) return;   // item listed for rows the tangle does not fill (yet)
    const int loc = I.tb * 256 + tid, i = base + loc;
    // a point whose links leave its tangle is skipped like an empty one (the
    // context's state is checked on upload; caller memory is not)
    const int fi0 = loc < n ? s.infront[i] - 1 : -1, bi0 = loc < n ? s.behind[i] - 1 : -1;
    const bool act = loc < n && fi0 >= base && fi0 < base + n && bi0 >= base && bi0 < base + n;
    double x0 = 0, x1 = 0, x2 = 0;
    int fi = -2, bi = -2;
    double best = 100.0;
    if (act) {
        x0 = s.x[i]; x1 = s.x[s.cap + i]; x2 = s.x[2 * s.cap + i];
        fi = fi0; bi = bi0;
        const int j2 = s.infront[fi] - 1;
        if (j2 >= base && j2 < base + n && j2 != i && j2 != fi && j2 != bi && s.infront[j2] != 0) {
            const double d0 = x0 - s.x[j2], d1 = x1 - s.x[s.cap + j2], d2 = x2 - s.x[2 * s.cap + j2];
            const double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            if (d < 100.0) best = fmin(100.0, d * (1.0 + 4.0e-16) + 1.0e-300);
        }
    }
    double bound = best * best * (1.0 + 8e-16);
    const int ntiles = (n + T - 1) / T;
    const int t0 = (int)((long)ntiles * I.c / S), t1 = (int)((long)ntiles * (I.c + 1) / S);
    int bj = -1;
    for (int t = t0; t < t1; ++t) {
        // segments of this tile the wave needs (wave-uniform bits)
        unsigned need = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const Sphere64 sp = seg[(base / T + t) * 4 + q];
            bool lane = false;
            if (act && sp.rad >= 0.0) {
                const double c0 = x0 - sp.c[0], c1 = x1 - sp.c[1], c2 = x2 - sp.c[2];
                const double dc = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
                lane = !((dc - sp.rad) * (1.0 - 1.0e-12) > best);
            }
            if (__any(lane)) need |= 1u << q;
        }
        if (!__syncthreads_or(need != 0)) continue;
        // empty rows get coordinates no distance can come near the bound from
        for (int q = tid; q < T; q += 256) {
            const int jl = t * T + q, j = base + jl;
            const bool v = jl < n && s.infront[j] != 0;
            tile[q * 4 + 0] = v ? s.x[j] : 1.0e150;
            tile[q * 4 + 1] = v ? s.x[s.cap + j] : 1.0e150;
            tile[q * 4 + 2] = v ? s.x[2 * s.cap + j] : 1.0e150;
        }
        __syncthreads();
        const int row0 = base + t * T;
        for (int q = 0; q < 4; ++q) {
            if (!(need & (1u << q))) continue;
            const int s0 = row0 + q * 64;
            // rows this lane excludes (itself, infront, behind) in the segment?
            const bool ex = act && ((unsigned)(i - s0) < 64u || (unsigned)(fi - s0) < 64u ||
                                    (unsigned)(bi - s0) < 64u);
            if (!act) continue;
            if (__any(ex)) {
                for (int jj = q * 64; jj < q * 64 + 64; ++jj) {
                    const double *r = &tile[jj * 4];
                    const int j = row0 + jj;
                    if (j == i || j == fi || j == bi) continue;
                    const double d0 = x0 - r[0], d1 = x1 - r[1], d2 = x2 - r[2];
                    const double dd = d0 * d0 + d1 * d1 + d2 * d2;
                    if (dd < bound) {
                        const double dist = sqrt(dd);
                        if (dist < best) {
                            best = dist;
                            bound = dist * dist * (1.0 + 8e-16);
                            bj = j;
                        }
                    }
                }
            } else {
                // no exclusion in the segment: a fused filter (within 1e-15 of
                // the reference's sum of squares), the exact sum for candidates
                double fbound = bound * (1.0 + 1e-15);
                for (int jj = q * 64; jj < q * 64 + 64; ++jj) {
                    const double *r = &tile[jj * 4];
                    const double d0 = x0 - r[0], d1 = x1 - r[1], d2 = x2 - r[2];
                    const double ddf = fma(d0, d0, fma(d1, d1, d2 * d2));
                    if (ddf < fbound) {
                        const double dd = d0 * d0 + d1 * d1 + d2 * d2;
                        if (dd < bound) {
                            const double dist = sqrt(dd);
                            if (dist < best) {
                                best = dist;
                                bound = dist * dist * (1.0 + 8e-16);
                                fbound = bound * (1.0 + 1e-15);
                                bj = row0 + jj;
                            }
                        }
                    }
                }
            }
        }
    }
    if (loc < capk) {
        const size_t o = tg.poff[I.k] + (size_t)I.c * capk + loc;
        pd[o] = best;
        pj[o] = bj;
    }
}

__global__ void k_pclose_fin(DState s, Tangles tg, int rows, const double *pd, const int *pj) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows || s.infront[i] == 0) return;
    const int k = tg.row2t ? tg.row2t[i] : 0;   // no row table: one tangle
    const int capk = tg.cap[k], loc = i - tg.base[k], S = tg.S[k];
    double best = 100.0;
    int bj = -1;
    for (int c = 0; c < S; ++c) {
        const size_t o = tg.poff[k] + (size_t)c * capk + loc;
        const double d = pd[o];
        if (pj[o] >= 0 && d < best) {
            best = d;
            bj = pj[o];
        }
    }
    s.closestd[i] = best;
    if (bj >= 0) s.closest[i] = bj + 1;
}

// Tables and work list of the one-tangle close search in a caller-given
// block (nvk_pclose_flat): tangle 0 owns rows [0, n), S chunks per target.
struct FlatTab {
    int base, cap, n, S;
    long long poff;
};

__global__ void k_pclose_flat_setup(FlatTab *tab, Item *items, int n, int S, int nitems) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) {
        tab->base = 0;
        tab->cap = n;
        tab->n = n;
        tab->S = S;
        tab->poff = 0;
    }
    if (q < nitems) items[q] = Item{0, q / S, q % S, 0};
}

// ------------------------------------------------------- step begin / end
// src/run.f90 loop head of a recorded step: itime advances, the per-step
// counters and the work queue start at zero.
__global__ void k_step_begin(DCtl *ctl, int *tnact, int K, int *queue) {
    for (int k = threadIdx.x; k < K; k += blockDim.x) tnact[k] = 0;
    if (threadIdx.x == 0) {
        ctl->itime = ctl->itime + 1;
        ctl->nvel = 0;
        ctl->nactive = 0;
        ctl->n_old = 0;
        ctl->total_length = 0.0;
        ctl->empty_after = 0;
        queue[0] = 0;
    }
}

// enforce_boundary (src/boundary.f90:91-138), then t = t + dt, then the NaN
// test of the wrapped positions (src/general.f90:172-191), in one launch
__global__ void k_step_end(DState s, Phys p, int n, DCtl *ctl, double dt) {
    if (ctl->abort) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) {
        ctl->t = ctl->t + dt;
        ctl->nvel_total += (unsigned long long)ctl->nvel;
    }
    if (r >= n) return;
    double x[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) x[k] = s.x[k * s.cap + r];
    if (s.infront[r] != 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (!p.periodic[k]) continue;
            const double L = p.box[k];
            if (x[k] > (L / 2.0)) x[k] = x[k] - L;
            else if (x[k] < (-L / 2.0)) x[k] = x[k] + L;
            s.x[k * s.cap + r] = x[k];
        }
        if (p.walls && s.pin[r]) {
            pin_to_wall(s, p, r);
#pragma unroll
            for (int k = 0; k < 3; ++k) x[k] = s.x[k * s.cap + r];
        }
    }
    if (isnan(x[0]) || isnan(x[1]) || isnan(x[2])) atomicOr(&ctl->status, NVK_NAN);
}

// ------------------------------------------------------------ relayout
// Copy every row of the old layout to its tangle's new range; links (1-based
// global rows, 0 = none) move with their tangle.
__global__ void k_relayout(DState o, DState nu, const int *__restrict__ row2t, const int *__restrict__ base_old,
                           const int *__restrict__ base_new, int rows_old) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows_old) return;
    const int k = row2t[r];
    const int shift = base_new[k] - base_old[k];
    const int q = r + shift;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        nu.x[c * nu.cap + q] = o.x[c * o.cap + r];
        nu.u[c * nu.cap + q] = o.u[c * o.cap + r];
        nu.u1[c * nu.cap + q] = o.u1[c * o.cap + r];
        nu.u2[c * nu.cap + q] = o.u2[c * o.cap + r];
        nu.usup[c * nu.cap + q] = o.usup[c * o.cap + r];
        nu.umf[c * nu.cap + q] = o.umf[c * o.cap + r];
        nu.gi[c * nu.cap + q] = o.gi[c * o.cap + r];
        nu.gb[c * nu.cap + q] = o.gb[c * o.cap + r];
    }
    nu.closestd[q] = o.closestd[r];
    nu.curv[q] = o.curv[r];
    nu.stretch[q] = o.stretch[r];
    nu.pin[q] = o.pin[r];
#pragma unroll
    for (int c = 0; c < 3; ++c) nu.wp[c * nu.cap + q] = o.wp[c * o.cap + r];
    const int f = o.infront[r], b = o.behind[r], cl = o.closest[r];
    nu.infront[q] = f > 0 ? f + shift : 0;
    nu.behind[q] = b > 0 ? b + shift : 0;
    nu.closest[q] = cl > 0 ? cl + shift : 0;
}

// ------------------------------------------------------------ clock
__global__ void k_clock_set(DCtl *ctl, double t, int itime) {
    ctl->t = t;
    ctl->itime = itime;
}

// src/run.f90: the loop index advances before the body runs
__global__ void k_clock_begin(DCtl *ctl) { ctl->itime = ctl->itime + 1; }

// src/run.f90: t=t+dt after enforce_boundary; the step's moved points are
// added to the running total
__global__ void k_clock_end(DCtl *ctl, double dt) {
    ctl->t = ctl->t + dt;
    ctl->nvel_total += (unsigned long long)ctl->nvel;
}

// ------------------------------------------------------------ solid walls
// With solid walls the topology phases run as the reference's own loops, one
// thread per tangle over every row (pinned points change the ghosts an
// insertion refreshes and the loops' skips, which the staged kernels above
// do not model).  Links are global 1-based rows.

// src/reconnection.f90:155-182 loop_killer: the walk returns at a pinned point
__device__ void loop_killer_pinned(const DState &s, int particle, int n) {
    int counter = 1, next = particle;
    for (int it = 1; it <= n; ++it) {
        if (s.pin[next]) return;
        next = s.infront[next] - 1;
        if (next == particle) break;
        if (next < 0) return;   // a broken chain (the reference would read f(0))
        counter = counter + 1;
    }
    if (counter < 4) {
        next = particle;
        for (int it = 1; it <= n; ++it) {
            const int store_next = s.infront[next] - 1;
            clear_row(s, next);
            next = store_next;
            if (next == particle || next < 0) break;
        }
    }
}

// src/line.f90:18-96 pinsert's loop over rows [from, end) of tangle tk
// (end: the tangle's pcount at entry), from the tangle's current count; the
// lengths of the rows it visits join total_length when len_at_visit, else
// only those of points it inserts into empty slots it has still to visit (a
// loop resumed after the staged kernel, whose pre-filter counted the rest).
// Returns the tangle's count after it.
__device__ int pinsert_loop(const DState &s, const Phys &p, const Tangles &tg, DCtl *ctl, int tk, int from, int end,
                            bool len_at_visit) {
    const int base = tg.base[tk], tcap = tg.cap[tk];
    int n = tg.n[tk];
    const int cap = s.cap;
    double total = 0.0;
    int fe = base;   // the first empty row (slots are only filled here)
    for (int r = from; r < end; ++r) {
        if (s.infront[r] == 0) continue;
        if (s.pin[r] & NV_PINNED_I) continue;
        double x[3], gi[3], gb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = s.x[k * cap + r];
            gi[k] = s.gi[k * cap + r];
            gb[k] = s.gb[k * cap + r];
        }
        const double disti = dist_gen(x[0], x[1], x[2], gi[0], gi[1], gi[2]);
        if (len_at_visit) total = total + disti;
        if (!(disti > p.delta)) continue;
        while (fe < base + n && s.infront[fe] != 0) ++fe;
        int slot;
        if (fe < base + n) {
            slot = fe;
        } else {
            if (n >= tcap) {
                atomicOr(&ctl->status, NVK_CAPACITY);
                break;
            }
            slot = base + n;
            ++n;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s.usup[k * cap + slot] = 0.0;
                s.umf[k * cap + slot] = 0.0;
            }
            clear_row(s, slot);
        }
        double fd[3], fdd[3], xn[3];
        derivs(x, gi, gb, fd, fdd);
        double curv = sqrt(fdd[0] * fdd[0] + fdd[1] * fdd[1] + fdd[2] * fdd[2]);
        if (curv > 1e-5) {
            curv = 1.0 / curv;
            if (curv * curv - 0.25 * (disti * disti) > 0.0) {
                const double rr = (sqrt(curv * curv - 0.25 * (disti * disti)) - curv) * curv;
#pragma unroll
                for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (x[k] + gi[k]) + rr * fdd[k];
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (x[k] + gi[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) xn[k] = 0.5 * (x[k] + gi[k]);
        }
        const int fi = s.infront[r] - 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s.x[k * cap + slot] = xn[k];
            s.u[k * cap + slot] = 0.5 * (s.u[k * cap + r] + s.u[k * cap + fi]);
            s.u1[k * cap + slot] = 0.0;
            s.u2[k * cap + slot] = 0.0;
            s.wp[k * cap + slot] = 0;
        }
        s.pin[slot] = 0;
        s.behind[slot] = r + 1;
        s.infront[slot] = fi + 1;
        store_ghosts(s, p, slot);
        s.behind[fi] = slot + 1;
        store_ghosts(s, p, fi);
        s.infront[r] = slot + 1;
        store_ghosts(s, p, r);
        if (!len_at_visit && slot > r && slot < end) {
            double g[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) g[k] = s.gi[k * cap + slot];
            total = total + dist_gen(xn[0], xn[1], xn[2], g[0], g[1], g[2]);
        }
    }
    int empties = 0;
    for (int r = base; r < base + n; ++r) empties += s.infront[r] == 0;
    if (total != 0.0) atomicAdd(&ctl->total_length, total);
    atomicAdd(&ctl->empty_after, empties);
    tg.n[tk] = n;
    return n;
}

// pinsert as the reference's loop, one thread per tangle (solid walls)
__global__ void k_pinsert_walls(DState s, Phys p, Tangles tg, DCtl *ctl) {
    if (ctl->abort || threadIdx.x != 0) return;
    const int tk = blockIdx.x, base = tg.base[tk], n = tg.n[tk];
    int active = 0;
    for (int r = base; r < base + n; ++r) active += s.infront[r] != 0;
    if (active == 0) {
        atomicOr(&ctl->status, NVK_ZERO_LENGTH);
        return;
    }
    atomicAdd(&ctl->n_old, n);
    pinsert_loop(s, p, tg, ctl, tk, base, base + n, true);
}

// src/reconnection.f90:11-61 precon (the tests of recon_tests)
__global__ void k_precon_walls(DState s, Phys p, Tangles tg, DCtl *ctl, nv_recon_event *ev, int ev_cap) {
    if (ctl->abort || threadIdx.x != 0) return;
    const int tk = blockIdx.x, base = tg.base[tk], n = tg.n[tk];
    const double t = ctl->t;
    const int itime = ctl->itime;
    int nrec = 0;
    for (int r = base; r < base + n; ++r) {
        ReconIn in;
        load_recon(s, r, in);
        recon_tests(p, in);
        if (!in.go) continue;
        const int pari = in.pari, j = in.j, rj = j - 1, parjb = in.parjb;
        ++nrec;
        const int q = atomicAdd(&ctl->nevents, 1);
        if (q < ev_cap) {
            nv_recon_event e;
            e.itime = itime; e.i = r + 1 - base; e.j = j - base; e.tangle = tk; e.t = t;
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) e.mid[k2] = 0.5 * (in.x[k2] + in.xj[k2]);
            e.angle = acos(in.dot);
            e.dl = in.lb - in.la;
            ev[q] = e;
        }
        s.infront[parjb - 1] = pari;
        s.behind[pari - 1] = parjb;
        s.infront[r] = j;
        s.behind[rj] = r + 1;
        loop_killer_pinned(s, pari - 1, n);
        loop_killer_pinned(s, r, n);
    }
    if (nrec) tg.recon[tk] += nrec;
}

// src/reconnection.f90:66-150 wall_recon
__global__ void k_wall_recon(DState s, Phys p, Tangles tg, DCtl *ctl) {
    if (ctl->abort || threadIdx.x != 0) return;
    const int tk = blockIdx.x, base = tg.base[tk], n = tg.n[tk];
    const int cap = s.cap;
    int nwall = 0;
    for (int r = base; r < base + n; ++r) {
        if (s.infront[r] == 0) continue;
        // a pinned pair with nothing between (:75-78)
        if ((s.pin[r] & NV_PINNED_B) && (s.pin[s.infront[r] - 1] & NV_PINNED_I)) {
            clear_row(s, s.infront[r] - 1);
            clear_row(s, r);
        }
        if (s.pin[r]) continue;
        for (int k = 0; k < 3; ++k) {
            if (!p.solid[k]) continue;
            const double xk = s.x[k * cap + r];
            int side = 0;
            if (xk > 0.5 * (p.box[k] - p.delta)) side = 1;
            else if (xk < -0.5 * (p.box[k] - p.delta)) side = -1;
            if (!side) continue;
            const int pari = s.infront[r] - 1, parb = s.behind[r] - 1;
            s.pin[pari] |= NV_PINNED_B;
            s.behind[pari] = pari + 1;
            s.pin[parb] |= NV_PINNED_I;
            s.infront[parb] = parb + 1;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                s.wp[q * cap + pari] = q == k ? side : 0;
                s.wp[q * cap + parb] = q == k ? side : 0;
            }
            clear_row(s, r);
            if (s.pin[pari] == (NV_PINNED_I | NV_PINNED_B)) clear_row(s, pari);
            if (s.pin[parb] == (NV_PINNED_I | NV_PINNED_B)) clear_row(s, parb);
            ++nwall;
        }
    }
    if (nwall) tg.wall[tk] += nwall;
}

// src/line.f90:102-130 premove's loop over rows [from, base + pcount) of tangle tk
__device__ void premove_loop(const DState &s, const Phys &p, const Tangles &tg, int tk, int from) {
    const int base = tg.base[tk], n = tg.n[tk];
    const int cap = s.cap;
    int nrm = 0;
    for (int r = from; r < base + n; ++r) {
        if (s.infront[r] == 0) continue;
        if (s.pin[r] & NV_PINNED_I) continue;
        const int inf = s.infront[r] - 1;
        if (s.pin[inf] & NV_PINNED_I) continue;
        const int ii = s.infront[inf] - 1;
        const double distii = dist_gen(s.x[r], s.x[cap + r], s.x[2 * cap + r], s.x[ii], s.x[cap + ii], s.x[2 * cap + ii]);
        if (distii < .99 * p.delta) {
            s.behind[ii] = r + 1;
            s.infront[r] = ii + 1;
            clear_row(s, inf);
            loop_killer_pinned(s, r, n);
            ++nrm;
        }
        // two points on the boundary (:126-128)
        if (s.infront[r] != 0 && (s.pin[r] & NV_PINNED_B) && (s.pin[s.infront[r] - 1] & NV_PINNED_I)) {
            clear_row(s, s.infront[r] - 1);
            clear_row(s, r);
        }
    }
    if (nrm) tg.remove[tk] += nrm;
}

// premove as the reference's loop, one thread per tangle (solid walls)
__global__ void k_premove_walls(DState s, Phys p, Tangles tg, DCtl *ctl) {
    if (ctl->abort || threadIdx.x != 0) return;
    premove_loop(s, p, tg, blockIdx.x, tg.base[blockIdx.x]);
}

}  // namespace

// ------------------------------------------------------------------ launch
#define GRID(n) dim3(((n) + 255) / 256), dim3(256)

hipError_t nvk_ghosts(hipStream_t st, DState s, Phys p, int n) {
    if (n > 0) hipLaunchKernelGGL(k_ghosts, GRID(n), 0, st, s, p, n);
    return hipGetLastError();
}
hipError_t nvk_wrap(hipStream_t st, DState s, Phys p, int n) {
    if (n > 0) hipLaunchKernelGGL(k_wrap, GRID(n), 0, st, s, p, n);
    return hipGetLastError();
}
hipError_t nvk_nan(hipStream_t st, DState s, int n, DCtl *ctl) {
    if (n > 0) hipLaunchKernelGGL(k_nan, GRID(n), 0, st, s, n, ctl);
    return hipGetLastError();
}
hipError_t nvk_pinsert_pre(hipStream_t st, DState s, Phys p, Tangles tg, int n, int *cand, int *empty, DCtl *ctl) {
    if (n > 0) hipLaunchKernelGGL(k_pinsert_pre, GRID(n), 0, st, s, p, tg, n, cand, empty, ctl);
    return hipGetLastError();
}
hipError_t nvk_pinsert_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, const int *empty,
                              DCtl *ctl) {
    hipLaunchKernelGGL(k_pinsert_serial, dim3(tg.K), dim3(INS_CHUNK), 0, st, s, p, tg, cand, empty, ctl);
    return hipGetLastError();
}
hipError_t nvk_precon_pre(hipStream_t st, DState s, Phys p, int n, int *cand) {
    if (n > 0) hipLaunchKernelGGL(k_precon_pre, GRID(n), 0, st, s, p, n, cand);
    return hipGetLastError();
}
size_t nvk_recstage_bytes() { return sizeof(RecStage); }
hipError_t nvk_precon_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl,
                             nv_recon_event *ev, int ev_cap, void *rs, int rs_cap) {
    RecStage *r = (RecStage *)rs;
    if (rs_cap > 0) hipLaunchKernelGGL(k_precon_stage, GRID(rs_cap), 0, st, s, p, cand, ctl, r, rs_cap);
    hipLaunchKernelGGL(k_precon_serial, dim3(tg.K), dim3(REC_CHUNK), 0, st, s, p, tg, cand, ctl, ev, ev_cap, r,
                       rs_cap);
    return hipGetLastError();
}
hipError_t nvk_premove_pre(hipStream_t st, DState s, Phys p, int n, int *cand) {
    if (n > 0) hipLaunchKernelGGL(k_premove_pre, GRID(n), 0, st, s, p, n, cand);
    return hipGetLastError();
}
hipError_t nvk_premove_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl) {
    hipLaunchKernelGGL(k_premove_serial, dim3(tg.K), dim3(RM_CHUNK), 0, st, s, p, tg, cand, ctl);
    return hipGetLastError();
}
size_t nvk_sphere_bytes() { return sizeof(Sphere64); }

hipError_t nvk_pclose(hipStream_t st, DState s, Tangles tg, int rows, const Item *items, int nitems, void *spheres,
                      double *pd, int *pj) {
    Sphere64 *sp = (Sphere64 *)spheres;
    if (rows > 0) hipLaunchKernelGGL(k_seg_spheres, dim3((rows + 255) / 256), dim3(256), 0, st, s, rows, sp);
    if (nitems > 0) hipLaunchKernelGGL(k_pclose, dim3(nitems), dim3(256), 0, st, s, tg, items, sp, pd, pj);
    if (rows > 0) hipLaunchKernelGGL(k_pclose_fin, GRID(rows), 0, st, s, tg, rows, pd, pj);
    return hipGetLastError();
}

// The brute close search of one tangle whose arrays the caller holds:
// positions component-major with leading dimension ld (x[k*ld + r]), its
// working memory (tables, work list, segment spheres, chunk partials) laid
// out in `work`, which must hold nvk_pclose_flat_bytes(n, S) bytes.
static size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

size_t nvk_pclose_flat_bytes(int n, int S) {
    const size_t ntb = ((size_t)n + 255) / 256;
    return up256(sizeof(FlatTab)) + up256(sizeof(Item) * ntb * S) + up256(sizeof(Sphere64) * ntb * 4) +
           up256(sizeof(double) * (size_t)S * n) + up256(sizeof(int) * (size_t)S * n);
}

hipError_t nvk_pclose_flat(hipStream_t st, const double *x, int ld, const int *infront, const int *behind,
                           int *closest, double *closestd, int n, int S, void *work) {
    if (n <= 0) return hipSuccess;
    const int ntb = (n + 255) / 256, nitems = ntb * S;
    char *w = (char *)work;
    FlatTab *tab = (FlatTab *)w;
    w += up256(sizeof(FlatTab));
    Item *items = (Item *)w;
    w += up256(sizeof(Item) * (size_t)nitems);
    Sphere64 *sp = (Sphere64 *)w;
    w += up256(sizeof(Sphere64) * (size_t)ntb * 4);
    double *pd = (double *)w;
    w += up256(sizeof(double) * (size_t)S * n);
    int *pj = (int *)w;
    DState s{};
    s.cap = ld;
    s.x = const_cast<double *>(x);
    s.infront = const_cast<int *>(infront);
    s.behind = const_cast<int *>(behind);
    s.closest = closest;
    s.closestd = closestd;
    Tangles tg{};
    tg.K = 1;
    tg.base = &tab->base;
    tg.cap = &tab->cap;
    tg.n = &tab->n;
    tg.S = &tab->S;
    tg.poff = &tab->poff;
    tg.row2t = nullptr;
    hipLaunchKernelGGL(k_pclose_flat_setup, dim3((nitems + 255) / 256), dim3(256), 0, st, tab, items, n, S, nitems);
    hipLaunchKernelGGL(k_seg_spheres, dim3(ntb), dim3(256), 0, st, s, n, sp);
    hipLaunchKernelGGL(k_pclose, dim3(nitems), dim3(256), 0, st, s, tg, items, sp, pd, pj);
    hipLaunchKernelGGL(k_pclose_fin, GRID(n), 0, st, s, tg, n, pd, pj);
    return hipGetLastError();
}

int nvk_diag_parts(int n) { return ((n + 255) / 256) * DIAG_VALS; }

hipError_t nvk_diagnostics(hipStream_t st, DState s, int n, double *part, DDiag *out) {
    const int nblk = (n + 255) / 256;
    if (nblk > 0) hipLaunchKernelGGL(k_diag, dim3(nblk), dim3(256), 0, st, s, n, part);
    hipLaunchKernelGGL(k_diag_fin, dim3(1), dim3(256), 0, st, (const double *)part, nblk, out);
    return hipGetLastError();
}

hipError_t nvk_clock_set(hipStream_t st, DCtl *ctl, double t, int itime) {
    hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(1), 0, st, ctl, t, itime);
    return hipGetLastError();
}
hipError_t nvk_clock_begin_step(hipStream_t st, DCtl *ctl) {
    hipLaunchKernelGGL(k_clock_begin, dim3(1), dim3(1), 0, st, ctl);
    return hipGetLastError();
}
hipError_t nvk_clock_end_step(hipStream_t st, DCtl *ctl, double dt) {
    hipLaunchKernelGGL(k_clock_end, dim3(1), dim3(1), 0, st, ctl, dt);
    return hipGetLastError();
}

hipError_t nvk_relayout(hipStream_t st, DState o, DState nu, const int *row2t, const int *base_old,
                        const int *base_new, int rows_old) {
    if (rows_old > 0)
        hipLaunchKernelGGL(k_relayout, GRID(rows_old), 0, st, o, nu, row2t, base_old, base_new, rows_old);
    return hipGetLastError();
}

hipError_t nvk_step_begin(hipStream_t st, DCtl *ctl, int *tnact, int K, int *queue) {
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(256), 0, st, ctl, tnact, K, queue);
    return hipGetLastError();
}
hipError_t nvk_step_end(hipStream_t st, DState s, Phys p, int n, DCtl *ctl, double dt) {
    hipLaunchKernelGGL(k_step_end, dim3((std::max(n, 1) + 255) / 256), dim3(256), 0, st, s, p, n, ctl, dt);
    return hipGetLastError();
}

hipError_t nvk_pinsert_walls(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl) {
    hipLaunchKernelGGL(k_pinsert_walls, dim3(tg.K), dim3(64), 0, st, s, p, tg, ctl);
    return hipGetLastError();
}
hipError_t nvk_precon_walls(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl, nv_recon_event *ev, int ev_cap) {
    hipLaunchKernelGGL(k_precon_walls, dim3(tg.K), dim3(64), 0, st, s, p, tg, ctl, ev, ev_cap);
    return hipGetLastError();
}
hipError_t nvk_wall_recon(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl) {
    hipLaunchKernelGGL(k_wall_recon, dim3(tg.K), dim3(64), 0, st, s, p, tg, ctl);
    return hipGetLastError();
}
hipError_t nvk_premove_walls(hipStream_t st, DState s, Phys p, Tangles tg, DCtl *ctl) {
    hipLaunchKernelGGL(k_premove_walls, dim3(tg.K), dim3(64), 0, st, s, p, tg, ctl);
    return hipGetLastError();
}
