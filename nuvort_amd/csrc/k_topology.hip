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
    // src/general.f90:10-20 clear_particle (u_sup, u_mf untouched)
    const int cap = s.cap;
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
}

__global__ void k_nan(DState s, int n, DCtl *ctl) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    if (isnan(s.x[r]) || isnan(s.x[s.cap + r]) || isnan(s.x[2 * s.cap + r])) atomicOr(&ctl->status, NVK_NAN);
}

// ---------------------------------------------------------------- pinsert
__global__ void k_pinsert_pre(DState s, Phys p, Tangles tg, int n, int *cand, int *empty, DCtl *ctl) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = r < n && s.infront[r] != 0;
    if (r < n) {
        // a free slot of the reference's f(1:pcount): empty and below pcount
        const int k = tg.row2t[r];
        empty[r] = (!act && r - tg.base[k] < tg.n[k]) ? 1 : 0;
    }
    int c = 0;
    if (act) {
        double disti = dist_gen(s.x[r], s.x[s.cap + r], s.x[2 * s.cap + r], s.gi[r], s.gi[s.cap + r],
                                s.gi[2 * s.cap + r]);
        c = disti > p.delta ? 1 : 0;
    }
    if (r < n) cand[r] = c;
    // active points per tangle for the zero-length stop (src/line.f90:25-27);
    // tangles are 256-row aligned, so a wavefront lies in one tangle
    const unsigned long long m = __ballot(act);
    if ((threadIdx.x & 63) == 0 && m) {
        atomicAdd(&ctl->nactive, (int)__popcll(m));
        atomicAdd(&tg.nact[tg.row2t[r]], (int)__popcll(m));
    }
}

__global__ void k_pinsert_serial(DState s, Phys p, Tangles tg, const int *cand, const int *empty, DCtl *ctl) {
    if (threadIdx.x != 0) return;
    __shared__ Dirty d;
    d.hd = d.nd = 0;
    const int tk = blockIdx.x;
    const int base = tg.base[tk], tcap = tg.cap[tk];
    int nl = tg.n[tk];                     // local pcount
    const int n_old = base + nl;           // rows the reference loop visits end here
    // this tangle's slices of the candidate and empty-slot lists
    const int c0 = lower_row(cand, ctl->ncand, base), c1 = lower_row(cand, ctl->ncand, base + tcap);
    const int e0 = lower_row(empty, ctl->nempty, base), e1 = lower_row(empty, ctl->nempty, base + tcap);
    // src/line.f90:25-27: no active point in this tangle
    if (tg.nact[tk] == 0) { atomicOr(&ctl->status, NVK_ZERO_LENGTH); return; }
    const int *cl = cand + c0;
    const int ncand = c1 - c0;
    const int cap = s.cap;
    int ci = 0, ei = e0;
    while (true) {
        const int r = next_row(cl, ncand, ci, d);
        if (r == 0x7fffffff) break;
        if (s.infront[r] == 0) continue;
        double x[3], gi[3], gb[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = s.x[k * cap + r];
            gi[k] = s.gi[k * cap + r];
            gb[k] = s.gb[k * cap + r];
        }
        const double disti = dist_gen(x[0], x[1], x[2], gi[0], gi[1], gi[2]);
        if (!(disti > p.delta)) continue;
        // every other input of the insertion, loaded together: the three
        // ghost refreshes below are computed from these values, as
        // store_ghosts (src/boundary.f90:26-84) would from the updated state
        const int fi = s.infront[r] - 1, bb = s.behind[r] - 1;
        const int ffi = s.infront[fi] - 1;
        double xf[3], xff[3], xb[3], ur[3], uf[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xf[k] = s.x[k * cap + fi];
            xff[k] = s.x[k * cap + ffi];
            xb[k] = s.x[k * cap + bb];
            ur[k] = s.u[k * cap + r];
            uf[k] = s.u[k * cap + fi];
        }
        int slot;
        if (ei < e1) {
            slot = empty[ei++];
        } else {
            if (nl >= tcap) { atomicOr(&ctl->status, NVK_CAPACITY); break; }
            slot = base + nl;
            ++nl;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s.usup[k * cap + slot] = 0.0;
                s.umf[k * cap + slot] = 0.0;
            }
            clear_row(s, slot);
        }
        double fd[3], fdd[3];
        derivs(x, gi, gb, fd, fdd);
        double curv = sqrt(fdd[0] * fdd[0] + fdd[1] * fdd[1] + fdd[2] * fdd[2]);
        double xn[3];
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
        // the new point, its links, and the ghosts of the new point, of r and
        // of its old infront fi from the positions now in place (fi's
        // infront and r's behind are unchanged by this insertion)
        double gs_i[3], gs_b[3], gr_i[3], gr_b[3], gf_i[3], gf_b[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gs_i[k] = xf[k]; gs_b[k] = x[k];
            gr_i[k] = xn[k]; gr_b[k] = xb[k];
            gf_i[k] = xff[k]; gf_b[k] = xn[k];
            if (!p.periodic[k]) continue;
            const double L = p.box[k];
            wrap_ghost(xn[k], L, gs_i[k]);
            wrap_ghost(xn[k], L, gs_b[k]);
            wrap_ghost(x[k], L, gr_i[k]);
            wrap_ghost(x[k], L, gr_b[k]);
            wrap_ghost(xf[k], L, gf_i[k]);
            wrap_ghost(xf[k], L, gf_b[k]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s.x[k * cap + slot] = xn[k];
            s.u[k * cap + slot] = 0.5 * (ur[k] + uf[k]);
            s.u1[k * cap + slot] = 0.0;
            s.u2[k * cap + slot] = 0.0;
            s.gi[k * cap + slot] = gs_i[k];
            s.gb[k * cap + slot] = gs_b[k];
            // fi before r, the reference's refresh order (they coincide only
            // for a one-point loop)
            s.gi[k * cap + fi] = gf_i[k];
            s.gb[k * cap + fi] = gf_b[k];
            s.gi[k * cap + r] = gr_i[k];
            s.gb[k * cap + r] = gr_b[k];
        }
        s.behind[slot] = r + 1;
        s.infront[slot] = fi + 1;
        s.behind[fi] = slot + 1;
        s.infront[r] = slot + 1;
        // rows the reference loop still has to reach whose inputs changed
        if (fi > r && fi < n_old && !dirty_push(d, fi)) { atomicOr(&ctl->status, NVK_CAPACITY); break; }
        if (slot > r && slot < n_old && !dirty_push(d, slot)) { atomicOr(&ctl->status, NVK_CAPACITY); break; }
    }
    tg.n[tk] = nl;
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

__global__ void k_precon_serial(DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl, nv_recon_event *ev,
                                int ev_cap) {
    if (threadIdx.x != 0) return;
    const double t = ctl->t;
    const int itime = ctl->itime;
    const int tk = blockIdx.x;
    const int base = tg.base[tk];
    const int c0 = lower_row(cand, ctl->ncand, base), c1 = lower_row(cand, ctl->ncand, base + tg.cap[tk]);
    const int n = tg.n[tk];
    const int cap = s.cap;
    for (int ci = c0; ci < c1; ++ci) {
        const int r = cand[ci];
        if (s.infront[r] == 0) continue;
        const int pari = s.infront[r];
        const double cd = s.closestd[r];
        if (!((cd < p.delta / 2.0) && (cd > NV_EPS))) continue;
        const int j = s.closest[r];
        if (j == pari) continue;
        if (j == s.behind[r]) continue;
        if (j == 0) continue;
        const int rj = j - 1;
        if (s.infront[rj] == 0) continue;
        const int parjb = s.behind[rj];
        double t1[3], t2[3];
        norm_tanf(s, r, t1);
        norm_tanf(s, rj, t2);
        const double dot_val = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
        if (dot_val > 0.9) continue;
        const double l_before = dist_rows(s, s.x, r, s.gi, r) + dist_rows(s, s.x, r, s.gb, r) +
                                dist_rows(s, s.x, rj, s.gi, rj) + dist_rows(s, s.x, rj, s.gb, rj);
        const double l_after = dist_rows(s, s.x, r, s.gb, r) + dist_rows(s, s.x, r, s.x, rj) +
                               dist_rows(s, s.x, rj, s.gi, rj) + dist_rows(s, s.x, pari - 1, s.x, parjb - 1);
        if (l_after <= l_before) {
            tg.recon[tk] += 1;
            const int q = atomicAdd(&ctl->nevents, 1);
            if (q < ev_cap) {
                nv_recon_event e;
                e.itime = itime; e.i = r + 1 - base; e.j = j - base; e.tangle = tk; e.t = t;
#pragma unroll
                for (int k = 0; k < 3; ++k) e.mid[k] = 0.5 * (s.x[k * cap + r] + s.x[k * cap + rj]);
                e.angle = acos(dot_val);
                e.dl = l_before - l_after;
                ev[q] = e;
            }
            s.infront[parjb - 1] = pari;
            s.behind[pari - 1] = parjb;
            s.infront[r] = j;
            s.behind[rj] = r + 1;
            loop_killer(s, pari - 1, n);
            loop_killer(s, r, n);
        }
    }
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

__global__ void k_premove_serial(DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl) {
    if (threadIdx.x != 0) return;
    __shared__ Dirty d;
    d.hd = d.nd = 0;
    const int tk = blockIdx.x;
    const int base = tg.base[tk];
    const int c0 = lower_row(cand, ctl->ncand, base), c1 = lower_row(cand, ctl->ncand, base + tg.cap[tk]);
    const int *cl = cand + c0;
    const int ncand = c1 - c0;
    const int n = tg.n[tk];
    int ci = 0;
    while (true) {
        const int r = next_row(cl, ncand, ci, d);
        if (r == 0x7fffffff) break;
        if (s.infront[r] == 0) continue;
        const int inf = s.infront[r] - 1;
        const int ii = s.infront[inf] - 1;
        if (ii < 0) continue;
        const double distii = dist_rows(s, s.x, r, s.x, ii);
        if (distii < .99 * p.delta) {
            const int tinfront = ii;
            s.behind[tinfront] = r + 1;
            s.infront[r] = tinfront + 1;
            clear_row(s, inf);
            loop_killer(s, r, n);
            tg.remove[tk] += 1;
            if (s.infront[r] != 0) {
                const int b = s.behind[r] - 1;
                if (b > r && !dirty_push(d, b)) { atomicOr(&ctl->status, NVK_CAPACITY); break; }
            }
        }
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
    const bool act = loc < n && s.infront[i] != 0;
    double x0 = 0, x1 = 0, x2 = 0;
    int fi = -2, bi = -2;
    double best = 100.0;
    if (act) {
        x0 = s.x[i]; x1 = s.x[s.cap + i]; x2 = s.x[2 * s.cap + i];
        fi = s.infront[i] - 1; bi = s.behind[i] - 1;
        const int j2 = s.infront[fi] - 1;
        if (j2 >= 0 && j2 != i && j2 != fi && j2 != bi && s.infront[j2] != 0) {
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
    const int k = tg.row2t[i];
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

// ------------------------------------------------------- step begin / end
// src/run.f90 loop head of a recorded step: itime advances, the per-step
// counters and the work queue start at zero.
__global__ void k_step_begin(DCtl *ctl, int *tnact, int K, int *queue) {
    for (int k = threadIdx.x; k < K; k += blockDim.x) tnact[k] = 0;
    if (threadIdx.x == 0) {
        ctl->itime = ctl->itime + 1;
        ctl->nvel = 0;
        ctl->nactive = 0;
        queue[0] = 0;
    }
}

// enforce_boundary (src/boundary.f90:91-138), then t = t + dt, then the NaN
// test of the wrapped positions (src/general.f90:172-191), in one launch
__global__ void k_step_end(DState s, Phys p, int n, DCtl *ctl, double dt) {
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
    hipLaunchKernelGGL(k_pinsert_serial, dim3(tg.K), dim3(64), 0, st, s, p, tg, cand, empty, ctl);
    return hipGetLastError();
}
hipError_t nvk_precon_pre(hipStream_t st, DState s, Phys p, int n, int *cand) {
    if (n > 0) hipLaunchKernelGGL(k_precon_pre, GRID(n), 0, st, s, p, n, cand);
    return hipGetLastError();
}
hipError_t nvk_precon_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl,
                             nv_recon_event *ev, int ev_cap) {
    hipLaunchKernelGGL(k_precon_serial, dim3(tg.K), dim3(64), 0, st, s, p, tg, cand, ctl, ev, ev_cap);
    return hipGetLastError();
}
hipError_t nvk_premove_pre(hipStream_t st, DState s, Phys p, int n, int *cand) {
    if (n > 0) hipLaunchKernelGGL(k_premove_pre, GRID(n), 0, st, s, p, n, cand);
    return hipGetLastError();
}
hipError_t nvk_premove_serial(hipStream_t st, DState s, Phys p, Tangles tg, const int *cand, DCtl *ctl) {
    hipLaunchKernelGGL(k_premove_serial, dim3(tg.K), dim3(64), 0, st, s, p, tg, cand, ctl);
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
