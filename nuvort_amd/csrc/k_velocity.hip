// Direct Biot-Savart sum over straight segments with the 26 periodic images
// (src/timestep.f90:74-98 with biot_savart :151-172 and biot_savart_shift
// :175-198), fused with the local term, mutual friction and the
// Adams-Bashforth update (src/timestep.f90:16-45).
//
// Work decomposition.  For every tangle, targets are split into blocks of
// 256 lanes x TPT points; the (image, source tile) sequence of each target --
// image-major, tiles of 256 sources in ascending row order, the order the
// reference visits them -- is split into S contiguous chunks.  A work item is
// (tangle, target block, chunk).  One persistent launch runs all items of all
// tangles: every workgroup pulls the next item from an atomic queue (items
// are listed largest tangle first), stages each source tile in LDS with
// coalesced reads, and every lane then reads the same LDS record
// (broadcast), so the inner loop is pure VALU.  Chunk partial sums go to a
// per-tangle partial block that k_finish reduces in chunk order, so the
// result does not depend on which workgroup ran which item.  With S = 1 the
// summation order per target is the reference's, term for term.
#include "nv_kernels.h"
#include "nv_util.h"

namespace {

constexpr int REC64 = 8;   // x0 x1 x2 e0 e1 e2 c valid
constexpr int TB = 256;    // lanes per workgroup
constexpr int TS = 256;    // sources per tile

// a row's segment is a source of the direct sums: not empty, and not going
// into a wall (pinned in front, src/timestep.f90:161,187)
__device__ __forceinline__ bool is_source(const DState &s, int r) {
    return s.infront[r] != 0 && !(s.pin && (s.pin[r] & NV_PINNED_I));
}

// FP64 source records, one per row: x, e = ghosti - x, c = |e|^2
// (dist_gen_sq(ghosti, x), src/timestep.f90:164), valid = is_source.
__global__ void k_src64(DState s, int rows, double *__restrict__ rec) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double v[REC64] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (is_source(s, r)) {
        double e[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = ld(s.x, s.cap, k, r);
            e[k] = ld(s.gi, s.cap, k, r) - v[k];
            v[3 + k] = e[k];
        }
        v[6] = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        v[7] = 1.0;
    }
#pragma unroll
    for (int k = 0; k < REC64; ++k) rec[(size_t)r * REC64 + k] = v[k];
}

__device__ __forceinline__ int next_item(int *queue, int *slot) {
    __syncthreads();
    if (threadIdx.x == 0) *slot = atomicAdd(queue, 1);
    __syncthreads();
    return *slot;
}

// FP64 plain direct sum, the reference's arithmetic.
// Partial block of tangle k: [S][6][cap_k] (acc0 xyz, accs xyz).
__global__ __launch_bounds__(TB) void k_bs64(DState s, Tangles tg, const Item *__restrict__ items, int nitems,
                                             int *queue, const double *__restrict__ rec, Shifts sh, double kappa,
                                             double *__restrict__ part) {
    __shared__ double tile[TS * REC64];
    __shared__ int slot;
    const int tid = threadIdx.x;
    for (;;) {
        const int it = next_item(queue, &slot);
        if (it >= nitems) break;
        const Item I = items[it];
        const int base = tg.base[I.k], n = tg.n[I.k], capk = tg.cap[I.k], S = tg.S[I.k];
        if (I.tb * TB >= n) continue;   // item listed for rows the tangle does not fill (yet)
        const int loc = I.tb * TB + tid, i = base + loc;
        const bool act = (loc < n) && (s.infront[i] != 0);
        double xi0 = 0, xi1 = 0, xi2 = 0;
        int bi = -2;
        if (act) {
            xi0 = ld(s.x, s.cap, 0, i);
            xi1 = ld(s.x, s.cap, 1, i);
            xi2 = ld(s.x, s.cap, 2, i);
            bi = s.behind[i] - 1;
        }
        double a00 = 0, a01 = 0, a02 = 0, as0 = 0, as1 = 0, as2 = 0;
        const int ntiles = (n + TS - 1) / TS;
        const long U = (long)sh.n * ntiles;
        const long u0 = U * I.c / S, u1 = U * (I.c + 1) / S;
        for (long un = u0; un < u1; ++un) {
            const int m = (int)(un / ntiles), t = (int)(un % ntiles);
            const double s0 = sh.s[m][0], s1 = sh.s[m][1], s2 = sh.s[m][2];
            __syncthreads();
            for (int q = tid; q < TS * REC64; q += TB) tile[q] = rec[((size_t)base + (size_t)t * TS) * REC64 + q];
            __syncthreads();
            if (!act) continue;
            const int jb = base + t * TS;
            for (int jj = 0; jj < TS; ++jj) {
                const double *qv = &tile[jj * REC64];
                if (qv[7] == 0.0) continue;
                const int j = jb + jj;
                if (j == bi || (m == 0 && j == i)) continue;
                double out[3];
                if (!bs_term64((qv[0] + s0) - xi0, (qv[1] + s1) - xi1, (qv[2] + s2) - xi2, qv[3], qv[4], qv[5],
                               qv[6], kappa, out))
                    continue;
                // the reference's running-sum order, term by term
                if (m == 0) { a00 = a00 + out[0]; a01 = a01 + out[1]; a02 = a02 + out[2]; }
                else { as0 = as0 + out[0]; as1 = as1 + out[1]; as2 = as2 + out[2]; }
            }
        }
        if (loc < capk) {
            double *pp = part + tg.poff[I.k] + (size_t)I.c * 6 * capk;
            pp[0 * capk + loc] = a00; pp[1 * capk + loc] = a01; pp[2 * capk + loc] = a02;
            pp[3 * capk + loc] = as0; pp[4 * capk + loc] = as1; pp[5 * capk + loc] = as2;
        }
    }
}

// FP32 source records, one per row: position relative to its tile's origin
// (the first active row of the tile, FP64), c = |e|^2, xe = x x e (from the
// FP64 position and segment), e in FP32, and the row.  The far loop forms
// sum (x - t) x e fac as sum x x e fac - t x sum e fac.  Position and c come
// first: the near loop reads them (one 16-byte read) for every pair and the
// rest only for pairs beyond the near radius.
struct __align__(16) Rec32 {
    float x0, x1, x2, c;
    float xe0, xe1, xe2, e0;
    float e1, e2, pad;
    int j;
};
static_assert(sizeof(Rec32) == 48, "Rec32 is twelve words");

// Spatial order of the FP32 source records.  Within each tangle, active
// rows by the Morton code of their cell (10 bits per axis over the periodic
// box), then the inactive rows in row order; key = tangle << 32 | code.
// Any order is correct (records carry their row); a compact tile keeps the
// FP32 positions small relative to its origin and makes more tiles far.
__global__ void k_order_keys(DState s, Tangles tg, int rows, double bx, double by, double bz,
                             unsigned long long *__restrict__ keys, int *__restrict__ vals) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int k = tg.row2t[r];
    unsigned code;
    if (s.infront[r] != 0) {
        const double bb[3] = {bx, by, bz};
        unsigned c[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double u = ld(s.x, s.cap, a, r) / bb[a] + 0.5;
            const int q = (int)floor(u * 1024.0);
            c[a] = (unsigned)(q < 0 ? 0 : q > 1023 ? 1023 : q);
        }
        code = nv_spread10(c[0]) | (nv_spread10(c[1]) << 1) | (nv_spread10(c[2]) << 2);
    } else {
        code = 0x80000000u | (unsigned)(r - tg.base[k]);
    }
    keys[r] = ((unsigned long long)(unsigned)k << 32) | code;
    vals[r] = r;
}

__global__ void k_order_inverse(const int *__restrict__ perm, int rows, int *__restrict__ iperm) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < rows) iperm[perm[q]] = q;
}

// Per source tile: the FP64 origin its FP32 records are relative to, and
// a sphere (FP32, relative to the origin) holding its active sources;
// rad < 0 marks a tile without active sources.
constexpr int SUB = 16;     // sources per sub-block of a tile
template <int TSZ>
struct TileInfo {
    double o[3];
    float c[3], rad;
    // bounding sphere of each sub-block of SUB sources (centre xyz, radius;
    // radius -1 where the sub-block holds no active source)
    float sub[TSZ / SUB][4];
};

__device__ __forceinline__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

// One workgroup per tile of TSZ records (rows is a multiple of TSZ); record
// q of the tile sequence holds row perm[q].
template <int TSZ>
__global__ __launch_bounds__(TSZ) void k_src32(DState s, int rows, const int *__restrict__ perm,
                                               Rec32 *__restrict__ rec, TileInfo<TSZ> *__restrict__ tiles) {
    __shared__ float red[TSZ / 64][6];
    __shared__ int first;
    const int q0 = blockIdx.x * TSZ + threadIdx.x;
    const int r = q0 < rows ? perm[q0] : 0;
    const bool act = q0 < rows && is_source(s, r);
    // origin: the row of the tile's first active record
    if (threadIdx.x == 0) first = 0x7fffffff;
    __syncthreads();
    if (act) atomicMin(&first, q0);
    __syncthreads();
    double o[3] = {0.0, 0.0, 0.0};
    if (first != 0x7fffffff) {
        const int rf = perm[first];
        o[0] = ld(s.x, s.cap, 0, rf); o[1] = ld(s.x, s.cap, 1, rf); o[2] = ld(s.x, s.cap, 2, rf);
    }
    Rec32 v;
    if (act) {
        double x[3], e[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = ld(s.x, s.cap, k, r);
            e[k] = ld(s.gi, s.cap, k, r) - x[k];
        }
        const double xr[3] = {x[0] - o[0], x[1] - o[1], x[2] - o[2]};
        v.x0 = (float)xr[0]; v.x1 = (float)xr[1]; v.x2 = (float)xr[2];
        v.e0 = (float)e[0]; v.e1 = (float)e[1]; v.e2 = (float)e[2];
        v.c = (float)(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        v.j = r;
        v.xe0 = (float)(xr[1] * e[2] - xr[2] * e[1]);
        v.xe1 = (float)(xr[2] * e[0] - xr[0] * e[2]);
        v.xe2 = (float)(xr[0] * e[1] - xr[1] * e[0]);
    } else {
        // a neutral record: far away, zero segment, never matched
        v.x0 = v.x1 = v.x2 = 1.0e18f;
        v.e0 = v.e1 = v.e2 = 0.0f;
        v.c = 0.0f;
        v.j = -1;
        v.xe0 = v.xe1 = v.xe2 = 0.0f;
    }
    v.pad = 0.0f;
    if (q0 < rows) rec[q0] = v;
    // bounding box of the active sources, then its circumscribed sphere
    float lo[3], hi[3];
    lo[0] = act ? v.x0 : 3.0e38f; lo[1] = act ? v.x1 : 3.0e38f; lo[2] = act ? v.x2 : 3.0e38f;
    hi[0] = act ? v.x0 : -3.0e38f; hi[1] = act ? v.x1 : -3.0e38f; hi[2] = act ? v.x2 : -3.0e38f;
    {
        // the same for each sub-block of SUB consecutive records
        float sl[3], sh_[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            sl[k] = lo[k];
            sh_[k] = hi[k];
            for (int off = SUB / 2; off > 0; off >>= 1) {
                sl[k] = fminf(sl[k], __shfl_xor(sl[k], off));
                sh_[k] = fmaxf(sh_[k], __shfl_xor(sh_[k], off));
            }
        }
        if (q0 < rows && (threadIdx.x % SUB) == 0) {
            float *o4 = tiles[blockIdx.x].sub[threadIdx.x / SUB];
            if (sl[0] > sh_[0]) {
                o4[0] = o4[1] = o4[2] = 0.f;
                o4[3] = -1.f;
            } else {
                float d2 = 0.f;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    o4[k] = 0.5f * (sl[k] + sh_[k]);
                    const float h = sh_[k] - sl[k];
                    d2 += h * h;
                }
                o4[3] = 0.5f * sqrtf(d2) * 1.00001f;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
    }
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { red[w][k] = lo[k]; red[w][3 + k] = hi[k]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < TSZ / 64; ++q)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(lo[k], red[q][k]);
                hi[k] = fmaxf(hi[k], red[q][3 + k]);
            }
        TileInfo<TSZ> &ti = tiles[blockIdx.x];
        ti.o[0] = o[0]; ti.o[1] = o[1]; ti.o[2] = o[2];
        if (lo[0] > hi[0]) {
            ti.c[0] = ti.c[1] = ti.c[2] = 0.f;
            ti.rad = -1.f;
        } else {
            float d2 = 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                ti.c[k] = 0.5f * (lo[k] + hi[k]);
                const float h = hi[k] - lo[k];
                d2 += h * h;
            }
            // half the diagonal, rounded up by a relative 1e-5
            ti.rad = 0.5f * sqrtf(d2) * 1.00001f;
        }
    }
}

__device__ __forceinline__ void kahan_add(float &sum, float &comp, float v) {
    float y = v - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// One block of KB sources, far from every target of the wave: no pair needs
// the near-field test or an exclusion, every term is FP32.  With d = x - t,
// sum d x e fac = sum (x x e) fac - t x sum e fac: six FMAs per pair instead
// of the cross product's nine operations, and one cross product per target
// and block.
template <int TPT, int KB>
__device__ __forceinline__ void far_block(const Rec32 *tile, int jb, const float (&tt)[TPT][3], float (&f)[TPT][3],
                                          float (&k)[TPT][3]) {
    float A[TPT][3], B[TPT][3];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        A[q][0] = A[q][1] = A[q][2] = 0.f;
        B[q][0] = B[q][1] = B[q][2] = 0.f;
    }
#pragma unroll
    for (int jj = 0; jj < KB; ++jj) {
        const Rec32 r = tile[jb + jj];
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const float d0 = r.x0 - tt[q][0], d1 = r.x1 - tt[q][1], d2 = r.x2 - tt[q][2];
            const float r1sq = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
            const float de = fmaf(d0, r.e0, fmaf(d1, r.e1, d2 * r.e2));
            const float r2sq = fmaf(2.0f, de, r1sq + r.c);
            const float inv1 = __builtin_amdgcn_rsqf(r1sq);
            const float inv2 = __builtin_amdgcn_rsqf(r2sq);
            const float den = fmaf(r1sq * inv1, r2sq * inv2, r1sq + de);
            const float fac = (inv1 + inv2) * __builtin_amdgcn_rcpf(den);
            A[q][0] = fmaf(r.xe0, fac, A[q][0]);
            A[q][1] = fmaf(r.xe1, fac, A[q][1]);
            A[q][2] = fmaf(r.xe2, fac, A[q][2]);
            B[q][0] = fmaf(r.e0, fac, B[q][0]);
            B[q][1] = fmaf(r.e1, fac, B[q][1]);
            B[q][2] = fmaf(r.e2, fac, B[q][2]);
        }
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const float *t = tt[q];
        kahan_add(f[q][0], k[q][0], A[q][0] - fmaf(t[1], B[q][2], -(t[2] * B[q][1])));
        kahan_add(f[q][1], k[q][1], A[q][1] - fmaf(t[2], B[q][0], -(t[0] * B[q][2])));
        kahan_add(f[q][2], k[q][2], A[q][2] - fmaf(t[0], B[q][1], -(t[1] * B[q][0])));
    }
}

// FP32 compensated far field + FP64 near field.  Far pairs use the
// well-conditioned form u = k/4pi (d x e) (1/r1 + 1/r2) / (r1 r2 + R1.R2)
// with R1 = d, R2 = d + e (algebraically equal to the reference's
// expression, without its two cancellations); d is accurate to FP32
// rounding of |d| because positions are stored relative to their tile's
// origin.  Terms are accumulated in FP32, compensated (Kahan) every KB terms.
// Pairs closer than r_near, and the pairs the reference excludes, go
// through the reference's FP64 arithmetic instead.  Each lane carries TPT
// target points, so one LDS record read serves TPT pairs.
// Partial block of tangle k: [S][6][cap_k] = far xyz (times k/4pi), near xyz.
// Waves per SIMD the compiler is held to: 4 with two targets per lane (128
// VGPRs; a few spills outside the far loop), 5 with one (96 VGPRs).  On the
// 256-tangle ensemble: two per lane 943 ms at 3 waves, 915 ms at 4; one per
// lane 834 ms at 5 waves (and the same at 6, where it spills).
#ifndef NV_BS32_WAVES_PER_EU
#define NV_BS32_WAVES_PER_EU(TPT) ((TPT) == 1 ? 5 : 4)
#endif
template <int TPT, int KB, int TSZ>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(NV_BS32_WAVES_PER_EU(TPT)))) void k_bs32c(DState s, Tangles tg, const Item *__restrict__ items, int nitems,
                                              int *queue, const Rec32 *__restrict__ rec,
                                              const TileInfo<TSZ> *__restrict__ tiles, const int *__restrict__ iperm,
                                              const double *__restrict__ rec64,
                                              Shifts sh, double kappa, float rn2, double *__restrict__ part) {
    __shared__ Rec32 tile[TSZ];
    __shared__ int slot;
    const int tid = threadIdx.x;
    const float rnf = sqrtf(rn2);
    for (;;) {
        const int it = next_item(queue, &slot);
        if (it >= nitems) break;
        const Item I = items[it];
        const int base = tg.base[I.k], n = tg.n[I.k], capk = tg.cap[I.k], S = tg.S[I.k];
        if (I.tb * TB * TPT >= n) continue;   // item listed for rows the tangle does not fill (yet)
        int ii[TPT], loc[TPT], bi[TPT], bpos[TPT];
        bool act[TPT];
        double xi[TPT][3];
        float f[TPT][3], k[TPT][3];
        double nr[TPT][3];
        bool any = false;
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            loc[q] = I.tb * TB * TPT + q * TB + tid;
            ii[q] = base + loc[q];
            act[q] = (loc[q] < n) && (s.infront[ii[q]] != 0);
            bi[q] = -2;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                xi[q][c] = act[q] ? ld(s.x, s.cap, c, ii[q]) : 0.0;
                f[q][c] = 0.f;
                k[q][c] = 0.f;
                nr[q][c] = 0.0;
            }
            if (act[q]) bi[q] = s.behind[ii[q]] - 1;
            bpos[q] = bi[q] >= 0 ? iperm[bi[q]] : -1;   // record position of the behind row
            any = any || act[q];
        }
        const int ntiles = (n + TSZ - 1) / TSZ;
        const long U = (long)sh.n * ntiles;
        const long u0 = U * I.c / S, u1 = U * (I.c + 1) / S;
        for (long un = u0; un < u1; ++un) {
            const int m = (int)(un / ntiles), t = (int)(un % ntiles);
            const int gt = base / TSZ + t;
            __syncthreads();
            for (int q = tid; q < TSZ; q += TB) tile[q] = rec[(size_t)gt * TSZ + q];
            __syncthreads();
            if (!any) continue;
            const TileInfo<TSZ> &ti = tiles[gt];
            if (ti.rad < 0.f) continue;     // no active source in this tile
            const double s0 = sh.s[m][0], s1 = sh.s[m][1], s2 = sh.s[m][2];
            float tt[TPT][3];
            bool far = true;
#pragma unroll
            for (int q = 0; q < TPT; ++q) {
                tt[q][0] = (float)((xi[q][0] - s0) - ti.o[0]);
                tt[q][1] = (float)((xi[q][1] - s1) - ti.o[1]);
                tt[q][2] = (float)((xi[q][2] - s2) - ti.o[2]);
                // every source of the tile is at least rn away from this
                // target (with a relative margin far above FP32 rounding), and
                // the target's behind point is not in the tile
                const float c0 = tt[q][0] - ti.c[0], c1 = tt[q][1] - ti.c[1], c2 = tt[q][2] - ti.c[2];
                const float dd = c0 * c0 + c1 * c1 + c2 * c2;
                const float mag = fabsf(tt[q][0]) + fabsf(tt[q][1]) + fabsf(tt[q][2]) + ti.rad;
                const float lim = (ti.rad + rnf) * 1.0001f + 1.0e-6f * mag;
                const bool fq = dd > lim * lim && (unsigned)(bpos[q] - gt * TSZ) >= (unsigned)TSZ;
                far = far && (fq || !act[q]);
            }
            if (__all(far)) {
                // far tile for the whole wave: no pair needs the near-field
                // test or an exclusion, every term is FP32
                for (int jb = 0; jb < TSZ; jb += KB) far_block<TPT, KB>(tile, jb, tt, f, k);
                continue;
            }
            for (int jb = 0; jb < TSZ; jb += KB) {
                // A block of KB sources that is far for the whole wave (its
                // sub-block spheres tested as the tile's was) takes the far
                // path: most blocks of a tile that is near some lane are far
                // from every lane.
                bool bfar = true;
#pragma unroll
                for (int sb = jb / SUB; sb * SUB < jb + KB; ++sb) {
                    const float *si = ti.sub[sb];
                    const float srad = si[3];
                    if (srad < 0.f) continue;   // no active source
#pragma unroll
                    for (int q = 0; q < TPT; ++q) {
                        const float c0 = tt[q][0] - si[0], c1 = tt[q][1] - si[1], c2 = tt[q][2] - si[2];
                        const float dd = c0 * c0 + c1 * c1 + c2 * c2;
                        const float mag = fabsf(tt[q][0]) + fabsf(tt[q][1]) + fabsf(tt[q][2]) + srad;
                        const float lim = (srad + rnf) * 1.0001f + 1.0e-6f * mag;
                        const bool fq = dd > lim * lim && (unsigned)(bpos[q] - (gt * TSZ + sb * SUB)) >= (unsigned)SUB;
                        bfar = bfar && (fq || !act[q]);
                    }
                }
                if (__all(bfar)) {
                    far_block<TPT, KB>(tile, jb, tt, f, k);
                    continue;
                }
                float a[TPT][3];
#pragma unroll
                for (int q = 0; q < TPT; ++q) a[q][0] = a[q][1] = a[q][2] = 0.f;
#pragma unroll
                for (int jj = 0; jj < KB; ++jj) {
                    const Rec32 r = tile[jb + jj];
#pragma unroll
                    for (int q = 0; q < TPT; ++q) {
                        const float d0 = r.x0 - tt[q][0], d1 = r.x1 - tt[q][1], d2 = r.x2 - tt[q][2];
                        const float r1sq = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
                        if (r1sq < rn2 || r.j == bi[q]) {
                            // near field or excluded pair: reference arithmetic in FP64
                            const int j = r.j;
                            if (!act[q] || j < 0 || j == bi[q] || (m == 0 && j == ii[q])) continue;
                            const double *qv = &rec64[(size_t)j * REC64];
                            double out[3];
                            if (bs_term64((qv[0] + s0) - xi[q][0], (qv[1] + s1) - xi[q][1],
                                          (qv[2] + s2) - xi[q][2], qv[3], qv[4], qv[5], qv[6], kappa, out)) {
                                nr[q][0] += out[0]; nr[q][1] += out[1]; nr[q][2] += out[2];
                            }
                            continue;
                        }
                        const float de = fmaf(d0, r.e0, fmaf(d1, r.e1, d2 * r.e2));
                        const float r2sq = fmaf(2.0f, de, r1sq + r.c);
                        const float inv1 = __builtin_amdgcn_rsqf(r1sq);
                        const float inv2 = __builtin_amdgcn_rsqf(r2sq);
                        const float den = fmaf(r1sq * inv1, r2sq * inv2, r1sq + de);
                        const float fac = (inv1 + inv2) * __builtin_amdgcn_rcpf(den);
                        a[q][0] = fmaf(fmaf(d1, r.e2, -(d2 * r.e1)), fac, a[q][0]);
                        a[q][1] = fmaf(fmaf(d2, r.e0, -(d0 * r.e2)), fac, a[q][1]);
                        a[q][2] = fmaf(fmaf(d0, r.e1, -(d1 * r.e0)), fac, a[q][2]);
                    }
                }
#pragma unroll
                for (int q = 0; q < TPT; ++q) {
                    kahan_add(f[q][0], k[q][0], a[q][0]);
                    kahan_add(f[q][1], k[q][1], a[q][1]);
                    kahan_add(f[q][2], k[q][2], a[q][2]);
                }
            }
        }
        const double k4 = kappa / (4.0 * NV_PI);
        double *pp = part + tg.poff[I.k] + (size_t)I.c * 6 * capk;
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            if (loc[q] >= capk) continue;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                pp[cc * capk + loc[q]] = ((double)f[q][cc] - (double)k[q][cc]) * k4;
                pp[(3 + cc) * capk + loc[q]] = nr[q][cc];
            }
        }
    }
}

// Reduce chunk partials in chunk order and finish every point.
// mode 0: partials hold (acc0, accs) of k_bs64; mode 1: (far, near) of
// k_bs32c; mode 2: no partials (LIA); mode 3: the tree walks' [image][3][rows],
// image 0 the unshifted sum, the others the images' (added in image order,
// the same additions of the same values as mode 0 with the zero halves);
// mode 4: the FP32 walk's, image 0 as in mode 3 and images 1.. as floats
// ([image - 1][3][rows] after image 0's block).
__device__ __forceinline__ void finish_body(const DState &s, const Phys &p, const Tangles &tg, int rows,
                                            const double *__restrict__ part, int mode, int do_ab, DCtl *ctl,
                                            const double *__restrict__ mir) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = r < rows && s.infront[r] != 0;
    // active points counted per workgroup, one atomic each: atomics on one
    // address are serialised in L2, one per wave cost ~12 ns each
    __shared__ int wcount[4];
    const unsigned long long msk = __ballot(act);
    if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = (int)__popcll(msk);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int c = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        if (c) atomicAdd(&ctl->nvel, c);
    }
    if (r >= rows) return;
    if (!act) {
        if (do_ab) {
#pragma unroll
            for (int k = 0; k < 3; ++k) s.xn[k * s.cap + r] = s.x[k * s.cap + r];
        }
        return;
    }
    double ub0[3] = {0.0, 0.0, 0.0}, ubs[3] = {0.0, 0.0, 0.0};
    if (mode != 2) {
        const int kk = tg.row2t[r];
        const int capk = tg.cap[kk], loc = r - tg.base[kk], S = tg.S[kk];
        const double *pb = part + tg.poff[kk];
        if (mode == 3) {
#pragma unroll
            for (int k = 0; k < 3; ++k) ub0[k] = 0.0 + pb[k * capk + loc];
#pragma unroll 9
            for (int c = 1; c < S; ++c) {
                const double *pp = pb + (size_t)c * 3 * capk;
#pragma unroll
                for (int k = 0; k < 3; ++k) ubs[k] = ubs[k] + pp[k * capk + loc];
            }
        } else if (mode == 4) {
#pragma unroll
            for (int k = 0; k < 3; ++k) ub0[k] = 0.0 + pb[k * capk + loc];
            const float *pf = (const float *)(pb + (size_t)3 * capk);
#pragma unroll 9
            for (int c = 1; c < S; ++c) {
                const float *pp = pf + (size_t)(c - 1) * 3 * capk;
#pragma unroll
                for (int k = 0; k < 3; ++k) ubs[k] = ubs[k] + (double)pp[k * capk + loc];
            }
        } else if (mode == 0) {
            // unrolled so that the loads of several chunks are in flight; the
            // adds stay in chunk order
#pragma unroll 9
            for (int c = 0; c < S; ++c) {
                const double *pp = pb + (size_t)c * 6 * capk;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    ub0[k] = ub0[k] + pp[k * capk + loc];
                    ubs[k] = ubs[k] + pp[(3 + k) * capk + loc];
                }
            }
        } else {
            double far[3] = {0, 0, 0}, nr[3] = {0, 0, 0};
#pragma unroll 9
            for (int c = 0; c < S; ++c) {
                const double *pp = pb + (size_t)c * 6 * capk;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    far[k] += pp[k * capk + loc];
                    nr[k] += pp[(3 + k) * capk + loc];
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) ub0[k] = nr[k] + far[k];
        }
    }
    // the image vortices' sum of solid walls (src/timestep.f90:94-98), zero
    // without walls: the reference adds u_mir either way
    double ubm[3] = {0.0, 0.0, 0.0};
    if (mir) {
#pragma unroll
        for (int k = 0; k < 3; ++k) ubm[k] = mir[(size_t)k * rows + r];
    }
    const double noise = (p.nf == NV_NF_NOISEY_XFLOW && p.noise_n > 0) ? p.noise[ctl->noise_ctr] : 0.0;
    finish_point(s, p, r, ub0, ubs, ubm, ctl->t, noise, do_ab);
}

__global__ __launch_bounds__(256) void k_finish(DState s, Phys p, Tangles tg, int rows, const double *__restrict__ part,
                                                int mode, int do_ab, DCtl *ctl, const double *__restrict__ mir) {
    if (ctl->abort) return;
    finish_body(s, p, tg, rows, part, mode, do_ab, ctl, mir);
}

// The tree walks' layouts (modes 3 and 4, 27 chunks of 3 values) at eight waves per
// SIMD (64 VGPRs, no spills): a point streams about 0.9 KB through 36
// component arrays, and more waves hide the memory latency (389 -> 298 us at
// 1 M points against the compiler's 110 VGPRs and 4 waves). The chunked
// direct-sum layouts (hundreds of chunks per point) are slower so.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_finish_tree(
    DState s, Phys p, Tangles tg, int rows, const double *__restrict__ part, int mode, int do_ab, DCtl *ctl,
    const double *__restrict__ mir) {
    if (ctl->abort) return;   // (a recorded octree step whose build was incomplete)
    finish_body(s, p, tg, rows, part, mode, do_ab, ctl, mir);
}

// src/normal_fluid.f90:340-347 randomise_normal_fluid: the noise counter
// moves to the next value, back to the first after the last
__global__ void k_noise_tick(DCtl *ctl, int n) {
    if (ctl->abort) return;
    int c = ctl->noise_ctr + 1;
    if (c >= n) c = 0;
    ctl->noise_ctr = c;
}

// Reads every chunk partial of the active rows the way k_finish does, and
// writes nothing unless a sum is NaN: the reduction's cost, for the
// autotuner (the partial volume grows with the chunk count).
__global__ void k_reduce_probe(DState s, Tangles tg, int rows, const double *__restrict__ part,
                               double *__restrict__ sink) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows || s.infront[r] == 0) return;
    const int kk = tg.row2t[r];
    const int capk = tg.cap[kk], loc = r - tg.base[kk], S = tg.S[kk];
    const double *pb = part + tg.poff[kk];
    double acc = 0.0;
    for (int c = 0; c < S; ++c) {
        const double *pp = pb + (size_t)c * 6 * capk;
#pragma unroll
        for (int k = 0; k < 6; ++k) acc += pp[k * capk + loc];
    }
    if (acc != acc) sink[0] = acc;
}

}  // namespace

hipError_t nvk_reduce_probe(hipStream_t st, DState s, Tangles tg, int rows, const double *part, double *sink) {
    if (rows > 0)
        hipLaunchKernelGGL(k_reduce_probe, dim3((rows + 255) / 256), dim3(256), 0, st, s, tg, rows, part, sink);
    return hipGetLastError();
}

// ------------------------------------------------------------------ launch
size_t nvk_rec32_bytes() { return sizeof(Rec32); }

// resident workgroups per CU, which sizes the persistent grids
int nvk_occupancy_bs64() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_bs64, TB, 0) != hipSuccess || b < 1) b = 2;
    return b;
}

template <int TPT, int KB, int TSZ>
int occ_bs32c() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_bs32c<TPT, KB, TSZ>, TB, 0) != hipSuccess || b < 1) b = 2;
    return b;
}

// The FP32C variants: points per lane (1, 2), Kahan block (1, 2, 4, 8, 16,
// 32; 4 and up with tiles of 128 sources) and sources per LDS tile (128,
// 256).  F(TPT, KB, TSZ) is expanded for the variant asked for.
#define NV_BS32_DISPATCH(F, tpt, kb, ts)                                         \
    do {                                                                         \
        const int kb_ = (ts) == 128 && (kb) < 4 ? 4 : (kb);                      \
        if ((ts) == 128) {                                                       \
            if ((tpt) == 2) {                                                    \
                if (kb_ == 4) F(2, 4, 128); else if (kb_ == 16) F(2, 16, 128);   \
                else if (kb_ == 32) F(2, 32, 128); else F(2, 8, 128);            \
            } else {                                                             \
                if (kb_ == 4) F(1, 4, 128); else if (kb_ == 16) F(1, 16, 128);   \
                else if (kb_ == 32) F(1, 32, 128); else F(1, 8, 128);            \
            }                                                                    \
        } else if ((tpt) == 2) {                                                 \
            if (kb_ == 8) F(2, 8, 256); else if (kb_ == 16) F(2, 16, 256);       \
            else if (kb_ == 32) F(2, 32, 256); else if (kb_ == 4) F(2, 4, 256);  \
            else if (kb_ == 2) F(2, 2, 256); else F(2, 1, 256);                  \
        } else {                                                                 \
            if (kb_ == 8) F(1, 8, 256); else if (kb_ == 16) F(1, 16, 256);       \
            else if (kb_ == 32) F(1, 32, 256); else if (kb_ == 4) F(1, 4, 256);  \
            else if (kb_ == 2) F(1, 2, 256); else F(1, 1, 256);                  \
        }                                                                        \
    } while (0)

int nvk_occupancy_bs32c(int tpt, int kb, int ts) {
    int b = 2;
#define NV_OCC(T_, K_, S_) b = occ_bs32c<T_, K_, S_>()
    NV_BS32_DISPATCH(NV_OCC, tpt, kb, ts);
#undef NV_OCC
    return b;
}

hipError_t nvk_src64(hipStream_t st, DState s, int rows, double *rec) {
    if (rows > 0) hipLaunchKernelGGL(k_src64, dim3((rows + 255) / 256), dim3(256), 0, st, s, rows, rec);
    return hipGetLastError();
}

hipError_t nvk_bs64(hipStream_t st, DState s, Tangles tg, const Item *items, int nitems, int *queue, int grid,
                    const double *rec, const Shifts *sh, double kappa, double *part) {
    hipLaunchKernelGGL(k_bs64, dim3(grid), dim3(TB), 0, st, s, tg, items, nitems, queue, rec, *sh, kappa, part);
    return hipGetLastError();
}

// the tile records for `rows` rows with either tile size
size_t nvk_tileinfo_bytes(int rows) {
    const size_t a = sizeof(TileInfo<256>) * (size_t)(rows / 256 + 1);
    const size_t b = sizeof(TileInfo<128>) * (size_t)(rows / 128 + 1);
    return a > b ? a : b;
}

hipError_t nvk_order(hipStream_t st, DState s, Tangles tg, int rows, const double *box, unsigned long long *k0,
                     int *v0, unsigned long long *k1, int *v1, int *iperm, int *hist, int *scan_tmp) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_order_keys, dim3((rows + 255) / 256), dim3(256), 0, st, s, tg, rows, box[0], box[1], box[2],
                       k0, v0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int bits = tg.K <= 1 ? 32 : 48;
    if ((e = nv_radix_sort_pairs(st, k0, v0, k1, v1, rows, bits, hist, scan_tmp)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_order_inverse, dim3((rows + 255) / 256), dim3(256), 0, st, v0, rows, iperm);
    return hipGetLastError();
}

hipError_t nvk_src32(hipStream_t st, DState s, int rows, const int *perm, void *rec, void *tiles, int ts) {
    if (rows <= 0) return hipGetLastError();
    if (ts == 128)
        hipLaunchKernelGGL(k_src32<128>, dim3((rows + 127) / 128), dim3(128), 0, st, s, rows, perm, (Rec32 *)rec,
                           (TileInfo<128> *)tiles);
    else
        hipLaunchKernelGGL(k_src32<256>, dim3((rows + 255) / 256), dim3(256), 0, st, s, rows, perm, (Rec32 *)rec,
                           (TileInfo<256> *)tiles);
    return hipGetLastError();
}

hipError_t nvk_bs32c(hipStream_t st, DState s, Tangles tg, const Item *items, int nitems, int *queue, int grid,
                     const void *rec, const void *tiles, const int *iperm, const double *rec64, const Shifts *sh,
                     double kappa, float rn2, double *part, int tpt, int kb, int ts) {
    const Rec32 *r = (const Rec32 *)rec;
#define NV_BS32(T_, K_, S_)                                                                                     \
    hipLaunchKernelGGL((k_bs32c<T_, K_, S_>), dim3(grid), dim3(TB), 0, st, s, tg, items, nitems, queue, r,        \
                       (const TileInfo<S_> *)tiles, iperm, rec64, *sh, kappa, rn2, part)
    NV_BS32_DISPATCH(NV_BS32, tpt, kb, ts);
#undef NV_BS32
    return hipGetLastError();
}

hipError_t nvk_finish(hipStream_t st, DState s, Phys p, Tangles tg, int rows, const double *part, int mode,
                      int do_ab, DCtl *ctl, const double *mir) {
    if (rows > 0) {
        if (mode == 3 || mode == 4)
            hipLaunchKernelGGL(k_finish_tree, dim3((rows + 255) / 256), dim3(256), 0, st, s, p, tg, rows, part, mode,
                               do_ab, ctl, mir);
        else
            hipLaunchKernelGGL(k_finish, dim3((rows + 255) / 256), dim3(256), 0, st, s, p, tg, rows, part, mode,
                               do_ab, ctl, mir);
    }
    // the end of pmotion advances the noise table (src/timestep.f90:44)
    if (do_ab && p.nf == NV_NF_NOISEY_XFLOW && p.noise_n > 0)
        hipLaunchKernelGGL(k_noise_tick, dim3(1), dim3(1), 0, st, ctl, p.noise_n);
    return hipGetLastError();
}
