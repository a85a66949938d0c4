// Direct Biot-Savart sum over straight segments with the 26 periodic images
// (src/timestep.f90:74-98 with biot_savart :151-172 and biot_savart_shift
// :175-198), fused with the local term, mutual friction and the
// Adams-Bashforth update (src/timestep.f90:16-45).
//
// Work decomposition: targets are split into blocks of BT rows (one row per
// thread); the (image, source tile) sequence of each target -- image-major,
// tiles of T sources in ascending row order, exactly the order the
// reference visits them -- is split into S contiguous chunks, one per
// blockIdx.y.  Each block stages a tile of sources in LDS (coalesced SoA
// reads from HBM/L2), every lane then reads the same LDS record (broadcast),
// so the inner loop is pure VALU.  Chunk partial sums go to a partial
// buffer that k_finish reduces in chunk order.  With S = 1 the summation
// order per target is the reference's, term for term.
#include "nv_kernels.h"

namespace {

constexpr int REC64 = 8;   // x0 x1 x2 e0 e1 e2 c valid

// FP64 source records, one per row: x, e = ghosti - x, c = |e|^2
// (dist_gen_sq(ghosti, x), src/timestep.f90:164), valid = infront != 0.
__global__ void k_src64(DState s, int n, int npad, double *__restrict__ rec) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npad) return;
    double v[REC64] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (r < n && s.infront[r] != 0) {
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

// FP64 plain direct sum.  part layout: [S][6][npad] (acc0 xyz, accs xyz).
template <int BT, int T>
__global__ __launch_bounds__(BT) void k_bs64(DState s, int n, int npad, const double *__restrict__ rec, Shifts sh,
                                             int ntiles, int S, double kappa, double *__restrict__ part) {
    __shared__ double tile[T * REC64];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * BT + tid;
    const bool act = (i < n) && (s.infront[i] != 0);
    double xi0 = 0, xi1 = 0, xi2 = 0;
    int bi = -2;
    if (act) {
        xi0 = ld(s.x, s.cap, 0, i);
        xi1 = ld(s.x, s.cap, 1, i);
        xi2 = ld(s.x, s.cap, 2, i);
        bi = s.behind[i] - 1;
    }
    double a00 = 0, a01 = 0, a02 = 0, as0 = 0, as1 = 0, as2 = 0;
    const long U = (long)sh.n * ntiles;
    const int c = blockIdx.y;
    const long u0 = U * c / S, u1 = U * (c + 1) / S;
    for (long un = u0; un < u1; ++un) {
        const int m = (int)(un / ntiles), t = (int)(un % ntiles);
        const double s0 = sh.s[m][0], s1 = sh.s[m][1], s2 = sh.s[m][2];
        __syncthreads();
        for (int q = tid; q < T * REC64; q += BT) tile[q] = rec[(size_t)t * T * REC64 + q];
        __syncthreads();
        if (!act) continue;
        const int base = t * T;
        for (int jj = 0; jj < T; ++jj) {
            const double *qv = &tile[jj * REC64];
            if (qv[7] == 0.0) continue;
            const int j = base + jj;
            if (j == bi || (m == 0 && j == i)) continue;
            double out[3];
            if (!bs_term64((qv[0] + s0) - xi0, (qv[1] + s1) - xi1, (qv[2] + s2) - xi2, qv[3], qv[4], qv[5], qv[6],
                           kappa, out))
                continue;
            // keep the reference's running-sum order: accumulate straight
            // into the image accumulator, term by term
            if (m == 0) { a00 = a00 + out[0]; a01 = a01 + out[1]; a02 = a02 + out[2]; }
            else { as0 = as0 + out[0]; as1 = as1 + out[1]; as2 = as2 + out[2]; }
        }
    }
    if (i < npad) {
        double *pp = part + (size_t)c * 6 * npad;
        pp[0 * npad + i] = a00; pp[1 * npad + i] = a01; pp[2 * npad + i] = a02;
        pp[3 * npad + i] = as0; pp[4 * npad + i] = as1; pp[5 * npad + i] = as2;
    }
}

// FP32 source records, one per row: position relative to its tile's origin
// (the first row of the tile, FP64), e in FP32, c = |e|^2 in FP32, row.
struct __align__(16) Rec32 {
    float x0, x1, x2, e0, e1, e2, c;
    int j;
};

__global__ void k_src32(DState s, int n, int npad, int T, Rec32 *__restrict__ rec, double *__restrict__ origin) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npad) return;
    int t = r / T, r0 = t * T;
    double o[3] = {0.0, 0.0, 0.0};
    // origin: first active row of the tile (fall back to 0)
    for (int q = r0; q < r0 + T && q < n; ++q) {
        if (s.infront[q] != 0) {
            o[0] = ld(s.x, s.cap, 0, q); o[1] = ld(s.x, s.cap, 1, q); o[2] = ld(s.x, s.cap, 2, q);
            break;
        }
    }
    if (r == r0) { origin[t * 3 + 0] = o[0]; origin[t * 3 + 1] = o[1]; origin[t * 3 + 2] = o[2]; }
    Rec32 v;
    if (r < n && s.infront[r] != 0) {
        double x[3], e[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = ld(s.x, s.cap, k, r);
            e[k] = ld(s.gi, s.cap, k, r) - x[k];
        }
        v.x0 = (float)(x[0] - o[0]); v.x1 = (float)(x[1] - o[1]); v.x2 = (float)(x[2] - o[2]);
        v.e0 = (float)e[0]; v.e1 = (float)e[1]; v.e2 = (float)e[2];
        v.c = (float)(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        v.j = r;
    } else {
        // a neutral record: far away, zero segment, never matched
        v.x0 = v.x1 = v.x2 = 1.0e18f;
        v.e0 = v.e1 = v.e2 = 0.0f;
        v.c = 0.0f;
        v.j = -1;
    }
    rec[r] = v;
}

__device__ __forceinline__ void kahan_add(float &sum, float &comp, float v) {
    float y = v - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// FP32 compensated far field + FP64 near field.  Far pairs use the
// well-conditioned form u = k/4pi (d x e)(r1+r2) / (r1 r2 (r1 r2 + R1.R2))
// (algebraically equal to the reference's expression, without its two
// cancellations) with d accurate to FP32 rounding of |d| thanks to
// tile-local coordinates, and Kahan-compensated FP32 accumulation.  Pairs
// with |d| < r_near, and the reference's excluded pairs, go through the FP64
// reference arithmetic instead.  part: [S][6][npad] = far xyz (already times
// k/4pi), near xyz.
template <int BT, int T>
__global__ __launch_bounds__(BT) void k_bs32c(DState s, int n, int npad, const Rec32 *__restrict__ rec,
                                              const double *__restrict__ origin, const double *__restrict__ rec64,
                                              Shifts sh, int ntiles, int S, double kappa, float rn2,
                                              double *__restrict__ part) {
    __shared__ Rec32 tile[T];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * BT + tid;
    const bool act = (i < n) && (s.infront[i] != 0);
    double xi0 = 0, xi1 = 0, xi2 = 0;
    int bi = -2;
    if (act) {
        xi0 = ld(s.x, s.cap, 0, i);
        xi1 = ld(s.x, s.cap, 1, i);
        xi2 = ld(s.x, s.cap, 2, i);
        bi = s.behind[i] - 1;
    }
    float f0 = 0.f, f1 = 0.f, f2 = 0.f, k0 = 0.f, k1 = 0.f, k2 = 0.f;
    double n0 = 0, n1 = 0, n2 = 0;
    const long U = (long)sh.n * ntiles;
    const int c = blockIdx.y;
    const long u0 = U * c / S, u1 = U * (c + 1) / S;
    for (long un = u0; un < u1; ++un) {
        const int m = (int)(un / ntiles), t = (int)(un % ntiles);
        __syncthreads();
        for (int q = tid; q < T; q += BT) tile[q] = rec[(size_t)t * T + q];
        __syncthreads();
        if (!act) continue;
        // target relative to the tile origin, shifted image: d = xr_j - ti
        const float t0 = (float)((xi0 - sh.s[m][0]) - origin[t * 3 + 0]);
        const float t1 = (float)((xi1 - sh.s[m][1]) - origin[t * 3 + 1]);
        const float t2 = (float)((xi2 - sh.s[m][2]) - origin[t * 3 + 2]);
#pragma unroll 2
        for (int jj = 0; jj < T; ++jj) {
            const Rec32 q = tile[jj];
            const float d0 = q.x0 - t0, d1 = q.x1 - t1, d2 = q.x2 - t2;
            const float r1sq = fmaf(d0, d0, fmaf(d1, d1, d2 * d2));
            if (r1sq < rn2 || q.j == bi) {
                // near field or excluded pair: reference arithmetic in FP64
                const int j = q.j;
                if (j < 0 || j == bi || (m == 0 && j == i)) continue;
                const double *qv = &rec64[(size_t)j * 8];
                double out[3];
                if (bs_term64((qv[0] + sh.s[m][0]) - xi0, (qv[1] + sh.s[m][1]) - xi1, (qv[2] + sh.s[m][2]) - xi2,
                              qv[3], qv[4], qv[5], qv[6], kappa, out)) {
                    n0 += out[0]; n1 += out[1]; n2 += out[2];
                }
                continue;
            }
            const float de = fmaf(d0, q.e0, fmaf(d1, q.e1, d2 * q.e2));
            const float r2sq = fmaf(2.0f, de, r1sq + q.c);
            const float inv1 = __builtin_amdgcn_rsqf(r1sq);
            const float inv2 = __builtin_amdgcn_rsqf(r2sq);
            const float r1 = r1sq * inv1, r2 = r2sq * inv2;
            const float r1r2 = r1 * r2;
            const float den = r1r2 * (r1r2 + (r1sq + de));
            const float fac = (r1 + r2) * __builtin_amdgcn_rcpf(den);
            const float cx = fmaf(d1, q.e2, -(d2 * q.e1));
            const float cy = fmaf(d2, q.e0, -(d0 * q.e2));
            const float cz = fmaf(d0, q.e1, -(d1 * q.e0));
            kahan_add(f0, k0, cx * fac);
            kahan_add(f1, k1, cy * fac);
            kahan_add(f2, k2, cz * fac);
        }
    }
    if (i < npad) {
        const double k4 = kappa / (4.0 * NV_PI);
        double *pp = part + (size_t)c * 6 * npad;
        pp[0 * npad + i] = ((double)f0 - (double)k0) * k4;
        pp[1 * npad + i] = ((double)f1 - (double)k1) * k4;
        pp[2 * npad + i] = ((double)f2 - (double)k2) * k4;
        pp[3 * npad + i] = n0; pp[4 * npad + i] = n1; pp[5 * npad + i] = n2;
    }
}

// Reduce chunk partials in chunk order and finish every point.
// mode 0: part holds (acc0, accs) of k_bs64; mode 1: (far, near) of k_bs32c;
// mode 2: no partials (LIA).
__global__ void k_finish(DState s, Phys p, int n, int npad, const double *__restrict__ part, int S, int mode,
                         double t, int do_ab, DCtl *ctl) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = r < n && s.infront[r] != 0;
    const unsigned long long m = __ballot(act);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&ctl->nvel, (int)__popcll(m));
    if (r >= n) return;
    if (!act) {
        if (do_ab) {
#pragma unroll
            for (int k = 0; k < 3; ++k) s.xn[k * s.cap + r] = s.x[k * s.cap + r];
        }
        return;
    }
    double ub0[3] = {0.0, 0.0, 0.0}, ubs[3] = {0.0, 0.0, 0.0};
    if (mode == 0) {
        for (int c = 0; c < S; ++c) {
            const double *pp = part + (size_t)c * 6 * npad;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                ub0[k] = ub0[k] + pp[k * npad + r];
                ubs[k] = ubs[k] + pp[(3 + k) * npad + r];
            }
        }
    } else if (mode == 1) {
        double far[3] = {0, 0, 0}, nr[3] = {0, 0, 0};
        for (int c = 0; c < S; ++c) {
            const double *pp = part + (size_t)c * 6 * npad;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                far[k] += pp[k * npad + r];
                nr[k] += pp[(3 + k) * npad + r];
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) ub0[k] = nr[k] + far[k];
    }
    finish_point(s, p, r, ub0, ubs, t, do_ab);
}

}  // namespace

// ------------------------------------------------------------------ launch
size_t nvk_rec32_bytes() { return sizeof(Rec32); }

hipError_t nvk_src64(hipStream_t st, DState s, int n, int npad, double *rec) {
    hipLaunchKernelGGL(k_src64, dim3((npad + 255) / 256), dim3(256), 0, st, s, n, npad, rec);
    return hipGetLastError();
}

hipError_t nvk_bs64(hipStream_t st, DState s, int n, int npad, const double *rec, const Shifts *sh, int ntiles, int S,
                    double kappa, double *part) {
    dim3 grid(npad / 256, S);
    hipLaunchKernelGGL((k_bs64<256, 256>), grid, dim3(256), 0, st, s, n, npad, rec, *sh, ntiles, S, kappa, part);
    return hipGetLastError();
}

hipError_t nvk_src32(hipStream_t st, DState s, int n, int npad, void *rec, double *origin) {
    hipLaunchKernelGGL(k_src32, dim3((npad + 255) / 256), dim3(256), 0, st, s, n, npad, 256, (Rec32 *)rec, origin);
    return hipGetLastError();
}

hipError_t nvk_bs32c(hipStream_t st, DState s, int n, int npad, const void *rec, const double *origin,
                     const double *rec64, const Shifts *sh, int ntiles, int S, double kappa, float rn2, double *part) {
    dim3 grid(npad / 256, S);
    hipLaunchKernelGGL((k_bs32c<256, 256>), grid, dim3(256), 0, st, s, n, npad, (const Rec32 *)rec, origin, rec64,
                       *sh, ntiles, S, kappa, rn2, part);
    return hipGetLastError();
}

hipError_t nvk_finish(hipStream_t st, DState s, Phys p, int n, int npad, const double *part, int S, int mode, double t,
                      int do_ab, DCtl *ctl) {
    if (n > 0)
        hipLaunchKernelGGL(k_finish, dim3((n + 255) / 256), dim3(256), 0, st, s, p, n, npad, part, S, mode, t, do_ab,
                           ctl);
    return hipGetLastError();
}
