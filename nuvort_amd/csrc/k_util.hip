// Device-wide primitives: exclusive scan of int32 and stable compaction of
// row indices whose flag is set.  Deterministic (no atomics on the order).
#include <hip/hip_runtime.h>
#include "nv_util.h"

#include <utility>

namespace {

constexpr int SCAN_BT = 256;
constexpr int SCAN_IPT = 4;
constexpr int SCAN_TILE = SCAN_BT * SCAN_IPT;   // 1024 items per block

__device__ __forceinline__ int block_exclusive_scan(int v, int *lds, int *total) {
    const int tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BT; off <<= 1) {
        int add = tid >= off ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    int incl = lds[tid];
    *total = lds[SCAN_BT - 1];
    __syncthreads();
    return incl - v;
}

// in-place exclusive scan of each 1024-item tile; tile totals to sums[]
__global__ void k_scan_tiles(int *data, int n, int *sums) {
    __shared__ int lds[SCAN_BT];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IPT;
    int v[SCAN_IPT], loc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        v[k] = (base + k < n) ? data[base + k] : 0;
        loc += v[k];
    }
    int total;
    int off = block_exclusive_scan(loc, lds, &total);
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        if (base + k < n) data[base + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

// Exclusive scan, in place, of the nb tile totals k_scan_tiles wrote (one
// workgroup): afterwards the global exclusive prefix of item i of the tiled
// array is data[i] + sums[i / SCAN_TILE].
__global__ void k_scan_sums(int *sums, int nb) {
    __shared__ int lds[SCAN_BT];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += SCAN_TILE) {
        const int bb = b0 + threadIdx.x * SCAN_IPT;
        int w[SCAN_IPT], loc = 0;
#pragma unroll
        for (int k = 0; k < SCAN_IPT; ++k) {
            w[k] = bb + k < nb ? sums[bb + k] : 0;
            loc += w[k];
        }
        int tot;
        int o = block_exclusive_scan(loc, lds, &tot) + carry;
#pragma unroll
        for (int k = 0; k < SCAN_IPT; ++k) {
            if (bb + k < nb) sums[bb + k] = o;
            o += w[k];
        }
        carry += tot;
    }
}

__global__ void k_scan_add(int *data, int n, const int *sums) {
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IPT;
    const int add = sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k)
        if (base + k < n) data[base + k] += add;
}

__global__ void k_count_flags(const int *flags, int n, int *cnt) {
    __shared__ int lds[SCAN_BT];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IPT;
    int loc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) loc += (base + k < n && flags[base + k] != 0) ? 1 : 0;
    int total;
    block_exclusive_scan(loc, lds, &total);
    if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

__global__ void k_scatter_flags(const int *flags, int n, const int *offs, int nblk, int *out, int *nout) {
    __shared__ int lds[SCAN_BT];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_IPT;
    int f[SCAN_IPT], loc = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        f[k] = (base + k < n && flags[base + k] != 0) ? 1 : 0;
        loc += f[k];
    }
    int total;
    int off = block_exclusive_scan(loc, lds, &total) + offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k)
        if (f[k]) out[off++] = base + k;
    if (blockIdx.x == nblk - 1 && threadIdx.x == 0) *nout = offs[blockIdx.x] + total;
}

__global__ void k_fill_int(int *p, int v, int n) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) p[r] = v;
}

// Stable LSD radix sort, 8 bits per pass: per-block digit histograms, one
// exclusive scan over (digit, block), then a scatter in which each wave
// ranks equal digits with ballots (stable within and across waves).
constexpr int RS_BT = 256;
constexpr int RS_IPT = 16;
constexpr int RS_TILE = RS_BT * RS_IPT;

// dn: the item count on the device (at most n), or null for n
__device__ __forceinline__ int count_of(int n, const int *dn) { return dn ? min(*dn, n) : n; }

__global__ void k_rs_hist(const unsigned long long *keys, int n, const int *dn, int shift, int *hist, int nblk) {
    n = count_of(n, dn);
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_IPT; ++k) {
        int q = base + k * RS_BT + threadIdx.x;
        if (q < n) atomicAdd(&h[(keys[q] >> shift) & 0xFF], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

__global__ void k_rs_scatter(const unsigned long long *keys, const int *vals, int n, const int *dn, int shift,
                             const int *hist, const int *tsum, int nblk, unsigned long long *okeys, int *ovals) {
    n = count_of(n, dn);
    __shared__ int running[256];
    __shared__ int wcnt[RS_BT / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    running[tid] = hist[tid * nblk + blockIdx.x] + tsum[(tid * nblk + blockIdx.x) / SCAN_TILE];
    const int base = blockIdx.x * RS_TILE;
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int k = 0; k < RS_IPT; ++k) {
        for (int w = 0; w < RS_BT / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        const int q = base + k * RS_BT + tid;
        const bool ok = q < n;
        unsigned long long key = ok ? keys[q] : ~0ull;
        const int dg = ok ? (int)((key >> shift) & 0xFF) : 256;
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(ok);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dg >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        const bool last = ok && ((peers >> lane) >> 1) == 0ull;
        if (last) wcnt[wv][dg] = rank + 1;
        __syncthreads();
        if (ok) {
            int off = running[dg];
            for (int w = 0; w < wv; ++w) off += wcnt[w][dg];
            okeys[off + rank] = key;
            ovals[off + rank] = vals[q];
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < RS_BT / 64; ++w) add += wcnt[w][tid];
        running[tid] += add;
        __syncthreads();
    }
}

}  // namespace

// a fill kernel rather than hipMemsetAsync, so that code a caller may
// capture into a HIP graph records kernel nodes only
hipError_t nv_fill_int(hipStream_t st, int *p, int v, int n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

int nv_scan_tmp_ints(int n) {
    int t = 0;
    while (n > 1) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        t += n + 1;
    }
    return t + 2;
}

// exclusive scan of data[0..n) in place; tmp needs nv_scan_tmp_ints(n) ints
hipError_t nv_exclusive_scan(hipStream_t st, int *data, int n, int *tmp) {
    if (n <= 0) return hipSuccess;
    int nblk = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nblk == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_BT), 0, st, data, n, (int *)nullptr);
        return hipGetLastError();
    }
    int *sums = tmp;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nblk), dim3(SCAN_BT), 0, st, data, n, sums);
    hipError_t e = nv_exclusive_scan(st, sums, nblk, tmp + nblk + 1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3(nblk), dim3(SCAN_BT), 0, st, data, n, sums);
    return hipGetLastError();
}

// out[0..*nout) = ascending indices r < n with flags[r] != 0.
// tmp needs nv_scan_tmp_ints(n) + nblk ints.
hipError_t nv_compact_flags(hipStream_t st, const int *flags, int n, int *out, int *nout, int *tmp) {
    if (n <= 0) return nv_fill_int(st, nout, 0, 1);
    int nblk = (n + SCAN_TILE - 1) / SCAN_TILE;
    int *cnt = tmp;
    hipLaunchKernelGGL(k_count_flags, dim3(nblk), dim3(SCAN_BT), 0, st, flags, n, cnt);
    hipError_t e = nv_exclusive_scan(st, cnt, nblk, tmp + nblk);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scatter_flags, dim3(nblk), dim3(SCAN_BT), 0, st, flags, n, cnt, nblk, out, nout);
    return hipGetLastError();
}

int nv_compact_tmp_ints(int n) { return nv_scan_tmp_ints((n + SCAN_TILE - 1) / SCAN_TILE) + (n + SCAN_TILE - 1) / SCAN_TILE + 4; }

int nv_radix_blocks(int n) { return (n + RS_TILE - 1) / RS_TILE; }

// Two-launch scan of data[0, n): each tile scanned in place, then the tile
// totals (ceil(n / NV_SCAN_TILE) ints in sums) scanned by one workgroup;
// afterwards the global exclusive prefix of item i is
// data[i] + sums[i / NV_SCAN_TILE].
hipError_t nv_scan_tiles_sums(hipStream_t st, int *data, int n, int *sums) {
    if (n <= 0) return hipSuccess;
    const int nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(SCAN_BT), 0, st, data, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BT), 0, st, sums, nb);
    return hipGetLastError();
}

int nv_radix_hist_ints(int n) { return 256 * nv_radix_blocks(n); }

hipError_t nv_radix_sort_pairs(hipStream_t st, unsigned long long *k0, int *v0, unsigned long long *k1, int *v1, int n,
                               int bits, int *hist, int *scan_tmp) {
    return nv_radix_sort_pairs_dn(st, k0, v0, k1, v1, n, nullptr, bits, hist, scan_tmp);
}

// the same for the first *dn (<= n) pairs, a count on the device: the
// launches cover n, and blocks past the count take no part
hipError_t nv_radix_sort_pairs_dn(hipStream_t st, unsigned long long *k0, int *v0, unsigned long long *k1, int *v1,
                                  int n, const int *dn, int bits, int *hist, int *scan_tmp) {
    if (n <= 0) return hipSuccess;
    const int nblk = nv_radix_blocks(n);
    unsigned long long *ka = k0, *kb = k1;
    int *va = v0, *vb = v1;
    // histogram, tile scan, scan of the tile totals (in scan_tmp), scatter
    // adding the tile prefix: four launches per 8-bit pass
    const int nh = 256 * nblk;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(256), 0, st, ka, n, dn, shift, hist, nblk);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if ((e = nv_scan_tiles_sums(st, hist, nh, scan_tmp)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(RS_BT), 0, st, ka, va, n, dn, shift, (const int *)hist,
                           (const int *)scan_tmp, nblk, kb, vb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    return hipSuccess;
}
