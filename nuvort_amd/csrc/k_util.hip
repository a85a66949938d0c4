// Device-wide primitives: exclusive scan of int32 and stable compaction of
// row indices whose flag is set.  Deterministic (no atomics on the order).
#include <hip/hip_runtime.h>
#include "nv_util.h"

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
