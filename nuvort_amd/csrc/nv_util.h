#pragma once
#include <hip/hip_runtime.h>

int nv_scan_tmp_ints(int n);
int nv_compact_tmp_ints(int n);
hipError_t nv_exclusive_scan(hipStream_t st, int *data, int n, int *tmp);
hipError_t nv_compact_flags(hipStream_t st, const int *flags, int n, int *out, int *nout, int *tmp);
hipError_t nv_fill_int(hipStream_t st, int *p, int v, int n);
// stable LSD radix sort of (key, row) pairs by the low `bits` bits of the
// keys (a multiple of 16, so that the result ends in k0/v0); hist holds
// nv_radix_hist_ints(n) ints, scan_tmp nv_scan_tmp_ints(nv_radix_hist_ints(n))
int nv_radix_blocks(int n);
constexpr int NV_SCAN_TILE = 1024;
hipError_t nv_scan_tiles_sums(hipStream_t st, int *data, int n, int *sums);
int nv_radix_hist_ints(int n);
hipError_t nv_radix_sort_pairs(hipStream_t st, unsigned long long *k0, int *v0, unsigned long long *k1, int *v1, int n,
                               int bits, int *hist, int *scan_tmp);
// the same for the first *dn <= n pairs, a count held on the device
hipError_t nv_radix_sort_pairs_dn(hipStream_t st, unsigned long long *k0, int *v0, unsigned long long *k1, int *v1,
                                  int n, const int *dn, int bits, int *hist, int *scan_tmp);

// Wave votes on a lane condition. HIP's __ballot and __any take an int, so
// a condition already held as a lane mask is first turned into 0/1 in a
// vector register and compared again (two vector instructions); these read
// the mask directly.
__device__ __forceinline__ unsigned long long wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0ull; }
// this lane's bit of a wave-uniform mask, as a lane condition (no instruction:
// the mask is used as the execution or select mask directly)
__device__ __forceinline__ bool lane_in(unsigned long long m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// 10-bit Morton spreading: bit b of v moves to bit 3b
__device__ __forceinline__ unsigned nv_spread10(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
