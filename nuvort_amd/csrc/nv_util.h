#pragma once
#include <hip/hip_runtime.h>

int nv_scan_tmp_ints(int n);
int nv_compact_tmp_ints(int n);
hipError_t nv_exclusive_scan(hipStream_t st, int *data, int n, int *tmp);
hipError_t nv_compact_flags(hipStream_t st, const int *flags, int n, int *out, int *nout, int *tmp);
hipError_t nv_fill_int(hipStream_t st, int *p, int v, int n);
