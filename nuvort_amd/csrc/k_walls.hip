// Solid walls: the image vortices of the direct sum and the brute close
// search with pinned points.
//
// A solid wall at +-box/2 in a direction makes the reference add, for every
// target, the field of the whole tangle reflected in each wall plane and
// each combination of them (src/timestep.f90:94-98, :201-234; the
// reflections in src/boundary.f90:180-214 order).  A point pinned to a wall
// ends a filament there: its ghost is the mirror image of its neighbour
// (src/boundary.f90:31-38), the segment from a point pinned in front is not a
// source (:161, :187), and pinned points take no part in the close search
// (src/line.f90:143-151).
#include "nv_device.h"
#include "nv_kernels.h"

namespace {

constexpr int MB = 64;   // targets per workgroup, sources per LDS tile

// src/timestep.f90:201-234 biot_savart_mirror, summed over the reflections in
// order into one running sum per target, sources in index order: the
// reference's order, term by term.  Workgroups of 64 rows lie inside one
// tangle (tangle ranges are multiples of 256 rows); each source tile is
// reflected once into LDS for the workgroup.
__global__ __launch_bounds__(MB) void k_mirror_bs64(DState s, Phys p, Tangles tg, int rows, double *__restrict__ mir) {
    __shared__ double sgi[MB][3], sim[MB][3];   // reflected x_j (s_gi_img) and ghosti_j (s_img)
    __shared__ int sok[MB];
    const int r0 = blockIdx.x * MB, r = r0 + threadIdx.x;
    const int tk = tg.row2t[r0];
    const int base = tg.base[tk], n = tg.n[tk];
    const bool act = r < rows && r - base < n && s.infront[r] != 0;
    double u[3] = {0.0, 0.0, 0.0};
    if (__syncthreads_count(act) > 0) {
        const int cap = s.cap;
        double x[3] = {0.0, 0.0, 0.0};
        int pin = 0, beh = -1, w[3] = {0, 0, 0};
        if (act) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[k] = s.x[k * cap + r];
                w[k] = s.wp[k * cap + r];
            }
            pin = s.pin[r];
            beh = s.behind[r] - 1;
        }
        for (int m = 0; m < p.nmir; ++m) {
            const int rf[3] = {p.mir[m][0], p.mir[m][1], p.mir[m][2]};
            // the own-image test, sums of components as the reference writes
            // it (src/timestep.f90:215)
            const bool own = ((w[0] - rf[0]) + (w[1] - rf[1]) + (w[2] - rf[2])) == 0 && (w[0] + w[1] + w[2]) < 2;
            for (int t0 = base; t0 < base + n; t0 += MB) {
                __syncthreads();
                const int j = t0 + threadIdx.x;
                int ok = 0;
                if (j < base + n && s.infront[j] != 0 && !(s.pin[j] & NV_PINNED_I)) {
                    ok = 1;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        sgi[threadIdx.x][k] = reflect1(s.x[k * cap + j], rf[k], p.box[k]);
                        sim[threadIdx.x][k] = reflect1(s.gi[k * cap + j], rf[k], p.box[k]);
                    }
                }
                sok[threadIdx.x] = ok;
                __syncthreads();
                if (!act) continue;
                const int m_ = min(MB, base + n - t0);
                for (int q = 0; q < m_; ++q) {
                    if (!sok[q]) continue;
                    const int jj = t0 + q;
                    if (own) {
                        if ((pin & NV_PINNED_B) && jj == r) continue;
                        if ((pin & NV_PINNED_I) && jj == beh) continue;
                    }
                    double d[3], e[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        d[k] = sim[q][k] - x[k];
                        e[k] = sgi[q][k] - sim[q][k];
                    }
                    // c = dist_gen_sq(s_img, s_gi_img)
                    const double c0 = sim[q][0] - sgi[q][0], c1 = sim[q][1] - sgi[q][1], c2 = sim[q][2] - sgi[q][2];
                    const double c = c0 * c0 + c1 * c1 + c2 * c2;
                    double o[3];
                    if (bs_term64(d[0], d[1], d[2], e[0], e[1], e[2], c, p.kappa, o)) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) u[k] = u[k] + o[k];
                    }
                }
            }
        }
    }
    if (r < rows) {
#pragma unroll
        for (int k = 0; k < 3; ++k) mir[(size_t)k * rows + r] = u[k];
    }
}

// src/line.f90:135-160 pclose, pinned points included: one thread per target
// row, the sources in index order (strict <, the first minimum wins); a
// pinned target gets closestd = 100 and keeps its closest.
__global__ void k_pclose_walls(DState s, Tangles tg, int rows) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int tk = tg.row2t[r];
    const int base = tg.base[tk], n = tg.n[tk];
    if (r - base >= n || s.infront[r] == 0) return;
    s.closestd[r] = 100.0;
    if (s.pin[r]) return;
    const int cap = s.cap;
    const double x0 = s.x[r], x1 = s.x[cap + r], x2 = s.x[2 * cap + r];
    const int inf = s.infront[r] - 1, beh = s.behind[r] - 1;
    double best = 100.0;
    int cl = s.closest[r];
    for (int j = base; j < base + n; ++j) {
        if (s.infront[j] == 0) continue;
        if (j == r || j == inf || j == beh || s.pin[j]) continue;
        const double dist = dist_gen(x0, x1, x2, s.x[j], s.x[cap + j], s.x[2 * cap + j]);
        if (dist < best) {
            best = dist;
            cl = j + 1;
        }
    }
    s.closest[r] = cl;
    s.closestd[r] = best;
}

}  // namespace

hipError_t nvk_mirror_bs64(hipStream_t st, DState s, Phys p, Tangles tg, int rows, double *mir) {
    if (rows > 0) hipLaunchKernelGGL(k_mirror_bs64, dim3((rows + MB - 1) / MB), dim3(MB), 0, st, s, p, tg, rows, mir);
    return hipGetLastError();
}

hipError_t nvk_pclose_walls(hipStream_t st, DState s, Tangles tg, int rows) {
    if (rows > 0) hipLaunchKernelGGL(k_pclose_walls, dim3((rows + 255) / 256), dim3(256), 0, st, s, tg, rows);
    return hipGetLastError();
}
