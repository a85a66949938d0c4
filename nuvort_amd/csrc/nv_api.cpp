// Host side of the C-ABI declared in include/nuvort.h: the context, device
// buffers, and one function per reference procedure (src/run.f90:229-270
// call order).  All work is launched on the context's stream; every entry
// point synchronises before it returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/nuvort.h"
#include "nv_kernels.h"
#include "nv_tree.h"
#include "nv_util.h"

namespace {

struct Ctx {
    nv_params prm{};
    Phys phys{};
    Shifts sh{};
    hipStream_t st = nullptr;
    int cap = 0, n = 0;
    DState d{};
    DCtl *ctl = nullptr;      // device
    DCtl *ctl_h = nullptr;    // pinned host mirror
    int *cand = nullptr, *empty = nullptr, *lcand = nullptr, *lempty = nullptr, *tmp = nullptr;
    double *rec64 = nullptr, *origin = nullptr, *part = nullptr;
    void *rec32 = nullptr;
    size_t part_elems = 0;
    double *pd = nullptr;
    int *pj = nullptr;
    size_t pcl_elems = 0;
    nv_recon_event *ev = nullptr;
    int ev_cap = 1 << 16;
    double t = 0.0;
    int itime = 0;
    nv_stats stats{};
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    std::string err;
    TreeBuf *tree = nullptr;
    bool tree_valid = false;
    int64_t last_pair_evals = 0;
    int tpt = 1, kb = 1;         // FP32C kernel variant: targets per lane, compensation block
    int n_active = -1;           // active points, refreshed by nv_pinsert / nv_set_state
};

int fail(Ctx *c, int code, const std::string &msg) {
    c->err = msg;
    return code;
}

#define HIPCHK(c, x)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            return fail((c), NV_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));            \
    } while (0)

int npad_of(int n) { return ((n + 255) / 256) * 256; }

int free_state(Ctx *c) {
    double **dp[] = {&c->d.x, &c->d.xn, &c->d.u, &c->d.u1, &c->d.u2, &c->d.usup, &c->d.umf, &c->d.gi, &c->d.gb,
                     &c->d.closestd};
    for (double **p : dp)
        if (*p) { hipFree(*p); *p = nullptr; }
    int **ip[] = {&c->d.infront, &c->d.behind, &c->d.closest, &c->cand, &c->empty, &c->lcand, &c->lempty, &c->tmp};
    for (int **p : ip)
        if (*p) { hipFree(*p); *p = nullptr; }
    c->cap = 0;
    return NV_OK;
}

// (re)allocate device rows for `want` points, keeping the first c->n rows
int ensure_capacity(Ctx *c, int want) {
    if (want <= c->cap) return NV_OK;
    int ncap = std::max(want, c->cap + c->cap / 2);
    ncap = npad_of(ncap);
    DState o = c->d;
    int ocap = c->cap;
    DState d{};
    d.cap = ncap;
    double **dp[] = {&d.x, &d.xn, &d.u, &d.u1, &d.u2, &d.usup, &d.umf, &d.gi, &d.gb};
    double *op[] = {o.x, o.xn, o.u, o.u1, o.u2, o.usup, o.umf, o.gi, o.gb};
    for (int q = 0; q < 9; ++q) {
        HIPCHK(c, hipMalloc(dp[q], sizeof(double) * 3 * (size_t)ncap));
        HIPCHK(c, hipMemsetAsync(*dp[q], 0, sizeof(double) * 3 * (size_t)ncap, c->st));
        if (op[q] && c->n > 0)
            for (int k = 0; k < 3; ++k)
                HIPCHK(c, hipMemcpyAsync(*dp[q] + (size_t)k * ncap, op[q] + (size_t)k * ocap,
                                         sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->st));
    }
    HIPCHK(c, hipMalloc(&d.closestd, sizeof(double) * ncap));
    HIPCHK(c, hipMemsetAsync(d.closestd, 0, sizeof(double) * ncap, c->st));
    int **ip[] = {&d.infront, &d.behind, &d.closest};
    int *oi[] = {o.infront, o.behind, o.closest};
    for (int q = 0; q < 3; ++q) {
        HIPCHK(c, hipMalloc(ip[q], sizeof(int) * (size_t)ncap));
        HIPCHK(c, hipMemsetAsync(*ip[q], 0, sizeof(int) * (size_t)ncap, c->st));
        if (oi[q] && c->n > 0)
            HIPCHK(c, hipMemcpyAsync(*ip[q], oi[q], sizeof(int) * c->n, hipMemcpyDeviceToDevice, c->st));
    }
    if (o.closestd && c->n > 0)
        HIPCHK(c, hipMemcpyAsync(d.closestd, o.closestd, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    int nn = c->n;
    free_state(c);
    c->n = nn;
    c->d = d;
    c->cap = ncap;
    HIPCHK(c, hipMalloc(&c->cand, sizeof(int) * ncap));
    HIPCHK(c, hipMalloc(&c->empty, sizeof(int) * ncap));
    HIPCHK(c, hipMalloc(&c->lcand, sizeof(int) * ncap));
    HIPCHK(c, hipMalloc(&c->lempty, sizeof(int) * ncap));
    HIPCHK(c, hipMalloc(&c->tmp, sizeof(int) * (size_t)nv_compact_tmp_ints(ncap)));
    // velocity source records and chunk partials sized for ncap
    if (c->rec64) hipFree(c->rec64);
    if (c->rec32) hipFree(c->rec32);
    if (c->origin) hipFree(c->origin);
    HIPCHK(c, hipMalloc(&c->rec64, sizeof(double) * 8 * (size_t)ncap));
    HIPCHK(c, hipMalloc(&c->rec32, nvk_rec32_bytes() * (size_t)ncap));
    HIPCHK(c, hipMalloc(&c->origin, sizeof(double) * 3 * (size_t)(ncap / 256 + 1)));
    return NV_OK;
}

int ensure_part(Ctx *c, size_t elems) {
    if (elems <= c->part_elems) return NV_OK;
    if (c->part) hipFree(c->part);
    HIPCHK(c, hipMalloc(&c->part, sizeof(double) * elems));
    c->part_elems = elems;
    return NV_OK;
}

int read_ctl(Ctx *c) {
    HIPCHK(c, hipMemcpyAsync(c->ctl_h, c->ctl, sizeof(DCtl), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->n = c->ctl_h->n;
    int s = c->ctl_h->status;
    if (s & NVK_NAN) return fail(c, NV_ERR_NAN, "run.x: there is a NAN in the (filament) f%x array");
    if (s & NVK_ZERO_LENGTH) return fail(c, NV_ERR_ZERO_LENGTH, "line.mod:pinsert: vortex line length is 0, run over");
    if (s & NVK_SINGULAR) return fail(c, NV_ERR_SINGULAR, "tree.mod:tree_walk: singularity in BS (tree) velocity field");
    if (s & NVK_TREE_DEPTH) return fail(c, NV_ERR_TREE_DEPTH, "octree deeper than NV_TREE_MAX_LEVELS");
    if (s & NVK_CAPACITY) return fail(c, NV_ERR_CAPACITY, "device capacity exceeded");
    return NV_OK;
}

// chunks per target block so that the grid holds ~4 workgroups per CU
int pick_chunks(Ctx *c, int npad, long units) {
    if (c->prm.exact_order) return 1;
    int ntb = npad / 256;
    long want = (4L * 256 + ntb - 1) / ntb;
    if (want > units) want = units;
    if (want < 1) want = 1;
    return (int)want;
}

void swap_x(Ctx *c) { std::swap(c->d.x, c->d.xn); }

int velocity_impl(Ctx *c, int do_ab) {
    const int n = c->n;
    if (n == 0) return NV_OK;
    const int npad = npad_of(n);
    const int ntiles = npad / 256;
    HIPCHK(c, hipEventRecord(c->e0, c->st));
    int mode = 0, S = 1;
    bool timed_pair = false, timed_walk = false;
    if (c->prm.velocity == NV_VEL_LIA) {
        mode = 2;
    } else if (c->prm.velocity == NV_VEL_TREE) {
        if (!c->tree_valid) return fail(c, NV_ERR_STATE, "pmotion: tree velocity without construct_tree");
        if (int rc = ensure_part(c, (size_t)6 * npad)) return rc;
        HIPCHK(c, hipEventRecord(c->e2, c->st));
        int rc = nv_tree_walk_velocity(c->tree, c->st, c->d, c->phys, c->sh, n, npad, c->prm.precision, c->part,
                                       c->ctl, &c->stats);
        if (rc) return fail(c, NV_ERR_HIP, "tree walk launch failed");
        HIPCHK(c, hipEventRecord(c->e3, c->st));
        timed_walk = true;
        mode = 0;
        S = 1;
    } else {
        long units = (long)c->sh.n * ntiles;
        // FP32C: each workgroup covers tpt x 256 targets
        const int npadg = (c->prm.precision == NV_PREC_FP32C) ? ((npad + 256 * c->tpt - 1) / (256 * c->tpt)) * 256
                                                              : npad;
        S = pick_chunks(c, npadg, units);
        if (int rc = ensure_part(c, (size_t)S * 6 * npad)) return rc;
        HIPCHK(c, nvk_src64(c->st, c->d, n, npad, c->rec64));
        if (c->prm.precision == NV_PREC_FP32C) {
            HIPCHK(c, nvk_src32(c->st, c->d, n, npad, c->rec32, c->origin));
            double rn = std::max(c->prm.near_radius, 2.0) * c->prm.delta;
            HIPCHK(c, hipEventRecord(c->e2, c->st));
            HIPCHK(c, nvk_bs32c(c->st, c->d, n, npad, c->rec32, c->origin, c->rec64, &c->sh, ntiles, S,
                                c->phys.kappa, (float)(rn * rn), c->part, c->tpt, c->kb));
            HIPCHK(c, hipEventRecord(c->e3, c->st));
            mode = 1;
        } else {
            HIPCHK(c, hipEventRecord(c->e2, c->st));
            HIPCHK(c, nvk_bs64(c->st, c->d, n, npad, c->rec64, &c->sh, ntiles, S, c->phys.kappa, c->part));
            HIPCHK(c, hipEventRecord(c->e3, c->st));
            mode = 0;
        }
        timed_pair = true;
    }
    c->ctl_h->nvel = 0;
    HIPCHK(c, hipMemcpyAsync(&c->ctl->nvel, &c->ctl_h->nvel, sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, nvk_finish(c->st, c->d, c->phys, n, npad, c->part, S, mode, c->t, do_ab, c->ctl));
    HIPCHK(c, hipMemcpyAsync(&c->ctl_h->nvel, &c->ctl->nvel, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipEventRecord(c->e1, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->e0, c->e1);
    c->stats.ms_velocity = ms;
    const int64_t nv = c->ctl_h->nvel;
    if (do_ab) c->stats.point_updates += nv;
    // evaluated pairs: every active target against every active source in
    // every image, less j == i (unshifted) and j == behind(i) (all images)
    c->last_pair_evals = nv * ((int64_t)c->sh.n * nv - c->sh.n - 1);
    if (timed_pair) {
        float pm = 0.f;
        hipEventElapsedTime(&pm, c->e2, c->e3);
        c->stats.pair_ms_total += pm;
        c->stats.pair_launches += 1;
        c->stats.pair_evals_total += c->last_pair_evals;
    }
    if (timed_walk) {
        if (int rc = nv_tree_walk_counters(c->tree, c->st, &c->stats)) return fail(c, rc, "tree counters");
        float wm = 0.f;
        hipEventElapsedTime(&wm, c->e2, c->e3);
        c->stats.walk_ms_total += wm;
        c->stats.walk_launches += 1;
        c->stats.walk_nodes_total += c->stats.tree_nodes_visited;
    }
    if (do_ab) swap_x(c);
    return NV_OK;
}

int set_phys(Ctx *c) {
    const nv_params &p = c->prm;
    Phys &f = c->phys;
    f.delta = p.delta; f.dt = p.dt; f.kappa = p.quant_circ; f.corea = p.corea;
    for (int k = 0; k < 3; ++k) {
        f.box[k] = p.box_size[k];
        f.periodic[k] = p.boundary[k] == NV_BC_PERIODIC;
        f.sup[k] = p.super_velocity[k];
    }
    f.velocity = p.velocity; f.tree_theta = p.tree_theta; f.extra_corr = p.tree_extra_correction;
    f.nf = p.normal_velocity; f.xflow = p.norm_vel_xflow;
    f.alpha[0] = p.alpha[0]; f.alpha[1] = p.alpha[1];
    f.nf_cutoff = p.normal_fluid_cutoff; f.t_zero = p.t_zero_normal_fluid;
    // src/boundary.f90:143-175 create_periodic_loop_array; image 0 = no shift
    Shifts &s = c->sh;
    s.s[0][0] = s.s[0][1] = s.s[0][2] = 0.0;
    int m = 1;
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k) {
                double ii, jj, kk;
                if (f.periodic[0]) ii = i * f.box[0]; else { if (i != 0) continue; ii = 0.0; }
                if (f.periodic[1]) jj = j * f.box[1]; else { if (j != 0) continue; jj = 0.0; }
                if (f.periodic[2]) kk = k * f.box[2]; else { if (k != 0) continue; kk = 0.0; }
                if (ii == 0.0 && jj == 0.0 && kk == 0.0) continue;
                s.s[m][0] = ii; s.s[m][1] = jj; s.s[m][2] = kk;
                ++m;
            }
    s.n = m;
    return NV_OK;
}

}  // namespace

extern "C" {

int nv_create(const nv_params *params, void **out) {
    if (!params || !out) return NV_ERR_ARG;
    *out = nullptr;
    Ctx *c = new (std::nothrow) Ctx();
    if (!c) return NV_ERR_ARG;
    c->prm = *params;
    const nv_params &p = c->prm;
    int rc = NV_OK;
    if (!(p.delta > NV_EPS)) rc = fail(c, NV_ERR_ARG, "init.mod: delta must be set in run.in");
    else if (!(std::min(p.box_size[0], std::min(p.box_size[1], p.box_size[2])) > 0.0))
        rc = fail(c, NV_ERR_ARG, "init_setup: box size is less than or equal zero");
    else if (p.velocity < 0 || p.velocity > 2) rc = fail(c, NV_ERR_ARG, "velocity must be LIA, BS or Tree");
    else if (p.boundary[0] < 0 || p.boundary[0] > 1 || p.boundary[1] < 0 || p.boundary[1] > 1 || p.boundary[2] < 0 ||
             p.boundary[2] > 1)
        rc = fail(c, NV_ERR_UNSUPPORTED, "only periodic and open boundaries are supported");
    else if (p.tree_theta > 0 && (p.boundary[0] == 0 || p.boundary[1] == 0 || p.boundary[2] == 0))
        rc = fail(c, NV_ERR_ARG, "init.mod:init_setup: tree approximation cannot run with open boundaries");
    else if (p.normal_velocity < 0 || p.normal_velocity > 1)
        rc = fail(c, NV_ERR_UNSUPPORTED, "normal_velocity must be zero or xflow");
    else if (p.precision < 0 || p.precision > 1) rc = fail(c, NV_ERR_ARG, "precision must be FP64 or FP32C");
    else if (p.velocity == NV_VEL_TREE && !(p.tree_theta > 0))
        rc = fail(c, NV_ERR_ARG, "velocity Tree needs tree_theta > 0");
    else if (p.recon_shots < 1) rc = fail(c, NV_ERR_ARG, "recon_shots must be >= 1");
    else if (!p.overide_timestep_check) {
        // src/initial.f90:169-180 timestep_check
        double dmin = p.delta / 2.0;
        double dt_max = std::fabs((dmin * dmin) / (p.quant_circ * std::log(dmin / (NV_PI * p.corea))));
        if (!(p.dt < dt_max)) rc = fail(c, NV_ERR_TIMESTEP, "initial.mod:timestep_check: dt is too large");
    }
    if (rc) {
        // keep the message reachable through a context-less call
        static thread_local std::string last;
        last = c->err;
        delete c;
        return rc;
    }
    set_phys(c);
    c->tpt = p.targets_per_lane == 2 ? 2 : (p.targets_per_lane == 1 ? 1 : 2);
    // default: compensated add every 8 terms (measured: same error as every
    // term on the 16k-point ring tangle, 1.27x faster)
    c->kb = (p.kahan_block == 1 || p.kahan_block == 2 || p.kahan_block == 4 || p.kahan_block == 8) ? p.kahan_block : 8;
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->ctl, sizeof(DCtl)) != hipSuccess ||
        hipHostMalloc((void **)&c->ctl_h, sizeof(DCtl), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&c->ev, sizeof(nv_recon_event) * c->ev_cap) != hipSuccess || hipEventCreate(&c->e0) != hipSuccess ||
        hipEventCreate(&c->e1) != hipSuccess || hipEventCreate(&c->e2) != hipSuccess ||
        hipEventCreate(&c->e3) != hipSuccess) {
        delete c;
        return NV_ERR_HIP;
    }
    std::memset(c->ctl_h, 0, sizeof(DCtl));
    hipMemcpy(c->ctl, c->ctl_h, sizeof(DCtl), hipMemcpyHostToDevice);
    c->tree = nv_tree_new();
    *out = c;
    return NV_OK;
}

int nv_destroy(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    hipStreamSynchronize(c->st);
    free_state(c);
    if (c->rec64) hipFree(c->rec64);
    if (c->rec32) hipFree(c->rec32);
    if (c->origin) hipFree(c->origin);
    if (c->part) hipFree(c->part);
    if (c->pd) hipFree(c->pd);
    if (c->pj) hipFree(c->pj);
    if (c->ev) hipFree(c->ev);
    if (c->ctl) hipFree(c->ctl);
    if (c->ctl_h) hipHostFree(c->ctl_h);
    if (c->tree) nv_tree_free(c->tree);
    hipEventDestroy(c->e0); hipEventDestroy(c->e1); hipEventDestroy(c->e2); hipEventDestroy(c->e3);
    hipStreamDestroy(c->st);
    delete c;
    return NV_OK;
}

const char *nv_last_error(void *h) {
    Ctx *c = (Ctx *)h;
    return c ? c->err.c_str() : "no context";
}

int nv_set_state(void *h, const nv_state *hs) {
    Ctx *c = (Ctx *)h;
    if (!c || !hs || hs->pcount < 0 || hs->capacity < hs->pcount) return NV_ERR_ARG;
    const int n = hs->pcount;
    for (int r = 0; r < n; ++r) {
        if ((hs->pinnedi && hs->pinnedi[r]) || (hs->pinnedb && hs->pinnedb[r]) ||
            (hs->wpinned && (hs->wpinned[3 * r] | hs->wpinned[3 * r + 1] | hs->wpinned[3 * r + 2])))
            return fail(c, NV_ERR_UNSUPPORTED, "pinned points (solid walls) are not supported");
        int f = hs->infront[r], b = hs->behind[r];
        if (f < 0 || f > n || (f != 0 && (b < 1 || b > n))) return fail(c, NV_ERR_ARG, "link out of range");
    }
    c->n = 0;
    int want = std::max(c->prm.capacity, 2 * n + 1024);
    if (int rc = ensure_capacity(c, want)) return rc;
    c->n = n;
    // an existing tree stays valid: like the reference's vtree it holds its
    // own copies of the build-time data, so a close search after a state
    // upload sees the stale tree exactly as pclose_tree does
    std::vector<double> tmp((size_t)3 * c->cap, 0.0);
    const double *src[] = {hs->x, hs->u, hs->u1, hs->u2, hs->u_sup, hs->u_mf, hs->ghosti, hs->ghostb};
    double *dst[] = {c->d.x, c->d.u, c->d.u1, c->d.u2, c->d.usup, c->d.umf, c->d.gi, c->d.gb};
    for (int q = 0; q < 8; ++q) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        if (src[q])
            for (int r = 0; r < n; ++r)
                for (int k = 0; k < 3; ++k) tmp[(size_t)k * c->cap + r] = src[q][3 * r + k];
        HIPCHK(c, hipMemcpy(dst[q], tmp.data(), sizeof(double) * 3 * c->cap, hipMemcpyHostToDevice));
    }
    std::vector<int> zi(c->cap, 0);
    const int32_t *isrc[] = {hs->infront, hs->behind, hs->closest};
    int *idst[] = {c->d.infront, c->d.behind, c->d.closest};
    for (int q = 0; q < 3; ++q) {
        std::fill(zi.begin(), zi.end(), 0);
        if (isrc[q]) std::copy(isrc[q], isrc[q] + n, zi.begin());
        HIPCHK(c, hipMemcpy(idst[q], zi.data(), sizeof(int) * c->cap, hipMemcpyHostToDevice));
    }
    std::vector<double> cd(c->cap, 0.0);
    if (hs->closestd) std::copy(hs->closestd, hs->closestd + n, cd.begin());
    HIPCHK(c, hipMemcpy(c->d.closestd, cd.data(), sizeof(double) * c->cap, hipMemcpyHostToDevice));
    c->t = hs->t;
    c->itime = hs->itime;
    std::memset(c->ctl_h, 0, sizeof(DCtl));
    c->ctl_h->n = n;
    c->ctl_h->recon_count = hs->recon_count;
    c->ctl_h->remove_count = hs->remove_count;
    HIPCHK(c, hipMemcpy(c->ctl, c->ctl_h, sizeof(DCtl), hipMemcpyHostToDevice));
    return NV_OK;
}

int nv_get_state(void *h, nv_state *hs) {
    Ctx *c = (Ctx *)h;
    if (!c || !hs) return NV_ERR_ARG;
    if (int rc = read_ctl(c)) { (void)rc; }
    const int n = c->n;
    if (hs->capacity < n) {
        hs->pcount = n;
        return fail(c, NV_ERR_CAPACITY, "nv_get_state: caller arrays hold fewer rows than pcount");
    }
    hs->pcount = n;
    hs->t = c->t;
    hs->itime = c->itime;
    hs->recon_count = c->ctl_h->recon_count;
    hs->remove_count = c->ctl_h->remove_count;
    if (n == 0) return NV_OK;
    std::vector<double> tmp((size_t)3 * c->cap);
    double *dst[] = {hs->x, hs->u, hs->u1, hs->u2, hs->u_sup, hs->u_mf, hs->ghosti, hs->ghostb};
    const double *src[] = {c->d.x, c->d.u, c->d.u1, c->d.u2, c->d.usup, c->d.umf, c->d.gi, c->d.gb};
    for (int q = 0; q < 8; ++q) {
        if (!dst[q]) continue;
        HIPCHK(c, hipMemcpy(tmp.data(), src[q], sizeof(double) * 3 * c->cap, hipMemcpyDeviceToHost));
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < 3; ++k) dst[q][3 * r + k] = tmp[(size_t)k * c->cap + r];
    }
    int32_t *idst[] = {hs->infront, hs->behind, hs->closest};
    const int *isrc[] = {c->d.infront, c->d.behind, c->d.closest};
    for (int q = 0; q < 3; ++q)
        if (idst[q]) HIPCHK(c, hipMemcpy(idst[q], isrc[q], sizeof(int) * n, hipMemcpyDeviceToHost));
    if (hs->closestd) HIPCHK(c, hipMemcpy(hs->closestd, c->d.closestd, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (hs->pinnedi) std::memset(hs->pinnedi, 0, sizeof(int32_t) * n);
    if (hs->pinnedb) std::memset(hs->pinnedb, 0, sizeof(int32_t) * n);
    if (hs->wpinned) std::memset(hs->wpinned, 0, sizeof(int32_t) * 3 * n);
    return NV_OK;
}

int nv_set_time(void *h, double t, int32_t itime) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    c->t = t;
    c->itime = itime;
    return NV_OK;
}

int nv_ghostp(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    HIPCHK(c, nvk_ghosts(c->st, c->d, c->phys, c->n));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return NV_OK;
}

int nv_construct_tree(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    HIPCHK(c, hipEventRecord(c->e2, c->st));
    int rc = nv_tree_build(c->tree, c->st, c->d, c->phys, c->n, c->ctl, &c->stats);
    if (rc == NV_ERR_TREE_DEPTH) return fail(c, rc, "octree deeper than NV_TREE_MAX_LEVELS (coincident points?)");
    if (rc) return fail(c, rc, "tree build failed");
    HIPCHK(c, hipEventRecord(c->e3, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->e2, c->e3);
    c->stats.ms_tree_build = ms;
    c->tree_valid = true;
    return NV_OK;
}

int nv_velocity(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    return velocity_impl(c, 0);
}

int nv_pmotion(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    return velocity_impl(c, 1);
}

int nv_pinsert(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    // every visited row inserts at most once: 2 x pcount rows always suffice
    if (int rc = ensure_capacity(c, 2 * c->n + 256)) return rc;
    const int n = c->n;
    c->ctl_h->nactive = 0;
    HIPCHK(c, hipMemcpyAsync(&c->ctl->nactive, &c->ctl_h->nactive, sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, nvk_pinsert_pre(c->st, c->d, c->phys, n, c->cand, c->empty, c->ctl));
    HIPCHK(c, nv_compact_flags(c->st, c->cand, n, c->lcand, &c->ctl->ncand, c->tmp));
    HIPCHK(c, nv_compact_flags(c->st, c->empty, n, c->lempty, &c->ctl->nempty, c->tmp));
    HIPCHK(c, nvk_pinsert_serial(c->st, c->d, c->phys, n, c->lcand, c->lempty, c->ctl));
    return read_ctl(c);
}

int nv_pclose(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    const int n = c->n;
    if (n == 0) return NV_OK;
    const int npad = npad_of(n);
    int ntiles = npad / 256;
    int S = pick_chunks(c, npad, ntiles);
    size_t need = (size_t)S * npad;
    if (need > c->pcl_elems) {
        if (c->pd) hipFree(c->pd);
        if (c->pj) hipFree(c->pj);
        HIPCHK(c, hipMalloc(&c->pd, sizeof(double) * need));
        HIPCHK(c, hipMalloc(&c->pj, sizeof(int) * need));
        c->pcl_elems = need;
    }
    HIPCHK(c, hipEventRecord(c->e2, c->st));
    HIPCHK(c, nvk_pclose(c->st, c->d, n, npad, S, c->pd, c->pj));
    HIPCHK(c, hipEventRecord(c->e3, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->e2, c->e3);
    c->stats.ms_close = ms;
    return NV_OK;
}

int nv_pclose_tree(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (!c->tree_valid) return fail(c, NV_ERR_STATE, "pclose_tree without construct_tree");
    HIPCHK(c, hipEventRecord(c->e2, c->st));
    int rc = nv_tree_walk_close(c->tree, c->st, c->d, c->n, &c->stats);
    if (rc) return fail(c, NV_ERR_HIP, "tree close search launch failed");
    HIPCHK(c, hipEventRecord(c->e3, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->e2, c->e3);
    c->stats.ms_close = ms;
    return NV_OK;
}

int nv_precon(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    const int n = c->n;
    HIPCHK(c, nvk_precon_pre(c->st, c->d, c->phys, n, c->cand));
    HIPCHK(c, nv_compact_flags(c->st, c->cand, n, c->lcand, &c->ctl->ncand, c->tmp));
    HIPCHK(c, nvk_precon_serial(c->st, c->d, c->phys, c->lcand, c->ctl, c->ev, c->ev_cap, c->t, c->itime));
    return read_ctl(c);
}

int nv_premove(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    const int n = c->n;
    HIPCHK(c, nvk_premove_pre(c->st, c->d, c->phys, n, c->cand));
    HIPCHK(c, nv_compact_flags(c->st, c->cand, n, c->lcand, &c->ctl->ncand, c->tmp));
    HIPCHK(c, nvk_premove_serial(c->st, c->d, c->phys, c->lcand, c->ctl));
    return read_ctl(c);
}

int nv_enforce_boundary(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    HIPCHK(c, nvk_wrap(c->st, c->d, c->phys, c->n));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return NV_OK;
}

int nv_empty_tree(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    c->tree_valid = false;
    return NV_OK;
}

int nv_nan_check(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    HIPCHK(c, nvk_nan(c->st, c->d, c->n, c->ctl));
    return read_ctl(c);
}

int nv_step(void *h, int32_t nsteps) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    HIPCHK(c, hipEventRecord(c->e0, c->st));
    hipEvent_t s0, s1;
    hipEventCreate(&s0);
    hipEventCreate(&s1);
    hipEventRecord(s0, c->st);
    const bool use_tree = c->prm.tree_theta > 0;
    int rc = NV_OK;
    for (int k = 0; k < nsteps && rc == NV_OK; ++k) {
        c->itime += 1;
        if ((rc = nv_ghostp(c))) break;
        if (use_tree && (rc = nv_construct_tree(c))) break;
        if ((rc = nv_pmotion(c))) break;
        if ((rc = nv_pinsert(c))) break;
        if (c->itime % c->prm.recon_shots == 0) {
            bool tree_search = c->prm.close_search == NV_CLOSE_TREE ||
                               (c->prm.close_search == NV_CLOSE_AUTO && use_tree);
            if (tree_search) rc = nv_pclose_tree(c);
            else rc = nv_pclose(c);
            if (rc) break;
            if ((rc = nv_precon(c))) break;
            if ((rc = nv_premove(c))) break;
        }
        if ((rc = nv_enforce_boundary(c))) break;
        c->t = c->t + c->prm.dt;
        if (use_tree) nv_empty_tree(c);
        if ((rc = nv_nan_check(c))) break;
    }
    hipEventRecord(s1, c->st);
    hipEventSynchronize(s1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, s0, s1);
    c->stats.ms_step = nsteps > 0 ? ms / nsteps : 0.0;
    hipEventDestroy(s0);
    hipEventDestroy(s1);
    return rc;
}

int nv_get_recon_events(void *h, nv_recon_event *buf, int32_t cap, int32_t *n) {
    Ctx *c = (Ctx *)h;
    if (!c || !n) return NV_ERR_ARG;
    if (int rc = read_ctl(c)) return rc;
    int m = std::min(c->ctl_h->nevents, c->ev_cap);
    int take = std::min(m, (int)cap);
    if (take > 0 && buf) HIPCHK(c, hipMemcpy(buf, c->ev, sizeof(nv_recon_event) * take, hipMemcpyDeviceToHost));
    *n = c->ctl_h->nevents;
    c->ctl_h->nevents = 0;
    HIPCHK(c, hipMemcpy(&c->ctl->nevents, &c->ctl_h->nevents, sizeof(int), hipMemcpyHostToDevice));
    return NV_OK;
}

int nv_get_tree(void *h, nv_tree_node *buf, int32_t cap, int32_t *n) {
    Ctx *c = (Ctx *)h;
    if (!c || !n) return NV_ERR_ARG;
    if (!c->tree_valid) return fail(c, NV_ERR_STATE, "no tree (construct_tree not called)");
    int rc = nv_tree_dump(c->tree, c->st, buf, cap, n);
    if (rc) return fail(c, rc, "tree dump failed");
    return NV_OK;
}

int nv_get_stats(void *h, nv_stats *s) {
    Ctx *c = (Ctx *)h;
    if (!c || !s) return NV_ERR_ARG;
    *s = c->stats;
    return NV_OK;
}

}  // extern "C"
