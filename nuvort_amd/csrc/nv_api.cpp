// Host side of the C-ABI declared in include/nuvort.h: the context, device
// buffers, and one function per reference procedure (src/run.f90:24-65
// call order).  A context holds K >= 1 independent tangles (K = 1 for the
// reference's single filament system); all work is launched on the
// context's stream and every entry point synchronises before it returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/nuvort.h"
#include "nv_io.h"
#include "nv_kernels.h"
#include "nv_tree.h"
#include "nv_util.h"

namespace {

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t m) {
        if (m <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, sizeof(T) * (m > 0 ? m : 1));
        if (e == hipSuccess) n = m;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct Ctx {
    nv_params prm{};
    Phys phys{};
    Shifts sh{};
    hipStream_t st = nullptr;
    int ncu = 256;
    // tangle layout: host mirrors and device copies
    int K = 0, rows = 0;             // rows = sum of tangle capacities = the state arrays' stride
    std::vector<int> tbase, tcap, tn, trecon, tremove, twall;
    DevBuf<int> d_tbase, d_tcap, d_row2t;
    // per-tangle counters, one device block [pcount | recon | remove | wall | nact]
    // so that a step's host read is one copy into pinned memory
    DevBuf<int> d_tcnt;
    int *p_tn = nullptr, *p_trecon = nullptr, *p_tremove = nullptr, *p_twall = nullptr, *p_tnact = nullptr;
    int *h_tcnt = nullptr;           // pinned mirror of the first four
    int h_tcnt_cap = 0;
    DState d{};
    DCtl *ctl = nullptr;             // device
    DCtl *ctl_h = nullptr;           // pinned host mirror
    DevBuf<int> cand, empty, lcand, lempty, tmp;
    DevBuf<double> rec64, part, pd;
    DevBuf<char> rec32, tinfo, sph;
    DevBuf<char> rstage;             // precon's candidates staged before its pass
    int rs_cap = 0;                  // ... their number
    DevBuf<int> pj;
    DevBuf<double> noise;            // noisey_xflow table (nv_set_noise)
    DevBuf<double> diag;             // nv_diagnostics' partial reductions and result
    DevBuf<double> mir;              // solid walls: the image vortices' velocity, [3][rows]
    // work lists of the last direct-sum and close-search launches
    DevBuf<Item> items, pitems;
    DevBuf<int> d_S, d_pS, queue;
    DevBuf<long long> d_poff, d_ppoff;
    nv_recon_event *ev = nullptr;
    int ev_cap = 1 << 16;
    double t = 0.0;
    int itime = 0;
    nv_stats stats{};
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
    hipEvent_t e4 = nullptr, e5 = nullptr, e6 = nullptr, e7 = nullptr;   // the tree step's build and close search
    std::string err;
    TreeBuf *tree = nullptr;
    bool tree_valid = false;
    int tpt = 2, kb = 8, ts = 256;   // FP32C kernel variant: targets per lane, compensation block,
                                     // sources per LDS tile
    int ipc = 16;                    // direct-sum work items per CU
    int occ64 = 4, occ32 = 4;        // resident workgroups per CU of the persistent kernels
    // recorded steps, [x/xn parity][reconnection phase on], valid while
    // gen == layout_gen (a new layout moves every buffer)
    struct StepGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        long gen = -1;
        int tgen = -1;                           // octree steps: the tree buffers' generation ...
        std::array<const void *, 12> sig{};     // ... and the other buffers they hold
    } sg[2][2];
    long layout_gen = 0;
    double *xa = nullptr;            // the buffer alloc_state gave to x
    // spatial order of the FP32 source records (perm in operm, its inverse
    // in iperm), valid for the layout generation it was built in; rebuilt
    // every few steps as points move
    DevBuf<unsigned long long> okey0, okey1;
    DevBuf<int> operm, oval1, iperm, ohist, oscan;
    long state_gen = 0, order_gen = -1;
    int order_age = 0;
    // A batch of tangles with an octree (tree_theta > 0): one context per
    // tangle, each with its own octree, buffers and stream, to which this
    // context forwards every call (K and tn mirror theirs).
    std::vector<Ctx *> subs;
    std::vector<double> noise_h;     // the noise table, for contexts made after nv_set_noise
    bool replay_pending = false;     // a recorded octree step launched, not yet finished
};

constexpr int ORDER_EVERY = 8;       // steps between two source orderings

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

int round256(long v) { return (int)(((v + 255) / 256) * 256); }

Tangles tangles(Ctx *c, const int *S, const long long *poff) {
    Tangles g;
    g.K = c->K;
    g.base = c->d_tbase.p;
    g.cap = c->d_tcap.p;
    g.n = c->p_tn;
    g.recon = c->p_trecon;
    g.remove = c->p_tremove;
    g.wall = c->p_twall;
    g.nact = c->p_tnact;
    g.row2t = c->d_row2t.p;
    g.S = S;
    g.poff = poff;
    return g;
}

// (re)build the spatial order of the FP32 source records when it belongs to
// another layout or has aged (launches only, no synchronisation)
int order_sources(Ctx *c, bool force) {
    if (!force && c->order_gen == c->state_gen && c->order_age < ORDER_EVERY) return NV_OK;
    Tangles tg = tangles(c, nullptr, nullptr);
    HIPCHK(c, nvk_order(c->st, c->d, tg, c->rows, c->phys.box, c->okey0.p, c->operm.p, c->okey1.p, c->oval1.p,
                        c->iperm.p, c->ohist.p, c->oscan.p));
    c->order_gen = c->state_gen;
    c->order_age = 0;
    return NV_OK;
}

void free_dstate(DState &d) {
    double **dp[] = {&d.x, &d.xn, &d.u, &d.u1, &d.u2, &d.usup, &d.umf, &d.gi, &d.gb, &d.closestd, &d.curv, &d.stretch};
    for (double **p : dp)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    int **ip[] = {&d.infront, &d.behind, &d.closest, &d.pin, &d.wp};
    for (int **p : ip)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
}

void free_state(Ctx *c) { free_dstate(c->d); }

// state arrays for `rows` rows, zeroed
hipError_t alloc_dstate(hipStream_t st, int rows, DState *out) {
    DState d{};
    d.cap = rows;
    hipError_t e = hipSuccess;
    double **dp[] = {&d.x, &d.xn, &d.u, &d.u1, &d.u2, &d.usup, &d.umf, &d.gi, &d.gb};
    for (double **p : dp) {
        if ((e = hipMalloc(p, sizeof(double) * 3 * (size_t)rows)) != hipSuccess) break;
        if ((e = hipMemsetAsync(*p, 0, sizeof(double) * 3 * (size_t)rows, st)) != hipSuccess) break;
    }
    double **sp[] = {&d.closestd, &d.curv, &d.stretch};
    for (double **p : sp) {
        if (e != hipSuccess) break;
        if ((e = hipMalloc(p, sizeof(double) * rows)) == hipSuccess) e = hipMemsetAsync(*p, 0, sizeof(double) * rows, st);
    }
    int **ip[] = {&d.infront, &d.behind, &d.closest, &d.pin, &d.wp};
    for (int **p : ip) {
        if (e != hipSuccess) break;
        const size_t m = (p == &d.wp ? 3 : 1) * (size_t)rows;
        if ((e = hipMalloc(p, sizeof(int) * m)) != hipSuccess) break;
        e = hipMemsetAsync(*p, 0, sizeof(int) * m, st);
    }
    if (e != hipSuccess) {
        free_dstate(d);
        return e;
    }
    *out = d;
    return hipSuccess;
}

// row-sized scratch for c->rows rows
int alloc_scratch(Ctx *c) {
    const int rows = c->rows;
    HIPCHK(c, c->cand.reserve(rows));
    HIPCHK(c, c->empty.reserve(rows));
    HIPCHK(c, c->lcand.reserve(rows));
    HIPCHK(c, c->lempty.reserve(rows));
    HIPCHK(c, c->tmp.reserve((size_t)nv_compact_tmp_ints(rows)));
    HIPCHK(c, c->rec64.reserve((size_t)8 * rows));
    // precon stages at most this many candidates before its pass; the rest
    // are staged chunk by chunk from memory
    c->rs_cap = std::min(rows, 32768);
    HIPCHK(c, c->rstage.reserve(nvk_recstage_bytes() * (size_t)std::max(1, c->rs_cap)));
    HIPCHK(c, c->rec32.reserve(nvk_rec32_bytes() * (size_t)rows));
    HIPCHK(c, c->tinfo.reserve(nvk_tileinfo_bytes(rows)));
    HIPCHK(c, c->sph.reserve(nvk_sphere_bytes() * (size_t)(rows / 64 + 4)));
    HIPCHK(c, c->okey0.reserve(rows));
    HIPCHK(c, c->okey1.reserve(rows));
    HIPCHK(c, c->operm.reserve(rows));
    HIPCHK(c, c->oval1.reserve(rows));
    HIPCHK(c, c->iperm.reserve(rows));
    HIPCHK(c, c->ohist.reserve((size_t)nv_radix_hist_ints(rows)));
    HIPCHK(c, c->oscan.reserve((size_t)nv_scan_tmp_ints(nv_radix_hist_ints(rows))));
    if (c->phys.walls) HIPCHK(c, c->mir.reserve((size_t)3 * rows));
    return NV_OK;
}

// allocate the state arrays for c->rows rows (contents zeroed)
int alloc_state(Ctx *c) {
    free_state(c);
    DState d{};
    HIPCHK(c, alloc_dstate(c->st, c->rows, &d));
    c->d = d;
    c->xa = d.x;
    ++c->layout_gen;
    ++c->state_gen;
    if (int rc = alloc_scratch(c)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return NV_OK;
}

// upload the tangle tables (base, capacity, pcount, counters, row -> tangle)
int upload_tables(Ctx *c) {
    const int K = c->K;
    HIPCHK(c, c->d_tbase.reserve(K));
    HIPCHK(c, c->d_tcap.reserve(K));
    HIPCHK(c, c->d_tcnt.reserve((size_t)5 * K));
    c->p_tn = c->d_tcnt.p;
    c->p_trecon = c->d_tcnt.p + K;
    c->p_tremove = c->d_tcnt.p + 2 * K;
    c->p_twall = c->d_tcnt.p + 3 * K;
    c->p_tnact = c->d_tcnt.p + 4 * K;
    if (c->h_tcnt_cap < 4 * K) {
        if (c->h_tcnt) (void)hipHostFree(c->h_tcnt);
        c->h_tcnt = nullptr;
        c->h_tcnt_cap = 0;
        HIPCHK(c, hipHostMalloc((void **)&c->h_tcnt, sizeof(int) * 4 * (size_t)K, hipHostMallocDefault));
        c->h_tcnt_cap = 4 * K;
    }
    HIPCHK(c, c->d_row2t.reserve(c->rows));
    HIPCHK(c, hipMemcpy(c->d_tbase.p, c->tbase.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_tcap.p, c->tcap.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->p_tn, c->tn.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->p_trecon, c->trecon.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->p_tremove, c->tremove.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->p_twall, c->twall.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    std::vector<int> r2t(c->rows);
    for (int k = 0; k < K; ++k) std::fill(r2t.begin() + c->tbase[k], r2t.begin() + c->tbase[k] + c->tcap[k], k);
    HIPCHK(c, hipMemcpy(c->d_row2t.p, r2t.data(), sizeof(int) * c->rows, hipMemcpyHostToDevice));
    return NV_OK;
}

int read_ctl(Ctx *c) {
    HIPCHK(c, hipMemcpyAsync(c->ctl_h, c->ctl, sizeof(DCtl), hipMemcpyDeviceToHost, c->st));
    const int K = c->K;
    if (K > 0)
        HIPCHK(c, hipMemcpyAsync(c->h_tcnt, c->d_tcnt.p, sizeof(int) * 4 * (size_t)K, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (K > 0) {
        std::copy(c->h_tcnt, c->h_tcnt + K, c->tn.begin());
        std::copy(c->h_tcnt + K, c->h_tcnt + 2 * K, c->trecon.begin());
        std::copy(c->h_tcnt + 2 * K, c->h_tcnt + 3 * K, c->tremove.begin());
        std::copy(c->h_tcnt + 3 * K, c->h_tcnt + 4 * K, c->twall.begin());
    }
    int s = c->ctl_h->status;
    if (s & NVK_NAN) return fail(c, NV_ERR_NAN, "run.x: there is a NAN in the (filament) f%x array");
    if (s & NVK_ZERO_LENGTH) return fail(c, NV_ERR_ZERO_LENGTH, "line.mod:pinsert: vortex line length is 0, run over");
    if (s & NVK_SINGULAR) return fail(c, NV_ERR_SINGULAR, "tree.mod:tree_walk: singularity in BS (tree) velocity field");
    if (s & NVK_TREE_DEPTH) return fail(c, NV_ERR_TREE_DEPTH, "octree deeper than NV_TREE_MAX_LEVELS");
    if (s & NVK_CAPACITY) return fail(c, NV_ERR_CAPACITY, "device capacity exceeded");
    return NV_OK;
}

// --------------------------------------------------------- state transfer
// host arrays: the tangles concatenated, links local to each tangle (1-based)
int set_batch(Ctx *c, int K, const int32_t *pcounts, const nv_state *hs, const std::vector<int> *caps) {
    if (K < 1) return fail(c, NV_ERR_ARG, "at least one tangle");
    long total = 0;
    for (int k = 0; k < K; ++k) {
        if (pcounts[k] < 0) return fail(c, NV_ERR_ARG, "negative pcount");
        total += pcounts[k];
    }
    if (hs->capacity < total) return fail(c, NV_ERR_ARG, "state arrays smaller than the sum of pcounts");
    if (K > 1 && c->prm.tree_theta > 0) return fail(c, NV_ERR_UNSUPPORTED, "the octree path runs one tangle");
    // links, and pinned points only where there are walls, tangle by tangle
    long off = 0;
    for (int k = 0; k < K; ++k) {
        const int n = pcounts[k];
        for (int q = 0; q < n; ++q) {
            const long r = off + q;
            if (!c->phys.walls &&
                ((hs->pinnedi && hs->pinnedi[r]) || (hs->pinnedb && hs->pinnedb[r]) ||
                 (hs->wpinned && (hs->wpinned[3 * r] | hs->wpinned[3 * r + 1] | hs->wpinned[3 * r + 2]))))
                return fail(c, NV_ERR_ARG, "pinned points without a solid boundary");
            const int f = hs->infront[r], b = hs->behind[r];
            if (f < 0 || f > n || (f != 0 && (b < 1 || b > n))) return fail(c, NV_ERR_ARG, "link out of range");
        }
        off += n;
    }
    // layout: 256-row aligned ranges with room for at least one step of insertions
    c->K = K;
    c->tbase.assign(K, 0);
    c->tcap.assign(K, 0);
    c->tn.assign(pcounts, pcounts + K);
    c->trecon.assign(K, 0);
    c->tremove.assign(K, 0);
    c->twall.assign(K, 0);
    long rows = 0;
    for (int k = 0; k < K; ++k) {
        // the rows a move to a wider layout would give (3 x pcount + margin),
        // or the caller's count, never fewer than one step of insertions needs
        const long deflt = 3L * pcounts[k] + (K == 1 ? 1024 : 256);
        const long want = caps ? (long)(*caps)[k]
                               : c->prm.capacity > 0 ? std::max<long>(2L * pcounts[k] + 256, (long)c->prm.capacity)
                                                     : deflt;
        c->tbase[k] = (int)rows;
        c->tcap[k] = round256(std::max<long>(want, 256));
        rows += c->tcap[k];
    }
    if (rows > 0x7fffffffL / 8) return fail(c, NV_ERR_CAPACITY, "too many rows");
    c->trecon[0] = hs->recon_count;
    c->tremove[0] = hs->remove_count;
    c->twall[0] = hs->wall_recon_count;
    c->rows = (int)rows;
    if (int rc = alloc_state(c)) return rc;
    if (int rc = upload_tables(c)) return rc;
    const int R = c->rows;
    std::vector<double> tmp((size_t)3 * R);
    const double *src[] = {hs->x, hs->u, hs->u1, hs->u2, hs->u_sup, hs->u_mf, hs->ghosti, hs->ghostb};
    double *dst[] = {c->d.x, c->d.u, c->d.u1, c->d.u2, c->d.usup, c->d.umf, c->d.gi, c->d.gb};
    for (int q = 0; q < 8; ++q) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        if (src[q]) {
            long o = 0;
            for (int k = 0; k < K; ++k) {
                for (int r = 0; r < pcounts[k]; ++r)
                    for (int j = 0; j < 3; ++j) tmp[(size_t)j * R + c->tbase[k] + r] = src[q][3 * (o + r) + j];
                o += pcounts[k];
            }
        }
        HIPCHK(c, hipMemcpy(dst[q], tmp.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice));
    }
    std::vector<int> zi(R);
    const int32_t *isrc[] = {hs->infront, hs->behind, hs->closest};
    int *idst[] = {c->d.infront, c->d.behind, c->d.closest};
    for (int q = 0; q < 3; ++q) {
        std::fill(zi.begin(), zi.end(), 0);
        if (isrc[q]) {
            long o = 0;
            for (int k = 0; k < K; ++k) {
                for (int r = 0; r < pcounts[k]; ++r) {
                    const int v = isrc[q][o + r];
                    zi[c->tbase[k] + r] = v > 0 ? v + c->tbase[k] : 0;   // global 1-based rows
                }
                o += pcounts[k];
            }
        }
        HIPCHK(c, hipMemcpy(idst[q], zi.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    }
    const double *ssrc[] = {hs->closestd, hs->curv, hs->stretch};
    double *sdst[] = {c->d.closestd, c->d.curv, c->d.stretch};
    std::vector<double> cd(R);
    for (int q = 0; q < 3; ++q) {
        std::fill(cd.begin(), cd.end(), 0.0);
        if (ssrc[q]) {
            long o = 0;
            for (int k = 0; k < K; ++k) {
                for (int r = 0; r < pcounts[k]; ++r) cd[c->tbase[k] + r] = ssrc[q][o + r];
                o += pcounts[k];
            }
        }
        HIPCHK(c, hipMemcpy(sdst[q], cd.data(), sizeof(double) * R, hipMemcpyHostToDevice));
    }
    // pinned flags (bit 0 pinnedi, bit 1 pinnedb) and wpinned [3][R]
    std::vector<int> pin(R, 0), wp((size_t)3 * R, 0);
    {
        long o = 0;
        for (int k = 0; k < K; ++k) {
            for (int r = 0; r < pcounts[k]; ++r) {
                const long h = o + r;
                const int g = c->tbase[k] + r;
                pin[g] = (hs->pinnedi && hs->pinnedi[h] ? NV_PINNED_I : 0) | (hs->pinnedb && hs->pinnedb[h] ? NV_PINNED_B : 0);
                if (hs->wpinned)
                    for (int j = 0; j < 3; ++j) wp[(size_t)j * R + g] = hs->wpinned[3 * h + j];
            }
            o += pcounts[k];
        }
    }
    HIPCHK(c, hipMemcpy(c->d.pin, pin.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d.wp, wp.data(), sizeof(int) * 3 * (size_t)R, hipMemcpyHostToDevice));
    c->t = hs->t;
    c->itime = hs->itime;
    // a new point array; the noise table's counter is not part of it
    // (module state of normal_fluid, src/normal_fluid.f90:31)
    int noise_ctr = 0;
    HIPCHK(c, hipMemcpy(&noise_ctr, &c->ctl->noise_ctr, sizeof(int), hipMemcpyDeviceToHost));
    std::memset(c->ctl_h, 0, sizeof(DCtl));
    c->ctl_h->noise_ctr = noise_ctr;
    HIPCHK(c, hipMemcpy(c->ctl, c->ctl_h, sizeof(DCtl), hipMemcpyHostToDevice));
    return NV_OK;
}

int get_batch(Ctx *c, int32_t *pcounts, nv_state *hs) {
    if (int rc = read_ctl(c)) (void)rc;
    long total = 0;
    for (int k = 0; k < c->K; ++k) total += c->tn[k];
    if (pcounts)
        for (int k = 0; k < c->K; ++k) pcounts[k] = c->tn[k];
    if (hs->capacity < total) {
        hs->pcount = (int32_t)total;
        return fail(c, NV_ERR_CAPACITY, "nv_get_state: caller arrays hold fewer rows than pcount");
    }
    hs->pcount = (int32_t)total;
    hs->t = c->t;
    hs->itime = c->itime;
    hs->recon_count = std::accumulate(c->trecon.begin(), c->trecon.end(), 0);
    hs->remove_count = std::accumulate(c->tremove.begin(), c->tremove.end(), 0);
    hs->wall_recon_count = std::accumulate(c->twall.begin(), c->twall.end(), 0);
    if (total == 0) return NV_OK;
    const int R = c->rows;
    std::vector<double> tmp((size_t)3 * R);
    double *dst[] = {hs->x, hs->u, hs->u1, hs->u2, hs->u_sup, hs->u_mf, hs->ghosti, hs->ghostb};
    const double *src[] = {c->d.x, c->d.u, c->d.u1, c->d.u2, c->d.usup, c->d.umf, c->d.gi, c->d.gb};
    for (int q = 0; q < 8; ++q) {
        if (!dst[q]) continue;
        HIPCHK(c, hipMemcpy(tmp.data(), src[q], sizeof(double) * 3 * R, hipMemcpyDeviceToHost));
        long o = 0;
        for (int k = 0; k < c->K; ++k) {
            for (int r = 0; r < c->tn[k]; ++r)
                for (int j = 0; j < 3; ++j) dst[q][3 * (o + r) + j] = tmp[(size_t)j * R + c->tbase[k] + r];
            o += c->tn[k];
        }
    }
    std::vector<int> iv(R);
    int32_t *idst[] = {hs->infront, hs->behind, hs->closest};
    const int *isrc[] = {c->d.infront, c->d.behind, c->d.closest};
    for (int q = 0; q < 3; ++q) {
        if (!idst[q]) continue;
        HIPCHK(c, hipMemcpy(iv.data(), isrc[q], sizeof(int) * R, hipMemcpyDeviceToHost));
        long o = 0;
        for (int k = 0; k < c->K; ++k) {
            for (int r = 0; r < c->tn[k]; ++r) {
                const int v = iv[c->tbase[k] + r];
                idst[q][o + r] = v > 0 ? v - c->tbase[k] : 0;
            }
            o += c->tn[k];
        }
    }
    double *sdst[] = {hs->closestd, hs->curv, hs->stretch};
    const double *ssrc[] = {c->d.closestd, c->d.curv, c->d.stretch};
    for (int q = 0; q < 3; ++q) {
        if (!sdst[q]) continue;
        std::vector<double> cd(R);
        HIPCHK(c, hipMemcpy(cd.data(), ssrc[q], sizeof(double) * R, hipMemcpyDeviceToHost));
        long o = 0;
        for (int k = 0; k < c->K; ++k) {
            for (int r = 0; r < c->tn[k]; ++r) sdst[q][o + r] = cd[c->tbase[k] + r];
            o += c->tn[k];
        }
    }
    if (hs->pinnedi || hs->pinnedb || hs->wpinned) {
        std::vector<int> pin(R), wp((size_t)3 * R);
        HIPCHK(c, hipMemcpy(pin.data(), c->d.pin, sizeof(int) * R, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(wp.data(), c->d.wp, sizeof(int) * 3 * (size_t)R, hipMemcpyDeviceToHost));
        long o = 0;
        for (int k = 0; k < c->K; ++k) {
            for (int r = 0; r < c->tn[k]; ++r) {
                const int g = c->tbase[k] + r;
                if (hs->pinnedi) hs->pinnedi[o + r] = (pin[g] & NV_PINNED_I) ? 1 : 0;
                if (hs->pinnedb) hs->pinnedb[o + r] = (pin[g] & NV_PINNED_B) ? 1 : 0;
                if (hs->wpinned)
                    for (int j = 0; j < 3; ++j) hs->wpinned[3 * (o + r) + j] = wp[(size_t)j * R + g];
            }
            o += c->tn[k];
        }
    }
    return NV_OK;
}

// Give every tangle room for another step of insertions (pinsert inserts at
// most once per visited point, so 2 x pcount rows always suffice): a tangle
// that is short gets 3 x pcount + 1024 rows, and the state moves to the new
// layout on the device (rows copied tangle by tangle, links remapped).
int ensure_room(Ctx *c) {
    bool need = false;
    for (int k = 0; k < c->K; ++k)
        if (c->tcap[k] < 2 * c->tn[k] + 256) need = true;
    if (!need) return NV_OK;
    const int K = c->K;
    std::vector<int> nbase(K), ncap(K);
    long rows = 0;
    for (int k = 0; k < K; ++k) {
        ncap[k] = c->tcap[k] < 2 * c->tn[k] + 256 ? round256(3L * c->tn[k] + 1024) : c->tcap[k];
        nbase[k] = (int)rows;
        rows += ncap[k];
    }
    if (rows > 0x7fffffffL / 8) return fail(c, NV_ERR_CAPACITY, "too many rows");
    DState nd{};
    HIPCHK(c, alloc_dstate(c->st, (int)rows, &nd));
    DevBuf<int> d_nbase;
    if (d_nbase.reserve(K) != hipSuccess) {
        free_dstate(nd);
        return fail(c, NV_ERR_HIP, "relayout: hipMalloc failed");
    }
    hipError_t e = hipMemcpy(d_nbase.p, nbase.data(), sizeof(int) * K, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = nvk_relayout(c->st, c->d, nd, c->d_row2t.p, c->d_tbase.p, d_nbase.p, c->rows);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    d_nbase.release();
    if (e != hipSuccess) {
        free_dstate(nd);
        return fail(c, NV_ERR_HIP, std::string("relayout: ") + hipGetErrorString(e));
    }
    free_state(c);
    c->d = nd;
    c->xa = nd.x;
    c->rows = (int)rows;
    c->tbase = nbase;
    c->tcap = ncap;
    ++c->layout_gen;
    ++c->state_gen;
    if (int rc = alloc_scratch(c)) return rc;
    return upload_tables(c);
}

// ------------------------------------------------------------ work lists
// Direct-sum items, largest tangle first.  The chunk count of each tangle
// keeps items of similar size and gives the launch >= 16 items per CU.
// With cap_rows the list covers every row of each tangle's range, so a
// recorded launch stays valid while pcount grows; the kernels skip blocks
// past the tangle's current pcount.
int build_items(Ctx *c, int tpt, int ts, bool exact, bool cap_rows, int *nitems_out) {
    const int K = c->K;
    std::vector<int> order(K);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return c->tn[a] > c->tn[b]; });
    long work = 0;
    std::vector<int> ntb(K), U(K);
    for (int k = 0; k < K; ++k) {
        ntb[k] = (c->tn[k] + 256 * tpt - 1) / (256 * tpt);
        U[k] = c->sh.n * ((c->tn[k] + ts - 1) / ts);
        work += (long)ntb[k] * U[k];
    }
    const long want = (long)c->ipc * c->ncu;
    const long utarget = std::max<long>(4, std::min<long>(256, work / std::max<long>(1, want)));
    std::vector<int> S(K);
    std::vector<long long> poff(K);
    long long po = 0;
    for (int k = 0; k < K; ++k) {
        S[k] = exact ? 1 : (int)std::max<long>(1, std::min<long>(U[k], (U[k] + utarget - 1) / utarget));
        poff[k] = po;
        po += (long long)S[k] * 6 * c->tcap[k];
    }
    std::vector<Item> it;
    for (int k : order) {
        const int nb = cap_rows ? (c->tcap[k] + 256 * tpt - 1) / (256 * tpt) : ntb[k];
        for (int tb = 0; tb < nb; ++tb)
            for (int ch = 0; ch < S[k]; ++ch) it.push_back(Item{k, tb, ch, 0});
    }
    if (!cap_rows) ++c->layout_gen;   // recorded steps read these lists
    HIPCHK(c, c->items.reserve(std::max<size_t>(1, it.size())));
    HIPCHK(c, c->d_S.reserve(K));
    HIPCHK(c, c->d_poff.reserve(K));
    HIPCHK(c, c->part.reserve((size_t)std::max<long long>(po, 1)));
    if (!it.empty())
        HIPCHK(c, hipMemcpyAsync(c->items.p, it.data(), sizeof(Item) * it.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_S.p, S.data(), sizeof(int) * K, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_poff.p, poff.data(), sizeof(long long) * K, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    *nitems_out = (int)it.size();
    return NV_OK;
}

void swap_x(Ctx *c) { std::swap(c->d.x, c->d.xn); }

// evaluated pairs of one direct-sum launch: every active target against
// every active source of its tangle in every image, less j == i (unshifted)
// and j == behind(i) (all images); nv = active points of a single tangle,
// batch tangle sizes use pcount (empty slots are rare)
int64_t pair_evals(const Ctx *c, int64_t nv, const std::vector<int> &tn) {
    int64_t pe = 0;
    if (c->K == 1) pe = nv * ((int64_t)c->sh.n * nv - c->sh.n - 1);
    else
        for (int k = 0; k < c->K; ++k) pe += (int64_t)tn[k] * ((int64_t)c->sh.n * tn[k] - c->sh.n - 1);
    return pe;
}

// noisey_xflow reads its table in every pmotion that applies friction
bool noise_missing(const Ctx *c) {
    return c->prm.normal_velocity == NV_NF_NOISEY_XFLOW && !c->noise.p &&
           (std::fabs(c->prm.alpha[0]) > NV_EPS || std::fabs(c->prm.alpha[1]) > NV_EPS);
}

// pinsert's diagnostic sums start at zero (fills by kernel: recorded steps
// hold no memset nodes)
hipError_t clear_pinsert_stats(Ctx *c) {
    hipError_t e = nv_fill_int(c->st, &c->ctl->n_old, 0, 1);
    if (e == hipSuccess) e = nv_fill_int(c->st, (int *)&c->ctl->total_length, 0, 2);
    if (e == hipSuccess) e = nv_fill_int(c->st, &c->ctl->empty_after, 0, 1);
    return e;
}

// Solid walls: the image vortices' velocity into c->mir (launches only),
// the direct sum's (src/timestep.f90:94-98) or the octree's (:116-120; in a
// recorded step on the device's counts); LIA has none.
int enqueue_mirrors(Ctx *c, DState d, bool recorded) {
    if (!c->phys.walls || c->prm.velocity == NV_VEL_LIA) return NV_OK;
    if (c->prm.velocity == NV_VEL_BS) {
        HIPCHK(c, nvk_mirror_bs64(c->st, d, c->phys, tangles(c, nullptr, nullptr), c->rows, c->mir.p));
        return NV_OK;
    }
    const int rc = recorded ? nv_tree_graph_mirror(c->tree, c->st, d, c->phys, c->mir.p, c->ctl, c->p_tn)
                            : nv_tree_walk_mirror(c->tree, c->st, d, c->phys, c->tn[0], c->rows, c->mir.p, c->ctl);
    return rc ? fail(c, NV_ERR_HIP, "mirror tree walk launch failed") : NV_OK;
}

// the image vortices' sum k_finish adds, or none
const double *mirror_sum(const Ctx *c) {
    return c->phys.walls && c->prm.velocity != NV_VEL_LIA ? c->mir.p : nullptr;
}

int velocity_impl(Ctx *c, int do_ab) {
    if (c->K == 0) return NV_OK;
    if (noise_missing(c)) return fail(c, NV_ERR_STATE, "normal_velocity noisey_xflow without a noise table (nv_set_noise)");
    const int R = c->rows;
    HIPCHK(c, hipEventRecord(c->e0, c->st));
    int mode = 0;
    bool timed_pair = false, timed_walk = false;
    if (c->prm.velocity == NV_VEL_LIA) {
        mode = 2;
        ++c->layout_gen;
        HIPCHK(c, c->d_S.reserve(c->K));
        HIPCHK(c, c->d_poff.reserve(c->K));
    } else if (c->prm.velocity == NV_VEL_TREE) {
        if (!c->tree_valid) return fail(c, NV_ERR_STATE, "pmotion: tree velocity without construct_tree");
        // one chunk per image: partial block [images][3][rows] at offset 0
        const int nimg = c->sh.n;
        ++c->layout_gen;   // rewrites d_S / d_poff that recorded steps read
        HIPCHK(c, c->d_S.reserve(1));
        HIPCHK(c, c->d_poff.reserve(1));
        HIPCHK(c, c->part.reserve((size_t)nimg * 3 * R));
        HIPCHK(c, nv_fill_int(c->st, c->d_S.p, nimg, 1));
        HIPCHK(c, nv_fill_int(c->st, (int *)c->d_poff.p, 0, 2));   // one long long
        HIPCHK(c, hipEventRecord(c->e2, c->st));
        const double rn = std::max(c->prm.near_radius, 2.0) * c->prm.delta;
        int rc = nv_tree_walk_velocity(c->tree, c->st, c->d, c->phys, c->sh, c->tn[0], R, c->prm.precision,
                                       rn * rn, c->part.p, c->ctl, &c->stats);
        if (rc == NV_ERR_CAPACITY) return fail(c, rc, "FP32 tree walk: more records than 32-bit offsets reach");
        if (rc) return fail(c, NV_ERR_HIP, "tree walk launch failed");
        HIPCHK(c, hipEventRecord(c->e3, c->st));
        timed_walk = true;
        mode = nv_tree_walk_partial_mode(c->prm.precision, c->phys.tree_theta);
    } else {
        const bool fp32 = c->prm.precision == NV_PREC_FP32C;
        int nitems = 0;
        if (int rc = build_items(c, fp32 ? c->tpt : 1, fp32 ? c->ts : 256, c->prm.exact_order != 0, false, &nitems)) return rc;
        HIPCHK(c, c->queue.reserve(1));
        HIPCHK(c, nv_fill_int(c->st, c->queue.p, 0, 1));
        HIPCHK(c, nvk_src64(c->st, c->d, R, c->rec64.p));
        Tangles tg = tangles(c, c->d_S.p, c->d_poff.p);
        if (fp32) {
            if (int rc = order_sources(c, false)) return rc;
            ++c->order_age;
            HIPCHK(c, nvk_src32(c->st, c->d, R, c->operm.p, c->rec32.p, c->tinfo.p, c->ts));
            const double rn = std::max(c->prm.near_radius, 2.0) * c->prm.delta;
            const int grid = std::max(1, std::min(nitems, c->occ32 * c->ncu));
            HIPCHK(c, hipEventRecord(c->e2, c->st));
            HIPCHK(c, nvk_bs32c(c->st, c->d, tg, c->items.p, nitems, c->queue.p, grid, c->rec32.p, c->tinfo.p, c->iperm.p,
                                c->rec64.p, &c->sh, c->phys.kappa, (float)(rn * rn), c->part.p, c->tpt, c->kb, c->ts));
            HIPCHK(c, hipEventRecord(c->e3, c->st));
            mode = 1;
        } else {
            const int grid = std::max(1, std::min(nitems, c->occ64 * c->ncu));
            HIPCHK(c, hipEventRecord(c->e2, c->st));
            HIPCHK(c, nvk_bs64(c->st, c->d, tg, c->items.p, nitems, c->queue.p, grid, c->rec64.p, &c->sh,
                               c->phys.kappa, c->part.p));
            HIPCHK(c, hipEventRecord(c->e3, c->st));
            mode = 0;
        }
        timed_pair = true;
    }
    if (int rc = enqueue_mirrors(c, c->d, false)) return rc;
    HIPCHK(c, nv_fill_int(c->st, &c->ctl->nvel, 0, 1));
    HIPCHK(c, nvk_clock_set(c->st, c->ctl, c->t, c->itime));
    Tangles tg = tangles(c, c->d_S.p, c->d_poff.p);
    HIPCHK(c, nvk_finish(c->st, c->d, c->phys, tg, R, c->part.p, mode, do_ab, c->ctl, mirror_sum(c)));
    HIPCHK(c, hipMemcpyAsync(&c->ctl_h->nvel, &c->ctl->nvel, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipEventRecord(c->e1, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->e0, c->e1);
    c->stats.ms_velocity = ms;
    const int64_t nv = c->ctl_h->nvel;
    if (do_ab) c->stats.point_updates += nv;
    if (timed_pair) {
        const int64_t pe = pair_evals(c, nv, c->tn);
        float pm = 0.f;
        (void)hipEventElapsedTime(&pm, c->e2, c->e3);
        c->stats.pair_ms_total += pm;
        c->stats.pair_launches += 1;
        c->stats.pair_evals_total += pe;
    }
    if (timed_walk) {
        if (int rc = nv_tree_walk_counters(c->tree, c->st, &c->stats)) return fail(c, rc, "tree counters");
        float wm = 0.f;
        (void)hipEventElapsedTime(&wm, c->e2, c->e3);
        c->stats.walk_ms_total += wm;
        c->stats.walk_launches += 1;
        c->stats.walk_nodes_total += c->stats.tree_nodes_visited;
    }
    if (do_ab) swap_x(c);
    return NV_OK;
}

// Brute close-search items (tangle, target block, chunk of source tiles),
// ~16 per CU; cap_rows as for build_items.
int build_close_items(Ctx *c, bool cap_rows, int *nitems_out) {
    const int K = c->K;
    long work = 0;
    std::vector<int> ntb(K), S(K);
    for (int k = 0; k < K; ++k) {
        ntb[k] = (c->tn[k] + 255) / 256;
        work += (long)ntb[k] * ntb[k];
    }
    const long want = 16L * c->ncu;
    const long per = std::max<long>(4, work / std::max<long>(1, want));
    std::vector<long long> poff(K);
    long long po = 0;
    std::vector<Item> it;
    for (int k = 0; k < K; ++k) {
        S[k] = (int)std::max<long>(1, std::min<long>(ntb[k], (ntb[k] + per - 1) / per));
        poff[k] = po;
        po += (long long)S[k] * c->tcap[k];
        const int nb = cap_rows ? c->tcap[k] / 256 : ntb[k];
        for (int tb = 0; tb < nb; ++tb)
            for (int ch = 0; ch < S[k]; ++ch) it.push_back(Item{k, tb, ch, 0});
    }
    if (!cap_rows) ++c->layout_gen;
    HIPCHK(c, c->pitems.reserve(std::max<size_t>(1, it.size())));
    HIPCHK(c, c->d_pS.reserve(K));
    HIPCHK(c, c->d_ppoff.reserve(K));
    HIPCHK(c, c->pd.reserve((size_t)std::max<long long>(po, 1)));
    HIPCHK(c, c->pj.reserve((size_t)std::max<long long>(po, 1)));
    if (!it.empty())
        HIPCHK(c, hipMemcpyAsync(c->pitems.p, it.data(), sizeof(Item) * it.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_pS.p, S.data(), sizeof(int) * K, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->d_ppoff.p, poff.data(), sizeof(long long) * K, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    *nitems_out = (int)it.size();
    return NV_OK;
}

// ------------------------------------------------------- topology phases
// pinsert (src/line.f90:18-96) as launches only, its counters cleared by the
// caller: the staged kernels, or with solid walls the serial form
hipError_t enqueue_pinsert(Ctx *c, DState d) {
    hipStream_t st = c->st;
    const int R = c->rows;
    Tangles tg = tangles(c, nullptr, nullptr);
    if (c->phys.walls) return nvk_pinsert_walls(st, d, c->phys, tg, c->ctl);
    hipError_t e = nvk_pinsert_pre(st, d, c->phys, tg, R, c->cand.p, c->empty.p, c->ctl);
    if (e == hipSuccess) e = nv_compact_flags(st, c->cand.p, R, c->lcand.p, &c->ctl->ncand, c->tmp.p);
    if (e == hipSuccess) e = nv_compact_flags(st, c->empty.p, R, c->lempty.p, &c->ctl->nempty, c->tmp.p);
    if (e == hipSuccess) e = nvk_pinsert_serial(st, d, c->phys, tg, c->lcand.p, c->lempty.p, c->ctl);
    return e;
}

// precon, wall_recon and premove (src/run.f90:42-44) as launches only; the
// device clock stamps the reconnection log
hipError_t enqueue_recon_remove(Ctx *c, DState d) {
    hipStream_t st = c->st;
    const int R = c->rows;
    Tangles tg = tangles(c, nullptr, nullptr);
    hipError_t e;
    if (c->phys.walls) {
        e = nvk_precon_walls(st, d, c->phys, tg, c->ctl, c->ev, c->ev_cap);
        if (e == hipSuccess) e = nvk_wall_recon(st, d, c->phys, tg, c->ctl);
        if (e == hipSuccess) e = nvk_premove_walls(st, d, c->phys, tg, c->ctl);
        return e;
    }
    e = nvk_precon_pre(st, d, c->phys, R, c->cand.p);
    if (e == hipSuccess) e = nv_compact_flags(st, c->cand.p, R, c->lcand.p, &c->ctl->ncand, c->tmp.p);
    if (e == hipSuccess) e = nvk_precon_serial(st, d, c->phys, tg, c->lcand.p, c->ctl, c->ev, c->ev_cap, c->rstage.p, c->rs_cap);
    if (e == hipSuccess) e = nvk_premove_pre(st, d, c->phys, R, c->cand.p);
    if (e == hipSuccess) e = nv_compact_flags(st, c->cand.p, R, c->lcand.p, &c->ctl->ncand, c->tmp.p);
    if (e == hipSuccess) e = nvk_premove_serial(st, d, c->phys, tg, c->lcand.p, c->ctl);
    return e;
}

// ------------------------------------------------------- recorded steps
// One pass of the src/run.f90 loop body (direct sum or LIA) as kernel
// launches only: no copies, fills by kernel, no host reads, the step's
// itime and t advanced on the device.  Work lists cover every row of the
// current layout; the x / xn exchange of the AB update is resolved at
// record time (one recording per parity).
int enqueue_step(Ctx *c, DState d, bool recon, int nitems, int npitems) {
    hipStream_t st = c->st;
    const int R = c->rows;
    HIPCHK(c, nvk_step_begin(st, c->ctl, c->p_tnact, c->K, c->queue.p));
    HIPCHK(c, nvk_ghosts(st, d, c->phys, R));
    int mode = 2;
    Tangles tv = tangles(c, c->d_S.p, c->d_poff.p);
    if (c->prm.velocity == NV_VEL_BS) {
        const bool fp32 = c->prm.precision == NV_PREC_FP32C;
        HIPCHK(c, nvk_src64(st, d, R, c->rec64.p));
        if (fp32) {
            HIPCHK(c, nvk_src32(st, d, R, c->operm.p, c->rec32.p, c->tinfo.p, c->ts));
            const double rn = std::max(c->prm.near_radius, 2.0) * c->prm.delta;
            const int grid = std::max(1, std::min(nitems, c->occ32 * c->ncu));
            HIPCHK(c, nvk_bs32c(st, d, tv, c->items.p, nitems, c->queue.p, grid, c->rec32.p, c->tinfo.p, c->iperm.p,
                                c->rec64.p, &c->sh, c->phys.kappa, (float)(rn * rn), c->part.p, c->tpt, c->kb, c->ts));
            mode = 1;
        } else {
            const int grid = std::max(1, std::min(nitems, c->occ64 * c->ncu));
            HIPCHK(c, nvk_bs64(st, d, tv, c->items.p, nitems, c->queue.p, grid, c->rec64.p, &c->sh, c->phys.kappa,
                               c->part.p));
            mode = 0;
        }
    }
    if (int rc = enqueue_mirrors(c, d, true)) return rc;
    HIPCHK(c, nvk_finish(st, d, c->phys, tv, R, c->part.p, mode, 1, c->ctl, mirror_sum(c)));
    std::swap(d.x, d.xn);
    HIPCHK(c, enqueue_pinsert(c, d));
    if (recon) {
        if (c->phys.walls) {
            HIPCHK(c, nvk_pclose_walls(st, d, tangles(c, nullptr, nullptr), R));
        } else {
            Tangles tp = tangles(c, c->d_pS.p, c->d_ppoff.p);
            HIPCHK(c, nvk_pclose(st, d, tp, R, c->pitems.p, npitems, c->sph.p, c->pd.p, c->pj.p));
        }
        HIPCHK(c, enqueue_recon_remove(c, d));
    }
    HIPCHK(c, nvk_step_end(st, d, c->phys, R, c->ctl, c->prm.dt));
    return NV_OK;
}

// pmotion on the octree (launches only): the walk's partials
// [images][3][rows], then the AB update
int enqueue_tree_pmotion(Ctx *c) {
    hipStream_t st = c->st;
    const int R = c->rows, nimg = c->sh.n;
    ++c->layout_gen;   // rewrites d_S / d_poff that recorded steps read
    HIPCHK(c, c->d_S.reserve(1));
    HIPCHK(c, c->d_poff.reserve(1));
    HIPCHK(c, c->part.reserve((size_t)nimg * 3 * R));
    HIPCHK(c, hipEventRecord(c->e0, st));
    HIPCHK(c, nv_fill_int(st, c->d_S.p, nimg, 1));
    HIPCHK(c, nv_fill_int(st, (int *)c->d_poff.p, 0, 2));   // one long long
    const double rn = std::max(c->prm.near_radius, 2.0) * c->prm.delta;
    HIPCHK(c, hipEventRecord(c->e2, st));
    if (int rw = nv_tree_walk_velocity(c->tree, st, c->d, c->phys, c->sh, c->tn[0], R, c->prm.precision, rn * rn,
                                       c->part.p, c->ctl, &c->stats))
        return rw == NV_ERR_CAPACITY ? fail(c, rw, "FP32 tree walk: more records than 32-bit offsets reach")
                                     : fail(c, NV_ERR_HIP, "tree walk launch failed");
    HIPCHK(c, hipEventRecord(c->e3, st));
    if (int rc = enqueue_mirrors(c, c->d, false)) return rc;
    if (int rc = nv_tree_walk_counters_async(c->tree, st)) return fail(c, rc, "tree counters");
    HIPCHK(c, nv_fill_int(st, &c->ctl->nvel, 0, 1));
    HIPCHK(c, nvk_clock_set(st, c->ctl, c->t, c->itime));
    HIPCHK(c, nvk_finish(st, c->d, c->phys, tangles(c, c->d_S.p, c->d_poff.p), R, c->part.p,
                         nv_tree_walk_partial_mode(c->prm.precision, c->phys.tree_theta), 1, c->ctl,
                         mirror_sum(c)));
    HIPCHK(c, hipEventRecord(c->e1, st));
    swap_x(c);
    return NV_OK;
}

// the figures of enqueue_tree_pmotion once the stream has passed it and
// read_ctl has brought nvel
void take_tree_pmotion_stats(Ctx *c) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->e0, c->e1);
    c->stats.ms_velocity = ms;
    (void)hipEventElapsedTime(&ms, c->e2, c->e3);
    c->stats.walk_ms_total += ms;
    c->stats.walk_launches += 1;
    nv_tree_walk_counters_take(c->tree, &c->stats);
    c->stats.walk_nodes_total += c->stats.tree_nodes_visited;
    c->stats.point_updates += c->ctl_h->nvel;
}

// One pass of the src/run.f90 loop body on the octree path: the phases of
// the eager calls (ghostp, construct_tree, pmotion, pinsert, the close
// search, precon, premove, enforce_boundary, the NaN test) in their order,
// with three host synchronisations instead of one per phase: the build's
// read of its level sizes, the point count after pinsert (the close
// search's target count) and the step's end.  An error that a phase after
// pinsert raises on the device is reported at the step's end.
int tree_step(Ctx *c) {
    hipStream_t st = c->st;
    if (c->K != 1) return fail(c, NV_ERR_UNSUPPORTED, "the octree path runs one tangle");
    if (int rc = ensure_room(c)) return rc;
    c->itime += 1;
    const int R = c->rows;
    HIPCHK(c, nvk_ghosts(st, c->d, c->phys, R));
    HIPCHK(c, hipEventRecord(c->e4, st));
    int rc = nv_tree_build(c->tree, st, c->d, c->phys, c->tn[0], c->ctl, &c->stats);
    if (rc == NV_ERR_TREE_DEPTH) return fail(c, rc, "octree deeper than NV_TREE_MAX_LEVELS (coincident points?)");
    if (rc) return fail(c, rc, "tree build failed");
    HIPCHK(c, hipEventRecord(c->e5, st));
    c->tree_valid = true;
    // with another velocity the octree serves the close search only
    const bool walk = c->prm.velocity == NV_VEL_TREE;
    if ((rc = walk ? enqueue_tree_pmotion(c) : velocity_impl(c, 1))) return rc;
    // pinsert
    HIPCHK(c, nv_fill_int(st, &c->ctl->nactive, 0, 1));
    HIPCHK(c, nv_fill_int(st, c->p_tnact, 0, c->K));
    HIPCHK(c, clear_pinsert_stats(c));
    HIPCHK(c, enqueue_pinsert(c, c->d));
    const int rc_ins = read_ctl(c);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->e4, c->e5);
    c->stats.ms_tree_build = ms;
    if (walk) take_tree_pmotion_stats(c);
    if (rc_ins) return rc_ins;
    const bool search = c->itime % c->prm.recon_shots == 0;
    const bool brute = c->prm.close_search == NV_CLOSE_BRUTE;
    if (search) {
        if (brute) {
            if ((rc = nv_pclose(c))) return rc;
        } else {
            HIPCHK(c, hipEventRecord(c->e6, st));
            const double bsize = std::max(c->phys.box[0], std::max(c->phys.box[1], c->phys.box[2]));
            const int rw = c->phys.walls ? nv_tree_close_walls(c->tree, st, c->d, c->tn[0])
                                         : nv_tree_walk_close(c->tree, st, c->d, bsize, c->tn[0], &c->stats);
            if (rw) return fail(c, NV_ERR_HIP, "tree close search launch failed");
            HIPCHK(c, hipEventRecord(c->e7, st));
        }
        HIPCHK(c, nvk_clock_set(st, c->ctl, c->t, c->itime));
        HIPCHK(c, enqueue_recon_remove(c, c->d));
    }
    HIPCHK(c, nvk_wrap(st, c->d, c->phys, R));
    c->t = c->t + c->prm.dt;
    c->tree_valid = false;
    HIPCHK(c, nvk_nan(st, c->d, R, c->ctl));
    rc = read_ctl(c);
    if (search && !brute) {
        (void)hipEventElapsedTime(&ms, c->e6, c->e7);
        c->stats.ms_close = ms;
    }
    return rc;
}

// the buffers a recorded octree step holds besides the state and the tree
std::array<const void *, 12> tree_step_sig(const Ctx *c) {
    return {c->xa, c->part.p, c->d_S.p, c->d_poff.p, c->cand.p, c->empty.p, c->lcand.p, c->lempty.p, c->tmp.p,
            c->queue.p, c->ev, c->noise.p};
}

// One pass of tree_step as kernel launches only: the step's counts read on
// the device (pcount from the tangle table, the tree's sizes from its level
// table), itime and t advanced on the device, the x / xn exchange resolved at
// record time.  When the build outgrows the recorded plan it sets ctl->abort
// and every later phase does nothing.
int enqueue_tree_step(Ctx *c, DState d, bool recon) {
    hipStream_t st = c->st;
    const int R = c->rows, nimg = c->sh.n;
    const int *dn = c->p_tn;   // pcount of the one tangle
    HIPCHK(c, nvk_step_begin(st, c->ctl, c->p_tnact, c->K, c->queue.p));
    HIPCHK(c, nvk_ghosts(st, d, c->phys, R));
    if (nv_tree_graph_build(c->tree, st, d, c->phys, dn, c->ctl)) return fail(c, NV_ERR_HIP, "tree build launch failed");
    HIPCHK(c, nv_fill_int(st, c->d_S.p, nimg, 1));
    HIPCHK(c, nv_fill_int(st, (int *)c->d_poff.p, 0, 2));   // one long long
    const double rn = std::max(c->prm.near_radius, 2.0) * c->prm.delta;
    if (nv_tree_graph_walk(c->tree, st, d, c->phys, c->sh, c->prm.precision, rn * rn, c->part.p, c->ctl, dn))
        return fail(c, NV_ERR_HIP, "tree walk launch failed");
    if (int rc = enqueue_mirrors(c, d, true)) return rc;
    HIPCHK(c, nvk_finish(st, d, c->phys, tangles(c, c->d_S.p, c->d_poff.p), R, c->part.p,
                         nv_tree_walk_partial_mode(c->prm.precision, c->phys.tree_theta), 1, c->ctl,
                         mirror_sum(c)));
    std::swap(d.x, d.xn);
    // pinsert (its counters were cleared by nvk_step_begin)
    HIPCHK(c, enqueue_pinsert(c, d));
    if (recon) {
        const double bsize = std::max(c->phys.box[0], std::max(c->phys.box[1], c->phys.box[2]));
        const int rw = c->phys.walls ? nv_tree_graph_close_walls(c->tree, st, d, dn, c->ctl)
                                     : nv_tree_graph_close(c->tree, st, d, bsize, dn, c->ctl);
        if (rw) return fail(c, NV_ERR_HIP, "tree close search launch failed");
        HIPCHK(c, enqueue_recon_remove(c, d));
    }
    HIPCHK(c, nvk_step_end(st, d, c->phys, R, c->ctl, c->prm.dt));
    return NV_OK;
}

void drop_graph(Ctx::StepGraph &g) {
    if (g.e) (void)hipGraphExecDestroy(g.e);
    if (g.g) (void)hipGraphDestroy(g.g);
    g.e = nullptr;
    g.g = nullptr;
    g.gen = -1;
}

// Record every step variant from one set of work lists: both x / xn
// parities and, when recon_shots > 1, steps with and without the
// reconnection phase.
int record_steps(Ctx *c) {
    for (auto &row : c->sg)
        for (auto &g : row) drop_graph(g);
    int nitems = 0, npitems = 0;
    if (c->prm.velocity == NV_VEL_BS) {
        const bool fp32 = c->prm.precision == NV_PREC_FP32C;
        if (int rc = build_items(c, fp32 ? c->tpt : 1, fp32 ? c->ts : 256, c->prm.exact_order != 0, true, &nitems)) return rc;
        HIPCHK(c, c->queue.reserve(1));
    } else {
        HIPCHK(c, c->d_S.reserve(c->K));
        HIPCHK(c, c->d_poff.reserve(c->K));
        HIPCHK(c, c->queue.reserve(1));
    }
    if (int rc = build_close_items(c, true, &npitems)) return rc;
    DState d0 = c->d;
    if (d0.x != c->xa) std::swap(d0.x, d0.xn);
    for (int par = 0; par < 2; ++par) {
        DState d = d0;
        if (par) std::swap(d.x, d.xn);
        for (int recon = 0; recon < 2; ++recon) {
            if (!recon && c->prm.recon_shots == 1) continue;
            Ctx::StepGraph &g = c->sg[par][recon];
            HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_step(c, d, recon != 0, nitems, npitems);
            hipGraph_t graph = nullptr;
            const hipError_t ee = hipStreamEndCapture(c->st, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ee != hipSuccess)
                return fail(c, NV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
            g.g = graph;
            HIPCHK(c, hipGraphInstantiate(&g.e, graph, nullptr, nullptr, 0));
            g.gen = c->layout_gen;
        }
    }
    return NV_OK;
}

// nv_step for the direct sum and LIA: a recorded step replayed per step,
// one synchronisation per step (status, pcounts, then the capacity check)
int step_recorded(Ctx *c, int nsteps) {
    HIPCHK(c, nvk_clock_set(c->st, c->ctl, c->t, c->itime));
    HIPCHK(c, hipMemcpyAsync(&c->ctl_h->nvel_total, &c->ctl->nvel_total, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    unsigned long long pu = c->ctl_h->nvel_total;
    for (int k = 0; k < nsteps; ++k) {
        if (int rc = ensure_room(c)) return rc;
        if (c->d.x != c->xa && c->d.xn != c->xa) return fail(c, NV_ERR_STATE, "position buffers lost");
        const int par = c->d.x == c->xa ? 0 : 1;
        const bool recon = (c->itime + 1) % c->prm.recon_shots == 0;
        Ctx::StepGraph &g = c->sg[par][recon ? 1 : 0];
        if (g.gen != c->layout_gen)
            if (int rc = record_steps(c)) return rc;
        const bool fp32 = c->prm.velocity == NV_VEL_BS && c->prm.precision == NV_PREC_FP32C;
        if (fp32) {
            if (int rc = order_sources(c, false)) return rc;
            ++c->order_age;
        }
        HIPCHK(c, hipGraphLaunch(g.e, c->st));
        c->itime += 1;
        swap_x(c);
        const int rc = read_ctl(c);
        c->t = c->t + c->prm.dt;
        c->stats.point_updates += (int64_t)(c->ctl_h->nvel_total - pu);
        pu = c->ctl_h->nvel_total;
        if (rc) return rc;

    }
    return NV_OK;
}

// Record the octree step variants (both x / xn parities; with and without
// the reconnection phase when recon_shots > 1) on the plan of the last
// build.  NV_ERR_STATE (no error text) when there has been no build yet.
int record_tree_steps(Ctx *c) {
    for (auto &row : c->sg)
        for (auto &g : row) drop_graph(g);
    const int R = c->rows;
    if (int rc = nv_tree_graph_prepare(c->tree, c->st, R)) {
        if (rc == NV_ERR_STATE || rc == NV_ERR_UNSUPPORTED) return rc;
        return fail(c, rc, "recorded tree step: buffers");
    }
    HIPCHK(c, c->d_S.reserve(1));
    HIPCHK(c, c->d_poff.reserve(1));
    HIPCHK(c, c->queue.reserve(1));
    HIPCHK(c, c->part.reserve((size_t)c->sh.n * 3 * R));
    DState d0 = c->d;
    if (d0.x != c->xa) std::swap(d0.x, d0.xn);
    for (int par = 0; par < 2; ++par) {
        DState d = d0;
        if (par) std::swap(d.x, d.xn);
        for (int recon = 0; recon < 2; ++recon) {
            if (!recon && c->prm.recon_shots == 1) continue;
            Ctx::StepGraph &g = c->sg[par][recon];
            HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_tree_step(c, d, recon != 0);
            hipGraph_t graph = nullptr;
            const hipError_t ee = hipStreamEndCapture(c->st, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ee != hipSuccess)
                return fail(c, NV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
            g.g = graph;
            HIPCHK(c, hipGraphInstantiate(&g.e, graph, nullptr, nullptr, 0));
            g.gen = c->layout_gen;
            g.tgen = nv_tree_graph_gen(c->tree);
            g.sig = tree_step_sig(c);
        }
    }
    return NV_OK;
}

// nv_step on the octree: a recorded step replayed per step with one
// synchronisation (status, pcount, the tree's level table and counters).
// The first step, a step whose build outgrew the recorded plan (the replay
// then did nothing past the build) and any step that cannot be recorded run
// tree_step; the plan is recorded again from the sizes of the latest build.
// Launch one recorded octree step, or take it eagerly where it cannot be
// recorded; tree_replay_finish waits for a launched one.  Several contexts
// (the tangles of a batch) launch their steps before any waits, so the
// steps run concurrently.
int tree_replay_launch(Ctx *c) {
    hipStream_t st = c->st;
    c->replay_pending = false;
    const bool recorded = c->prm.step_launch == 0 && c->prm.velocity == NV_VEL_TREE &&
                          c->prm.close_search != NV_CLOSE_BRUTE && c->K == 1;
    if (!recorded) return tree_step(c);
    if (int rc = ensure_room(c)) return rc;
    if (c->d.x != c->xa && c->d.xn != c->xa) return fail(c, NV_ERR_STATE, "position buffers lost");
    const int par = c->d.x == c->xa ? 0 : 1;
    const bool recon = (c->itime + 1) % c->prm.recon_shots == 0;
    Ctx::StepGraph &g = c->sg[par][recon ? 1 : 0];
    if (g.gen != c->layout_gen || g.tgen != nv_tree_graph_gen(c->tree) || g.sig != tree_step_sig(c) || !g.e) {
        const int rc = record_tree_steps(c);
        if (rc == NV_ERR_STATE || rc == NV_ERR_UNSUPPORTED) return tree_step(c);
        if (rc) return rc;
    }
    HIPCHK(c, nvk_clock_set(st, c->ctl, c->t, c->itime));
    HIPCHK(c, hipGraphLaunch(g.e, st));
    if (int rc = nv_tree_graph_take_async(c->tree, st)) return fail(c, rc, "tree counters");
    c->replay_pending = true;
    return NV_OK;
}

int tree_replay_finish(Ctx *c) {
    if (!c->replay_pending) return NV_OK;
    c->replay_pending = false;
    hipStream_t st = c->st;
    const int rc = read_ctl(c);
    if (c->ctl_h->abort) {
        // the build outgrew the plan and the replay stopped after it: the
        // state is the step's start, taken again eagerly
        HIPCHK(c, nv_fill_int(st, &c->ctl->abort, 0, 1));
        ++c->stats.tree_redone;
        for (auto &row : c->sg)
            for (auto &gg : row) drop_graph(gg);
        return tree_step(c);
    }
    c->itime += 1;
    c->t = c->t + c->prm.dt;
    swap_x(c);
    c->tree_valid = false;
    c->stats.point_updates += c->ctl_h->nvel;
    ++c->stats.tree_replays;
    if (nv_tree_graph_take(c->tree, &c->stats))
        for (auto &row : c->sg)
            for (auto &gg : row) drop_graph(gg);
    return rc;
}

int step_tree_recorded(Ctx *c, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
        if (int rc = tree_replay_launch(c)) return rc;
        if (int rc = tree_replay_finish(c)) return rc;
    }
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
    f.shear_omega = p.norm_shear_omega;
    f.noise = c->noise.p;
    f.noise_n = c->noise.p ? (int)c->noise.n : 0;
    // src/boundary.f90:180-214 create_mirror_loop_array: the reflections in
    // the solid walls, i, j, k nested over -1..1 (a direction that is not
    // solid held at 0)
    f.walls = 0;
    f.nmir = 0;
    for (int k = 0; k < 3; ++k) {
        f.solid[k] = p.boundary[k] == NV_BC_SOLID;
        f.walls |= f.solid[k];
    }
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k) {
                if ((!f.solid[0] && i) || (!f.solid[1] && j) || (!f.solid[2] && k)) continue;
                if (!i && !j && !k) continue;
                f.mir[f.nmir][0] = (signed char)i;
                f.mir[f.nmir][1] = (signed char)j;
                f.mir[f.nmir][2] = (signed char)k;
                ++f.nmir;
            }
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

int compact(Ctx *c, DevBuf<int> &flags, DevBuf<int> &list, int *count) {
    HIPCHK(c, nv_compact_flags(c->st, flags.p, c->rows, list.p, count, c->tmp.p));
    return NV_OK;
}

// ------------------------------------------------------- octree batches
// A batch of tangles with an octree (tree_theta > 0) runs one context per
// tangle (Ctx::subs): each has its own octree, its own buffers and stream,
// and the batch context forwards every call to them in tangle order.  A
// step launches every tangle's recorded octree step before waiting for any,
// so the tangles' steps run concurrently on the device.
bool forest(const Ctx *c) { return !c->subs.empty(); }

void drop_subs(Ctx *c) {
    for (Ctx *s : c->subs) nv_destroy(s);
    c->subs.clear();
}

// a call on every tangle's context; an error names the tangle
template <class F>
int each_sub(Ctx *c, F f) {
    for (size_t k = 0; k < c->subs.size(); ++k) {
        const int rc = f(c->subs[k]);
        if (rc) return fail(c, rc, "tangle " + std::to_string(k) + ": " + c->subs[k]->err);
    }
    return NV_OK;
}

// tangle k's part of the host arrays: the tangles are concatenated, n rows
// from row o on
nv_state tangle_view(const nv_state *hs, long o, int n) {
    nv_state v = *hs;
    v.pcount = n;
    v.capacity = n;
    double **d3[] = {&v.x, &v.u, &v.u1, &v.u2, &v.u_sup, &v.u_mf, &v.ghosti, &v.ghostb};
    for (double **p : d3)
        if (*p) *p += 3 * o;
    if (v.closestd) v.closestd += o;
    if (v.curv) v.curv += o;
    if (v.stretch) v.stretch += o;
    int32_t **i1[] = {&v.infront, &v.behind, &v.closest, &v.pinnedi, &v.pinnedb};
    for (int32_t **p : i1)
        if (*p) *p += o;
    if (v.wpinned) v.wpinned += 3 * o;
    return v;
}

int set_forest(Ctx *c, int K, const int32_t *pcounts, const nv_state *hs) {
    long total = 0;
    for (int k = 0; k < K; ++k) {
        if (pcounts[k] < 0) return fail(c, NV_ERR_ARG, "negative pcount");
        total += pcounts[k];
    }
    if (hs->capacity < total) return fail(c, NV_ERR_ARG, "state arrays smaller than the sum of pcounts");
    drop_subs(c);
    for (auto &row : c->sg)
        for (auto &g : row) drop_graph(g);
    free_state(c);
    ++c->layout_gen;
    c->K = 0;
    c->tree_valid = false;
    for (int k = 0; k < K; ++k) {
        void *h = nullptr;
        int rc = nv_create(&c->prm, &h);
        if (rc) {
            drop_subs(c);
            return fail(c, rc, "tangle " + std::to_string(k) + ": context");
        }
        Ctx *s = (Ctx *)h;
        c->subs.push_back(s);
        if (!c->noise_h.empty() && (rc = nv_set_noise(s, (int32_t)c->noise_h.size(), c->noise_h.data()))) {
            const std::string m = s->err;
            drop_subs(c);
            return fail(c, rc, "tangle " + std::to_string(k) + ": " + m);
        }
    }
    long o = 0;
    for (int k = 0; k < K; ++k) {
        nv_state v = tangle_view(hs, o, pcounts[k]);
        // the batch's counters belong to its first tangle, as in a batch
        // without an octree
        if (k > 0) v.recon_count = v.remove_count = v.wall_recon_count = 0;
        if (int rc = nv_set_state(c->subs[k], &v)) {
            const std::string m = c->subs[k]->err;
            drop_subs(c);
            return fail(c, rc, "tangle " + std::to_string(k) + ": " + m);
        }
        o += pcounts[k];
    }
    c->K = K;
    c->tn.assign(pcounts, pcounts + K);
    c->t = hs->t;
    c->itime = hs->itime;
    return NV_OK;
}

int get_forest(Ctx *c, int32_t *pcounts, nv_state *hs) {
    const int K = (int)c->subs.size();
    long total = 0;
    for (int k = 0; k < K; ++k) {
        Ctx *s = c->subs[k];
        if (int rc = read_ctl(s)) (void)rc;
        c->tn[k] = s->tn[0];
        total += s->tn[0];
    }
    if (pcounts)
        for (int k = 0; k < K; ++k) pcounts[k] = c->tn[k];
    if (hs->capacity < total) {
        hs->pcount = (int32_t)total;
        return fail(c, NV_ERR_CAPACITY, "nv_get_state: caller arrays hold fewer rows than pcount");
    }
    long o = 0;
    int32_t rec = 0, rem = 0, wall = 0;
    for (int k = 0; k < K; ++k) {
        nv_state v = tangle_view(hs, o, c->tn[k]);
        if (int rc = nv_get_state(c->subs[k], &v)) return fail(c, rc, "tangle " + std::to_string(k) + ": " + c->subs[k]->err);
        rec += v.recon_count;
        rem += v.remove_count;
        wall += v.wall_recon_count;
        o += c->tn[k];
    }
    hs->pcount = (int32_t)total;
    hs->t = c->subs[0]->t;
    hs->itime = c->subs[0]->itime;
    hs->recon_count = rec;
    hs->remove_count = rem;
    hs->wall_recon_count = wall;
    return NV_OK;
}

// nsteps steps of every tangle: each step launched on every tangle's
// stream, then finished tangle by tangle
int step_forest(Ctx *c, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
        int rc = NV_OK;
        size_t q = 0;
        for (; q < c->subs.size(); ++q)
            if ((rc = tree_replay_launch(c->subs[q]))) break;
        // every launched step is finished, also after an error
        for (size_t u = 0; u < c->subs.size(); ++u) {
            const int rf = tree_replay_finish(c->subs[u]);
            if (!rc && rf) {
                rc = rf;
                q = u;
            }
        }
        if (rc) return fail(c, rc, "tangle " + std::to_string(q) + ": " + c->subs[q]->err);
    }
    for (size_t k = 0; k < c->subs.size(); ++k) c->tn[k] = c->subs[k]->tn[0];
    c->t = c->subs[0]->t;
    c->itime = c->subs[0]->itime;
    return NV_OK;
}

// the batch's statistics: counters summed over its tangles, the kernel
// variant of the first
void forest_stats(Ctx *c, nv_stats *out) {
    nv_stats a{};
    for (size_t k = 0; k < c->subs.size(); ++k) {
        nv_stats s{};
        nv_get_stats(c->subs[k], &s);
        if (k == 0) {
            a = s;
            continue;
        }
        a.pair_evals += s.pair_evals; a.near_evals += s.near_evals; a.tree_evals += s.tree_evals;
        a.tree_nodes_visited += s.tree_nodes_visited; a.tree_nodes += s.tree_nodes;
        a.tree_levels = std::max(a.tree_levels, s.tree_levels);
        a.ms_velocity += s.ms_velocity; a.ms_tree_build += s.ms_tree_build; a.ms_close += s.ms_close;
        a.ms_topology += s.ms_topology; a.point_updates += s.point_updates;
        a.pair_launches += s.pair_launches; a.pair_ms_total += s.pair_ms_total;
        a.pair_evals_total += s.pair_evals_total; a.walk_launches += s.walk_launches;
        a.walk_ms_total += s.walk_ms_total; a.walk_nodes_total += s.walk_nodes_total;
        a.tree_builds += s.tree_builds; a.tree_sync_builds += s.tree_sync_builds;
        a.tree_replays += s.tree_replays; a.tree_redone += s.tree_redone;
    }
    a.ms_step = c->stats.ms_step;
    *out = a;
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
    else if (p.boundary[0] < 0 || p.boundary[0] > 2 || p.boundary[1] < 0 || p.boundary[1] > 2 || p.boundary[2] < 0 ||
             p.boundary[2] > 2)
        rc = fail(c, NV_ERR_ARG, "init_setup: incorrect boundary parameter");
    else if (p.tree_theta > 0 && (p.boundary[0] == 0 || p.boundary[1] == 0 || p.boundary[2] == 0))
        rc = fail(c, NV_ERR_ARG, "init.mod:init_setup: tree approximation cannot run with open boundaries");
    else if (p.normal_velocity < NV_NF_ZERO || p.normal_velocity > NV_NF_CHANNEL)
        rc = fail(c, NV_ERR_UNSUPPORTED,
                  "normal_velocity must be zero, xflow, noisey_xflow, shear_xflow, ychannel or channel");
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
        delete c;
        return rc;
    }
    set_phys(c);
    c->tpt = p.targets_per_lane == 1 ? 1 : 2;
    // default: compensated add every 8 terms (measured: same error as every
    // term on the 16k-point ring tangle, 1.27x faster)
    c->kb = (p.kahan_block == 1 || p.kahan_block == 2 || p.kahan_block == 4 || p.kahan_block == 8 ||
             p.kahan_block == 16 || p.kahan_block == 32) ? p.kahan_block : 8;
    c->ts = p.source_tile == 128 ? 128 : 256;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->ctl, sizeof(DCtl)) != hipSuccess ||
        hipHostMalloc((void **)&c->ctl_h, sizeof(DCtl), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&c->ev, sizeof(nv_recon_event) * c->ev_cap) != hipSuccess || hipEventCreate(&c->e0) != hipSuccess ||
        hipEventCreate(&c->e1) != hipSuccess || hipEventCreate(&c->e2) != hipSuccess ||
        hipEventCreate(&c->e3) != hipSuccess || hipEventCreate(&c->e4) != hipSuccess ||
        hipEventCreate(&c->e5) != hipSuccess || hipEventCreate(&c->e6) != hipSuccess ||
        hipEventCreate(&c->e7) != hipSuccess) {
        delete c;
        return NV_ERR_HIP;
    }
    c->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    c->occ64 = nvk_occupancy_bs64();
    c->occ32 = nvk_occupancy_bs32c(c->tpt, c->kb, c->ts);
    std::memset(c->ctl_h, 0, sizeof(DCtl));
    (void)hipMemcpy(c->ctl, c->ctl_h, sizeof(DCtl), hipMemcpyHostToDevice);
    c->tree = nv_tree_new();
    *out = c;
    return NV_OK;
}

int nv_destroy(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    drop_subs(c);
    (void)hipStreamSynchronize(c->st);
    for (auto &row : c->sg)
        for (auto &g : row) drop_graph(g);
    free_state(c);
    if (c->h_tcnt) (void)hipHostFree(c->h_tcnt);
    DevBuf<int> *ib[] = {&c->d_tbase, &c->d_tcap, &c->d_tcnt, &c->d_row2t,
                         &c->cand, &c->empty, &c->lcand, &c->lempty, &c->tmp, &c->pj, &c->d_S, &c->d_pS, &c->queue};
    for (auto *b : ib) b->release();
    c->rec64.release(); c->tinfo.release(); c->sph.release(); c->part.release(); c->pd.release(); c->rec32.release(); c->rstage.release();
    c->items.release(); c->pitems.release(); c->d_poff.release(); c->d_ppoff.release();
    if (c->ev) (void)hipFree(c->ev);
    if (c->ctl) (void)hipFree(c->ctl);
    if (c->ctl_h) (void)hipHostFree(c->ctl_h);
    if (c->tree) nv_tree_free(c->tree);
    (void)hipEventDestroy(c->e0); (void)hipEventDestroy(c->e1);
    (void)hipEventDestroy(c->e2); (void)hipEventDestroy(c->e3);
    (void)hipEventDestroy(c->e4); (void)hipEventDestroy(c->e5);
    (void)hipEventDestroy(c->e6); (void)hipEventDestroy(c->e7);
    (void)hipStreamDestroy(c->st);
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
    // an existing tree stays valid: like the reference's vtree it holds its
    // own copies of the build-time data, so a close search after a state
    // upload sees the stale tree exactly as pclose_tree does
    const int32_t n = hs->pcount;
    drop_subs(c);
    return set_batch(c, 1, &n, hs, nullptr);
}

int nv_get_state(void *h, nv_state *hs) {
    Ctx *c = (Ctx *)h;
    if (!c || !hs) return NV_ERR_ARG;
    if (c->K > 1) return fail(c, NV_ERR_STATE, "a batch of tangles: use nv_get_state_batch");
    if (c->K == 0) {
        hs->pcount = 0;
        return NV_OK;
    }
    return get_batch(c, nullptr, hs);
}

int nv_set_state_batch(void *h, int32_t ntangles, const int32_t *pcounts, const nv_state *hs) {
    Ctx *c = (Ctx *)h;
    if (!c || !hs || !pcounts || ntangles < 1) return NV_ERR_ARG;
    if (ntangles > 1 && c->prm.tree_theta > 0) return set_forest(c, ntangles, pcounts, hs);
    drop_subs(c);
    return set_batch(c, ntangles, pcounts, hs, nullptr);
}

int nv_get_state_batch(void *h, int32_t ntangles, int32_t *pcounts, nv_state *hs) {
    Ctx *c = (Ctx *)h;
    if (!c || !hs || !pcounts) return NV_ERR_ARG;
    if (ntangles != c->K) return fail(c, NV_ERR_ARG, "ntangles differs from the uploaded batch");
    if (forest(c)) return get_forest(c, pcounts, hs);
    return get_batch(c, pcounts, hs);
}

int nv_set_time(void *h, double t, int32_t itime) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    c->t = t;
    c->itime = itime;
    if (forest(c)) return each_sub(c, [&](Ctx *s) { return nv_set_time(s, t, itime); });
    return NV_OK;
}

int nv_set_noise(void *h, int32_t n, const double *noise) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (n < 1 || !noise) return fail(c, NV_ERR_ARG, "nv_set_noise: the table needs at least one value");
    for (int32_t q = 0; q < n; ++q)
        if (std::isnan(noise[q])) return fail(c, NV_ERR_ARG, "nv_set_noise: NaN in the noise table");
    c->noise_h.assign(noise, noise + n);
    if (forest(c))
        if (int rc = each_sub(c, [&](Ctx *s) { return nv_set_noise(s, n, noise); })) return rc;
    c->noise.release();
    HIPCHK(c, c->noise.reserve((size_t)n));
    HIPCHK(c, hipMemcpy(c->noise.p, noise, sizeof(double) * n, hipMemcpyHostToDevice));
    const int zero = 0;
    HIPCHK(c, hipMemcpy(&c->ctl->noise_ctr, &zero, sizeof(int), hipMemcpyHostToDevice));
    set_phys(c);
    ++c->layout_gen;   // recorded steps hold the old table's address
    return NV_OK;
}

int nv_ghostp(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_ghostp(s); });
    HIPCHK(c, nvk_ghosts(c->st, c->d, c->phys, c->rows));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return NV_OK;
}

int nv_construct_tree(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_construct_tree(s); });
    if (c->K != 1) return fail(c, NV_ERR_UNSUPPORTED, "the octree path runs one tangle");
    HIPCHK(c, hipEventRecord(c->e2, c->st));
    int rc = nv_tree_build(c->tree, c->st, c->d, c->phys, c->tn[0], c->ctl, &c->stats);
    if (rc == NV_ERR_TREE_DEPTH) return fail(c, rc, "octree deeper than NV_TREE_MAX_LEVELS (coincident points?)");
    if (rc) return fail(c, rc, "tree build failed");
    HIPCHK(c, hipEventRecord(c->e3, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->e2, c->e3);
    c->stats.ms_tree_build = ms;
    c->tree_valid = true;
    return NV_OK;
}

int nv_velocity(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_velocity(s); });
    return velocity_impl(c, 0);
}

int nv_pmotion(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_pmotion(s); });
    return velocity_impl(c, 1);
}

int nv_pinsert(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_pinsert(s); });
    if (c->K == 0) return NV_OK;
    if (int rc = ensure_room(c)) return rc;
    Tangles tg = tangles(c, nullptr, nullptr);
    c->ctl_h->nactive = 0;
    HIPCHK(c, hipMemcpyAsync(&c->ctl->nactive, &c->ctl_h->nactive, sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, clear_pinsert_stats(c));
    HIPCHK(c, nv_fill_int(c->st, c->p_tnact, 0, c->K));
    if (c->phys.walls) {
        HIPCHK(c, nvk_pinsert_walls(c->st, c->d, c->phys, tg, c->ctl));
        return read_ctl(c);
    }
    HIPCHK(c, nvk_pinsert_pre(c->st, c->d, c->phys, tg, c->rows, c->cand.p, c->empty.p, c->ctl));
    if (int rc = compact(c, c->cand, c->lcand, &c->ctl->ncand)) return rc;
    if (int rc = compact(c, c->empty, c->lempty, &c->ctl->nempty)) return rc;
    HIPCHK(c, nvk_pinsert_serial(c->st, c->d, c->phys, tg, c->lcand.p, c->lempty.p, c->ctl));
    return read_ctl(c);
}

int nv_pclose(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_pclose(s); });
    if (c->K == 0) return NV_OK;
    if (c->phys.walls) {
        HIPCHK(c, hipEventRecord(c->e2, c->st));
        HIPCHK(c, nvk_pclose_walls(c->st, c->d, tangles(c, nullptr, nullptr), c->rows));
        HIPCHK(c, hipEventRecord(c->e3, c->st));
    } else {
        int nitems = 0;
        if (int rc = build_close_items(c, false, &nitems)) return rc;
        Tangles tg = tangles(c, c->d_pS.p, c->d_ppoff.p);
        HIPCHK(c, hipEventRecord(c->e2, c->st));
        HIPCHK(c, nvk_pclose(c->st, c->d, tg, c->rows, c->pitems.p, nitems, c->sph.p, c->pd.p, c->pj.p));
        HIPCHK(c, hipEventRecord(c->e3, c->st));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->e2, c->e3);
    c->stats.ms_close = ms;
    return NV_OK;
}

int nv_pclose_tree(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_pclose_tree(s); });
    if (!c->tree_valid) return fail(c, NV_ERR_STATE, "pclose_tree without construct_tree");
    if (c->K != 1) return fail(c, NV_ERR_UNSUPPORTED, "the octree path runs one tangle");
    HIPCHK(c, hipEventRecord(c->e2, c->st));
    const double bsize = std::max(c->phys.box[0], std::max(c->phys.box[1], c->phys.box[2]));
    int rc = c->phys.walls ? nv_tree_close_walls(c->tree, c->st, c->d, c->tn[0])
                           : nv_tree_walk_close(c->tree, c->st, c->d, bsize, c->tn[0], &c->stats);
    if (rc) return fail(c, NV_ERR_HIP, "tree close search launch failed");
    HIPCHK(c, hipEventRecord(c->e3, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->e2, c->e3);
    c->stats.ms_close = ms;
    return NV_OK;
}

int nv_precon(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_precon(s); });
    if (c->K == 0) return NV_OK;
    Tangles tg = tangles(c, nullptr, nullptr);
    if (c->phys.walls) {
        HIPCHK(c, nvk_clock_set(c->st, c->ctl, c->t, c->itime));
        HIPCHK(c, nvk_precon_walls(c->st, c->d, c->phys, tg, c->ctl, c->ev, c->ev_cap));
        return read_ctl(c);
    }
    HIPCHK(c, nvk_precon_pre(c->st, c->d, c->phys, c->rows, c->cand.p));
    if (int rc = compact(c, c->cand, c->lcand, &c->ctl->ncand)) return rc;
    HIPCHK(c, nvk_clock_set(c->st, c->ctl, c->t, c->itime));
    HIPCHK(c, nvk_precon_serial(c->st, c->d, c->phys, tg, c->lcand.p, c->ctl, c->ev, c->ev_cap, c->rstage.p,
                                 c->rs_cap));
    return read_ctl(c);
}

int nv_wall_recon(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_wall_recon(s); });
    // without solid walls there are no pinned points and no wall to reach:
    // the reference's loop changes nothing
    if (c->K == 0 || !c->phys.walls) return NV_OK;
    HIPCHK(c, nvk_wall_recon(c->st, c->d, c->phys, tangles(c, nullptr, nullptr), c->ctl));
    return read_ctl(c);
}

int nv_premove(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_premove(s); });
    if (c->K == 0) return NV_OK;
    Tangles tg = tangles(c, nullptr, nullptr);
    if (c->phys.walls) {
        HIPCHK(c, nvk_premove_walls(c->st, c->d, c->phys, tg, c->ctl));
        return read_ctl(c);
    }
    HIPCHK(c, nvk_premove_pre(c->st, c->d, c->phys, c->rows, c->cand.p));
    if (int rc = compact(c, c->cand, c->lcand, &c->ctl->ncand)) return rc;
    HIPCHK(c, nvk_premove_serial(c->st, c->d, c->phys, tg, c->lcand.p, c->ctl));
    return read_ctl(c);
}

int nv_enforce_boundary(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_enforce_boundary(s); });
    HIPCHK(c, nvk_wrap(c->st, c->d, c->phys, c->rows));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return NV_OK;
}

int nv_empty_tree(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_empty_tree(s); });
    c->tree_valid = false;
    return NV_OK;
}

int nv_nan_check(void *h) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (forest(c)) return each_sub(c, [](Ctx *s) { return nv_nan_check(s); });
    HIPCHK(c, nvk_nan(c->st, c->d, c->rows, c->ctl));
    return read_ctl(c);
}

int nv_step(void *h, int32_t nsteps) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    const int total = nsteps;
    if (forest(c)) {
        if (noise_missing(c)) return fail(c, NV_ERR_STATE, "normal_velocity noisey_xflow without a noise table (nv_set_noise)");
        const auto w0 = std::chrono::steady_clock::now();
        const int rc = step_forest(c, nsteps);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        c->stats.ms_step = total > 0 ? ms / total : 0.0;
        return rc;
    }
    if (noise_missing(c)) return fail(c, NV_ERR_STATE, "normal_velocity noisey_xflow without a noise table (nv_set_noise)");
    hipEvent_t s0, s1;
    HIPCHK(c, hipEventCreate(&s0));
    HIPCHK(c, hipEventCreate(&s1));
    HIPCHK(c, hipEventRecord(s0, c->st));
    const bool use_tree = c->prm.tree_theta > 0;
    int rc = NV_OK;
    if (!use_tree && c->prm.step_launch == 0 && c->K > 0 && nsteps > 0) {
        rc = step_recorded(c, nsteps);
        nsteps = 0;
    }
    if (use_tree) {
        const bool recorded = c->prm.step_launch == 0 && c->prm.velocity == NV_VEL_TREE &&
                              c->prm.close_search != NV_CLOSE_BRUTE && c->K == 1;
        if (recorded) rc = step_tree_recorded(c, nsteps);
        else
            for (int k = 0; k < nsteps && rc == NV_OK; ++k) rc = tree_step(c);
        nsteps = 0;
    }
    for (int k = 0; k < nsteps && rc == NV_OK; ++k) {
        // room for this step's insertions first, as the recorded step does,
        // so both make the same layout moves at the same points
        if ((rc = ensure_room(c))) break;
        c->itime += 1;
        if ((rc = nv_ghostp(c))) break;
        if (use_tree && (rc = nv_construct_tree(c))) break;
        if ((rc = nv_pmotion(c))) break;
        if ((rc = nv_pinsert(c))) break;
        if (c->itime % c->prm.recon_shots == 0) {
            const bool tree_search = c->prm.close_search == NV_CLOSE_TREE ||
                                     (c->prm.close_search == NV_CLOSE_AUTO && use_tree);
            if (tree_search) rc = nv_pclose_tree(c);
            else rc = nv_pclose(c);
            if (rc) break;
            if ((rc = nv_precon(c))) break;
            if ((rc = nv_wall_recon(c))) break;
            if ((rc = nv_premove(c))) break;
        }
        if ((rc = nv_enforce_boundary(c))) break;
        c->t = c->t + c->prm.dt;
        if (use_tree) nv_empty_tree(c);
        if ((rc = nv_nan_check(c))) break;
    }
    (void)hipEventRecord(s1, c->st);
    (void)hipEventSynchronize(s1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s0, s1);
    c->stats.ms_step = total > 0 ? ms / total : 0.0;
    (void)hipEventDestroy(s0);
    (void)hipEventDestroy(s1);
    return rc;
}

int nv_diagnostics(void *h, nv_diag *out) {
    Ctx *c = (Ctx *)h;
    if (!c || !out) return NV_ERR_ARG;
    if (c->K != 1) return fail(c, NV_ERR_UNSUPPORTED, "nv_diagnostics: one tangle (the reference's single f)");
    const int n = c->tn[0];
    HIPCHK(c, c->diag.reserve((size_t)nvk_diag_parts(n) + 16));
    DDiag *dd = (DDiag *)(c->diag.p + nvk_diag_parts(n));
    // src/diagnostic.f90:13: the ghosts of the current positions first
    HIPCHK(c, nvk_ghosts(c->st, c->d, c->phys, c->rows));
    HIPCHK(c, nvk_diagnostics(c->st, c->d, n, c->diag.p, dd));
    DDiag hd{};
    HIPCHK(c, hipMemcpyAsync(&hd, dd, sizeof(DDiag), hipMemcpyDeviceToHost, c->st));
    if (int rc = read_ctl(c)) return rc;   // synchronises; status and counts
    std::memset(out, 0, sizeof(*out));
    const DCtl &k = *c->ctl_h;
    out->active = hd.active;
    out->recon_count = c->trecon[0];
    out->remove_count = c->tremove[0];
    out->itime = c->itime;
    out->t = c->t;
    out->maxu = hd.maxu;
    out->maxdu = hd.maxdu;
    const double act = (double)hd.active;
    out->mean_u_sup = hd.sum_usup / act;
    out->mean_u_mf = hd.sum_umf / act;
    out->kappa_bar = hd.sum_curv / act;
    // minval of an empty selection is huge(0.) in Fortran
    out->kappa_min = hd.curv_min >= 1.0e300 ? 1.7976931348623157e308 : hd.curv_min;
    out->kappa_max = hd.curv_max;
    out->total_length = k.total_length;
    out->avg_sep = k.total_length / (double)(k.n_old - k.empty_after);
    out->tree_eval = c->stats.tree_evals;
    out->wall_recon_count = c->twall[0];
    return NV_OK;
}

int nv_data_dump(void *h, const char *path) {
    Ctx *c = (Ctx *)h;
    if (!c || !path) return NV_ERR_ARG;
    if (c->K != 1) return fail(c, NV_ERR_UNSUPPORTED, "nv_data_dump: one tangle (the reference's single f)");
    if (int rc = read_ctl(c)) return rc;
    const int n = c->tn[0];
    std::vector<double> v3((size_t)8 * 3 * n), v1((size_t)3 * n);
    std::vector<int32_t> i1((size_t)3 * n), ip((size_t)5 * n);
    nv_state hs{};
    hs.capacity = n;
    double **vp[] = {&hs.x, &hs.u, &hs.u1, &hs.u2, &hs.u_sup, &hs.u_mf, &hs.ghosti, &hs.ghostb};
    for (int q = 0; q < 8; ++q) *vp[q] = v3.data() + (size_t)q * 3 * n;
    hs.closestd = v1.data();
    hs.curv = v1.data() + n;
    hs.stretch = v1.data() + 2 * (size_t)n;
    hs.infront = i1.data();
    hs.behind = i1.data() + n;
    hs.closest = i1.data() + 2 * (size_t)n;
    hs.pinnedi = ip.data();
    hs.pinnedb = ip.data() + n;
    hs.wpinned = ip.data() + 2 * (size_t)n;
    if (int rc = get_batch(c, nullptr, &hs)) return rc;
    const int rc = nv_var_write(path, &hs);
    if (rc) return fail(c, rc, std::string("nv_data_dump: cannot write ") + path);
    return NV_OK;
}

int nv_data_restore(void *h, const char *path) {
    Ctx *c = (Ctx *)h;
    if (!c || !path) return NV_ERR_ARG;
    nv_state probe{};
    int rc = nv_var_read(path, &probe);
    if (rc != NV_ERR_CAPACITY && rc != NV_OK) return fail(c, NV_ERR_ARG, std::string("nv_data_restore: cannot read ") + path);
    const int n = probe.pcount;
    std::vector<double> v3((size_t)8 * 3 * n), v1((size_t)3 * n);
    std::vector<int32_t> i1((size_t)5 * n), w3((size_t)3 * n);
    nv_state hs{};
    hs.capacity = n;
    double **vp[] = {&hs.x, &hs.u, &hs.u1, &hs.u2, &hs.u_sup, &hs.u_mf, &hs.ghosti, &hs.ghostb};
    for (int q = 0; q < 8; ++q) *vp[q] = v3.data() + (size_t)q * 3 * n;
    hs.closestd = v1.data();
    hs.curv = v1.data() + n;
    hs.stretch = v1.data() + 2 * (size_t)n;
    hs.infront = i1.data();
    hs.behind = i1.data() + n;
    hs.closest = i1.data() + 2 * (size_t)n;
    hs.pinnedi = i1.data() + 3 * (size_t)n;
    hs.pinnedb = i1.data() + 4 * (size_t)n;
    hs.wpinned = w3.data();
    if ((rc = nv_var_read(path, &hs))) return fail(c, NV_ERR_ARG, std::string("nv_data_restore: cannot read ") + path);
    // the tree, like the reference's, is not part of the restart; the
    // restart replaces a batch as nv_set_state does
    c->tree_valid = false;
    drop_subs(c);
    const int32_t pc = hs.pcount;
    return set_batch(c, 1, &pc, &hs, nullptr);
}

int nv_printf(void *h, const char *path) {
    Ctx *c = (Ctx *)h;
    if (!c || !path) return NV_ERR_ARG;
    if (c->K != 1) return fail(c, NV_ERR_UNSUPPORTED, "nv_printf: one tangle (the reference's single f)");
    if (int rc = read_ctl(c)) return rc;
    const int n = c->tn[0];
    std::vector<double> v3((size_t)3 * 3 * n), v1((size_t)2 * n);
    std::vector<int32_t> i1((size_t)n);
    nv_state hs{};
    hs.capacity = n;
    hs.x = v3.data();
    hs.u = v3.data() + 3 * (size_t)n;
    hs.u_mf = v3.data() + 6 * (size_t)n;
    hs.curv = v1.data();
    hs.stretch = v1.data() + n;
    hs.infront = i1.data();
    if (int rc = get_batch(c, nullptr, &hs)) return rc;
    if (int rc = nv_var_log_write(path, &hs)) return fail(c, rc, std::string("nv_printf: cannot write ") + path);
    return NV_OK;
}

int nv_print_info(void *h, const nv_diag *d, int32_t shots, const char *path) {
    Ctx *c = (Ctx *)h;
    if (!c || !d || !path || shots < 1) return NV_ERR_ARG;
    FILE *f = std::fopen(path, "a");
    if (!f) return fail(c, NV_ERR_ARG, std::string("nv_print_info: cannot open ") + path);
    std::string s;
    if (d->itime == shots) s = nvio::ts_header();
    s += nvio::ts_line(shots, *d, c->prm.delta);
    const bool ok = std::fputs(s.c_str(), f) >= 0;
    if (std::fclose(f) != 0 || !ok) return fail(c, NV_ERR_ARG, std::string("nv_print_info: cannot write ") + path);
    return NV_OK;
}

int nv_tree_print(void *h, const char *path) {
    Ctx *c = (Ctx *)h;
    if (!c || !path) return NV_ERR_ARG;
    if (forest(c)) return fail(c, NV_ERR_UNSUPPORTED, "nv_tree_print: one tangle (a batch has one octree per tangle)");
    if (!c->tree_valid) return fail(c, NV_ERR_STATE, "nv_tree_print: no tree (construct_tree not called)");
    int n = 0;
    if (int rc = nv_tree_dump(c->tree, c->st, nullptr, 0, &n)) return fail(c, rc, "tree dump failed");
    std::vector<nv_tree_node> nodes((size_t)std::max(n, 1));
    if (int rc = nv_tree_dump(c->tree, c->st, nodes.data(), n, &n)) return fail(c, rc, "tree dump failed");
    FILE *f = std::fopen(path, "a");
    if (!f) return fail(c, NV_ERR_ARG, std::string("nv_tree_print: cannot open ") + path);
    bool ok = true;
    for (int q = 0; q < n && ok; ++q)
        if (nodes[q].pcount <= 1) ok = std::fputs(nvio::tree_leaf_line(nodes[q]).c_str(), f) >= 0;
    if (std::fclose(f) != 0 || !ok) return fail(c, NV_ERR_ARG, std::string("nv_tree_print: cannot write ") + path);
    return NV_OK;
}

int nv_get_recon_events(void *h, nv_recon_event *buf, int32_t cap, int32_t *n) {
    Ctx *c = (Ctx *)h;
    if (!c || !n) return NV_ERR_ARG;
    if (forest(c)) {
        // each tangle's log, its tangle index set, in the order below
        int32_t total = 0, taken = 0;
        for (size_t k = 0; k < c->subs.size(); ++k) {
            int32_t nk = 0;
            nv_recon_event *bk = buf ? buf + taken : nullptr;
            const int32_t room = buf ? std::max<int32_t>(0, cap - taken) : 0;
            if (int rc = nv_get_recon_events(c->subs[k], bk, room, &nk))
                return fail(c, rc, "tangle " + std::to_string(k) + ": " + c->subs[k]->err);
            const int32_t got = std::min(nk, room);
            for (int32_t q = 0; q < got; ++q) bk[q].tangle = (int32_t)k;
            taken += got;
            total += nk;
        }
        if (buf)
            std::stable_sort(buf, buf + taken, [](const nv_recon_event &a, const nv_recon_event &b) {
                if (a.itime != b.itime) return a.itime < b.itime;
                if (a.tangle != b.tangle) return a.tangle < b.tangle;
                return a.i < b.i;
            });
        *n = total;
        return NV_OK;
    }
    if (int rc = read_ctl(c)) return rc;
    const int m = std::min(c->ctl_h->nevents, c->ev_cap);
    const int take = std::min(m, (int)cap);
    if (take > 0 && buf) {
        HIPCHK(c, hipMemcpy(buf, c->ev, sizeof(nv_recon_event) * take, hipMemcpyDeviceToHost));
        // tangles log concurrently: order by (step, tangle, row); within a
        // tangle the serial pass visits rows in ascending order
        std::stable_sort(buf, buf + take, [](const nv_recon_event &a, const nv_recon_event &b) {
            if (a.itime != b.itime) return a.itime < b.itime;
            if (a.tangle != b.tangle) return a.tangle < b.tangle;
            return a.i < b.i;
        });
    }
    *n = c->ctl_h->nevents;
    c->ctl_h->nevents = 0;
    HIPCHK(c, hipMemcpy(&c->ctl->nevents, &c->ctl_h->nevents, sizeof(int), hipMemcpyHostToDevice));
    return NV_OK;
}

int nv_get_tree(void *h, nv_tree_node *buf, int32_t cap, int32_t *n) {
    Ctx *c = (Ctx *)h;
    if (!c || !n) return NV_ERR_ARG;
    if (forest(c)) return fail(c, NV_ERR_UNSUPPORTED, "nv_get_tree: one tangle (a batch has one octree per tangle)");
    if (!c->tree_valid) return fail(c, NV_ERR_STATE, "no tree (construct_tree not called)");
    int rc = nv_tree_dump(c->tree, c->st, buf, cap, n);
    if (rc) return fail(c, rc, "tree dump failed");
    return NV_OK;
}

int nv_get_stats(void *h, nv_stats *s) {
    Ctx *c = (Ctx *)h;
    if (!c || !s) return NV_ERR_ARG;
    if (forest(c)) {
        forest_stats(c, s);
        return NV_OK;
    }
    c->stats.targets_per_lane = c->prm.precision == NV_PREC_FP32C ? c->tpt : 1;
    c->stats.kahan_block = c->prm.precision == NV_PREC_FP32C ? c->kb : 0;
    c->stats.items_per_cu = c->ipc;
    c->stats.source_tile = c->prm.precision == NV_PREC_FP32C ? c->ts : 256;
    *s = c->stats;
    return NV_OK;
}

int nv_autotune(void *h, const char *cache_path) {
    Ctx *c = (Ctx *)h;
    if (!c) return NV_ERR_ARG;
    if (c->prm.velocity != NV_VEL_BS) return NV_OK;   // only the direct sum has variants
    if (forest(c)) return each_sub(c, [&](Ctx *s) { return nv_autotune(s, cache_path); });
    if (c->K == 0) return fail(c, NV_ERR_STATE, "nv_autotune: no state uploaded");
    if (cache_path && !cache_path[0]) cache_path = nullptr;   // "" = no cache
    const bool fp32 = c->prm.precision == NV_PREC_FP32C;
    long total = 0, largest = 0;
    for (int k = 0; k < c->K; ++k) {
        total += c->tn[k];
        largest = std::max<long>(largest, c->tn[k]);
    }
    auto lg = [](long v) { int b = 0; while (v > 1) { v >>= 1; ++b; } return b; };
    int dev = 0;
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipGetDeviceProperties(&prop, dev));
    std::string arch(prop.gcnArchName);
    arch = arch.substr(0, arch.find(':'));
    char key[160];
    std::snprintf(key, sizeof key, "%s|%s|K=%d|n=%d|m=%d", arch.c_str(), fp32 ? "fp32c" : "fp64",
                  c->K == 1 ? 1 : 1 << lg(c->K), lg(total), lg(largest));
    auto apply = [&](int tpt, int kb, int ipc, int ts) {
        c->tpt = tpt;
        c->kb = kb;
        c->ipc = ipc;
        c->ts = ts;
        c->occ32 = nvk_occupancy_bs32c(c->tpt, c->kb, c->ts);
        ++c->layout_gen;   // recorded steps hold the old variant's work lists
    };
    auto valid = [](int t, int k, int i, int ts) {
        return (t == 1 || t == 2) && (k == 1 || k == 2 || k == 4 || k == 8 || k == 16 || k == 32) && i >= 1 &&
               i <= 256 && (ts == 256 || (ts == 128 && k >= 4));
    };
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "r")) {
            char line[512];
            int ht = 0, hk = 0, hi = 0, hs = 256;
            double hm = 0;
            bool hit = false;
            while (std::fgets(line, sizeof line, f)) {
                char k2[160];
                int t2, b2, i2, s2 = 256;
                double m2;
                // lines written before the tile size was tuned have no ts=
                if ((std::sscanf(line, "%159s tpt=%d kb=%d ipc=%d ts=%d ms=%lf", k2, &t2, &b2, &i2, &s2, &m2) == 6 ||
                     (s2 = 256, std::sscanf(line, "%159s tpt=%d kb=%d ipc=%d ms=%lf", k2, &t2, &b2, &i2, &m2) == 5)) &&
                    std::strcmp(k2, key) == 0) {
                    ht = t2; hk = b2; hi = i2; hs = s2; hm = m2;
                    hit = true;   // the last matching line wins
                }
            }
            std::fclose(f);
            if (hit && valid(ht, hk, hi, hs)) {
                apply(ht, hk, hi, hs);
                c->stats.tuned = 2;
                c->stats.tune_ms = hm;
                return NV_OK;
            }
        }
    }
    // time one variant: warm-up launches, then reps timed launches
    int warm = 1, reps = 3;
    auto time_variant = [&](int tpt, int kb, int ipc, int ts, float *ms_out) -> int {
        apply(tpt, kb, ipc, ts);
        int nitems = 0;
        if (int rc = build_items(c, fp32 ? c->tpt : 1, fp32 ? c->ts : 256, c->prm.exact_order != 0, false, &nitems))
            return rc;
        HIPCHK(c, c->queue.reserve(1));
        HIPCHK(c, nvk_src64(c->st, c->d, c->rows, c->rec64.p));
        if (fp32) {
            if (int rc = order_sources(c, false)) return rc;
            HIPCHK(c, nvk_src32(c->st, c->d, c->rows, c->operm.p, c->rec32.p, c->tinfo.p, c->ts));
        }
        Tangles tg = tangles(c, c->d_S.p, c->d_poff.p);
        const double rn = std::max(c->prm.near_radius, 2.0) * c->prm.delta;
        HIPCHK(c, c->pd.reserve(1));
        float tot = 0.f;
        for (int r = 0; r < warm + reps; ++r) {
            HIPCHK(c, nv_fill_int(c->st, c->queue.p, 0, 1));
            HIPCHK(c, hipEventRecord(c->e2, c->st));
            if (fp32) {
                const int grid = std::max(1, std::min(nitems, c->occ32 * c->ncu));
                HIPCHK(c, nvk_bs32c(c->st, c->d, tg, c->items.p, nitems, c->queue.p, grid, c->rec32.p, c->tinfo.p, c->iperm.p,
                                    c->rec64.p, &c->sh, c->phys.kappa, (float)(rn * rn), c->part.p, c->tpt, c->kb, c->ts));
            } else {
                const int grid = std::max(1, std::min(nitems, c->occ64 * c->ncu));
                HIPCHK(c, nvk_bs64(c->st, c->d, tg, c->items.p, nitems, c->queue.p, grid, c->rec64.p, &c->sh,
                                   c->phys.kappa, c->part.p));
            }
            // and the chunk reduction that follows (its volume grows with the chunk count)
            HIPCHK(c, nvk_reduce_probe(c->st, c->d, tg, c->rows, c->part.p, c->pd.p));
            HIPCHK(c, hipEventRecord(c->e3, c->st));
            HIPCHK(c, hipEventSynchronize(c->e3));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->e2, c->e3));
            if (r >= warm) tot += ms;
        }
        *ms_out = tot / reps;
        return NV_OK;
    };
    const bool free_tpt = fp32 && c->prm.targets_per_lane == 0;
    const bool free_kb = fp32 && c->prm.kahan_block == 0;
    const bool free_ts = fp32 && c->prm.source_tile == 0;
    int bt = fp32 ? c->tpt : 1, bk = c->kb, bi = c->ipc, bs = fp32 ? c->ts : 256;
    float best = 0.f;
    if (int rc = time_variant(bt, bk, bi, bs, &best)) return rc;
    // keep the whole search to a few seconds on large batches
    if (best > 200.f) { warm = 0; reps = 1; }
    else if (best > 20.f) reps = 2;
    auto consider = [&](int t, int k, int i, int ts) -> int {
        if ((t == bt && k == bk && i == bi && ts == bs) || !valid(t, k, i, ts)) return NV_OK;
        float ms = 0.f;
        if (int rc = time_variant(t, k, i, ts, &ms)) return rc;
        if (ms < best) {
            best = ms;
            bt = t; bk = k; bi = i; bs = ts;
        }
        return NV_OK;
    };
    for (int i : {8, 16, 32})
        if (int rc = consider(bt, bk, i, bs)) return rc;
    if (free_tpt)
        for (int t : {1, 2})
            if (int rc = consider(t, bk, bi, bs)) return rc;
    if (free_kb)
        for (int k : {4, 8, 16, 32})
            if (int rc = consider(bt, k, bi, bs)) return rc;
    if (free_ts)
        for (int ts : {128, 256})
            if (int rc = consider(bt, bk, bi, ts)) return rc;
    apply(bt, bk, bi, bs);
    c->stats.tuned = 1;
    c->stats.tune_ms = best;
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "a")) {
            std::fprintf(f, "%s tpt=%d kb=%d ipc=%d ts=%d ms=%.4f\n", key, bt, bk, bi, bs, best);
            std::fclose(f);
        }
    }
    return NV_OK;
}

}  // extern "C"
