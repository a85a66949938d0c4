/*
 * Test infrastructure only: the oracle (a CPU restatement of the reference).
 * See nv_oracle.h for conventions.  Only tests/, the smoke check and the CPU
 * baseline leg of bench.py load this file's library; the product never does.
 *
 * Every expression below keeps the Fortran evaluation order of the line it
 * cites (left-to-right, explicit parentheses where the Fortran has implied
 * ones), compiled with -ffp-contract=off and without fast-math.
 */
#include "nv_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979324               /* src/cdata.f90:65 (real*8) */
#define EPS 2.220446049250313080847e-16      /* epsilon(0.) with -fdefault-real-8 */
#define ONE_HALF (1.0 / 2.0)                 /* src/cdata.f90:67 */
#define THREE_TWOS (3.0 / 2.0)               /* src/cdata.f90:68 */
#define TWENTY_THREE_TWELVE (23.0 / 12.0)    /* src/cdata.f90:69 */
#define FOUR_THIRDS (4.0 / 3.0)              /* src/cdata.f90:70 */
#define FIVE_TWELTHS (5.0 / 12.0)            /* src/cdata.f90:71 */

#define X(s, p, k) ((s)->x[((p) - 1) * 3 + (k)])
#define GI(s, p, k) ((s)->gi[((p) - 1) * 3 + (k)])
#define GB(s, p, k) ((s)->gb[((p) - 1) * 3 + (k)])
#define INF(s, p) ((s)->infront[(p) - 1])
#define BEH(s, p) ((s)->behind[(p) - 1])

static int64_t g_tree_eval = 0;
int64_t nvo_tree_eval_count(void) { return g_tree_eval; }

/* src/derivatives.f90:40-47 dist_gen */
static inline double dist_gen(const double *a, const double *b) {
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}
/* src/general.f90:64-71 dist_gen_sq (also distfsq :36-43) */
static inline double dist_gen_sq(const double *a, const double *b) {
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}
/* src/general.f90:152-159 cross_product */
static inline void cross(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* src/derivatives.f90:11-20 get_deriv_1 */
static void get_deriv_1(const nvo_state *s, int32_t i, double *sd) {
    const double *x = &X(s, i, 0), *gi = &GI(s, i, 0), *gb = &GB(s, i, 0);
    double disti = dist_gen(x, gi), distb = dist_gen(x, gb);
    for (int k = 0; k < 3; k++)
        sd[k] = distb * gi[k] + (disti - distb) * x[k] - disti * gb[k];
    double den = 2.0 * distb * disti;
    for (int k = 0; k < 3; k++) sd[k] = sd[k] / den;
}
/* src/derivatives.f90:28-36 get_deriv_2 */
static void get_deriv_2(const nvo_state *s, int32_t i, double *sdd) {
    const double *x = &X(s, i, 0), *gi = &GI(s, i, 0), *gb = &GB(s, i, 0);
    double disti = dist_gen(x, gi), distb = dist_gen(x, gb);
    for (int k = 0; k < 3; k++)
        sdd[k] = 2.0 * (gi[k] / (disti * (disti + distb)) - x[k] / (disti * distb) +
                        gb[k] / (distb * (disti + distb)));
}
/* src/general.f90:106-117 norm_tanf */
static void norm_tanf(const nvo_state *s, int32_t i, double *t) {
    get_deriv_1(s, i, t);
    double length = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int k = 0; k < 3; k++) t[k] = t[k] / length;
}
/* src/general.f90:10-20 clear_particle (u_sup, u_mf, curv, stretch untouched) */
static void clear_particle(nvo_state *s, int32_t i) {
    for (int k = 0; k < 3; k++) {
        X(s, i, k) = 0.0;
        s->u[(i - 1) * 3 + k] = 0.0;
        s->u1[(i - 1) * 3 + k] = 0.0;
        s->u2[(i - 1) * 3 + k] = 0.0;
        GI(s, i, k) = 0.0;
        GB(s, i, k) = 0.0;
        s->wpinned[(i - 1) * 3 + k] = 0;
    }
    INF(s, i) = 0;
    BEH(s, i) = 0;
    s->closest[i - 1] = 0;
    s->closestd[i - 1] = 0.0;
    s->pinnedi[i - 1] = 0;
    s->pinnedb[i - 1] = 0;
}

/* src/boundary.f90:143-175 create_periodic_loop_array */
int nvo_shift_table(const nvo_params *p, double shifts[26][3]) {
    int n = 0;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                double ii, jj, kk;
                if (p->periodic[0]) ii = i * p->box[0]; else { if (i != 0) continue; ii = 0.0; }
                if (p->periodic[1]) jj = j * p->box[1]; else { if (j != 0) continue; jj = 0.0; }
                if (p->periodic[2]) kk = k * p->box[2]; else { if (k != 0) continue; kk = 0.0; }
                if (ii == 0.0 && jj == 0.0 && kk == 0.0) continue;
                shifts[n][0] = ii; shifts[n][1] = jj; shifts[n][2] = kk;
                n++;
            }
    return n;
}

/* src/boundary.f90:180-214 create_mirror_loop_array: the reflections, i, j, k
 * nested over -1..1, a direction that is not solid held at 0 */
int nvo_mirror_table(const nvo_params *p, int32_t reflect[26][3]) {
    int n = 0;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                int ii, jj, kk;
                if (p->solid[0]) ii = i; else { if (i != 0) continue; ii = 0; }
                if (p->solid[1]) jj = j; else { if (j != 0) continue; jj = 0; }
                if (p->solid[2]) kk = k; else { if (k != 0) continue; kk = 0; }
                if (ii == 0 && jj == 0 && kk == 0) continue;
                reflect[n][0] = ii; reflect[n][1] = jj; reflect[n][2] = kk;
                n++;
            }
    return n;
}

/* the reference's reflection of v in the walls of r, v + 2*abs(r)*(0.5*r*box_size - v)
 * (src/timestep.f90:221-222, src/tree.f90:345-348) */
static inline void reflect_point(const double *v, const int32_t *r, const double *box, double *out) {
    for (int k = 0; k < 3; k++) out[k] = v[k] + (double)(2 * abs(r[k])) * ((0.5 * r[k]) * box[k] - v[k]);
}

/* the own-image test of biot_savart_mirror and tree_walk_mirror
 * (src/timestep.f90:215, src/tree.f90:369): sum(wpinned - reflect) == 0 and
 * sum(wpinned) < 2 (sums of components, as the reference writes them) */
static inline int own_image(const nvo_state *s, int32_t i, const int32_t *r) {
    const int32_t *w = &s->wpinned[(i - 1) * 3];
    return ((w[0] - r[0]) + (w[1] - r[1]) + (w[2] - r[2])) == 0 && (w[0] + w[1] + w[2]) < 2;
}

/* src/boundary.f90:26-84 get_ghost_p (pinned reflections at :31-38) */
static void get_ghost_p(const nvo_state *s, const nvo_params *p, int32_t i, double *gin, double *gbe) {
    const int32_t *w = &s->wpinned[(i - 1) * 3];
    if (s->pinnedi[i - 1]) {
        int32_t b = BEH(s, i);
        for (int k = 0; k < 3; k++) {
            gbe[k] = X(s, b, k);
            gin[k] = X(s, b, k) + (double)(2 * abs(w[k])) * (0.5 * w[k] * p->box[k] - X(s, b, k));
        }
    } else if (s->pinnedb[i - 1]) {
        int32_t f = INF(s, i);
        for (int k = 0; k < 3; k++) {
            gin[k] = X(s, f, k);
            gbe[k] = X(s, f, k) + (double)(2 * abs(w[k])) * (0.5 * w[k] * p->box[k] - X(s, f, k));
        }
    } else {
        int32_t f = INF(s, i), b = BEH(s, i);
        for (int k = 0; k < 3; k++) { gin[k] = X(s, f, k); gbe[k] = X(s, b, k); }
    }
    for (int k = 0; k < 3; k++) {
        if (!p->periodic[k]) continue;
        double L = p->box[k], xi = X(s, i, k);
        if ((xi - gin[k]) > (L / 2.0)) gin[k] = gin[k] + L;
        else if ((xi - gin[k]) < (-L / 2.0)) gin[k] = gin[k] - L;
        if ((xi - gbe[k]) > (L / 2.0)) gbe[k] = gbe[k] + L;
        else if ((xi - gbe[k]) < (-L / 2.0)) gbe[k] = gbe[k] - L;
    }
}

/* src/boundary.f90:10-19 ghostp */
void nvo_ghostp(nvo_state *s, const nvo_params *p) {
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) == 0) continue;
        get_ghost_p(s, p, i, &GI(s, i, 0), &GB(s, i, 0));
    }
}

/* ---------------------------------------------------------------- tree --- */
/* src/tree.f90:38-157 construct_tree / build_tree / new_tree.  Node points are
 * kept as lists of f-indices in ascending order, which is the order the
 * reference's parray copies hold them in. */
typedef struct { int32_t *idx; int32_t n; } plist;

static int tree_push(nvo_tree *tr) {
    if (tr->nnodes == tr->cap) {
        int32_t nc = tr->cap ? tr->cap * 2 : 1024;
        nvo_node *nn = (nvo_node *)realloc(tr->nodes, (size_t)nc * sizeof(nvo_node));
        if (!nn) return -1;
        tr->nodes = nn;
        tr->cap = nc;
    }
    return tr->nnodes++;
}

#define MAX_TREE_DEPTH 200

static int build_rec(const nvo_state *s, nvo_tree *tr, int32_t node, const plist *pl, int depth);

/* src/tree.f90:108-157 new_tree: filter the parent's list, accumulate moments */
static int new_tree(const nvo_state *s, nvo_tree *tr, double px, double py, double pz, double w,
                    const plist *parent, int depth, int32_t *out_node) {
    int32_t nd = tree_push(tr);
    if (nd < 0) return NVO_ERR_CAPACITY;
    *out_node = nd;
    nvo_node *n = &tr->nodes[nd];
    n->pos[0] = px; n->pos[1] = py; n->pos[2] = pz; n->width = w;
    double cx = 0.0, cy = 0.0, cz = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
    plist mine;
    mine.idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(parent->n > 0 ? parent->n : 1));
    mine.n = 0;
    for (int32_t q = 0; q < parent->n; q++) {
        int32_t i = parent->idx[q];
        if (INF(s, i) == 0) continue;
        if (s->pinnedi[i - 1]) continue;
        double x0 = X(s, i, 0), x1 = X(s, i, 1), x2 = X(s, i, 2);
        if ((x0 >= px) && (x0 < (px + w)) && (x1 >= py) && (x1 < (py + w)) && (x2 >= pz) &&
            (x2 < (pz + w))) {
            mine.idx[mine.n++] = i;
            cx = cx + x0; cy = cy + x1; cz = cz + x2;
            c0 = c0 + GI(s, i, 0) - x0;
            c1 = c1 + GI(s, i, 1) - x1;
            c2 = c2 + GI(s, i, 2) - x2;
        }
    }
    if (mine.n != 0) { cx = cx / mine.n; cy = cy / mine.n; cz = cz / mine.n; }
    n = &tr->nodes[nd];
    n->cent[0] = cx; n->cent[1] = cy; n->cent[2] = cz;
    n->circ[0] = c0; n->circ[1] = c1; n->circ[2] = c2;
    n->pcount = mine.n;
    n->leaf_infront = mine.n >= 1 ? INF(s, mine.idx[0]) : 0;
    for (int c = 0; c < 8; c++) n->child[c] = -1;
    int rc = build_rec(s, tr, nd, &mine, depth + 1);
    free(mine.idx);
    return rc;
}

/* src/tree.f90:64-104 build_tree: 8 children while pcount>1 in the order
 * fbl, fbr, ftl, ftr, bbl, bbr, btl, btr */
static int build_rec(const nvo_state *s, nvo_tree *tr, int32_t node, const plist *pl, int depth) {
    if (tr->nodes[node].pcount <= 1) return NVO_OK;
    if (depth > MAX_TREE_DEPTH) return NVO_ERR_TREE_DEPTH;
    double px = tr->nodes[node].pos[0], py = tr->nodes[node].pos[1], pz = tr->nodes[node].pos[2];
    double w = tr->nodes[node].width;
    /* children: c = 4*ybit + 2*zbit + xbit */
    for (int c = 0; c < 8; c++) {
        int xb = c & 1, zb = (c >> 1) & 1, yb = (c >> 2) & 1;
        double cx = xb ? px + (w / 2) : px;
        double cy = yb ? py + (w / 2) : py;
        double cz = zb ? pz + (w / 2) : pz;
        int32_t ch;
        int rc = new_tree(s, tr, cx, cy, cz, w / 2.0, pl, depth, &ch);
        if (rc) return rc;
        tr->nodes[node].child[c] = ch;
    }
    return NVO_OK;
}

/* src/tree.f90:38-59 construct_tree.  The root holds every slot (empty ones
 * included) and a centroid over the non-empty points (masked SUM). */
int nvo_tree_build(const nvo_state *s, const nvo_params *p, nvo_tree *tr) {
    tr->nnodes = 0;
    int32_t root = tree_push(tr);
    if (root < 0) return NVO_ERR_CAPACITY;
    double bsize = p->box[0];
    if (p->box[1] > bsize) bsize = p->box[1];
    if (p->box[2] > bsize) bsize = p->box[2];
    nvo_node *r = &tr->nodes[root];
    r->pos[0] = -bsize / 2.0; r->pos[1] = -bsize / 2.0; r->pos[2] = -bsize / 2.0;
    r->width = bsize;
    r->pcount = s->n;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int32_t cnt = 0;
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) > 0) { sx += X(s, i, 0); sy += X(s, i, 1); sz += X(s, i, 2); cnt++; }
    }
    r->cent[0] = sx / cnt; r->cent[1] = sy / cnt; r->cent[2] = sz / cnt;
    r->circ[0] = r->circ[1] = r->circ[2] = 0.0;
    r->leaf_infront = s->n >= 1 ? INF(s, 1) : 0;
    for (int c = 0; c < 8; c++) r->child[c] = -1;
    plist all;
    all.n = s->n;
    all.idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n > 0 ? s->n : 1));
    for (int32_t i = 0; i < s->n; i++) all.idx[i] = i + 1;
    int rc = build_rec(s, tr, root, &all, 0);
    free(all.idx);
    return rc;
}

void nvo_tree_free(nvo_tree *tr) {
    free(tr->nodes);
    tr->nodes = NULL;
    tr->nnodes = tr->cap = 0;
}

/* src/tree.f90:250-316 tree_walk */
static int tree_walk(const nvo_state *s, const nvo_params *p, const nvo_tree *tr, int32_t nd, int32_t i,
                     const double *shift, double *u, int64_t *evals) {
    const nvo_node *n = &tr->nodes[nd];
    if (n->pcount == 0) return NVO_OK;
    const double *xi = &X(s, i, 0);
    double e0 = xi[0] - (n->cent[0] + shift[0]);
    double e1 = xi[1] - (n->cent[1] + shift[1]);
    double e2 = xi[2] - (n->cent[2] + shift[2]);
    double dist = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    double theta;
    if (p->tree_extra_correction) {
        double g0 = n->cent[0] - (n->pos[0] + n->width / 2.0);
        double g1 = n->cent[1] - (n->pos[1] + n->width / 2.0);
        double g2 = n->cent[2] - (n->pos[2] + n->width / 2.0);
        double geo = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        if (dist - geo > 0.0) theta = n->width / (dist - geo);
        else theta = n->width / dist;
    } else {
        theta = n->width / dist;
    }
    if (n->pcount == 1 || theta < p->tree_theta) {
        if (n->pcount == 1) {
            int32_t j = BEH(s, n->leaf_infront);
            double ms = fabs(shift[0]);
            if (fabs(shift[1]) > ms) ms = fabs(shift[1]);
            if (fabs(shift[2]) > ms) ms = fabs(shift[2]);
            if (ms < EPS) {
                if (j == i) return NVO_OK;
                if (j == BEH(s, i)) return NVO_OK;
            }
            if (dist < EPS) return NVO_ERR_SINGULAR;
        }
        double v[3];
        v[0] = ((n->cent[0] + shift[0]) - xi[0]);
        v[1] = ((n->cent[1] + shift[1]) - xi[1]);
        v[2] = ((n->cent[2] + shift[2]) - xi[2]);
        double a = dist * dist;
        double b = 2.0 * dot3(v, n->circ);
        double c = dot3(n->circ, n->circ);
        double D = 4.0 * a * c - b * b;
        if (D < EPS) return NVO_OK;
        double ub[3];
        cross(v, n->circ, ub);
        double den = (2.0 * PI) * D;
        for (int k = 0; k < 3; k++) ub[k] = ub[k] * p->quant_circ / den;
        double fac = (2.0 * c + b) / sqrt(a + b + c) - (b / sqrt(a));
        for (int k = 0; k < 3; k++) u[k] = u[k] + ub[k] * fac;
        (*evals)++;
        return NVO_OK;
    }
    for (int c = 0; c < 8; c++) {
        int rc = tree_walk(s, p, tr, n->child[c], i, shift, u, evals);
        if (rc) return rc;
    }
    return NVO_OK;
}

/* src/tree.f90:321-392 tree_walk_mirror.  theta is width / dist with dist to
 * the reflected end centroid + circ: the extra correction's value is
 * overwritten at :356. */
static int tree_walk_mirror(const nvo_state *s, const nvo_params *p, const nvo_tree *tr, int32_t nd, int32_t i,
                            const int32_t *refl, double *u, int64_t *evals) {
    const nvo_node *n = &tr->nodes[nd];
    if (n->pcount == 0) return NVO_OK;
    const double *xi = &X(s, i, 0);
    double ce[3], simg[3], sgi[3];
    for (int k = 0; k < 3; k++) ce[k] = n->cent[k] + n->circ[k];
    reflect_point(n->cent, refl, p->box, sgi);
    reflect_point(ce, refl, p->box, simg);
    double dist = dist_gen(xi, simg);
    double theta = n->width / dist;
    if (n->pcount == 1 || theta < p->tree_theta) {
        if (n->pcount == 1) {
            int32_t j = BEH(s, n->leaf_infront);
            if (s->pinnedi[j - 1]) return NVO_OK;
            if (own_image(s, i, refl)) {
                if (s->pinnedb[i - 1] && i == j) return NVO_OK;
                if (s->pinnedi[i - 1] && j == BEH(s, i)) return NVO_OK;
            }
            if (dist < EPS) return NVO_ERR_SINGULAR;
        }
        double d[3], e[3];
        for (int k = 0; k < 3; k++) { d[k] = simg[k] - xi[k]; e[k] = sgi[k] - simg[k]; }
        double a = dist * dist;
        double b = 2.0 * dot3(d, e);
        double c = dist_gen_sq(simg, sgi);
        if (4.0 * a * c - b * b < EPS) return NVO_OK;
        double D = 4.0 * a * c - b * b;
        double ub[3];
        cross(d, e, ub);
        double den = (2.0 * PI) * D;
        for (int k = 0; k < 3; k++) ub[k] = ub[k] * p->quant_circ / den;
        double fac = (2.0 * c + b) / sqrt(a + b + c) - (b / sqrt(a));
        for (int k = 0; k < 3; k++) u[k] = u[k] + ub[k] * fac;
        (*evals)++;
        return NVO_OK;
    }
    for (int c = 0; c < 8; c++) {
        int rc = tree_walk_mirror(s, p, tr, n->child[c], i, refl, u, evals);
        if (rc) return rc;
    }
    return NVO_OK;
}

/* src/tree.f90:206-244 closest_tree */
static void closest_tree(nvo_state *s, const nvo_tree *tr, int32_t nd, int32_t i) {
    const nvo_node *n = &tr->nodes[nd];
    if (INF(s, i) == 0) return;
    if (n->pcount == 0) return;
    const double *xi = &X(s, i, 0);
    double e0 = xi[0] - n->cent[0], e1 = xi[1] - n->cent[1], e2 = xi[2] - n->cent[2];
    double dist = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    double theta = n->width / dist;
    if (n->pcount == 1 || theta < 0.5) {
        if (dist < s->closestd[i - 1]) {
            int32_t j = BEH(s, n->leaf_infront);
            if (n->pcount > 1) {
                s->closest[i - 1] = 0;
            } else if ((i != j) && (INF(s, i) != j) && (BEH(s, i) != j)) {
                if (s->pinnedi[j - 1] || s->pinnedb[j - 1]) {
                } else {
                    s->closest[i - 1] = j;
                    s->closestd[i - 1] = dist;
                }
            }
        }
        return;
    }
    for (int c = 0; c < 8; c++) closest_tree(s, tr, n->child[c], i);
}

/* src/tree.f90:190-203 pclose_tree */
void nvo_pclose_tree(nvo_state *s, const nvo_params *p, const nvo_tree *tr) {
    (void)p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 1; i <= s->n; i++) {
        s->closestd[i - 1] = 100.0;
        if (INF(s, i) == 0) continue;
        if (s->pinnedi[i - 1] || s->pinnedb[i - 1]) continue;
        closest_tree(s, tr, 0, i);
    }
}

/* ------------------------------------------------------------ velocity --- */
/* src/timestep.f90:151-172 biot_savart (shift == NULL) and :175-198
 * biot_savart_shift: accumulate into u the sum over source segments j. */
static void biot_savart(const nvo_state *s, const nvo_params *p, int32_t i, const double *shift, double *u) {
    const double *xi = &X(s, i, 0);
    int32_t bi = BEH(s, i);
    for (int32_t j = 1; j <= s->n; j++) {
        if (shift == NULL) {
            if ((INF(s, j) == 0) || (i == j) || (bi == j)) continue;
        } else {
            if ((INF(s, j) == 0) || (bi == j)) continue;
        }
        if (s->pinnedi[j - 1]) continue;
        const double *xj = &X(s, j, 0), *gj = &GI(s, j, 0);
        double d[3], e[3];
        if (shift == NULL) {
            for (int k = 0; k < 3; k++) d[k] = xj[k] - xi[k];
        } else {
            for (int k = 0; k < 3; k++) d[k] = (xj[k] + shift[k]) - xi[k];
        }
        for (int k = 0; k < 3; k++) e[k] = gj[k] - xj[k];
        double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double b = 2.0 * dot3(d, e);
        double c = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        double D = 4.0 * a * c - b * b;
        if (D == 0.0) continue;
        double ub[3];
        cross(d, e, ub);
        double den = (2.0 * PI) * D;
        for (int k = 0; k < 3; k++) ub[k] = ub[k] * p->quant_circ / den;
        double fac = (2.0 * c + b) / sqrt(a + b + c) - (b / sqrt(a));
        for (int k = 0; k < 3; k++) u[k] = u[k] + ub[k] * fac;
    }
}

/* src/timestep.f90:201-234 biot_savart_mirror: the image segment runs from
 * the reflected ghost to the reflected point */
static void biot_savart_mirror(const nvo_state *s, const nvo_params *p, int32_t i, const int32_t *refl, double *u) {
    const double *xi = &X(s, i, 0);
    const int own = own_image(s, i, refl);
    for (int32_t j = 1; j <= s->n; j++) {
        if ((INF(s, j) == 0) || s->pinnedi[j - 1]) continue;
        if (own) {
            if (s->pinnedb[i - 1] && (i == j)) continue;
            if (s->pinnedi[i - 1] && (j == BEH(s, i))) continue;
        }
        double sgi[3], simg[3];
        reflect_point(&X(s, j, 0), refl, p->box, sgi);
        reflect_point(&GI(s, j, 0), refl, p->box, simg);
        double a = dist_gen_sq(xi, simg);
        double d[3], e[3];
        for (int k = 0; k < 3; k++) { d[k] = simg[k] - xi[k]; e[k] = sgi[k] - simg[k]; }
        double b = 2.0 * dot3(d, e);
        double c = dist_gen_sq(simg, sgi);
        double D = 4.0 * a * c - b * b;
        if (D == 0.0) continue;
        double ub[3];
        cross(d, e, ub);
        double den = (2.0 * PI) * D;
        for (int k = 0; k < 3; k++) ub[k] = ub[k] * p->quant_circ / den;
        double fac = (2.0 * c + b) / sqrt(a + b + c) - (b / sqrt(a));
        for (int k = 0; k < 3; k++) u[k] = u[k] + ub[k] * fac;
    }
}

/* src/general.f90:46-51: the module's cosh, which hides the intrinsic in
 * normal_fluid (it `use`s general) */
static double ref_cosh(double x) { return 0.5 * (exp(x) + exp(-x)); }

/* src/normal_fluid.f90:83-118 get_normal_velocity at position x, time t;
 * the DNS models (laizetDNS, turbDNS) are not restated */
static void normal_velocity(const nvo_params *p, const double *x, double t, double *un) {
    un[0] = un[1] = un[2] = 0.0;
    switch (p->normal_velocity) {
    case 1: /* :96-97 xflow */
        un[0] = p->norm_vel_xflow;
        break;
    case 2: /* :98-99 noisey_xflow */
        un[0] = (p->noise && p->noise_counter ? p->noise[*p->noise_counter - 1] : 0.0) * p->norm_vel_xflow;
        break;
    case 3: /* :101-102 shear_xflow */
        un[0] = p->norm_vel_xflow * sin(2 * PI * x[2] / p->box[2] + t * p->norm_shear_omega);
        break;
    case 4: { /* :103-104 ychannel */
        double q = 2. * x[1] / p->box[1];
        un[0] = p->norm_vel_xflow * (1. - q * q);
        break;
    }
    case 5: { /* :105-112 channel */
        double cossum = 0.;
        for (int ii = 1; ii <= 10; ii++) {
            double ak = (double)(2 * ii - 1) * (PI / 2);
            int sgn = (ii % 2) ? -1 : 1; /* (-1)**ii */
            double ak3 = ak * ak * ak;
            cossum = cossum + 4 * ((double)sgn / ak3) *
                                  (ref_cosh(ak * 2 * x[1] / p->box[1]) * cos(ak * 2 * x[2] / p->box[2])) /
                                  ref_cosh(ak);
        }
        double q = 2 * x[2] / p->box[2];
        un[0] = p->norm_vel_xflow * (1. - q * q + cossum) * 1.2;
        break;
    }
    default:
        break;
    }
    for (int k = 0; k < 3; k++) un[k] = un[k] + p->super_velocity[k];
}

static int shift_table_cache(const nvo_params *p, double sh[26][3]) { return nvo_shift_table(p, sh); }

/* src/timestep.f90:49-141 calc_velocity (LIA :60-73, BS :74-98, Tree :99-121) */
static int calc_velocity(nvo_state *s, const nvo_params *p, const nvo_tree *tr, int32_t i, int nsh,
                         double sh[26][3], int nmir, int32_t mir[26][3], double *u, int64_t *evals) {
    double fd[3], fdd[3], beta;
    get_deriv_1(s, i, fd);
    get_deriv_2(s, i, fdd);
    if (p->velocity == 0) {
        double curv = sqrt(dot3(fdd, fdd));
        if (curv < EPS) curv = EPS;
        else curv = 1.0 / curv;
        beta = (p->quant_circ / (4.0 * PI)) * log(4.6 * curv / p->corea);
        cross(fd, fdd, u);
        for (int k = 0; k < 3; k++) u[k] = beta * u[k];
    } else {
        const double *xi = &X(s, i, 0);
        beta = (p->quant_circ / (4.0 * PI)) *
               log((1.1 / p->corea) * sqrt(dist_gen(xi, &GI(s, i, 0)) * dist_gen(xi, &GB(s, i, 0))));
        double c[3];
        cross(fd, fdd, c);
        for (int k = 0; k < 3; k++) u[k] = beta * c[k];
        double ub[3] = {0.0, 0.0, 0.0};
        static const double zero3[3] = {0.0, 0.0, 0.0};
        if (p->velocity == 1) {
            biot_savart(s, p, i, NULL, ub);
        } else {
            int rc = tree_walk(s, p, tr, 0, i, zero3, ub, evals);
            if (rc) return rc;
        }
        for (int k = 0; k < 3; k++) u[k] = u[k] + ub[k];
        ub[0] = ub[1] = ub[2] = 0.0;
        for (int m = 0; m < nsh; m++) {
            if (p->velocity == 1) {
                biot_savart(s, p, i, sh[m], ub);
            } else {
                int rc = tree_walk(s, p, tr, 0, i, sh[m], ub, evals);
                if (rc) return rc;
            }
        }
        for (int k = 0; k < 3; k++) u[k] = u[k] + ub[k];
        /* solid walls: image vortices (:94-98, :116-120) */
        double um[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < nmir; m++) {
            if (p->velocity == 1) {
                biot_savart_mirror(s, p, i, mir[m], um);
            } else {
                int rc = tree_walk_mirror(s, p, tr, 0, i, mir[m], um, evals);
                if (rc) return rc;
            }
        }
        for (int k = 0; k < 3; k++) u[k] = u[k] + um[k];
    }
    for (int k = 0; k < 3; k++) s->u_sup[(i - 1) * 3 + k] = u[k];
    if ((fabs(p->alpha[0]) > EPS) || (fabs(p->alpha[1]) > EPS)) {
        if (s->t < p->normal_fluid_cutoff) {
            double un[3], dv[3], c1[3], c2[3], mf[3];
            if (s->t < p->t_zero_normal_fluid) normal_velocity(p, &X(s, i, 0), s->t, un);
            else un[0] = un[1] = un[2] = 0.0;
            for (int k = 0; k < 3; k++) dv[k] = un[k] - u[k];
            cross(fd, dv, c1);
            cross(fd, c1, c2);
            for (int k = 0; k < 3; k++) mf[k] = p->alpha[0] * c1[k] - p->alpha[1] * c2[k];
            for (int k = 0; k < 3; k++) u[k] = u[k] + mf[k];
            for (int k = 0; k < 3; k++) s->u_mf[(i - 1) * 3 + k] = mf[k];
        }
    }
    for (int k = 0; k < 3; k++) u[k] = u[k] + p->super_velocity[k];
    return NVO_OK;
}

/* velocity of points [first, first+count) (1-based) into u_out[count][3];
 * empty points are skipped (their rows untouched).  Also stores u_sup/u_mf. */
int nvo_velocity(nvo_state *s, const nvo_params *p, const nvo_tree *tr, int32_t first, int32_t count,
                 double *u_out) {
    double sh[26][3];
    int nsh = shift_table_cache(p, sh);
    int32_t mir[26][3];
    int nmir = nvo_mirror_table(p, mir);
    int err = NVO_OK;
#ifdef _OPENMP
    if (p->threads > 0) omp_set_num_threads(p->threads);
#endif
    int64_t evals = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : evals)
    for (int32_t q = 0; q < count; q++) {
        int32_t i = first + q;
        if (INF(s, i) == 0) continue;
        double u[3];
        int rc = calc_velocity(s, p, tr, i, nsh, sh, nmir, mir, u, &evals);
        if (rc) {
#pragma omp critical
            err = rc;
        }
        for (int k = 0; k < 3; k++) u_out[q * 3 + k] = u[k];
    }
    g_tree_eval += evals;
    return err;
}

/* src/timestep.f90:28-43 the Adams-Bashforth loop */
void nvo_ab_update(nvo_state *s, const nvo_params *p) {
    double dt = p->dt;
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) == 0) continue;
        double *x = &s->x[(i - 1) * 3], *u = &s->u[(i - 1) * 3], *u1 = &s->u1[(i - 1) * 3],
               *u2 = &s->u2[(i - 1) * 3];
        double m1 = fabs(u1[0]);
        if (fabs(u1[1]) > m1) m1 = fabs(u1[1]);
        if (fabs(u1[2]) > m1) m1 = fabs(u1[2]);
        double m2 = fabs(u2[0]);
        if (fabs(u2[1]) > m2) m2 = fabs(u2[1]);
        if (fabs(u2[2]) > m2) m2 = fabs(u2[2]);
        if (m1 == 0.0) {
            for (int k = 0; k < 3; k++) x[k] = x[k] + dt * u[k];
        } else if (m2 == 0.0) {
            for (int k = 0; k < 3; k++) x[k] = x[k] + THREE_TWOS * dt * u[k] - ONE_HALF * dt * u1[k];
        } else {
            for (int k = 0; k < 3; k++)
                x[k] = x[k] + TWENTY_THREE_TWELVE * dt * u[k] - FOUR_THIRDS * dt * u1[k] +
                       FIVE_TWELTHS * dt * u2[k];
        }
        for (int k = 0; k < 3; k++) { u2[k] = u1[k]; u1[k] = u[k]; }
    }
}

/* src/timestep.f90:16-45 pmotion */
int nvo_pmotion(nvo_state *s, const nvo_params *p, const nvo_tree *tr) {
    if (p->threads > 0) {
#ifdef _OPENMP
        omp_set_num_threads(p->threads);
#endif
    }
    double *u = (double *)malloc(sizeof(double) * 3 * (size_t)(s->n > 0 ? s->n : 1));
    if (!u) return NVO_ERR_CAPACITY;
    memcpy(u, s->u, sizeof(double) * 3 * (size_t)s->n);
    int rc = nvo_velocity(s, p, tr, 1, s->n, u);
    memcpy(s->u, u, sizeof(double) * 3 * (size_t)s->n);
    free(u);
    if (rc) return rc;
    nvo_ab_update(s, p);
    /* src/normal_fluid.f90:340-347 randomise_normal_fluid (:44) */
    if (p->normal_velocity == 2 && p->noise_counter) {
        *p->noise_counter = *p->noise_counter + 1;
        if (*p->noise_counter > p->noise_n) *p->noise_counter = 1;
    }
    return NVO_OK;
}

/* ------------------------------------------------------------ topology --- */
/* src/line.f90:18-96 pinsert */
/* total_length and avg_sep of the last pinsert (module globals of the
 * reference, written by pinsert and printed by the diagnostics) */
static double g_total_length, g_avg_sep;
double nvo_total_length(void) { return g_total_length; }
double nvo_avg_sep(void) { return g_avg_sep; }

int nvo_pinsert(nvo_state *s, const nvo_params *p) {
    int32_t active = 0;
    for (int32_t i = 1; i <= s->n; i++) active += INF(s, i) > 0;
    if (active == 0) return NVO_ERR_ZERO_LENGTH;
    int32_t old = s->n;
    double total = 0.0;
    for (int32_t i = 1; i <= old; i++) {
        if (INF(s, i) == 0) continue;
        if (s->pinnedi[i - 1]) continue;
        double disti = dist_gen(&X(s, i, 0), &GI(s, i, 0));
        total = total + disti;
        if (disti > p->delta) {
            int32_t slot = 0;
            for (int32_t q = 1; q <= s->n; q++)
                if (INF(s, q) == 0) { slot = q; break; }
            if (slot == 0) {
                if (s->n >= s->cap) return NVO_ERR_CAPACITY;
                slot = s->n + 1;
                /* a grown slot: the reference leaves everything but infront
                 * undefined; zero it here */
                memset(&s->x[(slot - 1) * 3], 0, 3 * sizeof(double));
                memset(&s->u_sup[(slot - 1) * 3], 0, 3 * sizeof(double));
                memset(&s->u_mf[(slot - 1) * 3], 0, 3 * sizeof(double));
                s->closest[slot - 1] = 0;
                s->closestd[slot - 1] = 0.0;
                INF(s, slot) = 0;
                s->n = s->n + 1;
            }
            double fdd[3];
            get_deriv_2(s, i, fdd);
            double curv = sqrt(dot3(fdd, fdd));
            double *xn = &X(s, slot, 0), *xi = &X(s, i, 0), *gi = &GI(s, i, 0);
            if (curv > 1e-5) {
                curv = 1.0 / curv;
                if (curv * curv - 0.25 * (disti * disti) > 0.0) {
                    double r = (sqrt(curv * curv - 0.25 * (disti * disti)) - curv) * curv;
                    for (int k = 0; k < 3; k++) xn[k] = 0.5 * (xi[k] + gi[k]) + r * fdd[k];
                } else {
                    for (int k = 0; k < 3; k++) xn[k] = 0.5 * (xi[k] + gi[k]);
                }
            } else {
                for (int k = 0; k < 3; k++) xn[k] = 0.5 * (xi[k] + gi[k]);
            }
            int32_t fi = INF(s, i);
            for (int k = 0; k < 3; k++) {
                s->u[(slot - 1) * 3 + k] = 0.5 * (s->u[(i - 1) * 3 + k] + s->u[(fi - 1) * 3 + k]);
                s->u1[(slot - 1) * 3 + k] = 0.0;
                s->u2[(slot - 1) * 3 + k] = 0.0;
                s->wpinned[(slot - 1) * 3 + k] = 0;
            }
            s->pinnedi[slot - 1] = 0;
            s->pinnedb[slot - 1] = 0;
            BEH(s, slot) = i;
            INF(s, slot) = INF(s, i);
            get_ghost_p(s, p, slot, &GI(s, slot, 0), &GB(s, slot, 0));
            BEH(s, INF(s, i)) = slot;
            get_ghost_p(s, p, INF(s, i), &GI(s, INF(s, i), 0), &GB(s, INF(s, i), 0));
            INF(s, i) = slot;
            get_ghost_p(s, p, i, &GI(s, i, 0), &GB(s, i, 0));
        }
    }
    /* :95: the denominator is old_pcount less the empty slots at the end */
    int32_t empty = 0;
    for (int32_t q = 1; q <= s->n; q++) empty += INF(s, q) == 0;
    g_total_length = total;
    g_avg_sep = total / (double)(old - empty);
    return NVO_OK;
}

/* src/reconnection.f90:155-182 loop_killer */
void nvo_loop_killer(nvo_state *s, int32_t particle) {
    int32_t counter = 1, next = particle;
    for (int32_t it = 1; it <= s->n; it++) {
        if (s->pinnedi[next - 1] || s->pinnedb[next - 1]) return;
        next = INF(s, next);
        if (next == particle) break;
        if (next == 0) return; /* broken chain: the reference would read f(0) */
        counter = counter + 1;
    }
    if (counter < 4) {
        next = particle;
        for (int32_t it = 1; it <= s->n; it++) {
            int32_t store_next = INF(s, next);
            clear_particle(s, next);
            next = store_next;
            if (next == particle || next == 0) break;
        }
    }
}

/* src/line.f90:102-130 premove */
void nvo_premove(nvo_state *s, const nvo_params *p) {
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) == 0) continue;
        if (s->pinnedi[i - 1]) continue;
        int32_t infront = INF(s, i);
        if (s->pinnedi[infront - 1]) continue;
        int32_t ii = INF(s, INF(s, i));
        double distii = dist_gen(&X(s, i, 0), &X(s, ii, 0));
        if (distii < .99 * p->delta) {
            int32_t tinfront = INF(s, INF(s, i));
            BEH(s, tinfront) = i;
            INF(s, i) = tinfront;
            clear_particle(s, infront);
            nvo_loop_killer(s, i);
            s->remove_count = s->remove_count + 1;
        }
        if (INF(s, i) != 0 && s->pinnedb[i - 1] && s->pinnedi[INF(s, i) - 1]) {
            clear_particle(s, INF(s, i));
            clear_particle(s, i);
        }
    }
}

/* src/line.f90:135-160 pclose */
void nvo_pclose(nvo_state *s, const nvo_params *p) {
    (void)p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) == 0) continue;
        s->closestd[i - 1] = 100.0;
        if (s->pinnedi[i - 1] || s->pinnedb[i - 1]) continue;
        for (int32_t j = 1; j <= s->n; j++) {
            if (INF(s, j) == 0) continue;
            if ((i != j) && (INF(s, i) != j) && (BEH(s, i) != j) && !s->pinnedi[j - 1] &&
                !s->pinnedb[j - 1]) {
                double dist = dist_gen(&X(s, i, 0), &X(s, j, 0));
                if (dist < s->closestd[i - 1]) {
                    s->closest[i - 1] = j;
                    s->closestd[i - 1] = dist;
                }
            }
        }
    }
}

/* src/reconnection.f90:11-61 precon */
int nvo_precon(nvo_state *s, const nvo_params *p, nvo_recon_log *log) {
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) == 0) continue;
        int32_t pari = INF(s, i);
        if ((s->closestd[i - 1] < p->delta / 2.0) && (s->closestd[i - 1] > EPS)) {
            int32_t j = s->closest[i - 1];
            if (j == pari) continue;
            if (j == BEH(s, i)) continue;
            if (j == 0) continue;
            if (INF(s, j) == 0) continue;
            int32_t parjb = BEH(s, j);
            double t1[3], t2[3];
            norm_tanf(s, i, t1);
            norm_tanf(s, j, t2);
            double dot_val = dot3(t1, t2);
            if (dot_val > 0.9) continue;
            const double *xi = &X(s, i, 0), *xj = &X(s, j, 0);
            double l_before = dist_gen(xi, &GI(s, i, 0)) + dist_gen(xi, &GB(s, i, 0)) +
                              dist_gen(xj, &GI(s, j, 0)) + dist_gen(xj, &GB(s, j, 0));
            double l_after = dist_gen(xi, &GB(s, i, 0)) + dist_gen(xi, xj) + dist_gen(xj, &GI(s, j, 0)) +
                             dist_gen(&X(s, pari, 0), &X(s, parjb, 0));
            if (l_after <= l_before) {
                s->recon_count = s->recon_count + 1;
                if (log && log->n < log->cap) {
                    nvo_recon_event *e = &log->ev[log->n];
                    e->itime = s->itime; e->i = i; e->j = j; e->pad = 0;
                    e->t = s->t;
                    for (int k = 0; k < 3; k++) e->mid[k] = 0.5 * (xi[k] + xj[k]);
                    e->angle = acos(dot_val);
                    e->dl = l_before - l_after;
                }
                if (log) log->n++;
                INF(s, parjb) = pari;
                BEH(s, pari) = parjb;
                INF(s, i) = j;
                BEH(s, j) = i;
                nvo_loop_killer(s, pari);
                nvo_loop_killer(s, i);
            }
        }
    }
    return NVO_OK;
}

/* src/boundary.f90:91-138 enforce_boundary: the periodic wrap, then a
 * pinned point back onto its wall (the first component of largest |wpinned|,
 * maxloc) */
void nvo_enforce_boundary(nvo_state *s, const nvo_params *p) {
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) == 0) continue;
        for (int k = 0; k < 3; k++) {
            if (!p->periodic[k]) continue;
            double L = p->box[k];
            if (X(s, i, k) > (L / 2.0)) X(s, i, k) = X(s, i, k) - L;
            else if (X(s, i, k) < (-L / 2.0)) X(s, i, k) = X(s, i, k) + L;
        }
        if (s->pinnedi[i - 1] || s->pinnedb[i - 1]) {
            const int32_t *w = &s->wpinned[(i - 1) * 3];
            int pc = 0;
            for (int k = 1; k < 3; k++)
                if (abs(w[k]) > abs(w[pc])) pc = k;
            if (w[pc] > 0) X(s, i, pc) = 0.499999999 * p->box[pc];
            else if (w[pc] < 0) X(s, i, pc) = -0.499999999 * p->box[pc];
        }
    }
}

/* src/reconnection.f90:66-150 wall_recon: a point within delta/2 of a solid
 * wall is removed and its neighbours pinned to that wall (behind pinned
 * infront, infront pinned behind); a point pinned at both ends is cleared,
 * as is a pinned pair with nothing between. */
void nvo_wall_recon(nvo_state *s, const nvo_params *p) {
    for (int32_t i = 1; i <= s->n; i++) {
        if (INF(s, i) == 0) continue;
        if (s->pinnedb[i - 1] && s->pinnedi[INF(s, i) - 1]) {
            clear_particle(s, INF(s, i));
            clear_particle(s, i);
        }
        if (s->pinnedi[i - 1] || s->pinnedb[i - 1]) continue;
        for (int k = 0; k < 3; k++) {
            if (!p->solid[k]) continue;
            int side = 0;
            if (X(s, i, k) > 0.5 * (p->box[k] - p->delta)) side = 1;
            else if (X(s, i, k) < -0.5 * (p->box[k] - p->delta)) side = -1;
            if (!side) continue;
            int32_t pari = INF(s, i), parb = BEH(s, i);
            s->pinnedb[pari - 1] = 1;
            BEH(s, pari) = pari;
            s->pinnedi[parb - 1] = 1;
            INF(s, parb) = parb;
            for (int q = 0; q < 3; q++) {
                s->wpinned[(pari - 1) * 3 + q] = q == k ? side : 0;
                s->wpinned[(parb - 1) * 3 + q] = q == k ? side : 0;
            }
            clear_particle(s, i);
            if (s->pinnedi[pari - 1] && s->pinnedb[pari - 1]) clear_particle(s, pari);
            if (s->pinnedi[parb - 1] && s->pinnedb[parb - 1]) clear_particle(s, parb);
            s->wall_recon_count = s->wall_recon_count + 1;
        }
    }
}

/* src/general.f90:172-191 NAN_finder */
int nvo_nan_check(const nvo_state *s) {
    for (int32_t i = 0; i < s->n * 3; i++)
        if (isnan(s->x[i])) return NVO_ERR_NAN;
    return NVO_OK;
}

/* src/run.f90:24-65: one call = nsteps iterations of the main loop body
 * (diagnostics and file output excluded). */
int nvo_step(nvo_state *s, const nvo_params *p, int32_t nsteps, nvo_recon_log *log) {
    nvo_tree tr = {0};
    for (int32_t st = 0; st < nsteps; st++) {
        s->itime = s->itime + 1;
        nvo_ghostp(s, p);
        int rc;
        if (p->tree_theta > 0) {
            rc = nvo_tree_build(s, p, &tr);
            if (rc) { nvo_tree_free(&tr); return rc; }
        }
        rc = nvo_pmotion(s, p, p->tree_theta > 0 ? &tr : NULL);
        if (rc) { nvo_tree_free(&tr); return rc; }
        rc = nvo_pinsert(s, p);
        if (rc) { nvo_tree_free(&tr); return rc; }
        if (s->itime % p->recon_shots == 0) {
            if (p->tree_theta > 0) nvo_pclose_tree(s, p, &tr);
            else nvo_pclose(s, p);
            nvo_precon(s, p, log);
            nvo_wall_recon(s, p);
            nvo_premove(s, p);
        }
        nvo_enforce_boundary(s, p);
        s->t = s->t + p->dt;
        rc = nvo_nan_check(s);
        if (rc) { nvo_tree_free(&tr); return rc; }
    }
    nvo_tree_free(&tr);
    return NVO_OK;
}
