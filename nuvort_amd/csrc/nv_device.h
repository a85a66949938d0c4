// Device-side data layout and FP64 helpers shared by every kernel.
//
// Point rows are 0-based on the device (row r holds the reference's point
// r+1); link arrays keep the reference's 1-based values, 0 = empty slot.
// Vector fields are component-major: component k of row r is a[k*cap + r],
// so a wavefront reading consecutive rows reads consecutive addresses.
//
// All FP64 helpers evaluate their expressions in the order of the Fortran
// line they cite; the library is compiled with -ffp-contract=off so no
// multiply-add is fused behind their back.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NV_PI 3.14159265358979324              // src/cdata.f90:65
#define NV_EPS 2.220446049250313080847e-16     // epsilon(0.) for real*8

struct DState {
    int cap;
    double *x, *xn;                 // positions and the AB output buffer
    double *u, *u1, *u2, *usup, *umf, *gi, *gb;
    int *infront, *behind, *closest;
    double *closestd;
    double *curv, *stretch;         // diagnostics: |s''| and u . s'' (src/diagnostic.f90:37-74)
    int *pin;                       // solid walls: bit 0 pinnedi, bit 1 pinnedb (src/cdata.f90:28)
    int *wp;                        // wpinned, [3][cap] (src/cdata.f90:27)
};

constexpr int NV_PINNED_I = 1, NV_PINNED_B = 2;

// Independent tangles sharing the state arrays.  Tangle k owns the rows
// [base[k], base[k] + cap[k]) (both multiples of 256, so source tiles never
// straddle tangles); its first n[k] rows are its point array f(1:pcount),
// links hold global 1-based rows.  A single filament system is K = 1.
struct Tangles {
    int K;
    const int *base, *cap;
    int *n;                 // pcount of each tangle (updated on the device)
    int *recon, *remove;    // recon_count / remove_count of each tangle
    int *nact;              // active points of each tangle (pinsert pre-filter)
    int *wall;              // wall_recon_count of each tangle
    const int *row2t;       // tangle of every row
    const int *S;           // direct-sum chunks per target of each tangle (this launch)
    const long long *poff;  // offset of each tangle's partial block (this launch)
};

// one unit of direct-sum work: a block of targets of tangle k against the
// c-th of S contiguous chunks of its (image, source tile) sequence
struct Item {
    int k, tb, c, pad;
};

// run.in parameters the kernels read (src/cdata.f90:76-127)
struct Phys {
    double delta, dt, kappa, corea;
    double box[3];
    int periodic[3];
    int velocity;           // 0 LIA, 1 BS, 2 Tree
    double tree_theta;
    int extra_corr;
    int nf;                 // normal_velocity: NV_NF_* (0 zero, 1 xflow, 2 noisey_xflow, 3 shear_xflow,
                            // 4 ychannel, 5 channel)
    double xflow;
    double alpha[2];
    double sup[3];
    double nf_cutoff, t_zero;
    double shear_omega;     // norm_shear_omega
    const double *noise;    // noisey_xflow table (device), noise_n values
    int noise_n;
    int solid[3];           // solid walls at +-box/2 in that direction
    int walls;              // any solid wall: mirror images, pinned points, wall_recon
    int nmir;               // mirror images (src/boundary.f90:180-214 order)
    signed char mir[26][3];
};

struct Shifts {
    double s[27][3];        // s[0] = 0, s[1..26] = src/boundary.f90:143-175 order
    int n;                  // 1 + number of periodic images
};

__device__ __forceinline__ double ld(const double *a, int cap, int k, int r) { return a[k * cap + r]; }

// src/derivatives.f90:40-47 dist_gen
__device__ __forceinline__ double dist_gen(double a0, double a1, double a2, double b0, double b1, double b2) {
    double d0 = a0 - b0, d1 = a1 - b1, d2 = a2 - b2;
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// the reference's reflection of coordinate v in the wall of direction code r
// (-1, 0, 1): v + 2*abs(r)*(0.5*r*L - v) (src/boundary.f90:33-37,
// src/timestep.f90:221-222, src/tree.f90:345-348)
__device__ __forceinline__ double reflect1(double v, int r, double L) {
    return v + (double)(2 * abs(r)) * ((0.5 * (double)r) * L - v);
}

// src/boundary.f90:26-84 get_ghost_p of row r: a point pinned in front takes
// its behind's mirror image as its ghost in front, one pinned behind its
// infront's as its ghost behind (:31-38)
__device__ __forceinline__ void ghost_of(const DState &s, const Phys &p, int r, double gi[3], double gb[3]) {
    const int pin = p.walls ? s.pin[r] : 0;
    if (pin & NV_PINNED_I) {
        const int b = s.behind[r] - 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double xb = ld(s.x, s.cap, k, b);
            gb[k] = xb;
            gi[k] = reflect1(xb, s.wp[k * s.cap + r], p.box[k]);
        }
    } else if (pin & NV_PINNED_B) {
        const int f = s.infront[r] - 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double xf = ld(s.x, s.cap, k, f);
            gi[k] = xf;
            gb[k] = reflect1(xf, s.wp[k * s.cap + r], p.box[k]);
        }
    } else {
        const int f = s.infront[r] - 1, b = s.behind[r] - 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gi[k] = ld(s.x, s.cap, k, f);
            gb[k] = ld(s.x, s.cap, k, b);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (!p.periodic[k]) continue;
        double L = p.box[k], xi = ld(s.x, s.cap, k, r);
        if ((xi - gi[k]) > (L / 2.0)) gi[k] = gi[k] + L;
        else if ((xi - gi[k]) < (-L / 2.0)) gi[k] = gi[k] - L;
        if ((xi - gb[k]) > (L / 2.0)) gb[k] = gb[k] + L;
        else if ((xi - gb[k]) < (-L / 2.0)) gb[k] = gb[k] - L;
    }
}

// src/derivatives.f90:11-20 get_deriv_1 and :28-36 get_deriv_2
__device__ __forceinline__ void derivs(const double x[3], const double gi[3], const double gb[3], double fd[3],
                                       double fdd[3]) {
    double disti = dist_gen(x[0], x[1], x[2], gi[0], gi[1], gi[2]);
    double distb = dist_gen(x[0], x[1], x[2], gb[0], gb[1], gb[2]);
    double den = 2.0 * distb * disti;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = distb * gi[k] + (disti - distb) * x[k] - disti * gb[k];
        fd[k] = v / den;
        fdd[k] = 2.0 * (gi[k] / (disti * (disti + distb)) - x[k] / (disti * distb) +
                        gb[k] / (distb * (disti + distb)));
    }
}

// src/general.f90:152-159 cross_product
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// src/timestep.f90:151-198: one segment term of the direct sum, reference
// arithmetic.  d = x_j (+ shift) - x_i, e = ghosti_j - x_j, c = |e|^2.
// Returns false when the reference skips the pair (4ac - b^2 == 0).
__device__ __forceinline__ bool bs_term64(double d0, double d1, double d2, double e0, double e1, double e2, double c,
                                          double kappa, double out[3]) {
    double a = d0 * d0 + d1 * d1 + d2 * d2;
    double b = 2.0 * (d0 * e0 + d1 * e1 + d2 * e2);
    double D = 4.0 * a * c - b * b;
    if (D == 0.0) return false;
    double den = (2.0 * NV_PI) * D;
    double fac = (2.0 * c + b) / sqrt(a + b + c) - (b / sqrt(a));
    out[0] = ((d1 * e2 - d2 * e1) * kappa / den) * fac;
    out[1] = ((d2 * e0 - d0 * e2) * kappa / den) * fac;
    out[2] = ((d0 * e1 - d1 * e0) * kappa / den) * fac;
    return true;
}

// src/general.f90:46-51 cosh, which the channel model below calls (the
// module function hides the intrinsic): 0.5*(exp(x)+exp(-x))
__device__ __forceinline__ double ref_cosh(double x) { return 0.5 * (exp(x) + exp(-x)); }

// src/normal_fluid.f90:83-118 get_normal_velocity at position x and time t
// for the closed-form models; `noise` is the noisey_xflow table's current
// value.  Each expression keeps the Fortran line's evaluation order.
__device__ __forceinline__ void normal_velocity(const Phys &p, const double x[3], double t, double noise,
                                                double un[3]) {
    un[0] = un[1] = un[2] = 0.0;
    if (p.nf == 1) {
        un[0] = p.xflow;                                                       // :96-97
    } else if (p.nf == 2) {
        un[0] = noise * p.xflow;                                               // :98-99
    } else if (p.nf == 3) {
        un[0] = p.xflow * sin(2.0 * NV_PI * x[2] / p.box[2] + t * p.shear_omega);   // :101-102
    } else if (p.nf == 4) {
        const double q = 2.0 * x[1] / p.box[1];
        un[0] = p.xflow * (1.0 - q * q);                                       // :103-104
    } else if (p.nf == 5) {
        double cossum = 0.0;                                                   // :105-112
        for (int ii = 1; ii <= 10; ++ii) {
            const double ak = (double)(2 * ii - 1) * (NV_PI / 2.0);
            const double sgn = (ii & 1) ? -1.0 : 1.0;                          // (-1)**ii
            cossum = cossum + 4.0 * (sgn / (ak * ak * ak)) *
                                  (ref_cosh(ak * 2.0 * x[1] / p.box[1]) * cos(ak * 2.0 * x[2] / p.box[2])) /
                                  ref_cosh(ak);
        }
        const double q = 2.0 * x[2] / p.box[2];
        un[0] = p.xflow * (1.0 - q * q + cossum) * 1.2;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) un[k] = un[k] + p.sup[k];                      // :117
}

// AB coefficients, src/cdata.f90:67-71
#define NV_ONE_HALF (1.0 / 2.0)
#define NV_THREE_TWOS (3.0 / 2.0)
#define NV_23_12 (23.0 / 12.0)
#define NV_4_3 (4.0 / 3.0)
#define NV_5_12 (5.0 / 12.0)

// Velocity epilogue of one active point: local term (src/timestep.f90:60-80,
// 102-105), the non-local sums ub0 (unshifted), ubs (images) and ubm (solid
// walls' image vortices), superfluid
// velocity store (:123), mutual friction (:126-139, src/normal_fluid.f90:83-118),
// imposed flow (:140); `noise` is the noisey_xflow table's value for this
// call of pmotion; then, if do_ab, the Adams-Bashforth update of
// src/timestep.f90:28-43 into xn.
__device__ __forceinline__ void finish_point(const DState &s, const Phys &p, int r, const double ub0[3],
                                             const double ubs[3], const double ubm[3], double t, double noise,
                                             int do_ab) {
    const int cap = s.cap;
    double x[3], gi[3], gb[3], fd[3], fdd[3], c[3], u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = ld(s.x, cap, k, r);
        gi[k] = ld(s.gi, cap, k, r);
        gb[k] = ld(s.gb, cap, k, r);
    }
    derivs(x, gi, gb, fd, fdd);
    cross3(fd, fdd, c);
    if (p.velocity == 0) {
        double curv = sqrt(fdd[0] * fdd[0] + fdd[1] * fdd[1] + fdd[2] * fdd[2]);
        if (curv < NV_EPS) curv = NV_EPS;
        else curv = 1.0 / curv;
        double beta = (p.kappa / (4.0 * NV_PI)) * log(4.6 * curv / p.corea);
#pragma unroll
        for (int k = 0; k < 3; ++k) u[k] = beta * c[k];
    } else {
        double beta = (p.kappa / (4.0 * NV_PI)) *
                      log((1.1 / p.corea) * sqrt(dist_gen(x[0], x[1], x[2], gi[0], gi[1], gi[2]) *
                                                 dist_gen(x[0], x[1], x[2], gb[0], gb[1], gb[2])));
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            u[k] = beta * c[k];
            u[k] = u[k] + ub0[k];
            u[k] = u[k] + ubs[k];
            u[k] = u[k] + ubm[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) s.usup[k * cap + r] = u[k];
    if ((fabs(p.alpha[0]) > NV_EPS) || (fabs(p.alpha[1]) > NV_EPS)) {
        if (t < p.nf_cutoff) {
            double un[3] = {0.0, 0.0, 0.0};
            if (t < p.t_zero) normal_velocity(p, x, t, noise, un);
            double dv[3], c1[3], c2[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) dv[k] = un[k] - u[k];
            cross3(fd, dv, c1);
            cross3(fd, c1, c2);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double mf = p.alpha[0] * c1[k] - p.alpha[1] * c2[k];
                u[k] = u[k] + mf;
                s.umf[k * cap + r] = mf;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u[k] = u[k] + p.sup[k];
        s.u[k * cap + r] = u[k];
    }
    if (!do_ab) return;
    double u1[3], u2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        u1[k] = ld(s.u1, cap, k, r);
        u2[k] = ld(s.u2, cap, k, r);
    }
    double m1 = fmax(fmax(fabs(u1[0]), fabs(u1[1])), fabs(u1[2]));
    double m2 = fmax(fmax(fabs(u2[0]), fabs(u2[1])), fabs(u2[2]));
    const double dt = p.dt;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double xn;
        if (m1 == 0.0) xn = x[k] + dt * u[k];
        else if (m2 == 0.0) xn = x[k] + NV_THREE_TWOS * dt * u[k] - NV_ONE_HALF * dt * u1[k];
        else xn = x[k] + NV_23_12 * dt * u[k] - NV_4_3 * dt * u1[k] + NV_5_12 * dt * u2[k];
        s.xn[k * cap + r] = xn;
        s.u2[k * cap + r] = u1[k];
        s.u1[k * cap + r] = u[k];
    }
}
