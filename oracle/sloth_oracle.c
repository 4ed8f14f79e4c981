/*
 * sloth_oracle.c -- FP64 CPU restatement of SLOTH's ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against; it is never linked into, called by, or shipped
 * with the product library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * What it restates (reference = MAGPIE-ICL/SLOTH):
 *   orc_gradient_axis   src/core/propagator.py:39-93  precompute_gradients
 *                        (jnp.gradient with 1-D coordinate arrays: non-uniform
 *                        2nd-order interior, 1st-order one-sided edges)
 *   orc_kappa           src/core/propagator.py:116-164 kappa_inv_brems (NRL)
 *   orc_interp4         src/core/interpolator.py:15-106 RegularGridInterpolator
 *                        (searchsorted-1, clip to [0,n-2], fill 0 outside,
 *                        8 corners in itertools.product order)
 *   orc_trace           src/core/propagator.py:168-243 dsdt (RHS) integrated
 *                        by textbook fixed-step RK4 in real time, saving at the
 *                        schedule's segment ends (the save grid of
 *                        src/core/propagator.py:616-617 / :1113-1138).
 *
 * All arithmetic is IEEE binary64.  Inputs that the reference holds in fp32
 * (coordinates, ne, initial rays) are passed in already rounded to fp32 and
 * promoted, so the oracle sees exactly the numbers the reference sees.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_C 299792458.0 /* scipy.constants.c */

/* numpy.searchsorted(g, x, side='left'): number of g[i] < x. */
static int orc_searchsorted_left(const double* g, int n, double x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (g[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* ---------------------------------------------------------------------- */
/* precompute_gradients: gradient of f along one axis of an (nx,ny,nz)     */
/* C-ordered array, with the coordinate array of that axis.                */
/* Follows jnp/np.gradient(f, coord, axis) with edge_order=1:              */
/*   interior  a f[i-1] + b f[i] + c f[i+1]                                 */
/*       a = -dx2/(dx1(dx1+dx2)), b = (dx2-dx1)/(dx1 dx2), c = dx1/(dx2(dx1+dx2)) */
/*   edges     (f[1]-f[0])/(x1-x0),  (f[n-1]-f[n-2])/(x[n-1]-x[n-2])       */
/* ---------------------------------------------------------------------- */
void orc_gradient_axis(const double* f, const double* coord, int nx, int ny, int nz,
                       int axis, double* out) {
    const int n[3] = {nx, ny, nz};
    const int64_t stride[3] = {(int64_t)ny * nz, nz, 1};
    const int na = n[axis];
    const int64_t sa = stride[axis];
    const int64_t total = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (int64_t lin = 0; lin < total; ++lin) {
        int64_t rem = lin;
        int idx[3];
        idx[0] = (int)(rem / stride[0]); rem -= (int64_t)idx[0] * stride[0];
        idx[1] = (int)(rem / stride[1]); rem -= (int64_t)idx[1] * stride[1];
        idx[2] = (int)rem;
        const int i = idx[axis];
        double g;
        if (na < 2) {
            g = 0.0;
        } else if (i == 0) {
            g = (f[lin + sa] - f[lin]) / (coord[1] - coord[0]);
        } else if (i == na - 1) {
            g = (f[lin] - f[lin - sa]) / (coord[na - 1] - coord[na - 2]);
        } else {
            const double dx1 = coord[i] - coord[i - 1];
            const double dx2 = coord[i + 1] - coord[i];
            const double a = -dx2 / (dx1 * (dx1 + dx2));
            const double b = (dx2 - dx1) / (dx1 * dx2);
            const double c = dx1 / (dx2 * (dx1 + dx2));
            g = a * f[lin - sa] + b * f[lin] + c * f[lin + sa];
        }
        out[lin] = g;
    }
}

/* ---------------------------------------------------------------------- */
/* kappa_inv_brems (NRL p.58, classical b_min), per element.               */
/* Te and Z are either scalars (count 1) or full grids (count n).          */
/* cl_override > 0 replaces the Coulomb logarithm (extension; the          */
/* reference always computes max(2, ln(v_the/(w_max b_min)))).             */
/* ---------------------------------------------------------------------- */
void orc_kappa(const double* ne, int64_t n, const double* Te, int64_t n_te,
               const double* Z, int64_t n_z, double omega, double e_charge,
               double epsilon_0, double cl_override, double* out, double* cl_out) {
    const double pi = 3.14159265358979311600; /* numpy.pi */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double te = Te[n_te == 1 ? 0 : i];
        const double z = Z[n_z == 1 ? 0 : i];
        const double ne_cc = ne[i] * 1e-6;
        const double v_the = 4.19e5 * sqrt(te);
        const double o_pe = 5.64e4 * sqrt(ne_cc);
        const double o_max = o_pe > omega ? o_pe : omega;
        const double b_min = z * e_charge / (4.0 * pi * epsilon_0 * te);
        double cl = log(v_the / (o_max * b_min));
        if (!(cl >= 2.0)) cl = 2.0;
        if (cl_override > 0.0) cl = cl_override;
        const double r = ne_cc / omega;
        out[i] = 3.1e-5 * z * ORC_C * (r * r) * cl * pow(te, -1.5);
        if (cl_out) cl_out[i] = cl;
    }
}

/* ---------------------------------------------------------------------- */
/* Trilinear interpolation of up to 4 channels sharing one grid.           */
/* Returns 1 when the point is inside the closed box, 0 (all outputs 0)    */
/* otherwise -- the fill_value=0 rule of interpolator.py:86-87,102-104.    */
/* ---------------------------------------------------------------------- */
typedef struct {
    int n[3];
    const double* coord[3];
    const double* ch[4]; /* dndx, dndy, dndz, kappa (kappa may be NULL) */
} orc_grid;

static int orc_interp4(const orc_grid* G, const double p[3], double out[4]) {
    int idx[3];
    double t[3];
    int oob = 0;
    for (int a = 0; a < 3; ++a) {
        const double* g = G->coord[a];
        const int n = G->n[a];
        int i = orc_searchsorted_left(g, n, p[a]) - 1;
        if (i < 0) i = 0;
        if (i > n - 2) i = n - 2;
        idx[a] = i;
        t[a] = (p[a] - g[i]) / (g[i + 1] - g[i]);
        if (p[a] < g[0]) oob = 1;
        if (p[a] > g[n - 1]) oob = 1;
    }
    for (int c = 0; c < 4; ++c) out[c] = 0.0;
    if (oob) return 0;
    const int64_t sx = (int64_t)G->n[1] * G->n[2], sy = G->n[2];
    /* itertools.product([i,i+1],[j,j+1],[k,k+1]) order */
    for (int e = 0; e < 8; ++e) {
        const int ex = (e >> 2) & 1, ey = (e >> 1) & 1, ez = e & 1;
        double w = 1.0;
        w *= ex ? t[0] : 1.0 - t[0];
        w *= ey ? t[1] : 1.0 - t[1];
        w *= ez ? t[2] : 1.0 - t[2];
        const int64_t lin = (int64_t)(idx[0] + ex) * sx + (int64_t)(idx[1] + ey) * sy + (idx[2] + ez);
        for (int c = 0; c < 4; ++c)
            if (G->ch[c]) out[c] += G->ch[c][lin] * w;
    }
    return 1;
}

/* Exported single-point interpolation (for pinning against SciPy). */
void orc_interp_points(int nx, int ny, int nz, const double* cx, const double* cy,
                       const double* cz, const double* values, int64_t npts,
                       const double* pts /* npts x 3 */, double* out) {
    orc_grid G;
    G.n[0] = nx; G.n[1] = ny; G.n[2] = nz;
    G.coord[0] = cx; G.coord[1] = cy; G.coord[2] = cz;
    G.ch[0] = values; G.ch[1] = G.ch[2] = G.ch[3] = NULL;
    for (int64_t i = 0; i < npts; ++i) {
        double o[4];
        orc_interp4(&G, pts + 3 * i, o);
        out[i] = o[0];
    }
}

/* Does the straight segment p + s v, s in [0, h], touch the closed box?  */
static int orc_segment_touches_box(const orc_grid* G, const double p[3], const double v[3], double h) {
    double s0 = 0.0, s1 = h;
    for (int a = 0; a < 3; ++a) {
        const double lo = G->coord[a][0], hi = G->coord[a][G->n[a] - 1];
        if (v[a] == 0.0) {
            if (p[a] < lo || p[a] > hi) return 0;
        } else {
            double ta = (lo - p[a]) / v[a], tb = (hi - p[a]) / v[a];
            if (ta > tb) { double tmp = ta; ta = tb; tb = tmp; }
            if (ta > s0) s0 = ta;
            if (tb < s1) s1 = tb;
        }
    }
    /* generous tolerance: a false "touches" only costs an exact RK4 step */
    return s0 <= s1 * (1.0 + 1e-9) + 1e-30;
}

/* Distance from p to the surface of the closed box (inside: to the nearest face;  */
/* outside: Euclidean distance to the box).  A stage point closer than the gap    */
/* between two trajectories can be inside for one and outside for the other.     */
static double orc_face_gap(const orc_grid* G, const double p[3]) {
    double out2 = 0.0, in = INFINITY;
    int outside = 0;
    for (int a = 0; a < 3; ++a) {
        const double lo = G->coord[a][0], hi = G->coord[a][G->n[a] - 1];
        double e = 0.0;
        if (p[a] < lo) e = lo - p[a];
        if (p[a] > hi) e = p[a] - hi;
        if (e > 0.0) outside = 1;
        out2 += e * e;
        const double d = fmin(p[a] - lo, hi - p[a]);
        if (d < in) in = d;
    }
    return outside ? sqrt(out2) : in;
}

static void orc_rhs(const orc_grid* G, const double r[3], double a[3], double* kap, int* inside) {
    double o[4];
    int in = orc_interp4(G, r, o);
    a[0] = o[0]; a[1] = o[1]; a[2] = o[2];
    *kap = o[3];
    *inside |= in;
}

/* ---------------------------------------------------------------------- */
/* Fixed-step RK4 over a save schedule.                                    */
/*   s0      6 x n (row c = x,y,z,vx,vy,vz; column = ray)                  */
/*   nseg    number of save intervals; segment s has nsub[s] steps of h[s] */
/*   saves   (nseg+1) x 8 x n : rows x,y,z,vx,vy,vz,tau,path at each save  */
/*           (save 0 is the initial state)                                 */
/*   steps   in-box ray-steps per ray (a step counts when any of its four  */
/*           stage points lies inside the closed grid box)                 */
/*   naive   1: every step is a full RK4 step; 0: steps whose straight     */
/*           segment cannot touch the box are taken as exact linear moves  */
/*   face_gap  per ray (or NULL): the smallest orc_face_gap of any stage   */
/*           point of its RK4 steps, except the start point itself (which  */
/*           both a checked trace and this one take unrounded)             */
/* ---------------------------------------------------------------------- */
void orc_trace(int nx, int ny, int nz, const double* cx, const double* cy, const double* cz,
               const double* dndx, const double* dndy, const double* dndz, const double* kappa,
               int64_t n, const double* s0, int nseg, const int32_t* nsub, const double* h,
               int naive, double* saves, int64_t* steps, double* face_gap, int nthreads) {
    orc_grid G;
    G.n[0] = nx; G.n[1] = ny; G.n[2] = nz;
    G.coord[0] = cx; G.coord[1] = cy; G.coord[2] = cz;
    G.ch[0] = dndx; G.ch[1] = dndy; G.ch[2] = dndz; G.ch[3] = kappa;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int64_t plane = 8 * n;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ray = 0; ray < n; ++ray) {
        double r[3], v[3], tau = 0.0, path = 0.0, gap = INFINITY;
        int64_t nstep = 0;
        int moved = 0;
        for (int c = 0; c < 3; ++c) { r[c] = s0[c * n + ray]; v[c] = s0[(3 + c) * n + ray]; }
#define ORC_SAVE(sidx)                                                      \
        do {                                                                \
            double* base = saves + (int64_t)(sidx) * plane;                 \
            for (int c = 0; c < 3; ++c) { base[c * n + ray] = r[c]; base[(3 + c) * n + ray] = v[c]; } \
            base[6 * n + ray] = tau; base[7 * n + ray] = path;              \
        } while (0)
        ORC_SAVE(0);
        for (int s = 0; s < nseg; ++s) {
            const double hh = h[s];
            for (int k = 0; k < nsub[s]; ++k) {
                if (!naive && !orc_segment_touches_box(&G, r, v, hh)) {
                    const double sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    for (int c = 0; c < 3; ++c) r[c] += hh * v[c];
                    path += hh * sp;
                    moved = 1;
                    continue;
                }
                int inside = 0;
                double a1[3], a2[3], a3[3], a4[3], k1, k2, k3, k4;
                double r2[3], r3[3], r4[3], v2[3], v3[3], v4[3];
                orc_rhs(&G, r, a1, &k1, &inside);
                for (int c = 0; c < 3; ++c) { r2[c] = r[c] + 0.5 * hh * v[c]; v2[c] = v[c] + 0.5 * hh * a1[c]; }
                orc_rhs(&G, r2, a2, &k2, &inside);
                for (int c = 0; c < 3; ++c) { r3[c] = r[c] + 0.5 * hh * v2[c]; v3[c] = v[c] + 0.5 * hh * a2[c]; }
                orc_rhs(&G, r3, a3, &k3, &inside);
                for (int c = 0; c < 3; ++c) { r4[c] = r[c] + hh * v3[c]; v4[c] = v[c] + hh * a3[c]; }
                orc_rhs(&G, r4, a4, &k4, &inside);
                if (face_gap) {
                    if (moved) gap = fmin(gap, orc_face_gap(&G, r));
                    gap = fmin(gap, orc_face_gap(&G, r2));
                    gap = fmin(gap, orc_face_gap(&G, r3));
                    gap = fmin(gap, orc_face_gap(&G, r4));
                }
                moved = 1;
                double dr2 = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double dr = hh / 6.0 * (v[c] + 2.0 * v2[c] + 2.0 * v3[c] + v4[c]);
                    r[c] += dr;
                    dr2 += dr * dr;
                    v[c] += hh / 6.0 * (a1[c] + 2.0 * a2[c] + 2.0 * a3[c] + a4[c]);
                }
                tau += hh / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
                path += sqrt(dr2);
                nstep += inside;
            }
            ORC_SAVE(s + 1);
        }
#undef ORC_SAVE
        if (steps) steps[ray] = nstep;
        if (face_gap) face_gap[ray] = gap;
    }
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
