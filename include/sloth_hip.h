/*
 * sloth_hip.h -- C ABI of libsloth_hip.so, the MI355X (gfx950) ray tracer that
 * replaces the hot path of MAGPIE-ICL/SLOTH.
 *
 * Conventions
 *   - Every device pointer is caller-owned (e.g. a PyTorch tensor's data_ptr()).
 *     The library never frees caller memory and allocates none on these paths.
 *   - Every call is ordered on the given HIP stream (NULL = default stream) and
 *     returns asynchronously; 0 = success, < 0 = error (message in
 *     slt_last_error(), thread-local).  No C++ exception crosses this ABI.
 *   - Calls on distinct streams with distinct output buffers may run
 *     concurrently.
 *
 * The reference is pure Python/JAX: there is no FFI in it.  Each entry point
 * below names the reference function whose work it takes over; the Python
 * mirror in sloth_amd/ binds them with ctypes (sloth_amd/_native.py).
 */
#ifndef SLOTH_HIP_H
#define SLOTH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLT_ABI_VERSION 3

#define SLT_OK 0
#define SLT_ERR_INVALID (-1)     /* bad argument (shape, pointer, range) */
#define SLT_ERR_HIP (-2)         /* HIP runtime error (launch, memory)     */
#define SLT_ERR_UNSUPPORTED (-3) /* valid input outside what the kernels handle */

typedef void* slt_stream_t; /* hipStream_t */

/*
 * Layout of a packed grid: one 16-byte voxel (d/dx, d/dy, d/dz of the
 * refraction driving term, kappa) per grid point, addressed as
 *   voxel(i, j, k) = base + i*stride[0] + j*stride[1] + k*stride[2]
 * where (i, j, k) index the x, y, z coordinate arrays.  The probing axis is
 * laid out fastest so that consecutive steps of a ray reuse cache lines.
 */
typedef struct slt_grid_desc {
    int32_t n[3];      /* points per axis (x, y, z), each >= 2           */
    int32_t nchan;     /* 3: gradient only; 4: gradient + kappa          */
    int64_t stride[3]; /* voxel strides (units of 16-byte voxels)        */
    float lo[3];       /* first coordinate value x[0], y[0], z[0] (fp32) */
    float hi[3];       /* last coordinate value x[n-1], ...      (fp32)  */
    float inv_d[3];    /* (n-1) / (hi - lo)                              */
    int32_t probing_axis; /* 0, 1, 2 for x, y, z                         */
    int32_t uniform;   /* 1: every axis a linspace up to fp32 rounding   */
    int32_t reserved;
    const float* knots; /* uniform = 0: device array of the n[0] + n[1] + n[2]
                           grid points in grid units, (c[i] - lo) * inv_d
                           per axis (x, then y, then z), from slt_grid_knots;
                           NULL for a uniform grid                          */
} slt_grid_desc;

/* Library ABI version (SLT_ABI_VERSION). */
int slt_version(void);

/* Last error message of the calling thread ("" if none). */
const char* slt_last_error(void);

/*
 * Host-only: fill *out from the fp32 coordinate arrays (host memory).  Each
 * axis must be finite and strictly ascending (the reference interpolator's
 * requirement), else SLT_ERR_INVALID.  A grid whose axes are all uniformly
 * spaced up to fp32 rounding (any linspace is) gets uniform = 1 and index
 * arithmetic in the kernels.  Any other grid gets uniform = 0: the caller then
 * computes its knots with slt_grid_knots, copies them to the device and sets
 * out->knots before slt_trace, which locates cells by a per-axis binary search
 * over them; so do the adjoint entry points.  The serving entry points take
 * uniform grids only.
 * Replaces the coordinate handling of src/core/interpolator.py:72-87 (per-axis
 * searchsorted).
 */
int slt_grid_desc_init(const float* x, const float* y, const float* z, int32_t nx, int32_t ny,
                       int32_t nz, int32_t probing_axis, int32_t with_kappa,
                       slt_grid_desc* out);

/*
 * Host-only: the grid points of every axis in grid units, knots[off_a + i] =
 * (float)(((double)c_a[i] - lo_a) * (n_a - 1) / (hi_a - lo_a)) with off = 0,
 * n[0], n[0] + n[1] for x, y, z (n[0] + n[1] + n[2] floats, host memory).
 */
int slt_grid_knots(const float* x, const float* y, const float* z, const slt_grid_desc* g,
                   float* knots);

/* Bytes of the packed voxel buffer described by *g. */
int64_t slt_grid_bytes(const slt_grid_desc* g);

/*
 * k1: gradient stencil + inverse-bremsstrahlung rate + 16-byte voxel pack.
 * Replaces precompute_gradients (src/core/propagator.py:39-93) and
 * kappa_inv_brems (src/core/propagator.py:116-164).
 *   ne            device fp32 (nx, ny, nz) C-order, electron density [m^-3]
 *   cx, cy, cz    device fp32 coordinate arrays (the gradient uses the actual
 *                 coordinates, as numpy/jnp.gradient with coordinate arrays does)
 *   grad_coeff    -0.5 c^2 / (3.14207787e-4 omega^2), computed in fp64
 *   te, n_te      device fp32 electron temperature [eV], n_te = 1 or nx*ny*nz
 *   zz, n_z       device fp32 mean charge state,      n_z  = 1 or nx*ny*nz
 *                 (te / zz may be NULL when kappa is not wanted)
 *   omega, e_charge, epsilon_0, speed_of_light   physical constants (fp64)
 *   coulomb_log   > 0 overrides the Coulomb logarithm (extension); <= 0 uses
 *                 the reference's max(2, ln(v_the / (omega_max b_min)))
 *   packed        device voxel buffer (slt_grid_bytes(g) bytes) or NULL; voxel
 *                 = (d/dx * dx, d/dy * dy, d/dz * dz, kappa) with d_a the cell
 *                 size (hi - lo) / (n - 1) of g: the gradient in grid units, the
 *                 form the integrators step in
 *   planar[4]     optional device fp32 (nx,ny,nz) C-order outputs for
 *                 d/dx, d/dy, d/dz (physical units) and kappa (each may be NULL)
 * Arithmetic is fp64; results are rounded once to fp32.
 */
int slt_grid_pack(const float* ne, const float* cx, const float* cy, const float* cz,
                  double grad_coeff, const float* te, int64_t n_te, const float* zz,
                  int64_t n_z, double omega, double e_charge, double epsilon_0,
                  double speed_of_light, double coulomb_log, const slt_grid_desc* g,
                  void* packed, float* planar_dx, float* planar_dy, float* planar_dz,
                  float* planar_kappa, slt_stream_t stream);

/*
 * Element-wise inverse-bremsstrahlung rate [1/s] of n densities (fp64 math,
 * fp32 out).  Replaces kappa_inv_brems (src/core/propagator.py:116-164) when
 * called on its own; Te/Z have 1 or n elements; coulomb_log as in slt_grid_pack.
 */
int slt_kappa(const float* ne, int64_t n, const float* te, int64_t n_te, const float* zz,
              int64_t n_z, double omega, double e_charge, double epsilon_0,
              double speed_of_light, double coulomb_log, float* out, slt_stream_t stream);

/*
 * Save schedule: the integration runs from t = 0 through n_seg segments;
 * segment s has seg_nsub[s] RK4 steps of length seg_h[s] seconds and ends at
 * save point s+1 (save point 0 is the initial state).
 */
typedef struct slt_trace_args {
    int64_t n_rays;
    const float* s0;        /* device 6 x n_rays (rows x,y,z,vx,vy,vz), row stride n_rays */
    const int32_t* order;   /* device processing order (a permutation of 0..n-1) or NULL  */
    int32_t n_seg;          /* number of save intervals, >= 1                              */
    int32_t save_every;     /* write save points to save_state/save_tau: 1 = all, 0 = last only */
    const int32_t* seg_nsub;/* device [n_seg]                                              */
    const float* seg_h;     /* device [n_seg], seconds                                     */
    float* save_state;      /* device (n_seg+1) x 6 x n_rays (or 1 x 6 x n_rays if save_every == 0) */
    double* save_tau;       /* device, same leading shape as save_state, 1 row; or NULL    */
    double* path;           /* device [n_rays] path length [m] (polyline through the step end points) or NULL */
    uint32_t* steps;        /* device [n_rays] in-box RK4 steps per ray or NULL            */
    unsigned long long* total_steps; /* device [1], incremented by all in-box steps, or NULL */
} slt_trace_args;

/*
 * k2: fixed-step RK4 ray integrator (one ray per lane).  Replaces the vmapped
 * diffrax solve of src/core/propagator.py:726-793 and :1162-1198 with the RHS
 * dsdt (:168-243) and trilinear interpolation (src/core/interpolator.py:15-106):
 *   dr/dt = v,  dv/dt = interp(grad(coeff ne))(r),  dtau/dt = interp(kappa)(r)
 * with the interpolant 0 outside the closed grid box.  Position is an fp64 base
 * plus an fp32 offset, velocity the start velocity plus an fp32 change; tau and
 * path length accumulate in fp64.
 */
int slt_trace(const void* packed, const slt_grid_desc* g, const slt_trace_args* a,
              slt_stream_t stream);

/*
 * Jones-vector epilogue: state (6 x n, row stride ld) -> rf (4 x n, row stride n).
 * Replaces ray_to_Jonesvector(..., keep_current_plane=True)
 * (src/shared/propagation.py:4-90): rows [pos1, atan(v1/vp), pos2, atan(v2/vp)].
 * If amp != NULL and tau != NULL, also writes amp[i] = exp(-tau[i])
 * (src/core/propagator.py:324-326).
 */
int slt_jones(const float* state, int64_t ld, int64_t n, int32_t probing_axis, float* rf,
              const double* tau, double* amp, slt_stream_t stream);

/*
 * k3: detector histogram.  Replaces numpy.histogram2d inside
 * Diagnostic.histogram (src/processing/diagnostics.py:243-276) including its
 * H.T transpose: counts/wsum are (nyb, nxb) row-major.  Bin b holds
 * xedges[b] <= x < xedges[b+1]; x == xedges[nxb] goes to the last bin;
 * anything else outside, and NaN, is dropped (numpy's rule).  Values are
 * compared in fp64 against the fp64 edge arrays (device, nxb+1 / nyb+1 entries).
 * x, y are device arrays of fp32 (is_f64 = 0) or fp64 (is_f64 = 1) with the
 * given element strides (>= 0; 0 repeats one element) -- e.g. rows 0 and 2 of rf.
 * counts (uint32) and/or wsum (fp64, needs weights) are accumulated into, so
 * the caller zeroes them first.  mask: optional device uint8 per ray (0 = skip).
 * order: optional device int32 permutation of [0, n) (the Morton order the
 * rays were traced in); rays are then visited as order[0..n), else in index
 * order.  A wavefront combines its entries for each of its first few distinct
 * cells into one atomic, so rays visited in a spatially coherent order (Morton
 * order, or results already in that order) make few atomics on one address.
 * Counts do not depend on the visiting order; wsum only through fp64 summation
 * order.  Order entries outside [0, n) are skipped.  nxb * nyb must fit in int32.
 */
int slt_histogram2d(const void* x, int64_t x_stride, const void* y, int64_t y_stride,
                    int32_t is_f64, int64_t n, const int32_t* order, const uint8_t* mask,
                    const double* xedges, int32_t nxb, const double* yedges, int32_t nyb,
                    const double* weights, uint32_t* counts, double* wsum,
                    slt_stream_t stream);

/*
 * k5: ray-transfer-matrix optics between the plasma and the detector.
 * Replaces the element functions lens / sym_lens / travel /
 * circular_aperture / circular_stop / rect_aperture / knife_edge
 * (src/processing/diagnostics.py:28-128) as chained by the diagnostics'
 * solve methods (:284-406), with m_to_mm (:16-20) in front.
 *
 * rays: 4 x n (x [m], theta, y [m], phi), fp32 (rays_f64 = 0) or fp64, row
 * stride ld.  Rows 0 and 2 are multiplied by pos_scale in the input's
 * precision (1e3 for the reference's m -> mm), then the n_ops elements are
 * applied in order and the result is written as fp64 to out (4 x n, row
 * stride out_ld; may not alias rays).  A linear element is the product of
 * its 4x4 matrix with the ray, each component accumulated over the four
 * inputs with fused multiply-adds (NaN rays stay NaN, as in the reference's
 * matmul); a rejected ray becomes NaN in all four rows (r[:, filt] = None).
 * Squares are taken as R*R.  ops is a host array (copied at launch).
 */
#define SLT_MAX_OPTICS 32
enum {
    SLT_OPTIC_TRAVEL = 0,            /* a = d: [[1, d], [0, 1]] on both axes  (:52-65)    */
    SLT_OPTIC_LENS = 1,              /* a = f1 (x), b = f2 (y): [[1, 0], [-1/f, 1]] (:28-50) */
    SLT_OPTIC_CIRCULAR_APERTURE = 2, /* a = R: reject x^2 + y^2 > R^2        (:67-75)    */
    SLT_OPTIC_CIRCULAR_STOP = 3,     /* a = R: reject x^2 + y^2 < R^2        (:77-85)    */
    SLT_OPTIC_RECT_APERTURE = 4,     /* a = Lx, b = Ly: reject x^2 > Lx^2 and y^2 > Ly^2 (:98-109) */
    SLT_OPTIC_KNIFE_EDGE = 5         /* a = offset, b = direction (> 0: reject r[axis] > offset,
                                        < 0: reject r[axis] < offset), axis 0 (x) or 2 (y) (:111-128) */
};

typedef struct slt_optic {
    int32_t kind;
    int32_t axis;
    double a;
    double b;
} slt_optic;

int slt_rtm(const void* rays, int32_t rays_f64, int64_t ld, int64_t n, double pos_scale,
            const slt_optic* ops, int32_t n_ops, double* out, int64_t out_ld, slt_stream_t stream);

/*
 * Morton (Z-order) keys of the rays' start positions in the plane transverse
 * to the probing axis, quantised to 15 bits per axis over [lo, hi] of the grid
 * (clamped) and interleaved into a non-negative 30-bit int32 (4 radix passes
 * to sort).  Sorting rays by these keys keeps each wavefront's corner gathers
 * on a few cache lines.  No reference counterpart (the reference vmaps rays in
 * input order, src/core/propagator.py:792).
 */
int slt_morton_keys(const float* s0, int64_t n, const slt_grid_desc* g, int32_t* keys,
                    slt_stream_t stream);

/*
 * Ray permutation between the caller's order and a processing order (order[t] =
 * the caller's index of the ray processed t-th).  No reference counterpart (ray
 * ordering is this library's; the reference vmaps rays in input order,
 * src/core/propagator.py:792).  Both calls go through `work`, a caller-owned device
 * buffer of slt_ray_work_bytes(n) bytes (32 B per ray) that must not overlap the
 * data: a scattered 4-byte access moves a whole cache line, so each ray is moved as
 * one 32-byte record, with a coalesced transpose between rows and records.
 *
 * slt_gather_rays: s0_sorted[c][t] = s0[c][order[t]] for the six rows of a 6 x n
 * fp32 start array (row stride n); s0_sorted may be s0 itself.  An order entry
 * outside [0, n) gives a NaN row.  Tracing the gathered rays with order = NULL
 * reads and writes the state coalesced.
 *
 * slt_unpermute_rays: the inverse, for results in processing order:
 * dst[c][order[t]] = src[c][t] for `rows` fp32 rows (row stride n) and
 * dst64[order[t]] = src64[t] for an optional fp64 row; rows + 2 (with the fp64 row)
 * <= 8; dst may be src.  order must be a permutation of [0, n) (entries outside it
 * are skipped, leaving their columns undefined).
 */
int64_t slt_ray_work_bytes(int64_t n);
int slt_gather_rays(const float* s0, int64_t n, const int32_t* order, float* s0_sorted,
                    void* work, slt_stream_t stream);
int slt_unpermute_rays(const float* src, int32_t rows, const double* src64, int64_t n,
                       const int32_t* order, float* dst, double* dst64, void* work,
                       slt_stream_t stream);

/*
 * k6: probe beam made on the device (opt-in; no reference counterpart: the
 * reference draws rays on the host from numpy's RandomState,
 * src/core/beam.py:57-346, which the drop-in Beam keeps).  Ray i takes its
 * random numbers from Philox4x32-10 with counter (i, word) and key `seed`, so
 * each ray depends only on (seed, i).  Distributions as the reference's beam
 * types: circular (radius size_a, uniform on the disc), square (half-width
 * size_a), rectangular (half-widths size_a, size_b), linear (half-length size_a,
 * x-z plane); divergence is the standard deviation of the angle to the probing
 * axis; rays start at -ne_extent on the probing axis.  s0: 6 x n fp32 device.
 */
enum { SLT_BEAM_CIRCULAR = 0, SLT_BEAM_SQUARE = 1, SLT_BEAM_RECTANGULAR = 2, SLT_BEAM_LINEAR = 3 };

int slt_beam(int64_t n, int32_t beam_type, int32_t probing_axis, double size_a, double size_b,
             double divergence, double ne_extent, uint64_t seed, float* s0, slt_stream_t stream);

/* ------------------------------------------------------------------------- */
/* Serving: resumable ray state and multi-request step batches.              */
/* No reference counterpart (the reference traces one request per call).     */
/* ------------------------------------------------------------------------- */

#define SLT_MAX_JOBS 64

/* Bytes of resumable state for n rays (84 B per ray). */
int64_t slt_ray_state_bytes(int64_t n_rays);

/*
 * Initialise resumable state: row t takes ray order[t] of s0 (6 x n, row stride
 * n), or ray t if order is NULL.  The state is kept in the grid units of g (the
 * descriptor of the job's packed grid), so the same g goes into the job.
 */
int slt_ray_state_init(void* state, int64_t n_rays, const float* s0, const int32_t* order,
                       const slt_grid_desc* g, slt_stream_t stream);

/* One request's rays and the step range a step batch advances them through. */
typedef struct slt_job {
    const void* packed;        /* packed grid (slt_grid_pack)                        */
    slt_grid_desc grid;
    void* state;               /* slt_ray_state_bytes(n_rays) bytes (slt_ray_state_init) */
    const int32_t* order;      /* caller's index of state row t, or NULL (outputs
                                  then in state-row order)                         */
    int64_t n_rays;
    const int32_t* seg_nsub;   /* save schedule, as in slt_trace_args                */
    const float* seg_h;
    int32_t n_seg;
    int32_t reserved;
    int64_t sched_steps;       /* sum of seg_nsub                                    */
    int64_t step_begin;        /* advance every ray from this global step ...        */
    int64_t step_end;          /* ... to this one; == sched_steps finishes the job   */
    float* out_state;          /* 6 x n_rays final state, column order[t] for row t
                                  (row t without an order) (finishing)            */
    double* out_tau;           /* n_rays final tau, or NULL                          */
    double* out_path;          /* n_rays final path length, or NULL                  */
    uint32_t* out_steps;       /* n_rays in-box steps, or NULL                       */
    unsigned long long* step_counter; /* += in-box steps taken by this batch, or NULL */
} slt_job;

/* Unit queues of a step batch: one per XCD (a workgroup's XCD is its index mod 8). */
#define SLT_WORK_QUEUES 8

typedef struct slt_job_table {
    /* Work queues of slt_advance, one 128-byte line each so the XCDs do not share a
     * line: next_unit[q][0] is the next unit queue q hands out, waves_done[0] the
     * wavefronts that have drained every queue.  Zero when a launch starts
     * (slt_job_table_prepare zeroes them; the launch's last wavefront resets them). */
    uint32_t next_unit[SLT_WORK_QUEUES][32];
    uint32_t waves_done[32];
    int32_t n_jobs;
    int32_t n_units;                      /* filled by slt_job_table_prepare */
    int32_t unit_begin[SLT_MAX_JOBS + 1]; /* filled by slt_job_table_prepare */
    int32_t reserved;
    slt_job jobs[SLT_MAX_JOBS];
} slt_job_table;

/*
 * Host-only: validate every job (grid, pointers, step ranges), fill n_units /
 * unit_begin (one unit = 64 rays, one wavefront) and zero the work queues.
 * Call before copying the table to the device.
 */
int slt_job_table_prepare(slt_job_table* host_table);

/*
 * One step batch: advance every job of a device-resident, prepared table by
 * its [step_begin, step_end) range, in one persistent launch of `blocks`
 * workgroups (a positive multiple of 8).  Each wavefront takes units from its
 * XCD's queue (queue q holds the q-th eighth of the units, in ray order) and,
 * when that is empty, from the other queues, so the batch ends when the work
 * does, whatever the mix of jobs.  The kernel arguments do not depend on the
 * table's contents, so a launch can be captured in a HIP graph once and
 * replayed after each table update.  A request advanced in several batches
 * gives the same result as one slt_trace call (up to fp64 rounding of the
 * straight-line moves outside the grid, which are split at batch ends).
 * Launches on one table must not overlap.
 */
int slt_advance(slt_job_table* device_table, int32_t blocks, slt_stream_t stream);

/* ------------------------------------------------------------------------- */
/* Adjoint: d(loss on final ray states)/d(n_e).  No reference counterpart     */
/* (the reference has no gradient path; jax.grad through diffrax would be).  */
/* ------------------------------------------------------------------------- */

typedef struct slt_adjoint_args {
    int64_t n_rays;
    const float* s0;              /* 6 x n_rays start states (caller order)              */
    const int32_t* order;         /* processing order (Morton) or NULL                    */
    int32_t n_seg;                /* save schedule, as in slt_trace_args                  */
    int32_t ckpt_every;           /* K: checkpoint every K schedule steps (e.g. 16); 0:
                                     no checkpoints, the forward records the trajectory */
    const int32_t* seg_nsub;
    const float* seg_h;
    int64_t sched_steps;          /* sum of seg_nsub                                      */
    void* checkpoints;            /* slt_checkpoint_bytes(n_rays, sched_steps, K) bytes   */
    float* final_state;           /* forward out: 6 x n_rays final states (caller order)  */
    double* final_tau;            /* forward out: n_rays final tau, or NULL               */
    const float* seed_state;      /* backward in: d loss / d final state, 6 x n_rays      */
    const double* seed_tau;       /* backward in: d loss / d final tau, or NULL           */
    float* grid_adj;              /* backward out (accumulated): d loss / d voxel, packed layout, 4 floats per voxel */
    void* tape;                   /* backward scratch: slt_tape_bytes(n_rays, K) bytes;
                                     K = 0: the trajectory tape, slt_trajectory_tape_bytes
                                     (n_rays, sched_steps) bytes, written by slt_forward_ckpt
                                     and read by slt_adjoint (keep it between the two)   */
    unsigned long long* total_steps; /* forward: += in-box steps, or NULL                 */
} slt_adjoint_args;

int64_t slt_checkpoint_bytes(int64_t n_rays, int64_t sched_steps, int32_t ckpt_every);
int64_t slt_tape_bytes(int64_t n_rays, int32_t ckpt_every);
/*
 * Bytes of the whole-trajectory tape of ckpt_every = 0: one 32-byte record per
 * ray and schedule step at most (sched_steps <= INT32_MAX; 0 when the size does
 * not fit int64).  It trades memory for the reverse pass's replay: 2e6 rays over
 * 1443 steps take 92 GB, with no checkpoint replayed.  The checkpoint area is
 * then slt_checkpoint_bytes(n_rays, sched_steps, 0) = 4 * n_rays bytes (each
 * ray's record count).  No counterpart in the reference.
 */
int64_t slt_trajectory_tape_bytes(int64_t n_rays, int64_t sched_steps);

/* Forward trace (as slt_trace) that also stores every ray's state each K steps
 * (K = 0: records every move on a->tape instead). */
int slt_forward_ckpt(const void* packed, const slt_grid_desc* g, const slt_adjoint_args* a,
                     slt_stream_t stream);

/*
 * Reverse pass: from the seeds (adjoints of the final states and tau), replay
 * each K-step interval from its checkpoint and run the transposed RK4 steps,
 * accumulating d loss / d voxel into grid_adj (zero it first; float atomics).
 * K = 0: no replay; the transposed steps walk the forward's trajectory tape.
 */
int slt_adjoint(const void* packed, const slt_grid_desc* g, const slt_adjoint_args* a,
                slt_stream_t stream);

/*
 * slt_adjoint without the tape argument, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The reverse pass's tape lives in a library-owned device array
 * of SLT_ADJ_SCRATCH_BYTES, one per GPU, held from its first use until the
 * process ends; a->tape must be NULL.  Limit: slt_tape_bytes(n_rays,
 * ckpt_every) <= SLT_ADJ_SCRATCH_BYTES, i.e. n_rays * ckpt_every <= 4194304
 * (e.g. 262144 rays at K = 16); larger calls, and K = 0 (whose tape the
 * caller holds), return SLT_ERR_UNSUPPORTED before any device work.  The checkpoints, seeds and grid_adj stay
 * caller-owned, as in slt_adjoint, and the result is the same.  Every call on
 * one GPU shares that GPU's array, so calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller.  The
 * stream must belong to the current device.
 */
#define SLT_ADJ_SCRATCH_BYTES ((int64_t)128 << 20)
int slt_adjoint_noalloc(const void* packed, const slt_grid_desc* g, const slt_adjoint_args* a,
                        slt_stream_t stream);

/*
 * d loss / d n_e from the voxel adjoint: the transposed numpy.gradient stencil
 * (src/core/propagator.py:90-92) times grad_coeff, plus d kappa / d n_e
 * (propagator.py:147-164) times the kappa adjoint.  Arguments as in
 * slt_grid_pack; dne is (nx, ny, nz) C-order fp32.
 */
int slt_adjoint_unpack(const float* grid_adj, const slt_grid_desc* g, const float* ne,
                       const float* cx, const float* cy, const float* cz, double grad_coeff,
                       const float* te, int64_t n_te, const float* zz, int64_t n_z, double omega,
                       double e_charge, double epsilon_0, double speed_of_light,
                       double coulomb_log, float* dne, slt_stream_t stream);

/* ------------------------------------------------------------------------- */
/* Measurement.  No reference counterpart (the reference times only wall      */
/* clock around its solve, src/core/propagator.py:783,795).                   */
/* ------------------------------------------------------------------------- */

/* Independent chains per lane in the calibration kernels. */
#define SLT_CALIB_CHAINS 16

/* Instruction forms of the calibration kernels. */
enum {
    SLT_CALIB_FMA_SGPR = 0, /* v_fma_f32 (VOP3) with uniform (SGPR) multiplier and addend */
    SLT_CALIB_FMA_VGPR = 1, /* v_fma_f32 (VOP3) with per-lane (VGPR) operands            */
    SLT_CALIB_FMAC = 2,     /* v_fmac_f32 (VOP2), per-lane operands                        */
    SLT_CALIB_ADD = 3,      /* v_add_f32 (VOP2)                                            */
    SLT_CALIB_MUL = 4,      /* v_mul_f32 (VOP2)                                            */
    SLT_CALIB_FMA64 = 5     /* v_fma_f64                                                   */
};

/*
 * VALU-issue calibration: `blocks` workgroups of 256 lanes, each lane running
 * SLT_CALIB_CHAINS independent chains of instruction form `kind` for `iters`
 * iterations, so each wave64 issues SLT_CALIB_CHAINS * iters instructions of
 * that form (and a few dozen other VALU instructions); no memory traffic but
 * one float per lane stored to out (blocks * 256 floats, device).  The
 * fastest form's instruction rate is the VALU-issue peak that the trace
 * kernel's measured instruction rate is a fraction of.
 */
int slt_valu_calibrate(int32_t kind, int32_t blocks, int32_t iters, float* out,
                       slt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SLOTH_HIP_H */
