/*
 * synchrad_amd.h -- C ABI of the MI355X (gfx950) spectral radiation library.
 *
 * Every pointer named "device" is a HIP device pointer; every call is ordered
 * on the `stream` argument (a hipStream_t, NULL = the default stream) and
 * returns without synchronising.  The library never allocates: buffers are
 * owned by the caller (torch tensors in the Python host) and borrowed for the
 * duration of the call.  Functions return SR_OK (0) or a negative sr_status;
 * sr_last_error() gives a message for the calling thread's last failure.
 *
 * Spectra are laid out (nSnaps, nPhi, nTheta|nRadius, nOmega) with omega
 * fastest, as the reference's device buffers are (kernel_farfield.cl:102),
 * and every entry point ADDS its contribution to them, as the reference's
 * kernels do (`spectrum[...] +=`, kernel_farfield.cl:102).
 */
#ifndef SYNCHRAD_AMD_H
#define SYNCHRAD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SR_ABI_VERSION 3

typedef enum sr_status {
  SR_OK = 0,
  SR_ERR_INVALID = -1,     /* bad argument, shape or size */
  SR_ERR_UNSUPPORTED = -2, /* valid request this build does not implement */
  SR_ERR_HIP = -3,         /* a HIP runtime call failed */
  SR_ERR_WORKSPACE = -4    /* workspace missing or too small */
} sr_status;

/* field mode, synchrad/calc.py:358-359 ('far' / 'near') */
enum { SR_MODE_FAR = 0, SR_MODE_NEAR = 1 };

/* vector components, synchrad/calc.py:458-465 */
enum {
  SR_COMP_TOTAL = 0,             /* 1 output:  total                         */
  SR_COMP_CARTESIAN = 1,         /* 3 outputs: x, y, z                       */
  SR_COMP_CARTESIAN_COMPLEX = 2, /* 6 outputs: xre, xim, yre, yim, zre, zim  */
  SR_COMP_SPHERIC = 3,           /* 3 outputs: r, theta, phi   (far only)    */
  SR_COMP_SPHERIC_COMPLEX = 4    /* 6 outputs: rre .. phiim    (far only)    */
};

/*
 * Arithmetic.  SR_PREC_F64 is the reference's double kernels
 * (calc_input['dtype'] = 'double').  SR_PREC_SPLIT keeps the retarded-time
 * phase omega*(t - n.r) (far) or omega*(t + |P - r|) (near) in FP64, reduces
 * it modulo 2*pi in FP64, and runs amplitudes, sin/cos and the compensated
 * step sums in FP32 (calc_input['dtype'] = 'float').
 */
enum { SR_PREC_F64 = 0, SR_PREC_SPLIT = 1 };

/*
 * Sweep of the split path.  SR_SWEEP_DIRECT forms every phase in FP64.
 * SR_SWEEP_RECURRENCE, for grids that declare a constant omega_step, forms
 * exp(i omega tau) in FP64/FP32 for one omega per thread and advances it to the
 * thread's other omegas by complex multiplication, and sweeps only the steps
 * that keep some omega of each 64-omega run.  Both apply the |dphase| < pi
 * filter as the equivalent threshold |omega| < pi/|tau_s - tau_{s-1}|.
 * SR_SWEEP_AUTO selects the recurrence on linear grids and the direct sweep
 * otherwise.  The FP64 path always sweeps directly and filters
 * exactly as the reference does.
 */
enum { SR_SWEEP_AUTO = 0, SR_SWEEP_DIRECT = 1, SR_SWEEP_RECURRENCE = 2 };

/*
 * A batch of tracks packed CSR-style: particle p owns samples
 * [offsets[p], offsets[p+1]) of the x..uz arrays.  Replaces the per-particle
 * host loop of SynchRad.calculate_spectrum (synchrad/calc.py:257-267) and the
 * per-track uploads of _track_to_device (synchrad/calc.py:579-603).
 *
 * Time alignment follows _process_track (synchrad/calc.py:296-307): for
 * particle p the kernel runs iterations it < it_end[p]-1, the step is active
 * while it < n_p - 1, the global iteration is it_start[p] + it, and snapshots
 * are read from it_snaps + p*snaps_stride (snaps_stride = 0: one shared list,
 * the global it_range mode; snaps_stride = n_snaps: one list per particle,
 * the per-track mode where it_start = 0 and it_end = n_p).
 */
typedef struct sr_tracks {
  int64_t n_particles;
  int64_t n_points;            /* offsets[n_particles] */
  const int64_t *offsets;      /* device [n_particles + 1] */
  const int64_t *offsets_host; /* optional host copy, used to balance particle slices */
  const double *x, *y, *z;     /* device [n_points], coordinates (length unit) */
  const double *ux, *uy, *uz;  /* device [n_points], momenta (beta*gamma) */
  const double *weight;        /* device [n_particles] */
  const uint32_t *it_start;    /* device [n_particles] */
  const uint32_t *it_end;      /* device [n_particles], >= 1 */
  const uint32_t *it_snaps;    /* device, see above */
  uint32_t n_snaps;            /* >= 1 */
  uint32_t snaps_stride;       /* 0 or n_snaps */
  double dt;                   /* timeStep = c*dt, synchrad/calc.py:125-127 */
} sr_tracks;

/*
 * Spectral grid, as SynchRad._init_data uploads it (synchrad/calc.py:494-512).
 * Far: omega (already multiplied by 2*pi), sin/cos of theta and phi.
 * Near: omega, radius, sin/cos of phi and the screen distance; the observer
 * of (radius i, phi j) is (radius[i] cos phi_j, radius[i] sin phi_j, l_screen)
 * (kernel_nearfield.cl:40-42).  A Cartesian screen (no counterpart in the
 * reference, whose screen is polar) sets screen_y: the observer of (i, j) is
 * then (radius[i], screen_y[j], l_screen) -- radius holds the n_theta x
 * positions, screen_y the n_phi y positions -- and sin_phi, cos_phi are not
 * read.
 */
typedef struct sr_grid {
  int32_t mode;             /* SR_MODE_FAR / SR_MODE_NEAR */
  uint32_t n_omega;
  uint32_t n_theta;         /* near: number of radii */
  uint32_t n_phi;
  const double *omega;      /* device [n_omega] */
  const double *sin_theta;  /* device [n_theta]  (far) */
  const double *cos_theta;  /* device [n_theta]  (far) */
  const double *radius;     /* device [n_theta]  (near) */
  const double *sin_phi;    /* device [n_phi] */
  const double *cos_phi;    /* device [n_phi] */
  double l_screen;          /* near: distance to the screen */
  double omega_step;        /* omega[i+1] - omega[i] if constant (linear grids), else 0 */
  const double *screen_y;   /* near: device [n_phi] y positions of a Cartesian screen, or NULL */
} sr_grid;

typedef struct sr_opts {
  int32_t comp;             /* SR_COMP_* */
  int32_t precision;        /* SR_PREC_* */
  int32_t omega_per_lane;   /* 0 = tuned, else 1, 2 or 4; a value the variant lacks
                               falls back: split direct 1, 2, 4 far / 1, 2 near, else 1;
                               recurrence 2, 4 far / 2 near, else 2; f64 1, 2, else 1 */
  int32_t waves_per_block;  /* 0 = tuned, else 1, 2 or 4 */
  int32_t slices;           /* 0 = tuned, else 1..32 particle slices */
  int32_t persistent;       /* 0 = one block per tile; >0 = that many blocks loop over tiles */
  int32_t sweep;            /* SR_SWEEP_* */
  const double *form_factor;/* device [n_omega]; far cartesian_complex only */
  int32_t tile_order;       /* 0 = batch calls launch the far-field tiles heaviest first (estimated
                               from sampled steps) when the launch is large; 1 = index order.
                               Results are identical either way. */
} sr_opts;

int sr_abi_version(void);
const char *sr_last_error(void);

/* Number of output arrays for (mode, comp): 1, 3 or 6; negative if invalid. */
int sr_num_outputs(int32_t mode, int32_t comp);

/*
 * Workspace bytes sr_spectrum_batch needs for this request: the
 * direction-independent data of every sample (far: a and the interval's mean
 * velocity, 6 doubles; near: the velocity, 3 doubles), computed once per call
 * instead of once per spectral point, plus, when the plan uses particle
 * slices, slices*n_out*nSnaps*n_grid doubles of partial spectra summed in
 * slice order.  0 for an empty batch.
 */
int sr_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                       size_t *bytes);

/*
 * One spectrum of a whole CSR batch in one stream-ordered launch sequence
 * (a tile kernel plus, with particle slices, a fixed-order slice reduction).
 * Replaces SynchRad.calculate_spectrum's particle loop and its per-particle
 * launches (synchrad/calc.py:257-267, 292-353).  `outs` holds
 * sr_num_outputs() device arrays of nSnaps*n_phi*n_theta*n_omega doubles.
 */
int sr_spectrum_batch(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                      double *const *outs, void *workspace, size_t workspace_bytes,
                      void *stream);

/*
 * The plan sr_spectrum_batch resolves for a request: sweep, omegas per lane,
 * waves per block, particle slices, tiles, and whether the tiles are launched
 * heaviest first.  No counterpart in the reference (whose work-group size is
 * fixed, synchrad/calc.py:633-646).
 */
typedef struct sr_plan_info {
  int32_t sweep;            /* SR_SWEEP_DIRECT or SR_SWEEP_RECURRENCE */
  int32_t omega_per_lane;
  int32_t waves_per_block;
  int32_t slices;
  int64_t tiles;
  int32_t ordered;          /* 1: heaviest tiles first */
} sr_plan_info;
int sr_plan(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts, sr_plan_info *info);

/*
 * Launch-shape autotuning.  Times sr_spectrum_batch on the given request once
 * per candidate -- each candidate's plan fields (sweep, omega_per_lane,
 * waves_per_block, slices, persistent, tile_order) over base's comp,
 * precision and form_factor -- with `warmup` untimed runs and `reps` timed
 * ones, HIP events on `stream`; ms[i] is candidate i's mean time (INFINITY for a
 * plan the request cannot run or whose workspace does not fit) and *best the
 * fastest.  The caller chooses the request: the real one, or a representative
 * sample of it (the Python host samples large launches as full-length tracks
 * on a subset of the directions).  `outs` receive the sum of all runs (scratch
 * to the caller); the workspace must hold sr_autotune_workspace_bytes.  Unlike
 * every other entry point this one synchronises the stream (to read the
 * timers).  No counterpart in the reference.
 */
int sr_autotune_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *base,
                                const sr_opts *candidates, int n_candidates, size_t *bytes);
int sr_autotune(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *base,
                const sr_opts *candidates, int n_candidates, double *const *outs,
                void *workspace, size_t workspace_bytes, int warmup, int reps, void *stream,
                float *ms, int *best);

/*
 * sr_spectrum_batch without a workspace argument and without any allocation:
 * no counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The tile kernels compute the direction-independent
 * step data themselves, and the particle slices' partial spectra and the tile
 * launch order live in a static device array of SR_NOALLOC_SCRATCH_BYTES per
 * GPU, held from the library's load to the end of the process.  The plan is
 * sr_spectrum_batch's; with opts.slices = 0 the slice count is halved until
 * its partial spectra fit the array, and a request that still does not fit
 * (an explicit opts.slices) fails with SR_ERR_WORKSPACE.  Results equal
 * sr_spectrum_batch's with the same slice count.  Every call on one GPU uses
 * that GPU's one array: calls that overlap in time on one GPU (concurrent
 * streams, or several host threads) race and must be serialised by the caller.
 */
#define SR_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int sr_spectrum_batch_noalloc(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                              double *const *outs, void *stream);

/*
 * Adjoint of the far-field spectrum (comp SR_COMP_TOTAL or SR_COMP_CARTESIAN,
 * one snapshot).  No counterpart in the reference, which has no gradients.
 * Given dS = dL/dS (1 or 3 device arrays laid out like the spectra) it ADDS
 * dL/dx, dL/dy, dL/dz, dL/dux, dL/duy, dL/duz (device [n_points]) and dL/dw
 * (device [n_particles]).  The |dphase| < pi mask of the forward is held
 * fixed.  SR_PREC_F64 evaluates everything in FP64; SR_PREC_SPLIT forms
 * phases in FP64 and sin/cos and the omega sums in FP32.  opts.slices sets the
 * number of direction groups (0 = chosen); opts.waves_per_block the block size
 * of the transpose sweep (threads = steps).  The workspace holds the group
 * slabs, the particles' complex fields for a chunk of particles (at most
 * ~2 GiB; larger batches are processed chunk by chunk) and the step data.
 */
int sr_far_adjoint_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                                   size_t *bytes);
int sr_far_adjoint_batch(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                         const double *const *dS,
                         double *gx, double *gy, double *gz,
                         double *gux, double *guy, double *guz, double *gw,
                         void *workspace, size_t workspace_bytes, void *stream);

/*
 * The forward half of the adjoint, kept for its backward: every particle's
 * complex far field F_p(g) = sum_s A_s e^{i phase_s} (no weight, no dt), as
 * fields[6][n_particles][n_grid] doubles (x, y, z as re/im pairs, grid index
 * as in the spectra), and, when `outs` is not NULL, the spectra of comp
 * SR_COMP_TOTAL (1 array) or SR_COMP_CARTESIAN (3) ADDED from those fields
 * (w dt^2 |F_p|^2 particle by particle in particle order, the records'
 * expressions of synchrad/kernel_farfield.cl:100-106, 208-222).  No
 * counterpart in the reference.  Same restrictions as the adjoint (far field,
 * one snapshot); the workspace holds the step data (6 * n_points doubles) and
 * the tile launch order (8 bytes per tile).
 */
int sr_far_fields_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                                  size_t *bytes);
int sr_far_fields_batch(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                        double *fields, double *const *outs,
                        void *workspace, size_t workspace_bytes, void *stream);

/*
 * sr_far_adjoint_batch with the fields of sr_far_fields_batch (same tracks,
 * grid and precision) passed in instead of recomputed: the backward of a
 * forward that kept them.  The workspace holds only the group slabs and the
 * step data.  No counterpart in the reference.
 */
int sr_far_adjoint_fields_workspace_bytes(const sr_tracks *tracks, const sr_grid *grid,
                                          const sr_opts *opts, size_t *bytes);
int sr_far_adjoint_from_fields(const sr_tracks *tracks, const sr_grid *grid, const sr_opts *opts,
                               const double *const *dS, const double *fields,
                               double *gx, double *gy, double *gz,
                               double *gux, double *guy, double *guz, double *gw,
                               void *workspace, size_t workspace_bytes, void *stream);

/*
 * Per-particle entry points with the argument order of the reference's
 * OpenCL kernels, plus a precision selector and a stream.  One call adds one
 * particle's contribution, as one reference launch does.
 */

/* replaces far `total`, synchrad/kernel_farfield.cl:6-28 */
int sr_far_total(double *spectrum,
                 const double *x, const double *y, const double *z,
                 const double *ux, const double *uy, const double *uz,
                 double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                 const double *omega, const double *sinTheta, const double *cosTheta,
                 const double *sinPhi, const double *cosPhi,
                 uint32_t nOmega, uint32_t nTheta, uint32_t nPhi,
                 double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                 int32_t precision, void *stream);

/* replaces far `cartesian_comps`, synchrad/kernel_farfield.cl:111-135 */
int sr_far_cartesian(double *spectrum1, double *spectrum2, double *spectrum3,
                     const double *x, const double *y, const double *z,
                     const double *ux, const double *uy, const double *uz,
                     double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                     const double *omega, const double *sinTheta, const double *cosTheta,
                     const double *sinPhi, const double *cosPhi,
                     uint32_t nOmega, uint32_t nTheta, uint32_t nPhi,
                     double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                     int32_t precision, void *stream);

/* replaces near `total`, synchrad/kernel_nearfield.cl:5-27 */
int sr_near_total(double *spectrum,
                  const double *x, const double *y, const double *z,
                  const double *ux, const double *uy, const double *uz,
                  double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                  const double *omega, const double *radius,
                  const double *sinPhi, const double *cosPhi, double distanceToScreen,
                  uint32_t nOmega, uint32_t nRadius, uint32_t nPhi,
                  double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                  int32_t precision, void *stream);

/* replaces near `cartesian_comps`, synchrad/kernel_nearfield.cl:106-130 */
int sr_near_cartesian(double *spectrum1, double *spectrum2, double *spectrum3,
                      const double *x, const double *y, const double *z,
                      const double *ux, const double *uy, const double *uz,
                      double wp, uint32_t itStart, uint32_t itEnd, uint32_t nSteps,
                      const double *omega, const double *radius,
                      const double *sinPhi, const double *cosPhi, double distanceToScreen,
                      uint32_t nOmega, uint32_t nRadius, uint32_t nPhi,
                      double dt, uint32_t nSnaps, const uint32_t *itSnaps,
                      int32_t precision, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SYNCHRAD_AMD_H */
