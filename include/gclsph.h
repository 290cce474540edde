/*
 * libgclsph — C ABI of the MI355X (gfx950) cell-list + WCSPH step.
 *
 * Each entry point replaces one piece of the Julia API of GPUCellListSPH.jl
 * (exports at src/GPUCellListSPH.jl:9-17); the reference location is cited on
 * every declaration.  Conventions:
 *   - every function returns 0 on success and a negative GCLSPH_E* code on
 *     failure; gclsph_last_error() gives the message (thread-local).  The
 *     reference raises `error(...)` in the same situations;
 *   - arrays cross the boundary as host pointers and are copied; particle
 *     arrays are in the caller's particle order (index i = particle i+1 of the
 *     reference), vectors as D separate component arrays (the reference's
 *     NTuple{D,CuArray} SoA layout);
 *   - every call is blocking on return, like the reference's CUDA.@sync.
 *     Objects are not thread-safe; use one object from one thread at a time.
 *   - the library always runs on the GPU; it has no CPU fallback.  Creating
 *     an object without a usable gfx950 device fails with GCLSPH_ENODEVICE.
 */
#ifndef GCLSPH_H
#define GCLSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCLSPH_OK 0
#define GCLSPH_EARG (-1)        /* bad argument, dimension or length */
#define GCLSPH_EHIP (-2)        /* HIP runtime failure */
#define GCLSPH_ESTATE (-3)      /* not valid in the object's state */
#define GCLSPH_ENODEVICE (-4)   /* no usable GPU */
#define GCLSPH_EUNSUPPORTED (-5)

/* Precision of a cell list (its point type T) or of an SPH problem. */
#define GCLSPH_F64 64        /* Float64 everywhere (the reference's runs) */
#define GCLSPH_F32 32        /* GPUCellList{Float32} (cell list only) */
#define GCLSPH_F32SPLIT 33   /* SPH: FP32 pair terms summed in FP64, density as rho-rho0,
                                FP64 positions and FP64 dt reduction */

typedef struct gcl_s* gcl_t;
typedef struct sph_s* sph_t;
typedef struct sph_batch_s* sph_batch_t;

/* Message of the last failed call on this thread ("" if none). */
const char* gclsph_last_error(void);
/* Library version string. */
const char* gclsph_version(void);

/* ---------------------------------------------------------------- GPUCellList */

/* GPUCellList(points, cellsize, dist; MIN, MAX, mppcell, mpairs)
 * — src/celllist.jl:22-88.  pts: dim host arrays of n coordinates.
 * min/max: dim values or NULL (then the point extrema are used).
 * warn_flags (nullable): bit0 = a cell size was raised to dist (:32-39),
 * bit1 = the first pair list did not fit mpairs (:77-79). */
int gcl_create(int dim, int prec, int64_t n, const double* const* pts, const double* cellsize, double dist,
               const double* min_or_null, const double* max_or_null, int64_t mppcell, int64_t mpairs,
               int* warn_flags, gcl_t* out);
int gcl_destroy(gcl_t h);

/* update!(c, fillzero) — src/celllist.jl:95-134: re-bin the current points,
 * regrow the pair buffer by the reference's rule and rebuild the pair list. */
int gcl_update(gcl_t h, int fillzero);
/* partialupdate!(c, fillzero) — src/celllist.jl:141-145: pair list of the
 * current points over the cells of the last update!; no regrowth. */
int gcl_partialupdate(gcl_t h, int fillzero);

/* Replace / read the points (the reference mutates c.points in place).  On a
 * cell list owned by an SPH problem these are the problem's positions. */
int gcl_set_points(gcl_t h, const double* const* pts);
int gcl_get_points(gcl_t h, double* const* out);

/* c.pairsn and length(c.pairs) — src/celllist.jl:13-14. */
int gcl_pairsn(gcl_t h, int64_t* out);
int gcl_capacity(gcl_t h, int64_t* out);
/* neighborlist(c) — src/celllist.jl:152-154: the whole pair buffer,
 * 2*min(cap, capacity) int32 values (i, j), 1-based, i < j, zero padded.
 * neighborlistview(c) (:161-163) is the first pairsn of them; it fails with
 * GCLSPH_ESTATE when pairsn > capacity, as the reference's view would. */
int gcl_get_pairs(gcl_t h, int32_t* ij, int64_t cap, int64_t* written);

typedef struct {
  int32_t dim;
  int32_t prec;
  int64_t n;
  int64_t grid[3];
  double offset[3];   /* c.offset as stored in T */
  double cellsize[3]; /* c.cs after the dist clamp */
  double dist;
  int64_t mppcell;    /* size(c.celllist, 1) */
} gcl_info_t;
int gcl_info(gcl_t h, gcl_info_t* out);
/* c.cellpnum (column-major over the grid) and c.pcell (dim x n, 1-based). */
int gcl_get_cellpnum(gcl_t h, int32_t* out);
int gcl_get_pcell(gcl_t h, int32_t* out);
/* мaxpairs(c.cellpnum) — src/gpukernels_cls.jl:109-286. */
int gcl_maxpairs(gcl_t h, int64_t* out);

/* Neighbour list without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code that must not
 * call hipMalloc: stream-captured sequences, tight serving loops, harnesses
 * that forbid it).  It computes what GPUCellList(points, cellsize, dist;
 * MIN, MAX) followed by neighborlistview returns (src/celllist.jl:22-88,
 * :161-163): the same pairs in the same order as gcl_create + gcl_get_pairs
 * on the same Float64 points, 1-based (i, j) with i < j, with the cell grid
 * of the given MIN / MAX (host arrays of dim values; points outside clamp
 * into the edge cells).  pts_dev: dim DEVICE pointers to n Float64
 * coordinates.  pairs_dev: DEVICE buffer of 2*capacity int32, 8-byte aligned
 * (as hipMalloc returns it); the first
 * min(pairsn, capacity) pairs are written, entries after them are left as
 * they were.  pairsn (host): the total number of pairs.  stream: the
 * hipStream_t to run on (NULL: the null stream); the call returns when the
 * list is complete.  Intermediates live in a fixed device array of the
 * library (38 MB per device, held until the process ends), which bounds
 * n <= GCLSPH_NOALLOC_MAX_POINTS and the cell grid <=
 * GCLSPH_NOALLOC_MAX_CELLS cells (GCLSPH_EARG beyond).  Calls on one device
 * share that array: host threads are serialised, but calls that overlap on
 * one device in any other way race. */
#define GCLSPH_NOALLOC_MAX_POINTS 524288
#define GCLSPH_NOALLOC_MAX_CELLS 1048576
int gcl_neighborlist_noalloc(int dim, int64_t n, const double* const* pts_dev, const double* cellsize, double dist,
                             const double* min, const double* max, int32_t* pairs_dev, int64_t capacity,
                             int64_t* pairsn, void* stream);

/* ---------------------------------------------------------------- SPHProblem */

/* Parameter indices for sph_set_param / sph_get_param: the mutable fields of
 * SPHProblem (src/sphproblem.jl:62-112) plus Δt and the simulation time. */
enum {
  SPH_DX = 0, SPH_H, SPH_HH /* H */, SPH_RHO0, SPH_M0, SPH_DT, SPH_ALPHA, SPH_C0, SPH_GAMMA, SPH_DELTA,
  SPH_CFL, SPH_G, SPH_S, SPH_NU, SPH_DPC_L0, SPH_DPC_PMIN, SPH_DPC_PMAX, SPH_DPC_LAMBDA, SPH_XSPH_EPS,
  SPH_CSPMN, SPH_BOUND_D, SPH_BOUND_L, SPH_NUI, SPH_ETIME, SPH_NPARAMS
};

/* SPHProblem(system, dx, h, H, sphkernel, ρ, ptype, ρ₀, m₀, Δt, α, c₀, γ, δᵩ,
 * CFL; v, g, s) — src/sphproblem.jl:113-193.  params: SPH_NPARAMS doubles in
 * the enum order (γ and cspmn as integral values); the SPH kernel is
 * WendlandC2 of the system's dimension.  The problem takes over the system's
 * points (the reference aliases them), so `system` must outlive the problem.
 * v may be NULL (zeros). prec: GCLSPH_F64 or GCLSPH_F32SPLIT; the system
 * must be Float64. */
int sph_create(gcl_t system, int prec, const double* params, const double* rho, const int32_t* ptype,
               const double* const* v_or_null, sph_t* out);
int sph_destroy(sph_t h);
int sph_set_param(sph_t h, int which, double value);
int sph_get_param(sph_t h, int which, double* value);

/* stepsolve!(prob, n; timestepping, timelims) — src/sphproblem.jl:204-345.
 * Returns (updaten, maxcΔxout) through the two out pointers. */
int sph_stepsolve(sph_t h, int64_t n, int timestepping, double tl0, double tl1, int64_t* updaten,
                  double* maxcdx);

/* Fields for sph_get_field (vectors are dim consecutive blocks of n). */
enum {
  SPH_FIELD_X = 0,     /* get_points      src/sphproblem.jl:348 */
  SPH_FIELD_V = 1,     /* get_velocity    :352 */
  SPH_FIELD_RHO = 2,   /* get_density     :356 */
  SPH_FIELD_P = 3,     /* get_pressure    :360 (after the last call's last step) */
  SPH_FIELD_ACC = 4,   /* get_acceleration :364 */
  SPH_FIELD_SUMW = 5,  /* get_sumw        :380 */
  SPH_FIELD_SUMGRADW = 6, /* get_sumgradw :384 */
  SPH_FIELD_DRHO = 10, /* prob.∑∂ρ∂t of the last step */
  SPH_FIELD_PTYPE = 12, /* get_ptype :388 (as doubles) */
  SPH_FIELD_DPC = 13   /* get_dpccorr :368-370, which names a field the reference's SPHProblem lacks:
                          here the per-particle Δv_dpc of the last step's dpcreg! (zero without DPC) */
};
int sph_get_field(sph_t h, int field, double* out);

/* Batched ensemble (no counterpart in the reference: the batching pattern for
 * many independent scenes of uneven size).  A batch steps `count` existing
 * problems of one dimension and precision together: every stage of the step
 * is one launch over all of them, while each problem keeps its own Δt,
 * rebuild schedule, time and corrections, so every problem ends exactly as if
 * stepped alone with sph_stepsolve.  The problems stay usable on their own
 * (getters, parameters, single stepping); a problem cannot be destroyed while
 * a batch holds it. */
int sph_batch_create(const sph_t* problems, int32_t count, sph_batch_t* out);
int sph_batch_destroy(sph_batch_t b);
/* n steps of every problem (stepsolve! per problem, src/sphproblem.jl:207-345).
 * tl0 / tl1: per-problem Δt limits, or NULL for each problem's default
 * (sqrt(eps), CFL·H/(3c0)).  updaten / maxcdx: count outputs. */
int sph_batch_stepsolve(sph_batch_t b, int64_t n, int timestepping, const double* tl0, const double* tl1,
                        int64_t* updaten, double* maxcdx);

/* Per-stage GPU timing over the calls since the last reset, measured with HIP
 * events on the problem's stream (enabled by sph_set_timing(h, 1)).  Writes up
 * to cap entries of (total milliseconds, launches) for stages in this order:
 * rebuild, pass1, pass2, corrections, finalize. */
int sph_set_timing(sph_t h, int on);
int sph_get_timing(sph_t h, double* ms, int64_t* launches, int cap, int* nstages);

/* The pair-pass variants the autotuner chose for a problem (its own calls) or
 * a batch (no counterpart in the reference: the autotuner is this library's).
 * The first call of more than 15 steps times them on 15 of its first 31
 * steps, after up to 16 untimed ones (the device warming up).
 * out4: variant of pass 1, variant of pass 2, threads per block of pass 1,
 * of pass 2; -1 until tuned.  A variant is codegen + 4·b + 12·c, codegen in
 * [0, 4) (partner prefetch on/off, occupancy target), b the block size (0:
 * 256, 1: 128, 2: 512 threads) and c an occupancy cap (1: at most 4 waves per
 * SIMD); every variant gives identical results. */
int sph_get_pass_variants(sph_t h, int32_t* out4);
int sph_batch_get_pass_variants(sph_batch_t b, int32_t* out4);

#ifdef __cplusplus
}
#endif

#endif /* GCLSPH_H */
