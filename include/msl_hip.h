/*
 * msl_hip.h -- C ABI of the MI355X (gfx950) skeleton library libmsl_hip.so.
 *
 * This is the drop-in boundary under the msl::DA / DM / DC<T> class
 * templates (muesli4_amd/include/).  The reference (NinaHerrmann/muesli4) is
 * header-only C++/CUDA with no FFI; each entry point below names the
 * reference skeleton or runtime function whose device work it replaces.
 *
 * Conventions
 *  - Every function returns int: MSL_OK (0) or a negative MSL_E_* code
 *    (msl_server_poll may also return MSL_NOT_READY).
 *    msl_last_error() returns a message for the calling thread's last error.
 *    Nothing throws across this boundary.  (Reference: errors print and
 *    continue, src/muesli.cpp:292-294; fatal paths exit(), src/ds.cpp:231.)
 *  - Pointers named d_* are device pointers (from msl_malloc or any HIP
 *    allocation on the current device); h_* are host pointers.
 *  - All launches are stream-ordered on the library stream (msl_stream()).
 *    Entry points that return a value to the host synchronise that stream.
 *  - x = column, y = row, z = depth.  DM layout row*cols+col, DC layout
 *    (z*rows+y)*cols+x (reference include/nplmatrix.h:141-143,
 *    include/plcube.h:121-123).
 *  - Floating point: IEEE binary32, no FMA contraction, correctly rounded
 *    division; elementwise and stencil results are bit-identical to a scalar
 *    CPU loop with the same operation order.
 */
#ifndef MSL_HIP_H
#define MSL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSL_OK 0
#define MSL_E_INVALID (-1)   /* bad argument (shape, radius, null pointer)   */
#define MSL_E_HIP (-2)       /* HIP runtime error                             */
#define MSL_E_NOMEM (-3)     /* device allocation failed / budget exceeded    */
#define MSL_E_STATE (-4)     /* not initialised, or call not allowed now      */
#define MSL_E_TIMEOUT (-5)   /* bounded wait expired                          */
#define MSL_E_FULL (-6)      /* queue full (serving back-pressure)            */
#define MSL_NOT_READY 1      /* not an error: the polled job is not done yet  */

/* ------------------------------------------------------------------ */
/* Runtime (reference: msl::initSkeletons / terminateSkeletons /
 * syncStreams, src/muesli.cpp:63-158, 213-223; one stream per device,
 * src/muesli.cpp:78-85).                                              */
int msl_init(int device);
int msl_finalize(void);
const char *msl_last_error(void);
/* hipStream_t of the library stream, as void*. */
void *msl_stream(void);
int msl_synchronize(void);
/* Device properties: compute units, XCD count estimate, LDS bytes per CU. */
int msl_device_info(int *num_cus, int *warp_size, size_t *total_mem, char *name, int name_len);

/* ------------------------------------------------------------------ */
/* Residency (reference: DS constructors / updateDevice / updateHost /
 * gather, src/ds.cpp:63-131, 441-484, 538-591).                        */
int msl_malloc(void **d_ptr, size_t bytes);
int msl_free(void *d_ptr);
int msl_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes);
int msl_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes);
int msl_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);
/* The library's pinned result slot (replaces the device-to-host copy of a
 * fold result, src/ds.cpp:739-912): *h_slot is host memory that kernels may
 * store to directly; its contents are valid after msl_synchronize() and are
 * overwritten by the next call that uses the slot (small msl_memcpy_d2h,
 * DS::fold).  64 KiB.                                                   */
int msl_result_slot(void **h_slot, size_t *h_bytes);

/* ------------------------------------------------------------------ */
/* Fills (reference: DS::fill(const T&), src/ds.cpp:289-305; seeded
 * generators replace the examples' unseeded rand(),
 * examples/gameoflife.cpp:87-89).  Hash = SplitMix64(seed, index).   */
int msl_fill_f32(float *d_out, size_t n, float value);
int msl_fill_i32(int32_t *d_out, size_t n, int32_t value);
/* d_out[i] = (float)(hash(seed,i) >> 40) * scale */
int msl_fill_uniform_f32(float *d_out, size_t n, uint64_t seed, float scale);
/* d_out[i] = hash(seed,i) < threshold ? 1 : 0 */
int msl_fill_bernoulli_i32(int32_t *d_out, size_t n, uint64_t seed, uint64_t threshold);
/* d_out[i] = (uint32)(hash(seed,i) >> 32) */
int msl_fill_u32(uint32_t *d_out, size_t n, uint64_t seed);

/* ------------------------------------------------------------------ */
/* Elementwise skeletons.
 * msl_map_affine_f32: d_out[i] = d_in[i]*a + b (two roundings).
 *   Replaces DS::map / mapInPlace with functor x*a+b (src/ds.cpp:612-663,
 *   kernel src/map_kernels.cu:36-43).  d_out == d_in is mapInPlace.     */
int msl_map_affine_f32(float *d_out, const float *d_in, size_t n, float a, float b);
/* uint32 wrapping variant (DA<int> map(x*a+b)). */
int msl_map_affine_u32(uint32_t *d_out, const uint32_t *d_in, size_t n, uint32_t a, uint32_t b);
/* d_out[i] = d_a[i] + d_b[i].  Replaces DS::zip / zipInPlace with a+b
 * (src/ds.cpp:667-727, kernel src/zip_kernels.cu:37-43).               */
int msl_zip_add_f32(float *d_out, const float *d_a, const float *d_b, size_t n);
/* d_out[row*cols+col] = (float)(row*cols+col).  Replaces
 * DM::mapIndexInPlace with f(row,col,v) = row*cols+col
 * (src/dm.cpp:370-401, kernel src/map_kernels.cu:60-70).               */
int msl_mapindex_lin_f32(float *d_out, int rows, int cols);

/* ------------------------------------------------------------------ */
/* Folds (reference: DS::fold, src/ds.cpp:739-912; kernel
 * src/fold_kernels.cu:51-129).  Result stays on the device; the _host
 * forms copy the scalar back (one small D2H, not the whole array).
 * msl_fold_sum_f32: compensated FP32 (TwoSum per lane and across lanes),
 *   FP64 only in the final cross-block combine.  d_result receives
 *   {double sum} ; h_sum64 / h_sum32 receive the double and its float
 *   rounding.  Deterministic for a given device.                        */
int msl_fold_sum_f32(const float *d_in, size_t n, double *d_result);
int msl_fold_sum_f32_host(const float *d_in, size_t n, double *h_sum64, float *h_sum32);
/* msl_fold_sum_f32 without any allocation: no reference counterpart.  For
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The block
 * partials live in a fixed 32 KiB device array held per GPU until the
 * process ends (at most 4096 partials; the default launch uses two per CU,
 * and a launch shape that needs more is rejected with MSL_E_INVALID).  Same
 * kernels and launch shape as msl_fold_sum_f32, so the result is
 * bit-identical.  Calls that overlap on one GPU share the array and race.  */
int msl_fold_sum_f32_noalloc(const float *d_in, size_t n, double *d_result);
/* Wrapping uint32 sum (int fold with defined overflow). */
int msl_fold_sum_u32_host(const uint32_t *d_in, size_t n, uint32_t *h_sum);
/* Exact int32 sum into int64 (GoL live cells). */
int msl_fold_sum_i32_host(const int32_t *d_in, size_t n, int64_t *h_sum);
/* max over d_in (n > 0). */
int msl_fold_max_f32_host(const float *d_in, size_t n, float *h_max);

/* ------------------------------------------------------------------ */
/* 2-D stencils (reference: DM::mapStencil<f>, src/dm.cpp:1535-1591;
 * view include/nplmatrix.h:123-139; kernel
 * include/detail/map_kernels.cuh:159-168).  Neutral value nv outside.
 * msl_stencil2d_mean_f32: out(x,y) = (sum_dy (sum_dx in(x+dx,y+dy)))/(2r+1)^2
 *   with both sums starting at 0 and ascending; 1 <= r <= 16.            */
int msl_stencil2d_mean_f32(float *d_out, const float *d_in, int rows, int cols, int r, float nv);
/* iters ping-pong steps between d_a and d_b starting from d_a, captured
 * once into a HIP graph and replayed.  *h_result_in_b = 1 when the final
 * grid is in d_b (odd iters).  Launches may apply several generations
 * (r = 1: up to six; r = 2..7 on large grids with rows of 4 or more
 * cells: two) with results
 * bit-identical to one launch per generation; d_b then holds intermediate
 * generations other than those of the one-per-launch schedule.            */
int msl_stencil2d_mean_iterate_f32(float *d_a, float *d_b, int rows, int cols, int r, float nv, int iters,
                                   int *h_result_in_b);
/* The launch plan msl_stencil2d_mean_iterate_f32 would use for the same
 * buffers, shape, radius and count: *h_launches launches, h_gens[k] the
 * generations of launch k (at most `capacity` entries written).  No
 * reference counterpart (the reference launches one generation per
 * skeleton call); it lets callers and tests see which kernels run.      */
int msl_stencil2d_mean_iterate_plan(const float *d_a, const float *d_b, int rows, int cols, int r, int iters,
                                    int *h_gens, int capacity, int *h_launches);
/* Game of Life B3/S23, cells outside are 0 (reference:
 * examples/gameoflife.cpp:18-43 through DM::mapStencilMM,
 * src/dm.cpp:615-1208).                                                 */
int msl_gol_i32(int32_t *d_out, const int32_t *d_in, int rows, int cols);
int msl_gol_iterate_i32(int32_t *d_a, int32_t *d_b, int rows, int cols, int iters, int *h_result_in_b);

/* ------------------------------------------------------------------ */
/* 3-D stencils (reference: DC::mapStencil<f>, src/dc.cpp:559-592; view
 * include/plcube.h:105-119; kernel include/detail/map_kernels.cuh:140-149;
 * functor heat3D examples/heatdiffusion.cpp:55-65).
 * msl_heat7_f32: v = 0 + in(x-1) + in(x+1) + in(y-1) + in(y+1) + in(z-1)
 *   + in(z+1) (left to right), out = v / 6.                              */
int msl_heat7_f32(float *d_out, const float *d_in, int nx, int ny, int nz, float nv);
/* heat7 step fused with max |out - in| folded into *d_maxdiff (atomic max;
 * the caller zeroes it, e.g. with msl_fill_f32).                         */
int msl_heat7_residual_f32(float *d_out, const float *d_in, int nx, int ny, int nz, float nv, float *d_maxdiff);
/* iters steps ping-pong from d_a; every residual_every-th step (1-based)
 * is the fused residual step and writes h_residuals[k] for k-th such step
 * (h_residuals holds iters/residual_every floats; may be NULL when
 * residual_every == 0).  Captured once into a HIP graph and replayed.   */
int msl_heat7_iterate_f32(float *d_a, float *d_b, int nx, int ny, int nz, float nv, int iters, int residual_every,
                          float *h_residuals, int *h_result_in_b);
/* The launch plan of msl_heat7_iterate_f32 (generations per launch, in
 * order), as msl_stencil2d_mean_iterate_plan.  No reference counterpart. */
int msl_heat7_iterate_plan(int nx, int ny, int nz, int iters, int residual_every, int *h_gens, int capacity,
                           int *h_launches);

/* ------------------------------------------------------------------ */
/* Stream-ordered capture of any sequence of the calls above into a
 * replayable HIP graph (the reference launches and synchronises every
 * skeleton individually, src/muesli.cpp:213-223).                       */
int msl_graph_begin(void);
int msl_graph_end(void **graph);
int msl_graph_launch(void *graph);
int msl_graph_destroy(void *graph);

/* ------------------------------------------------------------------ */
/* Diagnostics (no reference counterpart).  The mean-blur kernels divide
 * by (2r+1)^2 with a multiply/FMA sequence instead of the IEEE division
 * instruction sequence; this runs both forms (x / d, and (0 + x) / d for
 * column sums formed without their leading +0) against IEEE division on
 * all 2^32 float inputs on the device and returns the number of bitwise
 * mismatches (0 expected).                                              */
int msl_selftest_div_odd_square(int r, uint64_t *h_mismatches);
/* The same check for the heat kernels' division by 6 (the multiply/FMA
 * sequence plus the ties-to-even rule for quotients exactly halfway
 * between two subnormals).                                              */
int msl_selftest_div_six(uint64_t *h_mismatches);

/* ------------------------------------------------------------------ */
/* Batched mean blur.  No single reference counterpart: it replaces a loop
 * of DM::mapStencil (src/dm.cpp:1535-1591) over many grids, one launch and
 * synchronisation per grid per step, followed by DS::fold
 * (src/ds.cpp:739-912) per grid.
 *
 * A batch is a fixed list of independent grids, each with its own shape,
 * radius (1..10) and neutral value, and two device buffers a and b (step 0
 * reads a and writes b, step 1 reads b and writes a, ...).  Every output
 * is bit-identical to msl_stencil2d_mean_f32 on that grid alone.
 *
 * msl_batch_run: the steps run in launch rounds, each a persistent launch
 * over an atomic work queue of (grid, strip, segment) items; grids of
 * radius <= 8 with 256-column strips advance two generations per item
 * (pairs of steps, one pair split when the number of rounds would not
 * have the parity of iters), the others one.  The rounds and the per-grid
 * fold are captured once into a HIP graph (per iteration count and
 * d_sums) and replayed.  d_sums (nullable; ngrids doubles) receives the
 * compensated fold(sum) of each grid's final state (iters >= 1); the
 * result is in b exactly when iters is odd (*h_result_in_b).
 * MSL_BATCH_PAIRS=0 in the environment at msl_batch_create turns the
 * two-generation items off; MSL_BATCH_PAIR_SEG sets their rows (128).
 * msl_batch_run_loop: the plain loop, one single-grid launch per grid per
 * step and one msl_fold_sum_f32 per grid (the comparison baseline).
 * Buffers must not be freed while the batch exists.                      */
typedef struct msl_batch_grid {
    float *a;   /* initial state (device)                     */
    float *b;   /* second buffer (device), same size          */
    int rows;
    int cols;
    int r;      /* mean-blur radius, 1..10                    */
    float nv;   /* neutral value read outside the grid        */
} msl_batch_grid;

int msl_batch_create(const msl_batch_grid *h_grids, int ngrids, void **h_batch);
int msl_batch_destroy(void *batch);
/* Tile shape of one (radius, width class) cell: width classes are
 * cols < 96, < 192, < 384 and >= 384; cols_per_lane 4 (256-column wave
 * strips) or 1 (64-column strips); seg = rows per work item.           */
int msl_batch_set_tile(void *batch, int r, int width_class, int cols_per_lane, int seg);
int msl_batch_get_tile(const void *batch, int r, int width_class, int *h_cols_per_lane, int *h_seg);
/* Persistent launch shape: workgroups per CU (capped by occupancy) and
 * queue mode: 0 static round-robin, 1 one atomic queue per XCD over a
 * cost-balanced item range, 2 one global atomic queue.                 */
int msl_batch_set_launch(void *batch, int blocks_per_cu, int queue_mode);
/* Autotuner: times every tile candidate per (radius, width class) present
 * in the batch, then the launch shape, and keeps the fastest.  With
 * cache_path, a table saved for this device is loaded instead (then
 * *h_entries is minus the number of entries loaded), else the tuned table
 * is saved there.                                                       */
int msl_batch_autotune(void *batch, const char *cache_path, int *h_entries);
int msl_batch_info(const void *batch, int *h_items, int *h_blocks, int *h_blocks_per_cu, int *h_dynamic);
int msl_batch_run(void *batch, int iters, double *d_sums, int *h_result_in_b);
/* The launch plan msl_batch_run would use (no device work): rounds,
 * kernel launches in its graph, grids that run two generations per item
 * and work items over all rounds; with_fold as d_sums != NULL.          */
int msl_batch_run_plan(const void *batch, int iters, int with_fold, int *h_rounds, int *h_launches,
                       int *h_pair_grids, int *h_items);
int msl_batch_run_loop(void *batch, int iters, double *d_sums, int *h_result_in_b);

/* ------------------------------------------------------------------ */
/* Serving engine.  No reference counterpart: a Muesli program serves one
 * request at a time, building containers (src/ds.cpp:63-131), running each
 * skeleton with a synchronisation after every call (src/muesli.cpp:213-223)
 * and folding (src/ds.cpp:739-912); msl_serve_plain_loop keeps that model
 * as the comparison baseline.
 *
 * Jobs (initial state generated on the device from `seed`):
 *   MSL_JOB_BLUR2D      DM<float> d0 x d1 (rows, cols), values uniform in
 *                       [0,255); `steps` mean-blur steps of radius r (1..10)
 *                       with neutral value nv; result: compensated fold(sum)
 *                       of the final state (sum).
 *   MSL_JOB_HEAT3D      DC<float> d0 x d1 x d2 (nx, ny, nz), values uniform
 *                       in [0,1); `steps` 7-point heat steps with neutral
 *                       value nv; result: fold(sum) of the final state.
 *   MSL_JOB_MAPFOLD_U32 DA<u32> of length n, bits from the seed;
 *                       x <- x*mul + add (wrapping), then the wrapping
 *                       fold(sum) (sum_u32).
 * The engine admits queued jobs in arrival order while their HBM bytes fit
 * its budget (one device allocation made at creation) and a slot is free,
 * keeps each job's state resident, and every round advances all active
 * jobs by one step with one persistent launch per job kind (continuous
 * batching).  Final states and sums are bit-identical whether a job runs
 * alone or in a mix.  Threading: msl_server_submit may be called from any
 * thread; step / drain / poll / run_trace from one thread.              */
#define MSL_JOB_BLUR2D 1
#define MSL_JOB_HEAT3D 2
#define MSL_JOB_MAPFOLD_U32 3

typedef struct msl_job {
    int kind;
    int steps;         /* BLUR2D, HEAT3D: stencil steps (>= 1)             */
    int r;             /* BLUR2D: radius 1..10                             */
    int d0, d1, d2;    /* BLUR2D: rows, cols; HEAT3D: nx, ny, nz           */
    int64_t n;         /* MAPFOLD_U32: length                              */
    float nv;          /* stencils: neutral value outside the grid         */
    uint32_t mul, add; /* MAPFOLD_U32: x <- x*mul + add                    */
    uint64_t seed;     /* initial state                                    */
    int keep_final;    /* keep the final state until msl_server_poll      */
} msl_job;

typedef struct msl_job_result {
    int64_t id;
    int kind;
    double sum;        /* BLUR2D, HEAT3D                                   */
    uint32_t sum_u32;  /* MAPFOLD_U32                                      */
    double t_arrival;  /* seconds (server clock, or since the trace start) */
    double t_admit;
    double t_done;     /* when the host saw the result                     */
    size_t bytes;      /* HBM held while active                            */
} msl_job_result;

int msl_server_create(size_t hbm_budget_bytes, int max_active, void **h_server);
int msl_server_destroy(void *server);
/* t_arrival < 0: now.  MSL_E_NOMEM if the job can never fit the budget. */
int msl_server_submit(void *server, const msl_job *job, double t_arrival, int64_t *h_id);
/* One round (admit, advance every active job one step, retire finished
 * rounds); *h_active / *h_pending after the round.                     */
int msl_server_step(void *server, int *h_active, int *h_pending);
/* Rounds until nothing is pending, active or in flight.                */
int msl_server_drain(void *server);
/* MSL_OK and the result once the job is done (its record is then
 * released), MSL_NOT_READY while it is not.  h_final (nullable) receives
 * the final state of a job submitted with keep_final.                  */
int msl_server_poll(void *server, int64_t id, msl_job_result *h_result, void *h_final, size_t final_bytes);
int msl_server_info(const void *server, int64_t *h_rounds, size_t *h_peak_bytes, double *h_mean_active,
                    int64_t *h_done_jobs);
/* Replays an arrival trace in real time (arrival_s ascending, seconds from
 * the call) on an idle server and returns every job's result.         */
int msl_server_run_trace(void *server, const msl_job *jobs, const double *arrival_s, int njobs,
                         msl_job_result *h_results);
/* Baseline: the same trace, one job at a time, one launch plus one stream
 * synchronisation per step, then a fold copied to the host.            */
int msl_serve_plain_loop(const msl_job *jobs, const double *arrival_s, int njobs, size_t hbm_budget_bytes,
                         msl_job_result *h_results);

/* ------------------------------------------------------------------ */
/* Event timing on the library stream (for benchmarks). */
int msl_timer_start(void);
int msl_timer_stop(float *h_ms);

#ifdef __cplusplus
}
#endif
#endif
