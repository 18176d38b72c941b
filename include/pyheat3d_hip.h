/* pyheat3d_hip.h -- C ABI of the MI355X (gfx950) stencil library libpyheat3d_hip.so.
 *
 * The reference has no C ABI: its native boundary is CuPy's RawModule, which compiles
 * src/pyheat3d/backends/stencil.cu with NVRTC and launches kernels with the argument tuple
 * (u, o, [nz,] ny, nx, c0, c1) (backends/cupy_backend.py:86-100, 183-186). Each entry point
 * below names the reference interface it replaces. Python binds them with ctypes
 * (pyheat3d_amd/_hip.py).
 *
 * Conventions
 *   - Every function returns 0 on success or a negative PH3_E* code; ph3_last_error() gives a
 *     thread-local message for the last failure on the calling thread.
 *   - Device pointers are plain addresses from ph3_malloc (or any hipMalloc'd memory).
 *   - Streams and events are opaque handles (hipStream_t / hipEvent_t); a NULL stream is the
 *     device's null stream.
 *   - Arrays are C order (z, y, x) or (y, x), x contiguous, no padding, like the reference.
 *   - storage: 0 = float32, 1 = float16, 2 = bfloat16 (bit pattern). Arithmetic is float32.
 *   - Kernel launches are asynchronous; only the *_sync functions and ph3_memcpy block.
 *   - Calls sharing one stream are not thread-safe against each other.
 */
#ifndef PYHEAT3D_HIP_H
#define PYHEAT3D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PH3_OK 0
#define PH3_EINVAL (-1)   /* bad argument (shape, storage, kind, alignment) */
#define PH3_EHIP (-2)     /* a HIP runtime call failed */
#define PH3_ENOMEM (-3)   /* device allocation failed */
#define PH3_EUNSUP (-4)   /* valid request that this library does not support */

/* sweep kinds (the reference's rungs, backends/cupy_backend.py:115-194) */
#define PH3_NAIVE 0   /* one site per thread            (stencil.cu:43-93 naive2d/naive3d) */
#define PH3_ZMARCH 1  /* 3D 2.5D z-march / 2D y-march   (stencil.cu:101-161 zmarch3d, tb2d K=1) */
#define PH3_TB 2      /* k fused steps per pass         (stencil.cu:178-385 tb3d/tb2d) */

/* ---- errors, version ------------------------------------------------------------------ */
const char* ph3_last_error(void);
int ph3_version(void);

/* ---- device query: replaces the libcuda probe registry.cuda_available()
 *      (src/pyheat3d/registry.py:164-176) and NVML machine_info (benchmark.py:103-117) ---- */
int ph3_device_count(int* n);
int ph3_device_info(int dev, char* name, int name_len, char* arch, int arch_len, int* cu_count,
                    size_t* total_mem, int* wave_size, int* lds_per_cu, int* mem_clock_khz,
                    int* bus_width_bits);
int ph3_set_device(int dev);
int ph3_device_sync(void);
int ph3_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ---- memory and copies: replace cp.asarray / dev.copy() / cp.asnumpy
 *      (backends/cupy_backend.py:129-135, 196-200) ------------------------------------------ */
int ph3_malloc(void** ptr, size_t bytes);
int ph3_free(void* ptr);
int ph3_host_alloc(void** ptr, size_t bytes);      /* pinned host memory */
int ph3_host_free(void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device. ph3_memcpy returns after the copy
 * has completed on the device (safe to follow with launches on any stream); the _async form is
 * ordered on `stream` only. */
int ph3_memcpy(void* dst, const void* src, size_t bytes, int kind);
int ph3_memcpy_async(void* dst, const void* src, size_t bytes, int kind, void* stream);

/* ---- streams and events: replace the default stream + deviceSynchronize
 *      (cupy_backend.py:202-203) and the CUDA events of roofline.py:20-26 -------------------- */
int ph3_stream_create(void** stream);
int ph3_stream_destroy(void* stream);
int ph3_stream_sync(void* stream);
int ph3_event_create(void** event);
int ph3_event_destroy(void* event);
int ph3_event_record(void* event, void* stream);
int ph3_event_sync(void* event);
int ph3_event_elapsed_ms(float* ms, void* start, void* stop);

/* ---- tile configuration (replaces the module-level tile dicts NAIVE*_BLOCK, ZMARCH, TB3D,
 *      TB2D of cupy_backend.py:49-70 that bench/tune.py mutates) ----------------------------- */
typedef struct ph3_tile {
    int ty;    /* zmarch/tb: rows (y) per wave; 0 = default */
    int zch;   /* planes (3D) or rows (2D) marched per workgroup; 0 = default (auto) */
    int vec;   /* sites per lane along x (naive: strided, zmarch: contiguous); 0 = default */
    int waves; /* waves per workgroup (zmarch/tb); 0 = default */
    int flags; /* bit 0: non-temporal stores of the output; other bits must be 0 */
} ph3_tile;

#define PH3_TILE_NT_STORE 1

/* ---- the sweep: replaces CupyStepper._launch / step (cupy_backend.py:183-194) with the
 *      kernels of stencil.cu. Advances u by k steps (kind PH3_TB) or 1 step (naive/zmarch)
 *      into o. u and o must not overlap. b (float32, same shape, may be NULL) is added after
 *      c0*u + c1*S in every step (the reference's NumPy source term, numpy_ref.py:44-45).
 *      Dirichlet: the boundary layer of o is never written (callers keep it equal in both
 *      ping-pong buffers). nz is ignored when dims == 2. ---------------------------------- */
int ph3_sweep(int kind, int k, int storage, int periodic, int dims, int nz, int ny, int nx,
              float c0, float c1, const void* u, void* o, const float* b,
              const ph3_tile* tile, void* stream);

/* The launch ph3_sweep would make for these arguments, without launching: the lane width,
 * rows, waves and march chunk it resolves (tuned tile, buffer alignment, occupancy), the tile
 * geometry and the workgroup count. This is the record behind the reference's `tile` field of
 * temporal-blocking results (benchmark.py:222-231), which it reads from its tile dicts. `u`,
 * `o` and `b` are the buffers the sweep would use (their alignment picks the lane width);
 * nothing is read or written through them. */
typedef struct ph3_plan {
    int kind;            /* PH3_NAIVE, PH3_ZMARCH or PH3_TB, as launched (k = 1 runs the march) */
    int k;               /* steps per launch */
    int vec, rows, waves, zch;   /* sites per lane, rows per wave, waves per workgroup, march
                                    chunk (planes in 3D, rows in 2D); 0 where not used */
    int tile_x, tile_y;  /* extended tile: sites along x, rows along y (1 for a 2D strip) */
    int out_x, out_y;    /* of which written: sites along x, rows along y */
    int threads;         /* lanes per workgroup */
    long long groups;    /* workgroups of the launch */
} ph3_plan;
int ph3_sweep_plan(int kind, int k, int storage, int periodic, int dims, int nz, int ny, int nx,
                   const void* u, void* o, const float* b, const ph3_tile* tile, ph3_plan* plan);

/* Steps per pass supported by kind PH3_TB for (dims, storage); nonzero if supported. */
int ph3_tb_supported(int dims, int k);

/* ---- storage codec on the device: replaces to_bf16_bits / astype on the host
 *      (cupy_backend.py:103-112, 138-142, 196-200) -------------------------------------------- */
int ph3_encode(int storage, const float* src, void* dst, size_t n, void* stream);
int ph3_decode(int storage, const void* src, float* dst, size_t n, void* stream);

/* ---- synthetic initial states on the device (no counterpart: the reference draws fields on
 *      the host with NumPy, benchmark.py:276-280). Site i of stream `seed` gets
 *      (splitmix64(seed * 0xD1B54A32D192ED03 + i) >> 40) * 2^-24, rounded to storage.
 *      kind 0: random inside, 0 on the outer layer; 1: random outer layer, 0 inside (Laplace
 *      boundary data); 2: random everywhere. ---------------------------------------------- */
int ph3_fill_random(int storage, void* dst, int dims, int nz, int ny, int nx,
                    unsigned long long seed, int kind, void* stream);

/* ---- measurement: replaces roofline.copy_bandwidth's torch copy_ (roofline.py:8-29) ---- */
int ph3_copy_kernel(const void* src, void* dst, size_t bytes, void* stream);

/* ---- FP64 Jacobi residual (no counterpart: the reference runs fixed sweep counts,
 *      validation.py:89-110). Writes sum over updated sites of (b + c0*u + c1*S(u) - u)^2,
 *      each value widened to double, into *out (device pointer). work must hold
 *      ph3_residual_work_bytes() bytes of device memory. Deterministic: the summation order
 *      depends only on the shape. ------------------------------------------------------- */
size_t ph3_residual_work_bytes(void);
int ph3_residual(int storage, int periodic, int dims, int nz, int ny, int nx, float c0,
                 float c1, const void* u, const float* b, double* out, void* work,
                 void* stream);

/* ---- the FP64 residual without any allocation. No counterpart in the reference: this entry
 *      point is for callers that cannot allocate device memory (code under HIP graph capture,
 *      tight serving loops, harnesses that forbid hipMalloc). Same arguments, kernels and result
 *      bits as ph3_residual, minus `work`: the pass-1 partials live in a static device array of
 *      ph3_residual_noalloc_scratch_bytes() bytes (8 KiB, one FP64 partial per pass-1
 *      workgroup; every grid shape fits, since that count is fixed) that each device holds
 *      until the process ends. The first call on a device looks up that device's copy. Calls
 *      never allocate; calls that overlap on one device race on the shared partials, so order
 *      them on one stream (or synchronise between streams). ------------------------------ */
size_t ph3_residual_noalloc_scratch_bytes(void);
int ph3_residual_noalloc(int storage, int periodic, int dims, int nz, int ny, int nx, float c0,
                         float c1, const void* u, const float* b, double* out, void* stream);

/* ---- HIP graphs: capture whatever the calling thread enqueues on `stream` between begin and
 *      end into an executable graph (no counterpart in the reference; its nearest analogue is
 *      Warp's pre-recorded launches, backends/warp_backend.py:144-153). ---------------------- */
int ph3_graph_begin(void* stream);
int ph3_graph_end(void* stream, void** graph_exec);
int ph3_graph_launch(void* graph_exec, void* stream);
int ph3_graph_destroy(void* graph_exec);

/* ---- batched solves: many independent uneven 2D/3D grids advanced together (no counterpart
 *      in the reference, which times one grid at a time: benchmark.py:276-304).
 *
 * A solve is either a fixed number of steps (mode 0, heat) or Jacobi sweeps until the FP64
 * residual falls below tol relative to the initial one (mode 1), checked every
 * PH3_BATCH_CHECK sweeps. The state lives in buf[(done / kpass) & 1]. Every decision (tier,
 * tile shape, check points, reduction order) depends only on the solve itself, so a solve
 * gives bit-identical results alone or in any batch.
 *
 * Tier 1 (grids of at most PH3_BATCH_SMALL_SITES sites: 64^2, 16^3): one workgroup owns the
 * grid and runs its sweeps back to back with workgroup barriers; a persistent kernel hands
 * grids to workgroups through an atomic counter. Larger grids advance faster as tier-2 tiles
 * (a 181^2 grid: ~40 us per sweep on one workgroup). Tier 2 (larger grids), per class of solves
 * (storage, dims, boundary, source, steps per pass, tile shape): each pass is one varlen launch
 * over the tiles of every active grid of the class (one workgroup per tile, or one wave per 2D
 * strip; a per-class tile prefix table), followed at each check point by a deterministic
 * per-grid residual.
 * ------------------------------------------------------------------------------------------ */
#define PH3_BATCH_CHECK 64
#define PH3_BATCH_SMALL_SITES 4096
#define PH3_STALL_RING 32

typedef struct ph3_solve_desc {
    /* problem: set by the caller */
    int dims, nz, ny, nx;        /* nz = 1 for 2D */
    int periodic;
    int storage;                 /* 0 float32, 1 float16, 2 bfloat16 */
    float c0, c1;
    void* buf[2];                /* ping-pong device buffers; the initial state is in buf[0] */
    const float* src;            /* optional float32 source b, or NULL */
    int mode;                    /* 0: run `steps` steps; 1: Jacobi to `tol` */
    int stall_window;            /* mode 1: stagnation window in checks (0 = off, max 31) */
    int kpass;                   /* steps fused per tier-2 pass: 1, or for float32 storage 2 / 4
                                    / 8 dividing PH3_BATCH_CHECK and `steps` (0 means 1). The
                                    state lives in buf[(done / kpass) & 1]. */
    int tile;                    /* tier-2 tile shape of fused passes: 0 = the library's rule;
                                    3D 1: 1-site lanes, 32 rows; 2: 2-site lanes, 32 rows;
                                    3: 1-site lanes, 64 rows; 4: 2-site lanes, 32 rows in 8
                                    waves; 5/6: shapes 1/3 with buffer stores on the stored
                                    planes only. 2D 1/2: 1/2-site lanes; 3/4: the same with a store
                                    on every step; 5: shape 4 with each site loaded and stored
                                    on its own (any width). Shapes 2 and 4 need an even width
                                    above 64 and 8-byte aligned buffers (else 2D shape 4 runs as
                                    5, the others as the 1-site shape). Never changes the
                                    result. */
    long long steps;             /* mode 0: steps; mode 1: sweep cap */
    double tol;                  /* mode 1: target ||r|| / ||r0|| */
    double log_rate;             /* mode 1: ln(Jacobi spectral radius) * PH3_BATCH_CHECK (< 0) */
    double stall_fraction;       /* mode 1: required fraction of the predicted decrease */
    /* state: zero-initialise; maintained by the library */
    long long done;              /* steps / sweeps done */
    double res0_sq;              /* FP64 squared residual of the initial state (mode 1) */
    double res_sq;               /* latest FP64 squared residual (mode 1) */
    int status;                  /* 0 running, 1 finished / converged, 2 stagnated, 3 capped */
    int checks;                  /* residual checks done */
    double ring[PH3_STALL_RING]; /* past residuals (squared) for the stagnation test */
} ph3_solve_desc;

typedef struct ph3_batch ph3_batch;

size_t ph3_solve_desc_bytes(void);  /* sizeof(ph3_solve_desc), to check foreign mirrors */
int ph3_batch_create(ph3_batch** batch, int capacity);
int ph3_batch_destroy(ph3_batch* batch);
/* March-chunk policy of the tier-2 classes of one (dims, kpass), taking effect at the next
 * ph3_batch_set: a class's grids are cut into about target_tiles tiles together (each grid into
 * at most 512), with chunks of at most max_chunk planes / rows. 0 restores the default (512
 * tiles per grid, chunks up to 64, or 128 for 8-step passes). Never changes the result (the
 * autotuner sets it: the reference's bench/tune.py:21-95 tunes single-grid tiles only). */
int ph3_batch_params(ph3_batch* batch, int dims, int kpass, int target_tiles, int max_chunk);
/* Replace the batch's solve table with descs[0..n) (host memory) and compute the initial
 * residual of every mode-1 solve whose `checks` is 0. Synchronous. */
int ph3_batch_set(ph3_batch* batch, const ph3_solve_desc* descs, int n, void* stream);
/* Enqueue one round: every running solve advances by up to `sweeps` steps (rounded up to a
 * multiple of PH3_BATCH_CHECK for mode 1), stopping early when finished. Asynchronous and
 * capturable into a graph (the launch shapes depend only on the table set last). Tier-1 work,
 * and the round of every tier-2 class after the first (its passes, residuals and checks), run
 * on streams of the batch, forked from and joined back to `stream` with events once per round,
 * so all of it is ordered on `stream` as seen by the caller. */
int ph3_batch_round(ph3_batch* batch, int sweeps, void* stream);
/* Copy the solve table back to descs[0..n) (host). Synchronous on `stream`. */
int ph3_batch_get(ph3_batch* batch, ph3_solve_desc* descs, int n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYHEAT3D_HIP_H */
