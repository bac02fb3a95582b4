/*
 * stencil_rt.h -- C ABI of the MI355X stencil runtime (libstencil_rt.so).
 *
 * The upstream project has no native library: its emitted Python module
 * calls cuTile/CuPy directly.  This ABI is the device boundary that the
 * emitted launcher module (and any other host language) binds instead:
 *
 *   upstream call site                               -> entry point here
 *   ct.launch(stream, grid, kernel, args)            -> st_launch
 *     (cutile/lowering/target_to_python.py:381-398, the body of every
 *      generated launch_<name>)
 *   cuTile JIT on first launch per constant set      -> st_compile / st_get_function
 *     (target_to_python.py:305-321; README.md:487-491)
 *   apply_boundary_<name>(u, stream) CuPy copies      -> st_fill_halo
 *     (target_to_python.py:199-297; semantics stencil_ref.py:148-201)
 *   regcheck.probe_resources (cuobjdump -res-usage)   -> st_function_resources
 *     (cutile/runtime/regcheck.py:136-212)
 *   cp.cuda.Event timing in benchmark()/autotune()    -> st_event_*
 *     (runtime/launcher.py:208-216, runtime/autotune.py:187-194)
 *   STREAM-copy peak in run_benchmarks.py:38-59       -> st_copy, st_copy_wide
 *   to_bricks / from_bricks (multigpu_emitter.py:370-434) -> st_bricks
 *   max-norm comparisons in validate()               -> st_diff_norms
 *     (runtime/launcher.py:127-131)
 *   (no upstream counterpart: FP64 residual norms, field checksums)
 *                                                    -> st_residual_norms(_static), st_field_digest
 *   (no upstream counterpart: hipGraph capture of step loops)
 *                                                    -> st_graph_*
 *   (no upstream counterpart: the serving loop's completion polls)
 *                                                    -> st_event_query
 *
 * Conventions: every function returns 0 on success and a non-zero code
 * (a hipError_t, or a negative ST_E* code) on failure; st_last_error()
 * returns a thread-local message for the last failure.  Nothing throws or
 * aborts across the ABI.  Device memory is owned by the caller; streams are
 * plain hipStream_t handles passed as void* (NULL = the null stream).
 */
#ifndef STENCIL_RT_H
#define STENCIL_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ST_OK = 0,
  ST_E_INVALID = -1,    /* bad argument */
  ST_E_COMPILE = -2,    /* hiprtc compilation failed (see st_last_error) */
  ST_E_NOTFOUND = -3,   /* kernel name not in module */
  ST_E_NODEVICE = -4,   /* no HIP device */
};

/* element types understood by the fixed kernels */
enum { ST_F64 = 0, ST_F32 = 1, ST_BF16 = 2 };

/* boundary kinds (stencil_ref.py:148-201) */
enum { ST_BC_NONE = 0, ST_BC_DIRICHLET = 1, ST_BC_NEUMANN = 2, ST_BC_PERIODIC = 3, ST_BC_REFLECTING = 4 };

typedef struct st_device_info {
  char name[256];
  char arch[64];          /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int compute_units;
  int max_threads_per_block;
  size_t lds_per_block;   /* bytes */
  size_t total_mem;       /* bytes */
  size_t free_mem;        /* bytes */
  int clock_khz;
  int mem_clock_khz;
  int mem_bus_width;
  int warp_size;
} st_device_info;

typedef struct st_resources {
  int num_regs;            /* VGPRs per thread */
  int local_bytes;         /* private (scratch) bytes per thread; > 0 means spills */
  int shared_bytes;        /* static LDS bytes per workgroup */
  int max_threads;         /* max threads per block for this kernel */
  int const_bytes;
} st_resources;

/* Thread-local description of the last error. */
const char* st_last_error(void);

/* Select device `dev` (hipSetDevice). */
int st_init(int dev);
int st_device_count(int* count);
int st_device_info_get(int dev, st_device_info* out);
int st_mem_info(size_t* free_bytes, size_t* total_bytes);

/* Compile HIP source with hiprtc (no device needed).  `opts` are hiprtc
 * options (e.g. "--offload-arch=gfx950", "-I<dir>").  On success the code
 * object is written to *code (malloc'ed; release with st_free) and its size
 * to *size.  The compiler log is available through st_last_error(). */
int st_compile_only(const char* src, const char* name, const char* const* opts, int nopts,
                    void** code, size_t* size);
void st_free(void* p);

/* Compile (or fetch from the in-process / on-disk cache under `cache_dir`,
 * may be NULL) and load a module on the current device.  The handle stays
 * valid until st_unload.  The cache key is the FNV-1a hash of the source and
 * of every option except "-I..." include paths, each option followed by the
 * byte 0x1f; the file is `<cache_dir>/<16 hex digits>.hsaco`. */
int st_compile(const char* src, const char* name, const char* const* opts, int nopts,
               const char* cache_dir, void** module);
int st_load_code(const void* code, size_t size, void** module);
int st_unload(void* module);
int st_get_function(void* module, const char* kernel, void** function);

/* Launch `function` with one by-value argument blob of `args_size` bytes. */
int st_launch(void* function, unsigned gx, unsigned gy, unsigned gz,
              unsigned bx, unsigned by, unsigned bz, unsigned dyn_lds,
              void* stream, const void* args, size_t args_size);

/* Register/scratch/LDS usage of a loaded kernel. */
int st_function_resources(void* function, st_resources* out);

/* Halo fill of one array (C-contiguous, element type `dtype`), dims in order
 * 0..ndim-1, low side then high side, each dim a separate launch so corner
 * cells see earlier dims' fills.  kinds/values hold 2*ndim entries
 * (low0, high0, low1, high1, ...). */
int st_fill_halo(void* ptr, int dtype, int ndim, const int64_t* shape, const int64_t* halo,
                 const int* kinds, const double* values, void* stream);

/* One halo-fill descriptor of a batch (device-resident array of these). */
typedef struct st_fill_desc {
  void* ptr;
  int64_t shape[4];
  int64_t halo[4];
  int kinds[8];       /* low0, high0, low1, high1, ... */
  double values[8];
  int ndim;
  int dtype;
  int64_t pitch;      /* row stride (elements) of the last axis; 0 = shape[ndim-1] */
} st_fill_desc;

/* Halo fill of `count` arrays in one launch per dimension (serving and
 * batched stepping; no upstream counterpart).  `descs` is DEVICE memory
 * holding `count` st_fill_desc; every array must have an interior at least
 * as wide as its halo in each dimension.  `max_slab` bounds the elements of
 * any one array's two slabs in any dimension (sizes the grid). */
int st_fill_halo_batched(const void* descs, int count, int ndim, int64_t max_slab, void* stream);

/* dst[i] = src[i] for nbytes bytes (16-byte vector copy kernel). */
int st_copy(void* dst, const void* src, size_t nbytes, void* stream);

/* The same copy with `vectors` (1, 2, 4 or 8) 16-byte vectors per thread,
 * every load issued before the stores, non-temporal loads and stores: a
 * second copy-bandwidth reference beside st_copy and hipMemcpyAsync.
 * Pointers and nbytes must be multiples of 16. */
int st_copy_wide(void* dst, const void* src, size_t nbytes, int vectors, void* stream);

/* Max-norms over the interior of two arrays of the same shape:
 * out[0] = max|a-b|, out[1] = max|b| (computed in FP64). `out` is host memory;
 * the call synchronizes the stream.  Replaces the max-norm comparison of
 * validate() (runtime/launcher.py:127-131).  The reduction runs in two
 * passes (partials of at most 65536 cells, then one workgroup) in a
 * workspace the library allocates once per device and grows on demand. */
int st_diff_norms(const void* a, int dtype_a, const void* b, int dtype_b, int ndim,
                  const int64_t* shape, const int64_t* halo, void* stream, double* out);

/* Residual norms of a time step (no upstream counterpart; the stepping API's
 * FP64 residuals): over the interior, out[0] = max|a-b|, out[1] = max|b|,
 * out[2] = ||a-b||_2, out[3] = ||b||_2, all accumulated in FP64 with a
 * summation order fixed by the shape (reproducible bit for bit).  NaN in
 * either array gives NaN in every output.  Synchronizes the stream. */
int st_residual_norms(const void* a, int dtype_a, const void* b, int dtype_b, int ndim,
                      const int64_t* shape, const int64_t* halo, void* stream, double* out);

/* st_residual_norms without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The partials live in a fixed
 * array in this library's device image (16385 x 40 B = 640 KiB per GPU,
 * held until the process ends): interiors needing more than 16384 partials
 * of 65536 cells (about 1.07e9 cells) are rejected with ST_E_INVALID.
 * Every call on a GPU shares that GPU's array, so calls that overlap on one
 * GPU (other threads, other streams) race; serialize them.  Same outputs,
 * bit for bit, as st_residual_norms. */
int st_residual_norms_static(const void* a, int dtype_a, const void* b, int dtype_b, int ndim,
                             const int64_t* shape, const int64_t* halo, void* stream, double* out);

/* Checksum of an array's interior (no upstream counterpart; pins device
 * states against the oracle at sizes where copying them out is too slow):
 * the sum modulo 2^64 over interior cells e (row-major interior index) of
 * mix64(bits(x_e) XOR e * 0x9e3779b97f4a7c15), where bits are the stored
 * bits of the element (dtype ST_F64 / ST_F32 / ST_BF16, zero-extended) and
 * mix64 is the SplitMix64 finaliser.  Synchronizes the stream. */
int st_field_digest(const void* a, int dtype, int ndim, const int64_t* shape, const int64_t* halo, void* stream,
                    uint64_t* out);

/* Flat <-> bricked layout (to_bricks / from_bricks): `flat` has `shape`,
 * `bricked` is the brick-major array of shape (nb0, .., nb_{d-1}, bs, .., bs)
 * with nb_i = ceil(shape_i / bs); padding cells are zero on the way in. */
int st_bricks(void* flat, void* bricked, int dtype, int ndim, const int64_t* shape, int bs,
              int to_bricks, void* stream);

/* Events (hipEvent) for timing on an arbitrary stream. */
int st_event_create(void** ev);
int st_event_record(void* ev, void* stream);
int st_event_elapsed_ms(void* start, void* stop, float* ms);
int st_event_destroy(void* ev);
int st_event_sync(void* ev);         /* wait until the event has completed */
int st_event_query(void* ev, int* done);   /* *done = 1 if the event has completed, else 0; never waits */
int st_stream_sync(void* stream);
int st_device_sync(void);

/* hipGraph capture of everything launched on `stream` between begin/end. */
int st_graph_begin(void* stream);
int st_graph_end(void* stream, void** graph_exec);
int st_graph_launch(void* graph_exec, void* stream);
int st_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif
#endif
