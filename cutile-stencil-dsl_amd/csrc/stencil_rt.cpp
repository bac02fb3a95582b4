// libstencil_rt.so -- C-ABI runtime of the MI355X stencil backend.
//
// Host side: hiprtc compilation with an in-process and on-disk code-object
// cache, module/function handles, by-value-argument launches, events,
// hipGraph capture, device queries and kernel resource queries.
// Device side (compiled ahead of time for gfx950): halo fill, vector copy,
// interior max-norm differences and flat<->brick layout conversion.
// See include/stencil_rt.h for the contract of every entry point.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <unistd.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/stencil_rt.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                             \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail((int)e_, "%s failed: %s", #expr, hipGetErrorString(e_));       \
  } while (0)

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 14695981039346656037ull) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

std::mutex g_mu;
std::map<std::string, hipModule_t> g_modules;  // key: device + hash

int do_compile(const char* src, const char* name, const char* const* opts, int nopts,
               std::vector<char>& code) {
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, src, name ? name : "stencil.hip", 0, nullptr, nullptr);
  if (r != HIPRTC_SUCCESS) return fail(ST_E_COMPILE, "hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
  r = hiprtcCompileProgram(prog, nopts, const_cast<const char**>(opts));
  size_t log_size = 0;
  hiprtcGetProgramLogSize(prog, &log_size);
  std::string log(log_size + 1, '\0');
  if (log_size) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    g_err = std::string("hiprtc compile failed: ") + hiprtcGetErrorString(r) + "\n" + log;
    return ST_E_COMPILE;
  }
  size_t size = 0;
  hiprtcGetCodeSize(prog, &size);
  code.resize(size);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  g_err = log;
  return ST_OK;
}

// ---------------------------------------------------------------------------
// fixed device kernels
// ---------------------------------------------------------------------------

template <class E>
__global__ void k_fill_side(E* u, int64_t outer, int64_t n, int64_t inner, int64_t h, int kind,
                            E value, int side_mask, int64_t only_layer) {
  // element e of the (side, k, o, i) slab space
  const int64_t per_side = outer * h * inner;
  const int64_t total = per_side * 2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int side = (int)(e / per_side);
    if (!((side_mask >> side) & 1)) continue;
    int64_t r = e % per_side;
    const int64_t o = r / (h * inner);
    r %= h * inner;
    const int64_t k = r / inner;
    const int64_t i = r % inner;
    if (only_layer >= 0 && k != only_layer) continue;
    int64_t dst, src = -1;
    if (side == 0) {
      dst = k;
      if (kind == ST_BC_NEUMANN) src = h;
      else if (kind == ST_BC_PERIODIC) src = n - 2 * h + k;
      else if (kind == ST_BC_REFLECTING) src = 2 * h - 1 - k;
    } else {
      dst = n - h + k;
      if (kind == ST_BC_NEUMANN) src = n - h - 1;
      else if (kind == ST_BC_PERIODIC) src = h + k;
      else if (kind == ST_BC_REFLECTING) src = n - h - 1 - k;
    }
    E* base = u + o * n * inner + i;
    base[dst * inner] = (kind == ST_BC_DIRICHLET) ? value : base[src * inner];
  }
}

// One element of the two halo slabs of dimension `dim` of a (possibly
// row-padded) array: e enumerates (side, layer k, other indices) with the
// other indices in row-major order over the logical shape.
template <class E>
__device__ inline void fill_one(E* u, const st_fill_desc& d, int dim, int64_t e) {
  const int64_t h = d.halo[dim], n = d.shape[dim];
  int64_t stride[4];
  stride[d.ndim - 1] = 1;
  const int64_t rowlen = d.pitch > 0 ? d.pitch : d.shape[d.ndim - 1];
  for (int k = d.ndim - 2; k >= 0; --k) stride[k] = stride[k + 1] * (k + 1 == d.ndim - 1 ? rowlen : d.shape[k + 1]);
  // decode: other dims (row-major), then layer k, then side
  int64_t r = e, off = 0;
  for (int k = d.ndim - 1; k >= 0; --k) {
    if (k == dim) continue;
    off += (r % d.shape[k]) * stride[k];
    r /= d.shape[k];
  }
  const int64_t layer = r % h;
  const int side = (int)(r / h);
  const int kind = d.kinds[2 * dim + side];
  if (kind == ST_BC_NONE) return;
  int64_t dst, src;
  if (side == 0) {
    dst = layer;
    src = kind == ST_BC_NEUMANN ? h : kind == ST_BC_PERIODIC ? n - 2 * h + layer : 2 * h - 1 - layer;
  } else {
    dst = n - h + layer;
    src = kind == ST_BC_NEUMANN ? n - h - 1 : kind == ST_BC_PERIODIC ? h + layer : n - h - 1 - layer;
  }
  E* base = u + off;
  if (kind == ST_BC_DIRICHLET) {
    const double v = d.values[2 * dim + side];
    E ev;
    if constexpr (sizeof(E) == 2) ev = (E)(unsigned short)v;  // pre-rounded bf16 bits
    else ev = (E)v;
    base[dst * stride[dim]] = ev;
  } else {
    base[dst * stride[dim]] = base[src * stride[dim]];
  }
}

__global__ void k_fill_batched(const st_fill_desc* descs, int dim) {
  const st_fill_desc& d = descs[blockIdx.y];
  if (dim >= d.ndim || d.halo[dim] == 0) return;
  int64_t total = 2 * d.halo[dim];
  for (int k = 0; k < d.ndim; ++k)
    if (k != dim) total *= d.shape[k];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (d.dtype == ST_F64) fill_one<double>((double*)d.ptr, d, dim, e);
    else if (d.dtype == ST_F32) fill_one<float>((float*)d.ptr, d, dim, e);
    else fill_one<unsigned short>((unsigned short*)d.ptr, d, dim, e);
  }
}

// 16-byte vectors, one per thread, one short-lived workgroup per 4 KiB and
// non-temporal stores: the copy-bandwidth reference kernel.  On MI355X this
// full-grid shape copies at about 6.3 TB/s where grid-stride loops with
// several vectors in flight per thread stay near 5 TB/s.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy16(u32x4* dst, const u32x4* src, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    __builtin_nontemporal_store(src[i], dst + i);
}
// V 16-byte vectors per thread, all loads issued before the stores: a copy
// with V-fold more reads in flight per wave than k_copy16 (the ceiling is
// reported for both shapes and for hipMemcpyAsync).  Vector k of a
// workgroup's block is k*256 + tid, so every load instruction of a wave is
// one contiguous 1 KiB span.
template <int V>
__global__ void __launch_bounds__(256) k_copy16_wide(u32x4* dst, const u32x4* src, size_t n16) {
  const size_t base = (size_t)blockIdx.x * 256 * V + threadIdx.x;
  u32x4 r[V];
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const size_t i = base + (size_t)k * 256;
    if (i < n16) r[k] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const size_t i = base + (size_t)k * 256;
    if (i < n16) __builtin_nontemporal_store(r[k], dst + i);
  }
}
__global__ void k_copy1(unsigned char* dst, const unsigned char* src, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__device__ inline double load_as_f64(const void* p, int dtype, int64_t i) {
  if (dtype == ST_F64) return static_cast<const double*>(p)[i];
  if (dtype == ST_F32) return (double)static_cast<const float*>(p)[i];
  unsigned short b = static_cast<const unsigned short*>(p)[i];
  return (double)__builtin_bit_cast(float, ((unsigned)b) << 16);
}

struct Shape4 {
  int64_t s[4];
  int64_t h[4];
  int64_t n[4];
};

// ---------------------------------------------------------------------------
// interior reductions (residual norms, field digest): two passes, no atomics
// ---------------------------------------------------------------------------
//
// The interior (padded to 4-D) is walked as rows of n[3] cells.  Partial p
// covers rows [p*rpp, (p+1)*rpp) with rpp = max(1, RED_CELLS / n[3]); pass 1
// reduces one partial per workgroup into parts[p], pass 2 reduces the
// partials in one workgroup.  The order of every reduction is a function of
// the shape alone, so the FP64 sums are reproducible bit for bit.
constexpr int64_t RED_CELLS = 65536;
constexpr int RED_THREADS = 256;

struct Interior {
  int64_t s[4], h[4], n[4];
  __host__ __device__ int64_t rows() const { return n[0] * n[1] * n[2]; }
  __device__ int64_t row_base(int64_t r) const {
    const int64_t c2 = r % n[2], r1 = r / n[2];
    const int64_t c1 = r1 % n[1], c0 = r1 / n[1];
    return (((c0 + h[0]) * s[1] + (c1 + h[1])) * s[2] + (c2 + h[2])) * s[3] + h[3];
  }
};

int64_t red_rows_per_part(const Interior& in) {
  const int64_t rpp = in.n[3] > 0 ? RED_CELLS / in.n[3] : 1;
  return rpp > 0 ? rpp : 1;
}
int64_t red_parts(const Interior& in) {
  const int64_t rows = in.rows();
  if (rows <= 0 || in.n[3] <= 0) return 0;
  const int64_t rpp = red_rows_per_part(in);
  return (rows + rpp - 1) / rpp;
}

// Norms of a - b: max|a-b|, max|b|, sum (a-b)^2, sum b^2 (FP64), NaN seen.
struct NormAcc {
  double v[5];
};
struct NormOp {
  typedef NormAcc Acc;
  const void* a;
  const void* b;
  int da, db;
  static __device__ Acc zero() { return Acc{{0.0, 0.0, 0.0, 0.0, 0.0}}; }
  __device__ void cell(Acc& acc, int64_t off, int64_t) const {
    const double x = load_as_f64(a, da, off), y = load_as_f64(b, db, off);
    const double d = x - y;
    const double ad = fabs(d), ay = fabs(y);
    if (ad != ad || ay != ay) acc.v[4] = 1.0;
    acc.v[0] = ad > acc.v[0] ? ad : acc.v[0];
    acc.v[1] = ay > acc.v[1] ? ay : acc.v[1];
    acc.v[2] += d * d;
    acc.v[3] += y * y;
  }
  static __device__ Acc join(const Acc& p, const Acc& q) {
    Acc r;
    r.v[0] = q.v[0] > p.v[0] ? q.v[0] : p.v[0];
    r.v[1] = q.v[1] > p.v[1] ? q.v[1] : p.v[1];
    r.v[2] = p.v[2] + q.v[2];
    r.v[3] = p.v[3] + q.v[3];
    r.v[4] = (p.v[4] != 0.0 || q.v[4] != 0.0) ? 1.0 : 0.0;
    return r;
  }
  static __device__ Acc shfl_down(const Acc& p, int off) {
    Acc r;
#pragma unroll
    for (int k = 0; k < 5; ++k) r.v[k] = __shfl_down(p.v[k], off);
    return r;
  }
};

// Position-dependent checksum of the stored bits: the sum, modulo 2^64, of
// mix64(bits(x_e) ^ e * 0x9e3779b97f4a7c15) over interior cells e (row-major
// interior index).  mix64 is the finaliser of Steele, Lea and Flood's
// SplitMix64 ("Fast splittable pseudorandom number generators", OOPSLA 2014).
__device__ inline uint64_t mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xbf58476d1ce4e5b9ull;
  z ^= z >> 27;
  z *= 0x94d049bb133111ebull;
  z ^= z >> 31;
  return z;
}
struct DigestOp {
  typedef unsigned long long Acc;
  const void* a;
  int da;
  static __device__ Acc zero() { return 0ull; }
  __device__ void cell(Acc& acc, int64_t off, int64_t e) const {
    uint64_t bits;
    if (da == ST_F64) bits = static_cast<const uint64_t*>(a)[off];
    else if (da == ST_F32) bits = static_cast<const uint32_t*>(a)[off];
    else bits = static_cast<const uint16_t*>(a)[off];
    acc += mix64(bits ^ ((uint64_t)e * 0x9e3779b97f4a7c15ull));
  }
  static __device__ Acc join(Acc p, Acc q) { return p + q; }
  static __device__ Acc shfl_down(Acc p, int off) { return __shfl_down(p, off); }
};

template <class Op>
__device__ typename Op::Acc block_reduce(typename Op::Acc acc) {
  __shared__ typename Op::Acc part[RED_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) acc = Op::join(acc, Op::shfl_down(acc, off));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  acc = part[0];
  for (int w = 1; w < RED_THREADS / 64; ++w) acc = Op::join(acc, part[w]);
  return acc;
}

template <class Op>
__global__ void __launch_bounds__(RED_THREADS) k_reduce_parts(Op op, Interior in, int64_t rpp,
                                                              typename Op::Acc* parts) {
  const int64_t p = blockIdx.x;
  const int64_t r0 = p * rpp;
  int64_t r1 = r0 + rpp;
  if (r1 > in.rows()) r1 = in.rows();
  typename Op::Acc acc = Op::zero();
  for (int64_t r = r0; r < r1; ++r) {
    const int64_t base = in.row_base(r);
    for (int64_t c = threadIdx.x; c < in.n[3]; c += RED_THREADS) op.cell(acc, base + c, r * in.n[3] + c);
  }
  acc = block_reduce<Op>(acc);
  if (threadIdx.x == 0) parts[p] = acc;
}

template <class Op>
__global__ void __launch_bounds__(RED_THREADS) k_reduce_final(const typename Op::Acc* parts, int64_t nparts,
                                                              typename Op::Acc* out) {
  typename Op::Acc acc = Op::zero();
  for (int64_t i = threadIdx.x; i < nparts; i += RED_THREADS) acc = Op::join(acc, parts[i]);
  acc = block_reduce<Op>(acc);
  if (threadIdx.x == 0) out[0] = acc;
}

// Enqueue both passes with the partials at `parts` and the result at `out`
// (device memory), then copy the result to `host` and wait for the stream.
template <class Op>
int run_reduction(const Op& op, const Interior& in, typename Op::Acc* parts, typename Op::Acc* out,
                  typename Op::Acc* host, hipStream_t s) {
  const int64_t np = red_parts(in);
  if (np == 0) {
    *host = typename Op::Acc{};
    return ST_OK;
  }
  if (np > 0x7fffffff) return fail(ST_E_INVALID, "interior too large for one reduction");
  hipLaunchKernelGGL(k_reduce_parts<Op>, dim3((unsigned)np), dim3(RED_THREADS), 0, s, op, in,
                     red_rows_per_part(in), parts);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final<Op>, dim3(1), dim3(RED_THREADS), 0, s, (const typename Op::Acc*)parts, np, out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host, out, sizeof(*host), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return ST_OK;
}

int make_interior(int ndim, const int64_t* shape, const int64_t* halo, Interior& in) {
  if (ndim < 1 || ndim > 4 || !shape || !halo) return fail(ST_E_INVALID, "bad interior description (ndim %d)", ndim);
  for (int d = 0; d < 4; ++d) {
    const int k = d - (4 - ndim);
    in.s[d] = k >= 0 ? shape[k] : 1;
    in.h[d] = k >= 0 ? halo[k] : 0;
    in.n[d] = in.s[d] - 2 * in.h[d];
    if (in.h[d] < 0 || in.n[d] < 0) return fail(ST_E_INVALID, "dim %d: negative halo or interior", k);
  }
  return ST_OK;
}

// Library-held reduction workspace: one buffer per device, grown on demand
// (hipMalloc; kept until the process ends).  Calls on one device from
// several threads take the lock for the whole enqueue-and-wait.
struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<int, Workspace> g_ws;

int workspace(size_t bytes, void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  Workspace& w = g_ws[dev];
  if (w.bytes < bytes) {
    if (w.ptr) HIPCHK(hipFree(w.ptr));
    w.ptr = nullptr;
    w.bytes = 0;
    size_t want = 4096;
    while (want < bytes) want *= 2;
    HIPCHK(hipMalloc(&w.ptr, want));
    w.bytes = want;
  }
  *out = w.ptr;
  return ST_OK;
}

// The allocation-free variant's scratch: a fixed array in this library's
// device image, one copy per GPU.  Host code reaches it only through the
// per-device address from hipGetSymbolAddress (taken in st_init, or on the
// first call on a device); the array's name is not a device address on the
// host.
constexpr int64_t STATIC_PARTS = 16384;
static __device__ __align__(256) NormAcc scratch[STATIC_PARTS + 1];
std::mutex g_scratch_mu;
std::map<int, NormAcc*> g_scratch;  // device -> address of `scratch` on it

int scratch_address(int dev, NormAcc** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it == g_scratch.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail((int)(e != hipSuccess ? e : hipErrorInvalidSymbol),
                  "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    it = g_scratch.emplace(dev, (NormAcc*)p).first;
  }
  *out = it->second;
  return ST_OK;
}

template <class E>
__global__ void k_bricks(E* flat, E* bricked, int ndim, Shape4 sh, int64_t nb0, int64_t nb1,
                         int64_t nb2, int bs, int to_bricks, int64_t total) {
  const int64_t nb[3] = {nb0, nb1, nb2};
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    // e indexes the bricked array: (b_0..b_{d-1}, o_0..o_{d-1})
    int64_t r = e;
    int64_t o[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    for (int d = ndim - 1; d >= 0; --d) {
      o[d] = r % bs;
      r /= bs;
    }
    for (int d = ndim - 1; d >= 0; --d) {
      b[d] = r % nb[d];
      r /= nb[d];
    }
    bool in = true;
    int64_t fi = 0;
    for (int d = 0; d < ndim; ++d) {
      const int64_t g = b[d] * bs + o[d];
      in = in && g < sh.s[d];
      fi = fi * sh.s[d] + g;
    }
    if (to_bricks) bricked[e] = in ? flat[fi] : E(0);
    else if (in) flat[fi] = bricked[e];
  }
}

unsigned grid_for(int64_t n, int bs = 256) {
  int64_t g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

size_t esize(int dtype) { return dtype == ST_F64 ? 8 : dtype == ST_F32 ? 4 : 2; }

}  // namespace

extern "C" {

const char* st_last_error(void) { return g_err.c_str(); }

int st_init(int dev) {
  HIPCHK(hipSetDevice(dev));
  NormAcc* p = nullptr;
  return scratch_address(dev, &p);
}

int st_device_count(int* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return ST_OK;
}

int st_device_info_get(int dev, st_device_info* out) {
  if (!out) return fail(ST_E_INVALID, "null output");
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  memset(out, 0, sizeof(*out));
  snprintf(out->name, sizeof(out->name), "%s", p.name);
  snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
  out->compute_units = p.multiProcessorCount;
  out->max_threads_per_block = p.maxThreadsPerBlock;
  out->lds_per_block = p.sharedMemPerBlock;
  out->total_mem = p.totalGlobalMem;
  out->clock_khz = p.clockRate;
  out->mem_clock_khz = p.memoryClockRate;
  out->mem_bus_width = p.memoryBusWidth;
  out->warp_size = p.warpSize;
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur == dev) {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) == hipSuccess) out->free_mem = f;
  }
  return ST_OK;
}

int st_mem_info(size_t* free_bytes, size_t* total_bytes) {
  HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
  return ST_OK;
}

int st_compile_only(const char* src, const char* name, const char* const* opts, int nopts,
                    void** code, size_t* size) {
  if (!src || !code || !size) return fail(ST_E_INVALID, "null argument");
  std::vector<char> buf;
  int rc = do_compile(src, name, opts, nopts, buf);
  if (rc) return rc;
  void* p = malloc(buf.size());
  if (!p) return fail(ST_E_INVALID, "out of host memory");
  memcpy(p, buf.data(), buf.size());
  *code = p;
  *size = buf.size();
  return ST_OK;
}

void st_free(void* p) { free(p); }

int st_load_code(const void* code, size_t size, void** module) {
  (void)size;
  hipModule_t m;
  HIPCHK(hipModuleLoadData(&m, code));
  *module = (void*)m;
  return ST_OK;
}

int st_compile(const char* src, const char* name, const char* const* opts, int nopts,
               const char* cache_dir, void** module) {
  if (!src || !module) return fail(ST_E_INVALID, "null argument");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  // key: the source and every option except include paths (the same tree
  // may sit at another path; header contents are keyed through a -D option)
  uint64_t h = fnv1a(src, strlen(src));
  for (int i = 0; i < nopts; ++i) {
    if (strncmp(opts[i], "-I", 2) == 0) continue;
    h = fnv1a(opts[i], strlen(opts[i]), h);
    h = fnv1a("\x1f", 1, h);
  }
  char key[64];
  snprintf(key, sizeof(key), "%016llx", (unsigned long long)h);
  const std::string mkey = std::to_string(dev) + ":" + key;
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_modules.find(mkey);
  if (it != g_modules.end()) {
    *module = (void*)it->second;
    return ST_OK;
  }
  std::vector<char> code;
  std::string path;
  if (cache_dir && *cache_dir) {
    path = std::string(cache_dir) + "/" + key + ".hsaco";
    FILE* f = fopen(path.c_str(), "rb");
    if (f) {
      fseek(f, 0, SEEK_END);
      long n = ftell(f);
      fseek(f, 0, SEEK_SET);
      if (n > 0) {
        code.resize((size_t)n);
        if (fread(code.data(), 1, (size_t)n, f) != (size_t)n) code.clear();
      }
      fclose(f);
    }
  }
  if (code.empty()) {
    int rc = do_compile(src, name, opts, nopts, code);
    if (rc) return rc;
    if (!path.empty()) {
      std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
      FILE* f = fopen(tmp.c_str(), "wb");
      if (f) {
        size_t w = fwrite(code.data(), 1, code.size(), f);
        fclose(f);
        if (w == code.size()) rename(tmp.c_str(), path.c_str());
        else remove(tmp.c_str());
      }
    }
  }
  hipModule_t m;
  HIPCHK(hipModuleLoadData(&m, code.data()));
  g_modules[mkey] = m;
  *module = (void*)m;
  return ST_OK;
}

int st_unload(void* module) {
  std::lock_guard<std::mutex> lock(g_mu);
  for (auto it = g_modules.begin(); it != g_modules.end(); ++it)
    if ((void*)it->second == module) {
      g_modules.erase(it);
      break;
    }
  HIPCHK(hipModuleUnload((hipModule_t)module));
  return ST_OK;
}

int st_get_function(void* module, const char* kernel, void** function) {
  hipFunction_t f;
  hipError_t e = hipModuleGetFunction(&f, (hipModule_t)module, kernel);
  if (e != hipSuccess) return fail(ST_E_NOTFOUND, "kernel '%s' not found: %s", kernel, hipGetErrorString(e));
  *function = (void*)f;
  return ST_OK;
}

int st_launch(void* function, unsigned gx, unsigned gy, unsigned gz, unsigned bx, unsigned by,
              unsigned bz, unsigned dyn_lds, void* stream, const void* args, size_t args_size) {
  if (!function) return fail(ST_E_INVALID, "null function");
  if (gx == 0 || gy == 0 || gz == 0) return ST_OK;  // empty launch: nothing to do
  size_t sz = args_size;
  void* config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void*>(args),
                    HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END};
  HIPCHK(hipModuleLaunchKernel((hipFunction_t)function, gx, gy, gz, bx, by, bz, dyn_lds,
                               (hipStream_t)stream, nullptr, config));
  return ST_OK;
}

int st_function_resources(void* function, st_resources* out) {
  if (!function || !out) return fail(ST_E_INVALID, "null argument");
  hipFunction_t f = (hipFunction_t)function;
  memset(out, 0, sizeof(*out));
  HIPCHK(hipFuncGetAttribute(&out->num_regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
  HIPCHK(hipFuncGetAttribute(&out->local_bytes, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f));
  HIPCHK(hipFuncGetAttribute(&out->shared_bytes, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f));
  HIPCHK(hipFuncGetAttribute(&out->max_threads, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, f));
  HIPCHK(hipFuncGetAttribute(&out->const_bytes, HIP_FUNC_ATTRIBUTE_CONST_SIZE_BYTES, f));
  return ST_OK;
}

int st_fill_halo(void* ptr, int dtype, int ndim, const int64_t* shape, const int64_t* halo,
                 const int* kinds, const double* values, void* stream) {
  if (!ptr || ndim < 1 || ndim > 4) return fail(ST_E_INVALID, "bad fill_halo arguments");
  hipStream_t s = (hipStream_t)stream;
  for (int d = 0; d < ndim; ++d) {
    const int64_t h = halo[d], n = shape[d];
    if (h == 0) continue;
    if (n < 2 * h) return fail(ST_E_INVALID, "dim %d: extent %lld < 2*halo", d, (long long)n);
    int64_t outer = 1, inner = 1;
    for (int k = 0; k < d; ++k) outer *= shape[k];
    for (int k = d + 1; k < ndim; ++k) inner *= shape[k];
    const int lk = kinds[2 * d], hk = kinds[2 * d + 1];
    // Sides of one dim read/write disjoint layers when the interior is at
    // least h wide; then one launch does both sides, else low then high.
    // An interior narrower than h lets a side read layers it writes: then
    // its layers go one launch each, in order, as the sequential definition.
    const bool narrow = (n - 2 * h) < h;
    const bool split = narrow || lk != hk || values[2 * d] != values[2 * d + 1];
    for (int pass = 0; pass < (split ? 2 : 1); ++pass) {
      const int kind = split ? (pass == 0 ? lk : hk) : lk;
      const int mask = split ? (1 << pass) : 3;
      const double val = split ? values[2 * d + pass] : values[2 * d];
      if (kind == ST_BC_NONE) continue;
      const unsigned g = grid_for(2 * outer * h * inner);
      for (int64_t layer = narrow ? 0 : -1; layer < (narrow ? h : 0); ++layer) {
        if (dtype == ST_F64) {
          hipLaunchKernelGGL(k_fill_side<double>, dim3(g), dim3(256), 0, s, (double*)ptr, outer, n,
                             inner, h, kind, val, mask, layer);
        } else if (dtype == ST_F32) {
          hipLaunchKernelGGL(k_fill_side<float>, dim3(g), dim3(256), 0, s, (float*)ptr, outer, n,
                             inner, h, kind, (float)val, mask, layer);
        } else {
          float fv = (float)val;
          unsigned u;
          memcpy(&u, &fv, 4);
          u += 0x7fffu + ((u >> 16) & 1u);
          hipLaunchKernelGGL(k_fill_side<unsigned short>, dim3(g), dim3(256), 0, s,
                             (unsigned short*)ptr, outer, n, inner, h, kind,
                             (unsigned short)(u >> 16), mask, layer);
        }
        HIPCHK(hipGetLastError());
      }
    }
  }
  return ST_OK;
}

int st_fill_halo_batched(const void* descs, int count, int ndim, int64_t max_slab, void* stream) {
  if (!descs || count < 0 || ndim < 1 || ndim > 4) return fail(ST_E_INVALID, "bad fill_halo_batched arguments");
  if (count == 0) return ST_OK;
  if (count > 65535) return fail(ST_E_INVALID, "too many arrays in one batched fill (%d)", count);
  unsigned gx = grid_for(max_slab);
  if (gx > 1024) gx = 1024;
  for (int d = 0; d < ndim; ++d) {
    hipLaunchKernelGGL(k_fill_batched, dim3(gx, (unsigned)count), dim3(256), 0, (hipStream_t)stream,
                       (const st_fill_desc*)descs, d);
    HIPCHK(hipGetLastError());
  }
  return ST_OK;
}

int st_copy(void* dst, const void* src, size_t nbytes, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const bool aligned = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0;
  size_t done = 0;
  if (aligned && nbytes >= 16) {
    const size_t n16 = nbytes / 16;
    int64_t blocks = (int64_t)(n16 + 255) / 256;
    if (blocks > (int64_t)1 << 30) blocks = (int64_t)1 << 30;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (u32x4*)dst, (const u32x4*)src, n16);
    done = n16 * 16;
  }
  if (done < nbytes) {
    hipLaunchKernelGGL(k_copy1, dim3(grid_for((int64_t)(nbytes - done))), dim3(256), 0, s,
                       (unsigned char*)dst + done, (const unsigned char*)src + done, nbytes - done);
  }
  HIPCHK(hipGetLastError());
  return ST_OK;
}

int st_copy_wide(void* dst, const void* src, size_t nbytes, int vectors, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if ((((uintptr_t)dst | (uintptr_t)src) & 15) != 0 || nbytes % 16 != 0)
    return fail(ST_E_INVALID, "st_copy_wide: pointers and size must be multiples of 16 bytes");
  if (vectors != 1 && vectors != 2 && vectors != 4 && vectors != 8)
    return fail(ST_E_INVALID, "st_copy_wide: vectors must be 1, 2, 4 or 8 (got %d)", vectors);
  const size_t n16 = nbytes / 16;
  if (n16 == 0) return ST_OK;
  const size_t per = (size_t)256 * vectors;
  const size_t blocks = (n16 + per - 1) / per;
  if (blocks > ((size_t)1 << 31) - 1) return fail(ST_E_INVALID, "st_copy_wide: buffer too large");
  const dim3 g((unsigned)blocks), b(256);
  switch (vectors) {
    case 1: hipLaunchKernelGGL(k_copy16_wide<1>, g, b, 0, s, (u32x4*)dst, (const u32x4*)src, n16); break;
    case 2: hipLaunchKernelGGL(k_copy16_wide<2>, g, b, 0, s, (u32x4*)dst, (const u32x4*)src, n16); break;
    case 4: hipLaunchKernelGGL(k_copy16_wide<4>, g, b, 0, s, (u32x4*)dst, (const u32x4*)src, n16); break;
    default: hipLaunchKernelGGL(k_copy16_wide<8>, g, b, 0, s, (u32x4*)dst, (const u32x4*)src, n16); break;
  }
  HIPCHK(hipGetLastError());
  return ST_OK;
}

static int norms_into(const void* a, int dtype_a, const void* b, int dtype_b, int ndim, const int64_t* shape,
                      const int64_t* halo, void* stream, double* out, int nout, bool no_alloc) {
  if (!a || !b || !out) return fail(ST_E_INVALID, "null argument");
  Interior in;
  int rc = make_interior(ndim, shape, halo, in);
  if (rc) return rc;
  const int64_t np = red_parts(in);
  NormAcc host{};
  hipStream_t s = (hipStream_t)stream;
  const NormOp op{a, b, dtype_a, dtype_b};
  if (no_alloc) {
    if (np > STATIC_PARTS)
      return fail(ST_E_INVALID,
                  "st_residual_norms_static: the interior needs %lld partials of at most %lld cells, more than "
                  "the %lld the static scratch holds (fields up to about %lld interior cells); use "
                  "st_residual_norms",
                  (long long)np, (long long)RED_CELLS, (long long)STATIC_PARTS,
                  (long long)(STATIC_PARTS * RED_CELLS));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    NormAcc* sp = nullptr;
    rc = scratch_address(dev, &sp);
    if (rc) return rc;
    rc = run_reduction(op, in, sp, sp + STATIC_PARTS, &host, s);
  } else {
    std::lock_guard<std::mutex> lock(g_ws_mu);
    void* ws = nullptr;
    rc = workspace((size_t)(np + 1) * sizeof(NormAcc), &ws);
    if (rc) return rc;
    rc = run_reduction(op, in, (NormAcc*)ws, (NormAcc*)ws + np, &host, s);
  }
  if (rc) return rc;
  const double r[4] = {host.v[0], host.v[1], sqrt(host.v[2]), sqrt(host.v[3])};
  for (int k = 0; k < nout; ++k) out[k] = host.v[4] != 0.0 ? __builtin_nan("") : r[k];
  return ST_OK;
}

int st_diff_norms(const void* a, int dtype_a, const void* b, int dtype_b, int ndim,
                  const int64_t* shape, const int64_t* halo, void* stream, double* out) {
  return norms_into(a, dtype_a, b, dtype_b, ndim, shape, halo, stream, out, 2, false);
}

int st_residual_norms(const void* a, int dtype_a, const void* b, int dtype_b, int ndim,
                      const int64_t* shape, const int64_t* halo, void* stream, double* out) {
  return norms_into(a, dtype_a, b, dtype_b, ndim, shape, halo, stream, out, 4, false);
}

int st_residual_norms_static(const void* a, int dtype_a, const void* b, int dtype_b, int ndim,
                             const int64_t* shape, const int64_t* halo, void* stream, double* out) {
  return norms_into(a, dtype_a, b, dtype_b, ndim, shape, halo, stream, out, 4, true);
}

int st_field_digest(const void* a, int dtype, int ndim, const int64_t* shape, const int64_t* halo, void* stream,
                    uint64_t* out) {
  if (!a || !out) return fail(ST_E_INVALID, "null argument");
  Interior in;
  int rc = make_interior(ndim, shape, halo, in);
  if (rc) return rc;
  const int64_t np = red_parts(in);
  unsigned long long host = 0;
  std::lock_guard<std::mutex> lock(g_ws_mu);
  void* ws = nullptr;
  rc = workspace((size_t)(np + 1) * sizeof(unsigned long long), &ws);
  if (rc) return rc;
  const DigestOp op{a, dtype};
  rc = run_reduction(op, in, (unsigned long long*)ws, (unsigned long long*)ws + np, &host, (hipStream_t)stream);
  if (rc) return rc;
  *out = host;
  return ST_OK;
}

int st_bricks(void* flat, void* bricked, int dtype, int ndim, const int64_t* shape, int bs,
              int to_bricks, void* stream) {
  if (ndim < 1 || ndim > 3 || bs < 1) return fail(ST_E_INVALID, "bad bricks arguments");
  Shape4 sh;
  int64_t nb[3] = {1, 1, 1};
  int64_t total = 1;
  for (int d = 0; d < ndim; ++d) {
    sh.s[d] = shape[d];
    nb[d] = (shape[d] + bs - 1) / bs;
    total *= nb[d] * bs;
  }
  hipStream_t s = (hipStream_t)stream;
  const unsigned g = grid_for(total);
  if (esize(dtype) == 8)
    hipLaunchKernelGGL(k_bricks<double>, dim3(g), dim3(256), 0, s, (double*)flat, (double*)bricked,
                       ndim, sh, nb[0], nb[1], nb[2], bs, to_bricks, total);
  else if (esize(dtype) == 4)
    hipLaunchKernelGGL(k_bricks<float>, dim3(g), dim3(256), 0, s, (float*)flat, (float*)bricked, ndim,
                       sh, nb[0], nb[1], nb[2], bs, to_bricks, total);
  else
    hipLaunchKernelGGL(k_bricks<unsigned short>, dim3(g), dim3(256), 0, s, (unsigned short*)flat,
                       (unsigned short*)bricked, ndim, sh, nb[0], nb[1], nb[2], bs, to_bricks, total);
  HIPCHK(hipGetLastError());
  return ST_OK;
}

int st_event_create(void** ev) {
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  *ev = (void*)e;
  return ST_OK;
}
int st_event_record(void* ev, void* stream) {
  HIPCHK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
  return ST_OK;
}
int st_event_elapsed_ms(void* start, void* stop, float* ms) {
  HIPCHK(hipEventSynchronize((hipEvent_t)stop));
  HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return ST_OK;
}
int st_event_destroy(void* ev) {
  HIPCHK(hipEventDestroy((hipEvent_t)ev));
  return ST_OK;
}
int st_event_sync(void* ev) {
  HIPCHK(hipEventSynchronize((hipEvent_t)ev));
  return ST_OK;
}
int st_event_query(void* ev, int* done) {
  const hipError_t e = hipEventQuery((hipEvent_t)ev);
  if (e == hipErrorNotReady) {
    *done = 0;
    return ST_OK;
  }
  HIPCHK(e);
  *done = 1;
  return ST_OK;
}
int st_stream_sync(void* stream) {
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return ST_OK;
}
int st_device_sync(void) {
  HIPCHK(hipDeviceSynchronize());
  return ST_OK;
}

int st_graph_begin(void* stream) {
  HIPCHK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return ST_OK;
}
int st_graph_end(void* stream, void** graph_exec) {
  hipGraph_t g;
  HIPCHK(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t ex;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return fail((int)e, "hipGraphInstantiate: %s", hipGetErrorString(e));
  *graph_exec = (void*)ex;
  return ST_OK;
}
int st_graph_launch(void* graph_exec, void* stream) {
  HIPCHK(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
  return ST_OK;
}
int st_graph_destroy(void* graph_exec) {
  HIPCHK(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return ST_OK;
}

}  // extern "C"
