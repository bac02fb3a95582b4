// Hand-written device templates for the MI355X (gfx950) stencil backend.
//
// Every JIT-built stencil module includes this header and supplies a traits
// struct P (generated from the stencil IR) describing the fields, the
// accesses and the per-cell expression.  The kernel template below does the
// data movement; the expression is the only generated arithmetic.
//
// Execution model (one workgroup = NT threads = NT/64 wave64s):
//   * Arrays are 3-D (s0, s1, s2), C-contiguous, interior (n0, n1, n2) at
//     offset (h0, h1, h2).  1-D/2-D stencils are mapped to 3-D by leading
//     unit dimensions, so axis 0 is the streaming axis and (1, 2) the plane.
//   * A workgroup owns a TY x TZ tile of the plane and a chunk of `xch`
//     planes along axis 0, and streams through the chunk plane by plane.
//   * Accesses with no in-plane offset ("column" accesses, any dx) come from
//     a per-thread register queue holding planes x+QLO .. x+QHI of the
//     thread's own cells, so every plane is fetched from HBM once.
//   * Accesses with an in-plane offset read an LDS tile of plane x+dx with a
//     (RY, RZ) halo; the tile is loaded with row-contiguous 16-byte vector
//     loads (RZ rounded up to the vector width so rows stay aligned).
//   * The next plane's queue entries and LDS vectors are prefetched into
//     registers before the current plane is computed, so HBM latency
//     overlaps the arithmetic.
//   * Workgroup ids are remapped so that each XCD (own L2) receives a
//     contiguous range of tiles: neighbouring tiles share halos in one L2.
//
// Expression arithmetic is evaluated exactly as written (contraction off),
// so results are bit-identical to NumPy on the same inputs.
#pragma once
#pragma clang fp contract(off)

#define ST_HD __device__ __forceinline__
#include "st_math.h"

#ifndef ST_MAXF
#define ST_MAXF 12
#endif

namespace st {

typedef long long i64;
typedef unsigned short bf16_t;

// ---------------------------------------------------------------------------
// storage <-> compute conversions
// ---------------------------------------------------------------------------
template <class S, class C> struct Cvt {
  static ST_HD C ld(S v) { return (C)v; }
  static ST_HD S st(C v) { return (S)v; }
};
template <> struct Cvt<bf16_t, float> {
  static ST_HD float ld(bf16_t v) { return __builtin_bit_cast(float, ((unsigned)v) << 16); }
  // round to nearest even; NaN kept quiet
  static ST_HD bf16_t st(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
  }
};

template <class T, int V> struct alignas(sizeof(T) * V) Vec { T v[V]; };

template <class T, int V> ST_HD Vec<T, V> vload(const T* p) { return *reinterpret_cast<const Vec<T, V>*>(p); }
template <class T, int V> ST_HD void vstore(T* p, const Vec<T, V>& x) { *reinterpret_cast<Vec<T, V>*>(p) = x; }

// ---------------------------------------------------------------------------
// launch arguments (one by-value struct; the host packs the same layout)
// ---------------------------------------------------------------------------
struct SweepArgs {
  const void* in[ST_MAXF];
  void* out[ST_MAXF];
  i64 s0, s1, s2;   // array extents
  int h0, h1, h2;   // array halo widths
  int n0, n1, n2;   // interior extents
  int ntz, nty, nch;  // tiles along z, y; chunks along x
  int xch;          // planes per chunk
  int nxcd;         // XCDs for the workgroup swizzle (1 = identity)
  int tmode;        // temporal blocking: 0 = zero halo between levels, 1 = recompute ghosts
  int yslab;        // tile order: 0 = XCD gets an x-slab (z, y, x-chunk), 1 = a y-slab (z, x-chunk, y)
  int wrap;         // bit d set: axis d is periodic; reads past the interior wrap around
                    // (the halo along it is neither read nor written)
  int nby, nbz;     // brick-major arrays: bricks along axes 1 and 2 (unused for flat arrays)
};

// Address map of an array: flat C-contiguous (B0 == 0), or brick-major with
// B0 x B1 x B2 bricks (nb0 x nby x nbz bricks in C order, each brick's cells
// contiguous and row-major).  X, Y, Z are array coordinates (halo included).
// base(X): element offset of plane X's first cell; ioff(Y, Z): offset of cell
// (Y, Z) from there; bytes(X): bytes from base(X) to the end of its plane
// (flat) or of its slab of bricks (brick-major).  A VZ-vector never crosses a
// brick when B2 % VZ == 0 and Z % VZ == 0.
template <int B0, int B1, int B2> struct Layout {
  i64 plane;
  int s2, nby, nbz;
  ST_HD i64 base(int X) const {
    if constexpr (B0 == 0) return (i64)X * plane;
    else return (i64)(X / B0) * ((i64)nby * nbz * (B0 * B1 * B2)) + (i64)(X % B0) * (B1 * B2);
  }
  ST_HD int ioff(int Y, int Z) const {
    if constexpr (B0 == 0) return Y * s2 + Z;
    else return ((Y / B1) * nbz + Z / B2) * (B0 * B1 * B2) + (Y % B1) * B2 + (Z % B2);
  }
  ST_HD int bytes(int X, int es) const {
    if constexpr (B0 == 0) return (int)(plane * es);
    else return (int)(((i64)nby * nbz * (B0 * B1 * B2) - (i64)(X % B0) * (B1 * B2)) * es);
  }
};

// Bijective remap of dispatch order -> logical tile so that XCD x (which
// receives workgroups x, x+nxcd, ...) owns a contiguous logical range.
ST_HD int xcd_swizzle(int i, int n, int nxcd) {
  if (nxcd <= 1) return i;
  int x = i % nxcd, j = i / nxcd;
  int q = n / nxcd, r = n % nxcd;
  return x * q + (x < r ? x : r) + j;
}

constexpr int pmod(int a, int n) { return ((a % n) + n) % n; }
constexpr int gcd_i(int a, int b) { return b == 0 ? a : gcd_i(b, a % b); }
constexpr int lcm_i(int a, int b) { return a / gcd_i(a, b) * b; }

// compile-time integer sequences (hiprtc has no <utility>)
template <int... I> struct iseq {};
template <int N, int... I> struct mkseq : mkseq<N - 1, N - 1, I...> {};
template <int... I> struct mkseq<0, I...> { using type = iseq<I...>; };
template <int V> struct ic { static constexpr int value = V; };
template <class F, int... I> ST_HD void for_each_ic(F&& f, iseq<I...>) { (f(ic<I>{}), ...); }

template <class P> struct Tab {
  static constexpr int NTZ = P::TZ / P::VZ;
  static constexpr int NTY = P::NT / NTZ > 0 ? P::NT / NTZ : 1;
  static constexpr int KY = (P::TY + NTY - 1) / NTY;
  static constexpr bool HASQ = P::QHI >= P::QLO;
  static constexpr int QN = HASQ ? (P::QHI - P::QLO + 1) : 1;
  static constexpr int rzv(int b) { return ((P::BRZ[b] + P::VZ - 1) / P::VZ) * P::VZ; }
  static constexpr int rows(int b) { return P::TY + 2 * P::BRY[b]; }
  static constexpr int cols(int b) { return P::TZ + 2 * rzv(b); }
  static constexpr int vpr(int b) { return cols(b) / P::VZ; }
  // LDS row pitch of buffer b, in elements.  A wave's lanes walk the tile
  // row by row, NTZ lanes a row.  Where a row ends inside a 32-lane span
  // (NTZ not a multiple of 32), the row is padded so that the pitch is
  // congruent to the tile width modulo 256 bytes: the lane after a row wrap
  // then continues in the bank after the previous lane's, and one LDS access
  // of a wave stays free of bank conflicts.  Unpadded, the halo columns
  // shift the next row's lanes onto the banks of the lanes before them.
  static constexpr int pitch(int b) {
    if (NTZ % 32 == 0) return cols(b);
    int p = cols(b);
    while (((p - P::TZ) * (int)sizeof(typename P::C)) % 256 != 0) ++p;
    return p;
  }
  static constexpr int svec(int b) { return rzv(b) / P::VZ; }
  // z-neighbours of my queue cells read across lanes (P::ZDPP): lane
  // distances 1 .. ndz(b) on each side, for buffers copied from the queue
  static constexpr bool zdpp(int b) { return P::ZDPP && fromq(b) && P::BRZ[b] > 0; }
  static constexpr int ndz(int b) { return zdpp(b) ? (P::BRZ[b] + P::VZ - 1) / P::VZ : 0; }
  static constexpr int ndz_max() { int m = 1; for (int k = 0; k < P::NB; ++k) m = ndz(k) > m ? ndz(k) : m; return m; }
  // the tile interior of buffer b is copied from the register queue (only
  // the halo ring is fetched from memory)
  static constexpr bool fromq(int b) {
    return HASQ && P::QF[P::BF[b]] && P::BDX[b] >= P::QLO && P::BDX[b] <= P::QHI;
  }
  static constexpr int gvecs(int b) {
    return fromq(b) ? 2 * P::BRY[b] * vpr(b) + P::TY * 2 * svec(b) : rows(b) * vpr(b);
  }
  static constexpr int nvb(int b) { return (gvecs(b) + P::NT - 1) / P::NT; }
  static constexpr int off(int b) { int o = 0; for (int k = 0; k < b; ++k) o += rows(k) * pitch(k); return o; }
  static constexpr int lds_elems() { int o = 0; for (int k = 0; k < P::NB; ++k) o += rows(k) * pitch(k); return o > 0 ? o : 1; }
  // LDS of a sweep: one set of plane buffers, or two (P::SDB) used by
  // alternate planes
  static constexpr int sweep_elems() { return (P::SDB ? 2 : 1) * lds_elems(); }
  static constexpr int max_nvb() { int m = 1; for (int k = 0; k < P::NB; ++k) m = nvb(k) > m ? nvb(k) : m; return m; }
  static constexpr int buf(int f, int dx) {
    for (int k = 0; k < P::NB; ++k) if (P::BF[k] == f && P::BDX[k] == dx) return k;
    return -1;
  }
  // (row, vector column) in buffer b of the i-th vector fetched from memory
  static ST_HD void gvec_rc(int b, int i, int& r, int& c) {
    const int w = vpr(b);
    if (!fromq(b)) { r = i / w; c = i % w; return; }
    const int band = P::BRY[b] * w;
    if (i < band) { r = i / w; c = i % w; return; }
    i -= band;
    if (i < band) { r = P::BRY[b] + P::TY + i / w; c = i % w; return; }
    i -= band;
    const int sv = svec(b);
    r = P::BRY[b] + i / (2 * sv);
    const int k = i % (2 * sv);
    c = k < sv ? k : w - 2 * sv + k;
  }
};

// Per-cell view handed to the generated expression.  The register queue is
// a ring: at rotation R, slot (i + R) % QN holds plane x + QLO + i.  KD is
// the number of cells per thread (the queue's third extent).
// bytes of LDS the largest of several sweep instantiations needs
template <class... P> constexpr int sweep_lds_bytes() {
  const int bytes[] = {16, (int)(Tab<P>::sweep_elems() * sizeof(typename P::C))...};
  int m = 0;
  for (int b : bytes) m = b > m ? b : m;
  return m;
}

// Value of x in the lane DIR lanes away within the wave (DIR = +1: lane
// l+1, DPP wave_shl:1; DIR = -1: lane l-1, wave_shr:1), 32 bits at a time.
// Lanes whose source lies outside the wave get 0; callers never use those.
template <int DIR, class C> ST_HD C lane_shift(C x) {
  constexpr int ctrl = DIR > 0 ? 0x130 : 0x138;
  if constexpr (sizeof(C) == 8) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)u, ctrl, 0xf, 0xf, false);
    const unsigned hi = __builtin_amdgcn_update_dpp(0u, (unsigned)(u >> 32), ctrl, 0xf, 0xf, false);
    return __builtin_bit_cast(C, ((unsigned long long)hi << 32) | lo);
  } else {
    static_assert(sizeof(C) == 4, "lane_shift: 4- or 8-byte values");
    return __builtin_bit_cast(C, __builtin_amdgcn_update_dpp(0u, __builtin_bit_cast(unsigned, x), ctrl, 0xf, 0xf,
                                                               false));
  }
}

// YA: the thread's KD cells lie on adjacent rows (cell k + d is d rows
// below cell k), so an in-plane neighbour DY rows away that is another of
// my cells is read from my registers, not from LDS.
template <class P, int R, int KD = Tab<P>::KY, bool YA = false> struct Cell {
  using C = typename P::C;
  using T = Tab<P>;
  const C (*q)[T::QN][KD][P::VZ];
  const C* lds;
  int k, v, ly, lz;
  // P::ZDPP: my queue vectors' neighbours across lanes (zn[b][j]: lane
  // +(j+1) for j < ndz(b), lane -(j-ndz(b)+1) after), my lane's column in
  // the tile row and my lane in the wave
  const C (*zn)[2 * T::ndz_max()][P::VZ] = nullptr;
  int tzl = 0, wl = 0;

  template <int F, int DX, int DY, int DZ> ST_HD C g() const {
    if constexpr (DY == 0 && DZ == 0 && T::HASQ && DX >= P::QLO && DX <= P::QHI) {
      return q[F][(DX - P::QLO + R) % T::QN][k][v];
    } else if constexpr (DY == 0 && DZ != 0 && P::ZDPP && T::HASQ && DX >= P::QLO && DX <= P::QHI &&
                         P::QF[F] && T::buf(F, DX) >= 0 && T::zdpp(T::buf(F, DX))) {
      // cell z+DZ: in my vector, else in the vector of the lane d lanes
      // away (same tile row, same wave), else (row edge) the LDS halo
      constexpr int b = T::buf(F, DX);
      const int t = v + DZ;
      if (t >= 0 && t < P::VZ) return q[F][(DX - P::QLO + R) % T::QN][k][t];
      const int d = t >= 0 ? t / P::VZ : -((P::VZ - 1 - t) / P::VZ);
      const int vv = t - d * P::VZ;
      const int j = d > 0 ? d - 1 : T::ndz(b) - d - 1;
      const bool lane_ok = (unsigned)(tzl + d) < (unsigned)T::NTZ && (unsigned)(wl + d) < 64u;
      C val = zn[b][j][vv];
      if (!lane_ok) val = lds[T::off(b) + (ly + P::BRY[b]) * T::pitch(b) + (lz + T::rzv(b) + DZ)];
      return val;
    } else if constexpr (DY == 0 && DZ != 0 && DZ > -P::VZ && DZ < P::VZ && T::HASQ &&
                         DX >= P::QLO && DX <= P::QHI && P::QF[F]) {
      // a neighbour along z inside my own vector is in the queue as well (the
      // LDS interior is copied from the same registers); v is a constant
      // once the per-vector loop is unrolled
      if (v + DZ >= 0 && v + DZ < P::VZ) return q[F][(DX - P::QLO + R) % T::QN][k][v + DZ];
      constexpr int b = T::buf(F, DX);
      return lds[T::off(b) + (ly + P::BRY[b]) * T::pitch(b) + (lz + T::rzv(b) + DZ)];
    } else {
      constexpr int b = T::buf(F, DX);
      static_assert(b >= 0, "access has no LDS buffer");
      if constexpr (YA && DY != 0 && DZ > -P::VZ && DZ < P::VZ && T::HASQ && DX >= P::QLO && DX <= P::QHI &&
                    P::QF[F] && T::fromq(b)) {
        // another of my cells (the LDS interior is copied from the queue);
        // k and v are constants once the cell loops are unrolled
        if (k + DY >= 0 && k + DY < KD && v + DZ >= 0 && v + DZ < P::VZ)
          return q[F][(DX - P::QLO + R) % T::QN][k + DY][v + DZ];
      }
      return lds[T::off(b) + (ly + P::BRY[b] + DY) * T::pitch(b) + (lz + T::rzv(b) + DZ)];
    }
  }
};

// ---------------------------------------------------------------------------
// buffer memory access
// ---------------------------------------------------------------------------
//
// Every global load and store of the sweep goes through a buffer resource
// covering one plane (base = the plane, range = its bytes).  A lane whose
// access must not happen gets the offset BUF_OOB instead of a branch: the
// hardware drops such stores and returns 0 for such loads.  With no
// branches around memory instructions their count per step is fixed, so the
// compiler waits only for the loads it needs (vmcnt(N)) and the stores of
// the previous plane stay in flight.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int BUF_OOB = (int)0x80000000;          // beyond any plane's range (planes < 2 GiB)
constexpr int BUF_CONFIG = 0x00020000;            // raw buffer, dword data format (gfx9 family)
constexpr int BUF_NT = 2;                         // non-temporal store

ST_HD Rsrc make_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, BUF_CONFIG);
}

template <int B> struct BufIO;
template <> struct BufIO<16> {
  typedef unsigned T __attribute__((ext_vector_type(4)));
  template <int AUX> static ST_HD T ld(Rsrc r, int off) { return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX); }
  template <int AUX> static ST_HD void st(T v, Rsrc r, int off) { __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, AUX); }
};
template <> struct BufIO<8> {
  typedef unsigned T __attribute__((ext_vector_type(2)));
  template <int AUX> static ST_HD T ld(Rsrc r, int off) { return __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX); }
  template <int AUX> static ST_HD void st(T v, Rsrc r, int off) { __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, AUX); }
};
template <> struct BufIO<4> {
  typedef unsigned T;
  template <int AUX> static ST_HD T ld(Rsrc r, int off) { return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX); }
  template <int AUX> static ST_HD void st(T v, Rsrc r, int off) { __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, AUX); }
};
template <> struct BufIO<2> {
  typedef unsigned short T;
  template <int AUX> static ST_HD T ld(Rsrc r, int off) { return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, AUX); }
  template <int AUX> static ST_HD void st(T v, Rsrc r, int off) { __builtin_amdgcn_raw_buffer_store_b16(v, r, off, 0, AUX); }
};

template <class S, int V, int AUX = 0> ST_HD Vec<S, V> bload(Rsrc r, int off) {
  using IO = BufIO<(int)sizeof(S) * V>;
  return __builtin_bit_cast(Vec<S, V>, IO::template ld<AUX>(r, off));
}
template <class S, int V, int AUX> ST_HD void bstore(const Vec<S, V>& x, Rsrc r, int off) {
  using IO = BufIO<(int)sizeof(S) * V>;
  IO::template st<AUX>(__builtin_bit_cast(typename IO::T, x), r, off);
}

// index along a periodic axis of interior extent n (|i| < n beyond either side)
ST_HD int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ---------------------------------------------------------------------------
// the plane-streaming sweep
// ---------------------------------------------------------------------------
//
// a.wrap bit d: axis d of the 3-D map is periodic.  Reads past the interior
// along it wrap around instead of reading the halo, so no halo fill is
// needed between steps (the output's halo is not written).  The host allows
// it when n_d >= h_d and, for the vector variant, n2 % VZ == 0.
// lds: the workgroup's LDS, at least Tab<P>::sweep_elems() elements (kernels
// that dispatch to several instantiations share one pool sized to the largest)
template <class P>
ST_HD void sweep_tile(const SweepArgs& a, int tile_z, int tile_y, int chunk, typename P::C* lds) {
  using S = typename P::S;
  using C = typename P::C;
  using T = Tab<P>;
  constexpr int NT = P::NT, TY = P::TY, TZ = P::TZ, VZ = P::VZ, NF = P::NF, NO = P::NO, NB = P::NB;
  constexpr int NOUT = P::NOUT;  // arrays written: NO results (+ NO compensations)
  constexpr int NTZ = T::NTZ, NTY = T::NTY, KY = T::KY, QN = T::QN, NVB = T::max_nvb();
  constexpr int ES = (int)sizeof(S);
  static_assert(P::NT % NTZ == 0, "NT must be a multiple of TZ/VZ");
  static_assert(P::BR0 == 0 || P::BR2 % VZ == 0, "a vector must not cross a brick");

  const int tid = threadIdx.x;
  const int tz = tid % NTZ, ty = tid / NTZ;
  const int z0 = tile_z * TZ, y0 = tile_y * TY;
  const int xb = chunk * a.xch;
  const int xe = (xb + a.xch < a.n0) ? xb + a.xch : a.n0;
  const int s1 = (int)a.s1, s2 = (int)a.s2;
  const Layout<P::BR0, P::BR1, P::BR2> lay{a.s1 * a.s2, s2, a.nby, a.nbz};
  const bool w0 = a.wrap & 1, w1 = a.wrap & 2, w2 = a.wrap & 4;
  const int zc = z0 + tz * VZ;                  // first interior z of my vector
  const bool zin = zc < a.n2;                   // my cells are interior (stored)
  const int zq = w2 ? wrapi(zc, a.n2) : zc;     // where my vector is read from
  const bool zload = a.h2 + zq + VZ <= s2;      // ... and it lies inside the array

  constexpr int PD = P::PD;   // planes prefetched ahead into registers
  // halo rings loaded RPD planes ahead (1: the next plane).  Fetching a
  // plane's ring when the neighbouring tiles fetch it as their queue head
  // (RPD = QHI + PD) lets the shared rows hit in the XCD's L2.
  constexpr int RPD = P::RPD < 1 ? 1 : P::RPD;
  C q[NF][QN][KY][VZ];
  C pq[PD][NF][KY][VZ];
  C pbq[RPD][NB > 0 ? NB : 1][NVB][VZ];

  auto in_rsrc = [&](int f, int x) -> Rsrc {
    const int px = w0 ? wrapi(x, a.n0) : x;
    return make_rsrc(reinterpret_cast<const S*>(a.in[f]) + lay.base(a.h0 + px), lay.bytes(a.h0 + px, ES));
  };

  // In-plane byte offsets do not depend on the plane: they are computed
  // once per workgroup (BUF_OOB where the access must not happen), so a
  // plane step spends no VALU on the address arithmetic of the ring.
  // qoff: my cells (queue loads); roff: the halo-ring (or whole-tile)
  // vectors this thread fetches for every LDS buffer.
  int qoff[KY];
#pragma unroll
  for (int k = 0; k < KY; ++k) {
    const int y = y0 + ty + k * NTY;
    const int yq = w1 ? wrapi(y, a.n1) : y;
    const bool ok = zload && a.h1 + yq < s1 && ty + k * NTY < TY;
    qoff[k] = ok ? lay.ioff(a.h1 + yq, a.h2 + zq) * ES : BUF_OOB;
  }
  int roff[NB > 0 ? NB : 1][NVB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int j = 0; j < NVB; ++j) {
      const int i = tid + j * NT;
      bool ok = (j < T::nvb(b)) && (i < T::gvecs(b));
      int rr = 0, c = 0;
      T::gvec_rc(b, ok ? i : 0, rr, c);
      int gy = y0 - P::BRY[b] + rr;           // interior coordinates
      int gz = z0 - T::rzv(b) + c * VZ;
      if (w1) gy = wrapi(gy, a.n1);
      if (w2) gz = wrapi(gz, a.n2);
      gy += a.h1;
      gz += a.h2;
      ok = ok && gy >= 0 && gy < s1 && gz >= 0 && gz + VZ <= s2;
      roff[b][j] = ok ? lay.ioff(gy, gz) * ES : BUF_OOB;
    }
  }

  // my cells of plane x of field f (cells outside the interior but inside
  // the array are loaded too: they are the halo other cells read)
  // (live = false issues the same loads out of range: zeros, no traffic)
  auto load_cells = [&](int f, int x, C (&dst)[KY][VZ], bool live) {
    const Rsrc r = in_rsrc(f, x);
#pragma unroll
    for (int k = 0; k < KY; ++k) {
      // a queue plane is read once per chunk: optionally non-temporal (P::NTL)
      const Vec<S, VZ> w = bload<S, VZ, P::NTL ? BUF_NT : 0>(r, live ? qoff[k] : BUF_OOB);
#pragma unroll
      for (int v = 0; v < VZ; ++v) dst[k][v] = Cvt<S, C>::ld(w.v[v]);
    }
  };

  // prefetch the memory part (halo ring, or whole tile) of every LDS buffer
  auto load_bufs = [&](int x, bool live, C (&pb)[NB > 0 ? NB : 1][NVB][VZ]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const Rsrc r = in_rsrc(P::BF[b], x + P::BDX[b]);
#pragma unroll
      for (int j = 0; j < NVB; ++j) {
        const Vec<S, VZ> w = bload<S, VZ>(r, live ? roff[b][j] : BUF_OOB);
#pragma unroll
        for (int v = 0; v < VZ; ++v) pb[b][j][v] = Cvt<S, C>::ld(w.v[v]);
      }
    }
  };

  auto store_bufs = [&](auto Rc, const C (&pb)[NB > 0 ? NB : 1][NVB][VZ], C* ldsb) {
    constexpr int R = decltype(Rc)::value;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      C* base = ldsb + T::off(b);
      if constexpr (T::HASQ) {
        if (T::fromq(b)) {
#pragma unroll
          for (int k = 0; k < KY; ++k) {
            if (ty + k * NTY < TY) {
              C* d = base + (P::BRY[b] + ty + k * NTY) * T::pitch(b) + T::rzv(b) + tz * VZ;
#pragma unroll
              for (int v = 0; v < VZ; ++v) d[v] = q[P::BF[b]][(P::BDX[b] - P::QLO + R) % QN][k][v];
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < NVB; ++j) {
        const int i = tid + j * NT;
        if (j < T::nvb(b) && i < T::gvecs(b)) {
          int r, c;
          T::gvec_rc(b, i, r, c);
          C* d = base + r * T::pitch(b) + c * VZ;
#pragma unroll
          for (int v = 0; v < VZ; ++v) d[v] = pb[b][j][v];
        }
      }
    }
  };

  // ---- prologue: queue planes xb+QLO .. xb+QHI-1, prefetch planes
  //      xb+QHI .. xb+QHI+PD-1 (those the chunk will install)
  if constexpr (T::HASQ) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (!P::QF[f]) continue;
#pragma unroll
      for (int i = 0; i + 1 < QN; ++i) load_cells(f, xb + P::QLO + i, q[f][i], true);
#pragma unroll
      for (int k = 0; k < PD; ++k) load_cells(f, xb + P::QHI + k, pq[k][f], xb + k < xe);
    }
  }
#pragma unroll
  for (int d = 0; d < RPD; ++d) load_bufs(xb + d, xb + d < xe, pbq[d]);

  // P::SDB: two sets of LDS plane buffers used by alternate planes, so a
  // plane's writes need no barrier after the previous plane's reads (one
  // barrier per plane instead of two)
  constexpr int NBUF = P::SDB ? 2 : 1;
  constexpr int UN = lcm_i(QN, NBUF);

  // one plane: rotation R = (x - xb) % QN names the queue slots, so the
  // queue never moves between registers (every use of R takes it modulo QN)
  auto step = [&](auto Rc, int x) {
    constexpr int R = decltype(Rc)::value;
    C* const ldsb = lds + (NBUF == 2 ? (R & 1) : 0) * T::lds_elems();
    // install plane x+QHI in the slot that held plane x-1+QLO
    if constexpr (T::HASQ) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if (!P::QF[f]) continue;
#pragma unroll
        for (int k = 0; k < KY; ++k)
#pragma unroll
          for (int v = 0; v < VZ; ++v) q[f][(R + QN - 1) % QN][k][v] = pq[0][f][k][v];
#pragma unroll
        for (int d = 0; d + 1 < PD; ++d)
#pragma unroll
          for (int k = 0; k < KY; ++k)
#pragma unroll
            for (int v = 0; v < VZ; ++v) pq[d][f][k][v] = pq[d + 1][f][k][v];
      }
    }
    if constexpr (NB > 0) {
      if constexpr (NBUF == 1) __syncthreads();
      store_bufs(Rc, pbq[0], ldsb);
      __syncthreads();
    }
    // prefetch plane x+PD of the queue and the LDS ring of plane x+RPD while
    // plane x is computed (a plane past the chunk is loaded as out of range)
    if constexpr (T::HASQ) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (P::QF[f]) load_cells(f, x + PD + P::QHI, pq[PD - 1][f], x + PD < xe);
    }
#pragma unroll
    for (int d = 0; d + 1 < RPD; ++d)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int j = 0; j < NVB; ++j)
#pragma unroll
          for (int v = 0; v < VZ; ++v) pbq[d][b][j][v] = pbq[d + 1][b][j][v];
    load_bufs(x + RPD, x + RPD < xe, pbq[RPD - 1]);

    Rsrc ro[NOUT];
#pragma unroll
    for (int m = 0; m < NOUT; ++m)
      ro[m] = make_rsrc(reinterpret_cast<S*>(a.out[m]) + lay.base(a.h0 + x), lay.bytes(a.h0 + x, ES));
#pragma unroll
    for (int k = 0; k < KY; ++k) {
      // rows past the tile (when NTY does not divide TY) compute on the
      // last row's LDS data and store nothing
      const int lyk = ty + k * NTY < TY ? ty + k * NTY : TY - 1;
      const int y = y0 + ty + k * NTY;
      const bool valid = ty + k * NTY < TY && y < a.n1 && zin;
      // P::ZDPP: neighbour lanes' queue vectors (all lanes active here)
      C zn[NB > 0 ? NB : 1][2 * T::ndz_max()][VZ];
      if constexpr (P::ZDPP) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (!T::zdpp(b)) continue;
          const int sl = (P::BDX[b] - P::QLO + R) % QN;
          C up[VZ], dn[VZ];
#pragma unroll
          for (int v = 0; v < VZ; ++v) up[v] = dn[v] = q[P::BF[b]][sl][k][v];
#pragma unroll
          for (int d = 0; d < T::ndz(b); ++d) {
#pragma unroll
            for (int v = 0; v < VZ; ++v) {
              up[v] = lane_shift<+1>(up[v]);
              dn[v] = lane_shift<-1>(dn[v]);
              zn[b][d][v] = up[v];
              zn[b][T::ndz(b) + d][v] = dn[v];
            }
          }
        }
      }
      C res[NOUT][VZ];
#pragma unroll
      for (int v = 0; v < VZ; ++v) {
        Cell<P, R> c{q, ldsb, k, v, lyk, tz * VZ + v, zn, tz, tid & 63};
        C o[NO];
        P::eval(c, o);
        if constexpr (P::UPD == 0) {
#pragma unroll
          for (int m = 0; m < NO; ++m) res[m][v] = o[m];
        } else {
          // time-update epilogue: output m is field TF[m] advanced by DT * rate
          for_each_ic([&](auto Mc) {
            constexpr int m = decltype(Mc)::value;
            const C base = c.template g<P::TF[m], 0, 0, 0>();
            if constexpr (P::UPD == 1) {
              res[m][v] = base + P::DT * o[m];
            } else {
              // Kahan: carry what the storage rounding drops in CF[m]
              const C y = P::DT * o[m] - c.template g<P::CF[m], 0, 0, 0>();
              const C t = Cvt<S, C>::ld(Cvt<S, C>::st(base + y));
              res[m][v] = t;
              res[NO + m][v] = (t - base) - y;
            }
          }, typename mkseq<NO>::type{});
        }
      }
      const int off = lay.ioff(a.h1 + y, a.h2 + zc) * ES;
      // one unconditional vector store per output.  With PARTIAL (rows whose
      // interior is not a multiple of VZ: batched instances with padded
      // rows) the lane holding a row's ragged end first reads the output's
      // cells past the interior and writes them back unchanged, so its
      // store stays a whole vector and no branch guards a memory access
      const bool ragged = P::PARTIAL && valid && zc + VZ > a.n2;
#pragma unroll
      for (int m = 0; m < NOUT; ++m) {
        Vec<S, VZ> w;
        if constexpr (P::PARTIAL) {
          const Vec<S, VZ> keep = bload<S, VZ>(ro[m], ragged ? off : BUF_OOB);
#pragma unroll
          for (int v = 0; v < VZ; ++v) w.v[v] = zc + v < a.n2 ? Cvt<S, C>::st(res[m][v]) : keep.v[v];
        } else {
#pragma unroll
          for (int v = 0; v < VZ; ++v) w.v[v] = Cvt<S, C>::st(res[m][v]);
        }
        bstore<S, VZ, P::NTS ? BUF_NT : 0>(w, ro[m], valid ? off : BUF_OOB);
      }
    }
  };

  for (int x = xb; x < xe; x += UN) {
    for_each_ic([&](auto Rc) {
      if (x + decltype(Rc)::value < xe) step(Rc, x + decltype(Rc)::value);
    }, typename mkseq<UN>::type{});
  }
}

// ---------------------------------------------------------------------------
// in-kernel temporal blocking: TT time levels per launch in registers + LDS
// ---------------------------------------------------------------------------
//
// A workgroup owns an output tile of (TY - 2*GY) x (TZ - 2*GZ) cells and a
// chunk of planes, and evaluates every level on the widened region TY x TZ:
// the output tile plus GY rows and GZ columns per side (GY >= (TT-1)*RY,
// GZ >= (TT-1)*RZ), so level l is exact on the output tile widened by
// (TT-l) radii, which is all level l+1 reads there.  Every level keeps the
// sweep's structure: a per-thread register queue of QN planes of the
// thread's own cells (x-neighbours, own-vector z-neighbours) and one LDS
// plane per (field, dx) buffer for the other in-plane neighbours.  Level 0
// is the input array: its queue head is loaded from memory and only its LDS
// halo ring is fetched.  Levels >= 1 never touch memory: a level's queue
// head is the value the thread has just computed and its LDS planes are
// copied from its queue.
//
// The levels are skewed by D = QHI + 1 planes along the stream axis: at step
// s, level l (1..TT) evaluates plane s - (l-1)*D from level-(l-1) planes
// s - (l-1)*D + QLO .. s - (l-1)*D + QHI, all produced in earlier steps.
// Levels are evaluated from TT down to 1, so each level's queue and LDS
// planes are read before this step's value of that level is pushed.
// Level TT is stored.  A chunk of planes [xb, xe) takes steps xb + (TT-1)*QLO
// .. xe - 1 + (TT-1)*D.
//
// Cells are assigned to threads flat: vector cell c = tid + k*NT (k < KC) of
// the TY x TZ/VZ region in row-major order, so the region's shape needs no
// relation to the workgroup size (lanes past the region duplicate the last
// cell and write nothing).
//
// Values of intermediate levels outside the interior:
//   tmode 0: 0 (T chained sweeps through zero-halo scratch arrays);
//   tmode 1: evaluated like interior cells (T steps with a copy-type halo
//            refill, when the halo is at least TT radii deep and was filled
//            before the launch);
//   a.wrap bit d: axis d is periodic; level-0 reads past the interior wrap
//            around, so intermediate levels there are exact periodic images
//            (no halo is read or written along that axis).

// index along a periodic axis of interior extent n (any distance)
ST_HD int wrapm(int i, int n) {
  const int r = i % n;
  return r < 0 ? r + n : r;
}

template <class P>
ST_HD void tblock_tile(const SweepArgs& a, int tile_z, int tile_y, int chunk) {
  using S = typename P::S;
  using C = typename P::C;
  using T = Tab<P>;
  constexpr int NT = P::NT, TY = P::TY, TZ = P::TZ, VZ = P::VZ, NB = P::NB, TT = P::TT;
  constexpr int QLO = P::QLO, QHI = P::QHI, QN = T::QN, NVB = T::max_nvb(), LE = T::lds_elems();
  constexpr int NTZ = TZ / VZ, CELLS = TY * NTZ, KC = (CELLS + NT - 1) / NT;
  constexpr int D = QHI + 1;
  constexpr int GY = P::GY, GZ = P::GZ, OTY = TY - 2 * GY, OTZ = TZ - 2 * GZ;
  constexpr int ES = (int)sizeof(S);
  constexpr int NF = P::NF;
  static_assert(P::NO == NF && P::UPD == 0, "temporal blocking: result m is the next state of field m");
  static_assert(T::HASQ && QLO <= 0 && QHI >= 0, "level queues must hold plane offset 0");
  static_assert(OTY > 0 && OTZ > 0 && GZ % VZ == 0 && TZ % VZ == 0, "temporal tile geometry");
  static_assert(TT >= 2, "at least two levels");

  // P::TBDB: two sets of LDS planes used by alternate steps, so a step's
  // plane writes need no barrier after the previous step's reads (one
  // barrier per step instead of two)
  constexpr int NBUF = P::TBDB ? 2 : 1;
  // P::TBPD: level-0 planes and rings are loaded PD steps ahead (1..4;
  // 2 by default: 1-5 % faster than 1 on MI355X, 3 and 4 no better)
  constexpr int PD = P::TBPD < 1 ? 1 : (P::TBPD > 4 ? 4 : P::TBPD);
  // steps unrolled per loop trip: a multiple of QN (queue rotation), of the
  // plane sets and of the prefetch stages
  constexpr int UN = lcm_i(QN, lcm_i(NBUF, PD));
  __shared__ C lds[NBUF * TT * LE];

  const int tid = threadIdx.x;
  const int y0 = tile_y * OTY - GY, z0 = tile_z * OTZ - GZ;   // region origin, interior coordinates
  const int xb = chunk * a.xch;
  const int xe = (xb + a.xch < a.n0) ? xb + a.xch : a.n0;
  const int s1 = (int)a.s1, s2 = (int)a.s2;
  const i64 plane = a.s1 * a.s2;
  const int pbytes = (int)(plane * ES);
  const bool w0 = a.wrap & 1, w1 = a.wrap & 2, w2 = a.wrap & 4;
  const bool zero_out = a.tmode == 0;

  // my cells, their in-plane load / store offsets and whether levels >= 1
  // keep their value there (tmode 0 zeroes cells outside the interior)
  int ly[KC], lz[KC], qoff[KC], soff[KC];
  bool act[KC], keep[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int c = tid + k * NT;
    act[k] = c < CELLS;
    const int cc = act[k] ? c : CELLS - 1;
    ly[k] = cc / NTZ;
    lz[k] = (cc % NTZ) * VZ;
    int y = y0 + ly[k], z = z0 + lz[k];
    const bool iny = y >= 0 && y < a.n1, inz = z >= 0 && z < a.n2;
    keep[k] = (w1 || iny) && (w2 || inz);
    const bool mine = act[k] && ly[k] >= GY && ly[k] < GY + OTY && lz[k] >= GZ && lz[k] < GZ + OTZ && iny && inz;
    soff[k] = mine ? ((a.h1 + y) * s2 + a.h2 + z) * ES : BUF_OOB;
    if (w1) y = wrapm(y, a.n1);
    if (w2) z = wrapm(z, a.n2);
    const bool ok = act[k] && a.h1 + y >= 0 && a.h1 + y < s1 && a.h2 + z >= 0 && a.h2 + z + VZ <= s2;
    qoff[k] = ok ? ((a.h1 + y) * s2 + a.h2 + z) * ES : BUF_OOB;
  }
  // in-plane offsets of the level-0 halo-ring vectors this thread fetches
  int roff[NB > 0 ? NB : 1][NVB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int j = 0; j < NVB; ++j) {
      const int i = tid + j * NT;
      bool ok = (j < T::nvb(b)) && (i < T::gvecs(b));
      int rr = 0, c = 0;
      T::gvec_rc(b, ok ? i : 0, rr, c);
      int gy = y0 - P::BRY[b] + rr, gz = z0 - T::rzv(b) + c * VZ;
      if (w1) gy = wrapm(gy, a.n1);
      if (w2) gz = wrapm(gz, a.n2);
      gy += a.h1;
      gz += a.h2;
      ok = ok && gy >= 0 && gy < s1 && gz >= 0 && gz + VZ <= s2;
      roff[b][j] = ok ? (gy * s2 + gz) * ES : BUF_OOB;
    }
  }

  // plane x of input field f: its index in the array, and whether it lies there
  // (the wrap is a compare-and-add loop: x is within a few interiors)
  auto in_plane = [&](int f, int x, bool& ok) -> Rsrc {
    int px = x;
    if (w0) {
      while (px < 0) px += a.n0;
      while (px >= a.n0) px -= a.n0;
    }
    ok = a.h0 + px >= 0 && a.h0 + px < a.s0;
    return make_rsrc(reinterpret_cast<const S*>(a.in[f]) + (i64)(a.h0 + (ok ? px : xb)) * plane, pbytes);
  };

  C q[TT][NF][QN][KC][VZ];       // level l's queue of every field (level 0: the inputs)
  C pq[PD][NF][KC][VZ];              // level-0 planes prefetched for the next PD steps
  C pb[PD][NB > 0 ? NB : 1][NVB][VZ];  // level-0 halo rings prefetched for the next PD steps

  auto load_q0 = [&](int f, int x, C (&dst)[KC][VZ], bool live) {
    bool ok;
    const Rsrc r = in_plane(f, x, ok);
    ok = ok && live;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const Vec<S, VZ> w = bload<S, VZ>(r, ok ? qoff[k] : BUF_OOB);
#pragma unroll
      for (int v = 0; v < VZ; ++v) dst[k][v] = Cvt<S, C>::ld(w.v[v]);
    }
  };
  auto load_ring = [&](int x, bool live, C (&dst)[NB > 0 ? NB : 1][NVB][VZ]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      bool ok;
      const Rsrc r = in_plane(P::BF[b], x + P::BDX[b], ok);
      ok = ok && live;
#pragma unroll
      for (int j = 0; j < NVB; ++j) {
        const Vec<S, VZ> w = bload<S, VZ>(r, ok ? roff[b][j] : BUF_OOB);
#pragma unroll
        for (int v = 0; v < VZ; ++v) dst[b][j][v] = Cvt<S, C>::ld(w.v[v]);
      }
    }
  };

  const int s_first = xb + (TT - 1) * QLO;
  const int s_last = xe - 1 + (TT - 1) * D;

  // ---- prologue: level-0 planes s_first+QLO .. s_first+QHI-1 in slots
  //      0 .. QN-2, planes s_first+QHI .. s_first+QHI+PD-1 and the rings of
  //      planes s_first .. s_first+PD-1 prefetched (a chunk takes at least
  //      TT steps); levels >= 1 and the LDS start at 0 (the halo rings of
  //      levels >= 1 are never written)
#pragma unroll
  for (int f = 0; f < NF; ++f) {
#pragma unroll
    for (int i = 0; i + 1 < QN; ++i) load_q0(f, s_first + QLO + i, q[0][f][i], true);
#pragma unroll
    for (int p = 0; p < PD; ++p) load_q0(f, s_first + p + QHI, pq[p][f], true);
  }
#pragma unroll
  for (int p = 0; p < PD; ++p) load_ring(s_first + p, true, pb[p]);
#pragma unroll
  for (int l = 1; l < TT; ++l)
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < QN; ++i)
#pragma unroll
        for (int k = 0; k < KC; ++k)
#pragma unroll
          for (int v = 0; v < VZ; ++v) q[l][f][i][k][v] = C(0);
  for (int i = tid; i < NBUF * TT * LE; i += NT) lds[i] = C(0);
  if constexpr (NBUF == 2) __syncthreads();      // the first step writes without a leading barrier

  // one step; (s - s_first) % UN = Rc, R = (s - s_first) % QN.  Plane p of
  // level l sits in queue slot (p - QLO - s_first) mod QN.
  auto step = [&](auto Rc, int s) {
    constexpr int R = decltype(Rc)::value % QN;
    constexpr int PI = decltype(Rc)::value % PD;     // prefetch stage of this step
    C* const ldsb = lds + (NBUF == 2 ? (decltype(Rc)::value & 1) : 0) * TT * LE;
    // level-0 head: plane s + QHI
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < KC; ++k)
#pragma unroll
        for (int v = 0; v < VZ; ++v) q[0][f][pmod(R - 1, QN)][k][v] = pq[PI][f][k][v];
    if constexpr (NBUF == 1) __syncthreads();
    // LDS planes: level l's buffer b holds plane s - l*D + BDX[b] of level l
    for_each_ic([&](auto Lc) {
      constexpr int L = decltype(Lc)::value;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        C* base = ldsb + L * LE + T::off(b);
        const int sl = pmod(R - L * D + P::BDX[b] - QLO, QN);
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          if (act[k]) {
            C* d = base + (P::BRY[b] + ly[k]) * T::pitch(b) + T::rzv(b) + lz[k];
#pragma unroll
            for (int v = 0; v < VZ; ++v) d[v] = q[L][P::BF[b]][sl][k][v];
          }
        }
        if constexpr (L == 0) {
#pragma unroll
          for (int j = 0; j < NVB; ++j) {
            const int i = tid + j * NT;
            if (j < T::nvb(b) && i < T::gvecs(b)) {
              int rr, c;
              T::gvec_rc(b, i, rr, c);
              C* d = base + rr * T::pitch(b) + c * VZ;
#pragma unroll
              for (int v = 0; v < VZ; ++v) d[v] = pb[PI][b][j][v];
            }
          }
        }
      }
    }, typename mkseq<TT>::type{});
    __syncthreads();
    // prefetch the next step's level-0 plane and ring while this step computes
#pragma unroll
    for (int f = 0; f < NF; ++f) load_q0(f, s + PD + QHI, pq[PI][f], s + PD <= s_last);
    load_ring(s + PD, s + PD <= s_last, pb[PI]);

    // levels TT .. 1
    for_each_ic([&](auto Mc) {
      constexpr int L = TT - decltype(Mc)::value;     // level evaluated
      constexpr int RL = pmod(R - (L - 1) * D, QN);   // rotation of level L-1's queue
      // level L is read only on the output tile widened by (TT-L) radii.
      // P::TBSKIP 1: a wave whose cells all lie outside skips the evaluation
      // (its values there are never read, 0 is pushed); 2: only levels whose
      // band is narrower than the region test it; 0 (default): every level
      // evaluates the whole region straight through.  On gfx950 the
      // branch-free form is the faster one (1-10% in 3-D, up to 18% for
      // ten 2-D levels): the per-wave test splits each plane step into
      // small blocks the scheduler cannot interleave.
      constexpr int WY = (TT - L) * P::RY, WZ = (TT - L) * P::RZ;
      constexpr bool SKIP = P::TBSKIP == 1 || (P::TBSKIP == 2 && (WY < GY || WZ < GZ));
      const int p = s - (L - 1) * D;                  // plane evaluated
      C res[NF][KC][VZ];
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const bool need = act[k] && ly[k] >= GY - WY && ly[k] < GY + OTY + WY &&
                          lz[k] + VZ > GZ - WZ && lz[k] < GZ + OTZ + WZ;
        if (!SKIP || __builtin_amdgcn_ballot_w64(need) != 0) {
          // P::ZDPP: z-neighbours of level L-1 from the neighbour lanes'
          // queue vectors (the LDS planes still serve the lanes at row ends)
          C zn[NB > 0 ? NB : 1][2 * T::ndz_max()][VZ];
          if constexpr (P::ZDPP) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              if (!T::zdpp(b)) continue;
              const int sl = (P::BDX[b] - QLO + RL) % QN;
              C up[VZ], dn[VZ];
#pragma unroll
              for (int v = 0; v < VZ; ++v) up[v] = dn[v] = q[L - 1][P::BF[b]][sl][k][v];
#pragma unroll
              for (int d = 0; d < T::ndz(b); ++d) {
#pragma unroll
                for (int v = 0; v < VZ; ++v) {
                  up[v] = lane_shift<+1>(up[v]);
                  dn[v] = lane_shift<-1>(dn[v]);
                  zn[b][d][v] = up[v];
                  zn[b][T::ndz(b) + d][v] = dn[v];
                }
              }
            }
          }
#pragma unroll
          for (int v = 0; v < VZ; ++v) {
            Cell<P, RL, KC> c{q[L - 1], ldsb + (L - 1) * LE, k, v, ly[k], lz[k] + v, zn, lz[k] / VZ, tid & 63};
            C o[NF];
            P::eval(c, o);
#pragma unroll
            for (int m = 0; m < NF; ++m) res[m][k][v] = o[m];
          }
        } else {
#pragma unroll
          for (int m = 0; m < NF; ++m)
#pragma unroll
            for (int v = 0; v < VZ; ++v) res[m][k][v] = C(0);
        }
      }
      if constexpr (L == TT) {
        const bool st = p >= xb;
#pragma unroll
        for (int m = 0; m < NF; ++m) {
          S* out = reinterpret_cast<S*>(a.out[m]);
          const Rsrc ro = make_rsrc(out + (i64)(a.h0 + (st ? p : xb)) * plane, pbytes);
#pragma unroll
          for (int k = 0; k < KC; ++k) {
            Vec<S, VZ> w;
#pragma unroll
            for (int v = 0; v < VZ; ++v) w.v[v] = Cvt<S, C>::st(res[m][k][v]);
            bstore<S, VZ, P::NTS ? BUF_NT : 0>(w, ro, st ? soff[k] : BUF_OOB);
          }
        }
      } else {
        constexpr int sl = pmod(R - (L - 1) * D - QLO, QN);
        const bool pin = w0 || (p >= 0 && p < a.n0);
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const bool zero = zero_out && !(pin && keep[k]);
#pragma unroll
          for (int m = 0; m < NF; ++m)
#pragma unroll
            for (int v = 0; v < VZ; ++v) q[L][m][sl][k][v] = zero ? C(0) : res[m][k][v];
        }
      }
    }, typename mkseq<TT>::type{});
  };

  for (int s = s_first; s <= s_last; s += UN) {
    for_each_ic([&](auto Rc) {
      if (s + decltype(Rc)::value <= s_last) step(Rc, s + decltype(Rc)::value);
    }, typename mkseq<UN>::type{});
  }
}

// ---------------------------------------------------------------------------
// two time levels per launch, one level per wave
// ---------------------------------------------------------------------------
//
// A workgroup owns an output tile of OTY x OTZ cells and a chunk of planes.
// Its threads split by whole waves into two roles:
//   * NA level-1 threads own the region (the output tile widened by GY rows
//     and GZ columns per side), KA vector cells each.  They keep a register
//     queue of QN level-0 planes of their cells, loaded from HBM, read the
//     in-plane neighbours from an LDS copy of the level-0 plane (its interior
//     written from their queues), and write level 1 into an LDS ring of NR
//     planes;
//   * the NT - NA level-2 threads own the output tile only, KB vector cells
//     each.  They keep a register queue of QN level-1 planes of their cells,
//     read from the ring, take the in-plane neighbours from the ring, store
//     level 2, and fetch the level-0 halo ring from HBM.
// Level 2 is thus evaluated on the output tile only (no lanes spent on the
// ghost ring), and no wave holds the queues of both levels, so a workgroup
// of 16 waves fits 4 waves per SIMD.
//
// Step s: the level-1 threads evaluate plane s, the level-2 threads plane
// s - D (D = QHI + 1), whose queue reaches plane s - 1.  The ring holds
// planes s - D .. s (NR = D + 1).  Level-0 planes alternate between two LDS
// buffers: plane s is read while plane s + 1 is written.  One barrier per
// step: ring slot s % NR and level-0 buffer (s + 1) % 2 were last read in
// step s - 1.  In-plane accesses must all lie in the plane being evaluated
// (dx = 0), so every LDS buffer is the plane's own.
//
// Values of level 1 outside the interior follow tblock_tile (tmode 0: 0;
// tmode 1: evaluated; wrapped axes: periodic images).

// Planes x, x + 1, ... of an array streamed one per step: the plane's
// element offset is kept incrementally (one add per plane, and a compare
// and subtract where a periodic stream axis wraps) instead of a wrap loop
// and a 64-bit multiply per plane.
struct PlaneWalk {
  i64 off;    // (h0 + px) * plane
  int px;     // plane index, wrapped into [0, n0) along a periodic stream axis
  ST_HD void init(int x, const SweepArgs& a, bool w0, i64 plane) {
    px = x;
    if (w0) {
      while (px < 0) px += a.n0;
      while (px >= a.n0) px -= a.n0;
    }
    off = (i64)(a.h0 + px) * plane;
  }
  ST_HD void next(const SweepArgs& a, bool w0, i64 plane, i64 span) {
    ++px;
    off += plane;
    if (w0 && px >= a.n0) {
      px -= a.n0;
      off -= span;
    }
  }
  ST_HD bool ok(const SweepArgs& a) const { return a.h0 + px >= 0 && a.h0 + px < a.s0; }
};

// Layout of the level-1 ring: row-major over the region, each row holding
// the NR planes (slots) one after the other, each plane the NF fields, each
// field PR elements (>= TZ).  A row then spans NR * NF * PR elements, so an
// access of the level-2 threads is a constant offset of at most about ten
// rows from their base (LDS instruction offsets reach 64 KiB).
// The level-2 threads read rows of OTZ / VZ lanes.  On gfx950 a 16-byte LDS
// read of a wave runs without bank conflicts when, within each half-wave,
// lane l's 16-byte slot (address / 16 mod 16) is l plus a constant
// (measured: rows of 16 lanes 320 B apart spent half their LDS cycles in
// conflicts, 256 B apart none).  So PR makes the row span congruent to the
// output width modulo 256 bytes, unless an output row fills whole
// half-waves.
template <class P> constexpr int ring_pitch() {
  constexpr int OTZ = P::TZ - 2 * P::GZ, NR = P::QHI + 2;
  if ((OTZ / P::VZ) % 32 == 0) return P::TZ;
  int p = P::TZ;
  while (((NR * P::NF * p - OTZ) * (int)sizeof(typename P::C)) % 256 != 0) ++p;
  return p;
}

// level-1 ring view handed to the expression: column accesses from the
// register queue (rotation R), in-plane neighbours from ring plane SL.
// `nb` points at the ring cell (row - RY, column - RZ) of my vector's first
// cell in slot 0, so every in-plane access is a non-negative constant
// offset from it (LDS instruction offsets are unsigned).
template <class P, int R, int KD, int SL, bool YA = false> struct RingCell {
  using C = typename P::C;
  using T = Tab<P>;
  static constexpr int PR = ring_pitch<P>();             // one field of one plane in a row
  static constexpr int SS = P::NF * PR;                  // one plane in a row
  static constexpr int RP = (P::QHI + 2) * SS;           // one row
  const C (*q)[T::QN][KD][P::VZ];
  const C* nb;
  int k, v;
  template <int F, int DX, int DY, int DZ> ST_HD C g() const {
    if constexpr (DY == 0 && DZ == 0) {
      static_assert(DX >= P::QLO && DX <= P::QHI, "column access outside the queue");
      return q[F][(DX - P::QLO + R) % T::QN][k][v];
    } else {
      static_assert(DX == 0, "in-plane accesses of the two-level kernel lie in the evaluated plane");
      static_assert(DY >= -P::RY && DY <= P::RY && DZ >= -P::RZ && DZ <= P::RZ, "access beyond the radius");
      if constexpr (DY == 0 && DZ > -P::VZ && DZ < P::VZ) {
        // a z-neighbour inside my own vector is in the queue
        if (v + DZ >= 0 && v + DZ < P::VZ) return q[F][(0 - P::QLO + R) % T::QN][k][v + DZ];
      } else if constexpr (YA && DZ > -P::VZ && DZ < P::VZ) {
        // a neighbour in another of my cells (adjacent rows): in the queue
        if (k + DY >= 0 && k + DY < KD && v + DZ >= 0 && v + DZ < P::VZ)
          return q[F][(0 - P::QLO + R) % T::QN][k + DY][v + DZ];
      }
      return nb[(DY + P::RY) * RP + SL * SS + F * PR + (v + DZ + P::RZ)];
    }
  }
};

template <class P>
ST_HD void tbws_tile(const SweepArgs& a, int tile_z, int tile_y, int xb, int xe) {
  using S = typename P::S;
  using C = typename P::C;
  using T = Tab<P>;
  constexpr int NT = P::NT, NA = P::NA, NBT = NT - NA, TY = P::TY, TZ = P::TZ, VZ = P::VZ, NB = P::NB;
  constexpr int NF = P::NF, QLO = P::QLO, QHI = P::QHI, QN = T::QN, D = QHI + 1, NR = D + 1;
  constexpr int GY = P::GY, GZ = P::GZ, OTY = TY - 2 * GY, OTZ = TZ - 2 * GZ;
  constexpr int NTZ = TZ / VZ, CELLS = TY * NTZ, KA = (CELLS + NA - 1) / NA;
  constexpr int ONTZ = OTZ / VZ, OCELLS = OTY * ONTZ, KB = (OCELLS + NBT - 1) / NBT;
  constexpr int ES = (int)sizeof(S);
  constexpr int LE = T::lds_elems();               // one level-0 plane (every buffer)
  constexpr int PR = ring_pitch<P>(), SS = NF * PR, RP = NR * SS;   // ring: field, plane, row
  constexpr int PD = P::TBPD < 1 ? 1 : (P::TBPD > 4 ? 4 : P::TBPD);
  // level-0 halo rings are loaded RB planes ahead of their use (RB = QHI +
  // PD: with the neighbouring tiles' queue heads of the same plane, so the
  // shared rows hit in the XCD's L2)
  constexpr int RB = P::TBRP < 1 ? PD : P::TBRP;
  constexpr int UN = lcm_i(QN, lcm_i(2, lcm_i(NR, lcm_i(PD, RB))));
  static_assert(P::TT == 2 && P::NO == NF && P::UPD == 0, "two levels; result m is the next state of field m");
  static_assert(NA % 64 == 0 && NBT > 0 && NBT % 64 == 0, "roles are whole waves");
  static_assert(!P::ZDPP, "cross-lane z-neighbours are an option of the sweep and the single-role kernel");
  // P::TBYA: a thread's cells lie on adjacent rows of one column (level 1:
  // KA rows of the region, level 2: KB rows of the output tile), so the
  // in-plane neighbours they share are read from LDS once and the ones that
  // are my own cells come from my queue
  constexpr bool YA = P::TBYA != 0;
  static_assert(!YA || (TY % KA == 0 && NA >= (TY / KA) * NTZ && OTY % KB == 0 && NBT >= (OTY / KB) * ONTZ),
                "adjacent-row cells: whole row groups, one thread per group and column");
  static_assert(QLO <= -1 && QHI >= 1, "column accesses on both sides of the plane");
  static_assert(OTY > 0 && OTZ > 0 && GZ % VZ == 0 && TZ % VZ == 0 && GY >= P::RY && GZ >= P::RZ,
                "two-level tile geometry");
  // every buffer is a plane's own (dx = 0) and the fields' buffers are distinct
  constexpr int NVB = [] {
    int m = 1;
    for (int b = 0; b < NB; ++b) {
      const int g = T::gvecs(b);
      m = (g + NBT - 1) / NBT > m ? (g + NBT - 1) / NBT : m;
    }
    return m;
  }();
  static_assert([] {
    for (int b = 0; b < NB; ++b)
      if (P::BDX[b] != 0 || !T::fromq(b)) return false;
    return true;
  }(), "in-plane accesses at dx = 0 only");
  __shared__ __attribute__((aligned(16))) C lds[2 * LE + TY * RP];
  C* const ring = lds + 2 * LE;

  const int tid = threadIdx.x;
  const int y0 = tile_y * OTY - GY, z0 = tile_z * OTZ - GZ;   // region origin, interior coordinates
  const int s2 = (int)a.s2, s1 = (int)a.s1;
  const i64 plane = a.s1 * a.s2;
  const int pbytes = (int)(plane * ES);
  const bool w0 = a.wrap & 1, w1 = a.wrap & 2, w2 = a.wrap & 4;
  const int s_first = xb + QLO, s_last = xe - 1 + D;
  const i64 span = (i64)a.n0 * plane;            // one wrap of a periodic stream axis

  // plane x of input field f (wrapped along a periodic stream axis)
  auto in_plane = [&](int f, int x, bool& ok) -> Rsrc {
    int px = x;
    if (w0) {
      while (px < 0) px += a.n0;
      while (px >= a.n0) px -= a.n0;
    }
    ok = a.h0 + px >= 0 && a.h0 + px < a.s0;
    return make_rsrc(reinterpret_cast<const S*>(a.in[f]) + (i64)(a.h0 + (ok ? px : xb)) * plane, pbytes);
  };

  for (int i = tid; i < 2 * LE + TY * RP; i += NT) lds[i] = C(0);
  __syncthreads();

  if (__builtin_amdgcn_readfirstlane(tid) < NA) {
    // ---------------- level 1 ----------------
    const int t = tid;
    int ly[KA], lz[KA], qoff[KA];
    bool act[KA], keep[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      if constexpr (YA) {
        const int grp = t / NTZ;
        act[k] = grp < TY / KA;
        ly[k] = (act[k] ? grp : TY / KA - 1) * KA + k;
        lz[k] = (t % NTZ) * VZ;
      } else {
        const int c = t + k * NA;
        act[k] = c < CELLS;
        const int cc = act[k] ? c : CELLS - 1;
        ly[k] = cc / NTZ;
        lz[k] = (cc % NTZ) * VZ;
      }
      int y = y0 + ly[k], z = z0 + lz[k];
      keep[k] = (w1 || (y >= 0 && y < a.n1)) && (w2 || (z >= 0 && z < a.n2));
      if (w1) y = wrapm(y, a.n1);
      if (w2) z = wrapm(z, a.n2);
      const bool ok = act[k] && a.h1 + y >= 0 && a.h1 + y < s1 && a.h2 + z >= 0 && a.h2 + z + VZ <= s2;
      qoff[k] = ok ? ((a.h1 + y) * s2 + a.h2 + z) * ES : BUF_OOB;
    }
    C q[NF][QN][KA][VZ];     // level-0 planes; plane p in slot (p - QLO - s_first) mod QN
    C pq[PD][NF][KA][VZ];    // queue heads loaded PD steps ahead
    auto load_q0 = [&](int f, int x, C (&dst)[KA][VZ], bool live) {
      bool ok;
      const Rsrc r = in_plane(f, x, ok);
      ok = ok && live;
#pragma unroll
      for (int k = 0; k < KA; ++k) {
        const Vec<S, VZ> w = bload<S, VZ>(r, ok ? qoff[k] : BUF_OOB);
#pragma unroll
        for (int v = 0; v < VZ; ++v) dst[k][v] = Cvt<S, C>::ld(w.v[v]);
      }
    };
    // the queue heads loaded in the step loop: planes s_first + QHI + PD, ...
    PlaneWalk wq;
    wq.init(s_first + QHI + PD, a, w0, plane);
    const S* xb_plane[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) xb_plane[f] = reinterpret_cast<const S*>(a.in[f]) + (i64)(a.h0 + xb) * plane;
    auto load_q0w = [&](int f, C (&dst)[KA][VZ], bool live) {
      const bool ok = live && wq.ok(a);
      const Rsrc r = make_rsrc(ok ? reinterpret_cast<const S*>(a.in[f]) + wq.off : xb_plane[f], pbytes);
#pragma unroll
      for (int k = 0; k < KA; ++k) {
        const Vec<S, VZ> w = bload<S, VZ>(r, ok ? qoff[k] : BUF_OOB);
#pragma unroll
        for (int v = 0; v < VZ; ++v) dst[k][v] = Cvt<S, C>::ld(w.v[v]);
      }
    };
    // the interior of level-0 plane (queue slot SL) into LDS plane set `dst`
    auto put_l0 = [&](C* dst, int sl) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int k = 0; k < KA; ++k) {
          if (act[k]) {
            C* d = dst + T::off(b) + (P::BRY[b] + ly[k]) * T::pitch(b) + T::rzv(b) + lz[k];
#pragma unroll
            for (int v = 0; v < VZ; ++v) d[v] = q[P::BF[b]][sl][k][v];
          }
        }
      }
    };
#pragma unroll
    for (int f = 0; f < NF; ++f) {
#pragma unroll
      for (int i = 0; i + 1 < QN; ++i) load_q0(f, s_first + QLO + i, q[f][i], true);
#pragma unroll
      for (int p = 0; p < PD; ++p) load_q0(f, s_first + QHI + p, pq[p][f], true);
    }
    put_l0(lds, -QLO);                   // plane s_first
    __syncthreads();

    auto step = [&](auto Rc, int s) {
      constexpr int r = decltype(Rc)::value;
      constexpr int R = r % QN, PI = r % PD;
      // queue head: plane s + QHI
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < KA; ++k)
#pragma unroll
          for (int v = 0; v < VZ; ++v) q[f][pmod(R - 1, QN)][k][v] = pq[PI][f][k][v];
#pragma unroll
      for (int f = 0; f < NF; ++f) load_q0w(f, pq[PI][f], s + PD <= s_last);
      wq.next(a, w0, plane, span);
      // level 1 at plane s
      const C* l0 = lds + (r & 1) * LE;
      C* const dst = ring + (r % NR) * SS;
      const bool zero_pl = a.tmode == 0 && !(w0 || (s >= 0 && s < a.n0));
      // every cell is evaluated before any is stored (the ring stores would
      // otherwise keep the loads of one cell from serving the next)
      C res[KA][NF][VZ];
#pragma unroll
      for (int k = 0; k < KA; ++k) {
#pragma unroll
        for (int v = 0; v < VZ; ++v) {
          Cell<P, R, KA, YA> c{q, l0, k, v, ly[k], lz[k] + v};
          C o[NF];
          P::eval(c, o);
#pragma unroll
          for (int m = 0; m < NF; ++m) res[k][m][v] = o[m];
        }
      }
#pragma unroll
      for (int k = 0; k < KA; ++k) {
        const bool zero = a.tmode == 0 && (zero_pl || !keep[k]);
        if (act[k]) {
#pragma unroll
          for (int m = 0; m < NF; ++m) {
            C* d = dst + ly[k] * RP + m * PR + lz[k];
#pragma unroll
            for (int v = 0; v < VZ; ++v) d[v] = zero ? C(0) : res[k][m][v];
          }
        }
      }
      // level-0 plane s + 1 into the other buffer
      put_l0(lds + ((r + 1) & 1) * LE, pmod(r + 1 - QLO, QN));
      __syncthreads();
    };
    for (int s = s_first; s <= s_last; s += UN) {
      for_each_ic([&](auto Rc) {
        if (s + decltype(Rc)::value <= s_last) step(Rc, s + decltype(Rc)::value);
      }, typename mkseq<UN>::type{});
    }
  } else {
    // ---------------- level 2 ----------------
    const int t = tid - NA;
    int soff[KB];
    const C* nb[KB];          // ring cell (row - RY, column - RZ) of my vector, plane slot 0
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      bool act;
      int ly, lz;
      if constexpr (YA) {
        const int grp = t / ONTZ;
        act = grp < OTY / KB;
        ly = GY + (act ? grp : OTY / KB - 1) * KB + k;
        lz = GZ + (t % ONTZ) * VZ;
      } else {
        const int c = t + k * NBT;
        act = c < OCELLS;
        const int cc = act ? c : OCELLS - 1;
        ly = GY + cc / ONTZ;
        lz = GZ + (cc % ONTZ) * VZ;
      }
      nb[k] = ring + (ly - P::RY) * RP + (lz - P::RZ);
      const int y = y0 + ly, z = z0 + lz;
      const bool mine = act && y >= 0 && y < a.n1 && z >= 0 && z < a.n2;
      soff[k] = mine ? ((a.h1 + y) * s2 + a.h2 + z) * ES : BUF_OOB;
    }
    // level-0 halo-ring vectors this thread fetches: in-plane offsets and
    // their place in the LDS plane
    int roff[NB > 0 ? NB : 1][NVB], loff[NB > 0 ? NB : 1][NVB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
#pragma unroll
      for (int j = 0; j < NVB; ++j) {
        const int i = t + j * NBT;
        bool ok = i < T::gvecs(b);
        int rr = 0, c = 0;
        T::gvec_rc(b, ok ? i : 0, rr, c);
        loff[b][j] = ok ? T::off(b) + rr * T::pitch(b) + c * VZ : -1;
        int gy = y0 - P::BRY[b] + rr, gz = z0 - T::rzv(b) + c * VZ;
        if (w1) gy = wrapm(gy, a.n1);
        if (w2) gz = wrapm(gz, a.n2);
        gy += a.h1;
        gz += a.h2;
        ok = ok && gy >= 0 && gy < s1 && gz >= 0 && gz + VZ <= s2;
        roff[b][j] = ok ? (gy * s2 + gz) * ES : BUF_OOB;
      }
    }
    C q[NF][QN][KB][VZ];                       // level-1 planes; plane p in slot (p - s_first) mod QN
    C pb[RB][NB > 0 ? NB : 1][NVB][VZ];        // level-0 halo rings loaded RB steps ahead
    auto load_ring = [&](int x, bool live, C (&dst)[NB > 0 ? NB : 1][NVB][VZ]) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        bool ok;
        const Rsrc r = in_plane(P::BF[b], x, ok);
        ok = ok && live;
#pragma unroll
        for (int j = 0; j < NVB; ++j) {
          const Vec<S, VZ> w = bload<S, VZ>(r, ok ? roff[b][j] : BUF_OOB);
#pragma unroll
          for (int v = 0; v < VZ; ++v) dst[b][j][v] = Cvt<S, C>::ld(w.v[v]);
        }
      }
    };
    // rings loaded in the step loop (planes s_first + 1 + RB, ...) and the
    // planes stored (s_first - D, ...; from xb on)
    PlaneWalk wr, wo;
    wr.init(s_first + 1 + RB, a, w0, plane);
    wo.init(s_first - D, a, false, plane);
    const i64 xb_off = (i64)(a.h0 + xb) * plane;
    auto load_ringw = [&](bool live, C (&dst)[NB > 0 ? NB : 1][NVB][VZ]) {
      const bool ok = live && wr.ok(a);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const Rsrc r = make_rsrc(reinterpret_cast<const S*>(a.in[P::BF[b]]) + (ok ? wr.off : xb_off), pbytes);
#pragma unroll
        for (int j = 0; j < NVB; ++j) {
          const Vec<S, VZ> w = bload<S, VZ>(r, ok ? roff[b][j] : BUF_OOB);
#pragma unroll
          for (int v = 0; v < VZ; ++v) dst[b][j][v] = Cvt<S, C>::ld(w.v[v]);
        }
      }
    };
    auto put_ring = [&](C* dst, const C (&src)[NB > 0 ? NB : 1][NVB][VZ]) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int j = 0; j < NVB; ++j)
          if (loff[b][j] >= 0) {
#pragma unroll
            for (int v = 0; v < VZ; ++v) dst[loff[b][j] + v] = src[b][j][v];
          }
    };
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int i = 0; i < QN; ++i)
#pragma unroll
        for (int k = 0; k < KB; ++k)
#pragma unroll
          for (int v = 0; v < VZ; ++v) q[f][i][k][v] = C(0);
#pragma unroll
    for (int p = 0; p < RB; ++p) load_ring(s_first + p, s_first + p <= s_last, pb[p]);
    put_ring(lds, pb[0]);                      // plane s_first
    load_ring(s_first + RB, s_first + RB <= s_last, pb[0]);
    __syncthreads();

    auto step = [&](auto Rc, int s) {
      constexpr int r = decltype(Rc)::value;
      constexpr int R = r % QN, PI = (r + 1) % RB;
      // queue head: level-1 plane s - 1 (ring slot (r - 1) mod NR)
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < KB; ++k)
#pragma unroll
          for (int v = 0; v < VZ; ++v)
            q[f][pmod(R - 1, QN)][k][v] = nb[k][P::RY * RP + pmod(r - 1, NR) * SS + f * PR + P::RZ + v];
      // level-0 halo ring of plane s + 1 into the other buffer, and the
      // ring of plane s + 1 + RB loaded in its place
      put_ring(lds + ((r + 1) & 1) * LE, pb[PI]);
      load_ringw(s + 1 + RB <= s_last, pb[PI]);
      wr.next(a, w0, plane, span);
      // level 2 at plane p = s - D (stored from the chunk's first plane on)
      const int p = s - D;
      const bool st = p >= xb;
      C res[NF][KB][VZ];
#pragma unroll
      for (int k = 0; k < KB; ++k)
#pragma unroll
        for (int v = 0; v < VZ; ++v) {
          RingCell<P, R, KB, pmod(r - D, NR), YA> c{q, nb[k], k, v};
          C o[NF];
          P::eval(c, o);
#pragma unroll
          for (int m = 0; m < NF; ++m) res[m][k][v] = o[m];
        }
      const i64 po = st ? wo.off : xb_off;
      wo.next(a, false, plane, span);
#pragma unroll
      for (int m = 0; m < NF; ++m) {
        const Rsrc ro = make_rsrc(reinterpret_cast<S*>(a.out[m]) + po, pbytes);
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          Vec<S, VZ> w;
#pragma unroll
          for (int v = 0; v < VZ; ++v) w.v[v] = Cvt<S, C>::st(res[m][k][v]);
          bstore<S, VZ, P::NTS ? BUF_NT : 0>(w, ro, st ? soff[k] : BUF_OOB);
        }
      }
      __syncthreads();
    };
    for (int s = s_first; s <= s_last; s += UN) {
      for_each_ic([&](auto Rc) {
        if (s + decltype(Rc)::value <= s_last) step(Rc, s + decltype(Rc)::value);
      }, typename mkseq<UN>::type{});
    }
  }
}

template <class P>
ST_HD void tblock_entry(const SweepArgs& a) {
  const int n = a.ntz * a.nty * a.nch;
  const int t = xcd_swizzle((int)blockIdx.x, n, a.nxcd);
  const int tile_z = t % a.ntz;
  const int rest = t / a.ntz;
  if constexpr (P::NA > 0) {
    const int xb = (rest / a.nty) * a.xch;
    tbws_tile<P>(a, tile_z, rest % a.nty, xb, xb + a.xch < a.n0 ? xb + a.xch : a.n0);
  } else {
    tblock_tile<P>(a, tile_z, rest % a.nty, rest / a.nty);
  }
}

// ---------------------------------------------------------------------------
// in-kernel temporal blocking (single field): T time levels per launch
// ---------------------------------------------------------------------------
//
// Level 0 is the input array (read through the caches); levels 1..T-1 live
// in LDS rings of NR = 2*RX+1 planes, level l covering the tile widened by
// (T-l)*(RY, RZ) so level l+1 can be evaluated on its own region; level T is
// the output.  The levels are skewed along the streaming axis: at step s,
// level l produces plane s - l*RX, which reads level l-1 planes
// s-(l+1)*RX .. s-(l-1)*RX, all resident in the ring.
//
// tmode 0 reproduces T back-to-back sweeps through zero-halo scratch arrays:
// level-l cells outside the interior are 0.  tmode 1 ("recompute") evaluates
// intermediate levels in the ghost zone too, which equals T steps with a
// periodic (or any copy-type) refill when the halo is at least T*r deep and
// was filled before the launch.
template <class P, int T> struct Temporal {
  using S = typename P::S;
  using C = typename P::C;
  static constexpr int RX = P::RX, RY = P::RY, RZ = P::RZ, NR = 2 * P::RX + 1;
  static constexpr int NT = P::NT, TY = P::TY, TZ = P::TZ;
  static constexpr int H(int l) { return TY + 2 * (T - l) * RY; }
  static constexpr int W(int l) { return TZ + 2 * (T - l) * RZ; }
  static constexpr int plane(int l) { return H(l) * W(l); }
  static constexpr int off(int l) { int o = 0; for (int k = 1; k < l; ++k) o += NR * plane(k); return o; }
  static constexpr int total() { int o = off(T); return o > 0 ? o : 1; }

  static ST_HD int slot(int p) { int r = p % NR; return r < 0 ? r + NR : r; }

  struct GAcc {  // level 0: the input array
    const S* u; i64 s1, s2, x, y, z;
    template <int F, int DX, int DY, int DZ> ST_HD C g() const {
      return Cvt<S, C>::ld(u[((x + DX) * s1 + (y + DY)) * s2 + (z + DZ)]);
    }
  };
  template <int L> struct RAcc {  // level L >= 1 in its LDS ring
    const C* ring; int p, ly, lz;  // (ly, lz) in level-L region coordinates
    template <int F, int DX, int DY, int DZ> ST_HD C g() const {
      return ring[off(L) + slot(p + DX) * plane(L) + (ly + DY) * W(L) + (lz + DZ)];
    }
  };

  template <int L>
  static ST_HD void level(const SweepArgs& a, C* lds, int p, int y0, int z0) {
    // produce level L (1 <= L < T) plane p over its widened region
    const int gy0 = y0 - (T - L) * RY, gz0 = z0 - (T - L) * RZ;
    C* dst = lds + off(L) + slot(p) * plane(L);
    for (int c = threadIdx.x; c < plane(L); c += NT) {
      const int ly = c / W(L), lz = c % W(L);
      const int gy = gy0 + ly, gz = gz0 + lz;
      bool valid;
      if (a.tmode == 0) {
        valid = p >= 0 && p < a.n0 && gy >= 0 && gy < a.n1 && gz >= 0 && gz < a.n2;
      } else {
        const i64 ax = (i64)p + a.h0, ay = (i64)gy + a.h1, az = (i64)gz + a.h2;
        valid = ax >= L * RX && ax < a.s0 - L * RX && ay >= L * RY && ay < a.s1 - L * RY &&
                az >= L * RZ && az < a.s2 - L * RZ;
      }
      C v = C(0);
      if (valid) {
        C o[1];
        if constexpr (L == 1) {
          GAcc acc{reinterpret_cast<const S*>(a.in[0]), a.s1, a.s2, (i64)p + a.h0, (i64)gy + a.h1,
                   (i64)gz + a.h2};
          P::eval(acc, o);
        } else {
          RAcc<L - 1> acc{lds, p, ly + RY, lz + RZ};
          P::eval(acc, o);
        }
        v = o[0];
      }
      dst[c] = v;
    }
  }

  template <int L>
  static ST_HD void levels(const SweepArgs& a, C* lds, int s, int xb, int xe, int y0, int z0) {
    if constexpr (L < T) {
      const int p = s - L * RX;
      if (p >= xb - (T - L) * RX && p <= xe - 1 + (T - L) * RX) level<L>(a, lds, p, y0, z0);
      __syncthreads();
      levels<L + 1>(a, lds, s, xb, xe, y0, z0);
    }
  }

  static ST_HD void tile(const SweepArgs& a, int tile_z, int tile_y, int chunk) {
    __shared__ C lds[total()];
    const int z0 = tile_z * TZ, y0 = tile_y * TY;
    const int xb = chunk * a.xch;
    const int xe = (xb + a.xch < a.n0) ? xb + a.xch : a.n0;
    S* out = reinterpret_cast<S*>(a.out[0]);
    for (int s = xb + (2 - T) * RX; s <= xe - 1 + T * RX; ++s) {
      levels<1>(a, lds, s, xb, xe, y0, z0);
      const int p = s - T * RX;
      if (p >= xb && p < xe) {
        for (int c = threadIdx.x; c < TY * TZ; c += NT) {
          const int ly = c / TZ, lz = c % TZ;
          const int gy = y0 + ly, gz = z0 + lz;
          if (gy < a.n1 && gz < a.n2) {
            RAcc<T - 1> acc{lds, p, ly + RY, lz + RZ};
            C o[1];
            P::eval(acc, o);
            out[(((i64)p + a.h0) * a.s1 + (gy + a.h1)) * a.s2 + (gz + a.h2)] = Cvt<S, C>::st(o[0]);
          }
        }
      }
      __syncthreads();
    }
  }
};

template <class P, int T>
ST_HD void temporal_entry(const SweepArgs& a) {
  const int n = a.ntz * a.nty * a.nch;
  const int t = xcd_swizzle((int)blockIdx.x, n, a.nxcd);
  const int tile_z = t % a.ntz;
  const int rest = t / a.ntz;
  Temporal<P, T>::tile(a, tile_z, rest % a.nty, rest / a.nty);
}

template <class P>
ST_HD void sweep_entry(const SweepArgs& a) {
  __shared__ typename P::C lds[Tab<P>::sweep_elems()];
  const int n = a.ntz * a.nty * a.nch;
  const int t = xcd_swizzle((int)blockIdx.x, n, a.nxcd);
  const int tile_z = t % a.ntz;
  const int rest = t / a.ntz;
  if (a.yslab) {
    sweep_tile<P>(a, tile_z, rest / a.nch, rest % a.nch, lds);
  } else {
    sweep_tile<P>(a, tile_z, rest % a.nty, rest / a.nty, lds);
  }
}

}  // namespace st
