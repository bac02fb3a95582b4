// Element-wise combination kernels (Runge-Kutta stage algebra, Kahan
// updates): out[i] = E(x0[i], x1[i], ..., scalars), evaluated in the
// accumulation type A (double for the time update even when the fields are
// stored in float or bfloat16), grid-stride over every element.
#pragma once
#pragma clang fp contract(off)

#include "stencil_core.h"

namespace st {

struct PointArgs {
  const void* in[ST_MAXF];
  void* out[4];
  double sc[8];   // runtime scalars (dt-dependent coefficients)
  i64 n;          // elements
};

template <class P>
ST_HD void pointwise_entry(const PointArgs& a) {
  using S = typename P::S;
  using A = typename P::A;
  const i64 stride = (i64)gridDim.x * blockDim.x;
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    A x[P::NI];
#pragma unroll
    for (int k = 0; k < P::NI; ++k) x[k] = (A)Cvt<S, typename P::C>::ld(reinterpret_cast<const S*>(a.in[k])[i]);
    A o[P::NO];
    P::eval(x, a.sc, o);
#pragma unroll
    for (int m = 0; m < P::NO; ++m)
      reinterpret_cast<S*>(a.out[m])[i] = Cvt<S, typename P::C>::st((typename P::C)o[m]);
  }
}

}  // namespace st
