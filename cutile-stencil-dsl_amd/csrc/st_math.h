// Element-wise helpers shared by the JIT-built device code and by the
// host-side (g++) expression harness used in tests.  Every helper is an
// exact IEEE operation or a select, so a stencil expression evaluates
// bit-identically on the GPU and in NumPy.
#pragma once

#ifndef ST_HD
#define ST_HD
#endif

namespace st {

// NumPy's maximum/minimum propagate NaN (elementwise.py:25-52 swaps the
// builtins for np.maximum/np.minimum); fmax/fmin would drop it.
template <class T> ST_HD T maxp(T a, T b) {
  return (a != a) ? a : ((b != b) ? b : (a >= b ? a : b));
}
template <class T> ST_HD T minp(T a, T b) {
  return (a != a) ? a : ((b != b) ? b : (a <= b ? a : b));
}
template <class T> ST_HD T absp(T a) { return a < T(0) ? -a : (a == T(0) ? T(0) : a); }

}  // namespace st
