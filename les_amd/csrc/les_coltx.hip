// les_coltx.hip -- the column-FFT kernels at the tile widths les_autotune can
// choose besides the default one (les_xform.h): FP32 8 and 32 columns,
// FP64 4 and 16 columns per block.  A width whose block or LDS tile does not
// fit the gfx950 limits (1024 threads, 160 KB) has no variant for that pair
// of lengths; the default then serves it.  The per-column arithmetic is the
// same at every width, so results do not depend on the choice.
#include "les_xform.h"

namespace les {

template <typename T, int TX>
static bool col_tx_dispatch(les_ctx* c, const cx<T>* in, const cx<T>* in2, cx<T>* o0, cx<T>* o1, cx<T>* o2,
                            const ColArgs& a, long nplanes, int ncols, const T* G0, const T* G1, const T* G2) {
#define LES_CASE(I, O)                                                                      \
  if (a.ny_in == I && a.ny_out == O) {                                                      \
    if constexpr (col_tile_fits<T, TX, I, O>()) {                                           \
      launch_col_n<T, TX, I, O>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);    \
      return true;                                                                          \
    }                                                                                       \
    return false;                                                                           \
  }
  LES_COL_PAIRS(LES_CASE)
#undef LES_CASE
  return false;  // lengths without a compile-time transform keep the default tile
}

template <>
bool launch_col_alt<float>(les_ctx* c, int tx, const cx<float>* in, const cx<float>* in2, cx<float>* o0,
                           cx<float>* o1, cx<float>* o2, const ColArgs& a, long nplanes, int ncols, const float* G0,
                           const float* G1, const float* G2) {
  switch (tx) {
    case 8: return col_tx_dispatch<float, 8>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
    case 32: return col_tx_dispatch<float, 32>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
    default: return false;
  }
}
template <>
bool launch_col_alt<double>(les_ctx* c, int tx, const cx<double>* in, const cx<double>* in2, cx<double>* o0,
                            cx<double>* o1, cx<double>* o2, const ColArgs& a, long nplanes, int ncols,
                            const double* G0, const double* G1, const double* G2) {
  switch (tx) {
    case 4: return col_tx_dispatch<double, 4>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
    case 16: return col_tx_dispatch<double, 16>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
    default: return false;
  }
}

}  // namespace les
