// les_coltx.hip -- the column-FFT kernels les_autotune can choose besides the
// default (les_xform.h): other tile widths (FP32 8 and 32 columns, FP64 4 and
// 16 per block) with the radix-4 plan, and the radix-8 plan (three LDS passes
// for 256 and 384 points instead of four and five) at FP32 8/16, FP64 4/8.
// A variant whose block or LDS tile does not fit the gfx950 limits (1024
// threads, 160 KB) does not exist for that pair of lengths; the default then
// serves it.  The tile width does not change the per-column arithmetic; the
// plan changes its rounding (another order of the same DFT).
#include "les_xform.h"

namespace les {

template <typename T, int TX, int PLAN>
static bool col_tx_dispatch(les_ctx* c, const cx<T>* in, const cx<T>* in2, cx<T>* o0, cx<T>* o1, cx<T>* o2,
                            const ColArgs& a, long nplanes, int ncols, const T* G0, const T* G1, const T* G2) {
#define LES_CASE(I, O)                                                                          \
  if (a.ny_in == I && a.ny_out == O) {                                                          \
    if constexpr (col_tile_fits<T, TX, I, O>()) {                                               \
      launch_col_n<T, TX, I, O, PLAN>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);  \
      return true;                                                                              \
    }                                                                                           \
    return false;                                                                               \
  }
  LES_COL_PAIRS(LES_CASE)
#undef LES_CASE
  return false;  // lengths without a compile-time transform keep the default tile
}

// (tile width, plan) variants: plan 0 at the non-default widths, plan 2
// (radix-8 passes) at the two narrower widths of each precision
template <>
bool launch_col_alt<float>(les_ctx* c, int tx, int plan, const cx<float>* in, const cx<float>* in2, cx<float>* o0,
                           cx<float>* o1, cx<float>* o2, const ColArgs& a, long nplanes, int ncols, const float* G0,
                           const float* G1, const float* G2) {
  if (plan == 0 && tx == 8) return col_tx_dispatch<float, 8, 0>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  if (plan == 0 && tx == 32) return col_tx_dispatch<float, 32, 0>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  if (plan == 2 && tx == 8) return col_tx_dispatch<float, 8, 2>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  if (plan == 2 && tx == 16) return col_tx_dispatch<float, 16, 2>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  return false;
}
template <>
bool launch_col_alt<double>(les_ctx* c, int tx, int plan, const cx<double>* in, const cx<double>* in2,
                            cx<double>* o0, cx<double>* o1, cx<double>* o2, const ColArgs& a, long nplanes,
                            int ncols, const double* G0, const double* G1, const double* G2) {
  if (plan == 0 && tx == 4) return col_tx_dispatch<double, 4, 0>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  if (plan == 0 && tx == 16) return col_tx_dispatch<double, 16, 0>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  if (plan == 2 && tx == 4) return col_tx_dispatch<double, 4, 2>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  if (plan == 2 && tx == 8) return col_tx_dispatch<double, 8, 2>(c, in, in2, o0, o1, o2, a, nplanes, ncols, G0, G1, G2);
  return false;
}

}  // namespace les
