// Entry points whose kernels are not written yet: they fail loudly.
#include "gx_internal.hpp"

using namespace gx;

extern "C" gx_status gx_pagerank_weighted(gx_graph*, const float*, const float*, double, double, uint32_t, float*,
                                          gx_stats*) {
  GX_API_BEGIN
  fail(GX_ERR_INVALID, "gx_pagerank_weighted: not implemented");
  GX_API_END
}
extern "C" gx_status gx_pagerank_adjoint(gx_graph*, const float*, const float*, const float*, double, double,
                                         uint32_t, float*, float*, gx_stats*) {
  GX_API_BEGIN
  fail(GX_ERR_INVALID, "gx_pagerank_adjoint: not implemented");
  GX_API_END
}
