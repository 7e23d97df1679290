// engine_inst.hip -- the engines' host code and kernels for one iterate type,
// SOVLP_T (double or float), set on the compiler's command line: the two
// precisions compile as separate translation units, side by side.
#include "engine.hpp"

#ifndef SOVLP_T
#error "build with -DSOVLP_T=double or -DSOVLP_T=float"
#endif

namespace sovh {
template int pick_width<SOVLP_T>(const sovlp_lp*, int, int*);
template size_t single_ws_T<SOVLP_T>(const sovlp_lp*, long long, int);
template int single_create_T<SOVLP_T>(const sovlp_lp*, const sovlp_params*, void*, int64_t,
                                      int64_t, int, hipStream_t, sovlp_pdlp*);
template size_t batch_ws_T<SOVLP_T>(int, const sovlp_lp*, long long);
template int batch_create_T<SOVLP_T>(int, const sovlp_lp*, const sovlp_params*, const int32_t*,
                                     void*, int64_t, int64_t, hipStream_t, sovlp_batch*);
template int batch_reset_T<SOVLP_T>(sovlp_batch*, hipStream_t);
template size_t server_ws_T<SOVLP_T>(int, int);
template size_t server_lp_ws_T<SOVLP_T>(const sovlp_lp*, long long);
template int server_create_T<SOVLP_T>(int, int, const sovlp_params*, int, int, void*, int64_t,
                                      hipStream_t, sovlp_server*);
template int server_admit_T<SOVLP_T>(sovlp_server*, int, const sovlp_lp*, const double*,
                                     const double*, double, void*, int64_t, int64_t,
                                     hipStream_t);
}  // namespace sovh
