// engine_run.hip -- one run kernel (pdlp_run_kernel) for the iterate type
// SOVLP_T and the workgroup size SOVLP_BLOCK, both set on the compiler's
// command line, so that the six run kernels compile side by side.
#include "engine.hpp"

#if !defined(SOVLP_T) || !defined(SOVLP_BLOCK)
#error "build with -DSOVLP_T=double|float and -DSOVLP_BLOCK=64|128|256"
#endif

namespace sovh {

template <class T, int BLOCK>
void launch_run_block(const LpDev<T>* lps, int count, const StepCtl* ctl, unsigned* done,
                      int phases, hipStream_t st) {
  constexpr int TN = run_tnnz<T>(BLOCK);
  pdlp_run_kernel<T, BLOCK, TN><<<count, BLOCK, 0, st>>>(lps, ctl, done, phases);
}

template void launch_run_block<SOVLP_T, SOVLP_BLOCK>(const LpDev<SOVLP_T>*, int, const StepCtl*,
                                                     unsigned*, int, hipStream_t);

}  // namespace sovh
