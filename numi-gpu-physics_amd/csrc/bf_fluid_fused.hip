// Instantiations of the fluid kernels with fused arithmetic, for every
// population storage and collision operator.
#include "bf_fluid.h"

namespace bf {

#define BF_FUSED_DISPATCH(storage, coll, CALL)                                \
    switch (storage * 4 + coll) {                                             \
        case 0: return CALL(COLL_TRT, STORE_FP32);                            \
        case 1: return CALL(COLL_TRT_SYMMETRIC_LIMITER, STORE_FP32);          \
        case 2: return CALL(COLL_REGULARIZED, STORE_FP32);                    \
        case 3: return CALL(COLL_RR3, STORE_FP32);                            \
        case 4: return CALL(COLL_TRT, STORE_FP16_OFFSET);                     \
        case 5: return CALL(COLL_TRT_SYMMETRIC_LIMITER, STORE_FP16_OFFSET);   \
        case 6: return CALL(COLL_REGULARIZED, STORE_FP16_OFFSET);             \
        case 7: return CALL(COLL_RR3, STORE_FP16_OFFSET);                     \
        case 8: return CALL(COLL_TRT, STORE_BF16_OFFSET);                     \
        case 9: return CALL(COLL_TRT_SYMMETRIC_LIMITER, STORE_BF16_OFFSET);   \
        case 10: return CALL(COLL_REGULARIZED, STORE_BF16_OFFSET);            \
        case 11: return CALL(COLL_RR3, STORE_BF16_OFFSET);                    \
        default: return hipErrorInvalidValue;                                 \
    }

hipError_t launch_fluid_fast_fused(int storage, int coll, const FluidBuffers& b, const StepParams& p,
                                   hipStream_t s) {
#define CALL(C, S) launch_fast_t<ARITH_FUSED, C, S>(b, p, s)
    BF_FUSED_DISPATCH(storage, coll, CALL)
#undef CALL
}

hipError_t launch_fluid_general_fused(int storage, int coll, const FluidBuffers& b, const StepParams& p,
                                      int general_blocks, hipStream_t s) {
#define CALL(C, S) launch_general_t<ARITH_FUSED, C, S>(b, p, general_blocks, s)
    BF_FUSED_DISPATCH(storage, coll, CALL)
#undef CALL
}

}  // namespace bf
