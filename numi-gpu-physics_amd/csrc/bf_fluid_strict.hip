// Instantiations of the fluid kernels with strict arithmetic (the
// reference's expression order, no contraction; bit-identical to the FP32
// oracle) for FP32 storage.
#include "bf_fluid.h"

namespace bf {

#define BF_STRICT_DISPATCH(coll, CALL)                        \
    switch (coll) {                                           \
        case 0: return CALL(COLL_TRT);                        \
        case 1: return CALL(COLL_TRT_SYMMETRIC_LIMITER);      \
        case 2: return CALL(COLL_REGULARIZED);                \
        case 3: return CALL(COLL_RR3);                        \
        default: return hipErrorInvalidValue;                 \
    }

hipError_t launch_fluid_fast_strict(int coll, const FluidBuffers& b, const StepParams& p, hipStream_t s) {
#define CALL(C) launch_fast_t<ARITH_STRICT, C, STORE_FP32>(b, p, s)
    BF_STRICT_DISPATCH(coll, CALL)
#undef CALL
}

hipError_t launch_fluid_general_strict(int coll, const FluidBuffers& b, const StepParams& p, int general_blocks,
                                       hipStream_t s) {
#define CALL(C) launch_general_t<ARITH_STRICT, C, STORE_FP32>(b, p, general_blocks, s)
    BF_STRICT_DISPATCH(coll, CALL)
#undef CALL
}

}  // namespace bf
