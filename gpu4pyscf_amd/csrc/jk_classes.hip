// Instantiates the class kernels of one part of the angular-class list;
// compiled once per part (-DJK_PART=0..4) so that the parts build in
// parallel.  Parts: 0 = li <= 2, then li = 3 split by lj = 0, 1, 2, 3.
#include "jk_class_kernels.h"

#ifndef JK_PART
#error "JK_PART must be defined (0..4)"
#endif

static constexpr bool in_part(int li, int lj)
{
    return JK_PART == 0 ? li <= 2 : (li == 3 && lj == JK_PART - 1);
}

template <int I, int J, int K, int L>
static hipError_t launch_if_in_part(const JKBuildArgs &a, hipStream_t stream)
{
    if constexpr (in_part(I, J)) return launch_class<I, J, K, L>(a, stream);
    else return hipErrorNotSupported;
}

#define JK_CASE(I, J, K, L) \
    case ((I)*125 + (J)*25 + (K)*5 + (L)): return launch_if_in_part<I, J, K, L>(a, stream);

#include "jk_dispatch.inc"

#define JK_CAT2(a, b) a##b
#define JK_CAT(a, b) JK_CAT2(a, b)

hipError_t JK_CAT(jk_dispatch_part, JK_PART)(int key, const JKBuildArgs &a, hipStream_t stream)
{
    return dispatch_class(key, a, stream);
}
