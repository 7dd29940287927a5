// One (padded DOF, padded EE count) instantiation of the LM kernels of
// pk_ik.h.  The build compiles this file once per pair, with -DPK_IK_MAXA and
// -DPK_IK_NEE, so the heavy instantiations compile in parallel.
#include "pk_ik.h"

#if !defined(PK_IK_MAXA) || !defined(PK_IK_NEE)
#error "compile with -DPK_IK_MAXA=<8|16> -DPK_IK_NEE=<1|2|4>"
#endif

#define PK_IK_NAME2(m, e) launch_ik_##m##_##e
#define PK_IK_NAME(m, e) PK_IK_NAME2(m, e)

namespace pk {

int PK_IK_NAME(PK_IK_MAXA, PK_IK_NEE)(void* stream, const IkArgs& a, bool batched, int grid_waves) {
  return launch_ik<PK_IK_MAXA, PK_IK_NEE>(stream, a, batched, grid_waves);
}

}  // namespace pk
