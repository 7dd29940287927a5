// One (padded DOF, padded EE count) instantiation of the JAX-convention IK
// kernel of pk_ik_jax.h, compiled once per pair with -DPK_IK_MAXA and
// -DPK_IK_NEE like pk_ik_inst.hip.
#include "pk_ik_jax.h"

#if !defined(PK_IK_MAXA) || !defined(PK_IK_NEE)
#error "compile with -DPK_IK_MAXA=<8|16> -DPK_IK_NEE=<1|2|4>"
#endif

#define PK_JAX_IK_NAME2(m, e) launch_jax_ik_##m##_##e
#define PK_JAX_IK_NAME(m, e) PK_JAX_IK_NAME2(m, e)

namespace pk {

int PK_JAX_IK_NAME(PK_IK_MAXA, PK_IK_NEE)(void* stream, const JaxIkArgs& a) {
  return launch_jax_ik<PK_IK_MAXA, PK_IK_NEE>(stream, a);
}

}  // namespace pk
