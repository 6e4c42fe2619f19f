// SECLR4 on the GPU -- placeholder until the eigensolver kernels land.
#include "w2k_internal.h"

namespace w2k {
int seclr4(Ctx &c, double el, double eu, int nume, int *ne, int *nbelw)
{
  (void)c; (void)el; (void)eu; (void)nume;
  *ne = 0;
  *nbelw = 0;
  return W2K_EINVAL;
}
}  // namespace w2k
