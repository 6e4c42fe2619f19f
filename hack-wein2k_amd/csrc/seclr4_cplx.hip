// SECLR4 for the complex build (placeholder)
#include "seclr4_shared.h"
namespace w2k {
int seclr4_complex(Ctx &c, double el, double eu, int nume, int *ne, int *nbelw)
{
  (void)c; (void)el; (void)eu; (void)nume; *ne = 0; *nbelw = 0;
  return W2K_EINVAL;
}
}  // namespace w2k
