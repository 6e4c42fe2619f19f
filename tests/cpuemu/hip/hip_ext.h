// CPU build (test infrastructure): hipExtLaunchKernelGGL as an ordinary
// emulated launch bracketed by the two events.
#pragma once
#include "hip_runtime.h"

#define hipExtLaunchKernelGGL(K, G, B, SH, S, E0, E1, FL, ...) \
  do {                                                         \
    if (E0) (void)hipEventRecord((E0), (S));                   \
    hipLaunchKernelGGL(K, G, B, SH, S, __VA_ARGS__);           \
    if (E1) (void)hipEventRecord((E1), (S));                   \
  } while (0)
