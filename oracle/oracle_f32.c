/*
 * Single-precision build of the FP64 oracle (oracle.c) -- TEST INFRASTRUCTURE
 * and the CPU baseline's FP32 leg only.  Every `double` of oracle.c becomes
 * `float` and tgmath.h maps the math calls to their float versions, so the
 * same restatement runs in FP32 arithmetic (the reference CUDA kernels' type);
 * the exported functions take float arrays and scalars.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#define double float
#include "oracle.c"
