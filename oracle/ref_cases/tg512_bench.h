/* FFLAGS: -O2 -march=x86-64-v3 */
/* Timing case: 512^2 Taylor-Green, Re 1000, RK4 + PST, h_eff = 3.1 dx (the
   single-core sample of the CPU baseline) */
#define RK 4
#define NUM_X 512
#define NUM_Y 512
#define VIS_REF 1.0d-3
#include "base.h"
