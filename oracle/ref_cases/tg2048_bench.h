/* FFLAGS: -O2 -march=x86-64-v3 */
/* Timing case: 2048^2 Taylor-Green, Re 1000, RK4 + PST, h_eff = 3.1 dx */
#define RK 4
#define NUM_X 2048
#define NUM_Y 2048
#define VIS_REF 1.0d-3
#include "base.h"
