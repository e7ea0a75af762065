/* SKIP: source/boundary/interpolate_LSSPH_B_mod.f90 */
/* (flang rejects its q=2 placeholder branch at compile time; WALL_MODEL 1
   never calls it) */
/* Taylor-Green, RK2, cubic spline, CSPH walls */
#define RK 2
#define KERNEL_TYPE 1
#define WALL_MODEL 1
#include "base.h"
