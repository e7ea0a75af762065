/* Taylor-Green on a 1 x 0.75 box, RK2, quintic spline (h_eff = 3h) */
#define RK 2
#define KERNEL_TYPE 2
#define COE_H 1.0d0
#define LEN_Y 0.75d0
#define NUM_Y 12
#include "base.h"
