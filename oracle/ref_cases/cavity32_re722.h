/* Lid-driven cavity 32 x 32, Re 722.16 (an instance of the 64-cavity
   ensemble drawn from seed 2030), RK2 + PST, COE_H 1.2: used to show that the
   reference's own scheme diverges near step 1800 at this resolution and
   Reynolds number. */
#define TARGET_PROBLEM 3
#define RK 2
#define COE_H 1.2d0
#define NUM_X 32
#define NUM_Y 32
#define VIS_REF 1.3847262298437829d-3
#define U_BOUNDARY_TOP 3
#define U_BOUNDARY_BOTTOM 2
#define U_BOUNDARY_LEFT 2
#define U_BOUNDARY_RIGHT 2
#include "base.h"
