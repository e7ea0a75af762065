/* Lid-driven cavity, RK2, COE_H 1.2 (3 wall layers) */
#define TARGET_PROBLEM 3
#define RK 2
#define COE_H 1.2d0
#define U_BOUNDARY_TOP 3
#define U_BOUNDARY_BOTTOM 2
#define U_BOUNDARY_LEFT 2
#define U_BOUNDARY_RIGHT 2
#include "base.h"
