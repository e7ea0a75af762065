/* Diffusion test, RK2, classical SPH Laplacian */
#define TARGET_PROBLEM 1
#define RK 2
#define SPH_MODEL 2
#define U_BOUNDARY_TOP 1
#define U_BOUNDARY_BOTTOM 2
#define U_BOUNDARY_LEFT 1
#define U_BOUNDARY_RIGHT 2
#define COE_DIF 0.1d0
#include "base.h"
