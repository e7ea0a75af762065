/* Boussinesq convection, RK2, classical SPH (sum model), LS-B q=3 walls */
#define TARGET_PROBLEM 4
#define RK 2
#define SPH_MODEL 1
#define U_BOUNDARY_TOP 2
#define U_BOUNDARY_BOTTOM 2
#define U_BOUNDARY_LEFT 2
#define U_BOUNDARY_RIGHT 2
#define VIS_REF 2.24d-3
#define K_TH_REF 3.16d-3
#define CP_REF 1.0d0
#define GRAVITY 50.0d0
#include "base.h"
