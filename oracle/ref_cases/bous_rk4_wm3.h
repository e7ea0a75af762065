/* Boussinesq convection, RK4, LS-B q=6 walls, VIM xi=1.2 */
#define TARGET_PROBLEM 4
#define RK 4
#define WALL_MODEL 3
#define U_BOUNDARY_TOP 2
#define U_BOUNDARY_BOTTOM 2
#define U_BOUNDARY_LEFT 2
#define U_BOUNDARY_RIGHT 2
#define VIS_REF 2.24d-3
#define K_TH_REF 3.16d-3
#define CP_REF 1.0d0
#define XI_VIM 1.2d0
#define GRAVITY 50.0d0
#include "base.h"
