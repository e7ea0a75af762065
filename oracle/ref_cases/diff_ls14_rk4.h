/* SKIP: source/shifting/shift_interpolate_mod.f90 source/shifting/PST_mod.f90 */
/* (flang rejects shift_interpolate_mod.f90 for SPH_MODEL 5 at compile time:
   its basis writes a_vec(15:20) past q = 14, shift_interpolate_mod.f90:174;
   the diffusion test never shifts particles) */
/* Diffusion test, RK4, LS-SPH with 4th-order Taylor expansion (q = 14) */
#define TARGET_PROBLEM 1
#define RK 4
#define SPH_MODEL 5
#define U_BOUNDARY_TOP 1
#define U_BOUNDARY_BOTTOM 2
#define U_BOUNDARY_LEFT 1
#define U_BOUNDARY_RIGHT 2
#define COE_DIF 0.1d0
#include "base.h"
