/* Default config.h macros for the reference harness builds; a case
   header defines its overrides and then includes this file. */
#ifndef TARGET_PROBLEM
#define TARGET_PROBLEM 2
#endif
#ifndef U_BOUNDARY_TOP
#define U_BOUNDARY_TOP 1
#endif
#ifndef U_BOUNDARY_BOTTOM
#define U_BOUNDARY_BOTTOM 1
#endif
#ifndef U_BOUNDARY_LEFT
#define U_BOUNDARY_LEFT 1
#endif
#ifndef U_BOUNDARY_RIGHT
#define U_BOUNDARY_RIGHT 1
#endif
#ifndef U_TOP
#define U_TOP 1.0d0
#endif
#ifndef TEM_BOUNDARY_TOP
#define TEM_BOUNDARY_TOP 1
#endif
#ifndef TEM_BOUNDARY_BOTTOM
#define TEM_BOUNDARY_BOTTOM 1
#endif
#ifndef TEM_BOUNDARY_LEFT
#define TEM_BOUNDARY_LEFT 2
#endif
#ifndef TEM_BOUNDARY_RIGHT
#define TEM_BOUNDARY_RIGHT 2
#endif
#ifndef TEM_TOP
#define TEM_TOP 0.0d0
#endif
#ifndef TEM_BOTTOM
#define TEM_BOTTOM 1.0d0
#endif
#ifndef TEM_LEFT
#define TEM_LEFT 0.0d0
#endif
#ifndef TEM_RIGHT
#define TEM_RIGHT 0.0d0
#endif
#ifndef SAVE_NAME
#define SAVE_NAME "harness"
#endif
#ifndef READ_NAME
#define READ_NAME "new"
#endif
#ifndef OMP_THREADS
#define OMP_THREADS 1
#endif
#ifndef LEN_X
#define LEN_X 1.0d0
#endif
#ifndef LEN_Y
#define LEN_Y 1.0d0
#endif
#ifndef NUM_X
#define NUM_X 16
#endif
#ifndef NUM_Y
#define NUM_Y 16
#endif
#ifndef RK
#define RK 2
#endif
#ifndef START_STEP
#define START_STEP 1
#endif
#ifndef END_STEP
#define END_STEP 100
#endif
#ifndef WRITE_STEP
#define WRITE_STEP 100
#endif
#ifndef THRESHOLD
#define THRESHOLD 1.0d-10
#endif
#ifndef COE_CFL
#define COE_CFL 0.5d0
#endif
#ifndef COE_DIF
#define COE_DIF 0.2d0
#endif
#ifndef ZETA_RSST
#define ZETA_RSST 1.0d0
#endif
#ifndef XI_VIM
#define XI_VIM 1.0d0
#endif
#ifndef RHO_REF
#define RHO_REF 1.0d0
#endif
#ifndef VIS_REF
#define VIS_REF 1.0d-2
#endif
#ifndef K_REF
#define K_REF 1.0d2
#endif
#ifndef K_TH_REF
#define K_TH_REF 1.0d-2
#endif
#ifndef CP_REF
#define CP_REF 1.0d3
#endif
#ifndef ALPHA_REF
#define ALPHA_REF 7.1d-3
#endif
#ifndef GRAVITY
#define GRAVITY 10.0d0
#endif
#ifndef POS_PERT
#define POS_PERT 0.0d0
#endif
#ifndef TG_A
#define TG_A 2
#endif
#ifndef TG_B
#define TG_B 2
#endif
#ifndef COE_H
#define COE_H 1.55d0
#endif
#ifndef PST_C
#define PST_C 1.5d0
#endif
#ifndef COE_DELTA_SPH
#define COE_DELTA_SPH 0.0d0
#endif
#ifndef KERNEL_TYPE
#define KERNEL_TYPE 3
#endif
#ifndef SPH_MODEL
#define SPH_MODEL 3
#endif
#ifndef WALL_MODEL
#define WALL_MODEL 2
#endif
