/* Default configuration of the Fortran program: the 64 x 64 Taylor-Green vortex
   (a = b = 2, Re = 100) in the closed box with free-slip ghost walls, RK2 with
   particle shifting, dt = 1e-3, 100 steps.  Macro names and meanings follow
   the reference's config.h; PERIODIC, PRECISION, DIFF_COEF, DIFF_SOURCE, DT,
   COND_MAX, JITTER_SEED, HEAT_MODE and USE_GRAPH are this program's additions
   (defaults in lssph_defaults.h). */
#define TARGET_PROBLEM 2
#define U_BOUNDARY_TOP    1
#define U_BOUNDARY_BOTTOM 1
#define U_BOUNDARY_LEFT   1
#define U_BOUNDARY_RIGHT  1
#define U_TOP 1.0d0
#define TEM_BOUNDARY_TOP    1
#define TEM_BOUNDARY_BOTTOM 1
#define TEM_BOUNDARY_LEFT   2
#define TEM_BOUNDARY_RIGHT  2
#define TEM_TOP    0.0d0
#define TEM_BOTTOM 1.0d0
#define TEM_LEFT   0.0d0
#define TEM_RIGHT  0.0d0
#define SAVE_NAME "tg64"
#define READ_NAME "new"
#define OMP_THREADS 1
#define LEN_X 1.0d0
#define LEN_Y 1.0d0
#define NUM_X 64
#define NUM_Y 64
#define RK 2
#define START_STEP 1
#define END_STEP 100
#define WRITE_STEP 10
#define THRESHOLD 1.0d-10
#define COE_CFL 0.64d0
#define COE_DIF 0.2d0
#define ZETA_RSST 1.0d0
#define XI_VIM    1.0d0
#define RHO_REF 1.0d0
#define VIS_REF 1.0d-2
#define K_REF 1.0d2
#define K_TH_REF 1.0d-2
#define CP_REF 1.0d3
#define ALPHA_REF 7.1d-3
#define GRAVITY 10.0d0
#define POS_PERT 0.0d0
#define TG_A 2
#define TG_B 2
#define COE_H 1.55d0
#define PST_C 1.5d0
#define COE_DELTA_SPH 0.0d0
#define KERNEL_TYPE 3
#define SPH_MODEL 3
#define WALL_MODEL 2
