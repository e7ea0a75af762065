/* Taylor-Green, RK4, delta-SPH density diffusion, RSST zeta=1.5 */
#define RK 4
#define COE_DELTA_SPH 0.1d0
#define ZETA_RSST 1.5d0
#include "base.h"
