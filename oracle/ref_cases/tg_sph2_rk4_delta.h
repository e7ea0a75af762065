/* Taylor-Green, RK4, classical SPH (difference model), delta-SPH */
#define RK 4
#define SPH_MODEL 2
#define COE_DELTA_SPH 0.1d0
#include "base.h"
