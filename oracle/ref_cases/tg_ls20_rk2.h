/* Taylor-Green, RK2, LS-SPH with 5th-order Taylor expansion (q = 20) */
#define RK 2
#define SPH_MODEL 6
#include "base.h"
