/* Taylor-Green, RK2, LS-SPH with 3rd-order Taylor expansion (q = 9) */
#define RK 2
#define SPH_MODEL 4
#include "base.h"
