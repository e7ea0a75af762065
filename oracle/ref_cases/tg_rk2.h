/* Taylor-Green, RK2, Wendland C2, LS-B q=3 free-slip walls */
#define RK 2
#define COE_CFL 0.64d0
#include "base.h"
