/* Taylor-Green, RK2, classical SPH (pressure gradient sum model) */
#define RK 2
#define SPH_MODEL 1
#include "base.h"
