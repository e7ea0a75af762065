/* Taylor-Green, RK4, Wendland C6 */
#define RK 4
#define KERNEL_TYPE 5
#include "base.h"
