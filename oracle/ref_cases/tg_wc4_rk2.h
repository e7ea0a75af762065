/* Taylor-Green, RK2, Wendland C4 */
#define RK 2
#define KERNEL_TYPE 4
#include "base.h"
