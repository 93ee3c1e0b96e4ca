/* The oracle's gwo_pv2.c on the counting scalar (test and measurement infrastructure). */
#include "counted.h"
#define double cd
#include "../gwo_pv2.c"
