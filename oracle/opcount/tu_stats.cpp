/* The oracle's gwo_stats.c on the counting scalar (test and measurement infrastructure). */
#include "counted.h"
#define double cd
#include "../gwo_stats.c"
