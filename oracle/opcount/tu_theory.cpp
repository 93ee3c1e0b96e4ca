/* The oracle's gwo_theory.c on the counting scalar (test and measurement infrastructure). */
#include "counted.h"
#define double cd
#include "../gwo_theory.c"
