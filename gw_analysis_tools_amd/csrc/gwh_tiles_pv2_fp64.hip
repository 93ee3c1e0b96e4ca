/* the per-bin likelihood kernels of IMRPhenomPv2, FP64 evaluation: every tiling class (gwh_tiles.h) */
#include "gwh_tiles.h"

template void gwh_launch_tiles<GWH_IMRPHENOMPV2, GWH_FP64>(int, int, long long, const int2 *, hipStream_t, const WorkMap &,
                                      const EventDesc *, const InstCoef *, double *, const int *, const int *);
template int gwh_tiles_load<GWH_IMRPHENOMPV2, GWH_FP64>();
