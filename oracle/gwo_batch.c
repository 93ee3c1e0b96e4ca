/* Reference CPU path over a batch of instances (TEST INFRASTRUCTURE / CPU baseline). */
#include <omp.h>
#include "gwo_internal.h"

int gwo_network_loglike_batch(const gwo_params *ps, int n_inst, int model, const double *f, int n,
                              const double *const *data, const double *const *psd, const int *dets, int ndet,
                              int nthreads, double *out)
{
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int i = 0; i < n_inst; i++) out[i] = gwo_network_loglike(&ps[i], model, f, n, data, psd, dets, ndet);
    return nthreads;
}
