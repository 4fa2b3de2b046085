/* A CFD host's per-point calls, for measurement and tests of the per-point
 * path (not part of libcfdwm): every thread of an OpenMP team takes a share of
 * the points and calls wave_SurfElev, wave_VeloX, wave_VeloY and wave_VeloZ
 * for each, as a solver filling its cells does (CFDwavemaker.h:29-34). */
#include <omp.h>
#include <stdint.h>

#include "../../include/cfdwm.h"

CFDWM_EXPORT int cfdwm_host_sweep(int nthreads, int64_t n, const double* x, const double* y, const double* z,
                                  double t, double* out4, double* seconds) {
    if (nthreads < 1 || n < 0 || (n && (!x || !y || !z || !out4))) return -1;
    const double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        out4[4 * i] = wave_SurfElev(x[i], y[i], t);
        out4[4 * i + 1] = wave_VeloX(x[i], y[i], z[i], t);
        out4[4 * i + 2] = wave_VeloY(x[i], y[i], z[i], t);
        out4[4 * i + 3] = wave_VeloZ(x[i], y[i], z[i], t);
    }
    if (seconds) *seconds = omp_get_wtime() - t0;
    return 0;
}
