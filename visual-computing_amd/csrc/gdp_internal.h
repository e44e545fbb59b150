// Host-side declarations shared by the translation units of libgdpoisson.
#pragma once
#include <hip/hip_runtime.h>

namespace gdp {

hipError_t launch_laplacian(const double* u, double* out, int m, int n, int planes,
                            double spacing, int bc, hipStream_t s);
hipError_t launch_gradient(const double* img, double* gx, double* gy, int h, int w, int planes,
                           hipStream_t s);
hipError_t launch_divergence(const double* gx, const double* gy, double* out, int h, int w,
                             int planes, hipStream_t s);
hipError_t launch_fold(const double* rhs, const double* bd, double* out, int h, int w,
                       int planes, double spacing, hipStream_t s);
hipError_t launch_smooth(double* u, const double* b, int m, int n, double spacing, int sweeps,
                         int bc, int order, hipStream_t s);
hipError_t launch_restrict(const double* f, double* c, int m, int n, int mc, int nc, int bc,
                           hipStream_t s);
hipError_t launch_prolong(const double* c, double* f, int m, int n, int mc, int nc, int bc,
                          hipStream_t s);
hipError_t launch_resnorm_partial(const double* u, const double* b, int m, int n, int planes,
                                  double spacing, int bc, double2* partial, int blocks,
                                  hipStream_t s);

void set_error(const char* fmt, ...);

}  // namespace gdp
