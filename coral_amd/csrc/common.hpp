// Shared device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace coral {

// A failed launch becomes an exception; the C ABI returns it as an error code
// with the message.  (hipGetLastError clears the error, so it is read here,
// once, right after the launch.)
inline void check_launch(const char *where) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        throw std::runtime_error(std::string("kernel launch failed in ") + where + ": " + hipGetErrorString(e));
}

// Largest dynamic LDS a workgroup may request on the current device.
inline size_t lds_limit_current() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        v = 65536;
    return size_t(v);
}

// Opts a kernel into `bytes` of dynamic LDS (needed above 64 KiB); throws when
// the runtime refuses.
inline void set_dynamic_lds(const void *fn, size_t bytes, const char *where) {
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
    if (e != hipSuccess)
        throw std::runtime_error(std::string("dynamic LDS of ") + std::to_string(bytes) + " bytes refused in " + where +
                                 ": " + hipGetErrorString(e));
}

__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// 1/z with scaling (Smith's method) so that huge or tiny entries do not overflow
__host__ __device__ inline double2 crecip(double2 z) {
    if (fabs(z.x) >= fabs(z.y)) {
        double r = z.y / z.x, d = z.x + z.y * r;
        return make_double2(1.0 / d, -r / d);
    }
    double r = z.x / z.y, d = z.x * r + z.y;
    return make_double2(r / d, -1.0 / d);
}

// single-precision complex (the FP32 transform chain)
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// Complex element types of the transform kernels: the real type and a maker.
template <class C>
struct CxT;
template <>
struct CxT<double2> {
    using R = double;
    __host__ __device__ static inline double2 make(double x, double y) { return make_double2(x, y); }
};
template <>
struct CxT<float2> {
    using R = float;
    __host__ __device__ static inline float2 make(float x, float y) { return make_float2(x, y); }
};
template <class C>
__host__ __device__ inline C mkc(typename CxT<C>::R x, typename CxT<C>::R y) {
    return CxT<C>::make(x, y);
}

// (i k)^a for integer a >= 0: the wavenumber factor px^a of
// recipe_to_prefactors_array (PL_IMEX_timestepping.f90:1558-1574); k^a is
// formed by repeated multiplication, as the complex power of (0, k) is.
__host__ __device__ inline double2 ipow(double k, int a) {
    double m = 1.0;
    for (int t = 0; t < a; ++t) m *= k;
    switch (a & 3) {
    case 0: return make_double2(m, 0.0);
    case 1: return make_double2(0.0, m);
    case 2: return make_double2(-m, 0.0);
    default: return make_double2(0.0, -m);
    }
}

__host__ __device__ inline double2 monomial(int a, int b, double kx, double ky) {
    return cmul(ipow(kx, a), ipow(ky, b));
}

// Geometry of the kept wavenumbers: s = ix * nky + jy.
struct KeptModes {
    int nkx, nky, NY;
    double kx_box, ky_box;
    __device__ inline int kx_int(int s) const { return s / nky; }
    __device__ inline int ky_int(int s) const {
        int jy = s % nky;
        return jy < NY / 2 ? jy : jy - (NY - 1);
    }
    __device__ inline double kx(int s) const { return double(kx_int(s)) * kx_box; }
    __device__ inline double ky(int s) const { return double(ky_int(s)) * ky_box; }
};

}  // namespace coral
