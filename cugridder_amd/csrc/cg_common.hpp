// Common definitions for the MI355X (gfx950) W-stacking imager.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <cmath>

namespace cg {

// Error codes returned through the C-ABI (0 = ok, matching the reference's
// ier convention; 1 = eps clamped warning as in setup_conv_opts).
enum : int {
    CG_OK = 0,
    CG_WARN_EPS = 1,
    CG_ERR_ARG = 2,
    CG_ERR_SIZE = 3,
    CG_ERR_HIP = 4,
    CG_ERR_ALLOC = 5,
};

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kC = 299792458.0;  // speed of light (m/s)
constexpr int kMaxKW = 16;

template <typename T> struct Cx;
template <> struct Cx<float> { using type = float2; };
template <> struct Cx<double> { using type = double2; };
template <typename T> using cx_t = typename Cx<T>::type;

__host__ __device__ inline float2 mkc(float a, float b) { return make_float2(a, b); }
__host__ __device__ inline double2 mkc(double a, double b) { return make_double2(a, b); }

// a store marked non-temporal (streaming), for large outputs written in
// whole lines that no cache holds until they are read again (the w-IDFT
// planes, the forward spectrum, the point records).  Partial-line scatters
// (radix passes, band rows) measured slower with it.
__device__ inline void store_nt(float2 *p, float2 v)
{
    typedef float v2f __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2f{v.x, v.y}, reinterpret_cast<v2f *>(p));
}
__device__ inline void store_nt(double2 *p, double2 v)
{
    typedef double v2d __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2d{v.x, v.y}, reinterpret_cast<v2d *>(p));
}
__device__ inline void store_nt(float4 *p, float4 v)
{
    typedef float v4f __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(v4f{v.x, v.y, v.z, v.w}, reinterpret_cast<v4f *>(p));
}
__device__ inline void store_nt(double4 *p, double4 v)
{
    typedef double v4d __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(v4d{v.x, v.y, v.z, v.w}, reinterpret_cast<v4d *>(p));
}

// rounding of a (non-negative) value up to a multiple
__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace cg

#define CG_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) {                                             \
            cg::report_hip_error(_e, #call, __FILE__, __LINE__);            \
            return cg::CG_ERR_HIP;                                          \
        }                                                                   \
    } while (0)

// propagates a non-zero status code
#define CG_TRY(x)              \
    do {                       \
        int _r = (x);          \
        if (_r) return _r;     \
    } while (0)

namespace cg {
void report_hip_error(hipError_t e, const char *what, const char *file, int line);
}
