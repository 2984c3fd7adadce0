// spectral.hip -- spectral sampling helpers of a polychromatic source on CDNA4.
//
// Replaces ch_opt.f90:28-152 (sample, lin_inter_1D, lin_inter_2D, search_1D,
// search_2D, mk_cdf).  A spectrum is the reference's table array(len, 2),
// column-major: wavelengths (ascending) in the first column, values in the
// second.  mk_cdf runs on the host (len terms); the per-draw work -- one
// uniform, a bisection over the CDF and the inverse interpolation -- runs one
// draw per lane, and so does the per-wavelength property lookup.
//
// The arithmetic is the reference's, in FP64 and in its evaluation order (no
// contraction), so a draw equals the reference's sample() for the same
// uniform.  The uniform comes from Philox4x32-10 (key = seed, counter =
// (draw index, 0, stream)) in place of ran2's sequential stream.
#include <cstring>
#include <vector>

#include "tam_internal.h"

#pragma clang fp contract(off)

namespace {

// search_1D / search_2D (ch_opt.f90:97-135): bisection on an ascending
// array; value > a(middle) moves the lower end; a(len) itself is never
// compared.  1-based result.
__device__ __forceinline__ int bisect(const double* __restrict__ a, int len, double value)
{
    int nup = len, nlow = 1;
    while (nup - nlow > 1) {
        const int middle = (nup + nlow) >> 1;
        if (value > a[middle - 1]) nlow = middle;
        else nup = middle;
    }
    return nlow;
}

__global__ __launch_bounds__(256) void sample_kernel(const double* __restrict__ table,
                                                     const double* __restrict__ cdf, int len, long n,
                                                     uint32_t k0, uint32_t k1, uint32_t stream,
                                                     double* __restrict__ wave)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint4 r = philox4x32_10(make_uint4((uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, stream), k0, k1);
    const double value = (double)u01(r.x);
    // sample: search_1D over the CDF, then lin_inter_1D (ch_opt.f90:28-76):
    // cdf(j) is the CDF at row j+1, so [cdf(nlow), cdf(nlow+1)] maps onto
    // rows nlow+1 .. nlow+2
    const int nlow = bisect(cdf, len, value);
    wave[i] = table[nlow] + (table[nlow + 1] - table[nlow]) * (value - cdf[nlow - 1]) / (cdf[nlow] - cdf[nlow - 1]);
}

__global__ __launch_bounds__(256) void interp_kernel(const double* __restrict__ table, int len,
                                                     const double* __restrict__ x, long n, double* __restrict__ y)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    // search_2D on the wavelengths, then lin_inter_2D (ch_opt.f90:79-94)
    const int nlow = bisect(table, len, v);
    const double* val = table + len;
    y[i] = val[nlow - 1] + (val[nlow] - val[nlow - 1]) * (v - table[nlow - 1]) / (table[nlow] - table[nlow - 1]);
}

unsigned blocks_for(long n) { return (unsigned)((n + 255) / 256); }
constexpr int64_t kMaxDraws = (int64_t)0x7fffffff * 256;    // one lane per element, at most 2^31 - 1 blocks

// device buffers of one call, freed on every path out
struct Bufs {
    std::vector<void*> p;
    ~Bufs()
    {
        for (void* q : p) hipFree(q);
    }
    template <class T>
    hipError_t get(T** out, size_t count)
    {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) p.push_back(q);
        *out = (T*)q;
        return e;
    }
};

int hip_err(const char* what, hipError_t e)
{
    return tam_fail_noctx(TAM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int check_table(const double* table, int32_t len, int32_t min_len)
{
    if (!table) return tam_fail_noctx(TAM_ERR_ARG, "null spectrum table");
    if (len < min_len) return tam_fail_noctx(TAM_ERR_ARG, "spectrum table too short");
    for (int32_t i = 1; i < len; ++i)
        if (!(table[i] > table[i - 1])) return tam_fail_noctx(TAM_ERR_ARG, "spectrum wavelengths must ascend");
    return TAM_OK;
}

struct DevGuard {                       // the caller's current device, restored on return
    int old = 0;
    explicit DevGuard(int d)
    {
        hipGetDevice(&old);
        hipSetDevice(d);
    }
    ~DevGuard() { hipSetDevice(old); }
};

}  // namespace

extern "C" {

int tam_spectrum_cdf(const double* table, int32_t len, double* cdf)
{
    const int rc = check_table(table, len, 2);
    if (rc != TAM_OK) return rc;
    if (!cdf) return tam_fail_noctx(TAM_ERR_ARG, "null cdf");
    // mk_cdf (ch_opt.f90:138-152): the trapezoid sums of rows 1..j+1, each
    // added in row order; every entry (the last one as passed) is then
    // divided by cdf(len-1)
    const double* x = table;
    const double* y = table + len;
    double summ = 0.0;
    for (int32_t j = 1; j <= len - 1; ++j) {
        summ = summ + 0.5 * (y[j] + y[j - 1]) * (x[j] - x[j - 1]);
        cdf[j - 1] = summ;
    }
    const double norm = cdf[len - 2];
    if (!(norm > 0.0)) return tam_fail_noctx(TAM_ERR_ARG, "spectrum has no positive area");
    for (int32_t j = 0; j < len; ++j) cdf[j] = cdf[j] / norm;
    return TAM_OK;
}

int tam_spectrum_sample(int32_t device, const double* table, const double* cdf, int32_t len, int64_t n,
                        uint64_t seed, uint32_t stream, double* wave)
{
    // lin_inter_1D reads row nlow+2 <= len: at least three rows
    int rc = check_table(table, len, 3);
    if (rc != TAM_OK) return rc;
    if (!cdf || (n > 0 && !wave) || n < 0 || n > kMaxDraws)
        return tam_fail_noctx(TAM_ERR_ARG, "bad cdf, output or count");
    for (int32_t j = 1; j < len - 1; ++j)
        if (cdf[j] < cdf[j - 1]) return tam_fail_noctx(TAM_ERR_ARG, "cdf must not decrease");
    if (n == 0) return TAM_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return tam_fail_noctx(TAM_ERR_NODEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return tam_fail_noctx(TAM_ERR_ARG, "device ordinal out of range");
    DevGuard g(device);
    Bufs b;
    double *dt = nullptr, *dc = nullptr, *dw = nullptr;
    hipError_t e = b.get(&dt, 2 * (size_t)len);
    if (e == hipSuccess) e = b.get(&dc, (size_t)len);
    if (e == hipSuccess) e = b.get(&dw, (size_t)n);
    if (e == hipSuccess) e = hipMemcpy(dt, table, sizeof(double) * 2 * (size_t)len, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dc, cdf, sizeof(double) * (size_t)len, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_err("spectrum sample set-up", e);
    hipLaunchKernelGGL(sample_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, dt, dc, (int)len, (long)n,
                       (uint32_t)seed, (uint32_t)(seed >> 32), stream, dw);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(wave, dw, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err("spectrum sample", e);
    return TAM_OK;
}

int tam_spectrum_interp(int32_t device, const double* table, int32_t len, const double* x, int64_t n, double* y)
{
    int rc = check_table(table, len, 2);
    if (rc != TAM_OK) return rc;
    if ((n > 0 && (!x || !y)) || n < 0 || n > kMaxDraws)
        return tam_fail_noctx(TAM_ERR_ARG, "bad input, output or count");
    if (n == 0) return TAM_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return tam_fail_noctx(TAM_ERR_NODEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return tam_fail_noctx(TAM_ERR_ARG, "device ordinal out of range");
    DevGuard g(device);
    Bufs b;
    double *dt = nullptr, *dx = nullptr, *dy = nullptr;
    hipError_t e = b.get(&dt, 2 * (size_t)len);
    if (e == hipSuccess) e = b.get(&dx, (size_t)n);
    if (e == hipSuccess) e = b.get(&dy, (size_t)n);
    if (e == hipSuccess) e = hipMemcpy(dt, table, sizeof(double) * 2 * (size_t)len, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_err("spectrum interp set-up", e);
    hipLaunchKernelGGL(interp_kernel, dim3(blocks_for(n)), dim3(256), 0, 0, dt, (int)len, dx, (long)n, dy);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(y, dy, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err("spectrum interp", e);
    return TAM_OK;
}

}  // extern "C"
