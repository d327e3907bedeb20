// Vector kernels: fill, copy/convert, and the additive halo sum over
// duplicated interface nodes.
#include "tng_internal.hpp"

namespace tng {

namespace {

template <typename T> __global__ void fill_kernel(T *p, int64_t n, T v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename TD, typename TS> __global__ void copy_kernel(TD *d, const TS *s, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (TD)s[i];
}

// communication/shell/communication_plan.hpp:48-71 exchange_and_reduce (SUM):
// every copy of an interface node receives the sum over all copies.  One
// thread per multi-copy node; copy indices are stored slot-major so that
// consecutive threads (consecutive radial levels of the same interface
// columns) read consecutive addresses.
template <typename T>
__global__ void halo_sum_kernel(T *d0, T *d1, T *d2, int ncomp, const int32_t *__restrict__ idx, int64_t n, int maxc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int32_t ci[16];
    int nc = 0;
    for (int j = 0; j < maxc; ++j) {
        const int32_t i = idx[(int64_t)j * n + t];
        if (i >= 0) ci[nc++] = i;
    }
    T *comp[3] = {d0, d1, d2};
    for (int c = 0; c < ncomp; ++c) {
        T *d = comp[c];
        T s = T(0);
        for (int j = 0; j < nc; ++j) s += d[ci[j]];
        for (int j = 0; j < nc; ++j) d[ci[j]] = s;
    }
}

unsigned grid_for(int64_t n, int bs = 256)
{
    int64_t g = (n + bs - 1) / bs;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

void launch_fill(tng_vector &v, double value, hipStream_t s)
{
    const int64_t n = v.comp_stride * v.ncomp;
    if (v.dtype == TNG_F64)
        hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(256), 0, s, v.buf.as<double>(), n, value);
    else
        hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, s, v.buf.as<float>(), n, (float)value);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_copy(tng_vector &dst, const tng_vector &src, hipStream_t s)
{
    TNG_REQUIRE(dst.mesh == src.mesh && dst.ncomp == src.ncomp, "copy: vectors must share mesh and component count");
    // component strides differ between dtypes: copy component by component
    const int64_t n = dst.mesh->nalloc;
    for (int c = 0; c < dst.ncomp; ++c) {
        if (dst.dtype == TNG_F64 && src.dtype == TNG_F64)
            hipLaunchKernelGGL((copy_kernel<double, double>), dim3(grid_for(n)), dim3(256), 0, s,
                               (double *)dst.comp(c), (const double *)src.comp(c), n);
        else if (dst.dtype == TNG_F32 && src.dtype == TNG_F64)
            hipLaunchKernelGGL((copy_kernel<float, double>), dim3(grid_for(n)), dim3(256), 0, s,
                               (float *)dst.comp(c), (const double *)src.comp(c), n);
        else if (dst.dtype == TNG_F64 && src.dtype == TNG_F32)
            hipLaunchKernelGGL((copy_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0, s,
                               (double *)dst.comp(c), (const float *)src.comp(c), n);
        else
            hipLaunchKernelGGL((copy_kernel<float, float>), dim3(grid_for(n)), dim3(256), 0, s, (float *)dst.comp(c),
                               (const float *)src.comp(c), n);
        TNG_HIP_CHECK(hipGetLastError());
    }
}

void launch_halo_sum(const tng_domain &m, tng_vector &v, hipStream_t s)
{
    if (m.halo_n == 0) return;
    TNG_REQUIRE(m.halo_maxc <= 16, "halo: more than 16 copies of one node");
    const unsigned g = (unsigned)((m.halo_n + 255) / 256);
    if (v.dtype == TNG_F64)
        hipLaunchKernelGGL(halo_sum_kernel<double>, dim3(g), dim3(256), 0, s, (double *)v.comp(0),
                           (double *)v.comp(v.ncomp > 1 ? 1 : 0), (double *)v.comp(v.ncomp > 2 ? 2 : 0), v.ncomp,
                           m.d_halo_idx.as<int32_t>(), m.halo_n, m.halo_maxc);
    else
        hipLaunchKernelGGL(halo_sum_kernel<float>, dim3(g), dim3(256), 0, s, (float *)v.comp(0),
                           (float *)v.comp(v.ncomp > 1 ? 1 : 0), (float *)v.comp(v.ncomp > 2 ? 2 : 0), v.ncomp,
                           m.d_halo_idx.as<int32_t>(), m.halo_n, m.halo_maxc);
    TNG_HIP_CHECK(hipGetLastError());
}

}  // namespace tng
