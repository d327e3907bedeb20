// Vector kernels: fill, copy/convert, and the additive halo sum over
// duplicated interface nodes.
#include <cfloat>

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
// One thread per (component, shared node): component-major, so consecutive
// threads read consecutive entries of the copy table; the copies are loaded
// before they are summed (in copy order).
template <typename T>
__global__ void halo_sum_kernel(T *d0, T *d1, T *d2, int ncomp, const int32_t *__restrict__ idx, int64_t n, int maxc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * ncomp) return;
    const int c = (int)(t / n);
    const int64_t g = t - (int64_t)c * n;
    T *d = c == 0 ? d0 : (c == 1 ? d1 : d2);
    int32_t ci[16];
    T v[16];
    int nc = 0;
    for (int j = 0; j < maxc; ++j) {
        const int32_t i = idx[(int64_t)j * n + g];
        if (i >= 0) ci[nc++] = i;
    }
    for (int j = 0; j < nc; ++j) v[j] = d[ci[j]];
    T s = T(0);
    for (int j = 0; j < nc; ++j) s += v[j];
    for (int j = 0; j < nc; ++j) d[ci[j]] = s;
}

// Shared node copies of a fused Chebyshev sub-step: A x summed over the
// copies (halo_sum_kernel's order), then the update of every copy with the
// rounding of the separate update kernel.
template <typename T>
__global__ void halo_cheb_kernel(const Comp3<T> ax, const Comp3<T> b, const Comp3<T> invd, const Comp3<T> d,
                                 const Comp3<T> x, const Comp3<T> xo, T az, T ad, int read_d,
                                 const int32_t *__restrict__ idx, int64_t n, int maxc)
{
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int32_t ci[16];
    int nc = 0;
    for (int j = 0; j < maxc; ++j) {
        const int32_t i = idx[(int64_t)j * n + t];
        if (i >= 0) ci[nc++] = i;
    }
    for (int c = 0; c < 3; ++c) {
        T sum = T(0);
        for (int j = 0; j < nc; ++j) sum += ax.p[c][ci[j]];
        for (int j = 0; j < nc; ++j) {
            const int32_t i = ci[j];
            const T r = b.p[c][i] - sum;
            const T z = r * invd.p[c][i];
            T tt = az * z;
            if (read_d) tt = fma(ad, d.p[c][i], tt);
            d.p[c][i] = tt;
            xo.p[c][i] = x.p[c][i] + tt;
        }
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
    const int64_t n = v.mesh->nalloc;
    for (int c = 0; c < v.ncomp; ++c) {
        if (v.dtype == TNG_F64)
            hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(256), 0, s, (double *)v.comp(c), n, value);
        else
            hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(256), 0, s, (float *)v.comp(c), n,
                               (float)value);
        TNG_HIP_CHECK(hipGetLastError());
    }
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
    const unsigned g = (unsigned)((m.halo_n * v.ncomp + 255) / 256);
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

void launch_halo_cheb(const tng_domain &m, const tng_vector &ax, const tng_vector &b, const tng_vector &invd,
                      tng_vector &d, const tng_vector &x, tng_vector &x_out, double az, double ad, bool read_d,
                      hipStream_t s)
{
    if (m.halo_n == 0) return;
    TNG_REQUIRE(m.halo_maxc <= 16, "halo: more than 16 copies of a node");
    const unsigned g = (unsigned)((m.halo_n + 255) / 256);
    auto run = [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL(halo_cheb_kernel<T>, dim3(g), dim3(256), 0, s, comps<T>(ax), comps<T>(b), comps<T>(invd),
                           comps<T>(d), comps<T>(x), comps<T>(x_out), (T)az, (T)ad, read_d ? 1 : 0,
                           m.d_halo_idx.as<int32_t>(), m.halo_n, m.halo_maxc);
    };
    if (x.dtype == TNG_F64)
        run(double());
    else
        run(float());
    TNG_HIP_CHECK(hipGetLastError());
}

}  // namespace tng

// ----------------------------------------------------------------------------
// BLAS-1 over all components (grid_operations.hpp:11-273) and reductions over
// OWNED node copies (grid_operations.hpp:275-733).  Reductions are two
// launches and deterministic: per-block partials, then one block sums them
// in a fixed order.
// ----------------------------------------------------------------------------
namespace tng {

namespace {

// The streaming kernels move V consecutive elements of a component per
// access: 16 bytes (V = 16 / sizeof(T)) when every pointer is 16-byte aligned,
// else V = 1; the n % V trailing elements take V = 1.  The pointers may alias
// as far as the compiler knows, so loads are not moved past stores by
// themselves: a chunk either loads all components before its first store
// (kLoadFirst: faster for FP32 vectors and for the Gram-Schmidt pass) or
// finishes one component before the next (measured faster for the FP64
// Chebyshev update and linear combination).
template <typename T, int V> struct Pack {
    T e[V];
};

template <int V, typename T> __device__ __forceinline__ Pack<T, V> ldv(const T *p, int64_t i)
{
    Pack<T, V> o;
    if constexpr (V == 1) {
        o.e[0] = p[i];
    } else {
        typedef T VT __attribute__((ext_vector_type(V)));
        const VT v = reinterpret_cast<const VT *>(p)[i];
#pragma unroll
        for (int k = 0; k < V; ++k) o.e[k] = v[k];
    }
    return o;
}

template <int V, typename T> __device__ __forceinline__ void stv(T *p, int64_t i, const Pack<T, V> &o)
{
    if constexpr (V == 1) {
        p[i] = o.e[0];
    } else {
        typedef T VT __attribute__((ext_vector_type(V)));
        VT v;
#pragma unroll
        for (int k = 0; k < V; ++k) v[k] = o.e[k];
        reinterpret_cast<VT *>(p)[i] = v;
    }
}

template <typename T> constexpr int kVec = 16 / (int)sizeof(T);
template <typename T> constexpr bool kLoadFirst = sizeof(T) == 4;

template <typename T> bool aligned16(const Comp3<T> &v, int ncomp)
{
    for (int c = 0; c < ncomp; ++c)
        if (reinterpret_cast<uintptr_t>(v.p[c]) & 15) return false;
    return true;
}

template <int V, typename T>
__device__ __forceinline__ void lincomb_chunk(const Comp3<T> &y, const Comp3<T> &x0, const Comp3<T> &x1,
                                              const Comp3<T> &x2, int nx, T a0, T a1, T a2, T a3, int ncomp,
                                              int64_t i)
{
    Pack<T, V> v[3][3];
    auto load = [&](int c) {
        if (nx > 0) v[c][0] = ldv<V>(x0.p[c], i);
        if (nx > 1) v[c][1] = ldv<V>(x1.p[c], i);
        if (nx > 2) v[c][2] = ldv<V>(x2.p[c], i);
    };
    auto update = [&](int c) {
        Pack<T, V> o;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            T t = a0;
            if (nx > 0) t += a1 * v[c][0].e[k];
            if (nx > 1) t += a2 * v[c][1].e[k];
            if (nx > 2) t += a3 * v[c][2].e[k];
            o.e[k] = t;
        }
        stv<V>(y.p[c], i, o);
    };
    if constexpr (kLoadFirst<T>) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (c < ncomp) load(c);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (c < ncomp) update(c);
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (c < ncomp) {
                load(c);
                update(c);
            }
    }
}

template <typename T, int V>
__global__ __launch_bounds__(256) void lincomb_kernel(const Comp3<T> y, const Comp3<T> x0, const Comp3<T> x1,
                                                      const Comp3<T> x2, int nx, double c0, double c1, double c2,
                                                      double c3, int ncomp, int64_t n)
{
    const T a0 = (T)c0, a1 = (T)c1, a2 = (T)c2, a3 = (T)c3;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nv = n / V;
    for (int64_t i = t0; i < nv; i += stride) lincomb_chunk<V>(y, x0, x1, x2, nx, a0, a1, a2, a3, ncomp, i);
    if constexpr (V > 1)
        for (int64_t i = nv * V + t0; i < n; i += stride) lincomb_chunk<1>(y, x0, x1, x2, nx, a0, a1, a2, a3, ncomp, i);
}

template <typename T> __global__ __launch_bounds__(256) void mult_kernel(const Comp3<T> y, const Comp3<T> x, int ncomp, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < ncomp; ++c) y.p[c][i] *= x.p[c][i];
}

// One Chebyshev sub-step (linalg/solvers/chebyshev.hpp:102-165) in one pass:
//   z = invd * (b - Ax)   (Ax absent: x = 0)
//   d = cz * z + cd * d   (cd absent: d is overwritten)
//   x = x + d             (x absent: x = 0, written without being read)
// with the rounding of the separate lincomb / scale kernels.
template <int V, typename T>
__device__ __forceinline__ void cheb_chunk(const Comp3<T> &x, const Comp3<T> &d, const Comp3<T> &b,
                                           const Comp3<T> &ax, const Comp3<T> &invd, int has_ax, int read_d,
                                           int read_x, T az, T ad, int64_t i)
{
    Pack<T, V> vb[3], va[3], vi[3], vd[3], vx[3];
    auto load = [&](int c) {
        vb[c] = ldv<V>(b.p[c], i);
        if (has_ax) va[c] = ldv<V>(ax.p[c], i);
        vi[c] = ldv<V>(invd.p[c], i);
        if (read_d) vd[c] = ldv<V>(d.p[c], i);
        if (read_x) vx[c] = ldv<V>(x.p[c], i);
    };
    auto update = [&](int c) {
#pragma clang fp contract(off)
        Pack<T, V> dn, xn;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            T r = vb[c].e[k];
            if (has_ax) r = r - va[c].e[k];
            const T z = r * vi[c].e[k];
            T t = az * z;
            if (read_d) t = fma(ad, vd[c].e[k], t);
            dn.e[k] = t;
            xn.e[k] = read_x ? vx[c].e[k] + t : t;
        }
        stv<V>(d.p[c], i, dn);
        stv<V>(x.p[c], i, xn);
    };
    if constexpr (kLoadFirst<T>) {
#pragma unroll
        for (int c = 0; c < 3; ++c) load(c);
#pragma unroll
        for (int c = 0; c < 3; ++c) update(c);
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            load(c);
            update(c);
        }
    }
}

template <typename T, int V>
__global__ __launch_bounds__(256) void cheb_update_kernel(const Comp3<T> x, const Comp3<T> d, const Comp3<T> b,
                                                          const Comp3<T> ax, const Comp3<T> invd, int has_ax,
                                                          int read_d, int read_x, double cz, double cd, int64_t n)
{
    const T az = (T)cz, ad = (T)cd;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nv = n / V;
    for (int64_t i = t0; i < nv; i += stride) cheb_chunk<V>(x, d, b, ax, invd, has_ax, read_d, read_x, az, ad, i);
    if constexpr (V > 1)
        for (int64_t i = nv * V + t0; i < n; i += stride)
            cheb_chunk<1>(x, d, b, ax, invd, has_ax, read_d, read_x, az, ad, i);
}

template <typename T> __global__ __launch_bounds__(256) void invert_kernel(const Comp3<T> y, int ncomp, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < ncomp; ++c) y.p[c][i] = T(1) / y.p[c][i];
}

__device__ __forceinline__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wmax(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

enum RedOp { RED_DOT = 0, RED_MAXABS = 1, RED_NONFINITE = 2 };

// op RED_DOT: sum over OWNED copies of sum_c x_c y_c (y may alias x)
// op RED_MAXABS: max |x| over all copies; RED_NONFINITE: count of non-finite entries
// Two launches: per-block partials, then one block combines them in a fixed
// order (deterministic for a given grid; no inter-block fences or atomics).
template <int OP> __device__ __forceinline__ double red_combine(double a, double b)
{
    return OP == RED_MAXABS ? fmax(a, b) : a + b;
}

template <int OP> __device__ __forceinline__ double block_reduce(double acc, double *red)
{
    acc = OP == RED_MAXABS ? wmax(acc) : wsum(acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < 4; ++w) s = red_combine<OP>(s, red[w]);
    return s;
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void reduce_partial_kernel(const Comp3<T> x, const Comp3<T> y, int ncomp,
                                                             const uint8_t *omask, int64_t n, double *partials)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (OP == RED_DOT) {
            const bool own = omask[i] == OWNED;
            double v = 0.0;
            for (int c = 0; c < ncomp; ++c) v += (double)x.p[c][i] * (double)y.p[c][i];
            acc += own ? v : 0.0;
        } else if (OP == RED_MAXABS) {
            for (int c = 0; c < ncomp; ++c) acc = fmax(acc, fabs((double)x.p[c][i]));
        } else {
            for (int c = 0; c < ncomp; ++c) acc += isfinite((double)x.p[c][i]) ? 0.0 : 1.0;
        }
    }
    const double s = block_reduce<OP>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <int OP> __global__ __launch_bounds__(256) void reduce_final_kernel(const double *partials, int nb, double *out)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc = red_combine<OP>(acc, partials[b]);
    const double s = block_reduce<OP>(acc, red);
    if (threadIdx.x == 0) *out = s;
}

// ----------------------------------------------------------------------------
// PCG with the scalars on the device (linalg/solvers/pcg.hpp:81-170, identity
// preconditioner, so z = r).  State, in doubles: rr, initial norm, alpha,
// beta, done, iterations, final relative residual.  A finished solve turns
// every later step into a no-op, so the host issues iterations without
// waiting for the dots.  Dots use the reduction grid and per-thread order of
// reduce_partial_kernel, so every scalar is bitwise that of the host-synced
// formulation.
// ----------------------------------------------------------------------------
enum { PCG_RR = 0, PCG_INIT, PCG_ALPHA, PCG_BETA, PCG_DONE, PCG_ITERS, PCG_REL, PCG_NSTATE };

__device__ __forceinline__ double block_sum_partials(const double *partials, int nb, double *red)
{
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[b];
    return block_reduce<RED_DOT>(acc, red);
}

__global__ __launch_bounds__(256) void pcg_begin_kernel(const double *partials, int nb, double *st, double atol)
{
    __shared__ double red[4];
    const double rr = block_sum_partials(partials, nb, red);
    if (threadIdx.x == 0) {
        const double init = sqrt(rr);
        st[PCG_RR] = rr;
        st[PCG_INIT] = init;
        st[PCG_ALPHA] = 0.0;
        st[PCG_BETA] = 0.0;
        st[PCG_DONE] = init < atol ? 1.0 : 0.0;
        st[PCG_ITERS] = 0.0;
        st[PCG_REL] = 1.0;
    }
}

__global__ __launch_bounds__(256) void pcg_alpha_kernel(const double *partials, int nb, double *st, int it)
{
    __shared__ double red[4];
    const double pap = block_sum_partials(partials, nb, red);
    if (threadIdx.x == 0 && st[PCG_DONE] == 0.0) {
        if (fabs(pap) < DBL_MIN) {
            st[PCG_DONE] = 1.0;
            st[PCG_ITERS] = it;
        } else {
            st[PCG_ALPHA] = st[PCG_RR] / pap;
        }
    }
}

// x += alpha p; r -= alpha ap; per-block partials of r.r (OWNED copies)
template <typename T>
__global__ __launch_bounds__(256) void pcg_update_kernel(const Comp3<T> x, const Comp3<T> r, const Comp3<T> p,
                                                         const Comp3<T> ap, int ncomp, const uint8_t *omask, int64_t n,
                                                         const double *st, double *partials)
{
    __shared__ double red[4];
    if (st[PCG_DONE] != 0.0) return;
    const T alpha = (T)st[PCG_ALPHA], malpha = (T)(-st[PCG_ALPHA]);
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool own = omask[i] == OWNED;
        double v = 0.0;
        for (int c = 0; c < ncomp; ++c) {
            x.p[c][i] = fma(alpha, p.p[c][i], x.p[c][i]);
            const T rn = fma(malpha, ap.p[c][i], r.p[c][i]);
            r.p[c][i] = rn;
            v += (double)rn * (double)rn;
        }
        acc += own ? v : 0.0;
    }
    const double sblk = block_reduce<RED_DOT>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = sblk;
}

__global__ __launch_bounds__(256) void pcg_beta_kernel(const double *partials, int nb, double *st, double rtol,
                                                       double atol, int it)
{
    __shared__ double red[4];
    if (st[PCG_DONE] != 0.0) return;  // uniform over the block
    const double rr_new = block_sum_partials(partials, nb, red);
    if (threadIdx.x == 0) {
        const double res = sqrt(rr_new), rel = res / st[PCG_INIT];
        st[PCG_REL] = rel;
        const double rr_old = st[PCG_RR];
        if (rel <= rtol || res < atol || fabs(rr_old) < DBL_MIN) {
            st[PCG_DONE] = 1.0;
            st[PCG_ITERS] = it;
        } else {
            st[PCG_BETA] = rr_new / rr_old;
        }
        st[PCG_RR] = rr_new;
    }
}

// p = r + beta p
template <typename T>
__global__ __launch_bounds__(256) void pcg_direction_kernel(const Comp3<T> p, const Comp3<T> r, int ncomp, int64_t n,
                                                            const double *st)
{
    if (st[PCG_DONE] != 0.0) return;
    const T beta = (T)st[PCG_BETA];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < ncomp; ++c) p.p[c][i] = fma(beta, p.p[c][i], r.p[c][i]);
}

// ----------------------------------------------------------------------------
// Modified Gram-Schmidt step of FGMRES (linalg/solvers/fgmres.hpp:211-224)
// on one block of a Stokes vector, in one pass:
//   w -= h_prev v_prev   (h_prev read from the device; skipped when absent)
//   partial of (w . v_next) over OWNED copies   (v_next absent: w . w)
// The dot uses the reduction grid and per-thread order of
// reduce_partial_kernel; mgs_final_kernel adds the velocity and pressure dots
// as the host formulation did.
// ----------------------------------------------------------------------------
template <int V, typename T>
__device__ __forceinline__ double mgs_chunk(const Comp3<T> &w, const Comp3<T> &vprev, const Comp3<T> &vnext,
                                            int ncomp, const uint8_t *omask, bool upd, T mh, int self, int64_t i)
{
    Pack<T, V> vw[3], vp[3], vn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= ncomp) break;
        vw[c] = ldv<V>(w.p[c], i);
        if (upd) vp[c] = ldv<V>(vprev.p[c], i);
        if (!self) vn[c] = ldv<V>(vnext.p[c], i);
    }
    double v[V];
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c >= ncomp) break;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            T wc = vw[c].e[k];
            if (upd) wc = fma(mh, vp[c].e[k], wc);
            vw[c].e[k] = wc;
            const T vc = self ? wc : vn[c].e[k];
            v[k] += (double)wc * (double)vc;
        }
        if (upd) stv<V>(w.p[c], i, vw[c]);
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < V; ++k) acc += omask[i * V + k] == OWNED ? v[k] : 0.0;
    return acc;
}

template <typename T, int V>
__global__ __launch_bounds__(256) void mgs_block_kernel(const Comp3<T> w, const Comp3<T> vprev, const Comp3<T> vnext,
                                                        int ncomp, const uint8_t *omask, int64_t n,
                                                        const double *hprev, int self, double *partials)
{
    __shared__ double red[4];
    const bool upd = hprev != nullptr;
    const T mh = upd ? (T)(-hprev[0]) : T(0);
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nv = n / V;
    double acc = 0.0;
    for (int64_t i = t0; i < nv; i += stride) acc += mgs_chunk<V>(w, vprev, vnext, ncomp, omask, upd, mh, self, i);
    if constexpr (V > 1)
        for (int64_t i = nv * V + t0; i < n; i += stride)
            acc += mgs_chunk<1>(w, vprev, vnext, ncomp, omask, upd, mh, self, i);
    const double sblk = block_reduce<RED_DOT>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = sblk;
}

__global__ __launch_bounds__(256) void mgs_final_kernel(const double *pu, int gu, const double *pp, int gp, double *out)
{
    __shared__ double red[4];
    const double su = block_sum_partials(pu, gu, red);
    __syncthreads();
    const double sp = block_sum_partials(pp, gp, red);
    if (threadIdx.x == 0) *out = su + sp;
}

constexpr int kRedBlocks = 1024;

unsigned gs(int64_t n)
{
    int64_t g = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

}  // namespace

void launch_lincomb(tng_vector &y, int nx, const double *c, const tng_vector *const *x, double c0, hipStream_t s)
{
    for (int i = 0; i < nx; ++i)
        TNG_REQUIRE(x[i]->mesh == y.mesh && x[i]->ncomp == y.ncomp && x[i]->dtype == y.dtype,
                    "lincomb: vectors must share mesh, component count and dtype");
    TNG_REQUIRE(nx >= 0 && nx <= 3, "lincomb: at most 3 input vectors");
    const int64_t n = y.mesh->nalloc;
    const double cc[3] = {nx > 0 ? c[0] : 0, nx > 1 ? c[1] : 0, nx > 2 ? c[2] : 0};
    const tng_vector &x0 = nx > 0 ? *x[0] : y, &x1 = nx > 1 ? *x[1] : y, &x2 = nx > 2 ? *x[2] : y;
    auto run = [&](auto tag) {
        using T = decltype(tag);
        const Comp3<T> py = comps<T>(y), p0 = comps<T>(x0), p1 = comps<T>(x1), p2 = comps<T>(x2);
        const int nc = y.ncomp;
        if (aligned16(py, nc) && aligned16(p0, nc) && aligned16(p1, nc) && aligned16(p2, nc))
            hipLaunchKernelGGL((lincomb_kernel<T, kVec<T>>), dim3(gs(n / kVec<T>)), dim3(256), 0, s, py, p0, p1, p2, nx,
                               c0, cc[0], cc[1], cc[2], nc, n);
        else
            hipLaunchKernelGGL((lincomb_kernel<T, 1>), dim3(gs(n)), dim3(256), 0, s, py, p0, p1, p2, nx, c0, cc[0],
                               cc[1], cc[2], nc, n);
    };
    if (y.dtype == TNG_F64)
        run(double());
    else
        run(float());
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_mult(tng_vector &y, const tng_vector &x, hipStream_t s)
{
    TNG_REQUIRE(x.mesh == y.mesh && x.dtype == y.dtype && (x.ncomp == y.ncomp), "scale: mismatched vectors");
    const int64_t n = y.mesh->nalloc;
    if (y.dtype == TNG_F64)
        hipLaunchKernelGGL(mult_kernel<double>, dim3(gs(n)), dim3(256), 0, s, comps<double>(y), comps<double>(x),
                           y.ncomp, n);
    else
        hipLaunchKernelGGL(mult_kernel<float>, dim3(gs(n)), dim3(256), 0, s, comps<float>(y), comps<float>(x), y.ncomp,
                           n);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_cheb_update(tng_vector &x, tng_vector &d, const tng_vector &b, const tng_vector *ax,
                        const tng_vector &invd, double cz, double cd, bool read_d, bool x_zero, hipStream_t s)
{
    const tng_vector *const others[3] = {&d, &b, &invd};
    for (const tng_vector *v : others)
        TNG_REQUIRE(v->mesh == x.mesh && v->dtype == x.dtype && v->ncomp == 3 && x.ncomp == 3,
                    "chebyshev: vectors must share mesh and dtype and have 3 components");
    if (ax) TNG_REQUIRE(ax->mesh == x.mesh && ax->dtype == x.dtype && ax->ncomp == 3, "chebyshev: mismatched A x");
    const int64_t n = x.mesh->nalloc;
    const tng_vector &axv = ax ? *ax : b;
    auto run = [&](auto tag) {
        using T = decltype(tag);
        const Comp3<T> px = comps<T>(x), pd = comps<T>(d), pb = comps<T>(b), pa = comps<T>(axv), pi = comps<T>(invd);
        if (aligned16(px, 3) && aligned16(pd, 3) && aligned16(pb, 3) && aligned16(pa, 3) && aligned16(pi, 3))
            hipLaunchKernelGGL((cheb_update_kernel<T, kVec<T>>), dim3(gs(n / kVec<T>)), dim3(256), 0, s, px, pd, pb,
                               pa, pi, ax ? 1 : 0, read_d ? 1 : 0, x_zero ? 0 : 1, cz, cd, n);
        else
            hipLaunchKernelGGL((cheb_update_kernel<T, 1>), dim3(gs(n)), dim3(256), 0, s, px, pd, pb, pa, pi, ax ? 1 : 0,
                               read_d ? 1 : 0, x_zero ? 0 : 1, cz, cd, n);
    };
    if (x.dtype == TNG_F64)
        run(double());
    else
        run(float());
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_invert(tng_vector &y, hipStream_t s)
{
    const int64_t n = y.mesh->nalloc;
    if (y.dtype == TNG_F64)
        hipLaunchKernelGGL(invert_kernel<double>, dim3(gs(n)), dim3(256), 0, s, comps<double>(y), y.ncomp, n);
    else
        hipLaunchKernelGGL(invert_kernel<float>, dim3(gs(n)), dim3(256), 0, s, comps<float>(y), y.ncomp, n);
    TNG_HIP_CHECK(hipGetLastError());
}

// Reduction into ctx.red_result[slot] on the device.
void launch_reduce(int op, const tng_vector &x, const tng_vector *y, int slot, hipStream_t s)
{
    tng_context &ctx = *x.mesh->ctx;
    if (y) TNG_REQUIRE(y->mesh == x.mesh && y->ncomp == x.ncomp && y->dtype == x.dtype, "dot: mismatched vectors");
    const int64_t n = x.mesh->nalloc;
    const unsigned g = std::min<unsigned>(kRedBlocks, gs(n));
    double *out = ctx.red_result.as<double>() + slot;
    const uint8_t *om = x.mesh->d_omask.as<uint8_t>();
    double *part = ctx.red_partials.as<double>();
    const tng_vector &yy = y ? *y : x;
#define TNG_RED(T, OPV)                                                                                            \
    do {                                                                                                           \
        hipLaunchKernelGGL((reduce_partial_kernel<T, OPV>), dim3(g), dim3(256), 0, s, comps<T>(x), comps<T>(yy),   \
                           x.ncomp, om, n, part);                                                                  \
        hipLaunchKernelGGL((reduce_final_kernel<OPV>), dim3(1), dim3(256), 0, s, part, (int)g, out);               \
    } while (0)
    if (x.dtype == TNG_F64) {
        if (op == RED_DOT) TNG_RED(double, RED_DOT);
        else if (op == RED_MAXABS) TNG_RED(double, RED_MAXABS);
        else TNG_RED(double, RED_NONFINITE);
    } else {
        if (op == RED_DOT) TNG_RED(float, RED_DOT);
        else if (op == RED_MAXABS) TNG_RED(float, RED_MAXABS);
        else TNG_RED(float, RED_NONFINITE);
    }
#undef TNG_RED
    TNG_HIP_CHECK(hipGetLastError());
}

double reduce_to_host(int op, const tng_vector &x, const tng_vector *y, hipStream_t s)
{
    tng_context &ctx = *x.mesh->ctx;
    launch_reduce(op, x, y, 0, s);
    double v = 0;
    TNG_HIP_CHECK(hipMemcpyAsync(&v, ctx.red_result.p, sizeof(double), hipMemcpyDeviceToHost, s));
    TNG_HIP_CHECK(hipStreamSynchronize(s));
    return v;
}

// PCG on the device: the state lives in ctx.red_result[kPcgSlot ...]
constexpr int kPcgSlot = 64;

static unsigned red_grid(int64_t n) { return std::min<unsigned>(kRedBlocks, gs(n)); }

void launch_pcg_begin(const tng_vector &r, double atol, hipStream_t s)
{
    tng_context &ctx = *r.mesh->ctx;
    const int64_t n = r.mesh->nalloc;
    const unsigned g = red_grid(n);
    double *part = ctx.red_partials.as<double>(), *st = ctx.red_result.as<double>() + kPcgSlot;
    const uint8_t *om = r.mesh->d_omask.as<uint8_t>();
    if (r.dtype == TNG_F64)
        hipLaunchKernelGGL((reduce_partial_kernel<double, RED_DOT>), dim3(g), dim3(256), 0, s, comps<double>(r),
                           comps<double>(r), r.ncomp, om, n, part);
    else
        hipLaunchKernelGGL((reduce_partial_kernel<float, RED_DOT>), dim3(g), dim3(256), 0, s, comps<float>(r),
                           comps<float>(r), r.ncomp, om, n, part);
    hipLaunchKernelGGL(pcg_begin_kernel, dim3(1), dim3(256), 0, s, part, (int)g, st, atol);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_pcg_alpha(const tng_vector &ap, const tng_vector &p, int it, hipStream_t s)
{
    tng_context &ctx = *p.mesh->ctx;
    const int64_t n = p.mesh->nalloc;
    const unsigned g = red_grid(n);
    double *part = ctx.red_partials.as<double>(), *st = ctx.red_result.as<double>() + kPcgSlot;
    const uint8_t *om = p.mesh->d_omask.as<uint8_t>();
    if (p.dtype == TNG_F64)
        hipLaunchKernelGGL((reduce_partial_kernel<double, RED_DOT>), dim3(g), dim3(256), 0, s, comps<double>(ap),
                           comps<double>(p), p.ncomp, om, n, part);
    else
        hipLaunchKernelGGL((reduce_partial_kernel<float, RED_DOT>), dim3(g), dim3(256), 0, s, comps<float>(ap),
                           comps<float>(p), p.ncomp, om, n, part);
    hipLaunchKernelGGL(pcg_alpha_kernel, dim3(1), dim3(256), 0, s, part, (int)g, st, it);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_pcg_update(tng_vector &x, tng_vector &r, const tng_vector &p, const tng_vector &ap, double rtol,
                       double atol, int it, hipStream_t s)
{
    tng_context &ctx = *p.mesh->ctx;
    const int64_t n = p.mesh->nalloc;
    const unsigned g = red_grid(n);
    double *part = ctx.red_partials.as<double>(), *st = ctx.red_result.as<double>() + kPcgSlot;
    const uint8_t *om = p.mesh->d_omask.as<uint8_t>();
    if (p.dtype == TNG_F64)
        hipLaunchKernelGGL(pcg_update_kernel<double>, dim3(g), dim3(256), 0, s, comps<double>(x), comps<double>(r),
                           comps<double>(p), comps<double>(ap), p.ncomp, om, n, (const double *)st, part);
    else
        hipLaunchKernelGGL(pcg_update_kernel<float>, dim3(g), dim3(256), 0, s, comps<float>(x), comps<float>(r),
                           comps<float>(p), comps<float>(ap), p.ncomp, om, n, (const double *)st, part);
    hipLaunchKernelGGL(pcg_beta_kernel, dim3(1), dim3(256), 0, s, part, (int)g, st, rtol, atol, it);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_pcg_direction(tng_vector &p, const tng_vector &r, hipStream_t s)
{
    tng_context &ctx = *p.mesh->ctx;
    const int64_t n = p.mesh->nalloc;
    const double *st = ctx.red_result.as<double>() + kPcgSlot;
    if (p.dtype == TNG_F64)
        hipLaunchKernelGGL(pcg_direction_kernel<double>, dim3(gs(n)), dim3(256), 0, s, comps<double>(p),
                           comps<double>(r), p.ncomp, n, st);
    else
        hipLaunchKernelGGL(pcg_direction_kernel<float>, dim3(gs(n)), dim3(256), 0, s, comps<float>(p),
                           comps<float>(r), p.ncomp, n, st);
    TNG_HIP_CHECK(hipGetLastError());
}

// MGS step on a Stokes vector (u, p): h_out = (w -= h_prev v_prev) . v_next
void launch_mgs_step(tng_vector &wu, tng_vector &wp, const tng_vector *vprev_u, const tng_vector *vprev_p,
                     const double *h_prev_dev, const tng_vector *vnext_u, const tng_vector *vnext_p, double *h_out_dev,
                     hipStream_t s)
{
    tng_context &ctx = *wu.mesh->ctx;
    double *part = ctx.red_partials.as<double>();
    unsigned g[2] = {0, 0};
    tng_vector *w[2] = {&wu, &wp};
    const tng_vector *vp[2] = {vprev_u, vprev_p}, *vn[2] = {vnext_u, vnext_p};
    for (int b = 0; b < 2; ++b) {
        const tng_vector &wb = *w[b];
        const int64_t n = wb.mesh->nalloc;
        const tng_vector &pv = vp[b] ? *vp[b] : wb, &nv = vn[b] ? *vn[b] : wb;
        TNG_REQUIRE(pv.mesh == wb.mesh && nv.mesh == wb.mesh && pv.dtype == wb.dtype && nv.dtype == wb.dtype,
                    "mgs: mismatched vectors");
        const uint8_t *om = wb.mesh->d_omask.as<uint8_t>();
        double *pb = part + b * kRedBlocks;
        auto run = [&](auto tag) {
            using T = decltype(tag);
            const Comp3<T> pw = comps<T>(wb), pp = comps<T>(pv), pn = comps<T>(nv);
            const int nc = wb.ncomp;
            if (aligned16(pw, nc) && aligned16(pp, nc) && aligned16(pn, nc)) {
                g[b] = red_grid(n / kVec<T>);
                hipLaunchKernelGGL((mgs_block_kernel<T, kVec<T>>), dim3(g[b]), dim3(256), 0, s, pw, pp, pn, nc, om, n,
                                   h_prev_dev, vn[b] ? 0 : 1, pb);
            } else {
                g[b] = red_grid(n);
                hipLaunchKernelGGL((mgs_block_kernel<T, 1>), dim3(g[b]), dim3(256), 0, s, pw, pp, pn, nc, om, n,
                                   h_prev_dev, vn[b] ? 0 : 1, pb);
            }
        };
        if (wb.dtype == TNG_F64)
            run(double());
        else
            run(float());
    }
    hipLaunchKernelGGL(mgs_final_kernel, dim3(1), dim3(256), 0, s, (const double *)part, (int)g[0],
                       (const double *)(part + kRedBlocks), (int)g[1], h_out_dev);
    TNG_HIP_CHECK(hipGetLastError());
}

double *device_slots(tng_context &ctx, int first) { return ctx.red_result.as<double>() + first; }

void pcg_read_state(const tng_vector &v, int *done, int *iters, double *final_rel, hipStream_t s)
{
    tng_context &ctx = *v.mesh->ctx;
    double st[PCG_NSTATE];
    TNG_HIP_CHECK(hipMemcpyAsync(st, ctx.red_result.as<double>() + kPcgSlot, sizeof(st), hipMemcpyDeviceToHost, s));
    TNG_HIP_CHECK(hipStreamSynchronize(s));
    if (done) *done = st[PCG_DONE] != 0.0;
    if (iters) *iters = (int)st[PCG_ITERS];
    if (final_rel) *final_rel = st[PCG_REL];
}

}  // namespace tng

// ----------------------------------------------------------------------------
// Seeded uniform [0, 1) fill (the power iteration's random start; the
// reference draws from a Kokkos XorShift64 pool, grid_operations.hpp:748-773).
// Every entry is SplitMix64 of (seed, component, linear index), so the fill is
// reproducible and independent of the launch shape.
// ----------------------------------------------------------------------------
namespace tng {

namespace {

__device__ __forceinline__ double splitmix_unit(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

template <typename T> __global__ __launch_bounds__(256) void randomize_kernel(T *p, int64_t n, uint64_t base)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (T)splitmix_unit(base + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull);
}

}  // namespace

void launch_randomize(tng_vector &v, uint64_t seed, hipStream_t s)
{
    const int64_t n = v.mesh->nalloc;
    for (int c = 0; c < v.ncomp; ++c) {
        const uint64_t base = seed * 0xD1B54A32D192ED03ull + (uint64_t)c * 0x8CB92BA72F3D8DD7ull;
        if (v.dtype == TNG_F64)
            hipLaunchKernelGGL(randomize_kernel<double>, dim3(grid_for(n)), dim3(256), 0, s, (double *)v.comp(c), n, base);
        else
            hipLaunchKernelGGL(randomize_kernel<float>, dim3(grid_for(n)), dim3(256), 0, s, (float *)v.comp(c), n, base);
        TNG_HIP_CHECK(hipGetLastError());
    }
}

}  // namespace tng
