// Gram-Schmidt of the FGMRES Arnoldi step with one global reduction per step
// (classical Gram-Schmidt with the re-orthogonalisation of each basis vector
// delayed to the next step, "DCGS2").  Two passes over the basis per step:
//
//   gs_dots:   one multi-dot over the Stokes vectors: Q^T u, u.u, Q^T w, u.w,
//              w.w (Q = the final basis q_0..q_{j-1}, u = the tentative vector
//              u_j, w = K z_j), OWNED-masked, as per-block partials;
//   gs_update: q_j = (u - Q s) / alpha and u_{j+1} = (w - Q c' - r u) / nu in
//              one pass that reads Q, u and w once.
//
// The reference orthogonalises by modified Gram-Schmidt
// (linalg/solvers/fgmres.hpp:211-224: j + 2 sequential passes per step, each
// with a host round-trip); the two-pass form moves about half the bytes.
#include <algorithm>

#include "tng_gs.hpp"
#include "tng_internal.hpp"

namespace tng {

namespace {

__device__ __forceinline__ double gs_wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Rows: the velocity part (3 components per node copy) then the pressure part
// (1 component).  Slot i of the basis table holds 4 component pointers.
template <int JB>
__global__ __launch_bounds__(256) void gs_dots_kernel(const GsArgs a, double *partials)
{
    constexpr int ND = 2 * JB + 3;
    __shared__ double red[4][ND];
    double acc[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) acc[k] = 0.0;
    const int j = a.j;
    const int64_t nrows = a.nv + a.np;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x) {
        const bool vel = r < a.nv;
        const int64_t n = vel ? r : r - a.nv;
        const bool own = (vel ? a.mask_v[n] : a.mask_p[n]) == OWNED;
        const int c0 = vel ? 0 : 3, c1 = vel ? 3 : 4;
        double v[ND];
#pragma unroll
        for (int k = 0; k < ND; ++k) v[k] = 0.0;
        for (int c = c0; c < c1; ++c) {
            const double wc = a.w[c][n];
            const double uc = a.V[4 * j + c][n];
#pragma unroll
            for (int i = 0; i < JB; ++i) {
                if (i < j) {
                    const double q = a.V[4 * i + c][n];
                    v[i] = fma(q, uc, v[i]);
                    v[JB + i] = fma(q, wc, v[JB + i]);
                }
            }
            v[2 * JB] = fma(uc, uc, v[2 * JB]);
            v[2 * JB + 1] = fma(uc, wc, v[2 * JB + 1]);
            v[2 * JB + 2] = fma(wc, wc, v[2 * JB + 2]);
        }
        if (own) {
#pragma unroll
            for (int k = 0; k < ND; ++k) acc[k] += v[k];
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        const double s = gs_wsum(acc[k]);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ND; k += blockDim.x)
        partials[(int64_t)blockIdx.x * ND + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// Sums the per-block partials of every dot in a fixed order.
__global__ __launch_bounds__(256) void gs_final_kernel(const double *partials, int nb, int nd, double *out)
{
    for (int k = threadIdx.x; k < nd; k += blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partials[(int64_t)b * nd + k];
        out[k] = s;
    }
}

template <int JB> __global__ __launch_bounds__(256) void gs_update_kernel(const GsArgs a, const GsCoef cf)
{
    const int j = a.j;
    const int64_t nrows = a.nv + a.np;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x) {
        const bool vel = r < a.nv;
        const int64_t n = vel ? r : r - a.nv;
        const int c0 = vel ? 0 : 3, c1 = vel ? 3 : 4;
        for (int c = c0; c < c1; ++c) {
            double ts = 0.0, tc = 0.0;
#pragma unroll
            for (int i = 0; i < JB; ++i) {
                if (i < j) {
                    const double q = a.V[4 * i + c][n];
                    ts = fma(cf.s[i], q, ts);
                    tc = fma(cf.c[i], q, tc);
                }
            }
            double *uslot = a.V[4 * j + c];
            const double u = uslot[n], wc = a.w[c][n];
            uslot[n] = (u - ts) * cf.inv_alpha;
            a.V[4 * (j + 1) + c][n] = ((wc - tc) - cf.r * u) * cf.inv_nu;
        }
    }
}

// y += sum_{i < k} coef_i X_i over the Stokes rows (X_i: table slots 0..k-1)
__global__ __launch_bounds__(256) void gs_axpy_kernel(double *const *X, int k, const GsAxpy cf, double *const *y,
                                                      int64_t nv, int64_t np)
{
    const int64_t nrows = nv + np;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * blockDim.x) {
        const bool vel = r < nv;
        const int64_t n = vel ? r : r - nv;
        const int c0 = vel ? 0 : 3, c1 = vel ? 3 : 4;
        for (int c = c0; c < c1; ++c) {
            double t = y[c][n];
            for (int i = 0; i < k; ++i) t = fma(cf.c[i], X[4 * i + c][n], t);
            y[c][n] = t;
        }
    }
}

// Rigid rotation e_axis x X at every node copy (X = radius * unit lateral position).
__global__ __launch_bounds__(256) void rotation_field_kernel(double *v0, double *v1, double *v2, const double *coords,
                                                             const double *radii, int axis, int nx, int nr, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t lcol = i / nr;
        const int l = (int)(i - lcol * nr);
        const int64_t sd = lcol / ((int64_t)nx * nx);
        const double r = radii[sd * nr + l];
        const double x = coords[3 * lcol] * r, y = coords[3 * lcol + 1] * r, z = coords[3 * lcol + 2] * r;
        double m0, m1, m2;
        if (axis == 0) { m0 = 0.0; m1 = -z; m2 = y; }
        else if (axis == 1) { m0 = z; m1 = 0.0; m2 = -x; }
        else { m0 = -y; m1 = x; m2 = 0.0; }
        v0[i] = m0; v1[i] = m1; v2[i] = m2;
    }
}

constexpr int kGsBlocks = 512;

template <int JB> void gs_dots_jb(const GsArgs &a, double *partials, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(gs_dots_kernel<JB>, dim3(kGsBlocks), dim3(256), 0, s, a, partials);
    hipLaunchKernelGGL(gs_final_kernel, dim3(1), dim3(256), 0, s, (const double *)partials, kGsBlocks, 2 * JB + 3,
                       out);
}

}  // namespace

int gs_bucket(int j) { return j <= 8 ? 8 : (j <= 16 ? 16 : 32); }

size_t gs_partials_doubles() { return (size_t)kGsBlocks * (2 * 32 + 3); }

void gs_dots(const GsArgs &a, double *partials, double *out, hipStream_t s)
{
    TNG_REQUIRE(a.j >= 0 && a.j <= 32, "Gram-Schmidt: at most 32 basis vectors");
    switch (gs_bucket(a.j)) {
    case 8: gs_dots_jb<8>(a, partials, out, s); break;
    case 16: gs_dots_jb<16>(a, partials, out, s); break;
    default: gs_dots_jb<32>(a, partials, out, s); break;
    }
    TNG_HIP_CHECK(hipGetLastError());
}

void gs_update(const GsArgs &a, const GsCoef &cf, hipStream_t s)
{
    TNG_REQUIRE(a.j >= 0 && a.j <= 32, "Gram-Schmidt: at most 32 basis vectors");
    const int64_t rows = a.nv + a.np;
    const unsigned g = (unsigned)std::min<int64_t>((rows + 255) / 256, 4096);
    switch (gs_bucket(a.j)) {
    case 8: hipLaunchKernelGGL(gs_update_kernel<8>, dim3(g), dim3(256), 0, s, a, cf); break;
    case 16: hipLaunchKernelGGL(gs_update_kernel<16>, dim3(g), dim3(256), 0, s, a, cf); break;
    default: hipLaunchKernelGGL(gs_update_kernel<32>, dim3(g), dim3(256), 0, s, a, cf); break;
    }
    TNG_HIP_CHECK(hipGetLastError());
}

void gs_axpy(double *const *X, int k, const GsAxpy &cf, double *const *y, int64_t nv, int64_t np, hipStream_t s)
{
    TNG_REQUIRE(k >= 0 && k <= 8, "multi-axpy: at most 8 vectors");
    const int64_t rows = nv + np;
    const unsigned g = (unsigned)std::min<int64_t>((rows + 255) / 256, 4096);
    hipLaunchKernelGGL(gs_axpy_kernel, dim3(g), dim3(256), 0, s, X, k, cf, y, nv, np);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_rotation_field(tng_vector &v, int axis, hipStream_t s)
{
    const tng_domain &m = *v.mesh;
    TNG_REQUIRE(v.dtype == TNG_F64 && v.ncomp == 3 && axis >= 0 && axis < 3, "rotation field: FP64 velocity");
    const unsigned g = (unsigned)std::min<int64_t>((m.nalloc + 255) / 256, 16384);
    hipLaunchKernelGGL(rotation_field_kernel, dim3(g), dim3(256), 0, s, (double *)v.comp(0), (double *)v.comp(1),
                       (double *)v.comp(2), m.d_coords.as<double>(), m.d_radii.as<double>(), axis, m.nx, m.nr,
                       m.nalloc);
    TNG_HIP_CHECK(hipGetLastError());
}

}  // namespace tng
