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
// (1 component), in two loops so that the component of every access is known
// at compile time and the basis pointers are wave-uniform (scalar loads).
// Slot i of the basis table holds 4 component pointers.  Workgroup row
// blockIdx.y takes the basis vectors 8 y .. 8 y + 7 (19 sums per thread); the
// first row also forms u.u, u.w, w.w.  A thread takes two consecutive rows
// per iteration with 16-byte loads (component arrays are 256-byte aligned);
// rows that are not OWNED enter with u = w = 0, so all loads of an iteration
// are independent of the mask.
constexpr int kGsChunk = 8, kGsND = 2 * kGsChunk + 3, kGsMaxChunks = 4;

__device__ __forceinline__ double2 ld2(const double *p, int64_t n) { return *reinterpret_cast<const double2 *>(p + n); }
__device__ __forceinline__ void st2(double *p, int64_t n, double2 v) { *reinterpret_cast<double2 *>(p + n) = v; }

template <int C0, int NC>
__device__ __forceinline__ void gs_dots_rows(const GsArgs &a, const uint8_t *mask, int64_t nrows, int i0, int nq,
                                             bool first, double (&acc)[kGsND])
{
    const int64_t npairs = nrows >> 1, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t pr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pr < npairs; pr += stride) {
        const int64_t n = 2 * pr;
        const unsigned m2 = *reinterpret_cast<const uint16_t *>(mask + n);
        const bool oa = (m2 & 0xffu) == OWNED, ob = (m2 >> 8) == OWNED;
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            const double2 w2 = ld2(a.w[c], n), u2 = ld2(a.V[4 * a.j + c], n);
            const double ua = oa ? u2.x : 0.0, ub = ob ? u2.y : 0.0;
            const double wa = oa ? w2.x : 0.0, wb = ob ? w2.y : 0.0;
#pragma unroll
            for (int i = 0; i < kGsChunk; ++i) {
                if (i < nq) {
                    const double2 q = ld2(a.V[4 * (i0 + i) + c], n);
                    acc[i] = fma(q.y, ub, fma(q.x, ua, acc[i]));
                    acc[kGsChunk + i] = fma(q.y, wb, fma(q.x, wa, acc[kGsChunk + i]));
                }
            }
            if (first) {
                acc[2 * kGsChunk] = fma(ub, ub, fma(ua, ua, acc[2 * kGsChunk]));
                acc[2 * kGsChunk + 1] = fma(ub, wb, fma(ua, wa, acc[2 * kGsChunk + 1]));
                acc[2 * kGsChunk + 2] = fma(wb, wb, fma(wa, wa, acc[2 * kGsChunk + 2]));
            }
        }
    }
    // an odd last row: the first thread of the grid
    if ((nrows & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const int64_t n = nrows - 1;
        const bool o = mask[n] == OWNED;
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            const double u = o ? a.V[4 * a.j + c][n] : 0.0, w = o ? a.w[c][n] : 0.0;
#pragma unroll
            for (int i = 0; i < kGsChunk; ++i) {
                if (i < nq) {
                    const double q = a.V[4 * (i0 + i) + c][n];
                    acc[i] = fma(q, u, acc[i]);
                    acc[kGsChunk + i] = fma(q, w, acc[kGsChunk + i]);
                }
            }
            if (first) {
                acc[2 * kGsChunk] = fma(u, u, acc[2 * kGsChunk]);
                acc[2 * kGsChunk + 1] = fma(u, w, acc[2 * kGsChunk + 1]);
                acc[2 * kGsChunk + 2] = fma(w, w, acc[2 * kGsChunk + 2]);
            }
        }
    }
}

__global__ __launch_bounds__(256) void gs_dots_kernel(const GsArgs a, double *partials)
{
    __shared__ double red[4][kGsND];
    const int i0 = blockIdx.y * kGsChunk;
    const int nq = a.j - i0;  // basis vectors of this row of workgroups: min(8, nq)
    const bool first = blockIdx.y == 0;
    double acc[kGsND];
#pragma unroll
    for (int k = 0; k < kGsND; ++k) acc[k] = 0.0;
    gs_dots_rows<0, 3>(a, a.mask_v, a.nv, i0, nq, first, acc);
    gs_dots_rows<3, 1>(a, a.mask_p, a.np, i0, nq, first, acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kGsND; ++k) {
        const double s = gs_wsum(acc[k]);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    double *o = partials + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * kGsND;
    for (int k = threadIdx.x; k < kGsND; k += blockDim.x) o[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// Sums the per-workgroup partials of every dot in a fixed order into
// out[0 .. 2B+2]: Q^T u (B slots), Q^T w (B slots), u.u, u.w, w.w.  One
// workgroup per output; each thread sums a fixed strided subset, then a fixed
// tree over the threads.
__global__ __launch_bounds__(256) void gs_final_kernel(const double *partials, int nbx, int nch, int B, double *out)
{
    __shared__ double red[4];
    const int o = blockIdx.x;
    int ch, k;
    if (o < B) {
        ch = o / kGsChunk;
        k = o % kGsChunk;
    } else if (o < 2 * B) {
        ch = (o - B) / kGsChunk;
        k = kGsChunk + (o - B) % kGsChunk;
    } else {
        ch = 0;
        k = 2 * kGsChunk + (o - 2 * B);
    }
    double s = 0.0;
    if (ch < nch)
        for (int b = threadIdx.x; b < nbx; b += blockDim.x) s += partials[((int64_t)ch * nbx + b) * kGsND + k];
    s = gs_wsum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[o] = (red[0] + red[1]) + (red[2] + red[3]);
}

// The basis vectors in runs of 8 (their pointers and coefficients are
// wave-uniform scalar loads per run); two consecutive rows per thread with
// 16-byte loads and stores, an odd last row by the first thread.
template <int C0, int NC>
__device__ __forceinline__ void gs_update_rows(const GsArgs &a, const GsCoef &cf, int64_t nrows)
{
    const int j = a.j;
    const int64_t npairs = nrows >> 1, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t pr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pr < npairs; pr += stride) {
        const int64_t n = 2 * pr;
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            double tsa = 0.0, tsb = 0.0, tca = 0.0, tcb = 0.0;
            for (int i0 = 0; i0 < j; i0 += 8) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int i = i0 + k;
                    if (i < j) {
                        const double2 q = ld2(a.V[4 * i + c], n);
                        tsa = fma(cf.s[i], q.x, tsa);
                        tsb = fma(cf.s[i], q.y, tsb);
                        tca = fma(cf.c[i], q.x, tca);
                        tcb = fma(cf.c[i], q.y, tcb);
                    }
                }
            }
            double *uslot = a.V[4 * j + c];
            const double2 u = ld2(uslot, n), w = ld2(a.w[c], n);
            st2(uslot, n, make_double2((u.x - tsa) * cf.inv_alpha, (u.y - tsb) * cf.inv_alpha));
            st2(a.V[4 * (j + 1) + c], n,
                make_double2(((w.x - tca) - cf.r * u.x) * cf.inv_nu, ((w.y - tcb) - cf.r * u.y) * cf.inv_nu));
        }
    }
    if ((nrows & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const int64_t n = nrows - 1;
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            double ts = 0.0, tc = 0.0;
            for (int i = 0; i < j; ++i) {
                const double q = a.V[4 * i + c][n];
                ts = fma(cf.s[i], q, ts);
                tc = fma(cf.c[i], q, tc);
            }
            double *uslot = a.V[4 * j + c];
            const double u = uslot[n], wc = a.w[c][n];
            uslot[n] = (u - ts) * cf.inv_alpha;
            a.V[4 * (j + 1) + c][n] = ((wc - tc) - cf.r * u) * cf.inv_nu;
        }
    }
}

__global__ __launch_bounds__(256) void gs_update_kernel(const GsArgs a, const GsCoef cf)
{
    gs_update_rows<0, 3>(a, cf, a.nv);
    gs_update_rows<3, 1>(a, cf, a.np);
}

// y += sum_{i < k} coef_i X_i over the Stokes rows (X_i: table slots 0..k-1):
// velocity rows then pressure rows (components known at compile time, the
// pointers wave-uniform), two consecutive rows per thread with 16-byte loads.
template <int C0, int NC>
__device__ __forceinline__ void gs_axpy_rows(double *const *X, int k, const GsAxpy &cf, double *const *y, int64_t nrows)
{
    const int64_t npairs = nrows >> 1, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t pr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pr < npairs; pr += stride) {
        const int64_t n = 2 * pr;
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            double2 t = ld2(y[c], n);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < k) {
                    const double2 q = ld2(X[4 * i + c], n);
                    t.x = fma(cf.c[i], q.x, t.x);
                    t.y = fma(cf.c[i], q.y, t.y);
                }
            }
            st2(y[c], n, t);
        }
    }
    if ((nrows & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const int64_t n = nrows - 1;
#pragma unroll
        for (int c = C0; c < C0 + NC; ++c) {
            double t = y[c][n];
            for (int i = 0; i < k; ++i) t = fma(cf.c[i], X[4 * i + c][n], t);
            y[c][n] = t;
        }
    }
}

__global__ __launch_bounds__(256) void gs_axpy_kernel(double *const *X, int k, const GsAxpy cf, double *const *y,
                                                      int64_t nv, int64_t np)
{
    gs_axpy_rows<0, 3>(X, k, cf, y, nv);
    gs_axpy_rows<3, 1>(X, k, cf, y, np);
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

#ifndef TNG_GS_BLOCKS
#define TNG_GS_BLOCKS 2048
#endif
constexpr int kGsBlocks = TNG_GS_BLOCKS;

}  // namespace

int gs_bucket(int j) { return j <= 8 ? 8 : (j <= 16 ? 16 : 32); }

size_t gs_partials_doubles() { return (size_t)kGsBlocks * kGsMaxChunks * kGsND; }

void gs_dots(const GsArgs &a, double *partials, double *out, hipStream_t s)
{
    TNG_REQUIRE(a.j >= 0 && a.j <= kGsChunk * kGsMaxChunks, "Gram-Schmidt: at most 32 basis vectors");
    const int nch = std::max(1, (a.j + kGsChunk - 1) / kGsChunk);
    // enough workgroups per chunk row for eight waves per SIMD: with fewer,
    // the pass is bound by the loads each wave has in flight, not by HBM
    const int64_t pairs = std::max(a.nv, a.np) / 2;
    const int nbx = (int)std::max<int64_t>(1, std::min<int64_t>(kGsBlocks, (pairs + 255) / 256));
    hipLaunchKernelGGL(gs_dots_kernel, dim3(nbx, nch), dim3(256), 0, s, a, partials);
    hipLaunchKernelGGL(gs_final_kernel, dim3(2 * gs_bucket(a.j) + 3), dim3(256), 0, s, (const double *)partials,
                       nbx, nch, gs_bucket(a.j), out);
    TNG_HIP_CHECK(hipGetLastError());
}

void gs_update(const GsArgs &a, const GsCoef &cf, hipStream_t s)
{
    TNG_REQUIRE(a.j >= 0 && a.j <= 32, "Gram-Schmidt: at most 32 basis vectors");
    const int64_t pairs = std::max(a.nv, a.np) / 2;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((pairs + 255) / 256, 4096));
    hipLaunchKernelGGL(gs_update_kernel, dim3(g), dim3(256), 0, s, a, cf);
    TNG_HIP_CHECK(hipGetLastError());
}

void gs_axpy(double *const *X, int k, const GsAxpy &cf, double *const *y, int64_t nv, int64_t np, hipStream_t s)
{
    TNG_REQUIRE(k >= 0 && k <= 8, "multi-axpy: at most 8 vectors");
    const int64_t pairs = std::max(nv, np) / 2;
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((pairs + 255) / 256, 4096));
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
