// Direct solve on the coarsest multigrid level (an option beside the
// reference's PCG, linalg/solvers/pcg.hpp): the level's operator, boundary
// rows and fs/fs penalty included, as a dense matrix over the unique node
// components, inverted once per viscosity update by Gauss-Jordan on the GPU
// (no pivoting: the operator is symmetric positive definite), and applied
// per V-cycle as one matrix-vector product.
//
//   unique numbering: uid[copy] = unique node of a node copy; rep[node] = one
//   copy of it.  Unknown 3 u + c is component c of unique node u.
#include <cmath>
#include <string>
#include <vector>

#include "tng_coarse.hpp"

namespace tng {

namespace {

// x = e_(node, comp) on every copy of the node
__global__ __launch_bounds__(256) void unit_vector_kernel(double *x0, double *x1, double *x2, const int32_t *uid,
                                                          int64_t ncopies, int node, int comp)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncopies) return;
    const bool on = uid[i] == node;
    x0[i] = on && comp == 0 ? 1.0 : 0.0;
    x1[i] = on && comp == 1 ? 1.0 : 0.0;
    x2[i] = on && comp == 2 ? 1.0 : 0.0;
}

// column `col` of the row-major n x n matrix from an applied unit vector
__global__ __launch_bounds__(256) void column_kernel(const double *y0, const double *y1, const double *y2,
                                                     const int32_t *rep, int nu, double *M, int n, int col)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int u = r / 3, c = r - 3 * u;
    const double *y = c == 0 ? y0 : (c == 1 ? y1 : y2);
    M[(int64_t)r * n + col] = y[rep[u]];
}

// Gauss-Jordan step k, part 1: row k and column k before the step
__global__ __launch_bounds__(256) void gj_copy_kernel(const double *M, int n, int k, double *rowk, double *colk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rowk[i] = M[(int64_t)k * n + i];
    colk[i] = M[(int64_t)i * n + k];
}

// part 2: every entry from the saved row and column (in place):
//   (k, k) -> 1 / p, (k, j) -> a_kj / p, (i, k) -> -a_ik / p,
//   (i, j) -> a_ij - a_ik a_kj / p
__global__ __launch_bounds__(256) void gj_update_kernel(double *M, int n, int k, const double *rowk,
                                                        const double *colk)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e - (int64_t)i * n);
    const double p = rowk[k];
    double v;
    if (i == k && j == k) v = 1.0 / p;
    else if (i == k) v = rowk[j] / p;
    else if (j == k) v = -colk[i] / p;
    else v = M[e] - colk[i] * (rowk[j] / p);
    M[e] = v;
}

// U = M Z and V = M^T W (n x 3 each, row-major), one thread per row
__global__ __launch_bounds__(256) void proj_vectors_kernel(const double *M, int n, const double *Z, const double *W,
                                                           double *U, double *V)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double u[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < n; ++j) {
        const double mij = M[(int64_t)i * n + j], mji = M[(int64_t)j * n + i];
        for (int k = 0; k < 3; ++k) {
            u[k] = fma(mij, Z[3 * j + k], u[k]);
            v[k] = fma(mji, W[3 * j + k], v[k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        U[3 * i + k] = u[k];
        V[3 * i + k] = v[k];
    }
}

struct Proj3 {
    double C[9];  // W^T M Z
};

// M <- P M P with P = I - Z W^T:  M - Z V^T - U W^T + Z C W^T
__global__ __launch_bounds__(256) void proj_update_kernel(double *M, int n, const double *Z, const double *W,
                                                          const double *U, const double *V, const Proj3 c)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e - (int64_t)i * n);
    double s = M[e];
    for (int k = 0; k < 3; ++k) s -= Z[3 * i + k] * V[3 * j + k] + U[3 * i + k] * W[3 * j + k];
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) s += Z[3 * i + k] * c.C[3 * k + l] * W[3 * j + l];
    M[e] = s;
}

__global__ __launch_bounds__(256) void diag_kernel(const double *M, int n, double *d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = M[(int64_t)i * n + i];
}

template <typename T>
__global__ __launch_bounds__(256) void gather_kernel(const T *r0, const T *r1, const T *r2, const int32_t *rep, int nu,
                                                     double *b)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int64_t i = rep[u];
    b[3 * u] = (double)r0[i];
    b[3 * u + 1] = (double)r1[i];
    b[3 * u + 2] = (double)r2[i];
}

// x = M b, one wave per row, lanes over the columns in a fixed order
__global__ __launch_bounds__(256) void gemv_kernel(const double *M, const double *b, double *x, int n)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    const double *m = M + (int64_t)row * n;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s = fma(m[j], b[j], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) x[row] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void scatter_kernel(T *e0, T *e1, T *e2, const int32_t *uid, int64_t ncopies,
                                                      const double *x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncopies) return;
    const int u = uid[i];
    e0[i] = (T)x[3 * u];
    e1[i] = (T)x[3 * u + 1];
    e2[i] = (T)x[3 * u + 2];
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void CoarseDirect::numbering(const tng_domain &m)
{
    const int64_t nc = m.nalloc;
    std::vector<int32_t> uid(nc), hidx((size_t)m.halo_maxc * m.halo_n);
    if (!hidx.empty())
        TNG_HIP_CHECK(hipMemcpy(hidx.data(), m.d_halo_idx.p, hidx.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nc; ++i) uid[i] = (int32_t)i;
    // every copy of a shared node points at the group's first copy
    for (int64_t t = 0; t < m.halo_n; ++t) {
        const int32_t first = hidx[t];
        for (int j = 1; j < m.halo_maxc; ++j) {
            const int32_t c = hidx[(size_t)j * m.halo_n + t];
            if (c >= 0) uid[c] = first;
        }
    }
    std::vector<int32_t> compact(nc, -1), rep;
    for (int64_t i = 0; i < nc; ++i)
        if (uid[i] == i) {
            compact[i] = (int32_t)rep.size();
            rep.push_back((int32_t)i);
        }
    for (int64_t i = 0; i < nc; ++i) uid[i] = compact[uid[i]];
    nu = (int)rep.size();
    n = 3 * nu;
    TNG_REQUIRE(n <= kMaxUnknowns, "direct coarse solve: the coarsest level has " + std::to_string(n) +
                                       " unknowns, more than " + std::to_string(kMaxUnknowns));
    ncopies = nc;
    d_uid.alloc(nc * sizeof(int32_t));
    d_rep.alloc(rep.size() * sizeof(int32_t));
    TNG_HIP_CHECK(hipMemcpy(d_uid.p, uid.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice));
    TNG_HIP_CHECK(hipMemcpy(d_rep.p, rep.data(), rep.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    d_M.alloc((size_t)n * n * sizeof(double));
    d_row.alloc(n * sizeof(double));
    d_col.alloc(n * sizeof(double));
    d_b.alloc(n * sizeof(double));
    d_x.alloc(n * sizeof(double));
}

void CoarseDirect::build(tng_operator &A, tng_vector &x, tng_vector &y, bool fp32_levels, hipStream_t s)
{
    TNG_REQUIRE(x.dtype == TNG_F64 && y.dtype == TNG_F64 && x.ncomp == 3 && y.ncomp == 3,
                "direct coarse solve: FP64 work vectors required");
    if (!d_uid.p) numbering(*A.mesh);
    double *M = d_M.as<double>();
    const int32_t *uid = d_uid.as<int32_t>(), *rep = d_rep.as<int32_t>();
    for (int col = 0; col < n; ++col) {
        hipLaunchKernelGGL(unit_vector_kernel, dim3(blocks(ncopies)), dim3(256), 0, s, (double *)x.comp(0),
                           (double *)x.comp(1), (double *)x.comp(2), uid, ncopies, col / 3, col % 3);
        op_apply(A, x, y);
        hipLaunchKernelGGL(column_kernel, dim3(blocks(n)), dim3(256), 0, s, (const double *)y.comp(0),
                           (const double *)y.comp(1), (const double *)y.comp(2), rep, nu, M, n, col);
    }
    for (int k = 0; k < n; ++k) {
        hipLaunchKernelGGL(gj_copy_kernel, dim3(blocks(n)), dim3(256), 0, s, (const double *)M, n, k,
                           d_row.as<double>(), d_col.as<double>());
        hipLaunchKernelGGL(gj_update_kernel, dim3(blocks((int64_t)n * n)), dim3(256), 0, s, M, n, k,
                           (const double *)d_row.p, (const double *)d_col.p);
    }
    TNG_HIP_CHECK(hipGetLastError());
    // the inverse of a symmetric positive definite matrix has a positive,
    // finite diagonal; a vanishing pivot leaves non-finite entries
    hipLaunchKernelGGL(diag_kernel, dim3(blocks(n)), dim3(256), 0, s, (const double *)M, n, d_b.as<double>());
    std::vector<double> diag(n);
    TNG_HIP_CHECK(hipMemcpyAsync(diag.data(), d_b.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    TNG_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i)
        TNG_REQUIRE(std::isfinite(diag[i]) && diag[i] > 0.0,
                    "direct coarse solve: the coarsest operator is not positive definite");
    if (A.penalty_active && fp32_levels) project_rotations(*A.mesh, s);
    ready = true;
}

// With the fs/fs penalty the three rigid rotations are near-null modes whose
// inverse gain is 1 / eps.  FP32 levels do not resolve the eps-sized
// penalty, so the rotational part of their restricted residual is rounding
// noise that an exact inverse would amplify: there the coarse correction is
// taken in the rotations' Euclidean complement, M <- P M P with
// P = I - Z (Z^T Z)^{-1} Z^T (Z: the rotations e_k x x at the unique nodes),
// and the rotations are left to the fine levels and the Krylov method.
void CoarseDirect::project_rotations(const tng_domain &m, hipStream_t s)
{
    std::vector<int32_t> rep(nu);
    TNG_HIP_CHECK(hipMemcpy(rep.data(), d_rep.p, nu * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<double> Z((size_t)n * 3), W((size_t)n * 3);
    for (int u = 0; u < nu; ++u) {
        const int64_t i = rep[u], lcol = i / m.nr;
        const int l = (int)(i - lcol * m.nr);
        const int64_t sd = lcol / ((int64_t)m.nx * m.nx);
        const double r = m.radii[sd * m.nr + l];
        const double x[3] = {m.coords[3 * lcol] * r, m.coords[3 * lcol + 1] * r, m.coords[3 * lcol + 2] * r};
        for (int k = 0; k < 3; ++k) {
            // e_k x x, component c at row 3 u + c
            const double e[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
            const double z[3] = {e[1] * x[2] - e[2] * x[1], e[2] * x[0] - e[0] * x[2], e[0] * x[1] - e[1] * x[0]};
            for (int c = 0; c < 3; ++c) Z[(size_t)(3 * u + c) * 3 + k] = z[c];
        }
    }
    // W = Z (Z^T Z)^{-1}
    double G[9] = {0}, Gi[9];
    for (int r = 0; r < n; ++r)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) G[3 * a + b] += Z[(size_t)r * 3 + a] * Z[(size_t)r * 3 + b];
    const double det = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) +
                       G[2] * (G[3] * G[7] - G[4] * G[6]);
    TNG_REQUIRE(std::isfinite(det) && det > 0.0, "direct coarse solve: degenerate rotation fields");
    Gi[0] = (G[4] * G[8] - G[5] * G[7]) / det;
    Gi[1] = (G[2] * G[7] - G[1] * G[8]) / det;
    Gi[2] = (G[1] * G[5] - G[2] * G[4]) / det;
    Gi[3] = (G[5] * G[6] - G[3] * G[8]) / det;
    Gi[4] = (G[0] * G[8] - G[2] * G[6]) / det;
    Gi[5] = (G[2] * G[3] - G[0] * G[5]) / det;
    Gi[6] = (G[3] * G[7] - G[4] * G[6]) / det;
    Gi[7] = (G[1] * G[6] - G[0] * G[7]) / det;
    Gi[8] = (G[0] * G[4] - G[1] * G[3]) / det;
    for (int r = 0; r < n; ++r)
        for (int a = 0; a < 3; ++a) {
            double v = 0.0;
            for (int b = 0; b < 3; ++b) v += Z[(size_t)r * 3 + b] * Gi[3 * b + a];
            W[(size_t)r * 3 + a] = v;
        }
    DevBuf dZ, dW, dU, dV;
    dZ.alloc(Z.size() * sizeof(double));
    dW.alloc(W.size() * sizeof(double));
    dU.alloc(Z.size() * sizeof(double));
    dV.alloc(Z.size() * sizeof(double));
    TNG_HIP_CHECK(hipMemcpyAsync(dZ.p, Z.data(), Z.size() * sizeof(double), hipMemcpyHostToDevice, s));
    TNG_HIP_CHECK(hipMemcpyAsync(dW.p, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice, s));
    double *M = d_M.as<double>();
    hipLaunchKernelGGL(proj_vectors_kernel, dim3(blocks(n)), dim3(256), 0, s, (const double *)M, n,
                       (const double *)dZ.p, (const double *)dW.p, dU.as<double>(), dV.as<double>());
    std::vector<double> U((size_t)n * 3);
    TNG_HIP_CHECK(hipMemcpyAsync(U.data(), dU.p, U.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    TNG_HIP_CHECK(hipStreamSynchronize(s));
    Proj3 c{};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double v = 0.0;
            for (int r = 0; r < n; ++r) v += W[(size_t)r * 3 + a] * U[(size_t)r * 3 + b];
            c.C[3 * a + b] = v;
        }
    hipLaunchKernelGGL(proj_update_kernel, dim3(blocks((int64_t)n * n)), dim3(256), 0, s, M, n, (const double *)dZ.p,
                       (const double *)dW.p, (const double *)dU.p, (const double *)dV.p, c);
    TNG_HIP_CHECK(hipGetLastError());
    TNG_HIP_CHECK(hipStreamSynchronize(s));
}

void CoarseDirect::solve(tng_vector &e0, const tng_vector &r0, hipStream_t s)
{
    TNG_REQUIRE(ready, "direct coarse solve: not built");
    const int32_t *uid = d_uid.as<int32_t>(), *rep = d_rep.as<int32_t>();
    double *b = d_b.as<double>(), *x = d_x.as<double>();
    if (r0.dtype == TNG_F64)
        hipLaunchKernelGGL(gather_kernel<double>, dim3(blocks(nu)), dim3(256), 0, s, (const double *)r0.comp(0),
                           (const double *)r0.comp(1), (const double *)r0.comp(2), rep, nu, b);
    else
        hipLaunchKernelGGL(gather_kernel<float>, dim3(blocks(nu)), dim3(256), 0, s, (const float *)r0.comp(0),
                           (const float *)r0.comp(1), (const float *)r0.comp(2), rep, nu, b);
    hipLaunchKernelGGL(gemv_kernel, dim3((n + 3) / 4), dim3(256), 0, s, (const double *)d_M.p, (const double *)b, x, n);
    if (e0.dtype == TNG_F64)
        hipLaunchKernelGGL(scatter_kernel<double>, dim3(blocks(ncopies)), dim3(256), 0, s, (double *)e0.comp(0),
                           (double *)e0.comp(1), (double *)e0.comp(2), uid, ncopies, (const double *)x);
    else
        hipLaunchKernelGGL(scatter_kernel<float>, dim3(blocks(ncopies)), dim3(256), 0, s, (float *)e0.comp(0),
                           (float *)e0.comp(1), (float *)e0.comp(2), uid, ncopies, (const double *)x);
    TNG_HIP_CHECK(hipGetLastError());
}

}  // namespace tng
