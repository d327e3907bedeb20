// Energy equation around the Stokes solve: the SUPG advection-diffusion
// operator of the implicit energy step, its Dirichlet data, the
// temperature-dependent viscosity and the advective time-step bound.
//
// SUPG operator (fe/wedge/operators/shell/unsteady_advection_diffusion_supg.hpp
// :19-49, :240-433; unsteady_advection_diffusion_supg_kerngen.hpp:1-60 for the
// collapsed form): per wedge and point q of the 6-point Felippa rule,
//   dst_i += w_q |det J_q| [ phi_i (m T(q) + dt u(q).grad T(q))
//                            + grad phi_i . (dt kappa grad T(q) + dt tau (u(q).grad T(q)) u(q)) ]
// with T(q), grad T(q) from the wedge's nodal values (boundary-layer values
// removed where the Dirichlet rows are treated), u(q) the interpolated
// velocity and tau the quadrature-weighted mean of supg_tau over the points
// (h = r_2 - r_1).  Boundary-layer rows keep their diagonal entry only; the
// diagonal mode keeps every row's diagonal entry; the lumped mass takes the
// row sums m sum_q w_q |det J_q| phi_i.  One thread per cell writes the 12
// wedge-node partials of the cell; a per-node gather sums them (no atomics).
#include <algorithm>

#include "tng_internal.hpp"

namespace tng {

namespace {

__constant__ int kEW[2][6][3] = {{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}},
                                 {{1, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 1}, {0, 1, 1}, {1, 0, 1}}};

__device__ double supg_tau_dev(double vel_norm, double kappa, double h, double pe_tol)
{
    const double kappa_virtual = fmax(kappa, 1e-8);
    const double eps_vel = 1e-12;
    if (vel_norm <= eps_vel) return 0.0;
    const double pe = vel_norm * h / (2.0 * kappa_virtual);
    if (pe <= pe_tol) return 0.0;
    const double tau_max = 10.0 * h / fmax(vel_norm, eps_vel);
    const double tau = (h / (2.0 * vel_norm)) * (1.0 / tanh(pe) - 1.0 / pe);
    return tau > tau_max ? tau_max : tau;
}

struct SupgArgs {
    const double *src, *u0, *u1, *u2;
    const double *coords, *radii;
    const uint8_t *bmask;
    double *partial;  // (cell, wedge, 6)
    double kappa, dt, mass;
    int treat_boundary, diagonal, lumped, supg;
    int nsd, nx, nr;
};

__global__ __launch_bounds__(128) void supg_cell_kernel(const SupgArgs a)
{
    const int nx = a.nx, nr = a.nr, nc = nx - 1, ncr = nr - 1;
    const int64_t ncell = (int64_t)a.nsd * nc * nc * ncr;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    const int rc = (int)(cell % ncr);
    int64_t t = cell / ncr;
    const int yc = (int)(t % nc);
    t /= nc;
    const int xc = (int)(t % nc);
    const int sd = (int)(t / nc);
    const double r1 = a.radii[(int64_t)sd * nr + rc], r2 = a.radii[(int64_t)sd * nr + rc + 1];
    const double h = r2 - r1;
    const uint8_t cmb = 2 | 4, surf = 2 | 8;
    const bool at_cmb = (a.bmask[(((int64_t)sd * nx + xc) * nx + yc) * nr + rc] & cmb) == cmb;
    const bool at_surf = (a.bmask[(((int64_t)sd * nx + xc) * nx + yc) * nr + rc + 1] & surf) == surf;
    const double qx[3] = {0.6666666666666666, 0.1666666666666667, 0.1666666666666667};
    const double qy[3] = {0.1666666666666667, 0.6666666666666666, 0.1666666666666667};
    const double qz[2] = {-0.5773502691896257, 0.5773502691896257};
    const double qw = 0.1666666666666667;
    for (int w = 0; w < 2; ++w) {
        double P[3][3], Tn[6], U[6][3];
        bool bnd[6];
        for (int n = 0; n < 6; ++n) {
            const int64_t lcol = ((int64_t)sd * nx + xc + kEW[w][n][0]) * nx + yc + kEW[w][n][1];
            const int64_t i = lcol * nr + rc + kEW[w][n][2];
            if (n < 3)
                for (int d = 0; d < 3; ++d) P[n][d] = a.coords[3 * lcol + d];
            Tn[n] = a.src[i];
            U[n][0] = a.u0[i];
            U[n][1] = a.u1[i];
            U[n][2] = a.u2[i];
            bnd[n] = a.treat_boundary && ((at_cmb && n < 3) || (at_surf && n >= 3));
        }
        // velocity at the points and tau
        double uq[6][3];
        double tau_acc = 0.0, wacc = 0.0;
        for (int q = 0; q < 6; ++q) {
            const double xi = qx[q % 3], eta = qy[q % 3], zeta = qz[q / 3];
            const double Nl[3] = {1.0 - xi - eta, xi, eta}, Nr[2] = {0.5 * (1.0 - zeta), 0.5 * (1.0 + zeta)};
            for (int d = 0; d < 3; ++d) {
                double s = 0.0;
                for (int n = 0; n < 6; ++n) s += U[n][d] * (Nl[n % 3] * Nr[n / 3]);
                uq[q][d] = s;
            }
            const double vn = sqrt(uq[q][0] * uq[q][0] + uq[q][1] * uq[q][1] + uq[q][2] * uq[q][2]);
            tau_acc += supg_tau_dev(vn, a.kappa, h, 1e-08) * qw;
            wacc += qw;
        }
        const double tau = a.supg ? (wacc > 0.0 ? tau_acc / wacc : 0.0) : 0.0;
        double out[6] = {0, 0, 0, 0, 0, 0}, diag[6] = {0, 0, 0, 0, 0, 0}, lmass[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 6; ++q) {
            const double xi = qx[q % 3], eta = qy[q % 3], zeta = qz[q / 3];
            const double Nl[3] = {1.0 - xi - eta, xi, eta}, Nr[2] = {0.5 * (1.0 - zeta), 0.5 * (1.0 + zeta)};
            const double r = r1 + 0.5 * (r2 - r1) * (1.0 + zeta), gr = 0.5 * (r2 - r1);
            double J[3][3];
            for (int d = 0; d < 3; ++d) {
                J[d][0] = (P[1][d] - P[0][d]) * r;
                J[d][1] = (P[2][d] - P[0][d]) * r;
                J[d][2] = (Nl[0] * P[0][d] + Nl[1] * P[1][d] + Nl[2] * P[2][d]) * gr;
            }
            const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                         c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
            const double id = 1.0 / det;
            // (J^{-T})_{ab}
            double iT[3][3];
            iT[0][0] = c00 * id; iT[0][1] = c01 * id; iT[0][2] = c02 * id;
            iT[1][0] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
            iT[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
            iT[1][2] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
            iT[2][0] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
            iT[2][1] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
            iT[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
            const double wd = qw * fabs(det);
            const double gxi[3] = {-1.0, 1.0, 0.0}, geta[3] = {-1.0, 0.0, 1.0}, grad_rad[2] = {-0.5, 0.5};
            double g[6][3], phi[6];
            for (int n = 0; n < 6; ++n) {
                const double gref[3] = {gxi[n % 3] * Nr[n / 3], geta[n % 3] * Nr[n / 3], Nl[n % 3] * grad_rad[n / 3]};
                for (int b = 0; b < 3; ++b) g[n][b] = iT[b][0] * gref[0] + iT[b][1] * gref[1] + iT[b][2] * gref[2];
                phi[n] = Nl[n % 3] * Nr[n / 3];
            }
            // trial values with the Dirichlet columns removed
            double Th = 0.0, gT[3] = {0.0, 0.0, 0.0};
            for (int n = 0; n < 6; ++n) {
                const double tv = bnd[n] ? 0.0 : Tn[n];
                Th += phi[n] * tv;
                for (int b = 0; b < 3; ++b) gT[b] += g[n][b] * tv;
            }
            const double ugT = uq[q][0] * gT[0] + uq[q][1] * gT[1] + uq[q][2] * gT[2];
            const double As = a.mass * Th + a.dt * ugT;
            const double As_nomass = a.dt * ugT;
            double Bv[3];
            for (int b = 0; b < 3; ++b) Bv[b] = a.dt * a.kappa * gT[b] + a.dt * tau * ugT * uq[q][b];
            for (int i = 0; i < 6; ++i) {
                const double gB = g[i][0] * Bv[0] + g[i][1] * Bv[1] + g[i][2] * Bv[2];
                out[i] += wd * (phi[i] * (a.lumped ? As_nomass : As) + gB);
                // diagonal entry (M_ii + A_ii)
                const double ugi = uq[q][0] * g[i][0] + uq[q][1] * g[i][1] + uq[q][2] * g[i][2];
                const double gg = g[i][0] * g[i][0] + g[i][1] * g[i][1] + g[i][2] * g[i][2];
                const double mii = a.mass * phi[i] * phi[i];
                diag[i] += wd * ((a.lumped ? 0.0 : mii) + a.dt * (a.kappa * gg + ugi * phi[i] + tau * ugi * ugi));
                lmass[i] += wd * a.mass * phi[i];
            }
        }
        double *o = a.partial + (cell * 2 + w) * 6;
        for (int i = 0; i < 6; ++i) {
            double v;
            if (a.diagonal || bnd[i]) v = (diag[i] + (a.lumped ? lmass[i] : 0.0)) * Tn[i];
            else v = out[i] + (a.lumped ? lmass[i] * Tn[i] : 0.0);
            o[i] = v;
        }
    }
}

// g: the Dirichlet temperatures at the boundary nodes (CMB first), zero elsewhere
__global__ __launch_bounds__(256) void boundary_values_kernel(double *g, const uint8_t *bmask, double t_cmb,
                                                              double t_surf, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t f = bmask[i], cmb = 2 | 4, surf = 2 | 8;
    g[i] = (f & cmb) == cmb ? t_cmb : ((f & surf) == surf ? t_surf : 0.0);
}

// b = src at the nodes carrying all bits of `flag`, b unchanged elsewhere
__global__ __launch_bounds__(256) void assign_masked_kernel(double *b, const double *src, const uint8_t *bmask,
                                                            uint8_t flag, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if ((bmask[i] & flag) == flag) b[i] = src[i];
}

// eta = rmu^(0.5 - T) (Frank-Kamenetskii, interpolators.hpp:219-244)
__global__ __launch_bounds__(256) void viscosity_fk_kernel(double *eta, const double *temp, double rmu, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    eta[i] = pow(rmu, 0.5 - temp[i]);
}

// per-block max of |u| over all copies; the last kernel combines
__global__ __launch_bounds__(256) void max_magnitude_kernel(const double *u0, const double *u1, const double *u2,
                                                            int64_t n, double *partials)
{
    __shared__ double red[4];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, sqrt(u0[i] * u0[i] + u1[i] * u1[i] + u2[i] * u2[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ __launch_bounds__(256) void max_final_kernel(const double *partials, int nb, double *out)
{
    __shared__ double red[4];
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmax(m, partials[b]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) *out = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

unsigned grid1d(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void supg_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const tng_domain &m = *op.mesh;
    TNG_REQUIRE(op.vel && op.vel->mesh == &m && op.vel->ncomp == 3 && op.vel->dtype == TNG_F64,
                "SUPG: the velocity must be an FP64 vector on the operator's mesh");
    TNG_REQUIRE(src.dtype == TNG_F64 && m.nr >= 2, "SUPG: FP64 temperature");
    const int64_t ncell = (int64_t)m.nsd * (m.nx - 1) * (m.nx - 1) * (m.nr - 1);
    const size_t bytes = (size_t)ncell * 12 * sizeof(double);
    if (op.scratch.bytes < bytes) op.scratch.alloc(bytes);
    SupgArgs a{};
    a.src = (const double *)src.comp(0);
    a.u0 = (const double *)op.vel->comp(0);
    a.u1 = (const double *)op.vel->comp(1);
    a.u2 = (const double *)op.vel->comp(2);
    a.coords = m.d_coords.as<double>();
    a.radii = m.d_radii.as<double>();
    a.bmask = m.d_bmask.as<uint8_t>();
    a.partial = op.scratch.as<double>();
    a.kappa = op.supg_kappa;
    a.dt = op.supg_dt;
    a.mass = op.supg_mass_scaling;
    a.treat_boundary = op.supg_treat_boundary ? 1 : 0;
    a.diagonal = op.diagonal ? 1 : 0;
    a.lumped = op.supg_lumped ? 1 : 0;
    a.supg = op.supg_enabled ? 1 : 0;
    a.nsd = m.nsd;
    a.nx = m.nx;
    a.nr = m.nr;
    hipLaunchKernelGGL(supg_cell_kernel, dim3((unsigned)((ncell + 127) / 128)), dim3(128), 0, s, a);
    TNG_HIP_CHECK(hipGetLastError());
    launch_wedge_gather(op.scratch.as<double>(), dst, op.apply_mode == TNG_ADD, s);
    if (op.comm_mode == TNG_COMMUNICATE_ADDITIVELY) launch_halo_sum(m, dst, s);
}

void launch_boundary_values(tng_vector &g, double t_cmb, double t_surf, hipStream_t s)
{
    const tng_domain &m = *g.mesh;
    hipLaunchKernelGGL(boundary_values_kernel, dim3(grid1d(m.nalloc)), dim3(256), 0, s, (double *)g.comp(0),
                       m.d_bmask.as<uint8_t>(), t_cmb, t_surf, m.nalloc);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_assign_masked(tng_vector &b, const tng_vector &src, uint8_t flag, hipStream_t s)
{
    const tng_domain &m = *b.mesh;
    hipLaunchKernelGGL(assign_masked_kernel, dim3(grid1d(m.nalloc)), dim3(256), 0, s, (double *)b.comp(0),
                       (const double *)src.comp(0), m.d_bmask.as<uint8_t>(), flag, m.nalloc);
    TNG_HIP_CHECK(hipGetLastError());
}

void launch_viscosity_fk(tng_vector &eta, const tng_vector &temp, double rmu, hipStream_t s)
{
    const tng_domain &m = *eta.mesh;
    hipLaunchKernelGGL(viscosity_fk_kernel, dim3(grid1d(m.nalloc)), dim3(256), 0, s, (double *)eta.comp(0),
                       (const double *)temp.comp(0), rmu, m.nalloc);
    TNG_HIP_CHECK(hipGetLastError());
}

double max_vector_magnitude(const tng_vector &u, hipStream_t s)
{
    const tng_domain &m = *u.mesh;
    TNG_REQUIRE(u.ncomp == 3 && u.dtype == TNG_F64, "max |u|: FP64 velocity");
    tng_context &ctx = *m.ctx;
    const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (m.nalloc + 255) / 256));
    double *part = ctx.red_partials.as<double>(), *out = ctx.red_result.as<double>() + 300;
    hipLaunchKernelGGL(max_magnitude_kernel, dim3(nb), dim3(256), 0, s, (const double *)u.comp(0),
                       (const double *)u.comp(1), (const double *)u.comp(2), m.nalloc, part);
    hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, s, (const double *)part, nb, out);
    TNG_HIP_CHECK(hipGetLastError());
    double v = 0.0;
    TNG_HIP_CHECK(hipMemcpyAsync(&v, out, sizeof(double), hipMemcpyDeviceToHost, s));
    TNG_HIP_CHECK(hipStreamSynchronize(s));
    return v;
}

}  // namespace tng
