// Coefficient gradient of the viscous operator, the element kernel of the
// viscosity adjoint:  g_i = d(lam^T A(k) u) / d log k_i.
//
// A is linear in the wedge coefficient k_w = mean of the six nodal k, so
//   g_i = (k_i / 6) sum_{wedges w of copy i} Phi_w,   Phi_w = lam_w^T A_w(k=1) u_w,
// summed afterwards over the copies of each node.  With one-point quadrature
// Phi_w = |det J| [2 eps(lam_t) : eps(u_t) - 2/3 div lam_t div u_t] on the
// boundary-modified trial and test values (Dirichlet: boundary values removed;
// free slip: tangential projection), plus the wedge's share of the boundary
// rows: Dirichlet  lam_b . diag_b u_b,  diag_b[d] = |det J| (|g_b|^2 + g_b[d]^2/3);
// free slip  (n.lam_b)(n.u_b) |det J| (|g_b|^2 + (g_b.n)^2/3).
// Two passes, no atomics: Phi per wedge into scratch, then a per-node gather.
#include "tng_internal.hpp"

namespace tng {

namespace {

__constant__ int kAdjWC[2][3][2] = {{{0, 0}, {1, 0}, {0, 1}}, {{1, 1}, {0, 1}, {1, 0}}};

struct PhiArgs {
    const double *u0, *u1, *u2, *l0, *l1, *l2;
    const double *coords, *radii, *latgeom;
    const uint8_t *sd_cmb, *sd_surf;
    double *phi;  // (nsd, nc, nc, ncr, 2)
    int nsd, nx, nr, bc_cmb, bc_surf;
};

__device__ __forceinline__ void unit(const double *c, double n[3])
{
    const double m = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    n[0] = c[0] / m; n[1] = c[1] / m; n[2] = c[2] / m;
}

__global__ __launch_bounds__(256) void phi_kernel(const PhiArgs a)
{
    const int nx = a.nx, nr = a.nr, nc = nx - 1, ncr = nr - 1;
    const int64_t n = (int64_t)a.nsd * nc * nc * ncr;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int rc = (int)(gid % ncr);
    int64_t t = gid / ncr;
    const int yc = (int)(t % nc); t /= nc;
    const int xc = (int)(t % nc);
    const int sd = (int)(t / nc);
    const bool has_cmb = a.sd_cmb[sd] != 0, has_surf = a.sd_surf[sd] != 0;
    int layer = -1, cbc = TNG_NEUMANN;
    if (rc == 0 && has_cmb) { layer = 0; cbc = a.bc_cmb; }
    else if (rc + 1 == nr - 1 && has_surf) { layer = 1; cbc = a.bc_surf; }
    if (cbc == TNG_NEUMANN) layer = -1;
    const double r0 = a.radii[(int64_t)sd * nr + rc], r1 = a.radii[(int64_t)sd * nr + rc + 1];
    const double rmid = 0.5 * (r0 + r1), h = 0.5 * (r1 - r0);
    const double s = 0.5 / rmid, dz = 1.0 / (6.0 * h);
    const double dxi[3] = {-1.0, 1.0, 0.0}, deta[3] = {-1.0, 0.0, 1.0};
    for (int w = 0; w < 2; ++w) {
        const double *G = a.latgeom + ((((int64_t)sd * nc + xc) * nc + yc) * 2 + w) * 10;
        const double detJ = G[9] * rmid * rmid * h;
        double U[3][3] = {{0}}, L[3][3] = {{0}};
        double bterm = 0.0;
        for (int j = 0; j < 6; ++j) {
            const int c = j % 3, e = j / 3;
            const int X = xc + kAdjWC[w][c][0], Y = yc + kAdjWC[w][c][1];
            const int64_t col = ((int64_t)sd * nx + X) * nx + Y;
            const int64_t i = col * nr + rc + e;
            double g[3];
            for (int d = 0; d < 3; ++d)
                g[d] = s * (dxi[c] * G[d] + deta[c] * G[3 + d]) + (e == 0 ? -dz : dz) * G[6 + d];
            double uu[3] = {a.u0[i], a.u1[i], a.u2[i]}, ll[3] = {a.l0[i], a.l1[i], a.l2[i]};
            if (e == layer) {
                const double gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
                if (cbc == TNG_DIRICHLET) {
                    for (int d = 0; d < 3; ++d) bterm += ll[d] * detJ * (gg + g[d] * g[d] * (1.0 / 3.0)) * uu[d];
                    continue;  // removed from the interior form
                }
                double nn[3];
                unit(a.coords + 3 * col, nn);
                const double un = nn[0] * uu[0] + nn[1] * uu[1] + nn[2] * uu[2];
                const double ln = nn[0] * ll[0] + nn[1] * ll[1] + nn[2] * ll[2];
                const double gn = g[0] * nn[0] + g[1] * nn[1] + g[2] * nn[2];
                bterm += ln * un * detJ * (gg + gn * gn * (1.0 / 3.0));
                for (int d = 0; d < 3; ++d) {
                    uu[d] -= un * nn[d];
                    ll[d] -= ln * nn[d];
                }
            }
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    U[p][q] += uu[p] * g[q];
                    L[p][q] += ll[p] * g[q];
                }
        }
        double dc = 0.0;
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) dc += L[p][q] * (U[p][q] + U[q][p]);
        const double divu = U[0][0] + U[1][1] + U[2][2], divl = L[0][0] + L[1][1] + L[2][2];
        a.phi[gid * 2 + w] = detJ * (dc - (2.0 / 3.0) * divl * divu) + bterm;
    }
}

__global__ __launch_bounds__(256) void phi_gather_kernel(const double *phi, const double *k, double *g, int nsd, int nx,
                                                         int nr)
{
    const int64_t n = (int64_t)nsd * nx * nx * nr;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int l = (int)(gid % nr);
    int64_t t = gid / nr;
    const int Y = (int)(t % nx); t /= nx;
    const int X = (int)(t % nx);
    const int sd = (int)(t / nx);
    const int nc = nx - 1, ncr = nr - 1;
    const int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};
    double s = 0.0;
    for (int dxc = -1; dxc <= 0; ++dxc) {
        const int xc = X + dxc;
        if (xc < 0 || xc >= nc) continue;
        for (int dyc = -1; dyc <= 0; ++dyc) {
            const int yc = Y + dyc;
            if (yc < 0 || yc >= nc) continue;
            for (int w = 0; w < 2; ++w) {
                if (corner_of[w][X - xc][Y - yc] < 0) continue;
                for (int rc = l - 1; rc <= l; ++rc) {
                    if (rc < 0 || rc >= ncr) continue;
                    s += phi[((((int64_t)sd * nc + xc) * nc + yc) * ncr + rc) * 2 + w];
                }
            }
        }
    }
    g[gid] = s * k[gid] * (1.0 / 6.0);
}

}  // namespace

void epsdivdiv_log_k_gradient(tng_operator &op, const tng_vector &lam, const tng_vector &u, tng_vector &g,
                              hipStream_t s)
{
    const tng_domain &m = *op.mesh;
    TNG_REQUIRE(op.kind == OP_EPSDIVDIV, "coefficient gradient of the viscous operator only");
    TNG_REQUIRE(op.k->dtype == TNG_F64 && lam.dtype == TNG_F64 && u.dtype == TNG_F64 && g.dtype == TNG_F64,
                "coefficient gradient: FP64 operator and vectors");
    TNG_REQUIRE(lam.mesh == &m && u.mesh == &m && g.mesh == &m && lam.ncomp == 3 && u.ncomp == 3 && g.ncomp == 1,
                "coefficient gradient: lam, u (velocity) and g (scalar) on the operator's mesh");
    TNG_REQUIRE(m.nr >= 3, "coefficient gradient: at least two radial layers");
    const int64_t ncell = (int64_t)m.nsd * (m.nx - 1) * (m.nx - 1) * (m.nr - 1);
    const size_t bytes = (size_t)ncell * 2 * sizeof(double);
    if (op.scratch.bytes < bytes) op.scratch.alloc(bytes);
    PhiArgs a{};
    a.u0 = (const double *)u.comp(0); a.u1 = (const double *)u.comp(1); a.u2 = (const double *)u.comp(2);
    a.l0 = (const double *)lam.comp(0); a.l1 = (const double *)lam.comp(1); a.l2 = (const double *)lam.comp(2);
    a.coords = m.d_coords.as<double>(); a.radii = m.d_radii.as<double>(); a.latgeom = m.d_latgeom.as<double>();
    a.sd_cmb = m.d_sd_cmb.as<uint8_t>(); a.sd_surf = m.d_sd_surf.as<uint8_t>();
    a.phi = op.scratch.as<double>();
    a.nsd = m.nsd; a.nx = m.nx; a.nr = m.nr; a.bc_cmb = op.bc_cmb; a.bc_surf = op.bc_surf;
    hipLaunchKernelGGL(phi_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s, a);
    TNG_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(phi_gather_kernel, dim3((unsigned)((m.nalloc + 255) / 256)), dim3(256), 0, s,
                       (const double *)op.scratch.p, (const double *)op.k->comp(0), (double *)g.comp(0), m.nsd, m.nx,
                       m.nr);
    TNG_HIP_CHECK(hipGetLastError());
    launch_halo_sum(m, g, s);
}

}  // namespace tng
