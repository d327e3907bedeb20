// Stored local matrices of the EpsilonDivDiv operator and the Galerkin
// coarse-grid approximation (GCA).
//
//  - local_matrix_kernel: the Neumann 18x18 matrix of every wedge
//    (epsilon_divdiv_kerngen.hpp:2315-2392 assemble_local_matrix):
//    A(i + 6 di, j + 6 dj) = k |det J| [ (g_i . g_j) d_{di,dj} + g_i[dj] g_j[di]
//                                        - 2/3 g_i[di] g_j[dj] ]
//    with g_n = J^{-T} grad N_n at the centroid and k the mean nodal k.
//  - gca_onthefly_kernel / gca_stored_kernel: per coarse cell, sum over its
//    8 fine cells x 2 wedges of P^T A_fine P (linalg/solvers/gca/gca.hpp
//    :153-397), A_fine assembled on the fly from k or read from storage.
//  - stored_apply_kernel: the slow path on stored matrices
//    (epsilon_divdiv_kerngen.hpp:1013-1242) as a per-node gather: a node
//    collects rows of the <= 12 wedges around it, so every matrix row is read
//    once and every output written once (no atomics).  Dirichlet: boundary
//    trial values removed, boundary rows = diagonal x src.  Free slip: trial
//    values projected on the tangent plane, test rows projected, plus the
//    normal-normal entry times the normal component (the reference's
//    R A R^T with masked normal couplings, rotated back).
#include "tng_internal.hpp"

namespace tng {

namespace {

// wedge node n = lateral corner n % 3 (kWN offsets), radial n / 3
__constant__ int kWN[2][3][2] = {{{0, 0}, {1, 0}, {0, 1}}, {{1, 1}, {0, 1}, {1, 0}}};

__host__ __device__ __forceinline__ int64_t lmat_index(int nc, int ncr, int sd, int xc, int yc, int rc, int w)
{
    return (((((int64_t)sd * nc + xc) * nc + yc) * ncr + rc) * 2 + w) * 324;
}

// Geometry of one wedge at the centroid: |det J| k_eval and the six physical
// gradients (wedge node order: bottom corners 0..2, top corners 3..5).
__device__ void wedge_gradients(const double *coords, const double *radii, const double *k, int nx, int nr, int sd,
                                int xc, int yc, int rc, int w, double g[6][3], double *jkw)
{
    double p[3][3];
    for (int c = 0; c < 3; ++c) {
        const double *q = coords + 3 * (((int64_t)sd * nx + xc + kWN[w][c][0]) * nx + yc + kWN[w][c][1]);
        p[c][0] = q[0]; p[c][1] = q[1]; p[c][2] = q[2];
    }
    const double r0 = radii[(int64_t)sd * nr + rc], r1 = radii[(int64_t)sd * nr + rc + 1];
    const double rmid = 0.5 * (r0 + r1), h = 0.5 * (r1 - r0);
    double J[3][3];
    for (int i = 0; i < 3; ++i) {
        J[i][0] = rmid * (p[1][i] - p[0][i]);
        J[i][1] = rmid * (p[2][i] - p[0][i]);
        J[i][2] = h * ((p[0][i] + p[1][i] + p[2][i]) * (1.0 / 3.0));
    }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double id = 1.0 / det;
    // (J^{-T})_{ab} = cof(J)_{ab} / det
    double iT[3][3];
    iT[0][0] = id * (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
    iT[0][1] = id * (J[1][2] * J[2][0] - J[1][0] * J[2][2]);
    iT[0][2] = id * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    iT[1][0] = id * (J[0][2] * J[2][1] - J[0][1] * J[2][2]);
    iT[1][1] = id * (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
    iT[1][2] = id * (J[0][1] * J[2][0] - J[0][0] * J[2][1]);
    iT[2][0] = id * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
    iT[2][1] = id * (J[0][2] * J[1][0] - J[0][0] * J[1][2]);
    iT[2][2] = id * (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    // reference gradients at (1/3, 1/3, 0)
    const double dxi[3] = {-0.5, 0.5, 0.0}, deta[3] = {-0.5, 0.0, 0.5};
    double ks = 0.0;
    for (int n = 0; n < 6; ++n) {
        const int c = n % 3, e = n / 3;
        const double dz = e == 0 ? -1.0 / 6.0 : 1.0 / 6.0;
        for (int a = 0; a < 3; ++a) g[n][a] = iT[a][0] * dxi[c] + iT[a][1] * deta[c] + iT[a][2] * dz;
        ks += k[(((int64_t)sd * nx + xc + kWN[w][c][0]) * nx + yc + kWN[w][c][1]) * nr + rc + e];
    }
    *jkw = (ks * (1.0 / 6.0)) * fabs(det);
}

__device__ __forceinline__ double local_entry(const double g[6][3], double jkw, int row, int col)
{
    const int di = row / 6, i = row % 6, dj = col / 6, j = col % 6;
    const double gg = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
    return jkw * ((di == dj ? gg : 0.0) + g[i][dj] * g[j][di] - (2.0 / 3.0) * g[i][di] * g[j][dj]);
}

struct LocalArgs {
    const double *coords, *radii, *k;
    double *out;
    int nsd, nx, nr;
};

// one thread per wedge
__global__ __launch_bounds__(128) void local_matrix_kernel(const LocalArgs a)
{
    const int nc = a.nx - 1, ncr = a.nr - 1;
    const int64_t n = (int64_t)a.nsd * nc * nc * ncr * 2;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int w = (int)(gid & 1);
    int64_t t = gid >> 1;
    const int rc = (int)(t % ncr); t /= ncr;
    const int yc = (int)(t % nc); t /= nc;
    const int xc = (int)(t % nc);
    const int sd = (int)(t / nc);
    double g[6][3], jkw;
    wedge_gradients(a.coords, a.radii, a.k, a.nx, a.nr, sd, xc, yc, rc, w, g, &jkw);
    double *o = a.out + gid * 324;
    for (int e = 0; e < 324; ++e) o[e] = local_entry(g, jkw, e / 18, e % 18);
}

// prolongation_constant_weight (fe/wedge/shell/grid_transfer_constant.hpp:24-87)
__device__ __forceinline__ double pc_weight(int xf, int yf, int rf, int xc, int yc, int rc)
{
    const int dx = xf - 2 * xc, dy = yf - 2 * yc, dr = rf - 2 * rc;
    double lat = 0.0;
    if (dx == 0 && dy == 0) lat = 1.0;
    else if ((dy == 0 && (dx == 1 || dx == -1)) || (dx == 0 && (dy == 1 || dy == -1)) || (dx == -1 && dy == 1) ||
             (dx == 1 && dy == -1))
        lat = 0.5;
    if (dr == 0) return lat;
    if (dr == 1 || dr == -1) return 0.5 * lat;
    return 0.0;
}

// Row n of the constant interpolation P (6 x 6) from the parent coarse wedge's
// nodes to node n of fine wedge w of the fine hex h of a coarse cell
// (gca.hpp:190-361).  The fine node sits at offsets (ox, oy, orr) in 0..2 from
// the coarse cell's first node, so P depends only on (h, w, n), on whether the
// coarse cell is the top cell of its subdomain, and on the top-face rule.
// The reference takes the coarse bottom layer of a fine node as r_fine / 2
// (gca.hpp:206-208), except at the subdomain's top node level, where it takes
// r_fine / 2 - 1: fine nodes on the top face of an interior coarse cell at odd
// lateral positions then draw their weights from the coarse top layer but
// store them in the bottom slots.  exact = true uses the cell's own bottom
// layer everywhere, which gives the Galerkin product R A P.
__device__ __forceinline__ void gca_prow(int h, int w, int n, bool top_cell, bool exact, double prow[6])
{
    const int ox = (h & 1) + kWN[w][n % 3][0], oy = ((h >> 1) & 1) + kWN[w][n % 3][1], orr = (h >> 2) + n / 3;
    for (int c = 0; c < 6; ++c) prow[c] = 0.0;
    if (ox % 2 == 0 && oy % 2 == 0 && orr % 2 == 0) {
        prow[n] = 1.0;
        return;
    }
    if (ox % 2 == 0 && oy % 2 == 0) {
        prow[n % 3] = 0.5;
        prow[n % 3 + 3] = 0.5;
        return;
    }
    const int rb = (orr == 2 && (top_cell || exact)) ? 0 : orr / 2;
    int x0, y0, x1, y1, l0, l1, l2, l3;
    if (ox % 2 == 0) {
        x0 = ox / 2; x1 = ox / 2; y0 = oy / 2; y1 = oy / 2 + 1; l0 = 0; l1 = 2; l2 = 3; l3 = 5;
    } else if (oy % 2 == 0) {
        x0 = ox / 2; x1 = ox / 2 + 1; y0 = oy / 2; y1 = oy / 2; l0 = 0; l1 = 1; l2 = 3; l3 = 4;
    } else {
        x0 = ox / 2 + 1; x1 = ox / 2; y0 = oy / 2; y1 = oy / 2 + 1; l0 = 1; l1 = 2; l2 = 4; l3 = 5;
    }
    prow[l0] = pc_weight(ox, oy, orr, x0, y0, rb);
    prow[l1] = pc_weight(ox, oy, orr, x1, y1, rb);
    prow[l2] = pc_weight(ox, oy, orr, x0, y0, rb + 1);
    prow[l3] = pc_weight(ox, oy, orr, x1, y1, rb + 1);
}

// The 8 fine wedges (2 h + w) that each coarse wedge collects, in the order
// h-major, w-minor (gca.hpp:374-391).
__constant__ int kGcaFrom[2][8] = {{0, 1, 2, 4, 8, 9, 10, 12}, {3, 5, 6, 7, 11, 13, 14, 15}};

struct GcaArgs {
    const double *lf;                   // stored fine matrices, or nullptr: assemble from k
    const double *coords, *radii, *k;   // fine mesh, fine k
    double *lc;                         // coarse matrices
    int nsd, nxf, nrf, nxc, nrc;
    int exact;                          // 1: Galerkin product R A P on the top faces too
    int64_t ncell;                      // coarse cells
};

constexpr int kGcaCells = 16;  // coarse cells per workgroup of the on-the-fly kernel

// Fine operator assembled on the fly (Neumann local matrices from k).  Every
// fine matrix entry is a bilinear form of two nodal gradients,
//   A_f((i,di),(j,dj)) = jkw [ (g_i . g_j) d_{di,dj} + g_i[dj] g_j[di] - 2/3 g_i[di] g_j[dj] ],
// so P^T A_f P is the same form of the combined gradients G_c = sum_n P[n][c] g_n.
// Phase 1: one thread per fine wedge of 16 coarse cells computes its geometry
// and G (6 x 3) into LDS.  Phase 2: the 16 x 648 coarse entries, each a sum
// over the 8 fine wedges of its coarse wedge, written contiguously.
__global__ __launch_bounds__(256) void gca_onthefly_kernel(const GcaArgs a)
{
    __shared__ double sG[kGcaCells][16][6][3];
    __shared__ double sJ[kGcaCells][16];
    const int ncc = a.nxc - 1, ncrc = a.nrc - 1;
    const int64_t cell0 = (int64_t)blockIdx.x * kGcaCells;
    const int ncl = (int)min<int64_t>(kGcaCells, a.ncell - cell0);
    {
        const int cl = threadIdx.x >> 4, fw = threadIdx.x & 15;
        if (cl < ncl) {
            int64_t t = cell0 + cl;
            const int Rr = (int)(t % ncrc); t /= ncrc;
            const int Y = (int)(t % ncc); t /= ncc;
            const int X = (int)(t % ncc);
            const int sd = (int)(t / ncc);
            const int h = fw >> 1, w = fw & 1;
            const int fx = 2 * X + (h & 1), fy = 2 * Y + ((h >> 1) & 1), fr = 2 * Rr + (h >> 2);
            double g[6][3], jkw;
            wedge_gradients(a.coords, a.radii, a.k, a.nxf, a.nrf, sd, fx, fy, fr, w, g, &jkw);
            double G[6][3];
            for (int c = 0; c < 6; ++c)
                for (int d = 0; d < 3; ++d) G[c][d] = 0.0;
            const bool top = Rr == ncrc - 1;
            for (int n = 0; n < 6; ++n) {
                double prow[6];
                gca_prow(h, w, n, top, a.exact != 0, prow);
                for (int c = 0; c < 6; ++c)
                    if (prow[c] != 0.0)
                        for (int d = 0; d < 3; ++d) G[c][d] = fma(prow[c], g[n][d], G[c][d]);
            }
            for (int c = 0; c < 6; ++c)
                for (int d = 0; d < 3; ++d) sG[cl][fw][c][d] = G[c][d];
            sJ[cl][fw] = jkw;
        }
    }
    __syncthreads();
    double *out = a.lc + cell0 * 648;
    for (int o = threadIdx.x; o < ncl * 648; o += 256) {
        const int cl = o / 648, rem = o - cl * 648;
        const int tw = rem >= 324 ? 1 : 0, e = rem - tw * 324;
        const int row = e / 18, col = e % 18;
        const int di = row / 6, ii = row % 6, dj = col / 6, jj = col % 6;
        double s = 0.0;
        for (int q = 0; q < 8; ++q) {
            const int fw = kGcaFrom[tw][q];
            const double *Gi = sG[cl][fw][ii], *Gj = sG[cl][fw][jj];
            const double gg = Gi[0] * Gj[0] + Gi[1] * Gj[1] + Gi[2] * Gj[2];
            s += sJ[cl][fw] * ((di == dj ? gg : 0.0) + Gi[dj] * Gj[di] - (2.0 / 3.0) * Gi[di] * Gj[dj]);
        }
        out[o] = s;
    }
}

// Fine operator from stored matrices: one workgroup per coarse cell loads the
// 16 fine matrices (41 KB) into LDS with every load issued before the first
// use, the 16 P matrices beside them, then forms the 648 coarse entries
// sum_{a,b} P[a][i] A_f((a,di),(b,dj)) P[b][j] over the 8 fine wedges of each
// coarse wedge, skipping zero weights.
__global__ __launch_bounds__(256) void gca_stored_kernel(const GcaArgs a)
{
    __shared__ double sA[16][324];
    __shared__ double sP[16][6][6];
    const int ncc = a.nxc - 1, ncrc = a.nrc - 1, ncf = a.nxf - 1, ncrf = a.nrf - 1;
    int64_t t = blockIdx.x;
    const int Rr = (int)(t % ncrc); t /= ncrc;
    const int Y = (int)(t % ncc); t /= ncc;
    const int X = (int)(t % ncc);
    const int sd = (int)(t / ncc);
    // the fine matrices of hexes h and h + 4 (radial neighbours), both
    // wedges, are 4 x 324 contiguous doubles
    for (int i = threadIdx.x; i < 16 * 324; i += 256) {
        const int blk = i / 1296, off = i - blk * 1296;  // blk = h & 3
        const int fx = 2 * X + (blk & 1), fy = 2 * Y + (blk >> 1);
        const double v = a.lf[lmat_index(ncf, ncrf, sd, fx, fy, 2 * Rr, 0) + off];
        // off = (dr * 2 + w) * 324 + e  ->  fine wedge fw = 2 (blk + 4 dr) + w
        const int dr = off / 648, w = (off / 324) & 1, e = off % 324;
        sA[2 * (blk + 4 * dr) + w][e] = v;
    }
    if (threadIdx.x < 96) {
        const int fw = threadIdx.x / 6, n = threadIdx.x % 6;
        double prow[6];
        gca_prow(fw >> 1, fw & 1, n, Rr == ncrc - 1, a.exact != 0, prow);
        for (int c = 0; c < 6; ++c) sP[fw][n][c] = prow[c];
    }
    __syncthreads();
    double *out = a.lc + (int64_t)blockIdx.x * 648;
    for (int o = threadIdx.x; o < 648; o += 256) {
        const int tw = o >= 324 ? 1 : 0, e = o - tw * 324;
        const int row = e / 18, col = e % 18;
        const int di = row / 6, ii = row % 6, dj = col / 6, jj = col % 6;
        double s = 0.0;
        for (int q = 0; q < 8; ++q) {
            const int fw = kGcaFrom[tw][q];
            const double *A = sA[fw];
            double sw = 0.0;
            for (int aa = 0; aa < 6; ++aa) {
                const double pa = sP[fw][aa][ii];
                if (pa == 0.0) continue;
                double r = 0.0;
                for (int b = 0; b < 6; ++b) {
                    const double pb = sP[fw][b][jj];
                    if (pb != 0.0) r += A[(aa + 6 * di) * 18 + b + 6 * dj] * pb;
                }
                sw += pa * r;
            }
            s += sw;
        }
        out[o] = s;
    }
}

template <typename TD, typename TS> __global__ void convert_kernel(TD *d, const TS *s, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = (TD)s[i];
}

// linalg/trafo/local_basis_trafo_normal_tangential.hpp:34-76 (rows n, t1, t2)
template <typename T> __device__ void trafo_rows(const T c[3], T R[3][3])
{
    T n[3] = {c[0], c[1], c[2]};
    const T mag = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (mag > T(1e-15))
        for (int i = 0; i < 3; ++i) n[i] /= mag;
    const T ax = fabs(n[0]), ay = fabs(n[1]), az = fabs(n[2]);
    T ref[3] = {T(0), T(0), T(0)};
    if (ax <= ay && ax <= az) ref[0] = T(1);
    else if (ay <= ax && ay <= az) ref[1] = T(1);
    else ref[2] = T(1);
    T t1[3] = {ref[1] * n[2] - ref[2] * n[1], ref[2] * n[0] - ref[0] * n[2], ref[0] * n[1] - ref[1] * n[0]};
    T m1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
    if (m1 > T(1e-15))
        for (int i = 0; i < 3; ++i) t1[i] /= m1;
    const T eps = sizeof(T) == 8 ? T(1e-15) : T(1e-6);
    if (sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]) < eps) {
        t1[0] = n[1]; t1[1] = -n[0]; t1[2] = T(0);
        m1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1]);
        if (m1 > T(1e-15))
            for (int i = 0; i < 3; ++i) t1[i] /= m1;
    }
    T t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
    const T m2 = sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
    if (m2 > T(1e-15))
        for (int i = 0; i < 3; ++i) t2[i] /= m2;
    for (int i = 0; i < 3; ++i) { R[0][i] = n[i]; R[1][i] = t1[i]; R[2][i] = t2[i]; }
}

template <typename T> struct StoredArgs {
    const T *lmat;
    Comp3<const T> src;
    Comp3<T> dst;
    const T *coords, *radii;
    const double *pen;  // penalty axis (3 doubles) or nullptr
    const uint8_t *sd_cmb, *sd_surf;
    int nsd, nx, nr, bc_cmb, bc_surf, diagonal, add;
};

// boundary kind of cell layer rc (as B/B^T and the reference's slow path)
__device__ __forceinline__ int stored_cell_bc(int rc, int nr, bool has_cmb, bool has_surf, int bc_cmb, int bc_surf,
                                              int &layer)
{
    if (rc == 0 && has_cmb) { layer = 0; return bc_cmb; }
    if (rc + 1 == nr - 1 && has_surf) { layer = 1; return bc_surf; }
    layer = -1;
    return TNG_NEUMANN;
}

template <typename T> __global__ __launch_bounds__(256) void stored_apply_kernel(const StoredArgs<T> a)
{
    const int nx = a.nx, nr = a.nr, nc = nx - 1, ncr = nr - 1;
    const int64_t n = (int64_t)a.nsd * nx * nx * nr;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int l = (int)(gid % nr);
    int64_t t = gid / nr;
    const int Y = (int)(t % nx); t /= nx;
    const int X = (int)(t % nx);
    const int sd = (int)(t / nx);
    const bool has_cmb = a.sd_cmb[sd] != 0, has_surf = a.sd_surf[sd] != 0;
    const int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};
    T out[3] = {T(0), T(0), T(0)};
    T own[3] = {a.src.p[0][gid], a.src.p[1][gid], a.src.p[2][gid]};
    // boundary handling of this node's rows (same BC in all adjacent boundary cells)
    int node_bc = TNG_NEUMANN;
    if (l == 0 && has_cmb) node_bc = a.bc_cmb;
    else if (l == nr - 1 && has_surf) node_bc = a.bc_surf;
    T Ann = T(0);        // free slip: sum over wedges of n^T A_nn n
    T Bd[3] = {T(0), T(0), T(0)};  // free-slip diagonal mode: rotated diagonal entries
    T R[3][3];
    const T *cn = a.coords + 3 * (((int64_t)sd * nx + X) * nx + Y);
    T nrm[3] = {cn[0], cn[1], cn[2]};
    {
        const T mag = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int i = 0; i < 3; ++i) nrm[i] /= mag;
    }
    if (node_bc == TNG_FREESLIP && a.diagonal) trafo_rows<T>(cn, R);
    for (int dxc = -1; dxc <= 0; ++dxc) {
        const int xc = X + dxc;
        if (xc < 0 || xc >= nc) continue;
        for (int dyc = -1; dyc <= 0; ++dyc) {
            const int yc = Y + dyc;
            if (yc < 0 || yc >= nc) continue;
            for (int w = 0; w < 2; ++w) {
                const int k = corner_of[w][X - xc][Y - yc];
                if (k < 0) continue;
                for (int rc = l - 1; rc <= l; ++rc) {
                    if (rc < 0 || rc >= ncr) continue;
                    const int e = rc == l ? 0 : 1;
                    const int nn = k + 3 * e;  // this node's index in the wedge
                    int layer;
                    const int cbc = stored_cell_bc(rc, nr, has_cmb, has_surf, a.bc_cmb, a.bc_surf, layer);
                    const bool on_bd = layer == e && cbc != TNG_NEUMANN;
                    const T *A = a.lmat + lmat_index(nc, ncr, sd, xc, yc, rc, w);
                    if (a.diagonal) {
                        if (on_bd && cbc == TNG_FREESLIP) {
                            // diag(R A_nn R^T) in the rotated basis
                            for (int q = 0; q < 3; ++q) {
                                T s = T(0);
                                for (int d1 = 0; d1 < 3; ++d1)
                                    for (int d2 = 0; d2 < 3; ++d2)
                                        s += R[q][d1] * A[(nn + 6 * d1) * 18 + nn + 6 * d2] * R[q][d2];
                                Bd[q] += s;
                            }
                        } else {
                            for (int d = 0; d < 3; ++d) out[d] += A[(nn + 6 * d) * 18 + nn + 6 * d] * own[d];
                        }
                        continue;
                    }
                    if (on_bd && cbc == TNG_DIRICHLET) {
                        for (int d = 0; d < 3; ++d) out[d] += A[(nn + 6 * d) * 18 + nn + 6 * d] * own[d];
                        continue;
                    }
                    // trial values of the wedge, boundary-modified
                    T y[3] = {T(0), T(0), T(0)};
                    for (int m = 0; m < 6; ++m) {
                        const int mc = m % 3, me = m / 3;
                        const int64_t idx = ((((int64_t)sd * nx + xc + kWN[w][mc][0]) * nx + yc + kWN[w][mc][1]) *
                                             nr) + rc + me;
                        T s[3] = {a.src.p[0][idx], a.src.p[1][idx], a.src.p[2][idx]};
                        if (layer == me && cbc == TNG_DIRICHLET) continue;
                        if (layer == me && cbc == TNG_FREESLIP) {
                            const T *cm = a.coords + 3 * (((int64_t)sd * nx + xc + kWN[w][mc][0]) * nx + yc +
                                                          kWN[w][mc][1]);
                            const T mg = sqrt(cm[0] * cm[0] + cm[1] * cm[1] + cm[2] * cm[2]);
                            const T n0 = cm[0] / mg, n1 = cm[1] / mg, n2 = cm[2] / mg;
                            const T dd = n0 * s[0] + n1 * s[1] + n2 * s[2];
                            s[0] -= dd * n0; s[1] -= dd * n1; s[2] -= dd * n2;
                        }
                        for (int d = 0; d < 3; ++d) {
                            const T *row = A + (nn + 6 * d) * 18;
                            y[d] += row[m] * s[0] + row[m + 6] * s[1] + row[m + 12] * s[2];
                        }
                    }
                    if (on_bd) {  // free slip: tangential rows, normal-normal entry
                        const T dv = nrm[0] * y[0] + nrm[1] * y[1] + nrm[2] * y[2];
                        for (int d = 0; d < 3; ++d) y[d] -= dv * nrm[d];
                        T s = T(0);
                        for (int d1 = 0; d1 < 3; ++d1)
                            for (int d2 = 0; d2 < 3; ++d2) s += nrm[d1] * A[(nn + 6 * d1) * 18 + nn + 6 * d2] * nrm[d2];
                        Ann += s;
                    }
                    for (int d = 0; d < 3; ++d) out[d] += y[d];
                }
            }
        }
    }
    if (node_bc == TNG_FREESLIP) {
        if (a.diagonal) {
            // R^T diag(B) R own
            T rs[3];
            for (int q = 0; q < 3; ++q) rs[q] = Bd[q] * (R[q][0] * own[0] + R[q][1] * own[1] + R[q][2] * own[2]);
            for (int d = 0; d < 3; ++d) out[d] += R[0][d] * rs[0] + R[1][d] * rs[1] + R[2][d] * rs[2];
        } else {
            const T un = Ann * (nrm[0] * own[0] + nrm[1] * own[1] + nrm[2] * own[2]);
            for (int d = 0; d < 3; ++d) out[d] += un * nrm[d];
        }
    }
    if (a.pen) {
        const T rad = a.radii[(int64_t)sd * nr + l];
        const T x0 = cn[0] * rad, x1 = cn[1] * rad, x2 = cn[2] * rad;
        const T b0 = (T)a.pen[0], b1 = (T)a.pen[1], b2 = (T)a.pen[2];
        out[0] += b1 * x2 - b2 * x1;
        out[1] += b2 * x0 - b0 * x2;
        out[2] += b0 * x1 - b1 * x0;
    }
    for (int d = 0; d < 3; ++d) {
        T *o = a.dst.p[d] + gid;
        *o = a.add ? *o + out[d] : out[d];
    }
}

// Adaptive GCA element selection (linalg/solvers/gca/gca_elements_collector.hpp
// :20-96): per fine wedge, the gradient of k in reference coordinates at the
// centroid (1/3, 1/3, 0); a norm above 10 flags the coarsest-level ancestor
// cell (cell index >> level_range in x, y and r).  With s_a = k_bottom + k_top
// and d_a = k_top - k_bottom at the three lateral corners the gradient is
// (s_1 - s_0, s_2 - s_0, (d_0 + d_1 + d_2) / 3) / 2.
__global__ __launch_bounds__(256) void gca_collect_kernel(const double *k, int nsd, int nx, int nr, int level_range,
                                                          uint8_t *flags, int ncx0, int ncr0)
{
    const int nc = nx - 1, ncr = nr - 1;
    const int64_t n = (int64_t)nsd * nc * nc * ncr;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int rc = (int)(gid % ncr);
    int64_t t = gid / ncr;
    const int yc = (int)(t % nc); t /= nc;
    const int xc = (int)(t % nc);
    const int sd = (int)(t / nc);
    bool hit = false;
    for (int w = 0; w < 2; ++w) {
        double sa[3], da = 0.0;
        for (int c = 0; c < 3; ++c) {
            const int64_t i = (((int64_t)sd * nx + xc + kWN[w][c][0]) * nx + yc + kWN[w][c][1]) * nr + rc;
            const double kb = k[i], kt = k[i + 1];
            sa[c] = kb + kt;
            da += kt - kb;
        }
        const double g0 = 0.5 * (sa[1] - sa[0]), g1 = 0.5 * (sa[2] - sa[0]), g2 = da * (1.0 / 6.0);
        if (sqrt(g0 * g0 + g1 * g1 + g2 * g2) > 10.0) hit = true;
    }
    if (hit)
        flags[(((int64_t)sd * ncx0 + (xc >> level_range)) * ncx0 + (yc >> level_range)) * ncr0 + (rc >> level_range)] = 1;
}

// Selective storage (local_matrix_storage.hpp:28-243, OperatorStoredMatrixMode::
// Selective): wedges of cells whose coarsest ancestor is not flagged take the
// on-the-fly local matrix (epsilon_divdiv_kerngen.hpp:1047-1060) instead of the
// Galerkin one.
__global__ __launch_bounds__(256) void gca_select_kernel(double *lmat, const double *local, const uint8_t *flags,
                                                         int nsd, int nc, int ncr, int level_range, int ncx0, int ncr0)
{
    const int64_t ncell = (int64_t)nsd * nc * nc * ncr;
    const int64_t n = ncell * 2 * 324;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cell = e / (2 * 324);
        const int rc = (int)(cell % ncr);
        int64_t t = cell / ncr;
        const int yc = (int)(t % nc); t /= nc;
        const int xc = (int)(t % nc);
        const int sd = (int)(t / nc);
        const uint8_t f =
            flags[(((int64_t)sd * ncx0 + (xc >> level_range)) * ncx0 + (yc >> level_range)) * ncr0 + (rc >> level_range)];
        if (!f) lmat[e] = local[e];
    }
}

}  // namespace

void epsdivdiv_local_matrices(const tng_operator &op, DevBuf &out, hipStream_t s)
{
    const tng_domain &m = *op.mesh;
    TNG_REQUIRE(op.k && op.k->dtype == TNG_F64, "local matrices are assembled from an FP64 coefficient");
    const int64_t nw = (int64_t)m.nsd * (m.nx - 1) * (m.nx - 1) * (m.nr - 1) * 2;
    const size_t bytes = (size_t)nw * 324 * sizeof(double);
    if (out.bytes != bytes) out.alloc(bytes);
    LocalArgs a{m.d_coords.as<double>(), m.d_radii.as<double>(), (const double *)op.k->comp(0), out.as<double>(),
                m.nsd, m.nx, m.nr};
    hipLaunchKernelGGL(local_matrix_kernel, dim3((unsigned)((nw + 127) / 128)), dim3(128), 0, s, a);
    TNG_HIP_CHECK(hipGetLastError());
}

size_t stored_matrix_count(const tng_domain &m) { return (size_t)m.nsd * (m.nx - 1) * (m.nx - 1) * (m.nr - 1) * 2; }

void gca_assemble(const tng_operator &fine, tng_operator &coarse, hipStream_t s, bool exact)
{
    const tng_domain &f = *fine.mesh, &c = *coarse.mesh;
    TNG_REQUIRE(f.nsd == c.nsd && f.nx - 1 == 2 * (c.nx - 1) && f.nr - 1 == 2 * (c.nr - 1),
                "GCA: operators must live on meshes one refinement apart");
    TNG_REQUIRE(c.nr >= 3, "GCA: the coarse mesh needs at least two radial layers");
    const double *lf = nullptr;
    if (fine.stored) {
        TNG_REQUIRE(fine.lmat_dtype == TNG_F64, "GCA: the fine stored matrices must be FP64");
        lf = fine.d_lmat.as<double>();
    } else {
        TNG_REQUIRE(fine.k && fine.k->dtype == TNG_F64, "GCA: the fine operator needs an FP64 coefficient");
    }
    const size_t bytes = stored_matrix_count(c) * 324 * sizeof(double);
    if (coarse.d_lmat.bytes != bytes) coarse.d_lmat.alloc(bytes);
    const int64_t ncell = (int64_t)c.nsd * (c.nx - 1) * (c.nx - 1) * (c.nr - 1);
    GcaArgs a{lf, f.d_coords.as<double>(), f.d_radii.as<double>(), fine.k ? (const double *)fine.k->comp(0) : nullptr,
              coarse.d_lmat.as<double>(), f.nsd, f.nx, f.nr, c.nx, c.nr, exact ? 1 : 0, ncell};
    if (lf)
        hipLaunchKernelGGL(gca_stored_kernel, dim3((unsigned)ncell), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(gca_onthefly_kernel, dim3((unsigned)((ncell + kGcaCells - 1) / kGcaCells)), dim3(256), 0,
                           s, a);
    TNG_HIP_CHECK(hipGetLastError());
    coarse.stored = true;
    coarse.lmat_dtype = TNG_F64;
}

void gca_collect(const tng_vector &k_fine, int level_range, const tng_domain &coarsest, DevBuf &flags, hipStream_t s)
{
    const tng_domain &f = *k_fine.mesh;
    TNG_REQUIRE(k_fine.dtype == TNG_F64 && k_fine.ncomp == 1, "GCA collector: FP64 scalar coefficient");
    TNG_REQUIRE(level_range >= 0 && level_range < 16 && f.nsd == coarsest.nsd &&
                    (f.nx - 1) == ((coarsest.nx - 1) << level_range) && (f.nr - 1) == ((coarsest.nr - 1) << level_range),
                "GCA collector: the coefficient's mesh must be level_range refinements above the coarsest mesh");
    const int ncx0 = coarsest.nx - 1, ncr0 = coarsest.nr - 1;
    const size_t bytes = (size_t)coarsest.nsd * ncx0 * ncx0 * ncr0;
    if (flags.bytes != bytes) flags.alloc(bytes);
    TNG_HIP_CHECK(hipMemsetAsync(flags.p, 0, bytes, s));
    const int64_t n = (int64_t)f.nsd * (f.nx - 1) * (f.nx - 1) * (f.nr - 1);
    hipLaunchKernelGGL(gca_collect_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       (const double *)k_fine.comp(0), f.nsd, f.nx, f.nr, level_range, flags.as<uint8_t>(), ncx0, ncr0);
    TNG_HIP_CHECK(hipGetLastError());
}

void gca_assemble_selective(const tng_operator &fine, tng_operator &coarse, int level_range,
                            const tng_domain &coarsest, const DevBuf &flags, hipStream_t s, bool exact)
{
    const tng_domain &c = *coarse.mesh;
    const int ncx0 = coarsest.nx - 1, ncr0 = coarsest.nr - 1;
    TNG_REQUIRE(flags.bytes == (size_t)coarsest.nsd * ncx0 * ncx0 * ncr0, "GCA: flags do not match the coarsest mesh");
    TNG_REQUIRE(level_range >= 0 && c.nsd == coarsest.nsd && (c.nx - 1) == (ncx0 << level_range) &&
                    (c.nr - 1) == (ncr0 << level_range),
                "GCA: the coarse operator's mesh must be level_range refinements above the coarsest mesh");
    gca_assemble(fine, coarse, s, exact);
    DevBuf local;
    epsdivdiv_local_matrices(coarse, local, s);
    const int64_t n = (int64_t)stored_matrix_count(c) * 324;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(gca_select_kernel, dim3(g), dim3(256), 0, s, coarse.d_lmat.as<double>(),
                       (const double *)local.p, flags.as<uint8_t>(), c.nsd, c.nx - 1, c.nr - 1, level_range, ncx0,
                       ncr0);
    TNG_HIP_CHECK(hipGetLastError());
    TNG_HIP_CHECK(hipStreamSynchronize(s));  // `local` is freed on return
}

void copy_stored_matrices(tng_operator &dst, const tng_operator &src, tng_dtype dtype, hipStream_t s)
{
    TNG_REQUIRE(src.stored && src.lmat_dtype == TNG_F64, "copy: the source operator has no FP64 stored matrices");
    TNG_REQUIRE(dst.mesh->nsd == src.mesh->nsd && dst.mesh->nx == src.mesh->nx && dst.mesh->nr == src.mesh->nr,
                "copy: operators on different meshes");
    const int64_t n = (int64_t)stored_matrix_count(*src.mesh) * 324;
    const size_t es = dtype == TNG_F64 ? 8 : 4;
    if (dst.d_lmat.bytes != (size_t)n * es) dst.d_lmat.alloc((size_t)n * es);
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    if (dtype == TNG_F64)
        hipLaunchKernelGGL((convert_kernel<double, double>), dim3(g), dim3(256), 0, s, dst.d_lmat.as<double>(),
                           src.d_lmat.as<double>(), n);
    else
        hipLaunchKernelGGL((convert_kernel<float, double>), dim3(g), dim3(256), 0, s, dst.d_lmat.as<float>(),
                           src.d_lmat.as<double>(), n);
    TNG_HIP_CHECK(hipGetLastError());
    dst.stored = true;
    dst.lmat_dtype = dtype;
}

template <typename T>
void stored_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, const double *pen, bool add, hipStream_t s)
{
    const tng_domain &m = *op.mesh;
    TNG_REQUIRE(op.stored && op.lmat_dtype == src.dtype, "stored apply: matrices missing or of another precision");
    TNG_REQUIRE(m.nr >= 3, "stored apply needs at least two radial layers");
    StoredArgs<T> a{};
    a.lmat = op.d_lmat.as<T>();
    a.src = {{(const T *)src.comp(0), (const T *)src.comp(1), (const T *)src.comp(2)}};
    a.dst = comps<T>(dst);
    a.coords = src.dtype == TNG_F64 ? (const T *)m.d_coords.p : (const T *)m.d_coords_f.p;
    a.radii = src.dtype == TNG_F64 ? (const T *)m.d_radii.p : (const T *)m.d_radii_f.p;
    a.pen = pen;
    a.sd_cmb = m.d_sd_cmb.as<uint8_t>();
    a.sd_surf = m.d_sd_surf.as<uint8_t>();
    a.nsd = m.nsd; a.nx = m.nx; a.nr = m.nr;
    a.bc_cmb = op.bc_cmb; a.bc_surf = op.bc_surf; a.diagonal = op.diagonal ? 1 : 0; a.add = add ? 1 : 0;
    hipLaunchKernelGGL(stored_apply_kernel<T>, dim3((unsigned)((m.nalloc + 255) / 256)), dim3(256), 0, s, a);
    TNG_HIP_CHECK(hipGetLastError());
}

template void stored_apply<double>(tng_operator &, const tng_vector &, tng_vector &, const double *, bool, hipStream_t);
template void stored_apply<float>(tng_operator &, const tng_vector &, tng_vector &, const double *, bool, hipStream_t);

}  // namespace tng
