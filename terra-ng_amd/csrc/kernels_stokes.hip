// Divergence B (fine velocity -> coarse pressure), gradient B^T (coarse
// pressure -> fine velocity), constant-weight prolongation / restriction and
// the lumped k-weighted pressure mass.
//
// All kernels map consecutive threads to consecutive radial levels (the
// fastest index of the node layout) and gather per output node, so every
// output entry is written once, without atomics, deterministically.
//
// Divergence / gradient (reference: divergence_kerngen.hpp:373-518,
// gradient_kerngen.hpp:339-496).  One-point rule at the fine wedge centroid:
//   B:   coarse node i of coarse wedge w  +=  -|det J| (div u) N^c_i
//   B^T: fine node j of the fine wedge     +=  -|det J| g_j p(centroid)
// where fine wedge w of every fine cell belongs to coarse wedge w of the
// coarse cell (x/2, y/2, r/2) and N^c_i are the coarse shape functions of the
// fine wedge's position in its coarse triangle (fine_lateral_wedge_idx,
// kernel_helpers.hpp:600-608; shape_coarse, integrands.hpp:266-393).  At a
// Dirichlet layer the boundary velocity dofs are removed, at a free-slip
// layer the velocity is projected onto the tangent plane (trial side for B,
// test side for B^T).
#include "tng_internal.hpp"

namespace tng {

namespace {

// Logical block index with contiguous ranges per XCD: the hardware deals
// blocks round-robin to the 8 XCDs, each with its own L2, and neighbouring
// node columns (which read the same cells) sit in neighbouring blocks.
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t nb = gridDim.x, x = blockIdx.x & 7, i = blockIdx.x >> 3, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + i;
}

__host__ __device__ __forceinline__ int fine_lat_idx(int x, int y, int w)
{
    // [w][y%2][x%2]
    return w == 0 ? ((y & 1) ? ((x & 1) ? 3 : 2) : ((x & 1) ? 1 : 0))
                  : ((y & 1) ? ((x & 1) ? 0 : 1) : ((x & 1) ? 2 : 3));
}

// Coarse shape function i (0..5) at the fine centroid (1/3, 1/3, 0).
template <typename T> __device__ __forceinline__ T shape_coarse_c(int i, int lat, int rad)
{
    const T third = T(1.0 / 3.0);
    T sl;
    switch (i % 3) {
    case 0: sl = lat == 0 ? T(-0.5) * third - T(0.5) * third + T(1) : (lat == 3 ? T(0.5) * third + T(0.5) * third
                                                                                 : T(-0.5) * third - T(0.5) * third + T(0.5)); break;
    case 1: sl = lat == 1 ? T(0.5) * third + T(0.5) : (lat == 3 ? T(0.5) - T(0.5) * third : T(0.5) * third); break;
    default: sl = lat == 2 ? T(0.5) * third + T(0.5) : (lat == 3 ? T(0.5) - T(0.5) * third : T(0.5) * third); break;
    }
    const T sr = (i / 3 == 0) ? (rad == 0 ? T(0.75) : T(0.25)) : (rad == 0 ? T(0.25) : T(0.75));
    return sl * sr;
}

// wedge corner (dx, dy) per wedge and local lateral index
__constant__ int kWC[2][3][2] = {{{0, 0}, {1, 0}, {0, 1}}, {{1, 1}, {0, 1}, {1, 0}}};

template <typename T> using T2v = T __attribute__((ext_vector_type(2)));

template <typename T> struct DivArgs {
    const T *u0, *u1, *u2;       // fine velocity
    const T *coords, *radii;     // fine mesh
    const T *latgeom;            // fine mesh (nsd, nc, nc, 2, 10)
    const uint8_t *sd_cmb, *sd_surf;
    T *partial;                  // (nsd, ncc, ncc, nrc-1, 2, 6) coarse-wedge partials
    int nsd, nx, nr;             // fine
    int nxc, nrc;                // coarse
    int bc_cmb, bc_surf;
    Index4 cidx;                 // coarse cells (nsd, nxc-1, nxc-1, nrc-1)
    Index4 fidx;                 // fine cells (nsd, nx-1, nx-1, nr-1)
};

// Boundary kind of fine cell layer r (boundary_setup of the reference's
// eval_cell): a cell whose bottom lies on the CMB is a CMB cell, otherwise a
// cell whose top lies on the surface is a surface cell.  Returns the BC of the
// cell (TNG_NEUMANN for interior cells) and the wedge node layer (0 bottom,
// 1 top) that the BC acts on.
__device__ __forceinline__ int cell_bc(int r, int nr, bool has_cmb, bool has_surf, int bc_cmb, int bc_surf, int &layer)
{
    if (r == 0 && has_cmb) {
        layer = 0;
        return bc_cmb;
    }
    if (r + 1 == nr - 1 && has_surf) {
        layer = 1;
        return bc_surf;
    }
    layer = -1;
    return TNG_NEUMANN;
}

// Trial-side boundary modification of a fine velocity value of column lcol:
// Dirichlet removes it, free-slip removes its normal component.
template <typename T> __device__ __forceinline__ void bc_project(const T *coords, int64_t lcol, int bc, T u[3])
{
    if (bc == TNG_DIRICHLET) {
        u[0] = u[1] = u[2] = T(0);
    } else if (bc == TNG_FREESLIP) {
        const T *n = coords + 3 * lcol;
        const T d = n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
        u[0] -= d * n[0]; u[1] -= d * n[1]; u[2] -= d * n[2];
    }
}

// Phase 1 of B, per fine cell: cdiv_w = -|det J| (div u) of both wedges.
// One thread per fine cell, r fastest.  UNI (radial cells a multiple of 64):
// every wave covers cells of one lateral column, so the wedge geometry is
// wave-uniform and read with scalar loads.
template <typename T, bool UNI>
__global__ __launch_bounds__(256) void div_cell_kernel(const DivArgs<T> a, T *cdiv)
{
    const int nx = a.nx, nr = a.nr, nc = nx - 1, ncr = nr - 1;
    const uint32_t n = (uint32_t)a.nsd * nc * nc * ncr;
    const uint32_t gid = xcd_block() * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    int sd, x, y, r;
    a.fidx.decode(gid, sd, x, y, r);
    const bool has_cmb = a.sd_cmb[sd] != 0, has_surf = a.sd_surf[sd] != 0;
    const T r0 = a.radii[(int64_t)sd * nr + r], r1 = a.radii[(int64_t)sd * nr + r + 1];
    const T rmid = T(0.5) * (r0 + r1), h = T(0.5) * (r1 - r0);
    const T sfac = T(1) / rmid, tz = T(1) / (T(6) * h);
    const T vol = rmid * rmid * h;
    int blayer;
    const int cbc = cell_bc(r, nr, has_cmb, has_surf, a.bc_cmb, a.bc_surf, blayer);
    // the 8 nodes of the cell: lateral corners (dx, dy) at the two levels
    T u[2][2][2][3];
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const int64_t lcol = ((int64_t)sd * nx + x + dx) * nx + y + dy;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int64_t i = lcol * nr + r + e;
                u[dx][dy][e][0] = a.u0[i];
                u[dx][dy][e][1] = a.u1[i];
                u[dx][dy][e][2] = a.u2[i];
            }
        }
    if (cbc != TNG_NEUMANN) {
#pragma unroll
        for (int dx = 0; dx < 2; ++dx)
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int e = 0; e < 2; ++e)
                    if (e == blayer) bc_project(a.coords, ((int64_t)sd * nx + x + dx) * nx + y + dy, cbc, u[dx][dy][e]);
    }
    const int64_t gb = ((((int64_t)sd * nc + x) * nc + y) * 2) * 10;
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        T g[10];
        if constexpr (UNI) {
            using CT = const T __attribute__((address_space(4)));
            const int64_t gu = ((int64_t)__builtin_amdgcn_readfirstlane((int)(gb >> 32)) << 32) |
                               (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)gb);
            const CT *gw = (const CT *)a.latgeom + gu + w * 10;
#pragma unroll
            for (int v = 0; v < 10; ++v) g[v] = gw[v];
        } else {
#pragma unroll
            for (int v = 0; v < 10; ++v) g[v] = a.latgeom[gb + w * 10 + v];
        }
        // div u = sum_j g_j . u_j over the 6 wedge nodes
        T div = T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int dx = w == 0 ? (k == 1 ? 1 : 0) : (k == 1 ? 0 : 1);
            const int dy = w == 0 ? (k == 2 ? 1 : 0) : (k == 2 ? 0 : 1);
            const T l0 = k == 0 ? T(-0.5) : (k == 1 ? T(0.5) : T(0));
            const T l1 = k == 0 ? T(-0.5) : (k == 1 ? T(0) : T(0.5));
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const T dz = e == 0 ? -tz : tz;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    div += (sfac * (l0 * g[d] + l1 * g[3 + d]) + dz * g[6 + d]) * u[dx][dy][e][d];
            }
        }
        cdiv[(size_t)gid * 2 + w] = -(g[9] * vol) * div;
    }
}

// Phase 2 of B, per coarse cell and coarse wedge: the 6 nodal sums over the
// 8 fine wedges it owns, cdiv N^c_i.  One thread per coarse cell, r fastest.
template <typename T>
__global__ __launch_bounds__(256) void div_coarse_kernel(const DivArgs<T> a, const T *cdiv)
{
    const int ncc = a.nxc - 1, nrcc = a.nrc - 1;
    const uint32_t n = (uint32_t)a.nsd * ncc * ncc * nrcc;
    const uint32_t gid = xcd_block() * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    int sd, xc, yc, rc;
    a.cidx.decode(gid, sd, xc, yc, rc);
    const int nc = a.nx - 1, ncr = a.nr - 1;
    T2v<T> f[2][2][2];
#pragma unroll
    for (int fx = 0; fx < 2; ++fx)
#pragma unroll
        for (int fy = 0; fy < 2; ++fy)
#pragma unroll
            for (int fr = 0; fr < 2; ++fr)
                f[fx][fy][fr] = *reinterpret_cast<const T2v<T> *>(
                    cdiv + ((((size_t)sd * nc + 2 * xc + fx) * nc + 2 * yc + fy) * ncr + 2 * rc + fr) * 2);
    T acc[2][6];
#pragma unroll
    for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int i = 0; i < 6; ++i) acc[w][i] = T(0);
#pragma unroll
    for (int fx = 0; fx < 2; ++fx)
#pragma unroll
        for (int fy = 0; fy < 2; ++fy)
#pragma unroll
            for (int fr = 0; fr < 2; ++fr)
#pragma unroll
                for (int w = 0; w < 2; ++w) {
                    const int lat = fine_lat_idx(fx, fy, w);
#pragma unroll
                    for (int i = 0; i < 6; ++i) acc[w][i] += f[fx][fy][fr][w] * shape_coarse_c<T>(i, lat, fr);
                }
    T *o = a.partial + (size_t)gid * 12;
#pragma unroll
    for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int i = 0; i < 6; ++i) o[w * 6 + i] = acc[w][i];
}

// Phase 2 of B: coarse node = sum of the coarse-wedge partials it belongs to.
template <typename T>
__global__ __launch_bounds__(256) void wedge_gather_kernel(const T *partial, T *dst, int add, int nsd, int nx, int nr,
                                                           const Index4 idx)
{
    const uint32_t n = (uint32_t)nsd * nx * nx * nr;
    const uint32_t gid = xcd_block() * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    int sd, X, Y, l;
    idx.decode(gid, sd, X, Y, l);
    const int nc = nx - 1, ncr = nr - 1;
    // lateral corner index of (dx, dy) in wedge w (-1: not a corner)
    const int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};
    T s = T(0);
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
                    const int node = k + (rc == l ? 0 : 3);
                    s += partial[((((int64_t)sd * nc + xc) * nc + yc) * ncr + rc) * 12 + w * 6 + node];
                }
            }
        }
    }
    dst[gid] = add ? dst[gid] + s : s;
}

template <typename T> struct GradArgs {
    const T *p;                   // coarse pressure
    T *d0, *d1, *d2;              // fine velocity
    const T *coords, *radii, *latgeom;
    const uint8_t *sd_cmb, *sd_surf;
    int nsd, nx, nr, nxc, nrc;
    int bc_cmb, bc_surf, add;
    Index4 fcell, fnode;          // fine cells (nsd, nx-1, nx-1, nr-1) and fine nodes
};

// Phase 1 of B^T: per fine wedge, f = -|det J| p(centroid), with p
// interpolated from coarse wedge w of the coarse cell.  One thread per fine
// cell (both wedges), r fastest.
template <typename T> __global__ __launch_bounds__(256) void grad_weight_kernel(const GradArgs<T> a, T *fw)
{
    const int nx = a.nx, nr = a.nr, nc = nx - 1, ncr = nr - 1;
    const uint32_t n = (uint32_t)a.nsd * nc * nc * ncr;
    const uint32_t gid = xcd_block() * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    int sd, xc, yc, rc;
    a.fcell.decode(gid, sd, xc, yc, rc);
    const T r0 = a.radii[(int64_t)sd * nr + rc], r1 = a.radii[(int64_t)sd * nr + rc + 1];
    const T rmid = T(0.5) * (r0 + r1), h = T(0.5) * (r1 - r0);
    const int cx = xc / 2, cy = yc / 2, cr = rc / 2;
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const T *g = a.latgeom + ((((int64_t)sd * nc + xc) * nc + yc) * 2 + w) * 10;
        const int lat = fine_lat_idx(xc, yc, w);
        T pi = T(0);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const int jx = cx + kWC[w][j % 3][0], jy = cy + kWC[w][j % 3][1], jr = cr + j / 3;
            pi += shape_coarse_c<T>(j, lat, rc & 1) * a.p[(((int64_t)sd * a.nxc + jx) * a.nxc + jy) * a.nrc + jr];
        }
        fw[(size_t)gid * 2 + w] = -(g[9] * rmid * rmid * h) * pi;
    }
}

// Phase 2 of B^T per fine node: sum over the <= 12 fine wedges containing the
// node of f_w g_node, with g_node = J^{-T} grad N at the centroid.
// STAGE (columns of >= 52 nodes): the block's threads cover at most
// kGradCols node columns; the wedge geometry of the 4 cells around each is
// staged in LDS once and read there by every node of the column (broadcast),
// instead of 72 vector loads per node.
constexpr int kGradCols = 6;
template <typename T, bool STAGE>
__global__ __launch_bounds__(256) void gradient_kernel(const GradArgs<T> a, const T *fw)
{
    __shared__ T sg[STAGE ? kGradCols : 1][4][2][9];
    const int nx = a.nx, nr = a.nr, nc = nx - 1, ncr = nr - 1;
    const uint32_t n = (uint32_t)a.nsd * nx * nx * nr;
    const uint32_t b0 = xcd_block() * blockDim.x;
    const uint32_t gid = b0 + threadIdx.x;
    uint32_t col0 = 0;
    if constexpr (STAGE) {
        col0 = a.fnode.r.div(b0);
        const uint32_t ncol = (uint32_t)a.nsd * nx * nx;
        for (int e = threadIdx.x; e < kGradCols * 72; e += blockDim.x) {
            const int slot = e / 72, rest = e - slot * 72, cell = rest / 18, wv = rest - cell * 18;
            const uint32_t col = col0 + slot;
            T val = T(0);
            if (col < ncol) {
                const uint32_t t2 = a.fnode.x.div(col), s3 = a.fnode.x.div(t2);
                const int cy = (int)(col - t2 * a.fnode.x.d), cx = (int)(t2 - s3 * a.fnode.x.d);
                const int xc = cx + (cell >> 1) - 1, yc = cy + (cell & 1) - 1;
                if (xc >= 0 && xc < nc && yc >= 0 && yc < nc)
                    val = a.latgeom[(((int64_t)s3 * nc + xc) * nc + yc) * 20 + (wv / 9) * 10 + wv % 9];
            }
            sg[slot][cell][wv / 9][wv % 9] = val;
        }
        __syncthreads();
    }
    if (gid >= n) return;
    int sd, X, Y, l;
    a.fnode.decode(gid, sd, X, Y, l);
    const int slot = STAGE ? (int)(a.fnode.r.div(gid) - col0) : 0;
    const int64_t lcol = ((int64_t)sd * nx + X) * nx + Y;
    const int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};
    const bool has_cmb = a.sd_cmb[sd] != 0, has_surf = a.sd_surf[sd] != 0;
    // radial factors of the (at most) two cells below / above the node
    T sfac[2], dzf[2];
    int cbc[2];
    bool onb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int rc = l - 1 + q;
        const int rcc = min(max(rc, 0), ncr - 1);
        const T r0 = a.radii[(int64_t)sd * nr + rcc], r1 = a.radii[(int64_t)sd * nr + rcc + 1];
        const T rmid = T(0.5) * (r0 + r1), h = T(0.5) * (r1 - r0);
        sfac[q] = T(1) / rmid;
        dzf[q] = (q == 1 ? T(-1) : T(1)) / (T(6) * h);
        int blayer;
        cbc[q] = cell_bc(rcc, nr, has_cmb, has_surf, a.bc_cmb, a.bc_surf, blayer);
        onb[q] = blayer == (q == 1 ? 0 : 1);
    }
    // Every weight the node can use, requested before any of it is needed
    // (clamped indices: always in bounds; the branches below only skip
    // arithmetic), so the loads are in flight together.
    typedef T T2 __attribute__((ext_vector_type(2)));
    T2 fq[2][2][2];  // [dxc + 1][dyc + 1][q] -> (wedge 0, wedge 1)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int xcc = min(max(X - 1 + i, 0), nc - 1), ycc = min(max(Y - 1 + j, 0), nc - 1);
            const int64_t cbq = (((int64_t)sd * nc + xcc) * nc + ycc) * ncr;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int rcc = min(max(l - 1 + q, 0), ncr - 1);
                fq[i][j][q] = *reinterpret_cast<const T2 *>(fw + (cbq + rcc) * 2);
            }
        }
    // o: interior contributions; ob: contributions of boundary cells in which
    // this node lies on the boundary layer (free-slip: projected at the end)
    T o[3] = {T(0), T(0), T(0)}, ob[3] = {T(0), T(0), T(0)};
    bool any_fs = false;
#pragma unroll
    for (int dxc = -1; dxc <= 0; ++dxc) {
        const int xc = X + dxc;
        if (xc < 0 || xc >= nc) continue;
#pragma unroll
        for (int dyc = -1; dyc <= 0; ++dyc) {
            const int yc = Y + dyc;
            if (yc < 0 || yc >= nc) continue;
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int k = corner_of[w][-dxc][-dyc];
                if (k < 0) continue;
                T g[9];
                if constexpr (STAGE) {
#pragma unroll
                    for (int v = 0; v < 9; ++v) g[v] = sg[slot][(dxc + 1) * 2 + (dyc + 1)][w][v];
                } else {
                    const T *gp = a.latgeom + ((((int64_t)sd * nc + xc) * nc + yc) * 2 + w) * 10;
#pragma unroll
                    for (int v = 0; v < 9; ++v) g[v] = gp[v];
                }
                const T l0 = k == 0 ? T(-0.5) : (k == 1 ? T(0.5) : T(0));
                const T l1 = k == 0 ? T(-0.5) : (k == 1 ? T(0) : T(0.5));
                const T gl[3] = {l0 * g[0] + l1 * g[3], l0 * g[1] + l1 * g[4], l0 * g[2] + l1 * g[5]};
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int rc = l - 1 + q;
                    if (rc < 0 || rc >= ncr) continue;
                    if (onb[q] && cbc[q] == TNG_DIRICHLET) continue;
                    const T f = fq[dxc + 1][dyc + 1][q][w];
                    const bool to_b = onb[q] && cbc[q] == TNG_FREESLIP;
                    any_fs = any_fs || to_b;
                    T *acc = to_b ? ob : o;
                    for (int d = 0; d < 3; ++d) acc[d] += f * (sfac[q] * gl[d] + dzf[q] * g[6 + d]);
                }
            }
        }
    }
    if (any_fs) {
        bc_project(a.coords, lcol, TNG_FREESLIP, ob);
        o[0] += ob[0]; o[1] += ob[1]; o[2] += ob[2];
    }
    if (a.add) {
        o[0] += a.d0[gid]; o[1] += a.d1[gid]; o[2] += a.d2[gid];
    }
    a.d0[gid] = o[0]; a.d1[gid] = o[1]; a.d2[gid] = o[2];
}

// Constant prolongation (fe/wedge/shell/grid_transfer_constant.hpp:9-90):
// P1 interpolation per fine node from the coarse nodes of its position.
template <typename T>
__global__ __launch_bounds__(256) void prolongate_kernel(const Comp3<T> src, const Comp3<T> dst, int ncomp, int add,
                                                         int nsd, int nxf, int nrf, int nxc, int nrc, const Index4 idx)
{
    const uint32_t n = (uint32_t)nsd * nxf * nxf * nrf;
    const uint32_t gid = xcd_block() * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    int sd, xf, yf, rf;
    idx.decode(gid, sd, xf, yf, rf);
    // lateral parents (up to 2) with weights
    int px[2], py[2], np = 0;
    T pw;
    if (!(xf & 1) && !(yf & 1)) { px[0] = xf / 2; py[0] = yf / 2; np = 1; pw = T(1); }
    else if ((xf & 1) && !(yf & 1)) { px[0] = xf / 2; py[0] = yf / 2; px[1] = xf / 2 + 1; py[1] = yf / 2; np = 2; pw = T(0.5); }
    else if (!(xf & 1) && (yf & 1)) { px[0] = xf / 2; py[0] = yf / 2; px[1] = xf / 2; py[1] = yf / 2 + 1; np = 2; pw = T(0.5); }
    else { px[0] = (xf + 1) / 2; py[0] = (yf - 1) / 2; px[1] = (xf - 1) / 2; py[1] = (yf + 1) / 2; np = 2; pw = T(0.5); }
    int pr[2], nrp;
    T rw;
    if (!(rf & 1)) { pr[0] = rf / 2; nrp = 1; rw = T(1); }
    else { pr[0] = rf / 2; pr[1] = rf / 2 + 1; nrp = 2; rw = T(0.5); }
    for (int c = 0; c < ncomp; ++c) {
        T s = T(0);
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < nrp; ++j)
                s += (pw * rw) * src.p[c][(((int64_t)sd * nxc + px[i]) * nxc + py[i]) * nrc + pr[j]];
        T *o = dst.p[c] + gid;
        *o = add ? *o + s : s;
    }
}

// Constant restriction (restriction_constant.hpp:133-250): owned-masked
// transpose stencil; the additive exchange runs afterwards.
template <typename T>
__global__ __launch_bounds__(256) void restrict_kernel(const Comp3<T> src, const Comp3<T> dst, int ncomp, int add,
                                                       const uint8_t *omask_f, int nsd, int nxf, int nrf, int nxc,
                                                       int nrc, const Index4 idx)
{
    const uint32_t n = (uint32_t)nsd * nxc * nxc * nrc;
    const uint32_t gid = xcd_block() * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    int sd, xc, yc, rc;
    idx.decode(gid, sd, xc, yc, rc);
    const int st[7][2] = {{-1, 0}, {-1, 1}, {0, 1}, {1, 0}, {1, -1}, {0, -1}, {0, 0}};
    T s[3] = {T(0), T(0), T(0)};
    for (int dr = -1; dr <= 1; ++dr) {
        const int rf = 2 * rc + dr;
        if (rf < 0 || rf >= nrf) continue;
        const T rw = dr == 0 ? T(1) : T(0.5);
        for (int k = 0; k < 7; ++k) {
            const int xf = 2 * xc + st[k][0], yf = 2 * yc + st[k][1];
            if (xf < 0 || xf >= nxf || yf < 0 || yf >= nxf) continue;
            const int64_t fi = (((int64_t)sd * nxf + xf) * nxf + yf) * nrf + rf;
            if (omask_f[fi] != OWNED) continue;
            const T w = rw * (k == 6 ? T(1) : T(0.5));
            for (int c = 0; c < ncomp; ++c) s[c] += w * src.p[c][fi];
        }
    }
    for (int c = 0; c < ncomp; ++c) {
        T *o = dst.p[c] + gid;
        *o = add ? *o + s[c] : s[c];
    }
}

// Lumped k-weighted mass (kmass.hpp:96-180, 6-point rule, row sums), per node.
template <typename T>
__global__ __launch_bounds__(256) void kmass_lumped_kernel(const T *k, const T *src, T *dst, const T *radii,
                                                           const T *latgeom, int nsd, int nx, int nr)
{
    const int64_t n = (int64_t)nsd * nx * nx * nr;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int l = (int)(gid % nr);
    int64_t t = gid / nr;
    const int Y = (int)(t % nx); t /= nx;
    const int X = (int)(t % nx);
    const int sd = (int)(t / nx);
    const int nc = nx - 1;
    const int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};
    // quadrature.hpp:66-101 (Felippa 3x2)
    const T qx[3] = {T(0.6666666666666666), T(0.1666666666666667), T(0.1666666666666667)};
    const T qy[3] = {T(0.1666666666666667), T(0.6666666666666666), T(0.1666666666666667)};
    const T qz[2] = {T(-0.5773502691896257), T(0.5773502691896257)};
    const T qw = T(0.1666666666666667);
    T acc = T(0);
    for (int dxc = -1; dxc <= 0; ++dxc) {
        const int xc = X + dxc;
        if (xc < 0 || xc >= nc) continue;
        for (int dyc = -1; dyc <= 0; ++dyc) {
            const int yc = Y + dyc;
            if (yc < 0 || yc >= nc) continue;
            for (int w = 0; w < 2; ++w) {
                const int kc = corner_of[w][X - xc][Y - yc];
                if (kc < 0) continue;
                const T detl = latgeom[((((int64_t)sd * nc + xc) * nc + yc) * 2 + w) * 10 + 9];
                for (int rc = l - 1; rc <= l; ++rc) {
                    if (rc < 0 || rc >= nr - 1) continue;
                    const int top = rc == l ? 0 : 1;
                    const T r1 = radii[(int64_t)sd * nr + rc], r2 = radii[(int64_t)sd * nr + rc + 1];
                    const T gr = T(0.5) * (r2 - r1);
                    T kv[6];
                    for (int j = 0; j < 6; ++j) {
                        const int jx = xc + kWC[w][j % 3][0], jy = yc + kWC[w][j % 3][1];
                        kv[j] = k[(((int64_t)sd * nx + jx) * nx + jy) * nr + rc + j / 3];
                    }
                    for (int qq = 0; qq < 6; ++qq) {
                        const T xi = qx[qq % 3], eta = qy[qq % 3], zeta = qz[qq / 3];
                        const T Nl[3] = {T(1) - xi - eta, xi, eta};
                        const T Nr[2] = {T(0.5) * (T(1) - zeta), T(0.5) * (T(1) + zeta)};
                        T ke = T(0);
                        for (int j = 0; j < 6; ++j) ke += Nl[j % 3] * Nr[j / 3] * kv[j];
                        const T r = r1 + T(0.5) * (r2 - r1) * (T(1) + zeta);
                        acc += qw * ke * (Nl[kc] * Nr[top] * r * r * gr * detl);
                    }
                }
            }
        }
    }
    dst[gid] = acc * src[gid];
}

// Consistent vector mass (vector_mass.hpp:66-130, 6-point rule), per node and
// component: sum over the <= 12 wedges around the node of
// sum_q w_q N_i(q) u_h(q) r(q)^2 (dr/2) |det L|.
template <typename T>
__global__ __launch_bounds__(256) void vector_mass_kernel(const Comp3<const T> src, const Comp3<T> dst, const T *radii,
                                                          const T *latgeom, int nsd, int nx, int nr)
{
    const int64_t n = (int64_t)nsd * nx * nx * nr;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int l = (int)(gid % nr);
    int64_t t = gid / nr;
    const int Y = (int)(t % nx); t /= nx;
    const int X = (int)(t % nx);
    const int sd = (int)(t / nx);
    const int nc = nx - 1;
    const int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};
    const T qx[3] = {T(0.6666666666666666), T(0.1666666666666667), T(0.1666666666666667)};
    const T qy[3] = {T(0.1666666666666667), T(0.6666666666666666), T(0.1666666666666667)};
    const T qz[2] = {T(-0.5773502691896257), T(0.5773502691896257)};
    const T qw = T(0.1666666666666667);
    T acc[3] = {T(0), T(0), T(0)};
    for (int dxc = -1; dxc <= 0; ++dxc) {
        const int xc = X + dxc;
        if (xc < 0 || xc >= nc) continue;
        for (int dyc = -1; dyc <= 0; ++dyc) {
            const int yc = Y + dyc;
            if (yc < 0 || yc >= nc) continue;
            for (int w = 0; w < 2; ++w) {
                const int kc = corner_of[w][X - xc][Y - yc];
                if (kc < 0) continue;
                const T detl = latgeom[((((int64_t)sd * nc + xc) * nc + yc) * 2 + w) * 10 + 9];
                for (int rc = l - 1; rc <= l; ++rc) {
                    if (rc < 0 || rc >= nr - 1) continue;
                    const int top = rc == l ? 0 : 1;
                    const T r1 = radii[(int64_t)sd * nr + rc], r2 = radii[(int64_t)sd * nr + rc + 1];
                    const T gr = T(0.5) * (r2 - r1);
                    T uv[6][3];
                    for (int j = 0; j < 6; ++j) {
                        const int jx = xc + kWC[w][j % 3][0], jy = yc + kWC[w][j % 3][1];
                        const int64_t ji = (((int64_t)sd * nx + jx) * nx + jy) * nr + rc + j / 3;
                        for (int c = 0; c < 3; ++c) uv[j][c] = src.p[c][ji];
                    }
                    for (int q = 0; q < 6; ++q) {
                        const T xi = qx[q % 3], eta = qy[q % 3], zeta = qz[q / 3];
                        const T Nl[3] = {T(1) - xi - eta, xi, eta};
                        const T Nr[2] = {T(0.5) * (T(1) - zeta), T(0.5) * (T(1) + zeta)};
                        const T r = r1 + T(0.5) * (r2 - r1) * (T(1) + zeta);
                        const T f = qw * Nl[kc] * Nr[top] * r * r * gr * detl;
                        for (int c = 0; c < 3; ++c) {
                            T uh = T(0);
                            for (int j = 0; j < 6; ++j) uh += Nl[j % 3] * Nr[j / 3] * uv[j][c];
                            acc[c] += f * uh;
                        }
                    }
                }
            }
        }
    }
    for (int c = 0; c < 3; ++c) dst.p[c][gid] = acc[c];
}

// Buoyancy of the app's RHS interpolator (apps/mantlecirculation/src/interpolators.hpp:99-133):
// u = Ra * T * x/|x| per node.
template <typename T>
__global__ __launch_bounds__(256) void buoyancy_kernel(const T *temp, const Comp3<T> u, const T *coords, T ra,
                                                       int64_t ncol, int nr)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= ncol * nr) return;
    const T *c = coords + 3 * (gid / nr);
    const T inv = T(1) / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const T v = ra * temp[gid];
    for (int d = 0; d < 3; ++d) u.p[d][gid] = v * c[d] * inv;
}

// one thread per element; the node kernels index with 32 bits
unsigned blocks_for(int64_t n, int bs = 256)
{
    TNG_REQUIRE(n < ((int64_t)1 << 31), "more than 2^31 elements per launch");
    return (unsigned)((n + bs - 1) / bs);
}

template <typename T> struct MeshP {
    const T *coords, *radii, *latgeom;
};
template <typename T> MeshP<T> mp(const tng_domain &m);
template <> MeshP<double> mp<double>(const tng_domain &m)
{
    return {m.d_coords.as<double>(), m.d_radii.as<double>(), m.d_latgeom.as<double>()};
}
template <> MeshP<float> mp<float>(const tng_domain &m)
{
    return {m.d_coords_f.as<float>(), m.d_radii_f.as<float>(), m.d_latgeom_f.as<float>()};
}

template <typename T> void div_impl(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const tng_domain &f = *op.mesh, &c = *op.mesh2;
    const int64_t ncell_c = (int64_t)c.nsd * (c.nx - 1) * (c.nx - 1) * (c.nr - 1);
    const int64_t ncell_f = (int64_t)f.nsd * (f.nx - 1) * (f.nx - 1) * (f.nr - 1);
    // coarse-wedge partials (12 per coarse cell), then the fine-cell values (2 per fine cell)
    const size_t bytes = (size_t)(ncell_c * 12 + ncell_f * 2) * sizeof(T);
    if (op.scratch.bytes < bytes) op.scratch.alloc(bytes);
    DivArgs<T> a{};
    const MeshP<T> m = mp<T>(f);
    a.u0 = (const T *)src.comp(0); a.u1 = (const T *)src.comp(1); a.u2 = (const T *)src.comp(2);
    a.coords = m.coords; a.radii = m.radii; a.latgeom = m.latgeom;
    a.sd_cmb = f.d_sd_cmb.as<uint8_t>(); a.sd_surf = f.d_sd_surf.as<uint8_t>();
    a.partial = op.scratch.as<T>();
    a.nsd = f.nsd; a.nx = f.nx; a.nr = f.nr; a.nxc = c.nx; a.nrc = c.nr;
    a.bc_cmb = op.bc_cmb; a.bc_surf = op.bc_surf;
    a.cidx = Index4(c.nx - 1, c.nr - 1);
    a.fidx = Index4(f.nx - 1, f.nr - 1);
    T *cdiv = op.scratch.as<T>() + ncell_c * 12;
    if ((f.nr - 1) % 64 == 0)
        hipLaunchKernelGGL((div_cell_kernel<T, true>), dim3(blocks_for(ncell_f)), dim3(256), 0, s, a, cdiv);
    else
        hipLaunchKernelGGL((div_cell_kernel<T, false>), dim3(blocks_for(ncell_f)), dim3(256), 0, s, a, cdiv);
    hipLaunchKernelGGL(div_coarse_kernel<T>, dim3(blocks_for(ncell_c)), dim3(256), 0, s, a, (const T *)cdiv);
    TNG_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(wedge_gather_kernel<T>, dim3(blocks_for(c.nalloc)), dim3(256), 0, s,
                       (const T *)op.scratch.p, (T *)dst.comp(0), op.apply_mode == TNG_ADD ? 1 : 0, c.nsd, c.nx, c.nr,
                       Index4(c.nx, c.nr));
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T> void grad_impl(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const tng_domain &f = *op.mesh, &c = *op.mesh2;
    GradArgs<T> a{};
    const MeshP<T> m = mp<T>(f);
    a.p = (const T *)src.comp(0);
    a.d0 = (T *)dst.comp(0); a.d1 = (T *)dst.comp(1); a.d2 = (T *)dst.comp(2);
    a.coords = m.coords; a.radii = m.radii; a.latgeom = m.latgeom;
    a.sd_cmb = f.d_sd_cmb.as<uint8_t>(); a.sd_surf = f.d_sd_surf.as<uint8_t>();
    a.nsd = f.nsd; a.nx = f.nx; a.nr = f.nr; a.nxc = c.nx; a.nrc = c.nr;
    a.bc_cmb = op.bc_cmb; a.bc_surf = op.bc_surf; a.add = op.apply_mode == TNG_ADD ? 1 : 0;
    a.fcell = Index4(f.nx - 1, f.nr - 1);
    a.fnode = Index4(f.nx, f.nr);
    const int64_t nfw = (int64_t)f.nsd * (f.nx - 1) * (f.nx - 1) * (f.nr - 1) * 2;
    const size_t bytes = (size_t)nfw * sizeof(T);
    if (op.scratch.bytes < bytes) op.scratch.alloc(bytes);
    hipLaunchKernelGGL(grad_weight_kernel<T>, dim3(blocks_for(nfw / 2)), dim3(256), 0, s, a, op.scratch.as<T>());
    // a block of 256 threads spans at most 256 / nr + 2 node columns
    if (f.nr >= 52)
        hipLaunchKernelGGL((gradient_kernel<T, true>), dim3(blocks_for(f.nalloc)), dim3(256), 0, s, a,
                           (const T *)op.scratch.p);
    else
        hipLaunchKernelGGL((gradient_kernel<T, false>), dim3(blocks_for(f.nalloc)), dim3(256), 0, s, a,
                           (const T *)op.scratch.p);
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T> void prolong_impl(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const tng_domain &c = *op.mesh2, &f = *op.mesh;  // mesh = fine (dst), mesh2 = coarse (src)
    hipLaunchKernelGGL(prolongate_kernel<T>, dim3(blocks_for(f.nalloc)), dim3(256), 0, s, comps<T>(src),
                       comps<T>(dst), src.ncomp, op.apply_mode == TNG_ADD ? 1 : 0, f.nsd, f.nx, f.nr, c.nx, c.nr,
                       Index4(f.nx, f.nr));
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T> void restrict_impl(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const tng_domain &f = *op.mesh, &c = *op.mesh2;  // mesh = fine (src), mesh2 = coarse (dst)
    hipLaunchKernelGGL(restrict_kernel<T>, dim3(blocks_for(c.nalloc)), dim3(256), 0, s, comps<T>(src),
                       comps<T>(dst), src.ncomp, op.apply_mode == TNG_ADD ? 1 : 0, f.d_omask.as<uint8_t>(), f.nsd,
                       f.nx, f.nr, c.nx, c.nr, Index4(c.nx, c.nr));
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T> void kmass_impl(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const tng_domain &m = *op.mesh;
    const MeshP<T> g = mp<T>(m);
    hipLaunchKernelGGL(kmass_lumped_kernel<T>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s, (const T *)op.k->comp(0),
                       (const T *)src.comp(0), (T *)dst.comp(0), g.radii, g.latgeom, m.nsd, m.nx, m.nr);
    TNG_HIP_CHECK(hipGetLastError());
}

}  // namespace

void stokes_block_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const bool f64 = src.dtype == TNG_F64;
    switch (op.kind) {
    case OP_DIVERGENCE:
        f64 ? div_impl<double>(op, src, dst, s) : div_impl<float>(op, src, dst, s);
        if (op.comm_mode == TNG_COMMUNICATE_ADDITIVELY) launch_halo_sum(*op.mesh2, dst, s);
        break;
    case OP_GRADIENT:
        f64 ? grad_impl<double>(op, src, dst, s) : grad_impl<float>(op, src, dst, s);
        if (op.comm_mode == TNG_COMMUNICATE_ADDITIVELY) launch_halo_sum(*op.mesh, dst, s);
        break;
    case OP_PROLONGATION:
        f64 ? prolong_impl<double>(op, src, dst, s) : prolong_impl<float>(op, src, dst, s);
        break;
    case OP_RESTRICTION:
        f64 ? restrict_impl<double>(op, src, dst, s) : restrict_impl<float>(op, src, dst, s);
        launch_halo_sum(*op.mesh2, dst, s);
        break;
    case OP_KMASS_LUMPED:
        f64 ? kmass_impl<double>(op, src, dst, s) : kmass_impl<float>(op, src, dst, s);
        if (op.comm_mode == TNG_COMMUNICATE_ADDITIVELY) launch_halo_sum(*op.mesh, dst, s);
        break;
    case OP_VECTOR_MASS: {
        const tng_domain &m = *op.mesh;
        if (f64)
            hipLaunchKernelGGL(vector_mass_kernel<double>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s,
                               Comp3<const double>{{(const double *)src.comp(0), (const double *)src.comp(1),
                                                    (const double *)src.comp(2)}},
                               comps<double>(dst), m.d_radii.as<double>(), m.d_latgeom.as<double>(), m.nsd, m.nx, m.nr);
        else
            hipLaunchKernelGGL(vector_mass_kernel<float>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s,
                               Comp3<const float>{{(const float *)src.comp(0), (const float *)src.comp(1),
                                                   (const float *)src.comp(2)}},
                               comps<float>(dst), m.d_radii_f.as<float>(), m.d_latgeom_f.as<float>(), m.nsd, m.nx, m.nr);
        TNG_HIP_CHECK(hipGetLastError());
        if (op.comm_mode == TNG_COMMUNICATE_ADDITIVELY) launch_halo_sum(*op.mesh, dst, s);
        break;
    }
    default:
        throw Error(TNG_INVALID_ARGUMENT, "stokes_block_apply: wrong operator kind");
    }
}

}  // namespace tng

// ----------------------------------------------------------------------------
// Injection of a node field onto the next coarser mesh (coarse node (x, y, r)
// is fine node (2x, 2y, 2r) of the same subdomain) and the strong RHS
// boundary enforcement of a velocity vector
// (fe/strong_algebraic_dirichlet_enforcement.hpp:114-124,
// fe/strong_algebraic_freeslip_enforcement.hpp:22-38).
// ----------------------------------------------------------------------------
namespace tng {

namespace {

template <typename T>
__global__ __launch_bounds__(256) void inject_kernel(const T *f, T *c, int nsd, int nxf, int nrf, int nxc, int nrc)
{
    const int64_t n = (int64_t)nsd * nxc * nxc * nrc;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int r = (int)(gid % nrc);
    int64_t t = gid / nrc;
    const int y = (int)(t % nxc); t /= nxc;
    const int x = (int)(t % nxc);
    const int sd = (int)(t / nxc);
    c[gid] = f[(((int64_t)sd * nxf + 2 * x) * nxf + 2 * y) * nrf + 2 * r];
}

// One thread per boundary-layer node: layer 0 of CMB subdomains and layer
// nr-1 of surface subdomains.
template <typename T>
__global__ __launch_bounds__(256) void rhs_enforce_kernel(const Comp3<T> u, const T *coords, const uint8_t *sd_cmb,
                                                          const uint8_t *sd_surf, int nsd, int nx, int nr, int zero_cmb,
                                                          int zero_surf, int fs_cmb, int fs_surf)
{
    const int64_t n = (int64_t)nsd * nx * nx * 2;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int side = (int)(gid & 1);
    const int64_t lcol = gid >> 1;
    const int sd = (int)(lcol / ((int64_t)nx * nx));
    if (side == 0 ? !sd_cmb[sd] : !sd_surf[sd]) return;
    const bool zero = side == 0 ? zero_cmb : zero_surf;
    const bool fs = side == 0 ? fs_cmb : fs_surf;
    if (!zero && !fs) return;
    const int64_t i = lcol * nr + (side == 0 ? 0 : nr - 1);
    if (zero) {
        u.p[0][i] = T(0); u.p[1][i] = T(0); u.p[2][i] = T(0);
        return;
    }
    const T *c = coords + 3 * lcol;
    const T inv = T(1) / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const T n0 = c[0] * inv, n1 = c[1] * inv, n2 = c[2] * inv;
    const T d = n0 * u.p[0][i] + n1 * u.p[1][i] + n2 * u.p[2][i];
    u.p[0][i] -= d * n0; u.p[1][i] -= d * n1; u.p[2][i] -= d * n2;
}

}  // namespace

void launch_inject(tng_vector &coarse, const tng_vector &fine, hipStream_t s)
{
    const tng_domain &c = *coarse.mesh, &f = *fine.mesh;
    TNG_REQUIRE(coarse.dtype == fine.dtype && coarse.ncomp == fine.ncomp, "inject: mismatched vectors");
    TNG_REQUIRE(f.nsd == c.nsd && f.nx - 1 == 2 * (c.nx - 1) && f.nr - 1 == 2 * (c.nr - 1),
                "inject: meshes must be one refinement apart");
    for (int k = 0; k < coarse.ncomp; ++k) {
        if (coarse.dtype == TNG_F64)
            hipLaunchKernelGGL(inject_kernel<double>, dim3(blocks_for(c.nalloc)), dim3(256), 0, s,
                               (const double *)fine.comp(k), (double *)coarse.comp(k), c.nsd, f.nx, f.nr, c.nx, c.nr);
        else
            hipLaunchKernelGGL(inject_kernel<float>, dim3(blocks_for(c.nalloc)), dim3(256), 0, s,
                               (const float *)fine.comp(k), (float *)coarse.comp(k), c.nsd, f.nx, f.nr, c.nx, c.nr);
        TNG_HIP_CHECK(hipGetLastError());
    }
}

void launch_rhs_enforce(tng_vector &u, uint8_t dirichlet_flag, uint8_t freeslip_flag, hipStream_t s)
{
    TNG_REQUIRE(u.ncomp == 3, "boundary enforcement needs a velocity vector");
    const tng_domain &m = *u.mesh;
    const int zc = dirichlet_flag == SBF_CMB, zs = dirichlet_flag == SBF_SURFACE;
    const int fc = freeslip_flag == SBF_CMB, fsf = freeslip_flag == SBF_SURFACE;
    const int64_t n = (int64_t)m.nsd * m.nx * m.nx * 2;
    if (u.dtype == TNG_F64)
        hipLaunchKernelGGL(rhs_enforce_kernel<double>, dim3(blocks_for(n)), dim3(256), 0, s, comps<double>(u),
                           m.d_coords.as<double>(), m.d_sd_cmb.as<uint8_t>(), m.d_sd_surf.as<uint8_t>(), m.nsd, m.nx,
                           m.nr, zc, zs, fc, fsf);
    else
        hipLaunchKernelGGL(rhs_enforce_kernel<float>, dim3(blocks_for(n)), dim3(256), 0, s, comps<float>(u),
                           m.d_coords_f.as<float>(), m.d_sd_cmb.as<uint8_t>(), m.d_sd_surf.as<uint8_t>(), m.nsd, m.nx,
                           m.nr, zc, zs, fc, fsf);
    TNG_HIP_CHECK(hipGetLastError());
}

}  // namespace tng

namespace tng {
void launch_buoyancy(const tng_vector &temp, double rayleigh, tng_vector &u, hipStream_t s)
{
    const tng_domain &m = *u.mesh;
    TNG_REQUIRE(temp.mesh == u.mesh && temp.ncomp == 1 && u.ncomp == 3 && temp.dtype == u.dtype,
                "buoyancy: temperature (scalar) and velocity on one mesh");
    if (u.dtype == TNG_F64)
        hipLaunchKernelGGL(buoyancy_kernel<double>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s,
                           (const double *)temp.comp(0), comps<double>(u), m.d_coords.as<double>(), rayleigh, m.nlat,
                           m.nr);
    else
        hipLaunchKernelGGL(buoyancy_kernel<float>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s,
                           (const float *)temp.comp(0), comps<float>(u), m.d_coords_f.as<float>(), (float)rayleigh,
                           m.nlat, m.nr);
    TNG_HIP_CHECK(hipGetLastError());
}
}  // namespace tng
