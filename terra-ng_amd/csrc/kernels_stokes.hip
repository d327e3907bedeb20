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

// DPP wave_shr:1 -- lane i receives lane i-1; lane 0 receives 0 (bound control).
__device__ __forceinline__ double wave_shr1_t(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float wave_shr1_t(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
// A wave-uniform value moved to scalar registers.
__device__ __forceinline__ double uniform_value(double v)
{
    return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                            __builtin_amdgcn_readfirstlane(__double2loint(v)));
}
__device__ __forceinline__ float uniform_value(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// DPP wave_shl:1 -- lane i receives lane i+1; lane 63 receives 0.
__device__ __forceinline__ double wave_shl1_t(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float wave_shl1_t(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
// DPP quad_perm [1, 0, 3, 2] -- lanes 2i and 2i+1 exchange values.
__device__ __forceinline__ double pair_swap_t(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xb1, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xb1, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float pair_swap_t(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xb1, 0xf, 0xf, true));
}

// Lateral part of coarse shape function k (0..2) at the centroid (1/3, 1/3)
// of fine lateral wedge lat (0..3: the three corner triangles, then the
// inverted centre one).
template <typename T> __device__ __forceinline__ T shape_lat_c(int k, int lat)
{
    const T third = T(1.0 / 3.0);
    switch (k % 3) {
    case 0: return lat == 0 ? T(-0.5) * third - T(0.5) * third + T(1) : (lat == 3 ? T(0.5) * third + T(0.5) * third
                                                                                   : T(-0.5) * third - T(0.5) * third + T(0.5));
    case 1: return lat == 1 ? T(0.5) * third + T(0.5) : (lat == 3 ? T(0.5) - T(0.5) * third : T(0.5) * third);
    default: return lat == 2 ? T(0.5) * third + T(0.5) : (lat == 3 ? T(0.5) - T(0.5) * third : T(0.5) * third);
    }
}

// Coarse shape function i (0..5) at the fine centroid (1/3, 1/3, 0).
template <typename T> __device__ __forceinline__ T shape_coarse_c(int i, int lat, int rad)
{
    const T sr = (i / 3 == 0) ? (rad == 0 ? T(0.75) : T(0.25)) : (rad == 0 ? T(0.25) : T(0.75));
    return shape_lat_c<T>(i, lat) * sr;
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

// B in one pass.  A workgroup (wx, chunk) owns the coarse node columns
// X0 = wx CW .. X0 + CW - 1 of a subdomain over a chunk of coarse rows
// Y0 .. Y1 - 1; its waves are those node columns Xc times RW radial parts of
// 64 fine cells (lane = fine cell).  Sweeping the coarse rows Y, a wave
// evaluates the four fine cells of coarse cell (Xc, Y) -- per fine wedge
// f = -|det J| vol div u, with
//   div u = (M0 . (U_1 - U_0) + M1 . (U_2 - U_0)) / (2 r_mid) + M2 . sum_k D_k / (6h)
// (U_k = bottom + top and D_k = top - bottom of the wedge's corner column k)
// -- and adds f times the lateral coarse shape value of each corner to the
// corner's node: (Xc, Y), (Xc + 1, Y), or carried to the next row (Xc, Y + 1),
// (Xc + 1, Y + 1).  Node (Xc, Y) is the wave's own sums plus the sums for
// column Xc + 1 of the wave to its left, passed through LDS.  The radial
// coarse weights (3/4, 1/4) split a lane's sum between the bottom and top
// coarse node of its coarse cell, a lane pair forms one coarse cell, and the
// cell below's top part arrives by a DPP shift (across radial parts from the
// lanes 62, 63 sums in LDS).  Every coarse cell is evaluated once: what the
// workgroup's cells add to the nodes of its neighbours (column X0 + CW, row
// Y1 and their corner) is written, as coarse node values, to the side buffer
// and added by div_fixup_kernel, which also sets the last node column and row
// where no workgroup owns them.  The fine node columns are swept along x,
// three rows at a time, the next x-column's loads in flight while one is used
// (uniform base address, lane offset).  The cell geometry (made wave-uniform)
// and the node normals come from LDS tiles.  Trial-side boundary values are
// projected on every lane with 0/1 lane factors (node level 0 belongs only to
// the CMB cell, level ncr only to the surface cell).
constexpr int kDivMaxRows = 8;  // coarse rows per workgroup (cy) at most
#ifndef TNG_DIV_CW
#define TNG_DIV_CW 4  // coarse node columns per workgroup of the one-pass B
#endif
// side buffer of the one-pass B: coarse node columns of nrc values
struct DivSide {
    int nxw, nchunk, cw, cy;
    __host__ __device__ int64_t right(int sd, int wx, int Y, int nxc) const { return ((int64_t)sd * nxw + wx) * nxc + Y; }
    __host__ __device__ int64_t bottom(int sd, int ch, int X, int nxc, int nsd) const
    {
        return (int64_t)nsd * nxw * nxc + ((int64_t)sd * nchunk + ch) * nxc + X;
    }
    __host__ __device__ int64_t corner(int sd, int wx, int ch, int nxc, int nsd) const
    {
        return (int64_t)nsd * (nxw + nchunk) * nxc + ((int64_t)sd * nxw + wx) * nchunk + ch;
    }
    __host__ __device__ int64_t columns(int nxc, int nsd) const { return (int64_t)nsd * ((nxw + nchunk) * (int64_t)nxc + nxw * nchunk); }
};
template <typename T, int RW, int CW>
__global__ __launch_bounds__(64 * RW * CW) void div_column_kernel(const DivArgs<T> a, T *dst, int add, const DivSide sb,
                                                                  T *side)
{
    __shared__ T xr[2][CW][RW][64];  // node column Xc + 1 sums of each wave, by row parity
    __shared__ T xl[2][CW][RW][2];   // lanes 62, 63 of the node column Xc sums
    // lateral geometry of the workgroup's fine cells: 2 CW columns x 2 cy rows
    __shared__ T geo[2 * CW][2 * kDivMaxRows][20];
    // unit normals of the workgroup's fine node columns (free-slip rows)
    __shared__ T nrm[2 * CW + 1][2 * kDivMaxRows + 1][3];
    const int nx = a.nx, nr = a.nr, nc = nx - 1, nxc = a.nxc, nrc = a.nrc;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int part = wv % RW, xs = wv / RW;
    const int cy = sb.cy;
    const int64_t wg = xcd_block();
    const int wx = (int)(wg % sb.nxw);
    const int chunk = (int)((wg / sb.nxw) % sb.nchunk);
    const int sd = (int)(wg / ((int64_t)sb.nxw * sb.nchunk));
    const int X0 = wx * CW, Xc = X0 + xs;
    const bool cells = Xc <= nxc - 2;
    const bool out = Xc <= nxc - 1;
    const bool right_side = X0 + CW <= nxc - 1;  // node column X0 + CW exists
    const int c = part * 64 + lane;              // fine radial cell of the lane
    const int bcb = part == 0 && a.sd_cmb[sd] != 0 ? a.bc_cmb : TNG_NEUMANN;
    const int bct = part == RW - 1 && a.sd_surf[sd] != 0 ? a.bc_surf : TNG_NEUMANN;
    T hs, tz, vol;
    {
        const T r0 = a.radii[(int64_t)sd * nr + c], r1 = a.radii[(int64_t)sd * nr + c + 1];
        const T rmid = T(0.5) * (r0 + r1), h = T(0.5) * (r1 - r0);
        hs = T(0.5) / rmid;
        tz = T(1) / (T(6) * h);
        vol = rmid * rmid * h;
    }
    const bool odd = lane & 1;
    const T srb = odd ? T(0.25) : T(0.75), srt = odd ? T(0.75) : T(0.25);
    const int Y0 = chunk * cy, Y1 = min(Y0 + cy, nxc);
    {
        // fine cells x = 2 X0 + i, y = 2 Y0 + j of existing coarse cells
        const int xf0 = 2 * X0, yf0 = 2 * Y0;
        const int nyf = 2 * (min(Y1 - 1, nxc - 2) - Y0 + 1);
        for (int e = threadIdx.x; e < 2 * CW * 2 * kDivMaxRows * 20; e += 64 * RW * CW) {
            const int v = e % 20, j = (e / 20) % (2 * kDivMaxRows), i = e / (20 * 2 * kDivMaxRows);
            const int xf = xf0 + i, yf = yf0 + j;
            if (xf < nc && j < nyf) geo[i][j][v] = a.latgeom[(((int64_t)sd * nc + xf) * nc + yf) * 20 + v];
        }
        if (bcb == TNG_FREESLIP || bct == TNG_FREESLIP || (a.bc_cmb == TNG_FREESLIP && a.sd_cmb[sd] != 0) ||
            (a.bc_surf == TNG_FREESLIP && a.sd_surf[sd] != 0)) {
            for (int e = threadIdx.x; e < (2 * CW + 1) * (2 * kDivMaxRows + 1) * 3; e += 64 * RW * CW) {
                const int v = e % 3, j = (e / 3) % (2 * kDivMaxRows + 1), i = e / (3 * (2 * kDivMaxRows + 1));
                const int xf = xf0 + i, yf = yf0 + j;
                if (xf < nx && yf < nx) nrm[i][j][v] = a.coords[(((int64_t)sd * nx + xf) * nx + yf) * 3 + v];
            }
        }
        __syncthreads();
    }
    const T *src[3] = {a.u0, a.u1, a.u2};
    // node copies are indexed in 32 bits (checked on the host); the lane's
    // byte offset is the only per-lane part of a load's address
    const uint32_t xstride = (uint32_t)nx * nr, ystride = (uint32_t)nr;
    const uint32_t sdbase = (uint32_t)sd * nx * xstride;
    const uint32_t boff = (uint32_t)c * (uint32_t)sizeof(T);
    // FP32 keeps the sums of a coarse row's last fine row in registers for the
    // next coarse row, where it is the first (level 7: 0.125 ms against 0.135);
    // in FP64 the 33 more VGPRs and the copies between coarse rows cost what
    // the saved loads gain
    constexpr bool KEEP = sizeof(T) == 4;
    constexpr int R0 = KEEP ? 1 : 0;  // first fine row loaded per coarse row
    // raw values of a node column x at rows y0 + R0 .. y0 + 2: levels c and
    // c + 1 of every lane; the next x-column's loads are in flight while one is used
    T pr[3 - R0][3], pt[3 - R0][3];
    auto issue = [&](int x, int y0) {
        const uint32_t cb = sdbase + (uint32_t)x * xstride + (uint32_t)y0 * ystride;
#pragma unroll
        for (int row = R0; row < 3; ++row) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const T *q = (const T *)((const char *)(src[d] + (cb + row * ystride)) + boff);
                pr[row - R0][d] = q[0];
                pt[row - R0][d] = q[1];
            }
        }
    };
    // trial-side boundary factors: the bottom value of lane 0 (CMB) and the
    // top value of lane 63 (surface); Dirichlet multiplies by z, free slip
    // removes f (n . u) n
    const bool fs_any = bcb == TNG_FREESLIP || bct == TNG_FREESLIP;
    const bool dir_any = bcb == TNG_DIRICHLET || bct == TNG_DIRICHLET;
    const T f_lo = (bcb == TNG_FREESLIP && lane == 0) ? T(1) : T(0);
    const T f_hi = (bct == TNG_FREESLIP && lane == 63) ? T(1) : T(0);
    const T z_lo = (bcb == TNG_DIRICHLET && lane == 0) ? T(0) : T(1);
    const T z_hi = (bct == TNG_DIRICHLET && lane == 63) ? T(0) : T(1);
    // one node column row (normal-tile entry ni, nj): the boundary rows
    // applied, then the radial sums U = bottom + top and differences
    // D = top - bottom of the lane's cell.  The sweep below repeats this
    // arithmetic inline: called there, the FP64 kernel is allocated 256 VGPRs
    // and spills (193 inline)
    auto node_row = [&](T (&lo)[3], T (&hi)[3], int ni, int nj, T (&u)[3], T (&dv)[3]) {
        if (fs_any) {
#pragma clang fp contract(off)
            const T *n = nrm[ni][nj];
            const T n0 = n[0], n1 = n[1], n2 = n[2];
            const T dl = f_lo * fma(n2, lo[2], fma(n1, lo[1], n0 * lo[0]));
            const T dh = f_hi * fma(n2, hi[2], fma(n1, hi[1], n0 * hi[0]));
            lo[0] = fma(-dl, n0, lo[0]);
            lo[1] = fma(-dl, n1, lo[1]);
            lo[2] = fma(-dl, n2, lo[2]);
            hi[0] = fma(-dh, n0, hi[0]);
            hi[1] = fma(-dh, n1, hi[1]);
            hi[2] = fma(-dh, n2, hi[2]);
        }
        if (dir_any) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                lo[d] *= z_lo;
                hi[d] *= z_hi;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            u[d] = lo[d] + hi[d];
            dv[d] = hi[d] - lo[d];
        }
    };
    // the last fine row of the previous coarse row, per column of the cell:
    // the first row of the next (FP32: each fine row is loaded and processed once)
    T U2[KEEP ? 3 : 1][3], Dv2[KEEP ? 3 : 1][3];
    if (cells && Y0 <= nxc - 2) {
        // the chunk's first fine row 2 Y0
#pragma unroll
        for (int xi = 0; xi < (KEEP ? 3 : 0); ++xi) {
            const uint32_t b = sdbase + (uint32_t)(2 * Xc + xi) * xstride + (uint32_t)(2 * Y0) * ystride;
            T lo[3], hi[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const T *q = (const T *)((const char *)(src[d] + b) + boff);
                lo[d] = q[0];
                hi[d] = q[1];
            }
            node_row(lo, hi, 2 * xs + xi, 0, U2[xi], Dv2[xi]);
        }
        issue(2 * Xc, 2 * Y0);
    }
    // a node column's radial sums (one per fine cell) as coarse node values:
    // t62, t63 are lanes 62, 63 of the part below; stored to col[0 .. nrc)
    auto put = [&](T tot, T t62, T t63, T *col, bool acc) {
        const T b = tot * srb, t = tot * srt;
        const T bs = b + pair_swap_t(b), ts = t + pair_swap_t(t);
        T below = wave_shr1_t(ts);
        if (part > 0 && lane == 0) below = t63 * T(0.75) + t62 * T(0.25);
        if (!odd) {
            T *p = col + part * 32 + (lane >> 1);
            const T v = bs + below;
            *p = acc ? *p + v : v;
        }
        if (part == RW - 1 && lane == 63) {
            T *p = col + part * 32 + 32;
            *p = acc ? *p + ts : ts;
        }
    };
    T Lc = T(0), Ln = T(0), Rc = T(0), Rn = T(0);
    T U[2][3][3], Dv[2][3][3];  // [x slot][row][component]
    for (int Y = Y0; Y < Y1; ++Y) {
        Lc = Ln;
        Rc = Rn;
        Ln = Rn = T(0);
        if (cells && Y <= nxc - 2) {
#pragma unroll
            for (int xi = 0; xi < 3; ++xi) {
                const int s = xi & 1;
                // the column's sums and differences (row 0 kept from the
                // previous coarse row), then the next column's loads
                if constexpr (KEEP) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        U[s][0][d] = U2[xi][d];
                        Dv[s][0][d] = Dv2[xi][d];
                    }
                }
#pragma unroll
                for (int row = R0; row < 3; ++row) {
                    T lo[3] = {pr[row - R0][0], pr[row - R0][1], pr[row - R0][2]};
                    T hi[3] = {pt[row - R0][0], pt[row - R0][1], pt[row - R0][2]};
                    if (fs_any) {
#pragma clang fp contract(off)
                        const T *n = nrm[2 * xs + xi][2 * (Y - Y0) + row];
                        const T n0 = n[0], n1 = n[1], n2 = n[2];
                        const T dl = f_lo * fma(n2, lo[2], fma(n1, lo[1], n0 * lo[0]));
                        const T dh = f_hi * fma(n2, hi[2], fma(n1, hi[1], n0 * hi[0]));
                        lo[0] = fma(-dl, n0, lo[0]);
                        lo[1] = fma(-dl, n1, lo[1]);
                        lo[2] = fma(-dl, n2, lo[2]);
                        hi[0] = fma(-dh, n0, hi[0]);
                        hi[1] = fma(-dh, n1, hi[1]);
                        hi[2] = fma(-dh, n2, hi[2]);
                    }
                    if (dir_any) {
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            lo[d] *= z_lo;
                            hi[d] *= z_hi;
                        }
                    }
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        U[s][row][d] = lo[d] + hi[d];
                        Dv[s][row][d] = hi[d] - lo[d];
                    }
                }
                if constexpr (KEEP) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        U2[xi][d] = U[s][2][d];
                        Dv2[xi][d] = Dv[s][2][d];
                    }
                }
                if (xi < 2)
                    issue(2 * Xc + xi + 1, 2 * Y);
                else if (Y + 1 < Y1 && Y + 1 <= nxc - 2)
                    issue(2 * Xc, 2 * Y + 2);
                if (xi == 0) continue;
                // the two fine cells (2 Xc + xi - 1, 2Y + fy): corners c00, c10,
                // c01, c11 = (sL, fy), (sR, fy), (sL, fy + 1), (sR, fy + 1);
                // wedge 0 = (c00, c10, c01), wedge 1 = (c11, c01, c10)
                const int fx = xi - 1, sL = (xi - 1) & 1, sR = s;
#pragma unroll
                for (int fy = 0; fy < 2; ++fy) {
                    // the cell's geometry as wave-uniform scalars (as vector
                    // operands it takes ~70 more VGPRs and spills)
                    const T *gl = geo[2 * xs + fx][2 * (Y - Y0) + fy];
                    T gw[20];
#pragma unroll
                    for (int v = 0; v < 20; ++v) gw[v] = uniform_value(gl[v]);
                    T dm[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) dm[d] = Dv[sR][fy][d] + Dv[sL][fy + 1][d];
                    T l0 = T(0), r0 = T(0);
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const T u00 = U[sL][fy][d], u10 = U[sR][fy][d], u01 = U[sL][fy + 1][d];
                        l0 += gw[d] * (u10 - u00) + gw[3 + d] * (u01 - u00);
                        r0 += gw[6 + d] * (Dv[sL][fy][d] + dm[d]);
                    }
                    const T f0 = (-gw[9] * vol) * (hs * l0 + tz * r0);
                    T l1 = T(0), r1 = T(0);
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const T u10 = U[sR][fy][d], u01 = U[sL][fy + 1][d], u11 = U[sR][fy + 1][d];
                        l1 += gw[10 + d] * (u01 - u11) + gw[13 + d] * (u10 - u11);
                        r1 += gw[16 + d] * (Dv[sR][fy + 1][d] + dm[d]);
                    }
                    const T f1 = (-gw[19] * vol) * (hs * l1 + tz * r1);
                    const int lat0 = fine_lat_idx(fx, fy, 0), lat1 = fine_lat_idx(fx, fy, 1);
                    Lc += shape_lat_c<T>(0, lat0) * f0;
                    Rc += shape_lat_c<T>(1, lat0) * f0;
                    Ln += shape_lat_c<T>(2, lat0) * f0;
                    Rn += shape_lat_c<T>(0, lat1) * f1;
                    Ln += shape_lat_c<T>(1, lat1) * f1;
                    Rc += shape_lat_c<T>(2, lat1) * f1;
                }
            }
        }
        // node (Xc, Y) = this wave's left sums + the left neighbour's right sums
        const int pb = Y & 1;
        xr[pb][xs][part][lane] = Rc;
        if (lane >= 62) xl[pb][xs][part][lane - 62] = Lc;
        __syncthreads();
        if (!out) continue;
        T tot = Lc, t62 = T(0), t63 = T(0);
        if (xs > 0) tot += xr[pb][xs - 1][part][lane];
        if (part > 0) {
            t62 = xl[pb][xs][part - 1][0];
            t63 = xl[pb][xs][part - 1][1];
            if (xs > 0) {
                t62 += xr[pb][xs - 1][part - 1][62];
                t63 += xr[pb][xs - 1][part - 1][63];
            }
        }
        put(tot, t62, t63, dst + (((int64_t)sd * nxc + Xc) * nxc + Y) * nrc, add != 0);
        // the right neighbour's node (X0 + CW, Y): the last wave's right sums
        if (xs == CW - 1 && right_side) {
            const T r62 = part > 0 ? xr[pb][xs][part - 1][62] : T(0);
            const T r63 = part > 0 ? xr[pb][xs][part - 1][63] : T(0);
            put(Rc, r62, r63, side + sb.right(sd, wx, Y, nxc) * nrc, false);
        }
    }
    // node row Y1 (the next chunk's) from cell row Y1 - 1; the buffer of parity
    // Y1 was last read before the barrier of row Y1 - 1
    if (Y1 <= nxc - 1) {
        const int pb = Y1 & 1;
        xr[pb][xs][part][lane] = Rn;
        if (lane >= 62) xl[pb][xs][part][lane - 62] = Ln;
        __syncthreads();
        if (out) {
            T tot = Ln, t62 = T(0), t63 = T(0);
            if (xs > 0) tot += xr[pb][xs - 1][part][lane];
            if (part > 0) {
                t62 = xl[pb][xs][part - 1][0];
                t63 = xl[pb][xs][part - 1][1];
                if (xs > 0) {
                    t62 += xr[pb][xs - 1][part - 1][62];
                    t63 += xr[pb][xs - 1][part - 1][63];
                }
            }
            put(tot, t62, t63, side + sb.bottom(sd, chunk, Xc, nxc, a.nsd) * nrc, false);
            if (xs == CW - 1 && right_side) {
                const T r62 = part > 0 ? xr[pb][xs][part - 1][62] : T(0);
                const T r63 = part > 0 ? xr[pb][xs][part - 1][63] : T(0);
                put(Rn, r62, r63, side + sb.corner(sd, wx, chunk, nxc, a.nsd) * nrc, false);
            }
        }
    }
}

// The side contributions of the one-pass B: node column X = k CW (k >= 1)
// receives the right sums of workgroup column k - 1, node row Y = k cy the
// bottom sums of chunk k - 1, and their crossing the corner sums.  The last
// node column (row) has no owner when CW (cy) divides the coarse cell count:
// there the side values are the node's whole sum and the apply mode applies.
// One thread per coarse node value, the radial level fastest.
template <typename T>
__global__ __launch_bounds__(256) void div_fixup_kernel(const T *side, T *dst, int add, const DivSide sb, int nsd,
                                                        int nxc, int nrc)
{
    const int ncl = (nxc - 1) / sb.cw, nrl = (nxc - 1) / sb.cy;
    const int64_t per_sd = (int64_t)(ncl + nrl) * nxc;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per_sd * nsd * nrc) return;
    const int k = (int)(t % nrc);
    const int64_t n = t / nrc;
    const int sd = (int)(n / per_sd);
    const int64_t j = n % per_sd;
    int X, Y;
    if (j < (int64_t)ncl * nxc) {
        X = (int)(j / nxc + 1) * sb.cw;
        Y = (int)(j % nxc);
    } else {
        const int64_t i = j - (int64_t)ncl * nxc;
        Y = (int)(i / nxc + 1) * sb.cy;
        X = (int)(i % nxc);
        if (X > 0 && X % sb.cw == 0) return;  // on a column line: done there
    }
    // node columns (rows) below xcov (ycov) belong to a workgroup column
    // (chunk), whose sweep wrote their right (bottom) sums
    const int xcov = min(sb.nxw * sb.cw, nxc), ycov = min(sb.nchunk * sb.cy, nxc);
    const bool col = X > 0 && X % sb.cw == 0, row = Y > 0 && Y % sb.cy == 0;
    T v = T(0);
    if (col && Y < ycov) v += side[sb.right(sd, X / sb.cw - 1, Y, nxc) * nrc + k];
    if (row && X < xcov) v += side[sb.bottom(sd, Y / sb.cy - 1, X, nxc, nsd) * nrc + k];
    if (col && row) v += side[sb.corner(sd, X / sb.cw - 1, Y / sb.cy - 1, nxc, nsd) * nrc + k];
    const bool owned = X < xcov && Y < ycov;
    T *p = dst + (((int64_t)sd * nxc + X) * nxc + Y) * nrc + k;
    *p = (owned || add) ? *p + v : v;
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

// B^T as one wave per fine node column (X, Y) with K cells per lane (64 K
// radial cells; the top node level is the last lane's extra output).  Per
// lane and cell c (the cell above node level c) and per wedge w that contains
// the column (corner k): f = -|det J| p(centroid), a = f / r_mid, b = f / (6h);
// node level c then receives, per wedge,
//   (a_{c-1} + a_c) g_k + (b_{c-1} - b_c) M2,   g_k = l0 M0 + l1 M1,
// the cell below's a, b coming from the lane below by a DPP shift.  Each lane
// loads the coarse pressure levels of its cells from the 3 x 3 coarse node
// columns around the column, all at once; the lateral geometry is
// wave-uniform.  The parities (PX, PY) of the column fix which coarse columns
// and lateral shape values each wedge uses, so they are compile-time.
// Boundary nodes: Dirichlet 0, free slip the tangential part.
// Coarse levels a lane reads per coarse column (K fine cells per lane).
template <int K> constexpr int grad_np() { return K == 1 ? 2 : K / 2 + 1; }

// Radial factors of the lane's K fine cells: 1 / r_mid, 1 / (6 h), r_mid^2 h.
template <typename T, int K>
__device__ __forceinline__ void grad_radial(const GradArgs<T> &a, int lane, int sd, T (&sfac)[K], T (&tz)[K],
                                            T (&vol)[K])
{
#pragma unroll
    for (int p = 0; p < K; ++p) {
        const int c = lane * K + p;
        const T r0 = a.radii[(int64_t)sd * a.nr + c], r1 = a.radii[(int64_t)sd * a.nr + c + 1];
        const T rmid = T(0.5) * (r0 + r1), h = T(0.5) * (r1 - r0);
        sfac[p] = T(1) / rmid;
        tz[p] = T(1) / (T(6) * h);
        vol[p] = rmid * rmid * h;
    }
}

// Streaming store of a lane's run of K consecutive values (K = 1, 2, 4; the
// run is aligned to one value only).
template <typename T, int K> __device__ __forceinline__ void store_run_nt(T *p, const T (&v)[K])
{
    if constexpr (K == 1) {
        __builtin_nontemporal_store(v[0], p);
    } else {
        typedef T pair_t __attribute__((ext_vector_type(2), aligned(sizeof(T))));
#pragma unroll
        for (int i = 0; i < K; i += 2) {
            const pair_t x = {v[i], v[i + 1]};
            __builtin_nontemporal_store(x, (pair_t *)(p + i));
        }
    }
}

// B^T at the fine node column (X, Y) (parities PX, PY) from the coarse values
// pcr[i][j][p] of the coarse columns (cx0 + i, cy0 + j), cx0 = floor((X - 1) / 2),
// cy0 = floor((Y - 1) / 2), interpolated radially to the centroid of the lane's
// cell p (the clamped columns are never used).  vh, vt: the lane's radial
// factors -vol / (2 r_mid) and -vol / (6 h).
template <typename T, int K, int PX, int PY>
__device__ __forceinline__ void gradient_node(const GradArgs<T> &a, int lane, int X, int Y, int sd, bool has_cmb,
                                              bool has_surf, const T (&pcr)[3][3][K], const T (&vh)[K],
                                              const T (&vt)[K])
{
    const int nx = a.nx, nr = a.nr, nc = nx - 1, ncr = nr - 1;
    constexpr int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};
    T o[K + 1][3];
#pragma unroll
    for (int p = 0; p <= K; ++p)
#pragma unroll
        for (int d = 0; d < 3; ++d) o[p][d] = T(0);
    using CT = const T __attribute__((address_space(4)));
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int xc = X - 1 + i;
        if (xc < 0 || xc >= nc) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int yc = Y - 1 + j;
            if (yc < 0 || yc >= nc) continue;
            const CT *gcell = (const CT *)a.latgeom + (((int64_t)sd * nc + xc) * nc + yc) * 20;
            // coarse column of the cell relative to (cx0, cy0); parities of the cell
            const int cxl = i == 0 ? 0 : (PX ? 0 : 1), cyl = j == 0 ? 0 : (PY ? 0 : 1);
            const int pxc = PX ^ 1 ^ i, pyc = PY ^ 1 ^ j;
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int k = corner_of[w][1 - i][1 - j];
                if (k < 0) continue;
                T g[10];
#pragma unroll
                for (int v = 0; v < 10; ++v) g[v] = gcell[w * 10 + v];
                // the corner's lateral shape gradient times 2: -(M0 + M1), M0, M1
                T gl[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) gl[d] = k == 0 ? -(g[d] + g[3 + d]) : (k == 1 ? g[d] : g[3 + d]);
                const int lat = fine_lat_idx(pxc, pyc, w);
                const T sl0 = shape_lat_c<T>(0, lat), sl1 = shape_lat_c<T>(1, lat), sl2 = shape_lat_c<T>(2, lat);
                // coarse columns of the coarse wedge's corners 0, 1, 2
                const int ax = cxl + (w == 0 ? 0 : 1), ay = cyl + (w == 0 ? 0 : 1);
                const int bx = cxl + (w == 0 ? 1 : 0), by = cyl + (w == 0 ? 0 : 1);
                const int dx = cxl + (w == 0 ? 0 : 1), dy = cyl + (w == 0 ? 1 : 0);
                T av[K], bv[K];
#pragma unroll
                for (int p = 0; p < K; ++p) {
                    const T pi = sl0 * pcr[ax][ay][p] + sl1 * pcr[bx][by][p] + sl2 * pcr[dx][dy][p];
                    const T gp = g[9] * pi;
                    av[p] = gp * vh[p];
                    bv[p] = gp * vt[p];
                }
                // the cell below the lane's first node level: the lane below's last cell
                const T ab = wave_shr1_t(av[K - 1]), bb = wave_shr1_t(bv[K - 1]);
#pragma unroll
                for (int p = 0; p < K; ++p) {
                    const T A = (p == 0 ? ab : av[p - 1]) + av[p];
                    const T Bv = (p == 0 ? bb : bv[p - 1]) - bv[p];
#pragma unroll
                    for (int d = 0; d < 3; ++d) o[p][d] += A * gl[d] + Bv * g[6 + d];
                }
                // the top node level (last lane): the last cell only
#pragma unroll
                for (int d = 0; d < 3; ++d) o[K][d] += av[K - 1] * gl[d] + bv[K - 1] * g[6 + d];
            }
        }
    }
    const int64_t lcol = ((int64_t)sd * nx + X) * nx + Y;
    // the node's coordinates through the scalar cache (a constant address
    // space: its loads need not wait for the stores of the previous node)
    const CT *cn = (const CT *)a.coords + 3 * lcol;
    T *dst[3] = {a.d0, a.d1, a.d2};
    // the test-side boundary rows: Dirichlet removes the node's row, free slip
    // its normal component
    auto boundary = [&](int bc, T (&v)[3]) {
        if (bc == TNG_DIRICHLET) {
            v[0] = v[1] = v[2] = T(0);
        } else if (bc == TNG_FREESLIP) {
            const T dd = cn[0] * v[0] + cn[1] * v[1] + cn[2] * v[2];
            v[0] -= dd * cn[0]; v[1] -= dd * cn[1]; v[2] -= dd * cn[2];
        }
    };
    // levels lane K .. lane K + K - 1: one contiguous run per component (the
    // CMB level is lane 0's first)
    T run[3][K];
#pragma unroll
    for (int p = 0; p < K; ++p)
#pragma unroll
        for (int d = 0; d < 3; ++d) run[d][p] = o[p][d];
    if (lane == 0 && has_cmb) {
        T v[3] = {run[0][0], run[1][0], run[2][0]};
        boundary(a.bc_cmb, v);
#pragma unroll
        for (int d = 0; d < 3; ++d) run[d][0] = v[d];
    }
    // the velocity is written once and read by a later kernel: streaming
    // (non-temporal) stores of the lanes' runs keep the writes from competing
    // for the caches (level 7 FP64: 0.180 ms against 0.237 with write-back
    // stores); the single-lane top level stays write-back, so that L2 merges
    // it with the runs around it (as a streaming store it cost FP64 0.20 ms
    // and FP32 0.147 against 0.114)
    const int64_t i0 = lcol * nr + lane * K;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (a.add) {
            T old[K];
            __builtin_memcpy(old, dst[d] + i0, sizeof(old));
#pragma unroll
            for (int p = 0; p < K; ++p) run[d][p] += old[p];
        }
        store_run_nt<T, K>(dst[d] + i0, run[d]);
    }
    // the top level (the surface level of the column): lane 63's last cell
    if (lane == 63) {
        T v[3] = {o[K][0], o[K][1], o[K][2]};
        if (has_surf) boundary(a.bc_surf, v);
        const int64_t i = lcol * nr + ncr;
#pragma unroll
        for (int d = 0; d < 3; ++d) dst[d][i] = a.add ? dst[d][i] + v[d] : v[d];
    }
}

constexpr int kGcolWaves = 4;
// Fine rows per wave: FP64 2, FP32 4 (level 7: FP64 0.259 ms against 0.271
// with 4 and 0.277 with 8 rows, the waves' FP64 work being the bound; FP32
// 0.129 ms against 0.158 with one row).
template <typename T> constexpr int grad_rows() { return sizeof(T) == 8 ? 2 : 4; }
// B^T with one wave per fine node column and chunk of grad_rows rows (lane =
// fine radial cells, K per lane).  The radial factors are formed once per
// wave; the coarse rows the chunk needs (its nodes' 3 x 3 coarse columns
// slide by one coarse row every two fine rows) are each loaded once, the
// next one requested a node ahead of its first use.
template <typename T, int K, int PX>
__device__ __forceinline__ void gradient_sweep(const GradArgs<T> &a, int lane, int X, int chunk, int sd)
{
    constexpr int NP = grad_np<K>(), NROW = grad_rows<T>(), NWIN = NROW / 2 + 3;
    const int nxc = a.nxc;
    const int Ya = chunk * NROW;
    const int cx0 = (X - 1) >> 1, cr0 = (Ya >> 1) - 1;  // window row j = coarse row cr0 + j
    const int cb = K == 1 ? lane / 2 : lane * (K / 2);     // first coarse level of the lane
    const bool has_cmb = a.sd_cmb[sd] != 0, has_surf = a.sd_surf[sd] != 0;
    T sfac[K], tz[K], vol[K], vh[K], vt[K];
    grad_radial<T, K>(a, lane, sd, sfac, tz, vol);
#pragma unroll
    for (int p = 0; p < K; ++p) {
        // 1 / (2 r_mid): the lateral shape gradients' factor 1/2 (exact) moves here
        vh[p] = -(vol[p] * (T(0.5) * sfac[p]));
        vt[p] = -(vol[p] * tz[p]);
    }
    T win[NWIN][3][NP];
    auto load_row = [&](int j) {
        const int cy = min(max(cr0 + j, 0), nxc - 1);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int cx = min(max(cx0 + i, 0), nxc - 1);
            const T *pp = a.p + (((int64_t)sd * nxc + cx) * nxc + cy) * a.nrc + cb;
#pragma unroll
            for (int q = 0; q < NP; ++q) win[j][i][q] = pp[q];
        }
    };
#pragma unroll
    for (int j = 0; j < (NWIN < 4 ? NWIN : 4); ++j) load_row(j);
    // the window rows interpolated radially to the centroid of each of the
    // lane's cells, once per row (the nodes and wedges share them)
    T pw[NWIN][3][K];
    auto interp_row = [&](int j) {
#pragma unroll
        for (int p = 0; p < K; ++p) {
            const int q = K == 1 ? 0 : p / 2;  // coarse level of the cell within the lane
            const bool top = K == 1 ? (lane & 1) != 0 : (p & 1) != 0;
            const T wb = top ? T(0.25) : T(0.75), wt = top ? T(0.75) : T(0.25);
#pragma unroll
            for (int i = 0; i < 3; ++i) pw[j][i][p] = wb * win[j][i][q] + wt * win[j][i][q + 1];
        }
    };
#pragma unroll
    for (int k = 0; k < NROW; ++k) {
        // node k uses window rows j0 .. j0 + 2; at odd k the row that node
        // k + 2 needs first is requested
        const int j0 = k == 0 ? 0 : (k - 1) / 2 + 1;
        if ((k & 1) == 1 && j0 + 3 < NWIN) load_row(j0 + 3);
        const int Y = Ya + k;
        if (Y >= a.nx) break;
        // rows that node k - 1 did not use (it used up to row j0(k - 1) + 2)
        const int j0prev = k <= 1 ? 0 : (k - 2) / 2 + 1;
        const int jnew = k == 0 ? 0 : j0prev + 3;
#pragma unroll
        for (int j = jnew; j <= j0 + 2; ++j) interp_row(j);
        T pcr[3][3][K];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                for (int p = 0; p < K; ++p) pcr[i][jj][p] = pw[j0 + jj][i][p];
        if (k & 1) gradient_node<T, K, PX, 1>(a, lane, X, Y, sd, has_cmb, has_surf, pcr, vh, vt);
        else gradient_node<T, K, PX, 0>(a, lane, X, Y, sd, has_cmb, has_surf, pcr, vh, vt);
    }
}

template <typename T, int K>
__global__ __launch_bounds__(64 * kGcolWaves) void gradient_column_kernel(const GradArgs<T> a)
{
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nyc = (a.nx + grad_rows<T>() - 1) / grad_rows<T>();
    const int64_t nw = (int64_t)a.nsd * a.nx * nyc;
    const int64_t w = (int64_t)xcd_block() * kGcolWaves + wv;
    if (w >= nw) return;
    const int chunk = (int)(w % nyc);
    const int X = (int)((w / nyc) % a.nx);
    const int sd = (int)(w / ((int64_t)nyc * a.nx));
    if (X & 1) gradient_sweep<T, K, 1>(a, lane, X, chunk, sd);
    else gradient_sweep<T, K, 0>(a, lane, X, chunk, sd);
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
                                                          const T *latgeom, int nsd, int nx, int nr, int ncomp)
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
                        for (int c = 0; c < 3; ++c) uv[j][c] = c < ncomp ? src.p[c][ji] : T(0);
                    }
                    for (int q = 0; q < 6; ++q) {
                        const T xi = qx[q % 3], eta = qy[q % 3], zeta = qz[q / 3];
                        const T Nl[3] = {T(1) - xi - eta, xi, eta};
                        const T Nr[2] = {T(0.5) * (T(1) - zeta), T(0.5) * (T(1) + zeta)};
                        const T r = r1 + T(0.5) * (r2 - r1) * (T(1) + zeta);
                        const T f = qw * Nl[kc] * Nr[top] * r * r * gr * detl;
                        for (int c = 0; c < ncomp; ++c) {
                            T uh = T(0);
                            for (int j = 0; j < 6; ++j) uh += Nl[j % 3] * Nr[j / 3] * uv[j][c];
                            acc[c] += f * uh;
                        }
                    }
                }
            }
        }
    }
    for (int c = 0; c < ncomp; ++c) dst.p[c][gid] = acc[c];
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
    DivArgs<T> a{};
    const MeshP<T> m = mp<T>(f);
    a.u0 = (const T *)src.comp(0); a.u1 = (const T *)src.comp(1); a.u2 = (const T *)src.comp(2);
    a.coords = m.coords; a.radii = m.radii; a.latgeom = m.latgeom;
    a.sd_cmb = f.d_sd_cmb.as<uint8_t>(); a.sd_surf = f.d_sd_surf.as<uint8_t>();
    a.nsd = f.nsd; a.nx = f.nx; a.nr = f.nr; a.nxc = c.nx; a.nrc = c.nr;
    a.bc_cmb = op.bc_cmb; a.bc_surf = op.bc_surf;
    a.cidx = Index4(c.nx - 1, c.nr - 1);
    a.fidx = Index4(f.nx - 1, f.nr - 1);
    // one pass for 128 radial cells, and for 64 in FP32: workgroups of CW
    // coarse node columns (times 2 radial parts for 128 cells) and a chunk of
    // coarse rows, then the side contributions between workgroups
    const int ncr = f.nr - 1;
    if ((ncr == 128 || (ncr == 64 && sizeof(T) == 4)) && f.nalloc < ((int64_t)1 << 31)) {
        constexpr int CW = TNG_DIV_CW;
        const int RW = ncr == 128 ? 2 : 1;
        DivSide sb{};
        sb.cw = CW;
        sb.nxw = (c.nx - 1 + CW - 1) / CW;
        // rows per workgroup: 8 unless that leaves fewer than 8192 waves
        sb.cy = kDivMaxRows;
        while (sb.cy > 2 && (int64_t)f.nsd * sb.nxw * ((c.nx - 1 + sb.cy - 1) / sb.cy) * CW * RW < 8192) sb.cy /= 2;
        sb.nchunk = (c.nx - 1 + sb.cy - 1) / sb.cy;
        const size_t bytes = (size_t)sb.columns(c.nx, f.nsd) * c.nr * sizeof(T);
        if (op.scratch.bytes < bytes) op.scratch.alloc(bytes);
        const unsigned g = (unsigned)(f.nsd * sb.nxw * sb.nchunk);
        const int add = op.apply_mode == TNG_ADD ? 1 : 0;
        if (ncr == 64)
            hipLaunchKernelGGL((div_column_kernel<T, 1, CW>), dim3(g), dim3(64 * CW), 0, s, a, (T *)dst.comp(0), add, sb,
                               op.scratch.as<T>());
        else
            hipLaunchKernelGGL((div_column_kernel<T, 2, CW>), dim3(g), dim3(128 * CW), 0, s, a, (T *)dst.comp(0), add, sb,
                               op.scratch.as<T>());
        TNG_HIP_CHECK(hipGetLastError());
        const int64_t nfix = (int64_t)f.nsd * ((c.nx - 1) / CW + (c.nx - 1) / sb.cy) * c.nx * c.nr;
        if (nfix > 0) {
            hipLaunchKernelGGL(div_fixup_kernel<T>, dim3(blocks_for(nfix)), dim3(256), 0, s, (const T *)op.scratch.p,
                               (T *)dst.comp(0), add, sb, f.nsd, c.nx, c.nr);
            TNG_HIP_CHECK(hipGetLastError());
        }
        return;
    }
    const size_t bytes = (size_t)(ncell_c * 12 + ncell_f * 2) * sizeof(T);
    if (op.scratch.bytes < bytes) op.scratch.alloc(bytes);
    a.partial = op.scratch.as<T>();
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
    // one wave per node column and chunk of rows: 64, 128 or 256 radial cells
    const int ncr = f.nr - 1;
    if (ncr == 64 || ncr == 128 || ncr == 256) {
        const int64_t nw = (int64_t)f.nsd * f.nx * ((f.nx + grad_rows<T>() - 1) / grad_rows<T>());
        const unsigned g = (unsigned)((nw + kGcolWaves - 1) / kGcolWaves);
        if (ncr == 64)
            hipLaunchKernelGGL((gradient_column_kernel<T, 1>), dim3(g), dim3(64 * kGcolWaves), 0, s, a);
        else if (ncr == 128)
            hipLaunchKernelGGL((gradient_column_kernel<T, 2>), dim3(g), dim3(64 * kGcolWaves), 0, s, a);
        else
            hipLaunchKernelGGL((gradient_column_kernel<T, 4>), dim3(g), dim3(64 * kGcolWaves), 0, s, a);
        TNG_HIP_CHECK(hipGetLastError());
        return;
    }
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
                               comps<double>(dst), m.d_radii.as<double>(), m.d_latgeom.as<double>(), m.nsd, m.nx, m.nr,
                               3);
        else
            hipLaunchKernelGGL(vector_mass_kernel<float>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s,
                               Comp3<const float>{{(const float *)src.comp(0), (const float *)src.comp(1),
                                                   (const float *)src.comp(2)}},
                               comps<float>(dst), m.d_radii_f.as<float>(), m.d_latgeom_f.as<float>(), m.nsd, m.nx, m.nr,
                               3);
        TNG_HIP_CHECK(hipGetLastError());
        if (op.comm_mode == TNG_COMMUNICATE_ADDITIVELY) launch_halo_sum(*op.mesh, dst, s);
        break;
    }
    default:
        throw Error(TNG_INVALID_ARGUMENT, "stokes_block_apply: wrong operator kind");
    }
}

// Consistent scalar mass (fe/wedge/operators/shell/mass.hpp), FP64, additive exchange.
void launch_scalar_mass(const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const tng_domain &m = *src.mesh;
    TNG_REQUIRE(src.ncomp == 1 && dst.ncomp == 1 && src.dtype == TNG_F64 && dst.dtype == TNG_F64 && dst.mesh == &m,
                "scalar mass: FP64 scalars on one mesh");
    const double *sp = (const double *)src.comp(0);
    double *dp = (double *)dst.comp(0);
    hipLaunchKernelGGL(vector_mass_kernel<double>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s,
                       Comp3<const double>{{sp, sp, sp}}, Comp3<double>{{dp, dp, dp}}, m.d_radii.as<double>(),
                       m.d_latgeom.as<double>(), m.nsd, m.nx, m.nr, 1);
    TNG_HIP_CHECK(hipGetLastError());
    launch_halo_sum(m, dst, s);
}

// Node sums of per-cell wedge partials (cell, wedge, 6 wedge nodes), FP64 scalar.
void launch_wedge_gather(const double *partial, tng_vector &dst, bool add, hipStream_t s)
{
    const tng_domain &m = *dst.mesh;
    hipLaunchKernelGGL(wedge_gather_kernel<double>, dim3(blocks_for(m.nalloc)), dim3(256), 0, s, partial,
                       (double *)dst.comp(0), add ? 1 : 0, m.nsd, m.nx, m.nr, Index4(m.nx, m.nr));
    TNG_HIP_CHECK(hipGetLastError());
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
