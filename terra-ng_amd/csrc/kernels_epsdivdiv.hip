// EpsilonDivDiv viscous operator (2 eta eps(u):eps(v) - 2/3 eta div u div v) on
// Q1 wedges of the thick spherical shell, one-point quadrature at the wedge
// centroid (reference: fe/wedge/operators/shell/epsilon_divdiv_kerngen.hpp,
// fast paths :1245-1607 (Dirichlet/Neumann) and :1638-2198 (free-slip)).
//
// Formulation used here.  A wedge with lateral unit-sphere corners p0, p1, p2
// between shells r0 < r1 has, at the centroid, the Jacobian
//     J = L * diag(r_mid, r_mid, h),  L = [p1-p0, p2-p0, (p0+p1+p2)/3],
//     r_mid = (r0+r1)/2, h = (r1-r0)/2,
// so J^{-T} = L^{-T} diag(1/r_mid, 1/r_mid, 1/h) and |det J| = |det L| r_mid^2 h.
// L^{-T} (columns M0, M1, M2) and |det L| depend only on the lateral triangle
// and are precomputed once per mesh.  With sigma_k = u_k(bottom) + u_k(top)
// and delta_k = u_k(top) - u_k(bottom) at the three lateral corners, the
// velocity gradient at the centroid is
//     grad u = (1/(2 r_mid)) [(sigma_1-sigma_0) M0^T + (sigma_2-sigma_0) M1^T]
//            + (1/(6 h)) (delta_0+delta_1+delta_2) M2^T,
// the stress is tau = k |det J| (2 sym(grad u) - 2/3 div u I) with k the mean
// of the six nodal coefficients, and the test-side contribution to corner k is
// tau g_k with g_k = J^{-T} grad N_k.  This is the reference's arithmetic
// regrouped (about 130 flops per wedge instead of about 450).
//
// Work decomposition.  Lanes run along r (one radial cell per lane; 128
// layers as two radial chunks of one wave each; K cells per lane beyond) so
// every global access is a contiguous run of the r-fastest layout.  A
// "virtual column" is a group of R lanes that owns one lateral cell column
// (xc) and sweeps a segment of y.  The VX virtual columns of a workgroup cover
// consecutive cell columns (xc = -1 stands for the empty column left of a
// subdomain; with one column per wave the strips run across subdomains);
// virtual column v outputs node column xc+1 for v < VX-1, so neighbouring
// strips overlap by one column.  The lateral data of the next step (geometry,
// normals, boundary rows) is one vector load of a per-wave record, broadcast
// with readlane.  The contributions of a cell to its left corners
// belong to the neighbour on the left and are passed through LDS once per
// y-step, already combined radially; radial neighbours are combined with a
// DPP wave shift.  The next row of node columns is prefetched one step ahead.
// Boundary rows (Dirichlet, free-slip) are handled with per-lane factors, not
// branches.  Every dst entry is written exactly once with a plain store: no
// atomics, no zeroing pass, deterministic results.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>

#include "tng_internal.hpp"

#ifndef TNG_APPLY_SMEM_GEOM
#define TNG_APPLY_SMEM_GEOM 1
#endif

namespace tng {

namespace {

template <typename T> struct ApplyArgs {
    const T *src0, *src1, *src2;
    T *dst0, *dst1, *dst2;
    const T *k;
    const T *radii;     // (nsd, nr)
    const T *latgeom;   // (nsd, nc, nc, 2, 10)
    const T *coords;    // (nsd, nx, nx, 3) unit vectors (the radial normal)
    const T *btab;      // (nsd, nx, nx, 2, 3) boundary rows
    const double *pen;  // 3 doubles (penalty axis b), or nullptr
    // or: the per-block rotation moments, combined by every workgroup
    const double *pen_part;
    int pen_nb;
    double pen_eps, pen_ginv[9];
    const uint8_t *sd_cmb, *sd_surf;
    int nsd, nx, nr, ncells;
    int bc_cmb, bc_surf;
    int add;
    int strips, groups;  // strips per subdomain row (or of all subdomains), lateral groups
    int segs, seg_len;   // one-segment launches: segments per strip, rows per segment
    // rsplit > 1: every subdomain is swept as rsplit radial parts of
    // ncells / rsplit cells, one after the other like separate subdomains;
    // the node levels between two parts are written as two partial sums to
    // iface ((interface, lower|upper, component) planes of nsd*nx*nx) and
    // completed by epsdivdiv_iface_kernel
    int rsplit;
    T *iface;
    // fused Chebyshev sub-step (FUSE kernels): on node copies no other
    // subdomain shares, d = az * (b - A src) * invd [+ ad * d] and
    // xo = src + d replace the store of A src
    const T *cb0, *cb1, *cb2, *ci0, *ci1, *ci2;
    T *cd0, *cd1, *cd2, *cx0, *cx1, *cx2;
    T caz, cad;
    int cread_d;
};

template <typename T> __device__ __forceinline__ void normal_of(const T *c, T n[3])
{
    const T n2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const T inv = T(1) / sqrt(n2);
    n[0] = c[0] * inv;
    n[1] = c[1] * inv;
    n[2] = c[2] * inv;
}

// DPP wave_shr:1 -- lane i receives lane i-1; lane 0 has no source and the
// bound control writes 0 there (no initialisation of the destination).
__device__ __forceinline__ double wave_shr1(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float wave_shr1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

// Value of lane j (a compile-time constant) as a wave-uniform scalar.
__device__ __forceinline__ double lane_value(double v, int j)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// Value of lane src (per-lane index) through the LDS crossbar.
__device__ __forceinline__ double lane_fetch(double v, int src)
{
    const int lo = __builtin_amdgcn_ds_bpermute(src * 4, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src * 4, __double2hiint(v));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_fetch(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src * 4, __float_as_int(v)));
}

// Lateral record of one y-step of a wave that owns one cell column (R = 64):
// lane j holds element j, loaded one step ahead with a single vector load.
enum : int {
    REC_GEOM = 0,       // 20: geometry of cell (xc, yc), wedge 0 then wedge 1
    REC_NT0 = 20,       // 3: unit normal of node (xc, yc+1)
    REC_NT1 = 23,       // 3: unit normal of node (xc+1, yc+1)
    REC_NOUT = 26,      // 3: unit normal of the output node (xc+1, yc)
    REC_BTAB = 29,      // 6: boundary rows of the output node (CMB 3, surface 3)
    REC_UCMB = 35,      // 3: src of the output node at the CMB level
    REC_USURF = 38,     // 3: src of the output node at the surface level
    REC_LEN = 41,
    // fused Chebyshev step: b, invd, d of the output node at the surface level
    REC_SB = 41,
    REC_SI = 44,
    REC_SD = 47,
    REC_LEN_FUSE = 50
};

// Raw node values of one node column at the K+1 levels spanned by a lane.
template <typename T, int K> struct RawCol {
    T u[K + 1][3];
    T k[K + 1];
};

// Processed node column: per cell of the lane sums and differences of the
// bottom and top values (after the boundary modification of the trial side).
template <typename T, int K> struct NodeCol {
    T su[K][3];
    T du[K][3];
    T kk[K];
};

template <typename T, int K>
__device__ __forceinline__ void load_raw(const ApplyArgs<T> &a, int64_t base, int lvl0, RawCol<T, K> &o)
{
#pragma unroll
    for (int q = 0; q <= K; ++q) {
        const int64_t i = base + lvl0 + q;
        o.u[q][0] = a.src0[i];
        o.u[q][1] = a.src1[i];
        o.u[q][2] = a.src2[i];
        o.k[q] = a.k[i];
    }
}

// Per-lane boundary factors of the trial side: the bottom value of the
// lane's first cell (level 0 at the CMB) and the top value of its last cell
// (level nr-1 at the surface).  Dirichlet: multiply by z (0 at the boundary),
// free-slip: remove f * (n.u) n (f = 1 at the boundary).
template <typename T> struct LaneBC {
    T z_lo, f_lo, z_hi, f_hi;
};

constexpr int BCM_RUNTIME = 7;

template <typename T, int K, int BCM>
__device__ __forceinline__ void process(const RawCol<T, K> &r, const T n[3], const LaneBC<T> &bc, bool any_fs,
                                        bool any_dir, NodeCol<T, K> &o)
{
    // BCM: bit 0 free-slip, bit 1 Dirichlet compiled in and applied to every
    // lane (the factors are neutral away from the boundary); BCM_RUNTIME:
    // decided per subdomain
    constexpr bool rt = BCM == BCM_RUNTIME;
    T lo[3] = {r.u[0][0], r.u[0][1], r.u[0][2]};
    T hi[3] = {r.u[K][0], r.u[K][1], r.u[K][2]};
    if (rt ? any_fs : (BCM & 1) != 0) {
        // explicit fused products: the rounding does not depend on what the
        // compiler contracts in a given instantiation
#pragma clang fp contract(off)
        const T dl = bc.f_lo * fma(n[2], lo[2], fma(n[1], lo[1], n[0] * lo[0]));
        const T dh = bc.f_hi * fma(n[2], hi[2], fma(n[1], hi[1], n[0] * hi[0]));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[c] = fma(-dl, n[c], lo[c]);
            hi[c] = fma(-dh, n[c], hi[c]);
        }
    }
    if (rt ? any_dir : (BCM & 2) != 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lo[c] *= bc.z_lo;
            hi[c] *= bc.z_hi;
        }
    }
#pragma unroll
    for (int p = 0; p < K; ++p) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T b = p == 0 ? lo[c] : r.u[p][c];
            const T t = p == K - 1 ? hi[c] : r.u[p + 1][c];
            o.su[p][c] = b + t;
            o.du[p][c] = t - b;
        }
        o.kk[p] = r.k[p] + r.k[p + 1];
    }
}

// One wedge: corners (P0, P1, P2) in wedge order.  Returns the lateral parts
// A0 (corner P1), A1 (corner P2), -(A0+A1) (corner P0) and the radial part Cc
// (subtracted at the bottom node, added at the top node of every corner).
//
// The trial gradient is formed without its lateral factor hs = 1/(2 r_mid):
// G' = grad u / hs = (s1-s0) M0^T + (s2-s0) M1^T + q (d0+d1+d2) M2^T with
// q = (1/(6h)) / hs, and hs is folded into the stress scale
// (cvol = hs^2 * r_mid^2 h / 6 per lane).
template <typename T>
__device__ __forceinline__ void wedge_stress(const T *g /*M0,M1,M2,detL*/, const T su0[3], const T su1[3],
                                             const T su2[3], const T du0[3], const T du1[3], const T du2[3],
                                             T kksum, T q, T cvol, T A0[3], T A1[3], T Cc[3])
{
    T e1[3], e2[3], D[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        e1[d] = su1[d] - su0[d];
        e2[d] = su2[d] - su0[d];
        D[d] = q * (du0[d] + du1[d] + du2[d]);
    }
    const T *M0 = g, *M1 = g + 3, *M2 = g + 6;
    T G[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int e = 0; e < 3; ++e) G[d][e] = e1[d] * M0[e] + e2[d] * M1[e] + D[d] * M2[e];
    const T div = G[0][0] + G[1][1] + G[2][2];
    // f = hs^2 * k_eval * |det J| = hs^2 * (kksum / 6) * |det L| * r_mid^2 h
    const T f = cvol * kksum * g[9];
    const T f2 = T(2) * f;
    const T fd = T(-2.0 / 3.0) * f * div;
    T tau[3][3];
    tau[0][0] = f2 * G[0][0] + fd;
    tau[1][1] = f2 * G[1][1] + fd;
    tau[2][2] = f2 * G[2][2] + fd;
    tau[0][1] = tau[1][0] = f * (G[0][1] + G[1][0]);
    tau[0][2] = tau[2][0] = f * (G[0][2] + G[2][0]);
    tau[1][2] = tau[2][1] = f * (G[1][2] + G[2][1]);
    const T qM2[3] = {q * M2[0], q * M2[1], q * M2[2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        A0[d] = tau[d][0] * M0[0] + tau[d][1] * M0[1] + tau[d][2] * M0[2];
        A1[d] = tau[d][0] * M1[0] + tau[d][1] * M1[1] + tau[d][2] * M1[2];
        Cc[d] = tau[d][0] * qM2[0] + tau[d][1] * qM2[1] + tau[d][2] * qM2[2];
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// fs/fs penalty axis b = eps G^{-1} m, m the sum of nb per-block moments
// (3 doubles each), in a fixed order: threads 0..255 take every 256th
// partial, a wave tree, then the four waves in order.  Every thread of the
// block calls it (two barriers); out (3 doubles) is valid for all on return.
__device__ __forceinline__ void penalty_combine(const double *partials, int nb, double eps, const double *ginv,
                                                double *red /* 12, shared */, double *out)
{
#pragma clang fp contract(off)
    double s[3] = {0.0, 0.0, 0.0};
    if (threadIdx.x < 256)
        for (int b = threadIdx.x; b < nb; b += 256)
            for (int c = 0; c < 3; ++c) s[c] += partials[3 * b + c];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int c = 0; c < 3; ++c) s[c] = wave_sum(s[c]);
    if (lane == 0 && wave < 4)
        for (int c = 0; c < 3; ++c) red[3 * wave + c] = s[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        double m[3] = {0.0, 0.0, 0.0};
        for (int w = 0; w < 4; ++w)
            for (int c = 0; c < 3; ++c) m[c] += red[3 * w + c];
        for (int r = 0; r < 3; ++r)
            out[r] = eps * fma(ginv[3 * r + 2], m[2], fma(ginv[3 * r + 1], m[1], ginv[3 * r] * m[0]));
    }
    __syncthreads();
}

// The Chebyshev sub-step of one node copy, with the rounding of the separate
// update kernel (no contraction: every product is rounded before it is added).
// in: b, invd, d (if read), x at the node.
template <typename T> __device__ __forceinline__ void cheb_load(const ApplyArgs<T> &a, int64_t i, T in[4][3])
{
    const T *cb[3] = {a.cb0, a.cb1, a.cb2}, *ci[3] = {a.ci0, a.ci1, a.ci2}, *xs[3] = {a.src0, a.src1, a.src2};
    const T *cd[3] = {a.cd0, a.cd1, a.cd2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        in[0][c] = cb[c][i];
        in[1][c] = ci[c][i];
        in[2][c] = a.cread_d ? cd[c][i] : T(0);
        in[3][c] = xs[c][i];
    }
}

template <typename T>
__device__ __forceinline__ void cheb_update(const ApplyArgs<T> &a, int64_t i, const T val[3], const T in[4][3])
{
#pragma clang fp contract(off)
    T *cd[3] = {a.cd0, a.cd1, a.cd2}, *cx[3] = {a.cx0, a.cx1, a.cx2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T r = in[0][c] - val[c];
        const T z = r * in[1][c];
        T t = a.caz * z;
        if (a.cread_d) t = fma(a.cad, in[2][c], t);
        cd[c][i] = t;
        cx[c][i] = in[3][c] + t;
    }
}

template <typename T, int R, int K, int NW, int RG, int BCM, bool FUSE, bool SPLIT>
__global__ __launch_bounds__(NW * 64) void epsdivdiv_apply_kernel(const ApplyArgs<T> a)
{
    // RG > 1: the radial cells of a column are split into RG chunks of 64,
    // one wave each (waves hc * VX + v); the node between two chunks is
    // output by the upper chunk, which receives the lower chunk's partial
    // sums through LDS.
    static_assert(RG == 1 || (R == 64 && K == 1), "radial chunks need one cell per lane");
    static_assert(!SPLIT || (R == 64 && K == 1 && !FUSE), "radial parts need the one-column kernels");
    const int rsplit = SPLIT ? a.rsplit : 1;  // radial parts per subdomain
    constexpr int VPW = 64 / R;
    constexpr int VX = NW / RG * VPW;
    constexpr int NV = VX * RG;
    constexpr int KL = K + 1;
    // left-corner contributions of every cell for rows yc (corner 00) and
    // yc+1 (corner 01), lateral part L and radial part C apart, per virtual
    // column, double-buffered by step parity (one cell per lane; more cells
    // per lane use one buffer and a second barrier); the last cell's own L
    // and C of a lower chunk's top lane (before the lateral exchange)
    constexpr int NB = K == 1 ? 2 : 1;
    __shared__ T xch[NB][NV][2][2][K][3][R];
    __shared__ T xtop[NB][RG > 1 ? NV : 1][2][3];

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int vw = lane / R, rl = lane % R;
    const int v = RG > 1 ? wave % VX : wave * VPW + vw;  // virtual column within the strip
    const int hc = RG > 1 ? wave / VX : 0;               // radial chunk
    const int vi = hc * VX + v;                          // LDS slot

    // Blocks are dispatched round-robin over the 8 XCDs; give every XCD a
    // contiguous range of logical blocks so that neighbouring strips and
    // segments (which share node columns) meet in one L2.
    int b;
    {
        const int nb = gridDim.x, x = blockIdx.x & 7, i = blockIdx.x >> 3, q = nb >> 3, r = nb & 7;
        b = x * q + min(x, r) + i;
    }
    const int nx = a.nx, nc = nx - 1, nr = a.nr, ncells = a.ncells;
    // Work: the node rows Y = 0 .. nx-1 of every lateral group (strip), one
    // strip after the other.  BAL: workgroup b sweeps an equal share [q, q1)
    // of them, so the whole grid runs in one round, and a share that crosses
    // into the next strip starts a new sweep there with its own warm-up row.
    // Otherwise one segment of seg_len rows of one strip per workgroup (the
    // loop costs registers that FP32 and the fused kernels do not have).
    constexpr bool BAL = sizeof(T) == 8 && !FUSE;
    int64_t q, q1;
    if constexpr (BAL) {
        const int64_t rows_total = (int64_t)a.groups * nx;
        q = rows_total * b / gridDim.x;
        q1 = rows_total * (b + 1) / gridDim.x;
    } else {
        const int gb = b / a.segs, sb = b % a.segs;
        q = (int64_t)gb * nx + sb * a.seg_len;
        q1 = (int64_t)gb * nx + min(nx, (sb + 1) * a.seg_len);
    }
    const bool pen_on = a.pen != nullptr || a.pen_part != nullptr;
    __shared__ double pen_red[12], pen_b[3];
    if (a.pen_part) penalty_combine(a.pen_part, a.pen_nb, a.pen_eps, a.pen_ginv, pen_red, pen_b);
    do {
        const int g = (int)(q / nx);
        const int Y0 = (int)(q - (int64_t)g * nx);
        const int Yend = (int)min((int64_t)nx, Y0 + (q1 - q));
        q += Yend - Y0;
        int sd, xc, part = 0;
        bool col_ok = true;  // the virtual column exists
        if constexpr (R == 64) {
            // One wave per column: strips run over the virtual columns of all
            // subdomains in a row (xc = -1 .. nc-1 per subdomain; xc = -1 has no
            // cells and only outputs node column 0), so only the last strip is
            // ragged.  A neighbour in the next subdomain has no cells there and
            // passes zeros.
            const int nvc = a.nsd * rsplit * nx;
            const int vcol = g * (VX - 1) + v;
            col_ok = vcol < nvc;
            const int vc = min(vcol, nvc - 1);
            const int psd = vc / nx;
            sd = SPLIT ? psd / rsplit : psd;
            part = psd - sd * rsplit;
            // past the last subdomain: a column without cells (zero contributions)
            xc = col_ok ? vc - psd * nx - 1 : nc;
        } else {
            const int strip = g % a.strips;
            sd = g / a.strips;
            xc = strip * (VX - 1) - 1 + v;
        }
        const int X = xc + 1;
        const bool has_cmb = a.sd_cmb[sd] != 0 && part == 0;
        const bool has_surf = a.sd_surf[sd] != 0 && part == rsplit - 1;
        const bool any_fs = (has_cmb && a.bc_cmb == TNG_FREESLIP) || (has_surf && a.bc_surf == TNG_FREESLIP);
        const bool any_dir = (has_cmb && a.bc_cmb == TNG_DIRICHLET) || (has_surf && a.bc_surf == TNG_DIRICHLET);

        // lanes: cells c0 + rl*K .. +K-1 (nch is either <= R with K = 1, or R*K)
        const int nch = ncells / (RG * rsplit);
        const int c0 = (part * RG + hc) * nch;
        const bool lane_has = rl * K < nch;
        const bool ovf_lane = rl * K + K == nch;  // lane holding the last cell of the chunk
        const bool bot_chunk = hc == 0, top_chunk = hc == RG - 1;
        const int lvl0 = c0 + (lane_has ? rl * K : nch - K);  // clamped first level (loads stay in bounds)
        LaneBC<T> lbc;
        lbc.z_lo = (has_cmb && bot_chunk && rl == 0 && a.bc_cmb == TNG_DIRICHLET) ? T(0) : T(1);
        lbc.f_lo = (has_cmb && bot_chunk && rl == 0 && a.bc_cmb == TNG_FREESLIP) ? T(1) : T(0);
        lbc.z_hi = (has_surf && top_chunk && ovf_lane && a.bc_surf == TNG_DIRICHLET) ? T(0) : T(1);
        lbc.f_hi = (has_surf && top_chunk && ovf_lane && a.bc_surf == TNG_FREESLIP) ? T(1) : T(0);

        T qq[K], cv[K];
    #pragma unroll
        for (int p = 0; p < K; ++p) {
            const int l = lvl0 + p;
            const T r0 = a.radii[(int64_t)sd * nr + l], r1 = a.radii[(int64_t)sd * nr + l + 1];
            const T rmid = T(0.5) * (r0 + r1), h = T(0.5) * (r1 - r0);
            const T hs = T(0.5) / rmid, tt = T(1) / (T(6) * h);
            qq[p] = tt / hs;
            cv[p] = lane_has ? hs * hs * rmid * rmid * h * T(1.0 / 6.0) : T(0);
        }

        auto colbase = [&](int Xq, int Yq) -> int64_t {
            const int xq = min(max(Xq, 0), nx - 1), yq = min(max(Yq, 0), nx - 1);
            return (((int64_t)sd * nx + xq) * nx + yq);
        };
        auto load_normal = [&](int Xq, int Yq, T n[3]) {
            const T *c = a.coords + 3 * colbase(Xq, Yq);
            n[0] = c[0];
            n[1] = c[1];
            n[2] = c[2];
        };

        // R == 64: lateral data of the next step, one element per lane
        // Per-lane address of the record element: base + stride * clamp(yc + dy,
        // 0, ymax), fixed for the whole sweep (every lane loads unconditionally
        // from a valid address, which keeps the wait counts of the pipeline exact).
        const T *rec_base = a.latgeom;
        int rec_stride = 0, rec_dy = 0, rec_ymax = 0;
        if constexpr (R == 64) {
            const int j = lane;
            const int xcc = min(max(xc, 0), nc - 1);
            const int xq0 = min(max(xc, 0), nx - 1), xq1 = min(max(xc + 1, 0), nx - 1);
            const int64_t col0 = ((int64_t)sd * nx + xq0) * nx, col1 = ((int64_t)sd * nx + xq1) * nx;
            if (j < REC_NT0) {
                rec_base = a.latgeom + ((int64_t)sd * nc + xcc) * nc * 20 + j;
                rec_stride = 20, rec_ymax = nc - 1;
            } else if (j < REC_NT1) {
                rec_base = a.coords + 3 * col0 + (j - REC_NT0);
                rec_stride = 3, rec_dy = 1, rec_ymax = nx - 1;
            } else if (j < REC_NOUT) {
                rec_base = a.coords + 3 * col1 + (j - REC_NT1);
                rec_stride = 3, rec_dy = 1, rec_ymax = nx - 1;
            } else if (j < REC_BTAB) {
                rec_base = a.coords + 3 * col1 + (j - REC_NOUT);
                rec_stride = 3, rec_ymax = nx - 1;
            } else if (j < REC_UCMB) {
                if (a.btab) {
                    rec_base = a.btab + 6 * col1 + (j - REC_BTAB);
                    rec_stride = 6, rec_ymax = nx - 1;
                }
            } else if (j < REC_LEN) {
                const int c = (j - REC_UCMB) % 3;
                const T *sc = c == 0 ? a.src0 : (c == 1 ? a.src1 : a.src2);
                rec_base = sc + col1 * nr + (j < REC_USURF ? 0 : nr - 1);
                rec_stride = nr, rec_ymax = nx - 1;
            } else if (FUSE && j < REC_LEN_FUSE) {
                const int c = (j - REC_SB) % 3, w = (j - REC_SB) / 3;
                const T *vc;  // selects, not a table: a table indexed per lane lives in scratch
                if (w == 0) vc = c == 0 ? a.cb0 : (c == 1 ? a.cb1 : a.cb2);
                else if (w == 1) vc = c == 0 ? a.ci0 : (c == 1 ? a.ci1 : a.ci2);
                else vc = c == 0 ? a.cd0 : (c == 1 ? a.cd1 : a.cd2);
                rec_base = vc + col1 * nr + nr - 1;
                rec_stride = nr, rec_ymax = nx - 1;
            }
        }
        auto load_record = [&](int yc) -> T {
            const int y = min(max(yc + rec_dy, 0), rec_ymax);
            // lanes past the record load from a valid address and are never read
            return rec_base[(int64_t)y * rec_stride];
        };

        NodeCol<T, K> B0, B1;
        RawCol<T, K> P0, P1;
        T rec = T(0);
        {
            T n0[3] = {T(0), T(0), T(0)}, n1[3] = {T(0), T(0), T(0)};
            load_raw<T, K>(a, colbase(xc, Y0 - 1) * nr, lvl0, P0);
            load_raw<T, K>(a, colbase(xc + 1, Y0 - 1) * nr, lvl0, P1);
            if (any_fs) {
                load_normal(xc, Y0 - 1, n0);
                load_normal(xc + 1, Y0 - 1, n1);
            }
            process<T, K, BCM>(P0, n0, lbc, any_fs, any_dir, B0);
            process<T, K, BCM>(P1, n1, lbc, any_fs, any_dir, B1);
            load_raw<T, K>(a, colbase(xc, Y0) * nr, lvl0, P0);
            load_raw<T, K>(a, colbase(xc + 1, Y0) * nr, lvl0, P1);
            if constexpr (R == 64) rec = load_record(Y0 - 1);
        }

        // loop invariants of the output phase
        T pb[3] = {T(0), T(0), T(0)};
        if (a.pen_part) {
            pb[0] = (T)pen_b[0];
            pb[1] = (T)pen_b[1];
            pb[2] = (T)pen_b[2];
        } else if (a.pen) {
            pb[0] = (T)a.pen[0];
            pb[1] = (T)a.pen[1];
            pb[2] = (T)a.pen[2];
        }
        T radl[KL];
    #pragma unroll
        for (int p = 0; p < KL; ++p) radl[p] = a.radii[(int64_t)sd * nr + min(c0 + rl * K + p, nr - 1)];

        // Accumulators per cell of the lane: lateral part L and radial part C of
        // every contribution to the cell's corner column of the output row.
        // Node level l of the column receives L - C from cell l (as its bottom)
        // and L + C from cell l - 1 (as its top); the store forms that sum once.
        T aLA[K][3], aCA[K][3], aLB[K][3], aCB[K][3];
    #pragma unroll
        for (int i = 0; i < K; ++i)
    #pragma unroll
            for (int c = 0; c < 3; ++c) aLA[i][c] = aCA[i][c] = aLB[i][c] = aCB[i][c] = T(0);

        // One y-step: bottom rows B0/B1 (processed), the prefetched raw row yc+1
        // becomes the top rows T0/T1; (aL_cur, aC_cur) collect node row yc,
        // (aL_next, aC_next) row yc+1.  The caller alternates the roles of the
        // two row pairs and the two accumulator pairs, so nothing is copied
        // between steps.
        auto do_step = [&](int step, NodeCol<T, K> &B0, NodeCol<T, K> &B1, NodeCol<T, K> &T0, NodeCol<T, K> &T1,
                           T (&aL_cur)[K][3], T (&aC_cur)[K][3], T (&aL_next)[K][3], T (&aC_next)[K][3], T &rec,
                           T &rec_next) {
            const int yc = Y0 - 1 + step;
            // one cell per lane and column per wave: the cell's geometry is
            // wave-uniform; scalar loads issued here have the processing of the
            // new row to arrive (more cells per lane run out of scalar registers
            // and take it from the record instead)
            T sg[20];
            if constexpr (R == 64 && K == 1 && TNG_APPLY_SMEM_GEOM) {
                using CT = const T __attribute__((address_space(4)));
                const int xcc = min(max(xc, 0), nc - 1), ycc = min(max(yc, 0), nc - 1);
                const int gi = __builtin_amdgcn_readfirstlane(((sd * nc + xcc) * nc + ycc) * 20);
                const CT *gw = (const CT *)a.latgeom + gi;
    #pragma unroll
                for (int i = 0; i < 20; ++i) sg[i] = gw[i];
            }
            // fused Chebyshev step.  FP64: the inputs of this step's output node
            // are requested ahead of the next row's prefetch, so that waiting for
            // them in the store phase does not wait for the prefetch (FP32 loads
            // them in the store phase: the registers would halve its occupancy).
            // The surface level comes with the record.
            constexpr bool early = FUSE && sizeof(T) == 8;
            T cin[4][3];
            bool fuse_here = false;
            if constexpr (FUSE) {
                fuse_here = X > 0 && X < nx - 1 && yc > 0 && yc < nx - 1;
                if (early && fuse_here)
                    cheb_load<T>(a, (((int64_t)sd * nx + X) * nx + yc) * nr + min(c0 + rl * K, nr - 1), cin);
            }
            // the prefetched row yc+1 becomes the top row; prefetch row yc+2
            {
                T n0[3] = {T(0), T(0), T(0)}, n1[3] = {T(0), T(0), T(0)};
                if (BCM == BCM_RUNTIME ? any_fs : (BCM & 1) != 0) {
                    if constexpr (R == 64) {
    #pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            n0[c] = lane_value(rec, REC_NT0 + c);
                            n1[c] = lane_value(rec, REC_NT1 + c);
                        }
                    } else {
                        load_normal(xc, yc + 1, n0);
                        load_normal(xc + 1, yc + 1, n1);
                    }
                }
                process<T, K, BCM>(P0, n0, lbc, any_fs, any_dir, T0);
                process<T, K, BCM>(P1, n1, lbc, any_fs, any_dir, T1);
            }
            if constexpr (R == 64) rec_next = load_record(yc + 1);
            load_raw<T, K>(a, colbase(xc, yc + 2) * nr, lvl0, P0);
            load_raw<T, K>(a, colbase(xc + 1, yc + 2) * nr, lvl0, P1);

            const int buf = NB == 2 ? (step & 1) : 0;
            const bool cell_ok = xc >= 0 && xc < nc && yc >= 0 && yc < nc;
            {
                T g0[10], g1[10];
                if constexpr (R == 64 && K == 1 && TNG_APPLY_SMEM_GEOM) {
    #pragma unroll
                    for (int i = 0; i < 10; ++i) {
                        g0[i] = sg[i];
                        g1[i] = sg[10 + i];
                    }
                } else if constexpr (R == 64) {
    #pragma unroll
                    for (int i = 0; i < 10; ++i) {
                        g0[i] = lane_value(rec, REC_GEOM + i);
                        g1[i] = lane_value(rec, REC_GEOM + 10 + i);
                    }
                } else {
                    const int xcc = min(max(xc, 0), nc - 1), ycc = min(max(yc, 0), nc - 1);
                    const T *gw = a.latgeom + ((((int64_t)sd * nc + xcc) * nc + ycc) * 2) * 10;
    #pragma unroll
                    for (int i = 0; i < 10; ++i) {
                        g0[i] = gw[i];
                        g1[i] = gw[10 + i];
                    }
                }
    #pragma unroll
                for (int p = 0; p < K; ++p) {
                    const T cvp = cell_ok ? cv[p] : T(0);
                    T A0[3], A1[3], Cc[3], E0[3], E1[3], Cd[3];
                    // wedge 0: (c00, c10, c01)
                    wedge_stress<T>(g0, B0.su[p], B1.su[p], T0.su[p], B0.du[p], B1.du[p], T0.du[p],
                                    B0.kk[p] + B1.kk[p] + T0.kk[p], qq[p], cvp, A0, A1, Cc);
                    // wedge 1: (c11, c01, c10)
                    wedge_stress<T>(g1, T1.su[p], T0.su[p], B1.su[p], T1.du[p], T0.du[p], B1.du[p],
                                    T1.kk[p] + T0.kk[p] + B1.kk[p], qq[p], cvp, E0, E1, Cd);
                    // corners: c00 and c11 get one wedge each, c10 and c01 both.
                    // The own corners (c10: row yc, c11: row yc+1) go to the
                    // accumulators, the left corners (c00, c01) to the neighbour
                    // v-1 through LDS, L and C apart.
    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const T x00 = -(A0[c] + A1[c]), x11 = -(E0[c] + E1[c]);
                        const T l10 = A0[c] + E1[c], l01 = A1[c] + E0[c], c2 = Cc[c] + Cd[c];
                        aL_cur[p][c] += l10;
                        aC_cur[p][c] += c2;
                        aL_next[p][c] += x11;
                        aC_next[p][c] += Cd[c];
                        xch[buf][vi][0][0][p][c][rl] = x00;
                        xch[buf][vi][0][1][p][c][rl] = Cc[c];
                        xch[buf][vi][1][0][p][c][rl] = l01;
                        xch[buf][vi][1][1][p][c][rl] = c2;
                    }
                }
            }
            if constexpr (RG > 1) {
                // the lower chunk's top lane: its last cell's own L and C
                if (ovf_lane && !top_chunk) {
    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        xtop[buf][vi][0][c] = aL_cur[K - 1][c];
                        xtop[buf][vi][1][c] = aC_cur[K - 1][c];
                    }
                }
            }
            __syncthreads();
            T below[3] = {T(0), T(0), T(0)};
            if constexpr (RG > 1) {
                // lane 0 of an upper chunk: the top level of the lower chunk's
                // last cell, with the neighbour's part added as that lane adds it
                if (rl == 0 && !bot_chunk && v + 1 < VX) {
    #pragma unroll
                    for (int c = 0; c < 3; ++c)
                        below[c] = (xtop[buf][vi - VX][0][c] + xch[buf][vi - VX + 1][0][0][K - 1][c][R - 1]) +
                                   (xtop[buf][vi - VX][1][c] + xch[buf][vi - VX + 1][0][1][K - 1][c][R - 1]);
                }
            }
            if (v + 1 < VX) {
    #pragma unroll
                for (int c = 0; c < 3; ++c) {
    #pragma unroll
                    for (int p = 0; p < K; ++p) {
                        aL_cur[p][c] += xch[buf][vi + 1][0][0][p][c][rl];
                        aC_cur[p][c] += xch[buf][vi + 1][0][1][p][c][rl];
                        aL_next[p][c] += xch[buf][vi + 1][1][0][p][c][rl];
                        aC_next[p][c] += xch[buf][vi + 1][1][1][p][c][rl];
                    }
                }
            }
            if constexpr (NB == 1) __syncthreads();  // the buffer is written again in the next step

            // node row yc is complete: the top level of every cell, L + C, goes
            // to the node above; the first level of the lane takes the lane below's
            T topv[K][3];
    #pragma unroll
            for (int p = 0; p < K; ++p)
    #pragma unroll
                for (int c = 0; c < 3; ++c) topv[p][c] = aL_cur[p][c] + aC_cur[p][c];
    #pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T up = wave_shr1(topv[K - 1][c]);
                if (RG == 1 || rl > 0 || bot_chunk) below[c] = (R == 64 || rl > 0) ? up : T(0);
            }
            const int Y = yc;
            if (Y >= Y0 && v < VX - 1 && X < nx && col_ok && lane_has) {
                const int64_t lcol = ((int64_t)sd * nx + X) * nx + Y;
                const int64_t base = lcol * nr;
                // unit normal of the output node (free-slip rows, penalty)
                T cc[3] = {T(0), T(0), T(0)};
                if (pen_on || (BCM == BCM_RUNTIME ? any_fs : (BCM & 1) != 0)) {
                    if constexpr (R == 64) {
    #pragma unroll
                        for (int c = 0; c < 3; ++c) cc[c] = lane_value(rec, REC_NOUT + c);
                    } else {
                        cc[0] = a.coords[3 * lcol];
                        cc[1] = a.coords[3 * lcol + 1];
                        cc[2] = a.coords[3 * lcol + 2];
                    }
                }
                // boundary row and source of this lane's boundary node (lane
                // 0: CMB, the top lane: surface), fetched while every lane is
                // active: a lane-dependent source lane keeps the fetch one LDS
                // permute instead of VALU broadcasts and selects (FP64 kernels
                // of eight one-column waves without radial chunks or the fused
                // step: elsewhere the extra registers spill or lower the
                // occupancy)
                constexpr bool hoist = R == 64 && !FUSE && sizeof(T) == 8 && NW == 8 && RG == 1;
                T bt_l[3] = {T(0), T(0), T(0)}, u_l[3] = {T(0), T(0), T(0)};
                if constexpr (hoist) {
                    const int off = rl == 0 ? 0 : 3;
    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        bt_l[c] = lane_fetch(rec, REC_BTAB + off + c);
                        u_l[c] = lane_fetch(rec, REC_UCMB + off + c);
                    }
                }
                // free-slip kernels add the penalty term unconditionally: with
                // the penalty off its axis is zero and r (b x n) = 0 leaves val
                // unchanged, while a runtime condition becomes per-lane selects
                constexpr bool pen_uncond = BCM != BCM_RUNTIME && (BCM & 1) != 0;
                T pbn[3] = {T(0), T(0), T(0)};
                if (pen_uncond || pen_on) {
    #pragma clang fp contract(off)
                    pbn[0] = fma(pb[1], cc[2], -(pb[2] * cc[1]));
                    pbn[1] = fma(pb[2], cc[0], -(pb[0] * cc[2]));
                    pbn[2] = fma(pb[0], cc[1], -(pb[1] * cc[0]));
                }
    #pragma unroll
                for (int p = 0; p < KL; ++p) {
                    if (p == K && !(ovf_lane && top_chunk)) continue;
                    const int l = c0 + rl * K + p;
                    T val[3];
    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        if (p == K) val[c] = topv[K - 1][c];
                        else val[c] = (aL_cur[p][c] - aC_cur[p][c]) + (p == 0 ? below[c] : topv[p > 0 ? p - 1 : 0][c]);
                    }
                    // level shared with the radial part above (lower partial)
                    // or below (upper partial): completed by the interface kernel
                    const bool if_lo = SPLIT && p == K && part < rsplit - 1;
                    const bool if_up = SPLIT && p == 0 && rl == 0 && bot_chunk && part > 0;
                    if (if_lo || if_up) {
                        const int64_t plane = (int64_t)a.nsd * nx * nx;
                        T *o = a.iface + ((int64_t)(((if_up ? part - 1 : part) * 2 + (if_up ? 1 : 0)) * 3)) * plane + lcol;
                        o[0] = val[0];
                        o[plane] = val[1];
                        o[2 * plane] = val[2];
                        continue;
                    }
                    const bool at_cmb = p == 0 && rl == 0 && bot_chunk && has_cmb && a.bc_cmb != TNG_NEUMANN;
                    const bool at_surf = p == K && top_chunk && has_surf && a.bc_surf != TNG_NEUMANN;
                    if constexpr (hoist && BCM == 1) {
                        // free slip only: the boundary row on every lane, scaled by
                        // g (1 on the boundary lane, 0 elsewhere) instead of a branch
                        // and its merges.  g = 1 gives the branch's bits, g = 0
                        // leaves val unchanged.
    #pragma clang fp contract(off)
                        const T g = (at_cmb || at_surf) ? T(1) : T(0);
                        const T dv = g * fma(cc[2], val[2], fma(cc[1], val[1], cc[0] * val[0]));
                        const T un = g * (bt_l[0] * fma(cc[2], u_l[2], fma(cc[1], u_l[1], cc[0] * u_l[0])));
    #pragma unroll
                        for (int c = 0; c < 3; ++c) val[c] = fma(un, cc[c], fma(-dv, cc[c], val[c]));
                    } else if (at_cmb || at_surf) {
                        const int bcv = at_cmb ? a.bc_cmb : a.bc_surf;
                        T bt[3], u[3];
                        if constexpr (R == 64 && !hoist) {
                            // lane 0 takes the CMB entries, the top lane the surface ones:
                            // wave-uniform reads of the record lanes (a lane permute here
                            // would read lanes that the branch has switched off)
    #pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                const T btc = lane_value(rec, REC_BTAB + c), bts = lane_value(rec, REC_BTAB + 3 + c);
                                const T uc = lane_value(rec, REC_UCMB + c), us = lane_value(rec, REC_USURF + c);
                                bt[c] = at_cmb ? btc : bts;
                                u[c] = at_cmb ? uc : us;
                            }
                        } else if constexpr (R == 64) {
    #pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                bt[c] = bt_l[c];
                                u[c] = u_l[c];
                            }
                        } else {
                            const T *btp = a.btab + (lcol * 2 + (at_cmb ? 0 : 1)) * 3;
                            bt[0] = btp[0];
                            bt[1] = btp[1];
                            bt[2] = btp[2];
                            u[0] = a.src0[base + l];
                            u[1] = a.src1[base + l];
                            u[2] = a.src2[base + l];
                        }
                        constexpr bool rt = BCM == BCM_RUNTIME, has_d = rt || (BCM & 2), has_f = rt || (BCM & 1);
                        if (has_d && (!has_f || bcv == TNG_DIRICHLET)) {
                            val[0] = bt[0] * u[0];
                            val[1] = bt[1] * u[1];
                            val[2] = bt[2] * u[2];
                        } else if (has_f) {
    #pragma clang fp contract(off)
                            const T dv = fma(cc[2], val[2], fma(cc[1], val[1], cc[0] * val[0]));
                            const T un = bt[0] * fma(cc[2], u[2], fma(cc[1], u[1], cc[0] * u[0]));
    #pragma unroll
                            for (int c = 0; c < 3; ++c) val[c] = fma(un, cc[c], fma(-dv, cc[c], val[c]));
                        }
                    }
                    if (pen_uncond || pen_on) {
    #pragma clang fp contract(off)
                        // b x (r n) = r (b x n), b x n once per output column
                        const T rad = radl[p];
                        val[0] = fma(rad, pbn[0], val[0]);
                        val[1] = fma(rad, pbn[1], val[1]);
                        val[2] = fma(rad, pbn[2], val[2]);
                    }
                    if (a.add) {
                        val[0] += a.dst0[base + l];
                        val[1] += a.dst1[base + l];
                        val[2] += a.dst2[base + l];
                    }
                    if constexpr (FUSE) {
                        // columns off the subdomain's lateral edges are not shared
                        // (every subdomain spans the whole shell radially here)
                        if (fuse_here) {
                            if (p == K) {
                                // the top lane's extra (surface) level
                                T in[4][3];
    #pragma unroll
                                for (int c = 0; c < 3; ++c) {
                                    in[0][c] = lane_value(rec, REC_SB + c);
                                    in[1][c] = lane_value(rec, REC_SI + c);
                                    in[2][c] = a.cread_d ? lane_value(rec, REC_SD + c) : T(0);
                                    in[3][c] = lane_value(rec, REC_USURF + c);
                                }
                                cheb_update<T>(a, base + l, val, in);
                            } else if constexpr (early) {
                                cheb_update<T>(a, base + l, val, cin);
                            } else {
                                T in[4][3];
                                cheb_load<T>(a, base + l, in);
                                cheb_update<T>(a, base + l, val, in);
                            }
                            continue;
                        }
                    }
                    a.dst0[base + l] = val[0];
                    a.dst1[base + l] = val[1];
                    a.dst2[base + l] = val[2];
                }
            }
    #pragma unroll
            for (int i = 0; i < K; ++i)
    #pragma unroll
                for (int c = 0; c < 3; ++c) aL_cur[i][c] = aC_cur[i][c] = T(0);
        };

        const int nsteps = Yend - (Y0 - 1);
        NodeCol<T, K> C0, C1;
        T rec2 = T(0);
        if constexpr (R == 64 && K == 1) {
            // two steps per iteration with alternating roles (no register copies)
            for (int step = 0; step < nsteps; step += 2) {
                do_step(step, B0, B1, C0, C1, aLA, aCA, aLB, aCB, rec, rec2);
                if (step + 1 < nsteps) do_step(step + 1, C0, C1, B0, B1, aLB, aCB, aLA, aCA, rec2, rec);
            }
        } else {
            for (int step = 0; step < nsteps; ++step) {
                do_step(step, B0, B1, C0, C1, aLA, aCA, aLB, aCB, rec, rec2);
                B0 = C0;
                B1 = C1;
                rec = rec2;
    #pragma unroll
                for (int i = 0; i < K; ++i)
    #pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        aLA[i][c] = aLB[i][c];
                        aCA[i][c] = aCB[i][c];
                        aLB[i][c] = aCB[i][c] = T(0);
                    }
            }
        }
        if constexpr (BAL) __syncthreads();  // the next sweep reuses the exchange buffers
    } while (BAL && q < q1);
}

// ----------------------------------------------------------------------------
// Diagonal (per node, gather over the <= 12 adjacent wedges).
//   mode 0: dst = D(src) for every node (Diagonal mode of the reference):
//           D = sum kwJ (|g|^2 + g_d^2/3) per Cartesian component; at
//           free-slip nodes R^T diag(R A_nn R^T) R (reference :2018-2073).
//   mode 1: boundary rows for the matvec: per lateral node and side
//           (0 = CMB at r = 0, 1 = surface at r = nr-1): Dirichlet diagonal
//           (3 values) or free-slip normal-normal entry Ann (value 0).
// ----------------------------------------------------------------------------
template <typename T> struct DiagArgs {
    const T *src0, *src1, *src2;
    T *dst0, *dst1, *dst2;
    const T *k, *radii, *latgeom, *coords;
    T *btab;
    const uint8_t *sd_cmb, *sd_surf;
    int nsd, nx, nr, ncells, bc_cmb, bc_surf, add, mode;
};

// linalg/trafo/local_basis_trafo_normal_tangential.hpp:34-76
template <typename T> __device__ void trafo_nt(const T nin[3], T R[3][3])
{
    T n[3] = {nin[0], nin[1], nin[2]};
    {
        T s = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        T mag = sqrt(s);
        if (mag > T(1e-15))
            for (int i = 0; i < 3; ++i) n[i] /= mag;
    }
    const T ax = fabs(n[0]), ay = fabs(n[1]), az = fabs(n[2]);
    T ref[3] = {T(0), T(0), T(0)};
    if (ax <= ay && ax <= az) ref[0] = T(1);
    else if (ay <= ax && ay <= az) ref[1] = T(1);
    else ref[2] = T(1);
    auto cross = [](const T *u, const T *w, T *o) {
        o[0] = u[1] * w[2] - u[2] * w[1];
        o[1] = u[2] * w[0] - u[0] * w[2];
        o[2] = u[0] * w[1] - u[1] * w[0];
    };
    auto normalize = [](T *v) {
        T s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        T mag = sqrt(s);
        if (mag > T(1e-15))
            for (int i = 0; i < 3; ++i) v[i] /= mag;
    };
    T t1[3], t2[3];
    cross(ref, n, t1);
    normalize(t1);
    const T eps = sizeof(T) == 8 ? T(1e-15) : T(1e-6);
    if (sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]) < eps) {
        t1[0] = n[1]; t1[1] = -n[0]; t1[2] = T(0);
        normalize(t1);
    }
    cross(n, t1, t2);
    normalize(t2);
    for (int i = 0; i < 3; ++i) { R[0][i] = n[i]; R[1][i] = t1[i]; R[2][i] = t2[i]; }
}

template <typename T> __global__ void epsdivdiv_diag_kernel(const DiagArgs<T> a)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = a.nx, nr = a.nr, nc = nx - 1;
    int sd, X, Y, l;
    int side = -1;
    if (a.mode == 0) {
        const int64_t n = (int64_t)a.nsd * nx * nx * nr;
        if (gid >= n) return;
        l = (int)(gid % nr);
        int64_t t = gid / nr;
        Y = (int)(t % nx); t /= nx;
        X = (int)(t % nx);
        sd = (int)(t / nx);
    } else {
        const int64_t n = (int64_t)a.nsd * nx * nx * 2;
        if (gid >= n) return;
        side = (int)(gid % 2);
        int64_t t = gid / 2;
        Y = (int)(t % nx); t /= nx;
        X = (int)(t % nx);
        sd = (int)(t / nx);
        l = side == 0 ? 0 : nr - 1;
    }
    const bool at_cmb = l == 0 && a.sd_cmb[sd];
    const bool at_surf = l == nr - 1 && a.sd_surf[sd];
    const int bc = at_cmb ? a.bc_cmb : (at_surf ? a.bc_surf : TNG_NEUMANN);
    const bool fs = (at_cmb || at_surf) && bc == TNG_FREESLIP;
    const int64_t lcol = ((int64_t)sd * nx + X) * nx + Y;
    T R[3][3] = {{T(1), T(0), T(0)}, {T(0), T(1), T(0)}, {T(0), T(0), T(1)}};
    if (fs) {
        T n[3];
        normal_of(a.coords + 3 * lcol, n);
        trafo_nt(n, R);
    }
    T dd[3] = {T(0), T(0), T(0)}, dr[3] = {T(0), T(0), T(0)};
    // lateral corner index of (dx, dy) in wedge w; -1 if not a corner
    const int corner_of[2][2][2] = {{{0, 2}, {1, -1}}, {{-1, 1}, {2, 0}}};  // [w][dx][dy]
    const int wc[2][3][2] = {{{0, 0}, {1, 0}, {0, 1}}, {{1, 1}, {0, 1}, {1, 0}}};
    for (int dxc = -1; dxc <= 0; ++dxc) {
        const int xc = X + dxc;
        if (xc < 0 || xc >= nc) continue;
        for (int dyc = -1; dyc <= 0; ++dyc) {
            const int yc = Y + dyc;
            if (yc < 0 || yc >= nc) continue;
            for (int w = 0; w < 2; ++w) {
                const int kc = corner_of[w][X - xc][Y - yc];
                if (kc < 0) continue;
                const T *g = a.latgeom + ((((int64_t)sd * nc + xc) * nc + yc) * 2 + w) * 10;
                const T l0 = kc == 0 ? T(-0.5) : (kc == 1 ? T(0.5) : T(0));
                const T l1 = kc == 0 ? T(-0.5) : (kc == 1 ? T(0) : T(0.5));
                for (int rc = l - 1; rc <= l; ++rc) {
                    if (rc < 0 || rc >= a.ncells) continue;
                    const T r0 = a.radii[(int64_t)sd * nr + rc], r1 = a.radii[(int64_t)sd * nr + rc + 1];
                    const T rmid = T(0.5) * (r0 + r1), h = T(0.5) * (r1 - r0);
                    const T s = T(1) / rmid;
                    const T dz = (rc == l) ? T(-1.0 / 6.0) : T(1.0 / 6.0);
                    T gv[3];
                    for (int d = 0; d < 3; ++d) gv[d] = s * (l0 * g[d] + l1 * g[3 + d]) + (dz / h) * g[6 + d];
                    T ks = T(0);
                    for (int q = 0; q < 3; ++q) {
                        const int64_t cb = (((int64_t)sd * nx + xc + wc[w][q][0]) * nx + yc + wc[w][q][1]) * nr;
                        ks += a.k[cb + rc] + a.k[cb + rc + 1];
                    }
                    const T kwJ = ks * T(1.0 / 6.0) * g[9] * rmid * rmid * h;
                    const T gg = gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2];
                    for (int d = 0; d < 3; ++d) dd[d] += kwJ * (gg + T(1.0 / 3.0) * gv[d] * gv[d]);
                    if (fs)
                        for (int c = 0; c < 3; ++c) {
                            const T rg = R[c][0] * gv[0] + R[c][1] * gv[1] + R[c][2] * gv[2];
                            dr[c] += kwJ * (gg + T(1.0 / 3.0) * rg * rg);
                        }
                }
            }
        }
    }
    if (a.mode == 1) {
        T *o = a.btab + (lcol * 2 + side) * 3;
        if (fs) {
            o[0] = dr[0]; o[1] = T(0); o[2] = T(0);
        } else {
            o[0] = dd[0]; o[1] = dd[1]; o[2] = dd[2];
        }
        return;
    }
    const int64_t i = lcol * nr + l;
    const T s0 = a.src0[i], s1 = a.src1[i], s2 = a.src2[i];
    T out[3];
    if (fs) {
        T rs[3];
        for (int c = 0; c < 3; ++c) rs[c] = dr[c] * (R[c][0] * s0 + R[c][1] * s1 + R[c][2] * s2);
        for (int d = 0; d < 3; ++d) out[d] = R[0][d] * rs[0] + R[1][d] * rs[1] + R[2][d] * rs[2];
    } else {
        out[0] = dd[0] * s0; out[1] = dd[1] * s1; out[2] = dd[2] * s2;
    }
    if (a.add) {
        out[0] += a.dst0[i]; out[1] += a.dst1[i]; out[2] += a.dst2[i];
    }
    a.dst0[i] = out[0]; a.dst1[i] = out[1]; a.dst2[i] = out[2];
}

// ----------------------------------------------------------------------------
// fs/fs rotation penalty, first half: m = sum_OWNED x cross u, then
// b = eps * G^{-1} m (G = Gram matrix of the rotation fields).  Two launches:
// per-block partials, then one block sums them in a fixed order, so the
// result is deterministic.
// ----------------------------------------------------------------------------
template <typename T> struct PenArgs {
    const T *src0, *src1, *src2;
    const T *coords, *radii;
    const uint8_t *omask;
    double *partials;
    double *out_b;
    double eps;
    double ginv[9];
    int nsd, nx, nr;
    int64_t n;
    FastDiv div_nr, div_nxx;  // node copy -> column, column -> subdomain
};

// Rotation moments m = sum over OWNED copies of x (x) u (x = radius * unit
// coordinate): per-block partials, flat over the node copies (coalesced
// streams of src and the mask), four independent nodes per thread and step.
constexpr int kPenThreads = 512;

template <typename T> __global__ __launch_bounds__(kPenThreads) void penalty_partial_kernel(const PenArgs<T> a)
{
    constexpr int NWV = kPenThreads / 64;
    __shared__ double red[NWV][3];
    double m0 = 0, m1 = 0, m2 = 0;
    const int n = (int)a.n, nr = a.nr;
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = i0 + q * stride;
            if (i < n) {
                const int col = (int)a.div_nr.div((uint32_t)i), l = i - col * nr;
                const int sd = (int)a.div_nxx.div((uint32_t)col);
                const double w = a.omask[i] == OWNED ? 1.0 : 0.0;
                const double rad = (double)a.radii[sd * nr + l] * w;
                const double x0 = (double)a.coords[3 * col] * rad, x1 = (double)a.coords[3 * col + 1] * rad,
                             x2 = (double)a.coords[3 * col + 2] * rad;
                const double u0 = a.src0[i], u1 = a.src1[i], u2 = a.src2[i];
                m0 += x1 * u2 - x2 * u1;
                m1 += x2 * u0 - x0 * u2;
                m2 += x0 * u1 - x1 * u0;
            }
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    m0 = wave_sum(m0); m1 = wave_sum(m1); m2 = wave_sum(m2);
    if (lane == 0) { red[wave][0] = m0; red[wave][1] = m1; red[wave][2] = m2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[3] = {0, 0, 0};
        for (int w = 0; w < NWV; ++w)
            for (int c = 0; c < 3; ++c) s[c] += red[w][c];
        for (int c = 0; c < 3; ++c) a.partials[3 * blockIdx.x + c] = s[c];
    }
}

// One block: the partials in a fixed order, then b = eps G^{-1} m.
template <typename T> __global__ __launch_bounds__(256) void penalty_final_kernel(const PenArgs<T> a, int nb)
{
    __shared__ double red[12], b[3];
    penalty_combine(a.partials, nb, a.eps, a.ginv, red, b);
    if (threadIdx.x < 3) a.out_b[threadIdx.x] = b[threadIdx.x];
}

// The node levels between the radial parts of a split sweep: upper partial +
// lower partial (the order in which a radial chunk adds the chunk below it),
// then the penalty term and the Add mode as in the apply's store phase.  No
// boundary rows: parts meet strictly inside the shell.
template <typename T> __global__ __launch_bounds__(256) void epsdivdiv_iface_kernel(const ApplyArgs<T> a)
{
#pragma clang fp contract(off)
    __shared__ double pen_red[12], pen_b[3];
    if (a.pen_part) penalty_combine(a.pen_part, a.pen_nb, a.pen_eps, a.pen_ginv, pen_red, pen_b);
    const bool pen_on = a.pen != nullptr || a.pen_part != nullptr;
    T pb[3] = {T(0), T(0), T(0)};
    for (int c = 0; c < 3; ++c) pb[c] = a.pen_part ? (T)pen_b[c] : (a.pen ? (T)a.pen[c] : T(0));
    const int nx = a.nx, nr = a.nr, ncp = a.ncells / a.rsplit;
    const int64_t plane = (int64_t)a.nsd * nx * nx, n = plane * (a.rsplit - 1);
    T *dst[3] = {a.dst0, a.dst1, a.dst2};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(i / plane);  // between parts j and j + 1
        const int64_t lcol = i - j * plane;
        const int sd = (int)(lcol / ((int64_t)nx * nx));
        const int l = (j + 1) * ncp;
        const T *lo = a.iface + (int64_t)(j * 2) * 3 * plane + lcol;
        const T *up = lo + 3 * plane;
        T val[3];
        for (int c = 0; c < 3; ++c) val[c] = up[c * plane] + lo[c * plane];
        if (pen_on) {
            const T cc[3] = {a.coords[3 * lcol], a.coords[3 * lcol + 1], a.coords[3 * lcol + 2]};
            const T pbn[3] = {fma(pb[1], cc[2], -(pb[2] * cc[1])), fma(pb[2], cc[0], -(pb[0] * cc[2])),
                              fma(pb[0], cc[1], -(pb[1] * cc[0]))};
            const T rad = a.radii[(int64_t)sd * nr + l];
            for (int c = 0; c < 3; ++c) val[c] = fma(rad, pbn[c], val[c]);
        }
        const int64_t node = lcol * nr + l;
        for (int c = 0; c < 3; ++c) {
            if (a.add) val[c] += dst[c][node];
            dst[c][node] = val[c];
        }
    }
}

constexpr int kPenBlocks = 1024;

template <typename T> struct MeshPtrs {
    const T *coords, *radii, *latgeom;
};
template <typename T> MeshPtrs<T> mesh_ptrs(const tng_domain &m);
template <> MeshPtrs<double> mesh_ptrs<double>(const tng_domain &m)
{
    return {m.d_coords.as<double>(), m.d_radii.as<double>(), m.d_latgeom.as<double>()};
}
template <> MeshPtrs<float> mesh_ptrs<float>(const tng_domain &m)
{
    return {m.d_coords_f.as<float>(), m.d_radii_f.as<float>(), m.d_latgeom_f.as<float>()};
}

// Workgroups resident at once on the current device for one kernel.
int resident_groups(const void *kernel, int threads)
{
    int dev = 0, cus = 0, per_cu = 0;
    TNG_HIP_CHECK(hipGetDevice(&dev));
    TNG_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    TNG_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    return std::max(1, per_cu) * std::max(1, cus);
}

// One-segment launches: a workgroup sweeps seg rows plus one warm-up row,
// after a prologue of about 1.5 row-steps; the grid runs in
// ceil(blocks / resident) rounds.  Pick the length with the least
// rounds * (seg + 2.5).
int pick_segment(int nx, int64_t lateral_groups, int resident)
{
    int best = nx;
    double best_cost = 1e300;
    for (int seg = std::min(4, nx); seg <= nx; ++seg) {
        const int64_t segs = (nx + seg - 1) / seg;
        const int64_t rounds = (lateral_groups * segs + resident - 1) / resident;
        const double cost = (double)rounds * (seg + 2.5);
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = seg;
        }
    }
    return best;
}

template <typename T, int R, int K, int NW, int RG, int BCM = BCM_RUNTIME, bool FUSE = false, bool SPLIT = false>
void launch_apply_rk(ApplyArgs<T> a, int64_t lateral_groups, hipStream_t s, const tng_tuning &tu)
{
    const void *kern = reinterpret_cast<const void *>(&epsdivdiv_apply_kernel<T, R, K, NW, RG, BCM, FUSE, SPLIT>);
    static int cached_dev = -1, cached_resident = 0;
    int dev = 0;
    TNG_HIP_CHECK(hipGetDevice(&dev));
    if (dev != cached_dev) {
        cached_resident = resident_groups(kern, NW * 64);
        cached_dev = dev;
    }
    a.groups = (int)lateral_groups;
    int blocks;
    if (sizeof(T) == 8 && !FUSE) {
        // one round of resident workgroups, each sweeping an equal share of
        // the node rows (the apply_blocks setting overrides the count)
        const int64_t rows = lateral_groups * a.nx;
        int64_t want = tu.apply_blocks;
        if (want <= 0) want = cached_resident;
        blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, rows));
    } else {
        int seg = tu.apply_seg;
        if (seg <= 0) seg = pick_segment(a.nx, lateral_groups, cached_resident);
        a.seg_len = seg;
        a.segs = (a.nx + seg - 1) / seg;
        blocks = (int)(lateral_groups * a.segs);
    }
    hipLaunchKernelGGL((epsdivdiv_apply_kernel<T, R, K, NW, RG, BCM, FUSE, SPLIT>), dim3(blocks), dim3(NW * 64), 0, s, a);
}

// Shape of the apply launch for a mesh: R lanes per column, K cells per lane,
// RG radial chunks, NW waves per workgroup, S radial parts swept separately.
struct ApplyShape {
    int R, K, RG, NW, S;
};

ApplyShape apply_shape(int ncells, const tng_tuning &tu, bool fused = false)
{
    ApplyShape sh{64, 1, 1, 8, 1};
    const int chunks = tu.apply_radial_chunks;
    const bool split = tu.apply_radial_split != 0;
    if (ncells >= 128 && chunks == 0 && split && !fused) {
        // radial parts of 64 cells swept like separate subdomains (eight
        // one-column waves per workgroup, seven new columns each), their
        // shared levels completed by the interface kernel
        sh.S = ncells / 64;
    } else if (ncells <= 64) {
        sh.R = 4;
        while (sh.R < ncells) sh.R *= 2;
    } else if ((ncells == 128 || ncells == 256) && chunks != 1) {
        // one wave per radial chunk of 64 cells; the K-cells-per-lane kernels
        // (apply_radial_chunks = 1) exceed the register file at K = 4
        sh.RG = ncells / 64;
    } else {
        sh.K = ncells / 64;
        sh.NW = 4;
    }
    if (sh.K == 1 && sh.R == 64 && sh.RG == 1 && sh.S == 1 && tu.apply_waves == 4) sh.NW = 4;
    return sh;
}

template <typename T, bool FUSE = false> void launch_apply(ApplyArgs<T> a, hipStream_t s, const tng_tuning &tu)
{
    TNG_REQUIRE(a.ncells >= 1 && (a.ncells & (a.ncells - 1)) == 0, "radial layers per subdomain must be a power of 2");
    TNG_REQUIRE((int64_t)a.nsd * a.nx * a.nx * a.nr < ((int64_t)1 << 31), "node copies exceed 32-bit indexing");
    // shapes: up to 64 radial cells one cell per lane (R lanes per column);
    // more as radial parts of 64 cells, or (128 and 256 cells) radial chunks
    // of one wave each or K cells per lane
    const ApplyShape sh = apply_shape(a.ncells, tu, FUSE);
    const int R = sh.R, K = sh.K, RG = sh.RG, NW = sh.NW;
    TNG_REQUIRE(K <= 4, "more than 256 radial layers per subdomain need the radial parts");
    a.rsplit = sh.S;
    TNG_REQUIRE(sh.S == 1 || (R == 64 && K == 1 && !FUSE && a.iface), "radial parts need the one-column kernels");
    TNG_REQUIRE(!FUSE || (R == 64 && K == 1 && NW == 8 && RG <= 2),
                "the fused Chebyshev step needs the one-column kernels with at most 128 radial layers");
    const int VX = NW / RG * (64 / R);
    // R = 64: strips over the columns of all subdomains in a row; otherwise
    // strips per subdomain
    a.strips = R == 64 ? (a.nsd * a.rsplit * a.nx + (VX - 1) - 1) / (VX - 1) : (a.nx + (VX - 1) - 1) / (VX - 1);
    const int64_t lg = R == 64 ? (int64_t)a.strips : (int64_t)a.nsd * a.strips;
    // the one-column kernels are compiled per boundary-condition mix
    const int bcm = ((a.bc_cmb == TNG_FREESLIP || a.bc_surf == TNG_FREESLIP) ? 1 : 0) |
                    ((a.bc_cmb == TNG_DIRICHLET || a.bc_surf == TNG_DIRICHLET) ? 2 : 0);
    if (sh.S > 1) {
        if constexpr (FUSE) {
            // unreachable: rejected above
        } else {
            switch (bcm) {
            case 0: launch_apply_rk<T, 64, 1, 8, 1, 0, false, true>(a, lg, s, tu); break;
            case 1: launch_apply_rk<T, 64, 1, 8, 1, 1, false, true>(a, lg, s, tu); break;
            case 2: launch_apply_rk<T, 64, 1, 8, 1, 2, false, true>(a, lg, s, tu); break;
            default: launch_apply_rk<T, 64, 1, 8, 1, 3, false, true>(a, lg, s, tu); break;
            }
        }
    } else if (RG == 4) {
        if constexpr (FUSE) {
            // unreachable: the fused step is limited to one or two chunks
        } else {
            switch (bcm) {
            case 0: launch_apply_rk<T, 64, 1, 8, 4, 0>(a, lg, s, tu); break;
            case 1: launch_apply_rk<T, 64, 1, 8, 4, 1>(a, lg, s, tu); break;
            case 2: launch_apply_rk<T, 64, 1, 8, 4, 2>(a, lg, s, tu); break;
            default: launch_apply_rk<T, 64, 1, 8, 4, 3>(a, lg, s, tu); break;
            }
        }
    } else if (RG == 2) {
        switch (bcm) {
        case 0: launch_apply_rk<T, 64, 1, 8, 2, 0, FUSE>(a, lg, s, tu); break;
        case 1: launch_apply_rk<T, 64, 1, 8, 2, 1, FUSE>(a, lg, s, tu); break;
        case 2: launch_apply_rk<T, 64, 1, 8, 2, 2, FUSE>(a, lg, s, tu); break;
        default: launch_apply_rk<T, 64, 1, 8, 2, 3, FUSE>(a, lg, s, tu); break;
        }
    } else if (K == 1 && R == 64 && NW == 8) {
        switch (bcm) {
        case 0: launch_apply_rk<T, 64, 1, 8, 1, 0, FUSE>(a, lg, s, tu); break;
        case 1: launch_apply_rk<T, 64, 1, 8, 1, 1, FUSE>(a, lg, s, tu); break;
        case 2: launch_apply_rk<T, 64, 1, 8, 1, 2, FUSE>(a, lg, s, tu); break;
        default: launch_apply_rk<T, 64, 1, 8, 1, 3, FUSE>(a, lg, s, tu); break;
        }
    } else if constexpr (FUSE) {
        // unreachable: rejected above
    } else if (K == 1) {
        switch (R) {
        case 4: launch_apply_rk<T, 4, 1, 8, 1>(a, lg, s, tu); break;
        case 8: launch_apply_rk<T, 8, 1, 8, 1>(a, lg, s, tu); break;
        case 16: launch_apply_rk<T, 16, 1, 8, 1>(a, lg, s, tu); break;
        case 32: launch_apply_rk<T, 32, 1, 8, 1>(a, lg, s, tu); break;
        default: launch_apply_rk<T, 64, 1, 4, 1>(a, lg, s, tu); break;
        }
    } else if (K == 2) {
        launch_apply_rk<T, 64, 2, 4, 1>(a, lg, s, tu);
    } else {
        launch_apply_rk<T, 64, 4, 4, 1>(a, lg, s, tu);
    }
    if (a.rsplit > 1) {
        const int64_t n = (int64_t)a.nsd * a.nx * a.nx * (a.rsplit - 1);
        const int blocks = (int)std::min<int64_t>(1024, (n + 255) / 256);
        hipLaunchKernelGGL(epsdivdiv_iface_kernel<T>, dim3(blocks), dim3(256), 0, s, a);
    }
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T> void run_diag_kernel(tng_operator &op, const tng_vector *src, tng_vector *dst, int mode, int add,
                                           hipStream_t s)
{
    const tng_domain &m = *op.mesh;
    DiagArgs<T> a{};
    const MeshPtrs<T> mp = mesh_ptrs<T>(m);
    if (src) {
        a.src0 = (const T *)src->comp(0); a.src1 = (const T *)src->comp(1); a.src2 = (const T *)src->comp(2);
    }
    if (dst) {
        a.dst0 = (T *)dst->comp(0); a.dst1 = (T *)dst->comp(1); a.dst2 = (T *)dst->comp(2);
    }
    a.k = (const T *)op.k->comp(0);
    a.radii = mp.radii; a.latgeom = mp.latgeom; a.coords = mp.coords;
    a.btab = op.d_btab.as<T>();
    a.sd_cmb = m.d_sd_cmb.as<uint8_t>(); a.sd_surf = m.d_sd_surf.as<uint8_t>();
    a.nsd = m.nsd; a.nx = m.nx; a.nr = m.nr; a.ncells = m.nr - 1;
    a.bc_cmb = op.bc_cmb; a.bc_surf = op.bc_surf; a.add = add; a.mode = mode;
    const int64_t n = mode == 0 ? m.nalloc : (int64_t)m.nlat * 2;
    const int bs = 256;
    hipLaunchKernelGGL(epsdivdiv_diag_kernel<T>, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, s, a);
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T> PenArgs<T> pen_args(tng_operator &op, const tng_vector &src);

// Scratch of the allocation-free apply: the penalty's per-block rotation
// moments, then the partial sums of the levels between radial parts.  One
// copy per GPU, resident until the process ends; every entry a call reads
// was written earlier in the same call, so nothing needs clearing.  Calls
// that overlap on one GPU race on it.
constexpr size_t kNoAllocScratchBytes = (size_t)8 << 20;
constexpr size_t kNoAllocPenDoubles = 3 * 1024;  // kPenBlocks x 3, rounded to a 256-byte boundary
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// fs/fs penalty: the per-block rotation moments of src into the context's
// partials (kPenBlocks x 3 doubles)
template <typename T> void penalty_partials(tng_operator &op, const tng_vector &src, hipStream_t s, double *partials)
{
    PenArgs<T> p = pen_args<T>(op, src);
    p.partials = partials;
    hipLaunchKernelGGL(penalty_partial_kernel<T>, dim3(kPenBlocks), dim3(kPenThreads), 0, s, p);
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T> PenArgs<T> pen_args(tng_operator &op, const tng_vector &src)
{
    const tng_domain &m = *op.mesh;
    PenArgs<T> p{};
    const MeshPtrs<T> mp = mesh_ptrs<T>(m);
    p.src0 = (const T *)src.comp(0); p.src1 = (const T *)src.comp(1); p.src2 = (const T *)src.comp(2);
    p.coords = mp.coords; p.radii = mp.radii; p.omask = m.d_omask.as<uint8_t>();
    p.partials = m.ctx->red_partials.as<double>();
    p.out_b = op.d_penalty_b.as<double>();
    p.eps = op.penalty_eps;
    for (int i = 0; i < 9; ++i) p.ginv[i] = m.rot_gram_inv[i];
    p.nsd = m.nsd; p.nx = m.nx; p.nr = m.nr; p.n = m.nalloc;
    p.div_nr = FastDiv((uint32_t)m.nr);
    p.div_nxx = FastDiv((uint32_t)(m.nx * m.nx));
    TNG_REQUIRE(m.nalloc < (int64_t)1 << 31, "penalty moment: node copies exceed 32-bit indexing");
    return p;
}

// fs/fs penalty: the rotation moment b of src into op.d_penalty_b
template <typename T> void penalty_moment(tng_operator &op, const tng_vector &src, hipStream_t s, double *partials)
{
    PenArgs<T> p = pen_args<T>(op, src);
    p.partials = partials;
    hipLaunchKernelGGL(penalty_partial_kernel<T>, dim3(kPenBlocks), dim3(kPenThreads), 0, s, p);
    hipLaunchKernelGGL(penalty_final_kernel<T>, dim3(1), dim3(256), 0, s, p, kPenBlocks);
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename T>
ApplyArgs<T> apply_args(const tng_operator &op, const tng_vector &src, tng_vector &dst, bool pen, bool add)
{
    const tng_domain &m = *op.mesh;
    ApplyArgs<T> a{};
    const MeshPtrs<T> mp = mesh_ptrs<T>(m);
    a.src0 = (const T *)src.comp(0); a.src1 = (const T *)src.comp(1); a.src2 = (const T *)src.comp(2);
    a.dst0 = (T *)dst.comp(0); a.dst1 = (T *)dst.comp(1); a.dst2 = (T *)dst.comp(2);
    a.k = (const T *)op.k->comp(0);
    a.radii = mp.radii; a.latgeom = mp.latgeom; a.coords = mp.coords;
    a.btab = op.d_btab.as<T>();
    a.pen = pen ? op.d_penalty_b.as<double>() : nullptr;
    a.sd_cmb = m.d_sd_cmb.as<uint8_t>(); a.sd_surf = m.d_sd_surf.as<uint8_t>();
    a.nsd = m.nsd; a.nx = m.nx; a.nr = m.nr; a.ncells = m.nr - 1;
    a.bc_cmb = op.bc_cmb; a.bc_surf = op.bc_surf; a.add = add;
    return a;
}

// ws: the scratch array of the allocation-free apply, or nullptr (the
// context's partials and the operator's interface buffer)
template <typename T>
void apply_impl(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s, double *ws = nullptr)
{
    tng_domain &m = *op.mesh;
    tng_context &ctx = *m.ctx;
    const bool add = op.apply_mode == TNG_ADD;
    hipEvent_t *ev = ctx.profiling && !ws ? ctx.prof_next_set() : nullptr;
    if (ev) TNG_HIP_CHECK(hipEventRecord(ev[0], s));
    double *partials = ws ? ws : m.ctx->red_partials.as<double>();
    auto penalty_moment = [&]() { tng::penalty_moment<T>(op, src, s, partials); };
    if (op.stored) {
        // slow path on stored matrices (Dirichlet/free-slip handled per row)
        const bool pen = op.penalty_active && !op.diagonal;
        if (pen) penalty_moment();
        if (ev) TNG_HIP_CHECK(hipEventRecord(ev[1], s));
        stored_apply<T>(op, src, dst, pen ? op.d_penalty_b.as<double>() : nullptr, add, s);
        if (ev) TNG_HIP_CHECK(hipEventRecord(ev[2], s));
    } else if (op.diagonal) {
        if (ev) TNG_HIP_CHECK(hipEventRecord(ev[1], s));
        run_diag_kernel<T>(op, &src, &dst, 0, add, s);
        if (ev) TNG_HIP_CHECK(hipEventRecord(ev[2], s));
    } else {
        if (!op.btab_valid) epsdivdiv_update_boundary_table(op, s);
        // penalty: the per-block moments only; every apply workgroup combines them
        const bool pen = op.penalty_active;
        ApplyArgs<T> a = apply_args<T>(op, src, dst, false, add);
        const ApplyShape sh = apply_shape(m.nr - 1, m.ctx->tuning);
        if (sh.S > 1) {
            // partial sums of the levels between radial parts
            const size_t bytes = sizeof(T) * 6 * (sh.S - 1) * (size_t)m.nsd * m.nx * m.nx;
            if (ws) {
                TNG_REQUIRE(bytes <= kNoAllocScratchBytes - kNoAllocPenDoubles * sizeof(double),
                            "allocation-free apply: the partial sums between radial parts need " +
                                std::to_string(bytes) + " bytes, more than the " +
                                std::to_string(kNoAllocScratchBytes - kNoAllocPenDoubles * sizeof(double)) +
                                " bytes of scratch");
                a.iface = reinterpret_cast<T *>(ws + kNoAllocPenDoubles);
            } else {
                if (op.d_iface.bytes < bytes) op.d_iface.alloc(bytes);
                a.iface = op.d_iface.as<T>();
            }
        }
        if (pen) {
            penalty_partials<T>(op, src, s, partials);
            a.pen_part = partials;
            a.pen_nb = kPenBlocks;
            a.pen_eps = op.penalty_eps;
            for (int i = 0; i < 9; ++i) a.pen_ginv[i] = m.rot_gram_inv[i];
        }
        if (ev) TNG_HIP_CHECK(hipEventRecord(ev[1], s));
        launch_apply<T>(a, s, m.ctx->tuning);
        if (ev) TNG_HIP_CHECK(hipEventRecord(ev[2], s));
    }
    if (op.comm_mode == TNG_COMMUNICATE_ADDITIVELY) launch_halo_sum(m, dst, s);
    if (ev) TNG_HIP_CHECK(hipEventRecord(ev[3], s));
}

// One Chebyshev sub-step with the update fused into the apply: node copies
// that no other subdomain shares are updated in the apply kernel's store
// phase, the shared ones (lateral subdomain edges) after their A x has been
// summed over the copies.  Returns false (nothing launched) where the fused
// kernels do not apply.
template <typename T>
bool apply_cheb_impl(tng_operator &op, const tng_vector &x, tng_vector &ax, tng_vector &x_out, tng_vector &d,
                     const tng_vector &b, const tng_vector &invd, double az, double ad, bool read_d, hipStream_t s)
{
    tng_domain &m = *op.mesh;
    if (op.stored || op.diagonal || op.apply_mode != TNG_REPLACE || op.comm_mode != TNG_COMMUNICATE_ADDITIVELY ||
        m.rad_sdr != 0)
        return false;
    const ApplyShape sh = apply_shape(m.nr - 1, m.ctx->tuning, true);
    if (!(sh.R == 64 && sh.K == 1 && sh.NW == 8 && sh.RG <= 2 && sh.S == 1)) return false;
    tng_context &ctx = *m.ctx;
    hipEvent_t *ev = ctx.profiling ? ctx.prof_next_set() : nullptr;
    if (ev) TNG_HIP_CHECK(hipEventRecord(ev[0], s));
    if (!op.btab_valid) epsdivdiv_update_boundary_table(op, s);
    const bool pen = op.penalty_active;
    if (pen) penalty_moment<T>(op, x, s, m.ctx->red_partials.as<double>());
    ApplyArgs<T> a = apply_args<T>(op, x, ax, pen, false);
    a.cb0 = (const T *)b.comp(0); a.cb1 = (const T *)b.comp(1); a.cb2 = (const T *)b.comp(2);
    a.ci0 = (const T *)invd.comp(0); a.ci1 = (const T *)invd.comp(1); a.ci2 = (const T *)invd.comp(2);
    a.cd0 = (T *)d.comp(0); a.cd1 = (T *)d.comp(1); a.cd2 = (T *)d.comp(2);
    a.cx0 = (T *)x_out.comp(0); a.cx1 = (T *)x_out.comp(1); a.cx2 = (T *)x_out.comp(2);
    a.caz = (T)az;
    a.cad = (T)ad;
    a.cread_d = read_d ? 1 : 0;
    if (ev) TNG_HIP_CHECK(hipEventRecord(ev[1], s));
    launch_apply<T, true>(a, s, m.ctx->tuning);
    if (ev) TNG_HIP_CHECK(hipEventRecord(ev[2], s));
    launch_halo_cheb(m, ax, b, invd, d, x, x_out, az, ad, read_d, s);
    if (ev) TNG_HIP_CHECK(hipEventRecord(ev[3], s));
    return true;
}

}  // namespace

void epsdivdiv_update_boundary_table(tng_operator &op, hipStream_t s)
{
    const tng_domain &m = *op.mesh;
    const size_t es = op.k->elem_size();
    const size_t bytes = (size_t)m.nlat * 2 * 3 * es;
    if (op.d_btab.bytes != bytes) op.d_btab.alloc(bytes);
    if (op.k->dtype == TNG_F64) run_diag_kernel<double>(op, nullptr, nullptr, 1, 0, s);
    else run_diag_kernel<float>(op, nullptr, nullptr, 1, 0, s);
    op.btab_valid = true;
}

void epsdivdiv_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    if (src.dtype == TNG_F64) apply_impl<double>(op, src, dst, s);
    else apply_impl<float>(op, src, dst, s);
}

void noalloc_scratch_init(int device)
{
    TNG_REQUIRE(device >= 0 && device < kMaxDevices, "allocation-free apply: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    TNG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TNG_REQUIRE(p != nullptr, "allocation-free apply: no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<double *>(p);
}

void epsdivdiv_apply_noalloc(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s)
{
    const int dev = op.mesh->ctx->device;
    double *ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (dev >= 0 && dev < kMaxDevices) ws = g_scratch_ptr[dev];
    }
    TNG_REQUIRE(ws != nullptr, "allocation-free apply: the context's device has no scratch address");
    const size_t btab_bytes = (size_t)op.mesh->nlat * 2 * 3 * op.k->elem_size();
    TNG_REQUIRE(op.btab_valid || op.d_btab.bytes == btab_bytes,
                "allocation-free apply: the boundary table would need a new allocation");
    if (src.dtype == TNG_F64) apply_impl<double>(op, src, dst, s, ws);
    else apply_impl<float>(op, src, dst, s, ws);
}

bool epsdivdiv_apply_cheb(tng_operator &op, const tng_vector &x, tng_vector &ax, tng_vector &x_out, tng_vector &d,
                          const tng_vector &b, const tng_vector &invd, double az, double ad, bool read_d,
                          hipStream_t s)
{
    if (op.kind != OP_EPSDIVDIV) return false;
    const tng_vector *vs[6] = {&x, &ax, &x_out, &d, &b, &invd};
    for (const tng_vector *v : vs)
        if (v->mesh != op.mesh || v->dtype != x.dtype || v->ncomp != 3) return false;
    if (x.comp(0) == x_out.comp(0)) return false;
    if (x.dtype == TNG_F64) return apply_cheb_impl<double>(op, x, ax, x_out, d, b, invd, az, ad, read_d, s);
    return apply_cheb_impl<float>(op, x, ax, x_out, d, b, invd, az, ad, read_d, s);
}

}  // namespace tng
