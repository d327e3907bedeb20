// Host-side construction of one level of the icosahedral thick spherical
// shell: node coordinates, radii, boundary and ownership masks, the halo plan
// for duplicated interface nodes, and the lateral wedge geometry table that
// the apply kernels read.
//
// Floating-point: the coordinate recursion must reproduce the reference bit
// for bit, so this file is compiled without FMA contraction and evaluates
// every expression in the reference's order.
#pragma clang fp contract(off)

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>

#include "tng_internal.hpp"

namespace tng {
namespace {

struct V3 {
    double v[3];
};

// dense/vec.hpp:88-118 norm() / normalized(): sum of squares in index order,
// sqrt, divide only when the norm exceeds 1e-15.
V3 normalized(V3 a)
{
    double s = 0;
    for (int i = 0; i < 3; ++i) s += a.v[i] * a.v[i];
    const double mag = std::sqrt(s);
    if (mag > 1e-15)
        for (int i = 0; i < 3; ++i) a.v[i] /= mag;
    return a;
}

// spherical_shell.hpp:465-555: the twelve icosahedral base nodes.
void icosahedron(double n[12][3])
{
    const double fifthpi = 0.4 * std::asin(1.0);
    const double w = 2.0 * std::acos(1.0 / (2.0 * std::sin(fifthpi)));
    const double cosw = std::cos(w), sinw = std::sin(w);
    n[0][0] = 0.0; n[0][1] = 0.0; n[0][2] = 1.0;
    n[11][0] = 0.0; n[11][1] = 0.0; n[11][2] = -1.0;
    for (int k = 1; k <= 5; k++) {
        const double phi = 2.0 * ((double)k - 0.5) * fifthpi;
        n[k][0] = sinw * std::cos(phi);
        n[k][1] = sinw * std::sin(phi);
        n[k][2] = cosw;
    }
    for (int k = 1; k <= 5; k++) {
        const double phi = 2.0 * ((double)k - 1) * fifthpi;
        n[k + 5][0] = sinw * std::cos(phi);
        n[k + 5][1] = sinw * std::sin(phi);
        n[k + 5][2] = -cosw;
    }
}

// spherical_shell.hpp:557-604: diamond -> (pole, left/right, opposite) nodes.
constexpr int kDiamondNodes[10][4] = {{0, 5, 6, 1},  {0, 1, 7, 2},  {0, 2, 8, 3},  {0, 3, 9, 4},  {0, 4, 10, 5},
                                      {11, 7, 1, 6}, {11, 8, 2, 7}, {11, 9, 3, 8}, {11, 10, 4, 9}, {11, 6, 5, 10}};

// Whole-diamond lateral node grid, (N+1) x (N+1), built coarse to fine.  At
// stride l every new node is the normalised sum of two nodes of the stride-l
// grid (the three creation rules of compute_node_recursive,
// spherical_shell.hpp:276-392), so filling level by level yields the same
// values as the memoised recursion.
std::vector<V3> diamond_grid(int diamond, int N, const double ico[12][3])
{
    const int n = N + 1;
    std::vector<V3> g((size_t)n * n);
    auto at = [&](int i, int j) -> V3 & { return g[(size_t)i * n + j]; };
    const int Lc = diamond < 5 ? 1 : 3, Rc = diamond < 5 ? 3 : 1;
    for (int d = 0; d < 3; ++d) {
        at(0, 0).v[d] = ico[kDiamondNodes[diamond][0]][d];
        at(N, 0).v[d] = ico[kDiamondNodes[diamond][Lc]][d];
        at(N, N).v[d] = ico[kDiamondNodes[diamond][2]][d];
        at(0, N).v[d] = ico[kDiamondNodes[diamond][Rc]][d];
    }
    auto mid = [](const V3 &a, const V3 &b) {
        V3 s;
        for (int d = 0; d < 3; ++d) s.v[d] = a.v[d] + b.v[d];
        return normalized(s);
    };
    for (int l = N; l >= 2; l /= 2) {
        const int h = l / 2;
        for (int i = 0; i <= N; i += l)
            for (int j = h; j <= N; j += l) at(i, j) = mid(at(i, j - h), at(i, j + h));
        for (int i = h; i <= N; i += l)
            for (int j = 0; j <= N; j += l) at(i, j) = mid(at(i - h, j), at(i + h, j));
        for (int i = h; i <= N; i += l)
            for (int j = h; j <= N; j += l) at(i, j) = mid(at(i - h, j + h), at(i + h, j - h));
    }
    return g;
}

// inverse transpose of a 3x3 matrix given by its columns a, b, c; returns det.
double inv_transpose_cols(const double a[3], const double b[3], const double c[3], double M[3][3])
{
    // L = [a b c] (columns).  L^{-T} = cof(L) / det(L).
    const double L[3][3] = {{a[0], b[0], c[0]}, {a[1], b[1], c[1]}, {a[2], b[2], c[2]}};
    const double det = L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) -
                       L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0]) +
                       L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
    const double id = 1.0 / det;
    double C[3][3];  // cofactor matrix == det * L^{-T}
    C[0][0] = (L[1][1] * L[2][2] - L[1][2] * L[2][1]);
    C[0][1] = (L[1][2] * L[2][0] - L[1][0] * L[2][2]);
    C[0][2] = (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
    C[1][0] = (L[0][2] * L[2][1] - L[0][1] * L[2][2]);
    C[1][1] = (L[0][0] * L[2][2] - L[0][2] * L[2][0]);
    C[1][2] = (L[0][1] * L[2][0] - L[0][0] * L[2][1]);
    C[2][0] = (L[0][1] * L[1][2] - L[0][2] * L[1][1]);
    C[2][1] = (L[0][2] * L[1][0] - L[0][0] * L[1][2]);
    C[2][2] = (L[0][0] * L[1][1] - L[0][1] * L[1][0]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i][j] = C[i][j] * id;
    return det;
}

}  // namespace

void mesh_build(tng_domain &m, int lat_level, int rad_level, double r_min, double r_max, int lat_sdr, int rad_sdr,
                bool device)
{
    TNG_REQUIRE(lat_level >= 1 && lat_level <= 10, "lateral level must be in [1, 10]");
    TNG_REQUIRE(rad_level >= 1 && rad_level <= 12, "radial level must be in [1, 12]");
    TNG_REQUIRE(lat_sdr >= 0 && lat_sdr <= lat_level - 1, "lateral subdomain level must be in [0, level-1]");
    TNG_REQUIRE(rad_sdr >= 0 && rad_sdr <= rad_level - 1, "radial subdomain level must be in [0, radial level-1]");
    TNG_REQUIRE(r_min < r_max && r_min > 0, "need 0 < r_min < r_max");
    m.lat_level = lat_level; m.rad_level = rad_level; m.lat_sdr = lat_sdr; m.rad_sdr = rad_sdr;
    m.r_min = r_min; m.r_max = r_max;
    m.S = 1 << lat_sdr; m.Sr = 1 << rad_sdr;
    const int E = 1 << lat_level, Lr = 1 << rad_level;
    m.nsd = 10 * m.S * m.S * m.Sr;
    m.nx = E / m.S + 1;
    m.nr = Lr / m.Sr + 1;
    m.nlat = (int64_t)m.nsd * m.nx * m.nx;
    m.nalloc = m.nlat * m.nr;
    TNG_REQUIRE(m.nalloc < (int64_t)INT32_MAX, "mesh too large for 32-bit halo indices");

    // local subdomain enumeration: diamond, x, y, r (spherical_shell.hpp:1034-1063, 2633-2639)
    m.sd_info.clear();
    for (int d = 0; d < 10; ++d)
        for (int x = 0; x < m.S; ++x)
            for (int y = 0; y < m.S; ++y)
                for (int r = 0; r < m.Sr; ++r) m.sd_info.insert(m.sd_info.end(), {d, x, y, r});

    // coordinates
    double ico[12][3];
    icosahedron(ico);
    const int nx = m.nx, nr = m.nr;
    m.coords.assign((size_t)m.nlat * 3, 0.0);
    for (int d = 0; d < 10; ++d) {
        const std::vector<V3> g = diamond_grid(d, E, ico);
        for (int sd = 0; sd < m.nsd; ++sd) {
            if (m.sd_info[4 * sd] != d) continue;
            const int i0 = m.sd_info[4 * sd + 1] * (E / m.S), j0 = m.sd_info[4 * sd + 2] * (E / m.S);
            for (int x = 0; x < nx; ++x)
                for (int y = 0; y < nx; ++y) {
                    const V3 &p = g[(size_t)(i0 + x) * (E + 1) + (j0 + y)];
                    std::memcpy(&m.coords[3 * (((size_t)sd * nx + x) * nx + y)], p.v, 3 * sizeof(double));
                }
        }
    }

    // radii: uniform_shell_radii (spherical_shell.hpp:27-52) and the per-subdomain slice (:2862-2887)
    std::vector<double> all_r(Lr + 1);
    const double r_step = (r_max - r_min) / (Lr);
    for (int i = 0; i <= Lr; ++i) all_r[i] = r_min + i * r_step;
    all_r[Lr] = r_max;
    m.radii.assign((size_t)m.nsd * nr, 0.0);
    for (int sd = 0; sd < m.nsd; ++sd)
        for (int j = 0; j < nr; ++j) m.radii[(size_t)sd * nr + j] = all_r[m.sd_info[4 * sd + 3] * (nr - 1) + j];

    // boundary flags (grid/shell/bit_masks.hpp:82-136)
    m.bmask.assign((size_t)m.nalloc, SBF_INNER);
    m.sd_has_cmb.assign(m.nsd, 0);
    m.sd_has_surf.assign(m.nsd, 0);
    for (int sd = 0; sd < m.nsd; ++sd) {
        const int sr = m.sd_info[4 * sd + 3];
        m.sd_has_cmb[sd] = sr == 0;
        m.sd_has_surf[sd] = sr == m.Sr - 1;
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < nx; ++y) {
                if (sr == 0) m.bmask[m.lin(sd, x, y, 0)] = SBF_CMB;
                if (sr == m.Sr - 1) m.bmask[m.lin(sd, x, y, nr - 1)] = SBF_SURFACE;
            }
    }

    // Duplicates: copies of a node share bit-identical lateral coordinates and
    // the global shell index.  Lateral points are identified by their
    // coordinate bits, nodes by (lateral point, global shell).
    std::vector<std::pair<std::array<uint64_t, 3>, int64_t>> keys((size_t)m.nlat);
    for (int64_t i = 0; i < m.nlat; ++i) {
        std::memcpy(keys[i].first.data(), &m.coords[3 * i], 3 * sizeof(double));
        keys[i].second = i;
    }
    std::sort(keys.begin(), keys.end());
    std::vector<int64_t> lat_id((size_t)m.nlat);
    std::vector<std::vector<int64_t>> lat_groups;  // lateral point -> lateral columns
    for (int64_t i = 0; i < m.nlat; ++i) {
        if (i == 0 || keys[i].first != keys[i - 1].first) lat_groups.emplace_back();
        lat_groups.back().push_back(keys[i].second);
        lat_id[keys[i].second] = (int64_t)lat_groups.size() - 1;
    }
    // ownership: MIN global subdomain id among the copies (grid/bit_masks.hpp:29-85;
    // SubdomainInfo::global_id, spherical_shell.hpp:696-722)
    auto gid = [&](int sd) {
        const int *s = &m.sd_info[4 * sd];
        return ((int64_t)s[0] << 57) | ((int64_t)s[3] << 38) | ((int64_t)s[2] << 19) | (int64_t)s[1];
    };
    m.omask.assign((size_t)m.nalloc, 0);
    const int nshell = Lr + 1;
    // for every lateral point, gather copies per global shell
    std::vector<std::vector<int32_t>> multi;  // list of copy sets
    int maxc = 0;
    int64_t n_interface = 0, n_owned = 0;
    std::vector<std::vector<int64_t>> per_shell(nshell);
    for (const auto &cols : lat_groups) {
        for (auto &v : per_shell) v.clear();
        for (int64_t col : cols) {
            const int sd = (int)(col / ((int64_t)nx * nx));
            const int sr = m.sd_info[4 * sd + 3];
            for (int r = 0; r < nr; ++r) per_shell[sr * (nr - 1) + r].push_back(col * nr + r);
        }
        for (int g = 0; g < nshell; ++g) {
            auto &cp = per_shell[g];
            if (cp.empty()) continue;
            std::sort(cp.begin(), cp.end());
            int64_t best = cp[0];
            int64_t best_gid = gid((int)(cp[0] / ((int64_t)nx * nx * nr)));
            for (int64_t c : cp) {
                const int64_t gg = gid((int)(c / ((int64_t)nx * nx * nr)));
                if (gg < best_gid) { best_gid = gg; best = c; }
            }
            m.omask[best] = OWNED;
            n_owned++;
            if (cp.size() >= 2) {
                n_interface += (int64_t)cp.size();
                multi.emplace_back(cp.begin(), cp.end());
                maxc = std::max<int>(maxc, (int)cp.size());
            }
        }
    }
    m.num_owned = n_owned;
    m.num_interface = n_interface;
    m.halo_maxc = maxc;
    m.halo_n = (int64_t)multi.size();
    std::vector<int32_t> hidx((size_t)maxc * m.halo_n, -1);
    for (int64_t t = 0; t < m.halo_n; ++t)
        for (size_t j = 0; j < multi[t].size(); ++j) hidx[j * m.halo_n + t] = multi[t][j];

    // Gram matrix of the rotation fields rho_j = e_j x x over OWNED nodes:
    // G_jk = sum (|x|^2 delta_jk - x_j x_k).  Its inverse turns the owned dots
    // <rho_j, u> into the coefficients of the orthogonal projection onto
    // span{rho_j}, which is what the Gram-Schmidt-orthonormalised penalty modes
    // of the reference represent (epsilon_divdiv_kerngen.hpp:324-445).
    long double G[3][3] = {{0}};
    for (int sd = 0; sd < m.nsd; ++sd)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < nx; ++y) {
                const double *c = &m.coords[3 * (((size_t)sd * nx + x) * nx + y)];
                for (int r = 0; r < nr; ++r) {
                    if (m.omask[m.lin(sd, x, y, r)] != OWNED) continue;
                    const double rad = m.radii[(size_t)sd * nr + r];
                    const long double X[3] = {c[0] * rad, c[1] * rad, c[2] * rad};
                    const long double n2 = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) G[a][b] += (a == b ? n2 : 0.0L) - X[a] * X[b];
                }
            }
    {
        const long double det = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) -
                                G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                                G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
        long double Ci[3][3];
        Ci[0][0] = (G[1][1] * G[2][2] - G[1][2] * G[2][1]);
        Ci[0][1] = (G[0][2] * G[2][1] - G[0][1] * G[2][2]);
        Ci[0][2] = (G[0][1] * G[1][2] - G[0][2] * G[1][1]);
        Ci[1][0] = (G[1][2] * G[2][0] - G[1][0] * G[2][2]);
        Ci[1][1] = (G[0][0] * G[2][2] - G[0][2] * G[2][0]);
        Ci[1][2] = (G[0][2] * G[1][0] - G[0][0] * G[1][2]);
        Ci[2][0] = (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
        Ci[2][1] = (G[0][1] * G[2][0] - G[0][0] * G[2][1]);
        Ci[2][2] = (G[0][0] * G[1][1] - G[0][1] * G[1][0]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) m.rot_gram_inv[3 * a + b] = (double)(Ci[a][b] / det);
    }

    if (!device) return;
    m.d_halo_idx.alloc(hidx.size() * sizeof(int32_t));
    if (!hidx.empty())
        TNG_HIP_CHECK(hipMemcpy(m.d_halo_idx.p, hidx.data(), hidx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
}

void mesh_upload(tng_domain &m)
{
    const int nx = m.nx, nr = m.nr;
    auto up = [](DevBuf &b, const void *src, size_t bytes) {
        b.alloc(bytes);
        TNG_HIP_CHECK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    up(m.d_coords, m.coords.data(), m.coords.size() * sizeof(double));
    up(m.d_radii, m.radii.data(), m.radii.size() * sizeof(double));
    up(m.d_omask, m.omask.data(), m.omask.size());
    up(m.d_bmask, m.bmask.data(), m.bmask.size());
    up(m.d_sd_cmb, m.sd_has_cmb.data(), m.sd_has_cmb.size());
    up(m.d_sd_surf, m.sd_has_surf.data(), m.sd_has_surf.size());
    std::vector<float> cf(m.coords.begin(), m.coords.end()), rf(m.radii.begin(), m.radii.end());
    up(m.d_coords_f, cf.data(), cf.size() * sizeof(float));
    up(m.d_radii_f, rf.data(), rf.size() * sizeof(float));

    // lateral wedge geometry (sd, xc, yc, w, 10)
    const int nc = nx - 1;
    std::vector<double> lg((size_t)m.nsd * nc * nc * 20);
    // wedge corner (dx, dy): kernel_helpers.hpp:25-56 wedge_surface_physical_coords
    static const int WC[2][3][2] = {{{0, 0}, {1, 0}, {0, 1}}, {{1, 1}, {0, 1}, {1, 0}}};
    for (int sd = 0; sd < m.nsd; ++sd)
        for (int xc = 0; xc < nc; ++xc)
            for (int yc = 0; yc < nc; ++yc)
                for (int w = 0; w < 2; ++w) {
                    const double *p[3];
                    for (int k = 0; k < 3; ++k)
                        p[k] = &m.coords[3 * (((size_t)sd * nx + xc + WC[w][k][0]) * nx + yc + WC[w][k][1])];
                    double a[3], b[3], c[3];
                    for (int d = 0; d < 3; ++d) {
                        a[d] = p[1][d] - p[0][d];
                        b[d] = p[2][d] - p[0][d];
                        c[d] = (1.0 / 3.0) * (p[0][d] + p[1][d] + p[2][d]);
                    }
                    double M[3][3];
                    const double det = inv_transpose_cols(a, b, c, M);
                    double *o = &lg[((((size_t)sd * nc + xc) * nc + yc) * 2 + w) * 10];
                    // M column i = L^{-T} e_i  -> store columns M0, M1, M2
                    for (int col = 0; col < 3; ++col)
                        for (int d = 0; d < 3; ++d) o[3 * col + d] = M[d][col];
                    o[9] = std::fabs(det);
                }
    up(m.d_latgeom, lg.data(), lg.size() * sizeof(double));
    std::vector<float> lgf(lg.begin(), lg.end());
    up(m.d_latgeom_f, lgf.data(), lgf.size() * sizeof(float));
    (void)nr;
}

}  // namespace tng
