// Internal declarations of the tng library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "tng.h"

namespace tng {

// Status-carrying exception used inside the library; converted to tng_status
// at the C ABI boundary.
struct Error : std::runtime_error {
    tng_status status;
    Error(tng_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

#define TNG_HIP_CHECK(expr)                                                                                    \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            throw ::tng::Error(e_ == hipErrorOutOfMemory ? TNG_OUT_OF_MEMORY : TNG_DEVICE_ERROR,               \
                               std::string(#expr) + ": " + hipGetErrorString(e_));                            \
    } while (0)

#define TNG_REQUIRE(cond, msg)                                                                                 \
    do {                                                                                                       \
        if (!(cond)) throw ::tng::Error(TNG_INVALID_ARGUMENT, msg);                                            \
    } while (0)

// Shell boundary flags (grid/shell/bit_masks.hpp:11-19) and ownership flag
// (grid/bit_masks.hpp:16-20), same bit values as the reference.
enum : uint8_t { SBF_INNER = 1, SBF_BOUNDARY = 2, SBF_CMB = 2 | 4, SBF_SURFACE = 2 | 8 };
enum : uint8_t { OWNED = 1 };

// Division by a launch-constant divisor through a multiply-high (round-up
// method of Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994): q = (mulhi(m, n) + n) >> l with
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1.  Exact for n < 2^31
// (the sum then fits 32 bits); index decoding in the node kernels uses it
// instead of 64-bit division.
struct FastDiv {
    uint32_t d = 1, m = 1;
    int l = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t div) : d(div)
    {
        l = 0;
        while (((uint64_t)1 << l) < div) ++l;
        m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - div)) / div + 1);
    }
    __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t t = __umulhi(m, n);
#else
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
#endif
        return (t + n) >> l;
    }
};

// (sd, x, y, r) of a linear index over a (nsd, nx, nx, nr) grid, r fastest
struct Index4 {
    FastDiv r, x;
    Index4() = default;
    Index4(int nx, int nr) : r((uint32_t)nr), x((uint32_t)nx) {}
    __host__ __device__ __forceinline__ void decode(uint32_t i, int &sd, int &X, int &Y, int &l) const
    {
        const uint32_t t = r.div(i);
        l = (int)(i - t * r.d);
        const uint32_t t2 = x.div(t);
        Y = (int)(t - t2 * x.d);
        const uint32_t t3 = x.div(t2);
        X = (int)(t2 - t3 * x.d);
        sd = (int)t3;
    }
};

// Device buffer RAII.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    void alloc(size_t n)
    {
        release();
        if (n) TNG_HIP_CHECK(hipMalloc(&p, n));
        bytes = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace tng

// Launch-shape settings of the apply kernels (tng_ctx_set_tuning); every
// setting computes the same operator.
struct tng_tuning {
    int apply_blocks = 0;        // FP64 applies: workgroups sharing the node rows (0: resident count)
    int apply_seg = 0;           // FP32 and fused applies: node rows per workgroup (0: chosen per mesh)
    int apply_radial_chunks = 0; // 1: 128 and 256 layers as two or four cells per lane
    int apply_radial_split = 1;  // 0: 128 and 256 layers as radial chunks in one workgroup
    int apply_waves = 8;         // 4: four-wave workgroups at 64 layers
    int cheb_fuse = 1;           // 0: Chebyshev sub-steps as separate kernels (read at solver creation)
};

struct tng_context {
    int device = 0;
    tng_tuning tuning;
    hipStream_t stream = nullptr;
    bool owns_stream = true;  // false: the caller's stream (tng_ctx_create_with_stream)
    std::string last_error;
    bool profiling = false;
    // Scratch for two-stage reductions: partial sums + a ticket counter.
    tng::DevBuf red_partials;   // doubles
    tng::DevBuf red_result;     // small device result slots (doubles)
    // Profiling: per apply four events (begin, main kernel begin/end, end)
    // recorded on the stream without synchronising; read back in bulk.
    std::vector<hipEvent_t> prof_events;
    size_t prof_used = 0;  // event sets recorded since the last read
    hipEvent_t *prof_next_set();
};

// One multigrid level of the thick spherical shell: DistributedDomain plus
// coordinates, radii, masks, halo plan and the precomputed lateral wedge
// geometry.
struct tng_domain {
    tng_context *ctx = nullptr;
    int lat_level = 0, rad_level = 0, lat_sdr = 0, rad_sdr = 0;
    int S = 1, Sr = 1;          // subdomains per diamond side / radially
    int nsd = 0, nx = 0, nr = 0; // nx == ny
    int64_t nlat = 0;           // nsd * nx * nx
    int64_t nalloc = 0;         // nlat * nr
    int64_t num_owned = 0, num_interface = 0;
    double r_min = 0, r_max = 0;
    bool solver_owned = false;  // a Stokes solver's level mesh (not destroyable through the ABI)

    std::vector<int> sd_info;      // (diamond, sx, sy, sr) per local subdomain
    std::vector<double> coords;    // (sd, x, y, 3)
    std::vector<double> radii;     // (sd, r)
    std::vector<uint8_t> bmask;    // (sd, x, y, r)
    std::vector<uint8_t> omask;    // (sd, x, y, r)
    std::vector<uint8_t> sd_has_cmb, sd_has_surf;

    // device copies
    tng::DevBuf d_coords, d_radii, d_omask, d_bmask, d_sd_cmb, d_sd_surf;
    tng::DevBuf d_coords_f, d_radii_f;
    // lateral wedge geometry: (sd, xc, yc, wedge, 10) = (M0, M1, M2, |det L|),
    // M = columns of L^{-T}, L = [p1-p0, p2-p0, (p0+p1+p2)/3]
    tng::DevBuf d_latgeom, d_latgeom_f;
    // halo plan: one entry per node that has >= 2 copies; copy slots in SoA
    // [maxc][nmulti] of int32 linear indices (-1 = unused).
    int halo_maxc = 0;
    int64_t halo_n = 0;
    tng::DevBuf d_halo_idx;
    // Gram matrix of the rotation fields e_j x x over OWNED nodes (inverse).
    double rot_gram_inv[9] = {0};

    int64_t lin(int sd, int x, int y, int r) const { return (((int64_t)sd * nx + x) * nx + y) * nr + r; }
};

// A node field of 1 or 3 components, each a separate (sd, x, y, r) array.
// Owned vectors keep their components in one allocation, each starting on a
// 256-byte boundary; wrapped vectors point at caller memory.
struct tng_vector {
    tng_domain *mesh = nullptr;
    tng_dtype dtype = TNG_F64;
    int ncomp = 1;
    void *cp[3] = {nullptr, nullptr, nullptr};  // component base addresses
    tng::DevBuf buf;                            // storage of an owned vector (empty when wrapped)
    bool swappable = false;                     // internal work vector: its storage may be exchanged
    size_t elem_size() const { return dtype == TNG_F64 ? 8 : 4; }
    void *comp(int c) const { return cp[c]; }
};

namespace tng {
// Pinned host staging memory (hipHostMalloc), grown on demand.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void reserve(size_t n);
    void release();
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
};
}  // namespace tng

// XDMF output of node fields of one mesh (io/xdmf.hpp:148-513).
struct tng_xdmf_writer {
    struct Item {
        tng_vector *v;
        std::string label;
        int bytes;
    };
    tng_domain *mesh = nullptr;
    std::string dir;
    int points_bytes = 4, conn_bytes = 8;
    std::vector<Item> items;
    int counter = 0;
    bool first_done = false;
    tng::DevBuf dev;
    tng::PinnedBuf host;
    double last_pack_seconds = 0.0, last_write_seconds = 0.0;
    std::vector<const Item *> ordered() const;
    void add(tng_vector *v, const std::string &label, int bytes);
    void emit(const std::string &path, size_t bytes, hipStream_t s);
    void write_metadata();
    void write();
};

namespace tng {
// Component pointers passed to kernels by value.
template <typename T> struct Comp3 {
    T *p[3];
};
template <typename T> inline Comp3<T> comps(const tng_vector &v)
{
    return {{static_cast<T *>(v.cp[0]), static_cast<T *>(v.cp[1]), static_cast<T *>(v.cp[2])}};
}
}  // namespace tng

enum OpKind {
    OP_EPSDIVDIV = 1,
    OP_DIVERGENCE,
    OP_GRADIENT,
    OP_PROLONGATION,
    OP_RESTRICTION,
    OP_KMASS_LUMPED,
    OP_VECTOR_MASS,
    OP_SUPG,    // UnsteadyAdvectionDiffusionSUPG (scalar temperature operator)
    OP_STOKES,  // EpsDivDivStokes: owns its blocks A, B^T, B and the zero block
    OP_ZERO     // Zero (pressure -> pressure): Replace writes 0, Add leaves dst
};

struct tng_operator {
    OpKind kind = OP_EPSDIVDIV;
    tng_domain *mesh = nullptr;   // fine / own mesh
    tng_domain *mesh2 = nullptr;  // coarse mesh of the two-level operators
    tng::DevBuf scratch;          // per-operator work space (coarse-wedge partials)
    tng_vector *k = nullptr;
    int bc_cmb = TNG_NEUMANN, bc_surf = TNG_NEUMANN;
    bool diagonal = false;
    int apply_mode = TNG_REPLACE, comm_mode = TNG_COMMUNICATE_ADDITIVELY;
    bool penalty_active = false;
    double penalty_eps = 1e-7;
    // boundary-row data (sd, x, y, side, 3): Dirichlet diagonal or free-slip Ann
    tng::DevBuf d_btab;
    bool btab_valid = false;
    tng::DevBuf d_penalty_b;  // 3 doubles: eps * G^{-1} m, written on device
    tng::DevBuf d_iface;      // partial sums between radial parts (apply sweeps of 256 layers)
    // OP_SUPG (UnsteadyAdvectionDiffusionSUPG): velocity (captured by reference), diffusivity,
    // dt, mass scaling and flags
    tng_vector *vel = nullptr;
    double supg_kappa = 0.0, supg_dt = 0.0, supg_mass_scaling = 1.0;
    bool supg_treat_boundary = true, supg_lumped = false, supg_enabled = true;
    // stored local matrices (OperatorStoredMatrixMode::Full): (sd, xc, yc, rc, wedge, 18, 18)
    bool stored = false;
    tng_dtype lmat_dtype = TNG_F64;
    tng::DevBuf d_lmat;
    // OP_STOKES: blocks 11, 12, 21, 22 (owned; handed out as borrowed handles)
    tng_operator *blocks[4] = {nullptr, nullptr, nullptr, nullptr};
    tng_operator *parent = nullptr;  // a block of a Stokes operator (not destroyable on its own)
    ~tng_operator();
};

namespace tng {

void mesh_build(tng_domain &m, int lat_level, int rad_level, double r_min, double r_max, int lat_sdr, int rad_sdr,
                bool device);
void mesh_upload(tng_domain &m);

// kernel launchers (kernels_*.hip)
void launch_fill(tng_vector &v, double value, hipStream_t s);
void launch_copy(tng_vector &dst, const tng_vector &src, hipStream_t s);
void launch_halo_sum(const tng_domain &m, tng_vector &v, hipStream_t s);
void epsdivdiv_update_boundary_table(tng_operator &op, hipStream_t s);
void epsdivdiv_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s);
// The same apply with its scratch (penalty partials, partial sums between
// radial parts) in a fixed per-GPU device array: no device allocation.
// noalloc_scratch_init records the array's address on the current device.
void noalloc_scratch_init(int device);
void epsdivdiv_apply_noalloc(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s);
// One Chebyshev sub-step fused into the apply: ax = A x on shared node copies
// (scratch), d = az (b - A x) invd [+ ad d], x_out = x + d everywhere.
// Returns false, launching nothing, where the fused kernels do not apply.
bool epsdivdiv_apply_cheb(tng_operator &op, const tng_vector &x, tng_vector &ax, tng_vector &x_out, tng_vector &d,
                          const tng_vector &b, const tng_vector &invd, double az, double ad, bool read_d,
                          hipStream_t s);
// XDMF field layout: pack (device layout -> file layout, out_bytes 4 or 8)
// and unpack; checkpoint read of one field (io/xdmf.hpp:1199-1441).
void xdmf_pack(const tng_vector &v, void *dst_device, int out_bytes, hipStream_t s);
void xdmf_unpack(const void *src_device, int in_bytes, tng_vector &v, hipStream_t s);
void xdmf_checkpoint_read(const std::string &dir, const std::string &label, int step, tng_vector &v);
void launch_halo_cheb(const tng_domain &m, const tng_vector &ax, const tng_vector &b, const tng_vector &invd,
                      tng_vector &d, const tng_vector &x, tng_vector &x_out, double az, double ad, bool read_d,
                      hipStream_t s);
void stokes_block_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s);

// BLAS-1 and reductions (kernels_vec.hip)
void launch_lincomb(tng_vector &y, int nx, const double *c, const tng_vector *const *x, double c0, hipStream_t s);
void launch_mult(tng_vector &y, const tng_vector &x, hipStream_t s);
void launch_invert(tng_vector &y, hipStream_t s);
void launch_cheb_update(tng_vector &x, tng_vector &d, const tng_vector &b, const tng_vector *ax,
                        const tng_vector &invd, double cz, double cd, bool read_d, bool x_zero, hipStream_t s);
enum { REDUCE_DOT = 0, REDUCE_MAXABS = 1, REDUCE_NONFINITE = 2 };
void launch_reduce(int op, const tng_vector &x, const tng_vector *y, int slot, hipStream_t s);
double reduce_to_host(int op, const tng_vector &x, const tng_vector *y, hipStream_t s);
// PCG with device-side scalars (identity preconditioner)
void launch_pcg_begin(const tng_vector &r, double atol, hipStream_t s);
void launch_pcg_alpha(const tng_vector &ap, const tng_vector &p, int it, hipStream_t s);
void launch_pcg_update(tng_vector &x, tng_vector &r, const tng_vector &p, const tng_vector &ap, double rtol,
                       double atol, int it, hipStream_t s);
void launch_pcg_direction(tng_vector &p, const tng_vector &r, hipStream_t s);
void pcg_read_state(const tng_vector &v, int *done, int *iters, double *final_rel, hipStream_t s);
// FGMRES modified Gram-Schmidt step on a Stokes vector; h values live on the device
void launch_mgs_step(tng_vector &wu, tng_vector &wp, const tng_vector *vprev_u, const tng_vector *vprev_p,
                     const double *h_prev_dev, const tng_vector *vnext_u, const tng_vector *vnext_p, double *h_out_dev,
                     hipStream_t s);
double *device_slots(tng_context &ctx, int first);  // ctx.red_result + first (doubles, 512 in total)

// generic operator apply (dispatch on kind)
void op_apply(tng_operator &op, const tng_vector &src, tng_vector &dst);
// EpsDivDivStokes apply: [u; p] -> [A u + B^T p; B u]
void stokes_op_apply(tng_operator &op, const tng_vector &u, const tng_vector &p, tng_vector &out_u, tng_vector &out_p);

// object factories (throw tng::Error; the caller owns the result)
tng_domain *new_mesh(tng_context *ctx, int lat_level, int rad_level, double r_min, double r_max, int lat_sdr,
                     int rad_sdr);
tng_vector *new_vector(tng_domain *mesh, tng_dtype dtype, int ncomp);
tng_operator *new_epsdivdiv(tng_domain *mesh, tng_vector *k, int bc_cmb, int bc_surf, bool diagonal);
tng_operator *new_two_level(OpKind kind, tng_domain *fine, tng_domain *coarse, int bc_cmb, int bc_surf);
tng_operator *new_kmass(tng_domain *mesh, tng_vector *k);

// stored local matrices and GCA (kernels_gca.hip)
size_t stored_matrix_count(const tng_domain &m);  // wedges = matrices
void epsdivdiv_local_matrices(const tng_operator &op, DevBuf &out, hipStream_t s);  // FP64, from k
// exact = true: the Galerkin product R A P (P interpolates top-face nodes of
// interior coarse cells from the coarse top layer); false: the reference's rule
void gca_assemble(const tng_operator &fine, tng_operator &coarse, hipStream_t s, bool exact = false);
// adaptive GCA: flags of the coarsest cells from the fine coefficient, and the
// selective assembly (Galerkin on flagged cells, on-the-fly matrices elsewhere)
void gca_collect(const tng_vector &k_fine, int level_range, const tng_domain &coarsest, DevBuf &flags, hipStream_t s);
void gca_assemble_selective(const tng_operator &fine, tng_operator &coarse, int level_range,
                            const tng_domain &coarsest, const DevBuf &flags, hipStream_t s, bool exact = false);
void copy_stored_matrices(tng_operator &dst, const tng_operator &src, tng_dtype dtype, hipStream_t s);
template <typename T>
void stored_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, const double *pen, bool add, hipStream_t s);

// energy equation (kernels_energy.hip, kernels_stokes.hip)
void supg_apply(tng_operator &op, const tng_vector &src, tng_vector &dst, hipStream_t s);
void launch_scalar_mass(const tng_vector &src, tng_vector &dst, hipStream_t s);
void launch_wedge_gather(const double *partial, tng_vector &dst, bool add, hipStream_t s);
void launch_boundary_values(tng_vector &g, double t_cmb, double t_surf, hipStream_t s);
void launch_assign_masked(tng_vector &b, const tng_vector &src, uint8_t flag, hipStream_t s);
void launch_viscosity_fk(tng_vector &eta, const tng_vector &temp, double rmu, hipStream_t s);
double max_vector_magnitude(const tng_vector &u, hipStream_t s);

// viscosity adjoint kernel (kernels_adjoint.hip): g = d(lam^T A(k) u)/d log k
void epsdivdiv_log_k_gradient(tng_operator &op, const tng_vector &lam, const tng_vector &u, tng_vector &g,
                              hipStream_t s);

// u = Ra T x/|x| (the app's buoyancy interpolator)
void launch_buoyancy(const tng_vector &temp, double rayleigh, tng_vector &u, hipStream_t s);

// solver support kernels
void launch_randomize(tng_vector &v, uint64_t seed, hipStream_t s);
void launch_inject(tng_vector &coarse, const tng_vector &fine, hipStream_t s);
// RHS boundary enforcement of a velocity vector: zero at the nodes of
// boundary `dirichlet_flag`, normal component removed at the nodes of
// boundary `freeslip_flag` (flags: 0 none, SBF_CMB or SBF_SURFACE)
void launch_rhs_enforce(tng_vector &u, uint8_t dirichlet_flag, uint8_t freeslip_flag, hipStream_t s);

}  // namespace tng
