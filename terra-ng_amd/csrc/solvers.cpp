// Solver layer of the Stokes path, host C++ driving the library's kernels on
// the context stream:
//   - Chebyshev-accelerated Jacobi smoother + power iteration
//     (linalg/solvers/chebyshev.hpp:51-185, power_iteration.hpp:13-58)
//   - PCG with the identity preconditioner (linalg/solvers/pcg.hpp:28-188)
//   - geometric multigrid V-cycle (linalg/solvers/multigrid.hpp:216-432)
//   - the Stokes block operator K = [A B^T; B 0]
//     (fe/wedge/operators/shell/epsilon_divdiv_stokes.hpp:55-84)
//   - the block-triangular preconditioner
//     (linalg/solvers/block_preconditioner_2x2.hpp:205-221)
//   - FGMRES(m) with modified Gram-Schmidt and Givens rotations
//     (linalg/solvers/fgmres.hpp:109-373)
//   - the solver set-up of the mantle-circulation app
//     (apps/mantlecirculation/src/stokes_solver.hpp:143-691).
// The multigrid levels can run in FP32 (vectors, operators, smoothers and
// transfers) while FGMRES, the Stokes operator, the Schur scaling and the
// coarsest-level PCG stay in FP64.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "tng_gs.hpp"
#include "tng_internal.hpp"
#include "tng_solvers.hpp"

namespace tng {

namespace {

using VecPtr = std::unique_ptr<tng_vector>;
using OpPtr = std::unique_ptr<tng_operator>;
using MeshPtr = std::unique_ptr<tng_domain>;

hipStream_t stream_of(const tng_vector &v) { return v.mesh->ctx->stream; }

// -- vector helpers (linalg/vector.hpp) --------------------------------------
void lincomb(tng_vector &y, std::initializer_list<double> c, std::initializer_list<const tng_vector *> x,
             double c0 = 0.0)
{
    const tng_vector *xs[3] = {nullptr, nullptr, nullptr};
    double cs[3] = {0, 0, 0};
    int n = 0;
    for (const tng_vector *v : x) xs[n++] = v;
    n = 0;
    for (double v : c) cs[n++] = v;
    launch_lincomb(y, (int)x.size(), cs, xs, c0, stream_of(y));
}
void assign(tng_vector &y, double c) { launch_fill(y, c, stream_of(y)); }
void assign(tng_vector &y, const tng_vector &x) { launch_copy(y, x, stream_of(y)); }
void scale_in_place(tng_vector &y, const tng_vector &x) { launch_mult(y, x, stream_of(y)); }
double dot(const tng_vector &x, const tng_vector &y) { return reduce_to_host(REDUCE_DOT, x, &y, stream_of(x)); }
double norm_2(const tng_vector &x) { return std::sqrt(dot(x, x)); }
bool has_nan_or_inf(const tng_vector &x) { return reduce_to_host(REDUCE_NONFINITE, x, nullptr, stream_of(x)) > 0; }
void apply(tng_operator &op, const tng_vector &src, tng_vector &dst) { op_apply(op, src, dst); }

tng_vector *vec(std::vector<VecPtr> &pool, tng_domain *m, tng_dtype dt, int ncomp)
{
    pool.emplace_back(new_vector(m, dt, ncomp));
    pool.back()->swappable = true;  // solver work vectors are only reached through their objects
    return pool.back().get();
}

// Exchange the storage of two owned work vectors of one shape.
void swap_storage(tng_vector &a, tng_vector &b)
{
    TNG_REQUIRE(a.mesh == b.mesh && a.dtype == b.dtype && a.ncomp == b.ncomp && a.buf.p && b.buf.p &&
                    a.buf.bytes == b.buf.bytes,
                "swap: vectors of different shape or not owned");
    std::swap(a.cp, b.cp);
    std::swap(a.buf.p, b.buf.p);
}

}  // namespace

// ---------------------------------------------------------------------------
// power_iteration (power_iteration.hpp:13-58) on D^{-1} A
// ---------------------------------------------------------------------------
double power_iteration(tng_operator &A, const tng_vector &inv_diag, tng_vector &it, tng_vector &aux, int iterations,
                       uint64_t seed)
{
    launch_randomize(it, seed, stream_of(it));
    double norm = norm_2(it);
    if (norm < std::numeric_limits<double>::min()) return 1.0;
    lincomb(it, {1.0 / norm}, {&it});
    apply(A, it, aux);
    scale_in_place(aux, inv_diag);
    double radius = 0.0;
    for (int i = 0; i < iterations; ++i) {
        norm = norm_2(aux);
        if (norm < std::numeric_limits<double>::min()) return radius > 0.0 ? radius : 1.0;
        lincomb(it, {1.0 / norm}, {&aux});
        apply(A, it, aux);
        scale_in_place(aux, inv_diag);
        radius = dot(it, aux);
    }
    return radius;
}

// ---------------------------------------------------------------------------
// Chebyshev smoother (chebyshev.hpp:102-165)
// ---------------------------------------------------------------------------
void Chebyshev::solve(tng_operator &A, tng_vector &x, const tng_vector &b, bool x_zero)
{
    if (need_estimate) {
        lmax_estimate = power_iteration(A, *inv_diag, *d, *z, power_iterations, seed);
        need_estimate = false;
    }
    if (!std::isfinite(lmax_estimate) || lmax_estimate <= 0.0) lmax_estimate = 1.0;
    hipStream_t s = stream_of(x);
    if (jacobi) {
        // x += omega D^{-1} (b - A x), by default omega = 2 / (1.3 lambda_max)
        const double omega = this->omega > 0.0 ? this->omega : 2.0 / (1.3 * lmax_estimate);
        for (int iter = 0; iter < iterations; ++iter) {
            if (x_zero && iter == 0)
                launch_cheb_update(x, *d, b, nullptr, *inv_diag, omega, 0.0, false, true, s);
            else if (xb && x.swappable && epsdivdiv_apply_cheb(A, x, *z, *xb, *d, b, *inv_diag, omega, 0.0, false, s))
                swap_storage(x, *xb);
            else {
                apply(A, x, *z);
                launch_cheb_update(x, *d, b, z, *inv_diag, omega, 0.0, false, false, s);
            }
        }
        return;
    }
    const double lambda_max = 1.5 * lmax_estimate, lambda_min = 0.1 * lmax_estimate;
    const double theta = 0.5 * (lambda_max + lambda_min), delta = 0.5 * (lambda_max - lambda_min);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;  // carried across iterations, as in the reference
    // d = az D^{-1}(b - A x) [+ ad d]; x += d.  With a spare vector the update
    // runs inside the apply and writes x + d to the spare, whose storage then
    // becomes x's (the apply still reads the old x).
    auto sub_step = [&](double az, double ad, bool read_d) {
        if (xb && x.swappable && epsdivdiv_apply_cheb(A, x, *z, *xb, *d, b, *inv_diag, az, ad, read_d, s)) {
            swap_storage(x, *xb);
            return;
        }
        apply(A, x, *z);
        launch_cheb_update(x, *d, b, z, *inv_diag, az, ad, read_d, false, s);
    };
    for (int iter = 0; iter < iterations; ++iter) {
        // first sub-step: d = z / theta.  With x = 0 on entry A x is exactly
        // zero and is not applied.
        if (x_zero && iter == 0)
            launch_cheb_update(x, *d, b, nullptr, *inv_diag, 1.0 / theta, 0.0, false, true, s);
        else
            sub_step(1.0 / theta, 0.0, false);
        for (int i = 2; i <= order; ++i) {
            const double rho_new = 1.0 / (2.0 * sigma - rho);
            const double beta = rho_new * rho, alpha = 2.0 * rho_new / delta;
            sub_step(alpha, beta, true);
            rho = rho_new;
        }
    }
}

// ---------------------------------------------------------------------------
// PCG, identity preconditioner (pcg.hpp:81-170)
// ---------------------------------------------------------------------------
int pcg_solve(tng_operator &A, tng_vector &x, const tng_vector &b, tng_vector &r, tng_vector &p, tng_vector &ap,
              tng_vector & /*z: identity preconditioner, z = r*/, int max_iterations, double rtol, double atol,
              double *final_rel, bool x_zero)
{
    // The scalars (dots, alpha, beta, the stopping test) stay on the device;
    // the host only checks every few iterations whether the solve finished.
    constexpr int kCheckEvery = 8;
    hipStream_t s = stream_of(x);
    if (x_zero) {
        assign(r, b);  // A x = 0 exactly
    } else {
        apply(A, x, r);
        lincomb(r, {1.0, -1.0}, {&b, &r});
    }
    assign(p, r);  // z = r, p = z
    launch_pcg_begin(r, atol, s);
    int done = 0, iters = 0;
    double rel = 1.0;
    int it = 1;
    for (; it <= max_iterations; ++it) {
        apply(A, p, ap);
        launch_pcg_alpha(ap, p, it, s);
        launch_pcg_update(x, r, p, ap, rtol, atol, it, s);
        launch_pcg_direction(p, r, s);
        if (it % kCheckEvery == 0 && it < max_iterations) {
            pcg_read_state(x, &done, &iters, &rel, s);
            if (done) break;
        }
    }
    pcg_read_state(x, &done, &iters, &rel, s);
    if (final_rel) *final_rel = rel;
    return done ? iters : max_iterations;
}

// ---------------------------------------------------------------------------
// Multigrid V-cycle (multigrid.hpp:216-432)
// ---------------------------------------------------------------------------
void Multigrid::coarse_solve(tng_vector &e0, const tng_vector &r0, bool x_zero)
{
    if (use_direct) {
        // exact solve: no initial guess to use
        direct.solve(e0, r0, stream_of(e0));
        coarse_iterations += 1;
        return;
    }
    // The coarsest solve runs in FP64 whatever the precision of the levels.
    tng_vector &x = e0.dtype == TNG_F64 ? e0 : *cg_x;
    const tng_vector *b = &r0;
    if (e0.dtype != TNG_F64) {
        assign(*cg_x, e0);
        assign(*cg_b, r0);
        b = cg_b;
    }
    coarse_iterations += pcg_solve(*coarse_op, x, *b, *cg_tmp[0], *cg_tmp[1], *cg_tmp[2], *cg_tmp[3], coarse_max_it,
                                   coarse_rtol, coarse_atol, nullptr, x_zero);
    if (e0.dtype != TNG_F64) assign(e0, *cg_x);
}

void Multigrid::recurse(int level, tng_vector &x, const tng_vector &b, bool x_zero)
{
    if (level == 0) {
        coarse_solve(x, b, x_zero);
        return;
    }
    Level &L = levels[level];
    pre[level].solve(*L.A, x, b, x_zero);
    apply(*L.A, x, *L.tmp);
    lincomb(*L.tmp, {1.0, -1.0}, {&b, L.tmp});
    Level &C = levels[level - 1];
    apply(*L.R, *L.tmp, *C.r);
    assign(*C.e, 0.0);
    recurse(level - 1, *C.e, *C.r, true);
    apply(*L.P, *C.e, x);  // Add mode
    post[level].solve(*L.A, x, b, false);
}

void Multigrid::solve(tng_vector &x, const tng_vector &b, bool x_zero)
{
    for (int c = 0; c < num_cycles; ++c) recurse((int)levels.size() - 1, x, b, x_zero && c == 0);
}

// ---------------------------------------------------------------------------
// Stokes solver (stokes_solver.hpp:143-691)
// ---------------------------------------------------------------------------
StokesSolver::StokesSolver(tng_context *ctx, int min_level, int max_level, int radial_extra, double r_min,
                           double r_max, int lat_sdr, int rad_sdr, int bc_cmb_, int bc_surf_,
                           const tng_stokes_params &prm)
    : params(prm), bc_cmb(bc_cmb_), bc_surf(bc_surf_)
{
    TNG_REQUIRE(min_level >= 1 && max_level >= min_level + 1, "need at least two levels (velocity and pressure)");
    TNG_REQUIRE(prm.krylov_restart >= 1 && prm.krylov_max_iterations >= 0, "invalid Krylov parameters");
    TNG_REQUIRE(prm.viscous_pc_chebyshev_order >= 1, "Chebyshev order must be at least 1");
    TNG_REQUIRE(prm.viscous_pc_dtype == TNG_F64 || prm.viscous_pc_dtype == TNG_F32, "invalid multigrid precision");
    TNG_REQUIRE(prm.gca >= 0 && prm.gca <= 2,
                "gca must be 0 (re-discretised), 1 (Galerkin on all coarse levels) or 2 (adaptive Galerkin)");
    TNG_REQUIRE(prm.viscous_pc_smoother == 0 || prm.viscous_pc_smoother == 1,
                "viscous_pc_smoother must be 0 (Chebyshev) or 1 (damped Jacobi)");
    TNG_REQUIRE(prm.krylov_orthogonalization == 0 || prm.krylov_orthogonalization == 1,
                "krylov_orthogonalization must be 0 (modified Gram-Schmidt) or 1 (CGS2)");
    TNG_REQUIRE(prm.rotation_correction == 0 ||
                    (prm.rotation_correction == 1 && bc_cmb_ == TNG_FREESLIP && bc_surf_ == TNG_FREESLIP),
                "rotation_correction needs free slip on both surfaces");
    TNG_REQUIRE(prm.coarse_solver == 0 || prm.coarse_solver == 1, "coarse_solver must be 0 (PCG) or 1 (direct)");
    TNG_REQUIRE(prm.rotation_correction == 0 || prm.krylov_restart <= 64,
                "rotation_correction supports krylov_restart <= 64");
    for (int l = min_level; l <= max_level; ++l) {
        meshes.emplace_back(new_mesh(ctx, l, l + radial_extra, r_min, r_max, lat_sdr, rad_sdr));
        meshes.back()->solver_owned = true;
    }
    const int nl = (int)meshes.size();
    vel = meshes[nl - 1].get();
    pre_mesh = meshes[nl - 2].get();
    const tng_dtype mgt = (tng_dtype)prm.viscous_pc_dtype;

    // viscosity on every level (FP64), plus the multigrid-precision copies
    for (int l = 0; l < nl; ++l) eta.push_back(vec(pool, meshes[l].get(), TNG_F64, 1));
    for (int l = 0; l < nl; ++l) eta_mg.push_back(mgt == TNG_F64 ? eta[l] : vec(pool, meshes[l].get(), mgt, 1));
    for (int l = 0; l < nl; ++l) assign(*eta[l], 1.0);

    // Stokes operator blocks (FP64)
    A.reset(new_epsdivdiv(vel, eta[nl - 1], bc_cmb, bc_surf, false));
    Bt.reset(new_two_level(OP_GRADIENT, vel, pre_mesh, bc_cmb, bc_surf));
    B.reset(new_two_level(OP_DIVERGENCE, vel, pre_mesh, bc_cmb, bc_surf));

    // Galerkin coarse operators (FP64 stored matrices), assembled top-down in viscosity_updated()
    if (prm.gca >= 1) {
        TNG_REQUIRE(meshes[0]->nr >= 3, "GCA needs at least two radial layers on the coarsest level");
        for (int l = 0; l < nl - 1; ++l)
            gca64.emplace_back(new_epsdivdiv(meshes[l].get(), eta[l], bc_cmb, bc_surf, false));
    }

    // multigrid hierarchy in the multigrid precision
    mg.num_cycles = prm.viscous_pc_num_vcycles;
    mg.levels.resize(nl);
    for (int l = 0; l < nl; ++l) {
        Multigrid::Level &L = mg.levels[l];
        tng_domain *m = meshes[l].get();
        if (l == nl - 1 && mgt == TNG_F64) {
            L.A = A.get();
        } else if (galerkin_level(l) && mgt == TNG_F64) {
            L.A = gca64[l].get();
        } else {
            mg_ops.emplace_back(new_epsdivdiv(m, eta_mg[l], bc_cmb, bc_surf, false));
            L.A = mg_ops.back().get();
        }
        L.tmp = vec(pool, m, mgt, 3);
        L.inv_diag = vec(pool, m, mgt, 3);
        if (l < nl - 1) {
            L.r = vec(pool, m, mgt, 3);
            L.e = vec(pool, m, mgt, 3);
        }
        if (l > 0) {
            mg_ops.emplace_back(new_two_level(OP_PROLONGATION, m, meshes[l - 1].get(), TNG_NEUMANN, TNG_NEUMANN));
            L.P = mg_ops.back().get();
            L.P->apply_mode = TNG_ADD;
            mg_ops.emplace_back(new_two_level(OP_RESTRICTION, m, meshes[l - 1].get(), TNG_NEUMANN, TNG_NEUMANN));
            L.R = mg_ops.back().get();
        }
    }
    mg.pre.resize(nl);
    mg.post.resize(nl);
    // spare solution vectors for the fused Chebyshev sub-steps (one-column
    // apply kernels: 64 or 128 radial cells); the cheb_fuse setting 0 turns them off
    const bool fuse = vel->ctx->tuning.cheb_fuse != 0;
    for (int l = 0; l < nl; ++l) {
        const int ncells = meshes[l]->nr - 1;
        if (fuse && (ncells == 64 || ncells == 128) && meshes[l]->rad_sdr == 0)
            mg.pre[l].xb = mg.post[l].xb = vec(pool, meshes[l].get(), mgt, 3);
    }
    for (int l = 0; l < nl; ++l)
        for (Chebyshev *s : {&mg.pre[l], &mg.post[l]}) {
            s->jacobi = prm.viscous_pc_smoother == 1;
            s->order = prm.viscous_pc_chebyshev_order;
            s->iterations = prm.viscous_pc_num_smoothing_steps_prepost;
            s->power_iterations = prm.viscous_pc_num_power_iterations;
            s->inv_diag = mg.levels[l].inv_diag;
            s->d = vec(pool, meshes[l].get(), mgt, 3);
            s->z = vec(pool, meshes[l].get(), mgt, 3);
            s->seed = prm.seed + 1000 * (uint64_t)l + (s == &mg.post[l] ? 1 : 0);
        }
    // FP64 coarsest solve
    mg.coarse_max_it = prm.coarse_max_iterations;
    mg.use_direct = prm.coarse_solver == 1;
    mg.coarse_rtol = prm.coarse_relative_tolerance;
    mg.coarse_atol = prm.coarse_absolute_tolerance;
    if (mgt == TNG_F64) {
        mg.coarse_op = mg.levels[0].A;
    } else if (prm.gca >= 1) {
        mg.coarse_op = gca64[0].get();
        mg.cg_x = vec(pool, meshes[0].get(), TNG_F64, 3);
        mg.cg_b = vec(pool, meshes[0].get(), TNG_F64, 3);
    } else {
        coarse_op64.reset(new_epsdivdiv(meshes[0].get(), eta[0], bc_cmb, bc_surf, false));
        mg.coarse_op = coarse_op64.get();
        mg.cg_x = vec(pool, meshes[0].get(), TNG_F64, 3);
        mg.cg_b = vec(pool, meshes[0].get(), TNG_F64, 3);
    }
    for (int i = 0; i < 4; ++i) mg.cg_tmp[i] = vec(pool, meshes[0].get(), TNG_F64, 3);

    // Schur complement approximation: lumped (1/eta)-weighted pressure mass
    k_pm = vec(pool, pre_mesh, TNG_F64, 1);
    pmass.reset(new_kmass(pre_mesh, k_pm));
    inv_lumped_pmass = vec(pool, pre_mesh, TNG_F64, 1);
    ones_p = vec(pool, pre_mesh, TNG_F64, 1);
    assign(*ones_p, 1.0);

    // block-triangular preconditioner and FGMRES work space (FP64)
    prec_tmp_u = vec(pool, vel, TNG_F64, 3);
    if (mgt != TNG_F64) {
        mg_x = vec(pool, vel, mgt, 3);
        mg_b = vec(pool, vel, mgt, 3);
    }
    const int nt = 2 * prm.krylov_restart + 4;
    for (int i = 0; i < nt; ++i) fg.push_back({vec(pool, vel, TNG_F64, 3), vec(pool, pre_mesh, TNG_F64, 1)});
    ones_v = nullptr;
    num_dofs_pressure = pre_mesh->num_owned;
    viscosity_updated();
}

void StokesSolver::set_viscosity(const tng_vector &eta_fine)
{
    TNG_REQUIRE(eta_fine.mesh == vel && eta_fine.ncomp == 1, "viscosity must be a scalar field on the velocity mesh");
    const int nl = (int)meshes.size();
    assign(*eta[nl - 1], eta_fine);
    refresh_viscosity_hierarchy();
}

// coarse levels: injection of the velocity-level viscosity (nested nodes), then
// every derived quantity rebuilt
void StokesSolver::refresh_viscosity_hierarchy()
{
    const int nl = (int)meshes.size();
    for (int l = nl - 2; l >= 0; --l) launch_inject(*eta[l], *eta[l + 1], stream_of(*eta[l]));
    viscosity_updated();
}

void StokesSolver::viscosity_updated()
{
    rc_dirty = true;
    const int nl = (int)meshes.size();
    for (int l = 0; l < nl; ++l)
        if (eta_mg[l] != eta[l]) assign(*eta_mg[l], *eta[l]);
    A->btab_valid = false;
    if (coarse_op64) coarse_op64->btab_valid = false;
    for (auto &op : mg_ops)
        if (op->kind == OP_EPSDIVDIV) op->btab_valid = false;
    if (params.gca >= 1) {
        // TwoGridGCA top-down (stokes_solver.hpp:375-388): the finest step
        // uses the Neumann local matrices of the velocity operator.  gca = 2
        // (stokes_solver.hpp:278-287, 361-365): coarsest cells flagged from
        // the gradient of the velocity-level viscosity keep the Galerkin
        // matrices of all their descendants, the rest the re-discretised ones.
        hipStream_t s = vel->ctx->stream;
        if (params.gca == 2) gca_collect(*eta[nl - 1], nl - 1, *meshes[0], gca_flags, s);
        const bool exact = params.gca_exact != 0;
        for (int l = nl - 2; l >= 0; --l) {
            const tng_operator &fine = l == nl - 2 ? *A : *gca64[l + 1];
            if (params.gca == 2)
                gca_assemble_selective(fine, *gca64[l], l, *meshes[0], gca_flags, s, exact);
            else
                gca_assemble(fine, *gca64[l], s, exact);
            // coarsest-only: the finer level's matrices were an intermediate
            if (params.gca_coarsest_only && l + 1 < nl - 1) {
                TNG_HIP_CHECK(hipStreamSynchronize(s));
                gca64[l + 1]->d_lmat.release();
                gca64[l + 1]->stored = false;
            }
        }
        if (params.viscous_pc_dtype != TNG_F64)
            for (int l = 0; l < nl - 1; ++l)
                if (galerkin_level(l))
                    copy_stored_matrices(*mg.levels[l].A, *gca64[l], (tng_dtype)params.viscous_pc_dtype, s);
    }
    // inverse diagonals (stokes_solver.hpp:390-417)
    for (int l = 0; l < nl; ++l) {
        Multigrid::Level &L = mg.levels[l];
        tng_vector &ones = *L.tmp;
        assign(ones, 1.0);
        L.A->diagonal = true;
        apply(*L.A, ones, *L.inv_diag);
        L.A->diagonal = false;
        launch_invert(*L.inv_diag, stream_of(*L.inv_diag));
    }
    for (int l = 0; l < nl; ++l) {
        mg.pre[l].need_estimate = true;
        mg.post[l].need_estimate = true;
    }
    if (mg.use_direct)
        mg.direct.build(*mg.coarse_op, *mg.cg_tmp[0], *mg.cg_tmp[1], params.viscous_pc_dtype != TNG_F64,
                        stream_of(*mg.cg_tmp[0]));
    // k_pm = 1 / eta at the pressure level; lumped diagonal, inverted (DiagonalSolver)
    assign(*k_pm, *eta[nl - 2]);
    launch_invert(*k_pm, stream_of(*k_pm));
    apply(*pmass, *ones_p, *inv_lumped_pmass);
    launch_invert(*inv_lumped_pmass, stream_of(*inv_lumped_pmass));
}

void StokesSolver::update_viscosity_velocity_level(const tng_vector &temp, double rmu)
{
    const int nl = (int)meshes.size();
    hipStream_t s = vel->ctx->stream;
    launch_viscosity_fk(*eta[nl - 1], temp, rmu, s);
    A->btab_valid = false;
    if (eta_mg[nl - 1] != eta[nl - 1]) assign(*eta_mg[nl - 1], *eta[nl - 1]);
    mg.levels[nl - 1].A->btab_valid = false;
    rc_dirty = true;
}

tng_vector *StokesSolver::make_vec(tng_domain *m, int ncomp) { return vec(pool, m, TNG_F64, ncomp); }

// K = [A B^T; B 0]: A (Replace, Skip) -> B^T (Add, Communicate) -> B (Replace, Communicate)
void StokesSolver::apply_stokes(const SVec &src, SVec &dst)
{
    A->apply_mode = TNG_REPLACE;
    A->comm_mode = TNG_SKIP_COMMUNICATION;
    Bt->apply_mode = TNG_ADD;
    Bt->comm_mode = TNG_COMMUNICATE_ADDITIVELY;
    B->apply_mode = TNG_REPLACE;
    B->comm_mode = TNG_COMMUNICATE_ADDITIVELY;
    apply(*A, *src.u, *dst.u);
    apply(*Bt, *src.p, *dst.u);
    apply(*B, *src.u, *dst.p);
    A->comm_mode = TNG_COMMUNICATE_ADDITIVELY;
    Bt->apply_mode = TNG_REPLACE;
}

// velocity block preconditioner: V-cycle(s) on A x = b, x as given
void StokesSolver::precondition_velocity(tng_vector &x, const tng_vector &b, bool x_zero)
{
    if (!mg_x) {
        mg.solve(x, b, x_zero);
        return;
    }
    if (x_zero) assign(*mg_x, 0.0);
    else assign(*mg_x, x);
    assign(*mg_b, b);
    mg.solve(*mg_x, *mg_b, x_zero);
    assign(x, *mg_x);
}

// block-triangular preconditioner (block_preconditioner_2x2.hpp:205-221)
void StokesSolver::precondition(SVec &x, const SVec &b, bool x_zero)
{
    assign(*x.p, *b.p);
    scale_in_place(*x.p, *inv_lumped_pmass);
    apply(*Bt, *x.p, *prec_tmp_u);  // Replace, Communicate
    lincomb(*prec_tmp_u, {1.0, 1.0}, {prec_tmp_u, b.u});
    lincomb(*x.p, {-1.0}, {x.p});
    precondition_velocity(*x.u, *prec_tmp_u, x_zero);
}

namespace {
void s_lincomb(StokesSolver::SVec &y, std::initializer_list<double> c,
               std::initializer_list<const StokesSolver::SVec *> x, double c0 = 0.0)
{
    const tng_vector *xu[3], *xp[3];
    double cs[3] = {0, 0, 0};
    int n = 0;
    for (const StokesSolver::SVec *v : x) {
        xu[n] = v->u;
        xp[n] = v->p;
        ++n;
    }
    n = 0;
    for (double v : c) cs[n++] = v;
    launch_lincomb(*y.u, (int)x.size(), cs, xu, c0, stream_of(*y.u));
    launch_lincomb(*y.p, (int)x.size(), cs, xp, c0, stream_of(*y.p));
}
double s_dot(const StokesSolver::SVec &a, const StokesSolver::SVec &b) { return dot(*a.u, *b.u) + dot(*a.p, *b.p); }
void s_assign(StokesSolver::SVec &y, double c)
{
    assign(*y.u, c);
    assign(*y.p, c);
}
void s_assign(StokesSolver::SVec &y, const StokesSolver::SVec &x)
{
    assign(*y.u, *x.u);
    assign(*y.p, *x.p);
}
bool s_nonfinite(const StokesSolver::SVec &x) { return has_nan_or_inf(*x.u) || has_nan_or_inf(*x.p); }
}  // namespace

// Rotation correction (free slip on both surfaces).  The three rigid
// rotations m_i = e_i x X are the near-null modes of the penalised viscous
// block (A m_i is of the order of the penalty eps, 1e-7), which the V-cycle
// does not resolve; left to FGMRES they stall the restarted iteration.  After
// the block preconditioner z = M v, with w = K z already formed:
//   c = E^{-1} Z^T (v - w),  z += Z c,  w += (K Z) c,   E = Z^T K Z (3 x 3),
// Z = [m_i / |m_i|; 0].  The operator and the stopping test are unchanged.
void StokesSolver::rotation_setup()
{
    hipStream_t s = vel->ctx->stream;
    const int m = params.krylov_restart;
    if (rc_z.empty()) {
        for (int i = 0; i < 3; ++i) {
            rc_z.push_back({vec(pool, vel, TNG_F64, 3), vec(pool, pre_mesh, TNG_F64, 1)});
            rc_kz.push_back({vec(pool, vel, TNG_F64, 3), vec(pool, pre_mesh, TNG_F64, 1)});
        }
        // device pointer tables: per basis slot j: [Z0, Z1, Z2, V_j]; [KZ0..2]; per j: [Z_j]; [w]
        const int offV = 1, offZ = offV + m + 1, idxW = offZ + m;
        auto put = [](std::vector<double *> &t, const SVec &v) {
            for (int c = 0; c < 3; ++c) t.push_back((double *)v.u->comp(c));
            t.push_back((double *)v.p->comp(0));
        };
        std::vector<double *> tab;
        for (int j = 0; j < m; ++j) {
            for (int i = 0; i < 3; ++i) put(tab, rc_z[i]);
            put(tab, fg[offV + j]);
        }
        for (int i = 0; i < 3; ++i) put(tab, rc_kz[i]);
        for (int j = 0; j < m; ++j) put(tab, fg[offZ + j]);
        put(tab, fg[idxW]);
        rc_ptrs.alloc(tab.size() * sizeof(double *));
        TNG_HIP_CHECK(hipMemcpyAsync(rc_ptrs.p, tab.data(), tab.size() * sizeof(double *), hipMemcpyHostToDevice, s));
        TNG_HIP_CHECK(hipStreamSynchronize(s));
        const size_t np = gs_partials_doubles();
        if (gs_part.bytes < np * sizeof(double)) gs_part.alloc(np * sizeof(double));
        if (gs_out.bytes < 128 * sizeof(double)) gs_out.alloc(128 * sizeof(double));
    }
    for (int i = 0; i < 3; ++i) {
        launch_rotation_field(*rc_z[i].u, i, s);
        assign(*rc_z[i].p, 0.0);
        const double nrm = std::sqrt(dot(*rc_z[i].u, *rc_z[i].u));
        lincomb(*rc_z[i].u, {1.0 / nrm}, {rc_z[i].u});
        apply_stokes(rc_z[i], rc_kz[i]);
    }
    double E[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) E[3 * i + j] = s_dot(rc_z[i], rc_kz[j]);
    // LU with partial pivoting
    for (int i = 0; i < 9; ++i) rc_lu[i] = E[i];
    for (int k = 0; k < 3; ++k) {
        int p = k;
        for (int i = k + 1; i < 3; ++i)
            if (std::abs(rc_lu[3 * i + k]) > std::abs(rc_lu[3 * p + k])) p = i;
        rc_piv[k] = p;
        if (p != k)
            for (int c = 0; c < 3; ++c) std::swap(rc_lu[3 * k + c], rc_lu[3 * p + c]);
        TNG_REQUIRE(rc_lu[3 * k + k] != 0.0 && std::isfinite(rc_lu[3 * k + k]),
                    "rotation correction: Z^T K Z is singular");
        for (int i = k + 1; i < 3; ++i) {
            rc_lu[3 * i + k] /= rc_lu[3 * k + k];
            for (int c = k + 1; c < 3; ++c) rc_lu[3 * i + c] -= rc_lu[3 * i + k] * rc_lu[3 * k + c];
        }
    }
    rc_dirty = false;
}

void StokesSolver::rotation_correct(int j, SVec &v, SVec &z, SVec &w)
{
    (void)v;
    hipStream_t s = stream_of(*w.u);
    const int m = params.krylov_restart;
    double *const *tab = rc_ptrs.as<double *const>();
    double *const *t_dots = tab + 16 * j;           // [Z0, Z1, Z2, V_j]
    double *const *t_kz = tab + 16 * m;              // [KZ0, KZ1, KZ2]
    double *const *t_zj = tab + 16 * m + 12 + 4 * j; // [Z_j]
    // the current storage of z: the fused Chebyshev smoother exchanges the
    // storage of the vectors it updates (uploaded before the synchronisation
    // below, which also keeps zp alive until the copy is done)
    double *zp[4] = {(double *)z.u->comp(0), (double *)z.u->comp(1), (double *)z.u->comp(2), (double *)z.p->comp(0)};
    TNG_HIP_CHECK(hipMemcpyAsync(rc_ptrs.as<double *>() + 16 * m + 12 + 4 * j, zp, sizeof(zp), hipMemcpyHostToDevice, s));
    GsArgs ga{};
    ga.V = t_dots;
    for (int c = 0; c < 3; ++c) ga.w[c] = (const double *)w.u->comp(c);
    ga.w[3] = (const double *)w.p->comp(0);
    ga.mask_v = vel->d_omask.as<uint8_t>();
    ga.mask_p = pre_mesh->d_omask.as<uint8_t>();
    ga.nv = vel->nalloc;
    ga.np = pre_mesh->nalloc;
    ga.j = 3;
    gs_dots(ga, gs_part.as<double>(), gs_out.as<double>(), s);
    const int B = gs_bucket(3);
    double d[2 * 8 + 3];
    TNG_HIP_CHECK(hipMemcpyAsync(d, gs_out.p, sizeof(double) * (2 * B + 3), hipMemcpyDeviceToHost, s));
    TNG_HIP_CHECK(hipStreamSynchronize(s));
    // t = Z^T v - Z^T w, c = E^{-1} t
    double c[3];
    for (int i = 0; i < 3; ++i) c[i] = d[i] - d[B + i];
    for (int k = 0; k < 3; ++k) std::swap(c[k], c[rc_piv[k]]);
    for (int i = 1; i < 3; ++i)
        for (int k = 0; k < i; ++k) c[i] -= rc_lu[3 * i + k] * c[k];
    for (int i = 2; i >= 0; --i) {
        for (int k = i + 1; k < 3; ++k) c[i] -= rc_lu[3 * i + k] * c[k];
        c[i] /= rc_lu[3 * i + i];
    }
    GsAxpy cf{};
    for (int i = 0; i < 3; ++i) cf.c[i] = c[i];
    gs_axpy(t_dots, 3, cf, t_zj, ga.nv, ga.np, s);
    gs_axpy(t_kz, 3, cf, tab + 16 * m + 12 + 4 * m, ga.nv, ga.np, s);
}

// FGMRES(m) (fgmres.hpp:109-373)
void StokesSolver::fgmres(SVec &x, const SVec &b, tng_solve_report &rep)
{
    const int m = params.krylov_restart;
    const double rtol = params.krylov_relative_tolerance, atol = params.krylov_absolute_tolerance;
    SVec &r = fg[0];
    apply_stokes(x, r);
    s_lincomb(r, {1.0, -1.0}, {&b, &r});
    double beta0 = std::sqrt(s_dot(r, r));
    const double initial = beta0;
    rep.initial_residual = initial;
    rep.final_residual = beta0;
    rep.relative_residual = 1.0;
    record(rep, beta0);
    if (beta0 <= atol) {
        rep.converged = 1;
        return;
    }
    const int offV = 1, offZ = offV + m + 1, idxW = offZ + m, idxAcc = idxW + 1;
    std::vector<double> H((size_t)(m + 1) * m), g(m + 1), cs(m + 1), sn(m + 1), y, hcol(m + 2);
    TNG_REQUIRE(m + 2 <= 256, "krylov_restart above 254 is not supported");
    double *hdev = device_slots(*x.u->mesh->ctx, 256);
    auto h = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
    int total = 0;
    while (total < params.krylov_max_iterations) {
        SVec &V0 = fg[offV];
        s_lincomb(V0, {1.0 / beta0}, {&r});
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        std::fill(cs.begin(), cs.end(), 0.0);
        std::fill(sn.begin(), sn.end(), 0.0);
        g[0] = beta0;
        int inner = 0;
        for (int j = 0; j < m && total < params.krylov_max_iterations; ++j, ++total) {
            SVec &vj = fg[offV + j], &zj = fg[offZ + j];
            rep.iterations += 1;  // preconditioned Arnoldi steps
            s_assign(zj, 0.0);
            precondition(zj, vj, true);
            if (s_nonfinite(zj)) {
                rep.nonfinite_preconditioner += 1;
                s_assign(zj, vj);  // skip the preconditioner in this iteration
            }
            SVec &w = fg[idxW];
            apply_stokes(zj, w);
            if (params.rotation_correction) rotation_correct(j, vj, zj, w);
            // modified Gram-Schmidt: h_ij = w . v_i, then w -= h_ij v_i, one
            // fused pass per i with the h values kept on the device; one copy
            // to the host per Arnoldi step
            for (int i = 0; i <= j + 1; ++i) {
                const SVec *vprev = i > 0 ? &fg[offV + i - 1] : nullptr;
                const SVec *vnext = i <= j ? &fg[offV + i] : nullptr;
                launch_mgs_step(*w.u, *w.p, vprev ? vprev->u : nullptr, vprev ? vprev->p : nullptr,
                                i > 0 ? hdev + i - 1 : nullptr, vnext ? vnext->u : nullptr,
                                vnext ? vnext->p : nullptr, hdev + i, stream_of(*w.u));
            }
            TNG_HIP_CHECK(hipMemcpyAsync(hcol.data(), hdev, sizeof(double) * (j + 2), hipMemcpyDeviceToHost,
                                         stream_of(*w.u)));
            TNG_HIP_CHECK(hipStreamSynchronize(stream_of(*w.u)));
            for (int i = 0; i <= j; ++i) h(i, j) = hcol[i];
            const double hj1 = std::sqrt(hcol[j + 1]);
            if (hj1 < std::numeric_limits<double>::epsilon() * initial) {
                h(j + 1, j) = 0;
                inner = j + 1;
                rep.breakdowns += 1;
                break;
            }
            h(j + 1, j) = hj1;
            if (hj1 > 0) s_lincomb(fg[offV + j + 1], {1.0 / hj1}, {&w});
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * h(i, j) + sn[i] * h(i + 1, j);
                h(i + 1, j) = -sn[i] * h(i, j) + cs[i] * h(i + 1, j);
                h(i, j) = t;
            }
            {
                const double a = h(j, j), bb = h(j + 1, j);
                const double rab = std::hypot(a, bb);
                cs[j] = rab == 0.0 ? 1.0 : a / rab;
                sn[j] = rab == 0.0 ? 0.0 : bb / rab;
                h(j, j) = rab;
                h(j + 1, j) = 0;
                if (std::abs(h(j, j)) < std::numeric_limits<double>::epsilon()) {
                    inner = j;
                    rep.breakdowns += 1;
                    break;
                }
                const double gj = g[j], gj1 = g[j + 1];
                g[j] = cs[j] * gj + sn[j] * gj1;
                g[j + 1] = -sn[j] * gj + cs[j] * gj1;
            }
            const double abs_res = std::abs(g[j + 1]);
            const double rel = abs_res / initial;
            record(rep, abs_res);
            if (!std::isfinite(abs_res)) {
                inner = j;
                ++total;
                rep.nonfinite_residual += 1;
                break;
            }
            inner = j + 1;
            if (rel <= rtol || abs_res < atol) break;
        }
        // back substitution and update x += sum y_i z_i
        y.assign(inner, 0.0);
        for (int row = inner - 1; row >= 0; --row) {
            double s = g[row];
            for (int col = row + 1; col < inner; ++col) s -= h(row, col) * y[col];
            y[row] = s / h(row, row);
        }
        SVec &acc = fg[idxAcc];
        s_assign(acc, 0.0);
        for (int i = 0; i < inner; ++i) s_lincomb(acc, {1.0, y[i]}, {&acc, &fg[offZ + i]});
        s_lincomb(x, {1.0, 1.0}, {&x, &acc});
        rep.restarts += 1;
        apply_stokes(x, r);
        s_lincomb(r, {1.0, -1.0}, {&b, &r});
        beta0 = std::sqrt(s_dot(r, r));
        rep.final_residual = beta0;
        rep.relative_residual = beta0 / initial;
        if (!std::isfinite(beta0)) {
            rep.nonfinite_residual += 1;
            return;
        }
        if (beta0 <= atol || beta0 / initial <= rtol) {
            rep.converged = 1;
            return;
        }
    }
}

// FGMRES(m) with the two-pass Gram-Schmidt of kernels_gs.hip (classical
// Gram-Schmidt, each basis vector re-orthogonalised one step later).  Step j:
// z_j = M u_j (u_j the tentative basis vector), w = K z_j, one multi-dot
// gives s = Q^T u_j, u_j.u_j, c = Q^T w, u_j.w, w.w; then
//   alpha = |u_j - Q s|,  q_j = (u_j - Q s) / alpha,
//   rho = q_j.w = (u_j.w - s.c) / alpha,
//   nu = |w - Q c - rho q_j| = sqrt(w.w - c.c - rho^2),
//   u_{j+1} = (w - Q c - rho q_j) / nu   (second pass).
// Arnoldi: K z_j = Q c + rho q_j + nu u_{j+1}; when u_{j+1} = Q' s' + alpha'
// q_{j+1} is known one step later, column j of H becomes (c + nu s', rho +
// nu s'_j, nu alpha').  The residual estimate of step j uses column j before
// that correction (which is of the order of the loss of orthogonality); the
// Givens factorisation is redone from the corrected H every step.
void StokesSolver::fgmres_cgs2(SVec &x, const SVec &b, tng_solve_report &rep)
{
    const int m = params.krylov_restart;
    const double rtol = params.krylov_relative_tolerance, atol = params.krylov_absolute_tolerance;
    TNG_REQUIRE(m <= 32, "CGS2 orthogonalisation supports krylov_restart <= 32");
    hipStream_t s = stream_of(*x.u);
    SVec &r = fg[0];
    apply_stokes(x, r);
    s_lincomb(r, {1.0, -1.0}, {&b, &r});
    double beta0 = std::sqrt(s_dot(r, r));
    const double initial = beta0;
    rep.initial_residual = initial;
    rep.final_residual = beta0;
    rep.relative_residual = 1.0;
    record(rep, beta0);
    if (beta0 <= atol) {
        rep.converged = 1;
        return;
    }
    const int offV = 1, offZ = offV + m + 1, idxW = offZ + m, idxAcc = idxW + 1;
    // device table of component pointers: the basis slots V_0..V_m, then the
    // preconditioned directions Z_0..Z_{m-1}, then the update accumulator
    {
        std::vector<double *> tab(4 * (size_t)(2 * m + 2));
        auto put = [&](int slot, const SVec &v) {
            for (int c = 0; c < 3; ++c) tab[4 * (size_t)slot + c] = (double *)v.u->comp(c);
            tab[4 * (size_t)slot + 3] = (double *)v.p->comp(0);
        };
        for (int i = 0; i <= m; ++i) put(i, fg[offV + i]);
        for (int i = 0; i < m; ++i) put(m + 1 + i, fg[offZ + i]);
        put(2 * m + 1, fg[idxAcc]);
        if (gs_table.bytes < tab.size() * sizeof(double *)) gs_table.alloc(tab.size() * sizeof(double *));
        TNG_HIP_CHECK(hipMemcpyAsync(gs_table.p, tab.data(), tab.size() * sizeof(double *), hipMemcpyHostToDevice, s));
        const size_t np = gs_partials_doubles();
        if (gs_part.bytes < np * sizeof(double)) gs_part.alloc(np * sizeof(double));
        if (gs_out.bytes < 128 * sizeof(double)) gs_out.alloc(128 * sizeof(double));
    }
    GsArgs ga{};
    ga.V = gs_table.as<double *const>();
    {
        const SVec &w = fg[idxW];
        for (int c = 0; c < 3; ++c) ga.w[c] = (const double *)w.u->comp(c);
        ga.w[3] = (const double *)w.p->comp(0);
    }
    ga.mask_v = vel->d_omask.as<uint8_t>();
    ga.mask_p = pre_mesh->d_omask.as<uint8_t>();
    ga.nv = vel->nalloc;
    ga.np = pre_mesh->nalloc;
    std::vector<double> H((size_t)(m + 1) * m), R, g(m + 1), cs(m + 1), sn(m + 1), y, dots(128);
    auto h = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
    // Givens QR of columns 0..ncol-1 of H into R, g from g0 e_1; returns false
    // on a vanishing diagonal (stagnation)
    auto factor = [&](int ncol, double g0, int *bad) -> bool {
        R = H;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = g0;
        auto rr = [&](int i, int j) -> double & { return R[(size_t)i * m + j]; };
        for (int j = 0; j < ncol; ++j) {
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * rr(i, j) + sn[i] * rr(i + 1, j);
                rr(i + 1, j) = -sn[i] * rr(i, j) + cs[i] * rr(i + 1, j);
                rr(i, j) = t;
            }
            const double a = rr(j, j), bb = rr(j + 1, j);
            const double rab = std::hypot(a, bb);
            cs[j] = rab == 0.0 ? 1.0 : a / rab;
            sn[j] = rab == 0.0 ? 0.0 : bb / rab;
            rr(j, j) = rab;
            rr(j + 1, j) = 0;
            if (std::abs(rr(j, j)) < std::numeric_limits<double>::epsilon()) {
                *bad = j;
                return false;
            }
            const double gj = g[j], gj1 = g[j + 1];
            g[j] = cs[j] * gj + sn[j] * gj1;
            g[j + 1] = -sn[j] * gj + cs[j] * gj1;
        }
        return true;
    };
    int total = 0;
    while (total < params.krylov_max_iterations) {
        SVec &V0 = fg[offV];
        s_lincomb(V0, {1.0 / beta0}, {&r});
        std::fill(H.begin(), H.end(), 0.0);
        double g0 = beta0;
        double nu_prev = 0.0;
        int inner = 0;
        for (int j = 0; j < m && total < params.krylov_max_iterations; ++j, ++total) {
            SVec &uj = fg[offV + j], &zj = fg[offZ + j];
            rep.iterations += 1;
            s_assign(zj, 0.0);
            precondition(zj, uj, true);
            if (s_nonfinite(zj)) {
                rep.nonfinite_preconditioner += 1;
                s_assign(zj, uj);
            }
            apply_stokes(zj, fg[idxW]);
            if (params.rotation_correction) rotation_correct(j, uj, zj, fg[idxW]);
            ga.j = j;
            gs_dots(ga, gs_part.as<double>(), gs_out.as<double>(), s);
            const int B = gs_bucket(j);
            TNG_HIP_CHECK(hipMemcpyAsync(dots.data(), gs_out.p, sizeof(double) * (2 * B + 3), hipMemcpyDeviceToHost, s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));
            const double *sv = dots.data(), *cv = dots.data() + B;
            const double uu = dots[2 * B], uw = dots[2 * B + 1], ww = dots[2 * B + 2];
            double ss = 0.0, sc = 0.0, cc = 0.0;
            for (int i = 0; i < j; ++i) {
                ss += sv[i] * sv[i];
                sc += sv[i] * cv[i];
                cc += cv[i] * cv[i];
            }
            const double alpha = std::sqrt(std::max(uu - ss, 0.0));
            if (!(alpha > 0.0) || !std::isfinite(alpha)) {
                inner = j;
                rep.breakdowns += 1;
                break;
            }
            // delayed correction of column j - 1 (u_j = Q s + alpha q_j)
            if (j == 0) {
                g0 = beta0 * alpha;  // r = beta0 u_0 = beta0 alpha q_0
            } else {
                for (int i = 0; i < j; ++i) h(i, j - 1) += nu_prev * sv[i];
                h(j, j - 1) = nu_prev * alpha;
            }
            const double rho = (uw - sc) / alpha;
            const double nu2 = ww - cc - rho * rho;
            GsCoef cf{};
            for (int i = 0; i < j; ++i) {
                cf.s[i] = sv[i];
                cf.c[i] = cv[i] - (rho / alpha) * sv[i];
            }
            cf.inv_alpha = 1.0 / alpha;
            cf.r = rho / alpha;
            double nu;
            if (nu2 > 1e-8 * ww) {
                nu = std::sqrt(nu2);
                cf.inv_nu = 1.0 / nu;
                gs_update(ga, cf, s);
            } else {
                // cancellation in w.w - c.c - rho^2: form the projection, then its norm
                cf.inv_nu = 1.0;
                gs_update(ga, cf, s);
                SVec &un = fg[offV + j + 1];
                nu = std::sqrt(std::max(s_dot(un, un), 0.0));
                if (nu > 0.0) s_lincomb(un, {1.0 / nu}, {&un});
            }
            for (int i = 0; i < j; ++i) h(i, j) = cv[i];
            h(j, j) = rho;
            h(j + 1, j) = nu;
            nu_prev = nu;
            if (nu < std::numeric_limits<double>::epsilon() * initial) {
                h(j + 1, j) = 0;
                inner = j + 1;
                rep.breakdowns += 1;
                int bad = -1;
                if (!factor(inner, g0, &bad)) inner = bad;
                break;
            }
            int bad = -1;
            if (!factor(j + 1, g0, &bad)) {
                inner = bad;
                rep.breakdowns += 1;
                break;
            }
            const double abs_res = std::abs(g[j + 1]);
            const double rel = abs_res / initial;
            record(rep, abs_res);
            if (!std::isfinite(abs_res)) {
                inner = j;
                ++total;
                rep.nonfinite_residual += 1;
                break;
            }
            inner = j + 1;
            if (rel <= rtol || abs_res < atol) break;
        }
        // back substitution on the factor of the last (uncorrected) column set
        if (inner > 0) {
            int bad = -1;
            if (!factor(inner, g0, &bad)) inner = bad;
        }
        auto rr = [&](int i, int j) -> double { return R[(size_t)i * m + j]; };
        y.assign(inner, 0.0);
        for (int row = inner - 1; row >= 0; --row) {
            double sum = g[row];
            for (int col = row + 1; col < inner; ++col) sum -= rr(row, col) * y[col];
            y[row] = sum / rr(row, row);
        }
        // x += Z y, the directions summed in runs of 8 per pass.  The table's
        // Z entries are refreshed first: the fused Chebyshev smoother exchanges
        // the storage of the vectors it updates, so a preconditioned
        // direction's components may have moved since the table was built.
        SVec &acc = fg[idxAcc];
        s_assign(acc, 0.0);
        {
            std::vector<double *> zt(4 * (size_t)(m + 1));
            for (int i = 0; i <= m; ++i) {
                const SVec &v = i < m ? fg[offZ + i] : acc;
                for (int c = 0; c < 3; ++c) zt[4 * (size_t)i + c] = (double *)v.u->comp(c);
                zt[4 * (size_t)i + 3] = (double *)v.p->comp(0);
            }
            TNG_HIP_CHECK(hipMemcpyAsync(gs_table.as<double *>() + 4 * (m + 1), zt.data(), zt.size() * sizeof(double *),
                                         hipMemcpyHostToDevice, s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));  // zt is released at the end of this block
            double *const *tabd = gs_table.as<double *const>();
            for (int i0 = 0; i0 < inner; i0 += 8) {
                GsAxpy cf{};
                const int k = std::min(8, inner - i0);
                for (int i = 0; i < k; ++i) cf.c[i] = y[i0 + i];
                gs_axpy(tabd + 4 * (m + 1 + i0), k, cf, tabd + 4 * (2 * m + 1), ga.nv, ga.np, s);
            }
        }
        s_lincomb(x, {1.0, 1.0}, {&x, &acc});
        rep.restarts += 1;
        apply_stokes(x, r);
        s_lincomb(r, {1.0, -1.0}, {&b, &r});
        beta0 = std::sqrt(s_dot(r, r));
        rep.final_residual = beta0;
        rep.relative_residual = beta0 / initial;
        if (!std::isfinite(beta0)) {
            rep.nonfinite_residual += 1;
            return;
        }
        if (beta0 <= atol || beta0 / initial <= rtol) {
            rep.converged = 1;
            return;
        }
    }
}

void StokesSolver::record(tng_solve_report &rep, double abs_res)
{
    if (rep.history && rep.history_len < rep.history_capacity) rep.history[rep.history_len] = abs_res;
    rep.history_len += 1;
}

void StokesSolver::solve(SVec &x, const SVec &b, tng_solve_report &rep)
{
    TNG_REQUIRE(x.u->mesh == vel && b.u->mesh == vel && x.p->mesh == pre_mesh && b.p->mesh == pre_mesh,
                "solve: vectors must live on the solver's velocity and pressure meshes");
    TNG_REQUIRE(x.u->dtype == TNG_F64 && x.p->dtype == TNG_F64 && b.u->dtype == TNG_F64 && b.p->dtype == TNG_F64,
                "solve: the Stokes vectors are FP64");
    double *hist = rep.history;
    const int cap = rep.history_capacity;
    std::memset(&rep, 0, sizeof(rep));
    rep.history = hist;
    rep.history_capacity = cap;
    mg.coarse_iterations = 0;
    if (params.rotation_correction && rc_dirty) rotation_setup();
    if (params.krylov_orthogonalization == 1) fgmres_cgs2(x, b, rep);
    else fgmres(x, b, rep);
    rep.coarse_iterations = mg.coarse_iterations;
    // subtract the pressure mean over the unique pressure nodes (stokes_solver.hpp:680-689)
    const double avg = dot(*x.p, *ones_p) / (double)num_dofs_pressure;
    lincomb(*x.p, {1.0}, {x.p}, -avg);
}

void StokesSolver::adjoint_gradient(const tng_vector &u, const tng_vector &dJdu, tng_vector &g, tng_solve_report &rep)
{
    TNG_REQUIRE(u.mesh == vel && dJdu.mesh == vel && g.mesh == vel && u.ncomp == 3 && dJdu.ncomp == 3 &&
                    g.ncomp == 1 && u.dtype == TNG_F64 && dJdu.dtype == TNG_F64 && g.dtype == TNG_F64,
                "adjoint: u, dJ/du (FP64 velocity) and g (FP64 scalar) on the velocity mesh");
    if (!lam_u) {
        lam_u = vec(pool, vel, TNG_F64, 3);
        lam_p = vec(pool, pre_mesh, TNG_F64, 1);
        zero_p = vec(pool, pre_mesh, TNG_F64, 1);
    }
    assign(*lam_u, 0.0);
    assign(*lam_p, 0.0);
    assign(*zero_p, 0.0);
    SVec x{lam_u, lam_p}, b{const_cast<tng_vector *>(&dJdu), zero_p};
    solve(x, b, rep);  // K is symmetric: K^T lam = K lam
    epsdivdiv_log_k_gradient(*A, *lam_u, u, g, stream_of(g));
    lincomb(g, {-1.0}, {&g});
}

void StokesSolver::buoyancy_rhs(const tng_vector &temp, double rayleigh, tng_vector &f_u)
{
    TNG_REQUIRE(temp.mesh == vel && f_u.mesh == vel && temp.ncomp == 1 && f_u.ncomp == 3 && temp.dtype == TNG_F64 &&
                    f_u.dtype == TNG_F64,
                "buoyancy: FP64 temperature (scalar) and RHS (velocity) on the velocity mesh");
    if (!vmass) {
        vmass.reset(new tng_operator());
        vmass->kind = OP_VECTOR_MASS;
        vmass->mesh = vel;
        rhs_tmp = vec(pool, vel, TNG_F64, 3);
    }
    launch_buoyancy(temp, rayleigh, *rhs_tmp, stream_of(f_u));
    apply(*vmass, *rhs_tmp, f_u);
    enforce_rhs(f_u);
}

// RHS boundary enforcement as the app applies it (stokes_solver.hpp:664-674):
// the Dirichlet and free-slip flags are those of the FIRST boundary carrying
// that condition (grid/shell/bit_masks.hpp:68-76), CMB before surface.
void StokesSolver::enforce_rhs(tng_vector &f_u)
{
    auto first = [&](int bc) -> uint8_t {
        if (bc_cmb == bc) return SBF_CMB;
        if (bc_surf == bc) return SBF_SURFACE;
        return 0;
    };
    launch_rhs_enforce(f_u, first(TNG_DIRICHLET), first(TNG_FREESLIP), stream_of(f_u));
}

}  // namespace tng

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

template <class F> tng_status sguard(tng_context *ctx, F &&f)
{
    try {
        TNG_HIP_CHECK(hipSetDevice(ctx->device));
        f();
        return TNG_OK;
    } catch (const tng::Error &e) {
        ctx->last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        ctx->last_error = "host allocation failed";
        return TNG_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        return TNG_DEVICE_ERROR;
    }
}

struct Scratch {
    std::vector<std::unique_ptr<tng_vector>> v;
    tng_vector *like(const tng_vector &x)
    {
        v.emplace_back(tng::new_vector(x.mesh, x.dtype, x.ncomp));
        return v.back().get();
    }
};

void check_square(const tng_operator &op, const tng_vector &x, const tng_vector &b)
{
    TNG_REQUIRE(op.kind == OP_EPSDIVDIV || op.kind == OP_KMASS_LUMPED, "the solver needs a square operator");
    TNG_REQUIRE(x.mesh == op.mesh && b.mesh == op.mesh && x.ncomp == b.ncomp && x.dtype == b.dtype,
                "vectors must live on the operator's mesh with one shape");
}

}  // namespace

extern "C" {

tng_status tng_pcg_solve(tng_op op, tng_vec x, tng_vec b, int max_iterations, double rtol, double atol,
                         int *iterations, double *relative_residual)
{
    if (!op || !x || !b || max_iterations < 0) return TNG_INVALID_ARGUMENT;
    return sguard(op->mesh->ctx, [&] {
        check_square(*op, *x, *b);
        Scratch s;
        double rel = 0;
        const int it = tng::pcg_solve(*op, *x, *b, *s.like(*x), *s.like(*x), *s.like(*x), *s.like(*x), max_iterations,
                                      rtol, atol, &rel);
        if (iterations) *iterations = it;
        if (relative_residual) *relative_residual = rel;
    });
}

tng_status tng_power_iteration(tng_op op, tng_vec inv_diag, int iterations, unsigned long long seed, double *estimate)
{
    if (!op || !inv_diag || !estimate || iterations < 0) return TNG_INVALID_ARGUMENT;
    return sguard(op->mesh->ctx, [&] {
        check_square(*op, *inv_diag, *inv_diag);
        Scratch s;
        *estimate = tng::power_iteration(*op, *inv_diag, *s.like(*inv_diag), *s.like(*inv_diag), iterations, seed);
    });
}

tng_status tng_chebyshev_smooth(tng_op op, tng_vec inv_diag, int order, int iterations, double lambda_max_estimate,
                                tng_vec x, tng_vec b)
{
    if (!op || !inv_diag || !x || !b || order < 1 || iterations < 0) return TNG_INVALID_ARGUMENT;
    return sguard(op->mesh->ctx, [&] {
        check_square(*op, *x, *b);
        TNG_REQUIRE(inv_diag->mesh == x->mesh && inv_diag->ncomp == x->ncomp && inv_diag->dtype == x->dtype,
                    "inverse diagonal must match the vectors");
        Scratch s;
        tng::Chebyshev c;
        c.order = order;
        c.iterations = iterations;
        c.inv_diag = inv_diag;
        c.d = s.like(*x);
        c.z = s.like(*x);
        c.need_estimate = false;
        c.lmax_estimate = lambda_max_estimate;
        c.solve(*op, *x, *b);
    });
}

tng_status tng_jacobi_smooth(tng_op op, tng_vec inv_diag, int iterations, double omega, tng_vec x, tng_vec b)
{
    if (!op || !inv_diag || !x || !b || iterations < 0 || !(omega > 0.0)) return TNG_INVALID_ARGUMENT;
    return sguard(op->mesh->ctx, [&] {
        check_square(*op, *x, *b);
        TNG_REQUIRE(inv_diag->mesh == x->mesh && inv_diag->ncomp == x->ncomp && inv_diag->dtype == x->dtype,
                    "inverse diagonal must match the vectors");
        Scratch s;
        tng::Chebyshev c;
        c.jacobi = true;
        c.omega = omega;
        c.iterations = iterations;
        c.inv_diag = inv_diag;
        c.d = s.like(*x);
        c.z = s.like(*x);
        c.need_estimate = false;
        c.lmax_estimate = 1.0;
        c.solve(*op, *x, *b);
    });
}

void tng_stokes_params_default(tng_stokes_params *p)
{
    if (!p) return;
    p->krylov_restart = 10;
    p->krylov_max_iterations = 10;
    p->krylov_relative_tolerance = 1e-6;
    p->krylov_absolute_tolerance = 1e-12;
    p->viscous_pc_num_vcycles = 1;
    p->viscous_pc_chebyshev_order = 2;
    p->viscous_pc_num_smoothing_steps_prepost = 2;
    p->viscous_pc_num_power_iterations = 10;
    p->gca = 0;
    p->viscous_pc_dtype = TNG_F64;
    p->coarse_max_iterations = 50;
    p->coarse_relative_tolerance = 1e-6;
    p->coarse_absolute_tolerance = 1e-16;
    p->seed = 12345;
    p->gca_coarsest_only = 0;
    p->gca_exact = 0;
    p->viscous_pc_smoother = 0;
    p->krylov_orthogonalization = 0;
    p->rotation_correction = 0;
    p->coarse_solver = 0;
}

tng_status tng_stokes_solver_create(tng_ctx ctx, int min_level, int max_level, int radial_extra_levels, double r_min,
                                    double r_max, int lateral_sdr, int radial_sdr, tng_bc bc_cmb, tng_bc bc_surface,
                                    const tng_stokes_params *params, tng_stokes_solver *out)
{
    if (!ctx || !out) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    tng_stokes_params prm;
    tng_stokes_params_default(&prm);
    if (params) prm = *params;
    return sguard(ctx, [&] {
        TNG_REQUIRE(bc_cmb >= TNG_NEUMANN && bc_cmb <= TNG_FREESLIP && bc_surface >= TNG_NEUMANN &&
                        bc_surface <= TNG_FREESLIP,
                    "invalid boundary condition");
        std::unique_ptr<tng_stokes_solver_s> s(new tng_stokes_solver_s());
        s->s.reset(new tng::StokesSolver(ctx, min_level, max_level, radial_extra_levels, r_min, r_max, lateral_sdr,
                                         radial_sdr, bc_cmb, bc_surface, prm));
        TNG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        *out = s.release();
    });
}

tng_status tng_stokes_solver_destroy(tng_stokes_solver solver)
{
    if (!solver) return TNG_INVALID_ARGUMENT;
    delete solver;  // hipFree synchronises; the context is not touched
    return TNG_OK;
}

tng_status tng_stokes_solver_num_levels(tng_stokes_solver solver, int *n)
{
    if (!solver || !n) return TNG_INVALID_ARGUMENT;
    *n = (int)solver->s->meshes.size();
    return TNG_OK;
}

tng_status tng_stokes_solver_mesh(tng_stokes_solver solver, int level, tng_mesh *mesh)
{
    if (!solver || !mesh || level < 0 || level >= (int)solver->s->meshes.size()) return TNG_INVALID_ARGUMENT;
    *mesh = solver->s->meshes[level].get();
    return TNG_OK;
}

tng_status tng_stokes_solver_set_viscosity(tng_stokes_solver solver, tng_vec eta)
{
    if (!solver || !eta) return TNG_INVALID_ARGUMENT;
    return sguard(solver->s->vel->ctx, [&] { solver->s->set_viscosity(*eta); });
}

tng_status tng_stokes_apply(tng_stokes_solver solver, tng_vec u, tng_vec p, tng_vec out_u, tng_vec out_p)
{
    if (!solver || !u || !p || !out_u || !out_p) return TNG_INVALID_ARGUMENT;
    tng::StokesSolver &S = *solver->s;
    return sguard(S.vel->ctx, [&] {
        TNG_REQUIRE(u->mesh == S.vel && out_u->mesh == S.vel && p->mesh == S.pre_mesh && out_p->mesh == S.pre_mesh &&
                        u->ncomp == 3 && out_u->ncomp == 3 && p->ncomp == 1 && out_p->ncomp == 1,
                    "Stokes vectors must be (velocity, pressure) on the solver's meshes");
        TNG_REQUIRE(u != out_u && p != out_p, "apply: src and dst must differ");
        tng::StokesSolver::SVec src{u, p}, dst{out_u, out_p};
        S.apply_stokes(src, dst);
    });
}

tng_status tng_stokes_precondition_velocity(tng_stokes_solver solver, tng_vec x, tng_vec b)
{
    if (!solver || !x || !b) return TNG_INVALID_ARGUMENT;
    tng::StokesSolver &S = *solver->s;
    return sguard(S.vel->ctx, [&] {
        TNG_REQUIRE(x->mesh == S.vel && b->mesh == S.vel && x->ncomp == 3 && b->ncomp == 3 &&
                        x->dtype == TNG_F64 && b->dtype == TNG_F64,
                    "velocity vectors (FP64) on the solver's velocity mesh expected");
        S.precondition_velocity(*x, *b);
    });
}

tng_status tng_stokes_buoyancy_rhs(tng_stokes_solver solver, tng_vec temperature, double rayleigh, tng_vec f_u)
{
    if (!solver || !temperature || !f_u) return TNG_INVALID_ARGUMENT;
    tng::StokesSolver &S = *solver->s;
    return sguard(S.vel->ctx, [&] { S.buoyancy_rhs(*temperature, rayleigh, *f_u); });
}

tng_status tng_stokes_enforce_rhs(tng_stokes_solver solver, tng_vec f_u)
{
    if (!solver || !f_u) return TNG_INVALID_ARGUMENT;
    tng::StokesSolver &S = *solver->s;
    return sguard(S.vel->ctx, [&] {
        TNG_REQUIRE(f_u->mesh == S.vel && f_u->ncomp == 3, "velocity vector on the solver's velocity mesh expected");
        S.enforce_rhs(*f_u);
    });
}

tng_status tng_epsdivdiv_log_coefficient_gradient(tng_op op, tng_vec lam, tng_vec u, tng_vec g)
{
    if (!op || !lam || !u || !g) return TNG_INVALID_ARGUMENT;
    return sguard(op->mesh->ctx, [&] { tng::epsdivdiv_log_k_gradient(*op, *lam, *u, *g, op->mesh->ctx->stream); });
}

tng_status tng_stokes_adjoint_gradient(tng_stokes_solver solver, tng_vec u, tng_vec dJdu, tng_vec g_log_eta,
                                       tng_solve_report *report)
{
    if (!solver || !u || !dJdu || !g_log_eta) return TNG_INVALID_ARGUMENT;
    tng::StokesSolver &S = *solver->s;
    tng_solve_report local{};
    tng_solve_report &rep = report ? *report : local;
    return sguard(S.vel->ctx, [&] { S.adjoint_gradient(*u, *dJdu, *g_log_eta, rep); });
}

tng_status tng_stokes_solve(tng_stokes_solver solver, tng_vec u, tng_vec p, tng_vec f_u, tng_vec f_p,
                            tng_solve_report *report)
{
    if (!solver || !u || !p || !f_u || !f_p) return TNG_INVALID_ARGUMENT;
    tng::StokesSolver &S = *solver->s;
    tng_solve_report local{};
    tng_solve_report &rep = report ? *report : local;
    return sguard(S.vel->ctx, [&] {
        TNG_REQUIRE(u->ncomp == 3 && f_u->ncomp == 3 && p->ncomp == 1 && f_p->ncomp == 1,
                    "Stokes vectors must be (velocity, pressure)");
        tng::StokesSolver::SVec x{u, p}, b{f_u, f_p};
        S.solve(x, b, rep);
    });
}

}  // extern "C"
