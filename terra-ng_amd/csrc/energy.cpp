// The caller around the Stokes solve in the mantle-circulation app: the
// implicit SUPG energy step and one step of the time loop.
//
//  - EnergySUPG mirrors SUPGSolver (apps/mantlecirculation/src/energy_solver.hpp
//    :85-276): A = M + dt (K_diff + K_adv + K_supg) with Dirichlet rows, the
//    operator without boundary treatment (Neumann) and its diagonal, the
//    consistent mass M; compute_dt = dt_scaling * h / max |u| (h the smallest
//    radial spacing); step: inverse diagonal of the Neumann operator at the
//    step's dt, then per substep q = M T, g = Dirichlet temperatures, the
//    strong elimination (strong_algebraic_dirichlet_enforcement.hpp:24-47)
//    and FGMRES with the diagonal (Jacobi) preconditioner on A T = q.
//  - tng_timestep mirrors the loop body of apps/mantlecirculation/
//    mantlecirculation.cpp:413-445: T snapshot, dt from the current velocity,
//    then per Picard iteration (T restored after the first) the Stokes solve
//    with the buoyancy of T, the energy step, and the viscosity update from T
//    (stokes_solver.hpp:623-638: Frank-Kamenetskii eta = rmu^(0.5 - T) on the
//    velocity level only).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "tng_internal.hpp"
#include "tng_solvers.hpp"

struct tng_energy_s {
    tng_energy_params prm{};
    tng::StokesSolver *stokes = nullptr;
    tng_vector *vel = nullptr, *temp = nullptr;
    std::unique_ptr<tng_operator> A, A_neumann, A_neumann_diag;
    std::vector<std::unique_ptr<tng_vector>> pool;
    tng_vector *g = nullptr, *tmp = nullptr, *q = nullptr, *diag = nullptr, *ones = nullptr, *t_backup = nullptr;
    std::vector<tng_vector *> fg;  // FGMRES work space: 2m + 4 scalars
    double h = 0.0;
};

namespace tng {

namespace {

hipStream_t stream_of(const tng_vector &v) { return v.mesh->ctx->stream; }

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
double dot(const tng_vector &x, const tng_vector &y) { return reduce_to_host(REDUCE_DOT, x, &y, stream_of(x)); }

tng_operator *new_supg(tng_domain *mesh, tng_vector *vel, double kappa, double dt, bool treat_boundary,
                       bool diagonal, double mass_scaling, bool lumped)
{
    TNG_REQUIRE(vel && vel->mesh == mesh && vel->ncomp == 3 && vel->dtype == TNG_F64,
                "SUPG: the velocity must be an FP64 vector on the operator's mesh");
    TNG_REQUIRE(mesh->nr >= 2, "SUPG: at least one radial layer");
    std::unique_ptr<tng_operator> op(new tng_operator());
    op->kind = OP_SUPG;
    op->mesh = mesh;
    op->vel = vel;
    op->supg_kappa = kappa;
    op->supg_dt = dt;
    op->supg_treat_boundary = treat_boundary;
    op->diagonal = diagonal;
    op->supg_mass_scaling = mass_scaling;
    op->supg_lumped = lumped;
    return op.release();
}

tng_vector *evec(tng_energy_s &e, tng_domain *m)
{
    e.pool.emplace_back(new_vector(m, TNG_F64, 1));
    return e.pool.back().get();
}

// Scalar FGMRES(m) with the diagonal preconditioner z = inv_diag * v and
// modified Gram-Schmidt (linalg/solvers/fgmres.hpp:109-373).
void energy_fgmres(tng_energy_s &e, tng_operator &A, tng_vector &x, const tng_vector &b, tng_solve_report &rep)
{
    const int m = e.prm.krylov_restart;
    const double rtol = e.prm.krylov_relative_tolerance, atol = e.prm.krylov_absolute_tolerance;
    tng_vector &r = *e.fg[0];
    op_apply(A, x, r);
    lincomb(r, {1.0, -1.0}, {&b, &r});
    double beta0 = std::sqrt(dot(r, r));
    const double initial = beta0;
    rep.initial_residual = initial;
    rep.final_residual = beta0;
    rep.relative_residual = 1.0;
    StokesSolver::record(rep, beta0);
    if (beta0 <= atol) {
        rep.converged = 1;
        return;
    }
    const int offV = 1, offZ = offV + m + 1, idxW = offZ + m, idxAcc = idxW + 1;
    std::vector<double> H((size_t)(m + 1) * m), g(m + 1), cs(m + 1), sn(m + 1), y;
    auto h = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
    int total = 0;
    while (total < e.prm.krylov_max_iterations) {
        lincomb(*e.fg[offV], {1.0 / beta0}, {&r});
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta0;
        int inner = 0;
        for (int j = 0; j < m && total < e.prm.krylov_max_iterations; ++j, ++total) {
            tng_vector &vj = *e.fg[offV + j], &zj = *e.fg[offZ + j], &w = *e.fg[idxW];
            rep.iterations += 1;
            launch_copy(zj, vj, stream_of(zj));
            launch_mult(zj, *e.diag, stream_of(zj));
            op_apply(A, zj, w);
            for (int i = 0; i <= j; ++i) {
                h(i, j) = dot(w, *e.fg[offV + i]);
                lincomb(w, {1.0, -h(i, j)}, {&w, e.fg[offV + i]});
            }
            const double hj1 = std::sqrt(dot(w, w));
            if (hj1 < std::numeric_limits<double>::epsilon() * initial) {
                h(j + 1, j) = 0;
                inner = j + 1;
                rep.breakdowns += 1;
                break;
            }
            h(j + 1, j) = hj1;
            lincomb(*e.fg[offV + j + 1], {1.0 / hj1}, {&w});
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * h(i, j) + sn[i] * h(i + 1, j);
                h(i + 1, j) = -sn[i] * h(i, j) + cs[i] * h(i + 1, j);
                h(i, j) = t;
            }
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
            const double abs_res = std::abs(g[j + 1]);
            StokesSolver::record(rep, abs_res);
            if (!std::isfinite(abs_res)) {
                inner = j;
                ++total;
                rep.nonfinite_residual += 1;
                break;
            }
            inner = j + 1;
            if (abs_res / initial <= rtol || abs_res < atol) break;
        }
        y.assign(inner, 0.0);
        for (int row = inner - 1; row >= 0; --row) {
            double s = g[row];
            for (int col = row + 1; col < inner; ++col) s -= h(row, col) * y[col];
            y[row] = s / h(row, row);
        }
        tng_vector &acc = *e.fg[idxAcc];
        launch_fill(acc, 0.0, stream_of(acc));
        for (int i = 0; i < inner; ++i) lincomb(acc, {1.0, y[i]}, {&acc, e.fg[offZ + i]});
        lincomb(x, {1.0, 1.0}, {&x, &acc});
        rep.restarts += 1;
        op_apply(A, x, r);
        lincomb(r, {1.0, -1.0}, {&b, &r});
        beta0 = std::sqrt(dot(r, r));
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

void reset_report(tng_solve_report &rep)
{
    double *hist = rep.history;
    const int cap = rep.history_capacity;
    std::memset(&rep, 0, sizeof(rep));
    rep.history = hist;
    rep.history_capacity = cap;
}

}  // namespace

double energy_compute_dt(tng_energy_s &e)
{
    const double max_vel = max_vector_magnitude(*e.vel, stream_of(*e.vel));
    const double dt_adv = max_vel > 1e-12 ? e.h / max_vel : 1e-3;
    return e.prm.dt_scaling * dt_adv;
}

void energy_step(tng_energy_s &e, double dt, tng_solve_report &rep)
{
    hipStream_t s = stream_of(*e.temp);
    e.A->supg_dt = dt;
    e.A_neumann->supg_dt = dt;
    e.A_neumann_diag->supg_dt = dt;
    // inverse diagonal at the step's dt
    op_apply(*e.A_neumann_diag, *e.ones, *e.diag);
    launch_invert(*e.diag, s);
    reset_report(rep);
    for (int i = 0; i < e.prm.energy_substeps; ++i) {
        launch_scalar_mass(*e.temp, *e.q, s);                                   // q = M T^n
        launch_boundary_values(*e.g, e.prm.temperature_cmb, e.prm.temperature_surface, s);
        op_apply(*e.A_neumann, *e.g, *e.tmp);                                   // g_A = A_N g
        lincomb(*e.q, {1.0, -1.0}, {e.q, e.tmp});                               // q -= g_A
        op_apply(*e.A_neumann_diag, *e.g, *e.tmp);                              // g_D = diag(A_N) g
        launch_assign_masked(*e.q, *e.tmp, SBF_BOUNDARY, s);                    // q = g_D on the boundary
        tng_solve_report sub{};
        sub.history = rep.history ? rep.history + std::min(rep.history_len, rep.history_capacity) : nullptr;
        sub.history_capacity = rep.history ? std::max(0, rep.history_capacity - rep.history_len) : 0;
        energy_fgmres(e, *e.A, *e.temp, *e.q, sub);
        rep.iterations += sub.iterations;
        rep.restarts += sub.restarts;
        rep.breakdowns += sub.breakdowns;
        rep.nonfinite_residual += sub.nonfinite_residual;
        rep.converged = sub.converged;
        rep.initial_residual = sub.initial_residual;
        rep.final_residual = sub.final_residual;
        rep.relative_residual = sub.relative_residual;
        rep.history_len += sub.history_len;
    }
}

}  // namespace tng

using namespace tng;

namespace {

template <class F> tng_status eguard(tng_context *ctx, F &&f)
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

}  // namespace

extern "C" {

tng_status tng_supg_create(tng_mesh mesh, tng_vec velocity, double diffusivity, double dt, int treat_boundary,
                           int diagonal, double mass_scaling, int lumped_mass, tng_op *out)
{
    if (!mesh || !velocity || !out || !mesh->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return eguard(mesh->ctx, [&] {
        *out = new_supg(mesh, velocity, diffusivity, dt, treat_boundary != 0, diagonal != 0, mass_scaling,
                        lumped_mass != 0);
    });
}

tng_status tng_supg_set_dt(tng_op op, double dt)
{
    if (!op || op->kind != OP_SUPG) return TNG_INVALID_ARGUMENT;
    op->supg_dt = dt;
    return TNG_OK;
}

tng_status tng_supg_set_enabled(tng_op op, int on)
{
    if (!op || op->kind != OP_SUPG) return TNG_INVALID_ARGUMENT;
    op->supg_enabled = on != 0;
    return TNG_OK;
}

tng_status tng_mass_apply(tng_vec src, tng_vec dst)
{
    if (!src || !dst || src == dst) return TNG_INVALID_ARGUMENT;
    return eguard(src->mesh->ctx, [&] { launch_scalar_mass(*src, *dst, src->mesh->ctx->stream); });
}

void tng_energy_params_default(tng_energy_params *p)
{
    if (!p) return;
    p->krylov_restart = 5;
    p->krylov_max_iterations = 100;
    p->krylov_relative_tolerance = 1e-6;
    p->krylov_absolute_tolerance = 1e-12;
    p->diffusivity = 1.0;
    p->temperature_cmb = 1.0;
    p->temperature_surface = 0.0;
    p->dt_scaling = 0.5;
    p->energy_substeps = 1;
}

tng_status tng_energy_create(tng_stokes_solver solver, tng_vec velocity, tng_vec temperature,
                             const tng_energy_params *params, tng_energy *out)
{
    if (!solver || !solver->s || !velocity || !temperature || !out) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    StokesSolver &S = *solver->s;
    return eguard(S.vel->ctx, [&] {
        TNG_REQUIRE(velocity->mesh == S.vel && temperature->mesh == S.vel && velocity->ncomp == 3 &&
                        temperature->ncomp == 1 && velocity->dtype == TNG_F64 && temperature->dtype == TNG_F64,
                    "energy: FP64 velocity and temperature on the solver's velocity mesh");
        std::unique_ptr<tng_energy_s> e(new tng_energy_s());
        tng_energy_params_default(&e->prm);
        if (params) e->prm = *params;
        TNG_REQUIRE(e->prm.krylov_restart >= 1 && e->prm.krylov_max_iterations >= 0 && e->prm.energy_substeps >= 1,
                    "energy: invalid Krylov or substep parameters");
        e->stokes = &S;
        e->vel = velocity;
        e->temp = temperature;
        tng_domain *m = S.vel;
        // smallest radial spacing over the subdomains (min_radial_h)
        double h = std::numeric_limits<double>::infinity();
        for (int sd = 0; sd < m->nsd; ++sd)
            for (int r = 0; r + 1 < m->nr; ++r) h = std::min(h, m->radii[(size_t)sd * m->nr + r + 1] - m->radii[(size_t)sd * m->nr + r]);
        e->h = h;
        const double kappa = e->prm.diffusivity;
        e->A.reset(new_supg(m, velocity, kappa, 0.0, true, false, 1.0, false));
        e->A_neumann.reset(new_supg(m, velocity, kappa, 0.0, false, false, 1.0, false));
        e->A_neumann_diag.reset(new_supg(m, velocity, kappa, 0.0, false, true, 1.0, false));
        e->g = evec(*e, m);
        e->tmp = evec(*e, m);
        e->q = evec(*e, m);
        e->diag = evec(*e, m);
        e->ones = evec(*e, m);
        e->t_backup = evec(*e, m);
        launch_fill(*e->ones, 1.0, m->ctx->stream);
        const int nt = 2 * e->prm.krylov_restart + 4;
        for (int i = 0; i < nt; ++i) e->fg.push_back(evec(*e, m));
        *out = e.release();
    });
}

tng_status tng_energy_destroy(tng_energy e)
{
    if (!e) return TNG_INVALID_ARGUMENT;
    delete e;
    return TNG_OK;
}

tng_status tng_energy_compute_dt(tng_energy e, double *dt)
{
    if (!e || !dt) return TNG_INVALID_ARGUMENT;
    return eguard(e->vel->mesh->ctx, [&] { *dt = energy_compute_dt(*e); });
}

tng_status tng_energy_step(tng_energy e, double dt, tng_solve_report *report)
{
    if (!e) return TNG_INVALID_ARGUMENT;
    return eguard(e->vel->mesh->ctx, [&] {
        tng_solve_report local{};
        energy_step(*e, dt, report ? *report : local);
    });
}

tng_status tng_stokes_solver_update_viscosity(tng_stokes_solver solver, tng_vec temperature, double rmu)
{
    if (!solver || !solver->s || !temperature) return TNG_INVALID_ARGUMENT;
    StokesSolver &S = *solver->s;
    return eguard(S.vel->ctx, [&] {
        TNG_REQUIRE(temperature->mesh == S.vel && temperature->ncomp == 1 && temperature->dtype == TNG_F64,
                    "viscosity update: FP64 temperature on the velocity mesh");
        S.update_viscosity_velocity_level(*temperature, rmu);
    });
}

tng_status tng_timestep(tng_stokes_solver solver, tng_energy energy, tng_vec u, tng_vec p,
                        const tng_timestep_params *params, double *dt_out, tng_solve_report *stokes_report,
                        tng_solve_report *energy_report)
{
    if (!solver || !solver->s || !energy || !u || !p || !params) return TNG_INVALID_ARGUMENT;
    StokesSolver &S = *solver->s;
    return eguard(S.vel->ctx, [&] {
        TNG_REQUIRE(params->picard_iterations >= 1, "time step: at least one Picard iteration");
        TNG_REQUIRE(params->viscosity_law == 0 || params->viscosity_law == 1, "time step: viscosity law 0 or 1");
        TNG_REQUIRE(u == energy->vel, "time step: u must be the velocity the energy solver advects with");
        hipStream_t s = S.vel->ctx->stream;
        tng_vector &T = *energy->temp;
        launch_copy(*energy->t_backup, T, s);          // snapshot_for_picard
        const double dt = energy_compute_dt(*energy);  // dt from the current velocity
        if (dt_out) *dt_out = dt;
        tng_solve_report srep{}, erep{};
        tng_solve_report &sr = stokes_report ? *stokes_report : srep;
        tng_solve_report &er = energy_report ? *energy_report : erep;
        for (int it = 0; it < params->picard_iterations; ++it) {
            if (it > 0) launch_copy(T, *energy->t_backup, s);  // restore_for_picard
            // Stokes solve with the buoyancy of T (stokes_solver.hpp:644-691), from the current (u, p)
            if (!S.rhs_f) S.rhs_f = S.make_vec(S.vel, 3);
            if (!S.rhs_g) S.rhs_g = S.make_vec(S.pre_mesh, 1);
            S.buoyancy_rhs(T, params->rayleigh_number, *S.rhs_f);
            launch_fill(*S.rhs_g, 0.0, s);
            StokesSolver::SVec x{u, p}, b{S.rhs_f, S.rhs_g};
            S.solve(x, b, sr);
            energy_step(*energy, dt, er);
            if (params->viscosity_law == 1) {
                S.update_viscosity_velocity_level(T, params->rmu);
                if (params->refresh_hierarchy) S.refresh_viscosity_hierarchy();
            }
        }
    });
}

}  // extern "C"
