// Solver objects of the Stokes path (host side; see solvers.cpp).
#pragma once

#include <memory>
#include <vector>

#include "tng_coarse.hpp"
#include "tng_internal.hpp"

namespace tng {

double power_iteration(tng_operator &A, const tng_vector &inv_diag, tng_vector &it, tng_vector &aux, int iterations,
                       uint64_t seed);

int pcg_solve(tng_operator &A, tng_vector &x, const tng_vector &b, tng_vector &r, tng_vector &p, tng_vector &ap,
              tng_vector &z, int max_iterations, double rtol, double atol, double *final_rel,
              bool x_zero = false);

// Chebyshev-accelerated Jacobi smoother on D^{-1} A, or (jacobi = true) the
// damped Jacobi smoother x += omega D^{-1} (b - A x) with
// omega = 2 / (1.3 lambda_max) (jacobi.hpp:16-60).
struct Chebyshev {
    bool jacobi = false;
    double omega = 0.0;  // Jacobi relaxation; 0: 2 / (1.3 lambda_max) from the estimate
    int order = 2, iterations = 1, power_iterations = 10;
    tng_vector *inv_diag = nullptr;
    tng_vector *d = nullptr, *z = nullptr;  // work vectors
    tng_vector *xb = nullptr;               // spare solution vector (fused sub-steps), optional
    bool need_estimate = true;
    double lmax_estimate = 0.0;  // estimate of the largest eigenvalue of D^{-1} A
    uint64_t seed = 12345;
    void solve(tng_operator &A, tng_vector &x, const tng_vector &b, bool x_zero = false);
};

// Geometric multigrid V-cycle; level 0 is the coarsest.
struct Multigrid {
    struct Level {
        tng_operator *A = nullptr;
        tng_operator *P = nullptr;  // level-1 -> level (Add mode); null on level 0
        tng_operator *R = nullptr;  // level -> level-1; null on level 0
        tng_vector *tmp = nullptr, *r = nullptr, *e = nullptr, *inv_diag = nullptr;
    };
    std::vector<Level> levels;
    std::vector<Chebyshev> pre, post;
    int num_cycles = 1;
    // FP64 PCG on the coarsest level
    tng_operator *coarse_op = nullptr;
    tng_vector *cg_x = nullptr, *cg_b = nullptr, *cg_tmp[4] = {nullptr, nullptr, nullptr, nullptr};
    int coarse_max_it = 50;
    double coarse_rtol = 1e-6, coarse_atol = 1e-16;
    long coarse_iterations = 0;
    // direct coarse solve (coarse_solver = 1), built with the viscosity
    bool use_direct = false;
    CoarseDirect direct;

    void solve(tng_vector &x, const tng_vector &b, bool x_zero = false);
    void recurse(int level, tng_vector &x, const tng_vector &b, bool x_zero);
    void coarse_solve(tng_vector &e0, const tng_vector &r0, bool x_zero);
};

struct StokesSolver {
    struct SVec {
        tng_vector *u, *p;
    };
    tng_stokes_params params;
    int bc_cmb, bc_surf;
    std::vector<std::unique_ptr<tng_domain>> meshes;  // coarsest .. velocity level
    tng_domain *vel = nullptr, *pre_mesh = nullptr;
    std::vector<std::unique_ptr<tng_vector>> pool;   // every vector the solver owns
    std::vector<tng_vector *> eta, eta_mg;           // per level: FP64 viscosity, multigrid-precision copy
    std::unique_ptr<tng_operator> A, Bt, B, coarse_op64, pmass;
    std::vector<std::unique_ptr<tng_operator>> mg_ops;
    std::vector<std::unique_ptr<tng_operator>> gca64;  // FP64 Galerkin operators of the coarse levels (gca >= 1)
    DevBuf gca_flags;  // gca = 2: one byte per coarsest-level cell, 1 = Galerkin hierarchy
    Multigrid mg;
    tng_vector *k_pm = nullptr, *inv_lumped_pmass = nullptr, *ones_p = nullptr, *ones_v = nullptr;
    tng_vector *prec_tmp_u = nullptr, *mg_x = nullptr, *mg_b = nullptr;
    std::vector<SVec> fg;  // FGMRES work space: 2m + 4 Stokes vectors
    long long num_dofs_pressure = 0;

    StokesSolver(tng_context *ctx, int min_level, int max_level, int radial_extra, double r_min, double r_max,
                 int lat_sdr, int rad_sdr, int bc_cmb, int bc_surf, const tng_stokes_params &prm);
    void set_viscosity(const tng_vector &eta_fine);
    void refresh_viscosity_hierarchy();
    void viscosity_updated();
    void apply_stokes(const SVec &src, SVec &dst);
    // level l's multigrid operator uses Galerkin matrices
    bool galerkin_level(int l) const
    {
        return params.gca >= 1 && l < (int)meshes.size() - 1 && (l == 0 || !params.gca_coarsest_only);
    }
    void precondition_velocity(tng_vector &x, const tng_vector &b, bool x_zero = false);
    void precondition(SVec &x, const SVec &b, bool x_zero = false);
    void fgmres(SVec &x, const SVec &b, tng_solve_report &rep);
    void fgmres_cgs2(SVec &x, const SVec &b, tng_solve_report &rep);
    DevBuf gs_table, gs_part, gs_out;  // CGS2: basis pointer table, multi-dot partials and results
    // rotation correction: Z (3 rotations, pressure 0), K Z, E = Z^T K Z (LU), pointer tables
    std::vector<SVec> rc_z, rc_kz;
    bool rc_dirty = true;
    double rc_lu[9] = {0};
    int rc_piv[3] = {0, 1, 2};
    DevBuf rc_ptrs;
    void rotation_setup();
    // the app's viscosity update (stokes_solver.hpp:623-638): eta = rmu^(0.5 - T)
    // on the velocity level only
    void update_viscosity_velocity_level(const tng_vector &temp, double rmu);
    tng_vector *make_vec(tng_domain *m, int ncomp);
    tng_vector *rhs_f = nullptr, *rhs_g = nullptr;  // the time step's Stokes RHS
    void rotation_correct(int j, SVec &v, SVec &z, SVec &w);
    void solve(SVec &x, const SVec &b, tng_solve_report &rep);
    void enforce_rhs(tng_vector &f_u);
    // f_u = M (Ra T x/|x|) with the consistent vector mass, then enforce_rhs (stokes_solver.hpp:650-674)
    void buoyancy_rhs(const tng_vector &temp, double rayleigh, tng_vector &f_u);
    std::unique_ptr<tng_operator> vmass;
    tng_vector *rhs_tmp = nullptr;
    // dJ/dlog eta for J(u) with K [u; p] = f: lam solves K [lam_u; lam_p] = [dJdu; 0],
    // g = -d(lam_u^T A(eta) u)/dlog eta (B, B^T, the penalty and the preconditioner do not depend on eta)
    void adjoint_gradient(const tng_vector &u, const tng_vector &dJdu, tng_vector &g, tng_solve_report &rep);
    tng_vector *lam_u = nullptr, *lam_p = nullptr, *zero_p = nullptr;
    static void record(tng_solve_report &rep, double abs_res);
};

}  // namespace tng

struct tng_energy_s;
struct tng_stokes_solver_s {
    std::unique_ptr<tng::StokesSolver> s;
};
