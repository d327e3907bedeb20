/*
 * tng.h -- C ABI of the MI355X-native TERRA-NG Stokes hot path.
 *
 * Opaque handles over library-owned device buffers.  Every entry point returns
 * a tng_status; nothing throws across the ABI.  A context is single-threaded
 * (the caller serialises calls); all device work is ordered on the context's
 * HIP stream.  Calls that return host scalars synchronise the stream; every
 * other call is asynchronous with respect to the host.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the TERRA-NG source tree).
 *
 * Data layout.  A node field has the reference's logical indexing
 * (subdomain, x, y, r) with r fastest (grid/grid_types.hpp:44-113); a vector
 * field stores its components separately (SoA).  tng_vec_upload/download take
 * host arrays in exactly that order: component-major, then (sd, x, y, r).
 * Interface nodes are duplicated per subdomain as in the reference.
 */
#ifndef TNG_H
#define TNG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TNG_OK = 0,
    TNG_INVALID_ARGUMENT = 1,
    TNG_OUT_OF_MEMORY = 2,
    TNG_DEVICE_ERROR = 3,
    TNG_NOT_CONVERGED = 4,
    TNG_NONFINITE = 5,
    TNG_UNSUPPORTED = 6,
    TNG_IO_ERROR = 7
} tng_status;

typedef enum { TNG_F64 = 0, TNG_F32 = 1 } tng_dtype;

/* grid/shell/bit_masks.hpp:31-36 BoundaryConditionFlag (same numeric values) */
typedef enum { TNG_NEUMANN = 0, TNG_DIRICHLET = 1, TNG_FREESLIP = 2 } tng_bc;

/* linalg/operator.hpp:29-43 OperatorApplyMode / OperatorCommunicationMode */
typedef enum { TNG_REPLACE = 0, TNG_ADD = 1 } tng_apply_mode;
typedef enum { TNG_SKIP_COMMUNICATION = 0, TNG_COMMUNICATE_ADDITIVELY = 1 } tng_comm_mode;

typedef struct tng_context *tng_ctx;
typedef struct tng_domain *tng_mesh;
typedef struct tng_vector *tng_vec;
typedef struct tng_operator *tng_op;
typedef struct tng_xdmf_writer *tng_xdmf;

/* ---------------------------------------------------------------- context */

/* util/init.hpp:75-80 terra_initialize (device selection + stream) */
tng_status tng_ctx_create(int device_ordinal, tng_ctx *out);
/* Same, on the caller's HIP stream (a hipStream_t passed as void*), e.g. the
 * stream of a Kokkos::HIP execution space instance, so library work orders
 * with the caller's kernels without extra fences.  The stream is not
 * destroyed with the context. */
tng_status tng_ctx_create_with_stream(int device_ordinal, void *hip_stream, tng_ctx *out);
tng_status tng_ctx_destroy(tng_ctx ctx);
/* No counterpart in the reference (Kokkos picks its own launch shapes).
 * Launch-shape settings of the EpsilonDivDiv apply on this context; every
 * setting computes the same operator, the defaults are the measured best.
 * name / default / meaning:
 *   "apply_blocks"        0  FP64 applies: workgroups sharing the node rows (0: one resident round)
 *   "apply_seg"           0  FP32 and fused applies: node rows per workgroup (0: chosen per mesh)
 *   "apply_radial_split"  1  0: 128 and 256 layers as radial chunks of one workgroup
 *   "apply_radial_chunks" 0  1: 128 and 256 layers as two or four cells per lane
 *   "apply_waves"         8  4: four-wave workgroups at 64 layers
 *   "cheb_fuse"           1  0: Chebyshev sub-steps as separate kernels (read when a
 *                            Stokes solver is created)
 * A negative value restores the default; an unknown name is refused. */
tng_status tng_ctx_set_tuning(tng_ctx ctx, const char *name, int value);
/* Kokkos::fence() */
tng_status tng_ctx_synchronize(tng_ctx ctx);
/* Message of the last failed call on this context ("" if none). */
const char *tng_last_error(tng_ctx ctx);
/* Library version string. */
const char *tng_version(void);

/* ------------------------------------------------------------------- mesh */

typedef struct {
    int num_subdomains;       /* local subdomains (all of them on one GPU)   */
    int nodes_lateral;        /* nodes per subdomain side (x and y)          */
    int nodes_radial;         /* nodes per subdomain radially                */
    int lateral_level;        /* diamond lateral refinement level            */
    int radial_level;         /* log2(number of radial layers)               */
    int lateral_sdr;          /* log2(subdomains per diamond side)           */
    int radial_sdr;           /* log2(radial subdomains)                     */
    long long num_alloc;      /* allocated node copies = nsd*nl*nl*nr        */
    long long num_owned;      /* unique nodes (OWNED copies)                 */
    long long num_interface;  /* node copies that have duplicates            */
} tng_mesh_info;

/* grid/shell/spherical_shell.hpp:2599-2640 DistributedDomain::create_uniform
 * with uniform_shell_radii (:27-52), subdomain_unit_sphere_single_shell_coords
 * (:2827-2853), subdomain_shell_radii (:2862-2887), setup_boundary_mask_data
 * (grid/shell/bit_masks.hpp:82-136), setup_node_ownership_mask_data
 * (grid/bit_masks.hpp:29-85) and the additive halo plan
 * (communication/shell/communication_plan.hpp:31-71), one rank holding every
 * subdomain.  With ctx == NULL the mesh is built on the host only (no device
 * buffers): the copy/info queries work, vectors and operators cannot be
 * created on it. */
tng_status tng_mesh_create_uniform(tng_ctx ctx, int lateral_level, int radial_level, double r_min, double r_max,
                                   int lateral_sdr, int radial_sdr, tng_mesh *out);
tng_status tng_mesh_destroy(tng_mesh mesh);
tng_status tng_mesh_get_info(tng_mesh mesh, tng_mesh_info *info);
/* Host copies of the setup data, reference layouts: coords (sd, x, y, 3),
 * radii (sd, r), masks (sd, x, y, r). */
tng_status tng_mesh_copy_coords(tng_mesh mesh, double *host_out);
tng_status tng_mesh_copy_radii(tng_mesh mesh, double *host_out);
tng_status tng_mesh_copy_boundary_mask(tng_mesh mesh, uint8_t *host_out);
tng_status tng_mesh_copy_ownership_mask(tng_mesh mesh, uint8_t *host_out);

/* ---------------------------------------------------------------- vectors */

/* linalg/vector_q1.hpp:19-155 VectorQ1Scalar (ncomp = 1) and :166-304
 * VectorQ1Vec (ncomp = 3); storage grid/grid_types.hpp:44-113. */
tng_status tng_vec_create(tng_mesh mesh, tng_dtype dtype, int ncomp, tng_vec *out);
/* Non-owning view of caller device memory: ncomp component arrays, each
 * num_alloc contiguous elements in (sd, x, y, r) order -- the layout of a
 * Kokkos-HIP Grid4DDataVec component (grid/grid_types.hpp:44-113), so a
 * TERRA-NG vector can be passed without copies.  The memory must outlive the
 * view; destroying the view does not free it. */
tng_status tng_vec_wrap(tng_mesh mesh, tng_dtype dtype, int ncomp, void *const *component_ptrs, tng_vec *out);
tng_status tng_vec_destroy(tng_vec v);
/* Host <-> device in the reference order (comp, sd, x, y, r); dtype of v. */
tng_status tng_vec_upload(tng_vec v, const void *host);
tng_status tng_vec_download(tng_vec v, void *host);
/* Device address of component comp (non-owning view for interop). */
tng_status tng_vec_device_ptr(tng_vec v, int comp, void **ptr);
/* linalg::assign(v, c) (linalg/vector.hpp:71-202, grid_operations.hpp:11-34) */
tng_status tng_vec_fill(tng_vec v, double value);
/* linalg::assign(dst, src) incl. FP64 <-> FP32 conversion */
tng_status tng_vec_copy(tng_vec dst, tng_vec src);

/* -------------------------------------------------------------- operators */

/* fe/wedge/operators/shell/epsilon_divdiv_kerngen.hpp:203-323
 * EpsilonDivDivKerngen(domain, coords, radii, boundary mask, k, bcs, diagonal,
 * Replace, CommunicateAdditively, stored matrices Off).  k is a scalar node
 * field captured by reference: after changing its values call
 * tng_op_coefficient_updated.  The fs/fs rotation penalty (:324-445) is
 * armed at creation when both boundaries are free-slip, as in the reference. */
tng_status tng_epsdivdiv_create(tng_mesh mesh, tng_vec k, tng_bc bc_cmb, tng_bc bc_surface, int diagonal, tng_op *out);

/* epsilon_divdiv_kerngen.hpp:447-453 set_operator_apply_and_communication_modes */
tng_status tng_op_set_modes(tng_op op, tng_apply_mode apply_mode, tng_comm_mode comm_mode);
/* epsilon_divdiv_kerngen.hpp:455 set_diagonal */
tng_status tng_op_set_diagonal(tng_op op, int diagonal);
/* epsilon_divdiv_kerngen.hpp:474-481 set_boundary_conditions */
tng_status tng_op_set_boundary_conditions(tng_op op, tng_bc bc_cmb, tng_bc bc_surface);
/* epsilon_divdiv_kerngen.hpp:315 set_penalty_epsilon (default 1e-7, :168) */
tng_status tng_op_set_penalty_epsilon(tng_op op, double eps);
/* Re-derive the coefficient-dependent boundary rows after k changed in
 * place (the reference reads k at every apply; this library caches the
 * boundary-row diagonals derived from it). */
tng_status tng_op_coefficient_updated(tng_op op);
/* linalg::apply (linalg/operator.hpp:155-159) -> apply_impl
 * (epsilon_divdiv_kerngen.hpp:577-792). */
tng_status tng_op_apply(tng_op op, tng_vec src, tng_vec dst);
/* No counterpart in the reference.  tng_op_apply of an EpsilonDivDiv operator
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * kernels and results, with the fs/fs penalty's per-block moments and the
 * partial sums between radial parts in a fixed scratch array of 8 MiB per
 * GPU instead of library allocations.  A mesh whose radial-part sums exceed
 * the scratch (8 MiB less 24 KiB: e.g. more than three parts of 64 layers at
 * lateral level 7) is rejected with TNG_INVALID_ARGUMENT.  The operator's
 * boundary-row table is allocated when the operator is created, so the first
 * apply may be this one.  No allocation, no host synchronisation; profiling
 * events are not recorded.  Every call on a GPU shares that GPU's scratch:
 * calls that overlap on one GPU (two streams, two host threads) race. */
tng_status tng_epsdivdiv_apply_noalloc(tng_op op, tng_vec src, tng_vec dst);
tng_status tng_op_destroy(tng_op op);

/* Stored local matrices (epsilon_divdiv_kerngen.hpp:503-575
 * set_stored_matrix_mode(Full) + set_local_matrix / get_local_matrix, storage
 * linalg/solvers/gca/local_matrix_storage.hpp:28-243).  Layout of the host
 * arrays: (sd, x_cell, y_cell, r_cell, wedge, 18, 18) doubles, dof index
 * node + 6 * dimension.  With stored matrices the operator applies the slow
 * path (epsilon_divdiv_kerngen.hpp:1013-1242) on them; the mesh needs at
 * least two radial layers.  get returns the stored matrices, or the Neumann
 * matrices assembled from k (assemble_local_matrix, :2315-2392) when none
 * are stored. */
tng_status tng_op_get_local_matrices(tng_op op, double *host_out);
tng_status tng_op_set_local_matrices(tng_op op, const double *host_in);
tng_status tng_op_clear_local_matrices(tng_op op);
/* set_stored_matrix_mode (epsilon_divdiv_kerngen.hpp:503-530) with
 * OperatorStoredMatrixMode::Full (full = 1: zero matrices of the operator's
 * precision, as LocalMatrixStorage allocates them) or Off (full = 0). */
tng_status tng_op_set_stored_matrix_mode(tng_op op, int full);
/* GCACapable per-element access (linalg/operator.hpp:80-112;
 * epsilon_divdiv_kerngen.hpp:533-575): get_local_matrix(sd, x_cell, y_cell,
 * r_cell, wedge) returns the stored matrix, or the Neumann matrix assembled
 * from k without storage; set_local_matrix needs stored mode (the reference
 * asserts it).  mat: 324 doubles, row-major 18 x 18, dof index node + 6 *
 * dimension (kernel_helpers.hpp:617-651). */
tng_status tng_op_get_local_matrix(tng_op op, int sd, int x_cell, int y_cell, int r_cell, int wedge, double *mat);
tng_status tng_op_set_local_matrix(tng_op op, int sd, int x_cell, int y_cell, int r_cell, int wedge,
                                   const double *mat);
/* Copy (and convert to dst's precision) the stored FP64 matrices of src. */
tng_status tng_op_copy_local_matrices(tng_op dst, tng_op src);
/* linalg/solvers/gca/gca.hpp:41-531 TwoGridGCA(fine_op, coarse_op) with
 * constant interpolation on all elements (gca = 1): the coarse operator (FP64)
 * receives, per coarse wedge, the sum of P^T A_fine P over its fine wedges.
 * A_fine is the fine operator's stored matrices, or its Neumann local
 * matrices assembled from its (FP64) k when it stores none. */
tng_status tng_gca_assemble(tng_op fine, tng_op coarse);
/* As tng_gca_assemble with the Galerkin product R A P: fine nodes on the top
 * face of an interior coarse cell draw their weights from the coarse top
 * layer into the coarse wedge's top slots.  The reference's rule
 * (gca.hpp:206-208, r_idx_coarse_bot = r_fine / 2) places those weights in the
 * bottom slots; this variant has no counterpart in the reference. */
tng_status tng_gca_assemble_exact(tng_op fine, tng_op coarse);
/* Adaptive GCA (gca = 2).  linalg/solvers/gca/gca_elements_collector.hpp:20-96
 * GCAElementsCollector(fine_domain, k, level_range, GCAElements): per fine
 * wedge, |grad k| in reference coordinates at the centroid; above 10 the
 * coarsest-level ancestor cell (cell index >> level_range) is flagged.
 * host_flags receives one byte per coarsest cell, layout (sd, x_cell, y_cell,
 * r_cell) of `coarsest`, 1 = flagged; k is an FP64 scalar on a mesh
 * level_range refinements above `coarsest`. */
tng_status tng_gca_elements_collect(tng_vec k, int level_range, tng_mesh coarsest, unsigned char *host_flags);
/* TwoGridGCA with element selection (gca.hpp:150-175) into an operator in
 * Selective stored-matrix mode (epsilon_divdiv_kerngen.hpp:503-575,
 * 1040-1075): wedges of coarse cells whose coarsest ancestor
 * (cell >> level_range) is flagged in host_flags receive the Galerkin matrix,
 * all others the Neumann matrix assembled from the coarse operator's own k.
 * The coarse mesh is level_range refinements above `coarsest`. */
tng_status tng_gca_assemble_selective(tng_op fine, tng_op coarse, int level_range, tng_mesh coarsest,
                                      const unsigned char *host_flags);

/* fe/wedge/operators/shell/divergence_kerngen.hpp:121-257 DivergenceKerngen
 * (fine velocity mesh -> coarse pressure mesh, one refinement apart): B u,
 * Dirichlet test rows zeroed, free-slip trial projection, Replace +
 * CommunicateAdditively by default (tng_op_set_modes applies). */
tng_status tng_divergence_create(tng_mesh fine, tng_mesh coarse, tng_bc bc_cmb, tng_bc bc_surface, tng_op *out);
/* fe/wedge/operators/shell/gradient_kerngen.hpp:126-258 GradientKerngen
 * (coarse pressure -> fine velocity): B^T p with the same boundary handling. */
tng_status tng_gradient_create(tng_mesh fine, tng_mesh coarse, tng_bc bc_cmb, tng_bc bc_surface, tng_op *out);
/* fe/wedge/operators/shell/prolongation_constant.hpp:14-200
 * ProlongationConstant / ProlongationVecConstant (coarse -> fine, 1 or 3
 * components chosen by the vectors); no communication. */
tng_status tng_prolongation_create(tng_mesh coarse, tng_mesh fine, tng_op *out);
/* fe/wedge/operators/shell/restriction_constant.hpp:18-230
 * RestrictionConstant / RestrictionVecConstant (fine -> coarse): transpose
 * of the prolongation over OWNED fine nodes, additive exchange. */
tng_status tng_restriction_create(tng_mesh fine, tng_mesh coarse, tng_op *out);
/* fe/wedge/operators/shell/kmass.hpp:44-190 KMass with lumped_diagonal =
 * true (the Schur complement approximation of the Stokes preconditioner):
 * dst = diag(row sums of the k-weighted mass matrix) src, additive exchange. */
tng_status tng_kmass_lumped_create(tng_mesh mesh, tng_vec k, tng_op *out);

/* fe/wedge/operators/shell/vector_mass.hpp:17-130 VectorMass (consistent,
 * 6-point rule, per component; Replace + CommunicateAdditively). */
tng_status tng_vector_mass_create(tng_mesh mesh, tng_op *out);

/* ------------------------------------------------- vector linear algebra */

/* linalg/vector.hpp:72-95 lincomb: y = sum_i c[i] x[i] + c0, 0 <= n <= 3;
 * x[i] may alias y. */
tng_status tng_vec_lincomb(tng_vec y, int n, const double *c, const tng_vec *x, double c0);
/* linalg/vector.hpp:137-144 scale_in_place: y *= x entrywise */
tng_status tng_vec_scale_with_vector(tng_vec y, tng_vec x);
/* linalg/vector.hpp:127-135 invert_entries: y = 1 / y entrywise */
tng_status tng_vec_invert(tng_vec y);
/* linalg/vector_q1.hpp:96-102, 245-251 dot over OWNED nodes (synchronises) */
tng_status tng_vec_dot(tng_vec x, tng_vec y, double *out);
/* linalg/vector_q1.hpp:121, 270 max_abs_entry over all copies (synchronises) */
tng_status tng_vec_max_abs(tng_vec x, double *out);
/* linalg/vector.hpp:189-197 has_nan_or_inf (synchronises) */
tng_status tng_vec_has_nan_or_inf(tng_vec x, int *out);
/* norm_2 (linalg/vector.hpp:7-42 via dot_impl; vector_q1.hpp:245-250):
 * sqrt of the OWNED-masked dot of x with itself. */
tng_status tng_vec_norm2(tng_vec x, double *out);
/* dot_impl with the result left on the device: *out_device (a device address,
 * 8 bytes) receives the OWNED-masked dot, stream-ordered, without a host
 * synchronisation. */
tng_status tng_vec_dot_device(tng_vec x, tng_vec y, double *out_device);
/* randomize_impl (linalg/vector.hpp:7-42; kernels/common/grid_operations.hpp
 * :748-773 rand): uniform [0, 1) entries on every copy.  The reference draws
 * from a Kokkos XorShift64 pool with seed 12345 (order not reproducible across
 * backends); here every entry is SplitMix64 of (seed, component, linear index),
 * reproducible and independent of the launch. */
tng_status tng_vec_fill_seeded(tng_vec x, unsigned long long seed);

/* fe/wedge/operators/shell/epsilon_divdiv_stokes.hpp:18-133 EpsDivDivStokes
 * (domain_fine, domain_coarse, ..., k, bcs, diagonal): K = [A B^T; B 0] with
 * A = EpsilonDivDivKerngen on the fine (velocity) mesh with coefficient k,
 * B^T = GradientKerngen, B = DivergenceKerngen (coarse = pressure mesh), and
 * the Zero block (zero.hpp).  The Stokes operator owns its blocks. */
tng_status tng_stokes_create(tng_mesh fine, tng_mesh coarse, tng_vec k, tng_bc bc_cmb, tng_bc bc_surface,
                             int diagonal, tng_op *out);
/* EpsDivDivStokes::apply_impl (epsilon_divdiv_stokes.hpp:55-84):
 * out_u = A u + B^T p, out_p = B u; A (Replace, SkipCommunication), B^T (Add,
 * CommunicateAdditively), B (Replace, CommunicateAdditively), then every
 * block is reset to Replace + CommunicateAdditively.  Diagonal operator:
 * out_u = diag(A) u, out_p untouched. */
tng_status tng_stokes_op_apply(tng_op op, tng_vec u, tng_vec p, tng_vec out_u, tng_vec out_p);
/* block_11() / block_12() / block_21() / block_22() (epsilon_divdiv_stokes.hpp
 * :86-133; Block2x2OperatorLike, linalg/operator.hpp:117-141): row, col in
 * {1, 2}.  The handle is borrowed: it stays valid while the Stokes operator
 * lives and tng_op_destroy refuses it.  Blocks 12 and 21 of a diagonal
 * operator are refused, as the reference throws.  Block 22 is Zero: Replace
 * writes 0, Add leaves dst. */
tng_status tng_stokes_block(tng_op op, int row, int col, tng_op *block);

/* ---------------------------------------------------------------- solvers */

/* linalg/solvers/pcg.hpp:28-188 PCG with the identity preconditioner
 * (identity_solver.hpp:20-40) on any square operator (src and dst on the same
 * mesh).  x holds the initial guess.  Returns the iterations done and the
 * last relative residual ||r_k|| / ||r_0||. */
tng_status tng_pcg_solve(tng_op op, tng_vec x, tng_vec b, int max_iterations, double relative_tolerance,
                         double absolute_tolerance, int *iterations, double *relative_residual);
/* linalg/solvers/power_iteration.hpp:13-58 on D^{-1} A (D^{-1} = inv_diag),
 * seeded uniform start vector. */
tng_status tng_power_iteration(tng_op op, tng_vec inv_diag, int iterations, unsigned long long seed,
                               double *estimate);
/* linalg/solvers/chebyshev.hpp:102-165 with a given estimate of the largest
 * eigenvalue of D^{-1} A (lambda_max = 1.5 est, lambda_min = 0.1 est);
 * `iterations` sweeps of order `order` on A x = b. */
/* linalg/solvers/jacobi.hpp:16-60 Jacobi(inverse_diagonal, iterations, tmp,
 * omega).solve(A, x, b): `iterations` times x += omega D^{-1} (b - A x). */
tng_status tng_jacobi_smooth(tng_op op, tng_vec inv_diag, int iterations, double omega, tng_vec x, tng_vec b);
tng_status tng_chebyshev_smooth(tng_op op, tng_vec inv_diag, int order, int iterations, double lambda_max_estimate,
                                tng_vec x, tng_vec b);

/* apps/mantlecirculation/src/parameters.hpp:155-185 StokesSolverParameters
 * (same field names and defaults) plus the multigrid precision, the coarse
 * PCG parameters (stokes_solver.hpp:469-470) and the power-iteration seed. */
typedef struct {
    int krylov_restart;                         /* 10 */
    int krylov_max_iterations;                  /* 10 */
    double krylov_relative_tolerance;           /* 1e-6 */
    double krylov_absolute_tolerance;           /* 1e-12 */
    int viscous_pc_num_vcycles;                 /* 1 */
    int viscous_pc_chebyshev_order;             /* 2 */
    int viscous_pc_num_smoothing_steps_prepost; /* 2 */
    int viscous_pc_num_power_iterations;        /* 10 */
    int gca;                                    /* 0 (re-discretised coarse operators) */
    int viscous_pc_dtype;                       /* TNG_F64, or TNG_F32 for an FP32 V-cycle */
    int coarse_max_iterations;                  /* 50 */
    double coarse_relative_tolerance;           /* 1e-6 */
    double coarse_absolute_tolerance;           /* 1e-16 */
    unsigned long long seed;                    /* 12345 */
    /* Options of this library (no counterpart in StokesSolverParameters): */
    int gca_coarsest_only;   /* 0; 1: with gca = 1 or 2 the Galerkin products are formed top-down through
                                every level, but only the coarsest level's operator uses them (the levels
                                between keep re-discretised matrix-free operators; the intermediate
                                matrices are released after the set-up) */
    int gca_exact;           /* 0: the reference's interpolation rule (gca.hpp:206-208), which takes the
                                weights of fine nodes on the top face of an interior coarse cell into the
                                coarse wedge's bottom slots; 1: the Galerkin product R A P */
    int viscous_pc_smoother; /* 0: Chebyshev (chebyshev.hpp); 1: damped Jacobi (jacobi.hpp:16-60) with
                                omega = 2 / (1.3 lambda_max), lambda_max of D^{-1} A from the power
                                iteration (the reference tests' choice, test_epsilon_divdiv_stokes.cpp:563) */
    int krylov_orthogonalization; /* 0: modified Gram-Schmidt (fgmres.hpp:211-224); 1: classical
                                     Gram-Schmidt with one re-orthogonalisation (CGS2), the projections of
                                     a pass fused into one multi-dot */
    int rotation_correction; /* 0; 1 (free slip on both surfaces): after the block preconditioner, a
                                correction on the three rigid rotations, the near-null modes of the
                                penalised free-slip operator: z += Z E^{-1} Z^T (v - K z), E = Z^T K Z */
    int coarse_solver;       /* 0: PCG on the coarsest level (pcg.hpp, the coarse_* fields); 1: direct,
                                the coarsest operator (boundary rows and penalty included) inverted
                                densely over the unique node components when the viscosity is set (at
                                most 6000 unknowns: a lateral level-2 coarsest level has 2430) */
} tng_stokes_params;

/* Outcome of a solve.  history (caller memory, may be NULL) receives the
 * FGMRES residual estimates, history_len counts all of them. */
typedef struct {
    int iterations;
    int restarts;
    int converged;
    int breakdowns;               /* Arnoldi breakdown / Givens stagnation restarts */
    int nonfinite_preconditioner; /* preconditioner outputs replaced by their input */
    int nonfinite_residual;
    long long coarse_iterations;  /* PCG iterations on the coarsest level */
    double initial_residual;
    double final_residual;        /* true residual ||f - K x|| at the last restart */
    double relative_residual;
    double *history;
    int history_capacity;
    int history_len;
} tng_solve_report;

typedef struct tng_stokes_solver_s *tng_stokes_solver;

void tng_stokes_params_default(tng_stokes_params *params);

/* apps/mantlecirculation/src/stokes_solver.hpp:143-642 StokesContext set-up
 * for one rank: meshes of lateral levels min_level..max_level (radial level =
 * lateral level + radial_extra_levels), velocity on max_level, pressure on
 * max_level - 1, K = EpsDivDivStokes (epsilon_divdiv_stokes.hpp:18-133),
 * multigrid (Chebyshev smoothers, constant P/R, coarse PCG), lumped-KMass
 * Schur approximation, block-triangular preconditioner and FGMRES.  The
 * viscosity starts at 1 on every level. */
tng_status tng_stokes_solver_create(tng_ctx ctx, int min_level, int max_level, int radial_extra_levels, double r_min,
                                    double r_max, int lateral_sdr, int radial_sdr, tng_bc bc_cmb, tng_bc bc_surface,
                                    const tng_stokes_params *params, tng_stokes_solver *out);
tng_status tng_stokes_solver_destroy(tng_stokes_solver solver);
tng_status tng_stokes_solver_num_levels(tng_stokes_solver solver, int *num_levels);
/* Mesh of multigrid level `level` (0 = coarsest; num_levels - 1 = velocity,
 * num_levels - 2 = pressure).  The solver keeps ownership: do not destroy. */
tng_status tng_stokes_solver_mesh(tng_stokes_solver solver, int level, tng_mesh *mesh);
/* Viscosity on the velocity mesh; coarse levels receive its injection.
 * Rebuilds the inverse diagonals, the Schur scaling (k = 1/eta on the
 * pressure level) and re-arms the Chebyshev eigenvalue estimates. */
tng_status tng_stokes_solver_set_viscosity(tng_stokes_solver solver, tng_vec eta);
/* K [u; p] (epsilon_divdiv_stokes.hpp:55-84) */
tng_status tng_stokes_apply(tng_stokes_solver solver, tng_vec u, tng_vec p, tng_vec out_u, tng_vec out_p);
/* The velocity preconditioner alone: num_vcycles V-cycles on A x = b, x as given. */
tng_status tng_stokes_precondition_velocity(tng_stokes_solver solver, tng_vec x, tng_vec b);
/* RHS enforcement of StokesContext::solve (stokes_solver.hpp:664-674) on a
 * velocity RHS: zero at the Dirichlet boundary, normal component removed at
 * the free-slip boundary; as in the reference, only the first boundary (CMB,
 * then surface) carrying each condition is treated
 * (grid/shell/bit_masks.hpp:68-76). */
tng_status tng_stokes_enforce_rhs(tng_stokes_solver solver, tng_vec f_u);
/* The RHS assembly of StokesContext::solve (stokes_solver.hpp:650-674):
 * f_u = M (Ra T x/|x|) (interpolators.hpp:99-133, vector_mass.hpp), then the
 * boundary enforcement of tng_stokes_enforce_rhs.  T: scalar field on the
 * velocity mesh. */
tng_status tng_stokes_buoyancy_rhs(tng_stokes_solver solver, tng_vec temperature, double rayleigh, tng_vec f_u);
/* FGMRES solve of K [u; p] = [f_u; f_p] from the initial guess in (u, p),
 * then the pressure mean over unique nodes is removed
 * (stokes_solver.hpp:676-689).  Returns TNG_OK also when the iteration limit
 * is reached (see report->converged), as the reference logs and continues. */
tng_status tng_stokes_solve(tng_stokes_solver solver, tng_vec u, tng_vec p, tng_vec f_u, tng_vec f_p,
                            tng_solve_report *report);

/* ---------------------------------------------------------------- adjoint */

/* No counterpart in the reference (the viscosity adjoint is new to this
 * path).  g[i] = d(lam^T A(k) u) / d log k_i for the viscous operator op
 * (FP64), with A's Dirichlet / free-slip rows and without the fs/fs penalty
 * (independent of k); summed over the copies of every node, so all copies
 * hold the derivative with respect to the unique nodal value.  lam, u:
 * velocity fields, g: scalar field, all on op's mesh. */
tng_status tng_epsdivdiv_log_coefficient_gradient(tng_op op, tng_vec lam, tng_vec u, tng_vec g);
/* Gradient of J(u) with respect to the nodal log-viscosity of the velocity
 * level, u the solution of K [u; p] = f: solves K [lam; mu] = [dJ/du; 0]
 * with the solver's FGMRES (K is symmetric; the fs/fs penalty makes it so
 * only up to its multiplicity weighting, at the penalty's 1e-7 scale) and
 * returns g = -d(lam^T A(eta) u)/d log eta.  dJdu must be consistent across
 * copies (the derivative with respect to the unique nodal values). */
tng_status tng_stokes_adjoint_gradient(tng_stokes_solver solver, tng_vec u, tng_vec dJdu, tng_vec g_log_eta,
                                       tng_solve_report *report);

/* ---------------------------------------------------------------- timing */

/* Kernel timing with HIP events on the context stream (util/timer.hpp:22-389
 * times the same phases on the host).  While profiling is on, every operator
 * apply records events around itself and around its main kernel without
 * synchronising; tng_ctx_profile_read synchronises, returns the number of
 * applies and the average durations since the last read, and resets. */
tng_status tng_ctx_set_profiling(tng_ctx ctx, int on);
tng_status tng_ctx_profile_read(tng_ctx ctx, int *num_applies, double *avg_main_kernel_ms, double *avg_total_ms);

/* ------------------------------------------------------ XDMF output / checkpoints */

/* io/xdmf.hpp:148-201 XDMFOutput(directory, domain, coords_shell, coords_radii,
 * output_type_points, output_type_connectivity): a writer of node fields of
 * `mesh` into `directory` (created on the first write if missing; existing
 * files are overwritten).  points_bytes: 4 or 8 (node coordinates, default of
 * the reference 4); connectivity_bytes: 4 or 8 (default 8). */
tng_status tng_xdmf_create(tng_mesh mesh, const char *directory, int points_bytes, int connectivity_bytes,
                           tng_xdmf *out);
/* io/xdmf.hpp:208-281 add(data, output_type): a scalar or 3-component field,
 * written with `label` (the reference's Kokkos::View label) in output_bytes
 * (4 or 8) precision on every write().  Not allowed after the first write().
 * The writer keeps the handle: the field must outlive the writer (the
 * reference holds a reference-counted view instead). */
tng_status tng_xdmf_add(tng_xdmf writer, tng_vec field, const char *label, int output_bytes);
/* io/xdmf.hpp:203-207 set_write_counter */
tng_status tng_xdmf_set_write_counter(tng_xdmf writer, int counter);
/* io/xdmf.hpp:283-513 write(): first call geometry.bin, topology.bin,
 * checkpoint_metadata.bin, README.md; every call <label>_<step>.bin per field
 * and step_<step>.xmf last.  The fields are transposed to the file layout on
 * the device.  pack_seconds / total_seconds (either may be NULL) receive the
 * device time of the layout kernels and the wall time of the call. */
tng_status tng_xdmf_write(tng_xdmf writer, double *pack_seconds, double *total_seconds);
tng_status tng_xdmf_destroy(tng_xdmf writer);
/* io/xdmf.hpp:1199-1441 read_xdmf_checkpoint_grid(directory, label, step,
 * domain, grid): reads field `label` of write step `step` into `out` (scalar
 * or 3 components, FP64 or FP32), locating every subdomain of out's mesh by
 * its global subdomain id in the checkpoint metadata (:1094-1186).  The
 * checkpoint must have the mesh's subdomain sizes.  TNG_IO_ERROR on a
 * missing or malformed file. */
tng_status tng_xdmf_checkpoint_read(const char *directory, const char *label, int step, tng_vec out);

/* ------------------------------------------------------------------------- */
/* The caller around the Stokes solve: the implicit SUPG energy step and one  */
/* step of the mantle-circulation time loop.                                  */
/* ------------------------------------------------------------------------- */
/* fe/wedge/operators/shell/unsteady_advection_diffusion_supg.hpp:128-206
 * UnsteadyAdvectionDiffusionSUPG(domain, ..., velocity, diffusivity, dt,
 * treat_boundary, diagonal, mass_scaling, lumped_mass): the scalar operator
 * mass_scaling M + dt (K_diff + K_adv + tau K_streamline), 6-point rule,
 * tau from supg_tau (:19-49); treat_boundary keeps only the diagonal entry of
 * boundary rows and removes boundary columns; Replace + CommunicateAdditively
 * by default (tng_op_set_modes applies).  The velocity is captured by
 * reference (FP64, 3 components, same mesh). */
tng_status tng_supg_create(tng_mesh mesh, tng_vec velocity, double diffusivity, double dt, int treat_boundary,
                           int diagonal, double mass_scaling, int lumped_mass, tng_op *out);
/* dt() (:206-207) and set_supg_enabled (:209-215) */
tng_status tng_supg_set_dt(tng_op op, double dt);
tng_status tng_supg_set_enabled(tng_op op, int on);
/* fe/wedge/operators/shell/mass.hpp: dst = M src, consistent scalar mass,
 * 6-point rule, Replace + CommunicateAdditively (FP64). */
tng_status tng_mass_apply(tng_vec src, tng_vec dst);

/* apps/mantlecirculation/src/parameters.hpp:186-201 EnergySolverParameters
 * (Krylov fields, same defaults), PhysicsParameters::diffusivity (1),
 * BoundaryConditionsParameters temperature_cmb (1) / temperature_surface (0),
 * TimeSteppingParameters dt_scaling (0.5) and energy_substeps (1). */
typedef struct {
    int krylov_restart;               /* 5 */
    int krylov_max_iterations;        /* 100 */
    double krylov_relative_tolerance; /* 1e-6 */
    double krylov_absolute_tolerance; /* 1e-12 */
    double diffusivity;               /* 1 */
    double temperature_cmb;           /* 1 */
    double temperature_surface;       /* 0 */
    double dt_scaling;                /* 0.5 */
    int energy_substeps;              /* 1 */
} tng_energy_params;
typedef struct tng_energy_s *tng_energy;
void tng_energy_params_default(tng_energy_params *params);
/* apps/mantlecirculation/src/energy_solver.hpp:85-276 SUPGSolver on the Stokes
 * solver's velocity mesh: advects with `velocity` (captured by reference) and
 * updates `temperature` in place.  h = the smallest radial spacing. */
tng_status tng_energy_create(tng_stokes_solver solver, tng_vec velocity, tng_vec temperature,
                             const tng_energy_params *params, tng_energy *out);
tng_status tng_energy_destroy(tng_energy energy);
/* SUPGSolver::compute_dt (:170-182): dt_scaling * h / max |u| (1e-3 for a
 * velocity below 1e-12). */
tng_status tng_energy_compute_dt(tng_energy energy, double *dt);
/* SUPGSolver::step (:194-245): inverse diagonal at dt, then per substep
 * q = M T, Dirichlet elimination (strong_algebraic_dirichlet_enforcement.hpp
 * :24-47 with the CMB / surface temperatures) and FGMRES with the diagonal
 * preconditioner on (M + dt A) T = q, T as the initial guess.  report
 * (may be NULL) sums the substeps' iterations. */
tng_status tng_energy_step(tng_energy energy, double dt, tng_solve_report *report);
/* StokesContext::update_viscosity (stokes_solver.hpp:623-638) with the
 * Frank-Kamenetskii law (interpolators.hpp:219-244): eta = rmu^(0.5 - T) on
 * the velocity level only; as in the reference, the coarse levels, inverse
 * diagonals and smoother estimates keep their values. */
tng_status tng_stokes_solver_update_viscosity(tng_stokes_solver solver, tng_vec temperature, double rmu);
/* One step of the time loop (apps/mantlecirculation/mantlecirculation.cpp
 * :413-445): snapshot T, dt from the current velocity, then per Picard
 * iteration (T restored after the first) the Stokes solve with the buoyancy of
 * T (Ra T r/|r|, from the current (u, p)), the energy step, and (law 1) the
 * viscosity update.  u must be the energy solver's velocity. */
typedef struct {
    int picard_iterations;  /* TimeSteppingParameters::picard_iterations (1) */
    int viscosity_law;      /* 0 constant, 1 Frank-Kamenetskii */
    double rmu;             /* ViscosityParameters::rmu (1) */
    double rayleigh_number; /* PhysicsParameters::rayleigh_number (1e5) */
    /* 0: as the reference, the new viscosity reaches the velocity level only;
     * 1: then also the coarse levels (injection) with the inverse diagonals,
     * Galerkin coarse operators, Schur scaling and smoother estimates rebuilt,
     * as tng_stokes_solver_set_viscosity does (no counterpart in the reference:
     * for large viscosity contrasts, where the stale hierarchy stalls FGMRES). */
    int refresh_hierarchy;
} tng_timestep_params;
tng_status tng_timestep(tng_stokes_solver solver, tng_energy energy, tng_vec u, tng_vec p,
                        const tng_timestep_params *params, double *dt, tng_solve_report *stokes_report,
                        tng_solve_report *energy_report);

#ifdef __cplusplus
}
#endif

#endif /* TNG_H */
