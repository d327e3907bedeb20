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
    TNG_UNSUPPORTED = 6
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

/* ---------------------------------------------------------------- context */

/* util/init.hpp:75-80 terra_initialize (device selection + stream) */
tng_status tng_ctx_create(int device_ordinal, tng_ctx *out);
tng_status tng_ctx_destroy(tng_ctx ctx);
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
tng_status tng_op_destroy(tng_op op);

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

/* ---------------------------------------------------------------- timing */

/* Kernel timing with HIP events on the context stream (util/timer.hpp:22-389
 * times the same phases on the host).  While profiling is on, every operator
 * apply records events around itself and around its main kernel without
 * synchronising; tng_ctx_profile_read synchronises, returns the number of
 * applies and the average durations since the last read, and resets. */
tng_status tng_ctx_set_profiling(tng_ctx ctx, int on);
tng_status tng_ctx_profile_read(tng_ctx ctx, int *num_applies, double *avg_main_kernel_ms, double *avg_total_ms);

#ifdef __cplusplus
}
#endif

#endif /* TNG_H */
