// C ABI entry points (include/tng.h).  Every call converts library errors
// into a tng_status and records the message on the context.
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "tng_internal.hpp"

using namespace tng;

namespace {

thread_local std::string g_orphan_error;

template <class F> tng_status guarded(tng_context *ctx, F &&f)
{
    try {
        f();
        return TNG_OK;
    } catch (const Error &e) {
        if (ctx) ctx->last_error = e.what();
        else g_orphan_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->last_error = "host allocation failed";
        return TNG_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        else g_orphan_error = e.what();
        return TNG_DEVICE_ERROR;
    }
}

void set_device(tng_context *ctx) { TNG_HIP_CHECK(hipSetDevice(ctx->device)); }

size_t dtype_size(tng_dtype t) { return t == TNG_F64 ? 8 : 4; }

}  // namespace

extern "C" {

const char *tng_version(void) { return "tng-mi355x 0.1 (gfx950)"; }

const char *tng_last_error(tng_ctx ctx) { return ctx ? ctx->last_error.c_str() : g_orphan_error.c_str(); }

tng_status tng_ctx_create(int device_ordinal, tng_ctx *out)
{
    if (!out) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    auto *ctx = new (std::nothrow) tng_context();
    if (!ctx) return TNG_OUT_OF_MEMORY;
    tng_status st = guarded(nullptr, [&] {
        int n = 0;
        TNG_HIP_CHECK(hipGetDeviceCount(&n));
        TNG_REQUIRE(device_ordinal >= 0 && device_ordinal < n, "device ordinal out of range");
        ctx->device = device_ordinal;
        set_device(ctx);
        TNG_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->red_partials.alloc(4096 * 3 * sizeof(double));
        ctx->red_result.alloc(4096);
        noalloc_scratch_init(device_ordinal);
        TNG_HIP_CHECK(hipDeviceSynchronize());
    });
    if (st != TNG_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return TNG_OK;
}

tng_status tng_ctx_create_with_stream(int device_ordinal, void *hip_stream, tng_ctx *out)
{
    if (!out || !hip_stream) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    auto *ctx = new (std::nothrow) tng_context();
    if (!ctx) return TNG_OUT_OF_MEMORY;
    tng_status st = guarded(nullptr, [&] {
        int n = 0;
        TNG_HIP_CHECK(hipGetDeviceCount(&n));
        TNG_REQUIRE(device_ordinal >= 0 && device_ordinal < n, "device ordinal out of range");
        ctx->device = device_ordinal;
        set_device(ctx);
        ctx->stream = static_cast<hipStream_t>(hip_stream);
        ctx->owns_stream = false;
        ctx->red_partials.alloc(4096 * 3 * sizeof(double));
        ctx->red_result.alloc(4096);
        noalloc_scratch_init(device_ordinal);
        TNG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
    if (st != TNG_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return TNG_OK;
}

tng_status tng_ctx_destroy(tng_ctx ctx)
{
    if (!ctx) return TNG_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->prof_events)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream && ctx->owns_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TNG_OK;
}

tng_status tng_ctx_synchronize(tng_ctx ctx)
{
    if (!ctx) return TNG_INVALID_ARGUMENT;
    return guarded(ctx, [&] { TNG_HIP_CHECK(hipStreamSynchronize(ctx->stream)); });
}

tng_status tng_ctx_set_tuning(tng_ctx ctx, const char *name, int value)
{
    if (!ctx || !name) return TNG_INVALID_ARGUMENT;
    const tng_tuning dflt;
    struct Entry {
        const char *name;
        int tng_tuning::*field;
    };
    static const Entry entries[] = {
        {"apply_blocks", &tng_tuning::apply_blocks},
        {"apply_seg", &tng_tuning::apply_seg},
        {"apply_radial_split", &tng_tuning::apply_radial_split},
        {"apply_radial_chunks", &tng_tuning::apply_radial_chunks},
        {"apply_waves", &tng_tuning::apply_waves},
        {"cheb_fuse", &tng_tuning::cheb_fuse},
    };
    for (const Entry &e : entries)
        if (std::strcmp(e.name, name) == 0) {
            ctx->tuning.*e.field = value < 0 ? dflt.*e.field : value;
            return TNG_OK;
        }
    ctx->last_error = std::string("unknown tuning setting: ") + name;
    return TNG_INVALID_ARGUMENT;
}

tng_status tng_ctx_set_profiling(tng_ctx ctx, int on)
{
    if (!ctx) return TNG_INVALID_ARGUMENT;
    ctx->profiling = on != 0;
    ctx->prof_used = 0;
    return TNG_OK;
}

tng_status tng_ctx_profile_read(tng_ctx ctx, int *num_applies, double *avg_main_kernel_ms, double *avg_total_ms)
{
    if (!ctx) return TNG_INVALID_ARGUMENT;
    return guarded(ctx, [&] {
        TNG_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        double sm = 0, st = 0;
        for (size_t i = 0; i < ctx->prof_used; ++i) {
            hipEvent_t *e = &ctx->prof_events[4 * i];
            float a = 0, b = 0;
            TNG_HIP_CHECK(hipEventElapsedTime(&a, e[1], e[2]));
            TNG_HIP_CHECK(hipEventElapsedTime(&b, e[0], e[3]));
            sm += a;
            st += b;
        }
        const size_t n = ctx->prof_used;
        if (num_applies) *num_applies = (int)n;
        if (avg_main_kernel_ms) *avg_main_kernel_ms = n ? sm / n : 0.0;
        if (avg_total_ms) *avg_total_ms = n ? st / n : 0.0;
        ctx->prof_used = 0;
    });
}

tng_status tng_mesh_create_uniform(tng_ctx ctx, int lateral_level, int radial_level, double r_min, double r_max,
                                   int lateral_sdr, int radial_sdr, tng_mesh *out)
{
    if (!out) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(ctx, [&] { *out = new_mesh(ctx, lateral_level, radial_level, r_min, r_max, lateral_sdr, radial_sdr); });
}

// Destruction touches no other object (the context, mesh or solver may be
// gone already, e.g. when a garbage collector finalises in any order):
// hipFree synchronises with the device by itself.
tng_status tng_mesh_destroy(tng_mesh mesh)
{
    if (!mesh || mesh->solver_owned) return TNG_INVALID_ARGUMENT;
    delete mesh;
    return TNG_OK;
}

tng_status tng_mesh_get_info(tng_mesh m, tng_mesh_info *info)
{
    if (!m || !info) return TNG_INVALID_ARGUMENT;
    info->num_subdomains = m->nsd;
    info->nodes_lateral = m->nx;
    info->nodes_radial = m->nr;
    info->lateral_level = m->lat_level;
    info->radial_level = m->rad_level;
    info->lateral_sdr = m->lat_sdr;
    info->radial_sdr = m->rad_sdr;
    info->num_alloc = m->nalloc;
    info->num_owned = m->num_owned;
    info->num_interface = m->num_interface;
    return TNG_OK;
}

tng_status tng_mesh_copy_coords(tng_mesh m, double *out)
{
    if (!m || !out) return TNG_INVALID_ARGUMENT;
    std::memcpy(out, m->coords.data(), m->coords.size() * sizeof(double));
    return TNG_OK;
}
tng_status tng_mesh_copy_radii(tng_mesh m, double *out)
{
    if (!m || !out) return TNG_INVALID_ARGUMENT;
    std::memcpy(out, m->radii.data(), m->radii.size() * sizeof(double));
    return TNG_OK;
}
tng_status tng_mesh_copy_boundary_mask(tng_mesh m, uint8_t *out)
{
    if (!m || !out) return TNG_INVALID_ARGUMENT;
    std::memcpy(out, m->bmask.data(), m->bmask.size());
    return TNG_OK;
}
tng_status tng_mesh_copy_ownership_mask(tng_mesh m, uint8_t *out)
{
    if (!m || !out) return TNG_INVALID_ARGUMENT;
    std::memcpy(out, m->omask.data(), m->omask.size());
    return TNG_OK;
}

tng_status tng_vec_create(tng_mesh mesh, tng_dtype dtype, int ncomp, tng_vec *out)
{
    if (!mesh || !out || (ncomp != 1 && ncomp != 3) || (dtype != TNG_F64 && dtype != TNG_F32))
        return TNG_INVALID_ARGUMENT;
    if (!mesh->ctx) return TNG_INVALID_ARGUMENT;  // host-only mesh
    *out = nullptr;
    return guarded(mesh->ctx, [&] { *out = new_vector(mesh, dtype, ncomp); });
}

tng_status tng_vec_wrap(tng_mesh mesh, tng_dtype dtype, int ncomp, void *const *component_ptrs, tng_vec *out)
{
    if (!mesh || !out || !component_ptrs || (ncomp != 1 && ncomp != 3) || (dtype != TNG_F64 && dtype != TNG_F32))
        return TNG_INVALID_ARGUMENT;
    if (!mesh->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    for (int c = 0; c < ncomp; ++c) {
        if (!component_ptrs[c]) return TNG_INVALID_ARGUMENT;
        if (reinterpret_cast<uintptr_t>(component_ptrs[c]) % dtype_size(dtype)) return TNG_INVALID_ARGUMENT;
    }
    auto *v = new (std::nothrow) tng_vector();
    if (!v) return TNG_OUT_OF_MEMORY;
    v->mesh = mesh;
    v->dtype = dtype;
    v->ncomp = ncomp;
    for (int c = 0; c < ncomp; ++c) v->cp[c] = component_ptrs[c];
    *out = v;
    return TNG_OK;
}

tng_status tng_vec_destroy(tng_vec v)
{
    if (!v) return TNG_INVALID_ARGUMENT;
    delete v;
    return TNG_OK;
}

tng_status tng_vec_upload(tng_vec v, const void *host)
{
    if (!v || !host) return TNG_INVALID_ARGUMENT;
    return guarded(v->mesh->ctx, [&] {
        set_device(v->mesh->ctx);
        const size_t bytes = (size_t)v->mesh->nalloc * v->elem_size();
        for (int c = 0; c < v->ncomp; ++c)
            TNG_HIP_CHECK(hipMemcpyAsync(v->comp(c), static_cast<const char *>(host) + c * bytes, bytes,
                                         hipMemcpyHostToDevice, v->mesh->ctx->stream));
        TNG_HIP_CHECK(hipStreamSynchronize(v->mesh->ctx->stream));
    });
}

tng_status tng_vec_download(tng_vec v, void *host)
{
    if (!v || !host) return TNG_INVALID_ARGUMENT;
    return guarded(v->mesh->ctx, [&] {
        set_device(v->mesh->ctx);
        const size_t bytes = (size_t)v->mesh->nalloc * v->elem_size();
        for (int c = 0; c < v->ncomp; ++c)
            TNG_HIP_CHECK(hipMemcpyAsync(static_cast<char *>(host) + c * bytes, v->comp(c), bytes,
                                         hipMemcpyDeviceToHost, v->mesh->ctx->stream));
        TNG_HIP_CHECK(hipStreamSynchronize(v->mesh->ctx->stream));
    });
}

tng_status tng_vec_device_ptr(tng_vec v, int comp, void **ptr)
{
    if (!v || !ptr || comp < 0 || comp >= v->ncomp) return TNG_INVALID_ARGUMENT;
    *ptr = v->comp(comp);
    return TNG_OK;
}

tng_status tng_vec_fill(tng_vec v, double value)
{
    if (!v) return TNG_INVALID_ARGUMENT;
    return guarded(v->mesh->ctx, [&] {
        set_device(v->mesh->ctx);
        launch_fill(*v, value, v->mesh->ctx->stream);
    });
}

tng_status tng_vec_copy(tng_vec dst, tng_vec src)
{
    if (!dst || !src) return TNG_INVALID_ARGUMENT;
    return guarded(dst->mesh->ctx, [&] {
        set_device(dst->mesh->ctx);
        launch_copy(*dst, *src, dst->mesh->ctx->stream);
    });
}

tng_status tng_epsdivdiv_create(tng_mesh mesh, tng_vec k, tng_bc bc_cmb, tng_bc bc_surface, int diagonal, tng_op *out)
{
    if (!mesh || !k || !out || !mesh->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(mesh->ctx, [&] { *out = new_epsdivdiv(mesh, k, bc_cmb, bc_surface, diagonal != 0); });
}

tng_status tng_op_set_modes(tng_op op, tng_apply_mode apply_mode, tng_comm_mode comm_mode)
{
    if (!op || (apply_mode != TNG_REPLACE && apply_mode != TNG_ADD) ||
        (comm_mode != TNG_SKIP_COMMUNICATION && comm_mode != TNG_COMMUNICATE_ADDITIVELY))
        return TNG_INVALID_ARGUMENT;
    op->apply_mode = apply_mode;
    op->comm_mode = comm_mode;
    return TNG_OK;
}

tng_status tng_op_set_diagonal(tng_op op, int diagonal)
{
    if (!op) return TNG_INVALID_ARGUMENT;
    op->diagonal = diagonal != 0;
    return TNG_OK;
}

tng_status tng_op_set_boundary_conditions(tng_op op, tng_bc bc_cmb, tng_bc bc_surface)
{
    if (!op || bc_cmb < TNG_NEUMANN || bc_cmb > TNG_FREESLIP || bc_surface < TNG_NEUMANN || bc_surface > TNG_FREESLIP)
        return TNG_INVALID_ARGUMENT;
    op->bc_cmb = bc_cmb;
    op->bc_surf = bc_surface;
    op->btab_valid = false;  // the penalty stays as armed at creation, as in the reference
    return TNG_OK;
}

tng_status tng_op_set_penalty_epsilon(tng_op op, double eps)
{
    if (!op) return TNG_INVALID_ARGUMENT;
    op->penalty_eps = eps;
    return TNG_OK;
}

tng_status tng_op_coefficient_updated(tng_op op)
{
    if (!op) return TNG_INVALID_ARGUMENT;
    op->btab_valid = false;
    return TNG_OK;
}

tng_status tng_op_apply(tng_op op, tng_vec src, tng_vec dst)
{
    if (!op || !src || !dst) return TNG_INVALID_ARGUMENT;
    tng_context *ctx = op->mesh->ctx;
    return guarded(ctx, [&] {
        set_device(ctx);
        op_apply(*op, *src, *dst);
    });
}

tng_status tng_epsdivdiv_apply_noalloc(tng_op op, tng_vec src, tng_vec dst)
{
    if (!op || !src || !dst) return TNG_INVALID_ARGUMENT;
    tng_context *ctx = op->mesh->ctx;
    return guarded(ctx, [&] {
        TNG_REQUIRE(op->kind == OP_EPSDIVDIV, "allocation-free apply: an EpsilonDivDiv operator is required");
        TNG_REQUIRE(src != dst, "apply: src and dst must be different vectors");
        TNG_REQUIRE(src->mesh == op->mesh && dst->mesh == op->mesh && src->ncomp == 3 && dst->ncomp == 3,
                    "apply: vectors do not live on the operator's mesh");
        TNG_REQUIRE(src->dtype == dst->dtype && op->k->dtype == src->dtype,
                    "apply: coefficient and vectors must share one dtype");
        set_device(ctx);
        epsdivdiv_apply_noalloc(*op, *src, *dst, ctx->stream);
    });
}

tng_status tng_op_destroy(tng_op op)
{
    if (!op) return TNG_INVALID_ARGUMENT;
    if (op->parent) {
        op->mesh->ctx->last_error = "destroy: a block of a Stokes operator is released with the Stokes operator";
        return TNG_INVALID_ARGUMENT;
    }
    delete op;
    return TNG_OK;
}

}  // extern "C"

tng_operator::~tng_operator()
{
    for (tng_operator *b : blocks) delete b;
}

hipEvent_t *tng_context::prof_next_set()
{
    if (4 * (prof_used + 1) > prof_events.size()) {
        const size_t old = prof_events.size();
        prof_events.resize(old + 4 * 64, nullptr);
        for (size_t i = old; i < prof_events.size(); ++i) TNG_HIP_CHECK(hipEventCreate(&prof_events[i]));
    }
    return &prof_events[4 * prof_used++];
}

namespace tng {

tng_domain *new_mesh(tng_context *ctx, int lat_level, int rad_level, double r_min, double r_max, int lat_sdr,
                     int rad_sdr)
{
    std::unique_ptr<tng_domain> m(new tng_domain());
    m->ctx = ctx;
    if (ctx) set_device(ctx);
    mesh_build(*m, lat_level, rad_level, r_min, r_max, lat_sdr, rad_sdr, ctx != nullptr);
    if (ctx) mesh_upload(*m);
    return m.release();
}

tng_vector *new_vector(tng_domain *mesh, tng_dtype dtype, int ncomp)
{
    TNG_REQUIRE(mesh && mesh->ctx, "vectors need a mesh with a device context");
    TNG_REQUIRE((ncomp == 1 || ncomp == 3) && (dtype == TNG_F64 || dtype == TNG_F32), "invalid vector shape");
    set_device(mesh->ctx);
    std::unique_ptr<tng_vector> v(new tng_vector());
    v->mesh = mesh;
    v->dtype = dtype;
    v->ncomp = ncomp;
    // components start on 256-byte boundaries
    const int64_t align = 256 / (int64_t)dtype_size(dtype);
    const int64_t stride = (mesh->nalloc + align - 1) / align * align;
    v->buf.alloc((size_t)stride * ncomp * dtype_size(dtype));
    for (int c = 0; c < ncomp; ++c) v->cp[c] = static_cast<char *>(v->buf.p) + (size_t)c * stride * dtype_size(dtype);
    launch_fill(*v, 0.0, mesh->ctx->stream);
    return v.release();
}

static void require_bc(int bc_cmb, int bc_surf)
{
    TNG_REQUIRE(bc_cmb >= TNG_NEUMANN && bc_cmb <= TNG_FREESLIP && bc_surf >= TNG_NEUMANN && bc_surf <= TNG_FREESLIP,
                "invalid boundary condition");
}

tng_operator *new_epsdivdiv(tng_domain *mesh, tng_vector *k, int bc_cmb, int bc_surf, bool diagonal)
{
    TNG_REQUIRE(k->mesh == mesh && k->ncomp == 1, "coefficient must be a scalar field on the operator's mesh");
    require_bc(bc_cmb, bc_surf);
    set_device(mesh->ctx);
    std::unique_ptr<tng_operator> op(new tng_operator());
    op->kind = OP_EPSDIVDIV;
    op->mesh = mesh;
    op->k = k;
    op->bc_cmb = bc_cmb;
    op->bc_surf = bc_surf;
    op->diagonal = diagonal;
    op->penalty_active = bc_cmb == TNG_FREESLIP && bc_surf == TNG_FREESLIP;
    // the boundary-row table, allocated here so that applies never allocate it
    op->d_btab.alloc((size_t)mesh->nlat * 2 * 3 * k->elem_size());
    op->d_penalty_b.alloc(4 * sizeof(double));
    TNG_HIP_CHECK(hipMemsetAsync(op->d_penalty_b.p, 0, 4 * sizeof(double), mesh->ctx->stream));
    epsdivdiv_update_boundary_table(*op, mesh->ctx->stream);
    return op.release();
}

tng_operator *new_two_level(OpKind kind, tng_domain *fine, tng_domain *coarse, int bc_cmb, int bc_surf)
{
    TNG_REQUIRE(fine->ctx && fine->ctx == coarse->ctx, "meshes must share one context");
    TNG_REQUIRE(fine->nsd == coarse->nsd && fine->nx - 1 == 2 * (coarse->nx - 1) &&
                    fine->nr - 1 == 2 * (coarse->nr - 1) && fine->lat_sdr == coarse->lat_sdr &&
                    fine->rad_sdr == coarse->rad_sdr,
                "fine and coarse meshes must be one uniform refinement apart with the same subdomains");
    require_bc(bc_cmb, bc_surf);
    std::unique_ptr<tng_operator> op(new tng_operator());
    op->kind = kind;
    op->mesh = fine;
    op->mesh2 = coarse;
    op->bc_cmb = bc_cmb;
    op->bc_surf = bc_surf;
    if (kind == OP_PROLONGATION) op->comm_mode = TNG_SKIP_COMMUNICATION;
    return op.release();
}

tng_operator *new_kmass(tng_domain *mesh, tng_vector *k)
{
    TNG_REQUIRE(k->mesh == mesh && k->ncomp == 1, "coefficient must be a scalar field on the operator's mesh");
    std::unique_ptr<tng_operator> op(new tng_operator());
    op->kind = OP_KMASS_LUMPED;
    op->mesh = mesh;
    op->k = k;
    return op.release();
}

void op_apply(tng_operator &op, const tng_vector &src, tng_vector &dst)
{
    TNG_REQUIRE(&src != &dst, "apply: src and dst must be different vectors");
    TNG_REQUIRE(src.dtype == dst.dtype, "apply: src and dst must share one dtype");
    tng_domain *in = nullptr, *out = nullptr;
    int nin = 3, nout = 3;
    switch (op.kind) {
    case OP_EPSDIVDIV: in = out = op.mesh; break;
    case OP_DIVERGENCE: in = op.mesh; out = op.mesh2; nout = 1; break;
    case OP_GRADIENT: in = op.mesh2; out = op.mesh; nin = 1; break;
    case OP_PROLONGATION: in = op.mesh2; out = op.mesh; nin = nout = src.ncomp; break;
    case OP_RESTRICTION: in = op.mesh; out = op.mesh2; nin = nout = src.ncomp; break;
    case OP_KMASS_LUMPED: in = out = op.mesh; nin = nout = 1; break;
    case OP_VECTOR_MASS: in = out = op.mesh; break;
    case OP_ZERO: in = out = op.mesh; nin = nout = 1; break;
    case OP_SUPG: in = out = op.mesh; nin = nout = 1; break;
    case OP_STOKES: throw Error(TNG_INVALID_ARGUMENT, "apply: a Stokes operator takes (u, p) pairs: use tng_stokes_op_apply");
    }
    TNG_REQUIRE(src.mesh == in && dst.mesh == out, "apply: vectors do not live on the operator's meshes");
    TNG_REQUIRE(src.ncomp == nin && dst.ncomp == nout, "apply: wrong number of vector components");
    if (op.k) TNG_REQUIRE(op.k->dtype == src.dtype, "apply: coefficient and vectors must share one dtype");
    hipStream_t s = op.mesh->ctx->stream;
    if (op.kind == OP_EPSDIVDIV) epsdivdiv_apply(op, src, dst, s);
    else if (op.kind == OP_SUPG) supg_apply(op, src, dst, s);
    else if (op.kind == OP_ZERO) {
        if (op.apply_mode == TNG_REPLACE) launch_fill(dst, 0.0, s);
    } else stokes_block_apply(op, src, dst, s);
}

// EpsDivDivStokes::apply_impl (epsilon_divdiv_stokes.hpp:55-84): A (Replace,
// SkipCommunication) -> B^T (Add, Communicate) -> B (Replace, Communicate),
// B^T and B skipped for the diagonal operator; every block is left in
// Replace + Communicate.
void stokes_op_apply(tng_operator &op, const tng_vector &u, const tng_vector &p, tng_vector &out_u, tng_vector &out_p)
{
    tng_operator &A = *op.blocks[0], &Bt = *op.blocks[1], &B = *op.blocks[2];
    TNG_REQUIRE(u.mesh == op.mesh && out_u.mesh == op.mesh && p.mesh == op.mesh2 && out_p.mesh == op.mesh2,
                "Stokes apply: velocity on the fine mesh, pressure on the coarse mesh");
    if (!op.diagonal) {
        A.apply_mode = TNG_REPLACE;
        A.comm_mode = TNG_SKIP_COMMUNICATION;
    }
    Bt.apply_mode = TNG_ADD;
    Bt.comm_mode = TNG_COMMUNICATE_ADDITIVELY;
    B.apply_mode = TNG_REPLACE;
    B.comm_mode = TNG_COMMUNICATE_ADDITIVELY;
    op_apply(A, u, out_u);
    if (!op.diagonal) {
        op_apply(Bt, p, out_u);
        op_apply(B, u, out_p);
    }
    for (tng_operator *b : {&A, &Bt, &B}) {
        b->apply_mode = TNG_REPLACE;
        b->comm_mode = TNG_COMMUNICATE_ADDITIVELY;
    }
}

}  // namespace tng

extern "C" {

tng_status tng_divergence_create(tng_mesh fine, tng_mesh coarse, tng_bc bc_cmb, tng_bc bc_surface, tng_op *out)
{
    if (!fine || !coarse || !out || !fine->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(fine->ctx, [&] { *out = new_two_level(OP_DIVERGENCE, fine, coarse, bc_cmb, bc_surface); });
}

tng_status tng_gradient_create(tng_mesh fine, tng_mesh coarse, tng_bc bc_cmb, tng_bc bc_surface, tng_op *out)
{
    if (!fine || !coarse || !out || !fine->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(fine->ctx, [&] { *out = new_two_level(OP_GRADIENT, fine, coarse, bc_cmb, bc_surface); });
}

tng_status tng_prolongation_create(tng_mesh coarse, tng_mesh fine, tng_op *out)
{
    if (!fine || !coarse || !out || !fine->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(fine->ctx, [&] { *out = new_two_level(OP_PROLONGATION, fine, coarse, TNG_NEUMANN, TNG_NEUMANN); });
}

tng_status tng_restriction_create(tng_mesh fine, tng_mesh coarse, tng_op *out)
{
    if (!fine || !coarse || !out || !fine->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(fine->ctx, [&] { *out = new_two_level(OP_RESTRICTION, fine, coarse, TNG_NEUMANN, TNG_NEUMANN); });
}

tng_status tng_kmass_lumped_create(tng_mesh mesh, tng_vec k, tng_op *out)
{
    if (!mesh || !k || !out || !mesh->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(mesh->ctx, [&] { *out = new_kmass(mesh, k); });
}

tng_status tng_vector_mass_create(tng_mesh mesh, tng_op *out)
{
    if (!mesh || !out || !mesh->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(mesh->ctx, [&] {
        std::unique_ptr<tng_operator> op(new tng_operator());
        op->kind = OP_VECTOR_MASS;
        op->mesh = mesh;
        *out = op.release();
    });
}

tng_status tng_vec_lincomb(tng_vec y, int n, const double *c, const tng_vec *x, double c0)
{
    if (!y || n < 0 || n > 3 || (n > 0 && (!c || !x))) return TNG_INVALID_ARGUMENT;
    return guarded(y->mesh->ctx, [&] {
        set_device(y->mesh->ctx);
        const tng_vector *xs[3] = {nullptr, nullptr, nullptr};
        for (int i = 0; i < n; ++i) {
            TNG_REQUIRE(x[i] != nullptr, "lincomb: null input vector");
            xs[i] = x[i];
        }
        launch_lincomb(*y, n, c, xs, c0, y->mesh->ctx->stream);
    });
}

tng_status tng_vec_scale_with_vector(tng_vec y, tng_vec x)
{
    if (!y || !x) return TNG_INVALID_ARGUMENT;
    return guarded(y->mesh->ctx, [&] {
        set_device(y->mesh->ctx);
        launch_mult(*y, *x, y->mesh->ctx->stream);
    });
}

tng_status tng_vec_invert(tng_vec y)
{
    if (!y) return TNG_INVALID_ARGUMENT;
    return guarded(y->mesh->ctx, [&] {
        set_device(y->mesh->ctx);
        launch_invert(*y, y->mesh->ctx->stream);
    });
}

tng_status tng_vec_dot(tng_vec x, tng_vec y, double *out)
{
    if (!x || !y || !out) return TNG_INVALID_ARGUMENT;
    return guarded(x->mesh->ctx, [&] {
        set_device(x->mesh->ctx);
        *out = reduce_to_host(REDUCE_DOT, *x, y, x->mesh->ctx->stream);
    });
}

tng_status tng_vec_max_abs(tng_vec x, double *out)
{
    if (!x || !out) return TNG_INVALID_ARGUMENT;
    return guarded(x->mesh->ctx, [&] {
        set_device(x->mesh->ctx);
        *out = reduce_to_host(REDUCE_MAXABS, *x, nullptr, x->mesh->ctx->stream);
    });
}

tng_status tng_vec_has_nan_or_inf(tng_vec x, int *out)
{
    if (!x || !out) return TNG_INVALID_ARGUMENT;
    return guarded(x->mesh->ctx, [&] {
        set_device(x->mesh->ctx);
        *out = reduce_to_host(REDUCE_NONFINITE, *x, nullptr, x->mesh->ctx->stream) > 0 ? 1 : 0;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// EpsDivDivStokes block operator
// ---------------------------------------------------------------------------
extern "C" {

tng_status tng_stokes_create(tng_mesh fine, tng_mesh coarse, tng_vec k, tng_bc bc_cmb, tng_bc bc_surface,
                             int diagonal, tng_op *out)
{
    if (!fine || !coarse || !k || !out || !fine->ctx) return TNG_INVALID_ARGUMENT;
    *out = nullptr;
    return guarded(fine->ctx, [&] {
        set_device(fine->ctx);
        std::unique_ptr<tng_operator> op(new tng_operator());
        op->kind = OP_STOKES;
        op->mesh = fine;
        op->mesh2 = coarse;
        op->bc_cmb = bc_cmb;
        op->bc_surf = bc_surface;
        op->diagonal = diagonal != 0;
        op->blocks[1] = new_two_level(OP_GRADIENT, fine, coarse, bc_cmb, bc_surface);
        op->blocks[2] = new_two_level(OP_DIVERGENCE, fine, coarse, bc_cmb, bc_surface);
        op->blocks[0] = new_epsdivdiv(fine, k, bc_cmb, bc_surface, diagonal != 0);
        std::unique_ptr<tng_operator> zero(new tng_operator());
        zero->kind = OP_ZERO;
        zero->mesh = coarse;
        op->blocks[3] = zero.release();
        for (tng_operator *b : op->blocks) b->parent = op.get();
        *out = op.release();
    });
}

tng_status tng_stokes_op_apply(tng_op op, tng_vec u, tng_vec p, tng_vec out_u, tng_vec out_p)
{
    if (!op || !u || !p || !out_u || !out_p) return TNG_INVALID_ARGUMENT;
    if (op->kind != OP_STOKES) {
        op->mesh->ctx->last_error = "tng_stokes_op_apply: not a Stokes operator";
        return TNG_INVALID_ARGUMENT;
    }
    return guarded(op->mesh->ctx, [&] {
        set_device(op->mesh->ctx);
        stokes_op_apply(*op, *u, *p, *out_u, *out_p);
    });
}

tng_status tng_stokes_block(tng_op op, int row, int col, tng_op *block)
{
    if (!op || !block || row < 1 || row > 2 || col < 1 || col > 2) return TNG_INVALID_ARGUMENT;
    *block = nullptr;
    if (op->kind != OP_STOKES) {
        op->mesh->ctx->last_error = "tng_stokes_block: not a Stokes operator";
        return TNG_INVALID_ARGUMENT;
    }
    if (op->diagonal && row != col) {
        op->mesh->ctx->last_error = "block_12() / block_21() are not available for a diagonal Stokes operator";
        return TNG_INVALID_ARGUMENT;
    }
    *block = op->blocks[(row - 1) * 2 + (col - 1)];
    return TNG_OK;
}

tng_status tng_vec_norm2(tng_vec x, double *out)
{
    if (!x || !out) return TNG_INVALID_ARGUMENT;
    return guarded(x->mesh->ctx, [&] {
        set_device(x->mesh->ctx);
        *out = std::sqrt(reduce_to_host(REDUCE_DOT, *x, x, x->mesh->ctx->stream));
    });
}

tng_status tng_vec_dot_device(tng_vec x, tng_vec y, double *out_device)
{
    if (!x || !y || !out_device) return TNG_INVALID_ARGUMENT;
    return guarded(x->mesh->ctx, [&] {
        set_device(x->mesh->ctx);
        hipStream_t s = x->mesh->ctx->stream;
        launch_reduce(REDUCE_DOT, *x, y, 0, s);
        TNG_HIP_CHECK(hipMemcpyAsync(out_device, x->mesh->ctx->red_result.p, sizeof(double),
                                     hipMemcpyDeviceToDevice, s));
    });
}

tng_status tng_vec_fill_seeded(tng_vec x, unsigned long long seed)
{
    if (!x) return TNG_INVALID_ARGUMENT;
    return guarded(x->mesh->ctx, [&] {
        set_device(x->mesh->ctx);
        launch_randomize(*x, seed, x->mesh->ctx->stream);
    });
}

tng_status tng_op_set_stored_matrix_mode(tng_op op, int full)
{
    if (!op || op->kind != OP_EPSDIVDIV || (full != 0 && full != 1)) return TNG_INVALID_ARGUMENT;
    return guarded(op->mesh->ctx, [&] {
        set_device(op->mesh->ctx);
        hipStream_t s = op->mesh->ctx->stream;
        TNG_HIP_CHECK(hipStreamSynchronize(s));
        if (!full) {
            op->stored = false;
            op->d_lmat.release();
            return;
        }
        TNG_REQUIRE(op->mesh->nr >= 3, "stored matrices need at least two radial layers");
        const size_t es = op->k->dtype == TNG_F64 ? 8 : 4;
        const size_t bytes = stored_matrix_count(*op->mesh) * 324 * es;
        op->d_lmat.alloc(bytes);
        TNG_HIP_CHECK(hipMemsetAsync(op->d_lmat.p, 0, bytes, s));
        op->stored = true;
        op->lmat_dtype = op->k->dtype;
    });
}

static size_t local_matrix_offset(const tng_domain &m, int sd, int x, int y, int r, int wedge)
{
    TNG_REQUIRE(sd >= 0 && sd < m.nsd && x >= 0 && x < m.nx - 1 && y >= 0 && y < m.nx - 1 && r >= 0 &&
                    r < m.nr - 1 && (wedge == 0 || wedge == 1),
                "local matrix: cell index out of range");
    return ((((size_t)sd * (m.nx - 1) + x) * (m.nx - 1) + y) * (m.nr - 1) + r) * 2 + wedge;
}

tng_status tng_op_get_local_matrix(tng_op op, int sd, int x, int y, int r, int wedge, double *mat)
{
    if (!op || !mat || op->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(op->mesh->ctx, [&] {
        set_device(op->mesh->ctx);
        hipStream_t s = op->mesh->ctx->stream;
        const size_t w = local_matrix_offset(*op->mesh, sd, x, y, r, wedge);
        if (op->stored) {
            if (op->lmat_dtype == TNG_F64) {
                TNG_HIP_CHECK(hipMemcpyAsync(mat, op->d_lmat.as<double>() + w * 324, 324 * sizeof(double),
                                             hipMemcpyDeviceToHost, s));
                TNG_HIP_CHECK(hipStreamSynchronize(s));
            } else {
                float tmp[324];
                TNG_HIP_CHECK(hipMemcpyAsync(tmp, op->d_lmat.as<float>() + w * 324, sizeof(tmp),
                                             hipMemcpyDeviceToHost, s));
                TNG_HIP_CHECK(hipStreamSynchronize(s));
                for (int i = 0; i < 324; ++i) mat[i] = tmp[i];
            }
            return;
        }
        // no storage: the Neumann matrix assembled from k (assemble_local_matrix)
        DevBuf all;
        epsdivdiv_local_matrices(*op, all, s);
        TNG_HIP_CHECK(hipMemcpyAsync(mat, all.as<double>() + w * 324, 324 * sizeof(double), hipMemcpyDeviceToHost,
                                     s));
        TNG_HIP_CHECK(hipStreamSynchronize(s));
    });
}

tng_status tng_op_set_local_matrix(tng_op op, int sd, int x, int y, int r, int wedge, const double *mat)
{
    if (!op || !mat || op->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(op->mesh->ctx, [&] {
        set_device(op->mesh->ctx);
        TNG_REQUIRE(op->stored, "set_local_matrix: the operator stores no matrices (tng_op_set_stored_matrix_mode)");
        hipStream_t s = op->mesh->ctx->stream;
        const size_t w = local_matrix_offset(*op->mesh, sd, x, y, r, wedge);
        if (op->lmat_dtype == TNG_F64) {
            TNG_HIP_CHECK(hipMemcpyAsync(op->d_lmat.as<double>() + w * 324, mat, 324 * sizeof(double),
                                         hipMemcpyHostToDevice, s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));
        } else {
            float tmp[324];
            for (int i = 0; i < 324; ++i) tmp[i] = (float)mat[i];
            TNG_HIP_CHECK(hipMemcpyAsync(op->d_lmat.as<float>() + w * 324, tmp, sizeof(tmp), hipMemcpyHostToDevice,
                                         s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));
        }
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// stored local matrices and GCA
// ---------------------------------------------------------------------------
extern "C" {

tng_status tng_op_get_local_matrices(tng_op op, double *host_out)
{
    if (!op || !host_out || op->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(op->mesh->ctx, [&] {
        set_device(op->mesh->ctx);
        hipStream_t s = op->mesh->ctx->stream;
        const size_t n = stored_matrix_count(*op->mesh) * 324;
        if (op->stored && op->lmat_dtype == TNG_F64) {
            TNG_HIP_CHECK(hipMemcpyAsync(host_out, op->d_lmat.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        } else if (op->stored) {
            std::vector<float> tmp(n);
            TNG_HIP_CHECK(hipMemcpyAsync(tmp.data(), op->d_lmat.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i) host_out[i] = tmp[i];
        } else {
            DevBuf tmp;
            epsdivdiv_local_matrices(*op, tmp, s);
            TNG_HIP_CHECK(hipMemcpyAsync(host_out, tmp.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));
        }
        TNG_HIP_CHECK(hipStreamSynchronize(s));
    });
}

tng_status tng_op_set_local_matrices(tng_op op, const double *host_in)
{
    if (!op || !host_in || op->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(op->mesh->ctx, [&] {
        set_device(op->mesh->ctx);
        hipStream_t s = op->mesh->ctx->stream;
        TNG_REQUIRE(op->mesh->nr >= 3, "stored matrices need at least two radial layers");
        const size_t n = stored_matrix_count(*op->mesh) * 324;
        if (op->k->dtype == TNG_F64) {
            if (op->d_lmat.bytes != n * sizeof(double)) op->d_lmat.alloc(n * sizeof(double));
            TNG_HIP_CHECK(hipMemcpyAsync(op->d_lmat.p, host_in, n * sizeof(double), hipMemcpyHostToDevice, s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));
        } else {
            std::vector<float> tmp(host_in, host_in + n);
            if (op->d_lmat.bytes != n * sizeof(float)) op->d_lmat.alloc(n * sizeof(float));
            TNG_HIP_CHECK(hipMemcpyAsync(op->d_lmat.p, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
            TNG_HIP_CHECK(hipStreamSynchronize(s));
        }
        op->stored = true;
        op->lmat_dtype = op->k->dtype;
    });
}

tng_status tng_op_clear_local_matrices(tng_op op)
{
    if (!op || op->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(op->mesh->ctx, [&] {
        TNG_HIP_CHECK(hipStreamSynchronize(op->mesh->ctx->stream));
        op->stored = false;
        op->d_lmat.release();
    });
}

tng_status tng_gca_assemble(tng_op fine, tng_op coarse)
{
    if (!fine || !coarse || fine->kind != OP_EPSDIVDIV || coarse->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(coarse->mesh->ctx, [&] {
        set_device(coarse->mesh->ctx);
        TNG_REQUIRE(coarse->k->dtype == TNG_F64, "GCA assembles FP64 matrices: the coarse operator must be FP64");
        gca_assemble(*fine, *coarse, coarse->mesh->ctx->stream);
    });
}

tng_status tng_gca_assemble_exact(tng_op fine, tng_op coarse)
{
    if (!fine || !coarse || fine->kind != OP_EPSDIVDIV || coarse->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(coarse->mesh->ctx, [&] {
        set_device(coarse->mesh->ctx);
        TNG_REQUIRE(coarse->k->dtype == TNG_F64, "GCA assembles FP64 matrices: the coarse operator must be FP64");
        gca_assemble(*fine, *coarse, coarse->mesh->ctx->stream, true);
    });
}

tng_status tng_gca_elements_collect(tng_vec k, int level_range, tng_mesh coarsest, unsigned char *host_flags)
{
    if (!k || !coarsest || !host_flags) return TNG_INVALID_ARGUMENT;
    return guarded(coarsest->ctx, [&] {
        set_device(coarsest->ctx);
        hipStream_t s = coarsest->ctx->stream;
        DevBuf flags;
        gca_collect(*k, level_range, *coarsest, flags, s);
        TNG_HIP_CHECK(hipMemcpyAsync(host_flags, flags.p, flags.bytes, hipMemcpyDeviceToHost, s));
        TNG_HIP_CHECK(hipStreamSynchronize(s));
    });
}

tng_status tng_gca_assemble_selective(tng_op fine, tng_op coarse, int level_range, tng_mesh coarsest,
                                      const unsigned char *host_flags)
{
    if (!fine || !coarse || !coarsest || !host_flags || fine->kind != OP_EPSDIVDIV || coarse->kind != OP_EPSDIVDIV)
        return TNG_INVALID_ARGUMENT;
    return guarded(coarse->mesh->ctx, [&] {
        set_device(coarse->mesh->ctx);
        TNG_REQUIRE(coarse->k->dtype == TNG_F64, "GCA assembles FP64 matrices: the coarse operator must be FP64");
        hipStream_t s = coarse->mesh->ctx->stream;
        DevBuf flags((size_t)coarsest->nsd * (coarsest->nx - 1) * (coarsest->nx - 1) * (coarsest->nr - 1));
        TNG_HIP_CHECK(hipMemcpyAsync(flags.p, host_flags, flags.bytes, hipMemcpyHostToDevice, s));
        gca_assemble_selective(*fine, *coarse, level_range, *coarsest, flags, s);
    });
}

tng_status tng_xdmf_create(tng_mesh mesh, const char *directory, int points_bytes, int connectivity_bytes,
                           tng_xdmf *out)
{
    if (!mesh || !mesh->ctx || !directory || !out) return TNG_INVALID_ARGUMENT;
    return guarded(mesh->ctx, [&] {
        TNG_REQUIRE(points_bytes == 4 || points_bytes == 8, "XDMF: points_bytes must be 4 or 8");
        TNG_REQUIRE(connectivity_bytes == 4 || connectivity_bytes == 8, "XDMF: connectivity_bytes must be 4 or 8");
        std::unique_ptr<tng_xdmf_writer> w(new tng_xdmf_writer());
        w->mesh = mesh;
        w->dir = directory;
        w->points_bytes = points_bytes;
        w->conn_bytes = connectivity_bytes;
        *out = w.release();
    });
}

tng_status tng_xdmf_add(tng_xdmf writer, tng_vec field, const char *label, int output_bytes)
{
    if (!writer || !field || !label) return TNG_INVALID_ARGUMENT;
    return guarded(writer->mesh->ctx, [&] { writer->add(field, label, output_bytes); });
}

tng_status tng_xdmf_set_write_counter(tng_xdmf writer, int counter)
{
    if (!writer || counter < 0) return TNG_INVALID_ARGUMENT;
    writer->counter = counter;
    return TNG_OK;
}

tng_status tng_xdmf_write(tng_xdmf writer, double *pack_seconds, double *total_seconds)
{
    if (!writer) return TNG_INVALID_ARGUMENT;
    return guarded(writer->mesh->ctx, [&] {
        set_device(writer->mesh->ctx);
        writer->write();
        if (pack_seconds) *pack_seconds = writer->last_pack_seconds;
        if (total_seconds) *total_seconds = writer->last_write_seconds;
    });
}

tng_status tng_xdmf_destroy(tng_xdmf writer)
{
    // like the other destroy calls: touches neither the mesh nor the context,
    // which may already be gone when a garbage collector finalises in any order
    if (!writer) return TNG_INVALID_ARGUMENT;
    delete writer;
    return TNG_OK;
}

tng_status tng_xdmf_checkpoint_read(const char *directory, const char *label, int step, tng_vec out)
{
    if (!directory || !label || !out || step < 0) return TNG_INVALID_ARGUMENT;
    return guarded(out->mesh->ctx, [&] {
        set_device(out->mesh->ctx);
        xdmf_checkpoint_read(directory, label, step, *out);
    });
}

tng_status tng_op_copy_local_matrices(tng_op dst, tng_op src)
{
    if (!dst || !src || dst->kind != OP_EPSDIVDIV || src->kind != OP_EPSDIVDIV) return TNG_INVALID_ARGUMENT;
    return guarded(dst->mesh->ctx, [&] {
        set_device(dst->mesh->ctx);
        copy_stored_matrices(*dst, *src, dst->k->dtype, dst->mesh->ctx->stream);
    });
}

}  // extern "C"
