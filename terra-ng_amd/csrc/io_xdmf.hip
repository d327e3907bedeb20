// XDMF output and checkpoints of node fields (reference: io/xdmf.hpp:15-1502,
// util/xml.hpp, util/filesystem.hpp:16-44).
//
// Directory layout written by write() (io/xdmf.hpp:293-513):
//   geometry.bin, topology.bin, checkpoint_metadata.bin, README.md  (first call)
//   <label>_<step>.bin per added field, then step_<step>.xmf          (every call)
// Binary layout of every data file (io/xdmf.hpp:74-97): per local subdomain,
// r slowest, then y, then x, then the vector component.  The device layout is
// (subdomain, x, y, r) with r fastest and one array per component, so writing
// a field is a transpose of each (subdomain, y) plane [x][r] -> [r][x] with
// the components interleaved and the precision converted: xdmf_pack_kernel
// stages 32 x 32 tiles in LDS so that both the reads (along r) and the
// writes (along x) are contiguous.  One device-to-host copy of the packed
// field into pinned memory follows, then one file write.  Reading a
// checkpoint is the inverse transpose (xdmf_unpack_kernel).
//
// Geometry (io/xdmf.hpp:578-604): the unit-sphere node coordinate times the
// shell radius, in the file order.  Topology (:610-651): per hex cell of the
// file order, wedge 0 = (v0, v1, v2, v4, v5, v6) and wedge 1 = (v3, v2, v1, v7,
// v6, v5) with v0 = node (x, y, r), v1 = x+1, v2 = y+1, v3 = x+1,y+1 and v4..v7
// the same one level up, as linear node indices of the file order.
#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "tng_internal.hpp"

namespace tng {

namespace {

constexpr int kTile = 32;

// src: device layout, one array per component; dst: file layout, components
// interleaved.  Block: one 32 (x) x 32 (r) tile of the plane (sd, y).
template <typename Tin, typename Tout, int NC>
__global__ __launch_bounds__(256) void xdmf_pack_kernel(const Comp3<const Tin> src, Tout *dst, int nx, int ny, int nr)
{
    __shared__ Tin tile[NC][kTile][kTile + 1];
    const int tr = blockIdx.x % ((nr + kTile - 1) / kTile), tx = blockIdx.x / ((nr + kTile - 1) / kTile);
    const int y = blockIdx.y, sd = blockIdx.z;
    const int x0 = tx * kTile, r0 = tr * kTile;
    const int t = threadIdx.x;
    // load: 32 consecutive r per row x (8 rows per pass)
    for (int i = t; i < kTile * kTile; i += blockDim.x) {
        const int xl = i / kTile, rl = i % kTile, x = x0 + xl, r = r0 + rl;
        if (x < nx && r < nr) {
            const int64_t a = (((int64_t)sd * nx + x) * ny + y) * nr + r;
#pragma unroll
            for (int c = 0; c < NC; ++c) tile[c][xl][rl] = src.p[c][a];
        }
    }
    __syncthreads();
    // store: per row r, 32 x * NC consecutive values
    const int nxt = min(kTile, nx - x0);
    for (int i = t; i < kTile * kTile * NC; i += blockDim.x) {
        const int rl = i / (kTile * NC), k = i % (kTile * NC), xl = k / NC, c = k % NC, r = r0 + rl;
        if (xl < nxt && r < nr) {
            const int64_t b = ((((int64_t)sd * nr + r) * ny + y) * nx + x0 + xl) * NC + c;
            dst[b] = static_cast<Tout>(tile[c][xl][rl]);
        }
    }
}

// Inverse: src in file layout (type Tin), dst in device layout (type Tout).
template <typename Tin, typename Tout, int NC>
__global__ __launch_bounds__(256) void xdmf_unpack_kernel(const Tin *src, const Comp3<Tout> dst, int nx, int ny, int nr)
{
    __shared__ Tin tile[NC][kTile][kTile + 1];
    const int tr = blockIdx.x % ((nr + kTile - 1) / kTile), tx = blockIdx.x / ((nr + kTile - 1) / kTile);
    const int y = blockIdx.y, sd = blockIdx.z;
    const int x0 = tx * kTile, r0 = tr * kTile;
    const int t = threadIdx.x;
    const int nxt = min(kTile, nx - x0);
    for (int i = t; i < kTile * kTile * NC; i += blockDim.x) {
        const int rl = i / (kTile * NC), k = i % (kTile * NC), xl = k / NC, c = k % NC, r = r0 + rl;
        if (xl < nxt && r < nr)
            tile[c][xl][rl] = src[((((int64_t)sd * nr + r) * ny + y) * nx + x0 + xl) * NC + c];
    }
    __syncthreads();
    for (int i = t; i < kTile * kTile; i += blockDim.x) {
        const int xl = i / kTile, rl = i % kTile, x = x0 + xl, r = r0 + rl;
        if (x < nx && r < nr) {
            const int64_t a = (((int64_t)sd * nx + x) * ny + y) * nr + r;
#pragma unroll
            for (int c = 0; c < NC; ++c) dst.p[c][a] = static_cast<Tout>(tile[c][xl][rl]);
        }
    }
}

// Node coordinates in the file order: coords(sd, x, y) * radius(sd, r),
// computed in double and converted once (grid/shell/spherical_shell.hpp:2917-2940).
template <typename Tout>
__global__ __launch_bounds__(256) void xdmf_geometry_kernel(const double *coords, const double *radii, Tout *dst,
                                                            int nsd, int nx, int nr)
{
    const int64_t n = (int64_t)nsd * nr * nx * nx;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nx);
    int64_t q = i / nx;
    const int y = (int)(q % nx);
    q /= nx;
    const int r = (int)(q % nr);
    const int sd = (int)(q / nr);
    const double rad = radii[(int64_t)sd * nr + r];
    const double *c = coords + 3 * (((int64_t)sd * nx + x) * nx + y);
#pragma unroll
    for (int d = 0; d < 3; ++d) dst[3 * i + d] = static_cast<Tout>(c[d] * rad);
}

// Wedge connectivity, 12 indices per hex cell in the file order.
template <typename Tint>
__global__ __launch_bounds__(256) void xdmf_topology_kernel(Tint *dst, int nsd, int nx, int nr)
{
    const int nc = nx - 1, ncr = nr - 1;
    const int64_t n = (int64_t)nsd * ncr * nc * nc;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nc);
    int64_t q = i / nc;
    const int y = (int)(q % nc);
    q /= nc;
    const int r = (int)(q % ncr);
    const int sd = (int)(q / ncr);
    const Tint s0 = (Tint)nx * nx * nr, s1 = (Tint)nx * nx, s2 = (Tint)nx;
    const Tint v0 = (Tint)sd * s0 + (Tint)r * s1 + (Tint)y * s2 + x, v1 = v0 + 1, v2 = v0 + nx, v3 = v0 + nx + 1;
    const Tint v4 = v0 + s1, v5 = v4 + 1, v6 = v4 + nx, v7 = v4 + nx + 1;
    Tint *o = dst + 12 * i;
    o[0] = v0; o[1] = v1; o[2] = v2; o[3] = v4; o[4] = v5; o[5] = v6;
    o[6] = v3; o[7] = v2; o[8] = v1; o[9] = v7; o[10] = v6; o[11] = v5;
}

dim3 tile_grid(const tng_domain &m) { return dim3((unsigned)(((m.nr + kTile - 1) / kTile) * ((m.nx + kTile - 1) / kTile)), (unsigned)m.nx, (unsigned)m.nsd); }

template <typename Tin, typename Tout>
void pack_typed(const tng_vector &v, void *dst, hipStream_t s)
{
    const tng_domain &m = *v.mesh;
    const Comp3<const Tin> src{{(const Tin *)v.comp(0), (const Tin *)v.comp(v.ncomp > 1 ? 1 : 0),
                                (const Tin *)v.comp(v.ncomp > 2 ? 2 : 0)}};
    if (v.ncomp == 3)
        hipLaunchKernelGGL((xdmf_pack_kernel<Tin, Tout, 3>), tile_grid(m), dim3(256), 0, s, src, (Tout *)dst, m.nx,
                           m.nx, m.nr);
    else
        hipLaunchKernelGGL((xdmf_pack_kernel<Tin, Tout, 1>), tile_grid(m), dim3(256), 0, s, src, (Tout *)dst, m.nx,
                           m.nx, m.nr);
    TNG_HIP_CHECK(hipGetLastError());
}

template <typename Tin, typename Tout> void unpack_typed(const void *src, tng_vector &v, hipStream_t s)
{
    const tng_domain &m = *v.mesh;
    const Comp3<Tout> dst{{(Tout *)v.comp(0), (Tout *)v.comp(v.ncomp > 1 ? 1 : 0), (Tout *)v.comp(v.ncomp > 2 ? 2 : 0)}};
    if (v.ncomp == 3)
        hipLaunchKernelGGL((xdmf_unpack_kernel<Tin, Tout, 3>), tile_grid(m), dim3(256), 0, s, (const Tin *)src, dst,
                           m.nx, m.nx, m.nr);
    else
        hipLaunchKernelGGL((xdmf_unpack_kernel<Tin, Tout, 1>), tile_grid(m), dim3(256), 0, s, (const Tin *)src, dst,
                           m.nx, m.nx, m.nr);
    TNG_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

[[noreturn]] void io_error(const std::string &msg) { throw Error(TNG_IO_ERROR, msg); }

// util/xml.hpp: attributes sorted by name, two spaces per level, content
// escaped, "<Name ... />" for an empty element.
struct Xml {
    std::string name, content;
    std::map<std::string, std::string> attrs;
    std::vector<Xml> children;
    static std::string esc(const std::string &d)
    {
        std::string o;
        for (char c : d) {
            switch (c) {
            case '&': o += "&amp;"; break;
            case '"': o += "&quot;"; break;
            case '\'': o += "&apos;"; break;
            case '<': o += "&lt;"; break;
            case '>': o += "&gt;"; break;
            default: o += c;
            }
        }
        return o;
    }
    std::string str(int indent = 0) const
    {
        const std::string ind(indent, ' ');
        std::string o = ind + "<" + name;
        for (const auto &a : attrs) o += " " + a.first + "=\"" + esc(a.second) + "\"";
        if (children.empty() && content.empty()) return o + " />\n";
        o += ">";
        if (!content.empty()) o += esc(content);
        if (!children.empty()) {
            o += "\n";
            for (const Xml &c : children) o += c.str(indent + 2);
            o += ind;
        }
        return o + "</" + name + ">\n";
    }
};

Xml data_item(const std::string &type, int precision, const std::string &dims, const std::string &file)
{
    Xml d{"DataItem", file, {}, {}};
    d.attrs = {{"Format", "Binary"}, {"DataType", type}, {"Precision", std::to_string(precision)},
               {"Endian", "Little"}, {"Dimensions", dims}};
    return d;
}

void write_file(const std::string &path, const void *data, size_t bytes)
{
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) io_error("cannot open " + path + " for writing: " + std::strerror(errno));
    const size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    const int rc = std::fclose(f);
    if (w != bytes || rc != 0) io_error("short write to " + path);
}

std::vector<char> read_file(const std::string &path)
{
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) io_error("cannot open " + path + ": " + std::strerror(errno));
    std::vector<char> b;
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + k);
    std::fclose(f);
    return b;
}

// util/filesystem.hpp:16-44: create the directory (and parents) if missing
void make_dirs(const std::string &path)
{
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        if (::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) io_error("cannot create directory " + cur);
    }
}

int64_t global_subdomain_id(const tng_domain &m, int sd)
{
    // grid/shell/spherical_shell.hpp:700-722: x bits 0-18, y 19-37, r 38-56, diamond 57-63
    const int *s = &m.sd_info[4 * sd];
    return ((int64_t)s[0] << 57) | ((int64_t)s[3] << 38) | ((int64_t)s[2] << 19) | (int64_t)s[1];
}

std::vector<double> global_radii(const tng_domain &m)
{
    // all radial subdomains in order; shared levels once
    std::vector<double> r((size_t)m.Sr * (m.nr - 1) + 1, 0.0);
    for (int sd = 0; sd < m.nsd; ++sd) {
        const int sr = m.sd_info[4 * sd + 3];
        for (int j = 0; j < m.nr; ++j) r[(size_t)sr * (m.nr - 1) + j] = m.radii[(size_t)sd * m.nr + j];
    }
    return r;
}

}  // namespace

void PinnedBuf::reserve(size_t n)
{
    if (n <= bytes) return;
    release();
    TNG_HIP_CHECK(hipHostMalloc(&p, n, hipHostMallocDefault));
    bytes = n;
}

void PinnedBuf::release()
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

void xdmf_pack(const tng_vector &v, void *dst_device, int out_bytes, hipStream_t s)
{
    if (v.dtype == TNG_F64) {
        if (out_bytes == 8) pack_typed<double, double>(v, dst_device, s);
        else pack_typed<double, float>(v, dst_device, s);
    } else {
        if (out_bytes == 8) pack_typed<float, double>(v, dst_device, s);
        else pack_typed<float, float>(v, dst_device, s);
    }
}

void xdmf_unpack(const void *src_device, int in_bytes, tng_vector &v, hipStream_t s)
{
    if (v.dtype == TNG_F64) {
        if (in_bytes == 8) unpack_typed<double, double>(src_device, v, s);
        else unpack_typed<float, double>(src_device, v, s);
    } else {
        if (in_bytes == 8) unpack_typed<double, float>(src_device, v, s);
        else unpack_typed<float, float>(src_device, v, s);
    }
}

}  // namespace tng

// io/xdmf.hpp:148-513 XDMFOutput

// scalar float, scalar double, vector float, vector double (io/xdmf.hpp:475-497, 954-1000)
std::vector<const tng_xdmf_writer::Item *> tng_xdmf_writer::ordered() const
{
    std::vector<const Item *> o;
    for (int g = 0; g < 4; ++g)
        for (const Item &it : items) {
            const int gi = (it.v->ncomp == 3 ? 2 : 0) + (it.v->dtype == TNG_F64 ? 1 : 0);
            if (gi == g) o.push_back(&it);
        }
    return o;
}

void tng_xdmf_writer::add(tng_vector *v, const std::string &label, int bytes)
{
    TNG_REQUIRE(!first_done, "XDMF: cannot add data after write() has been called");
    TNG_REQUIRE(v && v->mesh == mesh, "XDMF: the field must live on the writer's mesh");
    TNG_REQUIRE(bytes == 4 || bytes == 8, "XDMF: output precision must be 4 or 8 bytes");
    TNG_REQUIRE(!label.empty(), "XDMF: empty label");
    for (const Item &it : items) TNG_REQUIRE(it.label != label, "XDMF: label '" + label + "' is taken");
    items.push_back({v, label, bytes});
}

// device -> pinned host -> file
void tng_xdmf_writer::emit(const std::string &path, size_t bytes, hipStream_t s)
{
    host.reserve(bytes);
    TNG_HIP_CHECK(hipMemcpyAsync(host.p, dev.p, bytes, hipMemcpyDeviceToHost, s));
    TNG_HIP_CHECK(hipStreamSynchronize(s));
    tng::write_file(path, host.p, bytes);
}

void tng_xdmf_writer::write_metadata()
{
    std::string b;
    auto i32 = [&](int32_t v) { b.append((const char *)&v, 4); };
    auto i64 = [&](int64_t v) { b.append((const char *)&v, 8); };
    auto f64 = [&](double v) { b.append((const char *)&v, 8); };
    const tng_domain &m = *mesh;
    i32(1);  // version
    i32(m.S);
    i32(m.Sr);
    i32(m.nx);
    i32(m.nx);
    i32(m.nr);
    for (double r : tng::global_radii(m)) f64(r);
    i32(points_bytes);
    const auto ord = ordered();
    i32((int32_t)ord.size());
    for (const Item *it : ord) {
        i32((int32_t)it->label.size());
        b.append(it->label);
        i32(2);  // floating point
        i32(it->bytes);
        i32(it->v->ncomp);
    }
    i32(0);  // checkpoint_subdomain_ordering: ids in the order of the data
    for (int sd = 0; sd < m.nsd; ++sd) i64(tng::global_subdomain_id(m, sd));
    tng::write_file(dir + "/checkpoint_metadata.bin", b.data(), b.size());
}

void tng_xdmf_writer::write()
{
    tng_domain &m = *mesh;
    hipStream_t s = m.ctx->stream;
    const int64_t nodes = m.nalloc, elements = (int64_t)m.nsd * (m.nx - 1) * (m.nx - 1) * (m.nr - 1) * 2;
    const auto t0 = std::chrono::steady_clock::now();
    double pack = 0.0;
    auto timed_sync = [&](hipEvent_t a, hipEvent_t z) {
        TNG_HIP_CHECK(hipEventSynchronize(z));
        float ms = 0.f;
        TNG_HIP_CHECK(hipEventElapsedTime(&ms, a, z));
        pack += ms * 1e-3;
    };
    struct Events {  // released on every exit path
        hipEvent_t a = nullptr, z = nullptr;
        ~Events()
        {
            if (a) (void)hipEventDestroy(a);
            if (z) (void)hipEventDestroy(z);
        }
    } ev;
    TNG_HIP_CHECK(hipEventCreate(&ev.a));
    TNG_HIP_CHECK(hipEventCreate(&ev.z));
    hipEvent_t e0 = ev.a, e1 = ev.z;
    if (!first_done) {
        tng::make_dirs(dir);
        const std::string readme =
            "# This directory contains the output of the XDMF writer (for visualization and checkpointing).\n\n"
            "Overview:\n"
            "- `geometry.bin`: cartesian node coordinates\n"
            "- `topology.bin`: connectivity/topology/elements (whatever you want to call it)\n"
            "- `checkpoint_metadata.bin`: metadata for checkpointing\n"
            "- `step_*.xmf`: xmf file (open this with Paraview for visualization) for each write() step\n"
            "- `<some_name>_*.bin`: binary grid data (per write() step)\n";
        tng::write_file(dir + "/README.md", readme.data(), readme.size());
        // geometry
        const size_t gbytes = (size_t)nodes * 3 * points_bytes;
        if (dev.bytes < gbytes) dev.alloc(gbytes);
        const unsigned gb = (unsigned)((nodes + 255) / 256);
        TNG_HIP_CHECK(hipEventRecord(e0, s));
        if (points_bytes == 8)
            hipLaunchKernelGGL(tng::xdmf_geometry_kernel<double>, dim3(gb), dim3(256), 0, s,
                               m.d_coords.as<double>(), m.d_radii.as<double>(), dev.as<double>(), m.nsd, m.nx, m.nr);
        else
            hipLaunchKernelGGL(tng::xdmf_geometry_kernel<float>, dim3(gb), dim3(256), 0, s,
                               m.d_coords.as<double>(), m.d_radii.as<double>(), dev.as<float>(), m.nsd, m.nx, m.nr);
        TNG_HIP_CHECK(hipGetLastError());
        TNG_HIP_CHECK(hipEventRecord(e1, s));
        timed_sync(e0, e1);
        emit(dir + "/geometry.bin", gbytes, s);
        // topology
        const int64_t cells = elements / 2;
        TNG_REQUIRE(conn_bytes == 8 || nodes < ((int64_t)1 << 31), "XDMF: 32-bit connectivity overflows");
        const size_t tbytes = (size_t)elements * 6 * conn_bytes;
        if (dev.bytes < tbytes) dev.alloc(tbytes);
        const unsigned tb = (unsigned)((cells + 255) / 256);
        TNG_HIP_CHECK(hipEventRecord(e0, s));
        if (conn_bytes == 8)
            hipLaunchKernelGGL(tng::xdmf_topology_kernel<int64_t>, dim3(tb), dim3(256), 0, s, dev.as<int64_t>(),
                               m.nsd, m.nx, m.nr);
        else
            hipLaunchKernelGGL(tng::xdmf_topology_kernel<int32_t>, dim3(tb), dim3(256), 0, s, dev.as<int32_t>(),
                               m.nsd, m.nx, m.nr);
        TNG_HIP_CHECK(hipGetLastError());
        TNG_HIP_CHECK(hipEventRecord(e1, s));
        timed_sync(e0, e1);
        emit(dir + "/topology.bin", tbytes, s);
        write_metadata();
    }
    tng::Xml grid{"Grid", "", {{"Name", "Grid"}, {"GridType", "Uniform"}}, {}};
    tng::Xml geo{"Geometry", "", {{"Type", "XYZ"}}, {}};
    geo.children.push_back(
        tng::data_item("Float", points_bytes, std::to_string(nodes) + " 3", "geometry.bin"));
    grid.children.push_back(geo);
    tng::Xml topo{"Topology", "", {{"Type", "Wedge"}, {"NumberOfElements", std::to_string(elements)}}, {}};
    topo.children.push_back(tng::data_item("Int", conn_bytes, std::to_string(elements * 6), "topology.bin"));
    grid.children.push_back(topo);
    for (const Item *it : ordered()) {
        const std::string file = it->label + "_" + std::to_string(counter) + ".bin";
        const size_t bytes = (size_t)nodes * it->v->ncomp * it->bytes;
        if (dev.bytes < bytes) dev.alloc(bytes);
        TNG_HIP_CHECK(hipEventRecord(e0, s));
        tng::xdmf_pack(*it->v, dev.p, it->bytes, s);
        TNG_HIP_CHECK(hipEventRecord(e1, s));
        timed_sync(e0, e1);
        emit(dir + "/" + file, bytes, s);
        tng::Xml at{"Attribute", "",
               {{"Name", it->label}, {"AttributeType", it->v->ncomp == 3 ? "Vector" : "Scalar"}, {"Center", "Node"}},
               {}};
        at.children.push_back(tng::data_item(
            "Float", it->bytes, it->v->ncomp == 3 ? std::to_string(nodes) + " 3" : std::to_string(nodes), file));
        grid.children.push_back(at);
    }
    tng::Xml domain{"Domain", "", {}, {grid}};
    tng::Xml xdmf{"Xdmf", "", {{"Version", "2.0"}}, {domain}};
    const std::string text =
        "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n" + xdmf.str();
    tng::write_file(dir + "/step_" + std::to_string(counter) + ".xmf", text.data(), text.size());
    last_pack_seconds = pack;
    last_write_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++counter;
    first_done = true;
}

namespace tng {

// io/xdmf.hpp:1094-1186 read_xdmf_checkpoint_metadata + :1199-1441
// read_xdmf_checkpoint_grid, for one GPU holding every subdomain.
void xdmf_checkpoint_read(const std::string &dir, const std::string &label, int step, tng_vector &v)
{
    const std::vector<char> meta = read_file(dir + "/checkpoint_metadata.bin");
    size_t pos = 0;
    auto need = [&](size_t n) {
        if (n > meta.size() - pos) io_error("checkpoint metadata: unexpected end of file");
    };
    auto i32 = [&]() { need(4); int32_t x; std::memcpy(&x, meta.data() + pos, 4); pos += 4; return x; };
    auto i64 = [&]() { need(8); int64_t x; std::memcpy(&x, meta.data() + pos, 8); pos += 8; return x; };
    auto f64 = [&]() { need(8); double x; std::memcpy(&x, meta.data() + pos, 8); pos += 8; return x; };
    const int32_t version = i32();
    if (version != 0 && version != 1)
        io_error("supported checkpoint versions: 0, 1; this checkpoint has version " + std::to_string(version));
    const int32_t S = i32(), Sr = i32(), sx = i32(), sy = i32(), sr = i32();
    if (S <= 0 || Sr <= 0 || sr < 2 || sx < 2 || sy < 2)
        io_error("checkpoint metadata: invalid subdomain counts or extents");
    if (S != v.mesh->S || Sr != v.mesh->Sr)
        io_error("checkpoint metadata: the subdomain decomposition does not match the mesh");
    if ((int64_t)Sr * (sr - 1) + 1 > (int64_t)(meta.size() - pos) / 8)
        io_error("checkpoint metadata: unexpected end of file");
    double prev = 0.0;
    for (int i = 0; i < Sr * (sr - 1) + 1; ++i) {
        const double r = f64();
        if (i > 0 && r <= prev) io_error("checkpoint radii are not sorted");
        prev = r;
    }
    if (version > 0) (void)i32();  // grid_scalar_bytes
    const int32_t nfiles = i32();
    int32_t dtype = -1, bytes = 0, vec_dim = 0;
    for (int f = 0; f < nfiles; ++f) {
        const int32_t len = i32();
        if (len < 0) io_error("checkpoint metadata: negative name length");
        need((size_t)len);
        const std::string name(meta.data() + pos, (size_t)len);
        pos += (size_t)len;
        const int32_t dt = i32(), by = i32(), vd = i32();
        if (name == label && dtype < 0) {
            dtype = dt;
            bytes = by;
            vec_dim = vd;
        }
    }
    if (dtype < 0) io_error("could not find requested data (" + label + ") in checkpoint");
    const tng_domain &m = *v.mesh;
    if (sx != m.nx || sy != m.nx || sr != m.nr) io_error("grid data extents do not match metadata");
    if (i32() != 0) io_error("checkpoint ordering type is not 0; not supported");
    const int64_t nglobal = (int64_t)10 * S * S * Sr;
    if (nglobal < m.nsd) io_error("checkpoint metadata: fewer subdomains than the mesh has");
    need((size_t)nglobal * 8);
    std::vector<int64_t> ids((size_t)nglobal);
    for (auto &id : ids) id = i64();
    if (dtype != 2 || (bytes != 4 && bytes != 8)) io_error("unsupported data type in checkpoint");
    if (vec_dim != v.ncomp)
        io_error(vec_dim == 1 ? "checkpoint is scalar, the vector's component count does not match"
                              : "checkpoint is vector-valued, the vector's component count does not match");
    const std::string path = dir + "/" + label + "_" + std::to_string(step) + ".bin";
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) io_error("could not find checkpoint data file for requested label and step: " + path);
    const size_t chunk = (size_t)m.nx * m.nx * m.nr * vec_dim * bytes;
    PinnedBuf host;
    host.reserve(chunk * m.nsd);
    for (int sd = 0; sd < m.nsd; ++sd) {
        const int64_t gid = global_subdomain_id(m, sd);
        size_t idx = 0;
        while (idx < ids.size() && ids[idx] != gid) ++idx;
        if (idx == ids.size()) {
            std::fclose(f);
            io_error("subdomain missing from checkpoint");
        }
        if (std::fseek(f, (long)(idx * chunk), SEEK_SET) != 0 ||
            std::fread((char *)host.p + (size_t)sd * chunk, 1, chunk, f) != chunk) {
            std::fclose(f);
            io_error("short read from " + path);
        }
    }
    std::fclose(f);
    hipStream_t s = m.ctx->stream;
    DevBuf dev(chunk * m.nsd);
    TNG_HIP_CHECK(hipMemcpyAsync(dev.p, host.p, chunk * m.nsd, hipMemcpyHostToDevice, s));
    xdmf_unpack(dev.p, bytes, v, s);
    TNG_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace tng
