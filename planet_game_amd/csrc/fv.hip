// Finite-volume neighbour sweeps on the Voronoi mesh for gfx950: the
// least-squares gradient, edge-flux divergence and two-point Laplacian of
// sim/planet/operators/finite_volume.cpp:34-111, and the diffusion source
// of sim/planet/surface/heat_transport.cpp:28-36,579-593.
//
// One thread per cell, cells in mesh order (the space-filling order keeps
// neighbour gathers inside a few KB of the cell's own row, so they hit L2),
// per-entry geometry stored in CSR entry order so a wavefront's loads of
// slot k are contiguous. Each cell accumulates in double, in the mesh's
// cell-edge order, with the reference's operation order and no FMA
// contraction (-ffp-contract=off), so results are bit-identical to the
// reference for float and double fields. grid.y walks the fields of a
// multi-field call; the geometry of a cell is re-read per field from L2.
#include "device.hpp"
#include "internal.hpp"

#include <cstring>
#include <memory>
#include <vector>

namespace psim::dev {
namespace {

template <class T>
__global__ __launch_bounds__(kTile) void gradient_kernel(int n, const int32_t* __restrict__ row,
                                                         const int32_t* __restrict__ nb, const float* __restrict__ ne,
                                                         const float* __restrict__ nn, const double* __restrict__ dist,
                                                         const T* __restrict__ phi, T* __restrict__ east,
                                                         T* __restrict__ north) {
    const int i = blockIdx.x * kTile + threadIdx.x;
    if (i >= n) return;
    const size_t base = static_cast<size_t>(blockIdx.y) * n;
    const T* f = phi + base;
    const double own = static_cast<double>(f[i]);
    double mee = 0.0, men = 0.0, mnn = 0.0, re = 0.0, rn = 0.0;
    for (int k = row[i]; k < row[i + 1]; ++k) {
        const double e = static_cast<double>(ne[k]);
        const double q = static_cast<double>(nn[k]);
        const double slope = (static_cast<double>(f[nb[k]]) - own) / dist[k];
        mee += e * e;
        men += e * q;
        mnn += q * q;
        re += slope * e;
        rn += slope * q;
    }
    const double det = mee * mnn - men * men;
    east[base + i] = static_cast<T>((mnn * re - men * rn) / det);
    north[base + i] = static_cast<T>((mee * rn - men * re) / det);
}

template <class T>
__global__ __launch_bounds__(kTile) void divergence_kernel(int n, int edges, const int32_t* __restrict__ row,
                                                           const int32_t* __restrict__ edge,
                                                           const double* __restrict__ signed_length,
                                                           const double* __restrict__ area,
                                                           const T* __restrict__ flux, T* __restrict__ out) {
    const int i = blockIdx.x * kTile + threadIdx.x;
    if (i >= n) return;
    const T* f = flux + static_cast<size_t>(blockIdx.y) * edges;
    double net = 0.0;
    // orientation (+-1) x flux x length == flux x (+-length): a sign flip
    // commutes with round-to-nearest multiplication.
    for (int k = row[i]; k < row[i + 1]; ++k) net += static_cast<double>(f[edge[k]]) * signed_length[k];
    out[static_cast<size_t>(blockIdx.y) * n + i] = static_cast<T>(net / area[i]);
}

template <class T>
__global__ __launch_bounds__(kTile) void laplacian_kernel(int n, const int32_t* __restrict__ row,
                                                          const int32_t* __restrict__ nb,
                                                          const double* __restrict__ length,
                                                          const double* __restrict__ dist,
                                                          const double* __restrict__ area, const T* __restrict__ phi,
                                                          T* __restrict__ out) {
    const int i = blockIdx.x * kTile + threadIdx.x;
    if (i >= n) return;
    const size_t base = static_cast<size_t>(blockIdx.y) * n;
    const T* f = phi + base;
    const double own = static_cast<double>(f[i]);
    double sum = 0.0;
    for (int k = row[i]; k < row[i + 1]; ++k) sum += (static_cast<double>(f[nb[k]]) - own) * length[k] / dist[k];
    out[base + i] = static_cast<T>(sum / area[i]);
}

__global__ __launch_bounds__(kTile) void diffusion_kernel(int n, const int32_t* __restrict__ row,
                                                          const int32_t* __restrict__ col,
                                                          const double* __restrict__ weight,
                                                          const double* __restrict__ area, double conductance,
                                                          const double* __restrict__ t, double* __restrict__ out) {
    const int i = blockIdx.x * kTile + threadIdx.x;
    if (i >= n) return;
    const double own = t[i];
    double sum = 0.0;
    for (int k = row[i]; k < row[i + 1]; ++k) sum += (t[col[k]] - own) * weight[k];
    out[i] = conductance * sum / area[i];
}

}  // namespace

void launch_diffusion(int n, const int32_t* row, const int32_t* col, const double* weight, const double* area,
                      double conductance, const double* t, double* out, hipStream_t stream) {
    hipLaunchKernelGGL(diffusion_kernel, dim3((n + kTile - 1) / kTile), dim3(kTile), 0, stream, n, row, col, weight,
                       area, conductance, t, out);
    hip_check(hipGetLastError(), "diffusion_source launch");
}

}  // namespace psim::dev

// ------------------------------------------------------------------ C ABI

struct psim_fv {
    psim_ctx* ctx = nullptr;
    int32_t n = 0, edges = 0, directed = 0;
    std::vector<uint32_t> first, second;   // host copies for edge_orientation
    void* block = nullptr;                 // one device allocation for all arrays
    int32_t* row = nullptr;
    int32_t* nb = nullptr;
    int32_t* edge = nullptr;
    float* ne = nullptr;
    float* nn = nullptr;
    double* dist = nullptr;
    double* length = nullptr;
    double* signed_length = nullptr;
    double* area = nullptr;
};

namespace {

using psim::guarded;

void validate_ctx(const psim_ctx* ctx) {
    if (ctx == nullptr) throw std::invalid_argument("null context");
}

template <class T> T* carve(char*& cursor, size_t count) {
    T* p = reinterpret_cast<T*>(cursor);
    cursor += (count * sizeof(T) + 255) / 256 * 256;
    return p;
}

psim_fv* upload_fv(psim_ctx* ctx, uint64_t cells, uint64_t edges, const double* area, const uint32_t* edge_offset,
                   const uint8_t* edge_count, const uint32_t* neighbor, const uint32_t* edge, const float* ne,
                   const float* nn, const uint32_t* first, const uint32_t* second, const double* length,
                   const double* distance) {
    validate_ctx(ctx);
    if (cells == 0 || cells >= (1ULL << 30) || edges >= (1ULL << 30)) throw std::invalid_argument("mesh size out of range");
    auto fv = std::make_unique<psim_fv>();
    fv->ctx = ctx;
    fv->n = static_cast<int32_t>(cells);
    fv->edges = static_cast<int32_t>(edges);
    std::vector<int32_t> row(cells + 1);
    row[0] = 0;
    for (uint64_t c = 0; c < cells; ++c) {
        if (edge_offset[c] != static_cast<uint32_t>(row[c])) throw std::invalid_argument("cell edges are not contiguous");
        row[c + 1] = row[c] + edge_count[c];
    }
    const uint64_t d = static_cast<uint64_t>(row[cells]);
    fv->directed = static_cast<int32_t>(d);
    std::vector<int32_t> nb(d), ed(d);
    std::vector<double> dist(d), len(d), slen(d);
    for (uint64_t c = 0; c < cells; ++c)
        for (int32_t k = row[c]; k < row[c + 1]; ++k) {
            if (neighbor[k] >= cells || edge[k] >= edges) throw std::invalid_argument("mesh index out of range");
            nb[k] = static_cast<int32_t>(neighbor[k]);
            ed[k] = static_cast<int32_t>(edge[k]);
            dist[k] = distance[edge[k]];
            len[k] = length[edge[k]];
            // edge_orientation: +1 from the first cell, -1 from the second
            if (first[edge[k]] == c)
                slen[k] = length[edge[k]];
            else if (second[edge[k]] == c)
                slen[k] = -length[edge[k]];
            else
                throw std::invalid_argument("cell is not incident to edge");
        }
    fv->first.assign(first, first + edges);
    fv->second.assign(second, second + edges);
    const size_t bytes = 256 * 16 + (cells + 1) * 4 + 2 * d * 4 + 2 * d * 4 + 3 * d * 8 + cells * 8;
    psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    psim::hip_check(hipMalloc(&fv->block, bytes), "mesh upload allocation");
    char* cur = static_cast<char*>(fv->block);
    fv->row = carve<int32_t>(cur, cells + 1);
    fv->nb = carve<int32_t>(cur, d);
    fv->edge = carve<int32_t>(cur, d);
    fv->ne = carve<float>(cur, d);
    fv->nn = carve<float>(cur, d);
    fv->dist = carve<double>(cur, d);
    fv->length = carve<double>(cur, d);
    fv->signed_length = carve<double>(cur, d);
    fv->area = carve<double>(cur, cells);
    const auto put = [&](void* dst, const void* src, size_t b) {
        psim::hip_check(hipMemcpy(dst, src, b, hipMemcpyHostToDevice), "mesh upload copy");
    };
    put(fv->row, row.data(), (cells + 1) * 4);
    put(fv->nb, nb.data(), d * 4);
    put(fv->edge, ed.data(), d * 4);
    put(fv->ne, ne, d * 4);
    put(fv->nn, nn, d * 4);
    put(fv->dist, dist.data(), d * 8);
    put(fv->length, len.data(), d * 8);
    put(fv->signed_length, slen.data(), d * 8);
    put(fv->area, area, cells * 8);
    return fv.release();
}

void check_fields(const psim_fv* fv, int dtype, uint64_t fields) {
    if (fv == nullptr) throw std::invalid_argument("null mesh");
    if (dtype != PSIM_F64 && dtype != PSIM_F32) throw std::invalid_argument("unknown dtype");
    if (fields == 0 || fields > 65535) throw std::invalid_argument("field count out of range");
}

dim3 grid_of(int n, uint64_t fields) { return dim3((n + psim::dev::kTile - 1) / psim::dev::kTile, static_cast<unsigned>(fields)); }

}  // namespace

extern "C" {

int psim_fv_create(psim_ctx* ctx, const psim_mesh* mesh, psim_fv** out) {
    return guarded([&] {
        if (mesh == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        const psim::HostMesh& m = psim::host_mesh_of(mesh);
        *out = upload_fv(ctx, m.cells, m.edges, m.area.data(), m.edge_offset.data(), m.edge_count.data(),
                         m.neighbor.data(), m.edge.data(), m.normal_east.data(), m.normal_north.data(), m.first.data(),
                         m.second.data(), m.length.data(), m.distance.data());
    });
}

int psim_fv_create_from_arrays(psim_ctx* ctx, uint64_t cells, uint64_t edges, const double* area,
                               const uint32_t* edge_offset, const uint8_t* edge_count, const uint32_t* neighbor,
                               const uint32_t* edge, const float* ne, const float* nn, const uint32_t* first,
                               const uint32_t* second, const double* length, const double* distance, psim_fv** out) {
    return guarded([&] {
        if (!area || !edge_offset || !edge_count || !neighbor || !edge || !ne || !nn || !first || !second || !length ||
            !distance || !out)
            throw std::invalid_argument("null argument");
        *out = upload_fv(ctx, cells, edges, area, edge_offset, edge_count, neighbor, edge, ne, nn, first, second, length,
                         distance);
    });
}

void psim_fv_destroy(psim_fv* fv) {
    if (fv == nullptr) return;
    if (fv->block != nullptr) (void)hipFree(fv->block);
    delete fv;
}

int psim_edge_orientation(const psim_fv* fv, uint32_t cell, uint32_t edge, double* out) {
    return guarded([&] {
        if (fv == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        if (edge >= static_cast<uint32_t>(fv->edges)) throw std::out_of_range("edge id out of range");
        if (fv->first[edge] == cell)
            *out = 1.0;
        else if (fv->second[edge] == cell)
            *out = -1.0;
        else
            throw std::invalid_argument("cell is not incident to edge");
    });
}

int psim_gradient(psim_fv* fv, int dtype, uint64_t fields, const void* scalar, void* east, void* north, int where) {
    return guarded([&] {
        check_fields(fv, dtype, fields);
        const size_t es = dtype == PSIM_F64 ? 8 : 4, bytes = es * fields * static_cast<size_t>(fv->n);
        psim::Staging st;
        const void* s = where == PSIM_HOST ? st.in(scalar, bytes) : scalar;
        void* e = where == PSIM_HOST ? st.out(bytes) : east;
        void* no = where == PSIM_HOST ? st.out(bytes) : north;
        hipStream_t stream = fv->ctx->stream;
        if (dtype == PSIM_F64)
            hipLaunchKernelGGL(psim::dev::gradient_kernel<double>, grid_of(fv->n, fields), dim3(psim::dev::kTile), 0,
                               stream, fv->n, fv->row, fv->nb, fv->ne, fv->nn, fv->dist, static_cast<const double*>(s),
                               static_cast<double*>(e), static_cast<double*>(no));
        else
            hipLaunchKernelGGL(psim::dev::gradient_kernel<float>, grid_of(fv->n, fields), dim3(psim::dev::kTile), 0,
                               stream, fv->n, fv->row, fv->nb, fv->ne, fv->nn, fv->dist, static_cast<const float*>(s),
                               static_cast<float*>(e), static_cast<float*>(no));
        psim::hip_check(hipGetLastError(), "gradient launch");
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(east, e, bytes, hipMemcpyDeviceToHost, stream), "gradient copy");
            psim::hip_check(hipMemcpyAsync(north, no, bytes, hipMemcpyDeviceToHost, stream), "gradient copy");
            psim::hip_check(hipStreamSynchronize(stream), "gradient");
        }
    });
}

int psim_divergence(psim_fv* fv, int dtype, uint64_t fields, const void* flux, void* result, int where) {
    return guarded([&] {
        check_fields(fv, dtype, fields);
        const size_t es = dtype == PSIM_F64 ? 8 : 4;
        const size_t in_bytes = es * fields * static_cast<size_t>(fv->edges);
        const size_t out_bytes = es * fields * static_cast<size_t>(fv->n);
        psim::Staging st;
        const void* f = where == PSIM_HOST ? st.in(flux, in_bytes) : flux;
        void* o = where == PSIM_HOST ? st.out(out_bytes) : result;
        hipStream_t stream = fv->ctx->stream;
        if (dtype == PSIM_F64)
            hipLaunchKernelGGL(psim::dev::divergence_kernel<double>, grid_of(fv->n, fields), dim3(psim::dev::kTile), 0,
                               stream, fv->n, fv->edges, fv->row, fv->edge, fv->signed_length, fv->area,
                               static_cast<const double*>(f), static_cast<double*>(o));
        else
            hipLaunchKernelGGL(psim::dev::divergence_kernel<float>, grid_of(fv->n, fields), dim3(psim::dev::kTile), 0,
                               stream, fv->n, fv->edges, fv->row, fv->edge, fv->signed_length, fv->area,
                               static_cast<const float*>(f), static_cast<float*>(o));
        psim::hip_check(hipGetLastError(), "divergence launch");
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(result, o, out_bytes, hipMemcpyDeviceToHost, stream), "divergence copy");
            psim::hip_check(hipStreamSynchronize(stream), "divergence");
        }
    });
}

int psim_laplacian(psim_fv* fv, int dtype, uint64_t fields, const void* scalar, void* result, int where) {
    return guarded([&] {
        check_fields(fv, dtype, fields);
        const size_t es = dtype == PSIM_F64 ? 8 : 4, bytes = es * fields * static_cast<size_t>(fv->n);
        psim::Staging st;
        const void* s = where == PSIM_HOST ? st.in(scalar, bytes) : scalar;
        void* o = where == PSIM_HOST ? st.out(bytes) : result;
        hipStream_t stream = fv->ctx->stream;
        if (dtype == PSIM_F64)
            hipLaunchKernelGGL(psim::dev::laplacian_kernel<double>, grid_of(fv->n, fields), dim3(psim::dev::kTile), 0,
                               stream, fv->n, fv->row, fv->nb, fv->length, fv->dist, fv->area,
                               static_cast<const double*>(s), static_cast<double*>(o));
        else
            hipLaunchKernelGGL(psim::dev::laplacian_kernel<float>, grid_of(fv->n, fields), dim3(psim::dev::kTile), 0,
                               stream, fv->n, fv->row, fv->nb, fv->length, fv->dist, fv->area,
                               static_cast<const float*>(s), static_cast<float*>(o));
        psim::hip_check(hipGetLastError(), "laplacian launch");
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(result, o, bytes, hipMemcpyDeviceToHost, stream), "laplacian copy");
            psim::hip_check(hipStreamSynchronize(stream), "laplacian");
        }
    });
}

}  // extern "C"
