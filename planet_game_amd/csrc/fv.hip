// Finite-volume neighbour sweeps on the Voronoi mesh for gfx950: the
// least-squares gradient, edge-flux divergence and two-point Laplacian of
// sim/planet/operators/finite_volume.cpp:34-111, and the diffusion source
// of sim/planet/surface/heat_transport.cpp:28-36,579-593.
//
// Layout: the per-entry geometry is slot-major (ELL): slot s of cell i at
// s * n + i, so a wavefront's loads of one slot are consecutive; slots past
// a cell's degree hold neighbour -1 and are skipped. One thread per cell
// and per group of kFieldGroup fields: the thread loads its cell's geometry
// once, then runs the group's fields with all their loads in flight, each
// field accumulated in double in the mesh's cell-edge order with the
// reference's operation order and no FMA contraction (-ffp-contract=off),
// so results are bit-identical to the reference for float and double.
//
// Workgroup order is XCD-aware: consecutive workgroups go to the eight XCDs
// in turn, so workgroup b runs on XCD b % 8. Each XCD is given one
// contiguous eighth of the tiles, and its workgroups walk those tiles in
// order with all field groups of a tile in a row: the field groups re-read
// the tile's geometry from that XCD's L2, and the neighbour gathers, which
// reach a few hundred cells either side in the mesh's space-filling order,
// find the neighbouring tiles' values in the same L2 instead of each XCD
// fetching them again from HBM.
#include "device.hpp"
#include "internal.hpp"

#include <cstring>
#include <memory>
#include <vector>

namespace psim::dev {
namespace {

constexpr int kFieldGroup = 4;   // fields per thread
constexpr int kXcds = 8;

// Tile and field group of this workgroup (grid: 8 x ceil(tiles / 8) x
// groups workgroups; XCD x takes tiles [x * per_xcd, (x + 1) * per_xcd));
// false for the padding past the last tile.
__device__ inline bool tile_and_group(int tiles, int groups, int& tile, int& group) {
    const int b = static_cast<int>(blockIdx.x);
    const int per_xcd = (tiles + kXcds - 1) / kXcds;
    const int x = b % kXcds, m = b / kXcds;
    group = m % groups;
    tile = x * per_xcd + m / groups;
    return m / groups < per_xcd && tile < tiles;
}

template <class T, int W>
__global__ __launch_bounds__(kTile) void gradient_kernel(int n, int fields, const int32_t* __restrict__ nb,
                                                         const float* __restrict__ ne, const float* __restrict__ nn,
                                                         const double* __restrict__ dist, const T* __restrict__ phi,
                                                         T* __restrict__ east, T* __restrict__ north) {
    int tile, group;
    const int groups = (fields + kFieldGroup - 1) / kFieldGroup;
    if (!tile_and_group((n + kTile - 1) / kTile, groups, tile, group)) return;
    const int i = tile * kTile + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    int c[W];
    double e[W], q[W], d[W];
#pragma unroll
    for (int s = 0; s < W; ++s) {
        c[s] = nb[s * n + i];
        e[s] = static_cast<double>(ne[s * n + i]);
        q[s] = static_cast<double>(nn[s * n + i]);
        d[s] = dist[s * n + i];
    }
    // M = sum n n^T in edge order, once for the group's fields
    double mee = 0.0, men = 0.0, mnn = 0.0;
#pragma unroll
    for (int s = 0; s < W; ++s)
        if (c[s] >= 0) {
            mee += e[s] * e[s];
            men += e[s] * q[s];
            mnn += q[s] * q[s];
        }
    const double det = mee * mnn - men * men;
    const int f0 = group * kFieldGroup;
    T own[kFieldGroup], v[kFieldGroup][W];
#pragma unroll
    for (int g = 0; g < kFieldGroup; ++g) {
        const T* f = phi + static_cast<size_t>(min(f0 + g, fields - 1)) * n;
        own[g] = f[i];
#pragma unroll
        for (int s = 0; s < W; ++s) v[g][s] = f[c[s] >= 0 ? c[s] : i];
    }
#pragma unroll
    for (int g = 0; g < kFieldGroup; ++g) {
        if (f0 + g >= fields) break;
        const double o = static_cast<double>(own[g]);
        double re = 0.0, rn = 0.0;
#pragma unroll
        for (int s = 0; s < W; ++s)
            if (c[s] >= 0) {
                const double slope = (static_cast<double>(v[g][s]) - o) / d[s];
                re += slope * e[s];
                rn += slope * q[s];
            }
        const size_t at = static_cast<size_t>(f0 + g) * n + i;
        east[at] = static_cast<T>((mnn * re - men * rn) / det);
        north[at] = static_cast<T>((mee * rn - men * re) / det);
    }
}

template <class T, int W>
__global__ __launch_bounds__(kTile) void divergence_kernel(int n, int edges, int fields,
                                                           const int32_t* __restrict__ edge,
                                                           const double* __restrict__ signed_length,
                                                           const double* __restrict__ area,
                                                           const T* __restrict__ flux, T* __restrict__ out) {
    int tile, group;
    const int groups = (fields + kFieldGroup - 1) / kFieldGroup;
    if (!tile_and_group((n + kTile - 1) / kTile, groups, tile, group)) return;
    const int i = tile * kTile + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    int k[W];
    double sl[W];
#pragma unroll
    for (int s = 0; s < W; ++s) {
        k[s] = edge[s * n + i];
        sl[s] = signed_length[s * n + i];
    }
    const double a = area[i];
    const int f0 = group * kFieldGroup;
    T v[kFieldGroup][W];
#pragma unroll
    for (int g = 0; g < kFieldGroup; ++g) {
        const T* f = flux + static_cast<size_t>(min(f0 + g, fields - 1)) * edges;
#pragma unroll
        for (int s = 0; s < W; ++s) v[g][s] = k[s] >= 0 ? f[k[s]] : T(0);
    }
#pragma unroll
    for (int g = 0; g < kFieldGroup; ++g) {
        if (f0 + g >= fields) break;
        // orientation (+-1) x flux x length == flux x (+-length): a sign flip
        // commutes with round-to-nearest multiplication
        double net = 0.0;
#pragma unroll
        for (int s = 0; s < W; ++s)
            if (k[s] >= 0) net += static_cast<double>(v[g][s]) * sl[s];
        out[static_cast<size_t>(f0 + g) * n + i] = static_cast<T>(net / a);
    }
}

template <class T, int W>
__global__ __launch_bounds__(kTile) void laplacian_kernel(int n, int fields, const int32_t* __restrict__ nb,
                                                          const double* __restrict__ length,
                                                          const double* __restrict__ dist,
                                                          const double* __restrict__ area, const T* __restrict__ phi,
                                                          T* __restrict__ out) {
    int tile, group;
    const int groups = (fields + kFieldGroup - 1) / kFieldGroup;
    if (!tile_and_group((n + kTile - 1) / kTile, groups, tile, group)) return;
    const int i = tile * kTile + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    int c[W];
    double l[W], d[W];
#pragma unroll
    for (int s = 0; s < W; ++s) {
        c[s] = nb[s * n + i];
        l[s] = length[s * n + i];
        d[s] = dist[s * n + i];
    }
    const double a = area[i];
    const int f0 = group * kFieldGroup;
    T own[kFieldGroup], v[kFieldGroup][W];
#pragma unroll
    for (int g = 0; g < kFieldGroup; ++g) {
        const T* f = phi + static_cast<size_t>(min(f0 + g, fields - 1)) * n;
        own[g] = f[i];
#pragma unroll
        for (int s = 0; s < W; ++s) v[g][s] = f[c[s] >= 0 ? c[s] : i];
    }
#pragma unroll
    for (int g = 0; g < kFieldGroup; ++g) {
        if (f0 + g >= fields) break;
        const double o = static_cast<double>(own[g]);
        double sum = 0.0;
#pragma unroll
        for (int s = 0; s < W; ++s)
            if (c[s] >= 0) sum += (static_cast<double>(v[g][s]) - o) * l[s] / d[s];
        out[static_cast<size_t>(f0 + g) * n + i] = static_cast<T>(sum / a);
    }
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
    int32_t width = 0;                     // slots per cell in the device layout (6, 8 or 16)
    std::vector<uint32_t> first, second;   // host copies for edge_orientation
    void* block = nullptr;                 // one device allocation for all arrays
    // slot-major: slot s of cell i at s * n + i; padding: nb = edge = -1
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
    uint64_t d = 0;
    int degree = 0;
    for (uint64_t c = 0; c < cells; ++c) {
        if (edge_offset[c] != static_cast<uint32_t>(d)) throw std::invalid_argument("cell edges are not contiguous");
        d += edge_count[c];
        degree = std::max<int>(degree, edge_count[c]);
    }
    if (degree > 16) throw std::invalid_argument("cells with more than 16 edges are not supported");
    fv->width = degree <= 6 ? 6 : degree <= 8 ? 8 : 16;
    fv->directed = static_cast<int32_t>(d);
    const size_t W = static_cast<size_t>(fv->width), slots = W * cells;
    std::vector<int32_t> nb(slots, -1), ed(slots, -1);
    std::vector<float> e_ne(slots, 0.0f), e_nn(slots, 0.0f);
    std::vector<double> dist(slots, 1.0), len(slots, 0.0), slen(slots, 0.0);
    for (uint64_t c = 0; c < cells; ++c)
        for (uint32_t q = 0; q < edge_count[c]; ++q) {
            const uint64_t k = edge_offset[c] + q;
            const size_t at = q * cells + c;
            if (neighbor[k] >= cells || edge[k] >= edges) throw std::invalid_argument("mesh index out of range");
            nb[at] = static_cast<int32_t>(neighbor[k]);
            ed[at] = static_cast<int32_t>(edge[k]);
            e_ne[at] = ne[k];
            e_nn[at] = nn[k];
            dist[at] = distance[edge[k]];
            len[at] = length[edge[k]];
            // edge_orientation: +1 from the first cell, -1 from the second
            if (first[edge[k]] == c)
                slen[at] = length[edge[k]];
            else if (second[edge[k]] == c)
                slen[at] = -length[edge[k]];
            else
                throw std::invalid_argument("cell is not incident to edge");
        }
    fv->first.assign(first, first + edges);
    fv->second.assign(second, second + edges);
    const size_t bytes = 256 * 16 + slots * (4 + 4 + 4 + 4 + 8 + 8 + 8) + cells * 8;
    psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    psim::hip_check(hipMalloc(&fv->block, bytes), "mesh upload allocation");
    char* cur = static_cast<char*>(fv->block);
    fv->nb = carve<int32_t>(cur, slots);
    fv->edge = carve<int32_t>(cur, slots);
    fv->ne = carve<float>(cur, slots);
    fv->nn = carve<float>(cur, slots);
    fv->dist = carve<double>(cur, slots);
    fv->length = carve<double>(cur, slots);
    fv->signed_length = carve<double>(cur, slots);
    fv->area = carve<double>(cur, cells);
    const auto put = [&](void* dst, const void* src, size_t b) {
        psim::hip_check(hipMemcpy(dst, src, b, hipMemcpyHostToDevice), "mesh upload copy");
    };
    put(fv->nb, nb.data(), slots * 4);
    put(fv->edge, ed.data(), slots * 4);
    put(fv->ne, e_ne.data(), slots * 4);
    put(fv->nn, e_nn.data(), slots * 4);
    put(fv->dist, dist.data(), slots * 8);
    put(fv->length, len.data(), slots * 8);
    put(fv->signed_length, slen.data(), slots * 8);
    put(fv->area, area, cells * 8);
    return fv.release();
}

void check_fields(const psim_fv* fv, int dtype, uint64_t fields) {
    if (fv == nullptr) throw std::invalid_argument("null mesh");
    if (dtype != PSIM_F64 && dtype != PSIM_F32) throw std::invalid_argument("unknown dtype");
    if (fields == 0 || fields > 65535) throw std::invalid_argument("field count out of range");
}

// Workgroups of a sweep: the tiles rounded up to a multiple of the XCD
// count, times the field groups (decoded by tile_and_group).
dim3 grid_of(int n, uint64_t fields) {
    const uint64_t tiles = (static_cast<uint64_t>(n) + psim::dev::kTile - 1) / psim::dev::kTile;
    const uint64_t groups = (fields + psim::dev::kFieldGroup - 1) / psim::dev::kFieldGroup;
    const uint64_t blocks = (tiles + psim::dev::kXcds - 1) / psim::dev::kXcds * psim::dev::kXcds * groups;
    if (blocks > 0x7fffffffULL) throw std::invalid_argument("cells x fields too large for one sweep");
    return dim3(static_cast<unsigned>(blocks));
}

// Launch `kernel<W>` for the layout's slot count W (6, 8 or 16).
#define PSIM_FV_LAUNCH(KERNEL, T, W, ...)                                                                       \
    do {                                                                                                        \
        if ((W) == 6)                                                                                           \
            hipLaunchKernelGGL((KERNEL<T, 6>), __VA_ARGS__);                                                    \
        else if ((W) == 8)                                                                                      \
            hipLaunchKernelGGL((KERNEL<T, 8>), __VA_ARGS__);                                                    \
        else                                                                                                    \
            hipLaunchKernelGGL((KERNEL<T, 16>), __VA_ARGS__);                                                   \
    } while (0)

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
        const int F = static_cast<int>(fields);
        if (dtype == PSIM_F64)
            PSIM_FV_LAUNCH(psim::dev::gradient_kernel, double, fv->width, grid_of(fv->n, fields), dim3(psim::dev::kTile),
                           0, stream, fv->n, F, fv->nb, fv->ne, fv->nn, fv->dist, static_cast<const double*>(s),
                           static_cast<double*>(e), static_cast<double*>(no));
        else
            PSIM_FV_LAUNCH(psim::dev::gradient_kernel, float, fv->width, grid_of(fv->n, fields), dim3(psim::dev::kTile),
                           0, stream, fv->n, F, fv->nb, fv->ne, fv->nn, fv->dist, static_cast<const float*>(s),
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
        const int F = static_cast<int>(fields);
        if (dtype == PSIM_F64)
            PSIM_FV_LAUNCH(psim::dev::divergence_kernel, double, fv->width, grid_of(fv->n, fields),
                           dim3(psim::dev::kTile), 0, stream, fv->n, fv->edges, F, fv->edge, fv->signed_length, fv->area,
                           static_cast<const double*>(f), static_cast<double*>(o));
        else
            PSIM_FV_LAUNCH(psim::dev::divergence_kernel, float, fv->width, grid_of(fv->n, fields),
                           dim3(psim::dev::kTile), 0, stream, fv->n, fv->edges, F, fv->edge, fv->signed_length, fv->area,
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
        const int F = static_cast<int>(fields);
        if (dtype == PSIM_F64)
            PSIM_FV_LAUNCH(psim::dev::laplacian_kernel, double, fv->width, grid_of(fv->n, fields),
                           dim3(psim::dev::kTile), 0, stream, fv->n, F, fv->nb, fv->length, fv->dist, fv->area,
                           static_cast<const double*>(s), static_cast<double*>(o));
        else
            PSIM_FV_LAUNCH(psim::dev::laplacian_kernel, float, fv->width, grid_of(fv->n, fields),
                           dim3(psim::dev::kTile), 0, stream, fv->n, F, fv->nb, fv->length, fv->dist, fv->area,
                           static_cast<const float*>(s), static_cast<float*>(o));
        psim::hip_check(hipGetLastError(), "laplacian launch");
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(result, o, bytes, hipMemcpyDeviceToHost, stream), "laplacian copy");
            psim::hip_check(hipStreamSynchronize(stream), "laplacian");
        }
    });
}

}  // extern "C"
