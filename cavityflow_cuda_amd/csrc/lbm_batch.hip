// Many independent cavities stepped together: one launch per half-step covers
// every cell of every member solver, and runs of half-steps are captured in a
// HIP graph. Members keep their own arrays (created by lbm_create), so a
// member's result is bit-identical alone and in a batch: the per-cell code is
// the same alpha_cell / beta_cell.
//
// Work is cut into tiles of block * cells_per_thread consecutive cells of one
// member; tile_member[t] names the member of tile t. Either one workgroup per
// tile (optionally renumbered per XCD) or a persistent grid of
// waves_per_cu * CUs workgroups striding over the tiles. With the pair pass
// (lbm_pair.hpp) one launch runs alpha and beta of every member out of place
// between each member's two population buffers. An autotuner times the
// candidate shapes on the actual batch and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "lbm_abi.h"
#include "lbm_cell.hpp"
#include "lbm_pair.hpp"

namespace lbm {

struct MemberDesc {
    const char* plane[2][kQ];   // alpha bases, beta bases (shifted by c_l)
    const uint8_t* flags;
    void* rho;
    void* u;
    int64_t stride;
    int32_t n, nx, nxy;
    uint32_t tile_begin;
    int32_t store_rho, store_u;
    double inv_tau, lid, g0, g1, g2;
    uint32_t* status;
};

template <class C>
__device__ __forceinline__ StepArgs<C> member_args(const MemberDesc& d, int phase)
{
    StepArgs<C> a;
    a.dd = nullptr;
    a.flags = d.flags;
    a.rho = (C*)d.rho;
    a.u = (C*)d.u;
    a.stride = d.stride;
    a.n = d.n;
    a.nx = d.nx;
    a.nxy = d.nxy;
    a.inv_tau = (C)d.inv_tau;
    a.lid = (C)d.lid;
    a.g0 = (C)d.g0;
    a.g1 = (C)d.g1;
    a.g2 = (C)d.g2;
    a.store_rho = d.store_rho;
    a.store_u = d.store_u;
    a.xpack_on = 0;
    a.xpack[0] = a.xpack[1] = nullptr;
    a.xunpack[0] = a.xunpack[1] = nullptr;
    a.status = d.status;
#pragma unroll
    for (int l = 0; l < kQ; l++)
        a.plane[l] = d.plane[phase][l];
    return a;
}

// Tile iteration over [tile_lo, tile_hi): first tile and stride of this
// workgroup (one tile, or a persistent stride over the range).
__device__ __forceinline__ void tile_range(uint32_t tile_lo, uint32_t tile_hi, int persistent,
                                           int xcd, uint32_t& first, uint32_t& step)
{
    if (!persistent) {
        first = tile_lo + (xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
        step = tile_hi;   // one tile only
    } else {
        first = tile_lo + blockIdx.x;
        step = gridDim.x;
    }
}

// One kernel per phase (PHASE 0 = alpha, 1 = beta), so that each compiles to
// the registers of its own half-step only.
template <class P, int BLOCK, bool NT, int PHASE>
__global__ __launch_bounds__(BLOCK) void k_batch(const MemberDesc* __restrict__ mem,
                                                 const int32_t* __restrict__ tile_member,
                                                 uint32_t tile_lo, uint32_t tiles, int cpt,
                                                 int persistent, int xcd)
{
    using C = typename P::C;
    using S = typename P::S;
    uint32_t first, step;
    tile_range(tile_lo, tiles, persistent, xcd, first, step);
    const uint32_t tile_cells = (uint32_t)BLOCK * (uint32_t)cpt;
    // A persistent workgroup strides over the tiles; consecutive tiles of a
    // large member share its descriptor, which is then read only once.
    int cur = -1;
    StepArgs<C> a;
    uint32_t tile_begin = 0;
    const char* plane0 = nullptr;
    for (uint32_t t = first; t < tiles; t += step) {
        const int m = tile_member[t];
        if (m != cur) {
            const MemberDesc& d = mem[m];
            a = member_args<C>(d, PHASE);
            tile_begin = d.tile_begin;
            plane0 = d.plane[0][0];
            cur = m;
        }
        const uint32_t base = (t - tile_begin) * tile_cells;
        if constexpr (PHASE == 0) {
            for (int k = 0; k < cpt; k++) {
                const uint32_t id = base + (uint32_t)k * BLOCK + threadIdx.x;
                if (id < (uint32_t)a.n) {
                    const Acc32<S, NT> acc{a.plane, id * (uint32_t)sizeof(S)};
                    alpha_cell<P, kXfNone>(a, id, ((const LBM_GLOBAL uint8_t*)a.flags)[id], acc);
                }
            }
        } else {
            const int64_t reach = (int64_t)a.nxy + a.nx + 1;
            const int64_t last = std::min<int64_t>((int64_t)base + tile_cells, a.n) - 1;
            const bool wrap = (int64_t)base < reach || last + reach >= (int64_t)a.n;
            for (int k = 0; k < cpt; k++) {
                const uint32_t id = base + (uint32_t)k * BLOCK + threadIdx.x;
                if (id >= (uint32_t)a.n)
                    continue;
                if (wrap) {
                    // the wrap path addresses planes from the unshifted base
                    const Acc64<S, true> acc{(S*)plane0, a.stride, (int32_t)id, a.n};
                    beta_cell<P, kXfNone>(a, id, ((const LBM_GLOBAL uint8_t*)a.flags)[id], acc);
                } else {
                    const Acc32<S, false> acc{a.plane, id * (uint32_t)sizeof(S)};
                    beta_cell<P, kXfNone>(a, id, ((const LBM_GLOBAL uint8_t*)a.flags)[id], acc);
                }
            }
        }
        if (!persistent)
            break;
    }
}


}  // namespace lbm

using namespace lbm;

struct GraphEntry {
    int first_phase, half_steps;
    hipStream_t stream;
    hipGraphExec_t exec;
};

struct lbm_batch {
    std::vector<lbm_solver*> members;
    int precision, device, cus;
    lbm_batch_tuning tuning;
    std::vector<MemberDesc> host_desc;
    MemberDesc* desc = nullptr;
    int32_t* tile_member = nullptr;
    uint32_t tiles = 0;        // tiles of the current tuning
    // tile ranges [lo, hi) of the member groups (tuning.group_mb), in order
    std::vector<std::pair<uint32_t, uint32_t>> groups;
    ResidualMember* res_mem = nullptr;   // members' residual inputs
    double* res_part = nullptr;          // members * kReduceBlocks * 5 partial sums
    double* res_out = nullptr;           // members * 5 sums
    PairDesc* pair_desc = nullptr;     // pair pass tables (tuning.pairs)
    int32_t* pair_member = nullptr;
    uint32_t pair_tiles = 0;
    int* check_flag = nullptr;         // boundary check result
    std::vector<GraphEntry> graphs;
    std::vector<lbm_solver*> alt_owned;  // members whose dd_alt this batch allocated
};

namespace {

#define BATCH_HIP(call, what)                  \
    do {                                       \
        hipError_t e_ = (call);                \
        if (e_ != hipSuccess)                  \
            return hip_fail(e_, what);         \
    } while (0)

// Destroys the cached graph executables once every launch of them has
// finished (the tables their kernel nodes point to are about to be freed or
// replaced).
void drop_graphs(lbm_batch* b)
{
    for (auto& g : b->graphs) {
        (void)hipStreamSynchronize(g.stream);
        (void)hipGraphExecDestroy(g.exec);
    }
    b->graphs.clear();
}

// Frees the second population buffers this batch allocated for the pair pass
// (after the device has finished with them).
void free_alt(lbm_batch* b)
{
    if (b->alt_owned.empty())
        return;
    (void)hipDeviceSynchronize();
    for (lbm_solver* s : b->alt_owned) {
        (void)hipFree(s->dd_alt);
        s->dd_alt = nullptr;
    }
    b->alt_owned.clear();
}

int upload_tiles(lbm_batch* b, uint32_t tile_cells, std::vector<MemberDesc>& desc,
                 int32_t** member_dev, uint32_t* ntiles, MemberDesc** desc_dev)
{
    std::vector<int32_t> map;
    for (size_t m = 0; m < b->members.size(); m++) {
        desc[m].tile_begin = (uint32_t)map.size();
        const uint32_t t = (uint32_t)((b->members[m]->n + tile_cells - 1) / tile_cells);
        map.insert(map.end(), t, (int32_t)m);
    }
    if (*member_dev)
        BATCH_HIP(hipFree(*member_dev), "hipFree");
    if (*desc_dev)
        BATCH_HIP(hipFree(*desc_dev), "hipFree");
    *member_dev = nullptr;
    *desc_dev = nullptr;
    if (hipMalloc((void**)member_dev, map.size() * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)desc_dev, desc.size() * sizeof(MemberDesc)) != hipSuccess) {
        set_error("batch: cannot allocate the tile table");
        return LBM_ERR_NOMEM;
    }
    BATCH_HIP(hipMemcpy(*member_dev, map.data(), map.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice), "H2D tile table");
    BATCH_HIP(hipMemcpy(*desc_dev, desc.data(), desc.size() * sizeof(MemberDesc),
                        hipMemcpyHostToDevice), "H2D member table");
    *ntiles = (uint32_t)map.size();
    return LBM_OK;
}

// candidates besides the starting tuning that the autotuner's final round
// times again over longer runs
constexpr size_t kFinalists = 4;

constexpr int kBatchOptions =
    LBM_LAUNCH_NONTEMPORAL | LBM_LAUNCH_XCD_REMAP | LBM_LAUNCH_WAVES4 | LBM_LAUNCH_WAVES3 |
    LBM_LAUNCH_WAVES2;

// Launch options a batch starts from before any autotuning (kernel 0 =
// alpha, 1 = beta): the single-domain options of 256-thread workgroups
// measured with the exchange for FP32 / F64, nontemporal alpha and plain
// beta for the 16-bit formats.
static int batch_start_options(int precision, int kernel)
{
    if (precision == LBM_F32 || precision == LBM_F64)
        return default_launch(precision, kernel, true, false) & kBatchOptions;
    return kernel == 0 ? LBM_LAUNCH_NONTEMPORAL : 0;
}

// Unused LDS reserved per workgroup for LBM_LAUNCH_WAVES4 / WAVES3 / WAVES2:
// caps the resident workgroups of a CU's 160 KB at 4 / 3 / 2 waves per SIMD.
size_t waves_lds(int opts, int block)
{
    const int waves = (opts & LBM_LAUNCH_WAVES2)   ? 2
                      : (opts & LBM_LAUNCH_WAVES3) ? 3
                      : (opts & LBM_LAUNCH_WAVES4) ? 4 : 0;
    if (!waves)
        return 0;
    const int per_cu = std::max(1, waves * 4 / (block / 64));   // workgroups per CU
    return (size_t)(160 * 1024) / (size_t)(per_cu + 1) + 1024;
}

bool same_tuning(const lbm_batch_tuning& a, const lbm_batch_tuning& b)
{
    return a.block == b.block && a.cells_per_thread == b.cells_per_thread &&
           a.waves_per_cu == b.waves_per_cu && a.options[0] == b.options[0] &&
           a.options[1] == b.options[1] && a.pairs == b.pairs && a.group_mb == b.group_mb;
}

bool valid_tuning(const lbm_batch_tuning& t)
{
    return t.pairs >= 0 && t.pairs <= kPairShapes &&
           (t.block == 256 || t.block == 512 || t.block == 1024) &&
           t.group_mb >= 0 && !(t.pairs && t.group_mb) &&
           (t.cells_per_thread == 1 || t.cells_per_thread == 2 || t.cells_per_thread == 4) &&
           t.waves_per_cu >= 0 && t.waves_per_cu <= 64 &&
           !(t.options[0] & ~kBatchOptions) && !(t.options[1] & ~kBatchOptions);
}

// Pair pass set-up: precision and boundary checks, second population
// buffers, and the column-tile tables.
void pair_tile(int pairs, int* tx, int* ty)
{
    switch (pairs) {
    case 2: *tx = PairShape1::kTX; *ty = PairShape1::kTY; break;
    case 3: *tx = PairShape2::kTX; *ty = PairShape2::kTY; break;
    case 4: *tx = PairShape3::kTX; *ty = PairShape3::kTY; break;
    default: *tx = PairShape0::kTX; *ty = PairShape0::kTY; break;
    }
}

int prepare_pairs(lbm_batch* b, int pairs)
{
    int ptx, pty;
    pair_tile(pairs, &ptx, &pty);
    if (b->precision == LBM_F64) {
        set_error("batch pair pass: FP64 is not supported (LDS window too large)");
        return LBM_ERR_UNSUPPORTED;
    }
    for (const lbm_solver* s : b->members)
        if (s->p.store_densities || s->p.store_velocities) {
            set_error("batch pair pass: members must not store rho / u outputs");
            return LBM_ERR_UNSUPPORTED;
        }
    if (!b->check_flag && hipMalloc((void**)&b->check_flag, sizeof(int)) != hipSuccess) {
        set_error("batch pair pass: cannot allocate the check flag");
        return LBM_ERR_NOMEM;
    }
    std::vector<PairDesc> desc(b->members.size());
    std::vector<int32_t> map;
    for (size_t m = 0; m < b->members.size(); m++) {
        lbm_solver* s = b->members[m];
        const int nx = s->size_h[0], ny = s->size_h[1], nz = s->size_h[2];
        BATCH_HIP(hipMemset(b->check_flag, 0, sizeof(int)), "hipMemset");
        hipLaunchKernelGGL(k_boundary_check, dim3(std::min<int64_t>((s->n + 255) / 256, 4096)),
                           dim3(256), 0, nullptr, s->flags, nx, ny, nz, b->check_flag);
        BATCH_HIP(hipGetLastError(), "boundary check launch");
        int bad = 0;
        BATCH_HIP(hipMemcpy(&bad, b->check_flag, sizeof(int), hipMemcpyDeviceToHost),
                  "D2H copy");
        if (bad) {
            set_error("batch pair pass: a member has a non-OBSTACLE cell on its array boundary");
            return LBM_ERR_UNSUPPORTED;
        }
        if (!s->dd_alt) {
            if (hipMalloc(&s->dd_alt, (size_t)kQ * s->stride * s->elem_bytes) != hipSuccess) {
                s->dd_alt = nullptr;
                set_error("batch pair pass: cannot allocate the second population buffer");
                return LBM_ERR_NOMEM;
            }
            b->alt_owned.push_back(s);
        }
        PairDesc& d = desc[m];
        const MemberDesc& h = b->host_desc[m];
        d.buf[0] = (const char*)s->dd;
        d.buf[1] = (const char*)s->dd_alt;
        d.flags = s->flags;
        d.stride = s->stride;
        d.nx = nx;
        d.ny = ny;
        d.nz = nz;
        d.tiles_x = (nx + ptx - 1) / ptx;
        d.tiles_y = (ny + pty - 1) / pty;
        d.chunks = (nz + kPairZChunk - 1) / kPairZChunk;
        d.tile_begin = (uint32_t)map.size();
        d.inv_tau = h.inv_tau;
        d.lid = h.lid;
        d.g0 = h.g0;
        d.g1 = h.g1;
        d.g2 = h.g2;
        d.status = s->status;
        map.insert(map.end(), (size_t)d.tiles_x * d.tiles_y * d.chunks, (int32_t)m);
    }
    if (b->pair_member)
        BATCH_HIP(hipFree(b->pair_member), "hipFree");
    if (b->pair_desc)
        BATCH_HIP(hipFree(b->pair_desc), "hipFree");
    b->pair_member = nullptr;
    b->pair_desc = nullptr;
    if (hipMalloc((void**)&b->pair_member, map.size() * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&b->pair_desc, desc.size() * sizeof(PairDesc)) != hipSuccess) {
        set_error("batch pair pass: cannot allocate the tile table");
        return LBM_ERR_NOMEM;
    }
    BATCH_HIP(hipMemcpy(b->pair_member, map.data(), map.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice), "H2D tile table");
    BATCH_HIP(hipMemcpy(b->pair_desc, desc.data(), desc.size() * sizeof(PairDesc),
                        hipMemcpyHostToDevice), "H2D member table");
    b->pair_tiles = (uint32_t)map.size();
    return LBM_OK;
}

int apply_tuning(lbm_batch* b, const lbm_batch_tuning& t)
{
    if (t.block == 1024 && b->precision == LBM_F64) {
        // 1024 threads leave 128 VGPRs per thread: the FP64 kernels would spill
        set_error("batch tuning: FP64 members take workgroups of 256 or 512 threads");
        return LBM_ERR_UNSUPPORTED;
    }
    drop_graphs(b);
    if (t.pairs) {
        int rc = prepare_pairs(b, t.pairs);
        if (rc) {
            if (!b->tuning.pairs)
                free_alt(b);   // a failed check leaves no buffer behind
            return rc;
        }
    } else {
        free_alt(b);           // no pair pass: no second population buffers
    }
    b->tuning = t;
    const uint32_t tile_cells = (uint32_t)(t.block * t.cells_per_thread);
    int rc = upload_tiles(b, tile_cells, b->host_desc, &b->tile_member, &b->tiles, &b->desc);
    if (rc)
        return rc;
    // members in order, packed into groups of at most group_mb MiB of
    // populations (a larger member alone); one group when group_mb is 0
    b->groups.clear();
    const uint64_t cap = (uint64_t)t.group_mb << 20;
    uint64_t bytes = 0;
    uint32_t lo = 0, hi = 0;
    for (const lbm_solver* s : b->members) {
        const uint64_t mb = (uint64_t)kQ * s->stride * s->elem_bytes;
        const uint32_t mt = (uint32_t)((s->n + tile_cells - 1) / tile_cells);
        if (cap && hi > lo && bytes + mb > cap) {
            b->groups.push_back({lo, hi});
            lo = hi;
            bytes = 0;
        }
        hi += mt;
        bytes += mb;
    }
    b->groups.push_back({lo, hi});
    return LBM_OK;
}

template <class P, class G>
hipError_t launch_pair_g(const lbm_batch* b, int parity, hipStream_t st)
{
    const int xcd = (b->tuning.options[0] & LBM_LAUNCH_XCD_REMAP) ? 1 : 0;
    const dim3 grid(b->pair_tiles), block(G::kBlock);
    if (b->tuning.options[0] & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_pair<P, true, G>), grid, block, 0, st, b->pair_desc,
                           b->pair_member, b->pair_tiles, parity, xcd);
    else
        hipLaunchKernelGGL((k_pair<P, false, G>), grid, block, 0, st, b->pair_desc,
                           b->pair_member, b->pair_tiles, parity, xcd);
    return hipGetLastError();
}

template <class P>
hipError_t launch_pair_p(const lbm_batch* b, int parity, hipStream_t st)
{
    switch (b->tuning.pairs) {
    case 1: return launch_pair_g<P, PairShape0>(b, parity, st);
    case 2: return launch_pair_g<P, PairShape1>(b, parity, st);
    case 3: return launch_pair_g<P, PairShape2>(b, parity, st);
    case 4: return launch_pair_g<P, PairShape3>(b, parity, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pair(const lbm_batch* b, int parity, hipStream_t st)
{
    switch (b->precision) {
    case LBM_F32: return launch_pair_p<StoreF32>(b, parity, st);
    case LBM_F16S: return launch_pair_p<StoreF16S>(b, parity, st);
    case LBM_BF16S: return launch_pair_p<StoreBF16S>(b, parity, st);
    case LBM_F16C: return launch_pair_p<StoreF16C>(b, parity, st);
    default: return hipErrorInvalidValue;
    }
}

template <class P, int BLOCK>
hipError_t launch_tiles_t(const lbm_batch* b, uint32_t lo, uint32_t hi, int phase,
                          hipStream_t st)
{
    const lbm_batch_tuning& t = b->tuning;
    const int opts = t.options[phase];
    const int persistent = t.waves_per_cu > 0;
    const uint32_t n = hi - lo;
    const uint32_t grid = persistent ? std::min<uint32_t>(n, (uint32_t)t.waves_per_cu * b->cus) : n;
    const int xcd = (opts & LBM_LAUNCH_XCD_REMAP) ? 1 : 0;
    const size_t lds = waves_lds(opts, BLOCK);
    // nontemporal accesses only for alpha (beta's shifted pattern loses with them)
    if (phase != 0)
        hipLaunchKernelGGL((k_batch<P, BLOCK, false, 1>), dim3(grid), dim3(BLOCK), lds, st,
                           b->desc, b->tile_member, lo, hi, t.cells_per_thread, persistent,
                           xcd);
    else if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_batch<P, BLOCK, true, 0>), dim3(grid), dim3(BLOCK), lds, st,
                           b->desc, b->tile_member, lo, hi, t.cells_per_thread, persistent,
                           xcd);
    else
        hipLaunchKernelGGL((k_batch<P, BLOCK, false, 0>), dim3(grid), dim3(BLOCK), lds, st,
                           b->desc, b->tile_member, lo, hi, t.cells_per_thread, persistent,
                           xcd);
    return hipGetLastError();
}

template <class P>
hipError_t launch_tiles_p(const lbm_batch* b, uint32_t lo, uint32_t hi, int phase,
                          hipStream_t st)
{
    if constexpr (!std::is_same_v<P, StoreF64>)
        if (b->tuning.block == 1024)
            return launch_tiles_t<P, 1024>(b, lo, hi, phase, st);
    return b->tuning.block == 512 ? launch_tiles_t<P, 512>(b, lo, hi, phase, st)
                                  : launch_tiles_t<P, 256>(b, lo, hi, phase, st);
}

// One half-step of the members whose tiles are [lo, hi).
hipError_t launch_tiles(const lbm_batch* b, uint32_t lo, uint32_t hi, int phase, hipStream_t st)
{
    if (hi <= lo)
        return hipSuccess;
    switch (b->precision) {
    case LBM_F32: return launch_tiles_p<StoreF32>(b, lo, hi, phase, st);
    case LBM_F64: return launch_tiles_p<StoreF64>(b, lo, hi, phase, st);
    case LBM_F16S: return launch_tiles_p<StoreF16S>(b, lo, hi, phase, st);
    case LBM_BF16S: return launch_tiles_p<StoreBF16S>(b, lo, hi, phase, st);
    case LBM_F16C: return launch_tiles_p<StoreF16C>(b, lo, hi, phase, st);
    default: return hipErrorInvalidValue;
    }
}

// One half-step of every member (one launch per group).
hipError_t launch_phase(const lbm_batch* b, int phase, hipStream_t st)
{
    for (const auto& g : b->groups) {
        const hipError_t e = launch_tiles(b, g.first, g.second, phase, st);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// The launches of half_steps half-steps from first_phase: single half-steps,
// or one pair pass per alpha + beta pair alternating between the buffers.
// With several member groups, each group runs all its half-steps before the
// next starts (members are independent, so any order gives the same bits).
hipError_t enqueue(const lbm_batch* b, int first_phase, int half_steps, hipStream_t st)
{
    hipError_t e = hipSuccess;
    if (b->tuning.pairs) {
        for (int i = 0; i < half_steps / 2 && e == hipSuccess; i++)
            e = launch_pair(b, i & 1, st);
        return e;
    }
    for (const auto& g : b->groups)
        for (int i = 0; i < half_steps && e == hipSuccess; i++)
            e = launch_tiles(b, g.first, g.second, (first_phase + i) & 1, st);
    return e;
}

int set_dev(int device)
{
    int cur = -1;
    BATCH_HIP(hipGetDevice(&cur), "hipGetDevice");
    if (cur != device)
        BATCH_HIP(hipSetDevice(device), "hipSetDevice");
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_create(lbm_solver* const* members, int count, lbm_batch** out)
{
    if (!members || count <= 0 || !out) {
        set_error("lbm_batch_create: invalid argument");
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    const int prec = members[0]->p.precision, dev = members[0]->p.device;
    for (int i = 0; i < count; i++) {
        const lbm_solver* s = members[i];
        if (!s || s->p.precision != prec || s->p.device != dev) {
            set_error("lbm_batch_create: members must share precision and device");
            return LBM_ERR_ARG;
        }
        if ((uint64_t)s->stride * (uint64_t)s->elem_bytes >= (1ull << 32)) {
            set_error("lbm_batch_create: a member is too large for 32-bit plane offsets");
            return LBM_ERR_UNSUPPORTED;
        }
        for (int j = 0; j < i; j++)
            if (members[j] == s) {
                set_error("lbm_batch_create: a solver appears twice");
                return LBM_ERR_ARG;
            }
    }
    int rc = set_dev(dev);
    if (rc)
        return rc;
    lbm_batch* b = new lbm_batch();
    b->members.assign(members, members + count);
    b->precision = prec;
    b->device = dev;
    if (hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        b->cus = 256;
    b->host_desc.resize(count);
    for (int i = 0; i < count; i++) {
        lbm_solver* s = members[i];
        MemberDesc& d = b->host_desc[i];
        const int nx = s->size_h[0], nxy = s->size_h[0] * s->size_h[1];
        for (int l = 0; l < kQ; l++) {
            const int64_t base = (int64_t)l * s->stride * s->elem_bytes;
            const int delta = cx(l) + cy(l) * nx + cz(l) * nxy;
            d.plane[0][l] = (const char*)s->dd + base;
            d.plane[1][l] = (const char*)s->dd + base + (int64_t)delta * s->elem_bytes;
        }
        d.flags = s->flags;
        d.rho = s->rho;
        d.u = s->u;
        d.stride = s->stride;
        d.n = (int32_t)s->n;
        d.nx = nx;
        d.nxy = nxy;
        d.store_rho = s->p.store_densities;
        d.store_u = s->p.store_velocities;
        d.inv_tau = s->p.inv_tau;
        d.lid = s->p.lid_velocity;
        d.g0 = s->p.gravitation[0];
        d.g1 = s->p.gravitation[1];
        d.g2 = s->p.gravitation[2];
        d.status = s->status;
    }
    lbm_batch_tuning t;
    t.block = 256;
    t.cells_per_thread = 1;
    t.waves_per_cu = 0;
    // the batch kernels take their shape from the tuning (block,
    // cells_per_thread)
    t.options[0] = batch_start_options(b->precision, 0);
    t.options[1] = batch_start_options(b->precision, 1);
    t.pairs = 0;
    t.group_mb = 0;
    if ((rc = apply_tuning(b, t))) {
        lbm_batch_destroy(b);
        return rc;
    }
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch* b)
{
    if (!b)
        return LBM_OK;
    (void)set_dev(b->device);
    (void)hipDeviceSynchronize();
    drop_graphs(b);
    free_alt(b);
    void* bufs[] = {b->desc,    b->tile_member, b->res_mem,     b->res_part,
                    b->res_out, b->pair_desc,   b->pair_member, b->check_flag};
    for (void* p : bufs)
        if (p)
            (void)hipFree(p);
    delete b;
    return LBM_OK;
}

int lbm_batch_size(const lbm_batch* b) { return b ? (int)b->members.size() : 0; }

int lbm_batch_set_tuning(lbm_batch* b, const lbm_batch_tuning* t)
{
    if (!b || !t || !valid_tuning(*t)) {
        set_error("lbm_batch_set_tuning: invalid tuning");
        return LBM_ERR_ARG;
    }
    int rc = set_dev(b->device);
    if (rc)
        return rc;
    return apply_tuning(b, *t);
}

int lbm_batch_get_tuning(const lbm_batch* b, lbm_batch_tuning* t)
{
    if (!b || !t) {
        set_error("lbm_batch_get_tuning: invalid argument");
        return LBM_ERR_ARG;
    }
    *t = b->tuning;
    return LBM_OK;
}

int lbm_batch_step(lbm_batch* b, int phase, hipStream_t stream)
{
    if (!b || phase < 0 || phase > 1) {
        set_error("lbm_batch_step: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc = set_dev(b->device);
    if (rc)
        return rc;
    BATCH_HIP(launch_phase(b, phase, stream), "batch kernel launch");
    return LBM_OK;
}

int lbm_batch_launch(lbm_batch* b, int first_phase, int half_steps, hipStream_t stream)
{
    if (!b || first_phase < 0 || first_phase > 1 || half_steps < 0) {
        set_error("lbm_batch_launch: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc = set_dev(b->device);
    if (rc)
        return rc;
    if (b->tuning.pairs && (first_phase != 0 || half_steps % 4)) {
        set_error("lbm_batch_launch: the pair pass needs first_phase 0 and a multiple of 4 "
                  "half-steps");
        return LBM_ERR_ARG;
    }
    BATCH_HIP(enqueue(b, first_phase, half_steps, stream), "batch kernel launch");
    return LBM_OK;
}

int lbm_batch_run(lbm_batch* b, int first_phase, int half_steps, hipStream_t stream)
{
    if (!b || first_phase < 0 || first_phase > 1 || half_steps < 0) {
        set_error("lbm_batch_run: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc = set_dev(b->device);
    if (rc)
        return rc;
    if (b->tuning.pairs && (first_phase != 0 || half_steps % 4)) {
        set_error("lbm_batch_run: the pair pass needs first_phase 0 and a multiple of 4 "
                  "half-steps");
        return LBM_ERR_ARG;
    }
    if (half_steps == 0)
        return LBM_OK;
    if (!stream) {   // the legacy default stream cannot be captured
        BATCH_HIP(enqueue(b, first_phase, half_steps, nullptr), "batch kernel launch");
        return LBM_OK;
    }
    for (auto& g : b->graphs)
        if (g.first_phase == first_phase && g.half_steps == half_steps && g.stream == stream) {
            BATCH_HIP(hipGraphLaunch(g.exec, stream), "hipGraphLaunch");
            return LBM_OK;
        }
    hipGraph_t graph = nullptr;
    BATCH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    const hipError_t le = enqueue(b, first_phase, half_steps, stream);
    hipError_t ee = hipStreamEndCapture(stream, &graph);
    if (le != hipSuccess || ee != hipSuccess) {
        if (ee == hipSuccess && graph)
            (void)hipGraphDestroy(graph);
        return le != hipSuccess ? hip_fail(le, "batch kernel capture")
                                : hip_fail(ee, "hipStreamEndCapture");
    }
    GraphEntry g{first_phase, half_steps, stream, nullptr};
    hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess)
        return hip_fail(ie, "hipGraphInstantiate");
    b->graphs.push_back(g);
    BATCH_HIP(hipGraphLaunch(g.exec, stream), "hipGraphLaunch");
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch* b, int trial_steps, hipStream_t stream,
                       lbm_batch_tuning* chosen)
{
    if (!b || trial_steps < 2) {
        set_error("lbm_batch_autotune: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc = set_dev(b->device);
    if (rc)
        return rc;
    // Trials time whole steps (alpha then beta), since each phase's speed
    // depends on what the other leaves in the caches: first the workgroup
    // shape with the current launch options, then alpha's options with beta's
    // fixed, then beta's with the chosen alpha, then the pair pass. The
    // members' state advances by whole steps during the trials (an even
    // number of pair passes), exactly as in a run.
    // the two timing events live until every return path of the tuner
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x)
                    (void)hipEventDestroy(x);
        }
    } ev;
    BATCH_HIP(hipEventCreate(&ev.e[0]), "hipEventCreate");
    BATCH_HIP(hipEventCreate(&ev.e[1]), "hipEventCreate");
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    // ms per whole step of candidate c, timed the way a run executes:
    // lbm_batch_run of an even number of whole steps (graph-captured on a
    // stream), once untimed (capture, warm-up) and once between two events
    auto step_time = [&](const lbm_batch_tuning& c, int steps, float* per_step) -> int {
        int r = apply_tuning(b, c);
        if (r)
            return r;
        const int whole = steps + (steps & 1);
        if ((r = lbm_batch_run(b, 0, 2 * whole, stream)))
            return r;
        BATCH_HIP(hipEventRecord(e0, stream), "hipEventRecord");
        if ((r = lbm_batch_run(b, 0, 2 * whole, stream)))
            return r;
        BATCH_HIP(hipEventRecord(e1, stream), "hipEventRecord");
        BATCH_HIP(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        BATCH_HIP(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        *per_step = ms / (float)whole;
        return LBM_OK;
    };
    lbm_batch_tuning best = b->tuning;
    best.pairs = 0;
    float best_step = std::numeric_limits<float>::infinity();
    std::vector<std::pair<float, lbm_batch_tuning>> timed;   // every trial, for the final round
    auto consider = [&](const lbm_batch_tuning& c) -> int {
        float t = 0.f;
        const int r = step_time(c, trial_steps, &t);
        if (r)
            return r;
        timed.push_back({t, c});
        if (t < best_step) {
            best_step = t;
            best = c;
        }
        return LBM_OK;
    };
    const lbm_batch_tuning start = best;
    // 1024-thread workgroups: not for FP64 (register spills), and at most two
    // are resident per CU
    const int max_block = b->precision == LBM_F64 ? 512 : 1024;
    for (int blk : {256, 512, 1024})
        for (int cpt : {1, 2, 4})
            for (int wpc : {0, 1, 2, 3, 4, 8, 16}) {
                if (blk > max_block || (blk == 1024 && wpc > 2))
                    continue;
                lbm_batch_tuning cand = start;
                cand.block = blk;
                cand.cells_per_thread = cpt;
                cand.waves_per_cu = wpc;
                if ((rc = consider(cand)))
                    return rc;
            }
    for (int ph = 0; ph < 2; ph++) {
        const lbm_batch_tuning shape = best;
        for (int opt : {0, 1, 2, 3, 4, 5, 8, 9, 16, 17}) {
            lbm_batch_tuning cand = shape;
            cand.options[ph] = opt;
            if ((rc = consider(cand)))
                return rc;
        }
    }
    // member groups small enough for the last-level cache to keep their state
    // between consecutive steps
    {
        const lbm_batch_tuning ungrouped = best;
        for (int mb : {32, 64, 128, 192, 256, 384}) {
            lbm_batch_tuning cand = ungrouped;
            cand.group_mb = mb;
            if ((rc = consider(cand)))
                return rc;
        }
    }
    // with the state of a group in the cache the best workgroup shape can
    // differ from the streaming one: the shapes once more, grouped
    if (best.group_mb > 0) {
        const lbm_batch_tuning grouped = best;
        for (int blk : {256, 512, 1024})
            for (int cpt : {1, 2, 4}) {
                if (blk > max_block || (blk == 1024 && grouped.waves_per_cu > 2))
                    continue;
                lbm_batch_tuning cand = grouped;
                cand.block = blk;
                cand.cells_per_thread = cpt;
                if ((rc = consider(cand)))
                    return rc;
            }
    }
    // the pair pass, when the members allow it
    bool eligible = true;
    for (int shape = 1; shape <= kPairShapes && eligible; shape++)
        for (int opt = 0; opt < 4; opt++) {
            lbm_batch_tuning cand = best;
            cand.pairs = shape;
            cand.group_mb = 0;
            cand.options[0] = opt;
            if (apply_tuning(b, cand) != LBM_OK) {
                eligible = false;   // not eligible: keep the in-place launches
                break;
            }
            float t = 0.f;
            if ((rc = step_time(cand, trial_steps, &t)))
                return rc;
            timed.push_back({t, cand});
            if (t < best_step) {
                best_step = t;
                best = cand;
            }
        }
    // Short trials are noisy, and the more candidates, the likelier a lucky
    // trial wins: the few fastest distinct candidates and the starting tuning
    // run again over longer runs, interleaved, and the least total time wins
    // (the starting tuning on a tie).
    std::stable_sort(timed.begin(), timed.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    std::vector<lbm_batch_tuning> finalists{start};
    for (const auto& tc : timed) {
        if (finalists.size() > kFinalists)
            break;
        bool seen = false;
        for (const lbm_batch_tuning& f : finalists)
            seen = seen || same_tuning(f, tc.second);
        if (!seen)
            finalists.push_back(tc.second);
    }
    if (finalists.size() > 1) {
        const int long_steps = 4 * trial_steps;
        std::vector<float> total(finalists.size(), 0.f);
        for (int rep = 0; rep < 2; rep++)
            for (size_t i = 0; i < finalists.size(); i++) {
                float t = 0.f;
                if ((rc = step_time(finalists[i], long_steps, &t)))
                    return rc;
                total[i] += t;
            }
        size_t win = 0;
        for (size_t i = 1; i < finalists.size(); i++)
            if (total[i] < total[win])
                win = i;
        best = finalists[win];
    }
    if ((rc = apply_tuning(b, best)))
        return rc;
    if (chosen)
        *chosen = best;
    return LBM_OK;
}

int lbm_batch_residual(lbm_batch* b, int layout, double* residuals, hipStream_t stream)
{
    if (!b || !residuals || layout < 0 || layout > 1) {
        set_error("lbm_batch_residual: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc = set_dev(b->device);
    if (rc)
        return rc;
    const int count = (int)b->members.size();
    if (!b->res_out) {
        std::vector<ResidualMember> mem(count);
        for (int i = 0; i < count; i++) {
            lbm_solver* s = b->members[i];
            if (!s->residual_prev) {
                const size_t bytes = (size_t)3 * s->stride * s->out_bytes;
                if (hipMalloc(&s->residual_prev, bytes) != hipSuccess) {
                    set_error("batch residual: out of device memory");
                    return LBM_ERR_NOMEM;
                }
                BATCH_HIP(hipMemset(s->residual_prev, 0, bytes), "hipMemset");
                s->residual_valid = 0;
            }
            mem[i] = ResidualMember{s->dd, s->flags, s->residual_prev, s->stride,
                                    (uint32_t)s->n, s->status};
        }
        if (hipMalloc((void**)&b->res_mem, mem.size() * sizeof(ResidualMember)) != hipSuccess ||
            hipMalloc((void**)&b->res_part,
                      (size_t)count * kReduceBlocks * 5 * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&b->res_out, (size_t)count * 5 * sizeof(double)) != hipSuccess) {
            set_error("batch residual: out of device memory");
            return LBM_ERR_NOMEM;
        }
        BATCH_HIP(hipMemcpy(b->res_mem, mem.data(), mem.size() * sizeof(ResidualMember),
                            hipMemcpyHostToDevice), "H2D residual table");
    }
    // every member's sums exactly as lbm_residual computes them for it alone
    BATCH_HIP(launch_residual_members(b->precision, b->res_mem, count, layout, b->res_part,
                                      b->res_out, stream), "batch residual launch");
    std::vector<double> host(5 * (size_t)count);
    BATCH_HIP(hipMemcpyAsync(host.data(), b->res_out, host.size() * sizeof(double),
                             hipMemcpyDeviceToHost, stream), "D2H copy");
    BATCH_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    // the snapshot and its validity are the member's own, shared with
    // lbm_residual and cleared by lbm_reset
    for (int i = 0; i < count; i++) {
        lbm_solver* s = b->members[i];
        residuals[i] = first_residual(s->residual_valid, host[5 * i], host[5 * i + 1]);
        s->residual_valid = 1;
    }
    return LBM_OK;
}

}  // extern "C"
