// Box copies between the strided device fields and dense [c][z][y][x]
// buffers, on the device. One kernel launch moves every selected component
// of a box (the reference issues one cudaMemcpy3D per lattice vector,
// CLbmSolverGPU.cu:553-568). Used for halo pack/unpack, checkpointing and the
// rho/u/flag extraction surface.
#include "lbm_d3q19.hpp"
#include "lbm_internal.hpp"
#include "lbm_storage.hpp"

namespace lbm {

constexpr int kCopyBlock = 256;

struct BoxArgs {
    int64_t stride;
    int32_t nx, nxy;
    int32_t ox, oy, oz;
    uint32_t cells;
    FastDiv div_bx, div_by;
    uint32_t* status;   // raised by a population store out of the storage range
    // populations: the lattice vectors copied, in order (grid.y = nplanes);
    // buffers keep all 19 planes, [l][cell], the others left alone
    uint8_t plane[kQ];
    int nplanes;
};

__device__ __forceinline__ uint32_t box_index(const BoxArgs& a, uint32_t gid)
{
    const uint32_t t = a.div_bx.div(gid);
    const uint32_t x = gid - t * a.div_bx.d;
    const uint32_t z = a.div_by.div(t);
    const uint32_t y = t - z * a.div_by.d;
    return (uint32_t)((a.oz + (int32_t)z) * a.nxy + (a.oy + (int32_t)y) * a.nx + a.ox + (int32_t)x);
}

// One population copy of up to two boxes in one launch (grid.z = box): the two
// faces of an exchange axis, each with its own plane list and buffer.
struct BoxCopy {
    BoxArgs box[2];
    void* buf[2];
};

// populations: grid.y = index into the box's plane list; shifted storage is
// converted to f. NT: streaming accesses, for boxes one cell thick on some axis
// (the faces of a halo exchange), which the next step does not reuse from the
// caches.
template <class P, bool NT>
__global__ __launch_bounds__(kCopyBlock) void k_get_dd(const typename P::S* __restrict__ dd,
                                                       BoxCopy c)
{
    const BoxArgs& a = c.box[blockIdx.z];
    if ((int)blockIdx.y >= a.nplanes)
        return;
    const int l = a.plane[blockIdx.y];
    const uint32_t gid = blockIdx.x * kCopyBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    typename P::C* __restrict__ dst = (typename P::C*)c.buf[blockIdx.z];
    const uint32_t id = box_index(a, gid);
    const typename P::S v = NT ? __builtin_nontemporal_load(dd + (int64_t)l * a.stride + id)
                               : dd[(int64_t)l * a.stride + id];
    dst[(int64_t)l * a.cells + gid] = P::load(v, l);
}

template <class P, bool NT>
__global__ __launch_bounds__(kCopyBlock) void k_set_dd(typename P::S* __restrict__ dd,
                                                       BoxCopy c)
{
    const BoxArgs& a = c.box[blockIdx.z];
    if ((int)blockIdx.y >= a.nplanes)
        return;
    const int l = a.plane[blockIdx.y];
    const uint32_t gid = blockIdx.x * kCopyBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    const typename P::C* __restrict__ src = (const typename P::C*)c.buf[blockIdx.z];
    const uint32_t id = box_index(a, gid);
    uint32_t ov = 0;
    const typename P::S v = P::store_checked(src[(int64_t)l * a.cells + gid], l, ov);
    raise_overflow<P>(a.status, ov);
    if (NT)
        __builtin_nontemporal_store(v, dd + (int64_t)l * a.stride + id);
    else
        dd[(int64_t)l * a.stride + id] = v;
}

// The lattice vectors of a component mask, in order.
struct PlaneList {
    uint8_t plane[kQ];
    int n;
};

static PlaneList plane_list(uint32_t mask)
{
    PlaneList p{};
    for (int l = 0; l < kQ; l++)
        if ((mask >> l) & 1u)
            p.plane[p.n++] = (uint8_t)l;
    return p;
}

// Both x faces of a solver in one launch: thread (l, r), r = 0..rows, writes
// the high-face cell of row r - 1 and the low-face cell of row r. In the
// halo planes (x = nx - 1 of one row, x = 0 of the next) the two cells are
// adjacent, so a pair shares its cache line instead of each write dirtying
// one line of its own. Planes outside a face's mask are left alone. NT:
// streaming stores (for the send planes, where the two faces' planes differ).
template <class P, bool NT>
__global__ __launch_bounds__(kCopyBlock) void k_set_dd_xpair(typename P::S* __restrict__ dd,
                                                             const typename P::C* __restrict__ lo,
                                                             const typename P::C* __restrict__ hi,
                                                             int64_t stride, int32_t nx,
                                                             int32_t x_lo, int32_t x_hi,
                                                             uint32_t rows, uint32_t mask_lo,
                                                             uint32_t mask_hi, uint32_t* status,
                                                             PlaneList planes)
{
    const int l = planes.plane[blockIdx.y];
    const bool on_lo = (mask_lo >> l) & 1u, on_hi = (mask_hi >> l) & 1u;
    const uint32_t r = blockIdx.x * kCopyBlock + threadIdx.x;
    if (r > rows)
        return;
    typename P::S* p = dd + (int64_t)l * stride;
    auto put = [&](int64_t i, typename P::S v) {
        if (NT)
            __builtin_nontemporal_store(v, p + i);
        else
            p[i] = v;
    };
    uint32_t ov = 0;
    if (on_hi && r > 0)
        put((int64_t)(r - 1) * nx + x_hi, P::store_checked(hi[(int64_t)l * rows + r - 1], l, ov));
    if (on_lo && r < rows)
        put((int64_t)r * nx + x_lo, P::store_checked(lo[(int64_t)l * rows + r], l, ov));
    raise_overflow<P>(status, ov);
}

// rho / u fields (compute type, no conversion)
template <class C>
__global__ __launch_bounds__(kCopyBlock) void k_get_field(const C* __restrict__ f,
                                                          C* __restrict__ dst, BoxArgs a)
{
    const int c = blockIdx.y;
    const uint32_t gid = blockIdx.x * kCopyBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    dst[(int64_t)c * a.cells + gid] = f[(int64_t)c * a.stride + box_index(a, gid)];
}

template <class C>
__global__ __launch_bounds__(kCopyBlock) void k_set_field(C* __restrict__ f,
                                                          const C* __restrict__ src, BoxArgs a)
{
    const int c = blockIdx.y;
    const uint32_t gid = blockIdx.x * kCopyBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    f[(int64_t)c * a.stride + box_index(a, gid)] = src[(int64_t)c * a.cells + gid];
}

__global__ __launch_bounds__(kCopyBlock) void k_get_flags(const uint8_t* __restrict__ f,
                                                          int32_t* __restrict__ dst, BoxArgs a)
{
    const uint32_t gid = blockIdx.x * kCopyBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    dst[gid] = f[box_index(a, gid)];
}

__global__ __launch_bounds__(kCopyBlock) void k_set_flags(uint8_t* __restrict__ f,
                                                          const int32_t* __restrict__ src,
                                                          BoxArgs a)
{
    const uint32_t gid = blockIdx.x * kCopyBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    f[box_index(a, gid)] = (uint8_t)src[gid];
}

static BoxArgs box_args(const lbm_solver* s, const Box& b, uint32_t mask)
{
    BoxArgs a;
    a.stride = s->stride;
    a.nx = s->size_h[0];
    a.nxy = s->size_h[0] * s->size_h[1];
    a.ox = b.ox;
    a.oy = b.oy;
    a.oz = b.oz;
    a.cells = b.cells();
    a.div_bx = FastDiv((uint32_t)b.bx);
    a.div_by = FastDiv((uint32_t)b.by);
    a.status = s->status;
    const PlaneList p = plane_list(mask);
    for (int l = 0; l < kQ; l++)
        a.plane[l] = p.plane[l];
    a.nplanes = p.n;
    return a;
}

// The population copy of `nbox` boxes (1 or 2) in one launch; boxes one cell
// thick on some axis stream.
template <class P, bool GET>
static hipError_t copy_dd_t(const lbm_solver* s, const Box* b, void* const* buf,
                            const uint32_t* mask, int nbox, hipStream_t st)
{
    BoxCopy c{};
    uint32_t cells = 0;
    int nplanes = 0;
    bool thin = true;
    for (int k = 0; k < nbox; k++) {
        c.box[k] = box_args(s, b[k], mask[k]);
        c.buf[k] = buf[k];
        cells = c.box[k].cells > cells ? c.box[k].cells : cells;
        nplanes = c.box[k].nplanes > nplanes ? c.box[k].nplanes : nplanes;
        thin = thin && (b[k].bx == 1 || b[k].by == 1 || b[k].bz == 1);
    }
    if (!nplanes || !cells)
        return hipSuccess;
    const dim3 grid((cells + kCopyBlock - 1) / kCopyBlock, nplanes, nbox);
    if (GET) {
        const auto* dd = (const typename P::S*)s->dd;
        if (thin)
            hipLaunchKernelGGL((k_get_dd<P, true>), grid, dim3(kCopyBlock), 0, st, dd, c);
        else
            hipLaunchKernelGGL((k_get_dd<P, false>), grid, dim3(kCopyBlock), 0, st, dd, c);
    } else {
        auto* dd = (typename P::S*)s->dd;
        if (thin)
            hipLaunchKernelGGL((k_set_dd<P, true>), grid, dim3(kCopyBlock), 0, st, dd, c);
        else
            hipLaunchKernelGGL((k_set_dd<P, false>), grid, dim3(kCopyBlock), 0, st, dd, c);
    }
    return hipGetLastError();
}

template <bool GET>
static hipError_t copy_dd(const lbm_solver* s, const Box* b, void* const* buf,
                          const uint32_t* mask, int nbox, hipStream_t st)
{
    switch (s->p.precision) {
    case LBM_F32: return copy_dd_t<StoreF32, GET>(s, b, buf, mask, nbox, st);
    case LBM_F64: return copy_dd_t<StoreF64, GET>(s, b, buf, mask, nbox, st);
    case LBM_F16S: return copy_dd_t<StoreF16S, GET>(s, b, buf, mask, nbox, st);
    case LBM_BF16S: return copy_dd_t<StoreBF16S, GET>(s, b, buf, mask, nbox, st);
    case LBM_F16C: return copy_dd_t<StoreF16C, GET>(s, b, buf, mask, nbox, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dd_get_pair(const lbm_solver* s, const Box* b, void* const* dst,
                              const uint32_t* mask, hipStream_t st)
{
    return copy_dd<true>(s, b, dst, mask, 2, st);
}

hipError_t launch_dd_set_pair(const lbm_solver* s, const Box* b, const void* const* src,
                              const uint32_t* mask, hipStream_t st)
{
    return copy_dd<false>(s, b, const_cast<void* const*>(src), mask, 2, st);
}

// Face exchanges between solvers of a device with no buffer between them,
// up to kXferBatch in one launch: plane l of a transfer's send box goes to the
// same cell of its receive box (grid.y = index into the plane list), through
// the compute type as a pack followed by an unpack takes it, the range of the
// stored value marked on the receiver as the unpack marks it. The two boxes
// of a transfer have the same extents; grid.x is the transfers' blocks one
// after the other, and the transfers of one launch touch disjoint elements.
struct XferBatch {
    const void* src[kXferBatch];
    void* dst[kXferBatch];
    BoxArgs sa[kXferBatch], da[kXferBatch];
    uint32_t block0[kXferBatch + 1];   // first block of each transfer, then the total
    int n;
};

template <class P>
__global__ __launch_bounds__(kCopyBlock) void k_xfer_dd(const XferBatch b)
{
    int t = 0;
    while (t + 1 < b.n && blockIdx.x >= b.block0[t + 1])
        t++;
    const BoxArgs& sa = b.sa[t];
    const BoxArgs& da = b.da[t];
    if ((int)blockIdx.y >= sa.nplanes)
        return;
    const int l = sa.plane[blockIdx.y];
    const uint32_t gid = (blockIdx.x - b.block0[t]) * kCopyBlock + threadIdx.x;
    if (gid >= sa.cells)
        return;
    const typename P::S* src = (const typename P::S*)b.src[t];
    typename P::S* dst = (typename P::S*)b.dst[t];
    const typename P::S v =
        __builtin_nontemporal_load(src + (int64_t)l * sa.stride + box_index(sa, gid));
    uint32_t ov = 0;
    const typename P::S w = P::store_checked(P::load(v, l), l, ov);
    raise_overflow<P>(da.status, ov);
    __builtin_nontemporal_store(w, dst + (int64_t)l * da.stride + box_index(da, gid));
}

template <class P>
static hipError_t xfer_dd_t(const XferDesc* list, int n, hipStream_t st)
{
    for (int first = 0; first < n; first += kXferBatch) {
        XferBatch b{};
        int nplanes = 0;
        uint32_t blocks = 0;
        for (int k = first; k < n && k < first + kXferBatch; k++) {
            const XferDesc& d = list[k];
            const BoxArgs sa = box_args(d.from, d.sb, d.mask);
            if (!sa.nplanes || !sa.cells)
                continue;
            b.src[b.n] = d.from->dd;
            b.dst[b.n] = d.to->dd;
            b.sa[b.n] = sa;
            b.da[b.n] = box_args(d.to, d.db, d.mask);
            b.block0[b.n] = blocks;
            blocks += (sa.cells + kCopyBlock - 1) / kCopyBlock;
            nplanes = sa.nplanes > nplanes ? sa.nplanes : nplanes;
            b.n++;
        }
        if (!b.n)
            continue;
        b.block0[b.n] = blocks;
        hipLaunchKernelGGL((k_xfer_dd<P>), dim3(blocks, nplanes), dim3(kCopyBlock), 0, st, b);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_dd_xfer(const XferDesc* list, int n, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    switch (list[0].from->p.precision) {
    case LBM_F32: return xfer_dd_t<StoreF32>(list, n, st);
    case LBM_F64: return xfer_dd_t<StoreF64>(list, n, st);
    case LBM_F16S: return xfer_dd_t<StoreF16S>(list, n, st);
    case LBM_BF16S: return xfer_dd_t<StoreBF16S>(list, n, st);
    case LBM_F16C: return xfer_dd_t<StoreF16C>(list, n, st);
    default: return hipErrorInvalidValue;
    }
}

template <class C>
static hipError_t field_t(const lbm_solver* s, int field, const Box& b, void* buf, bool get,
                          hipStream_t st)
{
    BoxArgs a = box_args(s, b, 0);
    const int ncomp = field == 1 ? 3 : 1;
    C* f = (C*)(field == 1 ? s->u : s->rho);
    dim3 grid((a.cells + kCopyBlock - 1) / kCopyBlock, ncomp);
    if (get)
        hipLaunchKernelGGL(k_get_field<C>, grid, dim3(kCopyBlock), 0, st, f, (C*)buf, a);
    else
        hipLaunchKernelGGL(k_set_field<C>, grid, dim3(kCopyBlock), 0, st, f, (const C*)buf, a);
    return hipGetLastError();
}

hipError_t launch_box_get(const lbm_solver* s, int field, const Box& b, void* dst,
                          uint32_t mask, hipStream_t st)
{
    if (field == 0)
        return copy_dd<true>(s, &b, &dst, &mask, 1, st);
    if (s->p.precision == LBM_F64)
        return field_t<double>(s, field, b, dst, true, st);
    return field_t<float>(s, field, b, dst, true, st);
}

hipError_t launch_box_set(const lbm_solver* s, int field, const Box& b, const void* src,
                          uint32_t mask, hipStream_t st)
{
    if (field == 0) {
        void* buf = const_cast<void*>(src);
        return copy_dd<false>(s, &b, &buf, &mask, 1, st);
    }
    if (s->p.precision == LBM_F64)
        return field_t<double>(s, field, b, const_cast<void*>(src), false, st);
    return field_t<float>(s, field, b, const_cast<void*>(src), false, st);
}

template <class P>
static hipError_t set_dd_xpair_t(const lbm_solver* s, const void* lo, const void* hi, int x_lo,
                                 int x_hi, uint32_t mask_lo, uint32_t mask_hi, bool nt,
                                 hipStream_t st)
{
    const uint32_t rows = (uint32_t)s->size_h[1] * (uint32_t)s->size_h[2];
    const PlaneList planes = plane_list(mask_lo | mask_hi);
    if (!planes.n)
        return hipSuccess;
    dim3 grid((rows + 1 + kCopyBlock - 1) / kCopyBlock, planes.n);
    if (nt)
        hipLaunchKernelGGL((k_set_dd_xpair<P, true>), grid, dim3(kCopyBlock), 0, st,
                           (typename P::S*)s->dd, (const typename P::C*)lo,
                           (const typename P::C*)hi, s->stride, s->size_h[0], x_lo, x_hi, rows,
                           mask_lo, mask_hi, s->status, planes);
    else
        hipLaunchKernelGGL((k_set_dd_xpair<P, false>), grid, dim3(kCopyBlock), 0, st,
                           (typename P::S*)s->dd, (const typename P::C*)lo,
                           (const typename P::C*)hi, s->stride, s->size_h[0], x_lo, x_hi, rows,
                           mask_lo, mask_hi, s->status, planes);
    return hipGetLastError();
}

hipError_t launch_set_dd_xpair(const lbm_solver* s, const void* lo, const void* hi, int x_lo,
                               int x_hi, uint32_t mask_lo, uint32_t mask_hi, bool nt,
                               hipStream_t st)
{
    switch (s->p.precision) {
    case LBM_F32:
        return set_dd_xpair_t<StoreF32>(s, lo, hi, x_lo, x_hi, mask_lo, mask_hi, nt, st);
    case LBM_F64:
        return set_dd_xpair_t<StoreF64>(s, lo, hi, x_lo, x_hi, mask_lo, mask_hi, nt, st);
    case LBM_F16S:
        return set_dd_xpair_t<StoreF16S>(s, lo, hi, x_lo, x_hi, mask_lo, mask_hi, nt, st);
    case LBM_BF16S:
        return set_dd_xpair_t<StoreBF16S>(s, lo, hi, x_lo, x_hi, mask_lo, mask_hi, nt, st);
    case LBM_F16C:
        return set_dd_xpair_t<StoreF16C>(s, lo, hi, x_lo, x_hi, mask_lo, mask_hi, nt, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_flags_get(const lbm_solver* s, const Box& b, int32_t* dst, hipStream_t st)
{
    BoxArgs a = box_args(s, b, 0);
    hipLaunchKernelGGL(k_get_flags, dim3((a.cells + kCopyBlock - 1) / kCopyBlock),
                       dim3(kCopyBlock), 0, st, s->flags, dst, a);
    return hipGetLastError();
}

hipError_t launch_flags_set(const lbm_solver* s, const Box& b, const int32_t* src,
                            hipStream_t st)
{
    BoxArgs a = box_args(s, b, 0);
    hipLaunchKernelGGL(k_set_flags, dim3((a.cells + kCopyBlock - 1) / kCopyBlock),
                       dim3(kCopyBlock), 0, st, s->flags, src, a);
    return hipGetLastError();
}

}  // namespace lbm
