// C ABI of the HIP solver: argument checks, device memory, host<->device
// staging, and dispatch to the kernels. No entry point calls exit(); errors
// come back as lbm_status codes with a per-thread message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "lbm_abi.h"
#include "lbm_d3q19.hpp"
#include "lbm_internal.hpp"

namespace lbm {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int hip_fail(hipError_t e, const char* what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return LBM_ERR_HIP;
}

}  // namespace lbm

using namespace lbm;

#define LBM_HIP(call, what)                       \
    do {                                          \
        hipError_t e_ = (call);                   \
        if (e_ != hipSuccess)                     \
            return hip_fail(e_, what);            \
    } while (0)

namespace {

bool valid_flag(int f)
{
    return f == LBM_OBSTACLE || f == LBM_FLUID || f == LBM_VELOCITY_INJECTION ||
           f == LBM_GHOST_LAYER;
}

int element_bytes(int precision)
{
    switch (precision) {
    case LBM_F32: return 4;
    case LBM_F64: return 8;
    case LBM_F16S: return 2;
    case LBM_BF16S: return 2;
    case LBM_F16C: return 2;
    default: return 0;
    }
}

// Resolves a (possibly NULL) box. whole: 0 -> full halo-inclusive domain,
// 1 -> interior. Reference semantic: assert origin >= 0, size > 0,
// origin + size <= size_h (CLbmSolverCPU.cpp:145-149).
int resolve_box(const lbm_solver* s, const int* origin, const int* size, int whole, Box* b)
{
    if (!s) {
        set_error("solver is NULL");
        return LBM_ERR_ARG;
    }
    if (!origin || !size) {
        if (whole == 0)
            *b = Box{0, 0, 0, s->size_h[0], s->size_h[1], s->size_h[2]};
        else
            *b = Box{1, 1, 1, s->p.size[0], s->p.size[1], s->p.size[2]};
        return LBM_OK;
    }
    for (int i = 0; i < 3; i++) {
        if (origin[i] < 0 || size[i] <= 0 || origin[i] + size[i] > s->size_h[i]) {
            char buf[256];
            snprintf(buf, sizeof buf,
                     "box origin (%d,%d,%d) size (%d,%d,%d) outside domain with halo (%d,%d,%d)",
                     origin[0], origin[1], origin[2], size[0], size[1], size[2], s->size_h[0],
                     s->size_h[1], s->size_h[2]);
            set_error(buf);
            return LBM_ERR_BOX;
        }
    }
    *b = Box{origin[0], origin[1], origin[2], size[0], size[1], size[2]};
    return LBM_OK;
}

int set_device(const lbm_solver* s)
{
    int cur = -1;
    LBM_HIP(hipGetDevice(&cur), "hipGetDevice");
    if (cur != s->p.device)
        LBM_HIP(hipSetDevice(s->p.device), "hipSetDevice");
    return LBM_OK;
}

int ensure_staging(lbm_solver* s, size_t bytes)
{
    if (s->staging_bytes >= bytes)
        return LBM_OK;
    if (s->staging)
        LBM_HIP(hipFree(s->staging), "hipFree(staging)");
    s->staging = nullptr;
    s->staging_bytes = 0;
    if (hipMalloc(&s->staging, bytes) != hipSuccess) {
        set_error("cannot allocate device staging buffer");
        return LBM_ERR_NOMEM;
    }
    s->staging_bytes = bytes;
    return LBM_OK;
}

// Host <-> device box copy of one field through the staging buffer.
int copy_get(lbm_solver* s, int field, const Box& b, int ncomp, void* dst, int dst_on_device,
             hipStream_t st)
{
    const size_t bytes = (size_t)b.cells() * ncomp * s->out_bytes;
    if (dst_on_device) {
        LBM_HIP(launch_box_get(s, field, b, dst, 0x7FFFFu, st), "box get kernel");
        return LBM_OK;
    }
    int rc = ensure_staging(s, bytes);
    if (rc)
        return rc;
    LBM_HIP(launch_box_get(s, field, b, s->staging, 0x7FFFFu, st), "box get kernel");
    LBM_HIP(hipMemcpyAsync(dst, s->staging, bytes, hipMemcpyDeviceToHost, st), "D2H copy");
    LBM_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    return LBM_OK;
}

int copy_set(lbm_solver* s, int field, const Box& b, int ncomp, const void* src,
             int src_on_device, uint32_t mask, hipStream_t st)
{
    const size_t bytes = (size_t)b.cells() * ncomp * s->out_bytes;
    if (src_on_device) {
        LBM_HIP(launch_box_set(s, field, b, src, mask, st), "box set kernel");
        return LBM_OK;
    }
    int rc = ensure_staging(s, bytes);
    if (rc)
        return rc;
    LBM_HIP(hipMemcpyAsync(s->staging, src, bytes, hipMemcpyHostToDevice, st), "H2D copy");
    LBM_HIP(launch_box_set(s, field, b, s->staging, mask, st), "box set kernel");
    LBM_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    return LBM_OK;
}

// norm . c_l > 0 mask of setDensityDistributions(..., direction, ...)
// (CLbmSolverCPU.cpp:564-619).
uint32_t direction_mask(int direction)
{
    int norm[3] = {0, 0, 0};
    switch (direction) {
    case LBM_LEFT: norm[0] = 1; break;
    case LBM_RIGHT: norm[0] = -1; break;
    case LBM_BOTTOM: norm[1] = 1; break;
    case LBM_TOP: norm[1] = -1; break;
    case LBM_BACK: norm[2] = 1; break;
    case LBM_FRONT: norm[2] = -1; break;
    default: return 0;
    }
    uint32_t m = 0;
    for (int l = 0; l < kQ; l++)
        if (norm[0] * cx(l) + norm[1] * cy(l) + norm[2] * cz(l) > 0)
            m |= 1u << l;
    return m;
}

// The lattice vectors a face exchange moves (lbm_pack_face / lbm_unpack_face).
// `face` is the sending solver's face. In the beta step the slot of plane l at
// cell p is read and written only by cell p - c_l, so of a halo plane only the
// 5 planes whose vectors point out through its face (read by the interior
// cells next to it) reach the interior; the other 14 stay in the halo or are
// moved (by the linear wrap, or by the halo cells' own pass-through) into
// interior slots that the phase-1 unpack overwrites before anything reads
// them. Phase 0 therefore sends those 5 planes of the send plane (vectors
// pointing back into the sender: direction_mask(face)), phase 1 the 5 planes
// of the halo plane the receiver's direction filter applies
// (direction_mask(face ^ 1)). The reference moves all 19 planes after alpha
// (CController.cpp:369-570); the interiors are the same bit for bit.
uint32_t exchange_mask(int face, int phase)
{
    return direction_mask(phase == 0 ? face : face ^ 1);
}

// The CLbmSolver constructor's time-step adaptation (CLbmSolver.cpp:75-281)
// in arithmetic type T. REYNOLDS_DEFAULT and TAU_DEFAULT are double literals
// in the reference (constants.h:26-27) and stay double here.
template <class T>
int compute_parameters_t(const lbm_physics* in, lbm_derived* out)
{
    constexpr double kReynoldsDefault = 1000.0;
    constexpr double kTauDefault = 0.85;
    const T eps = std::numeric_limits<T>::epsilon();
    const T vel[3] = {(T)in->velocity[0], (T)in->velocity[1], (T)in->velocity[2]};
    const T acc[3] = {(T)in->acceleration[0], (T)in->acceleration[1], (T)in->acceleration[2]};
    auto len = [](const T v[3]) { return (T)std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
    const T max_vel = (T)in->max_velocity, max_acc = (T)in->max_acceleration;
    const T cell_length = (T)in->length_x / (T)in->size_x;
    T viscosity = (T)in->viscosity;
    T timestep = (T)in->timestep;
    T vd[3], ad[3], visc_d = 0, tau = 0, old_timestep;
    bool by_velocity = false, by_acceleration = false;

    if (viscosity <= (T)0)
        viscosity = (T)in->length_max * vel[0] / kReynoldsDefault;

    for (int iter = 0;; iter++) {
        if (iter >= 64) {
            set_error("time-step adaptation did not settle");
            return LBM_ERR_PARAMS;
        }
        T sc = timestep / cell_length;
        for (int i = 0; i < 3; i++)
            vd[i] = vel[i] * sc;
        sc = (timestep * timestep) / cell_length;
        for (int i = 0; i < 3; i++)
            ad[i] = acc[i] * sc;
        visc_d = viscosity * (timestep / (cell_length * cell_length));
        if (len(vd) > max_vel + eps) {
            by_velocity = true;
            old_timestep = timestep;
            timestep = (max_vel * cell_length) / len(vel);
            continue;
        }
        if (len(ad) > max_acc + eps) {
            by_acceleration = true;
            old_timestep = timestep;
            timestep = (T)std::sqrt((max_acc * cell_length) / len(acc));
            continue;
        }
        tau = (T)0.5 * ((T)6 * visc_d + (T)1);
        if (tau - eps < (T)0.51 || tau + eps > (T)1.9) {
            old_timestep = timestep;
            timestep = (cell_length * cell_length) * (((T)2 * kTauDefault - (T)1) / ((T)6 * viscosity));
            if ((by_velocity || by_acceleration) && timestep > old_timestep) {
                set_error("no valid timestep size satisfies viscosity, max velocity, "
                          "max acceleration and tau limits");
                return LBM_ERR_PARAMS;
            }
            continue;
        }
        break;
    }
    for (int i = 0; i < 3; i++) {
        out->velocity_dimless[i] = vd[i];
        out->acceleration_dimless[i] = ad[i];
    }
    out->viscosity = viscosity;
    out->viscosity_dimless = visc_d;
    out->tau = tau;
    out->inv_tau = (T)1 / tau;
    out->timestep = timestep;
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_abi_version(void) { return 1; }

const char* lbm_last_error(void) { return g_error.c_str(); }

const char* lbm_status_string(int status)
{
    switch (status) {
    case LBM_OK: return "ok";
    case LBM_ERR_ARG: return "invalid argument";
    case LBM_ERR_BOX: return "box outside the domain";
    case LBM_ERR_HIP: return "HIP runtime error";
    case LBM_ERR_NOMEM: return "out of device memory";
    case LBM_ERR_PARAMS: return "no valid time step";
    case LBM_ERR_STATE: return "invalid solver state";
    case LBM_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

int lbm_compute_parameters(const lbm_physics* in, int precision, lbm_derived* out)
{
    if (!in || !out || in->size_x <= 0 || !(in->length_x > 0) || !(in->timestep > 0) ||
        !(in->max_velocity > 0) || !(in->max_acceleration > 0)) {
        set_error("lbm_compute_parameters: invalid physics input");
        return LBM_ERR_ARG;
    }
    if (precision == LBM_F64)
        return compute_parameters_t<double>(in, out);
    return compute_parameters_t<float>(in, out);
}

namespace {

// Shared validation of lbm_create / lbm_device_bytes: cells with halo.
int check_params(const lbm_params& p, const char* who, int64_t* cells)
{
    if (!element_bytes(p.precision)) {
        set_error(std::string(who) + ": unknown precision");
        return LBM_ERR_ARG;
    }
    for (int i = 0; i < 3; i++)
        if (p.size[i] < 1) {
            set_error(std::string(who) + ": size must be >= 1 on every axis");
            return LBM_ERR_ARG;
        }
    for (int i = 0; i < 6; i++)
        if (!valid_flag(p.bc[i])) {
            set_error(std::string(who) + ": boundary condition is not a Flag value");
            return LBM_ERR_ARG;
        }
    const int64_t nx = p.size[0] + 2, ny = p.size[1] + 2, nz = p.size[2] + 2;
    const int64_t n = nx * ny * nz;
    // 32-bit cell indices with room for the largest neighbour offset
    if (n + nx * ny + nx + 1 >= (int64_t)INT32_MAX) {
        set_error(std::string(who) + ": domain too large for 32-bit cell indexing");
        return LBM_ERR_ARG;
    }
    *cells = n;
    return LBM_OK;
}

}  // namespace

int lbm_device_bytes(const lbm_params* params, int pair_buffer, uint64_t* bytes)
{
    if (!params || !bytes) {
        set_error("lbm_device_bytes: NULL argument");
        return LBM_ERR_ARG;
    }
    int64_t n = 0;
    int rc = check_params(*params, "lbm_device_bytes", &n);
    if (rc)
        return rc;
    const uint64_t stride = (uint64_t)(n + 255) / 256 * 256;
    const uint64_t eb = (uint64_t)element_bytes(params->precision);
    const uint64_t ob = params->precision == LBM_F64 ? 8 : 4;
    uint64_t b = (uint64_t)kQ * stride * eb + stride;                  // populations, flags
    b += params->store_densities ? stride * ob : 0;
    b += params->store_velocities ? 3 * stride * ob : 0;
    b += kPartialsBytes;                                               // reduction partials
    b += 3 * stride * ob;                                              // residual snapshot
    b += pair_buffer ? (uint64_t)kQ * stride * eb : 0;                 // second populations
    *bytes = b;
    return LBM_OK;
}

int lbm_create(const lbm_params* params, lbm_solver** out)
{
    if (!params || !out) {
        set_error("lbm_create: NULL argument");
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    const lbm_params& p = *params;
    int64_t n = 0;
    int rc0 = check_params(p, "lbm_create", &n);
    if (rc0)
        return rc0;
    const int eb = element_bytes(p.precision);
    const int64_t nx = p.size[0] + 2, ny = p.size[1] + 2, nz = p.size[2] + 2;
    lbm_solver* s = new lbm_solver();
    std::memset(s, 0, sizeof(*s));
    s->p = p;
    s->size_h[0] = (int)nx;
    s->size_h[1] = (int)ny;
    s->size_h[2] = (int)nz;
    s->n = n;
    s->stride = (n + 255) / 256 * 256;
    s->elem_bytes = eb;
    s->out_bytes = p.precision == LBM_F64 ? 8 : 4;
    s->launch[0] = default_launch(p.precision, 0, false, s->n >= kWideCells);
    s->launch[1] = default_launch(p.precision, 1, false, s->n >= kWideCells);
    s->launch_set[0] = s->launch_set[1] = 0;
    // round inv_tau / lid / gravity to the arithmetic type once
    if (p.precision != LBM_F64) {
        s->p.inv_tau = (float)p.inv_tau;
        s->p.lid_velocity = (float)p.lid_velocity;
        for (int i = 0; i < 3; i++)
            s->p.gravitation[i] = (float)p.gravitation[i];
    }
    int rc = set_device(s);
    if (rc) {
        delete s;
        return rc;
    }
    auto fail = [&](int code, const char* what) {
        set_error(std::string("lbm_create: ") + what);
        lbm_destroy(s);
        return code;
    };
    if (hipMalloc(&s->dd, (size_t)kQ * s->stride * eb) != hipSuccess)
        return fail(LBM_ERR_NOMEM, "cannot allocate populations");
    if (hipMalloc((void**)&s->flags, (size_t)s->stride) != hipSuccess)
        return fail(LBM_ERR_NOMEM, "cannot allocate flags");
    if (p.store_densities && hipMalloc(&s->rho, (size_t)s->stride * s->out_bytes) != hipSuccess)
        return fail(LBM_ERR_NOMEM, "cannot allocate densities");
    if (p.store_velocities &&
        hipMalloc(&s->u, (size_t)3 * s->stride * s->out_bytes) != hipSuccess)
        return fail(LBM_ERR_NOMEM, "cannot allocate velocities");
    if (hipMalloc((void**)&s->partials, kPartialsBytes) != hipSuccess)
        return fail(LBM_ERR_NOMEM, "cannot allocate reduction buffer");
    // the overflow word after the partials; the init kernel clears it
    s->status = (uint32_t*)(s->partials + (kReduceBlocks + 1) * 5);
    hipError_t e = launch_init(s, nullptr, 1);
    if (e == hipSuccess)
        e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        hip_fail(e, "lbm_create: init kernel");
        lbm_destroy(s);
        return LBM_ERR_HIP;
    }
    *out = s;
    return LBM_OK;
}

int lbm_destroy(lbm_solver* s)
{
    if (!s)
        return LBM_OK;
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != s->p.device)
        (void)hipSetDevice(s->p.device);
    (void)hipDeviceSynchronize();
    void* bufs[] = {s->dd,       s->flags,    s->rho,           s->u,
                    s->staging,  s->partials, s->residual_prev, s->dd_alt};
    for (void* b : bufs)
        if (b)
            (void)hipFree(b);
    delete s;
    return LBM_OK;
}

int lbm_get_device_arrays(const lbm_solver* s, lbm_device_arrays* out)
{
    if (!s || !out) {
        set_error("lbm_get_device_arrays: NULL argument");
        return LBM_ERR_ARG;
    }
    out->dd = s->dd;
    out->flags = s->flags;
    out->densities = s->rho;
    out->velocities = s->u;
    out->cells = s->n;
    out->plane_stride = s->stride;
    for (int i = 0; i < 3; i++)
        out->size_h[i] = s->size_h[i];
    out->precision = s->p.precision;
    out->element_bytes = s->elem_bytes;
    return LBM_OK;
}

int lbm_set_launch_options(lbm_solver* s, int kernel, int options)
{
    if (!s || kernel < 0 || kernel > 1 ||
        (options & ~(LBM_LAUNCH_NONTEMPORAL | LBM_LAUNCH_XCD_REMAP | LBM_LAUNCH_WAVES4 |
                     LBM_LAUNCH_WAVES3 | LBM_LAUNCH_WAVES2 | LBM_LAUNCH_ADDR64 |
                     LBM_LAUNCH_WIDE | LBM_LAUNCH_LDS | LBM_LAUNCH_BIG))) {
        set_error("lbm_set_launch_options: invalid argument");
        return LBM_ERR_ARG;
    }
    s->launch[kernel] = options;
    s->launch_set[kernel] = 1;
    return LBM_OK;
}

int lbm_get_launch_options(const lbm_solver* s, int kernel, int* options)
{
    if (!s || !options || kernel < 0 || kernel > 1) {
        set_error("lbm_get_launch_options: invalid argument");
        return LBM_ERR_ARG;
    }
    *options = s->launch[kernel];
    return LBM_OK;
}

int lbm_simulation_step_alpha(lbm_solver* s, const int origin[3], const int size[3],
                              hipStream_t stream)
{
    Box b;
    int rc = resolve_box(s, origin, size, 0, &b);
    if (rc)
        return rc;
    if ((rc = set_device(s)))
        return rc;
    LBM_HIP(launch_alpha(s, b, stream), "alpha kernel launch");
    return LBM_OK;
}

int lbm_simulation_step_beta(lbm_solver* s, const int origin[3], const int size[3],
                             hipStream_t stream)
{
    Box b;
    int rc = resolve_box(s, origin, size, 0, &b);
    if (rc)
        return rc;
    if ((rc = set_device(s)))
        return rc;
    LBM_HIP(launch_beta(s, b, stream), "beta kernel launch");
    return LBM_OK;
}

int lbm_get_density_distributions(lbm_solver* s, const int origin[3], const int size[3],
                                  void* dst, int dst_on_device, hipStream_t stream)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    if (!dst) {
        set_error("destination is NULL");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    return copy_get(s, 0, b, kQ, dst, dst_on_device, stream);
}

int lbm_set_density_distributions(lbm_solver* s, const int origin[3], const int size[3],
                                  const void* src, int src_on_device, hipStream_t stream)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    if (!src) {
        set_error("source is NULL");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    return copy_set(s, 0, b, kQ, src, src_on_device, 0x7FFFFu, stream);
}

int lbm_set_density_distributions_dir(lbm_solver* s, const int origin[3], const int size[3],
                                      int direction, const void* src, int src_on_device,
                                      hipStream_t stream)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    const uint32_t mask = direction_mask(direction);
    if (!mask || !src) {
        set_error("invalid direction or NULL source");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    return copy_set(s, 0, b, kQ, src, src_on_device, mask, stream);
}

int lbm_get_moments(lbm_solver* s, int layout, const int origin[3], const int size[3],
                    void* dst, int dst_on_device, hipStream_t stream)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    if (!dst || layout < 0 || layout > 1) {
        set_error("lbm_get_moments: invalid layout or NULL destination");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    const size_t bytes = (size_t)b.cells() * 4 * s->out_bytes;
    if (dst_on_device) {
        LBM_HIP(launch_moments(s, b, layout, dst, stream), "moments kernel");
        return LBM_OK;
    }
    if ((rc = ensure_staging(s, bytes)))
        return rc;
    LBM_HIP(launch_moments(s, b, layout, s->staging, stream), "moments kernel");
    LBM_HIP(hipMemcpyAsync(dst, s->staging, bytes, hipMemcpyDeviceToHost, stream), "D2H copy");
    LBM_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_get_flags(lbm_solver* s, const int origin[3], const int size[3], int32_t* dst)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    if (!dst) {
        set_error("destination is NULL");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    const size_t bytes = (size_t)b.cells() * sizeof(int32_t);
    if ((rc = ensure_staging(s, bytes)))
        return rc;
    LBM_HIP(launch_flags_get(s, b, (int32_t*)s->staging, nullptr), "flags get kernel");
    LBM_HIP(hipMemcpy(dst, s->staging, bytes, hipMemcpyDeviceToHost), "D2H copy");
    return LBM_OK;
}

int lbm_set_flags(lbm_solver* s, const int origin[3], const int size[3], const int32_t* src)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    if (!src) {
        set_error("source is NULL");
        return LBM_ERR_ARG;
    }
    const uint32_t cells = b.cells();
    for (uint32_t i = 0; i < cells; i++)
        if (!valid_flag(src[i])) {
            set_error("setFlags: value is not one of OBSTACLE, FLUID, VELOCITY_INJECTION, "
                      "GHOST_LAYER");
            return LBM_ERR_ARG;
        }
    if ((rc = set_device(s)))
        return rc;
    const size_t bytes = (size_t)cells * sizeof(int32_t);
    if ((rc = ensure_staging(s, bytes)))
        return rc;
    LBM_HIP(hipMemcpy(s->staging, src, bytes, hipMemcpyHostToDevice), "H2D copy");
    LBM_HIP(launch_flags_set(s, b, (const int32_t*)s->staging, nullptr), "flags set kernel");
    LBM_HIP(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    return LBM_OK;
}

static int field_get(lbm_solver* s, int field, const int origin[3], const int size[3],
                     void* dst, int on_dev, hipStream_t st)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    if ((field == 1 && !s->u) || (field == 2 && !s->rho)) {
        set_error("field is not stored by this solver (store_velocities/store_densities)");
        return LBM_ERR_UNSUPPORTED;
    }
    if (!dst) {
        set_error("destination is NULL");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    return copy_get(s, field, b, field == 1 ? 3 : 1, dst, on_dev, st);
}

static int field_set(lbm_solver* s, int field, const int origin[3], const int size[3],
                     const void* src, int on_dev, hipStream_t st)
{
    Box b;
    int rc = resolve_box(s, origin, size, 1, &b);
    if (rc)
        return rc;
    if ((field == 1 && !s->u) || (field == 2 && !s->rho)) {
        set_error("field is not stored by this solver (store_velocities/store_densities)");
        return LBM_ERR_UNSUPPORTED;
    }
    if (!src) {
        set_error("source is NULL");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    return copy_set(s, field, b, field == 1 ? 3 : 1, src, on_dev, 0x7u, st);
}

int lbm_get_velocities(lbm_solver* s, const int origin[3], const int size[3], void* dst,
                       int dst_on_device, hipStream_t stream)
{
    return field_get(s, 1, origin, size, dst, dst_on_device, stream);
}
int lbm_set_velocities(lbm_solver* s, const int origin[3], const int size[3], const void* src,
                       int src_on_device, hipStream_t stream)
{
    return field_set(s, 1, origin, size, src, src_on_device, stream);
}
int lbm_get_densities(lbm_solver* s, const int origin[3], const int size[3], void* dst,
                      int dst_on_device, hipStream_t stream)
{
    return field_get(s, 2, origin, size, dst, dst_on_device, stream);
}
int lbm_set_densities(lbm_solver* s, const int origin[3], const int size[3], const void* src,
                      int src_on_device, hipStream_t stream)
{
    return field_set(s, 2, origin, size, src, src_on_device, stream);
}

// CComm send/recv boxes, CManager.cpp:137-202.
int lbm_face_box(const lbm_solver* s, int direction, int which, int origin[3], int size[3])
{
    if (!s || !origin || !size || direction < 0 || direction > 5 || which < 0 || which > 1) {
        set_error("lbm_face_box: invalid argument");
        return LBM_ERR_ARG;
    }
    const int axis = direction / 2;
    const bool high = direction & 1;
    for (int i = 0; i < 3; i++) {
        origin[i] = 0;
        size[i] = s->size_h[i];
    }
    size[axis] = 1;
    if (which == 0)
        origin[axis] = high ? s->size_h[axis] - 2 : 1;
    else
        origin[axis] = high ? s->size_h[axis] - 1 : 0;
    return LBM_OK;
}

int64_t lbm_face_elements(const lbm_solver* s, int direction)
{
    if (!s || direction < 0 || direction > 5)
        return 0;
    const int axis = direction / 2;
    int64_t cells = 1;
    for (int i = 0; i < 3; i++)
        if (i != axis)
            cells *= s->size_h[i];
    return cells * kQ;
}

int lbm_pack_face(lbm_solver* s, int direction, int phase, void* dev_buf, hipStream_t stream)
{
    int o[3], z[3];
    if (phase < 0 || phase > 1 || !dev_buf) {
        set_error("lbm_pack_face: invalid phase or NULL buffer");
        return LBM_ERR_ARG;
    }
    int rc = lbm_face_box(s, direction, phase == 0 ? 0 : 1, o, z);
    if (rc)
        return rc;
    if ((rc = set_device(s)))
        return rc;
    Box b{o[0], o[1], o[2], z[0], z[1], z[2]};
    LBM_HIP(launch_box_get(s, 0, b, dev_buf, exchange_mask(direction, phase), stream),
            "pack kernel");
    return LBM_OK;
}

int lbm_unpack_face(lbm_solver* s, int direction, int phase, const void* dev_buf,
                    hipStream_t stream)
{
    int o[3], z[3];
    if (phase < 0 || phase > 1 || !dev_buf) {
        set_error("lbm_unpack_face: invalid phase or NULL buffer");
        return LBM_ERR_ARG;
    }
    int rc = lbm_face_box(s, direction, phase == 0 ? 1 : 0, o, z);
    if (rc)
        return rc;
    if ((rc = set_device(s)))
        return rc;
    Box b{o[0], o[1], o[2], z[0], z[1], z[2]};
    // `direction` is the receiving face: the sender's is the opposite one
    LBM_HIP(launch_box_set(s, 0, b, dev_buf, exchange_mask(direction ^ 1, phase), stream),
            "unpack kernel");
    return LBM_OK;
}

int lbm_unpack_face_pair(lbm_solver* s, int axis, int phase, const void* lo_buf,
                         const void* hi_buf, hipStream_t stream)
{
    if (!s || axis < 0 || axis > 2 || phase < 0 || phase > 1 || !lo_buf || !hi_buf) {
        set_error("lbm_unpack_face_pair: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc;
    if (axis != 0) {
        // y / z: both faces' boxes in one launch (disjoint planes of the array,
        // or one plane with disjoint lattice vectors when the axis is one cell)
        Box b[2];
        uint32_t mask[2];
        for (int k = 0; k < 2; k++) {
            int o[3], z[3];
            if ((rc = lbm_face_box(s, 2 * axis + k, phase == 0 ? 1 : 0, o, z)))
                return rc;
            b[k] = Box{o[0], o[1], o[2], z[0], z[1], z[2]};
            mask[k] = exchange_mask((2 * axis + k) ^ 1, phase);
        }
        if ((rc = set_device(s)))
            return rc;
        const void* src[2] = {lo_buf, hi_buf};
        LBM_HIP(launch_dd_set_pair(s, b, src, mask, stream), "pair unpack kernel");
        return LBM_OK;
    }
    if ((rc = set_device(s)))
        return rc;
    // the boxes of lbm_face_box for LEFT / RIGHT
    const int nx = s->size_h[0];
    const int x_lo = phase == 0 ? 0 : 1, x_hi = phase == 0 ? nx - 1 : nx - 2;
    // the LEFT face receives from a RIGHT face, and vice versa; the two faces'
    // planes are disjoint, so their stores share no cache line: streaming
    const uint32_t mask_lo = exchange_mask(LBM_RIGHT, phase);
    const uint32_t mask_hi = exchange_mask(LBM_LEFT, phase);
    LBM_HIP(launch_set_dd_xpair(s, lo_buf, hi_buf, x_lo, x_hi, mask_lo, mask_hi, true, stream),
            "pair unpack kernel");
    return LBM_OK;
}

int lbm_pack_face_pair(lbm_solver* s, int axis, int phase, void* lo_buf, void* hi_buf,
                       hipStream_t stream)
{
    if (!s || axis < 0 || axis > 2 || phase < 0 || phase > 1 || !lo_buf || !hi_buf) {
        set_error("lbm_pack_face_pair: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc;
    Box b[2];
    uint32_t mask[2];
    for (int k = 0; k < 2; k++) {
        int o[3], z[3];
        if ((rc = lbm_face_box(s, 2 * axis + k, phase == 0 ? 0 : 1, o, z)))
            return rc;
        b[k] = Box{o[0], o[1], o[2], z[0], z[1], z[2]};
        mask[k] = exchange_mask(2 * axis + k, phase);
    }
    if ((rc = set_device(s)))
        return rc;
    void* dst[2] = {lo_buf, hi_buf};
    LBM_HIP(launch_dd_get_pair(s, b, dst, mask, stream), "pair pack kernel");
    return LBM_OK;
}

// Launch options the caller has not set follow the precision's defaults for
// the solver's size and fused face work.
static void refresh_launch_defaults(lbm_solver* s)
{
    const bool with_pack = s->xpack[0] || s->xpack[1];
    const bool with_unpack = s->xunpack[0] || s->xunpack[1];
    for (int k = 0; k < 2; k++)
        if (!s->launch_set[k])
            s->launch[k] = default_launch(s->p.precision, k, with_pack, s->n >= kWideCells,
                                          with_unpack, !plane_off32(s));
}

int lbm_set_step_pack(lbm_solver* s, void* left_buf, void* right_buf)
{
    if (!s) {
        set_error("lbm_set_step_pack: solver is NULL");
        return LBM_ERR_ARG;
    }
    s->xpack[0] = left_buf;
    s->xpack[1] = right_buf;
    refresh_launch_defaults(s);
    return LBM_OK;
}

int lbm_set_step_unpack(lbm_solver* s, const void* left_buf, const void* right_buf)
{
    if (!s) {
        set_error("lbm_set_step_unpack: solver is NULL");
        return LBM_ERR_ARG;
    }
    if ((left_buf || right_buf) && s->size_h[0] < 4) {
        // the kernels take each send-plane cell as one face's
        set_error("lbm_set_step_unpack: the x extent must be at least 2 interior cells");
        return LBM_ERR_UNSUPPORTED;
    }
    s->xunpack[0] = left_buf;
    s->xunpack[1] = right_buf;
    refresh_launch_defaults(s);
    return LBM_OK;
}

// lbm_pack_face(from, direction) followed by lbm_unpack_face(to, direction ^ 1)
// as one buffer-less transfer: the same boxes, planes and conversions
static int exchange_desc(const char* fn, lbm_solver* from, int direction, int phase,
                         lbm_solver* to, XferDesc* d)
{
    if (!from || !to || direction < 0 || direction > 5 || phase < 0 || phase > 1) {
        set_error(std::string(fn) + ": invalid argument");
        return LBM_ERR_ARG;
    }
    if (from->p.device != to->p.device || from->p.precision != to->p.precision) {
        set_error(std::string(fn) + ": solvers differ in device or precision");
        return LBM_ERR_ARG;
    }
    const int axis = direction / 2;
    for (int i = 0; i < 3; i++)
        if (i != axis && from->size_h[i] != to->size_h[i]) {
            set_error(std::string(fn) + ": face extents differ");
            return LBM_ERR_ARG;
        }
    int so[3], sz[3], ro[3], rz[3];
    int rc;
    if ((rc = lbm_face_box(from, direction, phase == 0 ? 0 : 1, so, sz)) ||
        (rc = lbm_face_box(to, direction ^ 1, phase == 0 ? 1 : 0, ro, rz)))
        return rc;
    d->from = from;
    d->sb = Box{so[0], so[1], so[2], sz[0], sz[1], sz[2]};
    d->to = to;
    d->db = Box{ro[0], ro[1], ro[2], rz[0], rz[1], rz[2]};
    d->mask = exchange_mask(direction, phase);
    return LBM_OK;
}

int lbm_exchange_face(lbm_solver* from, int direction, int phase, lbm_solver* to,
                      hipStream_t stream)
{
    XferDesc d;
    int rc = exchange_desc("lbm_exchange_face", from, direction, phase, to, &d);
    if (rc || (rc = set_device(from)))
        return rc;
    LBM_HIP(launch_dd_xfer(&d, 1, stream), "exchange kernel");
    return LBM_OK;
}

int lbm_exchange_faces(const lbm_face_transfer* transfers, int n, int phase,
                       hipStream_t stream)
{
    if (n < 0 || (n > 0 && !transfers)) {
        set_error("lbm_exchange_faces: invalid argument");
        return LBM_ERR_ARG;
    }
    if (n == 0)
        return LBM_OK;
    std::vector<XferDesc> list((size_t)n);
    std::vector<std::pair<uintptr_t, int>> written((size_t)n);
    for (int k = 0; k < n; k++) {
        const lbm_face_transfer& t = transfers[k];
        int rc = exchange_desc("lbm_exchange_faces", t.from, t.direction, phase, t.to, &list[k]);
        if (rc)
            return rc;
        const lbm_face_transfer& t0 = transfers[0];
        if (t.direction / 2 != t0.direction / 2 || t.from->p.device != t0.from->p.device ||
            t.from->p.precision != t0.from->p.precision) {
            set_error("lbm_exchange_faces: transfers differ in axis, device or precision");
            return LBM_ERR_ARG;
        }
        written[k] = {(uintptr_t)t.to, t.direction};
    }
    // one axis: sends read the planes one side of the faces, receives write
    // the others, so only two transfers into the same face collide
    std::sort(written.begin(), written.end());
    if (std::adjacent_find(written.begin(), written.end()) != written.end()) {
        set_error("lbm_exchange_faces: two transfers write the same face");
        return LBM_ERR_ARG;
    }
    int rc = set_device(transfers[0].from);
    if (rc)
        return rc;
    LBM_HIP(launch_dd_xfer(list.data(), n, stream), "exchange kernel");
    return LBM_OK;
}

int lbm_device_alloc(int device, size_t bytes, void** out)
{
    if (!out || !bytes) {
        set_error("lbm_device_alloc: invalid argument");
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    int cur = -1;
    LBM_HIP(hipGetDevice(&cur), "hipGetDevice");
    if (cur != device)
        LBM_HIP(hipSetDevice(device), "hipSetDevice");
    if (hipMalloc(out, bytes) != hipSuccess) {
        set_error("lbm_device_alloc: out of device memory");
        return LBM_ERR_NOMEM;
    }
    return LBM_OK;
}

int lbm_device_free(int device, void* ptr)
{
    if (!ptr)
        return LBM_OK;
    int cur = -1;
    LBM_HIP(hipGetDevice(&cur), "hipGetDevice");
    if (cur != device)
        LBM_HIP(hipSetDevice(device), "hipSetDevice");
    LBM_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    LBM_HIP(hipFree(ptr), "hipFree");
    return LBM_OK;
}

int lbm_reduce_box(lbm_solver* s, const int origin[3], const int size[3], double out[5],
                   hipStream_t stream)
{
    Box b;
    int rc = resolve_box(s, origin, size, 0, &b);
    if (rc)
        return rc;
    if (!out) {
        set_error("output is NULL");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    double* res = s->partials + kReduceBlocks * 5;
    LBM_HIP(launch_reduce_box(s, b, s->partials, res, stream), "reduce kernel");
    LBM_HIP(hipMemcpyAsync(out, res, 5 * sizeof(double), hipMemcpyDeviceToHost, stream),
            "D2H copy");
    LBM_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_reduce_boxes_noalloc(lbm_solver* const* solvers, const int* origins, const int* sizes,
                             int count, double* out_dev, hipStream_t stream)
{
    if (!solvers || !out_dev) {
        set_error("lbm_reduce_boxes_noalloc: solvers or out_dev is NULL");
        return LBM_ERR_ARG;
    }
    if (count < 1 || count > LBM_NOALLOC_MAX_BOXES) {
        char buf[160];
        snprintf(buf, sizeof buf,
                 "lbm_reduce_boxes_noalloc: count %d outside 1..%d (LBM_NOALLOC_MAX_BOXES, the "
                 "size of the static partials array)",
                 count, LBM_NOALLOC_MAX_BOXES);
        set_error(buf);
        return LBM_ERR_ARG;
    }
    Box boxes[LBM_NOALLOC_MAX_BOXES];
    for (int m = 0; m < count; m++) {
        const lbm_solver* s = solvers[m];
        if (!s) {
            set_error("lbm_reduce_boxes_noalloc: a solver is NULL");
            return LBM_ERR_ARG;
        }
        if (s->p.device != solvers[0]->p.device || s->p.precision != solvers[0]->p.precision) {
            set_error("lbm_reduce_boxes_noalloc: solvers differ in device or precision");
            return LBM_ERR_ARG;
        }
        const int* o = origins ? origins + 3 * m : nullptr;
        const int* z = sizes ? sizes + 3 * m : nullptr;
        int rc = resolve_box(s, o, z, 0, &boxes[m]);
        if (rc)
            return rc;
    }
    int rc = set_device(solvers[0]);
    if (rc)
        return rc;
    double* partials = nullptr;
    LBM_HIP(noalloc_scratch(solvers[0]->p.device, &partials), "hipGetSymbolAddress(scratch)");
    LBM_HIP(launch_reduce_boxes(solvers, boxes, count, partials, out_dev, stream),
            "reduce kernel");
    return LBM_OK;
}

int lbm_residual(lbm_solver* s, int layout, double* residual, hipStream_t stream)
{
    if (!s || !residual || layout < 0 || layout > 1) {
        set_error("lbm_residual: invalid argument");
        return LBM_ERR_ARG;
    }
    int rc = set_device(s);
    if (rc)
        return rc;
    if (!s->residual_prev) {
        const size_t bytes = (size_t)3 * s->stride * s->out_bytes;
        if (hipMalloc(&s->residual_prev, bytes) != hipSuccess) {
            set_error("cannot allocate residual snapshot");
            return LBM_ERR_NOMEM;
        }
        LBM_HIP(hipMemset(s->residual_prev, 0, bytes), "hipMemset");
        s->residual_valid = 0;
    }
    double* res = s->partials + kReduceBlocks * 5;
    double host[5];
    LBM_HIP(launch_residual(s, layout, s->partials, res, stream), "residual kernel");
    LBM_HIP(hipMemcpyAsync(host, res, 5 * sizeof(double), hipMemcpyDeviceToHost, stream),
            "D2H copy");
    LBM_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    *residual = first_residual(s->residual_valid, host[0], host[1]);
    s->residual_valid = 1;
    return LBM_OK;
}

int lbm_reset(lbm_solver* s)
{
    if (!s) {
        set_error("lbm_reset: solver is NULL");
        return LBM_ERR_ARG;
    }
    int rc = set_device(s);
    if (rc)
        return rc;
    LBM_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    LBM_HIP(launch_init(s, nullptr, 0), "init kernel");
    LBM_HIP(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    s->residual_valid = 0;
    return LBM_OK;
}

int lbm_state_bytes(const lbm_solver* s, uint64_t* bytes)
{
    if (!s || !bytes) {
        set_error("lbm_state_bytes: NULL argument");
        return LBM_ERR_ARG;
    }
    *bytes = (uint64_t)s->n * ((uint64_t)kQ * s->elem_bytes + 1);
    return LBM_OK;
}

namespace {

// Copies between the padded device planes and a dense host state buffer.
int state_copy(lbm_solver* s, void* host, uint64_t bytes, bool save)
{
    uint64_t need = 0;
    int rc = lbm_state_bytes(s, &need);
    if (rc)
        return rc;
    if (!host || bytes != need) {
        set_error("state copy: buffer size does not match lbm_state_bytes");
        return LBM_ERR_ARG;
    }
    if ((rc = set_device(s)))
        return rc;
    const size_t plane = (size_t)s->n * s->elem_bytes;
    const size_t pitch = (size_t)s->stride * s->elem_bytes;
    char* h = (char*)host;
    char* d = (char*)s->dd;
    if (save) {
        LBM_HIP(hipMemcpy2D(h, plane, d, pitch, plane, kQ, hipMemcpyDeviceToHost), "D2H copy");
        LBM_HIP(hipMemcpy(h + kQ * plane, s->flags, (size_t)s->n, hipMemcpyDeviceToHost),
                "D2H copy");
    } else {
        LBM_HIP(hipMemcpy2D(d, pitch, h, plane, plane, kQ, hipMemcpyHostToDevice), "H2D copy");
        LBM_HIP(hipMemcpy(s->flags, h + kQ * plane, (size_t)s->n, hipMemcpyHostToDevice),
                "H2D copy");
        // a saved state was stored by the same formats, hence in range
        LBM_HIP(hipMemset(s->status, 0, sizeof(uint32_t)), "hipMemset");
        s->residual_valid = 0;
    }
    return LBM_OK;
}

}  // namespace

int lbm_save_state(lbm_solver* s, void* dst, uint64_t bytes)
{
    if (!s) {
        set_error("lbm_save_state: solver is NULL");
        return LBM_ERR_ARG;
    }
    int rc = set_device(s);
    if (rc)
        return rc;
    LBM_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return state_copy(s, dst, bytes, true);
}

int lbm_load_state(lbm_solver* s, const void* src, uint64_t bytes)
{
    if (!s) {
        set_error("lbm_load_state: solver is NULL");
        return LBM_ERR_ARG;
    }
    int rc = set_device(s);
    if (rc)
        return rc;
    LBM_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return state_copy(s, const_cast<void*>(src), bytes, false);
}

int lbm_device_count(int* count)
{
    if (!count) {
        set_error("lbm_device_count: NULL argument");
        return LBM_ERR_ARG;
    }
    int n = 0;
    LBM_HIP(hipGetDeviceCount(&n), "hipGetDeviceCount");
    *count = n;
    return LBM_OK;
}

int lbm_sync(lbm_solver* s, hipStream_t stream)
{
    if (!s) {
        set_error("solver is NULL");
        return LBM_ERR_ARG;
    }
    int rc = set_device(s);
    if (rc)
        return rc;
    LBM_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return LBM_OK;
}

}  // extern "C"
