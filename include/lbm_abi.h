/*
 * C ABI of the MI355X-native D3Q19 alpha/beta A-A solver (libcavityflow_hip.so).
 *
 * This is the drop-in boundary for the hot path of saleemayman/cavityflow_cuda:
 * the CLbmSolver<T> operator surface (CLbmSolver.hpp:68-88, CLbmSolverGPU.cuh:56-97)
 * driven by CController / CManager. Each entry point below names the reference
 * method it replaces. Plain pointers and sizes only.
 *
 * Conventions (the reference's, CLbmSolverCPU.cpp:142-244, CLbmSolverGPU.cu:516-1328):
 *  - A solver owns one subdomain of size[3] interior cells plus a one-cell halo
 *    on each face: size_h = size + 2 per axis.
 *  - Boxes (origin[3], box[3]) are in halo coordinates, x fastest, and must
 *    satisfy origin >= 0, box > 0, origin + box <= size_h. A NULL origin means
 *    the reference's whole-domain overload: the full halo-inclusive domain for
 *    the steps (CLbmSolverCPU.cpp:270-281), the interior origin (1,1,1),
 *    box = size for the copies (CLbmSolverCPU.cpp:552-559).
 *  - Copy buffers are dense [component][z][y][x] over the box: 19 components
 *    for populations, 3 for velocities, 1 for densities and flags.
 *  - Population/velocity/density buffers hold `double` for LBM_F64 and `float`
 *    for every other precision (shifted 16-bit storage is converted).
 *  - Errors are returned as negative lbm_status codes; nothing calls exit().
 *    lbm_last_error() holds a message for the calling thread.
 *  - Streams: every call is ordered on `stream` (NULL = the default stream).
 *    Calls with host buffers are synchronous; calls with device buffers and
 *    the steps are asynchronous, as the reference GPU solver's are.
 *  - One solver is used from one host thread at a time.
 */
#ifndef CAVITYFLOW_LBM_ABI_H
#define CAVITYFLOW_LBM_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HIP's own opaque stream handle type (hip_runtime_api.h). */
typedef struct ihipStream_t* hipStream_t;

typedef struct lbm_solver lbm_solver;

/* enum Flag, common.h:54-61 */
enum lbm_flag {
    LBM_OBSTACLE = 1,
    LBM_FLUID = 2,
    LBM_VELOCITY_INJECTION = 4,
    LBM_GHOST_LAYER = 8
};

/* enum Direction, common.h:63-71 */
enum lbm_direction {
    LBM_LEFT = 0,
    LBM_RIGHT = 1,
    LBM_BOTTOM = 2,
    LBM_TOP = 3,
    LBM_BACK = 4,
    LBM_FRONT = 5
};

/* Population storage. F32 and F64 reproduce the reference bit for bit;
 * the 16-bit formats store x = f_l - w_l (w_l the lattice weight) with FP32
 * collision arithmetic:
 *   F16S : IEEE binary16 of x * 2^12;
 *   BF16S: bfloat16 of x;
 *   F16C : 1 sign, 4 exponent, 11 significand bits of x, normal range
 *          [2^-17, 2^-2) and subnormal step 2^-28 (FP16C of Lehmann et al.,
 *          Phys. Rev. E 106, 015308, 2022). It has no infinities or NaN: a
 *          store that does not fit (|x| >= 1/4 after rounding, or a
 *          non-finite value) marks the solver, after which lbm_residual,
 *          lbm_reduce_box and lbm_get_moments return NaN until lbm_reset or
 *          lbm_load_state. */
enum lbm_precision {
    LBM_F32 = 0,
    LBM_F64 = 1,
    LBM_F16S = 2,
    LBM_BF16S = 3,
    LBM_F16C = 4
};

enum lbm_status {
    LBM_OK = 0,
    LBM_ERR_ARG = -1,         /* bad argument (NULL, bad enum, bad flag value) */
    LBM_ERR_BOX = -2,         /* box outside the halo-inclusive domain (reference: assert) */
    LBM_ERR_HIP = -3,         /* a HIP runtime call failed */
    LBM_ERR_NOMEM = -4,       /* device allocation failed */
    LBM_ERR_PARAMS = -5,      /* no valid time step (reference: exit, CLbmSolver.cpp:263-272) */
    LBM_ERR_STATE = -6,       /* operation not available in this solver state */
    LBM_ERR_UNSUPPORTED = -7  /* e.g. outputs requested but not stored */
};

/* Construction parameters of one subdomain solver: the quantities the
 * CLbmSolver constructor derives (CLbmSolver.cpp:28-281) and the
 * CLbmSolverGPU constructor consumes (CLbmSolverGPU.cu:31-161). */
typedef struct lbm_params {
    int size[3];            /* CDomain::getSize(), interior cells per axis (>= 1) */
    int bc[6];              /* boundaryConditions: LEFT, RIGHT, BOTTOM, TOP, BACK, FRONT */
    int precision;          /* enum lbm_precision */
    double inv_tau;         /* CLbmSolver::tauInv, rounded to the arithmetic type */
    double lid_velocity;    /* CLbmSolver::velocityDimLess[0] */
    double gravitation[3];  /* CLbmSolver::accelerationDimLess */
    int store_densities;    /* CLbmSolver::storeDensities */
    int store_velocities;   /* CLbmSolver::storeVelocities */
    int device;             /* HIP device ordinal (reference: id % device count) */
} lbm_params;

/* Input of the CLbmSolver constructor's time-step adaptation. */
typedef struct lbm_physics {
    double length_x;        /* CDomain::getLength()[0] of this solver's domain */
    int size_x;             /* CDomain::getSize()[0] */
    double length_max;      /* CConfiguration::domainLength.max() */
    double velocity[3];     /* CConfiguration::velocity */
    double acceleration[3]; /* CConfiguration::acceleration */
    double viscosity;       /* <= 0 selects the default-Reynolds viscosity */
    double timestep;
    double max_velocity;     /* CConfiguration::maxVelocityDimLess */
    double max_acceleration; /* CConfiguration::maxAccelerationDimLess */
} lbm_physics;

typedef struct lbm_derived {
    double velocity_dimless[3];
    double acceleration_dimless[3];
    double viscosity;
    double viscosity_dimless;
    double tau;
    double inv_tau;
    double timestep;
} lbm_derived;

/* Device arrays of a solver, for callers that fuse their own kernels. */
typedef struct lbm_device_arrays {
    void* dd;            /* populations, plane l at dd + l * plane_stride elements */
    uint8_t* flags;      /* one byte per cell */
    void* densities;     /* NULL unless store_densities */
    void* velocities;    /* component c at velocities + c * plane_stride; NULL unless stored */
    int64_t cells;       /* N_h = size_h[0] * size_h[1] * size_h[2] */
    int64_t plane_stride;/* elements between population planes (>= cells, 256-aligned) */
    int size_h[3];
    int precision;
    int element_bytes;   /* bytes per stored population */
} lbm_device_arrays;

/* ---- lifecycle ---------------------------------------------------------- */

/* CLbmSolver<T>::CLbmSolver time-step adaptation (CLbmSolver.cpp:75-281),
 * in float for every precision except LBM_F64 (double). Returns LBM_OK or
 * LBM_ERR_PARAMS where the reference exits. */
int lbm_compute_parameters(const lbm_physics* in, int precision, lbm_derived* out);

/* CLbmSolverGPU<T>::CLbmSolverGPU (CLbmSolverGPU.cu:31-161): selects the
 * device, allocates the SoA arrays, runs the init kernel (lbm_init.cu:41-208:
 * flags from the six face conditions, x fastest; populations at rest
 * equilibrium; u = 0, rho = 1). Synchronous. */
int lbm_create(const lbm_params* params, lbm_solver** out);

/* CLbmSolverGPU<T>::~CLbmSolverGPU (CLbmSolverGPU.cu:163-172). */
int lbm_destroy(lbm_solver* solver);

int lbm_get_device_arrays(const lbm_solver* solver, lbm_device_arrays* out);

/* ---- time steps --------------------------------------------------------- */

/* simulationStepAlpha() / simulationStepAlpha(origin, size[, stream])
 * (CLbmSolver.hpp:68-69, CLbmSolverGPU.cu:174-338; CPU arithmetic
 * CLbmAlphaCPU.cpp:40-502). Even A-A step, in place. */
int lbm_simulation_step_alpha(lbm_solver* solver, const int origin[3], const int size[3],
                              hipStream_t stream);

/* simulationStepBeta() / simulationStepBeta(origin, size[, stream])
 * (CLbmSolver.hpp:70-71, CLbmSolverGPU.cu:340-514; CPU arithmetic
 * CLbmBetaCPU.cpp:52-459). Odd A-A step, in place, linear-index wrap. */
int lbm_simulation_step_beta(lbm_solver* solver, const int origin[3], const int size[3],
                             hipStream_t stream);

/* Launch variants of the step kernels; they change speed, never results.
 * kernel: 0 = alpha, 1 = beta. Defaults (measured on MI355X): F32 beta
 * WIDE | WAVES2 on domains of 6 Mi cells or more (halo included: from 183^3
 * interior cells), BIG | WAVES4 there with a step pack or unpack
 * (lbm_set_step_pack / lbm_set_step_unpack) except WIDE | WAVES3 with an
 * unpack on domains whose planes exceed 4 GiB, and WAVES2 below; F32 alpha
 * NONTEMPORAL | WAVES2 plus WIDE on those large domains, NONTEMPORAL | WAVES3
 * with a step unpack, or with a step pack on smaller domains; F64 alpha
 * NONTEMPORAL | WAVES2, beta LDS | WAVES2; F16S / BF16S alpha none, beta BIG
 * on those large domains without a step unpack (else none); F16C none.
 * Batches start from F32 / F64 alpha and beta as for 256-thread workgroups
 * with a step pack, 16-bit alpha NONTEMPORAL and beta none. No reference
 * counterpart: the reference takes its block shapes from the XML
 * configuration (default.xml:93-115). */
enum lbm_launch_option {
    LBM_LAUNCH_NONTEMPORAL = 1,  /* nontemporal population loads/stores */
    LBM_LAUNCH_XCD_REMAP = 2,    /* one contiguous range of cells per XCD */
    LBM_LAUNCH_WAVES4 = 4,       /* at most 4 waves per SIMD (LDS reserved per
                                    workgroup): fewer concurrent plane streams */
    LBM_LAUNCH_WAVES3 = 8,       /* at most 3 waves per SIMD (wins over WAVES4) */
    LBM_LAUNCH_WAVES2 = 16,      /* at most 2 waves per SIMD (wins over both) */
    LBM_LAUNCH_ADDR64 = 32,      /* 64-bit element indexing, as domains whose planes
                                    exceed 4 GiB always use (single-domain kernels
                                    only; results identical) */
    LBM_LAUNCH_WIDE = 64,        /* 512-thread workgroups, two cells per thread
                                    (1024 consecutive cells per workgroup; with it
                                    WAVES3 caps at 2 waves per SIMD, as WAVES2;
                                    single-domain kernels only) */
    LBM_LAUNCH_LDS = 128,        /* beta of the whole domain stages its shifted
                                    plane ranges through LDS with line-aligned
                                    accesses (wins over WIDE and BIG; alpha and box
                                    steps ignore it) */
    LBM_LAUNCH_BIG = 256         /* 1024-thread workgroups, one cell per thread
                                    (wins over WIDE; a workgroup alone fills 4
                                    waves per SIMD, so any WAVES cap holds one per
                                    CU; single-domain kernels only; beta spills
                                    registers in this shape for F64, and with a
                                    step unpack and 64-bit indexing for the
                                    other formats) */
};
int lbm_set_launch_options(lbm_solver* solver, int kernel, int options);
int lbm_get_launch_options(const lbm_solver* solver, int kernel, int* options);

/* ---- box copies (halo exchange, checkpoint, extraction) ------------------ */

/* getDensityDistributions(origin, size, dst[, stream]) / (dst)
 * (CLbmSolver.hpp:73-74, CLbmSolverGPU.cu:516-611, CLbmSolverCPU.cpp:497-559). */
int lbm_get_density_distributions(lbm_solver* solver, const int origin[3], const int size[3],
                                  void* dst, int dst_on_device, hipStream_t stream);

/* setDensityDistributions(origin, size, src[, stream]) / (src)
 * (CLbmSolver.hpp:75-76, CLbmSolverGPU.cu:723-820, CLbmSolverCPU.cpp:643-706). */
int lbm_set_density_distributions(lbm_solver* solver, const int origin[3], const int size[3],
                                  const void* src, int src_on_device, hipStream_t stream);

/* setDensityDistributions(origin, size, direction, src[, stream])
 * (CLbmSolverCPU.hpp:80, CLbmSolverGPU.cu:613-721, CLbmSolverCPU.cpp:562-641):
 * copies only the lattice vectors l with norm(direction) . c_l > 0. */
int lbm_set_density_distributions_dir(lbm_solver* solver, const int origin[3],
                                      const int size[3], int direction, const void* src,
                                      int src_on_device, hipStream_t stream);

/* getFlags / setFlags (CLbmSolver.hpp:77-80, CLbmSolverGPU.cu:822-986). Host
 * buffers of int32 Flag values; setFlags accepts only the four Flag values. */
int lbm_get_flags(lbm_solver* solver, const int origin[3], const int size[3], int32_t* dst);
int lbm_set_flags(lbm_solver* solver, const int origin[3], const int size[3],
                  const int32_t* src);

/* getVelocities / setVelocities (CLbmSolver.hpp:81-84, CLbmSolverGPU.cu:988-1160,
 * CLbmSolverCPU.cpp:833-960). LBM_ERR_UNSUPPORTED unless store_velocities. */
int lbm_get_velocities(lbm_solver* solver, const int origin[3], const int size[3], void* dst,
                       int dst_on_device, hipStream_t stream);
int lbm_set_velocities(lbm_solver* solver, const int origin[3], const int size[3],
                       const void* src, int src_on_device, hipStream_t stream);

/* getDensities / setDensities (CLbmSolver.hpp:85-88, CLbmSolverGPU.cu:1162-1328,
 * CLbmSolverCPU.cpp:962-1084). LBM_ERR_UNSUPPORTED unless store_densities. */
int lbm_get_densities(lbm_solver* solver, const int origin[3], const int size[3], void* dst,
                      int dst_on_device, hipStream_t stream);
int lbm_set_densities(lbm_solver* solver, const int origin[3], const int size[3],
                      const void* src, int src_on_device, hipStream_t stream);

/* Density and momentum computed from the current populations (no stored
 * outputs needed), in the reference accumulation order (CLbmAlphaCPU.cpp:87-182):
 * dst is dense [rho, jx, jy, jz][z][y][x] over the box (NULL origin: interior).
 * layout 0: populations in natural slots (after beta or init);
 * layout 1: swapped slots (after alpha). The velocity of the reference's
 * incompressible form is j itself. Complements getVelocities/getDensities,
 * which return what the last step stored. */
int lbm_get_moments(lbm_solver* solver, int layout, const int origin[3], const int size[3],
                    void* dst, int dst_on_device, hipStream_t stream);

/* ---- halo exchange faces (CComm shapes, CManager.cpp:135-202) ------------ */

/* Face boxes of the reference's CComm descriptors for face `direction`:
 * which = 0 -> sendOrigin/sendSize (plane at offset 1 from the face),
 * which = 1 -> recvOrigin/recvSize (the halo plane). Both span the full
 * halo-inclusive extent of the other two axes. */
int lbm_face_box(const lbm_solver* solver, int direction, int which, int origin[3],
                 int size[3]);

/* Elements (19 per cell) of one face buffer for `direction`. */
int64_t lbm_face_elements(const lbm_solver* solver, int direction);

/* Pack one face into a dense device buffer [l][z][y][x], the device-side
 * half of the exchange in CController::stepAlpha / stepBeta
 * (CController.cpp:369-570, 807-1008). Each phase moves the 5 planes the
 * receiver's interior depends on; the buffer keeps the 19-plane layout and
 * the other 14 planes are not written:
 *   phase 0 (after alpha): the send plane (offset 1), the planes whose lattice
 *            vectors point back into this solver through the face (the
 *            reference sends all 19, CController.cpp:380-387);
 *   phase 1 (after beta) : the halo plane, the planes whose vectors point out
 *            through the face, i.e. the populations the neighbour's
 *            direction-filtered unpack applies (the reference sends all 19
 *            and filters on receipt, :818-825, 849-874).
 * In the beta step plane l at cell p is read and written only by cell
 * p - c_l, so the other 14 planes of a halo plane never reach an interior
 * cell before the phase-1 unpack overwrites where they went: every interior
 * cell is bit-identical to the reference's 19-plane exchange. */
int lbm_pack_face(lbm_solver* solver, int direction, int phase, void* dev_buf,
                  hipStream_t stream);

/* Unpack a buffer packed by the neighbour across face `direction` (the
 * receiving face; the sender packed the opposite one), the same 5 planes:
 *   phase 0: into the halo plane, the planes whose vectors point out through
 *            the face (CController.cpp:411-435 writes all 19);
 *   phase 1: into the send plane, the populations pointing into the domain
 *            (setDensityDistributions with direction, :849-874). */
int lbm_unpack_face(lbm_solver* solver, int direction, int phase, const void* dev_buf,
                    hipStream_t stream);

/* Both faces of one axis (0 = x, 1 = y, 2 = z) in one launch: the same result
 * as lbm_unpack_face(2 * axis, phase, lo_buf) followed by
 * lbm_unpack_face(2 * axis + 1, phase, hi_buf). For x, the launch writes the
 * two faces row by row. No reference counterpart: the reference unpacks each
 * face with its own setDensityDistributions call (CController.cpp:411-435,
 * 849-874). */
int lbm_unpack_face_pair(lbm_solver* solver, int axis, int phase, const void* lo_buf,
                         const void* hi_buf, hipStream_t stream);

/* Both faces of one axis packed in one launch: the same buffers as
 * lbm_pack_face(2 * axis, phase, lo_buf) and lbm_pack_face(2 * axis + 1,
 * phase, hi_buf). No reference counterpart: the reference packs each face with
 * its own getDensityDistributions call (CController.cpp:369-409, 807-847). */
int lbm_pack_face_pair(lbm_solver* solver, int axis, int phase, void* lo_buf, void* hi_buf,
                       hipStream_t stream);

/* Fused halo packing of the two x faces into caller-owned device buffers of
 * lbm_face_elements(LEFT) elements each (NULL disables a side, both NULL
 * turns it off). While set, every whole-domain or box step also writes:
 *   alpha: the 5 populations of the send plane (x = 1 / x = nx_h - 2) that
 *          lbm_pack_face(phase 0) packs, at the same buffer positions;
 *   beta : the 5 populations that cross the face into the halo plane, at
 *          their halo-plane positions, i.e. the entries of
 *          lbm_pack_face(phase 1) that the neighbour's
 *          lbm_unpack_face(phase 1) reads (the other 14 are not written).
 * The x faces come first in the reference's exchange order
 * (CController.cpp:369-436, 807-874), so packing them inside the step is
 * exact; y and z faces keep lbm_pack_face. Launch options the caller has not
 * set follow the precision's defaults for packing or not packing steps.
 * No reference counterpart. */
int lbm_set_step_pack(lbm_solver* solver, void* left_buf, void* right_buf);

/* Fused halo unpacking of the two x faces from caller-owned device buffers
 * in the layout lbm_pack_face writes (NULL disables a side, both NULL turns it
 * off). While set, every whole-domain or box step takes the populations an x
 * face receives from these buffers, with exactly the results (populations,
 * rho / u, overflow flag) of running before it
 *   alpha: lbm_unpack_face(solver, LBM_LEFT / LBM_RIGHT, 1, left / right_buf),
 *   beta : lbm_unpack_face(solver, LBM_LEFT / LBM_RIGHT, 0, left / right_buf);
 * the caller skips those x-face unpacks. The send-plane cells read the
 * received values instead of the slots the unpack would have written (cells
 * that store nothing write them there). Each buffer must hold the neighbour's
 * pack of the previous half-step when a step runs. A phase-0 and a phase-1
 * pack of a face carry disjoint planes, so a neighbour's lbm_set_step_pack
 * buffer (same device) can serve as the receive buffer directly, with no
 * copy. Ordering: the reference unpacks x before it packs y and z, whose face
 * boxes include the x-halo cells, so in a decomposition also split along y
 * or z its y / z exchange carries the received x values on to the edge cells
 * of those neighbours; a fused unpack applies them only inside the next
 * step, after such exchanges. It reproduces the reference order where x is
 * the only split axis. LBM_ERR_UNSUPPORTED for an x extent below 2 interior
 * cells. No
 * reference counterpart: the reference unpacks with setDensityDistributions
 * (CController.cpp:410-436, 848-874). */
int lbm_set_step_unpack(lbm_solver* solver, const void* left_buf, const void* right_buf);

/* Single-GPU stand-in for the neighbour send/recv of one face: the result of
 * lbm_pack_face(from, direction, phase) followed by lbm_unpack_face into `to`
 * across the opposite face (e.g. from.RIGHT -> to.LEFT), in one launch that
 * copies the 5 populations from `from`'s face straight into `to`'s (no
 * buffer). Both solvers must live on the same device and have the same
 * halo-inclusive extent on the two axes of the face. */
int lbm_exchange_face(lbm_solver* from, int direction, int phase, lbm_solver* to,
                      hipStream_t stream);

/* One face transfer of lbm_exchange_faces: from's `direction` face into
 * to's opposite face. */
typedef struct lbm_face_transfer {
    lbm_solver* from;
    int direction;
    lbm_solver* to;
} lbm_face_transfer;

/* lbm_exchange_face for each of `n` transfers of one phase, with the same
 * result as those calls one after another, in ceil(n / 8) launches: all the
 * transfers of a decomposition along one axis (the faces between every pair
 * of neighbours on that axis, both ways). All transfers share one axis, device
 * and precision, and no two write the same face (same `to` and `direction`);
 * otherwise LBM_ERR_ARG and nothing runs. On one axis a phase's sends read
 * the face planes on one side of every face (phase 0 the last interior plane,
 * phase 1 the halo) and its receives write the planes on the other, so the
 * transfers are independent. The axes still go in order x, y, z (the y and z
 * faces carry x-halo cells). No reference counterpart: the reference
 * exchanges each face with its own MPI send/recv pair (CController.cpp:369-570,
 * 807-1008). */
int lbm_exchange_faces(const lbm_face_transfer* transfers, int n, int phase,
                       hipStream_t stream);

/* Device memory for face buffers and other caller-owned scratch. */
int lbm_device_alloc(int device, size_t bytes, void** out);
int lbm_device_free(int device, void* ptr);

/* ---- FP64 reductions (no reference counterpart; used by the steady-state
 *      loop and the mass-conservation checks) ------------------------------ */

/* Sum over the box of all 19 populations of every cell, in FP64:
 * out[0] = sum f, out[1..3] = sum c_l f (momentum in the layout the
 * populations are in), out[4] = number of cells. */
int lbm_reduce_box(lbm_solver* solver, const int origin[3], const int size[3], double out[5],
                   hipStream_t stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lbm_reduce_box of `count` solvers at once, allocating
 * nothing. Box m is origins[3m..3m+2] / sizes[3m..3m+2] (NULL origins or
 * sizes: every solver's whole halo-inclusive domain); the five sums of box m
 * go to out_dev[5m..5m+4], caller-owned device memory, bit-identical to what
 * lbm_reduce_box returns for that solver and box. The block partials live in
 * a static device array of the library sized for LBM_NOALLOC_MAX_BOXES boxes
 * (640 KiB per GPU, held until the process ends); count outside
 * 1..LBM_NOALLOC_MAX_BOXES is rejected with LBM_ERR_ARG. All solvers must
 * share one device and one precision. Asynchronous on `stream` (kernel
 * launches only: no allocation, copy or synchronisation). Calls that overlap
 * on one GPU share the partials array and race: order them on one stream. */
#define LBM_NOALLOC_MAX_BOXES 16
int lbm_reduce_boxes_noalloc(lbm_solver* const* solvers, const int* origins, const int* sizes,
                             int count, double* out_dev, hipStream_t stream);

/* Steady-state residual ||u - u_prev||_2 / ||u||_2 over FLUID cells, in FP64,
 * where u is the momentum computed from the populations (`layout` 0: after
 * beta, natural slots; 1: after alpha, swapped slots) and u_prev the value
 * saved by the previous call on this solver. Saves u for the next call.
 * The first call returns +inf; a non-finite state returns NaN. Synchronous. */
int lbm_residual(lbm_solver* solver, int layout, double* residual, hipStream_t stream);

/* ---- batched cavities ----------------------------------------------------
 * No reference counterpart. A batch steps many independent solvers (same
 * device and precision) with one launch per half-step over all their cells;
 * each member's result is bit-identical to stepping it alone. Members must
 * outlive the batch and must all be at the same step parity. */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_batch_tuning {
    int block;             /* threads per workgroup: 256, 512 or 1024 (not FP64) */
    int cells_per_thread;  /* 1, 2 or 4 (tile = block * cells_per_thread cells) */
    int waves_per_cu;      /* 0: one workgroup per tile; k > 0: persistent grid of
                              k workgroups per CU striding over the tiles */
    int options[2];        /* LBM_LAUNCH_* for the alpha and beta launches */
    int pairs;             /* 0: half-step launches. 1..4: one out-of-place pass per
                              alpha + beta pair with column tile 32x8, 32x4, 16x8 or
                              32x6 (two
                              time steps per sweep over HBM, a second population
                              buffer per member); needs F32 / F16S / BF16S, an
                              all-OBSTACLE array boundary and no stored rho / u
                              outputs. options[0] applies to the pass
                              (NONTEMPORAL: streaming stores of the new state).
                              The second buffers are the batch's: freed when a
                              tuning without the pass is applied or the batch
                              is destroyed, so a solver may be in the pair pass
                              of one batch at a time. */
    int group_mb;          /* 0: every launch covers all members. > 0: members are
                              packed, in order, into groups of at most group_mb MiB
                              of populations (a larger member forms its own group);
                              a half-step launches once per group, and a run
                              advances one group through all its half-steps before
                              the next, so a group's state can stay in the GPU's
                              last-level cache between steps. Members are
                              independent: the bits are the same in any schedule.
                              Not with pairs. */
} lbm_batch_tuning;

int lbm_batch_create(lbm_solver* const* members, int count, lbm_batch** out);
int lbm_batch_destroy(lbm_batch* batch);
int lbm_batch_size(const lbm_batch* batch);
int lbm_batch_set_tuning(lbm_batch* batch, const lbm_batch_tuning* tuning);
int lbm_batch_get_tuning(const lbm_batch* batch, lbm_batch_tuning* tuning);
/* One half-step (phase 0 alpha, 1 beta) of every member. */
int lbm_batch_step(lbm_batch* batch, int phase, hipStream_t stream);
/* half_steps alternating half-steps starting with first_phase. On a non-NULL
 * stream the sequence is captured once into a HIP graph (per first_phase,
 * half_steps and stream) and replayed on later calls. With tuning.pairs,
 * first_phase must be 0 and half_steps a multiple of 4 (the state returns
 * to each member's own buffer). */
int lbm_batch_run(lbm_batch* batch, int first_phase, int half_steps, hipStream_t stream);
/* The launches of lbm_batch_run issued directly on the stream, without graph
 * capture (same schedule, same bits). */
int lbm_batch_launch(lbm_batch* batch, int first_phase, int half_steps, hipStream_t stream);
/* Times candidate tunings (block, cells_per_thread, waves_per_cu, launch
 * options, member groups with the shapes again under the chosen group, and
 * the pair pass when the members allow it) as runs of trial_steps whole steps
 * (lbm_batch_run: graph-captured on a non-NULL stream, direct launches on the
 * NULL stream); the four fastest distinct candidates and the starting tuning
 * then run again over runs four times as long, interleaved, and the least
 * total time wins (the starting tuning on a tie).
 * The trials are real steps: the members advance; reset them afterwards
 * (lbm_reset) when their state matters. Synchronous. */
int lbm_batch_autotune(lbm_batch* batch, int trial_steps, hipStream_t stream,
                       lbm_batch_tuning* chosen);
/* lbm_residual for every member in one launch (each value bit-identical to
 * the member's own lbm_residual); residuals[count]. */
int lbm_batch_residual(lbm_batch* batch, int layout, double* residuals, hipStream_t stream);

/* Populations back to the rest equilibrium, u = 0, rho = 1, flags untouched
 * (the state lbm_create leaves, lid included). Synchronous. */
int lbm_reset(lbm_solver* solver);

/* Checkpoint of the whole A-A state: the populations of every cell, halo
 * included, as stored (F16S / BF16S keep their 16-bit codes, so a restore is
 * exact), followed by the flags: [19][cells with halo] elements of the
 * storage type, then [cells with halo] bytes. The argument-less
 * get/setDensityDistributions of the reference cover the interior only
 * (CLbmSolverCPU.cpp:552-559, 699-706), which loses the populations the A-A
 * pattern keeps in the wall cells. No counterpart in the reference.
 * lbm_state_bytes gives the size; save / load are synchronous host copies.
 * A state saved after alpha must be continued with beta. */
int lbm_state_bytes(const lbm_solver* solver, uint64_t* bytes);
int lbm_save_state(lbm_solver* solver, void* dst, uint64_t bytes);
int lbm_load_state(lbm_solver* solver, const void* src, uint64_t bytes);

/* Device bytes lbm_create allocates for these parameters, plus the residual
 * snapshot lbm_residual / lbm_batch_residual add on first use (no
 * counterpart in the reference; used for HBM-budget admission). Validates
 * params like lbm_create; touches no device. pair_buffer = 1 adds the second
 * population buffer of the batch pair pass. */
int lbm_device_bytes(const lbm_params* params, int pair_buffer, uint64_t* bytes);

/* ---- measurement --------------------------------------------------------
 * No reference counterpart: bandwidth probes of the step kernels' access
 * patterns without the collision, the ceilings that bound k_alpha / k_beta on
 * the device (bench.py --workload ceiling). They run on a solver's own
 * population array, elements of its storage size (4 bytes FP32, 8 FP64, 2 for
 * the 16-bit formats), one thread per cell: */
enum lbm_probe_mode {
    LBM_PROBE_RMW_ALIGNED = 0,  /* load + store the cell's own 19 slots (alpha's pattern) */
    LBM_PROBE_RMW_SHIFTED = 1,  /* load + store plane l at id + c_l (beta's pattern), cells
                                   whose neighbours lie inside the array */
    LBM_PROBE_READ = 2,         /* 19 loads per cell */
    LBM_PROBE_WRITE = 3,        /* 19 stores per cell (stores the rest state) */
    LBM_PROBE_RMW_SHIFTED_PITCHED = 4, /* beta's pattern with rows padded to 128 bytes:
                                   shifts c_x + c_y P + c_z P ny_h, P = nx_h rounded
                                   up to 32 */
    LBM_PROBE_RMW_SHIFTED_X = 5,  /* the x components of beta's shifts only */
    LBM_PROBE_RMW_SHIFTED_YZ = 6, /* the y and z components only */
    LBM_PROBE_RMW_LANE_SHIFTED = 7 /* FP32: beta's pattern over rows padded to 128 bytes
                                   with the x shifts taken across lanes: each wave
                                   loads and stores whole aligned 64-cell chunks (its
                                   LBM_PROBE_SPAN2 / 4 / 8 = that many consecutive
                                   chunks, default 1; SPAN8 not with BLOCK1024) plus
                                   one element past either end of its run */
};
/* options: LBM_LAUNCH_NONTEMPORAL and LBM_LAUNCH_WAVES4/3/2 (a cap must hold
 * one workgroup: WAVES3 / WAVES2 exclude BLOCK1024), and for the
 * read-modify-write patterns either the workgroup shape -- LBM_PROBE_BLOCK512
 * / BLOCK1024 threads (default 256) and LBM_PROBE_SPAN2 / 4 / 8 cells per
 * thread one block apart (default 1), so a workgroup covers block x span
 * consecutive cells -- or, on FP32 solvers, LBM_PROBE_VEC2 / LBM_PROBE_VEC4
 * (2 or 4 consecutive cells per thread, one vector access per plane); *cells
 * (may be NULL) receives the number of cells the launch covers.
 * Asynchronous. */
#define LBM_PROBE_VEC2 256
#define LBM_PROBE_VEC4 512
#define LBM_PROBE_BLOCK512 1024
#define LBM_PROBE_BLOCK1024 2048
#define LBM_PROBE_SPAN2 4096
#define LBM_PROBE_SPAN4 8192
#define LBM_PROBE_SPAN8 16384
int lbm_probe_pattern(lbm_solver* solver, int mode, int options, uint64_t* cells,
                      hipStream_t stream);

/* ---- misc ---------------------------------------------------------------- */

/* Number of visible GPUs: cudaGetDeviceCount in the CLbmSolverGPU constructor
 * (CLbmSolverGPU.cu:43-45), which places subdomain `id` on device
 * id % count. */
int lbm_device_count(int* count);

int lbm_sync(lbm_solver* solver, hipStream_t stream);
const char* lbm_last_error(void);
const char* lbm_status_string(int status);
int lbm_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CAVITYFLOW_LBM_ABI_H */
