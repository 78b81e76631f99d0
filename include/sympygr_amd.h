/* sympygr_amd -- BSSN right-hand side on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libsympygr_amd.so.  Plain pointers and sizes only; field arrays
 * use the reference's unzipped layout: 24 arrays in bssn::VAR order
 * (bssn/include/grDef.h:31), each holding every ghost-padded block
 * back to back, block b starting at its `offset`, point (i,j,k) of a block at
 * offset + i + nx*(j + ny*k) (bssn/include/derivs.h:8).  Outputs are written
 * at interior points [3, n-3)^3 only; ghost entries are left untouched
 * (bssn/src/rhs.cpp:206-369).
 *
 * Every entry point returns 0 on success and a negative SGR_E* code on
 * failure; sgr_last_error() gives the message of the calling thread's last
 * failure.  Nothing calls exit().
 */
#ifndef SYMPYGR_AMD_H
#define SYMPYGR_AMD_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGR_NUM_VARS 24

enum {
    SGR_OK = 0,
    SGR_E_INVALID = -1,     /* bad argument (null pointer, size, flag) */
    SGR_E_UNSUPPORTED = -2, /* valid for the reference, not handled here */
    SGR_E_HIP = -3,         /* HIP runtime error (message has the HIP text) */
    SGR_E_NOMEM = -4,       /* device or host allocation failed */
    SGR_E_NONFINITE = -5    /* non-finite input detected by sgr_check_finite */
};

enum { SGR_ADV_CENTRED = 0, SGR_ADV_UPWIND = 1 };
enum { SGR_PREC_FP64 = 0, SGR_PREC_MIXED = 1 };
enum { SGR_GAUGE_STANDARD = 0, SGR_GAUGE_ROCHESTER = 1 };
enum { SGR_ETA_CONST = 0, SGR_ETA_FUNC = 1 };
enum { SGR_FORM_BSSN = 0, SGR_FORM_CCZ4 = 1 };
enum { SGR_DISS_KO = 0, SGR_DISS_TVB3 = 1, SGR_DISS_TVB5 = 2 };

/* Fields of a CCZ4 plan: the 24 BSSN fields (the Gt slots hold the evolved
 * Gamma-hat^i) and Theta as field 24.  Every field-array argument of a CCZ4
 * plan's entry points points to SGR_NUM_VARS_CCZ4 arrays. */
#define SGR_NUM_VARS_CCZ4 25

/* Physics and method parameters.  The reference reads these from mutable
 * bssn:: globals at call time (bssn/src/parameters.cpp:99-109,
 * bssn/src/rhs.cpp:148-151, 215-219, 330); here they are explicit. */
typedef struct {
    uint32_t lambda[4];       /* BSSN_LAMBDA */
    double lambda_f[2];       /* BSSN_LAMBDA_F */
    double eta_const;         /* ETA_CONST */
    double eta_r0;            /* ETA_R0 */
    double eta_damping_exp;   /* ETA_DAMPING_EXP */
    double ko_sigma;          /* KO_DISS_SIGMA */
    int32_t adv_mode;         /* SGR_ADV_CENTRED (as shipped: deriv64adv_* = deriv64_*)
                                 or SGR_ADV_UPWIND (deriv42adv_*, bssn/src/derivs.cpp:495-906) */
    int32_t precision;        /* SGR_PREC_FP64, or SGR_PREC_MIXED: FP64 state, stencils,
                                 KO and boundary terms; FP32 pointwise algebra */
    /* Gauge variants the reference selects at compile time (bssn/src/rhs.cpp:224-243): */
    int32_t gauge;            /* SGR_GAUGE_STANDARD (bssneqs_eta_*_standard_gauge.cpp) or
                                 SGR_GAUGE_ROCHESTER (USE_ROCHESTER_GAUGE: GR/bssn.py:194-209,
                                 b_rhs = BSSN_XI[1] beta.ad beta + 3/4 BSSN_XI[2] Gt - eta beta,
                                 B not evolved) */
    int32_t eta_mode;         /* SGR_ETA_CONST (ETA_CONST, damped beyond ETA_R0) or SGR_ETA_FUNC
                                 (USE_ETA_FUNC: GR/bssn.py:52, eta = BSSN_ETA_R0
                                 sqrt(igt^ij d_i chi d_j chi) / (1 - chi^p0)^p1) */
    uint32_t xi[3];           /* BSSN_XI (unsigned in the reference) */
    double eta_func_r0;       /* BSSN_ETA_R0 */
    double eta_power[2];      /* BSSN_ETA_POWER: p0, p1 */
    /* The evolved system: SGR_FORM_BSSN (bssn/src/rhs.cpp) or SGR_FORM_CCZ4, the
       conformal covariant Z4 system of the reference's ccz4_eqns.py:195-431
       (its closures stated in sympygr_amd/codegen/ccz4_eqns.py; standard gauge
       only). */
    int32_t formulation;
    double kappa[3];          /* CCZ4 damping kappa1, kappa2, kappa3 (real; the
                                 reference declares them unsigned, ccz4_eqns.py:42-45) */
    /* DISSIPATION_TYPE (bssn/src/parameters.cpp:109, rhs.cpp:729-757), scaled by
       ko_sigma: SGR_DISS_KO (Kreiss-Oliger, ko_deriv42_*), SGR_DISS_TVB3 or
       SGR_DISS_TVB5 (disstvb3_* / disstvb5_*, derivs.cpp:1348-2054, weighted by
       the characteristic speeds of max_spacetime_speeds, rhs.cpp:984-1028).
       TVB: BSSN only, not in sgr_rhs_block_noalloc.  Where the reference's TVB
       result is undefined (speeds read at ghost nodes it never wrote, TVB3
       values left unwritten next to an unflagged face) the speeds are
       evaluated at ghost nodes too and TVB3 takes its interior form there. */
    int32_t dissipation;
} sgr_params;

/* Defaults of bssn/src/parameters.cpp:50-51,99-109,123 with adv_mode/precision
 * given: standard gauge, constant eta, BSSN, Kreiss-Oliger dissipation;
 * kappa = (0.02, 0, 0.5). */
void sgr_params_default(sgr_params *p, int32_t adv_mode, int32_t precision);

/* One ghost-padded block: the per-block arguments of the reference
 * bssnrhs(..., offset, ptmin, ptmax, sz, bflag) (bssn/include/rhs.h:33-36). */
typedef struct {
    uint64_t offset;   /* start of the block in every field array */
    uint32_t sz[3];    /* padded sizes (interior + 2*3), each >= 13 */
    uint32_t bflag;    /* bit d: face d is a domain boundary (LEFT RIGHT DOWN UP BACK FRONT) */
    double pmin[3];    /* coordinates of the first padded point */
    double pmax[3];    /* coordinates of the last padded point */
} sgr_block;

typedef struct sgr_plan sgr_plan;

typedef struct {
    uint64_t blocks;
    uint64_t tiles;              /* RHS work tiles: 8^3, or the shape tuned for the block's
                                    size (the library's tile table; SGR_TILE_SHAPE=tx,ty,tz
                                    forces one shape on every block) */
    uint64_t interior_points;
    uint64_t workspace_bytes;    /* device scratch held by the plan */
    uint32_t derivative_slots;   /* derivative values staged per point */
    uint32_t chunks;             /* launches per pass (workspace budget SGR_WS_GIB) */
    uint32_t pipelined;          /* 1: the chunks run pipelined (two workspace buffers; each
                                    chunk's stencil pass overlaps the previous chunk's
                                    pointwise pass on a second stream the plan owns): plans
                                    needing more than one chunk, and plans of at least 8192
                                    tiles without the tile queues */
    uint32_t tile_queue;         /* 1: the stencil pass is a persistent grid fed by per-XCD
                                    tile queues (plans where at least a tenth of the tiles
                                    reach a flagged face) */
} sgr_plan_info;

/* Replaces the per-block setup of the reference bssnRHS block loop
 * (bssn/src/rhs.cpp:41-61) and the workspace sizing of its GPU path
 * (bssn/cuda_gr/src/rhs_cuda.cu:113-135): validates the block table, uploads
 * it to `device` and allocates the derivative workspace once. */
int sgr_plan_create(sgr_plan **plan, const sgr_block *blocks, uint32_t nblocks, const sgr_params *params,
                    int device);
int sgr_plan_destroy(sgr_plan *plan);
int sgr_plan_get_info(const sgr_plan *plan, sgr_plan_info *info);

/* RHS of every block of the plan on `stream`; in/out are DEVICE pointers to
 * the 24 unzipped arrays.  Asynchronous; allocates nothing, so it can be
 * captured into a HIP graph.  A pipelined plan (sgr_plan_info) also runs
 * work on its own second stream, forked from and joined back into `stream`
 * through events, so the call is ordered on `stream` as one launch would
 * be.  Calls on one plan share its workspace (and second stream) and must
 * not overlap: issue them on one stream or serialise them.  Replaces the
 * reference's multi-block entry bssnRHS (bssn/include/rhs.h:31) and its GPU
 * plug-in cuda::computeRHS (bssn/cuda_gr/include/rhs_cuda.cuh:39). */
int sgr_rhs(sgr_plan *plan, double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS],
            hipStream_t stream);

/* Same as sgr_rhs, and records events[0..1] around the stencil pass and
 * events[2..3] around the pointwise kernel on `stream` (hipEventElapsedTime
 * then gives each pass's duration).  For a plan that runs in several
 * workspace chunks (pipelined or not), events[0..1] bracket the whole call
 * and events[2..3] are recorded at its end.  No reference counterpart: it
 * exists for measurement. */
int sgr_rhs_events(sgr_plan *plan, double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS],
                   hipStream_t stream, hipEvent_t events[4]);

/* Hamiltonian constraint (out[0]) and momentum constraints (out[1..3]) at
 * every interior point of the plan's blocks (GR/constraints.py:161-173, same
 * derivative operators as the RHS), computed and stored in FP64.  The
 * derivatives are staged in the plan's workspace type: FP64 for an FP64
 * plan; FP32 for a mixed-precision plan, whose constraints therefore carry
 * FP32 rounding of the derivatives (about 1e-7 relative) -- use an FP64 plan
 * for FP64 constraints.  Uses the plan's workspace: do not overlap with
 * sgr_rhs on the same plan.  The reference's generated counterpart
 * (bssn/CodeGen/physcon.cpp) does not compile. */
int sgr_constraints(sgr_plan *plan, const double *const in[SGR_NUM_VARS], double *const out[4], hipStream_t stream);

/* Same as sgr_constraints, and also the real (out[4]) and imaginary (out[5])
 * parts of the Weyl scalar Psi4 at every interior point (GR/constraints.py:
 * 46-155): tetrad built by Gram-Schmidt from the coordinate vectors r, theta,
 * phi at the node's coordinates (x, y, z) = pmin + (i, j, k) h of its block
 * (dendrosym/codegen.py:732).  Keeps the statement of GR/constraints.py as
 * written, including the sign of the half dchi terms in the fourth part of
 * the imaginary part (GR/constraints.py:150), so on a boosted slice of flat
 * space the imaginary part is not zero where those terms are not.  The
 * reference's counterpart is the generated physcon kernel that
 * sgr_constraints replaces. */
int sgr_constraints_psi4(sgr_plan *plan, const double *const in[SGR_NUM_VARS], double *const out[6],
                         hipStream_t stream);

/* Ghost refresh (the unzip step between RK stages), in place.  Every block
 * owns the box of its interior nodes widened by half a cell; a ghost node
 * inside another block's owned box takes that block's value: the coinciding
 * node where the grids align, otherwise the tensor-product 6-point Lagrange
 * interpolant over that block's interior nodes (neighbours may differ in
 * spacing, size and level).  Ghost nodes beyond flagged (domain-boundary)
 * faces are left as they are (no stencil reads them).  SGR_E_UNSUPPORTED if
 * some ghost node inside the domain has no owner, blocks overlap, or the
 * block coordinates share no common lattice.  No reference counterpart
 * (Dendro's unzip lives outside the reference repository). */
int sgr_ghost_fill(sgr_plan *plan, double *const u[SGR_NUM_VARS], hipStream_t stream);

/* Host-only check of the ghost refresh table sgr_ghost_fill would use for
 * these blocks (no GPU needed): stats[0] patches, stats[1] ghost nodes
 * refreshed, stats[2] of them interpolated (not plain copies).  Returns the
 * error sgr_ghost_fill would return.  No reference counterpart. */
int sgr_ghost_table_stats(const sgr_block *blocks, uint32_t nblocks, uint64_t stats[3]);

/* Plain device-to-device copy of `bytes` (a multiple of 16, 16-byte aligned
 * pointers) by one kernel, k_hbm_copy: reads and writes exactly `bytes`, the
 * reference point that scales hardware byte counters.  No reference
 * counterpart (measurement aid). */
int sgr_hbm_copy(void *dst, const void *src, uint64_t bytes, hipStream_t stream);

/* One classic-RK4 stage with the update fused into the RHS output store
 * (interior points only; a ghost refresh must precede the next stage):
 *   stage 0: k = RHS(u_stage); acc = u_n + dt/6 k; u_next = u_n + dt/2 k
 *   stage 1: k = RHS(u_stage); acc += dt/3 k;      u_next = u_n + dt/2 k
 *   stage 2: k = RHS(u_stage); acc += dt/3 k;      u_next = u_n + dt k
 *   stage 3: k = RHS(u_stage); u_next = acc + dt/6 k   (u_next may be u_n)
 * u_next must not alias u_stage.  The reference has only the RK enum
 * (bssn/include/grDef.h:25); the stage is the textbook tableau. */
int sgr_rk4_stage(sgr_plan *plan, int stage, double dt, const double *const u_n[SGR_NUM_VARS],
                  const double *const u_stage[SGR_NUM_VARS], double *const acc[SGR_NUM_VARS],
                  double *const u_next[SGR_NUM_VARS], hipStream_t stream);

/* A full RK4 step of u in place: ghost refresh + 4 fused stages, using the
 * caller's stage buffers s_a, s_b and accumulator acc (all same layout). */
int sgr_rk4_step(sgr_plan *plan, double dt, double *const u[SGR_NUM_VARS], double *const s_a[SGR_NUM_VARS],
                 double *const s_b[SGR_NUM_VARS], double *const acc[SGR_NUM_VARS], hipStream_t stream);

/* The same step captured once into a HIP graph (4 ghost refreshes + 4 stages =
 * 20 kernel nodes) and replayed with sgr_graph_launch. */
typedef struct sgr_graph sgr_graph;
int sgr_rk4_graph_create(sgr_plan *plan, double dt, double *const u[SGR_NUM_VARS], double *const s_a[SGR_NUM_VARS],
                         double *const s_b[SGR_NUM_VARS], double *const acc[SGR_NUM_VARS], sgr_graph **graph);
int sgr_graph_launch(sgr_graph *graph, hipStream_t stream);
int sgr_graph_destroy(sgr_graph *graph);

/* One block, DEVICE pointers, synchronous: bssnrhs (bssn/include/rhs.h:33-36)
 * with the arrays already on the device.  The per-block entries keep one
 * context per calling thread (a one-block plan, a stream and staging
 * buffers, rebuilt only when the device, block size, advection mode,
 * precision or formulation changes; held until the process ends), so calls
 * from different host threads proceed concurrently. */
int sgr_rhs_block(double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS], uint64_t offset,
                  const double pmin[3], const double pmax[3], const uint32_t sz[3], uint32_t bflag,
                  const sgr_params *params, int device);

/* The same as sgr_rhs_block on the CURRENT device, asynchronously on `stream`,
 * allocating nothing: no reference counterpart; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its scratch memory (the block's
 * descriptor and derivative workspace) is a fixed 32 MiB device array that
 * each GPU the library runs on holds until the process ends, shared by every
 * call on that GPU: a block may use at most 32 MiB of it, which holds 87
 * tiles of 8^3 interior points in FP64 BSSN (a 32^3 interior block has 64;
 * 139 in mixed precision); larger blocks are refused with SGR_E_INVALID.  Calls that overlap on one GPU (two
 * streams, or two threads) race on that memory; serialise them.  The entry
 * point itself calls no hipMalloc, hipFree or allocating routine. */
int sgr_rhs_block_noalloc(double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS], uint64_t offset,
                          const double pmin[3], const double pmax[3], const uint32_t sz[3], uint32_t bflag,
                          const sgr_params *params, hipStream_t stream);

/* One block, HOST pointers: stages the block's padded input to the device,
 * computes, and copies the interior RHS back.  This is the backend of the
 * C++ drop-in bssnrhs(...) (bssn/include/rhs.h:33-36). */
int sgr_bssnrhs_host(double *const out[SGR_NUM_VARS], const double *const in[SGR_NUM_VARS], uint64_t offset,
                     const double pmin[3], const double pmax[3], const uint32_t sz[3], uint32_t bflag,
                     const sgr_params *params, int device);

/* Returns SGR_E_NONFINITE if any padded input point of any plan block is NaN
 * or Inf (the reference propagates them silently). */
int sgr_check_finite(sgr_plan *plan, const double *const in[SGR_NUM_VARS], hipStream_t stream);

/* Copies the calling thread's last error message into buf (NUL-terminated)
 * and returns its full length. */
size_t sgr_last_error(char *buf, size_t len);

/* Library version string. */
const char *sgr_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SYMPYGR_AMD_H */
