/*
 * ssmd.h -- C ABI of the MI355X (gfx950) per-timestep hot path of the
 * ssmdLBM two-phase jet solver: regularized D3Q19 / D3Q27 stream+collide,
 * the D3Q7 conservative Allen-Cahn phase field (phi, normals, curvature,
 * surface-tension force) and the jet inlet / outflow / periodic boundaries.
 *
 * Reference interface this header replaces (file:line in brenogemelgo/ssmdLBM):
 *   - struct LBMFields               src/structs/LBMFields.cuh:26-85
 *   - host::FieldAllocate            src/fieldAllocate/FieldAllocate.cuh:42-162
 *   - kernel set k(LBMFields)        src/LBMIncludes.cuh:29-47, src/phaseField.cuh:31,54
 *   - graph::captureGraph            src/cuda/CUDAGraph.cuh:34-57
 *   - compile-time constants         src/constants.cuh:70-101, src/functions/globalFunctions.cuh:55-104,232-283
 *
 * Conventions
 *   - Every function returns 0 on success, a negative SSMD_E* code for an
 *     argument error, or a positive hipError_t value from the HIP runtime.
 *     Nothing aborts inside the library (the reference aborts in its
 *     checkCudaErrors macros, src/cuda/utils.cuh:101-161; a host driver that
 *     wants that behaviour aborts on a nonzero return).
 *   - All device work is asynchronous on the caller's stream (`stream` is a
 *     hipStream_t passed as void*; NULL selects the null stream).  One field
 *     set per stream; calls on one field set must not overlap.
 *   - Scalar fields are float32, x fastest: idx = x + nx*(y + ny*z), ghost
 *     layer = the outermost plane on every face (reference deviceFunctions.cuh:33-59).
 *   - Populations are Q-major planes, f[i*N + idx], stored as deviations
 *     fbar_i = f_i - w_i (reference D3Q19.cuh:139).  Unlike the reference's
 *     arrival-indexed push layout, the buffers here are SOURCE-indexed
 *     ("pull"): slot f[i][s] holds the post-collision population that leaves
 *     cell s along c_i, so the population arriving at cell a along c_i is
 *     f[i][a - c_i].  Ghost-cell slots hold what the boundaries emit.
 *   - The D3Q7 phase populations are stored in moment form: g holds 4
 *     planes (g_phi, o_x, o_y, o_z) per source cell s, and the population
 *     leaving s along c_k is w_k (g_phi[s] + c_k . o[s]) (w_0 = 1/4, else
 *     1/8).  The collision emits exactly this family (o = 4 phi u +
 *     gamma phi (1-phi) n), so the arriving sum is the same FP32 sum as
 *     with seven population planes, with 16 instead of 28 bytes per cell.
 *   - f ping-pongs: ssmd_step swaps f<->f_next (and g<->g_next).  g needs no
 *     second buffer: its one reader, the phase update, runs before the
 *     lattice kernel of the same step rewrites it, so g_next may be the same
 *     buffer as g, as ssmd_alloc makes it (a one-step g then stays in the
 *     Infinity Cache between the two kernels).
 */
#ifndef SSMD_H
#define SSMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSMD_ABI_VERSION 6

#define SSMD_OK 0
#define SSMD_EINVAL (-1)       /* bad argument or inconsistent parameters */
#define SSMD_EUNSUPPORTED (-2) /* parameter combination not built */
#define SSMD_ENOTREADY (-3)    /* fields not allocated / graph not created */

/* Population storage formats.  F32 is the reference's default scalar_t
 * storage; F16 is its ENABLE_FP16 pop_t (cuda/utils.cuh:67-99, IEEE binary16
 * with round-to-nearest-even); BF16 is new. Arithmetic is always float32. */
enum { SSMD_POP_F32 = 0, SSMD_POP_F16 = 1, SSMD_POP_BF16 = 2 };

/* Periodic handling of an axis.
 *   NONE      : no wrap; populations leaving from ghost cells stay frozen.
 *   REFERENCE : the reference periodicX copy (BoundaryConditions.cuh:295-349):
 *               arrival planes 1 <-> n-2 exchange, phi/u ghost wrap.  For y
 *               it is the mirror image of that copy.
 *   TRUE      : populations re-enter with period n-2 (mass conserving);
 *               phi/u ghost wrap as REFERENCE. */
enum { SSMD_WRAP_NONE = 0, SSMD_WRAP_REFERENCE = 1, SSMD_WRAP_TRUE = 2 };

/* Subgrid closure of the hydrodynamic relaxation time (the reference has no
 * subgrid model; all three are applied in the interior collision only -- the
 * inlet/outflow reconstructions keep their fixed omco).  tau0 = 1/2 + 3 nu(phi).
 *   NONE        : tau = tau0 (the reference, lbm.cu:229-231)
 *   SMAGORINSKY : tau = (tau0 + sqrt(tau0^2 + 18 Cs^2 sqrt(2 Pi:Pi) / rho)) / 2,
 *                 Pi the non-equilibrium moments of the step (Hou et al. 1996)
 *   LEARNED     : tau = tau0 + tau_t[cell]; the step writes the network inputs
 *                 of every interior cell to feat and then refreshes tau_t
 *                 from them (ssmd_closure_update), so tau_t lags one step. */
enum { SSMD_CLOSURE_NONE = 0, SSMD_CLOSURE_SMAGORINSKY = 1, SSMD_CLOSURE_LEARNED = 2 };

/* ssmd_step flags */
#define SSMD_STEP_MACROS 1u /* write rho, u, Pi and F at every interior cell */

/* A circular velocity inlet on a ghost face.
 * water: face y = 0, disc coordinates (a, b) = (x, z)   (applyWaterInflow, BoundaryConditions.cuh:36-99)
 * oil:   face z = 0, disc coordinates (a, b) = (x, y)   (applyOilInflow,   BoundaryConditions.cuh:101-168) */
typedef struct ssmd_jet {
    int32_t enabled;
    int32_t diameter; /* mesh::diam_* (constants.cuh:77-78): nominal diameter in cells, used
                         for the info file's Reynolds and Weber numbers; 0 = 2 sqrt(r2) */
    double center_a; /* geometry::center_x() = (nx-1)/2 in the reference */
    double center_b; /* geometry::z_pos() / y_pos() in the reference */
    double r2;       /* geometry::R2_*(): (D/2)^2 with integer D/2 */
    double u;        /* inflow speed along the face normal (physics::u_*) */
    double phi;      /* phase carried in: 0 water, 1 oil */
} ssmd_jet;

typedef struct ssmd_params {
    uint32_t nx, ny, nz; /* mesh::nx, ny, nz (ghost layer included) */
    int32_t q;           /* 19 or 27: hydrodynamic lattice */
    int32_t pop_format;  /* SSMD_POP_* */
    int32_t outflow_y;   /* applyOutflowY on y = ny-1 */
    int32_t outflow_z;   /* applyOutflowZ on z = nz-1 */
    int32_t wrap_x, wrap_y, wrap_z; /* SSMD_WRAP_* */
    int32_t tile;        /* lattice tile shape: index into the built shapes
                            (0: 64x4, 1: 128x2, 2: 64x2, 3: 128x4; 4: 128x4 with the
                            tiles dealt to the XCDs in contiguous ranges); see ssmd_tune */
    int32_t lz;          /* z planes per lattice block; 0 = automatic */
    double visc_water;   /* relaxation::visc_water() */
    double visc_oil;     /* relaxation::visc_oil() */
    double sigma;        /* physics::sigma */
    double gamma;        /* physics::gamma (D3Q7 sharpening) */
    double perturb;      /* inlet velocity noise: each step, cell and component
                            adds perturb * u * U[-1,1) (0 in the reference) */
    uint64_t seed;       /* seed of that noise */
    ssmd_jet water;
    ssmd_jet oil;
    int32_t closure;     /* SSMD_CLOSURE_* */
    int32_t pad1;
    double cs;           /* Smagorinsky constant (0.1) */
    /* LEARNED: network inputs per cell, in order: the strain rate
     * S_ab = -3 Pi_ab / (2 rho tau) times strain_scale (xx, yy, zz, xy, xz, yz),
     * phi, and |grad phi-hat| (the masked gradient of the force) times
     * grad_scale; output tau_t = max(tau_floor, tau_scale * network(inputs)).
     * tau_floor = 0 (default) keeps tau_t >= 0 (no backscatter); a negative
     * floor admits backscatter and must keep tau above 1/2 with margin:
     * tau_floor >= -0.9 * 3 * min(visc_water, visc_oil), or SSMD_EINVAL. */
    double strain_scale;
    double grad_scale;
    double tau_scale;
    double tau_floor;
    /* LEARNED: subgrid flux of the phase field as gradient diffusion with the
     * eddy viscosity, J = -D_t grad(phi), D_t = max(flux_coeff tau_t / 3, -0.1)
     * (flux_coeff = 1/Sc_t, an inverse turbulent Schmidt number): the D3Q7
     * first moments of the step get -4 D_t grad(phi), i.e. the sharpening
     * coefficient gamma phi (1 - phi) becomes gamma phi (1 - phi) - 4 D_t
     * |grad phi| along the normal.  0 (default): no flux closure; nonzero
     * values require the LEARNED closure (else SSMD_EINVAL).  The D3Q7
     * update stays stable (von Neumann, every 3-D mode) for D_t <= 0.375;
     * the caller keeps flux_coeff * max(tau_t) / 3 below that. */
    double flux_coeff;
} ssmd_params;

/* Device buffers.  Member order follows struct LBMFields (LBMFields.cuh:26-85);
 * normx/normy/normz/ind are never allocated, as in the reference (main.cu:55-70).
 * rho, u and Pi are refreshed every step on the six planes next to the faces
 * (x,y,z in {1, n-2}: the boundary kernels read them there); everywhere else
 * only on steps run with SSMD_STEP_MACROS, which also writes fsx/fsy/fsz. */
typedef struct ssmd_fields {
    float *rho, *ux, *uy, *uz;
    float *pxx, *pyy, *pzz, *pxy, *pxz, *pyz;
    float *phi;
    float *normx, *normy, *normz, *ind;
    float *fsx, *fsy, *fsz;
    void *f;          /* populations the next step reads (pop_format elements) */
    float *g;         /* D3Q7 phase populations the next step reads (4 moment planes) */
    void *f_next;     /* buffers the next step writes */
    float *g_next;    /* may equal g (ssmd_alloc: one g buffer) */
    uint64_t *step_counter; /* device word: inlet-noise step index for graph replays */
    double *reduce;   /* SSMD_REDUCE_DOUBLES doubles of reduction workspace */
    uint64_t owned;   /* nonzero when ssmd_alloc owns the buffers */
    /* SSMD_CLOSURE_LEARNED only (ssmd_alloc allocates them for that closure): */
    float *tau_t;     /* eddy relaxation time per cell, read by the collision */
    uint16_t *feat;   /* 8 bf16 network inputs per cell (16-byte aligned rows) */
    void *mlp;        /* packed network, ssmd_mlp_packed_bytes() (ssmd_mlp_pack) */
} ssmd_fields;

#define SSMD_REDUCE_DOUBLES 8192

typedef struct ssmd_graph ssmd_graph;

/* Library / build identification. */
int ssmd_abi_version(void);
const char *ssmd_error_string(int code);

/* Fill *p with the reference's compile-time case (constants.cuh:70-101,
 * globalFunctions.cuh:170-283) for lattice q: 128x256x256, water jet on y=0,
 * oil jet on z=0, outflow on y and z, reference periodicX. */
int ssmd_params_reference_default(ssmd_params *p, int32_t q);

/* 0 if *p describes a case this build runs, else SSMD_EINVAL/SSMD_EUNSUPPORTED.
 * Size limits: 4 * (nx*ny*nz + nx*ny) < 2^32 (the kernels address one
 * population plane with 32-bit byte offsets), e.g. 1024 x 1024 x 1020, and
 * at most 65536 cells along each axis. */
int ssmd_validate(const ssmd_params *p);

/* Bytes of one buffer: which = 0 scalar field, 1 f, 2 g (4 planes). 0 on bad params. */
size_t ssmd_bytes(const ssmd_params *p, int32_t which);

/* Allocate and zero every buffer (FieldAllocate semantics, hipMalloc). */
int ssmd_alloc(const ssmd_params *p, ssmd_fields *d);
/* Free buffers allocated by ssmd_alloc; no-op for caller-owned buffers. */
int ssmd_free(ssmd_fields *d);

/* t = 0 state: setWaterJet, setOilJet, setInitialDensity, setDistros
 * (initialConditions.cu:33-140), written into both population buffers. */
int ssmd_init(const ssmd_params *p, ssmd_fields *d, void *stream);

/* One timestep: phase::computePhase, then one fused kernel for
 * phase::computeForces + lbm::computeMoments + lbm::streamCollide, then the
 * boundary kernels in reference order (main.cu:162-169). `step` indexes the
 * inlet noise. Swaps f<->f_next and g<->g_next in *d. */
int ssmd_step(const ssmd_params *p, ssmd_fields *d, uint64_t step, uint32_t flags, void *stream);

/* The three launches ssmd_step makes, for callers that time or interleave
 * them (the reference's separate kernels, LBMIncludes.cuh:29-47):
 *   ssmd_launch_phase       phase::computePhase (reads d->g, writes phi)
 *   ssmd_launch_lattice     computeForces + computeMoments + streamCollide
 *                           (reads d->f, phi; writes d->f_next, d->g_next)
 *   ssmd_launch_boundaries  K5-K9 and the y/z periodic copies on d->f_next,
 *                           d->g_next (use_counter != 0: the inlet noise takes
 *                           the step index from d->step_counter and the call
 *                           advances it by one;
 *                           SSMD_STEP_MACROS in flags also refreshes rho/u on
 *                           the ghost layer, as ssmd_step does on macro steps)
 * A step run as these calls plus the LEARNED closure's ssmd_closure_update is
 * the same step as ssmd_step with the same flags.
 *   ssmd_swap               host-side exchange f<->f_next, g<->g_next */
int ssmd_launch_phase(const ssmd_params *p, const ssmd_fields *d, void *stream);
int ssmd_launch_lattice(const ssmd_params *p, const ssmd_fields *d, uint32_t flags, void *stream);
int ssmd_launch_boundaries(const ssmd_params *p, const ssmd_fields *d, uint64_t step, int32_t use_counter,
                           uint32_t flags, void *stream);
int ssmd_swap(ssmd_fields *d);

/* nsteps calls of ssmd_step for steps step0 .. step0+nsteps-1. */
int ssmd_run(const ssmd_params *p, ssmd_fields *d, uint64_t step0, uint64_t nsteps, void *stream);

/* Capture the step (both buffer parities) into HIP graphs, once as a single
 * step and once as a block of 16 steps.  The boundary kernels are inside the
 * graphs; the inlet-noise step index comes from d->step_counter, advanced on
 * the device by the phase kernel that opens each step (ssmd_graph_launch sets
 * it to step0 - 1 first).  The face kernels of a captured step among {z
 * inlet, z outflow, x-y periodic ghost ring} run as one launch, with the
 * results of the separate kernels.  With the LEARNED closure the
 * network pass is a branch of the graph beside the boundary kernels and the
 * next phase update (the graph owns a second, default-priority stream for
 * the capture); results are those of ssmd_step. */
int ssmd_graph_create(const ssmd_params *p, const ssmd_fields *d, void *stream, ssmd_graph **out);
/* Replay nsteps steps starting at step index step0 (whole 16-step blocks,
 * then single steps); swaps *d as ssmd_step does. */
int ssmd_graph_launch(ssmd_graph *g, ssmd_fields *d, uint64_t step0, uint64_t nsteps, void *stream);
int ssmd_graph_destroy(ssmd_graph *g);

/* FP64 reductions over interior cells, copied to out_host[4] (synchronizes
 * the stream): sum(phi), sum(rho*ux), sum(rho*uy), sum(rho*uz).  The
 * momentum entries need rho/u from a step run with SSMD_STEP_MACROS. */
int ssmd_phase_mass_f64(const ssmd_params *p, const ssmd_fields *d, double *out_host, void *stream);

/* The populations the next step reads, in the reference's own layout, for
 * code written against the reference's LBMFields f and g (LBMFields.cuh:26-85,
 * arrival-indexed push layout, seven D3Q7 planes; no reference counterpart:
 * there these ARE the buffers).  f_out: q planes of nx*ny*nz elements in
 * p->pop_format, f_out[i][a] = d->f[i][a - c_i]; g_out: 7 float planes,
 * g_out[k][a] = w_k (g_phi + c_k . o) of source cell a - c_k, formed with the
 * FP32 operations of the phase update (so sum_k g_out[k][a] is its phi).
 * Periodic images are folded in as the pulls fold them, so interior arrivals
 * equal the reference's after its boundary kernels; an arrival with no source
 * in the box is 0 (the reference keeps its t = 0 value there, which no kernel
 * reads).  Either pointer may be NULL; neither may be one of d's buffers. */
int ssmd_populations_to_reference(const ssmd_params *p, const ssmd_fields *d, void *f_out, float *g_out,
                                  void *stream);

/* Bit-exact storage conversions used by the kernels, evaluated on the device
 * (n elements, device pointers): float32 -> format -> float32. */
int ssmd_quantize(int32_t pop_format, const float *in, float *out, uint64_t n, void *stream);

/* Autotuner: times the fused lattice kernel for every built tile shape and a
 * set of z-chunk lengths on the current state of *d (reps launches each) and
 * returns the fastest in *best_tile, *best_lz (ready for p->tile, p->lz).
 * The state the next step reads (f, g) is not changed; f_next, g_next, phi
 * and the face-plane moments are rewritten as a step's launches would. */
int ssmd_tune(const ssmd_params *p, const ssmd_fields *d, int32_t reps, void *stream, int32_t *best_tile,
              int32_t *best_lz);
int ssmd_tile_count(void);
/* The automatic z-chunk (what p->lz = 0 selects) for tile shape `tile` on
 * the box of *p: planes per lattice block, or a negative SSMD_E* code.  No
 * reference counterpart. */
int32_t ssmd_auto_lz(const ssmd_params *p, int32_t tile);
int ssmd_tile_shape(int32_t tile, int32_t *tx, int32_t *ty);

/* Ensembles (batching form): n independent field sets of uneven size and
 * parameters -- the same lattice, storage format and boundary set -- advanced
 * by ONE launch per kernel kind per step: a persistent lattice kernel striding
 * over the tiles of every member, and batched phase and face kernels.  Each
 * member's results are bit-identical to running it alone.  The ensemble
 * keeps its own copy of the field descriptors (swapped each step); read the
 * current ones with ssmd_ensemble_fields.  Graph replays advance a device
 * step counter owned by the ensemble. */
typedef struct ssmd_ensemble ssmd_ensemble;
int ssmd_ensemble_create(const ssmd_params *params, const ssmd_fields *fields, int32_t n, ssmd_ensemble **out);
int ssmd_ensemble_step(ssmd_ensemble *e, uint64_t step, uint32_t flags, void *stream);
int ssmd_ensemble_graph_launch(ssmd_ensemble *e, uint64_t step0, uint64_t nsteps, void *stream);
/* Autotune the ensemble's lattice tile shape (kernel-only timing, as ssmd_tune).
 * Besides the shapes 0 .. ssmd_tile_count()-1, an ensemble takes the tile
 * index ssmd_tile_count(): 128x4 blocks that run 64x8 tiles for the members
 * whose rows a 64-wide tile covers with fewer idle lanes (nx <= 64). */
int ssmd_ensemble_tune(ssmd_ensemble *e, int32_t reps, void *stream, int32_t *best_tile);
int ssmd_ensemble_set_tile(ssmd_ensemble *e, int32_t tile);
int ssmd_ensemble_fields(const ssmd_ensemble *e, int32_t member, ssmd_fields *out);
int ssmd_ensemble_destroy(ssmd_ensemble *e);

/* Output in the reference's formats (host code; synchronizes `stream`):
 *   ssmd_write_bin   <dir>/<name><step:06>.bin, raw float32 of a device field
 *                    (host::PostProcess::bin, postProcess/PostProcess.cuh:45-98);
 *                    a null field (normx .. ind are never allocated) is
 *                    skipped: SSMD_OK, no file (PostProcess.cuh:64-73)
 *   ssmd_write_vti   <dir>/step_<step:06>.vti from those .bin files: VTK
 *                    ImageData, appended raw, uint32 byte count per array;
 *                    missing or wrong-size .bin files are skipped
 *                    (host::PostProcess::vti, PostProcess.cuh:101-262)
 *   ssmd_write_info  <dir>/<id>_info.txt metadata
 *                    (host::generateSimulationInfoFile, functions/hostFunctions.cuh:47-87);
 *                    the Reynolds numbers u D / nu and the Weber number
 *                    u_oil^2 D_oil / sigma are formed from the parameters
 *                    (D = ssmd_jet.diameter) and print as integers when they are */
int ssmd_write_bin(const ssmd_params *p, const float *device_field, const char *dir, const char *name, uint64_t step,
                   void *stream);
int ssmd_write_vti(const ssmd_params *p, const char *dir, uint64_t step, const char *const *names, int32_t count);
int ssmd_write_info(const ssmd_params *p, const char *dir, const char *id, const char *velocity_set, uint64_t nsteps,
                    uint64_t save_every);

/* ---------------------------------------------------------------------------
 * Derived statistics (reference src/derivedFields/, compiled out by default
 * there, compile.sh:71-74).  A null pointer skips that field, as the
 * reference's per-field checks do (timeAverage.cuh:52-78). */
typedef struct ssmd_derived {
    float *avg_phi, *avg_ux, *avg_uy, *avg_uz;                              /* timeAverage.cuh:37-82 */
    float *avg_uxux, *avg_uyuy, *avg_uzuz, *avg_uxuy, *avg_uxuz, *avg_uyuz; /* reynoldsMoments.cuh:37-96 */
    float *vort_x, *vort_y, *vort_z, *vort_mag;                             /* vorticityFields.cuh:37-105 */
} ssmd_derived;

/* Update after a step run with SSMD_STEP_MACROS (phi, rho, u current on
 * every cell), where the reference launches them (main.cu:175):
 *   averages  avg <- avg + (v - avg) / t with t = step + 1 (every cell);
 *   vorticity central differences of u on interior cells, 0 on the ghost
 *             layer, and its magnitude. */
int ssmd_derived_update(const ssmd_params *p, const ssmd_fields *d, const ssmd_derived *s, uint64_t step,
                        void *stream);

/* ---------------------------------------------------------------------------
 * Checkpoint / restart (no counterpart in the reference).  The file holds the
 * state one step hands to the next -- the populations the next step reads
 * (f, g), phi (its ghost layer is boundary data), the moment fields and, for
 * the LEARNED closure, tau_t and the packed network -- with the step index
 * and the parameters.  ssmd_checkpoint_read returns SSMD_EINVAL unless *p
 * equals the file's parameters in everything but the launch shape (tile, lz);
 * it then initializes the fields (ssmd_init: ghost slots no kernel writes keep
 * their t = 0 values) and loads that state; continuing from `step` reproduces
 * the uninterrupted run bit for bit.  Both calls synchronize `stream`. */
int ssmd_checkpoint_write(const ssmd_params *p, const ssmd_fields *d, uint64_t step, const char *path, void *stream);
int ssmd_checkpoint_read(const ssmd_params *p, ssmd_fields *d, uint64_t *step, const char *path, void *stream);

/* ---------------------------------------------------------------------------
 * Learned closure network (no counterpart in the reference).
 * 8 inputs -> 64 -> 64 -> 64 (ReLU) -> 1, BF16 weights and activations on the
 * matrix cores with FP32 accumulation.  Parameters are SSMD_MLP_PARAMS floats
 * in this order: W1[64][8] (row = output unit), b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], W4[64], b4.  Inputs are bf16 bit patterns, 8 per sample,
 * rows 16-byte aligned.  All pointers are device pointers; every call is
 * asynchronous on `stream` and allocates nothing.
 */
#define SSMD_MLP_INPUTS 8
#define SSMD_MLP_HIDDEN 64
#define SSMD_MLP_PARAMS 8961

/* Bytes of the packed (matrix-core operand order) form of a network. */
size_t ssmd_mlp_packed_bytes(void);
/* params (SSMD_MLP_PARAMS floats) -> packed. */
int ssmd_mlp_pack(const float *params, void *packed, void *stream);
/* y[c] = network(x[c]), c < n. */
int ssmd_mlp_forward(const void *packed, const uint16_t *x, float *y, uint32_t n, void *stream);
/* Workspace bytes ssmd_mlp_grad needs for n samples. */
size_t ssmd_mlp_grad_work_bytes(uint32_t n);
/* Gradient (SSMD_MLP_PARAMS floats) and value (one float) of the weighted
 * mean squared error (1/n) sum_c w[c] (network(x[c]) - t[c])^2; w = NULL
 * weighs every sample 1.  Deterministic (fixed reduction order). */
int ssmd_mlp_grad(const void *packed, const uint16_t *x, const float *t, const float *w, uint32_t n, float *grad,
                  float *loss, void *work, size_t work_bytes, void *stream);
/* ssmd_mlp_grad without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight training loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * per-block partials go to a library-owned static device array sized for
 * n <= SSMD_MLP_GRAD_NOALLOC_MAX samples (9.2 MB per device, held until the
 * process ends); larger n returns SSMD_EINVAL.  Results are bit-identical to
 * ssmd_mlp_grad.  Every call on one device shares that array: calls that
 * overlap in time on one device (other streams, other threads) race. */
#define SSMD_MLP_GRAD_NOALLOC_MAX 65536u
int ssmd_mlp_grad_noalloc(const void *packed, const uint16_t *x, const float *t, const float *w, uint32_t n,
                          float *grad, float *loss, void *stream);
/* One Adam step (Kingma & Ba 2015, bias-corrected; step counts from 1) on
 * params, with first/second moment buffers m, v (SSMD_MLP_PARAMS floats each). */
int ssmd_mlp_adam(float *params, const float *grad, float *m, float *v, uint32_t step, float lr, float beta1,
                  float beta2, float eps, void *stream);
/* Training samples of the LEARNED closure from a fine field set whose last
 * step wrote every moment (SSMD_STEP_MACROS), box-filtered by 2 onto the
 * coarse grid `coarse` (coarse interior cell i covers fine cells 2i-1 and 2i
 * per axis; requires 2 (n_c - 2) <= n_f - 2 on each axis).
 * work receives 20 coarse planes (ssmd_sgs_work_bytes): rho, Favre-filtered
 * ux uy uz, phi, the filtered strain rate S (xx yy zz xy xz yz, per coarse
 * step), the subgrid stress T = <rho u u> - <rho> u~ u~ (same order) and
 * the subgrid flux of the phase field J = <phi u> - <phi> <u> (x y z; the
 * target of the flux_coeff closure), all in coarse lattice units.  The two runs must describe one flow: with
 * dx_c = 2 dx_f, their water viscosities give the time ratio rt = dt_c / dt_f
 * = 4 nu_c / nu_f (2 for acoustic scaling, nu_f = 2 nu_c; 4 for diffusive
 * scaling, nu_f = nu_c; 1 <= rt <= 16 or SSMD_EINVAL), and fine velocities,
 * strain rates, stresses and fluxes are scaled by rt / 2, rt, (rt / 2)^2 and rt / 2.
 * Per coarse cell: feat = the 8 network inputs as the LEARNED step forms them
 * (coarse->strain_scale, grad_scale), target = 3 nu_t / coarse->tau_scale with
 * nu_t = -(dev T : S) / (2 rho S:S) the least-squares eddy viscosity, and
 * weight = S:S (0 on the ghost layer), so that the weighted squared error of
 * tau_t is the squared residual of the modelled subgrid stress. */
size_t ssmd_sgs_work_bytes(const ssmd_params *coarse);
int ssmd_sgs_samples(const ssmd_params *fine, const ssmd_fields *fd, const ssmd_params *coarse, float *work,
                     uint16_t *feat, float *target, float *weight, void *stream);
/* LEARNED closure: tau_t[c] = max(tau_floor, tau_scale * network(feat[c])) for every
 * cell (ssmd_step and the step graphs call it after the boundaries). */
int ssmd_closure_update(const ssmd_params *p, const ssmd_fields *d, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SSMD_H */
