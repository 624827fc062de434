/* plp_hip.h -- C ABI of the MI355X (gfx950) 3D pallet-loading decoder.
 *
 * Plain pointers, sizes and status codes only; no C++ or torch types cross
 * this boundary.  Every entry point returns PLP_OK (0) or a PLP_E_* code;
 * plp_last_error() gives the message of the last failure on the calling
 * thread.  Buffers are caller-owned host memory unless stated otherwise.
 *
 * Conventions (same as the reference, gpu/plp_common.hpp):
 *   coordinates and dimensions are int cm;
 *   an extreme point (candidate) is 3 ints (x, y, z);
 *   an orientation is 3 ints (w, d, h);
 *   a placed-box AABB is 6 ints (x, y, z, x2, y2, z2);
 *   a placement row is 7 ints (item_type, x, y, z, w, d, h), the JSON order
 *   of plp_common.hpp:602-607;
 *   an item type is 7 ints (l, lf, w, wf, h, hf, qty), file order.
 */
#ifndef PLP_HIP_H
#define PLP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define PLP_OK 0
#define PLP_E_ARG 1      /* invalid argument */
#define PLP_E_HIP 2      /* HIP runtime failure (message has the HIP error) */
#define PLP_E_NOMEM 3    /* device memory exhausted */
#define PLP_E_EPCAP 4    /* extreme-point list exceeded its capacity (reference device err 2, plp_gpu.cu:969-990) */
#define PLP_E_RANGE 5    /* container too large for packed keys (reference exit 2, plp_gpu.cu:1263-1268) */
#define PLP_E_FORMAT 6   /* instance text violates the thpack format (reference InstanceFormatError) */
#define PLP_E_BUSY 7     /* no free slot / memory for a serving request right now */

const char *plp_status_string(int status);
const char *plp_last_error(void);

/* Number of HIP devices visible to the process. */
int plp_device_count(int32_t *count);

/* ------------------------------------------------------- instance parser */
/* thpack "original" parser with the reference's error cases
 * (gpu/plp_common.hpp:68-151, plp/parser.py:67-139).  Replaces the host
 * call plp::load_instance (plp_common.hpp:143) for callers that hold text.
 * Requires exactly one instance.  types_out receives n_types * 7 ints; when
 * types_cap is too small PLP_E_ARG is returned and *n_types holds the need. */
int plp_parse_instance(const char *text, int32_t *L, int32_t *W, int32_t *H,
                       int32_t *n_types, int32_t *types_out, int32_t types_cap,
                       char *err, int32_t errlen);

/* --------------------------------------------------------------- context */
/* One device, one HIP stream, grow-only device scratch.  Not reentrant: one
 * host thread per context (the reference backends are single-threaded,
 * plp_gpu.cu:415-422). */
typedef struct plp_ctx plp_ctx;

int plp_ctx_create(int32_t device, plp_ctx **out);
/* Destroy every engine created on the context first: engines use its stream. */
void plp_ctx_destroy(plp_ctx *ctx);

/* ----------------------------------------------- seam 1: feasibility step */
/* Replaces FeasibilityBackend::check_all (gpu/plp_common.hpp:326-332), as
 * implemented by CudaBackend::check_all (gpu/plp_gpu.cu:443-502):
 * bits_out[ci * n_ors + oi] = 1 iff (cand ci, orientation oi) passes C2, C3
 * and C1 (plp_common.hpp:297-321) against the n_aabb placed boxes.  One
 * H2D + one launch + one D2H per call: the plain one-step-per-launch form. */
int plp_feas_check_all(plp_ctx *ctx, const int32_t *cand_xyz, int32_t n_cand,
                       const int32_t *ors_wdh, int32_t n_ors,
                       const int32_t *aabb6, int32_t n_aabb,
                       int32_t L, int32_t W, int32_t H, uint8_t *bits_out);

/* plp_feas_check_all without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as plp_feas_check_all, but inputs,
 * row tables and output bits are staged in a 16 MiB device array that each
 * device holds for the life of the process, and copied straight from and
 * to the caller's buffers: the call makes no hipMalloc, hipHostMalloc or
 * other allocation.  A step needing more than 16 MiB (12 B per candidate,
 * 24 B per box, 1 B per pair, plus about 1 KB) returns PLP_E_ARG.  Calls on
 * one device share the array, so calls that overlap in time on one device
 * race; the array keeps the last call's contents. */
int plp_feas_check_all_noalloc(plp_ctx *ctx, const int32_t *cand_xyz, int32_t n_cand,
                               const int32_t *ors_wdh, int32_t n_ors,
                               const int32_t *aabb6, int32_t n_aabb,
                               int32_t L, int32_t W, int32_t H, uint8_t *bits_out);

/* ------------------------------------------ seam 2: batched per-step form */
/* Replaces CudaBatchBackend::check_all_batch (gpu/plp_gpu.cu:174-255): K
 * rows that share one orientation table, each with its own candidates and
 * boxes; bits_out[k] gets n_cand[k] * n_ors bytes.  No K limit (the
 * reference caps K at 16, plp_gpu.cu:106). */
int plp_feas_check_batch(plp_ctx *ctx, int32_t K,
                         const int32_t *const *cand_xyz, const int32_t *n_cand,
                         const int32_t *ors_wdh, int32_t n_ors,
                         const int32_t *const *aabb6, const int32_t *n_aabb,
                         int32_t L, int32_t W, int32_t H, uint8_t *const *bits_out);

/* ------------------------------------------------------------ FSR on GPU */
/* Replaces plp::compute_fsr / plp::stability_score (plp_common.hpp:356-399)
 * and the device stability_kernel (plp_gpu.cu:1121-1221) for n_packs
 * packings given as placement rows (7 ints, JSON order) concatenated, with
 * pack_off[n_packs + 1] row offsets.  fsr_out (may be NULL) receives one
 * double per row; score_out one double per packing.  mu is the friction
 * coefficient (reference plp::MU, default 0.2). */
int plp_fsr_eval(plp_ctx *ctx, const int32_t *rows7, const int32_t *pack_off,
                 int32_t n_packs, double mu, double *fsr_out, double *score_out);

/* ------------------------------------------------------- device MT19937 */
/* The --rngdump probe (gpu/plp_gpu.cu:1479-1536) generalised: for each of
 * n_seeds seeds, run the op script on a device MT19937 seeded like CPython
 * random.Random(seed) (plp_common.hpp:164-250).  ops[i] = 0 -> random()
 * (double into out_f64), 1 -> getrandbits(arg[i]), 2 -> randrange(arg[i])
 * (into out_u32).  Outputs are n_seeds * n_ops, seed-major; unused slots 0. */
int plp_rng_script(plp_ctx *ctx, const uint64_t *seeds, int32_t n_seeds,
                   const int32_t *ops, const int32_t *args, int32_t n_ops,
                   double *out_f64, uint32_t *out_u32);

/* ------------------------------ seam 3: device-resident lockstep decoders */
/* Replaces CudaDeviceBackend init/run_pass/fetch (gpu/plp_gpu.cu:1225-1390)
 * and the D4' selection of run_device (plp_gpu.cu:1395-1473), extended to
 * many instances: decoder d decodes instance dec_inst[d] with
 * random.Random(dec_seed[d]) (the reference seeds row k with seed + k,
 * plp_gpu.cu:350).  Each decoder runs whole passes in one persistent
 * wavefront; RNG, selection state and the best packings stay in device
 * memory between calls. */

typedef struct {
    int32_t L, W, H;
    int32_t n_types;
    const int32_t *types;   /* n_types * 7: l, lf, w, wf, h, hf, qty */
} plp_instance_desc;

typedef struct {
    double eps;          /* eps-greedy probability (reference --eps, default 0.3) */
    double theta;        /* D4' acceptance threshold (--theta, default 0.9) */
    double mu;           /* FSR friction coefficient (--mu / PLP_MU, default 0.2) */
    int32_t use_theta;   /* 0 = --no-theta (accept every packing) */
    int32_t first_fit;   /* --first-fit */
    int32_t order;       /* 0 = volume, 1 = input (--order) */
    int32_t reserved;
} plp_params;

typedef struct {
    int32_t waves_per_decoder;   /* wavefronts per decoder block: 1, 2, 4, 8; 0 = automatic
                                    (4 for the item decoder, 1 for the block engine, or what
                                    plp_engine_autotune chose) */
    int32_t ep_cap;              /* extreme-point capacity per decoder, 0 = automatic */
    int32_t blocks_per_cu;       /* persistent blocks per CU, 0 = as many as fit */
    int32_t use_graph;           /* capture the launch sequence in a hipGraph */
    int32_t count_work;          /* 1: also keep the dead extreme points so that rows report the
                                    reference's list size (max_ep) and brute-force work
                                    (work_pair_box); costs a third list of ep_cap entries */
    int32_t box_cap;             /* placed boxes a decoder holds in LDS (block engine: items
                                    its placed blocks hold), 0 = every item of the largest
                                    instance; a decoder that would pass it is rerun with room
                                    for all (results do not depend on it) */
    int32_t verdicts;            /* item decoder's verdicts: 1 = pair states of the last two
                                    steps, reused when the type repeats, other steps scanning
                                    the boxes; 2 = one orientation-free state per extreme
                                    point, kept current as boxes are placed; 0 = automatic
                                    (2, and plp_engine_autotune tries both).  Results do not
                                    depend on it. */
    int32_t share_prefix;        /* item decoder: 2 = every run first decodes each instance's
                                    trajectory (feas[0] at every step) once, and each decoder
                                    replays its own RNG along it and resumes where its first
                                    eps pick with an index other than 0 leaves it (or takes
                                    its packing when none does); 1 = every decoder decodes
                                    from the origin; 0 = automatic: 2 when the RNG is unused
                                    or eps * (largest item count) <= 8, and the decoders are
                                    at least three rounds of the persistent grid (with
                                    fewer, a launch lasts as long as its decoders that
                                    diverge early, which the recording only delays), else 1;
                                    plp_engine_autotune tries both.  Never with count_work.
                                    Results do not depend on it. */
    int32_t trace_waves;         /* with shared prefixes: wavefronts of the recording launch's
                                    blocks (one block per instance), 4 or 8; 0 = the
                                    decoders' count, at least 4 */
    int32_t trace_verdicts;      /* with shared prefixes: the recording launch's verdict kind,
                                    as verdicts (1 pair states, 2 point states); 0 = the
                                    decoders'.  plp_engine_autotune picks both when
                                    share_prefix, trace_waves and trace_verdicts are 0 and it
                                    keeps sharing on, and returns them. */
} plp_engine_config;

typedef struct {
    int32_t n_boxes;
    int32_t fallback;     /* D4': no packing reached theta */
    int32_t n_passes;
    int32_t err;          /* 0, or PLP_E_EPCAP for a decoder whose EP list overflowed */
    int64_t volume;       /* volume_cm3 */
    double score;         /* stability_score of the selected packing */
    uint64_t seed;
    uint64_t work_pair_box;  /* sum over steps of n_cand * n_ors * n_boxes (brute-force AABB tests);
                                engine rows: only with count_work, else 0 */
    uint64_t work_live_pair_box;  /* the same over the steps and live points the kernel evaluates;
                                     engine rows: only with count_work, else 0 */
    int32_t max_ep;          /* largest extreme-point list of the reference (engine: count_work) */
    int32_t ep_need;         /* engine: the extreme-point capacity this decoder needed */
} plp_row_result;

typedef struct plp_engine plp_engine;

/* --------------------------------------------- plain one-step-per-launch */
/* The reference's default GPU form (plp_common.hpp:471-523 driving
 * CudaBackend, plp_gpu.cu:414-502): the decode loop, RNG and extreme-point
 * update run on the host; every item step makes one plp_feas_check_all call
 * (H2D + launch + D2H + sync); FSR runs through plp_fsr_eval.  Fixed pass
 * count, D4' selection as in run().  placements gets n_boxes * 7 ints (cap:
 * total item count * 7). */
int plp_plain_run(plp_ctx *ctx, const plp_instance_desc *inst, uint64_t seed,
                  int32_t n_passes, const plp_params *params, plp_row_result *row,
                  int32_t *placements);


int plp_engine_create(plp_ctx *ctx, const plp_engine_config *cfg, plp_engine **out);
void plp_engine_destroy(plp_engine *eng);

/* Upload instances and decoders; resets every decoder (RNG seeded, no
 * packing yet).  Every side must be at most 2^21 - 1 (PLP_E_RANGE
 * otherwise), the device form's range (gpu/plp_gpu.cu:1263-1268).  When some
 * instance has bitlen(L) + bitlen(W) + bitlen(H) > 31 the batch runs with
 * 64-bit extreme-point keys and support sums, else with 32-bit ones; the
 * results are the same. */
int plp_engine_load(plp_engine *eng, const plp_instance_desc *insts, int32_t n_inst,
                    const int32_t *dec_inst, const uint64_t *dec_seed, int32_t n_dec,
                    const plp_params *params);

/* ----------------------------- block-building engine (blockproof/) */
/* The reference's second search unit (blockproof/plp_block.cpp:95-278): a x b
 * bases of one item type and orientation stacked k high, planned per sweep
 * and placed at extreme points, then decomposed into item boxes for FSR. */
typedef struct {
    int32_t amax, bmax, kmax;   /* block bounds (--amax 3 --bmax 3 --kmax 4, plp_block.cpp:297) */
    int32_t topk;               /* eps-greedy shape choice among the topk largest (--topk 8) */
    int32_t pick;               /* 0 first, 1 support, 2 contact (--pick, plp_block.cpp:56) */
    int32_t reserved[3];
} plp_block_params;

/* Replaces run_block (plp_block.cpp:220-278) for many (instance, seed)
 * decoders: like plp_engine_load, but the engine then runs the block decoder
 * (fixed pass count, D4' per plp_block.cpp:228-275); run, sync, fetch, reset
 * and autotune work unchanged, and rows hold the decomposed item boxes.
 * params->eps, theta, use_theta and mu apply; order and first_fit do not
 * exist in the block engine and are ignored.  count_work is not supported. */
int plp_engine_load_blocks(plp_engine *eng, const plp_instance_desc *insts, int32_t n_inst,
                           const int32_t *dec_inst, const uint64_t *dec_seed, int32_t n_dec,
                           const plp_params *params, const plp_block_params *blocks);

/* Return every decoder to its freshly loaded state (RNG reseeded on the next
 * run, no packing kept) without re-uploading the tables. */
int plp_engine_reset(plp_engine *eng);

/* Enqueue n_passes more passes for every decoder (asynchronous). */
int plp_engine_run(plp_engine *eng, int32_t n_passes);
/* Wait for the engine's work.  The host reads one word, the count of
 * decoders that ended in an error; only when it is non-zero are the decoder
 * states copied over, and decoders whose extreme-point list outgrew the
 * capacity are rerun from their seeds at the largest capacity a block holds
 * (results do not depend on it).  Returns PLP_E_EPCAP if a decoder needs
 * more than that (the affected rows carry err; see
 * plp_engine_config.ep_cap).  Replaces the error check after the launch at
 * plp_gpu.cu:1417-1422. */
int plp_engine_sync(plp_engine *eng);
/* Largest placement count of any decoder (row stride for fetch). */
int plp_engine_max_items(plp_engine *eng, int32_t *max_items);
/* Copy results: rows[n_dec]; placements (may be NULL) gets max_items * 7
 * ints per decoder, the selected packing in JSON row order (rows past a
 * decoder's n_boxes are left as they were).  Replaces the slot-slab read of
 * CudaDeviceBackend::fetch (gpu/plp_gpu.cu:1353-1379).  The decoders write
 * their results into the engine's pinned host memory while they run, so this
 * is a host copy after the stream synchronises, with no transfer of its own. */
int plp_engine_fetch(plp_engine *eng, plp_row_result *rows, int32_t *placements);
/* plp_engine_sync then plp_engine_fetch, with the same results and errors,
 * but the copy overlaps the decode: each decoder publishes a completion word
 * in pinned memory after its row and packing, and host threads copy every
 * decoder's results as soon as its word is set, so that only the last
 * decoders' copies follow the kernel.  Replaces the per-pass slot copy of
 * run_device's loop (gpu/plp_gpu.cu:1353-1379, 1395-1437), overlapped.  With
 * more than one run enqueued since the last sync, count_work or decoders
 * rerun at a larger capacity it copies after the sync, as plp_engine_fetch
 * does. */
int plp_engine_sync_fetch(plp_engine *eng, plp_row_result *rows, int32_t *placements);
/* The same results without a copy: *rows points at n_dec records and
 * *placements (may be NULL) at n_dec * row_stride ints (decoder d's packing
 * at d * row_stride, JSON row order), both in the engine's pinned host
 * memory, valid until the next load, reset, run or destroy.  No reference
 * counterpart: the reference copies its slot slab into host vectors
 * (gpu/plp_gpu.cu:1353-1379). */
int plp_engine_results(plp_engine *eng, const plp_row_result **rows, const int32_t **placements,
                       int32_t *row_stride);
/* Milliseconds of device time of the last plp_engine_run (HIP events on the
 * engine's stream). */
int plp_engine_last_ms(plp_engine *eng, float *ms);
/* Launch shape chosen for the loaded batch. */
int plp_engine_describe(plp_engine *eng, int32_t *waves_per_decoder, int32_t *ep_cap,
                        int32_t *lds_bytes, int32_t *grid_blocks);

/* Autotuner: time candidate launch shapes (waves per decoder, EP capacity,
 * blocks per CU) on the loaded batch with n_passes passes each, keep the
 * fastest, reset the decoders and return its shape.  The decoders' results
 * do not depend on the shape. */
int plp_engine_autotune(plp_engine *eng, int32_t n_passes, plp_engine_config *best,
                        float *best_ms);

/* ------------------------------------------------------- serving loop */
/* Continuous batching: a request is one instance decoded by n_decoders
 * seeds.  Admitted requests queue; each poll or query collects finished
 * launches and then, while one of the max_inflight slots is free, launches
 * the queued requests that share the head's pass count and parameters
 * together as one multi-instance engine run (up to 65536 decoders) on that
 * slot's own stream: the requests submitted since the last poll share a
 * launch, and launches on different slots overlap on the device.  A request's
 * results depend only on its own instance, seeds and parameters.  No
 * reference counterpart (the reference runs one instance per process). */
typedef struct plp_server plp_server;

int plp_server_create(int32_t device, int32_t max_inflight, int64_t mem_budget_bytes,
                      plp_server **out);
void plp_server_destroy(plp_server *srv);
/* Admit a request; PLP_E_BUSY when the memory budget (the footprints of
 * the requests admitted and not yet returned), the device's free memory or
 * the queue bound (4096 waiting requests) does not allow it now.  *req_id
 * identifies it.  Seeds are seed0 .. seed0 + n_decoders - 1. */
int plp_server_submit(plp_server *srv, const plp_instance_desc *inst, int32_t n_decoders,
                      uint64_t seed0, int32_t n_passes, const plp_params *params,
                      int64_t *req_id);
/* Poll: *done = 1 when the request finished; then rows (n_decoders) and
 * placements (may be NULL, max_items * 7 ints per row) are filled and the
 * request is released.  *max_items receives the row stride: the request's
 * item count (sum of qty, at least 1).  Polling (and plp_server_query) also
 * collects finished launches and starts queued requests: a submitted
 * request starts at the next poll or query of any request. */
int plp_server_poll(plp_server *srv, int64_t req_id, int32_t *done, plp_row_result *rows,
                    int32_t *placements, int32_t placements_cap_rows, int32_t *max_items);
/* Requests currently admitted. */
int plp_server_inflight(plp_server *srv, int32_t *n);
/* *done = 1 when the request's decode has finished (its results wait for
 * plp_server_poll); does not release anything. */
int plp_server_query(plp_server *srv, int64_t req_id, int32_t *done);
/* Launches so far, and the requests and decoders they carried (requests
 * per launch > 1 means arrivals were coalesced). */
int plp_server_stats(plp_server *srv, int64_t *launches, int64_t *requests, int64_t *decoders);

#ifdef __cplusplus
}
#endif

#endif
