/*
 * empi_hip.h -- C ABI of the MI355X Matching Pursuit engine.
 *
 * Plain C types only (no C++, no torch).  Two granularities:
 *
 *  1. empi_ws_*: the reference's Workspace/WorkspaceBuilder operator API
 *     (inc/classes.hpp:65-80), one call per reference virtual, so a host can
 *     run the reference's own Decomposition loop on top of it.
 *  2. empi_plan_*: the whole decomposition loop (inc/classes.hpp:89-107,
 *     src/classes.cpp:26-87, src/mmp.cpp:9-45) for a batch of signals of
 *     any lengths and variants, device-resident from the full map build to
 *     the last atom.
 *
 * Status codes map 1:1 onto the reference's exception strings
 * (empi_strerror returns the camelCase text, src/empi.cpp:169-193).
 * Handles are single-threaded; different handles may run concurrently.
 */
#ifndef EMPI_HIP_H
#define EMPI_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int empi_status;

enum {
    EMPI_OK = 0,
    EMPI_SIGNAL_IS_EMPTY = 1,        /* "signalIsEmpty"       src/classes.cpp:37-39 */
    EMPI_ENERGY_MAP_IS_EMPTY = 2,    /* "energyMapIsEmpty"    src/gabor.cpp:108-110 */
    EMPI_TOO_LONG = 3,               /* "signalFileIsTooLongForThisDecomposition" src/gabor.cpp:239-241 */
    EMPI_INVALID_ARGUMENT = 4,       /* "invalidArgument" */
    EMPI_NO_MEMORY = 5,              /* "insufficientMemory"  src/empi.cpp:189-190 */
    EMPI_INTERNAL = 6,               /* "internalError"       src/empi.cpp:191-192 */
    EMPI_DEVICE = 7,                 /* "deviceError": a HIP runtime call failed */
    EMPI_UNSUPPORTED = 8             /* "unsupportedSize": a signal beyond 2^24 samples or an FFT beyond 2^24 points, or an unsupported proposal batch */
};

/* Decomposition variants (src/empi.cpp:114-130). */
enum { EMPI_SMP = 0, EMPI_MMP1 = 1, EMPI_MMP2 = 2, EMPI_MMP3 = 3 };

/* Multichannel constraint of a workspace (inc/classes.hpp:63). */
enum { EMPI_CONSTRAINT_NONE = 0, EMPI_CONSTRAINT_MMP1 = 1 };

/* GaborWorkspaceBuilder(energyError, scaleMin, scaleMax, freqMax)
 * (inc/gabor.hpp:67-78); use 0 / INFINITY / INFINITY for the defaults of
 * src/empi.cpp:63-68. */
typedef struct {
    double energyError;
    double scaleMin;
    double scaleMax;
    double freqMax;
} empi_dict_params;

/* Atom record: Atom{type=13, params[6]} (inc/base.hpp:24-29) plus the
 * dictionary indices it was selected at. params = modulus, amplitude,
 * position [samples], scale [samples], frequency [x Nyquist], phase. */
typedef struct {
    double params[6];
    int scale;      /* scale index k */
    int freq;       /* frequency bin */
    int time;       /* time index */
    int reserved;
} empi_atom;

/* One scale of the optimal Gabor dictionary (src/gabor.cpp:228-249). */
typedef struct {
    double scale;    /* s [seconds] */
    int Nfft;
    int tCount;
    int fCount;
    int reserved;
} empi_scale_info;

/* Host-only (no device call): the dictionary GaborWorkspaceBuilder would
 * build for fs and samples.  Writes up to max_scales entries, the scale
 * count to *count and sum F*T (Workspace::getAtomCount) to *entries. */
empi_status empi_dict_info(const empi_dict_params* dict, double fs, int samples, empi_scale_info* scales,
                           int max_scales, int* count, long long* entries);

/* ------------------------------------------------------------ workspace */

typedef struct empi_ws empi_ws;

/* GaborWorkspaceBuilder::prepareWorkspace (src/gabor.cpp:213-252). */
empi_status empi_ws_create(const empi_dict_params* dict, double fs, int channels, int samples,
                           int constraint, empi_ws** out);
/* Workspace::compute: full map build of signal [channels][samples] FP64 (src/gabor.cpp:75-83). */
empi_status empi_ws_compute(empi_ws* ws, const double* signal);
/* Workspace::findBestMatch: one atom per channel into out[channels] (src/gabor.cpp:85-133). */
empi_status empi_ws_find_best(empi_ws* ws, empi_atom* out);
/* Workspace::getAtomCount: dictionary size sum F*T (src/gabor.cpp:135-141). */
size_t      empi_ws_atom_count(const empi_ws* ws);
/* Workspace::subtractAtom: subtract from residual[samples] (host FP64, updated
 * in place) and re-evaluate the touched map region (src/gabor.cpp:189-209). */
empi_status empi_ws_subtract(empi_ws* ws, const empi_atom* atom, int channel, double* residual);
void        empi_ws_destroy(empi_ws* ws);

/* Workspace::subtractAtomFromSignal (src/classes.cpp:15-22, src/gabor.cpp:143-187)
 * on a host FP64 signal; fit != 0 refits params[0..1] by least squares. */
empi_status empi_subtract_atom_from_signal(double fs, empi_atom* atom, double* signal,
                                           int samples, int fit);

/* ------------------------------------------------------------ batched plan */

/* One signal of a batch: the settings of one reference run
 * (src/empi.cpp:59-130, DecompositionSettings inc/classes.hpp:84-87). */
typedef struct {
    int channels;
    int samples;
    double fs;
    int variant;              /* EMPI_SMP .. EMPI_MMP3 */
    int iterationMax;
    double energyPercent;     /* residualEnergy = 1 - 0.01 * energyPercent */
    empi_dict_params dict;
} empi_item;

typedef struct empi_plan empi_plan;

/* Builds every dictionary (shared between items with equal parameters),
 * sizes and allocates all device buffers. */
empi_status empi_plan_create(const empi_item* items, int count, empi_plan** out);

/* Decomposes the batch.  signal: float32 samples of every item, item after
 * item, each [channels][samples]; on the device when signal_on_device != 0,
 * else host memory; NULL (with signal_on_device 0): the input staged by
 * empi_plan_stage.  stream: a hipStream_t or NULL for the null stream.
 * Returns after the work is complete. */
empi_status empi_plan_run(empi_plan* plan, const float* signal, int signal_on_device, void* stream);

/* Streamed input (the streamed host-to-device path of the reference's epoch
 * reader, src/empi.cpp:145-164, one batch ahead): copies the next batch's
 * host samples (the layout of empi_plan_run) into the plan's second input
 * buffer on an upload stream of its own and returns when the copy is done.
 * It may be called from another thread while empi_plan_run decomposes the
 * previous input on the same plan; the next empi_plan_run(plan, NULL, 0, s)
 * decomposes the staged input.  One staged input at a time. */
empi_status empi_plan_stage(empi_plan* plan, const float* signal);

/* Results of item i after empi_plan_run:
 *   atoms    [channels][iterationMax]   (row c holds counts[c] atoms)
 *   counts   [channels]
 *   energies [rows][iterationMax + 1]   residual energy before each
 *            iteration (entry 0 = total energy), then after the last
 *            subtraction, NaN-padded; rows = channels for SMP, else 1.
 * Any pointer may be NULL.  Returns the item's own status. */
empi_status empi_plan_item_result(const empi_plan* plan, int item, empi_atom* atoms, int* counts,
                                  double* energies);

/* Allocation-free decomposition of one item (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same work and results as empi_plan_create({item}) + empi_plan_run +
 * empi_plan_item_result + empi_plan_destroy, but every device buffer is carved
 * from a fixed 64 MiB scratch array that the library holds on each GPU until
 * the process ends: no hipMalloc, hipFree or stream creation, and no graph.
 * Work runs on `stream` (a hipStream_t, NULL for the null stream); returns
 * after it is complete.  Items whose buffers need more than 64 MiB (about
 * 40 channels of 5120 samples; the learned proposal is not available here)
 * return EMPI_UNSUPPORTED and print the size needed.  Calls on one GPU share
 * the scratch and keep the last call's contents: calls that overlap on one
 * GPU race. */
empi_status empi_decompose_noalloc(const empi_item* item, const float* signal, int signal_on_device,
                                   empi_atom* atoms, int* counts, double* energies, void* stream);

/* Dictionary size (sum F*T) of item i; 0 on a bad index. */
size_t      empi_plan_item_atom_count(const empi_plan* plan, int item);

/* Map entries (scale, bin, time) re-evaluated by the last run, full build
 * included; each entry covers every channel of its unit. */
double      empi_plan_entries_evaluated(const empi_plan* plan);

/* Kernel accounting.  With profiling on, every launch of the fused column
 * kernel of each sub-transform class is bracketed by HIP events on the
 * launch stream; empi_plan_kernel_stats reports, for class cls of the last
 * run: the class's sub-transform length, launches, summed kernel time and
 * the map entries (and entry-channels) it re-evaluated. */
empi_status empi_plan_set_profiling(empi_plan* plan, int on);
int         empi_plan_kernel_classes(void);
empi_status empi_plan_kernel_stats(const empi_plan* plan, int cls, int* sub_length, long long* launches,
                                   double* total_ms, double* entries, double* entry_channels);

void        empi_plan_destroy(empi_plan* plan);

/* Autotuner (no counterpart in the reference).  Each column-kernel class has
 * variants: waves per SIMD the register allocator must allow (4, 3 or 2 for
 * single-channel units, the mode's default plus 0, 1 or 2 for multichannel
 * ones), and two persistent grid sizes.  empi_plan_autotune
 * times every variant of every class on the full map build of this batch
 * (signal as for empi_plan_run; best of `repeats`, HIP events on the plan's
 * stream) and keeps the fastest; the table, empi_plan_tuning_size() ints
 * ([mode][class]: wave variant + 3 * grid variant), can be read and set.
 * On a plan with the proposal set, the single-channel classes are timed as
 * the low-bin kernels the proposal runs. */
int         empi_plan_tuning_size(void);
empi_status empi_plan_autotune(empi_plan* plan, const float* signal, int signal_on_device, int repeats,
                               void* stream);
empi_status empi_plan_get_tuning(const empi_plan* plan, int* out);
empi_status empi_plan_set_tuning(empi_plan* plan, const int* in);

/* ------------------------------------------------------------ learned proposal */

/* Region-scoring network (no counterpart in the reference): a 3-layer
 * perceptron input_dim -> 256 -> 256 -> 1 with ReLU, FP16 weights and FP32
 * accumulation, trained on the device with Adam on a mean-squared-error loss.
 * Parameters are one FP32 block: W1 [256][input_dim], b1 [256],
 * W2 [256][256], b2 [256], w3 [256], b3 [1]; passes use their FP16 copies. */
typedef struct empi_mlp empi_mlp;
empi_status empi_mlp_create(int input_dim, unsigned seed, empi_mlp** out);
void        empi_mlp_destroy(empi_mlp* mlp);
size_t      empi_mlp_param_count(const empi_mlp* mlp);
empi_status empi_mlp_get_params(const empi_mlp* mlp, float* out);
empi_status empi_mlp_set_params(empi_mlp* mlp, const float* in);
/* y[rows] from X[rows][input_dim]; both on the device when on_device != 0. */
empi_status empi_mlp_forward(empi_mlp* mlp, const float* X, int rows, float* y, int on_device);
/* The inference pass the proposal runs: inputs and the first hidden layer
 * rounded to FP16, products on the matrix cores with FP32 accumulation. */
empi_status empi_mlp_forward16(empi_mlp* mlp, const float* X, int rows, float* y, int on_device);
/* Gradients of the mean squared error over one batch (host arrays), in the
 * parameter layout; *loss receives the mean squared error. */
empi_status empi_mlp_gradients(empi_mlp* mlp, const float* X, const float* target, int rows, float* grads,
                               float* loss);
/* Adam training over shuffled minibatches of contiguous rows. */
empi_status empi_mlp_train(empi_mlp* mlp, const float* X, const float* target, int rows, int on_device,
                           int epochs, int batch, float lr, unsigned seed, float* final_loss);

/* Proposal mode of a plan (no counterpart in the reference).  A region is one
 * scale x one 256-sample block of one signal.  Subtracting an atom marks the
 * regions it overlaps (the reference's |t - tA| <= 3 (s + sA) rule) dirty.
 * Each iteration then evaluates exactly: the top_k dirty candidate regions by
 * network score (by their last exact maximum when net is NULL), and, as the
 * full-search fallback, every dirty region whose certified bound exceeds
 * P / (1 - delta), P the best evaluated product.  The bound holds for every
 * bin: it is the smaller of csnorm * (residual energy under the region's
 * windows) and a bound from the region's last exact maximum grown by the
 * atoms subtracted since (their discrete overlap with the region's atoms, by
 * Poisson summation, times the Lipschitz constant of the reference's product
 * in the map value, which is large at the first bins of each scale where the
 * product's normalisation depends on the phase).  The chosen atom is
 * therefore within (1 - delta) of the exact maximum.  Single-channel units
 * only (SMP, MMP2 sums, one-channel items); EMPI_UNSUPPORTED otherwise.
 *   mode EMPI_PROPOSAL_OFF      the plain method (default)
 *   mode EMPI_PROPOSAL_ON       as above
 *   mode EMPI_PROPOSAL_COLLECT  every dirty region is evaluated (the plain
 *                               method's atoms) and each dirty region's
 *                               network input with its new exact maximum is
 *                               recorded as a training row
 * The network must outlive the plan's runs and have input dimension
 * empi_proposal_feature_dim(). */
#define EMPI_PROPOSAL_OFF 0
#define EMPI_PROPOSAL_ON 1
#define EMPI_PROPOSAL_COLLECT 2
int         empi_proposal_feature_dim(void);
/* Host-only: the factor c of the certified bound P <= c * (residual energy
 * under the window) for a scale of s seconds with FFT length Nfft. */
double      empi_proposal_bound_factor(double scale, double fs, int Nfft);
empi_status empi_plan_set_proposal(empi_plan* plan, int mode, empi_mlp* net, int top_k, double delta);
/* Counters of the last run: [0] unit-iterations that fell back to exact
 * evaluation of violating regions, [1] violating regions evaluated,
 * [2] certified unit-iterations, [3] dirty regions checked, [4] regions
 * evaluated in the first (proposed) pass. */
empi_status empi_plan_proposal_stats(const empi_plan* plan, unsigned long long out[5]);
/* Training rows of COLLECT runs: capacity (rows kept on the device across
 * runs until read) and sampling (one dirty region in stride, by a hash),
 * then a copy to host arrays X [rows][feature_dim], target [rows] (log of
 * the region's new exact maximum over the residual energy); *rows receives
 * the number copied; reading resets the store. */
empi_status empi_plan_set_training_capacity(empi_plan* plan, size_t rows, unsigned stride);
empi_status empi_plan_training_rows(empi_plan* plan, float* X, float* target, size_t capacity, size_t* rows);

/* Test hook (no counterpart in the reference; for tests of the argmax order):
 * runs the engine's argmax keys -- per cell the packed (product, bin) key,
 * per tile the (product, bin, cell) key, and the selection's (product,
 * scale, bin, time) order -- over a planted map on the device.  scales maps;
 * map l has T[l] time indices, G[l] cells per time index, ncell[l] cells per
 * tile (at most 256) and bins_per_cell[l] bins per cell; bins holds, map
 * after map, [G][bins_per_cell] bin indices (below 2^24), products, map after
 * map, [T][G][bins_per_cell] FP32 products.  out receives (scale, bin, time)
 * of the maximum product, exact ties going to the smallest scale, then bin,
 * then time, and *best that product (zero when no product is positive). */
empi_status empi_debug_argmax(int scales, const int* T, const int* G, const int* ncell, const int* bins_per_cell,
                              const int* bins, const float* products, int out[3], float* best);

/* Camel-case reference error text for a status code. */
const char* empi_strerror(int status);

/* Library version string. */
const char* empi_version(void);

#ifdef __cplusplus
}
#endif

#endif
