/*
 * elp_gpu.h -- C ABI of the MI355X (gfx950) ELP lattice Monte-Carlo move engine.
 *
 * Library: libelp.so (HIP, built from mc-polymer-simulation_amd/csrc).
 * Plain C types only: opaque handles, pointers, sizes.  Every entry point
 * returns 0 on success or a negative ELP_E* code; elp_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Lattice formats (host and device, row-major, t fastest):
 *   ring   : uint8  per site  index t + LT*(u + LU*v)
 *            bits 0-3 bond code to the next site (0 = empty site)
 *            bit  4   label bit of the monomer carrying the bond
 *            bit  5   label bit of the stored-length monomer
 *            bit  6   stored length (a second monomer on this site)
 *            (same site byte as the reference, gpupol2/oclpol.c:68-124)
 *   linear : uint16 per site  bits 0-3 next bond, 4-7 prev bond,
 *            bit 8 stored length, bits 9-10 labels
 *            (same site word as the reference, gpupol/oclpol.c:19-42)
 *
 * One "time unit" = one pass of the move kernel over every tile of the box,
 * with the tile grid origin given per time unit (the reference draws that
 * shift on the host: gpupol2/gpupol_cuda.cu:133-144, gpupol/gpupol_cuda.cu:99-105).
 *
 * Randomness on the device is counter-based (Philox4x32-10): key (seed,
 * stream), counter (time unit, tile, slot, block).  Results therefore depend
 * only on (lattice, schedule, origins, tau, seed, stream), never on launch
 * geometry, occupancy or replica batching.
 */
#ifndef ELP_GPU_H
#define ELP_GPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    ELP_OK = 0,
    ELP_E_ARG = -1,      /* invalid argument / unsupported schedule */
    ELP_E_HIP = -2,      /* HIP runtime error */
    ELP_E_NODEV = -3,    /* no usable GPU */
    ELP_E_STATE = -4     /* call out of order (e.g. run before upload) */
};

/* Sub-lattice of sites attempted together in one substep. */
enum {
    ELP_SUBLAT_CUBIC = 0,  /* one site per 4x4x4 cell (reference ringpol3/linpol) */
    ELP_SUBLAT_DENSE = 1   /* four sites per 4x4x4 cell: 4Z^3 + <(1,3,2)>, index 16 */
};

typedef struct elp_engine elp_engine;   /* opaque */

/* Last error message of the calling thread (never NULL). */
const char* elp_last_error(void);

/* Number of visible GPUs (0 if none); replaces the discovery in
   GPULibInit (gpupol2/gpupol_cuda.cu:6-27). */
int elp_device_count(int* n);

/*
 * Ring engine (replaces the buffers of GPULibLoadBuffers and the kernel of
 * gpupol2/ringpol3.cu).  Box LT x LU x LV; tiles tt x tu x tv (each a
 * multiple of 4 in [8, 32] dividing the box edge); n_iter iterations of
 * {TransForw, DiffuseSL, TransForw, TransBack} per time unit
 * (reference: 51, gpupol2/gpupol_init.c:76).
 */
int elp_ring_create(elp_engine** out, int device, int LT, int LU, int LV,
                    int tt, int tu, int tv, int sublattice, int n_iter);

/*
 * Linear-chain engine (replaces gpupol/gpupol_cuda.cu buffers and the kernel
 * of gpupol/linpol.cu).  n_iter iterations of {DiffuseSL, TransMove} per
 * time unit (reference: 64, gpupol/gpupol_init.c:79).
 */
int elp_lin_create(elp_engine** out, int device, int LT, int LU, int LV,
                   int tt, int tu, int tv, int sublattice, int n_iter);

int elp_destroy(elp_engine* e);

/* Copy the whole lattice host->device / device->host (synchronous).
   nbytes must equal LT*LU*LV*bytes_per_site (1 ring, 2 linear). */
int elp_upload(elp_engine* e, const void* lattice, uint64_t nbytes);
int elp_download(elp_engine* e, void* lattice, uint64_t nbytes);

/*
 * Run n_time time units.  origins[3*i+0..2] = (ot, ou, ov) tile-grid origin
 * of time unit i, each in [0, tile edge).  Time unit i uses Philox counter
 * tau0 + i.  Asynchronous on the engine's stream: follow with elp_sync (or
 * elp_download) before reading results.
 */
int elp_run(elp_engine* e, int n_time, const int32_t* origins, uint64_t tau0,
            uint32_t seed, uint32_t stream);

/* Wait for all work on the engine's stream. */
int elp_sync(elp_engine* e);

/* GPU time (ms) of the move kernels launched by the last elp_run, measured
   with HIP events on the engine's stream (valid after elp_sync). */
int elp_last_run_ms(elp_engine* e, float* ms);

/* Kernel launch geometry chosen for this engine (for reports). */
int elp_launch_info(elp_engine* e, int* threads_per_block, int* blocks_per_time_unit, int* lds_bytes);

/* Move counting (no reference counterpart).  With enable = 1 the following
   elp_run calls launch the counting build of the kernel, which moves exactly
   as the plain one and classifies every attempt: at an empty site; leaving
   the tile (the move's sites cross a tile face after its random gates
   passed); accepted; or rejected otherwise.  enable = 1 also clears the
   counts; 0 returns to the plain kernel.  elp_move_counts gives, since the
   last enable: counts[0] attempts, [1] at empty sites, [2] leaving the tile,
   [3] accepted (it waits for the engine's stream). */
int elp_count_moves(elp_engine* e, int enable);
int elp_move_counts(elp_engine* e, uint64_t counts[4]);

/* Threads per workgroup (each thread runs tile slots / threads slots of a
   substep); ELP_E_ARG if that launch shape is not compiled in.  Results do
   not depend on it. */
int elp_set_threads(elp_engine* e, int threads);

/*
 * Move rates of the ring engine.
 *   ELP_RULES_REFERENCE (default): ringpol3's rates (gpupol2/ringpol3.cu:92-218).
 *     A TransForw that splits stored length off a bond whose two ends both
 *     hold it has rate 1/4 per iteration, the forced TransBack that undoes it
 *     rate 1; the move set is not reversible, and its equilibrium depends on
 *     the schedule (cubic vs dense sub-lattice).
 *   ELP_RULES_BALANCED: the forced TransBack is accepted with probability 1/4
 *     (random bits 1-2 of its byte zero), and every substep of every tile
 *     draws its move from bits 6-7 of the tile's site-offset byte
 *     (TransForw 1/2, DiffuseSL 1/4, TransBack 1/4) instead of the fixed
 *     sequence.  Every move then has the probability of its reverse within
 *     one substep, each substep preserves the uniform measure over lattice
 *     states, and every schedule samples it.
 * The linear engine has the reference rates only.
 */
enum { ELP_RULES_REFERENCE = 0, ELP_RULES_BALANCED = 1 };
int elp_set_rules(elp_engine* e, int rules);

/*
 * Batched replicas (no reference counterpart: the reference runs one
 * independent system per process or per device, gpupol2/gpupol_init.c:14-17,
 * gpupol2/gpupol2.c:52-55).  n_rep independent lattices of any boxes
 * boxes[3*r+0..2] (each divisible by the common tile tt x tu x tv) share one
 * device buffer; one launch per time unit runs every tile of every replica,
 * and elp_batch_run can capture units_per_graph consecutive launches in a
 * HIP graph that it replays.  Replica r evolves exactly as an elp_engine of
 * the same box and schedule run with the same origins, tau and key.
 */
typedef struct elp_batch elp_batch;     /* opaque */

int elp_ring_batch_create(elp_batch** out, int device, int n_rep, const int32_t* boxes,
                          int tt, int tu, int tv, int sublattice, int n_iter);
int elp_lin_batch_create(elp_batch** out, int device, int n_rep, const int32_t* boxes,
                         int tt, int tu, int tv, int sublattice, int n_iter);
int elp_batch_destroy(elp_batch* b);

/* Copy replica rep's lattice host->device / device->host (synchronous). */
int elp_batch_upload(elp_batch* b, int rep, const void* lattice, uint64_t nbytes);
int elp_batch_download(elp_batch* b, int rep, void* lattice, uint64_t nbytes);

/*
 * Run n_time time units of every replica.  origins[3*(i*n_rep + r) + 0..2] =
 * tile-grid origin of replica r in time unit i; replica r uses Philox counter
 * tau0[r] + i and key (keys[2r], keys[2r+1]).  units_per_graph > 0 captures
 * that many launches in a HIP graph (kept for later calls) and replays it
 * n_time / units_per_graph times; the remainder is launched directly.
 * 0 launches every time unit directly.  Asynchronous on the batch's stream.
 */
int elp_batch_run(elp_batch* b, int n_time, const int32_t* origins, const uint64_t* tau0,
                  const uint32_t* keys, int units_per_graph);
int elp_batch_sync(elp_batch* b);
int elp_batch_last_run_ms(elp_batch* b, float* ms);
int elp_batch_info(elp_batch* b, int* threads_per_block, int* blocks_per_time_unit, uint64_t* sites);
int elp_batch_set_threads(elp_batch* b, int threads);
int elp_batch_set_rules(elp_batch* b, int rules);
/* Move counting over all replicas, as elp_count_moves / elp_move_counts. */
int elp_batch_count_moves(elp_batch* b, int enable);
int elp_batch_move_counts(elp_batch* b, uint64_t counts[4]);

/*
 * Schedule autotuner (no reference counterpart: the reference fixes its
 * tiles and iterations at compile time, gpupol2/gpupol_init.c:76).
 * A schedule = tile shape, sub-lattice, iterations per time unit, threads
 * per workgroup and time units per graph launch; every choice of it gives a
 * trajectory that the C replay reproduces for that schedule.
 */
typedef struct elp_schedule {
    int tt, tu, tv;
    int sublattice;
    int n_iter;
    int threads;
    int units_per_graph;
    double attempts_per_s;   /* measured by the tuner */
    int rules;               /* ELP_RULES_* the schedule was tuned for */
    double accepted_per_s;   /* accepted moves per second, measured by the tuner */
} elp_schedule;

/* Tune a batched schedule on scratch copies of lattices[r] (kind 0 ring, 1
   linear; the lattices are not modified): every compiled configuration whose
   tile divides all boxes is timed over time_units time units and run again
   with move counting, and ranked by accepted moves per second; then the time
   units per graph launch of the winner are timed.  Candidates must sample the
   same statics: with ELP_RULES_REFERENCE (not reversible, so the stationary
   law depends on the sub-lattice) only the given sublattice competes; with
   ELP_RULES_BALANCED sublattice = -1 lets both compete.  log (optional)
   receives one line per candidate. */
int elp_autotune(int device, int kind, int n_rep, const int32_t* boxes, const void* const* lattices,
                 int time_units, int rules, int sublattice, elp_schedule* best, char* log, uint64_t log_len);
/* One-line text record of a schedule, and its parser (round trip exact
   except the two rates, printed with 7 significant digits; records without
   rules / accepted_per_s parse as reference rates, 0). */
int elp_schedule_format(const elp_schedule* s, char* buf, uint64_t len);
int elp_schedule_parse(const char* text, elp_schedule* s);

/*
 * Ring recovery on the device (replaces the host walk GetAllRingPolymers /
 * FindPolymerStart / GetRingPolymer / UpdatePolymerWithLabels,
 * gpupol2/oclpol.c:183-376, and the writer WriteLatticeFile,
 * gpupol2/gpupol_io.c:167-203, with byte-identical output).  label2id: the
 * host's table of 2^18 entries (18 id bits of a ring -> polymer id, -1 if
 * unused); n_pol rings of at most max_len monomers.
 * elp_ring_chains: per polymer id its bond string (0 = stored length) in
 *   bonds[id*max_len ..], its length and its start site (3 ints); any output
 *   may be NULL.
 * elp_ring_tuv: the whole tuv file text (header with the given maxPolLength,
 *   then one line per ring in id order); with text == NULL only *len is set.
 * Both return ELP_E_STATE when the lattice does not hold n_pol labelled rings.
 */
int elp_ring_chains(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds,
                    int32_t* lens, int32_t* starts);
int elp_ring_tuv(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int header_max_pol_length,
                 char* text, uint64_t cap, uint64_t* len);
/* The same for a linear-chain engine (replaces GetAllRingPolymers /
   FindPolymerStart / GetLinPolymer / UpdatePolymerWithLabels,
   gpupol/oclpol.c:187-287 and :700-730, and WriteLatticeFile,
   gpupol/gpupol_io.c): a chain starts on its monomer without a prev bond;
   per site its stored-length monomer (bond 0) comes before its carrier; the
   last monomer's bond is 0xf. */
int elp_lin_chains(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, uint8_t* bonds,
                   int32_t* lens, int32_t* starts);
int elp_lin_tuv(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int header_max_pol_length,
                char* text, uint64_t cap, uint64_t* len);
/* Per polymer id, from the same device-side walk: squared radius of gyration
   in bond units (analysis/mode_compute.c:191-217) and the number of
   stored-length bonds (ComputeSL, :145-153); either output may be NULL. */
int elp_ring_chain_stats(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, double* rg2,
                         int32_t* n_sl);

/* One frame of the statistics of analysis/mode_compute.c reduced on the
   device, per polymer id c (same walk as elp_ring_chains):
     start3[3c..]  ring start site (t,u,v) in the box,
     sum3[3c..]    sum over its monomers of their unwrapped (t,u,v) relative
                   to the start (the chain's centre of mass is start + sum/n),
     mid3[3c..]    unwrapped (t,u,v) of monomer n/2 relative to the start,
     rg2[c], n_sl[c] as elp_ring_chain_stats (either may be NULL),
   and contacts[g], g < max_len: number of monomer pairs g apart along a ring
   on one site or on nearest-neighbour sites in the ring's own unwrapped frame,
   summed over rings (ContactProbability, :429-504; may be NULL).  Rings of up
   to 8192 monomers.  libelp_stats turns a series of these frames into the
   statics and mean-square displacements (elp_series_new_summary). */
int elp_ring_frame_summary(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                           int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts);
/* The same frame summary of a linear-chain engine (chains recovered as
   elp_lin_chains; the middle monomer is (n-1)/2, the analysis' polSize/2 with
   polSize = n-1 bonds, and a chain's n-1 bonds are its stored-length
   candidates, analysis/mode_compute.c:145-153). */
int elp_lin_frame_summary(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                          int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts);
/* The same frame summary of replica `rep` of a batch (ring or linear, as the
   batch), read in place in the batch's lattice buffer on the batch's stream:
   a series of a batched run needs no lattice download per frame. */
int elp_batch_frame_summary(elp_batch* b, int rep, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                            int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts);

/* elp_ring_frame_summary without allocation.  No reference counterpart: for
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving or analysis loops).  Same arguments and results; every
   device buffer of the computation (label table, ring starts, counters, bond
   strings, per-ring results, contact histogram) is carved from one static
   32 MiB device array per GPU, so n_pol rings of up to max_len monomers must
   fit: about 1 MiB + n_pol * (max_len + 72) + 8 * max_len bytes (ELP_E_ARG
   beyond that; e.g. L=128 N=1000, 2517 rings, takes 3.8 MB).  The array is
   shared by every call on that GPU and keeps the last call's contents (what
   must start at zero is cleared by a kernel); calls that overlap on one GPU
   (other threads, other engines' streams) race. */
int elp_ring_frame_summary_noalloc(elp_engine* e, const int32_t* label2id, int n_pol, int max_len, int32_t* start3,
                                   int64_t* sum3, int32_t* mid3, double* rg2, int32_t* n_sl, uint64_t* contacts);

/*
 * Learned proposal for ring equilibration (no reference counterpart).
 * A small network (144 int8 features of a site's 5^3 neighbourhood and bond
 * -> 64 -> 64 -> 11 move logits; bf16 weights on MFMA, fp32 accumulation,
 * fp32 master parameters) proposes one of the 11 moves of a site; a
 * Metropolis-Hastings test with the reverse proposal's probability keeps the
 * uniform measure (the equilibrium of ELP_RULES_BALANCED) stationary.
 * Parameter vector (14416 floats): W1[144][64] W2[64][64] b2[64] W3[64][16] b3[16].
 * Moves m of a site P: 0-3 TransForw through common neighbour m (stored
 * length of P if it has any, else of the partner), 4-7 TransForw moving the
 * partner's stored length while P holds stored length too, 8 TransBack into P,
 * 9 TransBack into the partner, 10 DiffuseSL.
 */
typedef struct elp_net elp_net;     /* opaque */

int elp_net_create(elp_net** out, int device, uint32_t seed);   /* Glorot-uniform init */
int elp_net_destroy(elp_net* n);
int elp_net_params(elp_net* n, float* params, uint64_t count);
int elp_net_set_params(elp_net* n, const float* params, uint64_t count);
/* Move probabilities of count feature rows (x: count x 144 int8, prob: count x 11). */
int elp_net_forward(elp_net* n, const int8_t* x, int count, float* prob);
/* Gradient of the mean cross-entropy of moves y (count) given features x;
   grad: 14416 floats, loss: mean loss. */
int elp_net_gradient(elp_net* n, const int8_t* x, const int32_t* y, int count, float* grad, float* loss);
/* Append count training samples drawn from the ring engine's lattice: a random
   site with at least one possible move, and a move drawn with the probability
   the plain engine (balanced rates) makes it there. */
int elp_net_harvest(elp_net* n, elp_engine* e, int count, uint32_t seed);
/* Number of samples, and optionally a copy of them (x: count x 144, y: count). */
int elp_net_samples(elp_net* n, int* count, int8_t* x, int32_t* y);
/* Adam (beta 0.9 / 0.999) over the samples in mini-batches; *last_loss = mean
   loss of the last epoch. */
int elp_net_train(elp_net* n, int epochs, int batch, float lr, float* last_loss);
/* n_time time units of the learned-proposal engine on a ring engine with 16^3
   tiles and the cubic sub-lattice (n_iter iterations of 4 substeps, one
   proposal per slot per substep); synchronous; counts proposals and
   accepted moves. */
int elp_run_learned(elp_engine* e, elp_net* n, int n_time, const int32_t* origins, uint64_t tau0,
                    uint32_t seed, uint32_t stream, uint64_t* proposals, uint64_t* accepted);
/* Test entry points: move m at the centre of 5^3 patches (0xFF = outside the
   tile), and the features of patches. */
int elp_net_patch_moves(int device, const uint8_t* patches, const int32_t* moves, int count, uint8_t* out,
                        int32_t* ok, int32_t* rev);
int elp_net_patch_features(int device, const uint8_t* patches, int count, int8_t* x);
/* Test entry point: checks the learned engine's tile path (features from the
   occupancy bitmap, moves applied to a 16^3 tile, bitmap flips, revert)
   against the patch functions above; cases = count x (tile, t, u, v, move),
   result = count x (failed-check bits, move possible). */
int elp_net_tile_check(int device, const uint8_t* tiles, int n_tiles, const int32_t* cases, int count,
                       int32_t* result);

/* Diagnostic (no reference counterpart): device-to-device copy of nbytes
   (multiple of 16) with a plain streaming kernel, reps times; *ms = average
   time per copy.  Used to calibrate HBM byte counters and to measure the
   achievable copy bandwidth next to the published peak. */
int elp_bench_copy(int device, uint64_t nbytes, int reps, float* ms);

#ifdef __cplusplus
}
#endif
#endif
