/* sunaengine.h — C ABI of the MI355X-native deterministic PBF substep.
 *
 * The drop-in boundary for the reference engine's hot path: the JS `Engine`
 * object of src/engine.js (createEngine / loadScene / step / digest / chainHex /
 * readState / readDerived / applyImpulse / spawnFluid / debugCounters /
 * destroy) and the replay/verify loops built on it.  Plain pointers and
 * sizes only.  Every call is synchronous with respect to the caller except
 * sn_step*, which enqueue device work; every read call waits for it.
 *
 * Error convention: functions return 0 on success and a negative SN_E* code
 * on failure; sn_last_error() then holds a message starting with "[aether]"
 * for contract violations, mirroring the reference's thrown Error text.
 *
 * A "world" is one or more independent scenes resident in HBM and advanced
 * together (one scene reproduces the reference engine exactly; many scenes
 * are the batched mode).  Each scene's H_set / H_slot / chain depend only on
 * that scene's own particles, so a scene's results are identical alone and
 * inside any batch.
 */
#ifndef SUNAENGINE_H
#define SUNAENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SN_OK 0
#define SN_E_ARG -1         /* invalid argument / contract violation (reference: assert throw) */
#define SN_E_CAPACITY -2    /* exceeds maxParticles / maxCells / maxScenes */
#define SN_E_DEVICE -3      /* HIP runtime error */
#define SN_E_UNSUPPORTED -4 /* a scene reached a material path this build does not run */
#define SN_E_STATE -5       /* call out of order (e.g. step before load) */

typedef struct sn_engine sn_engine;

/* createEngine options (src/engine.js:2786-2820). */
typedef struct sn_opts {
  uint32_t max_particles; /* total particle slots over all scenes of a world */
  uint32_t max_cells;     /* total grid cells over all scenes */
  uint32_t max_scenes;    /* scenes per world (1 for the reference engine) */
  uint32_t iters;         /* density iterations per substep, 1..16 (ITERS = 4) */
  uint32_t stream_cap;    /* per-scene H ring capacity in substeps (0 = 4096) */
  uint32_t device;        /* HIP device ordinal */
  uint32_t no_debug;      /* 1 = createEngine({debug: false}) (src/engine.js:2791,2884-2894):
                             debug counters stay zero, and sweeps skip work that
                             only feeds them; hashes and state are unchanged */
  uint32_t materials;     /* 0: the friction, containment and bond passes
                             (src/sim.wgsl:1673-2235,2708-3097) switch on when a
                             loaded scene has a fluid row with friction or bonds;
                             a water-only world skips them (they are bit-inert for
                             it) and reports SN_E_UNSUPPORTED if a transition later
                             reaches such a row.  1: they run from creation.  With
                             them on, a scene holds at most 131071 particles
                             (MAX_PARTICLES_BONDED, src/engine.js:261) */
} sn_opts;

/* One scene as loadScene writes it (src/engine.js:3888-3937): n records of 8
 * int32 words {pos.x, pos.y, vel.x, vel.y, matId, flags, temp, fuel}, already
 * validated and built by the host layer, fluid slots first. */
typedef struct sn_scene {
  uint32_t n;
  uint32_t n_fluid;
  const int32_t* state8; /* n * 8 words */
  int32_t dom_w;
  int32_t dom_h;
  uint32_t cell_shift;
  const int32_t* st_acc; /* optional n * 2 words of carried surface-tension
                            acceleration (Derived.stAcc); NULL = zero */
} sn_scene;

/* createEngine: allocate every device buffer; luts = the 16010-word frozen
 * kernel-table image, mats = 64 x 48 material words (src/engine.js:2991-3376). */
int sn_create(const sn_opts* opts, const int32_t* luts, const int32_t* mats, sn_engine** out);
void sn_destroy(sn_engine* e);
const char* sn_last_error(const sn_engine* e);

/* loadScene (src/engine.js:3704-3951): replaces the world with one scene. */
int sn_load_scene(sn_engine* e, const sn_scene* scene);
/* Batched form: replaces the world with `count` independent scenes. */
int sn_load_scenes(sn_engine* e, const sn_scene* scenes, uint32_t count);

/* step(count) (src/engine.js:3963-3999): enqueue exactly `count` substeps of
 * every scene of the world.  Each substep's 8-word digest of every scene is
 * appended to that scene's H ring. */
int sn_step(sn_engine* e, uint32_t count);
/* Wait for enqueued work; returns SN_E_UNSUPPORTED if a scene left the
 * supported material paths (the run is then void and must be reloaded). */
int sn_sync(sn_engine* e);
/* Non-blocking: 0 when all enqueued work has finished, 1 while some is still
 * running (no counterpart in the reference; for host loops that keep several
 * worlds busy from one thread). */
int sn_query(sn_engine* e);

/* digest() (src/engine.js:4453-4466) of scene `scene`: acc[0..3] = H_set,
 * acc[4..7] = H_slot, computed from the current state. */
int sn_read_hash(sn_engine* e, uint32_t scene, uint32_t acc_out[8]);
/* The per-substep digests appended since the last sn_reset_stream, oldest
 * first: out holds max_substeps * 8 words, *n_out receives the count copied. */
int sn_read_hash_stream(sn_engine* e, uint32_t scene, uint32_t* out, size_t max_substeps,
                        size_t* n_out);
int sn_reset_stream(sn_engine* e);
/* chainHex() words (src/engine.js:4406). */
int sn_read_chain(sn_engine* e, uint32_t scene, uint32_t chain_out[8]);
/* readState() (src/engine.js:4286-4289): n * 8 words in slot order. */
int sn_read_state(sn_engine* e, uint32_t scene, int32_t* out, size_t words);
/* readDerived() (src/engine.js:4290): n * 16 words in slot order. */
int sn_read_derived(sn_engine* e, uint32_t scene, int32_t* out, size_t words);
/* No reference counterpart (the batched and serving worlds): tune the
 * world's launch shape on the world itself.  Each neighbour-tile width
 * (1, 2, 4, 8 cells) x hipGraph length (4, 8, 16 substeps) is timed over
 * `trial_substeps` (0 = 16) graph-replayed substeps, all starting from the
 * world advanced `lookahead` substeps (0 = now); the world is restored to
 * its state before the call afterwards, and the fastest shape is kept.
 * Results do not depend on the shape: state, streams and counters are as
 * before the call.  Writes up to `max_trials` trials, their count to
 * *n_trials and the kept shape to *chosen.  Loading scenes again resets the
 * tile width to its occupancy heuristic. */
typedef struct sn_tune_trial {
  uint32_t tile_cells;      /* neighbour tile width in cells */
  uint32_t graph_substeps;  /* substeps per captured hipGraph */
  float ms_per_substep;     /* measured on the world */
} sn_tune_trial;
int sn_autotune(sn_engine* e, uint32_t trial_substeps, uint32_t lookahead, sn_tune_trial* trials,
                uint32_t max_trials, uint32_t* n_trials, sn_tune_trial* chosen);

/* The buffer writes of restore(snap) with `derived` (src/demo_boot.js:308-321,
 * after a loadScene): n * 8 state words and, optionally, n * 16 derived words
 * in slot order.  Only the derived stAcc words (8..9) carry into the next
 * substep; the engine-side counters are kept. */
int sn_write_state_derived(sn_engine* e, uint32_t scene, const int32_t* st8, size_t st_words, const int32_t* dv16,
                           size_t dv_words);
/* readBond() (src/engine.js:4295): n * 48 words in slot order, each a packed
 * bond {partner slot:17, L0q:13, tombstone:1} (src/sim.wgsl:1590-1625);
 * zero rows in a world that does not run the material passes. */
int sn_read_bond(sn_engine* e, uint32_t scene, uint32_t* out, size_t words);
/* debugCounters() (src/engine.js:4293): 16 words. */
int sn_debug_counters(sn_engine* e, uint32_t scene, uint32_t out[16]);
/* Substeps encoded since load for `scene` (the reference's `_substep`). */
int sn_substep_count(sn_engine* e, uint32_t scene, uint32_t* out);

/* applyImpulse(cx, cy, r, ix, iy, limit) (src/engine.js:6077-6098) on the
 * exact current state of `scene`: fluid slots in ascending order whose
 * centre is within r (IEEE double test), at most `limit`, get (ix, iy) added
 * to their velocity, clamped to +-VMAX.  *hits receives the count. */
int sn_apply_impulse(sn_engine* e, uint32_t scene, int32_t cx, int32_t cy, int32_t r, int32_t ix,
                     int32_t iy, uint32_t limit, uint32_t* hits);

/* spawnFluid(list, {noWeld, ownGroup}) (src/engine.js:5901-6020) on a
 * one-scene world: k new fluid records (8 words each, flags already holding
 * the caller's FLAG_BURNING/FLAG_ANCHOR and the material-derived FLAG_DEAD /
 * FLAG_SEAL, as the host builds them) are appended after the fluid block,
 * and the world's own boundary block moves up by k slots, rewritten as the
 * reference rewrites it (position, material, temperature and fuel kept;
 * velocity and flags zero).  The engine owns the weld-group sequence: each
 * spawn takes the next 12-bit group id (loadScene restarts it), written into
 * the flags of every record whose row welds on spawn (bondGroup >= 0,
 * bondReform == 0) unless SN_SPAWN_NO_WELD is set.  SN_SPAWN_OWN_GROUP is
 * accepted and changes nothing: this build opens no weld gesture, so every
 * spawn already owns its group.  *n_out receives the new n, unchanged when
 * capacity would be exceeded (no error, as the reference). */
#define SN_SPAWN_NO_WELD 1u   /* {noWeld: true} */
#define SN_SPAWN_OWN_GROUP 2u /* {ownGroup: true} */
int sn_spawn_fluid(sn_engine* e, const int32_t* fluid8, uint32_t k, uint32_t flags, uint32_t* n_out);

/* Device-resident input schedule for batched / served runs: `count` stir
 * events {scene, substep, cx, cy, r, ix, iy, limit}, sorted by (scene,
 * substep, arming order).  Each is applied as applyImpulse immediately before
 * its scene's substep `substep` (the push timing of src/demo_boot.js:244-272). */
typedef struct sn_event {
  uint32_t scene;
  uint32_t substep;
  int32_t cx, cy, r, ix, iy;
  uint32_t limit;
} sn_event;
int sn_set_events(sn_engine* e, const sn_event* events, uint32_t count);
/* The particles moved so far by stir events and sn_apply_impulse in scene
 * `scene` since its load (a wrapping u32 count; each event contributes its
 * hit count, at most its limit).  The harness's pushHits
 * (src/demo_boot.js:137,256) for pushes run on the device. */
int sn_read_push_hits(sn_engine* e, uint32_t scene, uint32_t* out);

/* Scene activity mask for continuous batching: inactive scenes keep their
 * state and are not advanced by sn_step. */
int sn_set_active(sn_engine* e, const uint8_t* active, uint32_t count);

/* Slotted worlds for continuous batching (no counterpart in the reference,
 * whose engine holds one scene).  sn_reserve_slots replaces the world with
 * `count` empty, inactive partitions of cap_particles[s] particle slots and
 * cap_cells[s] grid cells.  sn_load_slot places one scene (validated as
 * sn_load_scene validates) into partition `slot` and activates it, with its
 * own substep counter, chain and H ring starting afresh; the other
 * partitions keep their state.  sn_set_active(..., 0) retires a partition.
 * SN_E_CAPACITY if the scene does not fit the partition. */
int sn_reserve_slots(sn_engine* e, uint32_t count, const uint32_t* cap_particles, const uint32_t* cap_cells);
int sn_load_slot(sn_engine* e, uint32_t slot, const sn_scene* scene);

/* Profiling form of sn_step (no counterpart in the reference, which has no
 * device timers): enqueue `count` substeps eagerly with a HIP event pair
 * around every launch of the kernel classes selected by `class_mask` (bit
 * c = slot c + 1; ~0u = all) on the engine's stream, wait, and return per
 * class the summed device ms and the launch count.  Slot 0 is the span from
 * the first timed launch's start to the last one's end (calls = substeps);
 * slots 1.. are: [1] predict + cell count, [2] cell scan, [3] scatter,
 * [4] canonical gather, [5] neighbour build + iteration-0 density,
 * [6] density iterations 1.., [7] solveB per colour, [8] xsph + normals +
 * conduct, [9] surface tension + thermal + hash, [10] chain fold, [11] stir
 * events, [12] work-list build, [13] applyDp per colour (finalize on the
 * last iteration). */
#define SN_PROF_SLOTS 16
int sn_profile_step(sn_engine* e, uint32_t count, uint32_t class_mask, float out_ms[SN_PROF_SLOTS],
                    uint32_t out_calls[SN_PROF_SLOTS]);

/* ============================================================ state digest
 * digestState(state, {particleCount, strideWords, hashedWords, nFluid, bond,
 * bondCap}) (src/hash/statehash.js:80-159) over DEVICE-resident words, with
 * no engine: `state` (particle_count records of stride_words u32 words, of
 * which the first hashed_words are hashed) and the optional `bond` rows
 * (bond_cap words per fluid row; NULL = no bond fold, as an absent p.bond)
 * are device pointers; state_words is the length of `state`.  acc_out is a
 * host array: acc[0..3] = H_set, acc[4..7] = H_slot.  The work runs on
 * `stream` (a hipStream_t; NULL = the null stream) and the call returns when
 * it is done.  The reference's thrown errors ("hashedWords > strideWords",
 * "state buffer too small", bond without bondCap) return SN_E_ARG; the
 * message of the last failed digest call on this thread is
 * sn_digest_last_error().  The 2D engine's H of a loaded scene is this digest
 * with stride 8 = hashed 8 over readState's words and, in a world that runs
 * the material passes, readBond's rows with bond_cap 48; the 3D H_slot is it
 * with stride 16, hashed 8. */
const char* sn_digest_last_error(void);
int sn_digest_state(const uint32_t* state, size_t state_words, uint32_t particle_count, uint32_t stride_words,
                    uint32_t hashed_words, uint32_t n_fluid, const uint32_t* bond, uint32_t bond_cap,
                    uint32_t acc_out[8], void* stream);
/* No counterpart in the reference.  sn_digest_state for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, same result and errors,
 * but the partial sums live in a fixed device array of the library (256 KiB
 * per GPU, held until the process ends) and the call makes no allocation.
 * particle_count above SN_DIGEST_NOALLOC_MAX_PARTICLES returns
 * SN_E_CAPACITY.  All calls on one GPU share that array: calls that overlap
 * on one GPU (other host threads or streams) race and must be serialised by
 * the caller. */
#define SN_DIGEST_NOALLOC_MAX_PARTICLES 8388608u
int sn_digest_state_noalloc(const uint32_t* state, size_t state_words, uint32_t particle_count,
                            uint32_t stride_words, uint32_t hashed_words, uint32_t n_fluid, const uint32_t* bond,
                            uint32_t bond_cap, uint32_t acc_out[8], void* stream);

/* ===================================================================== 3D
 * The 3D solver (3d/sim3d.wgsl, host 3d/app.js:606-743): App.runSubsteps(k)
 * and App.chain().  A 3D record is 16 int32 words {p.xyzw, v.xyzw, pr.xyzw,
 * dp.xyzw} with the material index in p.w (3d/sim3d.wgsl:46-51); the world
 * is gx x gy x gz cells of one kernel radius (2 wu).  H_set/H_slot per
 * substep is digestState over p.xyzw, v.xyzw (the reference has no 3D state
 * hash of its own); sn3_fnv_digest returns the reference's striped FNV. */
typedef struct sn3_engine sn3_engine;
typedef struct sn3_scene {
  uint32_t n;
  const int32_t* parts16; /* n * 16 words */
  uint32_t gx, gy, gz;
} sn3_scene;
/* luts3 = the 3D kernel/wall/ball/material image (3d/app.js:85-231). */
int sn3_create(const sn_opts* opts, const int32_t* luts3, uint32_t lut_words, sn3_engine** out);
void sn3_destroy(sn3_engine* e);
const char* sn3_last_error(const sn3_engine* e);
int sn3_load(sn3_engine* e, const sn3_scene* scenes, uint32_t count);
/* Params of App.writeParams (3d/app.js:606-630): stir amplitude and phase,
 * cursor ball on/off, position and reach (all Q16.16 ticks). */
int sn3_set_inputs(sn3_engine* e, int32_t stir_amp, int32_t stir_phase, uint32_t ball_on, const int32_t* ball3,
                   int32_t ball_reach);
int sn3_step(sn3_engine* e, uint32_t count);
int sn3_sync(sn3_engine* e);
int sn3_query(sn3_engine* e);   /* as sn_query */
int sn3_read_hash(sn3_engine* e, uint32_t scene, uint32_t acc_out[8]);
int sn3_read_stream(sn3_engine* e, uint32_t scene, uint32_t* out, size_t max_substeps, size_t* n_out);
int sn3_reset_stream(sn3_engine* e);
int sn3_read_chain(sn3_engine* e, uint32_t scene, uint32_t chain_out[8]);
int sn3_read_state(sn3_engine* e, uint32_t scene, int32_t* out16, size_t words);
int sn3_debug_counters(sn3_engine* e, uint32_t scene, uint32_t out[16]);
/* App.chain() (3d/app.js:717-743): the 8 FNV lanes of the striped digest. */
int sn3_fnv_digest(sn3_engine* e, uint32_t scene, uint32_t lanes_out[8]);
/* 3D stir events (this build's definition — the reference has no 3D
 * impulse): fluid slots in ascending order within r of (cx, cy, cz), at most
 * `limit`, get (ix, iy, iz) added to v, clamped to +-65536, immediately
 * before the scene's substep `substep`. */
typedef struct sn3_event {
  uint32_t scene;
  uint32_t substep;
  int32_t cx, cy, cz, r, ix, iy, iz;
  uint32_t limit;
} sn3_event;
int sn3_set_events(sn3_engine* e, const sn3_event* events, uint32_t count);
int sn3_set_active(sn3_engine* e, const uint8_t* active, uint32_t count);
/* Slotted 3D worlds, as sn_reserve_slots / sn_load_slot (no reference
 * counterpart). */
int sn3_reserve_slots(sn3_engine* e, uint32_t count, const uint32_t* cap_particles, const uint32_t* cap_cells);
int sn3_load_slot(sn3_engine* e, uint32_t slot, const sn3_scene* scene);
/* out: [0] substep, [1] predict+sort, [2] neighbour build + density 0,
 * [3] solve iterations, [4] chain fold (device ms summed over count). */
int sn3_profile_step(sn3_engine* e, uint32_t count, float out[5]);

#ifdef __cplusplus
}
#endif
#endif /* SUNAENGINE_H */
