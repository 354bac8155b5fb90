/*
 * nebula_hip.h -- C ABI of libnebula_hip.so, the MI355X (gfx950) SPH step.
 *
 * Every entry point is stream-ordered: device pointers in, device pointers out, on the
 * stream given (a hipStream_t passed as void*; NULL = the default stream). Pointers to
 * device memory are owned by the caller; the library owns only the scratch space inside an
 * nb_ctx. Entry points that must report a size to the host (list totals, reductions)
 * synchronize the stream before returning.
 *
 * Return codes: 0 = ok; 1 = NB_RETRY_CAPACITY (output buffer too small, required size
 * reported, state kept for a follow-up call); 2 = NB_NOT_GRID_ORDER (nb_hydro_density64 only);
 * negative = error, message from
 * nb_last_error(ctx): -1 bad argument / failure, -2 HIP allocation, -3 kernel support larger
 * than half a periodic box (the reference's ValueError), -4 nothing built.
 *
 * Layouts: positions, velocities, accelerations (N,3) float64 row-major; per-particle
 * scalars (N,) float64; row subsets ``rows`` (R,) int64 or NULL for all particles (results
 * then come back in rows order, as in the reference). Neighbour lists (nb_lists): rows sorted
 * ascending and including the particle itself, either CSR (int64 offsets (N+1), int32
 * indices: the reference's Neighbours arrays) or 64-row blocks (int32 lengths (N) and entries
 * at ((i / 64) * row_cap + k) * 64 + i % 64, the layout the step uses internally: lanes of a
 * wavefront walking their own rows read contiguous memory).
 *
 * Each entry point names the reference function it replaces (paths in the reference
 * repository: src/nebula/sim/...). The Python binding is blackhole-nebula-simulator_amd/
 * nebula/_hip.py; the reference-named wrappers are nebula/sim/{neighbours,sph,integrator,
 * fastphys,coupled}.py.
 */
#ifndef NEBULA_HIP_H
#define NEBULA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nb_ctx nb_ctx;

/* Shock-capturing and kernel settings of the force pass (sph.py:409-432 ShockCapturing,
 * sph.py:655-690 argument meaning). */
typedef struct nb_force_params {
    int32_t kid;          /* 0 cubic spline, 1 Wendland C4 (kernels_nb.py:12) */
    int32_t use_visc;     /* alpha and c_s given */
    int32_t use_cond;     /* u given as well */
    int32_t cond_signal;  /* 0 pressure, 1 velocity (sph.py:404-405) */
    int32_t use_recon;    /* velocity reconstruction (needs grad_v) */
    int32_t closeness;
    int32_t gate;         /* 0 both, 1 published, 2 min-mod (sph.py:406) */
    int32_t precision;    /* 0 FP64; 1 mixed (FP32 pair terms, compensated sums; FP64 with use_recon) */
    double cnorm;
    double beta;
    double alpha_u;
    double eta_crit;
    double eta_fold;
} nb_force_params;

/* A neighbour list: CSR when length == NULL; else rows in 64-row blocks (row_major 0), of
 * capacity row_cap (offset NULL: entry k of row i at ((i / 64) row_cap + k) 64 + i % 64) or of
 * capacities of their own (offset = N / 64 + 1 block starts, the FP64 step's layout: entry k of
 * row i at offset[i / 64] + 64 k + i % 64), or rows of their own capacity (row_major 1: entry k
 * of row i at offset[i] + k, k < length[i]; the mixed-precision step's layout, rows not sorted,
 * row_cap unused).
 * n_entries: CSR offset[N], or the size of the entry array (the size of per-entry scratch).
 * n_rows: N, the particles the lists cover (0 = not given; the mixed-precision passes need it
 * when they run on a row subset). */
typedef struct nb_lists {
    const int64_t *offset;
    const int32_t *index;
    const int32_t *length;
    int64_t row_cap;
    int64_t n_entries;
    int32_t row_major;
    int32_t n_rows;
} nb_lists;

/* ------------------------------------------------------------------ context */
nb_ctx *nb_ctx_create(int32_t device);
void nb_ctx_destroy(nb_ctx *ctx);
const char *nb_last_error(const nb_ctx *ctx);
int32_t nb_abi_version(void);
/* bytes of scratch the context currently holds */
int64_t nb_ctx_scratch_bytes(const nb_ctx *ctx);
/* Scratch generation: incremented whenever the context frees a scratch buffer to grow it.
 * A hipGraph recorded over library calls holds the scratch addresses of its recording, so it
 * may be replayed only while this number is what it was when the recording ended (any eager
 * call in between may have grown a buffer). No counterpart in the reference (its loops are
 * numba; nothing is recorded). */
int64_t nb_ctx_generation(const nb_ctx *ctx);
/* Release every scratch buffer of the context (after synchronising its device), for a caller
 * that wants the memory back between phases of a run. Passes that read an earlier pass's
 * scratch (lists, the fixed-capacity reservation) must be redone; the generation advances.
 * No counterpart in the reference. */
int32_t nb_ctx_trim(nb_ctx *ctx);
/* Workgroup size (64, 128 or 256 threads = 4, 8 or 16 rows) of the mixed-precision lane-group
 * kernels of this context (no counterpart in the reference; set by the autotuner). */
int32_t nb_set_tile(nb_ctx *ctx, int32_t tile);
/* Kernel timing (no counterpart in the reference; for benchmarks): with enable != 0 the library
 * records HIP events around its main kernels, on the stream each runs on, and accumulates their
 * durations; every call clears the totals. */
int32_t nb_timing(nb_ctx *ctx, int32_t enable);
/* The accumulated kernel times as "kernel launches total_ms" lines in buf (cap bytes,
 * NUL-terminated; waits for the pending events). Returns the bytes the full report needs. */
int64_t nb_timing_report(nb_ctx *ctx, char *buf, int64_t cap);

/* ------------------------------------------------------------------ neighbour lists */
/* find_neighbours_numba (neighbours.py:240-281): symmetric lists {j : |x_i - x_j| <
 * 2 max(h_i, h_j)} at search lengths h. Writes out_offset (N+1) always; writes out_index when
 * its capacity holds the total, else returns NB_RETRY_CAPACITY with *total_out set and keeps
 * the lists for nb_lists_emit. box3: periodic lengths (host array of 3, 0 = open) or NULL. */
int32_t nb_lists_build(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n,
                       const double *box3, int64_t *out_offset, int32_t *out_index, int64_t capacity,
                       int64_t *total_out);
int32_t nb_lists_emit(nb_ctx *ctx, void *stream, int32_t *out_index, int64_t capacity);

/* The same lists in blocked rows of capacity row_cap (lengths (N), entries (64 row_cap ceil(N/64))).
 * *need_cap_out = longest row; NB_RETRY_CAPACITY (lists kept) if it exceeds row_cap, then call
 * nb_lists_emit_blocked with a large enough buffer. *total_out = number of entries. */
int32_t nb_lists_build_blocked(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n,
                               const double *box3, int32_t *out_len, int32_t *out_index, int64_t row_cap,
                               int64_t *need_cap_out, int64_t *total_out);
int32_t nb_lists_emit_blocked(nb_ctx *ctx, void *stream, int32_t *out_len, int32_t *out_index, int64_t row_cap);
/* statistics of the last build (host): [entries, longest row, rows that needed sorting, grid cells] */
int32_t nb_lists_stats(const nb_ctx *ctx, int64_t *out4);
/* The particle order of the neighbour grid for these positions and search lengths: particles by
 * cell (raster order), ascending index inside a cell. Storing particles in this order makes the
 * gathered rows come out sorted. (A layout choice of this implementation: the reference keeps
 * the caller's order.) cell_factor scales the cell edge (1 for the FP64 passes; the mixed passes
 * may take another, which their density pass must then use too). */
int32_t nb_grid_order(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n, const double *box3,
                      double cell_factor, int32_t *order_out);
/* The same for an ensemble of independent systems (no counterpart in the reference): system k holds
 * particles seg_off[k] .. seg_off[k+1] - 1 (seg_off: host array of nseg + 1), cloud[i] (device) the
 * system of particle i; each system gets its own grid, cells numbered system after system. */
int32_t nb_grid_order_seg(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n,
                          const int32_t *cloud, const int64_t *seg_off, int32_t nseg, double cell_factor,
                          int32_t *order_out);
/* dst row i = src row order[i], rows of `width` doubles (or int64 values). */
int32_t nb_permute(nb_ctx *ctx, void *stream, const double *src, const int32_t *order, int64_t n, int32_t width,
                   double *dst);
int32_t nb_permute_i64(nb_ctx *ctx, void *stream, const int64_t *src, const int32_t *order, int64_t n, int64_t *dst);
/* The same for count (<= 16) arrays of 8-byte values (doubles or int64) in one launch: dst[k] row i
 * = src[k] row order[i], rows of width[k] values (src, dst, width: host arrays of count). */
int32_t nb_permute_many(nb_ctx *ctx, void *stream, int32_t count, const void *const *src, void *const *dst,
                        const int32_t *width, const int32_t *order, int64_t n);
/* nb_permute_many where some destinations are kicks of the moved elements (no counterpart in
 * the reference: coupled.py:231-233's predictor kick fused into the step's reordering):
 * src2[k] (device, may be NULL per array) makes dst[k] = src[k] + s src2[k] as doubles, with
 * s = scale * s_dev[0] read on the device; the values of nb_axpy_dev followed by the move. */
int32_t nb_permute_kick_many(nb_ctx *ctx, void *stream, int32_t count, const void *const *src, const void *const *src2,
                             void *const *dst, const int32_t *width, const double *s_dev, double scale,
                             const int32_t *order, int64_t n);

/* Reuse of cached neighbour lists (sph.py:62-147, NeighbourCache / _reuse_limits /
 * _fast_rows_complete): the exact test that lists built at x_ref with search lengths h_search
 * still hold every interacting pair.
 * nb_reuse_displacements: disp[i] = |minimum_image(pos_i - x_ref_i)| (box3 NULL: open).
 * nb_reuse_limits: h_eff = h_search - (disp + d_rest) / 2; bad_dev (two int32, zeroed by the
 *   caller) counts solved rows (role bit 0) with h_eff <= h_guess / 2 and the other particles
 *   (role bit 1) with h_guess > h_eff (sph.py:134-137); h_lim = h_guess for role bit 1, else
 *   h_eff (sph.py:142-143).
 * nb_reuse_fast_rows: for each particle j of fast (int64, nfast) and each k with
 *   r_jk <= 2 max(h_lim_j, h_lim_k) now, k must be in j's sorted CSR row; *missing_dev (int32,
 *   zeroed by the caller) counts the pairs that are not (sph.py:86-111). */
int32_t nb_reuse_displacements(nb_ctx *ctx, void *stream, const double *pos, const double *x_ref, const double *box3,
                               int64_t n, double *disp);
int32_t nb_reuse_limits(nb_ctx *ctx, void *stream, const double *h_search, const double *disp, double d_rest,
                        const double *h_guess, const uint8_t *role, int64_t n, double *h_eff, double *h_lim,
                        int32_t *bad_dev);
int32_t nb_reuse_fast_rows(nb_ctx *ctx, void *stream, const double *pos, const int64_t *fast, int64_t nfast,
                           const double *h_lim, const double *box3, const int64_t *offset, const int32_t *index,
                           int64_t n, int32_t *missing_dev);

/* ------------------------------------------------------------------ block time steps
 * Individual power-of-two steps on an int64 timeline (blocks.py, timesteps.py:21-88; levels
 * k <= 40, a step of 2^(40 - k) ticks). Rows are int64 particle indices, levels int64.
 * nb_block_elapsed: s = (n_next - n_beg) tick and L/2 = 0.5 (n_end - n_beg) tick per particle.
 * nb_block_predict: the predictions of a system step at tick n_next (blocks.py:502-530): p_out =
 *   p_flow + (s - L/2) a, u_out = u_half + (s - L/2) du/dt, with beta cooling (beta finite; floor
 *   u_floor, or u_floors per particle) of u_half over s at the mid radius; the active particles
 *   (n_end = n_next) keep their cooled u_half and get e_part = m du (0 elsewhere), luminosity
 *   and heat_cool.
 * nb_block_close: closing kicks of the rows (blocks.py:543-546) and their heat budget.
 * nb_block_levels: new levels of the rows (blocks.py:318-335, 367-372): step criteria (hydro
 *   0: orbit only), level_for_dt, forced offset, allowed_level; nu may be NULL; eta_orb <= 0
 *   without the hole. bad_dev[0] counts criteria not positive, bad_dev[1] levels beyond 40.
 * nb_block_limit: one sweep of the neighbour limiter over the rows (blocks.py:338-353) on CSR
 *   lists; *changed_dev set when a level rose (repeat until it stays 0).
 * nb_block_land: levels raised until steps opened at n_now end by n_stop (blocks.py:383-393);
 *   *bad_dev counts levels beyond 40.
 * nb_block_wake: wake-up (blocks.py:357-365, 396-431): need (N, filled with -1 by the caller)
 *   gets max over the rows' list entries of level - gap; running steps longer than need allows
 *   are cut, their opening kicks corrected; *woken_dev counts the cuts.
 * nb_block_open: opening kicks of the rows at n_now (blocks.py:443-452).
 * nb_block_mark: mark[i] = 1 for the rows and every entry of their lists (blocks.py:202-214). */
int32_t nb_block_elapsed(nb_ctx *ctx, void *stream, const int64_t *n_beg, const int64_t *n_end, int64_t n_next,
                         double tick, int64_t n, double *s_out, double *half_out);
int32_t nb_block_predict(nb_ctx *ctx, void *stream, const int64_t *n_beg, const int64_t *n_end, int64_t n_next,
                         double tick, const double *p_flow, const double *a, double *u_half, const double *dudt,
                         const double *x_beg, const double *x, const double *mass, double u_floor,
                         const double *u_floors, double beta, int64_t n, double *p_out, double *u_out,
                         double *e_part, double *lum, double *heat_cool);
int32_t nb_block_close(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *n_beg,
                       const int64_t *n_end, double tick, const double *p_flow, const double *a, const double *u_half,
                       const double *dudt, const double *visc, const double *cond, const double *disk, double *p,
                       double *u, double *heat_visc, double *heat_cond, double *heat_disk, double *heat_pdv);
int32_t nb_block_levels(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const double *x,
                        const double *h, const double *vsig, const double *a, const double *nu, double courant,
                        double c_force, double eta_orb, int32_t hydro, double dt_max, const int64_t *level_offset,
                        int64_t n_now, int64_t *level, int32_t *bad_dev);
int32_t nb_block_limit(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                       const int32_t *index, int64_t gap, int64_t *level, int32_t *changed_dev);
int32_t nb_block_land(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, int64_t n_now, int64_t n_stop,
                      int64_t *level, int32_t *bad_dev);
int32_t nb_block_wake(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                      const int32_t *index, int64_t gap, int64_t n, int64_t n_now, double tick, int64_t *need,
                      int64_t *level, int64_t *n_end, const double *a, const double *dudt, const double *visc,
                      const double *cond, const double *disk, double *p_half, double *u_half, double *heat_visc,
                      double *heat_cond, double *heat_disk, double *heat_pdv, int32_t *woken_dev);
int32_t nb_block_open(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, int64_t n_now, double tick,
                      const int64_t *level, int64_t *n_beg, int64_t *n_end, const double *x, const double *p,
                      const double *u, const double *a, const double *dudt, const double *visc, const double *cond,
                      const double *disk, double *x_beg, double *p_half, double *u_half, double *heat_visc,
                      double *heat_cond, double *heat_disk, double *heat_pdv);
int32_t nb_block_mark(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                      const int32_t *index, uint8_t *mark);

/* tighten (neighbours.py:306-323): keep pairs with |x_i - x_j| < 2 max(h_i, h_j); rows not in
 * ``rows`` are left empty. Same capacity protocol (the input list total always suffices). */
int32_t nb_tighten(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n,
                   const double *box3, const int64_t *offset, const int32_t *index, const int64_t *rows,
                   int64_t nrows, int64_t *out_offset, int32_t *out_index, int64_t capacity,
                   int64_t *total_out);
/* nb_tighten with the boundary included (|x_i - x_j| <= 2 max(h_i, h_j)): the pair set of the
 * reference's cKDTree twin find_neighbours (neighbours.py:53-71; query_ball_point is inclusive),
 * taken from lists built at a slightly larger search length. */
int32_t nb_tighten_inclusive(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n,
                             const double *box3, const int64_t *offset, const int32_t *index, const int64_t *rows,
                             int64_t nrows, int64_t *out_offset, int32_t *out_index, int64_t capacity,
                             int64_t *total_out);

/* tighten of blocked rows (all particles): one pass into blocked rows of the same capacity;
 * *total_out = entries kept. */
int32_t nb_tighten_blocked(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n,
                           const double *box3, const int32_t *in_len, const int32_t *in_index, int64_t row_cap,
                           int32_t *out_len, int32_t *out_index, int64_t *total_out);

/* ------------------------------------------------------------------ fused passes of the step */
/* The density pass of coupled.step (compute_density's first round, sph.py:329-387, open box):
 * cell grid at search lengths h_search, gather of every row, and each row's h-rho solve right
 * after it, summing over the gathered d2 (FP64 in the reference's order for precision 0, FP32
 * with compensated sums for precision 1); the lists stay in the context for nb_hydro_tight.
 * The gathered rows sit in 64-row blocks with capacities of their own, sized from row_hint
 * (N x 2 int32, the previous pass's row_counts of the same particles; NULL: a uniform first
 * guess) and gathered again with exact capacities when a row does not fit. row_counts (N x 2
 * int32, may be NULL) receives this pass's (gather row, extras) counts per particle.
 * Outputs per particle; stats4 (host) = [longest gather row, rows gathered unsorted, search-list
 * entries, entries of the tight rows' blocks (the index capacity nb_hydro_tight needs)]. Rows
 * flagged 2 must be redone by the caller (sph.py:382-386). */
int32_t nb_hydro_density(nb_ctx *ctx, void *stream, const double *pos, const double *mass, double m_uniform,
                         const double *h_start, const double *h_search, int64_t n, const int32_t *row_hint,
                         int32_t *row_counts, int32_t kid, double cnorm, double eta, double tol, double h_max,
                         int32_t precision, double *h_out, double *rho, double *omega, int32_t *iters,
                         int32_t *flags, int64_t *stats4);
/* ------------------------------------------------------------------ block time steps
 * Individual power-of-two steps on an int64 timeline (blocks.py, timesteps.py:21-88; levels
 * k <= 40, a step of 2^(40 - k) ticks). Rows are int64 particle indices, levels int64.
 * nb_block_elapsed: s = (n_next - n_beg) tick and L/2 = 0.5 (n_end - n_beg) tick per particle.
 * nb_block_predict: the predictions of a system step at tick n_next (blocks.py:502-530): p_out =
 *   p_flow + (s - L/2) a, u_out = u_half + (s - L/2) du/dt, with beta cooling (beta finite; floor
 *   u_floor, or u_floors per particle) of u_half over s at the mid radius; the active particles
 *   (n_end = n_next) keep their cooled u_half and get e_part = m du (0 elsewhere), luminosity
 *   and heat_cool.
 * nb_block_close: closing kicks of the rows (blocks.py:543-546) and their heat budget.
 * nb_block_levels: new levels of the rows (blocks.py:318-335, 367-372): step criteria (hydro
 *   0: orbit only), level_for_dt, forced offset, allowed_level; nu may be NULL; eta_orb <= 0
 *   without the hole. bad_dev[0] counts criteria not positive, bad_dev[1] levels beyond 40.
 * nb_block_limit: one sweep of the neighbour limiter over the rows (blocks.py:338-353) on CSR
 *   lists; *changed_dev set when a level rose (repeat until it stays 0).
 * nb_block_land: levels raised until steps opened at n_now end by n_stop (blocks.py:383-393);
 *   *bad_dev counts levels beyond 40.
 * nb_block_wake: wake-up (blocks.py:357-365, 396-431): need (N, filled with -1 by the caller)
 *   gets max over the rows' list entries of level - gap; running steps longer than need allows
 *   are cut, their opening kicks corrected; *woken_dev counts the cuts.
 * nb_block_open: opening kicks of the rows at n_now (blocks.py:443-452).
 * nb_block_mark: mark[i] = 1 for the rows and every entry of their lists (blocks.py:202-214). */
int32_t nb_block_elapsed(nb_ctx *ctx, void *stream, const int64_t *n_beg, const int64_t *n_end, int64_t n_next,
                         double tick, int64_t n, double *s_out, double *half_out);
int32_t nb_block_predict(nb_ctx *ctx, void *stream, const int64_t *n_beg, const int64_t *n_end, int64_t n_next,
                         double tick, const double *p_flow, const double *a, double *u_half, const double *dudt,
                         const double *x_beg, const double *x, const double *mass, double u_floor,
                         const double *u_floors, double beta, int64_t n, double *p_out, double *u_out,
                         double *e_part, double *lum, double *heat_cool);
int32_t nb_block_close(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *n_beg,
                       const int64_t *n_end, double tick, const double *p_flow, const double *a, const double *u_half,
                       const double *dudt, const double *visc, const double *cond, const double *disk, double *p,
                       double *u, double *heat_visc, double *heat_cond, double *heat_disk, double *heat_pdv);
int32_t nb_block_levels(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const double *x,
                        const double *h, const double *vsig, const double *a, const double *nu, double courant,
                        double c_force, double eta_orb, int32_t hydro, double dt_max, const int64_t *level_offset,
                        int64_t n_now, int64_t *level, int32_t *bad_dev);
int32_t nb_block_limit(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                       const int32_t *index, int64_t gap, int64_t *level, int32_t *changed_dev);
int32_t nb_block_land(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, int64_t n_now, int64_t n_stop,
                      int64_t *level, int32_t *bad_dev);
int32_t nb_block_wake(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                      const int32_t *index, int64_t gap, int64_t n, int64_t n_now, double tick, int64_t *need,
                      int64_t *level, int64_t *n_end, const double *a, const double *dudt, const double *visc,
                      const double *cond, const double *disk, double *p_half, double *u_half, double *heat_visc,
                      double *heat_cond, double *heat_disk, double *heat_pdv, int32_t *woken_dev);
int32_t nb_block_open(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, int64_t n_now, double tick,
                      const int64_t *level, int64_t *n_beg, int64_t *n_end, const double *x, const double *p,
                      const double *u, const double *a, const double *dudt, const double *visc, const double *cond,
                      const double *disk, double *x_beg, double *p_half, double *u_half, double *heat_visc,
                      double *heat_cond, double *heat_disk, double *heat_pdv);
int32_t nb_block_mark(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                      const int32_t *index, uint8_t *mark);

/* tighten (neighbours.py:306-323) of the last nb_hydro_density lists at the final smoothing
 * lengths, straight into 64-row blocks with capacities of their own: out_off (N / 64 + 1 int64,
 * device) receives the blocks' starts (the nb_lists offset of that layout), out_index holds
 * index_cap >= stats4[3] entries. */
int32_t nb_hydro_tight(nb_ctx *ctx, void *stream, const double *h_final, int64_t *out_off, int32_t *out_len,
                       int32_t *out_index, int64_t index_cap, int64_t *total_out);

/* ------------------------------------------------------------------ mixed-precision step passes */
/* The density pass of coupled.step for the mixed-precision mode (open box): the same search
 * lists as nb_hydro_density (FP64 predicate, exact), gathered by 16-lane groups into rows of their
 * own capacity holding neighbour indices only; the h-rho solve sums in FP32 with compensated
 * accumulation (d2 formed in FP64 and rounded) and takes its Newton steps in FP64. cloud / seg_off / nseg: an ensemble of independent systems as in
 * nb_grid_order_seg (cloud NULL: one system); cell_factor as in nb_grid_order.
 * Row capacities come from row_hint (device, 2 N: gather-row and extras counts of each particle
 * from a previous pass, e.g. row_counts of the last step carried through the particles'
 * permutation) with headroom, or are uniform when row_hint is NULL; a row that outgrows its
 * capacity makes the pass repeat with the exact counts. row_counts (device, 2 N, may be NULL)
 * receives this pass's counts. Writes h_out, iters and flags of every row; rho and omega of the
 * rows that needed more than one Newton step (the others are completed by nb_hydro_tight_mixed,
 * which evaluates them at the final h while it reads the rows). stats4 (host) = [gather-row
 * entries allocated, rows needing more than one step, search-list entries (the index capacity
 * nb_hydro_tight_mixed needs), longest merged row]. */
int32_t nb_hydro_density_mixed(nb_ctx *ctx, void *stream, const double *pos, const double *mass, double m_uniform,
                               const double *h_start, const double *h_search, int64_t n, const int32_t *cloud,
                               const int64_t *seg_off, int32_t nseg, double cell_factor, const int32_t *row_hint,
                               int32_t *row_counts, int32_t kid, double cnorm, double eta, double tol, double h_max,
                               double *h_out, double *rho, double *omega, int32_t *iters, int32_t *flags,
                               int64_t *stats4);
/* ------------------------------------------------------------------ block time steps
 * Individual power-of-two steps on an int64 timeline (blocks.py, timesteps.py:21-88; levels
 * k <= 40, a step of 2^(40 - k) ticks). Rows are int64 particle indices, levels int64.
 * nb_block_elapsed: s = (n_next - n_beg) tick and L/2 = 0.5 (n_end - n_beg) tick per particle.
 * nb_block_predict: the predictions of a system step at tick n_next (blocks.py:502-530): p_out =
 *   p_flow + (s - L/2) a, u_out = u_half + (s - L/2) du/dt, with beta cooling (beta finite; floor
 *   u_floor, or u_floors per particle) of u_half over s at the mid radius; the active particles
 *   (n_end = n_next) keep their cooled u_half and get e_part = m du (0 elsewhere), luminosity
 *   and heat_cool.
 * nb_block_close: closing kicks of the rows (blocks.py:543-546) and their heat budget.
 * nb_block_levels: new levels of the rows (blocks.py:318-335, 367-372): step criteria (hydro
 *   0: orbit only), level_for_dt, forced offset, allowed_level; nu may be NULL; eta_orb <= 0
 *   without the hole. bad_dev[0] counts criteria not positive, bad_dev[1] levels beyond 40.
 * nb_block_limit: one sweep of the neighbour limiter over the rows (blocks.py:338-353) on CSR
 *   lists; *changed_dev set when a level rose (repeat until it stays 0).
 * nb_block_land: levels raised until steps opened at n_now end by n_stop (blocks.py:383-393);
 *   *bad_dev counts levels beyond 40.
 * nb_block_wake: wake-up (blocks.py:357-365, 396-431): need (N, filled with -1 by the caller)
 *   gets max over the rows' list entries of level - gap; running steps longer than need allows
 *   are cut, their opening kicks corrected; *woken_dev counts the cuts.
 * nb_block_open: opening kicks of the rows at n_now (blocks.py:443-452).
 * nb_block_mark: mark[i] = 1 for the rows and every entry of their lists (blocks.py:202-214). */
int32_t nb_block_elapsed(nb_ctx *ctx, void *stream, const int64_t *n_beg, const int64_t *n_end, int64_t n_next,
                         double tick, int64_t n, double *s_out, double *half_out);
int32_t nb_block_predict(nb_ctx *ctx, void *stream, const int64_t *n_beg, const int64_t *n_end, int64_t n_next,
                         double tick, const double *p_flow, const double *a, double *u_half, const double *dudt,
                         const double *x_beg, const double *x, const double *mass, double u_floor,
                         const double *u_floors, double beta, int64_t n, double *p_out, double *u_out,
                         double *e_part, double *lum, double *heat_cool);
int32_t nb_block_close(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *n_beg,
                       const int64_t *n_end, double tick, const double *p_flow, const double *a, const double *u_half,
                       const double *dudt, const double *visc, const double *cond, const double *disk, double *p,
                       double *u, double *heat_visc, double *heat_cond, double *heat_disk, double *heat_pdv);
int32_t nb_block_levels(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const double *x,
                        const double *h, const double *vsig, const double *a, const double *nu, double courant,
                        double c_force, double eta_orb, int32_t hydro, double dt_max, const int64_t *level_offset,
                        int64_t n_now, int64_t *level, int32_t *bad_dev);
int32_t nb_block_limit(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                       const int32_t *index, int64_t gap, int64_t *level, int32_t *changed_dev);
int32_t nb_block_land(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, int64_t n_now, int64_t n_stop,
                      int64_t *level, int32_t *bad_dev);
int32_t nb_block_wake(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                      const int32_t *index, int64_t gap, int64_t n, int64_t n_now, double tick, int64_t *need,
                      int64_t *level, int64_t *n_end, const double *a, const double *dudt, const double *visc,
                      const double *cond, const double *disk, double *p_half, double *u_half, double *heat_visc,
                      double *heat_cond, double *heat_disk, double *heat_pdv, int32_t *woken_dev);
int32_t nb_block_open(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, int64_t n_now, double tick,
                      const int64_t *level, int64_t *n_beg, int64_t *n_end, const double *x, const double *p,
                      const double *u, const double *a, const double *dudt, const double *visc, const double *cond,
                      const double *disk, double *x_beg, double *p_half, double *u_half, double *heat_visc,
                      double *heat_cond, double *heat_disk, double *heat_pdv);
int32_t nb_block_mark(nb_ctx *ctx, void *stream, const int64_t *rows, int64_t nrows, const int64_t *offset,
                      const int32_t *index, uint8_t *mark);

/* tighten (neighbours.py:306-323) of the last nb_hydro_density_mixed lists at h_final (its
 * h_out), into rows of their own capacity: row i at out_off[i] (N + 1 entries written; row i's
 * capacity is its merged search row) with out_len[i] entries; out_index holds index_cap >= the
 * search-list entries (stats4[2]). pos as in the density pass. Rows are exactly tighten() of the
 * search lists (d2 formed in FP64 from pos as the search formed it). Completes rho and omega (the
 * same arrays given to the density pass). */
int32_t nb_hydro_tight_mixed(nb_ctx *ctx, void *stream, const double *pos, const double *h_final, double *rho,
                             double *omega, int64_t *out_off, int32_t *out_len, int32_t *out_index,
                             int64_t index_cap, int64_t *total_out);
/* nb_hydro_density with its outputs, statistics and bits (sph.py:251-297, compute_density at
 * sph.py:329-387 per build), gathered by 16-lane groups over trimmed column runs into rows of
 * their own capacity (the lists of nb_hydro_density_mixed) and solved in FP64 with every sum in
 * the reference's row order. Needs the particles stored in grid order (the coupled step's own
 * permutation): otherwise nothing is computed and NB_NOT_GRID_ORDER (2) is returned, for the
 * caller to take nb_hydro_density. stats4 as nb_hydro_density's (stats4[1] = 0). */
int32_t nb_hydro_density64(nb_ctx *ctx, void *stream, const double *pos, const double *mass, double m_uniform,
                           const double *h_start, const double *h_search, int64_t n, const int32_t *row_hint,
                           int32_t *row_counts, int32_t kid, double cnorm, double eta, double tol, double h_max,
                           double *h_out, double *rho, double *omega, int32_t *iters, int32_t *flags,
                           int64_t *stats4);
/* nb_hydro_tight after nb_hydro_density64 (neighbours.py:284-323 at h_final): the same rows in
 * the same 64-row blocks (out_off from the density pass's block capacities, index_cap at least
 * stats4[3]), each tight row ascending. pos as given to the density pass. */
int32_t nb_hydro_tight64(nb_ctx *ctx, void *stream, const double *pos, const double *h_final, int64_t *out_off,
                         int32_t *out_len, int32_t *out_index, int64_t index_cap, int64_t *total_out);
/* Fixed-capacity passes for steps recorded in a hipGraph (no counterpart in the reference; for
 * callers that replay a captured step). nb_fixed_reserve sizes every buffer of the passes below
 * once (N particles, nseg systems, gather-row, extras and tight-index entries, grid cells;
 * masses != 0: per-particle masses). The passes then make no host synchronisation, allocation or
 * memset; what does not fit, or must be redone eagerly, is flagged in the int32 device status
 * words: [0] gather rows, [1] extras, [2] tight index, [3] grid cells (the grid was coarsened:
 * valid lists, other FP32 sums), [4] rows whose h outgrew the search, [5] particles not stored in
 * cell order. nb_glf_step and nb_eos_ideal given status + 6 / status + 7 add theirs. */
int32_t nb_fixed_reserve(nb_ctx *ctx, int64_t n, int32_t nseg, int64_t gather_cap, int64_t extras_cap,
                         int64_t index_cap, int64_t cells_cap, int32_t masses);
/* out6 = [gather-row entries, extras entries and search-list entries of the last mixed density
 * pass, cells of the last grid, its N, fixed capacities reserved (0/1)]: sizes for nb_fixed_reserve. */
int32_t nb_ctx_sizes(const nb_ctx *ctx, int64_t *out6);
/* nb_grid_order_seg with the grid planned on the device (seg_off_dev: device, nseg + 1). */
int32_t nb_grid_order_seg_fixed(nb_ctx *ctx, void *stream, const double *pos, const double *h, int64_t n,
                                const int32_t *cloud, const int64_t *seg_off_dev, int32_t nseg, double cell_factor,
                                int32_t *status, int32_t *order_out);
/* nb_hydro_density_mixed of a segmented system stored in cell order, with the fixed capacities:
 * row capacities from row_hint (required), one search round (rows that need more are flagged). */
int32_t nb_hydro_density_mixed_fixed(nb_ctx *ctx, void *stream, const double *pos, const double *mass,
                                     double m_uniform, const double *h_start, const double *h_search, int64_t n,
                                     const int32_t *cloud, const int64_t *seg_off_dev, int32_t nseg,
                                     double cell_factor, const int32_t *row_hint, int32_t *row_counts, int32_t kid,
                                     double cnorm, double eta, double tol, double h_max, double *h_out, double *rho,
                                     double *omega, int32_t *iters, int32_t *flags, int32_t *status);
/* nb_hydro_tight_mixed after nb_hydro_density_mixed_fixed; out_index holds the reserved index
 * capacity. */
int32_t nb_hydro_tight_mixed_fixed(nb_ctx *ctx, void *stream, const double *pos, const double *h_final, double *rho,
                                   double *omega, int64_t *out_off, int32_t *out_len, int32_t *out_index,
                                   int32_t *status);
/* nb_dt_minima_seg into device memory (out_dev: nseg x 4). */
int32_t nb_dt_minima_seg_dev(nb_ctx *ctx, void *stream, const double *h, const double *vsig, const double *acc,
                             const double *nu, double c_visc, const double *x, const int64_t *seg_off_dev,
                             int32_t nseg, double *out_dev);
/* Node counts by hipGraphNodeType (16 entries) of a recorded graph (hipGraph_t). */
int32_t nb_graph_node_counts(void *graph, int32_t *counts16);

/* gradients_numba (sph.py:837-884) in mixed precision: FP32 pair terms over 16-lane row
 * groups, compensated sums, partials combined and the 3x3 inverse taken in FP64. Arguments as
 * nb_gradients. (The force pass takes its precision from nb_force_params.precision.) */
int32_t nb_gradients_mixed(nb_ctx *ctx, void *stream, const double *pos, const double *vel, const double *acc,
                           const double *mass, const double *h, const nb_lists *lists, const double *box3,
                           const int64_t *rows, int64_t nrows, int32_t kid, double cnorm, double *div, double *curl,
                           double *ddiv);

/* ------------------------------------------------------------------ pair passes */
/* density_numba (sph.py:300-326 -> _density_kernel sph.py:251-297): Newton-bisection h-rho
 * solve per row. m_uniform > 0 declares that every mass equals it (masses then not cached).
 * FP64, bit-identical to the reference's arithmetic. flags: 0 ok, 1 at h_max, 2 search too small, 3 no convergence. */
int32_t nb_density(nb_ctx *ctx, void *stream, const double *pos, const double *mass, double m_uniform,
                   const double *h_start, const double *h_search, const nb_lists *lists,
                   const double *box3, const int64_t *rows,
                   int64_t nrows, int32_t kid, double cnorm, double eta, double tol, double h_max,
                   double *h_out, double *rho, double *omega, int32_t *iters, int32_t *flags);

/* forces_numba (sph.py:655-690 -> _forces_kernel sph.py:525-652). u, c_s, alpha, grad_v may be
 * NULL when the corresponding term is off. vheat / cheat may be NULL. */
int32_t nb_forces(nb_ctx *ctx, void *stream, const double *pos, const double *vel, const double *mass,
                  const double *h, const double *rho, const double *omega, const double *pressure,
                  const double *u, const double *c_s, const double *alpha, const double *grad_v,
                  const nb_lists *lists, const double *box3, const int64_t *rows,
                  int64_t nrows, const nb_force_params *params, double *acc, double *dudt,
                  double *vsig_max, double *vheat, double *cheat);

/* gradients_numba (sph.py:887-902 -> _gradients_kernel sph.py:837-884); acc NULL = zeros. */
int32_t nb_gradients(nb_ctx *ctx, void *stream, const double *pos, const double *vel, const double *acc,
                     const double *mass, const double *h, const nb_lists *lists,
                     const double *box3, const int64_t *rows, int64_t nrows, int32_t kid, double cnorm,
                     double *div, double *curl, double *ddiv);

/* velocity_gradients_numba (sph.py:818-834 -> sph.py:785-815): out (N,3,3), rows written. */
int32_t nb_velocity_gradients(nb_ctx *ctx, void *stream, const double *pos, const double *vel,
                              const double *mass, const double *h, const nb_lists *lists,
                              const double *box3, const int64_t *rows, int64_t nrows,
                              int32_t kid, double cnorm, double *out);

/* ns_gradients_numba (sph.py:984-999 -> sph.py:957-981): out (N,3,3), rows written. */
int32_t nb_ns_gradients(nb_ctx *ctx, void *stream, const double *pos, const double *vel, const double *mass,
                        const double *h, const double *rho, const double *omega, const nb_lists *lists,
                        const double *box3, const int64_t *rows, int64_t nrows,
                        int32_t kid, double cnorm, double *out);

/* ns_forces_numba (sph.py:1030-1045 -> sph.py:1002-1027): acc (R,3). */
int32_t nb_ns_forces(nb_ctx *ctx, void *stream, const double *pos, const double *mass, const double *h,
                     const double *rho, const double *omega, const double *stress, const nb_lists *lists,
                     const double *box3, const int64_t *rows, int64_t nrows,
                     int32_t kid, double cnorm, double *acc);

/* disc_forces_numba (sph.py:1120-1137 -> _disc_kernel sph.py:1084-1117). */
int32_t nb_disc_forces(nb_ctx *ctx, void *stream, const double *pos, const double *vel, const double *mass,
                       const double *h, const double *rho, const double *omega, const double *c_s,
                       double alpha_av, const nb_lists *lists, const double *box3,
                       const int64_t *rows, int64_t nrows, int32_t kid, double cnorm, double *acc,
                       double *dudt);

/* ------------------------------------------------------------------ GN gravity and the step */
/* glf_step_many (integrator.py:188-198 -> _glf_one integrator.py:139-176): dt_arr (N) or NULL
 * with dt_scalar. iters: count or -1 / -2 / -3 (integrator.py:38-41); nfail_dev (device int32,
 * optional) is incremented per failed particle. */
int32_t nb_glf_step(nb_ctx *ctx, void *stream, const double *x, const double *p, const double *dt_arr,
                    double dt_scalar, int64_t n, double *x_out, double *p_out, int32_t *iters,
                    int32_t *nfail_dev);

/* gn_v_from_p_nb (fastphys.py:35-42). */
int32_t nb_gn_v_from_p(nb_ctx *ctx, void *stream, const double *x, const double *p, int64_t n, double *out);

/* out = y + s a over m doubles (the half kicks of coupled.py:229-258). */
int32_t nb_axpy(nb_ctx *ctx, void *stream, const double *y, const double *a, double s, int64_t m,
                double *out);
/* out = y + (scale s_dev[0]) a, the scalar read from device memory (no counterpart in the
 * reference: the kicks of a step recorded for replay). */
int32_t nb_axpy_dev(nb_ctx *ctx, void *stream, const double *y, const double *a, const double *s_dev, double scale,
                    int64_t m, double *out);
/* nb_axpy_dev on a momentum (n x 3) and an energy (n) pair with the same dt, in one launch
 * (coupled.py:231-233, 262-263 and the predictor kick; no separate counterpart in the reference):
 * out3 = y3 + scale s_dev[0] a3, out1 = y1 + scale s_dev[0] a1, the same values as two nb_axpy_dev
 * calls. */
int32_t nb_kick2_dev(nb_ctx *ctx, void *stream, const double *y3, const double *a3, const double *y1, const double *a1,
                     const double *s_dev, double scale, int64_t n, double *out3, double *out1);

/* eos_state, ideal branch (hydro.py:121-139; eos.py:21-28); nbad_dev counts u <= 0. */
int32_t nb_eos_ideal(nb_ctx *ctx, void *stream, const double *rho, const double *u, double gamma, int64_t n,
                     double *pressure, double *cs, int32_t *nbad_dev);

/* beta cooling inside the step (coupled.py:244-251; cooling.py:28-40); e_part = m (u - u_new). */
int32_t nb_cool_beta(nb_ctx *ctx, void *stream, const double *u, const double *x_old, const double *x_new,
                     const double *mass, double u_floor, double dt, double beta, int64_t n, double *u_out,
                     double *e_part);
/* nb_cool_beta with a floor per particle (the gas + radiation floor depends on the density,
 * hydro.py:155-163). */
int32_t nb_cool_beta_floors(nb_ctx *ctx, void *stream, const double *u, const double *x_old, const double *x_new,
                            const double *mass, const double *u_floor, double dt, double beta, int64_t n,
                            double *u_out, double *e_part);
/* Gas + radiation equation of state (fastphys.py:45-111, eos.py:76-148; hydro.py:133-135): P and
 * c_s (code units) and T [K] (t_out may be NULL) from code rho and u, dens_kg_m3 being one code
 * density unit (GasUnits.density_kg_m3). T by Newton on log T in the bracket [1, 1e13] K to
 * 1e-12. bad_dev (two int32, zeroed by the caller) counts u outside the bracket (the reference's
 * ValueError) and solves that did not converge (its RuntimeError). */
int32_t nb_eos_gasrad(nb_ctx *ctx, void *stream, const double *rho, const double *u, double dens_kg_m3, int64_t n,
                      double *pressure, double *cs, double *t_out, int32_t *bad_dev);
/* The code u of gas + radiation at temperature t_k for each code density (hydro.py:155-163,
 * floor_energy). */
int32_t nb_gasrad_u(nb_ctx *ctx, void *stream, const double *rho, double dens_kg_m3, double t_k, int64_t n,
                    double *u_out);

/* optically thin power-law cooling du/dt = -k rho sqrt(u) above u_floor, exact at fixed rho
 * (the reference has no such law; same place in the split as beta cooling). */
int32_t nb_cool_powerlaw(nb_ctx *ctx, void *stream, const double *u, const double *rho, const double *mass,
                         double u_floor, double k, double dt, int64_t n, double *u_out, double *e_part);
/* nb_cool_powerlaw with dt read from device memory (dt_dev[0]; no counterpart in the reference:
 * a step recorded for replay, whose dt the host writes between replays). */
int32_t nb_cool_powerlaw_dev(nb_ctx *ctx, void *stream, const double *u, const double *rho, const double *mass,
                             double u_floor, double k, const double *dt_dev, int64_t n, double *u_out,
                             double *e_part);

/* stable_dt minima (coupled.py:213-226): out4 (host) = [min h/v_sig, min sqrt(h/|a|) over |a|>0,
 * min c_visc h^2/nu over nu>0 (nu may be NULL), min |x|]. */
int32_t nb_dt_minima(nb_ctx *ctx, void *stream, const double *h, const double *vsig, const double *acc,
                     const double *nu, double c_visc, const double *x, int64_t n, double *out4);
/* nb_dt_minima into device memory (out4_dev), no host synchronisation: a step driver reads it
 * with its other per-step words in one copy. */
int32_t nb_dt_minima_dev(nb_ctx *ctx, void *stream, const double *h, const double *vsig, const double *acc,
                         const double *nu, double c_visc, const double *x, int64_t n, double *out4_dev);
/* nb_dt_minima_dev that allocates nothing. No counterpart in the reference: it is for callers
 * that cannot allocate (code under hipGraph capture, serving loops that forbid hipMalloc). Its
 * partial minima live in a fixed device array of the library (1025 x 4 doubles = 32.8 KB per
 * GPU, held until the process ends) instead of the context's scratch arena; a launch writes at
 * most 1024 partial rows whatever n, so every n >= 1 fits (n above INT64_MAX / 3 is refused).
 * Same results as nb_dt_minima_dev, bit for bit. Every call on one GPU shares its array: calls
 * that overlap on one GPU (concurrent streams) race; calls in stream order do not. */
int32_t nb_dt_minima_dev_static(nb_ctx *ctx, void *stream, const double *h, const double *vsig,
                                const double *acc, const double *nu, double c_visc, const double *x, int64_t n,
                                double *out4_dev);
/* The next step's dt from the minima of nb_dt_minima_dev, on the device (replaces the host half
 * of coupled.py:213-226 stable_dt, as coupled.dt_from_minima computes it: courant min_0, then
 * c_force min_1 when finite, min_2 when viscous, eta_orb min_3^1.5 with the hole, each kept only
 * where strictly smaller): dts2_dev = (dt, dt_prev) becomes (new dt, old dt). For a step driver
 * that launches the next step before it reads this one back. */
int32_t nb_dt_advance(nb_ctx *ctx, void *stream, const double *minima4_dev, double courant, double c_force,
                      int32_t viscous, int32_t hole, double eta_orb, double *dts2_dev);

/* What a step driver that launches the next step before reading this one back reads back, in
 * one device buffer (no counterpart in the reference: the sums coupled.step's caller forms on
 * the host, GasState.n_search_entries and n_tight_entries, and the gate's count of rows that
 * took more than one Newton step): out (device, nstatus + 9 doubles) = [status words, minima4,
 * sum of row_counts (2 int32 per particle), sum of nneigh, dts2, count of iters > 1]. The status
 * words are cleared for the next step; acc3 (device, 3 uint64, zero on entry) is left zero. */
int32_t nb_step_readback(nb_ctx *ctx, void *stream, int32_t *status, int32_t nstatus, const double *minima4,
                         const int32_t *row_counts, const double *nneigh, const int32_t *iters, int64_t n,
                         const double *dts2, unsigned long long *acc3, double *out);

/* Ensemble helpers (no counterpart in the reference; the systems of nb_grid_order_seg):
 * nb_dt_minima_seg: nb_dt_minima per system, out (host) nseg x 4; nb_axpy_rows: out row i = y row i
 * + s_rows[i] a row i (rows of width values, a kick with a step per particle); nb_expand_seg:
 * out[i] = scale per_seg[cloud[i]] (per_seg on the device). */
int32_t nb_dt_minima_seg(nb_ctx *ctx, void *stream, const double *h, const double *vsig, const double *acc,
                         const double *nu, double c_visc, const double *x, const int64_t *seg_off, int32_t nseg,
                         double *out);
int32_t nb_axpy_rows(nb_ctx *ctx, void *stream, const double *y, const double *a, const double *s_rows, int64_t n,
                     int32_t width, double *out);
int32_t nb_expand_seg(nb_ctx *ctx, void *stream, const double *per_seg, const int32_t *cloud, int64_t n, double scale,
                      double *out);

/* ledger sums (coupled.py:261-270): out6 (host) = [sum m (H + u), sum m |H|, sum m (x cross p)
 * (3), sum e_part]. gn = 1: GN Hamiltonian (gravity.py:92-102); 0: |p|^2/2. */
int32_t nb_ledger(nb_ctx *ctx, void *stream, const double *x, const double *p, const double *u,
                  const double *mass, const double *e_part, int32_t gn, int64_t n, double *out6);

/* ------------------------------------------------------------------ element-wise step glue */
/* search lengths min(min(h, h_max) * skin, box_cap) (sph.py:371). */
int32_t nb_search_lengths(nb_ctx *ctx, void *stream, const double *h, double h_max, double skin, double box_cap,
                          int64_t n, double *out);
/* the rebuild round of compute_density (sph.py:382-386): flagged rows (flag 2) double their search
 * (capped), every row's next starting h is its result; *grown_out = number of flagged rows. */
int32_t nb_grow_search(nb_ctx *ctx, void *stream, const int32_t *flags, const double *h_res, const int64_t *rows,
                       int64_t nrows, double box_cap, double *h_search, double *h_guess, int64_t *grown_out);
int32_t nb_fill(nb_ctx *ctx, void *stream, double *p, double v, int64_t n);
/* Balsara limiter on the viscosity's sound-speed part (no counterpart in the reference, which
 * has only the Cullen-Dehnen switch): alpha = alpha_const |div v| / (|div v| + |curl v| + eps c_s/h). */
int32_t nb_alpha_balsara(nb_ctx *ctx, void *stream, const double *div, const double *curl, const double *cs,
                         const double *h, double alpha_const, double eps, int64_t n, double *alpha);
/* Cullen-Dehnen switch update (viscosity.py:82-106 as used at coupled.py:159-162). */
int32_t nb_alpha_switch(nb_ctx *ctx, void *stream, const double *div, const double *curl, const double *ddiv,
                        const double *cs, const double *h, const double *alpha_in, double dt, double alpha_max,
                        double sigma, int64_t n, double *alpha_out);
/* nb_alpha_switch with a step per element (blocks.py:266-270: each active row its own step). */
int32_t nb_alpha_switch_dt(nb_ctx *ctx, void *stream, const double *div, const double *curl, const double *ddiv,
                           const double *cs, const double *h, const double *alpha_in, const double *dt,
                           double alpha_max, double sigma, int64_t n, double *alpha_out);
/* Navier-Stokes glue (navier_stokes.py:29-50 and hydro.py:221-230): nu, viscous stress (N,3,3),
 * viscous heating (N). alpha_ss > 0: nu = alpha_ss c_s^2 |x|^1.5, else nu = nu_const. */
int32_t nb_ns_stress(nb_ctx *ctx, void *stream, const double *g, const double *rho, const double *cs,
                     const double *x, double alpha_ss, double nu_const, int64_t n, double *nu, double *stress,
                     double *heat);
/* nu_AV = alpha_AV c_s h / 10 (disc_viscosity.py:46-48). */
int32_t nb_nu_av(nb_ctx *ctx, void *stream, const double *cs, const double *h, double alpha_av, int64_t n,
                 double *nu);
int32_t nb_add(nb_ctx *ctx, void *stream, const double *a, const double *b, int64_t m, double *out);
/* acc_dev[0] += sum(x) on the device (no host synchronization). */
int32_t nb_sum_into(nb_ctx *ctx, void *stream, const double *x, int64_t n, double *acc_dev);

/* ------------------------------------------------------------------ learned starting h
 * No counterpart in the reference: the learned-h mode of this implementation. A small MLP
 * (8 -> 64 -> 64 -> 1, tanh) predicts delta = log h_final - log h_prev from the previous step's
 * state; the Newton iteration starts from h_prev exp(delta) and still converges to the same
 * tolerance (it stays the arbiter). params: nb_mlp_param_count() floats, W1[8][64], b1[64],
 * W2[64][64], b2[64], w3[64], b3. norm: 18 floats, feature means[8], feature 1/std[8], target
 * mean, target scale. Features (raw, float32, 8 per row): log h, log rho, div v dt, |curl v| dt,
 * log(neighbour count), log r, log u, dt / dt_prev. */
int32_t nb_mlp_param_count(void);
/* raw features of particles 0, stride, 2 stride, ... (nneigh: neighbour counts as doubles; div_v,
 * curl_v, nneigh may be NULL: 0, 0, 1) */
int32_t nb_mlp_features(nb_ctx *ctx, void *stream, const double *h_prev, const double *rho_prev, const double *div_v,
                        const double *curl_v, const double *nneigh, const double *x, const double *u, int64_t n,
                        int64_t stride, double dt, double dt_prev, float *out);
/* nb_mlp_features with (dt, dt_prev) read from device memory (dts_dev[0], dts_dev[1]). */
int32_t nb_mlp_features_dev(nb_ctx *ctx, void *stream, const double *h_prev, const double *rho_prev,
                            const double *div_v, const double *curl_v, const double *nneigh, const double *x,
                            const double *u, int64_t n, int64_t stride, const double *dts_dev, float *out);
/* training target log h_final - log h_prev of the same rows */
int32_t nb_mlp_target(nb_ctx *ctx, void *stream, const double *h_prev, const double *h_final, int64_t n,
                      int64_t stride, float *out);
/* delta for n feature rows: BF16 MFMA (FP32 accumulation) */
int32_t nb_mlp_infer(nb_ctx *ctx, void *stream, const float *params, const float *norm, const float *feat, int64_t n,
                     float *out);
/* Newton starting h = h_prev exp(delta). The previous h where |delta| < gate (a change below the
 * Newton tolerance converges in one step from it anyway), and where delta is not finite or
 * |delta| >= ln 2 (fallbacks: their number, written to the device int64 *fallbacks_dev; no host
 * synchronisation). */
int32_t nb_mlp_h_start(nb_ctx *ctx, void *stream, const double *h_prev, const float *delta, int64_t n, float gate,
                       double *h_start, long long *fallbacks_dev);
/* FP32 gradient of the mean squared error (normalised target) over a batch, and the loss */
int32_t nb_mlp_grad(nb_ctx *ctx, void *stream, const float *params, const float *norm, const float *feat,
                    const float *target, int64_t batch, float *grad_out, float *loss_out);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8; step counts from 1) on a batch; loss before the step */
int32_t nb_mlp_train_step(nb_ctx *ctx, void *stream, float *params, float *adam_m, float *adam_v, int64_t step,
                          const float *norm, const float *feat, const float *target, int64_t batch, float lr,
                          float *loss_out);

#ifdef __cplusplus
}
#endif

#endif /* NEBULA_HIP_H */
