// Library context: the workspace arena behind the C ABI (include/nebula_hip.h).
//
// Device memory for inputs and outputs belongs to the caller; the context owns only
// scratch space (grid, padded gather rows, scan temporaries, reduction partials), which
// grows on demand and is reused between calls. One context per stream of work: calls that
// share a context must be issued in order (they share the scratch).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "nb_common.h"

// one timed launch: HIP events recorded around it on its stream
struct nb_timed {
    const char* name;
    hipEvent_t start, stop;
};

struct nb_kernel_time {
    std::string name;
    long long launches = 0;
    double ms = 0.0;
};

struct nb_ctx {
    int device = 0;
    std::string err;
    // bumped whenever ctx_reserve moves a buffer it had handed out: a recorded graph holds raw
    // pointers into the arena, so a caller replays one only while the generation it was
    // recorded at is still current (nb_ctx_generation)
    uint64_t generation = 0;
    // kernel timing (nb_timing)
    int timing = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<nb_timed> ev_open;
    std::vector<nb_kernel_time> ktimes;
    // neighbour grid
    uint32_t* cell_id = nullptr; size_t cell_id_bytes = 0;
    uint32_t* keys_out = nullptr; size_t keys_out_bytes = 0;
    uint32_t* iota = nullptr; size_t iota_bytes = 0;
    uint32_t* order = nullptr; size_t order_bytes = 0;
    uint32_t* cell_count = nullptr; size_t cell_count_bytes = 0;
    uint32_t* cell_start = nullptr; size_t cell_start_bytes = 0;
    // gather rows and the transposed extras
    uint32_t* gbuf = nullptr; size_t gbuf_bytes = 0;
    int32_t* rowcnt = nullptr; size_t rowcnt_bytes = 0;
    uint8_t* unsorted = nullptr; size_t unsorted_bytes = 0;
    uint8_t* pending = nullptr; size_t pending_bytes = 0;
    double* edbuf = nullptr; size_t edbuf_bytes = 0;
    unsigned long long* ecount = nullptr; size_t ecount_bytes = 0;
    unsigned long long* eoff = nullptr; size_t eoff_bytes = 0;
    unsigned int* efill = nullptr; size_t efill_bytes = 0;
    uint32_t* ebuf = nullptr; size_t ebuf_bytes = 0;
    unsigned long long* tot = nullptr; size_t tot_bytes = 0;
    unsigned long long* offs = nullptr; size_t offs_bytes = 0;
    unsigned long long* tcount = nullptr; size_t tcount_bytes = 0;
    // density distance cache (one double per search-list entry)
    double* dscratch = nullptr; size_t dscratch_bytes = 0;
    float* fscratch = nullptr; size_t fscratch_bytes = 0;
    void* pack = nullptr; size_t pack_bytes = 0;  // per-particle FP32 records of the mixed passes
    long long* part = nullptr; size_t part_bytes = 0;  // per-workgroup partials of reductions
    // mixed lists: per-row capacities and offsets (gather rows, extras segments, tight rows)
    unsigned long long* rm_gc = nullptr; size_t rm_gc_bytes = 0;
    unsigned long long* rm_go = nullptr; size_t rm_go_bytes = 0;
    unsigned long long* rm_ec = nullptr; size_t rm_ec_bytes = 0;
    unsigned long long* rm_eo = nullptr; size_t rm_eo_bytes = 0;
    unsigned long long* rm_to = nullptr; size_t rm_to_bytes = 0;
    // FP64 step lists: 64-row blocks with capacities of their own (sizes, gather and tight offsets)
    unsigned long long* fb_sz = nullptr; size_t fb_sz_bytes = 0;
    unsigned long long* fb_go = nullptr; size_t fb_go_bytes = 0;
    unsigned long long* fb_to = nullptr; size_t fb_to_bytes = 0;
    int64_t fb_gtotal = 0, fb_ttotal = 0;
    uint32_t* ebuf_s = nullptr; size_t ebuf_s_bytes = 0;  // extras in index order
    uint32_t* xlist = nullptr; size_t xlist_bytes = 0;    // work lists of the longer extras segments
    int64_t rm_gtotal = 0;  // gather-row entries of the last mixed pass
    int64_t rm_etotal = 0;  // extras entries of the last mixed pass
    // fixed capacities of the passes that a captured step records (nb_fixed_reserve): no host
    // synchronisation or allocation inside them; what does not fit is reported in a status word
    int fx_ready = 0;
    int64_t fx_n = 0, fx_nseg = 0, fx_gather = 0, fx_extras = 0, fx_index = 0, fx_cells = 0;
    int fx_bits = 32;  // radix-sort key bits of fx_cells
    int fx_grid_from_order = 0;  // the last fixed order pass's grid awaits its density pass
    void* segs = nullptr; size_t segs_bytes = 0;       // grid table of an ensemble's segments
    void* segpart = nullptr; size_t segpart_bytes = 0; // per-workgroup partials of the segment boxes
    std::vector<char> host_segs;                       // its host copy
    void* posh = nullptr; size_t posh_bytes = 0;  // (x, y, z, h_search) records of the mixed gather
    void* mlp_scratch = nullptr; size_t mlp_scratch_bytes = 0;  // training activations and partials
    void* mlp_pack = nullptr; size_t mlp_pack_bytes = 0;        // BF16 MFMA operands of the network
    long long* mlp_part = nullptr; size_t mlp_part_bytes = 0;   // fallback counts per workgroup
    // reductions
    double* red = nullptr; size_t red_bytes = 0;
    int64_t* scal = nullptr; size_t scal_bytes = 0;
    void* cub_tmp = nullptr; size_t cub_tmp_bytes = 0;
    void* fx_tmp = nullptr; size_t fx_tmp_bytes = 0;  // scan partials and sort histograms of the fixed passes
    int32_t row_cap = 0;
    int32_t row_cap_rm = 0;
    int64_t ecap_rm = 0;  // extras segment capacity of the mixed lists
    int64_t lists_n = 0, lists_total = 0, lists_maxrow = 0, last_unsorted = 0;
    int lists_ready = 0;
    int hydro_ready = 0;
    int hydro_mixed_ready = 0;
    int hydro64_ready = 0;  // nb_hydro_density64's lists await nb_hydro_tight64
    int order_ident = 0;  // last grid: particles already stored in cell order (no order array)
    int tile = 256;       // workgroup size of the mixed-precision lane-group kernels (nb_set_tile)
    // settings of the last mixed density pass, for its emission pass
    int32_t mix_kid = 0;
    double mix_cnorm = 0.0, mix_m_uniform = 0.0;
    int64_t last_grid_cells = 0;
    double last_cell = 0.0;
};

// Look up, once per device, the address of the fixed device array behind the no-allocation
// entry point (step.hip); called by nb_ctx_create with that device current. Nonzero on failure.
int static_scratch_init(int device, std::string* err);

int32_t ctx_fail(nb_ctx* ctx, const char* msg);
int32_t ctx_fail_code(nb_ctx* ctx, int32_t code, const char* msg);
// grow *p to at least `bytes` (contents are not kept); nonzero on failure
int ctx_reserve(nb_ctx* ctx, void** p, size_t* have, size_t bytes);

// Kernel timing: with ctx->timing set, events around a launch (same stream); their durations are
// added to ctx->ktimes when the events are read (ktime_flush, or when many are open).
int ktime_begin(nb_ctx* ctx, hipStream_t s, const char* name);
void ktime_end(nb_ctx* ctx, hipStream_t s, int slot);
void ktime_flush(nb_ctx* ctx);

#define NB_TIMED(CTX, STREAM, NAME, ...)                      \
    do {                                                      \
        const int ts_ = ktime_begin((CTX), (STREAM), (NAME)); \
        hipLaunchKernelGGL(__VA_ARGS__);                      \
        ktime_end((CTX), (STREAM), ts_);                      \
    } while (0)
