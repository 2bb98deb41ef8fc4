// xts_internal.hpp -- shared pieces of the HIP implementation: context,
// error reporting, device arena, separation-grid constants and the FP64 WCS
// math used both by kernels and by host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/xtsrcmaps_amd.h"
#include "xts_wcs.hpp"

namespace xts {

// ---------------------------------------------------------------- constants
// Separation grid (psf/psf.hxx:24-31, misc/misc.hxx:22-39): node 0 at 0 and
// node i >= 1 at 1e-4 deg * exp((i - 1) ln(7e5) / 399).
constexpr int kND = 401;
constexpr double kSepStep = 0.033731417579011382769913057686378448039;
constexpr double kRecipStep = 29.645952402019046699874076387601854602589;
constexpr double kRedm1 = 29.148763300179999028597222671574207003288;
constexpr double kLn1e4 = 9.2103403719761827360719658187374568304044;

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define XTS_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return ::xts::fail(e_ == hipErrorOutOfMemory ? XTS_ERR_ALLOC     \
                                                         : XTS_ERR_DEVICE,   \
                               "%s: %s (%s:%d)", #call,                      \
                               hipGetErrorString(e_), __FILE__, __LINE__);   \
    } while (0)

#define XTS_CHECK(cond, ...)                                                 \
    do {                                                                     \
        if (!(cond)) return ::xts::fail(XTS_ERR_INVALID, __VA_ARGS__);       \
    } while (0)

#define XTS_TRY(expr)                                                        \
    do {                                                                     \
        int rc_ = (expr);                                                    \
        if (rc_ != XTS_OK) return rc_;                                       \
    } while (0)

// ---------------------------------------------------------------- arena
// A bump allocator over one device buffer per slot.  Slots keep their
// buffer between calls and grow (free + malloc) only when a call needs more,
// so steady-state calls allocate nothing.
struct Slot {
    void* ptr = nullptr;
    size_t bytes = 0;
};

enum SlotId {
    kSlotIn0 = 0, kSlotIn1, kSlotIn2, kSlotIn3, kSlotIn4, kSlotIn5,
    kSlotIn6, kSlotIn7, kSlotOut0, kSlotOut1, kSlotOut2, kSlotOut3,
    kSlotGeom, kSlotLut, kSlotAcc, kSlotWork0, kSlotWork1, kSlotWork2,
    kSlotRows, kSlotConst, kSlotRule, kSlotSrcGeo, kSlotNearPool, kSlotBilerp,
    kSlotMlpState, kSlotMlpPacked, kSlotSurrogate, kSlotSurList, kSlotSumFac,
    kSlotSurrogate8, kSlotSurrogateH, kSlotDiffSep, kSlotDiffSpec, kSlotDiffKspec, kSlotDiffConv, kSlotCount
};

}  // namespace xts

struct xts_timing_rec {
    const char* name;
    hipEvent_t start, stop;
};

// Device copies of constant tables, uploaded once per context.
struct xts_const_tables {
    double* sep = nullptr;       // [401] separations (deg)
    double* inv64 = nullptr;     // [401] 2 R2D / sep (0 at node 0)
    float* inv32 = nullptr;      // [401] the same in FP32
    double* theta = nullptr;     // [2 * 401] theta (rad), then sin theta
};

// The last cubature rule uploaded, with the quantities derived from it.
struct xts_rule_cache {
    std::vector<double> host;    // [N, 3] as given
    void* dev = nullptr;         // per-point records on the device, then FP32 arrays
    size_t lean_off = 0;         // byte offset of the FP32 arrays in dev
    double rho_scale = 0.0, wsum = 0.0;
    int nu = 0, nv = 0;
};

struct xts_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    xts::Slot slots[xts::kSlotCount];
    int cu_count = 0;
    int timing = 0;                        // record events around kernels
    int tune_tiles_per_block = 0;          // 0: chosen per call (row-weight pass)
    int64_t tune_rw_blocks = 0;            // >0: the row-weight pass on this many resident blocks
    std::vector<xts_timing_rec> timing_recs;
    std::vector<hipEvent_t> event_pool;
    bool consts_ready = false;
    xts_const_tables consts;
    xts_rule_cache rule;
    bool geom_cache = true;                // keep the last image's geometry (tuning key)
    bool geom_valid = false;               // pixel geometry in kSlotGeom is for geom_wcs
    xts_wcs geom_wcs{};
    // A caller-provided device region that slot_reserve bump-allocates from
    // instead of the slots while base is set (the no-allocation entry
    // point); nothing taken from it is cached past the call.
    struct {
        char* base = nullptr;
        size_t cap = 0, used = 0;
    } arena;
    xts_const_tables arena_consts;         // the constant tables inside the arena
    // the surrogate network in its matrix-core record (xts_ctx_set_surrogate)
    void* surrogate = nullptr;
    void* surrogate8 = nullptr;   // the same network's FP8 record
    void* surrogate_h = nullptr;  // its BF16 record without the weights' low parts
    bool surrogate_ready = false;
    // batched engine: the second context its odd batches run on (own slots,
    // own stream), fork/join events, the tuned launch shapes per workload
    // and the captured graph of the last batched call
    xts_ctx* aux = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // the batched engine's pipeline: this context's last row-weight half and
    // last second half done
    hipEvent_t rows_ev = nullptr, fin_ev = nullptr;
    struct Tuned {
        std::string key;
        int64_t batch_sources = 0, rw_blocks = 0;
        double ms = 0.0;
    };
    std::vector<Tuned> tuned;
    int autotune = 0, use_graph = 0;
    int64_t tune_batch_sources = 0;        // >0: sources per batch of the batched engine
    int64_t tune_sum_batch = 0;            // >0: sources per batch of a summed model
    int64_t tune_sum_group = 4;            // sources per staging round of the summed model (4, 8)
    int64_t tune_pipe_batches = 0;         // source batches of the pipelined map (0 default, 1 off)
    int64_t tune_rw_frames = 0;            // >0: row weights by tile frames, this many sources a block
    int64_t tune_rw_frames_cu = 0;         // >0: resident tile-frame blocks per CU
    int64_t tune_emit_prio = 0;            // wave priority of the pipelined map's emit (0-3)
    int64_t tune_emit_ec = 0;              // energies per block of the pipelined map's emit (20, 40)
    int64_t tune_rw_legacy = 0;            // 1: FP32 row weights by per-source tile blocks
    int64_t tune_emit_wide_apart = 0;      // 1: the pipeline's wide pixels in a launch of their own
    int64_t tune_near_emit_gx = 0;         // >0: the pipeline's near-pixel blocks per source
    int64_t tune_sur_lo = 0;               // 0: BF16 network with plain BF16 hidden weights,
                                           // 1: with their high and low parts
    int64_t tune_sur_beside = 4;           // the pipeline's network pass: 0 on the context stream,
                                           // n on the second stream, n resident blocks per CU
    // the pipelined map's second stream (row weights of the next batch) and
    // its events: [0, B) row weights done, [B, 2B) batch emitted, 2B fork
    hipStream_t pipe_stream = nullptr;
    std::vector<hipEvent_t> pipe_ev;
    int64_t last_batch_sources = 0, last_rw_blocks = 0, last_batches = 0;  // last batched call
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;
    // advanced by every buffer (re)allocation and every upload of a rule or
    // network: a captured graph is replayed only while it has not changed
    uint64_t generation = 0;
    int64_t graph_replays = 0, graph_captures = 0;
    std::vector<unsigned char> arena_rule_blob;  // host copy of the rule uploaded into it
    // arguments of a model map whose row-weight half has been enqueued and
    // whose second half has not
    struct {
        bool valid = false;
        xts_wcs wcs{};
        const double* src = nullptr;
        const double* rule = nullptr;
        int64_t S = 0, E = 0, N = 0;
        int32_t flags = 0;
    } pending_rows;
};

namespace xts {

int slot_reserve(xts_ctx* ctx, int slot, size_t bytes, void** out);
// Bytes slot_reserve takes from an arena for a request of `bytes`, and the
// arena alignment: every request starts on a 256-byte boundary.
constexpr size_t kArenaAlign = 256;
inline size_t arena_bytes(size_t bytes) {
    if (bytes == 0) bytes = 256;
    return (bytes + kArenaAlign - 1) & ~(kArenaAlign - 1);
}
// Device address of the no-allocation scratch on the current device, taken
// once per device at context creation (xts_modelmap.hip).
int noalloc_scratch_init(int device);
// The device model map of xts_point_src_model_map_wcs_dev, for the batched
// engine's batches (xts_batch.hip).
// phase: bits 1 the row-weight half (geometry, radii, row weights, near
// rows), 2 the PSF table and the factors, 4 the emit (after calls with bits 1
// and 2 and the same arguments on the same context); 7 the whole map.
int model_map_dev_call(xts_ctx* ctx, const xts_wcs* wcs, const double* src_sph, int64_t S,
                       const double* expo, const double* upsf, const double* parint, int64_t E,
                       const double* rule, int64_t N, int32_t flags, void* out, int phase);
int noalloc_scratch(int device, void** ptr, size_t* bytes);
// Per-pixel solid angle [naxis1 * naxis2] of an image on the device, in the
// geometry slot (xts_modelmap.hip).
int image_solid_angle(xts_ctx* ctx, const xts_wcs* wcs, const double** omega);
// Raise a kernel's dynamic LDS limit to at least `bytes` (once per kernel
// and size: the attribute belongs to the function).
int allow_lds(const void* fn, size_t bytes);

// Constant tables on the device (uploaded on the first call; into the arena
// on every call while one is set).
int const_tables(xts_ctx* ctx, const xts_const_tables** out);
size_t const_tables_bytes();

// Brackets kernel launches with events on the context stream when timing is
// enabled (xts_ctx_enable_timing); no effect otherwise.
hipEvent_t pooled_event(xts_ctx* ctx);
struct TimedRegion {
    xts_ctx* ctx;
    const char* name;
    hipStream_t st;
    hipEvent_t start = nullptr;
    bool open = false;
    TimedRegion(xts_ctx* c, const char* n) : ctx(c), name(n), st(c->stream) {
        if (ctx->timing) {
            start = pooled_event(ctx);
            if (start && hipEventRecord(start, st) == hipSuccess) open = true;
        }
    }
    void stop() {
        if (!open) return;
        hipEvent_t e = pooled_event(ctx);
        if (e && hipEventRecord(e, st) == hipSuccess)
            ctx->timing_recs.push_back({name, start, e});
        open = false;
    }
    ~TimedRegion() { stop(); }
};
int to_device(xts_ctx* ctx, int slot, const void* host, size_t bytes, void** dev);
int activate(xts_ctx* ctx);

// ---------------------------------------------------------------- helpers
inline unsigned div_up(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// The reference's production cubature rule: Festa-Sommariva degree-7,
// 12-point rule on [-1, 1]^2 (mm_cubature/cubature.hpp:209-225).
extern const double kSquareDeg7[36];

}  // namespace xts
