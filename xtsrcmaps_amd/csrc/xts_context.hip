// xts_context.hip -- context, device arena, error reporting and WCS setup.
#include <cstdarg>
#include <cstring>

#include <cmath>
#include <mutex>
#include <utility>
#include <vector>

#include "xts_internal.hpp"

namespace xts {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int activate(xts_ctx* ctx) {
    XTS_CHECK(ctx != nullptr, "null context");
    XTS_HIP(hipSetDevice(ctx->device));
    return XTS_OK;
}

int slot_reserve(xts_ctx* ctx, int slot, size_t bytes, void** out) {
    if (ctx->arena.base) {
        // no allocation: the next aligned piece of the arena
        const size_t off = ctx->arena.used;
        const size_t n = arena_bytes(bytes);
        if (off > ctx->arena.cap || n > ctx->arena.cap - off)
            return fail(XTS_ERR_INVALID, "call needs more than the %zu MiB no-allocation scratch",
                        ctx->arena.cap >> 20);
        *out = ctx->arena.base + off;
        ctx->arena.used = off + n;
        return XTS_OK;
    }
    Slot& s = ctx->slots[slot];
    if (bytes == 0) bytes = 256;
    if (s.bytes < bytes) {
        if (s.ptr) {
            XTS_HIP(hipStreamSynchronize(ctx->stream));
            XTS_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        // round up so that slowly growing calls do not reallocate each time
        size_t want = bytes + bytes / 4;
        want = (want + 4095) & ~size_t(4095);
        XTS_HIP(hipMalloc(&s.ptr, want));
        s.bytes = want;
        ++ctx->generation;
    }
    *out = s.ptr;
    return XTS_OK;
}

// separations (psf/psf.hxx:24-31), 2 R2D / sep, theta and sin theta, then
// 2 R2D / sep in FP32 (half the doubles)
static const std::vector<double>& host_const_tables() {
    static const std::vector<double> h = [] {
        std::vector<double> v(kND * 4 + kND / 2 + 1);
        double* sep = v.data();
        double* inv = sep + kND;
        double* th = inv + kND;
        float* inv32 = reinterpret_cast<float*>(th + 2 * kND);
        const double xstep = std::log(70.0 / 1e-4) / 399.0;
        sep[0] = 0.0;
        for (int i = 0; i < 400; ++i) sep[i + 1] = 1e-4 * std::exp(i * xstep);
        for (int i = 0; i < kND; ++i) {
            inv[i] = sep[i] > 0 ? 2.0 * kR2D / sep[i] : 0.0;
            inv32[i] = (float)inv[i];
            th[i] = kD2R * sep[i];
            th[kND + i] = std::sin(th[i]);
        }
        return v;
    }();
    return h;
}

size_t const_tables_bytes() { return host_const_tables().size() * 8; }

static void set_const_ptrs(xts_const_tables* c, double* dd) {
    c->sep = dd;
    c->inv64 = dd + kND;
    c->theta = dd + 2 * kND;
    c->inv32 = reinterpret_cast<float*>(dd + 4 * kND);
}

int const_tables(xts_ctx* ctx, const xts_const_tables** out) {
    const std::vector<double>& h = host_const_tables();
    if (ctx->arena.base) {
        // inside the arena, uploaded on the context stream for this call only
        void* d;
        XTS_TRY(slot_reserve(ctx, kSlotConst, h.size() * 8, &d));
        XTS_HIP(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        set_const_ptrs(&ctx->arena_consts, (double*)d);
        *out = &ctx->arena_consts;
        return XTS_OK;
    }
    if (!ctx->consts_ready) {
        void* d;
        XTS_TRY(slot_reserve(ctx, kSlotConst, h.size() * 8, &d));
        // a blocking copy: done once, before any kernel reads the tables
        XTS_HIP(hipMemcpy(d, h.data(), h.size() * 8, hipMemcpyHostToDevice));
        set_const_ptrs(&ctx->consts, (double*)d);
        ctx->consts_ready = true;
    }
    *out = &ctx->consts;
    return XTS_OK;
}

hipEvent_t pooled_event(xts_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int allow_lds(const void* fn, size_t bytes) {
    static std::mutex mu;
    static std::vector<std::pair<const void*, size_t>> done;
    std::lock_guard<std::mutex> lock(mu);
    for (auto& d : done)
        if (d.first == fn && d.second >= bytes) return XTS_OK;
    XTS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    done.emplace_back(fn, bytes);
    return XTS_OK;
}

int to_device(xts_ctx* ctx, int slot, const void* host, size_t bytes, void** dev) {
    XTS_TRY(slot_reserve(ctx, slot, bytes, dev));
    if (bytes) XTS_HIP(hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    return XTS_OK;
}

// Published Festa-Sommariva low-cardinality square rule of degree 7
// (cardinality 12, all points inside, positive weights), as used by the
// reference's production cubature (mm_cubature/cubature.hpp:209-225).
const double kSquareDeg7[36] = {
    3.9383136101878902e-01, -3.6699793434042738e-01, 5.2092862230230441e-01,
    8.0196847620659251e-01, 8.0992592975809186e-01, 2.3740240561379911e-01,
    -1.8327327664247231e-02, 9.2596092593108370e-01, 2.4166897208389651e-01,
    3.6699793434042738e-01, 3.9383136101878902e-01, 5.2092862230230441e-01,
    -8.0992592975809186e-01, 8.0196847620659251e-01, 2.3740240561379911e-01,
    -9.2596092593108370e-01, -1.8327327664247231e-02, 2.4166897208389651e-01,
    -3.9383136101878902e-01, 3.6699793434042738e-01, 5.2092862230230441e-01,
    -8.0196847620659251e-01, -8.0992592975809186e-01, 2.3740240561379911e-01,
    1.8327327664247231e-02, -9.2596092593108370e-01, 2.4166897208389651e-01,
    -3.6699793434042738e-01, -3.9383136101878902e-01, 5.2092862230230441e-01,
    8.0992592975809186e-01, -8.0196847620659251e-01, 2.3740240561379911e-01,
    9.2596092593108370e-01, 1.8327327664247231e-02, 2.4166897208389651e-01,
};

}  // namespace xts

using namespace xts;

extern "C" {

const char* xts_last_error(void) { return g_err; }

int xts_abi_version(void) { return XTS_ABI_VERSION; }

int xts_ctx_create(int device, xts_ctx** out) {
    XTS_CHECK(out != nullptr, "xts_ctx_create: null out");
    *out = nullptr;
    int n = 0;
    XTS_HIP(hipGetDeviceCount(&n));
    XTS_CHECK(device >= 0 && device < n, "xts_ctx_create: device %d of %d", device, n);
    xts_ctx* c = new xts_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        delete c;
        return fail(XTS_ERR_DEVICE, "xts_ctx_create: %s", hipGetErrorString(e));
    }
    c->stream = c->own_stream;
    // the no-allocation entry point's scratch address on this device (its
    // status is kept and reported by that entry point)
    noalloc_scratch_init(device);
    *out = c;
    return XTS_OK;
}

int xts_ctx_destroy(xts_ctx* ctx) {
    if (!ctx) return XTS_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto& r : ctx->timing_recs) {
        hipEventDestroy(r.start);
        hipEventDestroy(r.stop);
    }
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) hipEventDestroy(ctx->join_ev);
    if (ctx->rows_ev) hipEventDestroy(ctx->rows_ev);
    if (ctx->fin_ev) hipEventDestroy(ctx->fin_ev);
    if (ctx->pipe_stream) {
        hipStreamSynchronize(ctx->pipe_stream);
        hipStreamDestroy(ctx->pipe_stream);
    }
    for (auto e : ctx->pipe_ev) hipEventDestroy(e);
    if (ctx->aux) xts_ctx_destroy(ctx->aux);
    hipSetDevice(ctx->device);
    for (auto& s : ctx->slots)
        if (s.ptr) hipFree(s.ptr);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return XTS_OK;
}

int xts_ctx_set_stream(xts_ctx* ctx, void* stream) {
    XTS_CHECK(ctx != nullptr, "null context");
    hipStream_t next = stream ? (hipStream_t)stream : ctx->own_stream;
    if (next != ctx->stream) {
        // cached geometry and rule were written in the old stream's order:
        // rebuild them in the new one's
        ctx->geom_valid = false;
        ctx->rule.dev = nullptr;
    }
    // a pending first half of a model map was enqueued in the old stream's
    // order; its second half is refused from now on
    ctx->pending_rows.valid = false;
    ctx->stream = next;
    return XTS_OK;
}

int xts_ctx_set_tuning(xts_ctx* ctx, const char* key, int64_t value) {
    XTS_CHECK(ctx != nullptr && key != nullptr, "null argument");
    XTS_CHECK(value >= 0 && value <= 1 << 20, "tuning value %lld out of range", (long long)value);
    // a pending first half of a split model map was shaped by the old
    // settings: its second half is refused from now on
    ctx->pending_rows.valid = false;
    if (strcmp(key, "rowweight_frames") == 0) {
        ctx->tune_rw_frames = value;
        return XTS_OK;
    }
    if (strcmp(key, "emit_wide_apart") == 0) {
        XTS_CHECK(value == 0 || value == 1, "emit_wide_apart is 0 or 1");
        ctx->tune_emit_wide_apart = value;
        return XTS_OK;
    }
    if (strcmp(key, "near_emit_blocks") == 0) {
        ctx->tune_near_emit_gx = value;
        return XTS_OK;
    }
    if (strcmp(key, "rowweight_legacy") == 0) {
        XTS_CHECK(value == 0 || value == 1, "rowweight_legacy is 0 or 1");
        ctx->tune_rw_legacy = value;
        return XTS_OK;
    }
    if (strcmp(key, "emit_energies") == 0) {
        XTS_CHECK(value == 0 || value == 20 || value == 40, "emit_energies is 20 or 40");
        ctx->tune_emit_ec = value;
        return XTS_OK;
    }
    if (strcmp(key, "emit_priority") == 0) {
        XTS_CHECK(value <= 3, "emit_priority is 0 .. 3");
        ctx->tune_emit_prio = value;
        return XTS_OK;
    }
    if (strcmp(key, "rowweight_frames_cu") == 0) {
        XTS_CHECK(value <= 16, "rowweight_frames_cu at most 16");
        ctx->tune_rw_frames_cu = value;
        return XTS_OK;
    }
    if (strcmp(key, "surrogate_bf16_lo") == 0) {
        XTS_CHECK(value == 0 || value == 1, "surrogate_bf16_lo is 0 or 1");
        ctx->tune_sur_lo = value;
        return XTS_OK;
    }
    if (strcmp(key, "surrogate_beside") == 0) {
        XTS_CHECK(value >= 0 && value <= 8, "surrogate_beside is 0 .. 8");
        ctx->tune_sur_beside = value;
        return XTS_OK;
    }
    if (strcmp(key, "pipeline_batches") == 0) {
        ctx->tune_pipe_batches = value;
        return XTS_OK;
    }
    if (strcmp(key, "geometry_cache") == 0) {
        XTS_CHECK(value == 0 || value == 1, "geometry_cache is 0 or 1");
        ctx->geom_cache = value != 0;
        ctx->geom_valid = false;
        return XTS_OK;
    }
    if (strcmp(key, "rowweight_blocks") == 0) {
        ctx->tune_rw_blocks = value;
        return XTS_OK;
    }
    if (strcmp(key, "batch_sources") == 0) {
        ctx->tune_batch_sources = value;
        return XTS_OK;
    }
    if (strcmp(key, "sum_batch_sources") == 0) {
        XTS_CHECK(value >= 0, "sum_batch_sources >= 0");
        ctx->tune_sum_batch = value;
        return XTS_OK;
    }
    if (strcmp(key, "sum_group") == 0) {
        XTS_CHECK(value == 4 || value == 8, "sum_group is 4 or 8");
        ctx->tune_sum_group = value;
        return XTS_OK;
    }
    if (strcmp(key, "autotune") == 0) {
        XTS_CHECK(value == 0 || value == 1, "autotune is 0 or 1");
        ctx->autotune = (int)value;
        return XTS_OK;
    }
    if (strcmp(key, "graph") == 0) {
        XTS_CHECK(value == 0 || value == 1, "graph is 0 or 1");
        ctx->use_graph = (int)value;
        return XTS_OK;
    }
    if (strcmp(key, "tiles_per_block") == 0) ctx->tune_tiles_per_block = (int)value;
    else return fail(XTS_ERR_INVALID, "unknown tuning key '%s'", key);
    return XTS_OK;
}

int xts_ctx_get_tuning(xts_ctx* ctx, const char* key, int64_t* value) {
    XTS_CHECK(ctx != nullptr && key != nullptr && value != nullptr, "null argument");
    if (strcmp(key, "batch_sources") == 0) *value = ctx->last_batch_sources;
    else if (strcmp(key, "rowweight_blocks") == 0) *value = ctx->last_rw_blocks;
    else if (strcmp(key, "batches") == 0) *value = ctx->last_batches;
    else if (strcmp(key, "graph_replays") == 0) *value = ctx->graph_replays;
    else if (strcmp(key, "graph_captures") == 0) *value = ctx->graph_captures;
    else return fail(XTS_ERR_INVALID, "unknown tuning key '%s'", key);
    return XTS_OK;
}

int xts_ctx_enable_timing(xts_ctx* ctx, int enable) {
    XTS_CHECK(ctx != nullptr, "null context");
    ctx->timing = enable ? 1 : 0;
    return XTS_OK;
}

int xts_ctx_kernel_times(xts_ctx* ctx, const char* name, double* total_ms, int64_t* count) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(name && total_ms && count, "null argument");
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    int64_t n = 0;
    std::vector<xts_timing_rec> keep;
    for (auto& r : ctx->timing_recs) {
        if (strcmp(r.name, name) == 0) {
            float ms = 0.f;
            XTS_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
            tot += ms;
            ++n;
            ctx->event_pool.push_back(r.start);
            ctx->event_pool.push_back(r.stop);
        } else {
            keep.push_back(r);
        }
    }
    ctx->timing_recs.swap(keep);
    *total_ms = tot;
    *count = n;
    return XTS_OK;
}

int xts_ctx_synchronize(xts_ctx* ctx) {
    XTS_TRY(activate(ctx));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    return XTS_OK;
}

int xts_costhetas(int64_t nbins, double cosmin, int32_t thetabin, double* out) {
    XTS_CHECK(out != nullptr && nbins > 0, "xts_costhetas: bad arguments");
    for (int64_t i = 0; i < nbins; ++i) {
        double f = (i + 0.5) / (double)nbins;
        f = thetabin ? f * f : f;
        out[i] = 1. - f * (1. - cosmin);
    }
    return XTS_OK;
}

int xts_livetime_efficiency_factors(const double* logEs, int64_t E, const float* p0,
                                    const float* p1, double* ltf) {
    XTS_CHECK(logEs && p0 && p1 && ltf && E > 0, "xts_livetime_efficiency_factors: bad arguments");
    const float* ps[2] = {p1, p0};  // row 0 from p1, row 1 from p0
    for (int r = 0; r < 2; ++r) {
        const float* p = ps[r];
        const float a0 = p[0], b0 = p[1], a1 = p[2], lb1 = p[3], a2 = p[4], lb2 = p[5];
        const float b1 = (a0 - a1) * lb1 + b0;
        const float b2 = (a1 - a2) * lb2 + b1;
        for (int64_t e = 0; e < E; ++e) {
            const double x = logEs[e];
            ltf[r * E + e] = x < lb1 ? a0 * x + b0 : (x < lb2 ? a1 * x + (double)b1 : a2 * x + (double)b2);
        }
    }
    return XTS_OK;
}

}  // extern "C"
