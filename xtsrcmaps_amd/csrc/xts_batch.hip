// xts_batch.hip -- point-source model maps of many sources over several
// images (regions of interest) in one call: the batched engine.
//
// The reference computes one image's map for all its sources in one dense
// OpenMP loop (mm_cubature/cubature.implementation.cxx:33-53); here:
//   footprint_kernel  per source, the separation at which the normalised
//                     cumulative PSF integral (psf.partial_total_integral.cxx:
//                     23-48) reaches the containment fraction in each energy,
//                     their maximum plus the pixel half-diagonal, and the
//                     pixel box of that disk on the source's image
//   batches           a work list of (image, run of sources) batches ordered
//                     by cost (pixels x sources, longest first); each batch is
//                     one device model map in two calls (geometry, row
//                     weights, near pixels, correction factor; then the
//                     emit).  Consecutive batches alternate between the
//                     context and a second context with its own stream and
//                     buffers; a batch's first call waits for the previous
//                     batch's first call and its emit for the previous emit,
//                     so the row-weight pass (VALU-bound) of one batch runs
//                     beside the emit (HBM-store-bound) of the previous one;
//                     the row-weight pass can run on a limited grid of
//                     resident blocks ("rowweight_blocks") so that the emit's
//                     blocks find room
//                     on every compute unit
//   autotune          the batch size and row-weight grid, timed on the call's
//                     own workload the first time a shape is seen
//   graph             the whole multi-stream launch sequence captured once in
//                     a HIP graph and replayed by later identical calls
// Output: per source, its box's [E][h1 - h0][w1 - w0] FP32 block at the
// box's offset in one flat buffer (a box covering its whole image is the
// [E][H][W] plane of the dense [S, E, H, W] layout).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "xts_internal.hpp"

namespace xts {

__device__ inline double sep_node(int k) {
    return k == 0 ? 0.0 : 1e-4 * exp((k - 1) * kSepStep);
}

// One block per source; box_out [S][4] = h0, h1, w0, w1; radius_out [S] deg.
__global__ void __launch_bounds__(256)
footprint_kernel(const Wcs* __restrict__ imgs, const int* __restrict__ src_img,
                 const int* __restrict__ img_hw, const double* __restrict__ img_margin,
                 const double* __restrict__ src_sph,
                 const double* __restrict__ parint, int E, double containment,
                 double* __restrict__ radius_out, long long* __restrict__ box_out) {
    __shared__ double red[256];
    __shared__ double bb[4][256];
    __shared__ int bad_s;
    const int s = blockIdx.x, tid = threadIdx.x;
    const int img = src_img[s];
    const Wcs w = imgs[img];
    const int H = img_hw[2 * img], W = img_hw[2 * img + 1];
    const double* P = parint + (size_t)s * kND * E;
    double r = 0.0;
    for (int e = tid; e < E; e += 256) {
        double prev = P[e], re = 70.0;
        for (int k = 1; k < kND; ++k) {
            const double v = P[(size_t)k * E + e];
            if (v >= containment) {
                const double s0 = sep_node(k - 1), s1 = sep_node(k);
                re = v > prev ? s0 + (containment - prev) / (v - prev) * (s1 - s0) : s1;
                break;
            }
            prev = v;
        }
        r = fmax(r, re);
    }
    red[tid] = r;
    if (tid == 0) bad_s = 0;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k) red[tid] = fmax(red[tid], red[tid + k]);
        __syncthreads();
    }
    const double R = red[0] + img_margin[img];
    const double lon = src_sph[2 * s], lat = src_sph[2 * s + 1];
    const bool full = R >= 60.0 || fabs(lat) + R > 80.0;
    if (!full) {
        // a point of the boundary circle, bearing theta from north
        const double th = 2.0 * 3.14159265358979323846 * tid / 256.0;
        const double cl = cos(lon * kD2R), sl = sin(lon * kD2R), cb = cos(lat * kD2R),
                     sb = sin(lat * kD2R);
        const double c[3] = {cl * cb, sl * cb, sb};
        const double nv[3] = {-sb * cl, -sb * sl, cb}, ev[3] = {-sl, cl, 0.0};
        const double cr = cos(R * kD2R), sr = sin(R * kD2R), ct = cos(th), st = sin(th);
        double q[3];
        for (int k = 0; k < 3; ++k) q[k] = cr * c[k] + sr * (ct * nv[k] + st * ev[k]);
        const double plon = atan2(q[1], q[0]) * kR2D;
        const double plat = asin(fmin(1.0, fmax(-1.0, q[2]))) * kR2D;
        double p1, p2;
        wcs_sph2pix(w, plon < 0 ? plon + 360.0 : plon, plat, &p1, &p2);
        if (!isfinite(p1) || !isfinite(p2)) atomicOr(&bad_s, 1);
        bb[0][tid] = p1;
        bb[1][tid] = p1;
        bb[2][tid] = p2;
        bb[3][tid] = p2;
        if (tid == 0) {
            double q1, q2;
            wcs_sph2pix(w, lon, lat, &q1, &q2);
            if (!isfinite(q1) || !isfinite(q2)) atomicOr(&bad_s, 1);
            bb[0][0] = fmin(bb[0][0], q1);
            bb[1][0] = fmax(bb[1][0], q1);
            bb[2][0] = fmin(bb[2][0], q2);
            bb[3][0] = fmax(bb[3][0], q2);
        }
    }
    __syncthreads();
    if (!full) {
        for (int k = 128; k > 0; k >>= 1) {
            if (tid < k) {
                bb[0][tid] = fmin(bb[0][tid], bb[0][tid + k]);
                bb[1][tid] = fmax(bb[1][tid], bb[1][tid + k]);
                bb[2][tid] = fmin(bb[2][tid], bb[2][tid + k]);
                bb[3][tid] = fmax(bb[3][tid], bb[3][tid + k]);
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        long long h0 = 0, h1 = H, w0 = 0, w1 = W;
        if (!full && !bad_s) {
            // pixel index = FITS coordinate - 1 at the centre; two pixels of margin
            h0 = std::max(0LL, (long long)floor(bb[0][0] - 1.0) - 2);
            h1 = std::min((long long)H, (long long)ceil(bb[1][0] - 1.0) + 3);
            w0 = std::max(0LL, (long long)floor(bb[2][0] - 1.0) - 2);
            w1 = std::min((long long)W, (long long)ceil(bb[3][0] - 1.0) + 3);
            if (h1 <= h0 || w1 <= w0) h0 = h1 = w0 = w1 = 0;   // disk off the image
        }
        radius_out[s] = R;
        box_out[4 * s] = h0;
        box_out[4 * s + 1] = h1;
        box_out[4 * s + 2] = w0;
        box_out[4 * s + 3] = w1;
    }
}

// dense [n][E][H][W] -> each source's box block at its offset
__global__ void box_copy_kernel(const float* __restrict__ dense, int n, int E, int H, int W,
                                const xts_footprint* __restrict__ boxes, float* __restrict__ out) {
    const int s = blockIdx.y;
    const xts_footprint b = boxes[s];
    const long long bh = b.h1 - b.h0, bw = b.w1 - b.w0;
    const long long m = (long long)E * bh * bw;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
         k += (long long)gridDim.x * blockDim.x) {
        const long long e = k / (bh * bw), r = k % (bh * bw);
        const long long h = b.h0 + r / bw, w = b.w0 + r % bw;
        out[b.offset + k] = dense[(((size_t)s * E + e) * H + h) * W + w];
    }
}

struct Batch {
    int img;
    int64_t b0, n;
    bool direct;
    double cost;
};

static bool full_box(const xts_footprint& b, const xts_wcs& w) {
    return b.h0 == 0 && b.w0 == 0 && b.h1 == w.naxis1 && b.w1 == w.naxis2;
}

static int ensure_aux(xts_ctx* ctx) {
    if (!ctx->aux) {
        XTS_TRY(xts_ctx_create(ctx->device, &ctx->aux));
        XTS_HIP(hipSetDevice(ctx->device));
    }
    if (!ctx->fork_ev) XTS_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    if (!ctx->join_ev) XTS_HIP(hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming));
    xts_ctx* a = ctx->aux;
    for (xts_ctx* x : {ctx, a}) {
        if (!x->rows_ev) XTS_HIP(hipEventCreateWithFlags(&x->rows_ev, hipEventDisableTiming));
        if (!x->fin_ev) XTS_HIP(hipEventCreateWithFlags(&x->fin_ev, hipEventDisableTiming));
    }
    // batches switch images on both contexts, and a captured graph must not
    // depend on which image a context's geometry buffer held before: the
    // geometry of every batch is recomputed (the second context's setting
    // is kept; the call's own context is set for the call by the caller)
    a->geom_cache = false;
    a->tune_tiles_per_block = ctx->tune_tiles_per_block;
    a->surrogate = ctx->surrogate;
    a->surrogate8 = ctx->surrogate8;
    a->surrogate_h = ctx->surrogate_h;
    a->tune_sur_lo = ctx->tune_sur_lo;
    a->surrogate_ready = ctx->surrogate_ready;
    return XTS_OK;
}

struct BatchCall {
    int64_t n_images;
    const xts_wcs* wcs;
    const int64_t* begin;
    const double *src, *expo, *upsf, *parint;
    int64_t E;
    const double* rule;
    int64_t N;
    int32_t flags;
    const xts_footprint* boxes;
    float* out;
};

static std::vector<Batch> plan(const BatchCall& c, int64_t batch_sources) {
    double total = 0.0;
    for (int64_t i = 0; i < c.n_images; ++i)
        total += (double)(c.begin[i + 1] - c.begin[i]) * c.wcs[i].naxis1 * c.wcs[i].naxis2;
    std::vector<Batch> out;
    for (int64_t i = 0; i < c.n_images; ++i) {
        const double hw = (double)c.wcs[i].naxis1 * c.wcs[i].naxis2;
        const int64_t Si = c.begin[i + 1] - c.begin[i];
        // default: about six batches over the whole call
        int64_t B = batch_sources > 0 ? batch_sources
                                      : (int64_t)std::ceil(total / 6.0 / std::max(hw, 1.0));
        B = std::max<int64_t>(1, std::min(B, Si));
        for (int64_t b0 = c.begin[i]; b0 < c.begin[i + 1]; b0 += B) {
            const int64_t n = std::min(B, c.begin[i + 1] - b0);
            bool direct = true;
            for (int64_t s = b0; s < b0 + n && direct; ++s)
                direct = full_box(c.boxes[s], c.wcs[i]) &&
                         c.boxes[s].offset ==
                             c.boxes[b0].offset + (s - b0) * c.E * (int64_t)hw;
            out.push_back({(int)i, b0, n, direct, (double)n * hw});
        }
    }
    std::stable_sort(out.begin(), out.end(),
                     [](const Batch& a, const Batch& b) { return a.cost > b.cost; });
    return out;
}

// Enqueue every batch as a two-stage pipeline over two contexts: even
// batches on ctx's stream, odd ones on the second context's, which forks
// from and joins back into ctx's stream.  Each batch's row-weight half waits
// for the previous batch's (the PSF table and the factors follow the row
// weights), and its emit for the previous batch's emit, so the compute-bound
// row weights of batch k + 1 run beside the store-bound emit of batch k.
// A context's
// batches are in its stream's order, so batch k + 2 reuses batch k's
// buffers only after batch k is done.
static int run(xts_ctx* ctx, const BatchCall& c, int64_t batch_sources, int64_t rw_blocks,
               const xts_footprint* boxes_dev) {
    const std::vector<Batch> bs = plan(c, batch_sources);
    ctx->last_batch_sources = batch_sources;
    ctx->last_rw_blocks = rw_blocks;
    ctx->last_batches = (int64_t)bs.size();
    XTS_TRY(ensure_aux(ctx));
    xts_ctx* ctxs[2] = {ctx, ctx->aux};
    ctx->tune_rw_blocks = ctx->aux->tune_rw_blocks = rw_blocks;
    XTS_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
    XTS_HIP(hipStreamWaitEvent(ctx->aux->stream, ctx->fork_ev, 0));
    for (size_t k = 0; k < bs.size(); ++k) {
        const Batch& b = bs[k];
        xts_ctx* x = ctxs[k & 1];
        xts_ctx* y = ctxs[(k + 1) & 1];
        const xts_wcs& w = c.wcs[b.img];
        const size_t plane = (size_t)w.naxis1 * w.naxis2;
        const double* src = c.src + 2 * b.b0;
        const double* ex = c.expo + b.b0 * c.E;
        const double* up = c.upsf + (size_t)b.b0 * kND * c.E;
        const double* pa = c.parint + (size_t)b.b0 * kND * c.E;
        void* dst = c.out + c.boxes[b.b0].offset;
        if (!b.direct) XTS_TRY(slot_reserve(x, kSlotOut3, (size_t)b.n * c.E * plane * 4, &dst));
        if (k > 0) XTS_HIP(hipStreamWaitEvent(x->stream, y->rows_ev, 0));
        XTS_TRY(model_map_dev_call(x, &w, src, b.n, ex, up, pa, c.E, c.rule, c.N, c.flags, dst,
                                   3));
        XTS_HIP(hipEventRecord(x->rows_ev, x->stream));
        if (k > 0) XTS_HIP(hipStreamWaitEvent(x->stream, y->fin_ev, 0));
        XTS_TRY(model_map_dev_call(x, &w, src, b.n, ex, up, pa, c.E, c.rule, c.N, c.flags, dst,
                                   4));
        if (!b.direct) {
            XTS_TRY(activate(ctx));
            box_copy_kernel<<<dim3(256, (unsigned)b.n), 256, 0, x->stream>>>(
                (const float*)dst, (int)b.n, (int)c.E, (int)w.naxis1, (int)w.naxis2,
                boxes_dev + b.b0, c.out);
            XTS_HIP(hipGetLastError());
        }
        XTS_HIP(hipEventRecord(x->fin_ev, x->stream));
    }
    XTS_TRY(activate(ctx));
    XTS_HIP(hipEventRecord(ctx->join_ev, ctx->aux->stream));
    XTS_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0));
    return XTS_OK;
}

static std::string shape_key(const BatchCall& c) {
    std::string k = "E" + std::to_string(c.E) + "N" + std::to_string(c.N) + "f" +
                    std::to_string(c.flags);
    for (int64_t i = 0; i < c.n_images; ++i)
        k += ":" + std::to_string(c.begin[i + 1] - c.begin[i]) + "x" +
             std::to_string(c.wcs[i].naxis1) + "x" + std::to_string(c.wcs[i].naxis2);
    return k;
}

static int timed_run(xts_ctx* ctx, const BatchCall& c, int64_t bsrc, int64_t rwb,
                     const xts_footprint* boxes_dev, double* ms) {
    hipEvent_t a, b;
    XTS_HIP(hipEventCreate(&a));
    XTS_HIP(hipEventCreate(&b));
    int rc = run(ctx, c, bsrc, rwb, boxes_dev);   // warm: buffers sized
    if (rc == XTS_OK) {
        hipEventRecord(a, ctx->stream);
        rc = run(ctx, c, bsrc, rwb, boxes_dev);
        hipEventRecord(b, ctx->stream);
    }
    float f = 0.f;
    if (rc == XTS_OK && hipEventSynchronize(b) == hipSuccess) hipEventElapsedTime(&f, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = f;
    return rc;
}

// The tuned (batch sources, row-weight blocks) of the call's shape: the
// first call of a shape times candidates on its own inputs (every candidate
// writes the same output) and keeps the fastest.
static int tuned_shape(xts_ctx* ctx, const BatchCall& c, const xts_footprint* boxes_dev,
                       int64_t* bsrc, int64_t* rwb) {
    const std::string key = shape_key(c);
    for (const auto& t : ctx->tuned)
        if (t.key == key) {
            *bsrc = t.batch_sources;
            *rwb = t.rw_blocks;
            return XTS_OK;
        }
    int64_t maxS = 1;
    double total = 0.0, maxhw = 1.0;
    for (int64_t i = 0; i < c.n_images; ++i) {
        maxS = std::max(maxS, c.begin[i + 1] - c.begin[i]);
        const double hw = (double)c.wcs[i].naxis1 * c.wcs[i].naxis2;
        total += hw * (c.begin[i + 1] - c.begin[i]);
        maxhw = std::max(maxhw, hw);
    }
    std::vector<int64_t> bcands = {maxS};
    for (int nb : {2, 4, 8, 16}) {
        const int64_t B = std::max<int64_t>(1, (int64_t)std::ceil(total / nb / maxhw));
        if (std::find(bcands.begin(), bcands.end(), B) == bcands.end()) bcands.push_back(B);
    }
    const int64_t cu = std::max(ctx->cu_count, 1);
    xts_ctx::Tuned best{key, maxS, 0, 1e30};
    for (int64_t B : bcands)
        for (int64_t r : {(int64_t)0, 2 * cu, 4 * cu, 6 * cu}) {
            double ms;
            XTS_TRY(timed_run(ctx, c, B, r, boxes_dev, &ms));
            if (ms > 0 && ms < best.ms) best = {key, B, r, ms};
        }
    ctx->tuned.push_back(best);
    *bsrc = best.batch_sources;
    *rwb = best.rw_blocks;
    return XTS_OK;
}

}  // namespace xts

using namespace xts;

extern "C" {

int xts_model_map_footprints(xts_ctx* ctx, int64_t n_images, const xts_wcs* wcs,
                             const int64_t* src_begin, const double* src_sph,
                             const double* parint, int64_t E, double containment,
                             xts_footprint* boxes, double* radius, int64_t* total) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(n_images > 0 && n_images <= 4096 && wcs && src_begin && src_sph && parint &&
                  boxes && total,
              "xts_model_map_footprints: bad arguments");
    XTS_CHECK(E > 0 && E <= 4096 && containment > 0.0 && containment <= 1.0,
              "bad E or containment");
    XTS_CHECK(src_begin[0] == 0, "src_begin[0] must be 0");
    for (int64_t i = 0; i < n_images; ++i) {
        XTS_CHECK(src_begin[i + 1] >= src_begin[i], "src_begin not ascending");
        XTS_CHECK(wcs[i].naxis1 > 0 && wcs[i].naxis2 > 0 && wcs[i].naxis1 <= 65536 &&
                      wcs[i].naxis2 <= 65536 &&
                      (wcs[i].proj == XTS_PROJ_CAR || wcs[i].proj == XTS_PROJ_AIT),
                  "bad image %lld", (long long)i);
    }
    const int64_t S = src_begin[n_images];
    XTS_CHECK(S > 0 && S <= (1 << 24), "bad source count %lld", (long long)S);
    std::vector<Wcs> w((size_t)n_images);
    std::vector<int> hw((size_t)n_images * 2), simg((size_t)S);
    std::vector<double> margin((size_t)n_images);
    for (int64_t i = 0; i < n_images; ++i) {
        w[i] = wcs_setup(wcs[i]);
        // the pixel half-diagonal and a half more (degrees)
        margin[i] = 0.75 * std::sqrt(wcs[i].cdelt1 * wcs[i].cdelt1 + wcs[i].cdelt2 * wcs[i].cdelt2);
        hw[2 * i] = (int)wcs[i].naxis1;
        hw[2 * i + 1] = (int)wcs[i].naxis2;
        for (int64_t s = src_begin[i]; s < src_begin[i + 1]; ++s) simg[s] = (int)i;
    }
    void *dw, *dh, *ds, *dr, *db, *dm;
    XTS_TRY(to_device(ctx, kSlotIn3, margin.data(), margin.size() * sizeof(double), &dm));
    XTS_TRY(to_device(ctx, kSlotIn5, w.data(), w.size() * sizeof(Wcs), &dw));
    XTS_TRY(to_device(ctx, kSlotIn6, hw.data(), hw.size() * sizeof(int), &dh));
    XTS_TRY(to_device(ctx, kSlotIn7, simg.data(), simg.size() * sizeof(int), &ds));
    XTS_TRY(slot_reserve(ctx, kSlotWork2, (size_t)S * 8 * 5, &dr));
    db = (char*)dr + (size_t)S * 8;
    footprint_kernel<<<(unsigned)S, 256, 0, ctx->stream>>>(
        (const Wcs*)dw, (const int*)ds, (const int*)dh, (const double*)dm, src_sph, parint,
        (int)E, containment,
        (double*)dr, (long long*)db);
    XTS_HIP(hipGetLastError());
    std::vector<long long> hb((size_t)S * 4);
    std::vector<double> hr((size_t)S);
    XTS_HIP(hipMemcpyAsync(hb.data(), db, hb.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipMemcpyAsync(hr.data(), dr, hr.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    int64_t off = 0;
    for (int64_t s = 0; s < S; ++s) {
        xts_footprint& b = boxes[s];
        b.h0 = hb[4 * s];
        b.h1 = hb[4 * s + 1];
        b.w0 = hb[4 * s + 2];
        b.w1 = hb[4 * s + 3];
        b.offset = off;
        off += E * (b.h1 - b.h0) * (b.w1 - b.w0);
        if (radius) radius[s] = hr[s];
    }
    *total = off;
    return XTS_OK;
}

int xts_point_src_model_maps_batched_dev(xts_ctx* ctx, int64_t n_images, const xts_wcs* wcs,
                                         const int64_t* src_begin, const double* src_sph,
                                         const double* exposure, const double* upsf,
                                         const double* parint, int64_t E, const double* rule,
                                         int64_t N, int32_t flags, const xts_footprint* boxes,
                                         float* out, int64_t out_elems) {
    XTS_TRY(activate(ctx));
    XTS_CHECK(n_images > 0 && n_images <= 4096 && wcs && src_begin && src_sph && exposure &&
                  upsf && parint && boxes && out,
              "xts_point_src_model_maps_batched_dev: bad arguments");
    XTS_CHECK((flags & XTS_LAYOUT_SEHW) && (flags & XTS_OUT_F32) && !(flags & XTS_COMPUTE_F64),
              "the batched engine writes FP32 [E][h][w] blocks (XTS_LAYOUT_SEHW | XTS_OUT_F32, "
              "FP32 cubature)");
    XTS_CHECK(src_begin[0] == 0, "src_begin[0] must be 0");
    const int64_t S = src_begin[n_images];
    XTS_CHECK(S > 0 && S <= (1 << 24) && E > 0 && E <= 4096, "bad sizes");
    for (int64_t i = 0; i < n_images; ++i) {
        XTS_CHECK(src_begin[i + 1] >= src_begin[i], "src_begin not ascending");
        for (int64_t s = src_begin[i]; s < src_begin[i + 1]; ++s) {
            const xts_footprint& b = boxes[s];
            XTS_CHECK(b.h0 >= 0 && b.h0 <= b.h1 && b.h1 <= wcs[i].naxis1 && b.w0 >= 0 &&
                          b.w0 <= b.w1 && b.w1 <= wcs[i].naxis2 && b.offset >= 0 &&
                          b.offset + E * (b.h1 - b.h0) * (b.w1 - b.w0) <= out_elems,
                      "source %lld: box outside its image or the output", (long long)s);
        }
    }
    BatchCall c{n_images, wcs, src_begin, src_sph, exposure, upsf, parint, E, rule, N, flags,
                boxes, out};
    // the boxes on the device, for batches that copy into boxes
    void* dbx;
    XTS_TRY(slot_reserve(ctx, kSlotIn4, (size_t)S * sizeof(xts_footprint), &dbx));
    bool need_dev_boxes = false;
    for (int64_t i = 0; i < n_images && !need_dev_boxes; ++i)
        for (int64_t s = src_begin[i]; s < src_begin[i + 1] && !need_dev_boxes; ++s)
            need_dev_boxes = !full_box(boxes[s], wcs[i]);
    if (need_dev_boxes)
        XTS_HIP(hipMemcpyAsync(dbx, boxes, (size_t)S * sizeof(xts_footprint),
                               hipMemcpyHostToDevice, ctx->stream));
    // every batch recomputes its image's geometry (see ensure_aux); the
    // context's own setting is restored on return
    struct Keep {
        xts_ctx* c;
        int64_t rw;
        bool geom;
        ~Keep() {
            c->tune_rw_blocks = rw;
            c->geom_cache = geom;
            c->geom_valid = false;
        }
    } keep{ctx, ctx->tune_rw_blocks, ctx->geom_cache};
    ctx->geom_cache = false;
    int64_t bsrc = ctx->tune_batch_sources, rwb = ctx->tune_rw_blocks;
    if (ctx->autotune) XTS_TRY(tuned_shape(ctx, c, (const xts_footprint*)dbx, &bsrc, &rwb));
    if (!ctx->use_graph || need_dev_boxes) return run(ctx, c, bsrc, rwb, (const xts_footprint*)dbx);
    XTS_TRY(ensure_aux(ctx));
    // graph: the call's key is its shape, tuning and every pointer it reads
    // or writes; a new key is run eagerly once (buffers sized), captured,
    // and replayed from then on
    char ptrs[256];
    std::snprintf(ptrs, sizeof ptrs, "|%p|%p|%p|%p|%p|%p|%lld|%lld", (const void*)src_sph,
                  (const void*)exposure, (const void*)upsf, (const void*)parint, (void*)out,
                  (const void*)rule, (long long)bsrc, (long long)rwb);
    std::string key = shape_key(c) + ptrs;
    for (int64_t s = 0; s < S; ++s) key += ":" + std::to_string(boxes[s].offset);
    for (int64_t i = 0; i < n_images; ++i) {
        char wbuf[sizeof(xts_wcs) * 2 + 1];
        const unsigned char* p = (const unsigned char*)&wcs[i];
        for (size_t k = 0; k < sizeof(xts_wcs); ++k) std::snprintf(wbuf + 2 * k, 3, "%02x", p[k]);
        key += wbuf;
    }
    // the buffers and uploads the graph's kernels were captured with
    auto gens = [&] {
        return "|g" + std::to_string(ctx->generation) + "." + std::to_string(ctx->aux->generation);
    };
    if (ctx->graph_exec && ctx->graph_key == key + gens()) {
        XTS_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
        ++ctx->graph_replays;
        return XTS_OK;
    }
    if (ctx->graph_exec) {
        hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
    }
    XTS_CHECK(!ctx->timing, "graph capture with kernel timing enabled");
    // eagerly once (this call's result; buffers sized), then captured
    XTS_TRY(run(ctx, c, bsrc, rwb, (const xts_footprint*)dbx));
    XTS_HIP(hipStreamSynchronize(ctx->stream));
    const std::string gen_before = gens();
    hipGraph_t g = nullptr;
    XTS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    const int rc = run(ctx, c, bsrc, rwb, (const xts_footprint*)dbx);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc == XTS_OK && gens() != gen_before) {
        if (g) hipGraphDestroy(g);
        return fail(XTS_ERR_DEVICE, "batched model map: buffers changed during graph capture");
    }
    if (rc != XTS_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    XTS_HIP(e);
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    XTS_HIP(ie);
    ctx->graph_exec = ex;
    ctx->graph_key = key + gens();
    ++ctx->graph_captures;
    return XTS_OK;
}

}  // extern "C"
