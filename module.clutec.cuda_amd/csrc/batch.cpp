// C ABI of the batched matcher: many stereo pairs of uneven size and disparity range
// in one persistent work-list launch, captured into a HIP graph and replayed.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace capi;

namespace {

constexpr int kMinChunk = 32;  // smallest work item the workspace is sized for

struct BatchExt : DriverExt {
    stereo_context* ctx = nullptr;
    int npairs = 0;
    stereo_matcher_params mp{};
    int kind = 0;
    size_t elem = 1;
    stereo_batch_tuning tuning{};
    // work-item length per pair (0: tuning.chunk); set per class by the autotuner or a
    // tuning table
    std::vector<int> pair_chunk;
    hipEvent_t ev_done = nullptr;  // recorded after every process call's last launch
    bool done_recorded = false;
    // the inputs of one call: left and right image of every pair into its padded buffer,
    // as one batched census or copy launch (base: census tiles or rows before the image)
    std::vector<stereo::ImageJob> jobs;
    stereo::ImageJob* jobs_dev = nullptr;
    int job_units = 0;  // census tiles or rows of all images
    std::vector<uint8_t*> padL, padR;
    std::vector<stereo::BmFrame> frames;  // host copy of the device frame table
    stereo::BmFrame* frames_dev = nullptr;
    stereo::BmTile* tiles_dev = nullptr;
    stereo::BmItem* items_dev = nullptr;
    int* sched_dev = nullptr;
    int* counter_dev = nullptr;  // item counter of the list matcher (reset by its finish)
    int* nitems_dev = nullptr;
    int ntiles = 0, nitems = 0, max_items = 0;
    // key planes of every pair (left, right), zero at every matcher launch: cleared at
    // configure and by every merge; planes_dirty when a matcher ran without its merge
    uint32_t* lplane = nullptr;
    uint32_t* rplane = nullptr;
    size_t plane_elems = 0;
    bool planes_dirty = false;
    int total_rows = 0, max_width = 0;
    hipStream_t cap = nullptr;  // capture stream
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> graph_key;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_key.clear();
    }

    ~BatchExt() override
    {
        drop_graph();
        if (cap) (void)hipStreamDestroy(cap);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_done) (void)hipEventDestroy(ev_done);
        for (uint8_t* p : padL)
            if (p) (void)hipFree(p);
        for (uint8_t* p : padR)
            if (p) (void)hipFree(p);
        for (void* p : {(void*)frames_dev, (void*)tiles_dev, (void*)items_dev, (void*)sched_dev, (void*)counter_dev, (void*)nitems_dev,
                        (void*)lplane, (void*)rplane, (void*)jobs_dev})
            if (p) (void)hipFree(p);
    }
};

// Tuning classes of the pairs: the disparity count and the tile count (small below 64
// tiles of 64 x 56, medium below 256, large), the two sizes that decide how a pair's
// work splits into items.
const char* const kSizeClass[3] = {"small", "medium", "large"};
int size_class(const stereo::BmFrame& f)
{
    const int t = f.ntx * f.nty;
    return t < 64 ? 0 : (t < 256 ? 1 : 2);
}

int chunk_of(const BatchExt* x, int f)
{
    const int c = x->pair_chunk.empty() ? 0 : x->pair_chunk[f];
    return c ? c : x->tuning.chunk;
}

// Items of every tile of every pair, tile by tile (the slots the merge reads), the
// tile table, and the schedule: 0 = longest items of the largest frames first,
// 1 = table order.
stereo_status upload_work(BatchExt* x)
{
    HIP_TRY(x->ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");  // no launch of the old tables in flight
    std::vector<stereo::BmItem> items;
    std::vector<stereo::BmTile> tiles;
    for (int f = 0; f < x->npairs; ++f) {
        const stereo::BmFrame& fr = x->frames[f];
        const int chunk = chunk_of(x, f);
        for (int ty = 0; ty < fr.nty; ++ty)
            for (int tx = 0; tx < fr.ntx; ++tx) {
                stereo::BmTile t;
                t.lo = 0;
                t.hi = fr.D;
                t.first_item = (int)items.size();
                t.nitems = 0;
                for (int ds = 0; ds < fr.D; ds += chunk) {
                    stereo::BmItem it;
                    it.frame = f;
                    it.tx = tx;
                    it.ty = ty;
                    it.ds = ds;
                    it.de = std::min(fr.D, ds + chunk);
                    items.push_back(it);
                    ++t.nitems;
                }
                tiles.push_back(t);
            }
    }
    const int n = (int)items.size();
    if (n > x->max_items)
        return fail(x->ctx, STEREO_ERROR_INVALID_CONFIG, "batch: %d work items, more than the %d configured", n,
                    x->max_items);
    std::vector<int> sched(n);
    std::iota(sched.begin(), sched.end(), 0);
    if (x->tuning.order == 0)
        std::stable_sort(sched.begin(), sched.end(), [&](int a, int b) {
            const int la = items[a].de - items[a].ds, lb = items[b].de - items[b].ds;
            if (la != lb) return la > lb;
            const stereo::BmFrame& fa = x->frames[items[a].frame];
            const stereo::BmFrame& fb = x->frames[items[b].frame];
            return (long)fa.W * fa.H > (long)fb.W * fb.H;
        });
    x->nitems = n;
    x->ntiles = (int)tiles.size();
    HIP_TRY(x->ctx, hipMemcpy(x->items_dev, items.data(), n * sizeof(stereo::BmItem), hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(x->ctx, hipMemcpy(x->sched_dev, sched.data(), n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    HIP_TRY(x->ctx, hipMemcpy(x->tiles_dev, tiles.data(), tiles.size() * sizeof(stereo::BmTile), hipMemcpyHostToDevice),
            "hipMemcpy");
    HIP_TRY(x->ctx, hipMemcpy(x->nitems_dev, &n, sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    const int zero = 0;  // the item counter starts every matcher launch at 0
    HIP_TRY(x->ctx, hipMemcpy(x->counter_dev, &zero, sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    x->drop_graph();
    return STEREO_OK;
}

// The job table of the call's input images; rewritten (after the previous call's
// launches, which read it) when the images change.  Before any capture or launch.
stereo_status upload_jobs(BatchExt* x, const stereo_image* left, const stereo_image* right)
{
    std::vector<stereo::ImageJob> jobs(2 * x->npairs);
    int base = 0;
    for (int i = 0; i < x->npairs; ++i) {
        const stereo::BmFrame& f = x->frames[i];
        for (int k = 0; k < 2; ++k) {
            stereo::ImageJob& j = jobs[2 * i + k];
            j.src = view_of(k ? &right[i] : &left[i]);
            j.dst = k ? f.R : f.L;
            j.base = base;
            base += x->kind == stereo::BM_COST_CENSUS ? stereo::census_tiles(f.W, f.H) : f.H;
        }
    }
    const bool same = jobs.size() == x->jobs.size() &&
                      std::memcmp(jobs.data(), x->jobs.data(), jobs.size() * sizeof(stereo::ImageJob)) == 0;
    if (same) return STEREO_OK;
    if (x->done_recorded) HIP_TRY(x->ctx, hipEventSynchronize(x->ev_done), "hipEventSynchronize");
    HIP_TRY(x->ctx, hipMemcpy(x->jobs_dev, jobs.data(), jobs.size() * sizeof(stereo::ImageJob), hipMemcpyHostToDevice),
            "hipMemcpy");
    x->jobs = jobs;
    x->job_units = base;
    return STEREO_OK;
}

stereo_status enqueue(stereo_driver* d, BatchExt* x, const stereo_image* left, const stereo_image* right, hipStream_t s)
{
    stereo_context* ctx = d->ctx;
    (void)left;
    (void)right;  // read through the job table
    if (x->kind == stereo::BM_COST_CENSUS)
        HIP_TRY(ctx, stereo::launch_census_batch(x->jobs_dev, 2 * x->npairs, x->job_units, x->mp.census_w, x->mp.census_h, s),
                "census launch");
    else
        HIP_TRY(ctx, stereo::launch_pad_copy_batch(x->jobs_dev, 2 * x->npairs, x->job_units, s), "copy launch");
    stereo::BmListParams lp;
    lp.frames = x->frames_dev;
    lp.tiles = x->tiles_dev;
    lp.items = x->items_dev;
    lp.nitems = x->nitems_dev;
    lp.sched = x->sched_dev;
    lp.counter = x->counter_dev;
    lp.lplane = x->lplane;
    lp.rplane = x->rplane;
    lp.nframes = x->npairs;
    lp.total_rows = x->total_rows;
    const int grid = std::max(1, std::min(x->tuning.grid, x->nitems));
    x->planes_dirty = true;  // until the merge is enqueued behind the matcher
    HIP_TRY(ctx, stereo::launch_bm_list(x->kind, x->mp.window_w, lp, grid, s), "list matcher launch");
    HIP_TRY(ctx, stereo::launch_bm_list_finish(lp, x->max_width, s), "list finish launch");
    x->planes_dirty = false;
    return STEREO_OK;
}

// A matcher launch without its merge leaves keys in the planes: clear them (a kernel, so
// that the call may also run under a caller's graph capture).
stereo_status clear_if_dirty(BatchExt* x, hipStream_t s)
{
    if (!x->planes_dirty) return STEREO_OK;
    HIP_TRY(x->ctx, stereo::launch_bm_clear(x->lplane, x->plane_elems, s), "clear launch");
    HIP_TRY(x->ctx, stereo::launch_bm_clear(x->rplane, x->plane_elems, s), "clear launch");
    x->planes_dirty = false;
    return STEREO_OK;
}

// Outputs are part of the frame table; returns true when it changed.
bool set_outputs(BatchExt* x, stereo_image* disp, stereo_image* disp_lr, stereo_image* disp_rl)
{
    bool changed = false;
    for (int i = 0; i < x->npairs; ++i) {
        stereo::BmFrame& f = x->frames[i];
        const View a = view_of(&disp[i]);
        const View b = (disp_lr && disp_lr[i].data) ? view_of(&disp_lr[i]) : null_view();
        const View c = (disp_rl && disp_rl[i].data) ? view_of(&disp_rl[i]) : null_view();
        if (f.disp.ptr != a.ptr || f.disp.pitch != a.pitch || f.disp_l.ptr != b.ptr || f.disp_l.pitch != b.pitch ||
            f.disp_r.ptr != c.ptr || f.disp_r.pitch != c.pitch) {
            f.disp = a;
            f.disp_l = b;
            f.disp_r = c;
            changed = true;
        }
    }
    return changed;
}

}  // namespace

extern "C" {

stereo_status stereo_batch_configure(stereo_context* ctx, stereo_driver** out, int32_t npairs,
                                     const stereo_image_format* fmts, const int32_t* num_disparities,
                                     const stereo_matcher_params* params, const stereo_batch_tuning* tuning)
{
    if (!ctx || !out || !fmts || !num_disparities || !params || npairs < 1 || npairs > 4096)
        return STEREO_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (params->algorithm != STEREO_ALGO_SAD && params->algorithm != STEREO_ALGO_CENSUS)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "batch: SAD or census matcher only");
    stereo_batch_tuning tn{64, 512, 0, 1};
    if (tuning) {
        tn = *tuning;
        if (tn.chunk == 0) tn.chunk = 64;
        if (tn.grid == 0) tn.grid = 512;
    }
    if (tn.chunk < kMinChunk || tn.chunk > 256 || tn.chunk % 8 || tn.grid < 1 || tn.grid > 65535 || tn.order < 0 ||
        tn.order > 1)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "batch tuning: chunk %d (32..256, multiple of 8), grid %d, order %d",
                    tn.chunk, tn.grid, tn.order);
    char msg[256];
    for (int i = 0; i < npairs; ++i) {
        stereo_matcher_params p = *params;
        p.num_disparities = num_disparities[i];
        stereo_status st = stereo_matcher_validate(&fmts[i], &p, msg, sizeof(msg));
        if (st) return fail(ctx, st, "pair %d: %s", i, msg);
    }
    stereo_status st = set_device(ctx);
    if (st) return st;
    stereo_driver* d = new stereo_driver(stereo_driver::BATCH);
    d->ctx = ctx;
    d->mp = *params;
    d->fmt = fmts[0];
    d->name = params->algorithm == STEREO_ALGO_CENSUS ? "Batched Block Matcher Census" : "Batched Block Matcher SAD";
    BatchExt* x = new BatchExt();
    d->ext = x;
    x->ctx = ctx;
    x->npairs = npairs;
    x->mp = *params;
    x->kind = params->algorithm == STEREO_ALGO_CENSUS ? stereo::BM_COST_CENSUS : stereo::BM_COST_SAD;
    x->elem = x->kind == stereo::BM_COST_CENSUS ? 8 : 1;
    x->tuning = tn;
    auto cleanup = [&](stereo_status s2) {
        driver_release(d);
        delete d;
        return s2;
    };
    if ((st = driver_events(d))) return cleanup(st);
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&x->cap, hipStreamNonBlocking)) || (e = hipEventCreate(&x->ev_fork)) ||
        (e = hipEventCreate(&x->ev_join)) || (e = hipEventCreate(&x->ev_done)))
        return cleanup(hip_fail(ctx, e, "hipStreamCreate"));
    d->bx = params->window_w / 2 + (x->kind == stereo::BM_COST_CENSUS ? params->census_w / 2 : 0);
    d->by = params->window_h / 2 + (x->kind == stereo::BM_COST_CENSUS ? params->census_h / 2 : 0);
    x->padL.assign(npairs, nullptr);
    x->padR.assign(npairs, nullptr);
    int tiles = 0, rows = 0;
    size_t max_items = 0;
    for (int i = 0; i < npairs; ++i) {
        const int W = fmts[i].width, H = fmts[i].height;
        const size_t cols = (size_t)stereo::BM_PAD_LEFT + W + stereo::BM_PAD_RIGHT;
        const size_t pitch = (cols * x->elem + 255) & ~(size_t)255;
        if ((e = hipMalloc((void**)&x->padL[i], pitch * H)) || (e = hipMalloc((void**)&x->padR[i], pitch * H)) ||
            (e = hipMemset(x->padL[i], 0, pitch * H)) || (e = hipMemset(x->padR[i], 0, pitch * H)))
            return cleanup(hip_fail(ctx, e, "hipMalloc"));
        stereo::BmFrame f;
        std::memset(&f, 0, sizeof(f));
        f.L.ptr = x->padL[i] + (size_t)stereo::BM_PAD_LEFT * x->elem;
        f.R.ptr = x->padR[i] + (size_t)stereo::BM_PAD_LEFT * x->elem;
        f.L.pitch = f.R.pitch = pitch;
        f.L.w = f.R.w = W;
        f.L.h = f.R.h = H;
        f.W = W;
        f.H = H;
        f.D = num_disparities[i];
        f.bx = d->bx;
        f.by = d->by;
        f.tol = params->lr_tolerance;
        f.ntx = (W + stereo::BM_TILE_W - 1) / stereo::BM_TILE_W;
        f.nty = (H + stereo::BM_TILE_H - 1) / stereo::BM_TILE_H;
        f.tile_base = tiles;
        f.row_base = rows;
        f.plane_off = x->plane_elems;
        x->plane_elems += (size_t)W * H;
        tiles += f.ntx * f.nty;
        rows += H;
        max_items += (size_t)f.ntx * f.nty * ((f.D + kMinChunk - 1) / kMinChunk);
        x->max_width = std::max(x->max_width, W);
        x->frames.push_back(f);
    }
    if (x->max_width > stereo::BM_LIST_MAX_WIDTH)
        return cleanup(fail(ctx, STEREO_ERROR_INVALID_FORMAT, "batch: width above %d", stereo::BM_LIST_MAX_WIDTH));
    x->total_rows = rows;
    x->max_items = (int)max_items;
    x->pair_chunk.assign(npairs, 0);
    if ((e = hipMalloc((void**)&x->frames_dev, npairs * sizeof(stereo::BmFrame))) ||
        (e = hipMalloc((void**)&x->tiles_dev, tiles * sizeof(stereo::BmTile))) ||
        (e = hipMalloc((void**)&x->items_dev, max_items * sizeof(stereo::BmItem))) ||
        (e = hipMalloc((void**)&x->sched_dev, max_items * sizeof(int))) ||
        (e = hipMalloc((void**)&x->nitems_dev, sizeof(int))) ||
        (e = hipMalloc((void**)&x->counter_dev, sizeof(int))) ||
        (e = hipMalloc((void**)&x->lplane, x->plane_elems * 4)) || (e = hipMalloc((void**)&x->rplane, x->plane_elems * 4)) ||
        (e = hipMalloc((void**)&x->jobs_dev, 2 * npairs * sizeof(stereo::ImageJob))) ||
        (e = hipMemset(x->lplane, 0, x->plane_elems * 4)) || (e = hipMemset(x->rplane, 0, x->plane_elems * 4)) ||
        (e = hipMemset(x->counter_dev, 0, sizeof(int))) || (e = hipDeviceSynchronize()))
        return cleanup(hip_fail(ctx, e, "hipMalloc"));
    if ((e = hipMemcpy(x->frames_dev, x->frames.data(), npairs * sizeof(stereo::BmFrame), hipMemcpyHostToDevice)))
        return cleanup(hip_fail(ctx, e, "hipMemcpy"));
    if ((st = upload_work(x))) return cleanup(st);
    *out = d;
    return STEREO_OK;
}

stereo_status stereo_batch_process(stereo_driver* d, stereo_image* disp, stereo_image* disp_lr, stereo_image* disp_rl,
                                   const stereo_image* left, const stereo_image* right, stereo_stream stream)
{
    if (!d || d->kind != stereo_driver::BATCH || !disp || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    BatchExt* x = static_cast<BatchExt*>(d->ext);
    stereo_context* ctx = d->ctx;
    stereo_status st;
    for (int i = 0; i < x->npairs; ++i) {
        const int W = x->frames[i].W, H = x->frames[i].H;
        if ((st = check_image(ctx, &left[i], W, H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
        if ((st = check_image(ctx, &right[i], W, H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
        if ((st = check_image(ctx, &disp[i], W, H, STEREO_PIXEL_LUM_I16, "disparity output"))) return st;
        if (disp_lr && disp_lr[i].data && (st = check_image(ctx, &disp_lr[i], W, H, STEREO_PIXEL_LUM_I16, "left-to-right output")))
            return st;
        if (disp_rl && disp_rl[i].data && (st = check_image(ctx, &disp_rl[i], W, H, STEREO_PIXEL_LUM_I16, "right-to-left output")))
            return st;
    }
    if ((st = set_device(ctx))) return st;
    GUARD(ctx, "stereo_batch_process");
    const hipStream_t s = (hipStream_t)stream;
    if (set_outputs(x, disp, disp_lr, disp_rl)) {
        // the device frame table is read by the previous call's launches, which may still
        // run on another stream: wait for their end before rewriting it
        if (x->done_recorded) HIP_TRY(ctx, hipEventSynchronize(x->ev_done), "hipEventSynchronize");
        HIP_TRY(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
        HIP_TRY(ctx, hipMemcpy(x->frames_dev, x->frames.data(), x->npairs * sizeof(stereo::BmFrame), hipMemcpyHostToDevice),
                "hipMemcpy");
        x->drop_graph();
    }
    if ((st = clear_if_dirty(x, s))) return st;
    if ((st = upload_jobs(x, left, right))) return st;
    if ((st = begin(d, s))) return st;
    if (!x->tuning.use_graph) {
        if ((st = enqueue(d, x, left, right, s))) return st;
        HIP_TRY(ctx, hipEventRecord(x->ev_done, s), "hipEventRecord");
        x->done_recorded = true;
        return end(d, s);
    }
    std::vector<const void*> key;
    key.reserve(2 * x->npairs);
    for (int i = 0; i < x->npairs; ++i) {
        key.push_back(left[i].data);
        key.push_back(right[i].data);
    }
    if (!x->exec || key != x->graph_key) {
        x->drop_graph();
        // capture on the driver's own stream (the default stream cannot be captured),
        // then launch the graph on the caller's stream
        HIP_TRY(ctx, hipStreamBeginCapture(x->cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const stereo_status cst = enqueue(d, x, left, right, x->cap);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(x->cap, &g);
        if (cst) {
            if (g) (void)hipGraphDestroy(g);
            return cst;
        }
        HIP_TRY(ctx, ee, "hipStreamEndCapture");
        x->graph = g;
        HIP_TRY(ctx, hipGraphInstantiate(&x->exec, x->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        x->graph_key = key;
    }
    HIP_TRY(ctx, hipGraphLaunch(x->exec, s), "hipGraphLaunch");
    HIP_TRY(ctx, hipEventRecord(x->ev_done, s), "hipEventRecord");
    x->done_recorded = true;
    return end(d, s);
}

stereo_status stereo_batch_get_tuning(const stereo_driver* d, stereo_batch_tuning* t)
{
    if (!d || d->kind != stereo_driver::BATCH || !t) return STEREO_ERROR_INVALID_ARGUMENT;
    *t = static_cast<const BatchExt*>(d->ext)->tuning;
    return STEREO_OK;
}

stereo_status stereo_batch_autotune(stereo_driver* d, const stereo_image* left, const stereo_image* right,
                                    stereo_batch_tuning* best)
{
    if (!d || d->kind != stereo_driver::BATCH || !left || !right) return STEREO_ERROR_INVALID_ARGUMENT;
    BatchExt* x = static_cast<BatchExt*>(d->ext);
    stereo_context* ctx = d->ctx;
    stereo_status st;
    for (int i = 0; i < x->npairs; ++i) {
        if ((st = check_image(ctx, &left[i], x->frames[i].W, x->frames[i].H, STEREO_PIXEL_LUM_U8, "left image"))) return st;
        if ((st = check_image(ctx, &right[i], x->frames[i].W, x->frames[i].H, STEREO_PIXEL_LUM_U8, "right image"))) return st;
    }
    if ((st = set_device(ctx))) return st;
    // the timed runs need no outputs beyond what the frame table holds; point every
    // frame at a scratch map so that no caller image is written
    std::vector<stereo::BmFrame> saved = x->frames;
    int16_t* scratch = nullptr;
    size_t spitch = ((size_t)x->max_width * 2 + 255) & ~(size_t)255;
    int maxh = 0;
    for (const stereo::BmFrame& f : x->frames) maxh = std::max(maxh, f.H);
    HIP_TRY(ctx, hipDeviceSynchronize(), "hipDeviceSynchronize");  // earlier calls may still read the frame table
    HIP_TRY(ctx, hipMalloc((void**)&scratch, spitch * maxh), "hipMalloc");
    for (stereo::BmFrame& f : x->frames) {
        f.disp.ptr = reinterpret_cast<uint8_t*>(scratch);
        f.disp.pitch = spitch;
        f.disp.w = f.W;
        f.disp.h = f.H;
        f.disp_l = null_view();
        f.disp_r = null_view();
    }
    HIP_TRY(ctx, hipMemcpy(x->frames_dev, x->frames.data(), x->npairs * sizeof(stereo::BmFrame), hipMemcpyHostToDevice),
            "hipMemcpy");
    const stereo_batch_tuning keep = x->tuning;
    stereo_batch_tuning win = keep;
    float win_ms = 1e30f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(ctx, hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(ctx, hipEventCreate(&e1), "hipEventCreate");
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    const int chunks[] = {32, 64, 128, 256};
    const int grids[] = {cus, 2 * cus, 3 * cus};
    const std::vector<int> keep_chunks = x->pair_chunk;
    std::fill(x->pair_chunk.begin(), x->pair_chunk.end(), 0);
    if ((st = upload_jobs(x, left, right))) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(scratch);
        x->frames = saved;
        (void)hipMemcpy(x->frames_dev, x->frames.data(), x->npairs * sizeof(stereo::BmFrame), hipMemcpyHostToDevice);
        return st;
    }
    st = STEREO_OK;
    // time the current work list: warm-up, then three runs
    auto time_it = [&](float* ms) -> stereo_status {
        stereo_status s2 = upload_work(x);
        if (s2) return s2;
        if ((s2 = clear_if_dirty(x, nullptr))) return s2;
        if ((s2 = enqueue(d, x, left, right, nullptr))) return s2;
        hipError_t e = hipEventRecord(e0, nullptr);
        for (int r = 0; r < 3 && e == hipSuccess && !s2; ++r) s2 = enqueue(d, x, left, right, nullptr);
        if (s2) return s2;
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
        return e == hipSuccess ? STEREO_OK : hip_fail(ctx, e, "autotune timing");
    };
    // one item length for every pair, with the grid size and the item order
    for (int c : chunks)
        for (int g : grids)
            for (int o = 0; o < 2 && !st; ++o) {
                x->tuning.chunk = c;
                x->tuning.grid = g;
                x->tuning.order = o;
                x->tuning.use_graph = 0;
                float ms = 0.f;
                if ((st = time_it(&ms))) break;
                if (ms < win_ms) {
                    win_ms = ms;
                    win = x->tuning;
                }
            }
    // then the item length of each class of pairs (disparity count x tile count), the
    // classes with the most evaluations first
    std::vector<int> win_chunks(x->npairs, 0);
    if (!st) {
        x->tuning = win;
        x->tuning.use_graph = 0;
        std::vector<std::pair<long long, std::pair<int, int>>> classes;  // (evaluations, (D, size class))
        for (int f = 0; f < x->npairs; ++f) {
            const stereo::BmFrame& fr = x->frames[f];
            const std::pair<int, int> key(fr.D, size_class(fr));
            auto it = std::find_if(classes.begin(), classes.end(), [&](const auto& c) { return c.second == key; });
            if (it == classes.end()) classes.push_back({0, key}), it = classes.end() - 1;
            it->first += (long long)fr.W * fr.H * fr.D;
        }
        std::sort(classes.begin(), classes.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        for (const auto& cl : classes) {
            int best_c = 0;
            float best_ms = win_ms;
            for (int c : chunks) {
                if (c == win.chunk || c > std::max(32, cl.second.first)) continue;
                for (int f = 0; f < x->npairs; ++f) {
                    const stereo::BmFrame& fr = x->frames[f];
                    if (fr.D == cl.second.first && size_class(fr) == cl.second.second) x->pair_chunk[f] = c;
                }
                float ms = 0.f;
                if ((st = time_it(&ms))) break;
                if (ms < best_ms) {
                    best_ms = ms;
                    best_c = c;
                }
                for (int f = 0; f < x->npairs; ++f) x->pair_chunk[f] = win_chunks[f];
            }
            if (st) break;
            if (best_c) {
                for (int f = 0; f < x->npairs; ++f) {
                    const stereo::BmFrame& fr = x->frames[f];
                    if (fr.D == cl.second.first && size_class(fr) == cl.second.second) win_chunks[f] = best_c;
                }
                win_ms = best_ms;
            }
            x->pair_chunk = win_chunks;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipDeviceSynchronize();
    (void)hipFree(scratch);
    x->frames = saved;
    x->tuning = st ? keep : win;
    x->tuning.use_graph = keep.use_graph;
    x->pair_chunk = st ? keep_chunks : win_chunks;
    HIP_TRY(ctx, hipMemcpy(x->frames_dev, x->frames.data(), x->npairs * sizeof(stereo::BmFrame), hipMemcpyHostToDevice),
            "hipMemcpy");
    const stereo_status st2 = upload_work(x);
    if (st) return st;
    if (st2) return st2;
    if (best) *best = x->tuning;
    return STEREO_OK;
}

stereo_status stereo_batch_tuning_table(const stereo_driver* d, char* buf, size_t n, size_t* needed)
{
    if (!d || d->kind != stereo_driver::BATCH) return STEREO_ERROR_INVALID_ARGUMENT;
    const BatchExt* x = static_cast<const BatchExt*>(d->ext);
    std::string t = "{\"matcher\": \"";
    t += x->kind == stereo::BM_COST_CENSUS ? "census" : "sad";
    char tmp[256];
    std::snprintf(tmp, sizeof(tmp),
                  "\", \"window\": %d, \"grid\": %d, \"order\": %d, \"use_graph\": %d, \"chunk\": %d, \"classes\": [",
                  x->mp.window_w, x->tuning.grid, x->tuning.order, x->tuning.use_graph, x->tuning.chunk);
    t += tmp;
    std::vector<std::pair<int, int>> seen;
    for (int f = 0; f < x->npairs; ++f) {
        const std::pair<int, int> key(x->frames[f].D, size_class(x->frames[f]));
        if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
        seen.push_back(key);
        std::snprintf(tmp, sizeof(tmp), "%s{\"disparities\": %d, \"tiles\": \"%s\", \"tile\": \"%dx%d\", \"chunk\": %d}",
                      seen.size() > 1 ? ", " : "", key.first, kSizeClass[key.second], stereo::BM_TILE_W, stereo::BM_TILE_H,
                      chunk_of(x, f));
        t += tmp;
    }
    t += "]}";
    if (needed) *needed = t.size() + 1;
    if (buf && n) {
        const size_t k = std::min(n - 1, t.size());
        std::memcpy(buf, t.data(), k);
        buf[k] = 0;
        if (k < t.size()) return fail(d->ctx, STEREO_ERROR_INVALID_ARGUMENT, "tuning table: buffer of %zu bytes, %zu needed",
                                      n, t.size() + 1);
    }
    return STEREO_OK;
}

namespace {

// The integer after "key": in [from, to); false when absent.
bool json_int(const char* from, const char* to, const char* key, int* v)
{
    const std::string k = std::string("\"") + key + "\":";
    const char* p = std::search(from, to, k.begin(), k.end());
    if (p == to) return false;
    return std::sscanf(p + k.size(), "%d", v) == 1;
}

bool json_word(const char* from, const char* to, const char* key, char* out, size_t n)
{
    const std::string k = std::string("\"") + key + "\":";
    const char* p = std::search(from, to, k.begin(), k.end());
    if (p == to) return false;
    p += k.size();
    while (p < to && *p == ' ') ++p;
    if (p >= to || *p != '"') return false;
    size_t i = 0;
    for (++p; p < to && *p != '"' && i + 1 < n; ++p) out[i++] = *p;
    out[i] = 0;
    return p < to && *p == '"';
}

}  // namespace

stereo_status stereo_batch_apply_tuning_table(stereo_driver* d, const char* json)
{
    if (!d || d->kind != stereo_driver::BATCH || !json) return STEREO_ERROR_INVALID_ARGUMENT;
    BatchExt* x = static_cast<BatchExt*>(d->ext);
    stereo_context* ctx = d->ctx;
    // the top-level keys are read from the text without the class list
    const char* cls = std::strstr(json, "\"classes\"");
    const char* cls_end = cls ? std::strchr(cls, ']') : nullptr;
    if (cls && !cls_end) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "tuning table: unterminated class list");
    const std::string top = cls ? std::string(json, cls) + std::string(cls_end + 1) : std::string(json);
    const char* t0 = top.c_str();
    const char* t1 = t0 + top.size();
    stereo_batch_tuning t = x->tuning;
    int v;
    if (json_int(t0, t1, "grid", &v)) t.grid = v;
    if (json_int(t0, t1, "order", &v)) t.order = v;
    if (json_int(t0, t1, "use_graph", &v)) t.use_graph = v;
    if (json_int(t0, t1, "chunk", &v)) t.chunk = v;
    auto chunk_ok = [](int c) { return c >= kMinChunk && c <= 256 && c % 8 == 0; };
    if (!chunk_ok(t.chunk) || t.grid < 1 || t.grid > 65535 || t.order < 0 || t.order > 1 || t.use_graph < 0 ||
        t.use_graph > 1)
        return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "tuning table: grid %d, order %d, use_graph %d, chunk %d", t.grid,
                    t.order, t.use_graph, t.chunk);
    std::vector<int> chunks(x->npairs, 0);
    for (const char* p = cls ? std::strchr(cls, '{') : nullptr; p && p < cls_end; p = std::strchr(p + 1, '{')) {
        const char* q = std::strchr(p, '}');
        if (!q || q > cls_end) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "tuning table: unterminated class entry");
        int D, c;
        char size[16];
        if (!json_int(p, q, "disparities", &D) || !json_int(p, q, "chunk", &c) || !json_word(p, q, "tiles", size, 16))
            return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "tuning table: a class needs disparities, tiles and chunk");
        if (!chunk_ok(c)) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "tuning table: chunk %d", c);
        int sc = -1;
        for (int i = 0; i < 3; ++i)
            if (!std::strcmp(size, kSizeClass[i])) sc = i;
        if (sc < 0) return fail(ctx, STEREO_ERROR_INVALID_CONFIG, "tuning table: tiles \"%s\"", size);
        for (int f = 0; f < x->npairs; ++f)
            if (x->frames[f].D == D && size_class(x->frames[f]) == sc) chunks[f] = c;
    }
    stereo_status st = set_device(ctx);
    if (st) return st;
    if (x->done_recorded) HIP_TRY(ctx, hipEventSynchronize(x->ev_done), "hipEventSynchronize");
    const stereo_batch_tuning keep = x->tuning;
    const std::vector<int> keep_chunks = x->pair_chunk;
    x->tuning = t;
    x->pair_chunk = chunks;
    st = upload_work(x);
    if (st) {
        x->tuning = keep;
        x->pair_chunk = keep_chunks;
        (void)upload_work(x);
        return st;
    }
    return STEREO_OK;
}

}  // extern "C"
