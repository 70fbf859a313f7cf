// xg_serve.hip -- serving: continuous batching of frames from many streams.
//
// Each admitted stream owns two padded pictures in HBM (the current frame and
// its reference, which swap roles after every frame) and one prediction
// picture.  Submitted frames are copied to pinned host memory and queued per
// stream.  A dispatcher thread keeps up to two batches in flight, one per
// device session (each with its own HIP stream): whenever a session is free it
// takes the oldest pending frame of every stream that has no frame in flight,
// uploads the frames, pads them on the GPU, and enqueues one integer-search
// launch and one refinement+MC launch over all their blocks.  Frames arriving
// while a batch runs join the next one; with maxDelayUs > 0 a free session
// also waits that long after the oldest waiting frame arrived.  Results are copied back to the
// caller's buffers when the batch's event completes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string.h>
#include <thread>
#include <vector>

#include "xg_runtime.hpp"

namespace xg {

static int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Edge replication of the padding ring of one plane (data points at sample (0,0)).
__global__ void k_pad_plane(uint16_t *data, intptr_t stride, int w, int h, int pad)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x) - pad;
    const int y = (int)blockIdx.y - pad;
    if (x >= w + pad) return;
    if (x >= 0 && x < w && y >= 0 && y < h) return;
    const int sx = min(max(x, 0), w - 1), sy = min(max(y, 0), h - 1);
    data[(intptr_t)y * stride + x] = data[(intptr_t)sy * stride + sx];
}

namespace {

constexpr int kMaxStreams = kSessionPicSlots / 2;
constexpr int kSessions = 2;

struct Request {
    int stream = 0;
    uint64_t tag = 0;
    int64_t submitNs = 0;
    std::vector<xg_block> blocks;
    uint16_t *pinned = nullptr;        // Y | Cb | Cr packed; reused for the prediction read-back
    size_t pinnedBytes = 0;
    xg_block_result *results = nullptr;
    xg_pixel *pred[3] = {nullptr, nullptr, nullptr};
    intptr_t predStride[3] = {0, 0, 0};
};

struct Stream {
    bool open = false, closing = false, inFlight = false, hasRef = false;
    int w = 0, h = 0, cur = 0;
    void *mem = nullptr;
    size_t bytes = 0;
    xg_picture pic[2];
    xg_picture pred;
    std::deque<Request *> pending;
    int64_t frames = 0;
};

struct Batch {
    bool active = false;
    std::vector<Request *> reqs;
    std::vector<int> off;
    hipEvent_t done = nullptr;
    int64_t startNs = 0;
    xg_block_result *pinnedResults = nullptr;
};

}  // namespace
}  // namespace xg

using namespace xg;

struct xg_server {
    xg_server_cfg cfg;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::thread thread;
    bool stop = false;
    xg_session *sess[kSessions] = {nullptr, nullptr};
    Batch batch[kSessions];
    Stream streams[kMaxStreams];
    std::deque<xg_frame_result> completed;
    std::multimap<size_t, uint16_t *> pinnedPool;
    uint64_t bytesSessions = 0, bytesStreams = 0;
    int64_t framesDone = 0, framesDropped = 0, batches = 0;

    uint16_t *pinned_get(size_t bytes)
    {
        auto it = pinnedPool.find(bytes);
        if (it != pinnedPool.end()) {
            uint16_t *p = it->second;
            pinnedPool.erase(it);
            return p;
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        return (uint16_t *)p;
    }
    void pinned_put(uint16_t *p, size_t bytes) { pinnedPool.emplace(bytes, p); }

    void finish_request(Request *r, int32_t status, int batchFrames, int64_t startNs)
    {
        xg_frame_result fr;
        fr.stream = r->stream;
        fr.status = status;
        fr.nblocks = (int32_t)r->blocks.size();
        fr.batchFrames = batchFrames;
        fr.tag = r->tag;
        fr.submitNs = r->submitNs;
        fr.startNs = startNs;
        fr.doneNs = now_ns();
        completed.push_back(fr);
        if (status == XG_OK) framesDone++;
        else framesDropped++;
        if (r->pinned) pinned_put(r->pinned, r->pinnedBytes);
        delete r;
    }

    // Finalises a completed batch: results to the callers, reference swap.
    void finalize(int si)
    {
        Batch &b = batch[si];
        const int nf = (int)b.reqs.size();
        for (int i = 0; i < nf; i++) {
            Request *r = b.reqs[i];
            Stream &st = streams[r->stream];
            memcpy(r->results, b.pinnedResults + b.off[i], sizeof(xg_block_result) * r->blocks.size());
            if (r->pred[0]) {
                const int cw = st.w / 2, ch = st.h / 2;
                unpack2d(r->pred[0], r->predStride[0], r->pinned, st.w, st.h, 2);
                unpack2d(r->pred[1], r->predStride[1], r->pinned + (size_t)st.w * st.h, cw, ch, 2);
                unpack2d(r->pred[2], r->predStride[2], r->pinned + (size_t)st.w * st.h + (size_t)cw * ch, cw, ch, 2);
            }
            st.cur ^= 1;           // this frame's picture becomes the next frame's reference
            st.hasRef = true;
            st.inFlight = false;
            st.frames++;
            finish_request(r, XG_OK, nf, b.startNs);
        }
        b.reqs.clear();
        b.off.clear();
        b.active = false;
        cvDone.notify_all();
    }

    // Assembles and enqueues one batch on session si; returns false if nothing to do.
    bool launch(int si)
    {
        Batch &b = batch[si];
        std::vector<int> ids;
        for (int i = 0; i < kMaxStreams; i++)
            if (streams[i].open && !streams[i].closing && !streams[i].inFlight && !streams[i].pending.empty())
                ids.push_back(i);
        if (ids.empty()) return false;
        std::sort(ids.begin(), ids.end(), [&](int x, int y) {
            return streams[x].pending.front()->submitNs < streams[y].pending.front()->submitNs;
        });
        std::vector<xg_block> blocks;
        b.reqs.clear();
        b.off.clear();
        for (int id : ids) {
            Request *r = streams[id].pending.front();
            if (!b.reqs.empty() && (int)(blocks.size() + r->blocks.size()) > cfg.maxBatchBlocks) break;
            streams[id].pending.pop_front();
            Stream &st = streams[id];
            const int curSlot = 2 * id + st.cur, refSlot = st.hasRef ? 2 * id + (st.cur ^ 1) : curSlot;
            b.off.push_back((int)blocks.size());
            for (xg_block B : r->blocks) {
                B.pic = (uint8_t)curSlot;
                B.ref = (uint8_t)refSlot;
                B.mvx = B.mvy = 0;
                blocks.push_back(B);
            }
            b.reqs.push_back(r);
            st.inFlight = true;
        }
        xg_session *s = sess[si];
        hipStream_t hs = (hipStream_t)xg_session_stream(s);
        b.startNs = now_ns();
        int32_t rc = xg_session_set_blocks(s, blocks.data(), (int32_t)blocks.size());
        for (size_t i = 0; i < b.reqs.size() && rc == XG_OK; i++) {
            Request *r = b.reqs[i];
            Stream &st = streams[r->stream];
            const xg_picture &P = st.pic[st.cur];
            const uint16_t *src = r->pinned;
            for (int p = 0; p < 3 && rc == XG_OK; p++) {
                const xg_plane &pl = P.plane[p];
                if (hipMemcpy2DAsync(pl.data, pl.stride * 2, src, (size_t)pl.width * 2, (size_t)pl.width * 2,
                                     pl.height, hipMemcpyHostToDevice, hs) != hipSuccess)
                    rc = XG_ERR_HIP;
                dim3 grid((pl.width + 2 * pl.pad + 255) / 256, pl.height + 2 * pl.pad);
                k_pad_plane<<<grid, 256, 0, hs>>>(pl.data, pl.stride, pl.width, pl.height, pl.pad);
                src += (size_t)pl.width * pl.height;
            }
        }
        if (rc == XG_OK) rc = xg_search_blocks(s, cfg.searchRange, cfg.searchLambda, cfg.refine.ctuSize);
        if (rc == XG_OK) rc = xg_refine_frac(s, &cfg.refine, nullptr, XG_MODE_BATCHED);
        if (rc == XG_OK) {
            xg_block_result *dBest = nullptr;
            xg_session_outputs(s, nullptr, &dBest);
            if (hipMemcpyAsync(b.pinnedResults, dBest, sizeof(xg_block_result) * blocks.size(),
                               hipMemcpyDeviceToHost, hs) != hipSuccess)
                rc = XG_ERR_HIP;
        }
        for (size_t i = 0; i < b.reqs.size() && rc == XG_OK; i++) {
            Request *r = b.reqs[i];
            if (!r->pred[0]) continue;
            Stream &st = streams[r->stream];
            uint16_t *dst = r->pinned;
            for (int p = 0; p < 3; p++) {
                const xg_plane &pl = st.pred.plane[p];
                if (hipMemcpy2DAsync(dst, (size_t)pl.width * 2, pl.data, pl.stride * 2, (size_t)pl.width * 2, pl.height,
                                     hipMemcpyDeviceToHost, hs) != hipSuccess)
                    rc = XG_ERR_HIP;
                dst += (size_t)pl.width * pl.height;
            }
        }
        if (rc == XG_OK && hipEventRecord(b.done, hs) != hipSuccess) rc = XG_ERR_HIP;
        if (rc != XG_OK) {
            log_msg(0, "serve: batch of %zu frames failed (%d)", b.reqs.size(), rc);
            (void)hipStreamSynchronize(hs);
            for (Request *r : b.reqs) {
                streams[r->stream].inFlight = false;
                finish_request(r, rc, (int)b.reqs.size(), b.startNs);
            }
            b.reqs.clear();
            cvDone.notify_all();
            return true;
        }
        b.active = true;
        batches++;
        return true;
    }

    // Nanoseconds a free session should still wait to gather frames: > 0 while
    // the oldest launchable frame is younger than maxDelayUs and the launchable
    // frames do not fill a batch.
    int64_t gather_wait_ns() const
    {
        if (cfg.maxDelayUs <= 0) return 0;
        int64_t oldest = INT64_MAX, nblk = 0;
        for (int i = 0; i < kMaxStreams; i++) {
            const Stream &st = streams[i];
            if (!st.open || st.closing || st.inFlight || st.pending.empty()) continue;
            oldest = std::min(oldest, st.pending.front()->submitNs);
            nblk += (int64_t)st.pending.front()->blocks.size();
        }
        if (oldest == INT64_MAX || nblk >= cfg.maxBatchBlocks) return 0;
        return std::max<int64_t>(0, oldest + (int64_t)cfg.maxDelayUs * 1000 - now_ns());
    }

    void loop()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            for (int i = 0; i < kSessions; i++)
                if (batch[i].active && hipEventQuery(batch[i].done) == hipSuccess) finalize(i);
            const int64_t gather = stop ? 0 : gather_wait_ns();
            if (gather > 0) {
                cvWork.wait_for(lk, std::chrono::nanoseconds(std::min<int64_t>(gather, 100000)));
                continue;
            }
            bool launched = false;
            for (int i = 0; i < kSessions && !stop; i++)
                if (!batch[i].active && launch(i)) launched = true;
            if (launched) continue;
            int active = 0, oldest = -1;
            for (int i = 0; i < kSessions; i++)
                if (batch[i].active) {
                    active++;
                    if (oldest < 0 || batch[i].startNs < batch[oldest].startNs) oldest = i;
                }
            if (stop && active == 0) break;
            if (active == kSessions) {
                // nothing can launch until a batch completes
                hipEvent_t e = batch[oldest].done;
                lk.unlock();
                (void)hipEventSynchronize(e);
                lk.lock();
            } else if (active > 0) {
                cvWork.wait_for(lk, std::chrono::microseconds(100));
            } else {
                cvWork.wait(lk);
            }
        }
    }
};

extern "C" {

int32_t xg_server_create(xg_server **out, const xg_server_cfg *cfg)
{
    if (!out || !cfg || cfg->maxStreams < 1 || cfg->maxStreams > kMaxStreams || cfg->maxBatchBlocks < 1 ||
        cfg->searchRange < 0 || cfg->searchRange > 32 || cfg->maxDelayUs < 0 || cfg->maxDelayUs > 1000000 || cfg->refine.ncand < 1 || cfg->refine.ncand > XG_MAX_CAND) {
        set_error(XG_ERR_ARG, "xg_server_create: bad configuration");
        return XG_ERR_ARG;
    }
    *out = nullptr;
    if (!device_init()) return XG_ERR_HIP;
    xg_server *srv = new xg_server();
    srv->cfg = *cfg;
    // per session: blocks, tasks, costs, results and completion words per block
    srv->bytesSessions = (uint64_t)kSessions * (uint64_t)cfg->maxBatchBlocks *
                         (sizeof(xg_block) + 2 * 8 + XG_MAX_CAND * 4 + sizeof(xg_block_result) + 4 + 8);
    if (srv->bytesSessions > cfg->hbmBudget) {
        set_error(XG_ERR_BUDGET, "xg_server_create: sessions alone need %llu bytes",
                  (unsigned long long)srv->bytesSessions);
        delete srv;
        return XG_ERR_BUDGET;
    }
    for (int i = 0; i < kSessions; i++) {
        if (xg_session_create(&srv->sess[i], cfg->maxBatchBlocks) != XG_OK ||
            hipEventCreateWithFlags(&srv->batch[i].done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&srv->batch[i].pinnedResults, sizeof(xg_block_result) * cfg->maxBatchBlocks,
                          hipHostMallocDefault) != hipSuccess) {
            set_error(XG_ERR_NOMEM, "xg_server_create: session allocation failed");
            xg_server_destroy(srv);
            return XG_ERR_NOMEM;
        }
    }
    srv->thread = std::thread([srv] { srv->loop(); });
    *out = srv;
    return XG_OK;
}

int32_t xg_server_destroy(xg_server *srv)
{
    if (!srv) return XG_OK;
    if (srv->thread.joinable()) {
        {
            std::lock_guard<std::mutex> g(srv->mu);
            srv->stop = true;
        }
        srv->cvWork.notify_all();
        srv->thread.join();
    }
    for (int i = 0; i < kMaxStreams; i++) {
        Stream &st = srv->streams[i];
        for (Request *r : st.pending) {
            if (r->pinned) (void)hipHostFree(r->pinned);
            delete r;
        }
        st.pending.clear();
        if (st.mem) (void)hipFree(st.mem);
    }
    for (int i = 0; i < kSessions; i++) {
        if (srv->sess[i]) xg_session_destroy(srv->sess[i]);
        if (srv->batch[i].done) (void)hipEventDestroy(srv->batch[i].done);
        if (srv->batch[i].pinnedResults) (void)hipHostFree(srv->batch[i].pinnedResults);
    }
    for (auto &kv : srv->pinnedPool) (void)hipHostFree(kv.second);
    delete srv;
    return XG_OK;
}

int32_t xg_server_open_stream(xg_server *srv, int32_t width, int32_t height, int32_t *streamId)
{
    if (!srv || !streamId || width < 16 || height < 16 || (width & 7) || (height & 7) || width > 8192 ||
        height > 8192) {
        set_error(XG_ERR_ARG, "xg_server_open_stream: bad arguments (%dx%d)", width, height);
        return XG_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(srv->mu);
    int id = -1, open = 0;
    for (int i = 0; i < kMaxStreams; i++) {
        if (srv->streams[i].open) open++;
        else if (id < 0) id = i;
    }
    if (id < 0 || open >= srv->cfg.maxStreams) {
        set_error(XG_ERR_BUDGET, "xg_server_open_stream: all %d stream slots in use", srv->cfg.maxStreams);
        return XG_ERR_BUDGET;
    }
    const int pad = srv->cfg.refine.ctuSize + 16, cpad = pad / 2;
    const intptr_t ys = width + 2 * pad, cs = width / 2 + 2 * cpad;
    const size_t ypic = (size_t)ys * (height + 2 * pad), cpic = (size_t)cs * (height / 2 + 2 * cpad);
    const size_t picBytes = align_up((ypic + 2 * cpic) * 2, 4096);
    const size_t predBytes = align_up((size_t)width * height * 3, 4096);     // 1.5 samples x 2 bytes
    const size_t bytes = 2 * picBytes + predBytes;
    if (srv->bytesSessions + srv->bytesStreams + bytes > srv->cfg.hbmBudget) {
        set_error(XG_ERR_BUDGET, "xg_server_open_stream: %zu bytes do not fit the HBM budget", bytes);
        return XG_ERR_BUDGET;
    }
    Stream &st = srv->streams[id];
    if (!XG_HIP(hipMalloc(&st.mem, bytes))) return XG_ERR_NOMEM;
    char *base = (char *)st.mem;
    for (int k = 0; k < 2; k++) {
        uint16_t *p = (uint16_t *)(base + k * picBytes);
        st.pic[k].plane[0] = xg_plane{p + pad * ys + pad, ys, width, height, pad};
        uint16_t *c1 = p + ypic, *c2 = c1 + cpic;
        st.pic[k].plane[1] = xg_plane{c1 + cpad * cs + cpad, cs, width / 2, height / 2, cpad};
        st.pic[k].plane[2] = xg_plane{c2 + cpad * cs + cpad, cs, width / 2, height / 2, cpad};
    }
    uint16_t *pr = (uint16_t *)(base + 2 * picBytes);
    st.pred.plane[0] = xg_plane{pr, width, width, height, 0};
    st.pred.plane[1] = xg_plane{pr + (size_t)width * height, width / 2, width / 2, height / 2, 0};
    st.pred.plane[2] = xg_plane{pr + (size_t)width * height * 5 / 4, width / 2, width / 2, height / 2, 0};
    // the dispatcher holds no session call while we hold the lock
    for (int si = 0; si < kSessions; si++)
        for (int k = 0; k < 2; k++) {
            if (xg_session_set_picture(srv->sess[si], 2 * id + k, &st.pic[k]) != XG_OK ||
                xg_session_set_pred(srv->sess[si], 2 * id + k, &st.pred) != XG_OK) {
                (void)hipFree(st.mem);
                st.mem = nullptr;
                return XG_ERR_HIP;
            }
        }
    st.open = true;
    st.closing = st.inFlight = st.hasRef = false;
    st.w = width;
    st.h = height;
    st.cur = 0;
    st.bytes = bytes;
    st.frames = 0;
    srv->bytesStreams += bytes;
    *streamId = id;
    return XG_OK;
}

int32_t xg_server_close_stream(xg_server *srv, int32_t streamId)
{
    if (!srv || streamId < 0 || streamId >= kMaxStreams) return XG_ERR_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    Stream &st = srv->streams[streamId];
    if (!st.open) {
        set_error(XG_ERR_STATE, "xg_server_close_stream: stream %d is not open", streamId);
        return XG_ERR_STATE;
    }
    st.closing = true;
    while (!st.pending.empty()) {
        Request *r = st.pending.front();
        st.pending.pop_front();
        srv->finish_request(r, XG_ERR_STATE, 0, 0);
    }
    srv->cvDone.notify_all();
    srv->cvDone.wait(lk, [&] { return !st.inFlight; });
    (void)hipFree(st.mem);
    st.mem = nullptr;
    srv->bytesStreams -= st.bytes;
    st.open = st.closing = false;
    return XG_OK;
}

int32_t xg_server_submit(xg_server *srv, int32_t streamId, const xg_pixel *const planes[3], const intptr_t strides[3],
                         const xg_block *blocks, int32_t nblocks, xg_block_result *results, xg_pixel *const pred[3],
                         const intptr_t predStride[3], uint64_t tag)
{
    if (!srv || streamId < 0 || streamId >= kMaxStreams || !planes || !strides || !blocks || nblocks < 1 ||
        nblocks > srv->cfg.maxBatchBlocks || !results) {
        set_error(XG_ERR_ARG, "xg_server_submit: bad arguments");
        return XG_ERR_ARG;
    }
    int w, h;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        const Stream &st = srv->streams[streamId];
        if (!st.open || st.closing) {
            set_error(XG_ERR_STATE, "xg_server_submit: stream %d is not open", streamId);
            return XG_ERR_STATE;
        }
        w = st.w;
        h = st.h;
    }
    for (int i = 0; i < nblocks; i++) {   // malformed descriptors are rejected before anything is queued
        const xg_block &B = blocks[i];
        const bool okw = B.w >= 4 && B.w <= 64 && !(B.w & (B.w - 1)), okh = B.h >= 4 && B.h <= 64 && !(B.h & (B.h - 1));
        if (!okw || !okh || B.x < 0 || B.y < 0 || (B.x & 3) || (B.y & 3) || B.x + B.w > w || B.y + B.h > h) {
            set_error(XG_ERR_ARG, "xg_server_submit: block %d (%d,%d %dx%d) is malformed", i, B.x, B.y, B.w, B.h);
            return XG_ERR_ARG;
        }
    }
    Request *r = new Request();
    r->stream = streamId;
    r->tag = tag;
    r->blocks.assign(blocks, blocks + nblocks);
    r->results = results;
    if (pred)
        for (int p = 0; p < 3; p++) {
            r->pred[p] = pred[p];
            r->predStride[p] = predStride ? predStride[p] : (p ? w / 2 : w);
        }
    r->pinnedBytes = (size_t)w * h * 3;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        r->pinned = srv->pinned_get(r->pinnedBytes);
    }
    if (!r->pinned) {
        delete r;
        set_error(XG_ERR_NOMEM, "xg_server_submit: pinned staging allocation failed");
        return XG_ERR_NOMEM;
    }
    pack2d(r->pinned, planes[0], strides[0], w, h, 2);
    pack2d(r->pinned + (size_t)w * h, planes[1], strides[1], w / 2, h / 2, 2);
    pack2d(r->pinned + (size_t)w * h * 5 / 4, planes[2], strides[2], w / 2, h / 2, 2);
    {
        std::lock_guard<std::mutex> g(srv->mu);
        Stream &st = srv->streams[streamId];
        if (!st.open || st.closing) {
            srv->pinned_put(r->pinned, r->pinnedBytes);
            delete r;
            set_error(XG_ERR_STATE, "xg_server_submit: stream %d closed", streamId);
            return XG_ERR_STATE;
        }
        r->submitNs = now_ns();
        st.pending.push_back(r);
    }
    srv->cvWork.notify_one();
    return XG_OK;
}

int32_t xg_server_poll(xg_server *srv, xg_frame_result *out, int32_t cap, int32_t timeoutMs)
{
    if (!srv || !out || cap < 1) return -XG_ERR_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    if (srv->completed.empty() && timeoutMs > 0)
        srv->cvDone.wait_for(lk, std::chrono::milliseconds(timeoutMs), [&] { return !srv->completed.empty(); });
    int n = 0;
    while (n < cap && !srv->completed.empty()) {
        out[n++] = srv->completed.front();
        srv->completed.pop_front();
    }
    return n;
}

int32_t xg_server_stats_get(xg_server *srv, xg_server_stats *st)
{
    if (!srv || !st) return XG_ERR_ARG;
    std::lock_guard<std::mutex> g(srv->mu);
    st->framesDone = srv->framesDone;
    st->framesDropped = srv->framesDropped;
    st->batches = srv->batches;
    st->bytesInUse = srv->bytesSessions + srv->bytesStreams;
    st->bytesBudget = srv->cfg.hbmBudget;
    st->streamsOpen = 0;
    for (int i = 0; i < kMaxStreams; i++) st->streamsOpen += srv->streams[i].open ? 1 : 0;
    return XG_OK;
}

}  // extern "C"
