// xg_serve.hip -- serving: continuous batching of frames from many streams.
//
// Every admitted stream owns one HBM allocation that stays resident for its
// life: three padded pictures used in rotation (frame k is read into picture
// k mod 3 and refers to picture (k - 1) mod 3, so the next frame's upload
// never touches a picture in use), the prediction picture, a 4x4-unit MV field (the
// final MVs of its last frame, the source of the next frame's predictors),
// its block partition with the persistent kernel's task assignment, and the
// per-block outputs.  A partition is uploaded and assigned to workgroups only
// when a frame brings a new one.
//
// Frames are submitted without staging when the caller's planes are in
// page-locked host memory or on the device (xg_server_submit_frame): the
// dispatcher reads them by DMA.  A dispatcher thread keeps up to two batches
// in flight, one per device session (each with its own HIP stream), and each
// stream up to two frames.  When a session is free it takes the oldest pending
// frame of every stream with fewer than two frames in flight and enqueues, for
// all of them together: the plane copies and one padding launch; then, on the
// device, a wait for the other session's refinement (which wrote the MV
// fields, references and partitions this batch reads); then one AMVP launch
// (k_amvp_multi: each block's predictor chosen by prediction SAD among the
// AMVP candidates its stream's MV field gives it), one integer-search launch
// (k_search_groups, a workgroup per group of a CTU's blocks, +/-range around
// each block's predictor) and one refinement + MC launch (k_refine_multi,
// which also writes the winners into the MV fields), then the result copies.
// Results and predictions rotate with the pictures, so those copies never
// hold up the next batch.  The launches walk a per-batch table of per-stream
// descriptors, so a batch costs the host a few hundred bytes of descriptors,
// not a merged block list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <stddef.h>
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

// One plane whose padding ring is filled by edge replication.
struct PadPlane {
    uint16_t *data;                    // sample (0, 0)
    intptr_t stride;
    int w, h, pad;
};

// One workgroup per padded row of one plane (blockIdx.y = plane): rows above
// and below the picture are filled across their whole width, rows inside it
// only in their left and right padding.
__global__ void __launch_bounds__(256) k_pad_multi(const PadPlane *__restrict__ planes)
{
    const PadPlane P = planes[blockIdx.y];
    const int y = (int)blockIdx.x - P.pad;
    if (y >= P.h + P.pad) return;
    const int sy = min(max(y, 0), P.h - 1);
    uint16_t *row = P.data + (intptr_t)y * P.stride;
    const uint16_t *src = P.data + (intptr_t)sy * P.stride;
    if (y < 0 || y >= P.h) {
        for (int x = (int)threadIdx.x - P.pad; x < P.w + P.pad; x += 256) row[x] = src[min(max(x, 0), P.w - 1)];
    } else {
        for (int i = threadIdx.x; i < 2 * P.pad; i += 256) {
            const int x = i < P.pad ? i - P.pad : P.w + (i - P.pad);
            row[x] = src[i < P.pad ? 0 : P.w - 1];
        }
    }
}

// One device plane copied into a padded picture plane.
struct UploadPlane {
    const uint16_t *src;               // sample (0, 0), dword aligned, even stride
    intptr_t sstride;
    PadPlane dst;
};

// Copy and edge padding in one pass (device-resident frames): one workgroup
// per padded destination row (blockIdx.y = plane), each thread moving sample
// pairs; a pair left or right of the picture takes the clamped edge sample.
__global__ void __launch_bounds__(256) k_upload_pad(const UploadPlane *__restrict__ planes)
{
    const UploadPlane U = planes[blockIdx.y];
    const PadPlane &P = U.dst;
    const int y = (int)blockIdx.x - P.pad;
    if (y >= P.h + P.pad) return;
    const int sy = min(max(y, 0), P.h - 1);
    const uint16_t *src = U.src + (intptr_t)sy * U.sstride;
    uint32_t *row = reinterpret_cast<uint32_t *>(P.data + (intptr_t)y * P.stride - P.pad);
    const int pairs = (P.w + 2 * P.pad) >> 1;
    const uint32_t lo = src[0] * 0x10001u, hi = src[P.w - 1] * 0x10001u;
    for (int i = threadIdx.x; i < pairs; i += 256) {
        const int x = 2 * i - P.pad;
        row[i] = x < 0 ? lo : x >= P.w ? hi : *reinterpret_cast<const uint32_t *>(src + x);
    }
}

namespace {

constexpr int kPics = 3;                          // pictures per stream, in rotation
constexpr int kMaxStreams = 32;
static_assert(kMaxStreams <= 32, "a batch's streams are kept as a 32-bit mask");
static_assert(kMaxStreams * kPics <= kSessionPicSlots, "picture slots");
constexpr int kMaxInFlight = 2;                    // frames per stream on the device at once
constexpr int kSessions = 2;

struct Request {
    int stream = 0;
    uint64_t tag = 0;
    int64_t submitNs = 0;
    int planesKind = XG_MEM_HOST;
    const uint16_t *planes[3] = {nullptr, nullptr, nullptr};   // caller's (pinned or device)
    intptr_t strides[3] = {0, 0, 0};
    uint16_t *staged = nullptr;        // XG_MEM_HOST planes: Y | Cb | Cr packed at submit
    size_t stagedBytes = 0;
    uint16_t *predStage = nullptr;     // pageable prediction outputs: read back here first
    size_t predStageBytes = 0;
    xg_block_result *resStage = nullptr;   // pageable results: read back here first
    size_t resStageBytes = 0;
    bool newPartition = false;         // this frame brings a partition (uploaded at launch)
    std::vector<xg_block> blocks;
    std::vector<int2> tasksW;
    std::vector<int> wgBegin;
    std::vector<int> gidx;             // the partition's search groups (k_search_groups)
    std::vector<int2> groups;
    int nblocks = 0;
    int curPic = 0, refPic = 0;        // the stream's pictures this frame reads into / refers to
    int outputsKind = XG_MEM_HOST;
    xg_block_result *results = nullptr;
    xg_pixel *pred[3] = {nullptr, nullptr, nullptr};
    intptr_t predStride[3] = {0, 0, 0};
};

// Device state of a stream, carved from one allocation.
struct StreamDev {
    xg_picture pic[kPics];             // padded pictures in rotation
    xg_picture pred[kPics];            // unpadded prediction of the frame read into picture k
    int2 *field = nullptr;             // 4x4-unit MV field, fieldStride units per row
    int fieldStride = 0;
    xg_block *blocks = nullptr;        // partition (slots, predictor and base MV rewritten per frame)
    int2 *tasks = nullptr;             // workgroup-major task assignment
    int *wgBegin = nullptr;
    int *gidx = nullptr;               // search groups: block indices, group by group
    int2 *groups = nullptr;            //   and (first entry, count) per group
    uint32_t *candCost = nullptr;
    xg_block_result *best[kPics] = {}; // results of the frame read into picture k
    unsigned long long *bestKey = nullptr;
    int *done = nullptr;
};

struct Stream {
    bool open = false, closing = false;
    int inFlight = 0;                  // frames launched and not yet finalised
    int64_t launched = 0;              // frames launched so far: the next one reads picture launched % kPics
    int w = 0, h = 0;
    void *mem = nullptr;
    size_t bytes = 0;
    StreamDev d;
    int maxBlocks = 0, taskCap = 0;
    // partition last accepted at submit (host copy)
    std::vector<xg_block> lastBlocks;
    // partition currently on the device (set when a frame carrying it launches)
    int devBlocks = 0, devTasks = 0, devGroups = 0;
    std::deque<Request *> pending;
    int64_t frames = 0;
};

// Per-session descriptor tables of one batch: host (page-locked) and device.
struct Descs {
    SearchList search[kMaxStreams];
    RefineList refine[kMaxStreams];
    PadPlane pad[3 * kMaxStreams];     // planes copied by DMA, padded afterwards
    UploadPlane up[3 * kMaxStreams];   // device planes copied and padded in one pass
};

template <typename T>
T *dev_member(void *base, size_t offset)
{
    return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + offset);
}

struct Batch {
    bool active = false;
    std::vector<Request *> reqs;
    uint32_t streamMask = 0;           // the streams with a frame in this batch (kMaxStreams <= 32)
    hipEvent_t done = nullptr;         // after the batch's output copies
    hipEvent_t computed = nullptr;     // after its refinement: what the next batch's kernels depend on
    int64_t startNs = 0;
    Descs *hDesc = nullptr;
    Descs *dDesc = nullptr;
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
    xg_session *sess[kSessions] = {};
    Batch batch[kSessions];
    Stream streams[kMaxStreams];
    std::deque<xg_frame_result> completed;
    std::multimap<size_t, uint16_t *> pinnedPool;
    uint64_t bytesSessions = 0, bytesStreams = 0;
    int64_t framesDone = 0, framesDropped = 0, batches = 0;
    int grid = 1;

    uint16_t *pinned_get(size_t bytes)
    {
        auto it = pinnedPool.lower_bound(bytes);
        if (it != pinnedPool.end() && it->first == bytes) {
            uint16_t *p = it->second;
            pinnedPool.erase(it);
            return p;
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        return (uint16_t *)p;
    }
    void pinned_put(uint16_t *p, size_t bytes)
    {
        if (p) pinnedPool.emplace(bytes, p);
    }

    void finish_request(Request *r, int32_t status, int batchFrames, int64_t startNs)
    {
        xg_frame_result fr;
        fr.stream = r->stream;
        fr.status = status;
        fr.nblocks = r->nblocks;
        fr.batchFrames = batchFrames;
        fr.tag = r->tag;
        fr.submitNs = r->submitNs;
        fr.startNs = startNs;
        fr.doneNs = now_ns();
        completed.push_back(fr);
        if (status == XG_OK) framesDone++;
        else framesDropped++;
        pinned_put(r->staged, r->stagedBytes);
        pinned_put(r->predStage, r->predStageBytes);
        pinned_put(reinterpret_cast<uint16_t *>(r->resStage), r->resStageBytes);
        delete r;
    }

    // Finalises a completed batch: pageable outputs to the callers, reference swap.
    void finalize(int si)
    {
        Batch &b = batch[si];
        const int nf = (int)b.reqs.size();
        for (Request *r : b.reqs) {
            Stream &st = streams[r->stream];
            if (r->outputsKind == XG_MEM_HOST) {
                memcpy(r->results, r->resStage, sizeof(xg_block_result) * (size_t)r->nblocks);
                if (r->pred[0]) {
                    const int cw = st.w / 2, ch = st.h / 2;
                    unpack2d(r->pred[0], r->predStride[0], r->predStage, st.w, st.h, 2);
                    unpack2d(r->pred[1], r->predStride[1], r->predStage + (size_t)st.w * st.h, cw, ch, 2);
                    unpack2d(r->pred[2], r->predStride[2], r->predStage + (size_t)st.w * st.h + (size_t)cw * ch, cw,
                             ch, 2);
                }
            }
            st.inFlight--;
            st.frames++;
            finish_request(r, XG_OK, nf, b.startNs);
        }
        b.reqs.clear();
        b.active = false;
        cvDone.notify_all();
    }

    // A batch whose device work failed: every frame is reported with XG_ERR_HIP.
    void fail_batch(int si, int32_t rc)
    {
        Batch &b = batch[si];
        log_msg(0, "serve: batch of %zu frames failed (%d)", b.reqs.size(), rc);
        for (Request *r : b.reqs) {
            streams[r->stream].inFlight--;
            finish_request(r, rc, (int)b.reqs.size(), b.startNs);
        }
        b.reqs.clear();
        b.active = false;
        cvDone.notify_all();
    }

    static hipMemcpyKind kind_of(int memKind, bool toDevice)
    {
        if (memKind == XG_MEM_DEVICE) return hipMemcpyDeviceToDevice;
        return toDevice ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    }

    // Enqueues frame r's planes into its picture (padding is one launch for
    // the whole batch, afterwards).  The picture is one no frame in flight uses.
    // Device planes, dword aligned with even strides, take k_upload_pad.
    static bool fused_upload(const Request *r)
    {
        if (r->staged || r->planesKind != XG_MEM_DEVICE) return false;
        for (int p = 0; p < 3; p++)
            if (((uintptr_t)r->planes[p] & 3) || (r->strides[p] & 1)) return false;
        return true;
    }

    int32_t stage_planes(Request *r, hipStream_t hs)
    {
        if (fused_upload(r)) return XG_OK;
        Stream &st = streams[r->stream];
        const xg_picture &P = st.d.pic[r->curPic];
        const uint16_t *src = r->staged;
        for (int p = 0; p < 3; p++) {
            const xg_plane &pl = P.plane[p];
            const uint16_t *sp = r->staged ? src : r->planes[p];
            const size_t spitch = r->staged ? (size_t)pl.width * 2 : (size_t)r->strides[p] * 2;
            const hipMemcpyKind k = r->staged ? hipMemcpyHostToDevice : kind_of(r->planesKind, true);
            if (hipMemcpy2DAsync(pl.data, pl.stride * 2, sp, spitch, (size_t)pl.width * 2, pl.height, k, hs) !=
                hipSuccess)
                return XG_ERR_HIP;
            src += (size_t)pl.width * pl.height;
        }
        return XG_OK;
    }

    // Enqueues frame r's new partition, if it brings one (after the wait for
    // the frames that still use the old one).
    int32_t stage_partition(Request *r, hipStream_t hs)
    {
        if (!r->newPartition) return XG_OK;
        Stream &st = streams[r->stream];
        const size_t nb = r->blocks.size(), nt = r->tasksW.size();
        if (hipMemcpyAsync(st.d.blocks, r->blocks.data(), sizeof(xg_block) * nb, hipMemcpyHostToDevice, hs) !=
                hipSuccess ||
            (nt && hipMemcpyAsync(st.d.tasks, r->tasksW.data(), sizeof(int2) * nt, hipMemcpyHostToDevice, hs) !=
                       hipSuccess) ||
            hipMemcpyAsync(st.d.wgBegin, r->wgBegin.data(), sizeof(int) * r->wgBegin.size(), hipMemcpyHostToDevice,
                           hs) != hipSuccess ||
            hipMemcpyAsync(st.d.gidx, r->gidx.data(), sizeof(int) * r->gidx.size(), hipMemcpyHostToDevice, hs) !=
                hipSuccess ||
            hipMemcpyAsync(st.d.groups, r->groups.data(), sizeof(int2) * r->groups.size(), hipMemcpyHostToDevice,
                           hs) != hipSuccess)
            return XG_ERR_HIP;
        return XG_OK;
    }

    // Enqueues frame r's outputs: results and, when asked for, the prediction.
    int32_t read_outputs(Request *r, hipStream_t hs)
    {
        Stream &st = streams[r->stream];
        xg_block_result *dst = r->outputsKind == XG_MEM_HOST ? r->resStage : r->results;
        if (hipMemcpyAsync(dst, st.d.best[r->curPic], sizeof(xg_block_result) * (size_t)r->nblocks,
                           kind_of(r->outputsKind, false), hs) != hipSuccess)
            return XG_ERR_HIP;
        if (!r->pred[0]) return XG_OK;
        uint16_t *stage = r->predStage;
        for (int p = 0; p < 3; p++) {
            const xg_plane &pl = st.d.pred[r->curPic].plane[p];
            uint16_t *d = r->outputsKind == XG_MEM_HOST ? stage : r->pred[p];
            const size_t dpitch = r->outputsKind == XG_MEM_HOST ? (size_t)pl.width * 2 : (size_t)r->predStride[p] * 2;
            if (hipMemcpy2DAsync(d, dpitch, pl.data, pl.stride * 2, (size_t)pl.width * 2, pl.height,
                                 kind_of(r->outputsKind, false), hs) != hipSuccess)
                return XG_ERR_HIP;
            stage += (size_t)pl.width * pl.height;
        }
        return XG_OK;
    }

    bool launchable(const Stream &st) const
    {
        return st.open && !st.closing && st.inFlight < kMaxInFlight && !st.pending.empty();
    }

    // Assembles and enqueues one batch on session si; returns false if nothing to do.
    bool launch(int si)
    {
        Batch &b = batch[si];
        std::vector<int> ids;
        for (int i = 0; i < kMaxStreams; i++)
            if (launchable(streams[i])) ids.push_back(i);
        if (ids.empty()) return false;
        std::sort(ids.begin(), ids.end(), [&](int x, int y) {
            return streams[x].pending.front()->submitNs < streams[y].pending.front()->submitNs;
        });
        b.reqs.clear();
        b.streamMask = 0;
        int nblk = 0;
        for (int id : ids) {
            Stream &st = streams[id];
            Request *r = st.pending.front();
            if (!b.reqs.empty() && nblk + r->nblocks > cfg.maxBatchBlocks) break;
            st.pending.pop_front();
            r->curPic = (int)(st.launched % kPics);
            r->refPic = st.launched ? (int)((st.launched + kPics - 1) % kPics) : r->curPic;
            st.launched++;
            st.inFlight++;
            nblk += r->nblocks;
            b.reqs.push_back(r);
            b.streamMask |= 1u << id;
        }
        xg_session *s = sess[si];
        hipStream_t hs = (hipStream_t)xg_session_stream(s);
        b.startNs = now_ns();
        int32_t rc = XG_OK;
        for (Request *r : b.reqs)
            if (rc == XG_OK) rc = stage_planes(r, hs);
        // descriptor tables of the batch
        Descs &D = *b.hDesc;
        const int nf = (int)b.reqs.size();
        int begin = 0, blockBegin = 0, taskAcc = 0, nPad = 0, nUp = 0;
        for (int i = 0; i < nf && rc == XG_OK; i++) {
            Request *r = b.reqs[i];
            Stream &st = streams[r->stream];
            const int id = r->stream;
            if (r->newPartition) {        // the device partition changes with this frame
                st.devBlocks = (int)r->blocks.size();
                st.devTasks = (int)r->tasksW.size();
                st.devGroups = (int)r->groups.size();
            }
            D.search[i] = SearchList{st.d.blocks, st.d.field, st.d.fieldStride, begin,
                                     (uint8_t)(kPics * id + r->curPic), (uint8_t)(kPics * id + r->refPic),
                                     st.d.gidx, st.d.groups, blockBegin};
            begin += st.devGroups;
            blockBegin += st.devBlocks;
            // spread small partitions over the grid, in steps of whole XCD bands
            const int rot = (taskAcc / 8 * 8) % grid;
            taskAcc += st.devTasks;
            D.refine[i] = RefineList{st.d.blocks, st.d.tasks, st.d.wgBegin, st.d.candCost, st.d.best[r->curPic],
                                     st.d.bestKey, st.d.done, st.d.field, st.d.fieldStride, rot};
            for (int p = 0; p < 3; p++) {
                const xg_plane &pl = st.d.pic[r->curPic].plane[p];
                const PadPlane pp{pl.data, pl.stride, pl.width, pl.height, pl.pad};
                if (fused_upload(r)) D.up[nUp++] = UploadPlane{r->planes[p], r->strides[p], pp};
                else D.pad[nPad++] = pp;
            }
        }
        if (rc == XG_OK && hipMemcpyAsync(b.dDesc, b.hDesc, sizeof(Descs), hipMemcpyHostToDevice, hs) != hipSuccess)
            rc = XG_ERR_HIP;
        if (rc == XG_OK && nPad) {
            int rows = 0;
            for (int i = 0; i < nPad; i++) rows = std::max(rows, D.pad[i].h + 2 * D.pad[i].pad);
            k_pad_multi<<<dim3(rows, nPad), 256, 0, hs>>>(dev_member<PadPlane>(b.dDesc, offsetof(Descs, pad)));
            if (hipGetLastError() != hipSuccess) rc = XG_ERR_HIP;
        }
        if (rc == XG_OK && nUp) {
            int rows = 0;
            for (int i = 0; i < nUp; i++) rows = std::max(rows, D.up[i].dst.h + 2 * D.up[i].dst.pad);
            k_upload_pad<<<dim3(rows, nUp), 256, 0, hs>>>(dev_member<UploadPlane>(b.dDesc, offsetof(Descs, up)));
            if (hipGetLastError() != hipSuccess) rc = XG_ERR_HIP;
        }
        // if another session's batch holds the previous frames of any of these
        // streams, their MV fields, references and partitions come from it
        // (outputs rotate with the pictures, so its output copies need not be
        // waited for); batches of disjoint streams run side by side
        for (int oi = 0; oi < kSessions; oi++) {
            const Batch &o = batch[oi];
            if (rc == XG_OK && oi != si && o.active && (o.streamMask & b.streamMask) &&
                hipStreamWaitEvent(hs, o.computed, 0) != hipSuccess)
                rc = XG_ERR_HIP;
        }
        for (Request *r : b.reqs)
            if (rc == XG_OK) rc = stage_partition(r, hs);
        if (rc == XG_OK) {
            // AMVP predictors from the streams' fields, then the search around them
            SessionView v;
            session_view(s, &v);
            const SearchList *dl = dev_member<SearchList>(b.dDesc, offsetof(Descs, search));
            rc = amvp_multi_enqueue(hs, v.dPics, dl, nf, blockBegin, cfg.refine.ctuSize);
            if (rc == XG_OK)
                rc = search_groups_enqueue(hs, s, v.dPics, dl, nf, begin, cfg.searchRange, cfg.searchLambda,
                                           cfg.refine.ctuSize);
        }
        if (rc == XG_OK)
            rc = refine_multi_enqueue(s, &cfg.refine, dev_member<RefineList>(b.dDesc, offsetof(Descs, refine)), nf);
        if (rc == XG_OK && hipEventRecord(b.computed, hs) != hipSuccess) rc = XG_ERR_HIP;
        for (Request *r : b.reqs)
            if (rc == XG_OK) rc = read_outputs(r, hs);
        if (rc == XG_OK && hipEventRecord(b.done, hs) != hipSuccess) rc = XG_ERR_HIP;
        if (rc != XG_OK) {
            (void)hipStreamSynchronize(hs);
            fail_batch(si, rc);
            return true;
        }
        b.active = true;
        batches++;
        return true;
    }

    // Streams that may still bring a frame (open, not closing).
    int live_streams() const
    {
        int n = 0;
        for (int i = 0; i < kMaxStreams; i++) n += streams[i].open && !streams[i].closing;
        return n;
    }

    // Nanoseconds a free session should still wait to gather frames (INT64_MAX:
    // until more frames arrive).  Waits while fewer than minBatchFrames (capped
    // at the live streams) are launchable, and while the oldest launchable
    // frame is younger than maxDelayUs and the launchable frames do not fill a
    // batch; maxDelayUs > 0 also bounds the first wait.
    int64_t gather_wait_ns() const
    {
        int64_t oldest = INT64_MAX, nblk = 0;
        int n = 0;
        for (int i = 0; i < kMaxStreams; i++) {
            const Stream &st = streams[i];
            if (!launchable(st)) continue;
            oldest = std::min(oldest, st.pending.front()->submitNs);
            nblk += st.pending.front()->nblocks;
            n++;
        }
        if (n == 0 || nblk >= cfg.maxBatchBlocks) return 0;
        const int64_t left = cfg.maxDelayUs > 0 ? oldest + (int64_t)cfg.maxDelayUs * 1000 - now_ns() : 0;
        if (cfg.minBatchFrames > 0 && n < std::min(cfg.minBatchFrames, live_streams()))
            return cfg.maxDelayUs > 0 ? std::max<int64_t>(0, left) : INT64_MAX;
        return std::max<int64_t>(0, left);
    }

    void loop()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            for (int i = 0; i < kSessions; i++) {
                if (!batch[i].active) continue;
                const hipError_t q = hipEventQuery(batch[i].done);
                if (q == hipSuccess) finalize(i);
                else if (q != hipErrorNotReady) fail_batch(i, XG_ERR_HIP);
            }
            const int64_t gather = stop ? 0 : gather_wait_ns();
            bool anyActive = false;
            for (int i = 0; i < kSessions; i++) anyActive |= batch[i].active;
            if (gather > 0) {
                // wake for new frames, or to poll the batches in flight
                const int64_t w = anyActive ? std::min<int64_t>(gather, 100000) : std::min<int64_t>(gather, 50000000);
                cvWork.wait_for(lk, std::chrono::nanoseconds(w));
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
                const hipError_t se = hipEventSynchronize(e);
                lk.lock();
                if (se != hipSuccess && batch[oldest].active) fail_batch(oldest, XG_ERR_HIP);
            } else if (active > 0) {
                cvWork.wait_for(lk, std::chrono::microseconds(100));
            } else {
                cvWork.wait(lk);
            }
        }
    }

    // Validates a partition (cheap; under the lock).  Returns whether it
    // differs from the stream's last one, i.e. needs a task assignment.
    int32_t check_partition(int id, const xg_block *blocks, int n, bool *changed)
    {
        Stream &st = streams[id];
        for (int i = 0; i < n; i++) {   // malformed descriptors are rejected before anything is queued
            const xg_block &B = blocks[i];
            const bool okw = B.w >= 4 && B.w <= 64 && !(B.w & (B.w - 1)),
                       okh = B.h >= 4 && B.h <= 64 && !(B.h & (B.h - 1));
            if (!okw || !okh || B.x < 0 || B.y < 0 || (B.x & 3) || (B.y & 3) || B.x + B.w > st.w ||
                B.y + B.h > st.h) {
                set_error(XG_ERR_ARG, "xg_server_submit: block %d (%d,%d %dx%d) is malformed", i, B.x, B.y, B.w, B.h);
                return XG_ERR_ARG;
            }
        }
        if (n > st.maxBlocks) {
            set_error(XG_ERR_ARG, "xg_server_submit: %d blocks, more than the stream's %d", n, st.maxBlocks);
            return XG_ERR_ARG;
        }
        *changed = !(n == (int)st.lastBlocks.size() && memcmp(blocks, st.lastBlocks.data(), sizeof(xg_block) * n) == 0);
        return XG_OK;
    }

    // Queues r; a frame bringing a partition makes it the stream's last one.
    int32_t enqueue(int id, Request *r, const xg_block *blocks)
    {
        std::lock_guard<std::mutex> g(mu);
        Stream &st = streams[id];
        if (!st.open || st.closing) {
            set_error(XG_ERR_STATE, "xg_server_submit: stream %d closed", id);
            return XG_ERR_STATE;
        }
        if (r->newPartition) st.lastBlocks.assign(blocks, blocks + r->nblocks);
        r->submitNs = now_ns();
        st.pending.push_back(r);
        cvWork.notify_one();
        return XG_OK;
    }
};

namespace {

// Per-stream device layout: offsets of every part of the one allocation.
struct StreamLayout {
    size_t pic[kPics], pred[kPics], field, blocks, tasks, wgBegin, gidx, groups, candCost, best[kPics], bestKey,
        done, total;
    int pad, ys, cs, maxBlocks, taskCap, fieldStride;
    size_t ypic, cpic;
};

StreamLayout stream_layout(const xg_server_cfg &cfg, int width, int height, int grid, const int *taskPerBlock)
{
    StreamLayout L;
    L.pad = cfg.refine.ctuSize + 16;
    const int cpad = L.pad / 2;
    L.ys = width + 2 * L.pad;
    L.cs = width / 2 + 2 * cpad;
    L.ypic = (size_t)L.ys * (height + 2 * L.pad);
    L.cpic = (size_t)L.cs * (height / 2 + 2 * cpad);
    const size_t picBytes = align_up((L.ypic + 2 * L.cpic) * 2, 4096);
    L.maxBlocks = (width / 4) * (height / 4);
    L.taskCap = L.maxBlocks * *taskPerBlock;
    L.fieldStride = width / 4;
    size_t o = 0;
    for (int k = 0; k < kPics; k++) {
        L.pic[k] = o;
        o += picBytes;
    }
    for (int k = 0; k < kPics; k++) {
        L.pred[k] = o;
        o += align_up((size_t)width * height * 3, 4096);
    }
    L.field = o; o += align_up((size_t)L.maxBlocks * sizeof(int2), 256);
    L.blocks = o; o += align_up((size_t)L.maxBlocks * sizeof(xg_block), 256);
    L.tasks = o; o += align_up((size_t)L.taskCap * sizeof(int2), 256);
    L.wgBegin = o; o += align_up((size_t)(grid + 1) * sizeof(int), 256);
    L.gidx = o; o += align_up((size_t)L.maxBlocks * sizeof(int), 256);
    L.groups = o; o += align_up((size_t)L.maxBlocks * sizeof(int2), 256);
    L.candCost = o; o += align_up((size_t)L.maxBlocks * cfg.refine.ncand * sizeof(uint32_t), 256);
    for (int k = 0; k < kPics; k++) {
        L.best[k] = o;
        o += align_up((size_t)L.maxBlocks * sizeof(xg_block_result), 256);
    }
    L.bestKey = o; o += align_up((size_t)L.maxBlocks * sizeof(unsigned long long), 256);
    L.done = o; o += align_up((size_t)L.maxBlocks * sizeof(int), 256);
    L.total = o;
    return L;
}

// Tasks per block at most, over every block shape, for the session's tiles.
int max_tasks_per_block(xg_session *s, int ncand)
{
    int m = 1;
    for (int lw = 2; lw <= 6; lw++)
        for (int lh = 2; lh <= 6; lh++) {
            const int t = std::max(1, std::min(ncand, (int)xg_session_get_tile(s, lw, lh)));
            m = std::max(m, (ncand + t - 1) / t);
        }
    return m;
}

}  // namespace

extern "C" {

int32_t xg_server_create(xg_server **out, const xg_server_cfg *cfg)
{
    if (!out || !cfg || cfg->maxStreams < 1 || cfg->maxStreams > kMaxStreams || cfg->maxBatchBlocks < 1 ||
        cfg->searchRange < 0 || cfg->searchRange > 32 || cfg->maxDelayUs < 0 || cfg->maxDelayUs > 1000000 ||
        cfg->minBatchFrames < 0 || cfg->minBatchFrames > kMaxStreams || cfg->refine.ncand < 1 ||
        cfg->refine.ncand > XG_MAX_CAND || cfg->refine.ctuSize < 16 || cfg->refine.ctuSize > 128) {
        set_error(XG_ERR_ARG, "xg_server_create: bad configuration");
        return XG_ERR_ARG;
    }
    *out = nullptr;
    if (!device_init()) return XG_ERR_HIP;
    xg_server *srv = new xg_server();
    srv->cfg = *cfg;
    srv->bytesSessions = (uint64_t)kSessions * (sizeof(Descs) + 4096);
    if (srv->bytesSessions > cfg->hbmBudget) {
        set_error(XG_ERR_BUDGET, "xg_server_create: sessions alone need %llu bytes",
                  (unsigned long long)srv->bytesSessions);
        delete srv;
        return XG_ERR_BUDGET;
    }
    for (int i = 0; i < kSessions; i++) {
        if (xg_session_create(&srv->sess[i], 1) != XG_OK || xg_session_set_sched(srv->sess[i], 8, 16) != XG_OK ||
            hipEventCreateWithFlags(&srv->batch[i].done, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&srv->batch[i].computed, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&srv->batch[i].hDesc, sizeof(Descs), hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void **)&srv->batch[i].dDesc, sizeof(Descs)) != hipSuccess) {
            set_error(XG_ERR_NOMEM, "xg_server_create: session allocation failed");
            xg_server_destroy(srv);
            return XG_ERR_NOMEM;
        }
    }
    srv->grid = refine_grid(srv->sess[0]);
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
            if (r->staged) (void)hipHostFree(r->staged);
            if (r->predStage) (void)hipHostFree(r->predStage);
            if (r->resStage) (void)hipHostFree(r->resStage);
            delete r;
        }
        st.pending.clear();
        if (st.mem) (void)hipFree(st.mem);
    }
    for (int i = 0; i < kSessions; i++) {
        if (srv->sess[i]) xg_session_destroy(srv->sess[i]);
        if (srv->batch[i].done) (void)hipEventDestroy(srv->batch[i].done);
        if (srv->batch[i].computed) (void)hipEventDestroy(srv->batch[i].computed);
        if (srv->batch[i].hDesc) (void)hipHostFree(srv->batch[i].hDesc);
        if (srv->batch[i].dDesc) (void)hipFree(srv->batch[i].dDesc);
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
        else if (id < 0 && !srv->streams[i].mem) id = i;
    }
    if (id < 0 || open >= srv->cfg.maxStreams) {
        set_error(XG_ERR_BUDGET, "xg_server_open_stream: all %d stream slots in use", srv->cfg.maxStreams);
        return XG_ERR_BUDGET;
    }
    const int tpb = max_tasks_per_block(srv->sess[0], srv->cfg.refine.ncand);
    const StreamLayout L = stream_layout(srv->cfg, width, height, srv->grid, &tpb);
    if (srv->bytesSessions + srv->bytesStreams + L.total > srv->cfg.hbmBudget) {
        set_error(XG_ERR_BUDGET, "xg_server_open_stream: %zu bytes do not fit the HBM budget", L.total);
        return XG_ERR_BUDGET;
    }
    Stream &st = srv->streams[id];
    if (!XG_HIP(hipMalloc(&st.mem, L.total))) return XG_ERR_NOMEM;
    char *base = (char *)st.mem;
    const int pad = L.pad, cpad = pad / 2;
    for (int k = 0; k < kPics; k++) {
        uint16_t *p = (uint16_t *)(base + L.pic[k]);
        st.d.pic[k].plane[0] = xg_plane{p + pad * L.ys + pad, L.ys, width, height, pad};
        uint16_t *c1 = p + L.ypic, *c2 = c1 + L.cpic;
        st.d.pic[k].plane[1] = xg_plane{c1 + cpad * L.cs + cpad, L.cs, width / 2, height / 2, cpad};
        st.d.pic[k].plane[2] = xg_plane{c2 + cpad * L.cs + cpad, L.cs, width / 2, height / 2, cpad};
    }
    for (int k = 0; k < kPics; k++) {
        uint16_t *pr = (uint16_t *)(base + L.pred[k]);
        st.d.pred[k].plane[0] = xg_plane{pr, width, width, height, 0};
        st.d.pred[k].plane[1] = xg_plane{pr + (size_t)width * height, width / 2, width / 2, height / 2, 0};
        st.d.pred[k].plane[2] = xg_plane{pr + (size_t)width * height * 5 / 4, width / 2, width / 2, height / 2, 0};
        st.d.best[k] = (xg_block_result *)(base + L.best[k]);
    }
    st.d.field = (int2 *)(base + L.field);
    st.d.fieldStride = L.fieldStride;
    st.d.blocks = (xg_block *)(base + L.blocks);
    st.d.tasks = (int2 *)(base + L.tasks);
    st.d.wgBegin = (int *)(base + L.wgBegin);
    st.d.gidx = (int *)(base + L.gidx);
    st.d.groups = (int2 *)(base + L.groups);
    st.d.candCost = (uint32_t *)(base + L.candCost);
    st.d.bestKey = (unsigned long long *)(base + L.bestKey);
    st.d.done = (int *)(base + L.done);
    // MV field starts at zero; per-block completion words idle (all-ones / zero)
    bool ok = XG_HIP(hipMemset(st.d.field, 0, (size_t)L.maxBlocks * sizeof(int2))) &&
              XG_HIP(hipMemset(st.d.bestKey, 0xFF, (size_t)L.maxBlocks * sizeof(unsigned long long))) &&
              XG_HIP(hipMemset(st.d.done, 0, (size_t)L.maxBlocks * sizeof(int))) &&
              XG_HIP(hipDeviceSynchronize());
    // the dispatcher holds no session call while we hold the lock
    for (int si = 0; si < kSessions && ok; si++)
        for (int k = 0; k < kPics && ok; k++)
            ok = xg_session_set_picture(srv->sess[si], kPics * id + k, &st.d.pic[k]) == XG_OK &&
                 xg_session_set_pred(srv->sess[si], kPics * id + k, &st.d.pred[k]) == XG_OK;
    if (!ok) {
        (void)hipFree(st.mem);
        st.mem = nullptr;
        return XG_ERR_HIP;
    }
    st.open = true;
    st.closing = false;
    st.inFlight = 0;
    st.launched = 0;
    st.w = width;
    st.h = height;
    st.bytes = L.total;
    st.maxBlocks = L.maxBlocks;
    st.taskCap = L.taskCap;
    st.lastBlocks.clear();
    st.devBlocks = st.devTasks = st.devGroups = 0;
    st.frames = 0;
    srv->bytesStreams += L.total;
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
    srv->cvWork.notify_all();          // a batch held for more streams may go now
    srv->cvDone.wait(lk, [&] { return st.inFlight == 0; });
    (void)hipFree(st.mem);
    st.mem = nullptr;
    srv->bytesStreams -= st.bytes;
    st.open = st.closing = false;
    st.lastBlocks.clear();
    return XG_OK;
}

int32_t xg_server_submit_frame(xg_server *srv, int32_t streamId, const xg_server_frame *f)
{
    if (!srv || !f || streamId < 0 || streamId >= kMaxStreams || !f->results ||
        f->planesKind < XG_MEM_HOST || f->planesKind > XG_MEM_DEVICE || f->outputsKind < XG_MEM_HOST ||
        f->outputsKind > XG_MEM_DEVICE || !f->planes[0] || !f->planes[1] || !f->planes[2] ||
        (f->blocks && f->nblocks < 1)) {
        set_error(XG_ERR_ARG, "xg_server_submit_frame: bad arguments");
        return XG_ERR_ARG;
    }
    Request *r = new Request();
    int w, h, taskCap = 0;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        Stream &st = srv->streams[streamId];
        if (!st.open || st.closing) {
            delete r;
            set_error(XG_ERR_STATE, "xg_server_submit: stream %d is not open", streamId);
            return XG_ERR_STATE;
        }
        w = st.w;
        h = st.h;
        for (int p = 0; p < 3; p++)
            if (f->strides[p] < (p ? w / 2 : w)) {
                delete r;
                set_error(XG_ERR_ARG, "xg_server_submit: plane %d stride %ld below its width", p, (long)f->strides[p]);
                return XG_ERR_ARG;
            }
        int32_t rc;
        if (f->blocks) {
            rc = srv->check_partition(streamId, f->blocks, f->nblocks, &r->newPartition);
            r->nblocks = f->nblocks;
            taskCap = st.taskCap;
        } else if (st.lastBlocks.empty()) {
            set_error(XG_ERR_STATE, "xg_server_submit: stream %d has no partition yet", streamId);
            rc = XG_ERR_STATE;
        } else {
            r->nblocks = (int)st.lastBlocks.size();
            rc = XG_OK;
        }
        if (rc != XG_OK) {
            delete r;
            return rc;
        }
        if (f->outputsKind == XG_MEM_HOST) {
            r->resStageBytes = sizeof(xg_block_result) * (size_t)r->nblocks;
            r->resStage = reinterpret_cast<xg_block_result *>(srv->pinned_get(r->resStageBytes));
            if (f->pred[0]) {
                r->predStageBytes = (size_t)w * h * 3;
                r->predStage = srv->pinned_get(r->predStageBytes);
            }
            if (!r->resStage || (f->pred[0] && !r->predStage)) {
                srv->pinned_put(reinterpret_cast<uint16_t *>(r->resStage), r->resStageBytes);
                srv->pinned_put(r->predStage, r->predStageBytes);
                delete r;
                set_error(XG_ERR_NOMEM, "xg_server_submit: pinned staging allocation failed");
                return XG_ERR_NOMEM;
            }
        }
        if (f->planesKind == XG_MEM_HOST) {
            r->stagedBytes = (size_t)w * h * 3;
            r->staged = srv->pinned_get(r->stagedBytes);
            if (!r->staged) {
                srv->pinned_put(reinterpret_cast<uint16_t *>(r->resStage), r->resStageBytes);
                srv->pinned_put(r->predStage, r->predStageBytes);
                delete r;
                set_error(XG_ERR_NOMEM, "xg_server_submit: pinned staging allocation failed");
                return XG_ERR_NOMEM;
            }
        }
    }
    r->stream = streamId;
    r->tag = f->tag;
    r->planesKind = f->planesKind;
    for (int p = 0; p < 3; p++) {
        r->planes[p] = f->planes[p];
        r->strides[p] = f->strides[p];
    }
    r->outputsKind = f->outputsKind;
    r->results = f->results;
    if (f->pred[0])
        for (int p = 0; p < 3; p++) {
            r->pred[p] = f->pred[p];
            r->predStride[p] = f->predStride[p] ? f->predStride[p] : (p ? w / 2 : w);
        }
    auto drop = [&](int32_t code) {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->pinned_put(r->staged, r->stagedBytes);
        srv->pinned_put(r->predStage, r->predStageBytes);
        srv->pinned_put(reinterpret_cast<uint16_t *>(r->resStage), r->resStageBytes);
        delete r;
        return code;
    };
    if (r->newPartition) {   // the task assignment runs outside the server lock
        r->blocks.assign(f->blocks, f->blocks + f->nblocks);
        for (xg_block &B : r->blocks) B.mvx = B.mvy = B.mvpx = B.mvpy = 0;
        refine_assign(srv->sess[0], r->blocks.data(), f->nblocks, srv->cfg.refine.ncand, r->tasksW, r->wgBegin);
        search_groups(r->blocks.data(), f->nblocks, r->gidx, r->groups);
        if ((int)r->tasksW.size() > taskCap) {
            set_error(XG_ERR_ARG, "xg_server_submit: partition needs %zu tasks, more than %d", r->tasksW.size(),
                      taskCap);
            return drop(XG_ERR_ARG);
        }
    }
    if (r->staged) {      // pageable planes: copied before returning
        pack2d(r->staged, f->planes[0], f->strides[0], w, h, 2);
        pack2d(r->staged + (size_t)w * h, f->planes[1], f->strides[1], w / 2, h / 2, 2);
        pack2d(r->staged + (size_t)w * h * 5 / 4, f->planes[2], f->strides[2], w / 2, h / 2, 2);
    }
    const int32_t rc = srv->enqueue(streamId, r, f->blocks);
    return rc == XG_OK ? rc : drop(rc);
}

int32_t xg_server_submit(xg_server *srv, int32_t streamId, const xg_pixel *const planes[3], const intptr_t strides[3],
                         const xg_block *blocks, int32_t nblocks, xg_block_result *results, xg_pixel *const pred[3],
                         const intptr_t predStride[3], uint64_t tag)
{
    if (!planes || !strides || !blocks || nblocks < 1) {
        set_error(XG_ERR_ARG, "xg_server_submit: bad arguments");
        return XG_ERR_ARG;
    }
    xg_server_frame f;
    memset(&f, 0, sizeof(f));
    f.planesKind = XG_MEM_HOST;
    f.outputsKind = XG_MEM_HOST;
    for (int p = 0; p < 3; p++) {
        f.planes[p] = planes[p];
        f.strides[p] = strides[p];
        if (pred) {
            f.pred[p] = pred[p];
            f.predStride[p] = predStride ? predStride[p] : 0;
        }
    }
    f.blocks = blocks;
    f.nblocks = nblocks;
    f.results = results;
    f.tag = tag;
    return xg_server_submit_frame(srv, streamId, &f);
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
