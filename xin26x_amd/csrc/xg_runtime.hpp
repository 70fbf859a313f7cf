// xg_runtime.hpp -- host runtime shared by the C-ABI entry points: error
// state, logging, HIP checks and the per-thread call context (stream plus
// device and pinned staging buffers) used by the synchronous per-call
// primitives.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/xg266.h"

namespace xg {

void set_error(int32_t code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void log_msg(int32_t level, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// Returns true when e == hipSuccess; otherwise records XG_ERR_HIP.
bool hip_ok(hipError_t e, const char *what);

#define XG_HIP(call) ::xg::hip_ok((call), #call)

// Per-thread context of the synchronous per-call primitives.  The reference
// calls its primitives from every worker thread concurrently
// (h266_encode_section.c:63), so each host thread owns its stream and
// staging buffers; nothing mutable is shared between threads.
struct CallCtx {
    hipStream_t stream = nullptr;
    char *dev = nullptr;   size_t devCap = 0;
    char *host = nullptr;  size_t hostCap = 0;
};

// Returns the calling thread's context with at least devBytes of device
// scratch and hostBytes of pinned host staging, or nullptr (error recorded).
CallCtx *call_ctx(size_t devBytes, size_t hostBytes);

// Ensures the HIP runtime and device are usable and the device-side constant
// tables (transform matrices) are initialised.  Idempotent, thread-safe.
bool device_init();

// 2-D copies between strided host memory and packed staging memory.
void pack2d(void *dst, const void *src, intptr_t srcStride, int w, int h, size_t elem);
void unpack2d(void *dst, intptr_t dstStride, const void *src, int w, int h, size_t elem);

// Releases the forward-transform plan a session owns (xg_fdct.hip).
void fdct_release(xg_session *s);
// Releases the reconstruction plan a session owns (xg_recon.hip).
void recon_release(xg_session *s);
// Releases the DMVR plan a session owns (xg_dmvr.hip).
void dmvr_release(xg_session *s);
// Releases the half-resolution pictures of the lookahead search (xg_lookahead.hip).
void lookahead_release(xg_session *s);
// Releases the level pictures and fields of the MCTF search (xg_mctf.hip).
void mctf_release(xg_session *s);
// Releases the bi-prediction distortion plan a session owns (xg_bipred.hip).
void bipred_release(xg_session *s);
// Releases the quantiser plan a session owns (xg_quant.hip).
void quant_release(xg_session *s);

// Frame full search of xg_full_search_frame enqueued on `st`; `timer`, when
// not null, is the session whose stage-3 timer brackets the launch.
int32_t full_search_frame_on(hipStream_t st, xg_session *timer, const xg_plane *cur, const xg_plane *ref,
                             int32_t bw, int32_t bh, int32_t range, uint32_t sadLambda, int32_t ctuSize,
                             int16_t *bestMv, uint32_t *bestCost, uint32_t *bestSad);

void pipe_release(xg_session *s);
// Refinement of the session's current blocks, validated and uploaded but not
// launched, for callers that enqueue (or capture) it among their own kernels.
int32_t refine_prepare(xg_session *s, const xg_refine_cfg *cfg, const xg_picture *pred);
bool refine_enqueue_prepared(xg_session *s);
// Hash of everything the prepared refinement launch takes by value (grid,
// task lists, per-block state and output pointers, configuration): a graph
// that captured refine_enqueue_prepared() is valid only under the same key.
uint64_t refine_prepared_key(xg_session *s);
void *&session_pipe(xg_session *s);

// One frame's blocks in a multi-list refinement launch (serving): several
// frames of several streams share one persistent grid, each with its own
// resident block list, task assignment and outputs.
struct RefineList {
    const xg_block *blocks;            // device block list (slots, base MV, predictor)
    const int2 *tasks;                 // workgroup-major tasks for the session's grid
    const int *wgBegin;                // grid + 1 entries
    uint32_t *candCost;                // [n * ncand]
    xg_block_result *best;             // [n]
    unsigned long long *bestKey;       // [n], all-ones when idle
    int *done;                         // [n], zero when idle
    int2 *mvField;                     // 4x4-unit MV field the winners are written to (or null)
    int mvStride;                      // units per field row
    int rot;                           // workgroup g runs virtual workgroup (g + rot) mod grid
};
// One frame's blocks in a multi-list integer-search launch (serving).
struct SearchList {
    xg_block *blocks;                  // device; slots, predictor and base MV are written here
    const int2 *mvField;               // the stream's 4x4-unit MV field (AMVP candidate source)
    int mvStride;                      // units per field row
    int begin;                         // first workgroup of this list in the launch
    uint8_t curSlot, refSlot;          // picture slots of the frame and its reference
    const int *gidx;                   // k_search_groups: block indices, group by group
    const int2 *groups;                // k_search_groups: (first entry of gidx, count) per group
    int blockBegin;                    // first workgroup of this list in the k_amvp_multi launch
};
// Groups of a partition for k_search_groups: blocks of one 64x64 CTU, at most
// kSearchGroupMax per group, whose current samples span at most 64x64.
constexpr int kSearchGroupMax = 64;
void search_groups(const xg_block *blocks, int n, std::vector<int> &gidx, std::vector<int2> &groups);
// Enqueues k_search_groups over lists[0..nLists) (device) covering nGroups
// groups in all on `st`: every block searched +/-range around the predictor
// its list's blocks carry (mvpx, mvpy), with the results of search_block.
int32_t search_groups_enqueue(hipStream_t st, xg_session *timer, const xg_picture *dPics, const SearchList *dLists,
                              int nLists, int nGroups, int range, uint32_t lambda, int ctu);
// Enqueues k_amvp_multi over lists[0..nLists) (device) covering nBlocks blocks
// in all on `st`: each block takes its frame's picture slots and, as
// predictor, the AMVP choice (Xin266SelectMvp, quarter-pel AMVR) among the
// candidates its stream's MV field gives it (xg_amvp_from_field), written to
// the block's mvpx / mvpy in 1/16 pel.
int32_t amvp_multi_enqueue(hipStream_t st, const xg_picture *dPics, const SearchList *dLists, int nLists,
                           int nBlocks, int ctu);

// Persistent workgroups of the session's batched form (CUs x workgroups per CU).
int refine_grid(xg_session *s);
// Workgroup-major tasks of `n` blocks for the session's grid, tile table and
// bands, as the batched form builds them for its own block list.
void refine_assign(xg_session *s, const xg_block *blocks, int n, int ncand, std::vector<int2> &tasksW,
                   std::vector<int> &wgBegin);
// Enqueues one k_refine_multi launch over device lists[0..nLists) on the
// session stream, with the configuration's offsets and the session's picture
// and prediction tables.
int32_t refine_multi_enqueue(xg_session *s, const xg_refine_cfg *cfg, const RefineList *dLists, int nLists);

// Session kernel timers (HIP events on the session stream), by stage:
// 0 refinement (eager), 1 graph replay, 2 forward transform, 3 integer search,
// 4 a whole motion-search pipeline frame, 5 reconstruction, 6 DMVR,
// 7 bi-prediction distortion, 8 quantisation.
constexpr int kTimerStages = 11;
bool session_timer_begin(xg_session *s, int stage, int ev[2]);   // false when timing is off
void session_timer_end(xg_session *s, int stage, const int ev[2]);

// What other translation units need of a session (xg_refine.hip owns it).
struct SessionView {
    hipStream_t stream;
    const xg_picture *dPics;       // device picture table
    xg_block *dBlocks;             // device block list
    const xg_block *hBlocks;       // host copy of the block list
    int nblocks;
    const xg_picture *hPics;       // host picture table
    const bool *picSet;
    int numCUs;
    uint64_t blocksVersion;        // bumped by every xg_session_set_blocks
};
bool session_view(xg_session *s, SessionView *v);
constexpr int kSessionPicSlots = 128;

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace xg
