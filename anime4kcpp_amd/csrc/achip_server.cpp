// Serving loop of the HIP path (include/achip.h, achip_server_*).
//
// One scheduler thread per server.  Requests wait in arrival order until their device
// buffers fit the HBM budget.  One-off requests then move through upload -> colour split
// -> log2(factor) 2x passes -> chroma resize and merge.  The passes of all admitted one-offs
// are batched STEP by STEP into one persistent launch with a tile work-queue: each step
// takes, in admission order, tile-row ranges of the requests' current passes until the
// step's tile budget is spent, so a big pass spreads over several steps and a request
// admitted meanwhile joins the next step instead of waiting for the big one to finish.
// A request contributes at most one pass to a step (its next pass reads this one's
// output); all steps run on one stream, so stream order carries the dependencies.
// Video-stream frames (fixed geometry and format per stream) take a separate path: the luma
// shifts and pass and the chroma resizes of a frame are captured once per stream into a
// hipGraph and replayed on that stream's own HIP stream, overlapping the batched steps.
// Luma, YUV420 and the general YUV kind share one path: a frame format (planes, chroma
// subsampling, 8- or 16-bit words, LSB shift, chroma filter) as the reference's video
// module describes frames (video/src/Pipeline.cpp:495-566, cli/src/Main.cpp:183-205).
// u8 and u16 passes of one step go into separate persistent launches.
// Every kernel is one of the library's own (the same tiles and arithmetic as a request
// processed alone), so results do not depend on what else is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "achip.h"
#include "achip_internal.hpp"
#include "chunk_pool.hpp"

namespace {

using Clock = std::chrono::steady_clock;
constexpr int kU8 = 0x001;
constexpr int kU16 = 0x002;
constexpr int kMaxStreams = 64;
constexpr int kRing = 4;              // descriptor workspaces (steps in flight <= kRing - 2)
constexpr int kMaxStepItems = 512;    // descriptors per step
constexpr size_t kBlock = 2u << 20;   // allocation granularity

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t round_block(size_t v) { return (std::max<size_t>(v, 1) + kBlock - 1) / kBlock * kBlock; }

// hipMalloc / hipFree behind the serving loop's chunk pool
struct HipChunks {
    bool alloc(void** p, size_t bytes)
    {
        if (hipMalloc(p, bytes) == hipSuccess) return true;
        (void)hipGetLastError();
        return false;
    }
    void release(void* p) { (void)hipFree(p); }
};
using DevicePool = achip::ChunkPool<HipChunks>;

struct Plane {
    void* ptr = nullptr;
    long stride = 0;
    int w = 0, h = 0, ch = 1, es = 1;
    size_t bytes() const { return static_cast<size_t>(stride) * h; }
    long row() const { return static_cast<long>(w) * ch * es; }
};

// chroma planes of a frame format: count, size and interleaved channels
struct Chroma {
    int n = 0, w = 0, h = 0, ch = 1;
};
Chroma chroma_of(const ACHipFrameFormat& f, int W, int H)
{
    Chroma c;
    if (f.planes <= 1) return c;
    c.n = f.planes - 1;
    c.w = W >> f.chroma_shift_x;
    c.h = H >> f.chroma_shift_y;
    c.ch = f.planes == 2 ? 2 : 1;
    return c;
}
int chroma_mode(const ACHipFrameFormat& f)
{
    return f.chroma_filter == ACHIP_CHROMA_CATMULL_ROM ? achip::kResizeCatmullRom : achip::kResizeBilinear;
}
bool same_format(const ACHipFrameFormat& a, const ACHipFrameFormat& b)
{
    return a.planes == b.planes && a.chroma_shift_x == b.chroma_shift_x && a.chroma_shift_y == b.chroma_shift_y &&
           a.element_type == b.element_type && a.shift == b.shift && a.chroma_filter == b.chroma_filter;
}

enum class St { Pending, Uploading, Passing, Finishing, Done };

struct Req {
    unsigned long long ticket = 0;
    ACHipRequest r{};
    ACHipFrameFormat fmt{};           // luma/YUV kinds: the normalised frame format
    int type = kU8;                   // element type of every plane
    int passes = 0;
    size_t need = 0;                  // bytes reserved against the budget
    Clock::time_point t_submit, t_done;
    std::atomic<St> st{ St::Pending };   // written by the scheduler, read by waiters
    int err = ACHIP_OK;
    std::string msg;
    // one-off buffers: in (RGB), luma[0..passes], uv (RGB), chroma in c0 / out c1 (YUV), out (RGB)
    Plane in, luma[4], uv, c0[2], c1[2], out;
    Plane ws;                         // layer-by-layer nets: feature maps of the largest pass
    int pass = 0, next_row = 0;
    hipEvent_t ev_up = nullptr;       // upload done
    hipEvent_t ev_done = nullptr;     // everything done (own event, or a step's / a frame's)
    bool own_done_event = false;
    bool keep_output = false;         // dst all NULL: output stays until release
    long long out_pixels = 0;
};

struct Slot {                          // one video stream
    int w = 0, h = 0;
    hipStream_t stream = nullptr;
    Plane y0, c0[2], y1, c1[2];
    Plane ws;                          // layer-by-layer nets: feature maps of a frame's pass
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::deque<std::shared_ptr<Req>> inflight;
    // admitted frames not started yet: a stream's next frame starts once the one before has
    // completed, because a host-to-device copy queued behind kernel work can hold the
    // calling thread until that work drains (measured: ~10 ms per frame while pass steps
    // fill the device), and the scheduler thread must not wait
    std::deque<std::shared_ptr<Req>> waiting;
};

struct Step {
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<std::shared_ptr<Req>> finishing;
};

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    const size_t k = std::min(v.size() - 1, static_cast<size_t>(q * (v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
}

} // namespace

struct ACHipServer {
    ACHipNet* net = nullptr;
    ACHipServerConfig cfg{};
    int config = 0, tw = 64, th = 64;
    bool layered = false;             // ACNet<8> / ArtCNN / FSRCNNX: whole passes, not tile rows

    std::mutex mtx;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::thread thread;
    unsigned long long next_ticket = 1;
    std::deque<std::shared_ptr<Req>> pending;
    std::unordered_map<unsigned long long, std::shared_ptr<Req>> tickets;
    std::vector<std::shared_ptr<Req>> released;         // kept outputs to free (scheduler thread)
    size_t reserved = 0;
    int stream_w[kMaxStreams] = {}, stream_h[kMaxStreams] = {};
    ACHipFrameFormat stream_fmt[kMaxStreams] = {};
    bool slot_reserved[kMaxStreams] = {};

    // scheduler-thread state
    DevicePool pool;
    hipStream_t batch = nullptr, upload = nullptr, download = nullptr;
    std::vector<std::shared_ptr<Req>> active;           // admitted one-offs, admission order
    std::vector<std::shared_ptr<Req>> waiting_done;     // one-offs whose completion event is recorded
    Slot slots[kMaxStreams];
    Step steps[kRing];
    int step_next = 0;
    void* ws_dev[kRing] = {};
    void* ws_host[kRing] = {};
    void* ws_host_dev[kRing] = {};   // ws_host as the device sees it
    size_t ws_bytes = 0;

    // statistics (under mtx)
    std::vector<double> lat_frame, lat_oneoff;
    long failed = 0, nsteps = 0, replays = 0;
    double out_mpix = 0.0;
    bool have_first = false;
    Clock::time_point first_submit, last_done;

    void run();
    void admit_locked(std::vector<std::shared_ptr<Req>>& frames, std::vector<std::shared_ptr<Req>>& oneoffs);
    int start_frame(const std::shared_ptr<Req>& q);
    int start_oneoff(const std::shared_ptr<Req>& q);
    bool build_step();
    bool build_layered_step(Step& step);
    hipError_t launch_finish(const Req& q);
    void poll();
    void start_frames();
    void complete(const std::shared_ptr<Req>& q, int err, const std::string& msg);
    void free_req(Req& q);
    int alloc(Plane& p, int w, int h, int ch, int es = 1);
    int setup_slot(Slot& s, const std::shared_ptr<Req>& q, const Chroma& c);
    void reset_slot(Slot& s);
    int alloc_ws(Plane& p, int w, int h);
    hipError_t pass_launch(int precision_cfg, int T, const Plane& a, const Plane& b, const Plane& ws, hipStream_t st) const;
    bool busy_work() const;
};

namespace {

// device bytes a request needs (frames of an existing stream need none)
size_t request_bytes(const ACHipRequest& r, const ACHipFrameFormat& fmt, int type, int passes, size_t ws)
{
    const size_t W = r.width, H = r.height, f = static_cast<size_t>(1) << passes, es = type & 0xff;
    size_t b = ws ? round_block(ws) : 0;
    for (int k = 0; k <= passes; k++) b += round_block((W << k) * (H << k) * es);
    if (r.kind == ACHIP_REQ_RGB) {
        b += round_block(3 * W * H) + round_block(2 * W * H) + round_block(3 * f * W * f * H);
    } else {
        const Chroma c = chroma_of(fmt, r.width, r.height);
        const size_t cw = c.w, chh = c.h, cc = c.ch;
        b += c.n * (round_block(cw * chh * cc * es) + round_block(f * cw * f * chh * cc * es));
    }
    return b;
}

// p itself when [p, p + bytes) is page-locked host memory that the device sees at the same
// addresses (both ends checked), else nullptr (pageable or device memory, or another mapping).
void* host_mapping(const void* p, size_t bytes)
{
    const char* ends[2] = { static_cast<const char*>(p), static_cast<const char*>(p) + bytes - 1 };
    for (const char* q : ends) {
        hipPointerAttribute_t a{};
        if (hipPointerGetAttributes(&a, q) != hipSuccess) {
            (void)hipGetLastError();   // pageable memory the runtime has never seen
            return nullptr;
        }
        if (a.type != hipMemoryTypeHost || a.devicePointer != q) return nullptr;
    }
    return const_cast<void*>(p);
}

// With `by_kernel`, page-locked host planes move by a kernel through their device mapping, on
// the stream like any other launch.  Frames use it: the runtime's copy calls onto the frame
// streams blocked the scheduler thread for 6-7 ms at a time, now and then, with the device
// idle meanwhile (one call in several hundred; none of the one-off uploads did).  One-offs
// keep the copy engines, which move their large planes without holding CUs.  Pageable and
// device planes always take the runtime's copies.
int copy_plane(void* dst, long dstride, const void* src, long sstride, long row_bytes, int rows, hipStream_t st,
               bool by_kernel)
{
    if (rows <= 0 || row_bytes <= 0) return ACHIP_OK;
    const void* ms = by_kernel ? host_mapping(src, static_cast<size_t>(rows - 1) * sstride + row_bytes) : nullptr;
    void* md = by_kernel && !ms ? host_mapping(dst, static_cast<size_t>(rows - 1) * dstride + row_bytes) : nullptr;
    if (ms || md)
        return achip::launch_copy_rows(ms ? ms : src, sstride, md ? md : dst, dstride, row_bytes, rows, st) == hipSuccess
                   ? ACHIP_OK : ACHIP_EIO;
    hipError_t e;
    if (dstride == row_bytes && sstride == row_bytes)   // dense planes: one linear copy
        e = hipMemcpyAsync(dst, src, static_cast<size_t>(row_bytes) * rows, hipMemcpyDefault, st);
    else
        e = hipMemcpy2DAsync(dst, dstride, src, sstride, row_bytes, rows, hipMemcpyDefault, st);
    return e == hipSuccess ? ACHIP_OK : ACHIP_EIO;
}

} // namespace

int ACHipServer::alloc(Plane& p, int w, int h, int ch, int es)
{
    p.w = w;
    p.h = h;
    p.ch = ch;
    p.es = es;
    p.stride = p.row();
    return pool.get(p.bytes(), &p.ptr) ? ACHIP_OK : ACHIP_ENOMEM;
}

int ACHipServer::alloc_ws(Plane& p, int w, int h)
{
    const size_t bytes = achip_net_workspace_bytes(net, w, h);
    p.w = p.h = 0;
    p.stride = 0;
    return pool.get(bytes, &p.ptr) ? ACHIP_OK : ACHIP_ENOMEM;
}

// One 2x pass of the server's net from plane a into plane b: the fused ACNetLegacy kernel
// (configuration pcfg), or the layer launches of a layer-by-layer net with its feature maps
// in ws.  Kernels only, so frames can capture it into their graphs.
hipError_t ACHipServer::pass_launch(int pcfg, int T, const Plane& a, const Plane& b, const Plane& ws, hipStream_t st) const
{
    if (net->blocks && cfg.precision != ACHIP_FP32)
        return achip::launch_acnet8_16(cfg.precision == ACHIP_BF16, net->blocks, T, T, a.ptr, a.stride, a.w, a.h, b.ptr, b.stride,
                                       net->a8wts, cfg.precision == ACHIP_BF16 ? net->a8wabf16 : net->a8wa16, ws.ptr, st);
    if (net->blocks)
        return achip::launch_acnet8(net->blocks, T, T, a.ptr, a.stride, a.w, a.h, b.ptr, b.stride, net->a8wts, net->a8wa, ws.ptr, st);
    if (net->lfamily == kNetArtCNN)
        return achip::launch_artcnn(net->lf, net->lblocks, T, T, a.ptr, a.stride, a.w, a.h, b.ptr, b.stride, net->lk, net->lb,
                                    net->lwa, ws.ptr, st);
    if (net->lfamily == kNetFSRCNNX)
        return achip::launch_fsrcnnx(net->lf, net->lblocks, T, T, a.ptr, a.stride, a.w, a.h, b.ptr, b.stride, net->lk, net->lb,
                                     net->la, net->lwa, net->lw1, ws.ptr, st);
    return achip::launch_pass(net, cfg.precision, pcfg, T, T, a.ptr, a.stride, a.w, a.h, b.ptr, b.stride, st);
}

void ACHipServer::free_req(Req& q)
{
    for (Plane* p : { &q.in, &q.luma[0], &q.luma[1], &q.luma[2], &q.luma[3], &q.uv, &q.c0[0], &q.c0[1], &q.c1[0], &q.c1[1], &q.out,
                      &q.ws }) {
        pool.put(p->ptr);
        p->ptr = nullptr;
    }
    if (q.ev_up) (void)hipEventDestroy(q.ev_up);
    if (q.own_done_event && q.ev_done) (void)hipEventDestroy(q.ev_done);
    q.ev_up = q.ev_done = nullptr;
    std::lock_guard<std::mutex> lk(mtx);
    if (q.r.stream < 0) reserved -= q.need;
    q.need = 0;
}

void ACHipServer::complete(const std::shared_ptr<Req>& q, int err, const std::string& msg)
{
    const auto now = Clock::now();
    const bool frame = q->r.stream >= 0;
    {
        std::lock_guard<std::mutex> lk(mtx);
        q->t_done = now;
        q->err = err;
        q->msg = msg;
        q->st = St::Done;
        last_done = now;
        if (err) failed++;
        else {
            (frame ? lat_frame : lat_oneoff).push_back(ms_between(q->t_submit, now));
            out_mpix += q->out_pixels * 1e-6;
        }
    }
    if (frame || !q->keep_output || err) {
        // frames keep no buffers of their own; one-offs copied out (or failed) free theirs
        if (!frame) free_req(*q);
        else if (q->own_done_event && q->ev_done) {
            (void)hipEventDestroy(q->ev_done);
            q->ev_done = nullptr;
        }
    }
    cv_done.notify_all();
}

void ACHipServer::admit_locked(std::vector<std::shared_ptr<Req>>& frames, std::vector<std::shared_ptr<Req>>& oneoffs)
{
    bool blocked = false;    // one-offs are admitted strictly in order; frames may pass a blocked one
    for (auto it = pending.begin(); it != pending.end();) {
        const std::shared_ptr<Req>& q = *it;
        if (q->r.stream >= 0) {
            const size_t need = slot_reserved[q->r.stream] ? 0 : q->need;
            if (need == 0 || reserved + need <= cfg.hbm_budget) {
                reserved += need;
                slot_reserved[q->r.stream] = true;
                frames.push_back(q);
                it = pending.erase(it);
                continue;
            }
        } else if (!blocked) {
            if (reserved + q->need <= cfg.hbm_budget) {
                reserved += q->need;
                oneoffs.push_back(q);
                it = pending.erase(it);
                continue;
            }
            blocked = true;
        }
        ++it;
    }
}

// Stream, buffers and the captured graph of a stream's first frame.  The slot's geometry
// (s.w, s.h) is committed only once the graph is instantiated.
int ACHipServer::setup_slot(Slot& s, const std::shared_ptr<Req>& q, const Chroma& c)
{
    const ACHipRequest& r = q->r;
    const ACHipFrameFormat& f = q->fmt;
    const int T = q->type, es = T & 0xff;
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) {
        s.stream = nullptr;
        return ACHIP_EIO;
    }
    int e = alloc(s.y0, r.width, r.height, 1, es);
    if (!e) e = alloc(s.y1, 2 * r.width, 2 * r.height, 1, es);
    for (int i = 0; i < c.n && !e; i++) {
        e = alloc(s.c0[i], c.w, c.h, c.ch, es);
        if (!e) e = alloc(s.c1[i], 2 * c.w, 2 * c.h, c.ch, es);
    }
    if (!e && layered) e = alloc_ws(s.ws, r.width, r.height);
    if (e) return e;
    const int w = r.width, h = r.height;
    const int pcfg = achip::auto_config(net, cfg.precision, T, T, w, h);
    if (hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return ACHIP_EIO;
    hipError_t he = hipSuccess;
    if (f.shift) he = achip::launch_shift(T, 1, f.shift, s.y0.ptr, s.y0.stride, s.y0.ptr, s.y0.stride, w, h, s.stream);
    if (he == hipSuccess) he = pass_launch(pcfg, T, s.y0, s.y1, s.ws, s.stream);
    if (he == hipSuccess && f.shift)
        he = achip::launch_shift(T, 0, f.shift, s.y1.ptr, s.y1.stride, s.y1.ptr, s.y1.stride, s.y1.w, s.y1.h, s.stream);
    for (int i = 0; i < c.n && he == hipSuccess; i++)
        he = achip::launch_resize(T, c.ch, chroma_mode(f), s.c0[i].ptr, s.c0[i].stride, s.c0[i].w, s.c0[i].h, s.c1[i].ptr,
                                  s.c1[i].stride, s.c1[i].w, s.c1[i].h, s.stream);
    const hipError_t ce = hipStreamEndCapture(s.stream, &s.graph);
    if (ce != hipSuccess) s.graph = nullptr;
    if (he != hipSuccess || ce != hipSuccess) return ACHIP_EIO;
    if (hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0) != hipSuccess) {
        s.exec = nullptr;
        return ACHIP_EIO;
    }
    s.w = w;
    s.h = h;
    return ACHIP_OK;
}

// Everything a slot holds, returned; nothing of its own is in flight when this runs (its
// first frame failed before any launch, or the server is closing).
void ACHipServer::reset_slot(Slot& s)
{
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.exec) (void)hipGraphExecDestroy(s.exec);
    if (s.graph) (void)hipGraphDestroy(s.graph);
    for (Plane* p : { &s.y0, &s.c0[0], &s.c0[1], &s.y1, &s.c1[0], &s.c1[1], &s.ws }) {
        pool.put(p->ptr);
        *p = Plane{};
    }
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s.stream = nullptr;
    s.exec = nullptr;
    s.graph = nullptr;
    s.w = s.h = 0;
}

int ACHipServer::start_frame(const std::shared_ptr<Req>& q)
{
    const ACHipRequest& r = q->r;
    const Chroma c = chroma_of(q->fmt, r.width, r.height);
    Slot& s = slots[r.stream];
    if (!s.w) {    // first frame of this stream: buffers, stream and graph
        const int e = setup_slot(s, q, c);
        if (e) {
            reset_slot(s);   // the next frame of this stream starts again from nothing
            return e;
        }
    }
    int e = copy_plane(s.y0.ptr, s.y0.stride, r.src[0], r.src_stride[0], s.y0.row(), s.h, s.stream, true);
    for (int i = 0; i < c.n && !e; i++)
        e = copy_plane(s.c0[i].ptr, s.c0[i].stride, r.src[1 + i], r.src_stride[1 + i], s.c0[i].row(), s.c0[i].h, s.stream, true);
    if (e) return e;
    if (hipGraphLaunch(s.exec, s.stream) != hipSuccess) return ACHIP_EIO;
    {
        std::lock_guard<std::mutex> lk(mtx);
        replays++;
    }
    if (r.dst[0]) {
        e = copy_plane(r.dst[0], r.dst_stride[0], s.y1.ptr, s.y1.stride, s.y1.row(), s.y1.h, s.stream, true);
        for (int i = 0; i < c.n && !e; i++)
            e = copy_plane(r.dst[1 + i], r.dst_stride[1 + i], s.c1[i].ptr, s.c1[i].stride, s.c1[i].row(), s.c1[i].h, s.stream, true);
        if (e) return e;
    }
    if (hipEventCreateWithFlags(&q->ev_done, hipEventDisableTiming) != hipSuccess) return ACHIP_EIO;
    q->own_done_event = true;
    if (hipEventRecord(q->ev_done, s.stream) != hipSuccess) return ACHIP_EIO;
    q->out_pixels = 4LL * s.w * s.h;
    s.inflight.push_back(q);
    return ACHIP_OK;
}

int ACHipServer::start_oneoff(const std::shared_ptr<Req>& q)
{
    const ACHipRequest& r = q->r;
    const int W = r.width, H = r.height, P = q->passes, f = 1 << P, es = q->type & 0xff;
    int e = ACHIP_OK;
    for (int k = 0; k <= P && !e; k++) e = alloc(q->luma[k], W << k, H << k, 1, es);
    const Chroma c = r.kind == ACHIP_REQ_RGB ? Chroma{} : chroma_of(q->fmt, W, H);
    if (!e && r.kind == ACHIP_REQ_RGB) {
        e = alloc(q->in, W, H, 3);
        if (!e) e = alloc(q->uv, W, H, 2);
        if (!e) e = alloc(q->out, f * W, f * H, 3);
    }
    for (int i = 0; i < c.n && !e; i++) {
        e = alloc(q->c0[i], c.w, c.h, c.ch, es);
        if (!e) e = alloc(q->c1[i], f * c.w, f * c.h, c.ch, es);
    }
    if (!e && layered) e = alloc_ws(q->ws, W << (P - 1), H << (P - 1));   // the largest pass
    if (e) return e;
    if (r.kind == ACHIP_REQ_RGB) e = copy_plane(q->in.ptr, q->in.stride, r.src[0], r.src_stride[0], q->in.row(), H, upload, false);
    else e = copy_plane(q->luma[0].ptr, q->luma[0].stride, r.src[0], r.src_stride[0], q->luma[0].row(), H, upload, false);
    for (int i = 0; i < c.n && !e; i++)
        e = copy_plane(q->c0[i].ptr, q->c0[i].stride, r.src[1 + i], r.src_stride[1 + i], q->c0[i].row(), c.h, upload, false);
    if (e) return e;
    if (hipEventCreateWithFlags(&q->ev_up, hipEventDisableTiming) != hipSuccess) return ACHIP_EIO;
    if (hipEventRecord(q->ev_up, upload) != hipSuccess) return ACHIP_EIO;
    q->out_pixels = static_cast<long long>(f) * W * f * H;
    q->st = St::Uploading;
    active.push_back(q);
    return ACHIP_OK;
}

// Gather tile-row ranges of the admitted one-offs' current passes into one batched launch,
// then the finishing kernels of the requests whose last pass this step completes.
bool ACHipServer::build_step()
{
    Step& step = steps[step_next];
    if (step.busy) return false;
    // at most kRing - 2 steps in flight: the one before must not be two steps behind
    int inflight = 0;
    for (const Step& s : steps) inflight += s.busy ? 1 : 0;
    if (inflight >= kRing - 2) return false;

    if (layered) return build_layered_step(step);

    // descriptors of u8 passes first, then u16 passes: one persistent launch per element
    // type, each with its own tile counter (bytes 0 and 64 of the header) and tile numbering
    auto* hdr = static_cast<char*>(ws_host[step_next]);
    auto* desc = reinterpret_cast<achip::BatchDesc*>(hdr + 256);
    std::vector<achip::BatchDesc> d16;
    int n = 0;
    long total = 0, total16 = 0;
    long budget = cfg.step_tiles;
    std::vector<std::shared_ptr<Req>> finishing, members;
    for (const std::shared_ptr<Req>& q : active) {
        if (budget <= 0 || n + static_cast<int>(d16.size()) >= kMaxStepItems) break;
        if (q->st != St::Passing) continue;
        members.push_back(q);
        const Plane& a = q->luma[q->pass];
        const Plane& b = q->luma[q->pass + 1];
        const int tiles_x = (a.w + tw - 1) / tw, tiles_y = (a.h + th - 1) / th;
        const int rows = std::min<long>(tiles_y - q->next_row, std::max<long>(1, budget / tiles_x));
        long& tot = q->type == kU16 ? total16 : total;
        const achip::BatchDesc d{ a.ptr, b.ptr, a.stride, b.stride, a.w, a.h, tiles_x, static_cast<int>(tot), q->next_row };
        if (q->type == kU16) d16.push_back(d);
        else desc[n++] = d;
        tot += static_cast<long>(rows) * tiles_x;
        budget -= static_cast<long>(rows) * tiles_x;
        q->next_row += rows;
        if (q->next_row == tiles_y) {
            q->pass++;
            q->next_row = 0;
            if (q->pass == q->passes) {
                q->st = St::Finishing;
                finishing.push_back(q);
            }
        }
    }
    const int n16 = static_cast<int>(d16.size());
    if (n + n16 == 0) return false;
    std::memcpy(desc + n, d16.data(), static_cast<size_t>(n16) * sizeof(achip::BatchDesc));
    std::memset(hdr, 0, 256);   // tile counters
    const size_t bytes = 256 + static_cast<size_t>(n + n16) * sizeof(achip::BatchDesc);
    char* ws = static_cast<char*>(ws_dev[step_next]);
    const auto* ddesc = reinterpret_cast<const achip::BatchDesc*>(ws + 256);
    hipError_t e = achip::launch_copy_rows(ws_host_dev[step_next], static_cast<long>(bytes), ws, static_cast<long>(bytes),
                                           static_cast<long>(bytes), 1, batch);
    if (e == hipSuccess && n)
        e = achip::launch_batch(net, cfg.precision, config, kU8, kU8, ddesc, n, static_cast<int>(total), reinterpret_cast<int*>(ws),
                                cfg.max_workgroups, batch);
    if (e == hipSuccess && n16)
        e = achip::launch_batch(net, cfg.precision, config, kU16, kU16, ddesc + n, n16, static_cast<int>(total16),
                                reinterpret_cast<int*>(ws + 64), cfg.max_workgroups, batch);
    for (const std::shared_ptr<Req>& q : finishing) {
        if (e != hipSuccess) break;
        e = launch_finish(*q);
    }
    if (e == hipSuccess) e = hipEventRecord(step.ev, batch);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(batch);   // nothing of this step may still use the buffers
        for (const std::shared_ptr<Req>& q : members) complete(q, ACHIP_EIO, "step launch failed");
        active.erase(std::remove_if(active.begin(), active.end(), [](const std::shared_ptr<Req>& q) { return q->st == St::Done; }),
                     active.end());
        return false;
    }
    step.busy = true;
    step.finishing = std::move(finishing);
    step_next = (step_next + 1) % kRing;
    std::lock_guard<std::mutex> lk(mtx);
    nsteps++;
    return true;
}

// The kernels after a one-off's last pass: chroma resize fused with the merge (RGB), or
// the LSB shift and the chroma resizes (luma / YUV).
hipError_t ACHipServer::launch_finish(const Req& q)
{
    const int P = q.passes, T = q.type;
    const Plane& y = q.luma[P];
    if (q.r.kind == ACHIP_REQ_RGB)
        return achip::launch_resize_merge(kU8, 3, achip::kResizeCatmullRom, y.ptr, y.stride, q.uv.ptr, q.uv.stride, q.uv.w,
                                          q.uv.h, y.w, y.h, q.out.ptr, q.out.stride, batch);
    hipError_t e = hipSuccess;
    if (q.fmt.shift) e = achip::launch_shift(T, 0, q.fmt.shift, y.ptr, y.stride, y.ptr, y.stride, y.w, y.h, batch);
    const Chroma c = chroma_of(q.fmt, q.r.width, q.r.height);
    for (int i = 0; i < c.n && e == hipSuccess; i++)
        e = achip::launch_resize(T, c.ch, chroma_mode(q.fmt), q.c0[i].ptr, q.c0[i].stride, q.c0[i].w, q.c0[i].h, q.c1[i].ptr,
                                 q.c1[i].stride, q.c1[i].w, q.c1[i].h, batch);
    return e;
}

// A step of a layer-by-layer net: the admitted one-offs' next passes, each whole (its layer
// launches in turn on the batch stream), in admission order until the step's pixel budget
// (step_tiles x the tile area) is spent; then the finishing kernels of the requests whose
// last pass this step ran.
bool ACHipServer::build_layered_step(Step& step)
{
    long budget = static_cast<long>(cfg.step_tiles) * tw * th;
    std::vector<std::shared_ptr<Req>> finishing, members;
    hipError_t e = hipSuccess;
    for (const std::shared_ptr<Req>& q : active) {
        if (budget <= 0 || e != hipSuccess) break;
        if (q->st != St::Passing) continue;
        members.push_back(q);
        const Plane& a = q->luma[q->pass];
        e = pass_launch(0, q->type, a, q->luma[q->pass + 1], q->ws, batch);
        budget -= static_cast<long>(a.w) * a.h;
        if (++q->pass == q->passes) {
            q->st = St::Finishing;
            finishing.push_back(q);
        }
    }
    if (members.empty()) return false;
    for (const std::shared_ptr<Req>& q : finishing) {
        if (e != hipSuccess) break;
        e = launch_finish(*q);
    }
    if (e == hipSuccess) e = hipEventRecord(step.ev, batch);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(batch);   // nothing of this step may still use the buffers
        for (const std::shared_ptr<Req>& q : members) complete(q, ACHIP_EIO, "step launch failed");
        active.erase(std::remove_if(active.begin(), active.end(), [](const std::shared_ptr<Req>& q) { return q->st == St::Done; }),
                     active.end());
        return false;
    }
    step.busy = true;
    step.finishing = std::move(finishing);
    step_next = (step_next + 1) % kRing;
    std::lock_guard<std::mutex> lk(mtx);
    nsteps++;
    return true;
}

void ACHipServer::poll()
{
    // uploads finished: colour split (RGB) and into the pass phase
    for (const std::shared_ptr<Req>& q : active) {
        if (q->st != St::Uploading || hipEventQuery(q->ev_up) != hipSuccess) continue;
        if (q->r.kind == ACHIP_REQ_RGB &&
            achip::launch_color_split(kU8, 3, q->in.ptr, q->in.stride, q->in.w, q->in.h, q->luma[0].ptr, q->luma[0].stride,
                                      q->uv.ptr, q->uv.stride, batch) != hipSuccess) {
            complete(q, ACHIP_EIO, "colour split launch failed");
            continue;
        }
        const Plane& y0 = q->luma[0];
        if (q->r.kind != ACHIP_REQ_RGB && q->fmt.shift &&
            achip::launch_shift(q->type, 1, q->fmt.shift, y0.ptr, y0.stride, y0.ptr, y0.stride, y0.w, y0.h, batch) != hipSuccess) {
            complete(q, ACHIP_EIO, "luma shift launch failed");
            continue;
        }
        q->st = St::Passing;
    }
    // steps finished: their finishing requests are done (or copy out)
    for (Step& s : steps) {
        if (!s.busy || hipEventQuery(s.ev) != hipSuccess) continue;
        s.busy = false;
        for (const std::shared_ptr<Req>& q : s.finishing) {
            const ACHipRequest& r = q->r;
            if (!r.dst[0]) {
                q->keep_output = true;
                complete(q, ACHIP_OK, "");
                continue;
            }
            int e = ACHIP_OK;
            const Plane& y = q->luma[q->passes];
            if (r.kind == ACHIP_REQ_RGB) e = copy_plane(r.dst[0], r.dst_stride[0], q->out.ptr, q->out.stride, q->out.row(), q->out.h, download, false);
            else e = copy_plane(r.dst[0], r.dst_stride[0], y.ptr, y.stride, y.row(), y.h, download, false);
            const int nc = r.kind == ACHIP_REQ_RGB ? 0 : chroma_of(q->fmt, r.width, r.height).n;
            for (int i = 0; i < nc && !e; i++)
                e = copy_plane(r.dst[1 + i], r.dst_stride[1 + i], q->c1[i].ptr, q->c1[i].stride, q->c1[i].row(), q->c1[i].h, download, false);
            if (!e && hipEventCreateWithFlags(&q->ev_done, hipEventDisableTiming) != hipSuccess) e = ACHIP_EIO;
            if (!e) {
                q->own_done_event = true;
                if (hipEventRecord(q->ev_done, download) != hipSuccess) e = ACHIP_EIO;
            }
            if (e) complete(q, e, "output copy failed");
            else waiting_done.push_back(q);
        }
        s.finishing.clear();
    }
    // copies out finished
    for (auto it = waiting_done.begin(); it != waiting_done.end();) {
        if (hipEventQuery((*it)->ev_done) == hipSuccess) {
            complete(*it, ACHIP_OK, "");
            it = waiting_done.erase(it);
        } else {
            ++it;
        }
    }
    active.erase(std::remove_if(active.begin(), active.end(), [](const std::shared_ptr<Req>& q) { return q->st == St::Done; }),
                 active.end());
    // frames, in order per stream
    for (Slot& s : slots) {
        while (!s.inflight.empty() && hipEventQuery(s.inflight.front()->ev_done) == hipSuccess) {
            complete(s.inflight.front(), ACHIP_OK, "");
            s.inflight.pop_front();
        }
    }
}

bool ACHipServer::busy_work() const
{
    if (!active.empty() || !waiting_done.empty()) return true;
    for (const Step& s : steps)
        if (s.busy) return true;
    for (const Slot& s : slots)
        if (!s.inflight.empty() || !s.waiting.empty()) return true;
    return false;
}

// the next frame of every stream whose previous frame has completed
void ACHipServer::start_frames()
{
    for (Slot& s : slots) {
        while (s.inflight.empty() && !s.waiting.empty()) {
            const std::shared_ptr<Req> q = s.waiting.front();
            s.waiting.pop_front();
            const int e = start_frame(q);
            if (e) complete(q, e, "frame launch failed");
        }
    }
}

void ACHipServer::run()
{
    (void)hipSetDevice(net->device);
    for (;;) {
        std::vector<std::shared_ptr<Req>> frames, oneoffs, freed, refused;
        {
            std::unique_lock<std::mutex> lk(mtx);
            if (!busy_work() && pending.empty() && released.empty()) {
                if (stop) break;
                cv_work.wait(lk, [&] { return stop || !pending.empty() || !released.empty(); });
            }
            freed.swap(released);
            admit_locked(frames, oneoffs);
            // shutting down with nothing in flight: what still cannot be admitted never will be
            if (stop && frames.empty() && oneoffs.empty() && freed.empty() && !busy_work()) {
                refused.assign(pending.begin(), pending.end());
                pending.clear();
            }
        }
        for (const std::shared_ptr<Req>& q : freed) free_req(*q);
        for (const std::shared_ptr<Req>& q : refused) complete(q, ACHIP_ENOMEM, "not admitted before shutdown");
        for (const std::shared_ptr<Req>& q : frames) slots[q->r.stream].waiting.push_back(q);
        start_frames();
        // steps are built between one-off set-ups, so the passes of the first requests run
        // while the uploads of later ones are still being issued
        bool launched = false;
        for (const std::shared_ptr<Req>& q : oneoffs) {
            const int e = start_oneoff(q);
            if (e) complete(q, e, "request set-up failed");
            poll();
            start_frames();
            while (build_step()) launched = true;
        }
        poll();
        start_frames();
        while (build_step()) launched = true;
        if (!launched && busy_work()) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

extern "C" {

int achip_server_create(ACHipNet* net, const ACHipServerConfig* config, ACHipServer** out)
{
    if (!net || !out) return ACHIP_EINVAL;
    *out = nullptr;
    ACHipServerConfig c = config ? *config : ACHipServerConfig{};
    if (c.precision != ACHIP_FP32 && c.precision != ACHIP_FP16 && c.precision != ACHIP_BF16) return ACHIP_EINVAL;
    const bool layered = net->blocks || net->lfamily;
    if (net->lfamily && c.precision != ACHIP_FP32) return ACHIP_EINVAL;   // ArtCNN and FSRCNNX run in FP32 only
    if (c.hbm_budget == 0) c.hbm_budget = static_cast<size_t>(16) << 30;
    if (c.step_tiles <= 0) c.step_tiles = c.precision == ACHIP_FP32 ? 2048 : 12288;   // 4.2 / 25.2 Mpix
    if (c.max_workgroups < 0) return ACHIP_EINVAL;
    auto s = std::make_unique<ACHipServer>();
    s->net = net;
    s->cfg = c;
    s->pool.set_chunk_cap(c.hbm_budget / 4);
    s->layered = layered;
    s->config = 0;
    if (achip_config_info(c.precision, s->config, &s->tw, &s->th, nullptr, nullptr, nullptr) != ACHIP_OK) return ACHIP_EINVAL;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ACHIP_EIO;
    if (hipSetDevice(net->device) != hipSuccess) return ACHIP_ENODEV;
    bool ok = hipStreamCreateWithFlags(&s->batch, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&s->download, hipStreamNonBlocking) == hipSuccess;
    s->ws_bytes = 256 + static_cast<size_t>(kMaxStepItems) * sizeof(achip::BatchDesc);
    for (int i = 0; i < kRing && ok; i++) {
        ok = hipMalloc(&s->ws_dev[i], s->ws_bytes) == hipSuccess && hipHostMalloc(&s->ws_host[i], s->ws_bytes, 0) == hipSuccess &&
             hipHostGetDevicePointer(&s->ws_host_dev[i], s->ws_host[i], 0) == hipSuccess &&
             hipEventCreateWithFlags(&s->steps[i].ev, hipEventDisableTiming) == hipSuccess;
    }
    (void)hipSetDevice(dev);
    if (!ok) {
        (void)hipGetLastError();
        achip_server_destroy(s.release());
        return ACHIP_EIO;
    }
    ACHipServer* raw = s.release();
    raw->thread = std::thread([raw] { raw->run(); });
    *out = raw;
    return ACHIP_OK;
}

void achip_server_destroy(ACHipServer* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mtx);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->thread.joinable()) s->thread.join();
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(s->net->device);
    for (auto* st : { s->batch, s->upload, s->download })
        if (st) (void)hipStreamSynchronize(st);
    for (Slot& sl : s->slots) s->reset_slot(sl);
    for (auto& kv : s->tickets) s->free_req(*kv.second);
    s->tickets.clear();
    for (int i = 0; i < kRing; i++) {
        if (s->ws_dev[i]) (void)hipFree(s->ws_dev[i]);
        if (s->ws_host[i]) (void)hipHostFree(s->ws_host[i]);
        if (s->steps[i].ev) (void)hipEventDestroy(s->steps[i].ev);
    }
    for (auto* st : { s->batch, s->upload, s->download })
        if (st) (void)hipStreamDestroy(st);
    s->pool.trim();
    (void)hipSetDevice(dev);
    delete s;
}

int achip_server_submit(ACHipServer* s, const ACHipRequest* r, unsigned long long* ticket)
{
    if (!s || !r || !ticket) return ACHIP_EINVAL;
    if (r->kind < ACHIP_REQ_LUMA || r->kind > ACHIP_REQ_YUV) return ACHIP_EINVAL;
    if (r->width <= 0 || r->height <= 0 || r->width > (1 << 15) || r->height > (1 << 15)) return ACHIP_EINVAL;
    int passes = 0;
    switch (r->factor) {
    case 2: passes = 1; break;
    case 4: passes = 2; break;
    case 8: passes = 3; break;
    default: return ACHIP_EINVAL;
    }
    if ((static_cast<long>(r->width) << passes) > (1 << 16) || (static_cast<long>(r->height) << passes) > (1 << 16))
        return ACHIP_EINVAL;
    // the frame format: LUMA and YUV420 are shorthands of the general kind
    ACHipFrameFormat f{};
    if (r->kind == ACHIP_REQ_LUMA) f = ACHipFrameFormat{ 1, 0, 0, kU8, 0, ACHIP_CHROMA_BILINEAR };
    else if (r->kind == ACHIP_REQ_YUV420) f = ACHipFrameFormat{ 3, 1, 1, kU8, 0, ACHIP_CHROMA_BILINEAR };
    else if (r->kind == ACHIP_REQ_YUV) f = r->format;
    const bool rgb = r->kind == ACHIP_REQ_RGB;
    const int type = rgb ? kU8 : f.element_type;
    if (!rgb) {
        if (f.planes < 1 || f.planes > 3) return ACHIP_EINVAL;
        if (f.element_type != kU8 && f.element_type != kU16) return ACHIP_EINVAL;
        if (f.shift < 0 || f.shift > (f.element_type == kU8 ? 0 : 15)) return ACHIP_EINVAL;
        if (f.planes == 1) f.chroma_shift_x = f.chroma_shift_y = 0, f.chroma_filter = ACHIP_CHROMA_BILINEAR;
        if (f.chroma_shift_x < 0 || f.chroma_shift_x > 1 || f.chroma_shift_y < 0 || f.chroma_shift_y > 1) return ACHIP_EINVAL;
        if (f.chroma_filter != ACHIP_CHROMA_BILINEAR && f.chroma_filter != ACHIP_CHROMA_CATMULL_ROM) return ACHIP_EINVAL;
        if ((r->width & ((1 << f.chroma_shift_x) - 1)) || (r->height & ((1 << f.chroma_shift_y) - 1))) return ACHIP_EINVAL;
    }
    const int planes = rgb ? 1 : f.planes;
    const long es = type & 0xff;
    const Chroma c = rgb ? Chroma{} : chroma_of(f, r->width, r->height);
    auto row_of = [&](int p) { return p == 0 ? static_cast<long>(r->width) * (rgb ? 3 : 1) * es : static_cast<long>(c.w) * c.ch * es; };
    for (int p = 0; p < planes; p++)
        if (!r->src[p] || r->src_stride[p] < row_of(p)) return ACHIP_EINVAL;
    const bool has_dst = r->dst[0] != nullptr;
    for (int p = 0; p < planes; p++) {
        if (has_dst != (r->dst[p] != nullptr)) return ACHIP_EINVAL;
        if (has_dst && r->dst_stride[p] < (row_of(p) << passes)) return ACHIP_EINVAL;
    }
    if (r->stream >= kMaxStreams || r->stream < -1) return ACHIP_EINVAL;
    if (r->stream >= 0 && (rgb || r->factor != 2)) return ACHIP_EINVAL;
    auto q = std::make_shared<Req>();
    q->r = *r;
    q->fmt = f;
    q->type = type;
    q->passes = passes;
    const bool layered = s->net->blocks || s->net->lfamily;
    const size_t ws = layered ? achip_net_workspace_bytes(s->net, r->width << (passes - 1), r->height << (passes - 1)) : 0;
    q->need = r->stream >= 0 ? request_bytes(*r, f, type, 1, layered ? achip_net_workspace_bytes(s->net, r->width, r->height) : 0)
                             : request_bytes(*r, f, type, passes, ws);
    if (q->need > s->cfg.hbm_budget) return ACHIP_ENOMEM;
    std::lock_guard<std::mutex> lk(s->mtx);
    if (s->stop) return ACHIP_EIO;
    if (r->stream >= 0) {   // one geometry and format per stream
        int& sw = s->stream_w[r->stream];
        int& sh = s->stream_h[r->stream];
        ACHipFrameFormat& sf = s->stream_fmt[r->stream];
        if (sw && (sw != r->width || sh != r->height || !same_format(sf, f))) return ACHIP_EINVAL;
        sw = r->width;
        sh = r->height;
        sf = f;
    }
    q->ticket = s->next_ticket++;
    q->t_submit = Clock::now();
    if (!s->have_first) {
        s->have_first = true;
        s->first_submit = q->t_submit;
    }
    s->tickets[q->ticket] = q;
    s->pending.push_back(q);
    *ticket = q->ticket;
    s->cv_work.notify_one();
    return ACHIP_OK;
}

int achip_server_wait(ACHipServer* s, unsigned long long ticket, double timeout_ms, double* latency_ms)
{
    if (!s) return ACHIP_EINVAL;
    std::unique_lock<std::mutex> lk(s->mtx);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return ACHIP_EINVAL;
    std::shared_ptr<Req> q = it->second;
    auto done = [&] { return q->st == St::Done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done)) return ACHIP_EIO;
    if (latency_ms) *latency_ms = ms_between(q->t_submit, q->t_done);
    const int err = q->err;
    if (!q->keep_output) s->tickets.erase(q->ticket);   // nothing left to read
    return err;
}

int achip_server_output(ACHipServer* s, unsigned long long ticket, int plane, void** ptr, long* stride)
{
    if (!s || !ptr || !stride) return ACHIP_EINVAL;
    std::lock_guard<std::mutex> lk(s->mtx);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return ACHIP_EINVAL;
    const Req& q = *it->second;
    if (q.st != St::Done || !q.keep_output || q.err) return ACHIP_EINVAL;
    const Plane* p = nullptr;
    if (q.r.kind == ACHIP_REQ_RGB) p = plane == 0 ? &q.out : nullptr;
    else if (plane == 0) p = &q.luma[q.passes];
    else if (plane > 0 && plane <= chroma_of(q.fmt, q.r.width, q.r.height).n) p = &q.c1[plane - 1];
    if (!p) return ACHIP_EINVAL;
    *ptr = p->ptr;
    *stride = p->stride;
    return ACHIP_OK;
}

int achip_server_release(ACHipServer* s, unsigned long long ticket)
{
    if (!s) return ACHIP_EINVAL;
    std::shared_ptr<Req> q;
    {
        std::lock_guard<std::mutex> lk(s->mtx);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end()) return ACHIP_EINVAL;
        if (it->second->st != St::Done) return ACHIP_EINVAL;
        q = it->second;
        s->tickets.erase(it);
    }
    // buffers go back to the pool on the scheduler thread, which owns it
    if (q->keep_output) {
        std::lock_guard<std::mutex> lk(s->mtx);
        s->released.push_back(q);
    }
    s->cv_work.notify_one();
    return ACHIP_OK;
}

int achip_server_stats(ACHipServer* s, ACHipServerStats* out, int reset)
{
    if (!s || !out) return ACHIP_EINVAL;
    std::lock_guard<std::mutex> lk(s->mtx);
    out->completed_frames = static_cast<long>(s->lat_frame.size());
    out->completed_oneoffs = static_cast<long>(s->lat_oneoff.size());
    out->failed = s->failed;
    out->frame_p50_ms = percentile(s->lat_frame, 0.50);
    out->frame_p99_ms = percentile(s->lat_frame, 0.99);
    out->oneoff_p50_ms = percentile(s->lat_oneoff, 0.50);
    out->oneoff_p99_ms = percentile(s->lat_oneoff, 0.99);
    out->output_mpix = s->out_mpix;
    out->busy_s = s->have_first ? std::chrono::duration<double>(s->last_done - s->first_submit).count() : 0.0;
    out->steps = s->nsteps;
    out->graph_replays = s->replays;
    if (reset) {
        s->lat_frame.clear();
        s->lat_oneoff.clear();
        s->failed = 0;
        s->out_mpix = 0.0;
        s->nsteps = 0;
        s->replays = 0;
        s->have_first = false;
    }
    return ACHIP_OK;
}

} // extern "C"
