// engine.h — the wavefront scheduler shared by wr::Renderer and
// wr::RenderServer. One path pool holds paths of several requests (up to
// WRK_MAX_SLOTS admitted chunks, up to 4 resident scenes); every iteration
// traces, classifies and shades each scene's share of the merged queue and
// rebuilds the queue with one ordered partition. Each admitted chunk holds a
// contiguous pool range (queue-ordered state accesses stay coalesced) until
// its last path ends; new chunks fill freed ranges while the deep paths of
// other chunks finish. Replaces the host loop of
// libWR/src/internal/renderer.cpp:190-284.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "../../include/wr/wr.h"
#include "../../include/wrk.h"

namespace wr {
namespace detail {

void hip_check(hipError_t e, const char* what);
void wrk_check(int rc, const char* what);

// Owning device allocation.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; bytes = o.bytes; o.ptr = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release();
    void alloc(size_t n);
    void upload(const void* src, size_t n);
    uint64_t addr() const { return (uint64_t)(uintptr_t)ptr; }
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

wrk_camera camera_record(const Camera& c);

struct EngineConfig {
    uint32_t capacity = 1u << 26;       // paths in flight
    uint32_t maxChunkPaths = 0;          // 0 = capacity
    uint32_t minChunkPaths = 0;          // 0 = max(capacity / 32, 65536)
    uint64_t memoryBudgetBytes = 0;      // 0 = unlimited (beyond the pool itself)
    bool timeKernels = false;
    bool countTraversal = false;
    bool shortestFirst = false;          // admission order: fewest remaining paths first
    bool roundRobin = false;             // one chunk per pending request per round (else the first takes all that fit)
    uint64_t* sharedFootprint = nullptr; // budget counter shared with another engine (nullptr = this engine's own)
    // requests of at most this many paths render each admitted chunk to
    // completion in one fused launch (wrk_render_fused) instead of joining
    // the bounce iterations (0 = never)
    uint64_t fusedMaxPaths = 0;
    // in-order admission: while the oldest unfinished request has paths in
    // flight, the requests behind it take at most this many paths of the
    // pool in all (they fill what its draining tail leaves idle without
    // stretching its iterations); 0 = no limit. Fused requests are exempt.
    uint64_t backfillPaths = 0;
};

class Engine {
public:
    enum class Mode { Accumulate, Paths };
    struct RequestState {
        uint64_t id = 0;
        Scene* scene = nullptr;
        uint32_t sceneIdx = 0;
        wrk_camera cam{};
        wrk_render_params prm{};
        Mode mode = Mode::Accumulate;
        uint64_t firstPath = 0, endPath = 0, nextPath = 0;
        uint32_t spp = 1, npix = 0;
        float* pathsOut = nullptr;  // host, Mode::Paths
        DevBuf accum, image;
        bool guides = false;        // first-hit albedo / normal sums (trace.cu:267-286)
        DevBuf accAlb, accNrm, imgAlb, imgNrm;
        DevBuf accNoise, imgNoise;  // with guides: per-pixel (sum y, sum y^2) and the standard error of mean y
        // chunks finish in any order; their results enter the accumulation
        // buffer in sample order: accumNext is the first path not yet added,
        // held maps the first path of each finished chunk waiting for its
        // predecessors to its slot
        uint64_t accumNext = 0, finishedPaths = 0;
        std::map<uint64_t, uint32_t> held;
        uint32_t chunksInFlight = 0, chunks = 0;
        bool done = false, admitted = false;
        double tSubmit = 0, tAdmit = 0, tDone = 0;
        uint64_t footprint = 0;
    };

    explicit Engine(const EngineConfig& cfg);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void setConfig(const EngineConfig& cfg);
    void ensureCapacity(uint32_t cap) { reserve(cap); }
    const EngineConfig& config() const { return cfg_; }

    // Queues a request for paths [firstPath, firstPath + numPaths) of the
    // image described by params (numPaths = W*H*spp for a whole render).
    uint64_t submit(Scene* scene, const wrk_camera& cam, const wrk_render_params& prm, uint32_t spp,
                    uint64_t firstPath, uint64_t numPaths, Mode mode, float* pathsOut, bool guides = false);
    // Admission + one wavefront iteration + completions. Returns true while
    // any request is pending or in flight.
    bool step() {
        launch();
        return finish();
    }
    // The two halves of step(): launch() enqueues the iteration, the
    // admission of new chunks and the queue rebuild on this engine's streams
    // without waiting; finish() waits for them and records completions.
    // Another engine may run whole steps in between, on its own streams.
    void launch();
    bool finish();
    // true once the launched iteration has finished on the device
    bool iterationDone() const;
    bool isDone(uint64_t id) const;
    const RequestState& request(uint64_t id) const;
    size_t takeCompleted(uint64_t* ids, size_t maxIds) {
        const size_t n = std::min(maxIds, completed_.size());
        std::copy(completed_.begin(), completed_.begin() + n, ids);
        completed_.erase(completed_.begin(), completed_.begin() + n);
        return n;
    }
    // Copies the averaged image (W*H*3 floats) of a finished Accumulate
    // request to host memory and releases the request.
    // albedo / normal (W*H*3 floats each, averaged over spp) when the
    // request was submitted with guides; NULL skips them.
    // noise (W*H floats) with guides: the standard error of each pixel's
    // mean y = (r + g + b) / 3 (wrk_noise_stderr).
    void fetch(uint64_t id, float* out, float* albedo = nullptr, float* normal = nullptr, float* noise = nullptr);
    // The averaged images of a finished request, on the device (valid until
    // fetch or release).
    const float* deviceImage(uint64_t id) const;
    const float* deviceAlbedo(uint64_t id) const;
    const float* deviceNormal(uint64_t id) const;
    const float* deviceNoise(uint64_t id) const;
    hipStream_t stream() const { return stream_; }
    void release(uint64_t id);
    size_t numPending() const { return pending_.size(); }
    bool busy() const { return !pending_.empty() || numActiveSlots() != 0; }
    size_t numActiveSlots() const;
    RenderStats& stats() { return stats_; }
    static double now();

private:
    struct Slot {
        bool active = false;
        bool finished = false;  // all paths done, results held for ordered accumulation
        uint64_t request = 0;
        uint32_t poolBase = 0, count = 0;  // pool range
        uint64_t firstPath = 0;
    };
    void reserve(uint32_t cap);
    uint32_t sceneIndex(Scene* s);
    bool allocRange(uint32_t n, uint32_t* base);
    void freeRange(uint32_t base, uint32_t n);
    uint32_t largestFree() const;
    void complete(uint32_t slot);
    void retire(uint32_t slot);
    bool admitChunk(RequestState& r, uint32_t* added, uint32_t queueBase, uint64_t maxPaths = ~0ull);
    template <typename F> void timed(int category, hipStream_t stream, F&& launch);
    uint64_t& footprintUsed() { return cfg_.sharedFootprint ? *cfg_.sharedFootprint : footprintUsed_; }
    DevBuf takeBuf(size_t bytes);
    void giveBuf(DevBuf& b);  // keeps b's memory for later requests (or frees it past the pool's limit)
    // pool ranges are whole granules of 256 entries, so one byte per granule
    // (slotMap_) names the slot that owns an entry
    static uint32_t granuleRound(uint32_t n) { return (n + 255u) & ~255u; }

    EngineConfig cfg_;
    RenderStats stats_;
    hipStream_t stream_ = nullptr;
    // one stream per resident scene: the scenes' trace/classify/shade chains
    // of an iteration touch disjoint queue parts and run concurrently
    hipStream_t sceneStream_[4] = {};
    hipEvent_t stepStart_ = nullptr, sceneDone_[4] = {};
    uint32_t capacity_ = 0;
    DevBuf org_, dir_, thr_, rad_, st_, contrib_, savedOrg_, savedDir_, hit_, occl_, result_, guides_;
    DevBuf qA_, qB_, cat_, scratch_, offs_, prmTable_, fetch_, slotMap_;
    wrk_paths paths_{};
    uint32_t deadCount_ = 0;              // entries ended last iteration: qA_[0, deadCount_)
    bool launched_ = false;               // launch() done, finish() not yet
    uint32_t launchedItems_ = 0, launchedScenes_ = 0, launchedTotal_ = 0;
    uint32_t* hostOffs_ = nullptr;        // pinned: queue offsets, then the live counts
    uint32_t* hostLive_ = nullptr;        // hostOffs_ + 16: live paths per slot, read back each step
    uint32_t* hostLiveInit_ = nullptr;    // pinned staging: initial live count per slot
    wrk_render_params* hostPrm_ = nullptr;  // pinned staging for the slot table
    std::vector<Scene*> scenes_;
    // per-scene queue parts: trace only, trace + shadow, shadow only
    std::vector<uint32_t> segT_, segTS_, segS_;
    std::map<uint32_t, uint32_t> free_;   // free pool ranges: base -> length
    Slot slots_[WRK_MAX_SLOTS];
    std::vector<uint32_t> freeSlots_;
    std::deque<uint64_t> pending_;
    std::deque<uint64_t> completed_;      // finished, not yet taken by takeCompleted()
    std::map<uint64_t, std::unique_ptr<RequestState>> requests_;
    uint64_t nextId_ = 1;
    uint64_t footprintUsed_ = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks_;
    std::vector<hipEvent_t> eventPool_;
    // freed request buffers by size, reused by later requests (at most
    // kBufPoolBytes held; 512 MB is twenty 1080p accumulation buffers)
    static constexpr size_t kBufPoolBytes = size_t(512) << 20;
    std::multimap<size_t, DevBuf> bufPool_;
    size_t pooledBytes_ = 0;
    size_t eventNext_ = 0;
};

}  // namespace detail
}  // namespace wr
