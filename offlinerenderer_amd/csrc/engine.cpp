// engine.cpp — wavefront scheduler (see engine.h).
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace wr {
namespace detail {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
void wrk_check(int rc, const char* what) {
    if (rc != WRK_OK) throw std::runtime_error(std::string(what) + ": " + wrk_last_error());
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}
void DevBuf::alloc(size_t n) {
    if (n <= bytes && ptr) return;
    release();
    if (n == 0) return;
    hip_check(hipMalloc(&ptr, n), "hipMalloc");
    bytes = n;
}
void DevBuf::upload(const void* src, size_t n) {
    alloc(n);
    if (n) hip_check(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

// renderer.cpp:144-156
wrk_camera camera_record(const Camera& c) {
    Vec3 f = c.getForward(), r = c.getRight(), u = c.getUp();
    wrk_camera k;
    std::memset(&k, 0, sizeof(k));
    k.position[0] = c.position.x; k.position[1] = c.position.y; k.position[2] = c.position.z;
    k.forward[0] = f.x; k.forward[1] = f.y; k.forward[2] = f.z;
    k.right[0] = r.x; k.right[1] = r.y; k.right[2] = r.z;
    k.up[0] = u.x; k.up[1] = u.y; k.up[2] = u.z;
    k.tanHalfFovY = std::tan(c.fovY * 0.5f);
    k.aspect = c.aspect;
    k.focalDistance = c.focalDistance;
    k.lensRadius = c.lensRadius;
    return k;
}

double Engine::now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint32_t kMaxScenes = 4;
// final queue: [finished | per scene: trace, trace + shadow, shadow]
constexpr uint32_t kFinalCategories = 1 + 3 * kMaxScenes;

enum TimedCat { T_CLOSEST, T_ANY, T_CLASSIFY, T_PARTITION, T_SHADE, T_GENERATE, T_FUSED, T_ACCUM };

Engine::Engine(const EngineConfig& cfg) : cfg_(cfg) {
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    for (uint32_t s = 0; s < kMaxScenes; ++s) {
        hip_check(hipStreamCreateWithFlags(&sceneStream_[s], hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&sceneDone_[s], hipEventDisableTiming), "hipEventCreate");
    }
    hip_check(hipEventCreateWithFlags(&stepStart_, hipEventDisableTiming), "hipEventCreate");
    // one pinned block read back each step: the 16 words at offs_ (the
    // queue offsets) and the per-slot live counts behind them
    hip_check(hipHostMalloc((void**)&hostOffs_, (16 + WRK_MAX_SLOTS) * sizeof(uint32_t), hipHostMallocDefault),
              "hipHostMalloc");
    hostLive_ = hostOffs_ + 16;
    hip_check(hipHostMalloc((void**)&hostPrm_, WRK_MAX_SLOTS * sizeof(wrk_render_params), hipHostMallocDefault),
              "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&hostLiveInit_, WRK_MAX_SLOTS * sizeof(uint32_t), hipHostMallocDefault),
              "hipHostMalloc");
    std::memset(hostPrm_, 0, WRK_MAX_SLOTS * sizeof(wrk_render_params));
    // offs_: the queue offsets at 0, the per-slot live counts at 16
    offs_.alloc((16 + WRK_MAX_SLOTS) * sizeof(uint32_t));
    prmTable_.alloc(WRK_MAX_SLOTS * sizeof(wrk_render_params));
    fetch_.alloc(64);
    hip_check(hipMemsetAsync(fetch_.ptr, 0, 64, stream_), "hipMemsetAsync");
    for (uint32_t s = WRK_MAX_SLOTS; s-- > 0;) freeSlots_.push_back(s);
}

Engine::~Engine() {
    for (hipEvent_t e : eventPool_) (void)hipEventDestroy(e);
    if (hostOffs_) (void)hipHostFree(hostOffs_);
    if (hostPrm_) (void)hipHostFree(hostPrm_);
    if (hostLiveInit_) (void)hipHostFree(hostLiveInit_);
    for (uint32_t s = 0; s < kMaxScenes; ++s) {
        if (sceneDone_[s]) (void)hipEventDestroy(sceneDone_[s]);
        if (sceneStream_[s]) (void)hipStreamDestroy(sceneStream_[s]);
    }
    if (stepStart_) (void)hipEventDestroy(stepStart_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::setConfig(const EngineConfig& cfg) {
    if (numActiveSlots() != 0 || !pending_.empty())
        throw std::runtime_error("wavefront configuration changed while requests are in flight");
    cfg_ = cfg;
}

size_t Engine::numActiveSlots() const { return WRK_MAX_SLOTS - freeSlots_.size(); }

void Engine::reserve(uint32_t cap) {
    if (cap > 0xFFFFFF00u) throw std::runtime_error("path pool too large");
    cap = granuleRound(cap);
    if (cap <= capacity_) return;
    if (numActiveSlots() != 0) throw std::runtime_error("path pool cannot grow while paths are in flight");
    size_t c = cap;
    org_.alloc(c * 16); dir_.alloc(c * 16); thr_.alloc(c * 16); rad_.alloc(c * 16); st_.alloc(c * 16);
    contrib_.alloc(c * 16); savedOrg_.alloc(c * 16); savedDir_.alloc(c * 16); hit_.alloc(c * 8);
    occl_.alloc(c); result_.alloc(c * 16);
    qA_.alloc(c * 4); qB_.alloc(c * 4); cat_.alloc(c);
    slotMap_.alloc(c / 256);
    scratch_.alloc(wrk_partition_scratch_bytes(cap, kFinalCategories));
    capacity_ = cap;
    paths_.org = org_.addr(); paths_.dir = dir_.addr(); paths_.thr = thr_.addr(); paths_.rad = rad_.addr();
    paths_.state = st_.addr(); paths_.contrib = contrib_.addr(); paths_.savedOrg = savedOrg_.addr();
    paths_.savedDir = savedDir_.addr(); paths_.hit = hit_.addr(); paths_.occluded = occl_.addr();
    paths_.result = result_.addr(); paths_.capacity = cap;
    if (guides_.ptr) {
        guides_.alloc(c * 32);
        paths_.guides = guides_.addr();
    }
    deadCount_ = 0;
    free_.clear();
    free_[0] = cap;
}

uint32_t Engine::sceneIndex(Scene* s) {
    for (uint32_t i = 0; i < scenes_.size(); ++i)
        if (scenes_[i] == s) return i;
    // reuse the index of a scene with no requests left
    for (uint32_t i = 0; i < scenes_.size(); ++i) {
        bool used = false;
        for (auto& kv : requests_)
            if (kv.second->sceneIdx == i && !kv.second->done) used = true;
        if (!used && segT_[i] == 0 && segTS_[i] == 0 && segS_[i] == 0) {
            scenes_[i] = s;
            return i;
        }
    }
    if (scenes_.size() >= kMaxScenes) throw std::runtime_error("at most 4 scenes can be resident in one wavefront");
    scenes_.push_back(s);
    segT_.push_back(0);
    segTS_.push_back(0);
    segS_.push_back(0);
    return (uint32_t)scenes_.size() - 1;
}

bool Engine::allocRange(uint32_t n, uint32_t* base) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= n) {
            *base = it->first;
            uint32_t rest = it->second - n;
            uint32_t b = it->first;
            free_.erase(it);
            if (rest) free_[b + n] = rest;
            return true;
        }
    }
    return false;
}

void Engine::freeRange(uint32_t base, uint32_t n) {
    auto it = free_.emplace(base, n).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free_.erase(next);
    }
    if (it != free_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free_.erase(it);
        }
    }
}

uint32_t Engine::largestFree() const {
    uint32_t m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
}

uint64_t Engine::submit(Scene* scene, const wrk_camera& cam, const wrk_render_params& prm, uint32_t spp,
                        uint64_t firstPath, uint64_t numPaths, Mode mode, float* pathsOut, bool guides) {
    if (!scene) throw std::runtime_error("render: null scene");
    if (prm.width == 0 || prm.height == 0) throw std::runtime_error("render: empty image");
    if (numPaths == 0) throw std::runtime_error("render: no paths requested");
    scene->finalize();
    auto r = std::make_unique<RequestState>();
    r->id = nextId_++;
    r->scene = scene;
    r->sceneIdx = sceneIndex(scene);
    r->cam = cam;
    r->prm = prm;
    r->mode = mode;
    r->firstPath = firstPath;
    r->nextPath = firstPath;
    r->accumNext = firstPath;
    r->endPath = firstPath + numPaths;
    r->spp = spp;
    r->npix = prm.width * prm.height;
    r->pathsOut = pathsOut;
    r->tSubmit = now();
    r->guides = guides && mode == Mode::Accumulate;
    r->footprint = mode == Mode::Accumulate ? (uint64_t)r->npix * (r->guides ? 84u : 24u) : 0u;
    uint64_t id = r->id;
    requests_[id] = std::move(r);
    pending_.push_back(id);
    return id;
}

// Request buffers (accumulation, image and guide layers) come from a pool
// of freed ones: allocating and freeing them per request cost ~1 ms each
// (eight of them per denoised 1080p render).
DevBuf Engine::takeBuf(size_t bytes) {
    // the smallest pooled buffer that holds `bytes` without wasting more than half of it
    auto it = bufPool_.lower_bound(bytes);
    if (it != bufPool_.end() && it->first <= 2 * bytes) {
        DevBuf b = std::move(it->second);
        bufPool_.erase(it);
        pooledBytes_ -= b.bytes;
        return b;
    }
    DevBuf b;
    b.alloc(bytes);
    return b;
}

void Engine::giveBuf(DevBuf& b) {
    if (!b.ptr) return;
    if (pooledBytes_ + b.bytes > kBufPoolBytes) {
        b.release();
        return;
    }
    pooledBytes_ += b.bytes;
    const size_t n = b.bytes;
    bufPool_.emplace(n, std::move(b));
}

bool Engine::isDone(uint64_t id) const {
    auto it = requests_.find(id);
    if (it == requests_.end()) throw std::runtime_error("unknown request id");
    return it->second->done;
}

const Engine::RequestState& Engine::request(uint64_t id) const {
    auto it = requests_.find(id);
    if (it == requests_.end()) throw std::runtime_error("unknown request id");
    return *it->second;
}

void Engine::release(uint64_t id) {
    auto it = requests_.find(id);
    if (it == requests_.end()) return;
    if (!it->second->done) throw std::runtime_error("request still in flight");
    footprintUsed() -= it->second->admitted ? it->second->footprint : 0;
    RequestState& r = *it->second;
    for (DevBuf* b : {&r.accum, &r.image, &r.accAlb, &r.accNrm, &r.imgAlb, &r.imgNrm, &r.accNoise, &r.imgNoise})
        giveBuf(*b);
    requests_.erase(it);
}

void Engine::fetch(uint64_t id, float* out, float* albedo, float* normal, float* noise) {
    RequestState& r = *requests_.at(id);
    if (!r.done) throw std::runtime_error("request still in flight");
    if (r.mode != Mode::Accumulate) throw std::runtime_error("request has no image");
    if ((albedo || normal || noise) && !r.guides) throw std::runtime_error("request was submitted without guides");
    if (noise)
        hip_check(hipMemcpy(noise, r.imgNoise.ptr, (size_t)r.npix * sizeof(float), hipMemcpyDeviceToHost),
                  "hipMemcpy D2H");
    const size_t bytes = (size_t)r.npix * 3 * sizeof(float);
    if (out) hip_check(hipMemcpy(out, r.image.ptr, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    if (albedo) hip_check(hipMemcpy(albedo, r.imgAlb.ptr, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    if (normal) hip_check(hipMemcpy(normal, r.imgNrm.ptr, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    release(id);
}

const float* Engine::deviceImage(uint64_t id) const {
    const RequestState& r = request(id);
    if (!r.done || r.mode != Mode::Accumulate) throw std::runtime_error("request has no finished image");
    return r.image.as<float>();
}
const float* Engine::deviceAlbedo(uint64_t id) const {
    const RequestState& r = request(id);
    if (!r.done || !r.guides) throw std::runtime_error("request has no finished guides");
    return r.imgAlb.as<float>();
}
const float* Engine::deviceNormal(uint64_t id) const {
    const RequestState& r = request(id);
    if (!r.done || !r.guides) throw std::runtime_error("request has no finished guides");
    return r.imgNrm.as<float>();
}
const float* Engine::deviceNoise(uint64_t id) const {
    const RequestState& r = request(id);
    if (!r.done || !r.guides) throw std::runtime_error("request has no finished guides");
    return r.imgNoise.as<float>();
}

template <typename F> void Engine::timed(int category, hipStream_t stream, F&& launch) {
    if (!cfg_.timeKernels) {
        launch();
        return;
    }
    while (eventPool_.size() < eventNext_ + 2) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        eventPool_.push_back(e);
    }
    hipEvent_t a = eventPool_[eventNext_++], b = eventPool_[eventNext_++];
    hip_check(hipEventRecord(a, stream), "hipEventRecord");
    launch();
    hip_check(hipEventRecord(b, stream), "hipEventRecord");
    marks_.push_back({category, {a, b}});
}

// Frees a chunk's pool range and slot.
void Engine::retire(uint32_t s) {
    Slot& sl = slots_[s];
    freeRange(sl.poolBase, granuleRound(sl.count));
    sl.active = false;
    sl.finished = false;
    freeSlots_.push_back(s);
}

// Chunk completion. Paths mode copies the per-path results out at once;
// accumulate mode adds every finished chunk that continues the sample order
// (renderer.cpp:190: samples in order) and holds the others.
void Engine::complete(uint32_t s) {
    Slot& sl = slots_[s];
    RequestState& r = *requests_.at(sl.request);
    sl.finished = true;
    r.chunksInFlight--;
    r.finishedPaths += sl.count;
    if (r.mode == Mode::Paths) {
        float* dst = r.pathsOut + (sl.firstPath - r.firstPath) * 4;
        hip_check(hipMemcpyAsync(dst, result_.as<float4>() + sl.poolBase, (size_t)sl.count * 16,
                                 hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync");
        retire(s);
    } else {
        r.held[sl.firstPath] = s;
        while (!r.held.empty() && r.held.begin()->first == r.accumNext) {
            const uint32_t hs = r.held.begin()->second;
            Slot& h = slots_[hs];
            timed(T_ACCUM, stream_, [&] {
                wrk_check(wrk_accumulate(&paths_, h.poolBase, h.firstPath, h.count, r.npix, r.accum.as<float>(),
                                         r.guides ? r.accAlb.as<float>() : nullptr,
                                         r.guides ? r.accNrm.as<float>() : nullptr,
                                         r.guides ? r.accNoise.as<float>() : nullptr, stream_),
                          "wrk_accumulate");
            });
            r.accumNext += h.count;
            r.held.erase(r.held.begin());
            retire(hs);
        }
    }
    if (r.finishedPaths == r.endPath - r.firstPath) {
        if (r.mode == Mode::Accumulate)
            wrk_check(wrk_average(r.accum.as<float>(), r.npix, r.spp, r.image.as<float>(), stream_), "wrk_average");
        if (r.guides) {
            wrk_check(wrk_average(r.accAlb.as<float>(), r.npix, r.spp, r.imgAlb.as<float>(), stream_), "wrk_average");
            wrk_check(wrk_average(r.accNrm.as<float>(), r.npix, r.spp, r.imgNrm.as<float>(), stream_), "wrk_average");
            wrk_check(wrk_noise_stderr(r.accNoise.as<float>(), r.npix, r.spp, r.imgNoise.as<float>(), stream_),
                      "wrk_noise_stderr");
        }
        r.done = true;
        r.tDone = -1;  // stamped after the step's synchronisation
        completed_.push_back(r.id);
    }
}

// Admits the next chunk of r into a free pool range and slot and generates
// its camera rays behind the queue entries already placed. Returns false
// when no slot, no memory or no large enough range is free.
bool Engine::admitChunk(RequestState& r, uint32_t* added, uint32_t queueBase, uint64_t maxPaths) {
    if (freeSlots_.empty() || capacity_ == 0) return false;
    if (!r.admitted && cfg_.memoryBudgetBytes && footprintUsed() != 0 &&
        footprintUsed() + r.footprint > cfg_.memoryBudgetBytes)
        return false;  // back-pressure: wait for memory to be released
    const uint64_t want = r.endPath - r.nextPath;
    const uint64_t chunkCap = cfg_.maxChunkPaths ? std::min<uint64_t>(cfg_.maxChunkPaths, capacity_) : capacity_;
    const uint64_t full = std::min<uint64_t>(std::min<uint64_t>(want, chunkCap), maxPaths);  // the chunk an empty pool would take
    // smaller chunks wait for more paths to end, unless nothing is in flight
    const uint64_t floor = cfg_.minChunkPaths ? cfg_.minChunkPaths
                                              : std::max<uint64_t>(capacity_ / 32, std::min<uint64_t>(65536, capacity_ / 8));
    const uint64_t minChunk = numActiveSlots() == 0 ? 1 : std::min<uint64_t>(full, floor);
    // ranges are whole granules: a chunk of n paths holds granuleRound(n)
    // entries; free ranges are whole granules too, so when the rounded chunk
    // does not fit, the largest free range is smaller than n itself
    uint32_t n = (uint32_t)full;
    uint32_t base = 0;
    if (!allocRange(granuleRound(n), &base)) {
        const uint32_t m = largestFree();
        if (m == 0 || m < minChunk) return false;
        n = m;
        if (!allocRange(n, &base)) return false;
    }
    if (!r.admitted) {
        if (r.mode == Mode::Accumulate) {
            r.accum = takeBuf((size_t)r.npix * 3 * sizeof(float));
            r.image = takeBuf((size_t)r.npix * 3 * sizeof(float));
            hip_check(hipMemsetAsync(r.accum.ptr, 0, (size_t)r.npix * 3 * sizeof(float), stream_), "hipMemsetAsync");
        }
        if (r.guides) {
            const size_t bytes = (size_t)r.npix * 3 * sizeof(float);
            r.accAlb = takeBuf(bytes); r.accNrm = takeBuf(bytes); r.imgAlb = takeBuf(bytes); r.imgNrm = takeBuf(bytes);
            r.accNoise = takeBuf((size_t)r.npix * 2 * sizeof(float));
            r.imgNoise = takeBuf((size_t)r.npix * sizeof(float));
            hip_check(hipMemsetAsync(r.accAlb.ptr, 0, bytes, stream_), "hipMemsetAsync");
            hip_check(hipMemsetAsync(r.accNrm.ptr, 0, bytes, stream_), "hipMemsetAsync");
            hip_check(hipMemsetAsync(r.accNoise.ptr, 0, (size_t)r.npix * 2 * sizeof(float), stream_),
                      "hipMemsetAsync");
            if (!guides_.ptr) {
                // first request with guides: per-path guide records from now on
                guides_.alloc((size_t)capacity_ * 32);
                paths_.guides = guides_.addr();
            }
        }
        r.admitted = true;
        r.tAdmit = now();
        footprintUsed() += r.footprint;
    }
    const uint32_t s = freeSlots_.back();
    freeSlots_.pop_back();
    hip_check(hipMemsetAsync(slotMap_.as<uint8_t>() + base / 256, (int)s, granuleRound(n) / 256, stream_),
              "hipMemsetAsync");
    Slot& sl = slots_[s];
    sl.active = true;
    sl.finished = false;
    sl.request = r.id;
    sl.poolBase = base;
    sl.count = n;
    sl.firstPath = r.nextPath;
    // pinned staging: each slot's entries are written once per step and the
    // step ends with a stream synchronisation
    const bool fused = cfg_.fusedMaxPaths && r.endPath - r.firstPath <= cfg_.fusedMaxPaths;
    hostPrm_[s] = r.prm;
    hostLiveInit_[s] = fused ? 0u : n;  // a fused chunk has ended when its launch completes
    hip_check(hipMemcpyAsync(prmTable_.as<wrk_render_params>() + s, hostPrm_ + s, sizeof(wrk_render_params),
                             hipMemcpyHostToDevice, stream_),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(offs_.as<uint32_t>() + 16 + s, hostLiveInit_ + s, sizeof(uint32_t),
                             hipMemcpyHostToDevice, stream_),
              "hipMemcpyAsync");
    if (fused) {
        // the whole chunk in one launch; the slot completes at this step's end
        timed(T_FUSED, stream_, [&] {
            wrk_check(wrk_render_fused(&r.scene->getBVH(), &r.scene->getDeviceView(), &r.cam,
                                       prmTable_.as<wrk_render_params>(), s, &paths_, r.nextPath, n, base, stream_),
                      "wrk_render_fused");
        });
        stats_.fusedPaths += n;
    } else {
        const uint32_t at = queueBase + *added;
        timed(T_GENERATE, stream_, [&] {
            wrk_check(wrk_generate(&r.cam, &r.prm, &paths_, r.nextPath, n, base, s, qB_.as<uint32_t>() + at,
                                   cat_.as<uint8_t>() + at, 1 + 3 * r.sceneIdx, stream_),
                      "wrk_generate");
        });
        *added += n;
    }
    r.nextPath += n;
    r.chunksInFlight++;
    r.chunks++;
    stats_.batches++;
    return true;
}

void Engine::launch() {
    if (launched_) throw std::runtime_error("wavefront iteration launched twice");
    const uint32_t S = (uint32_t)scenes_.size();
    uint32_t* qa = qA_.as<uint32_t>();
    uint32_t* qb = qB_.as<uint32_t>();
    uint8_t* cat = cat_.as<uint8_t>();
    uint32_t* off = offs_.as<uint32_t>();
    const wrk_render_params* prmTable = prmTable_.as<wrk_render_params>();
    uint32_t* live = off + 16;
    unsigned long long* fc = fetch_.as<unsigned long long>();
    eventNext_ = 0;
    marks_.clear();

    // --- one iteration over the merged queue, scene by scene. A scene's
    // part is [trace | trace + shadow | shadow]: the closest-hit launch takes
    // the first two pieces, the any-hit launch (shadow rays in savedOrg /
    // savedDir) the last two.
    hip_check(hipEventRecord(stepStart_, stream_), "hipEventRecord");
    uint32_t total = 0;
    uint32_t* const qlive = qa + deadCount_;  // qa = [finished last iteration | live parts]
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t nT = segT_[s], nTS = segTS_[s], nS = segS_[s];
        const uint32_t a = total, n = nT + nTS + nS;
        total += n;
        if (n == 0) continue;
        hipStream_t ss = sceneStream_[s];
        hip_check(hipStreamWaitEvent(ss, stepStart_, 0), "hipStreamWaitEvent");
        const wrk_bvh& bvh = scenes_[s]->getBVH();
        const wrk_scene_view& view = scenes_[s]->getDeviceView();
        const uint32_t nC = nT + nTS, nA = nTS + nS;
        // Each traversal resolves its rays where they end (the decisions of
        // wrk_classify): shadow rays first, adding unoccluded contributions
        // and placing the shadow-only entries; then the closest-hit rays,
        // whose misses add the environment and flush after those
        // contributions (libWR's Shadow -> Shade -> Trace order), and whose
        // hits write their shade category.
        if (nA) {
            timed(T_ANY, ss, [&] {
                wrk_check(wrk_trace_any_resolve(&bvh, prmTable, &paths_, qlive + a + nT, nTS, nS, cat + a + nT,
                                                cfg_.countTraversal ? fc + 4 : nullptr, ss),
                          "wrk_trace_any_resolve");
            });
            stats_.launchesTraceAny++;
        }
        if (nC) {
            timed(T_CLOSEST, ss, [&] {
                wrk_check(wrk_trace_closest_resolve(&bvh, &view, prmTable, &paths_, qlive + a, nC, cat + a,
                                                    cfg_.countTraversal ? fc : nullptr, ss),
                          "wrk_trace_closest_resolve");
            });
            stats_.launchesTraceClosest++;
        }
        stats_.closestRays += nC;
        stats_.shadowRays += nA;
        // the material shaders in queue order, which keeps each admitted
        // chunk's paths in pool order
        timed(T_SHADE, ss, [&] {
            wrk_check(wrk_shade(&view, prmTable, &paths_, qlive + a, qb + a, n, 1 + 3 * s, 0, cat + a, ss),
                      "wrk_shade");
        });
        hip_check(hipEventRecord(sceneDone_[s], ss), "hipEventRecord");
        hip_check(hipStreamWaitEvent(stream_, sceneDone_[s], 0), "hipStreamWaitEvent");
    }
    if (total) stats_.iterations++;

    // --- admission: new chunks join behind the shaded paths, in the order of
    // the pending list (arrival, or fewest remaining paths first). By default
    // the first request takes every chunk that fits before the next one gets
    // any, so requests complete in order as they would served one after
    // another while the pool stays full; with roundRobin each pending
    // request takes one chunk per round. A request that does not fit (slots,
    // pool ranges or the memory budget) stops admission of those behind it.
    uint32_t added = 0;
    if (cfg_.shortestFirst && pending_.size() > 1) {
        std::stable_sort(pending_.begin(), pending_.end(), [&](uint64_t x, uint64_t y) {
            const RequestState& a = *requests_.at(x);
            const RequestState& b = *requests_.at(y);
            return a.endPath - a.nextPath < b.endPath - b.nextPath;
        });
    }
    if (!cfg_.roundRobin) {
        // backfill budget: requests behind the oldest unfinished one (the
        // head) share at most backfillPaths of the pool while it has paths
        // in flight, counted from the live counts of the last step
        uint64_t headId = 0, behindLive = 0;
        if (cfg_.backfillPaths) {
            for (auto& kv : requests_)
                if (!kv.second->done) {
                    headId = kv.first;
                    break;
                }
            for (uint32_t sl = 0; sl < WRK_MAX_SLOTS; ++sl)
                if (slots_[sl].active && !slots_[sl].finished && slots_[sl].request != headId)
                    behindLive += hostLive_[sl];
        }
        while (!pending_.empty()) {
            RequestState& r = *requests_.at(pending_.front());
            uint64_t cap = ~0ull;
            const bool fused = cfg_.fusedMaxPaths && r.endPath - r.firstPath <= cfg_.fusedMaxPaths;
            if (cfg_.backfillPaths && r.id != headId && !fused) {
                // no chunk smaller than a 64K-path share (or the rest of the request)
                const uint64_t least = std::min<uint64_t>(65536, r.endPath - r.nextPath);
                if (behindLive + least > cfg_.backfillPaths) break;
                cap = cfg_.backfillPaths - behindLive;
            }
            const uint64_t before = r.nextPath;
            if (!admitChunk(r, &added, total, cap)) break;
            if (r.id != headId && !fused) behindLive += r.nextPath - before;
            if (r.nextPath == r.endPath) pending_.erase(pending_.begin());
        }
    }
    for (bool progress = cfg_.roundRobin; progress && !pending_.empty();) {
        progress = false;
        for (auto it = pending_.begin(); it != pending_.end();) {
            RequestState& r = *requests_.at(*it);
            if (!admitChunk(r, &added, total)) {
                if (!r.admitted) break;
                ++it;
                continue;
            }
            progress = true;
            if (r.nextPath == r.endPath) it = pending_.erase(it);
            else ++it;
        }
    }
    if (numActiveSlots() == 0 && !pending_.empty()) {
        // an idle pool admits any request unless the memory budget holds it
        // back until the caller fetches finished images
        const RequestState& r = *requests_.at(pending_.front());
        const bool memoryBound = cfg_.memoryBudgetBytes && footprintUsed() != 0 &&
                                 footprintUsed() + r.footprint > cfg_.memoryBudgetBytes;
        if (!memoryBound) throw std::runtime_error("wavefront admission made no progress on an idle pool");
    }

    // --- rebuild the queue: [finished | per scene: trace, trace + shadow, shadow]
    const uint32_t S2 = (uint32_t)scenes_.size();
    const uint32_t K = 1 + 3 * S2;
    const uint32_t items = total + added;
    if (items) {
        timed(T_PARTITION, stream_, [&] {
            wrk_check(wrk_partition(cat, qb, items, K, qa, off, scratch_.ptr, scratch_.bytes, stream_),
                      "wrk_partition");
        });
        // live paths per slot: subtract the paths that ended
        timed(T_PARTITION, stream_, [&] {
            wrk_check(wrk_slot_finish_map(qa, items, off + 1, slotMap_.as<uint8_t>(), capacity_ / 256, live,
                                          stream_),
                      "wrk_slot_finish_map");
        });
    }
    // queue offsets and live counts in one read-back
    if (items || numActiveSlots())
        hip_check(hipMemcpyAsync(hostOffs_, off, (16 + WRK_MAX_SLOTS) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync");
    stats_.partitionItems += items;
    launchedTotal_ = total;
    launchedItems_ = items;
    launchedScenes_ = S2;
    launched_ = true;
}

bool Engine::iterationDone() const {
    const hipError_t e = hipStreamQuery(stream_);
    if (e == hipErrorNotReady) return false;
    hip_check(e, "hipStreamQuery");
    return true;
}

bool Engine::finish() {
    if (!launched_) throw std::runtime_error("wavefront iteration finished before it was launched");
    launched_ = false;
    const uint32_t items = launchedItems_, S2 = launchedScenes_;
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    stats_.shadeItems += launchedTotal_;
    deadCount_ = items ? hostOffs_[1] : 0;
    for (uint32_t s = 0; s < S2; ++s) {
        const uint32_t* o = hostOffs_ + 1 + 3 * s;
        segT_[s] = items ? o[1] - o[0] : 0;
        segTS_[s] = items ? o[2] - o[1] : 0;
        segS_[s] = items ? o[3] - o[2] : 0;
    }

    // --- completions
    bool any = false;
    for (uint32_t s = 0; s < WRK_MAX_SLOTS; ++s) {
        if (slots_[s].active && !slots_[s].finished && hostLive_[s] == 0) {
            complete(s);
            any = true;
        }
    }
    if (any || cfg_.timeKernels) hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    if (any) {
        double t = now();
        for (auto& kv : requests_)
            if (kv.second->done && kv.second->tDone < 0) kv.second->tDone = t;
    }
    for (auto& m : marks_) {
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, m.second.first, m.second.second), "hipEventElapsedTime");
        switch (m.first) {
            case T_CLOSEST: stats_.msTraceClosest += ms; break;
            case T_ANY: stats_.msTraceAny += ms; break;
            case T_CLASSIFY: stats_.msClassify += ms; break;
            case T_PARTITION: stats_.msPartition += ms; break;
            case T_SHADE: stats_.msShade += ms; break;
            case T_GENERATE: stats_.msGenerate += ms; break;
            case T_FUSED: stats_.msFused += ms; break;
            default: stats_.msAccumulate += ms; break;
        }
    }
    if (cfg_.countTraversal) {
        unsigned long long c[8];
        hip_check(hipMemcpy(c, fetch_.ptr, sizeof(c), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        stats_.closestNodeFetches = c[0]; stats_.closestTriFetches = c[1]; stats_.closestStackSpills = c[2];
        stats_.anyNodeFetches = c[4]; stats_.anyTriFetches = c[5]; stats_.anyStackSpills = c[6];
    }
    return !pending_.empty() || numActiveSlots() != 0;
}

}  // namespace detail
}  // namespace wr
