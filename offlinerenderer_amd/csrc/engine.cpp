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

enum TimedCat { T_CLOSEST, T_ANY, T_CLASSIFY, T_PARTITION, T_SHADE, T_GENERATE, T_ACCUM };

Engine::Engine(const EngineConfig& cfg) : cfg_(cfg) {
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipHostMalloc((void**)&hostOffs_, 64 * sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&hostLive_, WRK_MAX_SLOTS * sizeof(uint32_t), hipHostMallocDefault),
              "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&hostPrm_, WRK_MAX_SLOTS * sizeof(wrk_render_params), hipHostMallocDefault),
              "hipHostMalloc");
    hip_check(hipHostMalloc((void**)&hostRanges_, 2 * WRK_MAX_SLOTS * sizeof(uint32_t), hipHostMallocDefault),
              "hipHostMalloc");
    std::memset(hostPrm_, 0, WRK_MAX_SLOTS * sizeof(wrk_render_params));
    offs_.alloc(64 * sizeof(uint32_t));
    prmTable_.alloc(WRK_MAX_SLOTS * sizeof(wrk_render_params));
    live_.alloc(WRK_MAX_SLOTS * sizeof(uint32_t));
    ranges_.alloc(2 * WRK_MAX_SLOTS * sizeof(uint32_t));
    fetch_.alloc(64);
    hip_check(hipMemsetAsync(fetch_.ptr, 0, 64, stream_), "hipMemsetAsync");
    for (uint32_t s = WRK_MAX_SLOTS; s-- > 0;) freeSlots_.push_back(s);
}

Engine::~Engine() {
    for (hipEvent_t e : eventPool_) (void)hipEventDestroy(e);
    if (hostOffs_) (void)hipHostFree(hostOffs_);
    if (hostLive_) (void)hipHostFree(hostLive_);
    if (hostPrm_) (void)hipHostFree(hostPrm_);
    if (hostRanges_) (void)hipHostFree(hostRanges_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::setConfig(const EngineConfig& cfg) {
    if (numActiveSlots() != 0 || !pending_.empty())
        throw std::runtime_error("wavefront configuration changed while requests are in flight");
    cfg_ = cfg;
}

size_t Engine::numActiveSlots() const { return WRK_MAX_SLOTS - freeSlots_.size(); }

void Engine::reserve(uint32_t cap) {
    if (cap <= capacity_) return;
    if (numActiveSlots() != 0) throw std::runtime_error("path pool cannot grow while paths are in flight");
    size_t c = cap;
    org_.alloc(c * 16); dir_.alloc(c * 16); thr_.alloc(c * 16); rad_.alloc(c * 16); st_.alloc(c * 16);
    contrib_.alloc(c * 16); savedOrg_.alloc(c * 16); savedDir_.alloc(c * 16); hit_.alloc(c * 8);
    occl_.alloc(c); result_.alloc(c * 16);
    qA_.alloc(c * 4); qB_.alloc(c * 4); cat_.alloc(c);
    scratch_.alloc(wrk_partition_scratch_bytes(cap, 8));
    capacity_ = cap;
    paths_.org = org_.addr(); paths_.dir = dir_.addr(); paths_.thr = thr_.addr(); paths_.rad = rad_.addr();
    paths_.state = st_.addr(); paths_.contrib = contrib_.addr(); paths_.savedOrg = savedOrg_.addr();
    paths_.savedDir = savedDir_.addr(); paths_.hit = hit_.addr(); paths_.occluded = occl_.addr();
    paths_.result = result_.addr(); paths_.capacity = cap;
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
        if (!used && segC_[i] == 0 && segS_[i] == 0) {
            scenes_[i] = s;
            return i;
        }
    }
    if (scenes_.size() >= 4) throw std::runtime_error("at most 4 scenes can be resident in one wavefront");
    scenes_.push_back(s);
    segC_.push_back(0);
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
                        uint64_t firstPath, uint64_t numPaths, Mode mode, float* pathsOut) {
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
    r->endPath = firstPath + numPaths;
    r->spp = spp;
    r->npix = prm.width * prm.height;
    r->pathsOut = pathsOut;
    r->tSubmit = now();
    r->footprint = mode == Mode::Accumulate ? (uint64_t)r->npix * 24u : 0u;
    uint64_t id = r->id;
    requests_[id] = std::move(r);
    pending_.push_back(id);
    return id;
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
    footprintUsed_ -= it->second->admitted ? it->second->footprint : 0;
    requests_.erase(it);
}

void Engine::fetch(uint64_t id, float* out) {
    RequestState& r = *requests_.at(id);
    if (!r.done) throw std::runtime_error("request still in flight");
    if (r.mode != Mode::Accumulate) throw std::runtime_error("request has no image");
    hip_check(hipMemcpy(out, r.image.ptr, (size_t)r.npix * 3 * sizeof(float), hipMemcpyDeviceToHost),
              "hipMemcpy D2H");
    release(id);
}

template <typename F> void Engine::timed(int category, F&& launch) {
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
    hip_check(hipEventRecord(a, stream_), "hipEventRecord");
    launch();
    hip_check(hipEventRecord(b, stream_), "hipEventRecord");
    marks_.push_back({category, {a, b}});
}

// Chunk completion: accumulate (or copy out) the chunk's per-path results,
// free its pool range and slot, queue the request's next chunk or finish it.
void Engine::complete(uint32_t s) {
    Slot& sl = slots_[s];
    RequestState& r = *requests_.at(sl.request);
    if (r.mode == Mode::Accumulate) {
        timed(T_ACCUM, [&] {
            wrk_check(wrk_accumulate(&paths_, sl.poolBase, sl.firstPath, sl.count, r.npix, r.accum.as<float>(), stream_),
                      "wrk_accumulate");
        });
    } else {
        float* dst = r.pathsOut + (sl.firstPath - r.firstPath) * 4;
        hip_check(hipMemcpyAsync(dst, result_.as<float4>() + sl.poolBase, (size_t)sl.count * 16,
                                 hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync");
    }
    freeRange(sl.poolBase, sl.count);
    sl.active = false;
    freeSlots_.push_back(s);
    rangesDirty_ = true;
    r.inFlight = false;
    if (r.nextPath < r.endPath) {
        pending_.push_front(r.id);  // keeps its place ahead of later arrivals
    } else {
        if (r.mode == Mode::Accumulate)
            wrk_check(wrk_average(r.accum.as<float>(), r.npix, r.spp, r.image.as<float>(), stream_), "wrk_average");
        r.done = true;
        r.tDone = -1;  // stamped after the step's synchronisation
    }
}

bool Engine::step() {
    const uint32_t S = (uint32_t)scenes_.size();
    uint32_t* qa = qA_.as<uint32_t>();
    uint32_t* qb = qB_.as<uint32_t>();
    uint8_t* cat = cat_.as<uint8_t>();
    uint32_t* off = offs_.as<uint32_t>();
    const wrk_render_params* prmTable = prmTable_.as<wrk_render_params>();
    uint32_t* live = live_.as<uint32_t>();
    unsigned long long* fc = fetch_.as<unsigned long long>();
    eventNext_ = 0;
    marks_.clear();

    // --- one iteration over the merged queue, scene by scene
    uint32_t total = 0;
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t nC = segC_[s], nS = segS_[s];
        const uint32_t a = total;
        total += nC + nS;
        if (nC + nS == 0) continue;
        const wrk_bvh& bvh = scenes_[s]->getBVH();
        const wrk_scene_view& view = scenes_[s]->getDeviceView();
        if (nC) {
            timed(T_CLOSEST, [&] {
                if (cfg_.countTraversal)
                    wrk_check(wrk_trace_closest_counted(&bvh, org_.ptr, dir_.ptr, qa + a, nC, nullptr,
                                                        hit_.as<uint32_t>(), fc, stream_),
                              "wrk_trace_closest_counted");
                else
                    wrk_check(wrk_trace_closest(&bvh, org_.ptr, dir_.ptr, qa + a, nC, nullptr, hit_.as<uint32_t>(),
                                                stream_),
                              "wrk_trace_closest");
            });
            stats_.launchesTraceClosest++;
        }
        if (nS) {
            timed(T_ANY, [&] {
                if (cfg_.countTraversal)
                    wrk_check(wrk_trace_any_counted(&bvh, org_.ptr, dir_.ptr, qa + a + nC, nS, nullptr,
                                                    occl_.as<uint8_t>(), fc + 2, stream_),
                              "wrk_trace_any_counted");
                else
                    wrk_check(wrk_trace_any(&bvh, org_.ptr, dir_.ptr, qa + a + nC, nS, nullptr, occl_.as<uint8_t>(),
                                            stream_),
                              "wrk_trace_any");
            });
            stats_.launchesTraceAny++;
        }
        stats_.closestRays += nC;
        stats_.shadowRays += nS;
        timed(T_CLASSIFY, [&] {
            wrk_check(wrk_classify(&view, prmTable, &paths_, qa + a, nC, nS, cat + a, stream_), "wrk_classify");
        });
        timed(T_PARTITION, [&] {
            wrk_check(wrk_partition(cat + a, qa + a, nC + nS, 5, qb + a, off + 8 * s, scratch_.ptr, scratch_.bytes,
                                    stream_),
                      "wrk_partition");
        });
        timed(T_SHADE, [&] {
            wrk_check(wrk_shade(&view, prmTable, &paths_, qb + a, off + 8 * s, nC + nS, 2 * s, cat + a, stream_),
                      "wrk_shade");
        });
    }
    if (total) stats_.iterations++;

    // --- admission: new chunks join behind the shaded paths
    uint32_t added = 0;
    while (!pending_.empty() && !freeSlots_.empty()) {
        RequestState& r = *requests_.at(pending_.front());
        if (!r.admitted) {
            if (cfg_.memoryBudgetBytes && footprintUsed_ + r.footprint > cfg_.memoryBudgetBytes &&
                footprintUsed_ != 0)
                break;  // back-pressure: wait for memory to be released
        }
        uint64_t want = r.endPath - r.nextPath;
        uint64_t chunkCap = cfg_.maxChunkPaths ? cfg_.maxChunkPaths : capacity_;
        if (capacity_ == 0) break;
        uint32_t n = (uint32_t)std::min<uint64_t>(want, std::min<uint64_t>(chunkCap, capacity_));
        uint32_t base = 0;
        if (!allocRange(n, &base)) {
            if (numActiveSlots() != 0) break;  // wait for paths to drain
            n = largestFree();
            if (n == 0 || !allocRange(n, &base)) break;
        }
        if (!r.admitted) {
            if (r.mode == Mode::Accumulate) {
                r.accum.alloc((size_t)r.npix * 3 * sizeof(float));
                r.image.alloc((size_t)r.npix * 3 * sizeof(float));
                hip_check(hipMemsetAsync(r.accum.ptr, 0, (size_t)r.npix * 3 * sizeof(float), stream_),
                          "hipMemsetAsync");
            }
            r.admitted = true;
            r.tAdmit = now();
            footprintUsed_ += r.footprint;
        }
        uint32_t s = freeSlots_.back();
        freeSlots_.pop_back();
        Slot& sl = slots_[s];
        sl.active = true;
        sl.request = r.id;
        sl.poolBase = base;
        sl.count = n;
        sl.firstPath = r.nextPath;
        // pinned staging: each slot's entry is written once per step and the
        // step ends with a stream synchronisation
        hostPrm_[s] = r.prm;
        hip_check(hipMemcpyAsync(prmTable_.as<wrk_render_params>() + s, hostPrm_ + s, sizeof(wrk_render_params),
                                 hipMemcpyHostToDevice, stream_),
                  "hipMemcpyAsync");
        rangesDirty_ = true;
        timed(T_GENERATE, [&] {
            wrk_check(wrk_generate(&r.cam, &r.prm, &paths_, r.nextPath, n, base, s, qb + total + added,
                                   cat + total + added, 2 * r.sceneIdx, stream_),
                      "wrk_generate");
        });
        added += n;
        r.nextPath += n;
        r.inFlight = true;
        pending_.pop_front();
    }

    // --- rebuild the queue: [scene0 closest | scene0 shadow | scene1 ...]
    const uint32_t S2 = (uint32_t)scenes_.size();
    const uint32_t items = total + added;
    if (items) {
        timed(T_PARTITION, [&] {
            wrk_check(wrk_partition(cat, qb, items, 2 * S2, qa, off + 48, scratch_.ptr, scratch_.bytes, stream_),
                      "wrk_partition");
        });
        hip_check(hipMemcpyAsync(hostOffs_, off + 48, (2 * S2 + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync");
    }
    // live paths per slot = queue entries inside its pool range
    const uint32_t numRanges = (uint32_t)numActiveSlots();
    if (numRanges) {
        if (rangesDirty_) {
            uint32_t k = 0;
            std::vector<std::pair<uint32_t, uint32_t>> act;
            for (uint32_t s = 0; s < WRK_MAX_SLOTS; ++s)
                if (slots_[s].active) act.push_back({slots_[s].poolBase, s});
            std::sort(act.begin(), act.end());
            for (auto& e : act) {
                hostRanges_[k] = e.first;
                hostRanges_[numRanges + k] = e.second;
                ++k;
            }
            hip_check(hipMemcpyAsync(ranges_.ptr, hostRanges_, 2 * numRanges * sizeof(uint32_t),
                                     hipMemcpyHostToDevice, stream_),
                      "hipMemcpyAsync");
            rangesDirty_ = false;
        }
        timed(T_PARTITION, [&] {
            wrk_check(wrk_slot_count(qa, items, items ? off + 48 + 2 * S2 : nullptr, ranges_.as<uint32_t>(),
                                     numRanges, live, stream_),
                      "wrk_slot_count");
        });
        hip_check(hipMemcpyAsync(hostLive_, live, WRK_MAX_SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync");
    }
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (uint32_t s = 0; s < S2; ++s) {
        segC_[s] = items ? hostOffs_[2 * s + 1] - hostOffs_[2 * s] : 0;
        segS_[s] = items ? hostOffs_[2 * s + 2] - hostOffs_[2 * s + 1] : 0;
    }

    // --- completions
    bool any = false;
    for (uint32_t s = 0; s < WRK_MAX_SLOTS; ++s) {
        if (slots_[s].active && hostLive_[s] == 0) {
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
            default: stats_.msAccumulate += ms; break;
        }
    }
    if (cfg_.countTraversal) {
        unsigned long long c[4];
        hip_check(hipMemcpy(c, fetch_.ptr, sizeof(c), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        stats_.closestNodeFetches = c[0]; stats_.closestTriFetches = c[1];
        stats_.anyNodeFetches = c[2]; stats_.anyTriFetches = c[3];
    }
    return !pending_.empty() || numActiveSlots() != 0;
}

}  // namespace detail
}  // namespace wr
