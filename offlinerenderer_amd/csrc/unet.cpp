// unet.cpp — the learned denoiser network (replaces the closed OptiX AI
// denoiser of libWR/src/internal/denoiser.cpp:17-195, which takes the
// noisy colour plus albedo and normal guide layers).
//
// 4-level U-Net, 3x3 convolutions, channels 32 / 64 / 96 / 128, two
// convolutions per level and side, 2x2 average pooling, nearest upsampling,
// skip connections by concatenation ([upsampled | skip]), ReLU everywhere
// but the output convolution. Input: log(1 + colour), albedo, normal
// (9 channels, padded to 32). Output: channels 0..2 of a 32-channel tensor
// hold the correction added to the input's log(1 + colour). bf16 weights and activations, FP32 accumulation,
// FP32 master weights with Adam. All arithmetic runs in the kernels of
// denoise.hip; this file owns the graph and the buffers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/wrk.h"
#include "../../include/wrk_unet.h"

namespace {

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
void wrk_ok(int rc, const char* what) {
    if (rc != WRK_OK) throw std::runtime_error(std::string(what) + ": " + wrk_last_error());
}

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { if (p) (void)hipFree(p); }
    void need(size_t n) {
        if (n <= bytes) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hip_ok(hipMalloc(&p, n), "hipMalloc");
        bytes = n;
    }
    uint16_t* h() const { return static_cast<uint16_t*>(p); }
    float* f() const { return static_cast<float*>(p); }
};

// A tensor view: base pointer, channel pitch, channels, level (H >> level).
struct T {
    uint16_t* p;
    uint32_t pitch, c, lvl;
};

struct ConvSpec {
    uint32_t cin, cout, lvl;
    bool relu;
    uint64_t wOff, bOff;  // into the FP32 parameter vector
};

class UNet {
public:
    // head 0: channels 0..2 of a 32-channel output correct the input's
    // log(1 + colour); head 1: channels 0..48 of a 64-channel output are a
    // 7x7 filter per pixel applied to the noisy log colour and channels
    // 49..51 a per-channel offset (wrk_kpn_apply)
    // parameter count of a network with this head (without building one)
    static uint64_t paramsFor(uint32_t head) {
        if (head > 1) throw std::runtime_error("wrk_unet: unknown output head");
        return head ? 931904ull : 922656ull;
    }
    explicit UNet(uint32_t head) : head_(head) {
        if (head > 1) throw std::runtime_error("wrk_unet: unknown output head");
        // c0a c0b | c1a c1b | c2a c2b | c3a c3b | d2a d2b | d1a d1b | d0a d0b | out
        const uint32_t spec[15][3] = {{32, 32, 0},   {32, 32, 0},   {32, 64, 1},  {64, 64, 1},   {64, 96, 2},
                                      {96, 96, 2},   {96, 128, 3},  {128, 128, 3}, {224, 96, 2}, {96, 96, 2},
                                      {160, 64, 1},  {64, 64, 1},   {96, 32, 0},   {32, 32, 0},  {32, head ? 64u : 32u, 0}};
        uint64_t off = 0;
        for (int i = 0; i < 15; ++i) {
            ConvSpec c{spec[i][0], spec[i][1], spec[i][2], i != 14, 0, 0};
            c.wOff = off;
            off += 9ull * c.cin * c.cout;
            c.bOff = off;
            off += c.cout;
            conv_.push_back(c);
        }
        nParams_ = off;
        if (nParams_ != paramsFor(head)) throw std::runtime_error("wrk_unet: layer table and parameter count disagree");
        P_.need(off * 4); G_.need(off * 4); M_.need(off * 4); V_.need(off * 4);
        hip_ok(hipMemset(M_.p, 0, off * 4), "hipMemset");
        hip_ok(hipMemset(V_.p, 0, off * 4), "hipMemset");
        hip_ok(hipMemset(G_.p, 0, off * 4), "hipMemset");
        for (int i = 0; i < 15; ++i) {
            wb_.emplace_back(new Buf);
            wt_.emplace_back(new Buf);
            wb_[i]->need(9ull * conv_[i].cin * conv_[i].cout * 2);
            wt_[i]->need(9ull * conv_[i].cin * conv_[i].cout * 2);
        }
        init(1);
    }

    uint64_t numParams() const { return nParams_; }
    uint32_t head() const { return head_; }

    // He-normal weights over the real input channels, zero biases; the
    // output convolution starts small and only its first 3 channels live.
    void init(uint64_t seed) {
        std::vector<float> h(nParams_, 0.0f);
        std::mt19937_64 rng(seed);
        for (int i = 0; i < 15; ++i) {
            const ConvSpec& c = conv_[i];
            const uint32_t realIn = i == 0 ? 9 : c.cin;
            const uint32_t realOut = i == 14 ? (head_ ? wrk_kpn_taps() + 3u : 3u) : c.cout;
            std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * realIn)) * (i == 14 ? 0.1f : 1.0f));
            for (uint32_t tap = 0; tap < 9; ++tap)
                for (uint32_t co = 0; co < realOut; ++co)
                    for (uint32_t ci = 0; ci < realIn; ++ci)
                        h[c.wOff + ((uint64_t)tap * c.cout + co) * c.cin + ci] = nd(rng);
        }
        setParams(h.data());
        step_ = 0;
        hip_ok(hipMemset(M_.p, 0, nParams_ * 4), "hipMemset");
        hip_ok(hipMemset(V_.p, 0, nParams_ * 4), "hipMemset");
    }

    void setParams(const float* host) {
        hip_ok(hipMemcpy(P_.p, host, nParams_ * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        refreshBf16(nullptr);
        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }
    void getParams(float* host) const {
        hip_ok(hipMemcpy(host, P_.p, nParams_ * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    // input channel 9 of image k: log2(spp[k]) / 16
    void setSampleCounts(const uint32_t* spp, uint32_t count) {
        if (count && !spp) throw std::runtime_error("wrk_unet_set_sample_counts: null counts");
        std::vector<float> f(count);
        for (uint32_t k = 0; k < count; ++k) f[k] = spp[k] ? std::log2((float)spp[k]) / 16.0f : 0.0f;
        if (count) {
            feat_.need((size_t)count * 4);
            hip_ok(hipMemcpy(feat_.p, f.data(), (size_t)count * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        }
        featCount_ = count;
    }
    void setNoise(const float* noise) { noise_ = noise; }
    // Adam state for training checkpoints: both moment vectors (layout of
    // the parameters) and the step count
    void getState(float* m, float* v, uint32_t* step) const {
        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (m) hip_ok(hipMemcpy(m, M_.p, nParams_ * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if (v) hip_ok(hipMemcpy(v, V_.p, nParams_ * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        if (step) *step = step_;
    }
    void setState(const float* m, const float* v, uint32_t step) {
        if (!m || !v) throw std::runtime_error("wrk_unet_set_state: null moments");
        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_ok(hipMemcpy(M_.p, m, nParams_ * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        hip_ok(hipMemcpy(V_.p, v, nParams_ * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
        step_ = step;
    }
    void getGrads(float* host) const {
        hip_ok(hipMemcpy(host, G_.p, nParams_ * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }

    void forward(const float* color, const float* albedo, const float* normal, uint32_t srcH, uint32_t srcW,
                 const uint32_t* origins, uint32_t n, uint32_t h, uint32_t w, float* out, hipStream_t s) {
        shape(n, h, w, false);
        wrk_ok(wrk_unet_pack(color, albedo, normal, noise_, srcH, srcW, origins, featCount_ ? feat_.f() : nullptr,
                             featCount_, in_.buf.h(), n, h, w, s),
               "wrk_unet_pack");
        runForward(s, false, out);
        if (out && head_)
            wrk_ok(wrk_kpn_apply(pred_.buf.h(), pred_.c, color, srcH, srcW, origins, out, n, h, w, s), "wrk_kpn_apply");
    }

    float trainStep(const float* color, const float* albedo, const float* normal, const float* target,
                    const float* target2, uint32_t srcH,
                    uint32_t srcW, const uint32_t* origins, uint32_t n, uint32_t h, uint32_t w, float lr,
                    uint32_t lossMode, hipStream_t s) {
        shape(n, h, w, true);
        wrk_ok(wrk_unet_pack(color, albedo, normal, noise_, srcH, srcW, origins, featCount_ ? feat_.f() : nullptr,
                             featCount_, in_.buf.h(), n, h, w, s),
               "wrk_unet_pack");
        runForward(s, true, nullptr);
        if (head_)
            wrk_ok(wrk_kpn_loss(pred_.buf.h(), pred_.c, color, target, srcH, srcW, origins, target2, gPred_.buf.h(), n,
                                h, w, lossMode, lossPartial_.f(), s),
                   "wrk_kpn_loss");
        else
            wrk_ok(wrk_unet_loss(pred_.buf.h(), 32, target, srcH, srcW, origins, target2, in_.buf.h(), 32,
                                 gPred_.buf.h(), n, h, w, lossMode, lossPartial_.f(), s),
                   "wrk_unet_loss");
        runBackward(s);
        ++step_;
        wrk_ok(wrk_adam(P_.f(), G_.f(), M_.f(), V_.f(), nParams_, lr, 0.9f, 0.999f, 1e-8f, step_, s), "wrk_adam");
        refreshBf16(s);
        const uint64_t blocks = ((uint64_t)n * h * w + 255) / 256;
        std::vector<float> part(blocks);
        hip_ok(hipMemcpyAsync(part.data(), lossPartial_.p, blocks * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
        double sum = 0.0;
        for (float v : part) sum += v;
        return (float)(sum / (3.0 * (double)n * h * w));
    }

private:
    struct Act {
        Buf buf;
        uint32_t c = 0, lvl = 0;
    };

    T view(Act& a, uint32_t off = 0, uint32_t c = 0) {
        return T{a.buf.h() + off, a.c, c ? c : a.c - off, a.lvl};
    }

    void shape(uint32_t n, uint32_t h, uint32_t w, bool training) {
        if (h % 8 || w % 8 || n == 0) throw std::runtime_error("denoiser: image sides must be positive multiples of 8");
        n_ = n; h_ = h; w_ = w;
        auto alloc = [&](Act& a, uint32_t c, uint32_t lvl) {
            a.c = c;
            a.lvl = lvl;
            a.buf.need((size_t)n * (h >> lvl) * (w >> lvl) * c * 2);
        };
        alloc(in_, 32, 0); alloc(a0_, 32, 0); alloc(cat0_, 96, 0); alloc(t0_, 32, 0); alloc(d0_, 32, 0);
        alloc(pred_, conv_[14].cout, 0);
        alloc(p1_, 32, 1); alloc(a1_, 64, 1); alloc(cat1_, 160, 1); alloc(t1_, 64, 1); alloc(d1_, 64, 1);
        alloc(p2_, 64, 2); alloc(a2_, 96, 2); alloc(cat2_, 224, 2); alloc(t2_, 96, 2); alloc(d2_, 96, 2);
        alloc(p3_, 96, 3); alloc(a3_, 128, 3); alloc(b3_, 128, 3);
        if (!training) return;
        alloc(gPred_, conv_[14].cout, 0); alloc(g0a_, 32, 0); alloc(g0b_, 32, 0); alloc(gCat0_, 96, 0);
        alloc(g1a_, 64, 1); alloc(g1b_, 64, 1); alloc(gCat1_, 160, 1); alloc(gP1_, 32, 1);
        alloc(g2a_, 96, 2); alloc(g2b_, 96, 2); alloc(gCat2_, 224, 2); alloc(gP2_, 64, 2);
        alloc(g3a_, 128, 3); alloc(g3b_, 128, 3); alloc(gP3_, 96, 3);
        lossPartial_.need(wrk_unet_loss_scratch_bytes(n, h, w) + 16);
        uint64_t need = 0;
        for (const ConvSpec& c : conv_) {
            const uint64_t b = wrk_conv3x3_wgrad_scratch_bytes(c.cin, c.cout, n, h >> c.lvl, w >> c.lvl);
            need = b > need ? b : need;
        }
        wgradScratch_.need(need);
    }

    uint32_t H(uint32_t lvl) const { return h_ >> lvl; }
    uint32_t W(uint32_t lvl) const { return w_ >> lvl; }
    uint64_t npix(uint32_t lvl) const { return (uint64_t)n_ * H(lvl) * W(lvl); }

    void conv(int i, const T& x, const T& y, hipStream_t s) {
        const ConvSpec& c = conv_[i];
        if (i == 14 && head_ == 0) {
            // the residual head reads channels 0..2 of its output: compute
            // the first 16 (the others stay as they are)
            wrk_ok(wrk_conv3x3_fwd_first(x.p, x.pitch, c.cin, wb_[i]->h(), P_.f() + c.bOff, y.p, y.pitch, c.cout, 3u,
                                         c.relu ? 1u : 0u, n_, H(c.lvl), W(c.lvl), s),
                   "wrk_conv3x3_fwd_first");
            return;
        }
        wrk_ok(wrk_conv3x3_fwd(x.p, x.pitch, c.cin, wb_[i]->h(), P_.f() + c.bOff, y.p, y.pitch, c.cout,
                               c.relu ? 1u : 0u, n_, H(c.lvl), W(c.lvl), s),
               "wrk_conv3x3_fwd");
    }
    // gradients of conv i: dW/db from (x, dz); dx (if given) = data gradient
    void convBack(int i, const T& x, const T& dz, const T* dx, hipStream_t s) {
        const ConvSpec& c = conv_[i];
        wrk_ok(wrk_conv3x3_wgrad(x.p, x.pitch, c.cin, dz.p, dz.pitch, c.cout, n_, H(c.lvl), W(c.lvl),
                                 G_.f() + c.wOff, G_.f() + c.bOff, wgradScratch_.p, wgradScratch_.bytes, s),
               "wrk_conv3x3_wgrad");
        if (dx)
            wrk_ok(wrk_conv3x3_fwd(dz.p, dz.pitch, c.cout, wt_[i]->h(), nullptr, dx->p, dx->pitch, c.cin, 0u, n_,
                                   H(c.lvl), W(c.lvl), s),
                   "wrk_conv3x3_fwd (data gradient)");
    }
    void reluBack(const T& g, const T& y, hipStream_t s) {
        wrk_ok(wrk_relu_bwd(g.p, g.pitch, y.p, y.pitch, g.c, npix(g.lvl), s), "wrk_relu_bwd");
    }
    // encoder convolution i writing its skip y and, from the same epilogue,
    // the 2x2 average pool of y (the next level's input)
    void convPool(int i, const T& x, const T& y, const T& pooled, hipStream_t s) {
        const ConvSpec& c = conv_[i];
        wrk_ok(wrk_conv3x3_fwd_pool(x.p, x.pitch, c.cin, wb_[i]->h(), P_.f() + c.bOff, y.p, y.pitch, c.cout,
                                    c.relu ? 1u : 0u, pooled.p, pooled.pitch, n_, H(c.lvl), W(c.lvl), s),
               "wrk_conv3x3_fwd_pool");
    }
    void up(const T& x, const T& y, hipStream_t s) {
        wrk_ok(wrk_up2_fwd(x.p, x.pitch, x.c, y.p, y.pitch, n_, H(y.lvl), W(y.lvl), s), "wrk_up2_fwd");
    }
    // decoder convolution i over [2x upsampled low | skip] (cat holds the
    // skip after low.c channels): training materialises the upsampled half,
    // which the weight gradient reads; inference reads low directly
    void upConv(int i, const T& low, const T& cat, const T& y, bool training, hipStream_t s) {
        const ConvSpec& c = conv_[i];
        if (training) {
            up(low, T{cat.p, cat.pitch, low.c, cat.lvl}, s);
            conv(i, cat, y, s);
            return;
        }
        wrk_ok(wrk_conv3x3_fwd_up(cat.p, cat.pitch, c.cin, low.p, low.pitch, low.c, wb_[i]->h(), P_.f() + c.bOff, y.p,
                                  y.pitch, c.cout, c.relu ? 1u : 0u, n_, H(c.lvl), W(c.lvl), s),
               "wrk_conv3x3_fwd_up");
    }

    // rgb (inference, residual head): the output convolution writes the
    // unpacked image there instead of pred_
    void runForward(hipStream_t s, bool training, float* rgb = nullptr) {
        // encoder; each level's second convolution writes its skip slice
        conv(0, view(in_), view(a0_), s);
        const T e0 = view(cat0_, 64, 32);
        convPool(1, view(a0_), e0, view(p1_), s);
        conv(2, view(p1_), view(a1_), s);
        const T e1 = view(cat1_, 96, 64);
        convPool(3, view(a1_), e1, view(p2_), s);
        conv(4, view(p2_), view(a2_), s);
        const T e2 = view(cat2_, 128, 96);
        convPool(5, view(a2_), e2, view(p3_), s);
        conv(6, view(p3_), view(a3_), s);
        conv(7, view(a3_), view(b3_), s);
        // decoder: [upsampled | skip]
        upConv(8, view(b3_), view(cat2_), view(t2_), training, s);
        conv(9, view(t2_), view(d2_), s);
        upConv(10, view(d2_), view(cat1_), view(t1_), training, s);
        conv(11, view(t1_), view(d1_), s);
        upConv(12, view(d1_), view(cat0_), view(t0_), training, s);
        conv(13, view(t0_), view(d0_), s);
        if (rgb && head_ == 0) {
            const ConvSpec& c = conv_[14];
            wrk_ok(wrk_conv3x3_fwd_unpack(d0_.buf.h(), d0_.c, c.cin, wb_[14]->h(), P_.f() + c.bOff, c.cout, in_.buf.h(),
                                          32, rgb, n_, H(0), W(0), s),
                   "wrk_conv3x3_fwd_unpack");
        } else {
            conv(14, view(d0_), view(pred_), s);
        }
    }

    void runBackward(hipStream_t s) {
        T g;
        // output and level-0 decoder
        g = view(g0a_);
        convBack(14, view(d0_), view(gPred_), &g, s);
        reluBack(view(g0a_), view(d0_), s);
        g = view(g0b_);
        convBack(13, view(t0_), view(g0a_), &g, s);
        reluBack(view(g0b_), view(t0_), s);
        g = view(gCat0_);
        convBack(12, view(cat0_), view(g0b_), &g, s);
        wrk_ok(wrk_up2_bwd(gCat0_.buf.h(), 96, 64, g1a_.buf.h(), 64, n_, H(0), W(0), s), "wrk_up2_bwd");
        // level 1 decoder
        reluBack(view(g1a_), view(d1_), s);
        g = view(g1b_);
        convBack(11, view(t1_), view(g1a_), &g, s);
        reluBack(view(g1b_), view(t1_), s);
        g = view(gCat1_);
        convBack(10, view(cat1_), view(g1b_), &g, s);
        wrk_ok(wrk_up2_bwd(gCat1_.buf.h(), 160, 96, g2a_.buf.h(), 96, n_, H(1), W(1), s), "wrk_up2_bwd");
        // level 2 decoder
        reluBack(view(g2a_), view(d2_), s);
        g = view(g2b_);
        convBack(9, view(t2_), view(g2a_), &g, s);
        reluBack(view(g2b_), view(t2_), s);
        g = view(gCat2_);
        convBack(8, view(cat2_), view(g2b_), &g, s);
        wrk_ok(wrk_up2_bwd(gCat2_.buf.h(), 224, 128, g3a_.buf.h(), 128, n_, H(2), W(2), s), "wrk_up2_bwd");
        // bottleneck
        reluBack(view(g3a_), view(b3_), s);
        g = view(g3b_);
        convBack(7, view(a3_), view(g3a_), &g, s);
        reluBack(view(g3b_), view(a3_), s);
        g = view(gP3_);
        convBack(6, view(p3_), view(g3b_), &g, s);
        // level 2 encoder: pooled gradient + skip gradient
        wrk_ok(wrk_pool2_bwd(gP3_.buf.h(), 96, 96, gCat2_.buf.h() + 128, 224, g2a_.buf.h(), 96, n_, H(2), W(2), s),
               "wrk_pool2_bwd");
        reluBack(view(g2a_), view(cat2_, 128, 96), s);
        g = view(g2b_);
        convBack(5, view(a2_), view(g2a_), &g, s);
        reluBack(view(g2b_), view(a2_), s);
        g = view(gP2_);
        convBack(4, view(p2_), view(g2b_), &g, s);
        // level 1 encoder
        wrk_ok(wrk_pool2_bwd(gP2_.buf.h(), 64, 64, gCat1_.buf.h() + 96, 160, g1a_.buf.h(), 64, n_, H(1), W(1), s),
               "wrk_pool2_bwd");
        reluBack(view(g1a_), view(cat1_, 96, 64), s);
        g = view(g1b_);
        convBack(3, view(a1_), view(g1a_), &g, s);
        reluBack(view(g1b_), view(a1_), s);
        g = view(gP1_);
        convBack(2, view(p1_), view(g1b_), &g, s);
        // level 0 encoder
        wrk_ok(wrk_pool2_bwd(gP1_.buf.h(), 32, 32, gCat0_.buf.h() + 64, 96, g0a_.buf.h(), 32, n_, H(0), W(0), s),
               "wrk_pool2_bwd");
        reluBack(view(g0a_), view(cat0_, 64, 32), s);
        g = view(g0b_);
        convBack(1, view(a0_), view(g0a_), &g, s);
        reluBack(view(g0b_), view(a0_), s);
        convBack(0, view(in_), view(g0b_), nullptr, s);
    }

    void refreshBf16(hipStream_t s) {
        for (int i = 0; i < 15; ++i)
            wrk_ok(wrk_weights_bf16(P_.f() + conv_[i].wOff, conv_[i].cout, conv_[i].cin, wb_[i]->h(), wt_[i]->h(), s),
                   "wrk_weights_bf16");
    }

    uint32_t head_ = 0;
    std::vector<ConvSpec> conv_;
    uint64_t nParams_ = 0;
    uint32_t step_ = 0;
    Buf P_, G_, M_, V_;
    std::vector<std::unique_ptr<Buf>> wb_, wt_;
    uint32_t n_ = 0, h_ = 0, w_ = 0;
    Act in_, a0_, cat0_, t0_, d0_, pred_, p1_, a1_, cat1_, t1_, d1_, p2_, a2_, cat2_, t2_, d2_, p3_, a3_, b3_;
    Act gPred_, g0a_, g0b_, gCat0_, g1a_, g1b_, gCat1_, gP1_, g2a_, g2b_, gCat2_, gP2_, g3a_, g3b_, gP3_;
    Buf lossPartial_, wgradScratch_, feat_;
    uint32_t featCount_ = 0;
    const float* noise_ = nullptr;  // device noise layer of the source images, or none
};

thread_local std::string g_unetError;

template <typename F> int guard(F&& f) {
    try {
        f();
        return WRK_OK;
    } catch (const std::exception& e) {
        g_unetError = e.what();
        return WRK_ERR_INVALID;
    }
}

// A NULL network handle is an error, not a crash.
UNet* N(void* net) {
    if (!net) throw std::runtime_error("wrk_unet: null network handle");
    return static_cast<UNet*>(net);
}

}  // namespace

extern "C" {

const char* wrk_unet_last_error(void) { return g_unetError.c_str(); }

void* wrk_unet_create(void) { return wrk_unet_create_head(0); }
void* wrk_unet_create_head(uint32_t head) {
    UNet* u = nullptr;
    if (guard([&] { u = new UNet(head); }) != WRK_OK) return nullptr;
    return u;
}
uint32_t wrk_unet_head(void* net) {
    uint32_t h = 0;
    guard([&] { h = N(net)->head(); });
    return h;
}
uint64_t wrk_unet_params_for_head(uint32_t head) {
    uint64_t n = 0;
    guard([&] { n = UNet::paramsFor(head); });
    return n;
}
void wrk_unet_destroy(void* net) { delete static_cast<UNet*>(net); }
uint64_t wrk_unet_num_params(void* net) {
    uint64_t n = 0;
    guard([&] { n = N(net)->numParams(); });
    return n;
}
int wrk_unet_init(void* net, uint64_t seed) {
    return guard([&] { N(net)->init(seed); });
}
int wrk_unet_get_params(void* net, float* out) {
    return guard([&] { N(net)->getParams(out); });
}
int wrk_unet_set_params(void* net, const float* in) {
    return guard([&] { N(net)->setParams(in); });
}
int wrk_unet_get_state(void* net, float* m, float* v, uint32_t* step) {
    return guard([&] { N(net)->getState(m, v, step); });
}
int wrk_unet_set_state(void* net, const float* m, const float* v, uint32_t step) {
    return guard([&] { N(net)->setState(m, v, step); });
}
int wrk_unet_get_grads(void* net, float* out) {
    return guard([&] { N(net)->getGrads(out); });
}
int wrk_unet_set_sample_counts(void* net, const uint32_t* spp, uint32_t count) {
    return guard([&] { N(net)->setSampleCounts(spp, count); });
}
int wrk_unet_set_noise(void* net, const float* noise) {
    return guard([&] { N(net)->setNoise(noise); });
}
int wrk_unet_forward(void* net, const float* color, const float* albedo, const float* normal, uint32_t srcH,
                     uint32_t srcW, const uint32_t* origins, uint32_t n, uint32_t h, uint32_t w, float* out,
                     void* stream) {
    return guard([&] {
        N(net)->forward(color, albedo, normal, srcH, srcW, origins, n, h, w, out,
                                         static_cast<hipStream_t>(stream));
    });
}
int wrk_unet_train_step(void* net, const float* color, const float* albedo, const float* normal, const float* target,
                        const float* target2, uint32_t srcH, uint32_t srcW, const uint32_t* origins, uint32_t n,
                        uint32_t h, uint32_t w, float lr, uint32_t lossMode, float* loss, void* stream) {
    return guard([&] {
        float l = N(net)->trainStep(color, albedo, normal, target, target2, srcH, srcW, origins, n, h,
                                                     w, lr, lossMode, static_cast<hipStream_t>(stream));
        if (loss) *loss = l;
    });
}

}  // extern "C"
