/* wrk_unet.h — C ABI of the learned denoiser (no counterpart function in
 * libWR: it stands in for the closed OptiX AI denoiser that
 * libWR/src/internal/denoiser.cpp:17-195 drives with colour + albedo +
 * normal layers, renderer.cpp:289-298).
 *
 * Tensors are NHWC bf16 (passed as uint16_t*) with a channel pitch;
 * convolution weights are bf16 [9][cout][cin] (tap = 3*ky + kx), biases
 * FP32. Every function takes an explicit stream and returns WRK_OK or an
 * error code with wrk_last_error() (kernel level) / wrk_unet_last_error()
 * (network level).
 */
#ifndef WRK_UNET_H
#define WRK_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- kernels */
/* y[p][co] = act(bias[co] + sum_tap,ci w[tap][co][ci] x[p + (ky-1, kx-1)][ci]),
 * zero padding, act = ReLU when relu != 0. cin, cout multiples of 32. */
int wrk_conv3x3_fwd(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w, const float* bias,
                    uint16_t* y, uint32_t yPitch, uint32_t cout, uint32_t relu, uint32_t n, uint32_t h, uint32_t wd,
                    void* stream);
/* The same convolution, computing only output channels [0, c) with c =
 * coutUsed rounded up to 16 (at most cout); channels c.. of y are not
 * written. w has cout output channels. For an output layer whose consumer
 * reads its first few channels. */
int wrk_conv3x3_fwd_first(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w, const float* bias,
                          uint16_t* y, uint32_t yPitch, uint32_t cout, uint32_t coutUsed, uint32_t relu, uint32_t n,
                          uint32_t h, uint32_t wd, void* stream);
/* The same convolution over a decoder input [2x upsampled xUp | skip]:
 * input channels [0, cUp) are the nearest 2x upsampling of xUp ((n, h/2,
 * wd/2), channel pitch upPitch), read from xUp while the kernel stages its
 * input, and channels [cUp, cin) are read from x at those channel positions
 * (x's first cUp channels are not read). Equal to wrk_up2_fwd of xUp into
 * x's first cUp channels followed by wrk_conv3x3_fwd. h and wd even, cUp a
 * multiple of 32 in 32..cin. */
int wrk_conv3x3_fwd_up(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* xUp, uint32_t upPitch,
                       uint32_t cUp, const uint16_t* w, const float* bias, uint16_t* y, uint32_t yPitch, uint32_t cout,
                       uint32_t relu, uint32_t n, uint32_t h, uint32_t wd, void* stream);
/* wrk_conv3x3_fwd, and the 2x2 average pool of its bf16 output into yPool
 * ((n, h/2, wd/2), channel pitch poolPitch) from the same epilogue: equal
 * to wrk_conv3x3_fwd followed by wrk_pool2_fwd of y. h and wd even. */
int wrk_conv3x3_fwd_pool(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w, const float* bias,
                         uint16_t* y, uint32_t yPitch, uint32_t cout, uint32_t relu, uint16_t* yPool,
                         uint32_t poolPitch, uint32_t n, uint32_t h, uint32_t wd, void* stream);
/* The residual head's output convolution and wrk_unet_unpack in one pass:
 * out (n*h*wd, 3) FP32 = max(expm1(y_c + base_c), 0) for c < 3, y the bf16
 * convolution output (no activation) over output channels 0..15 of w
 * (cout output channels), base the packed input (channel pitch basePitch).
 * Equal to wrk_conv3x3_fwd_first (coutUsed 3, relu 0) followed by
 * wrk_unet_unpack; no tensor y is written. cin = 32. */
int wrk_conv3x3_fwd_unpack(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* w, const float* bias,
                           uint32_t cout, const uint16_t* base, uint32_t basePitch, float* out, uint32_t n,
                           uint32_t h, uint32_t wd, void* stream);
/* dw[tap][co][ci] = sum_p dz[p][co] x[p + off(tap)][ci] (FP32, overwritten);
 * db[co] = sum_p dz[p][co] when db != NULL. */
uint64_t wrk_conv3x3_wgrad_scratch_bytes(uint32_t cin, uint32_t cout, uint32_t n, uint32_t h, uint32_t w);
int wrk_conv3x3_wgrad(const uint16_t* x, uint32_t xPitch, uint32_t cin, const uint16_t* dz, uint32_t dzPitch,
                      uint32_t cout, uint32_t n, uint32_t h, uint32_t w, float* dw, float* db, void* scratch,
                      uint64_t scratchBytes, void* stream);
/* FP32 [9][cout][cin] -> bf16 w (same layout) and wt[8 - tap][ci][co]
 * (the data-gradient weights; may be NULL). */
int wrk_weights_bf16(const float* w32, uint32_t cout, uint32_t cin, uint16_t* w, uint16_t* wt, void* stream);
/* dy = y > 0 ? dy : 0 over npix pixels of c channels */
int wrk_relu_bwd(uint16_t* dy, uint32_t dyPitch, const uint16_t* y, uint32_t yPitch, uint32_t c, uint64_t npix,
                 void* stream);
/* 2x2 average pool of (n, h, w) -> (n, h/2, w/2); backward adds the
 * optional skip gradient: dx = dy / 4 + skip. */
int wrk_pool2_fwd(const uint16_t* x, uint32_t xPitch, uint32_t c, uint16_t* y, uint32_t yPitch, uint32_t n,
                  uint32_t h, uint32_t w, void* stream);
int wrk_pool2_bwd(const uint16_t* dy, uint32_t dyPitch, uint32_t c, const uint16_t* skip, uint32_t skipPitch,
                  uint16_t* dx, uint32_t dxPitch, uint32_t n, uint32_t h, uint32_t w, void* stream);
/* nearest 2x upsample to (n, h, w); backward sums each 2x2 block */
int wrk_up2_fwd(const uint16_t* x, uint32_t xPitch, uint32_t c, uint16_t* y, uint32_t yPitch, uint32_t n, uint32_t h,
                uint32_t w, void* stream);
int wrk_up2_bwd(const uint16_t* dy, uint32_t dyPitch, uint32_t c, uint16_t* dx, uint32_t dxPitch, uint32_t n,
                uint32_t h, uint32_t w, void* stream);
/* Network input (32 channels): log(1 + max(colour, 0)), albedo, normal,
 * a per-image feature (channel 9: feature[image] for image < numFeature,
 * else 0; the network's is log2(spp) / 16), a per-pixel noise level
 * (channel 10: log(1 + max(noise, 0) / (1 + max(y, 0))), y the mean of the
 * pixel's three colour channels; 0 when noise is NULL), zeros, from FP32
 * (img, srcH, srcW, 3) layers and the (img, srcH, srcW) noise layer (the
 * standard error of mean y, wrk_noise_stderr); origins (device, 3 per
 * crop: image, x0, y0) select n crops of h x w, or NULL for whole images, in
 * which case h >= srcH and w >= srcW: positions past the last source row or
 * column repeat it (edge padding). feature is a device array or NULL. */
int wrk_unet_pack(const float* color, const float* albedo, const float* normal, const float* noise, uint32_t srcH,
                  uint32_t srcW, const uint32_t* origins, const float* feature, uint32_t numFeature, uint16_t* out,
                  uint32_t n, uint32_t h, uint32_t w, void* stream);
/* out (npix, 3) FP32 = max(exp(q) - 1, 0) of channels 0..2, q = pred, or
 * pred + base (the packed input: a correction of its log colour) */
int wrk_unet_unpack(const uint16_t* pred, uint32_t pitch, const uint16_t* base, uint32_t basePitch, float* out,
                    uint64_t npix, void* stream);
/* per-256-pixel partial sums of the loss and its gradient dpred, with q as
 * in wrk_unet_unpack and y = exp(q) - 1; mode 0: (q - log(1 + t))^2,
 * mode 1: relMSE (y - t)^2 / (t^2 + 0.01), mode 2: (y - t)^2 / (y^2 + 0.01)
 * with the denominator constant in the gradient, mode 3: (y - t)^2 /
 * (t2^2 + 0.01) with t2 from target2 (a target over disjoint samples) */
uint64_t wrk_unet_loss_scratch_bytes(uint32_t n, uint32_t h, uint32_t w);
int wrk_unet_loss(const uint16_t* pred, uint32_t pitch, const float* target, uint32_t srcH, uint32_t srcW,
                  const uint32_t* origins, const float* target2, const uint16_t* base, uint32_t basePitch,
                  uint16_t* dpred, uint32_t n, uint32_t h, uint32_t w, uint32_t mode, float* partial, void* stream);
/* Kernel-predicting output (Bako et al. 2017), in the log domain:
 * channels 0 .. wrk_kpn_taps() - 1 of pred (bf16, channel pitch >=
 * wrk_kpn_taps() + 3) are the logits of a 7x7 filter per pixel (row-major,
 * dy then dx) and the next 3 channels a per-channel offset d. out (n, h, w,
 * 3) FP32 = max(exp(u) - 1, 0), u = sum_k softmax(logits)_k log(1 +
 * max(c_k, 0)) + d over the 7x7 source neighbourhood of each pixel of the
 * noisy colour (img, srcH, srcW, 3); origins as in wrk_unet_pack;
 * neighbours past the image edges repeat its edge pixels. */
uint32_t wrk_kpn_taps(void);
int wrk_kpn_apply(const uint16_t* pred, uint32_t pitch, const float* color, uint32_t srcH, uint32_t srcW,
                  const uint32_t* origins, float* out, uint32_t n, uint32_t h, uint32_t w, void* stream);
/* the losses of wrk_unet_loss (modes 0..3) with q = u of wrk_kpn_apply,
 * per-256-pixel partial sums, and their gradient with respect to the
 * logits and offsets in dpred (0 beyond them) */
int wrk_kpn_loss(const uint16_t* pred, uint32_t pitch, const float* color, const float* target, uint32_t srcH,
                 uint32_t srcW, const uint32_t* origins, const float* target2, uint16_t* dpred, uint32_t n,
                 uint32_t h, uint32_t w, uint32_t mode, float* partial, void* stream);
/* Adam on FP32 parameters (step counts from 1) */
int wrk_adam(float* p, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2, float eps,
             uint32_t step, void* stream);

/* ------------------------------------------------------------- network */
/* The 4-level U-Net (channels 32/64/96/128, ~0.9 M parameters). Device
 * pointers for all image data; params as one FP32 vector. */
const char* wrk_unet_last_error(void);
void* wrk_unet_create(void);
/* head 0 (wrk_unet_create): output channels 0..2 correct the input's
 * log(1 + colour) (922,656 parameters); head 1: a 64-channel output whose
 * first wrk_kpn_taps() + 3 channels are the per-pixel filter and offsets of
 * wrk_kpn_apply (931,904 parameters). */
void* wrk_unet_create_head(uint32_t head);
uint32_t wrk_unet_head(void* net);
/* parameter count of a network with this head (0 for an unknown head) */
uint64_t wrk_unet_params_for_head(uint32_t head);
void wrk_unet_destroy(void* net);
uint64_t wrk_unet_num_params(void* net);
int wrk_unet_init(void* net, uint64_t seed);
int wrk_unet_get_params(void* net, float* hostOut);
int wrk_unet_set_params(void* net, const float* hostIn);
/* Adam state, for training checkpoints: first and second moments (layout
 * of the parameters, NULL skips one) and the number of steps taken. No
 * counterpart in the reference, which runs OptiX's pretrained denoiser and
 * trains nothing; a resumed run continues bit for bit. */
int wrk_unet_get_state(void* net, float* mOut, float* vOut, uint32_t* stepOut);
int wrk_unet_set_state(void* net, const float* m, const float* v, uint32_t step);
/* gradient of the last wrk_unet_train_step (layout of the parameters) */
int wrk_unet_get_grads(void* net, float* hostOut);
/* Sample counts of the source images (host array, count entries) that the
 * next forward / training calls see: input channel 9 = log2(spp) / 16 of
 * each pixel's image (0 for images beyond count, or before any call). */
int wrk_unet_set_sample_counts(void* net, const uint32_t* spp, uint32_t count);
/* Noise layer of the source images (device, (img, srcH, srcW) FP32 as from
 * wrk_noise_stderr, or NULL) that the next forward / training calls pack
 * into input channel 10. Weights trained without it hold zeros for that
 * channel (its input was 0, so Adam never moved them) and are unaffected.
 * The pointer is kept, not copied: it must stay valid while in use. */
int wrk_unet_set_noise(void* net, const float* noise);
/* denoised colour (n, h, w, 3) FP32 into out; h, w multiples of 8. Without
 * origins, h x w may exceed srcH x srcW (edge padding, see wrk_unet_pack). */
int wrk_unet_forward(void* net, const float* color, const float* albedo, const float* normal, uint32_t srcH,
                     uint32_t srcW, const uint32_t* origins, uint32_t n, uint32_t h, uint32_t w, float* out,
                     void* stream);
/* one Adam step on the loss of wrk_unet_loss (lossMode) against target;
 * *loss (host) = the loss before the step */
int wrk_unet_train_step(void* net, const float* color, const float* albedo, const float* normal, const float* target,
                        const float* target2, uint32_t srcH, uint32_t srcW, const uint32_t* origins, uint32_t n,
                        uint32_t h, uint32_t w, float lr, uint32_t lossMode, float* loss, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WRK_UNET_H */
