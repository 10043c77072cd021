// C ABI of the learned density denoiser (pic_hip_denoiser_*, pic_hip_denoise,
// pic_hip_deposit_subsample; declared in include/pic_hip.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/pic_hip.h"
#include "api_util.hpp"
#include "context.hpp"
#include "pic_common.hpp"
#include "pic_kernels.hpp"

using namespace pichip;

struct pic_hip_denoiser {
    int H = 0, W = 0, max_batch = 0, max_pairs = 0;
    int dil[4] = {1, 3, 9, 27};  // balanced-ternary offsets: every shift in [-40, 40]^2 exactly once
    int ncu = 1;
    int step = 0;  // Adam steps taken
    // The network sees in_scale * delta and predicts out_scale * delta; both
    // are set from the training set's RMS at the first training (1 before).
    double in_scale = 1.0, out_scale = 1.0;
    bool scaled = false;
    hipStream_t stream = nullptr;
    CnnBuffers c{};
    float* inputs = nullptr;   // [max_pairs][H][W]
    float* targets = nullptr;  // [max_pairs][H][W]
    int n_in = 0, n_tg = 0;
    double* losses = nullptr;  // per-iteration loss of the last train call
    size_t losses_cap = 0;
    int4* items = nullptr;     // batch items of the last train call
    size_t items_cap = 0;
    double* diag_part = nullptr;  // rho sum scratch (5 * 1024)
    double* diag_out = nullptr;   // [8]
    double* scratch = nullptr;
    double* result = nullptr;     // [4]
    double* out64 = nullptr;      // [H][W]
    float2* spec_l = nullptr;     // [H][W/2+1] learned half spectrum of a step loop's solve
    double* h_result = nullptr;   // pinned [4]
    int* fallbacks = nullptr;     // gate fallbacks counted during pic_hip_run_learned
    // Work on the denoiser's buffers is enqueued on its own stream (training,
    // host calls) or on a context's stream (pic_hip_denoise); each hand-over
    // waits for the other side through these events.
    hipEvent_t ev_net = nullptr, ev_ctx = nullptr;

    ~pic_hip_denoiser() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (ev_net) (void)hipEventDestroy(ev_net);
        if (ev_ctx) (void)hipEventDestroy(ev_ctx);
        void* bufs[] = {c.w,       c.grad,      c.m,         c.v,       c.packed, c.x0,     c.out,     c.target,
                        c.dout,    c.a1,        c.a2,        c.a3,      c.d2,     c.d3,     c.wpart,   c.bias_part,
                        c.loss_part, inputs,    targets,     losses,    items,    diag_part, diag_out, scratch,
                        result,    out64,       fallbacks, spec_l};
        for (void* b : bufs) (void)hipFree(b);
        (void)hipHostFree(h_result);
        if (stream) (void)hipStreamDestroy(stream);
    }

    size_t plane() const { return (size_t)H * W; }
};

namespace {

pic_status check_net(pic_hip_denoiser* n) {
    if (!n) return fail(PIC_ERR_INVALID_ARGUMENT, "null denoiser");
    return PIC_OK;
}

pic_status check_pair(pic_hip_denoiser* n, pic_hip_ctx* ctx) {
    if (check_net(n) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!ctx) return fail(PIC_ERR_INVALID_ARGUMENT, "null context");
    if (ctx->dom.Nx != n->W || ctx->dom.Ny != n->H)
        return fail(PIC_ERR_INVALID_ARGUMENT, "context grid differs from the denoiser's");
    return PIC_OK;
}

CnnGeom geom(const pic_hip_denoiser* n, int batch) { return CnnGeom{batch, n->H, n->W}; }

// He-normal weights (std sqrt(2 / fan_in)), zero biases, last layer x 0.1.
std::vector<float> initial_params(uint64_t seed) {
    std::vector<float> w(cnn::kParams, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int count, int fan_in, double gain) {
        std::normal_distribution<double> nd(0.0, gain * std::sqrt(2.0 / fan_in));
        for (int i = 0; i < count; ++i) w[(size_t)(off + i)] = (float)nd(rng);
    };
    fill(cnn::kW1, 288, 9, 1.0);
    fill(cnn::kW2, 9216, 288, 1.0);
    fill(cnn::kW3, 9216, 288, 1.0);
    fill(cnn::kW4, 288, 288, 0.1);
    return w;
}

pic_status upload_params(pic_hip_denoiser* n, const float* p) {
    PIC_HIP_TRY(hipMemcpyAsync(n->c.w, p, cnn::kParams * sizeof(float), hipMemcpyHostToDevice, n->stream));
    launch_cnn_pack(n->c, n->stream);
    PIC_HIP_TRY(hipGetLastError());
    PIC_HIP_TRY(hipStreamSynchronize(n->stream));
    return PIC_OK;
}

// FP64 rho sum of the context into n->diag_out[0], then scale * delta into dst.
void enqueue_normalise(pic_hip_denoiser* n, pic_hip_ctx* ctx, float* dst, double scale) {
    launch_diagnostics(ctx->rho, nullptr, nullptr, ctx->G, nullptr, nullptr, nullptr, nullptr, 0, n->diag_part, 1024,
                       n->diag_out, ctx->stream);
    launch_normalize(ctx->rho, ctx->G, n->diag_out, scale, dst, ctx->stream);
}

double kappa_of(int scheme);

constexpr int kMomentBlocks = 128;

// Charge moments of the context's particles into net->scratch[0, 2 kMomentBlocks).
void enqueue_moments(pic_hip_denoiser* net, pic_hip_ctx* ctx) {
    launch_moments(ctx->s[4], (int64_t)ctx->n, net->scratch, kMomentBlocks, ctx->stream);
}

// Network, renormalisation and gate on the context's stream (kernel launches
// only).  A single call takes the charge moments and the sum of rho itself;
// inside the step loop (`in_run`) the moments were taken at the start of the
// run and the mean of rho comes from the particles' charge sum, which the
// deposit conserves.
void enqueue_denoise(pic_hip_ctx* ctx, pic_hip_denoiser* net, int scheme, double tolerance, int* fallback_count,
                     bool in_run) {
    const double cv = ctx->dev.dx * ctx->dev.dy;
    DenoiseArgs a{};
    a.rho = ctx->rho;
    a.out64 = net->out64;
    a.G = ctx->G;
    if (in_run) {
        a.sum = ctx->run_sums + 3;
        a.sum_scale = 1.0 / cv;
    } else {
        enqueue_moments(net, ctx);
        launch_diagnostics(ctx->rho, nullptr, nullptr, ctx->G, nullptr, nullptr, nullptr, nullptr, 0, net->diag_part,
                           1024, net->diag_out, ctx->stream);
        a.sum = net->diag_out;
        a.sum_scale = 1.0;
    }
    a.in_scale = net->in_scale;
    a.inv_out_scale = 1.0 / net->out_scale;
    a.moments = net->scratch;
    a.nbm = kMomentBlocks;
    a.cv = cv;
    a.kappa = kappa_of(scheme);
    a.tol = tolerance;
    a.stats = net->scratch + 2 * kMomentBlocks;
    a.result = net->result;
    a.fallback_count = fallback_count;
    // Inside a step loop whose solve runs in LDS, the gate and the write of the
    // learned density are left to the solve's first pass (ctx->pending_apply).
    const TileMerge* merge = nullptr;
    if (in_run && ctx->merge_pending) {
        merge = &ctx->pending_merge;
        ctx->merge_pending = false;
    }
    // Inside a step loop whose solve runs in LDS, its row pass also runs the
    // output layer where the grid allows.
    const bool defer = in_run && ctx->lds_fft;
    float2* spec_l = defer && poisson_rows_can_run_output_layer(net->W, net->H, net->dil[3]) ? net->spec_l : nullptr;
    launch_denoise(net->c, geom(net, 1), net->dil, net->ncu, a, ctx->stream, defer ? &ctx->pending_apply : nullptr,
                   merge, spec_l);
}

double kappa_of(int scheme) {
    switch (scheme) {
        case PIC_SCHEME_NGP: return 1.0;
        case PIC_SCHEME_TSC: return 0.55 * 0.55;
        default: return 4.0 / 9.0;  // CIC; the Esirkepov density of a short move is CIC-like
    }
}

}  // namespace

extern "C" {

pic_status pic_hip_denoiser_create(int32_t Nx, int32_t Ny, int32_t max_batch, int32_t max_pairs,
                                   const int32_t* dilation, uint64_t seed, pic_hip_denoiser** out) {
    if (!out) return fail(PIC_ERR_INVALID_ARGUMENT, "null output pointer");
    *out = nullptr;
    if (Nx < 32 || Ny < 4 || Nx % 32 != 0) return fail(PIC_ERR_INVALID_ARGUMENT, "need Nx a multiple of 32 and Ny >= 4");
    if ((int64_t)Nx * Ny > (1ll << 26)) return fail(PIC_ERR_INVALID_ARGUMENT, "grid above 2^26 cells");
    if (max_batch < 1 || max_batch > 256 || max_pairs < 0)
        return fail(PIC_ERR_INVALID_ARGUMENT, "max_batch must be 1..256 and max_pairs >= 0");
    std::unique_ptr<pic_hip_denoiser> n(new pic_hip_denoiser());
    n->H = Ny;
    n->W = Nx;
    n->max_batch = max_batch;
    n->max_pairs = max_pairs;
    for (int k = 0; k < 4; ++k) {
        if (dilation) n->dil[k] = dilation[k];
        if (n->dil[k] < 1 || n->dil[k] >= std::min(Nx, Ny))
            return fail(PIC_ERR_INVALID_ARGUMENT, "dilation must be >= 1 and below the grid size");
    }
    int dev = 0;
    hipDeviceProp_t prop;
    PIC_HIP_TRY(hipGetDevice(&dev));
    PIC_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    n->ncu = std::max(1, prop.multiProcessorCount);
    PIC_HIP_TRY(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    PIC_HIP_TRY(hipEventCreateWithFlags(&n->ev_net, hipEventDisableTiming));
    PIC_HIP_TRY(hipEventCreateWithFlags(&n->ev_ctx, hipEventDisableTiming));
    const size_t P = n->plane(), NB = P * (size_t)max_batch;
    CnnBuffers& c = n->c;
    PIC_HIP_TRY(dmalloc(&c.w, cnn::kParams));
    PIC_HIP_TRY(dmalloc(&c.grad, cnn::kParams));
    PIC_HIP_TRY(dmalloc(&c.m, cnn::kParams));
    PIC_HIP_TRY(dmalloc(&c.v, cnn::kParams));
    PIC_HIP_TRY(hipMemsetAsync(c.m, 0, cnn::kParams * sizeof(float), n->stream));
    PIC_HIP_TRY(hipMemsetAsync(c.v, 0, cnn::kParams * sizeof(float), n->stream));
    uint16_t* pk = nullptr;
    PIC_HIP_TRY(dmalloc(&pk, (size_t)cnn::kPackVecs * 8));
    c.packed = pk;
    PIC_HIP_TRY(dmalloc(&c.x0, NB));
    PIC_HIP_TRY(dmalloc(&c.out, NB));
    PIC_HIP_TRY(dmalloc(&c.target, NB));
    PIC_HIP_TRY(dmalloc(&c.dout, NB));
    uint16_t* acts[5] = {};
    for (auto& a : acts) PIC_HIP_TRY(dmalloc(&a, NB * 32));
    c.a1 = acts[0];
    c.a2 = acts[1];
    c.a3 = acts[2];
    c.d2 = acts[3];
    c.d3 = acts[4];
    PIC_HIP_TRY(dmalloc(&c.wpart, cnn_wpart_floats(n->ncu)));
    PIC_HIP_TRY(dmalloc(&c.bias_part, (size_t)256 * 32));
    PIC_HIP_TRY(dmalloc(&c.loss_part, 256));
    PIC_HIP_TRY(dmalloc(&n->inputs, P * std::max(max_pairs, 1)));
    PIC_HIP_TRY(dmalloc(&n->targets, P * std::max(max_pairs, 1)));
    PIC_HIP_TRY(dmalloc(&n->diag_part, (size_t)5 * 1024));
    PIC_HIP_TRY(dmalloc(&n->diag_out, 8));
    PIC_HIP_TRY(dmalloc(&n->scratch,
                        std::max<size_t>(4 * kMomentBlocks, 2 * kMomentBlocks + denoise_stats_doubles(n->ncu))));
    PIC_HIP_TRY(dmalloc(&n->result, 4));
    PIC_HIP_TRY(dmalloc(&n->out64, P));
    PIC_HIP_TRY(dmalloc(&n->spec_l, (size_t)n->H * (n->W / 2 + 1)));
    PIC_HIP_TRY(dmalloc(&n->fallbacks, 1));
    PIC_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&n->h_result), 4 * sizeof(double), hipHostMallocDefault));
    const std::vector<float> w0 = initial_params(seed);
    pic_status st = upload_params(n.get(), w0.data());
    if (st != PIC_OK) return st;
    *out = n.release();
    return PIC_OK;
}

pic_status pic_hip_denoiser_destroy(pic_hip_denoiser* net) {
    delete net;
    return PIC_OK;
}

pic_status pic_hip_denoiser_get_params(pic_hip_denoiser* net, float* params) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!params) return fail(PIC_ERR_INVALID_ARGUMENT, "null params");
    PIC_HIP_TRY(hipMemcpyAsync(params, net->c.w, cnn::kParams * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    PIC_HIP_TRY(hipStreamSynchronize(net->stream));
    return PIC_OK;
}

pic_status pic_hip_denoiser_set_params(pic_hip_denoiser* net, const float* params) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!params) return fail(PIC_ERR_INVALID_ARGUMENT, "null params");
    PIC_HIP_TRY(hipMemsetAsync(net->c.m, 0, cnn::kParams * sizeof(float), net->stream));
    PIC_HIP_TRY(hipMemsetAsync(net->c.v, 0, cnn::kParams * sizeof(float), net->stream));
    net->step = 0;
    return upload_params(net, params);
}

pic_status pic_hip_denoiser_forward_host(pic_hip_denoiser* net, const float* input, float* output, int32_t batch) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!input || !output) return fail(PIC_ERR_INVALID_ARGUMENT, "null image");
    if (batch < 1 || batch > net->max_batch) return fail(PIC_ERR_INVALID_ARGUMENT, "batch outside 1..max_batch");
    const size_t bytes = net->plane() * batch * sizeof(float);
    PIC_HIP_TRY(hipMemcpyAsync(net->c.x0, input, bytes, hipMemcpyHostToDevice, net->stream));
    launch_cnn_forward(net->c, geom(net, batch), net->dil, net->ncu, net->stream);
    PIC_HIP_TRY(hipGetLastError());
    PIC_HIP_TRY(hipMemcpyAsync(output, net->c.out, bytes, hipMemcpyDeviceToHost, net->stream));
    PIC_HIP_TRY(hipStreamSynchronize(net->stream));
    return PIC_OK;
}

pic_status pic_hip_denoiser_gradient_host(pic_hip_denoiser* net, const float* input, const float* target,
                                          int32_t batch, double* loss, float* grad) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!input || !target) return fail(PIC_ERR_INVALID_ARGUMENT, "null image");
    if (batch < 1 || batch > net->max_batch) return fail(PIC_ERR_INVALID_ARGUMENT, "batch outside 1..max_batch");
    const size_t bytes = net->plane() * batch * sizeof(float);
    const CnnGeom g = geom(net, batch);
    PIC_HIP_TRY(hipMemcpyAsync(net->c.x0, input, bytes, hipMemcpyHostToDevice, net->stream));
    PIC_HIP_TRY(hipMemcpyAsync(net->c.target, target, bytes, hipMemcpyHostToDevice, net->stream));
    launch_cnn_forward(net->c, g, net->dil, net->ncu, net->stream);
    launch_cnn_loss(net->c, g, net->diag_out, net->stream);
    launch_cnn_backward(net->c, g, net->dil, net->ncu, net->stream);
    PIC_HIP_TRY(hipGetLastError());
    if (loss) PIC_HIP_TRY(hipMemcpyAsync(loss, net->diag_out, sizeof(double), hipMemcpyDeviceToHost, net->stream));
    if (grad)
        PIC_HIP_TRY(hipMemcpyAsync(grad, net->c.grad, cnn::kParams * sizeof(float), hipMemcpyDeviceToHost,
                                   net->stream));
    PIC_HIP_TRY(hipStreamSynchronize(net->stream));
    return PIC_OK;
}

pic_status pic_hip_denoiser_add_sample(pic_hip_denoiser* net, pic_hip_ctx* ctx, int32_t is_target) {
    if (check_pair(net, ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    int& count = is_target ? net->n_tg : net->n_in;
    if (count >= net->max_pairs) return fail(PIC_ERR_INVALID_ARGUMENT, "sample set is full (max_pairs)");
    float* dst = (is_target ? net->targets : net->inputs) + (size_t)count * net->plane();
    PIC_HIP_TRY(hipEventRecord(net->ev_net, net->stream));
    PIC_HIP_TRY(hipStreamWaitEvent(ctx->stream, net->ev_net, 0));
    enqueue_normalise(net, ctx, dst, 1.0);
    PIC_HIP_TRY(hipGetLastError());
    PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ++count;
    return PIC_OK;
}

pic_status pic_hip_denoiser_get_scales(pic_hip_denoiser* net, double* in_scale, double* out_scale) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (in_scale) *in_scale = net->in_scale;
    if (out_scale) *out_scale = net->out_scale;
    return PIC_OK;
}

pic_status pic_hip_denoiser_set_scales(pic_hip_denoiser* net, double in_scale, double out_scale) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!(in_scale > 0.0) || !(out_scale > 0.0) || !std::isfinite(in_scale) || !std::isfinite(out_scale))
        return fail(PIC_ERR_INVALID_ARGUMENT, "scales must be positive and finite");
    net->in_scale = in_scale;
    net->out_scale = out_scale;
    net->scaled = true;
    return PIC_OK;
}

int32_t pic_hip_denoiser_pairs(pic_hip_denoiser* net) { return net ? std::min(net->n_in, net->n_tg) : 0; }

pic_status pic_hip_denoiser_clear_samples(pic_hip_denoiser* net) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    net->n_in = net->n_tg = 0;
    return PIC_OK;
}

pic_status pic_hip_denoiser_train(pic_hip_denoiser* net, int32_t iterations, int32_t batch, double lr, uint64_t seed,
                                  double* loss_history) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (iterations < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "iterations must be >= 1");
    if (batch < 1 || batch > net->max_batch) return fail(PIC_ERR_INVALID_ARGUMENT, "batch outside 1..max_batch");
    if (!(lr > 0.0) || !std::isfinite(lr)) return fail(PIC_ERR_INVALID_ARGUMENT, "learning rate must be positive");
    const int pairs = pic_hip_denoiser_pairs(net);
    if (pairs < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "no training pairs (pic_hip_denoiser_add_sample)");
    const size_t nitems = (size_t)iterations * batch;
    if (nitems > net->items_cap) {
        (void)hipFree(net->items);
        net->items = nullptr;
        net->items_cap = 0;
        PIC_HIP_TRY(dmalloc(&net->items, nitems));
        net->items_cap = nitems;
    }
    if ((size_t)iterations > net->losses_cap) {
        (void)hipFree(net->losses);
        net->losses = nullptr;
        net->losses_cap = 0;
        PIC_HIP_TRY(dmalloc(&net->losses, (size_t)iterations));
        net->losses_cap = (size_t)iterations;
    }
    if (!net->scaled) {
        // RMS of the input and target sets (FP64 sums on the device).
        const int nb = kMomentBlocks;
        const size_t len = (size_t)pairs * net->plane();
        double* part = net->scratch;  // 4 * kMomentBlocks doubles
        launch_moments(net->inputs, (int64_t)len, part, nb, net->stream);
        launch_moments(net->targets, (int64_t)len, part + 2 * nb, nb, net->stream);
        std::vector<double> h(4 * nb);
        PIC_HIP_TRY(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, net->stream));
        PIC_HIP_TRY(hipStreamSynchronize(net->stream));
        double s2[2] = {0.0, 0.0};
        for (int k = 0; k < 2; ++k)
            for (int b = 0; b < nb; ++b) s2[k] += h[(size_t)(2 * nb * k + 2 * b + 1)];
        const double rin = std::sqrt(s2[0] / len), rtg = std::sqrt(s2[1] / len);
        if (rin > 0.0 && rtg > 0.0) {
            net->in_scale = 1.0 / rin;
            net->out_scale = 1.0 / rtg;
        }
        net->scaled = true;
    }
    std::mt19937_64 rng(seed);
    std::vector<int4> h(nitems);
    for (auto& it : h) {
        it.x = (int)(rng() % (uint64_t)pairs);
        it.y = (int)(rng() % (uint64_t)net->H);
        it.z = (int)(rng() % (uint64_t)net->W);
        it.w = 0;
    }
    PIC_HIP_TRY(hipMemcpyAsync(net->items, h.data(), nitems * sizeof(int4), hipMemcpyHostToDevice, net->stream));
    const CnnGeom g = geom(net, batch);
    for (int it = 0; it < iterations; ++it) {
        const double frac = iterations > 1 ? (double)it / (iterations - 1) : 0.0;
        const double lr_it = lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac)));
        launch_cnn_batch(net->inputs, net->targets, net->items + (size_t)it * batch, g, (float)net->in_scale,
                         (float)net->out_scale, net->c.x0, net->c.target, net->stream);
        launch_cnn_forward(net->c, g, net->dil, net->ncu, net->stream);
        launch_cnn_loss(net->c, g, net->losses + it, net->stream);
        launch_cnn_backward(net->c, g, net->dil, net->ncu, net->stream);
        launch_cnn_adam(net->c, (float)lr_it, 0.9f, 0.999f, 1e-8f, ++net->step, net->stream);
        launch_cnn_pack(net->c, net->stream);
    }
    PIC_HIP_TRY(hipGetLastError());
    if (loss_history)
        PIC_HIP_TRY(hipMemcpyAsync(loss_history, net->losses, (size_t)iterations * sizeof(double),
                                   hipMemcpyDeviceToHost, net->stream));
    PIC_HIP_TRY(hipStreamSynchronize(net->stream));  // h (host items) stays alive until here
    return PIC_OK;
}

pic_status pic_hip_denoise(pic_hip_ctx* ctx, pic_hip_denoiser* net, int scheme, double tolerance, int32_t* fallback,
                           double* residual_ratio) {
    if (check_pair(net, ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (scheme < 0 || scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "unknown scheme");
    if (!(tolerance >= 0.0)) return fail(PIC_ERR_INVALID_ARGUMENT, "tolerance must be >= 0");
    hipStream_t s = ctx->stream;
    PIC_HIP_TRY(hipEventRecord(net->ev_net, net->stream));
    PIC_HIP_TRY(hipStreamWaitEvent(s, net->ev_net, 0));
    enqueue_denoise(ctx, net, scheme, tolerance, nullptr, false);
    PIC_HIP_TRY(hipGetLastError());
    PIC_HIP_TRY(hipEventRecord(net->ev_ctx, s));
    PIC_HIP_TRY(hipStreamWaitEvent(net->stream, net->ev_ctx, 0));
    if (fallback || residual_ratio) {
        PIC_HIP_TRY(hipMemcpyAsync(net->h_result, net->result, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
        PIC_HIP_TRY(hipStreamSynchronize(s));
        if (fallback) *fallback = net->h_result[1] != 0.0 ? 1 : 0;
        if (residual_ratio) *residual_ratio = net->h_result[0];
    }
    return PIC_OK;
}

pic_status pic_hip_run_learned(pic_hip_ctx* ctx, pic_hip_denoiser* net, const pic_step_config* cfg, double tolerance,
                               int32_t nsteps, int32_t record_from, double* field_energy, double* mode1_amplitude,
                               double* kinetic_energy, double* charge_error, int32_t* fallbacks) {
    if (check_pair(net, ctx) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!cfg) return fail(PIC_ERR_INVALID_ARGUMENT, "null step config");
    if (!(tolerance >= 0.0)) return fail(PIC_ERR_INVALID_ARGUMENT, "tolerance must be >= 0");
    PIC_HIP_TRY(hipEventRecord(net->ev_net, net->stream));
    PIC_HIP_TRY(hipStreamWaitEvent(ctx->stream, net->ev_net, 0));
    launch_set_int(net->fallbacks, 0, ctx->stream);
    enqueue_moments(net, ctx);  // the particles' charges do not change during the run
    const int scheme = cfg->scheme;
    const StepHook hook = [ctx, net, scheme, tolerance]() -> pic_status {
        enqueue_denoise(ctx, net, scheme, tolerance, net->fallbacks, true);
        PIC_HIP_TRY(hipGetLastError());
        return PIC_OK;
    };
    // The first layer merges the deposit's tile windows (one launch fewer per
    // step) when its dilation is 1.
    ctx->defer_merge = denoise_can_merge(geom(net, 1), net->dil);
    ctx->merge_pending = false;
    pic_status st = run_steps(ctx, cfg, nsteps, record_from, field_energy, mode1_amplitude, kinetic_energy,
                              charge_error, hook);
    ctx->defer_merge = false;
    ctx->merge_pending = false;
    if (st != PIC_OK) return st;
    if (fallbacks) {
        PIC_HIP_TRY(hipMemcpyAsync(fallbacks, net->fallbacks, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        PIC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    PIC_HIP_TRY(hipEventRecord(net->ev_ctx, ctx->stream));
    PIC_HIP_TRY(hipStreamWaitEvent(net->stream, net->ev_ctx, 0));
    return PIC_OK;
}

pic_status pic_hip_denoiser_get_output_f64(pic_hip_denoiser* net, double* rho) {
    if (check_net(net) != PIC_OK) return PIC_ERR_INVALID_ARGUMENT;
    if (!rho) return fail(PIC_ERR_INVALID_ARGUMENT, "null rho");
    // The net stream waits for the last pic_hip_denoise (ev_ctx).
    PIC_HIP_TRY(hipMemcpyAsync(rho, net->out64, net->plane() * sizeof(double), hipMemcpyDeviceToHost, net->stream));
    PIC_HIP_TRY(hipStreamSynchronize(net->stream));
    return PIC_OK;
}

pic_status pic_hip_deposit_subsample(pic_hip_ctx* ctx, int scheme, uint32_t stride, uint64_t salt) {
    if (!ctx) return fail(PIC_ERR_INVALID_ARGUMENT, "null context");
    if (scheme < 0 || scheme > 3) return fail(PIC_ERR_INVALID_ARGUMENT, "unknown scheme");
    if (stride < 1) return fail(PIC_ERR_INVALID_ARGUMENT, "stride must be >= 1");
    DepositArgs a;
    a.x = ctx->s[0];
    a.y = ctx->s[1];
    a.vx = ctx->s[2];
    a.vy = ctx->s[3];
    a.q = ctx->s[4];
    a.n = (int64_t)ctx->n;
    a.cell_off = nullptr;
    a.dom = ctx->dev;
    a.scheme = scheme;
    a.backend = PIC_BACKEND_GPU_ATOMICS;
    a.ppl = 4;
    a.pipe = 0;
    a.geom = ctx->geom[scheme];
    a.dt = 0.0f;
    a.rho = ctx->rho;
    a.partial = nullptr;
    a.acc = ctx->acc64;
    a.status = ctx->status;
    a.stream = ctx->stream;
    PIC_HIP_TRY(launch_deposit_subsample(a, stride, salt, (float)stride));
    return PIC_OK;
}

}  // extern "C"
