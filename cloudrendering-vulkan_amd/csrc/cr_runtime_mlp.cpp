// Host runtime of the learned in-scattering network (extension): parameter
// storage (fp32 masters + bf16 copies + Adam state), forward, one training
// step and the MFMA layout probe.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cr_internal_mlp.h"
#include "cr_runtime.h"

using namespace cr;

struct Mlp {
    uint32_t wOff[MLP_LAYERS], bOff[MLP_LAYERS], total = 0;
    float* params = nullptr;   // fp32 masters
    float* grad = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* W[MLP_LAYERS] = {};
    uint16_t* WT[MLP_LAYERS] = {};
    uint64_t step = 0;
    bool initialized = false;  // parameters set (cr_mlp_init / cr_mlp_set_params)
    // scratch of the last batch size
    uint16_t* X = nullptr;
    size_t XCap = 0;
    float* T = nullptr;
    size_t TCap = 0;
    float* Y = nullptr;
    size_t YCap = 0;
    uint16_t* tr = nullptr;    // feature-major layer inputs and output gradients of the last step
    size_t trCap = 0;
    float* loss = nullptr;     // the last training step's per-tile loss shares
    size_t lossCap = 0;
    unsigned lossBlocks = 0;
    float* slot = nullptr;     // gradient slots of the weight-gradient GEMM
    size_t slotCap = 0;
    // moving average of the parameters (allocated when first turned on)
    float emaDecay = 0.0f;     // 0: off
    uint64_t emaSteps = 0;     // steps averaged since the last restart
    float* ema = nullptr;
    float* emaHat = nullptr;   // bias-corrected average, used for inference
    uint16_t* EW[MLP_LAYERS] = {};
    uint16_t* EWT[MLP_LAYERS] = {};
};

namespace {

void layout(uint32_t wOff[MLP_LAYERS], uint32_t bOff[MLP_LAYERS], uint32_t* total)
{
    uint32_t o = 0;
    for (int l = 0; l < MLP_LAYERS; l++) {
        wOff[l] = o;
        o += MLP_K[l] * MLP_N[l];
        bOff[l] = o;
        o += MLP_N[l];
    }
    *total = o;
}

uint16_t f2bf_host(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Real (unpadded) output width of layer l.
uint32_t real_n(int l) { return l == MLP_LAYERS - 1 ? 3u : MLP_N[l]; }

cr_status mlp_get(cr_context* ctx, Mlp** out)
{
    if (!ctx->mlp) {
        Mlp* m = new Mlp();
        layout(m->wOff, m->bOff, &m->total);
        ctx->mlp = m;
        size_t cap = 0;
        cr_status st = ensure_buffer(ctx, (void**)&m->params, &cap, (size_t)m->total * 4, "mlp params");
        if (st == CR_OK) { cap = 0; st = ensure_buffer(ctx, (void**)&m->grad, &cap, (size_t)m->total * 4, "mlp grad"); }
        if (st == CR_OK) { cap = 0; st = ensure_buffer(ctx, (void**)&m->m, &cap, (size_t)m->total * 4, "mlp m"); }
        if (st == CR_OK) { cap = 0; st = ensure_buffer(ctx, (void**)&m->v, &cap, (size_t)m->total * 4, "mlp v"); }
        for (int l = 0; l < MLP_LAYERS && st == CR_OK; l++) {
            cap = 0;
            st = ensure_buffer(ctx, (void**)&m->W[l], &cap, (size_t)MLP_K[l] * MLP_N[l] * 2, "mlp W");
            cap = 0;
            if (st == CR_OK) st = ensure_buffer(ctx, (void**)&m->WT[l], &cap, (size_t)MLP_K[l] * MLP_N[l] * 2, "mlp WT");
        }
        if (st != CR_OK) return st;
    }
    *out = ctx->mlp;
    return CR_OK;
}

// Upload fp32 masters (padded layout, padding forced to zero) and their bf16 copies.
cr_status upload_params(cr_context* ctx, Mlp* m, const float* p)
{
    std::vector<float> h(p, p + m->total);
    for (int l = 0; l < MLP_LAYERS; l++) {
        const uint32_t K = MLP_K[l], N = MLP_N[l], rn = real_n(l);
        for (uint32_t k = 0; k < K; k++)
            for (uint32_t n = rn; n < N; n++) h[m->wOff[l] + k * N + n] = 0.0f;
        for (uint32_t n = rn; n < N; n++) h[m->bOff[l] + n] = 0.0f;
    }
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemcpyAsync(m->params, h.data(), (size_t)m->total * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(ctx, e, "upload params");
    for (int l = 0; l < MLP_LAYERS; l++) {
        const uint32_t K = MLP_K[l], N = MLP_N[l];
        std::vector<uint16_t> w((size_t)K * N), wt((size_t)K * N);
        for (uint32_t k = 0; k < K; k++)
            for (uint32_t n = 0; n < N; n++) {
                const uint16_t b = f2bf_host(h[m->wOff[l] + k * N + n]);
                w[(size_t)k * N + n] = b;
                wt[(size_t)n * K + k] = b;
            }
        e = hipMemcpyAsync(m->W[l], w.data(), w.size() * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(m->WT[l], wt.data(), wt.size() * 2, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(ctx, e, "upload weights");
    }
    return CR_OK;
}

cr_status ensure_batch(cr_context* ctx, Mlp* m, uint64_t B, bool train, bool hostX)
{
    cr_status st = CR_OK;
    if (hostX) st = ensure_buffer(ctx, (void**)&m->X, &m->XCap, B * 224 * 2, "mlp X");
    if (st == CR_OK && !train) st = ensure_buffer(ctx, (void**)&m->Y, &m->YCap, B * 16, "mlp Y");
    if (st == CR_OK && train) st = ensure_buffer(ctx, (void**)&m->loss, &m->lossCap, (size_t)mlp_tiles(B) * 4, "mlp loss");
    if (st == CR_OK && train)
        st = ensure_buffer(ctx, (void**)&m->slot, &m->slotCap, (size_t)mlp_slots(B) * m->total * 4, "mlp gradient slots");
    if (st == CR_OK && train && hostX) st = ensure_buffer(ctx, (void**)&m->T, &m->TCap, B * 12, "mlp targets");
    if (st == CR_OK && train)
        st = ensure_buffer(ctx, (void**)&m->tr, &m->trCap, (size_t)MLP_TR_ROWS * mlp_padded_rows(B) * 2, "mlp training activations");
    return st;
}

// Restart the moving average (cr_mlp_init, cr_mlp_set_params, cr_mlp_set_ema).
cr_status ema_restart(cr_context* ctx, Mlp* m)
{
    m->emaSteps = 0;
    if (!m->ema) return CR_OK;
    hipError_t e = zero_async(m->ema, (size_t)m->total * 4, ctx->stream);
    return e == hipSuccess ? CR_OK : hip_fail(ctx, e, "zero ema");
}

MlpArgs args_of(Mlp* m, const uint16_t* X, uint64_t B)
{
    MlpArgs a;
    memset(&a, 0, sizeof(a));
    a.X = X;
    a.B = (int64_t)B;
    a.Bp = mlp_padded_rows(B);
    for (int l = 0; l < MLP_LAYERS; l++) {
        a.WT[l] = m->WT[l];
        a.W[l] = m->W[l];
        a.bias[l] = m->params + m->bOff[l];
        a.wOff[l] = m->wOff[l];
        a.bOff[l] = m->bOff[l];
    }
    a.Y = m->Y;
    a.tr = m->tr;
    a.loss = m->loss;
    a.grad = m->grad;
    a.slot = m->slot;
    a.slots = mlp_slots(B);
    a.total = m->total;
    return a;
}

// Forward arguments for inference: the moving average once it holds a step, else the parameters.
MlpArgs inference_args(Mlp* m, const uint16_t* X, uint64_t B)
{
    MlpArgs a = args_of(m, X, B);
    if (m->emaDecay > 0.0f && m->emaSteps > 0) {
        for (int l = 0; l < MLP_LAYERS; l++) {
            a.WT[l] = m->EWT[l];
            a.W[l] = m->EW[l];
            a.bias[l] = m->emaHat + m->bOff[l];
        }
    }
    return a;
}

} // namespace

namespace cr {
bool mlp_initialized(cr_context* ctx) { return ctx->mlp && ctx->mlp->initialized; }

void mlp_free(cr_context* ctx)
{
    Mlp* m = ctx->mlp;
    if (!m) return;
    void* bufs[] = {m->params, m->grad, m->m, m->v, m->X, m->T, m->Y, m->tr, m->loss, m->slot, m->ema, m->emaHat};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (int l = 0; l < MLP_LAYERS; l++) {
        if (m->W[l]) hipFree(m->W[l]);
        if (m->WT[l]) hipFree(m->WT[l]);
        if (m->EW[l]) hipFree(m->EW[l]);
        if (m->EWT[l]) hipFree(m->EWT[l]);
    }
    delete m;
    ctx->mlp = nullptr;
}

// Forward pass on device-resident features; Y receives [B][4] fp32.
cr_status mlp_forward_device(cr_context* ctx, const uint16_t* X, uint64_t B, float* Y, const uint32_t* Bdev)
{
    Mlp* m;
    cr_status st = mlp_get(ctx, &m);
    if (st != CR_OK) return st;
    MlpArgs a = inference_args(m, X, B);
    a.Y = Y;
    a.Bdev = Bdev;
    hipError_t e = launch_mlp_infer(a, (unsigned)ctx->numCUs, ctx->stream);
    return e == hipSuccess ? CR_OK : hip_fail(ctx, e, "mlp forward");
}
} // namespace cr

#define CR_CHECK_CTX(ctx) \
    do { if (!(ctx)) return CR_ERR_INVALID_ARGUMENT; } while (0)
#define CR_HIP(ctx, call, what) \
    do { hipError_t _e = (call); if (_e != hipSuccess) return hip_fail((ctx), _e, (what)); } while (0)
#define CR_TRY(call) \
    do { cr_status _s = (call); if (_s != CR_OK) return _s; } while (0)

extern "C" {

void cr_mlp_param_layout(uint32_t K[5], uint32_t N[5], uint32_t wOffset[5], uint32_t bOffset[5], uint32_t* total)
{
    uint32_t w[MLP_LAYERS], b[MLP_LAYERS], t;
    layout(w, b, &t);
    for (int l = 0; l < MLP_LAYERS; l++) {
        if (K) K[l] = MLP_K[l];
        if (N) N[l] = MLP_N[l];
        if (wOffset) wOffset[l] = w[l];
        if (bOffset) bOffset[l] = b[l];
    }
    if (total) *total = t;
}

cr_status cr_mlp_init(cr_context* ctx, uint32_t seed)
{
    CR_CHECK_CTX(ctx);
    hipSetDevice(ctx->device);
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    std::vector<float> p(m->total, 0.0f);
    // splitmix64 + Box-Muller, He-normal std sqrt(2 / fan_in)
    uint64_t x = 0x9E3779B97F4A7C15ull * (uint64_t)(seed + 1u);
    auto next = [&x]() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < MLP_LAYERS; l++) {
        const uint32_t fanIn = l == 0 ? 224u : 128u, N = MLP_N[l];
        const double sd = std::sqrt(2.0 / fanIn);
        for (uint32_t k = 0; k < MLP_K[l]; k++)
            for (uint32_t n = 0; n < real_n(l); n++) {
                double u1 = next(), u2 = next();
                if (u1 < 1e-300) u1 = 1e-300;
                p[m->wOff[l] + k * N + n] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
    }
    CR_TRY(upload_params(ctx, m, p.data()));
    CR_HIP(ctx, zero_async(m->m, (size_t)m->total * 4, ctx->stream), "zero m");
    CR_HIP(ctx, zero_async(m->v, (size_t)m->total * 4, ctx->stream), "zero v");
    CR_TRY(ema_restart(ctx, m));
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    m->step = 0;
    m->initialized = true;
    learn_network_reset(ctx);
    return CR_OK;
}

cr_status cr_mlp_set_params(cr_context* ctx, const float* params, uint32_t n)
{
    CR_CHECK_CTX(ctx);
    if (!params) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null params");
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    if (n != m->total) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "parameter count must equal the padded layout total");
    CR_TRY(upload_params(ctx, m, params));
    CR_HIP(ctx, zero_async(m->m, (size_t)m->total * 4, ctx->stream), "zero m");
    CR_HIP(ctx, zero_async(m->v, (size_t)m->total * 4, ctx->stream), "zero v");
    CR_TRY(ema_restart(ctx, m));
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    m->step = 0;
    m->initialized = true;
    learn_network_reset(ctx);
    return CR_OK;
}

cr_status cr_mlp_set_ema(cr_context* ctx, float decay)
{
    CR_CHECK_CTX(ctx);
    if (!(decay >= 0.0f && decay < 1.0f)) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "decay must be in [0, 1)");
    hipSetDevice(ctx->device);
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    if (decay > 0.0f && !m->ema) {
        size_t cap = 0;
        CR_TRY(ensure_buffer(ctx, (void**)&m->ema, &cap, (size_t)m->total * 4, "mlp ema"));
        cap = 0;
        CR_TRY(ensure_buffer(ctx, (void**)&m->emaHat, &cap, (size_t)m->total * 4, "mlp ema (corrected)"));
        for (int l = 0; l < MLP_LAYERS; l++) {
            cap = 0;
            CR_TRY(ensure_buffer(ctx, (void**)&m->EW[l], &cap, (size_t)MLP_K[l] * MLP_N[l] * 2, "mlp ema W"));
            cap = 0;
            CR_TRY(ensure_buffer(ctx, (void**)&m->EWT[l], &cap, (size_t)MLP_K[l] * MLP_N[l] * 2, "mlp ema WT"));
        }
    }
    m->emaDecay = decay;
    CR_TRY(ema_restart(ctx, m));
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_mlp_get_ema(cr_context* ctx, float* params, uint32_t n)
{
    CR_CHECK_CTX(ctx);
    if (!params) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null params");
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    if (n != m->total) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "parameter count must equal the padded layout total");
    if (!(m->emaDecay > 0.0f) || m->emaSteps == 0) return fail(ctx, CR_ERR_STATE, "no moving average (cr_mlp_set_ema, then train)");
    CR_HIP(ctx, hipMemcpyAsync(params, m->emaHat, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy ema");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_mlp_get_params(cr_context* ctx, float* params, uint32_t n)
{
    CR_CHECK_CTX(ctx);
    if (!params) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null params");
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    if (!m->initialized) return fail(ctx, CR_ERR_STATE, "no network parameters (cr_mlp_init / cr_mlp_set_params)");
    if (n != m->total) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "parameter count must equal the padded layout total");
    CR_HIP(ctx, hipMemcpyAsync(params, m->params, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy params");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return CR_OK;
}

cr_status cr_mlp_forward(cr_context* ctx, const uint16_t* X, uint32_t B, float* Y)
{
    CR_CHECK_CTX(ctx);
    if (!X || !Y) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null input or output");
    if (B == 0) return CR_OK;
    hipSetDevice(ctx->device);
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    if (!m->initialized) return fail(ctx, CR_ERR_STATE, "no network parameters (cr_mlp_init / cr_mlp_set_params)");
    CR_TRY(ensure_batch(ctx, m, B, false, true));
    CR_HIP(ctx, hipMemcpyAsync(m->X, X, (size_t)B * 224 * 2, hipMemcpyHostToDevice, ctx->stream), "copy X");
    MlpArgs a = inference_args(m, m->X, B);
    CR_HIP(ctx, launch_mlp_infer(a, (unsigned)ctx->numCUs, ctx->stream), "mlp forward");
    std::vector<float> y((size_t)B * 4);
    CR_HIP(ctx, hipMemcpyAsync(y.data(), m->Y, y.size() * 4, hipMemcpyDeviceToHost, ctx->stream), "copy Y");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    for (uint32_t i = 0; i < B; i++)
        for (int c = 0; c < 3; c++) Y[(size_t)i * 3 + c] = y[(size_t)i * 4 + c];
    return CR_OK;
}

cr_status cr_mlp_train_step(cr_context* ctx, const uint16_t* X, const float* targets, uint32_t B, float lr, float* loss,
                            float* grads)
{
    CR_CHECK_CTX(ctx);
    if (!X || !targets) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null input or targets");
    if (B == 0) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "empty batch");
    hipSetDevice(ctx->device);
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    if (!m->initialized) return fail(ctx, CR_ERR_STATE, "no network parameters (cr_mlp_init / cr_mlp_set_params)");
    CR_TRY(ensure_batch(ctx, m, B, true, true));
    hipStream_t s = ctx->stream;
    CR_HIP(ctx, hipMemcpyAsync(m->X, X, (size_t)B * 224 * 2, hipMemcpyHostToDevice, s), "copy X");
    CR_HIP(ctx, hipMemcpyAsync(m->T, targets, (size_t)B * 12, hipMemcpyHostToDevice, s), "copy targets");
    CR_TRY(mlp_train_device(ctx, m->X, m->T, B, lr));
    if (loss) CR_TRY(mlp_last_loss(ctx, loss));
    if (grads) {
        CR_HIP(ctx, hipMemcpyAsync(grads, m->grad, (size_t)m->total * 4, hipMemcpyDeviceToHost, s), "copy grads");
    }
    CR_HIP(ctx, hipStreamSynchronize(s), "sync");
    return CR_OK;
}

cr_status cr_mlp_probe(cr_context* ctx, const uint16_t* A, const uint16_t* B, float* D)
{
    CR_CHECK_CTX(ctx);
    if (!A || !B || !D) return fail(ctx, CR_ERR_INVALID_ARGUMENT, "null matrix");
    hipSetDevice(ctx->device);
    void* buf = nullptr;
    CR_HIP(ctx, hipMalloc(&buf, 3072), "hipMalloc");
    uint16_t* dA = (uint16_t*)buf;
    uint16_t* dB = dA + 512;
    float* dD = (float*)(dB + 512);
    hipError_t e = hipMemcpyAsync(dA, A, 1024, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dB, B, 1024, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_mlp_probe(dA, dB, dD, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(D, dD, 1024, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(buf);
    return e == hipSuccess ? CR_OK : hip_fail(ctx, e, "mlp probe");
}

} // extern "C"

namespace cr {

// Loss of the last training step: its blocks' sums added in block order
// (synchronises the stream).
cr_status mlp_last_loss(cr_context* ctx, float* loss)
{
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    std::vector<float> part(m->lossBlocks);
    if (m->lossBlocks)
        CR_HIP(ctx, hipMemcpyAsync(part.data(), m->loss, part.size() * 4, hipMemcpyDeviceToHost, ctx->stream), "copy loss");
    CR_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    float sum = 0.0f;
    for (float v : part) sum += v;
    *loss = sum;
    return CR_OK;
}

// One training step on device-resident X and T, or on rows drawn from a data
// set (draw): the fused forward / loss / backward kernel, the weight-gradient
// GEMM, then Adam.
cr_status mlp_train_device(cr_context* ctx, const uint16_t* X, const float* T, uint64_t B, float lr, const MlpDraw* draw)
{
    Mlp* m;
    CR_TRY(mlp_get(ctx, &m));
    CR_TRY(ensure_batch(ctx, m, B, true, false));
    hipStream_t s = ctx->stream;
    MlpArgs a = args_of(m, X, B);
    a.T = T;
    if (draw) {
        a.X = nullptr;
        a.T = nullptr;
        a.dsX = draw->X;
        a.dsT = draw->T;
        a.dsN = draw->n;
        a.seed = draw->seed;
        a.step = draw->step;
    }
    CR_HIP(ctx, launch_mlp_train(a, s), "mlp train");
    m->lossBlocks = mlp_tiles(B);
    CR_HIP(ctx, launch_mlp_wgrad(a, s), "mlp weight gradients");
    m->step++;
    AdamArgs o;
    o.params = m->params;
    o.m = m->m;
    o.v = m->v;
    o.n = m->total;
    o.lr = lr;
    o.beta1 = 0.9f;
    o.beta2 = 0.999f;
    o.eps = 1e-8f;
    o.c1 = (float)(1.0 - std::pow(0.9, (double)m->step));
    o.c2 = (float)(1.0 - std::pow(0.999, (double)m->step));
    o.ema = nullptr;
    o.emaHat = nullptr;
    o.emaDecay = 0.0f;
    o.emaCorr = 1.0f;
    for (int l = 0; l < MLP_LAYERS; l++) {
        o.EW[l] = nullptr;
        o.EWT[l] = nullptr;
    }
    if (m->emaDecay > 0.0f) {
        m->emaSteps++;
        o.ema = m->ema;
        o.emaHat = m->emaHat;
        o.emaDecay = m->emaDecay;
        o.emaCorr = (float)(1.0 / (1.0 - std::pow((double)m->emaDecay, (double)m->emaSteps)));
        for (int l = 0; l < MLP_LAYERS; l++) {
            o.EW[l] = m->EW[l];
            o.EWT[l] = m->EWT[l];
        }
    }
    CR_HIP(ctx, launch_mlp_adam(a, o, s), "mlp adam");
    return CR_OK;
}

} // namespace cr
