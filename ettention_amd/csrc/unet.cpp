#include "unet.hpp"

#include <algorithm>
#include <cmath>

namespace etn {

namespace {

struct LayerSpec {
    const char* name;
    int cin, cout, level;
};

// c1a reads (x, d); c4a reads concat(upsample(c3b), c2b), c5a reads concat(upsample(c4b), c1b)
constexpr LayerSpec kLayers[] = {
    {"c1a", 2, 16, 0},  {"c1b", 16, 16, 0}, {"c2a", 16, 32, 1}, {"c2b", 32, 32, 1}, {"c3a", 32, 64, 2},
    {"c3b", 64, 64, 2}, {"c4a", 96, 32, 1}, {"c4b", 32, 32, 1}, {"c5a", 48, 16, 0}, {"c5b", 16, 16, 0},
};
constexpr int kLayerCount = 10;
constexpr int kHead = 16;
constexpr int kInputChannels = 8;   // (x, d) padded to one 16-byte channel group

}  // namespace

size_t UNet3D::parameterCount()
{
    size_t n = 0;
    for (const LayerSpec& l : kLayers) n += (size_t)l.cout * 27 * l.cin + l.cout;
    return n + kHead + 1;
}

UNet3D::UNet3D(int nx, int ny, int nz, hipStream_t stream) : stream_(stream)
{
    if (nx < 4 || ny < 4 || nz < 4 || nx % 4 || ny % 4 || nz % 4)
        fail("the correction network needs volume sizes that are multiples of 4");
    for (int l = 0; l < 3; ++l) {
        D_[l] = nz >> l;
        H_[l] = ny >> l;
        W_[l] = nx >> l;
    }
    size_t at = 0;
    for (const LayerSpec& s : kLayers) {
        Layer L{s.name, s.cin, s.cout, s.level, at, at + (size_t)s.cout * 27 * s.cin};
        at = L.b + s.cout;
        layers_.push_back(L);
    }
    head_ = at;
    const size_t P = parameterCount();
    params_.allocate(sizeof(float) * P);
    ETN_HIP(hipMemsetAsync(params_.get(), 0, sizeof(float) * P, stream_));
    for (const Layer& L : layers_) {
        wt_.emplace_back(sizeof(uint16_t) * L.cout * nn::kpad_for(std::max(L.cin, kInputChannels)));
        wb_.emplace_back(L.cin >= 8 ? sizeof(uint16_t) * L.cin * nn::kpad_for(L.cout) : 0);
        act_.emplace_back(sizeof(uint16_t) * voxels(L.level) * L.cout);
    }
    xin_.allocate(sizeof(uint16_t) * voxels(0) * kInputChannels);
    pool1_.allocate(sizeof(uint16_t) * voxels(1) * 16);
    pool2_.allocate(sizeof(uint16_t) * voxels(2) * 32);
    pred_.allocate(sizeof(float) * voxels(0));
    packAll();
}

void UNet3D::setParameters(const float* host, size_t count)
{
    if (count != parameterCount()) fail("parameter count mismatch");
    ETN_HIP(hipMemcpyAsync(params_.get(), host, sizeof(float) * count, hipMemcpyHostToDevice, stream_));
    ETN_HIP(hipStreamSynchronize(stream_));
    packAll();
    adam_t_ = 0;
    if (training_) {
        ETN_HIP(hipMemsetAsync(adam_m_.get(), 0, adam_m_.bytes(), stream_));
        ETN_HIP(hipMemsetAsync(adam_v_.get(), 0, adam_v_.bytes(), stream_));
    }
}

void UNet3D::getParameters(float* host, size_t count) const
{
    if (count != parameterCount()) fail("parameter count mismatch");
    ETN_HIP(hipMemcpyAsync(host, params_.get(), sizeof(float) * count, hipMemcpyDeviceToHost, stream_));
    ETN_HIP(hipStreamSynchronize(stream_));
}

void UNet3D::packAll()
{
    for (int i = 0; i < kLayerCount; ++i) {
        const Layer& L = layers_[i];
        nn::pack_weights(params_.as<float>() + L.w, L.cout, L.cin, std::max(L.cin, kInputChannels),
                         wt_[i].as<uint16_t>(), L.cin >= 8 ? wb_[i].as<uint16_t>() : nullptr, stream_);
    }
}

nn::ConvSrc UNet3D::layerInput(int i) const
{
    nn::ConvSrc s;
    switch (i) {
        case 0: s.a = xin_.as<uint16_t>(); s.ca = kInputChannels; break;
        case 2: s.a = pool1_.as<uint16_t>(); s.ca = 16; break;
        case 4: s.a = pool2_.as<uint16_t>(); s.ca = 32; break;
        case 6:
            s.a = layerOutput(5); s.ca = 64; s.a_half = 1;
            s.b = layerOutput(3); s.cb = 32;
            break;
        case 8:
            s.a = layerOutput(7); s.ca = 32; s.a_half = 1;
            s.b = layerOutput(1); s.cb = 16;
            break;
        default: s.a = layerOutput(i - 1); s.ca = layers_[i - 1].cout; break;
    }
    return s;
}

void UNet3D::forward(const float* x, const float* d, float scale, float dscale)
{
    nn::pack_input(x, scale, d, dscale, voxels(0), xin_.as<uint16_t>(), stream_);
    for (int i = 0; i < kLayerCount; ++i) {
        const Layer& L = layers_[i];
        if (i == 2) nn::pool2(layerOutput(1), D_[0], H_[0], W_[0], 16, pool1_.as<uint16_t>(), stream_);
        if (i == 4) nn::pool2(layerOutput(3), D_[1], H_[1], W_[1], 32, pool2_.as<uint16_t>(), stream_);
        nn::ConvArgs a;
        a.src = layerInput(i);
        a.D = D_[L.level];
        a.H = H_[L.level];
        a.W = W_[L.level];
        a.wt = wt_[i].as<uint16_t>();
        a.kpad = nn::kpad_for(std::max(L.cin, kInputChannels));
        a.bias = params_.as<float>() + L.b;
        a.cout = L.cout;
        a.relu = 1;
        a.out = act_[i].as<uint16_t>();
        nn::conv3_forward(a, stream_);
    }
    const float* p = params_.as<float>();
    nn::head_forward(layerOutput(9), x, scale, p + head_, p + head_ + kHead, voxels(0), pred_.as<float>(), stream_);
}

void UNet3D::predict(const float* x, const float* d, float scale, float dscale, float* out)
{
    if (!(scale > 0.0f) || !std::isfinite(scale) || !std::isfinite(dscale)) fail("scales must be finite, scale > 0");
    forward(x, d, scale, dscale);
    nn::unscale_prediction(pred_.as<float>(), x, scale, voxels(0), out, stream_);
}

void UNet3D::ensureTraining()
{
    if (training_) return;
    const size_t P = parameterCount();
    grads_.allocate(sizeof(float) * P);
    adam_m_.allocate(sizeof(float) * P);
    adam_v_.allocate(sizeof(float) * P);
    ETN_HIP(hipMemsetAsync(adam_m_.get(), 0, adam_m_.bytes(), stream_));
    ETN_HIP(hipMemsetAsync(adam_v_.get(), 0, adam_v_.bytes(), stream_));
    const size_t m0 = voxels(0), m1 = voxels(1), m2 = voxels(2);
    const size_t h = sizeof(uint16_t);
    gpred_.allocate(sizeof(float) * m0);
    gd1_.allocate(h * m0 * 16);
    gtmp_.allocate(h * std::max({m0 * 16, m1 * 32, m2 * 64}));
    gcat_.allocate(h * std::max(m0 * 48, m1 * 96));
    ge1_.allocate(h * m0 * 16);
    ge2_.allocate(h * m1 * 32);
    ge3_.allocate(h * m2 * 64);
    gd2_.allocate(h * m1 * 32);
    gp_.allocate(h * std::max(m2 * 32, m1 * 16));
    size_t wpart = 0, bpart = 0;
    for (const Layer& L : layers_) {
        const size_t chunks = (size_t)nn::wgrad_chunks(voxels(L.level));
        wpart = std::max(wpart, chunks * 27 * L.cout * ((L.cin + 15) / 16 * 16));
        bpart = std::max(bpart, chunks * L.cout);
    }
    wpart_.allocate(sizeof(float) * wpart);
    bpart_.allocate(sizeof(float) * bpart);
    hpart_.allocate(sizeof(float) * 1024 * 17);
    loss_.allocate(sizeof(double));
    lpart_.allocate(sizeof(double) * 1024);
    training_ = true;
}

void UNet3D::backward()
{
    float* g = grads_.as<float>();
    auto layer_back = [&](int i, uint16_t* gy, uint16_t* gx) {
        const Layer& L = layers_[i];
        const int D = D_[L.level], H = H_[L.level], W = W_[L.level];
        const size_t M = voxels(L.level);
        nn::relu_backward(gy, layerOutput(i), M * L.cout, stream_);
        nn::conv3_weight_grad(layerInput(i), D, H, W, gy, L.cout, wpart_.as<float>(), stream_);
        nn::conv3_weight_grad_reduce(wpart_.as<float>(), nn::wgrad_chunks(M), L.cout, L.cin, g + L.w, stream_);
        nn::channel_sum(gy, M, L.cout, bpart_.as<float>(), g + L.b, stream_);
        if (gx) {
            nn::ConvArgs a;
            a.src.a = gy;
            a.src.ca = L.cout;
            a.D = D;
            a.H = H;
            a.W = W;
            a.wt = wb_[i].as<uint16_t>();
            a.kpad = nn::kpad_for(L.cout);
            a.cout = L.cin;
            a.out = gx;
            nn::conv3_forward(a, stream_);
        }
    };
    uint16_t* gtmp = gtmp_.as<uint16_t>();
    uint16_t* gcat = gcat_.as<uint16_t>();
    uint16_t* gp = gp_.as<uint16_t>();
    const float* p = params_.as<float>();
    nn::head_backward(layerOutput(9), gpred_.as<float>(), p + head_, voxels(0), gd1_.as<uint16_t>(), hpart_.as<float>(),
                      g + head_, stream_);
    // decoder, full resolution
    layer_back(9, gd1_.as<uint16_t>(), gtmp);
    layer_back(8, gtmp, gcat);
    nn::up2_backward(gcat, 48, 0, D_[0], H_[0], W_[0], 32, gd2_.as<uint16_t>(), stream_);
    nn::take_channels(gcat, 48, 32, voxels(0), 16, ge1_.as<uint16_t>(), 0, stream_);
    // decoder, half resolution
    layer_back(7, gd2_.as<uint16_t>(), gtmp);
    layer_back(6, gtmp, gcat);
    nn::up2_backward(gcat, 96, 0, D_[1], H_[1], W_[1], 64, ge3_.as<uint16_t>(), stream_);
    nn::take_channels(gcat, 96, 64, voxels(1), 32, ge2_.as<uint16_t>(), 0, stream_);
    // bottleneck
    layer_back(5, ge3_.as<uint16_t>(), gtmp);
    layer_back(4, gtmp, gp);
    nn::pool2_backward(gp, D_[1], H_[1], W_[1], 32, ge2_.as<uint16_t>(), 1, stream_);
    // encoder, half resolution
    layer_back(3, ge2_.as<uint16_t>(), gtmp);
    layer_back(2, gtmp, gp);
    nn::pool2_backward(gp, D_[0], H_[0], W_[0], 16, ge1_.as<uint16_t>(), 1, stream_);
    // encoder, full resolution
    layer_back(1, ge1_.as<uint16_t>(), gtmp);
    layer_back(0, gtmp, nullptr);
}

double UNet3D::gradient(const float* x, const float* d, const float* target, float scale, float dscale)
{
    if (!(scale > 0.0f) || !std::isfinite(scale) || !std::isfinite(dscale)) fail("scales must be finite, scale > 0");
    ensureTraining();
    forward(x, d, scale, dscale);
    nn::mse_loss(pred_.as<float>(), target, scale, voxels(0), gpred_.as<float>(), lpart_.as<double>(),
                 loss_.as<double>(), stream_);
    backward();
    double loss = 0.0;
    ETN_HIP(hipMemcpyAsync(&loss, loss_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    ETN_HIP(hipStreamSynchronize(stream_));
    return loss;
}

void UNet3D::copyGradient(float* host, size_t count) const
{
    if (!training_) fail("no gradient has been computed");
    if (count != parameterCount()) fail("parameter count mismatch");
    ETN_HIP(hipMemcpyAsync(host, grads_.get(), sizeof(float) * count, hipMemcpyDeviceToHost, stream_));
    ETN_HIP(hipStreamSynchronize(stream_));
}

double UNet3D::trainStep(const float* x, const float* d, const float* target, float scale, float dscale, float lr)
{
    const double loss = gradient(x, d, target, scale, dscale);
    ++adam_t_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - std::pow(b1, (float)adam_t_), bc2 = 1.0f - std::pow(b2, (float)adam_t_);
    nn::adam_step(params_.as<float>(), grads_.as<float>(), adam_m_.as<float>(), adam_v_.as<float>(), parameterCount(),
                  lr, b1, b2, 1e-8f, bc1, bc2, stream_);
    packAll();
    return loss;
}

double UNet3D::timeForward(const float* x, float scale, double* flops)
{
    double f = 0.0;
    for (const Layer& L : layers_) f += 2.0 * 27.0 * L.cin * L.cout * (double)voxels(L.level);
    f += 2.0 * kHead * (double)voxels(0);
    if (flops) *flops = f;
    hipEvent_t a, b;
    ETN_HIP(hipEventCreate(&a));
    ETN_HIP(hipEventCreate(&b));
    ETN_HIP(hipEventRecord(a, stream_));
    forward(x, nullptr, scale, 0.0f);
    ETN_HIP(hipEventRecord(b, stream_));
    ETN_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    ETN_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

}  // namespace etn
