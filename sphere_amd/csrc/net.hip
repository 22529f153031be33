// net.hip -- the Darcy correction network: a four-level 3-D U-Net whose
// forward and backward passes run on the kernels of net_kernels.hip, behind
// the sph_net_* entry points (device pointers, caller's stream).
//
// Layers (3x3x3 convolutions, zero padding or wrap-around along periodic
// axes, ReLU except the last; levels L0 = the Darcy grid, L1..L3 halved by
// 2x average pooling):
//   x0 = c0(in)   3 -> 16 @L0      a3 = c4(x3)   64 -> 48 @L3
//   x1 = c1(pool x0) 16 -> 32 @L1   y2 = up(a3) + x2     @L2
//   x2 = c2(pool x1) 32 -> 48 @L2   a2 = c5(y2)   48 -> 32 @L2,  y1 = up(a2) + x1
//   x3 = c3(pool x2) 48 -> 64 @L3   a1 = c6(y1)   32 -> 16 @L1,  y0 = up(a1) + x0
//   h0 = c7(y0) 16 -> 16 @L0,   out = c8 . h0 + b8 (pointwise 16 -> 1, FP32)
// Parameters: per layer w[taps][cout][cin] then b[cout] (taps 27, 1 for
// c8), FP32, 284,977 in all.
#include "net.h"
#include "net_kernels.h"

#include <mutex>
#include <string>
#include <vector>

namespace {

struct LayerSpec {
    int cin, cout, level, taps;
};

const LayerSpec kLayers[9] = {{3, 16, 0, 27},  {16, 32, 1, 27}, {32, 48, 2, 27}, {48, 64, 3, 27},
                              {64, 48, 3, 27}, {48, 32, 2, 27}, {32, 16, 1, 27}, {16, 16, 0, 27},
                              {16, 1, 0, 1}};

int pad_in(int c) { return c <= 4 ? 4 : (c + 15) / 16 * 16; }
int pad_out(int c) { return (c + 15) / 16 * 16; }

std::string g_net_error;

}  // namespace

// Scratch of sph_net_conv3d_noalloc: the BF16 copy of one call's input
// followed by its packed weights.  Each GPU the library runs on holds its own
// copy of this array until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in): it uses the
// device address hipGetSymbolAddress returns, kept per device.
#define NET_NOALLOC_BYTES (16u << 20)
static __device__ __align__(256) unsigned char scratch[NET_NOALLOC_BYTES];

#define NET_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[NET_MAX_DEVICES];

// The scratch address on the current device, resolved on the first call
// there (nullptr with g_net_error set on failure).
static unsigned char *noalloc_scratch()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NET_MAX_DEVICES) {
        g_net_error = "no current HIP device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            g_net_error = std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_dev[dev] = (unsigned char *)ptr;
    }
    return g_scratch_dev[dev];
}

struct sph_net {
    int device = 0;
    NetGeom g[4];
    uint32_t V[4];
    size_t woff[9], boff[9];
    uint16_t *wf[9], *wt[9];
    std::vector<void *> allocs;
    // activations (BF16, channels-last)
    uint16_t *in4, *x0, *p0, *x1, *p1, *x2, *p2, *x3, *a3, *y2, *a2, *y1, *a1, *y0, *h0;
    // gradients (FP32) and the BF16 operand of the weight gradient
    float *d_h0, *d_y0, *d_a1, *d_y1, *d_a2, *d_y2, *d_a3, *d_x3, *d_p2, *d_p1, *d_p0;
    uint16_t *dz;
    float *gw;
};

static bool net_alloc(sph_net *n, void **p, size_t bytes)
{
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) return false;
    n->allocs.push_back(*p);
    return true;
}

extern "C" {

uint64_t sph_net_param_count(void)
{
    uint64_t n = 0;
    for (const LayerSpec &l : kLayers) n += (uint64_t)l.taps * l.cout * l.cin + l.cout;
    return n;
}

const char *sph_net_last_error(void) { return g_net_error.c_str(); }

void sph_net_destroy(sph_net *n)
{
    if (!n) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(n->device);
    for (void *p : n->allocs) (void)hipFree(p);
    (void)hipSetDevice(prev);
    delete n;
}

sph_net *sph_net_create(int device, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t periodic)
{
    if (nx == 0 || ny == 0 || nz == 0 || nx % 8 || ny % 8 || nz % 8) {
        g_net_error = "the network grid must be a positive multiple of 8 along each axis";
        return nullptr;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_net_error = "no HIP device available";
        return nullptr;
    }
    if (device < 0 || device >= count) {
        g_net_error = "device index out of range";
        return nullptr;
    }
    int prev = 0;
    (void)hipGetDevice(&prev);
    if (hipSetDevice(device) != hipSuccess) {
        g_net_error = "hipSetDevice failed";
        return nullptr;
    }
    sph_net *n = new sph_net();
    n->device = device;
    for (int l = 0; l < 4; ++l) {
        n->g[l] = NetGeom{nx >> l, ny >> l, nz >> l, periodic & 3u};
        n->V[l] = n->g[l].nx * n->g[l].ny * n->g[l].nz;
    }
    size_t off = 0;
    bool ok = true;
    for (int l = 0; l < 9; ++l) {
        const LayerSpec &s = kLayers[l];
        n->woff[l] = off;
        off += (size_t)s.taps * s.cout * s.cin;
        n->boff[l] = off;
        off += s.cout;
        n->wf[l] = n->wt[l] = nullptr;
        if (s.taps != 27) continue;  // the pointwise head reads its FP32 weights
        ok = ok && net_alloc(n, (void **)&n->wf[l], 27ull * pad_out(s.cout) * pad_in(s.cin) * 2);
        // adjoint: input = this layer's output gradient (cout channels,
        // padded to 4 or 16), output = cin channels
        ok = ok && net_alloc(n, (void **)&n->wt[l], 27ull * pad_out(s.cin) * pad_in(s.cout) * 2);
    }
    const uint32_t *V = n->V;
    auto A16 = [&](uint16_t **p, size_t count) { ok = ok && net_alloc(n, (void **)p, count * 2); };
    auto A32 = [&](float **p, size_t count) { ok = ok && net_alloc(n, (void **)p, count * 4); };
    A16(&n->in4, (size_t)V[0] * 4);
    A16(&n->x0, (size_t)V[0] * 16); A16(&n->p0, (size_t)V[1] * 16);
    A16(&n->x1, (size_t)V[1] * 32); A16(&n->p1, (size_t)V[2] * 32);
    A16(&n->x2, (size_t)V[2] * 48); A16(&n->p2, (size_t)V[3] * 48);
    A16(&n->x3, (size_t)V[3] * 64); A16(&n->a3, (size_t)V[3] * 48);
    A16(&n->y2, (size_t)V[2] * 48); A16(&n->a2, (size_t)V[2] * 32);
    A16(&n->y1, (size_t)V[1] * 32); A16(&n->a1, (size_t)V[1] * 16);
    A16(&n->y0, (size_t)V[0] * 16); A16(&n->h0, (size_t)V[0] * 16);
    A32(&n->d_h0, (size_t)V[0] * 16); A32(&n->d_y0, (size_t)V[0] * 16);
    A32(&n->d_a1, (size_t)V[1] * 16); A32(&n->d_y1, (size_t)V[1] * 32);
    A32(&n->d_a2, (size_t)V[2] * 32); A32(&n->d_y2, (size_t)V[2] * 48);
    A32(&n->d_a3, (size_t)V[3] * 48); A32(&n->d_x3, (size_t)V[3] * 64);
    A32(&n->d_p2, (size_t)V[3] * 48); A32(&n->d_p1, (size_t)V[2] * 32);
    A32(&n->d_p0, (size_t)V[1] * 16);
    A16(&n->dz, (size_t)V[0] * 16);
    A32(&n->gw, 27ull * 64 * 64);
    (void)hipSetDevice(prev);
    if (!ok) {
        g_net_error = "device allocation failed";
        sph_net_destroy(n);
        return nullptr;
    }
    return n;
}

}  // extern "C"

// BF16 operand copies of the FP32 parameters (forward and adjoint layouts)
void net_pack(sph_net *n, const float *params, hipStream_t s)
{
    for (int l = 0; l < 9; ++l) {
        const LayerSpec &L = kLayers[l];
        if (L.taps != 27) continue;
        launch_pack_weights(s, params + n->woff[l], L.cin, L.cout, n->wf[l], pad_in(L.cin), pad_out(L.cout),
                            n->wt[l], pad_in(L.cout), pad_out(L.cin));
    }
}

// the network's BF16 input rows (V x 4), which a caller may fill itself and
// then pass a null input to net_forward_packed
uint16_t *net_input_bf16(sph_net *n) { return n->in4; }

// forward pass with the weights of the last net_pack
int net_forward_packed(sph_net *n, const float *params, const float *input, float *output, hipStream_t s)
{
    const NetGeom *g = n->g;
    const float *b[9];
    for (int l = 0; l < 9; ++l) b[l] = params + n->boff[l];
    int rc = 0;
    if (input) launch_to_bf16_pad(s, n->V[0], input, 3, 4, n->in4);  // else in4 is filled
    rc |= launch_conv_fwd(s, g[0], n->in4, 4, n->wf[0], 16, b[0], 16, n->x0, NET_OUT_BF16_RELU);
    launch_pool_fwd(s, g[1], n->x0, 16, n->p0);
    rc |= launch_conv_fwd(s, g[1], n->p0, 16, n->wf[1], 32, b[1], 32, n->x1, NET_OUT_BF16_RELU);
    launch_pool_fwd(s, g[2], n->x1, 32, n->p1);
    rc |= launch_conv_fwd(s, g[2], n->p1, 32, n->wf[2], 48, b[2], 48, n->x2, NET_OUT_BF16_RELU);
    launch_pool_fwd(s, g[3], n->x2, 48, n->p2);
    rc |= launch_conv_fwd(s, g[3], n->p2, 48, n->wf[3], 64, b[3], 64, n->x3, NET_OUT_BF16_RELU);
    // decoder convolutions also write up(a) + skip on the finer grid
    rc |= launch_conv_fwd(s, g[3], n->x3, 64, n->wf[4], 48, b[4], 48, n->a3, NET_OUT_RELU_UP, nullptr,
                          nullptr, n->x2, n->y2);
    rc |= launch_conv_fwd(s, g[2], n->y2, 48, n->wf[5], 32, b[5], 32, n->a2, NET_OUT_RELU_UP, nullptr,
                          nullptr, n->x1, n->y1);
    rc |= launch_conv_fwd(s, g[1], n->y1, 32, n->wf[6], 16, b[6], 16, n->a1, NET_OUT_RELU_UP, nullptr,
                          nullptr, n->x0, n->y0);
    // c7 with the pointwise head c8 (weights then bias, contiguous) in its epilogue
    rc |= launch_conv_fwd(s, g[0], n->y0, 16, n->wf[7], 16, b[7], 16, n->h0, NET_OUT_PROJ,
                          params + n->woff[8], output);
    if (rc) {
        g_net_error = "unsupported layer shape";
        return SPH_ERR_VALUE;
    }
    if (hipGetLastError() != hipSuccess) {
        g_net_error = "kernel launch failed";
        return SPH_ERR_HIP;
    }
    return SPH_OK;
}

// ---------------------------------------------------------------------------
// Multi-box forward passes (ensembles)
// ---------------------------------------------------------------------------
struct NetMultiOp {  // one launch: a layer's boxes that share a kernel variant
    bool pool = false;
    int variant = 0, cin = 0, cout_pad = 0;
    void *tasks = nullptr;
    uint32_t *map = nullptr;
    uint32_t nbox = 0, blocks = 0;
};

struct NetMulti {
    int device = 0;
    std::vector<NetMultiOp> ops;
    std::vector<void *> allocs;
};

template <class T>
static bool net_multi_op(NetMulti *m, NetMultiOp op, const std::vector<T> &tasks,
                         const std::vector<uint32_t> &counts)
{
    if (tasks.empty()) return true;
    std::vector<uint32_t> map(2 * tasks.size());
    uint32_t at = 0;
    for (size_t k = 0; k < tasks.size(); ++k) {
        map[k] = at;
        map[tasks.size() + k] = counts[k];
        at += (counts[k] + 7u) & ~7u;
    }
    void *dt = nullptr, *dm = nullptr;
    if (hipMalloc(&dt, sizeof(T) * tasks.size()) != hipSuccess) return false;
    m->allocs.push_back(dt);
    if (hipMalloc(&dm, 4 * map.size()) != hipSuccess) return false;
    m->allocs.push_back(dm);
    if (hipMemcpy(dt, tasks.data(), sizeof(T) * tasks.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dm, map.data(), 4 * map.size(), hipMemcpyHostToDevice) != hipSuccess)
        return false;
    op.tasks = dt;
    op.map = (uint32_t *)dm;
    op.nbox = (uint32_t)tasks.size();
    op.blocks = at;
    m->ops.push_back(op);
    return true;
}

void net_multi_destroy(NetMulti *m)
{
    if (!m) return;
    for (void *p : m->allocs) (void)hipFree(p);
    delete m;
}

NetMulti *net_multi_create(sph_net *const *nets, const float *const *params, float *const *outputs, int n)
{
    NetMulti *m = new NetMulti();
    (void)hipGetDevice(&m->device);
    // the layers of net_forward_packed: (level, input, cin, layer, cout_pad,
    // output, mode, skip, fine output), with the pools between them
    struct Step {
        bool pool;
        int level, cin, layer, cout_pad, mode;
    };
    const Step steps[11] = {{false, 0, 4, 0, 16, NET_OUT_BF16_RELU}, {true, 1, 16, 0, 0, 0},
                            {false, 1, 16, 1, 32, NET_OUT_BF16_RELU}, {true, 2, 32, 0, 0, 0},
                            {false, 2, 32, 2, 48, NET_OUT_BF16_RELU}, {true, 3, 48, 0, 0, 0},
                            {false, 3, 48, 3, 64, NET_OUT_BF16_RELU}, {false, 3, 64, 4, 48, NET_OUT_RELU_UP},
                            {false, 2, 48, 5, 32, NET_OUT_RELU_UP},   {false, 1, 32, 6, 16, NET_OUT_RELU_UP},
                            {false, 0, 16, 7, 16, NET_OUT_PROJ}};
    for (int q = 0; q < 11; ++q) {
        const Step &st = steps[q];
        if (st.pool) {
            std::vector<PoolTask> t;
            std::vector<uint32_t> c;
            for (int b = 0; b < n; ++b) {
                const sph_net *nt = nets[b];
                const uint16_t *in = st.level == 1 ? nt->x0 : st.level == 2 ? nt->x1 : nt->x2;
                uint16_t *out = st.level == 1 ? nt->p0 : st.level == 2 ? nt->p1 : nt->p2;
                t.push_back(PoolTask{nt->g[st.level], in, st.cin, out});
                c.push_back(pool_blocks(nt->g[st.level], st.cin));
            }
            NetMultiOp op;
            op.pool = true;
            if (!net_multi_op(m, op, t, c)) {
                net_multi_destroy(m);
                return nullptr;
            }
            continue;
        }
        // group the boxes by the kernel variant their grid takes
        for (int v = 0; v < 5; ++v) {
            std::vector<ConvTask> t;
            std::vector<uint32_t> c;
            for (int b = 0; b < n; ++b) {
                const sph_net *nt = nets[b];
                const NetGeom &g = nt->g[st.level];
                if (conv_variant(g, st.cin, st.cout_pad) != v) continue;
                ConvTask k{};
                k.g = g;
                k.cin = st.cin;
                k.w = nt->wf[st.layer];
                k.bias = params[b] + nt->boff[st.layer];
                k.cout_real = kLayers[st.layer].cout;
                k.mode = st.mode;
                switch (st.layer) {
                case 0: k.in = nt->in4; k.out = nt->x0; break;
                case 1: k.in = nt->p0; k.out = nt->x1; break;
                case 2: k.in = nt->p1; k.out = nt->x2; break;
                case 3: k.in = nt->p2; k.out = nt->x3; break;
                case 4: k.in = nt->x3; k.out = nt->a3; k.skip = nt->x2; k.yfine = nt->y2; break;
                case 5: k.in = nt->y2; k.out = nt->a2; k.skip = nt->x1; k.yfine = nt->y1; break;
                case 6: k.in = nt->y1; k.out = nt->a1; k.skip = nt->x0; k.yfine = nt->y0; break;
                default:
                    k.in = nt->y0;
                    k.out = nt->h0;
                    k.proj = params[b] + nt->woff[8];
                    k.out2 = outputs[b];
                }
                t.push_back(k);
                c.push_back(conv_variant_blocks(v, g));
            }
            NetMultiOp op;
            op.variant = v;
            op.cin = st.cin;
            op.cout_pad = st.cout_pad;
            if (!net_multi_op(m, op, t, c)) {
                net_multi_destroy(m);
                return nullptr;
            }
        }
    }
    return m;
}

int net_multi_forward(NetMulti *m, hipStream_t s)
{
    for (const NetMultiOp &op : m->ops) {
        const MultiMap mm{op.map, op.nbox};
        if (op.pool) {
            launch_pool_fwd_multi(s, (const PoolTask *)op.tasks, mm, op.blocks);
        } else if (launch_conv_fwd_multi(s, op.variant, op.cin, op.cout_pad, (const ConvTask *)op.tasks, mm,
                                         op.blocks)) {
            g_net_error = "unsupported layer shape";
            return SPH_ERR_VALUE;
        }
    }
    return SPH_OK;
}

extern "C" {

// Forward pass: input[V0][3] FP32 -> output[V0] FP32.  params: FP32 device
// vector of sph_net_param_count() values.  Keeps the activations for
// sph_net_backward.
int sph_net_forward(sph_net *n, const float *params, const float *input, float *output, void *stream)
{
    if (!n || !params || !input || !output) {
        g_net_error = "null argument";
        return SPH_ERR_ARG;
    }
    net_pack(n, params, (hipStream_t)stream);
    return net_forward_packed(n, params, input, output, (hipStream_t)stream);
}

// Backward pass of the last forward: grad_output[V0] FP32 (dLoss/dout);
// adds dLoss/dparams into grad_params (FP32, param_count values).
int sph_net_backward(sph_net *n, const float *params, const float *grad_output, float *grad_params,
                     void *stream)
{
    if (!n || !params || !grad_output || !grad_params) {
        g_net_error = "null argument";
        return SPH_ERR_ARG;
    }
    // the adjoint weights were packed by the forward pass
    hipStream_t s = (hipStream_t)stream;
    const NetGeom *g = n->g;
    const uint32_t *V = n->V;
    int rc = 0;
    // weight and bias gradient of layer l from input x (cin_x channels) and
    // the BF16 output gradient in n->dz (cdz channels)
    auto wgrad = [&](int l, const uint16_t *x, int cin_x, int cdz, int level) {
        const LayerSpec &L = kLayers[l];
        launch_fill_f32(s, n->gw, 27ull * cdz * cin_x, 0.f);
        rc |= launch_conv_wgrad(s, g[level], x, cin_x, n->dz, cdz, n->gw);
        launch_unpack_wgrad(s, n->gw, L.cin, cin_x, L.cout, cdz, grad_params + n->woff[l]);
        launch_bias_grad(s, V[level], n->dz, cdz, L.cout, grad_params + n->boff[l]);
    };
    // c8: out = w8 . h0 + b8 (weights and bias contiguous)
    launch_proj_bwd(s, V[0], grad_output, n->h0, params + n->woff[8], n->d_h0, grad_params + n->woff[8]);
    // c7: h0 = relu(conv(y0))
    launch_relu_mask(s, V[0], n->d_h0, n->h0, 16, 16, n->dz);
    wgrad(7, n->y0, 16, 16, 0);
    rc |= launch_conv_fwd(s, g[0], n->dz, 16, n->wt[7], 16, nullptr, 16, n->d_y0, NET_OUT_F32);
    // y0 = up(a1) + x0: d_y0 is also the gradient of x0 (accumulated below)
    launch_up_bwd(s, g[1], n->d_y0, 16, n->d_a1);
    // c6: a1 = relu(conv(y1))
    launch_relu_mask(s, V[1], n->d_a1, n->a1, 16, 16, n->dz);
    wgrad(6, n->y1, 32, 16, 1);
    rc |= launch_conv_fwd(s, g[1], n->dz, 16, n->wt[6], 32, nullptr, 32, n->d_y1, NET_OUT_F32);
    launch_up_bwd(s, g[2], n->d_y1, 32, n->d_a2);
    // c5: a2 = relu(conv(y2))
    launch_relu_mask(s, V[2], n->d_a2, n->a2, 32, 32, n->dz);
    wgrad(5, n->y2, 48, 32, 2);
    rc |= launch_conv_fwd(s, g[2], n->dz, 32, n->wt[5], 48, nullptr, 48, n->d_y2, NET_OUT_F32);
    launch_up_bwd(s, g[3], n->d_y2, 48, n->d_a3);
    // c4: a3 = relu(conv(x3))
    launch_relu_mask(s, V[3], n->d_a3, n->a3, 48, 48, n->dz);
    wgrad(4, n->x3, 64, 48, 3);
    rc |= launch_conv_fwd(s, g[3], n->dz, 48, n->wt[4], 64, nullptr, 64, n->d_x3, NET_OUT_F32);
    // c3: x3 = relu(conv(p2))
    launch_relu_mask(s, V[3], n->d_x3, n->x3, 64, 64, n->dz);
    wgrad(3, n->p2, 48, 64, 3);
    rc |= launch_conv_fwd(s, g[3], n->dz, 64, n->wt[3], 48, nullptr, 48, n->d_p2, NET_OUT_F32);
    launch_pool_bwd(s, g[2], n->d_p2, 48, n->d_y2);  // x2's gradient: skip + pool paths
    // c2: x2 = relu(conv(p1))
    launch_relu_mask(s, V[2], n->d_y2, n->x2, 48, 48, n->dz);
    wgrad(2, n->p1, 32, 48, 2);
    rc |= launch_conv_fwd(s, g[2], n->dz, 48, n->wt[2], 32, nullptr, 32, n->d_p1, NET_OUT_F32);
    launch_pool_bwd(s, g[1], n->d_p1, 32, n->d_y1);
    // c1: x1 = relu(conv(p0))
    launch_relu_mask(s, V[1], n->d_y1, n->x1, 32, 32, n->dz);
    wgrad(1, n->p0, 16, 32, 1);
    rc |= launch_conv_fwd(s, g[1], n->dz, 32, n->wt[1], 16, nullptr, 16, n->d_p0, NET_OUT_F32);
    launch_pool_bwd(s, g[0], n->d_p0, 16, n->d_y0);
    // c0: x0 = relu(conv(in))
    launch_relu_mask(s, V[0], n->d_y0, n->x0, 16, 16, n->dz);
    wgrad(0, n->in4, 4, 16, 0);
    if (rc) {
        g_net_error = "unsupported layer shape";
        return SPH_ERR_VALUE;
    }
    if (hipGetLastError() != hipSuccess) {
        g_net_error = "kernel launch failed";
        return SPH_ERR_HIP;
    }
    return SPH_OK;
}

// One 3x3x3 convolution as the network runs it, for tests: in[V][cin]
// FP32 (cin 3 or 4: padded to 4 channels; else a multiple of 16), weights
// w[27][cout][cin] FP32, b[cout] or NULL, out[V][cout] FP32 (cout a
// multiple of 16 or less than 16).  Operands are rounded to BF16, sums are
// FP32.  Device pointers; allocates and frees its own temporaries.
int sph_net_conv3d(const float *in, int cin, const float *w, const float *b, int cout, uint32_t nx,
                   uint32_t ny, uint32_t nz, uint32_t periodic, float *out, void *stream)
{
    if (!in || !w || !out || cin <= 0 || cout <= 0 || cout > 64 || nx == 0 || ny == 0 || nz == 0) {
        g_net_error = "bad argument";
        return SPH_ERR_ARG;
    }
    const int cpad = pad_in(cin), opad = pad_out(cout);
    if (cpad > 64) {
        g_net_error = "too many input channels";
        return SPH_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const NetGeom g{nx, ny, nz, periodic & 3u};
    const uint32_t V = nx * ny * nz;
    uint16_t *inb = nullptr, *wf = nullptr;
    if (hipMalloc((void **)&inb, (size_t)V * cpad * 2) != hipSuccess ||
        hipMalloc((void **)&wf, 27ull * opad * cpad * 2) != hipSuccess) {
        (void)hipFree(inb);
        g_net_error = "device allocation failed";
        return SPH_ERR_HIP;
    }
    launch_to_bf16_pad(s, V, in, cin, cpad, inb);
    launch_pack_weights(s, w, cin, cout, wf, cpad, opad, nullptr, 0, 0);
    int rc = launch_conv_fwd(s, g, inb, cpad, wf, opad, b, cout, out, NET_OUT_F32);
    (void)hipStreamSynchronize(s);
    (void)hipFree(inb);
    (void)hipFree(wf);
    if (rc) {
        g_net_error = "unsupported shape";
        return SPH_ERR_VALUE;
    }
    return SPH_OK;
}

int sph_net_conv3d_noalloc(const float *in, int cin, const float *w, const float *b, int cout,
                           uint32_t nx, uint32_t ny, uint32_t nz, uint32_t periodic, float *out,
                           void *stream)
{
    if (!in || !w || !out || cin <= 0 || cout <= 0 || cout > 64 || nx == 0 || ny == 0 || nz == 0) {
        g_net_error = "bad argument";
        return SPH_ERR_ARG;
    }
    const int cpad = pad_in(cin), opad = pad_out(cout);
    if (cpad > 64) {
        g_net_error = "too many input channels";
        return SPH_ERR_ARG;
    }
    const uint64_t V64 = (uint64_t)nx * ny * nz;
    const uint64_t in_bytes = (V64 * (uint64_t)cpad * 2u + 255u) / 256u * 256u;
    const uint64_t w_bytes = 27ull * (uint64_t)opad * (uint64_t)cpad * 2u;
    if (V64 > 0xffffffffull || in_bytes + w_bytes > NET_NOALLOC_BYTES) {
        g_net_error = "sph_net_conv3d_noalloc: the input (" + std::to_string(V64) + " voxels x " +
                      std::to_string(cpad) + " channels) and weights need " +
                      std::to_string(in_bytes + w_bytes) + " bytes of scratch; at most " +
                      std::to_string(NET_NOALLOC_BYTES) + " are available";
        return SPH_ERR_ARG;
    }
    unsigned char *base = noalloc_scratch();
    if (!base) return SPH_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const NetGeom g{nx, ny, nz, periodic & 3u};
    uint16_t *inb = (uint16_t *)base;
    uint16_t *wf = (uint16_t *)(base + in_bytes);
    // the staging kernels write every element they reserve (padding as
    // zeros), so nothing in the scratch needs clearing between calls
    launch_to_bf16_pad(s, (uint32_t)V64, in, cin, cpad, inb);
    launch_pack_weights(s, w, cin, cout, wf, cpad, opad, nullptr, 0, 0);
    if (launch_conv_fwd(s, g, inb, cpad, wf, opad, b, cout, out, NET_OUT_F32)) {
        g_net_error = "unsupported shape";
        return SPH_ERR_VALUE;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_net_error = std::string("launch failed: ") + hipGetErrorString(e);
        return SPH_ERR_HIP;
    }
    return SPH_OK;
}

}  // extern "C"
