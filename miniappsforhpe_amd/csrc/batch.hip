// Batched fused TTI time steps (include/tti_batch.h): all shots of a batch advance by one step in one launch.
//
// Every shot keeps its own volumes, ranges and strides. The launch covers the concatenated tile lists of all shots
// (tiles of 64 x rows x zblock, as the single-volume kernels): block t looks its shot up in a tile-to-shot map and
// reads the shot's descriptor from a device table with scalar loads, then runs the same tile body as the
// single-volume kernel. The per-step values (ping-pong parity, time index for the source amplitude, snapshot slot)
// are kernel arguments, so a whole pass is a fixed sequence of launches that is captured once into a hipGraph.
//
// Uneven apertures need no padding: a 128^3 shot contributes 2 x 32 x 1 tiles, a 320^3 shot 5 x 80 x 3, and the
// hardware dispatcher fills the CUs from the combined list (small shots alone leave most of the 256 CUs idle).
#include "fused_tiles.h"
#include "tti_batch.h"

#include <array>
#include <map>
#include <new>
#include <vector>

namespace tti {
namespace fused {

struct ShotDev
{
    float* s1[2];
    float* s2[2];
    float* u;
    float* w;
    const float* m[8];
    const ab2* ab[3];
    Box g;
    unsigned src_off;
    const float* src_amp;
    Snap snap;       // slot 0
    long slot_bytes; // bytes between snapshot slots of snap.a / snap.b
    long scale_slot; // floats between slots of snap.sa / snap.sb
    double* grad;
    const long* rec_idx;
    const int* rec_ijk;
    int nrec;
    float* records;
    const float* residual;
    long nquad; // float4 words of every volume of the shot (planes 0 .. izend + 4)
    int bz, ntz, strip;
    int ntx[2], nty[2], tile0[2]; // [0]: forward tiles (64 x FWD_BY), [1]: adjoint tiles (ADJ_AX x ADJ_AY)
};

// Block -> (shot, tile) with the XCD-aware order of decode_tile over the concatenated tile list of pass P
// (0 forward, 1 adjoint).
template <int P>
__device__ __forceinline__ const ShotDev*
batch_tile(const ShotDev* __restrict__ shots, const int* __restrict__ tile_shot, int ntiles, int& bxi, int& byi,
           int& bzi)
{
    const int per = (ntiles + 7) / 8;
    const int t = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= ntiles)
        return nullptr;
    const ShotDev* S = shots + tile_shot[t];
    const int lt = t - S->tile0[P];
    const int per_chunk = S->ntx[P] * S->nty[P];
    bzi = lt / per_chunk;
    tile_xy(lt - bzi * per_chunk, S->ntx[P], S->nty[P], S->strip, bxi, byi);
    return S;
}

__device__ __forceinline__ Snap
slot_snap(const ShotDev& S, int slot)
{
    Snap s = S.snap;
    s.a = static_cast<char*>(s.a) + (long)slot * S.slot_bytes;
    s.b = static_cast<char*>(s.b) + (long)slot * S.slot_bytes;
    if (s.sa != nullptr) {
        s.sa += (long)slot * S.scale_slot;
        s.sb += (long)slot * S.scale_slot;
    }
    return s;
}

template <int SNAP, bool SPONGE>
__global__ __launch_bounds__(BX* FWD_BY, FWD_BY <= 4 ? 2 : 1) void k_fused_fwd_batch(const ShotDev* __restrict__ shots,
                                                                    const int* __restrict__ tile_shot, int ntiles,
                                                                    int par, int it, int slot, float dt, float invdx,
                                                                    float invdy, float invdz, float i4xy, float i4yz,
                                                                    float i4zx)
{
    int bxi, byi, bzi;
    const ShotDev* S = batch_tile<0>(shots, tile_shot, ntiles, bxi, byi, bzi);
    if (S == nullptr)
        return;
    const Snap sn = SNAP != TTI_SNAP_NONE ? slot_snap(*S, slot) : S->snap;
    const float sv = S->src_amp != nullptr ? S->src_amp[it] : 0.0f;
    fwd_tile<FWD_BY, SNAP, SPONGE>(S->s1[par], S->s2[par], S->s1[par ^ 1], S->s2[par ^ 1], S->u, S->w, S->m[0], S->m[1],
                                   S->m[2], S->m[3], S->m[4], S->m[5], S->m[6], S->m[7], S->ab[0], S->ab[1], S->ab[2],
                                   S->g, dt, invdx, invdy, invdz, i4xy, i4yz, i4zx, S->src_off, sv, sn, S->bz, bxi,
                                   byi, bzi);
}

template <int SNAP, bool SPONGE>
__global__ __launch_bounds__(ADJ_AX* ADJ_AY, ADJ_AX* ADJ_AY <= 256 ? 2 : 1) void k_fused_adj_batch(const ShotDev* __restrict__ shots,
                                                            const int* __restrict__ tile_shot, int ntiles, int par,
                                                            int slot, float dt, float invdx, float invdy, float invdz,
                                                            float i4xy, float i4yz, float i4zx)
{
    int bxi, byi, bzi;
    const ShotDev* S = batch_tile<1>(shots, tile_shot, ntiles, bxi, byi, bzi);
    if (S == nullptr)
        return;
    const Snap sn = SNAP != TTI_SNAP_NONE ? slot_snap(*S, slot) : S->snap;
    adj_tile<SNAP, SPONGE>(S->s1[par], S->s2[par], S->s1[par ^ 1], S->s2[par ^ 1], S->u, S->w, S->m[0], S->m[1],
                           S->m[2], S->m[3], S->m[4], S->m[5], S->m[6], S->m[7], S->ab[0], S->ab[1], S->ab[2], S->g, dt,
                           invdx, invdy, invdz, i4xy, i4yz, i4zx, sn, S->grad, S->bz, bxi, byi, bzi);
}

// receivers of every shot (blockIdx.y = shot) after forward step `it`: sigma is the buffer the step wrote
__global__ __launch_bounds__(256) void k_gather_batch(const ShotDev* __restrict__ shots, int par_out, int it)
{
    const ShotDev& S = shots[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < S.nrec)
        S.records[(long)it * S.nrec + k] = pressure_at(S.s1[par_out], S.s2[par_out], S.rec_idx[k]);
}

// residual row n-1 of every shot after the adjoint step that yields mu_n (slot >= 0: with its imaging term)
__global__ __launch_bounds__(256) void k_inject_batch(const ShotDev* __restrict__ shots, int par_out, int n, int slot,
                                                      float dt)
{
    const ShotDev& S = shots[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S.nrec)
        return;
    Snap sn = S.snap;
    if (slot >= 0)
        sn = slot_snap(S, slot);
    else
        sn.mode = TTI_SNAP_NONE;
    inject_point(S.u, S.w, S.s1[par_out], S.s2[par_out], S.rec_idx, S.rec_ijk, k, S.residual + (long)(n - 1) * S.nrec,
                 dt, S.ab[0], S.ab[1], S.ab[2], S.g, sn, S.grad);
}

// rest state: zero sigma (both buffers), u and w of every shot (blockIdx.y = 6 * shot + field)
__global__ __launch_bounds__(256) void k_zero_batch(const ShotDev* __restrict__ shots)
{
    const ShotDev& S = shots[blockIdx.y / 6];
    const int f = blockIdx.y % 6;
    float* p = f < 2 ? S.s1[f] : f < 4 ? S.s2[f - 2] : f == 4 ? S.u : S.w;
    float4* q = reinterpret_cast<float4*>(p);
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < S.nquad; i += (long)gridDim.x * blockDim.x)
        q[i] = z;
}

} // namespace fused
} // namespace tti

// ======================================================================================================================
// C ABI
// ======================================================================================================================
using namespace tti;
using namespace tti::fused;

struct tti_batch
{
    int nshots = 0, ntiles[2] = { 0, 0 }, mode = TTI_SNAP_NONE, maxrec = 0, nt_max = 0, slots = 0;
    bool sponge = false, have_grad = false, have_residual = false;
    long maxquad = 0;
    float dt = 0, invdx = 0, invdy = 0, invdz = 0, i4xy = 0, i4yz = 0, i4zx = 0;
    ShotDev* shots = nullptr;
    int* tile_shot = nullptr; // forward map (ntiles[0] entries) followed by the adjoint map
    hipStream_t cs = nullptr; // capture / replay stream (non-blocking, default priority)
    hipEvent_t ev[2] = { nullptr, nullptr };
    std::map<std::array<int, 3>, hipGraphExec_t> graphs; // (pass, nt, k) -> executable graph
};

namespace {

int
fwd_step(tti_batch* b, int par, int it, int slot, hipStream_t st)
{
    if (par < 0 || par > 1 || it < 0 || it >= b->nt_max || slot >= b->slots ||
        (slot >= 0 && b->mode == TTI_SNAP_NONE))
        return TTI_EINVAL;
    const dim3 grid(((b->ntiles[0] + 7) / 8) * 8), block(BX, FWD_BY);
    auto go = [&](auto kernel) {
        kernel<<<grid, block, 0, st>>>(b->shots, b->tile_shot, b->ntiles[0], par, it, slot, b->dt, b->invdx, b->invdy,
                                       b->invdz, b->i4xy, b->i4yz, b->i4zx);
    };
    const bool sp = b->sponge;
    switch (slot >= 0 ? b->mode : TTI_SNAP_NONE) {
        case TTI_SNAP_FP32: sp ? go(k_fused_fwd_batch<TTI_SNAP_FP32, true>) : go(k_fused_fwd_batch<TTI_SNAP_FP32, false>); break;
        case TTI_SNAP_FP16: sp ? go(k_fused_fwd_batch<TTI_SNAP_FP16, true>) : go(k_fused_fwd_batch<TTI_SNAP_FP16, false>); break;
        case TTI_SNAP_BF16: sp ? go(k_fused_fwd_batch<TTI_SNAP_BF16, true>) : go(k_fused_fwd_batch<TTI_SNAP_BF16, false>); break;
        default: sp ? go(k_fused_fwd_batch<TTI_SNAP_NONE, true>) : go(k_fused_fwd_batch<TTI_SNAP_NONE, false>); break;
    }
    int rc = launch_status();
    if (rc == TTI_OK && b->maxrec > 0) {
        k_gather_batch<<<dim3((b->maxrec + 255) / 256, b->nshots), dim3(256), 0, st>>>(b->shots, par ^ 1, it);
        rc = launch_status();
    }
    return rc;
}

int
adj_step(tti_batch* b, int par, int n, int slot, hipStream_t st)
{
    if (par < 0 || par > 1 || n < 1 || n > b->nt_max || slot >= b->slots ||
        (slot >= 0 && (b->mode == TTI_SNAP_NONE || !b->have_grad)))
        return TTI_EINVAL;
    if (b->maxrec > 0 && !b->have_residual)
        return TTI_EINVAL;
    const dim3 grid(((b->ntiles[1] + 7) / 8) * 8), block(ADJ_AX, ADJ_AY);
    auto go = [&](auto kernel) {
        kernel<<<grid, block, 0, st>>>(b->shots, b->tile_shot + b->ntiles[0], b->ntiles[1], par, slot, b->dt, b->invdx, b->invdy,
                                       b->invdz, b->i4xy, b->i4yz, b->i4zx);
    };
    const bool sp = b->sponge;
    switch (slot >= 0 ? b->mode : TTI_SNAP_NONE) {
        case TTI_SNAP_FP32: sp ? go(k_fused_adj_batch<TTI_SNAP_FP32, true>) : go(k_fused_adj_batch<TTI_SNAP_FP32, false>); break;
        case TTI_SNAP_FP16: sp ? go(k_fused_adj_batch<TTI_SNAP_FP16, true>) : go(k_fused_adj_batch<TTI_SNAP_FP16, false>); break;
        case TTI_SNAP_BF16: sp ? go(k_fused_adj_batch<TTI_SNAP_BF16, true>) : go(k_fused_adj_batch<TTI_SNAP_BF16, false>); break;
        default: sp ? go(k_fused_adj_batch<TTI_SNAP_NONE, true>) : go(k_fused_adj_batch<TTI_SNAP_NONE, false>); break;
    }
    int rc = launch_status();
    if (rc == TTI_OK && b->maxrec > 0) {
        k_inject_batch<<<dim3((b->maxrec + 255) / 256, b->nshots), dim3(256), 0, st>>>(b->shots, par ^ 1, n, slot,
                                                                                         b->dt);
        rc = launch_status();
    }
    return rc;
}

int
reset(tti_batch* b, hipStream_t st)
{
    const long blocks = (b->maxquad + 255) / 256;
    const unsigned gx = (unsigned)(blocks < 1024 ? blocks : 1024);
    k_zero_batch<<<dim3(gx, 6 * b->nshots), dim3(256), 0, st>>>(b->shots);
    return launch_status();
}

int
forward_pass(tti_batch* b, int nt, int k, hipStream_t st)
{
    int rc = reset(b, st);
    for (int it = 0; it < nt && rc == TTI_OK; ++it) {
        const int n = it + 1;
        const int slot = (b->mode != TTI_SNAP_NONE && n % k == 0) ? n / k - 1 : -1;
        rc = fwd_step(b, it & 1, it, slot, st);
    }
    return rc;
}

int
adjoint_pass(tti_batch* b, int nt, int k, hipStream_t st)
{
    int rc = reset(b, st);
    for (int i = 0; i < nt && rc == TTI_OK; ++i) {
        const int n = nt - i;
        const int slot = (b->mode != TTI_SNAP_NONE && b->have_grad && n % k == 0) ? n / k - 1 : -1;
        rc = adj_step(b, i & 1, n, slot, st);
    }
    return rc;
}

// Runs `pass` on `st` directly, or through the cached executable graph of (kind, nt, k) on the batch's own stream,
// ordered after the work already on `st` and before the work issued to it later.
template <class F>
int
run_pass(tti_batch* b, int kind, int nt, int k, bool use_graph, hipStream_t st, F&& pass)
{
    if (!use_graph)
        return pass(st);
    int rc;
    if (b->cs == nullptr) {
        if ((rc = hip_status(hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking))) != TTI_OK)
            return rc;
        for (auto& e : b->ev)
            if ((rc = hip_status(hipEventCreateWithFlags(&e, hipEventDisableTiming))) != TTI_OK)
                return rc;
    }
    const std::array<int, 3> key{ kind, nt, k };
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        if ((rc = hip_status(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal))) != TTI_OK)
            return rc;
        rc = pass(b->cs);
        hipGraph_t g = nullptr;
        const int rc_end = hip_status(hipStreamEndCapture(b->cs, &g));
        if (rc == TTI_OK)
            rc = rc_end;
        hipGraphExec_t x = nullptr;
        if (rc == TTI_OK)
            rc = hip_status(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
        if (g != nullptr)
            (void)hipGraphDestroy(g);
        if (rc != TTI_OK)
            return rc;
        it = b->graphs.emplace(key, x).first;
    }
    if ((rc = hip_status(hipEventRecord(b->ev[0], st))) != TTI_OK ||
        (rc = hip_status(hipStreamWaitEvent(b->cs, b->ev[0], 0))) != TTI_OK ||
        (rc = hip_status(hipGraphLaunch(it->second, b->cs))) != TTI_OK ||
        (rc = hip_status(hipEventRecord(b->ev[1], b->cs))) != TTI_OK)
        return rc;
    return hip_status(hipStreamWaitEvent(st, b->ev[1], 0));
}

void
release(tti_batch* b)
{
    for (auto& kv : b->graphs)
        (void)hipGraphExecDestroy(kv.second);
    for (auto e : b->ev)
        if (e != nullptr)
            (void)hipEventDestroy(e);
    if (b->cs != nullptr)
        (void)hipStreamDestroy(b->cs);
    if (b->shots != nullptr)
        (void)hipFree(b->shots);
    if (b->tile_shot != nullptr)
        (void)hipFree(b->tile_shot);
    delete b;
}

} // namespace

extern "C" {

int
tti_batch_create(const tti_shot* shots, int nshots, float dt, double dx, double dy, double dz, tti_batch** out)
{
    if (out == nullptr)
        return TTI_EINVAL;
    *out = nullptr;
    if (shots == nullptr || nshots < 1 || nshots > 10000 || !(dx > 0 && dy > 0 && dz > 0)) // grid.y <= 65535
        return TTI_EINVAL;
    std::vector<ShotDev> h(nshots);
    const int mode = shots[0].snap.mode;
    const bool sponge = shots[0].AB[0] != nullptr;
    long tiles[2] = { 0, 0 }, maxquad = 0;
    int maxrec = 0, nt_max = 1 << 30, slots = mode == TTI_SNAP_NONE ? 0 : 1 << 30;
    bool have_grad = true, have_res = true;
    for (int s = 0; s < nshots; ++s) {
        const tti_shot& q = shots[s];
        if (!q.field1[0] || !q.field1[1] || !q.field2[0] || !q.field2[1] || !q.field3 || !q.field4 ||
            q.field1[0] == q.field1[1] || q.field2[0] == q.field2[1])
            return TTI_EINVAL;
        for (const float* m : q.model)
            if (m == nullptr)
                return TTI_EINVAL;
        if (!box_ok(q.ixbeg, q.ixend, q.iybeg, q.iyend, q.izbeg, q.izend, q.ldimx, q.ldimy) ||
            !small_volume(q.ldimx, q.ldimy, q.izend))
            return TTI_EINVAL;
        if (!aligned16({ q.field1[0], q.field1[1], q.field2[0], q.field2[1], q.field3, q.field4, q.model[0],
                         q.model[1], q.model[2], q.model[3], q.model[4], q.model[5], q.model[6], q.model[7] }))
            return TTI_EINVAL;
        const bool sp = q.AB[0] != nullptr;
        if (sp != sponge || (sp && !(q.AB[1] && q.AB[2])) || (!sp && (q.AB[1] || q.AB[2])))
            return TTI_EINVAL;
        const int nx = q.ixend - q.ixbeg + 1, ny = q.iyend - q.iybeg + 1, nz = q.izend - q.izbeg + 1;
        if (q.snap.mode != mode || !snap_ok(&q.snap, nx, ny, nz) || q.snap.field1 != nullptr)
            return TTI_EINVAL; // (sigma snapshots for the density image are not batched)
        if (mode != TTI_SNAP_NONE &&
            (q.snap_slots < 1 || q.snap_slot_stride % 64 != 0 ||
             q.snap_slot_stride < (long)nz * q.snap.ldimy * q.snap.ldimx))
            return TTI_EINVAL;
        if (q.nt_max < 1 || q.src_index >= (long)q.ldimx * q.ldimy * (q.izend + R + 1) ||
            (q.src_index >= 0 && q.src_amplitude == nullptr))
            return TTI_EINVAL;
        if (q.nrec < 0 || (q.nrec > 0 && (!q.rec_index || !q.rec_ijk || !q.records)))
            return TTI_EINVAL;
        const int bz = q.zblock > 0 ? q.zblock : 128;
        ShotDev& d = h[s];
        d.s1[0] = q.field1[0];
        d.s1[1] = q.field1[1];
        d.s2[0] = q.field2[0];
        d.s2[1] = q.field2[1];
        d.u = q.field3;
        d.w = q.field4;
        for (int i = 0; i < 8; ++i)
            d.m[i] = q.model[i];
        for (int i = 0; i < 3; ++i)
            d.ab[i] = static_cast<const ab2*>(q.AB[i]);
        d.g = Box{ q.ixbeg, q.ixend, q.iybeg, q.iyend, q.izbeg, q.izend, (long)q.ldimx, (long)q.ldimy };
        d.src_off = q.src_index < 0 ? 0xffffffffu : (unsigned)(q.src_index * 4);
        d.src_amp = q.src_index < 0 ? nullptr : q.src_amplitude;
        d.snap = to_snap(&q.snap);
        const long esize = mode == TTI_SNAP_FP32 ? 4 : 2;
        d.slot_bytes = mode == TTI_SNAP_NONE ? 0 : q.snap_slot_stride * esize;
        d.scale_slot = mode == TTI_SNAP_FP16 ? q.snap_slot_stride / 64 : 0;
        d.grad = q.grad;
        d.rec_idx = q.rec_index;
        d.rec_ijk = q.rec_ijk;
        d.nrec = q.nrec;
        d.records = q.records;
        d.residual = q.residual;
        d.nquad = (long)q.ldimx * q.ldimy * (q.izend + R + 1) / 4;
        d.bz = bz;
        d.strip = q.strip;
        d.ntz = (nz + bz - 1) / bz;
        d.ntx[0] = (nx + BX - 1) / BX;
        d.nty[0] = (ny + FWD_BY - 1) / FWD_BY;
        d.ntx[1] = (nx + ADJ_AX - 1) / ADJ_AX;
        d.nty[1] = (ny + ADJ_AY - 1) / ADJ_AY;
        for (int pss = 0; pss < 2; ++pss) {
            d.tile0[pss] = (int)tiles[pss];
            tiles[pss] += (long)d.ntx[pss] * d.nty[pss] * d.ntz;
        }
        if (tiles[0] + tiles[1] >= (1L << 30))
            return TTI_EINVAL;
        maxquad = d.nquad > maxquad ? d.nquad : maxquad;
        maxrec = q.nrec > maxrec ? q.nrec : maxrec;
        nt_max = q.nt_max < nt_max ? q.nt_max : nt_max;
        if (mode != TTI_SNAP_NONE)
            slots = q.snap_slots < slots ? q.snap_slots : slots;
        have_grad = have_grad && q.grad != nullptr;
        have_res = have_res && (q.nrec == 0 || q.residual != nullptr);
    }
    std::vector<int> map(tiles[0] + tiles[1]);
    for (int s = 0; s < nshots; ++s)
        for (int pss = 0; pss < 2; ++pss) {
            const long n = (long)h[s].ntx[pss] * h[s].nty[pss] * h[s].ntz;
            const long base = (pss ? tiles[0] : 0) + h[s].tile0[pss];
            for (long t = 0; t < n; ++t)
                map[base + t] = s;
        }
    tti_batch* b = new (std::nothrow) tti_batch;
    if (b == nullptr)
        return TTI_EINVAL;
    b->nshots = nshots;
    b->ntiles[0] = (int)tiles[0];
    b->ntiles[1] = (int)tiles[1];
    b->mode = mode;
    b->maxrec = maxrec;
    b->nt_max = nt_max;
    b->slots = slots;
    b->sponge = sponge;
    b->have_grad = have_grad;
    b->have_residual = have_res;
    b->maxquad = maxquad;
    b->dt = dt;
    b->invdx = (float)(1.0 / dx);
    b->invdy = (float)(1.0 / dy);
    b->invdz = (float)(1.0 / dz);
    b->i4xy = (float)(1.0 / (4 * dx * dy));
    b->i4yz = (float)(1.0 / (4 * dy * dz));
    b->i4zx = (float)(1.0 / (4 * dx * dz));
    int rc = hip_status(hipMalloc(&b->shots, sizeof(ShotDev) * nshots));
    if (rc == TTI_OK)
        rc = hip_status(hipMalloc(&b->tile_shot, sizeof(int) * map.size()));
    if (rc == TTI_OK)
        rc = hip_status(hipMemcpy(b->shots, h.data(), sizeof(ShotDev) * nshots, hipMemcpyHostToDevice));
    if (rc == TTI_OK)
        rc = hip_status(hipMemcpy(b->tile_shot, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
    if (rc != TTI_OK) {
        release(b);
        return rc;
    }
    *out = b;
    return TTI_OK;
}

int
tti_batch_shot_bytes(void)
{
    return (int)sizeof(tti_shot);
}

int
tti_batch_destroy(tti_batch* batch)
{
    if (batch == nullptr)
        return TTI_EINVAL;
    release(batch);
    return TTI_OK;
}

int
tti_batch_tiles(const tti_batch* batch)
{
    return batch == nullptr ? TTI_EINVAL : batch->ntiles[0];
}

int
tti_batch_reset(tti_batch* batch, void* stream)
{
    return batch == nullptr ? TTI_EINVAL : reset(batch, (hipStream_t)stream);
}

int
tti_batch_fwd_step(tti_batch* batch, int parity, int it, int snap_slot, void* stream)
{
    return batch == nullptr ? TTI_EINVAL : fwd_step(batch, parity, it, snap_slot, (hipStream_t)stream);
}

int
tti_batch_adj_step(tti_batch* batch, int parity, int n, int snap_slot, void* stream)
{
    return batch == nullptr ? TTI_EINVAL : adj_step(batch, parity, n, snap_slot, (hipStream_t)stream);
}

int
tti_batch_forward(tti_batch* batch, int nt, int k, int use_graph, void* stream)
{
    if (batch == nullptr || nt < 1 || nt > batch->nt_max || k < 1 ||
        (batch->mode != TTI_SNAP_NONE && nt / k > batch->slots))
        return TTI_EINVAL;
    return run_pass(batch, 0, nt, k, use_graph != 0, (hipStream_t)stream,
                    [&](hipStream_t st) { return forward_pass(batch, nt, k, st); });
}

int
tti_batch_adjoint(tti_batch* batch, int nt, int k, int use_graph, void* stream)
{
    if (batch == nullptr || nt < 1 || nt > batch->nt_max || k < 1 ||
        (batch->mode != TTI_SNAP_NONE && batch->have_grad && nt / k > batch->slots))
        return TTI_EINVAL;
    if (batch->maxrec > 0 && !batch->have_residual)
        return TTI_EINVAL;
    return run_pass(batch, 1, nt, k, use_graph != 0, (hipStream_t)stream,
                    [&](hipStream_t st) { return adjoint_pass(batch, nt, k, st); });
}

} // extern "C"
