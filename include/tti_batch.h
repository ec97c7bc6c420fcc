/*
 * C ABI of the batched fused TTI time steps: many shots, each on its own subvolume of its own size, advance by one
 * forward (or adjoint) step in ONE kernel launch, their receivers in one more. The shot table lives on the device;
 * per-step arguments (ping-pong parity, time index, snapshot slot) are kernel arguments, so whole time loops are
 * captured once into a hipGraph and replayed. Conventions (pointers, ranges, status codes, stream) are those of
 * tti_hip.h and tti_fused.h.
 *
 * No one-to-one counterpart in the reference, which runs one shot per process: each step computes, for every shot,
 * exactly what tti_fused_fwd_step + tti_gather_pressure (forward) or tti_fused_adj_step + tti_inject_residual
 * (adjoint) compute for that shot alone, i.e. the step sequence of simulator2SetUp::execute
 * (simulator2/simulator2SetUp.cu:530-779) on the shot's volume.
 */
#ifndef TTI_BATCH_H
#define TTI_BATCH_H

#include "tti_fused.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One shot of a batch. All pointers are device pointers and must stay valid for the batch's lifetime. */
typedef struct
{
    float* field1[2];          /* sigma_H ping-pong pair: a step with parity p reads [p] and writes [1-p] */
    float* field2[2];          /* sigma_V pair */
    float* field3;             /* u (updated in place) */
    float* field4;             /* w */
    const float* model[8];     /* model1..model4, rx, ry, rz, Model11 (b = 1/rho), same layout as the fields */
    const void* AB[3];         /* float2 {A,B} sponge per x, y, z array index, or all NULL (A = B = 1); all shots of
                                  a batch have a sponge or none has */
    int ixbeg, ixend, iybeg, iyend, izbeg, izend; /* inclusive interior ranges (array coordinates) */
    int ldimx, ldimy;          /* row and plane strides of every volume of the shot */
    long src_index;            /* linear source index, or -1 */
    const float* src_amplitude; /* [nt_max]: added to u and w at the source at forward step it (already times dt) */
    int nt_max;                /* rows of records/residual and length of src_amplitude */
    tti_snapshot snap;         /* slot 0 of the shot's snapshot store; mode NONE: no snapshots/imaging. The mode is
                                  the same for every shot of a batch */
    long snap_slot_stride;     /* elements between consecutive slots of snap.field3/field4 (scale arrays: /64) */
    int snap_slots;            /* number of slots */
    double* grad;              /* FP64 image, interior nz x ny x nx, accumulated by adjoint steps with a slot */
    const long* rec_index;     /* [nrec] linear receiver indices (distinct points) */
    const int* rec_ijk;        /* [3 nrec] array coordinates x, y, z of each receiver */
    int nrec;
    float* records;            /* [nt_max][nrec]: pressure after forward step it goes to row it */
    const float* residual;     /* [nt_max][nrec]: row n-1 is injected after the adjoint step that yields mu_n */
    int zblock;                /* z planes per thread block (0 = 128) */
    int strip;                 /* block schedule: x tiles per strip (0 = full rows), as in tti_fused_fwd_step */
} tti_shot;

typedef struct tti_batch tti_batch;

/* sizeof(tti_shot), for bindings that mirror the struct to check their layout. */
int tti_batch_shot_bytes(void);

/* Validates every shot (as tti_fused_fwd_step does), uploads the shot table and the tile-to-shot map. */
int tti_batch_create(const tti_shot* shots, int nshots, float dt, double dx, double dy, double dz, tti_batch** out);
int tti_batch_destroy(tti_batch* batch);

/* Number of thread blocks of one batched forward step (all shots' tiles; the adjoint uses 64 x 8 tiles). */
int tti_batch_tiles(const tti_batch* batch);

/* Zeroes field1[0..1], field2[0..1], field3 and field4 of every shot (the rest state; fill kernels, no memset). */
int tti_batch_reset(tti_batch* batch, void* stream);

/* Forward step `it` with ping-pong parity `parity` for every shot, then the receiver pressures into records[it].
 * snap_slot >= 0 stores u, w of the step into that slot (the batch's snapshot mode). */
int tti_batch_fwd_step(tti_batch* batch, int parity, int it, int snap_slot, void* stream);

/* Adjoint step yielding mu_n for every shot (ping-pong parity `parity`), then the injection of residual[n-1];
 * snap_slot >= 0 adds the imaging condition with that slot to grad (FP64). */
int tti_batch_adj_step(tti_batch* batch, int parity, int n, int snap_slot, void* stream);

/* Whole passes from rest. Forward: reset, steps it = 0..nt-1 (parity it mod 2) with snapshots of u_n, w_n at
 * n = it+1 = k, 2k, ... into slot n/k - 1. Adjoint: reset, steps yielding mu_n for n = nt..1 (parity (nt-n) mod 2)
 * with imaging from slot n/k - 1 at n = k, 2k, ... (grad is accumulated, not zeroed). use_graph != 0 captures the
 * pass into a hipGraph on the first call with these (nt, k) and replays it afterwards, ordered after prior work on
 * `stream` and before later work on it. */
int tti_batch_forward(tti_batch* batch, int nt, int k, int use_graph, void* stream);
int tti_batch_adjoint(tti_batch* batch, int nt, int k, int use_graph, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TTI_BATCH_H */
