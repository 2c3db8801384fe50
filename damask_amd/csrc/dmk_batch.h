// Lockstep batches of RVEs (batch.cpp): the stress integrations of the
// residual evaluations of all RVEs in flight run as one launch.
#pragma once

#include <hip/hip_runtime.h>

struct DmkStressBatch;
struct DmkStressItem;
struct dmk_handle;

// Hand the stress integration of one residual evaluation (its items: one per
// phase, key = 2 nsl + quarter_pow) to the batch.  Called from the RVE's own
// host thread after its input F has been enqueued on `st`; returns once the
// batched launch holding these items is enqueued, with `st` made to wait for
// it.  Blocks until every RVE of the batch still running has submitted.
int dmk_stress_batch_submit(DmkStressBatch* b, int slot, const DmkStressItem* items, const int* key, int n,
                            hipStream_t st);
// attach (b != NULL) or detach a handle (solver.cpp)
void dmk_set_stress_batch(dmk_handle* h, DmkStressBatch* b, int slot);
// whether the batch takes whole residual evaluations of the basic scheme
bool dmk_stress_batch_resid(const DmkStressBatch* b);
struct DmkResidDev;
struct DmkFFT;
// Hand a request of a lockstep round to the batch (struct DmkResidDev: a
// whole residual evaluation with its stress items, an application of the
// adjoint's operator, or dot products; fft is the RVE's FFT configuration and
// must stay valid until the call returns).  With rq->rep_n > 0 it returns
// once the round holding it has finished on the device, with the rq->rep_n
// report doubles copied into report; without a report, once the round is
// launched.  Either way `st` is made to wait for the round; with done_out
// that wait is left to the caller, who gets the round's completion event
// (rounds complete in order: a later round's event covers it, and the event
// may be recorded again by a later round before the caller waits on it, which
// only lengthens the wait).  rq->idle: `st` has had no work since the
// caller's previous round, so the round does not wait for it.
int dmk_resid_batch_submit(DmkStressBatch* b, int slot, const DmkStressItem* items, const int* key, int n,
                           const DmkResidDev* rq, const DmkFFT* fft, hipStream_t st, double* report,
                           hipEvent_t* done_out);
// device bytes that the next reverse sweep of a handle's tape allocates for
// its work buffers (0 when they are allocated or there is no tape), and their
// release after a sweep, the tape kept (solver.cpp)
size_t dmk_adjoint_work_bytes(const dmk_handle* h);
void dmk_adjoint_work_release(dmk_handle* h);
