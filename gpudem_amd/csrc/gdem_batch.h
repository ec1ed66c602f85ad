// Launch interface of the batched kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "gdem_types.h"

namespace gdem {

// One batch step (rebuild phases of the scenes that requested one, then the
// contact step of every running scene) over `grid` workgroups of 256 slots.
template <class T, class H>
void launch_batch_step(const BBufs<H>& B, int grid, hipStream_t st);
// Raises a newly admitted scene's refresh flag of parity par and adds it to
// that parity's due list (in stream order after the earlier batch steps).
template <class H>
void launch_batch_admit(const BBufs<H>& B, int sid, int par, hipStream_t st);
// Scrubs an evicted scene's entries from both due lists.
template <class H>
void launch_batch_evict(const BBufs<H>& B, int sid, hipStream_t st);
template <class H>
void launch_batch_download(const BBufs<H>& B, int sid, int nseg, double* x, double* v, double* w, int* act,
                           hipStream_t st);
template <class H>
void launch_batch_counts(const BBufs<H>& B, int sid, int nseg, unsigned long long* acc, hipStream_t st);

}  // namespace gdem
