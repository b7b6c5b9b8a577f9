// symmetry::Backend on libmvs (MI355X).  Compiled together with the
// detector's own sources: it includes the detector's headers
// (include/symmetry/pipeline.h, cuda.h) and links libmvs.so.
#pragma once
#include <memory>
#include <vector>

#include "symmetry/cuda.h"
#include "symmetry/pipeline.h"

namespace symmetry {

// Same contract as make_cuda_backend (include/symmetry/cuda.h:6): throws
// std::runtime_error("HIP: no available device") when no GPU is visible.
// Every call borrows host views and returns host-owned results; device
// memory lives only for the duration of the call.  `device` selects the GPU.
std::unique_ptr<Backend> make_hip_backend(bool profile = false, int device = 0);

// Per-stage metrics of the calling thread's profiled calls, in the fields of
// CudaMetric (include/symmetry/cuda.h:7-15); drained on each call.
std::vector<CudaMetric> take_hip_metrics();

}  // namespace symmetry
