// FP32 instantiation of the step driver (fields, coefficients, UPML and Debye
// state in single precision; SIBC state and DFT accumulators stay FP64).
#include "vgpu_run.h"
template int vgpu::run_steps<float>(vgpu::Ctx *, uint64_t, hipStream_t, bool);
template int vgpu::prepare_sweeps<float>(vgpu::Ctx *);
