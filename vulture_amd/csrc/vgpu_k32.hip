// FP32 instantiation of the step driver (fields, coefficients, UPML and Debye
// state in single precision; SIBC state and DFT accumulators stay FP64).
#include "vgpu_run.h"
template int vgpu::run_steps<float>(vgpu::Ctx *, uint64_t, hipStream_t, bool, const vgpu::VisitRun *);
template int vgpu::prepare_sweeps<float>(vgpu::Ctx *);
template int vgpu::upload_pack<float>(vgpu::Ctx *);
template int vgpu::run_group_steps<float>(vgpu::Ctx *const *, int, uint64_t, hipStream_t);
template int vgpu::check_limits<float>(vgpu::Ctx *, const int (*)[6], vgpu_limits_report *, unsigned char *, void *,
                                         size_t);
